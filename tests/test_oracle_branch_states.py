"""The branch-driving states of tests/branch_states.py really exercise each guarded branch in the
oracle, and switching that branch off changes the result (so a GPU match on them is a match of
the branch, not of a path that never runs).  CPU only."""
import numpy as np
import pytest

from branch_states import KINDS, pocket_state
from oracle import Scheme

G = 5.0 / 3.0
NX, NY, NZ = 24, 20, 16


def scheme_for(orc, u, kind, **kw):
    k = KINDS[kind]
    pfl = k["smallr"] * k["smallc"] ** 2 / G
    dt = 0.4 * (1.0 / NX) / orc.ctot_max(u, G, k["smallr"], pfl)
    ch = 0.4 * (1.0 / NX) / dt / 3 * 0.25
    base = dict(gamma=G, smallr=k["smallr"], smallc=k["smallc"], dt=dt, dx=1.0 / NX, ch=ch, glm_fac=0.99,
                llf_dmin=k["llf_dmin"], llf_pmin=k["llf_pmin"], predictor="full", recon="ppm", riemann="hlld",
                mh_revert=True, face_guard="iflo", update_floor=True)
    base.update(kw)
    return Scheme(**base)


def run(orc, u, sc, n=3):
    for _ in range(n):
        u = orc.step(u, sc)
    return u


@pytest.mark.parametrize("kind,branch,off", [
    ("void", "mh_revert", dict(mh_revert=False)),
    ("void", "iflo_revert", dict(face_guard="none")),
    ("void", "llf_switch", dict(llf_dmin=0.0, llf_pmin=0.0)),
    ("degenerate", "update_floor", dict(update_floor=False)),
    ("degenerate", "hlld_degenerate", dict(riemann="hll")),
])
def test_branch_fires_and_matters(orc, kind, branch, off):
    u = pocket_state(orc, NX, NY, NZ, kind)
    sc = scheme_for(orc, u, kind)
    orc.branch_hits(reset=True)
    a = run(orc, u, sc)
    assert orc.branch_hits()[branch] > 0
    assert np.all(np.isfinite(a))
    b = run(orc, u, Scheme(**{**sc.__dict__, **off}))
    assert not np.array_equal(a, b)
