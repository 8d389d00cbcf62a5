"""The branch-driving states of tests/branch_states.py really exercise each guarded branch in the
oracle, and switching that branch off changes the result (so a GPU match on them is a match of
the branch, not of a path that never runs).  CPU only."""
import numpy as np
import pytest

from branch_states import CASES, KINDS, pocket_state
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
    ("floor", "update_floor", dict(update_floor=False)),
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


@pytest.fixture(scope="module")
def orc_fma():
    from oracle import Oracle
    return Oracle("f64fma")


@pytest.mark.parametrize("case", CASES, ids=lambda c: "-".join(c[:4]))
def test_gpu_branch_cases_are_well_conditioned(orc, orc_fma, case):
    """Every GPU branch-parity case fires its branches in the oracle, and the oracle built with
    FMA contraction agrees with the uncontracted build to 1e-14 relative L1 after the three steps
    the GPU test takes (psi measured against ch * sum |B|, as there): rounding differences alone
    stay two orders below that test's 1e-12 bar, so a GPU mismatch there is a scheme difference."""
    kind, pred, recon, riem, fires, off = case
    k = KINDS[kind]
    u = pocket_state(orc, NX, NY, NZ, kind)
    sc = scheme_for(orc, u, kind, predictor=pred, recon=recon, riemann=riem)
    orc.branch_hits(reset=True)
    a = run(orc, u, sc)
    hits = orc.branch_hits(reset=True)
    assert all(hits[b] > 0 for b in fires), hits
    b = run(orc_fma, u, sc)
    den = np.abs(a).reshape(9, -1).sum(axis=1)
    floor = np.full(9, 1e-8 * den[0])
    floor[8] = sc.ch * np.sqrt(a[5] ** 2 + a[6] ** 2 + a[7] ** 2).sum()
    err = np.abs(a - b).reshape(9, -1).sum(axis=1) / np.maximum(den, floor)
    assert np.all(err <= 1e-14), (case, err)
    assert k["smallr"] > 0
