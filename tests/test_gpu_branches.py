"""GPU parity on the scheme's guarded branches: the FP64 march against the oracle on states that
make each branch fire (tests/branch_states.py).

* ``void`` pockets: the predictor revert mh -> m0 (glmmhd_turb.jl:393-395), the iflo face revert
  (:618) and, with llf_dmin = 0.01 / llf_pmin = 1e-4, the HLLD -> LLF switch (:287-290).
* ``degenerate`` pockets: the HLLD star-state guard |denK| < 1e-12 (:227-236).
* ``floor`` pockets: the update's void floor (:448-451) and both reverts.

Each case checks three things: the oracle's branch counters show the branch fired on this very
run, the GPU result equals the oracle's within relative L1 1e-12 per conserved variable, and the
GPU result is far from the oracle run with that branch switched off (so the match is a match of
the branch, not of a path that never ran).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from branch_states import CASES, KINDS, pocket_state  # noqa: E402
from oracle import Scheme  # noqa: E402
from test_gpu_parity import TOL64, oracle_scheme, rel_l1, to_dev, to_np  # noqa: E402

from glmmhdturb_amd import Params, Solver, pfloor, step_args  # noqa: E402

G = 5.0 / 3.0
NX, NY, NZ = 24, 20, 16
STEPS = 3


def params_for(kind):
    k = KINDS[kind]
    return Params(N=NX, nx=NX, ny=NY, nz=NZ, gamma=G, courant=0.4, smallr=k["smallr"], smallc=k["smallc"],
                  switch_llf_dmin=k["llf_dmin"], switch_llf_pmin=k["llf_pmin"])


@pytest.mark.parametrize("kind,pred,recon,riem,fires,off", CASES)
def test_branch_parity_fp64(orc, kind, pred, recon, riem, fires, off):
    p = params_for(kind)
    s = Solver(p, "f64")
    u0 = pocket_state(orc, NX, NY, NZ, kind, gamma=G)
    dt = p.courant * (1.0 / NX) / orc.ctot_max(u0, G, p.smallr, pfloor(p))
    a = step_args(p, dt, 0.0, riemann=riem, recon=recon, predictor=pred, scheme="cube", want_cfl=False)
    sc = oracle_scheme(a)
    assert sc.llf_dmin == KINDS[kind]["llf_dmin"] and sc.mh_revert and sc.face_guard == "iflo" and sc.update_floor
    print(f"branch case {kind} {pred} {recon} {riem}", flush=True)

    ud, un = to_dev(s, u0), s.new_state()
    uo = u0.copy()
    orc.branch_hits(reset=True)
    for _ in range(STEPS):
        s.step(ud, un, a)
        ud, un = un, ud
        uo = orc.step(uo, sc)
    hits = orc.branch_hits(reset=True)
    for b in fires:
        assert hits[b] > 0, (b, hits)
    got = to_np(ud)
    assert np.all(np.isfinite(got))
    err = rel_l1(got, uo, ch=a.ch)
    assert np.all(err <= TOL64), (kind, pred, recon, riem, err)

    uoff = u0.copy()
    sc_off = Scheme(**{**sc.__dict__, **off})
    for _ in range(STEPS):
        uoff = orc.step(uoff, sc_off)
    assert np.max(rel_l1(got, uoff, ch=a.ch)) > 1e3 * TOL64, off
    s.close()


def test_branch_parity_fp64_zchunks(orc):
    """The void pockets through z-chunks of 1, 3 and 5 planes (the lead-in planes of every chunk
    cross the pockets) give the same bits as the unchunked march."""
    p = params_for("void")
    s = Solver(p, "f64")
    u0 = pocket_state(orc, NX, NY, NZ, "void", gamma=G)
    dt = p.courant * (1.0 / NX) / orc.ctot_max(u0, G, p.smallr, pfloor(p))
    outs = []
    for zc in (0, 1, 3, 5):
        a = step_args(p, dt, 0.0, riemann="hlld", recon="ppm", want_cfl=False, zchunk=zc)
        ud, un = to_dev(s, u0), s.new_state()
        s.step(ud, un, a)
        outs.append(to_np(un))
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
    s.close()
