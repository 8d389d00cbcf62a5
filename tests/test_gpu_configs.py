"""GPU runs of the benchmark configurations that are correctness cases, against the oracle or
against single-box runs.

* Orszag-Tang 128 x 128 x 1, FP64: full Hancock + PLM minmod + HLL + GLM, CFL 0.4, to t = 0.5,
  with the Dedner speed ch = the grid's max fast magnetosonic speed and cp^2/ch = 0.18
  (glmmhd_turb.jl:2041-2068 run_ot with the cleaning-speed rule of the Orszag-Tang benchmark).
  The GPU driver (turb.run_ot) is compared with the oracle stepped by the same driver logic.
* The 64-box ensemble of uneven sizes drawn from {32, 48, 64, 96, 128}^3 by seed 2024: three
  sweeps in one launch each, every box bit-identical to its own single-box run.
* OU-forced turbulence from a moving, magnetised state (not from rest), compared with the plain
  relative L1 (no scale floor on any variable).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle import OUForcing, Scheme  # noqa: E402
from test_gpu_parity import TOL64, oracle_scheme, smooth_state, to_dev, to_np  # noqa: E402

from glmmhdturb_amd import Ensemble, Params, Solver, pfloor, random_boxes, single_box_run, step_args  # noqa: E402
from glmmhdturb_amd.turb import dxof, run_ot  # noqa: E402


def plain_rel_l1(a, b):
    """sum |a_v - b_v| / sum |b_v| per variable; a variable that is zero in b (mz and Bz of a 2-D
    state) counts as 0 if it is zero in a too and as inf otherwise."""
    a = np.asarray(a, np.float64).reshape(a.shape[0], -1)
    b = np.asarray(b, np.float64).reshape(b.shape[0], -1)
    num, den = np.abs(a - b).sum(axis=1), np.abs(b).sum(axis=1)
    return np.where(den > 0, num / np.where(den > 0, den, 1.0), np.where(num > 0, np.inf, 0.0))


def test_orszag_tang_128_minmod_hll_maxfast_to_t05(orc):
    N, t_end = 128, 0.5
    d, ug, s = run_ot(N, t_end, nz=1, riemann="hll", recon="plm", predictor="full", slope="minmod",
                      scheme="cube", courant=0.4, verbose=False, ch_mode="maxfast")
    got = to_np(ug)
    s.close()

    p = Params(N=N, gamma=5.0 / 3.0, cs0=1.0, courant=0.4, nz=1, glm_ch_mode="maxfast")
    pfl = pfloor(p)
    uo = orc.ic_ot(N, N, 1, p.gamma)
    t, steps = 0.0, 0
    while t < t_end:
        dt = min(p.courant * dxof(p) / orc.ctot_max(uo, p.gamma, p.smallr, pfl), t_end - t)
        ch = orc.fast_max(uo, p.gamma, p.smallr, pfl)
        a = step_args(p, dt, t, riemann="hll", recon="plm", predictor="full", slope="minmod", scheme="cube",
                      want_cfl=False, ch=ch)
        uo = orc.step(uo, oracle_scheme(a))
        t += dt
        steps += 1
    assert steps > 300 and abs(t - t_end) < 1e-14
    err = plain_rel_l1(got, uo)
    assert np.all(err[:8] <= TOL64), err
    # psi starts at zero and stays at the level of the scheme's div B: measure it against ch |B|
    babs = np.sqrt(uo[5] ** 2 + uo[6] ** 2 + uo[7] ** 2).sum()
    assert np.abs(got[8] - uo[8]).sum() <= TOL64 * ch * babs
    dref = orc.diagnostics(uo, 1.0 / N, p.smallr)
    assert d["vrms"] == pytest.approx((dref["ke2"] / dref["mass"]) ** 0.5, rel=1e-12)
    assert d["finite"]


def test_ensemble_64_boxes_seed_2024_bit_identical():
    scheme = Params(gamma=5.0 / 3.0)
    boxes = random_boxes(64, seed=2024)
    sizes = sorted({b.N for b in boxes})
    assert sizes == [32, 48, 64, 96, 128]
    nsw = 3
    ens = Ensemble(boxes, scheme, "f64")
    ens.step(nsw)
    ens.sync()
    for b, spec in enumerate(ens.boxes):
        log = ens.log(b, nsw)
        s, u, rec = single_box_run(spec, scheme, nsw, "f64", log=log)
        for k in range(nsw):
            assert log[k]["dt"] == rec[k]["dt"] and log[k]["cmax"] == rec[k]["cmax"], (b, k)
        assert torch.equal(ens.current(b), u), f"box {b} ({spec.N}^3) differs from its single-box run"
        s.close()
    ens.close()


def test_forced_turbulence_parity_moving_state(orc):
    """OU kick on a moving state (rms v 0.3, B ~ 0.5): 10 forced PPM + HLLD steps, every variable
    within plain relative L1 1e-12 of the oracle, and the kick's effect on momentum well above
    that bar (so the comparison sees the forcing)."""
    N = 24
    p = Params(N=N, gamma=5.0 / 3.0, courant=0.3, turb_gs=16, turb_kmin=1.0, turb_kmax=3.0, comp_frac=0.0,
               turb_T=0.5, turb_Ndt=100, turb_seed=5, b0=0.5)
    s = Solver(p, "f64")
    f = s.setup_forcing(p, amp=200.0)
    fo = OUForcing(orc, g=16, kmin=1.0, kmax=3.0, T=0.5, ndt=100, comp_frac=0.0, amp=200.0, seed=5)
    u0 = smooth_state(orc, N, N, N, seed=9)
    uo, unf = u0.copy(), u0.copy()
    ud, un = to_dev(s, u0), s.new_state()
    t = 0.0
    for it in range(10):
        while t >= f.next_time:
            f.advance(); fo.advance()
        dt = min(p.courant * (1.0 / N) / orc.ctot_max(uo, p.gamma, p.smallr, pfloor(p)), f.next_time - t)
        a = step_args(p, dt, t + 0.2, do_turb=True, riemann="hlld", recon="ppm", want_cfl=False)
        s.step(ud, un, a)
        ud, un = un, ud
        uo = orc.step(uo, oracle_scheme(a, turb_gs=16), afield=fo.afield)
        unf = orc.step(unf, Scheme(**{**oracle_scheme(a, turb_gs=16).__dict__, "do_turb": False}))
        t += dt
    err = plain_rel_l1(to_np(ud), uo)
    assert np.all(err <= TOL64), err
    kick = plain_rel_l1(uo, unf)
    assert np.min(kick[1:4]) > 1e-6, kick
    s.close()
