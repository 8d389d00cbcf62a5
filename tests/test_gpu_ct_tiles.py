"""The CT step over staged 3-D tiles (glm_ct_step_tiles; the design of spike_ct2.cu:147-210 and
spike_ct3.cu:148-190, the physics of ct_mhd.jl:20-184) against the CPU oracle, against the
multi-pass CT step (glm_ct_step_passes, the same point functions) and against the fused CT march.

Bars: FP64 relative L1 <= 1e-12 per variable against the oracle; tiles vs passes and vs the march
<= 1e-13 in FP64 and 1e-6 in FP32.  Face div B stays at round-off across tile seams."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle import CTScheme, OUForcing  # noqa: E402
from test_gpu_ct_march import ct_args, ct_state, ct_state_np  # noqa: E402
from test_gpu_parity import TOL64, rel_l1, to_dev, to_np  # noqa: E402

from glmmhdturb_amd import Params, Solver  # noqa: E402


@pytest.mark.parametrize("riem,recon,fs,shape", [
    ("hlld", "plm", 0, (24, 20, 16)), ("hll", "plm", 1, (24, 20, 16)), ("llf", "ppm_spike", 0, (24, 20, 16)),
    ("hlld", "ppm_spike", 1, (20, 13, 11)), ("hll", "plm", 0, (17, 30, 9)), ("hlld", "plm", 0, (32, 32, 1))])
def test_ct_tiles_vs_oracle(orc, riem, recon, fs, shape):
    """Three steps on grids with partial tiles on every axis (and a one-cell-thick grid)."""
    nx, ny, nz = shape
    print(f"ct tiles {riem} {recon} fs={fs} {shape}", flush=True)
    p = Params(N=nx, nx=nx, ny=ny, nz=nz, gamma=5.0 / 3.0)
    s = Solver(p, "f64")
    u = ct_state(orc, nx, ny, nz)
    dx = 1.0 / nx
    dt = 0.3 * dx / orc.ct_ctot_max(u, p.gamma, 1e-6, 1e-6)
    a = ct_args(p, dt, riem, recon)
    a.update_floor = 0
    cs = CTScheme(gamma=p.gamma, smallr=1e-6, smallp=1e-6, dt=dt, dx=dx, face_slope=fs, riemann=riem, recon=recon)
    qd, qn = to_dev(s, u), s.new_state(8)
    qo = u.copy()
    for it in range(3):
        s.ct_step(qd, qn, a, fs, path="tiles")
        qd, qn = qn, qd
        qo = orc.ct_step(qo, cs)
    err = rel_l1(to_np(qd), qo, 8)
    assert np.all(err <= TOL64), err
    assert s.ct_divb_max(qd, dx) <= 1e-9
    s.close()


def test_ct_tiles_forced_floor_vs_oracle(orc):
    """The OU kick and the density floor of the tiles' update, against the oracle's CT step."""
    N = 24
    p = Params(N=N, gamma=5.0 / 3.0, courant=0.3, turb_gs=16, turb_kmin=1.0, turb_kmax=3.0, comp_frac=0.0,
               turb_T=0.5, turb_Ndt=100, turb_seed=5)
    s = Solver(p, "f64")
    f = s.setup_forcing(p, amp=50.0)
    fo = OUForcing(orc, g=16, kmin=1.0, kmax=3.0, T=0.5, ndt=100, comp_frac=0.0, amp=50.0, seed=5)
    u = ct_state(orc, N, N, N, seed=4)
    blk = (slice(8, 14), slice(6, 12), slice(4, 10))
    u[0][blk] = 1e-9
    u[1][blk] = 0.0; u[2][blk] = 0.0; u[3][blk] = 0.0
    u[4][blk] = 1e-3 / (p.gamma - 1.0) + 0.5 * (u[5][blk] ** 2 + u[6][blk] ** 2 + u[7][blk] ** 2)
    dx = 1.0 / N
    qd, qn = to_dev(s, u), s.new_state(8)
    qo = u.copy()
    t = 0.0
    for it in range(4):
        while t >= f.next_time:
            f.advance(); fo.advance()
        dt = min(0.3 * dx / orc.ct_ctot_max(qo, p.gamma, 1e-6, 1e-6), f.next_time - t)
        a = ct_args(p, dt, "hlld", "plm", do_turb=True, t=t + 0.05)
        a.update_floor = 1
        cs = CTScheme(gamma=p.gamma, smallr=a.smallr, smallp=1e-6, dt=dt, dx=dx, smallc=a.smallc,
                      turb_min_rho=a.turb_min_rho, ramp=a.ramp, riemann="hlld", recon="plm", do_turb=True,
                      turb_gs=16, update_floor=True)
        s.ct_step(qd, qn, a, 0, path="tiles")
        qd, qn = qn, qd
        qo = orc.ct_step(qo, cs, afield=fo.afield)
        t += dt
    err = rel_l1(to_np(qd), qo, 8)
    assert np.all(err <= TOL64), err
    s.close()


@pytest.mark.parametrize("prec,tol", [("f64", 1e-13), ("f32", 1e-6)])
@pytest.mark.parametrize("riem,recon", [("hlld", "plm"), ("hll", "ppm_spike"), ("llf", "plm")])
def test_ct_tiles_match_passes_and_march(orc, prec, tol, riem, recon):
    """Five forced steps of a 40 x 36 x 30 grid: the tiles, the passes and the march agree."""
    nx, ny, nz = 40, 36, 30
    p = Params(N=nx, nx=nx, ny=ny, nz=nz, gamma=5.0 / 3.0, turb_gs=16, turb_kmin=1.0, turb_kmax=3.0,
               comp_frac=0.0, turb_T=0.5, turb_seed=9)
    s = Solver(p, prec)
    f = s.setup_forcing(p, amp=30.0)
    f.advance()
    u = ct_state(orc, nx, ny, nz, seed=8)
    dx = 1.0 / nx
    dt = 0.3 * dx / orc.ct_ctot_max(u, p.gamma, 1e-6, 1e-6)
    out = {}
    for path in ("tiles", "passes", "march"):
        qd, qn = to_dev(s, u), s.new_state(8)
        for it in range(5):
            a = ct_args(p, dt, riem, recon, do_turb=True, t=0.05 + it * dt)
            s.ct_step(qd, qn, a, 0, path=path)
            qd, qn = qn, qd
        out[path] = to_np(qd)
    for other in ("passes", "march"):
        err = rel_l1(out["tiles"], out[other], 8)
        assert np.all(err <= tol), (other, err)
    s.close()


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_ct_tiles_divb_and_cfl(prec):
    """96^3: face div B stays at round-off over 10 tiled steps (every tile seam carries one EMF),
    and the CFL read-back after a step equals the standalone CT CFL of the result."""
    N = 96
    p = Params(N=N, gamma=5.0 / 3.0)
    s = Solver(p, prec)
    u = ct_state_np(N, p.gamma)
    dx = 1.0 / N
    qd, qn = to_dev(s, u), s.new_state(8)
    d0 = s.ct_divb_max(qd, dx)
    for it in range(10):
        dt = 0.3 * dx / s.ct_ctot_max(qd)
        a = ct_args(p, dt, "hlld", "plm")
        a.want_cfl = 1
        s.ct_step(qd, qn, a, 0, path="tiles")
        qd, qn = qn, qd
        c = s.ct_ctot_max(qd)
        assert abs(s.cfl_max() - c) <= 1e-14 * c
    lim = max(10 * d0, 1e-10) if prec == "f64" else 1e-5 / dx       # FP32: |B| <~ 1, ~170 ulp of it per dx
    assert s.ct_divb_max(qd, dx) <= lim
    assert bool(torch.isfinite(qd).all())
    s.close()
