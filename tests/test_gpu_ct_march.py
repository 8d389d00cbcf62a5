"""Fused CT z-march (glm_ct_step; spike_ctm.cu:250-298 ct_march, physics ct_mhd.jl:20-184)
against the CPU oracle and against the multi-pass CT step (glm_ct_step_passes).

Bars: FP64 relative L1 <= 1e-12 per variable against the oracle (reduction order and FMA
contraction only); fused vs multi-pass FP64 <= 1e-13 (same point functions, different kernels);
FP32 fused vs multi-pass <= 1e-6.  Face div B stays at round-off."""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle import CTScheme, OUForcing  # noqa: E402
from glmmhdturb_amd import Params, Solver, step_args  # noqa: E402
from test_gpu_parity import TOL64, rel_l1, smooth_state, to_dev, to_np  # noqa: E402


def ct_state(orc, nx, ny, nz, seed=11, b0=0.5):
    """Smooth hydro state with divergence-free face B = b0 x + curl A (A on edges)."""
    u = smooth_state(orc, nx, ny, nz, seed=seed)[:8].copy()
    x = np.arange(nx) / nx; y = np.arange(ny) / ny; z = np.arange(nz) / nz
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    tp = 2 * np.pi
    ax = 0.05 * np.sin(tp * (Y + Z)); ay = 0.05 * np.cos(tp * (X + 2 * Z)); az = 0.05 * np.sin(tp * (X - Y))
    dx = 1.0 / nx
    rl = lambda a, ax_: np.roll(a, -1, axis=ax_)  # noqa: E731  (index i+1)
    u[5] = b0 + ((rl(az, 1) - az) - (rl(ay, 0) - ay)) / dx
    u[6] = ((rl(ax, 0) - ax) - (rl(az, 2) - az)) / dx
    u[7] = ((rl(ay, 2) - ay) - (rl(ax, 1) - ax)) / dx
    return u


def ct_state_np(n, gamma=5.0 / 3.0, b0=0.5):
    """The same kind of state built with array arithmetic (for large grids): smooth rho, v, p and
    divergence-free face B; E from the cell-centred B."""
    x = (np.arange(n) + 0.5) / n
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    tp = 2 * np.pi
    rho = 1.0 + 0.15 * np.sin(tp * X) * np.cos(tp * Y) + 0.05 * np.cos(tp * Z)
    v = [0.3 * np.sin(tp * (Y + 0.2)), 0.3 * np.cos(tp * (Z + X)), 0.2 * np.sin(tp * (X - Z))]
    pr = 1.0 + 0.1 * np.cos(tp * (X + Y + Z))
    xe = np.arange(n) / n
    Ze, Ye, Xe = np.meshgrid(xe, xe, xe, indexing="ij")
    ax = 0.05 * np.sin(tp * (Ye + Ze)); ay = 0.05 * np.cos(tp * (Xe + 2 * Ze)); az = 0.05 * np.sin(tp * (Xe - Ye))
    rl = lambda a, ax_: np.roll(a, -1, axis=ax_)  # noqa: E731
    dx = 1.0 / n
    bx = b0 + ((rl(az, 1) - az) - (rl(ay, 0) - ay)) / dx
    by = ((rl(ax, 0) - ax) - (rl(az, 2) - az)) / dx
    bz = ((rl(ay, 2) - ay) - (rl(ax, 1) - ax)) / dx
    bc = [0.5 * (bx + rl(bx, 2)), 0.5 * (by + rl(by, 1)), 0.5 * (bz + rl(bz, 0))]
    e = pr / (gamma - 1.0) + 0.5 * rho * sum(c * c for c in v) + 0.5 * sum(c * c for c in bc)
    return np.stack([rho, rho * v[0], rho * v[1], rho * v[2], e, bx, by, bz])


def ct_args(p, dt, riem, recon, **kw):
    a = step_args(p, dt, kw.pop("t", 0.0), riemann=riem, recon=recon, want_cfl=False, **kw)
    a.pfl = 1e-6
    return a


@pytest.mark.parametrize("riem,recon,fs,shape", [
    ("hlld", "plm", 0, (24, 20, 16)), ("hll", "plm", 1, (24, 20, 16)), ("llf", "ppm_spike", 0, (24, 20, 16)),
    ("hlld", "ppm_spike", 1, (20, 13, 11)), ("hll", "plm", 0, (17, 30, 9))])
def test_ct_march_vs_oracle(orc, riem, recon, fs, shape):
    """Three steps of the fused march on grids with partial tiles and short z-chunks."""
    nx, ny, nz = shape
    p = Params(N=nx, nx=nx, ny=ny, nz=nz, gamma=5.0 / 3.0)
    s = Solver(p, "f64")
    u = ct_state(orc, nx, ny, nz)
    dx = 1.0 / nx
    assert orc.ct_divb_max(u, dx) < 1e-9
    dt = 0.3 * dx / orc.ct_ctot_max(u, p.gamma, 1e-6, 1e-6)
    a = ct_args(p, dt, riem, recon)
    a.update_floor = 0
    cs = CTScheme(gamma=p.gamma, smallr=1e-6, smallp=1e-6, dt=dt, dx=dx, face_slope=fs, riemann=riem, recon=recon)
    qd, qn = to_dev(s, u), s.new_state(8)
    qo = u.copy()
    for it in range(3):
        s.ct_step(qd, qn, a, fs)
        qd, qn = qn, qd
        qo = orc.ct_step(qo, cs)
    err = rel_l1(to_np(qd), qo, 8)
    assert np.all(err <= TOL64), err
    assert s.ct_divb_max(qd, dx) <= 1e-9
    s.close()


def test_ct_march_forced_floor_vs_oracle(orc):
    """OU kick and the density floor inside the fused update, against the oracle's CT step."""
    N = 24
    p = Params(N=N, gamma=5.0 / 3.0, courant=0.3, turb_gs=16, turb_kmin=1.0, turb_kmax=3.0, comp_frac=0.0,
               turb_T=0.5, turb_Ndt=100, turb_seed=5)
    s = Solver(p, "f64")
    f = s.setup_forcing(p, amp=50.0)
    fo = OUForcing(orc, g=16, kmin=1.0, kmax=3.0, T=0.5, ndt=100, comp_frac=0.0, amp=50.0, seed=5)
    u = ct_state(orc, N, N, N, seed=4)
    # a near-vacuum block at rest: its inner cells fall below smallr and are floored
    blk = (slice(None), slice(8, 14), slice(6, 12), slice(4, 10))
    u[0][blk[1:]] = 1e-9
    u[1][blk[1:]] = 0.0; u[2][blk[1:]] = 0.0; u[3][blk[1:]] = 0.0
    u[4][blk[1:]] = 1e-3 / (p.gamma - 1.0) + 0.5 * (u[5][blk[1:]] ** 2 + u[6][blk[1:]] ** 2 + u[7][blk[1:]] ** 2)
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
        s.ct_step(qd, qn, a, 0)
        qd, qn = qn, qd
        qo = orc.ct_step(qo, cs, afield=fo.afield)
        t += dt
    err = rel_l1(to_np(qd), qo, 8)
    assert np.all(err <= TOL64), err
    s.close()


@pytest.mark.parametrize("prec,tol", [("f64", 1e-13), ("f32", 1e-6)])
@pytest.mark.parametrize("riem,recon", [("hlld", "plm"), ("hll", "ppm_spike"), ("llf", "plm")])
def test_ct_march_matches_passes(orc, prec, tol, riem, recon):
    """Fused march and multi-pass kernels share the point functions: same results to rounding,
    over five steps of a 40 x 36 x 30 grid with forcing."""
    nx, ny, nz = 40, 36, 30
    p = Params(N=nx, nx=nx, ny=ny, nz=nz, gamma=5.0 / 3.0, turb_gs=16, turb_kmin=1.0, turb_kmax=3.0,
               comp_frac=0.0, turb_T=0.5, turb_seed=9)
    s = Solver(p, prec)
    f = s.setup_forcing(p, amp=30.0)
    f.advance()
    u = ct_state(orc, nx, ny, nz, seed=8)
    dx = 1.0 / nx
    dt = 0.3 * dx / orc.ct_ctot_max(u, p.gamma, 1e-6, 1e-6)
    out = []
    for passes in (False, True):
        qd, qn = to_dev(s, u), s.new_state(8)
        for it in range(5):
            a = ct_args(p, dt, riem, recon, do_turb=True, t=0.05 + it * dt)
            s.ct_step(qd, qn, a, 0, passes=passes)
            qd, qn = qn, qd
        out.append(to_np(qd))
    err = rel_l1(out[0], out[1], 8)
    assert np.all(err <= tol), err
    s.close()


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_ct_march_zchunk_invariance_and_tuner(orc, prec):
    """Each z-chunk recomputes the faces and EMFs at its boundary, so every chunking gives the same
    bits; the z-chunk tuner keeps one of them and changes nothing in the result."""
    nx, ny, nz = 24, 20, 40
    p = Params(N=nx, nx=nx, ny=ny, nz=nz, gamma=5.0 / 3.0)
    s = Solver(p, prec)
    u = ct_state(orc, nx, ny, nz, seed=6)
    dx = 1.0 / nx
    a = ct_args(p, 0.3 * dx / orc.ct_ctot_max(u, p.gamma, 1e-6, 1e-6), "hlld", "plm")
    qd = to_dev(s, u)
    outs = []
    for zc in (0, 1, 3, 7, 16, 40):
        b = ct_args(p, a.dt, "hlld", "plm", zchunk=zc)
        qn = s.new_state(8)
        s.ct_step(qd, qn, b, 0)
        outs.append(qn)
    for o in outs[1:]:
        assert torch.equal(o, outs[0])
    qn = s.new_state(8)
    r = s.tune_zchunk(qd, qn, a, reps=1, ct=True)
    assert s.ct_zchunk == r["zchunk"]
    qn.zero_()
    s.ct_step(qd, qn, a, 0)
    assert torch.equal(qn, outs[0])
    s.close()


def test_ct_march_largest_grid_fp32():
    """1300^3 FP32 CT (two 8-plane states: 141 GB of HBM; 2.2e9 cells per plane, past any 32-bit
    index).  A uniform state stays exactly uniform; from a density/pressure wave in a uniform field
    one fused step conserves mass and energy to half an FP32 ulp per cell and keeps face div B at
    the rounding of the six face values around a cell."""
    N = 1300
    if torch.cuda.get_device_properties(0).total_memory < 200 * 2 ** 30:
        pytest.skip("needs a GPU with >= 200 GiB")
    torch.cuda.empty_cache()
    p = Params(N=N, gamma=5.0 / 3.0)
    s = Solver(p, "f32")
    q, qn = s.new_state(8), s.new_state(8)
    assert q[0].numel() > 2 ** 31
    b = (0.3, 0.2, 0.5)
    q[0].fill_(1.0); q[1:4].zero_()
    q[5].fill_(b[0]); q[6].fill_(b[1]); q[7].fill_(b[2])
    q[4].fill_(1.0 / (p.gamma - 1.0) + 0.5 * sum(c * c for c in b))
    dx = 1.0 / N
    a = ct_args(p, 1e-4, "hlld", "plm")
    s.ct_step(q, qn, a, 0)
    for v in range(8):
        assert torch.equal(q[v], qn[v]), v
    x = (torch.arange(N, device=q.device, dtype=torch.float32) + 0.5) / N
    w = torch.sin(2 * math.pi * x)
    q[0].copy_((1.0 + 0.2 * w.view(N, 1, 1) * w.view(1, 1, N)).expand(N, N, N))
    q[4].copy_(((1.0 + 0.1 * w.view(1, N, 1)) / (p.gamma - 1.0) + 0.5 * sum(c * c for c in b)).expand(N, N, N))
    tot0 = [float(q[v].sum(dtype=torch.float64)) for v in (0, 4)]
    dt = 0.3 * dx / s.ct_ctot_max(q)
    s.ct_step(q, qn, ct_args(p, dt, "hlld", "plm"), 0)
    tot1 = [float(qn[v].sum(dtype=torch.float64)) for v in (0, 4)]
    for t0, t1 in zip(tot0, tot1):
        assert abs(t1 - t0) <= 3e-8 * abs(t0), (t0, t1)
    assert s.ct_divb_max(qn, dx) <= 6 * 2 ** -24 * max(b) * 1.05 / dx
    assert bool(torch.isfinite(qn[0]).all())
    s.close()
    del q, qn
    torch.cuda.empty_cache()


def test_ct_march_cfl_and_divb_large():
    """128^3 FP64: the fused step keeps face div B at round-off over 10 steps, and its fused
    CFL read-back equals the standalone CT CFL of the result."""
    N = 128
    p = Params(N=N, gamma=5.0 / 3.0)
    s = Solver(p, "f64")
    u = ct_state_np(N, p.gamma)
    dx = 1.0 / N
    qd, qn = to_dev(s, u), s.new_state(8)
    d0 = s.ct_divb_max(qd, dx)
    for it in range(10):
        dt = 0.3 * dx / s.ct_ctot_max(qd)
        a = ct_args(p, dt, "hlld", "plm")
        a.want_cfl = 1
        s.ct_step(qd, qn, a, 0)
        qd, qn = qn, qd
        assert abs(s.cfl_max() - s.ct_ctot_max(qd)) <= 1e-14 * s.ct_ctot_max(qd)
    assert s.ct_divb_max(qd, dx) <= max(10 * d0, 1e-10)
    assert bool(torch.isfinite(qd).all())
    s.close()
