"""GPU parity: the gfx950 kernels (through the C ABI) against the CPU oracle.

Bar (FP64): relative L1 <= 1e-12 per conserved variable after N steps, with only reduction-order
and FMA-contraction differences allowed.  FP32 paths are compared with the FP32 oracle at an
FP32 tolerance.
"""
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle import CTScheme, OUForcing, Scheme  # noqa: E402

from glmmhdturb_amd import Params, Solver, _lib, glm_speed, pfloor, step_args  # noqa: E402

TOL64 = 1e-12


def rel_l1(a, b, nv=9, ch=None):
    """Per-variable relative L1: sum|a_v - b_v| / max(sum|b_v|, floor_v).

    The floor only matters for a variable that is small compared with the fluxes that update
    it (momentum or transverse B generated from rest by the forcing, psi generated from
    round-off-level div B): its error is then measured against the scale the scheme works at,
    not against its own near-zero size.  Floors: momenta sum sqrt(rho |E|); field components
    sum |B|; psi ch * sum |B| (psi has units of ch*B); anything else 1e-8 * sum rho.
    For O(1) fields (every test state that starts away from rest) this is the plain relative L1."""
    a = np.asarray(a, dtype=np.float64).reshape(nv, -1)
    b = np.asarray(b, dtype=np.float64).reshape(nv, -1)
    floor = np.full(nv, 1e-8 * np.abs(b[0]).sum())
    babs = np.sqrt(b[5] ** 2 + b[6] ** 2 + b[7] ** 2).sum()
    floor[1:4] = np.maximum(floor[1:4], np.sqrt(np.abs(b[0] * b[4])).sum())
    floor[5:8] = np.maximum(floor[5:8], babs)
    if nv == 9 and ch is not None:
        floor[8] = max(floor[8], ch * babs)
    return np.array([np.abs(a[v] - b[v]).sum() / max(np.abs(b[v]).sum(), floor[v]) for v in range(nv)])


def smooth_state(orc, nx, ny, nz, seed=3, amp=0.3, bamp=0.4, gamma=5.0 / 3.0):
    """A smooth, fully 3-D periodic state with all nine variables active."""
    rng = np.random.default_rng(seed)
    x = (np.arange(nx) + 0.5) / nx
    y = (np.arange(ny) + 0.5) / ny
    z = (np.arange(nz) + 0.5) / nz
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    tp = 2 * np.pi

    def f(k):
        ph = rng.uniform(0, tp, 3)
        return np.sin(tp * X + ph[0]) * np.cos(tp * Y + ph[1]) + 0.5 * np.sin(tp * (Y + Z) + ph[2]) + 0.3 * np.cos(tp * k * Z)

    rho = 1.0 + amp * 0.5 * f(1)
    vx, vy, vz = amp * f(1), amp * f(2), amp * f(1)
    p = 1.0 + amp * 0.3 * f(2)
    bx, by, bz = 0.5 + bamp * f(1), bamp * f(2), bamp * f(1)
    psi = 0.01 * f(2)
    q = np.stack([rho, vx, vy, vz, p, bx, by, bz, psi])
    u = np.empty_like(q)
    for idx in np.ndindex(nz, ny, nx):
        u[(slice(None),) + idx] = orc.prim2cons(q[(slice(None),) + idx], gamma)
    return u


def to_dev(s, u):
    t = s.new_state(u.shape[0])
    t.copy_(torch.from_numpy(np.ascontiguousarray(u)).to(t.device, t.dtype))
    return t


def to_np(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy().astype(np.float64)


def oracle_scheme(a, **kw):
    """Oracle scheme from the C parameter block the GPU step receives."""
    inv = lambda d, v: [k for k, x in d.items() if x == v][0]  # noqa: E731
    return Scheme(gamma=a.gamma, smallr=a.smallr, smallc=a.smallc, pfl=a.pfl, dt=a.dt, dx=a.dx, ch=a.ch,
                  glm_fac=a.glm_fac, llf_dmin=a.llf_dmin, llf_pmin=a.llf_pmin, turb_min_rho=a.turb_min_rho,
                  ramp=a.ramp, predictor=inv(_lib.PREDICTOR, a.predictor), recon=inv(_lib.RECON, a.recon),
                  slope=inv(_lib.SLOPE, a.slope), riemann=inv(_lib.RIEMANN, a.riemann),
                  mh_revert=bool(a.mh_revert), face_guard=inv(_lib.GUARD, a.face_guard),
                  update_floor=bool(a.update_floor), do_turb=bool(a.do_turb), **kw)


@pytest.fixture(scope="module")
def grid3d(orc):
    nx, ny, nz = 24, 20, 16          # not multiples of the 16 x 12 tile: partial tiles on both axes
    p = Params(N=nx, nx=nx, ny=ny, nz=nz, gamma=5.0 / 3.0, courant=0.4)
    s = Solver(p, "f64")
    u0 = smooth_state(orc, nx, ny, nz)
    return p, s, u0


COMBOS = [(pr, rc, rm, sl) for pr in ("full", "oned") for rc in ("plm", "ppm", "ppm_spike")
          for rm in ("llf", "llf_bound", "hll", "hlld") for sl in ("moncen", "minmod")]


@pytest.mark.parametrize("pred,recon,riem,slope", COMBOS)
def test_step_matrix_fp64(orc, grid3d, pred, recon, riem, slope):
    p, s, u0 = grid3d
    dt = 0.3 * (1.0 / p.N) / orc.ctot_max(u0, p.gamma, p.smallr, pfloor(p))
    uo = u0.copy()
    ud = to_dev(s, u0)
    un = s.new_state()
    for it in range(3):
        a = step_args(p, dt, 0.0, riemann=riem, recon=recon, predictor=pred, slope=slope, scheme="cube", want_cfl=False)
        s.step(ud, un, a)
        ud, un = un, ud
        uo = orc.step(uo, oracle_scheme(a))
    err = rel_l1(to_np(ud), uo, ch=a.ch)
    assert np.all(err <= TOL64), (pred, recon, riem, slope, err)


@pytest.mark.parametrize("scheme", ["multi", "cpu_simd", "spike"])
def test_scheme_presets_fp64(orc, grid3d, scheme):
    p, s, u0 = grid3d
    dt = 0.3 * (1.0 / p.N) / orc.ctot_max(u0, p.gamma, p.smallr, pfloor(p))
    a = step_args(p, dt, 0.0, scheme=scheme, riemann="hlld", recon="ppm", want_cfl=False)
    ud, un = to_dev(s, u0), s.new_state()
    uo = u0.copy()
    for it in range(3):
        s.step(ud, un, a)
        ud, un = un, ud
        uo = orc.step(uo, oracle_scheme(a))
    assert np.all(rel_l1(to_np(ud), uo) <= TOL64)


def test_zchunk_invariance(orc, grid3d):
    """z-chunk boundaries recompute the same faces: results are bit-identical for any chunking."""
    p, s, u0 = grid3d
    dt = 1e-3
    outs = []
    for zc in (0, 1, 3, 5, 16):
        a = step_args(p, dt, 0.0, riemann="hlld", recon="ppm", want_cfl=False, zchunk=zc)
        ud, un = to_dev(s, u0), s.new_state()
        s.step(ud, un, a)
        outs.append(to_np(un))
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])


def test_zchunk_autotune_keeps_bits():
    """Solver.tune_zchunk times the march over several z-chunkings, keeps the fastest, and the
    tuned steps give the same bits as the library's own choice."""
    p = Params(N=40, nx=40, ny=30, nz=96, gamma=5.0 / 3.0, courant=0.3)
    s = Solver(p, "f64")
    u, un = s.new_state(), s.new_state()
    s.ic_gradient(u, p.gamma)
    a = step_args(p, 0.3 / 40 / s.ctot_max(u), 0.0, riemann="hlld", recon="ppm", want_cfl=False)
    s.step(u, un, a)
    ref = un.clone()
    r = s.tune_zchunk(u, un, a, reps=1)
    assert s.zchunk == r["zchunk"] and (r["zchunk"] == 0 or 8 <= r["zchunk"] <= 24)
    un.zero_()
    s.step(u, un, a)
    assert torch.equal(un, ref)
    assert a.zchunk == 0                                     # the caller's block is not modified
    s.close()


def test_cfl_and_diagnostics(orc, grid3d):
    p, s, u0 = grid3d
    ud = to_dev(s, u0)
    cm = s.ctot_max(ud)
    co = orc.ctot_max(u0, p.gamma, p.smallr, pfloor(p))
    assert abs(cm - co) <= 1e-14 * co
    d = s.diagnostics(ud)
    do = orc.diagnostics(u0, 1.0 / p.N, p.smallr)
    for k in ("mass", "ke2", "ekin", "emag", "cross_helicity", "etot", "psi2"):
        assert abs(d[k] - do[k]) <= 1e-13 * max(abs(do[k]), 1e-300), k
    assert abs(d["divbmax"] - do["divbmax"]) <= 1e-13 * do["divbmax"]
    assert d["rhomin"] == do["rhomin"] and d["finite"]
    # fused CFL of the step equals the standalone CFL of the result
    a = step_args(p, 1e-3, 0.0, riemann="hlld", recon="ppm", want_cfl=True)
    un = s.new_state()
    s.step(ud, un, a)
    fused = s.cfl_max()
    assert abs(fused - s.ctot_max(un)) <= 1e-14 * fused


def test_initial_conditions(orc):
    p = Params(N=32, nz=4, gamma=5.0 / 3.0)
    s = Solver(p, "f64")
    u = s.new_state()
    s.ic_ot(u, p.gamma)
    assert np.max(np.abs(to_np(u) - orc.ic_ot(32, 32, 4, p.gamma))) <= 1e-15
    s.ic_uniform(u, 1.0, 0.6, (0.5, 0.1, -0.2), p.gamma)
    assert np.array_equal(to_np(u), orc.ic_uniform(32, 32, 4, 1.0, 0.6, (0.5, 0.1, -0.2), p.gamma))
    s.ic_gradient(u, p.gamma)
    assert np.max(np.abs(to_np(u) - orc.ic_gradient(32, 32, 4, p.gamma))) <= 1e-15


def test_orszag_tang_multistep(orc):
    """Orszag-Tang 64x64x1, HLLD + PPM (cube scheme), 40 steps at the oracle's CFL dt."""
    N = 64
    p = Params(N=N, nz=1, gamma=5.0 / 3.0, courant=0.4)
    s = Solver(p, "f64")
    uo = orc.ic_ot(N, N, 1, p.gamma)
    ud, un = to_dev(s, uo), s.new_state()
    for it in range(40):
        dt = p.courant * (1.0 / N) / orc.ctot_max(uo, p.gamma, p.smallr, pfloor(p))
        a = step_args(p, dt, 0.0, riemann="hlld", recon="ppm", scheme="cube", want_cfl=False)
        s.step(ud, un, a)
        ud, un = un, ud
        uo = orc.step(uo, oracle_scheme(a))
    err = rel_l1(to_np(ud), uo, ch=a.ch)
    assert np.all(err <= TOL64), err


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("n,recon,riem,vrms,divb,rhomin", [
    (64, "plm", "hlld", 0.7887, 2.9, None), (128, "plm", "hlld", 0.7906, 4.5, None),
    (120, "ppm", "hll", 0.7910, 1.16, 0.0835)])
def test_orszag_tang_reference_figures(orc, prec, n, recon, riem, vrms, divb, rhomin):
    """The reference's own validation runs (validate_tiled_ot glmmhd_turb.jl:2278-2295: OT on an n^3
    grid, courant 0.7, t_end 0.2, a CFL dt each step) through the GPU step and its diagnostics.  Its
    printed figures (REPORT.md:24-26, BENCHMARKS.md:178; fp32 runs) hold to their digits in both
    precisions: vrms within half a printed unit plus fp32 noise, max|div B| within 1%, rho_min to its
    digits.  In FP64 the 3-D run also equals the oracle's n x n x 1 run (the state is z-invariant)."""
    from glmmhdturb_amd.turb import run_ot
    from test_oracle_physics import ot_oracle_run
    d, u, s = run_ot(n, 0.2, riemann=riem, recon=recon, scheme="cube", courant=0.7, precision=prec, verbose=False)
    s.close()
    assert abs(d["vrms"] - vrms) <= 6e-5, d["vrms"]
    assert d["divbmax"] == pytest.approx(divb, rel=0.01)
    if rhomin is not None:
        assert abs(d["rhomin"] - rhomin) <= 6e-5, d["rhomin"]
    assert d["finite"]
    if prec == "f64":
        do = ot_oracle_run(orc, n, recon, riem)
        assert d["vrms"] == pytest.approx(do["vrms"], rel=1e-10)
        assert d["rhomin"] == pytest.approx(do["rhomin"], rel=1e-10)
        assert d["divbmax"] == pytest.approx(do["divbmax"], rel=1e-6)


def brio_wu(n, axis, gamma=2.0):
    """Brio & Wu's shock tube (rho, p, B_t) = (1, 1, 1) | (0.125, 0.1, -1), B_n = 0.75, on a periodic
    line of n cells along `axis` (the grid is 1 cell wide across it), as a conserved state."""
    shape = [1, 1, 1]
    shape[2 - axis] = n                       # state order (var, z, y, x)
    left = (np.arange(n) < n // 2).reshape(shape)
    rho = np.where(left, 1.0, 0.125)
    p = np.where(left, 1.0, 0.1)
    bt = np.where(left, 1.0, -1.0)
    u = np.zeros((9, *shape))
    u[0] = rho
    u[5 + axis] = 0.75
    u[5 + (axis + 1) % 3] = bt
    u[4] = p / (gamma - 1.0) + 0.5 * (0.75 ** 2 + bt ** 2)
    return u


@pytest.mark.parametrize("riem,recon", [("hlld", "ppm"), ("hll", "plm"), ("llf", "plm")])
def test_brio_wu_tube_along_each_axis(orc, riem, recon):
    """Brio-Wu tube (shocks, a compound wave, a contact) along x, y and z on 1-cell-wide grids: 80
    steps at the oracle's CFL dt against the oracle, and the three directions against each other
    (the tube along y and z is the x tube with its components cycled)."""
    n, res = 128, []
    for axis in range(3):
        shape = [1, 1, 1]
        shape[axis] = n
        p = Params(N=n, nx=shape[0], ny=shape[1], nz=shape[2], gamma=2.0, courant=0.4)
        s = Solver(p, "f64")
        uo = brio_wu(n, axis)
        ud, un = to_dev(s, uo), s.new_state()
        for it in range(80):
            dt = p.courant * (1.0 / n) / orc.ctot_max(uo, p.gamma, p.smallr, pfloor(p))
            a = step_args(p, dt, 0.0, riemann=riem, recon=recon, want_cfl=False)
            s.step(ud, un, a)
            ud, un = un, ud
            uo = orc.step(uo, oracle_scheme(a))
        ug = to_np(ud)
        err = rel_l1(ug, uo, ch=a.ch)
        assert np.all(err <= TOL64), (axis, err)
        assert ug[0].min() < 0.3 and ug[0].max() > 0.9          # the tube has not smeared out
        perm = [0, 1 + axis, 1 + (axis + 1) % 3, 1 + (axis + 2) % 3, 4,
                5 + axis, 5 + (axis + 1) % 3, 5 + (axis + 2) % 3, 8]
        res.append(ug[perm].reshape(9, -1))                     # back to the x tube's components
        s.close()
    for axis in (1, 2):
        err = rel_l1(res[axis], res[0], ch=a.ch)
        assert np.all(err <= 1e-13), (axis, err)


def test_forced_turbulence_parity(orc):
    """OU-forced turbulence 24^3: forcing field and 12 steps (PPM + HLLD, kick on)."""
    N = 24
    p = Params(N=N, gamma=5.0 / 3.0, courant=0.3, turb_gs=16, turb_kmin=1.0, turb_kmax=3.0, comp_frac=0.0,
               turb_T=0.5, turb_Ndt=100, turb_seed=7, b0=0.5)
    s = Solver(p, "f64")
    f = s.setup_forcing(p, amp=50.0)
    fo = OUForcing(orc, g=16, kmin=1.0, kmax=3.0, T=0.5, ndt=100, comp_frac=0.0, amp=50.0, seed=7)
    uo = orc.ic_uniform(N, N, N, 1.0, 0.6, (0.5, 0, 0), p.gamma)
    ud, un = to_dev(s, uo), s.new_state()
    t = 0.0
    for it in range(12):
        while t >= f.next_time:
            f.advance(); fo.advance()
        afd = s.afield().cpu().numpy().reshape(-1)
        assert np.max(np.abs(afd - fo.afield)) <= 1e-12 * max(np.max(np.abs(fo.afield)), 1e-300)
        dt = p.courant * (1.0 / N) / orc.ctot_max(uo, p.gamma, p.smallr, pfloor(p))
        dt = min(dt, f.next_time - t)
        a = step_args(p, dt, t + 0.05, do_turb=True, riemann="hlld", recon="ppm", want_cfl=False)
        s.step(ud, un, a)
        ud, un = un, ud
        uo = orc.step(uo, oracle_scheme(a, turb_gs=16), afield=fo.afield)
        t += dt
    err = rel_l1(to_np(ud), uo, ch=a.ch)
    assert np.all(err <= TOL64), err


@pytest.mark.parametrize("pred,recon,riem,slope", COMBOS)
def test_fp32_step_vs_fp32_oracle(orc32, grid3d, orc, pred, recon, riem, slope):
    """The FP32 march (FP32 state, faces and fluxes; FP64 pressure recovery) against the oracle
    built in FP32, over the whole scheme matrix: 3 steps within 2e-6 relative L1."""
    p, _, u0 = grid3d
    s = Solver(p, "f32")
    dt = 0.3 * (1.0 / p.N) / orc.ctot_max(u0, p.gamma, p.smallr, pfloor(p))
    a = step_args(p, dt, 0.0, riemann=riem, recon=recon, predictor=pred, slope=slope, scheme="cube", want_cfl=False)
    u32 = u0.astype(np.float32)
    ud, un = to_dev(s, u32), s.new_state()
    uo = u32.copy()
    for it in range(3):
        s.step(ud, un, a)
        ud, un = un, ud
        uo = orc32.step(uo, oracle_scheme(a))
    err = rel_l1(to_np(ud), uo)
    assert np.all(err <= 2e-6), (pred, recon, riem, slope, err)
    s.close()


def test_legacy_march_run_dev(orc32, orc):
    """march_run_dev (float**, spike scheme) against the FP32 oracle with the spike preset."""
    from glmmhdturb_amd import march_bridge as MB
    m = MB.open_lib(_lib.LIB_PATH)
    assert m.NV == 9
    N = 32
    MB.set_grid(m, N, N, N)
    MB.set_gamma(m, 5.0 / 3.0)
    u0 = smooth_state(orc, N, N, N, seed=5).astype(np.float32)
    dtdx, ch, fac = 0.05, 1.2, 0.99
    MB.set_dtdx(m, dtdx)
    MB.set_glm(m, ch, fac)
    q = [torch.from_numpy(u0[v].copy()).cuda() for v in range(9)]
    o = [torch.empty_like(x) for x in q]
    ms = MB.run_(m, q, o, 3)
    assert ms >= 0.0
    sc = Scheme(gamma=5.0 / 3.0, smallr=1e-6, pfl=1e-6, dt=np.float32(dtdx), dx=1.0, ch=np.float32(ch),
                glm_fac=np.float32(fac), predictor="oned", recon="plm", slope="moncen", riemann="hll",
                mh_revert=False, face_guard="clamp", update_floor=False)
    uo = u0.copy()
    for it in range(3):
        uo = orc32.step(uo, sc)
    got = np.stack([x.cpu().numpy() for x in q])
    assert np.all(rel_l1(got, uo) <= 2e-6)


def test_legacy_ct_run_and_march_run_dev_f64(orc32, orc):
    """The other legacy entry points: ct_run (spike_ctm.cu:300-328: 8 float planes, the CT march
    with HLL + PLM) against the FP32 CT oracle, march_run_dev_f64 (9 double planes, the spike
    scheme) against the FP64 oracle, and the ct_regs / ct_shmem figures of the CT march."""
    import ctypes as C
    from test_gpu_ct_march import ct_state
    lib = _lib.load()
    N = 24
    g32, dtdx32 = float(np.float32(5.0 / 3.0)), float(np.float32(0.02))
    assert lib.march_set_grid(N, N, N) == 0
    lib.march_set_gamma(C.c_float(g32))
    lib.march_set_dtdx(C.c_float(dtdx32))
    lib.march_set_glm(C.c_float(1.0), C.c_float(1.0))
    # ct_run: 3 steps, result left in q (odd step count: copied back)
    u = ct_state(orc, N, N, N).astype(np.float32)
    q = [torch.from_numpy(u[v].copy()).cuda() for v in range(8)]
    o = [torch.empty_like(x) for x in q]
    cc = [torch.empty_like(x) for x in q[:3]]
    qa, kq = _lib.plane_array([x.data_ptr() for x in q])
    oa, ko = _lib.plane_array([x.data_ptr() for x in o])
    ca, kc = _lib.plane_array([x.data_ptr() for x in cc])
    assert lib.ct_run(qa, oa, ca, 3) >= 0.0
    cs = CTScheme(gamma=g32, smallr=1e-6, smallp=1e-6, dt=dtdx32, dx=1.0, riemann="hll", recon="plm", face_slope=0)
    uo = u.copy()
    for it in range(3):
        uo = orc32.ct_step(uo, cs)
    got = np.stack([x.cpu().numpy() for x in q])
    assert np.all(rel_l1(got, uo, 8) <= 2e-6)
    assert lib.ct_regs() > 0 and lib.ct_shmem() > 0
    # march_run_dev_f64: 2 steps (even: the result is in q without a copy)
    u64 = smooth_state(orc, N, N, N, seed=9)
    q = [torch.from_numpy(u64[v].copy()).cuda() for v in range(9)]
    o = [torch.empty_like(x) for x in q]
    qa, kq = _lib.plane_array([x.data_ptr() for x in q])
    oa, ko = _lib.plane_array([x.data_ptr() for x in o])
    assert lib.march_run_dev_f64(qa, oa, 2) >= 0.0
    sc = Scheme(gamma=g32, smallr=1e-6, pfl=1e-6, dt=dtdx32, dx=1.0, ch=1.0, glm_fac=1.0, predictor="oned",
                recon="plm", slope="moncen", riemann="hll", mh_revert=False, face_guard="clamp", update_floor=False)
    uo = u64.copy()
    for it in range(2):
        uo = orc.step(uo, sc)
    got = np.stack([x.cpu().numpy() for x in q])
    assert np.all(rel_l1(got, uo) <= TOL64)


@pytest.mark.parametrize("riem,recon,fs", [("hll", "plm", 0), ("hll", "plm", 1), ("hlld", "plm", 0),
                                           ("llf", "plm", 0), ("hll", "ppm_spike", 0)])
def test_ct_step_parity(orc, riem, recon, fs):
    nx, ny, nz = 24, 20, 16
    p = Params(N=nx, nx=nx, ny=ny, nz=nz, gamma=5.0 / 3.0)
    s = Solver(p, "f64")
    rng = np.random.default_rng(11)
    u = smooth_state(orc, nx, ny, nz, seed=11)[:8].copy()
    # face B from a vector potential so that the face divergence is zero to round-off
    x = np.arange(nx) / nx; y = np.arange(ny) / ny; z = np.arange(nz) / nz
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    tp = 2 * np.pi
    ax = 0.05 * np.sin(tp * (Y + Z)); ay = 0.05 * np.cos(tp * (X + 2 * Z)); az = 0.05 * np.sin(tp * (X - Y))
    dx = 1.0 / nx
    rl = lambda a, ax_: np.roll(a, -1, axis=ax_)  # noqa: E731  (index i+1)
    # B = curl A with A on edges: bx = dAz/dy - dAy/dz etc. (face-centred, divergence-free)
    u[5] = 0.5 + ((rl(az, 1) - az) - (rl(ay, 0) - ay)) / dx
    u[6] = ((rl(ax, 0) - ax) - (rl(az, 2) - az)) / dx
    u[7] = ((rl(ay, 2) - ay) - (rl(ax, 1) - ax)) / dx
    assert orc.ct_divb_max(u, dx) < 1e-10
    dt = 0.3 * dx / orc.ct_ctot_max(u, p.gamma, 1e-6, 1e-6)
    a = step_args(p, dt, 0.0, riemann=riem, recon=recon, want_cfl=False)
    a.pfl = 1e-6
    a.update_floor = 0
    cs = CTScheme(gamma=p.gamma, smallr=1e-6, smallp=1e-6, dt=dt, dx=dx, face_slope=fs, riemann=riem, recon=recon)
    qd, qn = to_dev(s, u), s.new_state(8)
    qo = u.copy()
    for it in range(3):
        s.ct_step(qd, qn, a, fs)
        qd, qn = qn, qd
        qo = orc.ct_step(qo, cs)
    got = to_np(qd)
    assert np.all(rel_l1(got, qo, 8) <= TOL64)
    assert s.ct_divb_max(qd, dx) <= 1e-10


def test_large_grid_invariants():
    """256^3 FP64 PPM+HLLD: a uniform state is an exact fixed point, and on a smooth state the
    totals of mass, momentum, energy and B are conserved to round-off (no floor or forcing hit)."""
    N = 256
    p = Params(N=N, gamma=5.0 / 3.0, courant=0.3)
    s = Solver(p, "f64")
    u = s.new_state()
    s.ic_uniform(u, 1.0, 0.6, (0.5, 0.0, 0.0), p.gamma)
    un = s.new_state()
    a = step_args(p, 1e-4, 0.0, riemann="hlld", recon="ppm", want_cfl=True)
    s.step(u, un, a)
    assert torch.equal(u[:8], un[:8])
    s.ic_gradient(u, p.gamma)
    tot0 = u.reshape(9, -1).sum(dim=1)
    dt = 0.3 * (1.0 / N) / s.ctot_max(u)
    s.step(u, un, step_args(p, dt, 0.0, riemann="hlld", recon="ppm", want_cfl=False))
    tot1 = un.reshape(9, -1).sum(dim=1)
    scale = u.reshape(9, -1).abs().sum(dim=1)
    for v in range(8):
        assert abs(float(tot1[v] - tot0[v])) <= 1e-12 * float(scale[v]) + 1e-12, v


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_runs_are_bit_reproducible(prec):
    """The step has no floating-point atomics and every face is solved by one lane, so two runs of
    the same forced trajectory (192^3, 30 steps from the gradient state, CFL dt folded into each
    step) give the same bits: a race between lanes or workgroups would show here."""
    from test_gpu_ct_march import ct_args
    N = 192
    p = Params(N=N, gamma=5.0 / 3.0, courant=0.3, turb_gs=64, turb_kmin=1.0, turb_kmax=3.0, comp_frac=0.0,
               turb_T=0.5, turb_Ndt=100, turb_seed=3)
    out = []
    for run in range(2):
        s = Solver(p, prec)
        u, un = s.new_state(), s.new_state()
        s.ic_gradient(u, p.gamma)
        f = s.setup_forcing(p, amp=1e4)
        t, cmax = 0.0, s.ctot_max(u)
        for it in range(30):
            while t >= f.next_time:
                f.advance()
            dt = min(p.courant / N / cmax, f.next_time - t)
            s.step(u, un, step_args(p, dt, t, do_turb=True, riemann="hlld", recon="ppm", want_cfl=True))
            u, un = un, u
            t += dt
            cmax = s.cfl_max()
        q, qn = s.new_state(8), s.new_state(8)
        q.copy_(u[:8])
        for it in range(5):
            s.ct_step(q, qn, ct_args(p, 0.2 * dt, "hlld", "plm", do_turb=True, t=t), 0)
            q, qn = qn, q
        out.append((u.clone(), q.clone(), cmax))
        s.close()
    assert torch.equal(out[0][0], out[1][0]) and out[0][2] == out[1][2]
    assert torch.equal(out[0][1], out[1][1])


def test_largest_grid_fp32():
    """1300^3 in FP32 (two states: 158 GB of HBM): each plane holds 2.2e9 cells, more than a 32-bit
    index reaches.  A uniform state stays exactly uniform through the step, and on the gradient
    state mass and total energy (the library's compensated FP64 sums) change by less than half an
    FP32 ulp per cell (3e-8 of the total: every face flux is shared bit for bit by its two cells,
    so only the per-cell rounding of the update is left), with the fused CFL equal to the
    standalone one."""
    N = 1300
    if torch.cuda.get_device_properties(0).total_memory < 200 * 2 ** 30:
        pytest.skip("needs a GPU with >= 200 GiB")
    torch.cuda.empty_cache()
    p = Params(N=N, gamma=5.0 / 3.0, courant=0.3)
    s = Solver(p, "f32")
    u, un = s.new_state(), s.new_state()
    assert u[0].numel() > 2 ** 31
    s.ic_uniform(u, 1.0, 0.6, (0.5, 0.0, 0.0), p.gamma)
    s.step(u, un, step_args(p, 1e-4, 0.0, riemann="hlld", recon="ppm", want_cfl=False))
    for v in range(9):
        assert torch.equal(u[v], un[v]), v
    s.ic_gradient(u, p.gamma)
    d0 = s.diagnostics(u)
    dt = 0.3 * (1.0 / N) / s.ctot_max(u)
    s.step(u, un, step_args(p, dt, 0.0, riemann="hlld", recon="ppm", want_cfl=True))
    d1 = s.diagnostics(un)
    assert abs(d1["mass"] - d0["mass"]) <= 3e-8 * d0["mass"]
    assert abs(d1["etot"] - d0["etot"]) <= 3e-8 * abs(d0["etot"])
    assert d1["finite"]
    assert abs(s.cfl_max() - s.ctot_max(un)) <= 1e-6 * s.cfl_max()
    s.close()
    del u, un
    torch.cuda.empty_cache()


def test_fp32_mode_diagnostics_vs_fp64(orc):
    """FP32 state/flux with FP64 pressure recovery and reductions: after 60 forced steps of a 48^3
    developed state, E_kin and E_mag agree with the FP64 run to 1e-5 relative."""
    N = 48
    p = Params(N=N, gamma=5.0 / 3.0, courant=0.3, turb_gs=16, turb_kmin=1.0, turb_kmax=3.0, comp_frac=0.0,
               turb_T=0.5, turb_seed=3)
    u0 = smooth_state(orc, N, N, N, seed=21)
    out = {}
    for prec in ("f64", "f32"):
        s = Solver(p, prec)
        f = s.setup_forcing(p, amp=5e3)
        ud, un = to_dev(s, u0), s.new_state()
        t = 0.0
        cmax = s.ctot_max(ud)
        for it in range(60):
            while t >= f.next_time:
                f.advance()
            dt = min(p.courant / N / cmax, f.next_time - t)
            s.step(ud, un, step_args(p, dt, t + 0.1, do_turb=True, riemann="hlld", recon="ppm", want_cfl=True))
            ud, un = un, ud
            t += dt
            cmax = s.cfl_max()
        out[prec] = s.diagnostics(ud)
        s.close()
    for k in ("ekin", "emag"):
        assert abs(out["f32"][k] - out["f64"][k]) <= 1e-5 * abs(out["f64"][k]), (k, out["f32"][k], out["f64"][k])
    assert abs(out["f32"]["cross_helicity"] - out["f64"]["cross_helicity"]) <= 1e-4 * abs(out["f64"]["cross_helicity"])
    assert out["f32"]["finite"]
