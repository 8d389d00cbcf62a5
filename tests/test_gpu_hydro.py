"""The 5-variable hydro line march (glm_hydro_step and the 5-plane MarchBridge library
libmarch480.so; spike_25d.cu:383-576 as build_march.sh builds it) against the CPU oracle
(oracle/hydro_oracle.c).

Bars: FP64 relative L1 <= 1e-12 per variable after several steps (FMA contraction and the
reciprocal/root sequences only); FP32 <= 2e-6; chunkings of z bit-identical; at sizes the oracle
cannot follow, conservation, the exact Sod solution and axis symmetry (to rounding: each axis has
its own contracted code on the GPU, where the oracle is bit-symmetric)."""
import ctypes as C
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle import HydroScheme  # noqa: E402
from glmmhdturb_amd import Params, Solver, hydro_args, _lib  # noqa: E402
from test_hydro_oracle import cons_of, exact_riemann, smooth_hydro  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
G = 1.4


def rel_l1(a, b):
    return np.array([np.abs(a[v] - b[v]).sum() / max(np.abs(b[v]).sum(), 1e-300) for v in range(5)])


def solver(nx, ny, nz, prec="f64"):
    return Solver(Params(N=nx, nx=nx, ny=ny, nz=nz, gamma=G), prec)


def dev(s, u):
    t = s.new_state(5)
    t.copy_(torch.from_numpy(np.ascontiguousarray(u)).to(t.device, t.dtype))
    return t


def run_gpu(s, u, args, steps):
    a, b = dev(s, u), s.new_state(5)
    for _ in range(steps):
        s.hydro_step(a, b, args)
        a, b = b, a
    torch.cuda.synchronize()
    return a.cpu().numpy().astype(np.float64)


def run_oracle(orc, u, sc, steps):
    for _ in range(steps):
        u = orc.hydro_step(u, sc)
    return u


def voids(nx, ny, nz, seed):
    """Smooth flow with low-density, low-pressure pockets: faces below the 1e-3 fallback
    floors, and supersonic jets into them."""
    u = smooth_hydro(nx, ny, nz, seed, amp=0.2)
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(np.arange(nz) / nz, np.arange(ny) / ny, np.arange(nx) / nx, indexing="ij")
    rho = u[0].copy()
    vx = u[1] / rho
    vy = u[2] / rho
    vz = u[3] / rho
    p = (G - 1) * (u[4] - 0.5 * rho * (vx ** 2 + vy ** 2 + vz ** 2))
    for _ in range(3):
        c = rng.uniform(0, 1, 3)
        r2 = ((x - c[0] + 0.5) % 1 - 0.5) ** 2 + ((y - c[1] + 0.5) % 1 - 0.5) ** 2 + ((z - c[2] + 0.5) % 1 - 0.5) ** 2
        m = np.minimum(1.5 * np.exp(-r2 / 0.01), 1.0)   # a plateau at the full depth
        rho = rho * (1 - 0.9995 * m)
        p = p * (1 - 0.9995 * m)
    vx = vx + 2.0 * np.sin(2 * np.pi * y)
    return cons_of(rho, vx, vy, vz, p)


CASES = [  # (nx, ny, nz, dtdx, fallback, state)
    (32, 8, 6, 0.12, None, "smooth"),
    (40, 20, 12, 0.12, None, "smooth"),     # partial tiles in x and y
    (7, 5, 3, 0.1, None, "smooth"),         # a tile larger than the grid
    (33, 17, 9, 0.05, (1e-3, 1e-3), "voids"),
    (33, 17, 9, 0.05, None, "voids"),
]


@pytest.mark.parametrize("k", range(len(CASES)))
def test_hydro_step_fp64_vs_oracle(orc, k):
    nx, ny, nz, dtdx, fb, kind = CASES[k]
    u = smooth_hydro(nx, ny, nz, 20 + k) if kind == "smooth" else voids(nx, ny, nz, 20 + k)
    sc = HydroScheme(gamma=G, dtdx=dtdx, fallback=fb is not None, fb_dmin=(fb or (0, 0))[0], fb_pmin=(fb or (0, 0))[1])
    s = solver(nx, ny, nz)
    got = run_gpu(s, u, hydro_args(dtdx, G, fallback=fb), 5)
    ref = run_oracle(orc, u, sc, 5)
    err = rel_l1(got, ref)
    print(f"[hydro f64] {CASES[k]}: rel-L1 {err.max():.2e}", flush=True)
    assert np.all(err <= 1e-12), err
    if fb is not None:   # the fallback is active in this state: the plain HLL answer differs
        alt = run_oracle(orc, u, HydroScheme(gamma=G, dtdx=dtdx), 5)
        assert rel_l1(alt, ref).max() > 1e-9
    s.close()


@pytest.mark.parametrize("k", [1, 3])
def test_hydro_step_fp32_vs_oracle(orc32, k):
    nx, ny, nz, dtdx, fb, kind = CASES[k]
    u = smooth_hydro(nx, ny, nz, 20 + k) if kind == "smooth" else voids(nx, ny, nz, 20 + k)
    u = u.astype(np.float32)
    sc = HydroScheme(gamma=float(np.float32(G)), dtdx=float(np.float32(dtdx)), fallback=fb is not None,
                     fb_dmin=(fb or (0, 0))[0], fb_pmin=(fb or (0, 0))[1])
    s = solver(nx, ny, nz, "f32")
    got = run_gpu(s, u, hydro_args(float(np.float32(dtdx)), float(np.float32(G)), fallback=fb), 3)
    ref = run_oracle(orc32, u, sc, 3).astype(np.float64)
    err = rel_l1(got, ref)
    print(f"[hydro f32] {CASES[k]}: rel-L1 {err.max():.2e}", flush=True)
    assert np.all(err <= 2e-6), err
    s.close()


def test_hydro_zchunk_is_bit_invariant(orc):
    nx, ny, nz = 40, 20, 24
    u = smooth_hydro(nx, ny, nz, 3)
    s = solver(nx, ny, nz)
    outs = [run_gpu(s, u, hydro_args(0.12, G, zchunk=zc), 3) for zc in (0, 1, 5, 24)]
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
    s.close()


def test_hydro_errors():
    s = solver(16, 8, 4)
    a, b = s.new_state(5), s.new_state(5)
    qa, kq = _lib.plane_array([a[v].data_ptr() for v in range(5)])
    with pytest.raises(_lib.GLMError, match="alias"):
        _lib.check(s.lib.glm_hydro_step(s.ctx, qa, qa, C.byref(hydro_args(0.1))))
    bad = hydro_args(0.1, 1.0)
    ob, ko = _lib.plane_array([b[v].data_ptr() for v in range(5)])
    with pytest.raises(_lib.GLMError, match="gamma"):
        _lib.check(s.lib.glm_hydro_step(s.ctx, qa, ob, C.byref(bad)))
    with pytest.raises(_lib.GLMError, match="nsp must be 0 or 2"):
        _lib.check(s.lib.glm_hydro_step_scalars(s.ctx, qa, ob, C.byref(hydro_args(0.1)), 1, 0))
    nulls = (C.c_void_p * 7)(*([a[v].data_ptr() for v in range(5)] + [None, None]))
    outs = (C.c_void_p * 7)(*([b[v].data_ptr() for v in range(5)] + [b[0].data_ptr(), b[1].data_ptr()]))
    with pytest.raises(_lib.GLMError, match="null plane"):
        _lib.check(s.lib.glm_hydro_step_scalars(s.ctx, C.cast(nulls, C.POINTER(C.c_void_p)),
                                                C.cast(outs, C.POINTER(C.c_void_p)), C.byref(hydro_args(0.1)), 2, 1))
    s.close()


def test_legacy_libmarch_run_dev(orc32):
    """libmarch480.so (build_march.sh): march_nv = 5, march_set_grid/dtdx/gamma, march_run_dev
    over 5 float planes with the result left in q after an odd step count (spike_25d.cu:563-575)."""
    from glmmhdturb_amd import march_bridge as mb
    m = mb.open_lib(os.path.join(ROOT, "glmmhdturb_amd", "libmarch480.so"))
    assert m.NV == 5 and m.N == 480 and m.f_glm is None
    nx = ny = nz = 40          # MarchBridge.jl:74-75 asserts N^3 planes
    mb.set_grid(m, nx, ny, nz)
    g, dtdx = float(np.float32(G)), float(np.float32(0.1))
    mb.set_gamma(m, g)
    mb.set_dtdx(m, dtdx)
    u = smooth_hydro(nx, ny, nz, 31).astype(np.float32)
    q = [torch.from_numpy(u[v].copy()).cuda() for v in range(5)]
    o = [torch.empty_like(x) for x in q]
    assert mb.run_(m, q, o, 3) >= 0.0
    ref = run_oracle(orc32, u, HydroScheme(gamma=g, dtdx=dtdx), 3).astype(np.float64)
    got = np.stack([x.cpu().numpy() for x in q]).astype(np.float64)
    assert np.all(rel_l1(got, ref) <= 2e-6)


def test_hydro_sod_large_grid_exact_and_symmetric():
    """Sod tubes along x, y and z through a 512 x 8 x 8 block of tubes at once per axis (sizes the
    oracle would take minutes over): the exact solution, the same result on every axis, identical
    tubes, and conservation of mass, momentum and energy to round-off."""
    n, dtdx = 512, 0.2
    dx = 2.0 / n
    x = (np.arange(n) + 0.5) * dx
    left = (x >= 0.5) & (x < 1.5)
    line = cons_of(np.where(left, 1.0, 0.125), 0 * x, 0 * x, 0 * x, np.where(left, 1.0, 0.1))
    steps = int(round(0.2 / (dtdx * dx)))
    res = []
    for axis in range(3):
        shape = [8, 8, 8]
        shape[2 - axis] = n
        u = np.empty((5, *shape))
        for v in range(5):
            u[v] = np.moveaxis(np.broadcast_to(line[v], (8, 8, n)), 2, 2 - axis)
        nz_, ny_, nx_ = shape
        s = solver(nx_, ny_, nz_)
        got = run_gpu(s, u, hydro_args(dtdx, G), steps)
        s.close()
        scale = np.abs(u[0]).sum() + np.abs(u[4]).sum()
        for v in range(5):
            assert abs(got[v].sum() - u[v].sum()) <= 1e-13 * scale
        lines = np.moveaxis(got, 3 - axis, -1).reshape(5, -1, n)
        assert np.all(lines == lines[:, :1, :])      # every tube of the block is the same
        res.append(lines[:, 0, :].copy())
    # the same tube along every axis, to the rounding of each axis's own (contracted) code
    for a_ in (1, 2):    # density, normal momentum and energy; the transverse momenta stay zero
        assert np.abs(res[a_][[0, 1 + a_, 4]] - res[0][[0, 1, 4]]).max() <= 1e-12
        assert not np.any(np.delete(res[a_], [0, 1 + a_, 4], axis=0))
    sel = (x > 1.0) & (x < 1.95)
    t = steps * dtdx * dx
    ex = exact_riemann(1.0, 0.0, 1.0, 0.125, 0.0, 0.1, (x[sel] - 1.5) / t)
    assert np.abs(res[0][0][sel] - ex[:, 0]).mean() < 0.0035


@pytest.mark.parametrize("cma,fb,prec", [(False, None, "f64"), (True, None, "f64"), (True, (1e-3, 1e-3), "f64"),
                                         (True, None, "f32"), (False, (1e-3, 1e-3), "f32")])
def test_hydro_scalars_vs_oracle(orc, orc32, cma, fb, prec):
    """Two passive scalars (glm_hydro_step_scalars; spike_25d.cu -DSCALARS, with -DCMA) against
    the oracle, on partial tiles and, with the fallback, near-vacuum pockets."""
    nx, ny, nz, dtdx = 33, 17, 9, 0.05
    u = voids(nx, ny, nz, 40) if fb else smooth_hydro(nx, ny, nz, 41)
    rng = np.random.default_rng(42)
    x0 = 0.5 + 0.3 * np.tanh(rng.normal(0, 1, u[0].shape))
    us = np.concatenate([u, (u[0] * x0)[None], (u[0] * (1 - x0))[None]])
    o_ = orc if prec == "f64" else orc32
    cast = (lambda v: v) if prec == "f64" else (lambda v: float(np.float32(v)))
    if prec == "f32":
        us = us.astype(np.float32)
    sc = HydroScheme(gamma=cast(G), dtdx=cast(dtdx), fallback=fb is not None, fb_dmin=(fb or (0, 0))[0],
                     fb_pmin=(fb or (0, 0))[1], nsp=2, cma=cma)
    s = solver(nx, ny, nz, prec)
    a, b = s.new_state(7), s.new_state(7)
    a.copy_(torch.from_numpy(np.ascontiguousarray(us)).to(a.device, a.dtype))
    args = hydro_args(cast(dtdx), cast(G), fallback=fb)
    for _ in range(4):
        s.hydro_step(a, b, args, nsp=2, cma=cma)
        a, b = b, a
    torch.cuda.synchronize()
    got = a.cpu().numpy().astype(np.float64)
    ref = run_oracle(o_, us, sc, 4).astype(np.float64)
    err = np.array([np.abs(got[v] - ref[v]).sum() / np.abs(ref[v]).sum() for v in range(7)])
    print(f"[hydro scalars {prec} cma={cma} fb={fb}] rel-L1 {err.max():.2e}", flush=True)
    assert np.all(err <= (1e-12 if prec == "f64" else 2e-6)), err
    # the hydro planes are those of the 5-variable march (to the rounding of its own kernel)
    h = run_gpu(s, us[:5], args, 4)
    assert np.all(np.abs(h - got[:5]) <= (1e-13 if prec == "f64" else 1e-5) * (np.abs(h) + 1))
    s.close()


def test_hydro_largest_grid_fp32():
    """1300^3 in FP32 (two 5-plane states: 88 GB of HBM; each plane holds 2.2e9 cells, more than a
    32-bit index reaches).  A uniform state stays uniform; on a state that varies in x and z only,
    every y row of the result is the same bit for bit, mass changes by less than half an FP32 ulp
    per cell, and the result is finite."""
    N = 1300
    if torch.cuda.get_device_properties(0).total_memory < 200 * 2 ** 30:
        pytest.skip("needs a GPU with >= 200 GiB")
    torch.cuda.empty_cache()
    s = solver(N, N, N, "f32")
    u, un = s.new_state(5), s.new_state(5)
    assert u[0].numel() > 2 ** 31
    for v, c in enumerate(cons_of(1.2, 0.3, -0.2, 0.1, 0.8)):
        u[v].fill_(float(c))
    args = hydro_args(0.1, G)
    s.hydro_step(u, un, args)
    assert torch.equal(un[0], u[0])
    for v in range(1, 5):
        assert bool((un[v] == un[v].view(-1)[0]).all()), v
    x = (torch.arange(N, device=u.device, dtype=torch.float64) + 0.5) / N
    X, Z = x[None, :], x[:, None]                       # (z, x) planes
    tp = 2 * np.pi
    rho = 1 + 0.2 * torch.sin(tp * X) + 0.1 * torch.cos(tp * Z)
    vx, vz = 0.3 * torch.sin(tp * Z) + 0 * X, 0.2 * torch.sin(tp * X) + 0 * Z
    p = 1 + 0.1 * torch.cos(tp * X) + 0 * Z
    cons = [rho, rho * vx, 0 * rho, rho * vz, p / (G - 1) + 0.5 * rho * (vx ** 2 + vz ** 2)]
    for v in range(5):
        u[v].copy_(cons[v][:, None, :].to(torch.float32).expand(N, N, N))
    del rho, vx, vz, p, cons
    m0 = float(u[0].sum(dtype=torch.float64))
    s.hydro_step(u, un, args)
    torch.cuda.synchronize()
    m1 = float(un[0].sum(dtype=torch.float64))
    assert abs(m1 - m0) <= 3e-8 * m0
    for v in range(5):
        assert bool(torch.isfinite(un[v]).all())
        for j in (1, 777, N - 1):
            assert torch.equal(un[v][:, j, :], un[v][:, 0, :]), (v, j)
    assert not bool(un[2].any())                        # no y momentum appears
    s.close()
    del u, un
    torch.cuda.empty_cache()


def test_legacy_species_library_run_dev(orc32):
    """libmarch480_cma.so (the -DSCALARS -DCMA build): 7 float planes through the MarchBridge
    mirror, 3 steps, against the oracle with two CMA scalars."""
    from glmmhdturb_amd import march_bridge as mb
    m = mb.open_lib(os.path.join(ROOT, "glmmhdturb_amd", "libmarch480_cma.so"))
    assert m.NV == 7 and m.f_glm is None
    n = 32
    mb.set_grid(m, n, n, n)
    g, dtdx = float(np.float32(G)), float(np.float32(0.1))
    mb.set_gamma(m, g)
    mb.set_dtdx(m, dtdx)
    u = smooth_hydro(n, n, n, 51)
    x0 = 0.5 + 0.3 * np.sin(2 * np.pi * np.arange(n) / n)[None, None, :] * np.ones_like(u[0])
    us = np.concatenate([u, (u[0] * x0)[None], (u[0] * (1 - x0))[None]]).astype(np.float32)
    q = [torch.from_numpy(us[v].copy()).cuda() for v in range(7)]
    o = [torch.empty_like(x) for x in q]
    assert mb.run_(m, q, o, 3) >= 0.0
    ref = run_oracle(orc32, us, HydroScheme(gamma=g, dtdx=dtdx, nsp=2, cma=True), 3).astype(np.float64)
    got = np.stack([x.cpu().numpy() for x in q]).astype(np.float64)
    err = np.array([np.abs(got[v] - ref[v]).sum() / np.abs(ref[v]).sum() for v in range(7)])
    assert np.all(err <= 2e-6), err
