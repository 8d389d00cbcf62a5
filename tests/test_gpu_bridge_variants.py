"""The scheme builds of the MarchBridge libraries: PPM=1 build_mhd.sh -> libmhd480_ppm.so and
PPM=1 build_ct.sh -> libct480_ppm.so (spike_mhd.cu / spike_ctm.cu -DPPM), SCHEME=ppm1d and
SCHEME=donor build_march.sh -> libmarch480_ppm1d.so / libmarch480_donor.so (spike_25d.cu -DPPM1D,
-DDONOR).  Each is opened as a caller of the reference opens it, stepped through its legacy entry
point, and compared bit for bit with the same scheme on the context API (whose parity with the
oracle is tested in test_gpu_parity.py, test_gpu_ct_march.py and test_gpu_hydro_variants.py)."""
import ctypes as C
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from test_gpu_parity import smooth_state  # noqa: E402
from test_hydro_variants_oracle import periodic_state  # noqa: E402

from glmmhdturb_amd import Params, Solver, _lib, hydro_args, march_args  # noqa: E402
from glmmhdturb_amd import march_bridge as MB  # noqa: E402

PKG = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "glmmhdturb_amd")
f32 = lambda x: float(np.float32(x))  # noqa: E731


def test_libmhd_ppm(orc):
    m = MB.open_lib(os.path.join(PKG, "libmhd480_ppm.so"))
    assert m.NV == 9
    N, dtdx, ch, fac = 24, 0.04, 1.1, 0.98
    MB.set_grid(m, N, N, N)
    MB.set_gamma(m, 5.0 / 3.0)
    MB.set_dtdx(m, dtdx)
    MB.set_glm(m, ch, fac)
    u0 = smooth_state(orc, N, N, N, seed=23).astype(np.float32)
    q = [torch.from_numpy(u0[v].copy()).cuda() for v in range(9)]
    o = [torch.empty_like(x) for x in q]
    print("[bridge libmhd480_ppm]", flush=True)
    assert MB.run_(m, q, o, 3) >= 0.0
    s = Solver(Params(N=N, gamma=5.0 / 3.0), "f32")
    a = march_args(f32(dtdx), f32(5.0 / 3.0), ch=f32(ch), glm_fac=f32(fac), recon="ppm_spike")
    u, un = torch.from_numpy(u0).cuda().contiguous(), s.new_state()
    for _ in range(3):
        s.step(u, un, a)
        u, un = un, u
    assert torch.equal(u, torch.stack(q))
    # and it is not the default build's scheme
    a0 = march_args(f32(dtdx), f32(5.0 / 3.0), ch=f32(ch), glm_fac=f32(fac))
    u, un = torch.from_numpy(u0).cuda().contiguous(), s.new_state()
    for _ in range(3):
        s.step(u, un, a0)
        u, un = un, u
    assert not torch.equal(u, torch.stack(q))
    s.close()


def test_libct_ppm(orc):
    from test_gpu_ct_march import ct_state
    lib = C.CDLL(os.path.join(PKG, "libct480_ppm.so"))
    lib.march_nv.restype = C.c_int
    lib.march_set_grid.restype = C.c_int
    lib.ct_run.restype = C.c_double
    lib.ct_run.argtypes = [C.POINTER(C.c_void_p)] * 3 + [C.c_int]
    assert lib.march_nv() == 8
    N, g = 24, f32(5.0 / 3.0)
    assert lib.march_set_grid(N, N, N) == 0
    lib.march_set_gamma(C.c_float(g))
    dtdx = f32(0.02)
    lib.march_set_dtdx(C.c_float(dtdx))
    u = ct_state(orc, N, N, N).astype(np.float32)
    q = [torch.from_numpy(u[v].copy()).cuda() for v in range(8)]
    o = [torch.empty_like(x) for x in q]
    cc = [torch.empty_like(x) for x in q[:3]]
    qa, kq = _lib.plane_array([x.data_ptr() for x in q])
    oa, ko = _lib.plane_array([x.data_ptr() for x in o])
    ca, kc = _lib.plane_array([x.data_ptr() for x in cc])
    print("[bridge libct480_ppm]", flush=True)
    assert lib.ct_run(qa, oa, ca, 3) >= 0.0
    s = Solver(Params(N=N, gamma=g), "f32")
    a = march_args(dtdx, g, recon="ppm_spike")
    uc = torch.from_numpy(u).cuda().contiguous()
    un = torch.empty_like(uc)
    for _ in range(3):
        s.ct_step(uc, un, a)
        uc, un = un, uc
    assert torch.equal(uc, torch.stack(q))
    s.close()


@pytest.mark.parametrize("scheme", ["ppm1d", "donor"])
def test_libmarch_scheme(scheme):
    m = MB.open_lib(os.path.join(PKG, f"libmarch480_{scheme}.so"))
    assert m.NV == 5
    nx = ny = nz = 36                    # the bridge takes N^3 planes; 36 leaves partial tiles in x and y
    dtdx = 0.05
    MB.set_grid(m, nx, ny, nz)
    MB.set_gamma(m, 1.4)
    MB.set_dtdx(m, dtdx)
    u0 = periodic_state(nx, ny, nz, seed=7).astype(np.float32)
    q = [torch.from_numpy(u0[v].copy()).cuda() for v in range(5)]
    o = [torch.empty_like(x) for x in q]
    print(f"[bridge libmarch480_{scheme}]", flush=True)
    assert MB.run_(m, q, o, 3) >= 0.0
    s = Solver(Params(N=nx, nx=nx, ny=ny, nz=nz, gamma=1.4), "f32")
    a = hydro_args(f32(dtdx), f32(1.4), ppm=scheme == "ppm1d")
    u, un = s.new_state(5), s.new_state(5)
    u.copy_(torch.from_numpy(u0).cuda())
    for _ in range(3):
        s.hydro_step(u, un, a, donor=scheme == "donor")
        u, un = un, u
    assert torch.equal(u, torch.stack(q))
    plain = s.new_state(5)
    plain.copy_(torch.from_numpy(u0).cuda())
    for _ in range(3):
        s.hydro_step(plain, un, hydro_args(f32(dtdx), f32(1.4)))
        plain, un = un, plain
    assert not torch.equal(plain, u)
    s.close()
