"""The MarchBridge libraries the build scripts produce (build_mhd.sh -> libmhd480.so, build_ct.sh
-> libct480.so), opened as a caller of the reference opens them and run on the GPU against the
FP32 oracle (the legacy ABI is float**; spike scheme: normal-only Hancock, PLM MonCen, HLL)."""
import ctypes as C
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from oracle import CTScheme, Scheme  # noqa: E402
from test_gpu_parity import rel_l1, smooth_state  # noqa: E402

from glmmhdturb_amd import _lib  # noqa: E402
from glmmhdturb_amd import march_bridge as MB  # noqa: E402

PKG = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "glmmhdturb_amd")
G32 = float(np.float32(5.0 / 3.0))


def test_libmhd_through_march_bridge(orc, orc32):
    m = MB.open_lib(os.path.join(PKG, "libmhd480.so"))
    assert m.NV == 9 and m.N == 480 and m.f_glm is not None
    N = 32
    MB.set_grid(m, N, N, N)
    MB.set_gamma(m, 5.0 / 3.0)
    dtdx, ch, fac = 0.04, 1.1, 0.98
    MB.set_dtdx(m, dtdx)
    MB.set_glm(m, ch, fac)
    u0 = smooth_state(orc, N, N, N, seed=21).astype(np.float32)
    q = [torch.from_numpy(u0[v].copy()).cuda() for v in range(9)]
    o = [torch.empty_like(x) for x in q]
    assert MB.run_(m, q, o, 3) >= 0.0
    sc = Scheme(gamma=G32, smallr=1e-6, pfl=1e-6, dt=np.float32(dtdx), dx=1.0, ch=np.float32(ch),
                glm_fac=np.float32(fac), predictor="oned", recon="plm", slope="moncen", riemann="hll",
                mh_revert=False, face_guard="clamp", update_floor=False)
    uo = u0.copy()
    for _ in range(3):
        uo = orc32.step(uo, sc)
    got = np.stack([x.cpu().numpy() for x in q])
    assert np.all(rel_l1(got, uo) <= 2e-6)
    # march_args is the same scheme on the context API: the same bits as the bridge library's steps
    from glmmhdturb_amd import Params, Solver, march_args
    s = Solver(Params(N=N, gamma=5.0 / 3.0), "f32")
    a = march_args(float(np.float32(dtdx)), float(np.float32(5.0 / 3.0)), ch=float(np.float32(ch)),
                   glm_fac=float(np.float32(fac)))
    u, un = torch.from_numpy(u0).cuda().contiguous(), s.new_state()
    for _ in range(3):
        s.step(u, un, a)
        u, un = un, u
    assert torch.equal(u, torch.stack(q))
    s.close()


def test_libct_ct_run(orc, orc32):
    from test_gpu_ct_march import ct_state
    path = os.path.join(PKG, "libct480.so")
    with pytest.raises(AttributeError):
        MB.open_lib(path)          # MarchBridge.open_lib needs march_run_dev, which a CT library lacks
    lib = C.CDLL(path)
    lib.march_nv.restype = C.c_int
    lib.march_set_grid.restype = C.c_int
    lib.ct_run.restype = C.c_double
    lib.ct_run.argtypes = [C.POINTER(C.c_void_p)] * 3 + [C.c_int]
    lib.ct_regs.restype = C.c_int
    lib.ct_shmem.restype = C.c_int
    assert lib.march_nv() == 8
    N = 24
    assert lib.march_set_grid(N, N, N) == 0
    lib.march_set_gamma(C.c_float(G32))
    dtdx = float(np.float32(0.02))
    lib.march_set_dtdx(C.c_float(dtdx))
    u = ct_state(orc, N, N, N).astype(np.float32)
    q = [torch.from_numpy(u[v].copy()).cuda() for v in range(8)]
    o = [torch.empty_like(x) for x in q]
    cc = [torch.empty_like(x) for x in q[:3]]
    qa, kq = _lib.plane_array([x.data_ptr() for x in q])
    oa, ko = _lib.plane_array([x.data_ptr() for x in o])
    ca, kc = _lib.plane_array([x.data_ptr() for x in cc])
    assert lib.ct_run(qa, oa, ca, 3) >= 0.0
    cs = CTScheme(gamma=G32, smallr=1e-6, smallp=1e-6, dt=dtdx, dx=1.0, riemann="hll", recon="plm", face_slope=0)
    uo = u.copy()
    for _ in range(3):
        uo = orc32.ct_step(uo, cs)
    got = np.stack([x.cpu().numpy() for x in q])
    assert np.all(rel_l1(got, uo, 8) <= 2e-6)
    assert lib.ct_regs() > 0 and lib.ct_shmem() > 0
