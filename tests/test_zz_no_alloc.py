"""glm_diagnostics_noalloc (the entry point that allocates nothing) against glm_diagnostics, through
the C ABI: two calls in a row on one GPU, the second on a smaller grid, each equal bit for bit to
the context entry point's result on the same state (same kernels, same partition of the cells).
Also: FP32 states, replays of a captured HIP graph of the call, and the rejection of grids past
the static scratch's 2^27-cell maximum before any GPU work."""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from test_gpu_parity import smooth_state, to_dev  # noqa: E402

from glmmhdturb_amd import Params, Solver, _lib, step_args  # noqa: E402

GLM_ERR_ARG = 1


def planes(t):
    arr = (C.c_void_p * 9)()
    for v in range(9):
        arr[v] = t[v].data_ptr()
    return arr


def call_noalloc(lib, u, prec, a, out):
    nz, ny, nx = u.shape[1:]
    st = torch.cuda.current_stream().cuda_stream
    rc = lib.glm_diagnostics_noalloc(planes(u), nx, ny, nz, prec, C.byref(a), C.c_void_p(st), C.c_void_p(out.data_ptr()))
    assert rc == 0, lib.glm_error_string(rc)


def noalloc(lib, u, prec, a):
    out = torch.full((10,), float("nan"), dtype=torch.float64, device=u.device)
    call_noalloc(lib, u, prec, a, out)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def with_context(u, p, prec_name):
    s = Solver(p, prec_name)
    out = (C.c_double * _lib.NDIAG)()
    a = step_args(p, 1.0, 0.0)
    assert s.lib.glm_diagnostics(s.ctx, planes(u), C.byref(a), out) == 0
    s.close()
    return np.array(list(out))


def states(orc, prec_name):
    """A 160 x 128 x 96 gradient state (120 partials) and then a smaller 24 x 20 x 16 smooth state."""
    p1 = Params(N=160, nx=160, ny=128, nz=96, gamma=5.0 / 3.0)
    s = Solver(p1, prec_name)
    u1 = s.new_state()
    s.ic_gradient(u1, p1.gamma)
    s.close()
    p2 = Params(N=24, nx=24, ny=20, nz=16, gamma=5.0 / 3.0)
    s = Solver(p2, prec_name)
    u2 = to_dev(s, smooth_state(orc, 24, 20, 16, seed=11))
    s.close()
    return [(p1, u1), (p2, u2)]


@pytest.mark.parametrize("prec_name,prec", [("f64", 8), ("f32", 4)])
def test_noalloc_diagnostics_equal_context_diagnostics(orc, prec_name, prec):
    lib = _lib.load()
    for n, (p, u) in enumerate(states(orc, prec_name)):
        print(f"noalloc diagnostics {prec_name} call {n} grid {tuple(u.shape[1:])}", flush=True)
        a = step_args(p, 1.0, 0.0)
        got = noalloc(lib, u, prec, a)
        ref = with_context(u, p, prec_name)
        assert np.all(np.isfinite(got)) and got[4] == 1.0
        assert np.array_equal(got, ref), (got, ref)


def test_noalloc_diagnostics_in_a_captured_graph(orc):
    """The call captured into a HIP graph (two kernels, nothing else) and replayed three times with
    eager work in between: every replay equals the eager result."""
    lib = _lib.load()
    (p, u), _ = states(orc, "f64")
    a = step_args(p, 1.0, 0.0)
    ref = noalloc(lib, u, 8, a)
    out = torch.zeros(10, dtype=torch.float64, device="cuda")
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        call_noalloc(lib, u, 8, a, out)                   # warm-up on the capture stream's side
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        call_noalloc(lib, u, 8, a, out)
    for k in range(3):
        out.fill_(-1.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy(), ref), k
        _ = (u[0] * 2.0).sum().item()                     # eager work of the program between replays


def test_noalloc_rejects_grids_past_the_static_scratch():
    lib = _lib.load()
    p = Params(N=8, gamma=5.0 / 3.0)
    a = step_args(p, 1.0, 0.0)
    dummy = torch.zeros(9, 8, device="cuda")
    out = torch.zeros(10, dtype=torch.float64, device="cuda")
    rc = lib.glm_diagnostics_noalloc(planes(dummy), 1024, 1024, 129, 8, C.byref(a), None, C.c_void_p(out.data_ptr()))
    assert rc == GLM_ERR_ARG
    assert b"maximum" in lib.glm_error_string(rc)
    assert lib.glm_diagnostics_noalloc(planes(dummy), 8, 8, 8, 8, C.byref(a), None, None) == GLM_ERR_ARG
    torch.cuda.synchronize()
    assert torch.count_nonzero(out).item() == 0        # nothing was launched
