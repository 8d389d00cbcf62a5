"""Error behaviour of the C ABI on the GPU (include/glmmhd.h, section "Error handling").

The reference has no error reporting (spike_mhd.cu's library path checks nothing; MarchBridge.jl
only asserts shapes, MarchBridge.jl:74-75).  Here every entry point validates its arguments before
it launches anything and returns a status code, with glm_last_error() / glm_error_string() naming
the problem.  Each case below is rejected on the host; the test checks the code, the message, that
the rejected call wrote nothing, and that the context still steps correctly afterwards."""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from glmmhdturb_amd import Params, Solver, _lib, step_args  # noqa: E402

OK, ARG, STATE = 0, 1, 4
N = 16


def planes(u, nv=9, null=None):
    ptrs = [u[v].data_ptr() for v in range(nv)]
    if null is not None:
        ptrs[null] = 0
    return _lib.plane_array(ptrs)


def expect(lib, code, want, words):
    assert code == want, (code, lib.glm_error_string(code))
    assert lib.glm_last_error() == want
    msg = lib.glm_error_string(code).decode()
    assert words in msg, msg


@pytest.fixture()
def ctx64():
    p = Params(N=N, gamma=5.0 / 3.0, courant=0.3)
    s = Solver(p, "f64")
    u, o = s.new_state(), s.new_state()
    s.ic_gradient(u, p.gamma)
    yield p, s, u, o
    s.close()


def test_step_rejects_bad_arguments_and_writes_nothing(ctx64):
    p, s, u, o = ctx64
    lib = s.lib
    o.fill_(7.0)
    a = step_args(p, 1e-3, 0.0, riemann="hlld", recon="ppm", want_cfl=False)
    q, kq = planes(u)
    expect(lib, lib.glm_step(s.ctx, planes(u, null=4)[0], planes(o)[0], C.byref(a)), ARG, "null plane pointer")
    expect(lib, lib.glm_step(s.ctx, q, planes(o, null=8)[0], C.byref(a)), ARG, "null plane pointer")
    expect(lib, lib.glm_step(s.ctx, q, q, C.byref(a)), ARG, "must differ")
    for field, bad in (("riemann", 9), ("recon", -1), ("predictor", 2), ("slope", 5), ("face_guard", 3)):
        b = step_args(p, 1e-3, 0.0, riemann="hlld", recon="ppm", want_cfl=False)
        setattr(b, field, bad)
        expect(lib, lib.glm_step(s.ctx, q, planes(o)[0], C.byref(b)), ARG, "scheme switch out of range")
    for field, bad in (("gamma", 1.0), ("dx", 0.0), ("ch", -1.0), ("smallr", 0.0), ("gamma", float("nan"))):
        b = step_args(p, 1e-3, 0.0, riemann="hlld", recon="ppm", want_cfl=False)
        setattr(b, field, bad)
        expect(lib, lib.glm_step(s.ctx, q, planes(o)[0], C.byref(b)), ARG, "need gamma > 1")
    b = step_args(p, 1e-3, 0.0, do_turb=True, riemann="hlld", recon="ppm", want_cfl=False)
    expect(lib, lib.glm_step(s.ctx, q, planes(o)[0], C.byref(b)), STATE, "glm_ou_setup")
    expect(lib, lib.glm_step(s.ctx, q, planes(o)[0], None), ARG, "null glm_params")
    torch.cuda.synchronize()
    assert bool((o == 7.0).all())
    # the context is unharmed: a valid step runs and its CFL slot is readable
    a.want_cfl = 1
    s.step(u, o, a)
    torch.cuda.synchronize()
    assert bool(torch.isfinite(o).all()) and s.cfl_max() > 0.0


def test_reductions_ic_and_forcing_reject_bad_arguments(ctx64):
    p, s, u, o = ctx64
    lib = s.lib
    a = step_args(p, 1e-3, 0.0)
    out = (C.c_double * _lib.NDIAG)()
    cm = C.c_double()
    expect(lib, lib.glm_cfl_max(s.ctx, C.byref(cm)), STATE, "no step with want_cfl")
    expect(lib, lib.glm_ctot_max(s.ctx, planes(u, null=3)[0], C.byref(a), C.byref(cm)), ARG, "null plane pointer")
    expect(lib, lib.glm_diagnostics(s.ctx, planes(u, null=0)[0], C.byref(a), out), ARG, "null plane pointer")
    args = (C.c_double * 1)(p.gamma)
    expect(lib, lib.glm_ic(s.ctx, planes(u, null=5)[0], 0, args), ARG, "null plane pointer")
    expect(lib, lib.glm_ic(s.ctx, planes(u)[0], 3, args), ARG, "unknown kind")
    expect(lib, lib.glm_ou_advance(s.ctx), STATE, "glm_ou_setup first")
    for g, T, ndt, kmin, kmax in ((1, 0.5, 10, 1.0, 3.0), (300, 0.5, 10, 1.0, 3.0), (16, 0.0, 10, 1.0, 3.0),
                                  (16, 0.5, 0, 1.0, 3.0), (16, 0.5, 10, 3.0, 1.0)):
        expect(lib, lib.glm_ou_setup(s.ctx, g, kmin, kmax, T, ndt, 0.0, 1.0, 1), ARG, "2 <= g <= 256")
    assert not lib.glm_create(0, N, N, 8)
    assert lib.glm_last_error() == ARG
    assert not lib.glm_create(N, N, N, 2)
    assert lib.glm_last_error() == ARG
    # a valid reduction still works on the same context
    d = s.diagnostics(u)
    assert d["finite"] and d["mass"] > 0.0


def test_adjoint_and_ct_reject_unsupported_schemes(ctx64):
    p, s, u, o = ctx64
    lib = s.lib
    lam1, lam0 = s.new_state(), s.new_state()
    lam0.fill_(3.0)
    a = step_args(p, 1e-3, 0.0, riemann="hlld", recon="ppm_spike", want_cfl=False)
    qu, qa, qb = planes(u)[0], planes(lam1)[0], planes(lam0)[0]
    expect(lib, lib.glm_step_adjoint(s.ctx, qu, qa, qb, C.byref(a), None), ARG, "PPM_SPIKE has no adjoint")
    a = step_args(p, 1e-3, 0.0, riemann="hlld", recon="ppm", want_cfl=False)
    expect(lib, lib.glm_step_adjoint(s.ctx, qu, qa, qa, C.byref(a), None), ARG, "must not alias")
    expect(lib, lib.glm_step_adjoint(s.ctx, qu, qa, qu, C.byref(a), None), ARG, "must not alias")
    b = step_args(p, 1e-3, 0.0, do_turb=True, riemann="hlld", recon="ppm", want_cfl=False)
    expect(lib, lib.glm_step_adjoint(s.ctx, qu, qa, qb, C.byref(b), None), STATE, "forcing field")
    torch.cuda.synchronize()
    assert bool((lam0 == 3.0).all())
    # CT: 8 planes, PLM / PPM_SPIKE and LLF / HLL / HLLD only
    c8 = s.new_state(8)
    d8 = s.new_state(8)
    for riem, recon in (("llf_bound", "plm"), ("hlld", "ppm")):
        c = step_args(p, 1e-3, 0.0, riemann=riem, recon=recon, want_cfl=False)
        expect(lib, lib.glm_ct_step(s.ctx, planes(c8, 8)[0], planes(d8, 8)[0], C.byref(c), 0), ARG,
               "recon PLM/PPM_SPIKE")
    c = step_args(p, 1e-3, 0.0, riemann="hlld", recon="plm", want_cfl=False)
    expect(lib, lib.glm_ct_step(s.ctx, planes(c8, 8)[0], planes(c8, 8)[0], C.byref(c), 0), ARG, "bad plane pointers")
    cm = C.c_double()
    expect(lib, lib.glm_ct_ctot_max(s.ctx, planes(c8, 8, null=7)[0], C.byref(c), C.byref(cm)), ARG,
           "null plane pointer")
    expect(lib, lib.glm_ct_divb_max(s.ctx, planes(c8, 8, null=6)[0], 1.0 / N, C.byref(cm)), ARG, "null plane pointer")
    expect(lib, lib.glm_ct_divb_max(s.ctx, planes(c8, 8)[0], 0.0, C.byref(cm)), ARG, "bad argument")


def test_fp32_context_has_no_adjoint():
    p = Params(N=N, gamma=5.0 / 3.0)
    s = Solver(p, "f32")
    u, l1, l0 = s.new_state(), s.new_state(), s.new_state()
    a = step_args(p, 1e-3, 0.0, riemann="hlld", recon="ppm", want_cfl=False)
    code = s.lib.glm_step_adjoint(s.ctx, planes(u)[0], planes(l1)[0], planes(l0)[0], C.byref(a), None)
    expect(s.lib, code, ARG, "FP64 only")
    with pytest.raises(_lib.GLMError, match="FP64 only"):
        s.step_adjoint(u, l1, l0, a)
    s.close()


def test_python_layer_raises_with_the_library_message(ctx64):
    p, s, u, o = ctx64
    a = step_args(p, 1e-3, 0.0, riemann="hlld", recon="ppm", want_cfl=False)
    a.riemann = 11
    with pytest.raises(_lib.GLMError, match="scheme switch out of range"):
        s.step(u, o, a)
    with pytest.raises(ValueError):
        s.step(u.float(), o, step_args(p, 1e-3, 0.0))
    with pytest.raises(_lib.GLMError, match="do_turb requires setup_forcing"):
        s.step_plm(u, o, p, 1e-3, 0.0, do_turb=True)
    assert np.isfinite(s.ctot_max(u))


def test_cross_plane_aliasing_is_rejected(ctx64):
    """An output plane that is a different input plane (o[w] == q[v], w != v) races in the z-chunked
    marches exactly as o[v] == q[v] does: every step entry point checks all pairs."""
    p, s, u, o = ctx64
    lib = s.lib
    o.fill_(5.0)
    a = step_args(p, 1e-3, 0.0, riemann="hlld", recon="ppm", want_cfl=False)
    optrs = [o[v].data_ptr() for v in range(9)]
    optrs[3] = u[5].data_ptr()
    expect(lib, lib.glm_step(s.ctx, planes(u)[0], _lib.plane_array(optrs)[0], C.byref(a)), ARG, "must differ")
    # adjoint: lam0 against every lam1 and u0 plane
    lam1, lam0 = s.new_state(), s.new_state()
    l0 = [lam0[v].data_ptr() for v in range(9)]
    l0[2] = lam1[7].data_ptr()
    expect(lib, lib.glm_step_adjoint(s.ctx, planes(u)[0], planes(lam1)[0], _lib.plane_array(l0)[0], C.byref(a), None),
           ARG, "must not alias")
    l0 = [lam0[v].data_ptr() for v in range(9)]
    l0[1] = u[4].data_ptr()
    expect(lib, lib.glm_step_adjoint(s.ctx, planes(u)[0], planes(lam1)[0], _lib.plane_array(l0)[0], C.byref(a), None),
           ARG, "must not alias")
    # CT: 8 planes
    c8, d8 = s.new_state(8), s.new_state(8)
    dp = [d8[v].data_ptr() for v in range(8)]
    dp[6] = c8[0].data_ptr()
    c = step_args(p, 1e-3, 0.0, riemann="hlld", recon="plm", want_cfl=False)
    expect(lib, lib.glm_ct_step(s.ctx, planes(c8, 8)[0], _lib.plane_array(dp)[0], C.byref(c), 0), ARG, "alias")
    torch.cuda.synchronize()
    assert bool((o == 5.0).all())


def test_forcing_field_is_validated_before_copying(ctx64):
    """set_afield / step_adjoint(afield=...) hand the library a raw address it reads 3 g^3 values
    from: a field of the wrong dtype, shape, layout or device is refused on the host."""
    p, s, u, o = ctx64
    s.setup_forcing(p, amp=1.0)
    g = p.turb_gs
    good = s.afield()
    s.set_afield(good)
    for bad in (good.float(), good[:, :-1].contiguous(), good.transpose(1, 3), good.cpu()):
        with pytest.raises(_lib.GLMError, match="forcing field"):
            s.set_afield(bad)
    lam1, lam0 = s.new_state(), s.new_state()
    a = step_args(p, 1e-3, 0.0, do_turb=True, riemann="hlld", recon="ppm", want_cfl=False)
    with pytest.raises(_lib.GLMError, match="forcing field"):
        s.step_adjoint(u, lam1, lam0, a, afield=torch.zeros((3, g, g, g - 1), dtype=torch.float64, device=u.device))
    assert torch.equal(s.afield(), good)
