"""The CPU baseline (oracle/cpu_stream.cpp, a restatement of cpu_simd.jl's chunked z-stream) computes
the same step as the oracle: one chunked, vectorised pass per step equals the oracle's multi-pass
step to FP64 rounding (FMA contraction allowed), for the literal cpu_simd.jl scheme and for the
GPU benchmark's PPM + HLLD scheme, with several chunkings.  CPU only."""
import ctypes as C
import math
import os

import numpy as np
import pytest

from oracle import BUILD, Scheme

G = 5.0 / 3.0


def cs_lib():
    lib = C.CDLL(os.path.join(BUILD, "libcpustream.so"))
    lib.cs_run.restype = C.c_double
    lib.cs_run.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int] + [C.c_double] * 7
    return lib


def smooth(orc, n, seed=3):
    from test_oracle_physics import smooth as sm
    return sm(orc, n, n, n, seed)


@pytest.mark.parametrize("scheme,nch", [(0, 0), (0, 3), (1, 0), (1, 5)])
def test_stream_step_equals_oracle_fp64(orc, scheme, nch):
    n = 16
    u = smooth(orc, n)
    dt, dx = 2e-3, 1.0 / n
    ch = 0.3 * dx / dt / 3 * 0.25
    fac = math.exp(-ch * dt / 0.18)
    pfl = 1e-6 * 1e-8 / G
    sc = Scheme(gamma=G, smallr=1e-6, pfl=pfl, dt=dt, dx=dx, ch=ch, glm_fac=fac, predictor="full",
                recon="plm" if scheme == 0 else "ppm", riemann="llf" if scheme == 0 else "hlld", slope="moncen",
                mh_revert=True, face_guard="none" if scheme == 0 else "iflo", update_floor=True)
    ref = u.copy()
    for _ in range(3):
        ref = orc.step(ref, sc)
    got = np.ascontiguousarray(u.copy())
    el = cs_lib().cs_run(got.ctypes.data, n, 3, 8, scheme, nch, G, 1e-6, pfl, dt, dx, ch, fac)
    assert el >= 0.0
    for v in range(9):
        err = np.abs(got[v] - ref[v]).sum() / np.abs(ref[v]).sum()
        assert err <= 1e-12, (v, err)


def test_stream_fp32_close_to_fp32_oracle(orc32):
    n = 16
    from oracle import Oracle
    u = smooth(Oracle("f64"), n).astype(np.float32)
    dt, dx = 2e-3, 1.0 / n
    ch = 0.3 * dx / dt / 3 * 0.25
    pfl = 1e-6 * 1e-8 / G
    sc = Scheme(gamma=G, smallr=1e-6, pfl=pfl, dt=dt, dx=dx, ch=ch, glm_fac=1.0, predictor="full", recon="plm",
                riemann="llf", mh_revert=True, face_guard="none", update_floor=True)
    ref = orc32.step(u.copy(), sc)
    got = np.ascontiguousarray(u.copy())
    assert cs_lib().cs_run(got.ctypes.data, n, 1, 4, 0, 0, G, 1e-6, pfl, dt, dx, ch, 1.0) >= 0.0
    for v in range(9):
        err = np.abs(got[v].astype(np.float64) - ref[v]).sum() / np.abs(ref[v]).sum()
        assert err <= 2e-6, (v, err)
