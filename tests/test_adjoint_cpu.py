"""CPU checks of the adjoint building blocks and of the checkpointing schedule (no GPU).

The hand-written reverse of every point function (glmmhdturb_amd/csrc/adjoint.hpp, compiled for
the host) is checked by the transpose identity <J v, w> = <v, J^T w>, with J v computed by the
same forward templates over dual numbers (oracle/dual.hpp) -- an independent derivative."""
import ctypes as C
import os
import subprocess

import pytest

from glmmhdturb_amd.adjoint import recompute_steps, split_point

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "glmmhdturb_amd", "_build", "libadjcheck.so")

NAMES = ["cons2prim", "prim2cons", "flux_x", "flux_y", "flux_z", "fast_speed", "half_slope_moncen",
         "half_slope_minmod", "ppm_edges", "llf", "llf_bound", "hll", "hlld", "riemann_hlld_x", "riemann_hlld_y",
         "riemann_hlld_z", "riemann_hll_y", "riemann_llf_z", "riemann_hlld_llf_switch", "faces_full_ppm",
         "faces_full_ppm_fallback", "faces_full_plm_minmod_clamp", "faces_oned_ppm", "faces_oned_plm",
         "hancock_moncen", "hancock_minmod", "epilogue_kick", "epilogue_floor"]


@pytest.fixture(scope="module")
def lib():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "native")], check=True)
    lib = C.CDLL(LIB)
    lib.adj_check.argtypes = [C.c_int, C.c_uint64, C.c_double, C.POINTER(C.c_double)]
    lib.adj_check.restype = C.c_int
    return lib


@pytest.mark.parametrize("which", range(len(NAMES)), ids=NAMES)
def test_point_adjoint_transpose_identity(lib, which):
    out = (C.c_double * 2)()
    nonzero = 0
    for bscale in (0.05, 0.7, 2.0):
        for seed in range(150):
            assert lib.adj_check(which, seed, bscale, out) == 0
            jvw, vjw = out[0], out[1]
            assert abs(jvw - vjw) <= 2e-11 * max(abs(jvw), abs(vjw), 1e-300), (which, seed, bscale, jvw, vjw)
            nonzero += abs(jvw) > 1e-8
    assert nonzero > 100      # the identity is not checked on zeros only


def test_unknown_function_is_rejected(lib):
    out = (C.c_double * 2)()
    assert lib.adj_check(999, 0, 1.0, out) == -1


def test_checkpoint_schedule():
    # enough snapshots: one forward sweep (each state computed once, the last step needs none)
    assert recompute_steps(20, 20) == 19
    # no snapshot: every state recomputed from the start
    assert recompute_steps(10, 0) == 45
    # binomial bound: with s snapshots and t sweeps, beta(s, t) = C(s+t, s) steps are reversible
    # at cost <= t * n (Griewank & Walther)
    import math
    for s in (1, 2, 3, 5):
        for t in (1, 2, 3):
            n = math.comb(s + t, s)
            assert recompute_steps(n, s) <= t * n
    assert 1 <= split_point(200, 16) < 200
    assert recompute_steps(200, 16) <= 3 * 200


@pytest.mark.parametrize("prec", [8, 4], ids=["f64", "f32"])
def test_spj_without_division(lib, prec):
    """The march's strong-pressure-jump test (glmmhd_turb.jl:926) compares max > 2 max(min, 1e-20)
    instead of dividing; it must agree with the correctly rounded quotient on every input,
    including ratios within a few ulps of 2 and minima below the 1e-20 floor."""
    lib.spj_check.argtypes = [C.c_long, C.c_uint64, C.c_int]
    lib.spj_check.restype = C.c_long
    assert lib.spj_check(400000, 7, prec) == 0
