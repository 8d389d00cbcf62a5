"""The select-free limiter forms the GPU kernels evaluate (csrc/physics.hpp, host-built in
tests/native/limiter_check.cpp) against the oracle's branch forms, which restate the reference:

* moncen_half(dl, dr) = 0.5 moncen(dl, dr) exactly (glmmhd_turb.jl:153-155);
* ppm_offsets, the CW84-limited parabola of glmmhd_turb.jl:917-923 in min/max form, and the spike
  ppm1 edges of spike_mhd.cu:164-172 in min/max form: equal to the branch forms up to rounding
  (a few units in the last place of the stencil's scale), also on inputs built to sit on each
  switch of the limiter (flat a b = 0, |a| = 2|b|, |b| = 2|a|), where the two forms may decide a
  switch differently and the limiter's continuity keeps them within rounding.

CPU only."""
import ctypes as C
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "glmmhdturb_amd", "_build", "liblimitercheck.so")
EPS = 2.0 ** -52


@pytest.fixture(scope="module")
def lc():
    if not os.path.exists(LIB):
        pytest.skip("build it with make -C tests/native")
    lib = C.CDLL(LIB)
    dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    lib.lc_moncen_half.argtypes = [dp, dp, dp, C.c_long]
    lib.lc_ppm_julia.argtypes = [dp, dp, dp, dp, dp, C.c_long]
    lib.lc_ppm_spike.argtypes = [dp, dp, dp, dp, dp, C.c_long]
    return lib


def random_triples(n, seed):
    rng = np.random.default_rng(seed)
    scale = 10.0 ** rng.uniform(-3, 3, n)
    q0 = scale * rng.uniform(-1, 1, n)
    return q0 + scale * rng.normal(0, 0.5, n), q0, q0 + scale * rng.normal(0, 0.5, n)


def from_slope_curve(q0, s, c, k):
    """(qm, q0, qp) with (qp - qm) = 4 s and (qm - 2 q0 + qp) = k c."""
    return q0 + 0.5 * k * c - 2.0 * s, q0, q0 + 0.5 * k * c + 2.0 * s


def switch_triples(n, seed, k):
    """Triples on the limiter's switches: for the offsets a = s - c', b = s + c' (Julia, c' = curve)
    or a = s + c', b = s - c' (spike), slopes at c' = 0, c' = +-s (flat: a b = 0) and c' = +-s/3
    (|a| = 2|b| or |b| = 2|a|), each nudged by a few ulps either way."""
    rng = np.random.default_rng(seed)
    q0 = 10.0 ** rng.uniform(-2, 2, n) * rng.choice([-1.0, 1.0], n)
    s = 10.0 ** rng.uniform(-3, 1, n) * rng.choice([-1.0, 1.0], n)
    ratio = rng.choice([0.0, 1.0, -1.0, 1.0 / 3.0, -1.0 / 3.0], n)
    nudge = 1.0 + rng.integers(-4, 5, n) * EPS
    return from_slope_curve(q0, s, s * ratio * nudge, k)


def close(a, b, qm, q0, qp, ulps=16):
    scale = np.maximum(np.maximum(np.abs(qm), np.abs(q0)), np.abs(qp))
    return np.abs(a - b) <= ulps * EPS * scale


def test_moncen_half_is_half_the_oracles_moncen(lc, orc):
    rng = np.random.default_rng(1)
    dl = rng.normal(0, 1, 20000) * 10.0 ** rng.uniform(-4, 4, 20000)
    dr = rng.normal(0, 1, 20000) * 10.0 ** rng.uniform(-4, 4, 20000)
    dl[:2000] = 0.0                                            # zero one-sided differences
    dr[2000:4000] = -2.0 * dl[2000:4000]                       # dl + dr = -dl: the centred slope's ties
    dr[4000:6000] = 0.5 * dl[4000:6000]
    out = np.empty_like(dl)
    lc.lc_moncen_half(dl, dr, out, dl.size)
    want = np.array([0.5 * orc.lib.o_moncen(a, b) for a, b in zip(dl, dr)])
    assert np.array_equal(out, want)


@pytest.mark.parametrize("kind", ["random", "switches"])
def test_julia_ppm_offsets_are_the_oracles_edges(lc, orc, kind):
    qm, q0, qp = random_triples(30000, 2) if kind == "random" else switch_triples(30000, 3, 12.0)
    lo, hi = np.empty_like(q0), np.empty_like(q0)
    lc.lc_ppm_julia(qm, q0, qp, lo, hi, q0.size)
    want = np.array([orc.ppm_edges(a, b, c) for a, b, c in zip(qm, q0, qp)])
    assert np.all(close(lo, want[:, 0], qm, q0, qp)) and np.all(close(hi, want[:, 1], qm, q0, qp))
    # the flat case is exact: both edges are q0
    flat = (want[:, 0] == q0) & (want[:, 1] == q0)
    assert flat.sum() > 100 and np.all(close(lo[flat], q0[flat], qm[flat], q0[flat], qp[flat], 2))


@pytest.mark.parametrize("kind", ["random", "switches"])
def test_spike_ppm1_edges_are_the_oracles(lc, orc, kind):
    qm, q0, qp = random_triples(30000, 4) if kind == "random" else switch_triples(30000, 5, 6.0)
    lo, hi = np.empty_like(q0), np.empty_like(q0)
    lc.lc_ppm_spike(qm, q0, qp, lo, hi, q0.size)
    wlo = np.array([orc.ppm1_spike(a, b, c, -1.0) for a, b, c in zip(qm, q0, qp)])
    whi = np.array([orc.ppm1_spike(a, b, c, 1.0) for a, b, c in zip(qm, q0, qp)])
    assert np.all(close(lo, wlo, qm, q0, qp)) and np.all(close(hi, whi, qm, q0, qp))
    # every branch of ppm1 is taken: flat, left reset, right reset, unlimited
    d = 0.25 * (qp - qm)
    cu = (qm - 2 * q0 + qp) / 6.0
    a, b = d + cu, d - cu
    assert (a * b <= 0).sum() > 100 and (np.abs(a) > 2 * np.abs(b) + 1e-9).sum() > 100
    assert (np.abs(b) > 2 * np.abs(a) + 1e-9).sum() > 100
