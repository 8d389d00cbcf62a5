"""The compact storage formats of the hydro march (spike_25d.cu -DF16STATE / -DU16STATE / -DU16SP,
:64-95) on the CPU: the uint16 codecs (:157-168) restated in numpy, and the oracle's two entry
points they need, the step from primitive planes (oh_step_prim) and the species held as log2 X
(HydroScheme.logsp).

* Every code round-trips exactly (encode(decode(u)) == u for all 65536 codes), decoding is
  monotonic, endpoints and clamping are the formulas', and a decoded value is within two float32
  roundings (at the scale of the range's ends) of the closed form a + u (b - a) / 65535.
* oh_step_prim is oh_step: from the primitive planes it equals the conserved step from the
  conserved state those primitives give, up to the rounding of that conversion.
* With log2 species the hydro planes are those of the run without species, every rho X is
  conserved, a uniform fraction stays uniform, and away from the jumps of fractions that are
  powers of two, piecewise constant in x (where log2 is exact and both slopes vanish), the step
  equals the linear-species step.
"""
import numpy as np
import pytest

from oracle import (HydroScheme, dec_log2, dec_log2s, dec_vel, enc_log2, enc_log2s, enc_vel)

U = np.arange(65536, dtype=np.uint16)


@pytest.mark.parametrize("dec,enc,lo,hi", [(dec_log2, enc_log2, -110.0, 0.0), (dec_vel, enc_vel, -8.0, 8.0),
                                           (dec_log2s, enc_log2s, -32.0, 32.0)])
def test_codecs_round_trip_and_closed_form(dec, enc, lo, hi):
    x = dec(U)
    assert x.dtype == np.float32
    assert np.array_equal(enc(x), U)
    assert np.all(np.diff(x.astype(np.float64)) > 0)
    assert x[0] == np.float32(lo)
    exact = lo + U.astype(np.float64) * (hi - lo) / 65535.0
    # one float32 rounding of the product and one of the sum, at the scale of the range's ends
    assert np.max(np.abs(x - exact)) <= 2.0 ** -23 * max(abs(lo), abs(hi))
    assert enc(np.float32(lo - 5.0)) == 0 and enc(np.float32(hi + 5.0)) == 65535
    # nearest code: a value a third of a step above code 1000 encodes to 1000, two thirds to 1001
    step = (hi - lo) / 65535.0
    assert enc(np.float32(x[1000] + step / 3)) == 1000 and enc(np.float32(x[1000] + 2 * step / 3)) == 1001


def hydro_state(n, seed=3):
    """The reference IC pattern (spike_25d.cu:530-545) with a smooth random modulation."""
    rng = np.random.default_rng(seed)
    i = np.arange(n ** 3)
    ph = (0.001 * (i % 911)).reshape(n, n, n)
    z, y, x = np.meshgrid(*(np.arange(n) / n,) * 3, indexing="ij")
    mod = 0.2 * np.sin(2 * np.pi * (x + rng.uniform())) * np.cos(2 * np.pi * y)
    rho = 1.0 + 0.3 * ph + mod
    m = [0.1 * ph + 0.1 * mod, -0.1 * ph, 0.05 * ph + 0.05 * np.cos(2 * np.pi * z)]
    e = 2.5 + 0.5 * ph + 0.3 * mod
    return np.stack([rho, m[0], m[1], m[2], e])


def test_step_from_primitive_planes_is_the_step(orc):
    sc = HydroScheme(gamma=1.4, dtdx=0.05)
    u = hydro_state(12)
    prim = np.stack([orc.hydro_c2p(u[:, k, j, i], sc) for k in range(12) for j in range(12) for i in range(12)],
                    axis=1).reshape(5, 12, 12, 12)
    a = orc.hydro_step_prim(prim, sc)
    b = orc.hydro_step(u, sc)
    assert np.max(np.abs(a - b) / np.abs(b).max(axis=(1, 2, 3))[:, None, None, None]) < 1e-14


def with_log_species(u, x1, x2):
    return np.concatenate([u, np.log2(x1)[None], np.log2(x2)[None]])


@pytest.mark.parametrize("fallback", [False, True])
def test_log2_species(orc, fallback):
    n = 12
    u = hydro_state(n)
    z, y, x = np.meshgrid(*(np.arange(n) / n,) * 3, indexing="ij")
    x1 = 0.3 + 0.2 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * z)
    x2 = 1e-20 * (1.0 + 0.5 * np.cos(2 * np.pi * y))          # a trace species, 20 dex down
    sc = HydroScheme(gamma=1.4, dtdx=0.05, nsp=2, cma=True, logsp=True, fallback=fallback, fb_dmin=1.05,
                     fb_pmin=1e-3)
    s0 = HydroScheme(gamma=1.4, dtdx=0.05, fallback=fallback, fb_dmin=1.05, fb_pmin=1e-3)
    q = with_log_species(u, x1, x2)
    out = orc.hydro_step(q, sc)                               # species out: conserved rho X
    assert np.array_equal(out[:5], orc.hydro_step(u, s0))     # no feedback
    rx_old = u[0] * np.stack([x1, x2])
    for v in (5, 6):                                          # conservation of rho X
        assert abs(out[v].sum() - rx_old[v - 5].sum()) <= 1e-13 * np.abs(rx_old[v - 5]).sum()
    assert np.all(out[6] > 0)                                 # the trace species stays positive
    qu = with_log_species(u, np.full_like(x1, 0.25), np.full_like(x1, 2.0 ** -60))
    ou = orc.hydro_step(qu, sc)
    assert np.max(np.abs(ou[5] / ou[0] - 0.25)) <= 1e-15 and np.max(np.abs(ou[6] / ou[0] / 2.0 ** -60 - 1)) <= 1e-14


def test_log2_species_equal_linear_species_where_log2_is_exact(orc):
    """Fractions that are powers of two, constant in x-slabs wider than the stencil: log2 X is
    exact, the slopes of log2 X and of X vanish together, and the edge values are the cell values,
    so the log2 and the linear reconstructions give the same fluxes."""
    n = 16
    u = hydro_state(n)
    z, y, x = np.meshgrid(*(np.arange(n),) * 3, indexing="ij")
    x1 = np.where(x < 8, 0.25, 0.5)
    x2 = np.where(x < 8, 2.0 ** -40, 2.0 ** -30)
    lin = HydroScheme(gamma=1.4, dtdx=0.05, nsp=2, cma=True)
    lg = HydroScheme(gamma=1.4, dtdx=0.05, nsp=2, cma=True, logsp=True)
    a = orc.hydro_step(np.concatenate([u, (u[0] * x1)[None], (u[0] * x2)[None]]), lin)
    b = orc.hydro_step(with_log_species(u, x1, x2), lg)
    assert np.array_equal(a[:5], b[:5])
    # away from the two jumps (x = 0|15 and 7|8, with the 2-cell stencil) the cells see no slope
    far = (x >= 3) & (x <= 4) | (x >= 11) & (x <= 12)
    assert np.allclose(a[5][far], b[5][far], rtol=1e-15, atol=0) and np.allclose(a[6][far], b[6][far], rtol=1e-15, atol=0)
