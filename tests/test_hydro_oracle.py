"""Pins of the 5-variable hydro oracle (oracle/hydro_oracle.c, spike_25d.cu:196-512) against
closed forms, invariants and the exact Riemann solution (CPU only)."""
from __future__ import annotations

import math
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
from oracle import HydroScheme, Oracle  # noqa: E402

G = 1.4


@pytest.fixture(scope="module")
def orc():
    return Oracle("f64")


def cons_of(rho, vx, vy, vz, p, g=G):
    return np.array([rho, rho * vx, rho * vy, rho * vz, p / (g - 1) + 0.5 * rho * (vx * vx + vy * vy + vz * vz)])


def test_c2p_tocons_roundtrip_and_floors(orc):
    sc = HydroScheme(gamma=G)
    rng = np.random.default_rng(1)
    for _ in range(50):
        q = np.array([rng.uniform(0.1, 3), *rng.normal(0, 1, 3), rng.uniform(0.1, 3)])
        c = orc.hydro_tocons(q, sc)
        assert np.allclose(c, cons_of(*q), rtol=1e-15, atol=1e-15)
        assert np.allclose(orc.hydro_c2p(c, sc), q, rtol=1e-13, atol=1e-14)
    # density below smallr: rho and 1/rho use the floor
    q = orc.hydro_c2p(np.array([1e-9, 1e-9, 0.0, 0.0, 1.0]), sc)
    assert q[0] == 1e-6 and q[1] == pytest.approx(1e-3, rel=1e-15)
    # internal energy below the pressure floor
    q = orc.hydro_c2p(cons_of(1.0, 2.0, 0.0, 0.0, 1e-12) - np.array([0, 0, 0, 0, 1e-9]), sc)
    assert q[4] == 1e-6


def test_moncen_table_and_symmetry(orc):
    m = orc.hydro_moncen
    assert m(0.0, 1.0, 2.0) == 1.0            # linear data: the central difference
    assert m(0.0, 1.0, 1.2) == pytest.approx(0.4)   # 2 * the smaller one-sided difference
    assert m(0.0, 0.1, 2.0) == pytest.approx(0.2)
    assert m(0.0, 1.0, 0.0) == 0.0            # extremum
    assert m(0.0, 0.0, 1.0) == 0.0            # flat on one side
    rng = np.random.default_rng(2)
    for a, b, c in rng.integers(-4, 5, (400, 3)).astype(float):
        s = m(a, b, c)
        assert m(c, b, a) == -s                                   # reflection
        assert abs(s) <= 2 * min(abs(b - a), abs(c - b)) + 1e-15  # TVD bound
        assert abs(s) <= abs(0.5 * (c - a)) + 1e-15
        assert s * (c - a) >= 0


def test_flux_closed_form_and_rotation(orc):
    sc = HydroScheme(gamma=G)
    q = np.array([1.3, 0.4, -0.7, 0.2, 2.1])
    rho, u, v, w, p = q
    E = p / (G - 1) + 0.5 * rho * (u * u + v * v + w * w)
    f = orc.hydro_flux(q, 0, sc)
    assert np.allclose(f, [rho * u, rho * u * u + p, rho * u * v, rho * u * w, (E + p) * u], rtol=1e-15)
    # y flux of q equals the x flux of the state with (u, v, w) -> (v, w, u), components rotated back
    qr = np.array([rho, v, w, u, p])
    fx = orc.hydro_flux(qr, 0, sc)
    fy = orc.hydro_flux(q, 1, sc)
    assert np.allclose(fy, [fx[0], fx[3], fx[1], fx[2], fx[4]], rtol=1e-15)


def test_hanc1d_matches_primitive_jacobian(orc):
    """mh = m - (dtdx/2) A_d s with the textbook primitive Euler matrix A_d."""
    dtdx = 0.3
    sc = HydroScheme(gamma=G, dtdx=dtdx)
    rng = np.random.default_rng(3)
    for d in range(3):
        m = np.array([rng.uniform(0.5, 2), *rng.normal(0, 1, 3), rng.uniform(0.5, 2)])
        s = rng.normal(0, 0.1, 5)
        rho, p, un = m[0], m[4], m[1 + d]
        A = un * np.eye(5)
        A[0, 1 + d] = rho
        A[1 + d, 4] = 1 / rho
        A[4, 1 + d] = G * p
        assert np.allclose(orc.hydro_hanc1d(m, s, d, sc), m - 0.5 * dtdx * A @ s, rtol=1e-14, atol=1e-15)
        # a flat state does not move
        assert np.array_equal(orc.hydro_hanc1d(m, np.zeros(5), d, sc), m)


def test_edge_and_hll(orc):
    sc = HydroScheme(gamma=G)
    b = np.array([1.0, 0.1, 0.2, 0.3, 1.0])
    s = np.array([0.2, 0.1, 0.0, -0.1, 0.4])
    assert np.array_equal(orc.hydro_edge(b, s, 1.0, sc), b + 0.5 * s)
    e = orc.hydro_edge(np.array([1e-7, 0, 0, 0, 1e-7]), np.zeros(5), -1.0, sc)
    assert e[0] == 1e-6 and e[4] == 1e-6
    rng = np.random.default_rng(4)
    for d in range(3):
        q = np.array([rng.uniform(0.5, 2), *rng.normal(0, 0.3, 3), rng.uniform(0.5, 2)])
        # consistency F(U, U) = F(U)
        assert np.allclose(orc.hydro_hll(q, q, d, sc), orc.hydro_flux(q, d, sc), rtol=1e-14, atol=1e-15)
        # supersonic to the right / left: pure upwinding
        L, R_ = q.copy(), q.copy()
        L[1 + d] = R_[1 + d] = 5.0
        R_[0] *= 1.1
        assert np.array_equal(orc.hydro_hll(L, R_, d, sc), orc.hydro_flux(L, d, sc))
        L[1 + d] = R_[1 + d] = -5.0
        assert np.array_equal(orc.hydro_hll(L, R_, d, sc), orc.hydro_flux(R_, d, sc))
        # subsonic: the HLL average between the two fluxes
        L, R_ = q.copy(), q.copy()
        R_[0] *= 1.3
        R_[4] *= 0.7
        aL, aR = math.sqrt(G * L[4] / L[0]), math.sqrt(G * R_[4] / R_[0])
        sL, sR = min(L[1 + d] - aL, R_[1 + d] - aR), max(L[1 + d] + aL, R_[1 + d] + aR)
        assert sL < 0 < sR
        UL, UR = orc.hydro_tocons(L, sc), orc.hydro_tocons(R_, sc)
        FL, FR = orc.hydro_flux(L, d, sc), orc.hydro_flux(R_, d, sc)
        f = orc.hydro_hll(L, R_, d, sc)
        # the HLL flux satisfies the integral conservation law over the fan:
        # sR UR - sL UL - (FR - FL) = (sR - sL) U_hll and F_hll = FL + sL (U_hll - UL)
        Uh = (sR * UR - sL * UL - (FR - FL)) / (sR - sL)
        assert np.allclose(f, FL + sL * (Uh - UL), rtol=1e-13, atol=1e-14)


def test_riemann_fb_switch(orc):
    sc = HydroScheme(gamma=G, fallback=True)
    L = np.array([1.0, 0.1, 0.0, 0.0, 1.0])
    R_ = np.array([0.5, -0.1, 0.0, 0.0, 0.4])
    assert np.array_equal(orc.hydro_riemann_fb(L, R_, 0, sc), orc.hydro_hll(L, R_, 0, sc))
    R_[0] = 5e-4   # below the 1e-3 fallback floor: Rusanov with the larger |u| + a
    f = orc.hydro_riemann_fb(L, R_, 0, sc)
    smax = max(abs(L[1]) + math.sqrt(G * L[4] / L[0]), abs(R_[1]) + math.sqrt(G * R_[4] / R_[0]))
    FL, FR = orc.hydro_flux(L, 0, sc), orc.hydro_flux(R_, 0, sc)
    UL, UR = orc.hydro_tocons(L, sc), orc.hydro_tocons(R_, sc)
    assert np.allclose(f, 0.5 * (FL + FR) - 0.5 * smax * (UR - UL), rtol=1e-14)
    assert not np.allclose(f, orc.hydro_hll(L, R_, 0, sc), rtol=1e-6)


def smooth_hydro(nx, ny, nz, seed, amp=0.2):
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(np.arange(nz) / nz, np.arange(ny) / ny, np.arange(nx) / nx, indexing="ij")
    ph = rng.uniform(0, 2 * np.pi, 8)
    rho = 1 + amp * np.sin(2 * np.pi * x + ph[0]) * np.cos(2 * np.pi * z + ph[1])
    vx = 0.5 * np.sin(2 * np.pi * y + ph[2])
    vy = 0.4 * np.cos(2 * np.pi * z + ph[3])
    vz = 0.3 * np.sin(2 * np.pi * x + ph[4])
    p = 1 + amp * np.cos(2 * np.pi * y + ph[5]) * np.sin(2 * np.pi * x + ph[6])
    return cons_of(rho, vx, vy, vz, p)


def test_step_conservation_uniform_and_shift(orc):
    sc = HydroScheme(gamma=G, dtdx=0.15)
    u = smooth_hydro(12, 10, 8, 5)
    u1 = orc.hydro_step(u, sc)
    for v in range(5):
        assert abs(u1[v].sum() - u[v].sum()) <= 1e-13 * np.abs(u[v]).sum()
    # integer periodic shifts commute with the step bit for bit
    for ax in (1, 2, 3):
        assert np.array_equal(orc.hydro_step(np.roll(u, 3, axis=ax), sc), np.roll(u1, 3, axis=ax))
    # a uniform state stays uniform (to the rounding of toCons(prim(U)))
    uu = np.broadcast_to(cons_of(1.2, 0.3, -0.2, 0.1, 0.8)[:, None, None, None], (5, 4, 4, 4)).copy()
    assert np.allclose(orc.hydro_step(uu, sc), uu, rtol=1e-15, atol=1e-16)


def test_step_axis_permutation(orc):
    """x -> y -> z relabelling of the grid and velocity gives the relabelled result."""
    sc = HydroScheme(gamma=G, dtdx=0.15)
    u = smooth_hydro(10, 8, 6, 7)
    u1 = orc.hydro_step(u, sc)
    # new axes (z', y', x') = (y, x, z): the old x axis becomes y', y becomes z', z becomes x'
    def perm(a):
        b = np.transpose(a, (0, 2, 3, 1)).copy()     # (v, y, x, z) -> arrays indexed [z'=y][y'=x][x'=z]
        return b[[0, 3, 1, 2, 4]]                    # (mx', my', mz') = (mz, mx, my)
    assert np.allclose(orc.hydro_step(perm(u), sc), perm(u1), rtol=1e-13, atol=1e-14)


def exact_riemann(rl, ul, pl, rr, ur, pr, xi, g=G):
    """Exact Euler Riemann solution sampled at x/t = xi (Toro, Riemann Solvers, ch. 4)."""
    cl, cr = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)

    def fk(p, rk, pk, ck):
        if p > pk:
            A, B = 2 / ((g + 1) * rk), (g - 1) / (g + 1) * pk
            return (p - pk) * math.sqrt(A / (p + B)), math.sqrt(A / (B + p)) * (1 - (p - pk) / (2 * (B + p)))
        return (2 * ck / (g - 1) * ((p / pk) ** ((g - 1) / (2 * g)) - 1),
                (p / pk) ** (-(g + 1) / (2 * g)) / (rk * ck))

    p = 0.5 * (pl + pr)
    for _ in range(100):
        f1, d1 = fk(p, rl, pl, cl)
        f2, d2 = fk(p, rr, pr, cr)
        dp = (f1 + f2 + ur - ul) / (d1 + d2)
        p = max(p - dp, 1e-12)
        if abs(dp) < 1e-14 * p:
            break
    us = 0.5 * (ul + ur) + 0.5 * (fk(p, rr, pr, cr)[0] - fk(p, rl, pl, cl)[0])
    out = np.empty((xi.size, 3))
    for n, s in enumerate(xi):
        if s < us:
            rk, uk, pk, ck, sg = rl, ul, pl, cl, -1
        else:
            rk, uk, pk, ck, sg = rr, ur, pr, cr, 1
        if p > pk:   # shock
            S = uk + sg * ck * math.sqrt((g + 1) / (2 * g) * p / pk + (g - 1) / (2 * g))
            rs = rk * ((p / pk + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * p / pk + 1))
            out[n] = (rk, uk, pk) if sg * (s - S) > 0 else (rs, us, p)
        else:        # rarefaction
            cs = ck * (p / pk) ** ((g - 1) / (2 * g))
            head, tail = uk + sg * ck, us + sg * cs
            if sg * (s - head) > 0:
                out[n] = (rk, uk, pk)
            elif sg * (s - tail) < 0:
                out[n] = (rk * (p / pk) ** (1 / g), us, p)
            else:
                c = 2 / (g + 1) * (ck - sg * (g - 1) / 2 * (uk - s))
                uu = 2 / (g + 1) * (-sg * ck + (g - 1) / 2 * uk + s)
                out[n] = (rk * (c / ck) ** (2 / (g - 1)), uu, pk * (c / ck) ** (2 * g / (g - 1)))
    return out


def sod_run(orc, n, axis, t_end=0.2, dtdx=0.2):
    """Sod tube on a periodic [0, 2) line along `axis`; left state on [0.5, 1.5)."""
    dx = 2.0 / n
    x = (np.arange(n) + 0.5) * dx
    left = (x >= 0.5) & (x < 1.5)
    rho = np.where(left, 1.0, 0.125)
    p = np.where(left, 1.0, 0.1)
    z = np.zeros(n)
    vel = [z, z, z]
    shape = [1, 1, 1]
    shape[2 - axis] = n
    u = cons_of(rho, *vel, p).reshape(5, *shape)
    steps = int(round(t_end / (dtdx * dx)))
    sc = HydroScheme(gamma=G, dtdx=dtdx)
    for _ in range(steps):
        u = orc.hydro_step(u, sc)
    u = u.reshape(5, n)
    t = steps * dtdx * dx
    sel = (x > 1.0) & (x < 1.95)
    ex = exact_riemann(1.0, 0.0, 1.0, 0.125, 0.0, 0.1, (x[sel] - 1.5) / t)
    err = np.abs(u[0, sel] - ex[:, 0]).mean()
    vel_other = [u[1 + c] for c in range(3) if c != axis]
    return err, u[1 + axis][sel] / u[0][sel], ex[:, 1], max(np.abs(v).max() for v in vel_other)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_sod_tube_against_exact_solution(orc, axis):
    e1, *_ = sod_run(orc, 200, axis)
    e2, vel, vex, other = sod_run(orc, 400, axis)
    assert e1 < 0.007 and e2 < 0.6 * e1, (e1, e2)
    assert np.abs(vel - vex).mean() < 0.02
    assert other == 0.0   # no transverse momentum appears


def test_sod_is_the_same_along_every_axis(orc):
    r = [sod_run(orc, 200, a)[0] for a in range(3)]
    assert r[0] == r[1] == r[2]


def sound_wave_error(orc, n, axis, eps=1e-4, rho0=2.0, p0=1.0, u0=0.5):
    """L1 density error of a right-going linear sound wave on a background moving at u0, after
    one crossing of the periodic unit line (exact: the initial state)."""
    c0 = math.sqrt(G * p0 / rho0)
    steps = int(2.5 * n)
    dtdx = n / ((u0 + c0) * steps)
    x = (np.arange(n) + 0.5) / n
    ph = eps * np.sin(2 * np.pi * x)
    rho, un, p = rho0 * (1 + ph), u0 + c0 * ph, p0 * (1 + G * ph)
    vel = [np.zeros(n)] * 3
    vel[axis] = un
    shape = [1, 1, 1]
    shape[2 - axis] = n
    u0_ = cons_of(rho, *vel, p).reshape(5, *shape)
    u = u0_
    sc = HydroScheme(gamma=G, dtdx=dtdx)
    for _ in range(steps):
        u = orc.hydro_step(u, sc)
    return np.abs(u[0] - u0_[0]).mean() / (rho0 * eps)


@pytest.mark.parametrize("axis", [0, 2])
def test_sound_wave_second_order(orc, axis):
    """The Hancock predictor makes the scheme second order in time and space: halving dx and
    dt cuts the error of a smooth acoustic wave by well over 2 (first-order terms would not)."""
    e32, e64 = sound_wave_error(orc, 32, axis), sound_wave_error(orc, 64, axis)
    assert e32 < 0.05 and e64 < 0.33 * e32, (e32, e64)


def with_scalars(u, s0, s1):
    return np.concatenate([u, (u[0] * s0)[None], (u[0] * s1)[None]])


@pytest.mark.parametrize("cma", [False, True])
def test_scalars_constant_fraction_and_conservation(orc, cma):
    """A uniform mass fraction stays uniform in any flow, the scalars are conserved, and the
    hydro variables do not see them (spike_25d.cu -DSCALARS / -DCMA)."""
    sc = HydroScheme(gamma=G, dtdx=0.15, nsp=2, cma=cma)
    u = smooth_hydro(12, 10, 8, 9)
    us = with_scalars(u, 0.3, 0.7)
    for _ in range(3):
        us = orc.hydro_step(us, sc)
    assert np.allclose(us[5] / us[0], 0.3, rtol=1e-13) and np.allclose(us[6] / us[0], 0.7, rtol=1e-13)
    ref = u
    for _ in range(3):
        ref = orc.hydro_step(ref, HydroScheme(gamma=G, dtdx=0.15))
    assert np.array_equal(us[:5], ref)


@pytest.mark.parametrize("cma", [False, True])
def test_scalars_sum_and_advection(orc, cma):
    """Fractions X and 1 - X keep summing to one; a scalar profile carried by a uniform flow comes
    back after one crossing with second-order convergence."""
    sc = HydroScheme(gamma=G, dtdx=0.15, nsp=2, cma=cma)
    u = smooth_hydro(12, 10, 8, 10)
    rng = np.random.default_rng(11)
    x0 = 0.5 + 0.3 * np.tanh(rng.normal(0, 1, u[0].shape))
    us = with_scalars(u, x0, 1 - x0)
    tot0 = us[5:].sum(axis=(1, 2, 3))
    for _ in range(4):
        us = orc.hydro_step(us, sc)
    assert np.allclose((us[5] + us[6]) / us[0], 1.0, rtol=1e-13)
    assert np.allclose(us[5:].sum(axis=(1, 2, 3)), tot0, rtol=1e-13)

    def err(n):
        x = (np.arange(n) + 0.5) / n
        prof = 0.5 + 0.2 * np.sin(2 * np.pi * x)
        line = cons_of(np.ones(n), np.full(n, 0.5), np.zeros(n), np.zeros(n), np.ones(n))
        v = with_scalars(line, prof, 1 - prof).reshape(7, 1, 1, n)
        steps = int(2.5 * n)
        s2 = HydroScheme(gamma=G, dtdx=n * 2.0 / steps, nsp=2, cma=cma)   # 0.5 * t = 1 crossing
        w = v
        for _ in range(steps):
            w = orc.hydro_step(w, s2)
        return np.abs(w[5] / w[0] - v[5] / v[0]).mean()
    e32, e64 = err(32), err(64)
    assert e32 < 0.02 and e64 < 0.33 * e32, (e32, e64)


def test_cma_flux_rides_the_mass_flux(orc):
    L = np.array([1.0, 0.2, 0.0, 0.0, 1.0, 0.3, 0.6])
    R_ = np.array([0.8, 0.1, 0.0, 0.0, 0.9, 0.9, 0.05])
    f = orc.hydro_hll(L, R_, 0, HydroScheme(gamma=G, nsp=2, cma=True))
    assert f[0] > 0 and f[5] == f[0] * L[5] and f[6] == f[0] * L[6]
    g = orc.hydro_hll(L, R_, 0, HydroScheme(gamma=G, nsp=2))
    assert np.array_equal(f[:5], g[:5]) and not np.isclose(f[5], g[5])
