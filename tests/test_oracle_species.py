"""Passive species in the oracle (o_step_sp: spike_mhd.cu -DNSP, the fraction X = rho X / rho riding
the mass flux), pinned by what the scheme's definition fixes.  CPU only.

* The species do not feed back: the 9 MHD planes equal o_step's bit for bit.
* Every rho X is conserved on the periodic grid (a flux difference per face).
* A uniform fraction stays uniform (the species flux is the mass flux times the fraction).
* Two complementary fractions (X2 = 1 - X1) keep summing to one: every limiter, reconstruction and
  the upwind choice are odd-symmetric, so the scheme is consistent across species.
* In a uniform flow (every MHD flux the same) the species step is MUSCL-Hancock for linear
  advection, written here in numpy from its textbook form, for u > 0 and u < 0.
* Where the update floors the density, rho X = smallr X_old.
"""
import numpy as np
import pytest

from branch_states import KINDS, pocket_state
from oracle import Scheme
from test_oracle_physics import smooth as smooth_state

G = 5.0 / 3.0
NX, NY, NZ = 12, 10, 8


def with_species(orc, u, x1):
    """Append rho X1 and rho X2 = rho (1 - X1) to a 9-plane state."""
    rho = u[0]
    return np.concatenate([u, (rho * x1)[None], (rho * (1.0 - x1))[None]])


def fraction_field(nz, ny, nx, seed=5):
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(*(np.arange(n) / n for n in (nz, ny, nx)), indexing="ij")
    tp = 2 * np.pi
    f = 0.5 + 0.3 * np.sin(tp * x + rng.uniform(0, tp)) * np.cos(tp * y) + 0.1 * np.cos(tp * z)
    f[(x > 0.3) & (x < 0.55)] = 0.95   # a step: the limiters' switches
    return f


def scheme(orc, u, **kw):
    dt = 0.3 * (1.0 / NX) / orc.ctot_max(u[:9], G, 1e-6, 1e-6 * 1e-8 / G)
    base = dict(gamma=G, dt=dt, dx=1.0 / NX, ch=0.3 / NX / dt / 3 * 0.25, glm_fac=0.98)
    base.update(kw)
    return Scheme(**base)


SCHEMES = [dict(predictor=p, recon=r, riemann=m, slope=sl) for p in ("full", "oned") for r in ("plm", "ppm", "ppm_spike")
           for m in ("hll", "hlld", "llf") for sl in ("moncen",)] + [dict(predictor="full", recon="plm", riemann="hlld",
                                                                       slope="minmod")]


@pytest.mark.parametrize("kw", SCHEMES, ids=lambda k: "-".join(k.values()))
def test_species_consistency(orc, kw):
    u9 = smooth_state(orc, NX, NY, NZ)
    x1 = fraction_field(NZ, NY, NX)
    u = with_species(orc, u9, x1)
    sc = scheme(orc, u9, **kw)
    a, b = u.copy(), u9.copy()
    for _ in range(4):
        a = orc.step_sp(a, sc, 2)
        b = orc.step(b, sc)
    assert np.array_equal(a[:9], b)                           # no feedback
    for v in (9, 10):                                         # every rho X is conserved (periodic)
        assert abs(a[v].sum() - u[v].sum()) <= 1e-13 * np.abs(u[v]).sum()
    s = (a[9] + a[10]) / a[0]
    assert np.max(np.abs(s - 1.0)) <= 1e-13                   # complementary fractions
    xu = with_species(orc, u9, np.full_like(x1, 0.3))         # uniform fraction
    for _ in range(4):
        xu = orc.step_sp(xu, sc, 2)
    assert np.max(np.abs(xu[9] / xu[0] - 0.3)) <= 1e-14
    assert np.max(np.abs(a[9] / a[0] - x1)) > 1e-3            # and the fraction did move


def moncen(dl, dr):
    dc = 0.5 * (dl + dr)
    m = np.minimum(2 * np.minimum(np.abs(dl), np.abs(dr)), np.abs(dc))
    return np.where(dl * dr <= 0, 0.0, np.sign(dc) * m)


def muscl_hancock_advection(x, c, steps):
    """MUSCL-Hancock for X_t + u X_x = 0 at Courant number c = u dt/dx (periodic): MonCen half
    slopes hs, half step X - c hs, faces X_h -/+ hs, upwind flux."""
    for _ in range(steps):
        hs = 0.5 * moncen(x - np.roll(x, 1), np.roll(x, -1) - x)
        xh = x - c * hs
        hi, lo = xh + hs, xh - hs
        face = hi if c >= 0 else np.roll(lo, -1)             # X at face i+1/2
        x = x - c * (face - np.roll(face, 1))
    return x


@pytest.mark.parametrize("u", [0.5, -0.7])
@pytest.mark.parametrize("predictor", ["full", "oned"])
def test_species_is_muscl_hancock_in_a_uniform_flow(orc, u, predictor):
    n = 64
    xf = (np.arange(n) + 0.5) / n
    x1 = 0.5 + 0.4 * np.sin(2 * np.pi * xf)
    x1[20:30] = 0.05
    q = np.zeros((9, 1, 1, n))
    q[0], q[1], q[4], q[5] = 1.0, u, 1.0, 0.2
    u9 = np.empty_like(q)
    for i in range(n):
        u9[:, 0, 0, i] = orc.prim2cons(q[:, 0, 0, i], G)
    uu = with_species(orc, u9, x1.reshape(1, 1, n))
    dt = 0.4 / n / (abs(u) + 2.0)
    sc = Scheme(gamma=G, dt=dt, dx=1.0 / n, ch=1.0, glm_fac=1.0, predictor=predictor, recon="plm", riemann="hll")
    for _ in range(30):
        uu = orc.step_sp(uu, sc, 2)
    got = (uu[9] / uu[0]).reshape(n)
    want = muscl_hancock_advection(x1, u * dt * n, 30)
    assert np.max(np.abs(got - want)) <= 1e-13
    assert np.max(np.abs(want - x1)) > 0.05


def test_floor_keeps_the_fraction(orc):
    k = KINDS["floor"]
    u9 = pocket_state(orc, 24, 20, 16, "floor")
    x1 = fraction_field(16, 20, 24)
    u = with_species(orc, u9, x1)
    pfl = k["smallr"] * k["smallc"] ** 2 / G
    dt = 0.4 / 24 / orc.ctot_max(u9, G, k["smallr"], pfl)
    sc = Scheme(gamma=G, smallr=k["smallr"], smallc=k["smallc"], dt=dt, dx=1.0 / 24, ch=0.4 / 24 / dt / 12,
                glm_fac=0.99, predictor="full", recon="ppm", riemann="hlld")
    orc.branch_hits(reset=True)
    a = orc.step_sp(u, sc, 2)
    assert orc.branch_hits(reset=True)["update_floor"] > 0
    fl = a[0] == k["smallr"]
    assert fl.sum() > 0
    xold = u[9] / np.maximum(u[0], k["smallr"])
    assert np.allclose(a[9][fl], k["smallr"] * xold[fl], rtol=1e-15, atol=0)
