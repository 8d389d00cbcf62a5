"""Passive species in the CT oracle (o_ct_step_sp: spike_ctm.cu -DNSP, the fraction X = rho X / rho
riding the mass flux), pinned by what the scheme's definition fixes.  CPU only.

* The species do not feed back: the 8 CT planes equal o_ct_step's bit for bit.
* Every rho X is conserved on the periodic grid.
* A uniform fraction stays uniform, and two complementary fractions (X2 = 1 - X1) keep summing to
  one: the limiters, the PPM edges and the upwind choice are odd-symmetric.
* In a uniform flow along any axis the species step is MUSCL-Hancock for linear advection with the
  normal-only predictor, written in numpy from its textbook form (test_oracle_species), for u > 0
  and u < 0, PLM.
* Where the update floors the density, rho X = smallr X_old.
"""
import numpy as np
import pytest

from oracle import CTScheme
from test_oracle_physics import smooth as smooth_state
from test_oracle_species import fraction_field, muscl_hancock_advection

G = 5.0 / 3.0
NX, NY, NZ = 12, 10, 8


def ct_state(orc, nx, ny, nz, seed=11, b0=0.5):
    """Smooth hydro state with divergence-free face B = b0 x + curl A (A on edges)."""
    u = smooth_state(orc, nx, ny, nz, seed=seed)[:8].copy()
    x, y, z = np.arange(nx) / nx, np.arange(ny) / ny, np.arange(nz) / nz
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    tp = 2 * np.pi
    ax = 0.05 * np.sin(tp * (Y + Z)); ay = 0.05 * np.cos(tp * (X + 2 * Z)); az = 0.05 * np.sin(tp * (X - Y))
    dx = 1.0 / nx
    rl = lambda a, ax_: np.roll(a, -1, axis=ax_)  # noqa: E731  (index i+1)
    u[5] = b0 + ((rl(az, 1) - az) - (rl(ay, 0) - ay)) / dx
    u[6] = ((rl(ax, 0) - ax) - (rl(az, 2) - az)) / dx
    u[7] = ((rl(ay, 2) - ay) - (rl(ax, 1) - ax)) / dx
    return u


def with_species(u8, x1):
    rho = u8[0]
    return np.concatenate([u8, (rho * x1)[None], (rho * (1.0 - x1))[None]])


def ct_scheme(orc, u8, **kw):
    dx = 1.0 / NX
    dt = 0.3 * dx / orc.ct_ctot_max(u8, G, 1e-6, 1e-6)
    base = dict(gamma=G, smallr=1e-6, smallp=1e-6, dt=dt, dx=dx)
    base.update(kw)
    return CTScheme(**base)


SCHEMES = [dict(riemann=m, recon=r, face_slope=fs) for m in ("hll", "hlld", "llf") for r in ("plm", "ppm_spike")
           for fs in (0, 1)]


@pytest.mark.parametrize("kw", SCHEMES, ids=lambda k: "-".join(str(v) for v in k.values()))
def test_ct_species_consistency(orc, kw):
    u8 = ct_state(orc, NX, NY, NZ)
    x1 = fraction_field(NZ, NY, NX)
    u = with_species(u8, x1)
    sc = ct_scheme(orc, u8, **kw)
    a, b = u.copy(), u8.copy()
    for _ in range(4):
        a = orc.ct_step_sp(a, sc, 2)
        b = orc.ct_step(b, sc)
    assert np.array_equal(a[:8], b)                           # no feedback
    for v in (8, 9):                                          # every rho X is conserved (periodic)
        assert abs(a[v].sum() - u[v].sum()) <= 1e-13 * np.abs(u[v]).sum()
    s = (a[8] + a[9]) / a[0]
    assert np.max(np.abs(s - 1.0)) <= 1e-13                   # complementary fractions
    xu = with_species(u8, np.full_like(x1, 0.3))              # uniform fraction
    for _ in range(4):
        xu = orc.ct_step_sp(xu, sc, 2)
    assert np.max(np.abs(xu[8] / xu[0] - 0.3)) <= 1e-14
    assert np.max(np.abs(a[8] / a[0] - x1)) > 1e-3            # and the fraction did move


@pytest.mark.parametrize("u", [0.5, -0.7])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_ct_species_is_muscl_hancock_in_a_uniform_flow(orc, u, axis):
    n = 48
    shape = [4, 4, 4]
    shape[2 - axis] = n                                       # planes are (z, y, x)
    nz, ny, nx = shape
    q = np.zeros((8, nz, ny, nx))
    q[0] = 1.0
    q[1 + axis] = u                                           # momentum = rho u with rho = 1
    q[5] = 0.2                                                # uniform face Bx
    q[4] = 1.0 / (G - 1.0) + 0.5 * u * u + 0.5 * 0.2 ** 2
    xf = (np.arange(n) + 0.5) / n
    x1 = 0.5 + 0.4 * np.sin(2 * np.pi * xf)
    x1[20:30] = 0.05
    prof = x1.reshape([n if d == 2 - axis else 1 for d in range(3)]) * np.ones((nz, ny, nx))
    uu = with_species(q, prof)
    dx = 1.0 / n
    dt = 0.4 * dx / (abs(u) + 2.0)
    sc = CTScheme(gamma=G, smallr=1e-6, smallp=1e-6, dt=dt, dx=dx, riemann="hll", recon="plm", face_slope=0)
    for _ in range(30):
        uu = orc.ct_step_sp(uu, sc, 2)
    got = np.moveaxis(uu[8] / uu[0], 2 - axis, -1)[0, 0]
    want = muscl_hancock_advection(x1, u * dt / dx, 30)
    assert np.max(np.abs(got - want)) <= 1e-13
    assert np.max(np.abs(want - x1)) > 0.05


def test_ct_floor_keeps_the_fraction(orc):
    """A near-void pocket (rho just above smallr) with a diverging x flow: the update floors its
    middle cells, and there every rho X becomes smallr X_old."""
    n = 16
    u8 = ct_state(orc, n, n, n, b0=0.2)
    zz, yy, xx = np.meshgrid(*(np.arange(n),) * 3, indexing="ij")
    pocket = (abs(xx - 7.5) < 3) & (abs(yy - 7.5) < 3) & (abs(zz - 7.5) < 3)
    rl = lambda a, ax_: np.roll(a, -1, axis=ax_)  # noqa: E731
    bc = [0.5 * (u8[5] + rl(u8[5], 2)), 0.5 * (u8[6] + rl(u8[6], 1)), 0.5 * (u8[7] + rl(u8[7], 0))]
    rho = np.where(pocket, 0.0101, 1.0)
    vx = np.where(pocket, 2.0 * np.sign(xx - 7.5), 0.1)
    pr = np.where(pocket, 0.01, 1.0)
    u8[0], u8[1], u8[2], u8[3] = rho, rho * vx, 0.0, 0.0
    u8[4] = pr / (G - 1.0) + 0.5 * rho * vx * vx + 0.5 * sum(c * c for c in bc)
    x1 = fraction_field(n, n, n)
    u = with_species(u8, x1)
    smallr = 0.01
    dx = 1.0 / n
    dt = 0.3 * dx / orc.ct_ctot_max(u8, G, smallr, 1e-6)
    sc = CTScheme(gamma=G, smallr=smallr, smallp=1e-6, dt=dt, dx=dx, riemann="hlld", recon="plm", update_floor=1)
    a = orc.ct_step_sp(u, sc, 2)
    fl = a[0] == smallr
    assert fl.sum() > 0
    xold = u[8:10] / np.maximum(u[0], smallr)
    for s_ in range(2):
        assert np.array_equal(a[8 + s_][fl], smallr * xold[s_][fl])
    assert np.array_equal(a[:8], orc.ct_step(u8, sc))
