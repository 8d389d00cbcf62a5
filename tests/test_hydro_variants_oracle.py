"""The parabolic-edge (-DPPM1D, spike_25d.cu:467-483) and donor-cell (-DDONOR, :438-454) variants
of the hydro march in the oracle (HydroScheme.ppm / .donor), pinned by what their definitions fix.
CPU only.

* PPM1D: on primitives linear in x (exact steps) the parabola is the line, its edges are the PLM
  edges and no limiter fires, so away from the periodic wrap the step equals the light march's to
  rounding (the edges are summed in another order); mass, momentum and energy are conserved; a state
  at rest stays at rest.
* DONOR: the faces on the 32 x 8 tile seams are first order, so the step equals the light march's
  in every cell none of whose x or y faces is a seam, differs in cells next to a seam, and is
  exactly conservative; on a state of constant 2 x 2 x 2 blocks every limited slope vanishes, the
  second-order faces are first order too, and the two steps are equal bit for bit everywhere (which
  fixes the seam flux as the HLL flux of the two cell states).
"""
import numpy as np
import pytest

from oracle import HydroScheme
from test_hydro_compact_oracle import hydro_state


def linear_in_x(nx, ny, nz):
    """Primitives rho, vx, p linear in x with power-of-two steps (exact differences), vy = vz = 0."""
    i = np.arange(nx, dtype=np.float64)
    rho = 1.0 + i * 2.0 ** -6
    vx = 0.25 + i * 2.0 ** -7
    p = 2.0 + i * 2.0 ** -5
    shape = (nz, ny, nx)
    rho, vx, p = (np.broadcast_to(a, shape).copy() for a in (rho, vx, p))
    return np.stack([rho, vx, np.zeros(shape), np.zeros(shape), p])


def test_ppm1d_is_the_light_march_on_linear_data(orc):
    q = linear_in_x(40, 4, 4)
    a = orc.hydro_step_prim(q, HydroScheme(gamma=1.4, dtdx=0.05, ppm=True))
    b = orc.hydro_step_prim(q, HydroScheme(gamma=1.4, dtdx=0.05))
    inner = slice(3, 37)   # cells whose 5-cell stencils do not cross the periodic wrap
    scale = np.maximum(np.abs(b).max(axis=(1, 2, 3)), 1.0)[:, None, None, None]   # vy = vz = 0 stay 0
    assert np.max(np.abs(a[..., inner] - b[..., inner]) / scale) <= 1e-15
    assert np.max(np.abs(a[..., :3] - b[..., :3]) / scale) > 1e-6   # the wrap's jump is treated differently


@pytest.mark.parametrize("nsp,cma", [(0, False), (2, True), (2, False)])
def test_ppm1d_conserves_and_keeps_rest(orc, nsp, cma):
    n = 12
    u = hydro_state(n)
    if nsp:
        z, y, x = np.meshgrid(*(np.arange(n) / n,) * 3, indexing="ij")
        u = np.concatenate([u, (u[0] * (0.3 + 0.2 * np.sin(2 * np.pi * x)))[None], (u[0] * 0.5)[None]])
    sc = HydroScheme(gamma=1.4, dtdx=0.05, ppm=True, nsp=nsp, cma=cma)
    out = orc.hydro_step(u, sc)
    for v in range(5 + nsp):
        assert abs(out[v].sum() - u[v].sum()) <= 1e-12 * np.abs(u[v]).sum()
    rest = np.zeros((5 + nsp, n, n, n))
    rest[0], rest[4] = 1.0, 2.5
    if nsp:
        rest[5:] = 0.5
    assert np.array_equal(orc.hydro_step(rest, sc), rest)


def periodic_state(nx, ny, nz, seed=5):
    """A smooth state periodic on the box, with random phases so that no limiter sees an exact
    extremum: every face has a nonzero slope on at least one side."""
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(np.arange(nz) / nz, np.arange(ny) / ny, np.arange(nx) / nx, indexing="ij")
    def wave(a):
        px, py, pz = rng.uniform(0, 1, 3)
        return a * (np.sin(2 * np.pi * (x + px)) + np.sin(2 * np.pi * (y + py)) + np.sin(2 * np.pi * (z + pz)))
    rho = 1.0 + wave(0.1)
    v = [wave(0.05) for _ in range(3)]
    p = 1.0 + wave(0.08)
    e = p / 0.4 + 0.5 * rho * sum(c * c for c in v)
    return np.stack([rho, rho * v[0], rho * v[1], rho * v[2], e])


def test_donor_seams(orc):
    nx, ny, nz = 70, 20, 6       # 70 = 2 x 32 + 6: a partial last tile; the periodic face is a seam
    u = periodic_state(nx, ny, nz)
    a = orc.hydro_step(u, HydroScheme(gamma=1.4, dtdx=0.05, donor=True))
    b = orc.hydro_step(u, HydroScheme(gamma=1.4, dtdx=0.05))
    i, j = np.arange(nx), np.arange(ny)
    seam_x = (i % 32 == 0) | (((i + 1) % nx) % 32 == 0)
    seam_y = (j % 8 == 0) | (((j + 1) % ny) % 8 == 0)
    far = ~seam_x[None, None, :] & ~seam_y[None, :, None] & np.ones((nz, 1, 1), bool)
    assert np.array_equal(a[:, far], b[:, far])
    near = ~far
    assert np.all(np.abs(a[0][near] - b[0][near]) > 0)              # every seam cell takes the first-order flux
    for v in range(5):                                               # and the scheme stays conservative
        assert abs(a[v].sum() - u[v].sum()) <= 1e-12 * np.abs(u[v]).sum()


def test_donor_is_the_light_march_without_slopes(orc):
    nx, ny, nz = 70, 20, 6
    u = periodic_state(nx // 2, ny // 2, nz // 2)
    u = u.repeat(2, axis=1).repeat(2, axis=2).repeat(2, axis=3)   # every cell has an equal neighbour per axis
    a = orc.hydro_step(u, HydroScheme(gamma=1.4, dtdx=0.05, donor=True))
    b = orc.hydro_step(u, HydroScheme(gamma=1.4, dtdx=0.05))
    assert np.array_equal(a, b) and not np.array_equal(b, u)
