"""Pins of the oracle's guarded branches and secondary solvers (no GPU).

Each test fixes a value by something other than the oracle's own formula: a function already
pinned by closed forms (the GLM flux, the fast speed, MonCen), a limiting case that reduces one
solver to another, a property the mathematics demands (monotone PPM parabola, physical flux of
the upwind state), or the reference's definition of a floor written as the expected state.
Each is chosen so that one plausible slip (a halved term, a dropped B^2, a wrong CW84 edge, a
floor that keeps momentum, a disabled revert, a wrong degenerate star) fails it.
"""
import math

import numpy as np
import pytest

from oracle import CTScheme, Scheme

G = 5.0 / 3.0
SMALLR = 1e-6
PFL = SMALLR * 1e-8 / G


def rand_prim(rng, bn=None):
    q = np.array([rng.uniform(0.5, 2.0), *rng.normal(0, 0.6, 3), rng.uniform(0.3, 2.0), *rng.normal(0, 0.8, 3),
                  rng.normal(0, 0.1)])
    if bn is not None:
        q[5] = bn
    return q


def glm_dir_flux(orc, q, d):
    return orc.rot_flux_from(orc.phys_flux_x(orc.rot_to(q, d), G), d)


# ----------------------------------------------------------------------------------------------
# CT flux (ct_mhd.jl:39-55) against the pinned GLM flux, and the CT step's 1-D reduction
# ----------------------------------------------------------------------------------------------
def ct_dflux(orc, q8, d):
    import ctypes as C
    q = orc.arr(q8, 8)
    f = orc.vec(8)
    orc.lib.o_ct_dflux(q.ctypes.data_as(C.c_void_p), d, G, f.ctypes.data_as(C.c_void_p))
    return f


def test_ct_flux_is_glm_flux_with_zero_normal_field(orc):
    """ct_mhd.jl's dflux is the ideal-MHD flux in direction d with F[Bn] = 0: component by
    component equal to the GLM directional flux (itself pinned by the flux Jacobian's
    eigenvalues and the textbook x-flux), for every direction."""
    rng = np.random.default_rng(11)
    for _ in range(40):
        q9 = rand_prim(rng)
        q9[8] = 0.0
        for d in range(3):
            f8 = ct_dflux(orc, q9[:8], d)
            f9 = glm_dir_flux(orc, q9, d)
            assert f8[5 + d] == 0.0
            assert np.allclose(f8, f9[:8], rtol=1e-14, atol=1e-14), (d, f8 - f9[:8])


@pytest.mark.parametrize("riem", ["hll", "hlld", "llf"])
def test_ct_step_reduces_to_glm_step_on_x_only_data(orc, riem):
    """A state that varies only along x, with a uniform normal field and psi = 0: the y and z
    fluxes cancel and the GLM pair leaves Bx and psi alone (Bn* = Bx, psi* = 0), so the hydro
    planes of one CT step (normal-only Hancock, PLM MonCen, clamped faces, the CT flux) equal those
    of the GLM step with the same scheme, and bx stays untouched."""
    n = 16
    x = (np.arange(n) + 0.5) / n
    tp = 2 * np.pi
    rho = 1 + 0.3 * np.sin(tp * x)
    prim = np.stack([rho, 0.4 * np.cos(tp * x), 0.2 * np.sin(2 * tp * x), -0.3 * np.cos(tp * x + 0.4),
                     1 + 0.2 * np.cos(tp * x), 0.7 + 0 * x, 0.5 * np.sin(tp * x + 1.0), 0.3 * np.cos(tp * x),
                     0 * x])
    u9 = np.zeros((9, 2, 3, n))
    for i in range(n):
        u9[:, :, :, i] = orc.prim2cons(prim[:, i], G)[:, None, None]
    dt, dx = 2e-3, 1.0 / n
    glm = orc.step(u9, Scheme(gamma=G, smallr=SMALLR, pfl=1e-6, dt=dt, dx=dx, ch=1.0, glm_fac=1.0,
                              predictor="oned", recon="plm", slope="moncen", riemann=riem, mh_revert=False,
                              face_guard="clamp", update_floor=False))
    ct = orc.ct_step(u9[:8].copy(), CTScheme(gamma=G, smallr=SMALLR, smallp=1e-6, dt=dt, dx=dx, face_slope=0,
                                             riemann=riem, recon="plm"))
    for v in range(5):
        assert np.allclose(ct[v], glm[v], rtol=1e-14, atol=1e-14 * np.abs(glm[v]).max()), (v, np.abs(ct[v] - glm[v]).max())
    assert np.array_equal(ct[5], u9[5])
    assert np.array_equal(glm[5], u9[5])
    assert not np.allclose(glm[1], u9[1], rtol=1e-6)   # the step moved momentum


# ----------------------------------------------------------------------------------------------
# LLF with the fast-speed upper bound (glmmhd_turb.jl:311-331)
# ----------------------------------------------------------------------------------------------
def test_llf_bound_speed(orc):
    """The bound is the fast speed with the normal field dropped, sqrt(a^2 + B^2/rho): with
    Bn* = 0 the LLF-bound flux equals the LLF flux; otherwise the two differ by
    (s_b - s)/2 * (U_R - U_L), one factor for all components, with s_b the speed built from
    fast_speed(q, bn = 0) and s the one built from fast_speed(q, Bn*)."""
    rng = np.random.default_rng(12)
    ch = 0.05
    for _ in range(30):
        L, R = rand_prim(rng), rand_prim(rng)
        L[5] = R[5] = 0.0
        L[8] = R[8] = 0.0
        f_b = orc.riemann(L, R, 0, "llf_bound", G, ch, SMALLR, PFL)
        f_l = orc.riemann(L, R, 0, "llf", G, ch, SMALLR, PFL)
        assert np.allclose(f_b, f_l, rtol=1e-13, atol=1e-13)
    for _ in range(30):
        L, R = rand_prim(rng), rand_prim(rng)
        R[5] = L[5]
        L[8] = R[8] = 0.0
        bn = L[5]
        f_b = orc.riemann(L, R, 0, "llf_bound", G, ch, SMALLR, PFL)
        f_l = orc.riemann(L, R, 0, "llf", G, ch, SMALLR, PFL)
        s = max(abs(L[1]) + orc.fast_speed(L, G, bn), abs(R[1]) + orc.fast_speed(R, G, bn), ch)
        s_b = max(abs(L[1]) + orc.fast_speed(L, G, 0.0), abs(R[1]) + orc.fast_speed(R, G, 0.0), ch)
        assert s_b >= s
        dU = orc.prim2cons(R, G) - orc.prim2cons(L, G)
        for v in (0, 1, 2, 3, 4, 6, 7):
            assert f_l[v] - f_b[v] == pytest.approx(0.5 * (s_b - s) * dU[v], rel=1e-9, abs=1e-12)


# ----------------------------------------------------------------------------------------------
# Spike PPM (spike_mhd.cu:164-172)
# ----------------------------------------------------------------------------------------------
SPIKE_CASES = [
    # (qm, q0, qp) -> (left edge, right edge), worked by hand from d = (qp - qm)/4,
    # cu = (qm - 2 q0 + qp)/6, qr = q0 + d - cu, ql = q0 - d - cu, then CW84
    ((0.0, 0.0, 1.0), (-1.0 / 6.0, 1.0 / 12.0)),      # step on the right: overshoot -> ql = 3 q0 - 2 qr
    ((1.0, 0.0, 0.0), (1.0 / 12.0, -1.0 / 6.0)),      # mirror: qr = 3 q0 - 2 ql
    ((0.0, 1.0, 1.0), (11.0 / 12.0, 7.0 / 6.0)),      # step on the left: qr = 3 q0 - 2 ql
    ((1.0, 2.0, 1.0), (2.0, 2.0)),                    # extremum: flat
    ((1.0, 1.5, 2.3), (1.125, 1.775)),              # smooth (d = 0.325, cu = 0.05): untouched
]


@pytest.mark.parametrize("stencil,edges", SPIKE_CASES)
def test_spike_ppm_case_table(orc, stencil, edges):
    qm, q0, qp = stencil
    assert orc.ppm1_spike(qm, q0, qp, -1.0) == pytest.approx(edges[0], rel=1e-14, abs=1e-15)
    assert orc.ppm1_spike(qm, q0, qp, 1.0) == pytest.approx(edges[1], rel=1e-14, abs=1e-15)


def test_spike_ppm_limited_parabola_is_monotone(orc):
    """CW84 (Colella & Woodward 1984, eq. 1.10): where an edge is reset, the parabola through the
    edges with cell mean q0 has its extremum exactly at the other edge, |a6| = |dq| with
    a6 = 6 (q0 - (ql + qr)/2), dq = qr - ql; elsewhere |a6| <= |dq| or the cell is flat."""
    rng = np.random.default_rng(13)
    reset = 0
    for _ in range(2000):
        qm, q0, qp = rng.normal(size=3)
        ql, qr = orc.ppm1_spike(qm, q0, qp, -1.0), orc.ppm1_spike(qm, q0, qp, 1.0)
        if ql == qr == q0:
            continue
        dq, a6 = qr - ql, 6 * (q0 - 0.5 * (ql + qr))
        assert abs(a6) <= abs(dq) * (1 + 1e-12) + 1e-14
        d, cu = 0.25 * (qp - qm), (qm - 2 * q0 + qp) / 6
        if not (ql == pytest.approx(q0 - d - cu, abs=1e-15) and qr == pytest.approx(q0 + d - cu, abs=1e-15)):
            reset += 1
            assert abs(a6) == pytest.approx(abs(dq), rel=1e-12, abs=1e-14)
    assert reset > 100


# ----------------------------------------------------------------------------------------------
# Update void floor (glmmhd_turb.jl:448-453)
# ----------------------------------------------------------------------------------------------
def test_update_void_floor_state(orc):
    """A cell whose updated density falls below smallr becomes (smallr, 0, 0, 0,
    pfl/(gamma-1) + B^2/2, B, psi * glm_fac): the magnetic field is kept, momentum is not."""
    sc = Scheme(gamma=G, smallr=SMALLR, pfl=3e-9, dt=1e-3, dx=0.1, glm_fac=0.9, update_floor=True)
    u = np.array([1e-5, 0.2, -0.1, 0.3, 2.0, 0.4, -0.6, 0.2, 0.05])
    z = np.zeros(9)
    flo = np.zeros(9); flo[0] = -1.0       # F_lo - F_hi = -1 on density: rho -> 1e-5 - 1e-2 < smallr
    r = orc.update_cell(u, flo, z, z, z, z, z, sc)
    emag = 0.5 * (0.4 ** 2 + 0.6 ** 2 + 0.2 ** 2)
    exp = np.array([SMALLR, 0, 0, 0, 3e-9 / (G - 1) + emag, 0.4, -0.6, 0.2, 0.05 * 0.9])
    assert np.allclose(r, exp, rtol=1e-15, atol=1e-300)
    r_off = orc.update_cell(u, flo, z, z, z, z, z, Scheme(**{**sc.__dict__, "update_floor": False}))
    assert r_off[0] < 0 and r_off[1] == 0.2


# ----------------------------------------------------------------------------------------------
# Predictor revert mh -> m0 (glmmhd_turb.jl:393-395)
# ----------------------------------------------------------------------------------------------
@pytest.mark.parametrize("pred", ["full", "oned"])
def test_predictor_revert_gives_plm_faces_about_m0(orc, pred):
    """A cell at the centre of a strong expansion: uniform rho, p, B and u = (-1, 0, 1) across it,
    so the MonCen half slope of u is 1/2 and the predicted density is rho (1 - dt/dx * 1/2) < 0
    for dt/dx = 3.  The predictor state (floored to smallr by cons2prim) reverts to m0, so the
    PLM faces are m0 -/+ s exactly; with the revert off they are built about the floored state."""
    nx = 5
    prim = np.zeros((9, 1, 1, nx))
    prim[0] = 1.0; prim[4] = 1.0; prim[5] = 0.3; prim[6] = 0.2
    prim[1, 0, 0, :] = [-2.0, -1.0, 0.0, 1.0, 2.0]
    sc = Scheme(gamma=G, smallr=SMALLR, dt=3.0, dx=1.0, predictor=pred, recon="plm", slope="moncen",
                mh_revert=True, face_guard="none")
    m0 = prim[:, 0, 0, 2]
    # the predictor really goes negative in x (checked with the pinned half-step functions)
    s = np.zeros(9); s[1] = 0.5 * orc.moncen(1.0, 1.0)
    if pred == "full":
        assert orc.hancock(m0, s, np.zeros(9), np.zeros(9), 3.0, G)[0] < 0
    else:
        assert orc.hanc1d(m0, s, 0, 3.0, G, SMALLR, sc.to_c().pfl)[0] == SMALLR   # floored by cons2prim
    fm, fp = orc.recon_faces(prim, 2, 0, 0, sc)
    assert np.array_equal(fm[0], m0 - s) and np.array_equal(fp[0], m0 + s)
    fm2, fp2 = orc.recon_faces(prim, 2, 0, 0, Scheme(**{**sc.__dict__, "mh_revert": False}))
    assert fm2[0][0] == SMALLR and fp2[0][0] == SMALLR


# ----------------------------------------------------------------------------------------------
# HLLD degenerate star (glmmhd_turb.jl:227-236)
# ----------------------------------------------------------------------------------------------
def test_hlld_degenerate_star_keeps_transverse_state(orc):
    """With Bn = 0 and a near-vacuum pair (rho ~ 1e-8, p ~ 1e-14, |B| ~ 1e-7) the star
    denominator rho (S - u)(S - SM) - Bn^2 is ~1e-13, below the 1e-12 guard, so the star keeps
    v_t and B_t of the outer state.  With Bn = 0 the inner fan collapses (S*_L = S_M = S*_R), so
    the flux is F(U_a) + S_a (U*_a - U_a) of the upwind side a, and its transverse-field
    components are exactly the physical ones, u_a B_t,a (for a star that rescales B_t they would
    carry an extra S_a (B*_t - B_t))."""
    L = np.array([1e-8, 1e-3, 2e-4, 0.0, 1e-14, 0.0, 1e-7, 5e-8, 0.0])
    R = np.array([1.2e-8, -5e-4, 0.0, 1e-4, 1.5e-14, 0.0, -8e-8, 1e-7, 0.0])
    orc.branch_hits(reset=True)
    f = orc.hlld_x(L, R, G, 0.0)
    assert orc.branch_hits()["hlld_degenerate"] == 2
    side = [(L[1] * L[6], L[1] * L[7]), (R[1] * R[6], R[1] * R[7])]
    assert any(f[6] == pytest.approx(a, rel=1e-12) and f[7] == pytest.approx(b, rel=1e-12) for a, b in side)
    # the same pair scaled to O(1) takes the regular star, whose B_t is compressed with rho
    orc.branch_hits(reset=True)
    Ls, Rs = L * 1.0, R * 1.0
    Ls[0] = Rs[0] = 1.0; Ls[4] = Rs[4] = 1.0; Ls[6:8] *= 1e7; Rs[6:8] *= 1e7
    fs = orc.hlld_x(Ls, Rs, G, 0.0)
    assert orc.branch_hits()["hlld_degenerate"] == 0
    assert np.all(np.isfinite(fs))
