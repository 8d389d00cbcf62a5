// adjoint.hpp -- reverse-mode (adjoint) versions of the GLM-MHD point physics of physics.hpp.
//
// Every function f(x) -> y here has an adjoint f_adj(x, y_bar, x_bar) that ADDS J_f(x)^T y_bar to
// x_bar.  The forward values are recomputed from x inside; branches (limiters, PPM monotonicity,
// floors, the HLLD fan, max/min selections) take the branch the forward evaluation takes, so the
// adjoint is exact wherever the forward map is differentiable.  At a switch point the one-sided
// derivative of the selected branch is returned.  Reciprocals are differentiated as exact 1/x.
//
// Supported schemes: full or normal-only Hancock predictor, PLM (MonCen or minmod) or PPM (the
// Julia form, with its strong-pressure-jump fallback), LLF / HLL / HLLD (with the LLF switch) and
// the GLM pair, face guards none / iflo / clamp, the update floor, psi damping and the OU kick.
#pragma once
#include "physics.hpp"

namespace glm {

// sign used for d|x|/dx (+1 at 0, matching tcopysign(m, +0))
template <typename T>
GLM_HD T sgn1(T x) { return x >= T(0) ? T(1) : T(-1); }

// d sqrt(x) with sqrt(x) = y
template <typename T>
GLM_HD T dsqrt(T y) { return y > T(0) ? T(0.5) / y : T(0); }

// ---- cons2prim (glmmhd_turb.jl:68-77) ------------------------------------------------------
template <typename T>
GLM_HD void cons2prim_adj(const T c[9], const T qb[9], T cb[9], const Phys<T>& s)
{
    const T rho = tmax(c[0], s.smallr);
    const T ir = trcp(rho);
    const T gm1 = s.gamma - T(1);
    const T ekin = T(0.5) * (c[1] * c[1] + c[2] * c[2] + c[3] * c[3]) * ir;
    const T emag = T(0.5) * (c[5] * c[5] + c[6] * c[6] + c[7] * c[7]);
    const bool pact = gm1 * (c[4] - ekin - emag) > s.pfl;
    T rho_b = qb[0], ir_b = T(0);
    for (int i = 1; i <= 3; ++i) {
        cb[i] += qb[i] * ir;
        ir_b += qb[i] * c[i];
    }
    if (pact) {
        const T g = gm1 * qb[4];
        cb[4] += g;
        for (int i = 1; i <= 3; ++i) cb[i] += -g * c[i] * ir;
        ir_b += -g * T(0.5) * (c[1] * c[1] + c[2] * c[2] + c[3] * c[3]);
        for (int j = 5; j <= 7; ++j) cb[j] += -g * c[j];
    }
    rho_b += -ir * ir * ir_b;
    if (c[0] > s.smallr) cb[0] += rho_b;
    for (int j = 5; j <= 8; ++j) cb[j] += qb[j];
}

// ---- prim2cons (glmmhd_turb.jl:79-86) ------------------------------------------------------
template <typename T>
GLM_HD void prim2cons_adj(const T q[9], const T cb[9], T qb[9], const Phys<T>& s)
{
    const T v2 = q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    qb[0] += cb[0] + cb[1] * q[1] + cb[2] * q[2] + cb[3] * q[3] + cb[4] * T(0.5) * v2;
    for (int i = 1; i <= 3; ++i) qb[i] += cb[i] * q[0] + cb[4] * q[0] * q[i];
    qb[4] += cb[4] * s.igm1;
    for (int j = 5; j <= 7; ++j) qb[j] += cb[j] + cb[4] * q[j];
    qb[8] += cb[8];
}

// ---- flux_dir (glmmhd_turb.jl:98-115 in direction D) --------------------------------------
template <typename T, int D>
GLM_HD void flux_dir_adj(const T q[9], const T fb[9], T qb[9], const Phys<T>& s)
{
    constexpr int A = D % 3, B = (D + 1) % 3, C = (D + 2) % 3;
    const T rho = q[0], un = q[1 + A], u1 = q[1 + B], u2 = q[1 + C], p = q[4];
    const T bn = q[5 + A], b1 = q[5 + B], b2_ = q[5 + C];
    const T b2 = bn * bn + b1 * b1 + b2_ * b2_;
    const T ptot = p + T(0.5) * b2;
    const T u2s = un * un + u1 * u1 + u2 * u2;
    const T E = p * s.igm1 + T(0.5) * rho * u2s + T(0.5) * b2;
    const T vdotb = un * bn + u1 * b1 + u2 * b2_;
    const T g0 = fb[0], gA = fb[1 + A], gB = fb[1 + B], gC = fb[1 + C], g4 = fb[4];
    const T hB = fb[5 + B], hC = fb[5 + C];
    qb[0] += g0 * un + gA * un * un + gB * un * u1 + gC * un * u2 + g4 * un * T(0.5) * u2s;
    qb[1 + A] += g0 * rho + gA * T(2) * rho * un + gB * rho * u1 + gC * rho * u2 +
                 g4 * ((E + ptot) + rho * un * un - bn * bn) + hB * b1 + hC * b2_;
    qb[1 + B] += gB * rho * un + g4 * (rho * un * u1 - bn * b1) - hB * bn;
    qb[1 + C] += gC * rho * un + g4 * (rho * un * u2 - bn * b2_) - hC * bn;
    qb[4] += gA + g4 * un * (s.igm1 + T(1));
    qb[5 + A] += -gA * bn - gB * b1 - gC * b2_ + g4 * (un * bn - vdotb) - hB * u1 - hC * u2;
    qb[5 + B] += gA * b1 - gB * bn + g4 * (T(2) * un * b1 - bn * u1) + hB * un;
    qb[5 + C] += gA * b2_ - gC * bn + g4 * (T(2) * un * b2_ - bn * u2) + hC * un;
}

// ---- fast_speed (glmmhd_turb.jl:89-95) ------------------------------------------------------
template <typename T>
GLM_HD void fast_speed_adj(T rho, T p, T b2, T bn, T cfb, T& rho_b, T& p_b, T& b2_b, T& bn_b, const Phys<T>& s)
{
    const T ir = trcp(rho);
    const T c2 = s.gamma * p * ir;
    const T bb = b2 * ir;
    const T d2 = T(0.5) * (bb + c2);
    const T disc = d2 * d2 - c2 * bn * bn * ir;
    const T sd = tsqrt(tmax(disc, T(0)));
    const T cf = tsqrt(d2 + sd);
    const T x_b = cfb * dsqrt(cf);
    T d2_b = x_b, c2_b = T(0), ir_b = T(0);
    if (disc > T(0)) {
        const T disc_b = x_b * dsqrt(sd);
        d2_b += disc_b * T(2) * d2;
        c2_b += -disc_b * bn * bn * ir;
        bn_b += -disc_b * T(2) * c2 * bn * ir;
        ir_b += -disc_b * c2 * bn * bn;
    }
    const T bb_b = T(0.5) * d2_b;
    c2_b += T(0.5) * d2_b;
    b2_b += bb_b * ir;
    ir_b += bb_b * b2;
    p_b += c2_b * s.gamma * ir;
    ir_b += c2_b * s.gamma * p;
    rho_b += -ir * ir * ir_b;
}

// ---- limiters and reconstructions -----------------------------------------------------------
// moncen (glmmhd_turb.jl:145-150): adjoint w.r.t. (dl, dr)
template <typename T>
GLM_HD void moncen_adj(T dl, T dr, T mb, T& dl_b, T& dr_b)
{
    if (dl * dr <= T(0)) return;
    const T dc = T(0.5) * (dl + dr);
    const T a = T(2) * tmin(tabs(dl), tabs(dr)), b = tabs(dc);
    const T g = mb * sgn1(dc);   // d copysign(m, dc)/dm
    if (a < b) {
        if (tabs(dl) <= tabs(dr)) dl_b += T(2) * sgn1(dl) * g;
        else dr_b += T(2) * sgn1(dr) * g;
    } else {
        const T dc_b = g * sgn1(dc);
        dl_b += T(0.5) * dc_b;
        dr_b += T(0.5) * dc_b;
    }
}
template <typename T>
GLM_HD void minmod_adj(T a, T b, T mb, T& a_b, T& b_b)
{
    if (a * b <= T(0)) return;
    if (tabs(a) < tabs(b)) a_b += mb;
    else b_b += mb;
}
// half slope h(l, m, r) (glmmhd_turb.jl:153-155, :302-303)
template <typename T, int SLOPE>
GLM_HD void half_slope_adj(T l, T m, T r, T hb, T& l_b, T& m_b, T& r_b)
{
    const T dl = m - l, dr = r - m;
    T dl_b = T(0), dr_b = T(0);
    if (SLOPE == SLOPE_MINMOD) minmod_adj(dl, dr, T(0.5) * hb, dl_b, dr_b);
    else moncen_adj(dl, dr, T(0.5) * hb, dl_b, dr_b);
    m_b += dl_b - dr_b;
    l_b -= dl_b;
    r_b += dr_b;
}

// ppm_edges (glmmhd_turb.jl:917-923)
template <typename T>
GLM_HD void ppm_edges_adj(T qm, T q0, T qp, T lo_b, T hi_b, T& qm_b, T& q0_b, T& qp_b)
{
    const T slope = T(0.25) * (qp - qm);
    const T curve = (qm - T(2) * q0 + qp) * T(1.0 / 12.0);
    const T ql = q0 - slope + curve, qr = q0 + slope + curve;
    const T dq = qr - ql;
    const T diff = (q0 - T(0.5) * (ql + qr)) * dq;
    const T lim = dq * dq * T(1.0 / 6.0);
    const bool flat = (qr - q0) * (q0 - ql) <= T(0);
    const bool hi_os = diff > lim;
    const bool lo_os = diff < -lim;
    T ql_b = T(0), qr_b = T(0);
    if (flat) {
        q0_b += lo_b + hi_b;
        return;
    }
    if (hi_os) { q0_b += T(3) * lo_b; qr_b -= T(2) * lo_b; }
    else ql_b += lo_b;
    if (hi_os) qr_b += hi_b;
    else if (lo_os) { q0_b += T(3) * hi_b; ql_b -= T(2) * hi_b; }
    else qr_b += hi_b;
    q0_b += ql_b + qr_b;
    const T slope_b = qr_b - ql_b, curve_b = (ql_b + qr_b) * T(1.0 / 12.0);
    qp_b += T(0.25) * slope_b + curve_b;
    qm_b += -T(0.25) * slope_b + curve_b;
    q0_b += -T(2) * curve_b;
}

// ---- Riemann solvers (x frame; L, R floored and GLM-cleaned) ---------------------------------
template <typename T, bool BOUND>
GLM_HD void llf_x_adj(const T L[9], const T R[9], const T fb[9], T Lb[9], T Rb[9], const Phys<T>& s)
{
    const T b2L = L[5] * L[5] + L[6] * L[6] + L[7] * L[7];
    const T b2R = R[5] * R[5] + R[6] * R[6] + R[7] * R[7];
    T cL, cR;
    if (BOUND) {
        cL = tsqrt((s.gamma * L[4] + b2L) * trcp(L[0]));
        cR = tsqrt((s.gamma * R[4] + b2R) * trcp(R[0]));
    } else {
        cL = fast_speed(L[0], L[4], b2L, L[5], s);
        cR = fast_speed(R[0], R[4], b2R, R[5], s);
    }
    const T aL = tabs(L[1]) + cL, aR = tabs(R[1]) + cR;
    const T smax = tmax(tmax(aL, aR), s.ch);
    T uL[9], uR[9];
    prim2cons(L, uL, s);
    prim2cons(R, uR, s);
    T fLb[9], uLb[9], uRb[9];
    T smax_b = T(0);
    for (int v = 0; v < 9; ++v) {
        fLb[v] = T(0.5) * fb[v];
        uRb[v] = -T(0.5) * smax * fb[v];
        uLb[v] = T(0.5) * smax * fb[v];
        smax_b += -T(0.5) * fb[v] * (uR[v] - uL[v]);
    }
    flux_dir_adj<T, 0>(L, fLb, Lb, s);
    flux_dir_adj<T, 0>(R, fLb, Rb, s);
    prim2cons_adj(L, uLb, Lb, s);
    prim2cons_adj(R, uRb, Rb, s);
    T cL_b = T(0), cR_b = T(0);
    const T mlr = tmax(aL, aR);
    if (mlr > s.ch) {
        if (aL >= aR) { Lb[1] += smax_b * sgn1(L[1]); cL_b = smax_b; }
        else { Rb[1] += smax_b * sgn1(R[1]); cR_b = smax_b; }
    }
    T b2L_b = T(0), b2R_b = T(0);
    if (BOUND) {
        if (cL > T(0)) {
            const T y_b = cL_b * dsqrt(cL), X = s.gamma * L[4] + b2L;
            Lb[4] += s.gamma * y_b / L[0]; b2L_b += y_b / L[0]; Lb[0] += -y_b * X / (L[0] * L[0]);
        }
        if (cR > T(0)) {
            const T y_b = cR_b * dsqrt(cR), X = s.gamma * R[4] + b2R;
            Rb[4] += s.gamma * y_b / R[0]; b2R_b += y_b / R[0]; Rb[0] += -y_b * X / (R[0] * R[0]);
        }
    } else {
        fast_speed_adj(L[0], L[4], b2L, L[5], cL_b, Lb[0], Lb[4], b2L_b, Lb[5], s);
        fast_speed_adj(R[0], R[4], b2R, R[5], cR_b, Rb[0], Rb[4], b2R_b, Rb[5], s);
    }
    for (int j = 5; j <= 7; ++j) {
        Lb[j] += T(2) * L[j] * b2L_b;
        Rb[j] += T(2) * R[j] * b2R_b;
    }
}

template <typename T>
GLM_HD void hll_x_adj(const T L[9], const T R[9], const T fb[9], T Lb[9], T Rb[9], const Phys<T>& s)
{
    const T bn = L[5];
    const T b2L = L[5] * L[5] + L[6] * L[6] + L[7] * L[7];
    const T b2R = R[5] * R[5] + R[6] * R[6] + R[7] * R[7];
    const T cfL = fast_speed(L[0], L[4], b2L, bn, s);
    const T cfR = fast_speed(R[0], R[4], b2R, bn, s);
    const T a = L[1] - cfL, b = R[1] - cfR, c = L[1] + cfL, d = R[1] + cfR;
    const T SL = tmin(tmin(a, b), T(0));
    const T SR = tmax(tmax(c, d), T(0));
    T fL[9], fR[9], uL[9], uR[9];
    flux_x(L, fL, s); flux_x(R, fR, s);
    prim2cons(L, uL, s); prim2cons(R, uR, s);
    const T ihd = trcp(SR - SL);
    const T slsr = SL * SR;
    T fLb[9], fRb[9], uLb[9], uRb[9];
    T SL_b = T(0), SR_b = T(0), ihd_b = T(0);
    for (int v = 0; v < 9; ++v) {
        const T du = uR[v] - uL[v];
        const T num = SR * fL[v] - SL * fR[v] + slsr * du;
        fLb[v] = fb[v] * SR * ihd;
        fRb[v] = -fb[v] * SL * ihd;
        uRb[v] = fb[v] * slsr * ihd;
        uLb[v] = -uRb[v];
        ihd_b += fb[v] * num;
        SR_b += fb[v] * ihd * (fL[v] + SL * du);
        SL_b += fb[v] * ihd * (-fR[v] + SR * du);
    }
    SR_b += -ihd * ihd * ihd_b;
    SL_b += ihd * ihd * ihd_b;
    flux_dir_adj<T, 0>(L, fLb, Lb, s);
    flux_dir_adj<T, 0>(R, fRb, Rb, s);
    prim2cons_adj(L, uLb, Lb, s);
    prim2cons_adj(R, uRb, Rb, s);
    T cfL_b = T(0), cfR_b = T(0);
    if (tmin(a, b) < T(0)) {
        if (a <= b) { Lb[1] += SL_b; cfL_b -= SL_b; }
        else { Rb[1] += SL_b; cfR_b -= SL_b; }
    }
    if (tmax(c, d) > T(0)) {
        if (c >= d) { Lb[1] += SR_b; cfL_b += SR_b; }
        else { Rb[1] += SR_b; cfR_b += SR_b; }
    }
    T b2L_b = T(0), b2R_b = T(0), bn_b = T(0);
    fast_speed_adj(L[0], L[4], b2L, bn, cfL_b, Lb[0], Lb[4], b2L_b, bn_b, s);
    fast_speed_adj(R[0], R[4], b2R, bn, cfR_b, Rb[0], Rb[4], b2R_b, bn_b, s);
    Lb[5] += bn_b;
    for (int j = 5; j <= 7; ++j) {
        Lb[j] += T(2) * L[j] * b2L_b;
        Rb[j] += T(2) * R[j] * b2R_b;
    }
}

// Star-state quantities of one HLLD side (glmmhd_turb.jl:227-239) and their adjoint.
template <typename T>
struct HlldStar {
    T denK, ik, fac, gb;
    bool sm;
    T vys, vzs, bys, bzs;
};
template <typename T>
GLM_HD HlldStar<T> hlld_star(T d, T a, T S, T SM, T u, T v, T w, T by, T bz, T bn)
{
    HlldStar<T> h;
    h.denK = d * a * (S - SM) - bn * bn;
    h.sm = tabs(h.denK) < T(1e-12);
    h.ik = trcp(h.denK);
    h.fac = bn * (SM - u) * h.ik;
    h.gb = (d * a * a - bn * bn) * h.ik;
    h.vys = h.sm ? v : v - by * h.fac;
    h.vzs = h.sm ? w : w - bz * h.fac;
    h.bys = h.sm ? by : by * h.gb;
    h.bzs = h.sm ? bz : bz * h.gb;
    return h;
}
template <typename T>
GLM_HD void hlld_star_adj(const HlldStar<T>& h, T d, T a, T S, T SM, T u, T v, T w, T by, T bz, T bn,
                          T vys_b, T vzs_b, T bys_b, T bzs_b,
                          T& d_b, T& a_b, T& S_b, T& SM_b, T& u_b, T& v_b, T& w_b, T& by_b, T& bz_b, T& bn_b)
{
    if (h.sm) {
        v_b += vys_b; w_b += vzs_b; by_b += bys_b; bz_b += bzs_b;
        return;
    }
    v_b += vys_b;
    by_b += -vys_b * h.fac;
    T fac_b = -vys_b * by;
    w_b += vzs_b;
    bz_b += -vzs_b * h.fac;
    fac_b += -vzs_b * bz;
    by_b += bys_b * h.gb;
    T gb_b = bys_b * by;
    bz_b += bzs_b * h.gb;
    gb_b += bzs_b * bz;
    bn_b += fac_b * (SM - u) * h.ik;
    SM_b += fac_b * bn * h.ik;
    u_b += -fac_b * bn * h.ik;
    T ik_b = fac_b * bn * (SM - u);
    d_b += gb_b * a * a * h.ik;
    a_b += gb_b * T(2) * d * a * h.ik;
    bn_b += -gb_b * T(2) * bn * h.ik;
    ik_b += gb_b * (d * a * a - bn * bn);
    const T den_b = -h.ik * h.ik * ik_b;
    d_b += den_b * a * (S - SM);
    a_b += den_b * d * (S - SM);
    S_b += den_b * d * a;
    SM_b += -den_b * d * a;
    bn_b += -den_b * T(2) * bn;
}

// glmmhd_turb.jl:193-272 hlld_x (the same sampled-side formulation as hlld_x in physics.hpp)
template <typename T>
GLM_HD void hlld_x_adj(const T L[9], const T R[9], const T fb[9], T Lb[9], T Rb[9], const Phys<T>& s)
{
    // ---- forward recompute
    const T bn = L[5];
    const T dL = L[0], uL = L[1], vL = L[2], wL = L[3], pL = L[4], byL = L[6], bzL = L[7];
    const T dR = R[0], uR = R[1], vR = R[2], wR = R[3], pR = R[4], byR = R[6], bzR = R[7];
    const T b2L = bn * bn + byL * byL + bzL * bzL;
    const T b2R = bn * bn + byR * byR + bzR * bzR;
    const T ptL = pL + T(0.5) * b2L, ptR = pR + T(0.5) * b2R;
    const T cfL = fast_speed(dL, pL, b2L, bn, s);
    const T cfR = fast_speed(dR, pR, b2R, bn, s);
    const T cfm = tmax(cfL, cfR);
    const T mn = tmin(uL, uR), mx = tmax(uL, uR);
    const T SL = tmin(mn - cfm, T(0));
    const T SR = tmax(mx + cfm, T(0));
    const T aL = SL - uL, aR = SR - uR;
    const T Dd = aR * dR - aL * dL;
    const T iden = trcp(Dd);
    const T Nsm = aR * dR * uR - aL * dL * uL - ptR + ptL;
    const T SM = Nsm * iden;
    const T Npts = aR * dR * ptL - aL * dL * ptR + dL * dR * aR * aL * (uR - uL);
    const T pts = Npts * iden;
    const T iL = trcp(SL - SM), iR = trcp(SR - SM);
    const T XL = dL * aL * iL, XR = dR * aR * iR;
    const T dLs = tmax(XL, T(1e-12));
    const T dRs = tmax(XR, T(1e-12));
    const T sqdLs = tsqrt(dLs), sqdRs = tsqrt(dRs);
    const T absbn = tabs(bn);
    const T rL = trcp(sqdLs), rR = trcp(sqdRs);
    const T SLs = SM - absbn * rL;
    const T SRs = SM + absbn * rR;
    const HlldStar<T> hL = hlld_star(dL, aL, SL, SM, uL, vL, wL, byL, bzL, bn);
    const HlldStar<T> hR = hlld_star(dR, aR, SR, SM, uR, vR, wR, byR, bzR, bn);
    const T sgn = bn >= T(0) ? T(1) : T(-1);
    const T invsum = trcp(sqdLs + sqdRs);
    const T Nvy = sqdLs * hL.vys + sqdRs * hR.vys + (hR.bys - hL.bys) * sgn;
    const T Nvz = sqdLs * hL.vzs + sqdRs * hR.vzs + (hR.bzs - hL.bzs) * sgn;
    const T Nby = sqdLs * hR.bys + sqdRs * hL.bys + sqdLs * sqdRs * (hR.vys - hL.vys) * sgn;
    const T Nbz = sqdLs * hR.bzs + sqdRs * hL.bzs + sqdLs * sqdRs * (hR.vzs - hL.vzs) * sgn;
    const T vyss = Nvy * invsum, vzss = Nvz * invsum, byss = Nby * invsum, bzss = Nbz * invsum;
    const T vdotbss = SM * bn + vyss * byss + vzss * bzss;
    const bool left = SM >= T(0);
    const T d = left ? dL : dR, u = left ? uL : uR, v = left ? vL : vR, w = left ? wL : wR;
    const T p = left ? pL : pR, by = left ? byL : byR, bz = left ? bzL : bzR;
    const T b2 = left ? b2L : b2R, pt = left ? ptL : ptR;
    const T S = left ? SL : SR, a = left ? aL : aR, iS = left ? iL : iR;
    const T ds = left ? dLs : dRs, sqd = left ? sqdLs : sqdRs, Ss = left ? SLs : SRs;
    const HlldStar<T>& hs = left ? hL : hR;
    const T vys = hs.vys, vzs = hs.vzs, bys = hs.bys, bzs = hs.bzs;
    const T E = p * s.igm1 + T(0.5) * d * (u * u + v * v + w * w) + T(0.5) * b2;
    const T vdotb = u * bn + v * by + w * bz;
    const T vdotbs = SM * bn + vys * bys + vzs * bzs;
    const T Nes = a * E - pt * u + pts * SM + bn * (vdotb - vdotbs);
    const T Es = Nes * iS;
    const T sg = left ? -sgn : sgn;
    const T Ess = Es + sqd * (vdotbs - vdotbss) * sg;
    const bool outer = left ? (SLs >= T(0)) : !(SRs >= T(0));
    const T q[9] = {d, u, v, w, p, bn, by, bz, T(0)};
    const T U[9] = {d, d * u, d * v, d * w, E, bn, by, bz, T(0)};
    const T Us[9] = {ds, ds * SM, ds * vys, ds * vzs, Es, bn, bys, bzs, T(0)};
    const T Uss[9] = {ds, ds * SM, ds * vyss, ds * vzss, Ess, bn, byss, bzss, T(0)};

    // ---- reverse
    T Ub[9], Usb[9], Ussb[9];
    T S_b = T(0), Ss_b = T(0);
    for (int k = 0; k < 9; ++k) {
        S_b += fb[k] * (Us[k] - U[k]);
        Ub[k] = -S * fb[k];
        if (outer) {
            Usb[k] = S * fb[k];
            Ussb[k] = T(0);
        } else {
            Ss_b += fb[k] * (Uss[k] - Us[k]);
            Ussb[k] = Ss * fb[k];
            Usb[k] = -(Ss - S) * fb[k];
        }
    }
    T qb[9] = {T(0), T(0), T(0), T(0), T(0), T(0), T(0), T(0), T(0)};
    flux_dir_adj<T, 0>(q, fb, qb, s);
    // side variables
    T d_b = qb[0], u_b = qb[1], v_b = qb[2], w_b = qb[3], p_b = qb[4], bn_b = qb[5], by_b = qb[6], bz_b = qb[7];
    T E_b = T(0), b2_b = T(0), pt_b = T(0), a_b = T(0), iS_b = T(0), ds_b = T(0), sqd_b = T(0);
    T vys_b = T(0), vzs_b = T(0), bys_b = T(0), bzs_b = T(0);
    T SM_b = T(0), pts_b = T(0), vyss_b = T(0), vzss_b = T(0), byss_b = T(0), bzss_b = T(0), Ess_b = T(0), Es_b = T(0);
    // Uss
    ds_b += Ussb[0] + Ussb[1] * SM + Ussb[2] * vyss + Ussb[3] * vzss;
    SM_b += Ussb[1] * ds;
    vyss_b += Ussb[2] * ds; vzss_b += Ussb[3] * ds;
    Ess_b += Ussb[4]; bn_b += Ussb[5]; byss_b += Ussb[6]; bzss_b += Ussb[7];
    // Us
    ds_b += Usb[0] + Usb[1] * SM + Usb[2] * vys + Usb[3] * vzs;
    SM_b += Usb[1] * ds;
    vys_b += Usb[2] * ds; vzs_b += Usb[3] * ds;
    Es_b += Usb[4]; bn_b += Usb[5]; bys_b += Usb[6]; bzs_b += Usb[7];
    // U
    d_b += Ub[0] + Ub[1] * u + Ub[2] * v + Ub[3] * w;
    u_b += Ub[1] * d; v_b += Ub[2] * d; w_b += Ub[3] * d;
    E_b += Ub[4]; bn_b += Ub[5]; by_b += Ub[6]; bz_b += Ub[7];
    // Ess = Es + sqd (vdotbs - vdotbss) sg
    Es_b += Ess_b;
    sqd_b += Ess_b * (vdotbs - vdotbss) * sg;
    T vdotbs_b = Ess_b * sqd * sg;
    T vdotbss_b = -Ess_b * sqd * sg;
    // Es = Nes iS
    const T Nes_b = Es_b * iS;
    iS_b += Es_b * Nes;
    a_b += Nes_b * E; E_b += Nes_b * a;
    pt_b += -Nes_b * u; u_b += -Nes_b * pt;
    pts_b += Nes_b * SM; SM_b += Nes_b * pts;
    bn_b += Nes_b * (vdotb - vdotbs);
    T vdotb_b = Nes_b * bn;
    vdotbs_b += -Nes_b * bn;
    // vdotbs, vdotb, E
    SM_b += vdotbs_b * bn; bn_b += vdotbs_b * SM;
    vys_b += vdotbs_b * bys; bys_b += vdotbs_b * vys;
    vzs_b += vdotbs_b * bzs; bzs_b += vdotbs_b * vzs;
    u_b += vdotb_b * bn; bn_b += vdotb_b * u;
    v_b += vdotb_b * by; by_b += vdotb_b * v;
    w_b += vdotb_b * bz; bz_b += vdotb_b * w;
    p_b += E_b * s.igm1;
    d_b += E_b * T(0.5) * (u * u + v * v + w * w);
    u_b += E_b * d * u; v_b += E_b * d * v; w_b += E_b * d * w;
    b2_b += T(0.5) * E_b;
    // vdotbss
    SM_b += vdotbss_b * bn; bn_b += vdotbss_b * SM;
    vyss_b += vdotbss_b * byss; byss_b += vdotbss_b * vyss;
    vzss_b += vdotbss_b * bzss; bzss_b += vdotbss_b * vzss;
    // per-side accumulators
    T dL_b = T(0), uL_b = T(0), vL_b = T(0), wL_b = T(0), pL_b = T(0), byL_b = T(0), bzL_b = T(0);
    T dR_b = T(0), uR_b = T(0), vR_b = T(0), wR_b = T(0), pR_b = T(0), byR_b = T(0), bzR_b = T(0);
    T b2L_b = T(0), b2R_b = T(0), ptL_b = T(0), ptR_b = T(0), SL_b = T(0), SR_b = T(0), aL_b = T(0), aR_b = T(0);
    T iL_b = T(0), iR_b = T(0), dLs_b = T(0), dRs_b = T(0), sqdLs_b = T(0), sqdRs_b = T(0), SLs_b = T(0), SRs_b = T(0);
    T vyLs_b = T(0), vzLs_b = T(0), byLs_b = T(0), bzLs_b = T(0), vyRs_b = T(0), vzRs_b = T(0), byRs_b = T(0), bzRs_b = T(0);
    if (left) {
        dL_b += d_b; uL_b += u_b; vL_b += v_b; wL_b += w_b; pL_b += p_b; byL_b += by_b; bzL_b += bz_b;
        b2L_b += b2_b; ptL_b += pt_b; SL_b += S_b; aL_b += a_b; iL_b += iS_b; dLs_b += ds_b; sqdLs_b += sqd_b;
        SLs_b += Ss_b; vyLs_b += vys_b; vzLs_b += vzs_b; byLs_b += bys_b; bzLs_b += bzs_b;
    } else {
        dR_b += d_b; uR_b += u_b; vR_b += v_b; wR_b += w_b; pR_b += p_b; byR_b += by_b; bzR_b += bz_b;
        b2R_b += b2_b; ptR_b += pt_b; SR_b += S_b; aR_b += a_b; iR_b += iS_b; dRs_b += ds_b; sqdRs_b += sqd_b;
        SRs_b += Ss_b; vyRs_b += vys_b; vzRs_b += vzs_b; byRs_b += bys_b; bzRs_b += bzs_b;
    }
    // double-star states
    const T Nvy_b = vyss_b * invsum, Nvz_b = vzss_b * invsum, Nby_b = byss_b * invsum, Nbz_b = bzss_b * invsum;
    T invsum_b = vyss_b * Nvy + vzss_b * Nvz + byss_b * Nby + bzss_b * Nbz;
    sqdLs_b += Nvy_b * hL.vys; vyLs_b += Nvy_b * sqdLs; sqdRs_b += Nvy_b * hR.vys; vyRs_b += Nvy_b * sqdRs;
    byRs_b += Nvy_b * sgn; byLs_b -= Nvy_b * sgn;
    sqdLs_b += Nvz_b * hL.vzs; vzLs_b += Nvz_b * sqdLs; sqdRs_b += Nvz_b * hR.vzs; vzRs_b += Nvz_b * sqdRs;
    bzRs_b += Nvz_b * sgn; bzLs_b -= Nvz_b * sgn;
    sqdLs_b += Nby_b * (hR.bys + sqdRs * (hR.vys - hL.vys) * sgn);
    sqdRs_b += Nby_b * (hL.bys + sqdLs * (hR.vys - hL.vys) * sgn);
    byRs_b += Nby_b * sqdLs; byLs_b += Nby_b * sqdRs;
    vyRs_b += Nby_b * sqdLs * sqdRs * sgn; vyLs_b -= Nby_b * sqdLs * sqdRs * sgn;
    sqdLs_b += Nbz_b * (hR.bzs + sqdRs * (hR.vzs - hL.vzs) * sgn);
    sqdRs_b += Nbz_b * (hL.bzs + sqdLs * (hR.vzs - hL.vzs) * sgn);
    bzRs_b += Nbz_b * sqdLs; bzLs_b += Nbz_b * sqdRs;
    vzRs_b += Nbz_b * sqdLs * sqdRs * sgn; vzLs_b -= Nbz_b * sqdLs * sqdRs * sgn;
    sqdLs_b += -invsum * invsum * invsum_b;
    sqdRs_b += -invsum * invsum * invsum_b;
    // star states
    T bn_b2 = T(0);
    hlld_star_adj(hL, dL, aL, SL, SM, uL, vL, wL, byL, bzL, bn, vyLs_b, vzLs_b, byLs_b, bzLs_b,
                  dL_b, aL_b, SL_b, SM_b, uL_b, vL_b, wL_b, byL_b, bzL_b, bn_b2);
    hlld_star_adj(hR, dR, aR, SR, SM, uR, vR, wR, byR, bzR, bn, vyRs_b, vzRs_b, byRs_b, bzRs_b,
                  dR_b, aR_b, SR_b, SM_b, uR_b, vR_b, wR_b, byR_b, bzR_b, bn_b2);
    bn_b += bn_b2;
    // SLs, SRs
    SM_b += SLs_b + SRs_b;
    T absbn_b = -SLs_b * rL + SRs_b * rR;
    sqdLs_b += SLs_b * absbn * rL * rL;
    sqdRs_b += -SRs_b * absbn * rR * rR;
    bn_b += absbn_b * sgn1(bn);
    // sqrt, floors
    dLs_b += sqdLs_b * dsqrt(sqdLs);
    dRs_b += sqdRs_b * dsqrt(sqdRs);
    if (XL > T(1e-12)) { dL_b += dLs_b * aL * iL; aL_b += dLs_b * dL * iL; iL_b += dLs_b * dL * aL; }
    if (XR > T(1e-12)) { dR_b += dRs_b * aR * iR; aR_b += dRs_b * dR * iR; iR_b += dRs_b * dR * aR; }
    SL_b += -iL * iL * iL_b; SM_b += iL * iL * iL_b;
    SR_b += -iR * iR * iR_b; SM_b += iR * iR * iR_b;
    // pts and SM (all consumers done)
    const T Npts_b = pts_b * iden;
    T iden_b = pts_b * Npts;
    const T du = uR - uL;
    aR_b += Npts_b * (dR * ptL + dL * dR * aL * du);
    dR_b += Npts_b * (aR * ptL + dL * aR * aL * du);
    ptL_b += Npts_b * aR * dR;
    aL_b += Npts_b * (-dL * ptR + dL * dR * aR * du);
    dL_b += Npts_b * (-aL * ptR + dR * aR * aL * du);
    ptR_b += -Npts_b * aL * dL;
    uR_b += Npts_b * dL * dR * aR * aL;
    uL_b -= Npts_b * dL * dR * aR * aL;
    const T Nsm_b = SM_b * iden;
    iden_b += SM_b * Nsm;
    aR_b += Nsm_b * dR * uR; dR_b += Nsm_b * aR * uR; uR_b += Nsm_b * aR * dR;
    aL_b -= Nsm_b * dL * uL; dL_b -= Nsm_b * aL * uL; uL_b -= Nsm_b * aL * dL;
    ptR_b -= Nsm_b; ptL_b += Nsm_b;
    const T Dd_b = -iden * iden * iden_b;
    aR_b += Dd_b * dR; dR_b += Dd_b * aR; aL_b -= Dd_b * dL; dL_b -= Dd_b * aL;
    // aL = SL - uL, aR = SR - uR
    SL_b += aL_b; uL_b -= aL_b;
    SR_b += aR_b; uR_b -= aR_b;
    // SL, SR, cfm
    T cfm_b = T(0);
    if (mn - cfm < T(0)) {
        if (uL <= uR) uL_b += SL_b; else uR_b += SL_b;
        cfm_b -= SL_b;
    }
    if (mx + cfm > T(0)) {
        if (uL >= uR) uL_b += SR_b; else uR_b += SR_b;
        cfm_b += SR_b;
    }
    T cfL_b = T(0), cfR_b = T(0);
    if (cfL >= cfR) cfL_b += cfm_b; else cfR_b += cfm_b;
    fast_speed_adj(dL, pL, b2L, bn, cfL_b, dL_b, pL_b, b2L_b, bn_b, s);
    fast_speed_adj(dR, pR, b2R, bn, cfR_b, dR_b, pR_b, b2R_b, bn_b, s);
    pL_b += ptL_b; b2L_b += T(0.5) * ptL_b;
    pR_b += ptR_b; b2R_b += T(0.5) * ptR_b;
    bn_b += T(2) * bn * (b2L_b + b2R_b);
    byL_b += T(2) * byL * b2L_b; bzL_b += T(2) * bzL * b2L_b;
    byR_b += T(2) * byR * b2R_b; bzR_b += T(2) * bzR * b2R_b;
    Lb[0] += dL_b; Lb[1] += uL_b; Lb[2] += vL_b; Lb[3] += wL_b; Lb[4] += pL_b; Lb[5] += bn_b; Lb[6] += byL_b; Lb[7] += bzL_b;
    Rb[0] += dR_b; Rb[1] += uR_b; Rb[2] += vR_b; Rb[3] += wR_b; Rb[4] += pR_b; Rb[6] += byR_b; Rb[7] += bzR_b;
}

// Riemann dispatch adjoint (physics.hpp riemann): floors, GLM pair, solver, rotation.
template <typename T, int RIEM, int D>
GLM_HD void riemann_adj(const T Lq[9], const T Rq[9], const T ob[9], T Lqb[9], T Rqb[9], const Phys<T>& s)
{
    constexpr int A = D % 3, B = (D + 1) % 3, C = (D + 2) % 3;
    T L[9] = {tmax(Lq[0], s.smallr), Lq[1 + A], Lq[1 + B], Lq[1 + C], tmax(Lq[4], s.pfl),
              Lq[5 + A], Lq[5 + B], Lq[5 + C], Lq[8]};
    T R[9] = {tmax(Rq[0], s.smallr), Rq[1 + A], Rq[1 + B], Rq[1 + C], tmax(Rq[4], s.pfl),
              Rq[5 + A], Rq[5 + B], Rq[5 + C], Rq[8]};
    const T L5 = L[5], R5 = R[5], L8 = L[8], R8 = R[8];
    const T bns = T(0.5) * (L5 + R5) - (R8 - L8) * s.half_ich;
    L[5] = bns; R[5] = bns;
    // solver output adjoint in the x frame (slots 5 and 8 of the solver output are replaced)
    const T fb[9] = {ob[0], ob[1 + A], ob[1 + B], ob[1 + C], ob[4], T(0), ob[5 + B], ob[5 + C], T(0)};
    T Lb[9] = {T(0), T(0), T(0), T(0), T(0), T(0), T(0), T(0), T(0)};
    T Rb[9] = {T(0), T(0), T(0), T(0), T(0), T(0), T(0), T(0), T(0)};
    if (RIEM == RIEM_LLF) llf_x_adj<T, false>(L, R, fb, Lb, Rb, s);
    else if (RIEM == RIEM_LLFB) llf_x_adj<T, true>(L, R, fb, Lb, Rb, s);
    else if (RIEM == RIEM_HLL) hll_x_adj(L, R, fb, Lb, Rb, s);
    else {
        const bool uself = (s.llf_dmin > T(0) && tmin(L[0], R[0]) < s.llf_dmin) ||
                           (s.llf_pmin > T(0) && tmin(L[4], R[4]) < s.llf_pmin);
        if (uself) llf_x_adj<T, false>(L, R, fb, Lb, Rb, s);
        else hlld_x_adj(L, R, fb, Lb, Rb, s);
    }
    const T psis_b = ob[5 + A];
    const T bns_b = s.ch2 * ob[8] + Lb[5] + Rb[5];
    T L5_b = T(0.5) * bns_b + T(0.5) * s.ch * psis_b;
    T R5_b = T(0.5) * bns_b - T(0.5) * s.ch * psis_b;
    T L8_b = bns_b * s.half_ich + T(0.5) * psis_b;
    T R8_b = -bns_b * s.half_ich + T(0.5) * psis_b;
    if (Lq[0] > s.smallr) Lqb[0] += Lb[0];
    if (Rq[0] > s.smallr) Rqb[0] += Rb[0];
    if (Lq[4] > s.pfl) Lqb[4] += Lb[4];
    if (Rq[4] > s.pfl) Rqb[4] += Rb[4];
    Lqb[1 + A] += Lb[1]; Lqb[1 + B] += Lb[2]; Lqb[1 + C] += Lb[3];
    Rqb[1 + A] += Rb[1]; Rqb[1 + B] += Rb[2]; Rqb[1 + C] += Rb[3];
    Lqb[5 + A] += L5_b; Lqb[5 + B] += Lb[6]; Lqb[5 + C] += Lb[7]; Lqb[8] += L8_b;
    Rqb[5 + A] += R5_b; Rqb[5 + B] += Rb[6]; Rqb[5 + C] += Rb[7]; Rqb[8] += R8_b;
}

// ---- predictors --------------------------------------------------------------------------
// Normal-only Hancock half step of the cell centre in direction D (spike_mhd.cu:149-161).
template <typename T, int D>
GLM_HD void hanc1d_fwd(const T c0[9], const T hs[9], T mh[9], T half_dtdx, const Phys<T>& s)
{
    T t[9], fa[9], fb[9], u[9];
    for (int v = 0; v < 9; ++v) t[v] = c0[v] - hs[v];
    flux_dir<T, D>(t, fa, s);
    for (int v = 0; v < 9; ++v) t[v] = c0[v] + hs[v];
    flux_dir<T, D>(t, fb, s);
    prim2cons(c0, u, s);
    for (int v = 0; v < 9; ++v) u[v] -= half_dtdx * (fb[v] - fa[v]);
    cons2prim(u, mh, s);
}
template <typename T, int D>
GLM_HD void hanc1d_adj(const T c0[9], const T hs[9], const T mhb[9], T c0b[9], T hsb[9], T half_dtdx,
                       const Phys<T>& s)
{
    T tm[9], tp[9], fa[9], fb[9], u[9];
    for (int v = 0; v < 9; ++v) { tm[v] = c0[v] - hs[v]; tp[v] = c0[v] + hs[v]; }
    flux_dir<T, D>(tm, fa, s);
    flux_dir<T, D>(tp, fb, s);
    prim2cons(c0, u, s);
    for (int v = 0; v < 9; ++v) u[v] -= half_dtdx * (fb[v] - fa[v]);
    T ub[9] = {T(0), T(0), T(0), T(0), T(0), T(0), T(0), T(0), T(0)};
    cons2prim_adj(u, mhb, ub, s);
    prim2cons_adj(c0, ub, c0b, s);
    T fbb[9], fab[9], tpb[9], tmb[9];
    for (int v = 0; v < 9; ++v) { fbb[v] = -half_dtdx * ub[v]; fab[v] = half_dtdx * ub[v]; tpb[v] = T(0); tmb[v] = T(0); }
    flux_dir_adj<T, D>(tp, fbb, tpb, s);
    flux_dir_adj<T, D>(tm, fab, tmb, s);
    for (int v = 0; v < 9; ++v) { c0b[v] += tpb[v] + tmb[v]; hsb[v] += tpb[v] - tmb[v]; }
}

// Full-transverse Hancock predictor of one cell (glmmhd_turb.jl:165-172, revert :391-395) from
// its six neighbours nb[2d] (low) and nb[2d+1] (high) in direction d.  Returns true if reverted.
template <typename T, int SLOPE>
GLM_HD bool hancock_fwd(const T p0[9], const T (*nb)[9], T mh[9], T half_dtdx, int mh_revert, const Phys<T>& s)
{
    T acc[9], t[9], f[9], hs[9];
    for (int v = 0; v < 9; ++v) acc[v] = T(0);
    for (int d = 0; d < 3; ++d) {
        for (int v = 0; v < 9; ++v) hs[v] = half_slope<T, SLOPE>(nb[2 * d][v], p0[v], nb[2 * d + 1][v]);
        for (int sgn = 0; sgn < 2; ++sgn) {
            for (int v = 0; v < 9; ++v) t[v] = sgn == 0 ? p0[v] + hs[v] : p0[v] - hs[v];
            if (d == 0) flux_dir<T, 0>(t, f, s);
            else if (d == 1) flux_dir<T, 1>(t, f, s);
            else flux_dir<T, 2>(t, f, s);
            for (int v = 0; v < 9; ++v) acc[v] = sgn == 0 ? acc[v] + f[v] : acc[v] - f[v];
        }
    }
    prim2cons(p0, t, s);
    for (int v = 0; v < 9; ++v) t[v] = t[v] - half_dtdx * acc[v];
    cons2prim(t, mh, s);
    if (mh_revert && (mh[0] <= s.smallr || mh[4] <= s.pfl)) {
        for (int v = 0; v < 9; ++v) mh[v] = p0[v];
        return true;
    }
    return false;
}
// The predictor's adjoint in two parts, so that a kernel can take the directions one at a time:
// the core (conversions and the revert) gives the adjoint of the flux sum, accb; each direction
// then adds its flux and slope adjoints.  Returns false if the predictor reverted (mh = p0).
template <typename T, int SLOPE, int D>
GLM_HD void hancock_acc_dir(const T p0[9], const T l[9], const T r[9], T acc[9], const Phys<T>& s)
{
    T t[9], f[9];
    for (int sgn = 0; sgn < 2; ++sgn) {
        for (int v = 0; v < 9; ++v) {
            const T h = half_slope<T, SLOPE>(l[v], p0[v], r[v]);
            t[v] = sgn == 0 ? p0[v] + h : p0[v] - h;
        }
        flux_dir<T, D>(t, f, s);
        for (int v = 0; v < 9; ++v) acc[v] = sgn == 0 ? acc[v] + f[v] : acc[v] - f[v];
    }
}
// Core from the flux sum acc of the three directions.
template <typename T>
GLM_HD bool hancock_core_from_acc(const T p0[9], const T acc[9], const T mhb[9], T p0b[9], T accb[9], T half_dtdx,
                                  int mh_revert, const Phys<T>& s)
{
    T u[9], mh[9];
    prim2cons(p0, u, s);
    for (int v = 0; v < 9; ++v) u[v] = u[v] - half_dtdx * acc[v];
    cons2prim(u, mh, s);
    if (mh_revert && (mh[0] <= s.smallr || mh[4] <= s.pfl)) {
        for (int v = 0; v < 9; ++v) p0b[v] += mhb[v];
        return false;
    }
    T ub[9] = {T(0), T(0), T(0), T(0), T(0), T(0), T(0), T(0), T(0)};
    cons2prim_adj(u, mhb, ub, s);
    prim2cons_adj(p0, ub, p0b, s);
    for (int v = 0; v < 9; ++v) accb[v] = -half_dtdx * ub[v];
    return true;
}
template <typename T, int SLOPE>
GLM_HD bool hancock_adj_core(const T p0[9], const T (*nb)[9], const T mhb[9], T p0b[9], T accb[9], T half_dtdx,
                             int mh_revert, const Phys<T>& s)
{
    T acc[9];
    for (int v = 0; v < 9; ++v) acc[v] = T(0);
    hancock_acc_dir<T, SLOPE, 0>(p0, nb[0], nb[1], acc, s);
    hancock_acc_dir<T, SLOPE, 1>(p0, nb[2], nb[3], acc, s);
    hancock_acc_dir<T, SLOPE, 2>(p0, nb[4], nb[5], acc, s);
    return hancock_core_from_acc(p0, acc, mhb, p0b, accb, half_dtdx, mh_revert, s);
}
template <typename T, int SLOPE, int D>
GLM_HD void hancock_adj_dir(const T p0[9], const T l[9], const T r[9], const T accb[9], T p0b[9], T lb[9], T rb[9],
                            const Phys<T>& s)
{
    T hs[9], hsb[9];
    for (int v = 0; v < 9; ++v) { hs[v] = half_slope<T, SLOPE>(l[v], p0[v], r[v]); hsb[v] = T(0); }
    for (int sgn = 0; sgn < 2; ++sgn) {
        T t[9], fb_[9], tb[9];
        for (int v = 0; v < 9; ++v) {
            t[v] = sgn == 0 ? p0[v] + hs[v] : p0[v] - hs[v];
            fb_[v] = sgn == 0 ? accb[v] : -accb[v];
            tb[v] = T(0);
        }
        flux_dir_adj<T, D>(t, fb_, tb, s);
        for (int v = 0; v < 9; ++v) {
            p0b[v] += tb[v];
            hsb[v] += sgn == 0 ? tb[v] : -tb[v];
        }
    }
    for (int v = 0; v < 9; ++v) half_slope_adj<T, SLOPE>(l[v], p0[v], r[v], hsb[v], lb[v], p0b[v], rb[v]);
}

template <typename T, int SLOPE>
GLM_HD void hancock_adj(const T p0[9], const T (*nb)[9], const T mhb[9], T p0b[9], T (*nbb)[9], T half_dtdx,
                        int mh_revert, const Phys<T>& s)
{
    T accb[9];
    if (!hancock_adj_core<T, SLOPE>(p0, nb, mhb, p0b, accb, half_dtdx, mh_revert, s)) return;
    hancock_adj_dir<T, SLOPE, 0>(p0, nb[0], nb[1], accb, p0b, nbb[0], nbb[1], s);
    hancock_adj_dir<T, SLOPE, 1>(p0, nb[2], nb[3], accb, p0b, nbb[2], nbb[3], s);
    hancock_adj_dir<T, SLOPE, 2>(p0, nb[4], nb[5], accb, p0b, nbb[4], nbb[5], s);
}

// ---- face states of one cell in direction D (march_impl.hpp faces_dir) ----------------------
struct FaceOpts {
    int mh_revert, face_guard;
};
template <typename T, int PRED, int RECON, int SLOPE, int D>
GLM_HD void faces_fwd(const T l[9], const T m0[9], const T r[9], const T mhf[9], bool up, T half_dtdx, FaceOpts o,
                      const Phys<T>& s, T lo[9], T hi[9])
{
    T hs[9], mh[9];
    if (RECON != RECON_PPM || PRED == PRED_ONED)
        for (int v = 0; v < 9; ++v) hs[v] = half_slope<T, SLOPE>(l[v], m0[v], r[v]);
    if (PRED == PRED_ONED) {
        hanc1d_fwd<T, D>(m0, hs, mh, half_dtdx, s);
        if (o.mh_revert && (mh[0] <= s.smallr || mh[4] <= s.pfl))
            for (int v = 0; v < 9; ++v) mh[v] = m0[v];
    } else {
        for (int v = 0; v < 9; ++v) mh[v] = mhf[v];
    }
    if (RECON == RECON_PPM) {
        for (int v = 0; v < 9; ++v) {
            T el, eh;
            ppm_edges(l[v], m0[v], r[v], el, eh);
            lo[v] = mh[v] + el - m0[v];
            hi[v] = mh[v] + eh - m0[v];
        }
        if (!up)
            for (int v = 0; v < 9; ++v) {
                T h = PRED == PRED_ONED ? hs[v] : half_slope<T, SLOPE>(l[v], m0[v], r[v]);
                lo[v] = mh[v] - h;
                hi[v] = mh[v] + h;
            }
    } else {
        for (int v = 0; v < 9; ++v) { lo[v] = mh[v] - hs[v]; hi[v] = mh[v] + hs[v]; }
    }
    for (int side = 0; side < 2; ++side) {
        T* f = side == 0 ? lo : hi;
        if (o.face_guard == GUARD_IFLO) {
            if (f[0] < s.smallr || f[4] < s.pfl)
                for (int v = 0; v < 9; ++v) f[v] = m0[v];
        } else if (o.face_guard == GUARD_CLAMP) {
            f[0] = tmax(f[0], s.smallr);
            f[4] = tmax(f[4], s.pfl);
        }
    }
}
template <typename T, int PRED, int RECON, int SLOPE, int D>
GLM_HD void faces_adj(const T l[9], const T m0[9], const T r[9], const T mhf[9], bool up, T half_dtdx, FaceOpts o,
                      const Phys<T>& s, const T lob_in[9], const T hib_in[9], T lb[9], T m0b[9], T rb[9], T mhfb[9])
{
    // forward recompute of the pre-guard faces
    T hs[9], mh[9], lo[9], hi[9];
    bool reverted = false;
    if (RECON != RECON_PPM || PRED == PRED_ONED)
        for (int v = 0; v < 9; ++v) hs[v] = half_slope<T, SLOPE>(l[v], m0[v], r[v]);
    if (PRED == PRED_ONED) {
        hanc1d_fwd<T, D>(m0, hs, mh, half_dtdx, s);
        if (o.mh_revert && (mh[0] <= s.smallr || mh[4] <= s.pfl)) {
            reverted = true;
            for (int v = 0; v < 9; ++v) mh[v] = m0[v];
        }
    } else {
        for (int v = 0; v < 9; ++v) mh[v] = mhf[v];
    }
    if (RECON == RECON_PPM) {
        for (int v = 0; v < 9; ++v) {
            T el, eh;
            ppm_edges(l[v], m0[v], r[v], el, eh);
            lo[v] = mh[v] + el - m0[v];
            hi[v] = mh[v] + eh - m0[v];
        }
        if (!up)
            for (int v = 0; v < 9; ++v) {
                T h = PRED == PRED_ONED ? hs[v] : half_slope<T, SLOPE>(l[v], m0[v], r[v]);
                lo[v] = mh[v] - h;
                hi[v] = mh[v] + h;
            }
    } else {
        for (int v = 0; v < 9; ++v) { lo[v] = mh[v] - hs[v]; hi[v] = mh[v] + hs[v]; }
    }
    // guards
    T lob[9], hib[9];
    for (int v = 0; v < 9; ++v) { lob[v] = lob_in[v]; hib[v] = hib_in[v]; }
    for (int side = 0; side < 2; ++side) {
        T* f = side == 0 ? lo : hi;
        T* fbv = side == 0 ? lob : hib;
        if (o.face_guard == GUARD_IFLO) {
            if (f[0] < s.smallr || f[4] < s.pfl)
                for (int v = 0; v < 9; ++v) { m0b[v] += fbv[v]; fbv[v] = T(0); }
        } else if (o.face_guard == GUARD_CLAMP) {
            if (!(f[0] > s.smallr)) fbv[0] = T(0);
            if (!(f[4] > s.pfl)) fbv[4] = T(0);
        }
    }
    // reconstruction
    T mhb[9], hsb[9];
    for (int v = 0; v < 9; ++v) { mhb[v] = lob[v] + hib[v]; hsb[v] = T(0); }
    if (RECON == RECON_PPM) {
        if (up) {
            for (int v = 0; v < 9; ++v) {
                m0b[v] -= lob[v] + hib[v];
                ppm_edges_adj(l[v], m0[v], r[v], lob[v], hib[v], lb[v], m0b[v], rb[v]);
            }
        } else {
            for (int v = 0; v < 9; ++v) {
                if (PRED == PRED_ONED) hsb[v] += hib[v] - lob[v];
                else half_slope_adj<T, SLOPE>(l[v], m0[v], r[v], hib[v] - lob[v], lb[v], m0b[v], rb[v]);
            }
        }
    } else {
        for (int v = 0; v < 9; ++v) hsb[v] += hib[v] - lob[v];
    }
    // predictor
    if (PRED == PRED_ONED) {
        if (reverted) for (int v = 0; v < 9; ++v) m0b[v] += mhb[v];
        else hanc1d_adj<T, D>(m0, hs, mhb, m0b, hsb, half_dtdx, s);
    } else {
        for (int v = 0; v < 9; ++v) mhfb[v] += mhb[v];
    }
    if (RECON != RECON_PPM || PRED == PRED_ONED)
        for (int v = 0; v < 9; ++v) half_slope_adj<T, SLOPE>(l[v], m0[v], r[v], hsb[v], lb[v], m0b[v], rb[v]);
}

// ---- update epilogue (march_impl.hpp update_cell; glmmhd_turb.jl:440-476) --------------------
struct UpdOpts {
    int update_floor, do_turb;
};
// Epilogue of r (the updated state before floor, damping and kick) with forcing acceleration a3
// (ramp already applied); returns the state after the step.
template <typename T>
GLM_HD void epilogue_fwd(T r[9], const T a3[3], T glm_fac, T smallc, T turb_min_rho, T dt, UpdOpts o,
                         const Phys<T>& s)
{
    if (o.update_floor && r[0] < s.smallr) {
        T emag = T(0.5) * (r[5] * r[5] + r[6] * r[6] + r[7] * r[7]);
        r[0] = s.smallr; r[1] = T(0); r[2] = T(0); r[3] = T(0);
        r[4] = s.pfl * s.igm1 + emag;
    }
    r[8] = r[8] * glm_fac;
    if (o.do_turb && r[0] >= turb_min_rho) {
        T rho = r[0], rhom = tmax(rho, s.smallr), irm = trcp(rhom), fl = rho * smallc * smallc;
        T e = r[4];
        e = tmax(e - T(0.5) * r[1] * r[1] * irm, fl);
        e = tmax(e - T(0.5) * r[2] * r[2] * irm, fl);
        e = tmax(e - T(0.5) * r[3] * r[3] * irm, fl);
        T m2 = r[1] + rhom * a3[0] * dt;
        T m3 = r[2] + rhom * a3[1] * dt;
        T m4 = r[3] + rhom * a3[2] * dt;
        e = tmax(e + T(0.5) * m2 * m2 * irm, fl);
        e = tmax(e + T(0.5) * m3 * m3 * irm, fl);
        e = tmax(e + T(0.5) * m4 * m4 * irm, fl);
        r[1] = m2; r[2] = m3; r[3] = m4; r[4] = e;
    }
}
// Adjoint of epilogue_fwd at r_in: adds to r_b (the adjoint of r_in) and a3_b.
template <typename T>
GLM_HD void epilogue_adj(const T r_in[9], const T a3[3], T glm_fac, T smallc, T turb_min_rho, T dt, UpdOpts o,
                         const Phys<T>& s, const T outb[9], T rb[9], T a3b[3])
{
    T r[9], b[9];
    for (int v = 0; v < 9; ++v) { r[v] = r_in[v]; b[v] = outb[v]; }
    const bool floored = o.update_floor && r[0] < s.smallr;
    if (floored) {
        T emag = T(0.5) * (r[5] * r[5] + r[6] * r[6] + r[7] * r[7]);
        r[0] = s.smallr; r[1] = T(0); r[2] = T(0); r[3] = T(0);
        r[4] = s.pfl * s.igm1 + emag;
    }
    T r8 = r[8];
    if (o.do_turb && r[0] >= turb_min_rho) {
        // forward with the branch record
        const T rho = r[0], rhom = tmax(rho, s.smallr), irm = trcp(rhom), fl = rho * smallc * smallc;
        T e[7];
        bool act[6];
        e[0] = r[4];
        T x;
        x = e[0] - T(0.5) * r[1] * r[1] * irm; act[0] = x > fl; e[1] = tmax(x, fl);
        x = e[1] - T(0.5) * r[2] * r[2] * irm; act[1] = x > fl; e[2] = tmax(x, fl);
        x = e[2] - T(0.5) * r[3] * r[3] * irm; act[2] = x > fl; e[3] = tmax(x, fl);
        const T m2 = r[1] + rhom * a3[0] * dt, m3 = r[2] + rhom * a3[1] * dt, m4 = r[3] + rhom * a3[2] * dt;
        x = e[3] + T(0.5) * m2 * m2 * irm; act[3] = x > fl; e[4] = tmax(x, fl);
        x = e[4] + T(0.5) * m3 * m3 * irm; act[4] = x > fl; e[5] = tmax(x, fl);
        x = e[5] + T(0.5) * m4 * m4 * irm; act[5] = x > fl; e[6] = tmax(x, fl);
        (void)e;
        // reverse
        T eb = b[4], m2b = b[1], m3b = b[2], m4b = b[3], rhob = b[0], irmb = T(0), flb = T(0), rhomb = T(0);
        T r1b = T(0), r2b = T(0), r3b = T(0);
        const T mm[3] = {m2, m3, m4};
        T* mb[3] = {&m2b, &m3b, &m4b};
        for (int k = 5; k >= 3; --k) {
            if (act[k]) {
                *mb[k - 3] += eb * mm[k - 3] * irm;
                irmb += eb * T(0.5) * mm[k - 3] * mm[k - 3];
            } else {
                flb += eb;
                eb = T(0);
            }
        }
        for (int c = 0; c < 3; ++c) {
            const T mbv = *mb[c];
            rhomb += mbv * a3[c] * dt;
            a3b[c] += mbv * rhom * dt;
        }
        r1b += m2b; r2b += m3b; r3b += m4b;
        const T rr[3] = {r[1], r[2], r[3]};
        T* rbp[3] = {&r1b, &r2b, &r3b};
        for (int k = 2; k >= 0; --k) {
            if (act[k]) {
                *rbp[k] += -eb * rr[k] * irm;
                irmb += -eb * T(0.5) * rr[k] * rr[k];
            } else {
                flb += eb;
                eb = T(0);
            }
        }
        // e[0] = r[4]
        b[4] = eb;
        b[1] = r1b; b[2] = r2b; b[3] = r3b;
        rhob += flb * smallc * smallc;
        rhomb += -irm * irm * irmb;
        if (rho > s.smallr) rhob += rhomb;
        b[0] = rhob;
    }
    (void)r8;
    b[8] = b[8] * glm_fac;
    if (floored) {
        for (int j = 5; j <= 7; ++j) b[j] += b[4] * r_in[j];
        b[0] = b[1] = b[2] = b[3] = b[4] = T(0);
    }
    for (int v = 0; v < 9; ++v) rb[v] += b[v];
}

}  // namespace glm
