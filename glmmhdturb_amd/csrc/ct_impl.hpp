// ct_impl.hpp -- constrained-transport MHD step on gfx950 (ct_mhd.jl:20-194, spike_ctm.cu).
//
// State: 8 SoA planes (rho, mx, my, mz, E, bxf, byf, bzf); bxf[i] is Bx on face i-1/2.
// Passes: (1) cell-centred B (bcc_kernel!, ct_mhd.jl:85-96) folded into the primitive pass
// (cellprim, :26-34); (2) the three lower-face fluxes of every cell, each solved once
// (flux_kernel!, :99-122; fflm, spike_ctm.cu:223-245); (3) edge EMFs (emf_kernel!, :139-154);
// (4) the hydro update (:124-136) and the face-B curl update (:156-171) in one kernel, with the
// optional floor / OU kick / next-step CFL folded in.
#pragma once
#include "physics.hpp"

namespace glm {

// Gas pressure (gamma-1)(E - rho v^2/2 - B^2/2), floored (ct_mhd.jl:26-34).  With FP32 state the
// subtraction is carried out in FP64, as in the GLM path: E can be dominated by the kinetic or
// magnetic part.
template <typename T>
GLM_DEV T ct_pressure(T E, T rho, T vx, T vy, T vz, T bx, T by, T bz, T gamma, T smallp)
{
    const double r = rho, u = vx, v = vy, w = vz, b1 = bx, b2 = by, b3 = bz;
    const double e = (double)E - 0.5 * r * (u * u + v * v + w * w) - 0.5 * (b1 * b1 + b2 * b2 + b3 * b3);
    return tmax((T)(((double)gamma - 1.0) * e), smallp);
}

constexpr int CT_NT = 256;

template <typename T>
struct CTArgs {
    Ptr9<const T> q;   // 8 used
    Ptr9<T> o;
    T* prim;           // 8n
    T* flx;            // 24n: flux at the lower-d face, plane (d*8 + t)
    T* emf;            // 3n: ex, ey, ez
    int nx, ny, nz;
    T gamma, igm1, smallr, smallp, dtdx, half_dtdx, dt, dx, lx, ly, lz, smallc, turb_min_rho, ramp;
    int face_slope, riemann, recon, update_floor, do_turb, tgs, want_cfl;
    const T* afield;
    unsigned long long* cmax;
};

GLM_DEV int cwrap(int i, int n)
{
    int m = i % n;
    return m < 0 ? m + n : m;
}

template <typename T>
GLM_DEV size_t cidx(const CTArgs<T>& a, int i, int j, int k)
{
    return (size_t)i + (size_t)a.nx * ((size_t)j + (size_t)a.ny * (size_t)k);
}

// ct_mhd.jl:39-55 dflux in direction D (8 components, F[Bn] = 0)
template <typename T, int D>
GLM_DEV void ct_dflux(const T q[8], T f[8], T igm1)
{
    const T rho = q[0], vx = q[1], vy = q[2], vz = q[3], p = q[4], bx = q[5], by = q[6], bz = q[7];
    T b2 = bx * bx + by * by + bz * bz, ptot = p + T(0.5) * b2;
    T E = p * igm1 + T(0.5) * rho * (vx * vx + vy * vy + vz * vz) + T(0.5) * b2;
    T vb = vx * bx + vy * by + vz * bz;
    const T un = D == 0 ? vx : (D == 1 ? vy : vz);
    const T bn = D == 0 ? bx : (D == 1 ? by : bz);
    f[0] = rho * un;
    f[1] = rho * un * vx - bn * bx + (D == 0 ? ptot : T(0));
    f[2] = rho * un * vy - bn * by + (D == 1 ? ptot : T(0));
    f[3] = rho * un * vz - bn * bz + (D == 2 ? ptot : T(0));
    f[4] = (E + ptot) * un - bn * vb;
    f[5] = D == 0 ? T(0) : un * bx - vx * bn;
    f[6] = D == 1 ? T(0) : un * by - vy * bn;
    f[7] = D == 2 ? T(0) : un * bz - vz * bn;
}

template <typename T>
GLM_DEV void ct_toC(const T q[8], T u[8], T igm1)
{
    const T rho = q[0], vx = q[1], vy = q[2], vz = q[3], p = q[4], bx = q[5], by = q[6], bz = q[7];
    u[0] = rho; u[1] = rho * vx; u[2] = rho * vy; u[3] = rho * vz;
    u[4] = p * igm1 + T(0.5) * rho * (vx * vx + vy * vy + vz * vz) + T(0.5) * (bx * bx + by * by + bz * bz);
    u[5] = bx; u[6] = by; u[7] = bz;
}

template <typename T>
GLM_DEV T ct_cf(const T q[8], T bn, T gamma)
{
    T ir = trcp(q[0]);
    T c2 = gamma * q[4] * ir, b2 = (q[5] * q[5] + q[6] * q[6] + q[7] * q[7]) * ir, dd = T(0.5) * (b2 + c2);
    return tsqrt_pos(dd + tsqrt_nn(tmax(dd * dd - c2 * bn * bn * ir, T(0))));
}

// ct_mhd.jl:74-82 / spike_ctm.cu:170-184 hanc1d with offset hs (face = c -/+ hs).  A is any
// argument block with igm1, half_dtdx, smallr, gamma and smallp (the multi-pass and the march).
template <typename T, int D, typename A>
GLM_DEV void ct_hanc1d(const T q[8], const T hs[8], T mh[8], const A& a)
{
    T t[8], FL[8], FR[8], U[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) t[k] = q[k] - hs[k];
    ct_dflux<T, D>(t, FL, a.igm1);
#pragma unroll
    for (int k = 0; k < 8; ++k) t[k] = q[k] + hs[k];
    ct_dflux<T, D>(t, FR, a.igm1);
    ct_toC(q, U, a.igm1);
#pragma unroll
    for (int k = 0; k < 8; ++k) U[k] = U[k] - a.half_dtdx * (FR[k] - FL[k]);
    T rho = tmax(U[0], a.smallr), ir = trcp(rho);
    T vx = U[1] * ir, vy = U[2] * ir, vz = U[3] * ir;
    T p = ct_pressure(U[4], rho, vx, vy, vz, U[5], U[6], U[7], a.gamma, a.smallp);
    mh[0] = rho; mh[1] = vx; mh[2] = vy; mh[3] = vz; mh[4] = p; mh[5] = U[5]; mh[6] = U[6]; mh[7] = U[7];
}

// The two face states of cell c in direction D from its neighbours l, r (ct_mhd.jl:99-122 /
// spike_ctm.cu:223-245, taken per cell): half slope fac * MonCen, the normal-only Hancock half
// step, then PLM (mh -/+ hs) or the spike PPM edges (edge + mh - c), floored.  lo is the state on
// the cell's lower D face, hi on its upper one; a face flux pairs hi of the lower cell with lo of
// the upper cell.
//
// NV = 8 + passive species fractions X_s (spike_ctm.cu -DNSP): their slopes and PPM edges are the
// primitives', and the normal-only half step advects them, X - (dt/dx / 2) un s with s = 2 hs the
// slope and un the cell's normal velocity (spike_ctm.cu:178-181).
//
// RC: the reconstruction fixed at compile time (RECON_PLM or RECON_PPM_SPIKE), or < 0 for a.recon.
template <typename T, int D, int NV = 8, int RC = -1, typename A>
GLM_DEV void ct_cell_states(const T l[NV], const T c[NV], const T r[NV], const A& a, T lo[NV], T hi[NV])
{
    T hs[NV], mh[NV];
    const T fac = a.face_slope == 1 ? T(0.25) : T(0.5);
#pragma unroll
    for (int k = 0; k < NV; ++k) hs[k] = fac * moncen(c[k] - l[k], r[k] - c[k]);
    ct_hanc1d<T, D>(c, hs, mh, a);
#pragma unroll
    for (int k = 8; k < NV; ++k) mh[k] = c[k] - a.half_dtdx * c[1 + D] * (T(2) * hs[k]);
    const bool ppm = RC < 0 ? a.recon == RECON_PPM_SPIKE : RC == RECON_PPM_SPIKE;
    if (ppm) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {   // the parabola's edges c - dl, c + dr plus mh - c
            T dl, dr;
            ppm_offsets_spike(l[k], c[k], r[k], dl, dr);
            hi[k] = mh[k] + dr;
            lo[k] = mh[k] - dl;
        }
    } else {
#pragma unroll
        for (int k = 0; k < NV; ++k) { hi[k] = mh[k] + hs[k]; lo[k] = mh[k] - hs[k]; }
    }
    hi[0] = tmax(hi[0], a.smallr); hi[4] = tmax(hi[4], a.smallp);
    lo[0] = tmax(lo[0], a.smallr); lo[4] = tmax(lo[4], a.smallp);
}

// Flux through a D face from the floored face states Lq (lower cell's hi) and Rq (upper cell's
// lo) with the face-normal field bn (ct_mhd.jl:114-122, spike_ctm.cu:235-245): F[Bn] = 0.
template <typename T, int RIEM, int D, typename A>
GLM_DEV void ct_riemann8(const T Lq[8], const T Rq[8], T bn, const A& a, T f[8])
{
    T L[8], R[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) { L[k] = Lq[k]; R[k] = Rq[k]; }
    L[5 + D] = bn; R[5 + D] = bn;
    if (RIEM == RIEM_HLLD) {
        // rotate D to x, Miyoshi-Kusano with the face-normal B and F[Bn] = 0, rotate back
        constexpr int A_ = D % 3, B_ = (D + 1) % 3, C_ = (D + 2) % 3;
        T Lr[9] = {L[0], L[1 + A_], L[1 + B_], L[1 + C_], L[4], L[5 + A_], L[5 + B_], L[5 + C_], T(0)};
        T Rr[9] = {R[0], R[1 + A_], R[1 + B_], R[1 + C_], R[4], R[5 + A_], R[5 + B_], R[5 + C_], T(0)};
        Phys<T> s;
        s.gamma = a.gamma; s.igm1 = a.igm1; s.smallr = a.smallr; s.pfl = a.smallp;
        s.ch = T(1); s.ch2 = T(1); s.half_ich = T(0.5); s.llf_dmin = T(0); s.llf_pmin = T(0);
        T fx[9];
        hlld_x(Lr, Rr, fx, s);
        f[0] = fx[0]; f[4] = fx[4];
        f[1 + A_] = fx[1]; f[1 + B_] = fx[2]; f[1 + C_] = fx[3];
        f[5 + A_] = T(0); f[5 + B_] = fx[6]; f[5 + C_] = fx[7];
        return;
    }
    T FL[8], FR[8], UL[8], UR[8];
    ct_dflux<T, D>(L, FL, a.igm1);
    ct_dflux<T, D>(R, FR, a.igm1);
    ct_toC(L, UL, a.igm1);
    ct_toC(R, UR, a.igm1);
    const T uL = L[1 + D], uR = R[1 + D];
    const T cL = ct_cf(L, L[5 + D], a.gamma), cR = ct_cf(R, R[5 + D], a.gamma);
    if (RIEM == RIEM_LLF) {
        T smax = tmax(tabs(uL) + cL, tabs(uR) + cR);
#pragma unroll
        for (int k = 0; k < 8; ++k) f[k] = T(0.5) * (FL[k] + FR[k]) - T(0.5) * smax * (UR[k] - UL[k]);
    } else {
        T SL = tmin(tmin(uL - cL, uR - cR), T(0)), SR = tmax(tmax(uL + cL, uR + cR), T(0));
        T ih = trcp(SR - SL);
#pragma unroll
        for (int k = 0; k < 8; ++k) f[k] = (SR * FL[k] - SL * FR[k] + SL * SR * (UR[k] - UL[k])) * ih;
    }
}

// ct_riemann8 plus the species fluxes of an NV-value state: F_s = F_rho X_s of the upwind face
// state (the consistent multi-fluid advection of spike_ctm.cu:240-243).
template <typename T, int RIEM, int D, int NV = 8, typename A>
GLM_DEV void ct_riemann(const T Lq[NV], const T Rq[NV], T bn, const A& a, T f[NV])
{
    ct_riemann8<T, RIEM, D>(Lq, Rq, bn, a, f);
#pragma unroll
    for (int k = 8; k < NV; ++k) f[k] = f[0] * (f[0] >= T(0) ? Lq[k] : Rq[k]);
}

// Flux at the lower-D face from the stencil m2, m1, c0, p1 (ct_mhd.jl:99-122 / spike_ctm.cu:223-245).
template <typename T, int D>
GLM_DEV void ct_face_flux(const T m2[8], const T m1[8], const T c0[8], const T p1[8], T bn, const CTArgs<T>& a, T f[8])
{
    T lo[8], L[8], R[8], hi[8];
    ct_cell_states<T, D>(m2, m1, c0, a, lo, L);
    ct_cell_states<T, D>(m1, c0, p1, a, R, hi);
    if (a.riemann == RIEM_HLLD) ct_riemann<T, RIEM_HLLD, D>(L, R, bn, a, f);
    else if (a.riemann == RIEM_LLF) ct_riemann<T, RIEM_LLF, D>(L, R, bn, a, f);
    else ct_riemann<T, RIEM_HLL, D>(L, R, bn, a, f);
}

// Update epilogue of one cell (ct_mhd.jl has none; the GLM rules of glmmhd_turb.jl:448-475): the void
// floor with the cell-centred field (bcx, bcy, bcz) and the OU kick with its energy floors.
template <typename T>
GLM_DEV void ct_epilogue(T r[5], T bcx, T bcy, T bcz, int i, int j, int k, const CTArgs<T>& a)
{
    if (a.update_floor && r[0] < a.smallr) {
        T emag = T(0.5) * (bcx * bcx + bcy * bcy + bcz * bcz);
        r[0] = a.smallr; r[1] = T(0); r[2] = T(0); r[3] = T(0); r[4] = a.smallp * a.igm1 + emag;
    }
    if (a.do_turb && r[0] >= a.turb_min_rho) {
        const int g = a.tgs;
        const T gg = T(g);
        T cx = ((T(i) + T(0.5)) * a.dx) / a.lx * gg, cy = ((T(j) + T(0.5)) * a.dx) / a.ly * gg;
        T cz = ((T(k) + T(0.5)) * a.dx) / a.lz * gg;
        T fx = floor(cx), fy = floor(cy), fz = floor(cz);
        int i0 = (int)fx, j0 = (int)fy, k0 = (int)fz;
        T wx = cx - fx, wy = cy - fy, wz = cz - fz;
        int i0w = cwrap(i0, g), i1w = cwrap(i0 + 1, g), j0w = cwrap(j0, g), j1w = cwrap(j0 + 1, g);
        int k0w = cwrap(k0, g), k1w = cwrap(k0 + 1, g);
        size_t g3 = (size_t)g * g * g;
        T acc[3];
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) {
            const T* af = a.afield + cc * g3;
            auto at = [&](int x, int y, int z) { return af[(size_t)x + (size_t)g * ((size_t)y + (size_t)g * z)]; };
            T t = (((at(i0w, j0w, k0w) * (T(1) - wx) + at(i1w, j0w, k0w) * wx) * (T(1) - wy) +
                    (at(i0w, j1w, k0w) * (T(1) - wx) + at(i1w, j1w, k0w) * wx) * wy) * (T(1) - wz) +
                   ((at(i0w, j0w, k1w) * (T(1) - wx) + at(i1w, j0w, k1w) * wx) * (T(1) - wy) +
                    (at(i0w, j1w, k1w) * (T(1) - wx) + at(i1w, j1w, k1w) * wx) * wy) * wz);
            acc[cc] = a.ramp * t;
        }
        T rho = r[0], rhom = tmax(rho, a.smallr), irm = trcp(rhom), fl = rho * a.smallc * a.smallc, e = r[4];
        e = tmax(e - T(0.5) * r[1] * r[1] * irm, fl);
        e = tmax(e - T(0.5) * r[2] * r[2] * irm, fl);
        e = tmax(e - T(0.5) * r[3] * r[3] * irm, fl);
        T m2 = r[1] + rhom * acc[0] * a.dt, m3 = r[2] + rhom * acc[1] * a.dt, m4 = r[3] + rhom * acc[2] * a.dt;
        e = tmax(e + T(0.5) * m2 * m2 * irm, fl);
        e = tmax(e + T(0.5) * m3 * m3 * irm, fl);
        e = tmax(e + T(0.5) * m4 * m4 * irm, fl);
        r[1] = m2; r[2] = m3; r[3] = m4; r[4] = e;
    }
}

// Pass 1: cell-centred B and primitives (ct_mhd.jl:85-96 bcc, :26-34 cellprim)
template <typename T>
__global__ void __launch_bounds__(CT_NT) ct_prim_kernel(CTArgs<T> a)
{
    size_t n = (size_t)a.nx * a.ny * a.nz;
    size_t c = (size_t)blockIdx.x * CT_NT + threadIdx.x;
    if (c >= n) return;
    int i = (int)(c % a.nx), j = (int)((c / a.nx) % a.ny), k = (int)(c / ((size_t)a.nx * a.ny));
    T bx = T(0.5) * (a.q[5][c] + a.q[5][cidx(a, cwrap(i + 1, a.nx), j, k)]);
    T by = T(0.5) * (a.q[6][c] + a.q[6][cidx(a, i, cwrap(j + 1, a.ny), k)]);
    T bz = T(0.5) * (a.q[7][c] + a.q[7][cidx(a, i, j, cwrap(k + 1, a.nz))]);
    T rho = tmax(a.q[0][c], a.smallr), ir = trcp(rho);
    T vx = a.q[1][c] * ir, vy = a.q[2][c] * ir, vz = a.q[3][c] * ir;
    T p = ct_pressure(a.q[4][c], rho, vx, vy, vz, bx, by, bz, a.gamma, a.smallp);
    T q8[8] = {rho, vx, vy, vz, p, bx, by, bz};
#pragma unroll
    for (int t = 0; t < 8; ++t) a.prim[(size_t)t * n + c] = q8[t];
}

// Pass 2: lower-face fluxes of each cell, one Riemann solve per face
template <typename T>
__global__ void __launch_bounds__(CT_NT) ct_flux_kernel(CTArgs<T> a)
{
    size_t n = (size_t)a.nx * a.ny * a.nz;
    size_t c = (size_t)blockIdx.x * CT_NT + threadIdx.x;
    if (c >= n) return;
    int i = (int)(c % a.nx), j = (int)((c / a.nx) % a.ny), k = (int)(c / ((size_t)a.nx * a.ny));
    auto ld = [&](size_t cc, T q[8]) {
#pragma unroll
        for (int t = 0; t < 8; ++t) q[t] = a.prim[(size_t)t * n + cc];
    };
    T m2[8], m1[8], c0[8], p1[8], f[8];
    ld(c, c0);
    // x
    ld(cidx(a, cwrap(i - 2, a.nx), j, k), m2); ld(cidx(a, cwrap(i - 1, a.nx), j, k), m1);
    ld(cidx(a, cwrap(i + 1, a.nx), j, k), p1);
    ct_face_flux<T, 0>(m2, m1, c0, p1, a.q[5][c], a, f);
#pragma unroll
    for (int t = 0; t < 8; ++t) a.flx[(size_t)(0 * 8 + t) * n + c] = f[t];
    // y
    ld(cidx(a, i, cwrap(j - 2, a.ny), k), m2); ld(cidx(a, i, cwrap(j - 1, a.ny), k), m1);
    ld(cidx(a, i, cwrap(j + 1, a.ny), k), p1);
    ct_face_flux<T, 1>(m2, m1, c0, p1, a.q[6][c], a, f);
#pragma unroll
    for (int t = 0; t < 8; ++t) a.flx[(size_t)(1 * 8 + t) * n + c] = f[t];
    // z
    ld(cidx(a, i, j, cwrap(k - 2, a.nz)), m2); ld(cidx(a, i, j, cwrap(k - 1, a.nz)), m1);
    ld(cidx(a, i, j, cwrap(k + 1, a.nz)), p1);
    ct_face_flux<T, 2>(m2, m1, c0, p1, a.q[7][c], a, f);
#pragma unroll
    for (int t = 0; t < 8; ++t) a.flx[(size_t)(2 * 8 + t) * n + c] = f[t];
}

// Pass 3: Balsara-Spicer edge EMFs (ct_mhd.jl:139-154)
template <typename T>
__global__ void __launch_bounds__(CT_NT) ct_emf_kernel(CTArgs<T> a)
{
    size_t n = (size_t)a.nx * a.ny * a.nz;
    size_t c = (size_t)blockIdx.x * CT_NT + threadIdx.x;
    if (c >= n) return;
    int i = (int)(c % a.nx), j = (int)((c / a.nx) % a.ny), k = (int)(c / ((size_t)a.nx * a.ny));
    size_t im = cidx(a, cwrap(i - 1, a.nx), j, k), jm = cidx(a, i, cwrap(j - 1, a.ny), k);
    size_t km = cidx(a, i, j, cwrap(k - 1, a.nz));
    const T* F = a.flx;
#define CF(d, t, cc) F[(size_t)((d) * 8 + (t)) * n + (cc)]
    a.emf[2 * n + c] = T(0.25) * (CF(1, 5, c) + CF(1, 5, im) - CF(0, 6, c) - CF(0, 6, jm));
    a.emf[c] = T(0.25) * (CF(2, 6, c) + CF(2, 6, jm) - CF(1, 7, c) - CF(1, 7, km));
    a.emf[n + c] = T(0.25) * (CF(0, 7, c) + CF(0, 7, km) - CF(2, 5, c) - CF(2, 5, im));
#undef CF
}

// Pass 4: hydro update + face-B curl update (+ floor, OU kick, next-step CFL)
template <typename T>
__global__ void __launch_bounds__(CT_NT) ct_update_kernel(CTArgs<T> a)
{
    size_t n = (size_t)a.nx * a.ny * a.nz;
    size_t c = (size_t)blockIdx.x * CT_NT + threadIdx.x;
    T ctv = T(0);
    if (c < n) {
        int i = (int)(c % a.nx), j = (int)((c / a.nx) % a.ny), k = (int)(c / ((size_t)a.nx * a.ny));
        size_t ip = cidx(a, cwrap(i + 1, a.nx), j, k), jp = cidx(a, i, cwrap(j + 1, a.ny), k);
        size_t kp = cidx(a, i, j, cwrap(k + 1, a.nz));
        const T* F = a.flx;
#define CF(d, t, cc) F[(size_t)((d) * 8 + (t)) * n + (cc)]
        T r[5];
#pragma unroll
        for (int t = 0; t < 5; ++t)
            r[t] = a.q[t][c] + a.dtdx * ((CF(0, t, c) - CF(0, t, ip)) + (CF(1, t, c) - CF(1, t, jp)) + (CF(2, t, c) - CF(2, t, kp)));
#undef CF
        const T *ex = a.emf, *ey = a.emf + n, *ez = a.emf + 2 * n;
        T bxn = a.q[5][c] - a.dtdx * ((ez[jp] - ez[c]) - (ey[kp] - ey[c]));
        T byn = a.q[6][c] - a.dtdx * ((ex[kp] - ex[c]) - (ez[ip] - ez[c]));
        T bzn = a.q[7][c] - a.dtdx * ((ey[ip] - ey[c]) - (ex[jp] - ex[c]));
        ct_epilogue(r, a.prim[5 * n + c], a.prim[6 * n + c], a.prim[7 * n + c], i, j, k, a);
#pragma unroll
        for (int t = 0; t < 5; ++t) a.o[t][c] = r[t];
        a.o[5][c] = bxn; a.o[6][c] = byn; a.o[7][c] = bzn;
    }
    (void)ctv;
}

// CFL of a CT state from cell-centred B (the GLM rule, glmmhd_turb.jl:1963-1971)
template <typename T>
__global__ void __launch_bounds__(CT_NT) ct_ctot_kernel(Ptr9<const T> q, int nx, int ny, int nz, T gamma, T smallr,
                                                        T smallp, unsigned long long* out)
{
    size_t n = (size_t)nx * ny * nz;
    T cmax = T(0);
    for (size_t c = (size_t)blockIdx.x * CT_NT + threadIdx.x; c < n; c += (size_t)gridDim.x * CT_NT) {
        int i = (int)(c % nx), j = (int)((c / nx) % ny), k = (int)(c / ((size_t)nx * ny));
        size_t pl = (size_t)nx * ny;
        T bx = T(0.5) * (q[5][c] + q[5][(size_t)cwrap(i + 1, nx) + (size_t)nx * j + pl * k]);
        T by = T(0.5) * (q[6][c] + q[6][(size_t)i + (size_t)nx * cwrap(j + 1, ny) + pl * k]);
        T bz = T(0.5) * (q[7][c] + q[7][(size_t)i + (size_t)nx * j + pl * cwrap(k + 1, nz)]);
        T rho = tmax(q[0][c], smallr), ir = trcp(rho);
        T vx = q[1][c] * ir, vy = q[2][c] * ir, vz = q[3][c] * ir;
        T p = tmax((gamma - T(1)) * (q[4][c] - T(0.5) * rho * (vx * vx + vy * vy + vz * vz) -
                                     T(0.5) * (bx * bx + by * by + bz * bz)),
                   smallp);
        T p8[8] = {rho, vx, vy, vz, p, bx, by, bz};
        T ct = (tabs(vx) + ct_cf(p8, bx, gamma)) + (tabs(vy) + ct_cf(p8, by, gamma)) + (tabs(vz) + ct_cf(p8, bz, gamma));
        if (!(ct <= cmax)) cmax = ct;
    }
    unsigned long long bits = (unsigned long long)__double_as_longlong((double)cmax);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        unsigned long long o = __shfl_xor(bits, off, 64);
        bits = o > bits ? o : bits;
    }
    if ((threadIdx.x & 63) == 0 && bits != 0ull) atomicMax(out, bits);
}

// max |div B| from face fields (ct_mhd.jl:191-194)
template <typename T>
__global__ void __launch_bounds__(CT_NT) ct_divb_kernel(Ptr9<const T> q, int nx, int ny, int nz, double dx,
                                                        unsigned long long* out)
{
    size_t n = (size_t)nx * ny * nz;
    double m = 0.0;
    for (size_t c = (size_t)blockIdx.x * CT_NT + threadIdx.x; c < n; c += (size_t)gridDim.x * CT_NT) {
        int i = (int)(c % nx), j = (int)((c / nx) % ny), k = (int)(c / ((size_t)nx * ny));
        size_t pl = (size_t)nx * ny;
        double d = (((double)q[5][(size_t)cwrap(i + 1, nx) + (size_t)nx * j + pl * k] - (double)q[5][c]) +
                    ((double)q[6][(size_t)i + (size_t)nx * cwrap(j + 1, ny) + pl * k] - (double)q[6][c]) +
                    ((double)q[7][(size_t)i + (size_t)nx * j + pl * cwrap(k + 1, nz)] - (double)q[7][c])) / dx;
        d = fabs(d);
        if (!(d <= m)) m = d;
    }
    unsigned long long bits = (unsigned long long)__double_as_longlong(m);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        unsigned long long o = __shfl_xor(bits, off, 64);
        bits = o > bits ? o : bits;
    }
    if ((threadIdx.x & 63) == 0 && bits != 0ull) atomicMax(out, bits);
}

}  // namespace glm
