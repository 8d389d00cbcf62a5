/*
 * glm_oracle.c -- CPU restatement of GLMMHDTurb's GLM-MHD line-march update.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle: it is built into
 * oracle/_build/liboracle_{f64,f32}.so and called (through oracle/oracle.py) only by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg.  The product
 * (glmmhdturb_amd/) never links, imports or executes it.
 *
 * It is deliberately written in a different shape from the GPU kernels: a plain
 * multi-pass step (c2p -> per-cell reconstruction and face states -> one Riemann solve per
 * face -> conservative update), as in the reference's validation path step_multi!
 * (glmmhd_turb.jl:1981-1992), with OpenMP over z-planes.  Each function cites the reference
 * lines it restates.  Paths are relative to the reference repository root.
 *
 * Precision: compiled twice, with -DREAL=double (FP64 oracle) and -DREAL=float.
 * Constants are written through R() so that the float build stays in float.
 *
 * State layout: 9 conserved planes, SoA, plane v at u + v*n, cell (i,j,k) at
 * i + nx*(j + ny*k) -- the bridge layout of march_bridge/cu/spike_mhd.cu:48 and the
 * memory layout of the Julia (N,N,N,9) array (glmmhd_turb.jl:2051).
 * Conserved order (rho, mx, my, mz, E, Bx, By, Bz, psi); primitive order
 * (rho, vx, vy, vz, p, Bx, By, Bz, psi)  (glmmhd_turb.jl:61-64).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <tgmath.h>

#ifndef REAL
#define REAL double
#endif
typedef REAL real;
#define R(x) ((real)(x))
#define NV 9

#if defined(_OPENMP)
#include <omp.h>
#endif

/* ------------------------------------------------------------------------- */
/* Scheme and run parameters.  All fields are double in the ABI; converted to   */
/* real at use.  Field meaning follows Params (glmmhd_turb.jl:28-55).          */
/* ------------------------------------------------------------------------- */
typedef struct {
    double gamma, smallr, smallc, pfl;
    double dt, dx, ch, glm_fac;
    double llf_dmin, llf_pmin;
    double turb_min_rho, ramp, boxlen;
    int predictor;    /* 0 full-transverse Hancock (glmmhd_turb.jl:165-172); 1 normal-only (spike_mhd.cu:149-161) */
    int recon;        /* 0 PLM; 1 PPM Julia form (glmmhd_turb.jl:917-926); 2 PPM spike form (spike_mhd.cu:162-192) */
    int slope;        /* 0 MonCen (glmmhd_turb.jl:145-155); 1 minmod (glmmhd_turb.jl:300-303) */
    int riemann;      /* 0 LLF (:182-190); 1 LLF upper-bound (:311-331); 2 HLL (:335-354); 3 HLLD (:193-292) */
    int mh_revert;    /* predictor revert mh->m0 when mh rho<=smallr or p<=pfl (glmmhd_turb.jl:393-395) */
    int face_guard;   /* 0 none (cpu_simd.jl); 1 iflo revert to m0 (glmmhd_turb.jl:618); 2 clamp (spike_mhd.cu:145-147) */
    int update_floor; /* static void-cell floor in the update (glmmhd_turb.jl:448-453) */
    int do_turb;      /* fused OU kick (glmmhd_turb.jl:456-475) */
    int turb_gs;      /* forcing grid size (TURB_GS, glmmhd_turb.jl:58) */
    int pad_;
} o_scheme;

/* Branch counters: how often each guarded branch of the scheme fired since the last reset.  Tests
 * read them to show that a run really exercised the branch they compare.
 * 0 HLLD degenerate star (|denK| < 1e-12), 1 predictor revert mh -> m0, 2 iflo face revert,
 * 3 update void floor, 4 HLLD -> LLF switch. */
#define O_NHITS 5
static long long o_hits[O_NHITS];
#define O_HIT(w) do { _Pragma("omp atomic") o_hits[w]++; } while (0)
long long o_branch_hits(int which) { return (which >= 0 && which < O_NHITS) ? o_hits[which] : -1; }
void o_branch_hits_reset(void) { for (int w = 0; w < O_NHITS; ++w) o_hits[w] = 0; }

/* Branch trace (built only with -DO_BRANCH_TRACE, liboracle_trace.so): for every decision point of
 * the step (limiter case, PPM case, wave-speed min/max choice, HLLD fan, floors, reverts) the count
 * of each outcome.  Two runs whose traces are equal took the same branches everywhere (up to
 * cancelling swaps), so a finite difference between them sees a smooth map.  Sites:
 *  0 c2p rho floor   1 c2p p floor   2 fast-speed radicand clamp   3 MonCen (0 zero, 1 2|dl|, 2 2|dr|, 3 |dc|)
 *  4 minmod (0 zero, 1 a, 2 b)   5 PPM CW84 (0 flat, 1 lo reset, 2 hi reset, 3 none)   6 PPM spj
 *  7 spike PPM (as 5)   8 LLF speed (0 L, 1 R, 2 ch)   9 LLF-bound speed (as 8)
 * 10 HLL SL (0 L, 1 R, 2 zero)  11 HLL SR (as 10)  12 HLLD SL clamped at 0  13 HLLD SR clamped
 * 14 HLLD min(uL,uR) side  15 HLLD max(cfL,cfR) side  16 HLLD rho*_L floor  17 HLLD rho*_R floor
 * 18 HLLD degenerate star   19 HLLD fan (0..3)   20 HLLD sign(Bn)   21 interface floors (0 rhoL,
 * 1 pL, 2 rhoR, 3 pR)   22 HLLD->LLF switch   23 predictor revert   24 iflo face revert
 * 25 clamped face floors   26 update floor   27 kick density threshold   28 kick energy floor
 * (removal: component)   29 kick energy floor (addition: component)   30 spike edge floors */
#define O_NSITES 32
#define O_NOUT 4
#ifdef O_BRANCH_TRACE
/* One row of counters per OpenMP thread (summed by o_trace_get), so no atomics are needed; threads
 * past the last private row share it with atomic increments. */
#define O_TRACE_THREADS 256
static long long o_trace[O_TRACE_THREADS][O_NSITES][O_NOUT];
static inline void o_br(int site, int k)
{
    const int t = omp_get_thread_num();
    if (t < O_TRACE_THREADS - 1) {
        o_trace[t][site][k]++;
    } else {
        _Pragma("omp atomic") o_trace[O_TRACE_THREADS - 1][site][k]++;
    }
}
#define O_BR(site, k) o_br((site), (k))
#else
#define O_BR(site, k) ((void)0)
#endif
void o_trace_get(long long* out)
{
    for (int a = 0; a < O_NSITES * O_NOUT; ++a) {
#ifdef O_BRANCH_TRACE
        long long s = 0;
        for (int t = 0; t < O_TRACE_THREADS; ++t) s += o_trace[t][a / O_NOUT][a % O_NOUT];
        out[a] = s;
#else
        out[a] = -1;
#endif
    }
}
void o_trace_reset(void)
{
#ifdef O_BRANCH_TRACE
    memset(o_trace, 0, sizeof o_trace);
#endif
}

/* ------------------------------------------------------------------------- */
/* Point physics                                                               */
/* ------------------------------------------------------------------------- */

/* glmmhd_turb.jl:68-77 cons2prim */
void o_cons2prim(const real* c, real* q, real gamma, real smallr, real pfl)
{
    real rho = fmax(c[0], smallr);
    O_BR(0, c[0] < smallr);
    real ir = R(1) / rho;
    real vx = c[1] * ir, vy = c[2] * ir, vz = c[3] * ir;
    real bx = c[5], by = c[6], bz = c[7];
    real ekin = R(0.5) * (c[1] * c[1] + c[2] * c[2] + c[3] * c[3]) * ir;
    real emag = R(0.5) * (bx * bx + by * by + bz * bz);
    real p = fmax((gamma - R(1)) * (c[4] - ekin - emag), pfl);
    O_BR(1, (gamma - R(1)) * (c[4] - ekin - emag) < pfl);
    q[0] = rho; q[1] = vx; q[2] = vy; q[3] = vz; q[4] = p;
    q[5] = bx; q[6] = by; q[7] = bz; q[8] = c[8];
}

/* glmmhd_turb.jl:79-86 prim2cons */
void o_prim2cons(const real* q, real* c, real gamma)
{
    real rho = q[0], vx = q[1], vy = q[2], vz = q[3], p = q[4];
    real bx = q[5], by = q[6], bz = q[7];
    real ekin = R(0.5) * rho * (vx * vx + vy * vy + vz * vz);
    real emag = R(0.5) * (bx * bx + by * by + bz * bz);
    c[0] = rho; c[1] = rho * vx; c[2] = rho * vy; c[3] = rho * vz;
    c[4] = p / (gamma - R(1)) + ekin + emag;
    c[5] = bx; c[6] = by; c[7] = bz; c[8] = q[8];
}

/* glmmhd_turb.jl:89-95 fast_speed (normal field bn) */
real o_fast_speed(const real* q, real gamma, real bn)
{
    real rho = q[0];
    real c2 = gamma * q[4] / rho;
    real b2 = (q[5] * q[5] + q[6] * q[6] + q[7] * q[7]) / rho;
    real d2 = R(0.5) * (b2 + c2);
    O_BR(2, d2 * d2 - c2 * bn * bn / rho < R(0));
    return sqrt(d2 + sqrt(fmax(d2 * d2 - c2 * bn * bn / rho, R(0))));
}

/* glmmhd_turb.jl:98-115 phys_flux_x; F[Bx] and F[psi] are zero (set by the GLM pair) */
void o_phys_flux_x(const real* q, real* f, real gamma)
{
    real rho = q[0], vx = q[1], vy = q[2], vz = q[3], p = q[4];
    real bx = q[5], by = q[6], bz = q[7];
    real b2 = bx * bx + by * by + bz * bz;
    real ptot = p + R(0.5) * b2;
    real ekin = R(0.5) * rho * (vx * vx + vy * vy + vz * vz);
    real E = p / (gamma - R(1)) + ekin + R(0.5) * b2;
    real vdotb = vx * bx + vy * by + vz * bz;
    f[0] = rho * vx;
    f[1] = rho * vx * vx + ptot - bx * bx;
    f[2] = rho * vx * vy - bx * by;
    f[3] = rho * vx * vz - bx * bz;
    f[4] = (E + ptot) * vx - bx * vdotb;
    f[5] = R(0);
    f[6] = vx * by - vy * bx;
    f[7] = vx * bz - vz * bx;
    f[8] = R(0);
}

/* glmmhd_turb.jl:119-127 rot_to: dir 0=x identity, 1=y ((x,y,z)->(y,z,x)), 2=z */
void o_rot_to(const real* q, real* r, int dir)
{
    memcpy(r, q, sizeof(real) * NV);
    if (dir == 1) {
        r[1] = q[2]; r[2] = q[3]; r[3] = q[1];
        r[5] = q[6]; r[6] = q[7]; r[7] = q[5];
    } else if (dir == 2) {
        r[1] = q[3]; r[2] = q[1]; r[3] = q[2];
        r[5] = q[7]; r[6] = q[5]; r[7] = q[6];
    }
}

/* glmmhd_turb.jl:130-138 rot_flux_from */
void o_rot_flux_from(const real* f, real* r, int dir)
{
    memcpy(r, f, sizeof(real) * NV);
    if (dir == 1) {
        r[1] = f[3]; r[2] = f[1]; r[3] = f[2];
        r[5] = f[7]; r[6] = f[5]; r[7] = f[6];
    } else if (dir == 2) {
        r[1] = f[2]; r[2] = f[3]; r[3] = f[1];
        r[5] = f[6]; r[6] = f[7]; r[7] = f[5];
    }
}

/* glmmhd_turb.jl:161-162 dir_flux */
static void dir_flux(const real* q, real* f, int dir, real gamma)
{
    real r[NV], fx[NV];
    o_rot_to(q, r, dir);
    o_phys_flux_x(r, fx, gamma);
    o_rot_flux_from(fx, f, dir);
}

/* glmmhd_turb.jl:145-150 moncen */
real o_moncen(real dl, real dr)
{
    real dc = R(0.5) * (dl + dr);
    real sgn = dc >= R(0) ? R(1) : R(-1);
    real val = sgn * fmin(R(2) * fmin(fabs(dl), fabs(dr)), fabs(dc));
    O_BR(3, (dl * dr <= R(0)) ? 0 : (R(2) * fmin(fabs(dl), fabs(dr)) < fabs(dc) ? (fabs(dl) <= fabs(dr) ? 1 : 2) : 3));
    return (dl * dr <= R(0)) ? R(0) : val;
}

/* glmmhd_turb.jl:300-301 minmod */
real o_minmod(real a, real b)
{
    O_BR(4, (a * b <= R(0)) ? 0 : (fabs(a) < fabs(b) ? 1 : 2));
    return (a * b <= R(0)) ? R(0) : (fabs(a) < fabs(b) ? a : b);
}

/* Half slope of a primitive vector: glmmhd_turb.jl:153-155 (MonCen) and :302-303 (minmod) */
static void prim_slope(const real* L, const real* M, const real* Rr, real* s, int lim)
{
    for (int v = 0; v < NV; ++v) {
        real dl = M[v] - L[v], dr = Rr[v] - M[v];
        s[v] = R(0.5) * (lim == 1 ? o_minmod(dl, dr) : o_moncen(dl, dr));
    }
}

/* glmmhd_turb.jl:165-172 hancock (full transverse predictor); returns conserved U_half */
void o_hancock(const real* m0, const real* sx, const real* sy, const real* sz,
               real dtdx, real gamma, real* uh)
{
    real u0[NV], a[NV], b[NV];
    real fxp[NV], fxm[NV], fyp[NV], fym[NV], fzp[NV], fzm[NV];
    o_prim2cons(m0, u0, gamma);
    real h = R(0.5) * dtdx;
    for (int v = 0; v < NV; ++v) { a[v] = m0[v] + sx[v]; b[v] = m0[v] - sx[v]; }
    dir_flux(a, fxp, 0, gamma); dir_flux(b, fxm, 0, gamma);
    for (int v = 0; v < NV; ++v) { a[v] = m0[v] + sy[v]; b[v] = m0[v] - sy[v]; }
    dir_flux(a, fyp, 1, gamma); dir_flux(b, fym, 1, gamma);
    for (int v = 0; v < NV; ++v) { a[v] = m0[v] + sz[v]; b[v] = m0[v] - sz[v]; }
    dir_flux(a, fzp, 2, gamma); dir_flux(b, fzm, 2, gamma);
    for (int v = 0; v < NV; ++v)
        uh[v] = u0[v] - h * ((fxp[v] - fxm[v]) + (fyp[v] - fym[v]) + (fzp[v] - fzm[v]));
}

/* spike_mhd.cu:149-161 hanc1d: normal-only half step in direction dir.  `s` is the half slope
 * (the .cu passes the full slope and evaluates c0 -/+ 0.5 s).  Returns the primitive mh
 * (floored by cons2prim, no revert). */
void o_hanc1d(const real* c0, const real* s, int dir, real dtdx, real gamma,
              real smallr, real pfl, real* mh)
{
    real a[NV], b[NV], fl[NV], fr[NV], u[NV];
    for (int v = 0; v < NV; ++v) { a[v] = c0[v] - s[v]; b[v] = c0[v] + s[v]; }
    dir_flux(a, fl, dir, gamma);
    dir_flux(b, fr, dir, gamma);
    o_prim2cons(c0, u, gamma);
    real h = R(0.5) * dtdx;
    for (int v = 0; v < NV; ++v) u[v] -= h * (fr[v] - fl[v]);
    o_cons2prim(u, mh, gamma, smallr, pfl);
}

/* glmmhd_turb.jl:917-922 ppm_mono (CW84 monotonize), returns (lo, hi) */
void o_ppm_mono(real ql, real q0, real qr, real* lo, real* hi)
{
    real dq = qr - ql;
    real diff = (q0 - R(0.5) * (ql + qr)) * dq;
    if ((qr - q0) * (q0 - ql) <= R(0)) { O_BR(5, 0); *lo = q0; *hi = q0; }
    else if (diff > dq * dq / R(6)) { O_BR(5, 1); *lo = R(3) * q0 - R(2) * qr; *hi = qr; }
    else if (diff < -(dq * dq / R(6))) { O_BR(5, 2); *lo = ql; *hi = R(3) * q0 - R(2) * ql; }
    else { O_BR(5, 3); *lo = ql; *hi = qr; }
}

/* glmmhd_turb.jl:923 ppm_edges */
void o_ppm_edges(real qm, real q0, real qp, real* lo, real* hi)
{
    real slope = R(0.25) * (qp - qm);
    real curve = (qm - R(2) * q0 + qp) / R(12);
    o_ppm_mono(q0 - slope + curve, q0, q0 + slope + curve, lo, hi);
}

/* glmmhd_turb.jl:926 spj (strong pressure jump -> PLM fallback) */
int o_spj(real a, real b, real c)
{
    real hi = fmax(a, fmax(b, c));
    real lo = fmax(fmin(a, fmin(b, c)), R(1e-20));
    O_BR(6, hi / lo > R(2));
    return hi / lo > R(2);
}

/* spike_mhd.cu:164-172 ppm1 (sgn>0 -> right edge, else left edge) */
real o_ppm1_spike(real qm, real q0, real qp, real sgn)
{
    real d = R(0.25) * (qp - qm), cu = (qm - R(2) * q0 + qp) * (R(1) / R(6));
    real qr = q0 + d - cu, ql = q0 - d - cu;
    if ((qr - q0) * (q0 - ql) <= R(0)) { O_BR(7, 0); qr = q0; ql = q0; }
    else {
        real dqe = qr - ql, d6 = R(6) * (q0 - R(0.5) * (ql + qr));
        O_BR(7, dqe * d6 > dqe * dqe ? 1 : (dqe * d6 < -dqe * dqe ? 2 : 3));
        if (dqe * d6 > dqe * dqe) ql = R(3) * q0 - R(2) * qr;
        if (dqe * d6 < -dqe * dqe) qr = R(3) * q0 - R(2) * ql;
    }
    return sgn > R(0) ? qr : ql;
}

/* glmmhd_turb.jl:175-179 glm_pair */
void o_glm_pair(real bnL, real bnR, real psiL, real psiR, real ch, real* bns, real* psis)
{
    *bns = R(0.5) * (bnL + bnR) - R(0.5) * (psiR - psiL) / ch;
    *psis = R(0.5) * (psiL + psiR) - R(0.5) * ch * (bnR - bnL);
}

/* glmmhd_turb.jl:182-190 llf_x (exact fast speed) */
static void llf_x(const real* L, const real* Rq, real gamma, real ch, real fbn, real fpsi, real* f)
{
    real fL[NV], fR[NV], uL[NV], uR[NV];
    o_phys_flux_x(L, fL, gamma); o_phys_flux_x(Rq, fR, gamma);
    o_prim2cons(L, uL, gamma); o_prim2cons(Rq, uR, gamma);
    real sa = fabs(L[1]) + o_fast_speed(L, gamma, L[5]), sb = fabs(Rq[1]) + o_fast_speed(Rq, gamma, Rq[5]);
    real smax = fmax(fmax(sa, sb), ch);
    O_BR(8, fmax(sa, sb) < ch ? 2 : (sa >= sb ? 0 : 1));
    for (int v = 0; v < NV; ++v) f[v] = R(0.5) * (fL[v] + fR[v]) - R(0.5) * smax * (uR[v] - uL[v]);
    f[5] = fbn; f[8] = fpsi;
}

/* glmmhd_turb.jl:311-314 fast_speed_bound and :316-324 llf_bound */
static real fast_speed_bound(const real* q, real gamma)
{
    real ir = R(1) / q[0];
    return sqrt((gamma * q[4] + q[5] * q[5] + q[6] * q[6] + q[7] * q[7]) * ir);
}
static void llf_bound(const real* L, const real* Rq, real gamma, real ch, real fbn, real fpsi, real* f)
{
    real fL[NV], fR[NV], uL[NV], uR[NV];
    o_phys_flux_x(L, fL, gamma); o_phys_flux_x(Rq, fR, gamma);
    o_prim2cons(L, uL, gamma); o_prim2cons(Rq, uR, gamma);
    real sa = fabs(L[1]) + fast_speed_bound(L, gamma), sb = fabs(Rq[1]) + fast_speed_bound(Rq, gamma);
    real smax = fmax(fmax(sa, sb), ch);
    O_BR(9, fmax(sa, sb) < ch ? 2 : (sa >= sb ? 0 : 1));
    for (int v = 0; v < NV; ++v) f[v] = R(0.5) * (fL[v] + fR[v]) - R(0.5) * smax * (uR[v] - uL[v]);
    f[5] = fbn; f[8] = fpsi;
}

/* glmmhd_turb.jl:335-343 hll_x */
static void hll_x(const real* L, const real* Rq, real gamma, real bn, real fbn, real fpsi, real* f)
{
    real cfL = o_fast_speed(L, gamma, bn), cfR = o_fast_speed(Rq, gamma, bn);
    real SL = fmin(fmin(L[1] - cfL, Rq[1] - cfR), R(0));
    real SR = fmax(fmax(L[1] + cfL, Rq[1] + cfR), R(0));
    O_BR(10, fmin(L[1] - cfL, Rq[1] - cfR) > R(0) ? 2 : (L[1] - cfL <= Rq[1] - cfR ? 0 : 1));
    O_BR(11, fmax(L[1] + cfL, Rq[1] + cfR) < R(0) ? 2 : (L[1] + cfL >= Rq[1] + cfR ? 0 : 1));
    real fL[NV], fR[NV], uL[NV], uR[NV];
    o_phys_flux_x(L, fL, gamma); o_phys_flux_x(Rq, fR, gamma);
    o_prim2cons(L, uL, gamma); o_prim2cons(Rq, uR, gamma);
    real ihd = R(1) / (SR - SL);
    for (int v = 0; v < NV; ++v) f[v] = (SR * fL[v] - SL * fR[v] + SL * SR * (uR[v] - uL[v])) * ihd;
    f[5] = fbn; f[8] = fpsi;
}

/* glmmhd_turb.jl:227-239 star (Miyoshi-Kusano 2005 eqs. 42-47 with a degenerate guard) */
static void hlld_star(real d, real u, real v, real w, real by, real bz, real S, real SM, real bn,
                      real* vy, real* vz, real* byS, real* bzS)
{
    real denK = d * (S - u) * (S - SM) - bn * bn;
    O_BR(18, fabs(denK) < R(1e-12));
    if (fabs(denK) < R(1e-12)) { O_HIT(0); *vy = v; *vz = w; *byS = by; *bzS = bz; }
    else {
        real fac = bn * (SM - u) / denK;
        *vy = v - by * fac;
        *vz = w - bz * fac;
        *byS = by * (d * (S - u) * (S - u) - bn * bn) / denK;
        *bzS = bz * (d * (S - u) * (S - u) - bn * bn) / denK;
    }
}

/* glmmhd_turb.jl:193-272 hlld_x (L, R primitive with Bx already the cleaned bn*) */
void o_hlld_x(const real* L, const real* Rq, real gamma, real bn, real fbn, real fpsi, real* f)
{
    real dL = L[0], uL = L[1], vL = L[2], wL = L[3], pL = L[4], byL = L[6], bzL = L[7];
    real dR = Rq[0], uR = Rq[1], vR = Rq[2], wR = Rq[3], pR = Rq[4], byR = Rq[6], bzR = Rq[7];
    real b2L = bn * bn + byL * byL + bzL * bzL;
    real b2R = bn * bn + byR * byR + bzR * bzR;
    real ptL = pL + R(0.5) * b2L, ptR = pR + R(0.5) * b2R;
    real EL = pL / (gamma - R(1)) + R(0.5) * dL * (uL * uL + vL * vL + wL * wL) + R(0.5) * b2L;
    real ER = pR / (gamma - R(1)) + R(0.5) * dR * (uR * uR + vR * vR + wR * wR) + R(0.5) * b2R;
    real qLc[NV] = {dL, uL, vL, wL, pL, bn, byL, bzL, R(0)};
    real qRc[NV] = {dR, uR, vR, wR, pR, bn, byR, bzR, R(0)};
    real cfL = o_fast_speed(L, gamma, bn), cfR = o_fast_speed(Rq, gamma, bn);
    real SL = fmin(fmin(uL, uR) - fmax(cfL, cfR), R(0));
    real SR = fmax(fmax(uL, uR) + fmax(cfL, cfR), R(0));
    O_BR(12, fmin(uL, uR) - fmax(cfL, cfR) > R(0));
    O_BR(13, fmax(uL, uR) + fmax(cfL, cfR) < R(0));
    O_BR(14, uL <= uR);
    O_BR(15, cfL >= cfR);
    real UL[NV] = {dL, dL * uL, dL * vL, dL * wL, EL, bn, byL, bzL, R(0)};
    real UR[NV] = {dR, dR * uR, dR * vR, dR * wR, ER, bn, byR, bzR, R(0)};
    real FL[NV], FR[NV];
    o_phys_flux_x(qLc, FL, gamma);
    o_phys_flux_x(qRc, FR, gamma);
    real den = (SR - uR) * dR - (SL - uL) * dL;
    real SM = ((SR - uR) * dR * uR - (SL - uL) * dL * uL - ptR + ptL) / den;
    real pts = ((SR - uR) * dR * ptL - (SL - uL) * dL * ptR + dL * dR * (SR - uR) * (SL - uL) * (uR - uL)) / den;
    real dLs = fmax(dL * (SL - uL) / (SL - SM), R(1e-12));
    real dRs = fmax(dR * (SR - uR) / (SR - SM), R(1e-12));
    O_BR(16, dL * (SL - uL) / (SL - SM) < R(1e-12));
    O_BR(17, dR * (SR - uR) / (SR - SM) < R(1e-12));
    real sqdLs = sqrt(dLs), sqdRs = sqrt(dRs);
    real SLs = SM - fabs(bn) / sqdLs;
    real SRs = SM + fabs(bn) / sqdRs;
    real vyLs, vzLs, byLs, bzLs, vyRs, vzRs, byRs, bzRs;
    hlld_star(dL, uL, vL, wL, byL, bzL, SL, SM, bn, &vyLs, &vzLs, &byLs, &bzLs);
    hlld_star(dR, uR, vR, wR, byR, bzR, SR, SM, bn, &vyRs, &vzRs, &byRs, &bzRs);
    real vdotbL = uL * bn + vL * byL + wL * bzL;
    real vdotbLs = SM * bn + vyLs * byLs + vzLs * bzLs;
    real ELs = ((SL - uL) * EL - ptL * uL + pts * SM + bn * (vdotbL - vdotbLs)) / (SL - SM);
    real vdotbR = uR * bn + vR * byR + wR * bzR;
    real vdotbRs = SM * bn + vyRs * byRs + vzRs * bzRs;
    real ERs = ((SR - uR) * ER - ptR * uR + pts * SM + bn * (vdotbR - vdotbRs)) / (SR - SM);
    real ULs[NV] = {dLs, dLs * SM, dLs * vyLs, dLs * vzLs, ELs, bn, byLs, bzLs, R(0)};
    real URs[NV] = {dRs, dRs * SM, dRs * vyRs, dRs * vzRs, ERs, bn, byRs, bzRs, R(0)};
    real sgn = bn >= R(0) ? R(1) : R(-1);
    O_BR(20, bn >= R(0));
    O_BR(19, SLs >= R(0) ? 0 : (SM >= R(0) ? 1 : (SRs >= R(0) ? 2 : 3)));
    real invsum = R(1) / (sqdLs + sqdRs);
    real vyss = (sqdLs * vyLs + sqdRs * vyRs + (byRs - byLs) * sgn) * invsum;
    real vzss = (sqdLs * vzLs + sqdRs * vzRs + (bzRs - bzLs) * sgn) * invsum;
    real byss = (sqdLs * byRs + sqdRs * byLs + sqdLs * sqdRs * (vyRs - vyLs) * sgn) * invsum;
    real bzss = (sqdLs * bzRs + sqdRs * bzLs + sqdLs * sqdRs * (vzRs - vzLs) * sgn) * invsum;
    real vdotbss = SM * bn + vyss * byss + vzss * bzss;
    real ELss = ELs - sqdLs * (vdotbLs - vdotbss) * sgn;
    real ERss = ERs + sqdRs * (vdotbRs - vdotbss) * sgn;
    real ULss[NV] = {dLs, dLs * SM, dLs * vyss, dLs * vzss, ELss, bn, byss, bzss, R(0)};
    real URss[NV] = {dRs, dRs * SM, dRs * vyss, dRs * vzss, ERss, bn, byss, bzss, R(0)};
    if (SLs >= R(0)) {
        for (int v = 0; v < NV; ++v) f[v] = FL[v] + SL * (ULs[v] - UL[v]);
    } else if (SM >= R(0)) {
        for (int v = 0; v < NV; ++v) f[v] = FL[v] + SLs * ULss[v] - (SLs - SL) * ULs[v] - SL * UL[v];
    } else if (SRs >= R(0)) {
        for (int v = 0; v < NV; ++v) f[v] = FR[v] + SRs * URss[v] - (SRs - SR) * URs[v] - SR * UR[v];
    } else {
        for (int v = 0; v < NV; ++v) f[v] = FR[v] + SR * (URs[v] - UR[v]);
    }
    f[5] = fbn; f[8] = fpsi;
}

/* Riemann dispatch in direction dir: rotate, floor interface states, GLM pair, solver, rotate
 * back.  riemann (glmmhd_turb.jl:275-292, with the LLF switch :287-290), riemann_hll
 * (:344-354), riemann_lean (:326-334).  kind: 0 llf, 1 llf_bound, 2 hll, 3 hlld. */
void o_riemann(const real* Lq, const real* Rq, int dir, int kind, real gamma, real ch,
               real smallr, real pfl, real llf_dmin, real llf_pmin, real* out)
{
    real L[NV], Rr[NV], f[NV];
    o_rot_to(Lq, L, dir);
    o_rot_to(Rq, Rr, dir);
    if (L[0] < smallr) O_BR(21, 0);
    if (L[4] < pfl) O_BR(21, 1);
    if (Rr[0] < smallr) O_BR(21, 2);
    if (Rr[4] < pfl) O_BR(21, 3);
    L[0] = fmax(L[0], smallr); L[4] = fmax(L[4], pfl);
    Rr[0] = fmax(Rr[0], smallr); Rr[4] = fmax(Rr[4], pfl);
    real bns, psis;
    o_glm_pair(L[5], Rr[5], L[8], Rr[8], ch, &bns, &psis);
    real fbn = psis, fpsi = ch * ch * bns;
    L[5] = bns; Rr[5] = bns;
    if (kind == 1) {
        llf_bound(L, Rr, gamma, ch, fbn, fpsi, f);
    } else if (kind == 2) {
        hll_x(L, Rr, gamma, bns, fbn, fpsi, f);
    } else {
        int uself = (llf_dmin > R(0) && fmin(L[0], Rr[0]) < llf_dmin) ||
                    (llf_pmin > R(0) && fmin(L[4], Rr[4]) < llf_pmin) || kind == 0;
        if (uself && kind == 3) O_HIT(4);
        if (kind == 3) O_BR(22, uself);
        if (uself) llf_x(L, Rr, gamma, ch, fbn, fpsi, f);
        else o_hlld_x(L, Rr, gamma, bns, fbn, fpsi, f);
    }
    o_rot_flux_from(f, out, dir);
}

/* glmmhd_turb.jl:483-505 turb_interp, 0-based cell index (the Julia 1-based i-0.5 is i0+0.5).
 * afield: 3 SoA planes of gs^3, plane c at a + c*gs^3, point (i,j,k) at i + gs*(j + gs*k). */
void o_turb_interp(const real* afield, int gs, int i, int j, int k, real dx,
                   real lx, real ly, real lz, real ramp, real* acc)
{
    real g = (real)gs;
    real cx = ((R(i) + R(0.5)) * dx) / lx * g;
    real cy = ((R(j) + R(0.5)) * dx) / ly * g;
    real cz = ((R(k) + R(0.5)) * dx) / lz * g;
    int i0 = (int)floor(cx), j0 = (int)floor(cy), k0 = (int)floor(cz);
    real wx = cx - R(i0), wy = cy - R(j0), wz = cz - R(k0);
    int i0w = ((i0 % gs) + gs) % gs, i1w = (((i0 + 1) % gs) + gs) % gs;
    int j0w = ((j0 % gs) + gs) % gs, j1w = (((j0 + 1) % gs) + gs) % gs;
    int k0w = ((k0 % gs) + gs) % gs, k1w = (((k0 + 1) % gs) + gs) % gs;
    size_t g3 = (size_t)gs * gs * gs;
#define AF(c, a, b, d) afield[(size_t)(c) * g3 + (size_t)(a) + (size_t)gs * ((size_t)(b) + (size_t)gs * (size_t)(d))]
    for (int c = 0; c < 3; ++c) {
        real a000 = AF(c, i0w, j0w, k0w), a100 = AF(c, i1w, j0w, k0w);
        real a010 = AF(c, i0w, j1w, k0w), a110 = AF(c, i1w, j1w, k0w);
        real a001 = AF(c, i0w, j0w, k1w), a101 = AF(c, i1w, j0w, k1w);
        real a011 = AF(c, i0w, j1w, k1w), a111 = AF(c, i1w, j1w, k1w);
        real t = (((a000 * (R(1) - wx) + a100 * wx) * (R(1) - wy) + (a010 * (R(1) - wx) + a110 * wx) * wy) * (R(1) - wz) +
                  ((a001 * (R(1) - wx) + a101 * wx) * (R(1) - wy) + (a011 * (R(1) - wx) + a111 * wx) * wy) * wz);
        acc[c] = ramp * t;
    }
#undef AF
}

/* ------------------------------------------------------------------------- */
/* Whole step (multi-pass)                                                     */
/* ------------------------------------------------------------------------- */
static inline size_t cidx(int i, int j, int k, int nx, int ny)
{
    return (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * (size_t)k);
}
static inline int wrapi(int i, int n) { int m = i % n; return m < 0 ? m + n : m; }

static void load9(const real* a, size_t n, size_t c, real* q)
{
    for (int v = 0; v < NV; ++v) q[v] = a[(size_t)v * n + c];
}
static void store9(real* a, size_t n, size_t c, const real* q)
{
    for (int v = 0; v < NV; ++v) a[(size_t)v * n + c] = q[v];
}

/* Reconstruction of one cell: the face states on its -d and +d faces, d = x,y,z.
 * Full predictor: glmmhd_turb.jl:378-403 (recon_kernel!), :955-973 (PPM with the spj fallback),
 * :618 (iflo).  Normal-only predictor: spike_mhd.cu:262-275 (edge/hanc1d, PPM variant). */
static void recon_cell(const real* prim, size_t n, int i, int j, int k, int nx, int ny, int nz,
                       const o_scheme* sc, real* fm /*3*NV*/, real* fp /*3*NV*/, int* rec /*10 or NULL*/)
{
    real gamma = R(sc->gamma), smallr = R(sc->smallr), pfl = R(sc->pfl);
    real dtdx = R(sc->dt) / R(sc->dx);
    real m0[NV], nb[6][NV], s[3][NV], mh[3][NV];
    load9(prim, n, cidx(i, j, k, nx, ny), m0);
    load9(prim, n, cidx(wrapi(i - 1, nx), j, k, nx, ny), nb[0]);
    load9(prim, n, cidx(wrapi(i + 1, nx), j, k, nx, ny), nb[1]);
    load9(prim, n, cidx(i, wrapi(j - 1, ny), k, nx, ny), nb[2]);
    load9(prim, n, cidx(i, wrapi(j + 1, ny), k, nx, ny), nb[3]);
    load9(prim, n, cidx(i, j, wrapi(k - 1, nz), nx, ny), nb[4]);
    load9(prim, n, cidx(i, j, wrapi(k + 1, nz), nx, ny), nb[5]);
    for (int d = 0; d < 3; ++d) prim_slope(nb[2 * d], m0, nb[2 * d + 1], s[d], sc->slope);
    if (sc->predictor == 0) {
        real uh[NV], q[NV];
        o_hancock(m0, s[0], s[1], s[2], dtdx, gamma, uh);
        o_cons2prim(uh, q, gamma, smallr, pfl);
        if (sc->mh_revert) O_BR(23, q[0] <= smallr || q[4] <= pfl);
        const int rv = sc->mh_revert && (q[0] <= smallr || q[4] <= pfl);
        if (rv) { O_HIT(1); memcpy(q, m0, sizeof q); }
        for (int d = 0; d < 3; ++d) memcpy(mh[d], q, sizeof q);
        if (rec) rec[0] = rec[1] = rec[2] = rv;
    } else {
        for (int d = 0; d < 3; ++d) {
            o_hanc1d(m0, s[d], d, dtdx, gamma, smallr, pfl, mh[d]);
            if (sc->mh_revert) O_BR(23, mh[d][0] <= smallr || mh[d][4] <= pfl);
            const int rv = sc->mh_revert && (mh[d][0] <= smallr || mh[d][4] <= pfl);
            if (rv) { O_HIT(1); memcpy(mh[d], m0, sizeof m0); }
            if (rec) rec[d] = rv;
        }
    }
    int up = 1;
    if (sc->recon == 1)
        up = !(o_spj(nb[0][4], m0[4], nb[1][4]) || o_spj(nb[2][4], m0[4], nb[3][4]) || o_spj(nb[4][4], m0[4], nb[5][4]));
    if (rec) rec[9] = up;
    for (int d = 0; d < 3; ++d) {
        real* lo = fm + d * NV;
        real* hi = fp + d * NV;
        if (sc->recon == 0 || (sc->recon == 1 && !up)) {
            for (int v = 0; v < NV; ++v) { lo[v] = mh[d][v] - s[d][v]; hi[v] = mh[d][v] + s[d][v]; }
        } else if (sc->recon == 1) {
            for (int v = 0; v < NV; ++v) {
                real el, eh;
                o_ppm_edges(nb[2 * d][v], m0[v], nb[2 * d + 1][v], &el, &eh);
                lo[v] = mh[d][v] + el - m0[v];
                hi[v] = mh[d][v] + eh - m0[v];
            }
        } else {
            /* spike_mhd.cu:174-192: edge floored (ppm_edge), then e + mh - c floored (padd_pred) */
            real el[NV], eh[NV];
            for (int v = 0; v < NV; ++v) {
                el[v] = o_ppm1_spike(nb[2 * d][v], m0[v], nb[2 * d + 1][v], R(-1));
                eh[v] = o_ppm1_spike(nb[2 * d][v], m0[v], nb[2 * d + 1][v], R(1));
            }
            if (el[0] < smallr || el[4] < pfl || eh[0] < smallr || eh[4] < pfl) O_BR(30, 0);
            el[0] = fmax(el[0], smallr); el[4] = fmax(el[4], pfl);
            eh[0] = fmax(eh[0], smallr); eh[4] = fmax(eh[4], pfl);
            for (int v = 0; v < NV; ++v) { lo[v] = el[v] + mh[d][v] - m0[v]; hi[v] = eh[v] + mh[d][v] - m0[v]; }
            if (lo[0] < smallr || lo[4] < pfl || hi[0] < smallr || hi[4] < pfl) O_BR(30, 1);
            lo[0] = fmax(lo[0], smallr); lo[4] = fmax(lo[4], pfl);
            hi[0] = fmax(hi[0], smallr); hi[4] = fmax(hi[4], pfl);
        }
        if (sc->face_guard == 1) {
            O_BR(24, lo[0] < smallr || lo[4] < pfl);
            O_BR(24, hi[0] < smallr || hi[4] < pfl);
            const int rl = lo[0] < smallr || lo[4] < pfl, rh = hi[0] < smallr || hi[4] < pfl;
            if (rl) { O_HIT(2); memcpy(lo, m0, sizeof m0); }
            if (rh) { O_HIT(2); memcpy(hi, m0, sizeof m0); }
            if (rec) { rec[3 + d] = rl; rec[6 + d] = rh; }
        } else if (rec) {
            rec[3 + d] = rec[6 + d] = 0;
        }
        if (sc->face_guard == 2) {
            if (lo[0] < smallr) O_BR(25, 0);
            if (lo[4] < pfl) O_BR(25, 1);
            if (hi[0] < smallr) O_BR(25, 2);
            if (hi[4] < pfl) O_BR(25, 3);
            lo[0] = fmax(lo[0], smallr); lo[4] = fmax(lo[4], pfl);
            hi[0] = fmax(hi[0], smallr); hi[4] = fmax(hi[4], pfl);
        }
    }
}

/* The face states of cell (i,j,k) from a primitive state (9 planes): fm/fp[d*9+v] on its -d/+d
 * face.  Exposes recon_cell to the tests. */
void o_recon_cell_faces(const real* prim, int nx, int ny, int nz, int i, int j, int k, const o_scheme* sc,
                        real* fm, real* fp)
{
    recon_cell(prim, (size_t)nx * ny * nz, i, j, k, nx, ny, nz, sc, fm, fp, NULL);
}

/* Fused update epilogue of one cell: glmmhd_turb.jl:440-476 (update_kernel!).
 * r: in = u0, out = new state.  acc_xy/acc_z are the flux differences (F_lo - F_hi). */
void o_update_cell(real* r, const real* fxl, const real* fxh, const real* fyl, const real* fyh,
                   const real* fzl, const real* fzh, const o_scheme* sc, const real* afield,
                   int i, int j, int k, int nx, int ny, int nz)
{
    real gamma = R(sc->gamma), smallr = R(sc->smallr), pfl = R(sc->pfl), smallc = R(sc->smallc);
    real dtdx = R(sc->dt) / R(sc->dx);
    for (int v = 0; v < NV; ++v) {
        real a = (fxl[v] - fxh[v]) + (fyl[v] - fyh[v]);
        a = a + (fzl[v] - fzh[v]);
        r[v] = r[v] + dtdx * a;
    }
    if (sc->update_floor) O_BR(26, r[0] < smallr);
    if (sc->update_floor && r[0] < smallr) {
        O_HIT(3);
        real emag = R(0.5) * (r[5] * r[5] + r[6] * r[6] + r[7] * r[7]);
        r[0] = smallr; r[1] = R(0); r[2] = R(0); r[3] = R(0);
        r[4] = pfl / (gamma - R(1)) + emag;
    }
    r[8] = r[8] * R(sc->glm_fac);
    if (sc->do_turb && afield) O_BR(27, r[0] >= R(sc->turb_min_rho));
    if (sc->do_turb && afield && r[0] >= R(sc->turb_min_rho)) {
        real dx = R(sc->dx), acc[3];
        o_turb_interp(afield, sc->turb_gs, i, j, k, dx, R(nx) * dx, R(ny) * dx, R(nz) * dx, R(sc->ramp), acc);
        real rho = r[0], rhom = fmax(rho, smallr), sc2 = smallc * smallc;
        real e = r[4], dt = R(sc->dt);
        for (int a = 1; a <= 3; ++a) {
            if (e - R(0.5) * r[a] * r[a] / rhom < rho * sc2) O_BR(28, a);
            e = fmax(e - R(0.5) * r[a] * r[a] / rhom, rho * sc2);
        }
        real m2 = r[1] + rhom * acc[0] * dt;
        real m3 = r[2] + rhom * acc[1] * dt;
        real m4 = r[3] + rhom * acc[2] * dt;
        const real mm[3] = {m2, m3, m4};
        for (int a = 0; a < 3; ++a) {
            if (e + R(0.5) * mm[a] * mm[a] / rhom < rho * sc2) O_BR(29, a + 1);
            e = fmax(e + R(0.5) * mm[a] * mm[a] / rhom, rho * sc2);
        }
        r[1] = m2; r[2] = m3; r[3] = m4; r[4] = e;
    }
}

/* One step unew = L(uold).  Three passes as step_multi! (glmmhd_turb.jl:1981-1992):
 * recon (per cell face states), flux (one Riemann solve per face, F_d at the +d face of each
 * cell), update.  scratch must hold 99*n reals (prim 9, face states 54, fluxes 27, plus 9). */
int o_step(const real* uold, real* unew, int nx, int ny, int nz, const o_scheme* sc,
           const real* afield, real* scratch)
{
    size_t n = (size_t)nx * ny * nz;
    real gamma = R(sc->gamma), smallr = R(sc->smallr), pfl = R(sc->pfl), ch = R(sc->ch);
    real* prim = scratch;                 /* 9n */
    real* fm = prim + (size_t)NV * n;     /* 27n: face states on the -d face, [d*9+v] planes */
    real* fp = fm + (size_t)3 * NV * n;   /* 27n: face states on the +d face */
    real* flx = fp + (size_t)3 * NV * n;  /* 27n: flux at the +d face of each cell */
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                size_t c = cidx(i, j, k, nx, ny);
                real u[NV], q[NV];
                load9(uold, n, c, u);
                o_cons2prim(u, q, gamma, smallr, pfl);
                store9(prim, n, c, q);
            }
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                size_t c = cidx(i, j, k, nx, ny);
                real a[3 * NV], b[3 * NV];
                recon_cell(prim, n, i, j, k, nx, ny, nz, sc, a, b, NULL);
                for (int w = 0; w < 3 * NV; ++w) { fm[(size_t)w * n + c] = a[w]; fp[(size_t)w * n + c] = b[w]; }
            }
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                size_t c = cidx(i, j, k, nx, ny);
                size_t nbr[3] = {cidx(wrapi(i + 1, nx), j, k, nx, ny), cidx(i, wrapi(j + 1, ny), k, nx, ny),
                                 cidx(i, j, wrapi(k + 1, nz), nx, ny)};
                for (int d = 0; d < 3; ++d) {
                    real L[NV], Rr[NV], f[NV];
                    for (int v = 0; v < NV; ++v) {
                        L[v] = fp[(size_t)(d * NV + v) * n + c];
                        Rr[v] = fm[(size_t)(d * NV + v) * n + nbr[d]];
                    }
                    o_riemann(L, Rr, d, sc->riemann, gamma, ch, smallr, pfl, R(sc->llf_dmin), R(sc->llf_pmin), f);
                    for (int v = 0; v < NV; ++v) flx[(size_t)(d * NV + v) * n + c] = f[v];
                }
            }
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                size_t c = cidx(i, j, k, nx, ny);
                size_t lo[3] = {cidx(wrapi(i - 1, nx), j, k, nx, ny), cidx(i, wrapi(j - 1, ny), k, nx, ny),
                                cidx(i, j, wrapi(k - 1, nz), nx, ny)};
                real f[6][NV], r[NV];
                for (int d = 0; d < 3; ++d)
                    for (int v = 0; v < NV; ++v) {
                        f[2 * d][v] = flx[(size_t)(d * NV + v) * n + lo[d]];
                        f[2 * d + 1][v] = flx[(size_t)(d * NV + v) * n + c];
                    }
                load9(uold, n, c, r);
                o_update_cell(r, f[0], f[1], f[2], f[3], f[4], f[5], sc, afield, i, j, k, nx, ny, nz);
                store9(unew, n, c, r);
            }
    return 0;
}

size_t o_step_scratch_reals(int nx, int ny, int nz) { return (size_t)99 * nx * ny * nz; }

/* ------------------------------------------------------------------------- */
/* Passive species riding the mass flux (march_bridge/cu/spike_mhd.cu -DNSP, :30-33, :55-61,   */
/* :128-131, :155-159, :177-188, :218-221, :278-288).                                          */
/* ------------------------------------------------------------------------- */
#define O_NSPMAX 2

/* One step of a state with nsp passive species: planes 0..8 the GLM-MHD state, planes 9..8+nsp the
 * conserved rho*X_s.  The MHD planes are o_step's result (the species do not feed back).  The
 * species, per cell:
 *   X = (rho X) / max(rho, smallr), with the cell's primitive density (spike_mhd.cu:243-244);
 *   half slopes with the scheme's limiter (:128-131, as every primitive);
 *   predictor: the advective half step X - dt/dx u_d hs_d (:155-159, hs the half slope) in each
 *     direction for the normal-only predictor, and X - dt/dx (u hs_x + v hs_y + w hs_z) for the
 *     full-transverse one (the same advection with all three transverse terms); X itself where
 *     the cell's MHD predictor reverts to m0;
 *   faces: the scheme's reconstruction of X around that half step (PLM, or PPM and its PLM
 *     fallback, or the spike PPM edges, :177-188), X itself on a face that the iflo guard reverts;
 *   flux: F_s = F_rho X_L if the face's mass flux F_rho >= 0, else F_rho X_R (CMA, :218-221);
 *   update: rho X += dt/dx sum_d (F_lo - F_hi) in o_update_cell's order (:278-288); where the MHD
 *     update floors the density (rho < smallr, glmmhd_turb.jl:448-451) rho X = smallr X_old.
 * scratch: o_step_sp_scratch_reals(nx, ny, nz, nsp) reals. */
size_t o_step_sp_scratch_reals(int nx, int ny, int nz, int nsp)
{
    return (size_t)(99 + 10 * nsp) * nx * ny * nz;
}

int o_step_sp(const real* uold, real* unew, int nx, int ny, int nz, int nsp, const o_scheme* sc,
              const real* afield, real* scratch)
{
    if (nsp < 0 || nsp > O_NSPMAX) return 1;
    int rc = o_step(uold, unew, nx, ny, nz, sc, afield, scratch);
    if (rc || nsp == 0) return rc;
    const size_t n = (size_t)nx * ny * nz;
    const real smallr = R(sc->smallr), dtdx = R(sc->dt) / R(sc->dx);
    const real* prim = scratch;                      /* o_step's primitives (9n) */
    const real* flx = scratch + (size_t)63 * n;      /* o_step's fluxes at the +d face (27n, after prim and fm/fp) */
    real* X = scratch + (size_t)99 * n;              /* nsp*n fractions */
    real* FS = X + (size_t)nsp * n;                  /* [d][lo,hi][s] face states: 6*nsp*n */
    real* FF = FS + (size_t)6 * nsp * n;             /* [d][s] species flux at the +d face: 3*nsp*n */
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                size_t c = cidx(i, j, k, nx, ny);
                for (int sp = 0; sp < nsp; ++sp) X[(size_t)sp * n + c] = uold[(size_t)(9 + sp) * n + c] / prim[c];
            }
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                size_t c = cidx(i, j, k, nx, ny);
                real fm[3 * NV], fp[3 * NV];
                int rec[10];
                recon_cell(prim, n, i, j, k, nx, ny, nz, sc, fm, fp, rec);
                size_t nb[6] = {cidx(wrapi(i - 1, nx), j, k, nx, ny), cidx(wrapi(i + 1, nx), j, k, nx, ny),
                                cidx(i, wrapi(j - 1, ny), k, nx, ny), cidx(i, wrapi(j + 1, ny), k, nx, ny),
                                cidx(i, j, wrapi(k - 1, nz), nx, ny), cidx(i, j, wrapi(k + 1, nz), nx, ny)};
                const real u3[3] = {prim[n + c], prim[2 * n + c], prim[3 * n + c]};
                for (int sp = 0; sp < nsp; ++sp) {
                    const real* Xs = X + (size_t)sp * n;
                    const real x0 = Xs[c];
                    real hs[3], xh[3];
                    for (int d = 0; d < 3; ++d) {
                        real dl = x0 - Xs[nb[2 * d]], dr = Xs[nb[2 * d + 1]] - x0;
                        hs[d] = R(0.5) * (sc->slope == 1 ? o_minmod(dl, dr) : o_moncen(dl, dr));
                    }
                    if (sc->predictor == 0) {
                        real t = x0 - dtdx * ((u3[0] * hs[0] + u3[1] * hs[1]) + u3[2] * hs[2]);
                        for (int d = 0; d < 3; ++d) xh[d] = rec[d] ? x0 : t;
                    } else {
                        for (int d = 0; d < 3; ++d) xh[d] = rec[d] ? x0 : x0 - (dtdx * u3[d]) * hs[d];
                    }
                    for (int d = 0; d < 3; ++d) {
                        const real xl = Xs[nb[2 * d]], xr = Xs[nb[2 * d + 1]];
                        real lo, hi;
                        if (sc->recon == 0 || (sc->recon == 1 && !rec[9])) {
                            lo = xh[d] - hs[d]; hi = xh[d] + hs[d];
                        } else if (sc->recon == 1) {
                            real el, eh;
                            o_ppm_edges(xl, x0, xr, &el, &eh);
                            lo = xh[d] + el - x0; hi = xh[d] + eh - x0;
                        } else {
                            real el = o_ppm1_spike(xl, x0, xr, R(-1)), eh = o_ppm1_spike(xl, x0, xr, R(1));
                            lo = el + xh[d] - x0; hi = eh + xh[d] - x0;
                        }
                        if (rec[3 + d]) lo = x0;
                        if (rec[6 + d]) hi = x0;
                        FS[((size_t)(d * 2 + 0) * nsp + sp) * n + c] = lo;
                        FS[((size_t)(d * 2 + 1) * nsp + sp) * n + c] = hi;
                    }
                }
            }
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                size_t c = cidx(i, j, k, nx, ny);
                size_t up[3] = {cidx(wrapi(i + 1, nx), j, k, nx, ny), cidx(i, wrapi(j + 1, ny), k, nx, ny),
                                cidx(i, j, wrapi(k + 1, nz), nx, ny)};
                for (int d = 0; d < 3; ++d) {
                    const real fr = flx[(size_t)(d * NV) * n + c];
                    for (int sp = 0; sp < nsp; ++sp) {
                        const real xL = FS[((size_t)(d * 2 + 1) * nsp + sp) * n + c];
                        const real xR = FS[((size_t)(d * 2 + 0) * nsp + sp) * n + up[d]];
                        FF[((size_t)d * nsp + sp) * n + c] = fr >= R(0) ? fr * xL : fr * xR;
                    }
                }
            }
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                size_t c = cidx(i, j, k, nx, ny);
                size_t lo[3] = {cidx(wrapi(i - 1, nx), j, k, nx, ny), cidx(i, wrapi(j - 1, ny), k, nx, ny),
                                cidx(i, j, wrapi(k - 1, nz), nx, ny)};
                /* the MHD update's density before its floor (o_update_cell's arithmetic) */
                real a0 = (flx[lo[0]] - flx[c]) + (flx[(size_t)NV * n + lo[1]] - flx[(size_t)NV * n + c]);
                a0 = a0 + (flx[(size_t)2 * NV * n + lo[2]] - flx[(size_t)2 * NV * n + c]);
                const int floored = sc->update_floor && uold[c] + dtdx * a0 < smallr;
                for (int sp = 0; sp < nsp; ++sp) {
                    const real* F = FF + (size_t)sp * n;
                    const size_t st = (size_t)nsp * n;
                    real a = (F[lo[0]] - F[c]) + (F[st + lo[1]] - F[st + c]);
                    a = a + (F[2 * st + lo[2]] - F[2 * st + c]);
                    real r = uold[(size_t)(9 + sp) * n + c] + dtdx * a;
                    if (floored) r = smallr * X[(size_t)sp * n + c];
                    unew[(size_t)(9 + sp) * n + c] = r;
                }
            }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* CFL, diagnostics, initial conditions                                        */
/* ------------------------------------------------------------------------- */

/* glmmhd_turb.jl:1963-1971 ctot_kernel! + :1976 maximum */
double o_ctot_max(const real* u, int nx, int ny, int nz, double gamma_, double smallr_, double pfl_)
{
    size_t n = (size_t)nx * ny * nz;
    real gamma = R(gamma_), smallr = R(smallr_), pfl = R(pfl_);
    double m = 0.0;
#pragma omp parallel for reduction(max : m) schedule(static)
    for (long long c = 0; c < (long long)n; ++c) {
        real cc[NV], q[NV];
        load9(u, n, (size_t)c, cc);
        o_cons2prim(cc, q, gamma, smallr, pfl);
        real ct = (fabs(q[1]) + o_fast_speed(q, gamma, q[5])) + (fabs(q[2]) + o_fast_speed(q, gamma, q[6])) +
                  (fabs(q[3]) + o_fast_speed(q, gamma, q[7]));
        if ((double)ct > m) m = (double)ct;
    }
    return m;
}

/* Largest fast magnetosonic speed of the grid, max over cells and axes of c_f(B_d) (the
 * "ch = max fast speed" Dedner cleaning speed; glmmhd_turb.jl:89-95 per axis). */
double o_fast_max(const real* u, int nx, int ny, int nz, double gamma_, double smallr_, double pfl_)
{
    size_t n = (size_t)nx * ny * nz;
    real gamma = R(gamma_), smallr = R(smallr_), pfl = R(pfl_);
    double m = 0.0;
#pragma omp parallel for reduction(max : m) schedule(static)
    for (long long c = 0; c < (long long)n; ++c) {
        real cc[NV], q[NV];
        load9(u, n, (size_t)c, cc);
        o_cons2prim(cc, q, gamma, smallr, pfl);
        real f = fmax(o_fast_speed(q, gamma, q[5]), fmax(o_fast_speed(q, gamma, q[6]), o_fast_speed(q, gamma, q[7])));
        if ((double)f > m) m = (double)f;
    }
    return m;
}

/* Diagnostics: glmmhd_turb.jl:1996-2018 (central max|div B|, mass, sum rho v^2, rhomin,
 * finite) plus the energy budget (E_kin, E_mag, cross helicity sum v.B, total E, sum psi^2),
 * all accumulated in double with Neumaier-compensated sums in fixed cell order.
 * out[0]=mass out[1]=sum m^2/max(rho,smallr) out[2]=divbmax out[3]=rhomin out[4]=finite
 * out[5]=Ekin (0.5 sum m^2/rho) out[6]=Emag (0.5 sum B^2) out[7]=cross helicity sum m.B/rho
 * out[8]=sum E out[9]=sum psi^2.  Sums are plain cell sums (multiply by dV for integrals). */
#define O_NDIAG 10
typedef struct { double s, c; } nsum;
static inline void nadd(nsum* a, double x)
{
    double t = a->s + x;
    if (fabs(a->s) >= fabs(x)) a->c += (a->s - t) + x;
    else a->c += (x - t) + a->s;
    a->s = t;
}
void o_diagnostics(const real* u, int nx, int ny, int nz, double dx, double smallr_, double* out)
{
    size_t n = (size_t)nx * ny * nz;
    nsum mass = {0, 0}, ke2 = {0, 0}, ek = {0, 0}, em = {0, 0}, hc = {0, 0}, et = {0, 0}, ps = {0, 0};
    double divb = 0.0, rmin = INFINITY;
    int finite = 1;
    const real *rho = u, *mx = u + n, *my = u + 2 * n, *mz = u + 3 * n, *E = u + 4 * n;
    const real *bx = u + 5 * n, *by = u + 6 * n, *bz = u + 7 * n, *psi = u + 8 * n;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                size_t c = cidx(i, j, k, nx, ny);
                double r = (double)rho[c], a = (double)mx[c], b = (double)my[c], d = (double)mz[c];
                double m2 = a * a + b * b + d * d;
                nadd(&mass, r);
                nadd(&ke2, m2 / fmax(r, smallr_));
                nadd(&ek, 0.5 * m2 / r);
                double Bx = (double)bx[c], By = (double)by[c], Bz = (double)bz[c];
                nadd(&em, 0.5 * (Bx * Bx + By * By + Bz * Bz));
                nadd(&hc, (a * Bx + b * By + d * Bz) / r);
                nadd(&et, (double)E[c]);
                nadd(&ps, (double)psi[c] * (double)psi[c]);
                if (r < rmin) rmin = r;
                /* glmmhd_turb.jl:1996-2004 divb_kernel! */
                double db = fabs(((double)bx[cidx(wrapi(i + 1, nx), j, k, nx, ny)] - (double)bx[cidx(wrapi(i - 1, nx), j, k, nx, ny)]) +
                                 ((double)by[cidx(i, wrapi(j + 1, ny), k, nx, ny)] - (double)by[cidx(i, wrapi(j - 1, ny), k, nx, ny)]) +
                                 ((double)bz[cidx(i, j, wrapi(k + 1, nz), nx, ny)] - (double)bz[cidx(i, j, wrapi(k - 1, nz), nx, ny)])) /
                            (2.0 * dx);
                if (db > divb) divb = db;
                for (int v = 0; v < NV; ++v)
                    if (!isfinite((double)u[(size_t)v * n + c])) finite = 0;
            }
    out[0] = mass.s + mass.c; out[1] = ke2.s + ke2.c; out[2] = divb; out[3] = rmin; out[4] = finite;
    out[5] = ek.s + ek.c; out[6] = em.s + em.c; out[7] = hc.s + hc.c; out[8] = et.s + et.c; out[9] = ps.s + ps.c;
}

/* glmmhd_turb.jl:2021-2036 ic_ot_kernel! (z-invariant); x = (i+0.5)/nx, y = (j+0.5)/ny */
void o_ic_ot(real* u, int nx, int ny, int nz, double gamma_)
{
    size_t n = (size_t)nx * ny * nz;
    real gamma = R(gamma_);
    real pi = R(3.14159265358979323846);
    real twopi = R(2) * pi;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                real x = (R(i) + R(0.5)) / R(nx), y = (R(j) + R(0.5)) / R(ny);
                real rho = R(25) / (R(36) * pi);
                real pr = R(5) / (R(12) * pi);
                real vx = -sin(twopi * y), vy = sin(twopi * x), vz = R(0);
                real sb = R(1) / sqrt(R(4) * pi);
                real bx = -sb * sin(twopi * y), by = sb * sin(R(2) * twopi * x), bz = R(0);
                real ekin = R(0.5) * rho * (vx * vx + vy * vy + vz * vz);
                real emag = R(0.5) * (bx * bx + by * by + bz * bz);
                size_t c = cidx(i, j, k, nx, ny);
                u[c] = rho; u[n + c] = rho * vx; u[2 * n + c] = rho * vy; u[3 * n + c] = rho * vz;
                u[4 * n + c] = pr / (gamma - R(1)) + ekin + emag;
                u[5 * n + c] = bx; u[6 * n + c] = by; u[7 * n + c] = bz; u[8 * n + c] = R(0);
            }
}

/* glmmhd_turb.jl:2164-2172 ic_uniform_kernel!, generalised to B0 = (b0x,b0y,b0z) */
void o_ic_uniform(real* u, int nx, int ny, int nz, double rho0, double p0, double b0x, double b0y,
                  double b0z, double gamma)
{
    size_t n = (size_t)nx * ny * nz;
    real emag = R(0.5) * (R(b0x) * R(b0x) + R(b0y) * R(b0y) + R(b0z) * R(b0z));
    real e = R(p0) / (R(gamma) - R(1)) + emag;
    for (size_t c = 0; c < n; ++c) {
        u[c] = R(rho0); u[n + c] = R(0); u[2 * n + c] = R(0); u[3 * n + c] = R(0);
        u[4 * n + c] = e; u[5 * n + c] = R(b0x); u[6 * n + c] = R(b0y); u[7 * n + c] = R(b0z);
        u[8 * n + c] = R(0);
    }
}

/* spike_mhd.cu:304-317 init: the gradient throughput IC, ph = 0.001*(i mod 911) over the
 * linear cell index. */
void o_ic_gradient(real* u, int nx, int ny, int nz, double gamma)
{
    size_t n = (size_t)nx * ny * nz;
    for (size_t c = 0; c < n; ++c) {
        real ph = R(0.001) * (real)(c % 911);
        real q[NV] = {R(1) + R(0.3) * ph, R(0.1) * ph, R(-0.1) * ph, R(0.05) * ph, R(2.5) + R(0.5) * ph,
                      R(0.2) + R(0.1) * ph, R(0.1) * ph, R(-0.05) * ph, R(0)};
        real cc[NV];
        o_prim2cons(q, cc, R(gamma));
        for (int v = 0; v < NV; ++v) u[(size_t)v * n + c] = cc[v];
    }
}
