/*
 * hydro_oracle.c -- CPU restatement of the 5-variable hydro line march
 * (TEST INFRASTRUCTURE ONLY; used by tests/ and smoke()).
 *
 * Follows march_bridge/cu/spike_25d.cu as build_march.sh builds it (-DHANCOCK1D, fp32
 * conserved SoA state rho, mx, my, mz, E): per cell and direction, MonCen slopes of the
 * primitive state, the normal-only Hancock half step, face states at mh -/+ s/2 clamped at
 * SMALLR/SMALLP, and HLL (or, with `fallback`, the unified HLL/LLF solver of -DLLF_FALLBACK);
 * U1 = toCons(prim(U0)) - dtdx * sum_d (F(+1/2) - F(-1/2)).  Periodic in all three axes.
 * Each face is solved once per adjacent cell, as in the reference (the two solves see the
 * same inputs).  Cell (i,j,k) at i + nx*(j + ny*k); plane v at v*nx*ny*nz.
 *
 * Passive scalars (-DSCALARS, nsp = 2): planes 5.. hold rho*s, reconstructed as s = (rho*s)/rho
 * and advected with the flow; their flux is the HLL (or LLF) flux of rho*s, or with `cma`
 * (-DCMA, consistent multi-fluid advection) the mass flux times the upwind face value of s.
 *
 * Compiled with -DREAL=double and -DREAL=float, -ffp-contract=off.
 */
#include <math.h>
#include <stddef.h>
#include <tgmath.h>

#ifndef REAL
#define REAL double
#endif
typedef REAL real;
#define R(x) ((real)(x))
#define HV 7   /* largest state: 5 hydro variables + 2 passive scalars */

typedef struct {
    double gamma, dtdx, smallr, smallp; /* spike_25d.cu:180 constants */
    int fallback;                       /* 0: HLL (:302-330); 1: riemann_fb (:335-376) */
    double fb_dmin, fb_pmin;            /* the fallback's floors (1e-3 in :339) */
    int nsp;                            /* passive scalars: 0 or 2 (-DSCALARS, :115-119) */
    int cma;                            /* scalar flux rides the mass flux (-DCMA, :321, :367) */
    int logsp;                          /* the scalar planes hold log2 X, reconstructed in log space
                                           (-DU16SP, :64-72, :188-191, :248-250, :317-319) */
    int ppm;                            /* parabolic edges (-DPPM1D, :467-483), HLL without fallback */
    int donor;                          /* donor-cell tile seams (-DDONOR, :438-454), HLL without fallback */
} HyScheme;

real o_ppm1_spike(real qm, real q0, real qp, real sgn);   /* glm_oracle.c: spike_25d.cu:264-273 ppm1 */

static int nvar(const HyScheme* s) { return 5 + s->nsp; }

/* spike_25d.cu:196-202 prim_from_cons (:183-194 with scalars) */
void oh_c2p(const real c[HV], real q[HV], const HyScheme* s)
{
    real g = R(s->gamma), sr = R(s->smallr), sp = R(s->smallp);
    real ir = R(1) / fmax(c[0], sr);
    q[0] = fmax(c[0], sr);
    q[1] = c[1] * ir;
    q[2] = c[2] * ir;
    q[3] = c[3] * ir;
    real ke = R(0.5) * q[0] * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    q[4] = fmax((g - R(1)) * (c[4] - ke), sp);
    for (int v = 5; v < nvar(s); ++v) q[v] = s->logsp ? c[v] : c[v] * ir;   /* :191; log2 X as is (:188-190) */
}

/* spike_25d.cu:207-213 moncen: the full monotonized-central slope */
real oh_moncen(real a, real b, real c)
{
    real dl = b - a, dr = c - b, dc = R(0.5) * (c - a);
    if (dl * dr <= R(0)) return R(0);
    real sg = dl > R(0) ? R(1) : R(-1);
    return sg * fmin(fabs(dc), fmin(R(2) * fabs(dl), R(2) * fabs(dr)));
}

/* spike_25d.cu:214-221 slopeP */
static void slope5(const real a[HV], const real b[HV], const real c[HV], real sl[HV], const HyScheme* s)
{
    for (int v = 0; v < nvar(s); ++v) sl[v] = oh_moncen(a[v], b[v], c[v]);
}

/* spike_25d.cu:222-230 edge: b + sign*s/2, clamped at SMALLR/SMALLP */
void oh_edge(const real b[HV], const real sl[HV], real sign, real e[HV], const HyScheme* s)
{
    for (int v = 0; v < nvar(s); ++v) e[v] = b[v] + sign * R(0.5) * sl[v];
    e[0] = fmax(e[0], R(s->smallr));
    e[4] = fmax(e[4], R(s->smallp));
}

/* spike_25d.cu:231-244 primFlux */
void oh_flux(const real q[HV], int dir, real f[HV], const HyScheme* s)
{
    real g = R(s->gamma);
    real un = q[1 + dir];
    real E = q[4] / (g - R(1)) + R(0.5) * q[0] * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    f[0] = q[0] * un;
    for (int c = 0; c < 3; ++c) f[1 + c] = q[0] * q[1 + c] * un + (dir == c ? q[4] : R(0));
    f[4] = (E + q[4]) * un;
    /* :241; with log2 X the fraction is decoded (the reference's supersonic return of this flux
     * multiplies by log2 X itself; here X is used on every face) */
    for (int v = 5; v < nvar(s); ++v) f[v] = q[0] * un * (s->logsp ? exp2(q[v]) : q[v]);
}

/* spike_25d.cu:245-255 toCons */
void oh_tocons(const real q[HV], real c[HV], const HyScheme* s)
{
    real g = R(s->gamma);
    c[0] = q[0];
    c[1] = q[0] * q[1];
    c[2] = q[0] * q[2];
    c[3] = q[0] * q[3];
    c[4] = q[4] / (g - R(1)) + R(0.5) * q[0] * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    for (int v = 5; v < nvar(s); ++v) c[v] = q[0] * (s->logsp ? exp2(q[v]) : q[v]);   /* :251, :248-250 */
}

/* spike_25d.cu:287-301 hanc1d: normal-only Hancock half step in direction dir */
void oh_hanc1d(const real m[HV], const real sl[HV], int dir, real mh[HV], const HyScheme* s)
{
    real g = R(s->gamma);
    real un = m[1 + dir], sn = sl[1 + dir];
    real ir = R(1) / m[0];
    real h = R(0.5) * R(s->dtdx);
    mh[0] = m[0] - h * (un * sl[0] + m[0] * sn);
    for (int c = 0; c < 3; ++c) mh[1 + c] = m[1 + c] - h * (un * sl[1 + c] + (dir == c ? sl[4] * ir : R(0)));
    mh[4] = m[4] - h * (un * sl[4] + g * m[4] * sn);
    for (int v = 5; v < nvar(s); ++v) mh[v] = m[v] - h * un * sl[v];   /* :298 */
}

/* spike_25d.cu:302-330 hll (the scalar part of the 5-variable build) */
void oh_hll(const real L[HV], const real Rt[HV], int dir, real F[HV], const HyScheme* s)
{
    real g = R(s->gamma);
    real unL = L[1 + dir], unR = Rt[1 + dir];
    real aL = sqrt(g * L[4] / L[0]), aR = sqrt(g * Rt[4] / Rt[0]);
    real sL = fmin(unL - aL, unR - aR), sR = fmax(unL + aL, unR + aR);
    real FL[HV], FR[HV], UL[HV], UR[HV];
    oh_flux(L, dir, FL, s);
    oh_flux(Rt, dir, FR, s);
    oh_tocons(L, UL, s);
    oh_tocons(Rt, UR, s);
    if (sL >= R(0)) { for (int v = 0; v < nvar(s); ++v) F[v] = FL[v]; return; }
    if (sR <= R(0)) { for (int v = 0; v < nvar(s); ++v) F[v] = FR[v]; return; }
    real inv = R(1) / (sR - sL);
    for (int v = 0; v < nvar(s); ++v) F[v] = (sR * FL[v] - sL * FR[v] + sL * sR * (UR[v] - UL[v])) * inv;
    if (s->cma)   /* :321 */
        for (int v = 5; v < nvar(s); ++v) {
            real xl = s->logsp ? exp2(L[v]) : L[v], xr = s->logsp ? exp2(Rt[v]) : Rt[v];   /* :317-319 */
            F[v] = F[0] >= R(0) ? F[0] * xl : F[0] * xr;
        }
}

/* spike_25d.cu:335-376 riemann_fb: LLF where a face state is below the fallback floors,
 * else HLL */
void oh_riemann_fb(const real L[HV], const real Rt[HV], int dir, real F[HV], const HyScheme* s)
{
    real g = R(s->gamma);
    real dmin = R(s->fb_dmin), pmin = R(s->fb_pmin);
    if (L[0] < dmin || L[4] < pmin || Rt[0] < dmin || Rt[4] < pmin) {
        real unL = L[1 + dir], unR = Rt[1 + dir];
        real aL = sqrt(g * L[4] / L[0]), aR = sqrt(g * Rt[4] / Rt[0]);
        real FL[HV], FR[HV], UL[HV], UR[HV];
        oh_flux(L, dir, FL, s);
        oh_flux(Rt, dir, FR, s);
        oh_tocons(L, UL, s);
        oh_tocons(Rt, UR, s);
        real sm = fmax(fabs(unL) + aL, fabs(unR) + aR);
        for (int v = 0; v < nvar(s); ++v) F[v] = R(0.5) * (FL[v] + FR[v]) - R(0.5) * sm * (UR[v] - UL[v]);
        return;
    }
    oh_hll(L, Rt, dir, F, s);
}

static void solve(const real L[HV], const real Rt[HV], int dir, real F[HV], const HyScheme* s)
{
    if (s->fallback) oh_riemann_fb(L, Rt, dir, F, s);
    else oh_hll(L, Rt, dir, F, s);
}

static size_t wrapi(long i, long n) { long m = i % n; return (size_t)(m < 0 ? m + n : m); }

/* Primitive state of every cell (spike_25d.cu:427 GP), `prim` holds 5 planes. */
static void prim_all(const real* u, real* prim, int nx, int ny, int nz, const HyScheme* s)
{
    size_t n = (size_t)nx * ny * nz;
#pragma omp parallel for schedule(static)
    for (long c = 0; c < (long)n; ++c) {
        real cc[HV], q[HV];
        for (int v = 0; v < nvar(s); ++v) cc[v] = u[(size_t)v * n + c];
        oh_c2p(cc, q, s);
        for (int v = 0; v < nvar(s); ++v) prim[(size_t)v * n + c] = q[v];
    }
}

/* Scratch reals oh_step needs: the primitive planes. */
size_t oh_step_scratch_reals(int nx, int ny, int nz) { return (size_t)HV * nx * ny * nz; }   /* enough for nsp = 2 */

/* One step of the march, spike_25d.cu:430-516 (HANCOCK1D branch :462-466, update :492,
 * store :512).  Returns 0. */
static int step_from_prim(const real* prim, real* out, int nx, int ny, int nz, const HyScheme* s);

int oh_step(const real* u, real* out, int nx, int ny, int nz, const HyScheme* s, real* scratch)
{
    prim_all(u, scratch, nx, ny, nz, s);
    return step_from_prim(scratch, out, nx, ny, nz, s);
}

/* The same step from primitive planes [rho, vx, vy, vz, p (, s...)] held as such (the compact
 * primitive storage of -DF16STATE / -DU16STATE, :74-95: the tile holds make_prim(...), rho and p
 * floored at SMALLR/SMALLP, :204-206, with no conversion from conserved).  out: conserved. */
int oh_step_prim(const real* prim_in, real* out, int nx, int ny, int nz, const HyScheme* s, real* scratch)
{
    size_t n = (size_t)nx * ny * nz;
    for (int v = 0; v < nvar(s); ++v)
        for (size_t c = 0; c < n; ++c) {
            real x = prim_in[(size_t)v * n + c];
            if (v == 0) x = fmax(x, R(s->smallr));
            if (v == 4) x = fmax(x, R(s->smallp));
            scratch[(size_t)v * n + c] = x;
        }
    return step_from_prim(scratch, out, nx, ny, nz, s);
}

static int step_from_prim(const real* prim, real* out, int nx, int ny, int nz, const HyScheme* s)
{
    size_t n = (size_t)nx * ny * nz;
    real dtdx = R(s->dtdx);
#pragma omp parallel for schedule(static) collapse(2)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                size_t c = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k);
                real c0[HV];
                for (int v = 0; v < nvar(s); ++v) c0[v] = prim[(size_t)v * n + c];
                real div[HV] = {0, 0, 0, 0, 0, 0, 0};
                for (int d = 0; d < 3; ++d) {
                    int ox = d == 0, oy = d == 1, oz = d == 2;
                    real P[5][HV];   /* cells at offsets -2..+2 along d */
                    for (int o = -2; o <= 2; ++o) {
                        size_t cc = wrapi(i + o * ox, nx) +
                                    (size_t)nx * (wrapi(j + o * oy, ny) + (size_t)ny * wrapi(k + o * oz, nz));
                        for (int v = 0; v < nvar(s); ++v) P[o + 2][v] = prim[(size_t)v * n + cc];
                    }
                    real sm1[HV], s0[HV], sp1[HV];
                    slope5(P[0], P[1], P[2], sm1, s);
                    slope5(P[1], P[2], P[3], s0, s);
                    slope5(P[2], P[3], P[4], sp1, s);
                    real mhm[HV], mh0[HV], mhp[HV];
                    oh_hanc1d(P[1], sm1, d, mhm, s);
                    oh_hanc1d(P[2], s0, d, mh0, s);
                    oh_hanc1d(P[3], sp1, d, mhp, s);
                    real a[HV], b[HV], Fm[HV], Fp[HV];
                    if (s->ppm) {   /* :476-481: the parabola's edge plus the half step's change, unclamped */
                        for (int v = 0; v < nvar(s); ++v) {
                            a[v] = o_ppm1_spike(P[0][v], P[1][v], P[2][v], R(1)) + mhm[v] - P[1][v];
                            b[v] = o_ppm1_spike(P[1][v], P[2][v], P[3][v], R(-1)) + mh0[v] - P[2][v];
                        }
                        oh_hll(a, b, d, Fm, s);
                        for (int v = 0; v < nvar(s); ++v) {
                            a[v] = o_ppm1_spike(P[1][v], P[2][v], P[3][v], R(1)) + mh0[v] - P[2][v];
                            b[v] = o_ppm1_spike(P[2][v], P[3][v], P[4][v], R(-1)) + mhp[v] - P[3][v];
                        }
                        oh_hll(a, b, d, Fp, s);
                    } else if (s->donor) {   /* :440-448: first order at the faces on the 32 x 8 tile seams */
                        int lob = d == 0 ? i % 32 == 0 : d == 1 ? j % 8 == 0 : 0;
                        int hib = d == 0 ? ((i + 1) % nx) % 32 == 0 : d == 1 ? ((j + 1) % ny) % 8 == 0 : 0;
                        if (lob) oh_hll(P[1], P[2], d, Fm, s);
                        else {
                            oh_edge(mhm, sm1, R(1), a, s);
                            oh_edge(mh0, s0, R(-1), b, s);
                            oh_hll(a, b, d, Fm, s);
                        }
                        if (hib) oh_hll(P[2], P[3], d, Fp, s);
                        else {
                            oh_edge(mh0, s0, R(1), a, s);
                            oh_edge(mhp, sp1, R(-1), b, s);
                            oh_hll(a, b, d, Fp, s);
                        }
                    } else {
                        oh_edge(mhm, sm1, R(1), a, s);
                        oh_edge(mh0, s0, R(-1), b, s);
                        solve(a, b, d, Fm, s);
                        oh_edge(mh0, s0, R(1), a, s);
                        oh_edge(mhp, sp1, R(-1), b, s);
                        solve(a, b, d, Fp, s);
                    }
                    for (int v = 0; v < nvar(s); ++v) div[v] += Fp[v] - Fm[v];
                }
                real U[HV];
                oh_tocons(c0, U, s);
                for (int v = 0; v < nvar(s); ++v) out[(size_t)v * n + c] = U[v] - dtdx * div[v];
            }
    return 0;
}
