/*
 * ct_oracle.c -- CPU restatement of the constrained-transport (CT) MHD step
 * (TEST INFRASTRUCTURE ONLY; used by tests/, smoke() and bench.py's cpu_baseline leg).
 *
 * Follows ct_mhd.jl:20-194 (bcc, flux, update_hydro, emf, faceB kernels and step!) and the
 * fused march spike_ctm.cu:192-298 (same scheme).  Multi-pass, like ct_mhd.jl's 7 kernels.
 *
 * State: 8 planes [rho, mx, my, mz, E, bxf, byf, bzf]; bxf[i,j,k] is Bx on the lower-x face
 * (i-1/2, j, k) of cell (i,j,k) (ct_mhd.jl:10-11).  Cell (i,j,k) at i + nx*(j + ny*k).
 *
 * Reference discrepancy kept as a switch (face_slope):
 *   0: spike_ctm.cu form -- full MonCen slope s (spike_ctm.cu:73-77, :86-93), predictor at
 *      c -/+ s/2 and faces at mh +/- s/2 (spike_ctm.cu:170-171, :229);
 *   1: ct_mhd.jl literal form -- pslope already returns s/2 (ct_mhd.jl:37) and hanc1d/flux then
 *      apply +/-0.5 of it again (ct_mhd.jl:75, :111-112), i.e. faces at mh +/- s/4.
 * Riemann: 2 = HLL (ct_mhd.jl:66-72, spike_ctm.cu:154-168); 3 = HLLD (Miyoshi-Kusano, the GLM
 * solver glmmhd_turb.jl:193-272 with the face-normal B and F[Bn]=0); 0 = LLF (:182-190).
 * recon: 0 PLM; 2 PPM spike form (spike_ctm.cu:95-122, used at :225-227).
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
#define NC 8

typedef struct {
    double gamma, smallr, smallp;
    double dt, dx;
    double smallc, turb_min_rho, ramp;
    int face_slope;   /* 0 spike_ctm form, 1 ct_mhd.jl literal */
    int riemann;      /* 0 llf, 2 hll, 3 hlld */
    int recon;        /* 0 plm, 2 ppm (spike form) */
    int do_turb;
    int turb_gs;
    int update_floor;
} o_ct_scheme;

/* from glm_oracle.c */
void o_hlld_x(const real* L, const real* Rq, real gamma, real bn, real fbn, real fpsi, real* f);
real o_moncen(real dl, real dr);
real o_ppm1_spike(real qm, real q0, real qp, real sgn);
void o_turb_interp(const real* afield, int gs, int i, int j, int k, real dx, real lx, real ly, real lz,
                   real ramp, real* acc);

static inline size_t cidx(int i, int j, int k, int nx, int ny)
{
    return (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * (size_t)k);
}
static inline int wrapi(int i, int n) { int m = i % n; return m < 0 ? m + n : m; }

/* ct_mhd.jl:26-34 cellprim (rho, vx, vy, vz, p, Bx, By, Bz) from conserved + Bcc */
void o_ct_cellprim(const real* u, const real* bcc, real gamma, real smallr, real smallp, real* q)
{
    real rho = fmax(u[0], smallr), ir = R(1) / rho;
    real vx = u[1] * ir, vy = u[2] * ir, vz = u[3] * ir;
    real bx = bcc[0], by = bcc[1], bz = bcc[2];
    real p = fmax((gamma - R(1)) * (u[4] - R(0.5) * rho * (vx * vx + vy * vy + vz * vz) -
                                    R(0.5) * (bx * bx + by * by + bz * bz)), smallp);
    q[0] = rho; q[1] = vx; q[2] = vy; q[3] = vz; q[4] = p; q[5] = bx; q[6] = by; q[7] = bz;
}

/* ct_mhd.jl:39-55 dflux (F[Bn] = 0) */
void o_ct_dflux(const real* q, int d, real gamma, real* f)
{
    real rho = q[0], vx = q[1], vy = q[2], vz = q[3], p = q[4], bx = q[5], by = q[6], bz = q[7];
    real b2 = bx * bx + by * by + bz * bz, ptot = p + R(0.5) * b2;
    real E = p / (gamma - R(1)) + R(0.5) * rho * (vx * vx + vy * vy + vz * vz) + R(0.5) * b2;
    real vb = vx * bx + vy * by + vz * bz;
    real un = d == 0 ? vx : (d == 1 ? vy : vz);
    real bn = d == 0 ? bx : (d == 1 ? by : bz);
    f[0] = rho * un;
    f[1] = rho * un * vx - bn * bx + (d == 0 ? ptot : R(0));
    f[2] = rho * un * vy - bn * by + (d == 1 ? ptot : R(0));
    f[3] = rho * un * vz - bn * bz + (d == 2 ? ptot : R(0));
    f[4] = (E + ptot) * un - bn * vb;
    f[5] = d == 0 ? R(0) : un * bx - vx * bn;
    f[6] = d == 1 ? R(0) : un * by - vy * bn;
    f[7] = d == 2 ? R(0) : un * bz - vz * bn;
}

/* ct_mhd.jl:56-60 fast_speed in direction d */
static real ct_fast_speed(const real* q, int d, real gamma)
{
    real rho = q[0], bn = d == 0 ? q[5] : (d == 1 ? q[6] : q[7]);
    real c2 = gamma * q[4] / rho, b2 = (q[5] * q[5] + q[6] * q[6] + q[7] * q[7]) / rho;
    real dd = R(0.5) * (b2 + c2);
    return sqrt(dd + sqrt(fmax(dd * dd - c2 * bn * bn / rho, R(0))));
}

/* ct_mhd.jl:61-64 toC */
static void ct_toC(const real* q, real gamma, real* u)
{
    real rho = q[0], vx = q[1], vy = q[2], vz = q[3], p = q[4], bx = q[5], by = q[6], bz = q[7];
    u[0] = rho; u[1] = rho * vx; u[2] = rho * vy; u[3] = rho * vz;
    u[4] = p / (gamma - R(1)) + R(0.5) * rho * (vx * vx + vy * vy + vz * vz) + R(0.5) * (bx * bx + by * by + bz * bz);
    u[5] = bx; u[6] = by; u[7] = bz;
}

/* ct_mhd.jl:66-72 hll */
static void ct_hll(const real* L, const real* Rq, int d, real gamma, real* f)
{
    real uL = L[1 + d], uR = Rq[1 + d];
    real cL = ct_fast_speed(L, d, gamma), cR = ct_fast_speed(Rq, d, gamma);
    real SL = fmin(fmin(uL - cL, uR - cR), R(0)), SR = fmax(fmax(uL + cL, uR + cR), R(0));
    real FL[NC], FR[NC], UL[NC], UR[NC];
    o_ct_dflux(L, d, gamma, FL); o_ct_dflux(Rq, d, gamma, FR);
    ct_toC(L, gamma, UL); ct_toC(Rq, gamma, UR);
    real ih = R(1) / (SR - SL);
    for (int t = 0; t < NC; ++t) f[t] = (SR * FL[t] - SL * FR[t] + SL * SR * (UR[t] - UL[t])) * ih;
}

/* LLF on the CT state (glmmhd_turb.jl:182-190 with the face-normal B and no psi) */
static void ct_llf(const real* L, const real* Rq, int d, real gamma, real* f)
{
    real FL[NC], FR[NC], UL[NC], UR[NC];
    o_ct_dflux(L, d, gamma, FL); o_ct_dflux(Rq, d, gamma, FR);
    ct_toC(L, gamma, UL); ct_toC(Rq, gamma, UR);
    real smax = fmax(fabs(L[1 + d]) + ct_fast_speed(L, d, gamma), fabs(Rq[1 + d]) + ct_fast_speed(Rq, d, gamma));
    for (int t = 0; t < NC; ++t) f[t] = R(0.5) * (FL[t] + FR[t]) - R(0.5) * smax * (UR[t] - UL[t]);
}

/* HLLD on the CT state: rotate d to x (glmmhd_turb.jl:119-138 cyclic permutation), solve with
 * bn = face B and F[Bn] = 0, rotate back. */
static void ct_hlld(const real* L, const real* Rq, int d, real gamma, real* f)
{
    static const int perm[3][3] = {{0, 1, 2}, {1, 2, 0}, {2, 0, 1}};
    real Lr[9], Rr[9], fx[9];
    for (int a = 0; a < 3; ++a) {
        Lr[1 + a] = L[1 + perm[d][a]]; Lr[5 + a] = L[5 + perm[d][a]];
        Rr[1 + a] = Rq[1 + perm[d][a]]; Rr[5 + a] = Rq[5 + perm[d][a]];
    }
    Lr[0] = L[0]; Lr[4] = L[4]; Lr[8] = R(0);
    Rr[0] = Rq[0]; Rr[4] = Rq[4]; Rr[8] = R(0);
    o_hlld_x(Lr, Rr, gamma, Lr[5], R(0), R(0), fx);
    f[0] = fx[0]; f[4] = fx[4];
    for (int a = 0; a < 3; ++a) { f[1 + perm[d][a]] = fx[1 + a]; f[5 + perm[d][a]] = fx[5 + a]; }
}

/* ct_mhd.jl:74-82 / spike_ctm.cu:170-184 hanc1d; `hs` is the offset applied at c -/+ hs. */
static void ct_hanc1d(const real* q, const real* hs, int d, real dtdx, real gamma, real smallr, real smallp,
                      real* mh)
{
    real a[NC], b[NC], FL[NC], FR[NC], U[NC], Uh[NC];
    for (int t = 0; t < NC; ++t) { a[t] = q[t] - hs[t]; b[t] = q[t] + hs[t]; }
    o_ct_dflux(a, d, gamma, FL); o_ct_dflux(b, d, gamma, FR);
    ct_toC(q, gamma, U);
    real h = R(0.5) * dtdx;
    for (int t = 0; t < NC; ++t) Uh[t] = U[t] - h * (FR[t] - FL[t]);
    real rho = fmax(Uh[0], smallr), ir = R(1) / rho;
    real vx = Uh[1] * ir, vy = Uh[2] * ir, vz = Uh[3] * ir, bx = Uh[5], by = Uh[6], bz = Uh[7];
    real p = fmax((gamma - R(1)) * (Uh[4] - R(0.5) * rho * (vx * vx + vy * vy + vz * vz) -
                                    R(0.5) * (bx * bx + by * by + bz * bz)), smallp);
    mh[0] = rho; mh[1] = vx; mh[2] = vy; mh[3] = vz; mh[4] = p; mh[5] = bx; mh[6] = by; mh[7] = bz;
}

/* Flux at the lower-d face of cell c from the 4-cell stencil m2, m1, c0, p1 (ct_mhd.jl:99-122,
 * spike_ctm.cu:223-245). */
void o_ct_face_flux(const real* m2, const real* m1, const real* c0, const real* p1, real bn, int d,
                    const o_ct_scheme* sc, real* f)
{
    real gamma = R(sc->gamma), smallr = R(sc->smallr), smallp = R(sc->smallp);
    real dtdx = R(sc->dt) / R(sc->dx);
    real sL[NC], sR[NC], hsL[NC], hsR[NC], mhL[NC], mhR[NC], L[NC], Rr[NC];
    for (int t = 0; t < NC; ++t) {
        sL[t] = o_moncen(m1[t] - m2[t], c0[t] - m1[t]);   /* full MonCen slope */
        sR[t] = o_moncen(c0[t] - m1[t], p1[t] - c0[t]);
        if (sc->face_slope == 1) { sL[t] = R(0.5) * sL[t]; sR[t] = R(0.5) * sR[t]; }   /* ct_mhd.jl:37 */
        hsL[t] = R(0.5) * sL[t]; hsR[t] = R(0.5) * sR[t];
    }
    ct_hanc1d(m1, hsL, d, dtdx, gamma, smallr, smallp, mhL);
    ct_hanc1d(c0, hsR, d, dtdx, gamma, smallr, smallp, mhR);
    if (sc->recon == 2) {
        for (int t = 0; t < NC; ++t) {
            L[t] = o_ppm1_spike(m2[t], m1[t], c0[t], R(1)) + mhL[t] - m1[t];
            Rr[t] = o_ppm1_spike(m1[t], c0[t], p1[t], R(-1)) + mhR[t] - c0[t];
        }
    } else {
        for (int t = 0; t < NC; ++t) { L[t] = mhL[t] + hsL[t]; Rr[t] = mhR[t] - hsR[t]; }
    }
    L[0] = fmax(L[0], smallr); L[4] = fmax(L[4], smallp);
    Rr[0] = fmax(Rr[0], smallr); Rr[4] = fmax(Rr[4], smallp);
    L[5 + d] = bn; Rr[5 + d] = bn;
    if (sc->riemann == 3) ct_hlld(L, Rr, d, gamma, f);
    else if (sc->riemann == 0) ct_llf(L, Rr, d, gamma, f);
    else ct_hll(L, Rr, d, gamma, f);
}

/* One CT step: bcc (ct_mhd.jl:85-96), 3 flux passes (:99-122), emf (:139-154), update_hydro
 * (:124-136), faceB (:156-171); order as step! (:174-184).  scratch: 3n (Bcc) + 8n (prim) +
 * 24n (fluxes) + 3n (EMF) = 38n reals. */
int o_ct_step(const real* qin, real* qout, int nx, int ny, int nz, const o_ct_scheme* sc,
              const real* afield, real* scratch)
{
    size_t n = (size_t)nx * ny * nz;
    real gamma = R(sc->gamma), smallr = R(sc->smallr), smallp = R(sc->smallp);
    real dtdx = R(sc->dt) / R(sc->dx);
    real* bcc = scratch;              /* 3n */
    real* prim = bcc + 3 * n;         /* 8n */
    real* F = prim + 8 * n;           /* 24n: F[d*8+t] plane, flux at the lower-d face */
    real* E = F + 24 * n;             /* 3n: ex, ey, ez */
    const real* bxf = qin + 5 * n;
    const real* byf = qin + 6 * n;
    const real* bzf = qin + 7 * n;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                size_t c = cidx(i, j, k, nx, ny);
                bcc[c] = R(0.5) * (bxf[c] + bxf[cidx(wrapi(i + 1, nx), j, k, nx, ny)]);
                bcc[n + c] = R(0.5) * (byf[c] + byf[cidx(i, wrapi(j + 1, ny), k, nx, ny)]);
                bcc[2 * n + c] = R(0.5) * (bzf[c] + bzf[cidx(i, j, wrapi(k + 1, nz), nx, ny)]);
            }
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                size_t c = cidx(i, j, k, nx, ny);
                real u[5] = {qin[c], qin[n + c], qin[2 * n + c], qin[3 * n + c], qin[4 * n + c]};
                real b[3] = {bcc[c], bcc[n + c], bcc[2 * n + c]}, q[NC];
                o_ct_cellprim(u, b, gamma, smallr, smallp, q);
                for (int t = 0; t < NC; ++t) prim[(size_t)t * n + c] = q[t];
            }
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                size_t c = cidx(i, j, k, nx, ny);
                for (int d = 0; d < 3; ++d) {
                    int o1 = d == 0, o2 = d == 1, o3 = d == 2;
                    size_t cm2 = cidx(wrapi(i - 2 * o1, nx), wrapi(j - 2 * o2, ny), wrapi(k - 2 * o3, nz), nx, ny);
                    size_t cm1 = cidx(wrapi(i - o1, nx), wrapi(j - o2, ny), wrapi(k - o3, nz), nx, ny);
                    size_t cp1 = cidx(wrapi(i + o1, nx), wrapi(j + o2, ny), wrapi(k + o3, nz), nx, ny);
                    real m2[NC], m1[NC], c0[NC], p1[NC], f[NC];
                    for (int t = 0; t < NC; ++t) {
                        m2[t] = prim[(size_t)t * n + cm2]; m1[t] = prim[(size_t)t * n + cm1];
                        c0[t] = prim[(size_t)t * n + c]; p1[t] = prim[(size_t)t * n + cp1];
                    }
                    real bn = qin[(size_t)(5 + d) * n + c];
                    o_ct_face_flux(m2, m1, c0, p1, bn, d, sc, f);
                    for (int t = 0; t < NC; ++t) F[(size_t)(d * NC + t) * n + c] = f[t];
                }
            }
#define FX(t, c) F[(size_t)(0 * NC + (t)) * n + (c)]
#define FY(t, c) F[(size_t)(1 * NC + (t)) * n + (c)]
#define FZ(t, c) F[(size_t)(2 * NC + (t)) * n + (c)]
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                size_t c = cidx(i, j, k, nx, ny);
                size_t im = cidx(wrapi(i - 1, nx), j, k, nx, ny), jm = cidx(i, wrapi(j - 1, ny), k, nx, ny);
                size_t km = cidx(i, j, wrapi(k - 1, nz), nx, ny);
                E[2 * n + c] = R(0.25) * (FY(5, c) + FY(5, im) - FX(6, c) - FX(6, jm));
                E[c] = R(0.25) * (FZ(6, c) + FZ(6, jm) - FY(7, c) - FY(7, km));
                E[n + c] = R(0.25) * (FX(7, c) + FX(7, km) - FZ(5, c) - FZ(5, im));
            }
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                size_t c = cidx(i, j, k, nx, ny);
                size_t ip = cidx(wrapi(i + 1, nx), j, k, nx, ny), jp = cidx(i, wrapi(j + 1, ny), k, nx, ny);
                size_t kp = cidx(i, j, wrapi(k + 1, nz), nx, ny);
                real r[5];
                for (int t = 0; t < 5; ++t)
                    r[t] = qin[(size_t)t * n + c] +
                           dtdx * ((FX(t, c) - FX(t, ip)) + (FY(t, c) - FY(t, jp)) + (FZ(t, c) - FZ(t, kp)));
                const real *ex = E, *ey = E + n, *ez = E + 2 * n;
                real bxn = bxf[c] - dtdx * ((ez[jp] - ez[c]) - (ey[kp] - ey[c]));
                real byn = byf[c] - dtdx * ((ex[kp] - ex[c]) - (ez[ip] - ez[c]));
                real bzn = bzf[c] - dtdx * ((ey[ip] - ey[c]) - (ex[jp] - ex[c]));
                if (sc->update_floor && r[0] < smallr) {
                    /* static void floor with the (pre-update) cell-centred B, as the GLM update */
                    real emag = R(0.5) * (bcc[c] * bcc[c] + bcc[n + c] * bcc[n + c] + bcc[2 * n + c] * bcc[2 * n + c]);
                    r[0] = smallr; r[1] = R(0); r[2] = R(0); r[3] = R(0); r[4] = smallp / (gamma - R(1)) + emag;
                }
                if (sc->do_turb && afield && r[0] >= R(sc->turb_min_rho)) {
                    real dx = R(sc->dx), acc[3], smallc = R(sc->smallc), dt = R(sc->dt);
                    o_turb_interp(afield, sc->turb_gs, i, j, k, dx, R(nx) * dx, R(ny) * dx, R(nz) * dx, R(sc->ramp), acc);
                    real rho = r[0], rhom = fmax(rho, smallr), sc2 = smallc * smallc, e = r[4];
                    e = fmax(e - R(0.5) * r[1] * r[1] / rhom, rho * sc2);
                    e = fmax(e - R(0.5) * r[2] * r[2] / rhom, rho * sc2);
                    e = fmax(e - R(0.5) * r[3] * r[3] / rhom, rho * sc2);
                    real m2 = r[1] + rhom * acc[0] * dt, m3 = r[2] + rhom * acc[1] * dt, m4 = r[3] + rhom * acc[2] * dt;
                    e = fmax(e + R(0.5) * m2 * m2 / rhom, rho * sc2);
                    e = fmax(e + R(0.5) * m3 * m3 / rhom, rho * sc2);
                    e = fmax(e + R(0.5) * m4 * m4 / rhom, rho * sc2);
                    r[1] = m2; r[2] = m3; r[3] = m4; r[4] = e;
                }
                for (int t = 0; t < 5; ++t) qout[(size_t)t * n + c] = r[t];
                qout[5 * n + c] = bxn; qout[6 * n + c] = byn; qout[7 * n + c] = bzn;
            }
#undef FX
#undef FY
#undef FZ
    return 0;
}

size_t o_ct_step_scratch_reals(int nx, int ny, int nz) { return (size_t)38 * nx * ny * nz; }

/* Passive species in the CT step (spike_ctm.cu -DNSP: fractions in the tile :217-219, :260-263;
 * slopes and PPM edges :88-90, :108-119; hanc1d's advective predictor X - h un dX (h = dt/dx / 2,
 * un the cell's normal velocity) :178-181; the CMA flux F_s = F_rho X of the upwind face state
 * :240-243; the update :289-292).  Species do not feed back on the MHD planes, so their face flux
 * needs only the fractions, the two cells' normal velocities and the MHD mass flux of the face. */
static real ct_species_flux(real m2x, real m1x, real c0x, real p1x, real unL, real unR, real f0,
                            const o_ct_scheme* sc)
{
    real dtdx = R(sc->dt) / R(sc->dx);
    real sL = o_moncen(m1x - m2x, c0x - m1x), sR = o_moncen(c0x - m1x, p1x - c0x);
    if (sc->face_slope == 1) { sL = R(0.5) * sL; sR = R(0.5) * sR; }
    real mhL = m1x - R(0.5) * dtdx * unL * sL, mhR = c0x - R(0.5) * dtdx * unR * sR;
    real L, Rr;
    if (sc->recon == 2) {
        L = o_ppm1_spike(m2x, m1x, c0x, R(1)) + mhL - m1x;
        Rr = o_ppm1_spike(m1x, c0x, p1x, R(-1)) + mhR - c0x;
    } else {
        L = mhL + R(0.5) * sL;
        Rr = mhR - R(0.5) * sR;
    }
    return f0 >= R(0) ? f0 * L : f0 * Rr;
}

/* One CT step of a (8 + nsp)-plane state: o_ct_step on the 8 MHD planes, then the nsp species
 * planes rho X_s with the mass fluxes that step computed.  Where the update floors the density the
 * species keep their fraction: rho X = smallr X_old, X_old = rho X / max(rho, smallr) before the step
 * (the floor of the GLM march's species).  scratch: o_ct_step_sp_scratch_reals. */
int o_ct_step_sp(const real* qin, real* qout, int nx, int ny, int nz, int nsp, const o_ct_scheme* sc,
                 const real* afield, real* scratch)
{
    if (nsp < 0 || nsp > 2) return 1;
    int e = o_ct_step(qin, qout, nx, ny, nz, sc, afield, scratch);
    if (e || nsp == 0) return e;
    size_t n = (size_t)nx * ny * nz;
    real smallr = R(sc->smallr), dtdx = R(sc->dt) / R(sc->dx);
    const real* prim = scratch + 3 * n;      /* o_ct_step's primitives: v_d in planes 1..3 */
    const real* F = scratch + 11 * n;         /* its fluxes: F[(d*8 + t) n + c] at the lower-d face */
    real* X = scratch + 38 * n;               /* nsp n fractions */
    real* FS = X + (size_t)nsp * n;           /* 3 nsp n species fluxes: FS[(d*nsp + s) n + c] */
    for (int sp = 0; sp < nsp; ++sp)
#pragma omp parallel for schedule(static)
        for (size_t c = 0; c < n; ++c) X[(size_t)sp * n + c] = qin[(size_t)(8 + sp) * n + c] / fmax(qin[c], smallr);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                size_t c = cidx(i, j, k, nx, ny);
                for (int d = 0; d < 3; ++d) {
                    int o1 = d == 0, o2 = d == 1, o3 = d == 2;
                    size_t cm2 = cidx(wrapi(i - 2 * o1, nx), wrapi(j - 2 * o2, ny), wrapi(k - 2 * o3, nz), nx, ny);
                    size_t cm1 = cidx(wrapi(i - o1, nx), wrapi(j - o2, ny), wrapi(k - o3, nz), nx, ny);
                    size_t cp1 = cidx(wrapi(i + o1, nx), wrapi(j + o2, ny), wrapi(k + o3, nz), nx, ny);
                    real unL = prim[(size_t)(1 + d) * n + cm1], unR = prim[(size_t)(1 + d) * n + c];
                    real f0 = F[(size_t)(d * 8) * n + c];
                    for (int sp = 0; sp < nsp; ++sp) {
                        const real* x = X + (size_t)sp * n;
                        FS[(size_t)(d * nsp + sp) * n + c] = ct_species_flux(x[cm2], x[cm1], x[c], x[cp1], unL, unR, f0, sc);
                    }
                }
            }
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                size_t c = cidx(i, j, k, nx, ny);
                size_t ip = cidx(wrapi(i + 1, nx), j, k, nx, ny), jp = cidx(i, wrapi(j + 1, ny), k, nx, ny);
                size_t kp = cidx(i, j, wrapi(k + 1, nz), nx, ny);
                /* the density update of o_ct_step, for its floor decision */
                real r0 = qin[c] + dtdx * ((F[c] - F[ip]) + (F[(size_t)8 * n + c] - F[(size_t)8 * n + jp]) +
                                           (F[(size_t)16 * n + c] - F[(size_t)16 * n + kp]));
                int floored = sc->update_floor && r0 < smallr;
                for (int sp = 0; sp < nsp; ++sp) {
                    const real* fx = FS + (size_t)(0 * nsp + sp) * n;
                    const real* fy = FS + (size_t)(1 * nsp + sp) * n;
                    const real* fz = FS + (size_t)(2 * nsp + sp) * n;
                    real r = qin[(size_t)(8 + sp) * n + c] + dtdx * ((fx[c] - fx[ip]) + (fy[c] - fy[jp]) + (fz[c] - fz[kp]));
                    if (floored) r = smallr * X[(size_t)sp * n + c];
                    qout[(size_t)(8 + sp) * n + c] = r;
                }
            }
    return 0;
}

size_t o_ct_step_sp_scratch_reals(int nx, int ny, int nz, int nsp) { return (size_t)(38 + 4 * nsp) * nx * ny * nz; }

/* ct_mhd.jl:191-194 divB_max (face divergence) */
double o_ct_divb_max(const real* q, int nx, int ny, int nz, double dx)
{
    size_t n = (size_t)nx * ny * nz;
    const real *bxf = q + 5 * n, *byf = q + 6 * n, *bzf = q + 7 * n;
    double m = 0.0;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                size_t c = cidx(i, j, k, nx, ny);
                double d = (((double)bxf[cidx(wrapi(i + 1, nx), j, k, nx, ny)] - (double)bxf[c]) +
                            ((double)byf[cidx(i, wrapi(j + 1, ny), k, nx, ny)] - (double)byf[c]) +
                            ((double)bzf[cidx(i, j, wrapi(k + 1, nz), nx, ny)] - (double)bzf[c])) / dx;
                if (fabs(d) > m) m = fabs(d);
            }
    return m;
}

/* CFL on the CT state: the GLM rule glmmhd_turb.jl:1963-1971 applied to (U, Bcc). */
double o_ct_ctot_max(const real* q, int nx, int ny, int nz, double gamma_, double smallr_, double smallp_)
{
    size_t n = (size_t)nx * ny * nz;
    real gamma = R(gamma_), smallr = R(smallr_), smallp = R(smallp_);
    double m = 0.0;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                size_t c = cidx(i, j, k, nx, ny);
                real u[5] = {q[c], q[n + c], q[2 * n + c], q[3 * n + c], q[4 * n + c]};
                real b[3] = {R(0.5) * (q[5 * n + c] + q[5 * n + cidx(wrapi(i + 1, nx), j, k, nx, ny)]),
                             R(0.5) * (q[6 * n + c] + q[6 * n + cidx(i, wrapi(j + 1, ny), k, nx, ny)]),
                             R(0.5) * (q[7 * n + c] + q[7 * n + cidx(i, j, wrapi(k + 1, nz), nx, ny)])};
                real p[NC];
                o_ct_cellprim(u, b, gamma, smallr, smallp, p);
                real ct = (fabs(p[1]) + ct_fast_speed(p, 0, gamma)) + (fabs(p[2]) + ct_fast_speed(p, 1, gamma)) +
                          (fabs(p[3]) + ct_fast_speed(p, 2, gamma));
                if ((double)ct > m) m = (double)ct;
            }
    return m;
}
