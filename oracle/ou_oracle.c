/*
 * ou_oracle.c -- CPU restatement of the OU turbulent forcing (TEST INFRASTRUCTURE ONLY;
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it).
 *
 * Follows glmmhd_turb.jl:2070-2161 (kfreq, ou_init_kernel!, ou_step_kernel!, Forcing,
 * ou_advance!) with two documented changes required by the task:
 *   - the Gaussian increments come from a seeded counter-based generator (Philox4x32-10,
 *     Salmon et al. SC'11, "Parallel random numbers: as easy as 1, 2, 3") instead of the
 *     reference's unseeded CUDA.randn! (glmmhd_turb.jl:2130);
 *   - the active band is a shell kmin <= |k| <= kmax (the reference fixes 1..n/2, :2083),
 *     so the solenoidal shells 1..3 of the turbulence workload can be expressed.
 * The inverse FFT (AbstractFFTs plan_bfft!, :2139/:2156) is replaced by a separable
 * unnormalised inverse DFT along x, then y, then z, with a twiddle table cos/sin(2 pi m/g).
 *
 * All OU arithmetic is double precision; the acceleration field is returned in `real`.
 * Layouts: Fhat is 3 components x g^3 modes, stored as re/im planes:
 *   fre[c*g3 + a + g*(b + g*d)], fim[...] for mode indices (a,b,d) along (x,y,z).
 * afield is 3 SoA planes of g^3 reals, point (i,j,k) at i + g*(j + g*k).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef REAL
#define REAL double
#endif
typedef REAL real;

/* Philox4x32-10 (Random123).  ctr: 4 words, key: 2 words. */
void o_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Six standard normals for one (mode, OU step): two Philox blocks, Box-Muller on word pairs. */
void o_ou_gauss6(uint64_t seed, uint64_t step, uint32_t mode, double g[6])
{
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    double z[8];
    for (int blk = 0; blk < 2; ++blk) {
        uint32_t ctr[4] = {mode, (uint32_t)step, (uint32_t)(step >> 32), (uint32_t)blk};
        uint32_t w[4];
        o_philox4x32_10(ctr, key, w);
        for (int p = 0; p < 2; ++p) {
            double u1 = ((double)w[2 * p] + 0.5) * 2.3283064365386963e-10;   /* 2^-32 */
            double u2 = ((double)w[2 * p + 1] + 0.5) * 2.3283064365386963e-10;
            double r = sqrt(-2.0 * log(u1));
            double th = 6.283185307179586 * u2;
            z[4 * blk + 2 * p] = r * cos(th);
            z[4 * blk + 2 * p + 1] = r * sin(th);
        }
    }
    for (int q = 0; q < 6; ++q) g[q] = z[q];
}

/* glmmhd_turb.jl:2075 kfreq (0-based index) */
static inline double kfreq(int i, int n) { return i <= n / 2 ? (double)i : (double)(i - n); }

/* glmmhd_turb.jl:2078-2091 ou_init_kernel!: sqrt(P)=|k|^-1 on the active shell, unit k. */
void o_ou_init(int g, double kmin, double kmax, double* sqp, double* unitk /*3*g^3*/)
{
    size_t g3 = (size_t)g * g * g;
    for (int d = 0; d < g; ++d)
        for (int b = 0; b < g; ++b)
            for (int a = 0; a < g; ++a) {
                size_t m = (size_t)a + (size_t)g * ((size_t)b + (size_t)g * d);
                double kx = kfreq(a, g), ky = kfreq(b, g), kz = kfreq(d, g);
                double kk = sqrt(kx * kx + ky * ky + kz * kz);
                if (kk >= kmin && kk <= kmax) {
                    sqp[m] = 1.0 / kk;
                    unitk[m] = kx / kk; unitk[g3 + m] = ky / kk; unitk[2 * g3 + m] = kz / kk;
                } else {
                    sqp[m] = 0.0;
                    unitk[m] = 0.0; unitk[g3 + m] = 0.0; unitk[2 * g3 + m] = 0.0;
                }
            }
}

/* glmmhd_turb.jl:2094-2113 ou_step_kernel! for every mode; the Gaussians of a mode are drawn
 * only where sqrt(P) != 0 (inactive modes would be multiplied by zero). */
void o_ou_step(int g, double* fre, double* fim, const double* sqp, const double* unitk, double alpha,
               double beta, double comp, double sol, uint64_t seed, uint64_t step)
{
    size_t g3 = (size_t)g * g * g;
    for (size_t m = 0; m < g3; ++m) {
        double s = sqp[m];
        double gg[6] = {0, 0, 0, 0, 0, 0};
        if (s != 0.0) o_ou_gauss6(seed, step, (uint32_t)m, gg);
        double wr[3] = {gg[0] * s, gg[2] * s, gg[4] * s};
        double wi[3] = {gg[1] * s, gg[3] * s, gg[5] * s};
        double ux = unitk[m], uy = unitk[g3 + m], uz = unitk[2 * g3 + m];
        double kr = ux * wr[0] + uy * wr[1] + uz * wr[2];
        double ki = ux * wi[0] + uy * wi[1] + uz * wi[2];
        double u[3] = {ux, uy, uz};
        for (int c = 0; c < 3; ++c) {
            double cr = u[c] * kr, ci = u[c] * ki;
            double pr = comp * cr + sol * (wr[c] - cr);
            double pi = comp * ci + sol * (wi[c] - ci);
            fre[c * g3 + m] = alpha * fre[c * g3 + m] + beta * pr;
            fim[c * g3 + m] = alpha * fim[c * g3 + m] + beta * pi;
        }
    }
}

/* One 1-D unnormalised inverse DFT pass along axis `ax` (0=x,1=y,2=z) of a g^3 complex field:
 * out[..x..] = sum_a in[..a..] * exp(+2 pi i a x / g), summed in a = 0..g-1 order. */
static void idft_pass(int g, int ax, const double* ir, const double* ii, double* orr, double* oi,
                      const double* cs, const double* sn)
{
    size_t g3 = (size_t)g * g * g;
    size_t stride = ax == 0 ? 1 : (ax == 1 ? (size_t)g : (size_t)g * g);
    for (size_t base = 0; base < g3; ++base) {
        int pos = (int)((base / stride) % (size_t)g);
        size_t line0 = base - (size_t)pos * stride;
        double sr = 0.0, si = 0.0;
        for (int a = 0; a < g; ++a) {
            int t = (a * pos) % g;
            double xr = ir[line0 + (size_t)a * stride], xi = ii[line0 + (size_t)a * stride];
            sr += xr * cs[t] - xi * sn[t];
            si += xr * sn[t] + xi * cs[t];
        }
        orr[base] = sr; oi[base] = si;
    }
}

/* glmmhd_turb.jl:2154-2158: for each component, inverse FFT then real part times amp/g^3.
 * cs/sn: twiddle table cos/sin(2 pi m / g), m = 0..g-1.  work: 4*g^3 doubles. */
void o_ou_afield(int g, const double* fre, const double* fim, double amp, const double* cs,
                 const double* sn, double* work, real* afield)
{
    size_t g3 = (size_t)g * g * g;
    double *ar = work, *ai = work + g3, *br = work + 2 * g3, *bi = work + 3 * g3;
    double scale = amp / ((double)g * (double)g * (double)g);
    for (int c = 0; c < 3; ++c) {
        idft_pass(g, 0, fre + c * g3, fim + c * g3, ar, ai, cs, sn);
        idft_pass(g, 1, ar, ai, br, bi, cs, sn);
        idft_pass(g, 2, br, bi, ar, ai, cs, sn);
        for (size_t m = 0; m < g3; ++m) afield[c * g3 + m] = (real)(ar[m] * scale);
    }
}
