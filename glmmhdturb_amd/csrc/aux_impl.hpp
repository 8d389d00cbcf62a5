// aux_impl.hpp -- reductions, initial conditions and OU forcing kernels for gfx950.
#pragma once
#include "physics.hpp"

namespace glm {

constexpr int AUX_NT = 256;

GLM_DEV int awrap(int i, int n)
{
    int m = i % n;
    return m < 0 ? m + n : m;
}

// ---------------------------------------------------------------------------------------------
// CFL: max over cells of sum_d(|v_d| + c_f(B_d))  (glmmhd_turb.jl:1963-1978).  One atomic per
// wavefront on the bit pattern of a non-negative double (NaN compares largest).
// ---------------------------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(AUX_NT) ctot_kernel(Ptr9<const T> q9, size_t n, Phys<T> s, unsigned long long* out)
{
    double cmax = 0.0;
    for (size_t c = (size_t)blockIdx.x * AUX_NT + threadIdx.x; c < n; c += (size_t)gridDim.x * AUX_NT) {
        T u[9], p[9];
#pragma unroll
        for (int v = 0; v < 9; ++v) u[v] = q9[v][c];
        cons2prim(u, p, s);
        double ct = cfl_speed(p, s);
        if (!(ct <= cmax)) cmax = ct;
    }
    unsigned long long bits = (unsigned long long)__double_as_longlong(cmax);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        unsigned long long o = __shfl_xor(bits, off, 64);
        bits = o > bits ? o : bits;
    }
    if ((threadIdx.x & 63) == 0 && bits != 0ull) atomicMax(out, bits);
}

__global__ void clear_u64_kernel(unsigned long long* p, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// ---------------------------------------------------------------------------------------------
// Diagnostics (glmmhd_turb.jl:1996-2018 plus the energy budget).  Deterministic: block b owns a
// fixed cell range; each thread keeps Neumaier-compensated double sums over a fixed stride; the
// block combines them in a fixed tree; a one-block kernel combines the block partials in order.
// ---------------------------------------------------------------------------------------------
constexpr int NDSUM = 7;   // mass, ke2, ekin, emag, cross helicity, total E, psi^2
constexpr int NDIAG = 10;  // + divbmax, rhomin, finite

struct KSum {
    double s, c;
};
GLM_DEV void kadd(KSum& a, double x)
{
    double t = a.s + x;
    if (fabs(a.s) >= fabs(x)) a.c += (a.s - t) + x;
    else a.c += (x - t) + a.s;
    a.s = t;
}
GLM_DEV KSum kmerge(KSum a, KSum b)
{
    KSum r = a;
    kadd(r, b.s);
    r.c += b.c;
    return r;
}

struct DiagPart {
    KSum sum[NDSUM];
    double divb, rhomin, finite;
};

template <typename T>
__global__ void __launch_bounds__(AUX_NT) diag_partial_kernel(Ptr9<const T> q9, int nx, int ny, int nz, double dx,
                                                              double smallr, size_t cells_per_block, DiagPart* parts)
{
    __shared__ DiagPart sh[AUX_NT];
    const size_t n = (size_t)nx * ny * nz;
    const size_t c0 = (size_t)blockIdx.x * cells_per_block;
    const size_t c1 = c0 + cells_per_block < n ? c0 + cells_per_block : n;
    DiagPart d;
#pragma unroll
    for (int k = 0; k < NDSUM; ++k) d.sum[k] = KSum{0.0, 0.0};
    d.divb = 0.0; d.rhomin = INFINITY; d.finite = 1.0;
    for (size_t c = c0 + threadIdx.x; c < c1; c += AUX_NT) {
        int i = (int)(c % nx), j = (int)((c / nx) % ny), k = (int)(c / ((size_t)nx * ny));
        double r = q9[0][c], a = q9[1][c], b = q9[2][c], e = q9[3][c];
        double m2 = a * a + b * b + e * e;
        double Bx = q9[5][c], By = q9[6][c], Bz = q9[7][c];
        kadd(d.sum[0], r);
        kadd(d.sum[1], m2 / fmax(r, smallr));
        kadd(d.sum[2], 0.5 * m2 / r);
        kadd(d.sum[3], 0.5 * (Bx * Bx + By * By + Bz * Bz));
        kadd(d.sum[4], (a * Bx + b * By + e * Bz) / r);
        kadd(d.sum[5], (double)q9[4][c]);
        double ps = q9[8][c];
        kadd(d.sum[6], ps * ps);
        d.rhomin = fmin(d.rhomin, r);
        const size_t pl = (size_t)nx * ny;
        double db = fabs(((double)q9[5][(size_t)awrap(i + 1, nx) + (size_t)nx * j + pl * k] -
                          (double)q9[5][(size_t)awrap(i - 1, nx) + (size_t)nx * j + pl * k]) +
                         ((double)q9[6][(size_t)i + (size_t)nx * awrap(j + 1, ny) + pl * k] -
                          (double)q9[6][(size_t)i + (size_t)nx * awrap(j - 1, ny) + pl * k]) +
                         ((double)q9[7][(size_t)i + (size_t)nx * j + pl * awrap(k + 1, nz)] -
                          (double)q9[7][(size_t)i + (size_t)nx * j + pl * awrap(k - 1, nz)])) /
                    (2.0 * dx);
        d.divb = fmax(d.divb, db);
#pragma unroll
        for (int v = 0; v < 9; ++v)
            if (!isfinite((double)q9[v][c])) d.finite = 0.0;
    }
    sh[threadIdx.x] = d;
    __syncthreads();
    for (int w = AUX_NT / 2; w >= 1; w >>= 1) {
        if ((int)threadIdx.x < w) {
            DiagPart x = sh[threadIdx.x], y = sh[threadIdx.x + w];
#pragma unroll
            for (int k = 0; k < NDSUM; ++k) x.sum[k] = kmerge(x.sum[k], y.sum[k]);
            x.divb = fmax(x.divb, y.divb);
            x.rhomin = fmin(x.rhomin, y.rhomin);
            x.finite = fmin(x.finite, y.finite);
            sh[threadIdx.x] = x;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) parts[blockIdx.x] = sh[0];
}

__global__ void diag_final_kernel(const DiagPart* parts, int nparts, double* out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    DiagPart d = parts[0];
    for (int b = 1; b < nparts; ++b) {
        DiagPart y = parts[b];
        for (int k = 0; k < NDSUM; ++k) d.sum[k] = kmerge(d.sum[k], y.sum[k]);
        d.divb = fmax(d.divb, y.divb);
        d.rhomin = fmin(d.rhomin, y.rhomin);
        d.finite = fmin(d.finite, y.finite);
    }
    // out: mass, ke2, divbmax, rhomin, finite, ekin, emag, cross helicity, etot, psi2
    out[0] = d.sum[0].s + d.sum[0].c;
    out[1] = d.sum[1].s + d.sum[1].c;
    out[2] = d.divb;
    out[3] = d.rhomin;
    out[4] = d.finite;
    out[5] = d.sum[2].s + d.sum[2].c;
    out[6] = d.sum[3].s + d.sum[3].c;
    out[7] = d.sum[4].s + d.sum[4].c;
    out[8] = d.sum[5].s + d.sum[5].c;
    out[9] = d.sum[6].s + d.sum[6].c;
}

// ---------------------------------------------------------------------------------------------
// Initial conditions
// ---------------------------------------------------------------------------------------------
// glmmhd_turb.jl:2021-2036 ic_ot_kernel! (z-invariant Orszag-Tang)
template <typename T>
__global__ void ic_ot_kernel(Ptr9<T> q9, int nx, int ny, int nz, T gamma)
{
    size_t n = (size_t)nx * ny * nz;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x) {
        int i = (int)(c % nx), j = (int)((c / nx) % ny);
        const T pi = T(3.14159265358979323846);
        const T twopi = T(2) * pi;
        T x = (T(i) + T(0.5)) / T(nx), y = (T(j) + T(0.5)) / T(ny);
        T rho = T(25) / (T(36) * pi);
        T pr = T(5) / (T(12) * pi);
        T vx = -sin(twopi * y), vy = sin(twopi * x), vz = T(0);
        T sb = T(1) / tsqrt(T(4) * pi);
        T bx = -sb * sin(twopi * y), by = sb * sin(T(2) * twopi * x), bz = T(0);
        T ekin = T(0.5) * rho * (vx * vx + vy * vy + vz * vz);
        T emag = T(0.5) * (bx * bx + by * by + bz * bz);
        q9[0][c] = rho; q9[1][c] = rho * vx; q9[2][c] = rho * vy; q9[3][c] = rho * vz;
        q9[4][c] = pr / (gamma - T(1)) + ekin + emag;
        q9[5][c] = bx; q9[6][c] = by; q9[7][c] = bz; q9[8][c] = T(0);
    }
}

// glmmhd_turb.jl:2164-2172 ic_uniform_kernel! with B0 = (b0x, b0y, b0z)
template <typename T>
__global__ void ic_uniform_kernel(Ptr9<T> q9, size_t n, T rho0, T p0, T bx, T by, T bz, T gamma)
{
    T emag = T(0.5) * (bx * bx + by * by + bz * bz);
    T e = p0 / (gamma - T(1)) + emag;
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x) {
        q9[0][c] = rho0; q9[1][c] = T(0); q9[2][c] = T(0); q9[3][c] = T(0); q9[4][c] = e;
        q9[5][c] = bx; q9[6][c] = by; q9[7][c] = bz; q9[8][c] = T(0);
    }
}

// spike_mhd.cu:304-317 init (the gradient throughput IC)
template <typename T>
__global__ void ic_gradient_kernel(Ptr9<T> q9, size_t n, T gamma)
{
    for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (size_t)gridDim.x * blockDim.x) {
        T ph = T(0.001) * (T)(c % 911);
        T r = T(1) + T(0.3) * ph, u = T(0.1) * ph, v = T(-0.1) * ph, w = T(0.05) * ph, p = T(2.5) + T(0.5) * ph;
        T bx = T(0.2) + T(0.1) * ph, by = T(0.1) * ph, bz = T(-0.05) * ph;
        T ekin = T(0.5) * r * (u * u + v * v + w * w), emag = T(0.5) * (bx * bx + by * by + bz * bz);
        q9[0][c] = r; q9[1][c] = r * u; q9[2][c] = r * v; q9[3][c] = r * w;
        q9[4][c] = p / (gamma - T(1)) + ekin + emag;
        q9[5][c] = bx; q9[6][c] = by; q9[7][c] = bz; q9[8][c] = T(0);
    }
}

// ---------------------------------------------------------------------------------------------
// OU forcing (glmmhd_turb.jl:2070-2161) with a seeded Philox4x32-10 generator and a
// kmin <= |k| <= kmax shell; all in double.  Fourier layout: component planes of g^3 modes,
// mode (a,b,d) at a + g*(b + g*d); re and im in separate arrays.
// ---------------------------------------------------------------------------------------------
GLM_DEV void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint32_t lo0 = 0xD2511F53u * c[0], hi0 = __umulhi(0xD2511F53u, c[0]);
        uint32_t lo1 = 0xCD9E8D57u * c[2], hi1 = __umulhi(0xCD9E8D57u, c[2]);
        uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
}

GLM_DEV void ou_gauss6(unsigned long long seed, unsigned long long step, uint32_t mode, double g[6])
{
    double z[8];
#pragma unroll
    for (int blk = 0; blk < 2; ++blk) {
        uint32_t c[4] = {mode, (uint32_t)step, (uint32_t)(step >> 32), (uint32_t)blk};
        philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            double u1 = ((double)c[2 * p] + 0.5) * 2.3283064365386963e-10;
            double u2 = ((double)c[2 * p + 1] + 0.5) * 2.3283064365386963e-10;
            double r = sqrt(-2.0 * log(u1));
            double th = 6.283185307179586 * u2;
            z[4 * blk + 2 * p] = r * cos(th);
            z[4 * blk + 2 * p + 1] = r * sin(th);
        }
    }
#pragma unroll
    for (int q = 0; q < 6; ++q) g[q] = z[q];
}

// glmmhd_turb.jl:2078-2091 ou_init_kernel! on a shell
__global__ void ou_init_kernel(int g, double kmin, double kmax, double* sqp, double* unitk)
{
    size_t g3 = (size_t)g * g * g;
    size_t m = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= g3) return;
    int a = (int)(m % g), b = (int)((m / g) % g), d = (int)(m / ((size_t)g * g));
    auto kf = [&](int i) { return i <= g / 2 ? (double)i : (double)(i - g); };
    double kx = kf(a), ky = kf(b), kz = kf(d);
    double kk = sqrt(kx * kx + ky * ky + kz * kz);
    bool on = kk >= kmin && kk <= kmax;
    sqp[m] = on ? 1.0 / kk : 0.0;
    unitk[m] = on ? kx / kk : 0.0;
    unitk[g3 + m] = on ? ky / kk : 0.0;
    unitk[2 * g3 + m] = on ? kz / kk : 0.0;
}

// glmmhd_turb.jl:2094-2113 ou_step_kernel!
__global__ void ou_step_kernel(int g, double* fre, double* fim, const double* sqp, const double* unitk, double alpha,
                               double beta, double comp, double sol, unsigned long long seed, unsigned long long step)
{
    size_t g3 = (size_t)g * g * g;
    size_t m = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= g3) return;
    double s = sqp[m];
    double gg[6] = {0, 0, 0, 0, 0, 0};
    if (s != 0.0) ou_gauss6(seed, step, (uint32_t)m, gg);
    double wr[3] = {gg[0] * s, gg[2] * s, gg[4] * s};
    double wi[3] = {gg[1] * s, gg[3] * s, gg[5] * s};
    double u[3] = {unitk[m], unitk[g3 + m], unitk[2 * g3 + m]};
    double kr = u[0] * wr[0] + u[1] * wr[1] + u[2] * wr[2];
    double ki = u[0] * wi[0] + u[1] * wi[1] + u[2] * wi[2];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double cr = u[c] * kr, ci = u[c] * ki;
        double pr = comp * cr + sol * (wr[c] - cr);
        double pi = comp * ci + sol * (wi[c] - ci);
        fre[c * g3 + m] = alpha * fre[c * g3 + m] + beta * pr;
        fim[c * g3 + m] = alpha * fim[c * g3 + m] + beta * pi;
    }
}

// One 1-D unnormalised inverse DFT pass along axis ax over 3 stacked g^3 components.
__global__ void idft_pass_kernel(int g, int ax, const double* ir, const double* ii, double* orr, double* oi,
                                 const double* cs, const double* sn)
{
    size_t g3 = (size_t)g * g * g;
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * g3) return;
    size_t comp = t / g3, base = t % g3;
    size_t stride = ax == 0 ? 1 : (ax == 1 ? (size_t)g : (size_t)g * g);
    int pos = (int)((base / stride) % (size_t)g);
    size_t line0 = comp * g3 + base - (size_t)pos * stride;
    double sr = 0.0, si = 0.0;
    for (int a = 0; a < g; ++a) {
        int tw = (a * pos) % g;
        double xr = ir[line0 + (size_t)a * stride], xi = ii[line0 + (size_t)a * stride];
        sr += xr * cs[tw] - xi * sn[tw];
        si += xr * sn[tw] + xi * cs[tw];
    }
    orr[t] = sr;
    oi[t] = si;
}

template <typename T>
__global__ void afield_kernel(int g, const double* re, double scale, T* afield)
{
    size_t n = 3 * (size_t)g * g * g;
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) afield[t] = (T)(re[t] * scale);
}

}  // namespace glm
