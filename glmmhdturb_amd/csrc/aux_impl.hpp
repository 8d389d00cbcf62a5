// aux_impl.hpp -- reductions and initial-condition kernels for gfx950.
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
// CFL: max over cells of sum_d(|v_d| + c_f(B_d))  (glmmhd_turb.jl:1963-1978), or with FAST the max
// over cells of max_d c_f(B_d) (the "max fast speed" cleaning speed).  One atomic per wavefront on
// the bit pattern of a non-negative double (NaN compares largest).
// ---------------------------------------------------------------------------------------------
template <typename T, bool FAST>
__global__ void __launch_bounds__(AUX_NT) ctot_kernel(Ptr9<const T> q9, size_t n, Phys<T> s, unsigned long long* out)
{
    double cmax = 0.0;
    for (size_t c = (size_t)blockIdx.x * AUX_NT + threadIdx.x; c < n; c += (size_t)gridDim.x * AUX_NT) {
        T u[9], p[9];
#pragma unroll
        for (int v = 0; v < 9; ++v) u[v] = q9[v][c];
        cons2prim(u, p, s);
        double ct = FAST ? fast_max_speed(p, s) : cfl_speed(p, s);
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

}  // namespace glm
