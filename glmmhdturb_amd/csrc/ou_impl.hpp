// ou_impl.hpp -- OU turbulent forcing (glmmhd_turb.jl:2070-2161) for gfx950, on the band of
// active modes only, batched over boxes.
//
// The forcing spectrum is zero outside the shell kmin <= |k| <= kmax, so every mode that can be
// non-zero has |kx|, |ky|, |kz| <= K = floor(kmax).  The OU state therefore lives on the band
// of nb = min(2K+1, g) mode indices per axis (nb^3 modes instead of g^3), and the unnormalised
// inverse DFT to the g^3 forcing grid is done as three separable passes that sum over band
// modes only:  x pass (nb*nb lines -> g points each), y pass, z pass (real part only, scaled and
// stored in the state precision).  Each output sums its terms in increasing mode index, the same
// order as a dense separable DFT, whose remaining terms are zero.
//
// One launch serves many boxes: blockIdx.y is the box, and a box advances either always (the
// single-box path passes its OU step counter) or when the ensemble step has marked it due.
#pragma once
#include "physics.hpp"
#include "ou_launch.hpp"

namespace glm {

// Grid mode index of band index e: 0..K, then g-K..g-1 (the identity when the band is the grid).
GLM_DEV int band_mode(int e, int nb, int K, int g) { return (nb == g || e <= K) ? e : g - (nb - e); }

// glmmhd_turb.jl:2072-2075 kfreq
GLM_DEV double kfreq(int i, int g) { return i <= g / 2 ? (double)i : (double)(i - g); }

GLM_DEV bool ou_active(const OUBox& b, long long step_all, unsigned long long& step)
{
    if (step_all >= 0) { step = (unsigned long long)step_all; return true; }
    step = b.step;
    return b.due != 0;
}

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

// Six standard normals for (seed, OU step, grid mode): two Philox4x32-10 blocks with counter
// (mode, step lo, step hi, block) and key (seed lo, seed hi), Box-Muller on word pairs.
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

// glmmhd_turb.jl:2078-2091 (sqrt(P) = 1/|k| and the unit vector on the shell, evaluated in place)
// and :2094-2113 ou_step_kernel! (Helmholtz split, OU update) on the band modes.
__global__ void ou_band_step_kernel(const OUBox* boxes, long long step_all)
{
    const OUBox& b = boxes[blockIdx.y];
    unsigned long long step;
    if (!ou_active(b, step_all, step)) return;
    const int g = b.g, nb = b.nb, K = b.K;
    const size_t n3 = (size_t)nb * nb * nb;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n3; e += (size_t)gridDim.x * blockDim.x) {
        const int ea = (int)(e % nb), eb = (int)((e / nb) % nb), ed = (int)(e / ((size_t)nb * nb));
        const int ma = band_mode(ea, nb, K, g), mb = band_mode(eb, nb, K, g), md = band_mode(ed, nb, K, g);
        const double kx = kfreq(ma, g), ky = kfreq(mb, g), kz = kfreq(md, g);
        const double kk = sqrt(kx * kx + ky * ky + kz * kz);
        const bool on = kk >= b.kmin && kk <= b.kmax;
        const double s = on ? 1.0 / kk : 0.0;
        double u[3] = {on ? kx / kk : 0.0, on ? ky / kk : 0.0, on ? kz / kk : 0.0};
        double gg[6] = {0, 0, 0, 0, 0, 0};
        const uint32_t m = (uint32_t)((size_t)ma + (size_t)g * ((size_t)mb + (size_t)g * md));
        if (s != 0.0) ou_gauss6(b.seed, step, m, gg);
        double wr[3] = {gg[0] * s, gg[2] * s, gg[4] * s};
        double wi[3] = {gg[1] * s, gg[3] * s, gg[5] * s};
        double kr = u[0] * wr[0] + u[1] * wr[1] + u[2] * wr[2];
        double ki = u[0] * wi[0] + u[1] * wi[1] + u[2] * wi[2];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double cr = u[c] * kr, ci = u[c] * ki;
            double pr = b.comp * cr + b.sol * (wr[c] - cr);
            double pi = b.comp * ci + b.sol * (wi[c] - ci);
            b.fre[c * n3 + e] = b.alpha * b.fre[c * n3 + e] + b.beta * pr;
            b.fim[c * n3 + e] = b.alpha * b.fim[c * n3 + e] + b.beta * pi;
        }
    }
}

// x pass: w1[c][ed][eb][i] = sum_ea F[c][ed][eb][ea] * exp(+2 pi i a i / g)
__global__ void ou_pass_x_kernel(const OUBox* boxes, long long step_all)
{
    const OUBox& b = boxes[blockIdx.y];
    unsigned long long step;
    if (!ou_active(b, step_all, step)) return;
    const int g = b.g, nb = b.nb, K = b.K;
    const size_t n = 3 * (size_t)nb * nb * g;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(t % g);
        const size_t line = t / g;   // (c, ed, eb)
        const double* xr = b.fre + line * nb;
        const double* xi = b.fim + line * nb;
        double sr = 0.0, si = 0.0;
        for (int ea = 0; ea < nb; ++ea) {
            const int tw = (band_mode(ea, nb, K, g) * i) % g;
            sr += xr[ea] * b.cs[tw] - xi[ea] * b.sn[tw];
            si += xr[ea] * b.sn[tw] + xi[ea] * b.cs[tw];
        }
        b.w1r[t] = sr;
        b.w1i[t] = si;
    }
}

// y pass: w2[c][ed][j][i] = sum_eb w1[c][ed][eb][i] * exp(+2 pi i b j / g)
__global__ void ou_pass_y_kernel(const OUBox* boxes, long long step_all)
{
    const OUBox& b = boxes[blockIdx.y];
    unsigned long long step;
    if (!ou_active(b, step_all, step)) return;
    const int g = b.g, nb = b.nb, K = b.K;
    const size_t gg = (size_t)g * g, n = 3 * (size_t)nb * gg;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(t % g), j = (int)((t / g) % g);
        const size_t ced = t / gg;   // c * nb + ed
        const size_t base = ced * nb * g + i;
        double sr = 0.0, si = 0.0;
        for (int eb = 0; eb < nb; ++eb) {
            const int tw = (band_mode(eb, nb, K, g) * j) % g;
            const double xr = b.w1r[base + (size_t)eb * g], xi = b.w1i[base + (size_t)eb * g];
            sr += xr * b.cs[tw] - xi * b.sn[tw];
            si += xr * b.sn[tw] + xi * b.cs[tw];
        }
        b.w2r[t] = sr;
        b.w2i[t] = si;
    }
}

// z pass, real part only: afield[c][k][j][i] = scale * Re sum_ed w2[c][ed][j][i] * exp(+2 pi i d k / g)
template <typename T>
__global__ void ou_pass_z_kernel(const OUBox* boxes, long long step_all)
{
    const OUBox& b = boxes[blockIdx.y];
    unsigned long long step;
    if (!ou_active(b, step_all, step)) return;
    const int g = b.g, nb = b.nb, K = b.K;
    const size_t gg = (size_t)g * g, g3 = gg * g, n = 3 * g3;
    T* af = (T*)b.afield;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
        const size_t ij = t % gg;
        const int k = (int)((t / gg) % g);
        const size_t c = t / g3;
        double sr = 0.0;
        for (int ed = 0; ed < nb; ++ed) {
            const int tw = (band_mode(ed, nb, K, g) * k) % g;
            const size_t at = (c * nb + ed) * gg + ij;
            sr += b.w2r[at] * b.cs[tw] - b.w2i[at] * b.sn[tw];
        }
        af[t] = (T)(sr * b.scale);
    }
}

}  // namespace glm
