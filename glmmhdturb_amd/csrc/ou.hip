// ou.hip -- host side of the band-limited OU forcing (ou_impl.hpp): per-box buffers and the
// batched advance.
#include <cmath>
#include <cstring>
#include <vector>

#include "ou_impl.hpp"

namespace glm {

static size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

int ou_band_width(int g, double kmax)
{
    int K = kmax >= 0.0 ? (int)std::floor(kmax) : 0;
    int nb = 2 * K + 1;
    return nb < g ? nb : g;
}

hipError_t ou_box_alloc(OUBox& b, int g, double kmin, double kmax, double T, int ndt, double comp, double amp,
                        unsigned long long seed, int prec, hipStream_t st)
{
    memset(&b, 0, sizeof b);
    b.g = g;
    b.nb = ou_band_width(g, kmax);
    b.K = kmax >= 0.0 ? (int)std::floor(kmax) : 0;
    b.kmin = kmin; b.kmax = kmax;
    const double dt = T / ndt;                 // glmmhd_turb.jl:2148-2150
    b.alpha = 1.0 - dt / T;
    b.beta = std::sqrt(2.0 * dt / T);
    b.comp = comp; b.sol = 1.0 - comp;
    b.scale = amp / ((double)g * (double)g * (double)g);
    b.seed = seed;
    const size_t nb = b.nb, gg = (size_t)g * g;
    const size_t n_f = 3 * nb * nb * nb, n_w1 = 3 * nb * nb * g, n_w2 = 3 * nb * gg, n_a = 3 * gg * g;
    const size_t sz_f = al256(n_f * 8), sz_w1 = al256(n_w1 * 8), sz_w2 = al256(n_w2 * 8), sz_t = al256((size_t)g * 8);
    const size_t total = 2 * sz_f + 2 * sz_w1 + 2 * sz_w2 + 2 * sz_t + al256(n_a * prec);
    char* base = nullptr;
    hipError_t e = hipMalloc(&base, total);
    if (e != hipSuccess) return e;
    char* p = base;
    b.fre = (double*)p; p += sz_f;
    b.fim = (double*)p; p += sz_f;
    b.w1r = (double*)p; p += sz_w1;
    b.w1i = (double*)p; p += sz_w1;
    b.w2r = (double*)p; p += sz_w2;
    b.w2i = (double*)p; p += sz_w2;
    double* cs = (double*)p; p += sz_t;
    double* sn = (double*)p; p += sz_t;
    b.cs = cs; b.sn = sn;
    b.afield = p;
    std::vector<double> hc(g), hs(g);
    for (int m = 0; m < g; ++m) {
        hc[m] = std::cos(2.0 * M_PI * m / g);
        hs[m] = std::sin(2.0 * M_PI * m / g);
    }
    if ((e = hipMemcpy(cs, hc.data(), g * 8, hipMemcpyHostToDevice)) != hipSuccess) return e;
    if ((e = hipMemcpy(sn, hs.data(), g * 8, hipMemcpyHostToDevice)) != hipSuccess) return e;
    if ((e = hipMemsetAsync(base, 0, 2 * sz_f, st)) != hipSuccess) return e;
    return hipMemsetAsync(b.afield, 0, n_a * prec, st);
}

void ou_box_free(OUBox& b)
{
    if (b.fre) (void)hipFree(b.fre);
    memset(&b, 0, sizeof b);
}

hipError_t ou_box_reset(const OUBox& b, int prec, hipStream_t st)
{
    const size_t nb = b.nb, gg = (size_t)b.g * b.g;
    hipError_t e = hipMemsetAsync(b.fre, 0, 3 * nb * nb * nb * 8, st);
    if (e == hipSuccess) e = hipMemsetAsync(b.fim, 0, 3 * nb * nb * nb * 8, st);
    if (e == hipSuccess) e = hipMemsetAsync(b.afield, 0, 3 * gg * b.g * prec, st);
    return e;
}

static unsigned blocks_for(size_t work, int nbox)
{
    size_t cap = (size_t)(2048 / (nbox > 0 ? nbox : 1));
    if (cap < 16) cap = 16;
    size_t bl = (work + 255) / 256;
    if (bl > cap) bl = cap;
    return (unsigned)(bl > 0 ? bl : 1);
}

hipError_t launch_ou_advance(int prec, const OUBox* d_boxes, int nbox, int gmax, int nbmax, long long step_all,
                             hipStream_t st)
{
    const size_t g = gmax, nb = nbmax;
    dim3 bl(256);
    hipLaunchKernelGGL(ou_band_step_kernel, dim3(blocks_for(nb * nb * nb, nbox), nbox), bl, 0, st, d_boxes, step_all);
    hipLaunchKernelGGL(ou_pass_x_kernel, dim3(blocks_for(3 * nb * nb * g, nbox), nbox), bl, 0, st, d_boxes, step_all);
    hipLaunchKernelGGL(ou_pass_y_kernel, dim3(blocks_for(3 * nb * g * g, nbox), nbox), bl, 0, st, d_boxes, step_all);
    dim3 gz(blocks_for(3 * g * g * g, nbox), nbox);
    if (prec == 8) hipLaunchKernelGGL((ou_pass_z_kernel<double>), gz, bl, 0, st, d_boxes, step_all);
    else hipLaunchKernelGGL((ou_pass_z_kernel<float>), gz, bl, 0, st, d_boxes, step_all);
    return hipGetLastError();
}

}  // namespace glm
