// aux.hip -- instantiations of the reduction / initial-condition / OU kernels (aux_impl.hpp).
#include "aux_impl.hpp"
#include "aux_launch.hpp"

namespace glm {

static unsigned grid_for(size_t n, unsigned nt = AUX_NT, unsigned cap = 8192)
{
    size_t g = (n + nt - 1) / nt;
    return (unsigned)(g < cap ? (g > 0 ? g : 1) : cap);
}

hipError_t launch_clear_u64(unsigned long long* p, int n, hipStream_t st)
{
    hipLaunchKernelGGL(clear_u64_kernel, dim3((n + 63) / 64), dim3(64), 0, st, p, n);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_ctot(Ptr9<const T> q9dev, size_t n, const Phys<T>& s, unsigned long long* out, hipStream_t st)
{
    hipLaunchKernelGGL((ctot_kernel<T>), dim3(grid_for(n, AUX_NT, 4096)), dim3(AUX_NT), 0, st, q9dev, n, s, out);
    return hipGetLastError();
}

size_t diag_nparts(size_t n) { return (n + DIAG_CELLS_PER_BLOCK - 1) / DIAG_CELLS_PER_BLOCK; }

template <typename T>
hipError_t launch_diag(Ptr9<const T> q9dev, int nx, int ny, int nz, double dx, double smallr, void* parts,
                       double* out, hipStream_t st)
{
    size_t n = (size_t)nx * ny * nz, np = diag_nparts(n);
    hipLaunchKernelGGL((diag_partial_kernel<T>), dim3((unsigned)np), dim3(AUX_NT), 0, st, q9dev, nx, ny, nz, dx, smallr,
                       (size_t)DIAG_CELLS_PER_BLOCK, (DiagPart*)parts);
    hipLaunchKernelGGL(diag_final_kernel, dim3(1), dim3(1), 0, st, (const DiagPart*)parts, (int)np, out);
    return hipGetLastError();
}

size_t diag_part_bytes() { return sizeof(DiagPart); }

template <typename T>
hipError_t launch_ic(Ptr9<T> q9dev, int nx, int ny, int nz, int kind, const double* args, hipStream_t st)
{
    size_t n = (size_t)nx * ny * nz;
    dim3 g(grid_for(n)), b(AUX_NT);
    if (kind == 0) hipLaunchKernelGGL((ic_ot_kernel<T>), g, b, 0, st, q9dev, nx, ny, nz, (T)args[0]);
    else if (kind == 1)
        hipLaunchKernelGGL((ic_uniform_kernel<T>), g, b, 0, st, q9dev, n, (T)args[0], (T)args[1], (T)args[2], (T)args[3],
                           (T)args[4], (T)args[5]);
    else if (kind == 2) hipLaunchKernelGGL((ic_gradient_kernel<T>), g, b, 0, st, q9dev, n, (T)args[0]);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_ou_init(int g, double kmin, double kmax, double* sqp, double* unitk, hipStream_t st)
{
    size_t g3 = (size_t)g * g * g;
    hipLaunchKernelGGL(ou_init_kernel, dim3((unsigned)((g3 + 255) / 256)), dim3(256), 0, st, g, kmin, kmax, sqp, unitk);
    return hipGetLastError();
}

hipError_t launch_ou_step(int g, double* fre, double* fim, const double* sqp, const double* unitk, double alpha,
                          double beta, double comp, double sol, unsigned long long seed, unsigned long long step,
                          hipStream_t st)
{
    size_t g3 = (size_t)g * g * g;
    hipLaunchKernelGGL(ou_step_kernel, dim3((unsigned)((g3 + 255) / 256)), dim3(256), 0, st, g, fre, fim, sqp, unitk,
                       alpha, beta, comp, sol, seed, step);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_ou_afield(int g, const double* fre, const double* fim, double* work, const double* cs,
                            const double* sn, double amp, T* afield, hipStream_t st)
{
    size_t g3 = (size_t)g * g * g, n3 = 3 * g3;
    double *ar = work, *ai = work + n3, *br = work + 2 * n3, *bi = work + 3 * n3;
    dim3 gr((unsigned)((n3 + 255) / 256)), bl(256);
    hipLaunchKernelGGL(idft_pass_kernel, gr, bl, 0, st, g, 0, fre, fim, ar, ai, cs, sn);
    hipLaunchKernelGGL(idft_pass_kernel, gr, bl, 0, st, g, 1, ar, ai, br, bi, cs, sn);
    hipLaunchKernelGGL(idft_pass_kernel, gr, bl, 0, st, g, 2, br, bi, ar, ai, cs, sn);
    double scale = amp / ((double)g * (double)g * (double)g);
    hipLaunchKernelGGL((afield_kernel<T>), gr, bl, 0, st, g, ar, scale, afield);
    return hipGetLastError();
}

#define GLM_INST(T)                                                                                                   \
    template hipError_t launch_ctot<T>(Ptr9<const T>, size_t, const Phys<T>&, unsigned long long*, hipStream_t);     \
    template hipError_t launch_diag<T>(Ptr9<const T>, int, int, int, double, double, void*, double*, hipStream_t);  \
    template hipError_t launch_ic<T>(Ptr9<T>, int, int, int, int, const double*, hipStream_t);                      \
    template hipError_t launch_ou_afield<T>(int, const double*, const double*, double*, const double*, const double*, \
                                            double, T*, hipStream_t);
GLM_INST(double)
GLM_INST(float)
#undef GLM_INST

}  // namespace glm
