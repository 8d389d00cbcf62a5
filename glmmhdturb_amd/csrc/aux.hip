// aux.hip -- instantiations of the reduction and initial-condition kernels (aux_impl.hpp).
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
hipError_t launch_ctot(Ptr9<const T> q9dev, size_t n, const Phys<T>& s, unsigned long long* out, hipStream_t st,
                       bool fast)
{
    if (fast) hipLaunchKernelGGL((ctot_kernel<T, true>), dim3(grid_for(n, AUX_NT, 4096)), dim3(AUX_NT), 0, st, q9dev, n, s, out);
    else hipLaunchKernelGGL((ctot_kernel<T, false>), dim3(grid_for(n, AUX_NT, 4096)), dim3(AUX_NT), 0, st, q9dev, n, s, out);
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

#define GLM_INST(T)                                                                                                   \
    template hipError_t launch_ctot<T>(Ptr9<const T>, size_t, const Phys<T>&, unsigned long long*, hipStream_t, bool); \
    template hipError_t launch_diag<T>(Ptr9<const T>, int, int, int, double, double, void*, double*, hipStream_t);  \
    template hipError_t launch_ic<T>(Ptr9<T>, int, int, int, int, const double*, hipStream_t);
GLM_INST(double)
GLM_INST(float)
#undef GLM_INST

}  // namespace glm
