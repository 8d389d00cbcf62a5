// adj.hip -- launches of the reverse step (adj_impl.hpp).
#include "adj_impl.hpp"
#include "adj_march.hpp"
#include "launch.hpp"

namespace glm {

constexpr unsigned ADJ_GRID_MAX = 2048;

static unsigned adj_grid(size_t n)
{
    size_t g = (n + ADJ_NT - 1) / ADJ_NT;
    return (unsigned)(g < ADJ_GRID_MAX ? (g > 0 ? g : 1) : ADJ_GRID_MAX);
}

size_t adj_scratch_bytes(size_t n) { return (size_t)ADJ_NPLANES * n * sizeof(double) + ADJ_GRID_MAX * sizeof(double); }

template <int PRED, int RECON, int SLOPE>
static void launch_cellwise(const AdjArgs& a, dim3 g, hipStream_t st, bool faces)
{
    if (faces) {
        hipLaunchKernelGGL((adj_faces_kernel<PRED, RECON, SLOPE>), g, dim3(ADJ_NT), 0, st, a);
        return;
    }
    hipLaunchKernelGGL((adj_cell_dir_kernel<PRED, RECON, SLOPE, 0>), g, dim3(ADJ_NT), 0, st, a);
    hipLaunchKernelGGL((adj_cell_dir_kernel<PRED, RECON, SLOPE, 1>), g, dim3(ADJ_NT), 0, st, a);
    hipLaunchKernelGGL((adj_cell_dir_kernel<PRED, RECON, SLOPE, 2>), g, dim3(ADJ_NT), 0, st, a);
    if (PRED == PRED_FULL) hipLaunchKernelGGL((adj_hancock_kernel<SLOPE>), g, dim3(ADJ_NT), 0, st, a);
}

static bool dispatch_cellwise(const AdjArgs& a, dim3 g, hipStream_t st, bool faces, int pred, int recon, int slope)
{
#define GLM_CASE(P, RC, SL)                                          \
    if (pred == P && recon == RC && slope == SL) {                   \
        launch_cellwise<P, RC, SL>(a, g, st, faces);                 \
        return true;                                                 \
    }
    GLM_CASE(PRED_FULL, RECON_PLM, SLOPE_MONCEN) GLM_CASE(PRED_FULL, RECON_PLM, SLOPE_MINMOD)
    GLM_CASE(PRED_FULL, RECON_PPM, SLOPE_MONCEN) GLM_CASE(PRED_FULL, RECON_PPM, SLOPE_MINMOD)
    GLM_CASE(PRED_ONED, RECON_PLM, SLOPE_MONCEN) GLM_CASE(PRED_ONED, RECON_PLM, SLOPE_MINMOD)
    GLM_CASE(PRED_ONED, RECON_PPM, SLOPE_MONCEN) GLM_CASE(PRED_ONED, RECON_PPM, SLOPE_MINMOD)
#undef GLM_CASE
    return false;
}

static bool dispatch_faceswise(const AdjArgs& a, dim3 g, hipStream_t st, bool upre, int riem)
{
#define GLM_CASE(R)                                                                                       \
    if (riem == R) {                                                                                      \
        if (upre) {                                                                                       \
            hipLaunchKernelGGL((adj_flux_kernel<R>), g, dim3(ADJ_NT), 0, st, a);                          \
            hipLaunchKernelGGL(adj_upre_kernel, g, dim3(ADJ_NT), 0, st, a);                               \
        } else {                                                                                          \
            hipLaunchKernelGGL((adj_riemann_kernel<R>), g, dim3(ADJ_NT), 0, st, a);                       \
        }                                                                                                 \
        return true;                                                                                      \
    }
    GLM_CASE(RIEM_LLF) GLM_CASE(RIEM_LLFB) GLM_CASE(RIEM_HLL) GLM_CASE(RIEM_HLLD)
#undef GLM_CASE
    return false;
}

static bool dispatch_back(const AdjArgs& a, int zchunk, hipStream_t st, int pred, int recon, int slope)
{
    int tx, ty, nzc, zc;
    march_tiling(a.nx, a.ny, a.nz, zchunk, 256, tx, ty, nzc, zc);   // one workgroup per CU
    const dim3 g((unsigned)(tx * ty * nzc));
#define GLM_CASE(P, RC, SL)                                                                          \
    if (pred == P && recon == RC && slope == SL) {                                                   \
        hipLaunchKernelGGL((adj_back_kernel<P, RC, SL>), g, dim3(NT), 0, st, a, tx, ty, zc);         \
        return true;                                                                                 \
    }
    GLM_CASE(PRED_FULL, RECON_PLM, SLOPE_MONCEN) GLM_CASE(PRED_FULL, RECON_PLM, SLOPE_MINMOD)
    GLM_CASE(PRED_FULL, RECON_PPM, SLOPE_MONCEN) GLM_CASE(PRED_FULL, RECON_PPM, SLOPE_MINMOD)
    GLM_CASE(PRED_ONED, RECON_PLM, SLOPE_MONCEN) GLM_CASE(PRED_ONED, RECON_PLM, SLOPE_MINMOD)
    GLM_CASE(PRED_ONED, RECON_PPM, SLOPE_MONCEN) GLM_CASE(PRED_ONED, RECON_PPM, SLOPE_MINMOD)
#undef GLM_CASE
    return false;
}

hipError_t launch_step_adjoint(AdjArgs a, void* scratch, int pred, int riem, int recon, int slope, double* fdot,
                               bool fused, int zchunk, hipStream_t st)
{
    if (recon == RECON_PPM_SPIKE) return hipErrorInvalidValue;
    const size_t n = (size_t)a.nx * a.ny * a.nz;
    a.scr = (double*)scratch;
    a.part = a.scr + (size_t)ADJ_NPLANES * n;
    const dim3 g(adj_grid(n));
    hipLaunchKernelGGL(adj_prim_kernel, g, dim3(ADJ_NT), 0, st, a);
    if (!dispatch_cellwise(a, g, st, true, pred, recon, slope)) return hipErrorInvalidValue;
    if (!dispatch_faceswise(a, g, st, true, riem)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(adj_fdot_kernel, dim3(1), dim3(ADJ_NT), 0, st, (const double*)a.part, (int)g.x, fdot);
    dispatch_faceswise(a, g, st, false, riem);
    if (fused) {
        dispatch_back(a, zchunk, st, pred, recon, slope);
    } else {
        dispatch_cellwise(a, g, st, false, pred, recon, slope);
        hipLaunchKernelGGL(adj_gather_kernel, g, dim3(ADJ_NT), 0, st, a);
    }
    return hipGetLastError();
}

}  // namespace glm
