// march_inst.hip -- one family of fused line-march kernels: precision GLM_T, predictor GLM_PRED,
// Riemann solver GLM_RIEM (set by the build), all reconstructions x limiters, for one grid
// (march_kernel) and for an ensemble of boxes (ens_march_kernel).
#include "launch.hpp"

#ifndef GLM_T
#error "build with -DGLM_T=double|float -DGLM_PRED=0|1 -DGLM_RIEM=0..3"
#endif

namespace glm {

template <>
hipError_t launch_march_pr<GLM_T, GLM_PRED, GLM_RIEM>(const MarchArgs<GLM_T>& a, int recon, int slope, hipStream_t st)
{
    dim3 grid((unsigned)(a.tiles_x * a.tiles_y * a.nzc)), block(NT);
#define GLM_CASE(RC, SL)                                                                             \
    if (recon == RC && slope == SL) {                                                                \
        hipLaunchKernelGGL((march_kernel<GLM_T, GLM_PRED, GLM_RIEM, RC, SL>), grid, block, 0, st, a); \
        return hipGetLastError();                                                                    \
    }
    GLM_CASE(RECON_PLM, SLOPE_MONCEN) GLM_CASE(RECON_PPM, SLOPE_MONCEN) GLM_CASE(RECON_PPM_SPIKE, SLOPE_MONCEN)
    GLM_CASE(RECON_PLM, SLOPE_MINMOD) GLM_CASE(RECON_PPM, SLOPE_MINMOD) GLM_CASE(RECON_PPM_SPIKE, SLOPE_MINMOD)
#undef GLM_CASE
    return hipErrorInvalidValue;
}

template <>
const void* march_kernel_ptr_pr<GLM_T, GLM_PRED, GLM_RIEM>(int recon, int slope)
{
#define GLM_CASE(RC, SL) \
    if (recon == RC && slope == SL) return (const void*)&march_kernel<GLM_T, GLM_PRED, GLM_RIEM, RC, SL>;
    GLM_CASE(RECON_PLM, SLOPE_MONCEN) GLM_CASE(RECON_PPM, SLOPE_MONCEN) GLM_CASE(RECON_PPM_SPIKE, SLOPE_MONCEN)
    GLM_CASE(RECON_PLM, SLOPE_MINMOD) GLM_CASE(RECON_PPM, SLOPE_MINMOD) GLM_CASE(RECON_PPM_SPIKE, SLOPE_MINMOD)
#undef GLM_CASE
    return nullptr;
}

// The same family with two passive species (11 planes).
template <>
hipError_t launch_march_pr_sp<GLM_T, GLM_PRED, GLM_RIEM>(const MarchArgs<GLM_T>& a, int recon, int slope, hipStream_t st)
{
    dim3 grid((unsigned)(a.tiles_x * a.tiles_y * a.nzc)), block(NT);
#define GLM_CASE(RC, SL)                                                                                \
    if (recon == RC && slope == SL) {                                                                   \
        hipLaunchKernelGGL((march_kernel<GLM_T, GLM_PRED, GLM_RIEM, RC, SL, 2>), grid, block, 0, st, a); \
        return hipGetLastError();                                                                       \
    }
    GLM_CASE(RECON_PLM, SLOPE_MONCEN) GLM_CASE(RECON_PPM, SLOPE_MONCEN) GLM_CASE(RECON_PPM_SPIKE, SLOPE_MONCEN)
    GLM_CASE(RECON_PLM, SLOPE_MINMOD) GLM_CASE(RECON_PPM, SLOPE_MINMOD) GLM_CASE(RECON_PPM_SPIKE, SLOPE_MINMOD)
#undef GLM_CASE
    return hipErrorInvalidValue;
}

template <>
const void* march_kernel_ptr_pr_sp<GLM_T, GLM_PRED, GLM_RIEM>(int recon, int slope)
{
#define GLM_CASE(RC, SL) \
    if (recon == RC && slope == SL) return (const void*)&march_kernel<GLM_T, GLM_PRED, GLM_RIEM, RC, SL, 2>;
    GLM_CASE(RECON_PLM, SLOPE_MONCEN) GLM_CASE(RECON_PPM, SLOPE_MONCEN) GLM_CASE(RECON_PPM_SPIKE, SLOPE_MONCEN)
    GLM_CASE(RECON_PLM, SLOPE_MINMOD) GLM_CASE(RECON_PPM, SLOPE_MINMOD) GLM_CASE(RECON_PPM_SPIKE, SLOPE_MINMOD)
#undef GLM_CASE
    return nullptr;
}

template <>
hipError_t launch_ens_pr<GLM_T, GLM_PRED, GLM_RIEM>(const EnsLaunch<GLM_T>& e, int recon, int slope, int grid,
                                                    hipStream_t st)
{
#define GLM_CASE(RC, SL)                                                                                 \
    if (recon == RC && slope == SL) {                                                                    \
        if (e.persistent)                                                                                \
            hipLaunchKernelGGL((ens_march_kernel<GLM_T, GLM_PRED, GLM_RIEM, RC, SL, true>), dim3((unsigned)grid), dim3(NT), 0, st, e); \
        else                                                                                             \
            hipLaunchKernelGGL((ens_march_kernel<GLM_T, GLM_PRED, GLM_RIEM, RC, SL, false>), dim3((unsigned)grid), dim3(NT), 0, st, e); \
        return hipGetLastError();                                                                        \
    }
    GLM_CASE(RECON_PLM, SLOPE_MONCEN) GLM_CASE(RECON_PPM, SLOPE_MONCEN) GLM_CASE(RECON_PPM_SPIKE, SLOPE_MONCEN)
    GLM_CASE(RECON_PLM, SLOPE_MINMOD) GLM_CASE(RECON_PPM, SLOPE_MINMOD) GLM_CASE(RECON_PPM_SPIKE, SLOPE_MINMOD)
#undef GLM_CASE
    return hipErrorInvalidValue;
}

template <>
const void* ens_kernel_ptr_pr<GLM_T, GLM_PRED, GLM_RIEM>(int recon, int slope)
{
#define GLM_CASE(RC, SL) \
    if (recon == RC && slope == SL) return (const void*)&ens_march_kernel<GLM_T, GLM_PRED, GLM_RIEM, RC, SL, true>;
    GLM_CASE(RECON_PLM, SLOPE_MONCEN) GLM_CASE(RECON_PPM, SLOPE_MONCEN) GLM_CASE(RECON_PPM_SPIKE, SLOPE_MONCEN)
    GLM_CASE(RECON_PLM, SLOPE_MINMOD) GLM_CASE(RECON_PPM, SLOPE_MINMOD) GLM_CASE(RECON_PPM_SPIKE, SLOPE_MINMOD)
#undef GLM_CASE
    return nullptr;
}

#if GLM_PRED == 0 && GLM_RIEM == 3   // the PRED_FULL, RIEM_HLLD object (the preprocessor sees no enums)
template <>
hipError_t launch_march_floor<GLM_T>(const MarchArgs<GLM_T>& a, hipStream_t st)
{
    hipLaunchKernelGGL((march_floor_kernel<GLM_T>), dim3((unsigned)(a.tiles_x * a.tiles_y * a.nzc)), dim3(NT), 0, st, a);
    return hipGetLastError();
}

template <>
const void* march_floor_ptr<GLM_T>()
{
    return (const void*)&march_floor_kernel<GLM_T>;
}
#endif

}  // namespace glm
