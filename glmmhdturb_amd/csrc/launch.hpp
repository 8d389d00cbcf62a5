// launch.hpp -- host-side launchers of the gfx950 kernels (defined in march_f64.hip,
// march_f32.hip, aux.hip, ct.hip) used by the C-ABI in capi.hip.
#pragma once
#include <hip/hip_runtime.h>
#include "march_impl.hpp"

namespace glm {

// One (precision, predictor, Riemann) family of march kernels; the 3 reconstructions x 2 limiters
// of a family are instantiated in one object built from march_inst.hip.
template <typename T, int PRED, int RIEM>
hipError_t launch_march_pr(const MarchArgs<T>& a, int recon, int slope, hipStream_t st);
template <typename T, int PRED, int RIEM>
const void* march_kernel_ptr_pr(int recon, int slope);
template <typename T, int PRED, int RIEM>
hipError_t launch_ens_pr(const EnsLaunch<T>& e, int recon, int slope, int grid, hipStream_t st);
template <typename T, int PRED, int RIEM>
const void* ens_kernel_ptr_pr(int recon, int slope);

template <typename T>
inline hipError_t launch_march(const MarchArgs<T>& a, int pred, int riem, int recon, int slope, hipStream_t st)
{
#define GLM_CASE(P, R) \
    if (pred == P && riem == R) return launch_march_pr<T, P, R>(a, recon, slope, st);
    GLM_CASE(PRED_FULL, RIEM_LLF) GLM_CASE(PRED_FULL, RIEM_LLFB) GLM_CASE(PRED_FULL, RIEM_HLL) GLM_CASE(PRED_FULL, RIEM_HLLD)
    GLM_CASE(PRED_ONED, RIEM_LLF) GLM_CASE(PRED_ONED, RIEM_LLFB) GLM_CASE(PRED_ONED, RIEM_HLL) GLM_CASE(PRED_ONED, RIEM_HLLD)
#undef GLM_CASE
    return hipErrorInvalidValue;
}

// Kernel symbol of one march variant (for resource queries); nullptr if unknown.
template <typename T>
inline const void* march_kernel_ptr(int pred, int riem, int recon, int slope)
{
#define GLM_CASE(P, R) \
    if (pred == P && riem == R) return march_kernel_ptr_pr<T, P, R>(recon, slope);
    GLM_CASE(PRED_FULL, RIEM_LLF) GLM_CASE(PRED_FULL, RIEM_LLFB) GLM_CASE(PRED_FULL, RIEM_HLL) GLM_CASE(PRED_FULL, RIEM_HLLD)
    GLM_CASE(PRED_ONED, RIEM_LLF) GLM_CASE(PRED_ONED, RIEM_LLFB) GLM_CASE(PRED_ONED, RIEM_HLL) GLM_CASE(PRED_ONED, RIEM_HLLD)
#undef GLM_CASE
    return nullptr;
}

template <typename T>
inline hipError_t launch_ens(const EnsLaunch<T>& e, int pred, int riem, int recon, int slope, int grid, hipStream_t st)
{
#define GLM_CASE(P, R) \
    if (pred == P && riem == R) return launch_ens_pr<T, P, R>(e, recon, slope, grid, st);
    GLM_CASE(PRED_FULL, RIEM_LLF) GLM_CASE(PRED_FULL, RIEM_LLFB) GLM_CASE(PRED_FULL, RIEM_HLL) GLM_CASE(PRED_FULL, RIEM_HLLD)
    GLM_CASE(PRED_ONED, RIEM_LLF) GLM_CASE(PRED_ONED, RIEM_LLFB) GLM_CASE(PRED_ONED, RIEM_HLL) GLM_CASE(PRED_ONED, RIEM_HLLD)
#undef GLM_CASE
    return hipErrorInvalidValue;
}

template <typename T>
inline const void* ens_kernel_ptr(int pred, int riem, int recon, int slope)
{
#define GLM_CASE(P, R) \
    if (pred == P && riem == R) return ens_kernel_ptr_pr<T, P, R>(recon, slope);
    GLM_CASE(PRED_FULL, RIEM_LLF) GLM_CASE(PRED_FULL, RIEM_LLFB) GLM_CASE(PRED_FULL, RIEM_HLL) GLM_CASE(PRED_FULL, RIEM_HLLD)
    GLM_CASE(PRED_ONED, RIEM_LLF) GLM_CASE(PRED_ONED, RIEM_LLFB) GLM_CASE(PRED_ONED, RIEM_HLL) GLM_CASE(PRED_ONED, RIEM_HLLD)
#undef GLM_CASE
    return nullptr;
}

// Workgroup tiling of the march for an (nx, ny, nz) grid: tiles and z-chunks.
inline void march_tiling(int nx, int ny, int nz, int zchunk_req, int& tiles_x, int& tiles_y, int& nzc, int& zchunk)
{
    tiles_x = (nx + OX - 1) / OX;
    tiles_y = (ny + OY - 1) / OY;
    zchunk = zchunk_req > 0 ? zchunk_req : 0;
    if (zchunk == 0) {
        // enough workgroups to fill 256 CUs a few times over, z-chunks no shorter than 16 planes
        long cols = (long)tiles_x * tiles_y;
        int want = (int)((2048 + cols - 1) / cols);
        int maxc = nz / 16 > 1 ? nz / 16 : 1;
        int c = want < maxc ? want : maxc;
        if (c < 1) c = 1;
        zchunk = (nz + c - 1) / c;
    }
    if (zchunk > nz) zchunk = nz;
    nzc = (nz + zchunk - 1) / zchunk;
}

}  // namespace glm
