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
hipError_t launch_march_pr_sp(const MarchArgs<T>& a, int recon, int slope, hipStream_t st);
template <typename T, int PRED, int RIEM>
const void* march_kernel_ptr_pr_sp(int recon, int slope);
template <typename T, int PRED, int RIEM>
hipError_t launch_ens_pr(const EnsLaunch<T>& e, int recon, int slope, int grid, hipStream_t st);
template <typename T, int PRED, int RIEM>
const void* ens_kernel_ptr_pr(int recon, int slope);
// The march's memory floor (march_floor_kernel; built in the full-predictor HLLD object).
template <typename T>
hipError_t launch_march_floor(const MarchArgs<T>& a, hipStream_t st);
template <typename T>
const void* march_floor_ptr();

// nsp = 0: the 9-plane march; nsp = 2: with two passive species (11 planes).
template <typename T>
inline hipError_t launch_march(const MarchArgs<T>& a, int pred, int riem, int recon, int slope, hipStream_t st,
                               int nsp = 0)
{
#define GLM_CASE(P, R) \
    if (pred == P && riem == R) \
        return nsp ? launch_march_pr_sp<T, P, R>(a, recon, slope, st) : launch_march_pr<T, P, R>(a, recon, slope, st);
    GLM_CASE(PRED_FULL, RIEM_LLF) GLM_CASE(PRED_FULL, RIEM_LLFB) GLM_CASE(PRED_FULL, RIEM_HLL) GLM_CASE(PRED_FULL, RIEM_HLLD)
    GLM_CASE(PRED_ONED, RIEM_LLF) GLM_CASE(PRED_ONED, RIEM_LLFB) GLM_CASE(PRED_ONED, RIEM_HLL) GLM_CASE(PRED_ONED, RIEM_HLLD)
#undef GLM_CASE
    return hipErrorInvalidValue;
}

// Kernel symbol of one march variant (for resource queries); nullptr if unknown.
template <typename T>
inline const void* march_kernel_ptr(int pred, int riem, int recon, int slope, int nsp = 0)
{
#define GLM_CASE(P, R) \
    if (pred == P && riem == R) \
        return nsp ? march_kernel_ptr_pr_sp<T, P, R>(recon, slope) : march_kernel_ptr_pr<T, P, R>(recon, slope);
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

// Workgroup tiling of the march for an (nx, ny, nz) grid: tiles and z-chunks.  With no requested
// z-chunk the number of chunks minimises the modelled run time on `slots` concurrent workgroups
// (CUs x resident workgroups per CU): rounds = ceil(workgroups / slots), each as long as one chunk
// plus about one plane of lead-in (the first plane's predictor and the chunk's extra z face).
// A grid whose workgroup count is a multiple of the slots leaves no partly filled last round.
inline void march_tiling(int nx, int ny, int nz, int zchunk_req, int slots, int& tiles_x, int& tiles_y, int& nzc,
                         int& zchunk)
{
    tiles_x = (nx + OX - 1) / OX;
    tiles_y = (ny + OY - 1) / OY;
    zchunk = zchunk_req > 0 ? zchunk_req : 0;
    if (zchunk == 0) {
        const long cols = (long)tiles_x * tiles_y;
        const long sl = slots > 0 ? slots : 256;
        const int maxc = nz / 8 > 1 ? nz / 8 : 1;   // chunks of at least 8 planes
        long best = -1;
        for (int c = 1; c <= maxc; ++c) {
            const int zc = (nz + c - 1) / c;
            const int n = (nz + zc - 1) / zc;      // chunks actually formed
            if (n != c) continue;
            const long cost = ((cols * n + sl - 1) / sl) * (long)(zc + 1);
            if (best < 0 || cost < best) { best = cost; zchunk = zc; }
        }
    }
    if (zchunk > nz) zchunk = nz;
    nzc = (nz + zchunk - 1) / zchunk;
}

}  // namespace glm
