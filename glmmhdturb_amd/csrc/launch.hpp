// launch.hpp -- host-side launchers of the gfx950 kernels (defined in march_f64.hip,
// march_f32.hip, aux.hip, ct.hip) used by the C-ABI in capi.hip.
#pragma once
#include <hip/hip_runtime.h>
#include "march_impl.hpp"

namespace glm {

template <typename T>
hipError_t launch_march(const MarchArgs<T>& a, int pred, int riem, hipStream_t st);
// Kernel symbol of one march variant (for resource queries); nullptr if unknown.
template <typename T>
const void* march_kernel_ptr(int pred, int riem);

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
