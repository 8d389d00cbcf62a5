// hydro_launch.hpp -- host launcher of the 5-variable hydro march (hydro.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace glm {

struct HydroArgsHost {
    int nx, ny, nz;
    double gamma, dtdx, smallr, smallp;
    int fallback;              // HLL with the LLF fallback below fb_dmin / fb_pmin
    double fb_dmin, fb_pmin;
    int zchunk;                // z planes per workgroup (0: the occupancy model)
    int nsp;                   // passive scalars: 0 or 2
    int cma;                   // scalar flux rides the mass flux
    int fmt;                   // storage format (hydro_march.hpp HY_*): 0 conserved planes of T, 1 f16
                               // primitive, 2 uint16 primitive, 3 uint16 log2 species (FP32 only)
    int scheme;                // 0 light march (-DHANCOCK1D), 1 parabolic edges (-DPPM1D), 2 donor-cell
                               // tile seams (-DDONOR); 1 and 2 take no fallback and conserved storage
};

constexpr int HYDRO_THREADS = 256;   // lanes of a hydro march workgroup (hydro_march.hpp HM_NT)
template <typename T>
hipError_t launch_hydro_march(void* const* q, void* const* o, const HydroArgsHost& h, int slots, hipStream_t st);
template <typename T>
const void* hydro_march_ptr(int fallback, int nsp, int cma, int fmt = 0, int scheme = 0);

}  // namespace glm
