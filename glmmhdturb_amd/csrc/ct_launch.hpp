// ct_launch.hpp -- host launchers of the constrained-transport kernels (ct.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace glm {

struct CTArgsHost {
    int nx, ny, nz;
    double gamma, smallr, smallp, dt, dx, smallc, turb_min_rho, ramp;
    int face_slope, riemann, recon, update_floor, do_turb, tgs, want_cfl;
    int zchunk;                    // z planes per march tile (0: the occupancy model)
    int nsp;                       // passive species planes after the 8 CT planes (0 or 2; fused march only)
    const void* afield;
    unsigned long long* cmax;
};

size_t ct_scratch_bytes(size_t n, int prec);
constexpr int CT_MARCH_THREADS = 252;   // lanes of a CT march workgroup (ct_march.hpp CM_NT)
template <typename T>
hipError_t launch_ct_step(void* const* q, void* const* o, const CTArgsHost& h, void* scratch, hipStream_t st);
template <typename T>
hipError_t launch_ct_ctot(void* const* q, int nx, int ny, int nz, double gamma, double smallr, double smallp,
                          unsigned long long* out, hipStream_t st);
template <typename T>
hipError_t launch_ct_divb(void* const* q, int nx, int ny, int nz, double dx, unsigned long long* out, hipStream_t st);
template <typename T>
const void* ct_kernel_ptr();
// Fused CT step (one z-march kernel) and its kernel for a Riemann solver (occupancy queries).
template <typename T>
hipError_t launch_ct_march(void* const* q, void* const* o, const CTArgsHost& h, int slots, hipStream_t st);
template <typename T>
const void* ct_march_ptr(int riemann, int nsp = 0, int recon = 0);
// The same step as one kernel over staged 3-D tiles (ct_tile.hpp) and that kernel (resource queries).
template <typename T>
hipError_t launch_ct_tiles(void* const* q, void* const* o, const CTArgsHost& h, hipStream_t st);
template <typename T>
const void* ct_tile_ptr();

}  // namespace glm
