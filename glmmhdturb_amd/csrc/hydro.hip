// hydro.hip -- launcher of the 5-variable hydro march (hydro_march.hpp).
#include "hydro_launch.hpp"
#include "hydro_march.hpp"

namespace glm {

static_assert(HYDRO_THREADS == HM_NT, "launch width of the hydro march");

template <typename T>
const void* hydro_march_ptr(int fallback)
{
    return fallback ? (const void*)&hydro_march_kernel<T, true> : (const void*)&hydro_march_kernel<T, false>;
}

template <typename T>
hipError_t launch_hydro_march(void* const* q, void* const* o, const HydroArgsHost& h, int slots, hipStream_t st)
{
    HydroArgs<T> a;
    for (int v = 0; v < 5; ++v) {
        a.q[v] = (const T*)q[v];
        a.o[v] = (T*)o[v];
    }
    a.nx = h.nx; a.ny = h.ny; a.nz = h.nz;
    hydro_tiling(h.nx, h.ny, h.nz, h.zchunk, slots, a.tiles_x, a.tiles_y, a.nzc, a.zchunk);
    a.gamma = (T)h.gamma; a.gm1 = (T)(h.gamma - 1.0); a.igm1 = (T)(1.0 / (h.gamma - 1.0));
    a.dtdx = (T)h.dtdx; a.half_dtdx = (T)(0.5 * h.dtdx);
    a.smallr = (T)h.smallr; a.smallp = (T)h.smallp;
    a.fb_dmin = (T)h.fb_dmin; a.fb_pmin = (T)h.fb_pmin;
    const long nblk = (long)a.tiles_x * a.tiles_y * a.nzc;
    if (nblk <= 0 || nblk > 0x7fffffffL) return hipErrorInvalidValue;
    dim3 g((unsigned)nblk), b(HM_NT);
    if (h.fallback) hipLaunchKernelGGL((hydro_march_kernel<T, true>), g, b, 0, st, a);
    else hipLaunchKernelGGL((hydro_march_kernel<T, false>), g, b, 0, st, a);
    return hipGetLastError();
}

template hipError_t launch_hydro_march<double>(void* const*, void* const*, const HydroArgsHost&, int, hipStream_t);
template hipError_t launch_hydro_march<float>(void* const*, void* const*, const HydroArgsHost&, int, hipStream_t);
template const void* hydro_march_ptr<double>(int);
template const void* hydro_march_ptr<float>(int);

}  // namespace glm
