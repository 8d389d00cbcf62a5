// hydro.hip -- launcher of the 5-variable hydro march (hydro_march.hpp).
#include "hydro_launch.hpp"
#include "hydro_march.hpp"

namespace glm {

static_assert(HYDRO_THREADS == HM_NT, "launch width of the hydro march");

template <typename T, bool FB>
static const void* ptr_fb(int nsp, int cma)
{
    if (nsp == 0) return (const void*)&hydro_march_kernel<T, FB, 0, false>;
    return cma ? (const void*)&hydro_march_kernel<T, FB, 2, true> : (const void*)&hydro_march_kernel<T, FB, 2, false>;
}

template <typename T>
const void* hydro_march_ptr(int fallback, int nsp, int cma)
{
    if (nsp != 0 && nsp != 2) return nullptr;
    return fallback ? ptr_fb<T, true>(nsp, cma) : ptr_fb<T, false>(nsp, cma);
}

template <typename T, bool FB, int NSP, bool CMA>
static void launch(const HydroArgs<T>& a, dim3 g, hipStream_t st)
{
    hipLaunchKernelGGL((hydro_march_kernel<T, FB, NSP, CMA>), g, dim3(HM_NT), 0, st, a);
}

template <typename T>
hipError_t launch_hydro_march(void* const* q, void* const* o, const HydroArgsHost& h, int slots, hipStream_t st)
{
    if (h.nsp != 0 && h.nsp != 2) return hipErrorInvalidValue;
    const int nv = 5 + h.nsp;
    HydroArgs<T> a;
    for (int v = 0; v < HM_NVMAX; ++v) {
        a.q[v] = v < nv ? (const T*)q[v] : nullptr;
        a.o[v] = v < nv ? (T*)o[v] : nullptr;
    }
    a.nx = h.nx; a.ny = h.ny; a.nz = h.nz;
    hydro_tiling(h.nx, h.ny, h.nz, h.zchunk, slots, a.tiles_x, a.tiles_y, a.nzc, a.zchunk);
    a.gamma = (T)h.gamma; a.gm1 = (T)(h.gamma - 1.0); a.igm1 = (T)(1.0 / (h.gamma - 1.0));
    a.dtdx = (T)h.dtdx; a.half_dtdx = (T)(0.5 * h.dtdx);
    a.smallr = (T)h.smallr; a.smallp = (T)h.smallp;
    a.fb_dmin = (T)h.fb_dmin; a.fb_pmin = (T)h.fb_pmin;
    const long nblk = (long)a.tiles_x * a.tiles_y * a.nzc;
    if (nblk <= 0 || nblk > 0x7fffffffL) return hipErrorInvalidValue;
    const dim3 g((unsigned)nblk);
    if (h.nsp == 0) {
        if (h.fallback) launch<T, true, 0, false>(a, g, st);
        else launch<T, false, 0, false>(a, g, st);
    } else if (h.cma) {
        if (h.fallback) launch<T, true, 2, true>(a, g, st);
        else launch<T, false, 2, true>(a, g, st);
    } else {
        if (h.fallback) launch<T, true, 2, false>(a, g, st);
        else launch<T, false, 2, false>(a, g, st);
    }
    return hipGetLastError();
}

template hipError_t launch_hydro_march<double>(void* const*, void* const*, const HydroArgsHost&, int, hipStream_t);
template hipError_t launch_hydro_march<float>(void* const*, void* const*, const HydroArgsHost&, int, hipStream_t);
template const void* hydro_march_ptr<double>(int, int, int);
template const void* hydro_march_ptr<float>(int, int, int);

}  // namespace glm
