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

// The compact storage formats (FP32 only): f16 and uint16 primitive without scalars, uint16 log2
// species with CMA.
template <typename T, bool FB>
static const void* ptr_fmt(int fmt)
{
    if constexpr (sizeof(T) == 4) {
        if (fmt == HY_F16P) return (const void*)&hydro_march_kernel<T, FB, 0, false, HY_F16P>;
        if (fmt == HY_U16P) return (const void*)&hydro_march_kernel<T, FB, 0, false, HY_U16P>;
        if (fmt == HY_U16SP) return (const void*)&hydro_march_kernel<T, FB, 2, true, HY_U16SP>;
    }
    return nullptr;
}

// The scheme variants (-DPPM1D, -DDONOR): conserved storage, HLL without the fallback.
template <typename T, int SV>
static const void* ptr_sv(int nsp, int cma)
{
    if (nsp == 0) return (const void*)&hydro_march_kernel<T, false, 0, false, HY_CONS, SV>;
    return cma ? (const void*)&hydro_march_kernel<T, false, 2, true, HY_CONS, SV>
               : (const void*)&hydro_march_kernel<T, false, 2, false, HY_CONS, SV>;
}

template <typename T>
const void* hydro_march_ptr(int fallback, int nsp, int cma, int fmt, int scheme)
{
    if (scheme == HY_PPM1D || scheme == HY_DONOR) {
        if (fmt != HY_CONS || (nsp != 0 && nsp != 2)) return nullptr;
        return scheme == HY_PPM1D ? ptr_sv<T, HY_PPM1D>(nsp, cma) : ptr_sv<T, HY_DONOR>(nsp, cma);
    }
    if (scheme != HY_PLM) return nullptr;
    if (fmt != HY_CONS) return fallback ? ptr_fmt<T, true>(fmt) : ptr_fmt<T, false>(fmt);
    if (nsp != 0 && nsp != 2) return nullptr;
    return fallback ? ptr_fb<T, true>(nsp, cma) : ptr_fb<T, false>(nsp, cma);
}

template <typename T, bool FB, int NSP, bool CMA, int FMT = HY_CONS, int SV = HY_PLM>
static void launch(const HydroArgs<T>& a, dim3 g, hipStream_t st)
{
    hipLaunchKernelGGL((hydro_march_kernel<T, FB, NSP, CMA, FMT, SV>), g, dim3(HM_NT), 0, st, a);
}

template <typename T, int SV>
static void launch_sv(const HydroArgs<T>& a, dim3 g, int nsp, int cma, hipStream_t st)
{
    if (nsp == 0) launch<T, false, 0, false, HY_CONS, SV>(a, g, st);
    else if (cma) launch<T, false, 2, true, HY_CONS, SV>(a, g, st);
    else launch<T, false, 2, false, HY_CONS, SV>(a, g, st);
}

template <typename T>
hipError_t launch_hydro_march(void* const* q, void* const* o, const HydroArgsHost& h, int slots, hipStream_t st)
{
    if (h.nsp != 0 && h.nsp != 2) return hipErrorInvalidValue;
    if (h.scheme < HY_PLM || h.scheme > HY_DONOR || (h.scheme != HY_PLM && h.fmt != HY_CONS)) return hipErrorInvalidValue;
    if (h.fmt != HY_CONS && (sizeof(T) != 4 || (h.fmt == HY_U16SP ? (h.nsp != 2 || !h.cma) : h.nsp != 0) ||
                             h.fmt < 0 || h.fmt > HY_U16SP))
        return hipErrorInvalidValue;
    const int nv = 5 + h.nsp;
    HydroArgs<T> a;
    for (int v = 0; v < HM_NVMAX; ++v) {
        // planes of T: every plane of the conserved format, the hydro planes of HY_U16SP;
        // 16-bit planes: every plane of the primitive formats, the species of HY_U16SP
        const bool wide = v < nv && (h.fmt == HY_CONS || (h.fmt == HY_U16SP && v < 5));
        const bool narrow = v < nv && !wide;
        a.q[v] = wide ? (const T*)q[v] : nullptr;
        a.o[v] = wide ? (T*)o[v] : nullptr;
        a.qh[v] = narrow ? (const unsigned short*)q[v] : nullptr;
        a.oh[v] = narrow ? (unsigned short*)o[v] : nullptr;
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
    if (h.scheme == HY_PPM1D) {
        launch_sv<T, HY_PPM1D>(a, g, h.nsp, h.cma, st);
        return hipGetLastError();
    }
    if (h.scheme == HY_DONOR) {
        launch_sv<T, HY_DONOR>(a, g, h.nsp, h.cma, st);
        return hipGetLastError();
    }
    if constexpr (sizeof(T) == 4) {
        if (h.fmt != HY_CONS) {
            if (h.fmt == HY_F16P) {
                if (h.fallback) launch<T, true, 0, false, HY_F16P>(a, g, st);
                else launch<T, false, 0, false, HY_F16P>(a, g, st);
            } else if (h.fmt == HY_U16P) {
                if (h.fallback) launch<T, true, 0, false, HY_U16P>(a, g, st);
                else launch<T, false, 0, false, HY_U16P>(a, g, st);
            } else {
                if (h.fallback) launch<T, true, 2, true, HY_U16SP>(a, g, st);
                else launch<T, false, 2, true, HY_U16SP>(a, g, st);
            }
            return hipGetLastError();
        }
    }
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
template const void* hydro_march_ptr<double>(int, int, int, int, int);
template const void* hydro_march_ptr<float>(int, int, int, int, int);

}  // namespace glm
