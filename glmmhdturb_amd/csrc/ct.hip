// ct.hip -- launchers of the constrained-transport kernels (ct_impl.hpp).
#include "ct_impl.hpp"
#include "ct_march.hpp"
#include "ct_tile.hpp"
#include "ct_launch.hpp"

namespace glm {

size_t ct_scratch_bytes(size_t n, int prec) { return (size_t)(8 + 24 + 3) * n * (size_t)prec; }

template <typename T>
static CTArgs<T> make_ct_args(void* const* q, void* const* o, const CTArgsHost& h, void* scratch)
{
    CTArgs<T> a;
    size_t n = (size_t)h.nx * h.ny * h.nz;
    for (int v = 0; v < 9; ++v) {
        a.q.p[v] = v < 8 ? (const T*)q[v] : nullptr;
        a.o.p[v] = v < 8 ? (T*)o[v] : nullptr;
    }
    a.prim = (T*)scratch;                       // the multi-pass step's scratch; null for the one-kernel steps
    a.flx = scratch ? a.prim + 8 * n : nullptr;
    a.emf = scratch ? a.flx + 24 * n : nullptr;
    a.nx = h.nx; a.ny = h.ny; a.nz = h.nz;
    a.gamma = (T)h.gamma; a.igm1 = (T)(1.0 / (h.gamma - 1.0)); a.smallr = (T)h.smallr; a.smallp = (T)h.smallp;
    double dtdx = h.dt / h.dx;
    a.dtdx = (T)dtdx; a.half_dtdx = (T)(0.5 * dtdx); a.dt = (T)h.dt; a.dx = (T)h.dx;
    a.lx = (T)(h.nx * h.dx); a.ly = (T)(h.ny * h.dx); a.lz = (T)(h.nz * h.dx);
    a.smallc = (T)h.smallc; a.turb_min_rho = (T)h.turb_min_rho; a.ramp = (T)h.ramp;
    a.face_slope = h.face_slope; a.riemann = h.riemann; a.recon = h.recon; a.update_floor = h.update_floor;
    a.do_turb = h.do_turb; a.tgs = h.tgs; a.want_cfl = h.want_cfl;
    a.afield = (const T*)h.afield;
    a.cmax = h.cmax;
    return a;
}

template <typename T>
hipError_t launch_ct_step(void* const* q, void* const* o, const CTArgsHost& h, void* scratch, hipStream_t st)
{
    CTArgs<T> a = make_ct_args<T>(q, o, h, scratch);
    size_t n = (size_t)h.nx * h.ny * h.nz;
    dim3 g((unsigned)((n + CT_NT - 1) / CT_NT)), b(CT_NT);
    hipLaunchKernelGGL((ct_prim_kernel<T>), g, b, 0, st, a);
    hipLaunchKernelGGL((ct_flux_kernel<T>), g, b, 0, st, a);
    hipLaunchKernelGGL((ct_emf_kernel<T>), g, b, 0, st, a);
    hipLaunchKernelGGL((ct_update_kernel<T>), g, b, 0, st, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !h.want_cfl) return e;
    Ptr9<const T> o9;
    for (int v = 0; v < 9; ++v) o9.p[v] = v < 8 ? (const T*)o[v] : nullptr;
    hipLaunchKernelGGL((ct_ctot_kernel<T>), dim3(2048), b, 0, st, o9, h.nx, h.ny, h.nz, (T)h.gamma, (T)h.smallr,
                       (T)h.smallp, h.cmax);
    return hipGetLastError();
}

// The CT step over staged 3-D tiles (ct_tile.hpp), then the next step's CFL max as the other paths.
template <typename T>
hipError_t launch_ct_tiles(void* const* q, void* const* o, const CTArgsHost& h, hipStream_t st)
{
    CTArgs<T> a = make_ct_args<T>(q, o, h, nullptr);
    const int tx = (h.nx + CTT_OX - 1) / CTT_OX, ty = (h.ny + CTT_OY - 1) / CTT_OY;
    const int tz = (h.nz + CTTile<T>::OZ - 1) / CTTile<T>::OZ;
    const long long nt = (long long)tx * ty * tz;
    if (nt > 0x7fffffffLL) return hipErrorInvalidValue;
    hipLaunchKernelGGL((ct_tile_kernel<T>), dim3((unsigned)nt), dim3(CTT_NT), 0, st, a, tx, ty, tz);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !h.want_cfl) return e;
    Ptr9<const T> o9;
    for (int v = 0; v < 9; ++v) o9.p[v] = v < 8 ? (const T*)o[v] : nullptr;
    hipLaunchKernelGGL((ct_ctot_kernel<T>), dim3(2048), dim3(CT_NT), 0, st, o9, h.nx, h.ny, h.nz, (T)h.gamma,
                       (T)h.smallr, (T)h.smallp, h.cmax);
    return hipGetLastError();
}

template <typename T>
const void* ct_tile_ptr()
{
    return (const void*)&ct_tile_kernel<T>;
}

template <typename T, int NS, int RC>
static const void* ct_march_fn(int riemann)
{
    if (riemann == RIEM_HLLD) return (const void*)&ct_march_kernel<T, RIEM_HLLD, NS, RC>;
    if (riemann == RIEM_HLL) return (const void*)&ct_march_kernel<T, RIEM_HLL, NS, RC>;
    if (riemann == RIEM_LLF) return (const void*)&ct_march_kernel<T, RIEM_LLF, NS, RC>;
    return nullptr;
}

// The march kernel for a Riemann solver, species count and reconstruction (RECON_PPM_SPIKE the
// parabolic edges, any other value MonCen PLM), or null.
template <typename T>
const void* ct_march_ptr(int riemann, int nsp, int recon)
{
    const bool ppm = recon == RECON_PPM_SPIKE;
    if (nsp == 2) return ppm ? ct_march_fn<T, 2, RECON_PPM_SPIKE>(riemann) : ct_march_fn<T, 2, RECON_PLM>(riemann);
    if (nsp == 0) return ppm ? ct_march_fn<T, 0, RECON_PPM_SPIKE>(riemann) : ct_march_fn<T, 0, RECON_PLM>(riemann);
    return nullptr;
}

template <typename T>
hipError_t launch_ct_march(void* const* q, void* const* o, const CTArgsHost& h, int slots, hipStream_t st)
{
    if (h.nsp != 0 && h.nsp != 2) return hipErrorInvalidValue;
    CTMArgs<T> a;
    for (int v = 0; v < 10; ++v) {
        a.q[v] = v < 8 + h.nsp ? (const T*)q[v] : nullptr;
        a.o[v] = v < 8 + h.nsp ? (T*)o[v] : nullptr;
    }
    a.nx = h.nx; a.ny = h.ny; a.nz = h.nz;
    ct_march_tiling(h.nx, h.ny, h.nz, h.zchunk, slots, a.tiles_x, a.tiles_y, a.nzc, a.zchunk);
    a.gamma = (T)h.gamma; a.igm1 = (T)(1.0 / (h.gamma - 1.0)); a.smallr = (T)h.smallr; a.smallp = (T)h.smallp;
    double dtdx = h.dt / h.dx;
    a.dtdx = (T)dtdx; a.half_dtdx = (T)(0.5 * dtdx); a.dt = (T)h.dt; a.dx = (T)h.dx;
    a.lx = (T)(h.nx * h.dx); a.ly = (T)(h.ny * h.dx); a.lz = (T)(h.nz * h.dx);
    a.smallc = (T)h.smallc; a.turb_min_rho = (T)h.turb_min_rho; a.ramp = (T)h.ramp;
    a.face_slope = h.face_slope; a.recon = h.recon; a.update_floor = h.update_floor;
    a.do_turb = h.do_turb; a.tgs = h.tgs;
    a.afield = (const T*)h.afield;
    const void* fn = ct_march_ptr<T>(h.riemann, h.nsp, h.recon);
    if (!fn) return hipErrorInvalidValue;
    void* args[] = {&a};
    hipError_t e = hipLaunchKernel(fn, dim3((unsigned)(a.tiles_x * a.tiles_y * a.nzc)), dim3(CM_NT), args, 0, st);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess || !h.want_cfl) return e;
    Ptr9<const T> o9;
    for (int v = 0; v < 9; ++v) o9.p[v] = v < 8 ? (const T*)o[v] : nullptr;
    hipLaunchKernelGGL((ct_ctot_kernel<T>), dim3(2048), dim3(CT_NT), 0, st, o9, h.nx, h.ny, h.nz, (T)h.gamma,
                       (T)h.smallr, (T)h.smallp, h.cmax);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_ct_ctot(void* const* q, int nx, int ny, int nz, double gamma, double smallr, double smallp,
                          unsigned long long* out, hipStream_t st)
{
    Ptr9<const T> q9;
    for (int v = 0; v < 9; ++v) q9.p[v] = v < 8 ? (const T*)q[v] : nullptr;
    hipLaunchKernelGGL((ct_ctot_kernel<T>), dim3(2048), dim3(CT_NT), 0, st, q9, nx, ny, nz, (T)gamma, (T)smallr,
                       (T)smallp, out);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_ct_divb(void* const* q, int nx, int ny, int nz, double dx, unsigned long long* out, hipStream_t st)
{
    Ptr9<const T> q9;
    for (int v = 0; v < 9; ++v) q9.p[v] = v < 8 ? (const T*)q[v] : nullptr;
    hipLaunchKernelGGL((ct_divb_kernel<T>), dim3(2048), dim3(CT_NT), 0, st, q9, nx, ny, nz, dx, out);
    return hipGetLastError();
}

template <typename T>
const void* ct_kernel_ptr()
{
    return (const void*)&ct_march_kernel<T, RIEM_HLL>;
}

#define GLM_CT_INST(T)                                                                                          \
    template hipError_t launch_ct_step<T>(void* const*, void* const*, const CTArgsHost&, void*, hipStream_t);   \
    template hipError_t launch_ct_ctot<T>(void* const*, int, int, int, double, double, double,                  \
                                          unsigned long long*, hipStream_t);                                    \
    template hipError_t launch_ct_divb<T>(void* const*, int, int, int, double, unsigned long long*, hipStream_t); \
    template const void* ct_kernel_ptr<T>();                                                                    \
    template const void* ct_march_ptr<T>(int, int, int);                                                        \
    template hipError_t launch_ct_march<T>(void* const*, void* const*, const CTArgsHost&, int, hipStream_t);   \
    template hipError_t launch_ct_tiles<T>(void* const*, void* const*, const CTArgsHost&, hipStream_t);        \
    template const void* ct_tile_ptr<T>();
GLM_CT_INST(double)
GLM_CT_INST(float)
#undef GLM_CT_INST

}  // namespace glm
