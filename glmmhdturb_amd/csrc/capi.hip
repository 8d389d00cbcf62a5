// capi.hip -- extern "C" entry points of libglmmhd.so (declared in include/glmmhd.h).
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/glmmhd.h"
#include "adj_launch.hpp"
#include "aux_launch.hpp"
#include "ct_launch.hpp"
#include "host_common.hpp"
#include "hydro_launch.hpp"
#include "launch.hpp"
#include "ou_launch.hpp"

#ifndef GLM_DEFAULT_N
#define GLM_DEFAULT_N 480
#endif
#ifndef GLM_LIB_NV
#define GLM_LIB_NV 9
#endif
#ifndef GLM_LIB_CMA
#define GLM_LIB_CMA 0
#endif
// The scheme of a bridge library's march_run_dev / ct_run, fixed when it is built as the reference
// fixes it with -D flags: 0 the default build; 1 the parabolic edges (spike_mhd.cu / spike_ctm.cu
// -DPPM, spike_25d.cu -DPPM1D); 2 the hydro march's donor-cell seams (spike_25d.cu -DDONOR).
#ifndef GLM_LIB_SCHEME
#define GLM_LIB_SCHEME 0
#endif
static_assert(GLM_LIB_SCHEME >= 0 && GLM_LIB_SCHEME <= 2 && (GLM_LIB_SCHEME != 2 || GLM_LIB_NV == 5 || GLM_LIB_NV == 7),
              "GLM_LIB_SCHEME: 0, 1 (PPM), or 2 (donor seams, hydro libraries only)");

using namespace glm;

// ------------------------------------------------------------------------------------------
// error state (thread local)
// ------------------------------------------------------------------------------------------
static thread_local int t_last_error = GLM_OK;
static thread_local char t_last_msg[256] = "";

namespace glm {
int set_err(int code, const char* msg)
{
    t_last_error = code;
    snprintf(t_last_msg, sizeof t_last_msg, "%s", msg);
    return code;
}
int hip_err(hipError_t e, const char* where)
{
    if (e == hipSuccess) return GLM_OK;
    char buf[256];
    snprintf(buf, sizeof buf, "%s: %s", where, hipGetErrorString(e));
    return set_err(GLM_ERR_HIP, buf);
}
}  // namespace glm

// True when any plane of a is one of the planes of b.  A step that writes o while other
// workgroups still read q (halo planes, the update base) races if the two share a plane.
static bool planes_alias(void* const* a, int na, void* const* b, int nb)
{
    for (int v = 0; v < na; ++v)
        for (int w = 0; w < nb; ++w)
            if (a[v] == b[w]) return true;
    return false;
}

// ------------------------------------------------------------------------------------------
// context
// ------------------------------------------------------------------------------------------
struct glm_ctx {
    int nx, ny, nz, prec;
    size_t n;
    hipStream_t stream;
    unsigned long long* d_cmax;  // [0],[1] march CFL ping-pong, [2] CT step CFL, [3] standalone reductions
    int cfl_parity;              // slot the next march step accumulates into
    int cfl_slot;                // slot holding the CFL max of the last step with want_cfl (-1: none)
    void* d_parts;
    double* d_diag;
    // OU forcing: one band-limited box (host copy of the descriptor, and the device copy)
    int g;
    long long nstep;
    OUBox ou;
    OUBox* d_ou;
    void* afield;
    // CT scratch (face fluxes / EMFs of the multi-pass CT path)
    void* ct_scratch;
    // reverse step: scratch and the forcing-adjoint accumulator
    void* adj_scratch;
    double* d_fdot;
    // concurrent march workgroups (CUs x resident workgroups per CU) of the last march variant
    int slots_key;
    int slots;
    int ct_slots_key;
    int ct_slots;
    int hydro_slots_key;
    int hydro_slots;
};

static void ctx_free(glm_ctx* c)
{
    if (!c) return;
    void* ptrs[] = {c->d_cmax, c->d_parts, c->d_diag, c->d_ou, c->ct_scratch, c->adj_scratch, c->d_fdot};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    ou_box_free(c->ou);
    delete c;
}

extern "C" {

void glm_params_default(glm_params* p)
{
    memset(p, 0, sizeof *p);
    p->gamma = 1.01;                 // Params.gamma (glmmhd_turb.jl:31)
    p->smallr = 1e-6;                // :33
    p->smallc = 1e-4;                // :34
    p->pfl = 0.0;                    // -> smallr*smallc^2/gamma (:55)
    p->dt = 1e-4;
    p->dx = 1.0 / 256;
    p->ch = 1.0;
    p->glm_fac = 1.0;
    p->llf_dmin = 0.0;               // :39-40
    p->llf_pmin = 0.0;
    p->turb_min_rho = 1e-5;          // :50
    p->ramp = 0.0;
    p->predictor = GLM_PRED_FULL;
    p->recon = GLM_RECON_PLM;
    p->slope = GLM_SLOPE_MONCEN;
    p->riemann = GLM_RIEMANN_HLLD;
    p->mh_revert = 1;
    p->face_guard = GLM_GUARD_IFLO;
    p->update_floor = 1;
    p->do_turb = 0;
    p->zchunk = 0;
    p->want_cfl = 1;
}

glm_ctx* glm_create(int nx, int ny, int nz, int precision)
{
    if (nx < 1 || ny < 1 || nz < 1 || (precision != 8 && precision != 4)) {
        set_err(GLM_ERR_ARG, "glm_create: grid must be >= 1 per axis and precision 8 or 4");
        return nullptr;
    }
    glm_ctx* c = new glm_ctx();
    memset(c, 0, sizeof *c);
    c->nx = nx; c->ny = ny; c->nz = nz; c->prec = precision;
    c->n = (size_t)nx * ny * nz;
    c->stream = nullptr;
    size_t np = diag_nparts(c->n);
    c->cfl_parity = 0;
    c->cfl_slot = -1;
    if (hipMalloc(&c->d_cmax, 4 * sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc(&c->d_parts, np * diag_part_bytes()) != hipSuccess ||
        hipMalloc(&c->d_diag, GLM_NDIAG * sizeof(double)) != hipSuccess) {
        set_err(GLM_ERR_ALLOC, "glm_create: device allocation failed");
        ctx_free(c);
        return nullptr;
    }
    if (hipMemset(c->d_cmax, 0, 4 * sizeof(unsigned long long)) != hipSuccess) {
        set_err(GLM_ERR_HIP, "glm_create: hipMemset failed");
        ctx_free(c);
        return nullptr;
    }
    return c;
}

void glm_destroy(glm_ctx* ctx) { ctx_free(ctx); }

int glm_set_stream(glm_ctx* ctx, void* stream)
{
    if (!ctx) return set_err(GLM_ERR_ARG, "glm_set_stream: null context");
    ctx->stream = (hipStream_t)stream;
    return GLM_OK;
}

}  // extern "C"

// One march step over 9 + nsp planes (nsp = 0 or 2 passive species); with floor, the march's
// memory floor (march_floor_kernel: the same data movement, no flux computation).
template <typename T>
static int step_impl(glm_ctx* c, void* const* q, void* const* o, const glm_params* p, int nsp = 0, bool floor = false)
{
    MarchArgs<T> a;
    memset(&a, 0, sizeof a);
    const int nv = 9 + nsp;
    for (int v = 0; v < nv; ++v) {
        if (!q[v] || !o[v]) return set_err(GLM_ERR_ARG, "glm_step: null plane pointer");
        a.q[v] = (const T*)q[v];
        a.o[v] = (T*)o[v];
    }
    if (planes_alias(q, nv, o, nv)) return set_err(GLM_ERR_ARG, "glm_step: input and output planes must differ");
    a.nx = c->nx; a.ny = c->ny; a.nz = c->nz;
    const int key = floor ? -1 : 1 + (((p->predictor * 4 + p->riemann) * 4 + p->recon) * 4 + p->slope) + 1000 * nsp;
    if (c->slots_key != key) {
        int dev = 0, cus = 0, occ = 0;
        const void* fn = floor ? march_floor_ptr<T>() : march_kernel_ptr<T>(p->predictor, p->riemann, p->recon, p->slope, nsp);
        if (!fn) return set_err(GLM_ERR_ARG, "glm_step: unknown scheme");
        HIPCHK(hipGetDevice(&dev));
        HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, NT, 0));
        c->slots = cus * (occ > 0 ? occ : 1);
        c->slots_key = key;
    }
    march_tiling(c->nx, c->ny, c->nz, p->zchunk, c->slots, a.tiles_x, a.tiles_y, a.nzc, a.zchunk);
    a.ph = make_phys<T>(p);
    double dtdx = p->dt / p->dx;
    a.dtdx = (T)dtdx;
    a.half_dtdx = (T)(0.5 * dtdx);
    a.glm_fac = (T)p->glm_fac;
    a.smallc = (T)p->smallc;
    a.turb_min_rho = (T)p->turb_min_rho;
    a.ramp = (T)p->ramp;
    a.dt = (T)p->dt;
    a.dx = (T)p->dx;
    a.lx = (T)(c->nx * p->dx); a.ly = (T)(c->ny * p->dx); a.lz = (T)(c->nz * p->dx);
    a.recon = p->recon; a.slope = p->slope; a.mh_revert = p->mh_revert;
    a.face_guard = p->face_guard; a.update_floor = p->update_floor;
    a.do_turb = p->do_turb;
    if (p->do_turb) {
        if (!c->afield) return set_err(GLM_ERR_STATE, "glm_step: do_turb set but glm_ou_setup was not called");
        a.afield = (const T*)c->afield;
        a.tgs = c->g;
    }
    a.want_cfl = p->want_cfl;
    a.cmax = c->d_cmax + c->cfl_parity;
    a.cmax_next = c->d_cmax + (1 - c->cfl_parity);
    if (floor) {
        HIPCHK(launch_march_floor<T>(a, c->stream));
        return GLM_OK;
    }
    HIPCHK(launch_march<T>(a, p->predictor, p->riemann, p->recon, p->slope, c->stream, nsp));
    if (p->want_cfl) {
        c->cfl_slot = c->cfl_parity;
        c->cfl_parity = 1 - c->cfl_parity;
    }
    return GLM_OK;
}

extern "C" {

int glm_step(glm_ctx* ctx, void* const* q, void* const* o, const glm_params* p)
{
    if (!ctx || !q || !o) return set_err(GLM_ERR_ARG, "glm_step: null argument");
    int e = check_params(p);
    if (e) return e;
    return ctx->prec == 8 ? step_impl<double>(ctx, q, o, p) : step_impl<float>(ctx, q, o, p);
}

int glm_march_floor(glm_ctx* ctx, void* const* q, void* const* o, const glm_params* p)
{
    if (!ctx || !q || !o) return set_err(GLM_ERR_ARG, "glm_march_floor: null argument");
    int e = check_params(p);
    if (e) return e;
    return ctx->prec == 8 ? step_impl<double>(ctx, q, o, p, 0, true) : step_impl<float>(ctx, q, o, p, 0, true);
}

int glm_step_species(glm_ctx* ctx, void* const* q, void* const* o, const glm_params* p, int nsp)
{
    if (!ctx || !q || !o) return set_err(GLM_ERR_ARG, "glm_step_species: null argument");
    if (nsp != 0 && nsp != 2) return set_err(GLM_ERR_ARG, "glm_step_species: nsp must be 0 or 2");
    int e = check_params(p);
    if (e) return e;
    return ctx->prec == 8 ? step_impl<double>(ctx, q, o, p, nsp) : step_impl<float>(ctx, q, o, p, nsp);
}

int glm_cfl_max(glm_ctx* ctx, double* cmax)
{
    if (!ctx || !cmax) return set_err(GLM_ERR_ARG, "glm_cfl_max: null argument");
    if (ctx->cfl_slot < 0) return set_err(GLM_ERR_STATE, "glm_cfl_max: no step with want_cfl has run");
    unsigned long long bits = 0;
    HIPCHK(hipMemcpyAsync(&bits, ctx->d_cmax + ctx->cfl_slot, sizeof bits, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    memcpy(cmax, &bits, sizeof bits);
    return GLM_OK;
}

}  // extern "C"

template <typename T>
static Ptr9<T> ptr9(void* const* q)
{
    Ptr9<T> r;
    for (int v = 0; v < 9; ++v) r.p[v] = (T*)q[v];
    return r;
}

template <typename T>
static Ptr9<const T> cptr9(void* const* q)
{
    Ptr9<const T> r;
    for (int v = 0; v < 9; ++v) r.p[v] = (const T*)q[v];
    return r;
}

extern "C" {

}  // extern "C"

static int ctot_impl(glm_ctx* ctx, void* const* q, const glm_params* p, double* cmax, bool fast, const char* who)
{
    char msg[128];
    if (!ctx || !cmax) {
        snprintf(msg, sizeof msg, "%s: null argument", who);
        return set_err(GLM_ERR_ARG, msg);
    }
    snprintf(msg, sizeof msg, "%s: null plane pointer", who);
    int e = check_planes(q, 9, msg);
    if (!e) e = check_params(p);
    if (e) return e;
    HIPCHK(launch_clear_u64(ctx->d_cmax + 3, 1, ctx->stream));
    if (ctx->prec == 8)
        HIPCHK(launch_ctot<double>(cptr9<double>(q), ctx->n, make_phys<double>(p), ctx->d_cmax + 3, ctx->stream, fast));
    else HIPCHK(launch_ctot<float>(cptr9<float>(q), ctx->n, make_phys<float>(p), ctx->d_cmax + 3, ctx->stream, fast));
    unsigned long long bits = 0;
    HIPCHK(hipMemcpyAsync(&bits, ctx->d_cmax + 3, sizeof bits, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    memcpy(cmax, &bits, sizeof bits);
    return GLM_OK;
}

extern "C" {

int glm_ctot_max(glm_ctx* ctx, void* const* q, const glm_params* p, double* cmax)
{
    return ctot_impl(ctx, q, p, cmax, false, "glm_ctot_max");
}

int glm_fast_max(glm_ctx* ctx, void* const* q, const glm_params* p, double* cmax)
{
    return ctot_impl(ctx, q, p, cmax, true, "glm_fast_max");
}

int glm_diagnostics(glm_ctx* ctx, void* const* q, const glm_params* p, double* out)
{
    if (!ctx || !out) return set_err(GLM_ERR_ARG, "glm_diagnostics: null argument");
    int e = check_planes(q, 9, "glm_diagnostics: null plane pointer");
    if (!e) e = check_params(p);
    if (e) return e;
    if (ctx->prec == 8)
        HIPCHK(launch_diag<double>(cptr9<double>(q), ctx->nx, ctx->ny, ctx->nz, p->dx, p->smallr, ctx->d_parts, ctx->d_diag, ctx->stream));
    else
        HIPCHK(launch_diag<float>(cptr9<float>(q), ctx->nx, ctx->ny, ctx->nz, p->dx, p->smallr, ctx->d_parts, ctx->d_diag, ctx->stream));
    HIPCHK(hipMemcpyAsync(out, ctx->d_diag, GLM_NDIAG * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return GLM_OK;
}

}  // extern "C"

// ------------------------------------------------------------------------------------------
// Diagnostics without allocation.  The per-workgroup partials of the diagnostics reduction live
// in `scratch`, a file-scope device array (one copy per GPU, held until the process ends).  In
// host code `scratch` names a host-side stand-in, so its device address is taken once per device
// with hipGetSymbolAddress and only that address is handed to the kernels.
// ------------------------------------------------------------------------------------------
constexpr size_t NOALLOC_MAX_PARTS = 8192;   // diagnostics workgroups of 16384 cells: grids up to 2^27 cells (512^3)
constexpr size_t NOALLOC_PART_BYTES = 136;   // one workgroup's partial (7 compensated sums, divB, rho_min, finite)
static __device__ __align__(256) unsigned char scratch[NOALLOC_MAX_PARTS * NOALLOC_PART_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[NOALLOC_MAX_DEVICES];

// Device address of `scratch` on the current device (looked up on the first call per device).
static int noalloc_scratch(void** out)
{
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return set_err(GLM_ERR_STATE, "glm_diagnostics_noalloc: device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return set_err(GLM_ERR_STATE, "glm_diagnostics_noalloc: no device address for the static scratch");
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return GLM_OK;
}

extern "C" {

int glm_diagnostics_noalloc(void* const* q, int nx, int ny, int nz, int precision, const glm_params* p, void* stream,
                            double* out)
{
    if (!out) return set_err(GLM_ERR_ARG, "glm_diagnostics_noalloc: null output");
    if (nx < 1 || ny < 1 || nz < 1 || (precision != 8 && precision != 4))
        return set_err(GLM_ERR_ARG, "glm_diagnostics_noalloc: grid must be >= 1 per axis and precision 8 or 4");
    int e = check_planes(q, 9, "glm_diagnostics_noalloc: null plane pointer");
    if (!e) e = check_params(p);
    if (e) return e;
    const size_t n = (size_t)nx * ny * nz;
    const size_t np = diag_nparts(n);
    if (np > NOALLOC_MAX_PARTS) {
        char msg[200];
        snprintf(msg, sizeof msg, "glm_diagnostics_noalloc: %zu cells exceed the static scratch's maximum of %zu cells",
                 n, NOALLOC_MAX_PARTS * DIAG_CELLS_PER_BLOCK);
        return set_err(GLM_ERR_ARG, msg);
    }
    if (diag_part_bytes() > NOALLOC_PART_BYTES) return set_err(GLM_ERR_STATE, "glm_diagnostics_noalloc: partial larger than its slot");
    void* parts = nullptr;
    if ((e = noalloc_scratch(&parts))) return e;
    hipStream_t st = (hipStream_t)stream;
    if (precision == 8) HIPCHK(launch_diag<double>(cptr9<double>(q), nx, ny, nz, p->dx, p->smallr, parts, out, st));
    else HIPCHK(launch_diag<float>(cptr9<float>(q), nx, ny, nz, p->dx, p->smallr, parts, out, st));
    return GLM_OK;
}

int glm_ic(glm_ctx* ctx, void* const* q, int kind, const double* args)
{
    if (!ctx || !args) return set_err(GLM_ERR_ARG, "glm_ic: null argument");
    if (int e = check_planes(q, 9, "glm_ic: null plane pointer")) return e;
    if (kind < 0 || kind > 2) return set_err(GLM_ERR_ARG, "glm_ic: unknown kind");
    if (ctx->prec == 8) HIPCHK(launch_ic<double>(ptr9<double>(q), ctx->nx, ctx->ny, ctx->nz, kind, args, ctx->stream));
    else HIPCHK(launch_ic<float>(ptr9<float>(q), ctx->nx, ctx->ny, ctx->nz, kind, args, ctx->stream));
    return GLM_OK;
}

int glm_ou_setup(glm_ctx* ctx, int g, double kmin, double kmax, double T, int ndt, double comp_frac, double amp,
                 unsigned long long seed)
{
    if (!ctx) return set_err(GLM_ERR_ARG, "glm_ou_setup: null context");
    if (g < 2 || g > 256 || !(T > 0.0) || ndt < 1 || !(kmax >= kmin))
        return set_err(GLM_ERR_ARG, "glm_ou_setup: need 2 <= g <= 256, T > 0, ndt >= 1, kmax >= kmin");
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ou_box_free(ctx->ou);
    ctx->afield = nullptr;
    ctx->g = g;
    ctx->nstep = 0;
    if (ou_box_alloc(ctx->ou, g, kmin, kmax, T, ndt, comp_frac, amp, seed, ctx->prec, ctx->stream) != hipSuccess)
        return set_err(GLM_ERR_ALLOC, "glm_ou_setup: device allocation failed");
    if (!ctx->d_ou && hipMalloc(&ctx->d_ou, sizeof(OUBox)) != hipSuccess)
        return set_err(GLM_ERR_ALLOC, "glm_ou_setup: device allocation failed");
    HIPCHK(hipMemcpyAsync(ctx->d_ou, &ctx->ou, sizeof(OUBox), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->afield = ctx->ou.afield;
    return GLM_OK;
}

int glm_ou_advance(glm_ctx* ctx)
{
    if (!ctx || !ctx->afield) return set_err(GLM_ERR_STATE, "glm_ou_advance: call glm_ou_setup first");
    HIPCHK(launch_ou_advance(ctx->prec, ctx->d_ou, 1, ctx->ou.g, ctx->ou.nb, ctx->nstep, ctx->stream));
    ctx->nstep += 1;
    return GLM_OK;
}

const void* glm_ou_afield(glm_ctx* ctx) { return ctx ? ctx->afield : nullptr; }

int glm_ou_copy_afield(glm_ctx* ctx, void* dst)
{
    if (!ctx || !ctx->afield || !dst) return set_err(GLM_ERR_STATE, "glm_ou_copy_afield: no forcing field or null dst");
    size_t bytes = 3 * (size_t)ctx->g * ctx->g * ctx->g * (size_t)ctx->prec;
    HIPCHK(hipMemcpyAsync(dst, ctx->afield, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return GLM_OK;
}
long long glm_ou_steps(glm_ctx* ctx) { return ctx ? ctx->nstep : -1; }

int glm_ou_set_afield(glm_ctx* ctx, const void* src)
{
    if (!ctx || !ctx->afield || !src) return set_err(GLM_ERR_STATE, "glm_ou_set_afield: no forcing field or null src");
    size_t bytes = 3 * (size_t)ctx->g * ctx->g * ctx->g * (size_t)ctx->prec;
    HIPCHK(hipMemcpyAsync(ctx->afield, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return GLM_OK;
}

static int step_adjoint_impl(glm_ctx* ctx, void* const* u0, void* const* lam1, void* const* lam0, const glm_params* p,
                             const void* afield, bool fused)
{
    if (!ctx || !u0 || !lam1 || !lam0) return set_err(GLM_ERR_ARG, "glm_step_adjoint: null argument");
    int e = check_params(p);
    if (e) return e;
    if (ctx->prec != 8) return set_err(GLM_ERR_ARG, "glm_step_adjoint: the reverse step is FP64 only");
    if (p->recon == GLM_RECON_PPM_SPIKE) return set_err(GLM_ERR_ARG, "glm_step_adjoint: PPM_SPIKE has no adjoint");
    if (p->do_turb && (!afield || ctx->g < 2))
        return set_err(GLM_ERR_STATE, "glm_step_adjoint: do_turb needs the step's forcing field (and glm_ou_setup)");
    for (int v = 0; v < 9; ++v)
        if (!u0[v] || !lam1[v] || !lam0[v])
            return set_err(GLM_ERR_ARG, "glm_step_adjoint: bad plane pointers (null)");
    if (planes_alias(lam0, 9, lam1, 9) || planes_alias(lam0, 9, u0, 9))
        return set_err(GLM_ERR_ARG, "glm_step_adjoint: bad plane pointers (lam0 must not alias lam1 or u0)");
    if (!ctx->adj_scratch && hipMalloc(&ctx->adj_scratch, adj_scratch_bytes(ctx->n)) != hipSuccess) {
        ctx->adj_scratch = nullptr;
        return set_err(GLM_ERR_ALLOC, "glm_step_adjoint: scratch allocation failed");
    }
    if (!ctx->d_fdot) {
        if (hipMalloc(&ctx->d_fdot, sizeof(double)) != hipSuccess) {
            ctx->d_fdot = nullptr;
            return set_err(GLM_ERR_ALLOC, "glm_step_adjoint: allocation failed");
        }
        HIPCHK(hipMemsetAsync(ctx->d_fdot, 0, sizeof(double), ctx->stream));
    }
    AdjArgs a;
    memset(&a, 0, sizeof a);
    for (int v = 0; v < 9; ++v) {
        a.u0[v] = (const double*)u0[v];
        a.lam1[v] = (const double*)lam1[v];
        a.lam0[v] = (double*)lam0[v];
    }
    a.nx = ctx->nx; a.ny = ctx->ny; a.nz = ctx->nz;
    a.tgs = ctx->g;
    a.ph = make_phys<double>(p);
    a.dtdx = p->dt / p->dx;
    a.half_dtdx = 0.5 * a.dtdx;
    a.glm_fac = p->glm_fac; a.smallc = p->smallc; a.turb_min_rho = p->turb_min_rho; a.ramp = p->ramp;
    a.dt = p->dt; a.dx = p->dx;
    a.lx = ctx->nx * p->dx; a.ly = ctx->ny * p->dx; a.lz = ctx->nz * p->dx;
    a.mh_revert = p->mh_revert; a.face_guard = p->face_guard; a.update_floor = p->update_floor;
    a.do_turb = p->do_turb;
    a.afield = (const double*)afield;
    hipError_t he = launch_step_adjoint(a, ctx->adj_scratch, p->predictor, p->riemann, p->recon, p->slope, ctx->d_fdot,
                                        fused, p->zchunk, ctx->stream);
    if (he == hipErrorInvalidValue) return set_err(GLM_ERR_ARG, "glm_step_adjoint: unsupported scheme");
    HIPCHK(he);
    return GLM_OK;
}

int glm_step_adjoint(glm_ctx* ctx, void* const* u0, void* const* lam1, void* const* lam0, const glm_params* p,
                     const void* afield)
{
    return step_adjoint_impl(ctx, u0, lam1, lam0, p, afield, true);
}

int glm_step_adjoint_passes(glm_ctx* ctx, void* const* u0, void* const* lam1, void* const* lam0, const glm_params* p,
                            const void* afield)
{
    return step_adjoint_impl(ctx, u0, lam1, lam0, p, afield, false);
}

int glm_adjoint_force_dot(glm_ctx* ctx, int reset, double* out)
{
    if (!ctx || !out) return set_err(GLM_ERR_ARG, "glm_adjoint_force_dot: null argument");
    *out = 0.0;
    if (!ctx->d_fdot) return GLM_OK;
    HIPCHK(hipMemcpyAsync(out, ctx->d_fdot, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (reset) HIPCHK(hipMemsetAsync(ctx->d_fdot, 0, sizeof(double), ctx->stream));
    return GLM_OK;
}

// CT step paths: the multi-pass kernel sequence, the fused z-march, the staged 3-D tiles.
enum { CT_PASSES = 0, CT_MARCH = 1, CT_TILES = 2 };

static int ct_step_impl(glm_ctx* ctx, void* const* q, void* const* o, const glm_params* p, int ct_face_slope,
                        int mode, const char* who, int nsp = 0)
{
    if (!ctx || !q || !o) return set_err(GLM_ERR_ARG, who);
    int e = check_params(p);
    if (e) return e;
    if (p->riemann == GLM_RIEMANN_LLF_BOUND || p->recon == GLM_RECON_PPM)
        return set_err(GLM_ERR_ARG, "glm_ct_step: riemann must be LLF/HLL/HLLD and recon PLM/PPM_SPIKE");
    if (nsp != 0 && (nsp != 2 || mode != CT_MARCH))
        return set_err(GLM_ERR_ARG, "glm_ct_step_species: nsp must be 0 or 2 (the fused march)");
    const int nv = 8 + nsp;
    for (int v = 0; v < nv; ++v)
        if (!q[v] || !o[v]) return set_err(GLM_ERR_ARG, "glm_ct_step: bad plane pointers");
    if (planes_alias(q, nv, o, nv)) return set_err(GLM_ERR_ARG, "glm_ct_step: bad plane pointers (input and output alias)");
    if (p->do_turb && !ctx->afield) return set_err(GLM_ERR_STATE, "glm_ct_step: do_turb without glm_ou_setup");
    if (mode == CT_PASSES && !ctx->ct_scratch) {
        if (hipMalloc(&ctx->ct_scratch, ct_scratch_bytes(ctx->n, ctx->prec)) != hipSuccess)
            return set_err(GLM_ERR_ALLOC, "glm_ct_step: scratch allocation failed");
    }
    CTArgsHost h;
    h.nx = ctx->nx; h.ny = ctx->ny; h.nz = ctx->nz;
    h.gamma = p->gamma; h.smallr = p->smallr;
    h.smallp = p->pfl > 0.0 ? p->pfl : p->smallr * p->smallc * p->smallc / p->gamma;
    h.dt = p->dt; h.dx = p->dx; h.smallc = p->smallc; h.turb_min_rho = p->turb_min_rho; h.ramp = p->ramp;
    h.face_slope = ct_face_slope; h.riemann = p->riemann; h.recon = p->recon;
    h.update_floor = p->update_floor; h.do_turb = p->do_turb; h.tgs = ctx->g; h.afield = ctx->afield;
    h.want_cfl = p->want_cfl; h.cmax = ctx->d_cmax + 2; h.zchunk = p->zchunk; h.nsp = nsp;
    if (p->want_cfl) HIPCHK(launch_clear_u64(ctx->d_cmax + 2, 1, ctx->stream));
    if (mode == CT_TILES) {
        if (ctx->prec == 8) HIPCHK(launch_ct_tiles<double>(q, o, h, ctx->stream));
        else HIPCHK(launch_ct_tiles<float>(q, o, h, ctx->stream));
    } else if (mode == CT_MARCH) {
        const int ppm = p->recon == GLM_RECON_PPM_SPIKE;
        const int key = 1000 + p->riemann + 10 * nsp + 100 * ppm;
        if (ctx->ct_slots_key != key) {
            int dev = 0, cus = 0, occ = 0;
            const void* fn = ctx->prec == 8 ? ct_march_ptr<double>(p->riemann, nsp, p->recon)
                                            : ct_march_ptr<float>(p->riemann, nsp, p->recon);
            if (!fn) return set_err(GLM_ERR_ARG, "glm_ct_step: unknown Riemann solver");
            HIPCHK(hipGetDevice(&dev));
            HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
            HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, CT_MARCH_THREADS, 0));
            ctx->ct_slots = cus * (occ > 0 ? occ : 1);
            ctx->ct_slots_key = key;
        }
        if (ctx->prec == 8) HIPCHK(launch_ct_march<double>(q, o, h, ctx->ct_slots, ctx->stream));
        else HIPCHK(launch_ct_march<float>(q, o, h, ctx->ct_slots, ctx->stream));
    } else {
        if (ctx->prec == 8) HIPCHK(launch_ct_step<double>(q, o, h, ctx->ct_scratch, ctx->stream));
        else HIPCHK(launch_ct_step<float>(q, o, h, ctx->ct_scratch, ctx->stream));
    }
    if (p->want_cfl) ctx->cfl_slot = 2;
    return GLM_OK;
}

int glm_ct_step(glm_ctx* ctx, void* const* q, void* const* o, const glm_params* p, int ct_face_slope)
{
    return ct_step_impl(ctx, q, o, p, ct_face_slope, CT_MARCH, "glm_ct_step: null argument");
}

int glm_ct_step_species(glm_ctx* ctx, void* const* q, void* const* o, const glm_params* p, int ct_face_slope, int nsp)
{
    return ct_step_impl(ctx, q, o, p, ct_face_slope, CT_MARCH, "glm_ct_step_species: null argument", nsp);
}

int glm_ct_step_passes(glm_ctx* ctx, void* const* q, void* const* o, const glm_params* p, int ct_face_slope)
{
    return ct_step_impl(ctx, q, o, p, ct_face_slope, CT_PASSES, "glm_ct_step_passes: null argument");
}

int glm_ct_step_tiles(glm_ctx* ctx, void* const* q, void* const* o, const glm_params* p, int ct_face_slope)
{
    return ct_step_impl(ctx, q, o, p, ct_face_slope, CT_TILES, "glm_ct_step_tiles: null argument");
}

int glm_ct_ctot_max(glm_ctx* ctx, void* const* q, const glm_params* p, double* cmax)
{
    if (!ctx || !cmax) return set_err(GLM_ERR_ARG, "glm_ct_ctot_max: null argument");
    int e = check_planes(q, 8, "glm_ct_ctot_max: null plane pointer");
    if (!e) e = check_params(p);
    if (e) return e;
    double smallp = p->pfl > 0.0 ? p->pfl : p->smallr * p->smallc * p->smallc / p->gamma;
    HIPCHK(launch_clear_u64(ctx->d_cmax + 3, 1, ctx->stream));
    if (ctx->prec == 8)
        HIPCHK(launch_ct_ctot<double>(q, ctx->nx, ctx->ny, ctx->nz, p->gamma, p->smallr, smallp, ctx->d_cmax + 3, ctx->stream));
    else
        HIPCHK(launch_ct_ctot<float>(q, ctx->nx, ctx->ny, ctx->nz, p->gamma, p->smallr, smallp, ctx->d_cmax + 3, ctx->stream));
    unsigned long long bits = 0;
    HIPCHK(hipMemcpyAsync(&bits, ctx->d_cmax + 3, sizeof bits, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    memcpy(cmax, &bits, sizeof bits);
    return GLM_OK;
}

int glm_ct_divb_max(glm_ctx* ctx, void* const* q, double dx, double* out)
{
    if (!ctx || !out || !(dx > 0.0)) return set_err(GLM_ERR_ARG, "glm_ct_divb_max: bad argument");
    if (int e = check_planes(q, 8, "glm_ct_divb_max: null plane pointer")) return e;
    HIPCHK(launch_clear_u64(ctx->d_cmax + 3, 1, ctx->stream));
    if (ctx->prec == 8) HIPCHK(launch_ct_divb<double>(q, ctx->nx, ctx->ny, ctx->nz, dx, ctx->d_cmax + 3, ctx->stream));
    else HIPCHK(launch_ct_divb<float>(q, ctx->nx, ctx->ny, ctx->nz, dx, ctx->d_cmax + 3, ctx->stream));
    unsigned long long bits = 0;
    HIPCHK(hipMemcpyAsync(&bits, ctx->d_cmax + 3, sizeof bits, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    memcpy(out, &bits, sizeof bits);
    return GLM_OK;
}

// The hydro march (spike_25d.cu:383-516 as build_march.sh builds it), with nsp = 0 or 2 passive
// scalars (-DSCALARS) whose flux rides the mass flux when cma is set (-DCMA).
static int hydro_step_impl(glm_ctx* ctx, void* const* q, void* const* o, const glm_params* p, int nsp, int cma,
                           const char* who, int fmt = 0, int donor = 0)
{
    char msg[160];
    if (!ctx || !q || !o || !p) {
        snprintf(msg, sizeof msg, "%s: null argument", who);
        return set_err(GLM_ERR_ARG, msg);
    }
    if (nsp != 0 && nsp != 2) {
        snprintf(msg, sizeof msg, "%s: nsp must be 0 or 2", who);
        return set_err(GLM_ERR_ARG, msg);
    }
    if (!(p->gamma > 1.0) || !(p->dx > 0.0) || !(p->smallr > 0.0) || !(p->dt >= 0.0)) {
        snprintf(msg, sizeof msg, "%s: need gamma > 1, dx > 0, dt >= 0, smallr > 0", who);
        return set_err(GLM_ERR_ARG, msg);
    }
    const int nv = 5 + nsp;
    for (int v = 0; v < nv; ++v)
        if (!q[v] || !o[v]) {
            snprintf(msg, sizeof msg, "%s: null plane pointer", who);
            return set_err(GLM_ERR_ARG, msg);
        }
    if (planes_alias(q, nv, o, nv)) {
        snprintf(msg, sizeof msg, "%s: input and output planes alias", who);
        return set_err(GLM_ERR_ARG, msg);
    }
    HydroArgsHost h;
    h.nx = ctx->nx; h.ny = ctx->ny; h.nz = ctx->nz;
    h.gamma = p->gamma; h.dtdx = p->dt / p->dx; h.smallr = p->smallr;
    h.smallp = p->pfl > 0.0 ? p->pfl : p->smallr * p->smallc * p->smallc / p->gamma;
    h.fallback = p->llf_dmin > 0.0 || p->llf_pmin > 0.0;
    h.fb_dmin = p->llf_dmin > 0.0 ? p->llf_dmin : 0.0;
    h.fb_pmin = p->llf_pmin > 0.0 ? p->llf_pmin : 0.0;
    h.zchunk = p->zchunk;
    h.nsp = nsp;
    h.cma = nsp > 0 && cma;
    h.fmt = fmt;
    // the scheme variant: donor-cell seams, or the parabolic edges when recon is PPM_SPIKE
    h.scheme = donor ? 2 : p->recon == GLM_RECON_PPM_SPIKE ? 1 : 0;
    if (h.scheme != 0 && fmt != 0) {
        snprintf(msg, sizeof msg, "%s: the PPM1D and donor variants take conserved planes", who);
        return set_err(GLM_ERR_ARG, msg);
    }
    if (h.scheme != 0) h.fallback = 0;   // those builds solve HLL without the fallback (spike_25d.cu:445-483)
    const int key = 2000 + h.fallback + 2 * nsp + 8 * h.cma + 16 * fmt + 128 * h.scheme;
    if (ctx->hydro_slots_key != key) {
        int dev = 0, cus = 0, occ = 0;
        const void* fn = ctx->prec == 8 ? hydro_march_ptr<double>(h.fallback, nsp, h.cma, fmt, h.scheme)
                                        : hydro_march_ptr<float>(h.fallback, nsp, h.cma, fmt, h.scheme);
        if (!fn) return set_err(GLM_ERR_ARG, "hydro step: no kernel for this format and precision");
        HIPCHK(hipGetDevice(&dev));
        HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, HYDRO_THREADS, 0));
        ctx->hydro_slots = cus * (occ > 0 ? occ : 1);
        ctx->hydro_slots_key = key;
    }
    if (ctx->prec == 8) HIPCHK(launch_hydro_march<double>(q, o, h, ctx->hydro_slots, ctx->stream));
    else HIPCHK(launch_hydro_march<float>(q, o, h, ctx->hydro_slots, ctx->stream));
    return GLM_OK;
}

int glm_hydro_step(glm_ctx* ctx, void* const* q, void* const* o, const glm_params* p)
{
    return hydro_step_impl(ctx, q, o, p, 0, 0, "glm_hydro_step");
}

int glm_hydro_step_compact(glm_ctx* ctx, void* const* q, void* const* o, const glm_params* p, int format)
{
    if (format != GLM_HYDRO_F16_PRIM && format != GLM_HYDRO_U16_PRIM && format != GLM_HYDRO_U16_LOG2_SPECIES)
        return set_err(GLM_ERR_ARG, "glm_hydro_step_compact: unknown storage format");
    if (ctx && ctx->prec != 4) return set_err(GLM_ERR_ARG, "glm_hydro_step_compact: the compact formats need an FP32 context");
    const int nsp = format == GLM_HYDRO_U16_LOG2_SPECIES ? 2 : 0;
    return hydro_step_impl(ctx, q, o, p, nsp, nsp > 0, "glm_hydro_step_compact", format);
}

int glm_hydro_step_donor(glm_ctx* ctx, void* const* q, void* const* o, const glm_params* p, int nsp, int cma)
{
    return hydro_step_impl(ctx, q, o, p, nsp, cma, "glm_hydro_step_donor", 0, 1);
}

int glm_hydro_step_scalars(glm_ctx* ctx, void* const* q, void* const* o, const glm_params* p, int nsp, int cma)
{
    return hydro_step_impl(ctx, q, o, p, nsp, cma, "glm_hydro_step_scalars");
}

int glm_last_error(void) { return t_last_error; }

const char* glm_error_string(int code)
{
    if (code == t_last_error && t_last_msg[0]) return t_last_msg;
    switch (code) {
    case GLM_OK: return "ok";
    case GLM_ERR_ARG: return "invalid argument";
    case GLM_ERR_HIP: return "HIP runtime error";
    case GLM_ERR_ALLOC: return "device allocation failed";
    case GLM_ERR_STATE: return "invalid state";
    default: return "unknown error";
    }
}

int glm_kernel_info(int precision, int predictor, int riemann, int recon, int slope, int* vgprs, int* lds_bytes)
{
    hipFuncAttributes fa;
    const void* fn = nullptr;
    if (precision == 8) fn = march_kernel_ptr<double>(predictor, riemann, recon, slope);
    else if (precision == 4) fn = march_kernel_ptr<float>(predictor, riemann, recon, slope);
    if (!fn) return set_err(GLM_ERR_ARG, "glm_kernel_info: unknown variant");
    HIPCHK(hipFuncGetAttributes(&fa, fn));
    if (vgprs) *vgprs = fa.numRegs;
    if (lds_bytes) *lds_bytes = (int)fa.sharedSizeBytes;
    return GLM_OK;
}

// ------------------------------------------------------------------------------------------
// Legacy MarchBridge ABI (spike_mhd.cu:319-341, spike_ctm.cu:299-329)
// ------------------------------------------------------------------------------------------
static std::mutex g_mu;
static int g_nx = GLM_DEFAULT_N, g_ny = GLM_DEFAULT_N, g_nz = GLM_DEFAULT_N;
static float g_dtdx = GLM_LIB_NV == 5 || GLM_LIB_NV == 7 ? 0.05f : 0.02f, g_ch = 1.0f, g_glmfac = 1.0f, g_gamma = 1.4f;   // spike_mhd.cu:52, spike_25d.cu:180
static glm_ctx* g_ctx32 = nullptr;
static glm_ctx* g_ctx64 = nullptr;

int march_nv(void) { return GLM_LIB_NV; }
int march_nx(void) { return g_nx; }
int march_nsp(void) { return GLM_LIB_NV == 7 || GLM_LIB_NV == 10 || GLM_LIB_NV == 11 ? 2 : 0; }
void march_set_dtdx(float v) { g_dtdx = v; }
#if GLM_LIB_NV == 9 || GLM_LIB_NV == 11
// Only the GLM-MHD libraries have a cleaning speed: MarchBridge.set_glm decides by this symbol's
// presence (MarchBridge.jl:37-38, :50-53).
void march_set_glm(float ch, float fac) { g_ch = ch; g_glmfac = fac; }
#endif
void march_set_gamma(float g) { g_gamma = g; }

int march_set_grid(int nx, int ny, int nz)
{
    if (nx < 1 || ny < 1 || nz < 1) return set_err(GLM_ERR_ARG, "march_set_grid: bad grid");
    std::lock_guard<std::mutex> lk(g_mu);
    g_nx = nx; g_ny = ny; g_nz = nz;
    glm_destroy(g_ctx32); g_ctx32 = nullptr;
    glm_destroy(g_ctx64); g_ctx64 = nullptr;
    return GLM_OK;
}

}  // extern "C"

// The spike march scheme (spike_mhd.cu:257-288): normal-only Hancock, PLM MonCen (or with
// GLM_LIB_SCHEME 1 the parabolic edges), HLL, GLM pair, psi *= GLMFAC, faces clamped at
// SMALLR/SMALLP = 1e-6, no predictor revert, no update floor.
static glm_params legacy_params()
{
    glm_params p;
    glm_params_default(&p);
    p.gamma = g_gamma;
    p.smallr = 1e-6;
    p.pfl = 1e-6;
    p.dx = 1.0;
    p.dt = g_dtdx;
    p.ch = g_ch;
    p.glm_fac = g_glmfac;
    p.predictor = GLM_PRED_ONED;
    p.recon = GLM_LIB_SCHEME == 1 ? GLM_RECON_PPM_SPIKE : GLM_RECON_PLM;
    p.slope = GLM_SLOPE_MONCEN;
    p.riemann = GLM_RIEMANN_HLL;
    p.mh_revert = 0;
    p.face_guard = GLM_GUARD_CLAMP;
    p.update_floor = 0;
    p.do_turb = 0;
    p.want_cfl = 0;
    return p;
}

template <typename T>
static double legacy_run(T* const* qv, T* const* ov, int nsteps, int nv, int kind)
{
    std::lock_guard<std::mutex> lk(g_mu);
    glm_ctx*& c = sizeof(T) == 8 ? g_ctx64 : g_ctx32;
    if (!c) c = glm_create(g_nx, g_ny, g_nz, (int)sizeof(T));
    if (!c || !qv || !ov || nsteps < 0) return -1.0;
    glm_params p = legacy_params();
    std::vector<void*> q(qv, qv + nv), o(ov, ov + nv);
    hipEvent_t t0, t1;
    if (hipEventCreate(&t0) != hipSuccess || hipEventCreate(&t1) != hipSuccess) return -1.0;
    (void)hipEventRecord(t0, c->stream);
    int err = GLM_OK;
    for (int it = 0; it < nsteps && !err; ++it) {
        err = kind == 1   ? glm_ct_step(c, q.data(), o.data(), &p, 0)
              : kind == 5 ? glm_ct_step_species(c, q.data(), o.data(), &p, 0, 2)
              : kind == 2 ? (GLM_LIB_SCHEME == 2 ? glm_hydro_step_donor(c, q.data(), o.data(), &p, 0, 0)
                                                 : glm_hydro_step(c, q.data(), o.data(), &p))
              : kind == 3 ? (GLM_LIB_SCHEME == 2 ? glm_hydro_step_donor(c, q.data(), o.data(), &p, 2, GLM_LIB_CMA)
                                                 : glm_hydro_step_scalars(c, q.data(), o.data(), &p, 2, GLM_LIB_CMA))
              : kind == 4 ? glm_step_species(c, q.data(), o.data(), &p, 2)
                          : glm_step(c, q.data(), o.data(), &p);
        std::swap(q, o);
    }
    (void)hipEventRecord(t1, c->stream);
    (void)hipEventSynchronize(t1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, t0, t1);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (err) return -1.0;
    if (q[0] != (void*)qv[0]) {   // odd step count: result is in ov; copy back (spike_mhd.cu:337)
        for (int v = 0; v < nv; ++v)
            if (hipMemcpyAsync(qv[v], q[v], c->n * sizeof(T), hipMemcpyDeviceToDevice, c->stream) != hipSuccess) return -1.0;
        (void)hipStreamSynchronize(c->stream);
    }
    return (double)ms;
}

extern "C" {

// Which legacy entry points a library exports follows the reference library of its plane count:
//   NV 9 (libglmmhd.so, build_mhd.sh): the GLM march (spike_mhd.cu:319-341) and, in libglmmhd.so,
//        also the CT entry points;
//   NV 8 (build_ct.sh): the CT march only (spike_ctm.cu:299-329) -- no march_run_dev, which would
//        read a ninth (psi) plane the caller's 8-entry arrays do not hold;
//   NV 5 / 7 (build_march.sh [-DSCALARS]): the hydro march (spike_25d.cu:559-576), no CT entry points;
//   NV 11 (NSP=2 build_mhd.sh): the GLM march with two passive species (spike_mhd.cu -DNSP=2);
//   NV 10 (NSP=2 build_ct.sh): the CT march with two passive species (spike_ctm.cu -DNSP=2).
#if GLM_LIB_NV == 9 || GLM_LIB_NV == 5 || GLM_LIB_NV == 7 || GLM_LIB_NV == 11
constexpr int LEGACY_KIND = GLM_LIB_NV == 5 ? 2 : GLM_LIB_NV == 7 ? 3 : GLM_LIB_NV == 11 ? 4 : 0;
double march_run_dev(float* const* qv, float* const* ov, int nsteps)
{
    return legacy_run<float>(qv, ov, nsteps, GLM_LIB_NV, LEGACY_KIND);
}
double march_run_dev_f64(double* const* qv, double* const* ov, int nsteps)
{
    return legacy_run<double>(qv, ov, nsteps, GLM_LIB_NV, LEGACY_KIND);
}
#endif
#if GLM_LIB_NV == 9 || GLM_LIB_NV == 8 || GLM_LIB_NV == 10
// qv/ov: the 8 CT planes, then (NV 10) the 2 species planes rho X_s (spike_ctm.cu:304-311)
double ct_run(float* const* qv, float* const* ov, float* const* cc, int nsteps)
{
    (void)cc;
    return GLM_LIB_NV == 10 ? legacy_run<float>(qv, ov, nsteps, 10, 5) : legacy_run<float>(qv, ov, nsteps, 8, 1);
}
int ct_regs(void)
{
    hipFuncAttributes fa;
    if (hipFuncGetAttributes(&fa, ct_kernel_ptr<float>()) != hipSuccess) return -1;
    return fa.numRegs;
}
int ct_shmem(void)
{
    hipFuncAttributes fa;
    if (hipFuncGetAttributes(&fa, ct_kernel_ptr<float>()) != hipSuccess) return -1;
    return (int)fa.sharedSizeBytes;
}
#endif

}  // extern "C"
