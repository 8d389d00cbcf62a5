// ens.hip -- ensemble of boxes: the device step control, the sweep (step control, batched OU
// advance, one march launch over every box's tiles), its HIP graph, the per-box log and timing,
// and the launch autotuner.  C entry points glm_ens_* (include/glmmhd.h, section 3).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "host_common.hpp"
#include "aux_launch.hpp"
#include "launch.hpp"
#include "ou_launch.hpp"

using namespace glm;

namespace glm {

// Per-box control state of the run_turb loop (glmmhd_turb.jl:2191-2241), in device memory.
struct EnsBox {
    double t, next_time, dt_ou, turb_T;
    double courant, dx, boxlen, ch_scale, cp_coef;
    unsigned long long cmax;   // CFL max of the current state (bits of a non-negative double);
                               // read and cleared by the step control, refilled by the march
    long long nstep, ou_steps;
    int do_turb, parity;       // u[parity] holds the current state
    void* u[2][9];
};

constexpr int ENS_MAX_BOXES = 1024;

// One thread per box.  For box b, as run_turb does before each step:
//   while t >= next_time: ou_advance!      (at most one advance per step here: dt <= next_time - t)
//   dt = courant*dx/cmax, clipped to next_time - t
//   ch, psi decay and ramp of step_plm! (:884-886)
// then point the box's march arguments at the current and the other buffer and flip them.
template <typename T>
__global__ void __launch_bounds__(ENS_MAX_BOXES) ens_ctl_kernel(EnsBox* eb, MarchArgs<T>* ma, OUBox* ou, int nbox,
                                                                long long* sweep, unsigned int* queue, double* log,
                                                                unsigned long long* t0, unsigned long long* t1)
{
    __shared__ long long s_sw;
    if (threadIdx.x == 0) {
        s_sw = *sweep + 1;
        *sweep = s_sw;
    }
    if (threadIdx.x < NXCD) queue[threadIdx.x] = 0u;
    __syncthreads();
    const int b = threadIdx.x;
    if (b >= nbox) return;
    const long long sw = s_sw;
    const int slot = (int)(sw % ENS_LOG);
    EnsBox& e = eb[b];
    const double cmax = __longlong_as_double((long long)e.cmax);
    e.cmax = 0ull;
    const double t = e.t;
    int due = 0;
    if (e.do_turb) {
        if (t >= e.next_time) {
            due = 1;
            ou[b].step = (unsigned long long)e.ou_steps;
            e.ou_steps += 1;
            e.next_time = e.next_time + e.dt_ou;
        }
        ou[b].due = due;
    }
    double dt = e.courant * e.dx / cmax;
    if (e.do_turb) dt = fmin(dt, e.next_time - t);
    const double ch = e.courant * e.dx / dt / 3.0 * e.ch_scale;
    const double fac = e.cp_coef > 0.0 ? exp(-(ch * ch / (e.cp_coef * e.boxlen * ch)) * dt) : 1.0;
    const double ramp = fmin(t / e.turb_T, 1.0);
    const double dtdx = dt / e.dx;
    MarchArgs<T>& m = ma[b];
    const int par = e.parity;
    for (int v = 0; v < 9; ++v) {
        m.q[v] = (const T*)e.u[par][v];
        m.o[v] = (T*)e.u[1 - par][v];
    }
    m.ph.ch = (T)ch;
    m.ph.ch2 = (T)(ch * ch);
    m.ph.half_ich = (T)(0.5 / ch);
    m.dtdx = (T)dtdx;
    m.half_dtdx = (T)(0.5 * dtdx);
    m.glm_fac = (T)fac;
    m.ramp = (T)ramp;
    m.dt = (T)dt;
    e.parity = 1 - par;
    e.t = t + dt;
    e.nstep += 1;
    double* lg = log + ((size_t)slot * nbox + b) * 8;
    lg[0] = (double)sw; lg[1] = t; lg[2] = dt; lg[3] = ch; lg[4] = fac; lg[5] = ramp; lg[6] = cmax; lg[7] = due;
    t0[(size_t)slot * nbox + b] = ~0ull;
    t1[(size_t)slot * nbox + b] = 0ull;
}

}  // namespace glm

struct glm_ens {
    int nbox, prec;
    glm_params scheme;
    std::vector<glm_ens_box> box;
    std::vector<OUBox> ou;
    hipStream_t stream;
    int any_turb, gmax, nbmax;
    int cu_count;
    double clock_khz;
    // device tables
    EnsBox* d_eb;
    void* d_ma;            // MarchArgs<T>[nbox]
    OUBox* d_ou;
    int* d_order;
    int* d_tile_off;
    unsigned int* d_queue;
    long long* d_sweep;
    double* d_log;
    unsigned long long *d_t0, *d_t1;
    // launch configuration
    int zchunk, persistent, order;
    int total, grid;
    // graph
    int use_graph;
    hipGraph_t graph;
    hipGraphExec_t exec;
};

namespace {

size_t ma_bytes(int prec) { return prec == 8 ? sizeof(MarchArgs<double>) : sizeof(MarchArgs<float>); }

void ens_free(glm_ens* e)
{
    if (!e) return;
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    void* ptrs[] = {e->d_eb, e->d_ma, e->d_ou, e->d_order, e->d_tile_off, e->d_queue, e->d_sweep, e->d_log,
                    e->d_t0, e->d_t1};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (OUBox& o : e->ou) ou_box_free(o);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

// z planes per tile of a box under the configured zchunk (0: automatic, about 32 planes).
int box_zchunk(int nz, int zreq)
{
    int z = zreq > 0 ? zreq : 32;
    if (z > nz) z = nz;
    const int nzc = (nz + z - 1) / z;
    return (nz + nzc - 1) / nzc;
}

// The flat tile list of a configuration: position order, tile offsets, per-box tiling.
struct Layout {
    std::vector<int> order, tile_off, tiles_x, tiles_y, nzc, zchunk;
    int total = 0;
};

Layout make_layout(const glm_ens* e, int zreq, int ord)
{
    Layout L;
    const int n = e->nbox;
    L.order.resize(n);
    for (int b = 0; b < n; ++b) L.order[b] = b;
    if (ord == 1) {
        auto cells = [&](int b) { return (double)e->box[b].nx * e->box[b].ny * e->box[b].nz; };
        std::stable_sort(L.order.begin(), L.order.end(), [&](int a, int b) { return cells(a) > cells(b); });
    }
    L.tiles_x.resize(n); L.tiles_y.resize(n); L.nzc.resize(n); L.zchunk.resize(n);
    L.tile_off.resize(n + 1);
    for (int b = 0; b < n; ++b) {
        const glm_ens_box& x = e->box[b];
        march_tiling(x.nx, x.ny, x.nz, box_zchunk(x.nz, zreq), 0, L.tiles_x[b], L.tiles_y[b], L.nzc[b], L.zchunk[b]);
    }
    int off = 0;
    for (int p = 0; p < n; ++p) {
        L.tile_off[p] = off;
        const int b = L.order[p];
        off += L.tiles_x[b] * L.tiles_y[b] * L.nzc[b];
    }
    L.tile_off[n] = off;
    L.total = off;
    return L;
}

// Host image of the per-box march arguments (the step control fills the per-step fields).
template <typename T>
std::vector<MarchArgs<T>> make_args(const glm_ens* e, const Layout& L, unsigned long long* cmax_slots)
{
    std::vector<MarchArgs<T>> v(e->nbox);
    for (int b = 0; b < e->nbox; ++b) {
        const glm_ens_box& x = e->box[b];
        MarchArgs<T>& a = v[b];
        memset(&a, 0, sizeof a);
        for (int k = 0; k < 9; ++k) {
            a.q[k] = (const T*)x.u0[k];
            a.o[k] = (T*)x.u1[k];
        }
        a.nx = x.nx; a.ny = x.ny; a.nz = x.nz;
        a.tiles_x = L.tiles_x[b]; a.tiles_y = L.tiles_y[b]; a.nzc = L.nzc[b]; a.zchunk = L.zchunk[b];
        glm_params p = e->scheme;
        const double dx = x.boxlen / x.nx;
        p.dx = dx;
        p.ch = 1.0;
        a.ph = make_phys<T>(&p);
        a.smallc = (T)p.smallc;
        a.turb_min_rho = (T)p.turb_min_rho;
        a.dx = (T)dx;
        a.lx = (T)(x.nx * dx); a.ly = (T)(x.ny * dx); a.lz = (T)(x.nz * dx);
        a.recon = p.recon; a.slope = p.slope; a.mh_revert = p.mh_revert;
        a.face_guard = p.face_guard; a.update_floor = p.update_floor;
        a.do_turb = x.do_turb;
        if (x.do_turb) {
            a.afield = (const T*)e->ou[b].afield;
            a.tgs = e->ou[b].g;
        }
        a.want_cfl = 1;
        a.cmax = cmax_slots + b;
        a.cmax_next = nullptr;
    }
    return v;
}

template <typename T>
EnsLaunch<T> make_launch(const glm_ens* e, const void* ma, const int* order, const int* tile_off, int total,
                         int persistent, unsigned long long* t0, unsigned long long* t1)
{
    EnsLaunch<T> l;
    l.boxes = (const MarchArgs<T>*)ma;
    l.order = order;
    l.tile_off = tile_off;
    l.npos = e->nbox;
    l.total = total;
    l.persistent = persistent > 0 ? 1 : 0;
    l.nbox = e->nbox;
    l.queue = e->d_queue;
    l.sweep = e->d_sweep;
    l.t0 = t0;
    l.t1 = t1;
    return l;
}

int grid_of(const glm_ens* e, int total, int persistent)
{
    if (persistent <= 0) return total;
    const int g = persistent * e->cu_count;
    return g < total ? g : total;
}

int apply_config(glm_ens* e, int zreq, int persistent, int ord)
{
    Layout L = make_layout(e, zreq, ord);
    if (L.total <= 0) return set_err(GLM_ERR_ARG, "glm_ens: empty tile list");
    HIPCHK(hipStreamSynchronize(e->stream));
    // the step control rewrites the per-step fields every sweep; the rest comes from here
    if (e->prec == 8) {
        auto v = make_args<double>(e, L, nullptr);
        for (int b = 0; b < e->nbox; ++b) v[b].cmax = &e->d_eb[b].cmax;
        HIPCHK(hipMemcpy(e->d_ma, v.data(), v.size() * sizeof v[0], hipMemcpyHostToDevice));
    } else {
        auto v = make_args<float>(e, L, nullptr);
        for (int b = 0; b < e->nbox; ++b) v[b].cmax = &e->d_eb[b].cmax;
        HIPCHK(hipMemcpy(e->d_ma, v.data(), v.size() * sizeof v[0], hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemcpy(e->d_order, L.order.data(), e->nbox * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->d_tile_off, L.tile_off.data(), (e->nbox + 1) * sizeof(int), hipMemcpyHostToDevice));
    e->zchunk = zreq; e->persistent = persistent; e->order = ord;
    e->total = L.total;
    e->grid = grid_of(e, L.total, persistent);
    return GLM_OK;
}

int enqueue_sweep(glm_ens* e)
{
    const int nt = ((e->nbox + 63) / 64) * 64;
    const glm_params& p = e->scheme;
    if (e->prec == 8) {
        hipLaunchKernelGGL((ens_ctl_kernel<double>), dim3(1), dim3(nt), 0, e->stream, e->d_eb,
                           (MarchArgs<double>*)e->d_ma, e->d_ou, e->nbox, e->d_sweep, e->d_queue, e->d_log, e->d_t0,
                           e->d_t1);
    } else {
        hipLaunchKernelGGL((ens_ctl_kernel<float>), dim3(1), dim3(nt), 0, e->stream, e->d_eb,
                           (MarchArgs<float>*)e->d_ma, e->d_ou, e->nbox, e->d_sweep, e->d_queue, e->d_log, e->d_t0,
                           e->d_t1);
    }
    HIPCHK(hipGetLastError());
    if (e->any_turb) HIPCHK(launch_ou_advance(e->prec, e->d_ou, e->nbox, e->gmax, e->nbmax, -1, e->stream));
    if (e->prec == 8)
        HIPCHK(launch_ens<double>(make_launch<double>(e, e->d_ma, e->d_order, e->d_tile_off, e->total, e->persistent,
                                                      e->d_t0, e->d_t1),
                                  p.predictor, p.riemann, p.recon, p.slope, e->grid, e->stream));
    else
        HIPCHK(launch_ens<float>(make_launch<float>(e, e->d_ma, e->d_order, e->d_tile_off, e->total, e->persistent,
                                                    e->d_t0, e->d_t1),
                                 p.predictor, p.riemann, p.recon, p.slope, e->grid, e->stream));
    return GLM_OK;
}

void drop_graph(glm_ens* e)
{
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->exec = nullptr;
    e->graph = nullptr;
}

int capture_graph(glm_ens* e)
{
    drop_graph(e);
    HIPCHK(hipStreamSynchronize(e->stream));
    HIPCHK(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    int err = enqueue_sweep(e);
    hipGraph_t g = nullptr;
    hipError_t ce = hipStreamEndCapture(e->stream, &g);
    if (err) {
        if (g) (void)hipGraphDestroy(g);
        return err;
    }
    HIPCHK(ce);
    e->graph = g;
    HIPCHK(hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
    return GLM_OK;
}

}  // namespace

extern "C" {

glm_ens* glm_ens_create(int nbox, const glm_ens_box* boxes, const glm_params* scheme, int precision)
{
    if (nbox < 1 || nbox > ENS_MAX_BOXES || !boxes || (precision != 8 && precision != 4)) {
        set_err(GLM_ERR_ARG, "glm_ens_create: need 1 <= nbox <= 1024, boxes, precision 8 or 4");
        return nullptr;
    }
    if (check_params(scheme)) return nullptr;
    for (int b = 0; b < nbox; ++b) {
        const glm_ens_box& x = boxes[b];
        bool ok = x.nx >= 1 && x.ny >= 1 && x.nz >= 1 && x.boxlen > 0.0 && x.courant > 0.0 && x.ch_scale > 0.0;
        for (int v = 0; v < 9 && ok; ++v) ok = x.u0[v] && x.u1[v] && x.u0[v] != x.u1[v];
        if (ok && x.do_turb) ok = x.g >= 2 && x.g <= 256 && x.turb_T > 0.0 && x.ndt >= 1 && x.kmax >= x.kmin;
        if (!ok) {
            set_err(GLM_ERR_ARG, "glm_ens_create: bad box (grid, boxlen, courant, ch_scale, planes or forcing)");
            return nullptr;
        }
    }
    glm_ens* e = new glm_ens();
    e->nbox = nbox;
    e->prec = precision;
    e->scheme = *scheme;
    e->box.assign(boxes, boxes + nbox);
    e->ou.assign(nbox, OUBox{});
    for (OUBox& o : e->ou) memset(&o, 0, sizeof o);
    auto fail = [&](int code, const char* msg) -> glm_ens* {
        set_err(code, msg);
        ens_free(e);
        return nullptr;
    };
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(GLM_ERR_HIP, "glm_ens_create: stream creation failed");
    int dev = 0, cus = 0, khz = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess)
        return fail(GLM_ERR_HIP, "glm_ens_create: device attribute query failed");
    e->cu_count = cus;
    e->clock_khz = khz > 0 ? (double)khz : 100000.0;
    const size_t logn = (size_t)ENS_LOG * nbox;
    if (hipMalloc(&e->d_eb, nbox * sizeof(EnsBox)) != hipSuccess ||
        hipMalloc(&e->d_ma, nbox * ma_bytes(precision)) != hipSuccess ||
        hipMalloc(&e->d_ou, nbox * sizeof(OUBox)) != hipSuccess ||
        hipMalloc(&e->d_order, nbox * sizeof(int)) != hipSuccess ||
        hipMalloc(&e->d_tile_off, (nbox + 1) * sizeof(int)) != hipSuccess ||
        hipMalloc(&e->d_queue, NXCD * sizeof(unsigned int)) != hipSuccess ||
        hipMalloc(&e->d_sweep, sizeof(long long)) != hipSuccess ||
        hipMalloc(&e->d_log, logn * 8 * sizeof(double)) != hipSuccess ||
        hipMalloc(&e->d_t0, logn * sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc(&e->d_t1, logn * sizeof(unsigned long long)) != hipSuccess)
        return fail(GLM_ERR_ALLOC, "glm_ens_create: device allocation failed");
    e->gmax = 2; e->nbmax = 1;
    for (int b = 0; b < nbox; ++b) {
        const glm_ens_box& x = boxes[b];
        if (!x.do_turb) continue;
        e->any_turb = 1;
        if (ou_box_alloc(e->ou[b], x.g, x.kmin, x.kmax, x.turb_T, x.ndt, x.comp_frac, x.amp, x.seed, precision,
                         e->stream) != hipSuccess)
            return fail(GLM_ERR_ALLOC, "glm_ens_create: forcing allocation failed");
        e->gmax = std::max(e->gmax, e->ou[b].g);
        e->nbmax = std::max(e->nbmax, e->ou[b].nb);
    }
    if (apply_config(e, 0, 0, 1)) {
        ens_free(e);
        return nullptr;
    }
    if (glm_ens_reset(e)) {
        ens_free(e);
        return nullptr;
    }
    return e;
}

void glm_ens_destroy(glm_ens* e) { ens_free(e); }

void* glm_ens_stream(glm_ens* e) { return e ? (void*)e->stream : nullptr; }

int glm_ens_reset(glm_ens* e)
{
    if (!e) return set_err(GLM_ERR_ARG, "glm_ens_reset: null ensemble");
    HIPCHK(hipStreamSynchronize(e->stream));
    std::vector<EnsBox> eb(e->nbox);
    for (int b = 0; b < e->nbox; ++b) {
        const glm_ens_box& x = e->box[b];
        EnsBox& h = eb[b];
        memset(&h, 0, sizeof h);
        h.turb_T = x.turb_T > 0.0 ? x.turb_T : 1.0;   // the ramp min(t/turb_T, 1) is logged for every box
        h.dt_ou = x.do_turb ? x.turb_T / x.ndt : 0.0;   // glmmhd_turb.jl:2148
        h.courant = x.courant;
        h.dx = x.boxlen / x.nx;
        h.boxlen = x.boxlen;
        h.ch_scale = x.ch_scale;
        h.cp_coef = x.cp_coef;
        h.do_turb = x.do_turb;
        for (int v = 0; v < 9; ++v) {
            h.u[0][v] = x.u0[v];
            h.u[1][v] = x.u1[v];
        }
    }
    HIPCHK(hipMemcpy(e->d_eb, eb.data(), eb.size() * sizeof(EnsBox), hipMemcpyHostToDevice));
    for (int b = 0; b < e->nbox; ++b)
        if (e->box[b].do_turb) HIPCHK(ou_box_reset(e->ou[b], e->prec, e->stream));
    HIPCHK(hipMemcpyAsync(e->d_ou, e->ou.data(), e->nbox * sizeof(OUBox), hipMemcpyHostToDevice, e->stream));
    const long long minus1 = -1;
    HIPCHK(hipMemcpyAsync(e->d_sweep, &minus1, sizeof minus1, hipMemcpyHostToDevice, e->stream));
    HIPCHK(hipMemsetAsync(e->d_log, 0, (size_t)ENS_LOG * e->nbox * 8 * sizeof(double), e->stream));
    for (int b = 0; b < e->nbox; ++b) {
        const glm_ens_box& x = e->box[b];
        const size_t n = (size_t)x.nx * x.ny * x.nz;
        glm_params p = e->scheme;
        p.ch = 1.0;
        if (e->prec == 8) {
            Ptr9<const double> q;
            for (int v = 0; v < 9; ++v) q.p[v] = (const double*)x.u0[v];
            HIPCHK(launch_ctot<double>(q, n, make_phys<double>(&p), &e->d_eb[b].cmax, e->stream));
        } else {
            Ptr9<const float> q;
            for (int v = 0; v < 9; ++v) q.p[v] = (const float*)x.u0[v];
            HIPCHK(launch_ctot<float>(q, n, make_phys<float>(&p), &e->d_eb[b].cmax, e->stream));
        }
    }
    HIPCHK(hipStreamSynchronize(e->stream));
    return GLM_OK;
}

int glm_ens_step(glm_ens* e, int nsweeps)
{
    if (!e || nsweeps < 0) return set_err(GLM_ERR_ARG, "glm_ens_step: bad argument");
    for (int s = 0; s < nsweeps; ++s) {
        if (e->use_graph) HIPCHK(hipGraphLaunch(e->exec, e->stream));
        else {
            int err = enqueue_sweep(e);
            if (err) return err;
        }
    }
    return GLM_OK;
}

int glm_ens_graph(glm_ens* e, int on)
{
    if (!e) return set_err(GLM_ERR_ARG, "glm_ens_graph: null ensemble");
    if (!on) {
        HIPCHK(hipStreamSynchronize(e->stream));
        drop_graph(e);
        e->use_graph = 0;
        return GLM_OK;
    }
    int err = capture_graph(e);
    if (err) return err;
    e->use_graph = 1;
    return GLM_OK;
}

int glm_ens_sync(glm_ens* e)
{
    if (!e) return set_err(GLM_ERR_ARG, "glm_ens_sync: null ensemble");
    HIPCHK(hipStreamSynchronize(e->stream));
    return GLM_OK;
}

long long glm_ens_sweeps(glm_ens* e)
{
    if (!e) return -1;
    long long sw = -1;
    if (hipMemcpyAsync(&sw, e->d_sweep, sizeof sw, hipMemcpyDeviceToHost, e->stream) != hipSuccess ||
        hipStreamSynchronize(e->stream) != hipSuccess)
        return -1;
    return sw + 1;
}

int glm_ens_current(glm_ens* e, int box, void** planes)
{
    if (!e || !planes || box < 0 || box >= e->nbox) return set_err(GLM_ERR_ARG, "glm_ens_current: bad argument");
    int par = 0;
    HIPCHK(hipMemcpyAsync(&par, (char*)(e->d_eb + box) + offsetof(EnsBox, parity), sizeof par,
                          hipMemcpyDeviceToHost, e->stream));
    HIPCHK(hipStreamSynchronize(e->stream));
    for (int v = 0; v < 9; ++v) planes[v] = par ? e->box[box].u1[v] : e->box[box].u0[v];
    return GLM_OK;
}

int glm_ens_log(glm_ens* e, int box, int count, double* out)
{
    if (!e || !out || box < 0 || box >= e->nbox || count < 0) return set_err(GLM_ERR_ARG, "glm_ens_log: bad argument");
    const long long n = glm_ens_sweeps(e);
    if (count > n || count > ENS_LOG) return set_err(GLM_ERR_ARG, "glm_ens_log: count exceeds the sweeps kept");
    std::vector<double> all((size_t)ENS_LOG * e->nbox * 8);
    HIPCHK(hipMemcpy(all.data(), e->d_log, all.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int k = 0; k < count; ++k) {
        const long long sw = n - count + k;
        const size_t at = ((size_t)(sw % ENS_LOG) * e->nbox + box) * 8;
        memcpy(out + (size_t)k * 8, all.data() + at, 8 * sizeof(double));
    }
    return GLM_OK;
}

int glm_ens_timing(glm_ens* e, int count, double* box_ms)
{
    if (!e || !box_ms || count < 0) return set_err(GLM_ERR_ARG, "glm_ens_timing: bad argument");
    const long long n = glm_ens_sweeps(e);
    if (count > n || count > ENS_LOG) return set_err(GLM_ERR_ARG, "glm_ens_timing: count exceeds the sweeps kept");
    const size_t logn = (size_t)ENS_LOG * e->nbox;
    std::vector<unsigned long long> a(logn), b(logn);
    HIPCHK(hipMemcpy(a.data(), e->d_t0, logn * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(b.data(), e->d_t1, logn * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    for (int k = 0; k < count; ++k) {
        const long long sw = n - count + k;
        for (int x = 0; x < e->nbox; ++x) {
            const size_t at = (size_t)(sw % ENS_LOG) * e->nbox + x;
            box_ms[(size_t)k * e->nbox + x] = b[at] > a[at] ? (double)(b[at] - a[at]) / e->clock_khz : 0.0;
        }
    }
    return GLM_OK;
}

int glm_ens_configure(glm_ens* e, int zchunk, int persistent, int order)
{
    if (!e || zchunk < 0 || persistent < 0 || persistent > 8 || order < 0 || order > 1)
        return set_err(GLM_ERR_ARG, "glm_ens_configure: bad argument");
    int err = apply_config(e, zchunk, persistent, order);
    if (err) return err;
    if (e->use_graph) return capture_graph(e);
    return GLM_OK;
}

int glm_ens_tune(glm_ens* e, int reps, int* best, double* best_ms)
{
    if (!e || reps < 1) return set_err(GLM_ERR_ARG, "glm_ens_tune: bad argument");
    HIPCHK(hipStreamSynchronize(e->stream));
    // current buffers and CFL maxima (the march of a tuning launch reads the current state and
    // writes the other buffer; its CFL max goes to a scratch slot)
    std::vector<EnsBox> eb(e->nbox);
    HIPCHK(hipMemcpy(eb.data(), e->d_eb, eb.size() * sizeof(EnsBox), hipMemcpyDeviceToHost));
    int occ = 1;
    {
        const void* fn = e->prec == 8 ? ens_kernel_ptr<double>(e->scheme.predictor, e->scheme.riemann,
                                                              e->scheme.recon, e->scheme.slope)
                                      : ens_kernel_ptr<float>(e->scheme.predictor, e->scheme.riemann,
                                                              e->scheme.recon, e->scheme.slope);
        if (!fn || hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, NT, 0) != hipSuccess || occ < 1) occ = 1;
    }
    void* ma = nullptr;
    int *order = nullptr, *toff = nullptr;
    unsigned long long *slots = nullptr, *t0 = nullptr, *t1 = nullptr;
    const size_t logn = (size_t)ENS_LOG * e->nbox;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int err = GLM_OK;
    double bms = 1e300;
    int bz = e->zchunk, bp = e->persistent, bo = e->order;
    if (hipMalloc(&ma, e->nbox * ma_bytes(e->prec)) != hipSuccess ||
        hipMalloc(&order, e->nbox * sizeof(int)) != hipSuccess ||
        hipMalloc(&toff, (e->nbox + 1) * sizeof(int)) != hipSuccess ||
        hipMalloc(&slots, e->nbox * sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc(&t0, logn * sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc(&t1, logn * sizeof(unsigned long long)) != hipSuccess)
        err = set_err(GLM_ERR_ALLOC, "glm_ens_tune: scratch allocation failed");
    if (!err && (hipEventCreate(&ev0) != hipSuccess || hipEventCreate(&ev1) != hipSuccess))
        err = set_err(GLM_ERR_HIP, "glm_ens_tune: event creation failed");
    const int zc[] = {16, 24, 32, 48, 64, 128};
    for (int zi = 0; zi < 6 && !err; ++zi)
        for (int pm = 0; pm <= occ && !err; ++pm)
            for (int od = 0; od <= 1 && !err; ++od) {
                Layout L = make_layout(e, zc[zi], od);
                auto fill = [&](auto tag) -> int {
                    using T = decltype(tag);
                    auto v = make_args<T>(e, L, slots);
                    for (int b = 0; b < e->nbox; ++b) {
                        const EnsBox& h = eb[b];
                        const double cmax = __builtin_bit_cast(double, h.cmax);
                        const double dt = cmax > 0.0 ? h.courant * h.dx / cmax : 1e-3 * h.dx;
                        const double ch = h.courant * h.dx / dt / 3.0 * h.ch_scale;
                        for (int k = 0; k < 9; ++k) {
                            v[b].q[k] = (const T*)h.u[h.parity][k];
                            v[b].o[k] = (T*)h.u[1 - h.parity][k];
                        }
                        v[b].ph.ch = (T)ch; v[b].ph.ch2 = (T)(ch * ch); v[b].ph.half_ich = (T)(0.5 / ch);
                        v[b].dtdx = (T)(dt / h.dx); v[b].half_dtdx = (T)(0.5 * (dt / h.dx));
                        v[b].glm_fac = (T)1; v[b].ramp = (T)0; v[b].dt = (T)dt;
                    }
                    HIPCHK(hipMemcpy(ma, v.data(), v.size() * sizeof v[0], hipMemcpyHostToDevice));
                    return GLM_OK;
                };
                err = e->prec == 8 ? fill(double()) : fill(float());
                if (err) break;
                if (hipMemcpy(order, L.order.data(), e->nbox * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
                    hipMemcpy(toff, L.tile_off.data(), (e->nbox + 1) * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
                    err = set_err(GLM_ERR_HIP, "glm_ens_tune: table upload failed");
                    break;
                }
                const int grid = grid_of(e, L.total, pm);
                std::vector<float> ms;
                for (int r = 0; r <= reps && !err; ++r) {   // one untimed warm-up launch
                    // the persistent queues are cleared by the step control in a sweep; here by a kernel
                    if (launch_clear_u64((unsigned long long*)e->d_queue, NXCD / 2, e->stream) != hipSuccess ||
                        hipEventRecord(ev0, e->stream) != hipSuccess) {
                        err = set_err(GLM_ERR_HIP, "glm_ens_tune: launch setup failed");
                        break;
                    }
                    hipError_t le = e->prec == 8
                        ? launch_ens<double>(make_launch<double>(e, ma, order, toff, L.total, pm, t0, t1), e->scheme.predictor,
                                             e->scheme.riemann, e->scheme.recon, e->scheme.slope, grid, e->stream)
                        : launch_ens<float>(make_launch<float>(e, ma, order, toff, L.total, pm, t0, t1), e->scheme.predictor,
                                            e->scheme.riemann, e->scheme.recon, e->scheme.slope, grid, e->stream);
                    if (le != hipSuccess || hipEventRecord(ev1, e->stream) != hipSuccess ||
                        hipEventSynchronize(ev1) != hipSuccess) {
                        err = hip_err(le != hipSuccess ? le : hipErrorUnknown, "glm_ens_tune: launch");
                        break;
                    }
                    float t = 0.f;
                    (void)hipEventElapsedTime(&t, ev0, ev1);
                    if (r > 0) ms.push_back(t);
                }
                if (err) break;
                std::sort(ms.begin(), ms.end());
                const double med = ms[ms.size() / 2];
                if (med < bms) { bms = med; bz = zc[zi]; bp = pm; bo = od; }
            }
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    void* ptrs[] = {ma, order, toff, slots, t0, t1};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (err) return err;
    err = glm_ens_configure(e, bz, bp, bo);
    if (err) return err;
    if (best) { best[0] = bz; best[1] = bp; best[2] = bo; }
    if (best_ms) *best_ms = bms;
    return GLM_OK;
}

}  // extern "C"
