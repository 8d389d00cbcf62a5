/*
 * glmmhd.h -- C ABI of libglmmhd.so, the gfx950 (MI355X) GLM-MHD line-march library.
 *
 * Plain C: pointers, sizes and scalars only.  Device pointers are HIP device addresses
 * (what AMDGPU.jl's pointer(::ROCArray), torch.Tensor.data_ptr() or hipMalloc return).
 *
 * State layout (all entry points): SoA planes, one contiguous plane per variable, cell (i,j,k)
 * at i + nx*(j + ny*k) -- march_bridge/cu/spike_mhd.cu:48, MarchBridge.jl:64-81, and the memory
 * layout of the Julia (N,N,N,9) array (glmmhd_turb.jl:2051).
 *   GLM: 9 planes (rho, rho*vx, rho*vy, rho*vz, E, Bx, By, Bz, psi)       glmmhd_turb.jl:61-64
 *   CT : 8 planes (rho, mx, my, mz, E, bxf, byf, bzf), bxf[i] = Bx on face i-1/2
 *                                                                           ct_mhd.jl:10-11
 * Paths in citations are relative to the reference repository root.
 */
#ifndef GLMMHD_H
#define GLMMHD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ===================================================================================== */
/* 1. Legacy MarchBridge ABI (single precision state, library-global parameters).         */
/*    Same symbols and C signatures as march_bridge/cu/spike_mhd.cu:319-341, so           */
/*    MarchBridge.open_lib/set_dtdx/set_glm/set_gamma/run! (MarchBridge.jl:34-81) bind     */
/*    unchanged.  The scheme is the spike march's: normal-only Hancock, PLM MonCen, HLL,   */
/*    GLM cleaning and psi damping, smallr = smallp = 1e-6 (spike_mhd.cu:52, :257-288).    */
/* ===================================================================================== */

/* spike_mhd.cu:321  number of SoA planes: 9 (GLM library), 11 (its -DNSP=2 build, spike_mhd.cu:30-33),
 * 8 (CT library, spike_ctm.cu:301), 10 (its -DNSP=2 build), 5 (hydro library, spike_25d.cu:560) or 7
 * (its -DSCALARS build, spike_25d.cu:115-119) */
int march_nv(void);
/* spike_mhd.cu:322  grid size N (compile-time NX in the reference; runtime here, see march_set_grid) */
int march_nx(void);
/* spike_mhd.cu:324  passive species count (2 in the 11-plane GLM, the 10-plane CT and the 7-plane hydro
 * libraries, else 0) */
int march_nsp(void);
/* spike_mhd.cu:324  fixed dt/dx of subsequent march_run_dev / ct_run calls */
void march_set_dtdx(float v);
/* spike_mhd.cu:325  Dedner cleaning speed ch and per-step psi decay factor */
void march_set_glm(float ch, float fac);
/* spike_mhd.cu:326  adiabatic index (default 1.4, spike_mhd.cu:52) */
void march_set_gamma(float g);
/* spike_mhd.cu:328-339  run nsteps periodic marches over device planes qv[0..8]; ov is scratch of
 * the same shape; the result is always left in qv.  Returns elapsed GPU time in ms (HIP events).
 * In a 5-plane library (build_march.sh) this is spike_25d.cu:563-575: the hydro march over
 * qv[0..4] = [rho, mx, my, mz, E] (normal-only Hancock, PLM MonCen, HLL, smallr = smallp = 1e-6,
 * default gamma 1.4 and dt/dx 0.05, spike_25d.cu:180). */
double march_run_dev(float* const* qv, float* const* ov, int nsteps);

/* No reference counterpart (the reference fixes NX/NY/NZ at compile time, build_mhd.sh:26-28):
 * set the runtime grid used by march_run_dev / march_run_dev_f64 / ct_run.  Returns 0 or an
 * error code. */
int march_set_grid(int nx, int ny, int nz);
/* FP64 twin of march_run_dev (the FP64 state path; no reference counterpart). */
double march_run_dev_f64(double* const* qv, double* const* ov, int nsteps);

/* spike_ctm.cu:305-325 CT march: qv/ov = 8 planes [U0..U4, bxf, byf, bzf], then in the 10-plane
 * library the 2 species planes rho X_s (spike_ctm.cu:304, :310-311); cc = 3 Bcc scratch planes
 * (accepted for ABI compatibility; the cell-centred B is formed in-kernel).  Uses the
 * march_set_dtdx / march_set_gamma globals.  Returns elapsed GPU time in ms. */
double ct_run(float* const* qv, float* const* ov, float* const* cc, int nsteps);
/* spike_ctm.cu:326  registers of the CT march kernel (gfx950 arch VGPRs + AGPRs) */
int ct_regs(void);
/* spike_ctm.cu:327  LDS bytes per workgroup of the CT march kernel */
int ct_shmem(void);

/* ===================================================================================== */
/* 2. Context API (re-entrant, per-context HIP stream, FP64 or FP32 state).               */
/*    Mirrors the Julia step API: step_plm! (glmmhd_turb.jl:878-908), compute_dt          */
/*    (:1973-1978), diagnostics (:2006-2018), the ICs (:2021-2036, :2164-2172) and the OU  */
/*    forcing (Forcing / ou_advance!, :2117-2161).                                         */
/* ===================================================================================== */

typedef struct glm_ctx glm_ctx;

/* scheme switches (values of glm_params fields) */
enum { GLM_PRED_FULL = 0, GLM_PRED_ONED = 1 };                                  /* :165-172 | spike_mhd.cu:149-161 */
enum { GLM_RECON_PLM = 0, GLM_RECON_PPM = 1, GLM_RECON_PPM_SPIKE = 2 };       /* :917-926 | spike_mhd.cu:162-192 */
enum { GLM_SLOPE_MONCEN = 0, GLM_SLOPE_MINMOD = 1 };                           /* :145-155 | :300-303 */
enum { GLM_RIEMANN_LLF = 0, GLM_RIEMANN_LLF_BOUND = 1, GLM_RIEMANN_HLL = 2, GLM_RIEMANN_HLLD = 3 };
enum { GLM_GUARD_NONE = 0, GLM_GUARD_IFLO = 1, GLM_GUARD_CLAMP = 2 };          /* cpu_simd.jl | :618 | spike_mhd.cu:145 */
enum { GLM_IC_ORSZAG_TANG = 0, GLM_IC_UNIFORM = 1, GLM_IC_GRADIENT = 2 };

/* One step's scalars and switches; names follow Params (glmmhd_turb.jl:28-55). */
typedef struct glm_params {
    double gamma, smallr, smallc;
    double pfl;                 /* pressure floor; <= 0 selects smallr*smallc^2/gamma (:55) */
    double dt, dx;              /* time step and (cubic) cell size; box length per axis = n*dx */
    double ch, glm_fac;         /* Dedner speed and psi decay factor (:884-885) */
    double llf_dmin, llf_pmin;  /* HLLD -> LLF switch thresholds, <= 0 disables (:287-290) */
    double turb_min_rho, ramp;  /* OU kick density floor and min(t/turb_T, 1) ramp (:886) */
    int predictor, recon, slope, riemann;
    int mh_revert;              /* predictor revert mh -> m0 (:393-395) */
    int face_guard;             /* GLM_GUARD_* */
    int update_floor;           /* static void-cell floor in the update (:448-453) */
    int do_turb;                /* fused OU kick with the context's forcing field (:456-475) */
    int zchunk;                 /* z planes per workgroup (0 = automatic) */
    int want_cfl;               /* fold the next step's CFL max into the step (read with glm_cfl_max) */
} glm_params;

/* Fill a glm_params with the reference defaults (Params, glmmhd_turb.jl:28-52) and the cube
 * scheme (full Hancock, PLM MonCen, HLLD, mh revert, iflo faces, update floor). */
void glm_params_default(glm_params* p);

/* Create a context for an nx*ny*nz grid; precision 8 (FP64 state) or 4 (FP32 state).
 * Returns NULL on failure (see glm_last_error). */
glm_ctx* glm_create(int nx, int ny, int nz, int precision);
void glm_destroy(glm_ctx* ctx);
/* Use a caller-owned HIP stream (hipStream_t passed as void*); NULL = the null stream. */
int glm_set_stream(glm_ctx* ctx, void* stream);

/* One step q -> o (step_plm!, glmmhd_turb.jl:878-908), asynchronous on the context stream.
 * q, o: 9 device plane pointers each (distinct buffers).  Returns 0 or an error code. */
int glm_step(glm_ctx* ctx, void* const* q, void* const* o, const glm_params* p);
/* glm_step with nsp passive species (spike_mhd.cu -DNSP, :55-61, :128-188, :218-220, :278-296):
 * q, o: 9 + nsp device planes, the 9 GLM-MHD planes followed by the conserved rho*X_s.  The
 * fractions X_s = rho*X_s / max(rho, smallr) are reconstructed and predicted like the primitives
 * (advected: X - dt/dx u_n dX_n for the one-dimensional predictor, summed over the three
 * directions for the full one), and each
 * face carries F_s = F_rho * X_s of the upwind side of F_rho.  The MHD planes come out bit-identical
 * to glm_step's.  nsp: 0 (same as glm_step) or 2. */
int glm_step_species(glm_ctx* ctx, void* const* q, void* const* o, const glm_params* p, int nsp);
/* The memory floor of glm_step (spike_mhd.cu -DMEMFLOOR, :260-299: the march's loads and stores
 * with the flux computation compiled out): the same launch, tiling, z-chunking, plane loads,
 * primitive conversion, LDS staging, update-base read and stores as glm_step, with no predictor,
 * faces, fluxes or epilogue, so o receives q unchanged (9 planes; p as for glm_step, of which the
 * grid chunking is used).  For measurement: the time of the march's data movement alone. */
int glm_march_floor(glm_ctx* ctx, void* const* q, void* const* o, const glm_params* p);
/* Max over cells of sum_d(|v_d| + c_f) of the state written by the last glm_step that had
 * want_cfl set (synchronises the stream).  dt = courant*dx/cmax (compute_dt, :1973-1978). */
int glm_cfl_max(glm_ctx* ctx, double* cmax);
/* Standalone CFL max of a state (ctot_kernel! + maximum, :1963-1976); synchronous. */
int glm_ctot_max(glm_ctx* ctx, void* const* q, const glm_params* p, double* cmax);
/* Max over cells of max_d c_f(B_d), the largest fast magnetosonic speed in a grid direction: the
 * "ch = max fast speed" Dedner cleaning speed (the reference's run_ot / glm_ch_scale rule at
 * glmmhd_turb.jl:884 is the other choice).  Synchronous. */
int glm_fast_max(glm_ctx* ctx, void* const* q, const glm_params* p, double* cmax);

#define GLM_NDIAG 10
/* Diagnostics (:1996-2018 + energy budget), FP64 with compensated sums; synchronous.
 * out: mass, sum m^2/max(rho,smallr), max|divB| (central), rho_min, finite, E_kin, E_mag,
 *      sum m.B/rho (cross helicity), sum E, sum psi^2   (plain cell sums; times dV for integrals) */
int glm_diagnostics(glm_ctx* ctx, void* const* q, const glm_params* p, double* out);

/* No reference counterpart.  For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): glm_diagnostics without a
 * context and without any allocation.  q: 9 device planes of an nx*ny*nz grid in the given
 * precision (8 or 4); out: 10 doubles in DEVICE memory, written in stream order on `stream`
 * (hipStream_t as void*, NULL = the null stream), same values and order as glm_diagnostics.
 * Uses p->dx and p->smallr.  Asynchronous: no host synchronisation.  The reduction's partials live
 * in a static device array of about 1 MiB (per GPU, held until the process ends), which limits the
 * grid to 2^27 cells (512^3); larger grids are rejected with GLM_ERR_ARG.  Every call on one GPU
 * shares that array: calls that overlap on one GPU (two streams at once) race. */
int glm_diagnostics_noalloc(void* const* q, int nx, int ny, int nz, int precision, const glm_params* p, void* stream,
                            double* out);

/* Initial conditions: GLM_IC_ORSZAG_TANG args {gamma} (:2021-2036); GLM_IC_UNIFORM args
 * {rho0, p0, b0x, b0y, b0z, gamma} (:2164-2172); GLM_IC_GRADIENT args {gamma}
 * (spike_mhd.cu:304-317).  Asynchronous. */
int glm_ic(glm_ctx* ctx, void* const* q, int kind, const double* args);

/* OU forcing (Forcing/ou_advance!, :2117-2161) on a g^3 Fourier grid with a shell
 * kmin <= |k| <= kmax, sqrt(P) = 1/|k|, compressive fraction comp_frac, correlation time T,
 * ndt OU substeps per T, amplitude amp, seeded Philox4x32-10 increments. */
int glm_ou_setup(glm_ctx* ctx, int g, double kmin, double kmax, double T, int ndt, double comp_frac, double amp,
                 unsigned long long seed);
/* One OU step of the Fourier field and a refresh of the real acceleration field (ou_advance!). */
int glm_ou_advance(glm_ctx* ctx);
/* Device pointer of the acceleration field (3 SoA planes of g^3, state precision). */
const void* glm_ou_afield(glm_ctx* ctx);
/* Copy the acceleration field (3*g^3 values, state precision) into the caller's device buffer dst,
 * in stream order on the context stream. */
int glm_ou_copy_afield(glm_ctx* ctx, void* dst);
/* Number of OU steps taken so far. */
long long glm_ou_steps(glm_ctx* ctx);

/* CT step as one fused z-march kernel (spike_ctm.cu:250-298 ct_march; physics of ct_mhd.jl
 * step!, :174-184): q, o = 8 planes.  Uses gamma, smallr, pfl (as SMALLP), dt, dx, riemann (HLL
 * or HLLD or LLF), recon (PLM or PPM_SPIKE), update_floor, do_turb, turb_min_rho, smallc, ramp,
 * want_cfl from p; ct_face_slope selects the spike_ctm.cu (0) or the ct_mhd.jl (1) face offset. */
int glm_ct_step(glm_ctx* ctx, void* const* q, void* const* o, const glm_params* p, int ct_face_slope);
/* glm_ct_step with nsp passive species (spike_ctm.cu -DNSP, :69-70, :80-119, :178-181, :240-243,
 * :289-292): q, o = 8 + nsp planes, the CT planes followed by the conserved rho*X_s.  The fractions
 * X_s = rho*X_s / max(rho, smallr) get the primitives' slopes and PPM edges, the normal-only half
 * step advects them (X - dt/dx/2 u_n s), and each face carries F_s = F_rho * X_s of the upwind side
 * of F_rho.  Where the update floor fires, rho*X_s = smallr * X_s(old).  The 8 CT planes come out
 * bit-identical to glm_ct_step's.  nsp: 0 (same as glm_ct_step) or 2. */
int glm_ct_step_species(glm_ctx* ctx, void* const* q, void* const* o, const glm_params* p, int ct_face_slope,
                        int nsp);
/* The same CT step as the kernel sequence of ct_mhd.jl step! (:174-184): cell-centred B and
 * primitives, face fluxes, edge EMFs, update, each its own pass over a 35-plane scratch.  Same
 * arguments and results (to rounding) as glm_ct_step. */
int glm_ct_step_passes(glm_ctx* ctx, void* const* q, void* const* o, const glm_params* p, int ct_face_slope);
/* The same CT step as one kernel over staged 3-D tiles (spike_ct2.cu:147-210 ct_step2 and
 * spike_ct3.cu:148-190 ct_step3 design): an 8 x 8 x 4 (FP32) or 8 x 8 x 2 (FP64) tile of owned cells
 * per workgroup, its primitives and every face flux it needs staged in LDS, then the hydro update
 * and the face-B curl from edge EMFs.  Same arguments and results (to rounding) as glm_ct_step. */
int glm_ct_step_tiles(glm_ctx* ctx, void* const* q, void* const* o, const glm_params* p, int ct_face_slope);
/* CFL max of a CT state (cell-centred B); synchronous. */
int glm_ct_ctot_max(glm_ctx* ctx, void* const* q, const glm_params* p, double* cmax);
/* max |div B| of a CT state from the face fields (ct_mhd.jl:191-194); synchronous. */
int glm_ct_divb_max(glm_ctx* ctx, void* const* q, double dx, double* out);

/* The 5-variable hydro line march (spike_25d.cu:383-516, the march kernel as build_march.sh
 * builds it with -DHANCOCK1D; -DLLF_FALLBACK when llf_dmin or llf_pmin > 0): q, o = 5 planes
 * [rho, mx, my, mz, E] (distinct buffers).  Uses gamma, dt/dx, smallr (SMALLR), pfl (SMALLP;
 * <= 0 selects smallr*smallc^2/gamma), llf_dmin/llf_pmin (the fallback's density and pressure
 * floors, 1e-3 in spike_25d.cu:339) and zchunk from p; recon = GLM_RECON_PPM_SPIKE selects the
 * parabolic-edge build (-DPPM1D, :467-483: the monotonized 3-point parabola's edges plus the half
 * step's change, unclamped, HLL without the fallback); the other fields are ignored.
 * Asynchronous on the context stream.  Returns 0 or an error code. */
int glm_hydro_step(glm_ctx* ctx, void* const* q, void* const* o, const glm_params* p);
/* The same march with nsp = 2 passive scalars (spike_25d.cu -DSCALARS, :115-119): q, o = 7 planes
 * [rho, mx, my, mz, E, rho*s0, rho*s1].  cma != 0 makes the scalar flux the mass flux times the
 * upwind face value of s (-DCMA, consistent multi-fluid advection, :321, :367); cma = 0 solves
 * rho*s with the same HLL/LLF as the other variables.  nsp = 0 is glm_hydro_step. */
int glm_hydro_step_scalars(glm_ctx* ctx, void* const* q, void* const* o, const glm_params* p, int nsp, int cma);
/* The donor-cell build (-DDONOR, spike_25d.cu:438-454): glm_hydro_step_scalars with the faces on the
 * seams of the 32 x 8 column tiles solved first order, HLL of the two cell states (x faces at
 * i - 1/2 with i mod 32 = 0, y faces at j - 1/2 with j mod 8 = 0, the periodic face after the last
 * cell included; z has none), the other faces as the light march; HLL without the fallback.
 * Exactly conservative.  nsp = 0 or 2 (cma as in glm_hydro_step_scalars); recon is ignored. */
int glm_hydro_step_donor(glm_ctx* ctx, void* const* q, void* const* o, const glm_params* p, int nsp, int cma);
/* The same march on the compact storage formats of spike_25d.cu's standalone builds (:64-95; the
 * reference's bridge library takes FP32 conserved planes only, so this entry has no counterpart in
 * its C ABI).  FP32 contexts only; computation in FP32.
 *   GLM_HYDRO_F16_PRIM (-DF16STATE): q, o = 5 planes of IEEE f16 bits [rho, vx, vy, vz, p];
 *   GLM_HYDRO_U16_PRIM (-DU16STATE): 5 uint16 planes, rho and p as log2 over [-32, 32], the
 *     velocities linear over [-8, 8] (:157-168);
 *   GLM_HYDRO_U16_LOG2_SPECIES (-DU16SP): 5 float conserved planes, then 2 uint16 planes of log2 X
 *     over [-110, 0]; the species ride the mass flux (CMA) and are reconstructed in log2 X.
 * Each stores its result in its own format (rounded to nearest).  Where a face is supersonic the
 * reference's -DU16SP build returns the one-sided flux with the mass flux times log2 X; here the
 * species flux is the mass flux times X on every face. */
enum { GLM_HYDRO_F16_PRIM = 1, GLM_HYDRO_U16_PRIM = 2, GLM_HYDRO_U16_LOG2_SPECIES = 3 };
int glm_hydro_step_compact(glm_ctx* ctx, void* const* q, void* const* o, const glm_params* p, int format);

/* ===================================================================================== */
/* 4. Adjoint (no reference counterpart: the reference computes no derivatives).             */
/*    The reverse of one glm_step for the scheme of its glm_params, FP64 contexts only:       */
/*    lam0 = (dU1/dU0)^T lam1 with the step's dt, ch, psi decay, ramp and forcing field held  */
/*    fixed.  Supported: full / normal-only predictor, PLM or PPM (Julia form), MonCen or     */
/*    minmod, LLF / LLF-bound / HLL / HLLD (+ LLF switch), every face guard, update floor,    */
/*    psi damping and the OU kick.  Checkpointing lives in the caller.                       */
/* ===================================================================================== */
/* u0: the 9 planes the forward step started from; lam1: adjoint of its result; lam0: output
 * (must not alias u0 or lam1); afield: the step's forcing field (3 g^3 doubles, as
 * glm_ou_copy_afield returns it) when p->do_turb, else NULL.  Also adds the step's
 * sum over cells of (kick acceleration adjoint) . (kick acceleration) to the context's
 * accumulator (glm_adjoint_force_dot): the derivative with respect to a factor on the forcing
 * amplitude.  Asynchronous on the context stream. */
int glm_step_adjoint(glm_ctx* ctx, void* const* u0, void* const* lam1, void* const* lam0, const glm_params* p,
                     const void* afield);
/* The same reverse step with its face-state and predictor adjoints and its gather as separate
 * passes over contribution planes instead of one z-march kernel: same arguments, same results. */
int glm_step_adjoint_passes(glm_ctx* ctx, void* const* u0, void* const* lam1, void* const* lam0,
                            const glm_params* p, const void* afield);
/* Read (and with reset != 0 clear) the forcing-amplitude accumulator; synchronous. */
int glm_adjoint_force_dot(glm_ctx* ctx, int reset, double* out);
/* Replace the context's forcing field by src (3 g^3 values in the state precision, device),
 * e.g. to recompute forward steps from a checkpoint with the field they used. */
int glm_ou_set_afield(glm_ctx* ctx, const void* src);

/* ===================================================================================== */
/* 3. Ensemble of boxes (no reference counterpart: glmmhd_turb.jl runs one box per run).   */
/*    Boxes of uneven size advance together.  A sweep is one step of every box, each with   */
/*    its own dt from its own CFL max and its own OU forcing -- the run_turb loop           */
/*    (glmmhd_turb.jl:2191-2241) evaluated on the device -- and one march launch over all   */
/*    boxes' tiles.  A sweep holds no host round trip, so it can be replayed as a HIP graph. */
/*    All boxes share one scheme (glm_params) and one precision.                            */
/* ===================================================================================== */
typedef struct glm_ens glm_ens;

typedef struct glm_ens_box {
    int nx, ny, nz;
    int do_turb;                /* OU kick with this box's forcing (below) */
    void* u0[9];                /* state planes holding the initial state */
    void* u1[9];                /* second state buffer of the same size */
    double boxlen;              /* box length along x: dx = boxlen / nx (dxof, :54) */
    double courant;             /* dt = courant*dx/cmax, clipped to the next OU time (:2222-2226) */
    double ch_scale, cp_coef;   /* ch = courant*dx/dt/3*ch_scale; psi *= exp(-(ch^2/(cp*boxlen*ch))*dt) (:884-885) */
    int g, ndt;                 /* OU forcing as glm_ou_setup: Fourier grid g, ndt OU steps per turb_T */
    double kmin, kmax, turb_T, comp_frac, amp;
    unsigned long long seed;
} glm_ens_box;

/* Create an ensemble of nbox (1..1024) boxes; precision 8 or 4.  The library keeps pointers to
 * the caller's state buffers.  Returns NULL on failure (see glm_last_error). */
glm_ens* glm_ens_create(int nbox, const glm_ens_box* boxes, const glm_params* scheme, int precision);
void glm_ens_destroy(glm_ens* e);
/* The ensemble's own HIP stream (hipStream_t as void*); all ensemble work runs on it. */
void* glm_ens_stream(glm_ens* e);
/* Start (or restart) every box at t = 0 from its u0: OU state zero, CFL max of u0 (ctot). */
int glm_ens_reset(glm_ens* e);
/* Enqueue nsweeps sweeps (asynchronous). */
int glm_ens_step(glm_ens* e, int nsweeps);
/* on = 1: capture one sweep into a HIP graph and replay it in glm_ens_step; 0: eager launches. */
int glm_ens_graph(glm_ens* e, int on);
int glm_ens_sync(glm_ens* e);
/* Sweeps done since the last reset (synchronous). */
long long glm_ens_sweeps(glm_ens* e);
/* The 9 device planes holding the box's current state (synchronous). */
int glm_ens_current(glm_ens* e, int box, void** planes);
#define GLM_ENS_NLOG 8
#define GLM_ENS_LOG_LEN 256
/* The box's step log of the last `count` sweeps (count <= min(sweeps, GLM_ENS_LOG_LEN)), oldest
 * first, GLM_ENS_NLOG values per sweep: sweep index, t at the start of the step, dt, ch,
 * glm_fac, ramp, the CFL max dt came from, OU advance in this step (0/1).  Synchronous. */
int glm_ens_log(glm_ens* e, int box, int count, double* out);
/* Per sweep and box, the time from the box's first tile start to its last tile end in the march
 * launch, in ms (device wall clock): count x nbox values, oldest sweep first.  Synchronous. */
int glm_ens_timing(glm_ens* e, int count, double* box_ms);
/* Launch configuration: zchunk = z planes per tile (0: automatic), persistent = 0 (one workgroup
 * per tile) or k >= 1 (k workgroups per CU take tiles from per-XCD queues), order = 0 (boxes
 * as given) or 1 (largest box first). */
int glm_ens_configure(glm_ens* e, int zchunk, int persistent, int order);
/* Autotuner: times `reps` march launches over the current states for each candidate
 * configuration, keeps the fastest and returns it in best[3] = {zchunk, persistent, order} with
 * its median launch time.  The states and the step log do not change (the second buffers are
 * used as scratch).  Synchronous. */
int glm_ens_tune(glm_ens* e, int reps, int* best, double* best_ms);

/* Error handling (the reference has none: spike_mhd.cu's AS_LIB path never checks errors). */
int glm_last_error(void);
const char* glm_error_string(int code);

/* Resource figures of the compiled march kernel for the given precision and scheme. */
int glm_kernel_info(int precision, int predictor, int riemann, int recon, int slope, int* vgprs, int* lds_bytes);

#define GLM_OK 0
#define GLM_ERR_ARG 1
#define GLM_ERR_HIP 2
#define GLM_ERR_ALLOC 3
#define GLM_ERR_STATE 4

#ifdef __cplusplus
}
#endif
#endif /* GLMMHD_H */
