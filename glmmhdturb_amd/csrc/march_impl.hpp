// march_impl.hpp -- fused GLM-MHD line-march step for gfx950 (one kernel per step).
//
// What it computes is the reference's one-step update  U1 = U0 + dt/dx * sum_d (F_lo - F_hi)
// with the positivity floor, GLM psi damping and the fused OU kick (glmmhd_turb.jl:436-480,
// :841-861), faces from PLM/PPM reconstruction with a MUSCL-Hancock predictor (full transverse,
// glmmhd_turb.jl:165-172, :378-403, :931-1014; or normal-only, spike_mhd.cu:149-161, :257-285),
// and LLF/HLL/HLLD Riemann fluxes with Dedner GLM cleaning (glmmhd_turb.jl:175-354).
//
// How it is laid out on the GPU:
//   * A workgroup of TX x TY = 18 x 14 threads owns a 16 x 12 column tile and a z-chunk (16
//     doubles = one 128-byte line per row, so row stores are whole lines).  Each thread owns one
//     (x,y) column of the tile grown by one halo cell and marches z.  Its column's primitives for
//     planes q-1 and q+1 live in registers; plane q is read back from LDS where it is needed.
//   * LDS holds the primitives of the current plane q over the tile grown by two cells
//     (20 x 16): the in-plane neighbours for slopes, PPM edges and the transverse predictor.
//   * Workgroup ids are renumbered so that each XCD (the hardware deals ids round-robin over the
//     8 XCDs) receives a contiguous run of tiles: neighbouring tiles then share their halo
//     reads in the same L2.
//   * Every face is solved once.  The z face q-1/2 is solved by the column owner from the z+
//     state it carried from plane q-1; x and y faces are solved after the +x/+y face states are
//     exchanged through LDS, and the fluxes are exchanged back the same way.
//   * The update of plane q-1 happens in the same pass, so face states and fluxes never touch
//     HBM: one read and one write of the 9 conserved planes per cell and step (plus the halo).
//     The update base U(q-1) is read again at the update; it was loaded two planes earlier by the
//     same workgroup and is normally still in L2.
//   * Budget: two workgroups per CU, i.e. two wavefronts per SIMD, so that one workgroup's LDS
//     waits, barriers and dependent FP64 chains are covered by the other's issue.  That takes
//     <= 80 KB of LDS (the three exchange planes and the forcing cache) and <= 256 registers per
//     lane: the next planes are prefetched only after the plane's Riemann solves, and plane q and
//     the predicted state are parked in LDS across the phases that do not use them.
//   * The epilogue also folds the next step's CFL signal speed (glmmhd_turb.jl:1963-1971) into a
//     per-thread max, reduced per wavefront and merged with one atomic per wavefront.
#pragma once
#include "force.hpp"
#include "gpu_util.hpp"
#include "physics.hpp"

namespace glm {

// Owned tile OX x OY (16 doubles = one 128-byte line per row).  Lanes 0..NO-1 own the tile row by
// row (4 rows per wavefront, so each row is stored by one instruction as a whole line); lanes
// NO..NT-1 are the tile's one-cell halo without its corners and form the last wavefront, which
// therefore skips the z solve and the update as a whole.
#ifndef GLM_OX
#define GLM_OX 16
#endif
#ifndef GLM_OY
#define GLM_OY 12
#endif
constexpr int OX = GLM_OX, OY = GLM_OY;
constexpr int NO = OX * OY;                 // owned lanes
constexpr int NH = 2 * OX + 2 * OY;         // halo lanes (one-cell ring without corners)
constexpr int NT = NO + NH;                 // threads per workgroup
constexpr int PX = OX + 4, PY = OY + 4;     // LDS primitive plane (two-cell ring)
constexpr int NPL = PX * PY;
constexpr int NHELP = 4 + NH;               // plane cells no lane owns: the 4 one-cell corners and the
                                            // two-cell ring without corners (in-plane neighbours of halo lanes)
static_assert(NT <= 256 && NT >= NHELP && NO % 64 == 0, "tile shape");

// Minimum resident wavefronts per SIMD requested from the register allocator.
#ifndef GLM_WAVES_F64
#define GLM_WAVES_F64 2
#endif
#ifndef GLM_WAVES_F32
#define GLM_WAVES_F32 2
#endif
#define GLM_WAVES_PER_EU(T) (sizeof(T) == 8 ? GLM_WAVES_F64 : GLM_WAVES_F32)
// With species the FP64 workgroup's 97 KB of LDS leaves one workgroup per CU: 512 registers per lane.
#define GLM_WAVES_PER_EU_NS(T, NS) ((NS) > 0 && sizeof(T) == 8 ? 1 : GLM_WAVES_PER_EU(T))
// FP32 with species: the normal-only builds with PLM or the spike PPM fit three wavefronts per SIMD
// (168 registers, no spills; the spike-PPM + HLL build took 169 at the default and ran at two:
// 12.3 -> 14.7 Gcell/s at 480^3 with three); the full-predictor and Julia-PPM builds would spill
// 16-44 B per lane there and keep two.  FP32 without species: four wavefronts per SIMD (40.75 KB
// of LDS per workgroup, four fit in 160 KB), held to 128 registers with 12-48 B of spills, beat
// three at ~155 registers: the full-predictor HLLD + PPM march 1.32 -> 1.27 ms at 256^3, the
// normal-only PLM + HLL 19.9 -> 21.3 Gcell/s at 480^3 (the spike PPM unchanged).
#ifndef GLM_WAVES_F32_NS0
#define GLM_WAVES_F32_NS0 4
#endif
#define GLM_WAVES_PER_EU_SCHEME(T, NS, PRED, RECON)                                              \
    ((NS) > 0 && sizeof(T) == 4 && (PRED) == PRED_ONED && (RECON) != RECON_PPM ? 3               \
     : (NS) == 0 && sizeof(T) == 4 ? GLM_WAVES_F32_NS0 : GLM_WAVES_PER_EU_NS(T, NS))

// Planes of a state: the 9 GLM-MHD planes and up to two passive species (rho X_s) after them.
constexpr int NVMAX = 11;

template <typename T>
struct MarchArgs {
    const T* q[NVMAX];
    T* o[NVMAX];
    int nx, ny, nz;
    int tiles_x, tiles_y, nzc, zchunk;
    Phys<T> ph;
    T dtdx, half_dtdx, glm_fac;
    T smallc, turb_min_rho, ramp, dt, dx, lx, ly, lz;
    int recon, slope, mh_revert, face_guard, update_floor;
    int do_turb, tgs, want_cfl;
    const T* afield;
    unsigned long long* cmax;       // this step's max CFL speed, as the bits of a non-negative double
    unsigned long long* cmax_next;  // the slot the next step accumulates into (zeroed here)
};

// Arguments used only at the loads, stores and the update epilogue.  They are staged in LDS at
// kernel entry so they do not hold scalar registers across the march loop.
template <typename T>
struct ColdArgs {
    const T* q[NVMAX];
    T* o[NVMAX];
    const T* afield;
    T dtdx, glm_fac, smallc, turb_min_rho, ramp, dt, dx, lx, ly, lz;
    int update_floor, do_turb, tgs, pad_;
};

// Cell (tx, ty) of lane t in the tile grown by one cell: tx in 0..OX+1, ty in 0..OY+1.
GLM_DEV void lane_cell(int t, int& tx, int& ty)
{
    if (t < NO) { tx = 1 + t % OX; ty = 1 + t / OX; return; }
    const int h = t - NO;
    if (h < OX) { tx = 1 + h; ty = 0; }
    else if (h < 2 * OX) { tx = 1 + h - OX; ty = OY + 1; }
    else if (h < 2 * OX + OY) { tx = 0; ty = 1 + h - 2 * OX; }
    else { tx = OX + 1; ty = 1 + h - 2 * OX - OY; }
}

// LDS plane position (px, py) of helper cell h (0 <= h < NHELP).
GLM_DEV void helper_cell(int h, int& px, int& py)
{
    if (h < 4) { px = (h & 1) ? PX - 2 : 1; py = (h & 2) ? PY - 2 : 1; return; }
    const int r = h - 4;
    if (r < OX) { px = 2 + r; py = 0; }
    else if (r < 2 * OX) { px = 2 + r - OX; py = PY - 1; }
    else if (r < 2 * OX + OY) { px = 0; py = 2 + r - 2 * OX; }
    else { px = PX - 1; py = 2 + r - 2 * OX - OY; }
}

template <int NV, typename T>
GLM_DEV void loadv(const ColdArgs<T>& a, size_t off, T c[NV])
{
#pragma unroll
    for (int v = 0; v < NV; ++v) c[v] = gload(a.q[v] + off);
}

// cons2prim plus the species fractions X = (rho X) / max(rho, smallr) (spike_mhd.cu:243-244).
template <int NV, typename T>
GLM_DEV void c2p(const T c[NV], T q[NV], const Phys<T>& s)
{
    cons2prim(c, q, s);
    if (NV > 9) {
        const T ir = trcp(q[0]);
#pragma unroll
        for (int v = 9; v < NV; ++v) q[v] = c[v] * ir;
    }
}

// Species fluxes riding the face's mass flux (CMA, spike_mhd.cu:218-221): F_s = F_rho X_L if
// F_rho >= 0, else F_rho X_R, with L and R the face states in the grid frame.
template <int NV, typename T>
GLM_DEV void species_flux(const T L[NV], const T R[NV], T f[NV])
{
#pragma unroll
    for (int v = 9; v < NV; ++v) f[v] = f[0] >= T(0) ? f[0] * L[v] : f[0] * R[v];
}

// Face guard on one face state (glmmhd_turb.jl:618 iflo; spike_mhd.cu:145-147 clamp).
template <int NV, typename T>
GLM_DEV void guard(T f[NV], const T m0[NV], int g, const Phys<T>& s)
{
    if (g == GUARD_IFLO) {
        if (__builtin_expect(f[0] < s.smallr || f[4] < s.pfl, 0)) {
#pragma unroll
            for (int v = 0; v < NV; ++v) f[v] = m0[v];
        }
    } else if (g == GUARD_CLAMP) {
        f[0] = tmax(f[0], s.smallr);
        f[4] = tmax(f[4], s.pfl);
    }
}

// Normal-only Hancock half step of the cell centre in direction D (spike_mhd.cu:149-161); hs is
// the half slope.  Returns the primitive state.
template <typename T, int D, int NV>
GLM_DEV void hanc1d(const T c0[NV], const T hs[NV], T mh[NV], T half_dtdx, const Phys<T>& s)
{
    T t[9], fa[9], fb[9], u[9];
#pragma unroll
    for (int v = 0; v < 9; ++v) t[v] = c0[v] - hs[v];
    flux_dir<T, D>(t, fa, s);
#pragma unroll
    for (int v = 0; v < 9; ++v) t[v] = c0[v] + hs[v];
    flux_dir<T, D>(t, fb, s);
    prim2cons(c0, u, s);
#pragma unroll
    for (int v = 0; v < 9; ++v) u[v] -= half_dtdx * (fb[v] - fa[v]);
    cons2prim(u, mh, s);
    // species: the advective half step X - dt/dx u_D hs (spike_mhd.cu:155-159, hs the half slope)
#pragma unroll
    for (int v = 9; v < NV; ++v) mh[v] = c0[v] - (T(2) * half_dtdx * c0[1 + D]) * hs[v];
}

// Face states of one cell in direction D from its left/right neighbours l, r, the cell m0 and
// the predicted mh (glmmhd_turb.jl:378-403, :955-973; spike_mhd.cu:262-275).
template <typename T, int PRED, int RECON, int SLOPE, int D, int NV>
GLM_DEV void faces_dir(const T l[NV], const T m0[NV], const T r[NV], const T mhf[NV], bool up,
                       const MarchArgs<T>& a, T lo[NV], T hi[NV])
{
    const Phys<T>& s = a.ph;
    // half slopes: the PLM faces, the normal-only predictor, and the PPM shock fallback
    T hs[NV];
    if (RECON != RECON_PPM || PRED == PRED_ONED) {
#pragma unroll
        for (int v = 0; v < NV; ++v) hs[v] = half_slope<T, SLOPE>(l[v], m0[v], r[v]);
    }
    T mh[NV];
    if (PRED == PRED_ONED) {
        hanc1d<T, D, NV>(m0, hs, mh, a.half_dtdx, s);
        if (a.mh_revert && __builtin_expect(mh[0] <= s.smallr || mh[4] <= s.pfl, 0)) {
#pragma unroll
            for (int v = 0; v < NV; ++v) mh[v] = m0[v];
        }
    } else {
#pragma unroll
        for (int v = 0; v < NV; ++v) mh[v] = mhf[v];
    }
    if (RECON == RECON_PPM) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            T dl, dr;
            ppm_offsets(l[v], m0[v], r[v], dl, dr);
            lo[v] = mh[v] - dl;
            hi[v] = mh[v] + dr;
        }
        if (__builtin_expect(!up, 0)) {
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                T h = PRED == PRED_ONED ? hs[v] : half_slope<T, SLOPE>(l[v], m0[v], r[v]);
                lo[v] = mh[v] - h;
                hi[v] = mh[v] + h;
            }
        }
    } else if (RECON == RECON_PPM_SPIKE) {
#pragma unroll
        for (int v = 0; v < NV; ++v) ppm_edges_spike(l[v], m0[v], r[v], lo[v], hi[v]);
        lo[0] = tmax(lo[0], s.smallr); lo[4] = tmax(lo[4], s.pfl);
        hi[0] = tmax(hi[0], s.smallr); hi[4] = tmax(hi[4], s.pfl);
#pragma unroll
        for (int v = 0; v < NV; ++v) { lo[v] = lo[v] + mh[v] - m0[v]; hi[v] = hi[v] + mh[v] - m0[v]; }
        lo[0] = tmax(lo[0], s.smallr); lo[4] = tmax(lo[4], s.pfl);
        hi[0] = tmax(hi[0], s.smallr); hi[4] = tmax(hi[4], s.pfl);
    } else {
#pragma unroll
        for (int v = 0; v < NV; ++v) { lo[v] = mh[v] - hs[v]; hi[v] = mh[v] + hs[v]; }
    }
    guard<NV>(lo, m0, a.face_guard, s);
    guard<NV>(hi, m0, a.face_guard, s);
}

// Update epilogue of one cell (glmmhd_turb.jl:440-476), CFL speed of the result
// (glmmhd_turb.jl:1963-1971).  acc = (Fx_lo - Fx_hi) + (Fy_lo - Fy_hi) + (Fz_lo - Fz_hi).
template <int NV, typename T>
GLM_DEV double update_cell(T r[NV], const T acc[NV], const ColdArgs<T>& a, const Phys<T>& s, bool want_cfl,
                           ForceCache<T, NT>& fc, int gx, int gy, int gz)
{
    T xold[NV > 9 ? NV - 9 : 1];   // species fractions at the start of the step (for the floor)
    if (NV > 9) {
        const T ir = trcp(tmax(r[0], s.smallr));
#pragma unroll
        for (int v = 9; v < NV; ++v) xold[v - 9] = r[v] * ir;
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) r[v] = r[v] + a.dtdx * acc[v];
    if (a.update_floor && r[0] < s.smallr) {
        T emag = T(0.5) * (r[5] * r[5] + r[6] * r[6] + r[7] * r[7]);
        r[0] = s.smallr; r[1] = T(0); r[2] = T(0); r[3] = T(0);
        r[4] = s.pfl * s.igm1 + emag;
#pragma unroll
        for (int v = 9; v < NV; ++v) r[v] = s.smallr * xold[v - 9];
    }
    r[8] = r[8] * a.glm_fac;
    if (a.do_turb && r[0] >= a.turb_min_rho) {
        T acc3[3];
        force_accel(fc, a, gx, gy, gz, acc3);
        T rho = r[0], rhom = tmax(rho, s.smallr), irm = trcp(rhom), fl = rho * a.smallc * a.smallc;
        T e = r[4];
        e = tmax(e - T(0.5) * r[1] * r[1] * irm, fl);
        e = tmax(e - T(0.5) * r[2] * r[2] * irm, fl);
        e = tmax(e - T(0.5) * r[3] * r[3] * irm, fl);
        T m2 = r[1] + rhom * acc3[0] * a.dt;
        T m3 = r[2] + rhom * acc3[1] * a.dt;
        T m4 = r[3] + rhom * acc3[2] * a.dt;
        e = tmax(e + T(0.5) * m2 * m2 * irm, fl);
        e = tmax(e + T(0.5) * m3 * m3 * irm, fl);
        e = tmax(e + T(0.5) * m4 * m4 * irm, fl);
        r[1] = m2; r[2] = m3; r[3] = m4; r[4] = e;
    }
    if (!want_cfl) return 0.0;
    T q[9];
    cons2prim(r, q, s);
    return cfl_speed(q, s);
}

// LDS of one march workgroup (NV planes: 9, or 11 with two species).
template <typename T, int NV = 9>
struct MarchLDS {
    T P[NV][NPL];    // primitives of the current plane over the tile grown by two cells
    T XB[NV][NPL];   // +x face states, then x fluxes, at the LDS plane position of the cell
    T YB[NV][NPL];
    T FCB[6][NT];   // forcing interpolation cache (ForceCache)
    ColdArgs<T> ca;
};

template <typename T>
GLM_DEV void stage_cold(const MarchArgs<T>& a, ColdArgs<T>& ca)
{
    for (int v = 0; v < NVMAX; ++v) { ca.q[v] = a.q[v]; ca.o[v] = a.o[v]; }
    ca.afield = a.afield;
    ca.dtdx = a.dtdx; ca.glm_fac = a.glm_fac; ca.smallc = a.smallc; ca.turb_min_rho = a.turb_min_rho;
    ca.ramp = a.ramp; ca.dt = a.dt; ca.dx = a.dx; ca.lx = a.lx; ca.ly = a.ly; ca.lz = a.lz;
    ca.update_floor = a.update_floor; ca.do_turb = a.do_turb; ca.tgs = a.tgs;
}

// The march of one tile: the workgroup's lanes march the columns of tile (tile_x, tile_y) through
// z-chunk zc, tile = tile_x + tiles_x * (tile_y + tiles_y * zc), and write the updated planes.
// L.ca must hold `a`'s cold arguments, published by a barrier.  Returns the lane's max CFL speed
// over the cells it updated.  MF (the memory floor, spike_mhd.cu -DMEMFLOOR): the same loads,
// primitive conversion, LDS staging, update-base read and stores, with no predictor, faces, fluxes
// or epilogue: each cell is stored as read.
template <typename T, int PRED, int RIEM, int RECON, int SLOPE, int NS = 0, bool MF = false, typename A = MarchArgs<T>>
GLM_DEV double march_tile(const A& a, MarchLDS<T, 9 + NS>& L, int tile)
{
    constexpr int NV = 9 + NS;   // the GLM-MHD planes and NS passive species
    T (&P)[NV][NPL] = L.P;
    T (&XB)[NV][NPL] = L.XB;
    T (&YB)[NV][NPL] = L.YB;
    const ColdArgs<T>& ca = L.ca;
    const Phys<T>& s = a.ph;
    const int tid = threadIdx.x;
    int tx, ty;
    lane_cell(tid, tx, ty);
    const int tile_x = tile % a.tiles_x; tile /= a.tiles_x;
    const int tile_y = tile % a.tiles_y; tile /= a.tiles_y;
    const int zs = tile * a.zchunk;
    const int ze = min(zs + a.zchunk, a.nz);
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const size_t plane = (size_t)nx * ny;
    const int gxu = tile_x * OX - 1 + tx, gyu = tile_y * OY - 1 + ty;
    const int gx = wrapi(gxu, nx), gy = wrapi(gyu, ny);
    const bool owned = tid < NO && gxu < nx && gyu < ny;
    const bool need_x = ty >= 1 && ty <= OY && tx >= 1;      // this lane solves the x face at tx - 1/2
    const bool need_y = tx >= 1 && tx <= OX && ty >= 1;      // this lane solves the y face at ty - 1/2
    const size_t col = (size_t)gx + (size_t)nx * gy;
    const int lo_own = (ty + 1) * PX + (tx + 1);
    const bool helper = tid < NHELP;
    int hx = 0, hy = 0;
    helper_cell(helper ? tid : 0, hx, hy);
    const size_t hcol = (size_t)wrapi(tile_x * OX - 2 + hx, nx) + (size_t)nx * wrapi(tile_y * OY - 2 + hy, ny);
    const int lo_ring = hy * PX + hx;


    // column window: primitives of planes q-1, q, q+1
    T pm[NV], pp[NV];
    {
        T c[NV], p0[NV];
        loadv<NV>(ca, col + plane * wrapi(zs - 2, nz), c); c2p<NV>(c, pm, s);
        loadv<NV>(ca, col + plane * wrapi(zs - 1, nz), c); c2p<NV>(c, p0, s);
        loadv<NV>(ca, col + plane * wrapi(zs, nz), c); c2p<NV>(c, pp, s);
#pragma unroll
        for (int v = 0; v < NV; ++v) P[v][lo_own] = p0[v];
        if (helper) {
            T h[NV];
            loadv<NV>(ca, hcol + plane * wrapi(zs - 1, nz), c); c2p<NV>(c, h, s);
#pragma unroll
            for (int v = 0; v < NV; ++v) P[v][lo_ring] = h[v];
        }
    }
    __syncthreads();

    // zprev: +z face state of plane q-1.  pend: (Fx_lo - Fx_hi) + (Fy_lo - Fy_hi) of plane q-1 plus
    // Fz(q-3/2), so the update of q-1 is U + dt/dx * (pend - Fz(q-1/2)).
    T zprev[NV], pend[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) { zprev[v] = T(0); pend[v] = T(0); }
    double cmax = 0.0;
    ForceCache<T, NT> fc;
    fc.k0 = -0x7fffffff;
    fc.b = &L.FCB[0][tid];

    for (int q = zs - 1; q <= ze; ++q) {
        const bool last = q == ze;
        const bool inplane = q >= zs && q < ze;
        const bool do_upd = owned && q >= zs + 1;

        T u0[NV];
        T p0[NV];   // this column's cell of plane q, read back from the LDS plane (not carried in registers)
#pragma unroll
        for (int v = 0; v < NV; ++v) p0[v] = P[v][lo_own];

        // ---- reconstruction of this column's cell in plane q --------------------------------
        bool up = true;   // PPM strong-pressure-jump fallback flag (glmmhd_turb.jl:959)
        if (RECON == RECON_PPM && !MF) {
            up = !(spj(P[4][lo_own - 1], p0[4], P[4][lo_own + 1]) || spj(P[4][lo_own - PX], p0[4], P[4][lo_own + PX]) ||
                   spj(pm[4], p0[4], pp[4]));
        }
        T mh[NV];
        if (PRED == PRED_FULL && !MF) {
            // glmmhd_turb.jl:165-172 hancock, :391-395 revert: U0 - dt/(2dx) sum_d [F_d(m0+s_d) - F_d(m0-s_d)];
            // species: the advective half step X - dt/dx (u hs_x + v hs_y + w hs_z)
            T acc[9], t[9], f[9], hs[NV], sadv[NS > 0 ? NS : 1];
#pragma unroll
            for (int v = 0; v < NV; ++v) hs[v] = half_slope<T, SLOPE>(P[v][lo_own - 1], p0[v], P[v][lo_own + 1]);
#pragma unroll
            for (int v = 9; v < NV; ++v) sadv[v - 9] = p0[1] * hs[v];
#pragma unroll
            for (int v = 0; v < 9; ++v) t[v] = p0[v] + hs[v];
            flux_dir<T, 0>(t, acc, s);
#pragma unroll
            for (int v = 0; v < 9; ++v) t[v] = p0[v] - hs[v];
            flux_dir<T, 0>(t, f, s);
#pragma unroll
            for (int v = 0; v < 9; ++v) acc[v] -= f[v];
#pragma unroll
            for (int v = 0; v < NV; ++v) hs[v] = half_slope<T, SLOPE>(P[v][lo_own - PX], p0[v], P[v][lo_own + PX]);
#pragma unroll
            for (int v = 9; v < NV; ++v) sadv[v - 9] = sadv[v - 9] + p0[2] * hs[v];
#pragma unroll
            for (int v = 0; v < 9; ++v) t[v] = p0[v] + hs[v];
            flux_dir<T, 1>(t, f, s);
#pragma unroll
            for (int v = 0; v < 9; ++v) acc[v] += f[v];
#pragma unroll
            for (int v = 0; v < 9; ++v) t[v] = p0[v] - hs[v];
            flux_dir<T, 1>(t, f, s);
#pragma unroll
            for (int v = 0; v < 9; ++v) acc[v] -= f[v];
#pragma unroll
            for (int v = 0; v < NV; ++v) hs[v] = half_slope<T, SLOPE>(pm[v], p0[v], pp[v]);
#pragma unroll
            for (int v = 9; v < NV; ++v) sadv[v - 9] = sadv[v - 9] + p0[3] * hs[v];
#pragma unroll
            for (int v = 0; v < 9; ++v) t[v] = p0[v] + hs[v];
            flux_dir<T, 2>(t, f, s);
#pragma unroll
            for (int v = 0; v < 9; ++v) acc[v] += f[v];
#pragma unroll
            for (int v = 0; v < 9; ++v) t[v] = p0[v] - hs[v];
            flux_dir<T, 2>(t, f, s);
#pragma unroll
            for (int v = 0; v < 9; ++v) acc[v] -= f[v];
            prim2cons(p0, t, s);
#pragma unroll
            for (int v = 0; v < 9; ++v) t[v] = t[v] - a.half_dtdx * acc[v];
            cons2prim(t, mh, s);
#pragma unroll
            for (int v = 9; v < NV; ++v) mh[v] = p0[v] - a.dtdx * sadv[v - 9];
            if (a.mh_revert && __builtin_expect(mh[0] <= s.smallr || mh[4] <= s.pfl, 0)) {
#pragma unroll
                for (int v = 0; v < NV; ++v) mh[v] = p0[v];
            }
        }

        // ---- z face q-1/2 and the update of plane q-1 ---------------------------------------
        T fz[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) fz[v] = T(0);
        if (owned && !MF) {
            T zlo[NV], zhi[NV];
            faces_dir<T, PRED, RECON, SLOPE, 2, NV>(pm, p0, pp, mh, up, a, zlo, zhi);
            if (q >= zs) {
                riemann<T, RIEM, 2>(zprev, zlo, fz, s);
                species_flux<NV>(zprev, zlo, fz);
            }
#pragma unroll
            for (int v = 0; v < NV; ++v) zprev[v] = zhi[v];
        }
        if (PRED == PRED_FULL && inplane && !MF) {
            // park the predicted state in this lane's own +x slot (free until its x face state is
            // stored) across the z solve and the update
#pragma unroll
            for (int v = 0; v < NV; ++v) XB[v][lo_own] = mh[v];
        }
        if (do_upd) {
#pragma unroll
            for (int v = 0; v < NV; ++v) u0[v] = gload(ca.q[v] + col + plane * wrapi(q - 1, nz));
            const int zq = wrapi(q - 1, nz);
            if (!MF) {
                T acc[NV];
#pragma unroll
                for (int v = 0; v < NV; ++v) acc[v] = pend[v] - fz[v];
                double ct = update_cell<NV>(u0, acc, ca, s, a.want_cfl != 0, fc, gx, gy, zq);
                if (!(ct <= cmax)) cmax = ct;
            }
            const size_t off = col + plane * zq;
#pragma unroll
            for (int v = 0; v < NV; ++v) gstore_nt(u0[v], ca.o[v] + off);
        }
        if (last) break;


        // ---- x and y faces of plane q (only for planes this chunk updates) ---------------------
        if (inplane && !MF) {
            T xlo[NV], ylo[NV];
            __asm__ volatile("" ::: "memory");   // re-read p0 here instead of keeping it live
#pragma unroll
            for (int v = 0; v < NV; ++v) p0[v] = P[v][lo_own];
            if (PRED == PRED_FULL) {
#pragma unroll
                for (int v = 0; v < NV; ++v) mh[v] = XB[v][lo_own];
            }
            {
                T l[NV], r[NV], hi[NV];
#pragma unroll
                for (int v = 0; v < NV; ++v) { l[v] = P[v][lo_own - 1]; r[v] = P[v][lo_own + 1]; }
                faces_dir<T, PRED, RECON, SLOPE, 0, NV>(l, p0, r, mh, up, a, xlo, hi);
#pragma unroll
                for (int v = 0; v < NV; ++v) XB[v][lo_own] = hi[v];
#pragma unroll
                for (int v = 0; v < NV; ++v) { l[v] = P[v][lo_own - PX]; r[v] = P[v][lo_own + PX]; }
                faces_dir<T, PRED, RECON, SLOPE, 1, NV>(l, p0, r, mh, up, a, ylo, hi);
#pragma unroll
                for (int v = 0; v < NV; ++v) YB[v][lo_own] = hi[v];
            }
            // From here to the plane's last barrier (face solves, flux exchange, the next planes'
            // loads and the LDS publish) this wavefront issues ahead of the other workgroup's on its
            // SIMD, which is then in its reconstruction: the workgroups reach their barriers sooner
            // (FP64 march -0.9%, FP32 -4% at 256^3).
            __builtin_amdgcn_s_setprio(1);
            __syncthreads();
            T fx[NV], fy[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) { fx[v] = T(0); fy[v] = T(0); }
            if (need_x) {
                T L[NV];
#pragma unroll
                for (int v = 0; v < NV; ++v) L[v] = XB[v][lo_own - 1];
                riemann<T, RIEM, 0>(L, xlo, fx, s);
                species_flux<NV>(L, xlo, fx);
            }
            if (need_y) {
                T L[NV];
#pragma unroll
                for (int v = 0; v < NV; ++v) L[v] = YB[v][lo_own - PX];
                riemann<T, RIEM, 1>(L, ylo, fy, s);
                species_flux<NV>(L, ylo, fy);
            }
            __syncthreads();
#pragma unroll
            for (int v = 0; v < NV; ++v) { XB[v][lo_own] = fx[v]; YB[v][lo_own] = fy[v]; }
            __syncthreads();
            if (owned) {
#pragma unroll
                for (int v = 0; v < NV; ++v)
                    pend[v] = ((fx[v] - XB[v][lo_own + 1]) + (fy[v] - YB[v][lo_own + PX])) + fz[v];
            }
        }

        if (!MF) __builtin_amdgcn_s_setprio(1);   // (set already when this plane's faces were solved)
        T cn[NV], hr[NV];
        loadv<NV>(ca, col + plane * wrapi(q + 2, nz), cn);
        if (helper) loadv<NV>(ca, hcol + plane * wrapi(q + 1, nz), hr);
        // ---- advance the column window and publish plane q+1 to LDS ------------------------------
        // The last reads of plane q's LDS cells precede the face-state barrier when this plane's
        // faces were solved; otherwise (the lead-in plane) they are the predictor's.
        if (!inplane || MF) __syncthreads();
#pragma unroll
        for (int v = 0; v < NV; ++v) { pm[v] = P[v][lo_own]; P[v][lo_own] = pp[v]; }
        c2p<NV>(cn, pp, s);
        if (helper) {
            T h[NV];
            c2p<NV>(hr, h, s);
#pragma unroll
            for (int v = 0; v < NV; ++v) P[v][lo_ring] = h[v];
        }
        __syncthreads();
        if (!MF) __builtin_amdgcn_s_setprio(0);   // the reconstruction runs at the base priority
    }

    return cmax;
}

// Workgroup max of the lanes' CFL speeds through LDS (the last wavefront may be partial), then one
// atomic on the bits of a non-negative double (NaN compares largest).
template <typename LDS>
GLM_DEV void cfl_merge(LDS& L, double cmax, unsigned long long* target)
{
    const int tid = threadIdx.x;
    __syncthreads();
    unsigned long long* red = reinterpret_cast<unsigned long long*>(&L.XB[0][0]);
    red[tid] = (unsigned long long)__double_as_longlong(cmax);
    __syncthreads();
    if (tid < 64) {
        unsigned long long bits = 0ull;
        for (int t = tid; t < NT; t += 64) bits = red[t] > bits ? red[t] : bits;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            unsigned long long o = __shfl_xor(bits, off, 64);
            bits = o > bits ? o : bits;
        }
        if (tid == 0 && bits != 0ull) atomicMax(target, bits);
    }
}

// One step of one grid: workgroup w marches tile xcd_block(w) (x fastest, then y, then z-chunk).
// NS = 0 or 2 passive species after the 9 GLM-MHD planes.
template <typename T, int PRED, int RIEM, int RECON, int SLOPE, int NS = 0>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(GLM_WAVES_PER_EU_SCHEME(T, NS, PRED, RECON)))) march_kernel(MarchArgs<T> a)
{
    __shared__ MarchLDS<T, 9 + NS> L;
    const auto& ak = kernarg_ref<MarchArgs<T>>();
    if (threadIdx.x == 0) stage_cold(a, L.ca);
    __syncthreads();
    const double cmax = march_tile<T, PRED, RIEM, RECON, SLOPE, NS>(ak, L, xcd_block((int)blockIdx.x, (int)gridDim.x));
    if (ak.want_cfl) {
        cfl_merge(L, cmax, ak.cmax);
        if (blockIdx.x == 0 && threadIdx.x == 0) *ak.cmax_next = 0ull;
    }
}

// The memory floor of the march (spike_mhd.cu -DMEMFLOOR): march_kernel's data movement and
// launch shape with no flux computation (march_tile with MF); o receives q unchanged.  The layout
// and LDS are those of the FP64/FP32 march without species; the scheme arguments only select them.
template <typename T>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(GLM_WAVES_PER_EU(T)))) march_floor_kernel(MarchArgs<T> a)
{
    __shared__ MarchLDS<T, 9> L;
    const auto& ak = kernarg_ref<MarchArgs<T>>();
    if (threadIdx.x == 0) stage_cold(a, L.ca);
    __syncthreads();
    (void)march_tile<T, PRED_FULL, RIEM_HLLD, RECON_PPM, SLOPE_MONCEN, 0, true>(ak, L, xcd_block((int)blockIdx.x, (int)gridDim.x));
}

// ---------------------------------------------------------------------------------------------
// Ensemble: many boxes of uneven size in one launch per sweep.  The boxes' tiles form one flat
// list (box positions in a chosen order, each box's tiles contiguous in its own x, y, z-chunk
// order); tile_off[pos] is the first flat tile of position pos.  Each box's MarchArgs live in
// device memory and are refreshed every sweep by the step-control kernel (ens.hip).
// ---------------------------------------------------------------------------------------------
constexpr int ENS_LOG = 256;   // sweeps kept in the per-box log and timing records

template <typename T>
struct EnsLaunch {
    const MarchArgs<T>* boxes;   // per box
    const int* order;            // flat position -> box
    const int* tile_off;         // [npos + 1]
    int npos, total;             // positions and flat tiles
    int persistent;              // 0: workgroup w marches flat tile xcd_block(w); 1: workgroups take
                                 // tiles from per-XCD queues and steal from the others when theirs is empty
    int nbox;
    unsigned int* queue;         // [NXCD] ticket counters, zeroed by the step-control kernel
    const long long* sweep;      // index of the current sweep
    unsigned long long* t0;      // [ENS_LOG][nbox] wall-clock start of a box's first tile in the sweep
    unsigned long long* t1;      // [ENS_LOG][nbox] wall-clock end of its last tile
};


// First flat tile of XCD x's contiguous range (the inverse view of xcd_block) and its length.
GLM_DEV int xcd_start(int x, int total)
{
    const int q = total / NXCD, r = total % NXCD;
    return x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q;
}
GLM_DEV int xcd_count(int x, int total) { return total / NXCD + (x < total % NXCD ? 1 : 0); }

// Thread 0 claims the workgroup's next tile (flat: its own, once; persistent: a ticket from the
// XCD queues, stealing when its own is empty), finds the box by binary search over the tile
// offsets and stages the box's cold arguments; the job is published in s_job after a barrier.
template <typename T, bool PERSIST>
GLM_DEV void ens_claim(const EnsLaunch<T>& e, int it, int& xq, unsigned long long* s_tb, MarchLDS<T>& L, int* s_job)
{
    int tile = -1;
    if (!PERSIST) {
        if (it == 0) tile = xcd_block((int)blockIdx.x, (int)gridDim.x);
    } else {
        for (int k = 0; k < NXCD && tile < 0; ++k) {
            const int x = (xq + k) % NXCD;
            const unsigned tk = atomicAdd(&e.queue[x], 1u);
            if (tk < (unsigned)xcd_count(x, e.total)) {
                tile = xcd_start(x, e.total) + (int)tk;
                xq = x;
            }
        }
    }
    int box = -1, lt = 0;
    if (tile >= 0) {
        int lo = 0, hi = e.npos;   // tile_off[lo] <= tile < tile_off[lo + 1]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (e.tile_off[mid] <= tile) lo = mid; else hi = mid;
        }
        box = e.order[lo];
        lt = tile - e.tile_off[lo];
        stage_cold(e.boxes[box], L.ca);
        *s_tb = (unsigned long long)wall_clock64();
    }
    s_job[0] = box;
    s_job[1] = lt;
}

// The march of one claimed tile, its CFL merge into the box's slot and the box's timing record
// (the start time waits in LDS, out of the registers the march needs).
template <typename T, int PRED, int RIEM, int RECON, int SLOPE>
GLM_DEV void ens_tile(const EnsLaunch<T>& e, MarchLDS<T>& L, int box, int lt, int slot, const unsigned long long* s_tb)
{
    const auto& a = constant_ref(e.boxes + box);
    const double cmax = march_tile<T, PRED, RIEM, RECON, SLOPE>(a, L, lt);
    cfl_merge(L, cmax, a.cmax);
    if (threadIdx.x == 0) {
        const size_t k = (size_t)slot * e.nbox + box;
        atomicMin(&e.t0[k], *s_tb);
        atomicMax(&e.t1[k], (unsigned long long)wall_clock64());
    }
}

// PERSIST = false: one workgroup per flat tile, straight-line (the same register budget as the
// single-grid kernel).  PERSIST = true: a loop over claimed tiles until the queues are empty.
template <typename T, int PRED, int RIEM, int RECON, int SLOPE, bool PERSIST>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(GLM_WAVES_PER_EU(T)))) ens_march_kernel(EnsLaunch<T> e)
{
    __shared__ MarchLDS<T> L;
    __shared__ int s_job[2];   // box (-1: no more work) and its local tile
    __shared__ unsigned long long s_tb;   // wall clock at the claim
    const int tid = threadIdx.x;
    const long long sw = *e.sweep;   // -1 before the first sweep (tuning launches)
    const int slot = (int)((sw < 0 ? 0 : sw) % ENS_LOG);
    int xq = (int)blockIdx.x % NXCD;
    if (!PERSIST) {
        if (tid == 0) ens_claim<T, false>(e, 0, xq, &s_tb, L, s_job);
        __syncthreads();
        const int box = __builtin_amdgcn_readfirstlane(s_job[0]);
        const int lt = __builtin_amdgcn_readfirstlane(s_job[1]);
        if (box < 0) return;
        ens_tile<T, PRED, RIEM, RECON, SLOPE>(e, L, box, lt, slot, &s_tb);
        return;
    }
    for (int it = 0;; ++it) {
        if (tid == 0) ens_claim<T, true>(e, it, xq, &s_tb, L, s_job);
        __syncthreads();
        const int box = __builtin_amdgcn_readfirstlane(s_job[0]);
        const int lt = __builtin_amdgcn_readfirstlane(s_job[1]);
        if (box < 0) break;
        ens_tile<T, PRED, RIEM, RECON, SLOPE>(e, L, box, lt, slot, &s_tb);
        __syncthreads();   // LDS (job, cold args, reduction) is free for the next tile
    }
}

}  // namespace glm
