// adj_impl.hpp -- reverse (adjoint) step of the GLM-MHD march for gfx950, FP64.
//
// Given the state u0 a step started from and the adjoint lam1 of the state it produced, the
// kernels below compute lam0 = (dU1/dU0)^T lam1 for that step with its scalars (dt, ch, psi
// decay, ramp, forcing field) held fixed.  The step is re-derived in passes, one thread per cell
// (every pass is a stencil over the periodic grid):
//   1 prim    W = c2p(u0)
//   2 faces   predictor (full Hancock) and the six face states of every cell
//   3 upre    the flux through every face, each solved once (FLX), then the step's flux
//             divergence -> the pre-epilogue state and the epilogue adjoint (kick, psi decay,
//             floor): LPRE = adjoint of the pre-epilogue state
//   4 riemann per lower face: lambda_F = dtdx (LPRE(c) - LPRE(c - e_d)), Riemann adjoint ->
//             adjoints of the two face states meeting there (each slot has one writer)
//   5 cell    face-state adjoints (one launch per direction) and the predictor adjoint of every
//             cell -> contributions to the primitive adjoints of the cell and its six
//             neighbours (7 x 9 values per cell)
//   6 gather  each cell sums the contributions aimed at it, then c2p adjoint: lam0
// Every pass writes only its own cell's slots (pass 4 writes the slots of the two face states of
// one face), so there are no atomics and the result is deterministic.
#pragma once
#include "adjoint.hpp"
#include "adj_launch.hpp"

namespace glm {

GLM_DEV int adj_wrap(int i, int n)
{
    int m = i % n;
    return m < 0 ? m + n : m;
}

struct Cell {
    size_t c;
    int i, j, k;
};
GLM_DEV Cell cell_of(size_t c, int nx, int ny)
{
    Cell r;
    r.c = c;
    r.i = (int)(c % nx);
    r.j = (int)((c / nx) % ny);
    r.k = (int)(c / ((size_t)nx * ny));
    return r;
}
// neighbour index in direction d (0..2), offset o (-1 or +1)
GLM_DEV size_t nbr(const Cell& x, int d, int o, int nx, int ny, int nz)
{
    int i = x.i, j = x.j, k = x.k;
    if (d == 0) i = adj_wrap(i + o, nx);
    else if (d == 1) j = adj_wrap(j + o, ny);
    else k = adj_wrap(k + o, nz);
    return (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k);
}

GLM_DEV double* plane(const AdjArgs& a, int p, size_t n) { return a.scr + (size_t)p * n; }
GLM_DEV void ld9(const double* base, size_t n, size_t c, double x[9])
{
#pragma unroll
    for (int v = 0; v < 9; ++v) x[v] = base[(size_t)v * n + c];
}
GLM_DEV void st9(double* base, size_t n, size_t c, const double x[9])
{
#pragma unroll
    for (int v = 0; v < 9; ++v) base[(size_t)v * n + c] = x[v];
}
// face-state slot: direction d, side 0 (low face) or 1 (high face)
GLM_DEV int fs_slot(int d, int side) { return (d * 2 + side) * 9; }

// ---- pass 1 ----------------------------------------------------------------------------------
__global__ void __launch_bounds__(ADJ_NT) adj_prim_kernel(AdjArgs a)
{
    const size_t n = (size_t)a.nx * a.ny * a.nz;
    for (size_t c = (size_t)blockIdx.x * ADJ_NT + threadIdx.x; c < n; c += (size_t)gridDim.x * ADJ_NT) {
        double u[9], w[9];
#pragma unroll
        for (int v = 0; v < 9; ++v) u[v] = a.u0[v][c];
        cons2prim(u, w, a.ph);
        st9(plane(a, ADJ_W, n), n, c, w);
    }
}

// PPM strong-pressure-jump flag of a cell from its pressure stencil (glmmhd_turb.jl:959)
GLM_DEV bool ppm_up(const double p0[9], const double (*nb)[9])
{
    return !(spj(nb[0][4], p0[4], nb[1][4]) || spj(nb[2][4], p0[4], nb[3][4]) || spj(nb[4][4], p0[4], nb[5][4]));
}

GLM_DEV void load_stencil(const AdjArgs& a, const Cell& x, size_t n, double p0[9], double nb[6][9])
{
    const double* W = plane(a, ADJ_W, n);
    ld9(W, n, x.c, p0);
    for (int d = 0; d < 3; ++d) {
        ld9(W, n, nbr(x, d, -1, a.nx, a.ny, a.nz), nb[2 * d]);
        ld9(W, n, nbr(x, d, +1, a.nx, a.ny, a.nz), nb[2 * d + 1]);
    }
}

// ---- pass 2 ----------------------------------------------------------------------------------
template <int PRED, int RECON, int SLOPE>
__global__ void __launch_bounds__(ADJ_NT) adj_faces_kernel(AdjArgs a)
{
    const size_t n = (size_t)a.nx * a.ny * a.nz;
    const FaceOpts o{a.mh_revert, a.face_guard};
    for (size_t c = (size_t)blockIdx.x * ADJ_NT + threadIdx.x; c < n; c += (size_t)gridDim.x * ADJ_NT) {
        const Cell x = cell_of(c, a.nx, a.ny);
        double p0[9], nb[6][9], mh[9], lo[9], hi[9];
        load_stencil(a, x, n, p0, nb);
        const bool up = RECON == RECON_PPM ? ppm_up(p0, nb) : true;
        if (PRED == PRED_FULL) {
            hancock_fwd<double, SLOPE>(p0, nb, mh, a.half_dtdx, a.mh_revert, a.ph);
            st9(plane(a, ADJ_MH, n), n, c, mh);
        }
        faces_fwd<double, PRED, RECON, SLOPE, 0>(nb[0], p0, nb[1], mh, up, a.half_dtdx, o, a.ph, lo, hi);
        st9(plane(a, ADJ_FS + fs_slot(0, 0), n), n, c, lo);
        st9(plane(a, ADJ_FS + fs_slot(0, 1), n), n, c, hi);
        faces_fwd<double, PRED, RECON, SLOPE, 1>(nb[2], p0, nb[3], mh, up, a.half_dtdx, o, a.ph, lo, hi);
        st9(plane(a, ADJ_FS + fs_slot(1, 0), n), n, c, lo);
        st9(plane(a, ADJ_FS + fs_slot(1, 1), n), n, c, hi);
        faces_fwd<double, PRED, RECON, SLOPE, 2>(nb[4], p0, nb[5], mh, up, a.half_dtdx, o, a.ph, lo, hi);
        st9(plane(a, ADJ_FS + fs_slot(2, 0), n), n, c, lo);
        st9(plane(a, ADJ_FS + fs_slot(2, 1), n), n, c, hi);
    }
}

// Forcing acceleration at the centre of cell (i,j,k) (glmmhd_turb.jl:483-505), ramp applied.
GLM_DEV void adj_force(const AdjArgs& a, int i, int j, int k, double a3[3])
{
    const int g = a.tgs;
    const double gg = (double)g;
    const double cx = (((double)i + 0.5) * a.dx) / a.lx * gg;
    const double cy = (((double)j + 0.5) * a.dx) / a.ly * gg;
    const double cz = (((double)k + 0.5) * a.dx) / a.lz * gg;
    const double fx = floor(cx), fy = floor(cy), fz = floor(cz);
    const int i0 = (int)fx, j0 = (int)fy, k0 = (int)fz;
    const double wx = cx - fx, wy = cy - fy, wz = cz - fz;
    const int i0w = adj_wrap(i0, g), i1w = adj_wrap(i0 + 1, g), j0w = adj_wrap(j0, g), j1w = adj_wrap(j0 + 1, g);
    const int k0w = adj_wrap(k0, g), k1w = adj_wrap(k0 + 1, g);
    const size_t g3 = (size_t)g * g * g;
    for (int cc = 0; cc < 3; ++cc) {
        const double* af = a.afield + cc * g3;
        auto at = [&](int ii, int jj, int kk) { return af[(size_t)ii + (size_t)g * ((size_t)jj + (size_t)g * kk)]; };
        const double b0 = (at(i0w, j0w, k0w) * (1.0 - wx) + at(i1w, j0w, k0w) * wx) * (1.0 - wy) +
                          (at(i0w, j1w, k0w) * (1.0 - wx) + at(i1w, j1w, k0w) * wx) * wy;
        const double b1 = (at(i0w, j0w, k1w) * (1.0 - wx) + at(i1w, j0w, k1w) * wx) * (1.0 - wy) +
                          (at(i0w, j1w, k1w) * (1.0 - wx) + at(i1w, j1w, k1w) * wx) * wy;
        a3[cc] = a.ramp * (b0 * (1.0 - wz) + b1 * wz);
    }
}

template <int RIEM, int D>
GLM_DEV void face_flux(const AdjArgs& a, size_t n, size_t cl, size_t cr, double f[9])
{
    double L[9], R[9];
    ld9(plane(a, ADJ_FS + fs_slot(D, 1), n), n, cl, L);
    ld9(plane(a, ADJ_FS + fs_slot(D, 0), n), n, cr, R);
    riemann<double, RIEM, D>(L, R, f, a.ph);
}

// ---- pass 3 ----------------------------------------------------------------------------------
// 3a: the flux through each cell's three lower faces, each face solved once -> FLX (27 planes)
template <int RIEM>
__global__ void __launch_bounds__(ADJ_NT) adj_flux_kernel(AdjArgs a)
{
    const size_t n = (size_t)a.nx * a.ny * a.nz;
    for (size_t c = (size_t)blockIdx.x * ADJ_NT + threadIdx.x; c < n; c += (size_t)gridDim.x * ADJ_NT) {
        const Cell x = cell_of(c, a.nx, a.ny);
        double f[9];
        face_flux<RIEM, 0>(a, n, nbr(x, 0, -1, a.nx, a.ny, a.nz), c, f);
        st9(plane(a, ADJ_FLX, n), n, c, f);
        face_flux<RIEM, 1>(a, n, nbr(x, 1, -1, a.nx, a.ny, a.nz), c, f);
        st9(plane(a, ADJ_FLX + 9, n), n, c, f);
        face_flux<RIEM, 2>(a, n, nbr(x, 2, -1, a.nx, a.ny, a.nz), c, f);
        st9(plane(a, ADJ_FLX + 18, n), n, c, f);
    }
}

// 3b: flux divergence from FLX, then the epilogue adjoint
__global__ void __launch_bounds__(ADJ_NT) adj_upre_kernel(AdjArgs a)
{
    __shared__ double red[ADJ_NT];
    const size_t n = (size_t)a.nx * a.ny * a.nz;
    const UpdOpts o{a.update_floor, a.do_turb};
    double fdot = 0.0;
    for (size_t c = (size_t)blockIdx.x * ADJ_NT + threadIdx.x; c < n; c += (size_t)gridDim.x * ADJ_NT) {
        const Cell x = cell_of(c, a.nx, a.ny);
        double acc[9], fl[9], fh[9];
        ld9(plane(a, ADJ_FLX, n), n, c, fl);
        ld9(plane(a, ADJ_FLX, n), n, nbr(x, 0, +1, a.nx, a.ny, a.nz), fh);
#pragma unroll
        for (int v = 0; v < 9; ++v) acc[v] = fl[v] - fh[v];
        ld9(plane(a, ADJ_FLX + 9, n), n, c, fl);
        ld9(plane(a, ADJ_FLX + 9, n), n, nbr(x, 1, +1, a.nx, a.ny, a.nz), fh);
#pragma unroll
        for (int v = 0; v < 9; ++v) acc[v] += fl[v] - fh[v];
        ld9(plane(a, ADJ_FLX + 18, n), n, c, fl);
        ld9(plane(a, ADJ_FLX + 18, n), n, nbr(x, 2, +1, a.nx, a.ny, a.nz), fh);
        double r[9], lam[9], rb[9];
#pragma unroll
        for (int v = 0; v < 9; ++v) {
            acc[v] += fl[v] - fh[v];
            r[v] = a.u0[v][c] + a.dtdx * acc[v];
            lam[v] = a.lam1[v][c];
            rb[v] = 0.0;
        }
        double a3[3] = {0.0, 0.0, 0.0}, a3b[3] = {0.0, 0.0, 0.0};
        if (a.do_turb) adj_force(a, x.i, x.j, x.k, a3);
        epilogue_adj<double>(r, a3, a.glm_fac, a.smallc, a.turb_min_rho, a.dt, o, a.ph, lam, rb, a3b);
        st9(plane(a, ADJ_LPRE, n), n, c, rb);
        fdot += a3b[0] * a3[0] + a3b[1] * a3[1] + a3b[2] * a3[2];
    }
    red[threadIdx.x] = fdot;
    __syncthreads();
    for (int s = ADJ_NT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) a.part[blockIdx.x] = red[0];
}

// ---- pass 4 ----------------------------------------------------------------------------------
template <int RIEM, int D>
GLM_DEV void face_adj(const AdjArgs& a, size_t n, const Cell& x)
{
    const size_t cl = nbr(x, D, -1, a.nx, a.ny, a.nz), cr = x.c;
    double L[9], R[9], lr[9], ll[9], ob[9];
    ld9(plane(a, ADJ_FS + fs_slot(D, 1), n), n, cl, L);
    ld9(plane(a, ADJ_FS + fs_slot(D, 0), n), n, cr, R);
    ld9(plane(a, ADJ_LPRE, n), n, cr, lr);
    ld9(plane(a, ADJ_LPRE, n), n, cl, ll);
    double Lb[9], Rb[9];
#pragma unroll
    for (int v = 0; v < 9; ++v) {
        ob[v] = a.dtdx * (lr[v] - ll[v]);
        Lb[v] = 0.0;
        Rb[v] = 0.0;
    }
    riemann_adj<double, RIEM, D>(L, R, ob, Lb, Rb, a.ph);
    st9(plane(a, ADJ_LFS + fs_slot(D, 1), n), n, cl, Lb);
    st9(plane(a, ADJ_LFS + fs_slot(D, 0), n), n, cr, Rb);
}
template <int RIEM>
__global__ void __launch_bounds__(ADJ_NT) adj_riemann_kernel(AdjArgs a)
{
    const size_t n = (size_t)a.nx * a.ny * a.nz;
    for (size_t c = (size_t)blockIdx.x * ADJ_NT + threadIdx.x; c < n; c += (size_t)gridDim.x * ADJ_NT) {
        const Cell x = cell_of(c, a.nx, a.ny);
        face_adj<RIEM, 0>(a, n, x);
        face_adj<RIEM, 1>(a, n, x);
        face_adj<RIEM, 2>(a, n, x);
    }
}

// ---- pass 5 ----------------------------------------------------------------------------------
// CON slot k of a cell: 0 the cell itself, 1 + 2d its -d neighbour, 2 + 2d its +d neighbour.
// 5d (one launch per direction D): face-state adjoints of the cell's two faces in D -> slots
// 1 + 2D and 2 + 2D, plus the cell's own slot 0 and the predictor adjoint MHB (the D = 0 launch
// stores, the others add).  5h (full predictor): the predictor adjoint -> slots 0..6 (added).
// One direction per launch keeps each thread's live set small.
GLM_DEV void load_p(const AdjArgs& a, size_t n, size_t c, double* out9) { ld9(plane(a, ADJ_W, n), n, c, out9); }

template <int PRED, int RECON, int SLOPE, int D>
__global__ void __launch_bounds__(ADJ_NT) adj_cell_dir_kernel(AdjArgs a)
{
    const size_t n = (size_t)a.nx * a.ny * a.nz;
    const FaceOpts o{a.mh_revert, a.face_guard};
    double* C = plane(a, ADJ_CON, n);
    double* MHB = plane(a, ADJ_MHB, n);
    for (size_t c = (size_t)blockIdx.x * ADJ_NT + threadIdx.x; c < n; c += (size_t)gridDim.x * ADJ_NT) {
        const Cell x = cell_of(c, a.nx, a.ny);
        double p0[9], l[9], r[9], mh[9], lob[9], hib[9], lb[9], m0b[9], rb[9], mhb[9];
        load_p(a, n, c, p0);
        load_p(a, n, nbr(x, D, -1, a.nx, a.ny, a.nz), l);
        load_p(a, n, nbr(x, D, +1, a.nx, a.ny, a.nz), r);
        bool up = true;
        if (RECON == RECON_PPM) {
            const double* P = plane(a, ADJ_W, n) + 4 * n;
            double pr[6];
            for (int d = 0; d < 3; ++d) {
                pr[2 * d] = P[nbr(x, d, -1, a.nx, a.ny, a.nz)];
                pr[2 * d + 1] = P[nbr(x, d, +1, a.nx, a.ny, a.nz)];
            }
            up = !(spj(pr[0], p0[4], pr[1]) || spj(pr[2], p0[4], pr[3]) || spj(pr[4], p0[4], pr[5]));
        }
        if (PRED == PRED_FULL) ld9(plane(a, ADJ_MH, n), n, c, mh);
        else for (int v = 0; v < 9; ++v) mh[v] = 0.0;
        ld9(plane(a, ADJ_LFS + fs_slot(D, 0), n), n, c, lob);
        ld9(plane(a, ADJ_LFS + fs_slot(D, 1), n), n, c, hib);
        for (int v = 0; v < 9; ++v) { lb[v] = 0.0; m0b[v] = 0.0; rb[v] = 0.0; mhb[v] = 0.0; }
        faces_adj<double, PRED, RECON, SLOPE, D>(l, p0, r, mh, up, a.half_dtdx, o, a.ph, lob, hib, lb, m0b, rb, mhb);
        st9(C + (size_t)(1 + 2 * D) * 9 * n, n, c, lb);
        st9(C + (size_t)(2 + 2 * D) * 9 * n, n, c, rb);
        if (D > 0) {
            double t[9];
            ld9(C, n, c, t);
            for (int v = 0; v < 9; ++v) m0b[v] += t[v];
            if (PRED == PRED_FULL) {
                ld9(MHB, n, c, t);
                for (int v = 0; v < 9; ++v) mhb[v] += t[v];
            }
        }
        st9(C, n, c, m0b);
        if (PRED == PRED_FULL) st9(MHB, n, c, mhb);
    }
}

template <int SLOPE, int D>
GLM_DEV void hancock_dir_add(const AdjArgs& a, size_t n, const Cell& x, const double p0[9], const double accb[9],
                             double p0b[9])
{
    double l[9], r[9], lb[9], rb[9], t[9];
    const size_t cl = nbr(x, D, -1, a.nx, a.ny, a.nz), cr = nbr(x, D, +1, a.nx, a.ny, a.nz);
    load_p(a, n, cl, l);
    load_p(a, n, cr, r);
    for (int v = 0; v < 9; ++v) { lb[v] = 0.0; rb[v] = 0.0; }
    hancock_adj_dir<double, SLOPE, D>(p0, l, r, accb, p0b, lb, rb, a.ph);
    double* C = plane(a, ADJ_CON, n);
    ld9(C + (size_t)(1 + 2 * D) * 9 * n, n, x.c, t);
    for (int v = 0; v < 9; ++v) t[v] += lb[v];
    st9(C + (size_t)(1 + 2 * D) * 9 * n, n, x.c, t);
    ld9(C + (size_t)(2 + 2 * D) * 9 * n, n, x.c, t);
    for (int v = 0; v < 9; ++v) t[v] += rb[v];
    st9(C + (size_t)(2 + 2 * D) * 9 * n, n, x.c, t);
}
template <int SLOPE, int D>
GLM_DEV void hancock_acc_load(const AdjArgs& a, size_t n, const Cell& x, const double p0[9], double acc[9])
{
    double l[9], r[9];
    load_p(a, n, nbr(x, D, -1, a.nx, a.ny, a.nz), l);
    load_p(a, n, nbr(x, D, +1, a.nx, a.ny, a.nz), r);
    hancock_acc_dir<double, SLOPE, D>(p0, l, r, acc, a.ph);
}

template <int SLOPE>
__global__ void __launch_bounds__(ADJ_NT) adj_hancock_kernel(AdjArgs a)
{
    const size_t n = (size_t)a.nx * a.ny * a.nz;
    double* C = plane(a, ADJ_CON, n);
    for (size_t c = (size_t)blockIdx.x * ADJ_NT + threadIdx.x; c < n; c += (size_t)gridDim.x * ADJ_NT) {
        const Cell x = cell_of(c, a.nx, a.ny);
        double p0[9], acc[9], mhb[9], p0b[9], accb[9];
        load_p(a, n, c, p0);
        ld9(plane(a, ADJ_MHB, n), n, c, mhb);
        for (int v = 0; v < 9; ++v) { acc[v] = 0.0; p0b[v] = 0.0; }
        hancock_acc_load<SLOPE, 0>(a, n, x, p0, acc);
        hancock_acc_load<SLOPE, 1>(a, n, x, p0, acc);
        hancock_acc_load<SLOPE, 2>(a, n, x, p0, acc);
        if (hancock_core_from_acc(p0, acc, mhb, p0b, accb, a.half_dtdx, a.mh_revert, a.ph)) {
            hancock_dir_add<SLOPE, 0>(a, n, x, p0, accb, p0b);
            hancock_dir_add<SLOPE, 1>(a, n, x, p0, accb, p0b);
            hancock_dir_add<SLOPE, 2>(a, n, x, p0, accb, p0b);
        }
        double t[9];
        ld9(C, n, c, t);
        for (int v = 0; v < 9; ++v) t[v] += p0b[v];
        st9(C, n, c, t);
    }
}

// ---- pass 6 ----------------------------------------------------------------------------------
__global__ void __launch_bounds__(ADJ_NT) adj_gather_kernel(AdjArgs a)
{
    const size_t n = (size_t)a.nx * a.ny * a.nz;
    const double* C = plane(a, ADJ_CON, n);
    for (size_t c = (size_t)blockIdx.x * ADJ_NT + threadIdx.x; c < n; c += (size_t)gridDim.x * ADJ_NT) {
        const Cell x = cell_of(c, a.nx, a.ny);
        double wb[9], t[9], u[9], lb[9];
        ld9(C, n, c, wb);
        for (int d = 0; d < 3; ++d) {
            // the -d neighbour's +d contribution and the +d neighbour's -d contribution aim at c
            ld9(C + (size_t)(2 + 2 * d) * 9 * n, n, nbr(x, d, -1, a.nx, a.ny, a.nz), t);
            for (int v = 0; v < 9; ++v) wb[v] += t[v];
            ld9(C + (size_t)(1 + 2 * d) * 9 * n, n, nbr(x, d, +1, a.nx, a.ny, a.nz), t);
            for (int v = 0; v < 9; ++v) wb[v] += t[v];
        }
#pragma unroll
        for (int v = 0; v < 9; ++v) u[v] = a.u0[v][c];
        ld9(plane(a, ADJ_LPRE, n), n, c, lb);
        cons2prim_adj(u, wb, lb, a.ph);
#pragma unroll
        for (int v = 0; v < 9; ++v) a.lam0[v][c] = lb[v];
    }
}

// Sum of the per-block partials of pass 3, added to *acc.  One block: thread t sums partials
// t, t + 256, ... in order, then a fixed tree; deterministic for a given grid.
__global__ void __launch_bounds__(ADJ_NT) adj_fdot_kernel(const double* part, int nparts, double* acc)
{
    __shared__ double red[ADJ_NT];
    double s = 0.0;
    for (int b = threadIdx.x; b < nparts; b += ADJ_NT) s += part[b];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = ADJ_NT / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *acc += red[0];
}

}  // namespace glm
