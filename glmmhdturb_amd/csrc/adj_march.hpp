// adj_march.hpp -- the back half of the reverse step as one z-march kernel (FP64).
//
// It replaces passes 5 and 6 of adj_impl.hpp (the face-state and predictor adjoints of every cell,
// their contributions to the primitive adjoints of the cell and its six neighbours, the gather and
// the c2p adjoint) with one pass that keeps those contributions on chip instead of writing and
// re-reading 63 + 9 contribution planes.  Inputs are the planes the front passes wrote: W = c2p(u0),
// the predicted states MH, the face-state adjoints LFS and LPRE; the output is lam0.
//
// Layout: the forward march's tile (march_impl.hpp): 248 lanes, a 16 x 12 owned tile and its
// one-cell ring, each lane one column, marching z; W of plane q over the tile grown by two cells in
// LDS.  Per plane q every lane takes the face-state adjoints of its cell in x, y, z and the
// predictor adjoint; the contributions to its in-plane neighbours go through LDS (one slot per
// direction, read back by the neighbour after a barrier), the one to the cell below finishes plane
// q-1 and the one to the cell above is carried to plane q+1.  The sums run in the order of the
// multi-pass gather, so both paths produce the same bits.
#pragma once
#include "adj_impl.hpp"
#include "gpu_util.hpp"
#include "march_impl.hpp"

namespace glm {

struct AdjBackLDS {
    double P[9][NPL];       // W of plane q (two-cell ring)
    double SX[2][9][NPL];   // each lane's contributions to its -x (0) and +x (1) neighbours
    double SY[2][9][NPL];   // and to its -y (0) and +y (1) neighbours
};

template <int PRED, int RECON, int SLOPE>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(1)))
adj_back_kernel(AdjArgs a, int tiles_x, int tiles_y, int zchunk)
{
    __shared__ AdjBackLDS L;
    double (&P)[9][NPL] = L.P;
    const int tid = threadIdx.x;
    int tile = xcd_block((int)blockIdx.x, (int)gridDim.x);
    int tx, ty;
    lane_cell(tid, tx, ty);
    const int tile_x = tile % tiles_x; tile /= tiles_x;
    const int tile_y = tile % tiles_y; tile /= tiles_y;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const int zs = tile * zchunk;
    const int ze = min(zs + zchunk, nz);
    const size_t n = (size_t)nx * ny * nz, pl = (size_t)nx * ny;
    const int gxu = tile_x * OX - 1 + tx, gyu = tile_y * OY - 1 + ty;
    const int gx = wrapi(gxu, nx), gy = wrapi(gyu, ny);
    const bool owned = tid < NO && gxu < nx && gyu < ny;
    const size_t col = (size_t)gx + (size_t)nx * gy;
    const int lo = (ty + 1) * PX + (tx + 1);
    const bool helper = tid < NHELP;
    int hx = 0, hy = 0;
    helper_cell(helper ? tid : 0, hx, hy);
    const size_t hcol = (size_t)wrapi(tile_x * OX - 2 + hx, nx) + (size_t)nx * wrapi(tile_y * OY - 2 + hy, ny);
    const int lo_ring = hy * PX + hx;
    const double* W = plane(a, ADJ_W, n);
    const FaceOpts o{a.mh_revert, a.face_guard};
    auto ldW = [&](size_t c, double w[9]) {
#pragma unroll
        for (int v = 0; v < 9; ++v) w[v] = gload(W + (size_t)v * n + c);
    };

    // W window: planes q-1 and q+1 in registers, plane q in LDS
    double pm[9], pp[9];
    {
        double p0[9];
        ldW(col + pl * wrapi(zs - 2, nz), pm);
        ldW(col + pl * wrapi(zs - 1, nz), p0);
        ldW(col + pl * wrapi(zs, nz), pp);
#pragma unroll
        for (int v = 0; v < 9; ++v) P[v][lo] = p0[v];
        if (helper) {
            double h[9];
            ldW(hcol + pl * wrapi(zs - 1, nz), h);
#pragma unroll
            for (int v = 0; v < 9; ++v) P[v][lo_ring] = h[v];
        }
    }
    __syncthreads();

    double part[9], szp[9];   // Wb of plane q-1 without plane q's share; plane q-1's share for plane q
#pragma unroll
    for (int v = 0; v < 9; ++v) { part[v] = 0.0; szp[v] = 0.0; }

    for (int q = zs - 1; q <= ze; ++q) {
        const size_t c = col + pl * wrapi(q, nz);
        const bool last = q == ze;
        // next planes, requested ahead of the cell's work
        double pn[9], hn[9];
        if (!last) {
            ldW(col + pl * wrapi(q + 2, nz), pn);
            if (helper) ldW(hcol + pl * wrapi(q + 1, nz), hn);
        }
        double p0[9];
#pragma unroll
        for (int v = 0; v < 9; ++v) p0[v] = P[v][lo];
        bool up = true;   // PPM strong-pressure-jump flag (glmmhd_turb.jl:959), stencil order of ppm_up
        if (RECON == RECON_PPM)
            up = !(spj(P[4][lo - 1], p0[4], P[4][lo + 1]) || spj(P[4][lo - PX], p0[4], P[4][lo + PX]) ||
                   spj(pm[4], p0[4], pp[4]));
        double mh[9];
        if (PRED == PRED_FULL) ld9(plane(a, ADJ_MH, n), n, c, mh);
        else {
#pragma unroll
            for (int v = 0; v < 9; ++v) mh[v] = 0.0;
        }

        // face-state adjoints, one direction at a time (adj_cell_dir_kernel's sums)
        double s0[9], mhb[9], szm[9], szn[9];
        {
            double l[9], r[9], lob[9], hib[9], lb[9], m0b[9], rb[9], mb[9];
#pragma unroll
            for (int v = 0; v < 9; ++v) {
                l[v] = P[v][lo - 1]; r[v] = P[v][lo + 1];
                lb[v] = 0.0; m0b[v] = 0.0; rb[v] = 0.0; mb[v] = 0.0;
            }
            ld9(plane(a, ADJ_LFS + fs_slot(0, 0), n), n, c, lob);
            ld9(plane(a, ADJ_LFS + fs_slot(0, 1), n), n, c, hib);
            faces_adj<double, PRED, RECON, SLOPE, 0>(l, p0, r, mh, up, a.half_dtdx, o, a.ph, lob, hib, lb, m0b, rb, mb);
#pragma unroll
            for (int v = 0; v < 9; ++v) { L.SX[0][v][lo] = lb[v]; L.SX[1][v][lo] = rb[v]; s0[v] = m0b[v]; mhb[v] = mb[v]; }
        }
        {
            double l[9], r[9], lob[9], hib[9], lb[9], m0b[9], rb[9], mb[9];
#pragma unroll
            for (int v = 0; v < 9; ++v) {
                l[v] = P[v][lo - PX]; r[v] = P[v][lo + PX];
                lb[v] = 0.0; m0b[v] = 0.0; rb[v] = 0.0; mb[v] = 0.0;
            }
            ld9(plane(a, ADJ_LFS + fs_slot(1, 0), n), n, c, lob);
            ld9(plane(a, ADJ_LFS + fs_slot(1, 1), n), n, c, hib);
            faces_adj<double, PRED, RECON, SLOPE, 1>(l, p0, r, mh, up, a.half_dtdx, o, a.ph, lob, hib, lb, m0b, rb, mb);
#pragma unroll
            for (int v = 0; v < 9; ++v) {
                L.SY[0][v][lo] = lb[v]; L.SY[1][v][lo] = rb[v];
                s0[v] = m0b[v] + s0[v]; mhb[v] = mb[v] + mhb[v];
            }
        }
        {
            double lob[9], hib[9], m0b[9], mb[9];
#pragma unroll
            for (int v = 0; v < 9; ++v) { szm[v] = 0.0; m0b[v] = 0.0; szn[v] = 0.0; mb[v] = 0.0; }
            ld9(plane(a, ADJ_LFS + fs_slot(2, 0), n), n, c, lob);
            ld9(plane(a, ADJ_LFS + fs_slot(2, 1), n), n, c, hib);
            faces_adj<double, PRED, RECON, SLOPE, 2>(pm, p0, pp, mh, up, a.half_dtdx, o, a.ph, lob, hib, szm, m0b, szn, mb);
#pragma unroll
            for (int v = 0; v < 9; ++v) { s0[v] = m0b[v] + s0[v]; mhb[v] = mb[v] + mhb[v]; }
        }

        // predictor adjoint (adj_hancock_kernel's sums: each direction's slots plus its share)
        if (PRED == PRED_FULL) {
            double acc[9], p0b[9], accb[9];
#pragma unroll
            for (int v = 0; v < 9; ++v) { acc[v] = 0.0; p0b[v] = 0.0; }
            {
                double l[9], r[9];
#pragma unroll
                for (int v = 0; v < 9; ++v) { l[v] = P[v][lo - 1]; r[v] = P[v][lo + 1]; }
                hancock_acc_dir<double, SLOPE, 0>(p0, l, r, acc, a.ph);
#pragma unroll
                for (int v = 0; v < 9; ++v) { l[v] = P[v][lo - PX]; r[v] = P[v][lo + PX]; }
                hancock_acc_dir<double, SLOPE, 1>(p0, l, r, acc, a.ph);
                hancock_acc_dir<double, SLOPE, 2>(p0, pm, pp, acc, a.ph);
            }
            if (hancock_core_from_acc(p0, acc, mhb, p0b, accb, a.half_dtdx, a.mh_revert, a.ph)) {
                double l[9], r[9], lb[9], rb[9];
#pragma unroll
                for (int v = 0; v < 9; ++v) { l[v] = P[v][lo - 1]; r[v] = P[v][lo + 1]; lb[v] = 0.0; rb[v] = 0.0; }
                hancock_adj_dir<double, SLOPE, 0>(p0, l, r, accb, p0b, lb, rb, a.ph);
#pragma unroll
                for (int v = 0; v < 9; ++v) { L.SX[0][v][lo] += lb[v]; L.SX[1][v][lo] += rb[v]; }
#pragma unroll
                for (int v = 0; v < 9; ++v) { l[v] = P[v][lo - PX]; r[v] = P[v][lo + PX]; lb[v] = 0.0; rb[v] = 0.0; }
                hancock_adj_dir<double, SLOPE, 1>(p0, l, r, accb, p0b, lb, rb, a.ph);
#pragma unroll
                for (int v = 0; v < 9; ++v) { L.SY[0][v][lo] += lb[v]; L.SY[1][v][lo] += rb[v]; lb[v] = 0.0; rb[v] = 0.0; }
                hancock_adj_dir<double, SLOPE, 2>(p0, pm, pp, accb, p0b, lb, rb, a.ph);
#pragma unroll
                for (int v = 0; v < 9; ++v) { szm[v] += lb[v]; szn[v] += rb[v]; }
            }
#pragma unroll
            for (int v = 0; v < 9; ++v) s0[v] += p0b[v];
        }

        // plane q-1 is complete with this plane's share: gather order, then the c2p adjoint
        if (owned && q >= zs + 1) {
            const size_t cu = col + pl * wrapi(q - 1, nz);
            double wb[9], u[9], lb[9];
#pragma unroll
            for (int v = 0; v < 9; ++v) {
                wb[v] = part[v] + szm[v];
                u[v] = gload(a.u0[v] + cu);
            }
            ld9(plane(a, ADJ_LPRE, n), n, cu, lb);
            cons2prim_adj(u, wb, lb, a.ph);
#pragma unroll
            for (int v = 0; v < 9; ++v) a.lam0[v][cu] = lb[v];
        }
        __syncthreads();   // the in-plane slots of plane q are written; plane q's LDS cells are read

        if (owned && q >= zs && q < ze) {
            // Wb of plane q from its own slot, the four in-plane neighbours and the cell below, in
            // the multi-pass gather's order (adj_gather_kernel)
#pragma unroll
            for (int v = 0; v < 9; ++v) {
                double w = s0[v];
                w += L.SX[1][v][lo - 1];
                w += L.SX[0][v][lo + 1];
                w += L.SY[1][v][lo - PX];
                w += L.SY[0][v][lo + PX];
                w += szp[v];
                part[v] = w;
            }
        }
#pragma unroll
        for (int v = 0; v < 9; ++v) szp[v] = szn[v];
        if (last) break;

        // publish plane q+1
#pragma unroll
        for (int v = 0; v < 9; ++v) { pm[v] = p0[v]; P[v][lo] = pp[v]; pp[v] = pn[v]; }
        if (helper) {
#pragma unroll
            for (int v = 0; v < 9; ++v) P[v][lo_ring] = hn[v];
        }
        __syncthreads();
    }
}

}  // namespace glm
