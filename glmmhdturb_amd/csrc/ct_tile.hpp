// ct_tile.hpp -- the constrained-transport step as one kernel over staged 3-D tiles (the design of
// march_bridge/cu/spike_ct2.cu:147-210 and spike_ct3.cu:148-190: stage a tile, solve every face
// flux it needs once into a shared flux tile, then update the owned cells and their faces).
//
// The step is the one of ct_mhd.jl step! (:174-184) that glm_ct_step_passes runs as four passes and
// glm_ct_step as a z-march: the same point functions (ct_impl.hpp), the same order of operations,
// so the three agree to rounding.  Per workgroup of 256 lanes and an 8 x 8 x OZ tile of owned cells:
//   1. primitives with the cell-centred field of the tile grown by 3 cells below and 2 above on
//      every axis (the 4-cell flux stencil of the outermost faces) -> LDS P;
//   2. the three lower-face fluxes of every cell of the tile grown by one cell on each side (the
//      faces of the owned cells and of their edge EMFs), each face solved once -> LDS FL, keeping
//      the five hydro components and the two transverse induction components;
//   3. per owned cell: the hydro update from six face fluxes, the Balsara-Spicer edge EMFs from the
//      stored induction fluxes, the face-B curl update, the floor and OU kick.  Neighbouring tiles
//      recompute their shared faces and edges from the same global data with the same code, so an
//      edge EMF has one value wherever it is used and div B is kept to rounding across tile seams.
// LDS: P is 8 x 13 x 13 x (OZ + 5) and FL 3 x 7 x 10 x 10 x (OZ + 2) values: OZ = 4 in FP32
// (99 KB), OZ = 2 in FP64 (143 KB), one workgroup per CU.  Workgroup ids are dealt to XCDs in
// contiguous runs of tiles (xcd_block), so neighbouring tiles share their halo reads in one L2.
#pragma once
#include "ct_impl.hpp"
#include "gpu_util.hpp"

namespace glm {

constexpr int CTT_OX = 8, CTT_OY = 8, CTT_NT = 256;
template <typename T> struct CTTile { static constexpr int OZ = sizeof(T) == 8 ? 2 : 4; };

template <typename T>
struct CTTileDims {
    static constexpr int OZ = CTTile<T>::OZ;
    static constexpr int PX = CTT_OX + 5, PY = CTT_OY + 5, PZ = OZ + 5;   // primitives: cells -3 .. O+1
    static constexpr int NP = PX * PY * PZ;
    static constexpr int FX = CTT_OX + 2, FY = CTT_OY + 2, FZ = OZ + 2;   // fluxes: cells -1 .. O
    static constexpr int NF = FX * FY * FZ;
};

// Primitives (rho, v, p) with the cell-centred field of global cell (i, j, k) (ct_mhd.jl:85-96 bcc,
// :26-34 cellprim; the arithmetic of ct_prim_kernel).
template <typename T>
GLM_DEV void ct_cell_prim(const CTArgs<T>& a, int i, int j, int k, T q8[8])
{
    const size_t c = cidx(a, i, j, k);
    T bx = T(0.5) * (a.q[5][c] + a.q[5][cidx(a, cwrap(i + 1, a.nx), j, k)]);
    T by = T(0.5) * (a.q[6][c] + a.q[6][cidx(a, i, cwrap(j + 1, a.ny), k)]);
    T bz = T(0.5) * (a.q[7][c] + a.q[7][cidx(a, i, j, cwrap(k + 1, a.nz))]);
    T rho = tmax(a.q[0][c], a.smallr), ir = trcp(rho);
    T vx = a.q[1][c] * ir, vy = a.q[2][c] * ir, vz = a.q[3][c] * ir;
    T p = ct_pressure(a.q[4][c], rho, vx, vy, vz, bx, by, bz, a.gamma, a.smallp);
    q8[0] = rho; q8[1] = vx; q8[2] = vy; q8[3] = vz; q8[4] = p; q8[5] = bx; q8[6] = by; q8[7] = bz;
}

template <typename T, int D>
GLM_DEV void ctt_flux(const CTArgs<T>& a, const T (*P)[CTTileDims<T>::NP], int lp, size_t gc, T out[7])
{
    using Dm = CTTileDims<T>;
    constexpr int st = D == 0 ? 1 : (D == 1 ? Dm::PX : Dm::PX * Dm::PY);
    T m2[8], m1[8], c0[8], p1[8], f[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        m2[t] = P[t][lp - 2 * st]; m1[t] = P[t][lp - st]; c0[t] = P[t][lp]; p1[t] = P[t][lp + st];
    }
    ct_face_flux<T, D>(m2, m1, c0, p1, a.q[5 + D][gc], a, f);
#pragma unroll
    for (int t = 0; t < 5; ++t) out[t] = f[t];
    // the two induction components transverse to D: x face (By, Bz), y face (Bx, Bz), z face (Bx, By)
    out[5] = D == 0 ? f[6] : f[5];
    out[6] = D == 2 ? f[6] : f[7];
}

template <typename T>
__global__ void __launch_bounds__(CTT_NT) ct_tile_kernel(CTArgs<T> a, int tiles_x, int tiles_y, int tiles_z)
{
    using Dm = CTTileDims<T>;
    constexpr int OZ = Dm::OZ, PX = Dm::PX, PY = Dm::PY, NP = Dm::NP, FX = Dm::FX, FY = Dm::FY, NF = Dm::NF;
    __shared__ T P[8][NP];
    __shared__ T FL[3][7][NF];
    const int tid = threadIdx.x;
    int tile = xcd_block((int)blockIdx.x, (int)gridDim.x);
    const int bx = tile % tiles_x; tile /= tiles_x;
    const int by = tile % tiles_y;
    const int bz = tile / tiles_y;
    const int x0 = bx * CTT_OX, y0 = by * CTT_OY, z0 = bz * OZ;
    (void)tiles_z;

    // 1. primitives of cells x0-3 .. x0+O+1 (per axis), periodic
    for (int c = tid; c < NP; c += CTT_NT) {
        const int lx = c % PX, ly = (c / PX) % PY, lz = c / (PX * PY);
        T q8[8];
        ct_cell_prim(a, cwrap(x0 - 3 + lx, a.nx), cwrap(y0 - 3 + ly, a.ny), cwrap(z0 - 3 + lz, a.nz), q8);
#pragma unroll
        for (int t = 0; t < 8; ++t) P[t][c] = q8[t];
    }
    __syncthreads();

    // 2. lower-face fluxes of cells x0-1 .. x0+O (per axis), one solve per face; work item w covers
    //    direction w / NF, so a wavefront's items share their direction
    for (int w = tid; w < 3 * NF; w += CTT_NT) {
        const int d = w / NF, f = w - d * NF;
        const int fx = f % FX, fy = (f / FX) % FY, fz = f / (FX * FY);
        const int lp = (fx + 2) + PX * ((fy + 2) + PY * (fz + 2));
        const size_t gc = cidx(a, cwrap(x0 - 1 + fx, a.nx), cwrap(y0 - 1 + fy, a.ny), cwrap(z0 - 1 + fz, a.nz));
        T out[7];
        if (d == 0) ctt_flux<T, 0>(a, P, lp, gc, out);
        else if (d == 1) ctt_flux<T, 1>(a, P, lp, gc, out);
        else ctt_flux<T, 2>(a, P, lp, gc, out);
#pragma unroll
        for (int t = 0; t < 7; ++t) FL[d][t][f] = out[t];
    }
    __syncthreads();

    // 3. owned cells: hydro update, edge EMFs and the face-B curl (ct_update_kernel's arithmetic)
    const int tx = tid % CTT_OX, ty = (tid / CTT_OX) % CTT_OY, tz = tid / (CTT_OX * CTT_OY);
    const int i = x0 + tx, j = y0 + ty, k = z0 + tz;
    if (tz >= OZ || i >= a.nx || j >= a.ny || k >= a.nz) return;
    const int fc = (tx + 1) + FX * ((ty + 1) + FY * (tz + 1));
    const int fxp = fc + 1, fyp = fc + FX, fzp = fc + FX * FY;
    const int fxm = fc - 1, fym = fc - FX, fzm = fc - FX * FY;
    const size_t c = cidx(a, i, j, k);
    T r[5];
#pragma unroll
    for (int t = 0; t < 5; ++t)
        r[t] = a.q[t][c] + a.dtdx * ((FL[0][t][fc] - FL[0][t][fxp]) + (FL[1][t][fc] - FL[1][t][fyp]) +
                                     (FL[2][t][fc] - FL[2][t][fzp]));
    // ez(e) = (Fy[Bx](e) + Fy[Bx](e - x) - Fx[By](e) - Fx[By](e - y)) / 4, ex and ey cyclically (ct_mhd.jl:139-154)
    auto EZ = [&](int e) { return T(0.25) * (FL[1][5][e] + FL[1][5][e - 1] - FL[0][5][e] - FL[0][5][e - FX]); };
    auto EX = [&](int e) { return T(0.25) * (FL[2][6][e] + FL[2][6][e - FX] - FL[1][6][e] - FL[1][6][e - FX * FY]); };
    auto EY = [&](int e) { return T(0.25) * (FL[0][6][e] + FL[0][6][e - FX * FY] - FL[2][5][e] - FL[2][5][e - 1]); };
    (void)fxm; (void)fym; (void)fzm;
    const T ez = EZ(fc), ex = EX(fc), ey = EY(fc);
    T bxn = a.q[5][c] - a.dtdx * ((EZ(fyp) - ez) - (EY(fzp) - ey));
    T byn = a.q[6][c] - a.dtdx * ((EX(fzp) - ex) - (EZ(fxp) - ez));
    T bzn = a.q[7][c] - a.dtdx * ((EY(fxp) - ey) - (EX(fyp) - ex));
    const int lo = (tx + 3) + PX * ((ty + 3) + PY * (tz + 3));
    ct_epilogue(r, P[5][lo], P[6][lo], P[7][lo], i, j, k, a);
#pragma unroll
    for (int t = 0; t < 5; ++t) a.o[t][c] = r[t];
    a.o[5][c] = bxn; a.o[6][c] = byn; a.o[7][c] = bzn;
}

}  // namespace glm
