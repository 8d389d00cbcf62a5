// hydro_march.hpp -- the 5-variable hydro line march (spike_25d.cu:383-516 as build_march.sh
// builds it: -DHANCOCK1D, optional -DLLF_FALLBACK) as one z-march kernel for gfx950.
//
// Scheme per cell and direction d: primitive state (:196-202), MonCen slopes of the cells at
// -1, 0, +1 (:207-221), the normal-only Hancock half step (:287-301), face states at mh -/+ s/2
// clamped at SMALLR/SMALLP (:222-230), HLL (:302-330) or HLL with the LLF fallback (:335-376),
// div += F(+1/2) - F(-1/2) in the order x, y, z (:486-487), U1 = toCons(prim(U0)) - dtdx*div
// (:432, :492).  Periodic in all three axes.
//
// Layout: 5 SoA planes [rho, mx, my, mz, E], cell (i,j,k) at i + nx*(j + ny*k); with NSP = 2
// passive scalars (-DSCALARS, :115-119) two more planes of rho*s, advected as s (:191, :241,
// :251, :298), their flux the HLL/LLF flux of rho*s or, with CMA (-DCMA), the HLL mass flux times
// the upwind face value of s (:321, :367).
//
// Kernel shape.  A workgroup owns a 32 x 8 column tile and a chunk of z planes; lane (tx, ty)
// owns one column.  Each plane of the tile plus a 2-cell halo (36 x 12 cells) is read from HBM
// once, converted to primitives once, and kept in a ring of 4 LDS planes: plane k+3 is fetched
// before plane k is updated and stored after it (one barrier per plane).  The x and y stencils read plane k; the z
// stencil is the lane's own column of planes k-1..k+2, kept in registers as it passes through
// the ring.  The z face k+1/2 solved for plane k is the face k+1/2 of plane k+1, so it is
// carried, not solved again; the first z face of a chunk is solved by the same code in a priming
// iteration, so every chunking gives the same bits.  x and y faces are solved once per adjacent
// cell, as in the reference.
#pragma once
#include "gpu_util.hpp"

namespace glm {

constexpr int HM_OX = 32, HM_OY = 8, HM_G = 2;
constexpr int HM_LX = HM_OX + 2 * HM_G, HM_LY = HM_OY + 2 * HM_G;
constexpr int HM_NPL = HM_LX * HM_LY;   // 432 cells of a haloed plane
constexpr int HM_NT = HM_OX * HM_OY;    // 256 lanes, one per owned column

constexpr int HM_NVMAX = 7;   // 5 hydro variables + 2 passive scalars (-DSCALARS)

// Storage formats of the state in HBM (spike_25d.cu:64-95, the compact-storage builds):
//   HY_CONS   conserved planes of T (the build_march.sh library);
//   HY_F16P   -DF16STATE: primitive rho, vx, vy, vz, p as f16;
//   HY_U16P   -DU16STATE: primitive, rho and p as uint16 log2 over [-32, 32], velocities uint16
//             linear over [-8, 8];
//   HY_U16SP  -DU16SP: conserved hydro planes of T, the two species fractions as uint16 log2 over
//             [-110, 0], reconstructed in log2 (the CMA flux and the conserved rho X decode X).
// The compact formats compute in FP32.
enum { HY_CONS = 0, HY_F16P = 1, HY_U16P = 2, HY_U16SP = 3 };

template <typename T>
struct HydroArgs {
    const T* q[HM_NVMAX];
    T* o[HM_NVMAX];
    const unsigned short* qh[HM_NVMAX];   // 16-bit planes of the compact formats
    unsigned short* oh[HM_NVMAX];
    int nx, ny, nz;
    int tiles_x, tiles_y, nzc, zchunk;
    T gm1, igm1, gamma, dtdx, half_dtdx, smallr, smallp, fb_dmin, fb_pmin;
};

// uint16 codecs (spike_25d.cu:157-168): log2 over [-110, 0] (species), velocity linear over
// [-8, 8], log2 over [-32, 32] (rho, p); encode rounds to nearest and clamps.  Not contracted:
// each is a float multiply and a float add, so a codec is a fixed function of its input.
GLM_DEV float hy_dec_log2(unsigned short u)
{
#pragma clang fp contract(off)
    return -110.f + (float)u * (110.f / 65535.f);
}
GLM_DEV unsigned short hy_enc_log2(float l2)
{
#pragma clang fp contract(off)
    float t = (l2 + 110.f) * (65535.f / 110.f);
    t = fminf(fmaxf(t, 0.f), 65535.f);
    return (unsigned short)(t + 0.5f);
}
GLM_DEV float hy_dec_vel(unsigned short u)
{
#pragma clang fp contract(off)
    return -8.f + (float)u * (16.f / 65535.f);
}
GLM_DEV unsigned short hy_enc_vel(float x)
{
#pragma clang fp contract(off)
    const float t = (fminf(fmaxf(x, -8.f), 8.f) + 8.f) * (65535.f / 16.f);
    return (unsigned short)(t + 0.5f);
}
GLM_DEV float hy_dec_log2s(unsigned short u)
{
#pragma clang fp contract(off)
    return -32.f + (float)u * (64.f / 65535.f);
}
GLM_DEV unsigned short hy_enc_log2s(float l2)
{
#pragma clang fp contract(off)
    float t = (l2 + 32.f) * (65535.f / 64.f);
    t = fminf(fmaxf(t, 0.f), 65535.f);
    return (unsigned short)(t + 0.5f);
}
GLM_DEV float hy_f16_to_f32(unsigned short h) { return (float)__builtin_bit_cast(_Float16, h); }
GLM_DEV unsigned short hy_f32_to_f16(float x) { return __builtin_bit_cast(unsigned short, (_Float16)x); }
template <typename T>
GLM_DEV T hy_exp2(T x) { return (T)exp2f((float)x); }

// spike_25d.cu:196-202 prim_from_cons (:183-194 with scalars: s = (rho s) / rho; with LG the scalar
// planes already hold log2 X, :188-190)
template <typename T, int NV, bool LG = false>
GLM_DEV void hy_c2p(const T c[NV], T q[NV], const HydroArgs<T>& a)
{
#pragma clang fp contract(off)   // the same bits at every call site of the kernel
    const T r = tmax(c[0], a.smallr);
    const T ir = trcp(r);
    q[0] = r;
    q[1] = c[1] * ir;
    q[2] = c[2] * ir;
    q[3] = c[3] * ir;
    const T ke = T(0.5) * r * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    q[4] = tmax(a.gm1 * (c[4] - ke), a.smallp);
#pragma unroll
    for (int v = 5; v < NV; ++v) q[v] = LG ? c[v] : c[v] * ir;
}

// spike_25d.cu:207-213 moncen (full slope).  The selects stay: a select-free form (the sign of dc
// applied as a factor) measured 2-3% slower in this kernel (480^3 FP32).
template <typename T>
GLM_DEV T hy_moncen(T l, T m, T r)
{
    const T dl = m - l, dr = r - m, dc = T(0.5) * (r - l);
    const T s = tmin(tabs(dc), tmin(T(2) * tabs(dl), T(2) * tabs(dr)));
    return dl * dr <= T(0) ? T(0) : (dl > T(0) ? s : -s);
}

// spike_25d.cu:231-244 primFlux and :245-255 toCons (with LG the scalars are log2 X, decoded for the
// flux and the conserved rho X, :248-250)
template <typename T, int D, int NV, bool LG = false>
GLM_DEV void hy_flux_cons(const T q[NV], T f[NV], T u[NV], const HydroArgs<T>& a)
{
    const T un = q[1 + D];
    const T E = q[4] * a.igm1 + T(0.5) * q[0] * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    const T mu = q[0] * un;
    f[0] = mu;
#pragma unroll
    for (int c = 0; c < 3; ++c) f[1 + c] = q[0] * q[1 + c] * un + (c == D ? q[4] : T(0));
    f[4] = (E + q[4]) * un;
    u[0] = q[0];
#pragma unroll
    for (int c = 0; c < 3; ++c) u[1 + c] = q[0] * q[1 + c];
    u[4] = E;
#pragma unroll
    for (int v = 5; v < NV; ++v) {
        const T x = LG ? hy_exp2(q[v]) : q[v];
        f[v] = mu * x;
        u[v] = q[0] * x;
    }
}

// spike_25d.cu:287-301 hanc1d: normal-only Hancock half step along D
template <typename T, int D, int NV>
GLM_DEV void hy_hanc1d(const T m[NV], const T s[NV], T mh[NV], const HydroArgs<T>& a)
{
    const T un = m[1 + D], sn = s[1 + D];
    const T ir = trcp(m[0]);
    const T h = a.half_dtdx;
    mh[0] = m[0] - h * (un * s[0] + m[0] * sn);
#pragma unroll
    for (int c = 0; c < 3; ++c) mh[1 + c] = m[1 + c] - h * (un * s[1 + c] + (c == D ? s[4] * ir : T(0)));
    mh[4] = m[4] - h * (un * s[4] + a.gamma * m[4] * sn);
    const T hu = h * un;
#pragma unroll
    for (int v = 5; v < NV; ++v) mh[v] = m[v] - hu * s[v];   // passive advection (:298)
}

// spike_25d.cu:222-230 edge: mh + sign*s/2, clamped at SMALLR/SMALLP
template <typename T, int NV>
GLM_DEV void hy_edge(const T mh[NV], const T s[NV], T sign, T e[NV], const HydroArgs<T>& a)
{
    const T hs = sign * T(0.5);
#pragma unroll
    for (int v = 0; v < NV; ++v) e[v] = mh[v] + hs * s[v];
    e[0] = tmax(e[0], a.smallr);
    e[4] = tmax(e[4], a.smallp);
}

// HLL (spike_25d.cu:302-330); with FB, LLF where a face state is below the fallback floors
// (riemann_fb, :335-376).  Both combine the same FL, FR, UL, UR.  With CMA the scalar fluxes of
// the HLL fan are the mass flux times the upwind face value (:321, :367).
template <typename T, int D, bool FB, int NV, bool CMA, bool LG = false>
GLM_DEV void hy_riemann(const T L[NV], const T R[NV], T F[NV], const HydroArgs<T>& a)
{
    const T unL = L[1 + D], unR = R[1 + D];
    const T aL = tsqrt_pos(a.gamma * L[4] * trcp(L[0])), aR = tsqrt_pos(a.gamma * R[4] * trcp(R[0]));
    T FL[NV], FR[NV], UL[NV], UR[NV];
    hy_flux_cons<T, D, NV, LG>(L, FL, UL, a);
    hy_flux_cons<T, D, NV, LG>(R, FR, UR, a);
    const T sL = tmin(unL - aL, unR - aR), sR = tmax(unL + aL, unR + aR);
    const T inv = trcp(sR - sL), slsr = sL * sR;
    const bool llf = FB && (L[0] < a.fb_dmin || L[4] < a.fb_pmin || R[0] < a.fb_dmin || R[4] < a.fb_pmin);
    const T sm = tmax(tabs(unL) + aL, tabs(unR) + aR);
    const T hll0 = (sR * FL[0] - sL * FR[0] + slsr * (UR[0] - UL[0])) * inv;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        T hll = v == 0 ? hll0 : (sR * FL[v] - sL * FR[v] + slsr * (UR[v] - UL[v])) * inv;
        if (CMA && v >= 5) hll = hll0 >= T(0) ? hll0 * (LG ? hy_exp2(L[v]) : L[v]) : hll0 * (LG ? hy_exp2(R[v]) : R[v]);
        const T h = sL >= T(0) ? FL[v] : (sR <= T(0) ? FR[v] : hll);
        F[v] = llf ? T(0.5) * (FL[v] + FR[v]) - T(0.5) * sm * (UR[v] - UL[v]) : h;
    }
}

template <typename T, int NV>
GLM_DEV void hy_slopes(const T l[NV], const T m[NV], const T r[NV], T s[NV])
{
#pragma unroll
    for (int v = 0; v < NV; ++v) s[v] = hy_moncen(l[v], m[v], r[v]);
}

// Scheme variants of the march (spike_25d.cu build flags): HY_PLM the light march (-DHANCOCK1D:
// MonCen slopes, normal-only Hancock half step, edges clamped at SMALLR/SMALLP, :462-466); HY_PPM1D
// the parabolic-edge variant (-DPPM1D, :467-483: the monotonized 3-point parabola's edge plus the
// half step's change mh - c, no clamp, HLL without the fallback); HY_DONOR (-DDONOR, :438-454): the
// light march with donor-cell HLL of the two cell states at the x and y faces of the 32 x 8 tiles
// (a face at x = i - 1/2 with i mod 32 = 0, at y = j - 1/2 with j mod 8 = 0; z has none), HLL
// without the fallback.
enum { HY_PLM = 0, HY_PPM1D = 1, HY_DONOR = 2 };

// The upper (SIGN = +1) or lower (-1) face state of cell c with neighbours l, r, slope sl and half
// step mh: the light march's edge mh +- s/2 clamped at SMALLR/SMALLP, or the PPM1D parabola's edge
// plus mh - c, unclamped (spike_25d.cu:476-479).
template <typename T, int NV, int SV, int SIGN>
GLM_DEV void hy_state(const T l[NV], const T c[NV], const T r[NV], const T sl[NV], const T mh[NV], T e[NV],
                      const HydroArgs<T>& a)
{
    if (SV == HY_PPM1D) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            T dl, dr;   // the edge is c - dl or c + dr; plus mh - c
            ppm_offsets_spike(l[v], c[v], r[v], dl, dr);
            e[v] = SIGN > 0 ? mh[v] + dr : mh[v] - dl;
        }
    } else {
        hy_edge<T, NV>(mh, sl, T(SIGN), e, a);
    }
}

// Flux through the face between cells c1 and c2 of the stencil c0 c1 | c2 c3 along D.
template <typename T, int D, bool FB, int NV, bool CMA, bool LG = false, int SV = HY_PLM>
GLM_DEV void hy_face(const T c0[NV], const T c1[NV], const T c2[NV], const T c3[NV], T F[NV], const HydroArgs<T>& a)
{
    T s[NV], mh[NV], L[NV], R[NV];
    hy_slopes<T, NV>(c0, c1, c2, s);
    hy_hanc1d<T, D, NV>(c1, s, mh, a);
    hy_state<T, NV, SV, 1>(c0, c1, c2, s, mh, L, a);
    hy_slopes<T, NV>(c1, c2, c3, s);
    hy_hanc1d<T, D, NV>(c2, s, mh, a);
    hy_state<T, NV, SV, -1>(c1, c2, c3, s, mh, R, a);
    hy_riemann<T, D, FB && SV == HY_PLM, NV, CMA, LG>(L, R, F, a);
}

// div += F(+1/2) - F(-1/2) along D for the cell c0 of the stencil m2 m1 c0 p1 p2: the centre
// cell's slope and half step are shared by its two faces (spike_25d.cu:462-466 forms them once
// per face).  lob / hib: the lower / upper face is a tile seam (HY_DONOR: donor-cell HLL of the
// two cell states, :445-448).
template <typename T, int D, bool FB, int NV, bool CMA, bool LG = false, int SV = HY_PLM>
GLM_DEV void hy_dir(const T m2[NV], const T m1[NV], const T c0[NV], const T p1[NV], const T p2[NV], T div[NV],
                    const HydroArgs<T>& a, bool lob = false, bool hib = false)
{
    constexpr bool fb = FB && SV == HY_PLM;
    T s[NV], mh[NV], L[NV], R[NV], Fm[NV], Fp[NV];
    hy_slopes<T, NV>(m2, m1, c0, s);
    hy_hanc1d<T, D, NV>(m1, s, mh, a);
    hy_state<T, NV, SV, 1>(m2, m1, c0, s, mh, L, a);
    hy_slopes<T, NV>(m1, c0, p1, s);
    hy_hanc1d<T, D, NV>(c0, s, mh, a);
    hy_state<T, NV, SV, -1>(m1, c0, p1, s, mh, R, a);
    if (SV == HY_DONOR && lob) hy_riemann<T, D, false, NV, CMA, LG>(m1, c0, Fm, a);
    else hy_riemann<T, D, fb, NV, CMA, LG>(L, R, Fm, a);
    hy_state<T, NV, SV, 1>(m1, c0, p1, s, mh, L, a);
    hy_slopes<T, NV>(c0, p1, p2, s);
    hy_hanc1d<T, D, NV>(p1, s, mh, a);
    hy_state<T, NV, SV, -1>(c0, p1, p2, s, mh, R, a);
    if (SV == HY_DONOR && hib) hy_riemann<T, D, false, NV, CMA, LG>(c0, p1, Fp, a);
    else hy_riemann<T, D, fb, NV, CMA, LG>(L, R, Fp, a);
#pragma unroll
    for (int v = 0; v < NV; ++v) div[v] += Fp[v] - Fm[v];
}

// z chunking: the chunk length minimising (waves of workgroups) x (planes + priming) for
// `slots` resident workgroups, or the requested length.
inline void hydro_tiling(int nx, int ny, int nz, int zchunk_req, int slots, int& tiles_x, int& tiles_y, int& nzc,
                         int& zchunk)
{
    tiles_x = (nx + HM_OX - 1) / HM_OX;
    tiles_y = (ny + HM_OY - 1) / HM_OY;
    if (zchunk_req > 0) {
        zchunk = zchunk_req < nz ? zchunk_req : nz;
        nzc = (nz + zchunk - 1) / zchunk;
        return;
    }
    const long cols = (long)tiles_x * tiles_y;
    const long sl = slots > 0 ? slots : 256;
    const int maxc = nz / 8 > 1 ? nz / 8 : 1;
    long best = -1;
    zchunk = nz;
    for (int c = 1; c <= maxc; ++c) {
        const int zc = (nz + c - 1) / c;
        const int n = (nz + zc - 1) / zc;
        if (n != c) continue;
        const long cost = ((cols * n + sl - 1) / sl) * (long)(zc + 2);
        if (best < 0 || cost < best) { best = cost; zchunk = zc; }
    }
    nzc = (nz + zchunk - 1) / zchunk;
}

// Resident waves per SIMD the register allocation is held to: as many as the LDS ring allows
// (FP64: 2 workgroups per CU of 69 KB, 1 of 97 KB with scalars; FP32: 4 of 35 KB), except FP32 with
// scalars, where 3 workgroups of 48 KB would hold it to 168 registers and spill: 2 workgroups and
// no spills run 1.4x faster (480^3, measured).
// The PPM1D and DONOR builds in FP32: 3 waves per SIMD without scalars (4 would hold them to 128
// registers and spill), 1 with scalars (PPM1D with scalars at 1: 20.8 Gcell/s, at 2: 19.9; 480^3).
#ifndef GLM_HYDRO_SV_WAVES0
#define GLM_HYDRO_SV_WAVES0 3
#endif
#ifndef GLM_HYDRO_SV_WAVES2
#define GLM_HYDRO_SV_WAVES2 1
#endif
#define GLM_HYDRO_WAVES_PER_EU(T, NSP, SV) \
    (sizeof(T) == 8 ? ((NSP) ? 1 : 2) : (SV) == HY_PLM ? ((NSP) ? 2 : 4) : ((NSP) ? GLM_HYDRO_SV_WAVES2 : GLM_HYDRO_SV_WAVES0))

template <typename T, bool FB, int NSP, bool CMA, int FMT = HY_CONS, int SV = HY_PLM>
__global__ void __launch_bounds__(HM_NT) __attribute__((amdgpu_waves_per_eu(GLM_HYDRO_WAVES_PER_EU(T, NSP, SV))))
hydro_march_kernel(HydroArgs<T> a)
{
    constexpr int NV = 5 + NSP;
    constexpr bool LG = FMT == HY_U16SP;      // species held as log2 X
    constexpr bool PRIM = FMT == HY_F16P || FMT == HY_U16P;   // primitive storage: no conversion on load
    static_assert(FMT == HY_CONS || (sizeof(T) == 4 && (FMT != HY_U16SP || (NSP == 2 && CMA)) &&
                                     (FMT == HY_U16SP || NSP == 0)), "compact formats: FP32; U16SP with 2 CMA species");
    // ring of 4 haloed primitive planes: plane j lives in buffer (j - k0 + 2) & 3
    __shared__ T P[4][NV][HM_NPL];
    const int tid = threadIdx.x;
    const int tx = tid % HM_OX, ty = tid / HM_OX;
    const int ntile = a.tiles_x * a.tiles_y;
    const int w = xcd_block((int)blockIdx.x, (int)gridDim.x);   // neighbouring tiles share an XCD's L2
    const int tile = w % ntile, zc = w / ntile;
    const int x0 = (tile % a.tiles_x) * HM_OX, y0 = (tile / a.tiles_x) * HM_OY;
    const int k0 = zc * a.zchunk;
    const int k1 = k0 + a.zchunk < a.nz ? k0 + a.zchunk : a.nz;
    const size_t plane = (size_t)a.nx * a.ny;
    // the lane's column; lanes past the grid edge compute a wrapped column and store nothing
    const int gx = x0 + tx, gy = y0 + ty;
    const bool own = gx < a.nx && gy < a.ny;
    const size_t col = (size_t)wrapi(gx, a.nx) + (size_t)a.nx * wrapi(gy, a.ny);
    const int li = (ty + HM_G) * HM_LX + tx + HM_G;

    // every primitive value of the kernel comes from hy_c2p, which is not contracted, so a cell
    // has the same bits in every stencil and every chunking.  A plane is fetched from HBM into
    // registers (`pre`, two cells per lane) at the top of an iteration and converted into its
    // LDS buffer at the bottom, so the fetch latency is hidden behind the plane's compute.
    T pre[2][NV];
    auto fetch = [&](int k) {
        const size_t zoff = (size_t)wrapi(k, a.nz) * plane;
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const int c = tid + r * HM_NT;
            if (c < HM_NPL) {
                const int lx = c % HM_LX, ly = c / HM_LX;
                const size_t g = zoff + wrapi(x0 - HM_G + lx, a.nx) + (size_t)a.nx * wrapi(y0 - HM_G + ly, a.ny);
                if (FMT == HY_CONS) {
#pragma unroll
                    for (int v = 0; v < NV; ++v) pre[r][v] = gload(a.q[v] + g);
                } else if (FMT == HY_F16P) {
#pragma unroll
                    for (int v = 0; v < 5; ++v) pre[r][v] = (T)hy_f16_to_f32(gload(a.qh[v] + g));
                } else if (FMT == HY_U16P) {   // spike_25d.cu:397-402
                    pre[r][0] = (T)exp2f(hy_dec_log2s(gload(a.qh[0] + g)));
#pragma unroll
                    for (int v = 1; v < 4; ++v) pre[r][v] = (T)hy_dec_vel(gload(a.qh[v] + g));
                    pre[r][4] = (T)exp2f(hy_dec_log2s(gload(a.qh[4] + g)));
                } else {                       // HY_U16SP, :409-411
#pragma unroll
                    for (int v = 0; v < 5; ++v) pre[r][v] = gload(a.q[v] + g);
#pragma unroll
                    for (int v = 5; v < NV; ++v) pre[r][v] = (T)hy_dec_log2(gload(a.qh[v] + g));
                }
            }
        }
    };
    auto commit = [&](int k) {
        const int b = (k - k0 + 2) & 3;
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const int c = tid + r * HM_NT;
            if (c < HM_NPL) {
                T qq[NV];
                if (PRIM) {   // make_prim (spike_25d.cu:204-206)
#pragma unroll
                    for (int v = 0; v < NV; ++v) qq[v] = pre[r][v];
                    qq[0] = tmax(qq[0], a.smallr);
                    qq[4] = tmax(qq[4], a.smallp);
                } else {
                    hy_c2p<T, NV, LG>(pre[r], qq, a);
                }
#pragma unroll
                for (int v = 0; v < NV; ++v) P[b][v][c] = qq[v];
            }
        }
    };
    auto load_plane = [&](int k) {
        fetch(k);
        commit(k);
    };
    auto own_cell = [&](int k, T qq[NV]) {
        const int b = (k - k0 + 2) & 3;
#pragma unroll
        for (int v = 0; v < NV; ++v) qq[v] = P[b][v][li];
    };

    T W[4][NV];   // the lane's column at planes k-1 .. k+2 (iteration k)
    T Fz[NV];     // z flux through face k-1/2
    load_plane(k0 - 2);
    load_plane(k0 - 1);
    load_plane(k0);
    load_plane(k0 + 1);
    __syncthreads();
    own_cell(k0 - 2, W[1]);
    own_cell(k0 - 1, W[2]);
    own_cell(k0, W[3]);
    __syncthreads();   // plane k0-2's buffer is reloaded in the first iteration

    // iteration k0-1 only computes the z face k0-1/2 (with the same code as every later face)
#pragma clang loop unroll(disable)
    for (int k = k0 - 1; k < k1; ++k) {
        const bool more = k + 3 <= k1 + 1;   // planes up to k1+1 are read
        if (more) fetch(k + 3);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            W[0][v] = W[1][v];
            W[1][v] = W[2][v];
            W[2][v] = W[3][v];
        }
        own_cell(k + 2, W[3]);
        T div[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) div[v] = T(0);
        const bool upd = k >= k0;
        if (upd) {
            const int b = (k - k0 + 2) & 3;
            T m2[NV], m1[NV], c0[NV], p1[NV], p2[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                m2[v] = P[b][v][li - 2]; m1[v] = P[b][v][li - 1]; c0[v] = P[b][v][li];
                p1[v] = P[b][v][li + 1]; p2[v] = P[b][v][li + 2];
            }
            // tile seams of HY_DONOR: faces at x = i - 1/2 with i mod 32 = 0, y = j - 1/2 with j mod 8 = 0
            // (the periodic face after the last cell included)
            const bool lox = SV == HY_DONOR && gx % HM_OX == 0, hix = SV == HY_DONOR && ((gx + 1) % a.nx) % HM_OX == 0;
            const bool loy = SV == HY_DONOR && gy % HM_OY == 0, hiy = SV == HY_DONOR && ((gy + 1) % a.ny) % HM_OY == 0;
            hy_dir<T, 0, FB, NV, CMA, LG, SV>(m2, m1, c0, p1, p2, div, a, lox, hix);
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                m2[v] = P[b][v][li - 2 * HM_LX]; m1[v] = P[b][v][li - HM_LX];
                p1[v] = P[b][v][li + HM_LX]; p2[v] = P[b][v][li + 2 * HM_LX];
            }
            hy_dir<T, 1, FB, NV, CMA, LG, SV>(m2, m1, c0, p1, p2, div, a, loy, hiy);
        }
        T Fp[NV];
        hy_face<T, 2, FB, NV, CMA, LG, SV>(W[0], W[1], W[2], W[3], Fp, a);   // face k+1/2
        if (upd) {
#pragma unroll
            for (int v = 0; v < NV; ++v) div[v] += Fp[v] - Fz[v];
            T U[NV], f[NV];
            hy_flux_cons<T, 0, NV, LG>(W[1], f, U, a);   // toCons(prim(U0)); the flux is unused
            if (own) {
                const size_t g = (size_t)k * plane + col;
#pragma unroll
                for (int v = 0; v < NV; ++v) U[v] = U[v] - a.dtdx * div[v];
                if (FMT == HY_CONS) {
#pragma unroll
                    for (int v = 0; v < NV; ++v) a.o[v][g] = U[v];
                } else if (PRIM) {   // c2p -> store primitive (spike_25d.cu:499-511)
                    const T ir = T(1) / tmax(U[0], a.smallr);
                    const T vx = U[1] * ir, vy = U[2] * ir, vz = U[3] * ir;
                    const T ke = T(0.5) * U[0] * (vx * vx + vy * vy + vz * vz);
                    const T pp = tmax(a.gm1 * (U[4] - ke), a.smallp);
                    if (FMT == HY_F16P) {
                        a.oh[0][g] = hy_f32_to_f16((float)U[0]);
                        a.oh[1][g] = hy_f32_to_f16((float)vx);
                        a.oh[2][g] = hy_f32_to_f16((float)vy);
                        a.oh[3][g] = hy_f32_to_f16((float)vz);
                        a.oh[4][g] = hy_f32_to_f16((float)pp);
                    } else {
                        a.oh[0][g] = hy_enc_log2s(log2f((float)tmax(U[0], a.smallr)));
                        a.oh[1][g] = hy_enc_vel((float)vx);
                        a.oh[2][g] = hy_enc_vel((float)vy);
                        a.oh[3][g] = hy_enc_vel((float)vz);
                        a.oh[4][g] = hy_enc_log2s(log2f((float)pp));
                    }
                } else {             // HY_U16SP: X_new = (rho X)_new / rho_new -> log2 -> uint16 (:514-517)
#pragma unroll
                    for (int v = 0; v < 5; ++v) a.o[v][g] = U[v];
                    const T rinv = T(1) / tmax(U[0], a.smallr);
#pragma unroll
                    for (int v = 5; v < NV; ++v) a.oh[v][g] = hy_enc_log2(log2f(fmaxf((float)(U[v] * rinv), 7.7e-34f)));
                }
            }
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) Fz[v] = Fp[v];
        if (more) commit(k + 3);   // into plane k-1's buffer, last read before the previous barrier
        __syncthreads();
    }
}

}  // namespace glm
