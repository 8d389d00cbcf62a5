// cpu_stream.cpp -- restatement of cpu_simd.jl's chunked z-stream CPU integrator
// (cpu_simd.jl:440-601: ChunkBufs, recon1!, faces1!, fzupd1!, cpu_stream_step!), the reference's
// fastest CPU path, used as the timed CPU baseline.
//
// TEST INFRASTRUCTURE ONLY.  Built into oracle/_build/libcpustream.so and called by tests/ and by
// bench.py's cpu_baseline leg; the product (glmmhdturb_amd/) never links it.
//
// The algorithm, as the reference runs it:
//   * the state is padded by one ghost column on each side in x (NP = N + 2), refreshed
//     periodically at the start of a step (cpu_simd.jl:571-573), so x-neighbour loads never wrap;
//     y and z wrap through row indices;
//   * one OpenMP region per step; thread c streams the contiguous k-chunk [c0, c1] (crange,
//     cpu_simd.jl:453) with its own plane buffers, first-touched by that thread
//     (alloc_chunkbufs, :542-549), recomputing one halo plane at each end;
//   * per plane kk = c0-1 .. c1+1: recon1! reconstructs the plane (7 cons2prim per cell, MonCen
//     half slopes, the full Hancock predictor with the revert, :462-490); faces1! solves both x
//     and both y faces of every cell (each in-plane face twice, once from each side) and forms the
//     in-plane flux difference and the two z face states (:492-514); fzupd1! solves the z face
//     kk-1/2 from the +z state of plane kk-1 and the -z state of plane kk, and updates plane kk-1
//     with the void floor and the psi decay (:516-540);
//   * SIMD across x: every x loop advances W cells at a time, the physics running on W-lane
//     vectors (8 floats or 4 doubles in one AVX2 register, the reference's Vec{8,Float32}) with
//     lane blends in place of branches (the reference's vifelse); N must be a multiple of W.
// Two schemes: the literal cpu_simd.jl one (full Hancock, PLM MonCen, LLF, no face guard) and
// the GPU benchmark's (full Hancock, PPM with the strong-jump PLM fallback, HLLD, iflo face
// guard).  Any precision: float (the reference's) or double.
//
// Physics follows glmmhd_turb.jl:61-367 (the same formulas as oracle/glm_oracle.c, re-written here
// as inline templates over scalars or lanes); tests/test_cpu_stream.py checks the steps against the
// oracle.
#include <immintrin.h>
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <vector>

namespace {

constexpr int NV = 9;

// W lanes of S across x: 8 floats or 4 doubles, one AVX2 register (the reference's Vec{8,Float32}).
template <typename S> struct VT;
template <> struct VT<float> { typedef float V __attribute__((vector_size(32))); typedef int M __attribute__((vector_size(32))); };
template <> struct VT<double> { typedef double V __attribute__((vector_size(32))); typedef long long M __attribute__((vector_size(32))); };

template <typename S>
struct Vec {
    static constexpr int W = 32 / sizeof(S);
    typedef typename VT<S>::V V;
    V x;
    Vec() = default;
    Vec(S c) { for (int i = 0; i < W; ++i) x[i] = c; }
    static Vec mk(V v) { Vec r; r.x = v; return r; }
    static Vec load(const S* p) { Vec r; std::memcpy(&r.x, p, sizeof(V)); return r; }
    void store(S* p) const { std::memcpy(p, &x, sizeof(V)); }
};
template <typename S>
struct Msk {
    typename VT<S>::M m;
};
template <typename S> inline Vec<S> operator+(Vec<S> a, Vec<S> b) { return Vec<S>::mk(a.x + b.x); }
template <typename S> inline Vec<S> operator-(Vec<S> a, Vec<S> b) { return Vec<S>::mk(a.x - b.x); }
template <typename S> inline Vec<S> operator*(Vec<S> a, Vec<S> b) { return Vec<S>::mk(a.x * b.x); }
template <typename S> inline Vec<S> operator/(Vec<S> a, Vec<S> b) { return Vec<S>::mk(a.x / b.x); }
template <typename S> inline Vec<S> operator-(Vec<S> a) { return Vec<S>::mk(-a.x); }
template <typename S> inline Msk<S> operator<(Vec<S> a, Vec<S> b) { return Msk<S>{a.x < b.x}; }
template <typename S> inline Msk<S> operator<=(Vec<S> a, Vec<S> b) { return Msk<S>{a.x <= b.x}; }
template <typename S> inline Msk<S> operator>(Vec<S> a, Vec<S> b) { return Msk<S>{a.x > b.x}; }
template <typename S> inline Msk<S> operator>=(Vec<S> a, Vec<S> b) { return Msk<S>{a.x >= b.x}; }
template <typename S> inline Msk<S> operator|(Msk<S> a, Msk<S> b) { return Msk<S>{a.m | b.m}; }
template <typename S> inline Msk<S> operator!(Msk<S> a) { return Msk<S>{~a.m}; }

// the same helpers on lanes (the physics below runs on Vec<S> only)
template <typename S> inline Vec<S> mx(Vec<S> a, Vec<S> b) { return Vec<S>::mk(a.x < b.x ? b.x : a.x); }
template <typename S> inline Vec<S> mn(Vec<S> a, Vec<S> b) { return Vec<S>::mk(b.x < a.x ? b.x : a.x); }
template <typename S> inline Vec<S> ab(Vec<S> a) { return Vec<S>::mk(a.x < S(0) ? -a.x : a.x); }
inline Vec<float> sq(Vec<float> a) { return Vec<float>::mk((Vec<float>::V)_mm256_sqrt_ps((__m256)a.x)); }
inline Vec<double> sq(Vec<double> a) { return Vec<double>::mk((Vec<double>::V)_mm256_sqrt_pd((__m256d)a.x)); }
template <typename S> inline Vec<S> sel(Msk<S> m, Vec<S> a, Vec<S> b) { return Vec<S>::mk(m.m ? a.x : b.x); }

template <typename T>
struct Sc {
    T gamma, smallr, pfl, dtdx, half_dtdx, ch, glm_fac;
};

template <typename T>
inline void c2p(const T* c, T* q, const Sc<T>& s)
{
    T rho = mx(c[0], s.smallr), ir = T(1) / rho;
    T ekin = T(0.5) * (c[1] * c[1] + c[2] * c[2] + c[3] * c[3]) * ir;
    T emag = T(0.5) * (c[5] * c[5] + c[6] * c[6] + c[7] * c[7]);
    q[0] = rho; q[1] = c[1] * ir; q[2] = c[2] * ir; q[3] = c[3] * ir;
    q[4] = mx((s.gamma - T(1)) * (c[4] - ekin - emag), s.pfl);
    q[5] = c[5]; q[6] = c[6]; q[7] = c[7]; q[8] = c[8];
}

template <typename T>
inline void p2c(const T* q, T* c, T gamma)
{
    T ekin = T(0.5) * q[0] * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    T emag = T(0.5) * (q[5] * q[5] + q[6] * q[6] + q[7] * q[7]);
    c[0] = q[0]; c[1] = q[0] * q[1]; c[2] = q[0] * q[2]; c[3] = q[0] * q[3];
    c[4] = q[4] / (gamma - T(1)) + ekin + emag;
    c[5] = q[5]; c[6] = q[6]; c[7] = q[7]; c[8] = q[8];
}

// ideal-MHD flux along axis D of a primitive state, F[Bn] = F[psi] = 0 (glmmhd_turb.jl:98-162)
template <typename T, int D>
inline void flux(const T* q, T* f, T gamma)
{
    constexpr int N1 = 1 + D, N2 = 1 + (D + 1) % 3, N3 = 1 + (D + 2) % 3;
    T rho = q[0], un = q[N1], u2 = q[N2], u3 = q[N3], p = q[4];
    T bn = q[4 + N1], b2_ = q[4 + N2], b3 = q[4 + N3];
    T b2 = bn * bn + b2_ * b2_ + b3 * b3, pt = p + T(0.5) * b2;
    T E = p / (gamma - T(1)) + T(0.5) * rho * (un * un + u2 * u2 + u3 * u3) + T(0.5) * b2;
    T vb = un * bn + u2 * b2_ + u3 * b3;
    f[0] = rho * un;
    f[N1] = rho * un * un + pt - bn * bn;
    f[N2] = rho * un * u2 - bn * b2_;
    f[N3] = rho * un * u3 - bn * b3;
    f[4] = (E + pt) * un - bn * vb;
    f[4 + N1] = T(0);
    f[4 + N2] = un * b2_ - u2 * bn;
    f[4 + N3] = un * b3 - u3 * bn;
    f[8] = T(0);
}

template <typename T>
inline T fast_speed(T rho, T p, T b2, T bn, T gamma)
{
    T c2 = gamma * p / rho, bb = b2 / rho, d2 = T(0.5) * (bb + c2);
    return sq(d2 + sq(mx(d2 * d2 - c2 * bn * bn / rho, T(0))));
}

// MonCen half slope (glmmhd_turb.jl:145-155)
template <typename T>
inline T half_moncen(T l, T m, T r)
{
    T dl = m - l, dr = r - m, dc = T(0.5) * (dl + dr);
    T sgn = sel(dc >= T(0), T(1), T(-1));
    T val = sgn * mn(T(2) * mn(ab(dl), ab(dr)), ab(dc));
    return T(0.5) * sel(dl * dr <= T(0), T(0), val);
}

// PPM edges (glmmhd_turb.jl:917-923)
template <typename T>
inline void ppm_edges(T qm, T q0, T qp, T& lo, T& hi)
{
    T slope = T(0.25) * (qp - qm), curve = (qm - T(2) * q0 + qp) / T(12);
    T ql = q0 - slope + curve, qr = q0 + slope + curve;
    T dq = qr - ql, diff = (q0 - T(0.5) * (ql + qr)) * dq, d6 = dq * dq / T(6);
    auto flat = (qr - q0) * (q0 - ql) <= T(0);
    auto lor = diff > d6, hir = diff < -d6;
    lo = sel(flat, q0, sel(lor, T(3) * q0 - T(2) * qr, ql));
    hi = sel(flat, q0, sel(lor, qr, sel(hir, T(3) * q0 - T(2) * ql, qr)));
}

template <typename T>
inline auto spj(T a, T b, T c)
{
    T hi = mx(a, mx(b, c)), lo = mx(mn(a, mn(b, c)), T(1e-20));
    return hi / lo > T(2);
}

// LLF (glmmhd_turb.jl:182-190) on a pair with the GLM pair applied
template <typename T, int D>
inline void llf(const T* L, const T* R, const Sc<T>& s, T bns, T psis, T* f)
{
    T fL[NV], fR[NV], uL[NV], uR[NV];
    flux<T, D>(L, fL, s.gamma); flux<T, D>(R, fR, s.gamma);
    p2c(L, uL, s.gamma); p2c(R, uR, s.gamma);
    T b2L = L[5] * L[5] + L[6] * L[6] + L[7] * L[7], b2R = R[5] * R[5] + R[6] * R[6] + R[7] * R[7];
    T smax = mx(mx(ab(L[1 + D]) + fast_speed(L[0], L[4], b2L, bns, s.gamma),
                   ab(R[1 + D]) + fast_speed(R[0], R[4], b2R, bns, s.gamma)), s.ch);
    for (int v = 0; v < NV; ++v) f[v] = T(0.5) * (fL[v] + fR[v]) - T(0.5) * smax * (uR[v] - uL[v]);
    f[5 + D] = psis;
    f[8] = s.ch * s.ch * bns;
}

// HLLD (Miyoshi & Kusano 2005; glmmhd_turb.jl:193-272) in the rotated frame: q[1] normal velocity,
// q[5] normal field (= bns), q[6], q[7] the transverse fields.  The degenerate star and the four
// fans are lane blends.
template <typename T>
inline void star(T d, T u, T v, T w, T by, T bz, T S, T SM, T bn, T& vy, T& vz, T& byS, T& bzS)
{
    T denK = d * (S - u) * (S - SM) - bn * bn;
    auto deg = ab(denK) < T(1e-12);
    T iden = T(1) / sel(deg, T(1), denK);
    T fac = bn * (SM - u) * iden, g = (d * (S - u) * (S - u) - bn * bn) * iden;
    vy = sel(deg, v, v - by * fac);
    vz = sel(deg, w, w - bz * fac);
    byS = sel(deg, by, by * g);
    bzS = sel(deg, bz, bz * g);
}

template <typename T>
inline void hlld_x(const T* L, const T* R, T gamma, T bn, T* f)
{
    T dL = L[0], uL = L[1], vL = L[2], wL = L[3], pL = L[4], byL = L[6], bzL = L[7];
    T dR = R[0], uR = R[1], vR = R[2], wR = R[3], pR = R[4], byR = R[6], bzR = R[7];
    T b2L = bn * bn + byL * byL + bzL * bzL, b2R = bn * bn + byR * byR + bzR * bzR;
    T ptL = pL + T(0.5) * b2L, ptR = pR + T(0.5) * b2R;
    T EL = pL / (gamma - T(1)) + T(0.5) * dL * (uL * uL + vL * vL + wL * wL) + T(0.5) * b2L;
    T ER = pR / (gamma - T(1)) + T(0.5) * dR * (uR * uR + vR * vR + wR * wR) + T(0.5) * b2R;
    T cfL = fast_speed(dL, pL, L[5] * L[5] + byL * byL + bzL * bzL, bn, gamma);
    T cfR = fast_speed(dR, pR, R[5] * R[5] + byR * byR + bzR * bzR, bn, gamma);
    T SL = mn(mn(uL, uR) - mx(cfL, cfR), T(0));
    T SR = mx(mx(uL, uR) + mx(cfL, cfR), T(0));
    T qLc[NV] = {dL, uL, vL, wL, pL, bn, byL, bzL, T(0)}, qRc[NV] = {dR, uR, vR, wR, pR, bn, byR, bzR, T(0)};
    T FL[NV], FR[NV];
    flux<T, 0>(qLc, FL, gamma); flux<T, 0>(qRc, FR, gamma);
    T UL[NV] = {dL, dL * uL, dL * vL, dL * wL, EL, bn, byL, bzL, T(0)};
    T UR[NV] = {dR, dR * uR, dR * vR, dR * wR, ER, bn, byR, bzR, T(0)};
    T den = (SR - uR) * dR - (SL - uL) * dL;
    T SM = ((SR - uR) * dR * uR - (SL - uL) * dL * uL - ptR + ptL) / den;
    T pts = ((SR - uR) * dR * ptL - (SL - uL) * dL * ptR + dL * dR * (SR - uR) * (SL - uL) * (uR - uL)) / den;
    T dLs = mx(dL * (SL - uL) / (SL - SM), T(1e-12)), dRs = mx(dR * (SR - uR) / (SR - SM), T(1e-12));
    T sqL = sq(dLs), sqR = sq(dRs);
    T SLs = SM - ab(bn) / sqL, SRs = SM + ab(bn) / sqR;
    T vyL, vzL, byLs, bzLs, vyR, vzR, byRs, bzRs;
    star(dL, uL, vL, wL, byL, bzL, SL, SM, bn, vyL, vzL, byLs, bzLs);
    star(dR, uR, vR, wR, byR, bzR, SR, SM, bn, vyR, vzR, byRs, bzRs);
    T vbL = uL * bn + vL * byL + wL * bzL, vbLs = SM * bn + vyL * byLs + vzL * bzLs;
    T ELs = ((SL - uL) * EL - ptL * uL + pts * SM + bn * (vbL - vbLs)) / (SL - SM);
    T vbR = uR * bn + vR * byR + wR * bzR, vbRs = SM * bn + vyR * byRs + vzR * bzRs;
    T ERs = ((SR - uR) * ER - ptR * uR + pts * SM + bn * (vbR - vbRs)) / (SR - SM);
    T sgn = sel(bn >= T(0), T(1), T(-1)), inv = T(1) / (sqL + sqR);
    T vyss = (sqL * vyL + sqR * vyR + (byRs - byLs) * sgn) * inv;
    T vzss = (sqL * vzL + sqR * vzR + (bzRs - bzLs) * sgn) * inv;
    T byss = (sqL * byRs + sqR * byLs + sqL * sqR * (vyR - vyL) * sgn) * inv;
    T bzss = (sqL * bzRs + sqR * bzLs + sqL * sqR * (vzR - vzL) * sgn) * inv;
    T vbss = SM * bn + vyss * byss + vzss * bzss;
    T ELss = ELs - sqL * (vbLs - vbss) * sgn, ERss = ERs + sqR * (vbRs - vbss) * sgn;
    T ULs[NV] = {dLs, dLs * SM, dLs * vyL, dLs * vzL, ELs, bn, byLs, bzLs, T(0)};
    T URs[NV] = {dRs, dRs * SM, dRs * vyR, dRs * vzR, ERs, bn, byRs, bzRs, T(0)};
    T ULss[NV] = {dLs, dLs * SM, dLs * vyss, dLs * vzss, ELss, bn, byss, bzss, T(0)};
    T URss[NV] = {dRs, dRs * SM, dRs * vyss, dRs * vzss, ERss, bn, byss, bzss, T(0)};
    auto f0 = SLs >= T(0), f1 = SM >= T(0), f2 = SRs >= T(0);
    for (int v = 0; v < NV; ++v) {
        T a0 = FL[v] + SL * (ULs[v] - UL[v]);
        T a1 = FL[v] + SLs * ULss[v] - (SLs - SL) * ULs[v] - SL * UL[v];
        T a2 = FR[v] + SRs * URss[v] - (SRs - SR) * URs[v] - SR * UR[v];
        T a3 = FR[v] + SR * (URs[v] - UR[v]);
        f[v] = sel(f0, a0, sel(f1, a1, sel(f2, a2, a3)));
    }
}

// Riemann solve at an axis-D face: interface floors, GLM pair, solver (glmmhd_turb.jl:275-292).
template <typename T, int D, bool HLLD>
inline void riemann(const T* Lq, const T* Rq, const Sc<T>& s, T* out)
{
    T L[NV], R[NV];
    for (int v = 0; v < NV; ++v) { L[v] = Lq[v]; R[v] = Rq[v]; }
    L[0] = mx(L[0], s.smallr); L[4] = mx(L[4], s.pfl);
    R[0] = mx(R[0], s.smallr); R[4] = mx(R[4], s.pfl);
    T bns = T(0.5) * (L[5 + D] + R[5 + D]) - T(0.5) * (R[8] - L[8]) / s.ch;
    T psis = T(0.5) * (L[8] + R[8]) - T(0.5) * s.ch * (R[5 + D] - L[5 + D]);
    L[5 + D] = bns; R[5 + D] = bns;
    if (!HLLD) {
        llf<T, D>(L, R, s, bns, psis, out);
        return;
    }
    // rotate axis D to x (cyclic), solve, rotate back
    constexpr int A1 = 1 + D, A2 = 1 + (D + 1) % 3, A3 = 1 + (D + 2) % 3;
    T Lr[NV] = {L[0], L[A1], L[A2], L[A3], L[4], L[4 + A1], L[4 + A2], L[4 + A3], T(0)};
    T Rr[NV] = {R[0], R[A1], R[A2], R[A3], R[4], R[4 + A1], R[4 + A2], R[4 + A3], T(0)};
    T f[NV];
    hlld_x(Lr, Rr, s.gamma, bns, f);
    out[0] = f[0]; out[4] = f[4];
    out[A1] = f[1]; out[A2] = f[2]; out[A3] = f[3];
    out[4 + A2] = f[6]; out[4 + A3] = f[7];
    out[4 + A1] = psis;
    out[8] = s.ch * s.ch * bns;
}

// full-transverse Hancock predictor (glmmhd_turb.jl:165-172) with the revert (:393-395)
template <typename T>
inline void hancock(const T* m0, const T* sx, const T* sy, const T* sz, const Sc<T>& s, T* mh)
{
    T a[NV], b[NV], fa[NV], fb[NV], acc[NV], u[NV];
    for (int v = 0; v < NV; ++v) { a[v] = m0[v] + sx[v]; b[v] = m0[v] - sx[v]; }
    flux<T, 0>(a, fa, s.gamma); flux<T, 0>(b, fb, s.gamma);
    for (int v = 0; v < NV; ++v) acc[v] = fa[v] - fb[v];
    for (int v = 0; v < NV; ++v) { a[v] = m0[v] + sy[v]; b[v] = m0[v] - sy[v]; }
    flux<T, 1>(a, fa, s.gamma); flux<T, 1>(b, fb, s.gamma);
    for (int v = 0; v < NV; ++v) acc[v] = acc[v] + (fa[v] - fb[v]);
    for (int v = 0; v < NV; ++v) { a[v] = m0[v] + sz[v]; b[v] = m0[v] - sz[v]; }
    flux<T, 2>(a, fa, s.gamma); flux<T, 2>(b, fb, s.gamma);
    for (int v = 0; v < NV; ++v) acc[v] = acc[v] + (fa[v] - fb[v]);
    p2c(m0, u, s.gamma);
    for (int v = 0; v < NV; ++v) u[v] = u[v] - s.half_dtdx * acc[v];
    T q[NV];
    c2p(u, q, s);
    auto bad = (q[0] <= s.smallr) | (q[4] <= s.pfl);
    for (int v = 0; v < NV; ++v) mh[v] = sel(bad, m0[v], q[v]);
}

// Per-thread plane buffers of one chunk (ChunkBufs, cpu_simd.jl:446-449).  Plane arrays are
// [component][j][NP], x fastest.  rec: the reconstruction of the current plane -- (mh, sx, sy, sz)
// for PLM as the reference stores it, or the six face states for PPM (lo/hi per axis).
template <typename S>
struct Bufs {
    std::vector<S> rec, cMz, cPz, cXY, pPz, pXY, pFz, Fz;
};

// S: the scalar type of the arrays; T = Vec<S>: the lanes the physics runs on.
template <typename S, bool PPM, bool HLLD>
struct Stream {
    typedef Vec<S> T;
    static constexpr int W = T::W;
    int N, NP;
    Sc<T> s;
    static constexpr int NREC = PPM ? 6 * NV : 4 * NV;

    size_t lidx(int x, int j, int k, int v) const { return (size_t)x + (size_t)NP * ((size_t)j + (size_t)N * ((size_t)k + (size_t)N * v)); }
    size_t pidx(int x, int j, int c) const { return (size_t)x + (size_t)NP * ((size_t)j + (size_t)N * c); }
    void ld9(const S* a, size_t base, size_t stride, T* q) const
    {
        for (int v = 0; v < NV; ++v) q[v] = T::load(a + base + stride * v);
    }

    // recon1!: reconstruct plane kk (cpu_simd.jl:462-490)
    void recon1(S* rec, const S* u, int kk) const
    {
        const int kc = ((kk % N) + N) % N, kdn = (((kk - 1) % N) + N) % N, kup = (((kk + 1) % N) + N) % N;
        const size_t vs = (size_t)NP * N * N, ps = (size_t)NP * N;
        for (int j = 0; j < N; ++j) {
            const int jm = j == 0 ? N - 1 : j - 1, jp = j == N - 1 ? 0 : j + 1;
            for (int x = 1; x <= N; x += W) {
                T c[NV], m0[NV], nb[6][NV];
                const int xs[6] = {x - 1, x + 1, x, x, x, x}, js[6] = {j, j, jm, jp, j, j}, ks[6] = {kc, kc, kc, kc, kdn, kup};
                ld9(u, lidx(x, j, kc, 0), vs, c);
                c2p(c, m0, s);
                for (int n = 0; n < 6; ++n) {
                    ld9(u, lidx(xs[n], js[n], ks[n], 0), vs, c);
                    c2p(c, nb[n], s);
                }
                T sl[3][NV];
                for (int d = 0; d < 3; ++d)
                    for (int v = 0; v < NV; ++v) sl[d][v] = half_moncen(nb[2 * d][v], m0[v], nb[2 * d + 1][v]);
                T mh[NV];
                hancock(m0, sl[0], sl[1], sl[2], s, mh);
                if (!PPM) {
                    for (int v = 0; v < NV; ++v) {
                        mh[v].store(rec + pidx(x, j, v));
                        sl[0][v].store(rec + pidx(x, j, NV + v));
                        sl[1][v].store(rec + pidx(x, j, 2 * NV + v));
                        sl[2][v].store(rec + pidx(x, j, 3 * NV + v));
                    }
                } else {
                    // PPM faces mh + edge - m0, PLM fallback at strong pressure jumps (glmmhd_turb.jl:955-973),
                    // and the iflo face revert (:618)
                    auto up = !(spj(nb[0][4], m0[4], nb[1][4]) | spj(nb[2][4], m0[4], nb[3][4]) |
                                spj(nb[4][4], m0[4], nb[5][4]));
                    for (int d = 0; d < 3; ++d) {
                        T lo[NV], hi[NV];
                        for (int v = 0; v < NV; ++v) {
                            T el, eh;
                            ppm_edges(nb[2 * d][v], m0[v], nb[2 * d + 1][v], el, eh);
                            lo[v] = sel(up, mh[v] + el - m0[v], mh[v] - sl[d][v]);
                            hi[v] = sel(up, mh[v] + eh - m0[v], mh[v] + sl[d][v]);
                        }
                        auto blo = (lo[0] < s.smallr) | (lo[4] < s.pfl), bhi = (hi[0] < s.smallr) | (hi[4] < s.pfl);
                        for (int v = 0; v < NV; ++v) {
                            sel(blo, m0[v], lo[v]).store(rec + pidx(x, j, (2 * d) * NV + v));
                            sel(bhi, m0[v], hi[v]).store(rec + pidx(x, j, (2 * d + 1) * NV + v));
                        }
                    }
                }
            }
            for (int c = 0; c < NREC; ++c) {   // periodic x ghosts of the plane buffer
                rec[pidx(0, j, c)] = rec[pidx(N, j, c)];
                rec[pidx(N + 1, j, c)] = rec[pidx(1, j, c)];
            }
        }
        (void)ps;
    }

    // the -/+ face state of axis d of the plane buffer at (x, j), W lanes
    void face(const S* rec, int x, int j, int d, bool hi, T* q) const
    {
        if (PPM) {
            for (int v = 0; v < NV; ++v) q[v] = T::load(rec + pidx(x, j, (2 * d + (hi ? 1 : 0)) * NV + v));
        } else {
            for (int v = 0; v < NV; ++v) {
                T m = T::load(rec + pidx(x, j, v)), sl = T::load(rec + pidx(x, j, (1 + d) * NV + v));
                q[v] = hi ? m + sl : m - sl;
            }
        }
    }

    // faces1!: in-plane fluxes (each face from both sides) and z face states (cpu_simd.jl:492-514)
    void faces1(S* cMz, S* cPz, S* cXY, const S* rec) const
    {
        for (int j = 0; j < N; ++j) {
            const int jp = j == N - 1 ? 0 : j + 1, jm = j == 0 ? N - 1 : j - 1;
            for (int x = 1; x <= N; x += W) {
                T a[NV], b[NV], Fxh[NV], Fxl[NV], Fyh[NV], Fyl[NV];
                face(rec, x, j, 0, true, a); face(rec, x + 1, j, 0, false, b); riemann<T, 0, HLLD>(a, b, s, Fxh);
                face(rec, x - 1, j, 0, true, a); face(rec, x, j, 0, false, b); riemann<T, 0, HLLD>(a, b, s, Fxl);
                face(rec, x, j, 1, true, a); face(rec, x, jp, 1, false, b); riemann<T, 1, HLLD>(a, b, s, Fyh);
                face(rec, x, jm, 1, true, a); face(rec, x, j, 1, false, b); riemann<T, 1, HLLD>(a, b, s, Fyl);
                face(rec, x, j, 2, false, a); face(rec, x, j, 2, true, b);
                for (int v = 0; v < NV; ++v) {
                    ((Fxl[v] - Fxh[v]) + (Fyl[v] - Fyh[v])).store(cXY + pidx(x, j, v));
                    a[v].store(cMz + pidx(x, j, v));
                    b[v].store(cPz + pidx(x, j, v));
                }
            }
        }
    }

    // fzupd1!: z face kupd+1/2 and the update of plane kupd (cpu_simd.jl:516-540)
    void fzupd1(S* unew, const S* uold, int kupd, bool doupd, const S* pPz, const S* cMz, const S* pXY,
                const S* pFz, S* Fz) const
    {
        const int kz = ((kupd % N) + N) % N;
        const size_t vs = (size_t)NP * N * N;
        for (int j = 0; j < N; ++j) {
            for (int x = 1; x <= N; x += W) {
                T pz[NV], mz[NV], fz[NV];
                for (int v = 0; v < NV; ++v) { pz[v] = T::load(pPz + pidx(x, j, v)); mz[v] = T::load(cMz + pidx(x, j, v)); }
                riemann<T, 2, HLLD>(pz, mz, s, fz);
                for (int v = 0; v < NV; ++v) fz[v].store(Fz + pidx(x, j, v));
                if (doupd) {
                    T r[NV];
                    const size_t b0 = lidx(x, j, kz, 0);
                    for (int v = 0; v < NV; ++v)
                        r[v] = T::load(uold + b0 + vs * v) +
                               s.dtdx * (T::load(pXY + pidx(x, j, v)) + (T::load(pFz + pidx(x, j, v)) - fz[v]));
                    const T emag = T(0.5) * (r[5] * r[5] + r[6] * r[6] + r[7] * r[7]);
                    const auto bad = r[0] < s.smallr;
                    const T fl5 = s.pfl / (s.gamma - T(1)) + emag;
                    sel(bad, s.smallr, r[0]).store(unew + b0);
                    sel(bad, T(0), r[1]).store(unew + b0 + vs);
                    sel(bad, T(0), r[2]).store(unew + b0 + 2 * vs);
                    sel(bad, T(0), r[3]).store(unew + b0 + 3 * vs);
                    sel(bad, fl5, r[4]).store(unew + b0 + 4 * vs);
                    for (int v = 5; v < 8; ++v) r[v].store(unew + b0 + vs * v);
                    (r[8] * s.glm_fac).store(unew + b0 + 8 * vs);
                }
            }
        }
    }

    // cpu_stream_step! (cpu_simd.jl:551-569)
    void step(S* unew, S* uold, std::vector<Bufs<S>>& bufs) const
    {
        const int nch = (int)bufs.size();
#pragma omp parallel for schedule(static)
        for (int k = 0; k < N; ++k)
            for (int v = 0; v < NV; ++v)
                for (int j = 0; j < N; ++j) {
                    uold[lidx(0, j, k, v)] = uold[lidx(N, j, k, v)];
                    uold[lidx(N + 1, j, k, v)] = uold[lidx(1, j, k, v)];
                }
#pragma omp parallel for schedule(static, 1)
        for (int c = 0; c < nch; ++c) {
            Bufs<S>& b = bufs[c];
            const int c0 = (int)((long long)c * N / nch), c1 = (int)((long long)(c + 1) * N / nch) - 1;
            S *cPz = b.cPz.data(), *pPz = b.pPz.data(), *cXY = b.cXY.data(), *pXY = b.pXY.data();
            S *Fz = b.Fz.data(), *pFz = b.pFz.data();
            for (int kk = c0 - 1; kk <= c1 + 1; ++kk) {
                recon1(b.rec.data(), uold, kk);
                faces1(b.cMz.data(), cPz, cXY, b.rec.data());
                fzupd1(unew, uold, kk - 1, kk >= c0 + 1, pPz, b.cMz.data(), pXY, pFz, Fz);
                std::swap(cPz, pPz); std::swap(cXY, pXY); std::swap(Fz, pFz);
            }
        }
    }
};

template <typename S>
std::vector<Bufs<S>> alloc_bufs(int nch, int N, int NP, int nrec)
{
    std::vector<Bufs<S>> bufs(nch);
    const size_t pl = (size_t)NP * N;
#pragma omp parallel for schedule(static, 1)
    for (int c = 0; c < nch; ++c) {   // first touch on the thread that streams the chunk
        Bufs<S>& b = bufs[c];
        b.rec.assign(pl * nrec, S(0));
        for (auto* v : {&b.cMz, &b.cPz, &b.cXY, &b.pPz, &b.pXY, &b.pFz, &b.Fz}) v->assign(pl * NV, S(0));
    }
    return bufs;
}

template <typename S, bool PPM, bool HLLD>
double run_impl(S* u, int N, int nsteps, int nch, double gamma, double smallr, double pfl, double dt, double dx,
                double ch, double glm_fac)
{
    typedef Vec<S> T;
    Stream<S, PPM, HLLD> st;
    st.N = N; st.NP = N + 2;
    st.s = Sc<T>{T((S)gamma), T((S)smallr), T((S)pfl), T((S)(dt / dx)), T((S)(0.5 * dt / dx)), T((S)ch), T((S)glm_fac)};
    const size_t n = (size_t)N * N * N, np = (size_t)st.NP * N * N;
    // padded state buffers, first-touched by chunk (cpu_simd.jl:579-590); + W for the last lanes' loads
    S* u1 = new S[np * NV + T::W];
    S* u2 = new S[np * NV + T::W];
#pragma omp parallel for schedule(static, 1)
    for (int c = 0; c < nch; ++c) {
        const int c0 = (int)((long long)c * N / nch), c1 = (int)((long long)(c + 1) * N / nch) - 1;
        for (int k = c0; k <= c1; ++k)
            for (int v = 0; v < NV; ++v)
                for (int j = 0; j < N; ++j) {
                    std::memset(&u2[st.lidx(0, j, k, v)], 0, sizeof(S) * st.NP);
                    u1[st.lidx(0, j, k, v)] = 0;
                    u1[st.lidx(N + 1, j, k, v)] = 0;
                    std::memcpy(&u1[st.lidx(1, j, k, v)], &u[(size_t)v * n + (size_t)N * (j + (size_t)N * k)], sizeof(S) * N);
                }
    }
    auto bufs = alloc_bufs<S>(nch, N, st.NP, Stream<S, PPM, HLLD>::NREC);
    const double t0 = omp_get_wtime();
    for (int it = 0; it < nsteps; ++it) {
        st.step(u2, u1, bufs);
        std::swap(u1, u2);
    }
    const double el = omp_get_wtime() - t0;
    for (int k = 0; k < N; ++k)
        for (int v = 0; v < NV; ++v)
            for (int j = 0; j < N; ++j)
                std::memcpy(&u[(size_t)v * n + (size_t)N * (j + (size_t)N * k)], &u1[st.lidx(1, j, k, v)], sizeof(S) * N);
    delete[] u1;
    delete[] u2;
    return el;
}

}  // namespace

extern "C" {

// nsteps steps of the chunked z-stream on an N^3 periodic state u (9 SoA planes, x fastest, in
// place).  scheme 0: cpu_simd.jl's (full Hancock, PLM MonCen, LLF); 1: PPM + HLLD with the iflo
// face revert.  prec 8 (double, 4 lanes) or 4 (float, 8 lanes); N a multiple of the lane count.
// nchunks <= 0: min(threads, N/2) as the reference.  Returns the wall time of the steps in seconds
// (buffer set-up excluded), or -1 on bad arguments.
double cs_run(void* u, int N, int nsteps, int prec, int scheme, int nchunks, double gamma, double smallr, double pfl,
              double dt, double dx, double ch, double glm_fac)
{
    if (!u || N < 2 || nsteps < 0 || (prec != 8 && prec != 4) || (scheme != 0 && scheme != 1)) return -1.0;
    if (N % (32 / prec) != 0) return -1.0;
    if (nchunks <= 0) nchunks = std::min(omp_get_max_threads(), std::max(1, N / 2));
    if (prec == 8) {
        return scheme == 0 ? run_impl<double, false, false>((double*)u, N, nsteps, nchunks, gamma, smallr, pfl, dt, dx, ch, glm_fac)
                           : run_impl<double, true, true>((double*)u, N, nsteps, nchunks, gamma, smallr, pfl, dt, dx, ch, glm_fac);
    }
    return scheme == 0 ? run_impl<float, false, false>((float*)u, N, nsteps, nchunks, gamma, smallr, pfl, dt, dx, ch, glm_fac)
                       : run_impl<float, true, true>((float*)u, N, nsteps, nchunks, gamma, smallr, pfl, dt, dx, ch, glm_fac);
}

int cs_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
