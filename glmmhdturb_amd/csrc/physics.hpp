// physics.hpp -- GLM-MHD point physics for the gfx950 kernels.
//
// Device functions templated on the scalar type T (double or float).  A direction D in {0,1,2}
// is a compile-time cyclic permutation (A,B,C) = (D, D+1, D+2) mod 3 of the vector components,
// so the reference's rot_to / rot_flux_from (glmmhd_turb.jl:119-138) become register renaming.
//
// Conserved order (rho, mx, my, mz, E, Bx, By, Bz, psi); primitive order
// (rho, vx, vy, vz, p, Bx, By, Bz, psi) -- glmmhd_turb.jl:61-64.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define GLM_DEV __device__ __forceinline__
// Point physics is also compiled for the host (CPU checks of the adjoint against dual numbers).
#define GLM_HD __host__ __device__ __forceinline__

namespace glm {

enum Pred { PRED_FULL = 0, PRED_ONED = 1 };
enum Recon { RECON_PLM = 0, RECON_PPM = 1, RECON_PPM_SPIKE = 2 };
enum Riem { RIEM_LLF = 0, RIEM_LLFB = 1, RIEM_HLL = 2, RIEM_HLLD = 3 };
enum Slope { SLOPE_MONCEN = 0, SLOPE_MINMOD = 1 };
enum FaceGuard { GUARD_NONE = 0, GUARD_IFLO = 1, GUARD_CLAMP = 2 };

// Scalars every point function needs (precomputed on the host in double, cast to T).
template <typename T>
struct Phys {
    T gamma, igm1;          // gamma, 1/(gamma-1)
    T smallr, pfl;          // density and pressure floors (pfl = smallr*smallc^2/gamma)
    T ch, ch2, half_ich;    // GLM speed, ch^2, 0.5/ch
    T llf_dmin, llf_pmin;   // LLF robustness switch (<= 0 disables), glmmhd_turb.jl:287-290
};

// Nine plane pointers passed by value in kernel arguments.
template <typename T>
struct Ptr9 {
    T* p[9];
    GLM_DEV T* operator[](int i) const { return p[i]; }
};

// fmax/fmin/fabs semantics (the C restatement's): one v_max/v_min instruction, |x| as a source
// modifier, copysign as one bit-field insert.
GLM_HD double tmax(double a, double b) { return __builtin_fmax(a, b); }
GLM_HD float tmax(float a, float b) { return __builtin_fmaxf(a, b); }
GLM_HD double tmin(double a, double b) { return __builtin_fmin(a, b); }
GLM_HD float tmin(float a, float b) { return __builtin_fminf(a, b); }
GLM_HD double tabs(double a) { return __builtin_fabs(a); }
GLM_HD float tabs(float a) { return __builtin_fabsf(a); }
GLM_HD double tcopysign(double a, double b) { return __builtin_copysign(a, b); }
GLM_HD float tcopysign(float a, float b) { return __builtin_copysignf(a, b); }

// Reciprocal from the hardware estimate refined by Newton steps (two for double, one for float):
// within about 1 ulp of 1/x for the normal-range operands of this solver, at half the cost of
// an IEEE divide sequence.  Used where the quotient feeds further arithmetic.  (The host build,
// used only by CPU checks, divides.)
GLM_HD double trcp(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
#else
    return 1.0 / x;
#endif
}
GLM_HD float trcp(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    float r = __builtin_amdgcn_rcpf(x);
    float e = __builtin_fmaf(-x, r, 1.0f);
    return __builtin_fmaf(r, e, r);
#else
    return 1.0f / x;
#endif
}
// Square root for the non-negative operands of this solver (speeds squared, densities, floored
// energies).  Device: the hardware reciprocal square root estimate (~25 bits), one Goldschmidt
// step for sqrt and 1/(2 sqrt), and one Newton correction of the root: within half an ulp of the
// correctly rounded root on 2^20 operands spanning 2^-30..2^31 (measured against long double),
// without the builtin's range scaling, which these operands never need.
// The float version is the hardware square root (1 ulp).  (The host build, used only by CPU
// checks, calls the library root.)
GLM_HD double tsqrt_core(double a, double y)
{
    double g = a * y, h = 0.5 * y;
    const double r = __builtin_fma(-g, h, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    const double d = __builtin_fma(-g, g, a);
    return __builtin_fma(d, h, g);
}
GLM_HD double tsqrt(double a)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const double s = tsqrt_core(a, __builtin_amdgcn_rsq(a));
    return a == 0.0 ? a : s;   // the estimate of 1/sqrt(0) is infinite; negatives give NaN as sqrt does
#else
    return __builtin_sqrt(a);
#endif
}
// The root of a positive operand (no zero test): the outer root of a fast speed, whose operand
// is at least half the squared sound speed.
GLM_HD double tsqrt_pos(double a)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return tsqrt_core(a, __builtin_amdgcn_rsq(a));
#else
    return __builtin_sqrt(a);
#endif
}
// The root of a non-negative operand that is 0 or far above 1e-300 (the inner root of a fast
// speed, a sum of squares): the estimate is taken at max(a, 1e-300), so a = 0 gives g = 0 and the
// iteration returns exactly 0 with no select; an operand below 1e-300 would get a root with a
// poor estimate, which next to the outer sum (>= the squared sound speed) is invisible.
GLM_HD double tsqrt_nn(double a)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return tsqrt_core(a, __builtin_amdgcn_rsq(__builtin_fmax(a, 1e-300)));
#else
    return __builtin_sqrt(a);
#endif
}
// Square root and reciprocal square root together (positive operands): the root as tsqrt, and
// 1/sqrt(a) from the same Goldschmidt step (2h, about an ulp from the correctly rounded value).
GLM_HD double tsqrt_rs(double a, double& rs)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const double y = __builtin_amdgcn_rsq(a);
    double g = a * y, h = 0.5 * y;
    const double r = __builtin_fma(-g, h, 0.5);
    g = __builtin_fma(g, r, g);
    h = __builtin_fma(h, r, h);
    rs = h + h;
    const double d = __builtin_fma(-g, g, a);
    return __builtin_fma(d, h, g);
#else
    const double q = __builtin_sqrt(a);
    rs = 1.0 / q;
    return q;
#endif
}
GLM_HD float tsqrt_rs(float a, float& rs)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const float q = __builtin_amdgcn_sqrtf(a);
    rs = __builtin_amdgcn_rsqf(a);
    return q;
#else
    const float q = __builtin_sqrtf(a);
    rs = 1.0f / q;
    return q;
#endif
}
// Any other scalar type (the dual numbers of the CPU checks): the root and its reciprocal.
template <typename T>
GLM_HD T tsqrt_rs(T a, T& rs)
{
    T q = tsqrt(a);
    rs = trcp(q);
    return q;
}
GLM_HD float tsqrt(float a)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_sqrtf(a);
#else
    return __builtin_sqrtf(a);
#endif
}
GLM_HD float tsqrt_pos(float a) { return tsqrt(a); }
GLM_HD float tsqrt_nn(float a) { return tsqrt(a); }
// Any other scalar type (the dual numbers of the CPU checks): the plain root.
template <typename T>
GLM_HD T tsqrt_pos(T a) { return tsqrt(a); }
template <typename T>
GLM_HD T tsqrt_nn(T a) { return tsqrt(a); }

// Gas pressure (gamma-1)(E - |m|^2/(2 rho) - |B|^2/2), floored (glmmhd_turb.jl:75).  With FP32 state
// the subtraction is carried out in FP64: in low-beta or high-Mach cells E is dominated by the
// magnetic or kinetic part and the FP32 difference would lose the thermal part.
GLM_HD double gas_pressure(const double c[9], double ir, double gm1, double pfl)
{
    double ekin = 0.5 * (c[1] * c[1] + c[2] * c[2] + c[3] * c[3]) * ir;
    double emag = 0.5 * (c[5] * c[5] + c[6] * c[6] + c[7] * c[7]);
    return __builtin_fmax(gm1 * (c[4] - ekin - emag), pfl);
}
GLM_HD float gas_pressure(const float c[9], float ir, float gm1, float pfl)
{
    double m1 = c[1], m2 = c[2], m3 = c[3], b1 = c[5], b2 = c[6], b3 = c[7];
    double ekin = 0.5 * (m1 * m1 + m2 * m2 + m3 * m3) * (double)ir;
    double emag = 0.5 * (b1 * b1 + b2 * b2 + b3 * b3);
    return __builtin_fmaxf((float)((double)gm1 * ((double)c[4] - ekin - emag)), pfl);
}

// glmmhd_turb.jl:68-77 cons2prim
template <typename T>
GLM_HD void cons2prim(const T c[9], T q[9], const Phys<T>& s)
{
    T rho = tmax(c[0], s.smallr);
    T ir = trcp(rho);
    q[0] = rho;
    q[1] = c[1] * ir; q[2] = c[2] * ir; q[3] = c[3] * ir;
    q[4] = gas_pressure(c, ir, s.gamma - T(1), s.pfl);
    q[5] = c[5]; q[6] = c[6]; q[7] = c[7]; q[8] = c[8];
}

// glmmhd_turb.jl:79-86 prim2cons
template <typename T>
GLM_HD void prim2cons(const T q[9], T c[9], const Phys<T>& s)
{
    T ekin = T(0.5) * q[0] * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    T emag = T(0.5) * (q[5] * q[5] + q[6] * q[6] + q[7] * q[7]);
    c[0] = q[0];
    c[1] = q[0] * q[1]; c[2] = q[0] * q[2]; c[3] = q[0] * q[3];
    c[4] = q[4] * s.igm1 + ekin + emag;
    c[5] = q[5]; c[6] = q[6]; c[7] = q[7]; c[8] = q[8];
}

// glmmhd_turb.jl:89-95 fast_speed; b2 = |B|^2 is frame independent.
template <typename T>
GLM_HD T fast_speed(T rho, T p, T b2, T bn, const Phys<T>& s)
{
    T ir = trcp(rho);
    T c2 = s.gamma * p * ir;
    T bb = b2 * ir;
    T d2 = T(0.5) * (bb + c2);
    return tsqrt_pos(d2 + tsqrt_nn(tmax(d2 * d2 - c2 * bn * bn * ir, T(0))));
}

// CFL signal speed sum_d (|v_d| + c_f(B_d)) of a primitive state (glmmhd_turb.jl:1963-1971),
// sharing 1/rho, a^2 and the half sum between the three directions.  Evaluated in the state
// precision (the fast speed has no cancellation that FP32 would lose: the inner root is of a sum
// of squares, the outer of a positive sum; the FP32 pressure it uses was recovered in FP64) and
// returned as FP64: the max over cells and the step size are FP64.
template <typename T>
GLM_DEV double cfl_speed(const T q[9], const Phys<T>& s)
{
    T ir = trcp(q[0]);
    T c2 = s.gamma * q[4] * ir;
    T bb = (q[5] * q[5] + q[6] * q[6] + q[7] * q[7]) * ir;
    T d2 = T(0.5) * (bb + c2);
    T dd = d2 * d2, k = c2 * ir;
    T fx = tsqrt_pos(d2 + tsqrt_nn(tmax(dd - k * q[5] * q[5], T(0))));
    T fy = tsqrt_pos(d2 + tsqrt_nn(tmax(dd - k * q[6] * q[6], T(0))));
    T fz = tsqrt_pos(d2 + tsqrt_nn(tmax(dd - k * q[7] * q[7], T(0))));
    return (double)((tabs(q[1]) + fx) + (tabs(q[2]) + fy) + (tabs(q[3]) + fz));
}

// Largest fast magnetosonic speed of a cell over the three grid directions, max_d c_f(B_d): the
// Dedner cleaning speed "ch = max fast speed" is the max of this over the grid.
template <typename T>
GLM_DEV double fast_max_speed(const T q[9], const Phys<T>& s)
{
    T ir = trcp(q[0]);
    T c2 = s.gamma * q[4] * ir;
    T bb = (q[5] * q[5] + q[6] * q[6] + q[7] * q[7]) * ir;
    T d2 = T(0.5) * (bb + c2);
    T dd = d2 * d2, k = c2 * ir;
    T fx = tsqrt_pos(d2 + tsqrt_nn(tmax(dd - k * q[5] * q[5], T(0))));
    T fy = tsqrt_pos(d2 + tsqrt_nn(tmax(dd - k * q[6] * q[6], T(0))));
    T fz = tsqrt_pos(d2 + tsqrt_nn(tmax(dd - k * q[7] * q[7], T(0))));
    return (double)tmax(fx, tmax(fy, fz));
}

// Ideal-MHD flux of a primitive state in direction D (glmmhd_turb.jl:98-115 evaluated in the
// rotated frame and rotated back, :161-162).  F[Bn] = F[psi] = 0.
template <typename T, int D>
GLM_HD void flux_dir(const T q[9], T f[9], const Phys<T>& s)
{
    constexpr int A = D % 3, B = (D + 1) % 3, C = (D + 2) % 3;
    const T rho = q[0], un = q[1 + A], u1 = q[1 + B], u2 = q[1 + C], p = q[4];
    const T bn = q[5 + A], b1 = q[5 + B], b2_ = q[5 + C];
    T b2 = bn * bn + b1 * b1 + b2_ * b2_;
    T ptot = p + T(0.5) * b2;
    T ekin = T(0.5) * rho * (un * un + u1 * u1 + u2 * u2);
    T E = p * s.igm1 + ekin + T(0.5) * b2;
    T vdotb = un * bn + u1 * b1 + u2 * b2_;
    f[0] = rho * un;
    f[1 + A] = rho * un * un + ptot - bn * bn;
    f[1 + B] = rho * un * u1 - bn * b1;
    f[1 + C] = rho * un * u2 - bn * b2_;
    f[4] = (E + ptot) * un - bn * vdotb;
    f[5 + A] = T(0);
    f[5 + B] = un * b1 - u1 * bn;
    f[5 + C] = un * b2_ - u2 * bn;
    f[8] = T(0);
}

// x-frame flux of a state already in the rotated (x-normal) order (phys_flux_x).
template <typename T>
GLM_HD void flux_x(const T q[9], T f[9], const Phys<T>& s) { flux_dir<T, 0>(q, f, s); }

// glmmhd_turb.jl:145-150 moncen and :300-301 minmod (each halved by the caller's prim_slope).
template <typename T>
GLM_HD T moncen(T dl, T dr)
{
    // sgn(dc) * m; at dc = -0 this gives -0 where the reference's sgn gives +0 (m = 0 there),
    // which no later sum can tell apart
    T dc = T(0.5) * (dl + dr);
    T m = tmin(T(2) * tmin(tabs(dl), tabs(dr)), tabs(dc));
    return (dl * dr <= T(0)) ? T(0) : tcopysign(m, dc);
}
template <typename T>
GLM_HD T minmod(T a, T b)
{
    return (a * b <= T(0)) ? T(0) : (tabs(a) < tabs(b) ? a : b);
}
// Half the MonCen slope without selects: with s the sign of dl + dr, a = s dl and b = s dr,
// 0.5 moncen(dl, dr) = s max(0, min(a, b, |dl + dr| / 4)).  When dl and dr share a sign, a and b are
// |dl| and |dr| and the min is 0.5 min(2|dl|, 2|dr|, |dc|) exactly (powers of two); otherwise one of
// a, b is <= 0 and the max gives 0 (possibly -0, as moncen's own sign convention).
template <typename T>
GLM_HD T moncen_half(T dl, T dr) { return T(0.5) * moncen(dl, dr); }
GLM_HD double moncen_half(double dl, double dr);
GLM_HD float moncen_half(float dl, float dr);
// Half slope (glmmhd_turb.jl:153-155 MonCen / :302-303 minmod).
template <typename T, int SLOPE>
GLM_HD T half_slope(T l, T m, T r)
{
    T dl = m - l, dr = r - m;
    if (SLOPE == SLOPE_MINMOD) return T(0.5) * minmod(dl, dr);
    return moncen_half(dl, dr);
}

// glmmhd_turb.jl:917-923 ppm_mono / ppm_edges (Julia form).
template <typename T>
GLM_HD void ppm_edges(T qm, T q0, T qp, T& lo, T& hi)
{
    T slope = T(0.25) * (qp - qm);
    T curve = (qm - T(2) * q0 + qp) * T(1.0 / 12.0);
    T ql = q0 - slope + curve, qr = q0 + slope + curve;
    T dq = qr - ql;
    T diff = (q0 - T(0.5) * (ql + qr)) * dq;
    T lim = dq * dq * T(1.0 / 6.0);
    bool flat = (qr - q0) * (q0 - ql) <= T(0);
    bool hi_os = diff > lim;
    bool lo_os = diff < -lim;
    lo = flat ? q0 : (hi_os ? T(3) * q0 - T(2) * qr : ql);
    // the reference's hi = flat ? q0 : (hi_os ? qr : (lo_os ? 3 q0 - 2 ql : qr)); hi_os and lo_os
    // exclude each other (lim >= 0, and NaN fails both), so the middle select is redundant
    hi = flat ? q0 : (lo_os ? T(3) * q0 - T(2) * ql : qr);
}

// The CW84-limited parabola of ppm_edges as edge offsets dl = q0 - lo, dr = hi - q0, in min/max
// form.  With a = slope - curve and b = slope + curve the unlimited offsets, ppm_edges' tests are
// (flat) a*b <= 0 -> both 0; (diff > lim) <=> |a| > 2|b| -> a = 2b; (diff < -lim) <=> |b| > 2|a|
// -> b = 2a (diff = (a^2 - b^2)/2 and lim = (a + b)^2/6, so diff > lim <=> (a + b)(a - 2b) > 0).
// With s the sign of a + b (that is, of the slope) a' = s max(0, min(s a, 2 s b)) and b' = s max(0, min(s b, 2 s a)):
// the same limited parabola up to rounding, with no selects (the limiter is continuous at each
// switch, so a switch decided differently by rounding moves the result by rounding only).
GLM_HD double flip_sign(double x, double ref)
{
    const unsigned long long m = __builtin_bit_cast(unsigned long long, ref) & 0x8000000000000000ull;
    return __builtin_bit_cast(double, __builtin_bit_cast(unsigned long long, x) ^ m);
}
GLM_HD float flip_sign(float x, float ref)
{
    const unsigned m = __builtin_bit_cast(unsigned, ref) & 0x80000000u;
    return __builtin_bit_cast(float, __builtin_bit_cast(unsigned, x) ^ m);
}
// The sign is applied by multiplying with s = +-1 rather than by flipping the sign bit: a product is
// a canonical float, so the min/max that follow need no NaN-quieting of their operands (a bit
// operation's result would get one).  Multiplying by +-1 is exact, so the values are the same.
GLM_HD double moncen_half(double dl, double dr)
{
    const double sum = dl + dr;
    const double sg = __builtin_copysign(1.0, sum);
    const double a = sg * dl, b = sg * dr;
    return sg * __builtin_fmax(__builtin_fmin(__builtin_fmin(a, b), 0.25 * __builtin_fabs(sum)), 0.0);
}
GLM_HD float moncen_half(float dl, float dr)
{
    const float sum = dl + dr;
    const float sg = __builtin_copysignf(1.0f, sum);
    const float a = sg * dl, b = sg * dr;
    return sg * __builtin_fmaxf(__builtin_fminf(__builtin_fminf(a, b), 0.25f * __builtin_fabsf(sum)), 0.0f);
}
template <typename T>
GLM_HD void ppm_offsets(T qm, T q0, T qp, T& dl, T& dr)
{
    T slope = T(0.25) * (qp - qm);
    T curve = (qm - T(2) * q0 + qp) * T(1.0 / 12.0);
    T a = slope - curve, b = slope + curve;
    // s = sign(slope): where a and b share a sign it is theirs (a + b = 2 slope up to rounding, which
    // keeps signs), and where they do not both offsets are 0 whatever s is.  s is applied as a
    // factor +-1 (exact), whose products need no NaN-quieting before the min/max.
    const T sg = tcopysign(T(1), slope);
    T sa = sg * a, sb = sg * b;
    dl = sg * tmax(tmin(sa, sb + sb), T(0));
    dr = sg * tmax(tmin(sb, sa + sa), T(0));
}

// spike_mhd.cu:164-172 ppm1 as edge offsets dl = q0 - lo, dr = hi - q0, in min/max form.  The
// unlimited edges are ql = q0 - a, qr = q0 + b with a = d + cu, b = d - cu (d = (qp - qm)/4,
// cu = (qm - 2 q0 + qp)/6).  ppm1's tests, with dqe = a + b and d6 = 3 (a - b): (flat)
// (qr - q0)(q0 - ql) = a b <= 0 -> both 0; dqe d6 > dqe^2 <=> (a + b)(a - 2b) > 0 -> ql = 3 q0 - 2 qr,
// i.e. a = 2b; dqe d6 < -dqe^2 <=> (a + b)(2a - b) < 0 -> qr = 3 q0 - 2 ql, i.e. b = 2a (the two
// exclude each other when a and b share a sign, so the order of ppm1's sequential ifs does not
// matter).  With s the sign of d (that of a + b = 2d where a and b share a sign):
// a' = s max(0, min(s a, 2 s b)) and b' = s max(0, min(s b, 2 s a)), the same limited parabola up
// to rounding (the limiter is continuous at each switch) and with no branch or select.
template <typename T>
GLM_HD void ppm_offsets_spike(T qm, T q0, T qp, T& dl, T& dr)
{
    const T d = T(0.25) * (qp - qm), cu = (qm - T(2) * q0 + qp) * (T(1) / T(6));
    const T a = d + cu, b = d - cu;
    const T sg = tcopysign(T(1), d);
    const T sa = sg * a, sb = sg * b;
    dl = sg * tmax(tmin(sa, sb + sb), T(0));
    dr = sg * tmax(tmin(sb, sa + sa), T(0));
}
// spike_mhd.cu:164-172 ppm1 (both edges).
template <typename T>
GLM_HD void ppm_edges_spike(T qm, T q0, T qp, T& lo, T& hi)
{
    T dl, dr;
    ppm_offsets_spike(qm, q0, qp, dl, dr);
    lo = q0 - dl;
    hi = q0 + dr;
}

// glmmhd_turb.jl:926 spj: the correctly rounded quotient hi / lo exceeds 2, decided without the
// division.  If hi > 2 lo, then hi >= 2 lo + ulp(2 lo), so hi / lo - 2 >= 2 ulp(lo) / lo > eps (the
// half ulp of 2) and the quotient rounds above 2; if hi <= 2 lo it does not exceed 2.  2 lo is
// exact (lo >= 1e-20) or, past half the largest finite value, infinite with the same answer.
template <typename T>
GLM_HD bool spj(T a, T b, T c)
{
    T hi = tmax(a, tmax(b, c));
    T lo = tmax(tmin(a, tmin(b, c)), T(1e-20));
    return hi > T(2) * lo;
}

// ----------------------------------------------------------------------------------------
// Riemann solvers in the x-normal frame.  L, R are floored, GLM-cleaned primitive states
// (Bx = bn*); fbn/fpsi are the GLM fluxes placed in slots 5 and 8.
// ----------------------------------------------------------------------------------------

// glmmhd_turb.jl:182-190 llf_x (exact fast speed) and :311-324 llf_bound.
template <typename T, bool BOUND>
GLM_HD void llf_x(const T L[9], const T R[9], T f[9], const Phys<T>& s)
{
    T b2L = L[5] * L[5] + L[6] * L[6] + L[7] * L[7];
    T b2R = R[5] * R[5] + R[6] * R[6] + R[7] * R[7];
    T cL, cR;
    if (BOUND) {
        cL = tsqrt_pos((s.gamma * L[4] + b2L) * trcp(L[0]));
        cR = tsqrt_pos((s.gamma * R[4] + b2R) * trcp(R[0]));
    } else {
        cL = fast_speed(L[0], L[4], b2L, L[5], s);
        cR = fast_speed(R[0], R[4], b2R, R[5], s);
    }
    T smax = tmax(tmax(tabs(L[1]) + cL, tabs(R[1]) + cR), s.ch);
    T fL[9], fR[9], uL[9], uR[9];
    flux_x(L, fL, s); flux_x(R, fR, s);
    prim2cons(L, uL, s); prim2cons(R, uR, s);
#pragma unroll
    for (int v = 0; v < 9; ++v) f[v] = T(0.5) * (fL[v] + fR[v]) - T(0.5) * smax * (uR[v] - uL[v]);
}

// glmmhd_turb.jl:335-343 hll_x
template <typename T>
GLM_HD void hll_x(const T L[9], const T R[9], T f[9], const Phys<T>& s)
{
    T bn = L[5];
    T cfL = fast_speed(L[0], L[4], L[5] * L[5] + L[6] * L[6] + L[7] * L[7], bn, s);
    T cfR = fast_speed(R[0], R[4], R[5] * R[5] + R[6] * R[6] + R[7] * R[7], bn, s);
    T SL = tmin(tmin(L[1] - cfL, R[1] - cfR), T(0));
    T SR = tmax(tmax(L[1] + cfL, R[1] + cfR), T(0));
    T fL[9], fR[9], uL[9], uR[9];
    flux_x(L, fL, s); flux_x(R, fR, s);
    prim2cons(L, uL, s); prim2cons(R, uR, s);
    T ihd = trcp(SR - SL);
    T slsr = SL * SR;
#pragma unroll
    for (int v = 0; v < 9; ++v) f[v] = (SR * fL[v] - SL * fR[v] + slsr * (uR[v] - uL[v])) * ihd;
}

// glmmhd_turb.jl:193-272 hlld_x (Miyoshi & Kusano 2005).  Only the side that the sampling
// selects (SM >= 0 -> left) has its 9-vectors formed; the scalar star/double-star quantities
// of both sides are computed as in the reference.
template <typename T>
GLM_HD void hlld_x(const T L[9], const T R[9], T f[9], const Phys<T>& s)
{
    const T bn = L[5];
    const T dL = L[0], uL = L[1], vL = L[2], wL = L[3], pL = L[4], byL = L[6], bzL = L[7];
    const T dR = R[0], uR = R[1], vR = R[2], wR = R[3], pR = R[4], byR = R[6], bzR = R[7];
    T b2L = bn * bn + byL * byL + bzL * bzL;
    T b2R = bn * bn + byR * byR + bzR * bzR;
    T ptL = pL + T(0.5) * b2L, ptR = pR + T(0.5) * b2R;
    T cfL = fast_speed(dL, pL, b2L, bn, s);
    T cfR = fast_speed(dR, pR, b2R, bn, s);
    T cfm = tmax(cfL, cfR);
    T SL = tmin(tmin(uL, uR) - cfm, T(0));
    T SR = tmax(tmax(uL, uR) + cfm, T(0));
    T aL = SL - uL, aR = SR - uR;
    T iden = trcp(aR * dR - aL * dL);
    T SM = (aR * dR * uR - aL * dL * uL - ptR + ptL) * iden;
    T pts = (aR * dR * ptL - aL * dL * ptR + dL * dR * aR * aL * (uR - uL)) * iden;
    T iL = trcp(SL - SM), iR = trcp(SR - SM);
    T dLs = tmax(dL * aL * iL, T(1e-12));
    T dRs = tmax(dR * aR * iR, T(1e-12));
    T rsL, rsR;
    T sqdLs = tsqrt_rs(dLs, rsL), sqdRs = tsqrt_rs(dRs, rsR);
    T absbn = tabs(bn);
    T SLs = SM - absbn * rsL;
    T SRs = SM + absbn * rsR;
    // star states (glmmhd_turb.jl:227-239)
    T vyLs, vzLs, byLs, bzLs, vyRs, vzRs, byRs, bzRs;
    {
        // the degenerate star (|denK| < 1e-12) keeps the side's state: fac = 0 and gb = 1 give it exactly
        T denK = dL * aL * (SL - SM) - bn * bn;
        bool sm = tabs(denK) < T(1e-12);
        T ik = trcp(denK);
        T fac = sm ? T(0) : bn * (SM - uL) * ik;
        T gb = sm ? T(1) : (dL * aL * aL - bn * bn) * ik;
        vyLs = vL - byL * fac;
        vzLs = wL - bzL * fac;
        byLs = byL * gb;
        bzLs = bzL * gb;
    }
    {
        T denK = dR * aR * (SR - SM) - bn * bn;
        bool sm = tabs(denK) < T(1e-12);
        T ik = trcp(denK);
        T fac = sm ? T(0) : bn * (SM - uR) * ik;
        T gb = sm ? T(1) : (dR * aR * aR - bn * bn) * ik;
        vyRs = vR - byR * fac;
        vzRs = wR - bzR * fac;
        byRs = byR * gb;
        bzRs = bzR * gb;
    }
    T sgn = bn >= T(0) ? T(1) : T(-1);
    T invsum = trcp(sqdLs + sqdRs);
    T vyss = (sqdLs * vyLs + sqdRs * vyRs + (byRs - byLs) * sgn) * invsum;
    T vzss = (sqdLs * vzLs + sqdRs * vzRs + (bzRs - bzLs) * sgn) * invsum;
    T byss = (sqdLs * byRs + sqdRs * byLs + sqdLs * sqdRs * (vyRs - vyLs) * sgn) * invsum;
    T bzss = (sqdLs * bzRs + sqdRs * bzLs + sqdLs * sqdRs * (vzRs - vzLs) * sgn) * invsum;
    T vdotbss = SM * bn + vyss * byss + vzss * bzss;
    // Select the sampled side: (SM >= 0) -> left fan, else right fan.
    const bool left = SM >= T(0);
    const T d = left ? dL : dR, u = left ? uL : uR, v = left ? vL : vR, w = left ? wL : wR;
    const T p = left ? pL : pR, by = left ? byL : byR, bz = left ? bzL : bzR;
    const T b2 = left ? b2L : b2R, pt = left ? ptL : ptR;
    const T S = left ? SL : SR, a = left ? aL : aR, iS = left ? iL : iR;
    const T ds = left ? dLs : dRs, sqd = left ? sqdLs : sqdRs, Ss = left ? SLs : SRs;
    const T vys = left ? vyLs : vyRs, vzs = left ? vzLs : vzRs;
    const T bys = left ? byLs : byRs, bzs = left ? bzLs : bzRs;
    T E = p * s.igm1 + T(0.5) * d * (u * u + v * v + w * w) + T(0.5) * b2;
    T vdotb = u * bn + v * by + w * bz;
    T vdotbs = SM * bn + vys * bys + vzs * bzs;
    T Es = (a * E - pt * u + pts * SM + bn * (vdotb - vdotbs)) * iS;
    T sg = left ? -sgn : sgn;                         // ELss = ELs - ..., ERss = ERs + ...
    T Ess = Es + sqd * (vdotbs - vdotbss) * sg;
    bool outer = left ? (SLs >= T(0)) : !(SRs >= T(0));
    T q[9] = {d, u, v, w, p, bn, by, bz, T(0)};
    T F[9];
    flux_x(q, F, s);
    T U[9] = {d, d * u, d * v, d * w, E, bn, by, bz, T(0)};
    T Us[9] = {ds, ds * SM, ds * vys, ds * vzs, Es, bn, bys, bzs, T(0)};
    T Uss[9] = {ds, ds * SM, ds * vyss, ds * vzss, Ess, bn, byss, bzss, T(0)};
    // F* = F + S (U* - U); F** = F + S** U** - (S** - S) U* - S U = F* + S** (U** - U*), and U** - U*
    // vanishes in rho, m_n and Bn (glmmhd_turb.jl:262-271)
    const T w2 = outer ? T(0) : Ss;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        T fo = F[k] + S * (Us[k] - U[k]);
        f[k] = (k == 0 || k == 1 || k == 5 || k == 8) ? fo : fo + w2 * (Uss[k] - Us[k]);
    }
}

// Riemann dispatch in direction D: floors (glmmhd_turb.jl:280-281), GLM pair (:175-179,
// :282-286), solver, rotation back.  RIEM: LLF / LLF-bound / HLL / HLLD (with the LLF switch).
template <typename T, int RIEM, int D>
GLM_HD void riemann(const T Lq[9], const T Rq[9], T out[9], const Phys<T>& s)
{
    constexpr int A = D % 3, B = (D + 1) % 3, C = (D + 2) % 3;
    T L[9] = {tmax(Lq[0], s.smallr), Lq[1 + A], Lq[1 + B], Lq[1 + C], tmax(Lq[4], s.pfl),
              Lq[5 + A], Lq[5 + B], Lq[5 + C], Lq[8]};
    T R[9] = {tmax(Rq[0], s.smallr), Rq[1 + A], Rq[1 + B], Rq[1 + C], tmax(Rq[4], s.pfl),
              Rq[5 + A], Rq[5 + B], Rq[5 + C], Rq[8]};
    T bns = T(0.5) * (L[5] + R[5]) - (R[8] - L[8]) * s.half_ich;
    T psis = T(0.5) * (L[8] + R[8]) - T(0.5) * s.ch * (R[5] - L[5]);
    L[5] = bns; R[5] = bns;
    T f[9];
    if (RIEM == RIEM_LLF) {
        llf_x<T, false>(L, R, f, s);
    } else if (RIEM == RIEM_LLFB) {
        llf_x<T, true>(L, R, f, s);
    } else if (RIEM == RIEM_HLL) {
        hll_x(L, R, f, s);
    } else {
        bool uself = (s.llf_dmin > T(0) && tmin(L[0], R[0]) < s.llf_dmin) ||
                     (s.llf_pmin > T(0) && tmin(L[4], R[4]) < s.llf_pmin);
        if (uself) llf_x<T, false>(L, R, f, s);
        else hlld_x(L, R, f, s);
    }
    out[0] = f[0];
    out[1 + A] = f[1]; out[1 + B] = f[2]; out[1 + C] = f[3];
    out[4] = f[4];
    out[5 + A] = psis; out[5 + B] = f[6]; out[5 + C] = f[7];
    out[8] = s.ch2 * bns;
}

}  // namespace glm
