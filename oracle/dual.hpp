// dual.hpp -- forward-mode dual numbers (value, tangent).  TEST INFRASTRUCTURE ONLY: the CPU
// oracle compiled with REAL=Dual (liboracle_dual.so) and the physics templates instantiated with
// Dual compute tangents J v exactly (to rounding), independently of the hand-written adjoint,
// which the tests check against them.
#pragma once
#include <cmath>

struct Dual {
    double v, d;
    Dual() = default;
    Dual(double x) : v(x), d(0) {}
    Dual(double x, double dx) : v(x), d(dx) {}
    explicit operator double() const { return v; }
    explicit operator int() const { return (int)v; }
};
inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
inline Dual operator-(Dual a) { return {-a.v, -a.d}; }
inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
inline Dual operator/(Dual a, Dual b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
inline Dual& operator+=(Dual& a, Dual b) { a = a + b; return a; }
inline Dual& operator-=(Dual& a, Dual b) { a = a - b; return a; }
inline Dual& operator*=(Dual& a, Dual b) { a = a * b; return a; }
inline bool operator<(Dual a, Dual b) { return a.v < b.v; }
inline bool operator<=(Dual a, Dual b) { return a.v <= b.v; }
inline bool operator>(Dual a, Dual b) { return a.v > b.v; }
inline bool operator>=(Dual a, Dual b) { return a.v >= b.v; }
inline bool operator==(Dual a, Dual b) { return a.v == b.v; }
inline bool operator!=(Dual a, Dual b) { return a.v != b.v; }

// Found by argument-dependent lookup from the physics templates (Dual is a global type).
inline Dual tmax(Dual a, Dual b) { return a.v >= b.v ? a : b; }
inline Dual tmin(Dual a, Dual b) { return a.v <= b.v ? a : b; }
inline Dual tabs(Dual a) { return a.v < 0 ? -a : a; }
inline Dual tcopysign(Dual m, Dual s) { return (std::signbit(m.v) != std::signbit(s.v)) ? -m : m; }
inline Dual trcp(Dual a) { return {1.0 / a.v, -a.d / (a.v * a.v)}; }
inline Dual tsqrt(Dual a)
{
    double r = std::sqrt(a.v);
    return {r, r > 0 ? a.d / (2 * r) : 0.0};
}
inline Dual gas_pressure(const Dual c[9], Dual ir, Dual gm1, Dual pfl)
{
    Dual ekin = Dual(0.5) * (c[1] * c[1] + c[2] * c[2] + c[3] * c[3]) * ir;
    Dual emag = Dual(0.5) * (c[5] * c[5] + c[6] * c[6] + c[7] * c[7]);
    return tmax(gm1 * (c[4] - ekin - emag), pfl);
}

// C math names used by the oracle (glm_oracle.c compiled as C++ with REAL=Dual)
inline Dual fmax(Dual a, Dual b) { return a.v >= b.v ? a : b; }
inline Dual fmin(Dual a, Dual b) { return a.v <= b.v ? a : b; }
inline Dual fabs(Dual a) { return a.v < 0 ? -a : a; }
inline Dual sqrt(Dual a) { return tsqrt(a); }
inline Dual floor(Dual a) { return Dual(std::floor(a.v)); }
inline Dual sin(Dual a) { return {std::sin(a.v), std::cos(a.v) * a.d}; }
