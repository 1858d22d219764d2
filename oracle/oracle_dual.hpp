// oracle_dual.hpp -- a forward-mode dual number for the parity oracle.
//
// TEST INFRASTRUCTURE ONLY (see partia_oracle.hpp).  Dual{v, d} carries a value
// and its derivative along one direction of the inputs; running the oracle's
// interpolation templates with T = Dual gives the outputs and their directional
// derivatives J * delta exactly as ForwardDiff would through the reference's
// generic Julia code (T <: AbstractFloat, KI/kernel_function/kernel.jl:141).
// Comparisons, integer truncation and NaN tests use the value only, so every
// branch (accepted neighbours, support cut-offs, table cell) is the float64
// run's.  The derivative rules are the textbook ones: (a b)' = a' b + a b',
// (a / b)' = (a' b - a b') / b^2, sqrt(a)' = a' / (2 sqrt(a)).
#pragma once

#include <cmath>
#include <limits>

namespace oracle {

struct Dual {
    double v = 0.0, d = 0.0;
    Dual() = default;
    Dual(double value) : v(value), d(0.0) {}  // constants: zero derivative
    Dual(double value, double deriv) : v(value), d(deriv) {}
    explicit operator double() const { return v; }
    explicit operator float() const { return (float)v; }
    explicit operator int() const { return (int)v; }
    Dual& operator+=(const Dual& b) { v += b.v; d += b.d; return *this; }
    Dual& operator-=(const Dual& b) { v -= b.v; d -= b.d; return *this; }
    Dual& operator*=(const Dual& b) { *this = Dual{v * b.v, d * b.v + v * b.d}; return *this; }
    Dual& operator/=(const Dual& b) { *this = Dual{v / b.v, (d * b.v - v * b.d) / (b.v * b.v)}; return *this; }
};

inline Dual operator+(Dual a, const Dual& b) { return a += b; }
inline Dual operator-(Dual a, const Dual& b) { return a -= b; }
inline Dual operator*(Dual a, const Dual& b) { return a *= b; }
inline Dual operator/(Dual a, const Dual& b) { return a /= b; }
inline Dual operator-(const Dual& a) { return Dual{-a.v, -a.d}; }
inline bool operator<(const Dual& a, const Dual& b) { return a.v < b.v; }
inline bool operator>(const Dual& a, const Dual& b) { return a.v > b.v; }
inline bool operator<=(const Dual& a, const Dual& b) { return a.v <= b.v; }
inline bool operator>=(const Dual& a, const Dual& b) { return a.v >= b.v; }
inline bool operator==(const Dual& a, const Dual& b) { return a.v == b.v; }
inline bool operator!=(const Dual& a, const Dual& b) { return a.v != b.v; }
inline Dual sqrt(const Dual& a) {
    const double r = std::sqrt(a.v);
    return Dual{r, a.d / (2.0 * r)};
}

}  // namespace oracle

namespace std {
template <> struct numeric_limits<oracle::Dual> {
    static constexpr bool is_specialized = true;
    static oracle::Dual infinity() { return oracle::Dual{numeric_limits<double>::infinity()}; }
    static oracle::Dual quiet_NaN() { return oracle::Dual{numeric_limits<double>::quiet_NaN()}; }
    static oracle::Dual max() { return oracle::Dual{numeric_limits<double>::max()}; }
    static oracle::Dual epsilon() { return oracle::Dual{numeric_limits<double>::epsilon()}; }
};
}  // namespace std
