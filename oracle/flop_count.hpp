// flop_count.hpp -- a scalar that counts floating-point operations (TEST INFRASTRUCTURE).
//
// Instantiating the oracle's templated algorithms with `Flop` yields the algorithmic
// operation count of the reference's operation sequence for one configuration: every
// +, -, *, / and every sqrt/sin/cos counts one operation (a multiply-add counts two).
// bench.py uses these counts for the roofline's algorithmic FLOP per configuration.
#pragma once

#include <cmath>
#include <limits>

namespace oracle {

struct Flop {
  double v = 0.0;
  Flop() = default;
  Flop(double x) : v(x) {}  // NOLINT: implicit, like a scalar
  explicit operator double() const { return v; }
  static long long& count() {
    static thread_local long long n = 0;
    return n;
  }
};
inline Flop operator+(Flop a, Flop b) { ++Flop::count(); return Flop(a.v + b.v); }
inline Flop operator-(Flop a, Flop b) { ++Flop::count(); return Flop(a.v - b.v); }
inline Flop operator*(Flop a, Flop b) { ++Flop::count(); return Flop(a.v * b.v); }
inline Flop operator/(Flop a, Flop b) { ++Flop::count(); return Flop(a.v / b.v); }
inline Flop operator-(Flop a) { return Flop(-a.v); }
inline Flop& operator+=(Flop& a, Flop b) { a = a + b; return a; }
inline Flop& operator-=(Flop& a, Flop b) { a = a - b; return a; }
inline Flop& operator*=(Flop& a, Flop b) { a = a * b; return a; }
inline Flop& operator/=(Flop& a, Flop b) { a = a / b; return a; }
inline bool operator<(Flop a, Flop b) { return a.v < b.v; }
inline bool operator>(Flop a, Flop b) { return a.v > b.v; }
inline bool operator<=(Flop a, Flop b) { return a.v <= b.v; }
inline bool operator>=(Flop a, Flop b) { return a.v >= b.v; }
inline bool operator==(Flop a, Flop b) { return a.v == b.v; }
inline bool operator!=(Flop a, Flop b) { return a.v != b.v; }
inline Flop sqrt(Flop a) { ++Flop::count(); return Flop(std::sqrt(a.v)); }
inline Flop sin(Flop a) { ++Flop::count(); return Flop(std::sin(a.v)); }
inline Flop cos(Flop a) { ++Flop::count(); return Flop(std::cos(a.v)); }

}  // namespace oracle

namespace std {
template<> struct numeric_limits<oracle::Flop> : numeric_limits<double> {
  static oracle::Flop epsilon() { return oracle::Flop(numeric_limits<double>::epsilon()); }
};
}  // namespace std
