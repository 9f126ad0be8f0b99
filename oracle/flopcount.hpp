/* Counting build of the oracle (test infrastructure): crimson_oracle.c is
 * compiled as C++ with every `double` replaced by cfo_real, a one-double
 * struct (the same size, alignment and calling convention as double) whose
 * arithmetic counts each +, -, *, / and sqrt as one floating-point
 * operation.  Negation, fabs and comparisons are not counted.  The values
 * computed are those of the plain build (each operation rounds once, as with
 * -ffp-contract=off). */
#pragma once
#include <math.h>
#include <stdlib.h>
#include <string.h>

extern unsigned long long cfo_nflop;

struct cfo_real {
  double v;
  cfo_real() = default;
  constexpr cfo_real(double x) : v(x) {}
  explicit operator int() const { return (int)v; }
  explicit operator long() const { return (long)v; }
  explicit operator long long() const { return (long long)v; }
  explicit operator unsigned long() const { return (unsigned long)v; }
  explicit operator float() const { return (float)v; }
  cfo_real &operator+=(cfo_real b) { ++cfo_nflop; v += b.v; return *this; }
  cfo_real &operator-=(cfo_real b) { ++cfo_nflop; v -= b.v; return *this; }
  cfo_real &operator*=(cfo_real b) { ++cfo_nflop; v *= b.v; return *this; }
  cfo_real &operator/=(cfo_real b) { ++cfo_nflop; v /= b.v; return *this; }
};
inline cfo_real operator+(cfo_real a, cfo_real b) { ++cfo_nflop; return cfo_real(a.v + b.v); }
inline cfo_real operator-(cfo_real a, cfo_real b) { ++cfo_nflop; return cfo_real(a.v - b.v); }
inline cfo_real operator*(cfo_real a, cfo_real b) { ++cfo_nflop; return cfo_real(a.v * b.v); }
inline cfo_real operator/(cfo_real a, cfo_real b) { ++cfo_nflop; return cfo_real(a.v / b.v); }
inline cfo_real operator-(cfo_real a) { return cfo_real(-a.v); }
inline cfo_real operator+(cfo_real a) { return a; }
inline bool operator<(cfo_real a, cfo_real b) { return a.v < b.v; }
inline bool operator>(cfo_real a, cfo_real b) { return a.v > b.v; }
inline bool operator<=(cfo_real a, cfo_real b) { return a.v <= b.v; }
inline bool operator>=(cfo_real a, cfo_real b) { return a.v >= b.v; }
inline bool operator==(cfo_real a, cfo_real b) { return a.v == b.v; }
inline bool operator!=(cfo_real a, cfo_real b) { return a.v != b.v; }
inline cfo_real sqrt(cfo_real a) { ++cfo_nflop; return cfo_real(::sqrt(a.v)); }
inline cfo_real fabs(cfo_real a) { return cfo_real(::fabs(a.v)); }
inline cfo_real log(cfo_real a) { return cfo_real(::log(a.v)); }
inline long lround(cfo_real a) { return ::lround(a.v); }

#define double cfo_real
