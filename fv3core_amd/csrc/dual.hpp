// Forward-mode derivative numbers for the linearisation of the PPM stencils.
//
// Dual<N> carries a value and its partial derivatives with respect to N
// chosen inputs.  The value part performs exactly the operations of the
// plain double code, in the same order, so templated stencil code run on
// Dual<N> returns the forward value bit for bit together with the
// derivatives at the branch decisions that value took.  Branching functions
// (fabs, fmin, fmax, comparisons) decide on the value; at a tie the
// derivative of the operand the double code would return is taken.
#pragma once

#include "common.hpp"

namespace fv3 {

// keep the double overloads visible next to the Dual ones below
using ::fabs;
using ::fmax;
using ::fmin;

template <int N>
struct Dual {
  double v;
  double d[N];
  __device__ __forceinline__ Dual() : v(0.0) {
#pragma unroll
    for (int n = 0; n < N; ++n) d[n] = 0.0;
  }
  __device__ __forceinline__ Dual(double x) : v(x) {  // a constant
#pragma unroll
    for (int n = 0; n < N; ++n) d[n] = 0.0;
  }
  // the value part, for code that stores flags as numbers and reads them back
  __device__ __forceinline__ explicit operator int() const { return (int)v; }
  // value x with unit derivative in direction m
  __device__ __forceinline__ static Dual seed(double x, int m) {
    Dual r(x);
#pragma unroll
    for (int n = 0; n < N; ++n) r.d[n] = n == m ? 1.0 : 0.0;
    return r;
  }
};

template <int N>
__device__ __forceinline__ Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a.v + b.v;
#pragma unroll
  for (int n = 0; n < N; ++n) r.d[n] = a.d[n] + b.d[n];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a.v - b.v;
#pragma unroll
  for (int n = 0; n < N; ++n) r.d[n] = a.d[n] - b.d[n];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N>& a) {
  Dual<N> r;
  r.v = -a.v;
#pragma unroll
  for (int n = 0; n < N; ++n) r.d[n] = -a.d[n];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a.v * b.v;
#pragma unroll
  for (int n = 0; n < N; ++n) r.d[n] = a.d[n] * b.v + a.v * b.d[n];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a.v / b.v;
#pragma unroll
  for (int n = 0; n < N; ++n) r.d[n] = (a.d[n] - r.v * b.d[n]) / b.v;
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N>& operator+=(Dual<N>& a, const Dual<N>& b) {
  a = a + b;
  return a;
}
template <int N>
__device__ __forceinline__ Dual<N>& operator-=(Dual<N>& a, const Dual<N>& b) {
  a = a - b;
  return a;
}
template <int N>
__device__ __forceinline__ bool operator<(const Dual<N>& a, const Dual<N>& b) {
  return a.v < b.v;
}
template <int N>
__device__ __forceinline__ bool operator>(const Dual<N>& a, const Dual<N>& b) {
  return a.v > b.v;
}
template <int N>
__device__ __forceinline__ bool operator<=(const Dual<N>& a, const Dual<N>& b) {
  return a.v <= b.v;
}
template <int N>
__device__ __forceinline__ bool operator>=(const Dual<N>& a, const Dual<N>& b) {
  return a.v >= b.v;
}
// c ? a : b, component by component (a conditional expression on two
// structures would be lowered through memory)
template <int N>
__device__ __forceinline__ Dual<N> pick(bool c, const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  r.v = c ? a.v : b.v;
#pragma unroll
  for (int n = 0; n < N; ++n) r.d[n] = c ? a.d[n] : b.d[n];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> fabs(const Dual<N>& a) {
  const bool neg = a.v < 0.0;
  Dual<N> r;
  r.v = ::fabs(a.v);  // the value is the double fabs, signed zeros included
#pragma unroll
  for (int n = 0; n < N; ++n) r.d[n] = neg ? -a.d[n] : a.d[n];
  return r;
}
// fmin/fmax of IEEE doubles return the other operand when one is NaN; the
// stencils never pass NaN, so the comparison form is equivalent here.
template <int N>
__device__ __forceinline__ Dual<N> fmin(const Dual<N>& a, const Dual<N>& b) {
  return pick(b.v < a.v, b, a);
}
template <int N>
__device__ __forceinline__ Dual<N> fmax(const Dual<N>& a, const Dual<N>& b) {
  return pick(b.v > a.v, b, a);
}

}  // namespace fv3
