// k_dual.h -- forward-mode dual numbers for the discrete adjoint: a value
// and K exact directional derivatives, carried through the same face
// arithmetic (FaceMath<T>) the residual uses.  Every operation applies its
// derivative rule to the K tangents; branches (limiters, max/min, the
// supersonic switches) follow the values, so each face's local Jacobian is
// the derivative of the branch the primal takes.
#pragma once
#include "blk.h"

namespace fest {

template <int K>
struct Dual {
  double v;
  double d[K];
  __device__ Dual() : v(0.0) {
#pragma unroll
    for (int k = 0; k < K; ++k) d[k] = 0.0;
  }
  // constants (and integers through double) convert implicitly
  __device__ Dual(double x) : v(x) {
#pragma unroll
    for (int k = 0; k < K; ++k) d[k] = 0.0;
  }

  friend __device__ __forceinline__ Dual operator+(const Dual &a, const Dual &b) {
    Dual r;
    r.v = a.v + b.v;
#pragma unroll
    for (int k = 0; k < K; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
  }
  friend __device__ __forceinline__ Dual operator-(const Dual &a, const Dual &b) {
    Dual r;
    r.v = a.v - b.v;
#pragma unroll
    for (int k = 0; k < K; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
  }
  friend __device__ __forceinline__ Dual operator-(const Dual &a) {
    Dual r;
    r.v = -a.v;
#pragma unroll
    for (int k = 0; k < K; ++k) r.d[k] = -a.d[k];
    return r;
  }
  friend __device__ __forceinline__ Dual operator*(const Dual &a, const Dual &b) {
    Dual r;
    r.v = a.v * b.v;
#pragma unroll
    for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
  }
  friend __device__ __forceinline__ Dual operator/(const Dual &a, const Dual &b) {
    Dual r;
    const double ib = 1.0 / b.v;
    r.v = a.v * ib;
#pragma unroll
    for (int k = 0; k < K; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
    return r;
  }
  friend __device__ __forceinline__ bool operator<(const Dual &a, const Dual &b) { return a.v < b.v; }
  friend __device__ __forceinline__ bool operator>(const Dual &a, const Dual &b) { return a.v > b.v; }
  friend __device__ __forceinline__ bool operator<=(const Dual &a, const Dual &b) { return a.v <= b.v; }
  friend __device__ __forceinline__ bool operator>=(const Dual &a, const Dual &b) { return a.v >= b.v; }

  // the face arithmetic's helpers (blk.h, k_flux.h) for dual arguments
  friend __device__ __forceinline__ Dual tdiv(const Dual &a, const Dual &b) { return a / b; }
  friend __device__ __forceinline__ Dual tsqrt(const Dual &a) {
    Dual r;
    r.v = ::sqrt(a.v);
    const double h = a.v > 0.0 ? 0.5 / r.v : 0.0;
#pragma unroll
    for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * h;
    return r;
  }
  friend __device__ __forceinline__ Dual tabs(const Dual &a) { return a.v < 0.0 ? -a : a; }
  friend __device__ __forceinline__ Dual fmx(const Dual &a, const Dual &b) { return a.v > b.v ? a : b; }
  friend __device__ __forceinline__ Dual fmn(const Dual &a, const Dual &b) { return a.v < b.v ? a : b; }
  // sign transfer |a| sign(b): the derivative of |a|, times sign(b)
  friend __device__ __forceinline__ Dual fsgn(const Dual &a, const Dual &b) {
    const Dual m = tabs(a);
    return signbit(b.v) ? -m : m;
  }
  friend __device__ __forceinline__ Dual sqd(const Dual &a) { return a * a; }
  // the names the ghost-cell rules call (k_bc.h riemann_state)
  friend __device__ __forceinline__ Dual sqrt(const Dual &a) { return tsqrt(a); }
  friend __device__ __forceinline__ Dual pow(const Dual &a, double e) {
    Dual r;
    r.v = ::pow(a.v, e);
    const double h = e * ::pow(a.v, e - 1.0);
#pragma unroll
    for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * h;
    return r;
  }
};

}  // namespace fest
