// sc_device.hpp -- device arithmetic shared by the kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace scamd {

template <typename T> struct Cx {
  T re, im;
};

template <typename T> __device__ __forceinline__ Cx<T> cx(T r, T i) {
  return Cx<T>{r, i};
}
template <typename T> __device__ __forceinline__ Cx<T> operator+(Cx<T> a, Cx<T> b) {
  return {a.re + b.re, a.im + b.im};
}
template <typename T> __device__ __forceinline__ Cx<T> operator-(Cx<T> a, Cx<T> b) {
  return {a.re - b.re, a.im - b.im};
}
template <typename T> __device__ __forceinline__ Cx<T> operator*(Cx<T> a, Cx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T> __device__ __forceinline__ Cx<T> scale(Cx<T> a, T s) {
  return {a.re * s, a.im * s};
}

// Square root and reciprocal: FP64 uses the IEEE operations; FP32 uses the
// hardware v_sqrt_f32 / v_rcp_f32 (1 ulp).
__device__ __forceinline__ double dsqrt(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float dsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double drcp(double x) { return 1.0 / x; }
__device__ __forceinline__ float drcp(float x) { return __builtin_amdgcn_rcpf(x); }

// Principal complex square root (Kahan's formulation): with
// t = sqrt((|x| + |z|) / 2), sqrt(z) = (t, y / 2t) for x >= 0 and
// (|y| / 2t, copysign(t, y)) otherwise.
template <typename T> __device__ __forceinline__ Cx<T> csqrt(Cx<T> z) {
  const T ax = fabs(z.re);
  const T mod = dsqrt(z.re * z.re + z.im * z.im);
  const T t = dsqrt(T(0.5) * (ax + mod));
  if (t == T(0)) return {T(0), z.im};
  const T q = z.im * drcp(T(2) * t);
  if (z.re >= T(0)) return {t, q};
  return {fabs(q), copysign(t, z.im)};
}

// a / b through the reciprocal of |b|^2 (|b| is of order one here).
template <typename T> __device__ __forceinline__ Cx<T> cdiv(Cx<T> a, Cx<T> b) {
  const T r = drcp(b.re * b.re + b.im * b.im);
  return {(a.re * b.re + a.im * b.im) * r, (a.im * b.re - a.re * b.im) * r};
}

// One implicit-midpoint step of the Riccati equation solved as the quadratic
// a y^2 + b y + c = 0 (RiccatiSolutions.h:65-94):
//   s = i h, a = s conj(D) / 4, b = 2 a y0 + s e - 1,
//   c = (a y0 + s e + 1) y0 + s D,  y = 2 c / (-b + sqrt(b^2 - 4 a c)).
template <typename T>
__device__ __forceinline__ Cx<T> riccatiStep(Cx<T> D, Cx<T> e, Cx<T> y0, T h) {
  const T qh = T(0.25) * h;
  const Cx<T> a = {qh * D.im, qh * D.re};
  const Cx<T> se = {-h * e.im, h * e.re};
  const Cx<T> ay = a * y0;
  const Cx<T> b = {T(2) * ay.re + se.re - T(1), T(2) * ay.im + se.im};
  const Cx<T> t = {ay.re + se.re + T(1), ay.im + se.im};
  const Cx<T> c = t * y0 + Cx<T>{-h * D.im, h * D.re};
  const Cx<T> bb = b * b;
  const Cx<T> ac = a * c;
  const Cx<T> disc = csqrt(Cx<T>{bb.re - T(4) * ac.re, bb.im - T(4) * ac.im});
  return cdiv(Cx<T>{T(2) * c.re, T(2) * c.im}, Cx<T>{disc.re - b.re, disc.im - b.im});
}

// Homogeneous (bulk) solutions (RiccatiSolutions.h:36-54).
template <typename T>
__device__ __forceinline__ void homogeneous(Cx<T> D, Cx<T> e, Cx<T> &g, Cx<T> &gt) {
  const Cx<T> e2 = e * e;
  const Cx<T> om = csqrt(Cx<T>{D.re * D.re + D.im * D.im - e2.re, T(0) - e2.im});
  const Cx<T> den = {e.re - om.im, e.im + om.re}; // e + i Omega
  const Cx<T> inv = cdiv(Cx<T>{T(1), T(0)}, den);
  g = Cx<T>{-(D.re * inv.re - D.im * inv.im), -(D.re * inv.im + D.im * inv.re)};
  gt = Cx<T>{D.re * inv.re + D.im * inv.im, D.re * inv.im - D.im * inv.re};
}

// Bilinear sample of a planar N x N grid at the rotated position of a site of
// a grid of side M (insertRotatedGrid, GridFunctions_GPU.h:436-489).  The
// coordinate arithmetic is kept free of contractions so that the integer
// cell decisions are taken on the same rounded values on every platform.
struct RotSample {
  int idx;      // Yi * N + Xi
  double fx, fy;
  bool valid;
};
__device__ __forceinline__ RotSample rotatedCell(int x, int y, int M, int N,
                                                 double c, double s) {
#pragma clang fp contract(off)
  const double xt = (N - M) / 2.0;
  const double sh = (N - 1) / 2.0;
  const double X = x + xt - sh;
  const double Y = y + xt - sh;
  const double Xr = X * c - Y * s + sh;
  const double Yr = X * s + Y * c + sh;
  const int Xi = static_cast<int>(Xr);
  const int Yi = static_cast<int>(Yr);
  RotSample r;
  r.valid = (Xi >= 0) && (Xi + 1 < N) && (Yi >= 0) && (Yi + 1 < N);
  r.fx = Xr - static_cast<double>(Xi);
  r.fy = Yr - static_cast<double>(Yi);
  r.idx = Yi * N + Xi;
  return r;
}
__device__ __forceinline__ double bilinear(const double *__restrict__ p, int N,
                                           const RotSample &r) {
#pragma clang fp contract(off)
  const double a = (1.0 - r.fx) * p[r.idx] + r.fx * p[r.idx + 1];
  const double b = (1.0 - r.fx) * p[r.idx + N] + r.fx * p[r.idx + N + 1];
  return (1.0 - r.fy) * a + r.fy * b;
}

} // namespace scamd
