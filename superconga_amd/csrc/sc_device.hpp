// sc_device.hpp -- device arithmetic shared by the kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <limits>

namespace scamd {

// Complex value aligned to its full size so that it moves as one vector
// access.
template <typename T> struct alignas(2 * sizeof(T)) Cx {
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

// Device pointers that reach a kernel inside a struct are generic: accesses
// through them compile to flat_* instructions, which also count against the
// LDS/scalar wait counter, so each wait for a lane shuffle or a scalar load
// would also wait for them.  Every buffer the kernels address lives in global
// memory; through global() the accesses are global_* instructions.
template <typename T> using GlobalPtr = __attribute__((address_space(1))) T *;
template <typename T> __device__ __forceinline__ GlobalPtr<T> global(T *p) {
  return (GlobalPtr<T>)p;
}

// Streaming access to the sweep scratch: every element is written once and
// read once, so the accesses carry the non-temporal hint and leave the cache
// to the grids the sweep samples.
template <typename T> struct VecOf;
template <> struct VecOf<double> { typedef double type __attribute__((ext_vector_type(2))); };
template <> struct VecOf<float> { typedef float type __attribute__((ext_vector_type(2))); };
template <typename T> __device__ __forceinline__ void storeStream(Cx<T> *p, Cx<T> v) {
  typedef typename VecOf<T>::type V;
  V w;
  w.x = v.re;
  w.y = v.im;
  __builtin_nontemporal_store(w, reinterpret_cast<V *>(p));
}
template <typename T> __device__ __forceinline__ Cx<T> loadStream(const Cx<T> *p) {
  typedef typename VecOf<T>::type V;
  const V w = __builtin_nontemporal_load(reinterpret_cast<const V *>(p));
  return Cx<T>{w.x, w.y};
}

// Reciprocal and reciprocal square root.  FP32: the hardware v_rcp_f32 /
// v_rsq_f32 (1 ulp).  FP64: the hardware v_rcp_f64 / v_rsq_f64 estimate
// refined by kNewtonSteps Newton-Raphson iterations (each doubles the number
// of correct bits of the roughly single-precision estimate; two reach the
// last bits of the double), instead of the longer IEEE division and square
// root sequences.
constexpr int kNewtonSteps = 2;
__device__ __forceinline__ float rcpT(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float rsqT(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double rcpT(double x) {
  double r = __builtin_amdgcn_rcp(x);
#pragma unroll
  for (int i = 0; i < kNewtonSteps; ++i) r = fma(r, fma(-x, r, 1.0), r);
  return r;
}
__device__ __forceinline__ double rsqT(double x) {
  // r += r (1 - x r^2) / 2, with the halving applied to x once: scaling by
  // 0.5 is exact, so this is the same arithmetic one multiply shorter
  const double hx = 0.5 * x;
  double r = __builtin_amdgcn_rsq(x);
#pragma unroll
  for (int i = 0; i < kNewtonSteps; ++i) {
    const double he = fma(-(hx * r), r, 0.5); // (1 - x r^2) / 2
    r = fma(r, he, r);
  }
  return r;
}
// Kept for the host-visible unit checks and the bulk solution.
__device__ __forceinline__ double drcp(double x) { return rcpT(x); }
__device__ __forceinline__ float drcp(float x) { return rcpT(x); }

// Principal complex square root (Kahan's formulation): with
// t = sqrt((|x| + |z|) / 2), sqrt(z) = (t, y / 2t) for x >= 0 and
// (|y| / 2t, copysign(t, y)) otherwise.  Both square roots come from
// reciprocal square roots: |z| = n rsq(n) with n = |z|^2, and with
// u = (|x| + |z|) / 2, t = u rsq(u) and y / 2t = y rsq(u) / 2.
// n is held at or above the smallest normal number, so z = 0 gives a root
// of order 1e-77 (1e-10 in FP32) instead of 0 -- negligible beside the O(1)
// terms it is added to -- with no branch and no select: the square roots of
// a lane's energies stay in one basic block and their latencies overlap.
template <typename T> __device__ __forceinline__ Cx<T> csqrt(Cx<T> z) {
  const T n = fmax(z.re * z.re + z.im * z.im, std::numeric_limits<T>::min());
  const T m = n * rsqT(n);
  const T u = T(0.5) * (fabs(z.re) + m);
  const T r = rsqT(u);
  const T t = u * r;
  const T q = (T(0.5) * z.im) * r;
  const bool right = z.re >= T(0);
  return {right ? t : fabs(q), right ? q : copysign(t, z.im)};
}

// a / b through the reciprocal of |b|^2 (|b| is of order one here).
template <typename T> __device__ __forceinline__ Cx<T> cdiv(Cx<T> a, Cx<T> b) {
  const T r = rcpT(b.re * b.re + b.im * b.im);
  return {(a.re * b.re + a.im * b.im) * r, (a.im * b.re - a.re * b.im) * r};
}

// One implicit-midpoint step of the Riccati equation solved as the quadratic
// a y^2 + b y + c = 0 (RiccatiSolutions.h:65-94):
//   s = i h, a = s conj(D) / 4, b = 2 a y0 + s e - 1,
//   c = (a y0 + s e + 1) y0 + s D,  y = 2 c / (-b + sqrt(b^2 - 4 a c)).
// The quantities that only depend on the site (D, the energy shift Re e and
// h) are formed once per site and shared by every Ozaki energy.
template <typename T> struct StepSite {
  Cx<T> a;  // s conj(D) / 4
  Cx<T> a4; // 4 a
  Cx<T> sD; // s D = (-4 a.re, 4 a.im)
  T sei;    // Im(s e) = h Re e
};
template <typename T>
__device__ __forceinline__ StepSite<T> stepSite(Cx<T> D, T shift, T h) {
  StepSite<T> p;
  p.a4 = {h * D.im, h * D.re};
  p.a = {T(0.25) * p.a4.re, T(0.25) * p.a4.im};
  p.sD = {-p.a4.re, p.a4.im};
  p.sei = h * shift;
  return p;
}
// ser = Re(s e) = -h Im e (the Matsubara part).
template <typename T>
__device__ __forceinline__ Cx<T> stepEnergy(const StepSite<T> &p, T ser, Cx<T> y0) {
  const Cx<T> ay = p.a * y0;
  const Cx<T> b = {T(2) * ay.re + (ser - T(1)), T(2) * ay.im + p.sei};
  const Cx<T> t = {ay.re + (ser + T(1)), ay.im + p.sei};
  const Cx<T> c = t * y0 + p.sD;
  const Cx<T> ac4 = p.a4 * c;
  const Cx<T> disc = {b.re * b.re - b.im * b.im - ac4.re, T(2) * b.re * b.im - ac4.im};
  const Cx<T> sq = csqrt(disc);
  return cdiv(c + c, sq - b);
}
// Workgroups are dispatched round robin over the eight XCDs, each with its
// own L2.  A grid of 8 * ceil(n / 8) blocks read through this mapping gives
// every XCD a contiguous range of the n logical blocks (spatially coherent
// sites or tiles share their L2 lines); -1 for the padding blocks.
constexpr int kNumXcd = 8;
__device__ __forceinline__ long long xcdBlock(long long nLogical) {
  const long long per = (nLogical + kNumXcd - 1) / kNumXcd;
  const long long b = blockIdx.x;
  const long long l = (b % kNumXcd) * per + b / kNumXcd;
  return l < nLogical ? l : -1;
}
inline unsigned xcdGrid(long long nLogical) {
  return static_cast<unsigned>(((nLogical + kNumXcd - 1) / kNumXcd) * kNumXcd);
}

// Real-frequency energies eps = E + i delta: sej = h E joins the site's
// Im(s e) (the Doppler shift enters the same way).
template <typename T>
__device__ __forceinline__ Cx<T> stepEnergy(const StepSite<T> &p, T ser, T sej, Cx<T> y0) {
  const T sei = p.sei + sej;
  const Cx<T> ay = p.a * y0;
  const Cx<T> b = {T(2) * ay.re + (ser - T(1)), T(2) * ay.im + sei};
  const Cx<T> t = {ay.re + (ser + T(1)), ay.im + sei};
  const Cx<T> c = t * y0 + p.sD;
  const Cx<T> ac4 = p.a4 * c;
  const Cx<T> disc = {b.re * b.re - b.im * b.im - ac4.re, T(2) * b.re * b.im - ac4.im};
  const Cx<T> sq = csqrt(disc);
  return cdiv(c + c, sq - b);
}
template <typename T>
__device__ __forceinline__ Cx<T> riccatiStep(Cx<T> D, Cx<T> e, Cx<T> y0, T h) {
  return stepEnergy(stepSite(D, e.re, h), -h * e.im, y0);
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
__device__ __forceinline__ double bilinear(const double *__restrict__ grid, int N,
                                           const RotSample &r) {
#pragma clang fp contract(off)
  const GlobalPtr<const double> p = global(grid);
  const double a = (1.0 - r.fx) * p[r.idx] + r.fx * p[r.idx + 1];
  const double b = (1.0 - r.fx) * p[r.idx + N] + r.fx * p[r.idx + N + 1];
  return (1.0 - r.fy) * a + r.fy * b;
}

} // namespace scamd
