// sc_adjoint.hip -- the gradient of the free energy with respect to the
// order parameter by a discrete adjoint (reverse mode) of one iteration.
// There is no counterpart in the reference: it serves gradient studies and
// torch.autograd in the Python frontend.
//
// F(Delta) is the free energy one iteration reports (ComputeFreeEnergy.h:
// 281-414 on the Green functions of the sweep) with the vector potential and
// the boundary in-values held fixed:
//   F = h^2 [ T sum_{inside g} fe(g) + sum_c |Delta_c|^2 (ln(T / Tc_c) + S)
//             + magnetic term ],
//   fe(g) = sum_k w_k bilinear_k(extend_k(gF_k))(g),
//   gF_k(s) = sum_e r_e [D.re (gt.im - g.im) + D.im (g.re + gt.re)]
// with g, gt the coherence functions of frame k at site s and D the rotated,
// extended order parameter sum_c b_c(k) Delta_c.  Reverse mode runs the
// stages backwards:
//   1. the adjoint seed rho_k(s) = dF/dgF_k(s): the transpose of the
//      back-rotation (bilinear taps of every inside site) and of the
//      extension of the frame planes;
//   2. per trajectory, the reverse of both Riccati recurrences: pass 1
//      integrates gamma~ and keeps it, pass 2 integrates gamma, keeps it and
//      carries the adjoint of gamma~ upwards, pass 3 carries the adjoint of
//      gamma downwards; each step's adjoint adds to dF/dD at its sites;
//   3. the transpose of the rotation and extension of the order parameter
//      and the basis functions, plus the explicit |Delta|^2 terms.
// With a penetration depth the induced vector potential A enters too: the
// Doppler shift es_k(s) = q v_k . (A(rotated s) + A_ext) is a real shift of
// every energy of every step, so the reverse sweep also returns dF/des per
// frame site, whose transpose through the bilinear taps of A gives dF/dA;
// the magnetic term (1 / 2 lambda^2) h^2 sum (curl lambda^2 A)^2 adds its
// own.  dF/dA is per unit of A as the setters take it (the state stores
// lambda^2 A).
// Complex adjoints follow zbar = dF/dRe z + i dF/dIm z; for y = f(x) with f
// holomorphic, xbar = conj(f'(x)) ybar.
#include "sc_engine.hpp"
#include "sc_device.hpp"
#include "sc_util.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

namespace scamd {

namespace {

constexpr int kAdjEL = kAdjointEL; // Ozaki energies per lane of the reverse sweep
constexpr int kAdjSlotsPerCU = 8; // persistent wavefronts per compute unit (2 per SIMD)

__device__ __forceinline__ Cx<double> conjc(Cx<double> a) { return {a.re, -a.im}; }

// The adjoints of one frame site's fields: dF/dD (re, im) and dF/des.
struct SiteBar {
  double re, im, es;
};
__device__ __forceinline__ void addBar(SiteBar &b, Cx<double> d, double e) {
  b.re += d.re;
  b.im += d.im;
  b.es += e;
}

// Adjoint of stepEnergy(p, ser, y0) = y.  y is the root of
//   F(y) = a (y + y0)^2 + (se - 1) y + (se + 1) y0 + sD = 0,
// a = i hs conj(D) / 4, sD = i hs D, se = ser + i sei (the StepSite of D
// and the shift es over the step length hs, sei = hs es), so by implicit
// differentiation
//   dy = q (F_y0 dy0 + (y + y0)^2 da + d(sD) + i (y + y0) dsei),
//   q = -1 / F_y,  F_y = 2 a (y + y0) + se - 1,  F_y0 = 2 a (y + y0) + se + 1.
// Returns y0bar; adds dF/dD of the step's D argument to dbar and dF/des of
// its shift to ebar.
__device__ __forceinline__ Cx<double> stepAdjoint(const StepSite<double> &p, double ser,
                                                  Cx<double> y0, Cx<double> y, Cx<double> ybar,
                                                  double hs, Cx<double> &dbar, double &ebar) {
  const Cx<double> w = y + y0;
  const Cx<double> aw = p.a * w;
  const Cx<double> fy = {2.0 * aw.re + ser - 1.0, 2.0 * aw.im + p.sei};
  const Cx<double> fy0 = {2.0 * aw.re + ser + 1.0, 2.0 * aw.im + p.sei};
  const double r = rcpT(fy.re * fy.re + fy.im * fy.im);
  const Cx<double> q = {-fy.re * r, fy.im * r};
  const Cx<double> abar = conjc(q * (w * w)) * ybar;
  const Cx<double> sdbar = conjc(q) * ybar;
  // a = (i hs / 4) conj(D): Dbar += conj(abar) i hs / 4;  sD = i hs D: Dbar += -i hs sdbar
  dbar.re += 0.25 * hs * abar.im + hs * sdbar.im;
  dbar.im += 0.25 * hs * abar.re - hs * sdbar.re;
  // a real parameter's adjoint is Re(conj(dy/dparam) ybar)
  const Cx<double> qiw = q * Cx<double>{-w.im, w.re};
  ebar += hs * (qiw.re * ybar.re + qiw.im * ybar.im);
  return conjc(q * fy0) * ybar;
}

// Sum over the energy groups of a column (lanes c, c + cols, ...) into the
// group-0 lane (the sweep's reduction).
__device__ __forceinline__ double groupSum(double v, int gp, int cols, int c) {
  if (gp * cols == 64) {
    for (int m = cols; m < 64; m <<= 1) v += __shfl_xor(v, m);
    return v;
  }
  double s = v;
  for (int m = 1; m < gp; ++m) s += __shfl(v, c + m * cols);
  return s;
}

// 1a. Transpose of the back-rotation (gatherKernel's taps): rho_k at the
// four taps of every inside reference site, times scale * w_k.
__global__ void feSeedKernel(GrainView v, const int *sites, int numSites, double scale,
                             double *rho) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= numSites) return;
  const int N0 = v.N0, N1 = v.N1;
  const int site = sites[t];
  const int x = site % N0, y = site / N0;
  const size_t n1 = size_t(N1) * N1;
  for (int k = 0; k < v.Na; ++k) {
    const RotSample r = rotatedCell(x, y, N0, N1, double(v.bkCos[k]), double(v.bkSin[k]));
    if (!r.valid) continue;
    const double w = scale * v.weight[k];
    const double wx0 = 1.0 - r.fx, wx1 = r.fx, wy0 = 1.0 - r.fy, wy1 = r.fy;
    double *p = rho + size_t(k) * n1 + r.idx;
    atomicAdd(p, w * (wy0 * wx0));
    atomicAdd(p + 1, w * (wy0 * wx1));
    atomicAdd(p + N1, w * (wy1 * wx0));
    atomicAdd(p + N1 + 1, w * (wy1 * wx1));
  }
}

// 1b / 3b. Transpose of an extension (target <- mean of its sources): a
// target hands its adjoint to its sources in equal parts and keeps none.
// Targets and sources are indices within a plane; sources are never targets.
__global__ void extendAdjointKernel(double *bar, const int *target, const int *offset,
                                    const int *source, int numTargets, int planes,
                                    long long planeStride) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= numTargets) return;
  const long long tg = target[t];
  const int o0 = offset[t], o1 = offset[t + 1];
  const double np = double(o1 - o0);
  for (int q = 0; q < planes; ++q) {
    double *p = bar + q * planeStride;
    const double v = p[tg];
    if (v == 0.0) continue;
    p[tg] = 0.0;
    const double share = v / np;
    for (int o = o0; o < o1; ++o) atomicAdd(p + source[o], share);
  }
}

// 3a. Transpose of the site fields, as a gather.  The site fields take the
// order parameter D_k(s) = sum_c b_c(k) bilinear(extDelta_c)(rotated s) and,
// with aBar, the Doppler shift es_k(s) = q v_k . bilinear(A)(rotated s)
// (siteFields).  One thread per reference site p collects, frame by frame,
// dF/dD_k(s) w(s, p) and dF/des_k(s) w(s, p) from the frame sites s whose
// bilinear taps include p: those lie within sqrt(2) of p's inverse-rotated
// position, inside a 4 x 4 window, and rotatedCell (the forward's own
// arithmetic) decides which of them tap p and with what weight.  At the
// grid's lower edges (p.x or p.y <= 1) rotatedCell's truncation toward zero
// also takes rotated positions in (-1, 0) to cell 0, up to 2 sites from p
// in that coordinate, so the window grows to 7 x 7 there.  Every output is
// written by one thread: no atomics, the same sum on every run.
__device__ __forceinline__ void gatherFrame(const double *const *planes, int nPlanes, int px,
                                            int py, int N0, int N1, double c, double s,
                                            double (&acc)[3]) {
  const double sh = 0.5 * double(N0 - 1), h1 = 0.5 * double(N1 - 1);
  const double Pcx = double(px) - sh, Pcy = double(py) - sh;
  const double xf = Pcx * c + Pcy * s + h1, yf = -Pcx * s + Pcy * c + h1;
  const bool edge = px <= 1 || py <= 1;
  const int lo = edge ? 3 : 1, hi = edge ? 3 : 2;
  const int x0 = max(int(floor(xf)) - lo, 0), x1 = min(int(floor(xf)) + hi, N1 - 1);
  const int y0 = max(int(floor(yf)) - lo, 0), y1 = min(int(floor(yf)) + hi, N1 - 1);
  for (int cy = y0; cy <= y1; ++cy)
    for (int cx = x0; cx <= x1; ++cx) {
      const size_t fi = size_t(cy) * N1 + cx;
      double val[3] = {0.0, 0.0, 0.0};
      bool any = false;
      for (int q = 0; q < nPlanes; ++q) {
        val[q] = planes[q][fi];
        any = any || val[q] != 0.0;
      }
      if (!any) continue;
      const RotSample r = rotatedCell(cx, cy, N1, N0, c, s);
      if (!r.valid) continue;
      const int Yi = r.idx / N0, Xi = r.idx - Yi * N0;
      const int dx = px - Xi, dy = py - Yi;
      if (dx < 0 || dx > 1 || dy < 0 || dy > 1) continue;
      const double w = (dy ? r.fy : 1.0 - r.fy) * (dx ? r.fx : 1.0 - r.fx);
      for (int q = 0; q < nPlanes; ++q) acc[q] += val[q] * w;
    }
}

// The frames are split into kFieldChunks ranges (blockIdx.y), each thread
// summing its range into its own slot of part ([chunk][plane][site], planes
// Re/Im of every component, then A_x and A_y): many more threads to hide
// the loads' latency, and fieldsAdjointSumKernel adds the ranges in order.
constexpr int kFieldChunks = 8;
__global__ void fieldsAdjointKernel(GrainView v, const double *dbar, long long framePlane,
                                    bool withA, double *part) {
  const long long p = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const int N0 = v.N0, N1 = v.N1, Na = v.Na, Nc = v.Nc;
  const size_t n0 = size_t(N0) * N0, n1 = size_t(N1) * N1;
  if (p >= (long long)n0) return;
  const int px = int(p % N0), py = int(p / N0);
  const int chunk = blockIdx.y;
  const int k0 = chunk * Na / kFieldChunks, k1 = (chunk + 1) * Na / kFieldChunks;
  double dr[4] = {0, 0, 0, 0}, di[4] = {0, 0, 0, 0};
  double ax = 0.0, ay = 0.0;
  for (int k = k0; k < k1; ++k) {
    const double *pl[2] = {dbar + size_t(k) * n1, dbar + framePlane + size_t(k) * n1};
    double acc[3] = {0.0, 0.0, 0.0};
    gatherFrame(pl, 2, px, py, N0, N1, double(v.opCos[k]), double(v.opSin[k]), acc);
    if (acc[0] != 0.0 || acc[1] != 0.0) {
      for (int c = 0; c < Nc && c < 4; ++c) {
        const double b = v.basis[c * Na + k];
        dr[c] += b * acc[0];
        di[c] += b * acc[1];
      }
    }
    if (withA) {
      const double *pe[1] = {dbar + 2 * framePlane + size_t(k) * n1};
      double ae[3] = {0.0, 0.0, 0.0};
      gatherFrame(pe, 1, px, py, N0, N1, double(v.aCos[k]), double(v.aSin[k]), ae);
      ax += double(v.charge) * v.vFx[k] * ae[0];
      ay += double(v.charge) * v.vFy[k] * ae[0];
    }
  }
  const int planes = 2 * Nc + 2;
  double *out = part + size_t(chunk) * planes * n0 + p;
  for (int c = 0; c < Nc && c < 4; ++c) {
    out[size_t(2 * c) * n0] = dr[c];
    out[size_t(2 * c + 1) * n0] = di[c];
  }
  out[size_t(2 * Nc) * n0] = ax;
  out[size_t(2 * Nc + 1) * n0] = ay;
}

// The frame ranges added in order: extBar (Re/Im of every component, the
// extension's transpose follows) and, with aBar, dF/dA's Doppler part.
__global__ void fieldsAdjointSumKernel(const double *part, int Nc, long long n0, double *extBar,
                                       double *aBar) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const int planes = 2 * Nc + 2;
  if (i >= planes * n0) return;
  const int q = int(i / n0);
  if (q >= 2 * Nc && !aBar) return;
  double sum = 0.0;
  for (int ch = 0; ch < kFieldChunks; ++ch) sum += part[size_t(ch) * planes * n0 + i];
  if (q < 2 * Nc)
    extBar[i] = sum;
  else
    aBar[i - 2 * Nc * n0] = sum;
}

// 3c. The magnetic term's gradient: with b_t the curl of the stored
// lambda^2 A at site t (curlSquareKernel's stencil), d/dA of
// (area / 2 lambda^2) sum_t b_t^2 is area sum_t b_t db_t/d(lambda^2 A) per
// unit of A.  One thread per site, its four stencil entries by atomics.
__global__ void curlAdjointKernel(const double *A, int N, double invH, double area,
                                  double *aBar) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N * N) return;
  const int x = t % N, y = t / N;
  const size_t n = size_t(N) * N;
  const int xn = min(x + 1, N - 1), xp = max(x - 1, 0);
  const int yn = min(y + 1, N - 1), yp = max(y - 1, 0);
  const double cx = invH / double(xn - xp), cy = invH / double(yn - yp);
  const double b = (A[n + size_t(y) * N + xn] - A[n + size_t(y) * N + xp]) * cx -
                   (A[size_t(yn) * N + x] - A[size_t(yp) * N + x]) * cy;
  const double ab = area * b;
  atomicAdd(aBar + n + size_t(y) * N + xn, ab * cx);
  atomicAdd(aBar + n + size_t(y) * N + xp, -ab * cx);
  atomicAdd(aBar + size_t(yn) * N + x, -ab * cy);
  atomicAdd(aBar + size_t(yp) * N + x, ab * cy);
}

// The explicit terms: grad = extBar + 2 h^2 factor_c Delta_c at inside sites.
struct ComponentFactors {
  double f[4];
};
__global__ void finishGradientKernel(const double *extBar, const double *delta,
                                     const uint8_t *domain, ComponentFactors fac, double area,
                                     long long n0, int Nc, double *grad) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= 2LL * Nc * n0) return;
  const int c = int(i / (2 * n0));
  const long long site = i % n0;
  const double e = domain[site] ? 2.0 * area * fac.f[c] * delta[i] : 0.0;
  grad[i] = extBar[i] + e;
}

struct AdjointArgs {
  GrainView view;
  const SweepItemDev *items;
  int numItems;
  int *counter;
  Cx<double> *scratch;    // per slot: gamma~ then gamma, [maxLen][EL][64] each
  long long slotStride;   // complex elements of one of the two buffers
  const double *rho;      // [Na][N1^2]
  const double *seed;     // residual mode: [3][Na][N1^2] adjoints of the Green
                          // sums per frame site (gap equation re, im; current)
  double *dbar;           // [3][Na][N1^2]: dF/dD re, im, dF/des
  long long framePlane;   // Na * N1^2
  double *partial;        // per item: its share of the free-energy sum
  int ck;                 // checkpoint segment length (0: every site kept)
  int ckSeg;              // segments per scratch buffer (ceil(maxLen / ck))
};

// Local adjoints of gamma and gamma~ at one site and energy through the
// Green sums of the gap equation and the current (the sweep's Green terms,
// LocalGreens.h:32-70): with p = g gt and Q = 2i / (1 + p)^2,
//   gD = -2i g / (1 + p) + conj(2i gt / (1 + p)),  Re gJ, gJ = -i (1 - p) / (1 + p),
//   dgD/dg = -Q and, conjugated, -Q gt^2;  dgD/dgt = Q g^2 and, conjugated, Q;
//   dgJ/dg = Q gt,  dgJ/dgt = Q g,
// seeded with db (the gap sum's adjoint, complex) and jb (the current's,
// real), times the residue r.
__device__ __forceinline__ void greensAdjoint(Cx<double> g, Cx<double> gt, Cx<double> db, double jb,
                                              double r, Cx<double> &gbar, Cx<double> &gtbar) {
  const Cx<double> z = {1.0 + (g.re * gt.re - g.im * gt.im), g.re * gt.im + g.im * gt.re};
  const Cx<double> z2 = z * z;
  const double rn = rcpT(z2.re * z2.re + z2.im * z2.im);
  // Q = 2i conj(z2) / |z2|^2
  const Cx<double> Q = {2.0 * z2.im * rn, 2.0 * z2.re * rn};
  const Cx<double> cQ = conjc(Q);
  const Cx<double> Qgt = Q * gt, Qg = Q * g;
  const Cx<double> jbc = {jb, 0.0};
  gbar = Cx<double>{-r, 0.0} * (cQ * db + conjc(Qgt * gt * db)) + Cx<double>{r, 0.0} * (conjc(Qgt) * jbc);
  gtbar = Cx<double>{r, 0.0} * (conjc(Qg * g) * db + conjc(Q * db) + conjc(Qg) * jbc);
}

// 2. The reverse sweep of one item (the sweep's items and lane layout).
// RES: the residual mode, seeded from a.seed through greensAdjoint instead
// of the free-energy density; pass 2 leaves the local adjoint of gamma in
// the gamma~ scratch (read there already) for pass 3.
template <int EL, bool RES>
__device__ void adjointItem(const AdjointArgs &a, const SweepItemDev &it, int itemIndex,
                            Cx<double> *scrT, Cx<double> *scrG) {
  const GrainView &v = a.view;
  const int lane = threadIdx.x;
  const int cols = v.cols, gp = v.groups;
  const int g = lane / cols;
  const int c = lane - g * cols;
  const int N1 = v.N1, Ne = v.Ne;
  const size_t n1 = size_t(N1) * N1;
  const double h = v.h;
  const int k = it.angle;
  const bool colActive = c < it.ncols;
  const int x = it.x0 + (colActive ? c : it.ncols - 1);
  const bool active = colActive && g < gp;
  const int8_t *lab = v.labels + size_t(k) * n1 + x;
  const double *F = v.fields + size_t(k - v.fieldBase) * n1 + x;
  const double *rhoK = a.rho + size_t(k) * n1 + x;
  const double *seedK = RES ? a.seed + size_t(k) * n1 + x : nullptr;
  double *dbK = a.dbar + size_t(k) * n1 + x;
  double ser[EL], res[EL];
  bool eOk[EL];
#pragma unroll
  for (int j = 0; j < EL; ++j) {
    const int e = g * EL + j;
    eOk[j] = active && e < Ne;
    ser[j] = eOk[j] ? -h * v.energy[e] : 0.0;
    res[j] = eOk[j] ? v.residue[e] : 0.0;
  }
  const int len = it.yhi - it.ylo + 1;
  const size_t col = size_t(k) * N1 + x;
  auto slot = [&](int t, int j) { return (size_t(t) * EL + j) * 64 + lane; };
  // dF/dD and dF/des of a finished site: the column's groups summed,
  // written (or added) by the group-0 lane
  auto writeDbar = [&](int y, SiteBar part, bool add) {
    const double re = groupSum(part.re, gp, cols, c);
    const double im = groupSum(part.im, gp, cols, c);
    const double es = groupSum(part.es, gp, cols, c);
    if (g == 0 && colActive) {
      double *p = dbK + size_t(y) * N1;
      if (add) {
        p[0] += re;
        p[a.framePlane] += im;
        p[2 * a.framePlane] += es;
      } else {
        p[0] = re;
        p[a.framePlane] = im;
        p[2 * a.framePlane] = es;
      }
    }
  };

  // With checkpoints (a.ck = K > 0) each scratch buffer holds, per segment
  // of K sites, the state before the segment's first site (EL values and
  // the entry count, per lane), then room for one segment's values: pass 1
  // keeps gamma~ only at the segments' starts and pass 2 integrates each
  // segment again from there when it reaches it; pass 2 leaves gamma's
  // state at the bottom of every segment and pass 3 integrates each segment
  // again.  The same arithmetic as keeping every site.
  const int K = a.ck;
  auto ckSlot = [&](int sg, int j) { return (size_t(sg) * (EL + 1) + j) * 64 + lane; };
  auto segSlot = [&](int i, int j) {
    return (size_t(a.ckSeg) * (EL + 1) + size_t(i) * EL + j) * 64 + lane;
  };
  const int baseInT0 = v.idxInGt[col], baseInG0 = v.idxInG[col];
  // one site of gamma~ from the top (the sweep's pass 1): t counts from it.yhi
  auto fwdT = [&](int t, Cx<double>(&coh)[EL], Cx<double> &dPrev, double &esPrev, int &nT) {
    const size_t o = size_t(it.yhi - t) * N1;
    const Cx<double> D = {-F[o], F[v.fieldPlane + o]};
    const double es = F[2 * v.fieldPlane + o];
    const int8_t lc = lab[o];
    if (active) {
      if (lc == 1 || lc == 2) {
        const StepSite<double> ps = stepSite(
            Cx<double>{(dPrev.re + D.re) * 0.5, (dPrev.im + D.im) * 0.5}, (esPrev + es) * 0.5, h);
        const double half = t > 0 ? 1.0 : 0.5;
#pragma unroll
        for (int j = 0; j < EL; ++j) coh[j] = stepEnergy(ps, ser[j] * half, coh[j]);
      } else if (lc == 3) {
        const int ci = baseInT0 + nT;
        const double dr = v.distInGt[ci];
        const StepSite<double> ps = stepSite(D, es, dr * h);
#pragma unroll
        for (int j = 0; j < EL; ++j) {
          const size_t si = size_t(ci) * Ne + g * EL + j;
          const Cx<double> b0 = eOk[j] ? Cx<double>{v.bInGt[si], v.bInGt[size_t(v.nInGt) * Ne + si]}
                                       : Cx<double>{0.0, 0.0};
          coh[j] = stepEnergy(ps, ser[j] * dr, b0);
        }
      } else {
#pragma unroll
        for (int j = 0; j < EL; ++j) coh[j] = {0.0, 0.0};
      }
    }
    if (lc == 3) ++nT;
    dPrev = D;
    esPrev = es;
  };
  // one site of gamma from the bottom (the sweep's pass 2): u counts from it.ylo
  auto fwdG = [&](int u, Cx<double>(&coh)[EL], Cx<double> &dPrev, double &esPrev, int &nG) {
    const size_t o = size_t(it.ylo + u) * N1;
    const Cx<double> D = {F[o], F[v.fieldPlane + o]};
    const double es = F[2 * v.fieldPlane + o];
    const int8_t lc = lab[o];
    if (active) {
      if (lc == 1 || lc == 3) {
        const StepSite<double> ps = stepSite(
            Cx<double>{(dPrev.re + D.re) * 0.5, (dPrev.im + D.im) * 0.5}, (esPrev + es) * 0.5, h);
        const double half = u > 0 ? 1.0 : 0.5;
#pragma unroll
        for (int j = 0; j < EL; ++j) coh[j] = stepEnergy(ps, ser[j] * half, coh[j]);
      } else if (lc == 2) {
        const int ci = baseInG0 + nG;
        const double dr = v.distInG[ci];
        const StepSite<double> ps = stepSite(D, es, dr * h);
#pragma unroll
        for (int j = 0; j < EL; ++j) {
          const size_t si = size_t(ci) * Ne + g * EL + j;
          const Cx<double> b0 = eOk[j] ? Cx<double>{v.bInG[si], v.bInG[size_t(v.nInG) * Ne + si]}
                                       : Cx<double>{0.0, 0.0};
          coh[j] = stepEnergy(ps, ser[j] * dr, b0);
        }
      } else {
#pragma unroll
        for (int j = 0; j < EL; ++j) coh[j] = {0.0, 0.0};
      }
    }
    if (lc == 2) ++nG;
    dPrev = D;
    esPrev = es;
  };
  // gamma~ of segment sg again, from its checkpoint, into the segment room
  auto recomputeT = [&](int sg) {
    const int t0 = sg * K, t1 = min(t0 + K, len);
    Cx<double> cc[EL];
#pragma unroll
    for (int j = 0; j < EL; ++j) cc[j] = active ? scrT[ckSlot(sg, j)] : Cx<double>{0.0, 0.0};
    int nT = active ? int(scrT[ckSlot(sg, EL)].re) : 0;
    Cx<double> dp = {0.0, 0.0};
    double ep = 0.0;
    if (t0 > 0) {
      const size_t o = size_t(it.yhi - (t0 - 1)) * N1;
      dp = {-F[o], F[v.fieldPlane + o]};
      ep = F[2 * v.fieldPlane + o];
    }
    for (int t = t0; t < t1; ++t) {
      fwdT(t, cc, dp, ep, nT);
      if (active) {
#pragma unroll
        for (int j = 0; j < EL; ++j) scrT[segSlot(t - t0, j)] = cc[j];
      }
    }
  };
  // gamma of segment sg again, from the state pass 2 left at its bottom
  auto recomputeG = [&](int sg) {
    const int t0 = sg * K, t1 = min(t0 + K, len) - 1;
    Cx<double> cc[EL];
#pragma unroll
    for (int j = 0; j < EL; ++j) cc[j] = active ? scrG[ckSlot(sg, j)] : Cx<double>{0.0, 0.0};
    int nG = active ? int(scrG[ckSlot(sg, EL)].re) : 0;
    const int u0 = len - 1 - t1;
    Cx<double> dp = {0.0, 0.0};
    double ep = 0.0;
    if (u0 > 0) {
      const size_t o = size_t(it.ylo + u0 - 1) * N1;
      dp = {F[o], F[v.fieldPlane + o]};
      ep = F[2 * v.fieldPlane + o];
    }
    for (int u = u0; u <= len - 1 - t0; ++u) {
      fwdG(u, cc, dp, ep, nG);
      if (active) {
#pragma unroll
        for (int j = 0; j < EL; ++j) scrG[segSlot((len - 1 - u) - t0, j)] = cc[j];
      }
    }
  };

  // ------- pass 1: gamma~ from the top, kept at every site or checkpointed -------
  int numEntryT = 0;
  {
    Cx<double> coh[EL];
#pragma unroll
    for (int j = 0; j < EL; ++j) coh[j] = {0.0, 0.0};
    Cx<double> dPrev = {0.0, 0.0};
    double esPrev = 0.0;
    for (int t = 0; t < len; ++t) {
      if (K > 0 && t % K == 0 && active) {
#pragma unroll
        for (int j = 0; j < EL; ++j) scrT[ckSlot(t / K, j)] = coh[j];
        scrT[ckSlot(t / K, EL)] = Cx<double>{double(numEntryT), 0.0};
      }
      fwdT(t, coh, dPrev, esPrev, numEntryT);
      if (K == 0 && active) {
#pragma unroll
        for (int j = 0; j < EL; ++j) scrT[slot(t, j)] = coh[j];
      }
    }
  }

  // ------- pass 2: gamma from the bottom, kept; the adjoint of gamma~ -------
  // (B = the site below, processed in the previous iteration)
  int numEntryG = 0;
  double omega = 0.0;
  {
    const int baseInG = v.idxInG[col], baseInT = v.idxInGt[col];
    int seenT = 0;
    Cx<double> coh[EL], mu[EL], gtB[EL];
#pragma unroll
    for (int j = 0; j < EL; ++j) coh[j] = mu[j] = gtB[j] = {0.0, 0.0};
    Cx<double> dPrev = {0.0, 0.0}, dtB = {0.0, 0.0};
    SiteBar dbB = {0.0, 0.0, 0.0};
    double esPrev = 0.0;
    bool havePrev = false;
    int8_t labB = 0;
    for (int u = 0; u < len; ++u) {
      const int t = len - 1 - u;
      const int y = it.ylo + u;
      const size_t o = size_t(y) * N1;
      const double dre = F[o], dim = F[v.fieldPlane + o], es = F[2 * v.fieldPlane + o];
      const Cx<double> D = {dre, dim}, Dt = {-dre, dim};
      const int8_t lc = lab[o];
      const double rho = RES ? 0.0 : rhoK[o];
      if (K > 0 && (t == len - 1 || t % K == K - 1)) {
        // the bottom of a segment: gamma's state before it, for pass 3, and
        // gamma~ of the segment again
        if (active) {
#pragma unroll
          for (int j = 0; j < EL; ++j) scrG[ckSlot(t / K, j)] = coh[j];
          scrG[ckSlot(t / K, EL)] = Cx<double>{double(numEntryG), 0.0};
        }
        recomputeT(t / K);
      }
      Cx<double> gt[EL];
#pragma unroll
      for (int j = 0; j < EL; ++j)
        gt[j] = active ? (K > 0 ? scrT[segSlot(t - (t / K) * K, j)] : scrT[slot(t, j)])
                       : Cx<double>{0.0, 0.0};
      SiteBar dbCur = {0.0, 0.0, 0.0};
      if (active) {
        // gamma into this site (the sweep's pass 2)
        if (lc == 1 || lc == 3) {
          const StepSite<double> ps = stepSite(
              Cx<double>{(dPrev.re + D.re) * 0.5, (dPrev.im + D.im) * 0.5}, (esPrev + es) * 0.5, h);
          const double half = havePrev ? 1.0 : 0.5;
#pragma unroll
          for (int j = 0; j < EL; ++j) coh[j] = stepEnergy(ps, ser[j] * half, coh[j]);
        } else if (lc == 2) {
          const int ci = baseInG + numEntryG;
          const double dr = v.distInG[ci];
          const StepSite<double> ps = stepSite(D, es, dr * h);
#pragma unroll
          for (int j = 0; j < EL; ++j) {
            const size_t si = size_t(ci) * Ne + g * EL + j;
            const Cx<double> b0 = eOk[j] ? Cx<double>{v.bInG[si], v.bInG[size_t(v.nInG) * Ne + si]}
                                         : Cx<double>{0.0, 0.0};
            coh[j] = stepEnergy(ps, ser[j] * dr, b0);
          }
        } else {
#pragma unroll
          for (int j = 0; j < EL; ++j) coh[j] = {0.0, 0.0};
        }
        if (K == 0) {
#pragma unroll
          for (int j = 0; j < EL; ++j) scrG[slot(t, j)] = coh[j];
        }
        if constexpr (!RES) {
          // the free-energy density and its own dependence on D
          double su = 0.0, sw = 0.0;
#pragma unroll
          for (int j = 0; j < EL; ++j) {
            if (!eOk[j]) continue;
            su += res[j] * (gt[j].im - coh[j].im);
            sw += res[j] * (coh[j].re + gt[j].re);
          }
          omega += rho * (dre * su + dim * sw);
          dbCur = {rho * su, rho * sw, 0.0};
        }
        // adjoint of gamma~ here: what the step into the site below carries
        // back, plus dF/dgt = rho r (D.im + i D.re)
        Cx<double> carry[EL];
#pragma unroll
        for (int j = 0; j < EL; ++j) carry[j] = {0.0, 0.0};
        if (u > 0 && (labB == 1 || labB == 2)) {
          const StepSite<double> ps = stepSite(
              Cx<double>{(Dt.re + dtB.re) * 0.5, (Dt.im + dtB.im) * 0.5}, (es + esPrev) * 0.5, h);
          Cx<double> db = {0.0, 0.0};
          double eb = 0.0;
#pragma unroll
          for (int j = 0; j < EL; ++j)
            if (eOk[j]) carry[j] = stepAdjoint(ps, ser[j], gt[j], gtB[j], mu[j], h, db, eb);
          // the step's D is -conj(D): dF/dD = -conj(dF/dD~), half to each site
          const Cx<double> dd = {-0.5 * db.re, 0.5 * db.im};
          addBar(dbCur, dd, 0.5 * eb);
          addBar(dbB, dd, 0.5 * eb);
        }
        if constexpr (RES) {
          const Cx<double> sd = {seedK[o], seedK[a.framePlane + o]};
          const double sj = seedK[2 * a.framePlane + o];
#pragma unroll
          for (int j = 0; j < EL; ++j) {
            Cx<double> lg = {0.0, 0.0}, lt = {0.0, 0.0};
            if (eOk[j]) greensAdjoint(coh[j], gt[j], sd, sj, res[j], lg, lt);
            mu[j] = eOk[j] ? carry[j] + lt : Cx<double>{0.0, 0.0};
            if (K == 0) scrT[slot(t, j)] = lg;
          }
        } else {
#pragma unroll
          for (int j = 0; j < EL; ++j)
            mu[j] = eOk[j] ? carry[j] + Cx<double>{rho * res[j] * dim, rho * res[j] * dre}
                           : Cx<double>{0.0, 0.0};
        }
        // gamma~ entered here from its in-value (the entries in reverse order)
        if (lc == 3) {
          const int ci = baseInT + (numEntryT - 1 - seenT);
          const double dr = v.distInGt[ci];
          const StepSite<double> ps = stepSite(Dt, es, dr * h);
          Cx<double> db = {0.0, 0.0};
          double eb = 0.0;
#pragma unroll
          for (int j = 0; j < EL; ++j) {
            if (!eOk[j]) continue;
            const size_t si = size_t(ci) * Ne + g * EL + j;
            const Cx<double> b0 = {v.bInGt[si], v.bInGt[size_t(v.nInGt) * Ne + si]};
            (void)stepAdjoint(ps, ser[j] * dr, b0, gt[j], mu[j], dr * h, db, eb);
          }
          addBar(dbCur, Cx<double>{-db.re, db.im}, eb);
        }
      }
      if (lc == 3) ++seenT;
      if (lc == 2) ++numEntryG;
      // the site below has all its contributions now
      if (u > 0) writeDbar(y - 1, dbB, false);
      dbB = dbCur;
#pragma unroll
      for (int j = 0; j < EL; ++j) gtB[j] = gt[j];
      dtB = Dt;
      labB = lc;
      dPrev = D;
      esPrev = es;
      havePrev = true;
    }
    // the first step of gamma~ (top site, from a zero start at half weight)
    if (active && (labB == 1 || labB == 2)) {
      const StepSite<double> ps =
          stepSite(Cx<double>{dtB.re * 0.5, dtB.im * 0.5}, esPrev * 0.5, h);
      Cx<double> db = {0.0, 0.0};
      double eb = 0.0;
#pragma unroll
      for (int j = 0; j < EL; ++j)
        if (eOk[j]) (void)stepAdjoint(ps, ser[j] * 0.5, Cx<double>{0.0, 0.0}, gtB[j], mu[j], h, db, eb);
      addBar(dbB, Cx<double>{-0.5 * db.re, 0.5 * db.im}, 0.5 * eb);
    }
    writeDbar(it.yhi, dbB, false);
  }

  // ------- pass 3: the adjoint of gamma from the top (A = the site above) -------
  {
    const int baseInG = v.idxInG[col];
    int seenG = 0;
    Cx<double> lam[EL], gA[EL];
#pragma unroll
    for (int j = 0; j < EL; ++j) lam[j] = gA[j] = {0.0, 0.0};
    Cx<double> dA = {0.0, 0.0};
    SiteBar dbA = {0.0, 0.0, 0.0};
    double esA = 0.0;
    int8_t labA = 0;
    for (int t = 0; t < len; ++t) {
      const int y = it.yhi - t;
      const size_t o = size_t(y) * N1;
      const double dre = F[o], dim = F[v.fieldPlane + o], es = F[2 * v.fieldPlane + o];
      const Cx<double> D = {dre, dim};
      const int8_t lc = lab[o];
      const double rho = RES ? 0.0 : rhoK[o];
      if (K > 0 && t % K == 0) {
        // the top of a segment: gamma of the segment again (and, for the
        // residual's local adjoints, gamma~)
        recomputeG(t / K);
        if constexpr (RES) recomputeT(t / K);
      }
      Cx<double> gc[EL];
#pragma unroll
      for (int j = 0; j < EL; ++j)
        gc[j] = active ? (K > 0 ? scrG[segSlot(t - (t / K) * K, j)] : scrG[slot(t, j)])
                       : Cx<double>{0.0, 0.0};
      SiteBar dbCur = {0.0, 0.0, 0.0};
      if (active) {
        Cx<double> carry[EL];
#pragma unroll
        for (int j = 0; j < EL; ++j) carry[j] = {0.0, 0.0};
        // the step of gamma from this site into the site above
        if (t > 0 && (labA == 1 || labA == 3)) {
          const StepSite<double> ps = stepSite(
              Cx<double>{(D.re + dA.re) * 0.5, (D.im + dA.im) * 0.5}, (es + esA) * 0.5, h);
          Cx<double> db = {0.0, 0.0};
          double eb = 0.0;
#pragma unroll
          for (int j = 0; j < EL; ++j)
            if (eOk[j]) carry[j] = stepAdjoint(ps, ser[j], gc[j], gA[j], lam[j], h, db, eb);
          const Cx<double> dd = {0.5 * db.re, 0.5 * db.im};
          addBar(dbCur, dd, 0.5 * eb);
          addBar(dbA, dd, 0.5 * eb);
        }
        if constexpr (RES) {
          // the local adjoint of gamma: left by pass 2 in the gamma~ scratch,
          // or formed here from the segment's gamma~ again
          const Cx<double> sd = {seedK[o], seedK[a.framePlane + o]};
          const double sj = seedK[2 * a.framePlane + o];
#pragma unroll
          for (int j = 0; j < EL; ++j) {
            Cx<double> lg = {0.0, 0.0};
            if (K == 0) {
              lg = scrT[slot(t, j)];
            } else if (eOk[j]) {
              Cx<double> lt;
              greensAdjoint(gc[j], scrT[segSlot(t - (t / K) * K, j)], sd, sj, res[j], lg, lt);
            }
            lam[j] = eOk[j] ? carry[j] + lg : Cx<double>{0.0, 0.0};
          }
        } else {
          // dF/dg = rho r (D.im - i D.re)
#pragma unroll
          for (int j = 0; j < EL; ++j)
            lam[j] = eOk[j] ? carry[j] + Cx<double>{rho * res[j] * dim, -rho * res[j] * dre}
                            : Cx<double>{0.0, 0.0};
        }
        // gamma entered here from its in-value (the entries in reverse order)
        if (lc == 2) {
          const int ci = baseInG + (numEntryG - 1 - seenG);
          const double dr = v.distInG[ci];
          const StepSite<double> ps = stepSite(D, es, dr * h);
          Cx<double> db = {0.0, 0.0};
          double eb = 0.0;
#pragma unroll
          for (int j = 0; j < EL; ++j) {
            if (!eOk[j]) continue;
            const size_t si = size_t(ci) * Ne + g * EL + j;
            const Cx<double> b0 = {v.bInG[si], v.bInG[size_t(v.nInG) * Ne + si]};
            (void)stepAdjoint(ps, ser[j] * dr, b0, gc[j], lam[j], dr * h, db, eb);
          }
          addBar(dbCur, db, eb);
        }
      }
      if (lc == 2) ++seenG;
      // the site above has all its contributions now
      if (t > 0) writeDbar(y + 1, dbA, true);
      dbA = dbCur;
#pragma unroll
      for (int j = 0; j < EL; ++j) gA[j] = gc[j];
      dA = D;
      esA = es;
      labA = lc;
    }
    // the first step of gamma (bottom site, from a zero start at half weight)
    if (active && (labA == 1 || labA == 3)) {
      const StepSite<double> ps = stepSite(Cx<double>{dA.re * 0.5, dA.im * 0.5}, esA * 0.5, h);
      Cx<double> db = {0.0, 0.0};
      double eb = 0.0;
#pragma unroll
      for (int j = 0; j < EL; ++j)
        if (eOk[j]) (void)stepAdjoint(ps, ser[j] * 0.5, Cx<double>{0.0, 0.0}, gA[j], lam[j], h, db, eb);
      addBar(dbA, Cx<double>{0.5 * db.re, 0.5 * db.im}, 0.5 * eb);
    }
    writeDbar(it.ylo, dbA, true);
  }

  // this item's share of the free-energy sum.  The butterfly leaves the
  // same bits in every lane (each level adds the same two operands, only
  // swapped), so every lane stores it: a store behind a lane-0 branch here
  // and the lane-0 fetch of the next item were threaded by the compiler
  // into one path that let lanes 1-63 re-enter the item loop without lane 0.
  for (int m = 1; m < 64; m <<= 1) omega += __shfl_xor(omega, m);
  a.partial[itemIndex] = omega;
}

template <int EL, bool RES>
__global__ __launch_bounds__(64, 2) void adjointKernel(AdjointArgs a) {
  Cx<double> *scrT = a.scratch + size_t(blockIdx.x) * 2 * a.slotStride;
  Cx<double> *scrG = scrT + a.slotStride;
  for (;;) {
    // every lane takes part in the fetch (lane 0 adds one, the others
    // zero), so no branch on the lane index surrounds it; lane 0's old
    // value is the wavefront's item
    const int item =
        __builtin_amdgcn_readfirstlane(atomicAdd(a.counter, threadIdx.x == 0 ? 1 : 0));
    if (item >= a.numItems) break;
    adjointItem<EL, RES>(a, a.items[item], item, scrT, scrG);
  }
}

__global__ void adjointResetCounter(int *c) { *c = 0; }

// ---------------------------------------------------------------------------
// The tangent-linear sweep.  The tangent of one midpoint step y = stepEnergy
// (p, ser, y0): from F(y) = a (y + y0)^2 + (se - 1) y + (se + 1) y0 + sD = 0,
//   dy = q (F_y0 dy0 + (y + y0)^2 da + d(sD) + i (y + y0) dsei),  q = -1 / F_y,
// where (da, d(sD), dsei) is the StepSite pt of the tangent fields over the
// same step length (all three are linear in D and es).
__device__ __forceinline__ Cx<double> stepTangent(const StepSite<double> &p,
                                                  const StepSite<double> &pt, double ser,
                                                  Cx<double> y0, Cx<double> y, Cx<double> dy0) {
  const Cx<double> w = y + y0;
  const Cx<double> aw = p.a * w;
  const Cx<double> fy = {2.0 * aw.re + ser - 1.0, 2.0 * aw.im + p.sei};
  const Cx<double> fy0 = {2.0 * aw.re + ser + 1.0, 2.0 * aw.im + p.sei};
  const double r = rcpT(fy.re * fy.re + fy.im * fy.im);
  const Cx<double> q = {-fy.re * r, fy.im * r};
  const Cx<double> dF = fy0 * dy0 + (w * w) * pt.a + pt.sD + Cx<double>{-w.im * pt.sei, w.re * pt.sei};
  return q * dF;
}

struct TangentArgs {
  GrainView view;
  const double *tfields;   // tangent site fields, laid out like view.fields
  const SweepItemDev *items;
  int numItems;
  int *counter;
  Cx<double> *scratch;     // per slot: gamma~ then its tangent, [maxLen][EL][64] each
  long long slotStride;
  const double *tInG, *tInGt; // tangents of the in-values (null: zero)
  double *tOutG, *tOutGt;     // tangents of the out-values
  double *tG;                 // tangents of the Green sums [Na][N1^2][4]
};

// One item: pass 1 gamma~ and its tangent from the top, kept; pass 2 gamma
// and its tangent from the bottom, with the tangents of the Green sums
// (greensAdjoint's derivatives applied forward) at every inside site.
template <int EL>
__device__ void tangentItem(const TangentArgs &a, const SweepItemDev &it, Cx<double> *scrT,
                            Cx<double> *scrD) {
  const GrainView &v = a.view;
  const int lane = threadIdx.x;
  const int cols = v.cols, gp = v.groups;
  const int g = lane / cols;
  const int c = lane - g * cols;
  const int N1 = v.N1, Ne = v.Ne;
  const size_t n1 = size_t(N1) * N1;
  const double h = v.h;
  const int k = it.angle;
  const bool colActive = c < it.ncols;
  const int x = it.x0 + (colActive ? c : it.ncols - 1);
  const bool active = colActive && g < gp;
  const int8_t *lab = v.labels + size_t(k) * n1 + x;
  const size_t fo = size_t(k - v.fieldBase) * n1 + x;
  const double *F = v.fields + fo;
  const double *T = a.tfields + fo;
  const long long fp = v.fieldPlane;
  double ser[EL], res[EL];
  bool eOk[EL];
#pragma unroll
  for (int j = 0; j < EL; ++j) {
    const int e = g * EL + j;
    eOk[j] = active && e < Ne;
    ser[j] = eOk[j] ? -h * v.energy[e] : 0.0;
    res[j] = eOk[j] ? v.residue[e] : 0.0;
  }
  const int len = it.yhi - it.ylo + 1;
  const size_t col = size_t(k) * N1 + x;
  auto slot = [&](int t, int j) { return (size_t(t) * EL + j) * 64 + lane; };

  // ---------------- pass 1: gamma~ and its tangent from the top ----------------
  {
    const int baseIn = v.idxInGt[col], baseOut = v.idxOutGt[col];
    int cntIn = 0, cntOut = 0;
    Cx<double> coh[EL], dc[EL];
#pragma unroll
    for (int j = 0; j < EL; ++j) coh[j] = dc[j] = {0.0, 0.0};
    Cx<double> dPrev = {0.0, 0.0}, tPrev = {0.0, 0.0};
    double esPrev = 0.0, tesPrev = 0.0;
    bool havePrev = false;
    for (int t = 0; t < len; ++t) {
      const size_t o = size_t(it.yhi - t) * N1;
      const Cx<double> D = {-F[o], F[fp + o]}, tD = {-T[o], T[fp + o]};
      const double es = F[2 * fp + o], tes = T[2 * fp + o];
      const int8_t lc = lab[o];
      if (active) {
        if (lc == 1 || lc == 2) {
          const StepSite<double> ps = stepSite(
              Cx<double>{(dPrev.re + D.re) * 0.5, (dPrev.im + D.im) * 0.5}, (esPrev + es) * 0.5, h);
          const StepSite<double> pt = stepSite(
              Cx<double>{(tPrev.re + tD.re) * 0.5, (tPrev.im + tD.im) * 0.5}, (tesPrev + tes) * 0.5,
              h);
          const double half = havePrev ? 1.0 : 0.5;
#pragma unroll
          for (int j = 0; j < EL; ++j) {
            const Cx<double> y = stepEnergy(ps, ser[j] * half, coh[j]);
            dc[j] = stepTangent(ps, pt, ser[j] * half, coh[j], y, dc[j]);
            coh[j] = y;
          }
          if (lc == 2) {
            const int ci = baseOut + cntOut;
            const double dr = v.distOutGt[ci];
            const StepSite<double> pe = stepSite(D, es, dr * h), pte = stepSite(tD, tes, dr * h);
#pragma unroll
            for (int j = 0; j < EL; ++j) {
              if (!eOk[j]) continue;
              const Cx<double> bv = stepEnergy(pe, ser[j] * dr, coh[j]);
              const Cx<double> tb = stepTangent(pe, pte, ser[j] * dr, coh[j], bv, dc[j]);
              const size_t si = size_t(ci) * Ne + g * EL + j;
              a.tOutGt[si] = tb.re;
              a.tOutGt[size_t(v.nOutGt) * Ne + si] = tb.im;
            }
          }
        } else if (lc == 3) {
          const int ci = baseIn + cntIn;
          const double dr = v.distInGt[ci];
          const StepSite<double> pe = stepSite(D, es, dr * h), pte = stepSite(tD, tes, dr * h);
#pragma unroll
          for (int j = 0; j < EL; ++j) {
            const size_t si = size_t(ci) * Ne + g * EL + j;
            const Cx<double> b0 = eOk[j] ? Cx<double>{v.bInGt[si], v.bInGt[size_t(v.nInGt) * Ne + si]}
                                         : Cx<double>{0.0, 0.0};
            const Cx<double> tb0 =
                eOk[j] && a.tInGt ? Cx<double>{a.tInGt[si], a.tInGt[size_t(v.nInGt) * Ne + si]}
                                  : Cx<double>{0.0, 0.0};
            const Cx<double> y = stepEnergy(pe, ser[j] * dr, b0);
            dc[j] = stepTangent(pe, pte, ser[j] * dr, b0, y, tb0);
            coh[j] = y;
          }
        } else {
#pragma unroll
          for (int j = 0; j < EL; ++j) coh[j] = dc[j] = {0.0, 0.0};
        }
#pragma unroll
        for (int j = 0; j < EL; ++j) {
          scrT[slot(t, j)] = coh[j];
          scrD[slot(t, j)] = dc[j];
        }
      }
      if (lc == 2) ++cntOut;
      if (lc == 3) ++cntIn;
      dPrev = D;
      tPrev = tD;
      esPrev = es;
      tesPrev = tes;
      havePrev = true;
    }
  }

  // ------ pass 2: gamma and its tangent from the bottom, the Green tangents ------
  {
    const int baseIn = v.idxInG[col], baseOut = v.idxOutG[col];
    int cntIn = 0, cntOut = 0;
    Cx<double> coh[EL], dc[EL];
#pragma unroll
    for (int j = 0; j < EL; ++j) coh[j] = dc[j] = {0.0, 0.0};
    Cx<double> dPrev = {0.0, 0.0}, tPrev = {0.0, 0.0};
    double esPrev = 0.0, tesPrev = 0.0;
    bool havePrev = false;
    for (int u = 0; u < len; ++u) {
      const int t = len - 1 - u;
      const int y = it.ylo + u;
      const size_t o = size_t(y) * N1;
      const Cx<double> D = {F[o], F[fp + o]}, tD = {T[o], T[fp + o]};
      const double es = F[2 * fp + o], tes = T[2 * fp + o];
      const int8_t lc = lab[o];
      double sgr = 0.0, sgi = 0.0, sgj = 0.0;
      if (active) {
        if (lc == 1 || lc == 3) {
          const StepSite<double> ps = stepSite(
              Cx<double>{(dPrev.re + D.re) * 0.5, (dPrev.im + D.im) * 0.5}, (esPrev + es) * 0.5, h);
          const StepSite<double> pt = stepSite(
              Cx<double>{(tPrev.re + tD.re) * 0.5, (tPrev.im + tD.im) * 0.5}, (tesPrev + tes) * 0.5,
              h);
          const double half = havePrev ? 1.0 : 0.5;
#pragma unroll
          for (int j = 0; j < EL; ++j) {
            const Cx<double> yv = stepEnergy(ps, ser[j] * half, coh[j]);
            dc[j] = stepTangent(ps, pt, ser[j] * half, coh[j], yv, dc[j]);
            coh[j] = yv;
          }
          if (lc == 3) {
            const int ci = baseOut + cntOut;
            const double dr = v.distOutG[ci];
            const StepSite<double> pe = stepSite(D, es, dr * h), pte = stepSite(tD, tes, dr * h);
#pragma unroll
            for (int j = 0; j < EL; ++j) {
              if (!eOk[j]) continue;
              const Cx<double> bv = stepEnergy(pe, ser[j] * dr, coh[j]);
              const Cx<double> tb = stepTangent(pe, pte, ser[j] * dr, coh[j], bv, dc[j]);
              const size_t si = size_t(ci) * Ne + g * EL + j;
              a.tOutG[si] = tb.re;
              a.tOutG[size_t(v.nOutG) * Ne + si] = tb.im;
            }
          }
        } else if (lc == 2) {
          const int ci = baseIn + cntIn;
          const double dr = v.distInG[ci];
          const StepSite<double> pe = stepSite(D, es, dr * h), pte = stepSite(tD, tes, dr * h);
#pragma unroll
          for (int j = 0; j < EL; ++j) {
            const size_t si = size_t(ci) * Ne + g * EL + j;
            const Cx<double> b0 = eOk[j] ? Cx<double>{v.bInG[si], v.bInG[size_t(v.nInG) * Ne + si]}
                                         : Cx<double>{0.0, 0.0};
            const Cx<double> tb0 =
                eOk[j] && a.tInG ? Cx<double>{a.tInG[si], a.tInG[size_t(v.nInG) * Ne + si]}
                                 : Cx<double>{0.0, 0.0};
            const Cx<double> yv = stepEnergy(pe, ser[j] * dr, b0);
            dc[j] = stepTangent(pe, pte, ser[j] * dr, b0, yv, tb0);
            coh[j] = yv;
          }
        } else {
#pragma unroll
          for (int j = 0; j < EL; ++j) coh[j] = dc[j] = {0.0, 0.0};
        }
        if (lc > 0) {
          // tangents of the Green sums: with Q = 2i / (1 + g gt)^2,
          // dgD = -Q dg + Q g^2 dgt + conj(-Q gt^2 dg + Q dgt),
          // dgJ = Re(Q (gt dg + g dgt))
#pragma unroll
          for (int j = 0; j < EL; ++j) {
            if (!eOk[j]) continue;
            const Cx<double> gt = scrT[slot(t, j)], tgt = scrD[slot(t, j)];
            const Cx<double> ga = coh[j], tga = dc[j];
            const Cx<double> z = Cx<double>{1.0, 0.0} + ga * gt;
            const Cx<double> z2 = z * z;
            const double rn = rcpT(z2.re * z2.re + z2.im * z2.im);
            const Cx<double> Q = {2.0 * z2.im * rn, 2.0 * z2.re * rn};
            const Cx<double> du = Q * (ga * ga * tgt - tga);
            const Cx<double> dv = Q * (tgt - gt * gt * tga);
            const Cx<double> dj = Q * (gt * tga + ga * tgt);
            sgr += res[j] * (du.re + dv.re);
            sgi += res[j] * (du.im - dv.im);
            sgj += res[j] * dj.re;
          }
        }
      }
      sgr = groupSum(sgr, gp, cols, c);
      sgi = groupSum(sgi, gp, cols, c);
      sgj = groupSum(sgj, gp, cols, c);
      if (g == 0 && colActive && lc > 0) {
        double *Gk = a.tG + (size_t(k) * n1 + o + x) * 4;
        Gk[0] = sgr;
        Gk[1] = sgi;
        Gk[2] = sgj;
        Gk[3] = 0.0;
      }
      if (lc == 3) ++cntOut;
      if (lc == 2) ++cntIn;
      dPrev = D;
      tPrev = tD;
      esPrev = es;
      tesPrev = tes;
      havePrev = true;
    }
  }
}

template <int EL>
__global__ __launch_bounds__(64, 2) void tangentKernel(TangentArgs a) {
  Cx<double> *scrT = a.scratch + size_t(blockIdx.x) * 2 * a.slotStride;
  Cx<double> *scrD = scrT + a.slotStride;
  for (;;) {
    // the branch-free fetch of adjointKernel
    const int item =
        __builtin_amdgcn_readfirstlane(atomicAdd(a.counter, threadIdx.x == 0 ? 1 : 0));
    if (item >= a.numItems) break;
    tangentItem<EL>(a, a.items[item], scrT, scrD);
  }
}

// Residual mode, 1. The residual r = G(u) - u per unit of the setters'
// variables: the order parameter at inside sites (G's is zero outside), the
// vector potential (stored as lambda^2 A) everywhere; and r_A / lambda^2,
// the seed of the Poisson solve's output.
__global__ void residualSeedKernel(const double *delta, const double *deltaNew,
                                   const uint8_t *domain, const double *A, const double *ANew,
                                   long long n0, int Nc, double invL2, double *rD, double *rA,
                                   double *poisSeed) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n0) return;
  const bool in = domain[i] != 0;
  for (int q = 0; q < 2 * Nc; ++q)
    rD[q * n0 + i] = in ? deltaNew[q * n0 + i] - delta[q * n0 + i] : 0.0;
  if (rA) {
    for (int q = 0; q < 2; ++q) {
      const double r = (ANew[q * n0 + i] - A[q * n0 + i]) * invL2;
      rA[q * n0 + i] = r;
      poisSeed[q * n0 + i] = r * invL2;
    }
  }
}

// 2. Transpose of the grain-mean subtraction and of the current's mask and
// factor: jbar = (sbar - sum(sbar) / Ng) jf at inside sites.
__global__ void currentBarKernel(const double *sbar, const double *mean, const uint8_t *domain,
                                 long long n0, double jf, double *jbar) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= 2 * n0) return;
  const long long site = i % n0;
  jbar[i] = domain[site] ? (sbar[i] - mean[i / n0]) * jf : 0.0;
}

// The sum of each of the two planes of x (block b: plane b) times scale.
__global__ void planeMeanKernel(const double *x, long long n, double scale, double *out) {
  const double *p = x + size_t(blockIdx.x) * n;
  double s = 0.0;
  for (long long i = threadIdx.x; i < n; i += blockDim.x) s += p[i];
  __shared__ double sh[256];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int m = blockDim.x / 2; m > 0; m >>= 1) {
    if (threadIdx.x < m) sh[threadIdx.x] += sh[threadIdx.x + m];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = sh[0] * scale;
}

// 3. Transpose of the back-rotation and angle integration (gatherAngle and
// gatherStore): every inside site scatters, per frame, its gap-equation
// adjoint w_k sum_c coupling_c b_c(k) rD_c and its current adjoint
// w_k v_k . jbar to the four taps of the frame's Green planes
// seed[0..2] (gap re, gap im, current).
__global__ void gatherAdjointKernel(GrainView v, const int *sites, int numSites,
                                    const double *coupling, const double *rD,
                                    const double *jbar, long long framePlane, double *seed) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= numSites) return;
  const int N0 = v.N0, N1 = v.N1, Na = v.Na, Nc = v.Nc;
  const size_t n0 = size_t(N0) * N0, n1 = size_t(N1) * N1;
  const int site = sites[t];
  const int x = site % N0, y = site / N0;
  double dr[4] = {0, 0, 0, 0}, di[4] = {0, 0, 0, 0};
  for (int c = 0; c < Nc && c < 4; ++c) {
    dr[c] = coupling[c] * rD[size_t(c) * 2 * n0 + site];
    di[c] = coupling[c] * rD[size_t(c) * 2 * n0 + n0 + site];
  }
  const double jx = jbar ? jbar[site] : 0.0, jy = jbar ? jbar[n0 + site] : 0.0;
  for (int k = 0; k < Na; ++k) {
    const RotSample r = rotatedCell(x, y, N0, N1, double(v.bkCos[k]), double(v.bkSin[k]));
    if (!r.valid) continue;
    const double w = v.weight[k];
    double sr = 0.0, si = 0.0;
    for (int c = 0; c < Nc && c < 4; ++c) {
      const double b = v.basis[c * Na + k];
      sr += dr[c] * b;
      si += di[c] * b;
    }
    const double vals[3] = {sr * w, si * w, (v.vFx[k] * jx + v.vFy[k] * jy) * w};
    const double w00 = (1.0 - r.fy) * (1.0 - r.fx), w10 = (1.0 - r.fy) * r.fx;
    const double w01 = r.fy * (1.0 - r.fx), w11 = r.fy * r.fx;
    for (int q = 0; q < 3; ++q) {
      if (vals[q] == 0.0) continue;
      double *p = seed + q * framePlane + size_t(k) * n1 + r.idx;
      atomicAdd(p, vals[q] * w00);
      atomicAdd(p + 1, vals[q] * w10);
      atomicAdd(p + N1, vals[q] * w01);
      atomicAdd(p + N1 + 1, vals[q] * w11);
    }
  }
}

// 4. The identity part of (J - I)^T r.
__global__ void residualFinishKernel(const double *extBar, const double *rD, long long n,
                                     double *grad, double *gradA, const double *rA, long long nA) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) grad[i] = extBar[i] - rD[i];
  if (gradA && i < nA) gradA[i] -= rA[i];
}

} // namespace

void tangentSweep(const GrainView &view, const double *tfields, const void *items, int numItems,
                  int *counter, double *scratch, long long slotStride, const double *tInG,
                  const double *tInGt, double *tOutG, double *tOutGt, double *tG, int slots,
                  hipStream_t stream) {
  TangentArgs a;
  a.view = view;
  a.tfields = tfields;
  a.items = reinterpret_cast<const SweepItemDev *>(items);
  a.numItems = numItems;
  a.counter = counter;
  a.scratch = reinterpret_cast<Cx<double> *>(scratch);
  a.slotStride = slotStride;
  a.tInG = tInG;
  a.tInGt = tInGt;
  a.tOutG = tOutG;
  a.tOutGt = tOutGt;
  a.tG = tG;
  adjointResetCounter<<<1, 1, 0, stream>>>(counter);
  SC_HIP(hipGetLastError());
  if (numItems > 0) {
    tangentKernel<kAdjEL><<<slots, 64, 0, stream>>>(a);
    SC_HIP(hipGetLastError());
  }
}

// The reverse sweep's items (the sweep's own, at kAdjEL energies per lane),
// its slots' scratch and the adjoint planes, once per grain.
void Grain::prepareAdjoint() {
  const Setup &s = *s_;
  const size_t n0 = size_t(s.N0) * s.N0, n1 = size_t(s.N1) * s.N1;
  const long long framePlane = static_cast<long long>(s.Na) * static_cast<long long>(n1);
  if (!adjItems_) {
    const GrainView gv = sweepView(kAdjEL);
    std::vector<SweepItem> items = sweepItems(s, gv.cols, &adjMaxLen_);
    std::vector<SweepItemDev> di(items.size());
    for (size_t i = 0; i < items.size(); ++i)
      di[i] = {0, items[i].angle, items[i].x0, items[i].ncols, items[i].ylo, items[i].yhi};
    adjNumItems_ = static_cast<int>(di.size());
    SC_HIP(hipMalloc(&adjItems_, std::max<size_t>(1, di.size()) * sizeof(SweepItemDev)));
    if (!di.empty()) toDevice(adjItems_, di.data(), di.size() * sizeof(SweepItemDev));
    int dev = 0, cus = 0;
    SC_HIP(hipGetDevice(&dev));
    SC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    adjSlots_ = std::max(1, std::min(adjNumItems_, kAdjSlotsPerCU * std::max(cus, 1)));
    dAdjPartial_ = devAlloc<double>(size_t(std::max(adjNumItems_, 1)));
    dAdjRho_ = devAlloc<double>(size_t(framePlane));
    dAdjDbar_ = devAlloc<double>(3 * size_t(framePlane));
    dAdjExt_ = devAlloc<double>(size_t(2 * s.Nc) * n0);
    dAdjGrad_ = devAlloc<double>(size_t(2 * s.Nc) * n0);
    dAdjGradA_ = devAlloc<double>(2 * n0);
    dAdjPart_ = devAlloc<double>(size_t(kFieldChunks) * size_t(2 * s.Nc + 2) * n0);
  }
  // the slots' scratch for the setting of the checkpoints: every site
  // (two buffers of maxLen sites), or segments of K sites (two buffers of
  // ceil(maxLen / K) checkpoints and one segment each)
  int K = 0;
  if (adjCk_ > 0) K = std::min(adjCk_, std::max(adjMaxLen_, 1));
  if (adjCk_ < 0) K = std::max(1, int(std::ceil(std::sqrt(double(std::max(adjMaxLen_, 1))))));
  if (!dAdjScratch_ || K != adjCkUsed_) {
    devFree(dAdjScratch_);
    adjCkUsed_ = K;
    adjCkSeg_ = K > 0 ? (adjMaxLen_ + K - 1) / K : 0;
    adjSlotStride_ = K > 0 ? (static_cast<long long>(adjCkSeg_) * (kAdjEL + 1) +
                              static_cast<long long>(K) * kAdjEL) * 64
                           : static_cast<long long>(adjMaxLen_) * kAdjEL * 64;
    // complex elements, two buffers per slot
    dAdjScratch_ = devAlloc<double>(size_t(adjSlots_) * 4 * size_t(adjSlotStride_));
  }
}

void Grain::setNewtonOptions(int restarts, double etaMax, bool fallback, bool fixPhase) {
  if (restarts < 1) throw std::runtime_error("restarts must be >= 1");
  if (!(etaMax > 0.0 && etaMax < 1.0)) throw std::runtime_error("eta_max must be in (0, 1)");
  nkRestarts_ = restarts;
  nkEtaMax_ = etaMax;
  nkFallback_ = fallback;
  nkFixPhase_ = fixPhase;
}

void Grain::setAdjointCheckpoints(int k) {
  if (k < -1) throw std::runtime_error("checkpoint segment length must be >= -1");
  adjCk_ = k;
}

double Grain::freeEnergyGradient(double *grad, double *gradA) {
  const Setup &s = *s_;
  if (s.keldysh) throw std::runtime_error("a spectroscopy context has no free-energy gradient");
  if (s.Nc > 4) throw std::runtime_error("at most 4 order-parameter components");
  const size_t n0 = size_t(s.N0) * s.N0, n1 = size_t(s.N1) * s.N1;
  const long long framePlane = static_cast<long long>(s.Na) * static_cast<long long>(n1);
  // the order parameter's extension and the site fields of every frame
  ensureBoundaryInit();
  extendDelta();
  prepareAdjoint();
  // A enters through the Doppler shift only with a penetration depth
  const bool withA = gradA != nullptr && s.lambda > 0.0;
  // (the seed of the free-energy density depends on the grain's geometry,
  // temperature and angle weights only: formed on the first call, kept)
  if (!adjRhoReady_)
    SC_HIP(hipMemsetAsync(dAdjRho_, 0, size_t(framePlane) * sizeof(double), stream_));
  SC_HIP(hipMemsetAsync(dAdjDbar_, 0, 3 * size_t(framePlane) * sizeof(double), stream_));
  if (withA) SC_HIP(hipMemsetAsync(dAdjGradA_, 0, 2 * n0 * sizeof(double), stream_));
  SC_HIP(hipMemsetAsync(dAdjExt_, 0, size_t(2 * s.Nc) * n0 * sizeof(double), stream_));
  const GrainView v = view();
  const double area = s.h * s.h;
  // 1. the adjoint seed of the free-energy density per frame site
  const int ns = static_cast<int>(s.domainSites.size());
  if (!adjRhoReady_) {
    if (ns > 0) {
      feSeedKernel<<<blocksFor(ns), 256, 0, stream_>>>(v, dDomainSites_, ns,
                                                        area * s.temperature, dAdjRho_);
      SC_HIP(hipGetLastError());
    }
    const int nr = static_cast<int>(s.extRot.target.size());
    if (nr > 0) {
      extendAdjointKernel<<<blocksFor(nr), 256, 0, stream_>>>(dAdjRho_, dExtRotT_, dExtRotO_,
                                                              dExtRotS_, nr, 1, 0);
      SC_HIP(hipGetLastError());
    }
    adjRhoReady_ = true;
  }
  // 2. the reverse sweep
  AdjointArgs a;
  a.view = sweepView(kAdjEL);
  a.items = reinterpret_cast<const SweepItemDev *>(adjItems_);
  a.numItems = adjNumItems_;
  a.counter = dCounter_;
  a.scratch = reinterpret_cast<Cx<double> *>(dAdjScratch_);
  a.slotStride = adjSlotStride_;
  a.ck = adjCkUsed_;
  a.ckSeg = adjCkSeg_;
  a.rho = dAdjRho_;
  a.seed = nullptr;
  a.dbar = dAdjDbar_;
  a.framePlane = framePlane;
  a.partial = dAdjPartial_;
  adjointResetCounter<<<1, 1, 0, stream_>>>(dCounter_);
  if (adjNumItems_ > 0) {
    adjointKernel<kAdjEL, false><<<adjSlots_, 64, 0, stream_>>>(a);
    SC_HIP(hipGetLastError());
  }
  // 3. the rotation, basis and extension of the order parameter
  fieldsAdjointKernel<<<dim3(blocksFor(static_cast<long long>(n0)), kFieldChunks), 256, 0,
                        stream_>>>(v, dAdjDbar_, framePlane, withA, dAdjPart_);
  SC_HIP(hipGetLastError());
  fieldsAdjointSumKernel<<<blocksFor(static_cast<long long>(2 * s.Nc + 2) * n0), 256, 0,
                           stream_>>>(dAdjPart_, s.Nc, static_cast<long long>(n0), dAdjExt_,
                                      withA ? dAdjGradA_ : nullptr);
  SC_HIP(hipGetLastError());
  if (withA) {
    curlAdjointKernel<<<blocksFor(n0), 256, 0, stream_>>>(dA_, s.N0, 1.0 / s.h, s.h * s.h,
                                                          dAdjGradA_);
    SC_HIP(hipGetLastError());
  }
  const int ne = static_cast<int>(s.extRef.target.size());
  if (ne > 0) {
    extendAdjointKernel<<<blocksFor(ne), 256, 0, stream_>>>(dAdjExt_, dExtRefT_, dExtRefO_,
                                                            dExtRefS_, ne, 2 * s.Nc,
                                                            static_cast<long long>(n0));
    SC_HIP(hipGetLastError());
  }
  ComponentFactors fac{};
  for (int c = 0; c < s.Nc; ++c)
    fac.f[c] = std::log(s.temperature / s.cfg.tc[c]) + s.sumInverseEnergy;
  const long long ng = 2LL * s.Nc * static_cast<long long>(n0);
  finishGradientKernel<<<blocksFor(ng), 256, 0, stream_>>>(dAdjExt_, dDelta_, dDomain_, fac,
                                                           area, static_cast<long long>(n0), s.Nc,
                                                           dAdjGrad_);
  SC_HIP(hipGetLastError());
  // the free energy: the sweep's share from the items, the explicit terms
  clearReductions();
  const int hSweep = queueSum(dAdjPartial_, nullptr, nullptr, adjNumItems_, 0);
  std::vector<int> hSq;
  int hCurl = -1;
  queueExplicitFreeEnergy(hSq, hCurl);
  fromDevice(grad, dAdjGrad_, size_t(ng) * sizeof(double));
  if (withA)
    fromDevice(gradA, dAdjGradA_, 2 * n0 * sizeof(double));
  else if (gradA)
    std::fill(gradA, gradA + 2 * n0, 0.0);
  flushReductions();
  double nonMag = 0.0;
  for (int c = 0; c < s.Nc; ++c)
    nonMag += (reductionSum(hSq[size_t(2 * c)]) + reductionSum(hSq[size_t(2 * c + 1)])) * fac.f[c];
  double mag = 0.0;
  if (hCurl >= 0) mag = 0.5 / (s.lambda * s.lambda) * reductionSum(hCurl);
  const double sweepPart = reductionSum(hSweep);
  clearReductions();
  return sweepPart + (nonMag + mag) * area;
}

double Grain::residualGradient(double *grad, double *gradA) {
  const Setup &s = *s_;
  if (s.keldysh) throw std::runtime_error("a spectroscopy context has no residual gradient");
  if (s.Nc > 4) throw std::runtime_error("at most 4 order-parameter components");
  const size_t n0 = size_t(s.N0) * s.N0, n1 = size_t(s.N1) * s.N1;
  const long long framePlane = static_cast<long long>(s.Na) * static_cast<long long>(n1);
  const bool doA = s.cfg.computeCurrent && s.lambda > 0.0;
  const bool withA = gradA != nullptr && doA;
  const size_t nd = size_t(2 * s.Nc) * n0;
  prepareAdjoint();
  // the state the forward map overwrites: G's order parameter and vector
  // potential, the current, the energy density and the boundary out-values
  const size_t nNew = nd + 2 * n0;
  const size_t nOutG = size_t(2 * s.nOutG * s.Ne), nOutGt = size_t(2 * s.nOutGt * s.Ne);
  const size_t nSave = nNew + 2 * n0 + n0 + nOutG + nOutGt;
  if (!dResD_) {
    dResD_ = devAlloc<double>(nd);
    dResA_ = devAlloc<double>(2 * n0);
    dResPois_ = devAlloc<double>(2 * n0);
    dResSbar_ = devAlloc<double>(2 * n0);
    dResJbar_ = devAlloc<double>(2 * n0);
    dResZero_ = devAlloc<double>(2);
    dResMean_ = devAlloc<double>(2);
    dResSeed_ = devAlloc<double>(3 * size_t(framePlane));
    dResSave_ = devAlloc<double>(nSave);
    SC_HIP(hipMemsetAsync(dResZero_, 0, 2 * sizeof(double), stream_));
  }
  struct Piece {
    double *p;
    size_t n;
  };
  const Piece pieces[] = {{dDeltaNew_, nNew}, {dCurrent_, 2 * n0}, {dEnergyDens_, n0},
                          {dBOutG_, nOutG}, {dBOutGt_, nOutGt}};
  {
    size_t off = 0;
    for (const Piece &pc : pieces) {
      if (pc.n > 0)
        SC_HIP(hipMemcpyAsync(dResSave_ + off, pc.p, pc.n * sizeof(double),
                              hipMemcpyDeviceToDevice, stream_));
      off += pc.n;
    }
  }
  // 1. G(u): one plain FP64 iteration without the accelerator and without
  // the boundary condition (the in-values stay as they are)
  ensureBoundaryInit();
  extendDelta(true);
  sweep(PREC_FP64, false);
  backRotate();
  if (doA) solveVectorPotential(dCurrent_, dANew_);
  clearReductions();
  const double invL2 = doA ? 1.0 / (s.lambda * s.lambda) : 1.0;
  residualSeedKernel<<<blocksFor(static_cast<long long>(n0)), 256, 0, stream_>>>(
      dDelta_, dDeltaNew_, dDomain_, dA_, dANew_, static_cast<long long>(n0), s.Nc, invL2, dResD_,
      doA ? dResA_ : nullptr, dResPois_);
  SC_HIP(hipGetLastError());
  const GrainView v = view();
  // 2. the current's adjoint: A' = -h^2 K * (j - mean_grain j) with the
  // kernel symmetric under reflection, so its transpose is the same solve
  // with zero means, then the transpose of the mean subtraction
  if (doA) {
    fft_->solve(dResPois_, dResZero_, dResSbar_);
    planeMeanKernel<<<2, 256, 0, stream_>>>(dResSbar_, static_cast<long long>(n0),
                                            1.0 / static_cast<double>(s.numGrain), dResMean_);
    SC_HIP(hipGetLastError());
    const double jf = 2.0 * static_cast<double>(s.cfg.chargeSign) * s.temperature;
    currentBarKernel<<<blocksFor(2 * static_cast<long long>(n0)), 256, 0, stream_>>>(
        dResSbar_, dResMean_, dDomain_, static_cast<long long>(n0), jf, dResJbar_);
    SC_HIP(hipGetLastError());
  }
  const int hD = queueSum(dResD_, nullptr, nullptr, static_cast<long long>(nd), 2);
  const int hA = doA ? queueSum(dResA_, nullptr, nullptr, 2 * static_cast<long long>(n0), 2) : -1;
  // 3. the seeds of every frame's Green sums: transposes of the
  // back-rotation and of the frames' extension
  SC_HIP(hipMemsetAsync(dResSeed_, 0, 3 * size_t(framePlane) * sizeof(double), stream_));
  const int ns = static_cast<int>(s.domainSites.size());
  if (ns > 0) {
    gatherAdjointKernel<<<blocksFor(ns), 256, 0, stream_>>>(
        v, dDomainSites_, ns, dCoupling_, dResD_, doA ? dResJbar_ : nullptr, framePlane,
        dResSeed_);
    SC_HIP(hipGetLastError());
  }
  const int nr = static_cast<int>(s.extRot.target.size());
  if (nr > 0) {
    extendAdjointKernel<<<blocksFor(nr), 256, 0, stream_>>>(dResSeed_, dExtRotT_, dExtRotO_,
                                                            dExtRotS_, nr, 3, framePlane);
    SC_HIP(hipGetLastError());
  }
  // 4. the reverse sweep in residual mode
  SC_HIP(hipMemsetAsync(dAdjDbar_, 0, 3 * size_t(framePlane) * sizeof(double), stream_));
  SC_HIP(hipMemsetAsync(dAdjExt_, 0, nd * sizeof(double), stream_));
  if (withA) SC_HIP(hipMemsetAsync(dAdjGradA_, 0, 2 * n0 * sizeof(double), stream_));
  AdjointArgs a;
  a.view = sweepView(kAdjEL);
  a.items = reinterpret_cast<const SweepItemDev *>(adjItems_);
  a.numItems = adjNumItems_;
  a.counter = dCounter_;
  a.scratch = reinterpret_cast<Cx<double> *>(dAdjScratch_);
  a.slotStride = adjSlotStride_;
  a.ck = adjCkUsed_;
  a.ckSeg = adjCkSeg_;
  a.rho = dAdjRho_;
  a.seed = dResSeed_;
  a.dbar = dAdjDbar_;
  a.framePlane = framePlane;
  a.partial = dAdjPartial_;
  adjointResetCounter<<<1, 1, 0, stream_>>>(dCounter_);
  if (adjNumItems_ > 0) {
    adjointKernel<kAdjEL, true><<<adjSlots_, 64, 0, stream_>>>(a);
    SC_HIP(hipGetLastError());
  }
  // 5. the transposes of the site fields and of the order parameter's
  // extension, then the identity part
  fieldsAdjointKernel<<<dim3(blocksFor(static_cast<long long>(n0)), kFieldChunks), 256, 0,
                        stream_>>>(v, dAdjDbar_, framePlane, withA, dAdjPart_);
  SC_HIP(hipGetLastError());
  fieldsAdjointSumKernel<<<blocksFor(static_cast<long long>(2 * s.Nc + 2) * n0), 256, 0,
                           stream_>>>(dAdjPart_, s.Nc, static_cast<long long>(n0), dAdjExt_,
                                      withA ? dAdjGradA_ : nullptr);
  SC_HIP(hipGetLastError());
  const int ne = static_cast<int>(s.extRef.target.size());
  if (ne > 0) {
    extendAdjointKernel<<<blocksFor(ne), 256, 0, stream_>>>(dAdjExt_, dExtRefT_, dExtRefO_,
                                                            dExtRefS_, ne, 2 * s.Nc,
                                                            static_cast<long long>(n0));
    SC_HIP(hipGetLastError());
  }
  residualFinishKernel<<<blocksFor(static_cast<long long>(nd)), 256, 0, stream_>>>(
      dAdjExt_, dResD_, static_cast<long long>(nd), dAdjGrad_, withA ? dAdjGradA_ : nullptr,
      dResA_, 2 * static_cast<long long>(n0));
  SC_HIP(hipGetLastError());
  fromDevice(grad, dAdjGrad_, nd * sizeof(double));
  if (withA)
    fromDevice(gradA, dAdjGradA_, 2 * n0 * sizeof(double));
  else if (gradA)
    std::fill(gradA, gradA + 2 * n0, 0.0);
  flushReductions();
  double R = 0.5 * reductionSum(hD);
  if (hA >= 0) R += 0.5 * reductionSum(hA);
  clearReductions();
  // the state as it was
  {
    size_t off = 0;
    for (const Piece &pc : pieces) {
      if (pc.n > 0)
        SC_HIP(hipMemcpyAsync(pc.p, dResSave_ + off, pc.n * sizeof(double),
                              hipMemcpyDeviceToDevice, stream_));
      off += pc.n;
    }
  }
  SC_HIP(hipStreamSynchronize(stream_));
  return R;
}

} // namespace scamd
