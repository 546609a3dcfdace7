// sc_newton.hip -- Jacobian-free Newton-Krylov solve of the self-consistency
// fixed point (declared in sc_engine.hpp).
//
// The fixed point of one iteration is G(u) = u, where the state u holds the
// order parameter, the scaled induced vector potential and the boundary
// in-values of both coherence functions (the reference iterates all three:
// Context.h:391-447 mixes the first two, the boundary condition updates the
// third).  Newton's method solves F(u) = G(u) - u = 0; each step solves
// J d = -F with restarted GMRES in the block-weighted inner product
// <a, b> = sum_blocks <a_b, b_b> / |u_b|^2, so every block is measured
// relative to its size like the reference's residuals.  Jacobian-vector
// products are finite differences of G: J v = (G(u + e v) - G(u)) / e - v.
// The Newton residual G(u) - u is always evaluated in FP64.  The
// Jacobian-vector products use the FP64 map, or (mixed precision) the map
// with FP32 stepping and FP64 sums on both ends of the difference, with a
// step e large enough that the FP32 rounding (~1e-7 relative) stays at
// ~1e-3 of the product: the inner solve is then an inexact Newton direction
// whose error the FP64 residuals of the outer iteration correct.  The sweep
// and everything after it are the same kernels as the plain iteration.
#include "sc_engine.hpp"
#include "sc_device.hpp"
#include "sc_util.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#define SC_NK_HIP(call)                                                             \
  do {                                                                              \
    const hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess)                                                           \
      throw std::runtime_error(std::string("HIP error in the Newton-Krylov solve: ") + \
                               hipGetErrorString(e_));                              \
  } while (0)

namespace scamd {
namespace {

constexpr int kMaxTerms = 64;
// GMRES cycles per Newton step (each of at most krylovDim directions)
constexpr int kRestarts = 4;
struct Terms {
  const double *v[kMaxTerms];
  double c[kMaxTerms];
};

// y = a * base + sum_i c_i v_i (base may alias y)
__global__ void combineKernel(double *y, const double *base, double a, Terms t, int m,
                              long long n) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double acc = base ? a * base[i] : 0.0;
  for (int q = 0; q < m; ++q) acc += t.c[q] * t.v[q][i];
  y[i] = acc;
}

unsigned blocks256(long long n) { return static_cast<unsigned>((n + 255) / 256); }

} // namespace

std::vector<Grain::NkBlock> Grain::nkBlocks() const {
  const Setup &s = *s_;
  const size_t n0 = size_t(s.N0) * s.N0;
  const bool doA = s.cfg.computeCurrent && s.lambda > 0.0;
  std::vector<NkBlock> b;
  size_t off = 0;
  for (int c = 0; c < s.Nc; ++c) {
    b.push_back({off, 2 * n0});
    off += 2 * n0;
  }
  if (doA) {
    b.push_back({off, 2 * n0});
    off += 2 * n0;
  }
  const size_t g = size_t(2 * s.nInG * s.Ne), gt = size_t(2 * s.nInGt * s.Ne);
  if (g) b.push_back({off, g});
  off += g;
  if (gt) b.push_back({off, gt});
  return b;
}

size_t Grain::nkLength() const {
  const Setup &s = *s_;
  const size_t n0 = size_t(s.N0) * s.N0;
  const bool doA = s.cfg.computeCurrent && s.lambda > 0.0;
  return size_t(2 * s.Nc + (doA ? 2 : 0)) * n0 + size_t(2 * s.nInG * s.Ne) +
         size_t(2 * s.nInGt * s.Ne);
}

void Grain::nkGet(double *u) {
  const Setup &s = *s_;
  const size_t n0 = size_t(s.N0) * s.N0;
  const bool doA = s.cfg.computeCurrent && s.lambda > 0.0;
  const size_t nx = size_t(2 * s.Nc + (doA ? 2 : 0)) * n0;
  const size_t g = size_t(2 * s.nInG * s.Ne), gt = size_t(2 * s.nInGt * s.Ne);
  SC_NK_HIP(hipMemcpyAsync(u, dDelta_, nx * sizeof(double), hipMemcpyDeviceToDevice, stream_));
  SC_NK_HIP(hipMemcpyAsync(u + nx, dBInG_, g * sizeof(double), hipMemcpyDeviceToDevice, stream_));
  SC_NK_HIP(hipMemcpyAsync(u + nx + g, dBInGt_, gt * sizeof(double), hipMemcpyDeviceToDevice,
                           stream_));
}

void Grain::nkSet(const double *u) {
  const Setup &s = *s_;
  const size_t n0 = size_t(s.N0) * s.N0;
  const bool doA = s.cfg.computeCurrent && s.lambda > 0.0;
  const size_t nx = size_t(2 * s.Nc + (doA ? 2 : 0)) * n0;
  const size_t g = size_t(2 * s.nInG * s.Ne), gt = size_t(2 * s.nInGt * s.Ne);
  SC_NK_HIP(hipMemcpyAsync(dDelta_, u, nx * sizeof(double), hipMemcpyDeviceToDevice, stream_));
  SC_NK_HIP(hipMemcpyAsync(dBInG_, u + nx, g * sizeof(double), hipMemcpyDeviceToDevice, stream_));
  SC_NK_HIP(hipMemcpyAsync(dBInGt_, u + nx + g, gt * sizeof(double), hipMemcpyDeviceToDevice,
                           stream_));
}

void Grain::nkMap(double *gu, int precision) {
  // One plain iteration of the current state without the accelerator and
  // without the compute objects' reductions.
  const Setup &s = *s_;
  const size_t n0 = size_t(s.N0) * s.N0;
  const bool doA = s.cfg.computeCurrent && s.lambda > 0.0;
  ensureBoundaryInit();
  extendDelta();
  sweep(precision, false);
  backRotate();
  int hb[2];
  queueBoundaryCondition(hb); // new in-values, in place
  if (doA) {
    solveVectorPotential(dCurrent_, dANew_);
    if (screen_) screenVectorPotential();
  }
  // the map needs none of the residual statistics: no read-back
  discardReductions();
  const size_t nx = size_t(2 * s.Nc + (doA ? 2 : 0)) * n0;
  const size_t g = size_t(2 * s.nInG * s.Ne), gt = size_t(2 * s.nInGt * s.Ne);
  SC_NK_HIP(hipMemcpyAsync(gu, dDeltaNew_, nx * sizeof(double), hipMemcpyDeviceToDevice,
                           stream_));
  SC_NK_HIP(hipMemcpyAsync(gu + nx, dBInG_, g * sizeof(double), hipMemcpyDeviceToDevice,
                           stream_));
  SC_NK_HIP(hipMemcpyAsync(gu + nx + g, dBInGt_, gt * sizeof(double), hipMemcpyDeviceToDevice,
                           stream_));
  ++nkMaps_;
}

NewtonReport Grain::newtonKrylov(int maxNewton, int krylovDim, double tol, int jvpPrecision) {
  if (jvpPrecision != PREC_FP64 && jvpPrecision != PREC_MIXED)
    throw std::runtime_error("Jacobian-vector products run in fp64 or mixed precision");
  const bool mixedJvp = jvpPrecision == PREC_MIXED;
  if (s_->keldysh) throw std::runtime_error("a spectroscopy context only runs the LDOS iteration");
  if (krylovDim < 1 || krylovDim > kMaxTerms - 1)
    throw std::runtime_error("Krylov dimension must be in [1, " + std::to_string(kMaxTerms - 1) +
                             "]");
  const std::vector<NkBlock> blocks = nkBlocks();
  const size_t L = nkLength();
  const long long nL = static_cast<long long>(L);
  // workspace: u, G(u), F, trial, w, the step d, the linear residual r,
  // G(u) in the products' precision, and the Krylov basis
  const int nv = 8 + krylovDim + 1;
  if (nkVecs_.size() != size_t(nv) || nkLen_ != L) {
    for (double *p : nkVecs_) devFree(p);
    nkVecs_.clear();
    for (int i = 0; i < nv; ++i) nkVecs_.push_back(devAlloc<double>(L));
    nkLen_ = L;
  }
  double *u = nkVecs_[0], *gu = nkVecs_[1], *f = nkVecs_[2], *trial = nkVecs_[3],
         *w = nkVecs_[4], *dstep = nkVecs_[5], *rlin = nkVecs_[6], *guJ = nkVecs_[7];
  std::vector<double *> V(nkVecs_.begin() + 8, nkVecs_.end());
  nkMaps_ = 0;
  NewtonReport rep;

  // block weights 1 / |u_b|^2 of the starting state (blocks at zero: 1 / size)
  ensureBoundaryInit();
  nkGet(u);
  std::vector<double> wt(blocks.size(), 1.0);
  for (size_t b = 0; b < blocks.size(); ++b) {
    const double uu = multiDot(u + blocks[b].offset, {u + blocks[b].offset},
                               static_cast<long long>(blocks[b].length))[0];
    wt[b] = uu > 0.0 ? 1.0 / uu : 1.0 / static_cast<double>(blocks[b].length);
  }
  // weighted dots of a with every vector of bs
  // (the dots of every block with one read-back)
  auto wdots = [&](const double *a, const std::vector<double *> &bs) {
    std::vector<DotGroup> groups;
    for (size_t b = 0; b < blocks.size(); ++b) {
      std::vector<double *> part(bs.size());
      for (size_t i = 0; i < bs.size(); ++i) part[i] = bs[i] + blocks[b].offset;
      groups.push_back({a + blocks[b].offset, part, static_cast<long long>(blocks[b].length)});
    }
    const std::vector<std::vector<double>> d = multiDotGroups(groups);
    std::vector<double> out(bs.size(), 0.0);
    for (size_t b = 0; b < blocks.size(); ++b)
      for (size_t i = 0; i < bs.size(); ++i) out[i] += wt[b] * d[b][i];
    return out;
  };
  auto wnorm = [&](double *a) { return std::sqrt(wdots(a, {a})[0]); };
  auto combine = [&](double *y, const double *base, double a,
                     const std::vector<const double *> &vs, const std::vector<double> &cs) {
    if (vs.size() > size_t(kMaxTerms)) throw std::runtime_error("too many terms");
    Terms t{};
    for (size_t q = 0; q < vs.size(); ++q) {
      t.v[q] = vs[q];
      t.c[q] = cs[q];
    }
    combineKernel<<<blocks256(nL), 256, 0, stream_>>>(y, base, a, t, static_cast<int>(vs.size()),
                                                      nL);
    SC_NK_HIP(hipGetLastError());
  };
  // F(x) into fout given x: writes G(x) into gout
  auto residual = [&](const double *x, double *gout, double *fout) {
    nkSet(x);
    nkMap(gout);
    combine(fout, gout, 1.0, {x}, {-1.0});
    return wnorm(fout);
  };

  const int m = krylovDim;
  double fn = residual(u, gu, f);
  double fnPrev = fn;
  rep.history.push_back(fn);
  double eta = 0.1;
  for (int it = 0; it < maxNewton && fn >= tol; ++it) {
    // restarted GMRES(m) on J d = -F from d = 0
    const double unorm = wnorm(u);
    const double e = (mixedJvp ? 3e-4 : 1e-7) * (1.0 + unorm);
    // the base point of the differences, in the products' precision
    const double *base = gu;
    if (mixedJvp) {
      nkSet(u);
      nkMap(guJ, PREC_MIXED);
      base = guJ;
    }
    combine(dstep, nullptr, 0.0, {}, {});
    combine(rlin, nullptr, 0.0, {f}, {-1.0});
    double rn = fn;
    for (int cycle = 0; cycle < kRestarts; ++cycle) {
      const double beta = rn;
      combine(V[0], nullptr, 0.0, {rlin}, {1.0 / beta});
      // H: Hessenberg matrix as built (for the restart residual), R: with
      // the Givens rotations applied
      std::vector<double> H(size_t(m + 1) * m, 0.0), R(size_t(m + 1) * m, 0.0),
          g(size_t(m + 1), 0.0), cs(size_t(m), 0.0), sn(size_t(m), 0.0);
      g[0] = beta;
      int k = 0;
      bool breakdown = false;
      for (; k < m; ++k) {
        combine(trial, u, 1.0, {V[size_t(k)]}, {e});
        nkSet(trial);
        nkMap(w, jvpPrecision);
        combine(w, w, 1.0 / e, {base, V[size_t(k)]}, {-1.0 / e, -1.0});
        // classical Gram-Schmidt with the norm of w read back with the
        // projections: |w - sum d_i v_i|^2 = |w|^2 - sum d_i^2 for the
        // orthonormal basis.  A second pass only when the first removed
        // more than half of |w|^2 (Kahan-Parlett: then the difference has
        // lost digits and the projection must be repeated).
        std::vector<double *> basis(V.begin(), V.begin() + k + 1);
        std::vector<const double *> cbasis(basis.begin(), basis.end());
        std::vector<double *> withW(basis);
        withW.push_back(w);
        std::vector<double> h(size_t(k) + 1, 0.0);
        double hn = 0.0;
        for (int pass = 0; pass < 2; ++pass) {
          const std::vector<double> d = wdots(w, withW);
          const double ww = d[size_t(k) + 1];
          double removed = 0.0;
          std::vector<double> neg(size_t(k) + 1);
          for (size_t i = 0; i <= size_t(k); ++i) {
            h[i] += d[i];
            neg[i] = -d[i];
            removed += d[i] * d[i];
          }
          combine(w, w, 1.0, cbasis, neg);
          const double rem = ww - removed;
          hn = std::sqrt(std::max(rem, 0.0));
          if (rem > 0.5 * ww) break;
        }
        for (int i = 0; i <= k; ++i) H[size_t(i) * m + k] = R[size_t(i) * m + k] = h[size_t(i)];
        H[size_t(k + 1) * m + k] = R[size_t(k + 1) * m + k] = hn;
        for (int i = 0; i < k; ++i) {
          const double a = R[size_t(i) * m + k], b = R[size_t(i + 1) * m + k];
          R[size_t(i) * m + k] = cs[size_t(i)] * a + sn[size_t(i)] * b;
          R[size_t(i + 1) * m + k] = -sn[size_t(i)] * a + cs[size_t(i)] * b;
        }
        const double a = R[size_t(k) * m + k], b = R[size_t(k + 1) * m + k];
        const double r = std::hypot(a, b);
        cs[size_t(k)] = r > 0.0 ? a / r : 1.0;
        sn[size_t(k)] = r > 0.0 ? b / r : 0.0;
        R[size_t(k) * m + k] = r;
        R[size_t(k + 1) * m + k] = 0.0;
        g[size_t(k + 1)] = -sn[size_t(k)] * g[size_t(k)];
        g[size_t(k)] = cs[size_t(k)] * g[size_t(k)];
        breakdown = !(hn > 0.0);
        if (!breakdown) combine(V[size_t(k + 1)], nullptr, 0.0, {w}, {1.0 / hn});
        if (std::fabs(g[size_t(k + 1)]) <= eta * fn || breakdown) {
          ++k;
          break;
        }
      }
      // back substitution: R y = g on the first k rows
      std::vector<double> y(size_t(k), 0.0);
      for (int i = k - 1; i >= 0; --i) {
        double acc = g[size_t(i)];
        for (int q = i + 1; q < k; ++q) acc -= R[size_t(i) * m + q] * y[size_t(q)];
        y[size_t(i)] = R[size_t(i) * m + i] != 0.0 ? acc / R[size_t(i) * m + i] : 0.0;
      }
      rep.krylovIterations += k;
      std::vector<const double *> vk(V.begin(), V.begin() + k);
      combine(dstep, dstep, 1.0, vk, y);
      rn = std::fabs(g[size_t(k)]);
      if (rn <= eta * fn || breakdown || cycle + 1 == kRestarts) break;
      // the linear residual -F - J d = V_(k+1) (beta e1 - H y)
      std::vector<double> c(size_t(k) + 1, 0.0);
      std::vector<const double *> vk1(V.begin(), V.begin() + k + 1);
      for (int i = 0; i <= k; ++i) {
        double acc = i == 0 ? beta : 0.0;
        for (int q = 0; q < k; ++q) acc -= H[size_t(i) * m + q] * y[size_t(q)];
        c[size_t(i)] = acc;
      }
      combine(rlin, nullptr, 0.0, vk1, c);
      rn = wnorm(rlin);
    }
    // backtracking line search on |F| along d
    double lambda = 1.0, fnew = 0.0;
    bool accepted = false;
    for (int ls = 0; ls < 6; ++ls) {
      combine(trial, u, 1.0, {dstep}, {lambda});
      fnew = residual(trial, w, rlin);
      if (fnew < (1.0 - 1e-4 * lambda) * fn) {
        accepted = true;
        break;
      }
      lambda *= 0.5;
    }
    if (!accepted) break; // no descent along the Newton direction
    std::swap(u, trial);
    std::swap(gu, w);
    std::swap(f, rlin);
    fnPrev = fn;
    fn = fnew;
    rep.history.push_back(fn);
    ++rep.newtonIterations;
    // Eisenstat-Walker forcing term (choice 2)
    eta = std::clamp(0.9 * (fn / fnPrev) * (fn / fnPrev), 1e-4, 0.3);
    nkVecs_[0] = u;
    nkVecs_[1] = gu;
    nkVecs_[2] = f;
    nkVecs_[3] = trial;
    nkVecs_[4] = w;
    nkVecs_[6] = rlin;
  }
  nkSet(u);
  SC_NK_HIP(hipStreamSynchronize(stream_));
  rep.residual = fn;
  rep.converged = fn < tol;
  rep.mapEvaluations = nkMaps_;
  // The iterate now lives in the state; start the accelerators afresh.
  resetAccelerator();
  return rep;
}

} // namespace scamd
