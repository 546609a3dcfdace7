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
// and everything after it are the same kernels as the plain iteration.  A
// solve stops on its tolerance, its step cap, a line search without descent
// or a stall (two accepted steps in a row that each keep more than 0.9 of
// the residual: the floor the map's rounding sets, about 6e-10 in the
// weighted norm on a 607 x 607 grain, below which no step makes progress).
#include "sc_engine.hpp"
#include "sc_device.hpp"
#include "sc_util.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>
#include <memory>
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
// A Newton solve ends after kStallSteps accepted steps in a row that each
// keep more than kStallRatio of the residual.
constexpr double kStallRatio = 0.9;
constexpr int kStallSteps = 2;
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

// The generator of the global phase rotation on one planar complex block
// of length 2 m: d/dtheta of exp(i sign theta) z = i sign z.
__global__ void phaseGeneratorKernel(double *n, const double *u, long long m, double sign) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= m) return;
  n[i] = -sign * u[m + i];
  n[m + i] = sign * u[i];
}

unsigned blocks256(long long n) { return static_cast<unsigned>((n + 255) / 256); }

} // namespace

std::vector<Grain::NkBlock> Grain::nkBlocks() const {
  const Setup &s = *s_;
  const size_t n0 = size_t(s.N0) * s.N0;
  const bool doA = s.cfg.computeCurrent && s.lambda > 0.0;
  std::vector<NkBlock> b;
  size_t off = 0;
  // gamma rotates with the order parameter, gamma~ against it, the vector
  // potential not at all
  for (int c = 0; c < s.Nc; ++c) {
    b.push_back({off, 2 * n0, 1});
    off += 2 * n0;
  }
  if (doA) {
    b.push_back({off, 2 * n0, 0});
    off += 2 * n0;
  }
  const size_t g = size_t(2 * s.nInG * s.Ne), gt = size_t(2 * s.nInGt * s.Ne);
  if (g) b.push_back({off, g, 1});
  off += g;
  if (gt) b.push_back({off, gt, -1});
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

void Grain::nkMapBegin(bool fields) {
  // One plain iteration of the current state without the accelerator and
  // without the compute objects' reductions: the order parameter's
  // extension and the site fields, then (the caller) the sweep, then
  // nkMapEnd.
  ensureBoundaryInit();
  extendDelta(fields);
}

void Grain::nkMapEnd(double *gu) {
  const Setup &s = *s_;
  const size_t n0 = size_t(s.N0) * s.N0;
  const bool doA = s.cfg.computeCurrent && s.lambda > 0.0;
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

// ---------------------------------------------------------------------------
// The solve, one map evaluation at a time.  Every point of the algorithm
// that needs G(x) ends a phase: next() names the point (and writes it into
// the grain's state), the caller evaluates the map, advance() continues.  A
// single grain drives it with nkMap; a batch evaluates the maps of all its
// solving grains with one sweep launch per precision.
// ---------------------------------------------------------------------------
NewtonKrylov::NewtonKrylov(Grain &g, int maxNewton, int krylovDim, double tol,
                           int jvpPrecision)
    : g_(g), maxNewton_(maxNewton), m_(krylovDim), tol_(tol), jvp_(jvpPrecision) {
  if (jvpPrecision != PREC_FP64 && jvpPrecision != PREC_MIXED && jvpPrecision != JVP_TANGENT)
    throw std::runtime_error("Jacobian-vector products run in fp64 or mixed precision, or "
                             "exactly by the tangent-linear sweep");
  if (g.s_->keldysh) throw std::runtime_error("a spectroscopy context only runs the LDOS iteration");
  eta_ = std::min(eta_, g.nkEtaMax_);
  if (krylovDim < 1 || krylovDim > kMaxTerms - 1)
    throw std::runtime_error("Krylov dimension must be in [1, " + std::to_string(kMaxTerms - 1) +
                             "]");
  if (maxNewton < 0) throw std::runtime_error("max_newton must be >= 0");
  blocks_ = g.nkBlocks();
  const size_t L = g.nkLength();
  nL_ = static_cast<long long>(L);
  // workspace: u, G(u), F, trial, w, the step d, the linear residual r,
  // G(u) in the products' precision, the phase generator, and the Krylov
  // basis
  const int nv = 9 + krylovDim + 1;
  if (g.nkVecs_.size() != size_t(nv) || g.nkLen_ != L) {
    for (double *p : g.nkVecs_) devFree(p);
    g.nkVecs_.clear();
    for (int i = 0; i < nv; ++i) g.nkVecs_.push_back(devAlloc<double>(L));
    g.nkLen_ = L;
  }
  u_ = g.nkVecs_[0];
  gu_ = g.nkVecs_[1];
  f_ = g.nkVecs_[2];
  trial_ = g.nkVecs_[3];
  w_ = g.nkVecs_[4];
  dstep_ = g.nkVecs_[5];
  rlin_ = g.nkVecs_[6];
  guJ_ = g.nkVecs_[7];
  gen_ = g.nkVecs_[8];
  V_.assign(g.nkVecs_.begin() + 9, g.nkVecs_.end());
  g.nkMaps_ = 0;
  // block weights 1 / |u_b|^2 of the starting state (blocks at zero: 1 / size)
  g.ensureBoundaryInit();
  g.nkGet(u_);
  wt_.assign(blocks_.size(), 1.0);
  for (size_t b = 0; b < blocks_.size(); ++b) {
    const double uu = g.multiDot(u_ + blocks_[b].offset, {u_ + blocks_[b].offset},
                                 static_cast<long long>(blocks_[b].length))[0];
    wt_[b] = uu > 0.0 ? 1.0 / uu : 1.0 / static_cast<double>(blocks_[b].length);
  }
  phase_ = INIT;
  want(u_, gu_, PREC_FP64);
}

std::vector<double> NewtonKrylov::wdots(const double *a, const std::vector<double *> &bs) {
  // weighted dots of a with every vector of bs (the dots of every block with
  // one read-back)
  std::vector<Grain::DotGroup> groups;
  for (size_t b = 0; b < blocks_.size(); ++b) {
    std::vector<double *> part(bs.size());
    for (size_t i = 0; i < bs.size(); ++i) part[i] = bs[i] + blocks_[b].offset;
    groups.push_back({a + blocks_[b].offset, part, static_cast<long long>(blocks_[b].length)});
  }
  const std::vector<std::vector<double>> d = g_.multiDotGroups(groups);
  std::vector<double> out(bs.size(), 0.0);
  for (size_t b = 0; b < blocks_.size(); ++b)
    for (size_t i = 0; i < bs.size(); ++i) out[i] += wt_[b] * d[b][i];
  return out;
}

double NewtonKrylov::wnorm(double *a) { return std::sqrt(wdots(a, {a})[0]); }

void NewtonKrylov::combine(double *y, const double *base, double a,
                           const std::vector<const double *> &vs, const std::vector<double> &cs) {
  if (vs.size() > size_t(kMaxTerms)) throw std::runtime_error("too many terms");
  Terms t{};
  for (size_t q = 0; q < vs.size(); ++q) {
    t.v[q] = vs[q];
    t.c[q] = cs[q];
  }
  combineKernel<<<blocks256(nL_), 256, 0, g_.stream_>>>(y, base, a, t,
                                                         static_cast<int>(vs.size()), nL_);
  SC_NK_HIP(hipGetLastError());
}

void NewtonKrylov::want(const double *x, double *out, int precision) {
  x_ = x;
  out_ = out;
  prec_ = precision;
}

bool NewtonKrylov::next(int &precision, bool fields) {
  // exact products need no map: each is the tangent-linear sweep at the
  // Newton point, until the solve asks for G again
  while (phase_ == KRYLOV && jvp_ == JVP_TANGENT) {
    g_.nkSet(u_);
    g_.jacobianVector(V_[size_t(k_)], w_, true, true);
    productDone();
  }
  if (phase_ == DONE) return false;
  g_.nkSet(x_);
  g_.nkMapBegin(fields);
  precision = prec_;
  return true;
}

void NewtonKrylov::startNewton() {
  // restarted GMRES(m) on J d = -F from d = 0
  const double unorm = wnorm(u_);
  e_ = (jvp_ == PREC_MIXED ? 3e-4 : 1e-7) * (1.0 + unorm);
  combine(dstep_, nullptr, 0.0, {}, {});
  combine(rlin_, nullptr, 0.0, {f_}, {-1.0});
  rn_ = fn_;
  cycle_ = 0;
  if (jvp_ == PREC_MIXED) {
    // the base point of the differences, in the products' precision
    phase_ = BASE;
    want(u_, guJ_, PREC_MIXED);
  } else {
    startCycle();
  }
}

void NewtonKrylov::startCycle() {
  beta_ = rn_;
  combine(V_[0], nullptr, 0.0, {rlin_}, {1.0 / beta_});
  // H: Hessenberg matrix as built (for the restart residual), R: with the
  // Givens rotations applied
  const size_t m = size_t(m_);
  H_.assign((m + 1) * m, 0.0);
  R_.assign((m + 1) * m, 0.0);
  gv_.assign(m + 1, 0.0);
  cs_.assign(m, 0.0);
  sn_.assign(m, 0.0);
  gv_[0] = beta_;
  k_ = 0;
  breakdown_ = false;
  startProduct();
}

void NewtonKrylov::startProduct() {
  phase_ = KRYLOV;
  if (jvp_ == JVP_TANGENT) return; // next() forms J v itself
  combine(trial_, u_, 1.0, {V_[size_t(k_)]}, {e_});
  want(trial_, w_, jvp_);
}

void NewtonKrylov::productDone() {
  const int m = m_, k = k_;
  if (jvp_ == JVP_TANGENT) {
    // w holds J v: (J - I) v
    combine(w_, w_, 1.0, {V_[size_t(k)]}, {-1.0});
  } else {
    const double *base = jvp_ == PREC_MIXED ? guJ_ : gu_;
    combine(w_, w_, 1.0 / e_, {base, V_[size_t(k)]}, {-1.0 / e_, -1.0});
  }
  // classical Gram-Schmidt with the norm of w read back with the
  // projections: |w - sum d_i v_i|^2 = |w|^2 - sum d_i^2 for the orthonormal
  // basis.  A second pass only when the first removed more than half of
  // |w|^2 (Kahan-Parlett: then the difference has lost digits and the
  // projection must be repeated).
  std::vector<double *> basis(V_.begin(), V_.begin() + k + 1);
  std::vector<const double *> cbasis(basis.begin(), basis.end());
  std::vector<double *> withW(basis);
  withW.push_back(w_);
  std::vector<double> h(size_t(k) + 1, 0.0);
  double hn = 0.0;
  for (int pass = 0; pass < 2; ++pass) {
    const std::vector<double> d = wdots(w_, withW);
    const double ww = d[size_t(k) + 1];
    double removed = 0.0;
    std::vector<double> neg(size_t(k) + 1);
    for (size_t i = 0; i <= size_t(k); ++i) {
      h[i] += d[i];
      neg[i] = -d[i];
      removed += d[i] * d[i];
    }
    combine(w_, w_, 1.0, cbasis, neg);
    const double rem = ww - removed;
    hn = std::sqrt(std::max(rem, 0.0));
    if (rem > 0.5 * ww) break;
  }
  for (int i = 0; i <= k; ++i) H_[size_t(i) * m + k] = R_[size_t(i) * m + k] = h[size_t(i)];
  H_[size_t(k + 1) * m + k] = R_[size_t(k + 1) * m + k] = hn;
  for (int i = 0; i < k; ++i) {
    const double a = R_[size_t(i) * m + k], b = R_[size_t(i + 1) * m + k];
    R_[size_t(i) * m + k] = cs_[size_t(i)] * a + sn_[size_t(i)] * b;
    R_[size_t(i + 1) * m + k] = -sn_[size_t(i)] * a + cs_[size_t(i)] * b;
  }
  const double a = R_[size_t(k) * m + k], b = R_[size_t(k + 1) * m + k];
  const double r = std::hypot(a, b);
  cs_[size_t(k)] = r > 0.0 ? a / r : 1.0;
  sn_[size_t(k)] = r > 0.0 ? b / r : 0.0;
  R_[size_t(k) * m + k] = r;
  R_[size_t(k + 1) * m + k] = 0.0;
  gv_[size_t(k + 1)] = -sn_[size_t(k)] * gv_[size_t(k)];
  gv_[size_t(k)] = cs_[size_t(k)] * gv_[size_t(k)];
  breakdown_ = !(hn > 0.0);
  if (!breakdown_) combine(V_[size_t(k + 1)], nullptr, 0.0, {w_}, {1.0 / hn});
  ++k_;
  if (std::fabs(gv_[size_t(k + 1)]) <= eta_ * fn_ || breakdown_ || k_ == m) endCycle();
  else startProduct();
}

void NewtonKrylov::endCycle() {
  const int m = m_, k = k_;
  // back substitution: R y = g on the first k rows
  std::vector<double> y(size_t(k), 0.0);
  for (int i = k - 1; i >= 0; --i) {
    double acc = gv_[size_t(i)];
    for (int q = i + 1; q < k; ++q) acc -= R_[size_t(i) * m + q] * y[size_t(q)];
    y[size_t(i)] = R_[size_t(i) * m + i] != 0.0 ? acc / R_[size_t(i) * m + i] : 0.0;
  }
  rep_.krylovIterations += k;
  std::vector<const double *> vk(V_.begin(), V_.begin() + k);
  combine(dstep_, dstep_, 1.0, vk, y);
  rn_ = std::fabs(gv_[size_t(k)]);
  if (rn_ <= eta_ * fn_ || breakdown_ || cycle_ + 1 == g_.nkRestarts_) {
    startLineSearch();
    return;
  }
  // the linear residual -F - J d = V_(k+1) (beta e1 - H y)
  std::vector<double> c(size_t(k) + 1, 0.0);
  std::vector<const double *> vk1(V_.begin(), V_.begin() + k + 1);
  for (int i = 0; i <= k; ++i) {
    double acc = i == 0 ? beta_ : 0.0;
    for (int q = 0; q < k; ++q) acc -= H_[size_t(i) * m + q] * y[size_t(q)];
    c[size_t(i)] = acc;
  }
  combine(rlin_, nullptr, 0.0, vk1, c);
  rn_ = wnorm(rlin_);
  ++cycle_;
  startCycle();
}

void NewtonKrylov::removePhase() {
  // The global phase of the order parameter is a symmetry of the map: every
  // rotation of a fixed point is one, so J - I is singular along the
  // rotation's generator n at u, and GMRES may put an arbitrary (and, in an
  // inconsistent least-squares solve, large) multiple of n into d, which
  // the line search then meets as a second-order growth of |F|.  The
  // component of d along n (in the weighted inner product) is removed.
  combine(gen_, nullptr, 0.0, {}, {});
  for (const Grain::NkBlock &b : blocks_) {
    if (b.phase == 0) continue;
    const long long m = static_cast<long long>(b.length / 2);
    phaseGeneratorKernel<<<blocks256(m), 256, 0, g_.stream_>>>(gen_ + b.offset, u_ + b.offset,
                                                               m, double(b.phase));
    SC_NK_HIP(hipGetLastError());
  }
  const std::vector<double> d = wdots(gen_, {dstep_, gen_});
  if (d[1] > 0.0) combine(dstep_, dstep_, 1.0, {gen_}, {-d[0] / d[1]});
}

void NewtonKrylov::startLineSearch() {
  // backtracking line search on |F| along d
  if (g_.nkFixPhase_) removePhase();
  lambda_ = 1.0;
  ls_ = 0;
  combine(trial_, u_, 1.0, {dstep_}, {lambda_});
  phase_ = LINE;
  want(trial_, w_, PREC_FP64);
}

void NewtonKrylov::lineDone() {
  combine(rlin_, w_, 1.0, {trial_}, {-1.0});
  const double fnew = wnorm(rlin_);
  if (fnew < (1.0 - 1e-4 * lambda_) * fn_) {
    std::swap(u_, trial_);
    std::swap(gu_, w_);
    std::swap(f_, rlin_);
    const double fnPrev = fn_;
    fn_ = fnew;
    rep_.history.push_back(fn_);
    ++rep_.newtonIterations;
    // Eisenstat-Walker forcing term (choice 2)
    eta_ = std::clamp(0.9 * (fn_ / fnPrev) * (fn_ / fnPrev), 1e-4, g_.nkEtaMax_);
    // Stagnation: Newton converges superlinearly near a regular solution,
    // so accepted steps that each keep more than kStallRatio of the
    // residual mean the map's rounding floor (or a singular direction) has
    // been reached; kStallSteps of them in a row end the solve.
    stalled_ = fn_ > kStallRatio * fnPrev ? stalled_ + 1 : 0;
    if (fn_ < tol_ || rep_.newtonIterations >= maxNewton_ || stalled_ >= kStallSteps)
      phase_ = DONE;
    else startNewton();
    return;
  }
  lambda_ *= 0.5;
  if (++ls_ == 6) {
    // no descent along the Newton direction: with FP32 products their
    // rounding is the likely limit, so the step again on FP64 products;
    // else the end
    if (jvp_ == PREC_MIXED && g_.nkFallback_ && rep_.newtonIterations < maxNewton_) {
      jvp_ = PREC_FP64;
      startNewton();
    } else {
      phase_ = DONE;
    }
    return;
  }
  combine(trial_, u_, 1.0, {dstep_}, {lambda_});
  want(trial_, w_, PREC_FP64);
}

void NewtonKrylov::advance() {
  switch (phase_) {
  case INIT:
    combine(f_, gu_, 1.0, {u_}, {-1.0});
    fn_ = wnorm(f_);
    rep_.history.push_back(fn_);
    if (fn_ < tol_ || maxNewton_ == 0) phase_ = DONE;
    else startNewton();
    break;
  case BASE:
    startCycle();
    break;
  case KRYLOV:
    productDone();
    break;
  case LINE:
    lineDone();
    break;
  case DONE:
    break;
  }
}

NewtonReport NewtonKrylov::finish() {
  // the workspace pointers were permuted by the accepted steps: hand the
  // pool back, then leave the last accepted u in the state
  g_.nkVecs_[0] = u_;
  g_.nkVecs_[1] = gu_;
  g_.nkVecs_[2] = f_;
  g_.nkVecs_[3] = trial_;
  g_.nkVecs_[4] = w_;
  g_.nkVecs_[6] = rlin_;
  g_.nkSet(u_);
  SC_NK_HIP(hipStreamSynchronize(g_.stream_));
  rep_.residual = fn_;
  rep_.converged = fn_ < tol_;
  rep_.mapEvaluations = g_.nkMaps_;
  // The iterate now lives in the state; start the accelerators afresh.
  g_.resetAccelerator();
  return rep_;
}

std::vector<NewtonReport> Batch::newtonKrylov(int maxNewton, int krylovDim, double tol,
                                              int jvpPrecision, const std::vector<int> &which) {
  // Every grain runs its own solve (its own Krylov bases, Hessenberg
  // matrices, line search and stopping), in lockstep: each round every
  // solve that still runs names the state it needs G of, one sweep launch
  // per precision covers all of them, each grain finishes its map on its
  // lane, and the solves advance.  Grains whose solve has ended leave the
  // launches, so the last rounds sweep only the slow grains.
  const size_t n = grains_.size();
  if (!which.empty() && which.size() != n)
    throw std::runtime_error("one selection flag per grain expected");
  std::vector<std::unique_ptr<NewtonKrylov>> nk(n);
  for (size_t g = 0; g < n; ++g)
    if (which.empty() || which[g])
      nk[g] = std::make_unique<NewtonKrylov>(*grains_[g], maxNewton, krylovDim, tol,
                                             jvpPrecision);
  for (;;) {
    std::vector<int> group[2], all;
    for (size_t g = 0; g < n; ++g) {
      int precision = PREC_FP64;
      if (!nk[g] || !nk[g]->next(precision, false)) continue;
      group[precision].push_back(static_cast<int>(g));
      all.push_back(static_cast<int>(g));
    }
    if (all.empty()) break;
    for (int p : {PREC_FP64, PREC_MIXED})
      if (!group[p].empty()) sweep(p, false, group[p], true);
    for (int g : all) grains_[size_t(g)]->nkMapEnd(nk[size_t(g)]->out());
    joinLanes();
    SC_NK_HIP(hipStreamSynchronize(stream_));
    for (int g : all) nk[size_t(g)]->advance();
  }
  std::vector<NewtonReport> reps(n);
  for (size_t g = 0; g < n; ++g)
    if (nk[g]) reps[g] = nk[g]->finish();
  return reps;
}

NewtonReport Grain::newtonKrylov(int maxNewton, int krylovDim, double tol, int jvpPrecision) {
  NewtonKrylov nk(*this, maxNewton, krylovDim, tol, jvpPrecision);
  int precision = PREC_FP64;
  while (nk.next(precision)) {
    sweep(precision, false);
    nkMapEnd(nk.out());
    nk.advance();
  }
  return nk.finish();
}

} // namespace scamd
