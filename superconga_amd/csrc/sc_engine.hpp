// sc_engine.hpp -- one grain on one GPU: device state and the iteration.
#pragma once

#include "sc_fft.hpp"
#include "sc_view.hpp"
#include "sc_setup.hpp"

#include <hip/hip_runtime.h>

#include <array>
#include <memory>
#include <string>
#include <vector>

namespace scamd {

enum Precision { PREC_FP64 = 0, PREC_MIXED = 1, PREC_MIXED_REFINE = 2 };
// Newton-Krylov's exact Jacobian-vector products (the tangent-linear sweep)
constexpr int JVP_TANGENT = 3;

// Energies per lane of the reverse and tangent-linear sweeps (sc_adjoint.hip).
constexpr int kAdjointEL = 4;

// The tangent-linear sweep (sc_adjoint.hip): on the items of the reverse
// sweep (SweepItemDev, kAdjointEL energies per lane, view = sweepView), both
// Riccati recurrences with their tangents for the tangent site fields
// tfields (laid out like view.fields) and the tangents of the in-values
// tInG, tInGt (null: zero); writes the tangents of the out-values at every
// exit crossing and, per inside frame site, the tangents of the Green sums
// into tG ([frame][site][4]: gap re, gap im, current, 0, the sweep's layout).
void tangentSweep(const GrainView &view, const double *tfields, const void *items, int numItems,
                  int *counter, double *scratch, long long slotStride, const double *tInG,
                  const double *tInGt, double *tOutG, double *tOutGt, double *tG, int slots,
                  hipStream_t stream);

struct Results {
  double residual = 0, resBoundary = 0, resDelta = 0, resCurrent = 0,
         resFreeEnergy = 0, resVectorPotential = 0;
  double freeEnergy = 0, freeEnergyPerArea = 0, magneticFreeEnergy = 0,
         magneticFreeEnergyPerArea = 0;
  int iterations = 0;
  int converged = 0;
};

struct Timing {
  double sweepMs = 0;  // last sweep kernel time (HIP events)
  double iterMs = 0;   // last whole iteration
  long long trajSteps = 0;
};

// Outcome of a Newton-Krylov solve: Newton steps, GMRES iterations, map
// evaluations (plain iterations) and the block-weighted residual norm after
// every Newton step.
struct NewtonReport {
  int newtonIterations = 0, krylovIterations = 0, mapEvaluations = 0;
  double residual = 0;
  bool converged = false;
  std::vector<double> history;
};

// State of a Keldysh (LDOS) context.
struct LdosStatus {
  int numEnergies = 0, perBlock = 0, blockStart = 0, blockCount = 0;
  int iterations = 0, angleChunk = 0;
  bool partiallyConverged = false, converged = false;
  double residual = 0, residualBoundary = 0, residualLdos = 0;
  double sweepMs = 0, iterationMs = 0;
  long long trajSteps = 0; // per iteration of the current block
};

class Grain {
public:
  // With a stream the grain runs on it (and does not own it); without, it
  // creates its own.
  explicit Grain(std::unique_ptr<Setup> setup, hipStream_t stream = nullptr);
  ~Grain();
  Grain(const Grain &) = delete;
  Grain &operator=(const Grain &) = delete;

  const Setup &setup() const { return *s_; }

  // host I/O (planar re/im, reference GridData layout)
  void setOrderParameter(int comp, const double *re, const double *im);
  void getOrderParameter(int which, int comp, double *re, double *im);
  void setVectorPotential(const double *ax, const double *ay, bool scaled);
  void getVectorPotential(int which, double *ax, double *ay, bool scaled);
  void getCurrent(double *jx, double *jy);
  void getBoundary(int inOut, int dir, double *re, double *im);
  void setBoundary(int dir, const double *re, const double *im);
  void getEnergyDensity(double *out);
  void initialGuess(); // Initial<T>::bulk / perturbedPhaseWindings

  // the self-consistency iteration (Context.h:377-447)
  void compute(int precision);
  void burnin(int precision);
  // The same iteration in two halves around the sweep, for callers that
  // sweep several grains in one launch: beginCompute(); <sweep>; endCompute().
  // endCompute = queueEndCompute (device work) + finishEndCompute (read-back),
  // so that a batch can queue every grain before it waits on any.
  void beginCompute();
  void queueBeginCompute();
  // fields = false leaves the site fields to the caller (a batch computes
  // those of all its grains in one launch)
  void finishBeginCompute(bool fields = true);
  void endCompute();
  void queueEndCompute();
  void finishEndCompute();
  void beginBurnin(bool fields = true);
  void endBurnin();
  void queueEndBurnin();
  void finishEndBurnin();
  // starts the transfer of the queued reductions (flushReductions waits)
  void queueReadback();
  bool run(int burnin, int itMin, int itMax, int precision);
  // The same iteration angle by angle, the way IntegrationIteratorOzaki::
  // compute drives it (IntegrationIteratorOzaki.h:300-420):
  //   beginIntegration(burnin);
  //   for consecutive angle ranges [k0, k1) covering [0, Na):
  //     sweepAngles(precision, burnin, k0, k1); integrateAngles(k0, k1);
  //   endIntegration(burnin);
  // equals compute() / burnin() bit for bit.  integrateAngles must see the
  // ranges in increasing order (the angle sums are added in that order);
  // a burn-in integrates nothing.
  void beginIntegration(bool burnin);
  void sweepAngles(int precision, bool burnin, int k0, int k1);
  void integrateAngles(int k0, int k1);
  void endIntegration(bool burnin);
  // The accelerator step on caller-supplied fields: x (the iterate) and g
  // (the map's result), each the concatenated field [Re D_0, Im D_0, ...,
  // (lambda^2 Ax, lambda^2 Ay when the grain has an induced potential)] of
  // numFields() N0^2 values; x receives the accelerated iterate.  The
  // accelerator's history is the grain's own.
  void acceleratorStepOn(double *x, const double *g);
  int numFields() const;
  // Replay the device work after the sweep (back-rotation and angle
  // integration, gap equation, current, boundary condition, Poisson solve
  // and their reductions) from a hipGraph captured on its second use, while
  // the reductions it was captured with keep the same offsets.  On by
  // default; off runs every launch eagerly.
  void setGraphs(bool on);
  // The screened vector-potential update (accelerator.vector_potential_
  // screening) switched on with superfluid density rho > 0, or off (0);
  // restarts the accelerator's history.
  void setScreening(double rho);
  bool graphs() const { return graphsOn_; }
  // individual stages
  // Reads the HIP-event time of the last sweep launch into the timing.
  void updateSweepTime();
  void sweep(int precision, bool burnin);
  void boundaryCondition();
  void solveVectorPotential(const double *dCurrent, double *dOut);
  // G~ for the vector-potential block: A_new <- A + P (A_new - A) with
  // P = I - K (-L + K)^-1 (the Meissner response preconditioner), when the
  // config enables it
  void screenVectorPotential();

  Results results() const { return res_; }
  Timing timing() const { return tim_; }
  // FP32-stepping iterations of the last run()
  int mixedIterations() const { return mixedIterations_; }
  // A batch that refines grain by grain counts their FP32 iterations here.
  void setMixedIterations(int n) { mixedIterations_ = n; }
  // Lane layout of the sweep for one precision: Ozaki energies per lane,
  // register budget (wavefronts per SIMD the kernel is compiled for) and
  // persistent wavefronts per compute unit.
  void setSweepShape(int precision, int energiesPerLane, int wavesPerSimd, int slotsPerCU);
  void sweepShape(int precision, int &energiesPerLane, int &wavesPerSimd,
                  int &slotsPerCU) const;
  // Times every compiled sweep variant on the current state and keeps the
  // fastest for `precision`; the sweeps it runs leave the state unchanged.
  void autotune(int precision);
  // Jacobian-free Newton-Krylov solve of the fixed point of one iteration
  // (state: order parameter, scaled vector potential, boundary in-values)
  // to a block-weighted residual below tol; leaves the solution as the
  // iterate and restarts the accelerators.
  NewtonReport newtonKrylov(int maxNewton, int krylovDim, double tol,
                            int jvpPrecision = PREC_FP64);
  // The free energy one iteration reports from the current state and its
  // gradient with respect to the order parameter (vector potential and
  // boundary in-values held fixed) by a discrete adjoint: grad holds
  // Nc x 2 x N0^2 values, dF/dRe Delta_c then dF/dIm Delta_c.  With gradA
  // (2 x N0^2: x then y) also dF/dA of the induced vector potential, per
  // unit of A as setVectorPotential takes it (zero without a penetration
  // depth).  The state is left as it was.
  double freeEnergyGradient(double *grad, double *gradA = nullptr);
  // The self-consistency residual R = 1/2 sum_inside |Delta' - Delta|^2 +
  // 1/2 sum |A' - A|^2 of one plain FP64 iteration u -> G(u) (boundary
  // in-values held fixed) and its gradient (J - I)^T (G(u) - u), grad and
  // gradA laid out as freeEnergyGradient's.  The state is left as it was.
  double residualGradient(double *grad, double *gradA = nullptr);
  // The Jacobian-vector product J v of the Newton-Krylov map (one plain FP64
  // iteration: gap equation, vector potential, with inValues also the new
  // in-values of the boundary condition; with screened the screened
  // vector-potential update as nkMapEnd applies it) at the current state, by
  // the tangent-linear sweep.  dv and djv are device vectors in the map's
  // layout ([Delta planes, lambda^2 A planes when the vector potential is
  // solved, with inValues the gamma then gamma~ in-values]).
  void jacobianVector(const double *dv, double *djv, bool inValues, bool screened);
  // The same on host vectors [Delta planes (Nc x 2 x N0^2), A planes
  // (2 x N0^2, in the setters' units)], in-values held fixed, unscreened.
  void jacobianVectorHost(const double *v, double *jv);
  // The reverse sweeps' storage of the coherence functions: 0 keeps every
  // site, k > 0 keeps checkpoints every k sites and integrates each segment
  // again when the reverse passes reach it, -1 (the default) picks
  // k = ceil(sqrt(longest trajectory)).  Results are the same either way.
  void setAdjointCheckpoints(int k);
  // Newton-Krylov's GMRES cycles per Newton step (each of at most
  // krylovDim directions; default 4), the cap of its Eisenstat-Walker
  // forcing term (the linear residual a Newton step asks for, relative to
  // the nonlinear one; default 0.3), and whether FP32 products hand over to
  // FP64 ones when a line search finds no descent (default) or the solve
  // ends there.
  void setNewtonOptions(int restarts, double etaMax, bool fallback = true, bool fixPhase = false);
  void resetAccelerator();

  // Keldysh spectroscopy (src/postprocess.cu), for a grain built by
  // buildSpectroscopySetup: the LDOS at real energies of the current order
  // parameter and vector potential.
  bool keldysh() const { return s_->keldysh; }
  // Context::compute / burnIn of the postprocess context
  // (IntegrationIteratorKeldysh.h:235-388, ComputeLDOS.h:110-192).
  void ldosCompute(int precision, bool burnin);
  // runPostprocess (run_simulation.h:44-108); true when every block converged.
  bool ldosRun(int burnin, int itMin, int itMax, int precision);
  void getLdos(double *out); // [numEnergies][N0][N0]
  LdosStatus ldosStatus() const;
  // Angles whose spectral weights are held at once (0: from free memory).
  void setLdosAngleChunk(int angles);
  // Boundary initial values (first call) and the extended order parameter:
  // the inputs of a sweep.
  void prepareSweepInputs() {
    ensureBoundaryInit();
    extendDelta();
  }
  hipStream_t stream() const { return stream_; }

  // Device view of the grain, and with the lane layout of `el` energies
  // per lane (for batched sweeps).
  GrainView view() const;
  GrainView sweepView(int el) const;

  // device pointers for zero-copy callers
  double *devDelta() { return dDelta_; }
  double *devVectorPotential() { return dA_; }

private:
  void upload();
  void uploadEnergies();
  // Keldysh helpers (sc_ldos.hip)
  void ldosBeginBlock();
  int ldosChunkAngles() const;
  void ldosEnsureBuffers(int precision);
  // extendValuesAtBoundary on `planes` values per site of frames [k0, k1)
  // whose data start at base (frame k at base + (k - k0) * blockStride).
  void extendFrames(double *base, int k0, int k1, int planes, long long planeStride,
                    long long blockStride, long long siteScale);
  // Site fields of frames [k0, k0 + frames) into F (plane stride `plane`).
  void computeFields(int k0, int frames, long long plane, double *F);
  // Keldysh sweeps chunk by chunk: site fields of the chunk, then its items.
  void ldosSweepChunk(int precision, int chunk, bool burnin);
  void ldosSweepAll(int precision, bool burnin);
  void toDevice(void *dst, const void *src, size_t bytes);
  void fromDevice(void *dst, const void *src, size_t bytes);
  void extendDelta(bool fields = true);
  void ensureBoundaryInit();
  void afterSweep(bool burnin);
  void backRotate();
  void finishIntegrate();
  void acceleratorStep();
  void acceleratorQueue();  // device work up to the read-back
  void acceleratorFinish(); // the update, once the reductions are back
  double residualOf(const double *oldv, const double *newv, long long n);
  double deviceSum(const double *a, const double *b, const uint8_t *mask,
                   long long n, int op);
  // deferred reductions: queue now, read back together at flushReductions()
  double *reserveReduction(int kind, long long blocks, long long n, int &handle);
  int queueSum(const double *a, const double *b, const uint8_t *mask, long long n, int op);
  int queueStats(const double *oldv, const double *newv, long long n);
  void flushReductions();
  void clearReductions();
  double reductionSum(int handle) const;
  double reductionResidual(int handle) const;
  void queueBoundaryCondition(int handles[2]);
  double boundaryResidual(const int handles[2]) const;
  void prepareSweep(int precision);
  std::vector<double> multiDot(const double *a, const std::vector<double *> &bs,
                               long long n);
  // Several multiDot groups (a, bs, n) with one read-back for all of them.
  struct DotGroup {
    const double *a;
    std::vector<double *> bs;
    long long n;
  };
  std::vector<std::vector<double>> multiDotGroups(const std::vector<DotGroup> &groups);
  // drops the queued reductions without reading them back
  void discardReductions();

  std::unique_ptr<Setup> s_;
  hipStream_t stream_ = nullptr;
  bool ownsStream_ = true;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr, ev2_ = nullptr;
  hipEvent_t evIterStart_ = nullptr, evIterStop_ = nullptr;
  // constant tables
  double *dEnergy_ = nullptr, *dResidue_ = nullptr, *dSaCos_ = nullptr,
         *dSaSin_ = nullptr, *dVFx_ = nullptr, *dVFy_ = nullptr,
         *dWeight_ = nullptr, *dBasis_ = nullptr, *dCoupling_ = nullptr;
  float *dTrig_ = nullptr; // opCos, opSin, aCos, aSin, bkCos, bkSin
  int8_t *dLabels_ = nullptr;
  uint8_t *dDomain_ = nullptr;
  int *dDomainSites_ = nullptr;
  int *dIdx_ = nullptr; // idxInG, idxOutG, idxInGt, idxOutGt
  int *dOutCross_ = nullptr; // colOutG, coordG, colOutGt, coordGt
  int *dFieldTiles_ = nullptr;
  double *dDist_ = nullptr;
  int *dBcOffG_ = nullptr, *dBcSrcG_ = nullptr, *dBcOffGt_ = nullptr,
      *dBcSrcGt_ = nullptr;
  double *dBcWG_ = nullptr, *dBcWsG_ = nullptr, *dBcWGt_ = nullptr,
         *dBcWsGt_ = nullptr;
  int *dExtRefT_ = nullptr, *dExtRefO_ = nullptr, *dExtRefS_ = nullptr;
  int *dExtRotT_ = nullptr, *dExtRotO_ = nullptr, *dExtRotS_ = nullptr;
  // state
  double *dDelta_ = nullptr, *dDeltaNew_ = nullptr, *dExtDelta_ = nullptr;
  double *dA_ = nullptr, *dANew_ = nullptr, *dCurrent_ = nullptr,
         *dCurrentOld_ = nullptr, *dEnergyDens_ = nullptr;
  double *dBInG_ = nullptr, *dBOutG_ = nullptr, *dBInGt_ = nullptr,
         *dBOutGt_ = nullptr;
  double *dG_ = nullptr;
  double *dFields_ = nullptr; // site fields of every frame, [3][Na][N1^2]
  double *dPartials_ = nullptr, *dWork_ = nullptr, *dWork2_ = nullptr;
  std::unique_ptr<PoissonFFT> fft_;
  std::unique_ptr<PoissonFFT> screen_; // (-L + K)^-1 of the preconditioner
  double screenK_ = 0.0;               // K = superfluid density / lambda^2
  double *dZero2_ = nullptr;           // two zeros: the screened solve's means
  // Keldysh state
  double *dEnergyRe_ = nullptr;
  double *dLdosAll_ = nullptr, *dLdosAcc_ = nullptr, *dLdosInterval_ = nullptr;
  double *dLdosG_ = nullptr;
  size_t ldosGCap_ = 0;
  double *dLdosFields_ = nullptr; // site fields of one angle chunk
  int ldosFieldFrames_ = 0;
  int ldosChunkReq_ = 0;
  int ldosBlockStart_ = 0, ldosLastBlock_ = -1, ldosIterations_ = 0;
  double ldosRes_ = 0, ldosSweepMs_ = 0, ldosIterMs_ = 0;
  bool ldosPartial_ = false, ldosConverged_ = false;
  std::vector<int> extRotFrameStart_; // first extRot entry of each frame (+ end)
  std::vector<hipEvent_t> ldosEvents_;
  // Newton-Krylov state
  struct NkBlock {
    size_t offset, length;
    // the block's charge under the global phase rotation of the order
    // parameter (planar complex values: +1 rotate with it, -1 against it,
    // 0 invariant)
    int phase;
  };
  std::vector<NkBlock> nkBlocks() const;
  size_t nkLength() const;
  void nkGet(double *u);
  void nkSet(const double *u);
  // one evaluation of the map G of the state: nkMapBegin, the sweep at the
  // chosen precision (the grain's own or a batch's), nkMapEnd(G)
  void nkMapBegin(bool fields = true);
  void nkMapEnd(double *gu);
  friend class NewtonKrylov;
  friend class Batch;
  std::vector<double *> nkVecs_;
  size_t nkLen_ = 0;
  int nkMaps_ = 0;
  struct Pending {
    int kind = 0; // 0: sum (one partial per block), 1: change statistics (six)
    long long blocks = 0, n = 0;
    size_t offset = 0;
    int handle = -1;
  };
  std::vector<Pending> pend_;
  std::vector<std::array<double, 6>> redOut_;
  std::vector<long long> redN_;
  double *dotPinned_ = nullptr; // pinned host copy of dPartials_ (multiDotGroups)
  size_t dotPinnedCap_ = 0;
  hipEvent_t evDots_ = nullptr;
  double *redPinned_ = nullptr; // pinned host copy of dRed_
  size_t redPinnedCap_ = 0;
  bool readbackQueued_ = false;
  hipEvent_t evReadback_ = nullptr;
  double *dRed_ = nullptr;
  size_t redCap_ = 0, redUsed_ = 0;
  size_t redReserved_ = 0; // elements reserved since construction (never reset)
  double *dMean_ = nullptr; // grain means of the current (Poisson solve)
  // handles of the reductions queued by integrate()
  std::vector<int> hDelta_, hDeltaSq_;
  int hCurrent_ = -1, hEnergy_ = -1, hCurl_ = -1, hBoundary_[2] = {-1, -1};
  bool pendingIntegrate_ = false;
  // angle-by-angle integration: running sums per inside site, the next
  // angle to integrate, and whether the current iteration runs that way
  double *dGatherAcc_ = nullptr;
  int nextAngle_ = 0;
  bool rangeMode_ = false, rangeBurnin_ = false;
  // the free-energy gradient (sc_adjoint.hip)
  void queueExplicitFreeEnergy(std::vector<int> &hSq, int &hCurl);
  void *adjItems_ = nullptr;
  int adjNumItems_ = 0, adjMaxLen_ = 0, adjSlots_ = 0;
  double *dAdjScratch_ = nullptr, *dAdjPartial_ = nullptr, *dAdjRho_ = nullptr,
         *dAdjDbar_ = nullptr, *dAdjExt_ = nullptr, *dAdjGrad_ = nullptr,
         *dAdjGradA_ = nullptr, *dAdjPart_ = nullptr;
  // the residual gradient's buffers: rD, rA, the Poisson seed and its
  // transpose, jbar, two zero means, two means, the frames' Green-sum seeds,
  // the saved state
  double *dResD_ = nullptr, *dResA_ = nullptr, *dResPois_ = nullptr, *dResSbar_ = nullptr,
         *dResJbar_ = nullptr, *dResZero_ = nullptr, *dResMean_ = nullptr, *dResSeed_ = nullptr,
         *dResSave_ = nullptr;
  void prepareAdjoint();
  int adjCk_ = -1, adjCkUsed_ = 0, adjCkSeg_ = 0;
  bool adjRhoReady_ = false; // dAdjRho_ holds the free-energy seed
  int nkRestarts_ = 4;
  double nkEtaMax_ = 0.3;
  bool nkFallback_ = true;
  bool nkFixPhase_ = false;
  long long adjSlotStride_ = 0;
  double *dTanScratch_ = nullptr; // the tangent sweep's own (every site kept)
  // the tangent-linear products' buffers
  double *dTanExt_ = nullptr, *dTanFields_ = nullptr, *dTanG_ = nullptr, *dTanJ_ = nullptr,
         *dTanE_ = nullptr, *dTanOutG_ = nullptr, *dTanOutGt_ = nullptr, *dTanV_ = nullptr,
         *dTanJV_ = nullptr;
  int hA_ = -1; // queued residual of the vector potential
  // the captured stages after the sweep (queueEndCompute)
  void queueEndStages();
  void dropEndGraph();
  bool graphsOn_ = true;
  bool endSeen_ = false;           // the stages ran eagerly once (sizes known)
  hipGraphExec_t endExec_ = nullptr;
  struct EndCapture {
    size_t redOutBefore = 0, redUsedBefore = 0, redUsedAfter = 0, need = 0;
    bool rangeMode = false;
    std::vector<Pending> pend;      // the reservations the stages make
    std::vector<long long> redN;    // and the element counts of their handles
    std::vector<int> hDelta, hDeltaSq;
    int hCurrent = -1, hEnergy = -1, hCurl = -1, hBoundary[2] = {-1, -1}, hA = -1;
  } endCap_;
  long long partialsCap_ = 0;
  // sweep plans, one per precision
  struct SweepPlan {
    int el = 8, waves = 2, slotsPerCU = 8;
    int cols = 0, groups = 0, maxLen = 0, numItems = 0, numSlots = 0;
    // Keldysh: items grouped by angle chunk (chunkAngles frames each),
    // chunk c's items at [chunkStart[c], chunkStart[c + 1])
    int chunkAngles = 0;
    std::vector<int> chunkStart;
    void *dItems = nullptr;
    // the same items ordered by angle (angle k's at [angleStart[k],
    // angleStart[k + 1])), for sweeps of angle ranges
    std::vector<int> angleStart;
    void *dItemsByAngle = nullptr;
    void *dView = nullptr; // device copy of the grain view with this lane layout
    bool ready = false;
  };
  SweepPlan plan_[2];
  void *dScratch_ = nullptr;
  size_t scratchBytes_ = 0;
  int *dCounter_ = nullptr;
  // accelerator state (history lives on the device)
  int accIter_ = 0;
  std::vector<double *> accX_, accD_;
  std::vector<double> accSq_;
  std::vector<double> accGram_; // rank x rank dot products of the differences
  std::vector<long long> accIdx_;
  double *dAccMean_ = nullptr, *dAccPrev_ = nullptr, *dAccVel_ = nullptr;
  std::vector<double> accStep_, accInertia_;
  std::vector<int> accH_; // reductions queued by acceleratorQueue
  double bbStep_ = 0;
  bool accInit_ = false;
  // Broyden: c_i in accX_, d_i in accD_ (the first brCount_ in use), the
  // previous iterate and residual, and the iterate difference
  int brCount_ = 0;
  double *dBrLastX_ = nullptr, *dBrLastF_ = nullptr, *dBrDx_ = nullptr;
  void broydenStep(double *X, const double *G, long long n);
  // bookkeeping
  bool boundaryInit_ = false, firstCompute_ = true, firstCurrent_ = true,
       firstFE_ = true;
  double fePerAreaOld_ = 0;
  int mixedIterations_ = 0;
  Results res_;
  Timing tim_;
  bool convDelta_ = false, convCurrent_ = false, convFE_ = false;
};

// Stand-alone ComputeVectorPotential solve with host buffers (2*n*n each);
// maxError > 0 uses the eigen-truncated kernel.
void poissonSolve(int n, double h, int numGrain, bool enforceMean,
                  const double *current, double *out, double maxError = 0.0);

// Many grains iterated together (a parameter sweep).  Every iteration runs
// the per-grain stages on one stream and sweeps all grains that are still
// iterating in one persistent launch whose work queue holds the items of all
// of them, longest first, so that small grains fill the gaps of large ones.
// One grain's Jacobian-free Newton-Krylov solve (sc_newton.hip), advanced
// one map evaluation at a time: next() writes the point to evaluate into the
// grain's state, starts its map (nkMapBegin) and names the sweep precision;
// the caller sweeps and calls grain.nkMapEnd(out()); advance() continues the
// algorithm.  Grain::newtonKrylov drives one; Batch::newtonKrylov drives one
// per grain with one sweep launch per precision for all of them.
class NewtonKrylov {
public:
  NewtonKrylov(Grain &g, int maxNewton, int krylovDim, double tol, int jvpPrecision);
  bool next(int &precision, bool fields = true);
  double *out() const { return out_; }
  void advance();
  bool done() const { return phase_ == DONE; }
  NewtonReport finish();

private:
  enum Phase { INIT, BASE, KRYLOV, LINE, DONE };
  std::vector<double> wdots(const double *a, const std::vector<double *> &bs);
  double wnorm(double *a);
  void combine(double *y, const double *base, double a, const std::vector<const double *> &vs,
               const std::vector<double> &cs);
  void want(const double *x, double *out, int precision);
  void startNewton();
  void startCycle();
  void startProduct();
  void productDone();
  void endCycle();
  void startLineSearch();
  void lineDone();
  void removePhase();

  Grain &g_;
  int maxNewton_, m_;
  double tol_;
  int jvp_;
  std::vector<Grain::NkBlock> blocks_;
  std::vector<double> wt_;
  long long nL_ = 0;
  double *u_, *gu_, *f_, *trial_, *w_, *dstep_, *rlin_, *guJ_, *gen_;
  std::vector<double *> V_;
  Phase phase_ = INIT;
  const double *x_ = nullptr;
  double *out_ = nullptr;
  int prec_ = PREC_FP64;
  NewtonReport rep_;
  double fn_ = 0, eta_ = 0.1, e_ = 0, rn_ = 0, beta_ = 0, lambda_ = 1;
  int cycle_ = 0, k_ = 0, ls_ = 0, stalled_ = 0;
  bool breakdown_ = false;
  std::vector<double> H_, R_, gv_, cs_, sn_;
};

class Batch {
public:
  explicit Batch(std::vector<std::unique_ptr<Setup>> setups);
  ~Batch();
  Batch(const Batch &) = delete;
  Batch &operator=(const Batch &) = delete;

  int size() const { return static_cast<int>(grains_.size()); }
  Grain &grain(int i) { return *grains_.at(size_t(i)); }
  // One iteration of every grain that is still iterating.
  void compute(int precision);
  // The same with a precision per grain (PREC_FP64 or PREC_MIXED, indexed
  // by grain): one sweep launch per precision that has grains.
  void computeEach(const std::vector<int> &precisionOf);
  void burnin(int precision);
  // runSimulation for every grain; a grain stops once converged with the
  // minimum reached, or at the maximum.  With PREC_MIXED_REFINE every grain
  // starts with FP32 stepping and switches to FP64 on its own (Grain::run's
  // rule); it only counts as converged on an FP64 iteration.  Returns the
  // number converged.
  int run(int nBurnin, int itMin, int itMax, int precision, std::vector<int> &converged);
  void setSweepShape(int precision, int el, int waves, int slotsPerCU);
  void sweepShape(int precision, int &el, int &waves, int &slotsPerCU) const;
  void autotune(int precision);
  double lastSweepMs() const { return sweepMs_; }
  double lastIterationMs() const { return iterMs_; }
  long long lastTrajSteps() const { return trajSteps_; }
  std::vector<int> active() const;
  // Newton-Krylov on many grains at once (Grain::newtonKrylov per grain),
  // the map evaluations of all solving grains in one sweep launch per
  // precision; one report per grain.
  // `which` (one flag per grain, empty: all) selects the grains to solve;
  // the others keep their state and report nothing.
  std::vector<NewtonReport> newtonKrylov(int maxNewton, int krylovDim, double tol,
                                         int jvpPrecision, const std::vector<int> &which = {});

private:
  struct Plan {
    int el = 8, waves = 2, slotsPerCU = 8;
    std::vector<int> key; // the active grains the items were built for
    int numItems = 0, numSlots = 0, maxLen = 0;
    void *dItems = nullptr, *dViews = nullptr;
    size_t itemsCap = 0; // items dItems can hold
    bool ready = false;
    // each grain's items at this layout (built once per layout)
    int cacheEl = -1;
    std::vector<std::vector<SweepItem>> itemsOf;
    std::vector<int> maxLenOf;
    // the site-field jobs of the active grains: grain index and first tile
    // (numJobs + 1 prefix entries), for fieldsBatchKernel
    int *dFieldJobs = nullptr;
    int numFieldJobs = 0, numFieldTiles = 0;
  };
  void prepare(int precision, const std::vector<int> &act);
  // fields: compute the site fields of the active grains (one launch) first
  void sweep(int precision, bool burnin, const std::vector<int> &act, bool fields = false);
  // elapsed time of the last sweep launch of a precision
  float sweepTime(int precision) const;
  // the main stream waits for every lane, every lane for the main stream
  void joinLanes();
  void forkLanes();

  // The sweep runs on the main stream; the per-grain stages run on a few
  // lanes (grain i on lane i % lanes) so that the small kernels of
  // different grains overlap.
  hipStream_t stream_ = nullptr;
  std::vector<hipStream_t> lanes_;
  std::vector<hipEvent_t> laneEv_;
  hipEvent_t evFork_ = nullptr;
  std::vector<std::unique_ptr<Grain>> grains_;
  std::vector<char> done_;
  Plan plan_[2];
  void *dScratch_ = nullptr;
  size_t scratchBytes_ = 0;
  int *dCounter_ = nullptr;
  // start and end of the last sweep launch per precision; the iteration
  hipEvent_t ev0_[2] = {nullptr, nullptr}, ev1_[2] = {nullptr, nullptr};
  hipEvent_t evIt0_ = nullptr, evIt1_ = nullptr;
  double sweepMs_ = 0, iterMs_ = 0;
  long long trajSteps_ = 0;
};

// One launch of the compiled sweep variant (precision, energies per lane,
// wavefronts per SIMD) over the items of `args`.
void launchSweepVariant(int precision, int el, int waves, const SweepArgs &args, int slots,
                        hipStream_t stream, bool keldysh = false);
bool sweepVariantExists(int el, int waves, bool keldysh = false);

} // namespace scamd
