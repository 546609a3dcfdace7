// sc_engine.hip -- device state of one grain and the self-consistency
// iteration (Context::compute / Context::burnIn, Context.h:377-447) on one
// MI355X.  All buffers stay resident in HBM between iterations; the host only
// reads back scalars (residuals, the accelerator's Gram matrix).

#include "sc_engine.hpp"
#include "sc_util.hpp"
#include "sc_kernels.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <stdexcept>

namespace scamd {


Grain::Grain(std::unique_ptr<Setup> setup, hipStream_t stream) : s_(std::move(setup)) {
  if (stream) {
    stream_ = stream;
    ownsStream_ = false;
  } else {
    SC_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  }
  SC_HIP(hipEventCreate(&ev0_));
  SC_HIP(hipEventCreate(&ev1_));
  SC_HIP(hipEventCreate(&ev2_));
  SC_HIP(hipEventCreate(&evIterStart_));
  SC_HIP(hipEventCreate(&evIterStop_));
  SC_HIP(hipEventCreateWithFlags(&evReadback_, hipEventDisableTiming));
  SC_HIP(hipEventCreateWithFlags(&evDots_, hipEventDisableTiming));
  upload();
  // Allocation clears and uploads ran on the null stream, which a
  // non-blocking stream does not wait for.
  SC_HIP(hipDeviceSynchronize());
  tim_.trajSteps = s_->trajSteps;
}

Grain::~Grain() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  for (double *p : accX_) devFree(p);
  for (double *p : accD_) devFree(p);
  devFree(dBrLastX_);
  devFree(dBrLastF_);
  devFree(dBrDx_);
  devFree(dAccMean_);
  devFree(dAccPrev_);
  devFree(dAccVel_);
  devFree(dEnergy_);
  devFree(dResidue_);
  devFree(dEnergyRe_);
  devFree(dLdosAll_);
  devFree(dLdosAcc_);
  devFree(dLdosInterval_);
  devFree(dLdosG_);
  for (hipEvent_t e : ldosEvents_) (void)hipEventDestroy(e);
  devFree(dSaCos_);
  devFree(dSaSin_);
  devFree(dVFx_);
  devFree(dVFy_);
  devFree(dWeight_);
  devFree(dBasis_);
  devFree(dCoupling_);
  devFree(dTrig_);
  devFree(dLabels_);
  devFree(dDomain_);
  devFree(dDomainSites_);
  devFree(dIdx_);
  devFree(dOutCross_);
  devFree(dFieldTiles_);
  devFree(dDist_);
  devFree(dBcOffG_);
  devFree(dBcSrcG_);
  devFree(dBcOffGt_);
  devFree(dBcSrcGt_);
  devFree(dBcWG_);
  devFree(dBcWsG_);
  devFree(dBcWGt_);
  devFree(dBcWsGt_);
  devFree(dExtRefT_);
  devFree(dExtRefO_);
  devFree(dExtRefS_);
  devFree(dExtRotT_);
  devFree(dExtRotO_);
  devFree(dExtRotS_);
  devFree(dDelta_); // also owns dA_
  devFree(dDeltaNew_); // also owns dANew_
  devFree(dExtDelta_);
  devFree(dCurrent_);
  devFree(dCurrentOld_);
  devFree(dEnergyDens_);
  devFree(dBInG_);
  devFree(dBOutG_);
  devFree(dBInGt_);
  devFree(dBOutGt_);
  devFree(dG_);
  devFree(dFields_);
  devFree(dLdosFields_);
  devFree(dPartials_);
  devFree(dRed_);
  devFree(dMean_);
  for (double *p : nkVecs_) devFree(p);
  devFree(dWork_);
  devFree(dWork2_);
  if (adjItems_) (void)hipFree(adjItems_);
  devFree(dAdjScratch_);
  devFree(dAdjPartial_);
  devFree(dAdjRho_);
  devFree(dAdjDbar_);
  devFree(dAdjExt_);
  devFree(dAdjGrad_);
  devFree(dAdjGradA_);
  devFree(dAdjPart_);
  for (double *p : {dResD_, dResA_, dResPois_, dResSbar_, dResJbar_, dResZero_, dResMean_,
                    dResSeed_, dResSave_, dTanExt_, dTanFields_, dTanG_, dTanJ_, dTanE_,
                    dTanOutG_, dTanOutGt_, dTanV_, dTanJV_, dTanScratch_})
    devFree(p);
  dropEndGraph();
  fft_.reset();
  screen_.reset();
  devFree(dZero2_);
  if (dScratch_) (void)hipFree(dScratch_);
  for (SweepPlan &pl : plan_) {
    if (pl.dItems) (void)hipFree(pl.dItems);
    if (pl.dItemsByAngle) (void)hipFree(pl.dItemsByAngle);
    if (pl.dView) (void)hipFree(pl.dView);
  }
  devFree(dGatherAcc_);
  devFree(dCounter_);
  if (ev0_) (void)hipEventDestroy(ev0_);
  if (ev1_) (void)hipEventDestroy(ev1_);
  if (ev2_) (void)hipEventDestroy(ev2_);
  if (evIterStart_) (void)hipEventDestroy(evIterStart_);
  if (evIterStop_) (void)hipEventDestroy(evIterStop_);
  if (evReadback_) (void)hipEventDestroy(evReadback_);
  if (evDots_) (void)hipEventDestroy(evDots_);
  if (redPinned_) (void)hipHostFree(redPinned_);
  if (dotPinned_) (void)hipHostFree(dotPinned_);
  if (stream_ && ownsStream_) (void)hipStreamDestroy(stream_);
}

void Grain::upload() {
  const Setup &s = *s_;
  const size_t n0 = size_t(s.N0) * s.N0, n1 = size_t(s.N1) * s.N1;
  {
    // Keldysh grains change their energies block by block: room for the
    // largest block
    const size_t cap = size_t(s.keldysh ? s.spec.perBlock : s.Ne);
    dEnergy_ = devAlloc<double>(cap);
    dResidue_ = devAlloc<double>(cap);
    if (s.keldysh) dEnergyRe_ = devAlloc<double>(cap);
    uploadEnergies();
  }
  dSaCos_ = devUpload(s.saCos);
  dSaSin_ = devUpload(s.saSin);
  dVFx_ = devUpload(s.vFx);
  dVFy_ = devUpload(s.vFy);
  dWeight_ = devUpload(s.weight);
  dBasis_ = devUpload(s.basis);
  dCoupling_ = devUpload(s.coupling);
  std::vector<float> trig;
  for (const auto *v : {&s.opCos, &s.opSin, &s.aCos, &s.aSin, &s.bkCos, &s.bkSin})
    trig.insert(trig.end(), v->begin(), v->end());
  dTrig_ = devUpload(trig);
  dLabels_ = devUpload(s.labels);
  dDomain_ = devUpload(s.domain);
  {
    // The gather visits the grain sites in 16 x 16 tiles so that the
    // threads of a block sample a compact rotated patch of every frame and
    // share its bilinear taps in the caches.
    std::vector<int> tiled(s.domainSites);
    const int N0 = s.N0;
    std::stable_sort(tiled.begin(), tiled.end(), [N0](int a, int b) {
      const int ax = a % N0, ay = a / N0, bx = b % N0, by = b / N0;
      const long long ka = (long long)(ay / 16) * N0 + ax / 16;
      const long long kb = (long long)(by / 16) * N0 + bx / 16;
      return ka < kb;
    });
    dDomainSites_ = devUpload(tiled);
  }
  std::vector<int> idx;
  for (const auto *v : {&s.idxInG, &s.idxOutG, &s.idxInGt, &s.idxOutGt})
    idx.insert(idx.end(), v->begin(), v->end());
  dIdx_ = devUpload(idx);
  {
    std::vector<int> oc;
    for (const auto *v : {&s.colOutG, &s.coordG, &s.colOutGt, &s.coordGt})
      oc.insert(oc.end(), v->begin(), v->end());
    dOutCross_ = devUpload(oc);
  }
  dFieldTiles_ = devUpload(s.fieldTiles);
  std::vector<double> dist;
  for (const auto *v : {&s.distInG, &s.distOutG, &s.distInGt, &s.distOutGt})
    dist.insert(dist.end(), v->begin(), v->end());
  dDist_ = devUpload(dist);
  dBcOffG_ = devUpload(s.bcG.offset);
  dBcSrcG_ = devUpload(s.bcG.src);
  dBcWG_ = devUpload(s.bcG.w);
  dBcWsG_ = devUpload(s.bcG.wsum);
  dBcOffGt_ = devUpload(s.bcGt.offset);
  dBcSrcGt_ = devUpload(s.bcGt.src);
  dBcWGt_ = devUpload(s.bcGt.w);
  dBcWsGt_ = devUpload(s.bcGt.wsum);
  dExtRefT_ = devUpload(s.extRef.target);
  dExtRefO_ = devUpload(s.extRef.offset);
  dExtRefS_ = devUpload(s.extRef.source);
  dExtRotT_ = devUpload(s.extRot.target);
  dExtRotO_ = devUpload(s.extRot.offset);
  dExtRotS_ = devUpload(s.extRot.source);
  // state: order parameter and vector potential are contiguous so that the
  // accelerators see the concatenated field [Re D_0, Im D_0, ..., Ax, Ay]
  // (accelerators/utils.h:58-186) without copies.
  const size_t nf = size_t(2 * s.Nc + 2) * n0;
  dDelta_ = devAlloc<double>(nf);
  dA_ = dDelta_ + size_t(2 * s.Nc) * n0;
  dDeltaNew_ = devAlloc<double>(nf);
  dANew_ = dDeltaNew_ + size_t(2 * s.Nc) * n0;
  dExtDelta_ = devAlloc<double>(size_t(2 * s.Nc) * n0);
  dCurrent_ = devAlloc<double>(2 * n0);
  dCurrentOld_ = devAlloc<double>(2 * n0);
  dEnergyDens_ = devAlloc<double>(n0);
  const long long neCap = s.keldysh ? s.spec.perBlock : s.Ne;
  dBInG_ = devAlloc<double>(size_t(2 * s.nInG * neCap));
  dBOutG_ = devAlloc<double>(size_t(2 * s.nOutG * neCap));
  dBInGt_ = devAlloc<double>(size_t(2 * s.nInGt * neCap));
  dBOutGt_ = devAlloc<double>(size_t(2 * s.nOutGt * neCap));
  // Keldysh grains hold spectral weights of an angle chunk instead
  if (!s.keldysh) {
    dG_ = devAlloc<double>(size_t(s.Na) * 4 * n1);
    dFields_ = devAlloc<double>(size_t(3) * s.Na * n1);
  }
  {
    // first extension entry of every frame (the list is built frame by frame)
    extRotFrameStart_.assign(size_t(s.Na) + 1, 0);
    const long long nt = static_cast<long long>(s.extRot.target.size());
    long long t = 0;
    for (int k = 0; k < s.Na; ++k) {
      extRotFrameStart_[size_t(k)] = static_cast<int>(t);
      while (t < nt && static_cast<long long>(s.extRot.target[size_t(t)]) / static_cast<long long>(n1) == k)
        ++t;
    }
    extRotFrameStart_[size_t(s.Na)] = static_cast<int>(nt);
    if (t != nt) throw std::runtime_error("extension list is not ordered by frame");
  }
  redCap_ = size_t(64) * 1024 * 6;
  dRed_ = devAlloc<double>(redCap_);
  dMean_ = devAlloc<double>(2);
  partialsCap_ = 4096;
  dPartials_ = devAlloc<double>(size_t(partialsCap_) * 6);
  dWork_ = devAlloc<double>(nf);
  dWork2_ = devAlloc<double>(nf);
  if (s.lambda > 0.0) fft_ = std::make_unique<PoissonFFT>(s.N0, s.h, stream_, s.cfg.vpError);
  if (s.lambda > 0.0 && s.cfg.computeCurrent && s.cfg.vpScreening > 0.0) {
    // j = -rho_s A in the bulk (A unscaled), and the iterate holds lambda^2 A
    screenK_ = s.cfg.vpScreening / (s.lambda * s.lambda);
    screen_ = std::make_unique<PoissonFFT>(s.N0, s.h, stream_, PoissonFFT::Screened{screenK_});
    dZero2_ = devAlloc<double>(2);
    SC_HIP(hipMemsetAsync(dZero2_, 0, 2 * sizeof(double), stream_));
  }
  dCounter_ = devAlloc<int>(1);
}

void Grain::uploadEnergies() {
  const Setup &s = *s_;
  const size_t b = size_t(s.Ne) * sizeof(double);
  SC_HIP(hipMemcpy(dEnergy_, s.energy.data(), b, hipMemcpyHostToDevice));
  SC_HIP(hipMemcpy(dResidue_, s.residue.data(), b, hipMemcpyHostToDevice));
  if (dEnergyRe_) SC_HIP(hipMemcpy(dEnergyRe_, s.energyRe.data(), b, hipMemcpyHostToDevice));
}

// Compiled sweep variants: (energies per lane, wavefronts per SIMD), for
// Matsubara energies and for real-frequency (Keldysh) energies.
#define SC_SWEEP_VARIANTS(X)                                                         \
  X(1, 4) X(2, 4) X(3, 3) X(4, 2) X(4, 3) X(4, 4) X(5, 2) X(6, 2) X(7, 1) X(7, 2) X(7, 4)  \
  X(8, 1) X(8, 2) X(8, 4)                                                            \
  X(14, 1) X(14, 2) X(16, 1) X(16, 2)
#define SC_KELDYSH_VARIANTS(X) X(2, 4) X(4, 2) X(4, 3) X(4, 4) X(8, 1) X(8, 2) X(8, 4) X(16, 1)

bool sweepVariantExists(int el, int waves, bool keldysh) {
#define SC_HAS(E, W) if (el == E && waves == W) return true;
  if (keldysh) {
    SC_KELDYSH_VARIANTS(SC_HAS)
  } else {
    SC_SWEEP_VARIANTS(SC_HAS)
  }
#undef SC_HAS
  return false;
}

void Grain::setSweepShape(int precision, int el, int waves, int slotsPerCU) {
  if (precision != PREC_FP64 && precision != PREC_MIXED)
    throw std::runtime_error("unknown precision");
  if (!sweepVariantExists(el, waves, s_->keldysh))
    throw std::runtime_error("no sweep variant with " + std::to_string(el) +
                             " energies per lane and " + std::to_string(waves) +
                             " wavefronts per SIMD");
  if (slotsPerCU < 1 || slotsPerCU > 32)
    throw std::runtime_error("slots per CU must be in [1, 32]");
  SweepPlan &pl = plan_[precision];
  if (el != pl.el || waves != pl.waves || slotsPerCU != pl.slotsPerCU) pl.ready = false;
  pl.el = el;
  pl.waves = waves;
  pl.slotsPerCU = slotsPerCU;
}

void Grain::sweepShape(int precision, int &el, int &waves, int &slotsPerCU) const {
  const SweepPlan &pl = plan_[precision == PREC_FP64 ? 0 : 1];
  el = pl.el;
  waves = pl.waves;
  slotsPerCU = pl.slotsPerCU;
}

void Grain::prepareSweep(int precision) {
  SweepPlan &pl = plan_[precision];
  if (pl.ready) return;
  // the buffers below may still be read by a sweep in flight
  SC_HIP(hipStreamSynchronize(stream_));
  const Setup &s = *s_;
  const GrainView gv = sweepView(pl.el);
  pl.groups = gv.groups;
  pl.cols = gv.cols;
  std::vector<SweepItem> items = sweepItems(s, pl.cols, &pl.maxLen);
  pl.numItems = static_cast<int>(items.size());
  // Keldysh grains sweep by angle chunks: the items of a chunk are
  // contiguous (still longest first within it) and name the view whose
  // angle base is the chunk's first angle.
  std::vector<GrainView> views(1, gv);
  pl.chunkStart.assign(2, 0);
  pl.chunkStart[1] = pl.numItems;
  pl.chunkAngles = s.Na;
  if (s.keldysh) {
    int ca = ldosChunkAngles();
    if (dLdosG_) {
      // never more angles than the weight and field buffers hold
      const size_t perAngle = size_t(s.N1) * s.N1 * size_t(s.spec.perBlock);
      ca = std::max(1, std::min(ca, static_cast<int>(ldosGCap_ / perAngle)));
      ca = std::min(ca, ldosFieldFrames_);
    }
    const int nch = (s.Na + ca - 1) / ca;
    pl.chunkAngles = ca;
    std::stable_sort(items.begin(), items.end(), [ca](const SweepItem &a, const SweepItem &b) {
      return a.angle / ca < b.angle / ca;
    });
    pl.chunkStart.assign(size_t(nch) + 1, 0);
    for (const SweepItem &it : items) ++pl.chunkStart[size_t(it.angle / ca) + 1];
    for (int c = 0; c < nch; ++c) pl.chunkStart[size_t(c) + 1] += pl.chunkStart[size_t(c)];
    views.assign(size_t(nch), gv);
    for (int c = 0; c < nch; ++c) {
      views[size_t(c)].angleBase = c * ca;
      views[size_t(c)].fieldBase = c * ca;
    }
  }
  std::vector<SweepItemDev> di(items.size());
  for (size_t i = 0; i < items.size(); ++i)
    di[i] = {s.keldysh ? items[i].angle / pl.chunkAngles : 0, items[i].angle, items[i].x0,
             items[i].ncols, items[i].ylo, items[i].yhi};
  if (pl.dItems) (void)hipFree(pl.dItems);
  pl.dItems = nullptr;
  SC_HIP(hipMalloc(&pl.dItems, std::max<size_t>(1, di.size()) * sizeof(SweepItemDev)));
  if (!di.empty()) toDevice(pl.dItems, di.data(), di.size() * sizeof(SweepItemDev));
  if (pl.dItemsByAngle) (void)hipFree(pl.dItemsByAngle);
  pl.dItemsByAngle = nullptr;
  pl.angleStart.clear();
  if (!s.keldysh) {
    // the same items grouped by angle (longest first within an angle)
    std::vector<SweepItemDev> byAngle(di);
    std::stable_sort(byAngle.begin(), byAngle.end(),
                     [](const SweepItemDev &a, const SweepItemDev &b) { return a.angle < b.angle; });
    pl.angleStart.assign(size_t(s.Na) + 1, 0);
    for (const SweepItemDev &it : byAngle) ++pl.angleStart[size_t(it.angle) + 1];
    for (int k = 0; k < s.Na; ++k) pl.angleStart[size_t(k) + 1] += pl.angleStart[size_t(k)];
    SC_HIP(hipMalloc(&pl.dItemsByAngle,
                     std::max<size_t>(1, byAngle.size()) * sizeof(SweepItemDev)));
    if (!byAngle.empty())
      toDevice(pl.dItemsByAngle, byAngle.data(), byAngle.size() * sizeof(SweepItemDev));
  }
  if (pl.dView) (void)hipFree(pl.dView);
  pl.dView = nullptr;
  SC_HIP(hipMalloc(&pl.dView, views.size() * sizeof(GrainView)));
  toDevice(pl.dView, views.data(), views.size() * sizeof(GrainView));
  int dev = 0, cus = 0;
  SC_HIP(hipGetDevice(&dev));
  SC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  pl.numSlots = std::max(1, std::min(pl.numItems, pl.slotsPerCU * std::max(cus, 1)));
  const size_t elemBytes = precision == PREC_FP64 ? sizeof(Cx<double>) : sizeof(Cx<float>);
  const size_t stride = size_t(pl.maxLen) * pl.el * 64; // complex elements per slot
  const size_t bytes = size_t(pl.numSlots) * stride * elemBytes;
  if (bytes > scratchBytes_) {
    if (dScratch_) (void)hipFree(dScratch_);
    dScratch_ = nullptr;
    scratchBytes_ = 0;
    SC_HIP(hipMalloc(&dScratch_, bytes));
    scratchBytes_ = bytes;
  }
  pl.ready = true;
}

static GrainView makeView(const Setup &s, const float *trig, const int *idx,
                          const double *dist, double *bInG, double *bOutG,
                          double *bInGt, double *bOutGt, const double *extDelta,
                          const double *A, double *G, const double *energy,
                          const double *residue, const double *saCos,
                          const double *saSin, const double *vFx,
                          const double *vFy, const double *weight,
                          const double *basis, const int8_t *labels) {
  GrainView v{};
  v.N0 = s.N0;
  v.N1 = s.N1;
  v.Ne = s.Ne;
  v.Na = s.Na;
  v.Nc = s.Nc;
  v.h = s.h;
  v.useInduced = s.lambda > 0.0 ? 1 : 0;
  v.lambdaInv2 = s.lambda > 0.0 ? 1.0 / (s.lambda * s.lambda) : 0.0;
  v.area = s.h * s.h * static_cast<double>(s.numGrain);
  v.fluxDensity = M_PI * s.cfg.fluxQuanta / v.area;
  v.useField = (v.useInduced || std::fabs(v.fluxDensity) > 0.0) ? 1 : 0;
  v.gauge = s.cfg.gauge;
  v.charge = s.cfg.chargeSign;
  v.energy = energy;
  v.residue = residue;
  v.energyRe = nullptr;
  v.angleBase = 0;
  const int Na = s.Na;
  v.opCos = trig;
  v.opSin = trig + Na;
  v.aCos = trig + 2 * Na;
  v.aSin = trig + 3 * Na;
  v.bkCos = trig + 4 * Na;
  v.bkSin = trig + 5 * Na;
  v.saCos = saCos;
  v.saSin = saSin;
  v.vFx = vFx;
  v.vFy = vFy;
  v.weight = weight;
  v.basis = basis;
  v.labels = labels;
  const size_t nc = size_t(Na) * s.N1;
  v.idxInG = idx;
  v.idxOutG = idx + nc;
  v.idxInGt = idx + 2 * nc;
  v.idxOutGt = idx + 3 * nc;
  v.distInG = dist;
  v.distOutG = dist + s.nInG;
  v.distInGt = dist + s.nInG + s.nOutG;
  v.distOutGt = dist + s.nInG + s.nOutG + s.nInGt;
  v.bInG = bInG;
  v.bOutG = bOutG;
  v.bInGt = bInGt;
  v.bOutGt = bOutGt;
  v.nInG = s.nInG;
  v.nOutG = s.nOutG;
  v.nInGt = s.nInGt;
  v.nOutGt = s.nOutGt;
  v.extDelta = extDelta;
  v.A = A;
  v.G = G;
  return v;
}

// ---------------------------------------------------------------------------
// host I/O
// ---------------------------------------------------------------------------
// Host transfers are ordered on the grain's stream and complete on return.
void Grain::toDevice(void *dst, const void *src, size_t bytes) {
  SC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream_));
  SC_HIP(hipStreamSynchronize(stream_));
}
void Grain::fromDevice(void *dst, const void *src, size_t bytes) {
  SC_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream_));
  SC_HIP(hipStreamSynchronize(stream_));
}

void Grain::setOrderParameter(int c, const double *re, const double *im) {
  const Setup &s = *s_;
  if (c < 0 || c >= s.Nc) throw std::runtime_error("component index out of range");
  const size_t n0 = size_t(s.N0) * s.N0;
  std::vector<double> buf(2 * n0);
  for (size_t i = 0; i < n0; ++i) {
    // OrderParameter::initialize masks the sites outside the grain.
    buf[i] = s.domain[i] ? re[i] : 0.0;
    buf[n0 + i] = s.domain[i] ? im[i] : 0.0;
  }
  toDevice(dDelta_ + size_t(c) * 2 * n0, buf.data(), 2 * n0 * sizeof(double));
}
void Grain::getOrderParameter(int which, int c, double *re, double *im) {
  const Setup &s = *s_;
  if (c < 0 || c >= s.Nc) throw std::runtime_error("component index out of range");
  const size_t n0 = size_t(s.N0) * s.N0;
  const double *src = (which == 0 ? dDelta_ : dDeltaNew_) + size_t(c) * 2 * n0;
  fromDevice(re, src, n0 * sizeof(double));
  fromDevice(im, src + n0, n0 * sizeof(double));
}
void Grain::setVectorPotential(const double *ax, const double *ay, bool scaled) {
  const Setup &s = *s_;
  const size_t n0 = size_t(s.N0) * s.N0;
  // The induced potential is stored scaled by lambda^2 (src/simulation.cu:369-376).
  const double f = scaled ? 1.0 : (s.lambda > 0.0 ? s.lambda * s.lambda : 0.0);
  std::vector<double> buf(2 * n0);
  for (size_t i = 0; i < n0; ++i) {
    buf[i] = ax[i] * f;
    buf[n0 + i] = ay[i] * f;
  }
  toDevice(dA_, buf.data(), 2 * n0 * sizeof(double));
}
void Grain::getVectorPotential(int which, double *ax, double *ay, bool scaled) {
  const Setup &s = *s_;
  const size_t n0 = size_t(s.N0) * s.N0;
  const double *src = which == 0 ? dA_ : dANew_;
  fromDevice(ax, src, n0 * sizeof(double));
  fromDevice(ay, src + n0, n0 * sizeof(double));
  if (!scaled && s.lambda > 0.0) {
    const double f = 1.0 / (s.lambda * s.lambda);
    for (size_t i = 0; i < n0; ++i) {
      ax[i] *= f;
      ay[i] *= f;
    }
  }
}
void Grain::getCurrent(double *jx, double *jy) {
  const size_t n0 = size_t(s_->N0) * s_->N0;
  fromDevice(jx, dCurrent_, n0 * sizeof(double));
  fromDevice(jy, dCurrent_ + n0, n0 * sizeof(double));
}
void Grain::getEnergyDensity(double *out) {
  const size_t n0 = size_t(s_->N0) * s_->N0;
  fromDevice(out, dEnergyDens_, n0 * sizeof(double));
}
void Grain::getBoundary(int inOut, int dir, double *re, double *im) {
  const Setup &s = *s_;
  const long long n = (dir == 0 ? (inOut == 0 ? s.nInG : s.nOutG)
                                : (inOut == 0 ? s.nInGt : s.nOutGt)) *
                      s.Ne;
  const double *p = dir == 0 ? (inOut == 0 ? dBInG_ : dBOutG_)
                             : (inOut == 0 ? dBInGt_ : dBOutGt_);
  fromDevice(re, p, n * sizeof(double));
  fromDevice(im, p + n, n * sizeof(double));
}
void Grain::setBoundary(int dir, const double *re, const double *im) {
  const Setup &s = *s_;
  const long long n = (dir == 0 ? s.nInG : s.nInGt) * s.Ne;
  double *p = dir == 0 ? dBInG_ : dBInGt_;
  toDevice(p, re, n * sizeof(double));
  toDevice(p + n, im, n * sizeof(double));
  boundaryInit_ = true;
}

// Initial<T>::bulk / ::random / ::perturbedPhaseWindings
// (OrderParameterInitial.h:186-238,286-427).  The noise is Gaussian from a
// seeded generator (the reference seeds with clock()).
void Grain::initialGuess() {
  const Setup &s = *s_;
  const int N0 = s.N0;
  const size_t n0 = size_t(N0) * N0;
  std::mt19937_64 rng(s.cfg.seed);
  std::normal_distribution<double> nd(0.0, 1.0);
  for (int c = 0; c < s.Nc; ++c) {
    std::vector<double> re(n0), im(n0);
    const double half = 0.5 * s.cfg.grainWidth;
    for (int j = 0; j < N0; ++j)
      for (int i = 0; i < N0; ++i) {
        double phase = s.cfg.phaseShift[c];
        for (const Vortex &vx : s.cfg.vortices[c]) {
          const double cx = (vx.cx - s.cfg.geomOffX) / half;
          const double cy = (vx.cy - s.cfg.geomOffY) / half;
          const double x = double(2 * i) / double(N0 - 1) - 1.0 - cx;
          const double y = double(2 * j) / double(N0 - 1) - 1.0 - cy;
          const double r = std::hypot(x, y);
          phase += (r > 0.0 ? std::atan2(y, x) : 0.0) * vx.winding;
        }
        const size_t q = size_t(j) * N0 + i;
        const double nr = s.cfg.noise[c] * nd(rng), ni = s.cfg.noise[c] * nd(rng);
        re[q] = s.deltaBulk[c] * std::cos(phase) + nr;
        im[q] = s.deltaBulk[c] * std::sin(phase) + ni;
      }
    setOrderParameter(c, re.data(), im.data());
  }
}

// ---------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------
namespace {
double residualFromStats(const double *st, long long numElements, int norm) {
  // ComputeResidual.h:69-106
  double numer, denom, eps = 1e-4;
  if (norm == NORM_L1) {
    numer = st[0];
    denom = st[1];
    eps *= static_cast<double>(numElements);
  } else if (norm == NORM_L2) {
    numer = std::sqrt(st[2]);
    denom = std::sqrt(st[3]);
    eps *= static_cast<double>(numElements);
  } else {
    numer = st[4];
    denom = st[5];
  }
  return numer / std::max(denom, eps);
}
} // namespace

// Deferred reductions.  Each queued reduction writes its block partials to
// its own range of dRed_; flushReductions() brings every pending range back
// with one transfer into pinned host memory and combines the partials in
// block order.  queueReadback() starts that transfer without waiting, so
// that a batch can start the transfers of all its grains before it waits on
// the first.
double *Grain::reserveReduction(int kind, long long blocks, long long n, int &handle) {
  const size_t need = size_t(blocks) * (kind == 0 ? 1 : 6);
  redReserved_ += need;
  if (readbackQueued_) flushReductions(); // the queued transfer would miss this one
  if (redUsed_ + need > redCap_) {
    flushReductions(); // keeps the results of everything queued so far
    if (need > redCap_) {
      devFree(dRed_);
      redCap_ = std::max(need, 2 * redCap_);
      dRed_ = devAlloc<double>(redCap_);
      dropEndGraph(); // its kernels write into the old buffer
    }
  }
  Pending p;
  p.kind = kind;
  p.blocks = blocks;
  p.n = n;
  p.offset = redUsed_;
  p.handle = static_cast<int>(redOut_.size());
  redOut_.push_back({0, 0, 0, 0, 0, 0});
  redN_.push_back(n);
  pend_.push_back(p);
  redUsed_ += need;
  handle = p.handle;
  return dRed_ + p.offset;
}

int Grain::queueSum(const double *a, const double *b, const uint8_t *mask, long long n,
                    int op) {
  const long long blocks = std::min<long long>(blocksFor(n), 1024);
  int h = -1;
  double *dst = reserveReduction(0, blocks, n, h);
  sumKernel<<<blocks, 256, 0, stream_>>>(a, b, mask, n, op, dst);
  SC_HIP(hipGetLastError());
  return h;
}

int Grain::queueStats(const double *oldv, const double *newv, long long n) {
  const long long blocks = std::min<long long>(blocksFor(n), 1024);
  int h = -1;
  double *dst = reserveReduction(1, blocks, n, h);
  statsKernel<<<blocks, 256, 0, stream_>>>(oldv, newv, n, dst);
  SC_HIP(hipGetLastError());
  return h;
}

void Grain::queueReadback() {
  if (pend_.empty() || readbackQueued_) return;
  if (redUsed_ > redPinnedCap_) {
    if (redPinned_) SC_HIP(hipHostFree(redPinned_));
    redPinned_ = nullptr;
    redPinnedCap_ = 0;
    SC_HIP(hipHostMalloc(reinterpret_cast<void **>(&redPinned_), redCap_ * sizeof(double),
                         hipHostMallocDefault));
    redPinnedCap_ = redCap_;
  }
  SC_HIP(hipMemcpyAsync(redPinned_, dRed_, redUsed_ * sizeof(double), hipMemcpyDeviceToHost,
                        stream_));
  SC_HIP(hipEventRecord(evReadback_, stream_));
  readbackQueued_ = true;
}

void Grain::flushReductions() {
  if (pend_.empty()) return;
  queueReadback();
  SC_HIP(hipEventSynchronize(evReadback_));
  readbackQueued_ = false;
  for (const Pending &p : pend_) {
    double *out = redOut_[size_t(p.handle)].data();
    const double *h = redPinned_ + p.offset;
    if (p.kind == 0) {
      double sum = 0.0;
      for (long long b = 0; b < p.blocks; ++b) sum += h[b];
      out[0] = sum;
    } else {
      double st[6] = {0, 0, 0, 0, 0, 0};
      for (long long b = 0; b < p.blocks; ++b)
        for (int q = 0; q < 6; ++q)
          st[q] = q < 4 ? st[q] + h[size_t(b) * 6 + q] : std::max(st[q], h[size_t(b) * 6 + q]);
      for (int q = 0; q < 6; ++q) out[q] = st[q];
    }
  }
  pend_.clear();
  redUsed_ = 0;
}

void Grain::clearReductions() {
  flushReductions();
  redOut_.clear();
  redN_.clear();
}

double Grain::reductionSum(int h) const { return redOut_[size_t(h)][0]; }

double Grain::reductionResidual(int h) const {
  return residualFromStats(redOut_[size_t(h)].data(), redN_[size_t(h)], s_->cfg.norm);
}

double Grain::residualOf(const double *oldv, const double *newv, long long n) {
  const int h = queueStats(oldv, newv, n);
  flushReductions();
  return reductionResidual(h);
}

double Grain::deviceSum(const double *a, const double *b, const uint8_t *mask,
                        long long n, int op) {
  const int h = queueSum(a, b, mask, n, op);
  flushReductions();
  return reductionSum(h);
}

std::vector<double> Grain::multiDot(const double *a, const std::vector<double *> &bs,
                                    long long n) {
  return multiDotGroups({DotGroup{a, bs, n}})[0];
}

std::vector<std::vector<double>> Grain::multiDotGroups(const std::vector<DotGroup> &groups) {
  // every group's kernels write their block partials to their own range of
  // dPartials_; one transfer into pinned host memory brings them all back,
  // and each dot product sums its partials in block order
  struct Launch {
    size_t group, first;
    int m;
    long long blocks;
    size_t offset;
  };
  std::vector<Launch> launches;
  size_t total = 0;
  for (size_t gi = 0; gi < groups.size(); ++gi) {
    const long long blocks = std::min<long long>(blocksFor(groups[gi].n), 512);
    for (size_t first = 0; first < groups[gi].bs.size(); first += kMaxDots) {
      const int m = static_cast<int>(std::min<size_t>(kMaxDots, groups[gi].bs.size() - first));
      launches.push_back({gi, first, m, blocks, total});
      total += size_t(blocks) * m;
    }
  }
  if (static_cast<long long>(total) > partialsCap_ * 6) {
    SC_HIP(hipStreamSynchronize(stream_));
    devFree(dPartials_);
    partialsCap_ = static_cast<long long>((total + 5) / 6);
    dPartials_ = devAlloc<double>(size_t(partialsCap_) * 6);
  }
  if (total > dotPinnedCap_) {
    SC_HIP(hipStreamSynchronize(stream_));
    if (dotPinned_) SC_HIP(hipHostFree(dotPinned_));
    dotPinned_ = nullptr;
    dotPinnedCap_ = 0;
    SC_HIP(hipHostMalloc(reinterpret_cast<void **>(&dotPinned_), total * sizeof(double),
                         hipHostMallocDefault));
    dotPinnedCap_ = total;
  }
  // up to kMaxDotGroups launches' worth of work per kernel
  for (size_t first = 0; first < launches.size(); first += kMaxDotGroups) {
    const size_t count = std::min<size_t>(kMaxDotGroups, launches.size() - first);
    DotGroupTable t{};
    t.numGroups = static_cast<int>(count);
    int blocks = 0;
    for (size_t q = 0; q < count; ++q) {
      const Launch &l = launches[first + q];
      const DotGroup &g = groups[l.group];
      t.a[q] = g.a;
      for (int i = 0; i < l.m; ++i) t.b[q][i] = g.bs[l.first + size_t(i)];
      t.m[q] = l.m;
      t.n[q] = g.n;
      t.offset[q] = static_cast<long long>(l.offset);
      t.blockStart[q] = blocks;
      blocks += static_cast<int>(l.blocks);
    }
    t.blockStart[count] = blocks;
    if (blocks == 0) continue; // empty vectors: their dots stay zero
    multiDotGroupsKernel<<<blocks, 256, 0, stream_>>>(t, dPartials_);
    SC_HIP(hipGetLastError());
  }
  std::vector<std::vector<double>> out(groups.size());
  for (size_t gi = 0; gi < groups.size(); ++gi) out[gi].assign(groups[gi].bs.size(), 0.0);
  if (total == 0) return out;
  SC_HIP(hipMemcpyAsync(dotPinned_, dPartials_, total * sizeof(double), hipMemcpyDeviceToHost,
                        stream_));
  SC_HIP(hipEventRecord(evDots_, stream_));
  SC_HIP(hipEventSynchronize(evDots_));
  for (const Launch &l : launches) {
    const double *h = dotPinned_ + l.offset;
    std::vector<double> &o = out[l.group];
    for (long long blk = 0; blk < l.blocks; ++blk)
      for (int i = 0; i < l.m; ++i) o[l.first + size_t(i)] += h[size_t(blk) * l.m + i];
  }
  return out;
}

void Grain::discardReductions() {
  // the queued partials are never read; the next reservations reuse dRed_
  // behind them in stream order
  if (readbackQueued_) flushReductions();
  pend_.clear();
  redUsed_ = 0;
  redOut_.clear();
  redN_.clear();
}

// ---------------------------------------------------------------------------
// the iteration
// ---------------------------------------------------------------------------
void Grain::extendDelta(bool fields) {
  const Setup &s = *s_;
  const size_t n0 = size_t(s.N0) * s.N0;
  SC_HIP(hipMemcpyAsync(dExtDelta_, dDelta_, size_t(2 * s.Nc) * n0 * sizeof(double),
                        hipMemcpyDeviceToDevice, stream_));
  const int nt = static_cast<int>(s.extRef.target.size());
  if (nt > 0) {
    extendKernel<<<blocksFor(nt), 256, 0, stream_>>>(
        dExtDelta_, dExtRefT_, dExtRefO_, dExtRefS_, nt, 2 * s.Nc,
        static_cast<long long>(n0), static_cast<long long>(n0), 0, 1, 0);
    SC_HIP(hipGetLastError());
  }
  // the site fields every sweep of this iteration reads (Keldysh grains
  // compute them per angle chunk; a batch computes those of all its grains
  // in one launch)
  if (!s.keldysh && fields)
    computeFields(0, s.Na, static_cast<long long>(s.Na) * s.N1 * s.N1, dFields_);
}

void Grain::computeFields(int k0, int frames, long long plane, double *F) {
  const Setup &s = *s_;
  const int t0 = s.fieldTileStart[size_t(k0)], t1 = s.fieldTileStart[size_t(k0 + frames)];
  if (t1 <= t0) return;
  GrainView v = view();
  fieldsKernel<<<xcdGrid(t1 - t0), 256, 0, stream_>>>(v, dFieldTiles_ + 4 * size_t(t0), t1 - t0,
                                                       k0, plane, F);
  SC_HIP(hipGetLastError());
}

GrainView Grain::view() const {
  GrainView v = makeView(*s_, dTrig_, dIdx_, dDist_, dBInG_, dBOutG_, dBInGt_, dBOutGt_,
                         dExtDelta_, dA_, s_->keldysh ? dLdosG_ : dG_, dEnergy_, dResidue_,
                         dSaCos_, dSaSin_, dVFx_, dVFy_, dWeight_, dBasis_, dLabels_);
  v.energyRe = dEnergyRe_;
  const long long n1 = static_cast<long long>(s_->N1) * s_->N1;
  v.fields = s_->keldysh ? dLdosFields_ : dFields_;
  v.fieldPlane = (s_->keldysh ? ldosFieldFrames_ : s_->Na) * n1;
  v.fieldBase = 0;
  v.fieldTiles = dFieldTiles_;
  v.numFieldTiles = s_->fieldTileStart[size_t(s_->Na)];
  return v;
}

GrainView Grain::sweepView(int el) const {
  GrainView v = view();
  // each column takes one lane per energy group; a wavefront carries as many
  // whole columns as fit in its 64 lanes (the rest of the lanes idle)
  v.groups = (s_->Ne + el - 1) / el;
  if (v.groups > 64)
    throw std::runtime_error("too many Ozaki energies for the lane layout; "
                             "raise the energies per lane");
  v.cols = 64 / v.groups;
  return v;
}

void Grain::ensureBoundaryInit() {
  if (boundaryInit_) return;
  const Setup &s = *s_;
  extendDelta();
  GrainView v = view();
  // homogeneous out-values at every exit crossing, then the boundary condition
  const int *colG = dOutCross_, *yG = dOutCross_ + s.nOutG;
  const int *colGt = dOutCross_ + 2 * s.nOutG, *yGt = colGt + s.nOutGt;
  if (s.nOutG > 0) {
    initialOutKernel<<<blocksFor(s.nOutG * s.Ne), 256, 0, stream_>>>(v, colG, yG, s.nOutG, 0,
                                                                      dBOutG_);
    SC_HIP(hipGetLastError());
  }
  if (s.nOutGt > 0) {
    initialOutKernel<<<blocksFor(s.nOutGt * s.Ne), 256, 0, stream_>>>(v, colGt, yGt, s.nOutGt,
                                                                        1, dBOutGt_);
    SC_HIP(hipGetLastError());
  }
  boundaryCondition();
  boundaryInit_ = true;
}

template <typename T, int EL, int W, bool KEL>
static void launchSweep(const SweepArgs &a, int slots, hipStream_t st) {
  sweepKernel<T, EL, W, KEL><<<slots, 64, 0, st>>>(a, a.views, a.items);
}

void launchSweepVariant(int precision, int el, int waves, const SweepArgs &a, int slots,
                        hipStream_t stream, bool keldysh) {
  const bool dp = precision == PREC_FP64;
#define SC_LAUNCH(E, W)                                                              \
  if (el == E && waves == W) {                                                       \
    if (dp) launchSweep<double, E, W, KEL>(a, slots, stream);                        \
    else launchSweep<float, E, W, KEL>(a, slots, stream);                            \
    SC_HIP(hipGetLastError());                                                       \
    return;                                                                          \
  }
  if (keldysh) {
    constexpr bool KEL = true;
    SC_KELDYSH_VARIANTS(SC_LAUNCH)
  } else {
    constexpr bool KEL = false;
    SC_SWEEP_VARIANTS(SC_LAUNCH)
  }
#undef SC_LAUNCH
  throw std::runtime_error("sweep variant not compiled");
}

void Grain::sweep(int precision, bool burnin) {
  if (s_->keldysh) {
    ldosSweepAll(precision, burnin);
    return;
  }
  prepareSweep(precision);
  const SweepPlan &pl = plan_[precision];
  SweepArgs a;
  a.views = reinterpret_cast<const GrainView *>(pl.dView);
  a.items = reinterpret_cast<const SweepItemDev *>(pl.dItems);
  a.numItems = pl.numItems;
  a.counter = dCounter_;
  a.scratch = dScratch_;
  a.slotStride = static_cast<long long>(pl.maxLen) * pl.el * 64;
  a.burnin = burnin ? 1 : 0;
  resetCounter<<<1, 1, 0, stream_>>>(dCounter_);
  SC_HIP(hipEventRecord(ev0_, stream_));
  launchSweepVariant(precision, pl.el, pl.waves, a, pl.numSlots, stream_, s_->keldysh);
  SC_HIP(hipEventRecord(ev1_, stream_));
}

void Grain::autotune(int precision) {
  if (precision != PREC_FP64 && precision != PREC_MIXED)
    throw std::runtime_error("unknown precision");
  if (s_->keldysh) {
    // the sweeps write spectral weights: the chunk buffer must exist
    if (ldosLastBlock_ != ldosBlockStart_ && ldosBlockStart_ < s_->spec.numEnergies)
      ldosBeginBlock();
    ldosEnsureBuffers(precision);
  }
  ensureBoundaryInit();
  extendDelta();
  SweepPlan &pl = plan_[precision];
  const int keepEl = pl.el, keepW = pl.waves, keepSlots = pl.slotsPerCU;
  float best = -1.f;
  int bestEl = keepEl, bestW = keepW, bestSlots = keepSlots;
  const int slotChoices[] = {4, 8, 16};
  auto tryShape = [&](int el, int w, int slots) {
    if ((s_->Ne + el - 1) / el > 64) return;
    setSweepShape(precision, el, w, slots);
    // the faster of two launches: one launch varies by a few per cent
    float ms = -1.f;
    for (int rep = 0; rep < 2; ++rep) {
      sweep(precision, false);
      updateSweepTime();
      const float t = static_cast<float>(tim_.sweepMs);
      if (ms < 0.f || t < ms) ms = t;
    }
    if (best < 0.f || ms < best) {
      best = ms;
      bestEl = el;
      bestW = w;
      bestSlots = slots;
    }
  };
#define SC_TRY(E, W)                                                                 \
  for (int slots : slotChoices)                                                       \
    if (slots >= 4 * W || slots == 16) tryShape(E, W, slots);
  if (s_->keldysh) {
    SC_KELDYSH_VARIANTS(SC_TRY)
  } else {
    SC_SWEEP_VARIANTS(SC_TRY)
  }
#undef SC_TRY
  setSweepShape(precision, bestEl, bestW, bestSlots);
  (void)keepEl;
  (void)keepW;
  (void)keepSlots;
}

void Grain::queueBoundaryCondition(int handles[2]) {
  // BoundaryConditionSpecular::computeBoundaryCondition (:361-486): new
  // in-values from the out-values of the reflected angles, with the change
  // statistics for the residual.
  const Setup &s = *s_;
  for (int dir = 0; dir < 2; ++dir) {
    handles[dir] = -1;
    const long long nIn = dir == 0 ? s.nInG : s.nInGt;
    const long long nOut = dir == 0 ? s.nOutG : s.nOutGt;
    const long long total = nIn * s.Ne;
    if (total == 0) continue;
    const long long blocks = blocksFor(total);
    double *dst = reserveReduction(1, blocks, 2 * total, handles[dir]);
    boundaryKernel<<<blocks, 256, 0, stream_>>>(
        dir == 0 ? dBInG_ : dBInGt_, dir == 0 ? dBOutG_ : dBOutGt_, nIn, nOut, s.Ne,
        dir == 0 ? dBcOffG_ : dBcOffGt_, dir == 0 ? dBcSrcG_ : dBcSrcGt_,
        dir == 0 ? dBcWG_ : dBcWGt_, dir == 0 ? dBcWsG_ : dBcWsGt_, dst);
    SC_HIP(hipGetLastError());
  }
}

double Grain::boundaryResidual(const int handles[2]) const {
  double r = 0.0;
  for (int dir = 0; dir < 2; ++dir)
    if (handles[dir] >= 0) r = std::max(r, reductionResidual(handles[dir]));
  return r;
}

void Grain::boundaryCondition() {
  int hb[2];
  queueBoundaryCondition(hb);
  flushReductions();
  res_.resBoundary = boundaryResidual(hb);
}

void Grain::extendFrames(double *base, int k0, int k1, int planes, long long planeStride,
                         long long blockStride, long long siteScale) {
  const size_t n1 = size_t(s_->N1) * s_->N1;
  const int t0 = extRotFrameStart_[size_t(k0)], t1 = extRotFrameStart_[size_t(k1)];
  if (t1 <= t0) return;
  extendKernel<<<blocksFor(t1 - t0), 256, 0, stream_>>>(
      base, dExtRotT_ + t0, dExtRotO_ + t0, dExtRotS_, t1 - t0, planes, planeStride,
      static_cast<long long>(n1), blockStride, siteScale, k0);
  SC_HIP(hipGetLastError());
}

void Grain::backRotate() {
  // Extension of the Green-function planes of every frame, then their
  // back-rotation, angle integration, gap equation and current
  // (ComputeOrderParameter.h:174-262, ComputeCurrent.h:142-222).
  const Setup &s = *s_;
  const size_t n1 = size_t(s.N1) * s.N1;
  const int nt = static_cast<int>(s.extRot.target.size());
  if (nt > 0) {
    extendKernel<<<blocksFor(nt), 256, 0, stream_>>>(
        dG_, dExtRotT_, dExtRotO_, dExtRotS_, nt, 4, 1, static_cast<long long>(n1),
        static_cast<long long>(4 * n1), 4, 0);
    SC_HIP(hipGetLastError());
  }
  GrainView v = view();
  const int ns = static_cast<int>(s.domainSites.size());
  const double jf = 2.0 * static_cast<double>(s.cfg.chargeSign) * s.temperature;
  gatherKernel<<<xcdGrid(blocksFor(ns)), 256, 0, stream_>>>(v, dDomainSites_, ns, dCoupling_, jf,
                                                   dDeltaNew_, dCurrent_, dEnergyDens_);
  SC_HIP(hipGetLastError());
}

void Grain::afterSweep(bool burnin) {
  // Back-rotation and angle integration of the swept Green functions, the
  // compute objects' reductions and the boundary condition, queued for one
  // read-back.
  const Setup &s = *s_;
  const size_t n0 = size_t(s.N0) * s.N0, n1 = size_t(s.N1) * s.N1;
  pendingIntegrate_ = !burnin;
  if (!burnin) {
    SC_HIP(hipMemcpyAsync(dCurrentOld_, dCurrent_, 2 * n0 * sizeof(double),
                          hipMemcpyDeviceToDevice, stream_));
    if (rangeMode_) {
      // the angle sums were added range by range: their final scaling
      GrainView v = view();
      const int ns = static_cast<int>(s.domainSites.size());
      const double jf = 2.0 * static_cast<double>(s.cfg.chargeSign) * s.temperature;
      if (ns > 0) {
        gatherFinishKernel<<<blocksFor(ns), 256, 0, stream_>>>(
            v, dDomainSites_, ns, dGatherAcc_, dCoupling_, jf, dDeltaNew_, dCurrent_,
            dEnergyDens_);
        SC_HIP(hipGetLastError());
      }
    } else {
      backRotate();
    }
    // ComputeOrderParameter::computePropertyFromGreens (:212-262)
    hDelta_.assign(size_t(s.Nc), -1);
    for (int c = 0; c < s.Nc; ++c)
      hDelta_[size_t(c)] = queueStats(dDelta_ + size_t(c) * 2 * n0,
                                      dDeltaNew_ + size_t(c) * 2 * n0,
                                      static_cast<long long>(2 * n0));
    // ComputeCurrent::computePropertyFromGreens (:175-222)
    hCurrent_ = s.cfg.computeCurrent
                    ? queueStats(dCurrentOld_, dCurrent_, static_cast<long long>(2 * n0))
                    : -1;
    // ComputeFreeEnergy::computePropertyFromGreens (:281-414)
    hEnergy_ = -1;
    hCurl_ = -1;
    hDeltaSq_.assign(size_t(2 * s.Nc), -1);
    if (s.cfg.computeFreeEnergy) {
      hEnergy_ = queueSum(dEnergyDens_, nullptr, nullptr, static_cast<long long>(n0), 0);
      for (int c = 0; c < s.Nc; ++c)
        for (int q = 0; q < 2; ++q)
          hDeltaSq_[size_t(2 * c + q)] =
              queueSum(dDelta_ + size_t(c) * 2 * n0 + size_t(q) * n0, nullptr, dDomain_,
                       static_cast<long long>(n0), 2);
      if (s.lambda > 0.0) {
        curlSquareKernel<<<blocksFor(n0), 256, 0, stream_>>>(dA_, s.N0, 1.0 / s.h, dWork_);
        SC_HIP(hipGetLastError());
        hCurl_ = queueSum(dWork_, nullptr, nullptr, static_cast<long long>(n0), 0);
      }
    }
  }
  queueBoundaryCondition(hBoundary_);
}

void Grain::finishIntegrate() {
  // Host side of the compute objects once the queued reductions are back.
  const Setup &s = *s_;
  res_.resBoundary = boundaryResidual(hBoundary_);
  if (!pendingIntegrate_) return;
  pendingIntegrate_ = false;
  double rd = 0.0;
  for (int c = 0; c < s.Nc; ++c) rd = std::max(rd, reductionResidual(hDelta_[size_t(c)]));
  res_.resDelta = rd;
  convDelta_ = rd < s.cfg.convergence;
  if (s.cfg.computeCurrent) {
    const double actual = reductionResidual(hCurrent_);
    res_.resCurrent = firstCurrent_ ? s.cfg.convergence * 0.1 : actual;
    convCurrent_ = res_.resCurrent < s.cfg.convergence;
    firstCurrent_ = false;
  }
  if (s.cfg.computeFreeEnergy) {
    double nonMag = s.temperature * reductionSum(hEnergy_);
    for (int c = 0; c < s.Nc; ++c) {
      const double f = std::log(s.temperature / s.cfg.tc[c]) + s.sumInverseEnergy;
      const double sq = reductionSum(hDeltaSq_[size_t(2 * c)]) +
                        reductionSum(hDeltaSq_[size_t(2 * c + 1)]);
      nonMag += sq * f;
    }
    double mag = 0.0;
    if (hCurl_ >= 0) mag = 0.5 / (s.lambda * s.lambda) * reductionSum(hCurl_);
    const double area = s.h * s.h;
    const double perArea = (nonMag + mag) / static_cast<double>(s.numGrain);
    const double actual =
        std::fabs(fePerAreaOld_ - perArea) / std::max(std::fabs(fePerAreaOld_), 1e-4);
    res_.resFreeEnergy = firstFE_ ? s.cfg.convergence * 0.1 : actual;
    convFE_ = res_.resFreeEnergy < s.cfg.convergence;
    res_.freeEnergy = (nonMag + mag) * area;
    res_.freeEnergyPerArea = perArea;
    res_.magneticFreeEnergy = mag * area;
    res_.magneticFreeEnergyPerArea = mag / static_cast<double>(s.numGrain);
    fePerAreaOld_ = perArea;
    firstFE_ = false;
  }
}

void Grain::updateSweepTime() {
  SC_HIP(hipEventSynchronize(ev1_));
  float ms = 0.f;
  SC_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
  tim_.sweepMs = ms;
}

void Grain::queueExplicitFreeEnergy(std::vector<int> &hSq, int &hCurl) {
  // the sums of |Delta_c|^2 over the grain (re and im planes) and of
  // (curl A)^2 that the free energy adds to the Green-function part
  const Setup &s = *s_;
  const size_t n0 = size_t(s.N0) * s.N0;
  hSq.clear();
  for (int c = 0; c < s.Nc; ++c)
    for (int q = 0; q < 2; ++q)
      hSq.push_back(queueSum(dDelta_ + size_t(c) * 2 * n0 + size_t(q) * n0, nullptr, dDomain_,
                             static_cast<long long>(n0), 2));
  hCurl = -1;
  if (s.lambda > 0.0) {
    curlSquareKernel<<<blocksFor(n0), 256, 0, stream_>>>(dA_, s.N0, 1.0 / s.h, dWork_);
    SC_HIP(hipGetLastError());
    hCurl = queueSum(dWork_, nullptr, nullptr, static_cast<long long>(n0), 0);
  }
}

void Grain::solveVectorPotential(const double *dj, double *dout) {
  // ComputeVectorPotential::solve (:647-745) with the exact Green function:
  // subtract the grain mean of j, correlate with G, scale by -h^2.  The
  // means are reduced on the device so that the solve needs no read-back.
  const Setup &s = *s_;
  const size_t n0 = size_t(s.N0) * s.N0;
  for (int d = 0; d < 2; ++d) {
    const long long blocks = std::min<long long>(blocksFor(static_cast<long long>(n0)), 1024);
    int h = -1;
    double *part = reserveReduction(0, blocks, static_cast<long long>(n0), h);
    sumKernel<<<blocks, 256, 0, stream_>>>(dj + size_t(d) * n0, nullptr, nullptr,
                                           static_cast<long long>(n0), 0, part);
    SC_HIP(hipGetLastError());
    finishMeanKernel<<<1, 1, 0, stream_>>>(part, blocks, static_cast<double>(s.numGrain),
                                           dMean_ + d);
    SC_HIP(hipGetLastError());
  }
  fft_->solve(dj, dMean_, dout);
}

void Grain::resetAccelerator() {
  SC_HIP(hipStreamSynchronize(stream_));
  for (double *p : accX_) devFree(p);
  for (double *p : accD_) devFree(p);
  devFree(dBrLastX_);
  devFree(dBrLastF_);
  devFree(dBrDx_);
  brCount_ = 0;
  // the running mean, previous step and velocity start from zero again
  devFree(dAccMean_);
  devFree(dAccPrev_);
  devFree(dAccVel_);
  accX_.clear();
  accD_.clear();
  accSq_.clear();
  accIdx_.clear();
  accGram_.clear();
  accIter_ = 0;
  accInit_ = false;
  // the next iteration starts without an accelerator step, like the first
  firstCompute_ = true;
}

void Grain::acceleratorStep() {
  acceleratorQueue();
  acceleratorFinish();
}

void Grain::acceleratorQueue() {
  // The device work of an accelerator step up to its read-back (the norms
  // and dot products of CongAcc and Barzilai-Borwein), queued so that a
  // batch can queue every grain's before it waits on one.
  const Setup &s = *s_;
  const size_t n0 = size_t(s.N0) * s.N0;
  const bool doA = s.cfg.computeCurrent && s.lambda > 0.0;
  const int F = 2 * s.Nc + (doA ? 2 : 0);
  const long long n = static_cast<long long>(F) * static_cast<long long>(n0);
  const double *X = dDelta_;
  const double *G = dDeltaNew_;
  accH_.clear();
  switch (s.cfg.accel) {
  case ACC_CONGACC: {
    if (!accInit_) {
      const double st0 = s.cfg.stepSize > 0 ? s.cfg.stepSize : 0.5;
      dAccMean_ = devAlloc<double>(size_t(n));
      accStep_.assign(F, st0);
      accInertia_.assign(F, 0.0);
      accInit_ = true;
    }
    diffKernel<<<blocksFor(n), 256, 0, stream_>>>(dWork_, G, X, n);
    SC_HIP(hipGetLastError());
    // |D_f|^2, |M_f|^2 and D_f . M_f of every field f
    for (int f = 0; f < F; ++f) {
      const double *Df = dWork_ + size_t(f) * n0, *Mf = dAccMean_ + size_t(f) * n0;
      accH_.push_back(queueSum(Df, nullptr, nullptr, static_cast<long long>(n0), 2));
      accH_.push_back(queueSum(Mf, nullptr, nullptr, static_cast<long long>(n0), 2));
      accH_.push_back(queueSum(Df, Mf, nullptr, static_cast<long long>(n0), 1));
    }
    break;
  }
  case ACC_BB: {
    if (!accInit_) {
      bbStep_ = s.cfg.stepSize > 0 ? s.cfg.stepSize : 1.0;
      dAccPrev_ = devAlloc<double>(size_t(n));
      accInit_ = true;
    }
    diffKernel<<<blocksFor(n), 256, 0, stream_>>>(dWork_, G, X, n);
    SC_HIP(hipGetLastError());
    if (accIter_ > 0) {
      diffKernel<<<blocksFor(n), 256, 0, stream_>>>(dWork2_, dWork_, dAccPrev_, n);
      SC_HIP(hipGetLastError());
      // the dot product per field, then the squared norm of the step
      for (int f = 0; f < F; ++f)
        accH_.push_back(queueSum(dAccPrev_ + size_t(f) * n0, dWork2_ + size_t(f) * n0,
                                 nullptr, static_cast<long long>(n0), 1));
      accH_.push_back((accIter_ % 2 == 0) ? queueSum(dWork2_, nullptr, nullptr, n, 2)
                                          : queueSum(dAccPrev_, nullptr, nullptr, n, 2));
    }
    break;
  }
  default:
    break;
  }
}

void Grain::acceleratorFinish() {
  const Setup &s = *s_;
  const size_t n0 = size_t(s.N0) * s.N0;
  const bool doA = s.cfg.computeCurrent && s.lambda > 0.0;
  const int F = 2 * s.Nc + (doA ? 2 : 0);
  const long long n = static_cast<long long>(F) * static_cast<long long>(n0);
  double *X = dDelta_;
  const double *G = dDeltaNew_;
  switch (s.cfg.accel) {
  case ACC_PICARD: {
    const double st = s.cfg.stepSize > 0 ? s.cfg.stepSize : 1.0;
    picardKernel<<<blocksFor(n), 256, 0, stream_>>>(X, G, n, st);
    break;
  }
  case ACC_ANDERSON: {
    // Anderson.h:267-355
    const size_t rank = s.cfg.rank > 0 ? size_t(s.cfg.rank) : 4;
    if (rank > size_t(kMaxHistory))
      throw std::runtime_error("Anderson history rank above " + std::to_string(kMaxHistory));
    const double st = s.cfg.stepSize > 0 ? s.cfg.stepSize : 1.0;
    size_t slot;
    if (accD_.size() == rank) {
      slot = size_t(std::min_element(accIdx_.begin(), accIdx_.end()) - accIdx_.begin());
    } else {
      accX_.push_back(devAlloc<double>(size_t(n)));
      accD_.push_back(devAlloc<double>(size_t(n)));
      accSq_.push_back(0.0);
      accIdx_.push_back(0);
      slot = accD_.size() - 1;
    }
    SC_HIP(hipMemcpyAsync(accX_[slot], X, size_t(n) * sizeof(double),
                          hipMemcpyDeviceToDevice, stream_));
    diffKernel<<<blocksFor(n), 256, 0, stream_>>>(accD_[slot], G, X, n);
    accIdx_[slot] = accIter_;
    const size_t m = accD_.size(), M = m + 1;
    // Only the replaced difference changes the Gram matrix: one pass of dot
    // products of it with every stored difference (itself included).
    // Field blocks: each order-parameter component (two planes) and the
    // vector potential (two planes).  Without block scaling one block spans
    // the whole field (Anderson.h:300-329).
    const int numBlocks = s.cfg.blockScaling ? (F / 2) : 1;
    const long long blockLen = s.cfg.blockScaling ? 2 * static_cast<long long>(n0) : n;
    if (accGram_.size() != rank * rank * size_t(numBlocks))
      accGram_.assign(rank * rank * size_t(numBlocks), 0.0);
    std::vector<double> weight(size_t(numBlocks), 1.0);
    for (int bk = 0; bk < numBlocks; ++bk) {
      std::vector<double *> ds(m);
      for (size_t i = 0; i < m; ++i) ds[i] = accD_[i] + size_t(bk) * size_t(blockLen);
      const std::vector<double> dots = multiDot(accD_[slot] + size_t(bk) * size_t(blockLen), ds,
                                                blockLen);
      double *gb = accGram_.data() + size_t(bk) * rank * rank;
      for (size_t i = 0; i < m; ++i) {
        gb[slot * rank + i] = dots[i];
        gb[i * rank + slot] = dots[i];
      }
      if (s.cfg.blockScaling) {
        const double xx =
            deviceSum(X + size_t(bk) * size_t(blockLen), nullptr, nullptr, blockLen, 2);
        weight[size_t(bk)] = xx > 0.0 ? 1.0 / xx : 0.0;
      }
    }
    for (size_t i = 0; i < m; ++i) {
      double d = 0.0;
      for (int bk = 0; bk < numBlocks; ++bk)
        d += weight[size_t(bk)] * accGram_[size_t(bk) * rank * rank + i * rank + i];
      accSq_[i] = d;
    }
    auto gram = [&](size_t i, size_t j) {
      double d = 0.0;
      for (int bk = 0; bk < numBlocks; ++bk)
        d += weight[size_t(bk)] * accGram_[size_t(bk) * rank * rank + i * rank + j];
      return d;
    };
    std::vector<double> A(M * M, 1.0), b(M, 0.0);
    double sr = 0.0;
    for (double v : accSq_) sr += v;
    const double scale = 1.0 / (sr / static_cast<double>(m));
    for (size_t i = 0; i < m; ++i) {
      A[i * M + i] = accSq_[i] * scale;
      for (size_t j = i + 1; j < m; ++j) {
        A[i * M + j] = gram(i, j) * scale;
        A[j * M + i] = gram(i, j) * scale;
      }
    }
    A[(M - 1) * M + (M - 1)] = 0.0;
    b[M - 1] = 1.0;
    // Gaussian elimination with partial pivoting on the bordered system.
    for (size_t c = 0; c < M; ++c) {
      size_t p = c;
      for (size_t r = c + 1; r < M; ++r)
        if (std::fabs(A[r * M + c]) > std::fabs(A[p * M + c])) p = r;
      if (p != c) {
        for (size_t q = 0; q < M; ++q) std::swap(A[c * M + q], A[p * M + q]);
        std::swap(b[c], b[p]);
      }
      if (A[c * M + c] == 0.0) continue;
      for (size_t r = c + 1; r < M; ++r) {
        const double f = A[r * M + c] / A[c * M + c];
        for (size_t q = c; q < M; ++q) A[r * M + q] -= f * A[c * M + q];
        b[r] -= f * b[c];
      }
    }
    std::vector<double> w(M, 0.0);
    for (size_t r = M; r-- > 0;) {
      double acc = b[r];
      for (size_t q = r + 1; q < M; ++q) acc -= A[r * M + q] * w[q];
      w[r] = A[r * M + r] != 0.0 ? acc / A[r * M + r] : 0.0;
    }
    AndersonTerms terms{};
    for (size_t q = 0; q < m; ++q) {
      terms.X[q] = accX_[q];
      terms.D[q] = accD_[q];
      terms.w[q] = w[q];
    }
    andersonCombine<<<blocksFor(n), 256, 0, stream_>>>(X, terms, static_cast<int>(m), n, st);
    SC_HIP(hipGetLastError());
    ++accIter_;
    break;
  }
  case ACC_CONGACC: {
    // CongAcc.h:271-419
    const double fac = s.cfg.stepFactor > 0 ? s.cfg.stepFactor : 1.234;
    const double minCos = s.cfg.cosMin > 0 ? s.cfg.cosMin : 0.7;
    const double smax = s.cfg.stepMax > 0 ? s.cfg.stepMax : 100.0;
    const double smin = s.cfg.stepMin > 0 ? s.cfg.stepMin : 0.001;
    if (F > kMaxFields) throw std::runtime_error("too many fields for CongAcc");
    flushReductions();
    for (int f = 0; f < F; ++f) {
      const double d = std::sqrt(reductionSum(accH_[size_t(3 * f)]));
      const double m = std::sqrt(reductionSum(accH_[size_t(3 * f + 1)]));
      const double den = d * m;
      if (den > 0.0) {
        const double dot = reductionSum(accH_[size_t(3 * f + 2)]);
        const double cs = dot / den;
        accStep_[f] *= std::pow(fac, cs);
        if (cs < minCos) {
          // bisection for the inertia (CongAcc.h:299-358)
          double eA = dot / (d * m) - minCos, xA = 1.0, eB = 1.0 - minCos, xB = 0.0;
          double xx = 0.5 * (xA + xB);
          bool ok = false;
          while (!ok) {
            const double y = 1.0 - xx;
            const double num = xx * dot + y * d * d;
            const double dd = d * std::sqrt(xx * xx * m * m + 2.0 * xx * y * dot + y * y * d * d);
            const double err = num / dd - minCos;
            ok = std::fabs(err) < 0.01;
            const int se = (0.0 < err) - (err < 0.0), sa = (0.0 < eA) - (eA < 0.0);
            if (se == sa) {
              eA = err;
              xA = xx;
            } else {
              eB = err;
              xB = xx;
            }
            const int sb = (0.0 < eB) - (eB < 0.0), sa2 = (0.0 < eA) - (eA < 0.0);
            if (sa2 == sb) break;
            xx = 0.5 * (xA + xB);
          }
          accInertia_[f] = xx;
        }
      }
      accStep_[f] = std::clamp(accStep_[f], smin, smax);
    }
    // the per-field step and inertia travel in the kernel arguments
    FieldCoefficients coef{};
    for (int f = 0; f < F; ++f) {
      coef.step[f] = accStep_[size_t(f)];
      coef.inertia[f] = accInertia_[size_t(f)];
    }
    congaccKernel<<<blocksFor(n), 256, 0, stream_>>>(X, dAccMean_, dWork_, coef,
                                                     static_cast<long long>(n0), n);
    break;
  }
  case ACC_BB: {
    // BarzilaiBorwein.h:209-243
    const double smax = s.cfg.stepMax > 0 ? s.cfg.stepMax : 100.0;
    const double smin = s.cfg.stepMin > 0 ? s.cfg.stepMin : 0.001;
    if (accIter_ > 0) {
      flushReductions();
      const int hS = accH_[size_t(F)];
      double dot = 0.0;
      for (int f = 0; f < F; ++f) dot += reductionSum(accH_[size_t(f)]);
      dot = std::fabs(dot);
      const double sc = (accIter_ % 2 == 0) ? dot / reductionSum(hS) : reductionSum(hS) / dot;
      bbStep_ *= sc;
    }
    bbStep_ = std::clamp(bbStep_, smin, smax);
    SC_HIP(hipMemcpyAsync(dAccPrev_, dWork_, size_t(n) * sizeof(double),
                          hipMemcpyDeviceToDevice, stream_));
    axpyKernel<<<blocksFor(n), 256, 0, stream_>>>(X, dWork_, n, bbStep_);
    ++accIter_;
    break;
  }
  case ACC_POLYAK: {
    // Polyak.h:216-265
    const double st = s.cfg.stepSize > 0 ? s.cfg.stepSize : 0.5;
    const double drag = s.cfg.drag > 0 ? s.cfg.drag : 0.5;
    if (!accInit_) {
      dAccVel_ = devAlloc<double>(size_t(n));
      accInit_ = true;
    }
    polyakKernel<<<blocksFor(n), 256, 0, stream_>>>(X, dAccVel_, G, n, st, drag);
    break;
  }
  case ACC_BROYDEN:
    broydenStep(X, G, n);
    break;
  }
  SC_HIP(hipGetLastError());
}

// Broyden's second method on F(x) = g(x) - x with a limited, restarted
// memory (no counterpart in the reference; the same iteration as
// scipy.optimize.broyden2 with reduction_method 'restart'): the inverse
// Jacobian is H = -alpha I + sum_i c_i d_i^T; after a step the pair
// (dx, dF) of iterate and residual differences adds
//   c = dx - H dF,  d = dF / |dF|^2
// (the secant condition H_new dF = dx), the memory first emptied when it
// already holds rank - 1 pairs; the step is x <- x - H F.  One read-back per
// step brings every dot product (dF and F against the d_i and dF).
void Grain::broydenStep(double *X, const double *G, long long n) {
  const Setup &s = *s_;
  const int rank = s.cfg.rank > 0 ? s.cfg.rank : 8;
  if (rank > kMaxHistory)
    throw std::runtime_error("Broyden rank above " + std::to_string(kMaxHistory));
  const double alpha = s.cfg.stepSize > 0 ? s.cfg.stepSize : 1.0;
  if (!accInit_) {
    dBrLastX_ = devAlloc<double>(size_t(n));
    dBrLastF_ = devAlloc<double>(size_t(n));
    dBrDx_ = devAlloc<double>(size_t(n));
    brCount_ = 0;
    accInit_ = true;
  }
  double *F = dWork_, *dF = dWork2_;
  diffKernel<<<blocksFor(n), 256, 0, stream_>>>(F, G, X, n);
  SC_HIP(hipGetLastError());
  std::vector<double> b; // d_i . F
  if (accIter_ > 0) {
    diffKernel<<<blocksFor(n), 256, 0, stream_>>>(dF, F, dBrLastF_, n);
    diffKernel<<<blocksFor(n), 256, 0, stream_>>>(dBrDx_, X, dBrLastX_, n);
    SC_HIP(hipGetLastError());
    if (brCount_ > rank - 1) brCount_ = 0;
    std::vector<double *> ds(accD_.begin(), accD_.begin() + brCount_);
    ds.push_back(dF);
    const auto dots = multiDotGroups({DotGroup{dF, ds, n}, DotGroup{F, ds, n}});
    const double dfn2 = dots[0][size_t(brCount_)];
    b.assign(dots[1].begin(), dots[1].begin() + brCount_);
    if (dfn2 > 0.0) {
      if (static_cast<int>(accX_.size()) <= brCount_) {
        accX_.push_back(devAlloc<double>(size_t(n)));
        accD_.push_back(devAlloc<double>(size_t(n)));
      }
      const size_t k = size_t(brCount_);
      // c = dx + alpha dF - sum_i (d_i . dF) c_i
      CombineTerms t{};
      int m = 0;
      t.v[m] = dF;
      t.c[m++] = alpha;
      for (size_t i = 0; i < k; ++i) {
        t.v[m] = accX_[i];
        t.c[m++] = -dots[0][i];
      }
      combineTermsKernel<<<blocksFor(n), 256, 0, stream_>>>(accX_[k], dBrDx_, 1.0, t, m, n);
      CombineTerms u{};
      u.v[0] = dF;
      u.c[0] = 1.0 / dfn2;
      combineTermsKernel<<<blocksFor(n), 256, 0, stream_>>>(accD_[k], nullptr, 0.0, u, 1, n);
      SC_HIP(hipGetLastError());
      b.push_back(dots[1][size_t(brCount_)] / dfn2);
      ++brCount_;
    }
  }
  SC_HIP(hipMemcpyAsync(dBrLastX_, X, size_t(n) * sizeof(double), hipMemcpyDeviceToDevice,
                        stream_));
  SC_HIP(hipMemcpyAsync(dBrLastF_, F, size_t(n) * sizeof(double), hipMemcpyDeviceToDevice,
                        stream_));
  // x <- x + alpha F - sum_i (d_i . F) c_i
  CombineTerms t{};
  int m = 0;
  t.v[m] = F;
  t.c[m++] = alpha;
  for (size_t i = 0; i < b.size(); ++i) {
    t.v[m] = accX_[i];
    t.c[m++] = -b[i];
  }
  combineTermsKernel<<<blocksFor(n), 256, 0, stream_>>>(X, X, 1.0, t, m, n);
  SC_HIP(hipGetLastError());
  ++accIter_;
}

void Grain::beginCompute() {
  queueBeginCompute();
  finishBeginCompute();
}

void Grain::queueBeginCompute() {
  // Context::compute (Context.h:391-447) up to the Riccati sweeps: the
  // accelerator's device reductions
  if (s_->keldysh) throw std::runtime_error("a spectroscopy context only runs the LDOS iteration");
  clearReductions();
  SC_HIP(hipEventRecord(evIterStart_, stream_));
  if (!firstCompute_) acceleratorQueue();
}

void Grain::finishBeginCompute(bool fields) {
  // the accelerator's update once its reductions are back, then the
  // extended order parameter and the site fields
  if (!firstCompute_) acceleratorFinish();
  ensureBoundaryInit();
  extendDelta(fields);
}

void Grain::queueEndStages() {
  const Setup &s = *s_;
  const size_t n0 = size_t(s.N0) * s.N0;
  afterSweep(false);
  hA_ = -1;
  if (s.cfg.computeCurrent && s.lambda > 0.0) {
    solveVectorPotential(dCurrent_, dANew_);
    hA_ = queueStats(dA_, dANew_, static_cast<long long>(2 * n0));
    if (screen_) screenVectorPotential();
  }
}

void Grain::dropEndGraph() {
  if (endExec_) (void)hipGraphExecDestroy(endExec_);
  endExec_ = nullptr;
}

void Grain::setGraphs(bool on) {
  graphsOn_ = on;
  if (!on) dropEndGraph();
}

void Grain::setScreening(double rho) {
  const Setup &s = *s_;
  if (rho < 0.0) throw std::runtime_error("superfluid density must be >= 0");
  if (rho > 0.0 && !(s.lambda > 0.0 && s.cfg.computeCurrent))
    throw std::runtime_error("screening needs an induced vector potential (penetration depth > 0)");
  SC_HIP(hipStreamSynchronize(stream_));
  dropEndGraph(); // the captured stages include (or omit) the screening
  screen_.reset();
  screenK_ = 0.0;
  if (rho > 0.0) {
    screenK_ = rho / (s.lambda * s.lambda);
    screen_ = std::make_unique<PoissonFFT>(s.N0, s.h, stream_, PoissonFFT::Screened{screenK_});
    if (!dZero2_) {
      dZero2_ = devAlloc<double>(2);
      SC_HIP(hipMemsetAsync(dZero2_, 0, 2 * sizeof(double), stream_));
    }
  }
  resetAccelerator();
}

void Grain::queueEndCompute() {
  // The rest of Context::compute after the sweeps: device work only, whose
  // launches never change between iterations (fixed buffers, fixed
  // reduction offsets once the accelerator's read-back has reset them).
  // The first run is eager and sizes the reductions; the second is captured
  // into a hipGraph (no memset nodes: every buffer it clears is cleared by
  // a kernel); later runs replay the graph and redo the host bookkeeping of
  // the reservations it was captured with.
  const bool keyOk = readbackQueued_ == false && redOut_.size() == endCap_.redOutBefore &&
                     redUsed_ == endCap_.redUsedBefore && rangeMode_ == endCap_.rangeMode;
  if (graphsOn_ && endExec_ && keyOk) {
    SC_HIP(hipGraphLaunch(endExec_, stream_));
    for (const Pending &p : endCap_.pend) pend_.push_back(p);
    for (long long n : endCap_.redN) {
      redOut_.push_back({0, 0, 0, 0, 0, 0});
      redN_.push_back(n);
    }
    redUsed_ = endCap_.redUsedAfter;
    pendingIntegrate_ = true;
    hDelta_ = endCap_.hDelta;
    hDeltaSq_ = endCap_.hDeltaSq;
    hCurrent_ = endCap_.hCurrent;
    hEnergy_ = endCap_.hEnergy;
    hCurl_ = endCap_.hCurl;
    hBoundary_[0] = endCap_.hBoundary[0];
    hBoundary_[1] = endCap_.hBoundary[1];
    hA_ = endCap_.hA;
  } else {
    // capture only when no reservation can flush or grow dRed_ (both would
    // synchronise inside the capture): the sizes of the eager run fit
    const bool capture = graphsOn_ && endSeen_ && !readbackQueued_ && !s_->keldysh &&
                         redUsed_ + endCap_.need <= redCap_ &&
                         rangeMode_ == endCap_.rangeMode;
    const size_t redOutBefore = redOut_.size(), redUsedBefore = redUsed_,
                 pendBefore = pend_.size(), reservedBefore = redReserved_;
    hipGraph_t graph = nullptr;
    // the bookkeeping below is rewritten: a graph of other offsets must go
    dropEndGraph();
    if (capture) SC_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    try {
      queueEndStages();
    } catch (...) {
      if (capture) {
        (void)hipStreamEndCapture(stream_, &graph);
        if (graph) (void)hipGraphDestroy(graph);
      }
      throw;
    }
    endCap_.redOutBefore = redOutBefore;
    endCap_.redUsedBefore = redUsedBefore;
    endCap_.redUsedAfter = redUsed_;
    endCap_.need = redReserved_ - reservedBefore;
    endCap_.rangeMode = rangeMode_;
    endCap_.pend.assign(pend_.begin() + static_cast<long>(pendBefore), pend_.end());
    endCap_.redN.assign(redN_.begin() + static_cast<long>(redOutBefore), redN_.end());
    endCap_.hDelta = hDelta_;
    endCap_.hDeltaSq = hDeltaSq_;
    endCap_.hCurrent = hCurrent_;
    endCap_.hEnergy = hEnergy_;
    endCap_.hCurl = hCurl_;
    endCap_.hBoundary[0] = hBoundary_[0];
    endCap_.hBoundary[1] = hBoundary_[1];
    endCap_.hA = hA_;
    endSeen_ = true;
    if (capture) {
      SC_HIP(hipStreamEndCapture(stream_, &graph));
      const hipError_t e = hipGraphInstantiate(&endExec_, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (e != hipSuccess) {
        endExec_ = nullptr;
        SC_HIP(e);
      }
      // the captured work has not run yet
      SC_HIP(hipGraphLaunch(endExec_, stream_));
    }
  }
  SC_HIP(hipEventRecord(evIterStop_, stream_));
}

void Grain::screenVectorPotential() {
  // The plain update B <- L^-1 j(B) (B = lambda^2 A, L^-1 the Poisson
  // solve) has the linear response j = j0 - K B inside the grain (K =
  // superfluid density / lambda^2), so its Jacobian is -L^-1 K: for every
  // wavelength longer than 1 / sqrt(K) the gain exceeds one and the plain
  // iteration oscillates.  The Newton step of that linear model is
  //   dB = (I + L^-1 K)^-1 R = (-L + K)^-1 (-L) R = R - K (-L + K)^-1 R
  // for the residual R = L^-1 j(B) - B, so the accelerator is handed
  //   B + R - K (-L + K)^-1 R
  // in place of L^-1 j(B).  Both have the same fixed points; the residual
  // statistics were taken on R before this.  K is uniform over the grid (the
  // grain's response is that of its interior), a preconditioner rather than
  // the exact Jacobian.
  const Setup &s = *s_;
  const long long n2 = 2LL * s.N0 * s.N0;
  double *R = dWork_, *S = dWork2_;
  diffKernel<<<blocksFor(n2), 256, 0, stream_>>>(R, dANew_, dA_, n2);
  SC_HIP(hipGetLastError());
  screen_->solve(R, dZero2_, S);
  CombineTerms t{};
  t.v[0] = S;
  t.c[0] = -screenK_;
  combineTermsKernel<<<blocksFor(n2), 256, 0, stream_>>>(dANew_, dANew_, 1.0, t, 1, n2);
  SC_HIP(hipGetLastError());
}

void Grain::endCompute() {
  queueEndCompute();
  finishEndCompute();
}

// ---------------------------------------------------------------------------
// Exact Jacobian-vector products of the plain iteration (tangent-linear
// mode).  Every stage after the sweep is linear (the extensions, the site
// fields without the external gauge field, the back-rotation and angle
// integration, the Poisson solve with its mean, the screened update, the
// boundary condition), so each runs on the tangents as it runs on the state;
// the sweep's tangent is tangentSweep.
// ---------------------------------------------------------------------------
void Grain::jacobianVector(const double *dv, double *djv, bool inValues, bool screened) {
  const Setup &s = *s_;
  if (s.keldysh) throw std::runtime_error("a spectroscopy context has no Jacobian products");
  if (s.Nc > 4) throw std::runtime_error("at most 4 order-parameter components");
  const size_t n0 = size_t(s.N0) * s.N0, n1 = size_t(s.N1) * s.N1;
  const long long framePlane = static_cast<long long>(s.Na) * static_cast<long long>(n1);
  const bool doA = s.cfg.computeCurrent && s.lambda > 0.0;
  const size_t nd = size_t(2 * s.Nc) * n0, nx = nd + (doA ? 2 * n0 : 0);
  const size_t nG = size_t(2 * s.nInG * s.Ne);
  prepareAdjoint();
  if (!dTanExt_) {
    dTanExt_ = devAlloc<double>(nd);
    dTanFields_ = devAlloc<double>(3 * size_t(framePlane));
    dTanG_ = devAlloc<double>(4 * size_t(framePlane));
    dTanJ_ = devAlloc<double>(2 * n0);
    dTanE_ = devAlloc<double>(n0);
    dTanOutG_ = devAlloc<double>(std::max<size_t>(1, size_t(2 * s.nOutG * s.Ne)));
    dTanOutGt_ = devAlloc<double>(std::max<size_t>(1, size_t(2 * s.nOutGt * s.Ne)));
    dTanScratch_ = devAlloc<double>(size_t(adjSlots_) * 4 * size_t(adjMaxLen_) * kAdjointEL * 64);
  }
  // the state's fields, then the tangents of the extended order parameter
  // and of the site fields (the induced vector potential's share of the
  // Doppler shift only: the external gauge field is fixed)
  ensureBoundaryInit();
  extendDelta(true);
  SC_HIP(hipMemcpyAsync(dTanExt_, dv, nd * sizeof(double), hipMemcpyDeviceToDevice, stream_));
  const int ne = static_cast<int>(s.extRef.target.size());
  if (ne > 0) {
    extendKernel<<<blocksFor(ne), 256, 0, stream_>>>(
        dTanExt_, dExtRefT_, dExtRefO_, dExtRefS_, ne, 2 * s.Nc, static_cast<long long>(n0),
        static_cast<long long>(n0), 0, 1, 0);
    SC_HIP(hipGetLastError());
  }
  GrainView tv = view();
  tv.extDelta = dTanExt_;
  tv.A = doA ? dv + nd : dA_;
  tv.fluxDensity = 0.0;
  tv.useInduced = doA ? 1 : 0;
  tv.useField = doA ? 1 : 0;
  SC_HIP(hipMemsetAsync(dTanFields_, 0, 3 * size_t(framePlane) * sizeof(double), stream_));
  const int nt = s.fieldTileStart[size_t(s.Na)];
  if (nt > 0) {
    fieldsKernel<<<xcdGrid(nt), 256, 0, stream_>>>(tv, dFieldTiles_, nt, 0, framePlane,
                                                    dTanFields_);
    SC_HIP(hipGetLastError());
  }
  // the sweep's tangent
  SC_HIP(hipMemsetAsync(dTanG_, 0, 4 * size_t(framePlane) * sizeof(double), stream_));
  tangentSweep(sweepView(kAdjointEL), dTanFields_, adjItems_, adjNumItems_, dCounter_,
               dTanScratch_, static_cast<long long>(adjMaxLen_) * kAdjointEL * 64,
               inValues ? dv + nx : nullptr, inValues ? dv + nx + nG : nullptr, dTanOutG_,
               dTanOutGt_, dTanG_, adjSlots_, stream_);
  // the back-rotation, angle integration and gap equation of the tangents
  const int nr = static_cast<int>(s.extRot.target.size());
  if (nr > 0) {
    extendKernel<<<blocksFor(nr), 256, 0, stream_>>>(
        dTanG_, dExtRotT_, dExtRotO_, dExtRotS_, nr, 4, 1, static_cast<long long>(n1),
        static_cast<long long>(4 * n1), 4, 0);
    SC_HIP(hipGetLastError());
  }
  // (the gather writes inside sites only; the Poisson solve reads them all)
  SC_HIP(hipMemsetAsync(djv, 0, nx * sizeof(double), stream_));
  SC_HIP(hipMemsetAsync(dTanJ_, 0, 2 * n0 * sizeof(double), stream_));
  GrainView gv = view();
  gv.G = dTanG_;
  const int ns = static_cast<int>(s.domainSites.size());
  const double jf = 2.0 * static_cast<double>(s.cfg.chargeSign) * s.temperature;
  if (ns > 0) {
    gatherKernel<<<xcdGrid(blocksFor(ns)), 256, 0, stream_>>>(gv, dDomainSites_, ns, dCoupling_,
                                                              jf, djv, dTanJ_, dTanE_);
    SC_HIP(hipGetLastError());
  }
  if (doA) {
    solveVectorPotential(dTanJ_, djv + nd);
    if (screened && screen_) {
      // the screened update's tangent: dA' - K (-L + K)^-1 (dA' - dA)
      const long long n2 = 2LL * static_cast<long long>(n0);
      double *R = dWork_, *S = dWork2_;
      diffKernel<<<blocksFor(n2), 256, 0, stream_>>>(R, djv + nd, dv + nd, n2);
      SC_HIP(hipGetLastError());
      screen_->solve(R, dZero2_, S);
      CombineTerms t{};
      t.v[0] = S;
      t.c[0] = -screenK_;
      combineTermsKernel<<<blocksFor(n2), 256, 0, stream_>>>(djv + nd, djv + nd, 1.0, t, 1, n2);
      SC_HIP(hipGetLastError());
    }
  }
  // the boundary condition's tangent: the same interpolation of the
  // out-values' tangents
  if (inValues) {
    for (int dir = 0; dir < 2; ++dir) {
      const long long nIn = dir == 0 ? s.nInG : s.nInGt;
      const long long nOut = dir == 0 ? s.nOutG : s.nOutGt;
      const long long total = nIn * s.Ne;
      if (total == 0) continue;
      const long long blocks = blocksFor(total);
      int h = -1;
      double *dst = reserveReduction(1, blocks, 2 * total, h);
      boundaryKernel<<<blocks, 256, 0, stream_>>>(
          djv + nx + (dir == 0 ? 0 : nG), dir == 0 ? dTanOutG_ : dTanOutGt_, nIn, nOut, s.Ne,
          dir == 0 ? dBcOffG_ : dBcOffGt_, dir == 0 ? dBcSrcG_ : dBcSrcGt_,
          dir == 0 ? dBcWG_ : dBcWGt_, dir == 0 ? dBcWsG_ : dBcWsGt_, dst);
      SC_HIP(hipGetLastError());
    }
  }
  discardReductions();
}

void Grain::jacobianVectorHost(const double *v, double *jv) {
  const Setup &s = *s_;
  const size_t n0 = size_t(s.N0) * s.N0;
  const bool doA = s.cfg.computeCurrent && s.lambda > 0.0;
  const size_t nd = size_t(2 * s.Nc) * n0, nx = nd + (doA ? 2 * n0 : 0);
  if (!dTanV_) {
    dTanV_ = devAlloc<double>(nd + 2 * n0);
    dTanJV_ = devAlloc<double>(nd + 2 * n0);
  }
  // the state holds lambda^2 A
  const double l2 = doA ? s.lambda * s.lambda : 1.0;
  std::vector<double> buf(v, v + nx);
  for (size_t i = nd; i < nx; ++i) buf[i] *= l2;
  toDevice(dTanV_, buf.data(), nx * sizeof(double));
  jacobianVector(dTanV_, dTanJV_, false, false);
  fromDevice(jv, dTanJV_, nx * sizeof(double));
  for (size_t i = nd; i < nx; ++i) jv[i] /= l2;
  for (size_t i = nx; i < nd + 2 * n0; ++i) jv[i] = 0.0;
}

void Grain::finishEndCompute() {
  // the read-back of the iteration's reductions and the residuals
  const Setup &s = *s_;
  const bool doA = s.cfg.computeCurrent && s.lambda > 0.0;
  flushReductions();
  finishIntegrate();
  if (doA) res_.resVectorPotential = reductionResidual(hA_);
  res_.iterations++;
  firstCompute_ = false;
  // Context::getResidual / isPartiallyConverged (Context.h:462-497)
  double r = doA ? res_.resVectorPotential : 0.0;
  r = std::max(r, res_.resBoundary);
  r = std::max(r, res_.resDelta);
  if (s.cfg.computeCurrent) r = std::max(r, res_.resCurrent);
  if (s.cfg.computeFreeEnergy) r = std::max(r, res_.resFreeEnergy);
  res_.residual = r;
  bool ok = r < s.cfg.convergence && convDelta_;
  if (s.cfg.computeCurrent) ok = ok && convCurrent_;
  if (s.cfg.computeFreeEnergy) ok = ok && convFE_;
  res_.converged = ok ? 1 : 0;
  float ms = 0.f;
  SC_HIP(hipEventElapsedTime(&ms, evIterStart_, evIterStop_));
  tim_.iterMs = ms;
}

void Grain::compute(int precision) {
  beginCompute();
  sweep(precision, false);
  endCompute();
  updateSweepTime();
}

void Grain::beginIntegration(bool burnin) {
  const Setup &s = *s_;
  if (s.keldysh) throw std::runtime_error("a spectroscopy context only runs the LDOS iteration");
  if (burnin) {
    beginBurnin();
  } else {
    beginCompute();
    const size_t ns = s.domainSites.size();
    if (!dGatherAcc_) dGatherAcc_ = devAlloc<double>(std::max<size_t>(1, size_t(2 * s.Nc + 3) * ns));
    if (ns > 0)
      SC_HIP(hipMemsetAsync(dGatherAcc_, 0, size_t(2 * s.Nc + 3) * ns * sizeof(double), stream_));
  }
  nextAngle_ = 0;
  rangeMode_ = true;
  rangeBurnin_ = burnin;
}

void Grain::sweepAngles(int precision, bool burnin, int k0, int k1) {
  const Setup &s = *s_;
  if (!rangeMode_ || burnin != rangeBurnin_)
    throw std::runtime_error("sweep of an angle range outside beginIntegration / endIntegration "
                             "(or with the other burn-in setting)");
  if (precision != PREC_FP64 && precision != PREC_MIXED)
    throw std::runtime_error("unknown precision");
  if (k0 < 0 || k1 > s.Na || k0 >= k1) throw std::runtime_error("angle range out of bounds");
  prepareSweep(precision);
  const SweepPlan &pl = plan_[precision];
  const int i0 = pl.angleStart[size_t(k0)], i1 = pl.angleStart[size_t(k1)];
  SweepArgs a;
  a.views = reinterpret_cast<const GrainView *>(pl.dView);
  a.items = reinterpret_cast<const SweepItemDev *>(pl.dItemsByAngle) + i0;
  a.numItems = i1 - i0;
  a.counter = dCounter_;
  a.scratch = dScratch_;
  a.slotStride = static_cast<long long>(pl.maxLen) * pl.el * 64;
  a.burnin = burnin ? 1 : 0;
  resetCounter<<<1, 1, 0, stream_>>>(dCounter_);
  SC_HIP(hipEventRecord(ev0_, stream_));
  if (a.numItems > 0)
    launchSweepVariant(precision, pl.el, pl.waves, a, std::min(pl.numSlots, a.numItems),
                       stream_, false);
  SC_HIP(hipEventRecord(ev1_, stream_));
}

void Grain::integrateAngles(int k0, int k1) {
  const Setup &s = *s_;
  if (!rangeMode_ || rangeBurnin_)
    throw std::runtime_error("angle integration outside beginIntegration / endIntegration "
                             "(or in a burn-in)");
  if (k0 != nextAngle_ || k1 <= k0 || k1 > s.Na)
    throw std::runtime_error("angles must be integrated in order: expected a range from " +
                             std::to_string(nextAngle_));
  const size_t n1 = size_t(s.N1) * s.N1;
  // the extension of the frames' Green-function planes, then their sums
  extendFrames(dG_ + size_t(k0) * 4 * n1, k0, k1, 4, 1, static_cast<long long>(4 * n1), 4);
  const int ns = static_cast<int>(s.domainSites.size());
  if (ns > 0) {
    gatherRangeKernel<<<blocksFor(ns), 256, 0, stream_>>>(view(), dDomainSites_, ns, k0, k1,
                                                         dGatherAcc_);
    SC_HIP(hipGetLastError());
  }
  nextAngle_ = k1;
}

void Grain::endIntegration(bool burnin) {
  if (!rangeMode_ || burnin != rangeBurnin_)
    throw std::runtime_error("endIntegration without a matching beginIntegration");
  if (!burnin && nextAngle_ != s_->Na)
    throw std::runtime_error("endIntegration before every angle was integrated (next angle " +
                             std::to_string(nextAngle_) + " of " + std::to_string(s_->Na) + ")");
  try {
    if (burnin) endBurnin();
    else endCompute();
  } catch (...) {
    rangeMode_ = false;
    throw;
  }
  rangeMode_ = false;
  updateSweepTime();
}

int Grain::numFields() const {
  const Setup &s = *s_;
  const bool doA = s.cfg.computeCurrent && s.lambda > 0.0;
  return 2 * s.Nc + (doA ? 2 : 0);
}

void Grain::acceleratorStepOn(double *x, const double *g) {
  const Setup &s = *s_;
  if (s.keldysh) throw std::runtime_error("a spectroscopy context has no accelerator");
  const size_t bytes = size_t(numFields()) * size_t(s.N0) * s.N0 * sizeof(double);
  toDevice(dDelta_, x, bytes);
  toDevice(dDeltaNew_, g, bytes);
  clearReductions();
  acceleratorStep();
  fromDevice(x, dDelta_, bytes);
}

void Grain::beginBurnin(bool fields) {
  if (s_->keldysh) throw std::runtime_error("a spectroscopy context only runs the LDOS iteration");
  clearReductions();
  ensureBoundaryInit();
  extendDelta(fields);
}

void Grain::endBurnin() {
  queueEndBurnin();
  finishEndBurnin();
}

void Grain::queueEndBurnin() { afterSweep(true); }

void Grain::finishEndBurnin() {
  flushReductions();
  finishIntegrate();
  res_.iterations++;
  res_.residual = res_.resBoundary;
}

void Grain::burnin(int precision) {
  // Context::burnIn (Context.h:377-389)
  beginBurnin();
  sweep(precision, true);
  endBurnin();
  updateSweepTime();
}

bool Grain::run(int nBurnin, int itMin, int itMax, int precision) {
  // runSimulation (run_simulation.h:124-187), save_frequency < 0.
  // PREC_MIXED_REFINE iterates with FP32 stepping until the residual reaches
  // the FP32 floor (below kRefineSwitch, or no new best residual for
  // kRefineStall iterations), then continues in FP64; convergence only
  // counts on an FP64 iteration, so the converged state is an FP64 fixed
  // point.
  constexpr double kRefineSwitch = 1e-5;
  constexpr int kRefineStall = 20;
  const bool refine = precision == PREC_MIXED_REFINE;
  int prec = refine ? PREC_MIXED : precision;
  // runBurnIn(n, n) (run_iteration.h:85-141): n iterations, or with n < 0
  // until the boundary residual is below the criterion
  if (nBurnin != 0) {
    for (int i = 1;; ++i) {
      burnin(prec);
      const bool bconv = res_.resBoundary < s_->cfg.convergence;
      if ((bconv && i >= nBurnin) || (nBurnin >= 0 && i >= nBurnin)) break;
    }
  }
  if (itMin == 0 && itMax == 0) return false;
  bool conv = false;
  double best = 0.0;
  int stall = 0;
  mixedIterations_ = 0;
  for (int i = 1;; ++i) {
    compute(prec);
    if (prec == PREC_MIXED) ++mixedIterations_;
    conv = res_.converged != 0 && prec == PREC_FP64;
    if (refine && prec == PREC_MIXED) {
      const double r = res_.residual;
      stall = (i > 1 && r >= best) ? stall + 1 : 0;
      best = (i == 1) ? r : std::min(best, r);
      if (r < kRefineSwitch || stall >= kRefineStall || res_.converged) prec = PREC_FP64;
    }
    const bool minReached = i >= itMin;
    const bool maxReached = itMax < 0 ? false : i >= itMax;
    if ((conv && minReached) || maxReached) break;
  }
  return conv;
}

// ---------------------------------------------------------------------------
// Batch: many grains, one sweep launch per iteration
// ---------------------------------------------------------------------------
Batch::Batch(std::vector<std::unique_ptr<Setup>> setups) {
  if (setups.empty()) throw std::runtime_error("a batch needs at least one grain");
  SC_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  for (hipEvent_t *e : {&ev0_[0], &ev0_[1], &ev1_[0], &ev1_[1], &evIt0_, &evIt1_, &evFork_})
    SC_HIP(hipEventCreate(e));
  // three lanes: with the main stream, the four hardware queues a process has
  const size_t numLanes = std::min<size_t>(3, setups.size());
  for (size_t i = 0; i < numLanes; ++i) {
    hipStream_t l = nullptr;
    SC_HIP(hipStreamCreateWithFlags(&l, hipStreamNonBlocking));
    lanes_.push_back(l);
    hipEvent_t e = nullptr;
    SC_HIP(hipEventCreate(&e));
    laneEv_.push_back(e);
  }
  for (size_t i = 0; i < setups.size(); ++i) {
    grains_.push_back(std::make_unique<Grain>(std::move(setups[i]), lanes_[i % numLanes]));
    grains_.back()->initialGuess();
  }
  done_.assign(grains_.size(), 0);
  dCounter_ = devAlloc<int>(1);
}

Batch::~Batch() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  for (hipStream_t l : lanes_) (void)hipStreamSynchronize(l);
  grains_.clear();
  for (Plan &pl : plan_) {
    if (pl.dItems) (void)hipFree(pl.dItems);
    if (pl.dViews) (void)hipFree(pl.dViews);
    if (pl.dFieldJobs) (void)hipFree(pl.dFieldJobs);
  }
  if (dScratch_) (void)hipFree(dScratch_);
  devFree(dCounter_);
  for (hipEvent_t e : {ev0_[0], ev0_[1], ev1_[0], ev1_[1], evIt0_, evIt1_, evFork_})
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : laneEv_) (void)hipEventDestroy(e);
  for (hipStream_t l : lanes_) (void)hipStreamDestroy(l);
  if (stream_) (void)hipStreamDestroy(stream_);
}

void Batch::joinLanes() {
  for (size_t i = 0; i < lanes_.size(); ++i) {
    SC_HIP(hipEventRecord(laneEv_[i], lanes_[i]));
    SC_HIP(hipStreamWaitEvent(stream_, laneEv_[i], 0));
  }
}

void Batch::forkLanes() {
  SC_HIP(hipEventRecord(evFork_, stream_));
  for (hipStream_t l : lanes_) SC_HIP(hipStreamWaitEvent(l, evFork_, 0));
}

std::vector<int> Batch::active() const {
  std::vector<int> a;
  for (size_t i = 0; i < grains_.size(); ++i)
    if (!done_[i]) a.push_back(static_cast<int>(i));
  return a;
}

void Batch::setSweepShape(int precision, int el, int waves, int slotsPerCU) {
  if (precision != PREC_FP64 && precision != PREC_MIXED)
    throw std::runtime_error("unknown precision");
  if (!sweepVariantExists(el, waves))
    throw std::runtime_error("no sweep variant with " + std::to_string(el) +
                             " energies per lane and " + std::to_string(waves) +
                             " wavefronts per SIMD");
  if (slotsPerCU < 1 || slotsPerCU > 32)
    throw std::runtime_error("slots per CU must be in [1, 32]");
  Plan &pl = plan_[precision];
  if (el != pl.el || waves != pl.waves || slotsPerCU != pl.slotsPerCU) pl.ready = false;
  pl.el = el;
  pl.waves = waves;
  pl.slotsPerCU = slotsPerCU;
}

void Batch::sweepShape(int precision, int &el, int &waves, int &slotsPerCU) const {
  const Plan &pl = plan_[precision == PREC_FP64 ? 0 : 1];
  el = pl.el;
  waves = pl.waves;
  slotsPerCU = pl.slotsPerCU;
}

void Batch::prepare(int precision, const std::vector<int> &act) {
  Plan &pl = plan_[precision];
  if (pl.ready && pl.key == act) return;
  SC_HIP(hipStreamSynchronize(stream_));
  // views of every grain with this lane layout (index = grain number)
  std::vector<GrainView> views;
  for (auto &g : grains_) views.push_back(g->sweepView(pl.el));
  // every grain's items at this layout, built once per layout
  if (pl.cacheEl != pl.el) {
    pl.itemsOf.assign(grains_.size(), {});
    pl.maxLenOf.assign(grains_.size(), 0);
    for (size_t gi = 0; gi < grains_.size(); ++gi)
      pl.itemsOf[gi] = sweepItems(grains_[gi]->setup(), views[gi].cols, &pl.maxLenOf[gi]);
    pl.cacheEl = pl.el;
  }
  // the items of the active grains, longest first
  std::vector<SweepItemDev> items;
  std::vector<int> lens;
  pl.maxLen = 0;
  for (int gi : act) {
    pl.maxLen = std::max(pl.maxLen, pl.maxLenOf[size_t(gi)]);
    for (const SweepItem &it : pl.itemsOf[size_t(gi)]) {
      items.push_back({gi, it.angle, it.x0, it.ncols, it.ylo, it.yhi});
      lens.push_back(it.yhi - it.ylo + 1);
    }
  }
  std::vector<size_t> order(items.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t a, size_t b) { return lens[a] > lens[b]; });
  std::vector<SweepItemDev> sorted(items.size());
  for (size_t i = 0; i < order.size(); ++i) sorted[i] = items[order[i]];
  // the site-field jobs of the active grains (fieldsBatchKernel)
  std::vector<int> jobs(2 * act.size() + 1, 0);
  int tiles = 0;
  for (size_t j = 0; j < act.size(); ++j) {
    jobs[j] = act[j];
    jobs[act.size() + j] = tiles;
    tiles += grains_[size_t(act[j])]->setup().fieldTileStart[size_t(
        grains_[size_t(act[j])]->setup().Na)];
  }
  jobs[2 * act.size()] = tiles;
  if (pl.dFieldJobs) (void)hipFree(pl.dFieldJobs);
  pl.dFieldJobs = nullptr;
  SC_HIP(hipMalloc(&pl.dFieldJobs, jobs.size() * sizeof(int)));
  SC_HIP(hipMemcpyAsync(pl.dFieldJobs, jobs.data(), jobs.size() * sizeof(int),
                        hipMemcpyHostToDevice, stream_));
  pl.numFieldJobs = static_cast<int>(act.size());
  pl.numFieldTiles = tiles;
  pl.numItems = static_cast<int>(sorted.size());
  if (sorted.size() > pl.itemsCap || !pl.dItems) {
    if (pl.dItems) (void)hipFree(pl.dItems);
    pl.dItems = nullptr;
    pl.itemsCap = std::max<size_t>(1, sorted.size());
    SC_HIP(hipMalloc(&pl.dItems, pl.itemsCap * sizeof(SweepItemDev)));
  }
  if (!pl.dViews) SC_HIP(hipMalloc(&pl.dViews, views.size() * sizeof(GrainView)));
  if (!sorted.empty())
    SC_HIP(hipMemcpyAsync(pl.dItems, sorted.data(), sorted.size() * sizeof(SweepItemDev),
                          hipMemcpyHostToDevice, stream_));
  SC_HIP(hipMemcpyAsync(pl.dViews, views.data(), views.size() * sizeof(GrainView),
                        hipMemcpyHostToDevice, stream_));
  int dev = 0, cus = 0;
  SC_HIP(hipGetDevice(&dev));
  SC_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  pl.numSlots = std::max(1, std::min(pl.numItems, pl.slotsPerCU * std::max(cus, 1)));
  const size_t elemBytes = precision == PREC_FP64 ? sizeof(Cx<double>) : sizeof(Cx<float>);
  const size_t bytes = size_t(pl.numSlots) * size_t(pl.maxLen) * pl.el * 64 * elemBytes;
  if (bytes > scratchBytes_) {
    if (dScratch_) (void)hipFree(dScratch_);
    dScratch_ = nullptr;
    scratchBytes_ = 0;
    SC_HIP(hipMalloc(&dScratch_, bytes));
    scratchBytes_ = bytes;
  }
  SC_HIP(hipStreamSynchronize(stream_));
  pl.key = act;
  pl.ready = true;
}

void Batch::sweep(int precision, bool burnin, const std::vector<int> &act, bool fields) {
  // the sweep reads what the grains' stages queued on their lanes
  joinLanes();
  prepare(precision, act);
  const Plan &pl = plan_[precision];
  if (fields && pl.numFieldTiles > 0) {
    // the site fields of every active grain in one launch (each grain's
    // extension of the order parameter is done on its lane)
    fieldsBatchKernel<<<xcdGrid(pl.numFieldTiles), 256, 0, stream_>>>(
        reinterpret_cast<const GrainView *>(pl.dViews), pl.dFieldJobs,
        pl.dFieldJobs + pl.numFieldJobs, pl.numFieldJobs, pl.numFieldTiles);
    SC_HIP(hipGetLastError());
  }
  SweepArgs a;
  a.views = reinterpret_cast<const GrainView *>(pl.dViews);
  a.items = reinterpret_cast<const SweepItemDev *>(pl.dItems);
  a.numItems = pl.numItems;
  a.counter = dCounter_;
  a.scratch = dScratch_;
  a.slotStride = static_cast<long long>(pl.maxLen) * pl.el * 64;
  a.burnin = burnin ? 1 : 0;
  resetCounter<<<1, 1, 0, stream_>>>(dCounter_);
  SC_HIP(hipEventRecord(ev0_[precision], stream_));
  launchSweepVariant(precision, pl.el, pl.waves, a, pl.numSlots, stream_);
  SC_HIP(hipEventRecord(ev1_[precision], stream_));
  // and the stages after it read what the sweep wrote
  forkLanes();
}

float Batch::sweepTime(int precision) const {
  float ms = 0.f;
  SC_HIP(hipEventElapsedTime(&ms, ev0_[precision], ev1_[precision]));
  return ms;
}

void Batch::compute(int precision) {
  if (precision != PREC_FP64 && precision != PREC_MIXED)
    throw std::runtime_error("unknown precision");
  computeEach(std::vector<int>(grains_.size(), precision));
}

void Batch::computeEach(const std::vector<int> &precisionOf) {
  if (precisionOf.size() != grains_.size())
    throw std::runtime_error("one precision per grain expected");
  const std::vector<int> act = active();
  if (act.empty()) return;
  // the grains of each precision: one sweep launch each
  std::vector<int> group[2];
  for (int gi : act) {
    const int p = precisionOf[size_t(gi)];
    if (p != PREC_FP64 && p != PREC_MIXED) throw std::runtime_error("unknown precision");
    group[p].push_back(gi);
  }
  SC_HIP(hipEventRecord(evIt0_, stream_));
  for (int gi : act) grains_[size_t(gi)]->queueBeginCompute();
  for (int gi : act) grains_[size_t(gi)]->queueReadback();
  for (int gi : act) grains_[size_t(gi)]->finishBeginCompute(false);
  for (int p : {PREC_FP64, PREC_MIXED})
    if (!group[p].empty()) sweep(p, false, group[p], true);
  // every grain's stages are queued before the first read-back waits, so
  // the lanes run them side by side
  for (int gi : act) grains_[size_t(gi)]->queueEndCompute();
  for (int gi : act) grains_[size_t(gi)]->queueReadback();
  for (int gi : act) grains_[size_t(gi)]->finishEndCompute();
  joinLanes();
  SC_HIP(hipEventRecord(evIt1_, stream_));
  SC_HIP(hipEventSynchronize(evIt1_));
  sweepMs_ = 0.0;
  for (int p : {PREC_FP64, PREC_MIXED})
    if (!group[p].empty()) sweepMs_ += sweepTime(p);
  float ms = 0.f;
  SC_HIP(hipEventElapsedTime(&ms, evIt0_, evIt1_));
  iterMs_ = ms;
  trajSteps_ = 0;
  for (int gi : act) trajSteps_ += grains_[size_t(gi)]->setup().trajSteps;
}

void Batch::burnin(int precision) {
  const std::vector<int> act = active();
  if (act.empty()) return;
  for (int gi : act) grains_[size_t(gi)]->beginBurnin(false);
  sweep(precision, true, act, true);
  for (int gi : act) grains_[size_t(gi)]->queueEndBurnin();
  for (int gi : act) grains_[size_t(gi)]->queueReadback();
  for (int gi : act) grains_[size_t(gi)]->finishEndBurnin();
  joinLanes();
  SC_HIP(hipStreamSynchronize(stream_));
}

int Batch::run(int nBurnin, int itMin, int itMax, int precision, std::vector<int> &converged) {
  // runSimulation (run_simulation.h:124-187) for every grain at once.
  if (precision != PREC_FP64 && precision != PREC_MIXED && precision != PREC_MIXED_REFINE)
    throw std::runtime_error("unknown precision");
  const bool refine = precision == PREC_MIXED_REFINE;
  const size_t n = grains_.size();
  // the precision of each grain's next iteration, and Grain::run's switch
  // rule per grain: FP32 stepping until the residual is below kRefineSwitch
  // or has not improved for kRefineStall iterations
  constexpr double kRefineSwitch = 1e-5;
  constexpr int kRefineStall = 20;
  std::vector<int> prec(n, refine ? PREC_MIXED : precision);
  std::vector<double> best(n, 0.0);
  std::vector<int> stall(n, 0), mixedIts(n, 0);
  for (auto &g : grains_) g->setMixedIterations(0);
  std::fill(done_.begin(), done_.end(), 0);
  for (int i = 0; i < nBurnin; ++i) burnin(refine ? PREC_MIXED : precision);
  converged.assign(n, 0);
  if (itMin == 0 && itMax == 0) return 0;
  for (int it = 1;; ++it) {
    computeEach(prec);
    for (size_t g = 0; g < n; ++g) {
      if (done_[g]) continue;
      const Results r = grains_[g]->results();
      if (prec[g] == PREC_MIXED) grains_[g]->setMixedIterations(++mixedIts[g]);
      const bool conv = r.converged != 0 && (!refine || prec[g] == PREC_FP64);
      if (refine && prec[g] == PREC_MIXED) {
        stall[g] = (it > 1 && r.residual >= best[g]) ? stall[g] + 1 : 0;
        best[g] = (it == 1) ? r.residual : std::min(best[g], r.residual);
        if (r.residual < kRefineSwitch || stall[g] >= kRefineStall || r.converged)
          prec[g] = PREC_FP64;
      }
      converged[g] = conv ? 1 : 0;
      const bool minReached = it >= itMin;
      const bool maxReached = itMax >= 0 && it >= itMax;
      if ((conv && minReached) || maxReached) done_[g] = 1;
    }
    if (active().empty()) break;
  }
  std::fill(done_.begin(), done_.end(), 0);
  int numConverged = 0;
  for (int c : converged) numConverged += c;
  return numConverged;
}

void Batch::autotune(int precision) {
  if (precision != PREC_FP64 && precision != PREC_MIXED)
    throw std::runtime_error("unknown precision");
  const std::vector<int> act = active();
  for (int gi : act) grains_[size_t(gi)]->prepareSweepInputs();
  int maxNe = 0;
  for (auto &g : grains_) maxNe = std::max(maxNe, g->setup().Ne);
  float best = -1.f;
  int bestEl = plan_[precision].el, bestW = plan_[precision].waves,
      bestSlots = plan_[precision].slotsPerCU;
  const int slotChoices[] = {4, 8, 16};
  auto tryShape = [&](int el, int w, int slots) {
    if ((maxNe + el - 1) / el > 64) return;
    setSweepShape(precision, el, w, slots);
    // the faster of two launches: one launch varies by a few per cent
    float ms = -1.f;
    for (int rep = 0; rep < 2; ++rep) {
      sweep(precision, false, act);
      SC_HIP(hipEventSynchronize(ev1_[precision]));
      const float t = sweepTime(precision);
      if (ms < 0.f || t < ms) ms = t;
    }
    if (best < 0.f || ms < best) {
      best = ms;
      bestEl = el;
      bestW = w;
      bestSlots = slots;
    }
  };
#define SC_TRY(E, W)                                                                 \
  for (int slots : slotChoices)                                                       \
    if (slots >= 4 * W || slots == 16) tryShape(E, W, slots);
  SC_SWEEP_VARIANTS(SC_TRY)
#undef SC_TRY
  setSweepShape(precision, bestEl, bestW, bestSlots);
}

void poissonSolve(int n, double h, int numGrain, bool enforce, const double *current,
                  double *out, double maxError) {
  const size_t n0 = size_t(n) * n;
  double mx = 0.0, my = 0.0;
  if (enforce) {
    for (size_t i = 0; i < n0; ++i) {
      mx += current[i];
      my += current[n0 + i];
    }
    mx /= numGrain;
    my /= numGrain;
  }
  hipStream_t stream = nullptr;
  SC_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  double *dj = nullptr, *dout = nullptr;
  try {
    PoissonFFT fft(n, h, stream, maxError);
    dj = devAlloc<double>(2 * n0 + 2);
    dout = devAlloc<double>(2 * n0);
    const double means[2] = {mx, my};
    SC_HIP(hipMemcpyAsync(dj, current, 2 * n0 * sizeof(double), hipMemcpyHostToDevice, stream));
    SC_HIP(hipMemcpyAsync(dj + 2 * n0, means, 2 * sizeof(double), hipMemcpyHostToDevice,
                          stream));
    fft.solve(dj, dj + 2 * n0, dout);
    SC_HIP(hipMemcpyAsync(out, dout, 2 * n0 * sizeof(double), hipMemcpyDeviceToHost, stream));
    SC_HIP(hipStreamSynchronize(stream));
  } catch (...) {
    devFree(dj);
    devFree(dout);
    (void)hipStreamDestroy(stream);
    throw;
  }
  devFree(dj);
  devFree(dout);
  SC_HIP(hipStreamDestroy(stream));
}

} // namespace scamd
