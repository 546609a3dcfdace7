// sc_ldos.hip -- Keldysh spectroscopy: the local density of states at real
// energies E + i delta of a converged state (src/postprocess.cu:64-472).
//
// The postprocess context iterates one block of energies at a time
// (IntegrationIteratorKeldysh.h:235-388): the Riccati sweep of every angle
// with the block's energies, the spectral weight per energy at every inside
// site, its back-rotation into the LDOS of the block, the specular boundary
// condition, and the residual of the block's LDOS against its previous
// iterate (ComputeLDOS.h:110-192).  Once both residuals are below the
// criterion the next block starts from the homogeneous boundary values.
//
// The spectral weights of all angles of a block do not fit in memory on
// large grains (720 angles x 1.7 M sites x 32 energies is 319 GB), so the
// angles are swept in chunks: sweep the chunk, extend its weights at the
// boundary, back-rotate them into per-site sums, next chunk.
#include "sc_engine.hpp"
#include "sc_device.hpp"
#include "sc_util.hpp"

#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>

namespace scamd {

// Back-rotation of the spectral weights of angles [k0, k1) (held at
// G + (k - angleBase) * N1^2 * Ne) into acc[i * Ne + e] for the inside site
// sites[i] (ComputeLDOS.h:133-151, insertRotatedGridAdditive_kernel
// GridFunctions_GPU.h:491-544): one thread per (site, energy), so that the
// threads of a site read the Ne weights of a tap together.  Angles are added
// in increasing order; the first chunk starts the sums.
__global__ void ldosGatherKernel(GrainView v, const int *sites, int numSites, int k0, int k1,
                                 int first, double *acc) {
  const int Ne = v.Ne;
  const long long total = (long long)numSites * Ne;
  const long long blk = xcdBlock((total + blockDim.x - 1) / blockDim.x);
  if (blk < 0) return;
  const long long t = blk * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int i = int(t / Ne);
  const int e = int(t - (long long)i * Ne);
  const int site = sites[i];
  const int x = site % v.N0, y = site / v.N0;
  const size_t n1 = size_t(v.N1) * v.N1;
  double sum = first ? 0.0 : acc[t];
  for (int k = k0; k < k1; ++k) {
    const RotSample r = rotatedCell(x, y, v.N0, v.N1, double(v.bkCos[k]), double(v.bkSin[k]));
    if (!r.valid) continue;
    const double *Gk = v.G + size_t(k - v.angleBase) * n1 * Ne + e;
    const size_t i00 = size_t(r.idx) * Ne, i10 = i00 + Ne, i01 = i00 + size_t(v.N1) * Ne,
                 i11 = i01 + Ne;
    double val;
    {
#pragma clang fp contract(off)
      const double wx0 = 1.0 - r.fx, wx1 = r.fx, wy0 = 1.0 - r.fy, wy1 = r.fy;
      val = wy0 * (wx0 * Gk[i00] + wx1 * Gk[i10]) + wy1 * (wx0 * Gk[i01] + wx1 * Gk[i11]);
    }
    sum += val * v.weight[k];
  }
  acc[t] = sum;
}

// The block's LDOS in the reference layout [e][N0^2] (zero outside the
// grain, ComputeLDOS.h:160-166).
__global__ void ldosFinishKernel(const int *sites, int numSites, int Ne, long long n0,
                                 const double *acc, double *interval) {
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= (long long)numSites * Ne) return;
  const int i = int(t / Ne);
  const int e = int(t - (long long)i * Ne);
  interval[size_t(e) * n0 + sites[i]] = acc[t];
}

__global__ void ldosResetCounter(int *c) { *c = 0; }

int Grain::ldosChunkAngles() const {
  const Setup &s = *s_;
  if (ldosChunkReq_ > 0) return std::min(ldosChunkReq_, s.Na);
  const size_t perAngle = size_t(s.N1) * s.N1 * size_t(s.spec.perBlock) * sizeof(double);
  size_t freeB = 0, totalB = 0;
  SC_HIP(hipMemGetInfo(&freeB, &totalB));
  // what the weights hold now is available to them again
  const size_t avail = freeB + ldosGCap_ * sizeof(double);
  const size_t budget = std::min(avail / 2, size_t(64) << 30);
  const size_t n = budget / perAngle;
  return static_cast<int>(std::max<size_t>(1, std::min<size_t>(n, size_t(s.Na))));
}

void Grain::setLdosAngleChunk(int angles) {
  if (!s_->keldysh) throw std::runtime_error("not a spectroscopy context");
  if (angles < 0) throw std::invalid_argument("angle chunk must be >= 0");
  if (angles != ldosChunkReq_) {
    ldosChunkReq_ = angles;
    plan_[0].ready = plan_[1].ready = false;
  }
}

void Grain::ldosEnsureBuffers(int precision) {
  const Setup &s = *s_;
  const size_t n0 = size_t(s.N0) * s.N0, n1 = size_t(s.N1) * s.N1;
  const size_t cap = size_t(s.spec.perBlock);
  if (!dLdosAll_) dLdosAll_ = devAlloc<double>(size_t(s.spec.numEnergies) * n0);
  if (!dLdosAcc_) dLdosAcc_ = devAlloc<double>(s.domainSites.size() * cap);
  if (!dLdosInterval_) dLdosInterval_ = devAlloc<double>(cap * n0);
  const int ca = plan_[precision].ready ? plan_[precision].chunkAngles : ldosChunkAngles();
  const size_t need = size_t(ca) * n1 * cap;
  if (need > ldosGCap_ || ca > ldosFieldFrames_) {
    SC_HIP(hipStreamSynchronize(stream_));
    devFree(dLdosG_);
    devFree(dLdosFields_);
    ldosGCap_ = 0;
    ldosFieldFrames_ = 0;
    dLdosG_ = devAlloc<double>(need); // cleared: every weight is finite
    ldosGCap_ = need;
    dLdosFields_ = devAlloc<double>(size_t(3) * size_t(ca) * n1);
    ldosFieldFrames_ = ca;
    // the device views hold the old addresses
    plan_[0].ready = plan_[1].ready = false;
  }
}

void Grain::ldosSweepChunk(int precision, int c, bool burnin) {
  // site fields of the chunk's frames, then the sweep of its items
  const Setup &s = *s_;
  const SweepPlan &pl = plan_[precision];
  const int k0 = c * pl.chunkAngles, k1 = std::min(s.Na, k0 + pl.chunkAngles);
  if (k1 - k0 > ldosFieldFrames_)
    throw std::runtime_error("angle chunk exceeds the site-field buffer");
  computeFields(k0, k1 - k0, static_cast<long long>(ldosFieldFrames_) * s.N1 * s.N1,
                dLdosFields_);
  SweepArgs a;
  a.views = reinterpret_cast<const GrainView *>(pl.dView);
  a.items = reinterpret_cast<const SweepItemDev *>(pl.dItems) + pl.chunkStart[size_t(c)];
  a.numItems = pl.chunkStart[size_t(c) + 1] - pl.chunkStart[size_t(c)];
  a.counter = dCounter_;
  a.scratch = dScratch_;
  a.slotStride = static_cast<long long>(pl.maxLen) * pl.el * 64;
  a.burnin = burnin ? 1 : 0;
  ldosResetCounter<<<1, 1, 0, stream_>>>(dCounter_);
  SC_HIP(hipGetLastError());
  if (a.numItems > 0) launchSweepVariant(precision, pl.el, pl.waves, a, pl.numSlots, stream_, true);
}

void Grain::ldosSweepAll(int precision, bool burnin) {
  // every chunk in turn (burn-in, autotuning)
  ldosEnsureBuffers(precision);
  prepareSweep(precision);
  const int nch = static_cast<int>(plan_[precision].chunkStart.size()) - 1;
  SC_HIP(hipEventRecord(ev0_, stream_));
  for (int c = 0; c < nch; ++c) ldosSweepChunk(precision, c, burnin);
  SC_HIP(hipEventRecord(ev1_, stream_));
}

void Grain::ldosBeginBlock() {
  // IntegrationIteratorKeldysh::compute on a new interval (:263-281):
  // energies of the block, homogeneous boundary values, the boundary
  // condition, compute objects not converged.
  Setup &s = *s_;
  const int oldNe = s.Ne;
  SC_HIP(hipStreamSynchronize(stream_));
  setKeldyshBlock(s, ldosBlockStart_);
  uploadEnergies();
  if (s.Ne != oldNe) plan_[0].ready = plan_[1].ready = false;
  tim_.trajSteps = s.trajSteps;
  boundaryInit_ = false;
  ensureBoundaryInit();
  ldosPartial_ = false;
  ldosLastBlock_ = ldosBlockStart_;
}

void Grain::ldosCompute(int precision, bool burnin) {
  if (!s_->keldysh) throw std::runtime_error("not a spectroscopy context");
  if (precision != PREC_FP64 && precision != PREC_MIXED)
    throw std::runtime_error("unknown precision");
  const Setup &s = *s_;
  if (ldosBlockStart_ < s.spec.numEnergies) {
    clearReductions();
    if (ldosLastBlock_ != ldosBlockStart_) ldosBeginBlock();
    SC_HIP(hipEventRecord(evIterStart_, stream_));
    extendDelta();
    if (burnin) {
      sweep(precision, true);
      SC_HIP(hipEventRecord(evIterStop_, stream_));
      boundaryCondition();
      float ms = 0.f;
      SC_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
      ldosSweepMs_ = ms;
    } else {
      ldosEnsureBuffers(precision);
      prepareSweep(precision);
      const SweepPlan &pl = plan_[precision];
      const int nch = static_cast<int>(pl.chunkStart.size()) - 1;
      const size_t n0 = size_t(s.N0) * s.N0, n1 = size_t(s.N1) * s.N1;
      const int ns = static_cast<int>(s.domainSites.size());
      if (size_t(pl.chunkAngles) * n1 * size_t(s.Ne) > ldosGCap_)
        throw std::runtime_error("angle chunk exceeds the spectral-weight buffer");
      while (ldosEvents_.size() < size_t(2 * nch)) {
        hipEvent_t e;
        SC_HIP(hipEventCreate(&e));
        ldosEvents_.push_back(e);
      }
      GrainView v = view();
      for (int c = 0; c < nch; ++c) {
        const int k0 = c * pl.chunkAngles, k1 = std::min(s.Na, k0 + pl.chunkAngles);
        SC_HIP(hipEventRecord(ldosEvents_[size_t(2 * c)], stream_));
        ldosSweepChunk(precision, c, false);
        SC_HIP(hipEventRecord(ldosEvents_[size_t(2 * c + 1)], stream_));
        // extendValuesAtBoundary on the chunk's weights, then back-rotation
        extendFrames(dLdosG_, k0, k1, s.Ne, 1, static_cast<long long>(n1) * s.Ne, s.Ne);
        v.angleBase = k0;
        const long long nt = static_cast<long long>(ns) * s.Ne;
        if (nt > 0) {
          ldosGatherKernel<<<xcdGrid(blocksFor(nt)), 256, 0, stream_>>>(v, dDomainSites_, ns, k0, k1,
                                                              c == 0 ? 1 : 0, dLdosAcc_);
          SC_HIP(hipGetLastError());
        }
      }
      int hb[2];
      queueBoundaryCondition(hb);
      // ComputeLDOS::computePropertyFromGreens (:153-192): the block's LDOS,
      // its residual against the previous iterate, then the copy into the
      // full LDOS.
      const long long nt = static_cast<long long>(ns) * s.Ne;
      if (nt > 0) {
        ldosFinishKernel<<<blocksFor(nt), 256, 0, stream_>>>(dDomainSites_, ns, s.Ne,
                                                             static_cast<long long>(n0),
                                                             dLdosAcc_, dLdosInterval_);
        SC_HIP(hipGetLastError());
      }
      double *slot = dLdosAll_ + size_t(ldosBlockStart_) * n0;
      const long long len = static_cast<long long>(s.Ne) * static_cast<long long>(n0);
      const int hl = queueStats(slot, dLdosInterval_, len);
      SC_HIP(hipMemcpyAsync(slot, dLdosInterval_, size_t(len) * sizeof(double),
                            hipMemcpyDeviceToDevice, stream_));
      SC_HIP(hipEventRecord(evIterStop_, stream_));
      flushReductions();
      res_.resBoundary = boundaryResidual(hb);
      ldosRes_ = reductionResidual(hl);
      double sweepMs = 0.0;
      for (int c = 0; c < nch; ++c) {
        float ms = 0.f;
        SC_HIP(hipEventElapsedTime(&ms, ldosEvents_[size_t(2 * c)], ldosEvents_[size_t(2 * c + 1)]));
        sweepMs += ms;
      }
      ldosSweepMs_ = sweepMs;
      ldosPartial_ = ldosRes_ < s.cfg.convergence;
      const bool boundaryConverged = res_.resBoundary < s.cfg.convergence;
      if (ldosPartial_ && boundaryConverged) {
        ldosBlockStart_ += s.spec.perBlock;
        if (ldosBlockStart_ >= s.spec.numEnergies) ldosConverged_ = true;
      }
    }
    float it = 0.f;
    SC_HIP(hipEventElapsedTime(&it, evIterStart_, evIterStop_));
    ldosIterMs_ = it;
  }
  ++ldosIterations_;
}

bool Grain::ldosRun(int burnin, int itMin, int itMax, int precision) {
  // runPostprocess (run_simulation.h:44-108) with runIterations
  // (run_iteration.h:85-141) for the burn-in and the compute phases.
  const Setup &s = *s_;
  auto iterations = [&](bool compute, int mn, int mx) {
    if (mn == 0 && mx == 0) return false;
    bool conv = false;
    for (int i = 1;; ++i) {
      ldosCompute(precision, !compute);
      const LdosStatus st = ldosStatus();
      conv = compute ? st.partiallyConverged : res_.resBoundary < s.cfg.convergence;
      const bool minReached = i >= mn;
      const bool maxReached = mx < 0 ? false : i >= mx;
      if ((conv && minReached) || maxReached) break;
    }
    return conv;
  };
  bool converged = false;
  while (!converged) {
    iterations(false, burnin, burnin);
    const bool blockConverged = iterations(true, itMin, itMax);
    converged = ldosStatus().converged;
    if (!blockConverged) break;
  }
  return converged;
}

LdosStatus Grain::ldosStatus() const {
  const Setup &s = *s_;
  LdosStatus st;
  st.numEnergies = s.spec.numEnergies;
  st.perBlock = s.spec.perBlock;
  st.blockStart = ldosBlockStart_;
  st.blockCount = s.Ne;
  st.iterations = ldosIterations_;
  st.angleChunk = plan_[0].ready ? plan_[0].chunkAngles
                                 : (plan_[1].ready ? plan_[1].chunkAngles : ldosChunkReq_);
  // Context::getResidual / isPartiallyConverged / isConverged (Context.h:462-511)
  st.residualBoundary = res_.resBoundary;
  st.residualLdos = ldosRes_;
  st.residual = std::max(res_.resBoundary, ldosRes_);
  st.partiallyConverged = st.residual < s.cfg.convergence && ldosPartial_;
  st.converged = st.residual < s.cfg.convergence && ldosConverged_;
  st.sweepMs = ldosSweepMs_;
  st.iterationMs = ldosIterMs_;
  st.trajSteps = s.trajSteps;
  return st;
}

void Grain::getLdos(double *out) {
  if (!s_->keldysh) throw std::runtime_error("not a spectroscopy context");
  const Setup &s = *s_;
  const size_t n = size_t(s.spec.numEnergies) * size_t(s.N0) * s.N0;
  if (!dLdosAll_) {
    std::fill(out, out + n, 0.0);
    return;
  }
  fromDevice(out, dLdosAll_, n * sizeof(double));
}

} // namespace scamd
