// sc_kernels.hpp -- HIP kernels of the self-consistency iteration (gfx950).
#pragma once

#include "sc_device.hpp"
#include "sc_view.hpp"

namespace scamd {

// Order parameter and Doppler shift at site (x, y) of frame k, from the
// extended order parameter and the vector potential on the N0 grid
// (OrderParameter.h:232-269, RiccatiSolver.h:127-176,
// VectorPotential_Kernel.h:21-99, Gauges.h:25-62).
__device__ __forceinline__ void siteFields(const GrainView &v, int k, int x,
                                           int y, double &dre, double &dim,
                                           double &es) {
  const int N0 = v.N0, N1 = v.N1;
  const size_t n0 = size_t(N0) * N0;
  dre = 0.0;
  dim = 0.0;
  const RotSample r =
      rotatedCell(x, y, N1, N0, double(global(v.opCos)[k]), double(global(v.opSin)[k]));
  if (r.valid) {
    for (int c = 0; c < v.Nc; ++c) {
      const double b = global(v.basis)[c * v.Na + k];
      const double *p = v.extDelta + size_t(c) * 2 * n0;
      dre += b * bilinear(p, N0, r);
      dim += b * bilinear(p + n0, N0, r);
    }
  }
  es = 0.0;
  if (v.useField) {
    double ax = 0.0, ay = 0.0;
    if (v.useInduced) {
      const RotSample q =
          rotatedCell(x, y, N1, N0, double(global(v.aCos)[k]), double(global(v.aSin)[k]));
      if (q.valid) {
        ax = bilinear(v.A, N0, q) * v.lambdaInv2;
        ay = bilinear(v.A + n0, N0, q) * v.lambdaInv2;
      }
    }
    const double off = -0.5 * double(N1 - 1);
    const double px0 = (double(x) + off) * v.h, py0 = (double(y) + off) * v.h;
    const double c = global(v.saCos)[k], s = global(v.saSin)[k];
    const double px = px0 * c - py0 * s, py = px0 * s + py0 * c;
    double ex, ey;
    if (v.gauge == 0) {
      ex = 0.0;
      ey = px * v.fluxDensity;
    } else if (v.gauge == 1) {
      const double f = 0.5 * v.fluxDensity;
      ex = -py * f;
      ey = px * f;
    } else {
      const double r2 = fmax(px * px + py * py, 1e-7);
      const double f = v.fluxDensity * v.area / (2.0 * M_PI * r2);
      ex = -py * f;
      ey = px * f;
    }
    es = double(v.charge) * (global(v.vFx)[k] * (ax + ex) + global(v.vFy)[k] * (ay + ey));
  }
}

__global__ void resetCounter(int *c) { *c = 0; }

// The site fields a sweep reads, for the tiles of frames [k0, ...): each
// block covers one tile (16 adjacent columns of one frame and the rows a
// sweep may read there), 16 columns x 16 rows per pass, so that the writes
// of a row of the tile are contiguous.  F is planar [3][frames][N1^2] with
// plane stride `plane`; frame k is at (k - k0) * N1^2.
__device__ __forceinline__ void fieldsTile(const GrainView &v, const int *tl, int k0,
                                           long long plane, double *Fout) {
  const GlobalPtr<const int> tg = global(tl);
  const int k = tg[0], x = tg[1] + int(threadIdx.x & 15), y0 = tg[2], y1 = tg[3];
  const GlobalPtr<double> F = global(Fout);
  const int N1 = v.N1;
  if (x >= N1) return;
  const size_t base = size_t(k - k0) * N1 * N1 + x;
  for (int y = y0 + int(threadIdx.x >> 4); y <= y1; y += 16) {
    double re, im, es;
    siteFields(v, k, x, y, re, im, es);
    const size_t i = base + size_t(y) * N1;
    F[i] = re;
    F[plane + i] = im;
    F[2 * plane + i] = es;
  }
}
__global__ void fieldsKernel(GrainView v, const int *tiles, int numTiles, int k0,
                             long long plane, double *F) {
  const long long tile = xcdBlock(numTiles);
  if (tile < 0) return;
  fieldsTile(v, tiles + 4 * size_t(tile), k0, plane, F);
}
// The site fields of several grains in one launch (a batch): grain j of the
// job list (view grainOf[j]) owns the blocks [tileStart[j], tileStart[j+1])
// of the numTiles in total, one tile each, every frame of the grain.
__global__ void fieldsBatchKernel(const GrainView *__restrict__ views,
                                  const int *__restrict__ grainOf,
                                  const int *__restrict__ tileStart, int numJobs, int numTiles) {
  const long long b = xcdBlock(numTiles);
  if (b < 0) return;
  int lo = 0, hi = numJobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tileStart[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  const GrainView &v = views[grainOf[lo]];
  // the view's site-field buffer is the grain's own, written here and read
  // by the sweep that follows
  fieldsTile(v, v.fieldTiles + 4 * (b - tileStart[lo]), 0, v.fieldPlane,
             const_cast<double *>(v.fields));
}

// One step for one energy: Matsubara energies (i omega) or, with KEL,
// real-frequency energies E + i delta.
template <bool KEL, typename T>
__device__ __forceinline__ Cx<T> stepE(const StepSite<T> &p, T ser, T sej, Cx<T> y0) {
  if constexpr (KEL) return stepEnergy(p, ser, sej, y0);
  else return stepEnergy(p, ser, y0);
}

// The Riccati sweep.  One wavefront handles an item: `cols` adjacent columns
// of one frame; each column is carried by `groups` lanes holding EL Ozaki
// energies each (lane = group * cols + column; the 64 - groups * cols lanes
// past the last group idle).  Pass 1 integrates gamma~ from the top of the
// item to the bottom and keeps it in the slot's scratch; pass 2 integrates
// gamma upwards and forms the energy-summed Green functions at every inside
// site.  Boundary in-values are read and out-values written as in
// RiccatiSolverConfined.h:129-312.
//
// The order parameter and Doppler shift of a row are computed once per
// column: the lanes of a column take consecutive rows of a chunk and pass
// the values round with __shfl.
template <typename T, int EL, bool KEL>
__device__ __forceinline__ void sweepItem(const SweepArgs &a, const GrainView &v,
                                          const SweepItemDev &it, Cx<T> *scr) {
  const int lane = threadIdx.x;
  const int cols = v.cols;
  const int gp = v.groups; // lanes per column
  const int g = lane / cols;
  const int c = lane - g * cols;
  const int N1 = v.N1, Ne = v.Ne;
  const size_t n1 = size_t(N1) * N1;
  const T h = T(v.h);
  const int k = it.angle;
  const bool colActive = c < it.ncols;
  const int x = it.x0 + (colActive ? c : it.ncols - 1);
  const bool active = colActive && (g < v.groups);
  const GlobalPtr<const int8_t> lab = global(v.labels + size_t(k) * n1 + x);
  T ser[EL], sej[EL], res[EL];
  bool eOk[EL];
#pragma unroll
  for (int j = 0; j < EL; ++j) {
    const int e = g * EL + j;
    eOk[j] = active && e < Ne;
    ser[j] = eOk[j] ? -h * T(global(v.energy)[e]) : T(0);
    sej[j] = (KEL && eOk[j]) ? h * T(global(v.energyRe)[e]) : T(0);
    res[j] = eOk[j] ? T(global(v.residue)[e]) : T(0);
  }
  const int len = it.yhi - it.ylo + 1;
  const size_t col = size_t(k) * N1 + x;

  // ---------------- pass 1: gamma~ along -y ----------------
  {
    const int baseIn = global(v.idxInGt)[col], baseOut = global(v.idxOutGt)[col];
    int cntIn = 0, cntOut = 0;
    Cx<T> coh[EL];
#pragma unroll
    for (int j = 0; j < EL; ++j) coh[j] = {T(0), T(0)};
    Cx<T> dPrev = {T(0), T(0)};
    T esPrev = T(0);
    bool havePrev = false;
    int8_t lcur = lab[size_t(it.yhi) * N1];
    // the fields and label of the lane's row of the next chunk, loaded while
    // the current chunk is swept
    double nre = 0.0, nim = 0.0, nes = 0.0;
    int nlab = 0;
    auto loadChunk = [&](int t0) {
      const int tt = t0 + g;
      if (colActive && g < gp && tt < len) {
        const int y = it.yhi - tt;
        const size_t fi = size_t(k - v.fieldBase) * n1 + size_t(y) * N1 + x;
        nre = global(v.fields)[fi];
        nim = global(v.fields)[v.fieldPlane + fi];
        nes = global(v.fields)[2 * v.fieldPlane + fi];
        nlab = lab[size_t(y) * N1];
      } else {
        nre = nim = nes = 0.0;
        nlab = 0;
      }
    };
    loadChunk(0);
    for (int t0 = 0; t0 < len; t0 += gp) {
      const double fre = nre, fim = nim, fes = nes;
      const int flab = nlab;
      const bool more = t0 + gp < len;
      if (more) loadChunk(t0 + gp);
      // at least one site (t0 < len, gp >= 1): with no zero-trip path every
      // way to the next chunk passes a site's memory operations, and the
      // wait for this chunk's loads there is counted past them
      const int tEnd = max(1, min(gp, len - t0));
      for (int i = 0; i < tEnd; ++i) {
        const int src = i * cols + c;
        const double dre = __shfl(fre, src), dim = __shfl(fim, src), e = __shfl(fes, src);
        const int t = t0 + i;
        // the next site's label: in this chunk, or the first row of the next
        const int ln = i + 1 < tEnd ? __shfl(flab, src + cols) : (more ? __shfl(nlab, c) : 0);
        const int8_t lnext = int8_t(ln);
        const Cx<T> D = {T(-dre), T(dim)}; // -conj(Delta) for gamma~
        const T es = T(e);
        if (active) {
          if (lcur == 1 || lcur == 2) {
            // inside (1) or exit of gamma~ (2): midpoint step
            const Cx<T> Dm = {(dPrev.re + D.re) * T(0.5), (dPrev.im + D.im) * T(0.5)};
            const StepSite<T> ps = stepSite(Dm, (esPrev + es) * T(0.5), h);
            const T half = havePrev ? T(1) : T(0.5);
#pragma unroll
            for (int j = 0; j < EL; ++j) coh[j] = stepE<KEL>(ps, ser[j] * half, sej[j] * half, coh[j]);
            if (lcur == 2) {
              const int ci = baseOut + cntOut;
              const T dr = T(global(v.distOutGt)[ci]), d = dr * h;
              const StepSite<T> pe = stepSite(D, es, d);
#pragma unroll
              for (int j = 0; j < EL; ++j) {
                if (!eOk[j]) continue;
                const Cx<T> bv = stepE<KEL>(pe, ser[j] * dr, sej[j] * dr, coh[j]);
                const size_t si = size_t(ci) * Ne + g * EL + j;
                global(v.bOutGt)[si] = double(bv.re);
                global(v.bOutGt)[size_t(v.nOutGt) * Ne + si] = double(bv.im);
              }
              ++cntOut;
            }
          } else if (lcur == 3) {
            // entry of gamma~: start from the boundary value
            const int ci = baseIn + cntIn;
            const T dr = T(global(v.distInGt)[ci]), d = dr * h;
            const StepSite<T> pe = stepSite(D, es, d);
#pragma unroll
            for (int j = 0; j < EL; ++j) {
              const size_t si = size_t(ci) * Ne + g * EL + j;
              const GlobalPtr<const double> bin = global(v.bInGt);
              const Cx<T> b0 = eOk[j] ? Cx<T>{T(bin[si]), T(bin[size_t(v.nInGt) * Ne + si])}
                                      : Cx<T>{T(0), T(0)};
              coh[j] = stepE<KEL>(pe, ser[j] * dr, sej[j] * dr, b0);
            }
            ++cntIn;
          } else {
#pragma unroll
            for (int j = 0; j < EL; ++j) coh[j] = {T(0), T(0)};
          }
        }
        // gamma~ into the slot's scratch on every path through a site (the
        // second pass reads it only where it is inside and not burning in):
        // the wait for the next chunk's loads is then counted past these
        // stores.  A burn-in sweep reads none, and all its lanes store into
        // one element.
        {
          const size_t at = a.burnin ? 0 : size_t(t) * EL;
          const int ln = a.burnin ? 0 : lane;
#pragma unroll
          for (int j = 0; j < EL; ++j)
            storeStream(scr + (at + (a.burnin ? 0 : j)) * 64 + ln, coh[j]);
        }
        dPrev = D;
        esPrev = es;
        havePrev = true;
        lcur = lnext;
      }
    }
  }

  // ---------------- pass 2: gamma along +y, Green functions ----------------
  {
    const int baseIn = global(v.idxInG)[col], baseOut = global(v.idxOutG)[col];
    int cntIn = 0, cntOut = 0;
    Cx<T> coh[EL];
#pragma unroll
    for (int j = 0; j < EL; ++j) coh[j] = {T(0), T(0)};
    Cx<T> dPrev = {T(0), T(0)};
    T esPrev = T(0);
    bool havePrev = false;
    int8_t lcur = lab[size_t(it.ylo) * N1];
    // gamma~ of the next site, loaded one site ahead of its use
    Cx<T> gtNext[EL];
    const bool green = active && !a.burnin;
#pragma unroll
    for (int j = 0; j < EL; ++j) gtNext[j] = {T(0), T(0)};
    if (green && lcur > 0) {
#pragma unroll
      for (int j = 0; j < EL; ++j)
        gtNext[j] = loadStream(scr + (size_t(len - 1) * EL + j) * 64 + lane);
    }
    double nre = 0.0, nim = 0.0, nes = 0.0;
    int nlab = 0;
    auto loadChunk = [&](int t0) {
      const int tt = t0 + g;
      if (colActive && g < gp && tt < len) {
        const int y = it.ylo + tt;
        const size_t fi = size_t(k - v.fieldBase) * n1 + size_t(y) * N1 + x;
        nre = global(v.fields)[fi];
        nim = global(v.fields)[v.fieldPlane + fi];
        nes = global(v.fields)[2 * v.fieldPlane + fi];
        nlab = lab[size_t(y) * N1];
      } else {
        nre = nim = nes = 0.0;
        nlab = 0;
      }
    };
    loadChunk(0);
    for (int t0 = 0; t0 < len; t0 += gp) {
      const double fre = nre, fim = nim, fes = nes;
      const int flab = nlab;
      const bool more = t0 + gp < len;
      if (more) loadChunk(t0 + gp);
      // at least one site (t0 < len, gp >= 1): with no zero-trip path every
      // way to the next chunk passes a site's memory operations, and the
      // wait for this chunk's loads there is counted past them
      const int tEnd = max(1, min(gp, len - t0));
      for (int i = 0; i < tEnd; ++i) {
        const int src = i * cols + c;
        const double dre = __shfl(fre, src), dim = __shfl(fim, src), e = __shfl(fes, src);
        const int t = t0 + i;
        const int y = it.ylo + t;
        const int ln = i + 1 < tEnd ? __shfl(flab, src + cols) : (more ? __shfl(nlab, c) : 0);
        const int8_t lnext = int8_t(ln);
        const Cx<T> D = {T(dre), T(dim)};
        const T es = T(e);
        Cx<T> gtCur[EL];
#pragma unroll
        for (int j = 0; j < EL; ++j) gtCur[j] = gtNext[j];
        // Every path through a site issues the same vector memory operations
        // (this prefetch, then the Green record's two stores): the wait for
        // the prefetched gamma~ is then counted past the stores that follow
        // it and leaves them in flight.  A branch around either would make
        // the count path-dependent, and the wait would also hold for the
        // store's completion at every site.  Sites that need no gamma~ read
        // a clamped, valid index of the slot's scratch.
        {
          const size_t tn = size_t(max(len - 2 - t, 0));
#pragma unroll
          for (int j = 0; j < EL; ++j) gtNext[j] = loadStream(scr + (tn * EL + j) * 64 + lane);
        }
        if (active) {
          if (lcur == 1 || lcur == 3) {
            const Cx<T> Dm = {(dPrev.re + D.re) * T(0.5), (dPrev.im + D.im) * T(0.5)};
            const StepSite<T> ps = stepSite(Dm, (esPrev + es) * T(0.5), h);
            const T half = havePrev ? T(1) : T(0.5);
#pragma unroll
            for (int j = 0; j < EL; ++j) coh[j] = stepE<KEL>(ps, ser[j] * half, sej[j] * half, coh[j]);
            if (lcur == 3) {
              const int ci = baseOut + cntOut;
              const T dr = T(global(v.distOutG)[ci]), d = dr * h;
              const StepSite<T> pe = stepSite(D, es, d);
#pragma unroll
              for (int j = 0; j < EL; ++j) {
                if (!eOk[j]) continue;
                const Cx<T> bv = stepE<KEL>(pe, ser[j] * dr, sej[j] * dr, coh[j]);
                const size_t si = size_t(ci) * Ne + g * EL + j;
                global(v.bOutG)[si] = double(bv.re);
                global(v.bOutG)[size_t(v.nOutG) * Ne + si] = double(bv.im);
              }
              ++cntOut;
            }
          } else if (lcur == 2) {
            const int ci = baseIn + cntIn;
            const T dr = T(global(v.distInG)[ci]), d = dr * h;
            const StepSite<T> pe = stepSite(D, es, d);
#pragma unroll
            for (int j = 0; j < EL; ++j) {
              const size_t si = size_t(ci) * Ne + g * EL + j;
              const GlobalPtr<const double> bin = global(v.bInG);
              const Cx<T> b0 = eOk[j] ? Cx<T>{T(bin[si]), T(bin[size_t(v.nInG) * Ne + si])}
                                      : Cx<T>{T(0), T(0)};
              coh[j] = stepE<KEL>(pe, ser[j] * dr, sej[j] * dr, b0);
            }
            ++cntIn;
          } else {
#pragma unroll
            for (int j = 0; j < EL; ++j) coh[j] = {T(0), T(0)};
          }
        }
        // Green functions at inside sites (GreensFunctionStatic.h:125-157,
        // LocalGreens.h:32-70, ComputeFreeEnergy.h:95-160).  Per energy, with
        // p = gamma gamma~ and F = 2 i / (1 + p):
        //   f + f~* = F (gamma~ - gamma) [real part], -F (gamma + gamma~) [imag]
        //   Re g = Im((1 - p) / (1 + p)) = -2 Im p / |1 + p|^2
        //   Re i(D* gamma - D gamma~) summed as D.re Su + D.im Sw with
        //   Su = sum r Im(gamma~ - gamma), Sw = sum r Re(gamma + gamma~).
        if constexpr (KEL) {
          // Spectral weight per energy (ComputeLDOS.h:126-151 with
          // LocalGreens.h:32-45): -Im(i X) with X = -(1 - p) / (1 + p),
          // i.e. Re((1 - p) / (1 + p)) = (1 - |p|^2) / |1 + p|^2.
          // Every lane stores its EL weights on every path, as the Green
          // records below; weights with no site or energy to go to land in
          // the consumed scratch.
          const bool record = !a.burnin && lcur > 0 && active;
          double *const row =
              v.G + (size_t(k - v.angleBase) * n1 + size_t(y) * N1 + x) * Ne + g * EL;
          double *const sink = reinterpret_cast<double *>(scr + size_t(len - 1) * EL * 64);
#pragma unroll
          for (int j = 0; j < EL; ++j) {
            double wgt = 0.0;
            if (record && eOk[j]) {
              const Cx<T> p = coh[j] * gtCur[j];
              const T pr = T(1) + p.re;
              const T den = pr * pr + p.im * p.im;
              wgt = double((T(1) - (p.re * p.re + p.im * p.im)) * rcpT(den));
            }
            *global(record && eOk[j] ? row + j : sink + j) = wgt;
          }
        } else {
          const bool record = !a.burnin && lcur > 0;
          double gdr = 0.0, gdi = 0.0, gj = 0.0, su = 0.0, sw = 0.0, gf = 0.0;
          if (record && active) {
#pragma unroll
            for (int j = 0; j < EL; ++j) {
              if (!eOk[j]) continue;
              const Cx<T> ga = coh[j];
              const Cx<T> gt = gtCur[j];
              const Cx<T> p = ga * gt;
              const T pr = T(1) + p.re;
              const T r2 = T(2) * rcpT(pr * pr + p.im * p.im);
              const Cx<T> F = {p.im * r2, pr * r2};
              const Cx<T> u = gt - ga;
              const Cx<T> w = ga + gt;
              const double r = double(res[j]);
              gdr += r * double(F.re * u.re - F.im * u.im);
              gdi -= r * double(F.re * w.im + F.im * w.re);
              gj -= r * double(p.im * r2);
              su += r * double(u.im);
              sw += r * double(w.re);
            }
          }
          if (record) {
            gf = dre * su + dim * sw;
            // sum the energy groups of this column (lanes c, c + cols, ...)
            // into lane c: a butterfly when the groups fill the wavefront,
            // else the other groups' lanes read in turn
            if (gp * cols == 64) {
              for (int m = cols; m < 64; m <<= 1) {
                gdr += __shfl_xor(gdr, m);
                gdi += __shfl_xor(gdi, m);
                gj += __shfl_xor(gj, m);
                gf += __shfl_xor(gf, m);
              }
            } else {
              const double g0r = gdr, g0i = gdi, g0j = gj, g0f = gf;
              for (int m = 1; m < gp; ++m) {
                const int src = c + m * cols;
                gdr += __shfl(g0r, src);
                gdi += __shfl(g0i, src);
                gj += __shfl(g0j, src);
                gf += __shfl(g0f, src);
              }
            }
          }
          // lanes with no record to write store into the slot scratch of
          // the first gamma~ this pass consumed (site len - 1, read before
          // the first store and never again)
          typedef typename VecOf<double>::type D2;
          const bool real = record && g == 0 && colActive;
          D2 *dst = real ? reinterpret_cast<D2 *>(v.G + (size_t(k) * n1 + size_t(y) * N1 + x) * 4)
                         : reinterpret_cast<D2 *>(scr + size_t(len - 1) * EL * 64);
          const GlobalPtr<D2> Gk = global(dst);
          Gk[0] = D2{gdr, gdi};
          Gk[1] = D2{gj, gf};
        }
        dPrev = D;
        esPrev = es;
        havePrev = true;
        lcur = lnext;
      }
    }
  }
}

// Persistent wavefronts pulling items from the work counter.  The second
// launch bound asks the compiler for a register budget that leaves room for
// WAVES wavefronts per SIMD.
// The views and items come as restrict-qualified parameters as well: the
// sweep never writes them, and with that known the compiler reads them with
// scalar loads once per item (in SGPRs, uniform across the wavefront) and
// reaches the arrays they point at with global (not flat) loads, instead of
// reloading each field after every store the sweep makes.
template <typename T, int EL, int WAVES, bool KEL>
__global__ __launch_bounds__(64, WAVES) void sweepKernel(SweepArgs a,
                                                          const GrainView *__restrict__ views,
                                                          const SweepItemDev *__restrict__ items) {
  Cx<T> *scr = reinterpret_cast<Cx<T> *>(a.scratch) + size_t(blockIdx.x) * a.slotStride;
  for (;;) {
    // every lane takes part in the fetch (lane 0 adds one, the others
    // zero): no branch on the lane index surrounds it, so the compiler
    // cannot merge it with a lane-0 branch of the item into a loop that
    // some lanes re-enter without lane 0.  Lane 0's old value is the item,
    // uniform across the wavefront (the item and its grain view are read
    // with scalar loads).
    const int item =
        __builtin_amdgcn_readfirstlane(atomicAdd(a.counter, threadIdx.x == 0 ? 1 : 0));
    if (item >= a.numItems) break;
    const SweepItemDev it = items[item];
    sweepItem<T, EL, KEL>(a, views[it.grain], it, scr);
  }
}

// Homogeneous coherence functions at the exit crossings of one direction
// (RiccatiSolverConfined.h:400-468, 673-767): one thread per (crossing,
// energy); crossing i lies in column col[i] = k * N1 + x at row y[i].  For
// gamma the out storage takes gamma, for gamma~ it takes gamma~.
__global__ void initialOutKernel(GrainView v, const int *col, const int *yOf,
                                 long long nOut, int tilde, double *bOut) {
  const long long tid = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const int Ne = v.Ne;
  if (tid >= nOut * Ne) return;
  const long long i = tid / Ne;
  const int e = int(tid - i * Ne);
  const int c = col[i];
  const int k = c / v.N1, x = c - k * v.N1, y = yOf[i];
  double dre, dim, es;
  siteFields(v, k, x, y, dre, dim, es);
  Cx<double> ga, gt;
  const double ere = v.energyRe ? v.energyRe[e] : 0.0;
  homogeneous(Cx<double>{dre, dim}, Cx<double>{ere + es, v.energy[e]}, ga, gt);
  const Cx<double> b = tilde ? gt : ga;
  bOut[tid] = b.re;
  bOut[nOut * Ne + tid] = b.im;
}

// Extension of values outwards from the boundary
// (geometry/domain_ops.h:11-57): target <- mean of its listed sources, for
// `planes` values per site.  A target index t addresses block t / siteStride
// and site t % siteStride; value q of a site lives at
// base + (block - blockBase) * blockStride + site * siteScale + q * planeStride.
__global__ void extendKernel(double *base, const int *target, const int *offset,
                             const int *source, int numTargets, int planes,
                             long long planeStride, long long siteStride,
                             long long blockStride, long long siteScale,
                             long long blockBase) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= numTargets) return;
  const long long tg = target[t];
  const long long blk = tg / siteStride;
  const long long site = tg - blk * siteStride;
  double *p0 = base + (blk - blockBase) * blockStride;
  const int o0 = offset[t], o1 = offset[t + 1];
  const double inv = double(o1 - o0);
  for (int q = 0; q < planes; ++q) {
    double *p = p0 + q * planeStride;
    double val = 0.0;
    for (int o = o0; o < o1; ++o) {
      const long long s = source[o] - blk * siteStride;
      val = (o == o0) ? p[s * siteScale] : val + p[s * siteScale];
    }
    p[site * siteScale] = val / inv;
  }
}

// The contribution of frame k to the angle sums of reference site (x, y):
// the bilinear back-rotation of the frame's four Green-function planes,
// weighted by the basis functions and the angle weight (gatherKernel and
// gatherRangeKernel share it so that both add the same terms in the same
// order).
__device__ __forceinline__ void gatherAngle(const GrainView &v, int k, int x, int y,
                                            double (&dre)[4], double (&dim)[4], double &jx,
                                            double &jy, double &fe) {
  const int N0 = v.N0, N1 = v.N1, Na = v.Na, Nc = v.Nc;
  const size_t n1 = size_t(N1) * N1;
  const RotSample r = rotatedCell(x, y, N0, N1, double(v.bkCos[k]), double(v.bkSin[k]));
  if (!r.valid) return;
  // four values per site: taps (x, y), (x+1, y), (x, y+1), (x+1, y+1)
  const double2 *Gk = reinterpret_cast<const double2 *>(v.G + size_t(k) * n1 * 4);
  const size_t i00 = size_t(r.idx) * 2, i10 = i00 + 2, i01 = i00 + size_t(N1) * 2,
               i11 = i01 + 2;
  const double2 a00 = Gk[i00], b00 = Gk[i00 + 1];
  const double2 a10 = Gk[i10], b10 = Gk[i10 + 1];
  const double2 a01 = Gk[i01], b01 = Gk[i01 + 1];
  const double2 a11 = Gk[i11], b11 = Gk[i11 + 1];
  double gre, gim, gj, gf;
  {
#pragma clang fp contract(off)
    const double wx0 = 1.0 - r.fx, wx1 = r.fx, wy0 = 1.0 - r.fy, wy1 = r.fy;
    gre = wy0 * (wx0 * a00.x + wx1 * a10.x) + wy1 * (wx0 * a01.x + wx1 * a11.x);
    gim = wy0 * (wx0 * a00.y + wx1 * a10.y) + wy1 * (wx0 * a01.y + wx1 * a11.y);
    gj = wy0 * (wx0 * b00.x + wx1 * b10.x) + wy1 * (wx0 * b01.x + wx1 * b11.x);
    gf = wy0 * (wx0 * b00.y + wx1 * b10.y) + wy1 * (wx0 * b01.y + wx1 * b11.y);
  }
  const double w = v.weight[k];
  for (int c = 0; c < Nc && c < 4; ++c) {
    const double b = v.basis[c * Na + k];
    dre[c] += (gre * b) * w;
    dim[c] += (gim * b) * w;
  }
  jx += (v.vFx[k] * w) * gj;
  jy += (v.vFy[k] * w) * gj;
  fe += w * gf;
}

// The gap equation's, the current's and the free-energy density's final
// scaling of the angle sums of one site.
__device__ __forceinline__ void gatherStore(const GrainView &v, int site,
                                            const double (&dre)[4], const double (&dim)[4],
                                            double jx, double jy, double fe,
                                            const double *coupling, double currentFactor,
                                            double *deltaOut, double *currentOut,
                                            double *energyDensity) {
  const size_t n0 = size_t(v.N0) * v.N0;
  for (int c = 0; c < v.Nc && c < 4; ++c) {
    deltaOut[size_t(c) * 2 * n0 + site] = dre[c] * coupling[c];
    deltaOut[size_t(c) * 2 * n0 + n0 + site] = dim[c] * coupling[c];
  }
  currentOut[site] = jx * currentFactor;
  currentOut[n0 + site] = jy * currentFactor;
  energyDensity[site] = fe;
}

// Back-rotation of the Green functions of every frame to the reference grid
// and the angle integration (ComputeOrderParameter.h:174-262,
// ComputeCurrent.h:142-222, ComputeFreeEnergy.h:260-279), one thread per
// inside site of the reference grain.
__global__ void gatherKernel(GrainView v, const int *sites, int numSites,
                             const double *coupling, double currentFactor,
                             double *deltaOut, double *currentOut,
                             double *energyDensity) {
  const long long blk = xcdBlock((numSites + blockDim.x - 1) / blockDim.x);
  if (blk < 0) return;
  const int t = int(blk * blockDim.x + threadIdx.x);
  if (t >= numSites) return;
  const int N0 = v.N0;
  const int site = sites[t];
  const int x = site % N0, y = site / N0;
  double dre[4] = {0, 0, 0, 0}, dim[4] = {0, 0, 0, 0};
  double jx = 0.0, jy = 0.0, fe = 0.0;
  for (int k = 0; k < v.Na; ++k) gatherAngle(v, k, x, y, dre, dim, jx, jy, fe);
  gatherStore(v, site, dre, dim, jx, jy, fe, coupling, currentFactor, deltaOut, currentOut,
              energyDensity);
}

// The angle sums of frames [k0, k1) added to running sums, one thread per
// inside site: acc holds, per inside site t, Re and Im of every component's
// sum, then jx, jy and the free-energy density (at acc[q * numSites + t]).
// Called for consecutive ranges from k0 = 0 it adds the same terms in the
// same order as gatherKernel.
__global__ void gatherRangeKernel(GrainView v, const int *sites, int numSites, int k0, int k1,
                                  double *acc) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= numSites) return;
  const int N0 = v.N0, Nc = v.Nc;
  const int site = sites[t];
  const int x = site % N0, y = site / N0;
  const size_t ns = size_t(numSites);
  double dre[4] = {0, 0, 0, 0}, dim[4] = {0, 0, 0, 0};
  for (int c = 0; c < Nc && c < 4; ++c) {
    dre[c] = acc[size_t(2 * c) * ns + t];
    dim[c] = acc[size_t(2 * c + 1) * ns + t];
  }
  double jx = acc[size_t(2 * Nc) * ns + t], jy = acc[size_t(2 * Nc + 1) * ns + t];
  double fe = acc[size_t(2 * Nc + 2) * ns + t];
  for (int k = k0; k < k1; ++k) gatherAngle(v, k, x, y, dre, dim, jx, jy, fe);
  for (int c = 0; c < Nc && c < 4; ++c) {
    acc[size_t(2 * c) * ns + t] = dre[c];
    acc[size_t(2 * c + 1) * ns + t] = dim[c];
  }
  acc[size_t(2 * Nc) * ns + t] = jx;
  acc[size_t(2 * Nc + 1) * ns + t] = jy;
  acc[size_t(2 * Nc + 2) * ns + t] = fe;
}

// The final scaling of the running sums of gatherRangeKernel.
__global__ void gatherFinishKernel(GrainView v, const int *sites, int numSites,
                                   const double *acc, const double *coupling,
                                   double currentFactor, double *deltaOut, double *currentOut,
                                   double *energyDensity) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= numSites) return;
  const int Nc = v.Nc;
  const size_t ns = size_t(numSites);
  double dre[4] = {0, 0, 0, 0}, dim[4] = {0, 0, 0, 0};
  for (int c = 0; c < Nc && c < 4; ++c) {
    dre[c] = acc[size_t(2 * c) * ns + t];
    dim[c] = acc[size_t(2 * c + 1) * ns + t];
  }
  gatherStore(v, sites[t], dre, dim, acc[size_t(2 * Nc) * ns + t],
              acc[size_t(2 * Nc + 1) * ns + t], acc[size_t(2 * Nc + 2) * ns + t], coupling,
              currentFactor, deltaOut, currentOut, energyDensity);
}

// Specular boundary condition as a precomputed stencil per entry crossing
// (BoundaryConditionSpecular.h:108-358), one thread per (crossing, energy).
// The previous in-values feed the residual statistics
// (ComputeResidual.h:54-106): per block sums of |d|, |old|, d^2, old^2 and
// maxima of |d|, |old|.
__global__ void boundaryKernel(double *bin, const double *bout, long long nIn,
                               long long nOut, int Ne, const int *offset,
                               const int *src, const double *w,
                               const double *wsum, double *partials) {
  const long long tid = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  double s[6] = {0, 0, 0, 0, 0, 0};
  const long long total = nIn * Ne;
  if (tid < total) {
    const long long i = tid / Ne;
    const int e = int(tid - i * Ne);
    double re = 0.0, im = 0.0;
    for (int o = offset[i]; o < offset[i + 1]; ++o) {
      const size_t si = size_t(src[o]) * Ne + e;
      re += bout[si] * w[o];
      im += bout[size_t(nOut) * Ne + si] * w[o];
    }
    const double ws = wsum[i];
    const double nre = ws > 0.0 ? re / ws : 0.0;
    const double nim = ws > 0.0 ? im / ws : 0.0;
    const double ore = bin[tid], oim = bin[total + tid];
    bin[tid] = nre;
    bin[total + tid] = nim;
    const double d1 = ore - nre, d2 = oim - nim;
    s[0] = fabs(d1) + fabs(d2);
    s[1] = fabs(ore) + fabs(oim);
    s[2] = d1 * d1 + d2 * d2;
    s[3] = ore * ore + oim * oim;
    s[4] = fmax(fabs(d1), fabs(d2));
    s[5] = fmax(fabs(ore), fabs(oim));
  }
  __shared__ double sh[6][256];
  for (int q = 0; q < 6; ++q) sh[q][threadIdx.x] = s[q];
  __syncthreads();
  for (int m = blockDim.x / 2; m > 0; m >>= 1) {
    if (threadIdx.x < m)
      for (int q = 0; q < 6; ++q)
        sh[q][threadIdx.x] = q < 4 ? sh[q][threadIdx.x] + sh[q][threadIdx.x + m]
                                   : fmax(sh[q][threadIdx.x], sh[q][threadIdx.x + m]);
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int q = 0; q < 6; ++q) partials[size_t(blockIdx.x) * 6 + q] = sh[q][0];
}

// Residual statistics of old/new arrays (ComputeResidual.h:54-106), same six
// per-block quantities as boundaryKernel.
__global__ void statsKernel(const double *oldv, const double *newv, long long n,
                            double *partials) {
  double s[6] = {0, 0, 0, 0, 0, 0};
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const double o = oldv[i];
    const double d = o - (newv ? newv[i] : 0.0);
    s[0] += fabs(d);
    s[1] += fabs(o);
    s[2] += d * d;
    s[3] += o * o;
    s[4] = fmax(s[4], fabs(d));
    s[5] = fmax(s[5], fabs(o));
  }
  __shared__ double sh[6][256];
  for (int q = 0; q < 6; ++q) sh[q][threadIdx.x] = s[q];
  __syncthreads();
  for (int m = blockDim.x / 2; m > 0; m >>= 1) {
    if (threadIdx.x < m)
      for (int q = 0; q < 6; ++q)
        sh[q][threadIdx.x] = q < 4 ? sh[q][threadIdx.x] + sh[q][threadIdx.x + m]
                                   : fmax(sh[q][threadIdx.x], sh[q][threadIdx.x + m]);
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int q = 0; q < 6; ++q) partials[size_t(blockIdx.x) * 6 + q] = sh[q][0];
}

// Generic block sums: partial[b] = sum_i f(i) with f selected by `op`:
// 0: a[i]; 1: a[i]*b[i]; 2: a[i]^2 masked by mask[i] (mask may be null).
__global__ void sumKernel(const double *a, const double *b, const uint8_t *mask,
                          long long n, int op, double *partials) {
  double s = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    if (mask && !mask[i]) continue;
    const double x = a[i];
    s += op == 0 ? x : (op == 1 ? x * b[i] : x * x);
  }
  __shared__ double sh[256];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int m = blockDim.x / 2; m > 0; m >>= 1) {
    if (threadIdx.x < m) sh[threadIdx.x] += sh[threadIdx.x + m];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = sh[0];
}

// B = dAy/dx - dAx/dy with one-sided differences at the edges
// (GridFunctions_GPU.h:326-353); writes B^2.
__global__ void curlSquareKernel(const double *A, int N, double invH, double *out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N * N) return;
  const int x = t % N, y = t / N;
  const size_t n = size_t(N) * N;
  const int xn = min(x + 1, N - 1), xp = max(x - 1, 0);
  const int yn = min(y + 1, N - 1), yp = max(y - 1, 0);
  const double dAy = (A[n + size_t(y) * N + xn] - A[n + size_t(y) * N + xp]) / double(xn - xp);
  const double dAx = (A[size_t(yn) * N + x] - A[size_t(yp) * N + x]) / double(yn - yp);
  const double b = (dAy - dAx) * invH;
  out[t] = b * b;
}

// ---------------- elementwise kernels of the accelerators ----------------
// x += (g - x) * s  (Picard.h:133-138)
__global__ void picardKernel(double *x, const double *g, long long n, double s) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) x[i] += (g[i] - x[i]) * s;
}
// d = g - x
__global__ void diffKernel(double *d, const double *g, const double *x, long long n) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) d[i] = g[i] - x[i];
}
// x = sum_i (X_i + D_i s) w_i over the history (Anderson.h:336-355)
// x = sum_q w_q (X_q + s D_q) (Anderson.h:329-352); the history pointers and
// weights travel in the kernel arguments.
constexpr int kMaxHistory = 16;
struct AndersonTerms {
  const double *X[kMaxHistory];
  const double *D[kMaxHistory];
  double w[kMaxHistory];
};
__global__ void andersonCombine(double *x, AndersonTerms t, int m, long long n, double s) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double acc = 0.0;
  for (int q = 0; q < m; ++q) acc += (t.X[q][i] + t.D[q][i] * s) * t.w[q];
  x[i] = acc;
}
// running mean and step per field (CongAcc.h:271-297); the per-field step
// and inertia travel by value (at most 4 order-parameter components and the
// vector potential: 10 fields)
constexpr int kMaxFields = 16;
struct FieldCoefficients {
  double step[kMaxFields];
  double inertia[kMaxFields];
};
__global__ void congaccKernel(double *x, double *mean, const double *d, FieldCoefficients c,
                              long long fieldSize, long long n) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int f = int(i / fieldSize);
  const double m = mean[i] * c.inertia[f] + d[i] * (1.0 - c.inertia[f]);
  mean[i] = m;
  x[i] += c.step[f] * m;
}
// x += d * s (BarzilaiBorwein.h:237-240)
__global__ void axpyKernel(double *x, const double *d, long long n, double s) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) x[i] += d[i] * s;
}
// y = a * base + sum_q c_q v_q (base may alias y; no v_q may alias y)
constexpr int kMaxCombine = 2 * kMaxHistory + 4;
struct CombineTerms {
  const double *v[kMaxCombine];
  double c[kMaxCombine];
};
__global__ void combineTermsKernel(double *y, const double *base, double a, CombineTerms t,
                                   int m, long long n) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double acc = base ? a * base[i] : 0.0;
  for (int q = 0; q < m; ++q) acc += t.c[q] * t.v[q][i];
  y[i] = acc;
}
// mean = (sum of the block partials, in block order) / count
__global__ void finishMeanKernel(const double *partials, long long blocks, double count,
                                 double *mean) {
  double s = 0.0;
  for (long long b = 0; b < blocks; ++b) s += partials[b];
  *mean = s / count;
}

// Dot products of one vector with up to kMaxDots others in one pass over it:
// partials[block * m + i] = block-partial of sum a * b_i.
constexpr int kMaxDots = 16;
struct DotOperands {
  const double *b[kMaxDots];
};
__global__ __launch_bounds__(256) void multiDotKernel(const double *a, DotOperands b,
                                                      int m, long long n,
                                                      double *partials) {
  double s[kMaxDots];
#pragma unroll
  for (int i = 0; i < kMaxDots; ++i) s[i] = 0.0;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x) {
    const double x = a[e];
#pragma unroll
    for (int i = 0; i < kMaxDots; ++i)
      if (i < m) s[i] += x * b.b[i][e];
  }
  __shared__ double sh[256];
  for (int i = 0; i < m; ++i) {
    sh[threadIdx.x] = s[i];
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
      if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) partials[size_t(blockIdx.x) * m + i] = sh[0];
    __syncthreads();
  }
}
// The same over several groups in one launch: block b belongs to the group
// g with blockStart[g] <= b < blockStart[g + 1] and writes its partials at
// offset[g] + (b - blockStart[g]) * m[g].
constexpr int kMaxDotGroups = 8;
struct DotGroupTable {
  const double *a[kMaxDotGroups];
  const double *b[kMaxDotGroups][kMaxDots];
  int m[kMaxDotGroups];
  long long n[kMaxDotGroups];
  long long offset[kMaxDotGroups];
  int blockStart[kMaxDotGroups + 1];
  int numGroups;
};
__global__ __launch_bounds__(256) void multiDotGroupsKernel(DotGroupTable t, double *partials) {
  int g = 0;
  while (g + 1 < t.numGroups && int(blockIdx.x) >= t.blockStart[g + 1]) ++g;
  const int blk = int(blockIdx.x) - t.blockStart[g];
  const int nb = t.blockStart[g + 1] - t.blockStart[g];
  const int m = t.m[g];
  const long long n = t.n[g];
  const double *a = t.a[g];
  double s[kMaxDots];
#pragma unroll
  for (int i = 0; i < kMaxDots; ++i) s[i] = 0.0;
  for (long long e = blk * (long long)blockDim.x + threadIdx.x; e < n;
       e += (long long)nb * blockDim.x) {
    const double x = a[e];
#pragma unroll
    for (int i = 0; i < kMaxDots; ++i)
      if (i < m) s[i] += x * t.b[g][i][e];
  }
  __shared__ double sh[256];
  for (int i = 0; i < m; ++i) {
    sh[threadIdx.x] = s[i];
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
      if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) partials[t.offset[g] + size_t(blk) * m + i] = sh[0];
    __syncthreads();
  }
}
// heavy ball (Polyak.h:216-240)
__global__ void polyakKernel(double *x, double *vel, const double *g,
                             long long n, double s, double drag) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double d = g[i] - x[i];
  const double u = vel[i] * (1.0 - drag) + d * s;
  vel[i] = u;
  x[i] += u;
}
// zero the sites outside the reference grain
__global__ void maskKernel(double *p, const uint8_t *dom, long long n0, int planes) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n0 * planes) return;
  if (!dom[i % n0]) p[i] = 0.0;
}

} // namespace scamd
