// sc_fft.hpp -- FFT solver of the vector-potential Poisson convolution.
//
// ComputeVectorPotential::solve (ComputeVectorPotential.h:647-745) with the
// exact lattice Green function (vector_potential_error = 0) computes
//     A(r) = -h^2 sum_r' G(|x - x'|, |y - y'|) (j(r') - <j>_grain)
// on the N0 x N0 grid, a direct O(N0^4) sum in the reference.  Here the
// correlation is a product in Fourier space on an M x M periodic grid,
// M = the smallest power of two >= 2 N0 - 1, so that no wrapped term reaches
// the N0 x N0 output window.  G is real and even in both directions, so its
// spectrum is real; jx and jy travel together as one complex field
// jx + i jy and come back as Ax + i Ay.
//
// Kernels (all FP64, radix-2 in LDS, one sequence per row or per group of
// adjacent columns):
//   rows forward  : N0 rows of the packed current, zero-padded to M
//   columns       : forward FFT of each column (first N0 rows non-zero),
//                   multiply by the kernel spectrum, inverse FFT, keep N0 rows
//   rows inverse  : N0 rows, scaled and written out as two N0 x N0 planes
// Work buffer: M x M complex doubles.  Kernel spectrum: M x M doubles.
// The buffers are allocated by the solver, or carved out of one block the
// caller provides (scratchBytes(n) bytes, 256-byte aligned), in which case
// the solver allocates no device memory at all.
#ifndef SC_FFT_HPP_
#define SC_FFT_HPP_

#include <hip/hip_runtime.h>

#include <vector>

namespace scamd {

class PoissonFFT {
public:
  // Largest transform length: one column of M complex doubles in 64 KiB LDS.
  static constexpr int kMaxLength = 4096;

  // maxError > 0 replaces the exact kernel by its truncated eigen-expansion
  // (the reference's separable filters, ComputeVectorPotential.h:319-395).
  // With `external` the buffers live there (scratchBytes(n) bytes).
  PoissonFFT(int n, double h, hipStream_t stream, double maxError = 0.0,
             void *external = nullptr);
  // The screened solver: solve() returns (-L_h + screening)^-1 (r - mean)
  // on the n x n window, L_h the five-point lattice Laplacian of spacing h,
  // as a zero-padded (open-boundary) convolution with the operator's lattice
  // Green function.  Its spectrum is the reciprocal of the transform of the
  // operator's stencil, so no real-space Green function is formed.  Used to
  // precondition the vector-potential update (Grain::screenVectorPotential).
  struct Screened {
    double screening; // > 0, in units of 1 / h^2's length unit squared
  };
  PoissonFFT(int n, double h, hipStream_t stream, Screened sc);
  ~PoissonFFT();
  // Transform length for grid side n, and the bytes of the buffers.
  static int lengthFor(int n);
  static size_t scratchBytes(int n);
  PoissonFFT(const PoissonFFT &) = delete;
  PoissonFFT &operator=(const PoissonFFT &) = delete;

  // out (2 planes n x n) = -h^2 G * (j - mean); j is 2 planes n x n and
  // mean points to the two means (x, y).  Device pointers; runs on the
  // stream given at construction.
  void solve(const double *dj, const double *dMean, double *dout);

  int length() const { return m_; }
  int n() const { return n_; }
  // rank-one filters of the truncated kernel (0: the exact kernel)
  int numFilters() const { return numFilters_; }

private:
  void allocate(void *external);
  // khat_ <- the forward transform of the even kernel g (n x n quadrant,
  // g[dy * n + dx]) on the periodic M x M grid
  void transformKernel(const std::vector<double> &g);
  void release();

  int n_, m_, logm_;
  double scale_ = 0.0; // of the inverse transform in solve()
  double h_;
  hipStream_t stream_;
  bool owns_ = true;
  int numFilters_ = 0;
  double *khat_ = nullptr;  // M x M real spectrum of the kernel
  double2 *work_ = nullptr; // M x M complex
  double2 *tw_ = nullptr;   // exp(-2 pi i k / M), k < M/2
};

} // namespace scamd

#endif
