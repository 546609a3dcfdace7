// sc_fft.hip -- FFT solver of the vector-potential Poisson convolution
// (declared in sc_fft.hpp).
//
// Every transform is radix-2 in LDS.  Forward transforms are decimation in
// frequency (natural order in, bit-reversed order out) and inverse transforms
// decimation in time (bit-reversed in, natural out), so no permutation pass
// is needed: the spectrum of the kernel is built by the same forward kernels
// and therefore lands in the same bit-reversed layout as the spectrum of the
// current it multiplies.
#include "sc_fft.hpp"

#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "sc_setup.hpp"

#define SC_FFT_HIP(call)                                                            \
  do {                                                                              \
    const hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess)                                                           \
      throw std::runtime_error(std::string("HIP error in Poisson FFT: ") +           \
                               hipGetErrorString(e_));                              \
  } while (0)

namespace scamd {
namespace {

constexpr int kThreads = 256;
constexpr int kLdsBytes = 65536;

__device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 cmulConj(double2 a, double2 w) { // a * conj(w)
  return make_double2(a.x * w.x + a.y * w.y, a.y * w.x - a.x * w.y);
}

// Forward DIF over `count` sequences of length M stored back to back in LDS.
__device__ void difForward(double2 *x, int M, int logM, int count, const double2 *tw) {
  const int perSeq = M >> 1;
  const int total = perSeq * count;
  for (int s = logM - 1; s >= 0; --s) {
    const int half = 1 << s;
    const int twStride = M >> (s + 1); // exp(-2 pi i pos / (2 half))
    for (int b = threadIdx.x; b < total; b += blockDim.x) {
      const int seq = b / perSeq, r = b - seq * perSeq;
      const int pos = r & (half - 1);
      const int i0 = seq * M + ((r >> s) << (s + 1)) + pos;
      const int i1 = i0 + half;
      const double2 a = x[i0], c = x[i1];
      x[i0] = make_double2(a.x + c.x, a.y + c.y);
      x[i1] = cmul(make_double2(a.x - c.x, a.y - c.y), tw[pos * twStride]);
    }
    __syncthreads();
  }
}

// Inverse DIT (unnormalised) over `count` sequences in bit-reversed order.
__device__ void ditInverse(double2 *x, int M, int logM, int count, const double2 *tw) {
  const int perSeq = M >> 1;
  const int total = perSeq * count;
  for (int s = 0; s < logM; ++s) {
    const int half = 1 << s;
    const int twStride = M >> (s + 1);
    for (int b = threadIdx.x; b < total; b += blockDim.x) {
      const int seq = b / perSeq, r = b - seq * perSeq;
      const int pos = r & (half - 1);
      const int i0 = seq * M + ((r >> s) << (s + 1)) + pos;
      const int i1 = i0 + half;
      const double2 a = x[i0];
      const double2 t = cmulConj(x[i1], tw[pos * twStride]);
      x[i0] = make_double2(a.x + t.x, a.y + t.y);
      x[i1] = make_double2(a.x - t.x, a.y - t.y);
    }
    __syncthreads();
  }
}

// Rows forward: row y < nRows of (re - sub[0]) + i (im - sub[1]), columns
// x < nCols (zero beyond), transformed and stored to work[y][.].
__global__ __launch_bounds__(kThreads) void rowsForwardKernel(
    const double *re, const double *im, int ld, int nRows, int nCols, const double *sub,
    int M, int logM, const double2 *tw, double2 *work) {
  extern __shared__ double2 lds[];
  const int y = blockIdx.x;
  if (y >= nRows) return;
  const double *rr = re + size_t(y) * ld;
  const double *ri = im ? im + size_t(y) * ld : nullptr;
  const double subRe = sub ? sub[0] : 0.0, subIm = sub ? sub[1] : 0.0;
  for (int x = threadIdx.x; x < M; x += blockDim.x) {
    double2 v = make_double2(0.0, 0.0);
    if (x < nCols) {
      v.x = rr[x] - subRe;
      v.y = ri ? ri[x] - subIm : 0.0;
    }
    lds[x] = v;
  }
  __syncthreads();
  difForward(lds, M, logM, 1, tw);
  double2 *out = work + size_t(y) * M;
  for (int x = threadIdx.x; x < M; x += blockDim.x) out[x] = lds[x];
}

// Columns: `cb` adjacent columns per block.  Rows >= nRowsIn are zero.  With
// khatOut set, stores the real part of the forward transform (the kernel
// spectrum); otherwise multiplies by khat, transforms back and stores rows
// < nRowsOut.
__global__ __launch_bounds__(kThreads) void columnsKernel(
    double2 *work, int nRowsIn, int nRowsOut, int M, int logM, int cb,
    const double2 *tw, const double *khat, double *khatOut) {
  extern __shared__ double2 lds[];
  const int c0 = blockIdx.x * cb;
  const int total = cb * M;
  for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
    const int y = idx / cb, c = idx - y * cb;
    lds[c * M + y] = y < nRowsIn ? work[size_t(y) * M + c0 + c] : make_double2(0.0, 0.0);
  }
  __syncthreads();
  difForward(lds, M, logM, cb, tw);
  if (khatOut) {
    for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
      const int y = idx / cb, c = idx - y * cb;
      khatOut[size_t(y) * M + c0 + c] = lds[c * M + y].x;
    }
    return;
  }
  for (int idx = threadIdx.x; idx < total; idx += blockDim.x) {
    const int y = idx / cb, c = idx - y * cb;
    const double k = khat[size_t(y) * M + c0 + c];
    double2 &v = lds[c * M + y];
    v = make_double2(v.x * k, v.y * k);
  }
  __syncthreads();
  ditInverse(lds, M, logM, cb, tw);
  for (int idx = threadIdx.x; idx < cb * nRowsOut; idx += blockDim.x) {
    const int y = idx / cb, c = idx - y * cb;
    work[size_t(y) * M + c0 + c] = lds[c * M + y];
  }
}

// Rows inverse: row y < n back to natural order, scaled, split into planes.
__global__ __launch_bounds__(kThreads) void rowsInverseKernel(
    const double2 *work, int n, int M, int logM, const double2 *tw, double scale,
    double *out) {
  extern __shared__ double2 lds[];
  const int y = blockIdx.x;
  if (y >= n) return;
  const double2 *row = work + size_t(y) * M;
  for (int x = threadIdx.x; x < M; x += blockDim.x) lds[x] = row[x];
  __syncthreads();
  ditInverse(lds, M, logM, 1, tw);
  const size_t plane = size_t(n) * n;
  for (int x = threadIdx.x; x < n; x += blockDim.x) {
    const double2 v = lds[x];
    out[size_t(y) * n + x] = v.x * scale;
    out[plane + size_t(y) * n + x] = v.y * scale;
  }
}

int columnsPerBlock(int M) {
  int cb = kLdsBytes / (M * int(sizeof(double2)));
  if (cb > 8) cb = 8;
  if (cb > M) cb = M;
  return cb < 1 ? 1 : cb;
}

// The kernel on the periodic M x M grid: G(|d|) for |d| < n in each
// direction, zero between (g is n x n, g[dy * n + dx]).
__global__ void kernelGridKernel(const double *g, int n, int M, double *grid) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= (long long)M * M) return;
  const int my = int(i / M), mx = int(i - (long long)my * M);
  const int dy = my < n ? my : (my > M - n ? M - my : -1);
  const int dx = mx < n ? mx : (mx > M - n ? M - mx : -1);
  grid[i] = (dy >= 0 && dx >= 0) ? g[size_t(dy) * n + dx] : 0.0;
}

// The Green function's spectrum from the operator's: 1 / lambda, elementwise
// (the screened operator's spectrum is positive everywhere).
__global__ void reciprocalKernel(double *s, long long n) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) s[i] = 1.0 / s[i];
}

size_t alignUp(size_t b) { return (b + 255) & ~size_t(255); }

} // namespace

int PoissonFFT::lengthFor(int n) {
  int m = 1;
  while (m < 2 * n - 1) m <<= 1;
  return m < 2 ? 2 : m;
}

size_t PoissonFFT::scratchBytes(int n) {
  const size_t M = size_t(lengthFor(n));
  return alignUp(sizeof(double) * M * M) + alignUp(sizeof(double2) * M * M) +
         alignUp(sizeof(double2) * (M / 2));
}

PoissonFFT::PoissonFFT(int n, double h, hipStream_t stream, double maxError, void *external)
    : n_(n), m_(1), logm_(0), h_(h), stream_(stream), owns_(external == nullptr) {
  if (n < 1) throw std::invalid_argument("Poisson grid must be non-empty");
  std::vector<double> g;
  numFilters_ = poissonKernel(n, h, g, maxError);
  // out = -h^2 G * (j - <j>): the inverse transform's 1 / M^2 folded in
  allocate(external);
  scale_ = -(h_ * h_) / (static_cast<double>(m_) * static_cast<double>(m_));
  try {
    transformKernel(g);
    SC_FFT_HIP(hipStreamSynchronize(stream_));
  } catch (...) {
    release();
    throw;
  }
}

PoissonFFT::PoissonFFT(int n, double h, hipStream_t stream, Screened sc)
    : n_(n), m_(1), logm_(0), h_(h), stream_(stream), owns_(true) {
  if (n < 2) throw std::invalid_argument("screened solver needs a grid of side >= 2");
  if (!(sc.screening > 0.0)) throw std::invalid_argument("screening must be positive");
  // the stencil of -L_h + screening: its transform is the operator's
  // spectrum 4 / h^2 (sin^2(pi p / M) + sin^2(pi q / M)) + screening, whose
  // reciprocal is the spectrum of the Green function
  std::vector<double> g(size_t(n) * n, 0.0);
  g[0] = 4.0 / (h * h) + sc.screening;
  g[1] = g[size_t(n)] = -1.0 / (h * h);
  allocate(nullptr);
  scale_ = 1.0 / (static_cast<double>(m_) * static_cast<double>(m_));
  try {
    transformKernel(g);
    const long long cells = static_cast<long long>(m_) * m_;
    reciprocalKernel<<<static_cast<unsigned>((cells + kThreads - 1) / kThreads), kThreads, 0,
                       stream_>>>(khat_, cells);
    SC_FFT_HIP(hipGetLastError());
    SC_FFT_HIP(hipStreamSynchronize(stream_));
  } catch (...) {
    release();
    throw;
  }
}

void PoissonFFT::allocate(void *external) {
  m_ = lengthFor(n_);
  logm_ = 0;
  while ((1 << logm_) < m_) ++logm_;
  if (m_ > kMaxLength)
    throw std::invalid_argument("Poisson FFT: grid side " + std::to_string(n_) +
                                " needs a transform longer than " +
                                std::to_string(kMaxLength));
  const int M = m_;
  // Twiddles exp(-2 pi i k / M), k < M/2, from long-double trigonometry.
  std::vector<double2> tw(M / 2);
  for (int k = 0; k < M / 2; ++k) {
    const long double a = 2.0L * 3.141592653589793238462643383279502884L *
                          static_cast<long double>(k) / static_cast<long double>(M);
    tw[k] = make_double2(static_cast<double>(cosl(a)), -static_cast<double>(sinl(a)));
  }
  try {
    if (owns_) {
      SC_FFT_HIP(hipMalloc(&tw_, sizeof(double2) * tw.size()));
      SC_FFT_HIP(hipMalloc(&khat_, sizeof(double) * size_t(M) * M));
      SC_FFT_HIP(hipMalloc(&work_, sizeof(double2) * size_t(M) * M));
    } else {
      char *base = static_cast<char *>(external);
      khat_ = reinterpret_cast<double *>(base);
      base += alignUp(sizeof(double) * size_t(M) * M);
      work_ = reinterpret_cast<double2 *>(base);
      base += alignUp(sizeof(double2) * size_t(M) * M);
      tw_ = reinterpret_cast<double2 *>(base);
    }
    SC_FFT_HIP(hipMemcpyAsync(tw_, tw.data(), sizeof(double2) * tw.size(),
                              hipMemcpyHostToDevice, stream_));
    SC_FFT_HIP(hipStreamSynchronize(stream_));
  } catch (...) {
    release();
    throw;
  }
}

void PoissonFFT::transformKernel(const std::vector<double> &g) {
  // The n x n kernel goes into the work buffer, its periodic M x M grid
  // into the spectrum buffer, whose forward transform (rows into the work
  // buffer, then columns) overwrites it with the spectrum.
  const int M = m_, n = n_;
  SC_FFT_HIP(hipMemcpyAsync(work_, g.data(), sizeof(double) * g.size(), hipMemcpyHostToDevice,
                            stream_));
  const long long cells = static_cast<long long>(M) * M;
  kernelGridKernel<<<static_cast<unsigned>((cells + kThreads - 1) / kThreads), kThreads, 0,
                     stream_>>>(reinterpret_cast<const double *>(work_), n, M, khat_);
  SC_FFT_HIP(hipGetLastError());
  const int cb = columnsPerBlock(M);
  rowsForwardKernel<<<M, kThreads, sizeof(double2) * M, stream_>>>(
      khat_, nullptr, M, M, M, nullptr, M, logm_, tw_, work_);
  SC_FFT_HIP(hipGetLastError());
  columnsKernel<<<M / cb, kThreads, sizeof(double2) * M * cb, stream_>>>(
      work_, M, 0, M, logm_, cb, tw_, nullptr, khat_);
  SC_FFT_HIP(hipGetLastError());
  // the host kernel must outlive the copy
  SC_FFT_HIP(hipStreamSynchronize(stream_));
}

void PoissonFFT::release() {
  if (!owns_) return;
  if (tw_) (void)hipFree(tw_);
  if (khat_) (void)hipFree(khat_);
  if (work_) (void)hipFree(work_);
  tw_ = nullptr;
  khat_ = nullptr;
  work_ = nullptr;
}

PoissonFFT::~PoissonFFT() { release(); }

void PoissonFFT::solve(const double *dj, const double *dMean, double *dout) {
  const int M = m_, n = n_;
  const int cb = columnsPerBlock(M);
  const size_t plane = size_t(n) * n;
  rowsForwardKernel<<<n, kThreads, sizeof(double2) * M, stream_>>>(
      dj, dj + plane, n, n, n, dMean, M, logm_, tw_, work_);
  SC_FFT_HIP(hipGetLastError());
  columnsKernel<<<M / cb, kThreads, sizeof(double2) * M * cb, stream_>>>(
      work_, n, n, M, logm_, cb, tw_, khat_, nullptr);
  SC_FFT_HIP(hipGetLastError());
  rowsInverseKernel<<<n, kThreads, sizeof(double2) * M, stream_>>>(work_, n, M, logm_, tw_,
                                                                  scale_, dout);
  SC_FFT_HIP(hipGetLastError());
}

} // namespace scamd
