// sc_noalloc.hip -- the vector-potential solve without device allocation.
//
// sc_poisson_solve_noalloc is sc_poisson_solve for callers that cannot call
// the device allocator (code under graph capture, tight serving loops,
// harnesses that forbid it).  The solver's transform buffers (kernel
// spectrum, work array, twiddles) and the staging of the current and of the
// result live in one static device array, `scratch`, of a fixed size that
// covers grids up to SC_NOALLOC_MAX_GRID sites a side.  Each GPU holds its
// own copy of the array from the load of the library until the process
// ends; its device address is looked up once per device, on the first call
// there, and only that address is handed to kernels and copies.  Every call
// on a GPU shares that GPU's array, so calls that overlap on one GPU (from
// several host threads) race; calls in sequence are independent, since
// every call rewrites all of the array it reads.
#include "../../include/superconga_amd.h"
#include "sc_fft.hpp"
#include "sc_util.hpp"

#include <hip/hip_runtime.h>

#include <cstddef>
#include <mutex>
#include <stdexcept>
#include <string>

namespace scamd {
namespace {

constexpr size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

constexpr int kMaxGrid = SC_NOALLOC_MAX_GRID;
constexpr size_t kMaxLength = 1024; // transform length for kMaxGrid: >= 2 n - 1
// PoissonFFT's buffers at that length: spectrum (M^2 doubles), work (M^2
// complex), twiddles (M / 2 complex), each 256-byte aligned
constexpr size_t kFftBytes = align256(sizeof(double) * kMaxLength * kMaxLength) +
                             align256(2 * sizeof(double) * kMaxLength * kMaxLength) +
                             align256(2 * sizeof(double) * (kMaxLength / 2));
constexpr size_t kPlane = size_t(kMaxGrid) * kMaxGrid;
// the current (two planes and the two means) and the result (two planes)
constexpr size_t kStageBytes = align256(sizeof(double) * (2 * kPlane + 2)) +
                               align256(sizeof(double) * 2 * kPlane);
constexpr size_t kScratchBytes = kFftBytes + kStageBytes;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
void *g_scratch[kMaxDevices] = {};

void hipOk(hipError_t e, const char *what) {
  if (e != hipSuccess)
    throw std::runtime_error(std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
}

// The device address of the current device's copy of `scratch`.
void *deviceScratch() {
  int dev = 0;
  hipOk(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices)
    throw std::runtime_error("device index " + std::to_string(dev) + " out of range");
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    hipOk(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress");
    if (!p) throw std::runtime_error("hipGetSymbolAddress returned a null address");
    g_scratch[dev] = p;
  }
  return g_scratch[dev];
}

} // namespace
} // namespace scamd

extern "C" int sc_poisson_solve_noalloc(int n, double h, int numGrain, int enforce,
                                        const double *current, double *out) {
  using scamd::reportError;
  if (!current || !out) return reportError(SC_ERR_ARGUMENT, "null pointer argument");
  if (n < 2 || !(h > 0.0) || numGrain < 1)
    return reportError(SC_ERR_ARGUMENT, "need n >= 2, h > 0 and num_grain >= 1");
  if (n > scamd::kMaxGrid)
    return reportError(SC_ERR_ARGUMENT,
                       "grid side " + std::to_string(n) + " exceeds the " +
                           std::to_string(scamd::kMaxGrid) +
                           " sites of sc_poisson_solve_noalloc's static scratch; use "
                           "sc_poisson_solve");
  try {
    if (scamd::PoissonFFT::scratchBytes(n) > scamd::kFftBytes)
      throw std::runtime_error("transform buffers exceed the static scratch");
    const size_t n0 = size_t(n) * n;
    double means[2] = {0.0, 0.0};
    if (enforce) {
      for (size_t i = 0; i < n0; ++i) {
        means[0] += current[i];
        means[1] += current[n0 + i];
      }
      means[0] /= numGrain;
      means[1] /= numGrain;
    }
    unsigned char *base = static_cast<unsigned char *>(scamd::deviceScratch());
    double *dj = reinterpret_cast<double *>(base + scamd::kFftBytes);
    double *dout = reinterpret_cast<double *>(
        base + scamd::kFftBytes + scamd::align256(sizeof(double) * (2 * scamd::kPlane + 2)));
    // everything runs on the null stream; the copies from and to pageable
    // host memory return when they are complete
    scamd::hipOk(hipMemcpy(dj, current, 2 * n0 * sizeof(double), hipMemcpyHostToDevice),
                 "hipMemcpy");
    scamd::hipOk(hipMemcpy(dj + 2 * n0, means, 2 * sizeof(double), hipMemcpyHostToDevice),
                 "hipMemcpy");
    scamd::PoissonFFT fft(n, h, nullptr, 0.0, base);
    fft.solve(dj, dj + 2 * n0, dout);
    scamd::hipOk(hipMemcpy(out, dout, 2 * n0 * sizeof(double), hipMemcpyDeviceToHost),
                 "hipMemcpy");
    return SC_OK;
  } catch (const std::exception &e) {
    return reportError(SC_ERR_DEVICE, e.what());
  }
}
