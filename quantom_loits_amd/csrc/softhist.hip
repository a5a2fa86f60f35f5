// softhist.hip -- differentiable soft-binning histograms of packed events.
//
// The discriminator input of the GAN-style training harness: for instance i
// with packed rows [row_start[i], row_start[i + 1]) and t = (v - lo) / bw
// (bw = (hi - lo) / bins),
//   hist[i][c] = sum_r exp(-(t_r - c - 1/2)^2 / (2 w^2))
// over the seven bins c nearest t_r (c in [0, bins); w = kernel width in bins,
// the weight beyond three bins is below 1e-7 for w = 0.5), and the backward
//   dv_r = sum_c ghist[i][c] exp(-d^2 / (2 w^2)) (-d / w^2) / bw,  d = t_r - c - 1/2.
// Float atomics make the sums order-dependent at float rounding.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "capi_internal.hpp"
#include "loits_hip.h"

namespace qloits {
namespace {

constexpr int kHistThreads = 256;
constexpr int kHistRows = 8192;     // rows per work-group
constexpr int kHistInst = 8;        // instances a work-group keeps in LDS
constexpr int kHistMaxBins = 128;
constexpr int kTaps = 3;            // bins on each side of the nearest one

struct HistParams {
  const float* values;
  const int64_t* row_start;  // [n_instances + 1]
  int64_t n_instances, n_rows;
  float lo, inv_bw, inv_2w2, inv_w2;
  int bins;
  float* hist;               // [n_instances, bins]
  const float* grad_hist;    // backward
  float* grad_values;        // backward
};

// instance of row r: last i with row_start[i] <= r (rows of empty instances skipped)
__device__ int64_t instance_of(const int64_t* row_start, int64_t n, int64_t r) {
  int64_t a = 0, b = n - 1;
  while (a < b) {
    const int64_t mid = (a + b + 1) >> 1;
    if (row_start[mid] <= r) a = mid; else b = mid - 1;
  }
  return a;
}

// Bin position t of a value and the first bin c0 = floor(t) of its taps.
// NaN values (and so NaN events) take no part: false.  t is clamped to
// [-(kTaps + 1), bins + kTaps + 1] first, so the float-to-int conversion is
// defined for every finite or infinite t; a value that far outside has no tap
// on a bin either way.
__device__ __forceinline__ bool tap_origin(float v, const HistParams& P, float* t, int* c0) {
  float u = (v - P.lo) * P.inv_bw;
  if (u != u) return false;
  const float lim_lo = -float(kTaps + 1), lim_hi = float(P.bins + kTaps + 1);
  u = u < lim_lo ? lim_lo : (u > lim_hi ? lim_hi : u);
  *t = u;
  *c0 = int(__builtin_floorf(u));
  return true;
}

__global__ void zero_hist(float* h, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) h[i] = 0.0f;
}

__global__ void __launch_bounds__(kHistThreads) hist_kernel(HistParams P) {
  __shared__ float sh[kHistInst * kHistMaxBins];
  __shared__ int64_t sh_inst0, sh_inst1;
  const int64_t r0 = int64_t(blockIdx.x) * kHistRows;
  const int64_t r1 = r0 + kHistRows < P.n_rows ? r0 + kHistRows : P.n_rows;
  if (threadIdx.x == 0) {
    sh_inst0 = instance_of(P.row_start, P.n_instances, r0);
    sh_inst1 = instance_of(P.row_start, P.n_instances, r1 - 1);
  }
  for (int i = threadIdx.x; i < kHistInst * kHistMaxBins; i += blockDim.x) sh[i] = 0.0f;
  __syncthreads();
  const int64_t inst0 = sh_inst0;
  const bool local = sh_inst1 - inst0 < kHistInst;
  int64_t inst = inst0;
  for (int64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
    while (inst + 1 < P.n_instances && P.row_start[inst + 1] <= r) ++inst;
    float t;
    int c0;
    if (!tap_origin(P.values[r], P, &t, &c0)) continue;
#pragma unroll
    for (int k = -kTaps; k <= kTaps; ++k) {
      const int c = c0 + k;
      if (c < 0 || c >= P.bins) continue;
      const float d = t - (float(c) + 0.5f);
      const float w = __expf(-(d * d) * P.inv_2w2);
      if (local)
        unsafeAtomicAdd(&sh[int(inst - inst0) * kHistMaxBins + c], w);
      else
        unsafeAtomicAdd(&P.hist[inst * P.bins + c], w);
    }
  }
  __syncthreads();
  if (!local) return;
  const int64_t n_local = sh_inst1 - inst0 + 1;
  for (int i = threadIdx.x; i < n_local * P.bins; i += blockDim.x) {
    const int li = i / P.bins, c = i - li * P.bins;
    const float v = sh[li * kHistMaxBins + c];
    if (v != 0.0f) unsafeAtomicAdd(&P.hist[(inst0 + li) * P.bins + c], v);
  }
}

__global__ void __launch_bounds__(kHistThreads) hist_backward_kernel(HistParams P) {
  __shared__ int64_t sh_inst0;
  const int64_t r0 = int64_t(blockIdx.x) * kHistRows;
  const int64_t r1 = r0 + kHistRows < P.n_rows ? r0 + kHistRows : P.n_rows;
  if (threadIdx.x == 0) sh_inst0 = instance_of(P.row_start, P.n_instances, r0);
  __syncthreads();
  int64_t inst = sh_inst0;
  for (int64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
    while (inst + 1 < P.n_instances && P.row_start[inst + 1] <= r) ++inst;
    float t;
    int c0;
    if (!tap_origin(P.values[r], P, &t, &c0)) {
      P.grad_values[r] = 0.0f;
      continue;
    }
    const float* gh = P.grad_hist + inst * P.bins;
    float acc = 0.0f;
#pragma unroll
    for (int k = -kTaps; k <= kTaps; ++k) {
      const int c = c0 + k;
      if (c < 0 || c >= P.bins) continue;
      const float d = t - (float(c) + 0.5f);
      const float w = __expf(-(d * d) * P.inv_2w2);
      acc += gh[c] * w * (-d * P.inv_w2);
    }
    P.grad_values[r] = acc * P.inv_bw;
  }
}

int check_hist(const float* values, const int64_t* row_start, int64_t n_instances, int64_t n_rows,
               float lo, float hi, int bins, float width) {
  if (n_instances <= 0 || n_rows < 0) return fail(LOITS_HIP_EINVAL, "soft_histogram: bad sizes");
  if (!row_start || (n_rows > 0 && !values))
    return fail(LOITS_HIP_EINVAL, "soft_histogram: NULL input");
  if (bins < 1 || bins > kHistMaxBins)
    return fail(LOITS_HIP_EINVAL, "soft_histogram: bins must be in [1, 128]");
  if (!(hi > lo) || !(width > 0.0f))
    return fail(LOITS_HIP_EINVAL, "soft_histogram: need hi > lo and width > 0");
  if ((n_rows + kHistRows - 1) / kHistRows > INT32_MAX)
    return fail(LOITS_HIP_EUNSUPPORTED, "soft_histogram: too many rows");
  return LOITS_HIP_OK;
}

HistParams params(const float* values, const int64_t* row_start, int64_t n_instances,
                  int64_t n_rows, float lo, float hi, int bins, float width) {
  HistParams P{};
  P.values = values;
  P.row_start = row_start;
  P.n_instances = n_instances;
  P.n_rows = n_rows;
  P.lo = lo;
  P.inv_bw = float(bins) / (hi - lo);
  P.inv_2w2 = 1.0f / (2.0f * width * width);
  P.inv_w2 = 1.0f / (width * width);
  P.bins = bins;
  return P;
}

}  // namespace
}  // namespace qloits

using namespace qloits;

extern "C" {

int loits_hip_soft_histogram(const float* values, const int64_t* row_start, int64_t n_instances,
                             int64_t n_rows, float lo, float hi, int bins, float width,
                             float* hist, void* stream) {
  int rc = check_hist(values, row_start, n_instances, n_rows, lo, hi, bins, width);
  if (rc != LOITS_HIP_OK) return rc;
  if (!hist) return fail(LOITS_HIP_EINVAL, "soft_histogram: hist is NULL");
  HistParams P = params(values, row_start, n_instances, n_rows, lo, hi, bins, width);
  P.hist = hist;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int64_t nh = n_instances * bins;
  hipLaunchKernelGGL(zero_hist, dim3(unsigned((nh + 255) / 256)), dim3(256), 0, st, hist, nh);
  rc = check_hip(hipGetLastError(), "soft_histogram zero launch");
  if (rc != LOITS_HIP_OK || n_rows == 0) return rc;
  hipLaunchKernelGGL(hist_kernel, dim3(unsigned((n_rows + kHistRows - 1) / kHistRows)),
                     dim3(kHistThreads), 0, st, P);
  return check_hip(hipGetLastError(), "soft_histogram launch");
}

int loits_hip_soft_histogram_backward(const float* values, const int64_t* row_start,
                                      int64_t n_instances, int64_t n_rows, float lo, float hi,
                                      int bins, float width, const float* grad_hist,
                                      float* grad_values, void* stream) {
  int rc = check_hist(values, row_start, n_instances, n_rows, lo, hi, bins, width);
  if (rc != LOITS_HIP_OK) return rc;
  if (n_rows == 0) return LOITS_HIP_OK;
  if (!grad_hist || !grad_values)
    return fail(LOITS_HIP_EINVAL, "soft_histogram_backward: NULL gradient");
  HistParams P = params(values, row_start, n_instances, n_rows, lo, hi, bins, width);
  P.grad_hist = grad_hist;
  P.grad_values = grad_values;
  hipLaunchKernelGGL(hist_backward_kernel, dim3(unsigned((n_rows + kHistRows - 1) / kHistRows)),
                     dim3(kHistThreads), 0, static_cast<hipStream_t>(stream), P);
  return check_hip(hipGetLastError(), "soft_histogram_backward launch");
}

}  // extern "C"
