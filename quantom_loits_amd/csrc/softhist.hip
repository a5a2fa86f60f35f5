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

#include <cmath>
#include <cstdint>

#include "capi_internal.hpp"
#include "loits_hip.h"

namespace qloits {
namespace {

constexpr int kHistThreads = 256;  // backward work-group
constexpr int kHistRows = 8192;     // rows per work-group
constexpr int kHistUnroll = 4;      // rows a thread loads before it uses them
constexpr int kHistMaxBins = 128;
constexpr int kTaps = 3;            // bins on each side of the nearest one
// Taps reach bins -(2 kTaps + 1) .. bins + 2 kTaps + 1 after the clamp of
// tap_origin: LDS rows are shifted by kPad so that every tap has a row.
constexpr int kPad = 2 * kTaps + 1;
// Widths with 1 / (2 w^2) up to this value (w >= 0.177 bins) take the factored
// taps below; their largest factor, exp(2 kTaps a |d0|), stays below 1e21.
constexpr float kFactoredMaxA = 16.0f;

struct HistParams {
  const float* values;
  const int64_t* row_start;  // [n_instances + 1]
  int64_t n_instances, n_rows;
  float lo, inv_bw, inv_2w2, inv_w2;
  float g[kTaps + 1];        // exp(-k^2 / (2 w^2)), k = 0..kTaps
  int bins;
  float* hist;               // [n_instances, bins]
  const float* grad_hist;    // backward
  float* grad_values;        // backward
};

// instance of row r: last i with row_start[i] <= r (0 for rows before row_start[0])
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

// Weights w[k + kTaps] = exp(-(d0 - k)^2 a) of bins c0 + k, d0 = t - c0 - 1/2,
// a = 1 / (2 w^2), and the distances d0 - k.  FACTORED: exp(-a d0^2)
// exp(2 a d0)^k exp(-a k^2), three exponentials instead of seven (d0 lies in
// [-1/2, 1/2), so the powers stay finite for a <= kFactoredMaxA).
template <bool FACTORED>
__device__ __forceinline__ void tap_weights(const HistParams& P, float t, int c0,
                                            float (&w)[2 * kTaps + 1], float (&d)[2 * kTaps + 1]) {
  const float d0 = t - (float(c0) + 0.5f);
#pragma unroll
  for (int k = -kTaps; k <= kTaps; ++k) d[k + kTaps] = d0 - float(k);
  if (FACTORED) {
    const float a = P.inv_2w2;
    const float e = __expf(-(d0 * d0) * a);
    const float up = __expf((2.0f * a) * d0), down = __expf((-2.0f * a) * d0);
    w[kTaps] = e;
    float pu = e, pd = e;
#pragma unroll
    for (int k = 1; k <= kTaps; ++k) {
      pu *= up;
      pd *= down;
      w[kTaps + k] = pu * P.g[k];
      w[kTaps - k] = pd * P.g[k];
    }
  } else {
#pragma unroll
    for (int j = 0; j < 2 * kTaps + 1; ++j) w[j] = __expf(-(d[j] * d[j]) * P.inv_2w2);
  }
}

// The work-group's rows [r0, r1) and the instances i0..i1 they fall in; the
// rows of instance i are [max(r0, row_start[i]), min(r1, row_start[i + 1])).
struct RowRange {
  int64_t r0, r1, i0, i1;
};

__device__ __forceinline__ RowRange row_range(const HistParams& P) {
  RowRange R;
  R.r0 = int64_t(blockIdx.x) * kHistRows;
  R.r1 = R.r0 + kHistRows < P.n_rows ? R.r0 + kHistRows : P.n_rows;
  R.i0 = instance_of(P.row_start, P.n_instances, R.r0);
  R.i1 = instance_of(P.row_start, P.n_instances, R.r1 - 1);
  return R;
}

// kHistUnroll rows r, r + STRIDE, ... below b, NaN past b (NaN takes no part)
template <int STRIDE>
__device__ __forceinline__ void load_rows(const float* values, int64_t r, int64_t b,
                                          float (&v)[kHistUnroll]) {
#pragma unroll
  for (int u = 0; u < kHistUnroll; ++u) {
    const int64_t q = r + int64_t(u) * STRIDE;
    v[u] = q < b ? values[q] : __builtin_nanf("");
  }
}

__global__ void zero_hist(float* h, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) h[i] = 0.0f;
}

// Forward, one wave per work-group.  Each lane owns column `lane` of an LDS
// histogram sh[(c + kPad) * 64 + lane] and adds its rows' taps with plain
// loads and stores: no atomics, and the 64 lanes of an access sit in 64
// distinct banks.  After each instance's rows lane j sums bin j's 64 columns
// (bin j + 64 next, for more than 64 bins), clears them and adds the sum to
// hist with one device atomic.  Rows outside bins 0..bins-1 are never summed.
template <bool FACTORED>
__global__ void __launch_bounds__(64) hist_kernel(HistParams P) {
  extern __shared__ float sh[];
  const int lane = threadIdx.x;
  const RowRange R = row_range(P);
  for (int c = 0; c < P.bins + 2 * kPad + 1; ++c) sh[c * 64 + lane] = 0.0f;
  for (int64_t i = R.i0; i <= R.i1; ++i) {
    const int64_t a = R.r0 > P.row_start[i] ? R.r0 : P.row_start[i];
    const int64_t b = R.r1 < P.row_start[i + 1] ? R.r1 : P.row_start[i + 1];
    if (a >= b) continue;
    for (int64_t r = a + lane; r < b; r += kHistUnroll * 64) {
      float v[kHistUnroll];
      load_rows<64>(P.values, r, b, v);
#pragma unroll
      for (int u = 0; u < kHistUnroll; ++u) {
        float t;
        int c0;
        if (!tap_origin(v[u], P, &t, &c0)) continue;
        float w[2 * kTaps + 1], d[2 * kTaps + 1];
        tap_weights<FACTORED>(P, t, c0, w, d);
        float* col = sh + (c0 - kTaps + kPad) * 64 + lane;
#pragma unroll
        for (int j = 0; j < 2 * kTaps + 1; ++j) col[j * 64] += w[j];
      }
    }
    __syncthreads();
    for (int bin = lane; bin < P.bins; bin += 64) {
      float* row = sh + (bin + kPad) * 64;
      float s = 0.0f;
      for (int j = 0; j < 64; ++j) {  // rotated by bin: the lanes read distinct banks
        const int col = (j + bin) & 63;
        s += row[col];
        row[col] = 0.0f;
      }
      if (s != 0.0f) unsafeAtomicAdd(&P.hist[i * P.bins + bin], s);
    }
    __syncthreads();
  }
}

// Backward: the instance's gradient row staged in LDS (zero on the pad rows),
// one row per thread at a time.  Rows before row_start[0] or from
// row_start[n_instances] on get 0.
template <bool FACTORED>
__global__ void __launch_bounds__(kHistThreads) hist_backward_kernel(HistParams P) {
  __shared__ float gh[kHistMaxBins + 2 * kPad + 1];
  const RowRange R = row_range(P);
  const int64_t first = P.row_start[0], last = P.row_start[P.n_instances];
  for (int64_t r = R.r0 + threadIdx.x; r < R.r1; r += kHistThreads)
    if (r < first || r >= last) P.grad_values[r] = 0.0f;
  for (int64_t i = R.i0; i <= R.i1; ++i) {
    const int64_t a = R.r0 > P.row_start[i] ? R.r0 : P.row_start[i];
    const int64_t b = R.r1 < P.row_start[i + 1] ? R.r1 : P.row_start[i + 1];
    if (a >= b) continue;
    __syncthreads();  // the previous instance's row is no longer read
    for (int j = threadIdx.x; j < P.bins + 2 * kPad + 1; j += kHistThreads) {
      const int c = j - kPad;
      gh[j] = c >= 0 && c < P.bins ? P.grad_hist[i * P.bins + c] : 0.0f;
    }
    __syncthreads();
    for (int64_t r = a + threadIdx.x; r < b; r += kHistUnroll * kHistThreads) {
      float v[kHistUnroll];
      load_rows<kHistThreads>(P.values, r, b, v);
#pragma unroll
      for (int u = 0; u < kHistUnroll; ++u) {
        const int64_t q = r + int64_t(u) * kHistThreads;
        if (q >= b) break;
        float t;
        int c0;
        float acc = 0.0f;
        if (tap_origin(v[u], P, &t, &c0)) {
          float w[2 * kTaps + 1], d[2 * kTaps + 1];
          tap_weights<FACTORED>(P, t, c0, w, d);
          const float* g = gh + (c0 - kTaps + kPad);
#pragma unroll
          for (int j = 0; j < 2 * kTaps + 1; ++j) acc += g[j] * w[j] * (-d[j] * P.inv_w2);
        }
        P.grad_values[q] = acc * P.inv_bw;
      }
    }
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
  for (int k = 0; k <= kTaps; ++k)
    P.g[k] = float(std::exp(-double(k * k) / (2.0 * double(width) * double(width))));
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
  const dim3 grid(unsigned((n_rows + kHistRows - 1) / kHistRows));
  const size_t lds = size_t(bins + 2 * kPad + 1) * 64 * sizeof(float);
  if (P.inv_2w2 <= kFactoredMaxA)
    hipLaunchKernelGGL(hist_kernel<true>, grid, dim3(64), lds, st, P);
  else
    hipLaunchKernelGGL(hist_kernel<false>, grid, dim3(64), lds, st, P);
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
  const dim3 grid(unsigned((n_rows + kHistRows - 1) / kHistRows));
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (P.inv_2w2 <= kFactoredMaxA)
    hipLaunchKernelGGL(hist_backward_kernel<true>, grid, dim3(kHistThreads), 0, st, P);
  else
    hipLaunchKernelGGL(hist_backward_kernel<false>, grid, dim3(kHistThreads), 0, st, P);
  return check_hip(hipGetLastError(), "soft_histogram_backward launch");
}

}  // extern "C"
