// batched.hip -- the batched 1-D LOITS path for gfx950.
//
// One x-curve per instance (the reference's 1-D mapping, nx = ny = 1).
//   tables_kernel   one work-group per instance: trapezoid norm, CDF (parallel
//                   blocked scan or the reference's sequential chain), integer
//                   thresholds, guide table and per-bin interpolation data
//                   (cpp/loits_core.cpp:55-165).
//   sample_kernel   one work-group per tile of an instance's slots: Philox
//                   uniforms (rng/philox.hpp), interval search over the LDS
//                   tables (cpp/loits_core.cpp:12-18), inverse interpolation
//                   (:190-195) and stable compaction (:239-266) with a
//                   per-tile ordered scan; writes a 1-bit keep mask per slot.
//   finalize_kernel per-tile kept counts -> row offsets, totals, dense flag.
//   pack_kernel     gathers tiles into contiguous rows when events were dropped.
//   vjp_kernel      recomputes Philox + search, accumulates the interpolation
//                   adjoint in an LDS fp64 histogram (ds_add_f64), then the CDF
//                   and density adjoints (cpp/loits_core.cpp:268-472) for the
//                   instance in the last tile to finish.
// All floating-point code is compiled with -ffp-contract=off; fused
// multiply-adds are explicit.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "batched.hpp"
#include "loits_device.hpp"

namespace qloits {

// ---------------------------------------------------------------------------
// Block-wide fp64 helpers (work-group of up to 1024 threads = 16 waves).
__device__ double block_sum(double v, double* scratch) {
  const int wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  v = wave_sum(v);
  __syncthreads();
  if (lane_id() == 0) scratch[wave] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < nwaves; ++w) t += scratch[w];
  __syncthreads();
  return t;
}

// Exclusive prefix (by thread index) of v; *total receives the sum.
__device__ double block_exclusive_scan(double v, double* scratch, double* total) {
  const int wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  const double inc = wave_inclusive_scan(v);
  __syncthreads();
  if (lane_id() == kWave - 1) scratch[wave] = inc;
  __syncthreads();
  double before = 0.0, all = 0.0;
  for (int w = 0; w < nwaves; ++w) {
    const double s = scratch[w];
    if (w < wave) before += s;
    all += s;
  }
  __syncthreads();
  *total = all;
  return before + inc - v;
}

// Suffix sum (by thread index) of v excluding this thread.
__device__ double block_exclusive_suffix(double v, double* scratch) {
  double total;
  const double excl = block_exclusive_scan(v, scratch, &total);
  return total - excl - v;
}

// Norm-sum fusion pattern of the reference build (see oracle/loits_oracle.c):
// the norm of nterms terms is fully fused when nterms <= 3, otherwise only an
// odd trailing term is fused.
__device__ __forceinline__ bool norm_term_fused(int64_t i, int64_t nterms) {
  return nterms <= 3 || ((nterms & 1) && i == nterms - 1);
}

template <typename Real>
struct Tables {
  TableHeader* hdr;
  uint32_t* thr;
  uint32_t* guide;
  typename CoefOf<Real>::type* coef;
  double* cdf;
  float* rlo;
  TableLayout L;
};

template <typename Real>
__device__ __forceinline__ Tables<Real> tables_of(uint8_t* base, const InstanceDesc& in) {
  using Coef = typename CoefOf<Real>::type;
  Tables<Real> t;
  t.L = table_layout(in.n_nodes, in.guide_log2, sizeof(Coef));
  uint8_t* blob = base + in.table_off;
  t.hdr = reinterpret_cast<TableHeader*>(blob);
  t.thr = reinterpret_cast<uint32_t*>(blob + t.L.thr);
  t.guide = reinterpret_cast<uint32_t*>(blob + t.L.guide);
  t.coef = reinterpret_cast<Coef*>(blob + t.L.coef);
  t.cdf = reinterpret_cast<double*>(blob + t.L.cdf);
  t.rlo = reinterpret_cast<float*>(blob + t.L.rlo);
  return t;
}

// ---------------------------------------------------------------------------
// Tables: one work-group per instance.
template <typename Real, bool SEQUENTIAL>
__global__ void __launch_bounds__(1024) tables_kernel(BatchedParams P) {
  __shared__ double scratch[16];
  __shared__ double sh_norm, sh_inv;
  const InstanceDesc in = P.inst[blockIdx.x];
  const int64_t n = in.n_nodes;
  const Real* x = static_cast<const Real*>(P.x) + in.node_off;
  const Real* s = static_cast<const Real*>(P.density) + in.node_off;
  Tables<Real> T = tables_of<Real>(P.tables, in);
  const double acc = in.accept ? 1.0 : 0.0;
  const int tid = threadIdx.x, nt = blockDim.x;

  if (tid == 0) P.done[blockIdx.x] = 0u;

  if (SEQUENTIAL) {
    // The reference's own order: cpp/loits_core.cpp:73-80 then :130-135.
    if (tid == 0) {
      double nrm = 0.0;
      const int64_t nterms = n - 1;
      for (int64_t t = 1; t < n; ++t) {
        const double h = 0.5 * (double(x[t]) - double(x[t - 1]));
        const double sum = double(s[t - 1]) + double(s[t]);
        nrm = norm_term_fused(t - 1, nterms) ? __builtin_fma(h, sum, nrm) : nrm + h * sum;
      }
      const double inv = 1.0 / nrm;
      double cum = 0.0, rho_prev = double(s[0]) * inv;
      T.cdf[0] = 0.0;
      for (int64_t t = 1; t < n; ++t) {
        const double rho = double(s[t]) * inv;
        cum = __builtin_fma(0.5 * (double(x[t]) - double(x[t - 1])), rho_prev + rho, cum);
        T.cdf[t] = cum * acc;
        rho_prev = rho;
      }
      sh_norm = nrm;
      sh_inv = inv;
    }
    __syncthreads();
  } else {
    // Blocked parallel scan: each thread owns a contiguous segment.
    const int64_t per = (n + nt - 1) / nt;
    const int64_t t0 = tid * per, t1 = t0 + per < n ? t0 + per : n;
    double local = 0.0;
    for (int64_t t = t0 > 1 ? t0 : 1; t < t1; ++t)
      local += 0.5 * (double(x[t]) - double(x[t - 1])) * (double(s[t - 1]) + double(s[t]));
    const double nrm = block_sum(local, scratch);
    const double inv = 1.0 / nrm;
    double part = 0.0;
    for (int64_t t = t0 > 1 ? t0 : 1; t < t1; ++t)
      part += 0.5 * (double(x[t]) - double(x[t - 1])) *
              (double(s[t - 1]) * inv + double(s[t]) * inv);
    double total;
    double run = block_exclusive_scan(part, scratch, &total);
    for (int64_t t = t0; t < t1; ++t) {
      if (t == 0) {
        T.cdf[0] = 0.0;
        continue;
      }
      run += 0.5 * (double(x[t]) - double(x[t - 1])) *
             (double(s[t - 1]) * inv + double(s[t]) * inv);
      T.cdf[t] = run * acc;
    }
    if (tid == 0) {
      sh_norm = nrm;
      sh_inv = inv;
    }
    __syncthreads();
  }
  const double nrm = sh_norm;

  // Thresholds, monotonicity, per-bin interpolation data.
  bool mono = true;
  for (int64_t t = tid; t < n; t += nt) {
    const double c = T.cdf[t];
    const uint32_t thr = threshold_of(c);
    T.thr[t] = thr;
    if (t > 0) mono = mono && thr >= threshold_of(T.cdf[t - 1]);
    if (P.cdf_out) P.cdf_out[in.node_off + t] = c;
    if (t + 1 < n) {
      const double c1 = T.cdf[t + 1];
      const double inv_d = 1.0 / (c1 - c + kEpsilon);       // cpp/loits_core.cpp:193
      const double m = (double(x[t + 1]) - double(x[t])) * inv_d;  // :194
      if constexpr (sizeof(Real) == 4) {
        const float r = static_cast<float>(double(thr) * 0x1.0p-24 - c);
        T.coef[t] = CoefF32{thr, r, static_cast<float>(m), static_cast<float>(x[t])};
        T.rlo[t] = r;
      } else {
        T.coef[t] = CoefF64{c, m, double(x[t]), thr, 0u};
      }
    }
  }
  mono = __syncthreads_and(mono);

  // Guide: guide[g] = #{t : T_t < g << shift} (monotone thresholds only).
  const int64_t G = int64_t(1) << in.guide_log2;
  const int shift = 24 - in.guide_log2;
  if (mono) {
    for (int64_t g = tid; g <= G; g += nt) {
      const uint64_t key = uint64_t(g) << shift;
      int64_t lo = 0, hi = n;
      while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (uint64_t(T.thr[mid]) < key) lo = mid + 1; else hi = mid;
      }
      T.guide[g] = static_cast<uint32_t>(lo);
    }
  }
  if (tid == 0) {
    T.hdr->norm = nrm;
    T.hdr->inv_n = sh_inv;
    T.hdr->flags = mono ? kMonotone : 0u;
    T.hdr->n = static_cast<uint32_t>(n);
    T.hdr->guide_log2 = static_cast<uint32_t>(in.guide_log2);
    if (P.norm_out) P.norm_out[blockIdx.x] = nrm;
  }
}

// ---------------------------------------------------------------------------
// Interval search: count(k) = #{t : T_t <= k}; j = clamp(count - 1, 0, n - 2).
__device__ __forceinline__ int find_bin(uint32_t k, const uint32_t* thr, const uint32_t* guide,
                                        int shift, int n, bool mono) {
  int count;
  if (mono) {
    const uint32_t g = k >> shift;
    int lo = static_cast<int>(guide[g]), hi = static_cast<int>(guide[g + 1]);
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (thr[mid] <= k) lo = mid + 1; else hi = mid;
    }
    count = lo;
  } else {
    count = 0;
    for (int t = 0; t < n; ++t) count += thr[t] <= k ? 1 : 0;
  }
  int j = count - 1;
  j = j < 0 ? 0 : j;
  return j > n - 2 ? n - 2 : j;
}

__device__ __forceinline__ bool keep_value(double v, const BatchedParams& P) {
  if (!(__builtin_isfinite(v) && v != 0.0)) return false;
  if (P.use_window && !(v >= P.win_lo && v <= P.win_hi)) return false;
  return true;
}

// The reference's fp64 interpolation for bin j (cpp/loits_core.cpp:190-195).
template <typename Real>
__device__ __forceinline__ double interp_reference(const double* cdf, const Real* x, int j,
                                                   uint32_t k) {
  const double c0 = cdf[j];
  const double inv_d = 1.0 / (cdf[j + 1] - c0 + kEpsilon);
  const double m = (double(x[j + 1]) - double(x[j])) * inv_d;
  return __builtin_fma(m, double(k) * 0x1.0p-24 - c0, double(x[j]));
}

// Value and keep flag of one slot.  fp32: single-precision evaluation of the
// same formula; results whose keep decision could differ from the fp64
// decision (near 0, near overflow, within 1e-6 of a window edge, non-finite)
// are re-evaluated in fp64 from the global tables.
template <typename Real>
__device__ __forceinline__ Real slot_value(uint32_t k, int j,
                                           const typename CoefOf<Real>::type* coef,
                                           const double* cdf_g, const Real* x_g,
                                           const BatchedParams& P, bool* keep) {
  if constexpr (sizeof(Real) == 4) {
    const CoefF32 cf = coef[j];
    const float t = static_cast<float>(static_cast<int32_t>(k - cf.thr)) * 0x1.0p-24f + cf.r;
    const float v = __builtin_fmaf(cf.m, t, cf.x);
    const float a = __builtin_fabsf(v);
    bool safe = a >= 1e-30f && a <= 1e30f;
    if (P.use_window) {
      const float lo = static_cast<float>(P.win_lo), hi = static_cast<float>(P.win_hi);
      const float margin = 2e-6f * a;
      safe = safe && __builtin_fabsf(v - lo) > margin && __builtin_fabsf(v - hi) > margin;
    }
    if (__builtin_expect(safe, 1)) {
      *keep = !P.use_window || (v >= static_cast<float>(P.win_lo) && v <= static_cast<float>(P.win_hi));
      return v;
    }
    const double v64 = interp_reference(cdf_g, x_g, j, k);
    *keep = keep_value(v64, P);
    return static_cast<float>(v64);
  } else {
    const CoefF64 cf = coef[j];
    const double v = __builtin_fma(cf.m, double(k) * 0x1.0p-24 - cf.c, cf.x);
    *keep = keep_value(v, P);
    return v;
  }
}

__device__ __forceinline__ void lds_copy(void* dst, const void* src, int64_t bytes) {
  const uint4* s = static_cast<const uint4*>(src);
  uint4* d = static_cast<uint4*>(dst);
  for (int64_t i = threadIdx.x; i < bytes / 16; i += blockDim.x) d[i] = s[i];
}

// Ordered work-group scan of per-thread counts (0..4): returns this thread's
// exclusive offset and the work-group total.  One barrier; the per-wave totals
// are double-buffered by `parity`.
__device__ __forceinline__ int wg_scan_counts(int cnt, int32_t (*wave_tot)[16], int parity,
                                              int* total) {
  const uint64_t lt = lanemask_lt();
  const uint64_t b0 = ballot(cnt & 1), b1 = ballot(cnt & 2), b2 = ballot(cnt & 4);
  const int pre = popc64(b0 & lt) + 2 * popc64(b1 & lt) + 4 * popc64(b2 & lt);
  const int wt = popc64(b0) + 2 * popc64(b1) + 4 * popc64(b2);
  const int wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  if (lane_id() == 0) wave_tot[parity][wave] = wt;
  __syncthreads();
  int base = 0, all = 0;
  for (int w = 0; w < nwaves; ++w) {
    const int v = wave_tot[parity][w];
    all += v;
    base += w < wave ? v : 0;
  }
  *total = all;
  return base + pre;
}

// ---------------------------------------------------------------------------
// Sampling: one work-group per tile (chunk).
template <typename Real>
__global__ void __launch_bounds__(1024) sample_kernel(BatchedParams P) {
  using Coef = typename CoefOf<Real>::type;
  extern __shared__ __align__(16) uint8_t lds[];
  __shared__ int32_t wave_tot[2][16];
  const int c = blockIdx.x;
  const ChunkDesc ch = P.chunks[c];
  const InstanceDesc in = P.inst[ch.instance];
  Tables<Real> T = tables_of<Real>(P.tables, in);
  const int n = static_cast<int>(in.n_nodes);
  const int shift = 24 - in.guide_log2;

  lds_copy(lds, T.thr, T.L.cdf - T.L.thr);
  __syncthreads();
  const uint32_t* s_thr = reinterpret_cast<const uint32_t*>(lds);
  const uint32_t* s_guide = reinterpret_cast<const uint32_t*>(lds + (T.L.guide - T.L.thr));
  const Coef* s_coef = reinterpret_cast<const Coef*>(lds + (T.L.coef - T.L.thr));
  const bool mono = (T.hdr->flags & kMonotone) != 0;
  const Real* x_g = static_cast<const Real*>(P.x) + in.node_off;

  const uint64_t p_lo = in.counter_base + uint64_t(ch.slot_lo);
  const uint64_t p_hi = in.counter_base + uint64_t(ch.slot_hi);
  const uint64_t q_begin = p_lo >> 2;
  const uint64_t q_end = ch.slot_hi > ch.slot_lo ? (p_hi + 3) >> 2 : q_begin;  // empty tile: no work
  Real* out = static_cast<Real*>(P.events) + in.out_off + ch.slot_lo;
  int32_t* bins_out = P.bins ? P.bins + in.out_off + ch.slot_lo : nullptr;
  uint64_t* mask = P.keepmask + ch.mask_off;
  const int wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  int64_t running = 0;
  int parity = 0, it = 0;
  for (uint64_t q0 = q_begin; q0 < q_end; q0 += blockDim.x, ++it) {
    const uint64_t q = q0 + threadIdx.x;
    const Words4 w = philox4x32_10(q, in.stream, P.seed);
    Real v[4];
    int jb[4];
    int keepbits = 0;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const uint64_t p = 4 * q + e;
      const bool live = q < q_end && p >= p_lo && p < p_hi;
      const uint32_t k = word_of(w, e) >> 8;
      const int j = find_bin(k, s_thr, s_guide, shift, n, mono);
      bool keep;
      v[e] = slot_value<Real>(k, j, s_coef, T.cdf, x_g, P, &keep);
      jb[e] = j;
      keepbits |= (live && keep) ? (1 << e) : 0;
    }
    // keep mask: 4 ballots per wave per iteration, one bit per slot
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const uint64_t b = ballot((keepbits >> e) & 1);
      if (lane_id() == e) mask[(int64_t(it) * nwaves + wave) * 4 + e] = b;
    }
    int total;
    const int off = wg_scan_counts(__popc(keepbits), wave_tot, parity, &total);
    parity ^= 1;
    int64_t row = running + off;
    if (keepbits == 0xF && ((row & 3) == 0) && ((in.out_off + ch.slot_lo) & 3) == 0) {
      if constexpr (sizeof(Real) == 4) {
        *reinterpret_cast<float4*>(out + row) = make_float4(v[0], v[1], v[2], v[3]);
      } else {
        reinterpret_cast<double2*>(out + row)[0] = make_double2(v[0], v[1]);
        reinterpret_cast<double2*>(out + row)[1] = make_double2(v[2], v[3]);
      }
      if (bins_out)
        *reinterpret_cast<int4*>(bins_out + row) = make_int4(jb[0], jb[1], jb[2], jb[3]);
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        if ((keepbits >> e) & 1) {
          out[row] = v[e];
          if (bins_out) bins_out[row] = jb[e];
          ++row;
        }
      }
    }
    running += total;
  }
  if (threadIdx.x == 0) P.chunk_kept[c] = running;
}

// ---------------------------------------------------------------------------
// Finalize: one work-group.  chunk_row = exclusive scan of chunk_kept; kept
// per instance; total; dense flag (every tile kept all of its slots).
__global__ void __launch_bounds__(1024) finalize_kernel(BatchedParams P) {
  __shared__ double scratch[16];
  __shared__ int sh_sparse;
  if (threadIdx.x == 0) sh_sparse = 0;
  __syncthreads();
  double carry = 0.0;  // counts < 2^53: exact in double
  bool sparse = false;
  for (int64_t base = 0; base < P.n_chunks; base += blockDim.x) {
    const int64_t c = base + threadIdx.x;
    double v = 0.0;
    if (c < P.n_chunks) {
      const ChunkDesc ch = P.chunks[c];
      v = double(P.chunk_kept[c]);
      sparse = sparse || P.chunk_kept[c] != ch.slot_hi - ch.slot_lo;
    }
    double tot;
    const double ex = block_exclusive_scan(v, scratch, &tot);
    if (c < P.n_chunks) P.chunk_row[c] = static_cast<int64_t>(carry + ex);
    carry += tot;
  }
  if (sparse) atomicOr(&sh_sparse, 1);
  __syncthreads();
  for (int64_t i = threadIdx.x; i < P.n_instances; i += blockDim.x) {
    const InstanceDesc in = P.inst[i];
    int64_t k = 0;
    for (int c = in.first_chunk; c < in.first_chunk + in.n_chunks; ++c) k += P.chunk_kept[c];
    P.kept[i] = k;
  }
  if (threadIdx.x == 0) {
    P.total[0] = static_cast<int64_t>(carry);
    P.flags[0] = sh_sparse ? 0 : 1;  // dense
  }
}

// Pack: one work-group per tile; a no-op when every tile is full.
template <typename Real>
__global__ void __launch_bounds__(256) pack_kernel(BatchedParams P, Real* packed,
                                                   int32_t* packed_bins) {
  if (P.flags[0]) return;
  const ChunkDesc ch = P.chunks[blockIdx.x];
  const InstanceDesc in = P.inst[ch.instance];
  const int64_t cnt = P.chunk_kept[blockIdx.x];
  const int64_t src = in.out_off + ch.slot_lo, dst = P.chunk_row[blockIdx.x];
  const Real* ev = static_cast<const Real*>(P.events);
  for (int64_t r = threadIdx.x; r < cnt; r += blockDim.x) {
    packed[dst + r] = ev[src + r];
    if (packed_bins && P.bins) packed_bins[dst + r] = P.bins[src + r];
  }
}

// ---------------------------------------------------------------------------
// Tail of the VJP for one instance, after its histogram (h0 = sum of
// cotangents, h1 = sum of cotangent * (u - C_j), per bin j) is complete in
// LDS.  gC (cpp/loits_core.cpp:294-301 summed per bin), the CDF adjoint
// (:355-365) and the density adjoint (:416-431) with blocked parallel scans.
template <typename Real, int MAXPER>
__device__ void vjp_tail(const BatchedParams& P, const InstanceDesc& in, const Tables<Real>& T,
                         const double* h0, const double* h1, Real* grad_density,
                         double* scratch) {
  const int64_t n = in.n_nodes;
  const Real* x = static_cast<const Real*>(P.x) + in.node_off;
  const Real* s = static_cast<const Real*>(P.density) + in.node_off;
  const double* cdf = T.cdf;
  const double acc = in.accept ? 1.0 : 0.0;
  const double inv_n = T.hdr->inv_n;
  const int nt = blockDim.x;
  const int64_t per = (n + nt - 1) / nt;
  const int64_t t0 = threadIdx.x * per;
  double gc[MAXPER];
#pragma unroll
  for (int k = 0; k < MAXPER; ++k) {
    const int64_t t = t0 + k;
    double g = 0.0;
    if (k < per && t < n) {
      if (t + 1 < n) {
        const double inv_d = 1.0 / (cdf[t + 1] - cdf[t] + kEpsilon);
        const double db = double(x[t + 1]) - double(x[t]);
        g += db * (inv_d * inv_d * h1[t] - inv_d * h0[t]);
      }
      if (t >= 1) {
        const double inv_d = 1.0 / (cdf[t] - cdf[t - 1] + kEpsilon);
        const double db = double(x[t]) - double(x[t - 1]);
        g -= db * inv_d * inv_d * h1[t - 1];
      }
    }
    gc[k] = g;
  }
  // local suffix sums, then the work-group suffix of the segment totals
  double seg = 0.0;
#pragma unroll
  for (int k = MAXPER - 1; k >= 0; --k) {
    seg += gc[k];
    gc[k] = seg;  // gc[k] now holds sum_{k' >= k} within the segment
  }
  const double after = block_exclusive_suffix(seg, scratch);
  // S_t = gc[k] + after; grho_t = acc * (h_t S_t [t>=1] + h_{t+1} S_{t+1} [t<=n-2])
  double grho[MAXPER];
  double local_norm = 0.0;
#pragma unroll
  for (int k = 0; k < MAXPER; ++k) {
    const int64_t t = t0 + k;
    double g = 0.0;
    if (k < per && t < n) {
      const double S_t = gc[k] + after;
      const double S_next = (k + 1 < MAXPER && k + 1 < per) ? gc[k + 1] + after : after;
      if (t >= 1) g += 0.5 * (double(x[t]) - double(x[t - 1])) * acc * S_t;
      if (t + 1 < n) g += 0.5 * (double(x[t + 1]) - double(x[t])) * acc * S_next;
      local_norm += g * double(s[t]);
    }
    grho[k] = g;
  }
  const double grad_norm = -block_sum(local_norm, scratch) * inv_n * inv_n;
  Real* out = grad_density + in.node_off;
#pragma unroll
  for (int k = 0; k < MAXPER; ++k) {
    const int64_t t = t0 + k;
    if (k < per && t < n) {
      double wgt = 0.0;
      if (t >= 1) wgt += 0.5 * (double(x[t]) - double(x[t - 1]));
      if (t + 1 < n) wgt += 0.5 * (double(x[t + 1]) - double(x[t]));
      out[t] = static_cast<Real>(grho[k] * inv_n + grad_norm * wgt);
    }
  }
}

// VJP: one work-group per tile; the tile that completes an instance runs the tail.
template <typename Real, int MAXPER>
__global__ void __launch_bounds__(1024) vjp_kernel(BatchedParams P, const Real* grad,
                                                   Real* grad_density) {
  extern __shared__ __align__(16) uint8_t lds[];
  __shared__ int32_t wave_tot[2][16];
  __shared__ double scratch[16];
  __shared__ int sh_last;
  const int c = blockIdx.x;
  const ChunkDesc ch = P.chunks[c];
  const InstanceDesc in = P.inst[ch.instance];
  Tables<Real> T = tables_of<Real>(P.tables, in);
  const int n = static_cast<int>(in.n_nodes);
  const int shift = 24 - in.guide_log2;

  // LDS: thr | guide | (r: fp32, C: fp64) | h0 | h1
  const int64_t search_bytes = T.L.coef - T.L.thr;
  const int64_t nb = pad4(n);
  lds_copy(lds, T.thr, search_bytes);
  uint8_t* lds_t = lds + search_bytes;
  if constexpr (sizeof(Real) == 4) lds_copy(lds_t, T.rlo, nb * 4);
  else lds_copy(lds_t, T.cdf, nb * 8);
  double* h0 = reinterpret_cast<double*>(lds_t + nb * sizeof(Real));
  double* h1 = h0 + nb;
  for (int64_t i = threadIdx.x; i < 2 * nb; i += blockDim.x) h0[i] = 0.0;
  __syncthreads();
  const uint32_t* s_thr = reinterpret_cast<const uint32_t*>(lds);
  const uint32_t* s_guide = reinterpret_cast<const uint32_t*>(lds + (T.L.guide - T.L.thr));
  const bool mono = (T.hdr->flags & kMonotone) != 0;

  const uint64_t p_lo = in.counter_base + uint64_t(ch.slot_lo);
  const uint64_t p_hi = in.counter_base + uint64_t(ch.slot_hi);
  const uint64_t q_begin = p_lo >> 2;
  const uint64_t q_end = ch.slot_hi > ch.slot_lo ? (p_hi + 3) >> 2 : q_begin;  // empty tile: no work
  const bool full = P.chunk_kept[c] == ch.slot_hi - ch.slot_lo;
  const int64_t row0 = P.chunk_row[c];
  const uint64_t* mask = P.keepmask + ch.mask_off;
  const int wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  int64_t running = 0;
  int parity = 0, it = 0;
  for (uint64_t q0 = q_begin; q0 < q_end; q0 += blockDim.x, ++it) {
    const uint64_t q = q0 + threadIdx.x;
    const Words4 w = philox4x32_10(q, in.stream, P.seed);
    int keepbits = 0;
    if (full) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const uint64_t p = 4 * q + e;
        keepbits |= (q < q_end && p >= p_lo && p < p_hi) ? (1 << e) : 0;
      }
    } else {
      const uint64_t* mw = mask + (int64_t(it) * nwaves + wave) * 4;
#pragma unroll
      for (int e = 0; e < 4; ++e) keepbits |= int((mw[e] >> lane_id()) & 1ull) << e;
    }
    int64_t row;
    if (full) {
      row = row0 + int64_t(4 * q) - int64_t(p_lo);  // first slot of this block
    } else {
      int total;
      const int off = wg_scan_counts(__popc(keepbits), wave_tot, parity, &total);
      parity ^= 1;
      row = row0 + running + off;
      running += total;
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      if (full) {
        if (!((keepbits >> e) & 1)) continue;
      } else if (!((keepbits >> e) & 1)) {
        continue;
      }
      const uint32_t k = word_of(w, e) >> 8;
      const int j = find_bin(k, s_thr, s_guide, shift, n, mono);
      double t;
      if constexpr (sizeof(Real) == 4) {
        const float* s_r = reinterpret_cast<const float*>(lds_t);
        t = double(static_cast<float>(static_cast<int32_t>(k - s_thr[j])) * 0x1.0p-24f + s_r[j]);
      } else {
        const double* s_c = reinterpret_cast<const double*>(lds_t);
        t = double(k) * 0x1.0p-24 - s_c[j];
      }
      const int64_t r = full ? row + e : row++;
      const double g = double(grad[r]);
      atomicAdd(&h0[j], g);
      atomicAdd(&h1[j], g * t);
    }
  }
  __syncthreads();

  if (in.n_chunks == 1) {
    vjp_tail<Real, MAXPER>(P, in, T, h0, h1, grad_density, scratch);
    return;
  }
  // multi-tile instance: publish partial sums; the last tile reduces them in order
  double* part = P.partials + ch.partial_off;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    part[i] = h0[i];
    part[nb + i] = h1[i];
  }
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint32_t old = atomicAdd(&P.done[ch.instance], 1u);
    sh_last = old == uint32_t(in.n_chunks - 1);
  }
  __syncthreads();
  if (!sh_last) return;
  __threadfence();
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    double a = 0.0, b = 0.0;
    for (int cc = in.first_chunk; cc < in.first_chunk + in.n_chunks; ++cc) {
      const double* pp = P.partials + P.chunks[cc].partial_off;
      a += __builtin_nontemporal_load(pp + i);
      b += __builtin_nontemporal_load(pp + nb + i);
    }
    h0[i] = a;
    h1[i] = b;
  }
  if (threadIdx.x == 0) P.done[ch.instance] = 0u;
  __syncthreads();
  vjp_tail<Real, MAXPER>(P, in, T, h0, h1, grad_density, scratch);
}

// ---------------------------------------------------------------------------
// Host-side launchers (called from capi.cpp).
template <typename Real>
hipError_t launch_tables(const BatchedParams& P, int threads, bool sequential,
                         hipStream_t stream) {
  if (sequential)
    hipLaunchKernelGGL((tables_kernel<Real, true>), dim3(P.n_instances), dim3(threads), 0,
                       stream, P);
  else
    hipLaunchKernelGGL((tables_kernel<Real, false>), dim3(P.n_instances), dim3(threads), 0,
                       stream, P);
  return hipGetLastError();
}

template <typename Real>
hipError_t launch_sample(const BatchedParams& P, int threads, int64_t lds, hipStream_t stream) {
  auto kern = sample_kernel<Real>;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kern, dim3(P.n_chunks), dim3(threads), size_t(lds), stream, P);
  return hipGetLastError();
}

hipError_t launch_finalize(const BatchedParams& P, hipStream_t stream) {
  hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(1024), 0, stream, P);
  return hipGetLastError();
}

template <typename Real>
hipError_t launch_pack(const BatchedParams& P, void* packed, int32_t* packed_bins,
                       hipStream_t stream) {
  hipLaunchKernelGGL((pack_kernel<Real>), dim3(P.n_chunks), dim3(256), 0, stream, P,
                     static_cast<Real*>(packed), packed_bins);
  return hipGetLastError();
}

template <typename Real>
hipError_t launch_vjp(const BatchedParams& P, int threads, int64_t lds, int maxper,
                      const void* grad, void* grad_density, hipStream_t stream) {
#define QL_VJP_CASE(MP)                                                                      \
  if (maxper <= MP) {                                                                        \
    auto kern = vjp_kernel<Real, MP>;                                                        \
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),                 \
                                       hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)); \
    if (e != hipSuccess) return e;                                                           \
    hipLaunchKernelGGL(kern, dim3(P.n_chunks), dim3(threads), size_t(lds), stream, P,       \
                       static_cast<const Real*>(grad), static_cast<Real*>(grad_density));    \
    return hipGetLastError();                                                                \
  }
  QL_VJP_CASE(2)
  QL_VJP_CASE(4)
  QL_VJP_CASE(8)
  QL_VJP_CASE(16)
  QL_VJP_CASE(32)
#undef QL_VJP_CASE
  return hipErrorInvalidValue;
}

template hipError_t launch_tables<float>(const BatchedParams&, int, bool, hipStream_t);
template hipError_t launch_tables<double>(const BatchedParams&, int, bool, hipStream_t);
template hipError_t launch_sample<float>(const BatchedParams&, int, int64_t, hipStream_t);
template hipError_t launch_sample<double>(const BatchedParams&, int, int64_t, hipStream_t);
template hipError_t launch_pack<float>(const BatchedParams&, void*, int32_t*, hipStream_t);
template hipError_t launch_pack<double>(const BatchedParams&, void*, int32_t*, hipStream_t);
template hipError_t launch_vjp<float>(const BatchedParams&, int, int64_t, int, const void*, void*,
                                      hipStream_t);
template hipError_t launch_vjp<double>(const BatchedParams&, int, int64_t, int, const void*, void*,
                                       hipStream_t);

}  // namespace qloits
