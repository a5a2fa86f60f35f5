// batched.hip -- the batched 1-D LOITS path for gfx950.
//
// One x-curve per instance (the reference's 1-D mapping, nx = ny = 1).
//   tables_kernel   one work-group per instance: trapezoid norm, CDF (parallel
//                   blocked scan or the reference's sequential chain), integer
//                   thresholds (sorted when the CDF is not monotone), a packed
//                   guide table and per-bin interpolation data
//                   (cpp/loits_core.cpp:55-165, 190-195).
//   sample_kernel   one work-group per chunk of an instance's slots; each wave
//                   owns a contiguous run of Philox blocks (a segment): Philox
//                   uniforms (rng/philox.hpp), interval search over the LDS
//                   tables (cpp/loits_core.cpp:12-18) with four searches in
//                   lockstep, inverse interpolation (:190-195) and a stable,
//                   barrier-free compaction of the segment (:239-266); writes
//                   a 1-bit keep mask per slot.
//   finalize_kernel segment kept counts -> packed row offsets, totals, dense flag.
//   pack_kernel     gathers segments into contiguous rows when events were dropped.
//   vjp_kernel      recomputes Philox + search, accumulates the interpolation
//                   adjoint in an LDS fp64 histogram (ds_add_f64), then the CDF
//                   and density adjoints (cpp/loits_core.cpp:268-472) for the
//                   instance in the last chunk to finish.
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

// Norm-sum fusion pattern of the reference build (oracle/loits_oracle.c,
// curve_density): all terms fused when there are at most three, otherwise
// only an odd trailing term.
__device__ __forceinline__ bool norm_term_fused(int64_t i, int64_t nterms) {
  return nterms <= 3 || ((nterms & 1) && i == nterms - 1);
}

__device__ __forceinline__ int ceil_log2_dev(int64_t v) {
  int l = 0;
  while ((int64_t(1) << l) < v) ++l;
  return l;
}

template <typename Real>
struct Tables {
  TableHeader* hdr;
  uint32_t* srt;
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
  t.srt = reinterpret_cast<uint32_t*>(blob + t.L.thr);
  t.guide = reinterpret_cast<uint32_t*>(blob + t.L.guide);
  t.coef = reinterpret_cast<Coef*>(blob + t.L.coef);
  t.cdf = reinterpret_cast<double*>(blob + t.L.cdf);
  t.rlo = reinterpret_cast<float*>(blob + t.L.rlo);
  return t;
}

// ---------------------------------------------------------------------------
// Tables: one work-group per instance.  Dynamic LDS: 4 * next_pow2(max n)
// bytes, used only to sort the thresholds of a non-monotone CDF.
template <typename Real, bool SEQUENTIAL>
__global__ void __launch_bounds__(1024) tables_kernel(BatchedParams P) {
  extern __shared__ __align__(16) uint32_t sort_lds[];
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

  // Thresholds, monotonicity, per-bin interpolation data, value-range check.
  bool mono = true, finite = true, pos = true, neg = true;
  for (int64_t t = tid; t < n; t += nt) {
    const double c = T.cdf[t];
    const uint32_t thr = threshold_of(c);
    T.srt[t] = thr;
    if (t > 0) mono = mono && thr >= threshold_of(T.cdf[t - 1]);
    if (P.cdf_out) P.cdf_out[in.node_off + t] = c;
    const double xt = double(x[t]);
    const double ax = __builtin_fabs(xt);
    finite = finite && __builtin_isfinite(c) && ax >= 1e-30 && ax <= 1e29;
    pos = pos && xt > 0.0;
    neg = neg && xt < 0.0;
    if (t + 1 < n) {
      const double c1 = T.cdf[t + 1];
      const double inv_d = 1.0 / (c1 - c + kEpsilon);              // cpp/loits_core.cpp:193
      const double m = (double(x[t + 1]) - xt) * inv_d;            // :194
      finite = finite && __builtin_isfinite(m);
      if constexpr (sizeof(Real) == 4) {
        const float r = static_cast<float>(double(thr) * 0x1.0p-24 - c);
        T.coef[t] = CoefF32{thr, r, static_cast<float>(m), static_cast<float>(xt)};
        T.rlo[t] = r;
      } else {
        T.coef[t] = CoefF64{c, m, xt, thr, 0u};
      }
    }
  }
  mono = __syncthreads_and(mono);
  finite = __syncthreads_and(finite);
  pos = __syncthreads_and(pos);
  neg = __syncthreads_and(neg);
  const bool safe = finite && (pos || neg);

  if (!mono) {
    // count(k) = #{t : T_t <= k} needs the thresholds in sorted order; bins
    // keep their own T_j in coef.  Bitonic sort in LDS.
    int64_t pw = 1;
    while (pw < n) pw <<= 1;
    for (int64_t i = tid; i < pw; i += nt) sort_lds[i] = i < n ? T.srt[i] : kNever;
    __syncthreads();
    for (int64_t k = 2; k <= pw; k <<= 1) {
      for (int64_t j = k >> 1; j > 0; j >>= 1) {
        for (int64_t i = tid; i < pw; i += nt) {
          const int64_t ixj = i ^ j;
          if (ixj > i) {
            const uint32_t a = sort_lds[i], b = sort_lds[ixj];
            const bool up = (i & k) == 0;
            if ((a > b) == up) {
              sort_lds[i] = b;
              sort_lds[ixj] = a;
            }
          }
        }
        __syncthreads();
      }
    }
    for (int64_t i = tid; i < n; i += nt) T.srt[i] = sort_lds[i];
    __syncthreads();
  }

  // Level-1 guide by scatter: srt position t is the lower bound of every cell
  // start key g << shift with srt_{t-1} < key <= srt_t (srt_{-1} = -inf,
  // srt_n = +inf).  Each cell word is written as two 16-bit halves.
  const int64_t G = int64_t(1) << in.guide_log2;
  const int shift = 24 - in.guide_log2;
  uint16_t* g16 = reinterpret_cast<uint16_t*>(T.guide);
  for (int64_t t = tid; t <= n; t += nt) {
    const int64_t ga = t == 0 ? 0 : int64_t(T.srt[t - 1] >> shift) + 1;
    int64_t gb = t == n ? G : int64_t(T.srt[t] >> shift);
    gb = gb > G ? G : gb;
    for (int64_t g = ga; g <= gb; ++g) {
      if (g < G) g16[2 * g] = static_cast<uint16_t>(t);      // lo of cell g
      if (g >= 1) g16[2 * g - 1] = static_cast<uint16_t>(t);  // hi of cell g - 1
    }
  }
  // sorted thresholds in LDS for the level-2 searches (the sort already left
  // them there when the CDF was not monotone)
  if (mono)
    for (int64_t i = tid; i < n; i += nt) sort_lds[i] = T.srt[i];
  __syncthreads();

  // Level 2: crowded cells (hi - lo >= 2), in cell order, get kSub sub-cells
  // each while the G words of level 2 last.
  const int64_t cap_cells = G >> kSubLog2;
  int32_t* crowd_list = reinterpret_cast<int32_t*>(sort_lds + pad4(int64_t(1) << ceil_log2_dev(n)));
  double carry = 0.0;
  for (int64_t base = 0; base < G; base += nt) {
    const int64_t g = base + tid;
    bool crowd = false;
    if (g < G) {
      const uint32_t w = T.guide[g];
      crowd = int(w >> 16) - int(w & 0xFFFFu) >= 2;
    }
    double tot;
    const int64_t rank = int64_t(carry + block_exclusive_scan(crowd ? 1.0 : 0.0, scratch, &tot));
    if (crowd && rank < cap_cells) crowd_list[rank] = int32_t(g);
    carry += tot;
  }
  __syncthreads();
  const int64_t n_assigned = int64_t(carry) < cap_cells ? int64_t(carry) : cap_cells;
  const int sub_shift = shift - kSubLog2;
  for (int64_t item = tid; item < n_assigned * kSub; item += nt) {
    const int64_t r = item >> kSubLog2, sub = item & (kSub - 1);
    const int64_t g = crowd_list[r];
    const uint32_t w = T.guide[g];
    const int lo0 = int(w & 0xFFFFu), hi0 = int(w >> 16);
    const uint64_t key_lo = (uint64_t(g) << shift) + (uint64_t(sub) << sub_shift);
    const uint64_t key_hi = key_lo + (uint64_t(1) << sub_shift);
    int a = lo0, b = hi0;  // first position with srt >= key_lo
    while (a < b) {
      const int mid = (a + b) >> 1;
      if (uint64_t(sort_lds[mid]) < key_lo) a = mid + 1; else b = mid;
    }
    int c2 = a, d2 = hi0;  // first position with srt >= key_hi
    while (c2 < d2) {
      const int mid = (c2 + d2) >> 1;
      if (uint64_t(sort_lds[mid]) < key_hi) c2 = mid + 1; else d2 = mid;
    }
    T.guide[G + r * kSub + sub] = uint32_t(a) | (uint32_t(c2) << 16);
  }
  __syncthreads();  // every level-1 word of an assigned cell has been read
  for (int64_t r = tid; r < n_assigned; r += nt)
    T.guide[crowd_list[r]] = kIndirect | uint32_t(G + r * kSub);

  if (tid == 0) {
    T.hdr->norm = nrm;
    T.hdr->inv_n = sh_inv;
    const bool last_full = T.srt[n - 1] >= 16777214u;
    T.hdr->flags = (mono ? kMonotone : 0u) | ((mono && safe && acc == 1.0 && last_full) ? kSafe : 0u);
    T.hdr->n = static_cast<uint32_t>(n);
    T.hdr->guide_log2 = static_cast<uint32_t>(in.guide_log2);
    if (P.norm_out) P.norm_out[blockIdx.x] = nrm;
  }
}

// ---------------------------------------------------------------------------
// Four interval searches: count(k) = #{t : srt_t <= k}, j = clamp(count - 1,
// 0, n - 2).  The two-level guide narrows each search to [lo, lo + len); one
// bisection step (always taken) settles len <= 1, and the loop after it runs
// only for the rare cells that still hold several thresholds.
__device__ __forceinline__ void bisect_step(uint32_t k, const uint32_t* srt, int n, int& lo,
                                            int& len) {
  const int half = len >> 1;
  int idx = lo + half;
  idx = idx < n ? idx : n - 1;
  const bool le = srt[idx] <= k;
  const bool act = len > 0;
  lo = (act && le) ? idx + 1 : lo;
  len = act ? (le ? len - half - 1 : half) : 0;
}

__device__ __forceinline__ void find_bins4(const uint32_t (&k)[4], const uint32_t* srt,
                                           const uint32_t* guide, int shift, int n,
                                           int (&j)[4]) {
  int lo[4], len[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    uint32_t w = guide[k[e] >> shift];
    if (w & kIndirect)
      w = guide[(w & ~kIndirect) + ((k[e] >> (shift - kSubLog2)) & (kSub - 1))];
    lo[e] = int(w & 0xFFFFu);
    len[e] = int((w >> 16) & 0x7FFFu) - lo[e];
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) bisect_step(k[e], srt, n, lo[e], len[e]);
  while ((len[0] | len[1] | len[2] | len[3]) != 0) {
#pragma unroll
    for (int e = 0; e < 4; ++e) bisect_step(k[e], srt, n, lo[e], len[e]);
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    int jj = lo[e] - 1;
    jj = jj < 0 ? 0 : jj;
    j[e] = jj > n - 2 ? n - 2 : jj;
  }
}

template <bool WINDOW>
__device__ __forceinline__ bool keep_value(double v, const BatchedParams& P) {
  if (!(__builtin_isfinite(v) && v != 0.0)) return false;
  if (WINDOW && !(v >= P.win_lo && v <= P.win_hi)) return false;
  return true;
}

// The reference's fp64 interpolation for bin j (cpp/loits_core.cpp:190-195).
template <typename Real>
__device__ __noinline__ double interp_reference(const double* cdf, const Real* x, int j,
                                                uint32_t k) {
  const double c0 = cdf[j];
  const double inv_d = 1.0 / (cdf[j + 1] - c0 + kEpsilon);
  const double m = (double(x[j + 1]) - double(x[j])) * inv_d;
  return __builtin_fma(m, double(k) * 0x1.0p-24 - c0, double(x[j]));
}

// Values and keep flags of a lane's four slots.  fp64: the reference's
// operands and operation order.  fp32: single-precision evaluation of the same
// formula, u - C_j = (k - T_j) 2^-24 + r_j; a value whose keep decision could
// differ from the fp64 decision (|v| < 1e-30 or > 1e30, non-finite, or within
// 2e-6 |v| of a window edge) is re-evaluated in fp64 from the global tables.
template <typename Real, bool WINDOW>
__device__ __forceinline__ int slot_values(const uint32_t (&k)[4], const int (&j)[4],
                                           const typename CoefOf<Real>::type* coef,
                                           const double* cdf_g, const Real* x_g,
                                           const BatchedParams& P, bool safe_curve,
                                           Real (&v)[4]) {
  using Coef = typename CoefOf<Real>::type;
  Coef cf[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) cf[e] = coef[j[e]];
  int keep = 0;
  if constexpr (sizeof(Real) == 4) {
    if (!WINDOW && safe_curve) {  // every value of this curve is kept (kSafe)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const float t =
            static_cast<float>(static_cast<int32_t>(k[e] - cf[e].thr)) * 0x1.0p-24f + cf[e].r;
        v[e] = __builtin_fmaf(cf[e].m, t, cf[e].x);
      }
      return 0xF;
    }
    const float lo = static_cast<float>(P.win_lo), hi = static_cast<float>(P.win_hi);
    int unsafe = 0;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const float t =
          static_cast<float>(static_cast<int32_t>(k[e] - cf[e].thr)) * 0x1.0p-24f + cf[e].r;
      v[e] = __builtin_fmaf(cf[e].m, t, cf[e].x);
      const float a = __builtin_fabsf(v[e]);
      bool safe = a >= 1e-30f && a <= 1e30f;
      bool in = true;
      if (WINDOW) {
        const float margin = 2e-6f * a;
        safe = safe && __builtin_fabsf(v[e] - lo) > margin && __builtin_fabsf(v[e] - hi) > margin;
        in = v[e] >= lo && v[e] <= hi;
      }
      keep |= (safe && in) ? (1 << e) : 0;
      unsafe |= safe ? 0 : (1 << e);
    }
    if (__builtin_expect(unsafe != 0, 0)) {
#pragma unroll 1
      for (int e = 0; e < 4; ++e) {
        if ((unsafe >> e) & 1) {
          const double v64 = interp_reference(cdf_g, x_g, j[e], k[e]);
          v[e] = static_cast<float>(v64);
          keep |= keep_value<WINDOW>(v64, P) ? (1 << e) : 0;
        }
      }
    }
  } else {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      v[e] = __builtin_fma(cf[e].m, double(k[e]) * 0x1.0p-24 - cf[e].c, cf[e].x);
      keep |= keep_value<WINDOW>(v[e], P) ? (1 << e) : 0;
    }
  }
  return keep;
}

__device__ __forceinline__ void lds_copy(void* dst, const void* src, int64_t bytes) {
  const uint4* s = static_cast<const uint4*>(src);
  uint4* d = static_cast<uint4*>(dst);
  for (int64_t i = threadIdx.x; i < bytes / 16; i += blockDim.x) d[i] = s[i];
}

// Exclusive prefix of per-lane counts (0..4) within a wave, and the wave total.
__device__ __forceinline__ int wave_scan_counts(int cnt, int* total) {
  const uint64_t lt = lanemask_lt();
  const uint64_t b0 = ballot(cnt & 1), b1 = ballot(cnt & 2), b2 = ballot(cnt & 4);
  *total = popc64(b0) + 2 * popc64(b1) + 4 * popc64(b2);
  return popc64(b0 & lt) + 2 * popc64(b1 & lt) + 4 * popc64(b2 & lt);
}

// Keep-mask words of a chunk: wave w's iteration `it` owns words
// [(w * per_wave + it) * 4, +4), one 64-bit ballot per slot position e.
__device__ __forceinline__ int64_t chunk_blocks(const InstanceDesc& in, const ChunkDesc& ch) {
  if (ch.slot_hi <= ch.slot_lo) return 0;
  return int64_t(((in.counter_base + uint64_t(ch.slot_hi) + 3) >> 2) -
                 ((in.counter_base + uint64_t(ch.slot_lo)) >> 2));
}

// ---------------------------------------------------------------------------
// Sampling: one work-group per chunk, one segment per wave.  While every slot
// of a segment so far was kept, rows are positional and no keep mask is
// written; the first iteration with a dropped slot is recorded, and masks are
// written from that iteration on.
template <typename Real, bool WINDOW>
__global__ void __launch_bounds__(1024) sample_kernel(BatchedParams P) {
  using Coef = typename CoefOf<Real>::type;
  extern __shared__ __align__(16) uint8_t lds[];
  const int c = blockIdx.x;
  const ChunkDesc ch = P.chunks[c];
  const InstanceDesc in = P.inst[ch.instance];
  Tables<Real> T = tables_of<Real>(P.tables, in);
  const int n = static_cast<int>(in.n_nodes);
  const int shift = 24 - in.guide_log2;

  lds_copy(lds, T.srt, T.L.cdf - T.L.thr);
  __syncthreads();
  const uint32_t* s_srt = reinterpret_cast<const uint32_t*>(lds);
  const uint32_t* s_guide = reinterpret_cast<const uint32_t*>(lds + (T.L.guide - T.L.thr));
  const Coef* s_coef = reinterpret_cast<const Coef*>(lds + (T.L.coef - T.L.thr));
  const Real* x_g = static_cast<const Real*>(P.x) + in.node_off;

  const int nw = blockDim.x >> 6, w = threadIdx.x >> 6, lane = lane_id();
  const Segment sg = segment_of(in.counter_base, ch.slot_lo, ch.slot_hi, w, nw);
  const int64_t seg = int64_t(c) * nw + w;
  const int64_t base_row = in.out_off + int64_t(sg.p_lo - in.counter_base);
  Real* out = static_cast<Real*>(P.events) + base_row;
  int32_t* bins_out = P.bins ? P.bins + base_row : nullptr;
  const int64_t per_wave = int64_t(blocks_per_wave(uint64_t(chunk_blocks(in, ch)), nw) / 64);
  uint64_t* mask = P.keepmask + ch.mask_off + int64_t(w) * per_wave * 4;
  const bool safe_curve = (T.hdr->flags & kSafe) != 0;
  // 32-bit slot positions relative to counter 4 qa: live slots are [off0, span)
  const uint32_t nblk = uint32_t(sg.qb - sg.qa);
  const uint32_t off0 = uint32_t(sg.p_lo - 4 * sg.qa);
  const uint32_t span = uint32_t(sg.p_hi - 4 * sg.qa);
  int64_t running = 0;
  int first_drop = int(sg.iters);  // iterations before this one kept every slot
  int it = 0;
  for (uint32_t i0 = 0; i0 < nblk; i0 += 64, ++it) {
    const uint32_t i = i0 + uint32_t(lane);
    const Words4 wd = philox4x32_10_uniform(sg.qa + i, in.stream, P.seed);
    const uint32_t k[4] = {wd.w0 >> 8, wd.w1 >> 8, wd.w2 >> 8, wd.w3 >> 8};
    int j[4];
    find_bins4(k, s_srt, s_guide, shift, n, j);
    Real v[4];
    const int keep = slot_values<Real, WINDOW>(k, j, s_coef, T.cdf, x_g, P, safe_curve, v);
    const bool full_iter = (i0 > 0 || off0 == 0) && 4 * (i0 + 64) <= span;
    int live = 0xF;
    if (!full_iter) {
      live = 0;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const uint32_t r = 4 * i + e;
        live |= (i < nblk && r >= off0 && r < span) ? (1 << e) : 0;
      }
    }
    const int keepbits = keep & live;
    int64_t row;
    int total;
    if (first_drop == int(sg.iters) && ballot(keepbits != live) == 0ull) {
      // every slot so far kept: slot 4i + e is row 4i + e - off0
      const int32_t r0 = int32_t(4 * i) - int32_t(off0);
      row = r0 < 0 ? 0 : r0;
      const uint32_t a = 4 * i0 > off0 ? 4 * i0 : off0;
      const uint32_t b = 4 * (i0 + 64) < span ? 4 * (i0 + 64) : span;
      total = b > a ? int(b - a) : 0;
    } else {
      if (first_drop == int(sg.iters)) first_drop = it;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const uint64_t b = ballot((keepbits >> e) & 1);
        if (lane == e) mask[int64_t(it) * 4 + e] = b;
      }
      row = running + wave_scan_counts(__popc(keepbits), &total);
    }
    if (keepbits == 0xF && ((base_row + row) & 3) == 0) {
      if constexpr (sizeof(Real) == 4) {
        *reinterpret_cast<float4*>(out + row) = make_float4(v[0], v[1], v[2], v[3]);
      } else {
        reinterpret_cast<double2*>(out + row)[0] = make_double2(v[0], v[1]);
        reinterpret_cast<double2*>(out + row)[1] = make_double2(v[2], v[3]);
      }
      if (bins_out) *reinterpret_cast<int4*>(bins_out + row) = make_int4(j[0], j[1], j[2], j[3]);
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        if ((keepbits >> e) & 1) {
          out[row] = v[e];
          if (bins_out) bins_out[row] = j[e];
          ++row;
        }
      }
    }
    running += total;
  }
  if (lane == 0) {
    P.seg_kept[seg] = running;
    P.seg_cap[seg] = int64_t(sg.p_hi - sg.p_lo);
    P.seg_drop[seg] = first_drop;
  }
}

// ---------------------------------------------------------------------------
// Finalize: one work-group.  seg_row = exclusive scan of seg_kept; kept per
// instance; total; dense flag (every segment kept all of its slots).
__global__ void __launch_bounds__(1024) finalize_kernel(BatchedParams P) {
  __shared__ double scratch[16];
  __shared__ int sh_sparse;
  if (threadIdx.x == 0) sh_sparse = 0;
  __syncthreads();
  double carry = 0.0;  // counts < 2^53: exact in double
  bool sparse = false;
  for (int64_t base = 0; base < P.n_segments; base += blockDim.x) {
    const int64_t sgi = base + threadIdx.x;
    double v = 0.0;
    if (sgi < P.n_segments) {
      v = double(P.seg_kept[sgi]);
      sparse = sparse || P.seg_kept[sgi] != P.seg_cap[sgi];
    }
    double tot;
    const double ex = block_exclusive_scan(v, scratch, &tot);
    if (sgi < P.n_segments) P.seg_row[sgi] = static_cast<int64_t>(carry + ex);
    carry += tot;
  }
  if (sparse) atomicOr(&sh_sparse, 1);
  __syncthreads();
  for (int64_t i = threadIdx.x; i < P.n_instances; i += blockDim.x) {
    const InstanceDesc in = P.inst[i];
    int64_t k = 0;
    const int64_t s0 = int64_t(in.first_chunk) * P.waves, s1 = s0 + int64_t(in.n_chunks) * P.waves;
    for (int64_t s = s0; s < s1; ++s) k += P.seg_kept[s];
    P.kept[i] = k;
  }
  if (threadIdx.x == 0) {
    P.total[0] = static_cast<int64_t>(carry);
    P.flags[0] = sh_sparse ? 0 : 1;  // dense
  }
}

// Pack: one work-group per chunk, one segment per wave; a no-op when dense.
template <typename Real>
__global__ void __launch_bounds__(1024) pack_kernel(BatchedParams P, Real* packed,
                                                    int32_t* packed_bins) {
  if (P.flags[0]) return;
  const ChunkDesc ch = P.chunks[blockIdx.x];
  const InstanceDesc in = P.inst[ch.instance];
  const int nw = blockDim.x >> 6, w = threadIdx.x >> 6;
  const Segment sg = segment_of(in.counter_base, ch.slot_lo, ch.slot_hi, w, nw);
  const int64_t seg = int64_t(blockIdx.x) * nw + w;
  const int64_t cnt = P.seg_kept[seg];
  const int64_t src = in.out_off + int64_t(sg.p_lo - in.counter_base), dst = P.seg_row[seg];
  const Real* ev = static_cast<const Real*>(P.events);
  for (int64_t r = lane_id(); r < cnt; r += 64) {
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
  double seg = 0.0;
#pragma unroll
  for (int k = MAXPER - 1; k >= 0; --k) {
    seg += gc[k];
    gc[k] = seg;  // sum over k' >= k within the segment
  }
  const double after = block_exclusive_suffix(seg, scratch);
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

// VJP: one work-group per chunk, one segment per wave; the chunk that
// completes an instance runs the tail.
template <typename Real, int MAXPER>
__global__ void __launch_bounds__(1024) vjp_kernel(BatchedParams P, const Real* grad,
                                                   Real* grad_density) {
  extern __shared__ __align__(16) uint8_t lds[];
  __shared__ double scratch[16];
  __shared__ int sh_last;
  const int c = blockIdx.x;
  const ChunkDesc ch = P.chunks[c];
  const InstanceDesc in = P.inst[ch.instance];
  Tables<Real> T = tables_of<Real>(P.tables, in);
  const int n = static_cast<int>(in.n_nodes);
  const int shift = 24 - in.guide_log2;

  // LDS: srt | guide | (r: fp32, C: fp64) | h0 | h1
  const int64_t search_bytes = T.L.coef - T.L.thr;
  const int64_t nb = pad4(n);
  lds_copy(lds, T.srt, search_bytes);
  uint8_t* lds_t = lds + search_bytes;
  if constexpr (sizeof(Real) == 4) lds_copy(lds_t, T.rlo, nb * 4);
  else lds_copy(lds_t, T.cdf, nb * 8);
  double* h0 = reinterpret_cast<double*>(lds_t + nb * sizeof(Real));
  double* h1 = h0 + nb;
  for (int64_t i = threadIdx.x; i < 2 * nb; i += blockDim.x) h0[i] = 0.0;
  __syncthreads();
  const uint32_t* s_srt = reinterpret_cast<const uint32_t*>(lds);
  const uint32_t* s_guide = reinterpret_cast<const uint32_t*>(lds + (T.L.guide - T.L.thr));
  const bool mono = (T.hdr->flags & kMonotone) != 0;

  const int nw = blockDim.x >> 6, w = threadIdx.x >> 6, lane = lane_id();
  const Segment sg = segment_of(in.counter_base, ch.slot_lo, ch.slot_hi, w, nw);
  const int64_t seg = int64_t(c) * nw + w;
  const int first_drop = P.seg_drop[seg];
  const int64_t row0 = P.seg_row[seg];
  const int64_t per_wave = int64_t(blocks_per_wave(uint64_t(chunk_blocks(in, ch)), nw) / 64);
  const uint64_t* mask = P.keepmask + ch.mask_off + int64_t(w) * per_wave * 4;
  const float* s_r = reinterpret_cast<const float*>(lds_t);
  const double* s_c = reinterpret_cast<const double*>(lds_t);
  // kept slots before the first masked iteration: every live slot before it
  const uint32_t nblk = uint32_t(sg.qb - sg.qa);
  const uint32_t off0 = uint32_t(sg.p_lo - 4 * sg.qa);
  const uint32_t span = uint32_t(sg.p_hi - 4 * sg.qa);
  // kept slots before the first masked iteration: every live slot before it
  int64_t running = int64_t(4 * 64 * int64_t(first_drop)) - int64_t(off0);
  running = running < 0 ? 0 : running;
  int it = 0;
  for (uint32_t i0 = 0; i0 < nblk; i0 += 64, ++it) {
    const uint32_t i = i0 + uint32_t(lane);
    const Words4 wd = philox4x32_10_uniform(sg.qa + i, in.stream, P.seed);
    const uint32_t k[4] = {wd.w0 >> 8, wd.w1 >> 8, wd.w2 >> 8, wd.w3 >> 8};
    const bool full_iter = (i0 > 0 || off0 == 0) && 4 * (i0 + 64) <= span;
    int keepbits = 0xF;
    int64_t row;
    if (it < first_drop) {  // every live slot of these iterations was kept
      if (!full_iter) {
        keepbits = 0;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const uint32_t r = 4 * i + e;
          keepbits |= (i < nblk && r >= off0 && r < span) ? (1 << e) : 0;
        }
      }
      const int32_t r0 = int32_t(4 * i) - int32_t(off0);
      row = row0 + (r0 < 0 ? 0 : r0);
    } else {
      const uint64_t* mw = mask + int64_t(it) * 4;
      keepbits = 0;
#pragma unroll
      for (int e = 0; e < 4; ++e) keepbits |= int((mw[e] >> lane) & 1ull) << e;
      int total;
      row = row0 + running + wave_scan_counts(__popc(keepbits), &total);
      running += total;
    }
    int j[4];
    find_bins4(k, s_srt, s_guide, shift, n, j);
    double t[4], g[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      if constexpr (sizeof(Real) == 4) {
        const uint32_t tj = mono ? s_srt[j[e]] : T.coef[j[e]].thr;
        t[e] = double(static_cast<float>(static_cast<int32_t>(k[e] - tj)) * 0x1.0p-24f + s_r[j[e]]);
      } else {
        t[e] = double(k[e]) * 0x1.0p-24 - s_c[j[e]];
      }
    }
    if (keepbits == 0xF) {
      if constexpr (sizeof(Real) == 4) {
        if ((row & 3) == 0) {
          const float4 g4 = *reinterpret_cast<const float4*>(grad + row);
          g[0] = g4.x; g[1] = g4.y; g[2] = g4.z; g[3] = g4.w;
        } else {
#pragma unroll
          for (int e = 0; e < 4; ++e) g[e] = double(grad[row + e]);
        }
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) g[e] = double(grad[row + e]);
      }
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        atomicAdd(&h0[j[e]], g[e]);
        atomicAdd(&h1[j[e]], g[e] * t[e]);
      }
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        if (!((keepbits >> e) & 1)) continue;
        const double ge = double(grad[row++]);
        atomicAdd(&h0[j[e]], ge);
        atomicAdd(&h1[j[e]], ge * t[e]);
      }
    }
  }
  __syncthreads();

  if (in.n_chunks == 1) {
    vjp_tail<Real, MAXPER>(P, in, T, h0, h1, grad_density, scratch);
    return;
  }
  // multi-chunk instance: publish partial sums; the last chunk reduces them in order
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
// Host-side launchers (called from capi.hip).
template <typename Real>
hipError_t launch_tables(const BatchedParams& P, int threads, bool sequential, int64_t lds,
                         hipStream_t stream) {
  auto kern = sequential ? tables_kernel<Real, true> : tables_kernel<Real, false>;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kern, dim3(P.n_instances), dim3(threads), size_t(lds), stream, P);
  return hipGetLastError();
}

template <typename Real>
hipError_t launch_sample(const BatchedParams& P, int threads, int64_t lds, hipStream_t stream) {
  auto kern = P.use_window ? sample_kernel<Real, true> : sample_kernel<Real, false>;
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
hipError_t launch_pack(const BatchedParams& P, int threads, void* packed, int32_t* packed_bins,
                       hipStream_t stream) {
  hipLaunchKernelGGL((pack_kernel<Real>), dim3(P.n_chunks), dim3(threads), 0, stream, P,
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

template hipError_t launch_tables<float>(const BatchedParams&, int, bool, int64_t, hipStream_t);
template hipError_t launch_tables<double>(const BatchedParams&, int, bool, int64_t, hipStream_t);
template hipError_t launch_sample<float>(const BatchedParams&, int, int64_t, hipStream_t);
template hipError_t launch_sample<double>(const BatchedParams&, int, int64_t, hipStream_t);
template hipError_t launch_pack<float>(const BatchedParams&, int, void*, int32_t*, hipStream_t);
template hipError_t launch_pack<double>(const BatchedParams&, int, void*, int32_t*, hipStream_t);
template hipError_t launch_vjp<float>(const BatchedParams&, int, int64_t, int, const void*, void*,
                                      hipStream_t);
template hipError_t launch_vjp<double>(const BatchedParams&, int, int64_t, int, const void*, void*,
                                       hipStream_t);

}  // namespace qloits
