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

#include <algorithm>
#include <cstddef>
#include <cmath>
#include <cstdint>

#include "batched.hpp"
#include "fast_tables.hpp"
#include "loits_device.hpp"

#ifndef QLOITS_TRIP
#define QLOITS_TRIP 2
#endif

namespace qloits {

// Philox blocks per lane per trip of the sampling and VJP loops.
constexpr int kTrip = QLOITS_TRIP;

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
template <typename T>
__device__ T block_exclusive_scan(T v, T* scratch, T* total) {
  const int wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  const T inc = wave_inclusive_scan(v);
  __syncthreads();
  if (lane_id() == kWave - 1) scratch[wave] = inc;
  __syncthreads();
  T before = T(0), all = T(0);
  for (int w = 0; w < nwaves; ++w) {
    const T s = scratch[w];
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
  return t;
}

// 16-byte copies by the whole work-group (LDS <-> global), eight loads per
// thread in flight before their stores (an index past the end reloads the
// last unit and stores nothing).
__device__ __forceinline__ void lds_copy(void* dst, const void* src, int64_t bytes) {
  const uint4* s = static_cast<const uint4*>(src);
  uint4* d = static_cast<uint4*>(dst);
  const int64_t units = bytes / 16, nt = blockDim.x;
  for (int64_t i = threadIdx.x; i < units; i += 8 * nt) {
    uint4 v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int64_t u = i + k * nt;
      v[k] = s[u < units ? u : units - 1];
    }
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (i + k * nt < units) d[i + k * nt] = v[k];
  }
}

// Bytes of a built guide in use: the G cell words and the level-2 words.
template <typename Real>
__device__ __forceinline__ int64_t guide_bytes(const Tables<Real>& T, const InstanceDesc& in) {
  return pad4((int64_t(1) << in.guide_log2) + int64_t(T.hdr->level2_words)) * 4;
}

// ---------------------------------------------------------------------------
// Table build: one work-group builds one instance's tables in LDS (layout
// lds_build_layout): trapezoid norm and CDF (parallel blocked scan or the
// reference's sequential chain), integer thresholds (sorted when the CDF is
// not monotone), the two-level guide and per-bin interpolation data
// (cpp/loits_core.cpp:55-165, 190-195).  A `writer` group also publishes them
// to the global blob for the VJP and the optional norm / CDF outputs.
struct BuildShared {
  double scratch[16];
  double norm, inv;
  uint32_t wtot[64];  // per-wave totals of the cell-table scans
  uint32_t bits;  // AND over nodes: 1 monotone, 2 finite, 4 x > 0, 8 x < 0
  uint32_t flags;
  int assigned, used;
};

template <typename Real>
struct LdsTables {
  uint32_t* srt;
  uint32_t* guide;
  typename CoefOf<Real>::type* coef;
  double* cdf;
  uint2* crowd;
};

template <typename Real>
__device__ __forceinline__ LdsTables<Real> lds_tables_of(uint8_t* lds, const InstanceDesc& in) {
  using Coef = typename CoefOf<Real>::type;
  const LdsLayout l = lds_build_layout(in.n_nodes, in.guide_log2, sizeof(Coef));
  LdsTables<Real> t;
  t.srt = reinterpret_cast<uint32_t*>(lds + l.srt);
  t.guide = reinterpret_cast<uint32_t*>(lds + l.guide);
  t.coef = reinterpret_cast<Coef*>(lds + l.coef);
  t.cdf = reinterpret_cast<double*>(lds + l.cdf);
  t.crowd = reinterpret_cast<uint2*>(lds + l.crowd);
  return t;
}

// Trapezoid norm and CDF of one curve into cdf[0..n) (LDS or global):
// the reference's sequential chain (cpp/loits_core.cpp:73-80 then :130-135) or
// a blocked parallel scan; sh.norm and sh.inv receive the norm and 1 / norm.
// Ends with a work-group barrier.
template <typename Real, bool SEQUENTIAL>
__device__ void build_norm_cdf(const Real* x, const Real* s, int64_t n, double acc, double* cdf,
                               BuildShared& sh) {
  const int tid = threadIdx.x, nt = blockDim.x;
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
      cdf[0] = 0.0;
      for (int64_t t = 1; t < n; ++t) {
        const double rho = double(s[t]) * inv;
        cum = __builtin_fma(0.5 * (double(x[t]) - double(x[t - 1])), rho_prev + rho, cum);
        cdf[t] = cum * acc;
        rho_prev = rho;
      }
      sh.norm = nrm;
      sh.inv = inv;
    }
  } else {
    // Blocked parallel scan: each thread owns a contiguous segment; segments
    // of up to four nodes keep their nodes in registers.
    const int64_t per = (n + nt - 1) / nt;
    const int64_t t0 = tid * per, t1 = t0 + per < n ? t0 + per : n;
    double nrm, inv, run;
    if (per <= 4) {
      double xv[5], sv[5];  // nodes t0 - 1 .. t0 + 3
#pragma unroll
      for (int q = 0; q < 5; ++q) {
        const int64_t t = t0 - 1 + q;
        const bool ok = t >= 0 && t < t1;
        xv[q] = ok ? double(x[t]) : 0.0;
        sv[q] = ok ? double(s[t]) : 0.0;
      }
      double local = 0.0;
#pragma unroll
      for (int q = 1; q < 5; ++q) {
        const int64_t t = t0 - 1 + q;
        if (t >= 1 && t < t1) local += 0.5 * (xv[q] - xv[q - 1]) * (sv[q - 1] + sv[q]);
      }
      nrm = block_sum(local, sh.scratch);
      inv = 1.0 / nrm;
      double terms[4], part = 0.0;
#pragma unroll
      for (int q = 1; q < 5; ++q) {
        const int64_t t = t0 - 1 + q;
        terms[q - 1] = 0.5 * (xv[q] - xv[q - 1]) * (sv[q - 1] * inv + sv[q] * inv);
        if (t >= 1 && t < t1) part += terms[q - 1];
      }
      double total;
      run = block_exclusive_scan(part, sh.scratch, &total);
#pragma unroll
      for (int q = 1; q < 5; ++q) {
        const int64_t t = t0 - 1 + q;
        if (t >= t1) continue;
        if (t == 0) {
          cdf[0] = 0.0;
          continue;
        }
        run += terms[q - 1];
        cdf[t] = run * acc;
      }
    } else {
      double local = 0.0;
      for (int64_t t = t0 > 1 ? t0 : 1; t < t1; ++t)
        local += 0.5 * (double(x[t]) - double(x[t - 1])) * (double(s[t - 1]) + double(s[t]));
      nrm = block_sum(local, sh.scratch);
      inv = 1.0 / nrm;
      double part = 0.0;
      for (int64_t t = t0 > 1 ? t0 : 1; t < t1; ++t)
        part += 0.5 * (double(x[t]) - double(x[t - 1])) *
                (double(s[t - 1]) * inv + double(s[t]) * inv);
      double total;
      run = block_exclusive_scan(part, sh.scratch, &total);
      for (int64_t t = t0; t < t1; ++t) {
        if (t == 0) {
          cdf[0] = 0.0;
          continue;
        }
        run += 0.5 * (double(x[t]) - double(x[t - 1])) *
               (double(s[t - 1]) * inv + double(s[t]) * inv);
        cdf[t] = run * acc;
      }
    }
    if (tid == 0) {
      sh.norm = nrm;
      sh.inv = inv;
    }
  }
  __syncthreads();
}

// GLOBAL: srt, coef and cdf are built in place in the global blob (instances
// whose LDS image does not fit); LDS then holds the crowded-cell list and the
// cell counts only.
template <typename Real, bool SEQUENTIAL, bool GLOBAL = false>
__device__ void build_tables(const BatchedParams& P, int inst, const InstanceDesc& in,
                             uint8_t* lds, bool writer, BuildShared& sh) {
  using Coef = typename CoefOf<Real>::type;
  const int64_t n = in.n_nodes;
  const Real* x = static_cast<const Real*>(P.x) + in.node_off;
  const Real* s = static_cast<const Real*>(P.density) + in.node_off;
  LdsTables<Real> L;
  uint32_t* hist;  // per-cell threshold counts (step 4)
  if constexpr (GLOBAL) {
    const Tables<Real> T = tables_of<Real>(P.tables, in);
    L.srt = T.srt;
    L.guide = T.guide;
    L.coef = T.coef;
    L.cdf = T.cdf;
    L.crowd = reinterpret_cast<uint2*>(lds);
    hist = reinterpret_cast<uint32_t*>(lds + ((int64_t(1) << in.guide_log2) / 4) * 8);
  } else {
    L = lds_tables_of<Real>(lds, in);
    hist = L.guide + (int64_t(1) << in.guide_log2);  // the level-2 region, free until step 6
  }
  const double acc = in.accept ? 1.0 : 0.0;
  const int tid = threadIdx.x, nt = blockDim.x;
  if (tid == 0) {
    sh.bits = 0xFu;
    sh.assigned = 0;
    sh.used = 0;
  }

  // 0. global tables: nodes and densities staged in LDS when they fit (the
  //    crowded-cell list and the counts take that LDS from step 4 on)
  const Real* xs = x;
  const Real* ds = s;
  if constexpr (GLOBAL) {
    if (global_build_stage_bytes(n, sizeof(Real)) > 0) {
      Real* sx = reinterpret_cast<Real*>(lds);
      Real* sd = sx + pad4(n);
      for (int64_t t0 = tid; t0 < n; t0 += 4 * nt) {  // four nodes per thread in flight
        Real a[4], b[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int64_t t = t0 + int64_t(u) * nt;
          a[u] = t < n ? x[t] : Real(0);
          b[u] = t < n ? s[t] : Real(0);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int64_t t = t0 + int64_t(u) * nt;
          if (t < n) {
            sx[t] = a[u];
            sd[t] = b[u];
          }
        }
      }
      __syncthreads();
      xs = sx;
      ds = sd;
    }
  }

  // 1. norm and CDF
  build_norm_cdf<Real, SEQUENTIAL>(xs, ds, n, acc, L.cdf, sh);

  // 2. thresholds, per-bin interpolation data, curve properties
  const int64_t pw = pow2_ceil(n);
  uint32_t bits = 0xFu;
  for (int64_t t = tid; t < pad4(pw); t += nt) {
    if (t >= n) {
      L.srt[t] = kNever;
      continue;
    }
    const double c = L.cdf[t];
    const uint32_t thr = threshold_of(c);
    L.srt[t] = thr;
    if (t > 0 && thr < threshold_of(L.cdf[t - 1])) bits &= ~1u;
    const double xt = double(xs[t]);
    const double ax = __builtin_fabs(xt);
    if (!(__builtin_isfinite(c) && ax >= 1e-30 && ax <= 1e29)) bits &= ~2u;
    if (!(xt > 0.0)) bits &= ~4u;
    if (!(xt < 0.0)) bits &= ~8u;
    if (t + 1 < n) {
      const double c1 = L.cdf[t + 1];
      const double inv_d = 1.0 / (c1 - c + kEpsilon);              // cpp/loits_core.cpp:193
      const double m = (double(xs[t + 1]) - xt) * inv_d;            // :194
      if (!__builtin_isfinite(m)) bits &= ~2u;
      if constexpr (sizeof(Real) == 4) {
        const double r = double(thr) * 0x1.0p-24 - c;
        const float M = static_cast<float>(m * 0x1.0p-24);
        const float A = static_cast<float>(__builtin_fma(m, r, xt));
        if constexpr (GLOBAL) {  // one 16-byte record per bin (CoefF32)
          reinterpret_cast<float4*>(L.coef)[t] =
              make_float4(M, A, __uint_as_float(thr), static_cast<float>(r));
        } else {
          const CoefF32View cv = coef_f32_view(L.coef, pad4(n));
          cv.ma[t] = make_float2(M, A);
          cv.thr[t] = thr;
          cv.r[t] = static_cast<float>(r);
        }
      } else {
        L.coef[t] = CoefF64{c, m, xt, thr, 0u};
      }
    }
  }
  {
    uint32_t wb = 0u;
#pragma unroll
    for (int q = 0; q < 4; ++q) wb |= ballot(!((bits >> q) & 1u)) == 0ull ? (1u << q) : 0u;
    if (lane_id() == 0 && wb != 0xFu) atomicAnd(&sh.bits, wb);
  }
  __syncthreads();
  const uint32_t all = sh.bits;
  const bool mono = (all & 1u) != 0;
  const bool safe = (all & 2u) != 0 && (all & 12u) != 0;

  // 3. sorted thresholds (count(k) = #{t : T_t <= k}); bins keep their own T_j
  if (!mono) {
    for (int64_t k = 2; k <= pw; k <<= 1) {
      for (int64_t j = k >> 1; j > 0; j >>= 1) {
        for (int64_t i = tid; i < pw; i += nt) {
          const int64_t ixj = i ^ j;
          if (ixj > i) {
            const uint32_t a = L.srt[i], b = L.srt[ixj];
            const bool up = (i & k) == 0;
            if ((a > b) == up) {
              L.srt[i] = b;
              L.srt[ixj] = a;
            }
          }
        }
        __syncthreads();
      }
    }
  }

  // 4. per-cell threshold counts h[g] = #{t : srt_t >> shift == g} (in the
  //    level-2 region, unused until step 6); thresholds >= 2^24 fall outside
  //    every cell.  Each thread owns `per` consecutive cells; an exclusive scan
  //    of the counts gives c0 = #{t : srt_t < g << shift} for its first cell.
  const int64_t G = int64_t(1) << in.guide_log2;
  const int shift = 24 - in.guide_log2;
  for (int64_t g = tid; g < G; g += nt) hist[g] = 0u;
  __syncthreads();
  for (int64_t t = tid; t < n; t += nt) {
    const uint32_t cell = L.srt[t] >> shift;
    if (cell < uint32_t(G)) atomicAdd(&hist[cell], 1u);
  }
  __syncthreads();
  const int per = int(G / nt);  // G >= 4096 and nt <= 1024 are powers of two
  const int64_t g0 = int64_t(tid) * per;
  const int nn = int(n);
  // Cells holding two or more thresholds are "crowded": they get 2^f level-2
  // sub-cells each, in cell order, while the G level-2 words last (any cell
  // may take the level-2 or search form; direct words are an optimisation).
  // One scan of (crowded cells | thresholds << 16 | level-2 words << 32).
  uint64_t local = 0u;
  for (int i = 0; i < per; i += 4) {
    const uint4 v = *reinterpret_cast<const uint4*>(hist + g0 + i);
    const uint32_t hv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      local += uint64_t(hv[q]) << 16;
      if (hv[q] >= 2u)
        local += 1u + (uint64_t(1u << guide_fanout_log2(int(hv[q]), shift)) << 32);
    }
  }
  uint64_t tot64;
  const uint64_t ex =
      block_exclusive_scan(local, reinterpret_cast<uint64_t*>(sh.scratch), &tot64);
  const uint32_t run0 = uint32_t(ex >> 16) & 0xFFFFu;
  uint32_t rank = uint32_t(ex) & 0xFFFFu, woff = uint32_t(ex >> 32);
  // Redirected cells are the crowded cells before the first one that would
  // overrun level 2; that cell's thread (or thread 0 when all fit) records
  // their number and the level-2 words they use.
  if (tid == 0 && (tot64 >> 32) <= uint64_t(G)) {
    sh.assigned = int(uint32_t(tot64) & 0xFFFFu);
    sh.used = int(tot64 >> 32);
  }
  auto jof = [&](uint32_t c) -> int {
    const int j = int(c) - 1;
    return j < 0 ? 0 : (j > nn - 2 ? nn - 2 : j);
  };
  uint32_t run;
  run = run0;
  for (int i = 0; i < per; i += 4) {
    const uint4 v = *reinterpret_cast<const uint4*>(hist + g0 + i);
    const uint32_t hv[4] = {v.x, v.y, v.z, v.w};
    uint32_t wv[4];
    uint32_t c = run;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      // guide_encode for a cell with at most one threshold
      const int j0 = jof(c), j1 = jof(c + hv[q]);
      const uint32_t t = L.srt[int(c) < nn ? int(c) : nn - 1];
      const uint32_t key_lo = uint32_t(g0 + i + q) << shift;
      wv[q] = uint32_t(j0) | ((j1 != j0 ? t - key_lo : kNoPoint) << 15);
      c += hv[q];
    }
    if ((v.x | v.y | v.z | v.w) > 1u) {  // some cell holds two or more
      c = run;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const uint32_t c0 = c, h = hv[q], g = uint32_t(g0 + i + q);
        if (h >= 2u) {
          const int fl = guide_fanout_log2(int(h), shift);
          if (woff + (1u << fl) <= uint32_t(G)) {
            L.crowd[rank] = make_uint2(c0 | ((c0 + h) << 16), woff | (g << 15));
            wv[q] = kSpecial | (uint32_t(fl) << 24) | uint32_t(G + woff);
          } else {
            if (woff <= uint32_t(G)) {  // the first cell that does not fit
              sh.assigned = int(rank);
              sh.used = int(woff);
            }
            wv[q] = kSearch | c0 | (h << 15);
          }
          woff += 1u << fl;
          ++rank;
        }
        c += h;
      }
    }
    *reinterpret_cast<uint4*>(L.guide + g0 + i) = make_uint4(wv[0], wv[1], wv[2], wv[3]);
    run = c;
  }
  __syncthreads();

  // 6. level-2 words (the counts are dead).  Each thread fills a contiguous
  //    run of words, walking the redirected cells in order; a sub-cell's
  //    upper bound is the next sub-cell's lower bound.  With global tables
  //    the sorted thresholds are first staged into the counts' LDS (n <= G
  //    words), so that the bisections read LDS instead of global memory.
  const uint32_t* srt6 = L.srt;
  if constexpr (GLOBAL) {
    if (n <= G) {
      lds_copy(hist, L.srt, pad4(n) * 4);
      __syncthreads();
      srt6 = hist;
    }
  }
  const int used = sh.used, n_assigned = sh.assigned;
  const int per_item = (used + nt - 1) / nt;
  int item = tid * per_item;
  const int item_end = item + per_item < used ? item + per_item : used;
  if (item < item_end) {
    int r = 0, hi_r = n_assigned - 1;  // last redirected cell with first word <= item
    while (r < hi_r) {
      const int mid = (r + hi_r + 1) >> 1;
      if (int(L.crowd[mid].y & 0x7FFFu) <= item) r = mid; else hi_r = mid - 1;
    }
    uint2 cr = L.crowd[r];
    int lo0 = int(cr.x & 0xFFFFu), hi0 = int(cr.x >> 16);
    int ss = shift - guide_fanout_log2(hi0 - lo0, shift);
    int cell_end = int(cr.y & 0x7FFFu) + (1 << (shift - ss));
    uint32_t key = ((cr.y >> 15) << shift) + (uint32_t(item - int(cr.y & 0x7FFFu)) << ss);
    int a = lo0, b = hi0;  // first position with srt >= key
    while (a < b) {
      const int mid = (a + b) >> 1;
      if (srt6[mid] < key) a = mid + 1; else b = mid;
    }
    for (; item < item_end; ++item) {
      if (item == cell_end) {
        cr = L.crowd[++r];
        lo0 = int(cr.x & 0xFFFFu);
        hi0 = int(cr.x >> 16);
        ss = shift - guide_fanout_log2(hi0 - lo0, shift);
        cell_end = item + (1 << (shift - ss));
        key = (cr.y >> 15) << shift;
        a = lo0;
      }
      const uint32_t key_hi = key + (1u << ss);
      int c2 = a, d2 = hi0;  // first position with srt >= key_hi
      while (c2 < d2) {
        const int mid = (c2 + d2) >> 1;
        if (srt6[mid] < key_hi) c2 = mid + 1; else d2 = mid;
      }
      L.guide[G + item] = guide_encode(a, c2, srt6, key, nn);
      a = c2;
      key = key_hi;
    }
  }
  if (tid == 0) {
    const bool last_full = L.srt[n - 1] >= 16777214u;
    sh.flags = (mono ? kMonotone : 0u) | ((mono && safe && acc == 1.0 && last_full) ? kSafe : 0u);
  }
  __syncthreads();

  // 7. publish: the guide words in use; the per-bin coefficients only where
  //    the VJP reads them (the fixed-point VJP of a monotone float32 curve
  //    derives its per-bin data from srt and cdf)
  if (writer) {
    Tables<Real> T = tables_of<Real>(P.tables, in);
    if constexpr (!GLOBAL) {
      lds_copy(T.srt, L.srt, pad4(n) * 4);
      lds_copy(T.guide, L.guide, pad4(G + used) * 4);
      if (!(sizeof(Real) == 4 && mono)) lds_copy(T.coef, L.coef, pad4(n) * int64_t(sizeof(Coef)));
      lds_copy(T.cdf, L.cdf, pad4(n) * 8);
    }
    if (P.cdf_out)
      for (int64_t t = tid; t < n; t += nt) P.cdf_out[in.node_off + t] = L.cdf[t];
    if (tid == 0) {
      T.hdr->norm = sh.norm;
      T.hdr->inv_n = sh.inv;
      T.hdr->flags = sh.flags;
      T.hdr->n = static_cast<uint32_t>(n);
      T.hdr->guide_log2 = static_cast<uint32_t>(in.guide_log2);
      T.hdr->level2_words = static_cast<uint32_t>(used);
      if (P.norm_out) P.norm_out[inst] = sh.norm;
      P.done[inst] = 0u;
    }
  }
}

// Ascending bitonic sort of a[0, pw) in LDS by the work-group (pw a power of two).
__device__ void bitonic_sort_u32(uint32_t* a, int64_t pw) {
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int64_t k = 2; k <= pw; k <<= 1) {
    for (int64_t j = k >> 1; j > 0; j >>= 1) {
      for (int64_t i = tid; i < pw; i += nt) {
        const int64_t ixj = i ^ j;
        if (ixj > i) {
          const uint32_t x = a[i], y = a[ixj];
          const bool up = (i & k) == 0;
          if ((x > y) == up) {
            a[i] = y;
            a[ixj] = x;
          }
        }
      }
      __syncthreads();
    }
  }
}

// Cell tables of a float32 curve (fast_tables.hpp) built by one work-group in
// LDS (layout fast_lds_forward): norm and CDF, thresholds and per-bin
// interpolation data, the sorted splits, the level-1 cell words and the level-2
// words of crowded cells.  A `writer` work-group publishes the CDF, the
// thresholds and the cell words (fast_blob_layout) for the VJP.
// cdf_copy (global, n doubles, private to the work-group) receives the CDF for
// the float64 re-evaluation of values near a keep boundary.
// Keep class of a bin of a monotone curve under a window [lo, hi], from its
// end nodes a, b: every fp64 value of the bin lies in [min(a, b), max(a, b)]
// (x_j + m (u - C_j) with 0 <= (u - C_j) m / (x_{j+1} - x_j) < 1), so
//   1: all kept (one sign, inside the window by a relative 4e-6),
//   2: none kept (outside it by as much),
//   0: decided per value.
__device__ __forceinline__ uint32_t window_class(double a, double b, double lo, double hi) {
  const double vlo = a < b ? a : b, vhi = a < b ? b : a;
  if (!(__builtin_isfinite(vlo) && __builtin_isfinite(vhi))) return 0u;
  constexpr double e = 4e-6;
  const double alo = __builtin_fabs(lo), ahi = __builtin_fabs(hi);
  const double avlo = __builtin_fabs(vlo), avhi = __builtin_fabs(vhi);
  if (vhi < lo - e * (alo > avhi ? alo : avhi) || vlo > hi + e * (ahi > avlo ? ahi : avlo))
    return 2u;
  if ((vlo > 0.0 || vhi < 0.0) && vlo > lo + e * (alo > avlo ? alo : avlo) &&
      vhi < hi - e * (ahi > avhi ? ahi : avhi))
    return 1u;
  return 0u;
}

template <bool SEQUENTIAL>
__device__ void build_fast(const BatchedParams& P, int inst, const InstanceDesc& in, uint8_t* lds,
                           bool writer, BuildShared& sh, double* cdf_copy) {
  const int64_t n = in.n_nodes;
  const int g = in.fast_log2, s = 24 - g;
  const uint32_t S = 1u << s;
  const int64_t G = int64_t(1) << g;
  const int l2cap = fast_l2_cap(n);
  const FastLdsForward L = fast_lds_forward(n, g);
  uint32_t* cells = reinterpret_cast<uint32_t*>(lds + L.cells);
  double* cdf = reinterpret_cast<double*>(lds + L.cells);  // until step 4
  uint2* list = reinterpret_cast<uint2*>(lds + L.list);
  float4* bins = reinterpret_cast<float4*>(lds + L.bins);
  uint32_t* srt = reinterpret_cast<uint32_t*>(lds + L.srt);
  const float* x = static_cast<const float*>(P.x) + in.node_off;
  const float* d = static_cast<const float*>(P.density) + in.node_off;
  const double acc = in.accept ? 1.0 : 0.0;
  const int tid = threadIdx.x, nt = blockDim.x;
  if (tid == 0) {
    sh.bits = 0xFu;
    sh.assigned = 0;
    sh.used = 0;
  }

  // 1. norm and CDF
  build_norm_cdf<float, SEQUENTIAL>(x, d, n, acc, cdf, sh);

  // 2. thresholds, per-bin {M, A, T}, curve properties (as build_tables)
  const int64_t pw = pow2_ceil(n);
  uint32_t bits = 0xFu;
  for (int64_t t = tid; t < pad4(pw); t += nt) {
    if (t >= n) {
      srt[t] = kNever;
      if (t < pad4(n)) bins[t] = make_float4(0.f, 0.f, 0.f, 0.f);
      continue;
    }
    const double c = cdf[t];
    const uint32_t thr = threshold_of(c);
    srt[t] = thr;
    if (t > 0 && thr < threshold_of(cdf[t - 1])) bits &= ~1u;
    const double xt = double(x[t]);
    const double ax = __builtin_fabs(xt);
    if (!(__builtin_isfinite(c) && ax >= 1e-30 && ax <= 1e29)) bits &= ~2u;
    if (!(xt > 0.0)) bits &= ~4u;
    if (!(xt < 0.0)) bits &= ~8u;
    float4 b = make_float4(0.f, 0.f, __uint_as_float(thr), 0.f);
    if (t + 1 < n) {
      const double inv_d = 1.0 / (cdf[t + 1] - c + kEpsilon);  // cpp/loits_core.cpp:193
      const double m = (double(x[t + 1]) - xt) * inv_d;        // :194
      if (!__builtin_isfinite(m)) bits &= ~2u;
      const double r = double(thr) * 0x1.0p-24 - c;
      b.x = static_cast<float>(m * 0x1.0p-24);
      b.y = static_cast<float>(__builtin_fma(m, r, xt));
      // keep class for a window (not the last bin: it also takes keys past C_{n-1})
      if (P.use_window && t + 2 < n)
        b.w = __uint_as_float(window_class(xt, double(x[t + 1]), P.win_lo, P.win_hi));
    }
    bins[t] = b;
  }
  {
    uint32_t wb = 0u;
#pragma unroll
    for (int q = 0; q < 4; ++q) wb |= ballot(!((bits >> q) & 1u)) == 0ull ? (1u << q) : 0u;
    if (lane_id() == 0 && wb != 0xFu) atomicAnd(&sh.bits, wb);
  }
  __syncthreads();
  const uint32_t all_bits = sh.bits;
  const bool mono = (all_bits & 1u) != 0;
  const bool safe = (all_bits & 2u) != 0 && (all_bits & 12u) != 0;

  // 3. publish the CDF and each bin's own threshold (before step 4 reuses the
  //    CDF's LDS), then sort the thresholds when the CDF is not monotone
  const FastBlobLayout B = fast_blob_layout(n, g);
  uint8_t* blob = P.tables + in.table_off;
  if (cdf_copy)
    for (int64_t t = tid; t < n; t += nt) cdf_copy[t] = cdf[t];
  if (writer) {
    double* gcdf = reinterpret_cast<double*>(blob + B.cdf);
    uint32_t* gown = reinterpret_cast<uint32_t*>(blob + B.own);
    for (int64_t t = tid; t < n; t += nt) {
      gcdf[t] = cdf[t];
      gown[t] = srt[t];
      if (P.cdf_out) P.cdf_out[in.node_off + t] = cdf[t];
    }
  }
  __syncthreads();
  if (!mono) bitonic_sort_u32(srt, pw);

  // 4. per-cell counts of the splits S_1 .. S_{n-2}: all (bits 0-15) and those
  //    at the cell's first key (bits 16-31); splits >= 2^24 lie in no cell.
  //    The level-2 words and the crowded-cell list start cleared.
  {
    uint4* c4 = reinterpret_cast<uint4*>(cells);
    for (int64_t i = tid; i < (G + l2cap) / 4; i += nt) c4[i] = make_uint4(0u, 0u, 0u, 0u);
    for (int i = tid; i < l2cap / 4; i += nt) list[i] = make_uint2(kNoCell, 0u);
  }
  __syncthreads();
  for (int64_t t = 1 + tid; t <= n - 2; t += nt) {
    const uint32_t key = srt[t];
    if (key < (1u << 24)) atomicAdd(&cells[key >> s], 1u + ((key & (S - 1u)) == 0u ? 0x10000u : 0u));
  }
  __syncthreads();

  // 5. level-1 words: a work-group exclusive scan, in cell order, of
  //    splits | level-2 words asked for << 16 (each total stays below 2^16: at
  //    most n - 2 splits and 4 (n - 2) words).  Pass r of thread t holds the
  //    four cells 4 (r nt + t) .. + 3 (16-byte LDS accesses, conflict-free);
  //    up to four passes are scanned together.  A crowded cell's list entry
  //    sits at its first level-2 word / 4 (every crowded cell asks for at least
  //    four words).
  const int nwaves = nt >> 6, lane = lane_id();
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  {
    const int passes = int(G / (4 * int64_t(nt)));
    uint4* c4 = reinterpret_cast<uint4*>(cells);
    uint32_t carry = 0u;
    uint32_t used_hi = 0u;  // end of this thread's last level-2 words
    for (int r0 = 0; r0 < passes; r0 += 4) {
      const int np = passes - r0 < 4 ? passes - r0 : 4;
      uint32_t v[4][4], a[4][4], gs[4], inc[4];
      bool crowd = false;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const uint4 q4 = r < np ? c4[(r0 + r) * nt + tid] : make_uint4(0u, 0u, 0u, 0u);
        v[r][0] = q4.x;
        v[r][1] = q4.y;
        v[r][2] = q4.z;
        v[r][3] = q4.w;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const uint32_t all = v[r][q] & 0xFFFFu, inside = all - (v[r][q] >> 16);
          a[r][q] = all;
          crowd = crowd || inside >= 2u;
        }
      }
      if (ballot(crowd) != 0ull) {  // some cell of this wave holds two or more inside splits
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const uint32_t inside = a[r][q] - (v[r][q] >> 16);
            if (inside >= 2u) a[r][q] += (1u << guide_fanout_log2(int(inside), s)) << 16;
          }
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        gs[r] = a[r][0] + a[r][1] + a[r][2] + a[r][3];
        inc[r] = wave_inclusive_scan(gs[r]);
      }
      if (lane == kWave - 1)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (r < np) sh.wtot[r * nwaves + wave] = inc[r];
      __syncthreads();
      // prefix over (pass, wave) in order, redone by every wave
      const uint32_t tv = lane < np * nwaves ? sh.wtot[lane] : 0u;
      const uint32_t tinc = wave_inclusive_scan(tv);
      const uint32_t tex = tinc - tv;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        if (r >= np) continue;
        uint32_t ex = carry + wave_lane(tex, r * nwaves + wave) + inc[r] - gs[r];
        uint32_t wv[4], exq[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          exq[q] = ex;
          ex += a[r][q];
          const uint32_t all = v[r][q] & 0xFFFFu, zero = v[r][q] >> 16;
          const uint32_t lo = exq[q] & 0xFFFFu;
          const uint32_t jl = lo + zero;  // splits <= L
          // no or one inside split: F = lo + all is jl or jl + 1
          const uint32_t nxt = srt[jl + 1];
          wv[q] = all - zero == 1u ? ((nxt << (32 - s)) | (lo + all)) : jl;
        }
        if (crowd) {
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const uint32_t all = v[r][q] & 0xFFFFu, zero = v[r][q] >> 16, inside = all - zero;
            if (inside < 2u) continue;
            const uint32_t jl = (exq[q] & 0xFFFFu) + zero, woff = exq[q] >> 16;
            const int f = guide_fanout_log2(int(inside), s);
            if (woff + (1u << f) <= uint32_t(l2cap)) {
              wv[q] = kCellMulti | (uint32_t(f) << 14) | (woff << 18);
              list[woff >> 2] = make_uint2(uint32_t(f), jl);
              used_hi = woff + (1u << f) > used_hi ? woff + (1u << f) : used_hi;
            } else {
              wv[q] = kCellMulti | kCellSearch | jl | (inside << 14);
            }
          }
        }
        c4[(r0 + r) * nt + tid] = make_uint4(wv[0], wv[1], wv[2], wv[3]);
      }
      carry += wave_lane(tinc, np * nwaves - 1);
      if (r0 + 4 < passes) __syncthreads();  // sh.wtot is reused
    }
    // level-2 words in use: one atomic per wave (those of the cells that fit
    // are a prefix in cell order)
    if (ballot(used_hi != 0u) != 0ull) {
      used_hi = wave_inclusive_max(used_hi);
      if (lane == kWave - 1) atomicMax(&sh.used, int(used_hi));
    }
  }
  __syncthreads();

  // 6. level-2 words of the crowded cells.  Per-sub-cell counts of the splits
  //    strictly inside the cell (all | at the sub-cell's first key << 16) are
  //    scattered by the splits; then one flat scan over the level-2 words in
  //    use, four per thread (one list slot): the exclusive sum P of the counts
  //    and, by a running maximum of the slots that start a cell, each word's
  //    cell.  A sub-cell's jL is its cell's jL + P - P(cell's first word) + the
  //    splits at the sub-cell's first key.
  {
    const int used = sh.used, R = used / 4;
    uint32_t* l2 = cells + G;
    uint4* l24 = reinterpret_cast<uint4*>(l2);
    for (int64_t t = 1 + tid; t <= n - 2; t += nt) {
      const uint32_t key = srt[t];
      if (key >= (1u << 24)) continue;
      const uint32_t w = cells[key >> s];
      const uint32_t off = key & (S - 1u);
      if ((w & (kCellMulti | kCellSearch)) != kCellMulti || off == 0u) continue;
      const int ss = s - int((w >> 14) & 0xFu);
      atomicAdd(&l2[(w >> 18) + (off >> ss)],
                1u + ((off & ((1u << ss) - 1u)) == 0u ? 0x10000u : 0u));
    }
    __syncthreads();
    uint32_t carry_sum = 0u, carry_max = 0u;
    for (int b0 = 0; b0 < R; b0 += nt) {
      const int slot = b0 + tid;
      const bool valid = slot < R;
      const uint4 q4 = valid ? l24[slot] : make_uint4(0u, 0u, 0u, 0u);
      const uint32_t v[4] = {q4.x, q4.y, q4.z, q4.w};
      const uint2 ent = valid ? list[slot] : make_uint2(kNoCell, 0u);
      const uint32_t start = ent.x != kNoCell ? uint32_t(slot + 1) : 0u;
      const uint32_t local = (v[0] & 0xFFFFu) + (v[1] & 0xFFFFu) + (v[2] & 0xFFFFu) + (v[3] & 0xFFFFu);
      const uint32_t isum = wave_inclusive_scan(local);
      const uint32_t imax = wave_inclusive_max(start);
      if (lane == kWave - 1) {
        sh.wtot[wave] = isum;
        sh.wtot[32 + wave] = imax;
      }
      __syncthreads();
      uint32_t before = carry_sum, owner1 = carry_max > imax ? carry_max : imax;
      uint32_t tot_sum = 0u, tot_max = 0u;
      for (int w2 = 0; w2 < nwaves; ++w2) {
        const uint32_t ts = sh.wtot[w2], tm = sh.wtot[32 + w2];
        if (w2 < wave) {
          before += ts;
          owner1 = tm > owner1 ? tm : owner1;
        }
        tot_sum += ts;
        tot_max = tm > tot_max ? tm : tot_max;
      }
      const uint32_t p0 = before + isum - local;  // P of the slot's first word
      if (valid && start != 0u) list[slot].y = ent.y - p0;  // jL - P(first word)
      __syncthreads();
      if (valid) {
        const uint2 oe = list[int(owner1) - 1];  // {f, jL - P(first word)} of the word's cell
        const int ss = s - int(oe.x);
        uint32_t p = p0, wv[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const uint32_t all = v[q] & 0xFFFFu, zero = v[q] >> 16, inside = all - zero;
          const uint32_t jl = oe.y + p + zero;
          const uint32_t nxt = srt[jl + 1];
          wv[q] = inside >= 2u ? kCellMulti | kCellSearch | jl | (inside << 14)
                               : cell_direct(jl, inside, nxt, ss);
          p += all;
        }
        l24[slot] = make_uint4(wv[0], wv[1], wv[2], wv[3]);
      }
      carry_sum += tot_sum;
      carry_max = tot_max > carry_max ? tot_max : carry_max;
      if (b0 + nt < R) __syncthreads();  // sh.wtot is reused
    }
  }
  if (tid == 0) {
    const bool last_full = srt[n - 1] >= 16777214u;
    sh.flags = kFastTables | (mono ? kMonotone : 0u) |
               ((mono && safe && acc == 1.0 && last_full) ? kSafe : 0u);
  }
  __syncthreads();

  // 7. publish the sorted splits and the cell words in use
  if (writer) {
    lds_copy(blob + B.srt, srt, pad4(n) * 4);
    lds_copy(blob + B.cells, cells, pad4(G + sh.used) * 4);
    if (tid == 0) {
      TableHeader* hdr = reinterpret_cast<TableHeader*>(blob);
      hdr->norm = sh.norm;
      hdr->inv_n = sh.inv;
      hdr->flags = sh.flags;
      hdr->n = static_cast<uint32_t>(n);
      hdr->guide_log2 = static_cast<uint32_t>(g);
      hdr->level2_words = static_cast<uint32_t>(sh.used);
      if (P.norm_out) P.norm_out[inst] = sh.norm;
      P.done[inst] = 0u;
    }
  }
}

// Standalone build: one work-group per instance.  Before sampling it builds
// the global-table instances (only_global); for the VJP's recompute path it
// rebuilds every instance's blob.
template <typename Real, bool SEQUENTIAL>
__global__ void __launch_bounds__(1024) tables_kernel(BatchedParams P, int only_global) {
  extern __shared__ __align__(16) uint8_t lds[];
  __shared__ BuildShared sh;
  // only_global: the global-table instances of global_list (most nodes first)
  const int i = only_global ? P.global_list[blockIdx.x] : int(blockIdx.x);
  const InstanceDesc in = P.inst[i];
  if (in.global_tables)
    build_tables<Real, SEQUENTIAL, true>(P, i, in, lds, true, sh);
  else if (only_global)
    return;
  else if (sizeof(Real) == 4 && in.fast_log2 > 0)
    build_fast<SEQUENTIAL>(P, i, in, lds, true, sh, nullptr);
  else
    build_tables<Real, SEQUENTIAL, false>(P, i, in, lds, true, sh);
}

// ---------------------------------------------------------------------------
// Interval search: j = clamp(#{t : srt_t <= k} - 1, 0, n - 2).  One guide word
// settles almost every key; keys in redirected cells read one level-2 word,
// and the rare crowded (sub-)cells bisect the sorted thresholds.
template <int NK>
__device__ __forceinline__ void find_bins(const uint32_t (&k)[NK], const uint32_t* guide,
                                          const uint32_t* srt, int shift, int n, int (&j)[NK]) {
  const uint32_t cmask = (1u << shift) - 1u;
  uint32_t w[NK];
#pragma unroll
  for (int e = 0; e < NK; ++e) w[e] = guide[k[e] >> shift];
  uint32_t any = 0u;
#pragma unroll
  for (int e = 0; e < NK; ++e) {
    j[e] = guide_direct_bin(w[e], k[e] & cmask);
    any |= w[e];
  }
  if (ballot((any & kSpecial) != 0) != 0ull) {
    // Level 2 for every key of the wave at once (a key outside crowded cells
    // reads word 0 and keeps its bin): one uniform step instead of a branch
    // per key, since most trips have some lane in a crowded cell.
#pragma unroll
    for (int e = 0; e < NK; ++e) {
      const bool ind = (w[e] & kSearch) == kSpecial;
      const int ss = shift - (ind ? int((w[e] >> 24) & 0xFu) : 0);
      const uint32_t idx = ind ? (w[e] & 0xFFFFFFu) + ((k[e] & cmask) >> ss) : 0u;
      const uint32_t w2 = guide[idx];
      j[e] = ind ? guide_direct_bin(w2, k[e] & ((1u << ss) - 1u)) : j[e];
      w[e] = ind ? w2 : w[e];
    }
    int lo[NK], len[NK], pending = 0;
#pragma unroll
    for (int e = 0; e < NK; ++e) {
      const bool search = (w[e] & kSearch) == kSearch;
      lo[e] = int(w[e] & 0x7FFFu);
      len[e] = search ? int((w[e] >> 15) & 0x7FFFu) : 0;
      pending |= len[e];
    }
    if (pending != 0) {
      while (pending != 0) {
        pending = 0;
#pragma unroll
        for (int e = 0; e < NK; ++e) {
          const int half = len[e] >> 1;
          const int idx = lo[e] + half;  // < lo + len <= n while len > 0
          const bool le = len[e] > 0 && srt[len[e] > 0 ? idx : 0] <= k[e];
          lo[e] = le ? idx + 1 : lo[e];
          len[e] = len[e] > 0 ? (le ? len[e] - half - 1 : half) : 0;
          pending |= len[e];
        }
      }
#pragma unroll
      for (int e = 0; e < NK; ++e) {
        if ((w[e] & kSearch) == kSearch) {
          const int jj = lo[e] - 1;
          j[e] = jj < 0 ? 0 : (jj > n - 2 ? n - 2 : jj);
        }
      }
    }
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
// tarr: the bins' own thresholds T_j (the sorted array when it is the same).
template <typename Real, bool WINDOW, int NK, bool AOS = false>
__device__ __forceinline__ int slot_values(const uint32_t (&k)[NK], const int (&j)[NK],
                                           const typename CoefOf<Real>::type* coef,
                                           const uint32_t* tarr, const double* cdf_g,
                                           const Real* x_g, const BatchedParams& P,
                                           bool safe_curve, Real (&v)[NK]) {
  int keep = 0;
  if constexpr (sizeof(Real) == 4) {
    float2 cm[NK];
    uint32_t ct[NK];
    if constexpr (AOS) {
      const float4* rec = reinterpret_cast<const float4*>(coef);
#pragma unroll
      for (int e = 0; e < NK; ++e) {
        const float4 q = rec[j[e]];
        cm[e] = make_float2(q.x, q.y);
        ct[e] = __float_as_uint(q.z);
      }
    } else {
      const float2* ma = reinterpret_cast<const float2*>(coef);
#pragma unroll
      for (int e = 0; e < NK; ++e) {
        cm[e] = ma[j[e]];
        ct[e] = tarr[j[e]];
      }
    }
    if (!WINDOW && safe_curve) {  // every value of this curve is kept (kSafe)
#pragma unroll
      for (int e = 0; e < NK; ++e)
        v[e] = __builtin_fmaf(cm[e].x, static_cast<float>(static_cast<int32_t>(k[e] - ct[e])),
                              cm[e].y);
      return (1 << NK) - 1;
    }
    const float lo = static_cast<float>(P.win_lo), hi = static_cast<float>(P.win_hi);
    int unsafe = 0;
#pragma unroll
    for (int e = 0; e < NK; ++e) {
      v[e] = __builtin_fmaf(cm[e].x, static_cast<float>(static_cast<int32_t>(k[e] - ct[e])),
                            cm[e].y);
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
      for (int e = 0; e < NK; ++e) {
        if ((unsafe >> e) & 1) {
          const double v64 = interp_reference(cdf_g, x_g, j[e], k[e]);
          v[e] = static_cast<float>(v64);
          keep |= keep_value<WINDOW>(v64, P) ? (1 << e) : 0;
        }
      }
    }
  } else {
    CoefF64 cf[NK];
#pragma unroll
    for (int e = 0; e < NK; ++e) cf[e] = coef[j[e]];
#pragma unroll
    for (int e = 0; e < NK; ++e) {
      v[e] = __builtin_fma(cf[e].m, double(k[e]) * 0x1.0p-24 - cf[e].c, cf[e].x);
      keep |= keep_value<WINDOW>(v[e], P) ? (1 << e) : 0;
    }
  }
  return keep;
}

// Exclusive prefix of per-lane counts (0..4) within a wave, and the wave total.
__device__ __forceinline__ int wave_scan_counts(int cnt, int* total) {
  const uint64_t lt = lanemask_lt();
  const uint64_t b0 = ballot(cnt & 1), b1 = ballot(cnt & 2), b2 = ballot(cnt & 4);
  *total = popc64(b0) + 2 * popc64(b1) + 4 * popc64(b2);
  return popc64(b0 & lt) + 2 * popc64(b1 & lt) + 4 * popc64(b2 & lt);
}

// A value every lane of the wave holds, in scalar registers (the segment
// words below are loaded per lane; without this the loops that depend on them
// count as divergent and keep their mask words and counters in vector registers).
__device__ __forceinline__ int uniform_i32(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint64_t uniform_u64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(v));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(uint32_t(v >> 32));
  return (uint64_t(hi) << 32) | lo;
}
__device__ __forceinline__ int64_t uniform_i64(int64_t v) { return int64_t(uniform_u64(uint64_t(v))); }

// Keep bits of a lane's four slots in iteration `it` of a segment of the VJP,
// and the segment-relative packed row of its first kept slot (its other kept
// slots follow in slot order).  Before the segment's first drop every live
// slot was kept: slot 4 i + e is row 4 i + e - off0.  From it on the
// iteration's keep masks count the kept slots of the lower lanes (mbcnt), and
// `run` (wave-uniform) the rows of the iterations before; it must start at
// the rows before the first drop.
__device__ __forceinline__ int kept_slots(const uint64_t* mask, int it, int first_drop,
                                          uint32_t i0, uint32_t nblk, uint32_t off0,
                                          uint32_t span, int64_t& run, int64_t* row) {
  const int lane = lane_id();
  const uint32_t i = i0 + uint32_t(lane);
  if (it < first_drop) {
    int live = 0xF;
    if (!((i0 > 0 || off0 == 0) && 4 * (i0 + 64) <= span)) {
      live = 0;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const uint32_t r = 4 * i + e;
        live |= (i < nblk && r >= off0 && r < span) ? (1 << e) : 0;
      }
    }
    const int32_t r0 = int32_t(4 * i) - int32_t(off0);
    *row = r0 < 0 ? 0 : r0;
    return live;
  }
  const uint64_t* mw = mask + int64_t(it) * 4;
  const uint64_t m0 = uniform_u64(mw[0]), m1 = uniform_u64(mw[1]), m2 = uniform_u64(mw[2]),
                 m3 = uniform_u64(mw[3]);
  const int kb = int((m0 >> lane) & 1ull) | (int((m1 >> lane) & 1ull) << 1) |
                 (int((m2 >> lane) & 1ull) << 2) | (int((m3 >> lane) & 1ull) << 3);
  uint32_t below = __builtin_amdgcn_mbcnt_lo(uint32_t(m0), 0u);
  below = __builtin_amdgcn_mbcnt_hi(uint32_t(m0 >> 32), below);
  below = __builtin_amdgcn_mbcnt_lo(uint32_t(m1), below);
  below = __builtin_amdgcn_mbcnt_hi(uint32_t(m1 >> 32), below);
  below = __builtin_amdgcn_mbcnt_lo(uint32_t(m2), below);
  below = __builtin_amdgcn_mbcnt_hi(uint32_t(m2 >> 32), below);
  below = __builtin_amdgcn_mbcnt_lo(uint32_t(m3), below);
  below = __builtin_amdgcn_mbcnt_hi(uint32_t(m3 >> 32), below);
  *row = run + int64_t(below);
  run += popc64(m0) + popc64(m1) + popc64(m2) + popc64(m3);
  return kb;
}

// Keep-mask words of a chunk: wave w's iteration `it` owns words
// [(w * per_wave + it) * 4, +4), one 64-bit ballot per slot position e.
__device__ __forceinline__ int64_t chunk_blocks(const InstanceDesc& in, const ChunkDesc& ch) {
  if (ch.slot_hi <= ch.slot_lo) return 0;
  return int64_t(((in.counter_base + uint64_t(ch.slot_hi) + 3) >> 2) -
                 ((in.counter_base + uint64_t(ch.slot_lo)) >> 2));
}

// Next segment for this wave from the work-group's LDS counter (uniform).
__device__ __forceinline__ int next_segment(int* counter) {
  int s = 0;
  if (lane_id() == 0) s = atomicAdd(counter, 1);
  return __builtin_amdgcn_readfirstlane(s);
}

// ---------------------------------------------------------------------------
// Finalize, by one work-group: chunk_row = exclusive scan of the chunks' kept
// counts (four chunks per thread per tile); kept per instance; total; dense
// flag (every chunk kept all of its slots).  chunk_kept holds kept + 1 for a
// chunk that dropped slots and 0 for one that kept them all (which writes
// nothing); the finalizer returns every entry to 0 for the next call.
__device__ void finalize_chunks(const BatchedParams& P) {
  __shared__ int64_t scratch[16];
  __shared__ int sh_sparse;
  const int tid = threadIdx.x, nt = blockDim.x;
  if (tid == 0) sh_sparse = 0;
  const int64_t nch = P.n_chunks;
  int64_t carry = 0;
  bool sparse = false;
  for (int64_t base = 0; base < nch; base += 4 * int64_t(nt)) {
    const int64_t c0 = base + 4 * int64_t(tid);
    int64_t v[4], local = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int64_t ci = c0 + q;
      v[q] = 0;
      if (ci < nch) {
        const ChunkDesc ch = P.chunks[ci];
        const int64_t stored = P.chunk_kept[ci];
        v[q] = stored == 0 ? ch.slot_hi - ch.slot_lo : stored - 1;
        sparse = sparse || stored != 0;
        if (stored != 0) P.chunk_kept[ci] = 0;
      }
      local += v[q];
    }
    int64_t tot;
    int64_t run = carry + block_exclusive_scan(local, scratch, &tot);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (c0 + q < nch) P.chunk_row[c0 + q] = run;
      run += v[q];
    }
    carry += tot;
  }
  if (sparse) atomicOr(&sh_sparse, 1);
  __syncthreads();  // chunk_row written above is visible to the whole work-group
  for (int64_t i = tid; i < P.n_instances; i += nt) {
    const InstanceDesc in = P.inst[i];
    const int64_t c1 = int64_t(in.first_chunk) + in.n_chunks;
    P.kept[i] = (c1 < nch ? P.chunk_row[c1] : carry) - P.chunk_row[in.first_chunk];
  }
  if (tid == 0) {
    P.total[0] = carry;
    P.flags[0] = sh_sparse ? 0 : 1;  // dense
  }
}

// Standalone finalize (a plan without chunks).
__global__ void __launch_bounds__(1024) finalize_kernel(BatchedParams P) { finalize_chunks(P); }

// Finalize of a forward in which every chunk kept every slot: rows are the
// capacity rows, so chunk row starts, kept counts and the total follow from the
// descriptors alone.
__device__ void finalize_dense(const BatchedParams& P) {
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int64_t c = tid; c < P.n_chunks; c += nt) {
    const ChunkDesc ch = P.chunks[c];
    P.chunk_row[c] = P.inst[ch.instance].out_off + ch.slot_lo;
  }
  for (int64_t i = tid; i < P.n_instances; i += nt) P.kept[i] = P.inst[i].events;
  if (tid == 0) {
    const InstanceDesc last = P.inst[P.n_instances - 1];
    P.total[0] = last.out_off + last.events;
    P.flags[0] = 1;  // dense
  }
}

// End of a sampling work-group's chunk c: the last of the n_chunks chunks to
// finish (a 64-bit counter in flags[2..3]: finished chunks in the low word,
// chunks that dropped a slot in the high word; the last resets it) finalizes.
// A chunk that kept every slot publishes nothing the finalizer reads, so it
// counts itself with a relaxed atomic and no fence; a chunk that dropped slots
// releases its chunk_kept (an L2 write-back), and the finalizer then acquires
// them.  An agent-scope fence in every work-group cost about 3.5 us at the end
// of each.
__device__ void finalize_if_last(const BatchedParams& P, int c) {
  __shared__ int sh_last;  // 0: not last; 1: last, every chunk dense; 2: last, some sparse
  __syncthreads();  // this group's chunk_kept and seg_row are written
  if (threadIdx.x == 0) {
    const bool dense = P.chunk_kept[c] == 0;
    unsigned long long* counter = reinterpret_cast<unsigned long long*>(P.flags + 2);
    const unsigned long long old =
        dense ? __hip_atomic_fetch_add(counter, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
              : __hip_atomic_fetch_add(counter, (1ull << 32) | 1ull, __ATOMIC_RELEASE,
                                       __HIP_MEMORY_SCOPE_AGENT);
    const bool last = uint32_t(old) == uint32_t(P.n_chunks - 1);
    const bool all_dense = dense && (old >> 32) == 0ull;
    sh_last = last ? (all_dense ? 1 : 2) : 0;
    if (last) {
      if (!all_dense) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      __hip_atomic_store(counter, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  __syncthreads();
  if (sh_last == 1)
    finalize_dense(P);
  else if (sh_last == 2)
    finalize_chunks(P);
}

// ---------------------------------------------------------------------------
// Sampling: one work-group per chunk; waves take segments from a counter.  While every slot
// of a segment so far was kept, rows are positional and no keep mask is
// written; the first iteration with a dropped slot is recorded, and masks are
// written from that iteration on.
// One chunk's sampling loop over tables in LDS or, for global-table instances,
// a guide in LDS and the per-bin arrays in global memory (forceinline keeps
// each call's address spaces known to the compiler).
// Interval search and values of a trip's keys over the two-level guide tables
// (find_bins, slot_values); returns the keep bits.
// AOS: float32 per-bin data as one {M, A, T, r} record per bin (the layout of
// global tables, one 16-byte read per key from L2) instead of the arrays.
template <typename Real, bool WINDOW, bool AOS = false>
struct GuideLookup {
  static constexpr int kBinBias = 0;  // returns j
  const uint32_t* guide;
  const uint32_t* srt;
  const typename CoefOf<Real>::type* coef;
  const uint32_t* thr;
  const double* cdf;  // global
  const Real* x;      // global
  int shift, n;
  bool safe;

  template <int NK>
  __device__ __forceinline__ int operator()(const uint32_t (&k)[NK], int (&j)[NK], Real (&v)[NK],
                                            const BatchedParams& P) const {
    find_bins<NK>(k, guide, srt, shift, n, j);
    return slot_values<Real, WINDOW, NK, AOS>(k, j, coef, thr, cdf, x, P, safe, v);
  }
};

// j + 1 of a trip's keys from the cell tables of fast_tables.hpp: one cell
// word per key; the level-2 step or bisection only at key positions where some
// lane of the wave needs it (each position's branch is wave-uniform).
template <int NK>
__device__ __forceinline__ void cell_bins(const uint32_t (&k)[NK], const uint32_t* cells,
                                          const uint32_t* l2, const uint32_t* srt, int s,
                                          int (&jj)[NK]) {
  // the ones below the key's offset, held in a vector register so that X is a
  // single shift-or (one scalar operand per vector instruction)
  uint32_t low = (1u << (32 - s)) - 1u;
  asm("v_mov_b32 %0, %1" : "=v"(low) : "s"(low));
  uint32_t w[NK], x[NK], t[NK];
#pragma unroll
  for (int e = 0; e < NK; ++e) w[e] = cells[k[e] >> s];
#pragma unroll
  for (int e = 0; e < NK; ++e) {
    x[e] = (k[e] << (32 - s)) | low;
    t[e] = w[e] & (kCellMulti | kCellField);
  }
#pragma unroll
  for (int e = 0; e < NK; ++e) {
    int r = int(t[e]) + (x[e] >= w[e] ? 1 : 0);
    const bool m = t[e] > kCellField;  // indirect or search word
    if (ballot(m) != 0ull)
      if (m) r = cell_multi_jj(w[e], x[e], k[e], l2, srt);
    jj[e] = r;
  }
}

// The same over the cell tables of fast_tables.hpp (float32): one cell word per
// key, the level-2 step only at key positions where some lane needs it, then
// one float4 {M, A, T} read per key.  Returns j + 1 (kBinBias).
template <bool WINDOW>
struct CellLookup {
  static constexpr int kBinBias = 1;
  const uint32_t* cells;
  const uint32_t* l2;
  const uint32_t* srt;
  const float4* bins1;  // bins - 1
  const double* cdf;    // global
  const float* x;       // global
  int s;
  bool safe;
  bool classed;  // window keep classes in the bins' fourth word (monotone curve)

  template <int NK>
  __device__ __forceinline__ int operator()(const uint32_t (&k)[NK], int (&j)[NK], float (&v)[NK],
                                            const BatchedParams& P) const {
    cell_bins<NK>(k, cells, l2, srt, s, j);
    float4 b[NK];
#pragma unroll
    for (int e = 0; e < NK; ++e) b[e] = bins1[j[e]];
    if (!WINDOW && safe) {  // every value of this curve is kept (kSafe)
#pragma unroll
      for (int e = 0; e < NK; ++e)
        v[e] = __builtin_fmaf(
            b[e].x, static_cast<float>(static_cast<int32_t>(k[e] - __float_as_uint(b[e].z))),
            b[e].y);
      return (1 << NK) - 1;
    }
    const float lo = static_cast<float>(P.win_lo), hi = static_cast<float>(P.win_hi);
    int keep = 0, unsafe = 0;
    // the float32 decision, or "unsafe" where it could differ from fp64's
    auto decide = [&](int e) {
      const float a = __builtin_fabsf(v[e]);
      bool ok = a >= 1e-30f && a <= 1e30f;
      bool in = true;
      if (WINDOW) {
        const float margin = 2e-6f * a;
        ok = ok && __builtin_fabsf(v[e] - lo) > margin && __builtin_fabsf(v[e] - hi) > margin;
        in = v[e] >= lo && v[e] <= hi;
      }
      keep |= (ok && in) ? (1 << e) : 0;
      unsafe |= ok ? 0 : (1 << e);
    };
#pragma unroll
    for (int e = 0; e < NK; ++e)
      v[e] = __builtin_fmaf(
          b[e].x, static_cast<float>(static_cast<int32_t>(k[e] - __float_as_uint(b[e].z))),
          b[e].y);
    if (WINDOW && classed) {  // the bin's class decides, except in straddling bins
      uint32_t least = 0xFFFFFFFFu;  // a zero class marks a straddling bin
#pragma unroll
      for (int e = 0; e < NK; ++e) {
        const uint32_t cls = __float_as_uint(b[e].w);
        keep |= int(cls & 1u) << e;
        least = cls < least ? cls : least;
      }
      if (ballot(least == 0u) != 0ull) {
#pragma unroll
        for (int e = 0; e < NK; ++e)
          if (__float_as_uint(b[e].w) == 0u) decide(e);
      }
    } else {
#pragma unroll
      for (int e = 0; e < NK; ++e) decide(e);
    }
    if (__builtin_expect(unsafe != 0, 0)) {
#pragma unroll 1
      for (int e = 0; e < NK; ++e) {
        if ((unsafe >> e) & 1) {
          const double v64 = interp_reference(cdf, x, j[e] - 1, k[e]);
          v[e] = static_cast<float>(v64);
          keep |= keep_value<WINDOW>(v64, P) ? (1 << e) : 0;
        }
      }
    }
    return keep;
  }
};

template <typename Real, bool WINDOW, typename Lookup>
__device__ __forceinline__ void sample_segment(const BatchedParams& P, int c, const ChunkDesc& ch,
                                               const InstanceDesc& in, int w, const Lookup& lookup,
                                               int64_t* sh_kept) {
  constexpr int kBias = Lookup::kBinBias;  // the lookup returns j + kBias
  const int nw = P.segs, lane = lane_id();  // segment w of nw in this chunk
  const Segment sg = segment_of(in.counter_base, ch.slot_lo, ch.slot_hi, w, nw);
  const int64_t seg = int64_t(c) * nw + w;
  const int64_t base_row = in.out_off + int64_t(sg.p_lo - in.counter_base);
  Real* out = static_cast<Real*>(P.events) + base_row;
  int32_t* bins_out = P.bins ? P.bins + base_row : nullptr;
  const int64_t per_wave = int64_t(blocks_per_wave(uint64_t(chunk_blocks(in, ch)), nw) / 64);
  uint64_t* mask = P.keepmask + ch.mask_off + int64_t(w) * per_wave * 4;
  // 32-bit slot positions relative to counter 4 qa: live slots are [off0, span)
  const uint32_t nblk = uint32_t(sg.qb - sg.qa);
  const uint32_t off0 = uint32_t(sg.p_lo - 4 * sg.qa);
  const uint32_t span = uint32_t(sg.p_hi - 4 * sg.qa);
  int64_t running = 0;
  int first_drop = int(sg.iters);  // iterations before this one kept every slot

  // One iteration (four slots per lane): the running kept count; the keep
  // masks from the segment's first iteration that dropped a live slot on; and
  // every live slot's value (and bin) at its own capacity row 4 i + e - off0,
  // kept or not, so that no row depends on a scan (pack_kernel gathers the
  // kept rows).  Whole blocks go out as one 16-byte store where the rows are
  // aligned.
  const bool vec_rows = ((base_row - int64_t(off0)) & 3) == 0;
  Real* out0 = out - off0;  // slot position 4 i + e of the segment
  int32_t* bins0 = bins_out ? bins_out - off0 : nullptr;
  auto emit = [&](int it, uint32_t i0, const Real* v, const int* j, int keep) {
    const uint32_t i = i0 + uint32_t(lane);
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
    if (first_drop == int(sg.iters) && ballot(keepbits != live) == 0ull) {
      const uint32_t a = 4 * i0 > off0 ? 4 * i0 : off0;
      const uint32_t b = 4 * (i0 + 64) < span ? 4 * (i0 + 64) : span;
      running += b > a ? int64_t(b - a) : 0;
    } else {
      if (first_drop == int(sg.iters)) first_drop = it;
      int total = 0;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const uint64_t bb = ballot((keepbits >> e) & 1);
        if (lane == e) mask[int64_t(it) * 4 + e] = bb;
        total += popc64(bb);
      }
      running += total;
    }
    if (vec_rows && live == 0xF) {
      if constexpr (sizeof(Real) == 4) {
        store_stream4(out0 + 4 * i, v[0], v[1], v[2], v[3]);
      } else {
        reinterpret_cast<double2*>(out0 + 4 * i)[0] = make_double2(v[0], v[1]);
        reinterpret_cast<double2*>(out0 + 4 * i)[1] = make_double2(v[2], v[3]);
      }
      if (bins0)
        *reinterpret_cast<int4*>(bins0 + 4 * i) =
            make_int4(j[0] - kBias, j[1] - kBias, j[2] - kBias, j[3] - kBias);
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        if ((live >> e) & 1) {
          out0[4 * i + e] = v[e];
          if (bins0) bins0[4 * i + e] = j[e] - kBias;
        }
      }
    }
  };

  // A curve whose every value is kept (kSafe, no window): slot 4 i + e of the
  // segment is row 4 i + e - off0, so rows need no keep bits, ballots or masks;
  // whole blocks go out as one 16-byte store when the rows are aligned, and
  // trips of whole blocks only (every trip but the segment's last) skip the
  // per-lane bounds.
  if (!WINDOW && lookup.safe) {
    const bool vec = off0 == 0u && (base_row & 3) == 0;
    const uint32_t nfull = vec ? span / 4 : 0u;  // blocks whose four rows all exist
    for (uint32_t i0 = 0; i0 < nblk; i0 += 64 * kTrip) {
      uint32_t k[4 * kTrip];
#pragma unroll
      for (int h = 0; h < kTrip; ++h) {
        const Words4 wd =
            philox4x32_10_uniform(sg.qa + i0 + 64 * h + uint32_t(lane), in.stream, P.seed);
        k[4 * h] = wd.w0 >> 8;
        k[4 * h + 1] = wd.w1 >> 8;
        k[4 * h + 2] = wd.w2 >> 8;
        k[4 * h + 3] = wd.w3 >> 8;
      }
      int j[4 * kTrip];
      Real v[4 * kTrip];
      lookup(k, j, v, P);
      if (i0 + 64 * kTrip <= nfull) {  // wave-uniform
#pragma unroll
        for (int h = 0; h < kTrip; ++h) {
          const uint32_t i = i0 + 64 * h + uint32_t(lane);
          if constexpr (sizeof(Real) == 4) {
            store_stream4(out + 4 * i, v[4 * h], v[4 * h + 1], v[4 * h + 2], v[4 * h + 3]);
          } else {
            reinterpret_cast<double2*>(out + 4 * i)[0] = make_double2(v[4 * h], v[4 * h + 1]);
            reinterpret_cast<double2*>(out + 4 * i)[1] = make_double2(v[4 * h + 2], v[4 * h + 3]);
          }
          if (bins_out)
            *reinterpret_cast<int4*>(bins_out + 4 * i) =
                make_int4(j[4 * h] - kBias, j[4 * h + 1] - kBias, j[4 * h + 2] - kBias,
                          j[4 * h + 3] - kBias);
        }
        continue;
      }
#pragma unroll
      for (int h = 0; h < kTrip; ++h) {
        const uint32_t i = i0 + 64 * h + uint32_t(lane);
        if (vec && 4 * i + 4 <= span) {
          if constexpr (sizeof(Real) == 4) {
            store_stream4(out + 4 * i, v[4 * h], v[4 * h + 1], v[4 * h + 2], v[4 * h + 3]);
          } else {
            reinterpret_cast<double2*>(out + 4 * i)[0] = make_double2(v[4 * h], v[4 * h + 1]);
            reinterpret_cast<double2*>(out + 4 * i)[1] = make_double2(v[4 * h + 2], v[4 * h + 3]);
          }
          if (bins_out)
            *reinterpret_cast<int4*>(bins_out + 4 * i) =
                make_int4(j[4 * h] - kBias, j[4 * h + 1] - kBias, j[4 * h + 2] - kBias,
                          j[4 * h + 3] - kBias);
        } else {
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const uint32_t r = 4 * i + e;
            if (i < nblk && r >= off0 && r < span) {
              out[r - off0] = v[4 * h + e];
              if (bins_out) bins_out[r - off0] = j[4 * h + e] - kBias;
            }
          }
        }
      }
    }
    if (lane == 0) {
      running = int64_t(span - off0);
      P.seg_kept[seg] = running;
      sh_kept[w] = running;
      P.seg_drop[seg] = first_drop;
    }
    return;
  }

  // kTrip iterations per trip: their Philox, search and interpolation chains
  // are independent, which hides LDS latency at four waves per SIMD.
  int it = 0;
  for (uint32_t i0 = 0; i0 < nblk; i0 += 64 * kTrip, it += kTrip) {
    uint32_t k[4 * kTrip];
#pragma unroll
    for (int h = 0; h < kTrip; ++h) {
      const Words4 wd =
          philox4x32_10_uniform(sg.qa + i0 + 64 * h + uint32_t(lane), in.stream, P.seed);
      k[4 * h] = wd.w0 >> 8;
      k[4 * h + 1] = wd.w1 >> 8;
      k[4 * h + 2] = wd.w2 >> 8;
      k[4 * h + 3] = wd.w3 >> 8;
    }
    int j[4 * kTrip];
    Real v[4 * kTrip];
    const int keep = lookup(k, j, v, P);
#pragma unroll
    for (int h = 0; h < kTrip; ++h)
      if (h == 0 || i0 + 64 * h < nblk)
        emit(it + h, i0 + 64 * h, v + 4 * h, j + 4 * h, (keep >> (4 * h)) & 0xF);
  }
  if (lane == 0) {
    P.seg_kept[seg] = running;
    sh_kept[w] = running;
    P.seg_drop[seg] = first_drop;
  }
}

// Rows of a chunk's segments relative to the chunk, and the chunk's total.
__device__ __forceinline__ void chunk_rows(const BatchedParams& P, int c, const int64_t* sh_kept) {
  __syncthreads();
  if (threadIdx.x < kWave) {
    const int lane = lane_id();
    const int64_t v = lane < P.segs ? sh_kept[lane] : 0;
    int64_t inc = v;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
      const uint32_t lo = __shfl_up(uint32_t(inc), d, kWave);
      const uint32_t hi = __shfl_up(uint32_t(uint64_t(inc) >> 32), d, kWave);
      if (lane >= d) inc += int64_t(uint64_t(lo) | (uint64_t(hi) << 32));
    }
    if (lane < P.segs) P.seg_row[int64_t(c) * P.segs + lane] = inc - v;
    if (lane == kWave - 1) {  // kept + 1 when the chunk dropped slots (finalize_chunks)
      const ChunkDesc ch = P.chunks[c];
      if (inc != ch.slot_hi - ch.slot_lo) P.chunk_kept[c] = inc + 1;
    }
  }
}

// Sampling of one chunk of a guide-table instance by the work-group; waves take
// segments from a counter.  Ends with the chunk's segment rows (chunk_rows).
template <typename Real, bool WINDOW, bool SEQUENTIAL>
__device__ __forceinline__ void guide_chunk(const BatchedParams& P, int c, uint8_t* lds,
                                            BuildShared& sh, int* sh_next, int64_t* sh_kept) {
  const ChunkDesc ch = P.chunks[c];
  const InstanceDesc in = P.inst[ch.instance];
  const int64_t nb = pad4(in.n_nodes);
  if (threadIdx.x == 0) *sh_next = 0;
  const Real* x_g = static_cast<const Real*>(P.x) + in.node_off;
  if (in.global_tables) {  // built beforehand (tables_kernel or a queue item); stage the guide
    const Tables<Real> T = tables_of<Real>(P.tables, in);
    lds_copy(lds, T.guide, guide_bytes(T, in));
    if (threadIdx.x == 0) sh.flags = T.hdr->flags;
    __syncthreads();
    // float32 global tables hold each bin's own threshold in its record
    const GuideLookup<Real, WINDOW, sizeof(Real) == 4> lk{
        reinterpret_cast<const uint32_t*>(lds), T.srt, T.coef, T.srt, T.cdf, x_g,
        24 - in.guide_log2, int(in.n_nodes), (sh.flags & kSafe) != 0};
    for (int sgi = next_segment(sh_next); sgi < P.segs; sgi = next_segment(sh_next))
      sample_segment<Real, WINDOW>(P, c, ch, in, sgi, lk, sh_kept);
  } else {
    build_tables<Real, SEQUENTIAL>(P, ch.instance, in, lds, ch.index == 0, sh);
    const LdsTables<Real> L = lds_tables_of<Real>(lds, in);
    const uint32_t* thr = (sh.flags & kMonotone) != 0 || sizeof(Real) != 4
                              ? L.srt
                              : coef_f32_view(L.coef, nb).thr;
    const GuideLookup<Real, WINDOW> lk{L.guide, L.srt, L.coef, thr, L.cdf, x_g,
                                       24 - in.guide_log2, int(in.n_nodes),
                                       (sh.flags & kSafe) != 0};
    for (int sgi = next_segment(sh_next); sgi < P.segs; sgi = next_segment(sh_next))
      sample_segment<Real, WINDOW>(P, c, ch, in, sgi, lk, sh_kept);
  }
  chunk_rows(P, c, sh_kept);
}

// Sampling of one chunk of a float32 cell-table instance (fast_tables.hpp).
template <bool WINDOW, bool SEQUENTIAL>
__device__ __forceinline__ void cell_chunk(const BatchedParams& P, int c, uint8_t* lds,
                                           BuildShared& sh, int* sh_next, int64_t* sh_kept) {
  const ChunkDesc ch = P.chunks[c];
  const InstanceDesc in = P.inst[ch.instance];
  if (threadIdx.x == 0) *sh_next = 0;
  // the chunk's VJP partial sums are free during the forward: they hold this
  // group's copy of the CDF
  double* cdf_copy = P.partials + ch.partial_off;
  build_fast<SEQUENTIAL>(P, ch.instance, in, lds, ch.index == 0, sh, cdf_copy);
  const FastLdsForward L = fast_lds_forward(in.n_nodes, in.fast_log2);
  const uint32_t* cells = reinterpret_cast<const uint32_t*>(lds + L.cells);
  const CellLookup<WINDOW> lk{cells, cells + (int64_t(1) << in.fast_log2),
                              reinterpret_cast<const uint32_t*>(lds + L.srt),
                              reinterpret_cast<const float4*>(lds + L.bins) - 1, cdf_copy,
                              static_cast<const float*>(P.x) + in.node_off, 24 - in.fast_log2,
                              (sh.flags & kSafe) != 0,
                              WINDOW && (sh.flags & kMonotone) != 0};
  for (int sgi = next_segment(sh_next); sgi < P.segs; sgi = next_segment(sh_next))
    sample_segment<float, WINDOW>(P, c, ch, in, sgi, lk, sh_kept);
  chunk_rows(P, c, sh_kept);
}

// Sampling of the guide-table instances: one work-group per chunk (the chunks
// after the cell-table ones in fwd_order).
template <typename Real, bool WINDOW, bool SEQUENTIAL>
__global__ void __launch_bounds__(1024) sample_kernel(BatchedParams P) {
  extern __shared__ __align__(16) uint8_t lds[];
  __shared__ BuildShared sh;
  __shared__ int sh_next;  // next unclaimed segment (read after the barriers below)
  __shared__ int64_t sh_kept[kWave];  // kept events per segment (at most 64 per chunk)
  const int c = P.fwd_order[P.n_cell_chunks + blockIdx.x];
  guide_chunk<Real, WINDOW, SEQUENTIAL>(P, c, lds, sh, &sh_next, sh_kept);
  finalize_if_last(P, c);
}

// Sampling of the float32 cell-table instances: one work-group per chunk (the
// first n_cell_chunks of fwd_order).  A kernel of its own, so that neither
// path's registers limit the other's occupancy.
template <bool WINDOW, bool SEQUENTIAL>
__global__ void __launch_bounds__(1024) sample_cell_kernel(BatchedParams P) {
  extern __shared__ __align__(16) uint8_t lds[];
  __shared__ BuildShared sh;
  __shared__ int sh_next;
  __shared__ int64_t sh_kept[kWave];
  const int c = P.fwd_order[blockIdx.x];
  cell_chunk<WINDOW, SEQUENTIAL>(P, c, lds, sh, &sh_next, sh_kept);
  finalize_if_last(P, c);
}

// Claim the next item of a persistent queue (counter flags[next]) for the work-group.
__device__ __forceinline__ int claim_item(const BatchedParams& P, int next, int* sh_item) {
  if (threadIdx.x == 0) *sh_item = atomicAdd(P.flags + next, 1);
  __syncthreads();
  return *sh_item;
}

// End of a persistent launch: the last work-group to leave returns the queue's
// counters to zero (every other group has made its last claim by then).
__device__ __forceinline__ void leave_queue(const BatchedParams& P, int next, int exited) {
  if (threadIdx.x == 0) {
    const int old = atomicAdd(P.flags + exited, 1);
    if (old == int(gridDim.x) - 1) {
      atomicExch(P.flags + next, 0);
      atomicExch(P.flags + exited, 0);
    }
  }
}

// Items of the persistent launches, each compiled as a function of its own
// (noinline) so that every kind keeps the register allocation it has in its
// own kernel; the work-group's dynamic LDS and shared words are declared here.
// A callee's arguments arrive in vector registers, so the kernel passes the
// address of its kernel-argument segment (BatchedParams first), which each
// callee makes wave-uniform and reads through the constant address space, and
// the item index, made wave-uniform too: both stay scalar.  (The segment
// pointer itself is not available inside a callee.)
using KernargPtr = const __attribute__((address_space(4))) uint8_t*;

__device__ __forceinline__ uint64_t kernarg_address() {
  return reinterpret_cast<uint64_t>(__builtin_amdgcn_kernarg_segment_ptr());
}

__device__ __forceinline__ KernargPtr uniform_kernarg(uint64_t a) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane(uint32_t(a));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(uint32_t(a >> 32));
  return (KernargPtr)((uint64_t(hi) << 32) | lo);
}

__device__ __forceinline__ const BatchedParams& launch_params(uint64_t kargs) {
  return *(const BatchedParams*)(const __attribute__((address_space(4))) BatchedParams*)
      uniform_kernarg(kargs);
}

template <bool WINDOW, bool SEQUENTIAL>
__device__ __noinline__ void persistent_cell_chunk(uint64_t kargs, int c_arg) {
  const BatchedParams& P = launch_params(kargs);
  const int c = __builtin_amdgcn_readfirstlane(c_arg);
  extern __shared__ __align__(16) uint8_t lds[];
  __shared__ BuildShared sh;
  __shared__ int sh_next;
  __shared__ int64_t sh_kept[kWave];
  cell_chunk<WINDOW, SEQUENTIAL>(P, c, lds, sh, &sh_next, sh_kept);
}

template <typename Real, bool WINDOW, bool SEQUENTIAL>
__device__ __noinline__ void persistent_guide_chunk(uint64_t kargs, int c_arg) {
  const BatchedParams& P = launch_params(kargs);
  const int c = __builtin_amdgcn_readfirstlane(c_arg);
  extern __shared__ __align__(16) uint8_t lds[];
  __shared__ BuildShared sh;
  __shared__ int sh_next;
  __shared__ int64_t sh_kept[kWave];
  guide_chunk<Real, WINDOW, SEQUENTIAL>(P, c, lds, sh, &sh_next, sh_kept);
}

// Persistent forward: resident work-groups take the chunks of fwd_queue,
// largest first, cell or guide tables alike.  The global-table instances are
// built by tables_kernel beforehand: a build inside the launch would hand its
// tables to work-groups on other XCDs, whose L2s are not coherent with the
// builder's within a launch, and the agent-scope release and acquire fences
// that takes (an L2 write-back per build, an invalidate per waiting chunk) cost
// more than the separate launch.  The last chunk to finish finalizes.
template <typename Real, bool WINDOW, bool SEQUENTIAL>
__global__ void __launch_bounds__(1024) forward_persistent_kernel(BatchedParams P) {
  __shared__ int sh_item;
  for (;;) {
    const int item = claim_item(P, kFlagFwdNext, &sh_item);
    if (item >= P.n_fwd_items) break;
    const int c = P.fwd_queue[item];
    bool cell = false;
    if constexpr (sizeof(Real) == 4) cell = P.inst[P.chunks[c].instance].fast_log2 > 0;
    if (cell) {
      if constexpr (sizeof(Real) == 4)
        persistent_cell_chunk<WINDOW, SEQUENTIAL>(kernarg_address(), c);
    } else {
      persistent_guide_chunk<Real, WINDOW, SEQUENTIAL>(kernarg_address(), c);
    }
    finalize_if_last(P, c);
    __syncthreads();  // the item's LDS and shared words are free
  }
  leave_queue(P, kFlagFwdNext, kFlagFwdExit);
}

// Pack: one work-group per chunk, segments round-robin over its waves; a no-op
// when dense.  The forward left slot 4 i + e of a segment at capacity row
// base + 4 i + e - off0 (kept or not): the rows before the segment's first
// drop are copied straight, the rest are gathered by their keep masks.
constexpr int kPackTrip = 2;  // mask iterations per wave with loads in flight

// One segment's rows of one array (values or bins) into `out` from row dst.
template <typename T>
__device__ __forceinline__ void pack_segment(const T* src_rows, T* out, int64_t dst,
                                             int64_t straight, const uint64_t* mask,
                                             int first_drop, int iters, int64_t off0) {
  const int lane = lane_id();
  for (int64_t r0 = 0; r0 < straight; r0 += 8 * kWave) {  // eight rows per lane in flight
    T v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int64_t r = r0 + u * kWave + lane;
      v[u] = r < straight ? src_rows[r] : T(0);
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int64_t r = r0 + u * kWave + lane;
      if (r < straight) out[dst + r] = v[u];
    }
  }
  if (first_drop >= iters) return;
  // From the first drop on, four positions p = 64 m + lane of each
  // iteration's 256 slots per lane: slot p of the forward's wave was lane
  // p / 4, word p % 4 of the iteration's keep masks.  Loads and stores are
  // coalesced: the kept positions of one m take consecutive rows, the
  // ballot's lower bits counting those before the lane.
  const T* pos0 = src_rows - off0;  // slot position 4 i + e of the segment
  int64_t running = dst + straight;
  const int sel = lane & 3, quad = lane >> 2;
  for (int it0 = first_drop; it0 < iters; it0 += kPackTrip) {
    T v[4 * kPackTrip];
    uint64_t bb[4 * kPackTrip];
#pragma unroll
    for (int h = 0; h < kPackTrip; ++h) {
      const bool in_seg = it0 + h < iters;
      const int it = in_seg ? it0 + h : it0;
      const uint64_t* mw = mask + int64_t(it) * 4;
      const uint64_t word = sel == 0 ? mw[0] : (sel == 1 ? mw[1] : (sel == 2 ? mw[2] : mw[3]));
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        const bool k = in_seg && ((word >> (16 * m + quad)) & 1ull);
        v[4 * h + m] = k ? pos0[int64_t(it) * 256 + 64 * m + lane] : T(0);
        bb[4 * h + m] = ballot(k);
      }
    }
#pragma unroll
    for (int q = 0; q < 4 * kPackTrip; ++q) {
      if ((bb[q] >> lane) & 1ull) out[running + popc64(bb[q] & lanemask_lt())] = v[q];
      running += popc64(bb[q]);
    }
  }
}

// Eight waves per SIMD for float32 (two work-groups per CU), four for float64.
template <typename Real>
__global__ void __launch_bounds__(1024, sizeof(Real) == 4 ? 8 : 4)
    pack_kernel(BatchedParams P, Real* packed, int32_t* packed_bins) {
  if (P.flags[0]) return;
  const ChunkDesc ch = P.chunks[blockIdx.x];
  const InstanceDesc in = P.inst[ch.instance];
  const int nwv = blockDim.x >> 6;
  const int64_t per_wave = int64_t(blocks_per_wave(uint64_t(chunk_blocks(in, ch)), P.segs) / 64);
  for (int w = threadIdx.x >> 6; w < P.segs; w += nwv) {
    const Segment sg = segment_of(in.counter_base, ch.slot_lo, ch.slot_hi, w, P.segs);
    const int64_t seg = int64_t(blockIdx.x) * P.segs + w;
    const int64_t src = in.out_off + int64_t(sg.p_lo - in.counter_base);
    const int64_t dst = uniform_i64(P.chunk_row[blockIdx.x] + P.seg_row[seg]);
    const int first_drop = uniform_i32(P.seg_drop[seg]);
    const uint32_t off0 = uint32_t(sg.p_lo - 4 * sg.qa);
    const uint32_t span = uint32_t(sg.p_hi - 4 * sg.qa);
    // rows of the iterations before the first drop: every live slot was kept
    const uint64_t lim = uint64_t(256) * uint64_t(first_drop);
    const int64_t end = int64_t(lim < span ? lim : span);
    const int64_t straight = end > int64_t(off0) ? end - int64_t(off0) : 0;
    const uint64_t* mask = P.keepmask + ch.mask_off + int64_t(w) * per_wave * 4;
    pack_segment<Real>(static_cast<const Real*>(P.events) + src, packed, dst, straight, mask,
                       first_drop, int(sg.iters), off0);
    if (packed_bins && P.bins)
      pack_segment<int32_t>(P.bins + src, packed_bins, dst, straight, mask, first_drop,
                            int(sg.iters), off0);
  }
}

// ---------------------------------------------------------------------------
// Tail of the VJP for one instance, after its histogram (h0 = sum of
// cotangents, h1 = sum of cotangent * (u - C_j), per bin j) is complete in
// LDS or in the chunk's partial sums.  gC (cpp/loits_core.cpp:294-301 summed
// per bin), the CDF adjoint (:355-365) and the density adjoint (:416-431).
// Nodes go right to left in tiles of kTailPer * blockDim.x, kTailPer
// consecutive nodes per thread, with the suffix sum carried from tile to tile;
// each node's rho adjoint replaces its h0 (read before by the same thread
// only), and a second pass writes the density gradient.  Registers do not grow
// with the node count.
constexpr int kTailPer = 4;

template <typename Real>
__device__ void vjp_tail(const BatchedParams& P, const InstanceDesc& in, const double* cdf,
                         double inv_n, double* h0, const double* h1, Real* grad_density,
                         double* scratch) {
  const int64_t n = in.n_nodes;
  const Real* x = static_cast<const Real*>(P.x) + in.node_off;
  const Real* s = static_cast<const Real*>(P.density) + in.node_off;
  const double acc = in.accept ? 1.0 : 0.0;
  const int64_t tile = int64_t(kTailPer) * blockDim.x;
  const int64_t mine = int64_t(threadIdx.x) * kTailPer;
  double carry = 0.0;  // sum of gC over the tiles to the right
  double local_norm = 0.0;
  for (int64_t base = ((n - 1) / tile) * tile; base >= 0; base -= tile) {
    const int64_t t0 = base + mine;
    double gc[kTailPer];
#pragma unroll
    for (int k = 0; k < kTailPer; ++k) {
      const int64_t t = t0 + k;
      double g = 0.0;
      if (t < n) {
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
    for (int k = kTailPer - 1; k >= 0; --k) {
      seg += gc[k];
      gc[k] = seg;  // sum over k' >= k within the thread's nodes
    }
    double total;
    const double excl = block_exclusive_scan(seg, scratch, &total);
    const double after = carry + (total - excl - seg);  // nodes right of this thread's
#pragma unroll
    for (int k = 0; k < kTailPer; ++k) {
      const int64_t t = t0 + k;
      if (t >= n) continue;
      const double S_t = gc[k] + after;
      const double S_next = k + 1 < kTailPer ? gc[k + 1] + after : after;
      double g = 0.0;
      if (t >= 1) g += 0.5 * (double(x[t]) - double(x[t - 1])) * acc * S_t;
      if (t + 1 < n) g += 0.5 * (double(x[t + 1]) - double(x[t])) * acc * S_next;
      local_norm += g * double(s[t]);
      h0[t] = g;
    }
    carry += total;
  }
  const double grad_norm = -block_sum(local_norm, scratch) * inv_n * inv_n;
  Real* out = grad_density + in.node_off;
  for (int64_t base = 0; base < n; base += tile) {
#pragma unroll
    for (int k = 0; k < kTailPer; ++k) {
      const int64_t t = base + mine + k;
      if (t >= n) continue;
      double wgt = 0.0;
      if (t >= 1) wgt += 0.5 * (double(x[t]) - double(x[t - 1]));
      if (t + 1 < n) wgt += 0.5 * (double(x[t + 1]) - double(x[t]));
      out[t] = static_cast<Real>(h0[t] * inv_n + grad_norm * wgt);
    }
  }
}

// A single-tile curve's share of one thread in the VJP tail (nodes t0 .. t0 + 3,
// t0 = kTailPer * threadIdx.x), loaded into registers right after the VJP's
// main loop so that the loads overlap the conversion and its barrier: nodes
// and CDF t0 - 1 .. t0 + 4 and the densities t0 .. t0 + 3 (zero past the ends).
template <typename Real>
struct TailNodes {
  double x[kTailPer + 2], c[kTailPer + 2], s[kTailPer];

  __device__ __forceinline__ void load(const BatchedParams& P, const InstanceDesc& in,
                                       const double* cdf) {
    const int64_t n = in.n_nodes, t0 = int64_t(threadIdx.x) * kTailPer;
    const Real* xg = static_cast<const Real*>(P.x) + in.node_off;
    const Real* sg = static_cast<const Real*>(P.density) + in.node_off;
#pragma unroll
    for (int q = 0; q < kTailPer + 2; ++q) {
      const int64_t t = t0 - 1 + q;
      const bool ok = t >= 0 && t < n;
      x[q] = ok ? double(xg[t]) : 0.0;
      c[q] = ok ? cdf[t] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < kTailPer; ++q) s[q] = t0 + q < n ? double(sg[t0 + q]) : 0.0;
  }
};

// vjp_tail for a curve of at most kTailPer * blockDim.x nodes from the values
// of TailNodes: the same operations in the same order (so the same bits), with
// each bin's 1 / (C_{j+1} - C_j + eps) computed once and the rho adjoints kept
// in registers.
template <typename Real>
__device__ void vjp_tail_regs(const BatchedParams& P, const InstanceDesc& in,
                              const TailNodes<Real>& tn, double inv_n, const double* h0,
                              const double* h1, Real* grad_density, double* scratch) {
  const int64_t n = in.n_nodes;
  const double acc = in.accept ? 1.0 : 0.0;
  const int64_t t0 = int64_t(threadIdx.x) * kTailPer;
  double inv_d[kTailPer + 1];  // bins t0 - 1 .. t0 + kTailPer - 1
#pragma unroll
  for (int q = 0; q <= kTailPer; ++q) {
    const int64_t j = t0 - 1 + q;
    inv_d[q] = j >= 0 && j + 1 < n ? 1.0 / (tn.c[q + 1] - tn.c[q] + kEpsilon) : 0.0;
  }
  double gc[kTailPer];
#pragma unroll
  for (int k = 0; k < kTailPer; ++k) {
    const int64_t t = t0 + k;
    double g = 0.0;
    if (t < n) {
      if (t + 1 < n) {
        const double db = tn.x[k + 2] - tn.x[k + 1];
        g += db * (inv_d[k + 1] * inv_d[k + 1] * h1[t] - inv_d[k + 1] * h0[t]);
      }
      if (t >= 1) {
        const double db = tn.x[k + 1] - tn.x[k];
        g -= db * inv_d[k] * inv_d[k] * h1[t - 1];
      }
    }
    gc[k] = g;
  }
  double seg = 0.0;
#pragma unroll
  for (int k = kTailPer - 1; k >= 0; --k) {
    seg += gc[k];
    gc[k] = seg;
  }
  double total;
  const double excl = block_exclusive_scan(seg, scratch, &total);
  const double carry = 0.0;
  const double after = carry + (total - excl - seg);
  double rho[kTailPer], local_norm = 0.0;
#pragma unroll
  for (int k = 0; k < kTailPer; ++k) {
    const int64_t t = t0 + k;
    rho[k] = 0.0;
    if (t >= n) continue;
    const double S_t = gc[k] + after;
    const double S_next = k + 1 < kTailPer ? gc[k + 1] + after : after;
    double g = 0.0;
    if (t >= 1) g += 0.5 * (tn.x[k + 1] - tn.x[k]) * acc * S_t;
    if (t + 1 < n) g += 0.5 * (tn.x[k + 2] - tn.x[k + 1]) * acc * S_next;
    local_norm += g * tn.s[k];
    rho[k] = g;
  }
  const double grad_norm = -block_sum(local_norm, scratch) * inv_n * inv_n;
  Real* out = grad_density + in.node_off;
#pragma unroll
  for (int k = 0; k < kTailPer; ++k) {
    const int64_t t = t0 + k;
    if (t >= n) continue;
    double wgt = 0.0;
    if (t >= 1) wgt += 0.5 * (tn.x[k + 1] - tn.x[k]);
    if (t + 1 < n) wgt += 0.5 * (tn.x[k + 2] - tn.x[k + 1]);
    out[t] = static_cast<Real>(rho[k] * inv_n + grad_norm * wgt);
  }
}

// Per-bin data of the fixed-point VJP, {alpha_j, beta_j}: t / (tau_j d_j) =
// float(k - T_j) * alpha_j + beta_j lies in [0, 1) for every key of bin j of a
// monotone curve (tau_j = 1 except for the last bin, which also takes the keys
// above T_{n-1}); T_j is the sorted threshold array itself.

// |g| bits of cotangent row r of a segment (0 past its kept rows): bit
// patterns order |g|, NaN above inf.
template <typename Real>
__device__ __forceinline__ uint32_t grad_bits(const Real* grad, int64_t kept, int64_t srow0,
                                              int64_t r) {
  return r < kept ? __builtin_bit_cast(uint32_t, float(grad[srow0 + r])) & 0x7FFFFFFFu : 0u;
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    const uint32_t o = uint32_t(__shfl_xor(int(v), d, kWave));
    v = o > v ? o : v;
  }
  return v;
}

// The fixed-point scale's sample of four of a wave's segments (s0, s0 + nwv,
// s0 + 2 nwv, s0 + 3 nwv): loaded before the work-group's set-up so that its
// descriptor and first-row loads overlap the set-up's copies.
struct ScaleProbe {
  int64_t kept[4], srow0[4];
  uint32_t bits[4];

  // Rows of four segments, then (load_bits) the first cotangent of each at this
  // lane; split so that other loads can be issued in between.
  __device__ __forceinline__ void load_rows(const BatchedParams& P, int c, int s0, int nwv) {
    const int nw = P.segs;
    const int64_t crow = P.chunk_row[c];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int sw = s0 + q * nwv;
      const int64_t sseg = int64_t(c) * nw + sw;
      kept[q] = sw < nw ? P.seg_kept[sseg] : 0;
      srow0[q] = sw < nw ? crow + P.seg_row[sseg] : 0;
    }
  }

  template <typename Real>
  __device__ __forceinline__ void load_bits(const Real* grad, int lane) {
#pragma unroll
    for (int q = 0; q < 4; ++q) bits[q] = grad_bits(grad, kept[q], srow0[q], lane);
  }

  template <typename Real>
  __device__ __forceinline__ void load(const BatchedParams& P, const Real* grad, int c, int s0,
                                       int nwv, int lane) {
    load_rows(P, c, s0, nwv);
    load_bits(grad, lane);
  }

  // Largest |g| bits over the four segments, each from its first 64-row block
  // holding a non-zero cotangent.
  template <typename Real>
  __device__ __forceinline__ uint32_t max(const Real* grad, int lane) const {
    uint32_t m_all = 0u;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      uint32_t m = wave_max_u32(bits[q]);
      for (int64_t r0 = kWave; r0 < kept[q] && m == 0u; r0 += kWave)
        m = wave_max_u32(grad_bits(grad, kept[q], srow0[q], r0 + lane));
      m_all = m > m_all ? m : m_all;
    }
    return m_all;
  }
};

__device__ __forceinline__ double vjp_tau_d_of(double cj, double cj1, bool last) {
  const double d = cj1 - cj + kEpsilon;
  if (!last) return d;
  const double tail = (1.0 - cj) / d;
  return (tail > 1.0 ? tail : 1.0) * d;
}

__device__ __forceinline__ double vjp_tau_d(const double* cdf, int64_t j, int64_t n) {
  return vjp_tau_d_of(cdf[j], cdf[j + 1], j == n - 2);
}

// {alpha_j, beta_j} from C_j, C_{j+1} and T_j.
__device__ __forceinline__ float2 vjp_alpha_beta(double cj, double cj1, uint32_t tj, bool last) {
  const double tau_d = vjp_tau_d_of(cj, cj1, last);
  const double r = double(tj) * 0x1.0p-24 - cj;
  return make_float2(static_cast<float>(0x1.0p-24 / tau_d), static_cast<float>(r / tau_d));
}

// VJP: one work-group per chunk, one segment per wave; the chunk that
// completes an instance runs the tail.
//
// float32 curves with monotone thresholds accumulate h0 = sum g and
// h1' = sum g t / (tau d) per bin as 64-bit integers (ds_add_u64, about half
// the LDS cost of ds_add_f64): g 2^e rounded to an integer, with 2^e chosen
// from the largest |g| the work-group sees first so that it maps below 2^24.
// A value that would reach 2^30 (64x that maximum) or is not finite goes to
// fp64 atomics in the chunk's global partial sums.  Other curves, and float64,
// accumulate fp64 directly (ds_add_f64).
// A load that skips the work-group's L1: values summed by device atomics at L2.
__device__ __forceinline__ double load_l2(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <bool GLOBAL>
__device__ __forceinline__ void add_f64(double* p, double v) {
  if constexpr (GLOBAL) unsafeAtomicAdd(p, v);
  else atomicAdd(p, v);
}

// GLOBAL (global-table instances): only the guide is staged in LDS; srt and
// coef are read from the blob and the fp64 histograms are the chunk's partial
// sums in global memory.
template <typename Real, bool GLOBAL>
__device__ __forceinline__ void vjp_body(const BatchedParams& P, const Real* grad,
                                         Real* grad_density, uint8_t* lds, int c) {
  using Coef = typename CoefOf<Real>::type;
  __shared__ double scratch[16];
  __shared__ int sh_last;
  __shared__ uint32_t sh_gmax;
  __shared__ int sh_outl;  // some cotangent took the fp64 overflow sums
  __shared__ int sh_next;  // next unclaimed segment
  const ChunkDesc ch = P.chunks[c];
  const InstanceDesc in = P.inst[ch.instance];
  Tables<Real> T = tables_of<Real>(P.tables, in);
  const int n = static_cast<int>(in.n_nodes);
  const int shift = 24 - in.guide_log2;
  const int tid = threadIdx.x, nt = blockDim.x;
  const bool fixed = !GLOBAL && sizeof(Real) == 4 && (T.hdr->flags & kMonotone) != 0;
  const int probe_wave = __builtin_amdgcn_readfirstlane(tid >> 6), probe_nwv = nt >> 6;
  ScaleProbe probe;
  if (fixed) probe.load(P, grad, c, probe_wave, probe_nwv, lane_id());

  // LDS: srt | guide | per-bin data (coef, or {alpha, beta} when fixed) | h0 | h1
  // (GLOBAL: guide only)
  const int64_t nb = pad4(n);
  double* outl = P.partials + ch.partial_off;  // [2 nb] fp64 overflow sums, then partials
  uint8_t* s_bin;
  double* h0;
  const uint32_t* s_srt;
  const uint32_t* s_guide;
  if constexpr (GLOBAL) {
    lds_copy(lds, T.guide, guide_bytes(T, in));
    s_bin = reinterpret_cast<uint8_t*>(T.coef);
    h0 = outl;
    s_srt = T.srt;
    s_guide = reinterpret_cast<const uint32_t*>(lds);
  } else {
    lds_copy(lds, T.srt, (T.L.guide - T.L.thr) + guide_bytes(T, in));  // srt | guide in use
    s_bin = lds + (T.L.coef - T.L.thr);
    h0 = reinterpret_cast<double*>(s_bin + nb * int64_t(sizeof(Coef)));
    s_srt = reinterpret_cast<const uint32_t*>(lds);
    s_guide = reinterpret_cast<const uint32_t*>(lds + (T.L.guide - T.L.thr));
  }
  double* h1 = h0 + nb;
  uint64_t* a0 = reinterpret_cast<uint64_t*>(h0);
  uint64_t* a1 = reinterpret_cast<uint64_t*>(h1);
  constexpr int kDepth = 4;  // bins per thread with loads in flight
  if (fixed) {
    float2* vb = reinterpret_cast<float2*>(s_bin);
    for (int64_t j0 = tid; j0 < n - 1; j0 += kDepth * nt) {
      double cj[kDepth], cj1[kDepth];
      uint32_t tj[kDepth];
#pragma unroll
      for (int k = 0; k < kDepth; ++k) {
        const int64_t j = j0 + k * nt, jj = j < n - 1 ? j : 0;
        cj[k] = T.cdf[jj];
        cj1[k] = T.cdf[jj + 1];
        tj[k] = T.srt[jj];
      }
#pragma unroll
      for (int k = 0; k < kDepth; ++k) {
        const int64_t j = j0 + k * nt;
        if (j < n - 1) vb[j] = vjp_alpha_beta(cj[k], cj1[k], tj[k], j == n - 2);
      }
    }
  } else if (!GLOBAL) {
    lds_copy(s_bin, T.coef, nb * int64_t(sizeof(Coef)));
  }
  for (int64_t i = tid; i < 2 * nb; i += nt) {
    h0[i] = 0.0;
    outl[i] = 0.0;
  }
  if (tid == 0) {
    sh_gmax = 0u;
    sh_outl = 0;
    sh_next = 0;
  }
  __syncthreads();
  const Coef* s_coef = reinterpret_cast<const Coef*>(s_bin);
  const float2* s_vb = reinterpret_cast<const float2*>(s_bin);

  const int nw = P.segs, lane = lane_id();

  // Fixed-point scale: the largest |g| over the first rows of each segment
  // holding a non-zero cotangent (bit patterns order |g|; NaN above inf).
  float scale = 0.0f;
  int scale_log2 = 0;
  if (fixed) {
    uint32_t m_all = probe.max(grad, lane);
    for (int s0 = probe_wave + 4 * probe_nwv; s0 < nw; s0 += 4 * probe_nwv) {
      probe.load(P, grad, c, s0, probe_nwv, lane);  // more than four segments per wave
      const uint32_t m = probe.max(grad, lane);
      m_all = m > m_all ? m : m_all;
    }
    if (lane == 0 && m_all != 0u) atomicMax(&sh_gmax, m_all);
    __syncthreads();
    const uint32_t gm = sh_gmax;
    const int ex = int(gm >> 23);  // |g| < 2^(ex - 126)
    if (ex == 255) {
      scale = __builtin_nanf("");  // inf / NaN cotangents: every value takes the fp64 path
    } else {
      scale_log2 = 24 - (ex - 126);
      scale_log2 = scale_log2 > 126 ? 126 : scale_log2;
      scale = __builtin_ldexpf(1.0f, scale_log2);
    }
  }

  // segments of the chunk, taken from the work-group's counter
  for (int w = next_segment(&sh_next); w < nw; w = next_segment(&sh_next)) {
    const Segment sg = segment_of(in.counter_base, ch.slot_lo, ch.slot_hi, w, nw);
    const int64_t seg = int64_t(c) * nw + w;
    const int first_drop = uniform_i32(P.seg_drop[seg]);
    const int64_t row0 = uniform_i64(P.chunk_row[c] + P.seg_row[seg]);
    const int64_t per_wave = int64_t(blocks_per_wave(uint64_t(chunk_blocks(in, ch)), nw) / 64);
    const uint64_t* mask = P.keepmask + ch.mask_off + int64_t(w) * per_wave * 4;
    const uint32_t nblk = uint32_t(sg.qb - sg.qa);
    const uint32_t off0 = uint32_t(sg.p_lo - 4 * sg.qa);
    const uint32_t span = uint32_t(sg.p_hi - 4 * sg.qa);
    // kept slots before the first masked iteration: every live slot before it
    int64_t running = int64_t(4 * 64 * int64_t(first_drop)) - int64_t(off0);
    running = running < 0 ? 0 : running;

    // one kept slot: bin j, cotangent g, key k
    auto add_fixed = [&](int jj, float g, float th) {
      const float gs = g * scale;
      if (__builtin_expect(__builtin_fabsf(gs) < 0x1.0p30f, 1)) {
        const int ia = static_cast<int>(__builtin_rintf(gs));
        const int ib = static_cast<int>(__builtin_rintf(gs * th));
        atomicAdd(reinterpret_cast<unsigned long long*>(a0 + jj),
                  static_cast<unsigned long long>(static_cast<int64_t>(ia)));
        atomicAdd(reinterpret_cast<unsigned long long*>(a1 + jj),
                  static_cast<unsigned long long>(static_cast<int64_t>(ib)));
      } else {
        sh_outl = 1;
        atomicAdd(outl + jj, double(g));
        atomicAdd(outl + nb + jj, double(g) * double(th));
      }
    };

    // keep bits and first packed row of a lane's four slots in iteration `it`
    auto rows_of = [&](int it, uint32_t i0, int& keepbits, int64_t& row) {
      keepbits = kept_slots(mask, it, first_drop, i0, nblk, off0, span, running, &row);
      row += row0;
    };

    // kTrip iterations per trip (independent Philox / search chains).
    int it = 0;
    for (uint32_t i0 = 0; i0 < nblk; i0 += 64 * kTrip, it += kTrip) {
      uint32_t k[4 * kTrip];
  #pragma unroll
      for (int h = 0; h < kTrip; ++h) {
        const Words4 wd =
            philox4x32_10_uniform(sg.qa + i0 + 64 * h + uint32_t(lane), in.stream, P.seed);
        k[4 * h] = wd.w0 >> 8;
        k[4 * h + 1] = wd.w1 >> 8;
        k[4 * h + 2] = wd.w2 >> 8;
        k[4 * h + 3] = wd.w3 >> 8;
      }
      int kb[kTrip];
      int64_t rw[kTrip];
  #pragma unroll
      for (int h = 0; h < kTrip; ++h) {
        kb[h] = 0;
        rw[h] = 0;
        if (h == 0 || i0 + 64 * h < nblk) rows_of(it + h, i0 + 64 * h, kb[h], rw[h]);
      }
      int j[4 * kTrip];
      find_bins<4 * kTrip>(k, s_guide, s_srt, shift, n, j);
      if (fixed) {
        float th[4 * kTrip];
  #pragma unroll
        for (int e = 0; e < 4 * kTrip; ++e) {
          const float2 ab = s_vb[j[e]];
          th[e] = __builtin_fmaf(static_cast<float>(static_cast<int32_t>(k[e] - s_srt[j[e]])), ab.x,
                                 ab.y);
        }
        float g[4 * kTrip];
  #pragma unroll
        for (int h = 0; h < kTrip; ++h) {
          if (kb[h] == 0xF && (rw[h] & 3) == 0) {
            const float4 g4 = *reinterpret_cast<const float4*>(grad + rw[h]);
            g[4 * h] = g4.x; g[4 * h + 1] = g4.y; g[4 * h + 2] = g4.z; g[4 * h + 3] = g4.w;
          } else {
            int64_t r = rw[h];
  #pragma unroll
            for (int e = 0; e < 4; ++e) g[4 * h + e] = ((kb[h] >> e) & 1) ? float(grad[r++]) : 0.0f;
          }
        }
  #pragma unroll
        for (int h = 0; h < kTrip; ++h) {
          if (kb[h] == 0xF) {
  #pragma unroll
            for (int e = 0; e < 4; ++e) add_fixed(j[4 * h + e], g[4 * h + e], th[4 * h + e]);
          } else {
  #pragma unroll
            for (int e = 0; e < 4; ++e)
              if ((kb[h] >> e) & 1) add_fixed(j[4 * h + e], g[4 * h + e], th[4 * h + e]);
          }
        }
      } else {
  #pragma unroll
        for (int h = 0; h < kTrip; ++h) {
          double t[4];
  #pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int je = j[4 * h + e];
            if constexpr (sizeof(Real) == 4) {
              uint32_t tj;
              float rj;
              if constexpr (GLOBAL) {  // global tables: {M, A, T, r} records
                const float4 q = reinterpret_cast<const float4*>(s_bin)[je];
                tj = __float_as_uint(q.z);
                rj = q.w;
              } else {
                const CoefF32View cv = coef_f32_view(const_cast<uint8_t*>(s_bin), nb);
                tj = cv.thr[je];
                rj = cv.r[je];
              }
              t[e] = double(static_cast<float>(static_cast<int32_t>(k[4 * h + e] - tj)) *
                                0x1.0p-24f +
                            rj);
            } else {
              t[e] = double(k[4 * h + e]) * 0x1.0p-24 - s_coef[je].c;
            }
          }
          int64_t r = rw[h];
  #pragma unroll
          for (int e = 0; e < 4; ++e) {
            if (!((kb[h] >> e) & 1)) continue;
            const double ge = double(grad[r++]);
            add_f64<GLOBAL>(&h0[j[4 * h + e]], ge);
            add_f64<GLOBAL>(&h1[j[4 * h + e]], ge * t[e]);
          }
        }
      }
    }
  }
  __syncthreads();
  if (GLOBAL) {  // the histograms were summed by device atomics: drop stale L1 lines
    if (tid == 0) __threadfence();
    __syncthreads();
  }

  if (fixed) {  // integers (and overflow sums) -> h0 = sum g, h1 = sum g t, in place
    const double inv_scale = __builtin_ldexp(1.0, -scale_log2);
    const bool any_outl = sh_outl != 0;
    for (int64_t i0 = tid; i0 < nb; i0 += kDepth * nt) {
      double td[kDepth];
#pragma unroll
      for (int k = 0; k < kDepth; ++k) {
        const int64_t i = i0 + k * nt, ii = i < n - 1 ? i : 0;
        td[k] = vjp_tau_d_of(T.cdf[ii], T.cdf[ii + 1], i == n - 2);
      }
#pragma unroll
      for (int k = 0; k < kDepth; ++k) {
        const int64_t i = i0 + k * nt;
        if (i >= nb) continue;
        double v0 = 0.0, v1 = 0.0;
        if (i < n - 1) {
          v0 = double(static_cast<int64_t>(a0[i])) * inv_scale +
               (any_outl ? load_l2(outl + i) : 0.0);
          v1 = (double(static_cast<int64_t>(a1[i])) * inv_scale +
                (any_outl ? load_l2(outl + nb + i) : 0.0)) *
               td[k];
        }
        h0[i] = v0;
        h1[i] = v1;
      }
    }
    __syncthreads();
  }

  if (in.n_chunks == 1) {
    vjp_tail<Real>(P, in, T.cdf, T.hdr->inv_n, h0, h1, grad_density, scratch);
    return;
  }
  // multi-chunk instance: publish partial sums; the last chunk reduces them in order
  if (!GLOBAL) {  // (GLOBAL: they are already there)
    double* part = P.partials + ch.partial_off;
    for (int64_t i = tid; i < n; i += nt) {
      part[i] = h0[i];
      part[nb + i] = h1[i];
    }
  }
  // one device-scope fence per work-group (each is an L2 write-back)
  __syncthreads();
  if (tid == 0) {
    __threadfence();
    const uint32_t old = atomicAdd(&P.done[ch.instance], 1u);
    sh_last = old == uint32_t(in.n_chunks - 1);
    if (sh_last) __threadfence();
  }
  __syncthreads();
  if (!sh_last) return;
  for (int64_t i = tid; i < n; i += nt) {  // each i read and written by one thread
    double a = 0.0, b = 0.0;
    for (int cc = in.first_chunk; cc < in.first_chunk + in.n_chunks; ++cc) {
      const double* pp = P.partials + P.chunks[cc].partial_off;
      a += __builtin_nontemporal_load(pp + i);
      b += __builtin_nontemporal_load(pp + nb + i);
    }
    h0[i] = a;
    h1[i] = b;
  }
  if (tid == 0) P.done[ch.instance] = 0u;
  __syncthreads();
  vjp_tail<Real>(P, in, T.cdf, T.hdr->inv_n, h0, h1, grad_density, scratch);
}

// VJP of a float32 curve with cell tables (fast_tables.hpp): one work-group per
// chunk.  The work-group stages the cell words and sorted splits, regenerates
// its keys and finds their bins as the forward did, and sums per bin
//   a0 = sum q,  a1 = sum q k,   q = round(g 2^e), |q| < 2^fixed_point_bits(keys)
// as 64-bit integers (exact, order-independent): with the bin's own threshold
// T_j, sum g (u - C_j) = 2^-24 sum g (k - T_j) + (T_j 2^-24 - C_j) sum g, and
// sum q (k - T_j) = a1 - T_j a0 exactly.  2^e is set from the largest |g| the
// work-group probes first, two bits below the limit; a cotangent beyond it (or
// not finite) goes to float64 sums of g and g (k - T_j) in the chunk's
// partials.  Then the interpolation adjoint per bin, the CDF and density
// adjoints (vjp_tail), in the last chunk of the instance to finish.
__device__ void fast_vjp_body(const BatchedParams& P, const float* grad, float* grad_density,
                              uint8_t* lds, int c) {
  __shared__ double scratch[16];
  __shared__ int sh_last;
  __shared__ uint32_t sh_gmax;
  __shared__ int sh_outl;
  __shared__ int sh_next;
  const ChunkDesc ch = P.chunks[c];
  const InstanceDesc in = P.inst[ch.instance];
  const int n = static_cast<int>(in.n_nodes);
  const int g = in.fast_log2, s = 24 - g;
  const int qbits = fixed_point_bits(ch.slot_hi - ch.slot_lo);
  const int64_t G = int64_t(1) << g;
  const uint8_t* blob = P.tables + in.table_off;
  const TableHeader* hdr = reinterpret_cast<const TableHeader*>(blob);
  const FastBlobLayout B = fast_blob_layout(n, g);
  const FastLdsVjp L = fast_lds_vjp(n, g);
  const double* cdf = reinterpret_cast<const double*>(blob + B.cdf);
  const int tid = threadIdx.x, nt = blockDim.x;
  const int probe_wave = __builtin_amdgcn_readfirstlane(tid >> 6), probe_nwv = nt >> 6;
  ScaleProbe probe;
  probe.load_rows(P, c, probe_wave, probe_nwv);

  uint32_t* cells = reinterpret_cast<uint32_t*>(lds + L.cells);
  uint32_t* srt = reinterpret_cast<uint32_t*>(lds + L.srt);
  uint64_t* a0 = reinterpret_cast<uint64_t*>(lds + L.a0);
  uint64_t* a1 = a0 + kFastMaxNodes;  // one address, an immediate offset for the second add
  uint64_t* a0m = a0 - 1;
  uint64_t* a1m = a1 - 1;
  const uint32_t* l2 = cells + G;
  const int64_t nb = pad4(n);
  double* outl = P.partials + ch.partial_off;  // [2 nb]: float64 sums, then partials
  // the sorted splits and every cell word (level 2 at its capacity, so that the
  // copy does not wait for the header) in one contiguous copy
  lds_copy(srt, blob + B.srt, B.own - B.srt);
  probe.load_bits(grad, lane_id());
  for (int64_t i = tid; i < nb; i += nt) {
    a0[i] = 0ull;
    a1[i] = 0ull;
  }
  for (int64_t i = tid; i < 2 * nb; i += nt) outl[i] = 0.0;
  if (tid == 0) {
    sh_gmax = 0u;
    sh_outl = 0;
    sh_next = 0;
  }
  const bool mono = (hdr->flags & kMonotone) != 0;
  const uint32_t* own = mono ? srt : reinterpret_cast<const uint32_t*>(blob + B.own);
  const int nw = P.segs, lane = lane_id();
  {
    uint32_t m_all = probe.max(grad, lane);
    for (int s0 = probe_wave + 4 * probe_nwv; s0 < nw; s0 += 4 * probe_nwv) {
      probe.load(P, grad, c, s0, probe_nwv, lane);
      const uint32_t m = probe.max(grad, lane);
      m_all = m > m_all ? m : m_all;
    }
    __syncthreads();  // the counters above are initialised
    if (lane == 0 && m_all != 0u) atomicMax(&sh_gmax, m_all);
  }
  __syncthreads();
  float scale;
  int scale_log2 = 0;
  {
    const uint32_t gm = sh_gmax;
    const int ex = int(gm >> 23);  // |g| < 2^(ex - 126)
    if (ex == 255) {
      scale = __builtin_nanf("");  // inf / NaN cotangents: every value takes the float64 sums
    } else {
      scale_log2 = (qbits - 2) - (ex - 126);
      scale_log2 = scale_log2 > 126 ? 126 : scale_log2;
      scale = __builtin_ldexpf(1.0f, scale_log2);
    }
  }
  const float kQLimit = float(1 << qbits);

  for (int w = next_segment(&sh_next); w < nw; w = next_segment(&sh_next)) {
    const Segment sg = segment_of(in.counter_base, ch.slot_lo, ch.slot_hi, w, nw);
    const int64_t seg = int64_t(c) * nw + w;
    const int first_drop = uniform_i32(P.seg_drop[seg]);
    const int64_t row0 = uniform_i64(P.chunk_row[c] + P.seg_row[seg]);
    const int64_t per_wave = int64_t(blocks_per_wave(uint64_t(chunk_blocks(in, ch)), nw) / 64);
    const uint64_t* mask = P.keepmask + ch.mask_off + int64_t(w) * per_wave * 4;
    const uint32_t nblk = uint32_t(sg.qb - sg.qa);
    const uint32_t off0 = uint32_t(sg.p_lo - 4 * sg.qa);
    const uint32_t span = uint32_t(sg.p_hi - 4 * sg.qa);
    int64_t running = int64_t(4 * 64 * int64_t(first_drop)) - int64_t(off0);
    running = running < 0 ? 0 : running;

    // jj = j + 1 (cell_bins): the arrays are addressed from one entry below
    auto add = [&](int jj, uint32_t k, float gv) {
      const float gs = gv * scale;
      if (__builtin_expect(__builtin_fabsf(gs) < kQLimit, 1)) {
        const int q = static_cast<int>(__builtin_rintf(gs));
        atomicAdd(reinterpret_cast<unsigned long long*>(a0m + jj),
                  static_cast<unsigned long long>(static_cast<int64_t>(q)));
        atomicAdd(reinterpret_cast<unsigned long long*>(a1m + jj),
                  static_cast<unsigned long long>(static_cast<int64_t>(q) * int64_t(k)));
      } else {
        sh_outl = 1;
        unsafeAtomicAdd(outl + (jj - 1), double(gv));
        unsafeAtomicAdd(outl + nb + (jj - 1),
                        double(gv) * double(static_cast<int32_t>(k - own[jj - 1])));
      }
    };

    if (first_drop == int(sg.iters)) {
      // every slot of the segment was kept: the cotangent of slot 4 i + e is row
      // base + 4 i + e; whole aligned blocks load as one 16-byte read
      const int64_t base = row0 - int64_t(off0);
      const bool vec = off0 == 0u && (base & 3) == 0;
      const uint32_t nfull = vec ? span / 4 : 0u;  // blocks whose four rows all exist
      for (uint32_t i0 = 0; i0 < nblk; i0 += 64 * kTrip) {
        uint32_t k[4 * kTrip];
#pragma unroll
        for (int h = 0; h < kTrip; ++h) {
          const Words4 wd =
              philox4x32_10_uniform(sg.qa + i0 + 64 * h + uint32_t(lane), in.stream, P.seed);
          k[4 * h] = wd.w0 >> 8;
          k[4 * h + 1] = wd.w1 >> 8;
          k[4 * h + 2] = wd.w2 >> 8;
          k[4 * h + 3] = wd.w3 >> 8;
        }
        float gv[4 * kTrip];
        int live[kTrip];
        const bool whole = i0 + 64 * kTrip <= nfull;  // wave-uniform: no per-lane bounds
#pragma unroll
        for (int h = 0; h < kTrip; ++h) {
          const uint32_t i = i0 + 64 * h + uint32_t(lane);
          if (whole || (vec && 4 * i + 4 <= span)) {
            const float4 g4 = load_stream4(grad + base + 4 * int64_t(i));
            gv[4 * h] = g4.x;
            gv[4 * h + 1] = g4.y;
            gv[4 * h + 2] = g4.z;
            gv[4 * h + 3] = g4.w;
            live[h] = 0xF;
          } else {
            live[h] = 0;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              const uint32_t r = 4 * i + e;
              const bool ok = i < nblk && r >= off0 && r < span;
              live[h] |= ok ? (1 << e) : 0;
              gv[4 * h + e] = ok ? grad[base + int64_t(r)] : 0.0f;
            }
          }
        }
        int j[4 * kTrip];
        cell_bins<4 * kTrip>(k, cells, l2, srt, s, j);
        if (whole) {
#pragma unroll
          for (int e = 0; e < 4 * kTrip; ++e) add(j[e], k[e], gv[e]);
          continue;
        }
#pragma unroll
        for (int h = 0; h < kTrip; ++h) {
          if (live[h] == 0xF) {
#pragma unroll
            for (int e = 0; e < 4; ++e) add(j[4 * h + e], k[4 * h + e], gv[4 * h + e]);
          } else {
#pragma unroll
            for (int e = 0; e < 4; ++e)
              if ((live[h] >> e) & 1) add(j[4 * h + e], k[4 * h + e], gv[4 * h + e]);
          }
        }
      }
      continue;
    }

    // Cotangents of the kept slots (0 for the others, whose adds are then
    // zero): no branch on the keep bits around the histogram updates.
    const float* gseg = grad + row0;
    int it = 0;
    for (uint32_t i0 = 0; i0 < nblk; i0 += 64 * kTrip, it += kTrip) {
      uint32_t k[4 * kTrip];
#pragma unroll
      for (int h = 0; h < kTrip; ++h) {
        const Words4 wd =
            philox4x32_10_uniform(sg.qa + i0 + 64 * h + uint32_t(lane), in.stream, P.seed);
        k[4 * h] = wd.w0 >> 8;
        k[4 * h + 1] = wd.w1 >> 8;
        k[4 * h + 2] = wd.w2 >> 8;
        k[4 * h + 3] = wd.w3 >> 8;
      }
      float gv[4 * kTrip];
#pragma unroll
      for (int h = 0; h < kTrip; ++h) {
        int kb = 0;
        int64_t rw = 0;
        if (h == 0 || i0 + 64 * h < nblk)
          kb = kept_slots(mask, it + h, first_drop, i0 + 64 * h, nblk, off0, span, running, &rw);
        if (kb == 0xF && ((row0 + rw) & 3) == 0) {
          const float4 g4 = load_stream4(gseg + rw);
          gv[4 * h] = g4.x;
          gv[4 * h + 1] = g4.y;
          gv[4 * h + 2] = g4.z;
          gv[4 * h + 3] = g4.w;
        } else {
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            gv[4 * h + e] = ((kb >> e) & 1) ? gseg[rw] : 0.0f;
            rw += (kb >> e) & 1;
          }
        }
      }
      int j[4 * kTrip];
      cell_bins<4 * kTrip>(k, cells, l2, srt, s, j);
#pragma unroll
      for (int e = 0; e < 4 * kTrip; ++e) add(j[e], k[e], gv[e]);
    }
  }
  // a single-chunk curve of at most kTailPer * nt nodes: the tail's nodes,
  // densities and CDF values load now, under the barrier and the conversion
  const bool regs_tail = in.n_chunks == 1 && int64_t(n) <= int64_t(kTailPer) * nt;
  TailNodes<float> tn;
  if (regs_tail) tn.load(P, in, cdf);
  __syncthreads();

  // integers (and float64 sums) -> h0 = sum g, h1 = sum g (u - C_j), in place
  double* h0 = reinterpret_cast<double*>(a0);
  double* h1 = reinterpret_cast<double*>(a1);
  {
    const double inv_scale = __builtin_ldexp(1.0, -scale_log2);
    const bool any_outl = sh_outl != 0;
    auto convert = [&](int64_t i, double C) {
      double v0 = 0.0, v1 = 0.0;
      if (i < n - 1) {
        const uint64_t A0 = a0[i], A1 = a1[i];
        const uint32_t T = own[i];
        const int64_t d = static_cast<int64_t>(A1 - uint64_t(T) * A0);  // sum q (k - T_j)
        v0 = double(static_cast<int64_t>(A0)) * inv_scale + (any_outl ? load_l2(outl + i) : 0.0);
        v1 = (double(d) * inv_scale + (any_outl ? load_l2(outl + nb + i) : 0.0)) * 0x1.0p-24 +
             (double(T) * 0x1.0p-24 - C) * v0;
      }
      h0[i] = v0;
      h1[i] = v1;
    };
    if (regs_tail) {  // this thread's own bins (nb <= kTailPer * nt), C from registers
      const int64_t t0 = int64_t(tid) * kTailPer;
#pragma unroll
      for (int k = 0; k < kTailPer; ++k)
        if (t0 + k < nb) convert(t0 + k, tn.c[k + 1]);
    } else {
      for (int64_t i = tid; i < nb; i += nt) convert(i, i < n - 1 ? cdf[i] : 0.0);
    }
  }
  __syncthreads();
  if (regs_tail) {
    vjp_tail_regs<float>(P, in, tn, hdr->inv_n, h0, h1, grad_density, scratch);
    return;
  }
  if (in.n_chunks == 1) {
    vjp_tail<float>(P, in, cdf, hdr->inv_n, h0, h1, grad_density, scratch);
    return;
  }
  // multi-chunk instance: publish partial sums; the last chunk reduces them in order
  {
    double* part = P.partials + ch.partial_off;
    for (int64_t i = tid; i < n; i += nt) {
      part[i] = h0[i];
      part[nb + i] = h1[i];
    }
  }
  __syncthreads();
  if (tid == 0) {
    __threadfence();
    const uint32_t old = atomicAdd(&P.done[ch.instance], 1u);
    sh_last = old == uint32_t(in.n_chunks - 1);
    if (sh_last) __threadfence();
  }
  __syncthreads();
  if (!sh_last) return;
  for (int64_t i = tid; i < n; i += nt) {
    double sa = 0.0, sb = 0.0;
    for (int cc = in.first_chunk; cc < in.first_chunk + in.n_chunks; ++cc) {
      const double* pp = P.partials + P.chunks[cc].partial_off;
      sa += __builtin_nontemporal_load(pp + i);
      sb += __builtin_nontemporal_load(pp + nb + i);
    }
    h0[i] = sa;
    h1[i] = sb;
  }
  if (tid == 0) P.done[ch.instance] = 0u;
  __syncthreads();
  vjp_tail<float>(P, in, cdf, hdr->inv_n, h0, h1, grad_density, scratch);
}

// Exclusive prefix maximum over the work-group (wmax: 16 words of LDS).
__device__ uint32_t block_exclusive_max(uint32_t v, uint32_t* wmax) {
  const int lane = lane_id(), w = threadIdx.x >> 6;
  uint32_t inc = v;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const uint32_t o = uint32_t(__shfl_up(int(inc), d, kWave));
    if (lane >= d) inc = o > inc ? o : inc;
  }
  uint32_t ex = uint32_t(__shfl_up(int(inc), 1, kWave));
  if (lane == 0) ex = 0u;
  if (lane == kWave - 1) wmax[w] = inc;
  __syncthreads();
  for (int i = 0; i < w; ++i) ex = wmax[i] > ex ? wmax[i] : ex;
  __syncthreads();
  return ex;
}

// Bin-range VJP of a monotone curve whose VJP tables and histograms do not fit
// LDS.  Work item (chunk c, range q) owns the bins [lo, hi), i.e. the keys k in
// [K_lo, K_hi) (K_lo = T_lo, or 0 for the first range; K_hi = T_hi, or 2^24 for
// the last): it regenerates the chunk's keys, skips the keys outside its
// interval and sums the others into LDS histograms of its bins.  The bin of a
// key is the last t with T_t <= k, clamped to n - 2 (cpp/loits_core.cpp:12-18
// for a sorted CDF): a cell table over the interval gives the bins at the
// start of the key's cell and of the next cell, and a bisection between them
// (usually zero or one step) finishes it.  Partial sums go to the chunk's
// partial slots of these bins; the last item of the instance reduces them.
template <typename Real>
__device__ void vjp_range_body(const BatchedParams& P, const Real* grad, Real* grad_density,
                               uint8_t* lds, int c, int q) {
  __shared__ double scratch[16];
  __shared__ int sh_last;
  __shared__ uint32_t sh_gmax;
  __shared__ uint32_t sh_wmax[16];
  __shared__ int sh_outl;  // some cotangent took the fp64 overflow sums
  __shared__ int sh_next;  // next unclaimed segment
  constexpr bool fixed = sizeof(Real) == 4;
  const ChunkDesc ch = P.chunks[c];
  const InstanceDesc in = P.inst[ch.instance];
  Tables<Real> T = tables_of<Real>(P.tables, in);
  const int n = static_cast<int>(in.n_nodes);
  const int R = in.vjp_ranges;
  const int qbits = fixed_point_bits(ch.slot_hi - ch.slot_lo);
  const int W = vjp_range_width(n, R);
  const int lo = q * W;
  const int hi = lo + W < n - 1 ? lo + W : n - 1;
  const int nbins = hi - lo;  // >= 1
  const int tid = threadIdx.x, nt = blockDim.x;
  const int64_t nb = pad4(n);
  double* outl = P.partials + ch.partial_off;  // [2 nb]: overflow sums, then partial sums
  const int probe_wave = __builtin_amdgcn_readfirstlane(tid >> 6), probe_nwv = nt >> 6;
  ScaleProbe probe;
  if constexpr (fixed) probe.load(P, grad, c, probe_wave, probe_nwv, lane_id());

  const uint32_t k_lo = lo == 0 ? 0u : T.srt[lo];
  const uint32_t k_hi = hi == n - 1 ? (1u << 24) : T.srt[hi];
  const uint32_t span = k_hi > k_lo ? k_hi - k_lo : 0u;
  const int span_log2 = span <= 1u ? 0 : 32 - __builtin_clz(span - 1u);
  const int cs = span_log2 > kRangeCellsLog2 ? span_log2 - kRangeCellsLog2 : 0;  // cell 2^cs keys
  const int ncell = span == 0u ? 1 : int(((span - 1u) >> cs) + 1u);             // <= kRangeCells

  uint32_t* cell_bin = reinterpret_cast<uint32_t*>(lds);  // [kRangeCells + 4]
  uint32_t* thr = cell_bin + (kRangeCells + 4);           // [W + 4]: T_lo .. T_hi
  // float64: C_j per bin; float32 keeps none (it sums q and q k exactly)
  uint8_t* bd = reinterpret_cast<uint8_t*>(thr + (W + 4));
  uint64_t* a0 = reinterpret_cast<uint64_t*>(bd + (fixed ? 0 : int64_t(8) * W));
  uint64_t* a1 = a0 + W;
  double* d0 = reinterpret_cast<double*>(a0);
  double* d1 = reinterpret_cast<double*>(a1);

  for (int t = tid; t <= nbins; t += nt) thr[t] = T.srt[lo + t];
  if constexpr (!fixed)
    for (int i = tid; i < nbins; i += nt) reinterpret_cast<double*>(bd)[i] = T.cdf[lo + i];
  for (int i = tid; i < 2 * W; i += nt) a0[i] = 0u;
  for (int i = lo + tid; i < hi; i += nt) {
    outl[i] = 0.0;
    outl[nb + i] = 0.0;
  }
  if (hi == n - 1)
    for (int64_t i = hi + tid; i < nb; i += nt) {
      outl[i] = 0.0;
      outl[nb + i] = 0.0;
    }
  for (int i = tid; i <= ncell; i += nt) cell_bin[i] = 0u;
  if (tid == 0) {
    sh_gmax = 0u;
    sh_outl = 0;
    sh_next = 0;
  }
  __syncthreads();
  // Bin t (relative) starts the cells whose first key lies in [T_t, T_next);
  // those cells are disjoint across bins, so each run's first cell gets one
  // mark and a prefix maximum fills the rest.
  for (int t = tid; t <= nbins; t += nt) {
    const uint32_t a = thr[t];
    const uint32_t b = lo + t + 1 < n ? T.srt[lo + t + 1] : (1u << 24);
    const uint32_t ca = a <= k_lo ? 0u : uint32_t((uint64_t(a - k_lo) + (1u << cs) - 1u) >> cs);
    const uint32_t cb = b <= k_lo ? 0u : uint32_t((uint64_t(b - k_lo) + (1u << cs) - 1u) >> cs);
    if (ca < cb && ca < uint32_t(ncell)) cell_bin[ca] = uint32_t(t);
  }
  __syncthreads();
  {
    const int per = (ncell + nt - 1) / nt;
    const int c0 = tid * per;
    uint32_t m = 0u;
    for (int i = 0; i < per; ++i)
      if (c0 + i < ncell) m = cell_bin[c0 + i] > m ? cell_bin[c0 + i] : m;
    uint32_t run = block_exclusive_max(m, sh_wmax);
    for (int i = 0; i < per; ++i) {
      if (c0 + i >= ncell) break;
      run = cell_bin[c0 + i] > run ? cell_bin[c0 + i] : run;
      cell_bin[c0 + i] = run;
    }
    if (tid == 0) cell_bin[ncell] = uint32_t(nbins);
  }

  const int nw = P.segs, lane = lane_id();

  // Fixed-point scale from the first rows of each segment holding a non-zero
  // cotangent (as in the LDS-resident VJP).
  float scale = 0.0f;
  int scale_log2 = 0;
  if constexpr (fixed) {
    uint32_t m_all = probe.max(grad, lane);
    for (int s0 = probe_wave + 4 * probe_nwv; s0 < nw; s0 += 4 * probe_nwv) {
      probe.load(P, grad, c, s0, probe_nwv, lane);  // more than four segments per wave
      const uint32_t m = probe.max(grad, lane);
      m_all = m > m_all ? m : m_all;
    }
    if (lane == 0 && m_all != 0u) atomicMax(&sh_gmax, m_all);
  }
  __syncthreads();
  if constexpr (fixed) {
    const uint32_t gm = sh_gmax;
    const int ex = int(gm >> 23);
    if (ex == 255) {
      scale = __builtin_nanf("");  // inf / NaN cotangents: every value takes the float64 sums
    } else {  // |round(g 2^e)| < 2^qbits, as in the cell-table VJP
      scale_log2 = (qbits - 2) - (ex - 126);
      scale_log2 = scale_log2 > 126 ? 126 : scale_log2;
      scale = __builtin_ldexpf(1.0f, scale_log2);
    }
  }
  const float kQLimit = float(1 << qbits);
  const int jmax = n - 2 - lo;  // relative clamp (only the last range reaches it)
  const double* s_c = reinterpret_cast<const double*>(bd);

  // segments of the chunk, taken from the work-group's counter
  for (int w = next_segment(&sh_next); w < nw; w = next_segment(&sh_next)) {
    const Segment sg = segment_of(in.counter_base, ch.slot_lo, ch.slot_hi, w, nw);
    const int64_t seg = int64_t(c) * nw + w;
    const int first_drop = uniform_i32(P.seg_drop[seg]);
    const int64_t row0 = uniform_i64(P.chunk_row[c] + P.seg_row[seg]);
    const int64_t per_wave = int64_t(blocks_per_wave(uint64_t(chunk_blocks(in, ch)), nw) / 64);
    const uint64_t* mask = P.keepmask + ch.mask_off + int64_t(w) * per_wave * 4;
    const uint32_t nblk = uint32_t(sg.qb - sg.qa);
    const uint32_t off0 = uint32_t(sg.p_lo - 4 * sg.qa);
    const uint32_t span_s = uint32_t(sg.p_hi - 4 * sg.qa);
    int64_t running = int64_t(4 * 64 * int64_t(first_drop)) - int64_t(off0);
    running = running < 0 ? 0 : running;

    auto rows_of = [&](int it, uint32_t i0, int& keepbits, int64_t& row) {
      keepbits = kept_slots(mask, it, first_drop, i0, nblk, off0, span_s, running, &row);
      row += row0;
    };

    // Per trip, 4 * kTrip keys per lane go through each step together (their
    // LDS reads are independent): cell bounds, bisection, per-bin data, then the
    // cotangents and the histogram updates of the kept keys inside the range.
    constexpr int NK = 4 * kTrip;
    int it = 0;
    for (uint32_t i0 = 0; i0 < nblk; i0 += 64 * kTrip, it += kTrip) {
      uint32_t k[NK];
  #pragma unroll
      for (int h = 0; h < kTrip; ++h) {
        const Words4 wd =
            philox4x32_10_uniform(sg.qa + i0 + 64 * h + uint32_t(lane), in.stream, P.seed);
        k[4 * h] = wd.w0 >> 8;
        k[4 * h + 1] = wd.w1 >> 8;
        k[4 * h + 2] = wd.w2 >> 8;
        k[4 * h + 3] = wd.w3 >> 8;
      }
      int kb[kTrip];
      int64_t rw[kTrip];
  #pragma unroll
      for (int h = 0; h < kTrip; ++h) {
        kb[h] = 0;
        rw[h] = 0;
        if (h == 0 || i0 + 64 * h < nblk) rows_of(it + h, i0 + 64 * h, kb[h], rw[h]);
      }
      uint32_t use = 0u;  // kept and inside [K_lo, K_hi)
      int a[NK], b[NK];
  #pragma unroll
      for (int e = 0; e < NK; ++e) {
        const bool in_range = ((kb[e >> 2] >> (e & 3)) & 1) && k[e] >= k_lo && k[e] < k_hi;
        use |= in_range ? (1u << e) : 0u;
        const uint32_t cell = in_range ? (k[e] - k_lo) >> cs : 0u;
        a[e] = int(cell_bin[cell]);
        b[e] = int(cell_bin[cell + 1]);
      }
      uint32_t pending = 0u;
  #pragma unroll
      for (int e = 0; e < NK; ++e) pending |= (((use >> e) & 1u) && a[e] < b[e]) ? (1u << e) : 0u;
      while (pending) {  // last t in [a, b] with T_t <= k (T_a <= k holds); usually one step
  #pragma unroll
        for (int e = 0; e < NK; ++e) {
          if (!((pending >> e) & 1u)) continue;
          const int mid = (a[e] + b[e] + 1) >> 1;
          if (thr[mid] <= k[e]) a[e] = mid; else b[e] = mid - 1;
          if (a[e] >= b[e]) pending &= ~(1u << e);
        }
      }
      double tc[NK];
  #pragma unroll
      for (int e = 0; e < NK; ++e) {
        a[e] = a[e] > jmax ? jmax : a[e];
        if constexpr (!fixed) tc[e] = double(k[e]) * 0x1.0p-24 - s_c[a[e]];
      }
  #pragma unroll
      for (int h = 0; h < kTrip; ++h) {
        const uint32_t uh = (use >> (4 * h)) & 0xFu;
        if (uh == 0u) continue;
        Real g[4];
        if (R == 1 && kb[h] == 0xF && (rw[h] & 3) == 0) {
          if constexpr (fixed) {
            const float4 g4 = load_stream4(reinterpret_cast<const float*>(grad + rw[h]));
            g[0] = g4.x; g[1] = g4.y; g[2] = g4.z; g[3] = g4.w;
          } else {
            const double2 ga = reinterpret_cast<const double2*>(grad + rw[h])[0];
            const double2 gb = reinterpret_cast<const double2*>(grad + rw[h])[1];
            g[0] = ga.x; g[1] = ga.y; g[2] = gb.x; g[3] = gb.y;
          }
        } else {
          int64_t r = rw[h];
  #pragma unroll
          for (int e = 0; e < 4; ++e) {
            g[e] = ((uh >> e) & 1u) ? grad[r] : Real(0);
            r += (kb[h] >> e) & 1;
          }
        }
  #pragma unroll
        for (int e = 0; e < 4; ++e) {
          if (!((uh >> e) & 1u)) continue;
          const int jr = a[4 * h + e];
          if constexpr (fixed) {  // a0 = sum q, a1 = sum q k (exact, < 2^62)
            const float gf = float(g[e]);
            const float gs = gf * scale;
            const uint32_t ke = k[4 * h + e];
            if (__builtin_expect(__builtin_fabsf(gs) < kQLimit, 1)) {
              const int qv = static_cast<int>(__builtin_rintf(gs));
              atomicAdd(reinterpret_cast<unsigned long long*>(a0 + jr),
                        static_cast<unsigned long long>(static_cast<int64_t>(qv)));
              atomicAdd(reinterpret_cast<unsigned long long*>(a1 + jr),
                        static_cast<unsigned long long>(static_cast<int64_t>(qv) * int64_t(ke)));
            } else {
              sh_outl = 1;
              unsafeAtomicAdd(outl + lo + jr, double(gf));
              unsafeAtomicAdd(outl + nb + lo + jr,
                              double(gf) * double(static_cast<int32_t>(ke - thr[jr])));
            }
          } else {
            const double ge = double(g[e]);
            atomicAdd(d0 + jr, ge);
            atomicAdd(d1 + jr, ge * tc[4 * h + e]);
          }
        }
      }
    }
  }
  __syncthreads();

  // histograms -> sums (h0 = sum g, h1 = sum g t): in LDS when this item is
  // the instance's only one (its tail reads them there), else into the
  // chunk's partial sums of bins [lo, hi)
  const bool any_outl = sh_outl != 0;
  const uint32_t arrivals = uint32_t(in.n_chunks) * uint32_t(R);
  {
    const double inv_scale = __builtin_ldexp(1.0, -scale_log2);
    for (int i = tid; i < nbins; i += nt) {
      const int j = lo + i;
      double v0, v1;
      if constexpr (fixed) {  // sum g (u - C_j) = 2^-24 sum g (k - T_j) + (T_j 2^-24 - C_j) sum g
        const uint64_t A0 = a0[i], A1 = a1[i];
        const uint32_t Tj = thr[i];
        v0 = double(static_cast<int64_t>(A0)) * inv_scale + (any_outl ? load_l2(outl + j) : 0.0);
        v1 = (double(static_cast<int64_t>(A1 - uint64_t(Tj) * A0)) * inv_scale +
              (any_outl ? load_l2(outl + nb + j) : 0.0)) *
                 0x1.0p-24 +
             (double(Tj) * 0x1.0p-24 - T.cdf[j]) * v0;
      } else {
        v0 = d0[i];
        v1 = d1[i];
      }
      if (arrivals == 1u) {  // in place (each i read and written by one thread)
        d0[i] = v0;
        d1[i] = v1;
      } else {
        outl[j] = v0;
        outl[nb + j] = v1;
      }
    }
  }
  if (arrivals == 1u) {
    __syncthreads();
    vjp_tail<Real>(P, in, T.cdf, T.hdr->inv_n, d0, d1, grad_density, scratch);
    return;
  }
  // one device-scope fence per work-group; the last of the instance's
  // n_chunks * R items sums the chunks' partials in chunk order
  __syncthreads();
  if (tid == 0) {
    __threadfence();
    const uint32_t old = atomicAdd(&P.done[ch.instance], 1u);
    sh_last = old == arrivals - 1u;
    if (sh_last) __threadfence();
  }
  __syncthreads();
  if (!sh_last) return;
  double* h0 = outl;
  double* h1 = outl + nb;
  for (int64_t i = tid; i < n; i += nt) {  // each i read and written by one thread
    double sa = 0.0, sb = 0.0;
    for (int cc = in.first_chunk; cc < in.first_chunk + in.n_chunks; ++cc) {
      const double* pp = P.partials + P.chunks[cc].partial_off;
      sa += load_l2(pp + i);
      sb += load_l2(pp + nb + i);
    }
    h0[i] = sa;
    h1[i] = sb;
  }
  if (tid == 0) P.done[ch.instance] = 0u;
  __syncthreads();
  vjp_tail<Real>(P, in, T.cdf, T.hdr->inv_n, h0, h1, grad_density, scratch);
}

// VJP of the guide-table instances without bin ranges: one work-group per chunk.
template <typename Real>
__global__ void __launch_bounds__(1024) vjp_kernel(BatchedParams P, const Real* grad,
                                                   Real* grad_density) {
  extern __shared__ __align__(16) uint8_t lds[];
  const int c = P.vjp_items[P.n_vjp_cell + blockIdx.x].x;
  if (P.inst[P.chunks[c].instance].global_tables)
    vjp_body<Real, true>(P, grad, grad_density, lds, c);
  else
    vjp_body<Real, false>(P, grad, grad_density, lds, c);
}

// VJP of the float32 cell-table instances: one work-group per chunk (the first
// n_vjp_cell items).
__global__ void __launch_bounds__(1024) vjp_cell_kernel(BatchedParams P, const float* grad,
                                                        float* grad_density) {
  extern __shared__ __align__(16) uint8_t lds[];
  fast_vjp_body(P, grad, grad_density, lds, P.vjp_items[blockIdx.x].x);
}

// VJP of the instances with bin ranges: one work-group per (chunk, range).
// A curve whose CDF turns out not to be monotone has no key interval per
// range; its range-0 items take the whole chunk with the global-table VJP.
template <typename Real>
__global__ void __launch_bounds__(1024) vjp_range_kernel(BatchedParams P, const Real* grad,
                                                         Real* grad_density) {
  extern __shared__ __align__(16) uint8_t lds[];
  const int2 item = P.vjp_items[P.n_vjp_plain + blockIdx.x];
  const InstanceDesc& in = P.inst[P.chunks[item.x].instance];
  const TableHeader* hdr = reinterpret_cast<const TableHeader*>(P.tables + in.table_off);
  if (hdr->flags & kMonotone)
    vjp_range_body<Real>(P, grad, grad_density, lds, item.x, item.y);
  else if (item.y == 0)
    vjp_body<Real, true>(P, grad, grad_density, lds, item.x);
}

// Persistent VJP: resident work-groups take items from vjp_queue, largest
// first, whatever their kind (cell tables, guide tables, bin range).
// The persistent VJP's arguments after BatchedParams.
template <typename Real>
struct VjpArgs {
  BatchedParams P;
  const Real* grad;
  Real* grad_density;
};

template <typename Real>
__device__ __forceinline__ const VjpArgs<Real>& vjp_launch_args(uint64_t kargs) {
  return *(const VjpArgs<Real>*)(const __attribute__((address_space(4))) VjpArgs<Real>*)
      uniform_kernarg(kargs);
}

template <typename Real>
__device__ __noinline__ void persistent_vjp_cell(uint64_t kargs, int c_arg) {
  extern __shared__ __align__(16) uint8_t lds[];
  const VjpArgs<Real>& A = vjp_launch_args<Real>(kargs);
  const int c = __builtin_amdgcn_readfirstlane(c_arg);
  if constexpr (sizeof(Real) == 4) fast_vjp_body(A.P, A.grad, A.grad_density, lds, c);
}

template <typename Real, bool GLOBAL>
__device__ __noinline__ void persistent_vjp_guide(uint64_t kargs, int c_arg) {
  extern __shared__ __align__(16) uint8_t lds[];
  const VjpArgs<Real>& A = vjp_launch_args<Real>(kargs);
  const int c = __builtin_amdgcn_readfirstlane(c_arg);
  vjp_body<Real, GLOBAL>(A.P, A.grad, A.grad_density, lds, c);
}

template <typename Real>
__device__ __noinline__ void persistent_vjp_range(uint64_t kargs, int c_arg, int q_arg) {
  extern __shared__ __align__(16) uint8_t lds[];
  const VjpArgs<Real>& A = vjp_launch_args<Real>(kargs);
  const int c = __builtin_amdgcn_readfirstlane(c_arg);
  const int q = __builtin_amdgcn_readfirstlane(q_arg);
  vjp_range_body<Real>(A.P, A.grad, A.grad_density, lds, c, q);
}

template <typename Real>
__global__ void __launch_bounds__(1024) vjp_persistent_kernel(BatchedParams P, const Real* grad,
                                                              Real* grad_density) {
  static_assert(offsetof(VjpArgs<Real>, grad) == sizeof(BatchedParams) &&
                    offsetof(VjpArgs<Real>, grad_density) == sizeof(BatchedParams) + sizeof(void*),
                "the kernel-argument segment holds P, grad, grad_density in this order");
  __shared__ int sh_item;
  for (;;) {
    const int item = claim_item(P, kFlagVjpNext, &sh_item);
    if (item >= P.n_vjp_items) break;
    const int idx = P.vjp_queue[item];
    const int2 it = P.vjp_items[idx];
    const InstanceDesc& in = P.inst[P.chunks[it.x].instance];
    if (idx < P.n_vjp_cell) {
      persistent_vjp_cell<Real>(kernarg_address(), it.x);
    } else if (idx < P.n_vjp_plain) {
      if (in.global_tables)
        persistent_vjp_guide<Real, true>(kernarg_address(), it.x);
      else
        persistent_vjp_guide<Real, false>(kernarg_address(), it.x);
    } else {
      const TableHeader* hdr = reinterpret_cast<const TableHeader*>(P.tables + in.table_off);
      if (hdr->flags & kMonotone)
        persistent_vjp_range<Real>(kernarg_address(), it.x, it.y);
      else if (it.y == 0)
        persistent_vjp_guide<Real, true>(kernarg_address(), it.x);
    }
    __syncthreads();  // the item's LDS and shared words are free
  }
  leave_queue(P, kFlagVjpNext, kFlagVjpExit);
}

// ---------------------------------------------------------------------------
// Host-side launchers (called from capi.hip).

// Work-groups of `kern` that fit the current device at once with this shape.
template <typename Kernel>
int resident_groups(Kernel kern, int threads, int64_t lds) {
  int dev = 0, cus = 0, per = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, threads, size_t(lds)) !=
      hipSuccess)
    return 0;
  return per * cus;
}

template <typename Real>
hipError_t launch_forward_persistent(const BatchedParams& P, int threads, bool sequential,
                                     int64_t lds, hipStream_t stream) {
  if (P.n_fwd_items == 0) return hipSuccess;
  auto kern = P.use_window
                  ? (sequential ? forward_persistent_kernel<Real, true, true>
                                : forward_persistent_kernel<Real, true, false>)
                  : (sequential ? forward_persistent_kernel<Real, false, true>
                                : forward_persistent_kernel<Real, false, false>);
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
  if (e != hipSuccess) return e;
  const int fit = resident_groups(kern, threads, lds);
  if (fit <= 0) return hipErrorInvalidConfiguration;
  const int64_t grid = std::min<int64_t>(fit, P.n_fwd_items);
  hipLaunchKernelGGL(kern, dim3(unsigned(grid)), dim3(threads), size_t(lds), stream, P);
  return hipGetLastError();
}

template <typename Real>
hipError_t launch_vjp_persistent(const BatchedParams& P, int threads, int64_t lds,
                                 const void* grad, void* grad_density, hipStream_t stream) {
  if (P.n_vjp_items == 0) return hipSuccess;
  auto kern = vjp_persistent_kernel<Real>;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
  if (e != hipSuccess) return e;
  const int fit = resident_groups(kern, threads, lds);
  if (fit <= 0) return hipErrorInvalidConfiguration;
  const int64_t grid = std::min<int64_t>(fit, P.n_vjp_items);
  hipLaunchKernelGGL(kern, dim3(unsigned(grid)), dim3(threads), size_t(lds), stream, P,
                     static_cast<const Real*>(grad), static_cast<Real*>(grad_density));
  return hipGetLastError();
}
template <typename Real>
hipError_t launch_tables(const BatchedParams& P, int threads, bool sequential, int64_t lds,
                         bool only_global, hipStream_t stream) {
  auto kern = sequential ? tables_kernel<Real, true> : tables_kernel<Real, false>;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
  if (e != hipSuccess) return e;
  const int64_t groups = only_global ? P.n_global : P.n_instances;
  if (groups == 0) return hipSuccess;
  hipLaunchKernelGGL(kern, dim3(unsigned(groups)), dim3(threads), size_t(lds), stream, P,
                     only_global ? 1 : 0);
  return hipGetLastError();
}

template <typename Real>
hipError_t launch_sample(const BatchedParams& P, int threads, bool sequential, int64_t lds,
                         hipStream_t stream) {
  if constexpr (sizeof(Real) == 4) {
    if (P.n_cell_chunks > 0) {
      auto kern = P.use_window ? (sequential ? sample_cell_kernel<true, true>
                                             : sample_cell_kernel<true, false>)
                               : (sequential ? sample_cell_kernel<false, true>
                                             : sample_cell_kernel<false, false>);
      hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
      if (e != hipSuccess) return e;
      hipLaunchKernelGGL(kern, dim3(unsigned(P.n_cell_chunks)), dim3(threads), size_t(lds),
                         stream, P);
      e = hipGetLastError();
      if (e != hipSuccess) return e;
    }
  }
  if (P.n_chunks - P.n_cell_chunks == 0) return hipSuccess;
  auto kern = P.use_window ? (sequential ? sample_kernel<Real, true, true>
                                         : sample_kernel<Real, true, false>)
                           : (sequential ? sample_kernel<Real, false, true>
                                         : sample_kernel<Real, false, false>);
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kern, dim3(unsigned(P.n_chunks - P.n_cell_chunks)), dim3(threads),
                     size_t(lds), stream, P);
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
hipError_t launch_vjp(const BatchedParams& P, int threads, int64_t lds, const void* grad,
                      void* grad_density, hipStream_t stream) {
  // cell-table items, other plain items, then bin-range items (independent
  // instances: any order)
  if constexpr (sizeof(Real) == 4) {
    if (P.n_vjp_cell > 0) {
      hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(vjp_cell_kernel),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
      if (e != hipSuccess) return e;
      hipLaunchKernelGGL(vjp_cell_kernel, dim3(unsigned(P.n_vjp_cell)), dim3(threads),
                         size_t(lds), stream, P, static_cast<const float*>(grad),
                         static_cast<float*>(grad_density));
      e = hipGetLastError();
      if (e != hipSuccess) return e;
    }
  }
  const int64_t grids[2] = {P.n_vjp_plain - P.n_vjp_cell, P.n_vjp_items - P.n_vjp_plain};
  for (int kind = 0; kind < 2; ++kind) {
    if (grids[kind] == 0) continue;
    auto kern = kind == 0 ? vjp_kernel<Real> : vjp_range_kernel<Real>;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, int(lds));
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kern, dim3(unsigned(grids[kind])), dim3(threads), size_t(lds), stream, P,
                       static_cast<const Real*>(grad), static_cast<Real*>(grad_density));
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

template hipError_t launch_tables<float>(const BatchedParams&, int, bool, int64_t, bool,
                                         hipStream_t);
template hipError_t launch_tables<double>(const BatchedParams&, int, bool, int64_t, bool,
                                          hipStream_t);
template hipError_t launch_sample<float>(const BatchedParams&, int, bool, int64_t, hipStream_t);
template hipError_t launch_sample<double>(const BatchedParams&, int, bool, int64_t, hipStream_t);
template hipError_t launch_pack<float>(const BatchedParams&, int, void*, int32_t*, hipStream_t);
template hipError_t launch_pack<double>(const BatchedParams&, int, void*, int32_t*, hipStream_t);
template hipError_t launch_vjp<float>(const BatchedParams&, int, int64_t, const void*, void*,
                                      hipStream_t);
template hipError_t launch_forward_persistent<float>(const BatchedParams&, int, bool, int64_t,
                                                     hipStream_t);
template hipError_t launch_forward_persistent<double>(const BatchedParams&, int, bool, int64_t,
                                                      hipStream_t);
template hipError_t launch_vjp_persistent<float>(const BatchedParams&, int, int64_t, const void*,
                                                 void*, hipStream_t);
template hipError_t launch_vjp_persistent<double>(const BatchedParams&, int, int64_t, const void*,
                                                  void*, hipStream_t);
template hipError_t launch_vjp<double>(const BatchedParams&, int, int64_t, const void*, void*,
                                       hipStream_t);

}  // namespace qloits
