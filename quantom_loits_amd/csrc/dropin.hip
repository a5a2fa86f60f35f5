// dropin.hip -- the reference drop-in (2-axis) path for gfx950.
//
// One call reproduces one reference `forward` / `backward`
// (cpp/bindings.cpp:24-208) on device memory:
//   alloc_kernel       counts = trunc(|w * n_events|), nmax, active (cpp/loits_core.cpp:22-37)
//   curve_kernel       one thread per x- or q-curve: the reference's sequential
//                      norm / rho / CDF chain (:55-165), integer thresholds and
//                      per-bin slopes m_j (:193-194)
//   row_start_kernel   first row of every cell: the scan of the cell counts
//   sample_kernel      launched twice, one launch doing the work: both Philox
//                      streams for every padded slot (:39-53,
//                      cpp/bindings.cpp:79-92), interval + interpolation for
//                      active slots (:167-237) and the stable compaction
//                      (:239-266) in one pass.  When every curve is drop-free
//                      (no active slot can be dropped) one-wave tiles of 256
//                      slots place their rows from the cell counts; otherwise
//                      1024-thread tiles of 4096 slots are ordered by a
//                      decoupled look-back
//   row_offsets_kernel openmp/loits_core.cpp:244-262 (the cell counts' scan
//                      when drop-free, else from the packed slots)
//   ivjp_cells_kernel  kx + kq <= 64: a lane per node of either axis, up to
//                      two cells per wave; each lane sums only the rows that
//                      touch its node, in row order (the accumulation order
//                      of openmp/loits_core.cpp:285-352), from per-node lists
//                      built per batch of rows; ivjp_wave_kernel (one cell per
//                      wave, kx, kq <= 64) and ivjp_kernel (one thread per
//                      cell) for larger curves
//   curve_vjp_kernel   one thread per curve: CDF and density adjoints (:340-472)
// The *_bounded entry points read nmax, active and the kept count on the
// device instead of taking them from the host.
// Built with -ffp-contract=off; every fused multiply-add of the reference
// build is an explicit fma, so fp64 inputs give the reference's bits.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>

#include "capi_internal.hpp"
#include "loits_device.hpp"
#include "loits_hip.h"

namespace qloits {
namespace {

// Tiles of the look-back order: 4096 slots (fewer chains); tiles of curves
// that are all drop-free: one wave of 256 slots (no ordering between tiles).
constexpr int kTileThreads = 1024;
constexpr int64_t kTileSlots = 4 * kTileThreads;
constexpr int kDirectThreads = 64;
constexpr int64_t kDirectSlots = 4 * kDirectThreads;
constexpr uint64_t kStatusShift = 62;
constexpr uint64_t kAggregate = 1ull << kStatusShift;
constexpr uint64_t kInclusive = 2ull << kStatusShift;
constexpr uint64_t kValueMask = (1ull << kStatusShift) - 1;
constexpr uint64_t kSpinLimit = 1ull << 28;

template <typename Real>
__device__ __forceinline__ double ld(const void* p, int64_t i) {
  return double(static_cast<const Real*>(p)[i]);
}

// ---------------------------------------------------------------------------
template <typename Real>
__global__ void __launch_bounds__(1024) alloc_kernel(const void* weights, int64_t n,
                                                     int64_t n_events, int64_t* counts,
                                                     int64_t* out) {
  __shared__ int64_t smax[16], ssum[16];
  const double scale = double(n_events);
  int64_t mx = 0, sum = 0;
  constexpr int kU = 4;  // weights loaded per thread at a time (latencies overlap)
  for (int64_t i0 = threadIdx.x; i0 < n; i0 += int64_t(kU) * blockDim.x) {
    double v[kU];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int64_t i = i0 + int64_t(u) * blockDim.x;
      v[u] = i < n ? __builtin_fabs(ld<Real>(weights, i) * scale) : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const int64_t i = i0 + int64_t(u) * blockDim.x;
      if (i >= n) continue;
      int64_t c;
      if (!(v[u] == v[u])) c = 0;
      else if (v[u] >= 4611686018427387904.0) c = int64_t(1) << 62;
      else c = static_cast<int64_t>(v[u]);
      counts[i] = c;
      mx = c > mx ? c : mx;
      sum += c;
    }
  }
  for (int d = 32; d >= 1; d >>= 1) {
    const int64_t om = __shfl_xor(mx, d, kWave), os = __shfl_xor(sum, d, kWave);
    mx = om > mx ? om : mx;
    sum += os;
  }
  const int wave = threadIdx.x >> 6;
  if (lane_id() == 0) {
    smax[wave] = mx;
    ssum[wave] = sum;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t m = 0, s = 0;
    for (int w = 0; w < int(blockDim.x >> 6); ++w) {
      m = smax[w] > m ? smax[w] : m;
      s += ssum[w];
    }
    out[0] = m;
    out[1] = s;
  }
}

// ---------------------------------------------------------------------------
// Curve tables.  Curve c < nxc is the x-curve stored at xsec_x[c] (c =
// (b*ny + iy)*nx + ix, nodes x_bins[b, ix]); otherwise the q-curve
// c - nxc stored at xsec_q (c' = (b*nx + ix)*ny + iy, nodes q_bins[b, iy]).
struct CurveParams {
  loits_hip_shape s;
  const void* x_bins;
  const void* xsec_x;
  const void* q_bins;
  const void* xsec_q;
  const uint8_t* acc;
  double* norm_x;
  double* norm_q;
  double* cdf_x;
  double* cdf_q;
  uint32_t* thr_x;
  uint32_t* thr_q;
  double* m_x;
  double* m_q;
  uint8_t* mono;     // [x curves | q curves]
  uint32_t* unsafe;  // set when some curve is not safe (below)
};

__device__ __forceinline__ bool norm_term_fused(int64_t i, int64_t nterms) {
  return nterms <= 3 || ((nterms & 1) && i == nterms - 1);
}

// The chain of one curve.  R > 0: its k <= R densities and nodes are loaded
// into registers first (every load in flight at once) and the CDF and slopes
// stay there for the checks; R == 0 reads them from memory.  The same
// expressions in the same order either way.
constexpr int kRegNodes = 16;

template <typename Real, int R>
__device__ __forceinline__ void curve_chain(const CurveParams& P, int64_t c, const void* xsec,
                                            int64_t off, const void* bins, int64_t bins_off,
                                            int64_t k, double acc, double* norm_out, double* cdf,
                                            uint32_t* thr, double* m) {
  constexpr int A = R > 0 ? R : 1;
  double xr[A], br[A], cr[A], mr[A];
  if constexpr (R > 0) {
#pragma unroll
    for (int t = 0; t < R; ++t) {
      xr[t] = t < k ? ld<Real>(xsec, off + t) : 0.0;
      br[t] = t < k ? ld<Real>(bins, bins_off + t) : 0.0;
    }
  }
  auto X = [&](int64_t t) -> double {
    if constexpr (R > 0) return xr[t]; else return ld<Real>(xsec, off + t);
  };
  auto Bn = [&](int64_t t) -> double {
    if constexpr (R > 0) return br[t]; else return ld<Real>(bins, bins_off + t);
  };
  auto Cd = [&](int64_t t) -> double {
    if constexpr (R > 0) return cr[t]; else return cdf[off + t];
  };
  auto Ms = [&](int64_t t) -> double {
    if constexpr (R > 0) return mr[t]; else return m[off + t];
  };
  const int64_t kl = R > 0 ? R : k;  // loop bound, a constant with registers
  // norm: cpp/loits_core.cpp:73-80 (fusion pattern of the reference build)
  double nrm = 0.0;
#pragma unroll
  for (int64_t t = 1; t < kl; ++t) {
    if (t >= k) break;
    const double h = 0.5 * (Bn(t) - Bn(t - 1));
    const double sum = X(t - 1) + X(t);
    nrm = norm_term_fused(t - 1, k - 1) ? __builtin_fma(h, sum, nrm) : nrm + h * sum;
  }
  *norm_out = nrm;
  const double inv = 1.0 / nrm;
  // CDF: cpp/loits_core.cpp:130-135
  double cum = 0.0, rho_prev = X(0) * inv, c_last = 0.0;
  cdf[off] = 0.0;
  if constexpr (R > 0) cr[0] = 0.0;
  uint32_t prev = 0u;
  bool mono = true;
  thr[off] = 0u;
#pragma unroll
  for (int64_t t = 1; t < kl; ++t) {
    if (t >= k) break;
    const double rho = X(t) * inv;
    cum = __builtin_fma(0.5 * (Bn(t) - Bn(t - 1)), rho_prev + rho, cum);
    const double cv = cum * acc;
    cdf[off + t] = cv;
    if constexpr (R > 0) cr[t] = cv;
    c_last = cv;
    const uint32_t tv = threshold_of(cv);
    thr[off + t] = tv;
    mono = mono && tv >= prev;
    prev = tv;
    rho_prev = rho;
  }
  // slopes: cpp/loits_core.cpp:192-194
#pragma unroll
  for (int64_t j = 0; j + 1 < kl; ++j) {
    if (j + 1 >= k) break;
    const double inv_d = 1.0 / (Cd(j + 1) - Cd(j) + kEpsilon);
    const double mj = (Bn(j + 1) - Bn(j)) * inv_d;
    m[off + j] = mj;
    if constexpr (R > 0) mr[j] = mj;
  }
  P.mono[c] = mono ? 1 : 0;
  // A safe curve gives every slot a finite value in [1e-100, 1e100]: monotone
  // thresholds put u at or above C_j (T_j <= k means C_j <= u, and T_0 = 0),
  // ascending nodes make m_j >= 0, so x_j <= fma(m_j, u - C_j, x_j) <=
  // x_j + m_j.  A cell of two safe curves then keeps every active slot
  // (cpp/loits_core.cpp:256), and rows follow from the counts alone.
  const double x0 = Bn(0);
  bool safe = mono && x0 >= 1e-100 && __builtin_isfinite(c_last);  // c_last = C_{k-1}
#pragma unroll
  for (int64_t j = 0; j + 1 < kl; ++j) {
    if (j + 1 >= k) break;
    const double xj = Bn(j), xn = Bn(j + 1);
    const double mj = Ms(j);
    safe = safe && xn >= xj && __builtin_isfinite(Cd(j)) && __builtin_isfinite(mj) &&
           xj + mj <= 1e100;
  }
  if (!safe) atomicOr(P.unsafe, 1u);
}

template <typename Real>
__global__ void __launch_bounds__(256) curve_kernel(CurveParams P) {
  const loits_hip_shape& s = P.s;
  const int64_t nxc = s.batch * s.ny * s.nx, nqc = s.batch * s.nx * s.ny;
  const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= nxc + nqc) return;
  const bool is_x = c < nxc;
  const int64_t cc = is_x ? c : c - nxc;
  int64_t b, ix, iy, k;
  const void* bins;
  const void* xsec;
  double *norm, *cdf, *m;
  uint32_t* thr;
  int64_t bins_off;
  if (is_x) {
    b = cc / (s.ny * s.nx);
    const int64_t r = cc - b * s.ny * s.nx;
    iy = r / s.nx;
    ix = r - iy * s.nx;
    k = s.kx;
    bins = P.x_bins;
    bins_off = (b * s.nx + ix) * k;
    xsec = P.xsec_x;
    norm = P.norm_x;
    cdf = P.cdf_x;
    thr = P.thr_x;
    m = P.m_x;
  } else {
    b = cc / (s.nx * s.ny);
    const int64_t r = cc - b * s.nx * s.ny;
    ix = r / s.ny;
    iy = r - ix * s.ny;
    k = s.kq;
    bins = P.q_bins;
    bins_off = (b * s.ny + iy) * k;
    xsec = P.xsec_q;
    norm = P.norm_q;
    cdf = P.cdf_q;
    thr = P.thr_q;
    m = P.m_q;
  }
  const int64_t off = cc * k;
  const double acc = P.acc[(b * s.nx + ix) * s.ny + iy] ? 1.0 : 0.0;
  if (k <= kRegNodes)
    curve_chain<Real, kRegNodes>(P, c, xsec, off, bins, bins_off, k, acc, norm + cc, cdf, thr, m);
  else
    curve_chain<Real, 0>(P, c, xsec, off, bins, bins_off, k, acc, norm + cc, cdf, thr, m);
}

// row0[g] = active slots of the cells before g (counts summed in cell
// order); row0[gcells] = all of them.  One work-group of 1024 threads, 8192
// cells at a time: coalesced loads into LDS, eight consecutive cells per
// thread, a block scan, coalesced stores.
constexpr int kRowPiece = 8192;

__global__ void __launch_bounds__(1024) row_start_kernel(const int64_t* counts, int64_t gcells,
                                                         int64_t* row0) {
  __shared__ int64_t buf[kRowPiece];
  __shared__ int64_t wave_tot[16];
  const int tid = threadIdx.x, wave = tid >> 6;
  constexpr int kPer = kRowPiece / 1024;
  int64_t carry = 0;
  for (int64_t base = 0; base < gcells; base += kRowPiece) {
    const int n = gcells - base < kRowPiece ? int(gcells - base) : kRowPiece;
    int64_t v[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int i = tid + k * 1024;
      v[k] = i < n ? counts[base + i] : 0;
    }
#pragma unroll
    for (int k = 0; k < kPer; ++k) buf[tid + k * 1024] = v[k];
    __syncthreads();
    int64_t local = 0;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      v[k] = buf[tid * kPer + k];
      local += v[k];
    }
    const int64_t inc = wave_inclusive_scan(local);
    if (lane_id() == kWave - 1) wave_tot[wave] = inc;
    __syncthreads();
    int64_t run = carry + inc - local, total = 0;
    for (int w = 0; w < 16; ++w) {
      run += w < wave ? wave_tot[w] : 0;
      total += wave_tot[w];
    }
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      buf[tid * kPer + k] = run;
      run += v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int i = tid + k * 1024;
      if (i < n) row0[base + i] = buf[i];
    }
    carry += total;
    __syncthreads();
  }
  if (tid == 0) row0[gcells] = carry;
}

// Curves of at most kSmallK nodes: the count of cpp/loits_core.cpp:12-18 as
// it is written, with all thresholds loaded at once (one cache line or two).
constexpr int kSmallK = 16;

// interval = clamp(count - 1, 0, K - 2)
__device__ __forceinline__ int64_t clamp_interval(int64_t count, int64_t K) {
  int64_t j = count - 1;
  j = j < 0 ? 0 : j;
  return j > K - 2 ? K - 2 : j;
}

// count(k) = #{t : T_t <= k}; interval = clamp(count - 1, 0, K - 2)
__device__ __forceinline__ int64_t interval_of(const uint32_t* thr, int64_t K, bool mono,
                                               uint32_t key) {
  int64_t count;
  if (K <= kSmallK) {
    int c = 0;
#pragma unroll
    for (int t = 0; t < kSmallK; ++t) c += (t < K && thr[t] <= key) ? 1 : 0;
    count = c;
  } else if (mono) {
    int64_t lo = 0, hi = K;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (thr[mid] <= key) lo = mid + 1; else hi = mid;
    }
    count = lo;
  } else {
    count = 0;
    for (int64_t t = 0; t < K; ++t) count += thr[t] <= key ? 1 : 0;
  }
  return clamp_interval(count, K);
}

// n / d for n >= 0, d > 0: a 32-bit division when both fit in 32 bits (every
// slot and cell index below 2^32), the 64-bit one otherwise.  The 64-bit
// division is a long VALU sequence and the sampling kernel is issue-bound.
__device__ __forceinline__ int64_t udiv(int64_t n, int64_t d) {
  if (((uint64_t(n) | uint64_t(d)) >> 32) == 0) return int64_t(uint32_t(n) / uint32_t(d));
  return n / d;
}

// ---------------------------------------------------------------------------
struct SampleParams {
  loits_hip_shape s;
  int64_t nmax, total_slots, n_tiles;
  uint64_t seed, sid_x, sid_q;
  const void* x_bins;
  const void* q_bins;
  const int64_t* counts;
  const double* cdf_x;
  const double* cdf_q;
  const uint32_t* thr_x;
  const uint32_t* thr_q;
  const double* m_x;
  const double* m_q;
  const uint8_t* mono;
  float* u_x;
  float* u_q;
  int16_t* interval_x;
  int16_t* interval_q;
  void* events;
  int64_t* packed;
  uint64_t* tile_status;
  uint32_t* tile_counter;
  int64_t* valid;  // [0] kept, [1] error flag
  // sync-free form: nmax and active read here; the values above are capacities
  const int64_t* dev_na;
  int64_t active_cap;
  const uint32_t* unsafe;    // 0: every cell safe, rows from row0 (no look-back)
  const int64_t* row0;       // [cells + 1] first row of each cell
};

// Launched twice per call: DIRECT (kDirectThreads) does the work when every
// curve is drop-free, the other (kTileThreads) when some curve is not; each
// returns at once in the other case.  Drop-free: every active slot is kept,
// a tile's first row is the count of active slots before it and the
// work-group index is the tile.  Otherwise tiles are taken from a counter, in
// the order the decoupled look-back needs.
// Five waves per SIMD: the register count fits 96 without spilling (98 left
// four).
template <typename Real, bool DIRECT>
__global__ void __launch_bounds__(DIRECT ? kDirectThreads : kTileThreads)
    __attribute__((amdgpu_waves_per_eu(5))) sample_kernel(SampleParams P0) {
  constexpr int kThreads = DIRECT ? kDirectThreads : kTileThreads;
  constexpr int64_t kSlots = 4 * kThreads;
  __shared__ int32_t wave_tot[16];
  __shared__ int64_t sh_tile, sh_excl;
  // one-wave tiles stage their kept rows so the stores below are coalesced
  __shared__ Real st_ev[DIRECT ? 2 * kSlots : 1];
  __shared__ int64_t st_pk[DIRECT ? kSlots : 1];
  if ((*P0.unsafe == 0) != DIRECT) return;  // the other launch's case
  SampleParams P = P0;
  const loits_hip_shape& s = P.s;
  if (P0.dev_na) {  // sizes from the device; the launch covers the capacities
    const int64_t nm = P0.dev_na[0], act = P0.dev_na[1];
    if (nm > P0.nmax || act > P0.active_cap) {
      if (blockIdx.x == 0 && threadIdx.x == 0)
        atomicOr(reinterpret_cast<unsigned long long*>(&P0.valid[1]), 2ull);
      return;
    }
    P.nmax = nm;
    P.total_slots = s.batch * s.cells * nm;
  }
  const int64_t n_tiles = (P.total_slots + kSlots - 1) / kSlots;
  constexpr bool direct = DIRECT;
  if (threadIdx.x == 0)
    sh_tile = direct ? int64_t(blockIdx.x) : int64_t(atomicAdd(P.tile_counter, 1u));
  __syncthreads();
  const int64_t tile = sh_tile;
  if (tile >= n_tiles) return;  // (capacity tiles past the live slots)
  const int64_t q = tile * kThreads + threadIdx.x;  // Philox block of this thread
  const Words4 wx = philox4x32_10_uniform(uint64_t(q), P.sid_x, P.seed);
  const Words4 wq = philox4x32_10_uniform(uint64_t(q), P.sid_q, P.seed);
  double vx[4], vq[4];
  int keepbits = 0;
  const int64_t nxc = s.batch * s.ny * s.nx;
  const int64_t p0 = 4 * q;
  // cell g and slot of p0 by one division; the next three slots step from there
  int64_t g = udiv(p0, P.nmax), slot = p0 - g * P.nmax;
  int64_t cur_g = -1, count = 0, cxi = 0, cqi = 0, ox = 0, bx = 0, oq = 0, bq = 0;
  int activebits = 0;
  bool mono_x = false, mono_q = false;
  float ufx[4], ufq[4];
  int16_t ix4[4] = {0, 0, 0, 0}, iq4[4] = {0, 0, 0, 0};
  auto enter_cell = [&](int64_t gg) {  // the curves of cell gg
    cur_g = gg;
    count = P.counts[gg];
    const int64_t b = udiv(gg, s.cells), cell = gg - b * s.cells;
    const int64_t ix = udiv(cell, s.ny), iy = cell - ix * s.ny;
    cxi = (b * s.ny + iy) * s.nx + ix;
    cqi = (b * s.nx + ix) * s.ny + iy;
    ox = cxi * s.kx;
    bx = (b * s.nx + ix) * s.kx;
    oq = cqi * s.kq;
    bq = (b * s.ny + iy) * s.kq;
    mono_x = P.mono[cxi] != 0;
    mono_q = P.mono[nxc + cqi] != 0;
  };
  if (p0 < P.total_slots) enter_cell(g);
  // Four active slots of one cell with small curves (the usual case): each
  // threshold is loaded once for the four keys.
  int cnt_x[4] = {-1, -1, -1, -1}, cnt_q[4] = {-1, -1, -1, -1};
  if (s.kx <= kSmallK && s.kq <= kSmallK && p0 + 3 < P.total_slots && slot + 3 < P.nmax &&
      slot + 3 < count) {
    const uint32_t* tx = P.thr_x + ox;
    const uint32_t* tq = P.thr_q + oq;
#pragma unroll
    for (int e = 0; e < 4; ++e) cnt_x[e] = cnt_q[e] = 0;
#pragma unroll
    for (int t = 0; t < kSmallK; ++t) {
      const uint32_t vx = t < s.kx ? tx[t] : kNever, vq = t < s.kq ? tq[t] : kNever;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        cnt_x[e] += vx <= (word_of(wx, e) >> 8) ? 1 : 0;
        cnt_q[e] += vq <= (word_of(wq, e) >> 8) ? 1 : 0;
      }
    }
  }
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    if (e > 0 && ++slot == P.nmax) {
      slot = 0;
      ++g;
    }
    const int64_t p = p0 + e;
    const uint32_t kx = word_of(wx, e) >> 8, kq = word_of(wq, e) >> 8;
    ufx[e] = float(kx) * 0x1.0p-24f;
    ufq[e] = float(kq) * 0x1.0p-24f;
    if (p >= P.total_slots) continue;
    if (g != cur_g) enter_cell(g);
    if (slot >= count) continue;
    activebits |= 1 << e;
    // x: cpp/loits_core.cpp:183-196
    const int64_t jx = cnt_x[e] >= 0 ? clamp_interval(cnt_x[e], s.kx)
                                     : interval_of(P.thr_x + ox, s.kx, mono_x, kx);
    const double ux = double(kx) * 0x1.0p-24;
    vx[e] = __builtin_fma(P.m_x[ox + jx], ux - P.cdf_x[ox + jx], ld<Real>(P.x_bins, bx + jx));
    // q: cpp/loits_core.cpp:219-232
    const int64_t jq = cnt_q[e] >= 0 ? clamp_interval(cnt_q[e], s.kq)
                                     : interval_of(P.thr_q + oq, s.kq, mono_q, kq);
    const double uq = double(kq) * 0x1.0p-24;
    vq[e] = __builtin_fma(P.m_q[oq + jq], uq - P.cdf_q[oq + jq], ld<Real>(P.q_bins, bq + jq));
    ix4[e] = int16_t(jx);
    iq4[e] = int16_t(jq);
    // keep rule: cpp/loits_core.cpp:256
    if (__builtin_isfinite(vx[e]) && __builtin_isfinite(vq[e]) && vx[e] * vq[e] != 0.0)
      keepbits |= 1 << e;
  }
  // every padded slot's uniforms and intervals (0 for inactive slots)
  if (p0 + 3 < P.total_slots) {
    *reinterpret_cast<float4*>(P.u_x + p0) = make_float4(ufx[0], ufx[1], ufx[2], ufx[3]);
    *reinterpret_cast<float4*>(P.u_q + p0) = make_float4(ufq[0], ufq[1], ufq[2], ufq[3]);
    *reinterpret_cast<short4*>(P.interval_x + p0) = make_short4(ix4[0], ix4[1], ix4[2], ix4[3]);
    *reinterpret_cast<short4*>(P.interval_q + p0) = make_short4(iq4[0], iq4[1], iq4[2], iq4[3]);
  } else {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      if (p0 + e >= P.total_slots) break;
      P.u_x[p0 + e] = ufx[e];
      P.u_q[p0 + e] = ufq[e];
      P.interval_x[p0 + e] = ix4[e];
      P.interval_q[p0 + e] = iq4[e];
    }
  }
  if (direct && keepbits != activebits)  // cannot happen for safe curves: report it
    atomicOr(reinterpret_cast<unsigned long long*>(&P.valid[1]), 4ull);
  // in-tile ordered offsets
  const int cnt = __popc(keepbits);
  const uint64_t lt = lanemask_lt();
  const uint64_t b0 = ballot(cnt & 1), b1 = ballot(cnt & 2), b2 = ballot(cnt & 4);
  const int pre = popc64(b0 & lt) + 2 * popc64(b1 & lt) + 4 * popc64(b2 & lt);
  const int wave = threadIdx.x >> 6;
  if (lane_id() == 0) wave_tot[wave] = popc64(b0) + 2 * popc64(b1) + 4 * popc64(b2);
  __syncthreads();
  int base = 0, agg = 0;
  for (int w = 0; w < kThreads / 64; ++w) {
    agg += wave_tot[w];
    base += w < wave ? wave_tot[w] : 0;
  }
  // decoupled look-back for the tile's first row: wave 0 reads the status
  // words of 64 predecessors at a time and sums back to the nearest
  // inclusive prefix (a predecessor without a status yet is waited for).
  // Each status word carries its own value, so relaxed device-scope atomics
  // suffice (no fence: nothing else written by a tile is read in this kernel).
  if (direct) {
    if (threadIdx.x == 0) {
      const int64_t p = tile * kSlots, g0 = udiv(p, P.nmax), s0 = p - g0 * P.nmax;
      const int64_t c0 = P.counts[g0];
      sh_excl = P.row0[g0] + (s0 < c0 ? s0 : c0);
      if (tile == n_tiles - 1) P.valid[0] = P.row0[s.batch * s.cells];
    }
  } else if (threadIdx.x < kWave) {
    const int lane = lane_id();
    uint64_t* st = P.tile_status;
    if (tile == 0) {
      if (lane == 0) {
        __hip_atomic_store(&st[0], kInclusive | uint64_t(agg), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        sh_excl = 0;
      }
    } else {
      if (lane == 0)
        __hip_atomic_store(&st[tile], kAggregate | uint64_t(agg), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
      int64_t excl = 0;
      int64_t end = tile - 1;
      uint64_t spins = 0;
      bool failed = false;
      while (true) {
        const int64_t idx = end - lane;
        const uint64_t v = idx >= 0 ? __hip_atomic_load(&st[idx], __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT)
                                    : kInclusive;
        const uint64_t flag = v & ~kValueMask;
        const uint64_t inc = ballot(flag == kInclusive);
        const uint64_t waiting = ballot(flag == 0);
        const int first = inc ? __builtin_ctzll(inc) : kWave;  // nearest inclusive
        const uint64_t need = first >= kWave - 1 ? ~0ull : ((2ull << first) - 1ull);
        if (waiting & need) {
          if (++spins > kSpinLimit) {  // a bounded wait: report instead of hanging
            failed = true;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
        uint64_t part = lane <= first ? (v & kValueMask) : 0ull;
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) {
          const uint32_t lo = __shfl_xor(uint32_t(part), d, kWave);
          const uint32_t hi = __shfl_xor(uint32_t(part >> 32), d, kWave);
          part += uint64_t(lo) | (uint64_t(hi) << 32);
        }
        excl += int64_t(part);
        if (first < kWave) break;
        end -= kWave;
      }
      if (lane == 0) {
        if (failed) atomicOr(reinterpret_cast<unsigned long long*>(&P.valid[1]), 1ull);
        __hip_atomic_store(&st[tile], kInclusive | uint64_t(excl + agg), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        sh_excl = excl;
      }
    }
    if (lane == 0 && tile == n_tiles - 1) P.valid[0] = sh_excl + agg;
  }
  Real* ev = static_cast<Real*>(P.events);
  if (direct) {  // one wave: base == 0, the tile's rows are [sh_excl, sh_excl + agg)
    int r = pre;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      if ((keepbits >> e) & 1) {
        st_ev[2 * r] = Real(vx[e]);
        st_ev[2 * r + 1] = Real(vq[e]);
        st_pk[r] = 4 * q + e;
        ++r;
      }
    }
    __syncthreads();
    const int64_t r0 = sh_excl;
    for (int i = threadIdx.x; i < 2 * agg; i += kThreads) ev[2 * r0 + i] = st_ev[i];
    for (int i = threadIdx.x; i < agg; i += kThreads) P.packed[r0 + i] = st_pk[i];
    return;
  }
  __syncthreads();
  int64_t row = sh_excl + base + pre;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    if ((keepbits >> e) & 1) {
      ev[2 * row] = Real(vx[e]);
      ev[2 * row + 1] = Real(vq[e]);
      P.packed[row] = 4 * q + e;
      ++row;
    }
  }
}

// Drop-free curves: the rows of cell g start at row0[g] (no search).
__global__ void row_offsets_kernel(const int64_t* packed, const int64_t* valid, int64_t nmax,
                                   int64_t gcells, int64_t* row_offsets, const int64_t* dev_na,
                                   const uint32_t* unsafe, const int64_t* row0) {
  const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g > gcells) return;
  if (*unsafe == 0) {
    row_offsets[g] = row0[g];
    return;
  }
  if (dev_na) nmax = dev_na[0];
  const int64_t key = g * nmax;
  int64_t lo = 0, hi = valid[0];
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (packed[mid] < key) lo = mid + 1; else hi = mid;
  }
  row_offsets[g] = lo;
}

// ---------------------------------------------------------------------------
// Backward.
struct VjpParams {
  loits_hip_shape s;
  int64_t nmax, nvalid;
  const void* grad;
  const void* x_bins;
  const void* xsec_x;
  const void* q_bins;
  const void* xsec_q;
  const uint8_t* acc;
  const double* norm_x;
  const double* norm_q;
  const double* cdf_x;
  const double* cdf_q;
  const float* u_x;
  const float* u_q;
  const int16_t* interval_x;
  const int16_t* interval_q;
  const int64_t* packed;
  double* gcdf_x;
  double* gcdf_q;
  double* grho;  // scratch per curve, same layout as gcdf (x then q)
  void* grad_x;
  void* grad_q;
  // sync-free form: nmax and nvalid read here
  const int64_t* dev_nmax;
  const int64_t* dev_nvalid;
  const int64_t* row_offsets;  // optional: rows of cell g (else searched in packed)
};

// Rows [lo, hi) of cell g from row_offsets, clamped to [0, nvalid).
__device__ __forceinline__ void cell_rows(const VjpParams& P, int64_t g, int64_t& lo,
                                          int64_t& hi) {
  int64_t a = P.row_offsets[g], b = P.row_offsets[g + 1];
  a = a < 0 ? 0 : (a > P.nvalid ? P.nvalid : a);
  b = b < a ? a : (b > P.nvalid ? P.nvalid : b);
  lo = a;
  hi = b;
}

__device__ __forceinline__ VjpParams with_device_sizes(const VjpParams& P0) {
  VjpParams P = P0;
  if (P0.dev_nmax) {
    P.nmax = *P0.dev_nmax;
    P.nvalid = *P0.dev_nvalid;
  }
  return P;
}

// cpp/loits_core.cpp:294-301, one row
__device__ __forceinline__ void ivjp_row(const double* curve, double* gcurve, double bj,
                                         double bj1, int64_t j, double uv, double gev) {
  const double c0 = curve[j];
  const double d = curve[j + 1] - c0 + kEpsilon;
  const double inv_d = 1.0 / d;
  const double t = uv - c0;
  const double upstream = gev * (bj1 - bj);
  gcurve[j] = __builtin_fma(upstream, __builtin_fma(t * inv_d, inv_d, -inv_d), gcurve[j]);
  gcurve[j + 1] = __builtin_fma(-(upstream * t * inv_d), inv_d, gcurve[j + 1]);
}

template <typename Real>
__global__ void __launch_bounds__(256) ivjp_kernel(VjpParams P0) {
  const VjpParams P = with_device_sizes(P0);
  const loits_hip_shape& s = P.s;
  const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g >= s.batch * s.cells) return;
  const int64_t b = g / s.cells, cell = g - b * s.cells;
  const int64_t ix = cell / s.ny, iy = cell - ix * s.ny;
  const int64_t ox = ((b * s.ny + iy) * s.nx + ix) * s.kx, bx = (b * s.nx + ix) * s.kx;
  const int64_t oq = ((b * s.nx + ix) * s.ny + iy) * s.kq, bq = (b * s.ny + iy) * s.kq;
  for (int64_t t = 0; t < s.kx; ++t) P.gcdf_x[ox + t] = 0.0;
  for (int64_t t = 0; t < s.kq; ++t) P.gcdf_q[oq + t] = 0.0;
  // rows of this cell: packed slots in [g*nmax, (g+1)*nmax)
  int64_t lo = 0, hi = P.nvalid;
  const int64_t key = g * P.nmax;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (P.packed[mid] < key) lo = mid + 1; else hi = mid;
  }
  for (int64_t row = lo; row < P.nvalid && P.packed[row] < key + P.nmax; ++row) {
    const int64_t p = P.packed[row];
    const int64_t jx = P.interval_x[p], jq = P.interval_q[p];
    ivjp_row(P.cdf_x + ox, P.gcdf_x + ox, ld<Real>(P.x_bins, bx + jx),
             ld<Real>(P.x_bins, bx + jx + 1), jx, double(P.u_x[p]), ld<Real>(P.grad, 2 * row));
    ivjp_row(P.cdf_q + oq, P.gcdf_q + oq, ld<Real>(P.q_bins, bq + jq),
             ld<Real>(P.q_bins, bq + jq + 1), jq, double(P.u_q[p]),
             ld<Real>(P.grad, 2 * row + 1));
  }
}

// First index i in [lo, hi) with a[i] >= key (hi if none) of an ascending
// array: a 65-way search, one probe per lane per step (wave-uniform result).
__device__ __forceinline__ int64_t wave_lower_bound(const int64_t* a, int64_t lo, int64_t hi,
                                                    int64_t key) {
  // the answer lies in [lo, hi]
  const int lane = lane_id();
  while (lo < hi) {
    const int64_t pos = lo + ((hi - lo) * (lane + 1)) / (kWave + 1);  // in [lo, hi), ascending
    const int c = popc64(ballot(a[pos] < key));  // the probes below key are lanes [0, c)
    const int64_t below = c > 0 ? wave_lane(pos, c - 1) : lo - 1;
    const int64_t above = c < kWave ? wave_lane(pos, c) : hi;
    lo = below + 1;
    hi = above;
  }
  return lo;
}

// Lane j receives the mask of the lanes r < nr whose v equals j (0 <= v < 64):
// one ballot per distinct value.
__device__ __forceinline__ uint64_t rows_with_value(int v, int nr) {
  const int lane = lane_id();
  uint64_t left = ballot(lane < nr), mine = 0ull;
  while (left != 0ull) {
    const int val = __builtin_amdgcn_readlane(v, __builtin_ctzll(left));
    const uint64_t m = ballot(v == val) & left;
    mine = lane == val ? m : mine;
    left &= ~m;
  }
  return mine;
}

__device__ __forceinline__ uint64_t shfl_up1_u64(uint64_t v) {
  const uint32_t lo = uint32_t(__shfl_up(int(uint32_t(v)), 1, kWave));
  const uint32_t hi = uint32_t(__shfl_up(int(uint32_t(v >> 32)), 1, kWave));
  return lane_id() == 0 ? 0ull : (uint64_t(lo) | (uint64_t(hi) << 32));
}

template <typename Real>
__global__ void __launch_bounds__(256) ivjp_wave_kernel(VjpParams P0) {
  const VjpParams P = with_device_sizes(P0);
  __shared__ unsigned long long masks[4 * 2 * kWave];
  __shared__ int2 offsets[4 * kWave];
  __shared__ double2 entries[4 * 4 * kWave];
  __shared__ double nodes[4 * 4 * kWave];  // the cell's C_x, x, C_q, q
  const loits_hip_shape& s = P.s;
  const int lane = lane_id();
  const int64_t g = int64_t(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (g >= s.batch * s.cells) return;  // whole waves (no barrier below)
  const int64_t b = g / s.cells, cell = g - b * s.cells;
  const int64_t ix = cell / s.ny, iy = cell - ix * s.ny;
  const int64_t ox = ((b * s.ny + iy) * s.nx + ix) * s.kx, bx = (b * s.nx + ix) * s.kx;
  const int64_t oq = ((b * s.nx + ix) * s.ny + iy) * s.kq, bq = (b * s.ny + iy) * s.kq;
  int64_t lo0, lo1;  // rows of this cell: [lo0, lo1)
  if (P.row_offsets) {
    cell_rows(P, g, lo0, lo1);
  } else {
    const int64_t key0 = g * P.nmax;
    lo0 = wave_lower_bound(P.packed, 0, P.nvalid, key0);
    const int64_t end = P.nvalid - lo0 < P.nmax ? P.nvalid : lo0 + P.nmax;  // at most nmax rows
    lo1 = wave_lower_bound(P.packed, lo0, end, key0 + P.nmax);
  }
  const int w = threadIdx.x >> 6;
  unsigned long long* mx = masks + w * 2 * kWave;  // rows with interval j, then touching bin j
  unsigned long long* mq = mx + kWave;
  int2* offs = offsets + w * kWave;                 // first list entry of bin j (x, q)
  double2* ex = entries + w * 4 * kWave;            // factor pairs, bin by bin, rows in order
  double2* eq = ex + 2 * kWave;
  double* cx = nodes + w * 4 * kWave;
  double* xn = cx + kWave;
  double* cq = xn + kWave;
  double* qn = cq + kWave;
  if (lane < s.kx) {
    cx[lane] = P.cdf_x[ox + lane];
    xn[lane] = ld<Real>(P.x_bins, bx + lane);
  }
  if (lane < s.kq) {
    cq[lane] = P.cdf_q[oq + lane];
    qn[lane] = ld<Real>(P.q_bins, bq + lane);
  }
  __builtin_amdgcn_wave_barrier();
  // The packed slot and cotangents of a batch's rows are loaded one batch ahead.
  auto row_of = [&](int64_t base) {
    const int64_t n = lo1 - base;
    return base + (lane < n ? lane : 0);
  };
  int64_t p = 0;
  double gx = 0.0, gq = 0.0;
  if (lo0 < lo1) {
    const int64_t row = row_of(lo0);
    p = P.packed[row];
    gx = ld<Real>(P.grad, 2 * row);
    gq = ld<Real>(P.grad, 2 * row + 1);
  }
  double accx = 0.0, accq = 0.0;
  for (int64_t base = lo0; base < lo1; base += kWave) {
    const int nr = lo1 - base < kWave ? int(lo1 - base) : kWave;
    const int jx = P.interval_x[p], jq = P.interval_q[p];
    const double ux = double(P.u_x[p]), uq = double(P.u_q[p]);
    const double gxb = gx, gqb = gq;
    if (base + kWave < lo1) {
      const int64_t row = row_of(base + kWave);
      p = P.packed[row];
      gx = ld<Real>(P.grad, 2 * row);
      gq = ld<Real>(P.grad, 2 * row + 1);
    }
    double upx, ax, bxm, idx, upq, aq, bqm, idq;
    {  // x: the factors of ivjp_row
      const double c0 = cx[jx];
      const double d = cx[jx + 1] - c0 + kEpsilon;
      const double inv_d = 1.0 / d;
      const double t = ux - c0;
      upx = gxb * (xn[jx + 1] - xn[jx]);
      ax = __builtin_fma(t * inv_d, inv_d, -inv_d);
      bxm = -(upx * t * inv_d);
      idx = inv_d;
    }
    {  // q
      const double c0 = cq[jq];
      const double d = cq[jq + 1] - c0 + kEpsilon;
      const double inv_d = 1.0 / d;
      const double t = uq - c0;
      upq = gqb * (qn[jq + 1] - qn[jq]);
      aq = __builtin_fma(t * inv_d, inv_d, -inv_d);
      bqm = -(upq * t * inv_d);
      idq = inv_d;
    }
    // Lane j holds bin j of x and of q.  Row r adds upx ax to bin jx and
    // bx idx to bin jx + 1, so bin j takes the rows with interval j or j - 1,
    // in row order: each row writes its two factor pairs into the lists of
    // those bins (list position = earlier rows touching the bin), then every
    // lane sums its own list.
    uint64_t tx = rows_with_value(jx, nr), tq = rows_with_value(jq, nr);  // interval == j
    tx |= shfl_up1_u64(tx);  // rows touching bin j
    tq |= shfl_up1_u64(tq);
    const int len_x = popc64(tx), len_q = popc64(tq);
    const int off_x = wave_inclusive_scan(len_x) - len_x;  // at most 2 nr entries in all
    const int off_q = wave_inclusive_scan(len_q) - len_q;
    __builtin_amdgcn_wave_barrier();
    mx[lane] = tx;
    mq[lane] = tq;
    offs[lane] = make_int2(off_x, off_q);
    __builtin_amdgcn_wave_barrier();
    if (lane < nr) {
      const uint64_t lt = lanemask_lt();
      const int2 oa = offs[jx], ob = offs[jx + 1], qa = offs[jq], qb = offs[jq + 1];
      ex[oa.x + popc64(mx[jx] & lt)] = make_double2(upx, ax);
      ex[ob.x + popc64(mx[jx + 1] & lt)] = make_double2(bxm, idx);
      eq[qa.y + popc64(mq[jq] & lt)] = make_double2(upq, aq);
      eq[qb.y + popc64(mq[jq + 1] & lt)] = make_double2(bqm, idq);
    }
    __builtin_amdgcn_wave_barrier();
    // four entries per list at a time, their reads issued together (a lane
    // past the end of its list reads an entry in range and keeps its sum)
    for (int i0 = 0; ballot(i0 < len_x || i0 < len_q) != 0ull; i0 += 4) {
      double2 fx[4], fq[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        // past the end of a list: entry 0, one broadcast address
        fx[u] = ex[i0 + u < len_x ? off_x + i0 + u : 0];
        fq[u] = eq[i0 + u < len_q ? off_q + i0 + u : 0];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const double nx = __builtin_fma(fx[u].x, fx[u].y, accx);
        const double nq = __builtin_fma(fq[u].x, fq[u].y, accq);
        accx = i0 + u < len_x ? nx : accx;
        accq = i0 + u < len_q ? nq : accq;
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
  if (lane < s.kx) P.gcdf_x[ox + lane] = accx;
  if (lane < s.kq) P.gcdf_q[oq + lane] = accq;
}

// Largest value over the wave.
__device__ __forceinline__ int64_t wave_max_i64(int64_t v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    const uint32_t lo = uint32_t(__shfl_xor(int(uint32_t(uint64_t(v))), d, kWave));
    const uint32_t hi = uint32_t(__shfl_xor(int(uint32_t(uint64_t(v) >> 32)), d, kWave));
    const int64_t o = int64_t(uint64_t(lo) | (uint64_t(hi) << 32));
    v = o > v ? o : v;
  }
  return v;
}

// Node offsets of cell g (x curve, x bins, q curve, q bins).
struct CellGeo {
  int64_t ox, bx, oq, bq;
};

__device__ __forceinline__ CellGeo cell_geo(const loits_hip_shape& s, int64_t g) {
  const int64_t b = g / s.cells, cell = g - b * s.cells;
  const int64_t ix = cell / s.ny, iy = cell - ix * s.ny;
  return CellGeo{((b * s.ny + iy) * s.nx + ix) * s.kx, (b * s.nx + ix) * s.kx,
                 ((b * s.nx + ix) * s.ny + iy) * s.kq, (b * s.ny + iy) * s.kq};
}

// Curves with kx + kq <= 64: C = min(64 / (kx + kq), 2) cells per wave (two
// measured faster than one or three at 9801 cells of 10 + 10 nodes), one lane
// per node of either axis.  Lanes [c R, c R + R)
// (R = 64 / C) take the rows of cell c, R per round; lane c L + j (L = kx + kq)
// owns bin j of x (j < kx) or bin j - kx of q of cell c and sums the factor
// pairs of the rows touching it, in row order (openmp/loits_core.cpp:285-352).
template <typename Real>
__global__ void __launch_bounds__(256) ivjp_cells_kernel(VjpParams P0) {
  const VjpParams P = with_device_sizes(P0);
  __shared__ unsigned long long touch[4 * kWave];
  __shared__ int offsets[4 * kWave];
  __shared__ double2 entries[4 * 4 * kWave];
  __shared__ double nodes[4 * 3 * kWave];
  const loits_hip_shape& s = P.s;
  const int lane = lane_id(), w = threadIdx.x >> 6;
  const int kx = int(s.kx), L = int(s.kx + s.kq), C = kWave / L < 2 ? kWave / L : 2;
  const int R = kWave / C;
  const int64_t gcells = s.batch * s.cells;
  const int64_t g0 = (int64_t(blockIdx.x) * (blockDim.x >> 6) + w) * C;
  if (g0 >= gcells) return;  // whole waves (no barrier below)
  unsigned long long* T = touch + w * kWave;  // rows with value j, then touching bin j
  int* offs = offsets + w * kWave;            // first list entry of bin j
  double2* ent = entries + w * 4 * kWave;     // factor pairs, bin by bin, rows in order
  // node of bin lane j: (C_j, 1 / (C_{j+1} - C_j + eps)) and x_{j+1} - x_j,
  // the factors every row of interval j uses (evaluated once per node)
  double2* cd = reinterpret_cast<double2*>(nodes + w * 3 * kWave);
  double* dxs = nodes + w * 3 * kWave + 2 * kWave;
  const int cb = lane / L, bb = lane - cb * L;  // this lane as a bin
  const int cr = lane / R, ir = lane - cr * R;  // this lane as a row
  const bool bin_live = cb < C && g0 + cb < gcells;
  const bool row_lane = cr < C;
  // rows of cell cr: [lo, hi); a cell holds at most nmax rows, so each
  // boundary after the first is searched within nmax of the previous one
  int64_t lo = 0, hi = 0;
  if (P.row_offsets) {
    if (row_lane && g0 + cr < gcells) cell_rows(P, g0 + cr, lo, hi);
  } else {
    int64_t a = wave_lower_bound(P.packed, 0, P.nvalid, g0 * P.nmax);
    for (int c = 0; c < C && g0 + c < gcells; ++c) {
      const int64_t end = P.nvalid - a < P.nmax ? P.nvalid : a + P.nmax;
      const int64_t b = wave_lower_bound(P.packed, a, end, (g0 + c + 1) * P.nmax);
      if (cr == c) {
        lo = a;
        hi = b;
      }
      a = b;
    }
  }
  if (bin_live) {
    const CellGeo q = cell_geo(s, g0 + cb);
    const bool xa = bb < kx;
    const int j = xa ? bb : bb - kx, kk = xa ? kx : int(s.kq);
    const double* cdf = xa ? P.cdf_x + q.ox : P.cdf_q + q.oq;
    const void* bins = xa ? P.x_bins : P.q_bins;
    const int64_t boff = xa ? q.bx : q.bq;
    const double c0 = cdf[j];
    double inv_d = 0.0, dx = 0.0;
    if (j + 1 < kk) {
      inv_d = 1.0 / (cdf[j + 1] - c0 + kEpsilon);
      dx = ld<Real>(bins, boff + j + 1) - ld<Real>(bins, boff + j);
    }
    cd[lane] = make_double2(c0, inv_d);
    dxs[lane] = dx;
  }
  __builtin_amdgcn_wave_barrier();
  const int64_t rounds = wave_max_i64(row_lane ? (hi - lo + R - 1) / R : 0);
  const int cl = row_lane ? cr * L : 0;  // first bin lane of this row's cell
  // A round's intervals, uniforms and cotangents are loaded one round ahead
  // and its packed slot two rounds ahead (the gathers at the slot then have
  // a whole round to arrive).
  int64_t pn = 0;  // packed slot of the next round's row
  int jxn = 0, jqn = 0;
  float uxn = 0.0f, uqn = 0.0f;
  double gxn = 0.0, gqn = 0.0;
  if (row_lane && lo + ir < hi) {
    const int64_t row = lo + ir;
    const int64_t p = P.packed[row];
    jxn = P.interval_x[p];
    jqn = P.interval_q[p];
    uxn = P.u_x[p];
    uqn = P.u_q[p];
    gxn = ld<Real>(P.grad, 2 * row);
    gqn = ld<Real>(P.grad, 2 * row + 1);
    if (row + R < hi) pn = P.packed[row + R];
  }
  double acc = 0.0;
  for (int64_t k = 0; k < rounds; ++k) {
    const int64_t row = lo + k * R + ir;
    const bool live = row_lane && row < hi;
    const int jx = live ? jxn : 0, jq = live ? jqn : 0;
    const double ux = live ? double(uxn) : 0.0, uq = live ? double(uqn) : 0.0;
    const double gxb = gxn, gqb = gqn;
    if (row_lane && row + R < hi) {
      const int64_t p = pn;
      jxn = P.interval_x[p];
      jqn = P.interval_q[p];
      uxn = P.u_x[p];
      uqn = P.u_q[p];
      gxn = ld<Real>(P.grad, 2 * (row + R));
      gqn = ld<Real>(P.grad, 2 * (row + R) + 1);
      if (row + 2 * R < hi) pn = P.packed[row + 2 * R];
    }
    const int la = cl + jx, qa = cl + kx + jq;  // bin lanes of this row's first nodes
    double upx, ax, bxm, idx, upq, aq, bqm, idq;
    {  // x: the factors of ivjp_row
      const double2 ci = cd[la];
      const double c0 = ci.x, inv_d = ci.y;
      const double t = ux - c0;
      upx = gxb * dxs[la];
      ax = __builtin_fma(t * inv_d, inv_d, -inv_d);
      bxm = -(upx * t * inv_d);
      idx = inv_d;
    }
    {  // q
      const double2 ci = cd[qa];
      const double c0 = ci.x, inv_d = ci.y;
      const double t = uq - c0;
      upq = gqb * dxs[qa];
      aq = __builtin_fma(t * inv_d, inv_d, -inv_d);
      bqm = -(upq * t * inv_d);
      idq = inv_d;
    }
    T[lane] = 0ull;
    __builtin_amdgcn_wave_barrier();
    if (live) {
      atomicOr(&T[la], 1ull << lane);
      atomicOr(&T[qa], 1ull << lane);
    }
    __builtin_amdgcn_wave_barrier();
    uint64_t t = T[lane];
    const uint64_t up = shfl_up1_u64(t);
    t |= (bin_live && bb != 0 && bb != kx) ? up : 0ull;  // rows with value j - 1 of this axis
    const int len = popc64(t);
    const int off = wave_inclusive_scan(len) - len;  // at most 4 rows' worth in all
    __builtin_amdgcn_wave_barrier();
    T[lane] = t;
    offs[lane] = off;
    __builtin_amdgcn_wave_barrier();
    if (live) {
      const uint64_t lt = lanemask_lt();
      ent[offs[la] + popc64(T[la] & lt)] = make_double2(upx, ax);
      ent[offs[la + 1] + popc64(T[la + 1] & lt)] = make_double2(bxm, idx);
      ent[offs[qa] + popc64(T[qa] & lt)] = make_double2(upq, aq);
      ent[offs[qa + 1] + popc64(T[qa + 1] & lt)] = make_double2(bqm, idq);
    }
    __builtin_amdgcn_wave_barrier();
    for (int i0 = 0; ballot(i0 < len) != 0ull; i0 += 4) {  // until every list is summed
      double2 f[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        // a lane past its list's end reads entry 0 with the others: one
        // broadcast address instead of a distinct (unused) entry per lane
        f[u] = ent[i0 + u < len ? off + i0 + u : 0];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const double nxt = __builtin_fma(f[u].x, f[u].y, acc);
        acc = i0 + u < len ? nxt : acc;
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
  if (bin_live) {
    const CellGeo q = cell_geo(s, g0 + cb);
    if (bb < kx) P.gcdf_x[q.ox + bb] = acc;
    else P.gcdf_q[q.oq + bb - kx] = acc;
  }
}

// CDF and density adjoints of one curve.  R > 0: the curve's k <= R values
// are loaded into registers first and the per-node gradient stays there
// (gr_mem is not touched); R == 0 keeps it in gr_mem.  The same expressions
// in the same order either way.
template <typename Real, int R>
__device__ __forceinline__ void curve_vjp_chain(const double* gc, const void* bins,
                                                int64_t bins_off, const void* xsec, int64_t off,
                                                int64_t k, double acc, double n, double* gr_mem,
                                                Real* gx) {
  constexpr int A = R > 0 ? R : 1;
  double gcr[A], br[A], xr[A], grr[A];
  if constexpr (R > 0) {
#pragma unroll
    for (int t = 0; t < R; ++t) {
      gcr[t] = t < k ? gc[t] : 0.0;
      br[t] = t < k ? ld<Real>(bins, bins_off + t) : 0.0;
      xr[t] = t < k ? ld<Real>(xsec, off + t) : 0.0;
      grr[t] = 0.0;
    }
  } else {
    for (int64_t t = 0; t < k; ++t) gr_mem[t] = 0.0;
  }
  auto GC = [&](int64_t t) -> double {
    if constexpr (R > 0) return gcr[t]; else return gc[t];
  };
  auto Bn = [&](int64_t t) -> double {
    if constexpr (R > 0) return br[t]; else return ld<Real>(bins, bins_off + t);
  };
  auto X = [&](int64_t t) -> double {
    if constexpr (R > 0) return xr[t]; else return ld<Real>(xsec, off + t);
  };
  auto GR = [&](int64_t t) -> double& {
    if constexpr (R > 0) return grr[t]; else return gr_mem[t];
  };
  const int64_t kl = R > 0 ? R : k;  // loop bound, a constant with registers
  // CDF adjoint: cpp/loits_core.cpp:355-365
  double suffix = 0.0;
#pragma unroll
  for (int64_t t = kl - 1; t > 0; --t) {
    if (t > k - 1) continue;
    suffix += GC(t);
    const double h = 0.5 * (Bn(t) - Bn(t - 1)) * acc;
    GR(t - 1) = __builtin_fma(h, suffix, GR(t - 1));
    GR(t) = __builtin_fma(h, suffix, GR(t));
  }
  // density adjoint: cpp/loits_core.cpp:416-431 (grad_norm fusion pattern of
  // the reference build: only an odd trailing term is fused)
  const double inv_n = 1.0 / n;
  const double inv_n2 = inv_n * inv_n;
  double grad_norm = 0.0;
#pragma unroll
  for (int64_t t = 0; t < kl; ++t) {
    if (t >= k) break;
    const double term = GR(t) * X(t);
    grad_norm = ((k & 1) && t == k - 1) ? __builtin_fma(-term, inv_n2, grad_norm)
                                        : grad_norm - term * inv_n2;
  }
  double prev = GR(0) * inv_n;
#pragma unroll
  for (int64_t t = 1; t < kl; ++t) {
    if (t >= k) break;
    const double h = 0.5 * (Bn(t) - Bn(t - 1));
    const double cur = GR(t) * inv_n;
    gx[t - 1] = Real(__builtin_fma(h, grad_norm, prev));
    prev = __builtin_fma(h, grad_norm, cur);
  }
  gx[k - 1] = Real(prev);
}

template <typename Real>
__global__ void __launch_bounds__(256) curve_vjp_kernel(VjpParams P) {
  const loits_hip_shape& s = P.s;
  const int64_t nxc = s.batch * s.ny * s.nx, nqc = s.batch * s.nx * s.ny;
  const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= nxc + nqc) return;
  const bool is_x = c < nxc;
  const int64_t cc = is_x ? c : c - nxc;
  int64_t b, ix, iy, k, bins_off;
  const void *bins, *xsec;
  const double *gc, *norm;
  void* out;
  double* gr;
  if (is_x) {
    b = cc / (s.ny * s.nx);
    const int64_t r = cc - b * s.ny * s.nx;
    iy = r / s.nx;
    ix = r - iy * s.nx;
    k = s.kx;
    bins = P.x_bins;
    bins_off = (b * s.nx + ix) * k;
    xsec = P.xsec_x;
    gc = P.gcdf_x + cc * k;
    norm = P.norm_x;
    out = P.grad_x;
    gr = P.grho + cc * k;
  } else {
    b = cc / (s.nx * s.ny);
    const int64_t r = cc - b * s.nx * s.ny;
    ix = r / s.ny;
    iy = r - ix * s.ny;
    k = s.kq;
    bins = P.q_bins;
    bins_off = (b * s.ny + iy) * k;
    xsec = P.xsec_q;
    gc = P.gcdf_q + cc * k;
    norm = P.norm_q;
    out = P.grad_q;
    gr = P.grho + nxc * s.kx + cc * k;
  }
  const int64_t off = cc * k;
  const double acc = P.acc[(b * s.nx + ix) * s.ny + iy] ? 1.0 : 0.0;
  Real* gx = static_cast<Real*>(out) + off;
  if (k <= kRegNodes)
    curve_vjp_chain<Real, kRegNodes>(gc, bins, bins_off, xsec, off, k, acc, norm[cc], gr, gx);
  else
    curve_vjp_chain<Real, 0>(gc, bins, bins_off, xsec, off, k, acc, norm[cc], gr, gx);
}

int64_t align16(int64_t v) { return (v + 15) & ~int64_t(15); }

struct FwdWorkspace {
  uint32_t* thr_x;
  uint32_t* thr_q;
  double* m_x;
  double* m_q;
  uint8_t* mono;
  uint64_t* tile_status;
  uint32_t* tile_counter;
  uint32_t* unsafe;  // in the cleared header after the tile counter
  int64_t* row0;
  int64_t bytes;
};

FwdWorkspace fwd_workspace(const loits_hip_shape& s, int64_t nmax, void* base) {
  const int64_t nxe = s.batch * s.ny * s.nx * s.kx, nqe = s.batch * s.nx * s.ny * s.kq;
  const int64_t curves = s.batch * s.cells * 2;
  const int64_t slots = s.batch * s.cells * nmax;
  const int64_t tiles = (slots + kTileSlots - 1) / kTileSlots;
  uint8_t* p = static_cast<uint8_t*>(base);
  FwdWorkspace w{};
  int64_t off = 0;
  w.tile_counter = reinterpret_cast<uint32_t*>(p + off);
  w.unsafe = w.tile_counter + 1;
  off += 64;
  w.tile_status = reinterpret_cast<uint64_t*>(p + off);
  off += align16(8 * (tiles > 0 ? tiles : 1));
  w.thr_x = reinterpret_cast<uint32_t*>(p + off);
  off += align16(4 * nxe);
  w.thr_q = reinterpret_cast<uint32_t*>(p + off);
  off += align16(4 * nqe);
  w.m_x = reinterpret_cast<double*>(p + off);
  off += align16(8 * nxe);
  w.m_q = reinterpret_cast<double*>(p + off);
  off += align16(8 * nqe);
  w.mono = p + off;
  off += align16(curves);
  w.row0 = reinterpret_cast<int64_t*>(p + off);
  off += align16(8 * (s.batch * s.cells + 1));
  w.bytes = off;
  return w;
}

__global__ void zero_words(uint64_t* p, int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

int check_shape(const loits_hip_shape& s) {
  if (s.batch < 1 || s.nx < 1 || s.ny < 1 || s.kx < 2 || s.kq < 2 || s.cells != s.nx * s.ny)
    return fail(LOITS_HIP_EINVAL, "shape: need batch, nx, ny >= 1, kx, kq >= 2, cells = nx*ny");
  if (s.kx > 32767 || s.kq > 32767)
    return fail(LOITS_HIP_EUNSUPPORTED, "shape: the int16 interval caps kx, kq at 32767");
  return LOITS_HIP_OK;
}

unsigned blocks_for(int64_t n, int threads) { return unsigned((n + threads - 1) / threads); }

}  // namespace
}  // namespace qloits

using namespace qloits;

extern "C" {

int loits_hip_allocate(const void* weights, int dtype, int64_t n, int64_t n_events,
                       int64_t* counts, int64_t* nmax_active, void* stream) {
  if (n < 0 || (n > 0 && (!weights || !counts)) || !nmax_active)
    return fail(LOITS_HIP_EINVAL, "allocate: bad arguments");
  if (n_events < 0) return fail(LOITS_HIP_EINVAL, "allocate: n_events must be >= 0");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == LOITS_HIP_F64)
    hipLaunchKernelGGL(alloc_kernel<double>, dim3(1), dim3(1024), 0, s, weights, n, n_events,
                       counts, nmax_active);
  else if (dtype == LOITS_HIP_F32)
    hipLaunchKernelGGL(alloc_kernel<float>, dim3(1), dim3(1024), 0, s, weights, n, n_events,
                       counts, nmax_active);
  else
    return fail(LOITS_HIP_EINVAL, "allocate: bad dtype");
  return check_hip(hipGetLastError(), "allocate launch");
}

int64_t loits_hip_forward_workspace_size(const loits_hip_shape* shape, int64_t nmax) {
  if (!shape || check_shape(*shape) != LOITS_HIP_OK || nmax < 0) return -1;
  return fwd_workspace(*shape, nmax, nullptr).bytes;
}

}  // extern "C"

namespace {

// dev_na: NULL (a->nmax, a->active are the sizes) or the device array the
// allocation wrote (a->nmax, a->active are then capacities).
int forward_impl(const loits_hip_forward_args* a, const int64_t* dev_na, void* stream) {
  if (!a) return fail(LOITS_HIP_EINVAL, "forward: args is NULL");
  const loits_hip_shape& s = a->shape;
  int rc = check_shape(s);
  if (rc != LOITS_HIP_OK) return rc;
  if (a->dtype != LOITS_HIP_F32 && a->dtype != LOITS_HIP_F64)
    return fail(LOITS_HIP_EINVAL, "forward: bad dtype");
  if (!a->x_bins || !a->xsec_x || !a->q_bins || !a->xsec_q || !a->acceptance || !a->counts ||
      !a->norm_x || !a->norm_q || !a->cdf_x || !a->cdf_q || !a->workspace || !a->valid)
    return fail(LOITS_HIP_EINVAL, "forward: a required pointer is NULL");
  if (a->nmax < 0 || a->active < 0) return fail(LOITS_HIP_EINVAL, "forward: bad nmax/active");
  const int64_t slots = s.batch * s.cells * a->nmax;
  if (slots > 0 && (!a->u_x || !a->u_q || !a->interval_x || !a->interval_q))
    return fail(LOITS_HIP_EINVAL, "forward: u/interval outputs are NULL");
  auto misaligned = [](const void* p, uintptr_t a) { return (reinterpret_cast<uintptr_t>(p) & (a - 1)) != 0; };
  if (slots > 0 && (misaligned(a->u_x, 16) || misaligned(a->u_q, 16) ||
                    misaligned(a->interval_x, 8) || misaligned(a->interval_q, 8)))
    return fail(LOITS_HIP_EINVAL,
                "forward: u_x/u_q must be 16-byte and interval_x/interval_q 8-byte aligned");
  if (a->active > 0 && (!a->events || !a->packed))
    return fail(LOITS_HIP_EINVAL, "forward: events/packed outputs are NULL");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const FwdWorkspace w = fwd_workspace(s, a->nmax, a->workspace);
  const int64_t tiles = (slots + kTileSlots - 1) / kTileSlots;
  const bool f32 = a->dtype == LOITS_HIP_F32;

  // zero: tile counter, safety flag, look-back status words, valid[0..1]
  const int64_t zwords = (64 + align16(8 * (tiles > 0 ? tiles : 1))) / 8;
  hipLaunchKernelGGL(zero_words, dim3(blocks_for(zwords, 256)), dim3(256), 0, st,
                     reinterpret_cast<uint64_t*>(w.tile_counter), zwords);
  hipLaunchKernelGGL(zero_words, dim3(1), dim3(64), 0, st,
                     reinterpret_cast<uint64_t*>(a->valid), int64_t(2));
  rc = check_hip(hipGetLastError(), "forward zero launch");
  if (rc != LOITS_HIP_OK) return rc;
  CurveParams C{s,          a->x_bins, a->xsec_x, a->q_bins, a->xsec_q, a->acceptance,
                a->norm_x,  a->norm_q, a->cdf_x,  a->cdf_q,  w.thr_x,   w.thr_q,
                w.m_x,      w.m_q,     w.mono,    w.unsafe};
  const int64_t curves = 2 * s.batch * s.cells;
  if (f32)
    hipLaunchKernelGGL(curve_kernel<float>, dim3(blocks_for(curves, 64)), dim3(64), 0, st, C);
  else
    hipLaunchKernelGGL(curve_kernel<double>, dim3(blocks_for(curves, 64)), dim3(64), 0, st, C);
  rc = check_hip(hipGetLastError(), "forward curve launch");
  if (rc != LOITS_HIP_OK) return rc;
  hipLaunchKernelGGL(row_start_kernel, dim3(1), dim3(1024), 0, st, a->counts,
                     s.batch * s.cells, w.row0);
  rc = check_hip(hipGetLastError(), "forward row-start launch");
  if (rc != LOITS_HIP_OK) return rc;
  if (tiles > 0) {
    if (tiles > INT32_MAX) return fail(LOITS_HIP_EUNSUPPORTED, "forward: too many slots");
    SampleParams S{s,          a->nmax,    slots,      tiles,       a->seed,
                   2 * a->sequence,        2 * a->sequence + 1,     a->x_bins,
                   a->q_bins,  a->counts,  a->cdf_x,   a->cdf_q,    w.thr_x,
                   w.thr_q,    w.m_x,      w.m_q,      w.mono,      a->u_x,
                   a->u_q,     a->interval_x,          a->interval_q,
                   a->events,  a->packed,  w.tile_status,           w.tile_counter,
                   a->valid,   dev_na,     a->active,  w.unsafe,    w.row0};
    const int64_t direct_tiles = (slots + kDirectSlots - 1) / kDirectSlots;
    if (direct_tiles > INT32_MAX) return fail(LOITS_HIP_EUNSUPPORTED, "forward: too many slots");
    if (f32) {
      hipLaunchKernelGGL((sample_kernel<float, true>), dim3(unsigned(direct_tiles)),
                         dim3(kDirectThreads), 0, st, S);
      hipLaunchKernelGGL((sample_kernel<float, false>), dim3(unsigned(tiles)), dim3(kTileThreads),
                         0, st, S);
    } else {
      hipLaunchKernelGGL((sample_kernel<double, true>), dim3(unsigned(direct_tiles)),
                         dim3(kDirectThreads), 0, st, S);
      hipLaunchKernelGGL((sample_kernel<double, false>), dim3(unsigned(tiles)),
                         dim3(kTileThreads), 0, st, S);
    }
    rc = check_hip(hipGetLastError(), "forward sample launch");
    if (rc != LOITS_HIP_OK) return rc;
  }
  if (a->row_offsets) {
    const int64_t gcells = s.batch * s.cells;
    hipLaunchKernelGGL(row_offsets_kernel, dim3(blocks_for(gcells + 1, 256)), dim3(256), 0, st,
                       a->packed, a->valid, a->nmax, gcells, a->row_offsets, dev_na, w.unsafe,
                       w.row0);
    rc = check_hip(hipGetLastError(), "forward row_offsets launch");
  }
  return rc;
}

}  // namespace

extern "C" {

int loits_hip_forward(const loits_hip_forward_args* a, void* stream) {
  return forward_impl(a, nullptr, stream);
}

int loits_hip_forward_bounded(const loits_hip_forward_args* a, const int64_t* nmax_active,
                              void* stream) {
  if (!nmax_active) return fail(LOITS_HIP_EINVAL, "forward_bounded: nmax_active is NULL");
  return forward_impl(a, nmax_active, stream);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free forward.  Its workspace is `scratch`, a file-scope device
// array: every GPU that loads the library holds one copy until the process
// ends.  Host code never passes `scratch` itself to a kernel (on the host the
// name is only a stand-in for the symbol); the device address of the current
// GPU's copy comes from hipGetSymbolAddress, once per device, and is kept.
namespace qloits {
namespace {

constexpr int64_t kNoAllocScratchBytes = LOITS_HIP_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxScratchDevices] = {};

// Device address of `scratch` on the GPU that runs `st` (the current device
// for the legacy default stream).
int scratch_address(hipStream_t st, void** out) {
  int dev = -1;
  int rc = check_hip(st ? hipStreamGetDevice(st, &dev) : hipGetDevice(&dev),
                     "forward_noalloc: device of the stream");
  if (rc != LOITS_HIP_OK) return rc;
  if (dev < 0 || dev >= kMaxScratchDevices)
    return fail(LOITS_HIP_EUNSUPPORTED, "forward_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    int cur = -1;
    rc = check_hip(hipGetDevice(&cur), "forward_noalloc: hipGetDevice");
    if (rc != LOITS_HIP_OK) return rc;
    if (cur != dev) {
      rc = check_hip(hipSetDevice(dev), "forward_noalloc: hipSetDevice");
      if (rc != LOITS_HIP_OK) return rc;
    }
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    rc = check_hip(e, "forward_noalloc: hipGetSymbolAddress(scratch)");
    if (rc != LOITS_HIP_OK) return rc;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return LOITS_HIP_OK;
}

}  // namespace
}  // namespace qloits

extern "C" {

int loits_hip_forward_noalloc(const loits_hip_forward_args* a, void* stream) {
  if (!a) return fail(LOITS_HIP_EINVAL, "forward_noalloc: args is NULL");
  if (a->workspace)
    return fail(LOITS_HIP_EINVAL,
                "forward_noalloc: workspace must be NULL (the library's own scratch is used)");
  int rc = check_shape(a->shape);
  if (rc != LOITS_HIP_OK) return rc;
  if (a->nmax < 0) return fail(LOITS_HIP_EINVAL, "forward_noalloc: bad nmax");
  const int64_t need = fwd_workspace(a->shape, a->nmax, nullptr).bytes;
  if (need > kNoAllocScratchBytes)
    return fail(LOITS_HIP_EINVAL,
                "forward_noalloc: this shape needs " + std::to_string(need) +
                    " bytes of workspace, more than the " +
                    std::to_string(kNoAllocScratchBytes) +
                    "-byte scratch of the allocation-free entry point; use loits_hip_forward");
  void* ws = nullptr;
  rc = scratch_address(static_cast<hipStream_t>(stream), &ws);
  if (rc != LOITS_HIP_OK) return rc;
  loits_hip_forward_args b = *a;
  b.workspace = ws;
  return forward_impl(&b, nullptr, stream);
}

int64_t loits_hip_backward_workspace_size(const loits_hip_shape* shape, int64_t nmax) {
  (void)nmax;
  if (!shape || check_shape(*shape) != LOITS_HIP_OK) return -1;
  const int64_t nxe = shape->batch * shape->ny * shape->nx * shape->kx;
  const int64_t nqe = shape->batch * shape->nx * shape->ny * shape->kq;
  return 2 * (align16(8 * nxe) + align16(8 * nqe));
}

}  // extern "C"

namespace {

int backward_impl(const loits_hip_backward_args* a, const int64_t* dev_nmax,
                  const int64_t* dev_nvalid, void* stream) {
  if (!a) return fail(LOITS_HIP_EINVAL, "backward: args is NULL");
  const loits_hip_shape& s = a->shape;
  int rc = check_shape(s);
  if (rc != LOITS_HIP_OK) return rc;
  if (a->dtype != LOITS_HIP_F32 && a->dtype != LOITS_HIP_F64)
    return fail(LOITS_HIP_EINVAL, "backward: bad dtype");
  if (!a->x_bins || !a->xsec_x || !a->q_bins || !a->xsec_q || !a->acceptance || !a->norm_x ||
      !a->norm_q || !a->cdf_x || !a->cdf_q || !a->grad_xsec_x || !a->grad_xsec_q ||
      !a->workspace)
    return fail(LOITS_HIP_EINVAL, "backward: a required pointer is NULL");
  if (a->nvalid > 0 && (!a->grad_events || !a->packed || !a->u_x || !a->u_q || !a->interval_x ||
                        !a->interval_q))
    return fail(LOITS_HIP_EINVAL, "backward: state pointers are NULL");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int64_t nxe = s.batch * s.ny * s.nx * s.kx, nqe = s.batch * s.nx * s.ny * s.kq;
  uint8_t* ws = static_cast<uint8_t*>(a->workspace);
  double* gcdf_x = reinterpret_cast<double*>(ws);
  double* gcdf_q = reinterpret_cast<double*>(ws + align16(8 * nxe));
  double* grho = reinterpret_cast<double*>(ws + align16(8 * nxe) + align16(8 * nqe));
  VjpParams V{s,         a->nmax,      a->nvalid,  a->grad_events, a->x_bins, a->xsec_x,
              a->q_bins, a->xsec_q,    a->acceptance, a->norm_x,   a->norm_q, a->cdf_x,
              a->cdf_q,  a->u_x,       a->u_q,     a->interval_x,  a->interval_q,
              a->packed, gcdf_x,       gcdf_q,     grho,           a->grad_xsec_x,
              a->grad_xsec_q, dev_nmax, dev_nvalid};
  V.row_offsets = a->row_offsets;
  const bool f32 = a->dtype == LOITS_HIP_F32;
  const int64_t gcells = s.batch * s.cells;
  if (s.kx + s.kq <= kWave) {
    const int64_t per_wave = kWave / (s.kx + s.kq) < 2 ? kWave / (s.kx + s.kq) : 2;
    const int64_t waves = (gcells + per_wave - 1) / per_wave;
    if (f32)
      hipLaunchKernelGGL(ivjp_cells_kernel<float>, dim3(blocks_for(waves, 4)), dim3(256), 0, st,
                         V);
    else
      hipLaunchKernelGGL(ivjp_cells_kernel<double>, dim3(blocks_for(waves, 4)), dim3(256), 0, st,
                         V);
  } else if (s.kx <= kWave && s.kq <= kWave) {
    if (f32)
      hipLaunchKernelGGL(ivjp_wave_kernel<float>, dim3(blocks_for(gcells, 4)), dim3(256), 0, st,
                         V);
    else
      hipLaunchKernelGGL(ivjp_wave_kernel<double>, dim3(blocks_for(gcells, 4)), dim3(256), 0, st,
                         V);
  } else if (f32) {
    hipLaunchKernelGGL(ivjp_kernel<float>, dim3(blocks_for(gcells, 256)), dim3(256), 0, st, V);
  } else {
    hipLaunchKernelGGL(ivjp_kernel<double>, dim3(blocks_for(gcells, 256)), dim3(256), 0, st, V);
  }
  rc = check_hip(hipGetLastError(), "backward interpolation launch");
  if (rc != LOITS_HIP_OK) return rc;
  const int64_t curves = 2 * gcells;
  if (f32)
    hipLaunchKernelGGL(curve_vjp_kernel<float>, dim3(blocks_for(curves, 64)), dim3(64), 0, st,
                       V);
  else
    hipLaunchKernelGGL(curve_vjp_kernel<double>, dim3(blocks_for(curves, 64)), dim3(64), 0, st,
                       V);
  return check_hip(hipGetLastError(), "backward curve launch");
}

}  // namespace

extern "C" {

int loits_hip_backward(const loits_hip_backward_args* a, void* stream) {
  return backward_impl(a, nullptr, nullptr, stream);
}

int loits_hip_backward_bounded(const loits_hip_backward_args* a, const int64_t* nmax_active,
                               const int64_t* valid, void* stream) {
  if (!nmax_active || !valid)
    return fail(LOITS_HIP_EINVAL, "backward_bounded: nmax_active or valid is NULL");
  return backward_impl(a, nmax_active, valid, stream);
}

}  // extern "C"
