// loits_device.hpp -- device building blocks of the gfx950 LOITS sampler.
//
// Philox4x32-10 (rng/philox.hpp:17-61), the exact integer form of the
// reference's "u >= C_t" comparison, wave64 helpers and the layout of the
// per-curve tables the batched kernels stage through LDS.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "loits_hip.h"

namespace qloits {

constexpr double kEpsilon = 1e-5;  // cpp/loits_core.hpp:15
constexpr int kWave = 64;

// ---------------------------------------------------------------------------
// Philox4x32-10, rng/philox.hpp:17-56.  Counter (block lo, block hi,
// stream lo, stream hi), key (seed lo, seed hi).  Each 32x32->64 product is a
// single v_mad_u64_u32; the key schedule is wave-uniform (SALU).
struct Words4 {
  uint32_t w0, w1, w2, w3;
};

__host__ __device__ __forceinline__ Words4 philox4x32_10(uint64_t block, uint64_t stream,
                                                         uint64_t seed) {
  uint32_t c0 = static_cast<uint32_t>(block), c1 = static_cast<uint32_t>(block >> 32);
  uint32_t c2 = static_cast<uint32_t>(stream), c3 = static_cast<uint32_t>(stream >> 32);
  uint32_t k0 = static_cast<uint32_t>(seed), k1 = static_cast<uint32_t>(seed >> 32);
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = static_cast<uint64_t>(0xD2511F53u) * c0;
    const uint64_t p1 = static_cast<uint64_t>(0xCD9E8D57u) * c2;
    const uint32_t n0 = static_cast<uint32_t>(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n2 = static_cast<uint32_t>(p0 >> 32) ^ c3 ^ k1;
    c1 = static_cast<uint32_t>(p1);
    c3 = static_cast<uint32_t>(p0);
    c0 = n0;
    c2 = n2;
    k0 += 0x9E3779B9u;  // the bump after the last round is never used
    k1 += 0xBB67AE85u;
  }
  return {c0, c1, c2, c3};
}

// a ^ b ^ c in one VALU instruction (gfx950 v_bitop3_b32, truth table 0x96;
// c wave-uniform, held in an SGPR)
__host__ __device__ __forceinline__ uint32_t xor3_su(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t r;
  asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0x96" : "=v"(r) : "v"(a), "v"(b), "s"(c));
  return r;
#else
  return a ^ b ^ c;
#endif
}

// Device Philox4x32-10 for a wave-uniform stream and seed: round 0 multiplies
// the uniform stream word on the scalar unit; the other rounds fold the two
// XORs into one v_bitop3_b32.
__host__ __device__ __forceinline__ Words4 philox4x32_10_uniform(uint64_t block, uint64_t stream,
                                                        uint64_t seed) {
  uint32_t c0 = static_cast<uint32_t>(block), c1 = static_cast<uint32_t>(block >> 32);
  uint32_t c2 = static_cast<uint32_t>(stream), c3 = static_cast<uint32_t>(stream >> 32);
  uint32_t k0 = static_cast<uint32_t>(seed), k1 = static_cast<uint32_t>(seed >> 32);
  {
    const uint64_t p0 = static_cast<uint64_t>(0xD2511F53u) * c0;
    const uint64_t p1 = static_cast<uint64_t>(0xCD9E8D57u) * c2;
    const uint32_t n0 = static_cast<uint32_t>(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n2 = static_cast<uint32_t>(p0 >> 32) ^ c3 ^ k1;
    c1 = static_cast<uint32_t>(p1);
    c3 = static_cast<uint32_t>(p0);
    c0 = n0;
    c2 = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
#pragma unroll
  for (int r = 1; r < 10; ++r) {
    const uint64_t p0 = static_cast<uint64_t>(0xD2511F53u) * c0;
    const uint64_t p1 = static_cast<uint64_t>(0xCD9E8D57u) * c2;
    const uint32_t n0 = xor3_su(static_cast<uint32_t>(p1 >> 32), c1, k0);
    const uint32_t n2 = xor3_su(static_cast<uint32_t>(p0 >> 32), c3, k1);
    c1 = static_cast<uint32_t>(p1);
    c3 = static_cast<uint32_t>(p0);
    c0 = n0;
    c2 = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return {c0, c1, c2, c3};
}

__host__ __device__ __forceinline__ uint32_t word_of(const Words4& w, int e) {
  return e == 0 ? w.w0 : e == 1 ? w.w1 : e == 2 ? w.w2 : w.w3;
}

// rng/philox.hpp:59-61: u = (bits >> 8) * 2^-24.  The sampler works on the
// 24-bit integer k = bits >> 8 and forms u only where it needs the value.
__host__ __device__ __forceinline__ float uniform_float(uint32_t bits) {
  return static_cast<float>(bits >> 8) * 0x1.0p-24f;
}

// ---------------------------------------------------------------------------
// Exact integer form of the comparison "u >= C" for u = k * 2^-24,
// k in [0, 2^24): u >= C  <=>  k >= threshold(C), with
//   NaN            -> 0xFFFFFFFF (never; NaN compares false)
//   C <= 0         -> 0          (always)
//   C * 2^24 >= 2^24 (incl. +inf) -> 2^24 (never)
//   otherwise      -> ceil(C * 2^24)  (scaling by 2^24 and ceil are exact)
// The reference's count-based interval (cpp/loits_core.cpp:12-18) is then
// count(k) = #{t : T_t <= k}.
constexpr uint32_t kNever = 0xFFFFFFFFu;

__host__ __device__ __forceinline__ uint32_t threshold_of(double c) {
  if (c != c) return kNever;
  if (c <= 0.0) return 0u;
  const double s = c * 16777216.0;
  if (s >= 16777216.0) return 16777216u;
  return static_cast<uint32_t>(__builtin_ceil(s));
}

// ---------------------------------------------------------------------------
// Wave64 helpers.
__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t lanemask_lt() {
  const int l = lane_id();
  return l == 0 ? 0ull : (~0ull >> (64 - l));
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

__device__ __forceinline__ int popc64(uint64_t v) { return __popcll(v); }

// Wave64 prefix sums with DPP row shifts and row broadcasts (no LDS traffic):
// Hillis-Steele within each row of 16 lanes, then row 0 / row 2 totals into
// rows 1 / 3 (row_bcast:15) and the row-1 total into rows 2-3 (row_bcast:31).
// Lanes without a source read 0, the identity.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ uint32_t dpp_shift_u32(uint32_t v) {
  return static_cast<uint32_t>(
      __builtin_amdgcn_update_dpp(0, static_cast<int>(v), CTRL, ROW_MASK, 0xF, false));
}

template <int CTRL, int ROW_MASK, typename T>
__device__ __forceinline__ T dpp_shift(T v) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, dpp_shift_u32<CTRL, ROW_MASK>(__builtin_bit_cast(uint32_t, v)));
  } else {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint64_t lo = dpp_shift_u32<CTRL, ROW_MASK>(static_cast<uint32_t>(b));
    const uint64_t hi = dpp_shift_u32<CTRL, ROW_MASK>(static_cast<uint32_t>(b >> 32));
    return __builtin_bit_cast(T, lo | (hi << 32));
  }
}

template <typename T>
__device__ __forceinline__ T wave_inclusive_scan(T v) {
  v = v + dpp_shift<0x111, 0xF>(v);  // row_shr:1
  v = v + dpp_shift<0x112, 0xF>(v);  // row_shr:2
  v = v + dpp_shift<0x114, 0xF>(v);  // row_shr:4
  v = v + dpp_shift<0x118, 0xF>(v);  // row_shr:8
  v = v + dpp_shift<0x142, 0xA>(v);  // row_bcast:15 -> rows 1, 3
  v = v + dpp_shift<0x143, 0xC>(v);  // row_bcast:31 -> rows 2, 3
  return v;
}

// Streaming (non-temporal) 16-byte store and load: events and cotangents are
// written or read once per launch, and keeping them out of the caches leaves
// room for the curve tables and the other operand.
typedef float f32x4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void store_stream4(float* p, float a, float b, float c, float d) {
  const f32x4_t v = {a, b, c, d};
  __builtin_nontemporal_store(v, reinterpret_cast<f32x4_t*>(p));
}

__device__ __forceinline__ float4 load_stream4(const float* p) {
  const f32x4_t v = __builtin_nontemporal_load(reinterpret_cast<const f32x4_t*>(p));
  return make_float4(v.x, v.y, v.z, v.w);
}

// Inclusive running maximum over the lanes of a wave (0 is the identity).
__device__ __forceinline__ uint32_t wave_inclusive_max(uint32_t v) {
  uint32_t t;
  t = dpp_shift_u32<0x111, 0xF>(v);
  v = t > v ? t : v;
  t = dpp_shift_u32<0x112, 0xF>(v);
  v = t > v ? t : v;
  t = dpp_shift_u32<0x114, 0xF>(v);
  v = t > v ? t : v;
  t = dpp_shift_u32<0x118, 0xF>(v);
  v = t > v ? t : v;
  t = dpp_shift_u32<0x142, 0xA>(v);
  v = t > v ? t : v;
  t = dpp_shift_u32<0x143, 0xC>(v);
  v = t > v ? t : v;
  return v;
}

template <typename T>
__device__ __forceinline__ T wave_lane(T v, int lane) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(
        T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
  } else {
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint64_t lo = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(b), lane));
    const uint64_t hi =
        static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(b >> 32), lane));
    return __builtin_bit_cast(T, lo | (hi << 32));
  }
}

// Sum over the 64 lanes (every lane gets it).
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  return wave_lane(wave_inclusive_scan(v), kWave - 1);
}

// ---------------------------------------------------------------------------
// Batched path: device descriptors and per-instance tables.

struct InstanceDesc {
  int64_t node_off;
  int64_t n_nodes;
  int64_t events;
  int64_t out_off;
  uint64_t counter_base;
  uint64_t stream;
  int64_t table_off;     // byte offset of this instance's tables
  int32_t accept;
  int32_t guide_log2;    // guide cells = 1 << guide_log2
  int32_t first_chunk;   // chunks of an instance are consecutive
  int32_t n_chunks;
  int32_t global_tables; // 1: per-bin tables too large for LDS stay in global memory
  int32_t vjp_ranges;    // 0: VJP tables and histograms in LDS; R > 0: R bin ranges
  int32_t fast_log2;     // > 0: float32 cell-table path with 2^fast_log2 cells (fast_tables.hpp)
  int32_t reserved;
};

// A chunk is the work of one work-group; each of its waves owns a contiguous
// run of Philox blocks (a "segment") and compacts it independently.
struct ChunkDesc {
  int64_t slot_lo;      // first slot (relative to the instance)
  int64_t slot_hi;      // one past the last slot
  int64_t mask_off;     // first 64-bit word of this chunk's keep mask
  int64_t partial_off;  // first double of this chunk's 2 pad4(n) VJP overflow / partial sums
  int32_t instance;
  int32_t index;        // position among the instance's chunks
};

// A chunk's Philox blocks are split into kSegsPerWave segments per wave of
// its work-group; waves take segments from an LDS counter until none is left,
// so a wave that runs fast takes more of them.
#ifndef QLOITS_SEGS_PER_WAVE
#define QLOITS_SEGS_PER_WAVE 4
#endif
constexpr int kSegsPerWave = QLOITS_SEGS_PER_WAVE;

// Blocks of segment w of nwaves in a chunk: [qa, qb), and the slot counters it covers.
struct Segment {
  uint64_t qa, qb;      // Philox blocks
  uint64_t p_lo, p_hi;  // counters (slots + counter_base), p_lo <= p_hi
  int64_t iters;        // 64-block iterations
};

__host__ __device__ __forceinline__ uint64_t blocks_per_wave(uint64_t nblk, int nwaves) {
  const uint64_t per = (nblk + uint64_t(64 * nwaves) - 1) / uint64_t(64 * nwaves);
  return per * 64;
}

__host__ __device__ __forceinline__ Segment segment_of(uint64_t counter_base, int64_t slot_lo,
                                                       int64_t slot_hi, int w, int nwaves) {
  Segment s{};
  const uint64_t p_lo = counter_base + uint64_t(slot_lo), p_hi = counter_base + uint64_t(slot_hi);
  const uint64_t q_begin = p_lo >> 2;
  const uint64_t q_end = slot_hi > slot_lo ? (p_hi + 3) >> 2 : q_begin;
  const uint64_t bw = blocks_per_wave(q_end - q_begin, nwaves);
  uint64_t qa = q_begin + bw * uint64_t(w);
  uint64_t qb = qa + bw;
  if (qa > q_end) qa = q_end;
  if (qb > q_end) qb = q_end;
  s.qa = qa;
  s.qb = qb;
  uint64_t lo = qa * 4, hi = qb * 4;
  lo = lo < p_lo ? p_lo : lo;
  hi = hi > p_hi ? p_hi : hi;
  s.p_lo = lo;
  s.p_hi = hi > lo ? hi : lo;
  s.iters = int64_t((qb - qa + 63) / 64);
  return s;
}

// Table blob of one instance (byte offsets from table_off), all 16-byte aligned:
//   TableHeader
//   uint32 srt[pad4(pow2(n))]   thresholds T_t sorted ascending (equal to T for a
//                               monotone CDF); read only by the rare bisection path
//   uint32 guide[G]             level 1: one word per cell of 2^s k values,
//                               s = 24 - guide_log2 (encoding below)
//   uint32 guide2[G]            level 2: sub-cell words of the redirected crowded cells
//   Coef   coef[n_pad4]         per-bin interpolation data (bin j uses nodes j, j+1)
//   double cdf[n_pad4]          C_t (fp64)
struct TableHeader {
  double norm;
  double inv_n;
  uint32_t flags;   // kMonotone | kSafe
  uint32_t n;
  uint32_t guide_log2;
  uint32_t level2_words;  // guide words [G, G + level2_words) are in use
  double reserved[4];
};
static_assert(sizeof(TableHeader) == 64, "header size");

constexpr uint32_t kMonotone = 1u;
// Every value the curve can produce is finite, non-zero and representable in
// float (monotone thresholds, acceptance 1, C_{n-1} within 2^-23 of 1, finite
// per-bin data, 1e-30 <= |x_t| <= 1e29 with one sign): the fp32 keep test
// without a window cut is then always true.
constexpr uint32_t kSafe = 2u;

// Guide words.  A cell covers keys [L, L + S).  With c0 = #{t : srt_t < L},
// c1 = #{t : srt_t < L + S} and j(c) = clamp(c - 1, 0, n - 2):
//   direct (bit 31 clear): bits 0-14 j0 = j(c0), bits 15-27 `off`; the bin of
//     key k in the cell is j0 + (k - L >= off).  Used when every threshold in
//     the cell sits at one key P with j(c1) - j0 <= 1 (off = P - L), or when
//     j(c1) == j0 (off = kNoPoint, larger than any in-cell offset).
//   search (bits 31, 30 set): bits 0-14 c0, bits 15-29 c1 - c0; bisect srt.
//   indirect (bit 31 set, bit 30 clear, level 1 only): bits 0-23 word index of
//     2^f level-2 words (direct or search) covering S / 2^f keys each, f in
//     bits 24-27.  A crowded cell with c thresholds gets 2^f = min(S, pow2(2c))
//     sub-cells while the G level-2 words last.
// LDS a single work-group may use on gfx950 (160 KiB per CU), less the
// kernels' static LDS.
constexpr int64_t kMaxLds = 160 * 1024 - 2048;

// Cells are at most 2^12 keys wide (guide_log2 >= kMinGuideLog2).
constexpr int kMinGuideLog2 = 12;
constexpr int kMaxGuideLog2 = 14;
constexpr uint32_t kSpecial = 0x80000000u;
constexpr uint32_t kSearch = 0xC0000000u;
constexpr uint32_t kNoPoint = 0x1000u;

// fp32 path: value = A_j + M_j * (k - T_j) with M_j = m_j 2^-24 and
// A_j = x_j + m_j (T_j 2^-24 - C_j), both rounded from fp64, which is
// x_j + m_j (u - C_j) for u = k 2^-24; r_j = T_j 2^-24 - C_j rounded to float
// gives u - C_j = (k - T_j) 2^-24 + r_j for the VJP.  Tables built in LDS
// store a structure of arrays over nb = pad4(n) bins (a random 8-byte LDS read
// costs what a 4-byte one does, a 12-byte one three times that):
//   float2 ma[nb] = {M_j, A_j} | uint32 thr[nb] = T_j | float r[nb] = r_j
// Tables built in global memory (read from L2 while sampling) store one
// 16-byte record {M_j, A_j, T_j (bits), r_j} per bin instead, so that a key
// costs one L2 access rather than two.
struct alignas(16) CoefF32 {  // one bin's share (sizing only)
  uint32_t w[4];
};
struct CoefF32View {
  float2* ma;
  uint32_t* thr;
  float* r;
};
__host__ __device__ __forceinline__ CoefF32View coef_f32_view(void* base, int64_t nb) {
  CoefF32View v;
  v.ma = static_cast<float2*>(base);
  v.thr = reinterpret_cast<uint32_t*>(v.ma + nb);
  v.r = reinterpret_cast<float*>(v.thr + nb);
  return v;
}
// fp64 path: the reference's own operands.
struct alignas(16) CoefF64 {
  double c;
  double m;
  double x;
  uint32_t thr;
  uint32_t pad;
};
static_assert(sizeof(CoefF32) == 16 && sizeof(CoefF64) == 32, "coef size");

template <typename Real>
struct CoefOf;
template <>
struct CoefOf<float> {
  using type = CoefF32;
};
template <>
struct CoefOf<double> {
  using type = CoefF64;
};

__host__ __device__ __forceinline__ int64_t pad4(int64_t n) { return (n + 3) & ~int64_t(3); }

__host__ __device__ __forceinline__ int64_t pow2_ceil(int64_t n) {
  int64_t p = 1;
  while (p < n) p <<= 1;
  return p;
}


struct TableLayout {
  int64_t thr, guide, coef, cdf, bytes;  // byte offsets within the instance blob
};

__host__ __device__ __forceinline__ TableLayout table_layout(int64_t n, int guide_log2,
                                                             int64_t coef_size) {
  TableLayout t;
  t.thr = sizeof(TableHeader);
  t.guide = t.thr + pad4(pow2_ceil(n)) * 4;  // room for an in-place bitonic sort
  t.coef = t.guide + 2 * (int64_t(1) << guide_log2) * 4;
  t.cdf = t.coef + pad4(n) * coef_size;
  t.bytes = t.cdf + pad4(n) * 8;
  return t;
}

// Bytes of LDS the sampling kernel stages for one instance (srt, guide, coef).
__host__ __device__ __forceinline__ int64_t lds_table_bytes(int64_t n, int guide_log2,
                                                            int64_t coef_size) {
  const TableLayout t = table_layout(n, guide_log2, coef_size);
  return t.cdf - t.thr;
}

// log2 of the level-2 fan-out of a crowded cell with cnt >= 2 thresholds.
__host__ __device__ __forceinline__ int guide_fanout_log2(int cnt, int shift) {
  const int l = 1 + (32 - __builtin_clz(uint32_t(cnt - 1)));  // 1 + ceil(log2 cnt)
  return l < shift ? l : shift;
}

// LDS image of one instance's tables while a work-group builds and samples it
// (byte offsets): srt[pad4(pow2(n))] | guide[2G] | coef[pad4(n)] | cdf[pad4(n)]
// | crowd[G / 4] (uint2 per redirected cell, used only while building).
struct LdsLayout {
  int64_t srt, guide, coef, cdf, crowd, bytes;
};

__host__ __device__ __forceinline__ LdsLayout lds_build_layout(int64_t n, int guide_log2,
                                                               int64_t coef_size) {
  LdsLayout l;
  const int64_t G = int64_t(1) << guide_log2;
  l.srt = 0;
  l.guide = pad4(pow2_ceil(n)) * 4;
  l.coef = l.guide + 2 * G * 4;
  l.cdf = l.coef + pad4(n) * coef_size;
  l.crowd = l.cdf + pad4(n) * 8;
  l.bytes = l.crowd + (G / 4) * 8;
  return l;
}

// Instances whose LDS image does not fit keep srt, coef and cdf in the global
// blob ("global tables"); a work-group then stages only the guide in LDS.
// Standalone build: crowded-cell list (G / 4 uint2) | cell counts (G uint32).
// A global-table build first stages the curve's nodes and densities in LDS
// (read from there by the norm, CDF and threshold steps) when both fit.
__host__ __device__ __forceinline__ int64_t global_build_stage_bytes(int64_t n, int64_t real_size) {
  const int64_t b = 2 * pad4(n) * real_size;
  return b <= kMaxLds ? b : 0;
}

__host__ __device__ __forceinline__ int64_t lds_global_build_bytes(int64_t n, int guide_log2,
                                                                   int64_t real_size) {
  const int64_t G = int64_t(1) << guide_log2;
  const int64_t lists = (G / 4) * 8 + G * 4;
  const int64_t stage = global_build_stage_bytes(n, real_size);
  return lists > stage ? lists : stage;
}

// Bin-range VJP (curves whose VJP tables and histograms do not fit LDS): the
// bins of a curve are split into R ranges; one work item (chunk, range) sums
// the keys of its range over a cell table of at most kRangeCells cells.  LDS:
// cell_bin[kRangeCells + 4] | thr[W + 4] | [float64: C_j per bin, 8 B] | two
// 64-bit histograms[W].  float32 sums exact integers (no per-bin data), so its
// ranges hold up to kRangeBinsF32 bins, float64's kRangeBins.
constexpr int kRangeBins = 5120;
constexpr int kRangeBinsF32 = 7168;
constexpr int kRangeCellsLog2 = 12;
constexpr int kRangeCells = 1 << kRangeCellsLog2;

__host__ __device__ __forceinline__ int vjp_range_count(int64_t n, bool f32) {
  const int w = f32 ? kRangeBinsF32 : kRangeBins;
  return int((n - 1 + w - 1) / w);
}

// Bins per range (a multiple of 4); every range is non-empty.
__host__ __device__ __forceinline__ int vjp_range_width(int64_t n, int ranges) {
  return int(pad4((n - 1 + ranges - 1) / ranges));
}

__host__ __device__ __forceinline__ int64_t lds_range_bytes(int width, bool f32) {
  return int64_t(kRangeCells + 4) * 4 + int64_t(width + 4) * 4 + int64_t(width) * (f32 ? 16 : 24);
}

// Bin of key k from a direct guide word; ko = k - L (the offset in the cell).
__host__ __device__ __forceinline__ int guide_direct_bin(uint32_t w, uint32_t ko) {
  return int(w & 0x7FFFu) + (ko >= ((w >> 15) & 0x1FFFu) ? 1 : 0);
}

// Guide word of a cell [key_lo, key_lo + S) with c0 / c1 thresholds below its
// ends; srt is the sorted threshold array.
__host__ __device__ __forceinline__ uint32_t guide_encode(int c0, int c1, const uint32_t* srt,
                                                          uint32_t key_lo, int n) {
  const int j0 = c0 - 1 < 0 ? 0 : (c0 - 1 > n - 2 ? n - 2 : c0 - 1);
  const int j1 = c1 - 1 < 0 ? 0 : (c1 - 1 > n - 2 ? n - 2 : c1 - 1);
  if (j1 == j0) return uint32_t(j0) | (kNoPoint << 15);
  if (j1 - j0 == 1 && srt[c0] == srt[c1 - 1]) return uint32_t(j0) | ((srt[c0] - key_lo) << 15);
  return kSearch | uint32_t(c0) | (uint32_t(c1 - c0) << 15);
}

}  // namespace qloits
