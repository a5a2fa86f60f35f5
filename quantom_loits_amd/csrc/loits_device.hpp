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

// Inclusive prefix sum across the 64 lanes of a wave (Hillis-Steele over
// lane shuffles; the compiler lowers the shuffles to DPP/ds_bpermute).
template <typename T>
__device__ __forceinline__ T wave_inclusive_scan(T v) {
  const int l = lane_id();
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const T o = __shfl_up(v, d, kWave);
    if (l >= d) v += o;
  }
  return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int d = kWave / 2; d >= 1; d >>= 1) v += __shfl_xor(v, d, kWave);
  return v;
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
};

// A chunk is the work of one work-group; each of its waves owns a contiguous
// run of Philox blocks (a "segment") and compacts it independently.
struct ChunkDesc {
  int64_t slot_lo;      // first slot (relative to the instance)
  int64_t slot_hi;      // one past the last slot
  int64_t mask_off;     // first 64-bit word of this chunk's keep mask
  int64_t partial_off;  // first double of this chunk's partial VJP sums (multi-chunk instances)
  int32_t instance;
  int32_t index;        // position among the instance's chunks
};

// Blocks of wave w of a chunk: [qa, qb), and the slot counters it covers.
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
//   uint32 srt[n_pad4]          thresholds T_t sorted ascending (the search array;
//                               equal to T for a monotone CDF)
//   uint32 guide[G]             level 1: cell g: lo | hi << 16 with lo = #{t : srt_t < g << s},
//                               hi = #{t : srt_t < (g + 1) << s}, s = 24 - guide_log2,
//                               or kIndirect | offset into level 2
//   uint32 guide2[G]            level 2 sub-cells (kSub per redirected cell)
//   Coef   coef[n_pad4]         per-bin interpolation data (bin j uses nodes j, j+1)
//   double cdf[n_pad4]          C_t (fp64)
//   float  rlo[n_pad4]          r_j = T_j * 2^-24 - C_j rounded to float (fp32 path)
struct TableHeader {
  double norm;
  double inv_n;
  uint32_t flags;   // kMonotone
  uint32_t n;
  uint32_t guide_log2;
  uint32_t pad;
  double reserved[4];
};
static_assert(sizeof(TableHeader) == 64, "header size");

constexpr uint32_t kMonotone = 1u;
// Every value the curve can produce is finite, non-zero and representable in
// float (monotone thresholds, acceptance 1, C_{n-1} within 2^-23 of 1, finite
// per-bin data, 1e-30 <= |x_t| <= 1e29 with one sign): the fp32 keep test
// without a window cut is then always true.
constexpr uint32_t kSafe = 2u;

// Two-level guide.  Level 1 has G = 2^guide_log2 cells over the 24-bit k
// range; a direct cell word is lo | hi << 16 (bit 31 clear).  A crowded cell
// (hi - lo >= 2) may redirect (bit 31 set, low bits = word offset into level
// 2) to kSub sub-cells of the same form covering its k range.
constexpr int kSubLog2 = 6;
constexpr int kSub = 1 << kSubLog2;
constexpr uint32_t kIndirect = 0x80000000u;

// fp32 path: u - C_j = (k - T_j) * 2^-24 + r_j with r_j = T_j * 2^-24 - C_j
// rounded to float; value = x_j + m_j * (u - C_j).
struct CoefF32 {
  uint32_t thr;
  float r;
  float m;
  float x;
};
// fp64 path: the reference's own operands.
struct CoefF64 {
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

struct TableLayout {
  int64_t thr, guide, coef, cdf, rlo, bytes;  // byte offsets within the instance blob
};

__host__ __device__ __forceinline__ TableLayout table_layout(int64_t n, int guide_log2,
                                                             int64_t coef_size) {
  TableLayout t;
  t.thr = sizeof(TableHeader);
  t.guide = t.thr + pad4(n) * 4;
  t.coef = t.guide + 2 * (int64_t(1) << guide_log2) * 4;
  t.cdf = t.coef + pad4(n) * coef_size;
  t.rlo = t.cdf + pad4(n) * 8;
  t.bytes = t.rlo + pad4(n) * 4;
  return t;
}

// Bytes of LDS the sampling kernel stages for one instance (thr, guide, coef).
__host__ __device__ __forceinline__ int64_t lds_table_bytes(int64_t n, int guide_log2,
                                                            int64_t coef_size) {
  const TableLayout t = table_layout(n, guide_log2, coef_size);
  return t.cdf - t.thr;
}

}  // namespace qloits
