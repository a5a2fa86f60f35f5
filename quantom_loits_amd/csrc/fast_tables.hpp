// fast_tables.hpp -- cell tables of the float32 batched path (curves of at
// most kFastMaxNodes nodes).
//
// The bin of key k (u = k 2^-24) is the reference's count-based interval
// (cpp/loits_core.cpp:12-18), j = clamp(#{t : T_t <= k} - 1, 0, n - 2) with
// T_t = ceil(C_t 2^24).  Since T_0 = 0 and the clamp drops the last
// threshold, j = #{i in 1..n-2 : S_i <= k} for the sorted thresholds S
// (S = T when the CDF is monotone): the "splits" S_1..S_{n-2}.
//
// The 2^24 keys are cut into G = 2^g cells of 2^s keys (s = 24 - g).  Cell c
// covers [L, L + 2^s), L = c 2^s.  Its word holds jL = #{splits <= L} and,
// when at most one split lies strictly inside (L, L + 2^s), that split's
// offset p.  Cells with two or more inside splits point to 2^f level-2 words
// covering 2^(s-f) keys each, in the same format (sub-cells of a single key
// never hold an inside split), or, past the level-2 budget, to a bisection of
// the sorted splits ("search" word).
//
// Word encodings (32 bits), for a cell of 2^s keys (2^(s-f) at level 2):
//   direct:   bit 12 clear.  Bits 0-11 F; bits (32 - s)..31 the offset p of the
//             single inside split (1 <= p < 2^s), or 0 when there is none; the
//             bits between are zero.  With the key's in-cell offset moved to the
//             top and ones below,  X = (k << (32 - s)) | (2^(32-s) - 1),
//                 j + 1 = F + (X >= w)
//             so F = jL + 1 when a split lies inside and F = jL otherwise.
//   indirect: bits 12 set, 13 clear; bits 14-17 f; bits 18-31 index of the
//             first level-2 word, counted from the end of level 1
//   search:   bits 12, 13 set; bits 0-11 jL; bits 14-31 the number of inside
//             splits (S_{jL+1} .. S_{jL+cnt})
// The level-2 word of an indirect cell is selected by the top f bits of the
// offset, X >> (32 - f), and read with X2 = (X << f) | (2^f - 1).
//
// Per-bin interpolation data, float4 {M_j, A_j, T_j (bits), 0}: the value of
// key k in bin j is fma(M_j, float(k - T_j), A_j) with M_j = m_j 2^-24 and
// A_j = x_j + m_j (T_j 2^-24 - C_j) rounded from float64 (the reference's
// x_j + m_j (u - C_j), cpp/loits_core.cpp:190-195, with the exact integer
// k - T_j).
#pragma once

#include <cstdint>

#include "loits_device.hpp"

namespace qloits {

constexpr int kFastMaxNodes = 4096;
constexpr int64_t kFastMaxChunk = int64_t(1) << 18;  // keeps the VJP's integer sums < 2^62
constexpr uint32_t kCellMulti = 1u << 12;
constexpr uint32_t kCellSearch = 1u << 13;
constexpr uint32_t kCellField = 0xFFFu;
constexpr uint32_t kNoCell = 0xFFFFFFFFu;  // empty slot of the crowded-cell list
constexpr uint32_t kFastTables = 4u;       // TableHeader flag: tables in the cell format
// VJP fixed point: |round(g 2^e)| < 2^bits with keys * 2^bits * 2^24 <= 2^62, so
// that sum q and sum q k over a chunk of `keys` keys below 2^24 cannot wrap: 20
// bits at 2^18 keys, up to 28 for chunks of at most 2^10 keys (the largest
// float32 cotangents of the chunk are then exact and the rest keep 26 bits
// below the largest).
__host__ __device__ __forceinline__ int fixed_point_bits(int64_t keys) {
  int l = 0;
  while ((int64_t(1) << l) < keys) ++l;
  return l >= 10 ? 38 - l : 28;
}

__host__ __device__ __forceinline__ int fast_l2_cap(int64_t n) { return n <= 2048 ? 4096 : 2048; }

// Blob of one instance after the TableHeader (byte offsets from table_off):
//   srt[pad4(pow2(n))] u32   sorted thresholds (room for a bitonic sort)
//   cells[G + l2cap]  u32    level-1 then level-2 words
//   own[pad4(n)]      u32    each bin's own threshold T_j (= srt when monotone)
//   cdf[pad4(n)]      f64
struct FastBlobLayout {
  int64_t srt, cells, own, cdf, bytes;
};

__host__ __device__ __forceinline__ FastBlobLayout fast_blob_layout(int64_t n, int g) {
  FastBlobLayout b;
  b.srt = sizeof(TableHeader);
  b.cells = b.srt + pad4(pow2_ceil(n)) * 4;
  b.own = b.cells + ((int64_t(1) << g) + fast_l2_cap(n)) * 4;
  b.cdf = b.own + pad4(n) * 4;
  b.bytes = b.cdf + pad4(n) * 8;
  return b;
}

// LDS of the forward (build + sampling), byte offsets:
//   cells[G + l2cap] u32 (the float64 CDF aliases its front during the build)
//   list[l2cap / 4]  uint2 (crowded cells by level-2 slot, used during the build)
//   bins[pad4(n)]    float4
//   srt[pad4(pow2(n))] u32
struct FastLdsForward {
  int64_t cells, list, bins, srt, bytes;
};

__host__ __device__ __forceinline__ FastLdsForward fast_lds_forward(int64_t n, int g) {
  FastLdsForward l;
  l.cells = 0;
  l.list = l.cells + ((int64_t(1) << g) + fast_l2_cap(n)) * 4;
  l.bins = l.list + (fast_l2_cap(n) / 4) * 8;
  l.srt = l.bins + pad4(n) * 16;
  l.bytes = l.srt + pad4(pow2_ceil(n)) * 4;
  return l;
}

// LDS of the VJP: srt[pad4(pow2(n))] | cells[G + l2cap] (the blob's order, one
// copy) | a0[kFastMaxNodes] | a1[kFastMaxNodes] (64-bit integer histograms at a
// fixed distance, one address for both adds).
struct FastLdsVjp {
  int64_t srt, cells, a0, a1, bytes;
};

__host__ __device__ __forceinline__ FastLdsVjp fast_lds_vjp(int64_t n, int g) {
  FastLdsVjp l;
  l.srt = 0;
  l.cells = pad4(pow2_ceil(n)) * 4;
  l.a0 = l.cells + ((int64_t(1) << g) + fast_l2_cap(n)) * 4;
  l.a0 = (l.a0 + 15) & ~int64_t(15);
  l.a1 = l.a0 + int64_t(kFastMaxNodes) * 8;
  l.bytes = l.a1 + int64_t(kFastMaxNodes) * 8;
  return l;
}

// Cells for n nodes: 2^(ceil log2 n + 3), within [2^12, 2^14].
__host__ __device__ __forceinline__ int fast_guide_log2(int64_t n) {
  int l = 0;
  while ((int64_t(1) << l) < n) ++l;
  l += 3;
  return l < 12 ? 12 : (l > 14 ? 14 : l);
}

// X of key k for cells of 2^s keys (1 <= s <= 31).
__host__ __device__ __forceinline__ uint32_t cell_x(uint32_t k, int s) {
  return (k << (32 - s)) | ((1u << (32 - s)) - 1u);
}

// Direct word of a cell of 2^s keys with jl splits at or below its first key
// and `inside` (0 or 1) splits strictly inside; `split` is that split's key
// (only its low s bits are kept, which is its offset in the cell).
__host__ __device__ __forceinline__ uint32_t cell_direct(uint32_t jl, uint32_t inside,
                                                         uint32_t split, int s) {
  return inside == 0u ? jl : ((split << ((32 - s) & 31)) | (jl + 1u));
}

// j + 1 for a direct word w and the key's X.
__host__ __device__ __forceinline__ int cell_jj(uint32_t w, uint32_t x) {
  return int(w & (kCellMulti | kCellField)) + (x >= w ? 1 : 0);
}

// Number of entries of the ascending a[lo, lo + len) that are <= key.
__host__ __device__ __forceinline__ int count_le(const uint32_t* a, int lo, int len,
                                                 uint32_t key) {
  int c = 0;
  while (len > 0) {
    const int half = len >> 1;
    if (a[lo + half] <= key) {
      lo += half + 1;
      c += half + 1;
      len -= half + 1;
    } else {
      len = half;
    }
  }
  return c;
}

// j + 1 for key k whose level-1 word w has bit 12 set (indirect or search);
// l2 = the level-2 words, srt = the sorted splits.
__host__ __device__ __forceinline__ int cell_multi_jj(uint32_t w, uint32_t x, uint32_t k,
                                                      const uint32_t* l2, const uint32_t* srt) {
  if ((w & kCellSearch) == 0u) {  // indirect: one level-2 word
    const uint32_t f = (w >> 14) & 0xFu;
    const uint32_t w2 = l2[(w >> 18) + (x >> (32u - f))];
    const uint32_t t2 = w2 & (kCellMulti | kCellField);
    if (t2 <= kCellField) return int(t2) + (((x << f) | ((1u << f) - 1u)) >= w2 ? 1 : 0);
    w = w2;
  }
  const int jl = int(w & kCellField);
  return jl + 1 + count_le(srt, jl + 1, int(w >> 14), k);
}

}  // namespace qloits
