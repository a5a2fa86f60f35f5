// batched.hpp -- parameters and launchers of the batched 1-D kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "loits_device.hpp"

namespace qloits {

struct BatchedParams {
  const InstanceDesc* inst;
  const ChunkDesc* chunks;
  const int2* vjp_items;  // VJP work items {chunk, bin range}: cell-table items, other plain
                          // items, then ranged
  const int32_t* fwd_order;  // chunks of the cell-table instances first, then the others
  const int32_t* fwd_queue;  // persistent forward: chunks, largest first
  const int32_t* vjp_queue;  // persistent VJP items: indices into vjp_items
  const int32_t* global_list;  // global-table instances, most nodes first (their build launch)
  int64_t n_fwd_items, n_global;
  int64_t n_instances, n_chunks, n_vjp_items, n_vjp_plain;
  int64_t n_cell_chunks, n_vjp_cell;
  const void* x;
  const void* density;
  uint8_t* tables;
  uint64_t seed;
  int32_t use_window;
  int32_t pad;
  double win_lo, win_hi;
  void* events;          // capacity layout
  int32_t* bins;         // optional, capacity layout
  int64_t* kept;         // [n_instances]
  int64_t* total;        // [1]
  double* norm_out;      // optional [n_instances]
  double* cdf_out;       // optional [total_nodes]
  int64_t n_segments;    // n_chunks * segs
  int32_t segs;          // segments per chunk (waves per work-group * kSegsPerWave)
  int32_t pad2;
  // scratch
  int32_t* flags;        // [0] = dense
  int64_t* seg_kept;     // [n_segments]
  int64_t* chunk_kept;   // [n_chunks] kept + 1 of a chunk that dropped slots, else 0
  int64_t* chunk_row;    // [n_chunks] first packed row of each chunk
  int64_t* seg_row;      // [n_segments] first packed row within its chunk
  int32_t* seg_drop;     // [n_segments] first iteration with a dropped slot
  uint32_t* done;        // [n_instances]
  uint64_t* keepmask;    // per-segment keep bits
  double* partials;      // per-chunk partial VJP sums
};

// Scratch flag words (P.flags, 16 int32): [0] dense, [2..3] the forward's 64-bit
// chunk counter (finished | dropped-slot chunks << 32), [4]/[5] next item /
// exited work-groups of the persistent forward,
// [6]/[7] the same for the persistent VJP.  Every counter returns to zero.
constexpr int kFlagFwdNext = 4, kFlagFwdExit = 5, kFlagVjpNext = 6, kFlagVjpExit = 7;

template <typename Real>
hipError_t launch_tables(const BatchedParams& P, int threads, bool sequential, int64_t lds,
                         bool only_global, hipStream_t stream);
template <typename Real>
hipError_t launch_sample(const BatchedParams& P, int threads, bool sequential, int64_t lds,
                         hipStream_t stream);
hipError_t launch_finalize(const BatchedParams& P, hipStream_t stream);
template <typename Real>
hipError_t launch_pack(const BatchedParams& P, int threads, void* packed, int32_t* packed_bins,
                       hipStream_t stream);
template <typename Real>
hipError_t launch_vjp(const BatchedParams& P, int threads, int64_t lds, const void* grad,
                      void* grad_density, hipStream_t stream);
// One persistent launch each: as many work-groups as fit the device at once
// (at most one per item), taking items from the queues in P.
template <typename Real>
hipError_t launch_forward_persistent(const BatchedParams& P, int threads, bool sequential,
                                     int64_t lds, hipStream_t stream);
template <typename Real>
hipError_t launch_vjp_persistent(const BatchedParams& P, int threads, int64_t lds,
                                 const void* grad, void* grad_density, hipStream_t stream);

}  // namespace qloits
