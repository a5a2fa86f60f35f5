// capi.hip -- the extern "C" boundary (include/loits_hip.h): argument
// checks, layouts, thread-local error reporting and kernel launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "batched.hpp"
#include "capi_internal.hpp"
#include "fast_tables.hpp"
#include "loits_device.hpp"
#include "loits_hip.h"

namespace qloits {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int status, const std::string& msg) {
  g_last_error = msg;
  return status;
}

int check_hip(hipError_t e, const char* what) {
  if (e == hipSuccess) return LOITS_HIP_OK;
  return fail(LOITS_HIP_ERUNTIME, std::string(what) + ": " + hipGetErrorString(e));
}

// ---------------------------------------------------------------------------
// Philox probe kernels (rng/philox.hpp:17-61, cpp/loits_core.cpp:39-53).
__global__ void philox_fill_kernel(float* values, int64_t count, uint64_t seed, uint64_t sid) {
  const int64_t block = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t base = block * 4;
  if (base >= count) return;
  const Words4 w = philox4x32_10(uint64_t(block), sid, seed);
  if (base + 3 < count) {
    const float4 v = make_float4(uniform_float(w.w0), uniform_float(w.w1), uniform_float(w.w2),
                                 uniform_float(w.w3));
    if ((reinterpret_cast<uintptr_t>(values) & 15) == 0) {
      reinterpret_cast<float4*>(values)[block] = v;
      return;
    }
  }
#pragma unroll
  for (int e = 0; e < 4; ++e)
    if (base + e < count) values[base + e] = uniform_float(word_of(w, e));
}

__global__ void philox_words_kernel(uint32_t* words, int64_t first, int64_t blocks,
                                    uint64_t sid, uint64_t seed) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= blocks) return;
  const Words4 w = philox4x32_10(uint64_t(first + i), sid, seed);
  reinterpret_cast<uint4*>(words)[i] = make_uint4(w.w0, w.w1, w.w2, w.w3);
}

// ---------------------------------------------------------------------------
// Batched layout.
namespace {

// Work of one node's share of a table build or staging, in keys (a chunk's
// sampling loop costs one unit per key): about 15 us per 2048-node build
// against about 0.75 ns per key per CU.
constexpr int64_t kBuildKeysPerNode = 10;

int ceil_log2(int64_t v) {
  int l = 0;
  while ((int64_t(1) << l) < v) ++l;
  return l;
}

int64_t align16(int64_t v) { return (v + 15) & ~int64_t(15); }

struct VjpItem {
  int32_t chunk, range;
};

struct Plan {
  std::vector<InstanceDesc> inst;
  std::vector<ChunkDesc> chunks;
  std::vector<VjpItem> items;  // VJP work items: one per chunk (cell-table instances first),
                               // then one per (chunk, range)
  std::vector<int32_t> fwd_order;  // chunks of the cell-table instances first
  std::vector<int32_t> fwd_queue;  // persistent forward: chunks, largest first
  std::vector<int32_t> vjp_queue;  // persistent VJP: indices into items
  std::vector<int32_t> global_list;  // global-table instances, most nodes first
  int64_t n_plain = 0;
  loits_hip_layout layout;
};

int make_plan(const loits_hip_instance* instances, int64_t n, int dtype,
              const loits_hip_tuning* tuning, Plan* plan) {
  if (!instances || n <= 0) return fail(LOITS_HIP_EINVAL, "batched: need at least one instance");
  if (dtype != LOITS_HIP_F32 && dtype != LOITS_HIP_F64)
    return fail(LOITS_HIP_EINVAL, "batched: dtype must be LOITS_HIP_F32 or LOITS_HIP_F64");
  if (!tuning) return fail(LOITS_HIP_EINVAL, "batched: tuning is NULL");
  const int threads = tuning->threads;
  if (threads != 256 && threads != 512 && threads != 1024)
    return fail(LOITS_HIP_EINVAL, "batched: threads must be 256, 512 or 1024");
  if (tuning->guide_shift < 0 || tuning->guide_shift > 3)
    return fail(LOITS_HIP_EINVAL, "batched: guide_shift must be in [0, 3]");
  if (tuning->persistent != 0 && tuning->persistent != 1)
    return fail(LOITS_HIP_EINVAL, "batched: persistent must be 0 or 1");
  if (tuning->reserved != 0) return fail(LOITS_HIP_EINVAL, "batched: tuning.reserved must be 0");
  for (int b = 0; b < 16; ++b) {
    const int cl = tuning->bucket_chunk_log2[b], gs = tuning->bucket_guide_shift[b];
    if (cl != 0 && (cl < 10 || cl > 30))
      return fail(LOITS_HIP_EINVAL, "batched: bucket_chunk_log2 entries must be 0 or in [10, 30]");
    if (gs < 0 || gs > 4)
      return fail(LOITS_HIP_EINVAL, "batched: bucket_guide_shift entries must be in [0, 4]");
  }
  if (tuning->table_format != 0 && tuning->table_format != 1)
    return fail(LOITS_HIP_EINVAL, "batched: table_format must be 0 or 1");
  const int64_t per_iter = int64_t(4) * threads;
  int64_t chunk_events = tuning->chunk_events;
  if (chunk_events <= 0 || chunk_events > (int64_t(1) << 30)) chunk_events = int64_t(1) << 30;
  chunk_events = (chunk_events + per_iter - 1) / per_iter * per_iter;
  const int64_t coef_size = dtype == LOITS_HIP_F32 ? 16 : 32;
  const int nwaves = threads / 64;
  const int nsegs = nwaves * kSegsPerWave;  // segments per chunk

  loits_hip_layout& L = plan->layout;
  std::memset(&L, 0, sizeof(L));
  L.n_instances = n;
  L.dtype = dtype;
  L.threads = threads;
  L.scan_mode = tuning->scan_mode;
  L.persistent = tuning->persistent;
  int64_t out = 0, table_off = 0, mask_words = 0, partial_doubles = 0;
  int64_t lds_f = 0, lds_b = 0, lds_t = 0;
  int32_t n_global = 0, n_ranged = 0, n_cell = 0;
  std::vector<VjpItem> ranged;  // appended after the plain items
  std::vector<VjpItem> guide_items;  // plain items of guide-table instances (after the cell items)
  std::vector<int32_t> guide_chunks;
  plan->inst.resize(size_t(n));
  plan->chunks.clear();
  plan->items.clear();
  plan->fwd_order.clear();
  plan->fwd_queue.clear();
  plan->vjp_queue.clear();
  plan->global_list.clear();
  for (int64_t i = 0; i < n; ++i) {
    const loits_hip_instance& s = instances[i];
    if (s.n_nodes < 2 || s.n_nodes > 32767)
      return fail(LOITS_HIP_EINVAL, "batched: instance " + std::to_string(i) +
                                        " needs 2 <= n_nodes <= 32767");
    if (s.events < 0 || s.node_off < 0)
      return fail(LOITS_HIP_EINVAL, "batched: instance " + std::to_string(i) +
                                        " has a negative offset or event count");
    if (s.out_off != out)
      return fail(LOITS_HIP_EINVAL, "batched: instance " + std::to_string(i) +
                                        " out_off must equal the events of the instances before it");
    if (s.counter_base > ~uint64_t(0) - uint64_t(s.events))
      return fail(LOITS_HIP_EINVAL, "batched: Philox counter range overflows 64 bits");
    // this curve's grid-size bucket may override the chunk size and guide shift
    const int bucket = std::min(15, ceil_log2(s.n_nodes));
    int64_t chunk_i = chunk_events;
    if (tuning->bucket_chunk_log2[bucket] != 0)
      chunk_i = (((int64_t(1) << tuning->bucket_chunk_log2[bucket]) + per_iter - 1) / per_iter) *
                per_iter;
    const int shift_i = tuning->bucket_guide_shift[bucket] != 0
                            ? tuning->bucket_guide_shift[bucket] - 1
                            : tuning->guide_shift;
    InstanceDesc& d = plan->inst[size_t(i)];
    d.node_off = s.node_off;
    d.n_nodes = s.n_nodes;
    d.events = s.events;
    d.out_off = s.out_off;
    d.counter_base = s.counter_base;
    d.stream = s.stream;
    d.accept = s.accept ? 1 : 0;
    // float32 curves of at most kFastMaxNodes nodes take the cell tables
    // (fast_tables.hpp) when both kernels' LDS images fit
    if (tuning->table_format == 0 && dtype == LOITS_HIP_F32 && s.n_nodes <= kFastMaxNodes &&
        chunk_i <= kFastMaxChunk) {
      int g = fast_guide_log2(s.n_nodes);
      while (g > kMinGuideLog2 && std::max(fast_lds_forward(s.n_nodes, g).bytes,
                                           fast_lds_vjp(s.n_nodes, g).bytes) > kMaxLds)
        --g;
      if (std::max(fast_lds_forward(s.n_nodes, g).bytes, fast_lds_vjp(s.n_nodes, g).bytes) <=
          kMaxLds)
        d.fast_log2 = g;
    }
    if (d.fast_log2 > 0) {
      ++n_cell;
      d.guide_log2 = d.fast_log2;
      d.table_off = table_off;
      table_off += align16(fast_blob_layout(s.n_nodes, d.fast_log2).bytes);
      lds_f = std::max(lds_f, fast_lds_forward(s.n_nodes, d.fast_log2).bytes);
      lds_t = std::max(lds_t, fast_lds_forward(s.n_nodes, d.fast_log2).bytes);
      lds_b = std::max(lds_b, fast_lds_vjp(s.n_nodes, d.fast_log2).bytes);
    } else {
      // Forward: the finest guide whose LDS build fits; when none does, the
      // instance keeps its per-bin tables in global memory (built by a pre-pass)
      // and only the guide is staged in LDS.  VJP: tables and histograms in LDS
      // at the finest guide that fits both kernels; otherwise bin ranges.
      const int g_want = std::min(kMaxGuideLog2,
                                  std::max(kMinGuideLog2, ceil_log2(s.n_nodes) + shift_i));
      auto build_bytes = [&](int g) { return lds_build_layout(s.n_nodes, g, coef_size).bytes; };
      auto vjp_bytes = [&](int g) {
        return lds_table_bytes(s.n_nodes, g, coef_size) + pad4(s.n_nodes) * 16;
      };
      d.guide_log2 = g_want;
      while (d.guide_log2 > kMinGuideLog2 &&
             std::max(build_bytes(d.guide_log2), vjp_bytes(d.guide_log2)) > kMaxLds)
        --d.guide_log2;
      if (vjp_bytes(d.guide_log2) > kMaxLds) {  // the VJP takes bin ranges; fit the build only
        d.guide_log2 = g_want;
        while (d.guide_log2 > kMinGuideLog2 && build_bytes(d.guide_log2) > kMaxLds) --d.guide_log2;
        d.vjp_ranges = vjp_range_count(s.n_nodes, dtype == LOITS_HIP_F32);
      }
      d.global_tables = build_bytes(d.guide_log2) > kMaxLds ? 1 : 0;
      if (d.global_tables) d.guide_log2 = g_want;
      d.table_off = table_off;
      const TableLayout tl = table_layout(s.n_nodes, d.guide_log2, coef_size);
      table_off += align16(tl.bytes);
      const int64_t guide_bytes = (int64_t(2) << d.guide_log2) * 4;
      if (d.global_tables) {
        ++n_global;
        plan->global_list.push_back(int32_t(i));
        lds_f = std::max(lds_f, guide_bytes);
        lds_t = std::max(lds_t, lds_global_build_bytes(s.n_nodes, d.guide_log2,
                                                        dtype == LOITS_HIP_F32 ? 4 : 8));
      } else {
        lds_f = std::max(lds_f, build_bytes(d.guide_log2));
        lds_t = std::max(lds_t, build_bytes(d.guide_log2));
      }
      if (d.vjp_ranges > 0) {
        // (a curve that turns out not to be monotone takes the global-table VJP,
        // which stages the guide)
        ++n_ranged;
        lds_b = std::max(lds_b, std::max(guide_bytes,
                                         lds_range_bytes(vjp_range_width(s.n_nodes, d.vjp_ranges),
                                                         dtype == LOITS_HIP_F32)));
      } else {
        lds_b = std::max(lds_b, (tl.cdf - tl.thr) + pad4(s.n_nodes) * 16);
      }
    }
    L.max_nodes = std::max(L.max_nodes, s.n_nodes);
    L.total_nodes = std::max(L.total_nodes, s.node_off + s.n_nodes);
    // float32 bin-range VJPs sum q k as 64-bit integers: at most kFastMaxChunk
    // keys per chunk keep the sums below 2^62
    const int64_t ce = (dtype == LOITS_HIP_F32 && d.vjp_ranges > 0)
                           ? std::min<int64_t>(chunk_i, kFastMaxChunk)
                           : chunk_i;
    const int64_t nch = std::max<int64_t>(1, (s.events + ce - 1) / ce);
    if (plan->chunks.size() + size_t(nch) > size_t(INT32_MAX))
      return fail(LOITS_HIP_EUNSUPPORTED, "batched: too many tiles");
    d.first_chunk = int32_t(plan->chunks.size());
    d.n_chunks = int32_t(nch);
    for (int64_t c = 0; c < nch; ++c) {
      ChunkDesc ch{};
      ch.slot_lo = c * ce;
      ch.slot_hi = std::min(s.events, (c + 1) * ce);
      ch.instance = int32_t(i);
      ch.index = int32_t(c);
      const uint64_t p_lo = s.counter_base + uint64_t(ch.slot_lo);
      const uint64_t p_hi = s.counter_base + uint64_t(ch.slot_hi);
      const uint64_t blocks = ch.slot_hi > ch.slot_lo ? ((p_hi + 3) >> 2) - (p_lo >> 2) : 0;
      const int64_t per_seg = int64_t(blocks_per_wave(blocks, nsegs) / 64);
      ch.mask_off = mask_words;
      mask_words += per_seg * nsegs * 4;
      ch.partial_off = partial_doubles;  // VJP overflow sums, then partials
      partial_doubles += 2 * pad4(s.n_nodes);
      const int32_t ci = int32_t(plan->chunks.size());
      if (d.fast_log2 > 0) {
        plan->items.push_back(VjpItem{ci, 0});
        plan->fwd_order.push_back(ci);
      } else {
        guide_chunks.push_back(ci);
        if (d.vjp_ranges == 0)
          guide_items.push_back(VjpItem{ci, 0});
        else
          for (int r = 0; r < d.vjp_ranges; ++r) ranged.push_back(VjpItem{ci, r});
      }
      plan->chunks.push_back(ch);
    }
    out += s.events;
  }
  // Work-groups are dispatched in index order as CUs free up, so each launch's
  // chunks (and VJP items) go largest first: the long ones start early and the
  // short ones fill the end instead of one long chunk running alone.  Cost in
  // key units: the chunk's keys plus a table build or staging of kBuildKeysPerNode
  // keys per node.  Stable, so equal costs keep instance order.
  auto cost = [&](int32_t ci) {
    const ChunkDesc& ch = plan->chunks[size_t(ci)];
    return (ch.slot_hi - ch.slot_lo) + kBuildKeysPerNode * plan->inst[size_t(ch.instance)].n_nodes;
  };
  {
    auto by_cost = [&](int32_t a, int32_t b) { return cost(a) > cost(b); };
    auto item_by_cost = [&](const VjpItem& a, const VjpItem& b) {
      return cost(a.chunk) > cost(b.chunk);
    };
    std::stable_sort(plan->fwd_order.begin(), plan->fwd_order.end(), by_cost);
    std::stable_sort(guide_chunks.begin(), guide_chunks.end(), by_cost);
    std::stable_sort(plan->items.begin(), plan->items.end(), item_by_cost);
    std::stable_sort(guide_items.begin(), guide_items.end(), item_by_cost);
    std::stable_sort(ranged.begin(), ranged.end(), item_by_cost);
  }
  L.capacity = out;
  L.n_chunks = int64_t(plan->chunks.size());
  L.cell_chunks = int64_t(plan->fwd_order.size());
  L.backward_cell = int64_t(plan->items.size());
  plan->fwd_order.insert(plan->fwd_order.end(), guide_chunks.begin(), guide_chunks.end());
  plan->items.insert(plan->items.end(), guide_items.begin(), guide_items.end());
  plan->n_plain = int64_t(plan->items.size());
  plan->items.insert(plan->items.end(), ranged.begin(), ranged.end());
  if (plan->items.size() > size_t(INT32_MAX))
    return fail(LOITS_HIP_EUNSUPPORTED, "batched: too many VJP work items");
  // Persistent queues: every forward chunk and every VJP item, largest first.
  if (L.persistent) {
    plan->fwd_queue = plan->fwd_order;
    std::stable_sort(plan->fwd_queue.begin(), plan->fwd_queue.end(),
                     [&](int32_t a, int32_t b) { return cost(a) > cost(b); });
    plan->vjp_queue.resize(plan->items.size());
    for (size_t k = 0; k < plan->items.size(); ++k) plan->vjp_queue[k] = int32_t(k);
    std::stable_sort(plan->vjp_queue.begin(), plan->vjp_queue.end(), [&](int32_t a, int32_t b) {
      return cost(plan->items[size_t(a)].chunk) > cost(plan->items[size_t(b)].chunk);
    });
  }
  L.forward_items = int64_t(plan->fwd_queue.size());
  // the global-table builds (a launch of their own): largest first
  std::stable_sort(plan->global_list.begin(), plan->global_list.end(), [&](int32_t a, int32_t b) {
    return plan->inst[size_t(a)].n_nodes > plan->inst[size_t(b)].n_nodes;
  });
  L.desc_bytes = align16(n * int64_t(sizeof(InstanceDesc))) +
                 align16(L.n_chunks * int64_t(sizeof(ChunkDesc))) +
                 align16(int64_t(plan->items.size()) * int64_t(sizeof(VjpItem))) +
                 align16(L.n_chunks * int64_t(sizeof(int32_t))) +
                 align16(L.forward_items * int64_t(sizeof(int32_t))) +
                 align16(int64_t(plan->vjp_queue.size()) * int64_t(sizeof(int32_t))) +
                 int64_t(plan->global_list.size()) * int64_t(sizeof(int32_t));
  L.table_bytes = std::max<int64_t>(table_off, 16);
  L.lds_forward = lds_f;
  L.lds_backward = lds_b;
  L.lds_tables = lds_t;
  // the persistent forward samples every kind of chunk (the global-table builds
  // stay a launch of their own)
  L.lds_persistent_forward = lds_f;
  L.n_segments = L.n_chunks * nsegs;
  if (lds_f > kMaxLds || lds_t > kMaxLds)
    return fail(LOITS_HIP_EUNSUPPORTED,
                "batched: curve tables of " + std::to_string(L.max_nodes) +
                    " nodes exceed the LDS budget (" + std::to_string(std::max(lds_f, lds_b)) +
                    " > " + std::to_string(kMaxLds) + " bytes)");
  if ((L.max_nodes + threads - 1) / threads > 32)
    return fail(LOITS_HIP_EUNSUPPORTED, "batched: n_nodes / threads must be <= 32");
  L.mask_words = mask_words;
  L.partial_doubles = partial_doubles;
  // scratch: flags | seg_kept | chunk_kept, chunk_row (in n_segments words) | seg_row |
  // seg_drop | done | keepmask | partials
  L.scratch_bytes = 64 + align16(8 * L.n_segments) * 3 + align16(4 * L.n_segments) + align16(4 * n) +
                    align16(8 * mask_words) + align16(8 * partial_doubles);
  L.grid_forward = int32_t(L.n_chunks);
  L.grid_backward = int32_t(plan->items.size());
  L.backward_plain = plan->n_plain;
  L.global_instances = n_global;
  L.range_instances = n_ranged;
  L.cell_instances = n_cell;
  return LOITS_HIP_OK;
}

BatchedParams params_of(const loits_hip_batched_args* a) {
  const loits_hip_layout& L = *a->layout;
  BatchedParams P{};
  const uint8_t* desc = static_cast<const uint8_t*>(a->desc);
  P.inst = reinterpret_cast<const InstanceDesc*>(desc);
  const int64_t chunks_at = align16(L.n_instances * int64_t(sizeof(InstanceDesc)));
  P.chunks = reinterpret_cast<const ChunkDesc*>(desc + chunks_at);
  const int64_t items_at = chunks_at + align16(L.n_chunks * int64_t(sizeof(ChunkDesc)));
  P.vjp_items = reinterpret_cast<const int2*>(desc + items_at);
  const int64_t order_at = items_at + align16(int64_t(L.grid_backward) * int64_t(sizeof(VjpItem)));
  P.fwd_order = reinterpret_cast<const int32_t*>(desc + order_at);
  const int64_t fq_at = order_at + align16(L.n_chunks * int64_t(sizeof(int32_t)));
  P.fwd_queue = reinterpret_cast<const int32_t*>(desc + fq_at);
  const int64_t vq_at = fq_at + align16(L.forward_items * int64_t(sizeof(int32_t)));
  P.vjp_queue = reinterpret_cast<const int32_t*>(desc + vq_at);
  // the VJP queue holds every item in persistent mode, none otherwise
  const int64_t vjp_queue_len = L.persistent ? int64_t(L.grid_backward) : 0;
  P.global_list = reinterpret_cast<const int32_t*>(
      desc + vq_at + align16(vjp_queue_len * int64_t(sizeof(int32_t))));
  P.n_global = L.global_instances;
  P.n_fwd_items = L.forward_items;
  P.n_instances = L.n_instances;
  P.n_chunks = L.n_chunks;
  P.n_vjp_items = L.grid_backward;
  P.n_vjp_plain = L.backward_plain;
  P.n_cell_chunks = L.cell_chunks;
  P.n_vjp_cell = L.backward_cell;
  P.x = a->x;
  P.density = a->density;
  P.tables = static_cast<uint8_t*>(a->tables);
  P.seed = a->seed;
  P.use_window = a->use_window;
  P.win_lo = a->window_lo;
  P.win_hi = a->window_hi;
  P.events = a->events;
  P.bins = a->bins;
  P.kept = a->kept;
  P.total = a->total;
  P.norm_out = a->norm;
  P.cdf_out = a->cdf;
  uint8_t* s = static_cast<uint8_t*>(a->scratch);
  P.n_segments = L.n_segments;
  P.segs = (L.threads / 64) * kSegsPerWave;
  P.flags = reinterpret_cast<int32_t*>(s);
  s += 64;
  P.seg_kept = reinterpret_cast<int64_t*>(s);
  s += align16(8 * L.n_segments);
  P.chunk_kept = reinterpret_cast<int64_t*>(s);  // n_segments >= 2 n_chunks (>= 16 per chunk)
  P.chunk_row = P.chunk_kept + L.n_chunks;
  s += align16(8 * L.n_segments);
  P.seg_row = reinterpret_cast<int64_t*>(s);
  s += align16(8 * L.n_segments);
  P.seg_drop = reinterpret_cast<int32_t*>(s);
  s += align16(4 * L.n_segments);
  P.done = reinterpret_cast<uint32_t*>(s);
  s += align16(4 * L.n_instances);
  P.keepmask = reinterpret_cast<uint64_t*>(s);
  s += align16(8 * L.mask_words);
  P.partials = reinterpret_cast<double*>(s);
  return P;
}

int check_args(const loits_hip_batched_args* a) {
  if (!a || !a->layout) return fail(LOITS_HIP_EINVAL, "batched: args or layout is NULL");
  if (!a->desc || !a->x || !a->density || !a->tables || !a->scratch || !a->events || !a->kept ||
      !a->total)
    return fail(LOITS_HIP_EINVAL, "batched: a required pointer is NULL");
  if (a->use_window && !(a->window_lo <= a->window_hi))
    return fail(LOITS_HIP_EINVAL, "batched: window_lo must be <= window_hi");
  return LOITS_HIP_OK;
}

}  // namespace

}  // namespace qloits

using namespace qloits;

extern "C" {

const char* loits_hip_last_error(void) { return g_last_error.c_str(); }

const char* loits_hip_version(void) { return "quantom-loits-hip 0.1 gfx950"; }

int loits_hip_device_info(int device, char* arch, int arch_len, int* cu_count) {
  hipDeviceProp_t prop;
  const hipError_t e = hipGetDeviceProperties(&prop, device);
  if (e != hipSuccess) return check_hip(e, "hipGetDeviceProperties");
  if (arch && arch_len > 0) {
    std::strncpy(arch, prop.gcnArchName, size_t(arch_len - 1));
    arch[arch_len - 1] = '\0';
  }
  if (cu_count) *cu_count = prop.multiProcessorCount;
  return LOITS_HIP_OK;
}

int loits_hip_synchronize(void* stream) {
  return check_hip(hipStreamSynchronize(static_cast<hipStream_t>(stream)), "hipStreamSynchronize");
}

int loits_hip_philox_fill(float* values, int64_t count, uint64_t seed, uint64_t stream_id,
                          void* stream) {
  if (count < 0 || (count > 0 && !values))
    return fail(LOITS_HIP_EINVAL, "philox_fill: bad count or NULL output");
  if (count == 0) return LOITS_HIP_OK;
  const int64_t blocks = (count + 3) / 4;
  const int64_t grid = (blocks + 255) / 256;
  if (grid > INT32_MAX) return fail(LOITS_HIP_EUNSUPPORTED, "philox_fill: count too large");
  hipLaunchKernelGGL(philox_fill_kernel, dim3(unsigned(grid)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), values, count, seed, stream_id);
  return check_hip(hipGetLastError(), "philox_fill launch");
}

int loits_hip_philox_words(uint32_t* words, int64_t first_block, int64_t blocks,
                           uint64_t stream_id, uint64_t seed, void* stream) {
  if (blocks < 0 || (blocks > 0 && !words) || first_block < 0)
    return fail(LOITS_HIP_EINVAL, "philox_words: bad arguments");
  if (blocks == 0) return LOITS_HIP_OK;
  if ((reinterpret_cast<uintptr_t>(words) & 15) != 0)
    return fail(LOITS_HIP_EINVAL, "philox_words: output must be 16-byte aligned");
  const int64_t grid = (blocks + 255) / 256;
  if (grid > INT32_MAX) return fail(LOITS_HIP_EUNSUPPORTED, "philox_words: too many blocks");
  hipLaunchKernelGGL(philox_words_kernel, dim3(unsigned(grid)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), words, first_block, blocks, stream_id,
                     seed);
  return check_hip(hipGetLastError(), "philox_words launch");
}

int loits_hip_batched_layout(const loits_hip_instance* instances, int64_t n, int dtype,
                             const loits_hip_tuning* tuning, loits_hip_layout* layout) {
  try {
    if (!layout) return fail(LOITS_HIP_EINVAL, "batched_layout: layout is NULL");
    Plan plan;
    const int st = make_plan(instances, n, dtype, tuning, &plan);
    if (st != LOITS_HIP_OK) return st;
    *layout = plan.layout;
    return LOITS_HIP_OK;
  } catch (const std::exception& e) {
    return fail(LOITS_HIP_ERUNTIME, std::string("batched_layout: ") + e.what());
  }
}

int loits_hip_batched_encode(const loits_hip_instance* instances, int64_t n,
                             const loits_hip_tuning* tuning, const loits_hip_layout* layout,
                             void* host_desc) {
  try {
    if (!layout || !host_desc) return fail(LOITS_HIP_EINVAL, "batched_encode: NULL argument");
    Plan plan;
    const int st = make_plan(instances, n, layout->dtype, tuning, &plan);
    if (st != LOITS_HIP_OK) return st;
    if (std::memcmp(&plan.layout, layout, sizeof(loits_hip_layout)) != 0)
      return fail(LOITS_HIP_EINVAL, "batched_encode: layout does not match instances/tuning");
    uint8_t* out = static_cast<uint8_t*>(host_desc);
    std::memset(out, 0, size_t(layout->desc_bytes));
    std::memcpy(out, plan.inst.data(), plan.inst.size() * sizeof(InstanceDesc));
    const int64_t chunks_at = align16(n * int64_t(sizeof(InstanceDesc)));
    std::memcpy(out + chunks_at, plan.chunks.data(), plan.chunks.size() * sizeof(ChunkDesc));
    const int64_t items_at =
        chunks_at + align16(int64_t(plan.chunks.size() * sizeof(ChunkDesc)));
    std::memcpy(out + items_at, plan.items.data(), plan.items.size() * sizeof(VjpItem));
    const int64_t order_at = items_at + align16(int64_t(plan.items.size() * sizeof(VjpItem)));
    std::memcpy(out + order_at, plan.fwd_order.data(), plan.fwd_order.size() * sizeof(int32_t));
    const int64_t fq_at = order_at + align16(int64_t(plan.fwd_order.size() * sizeof(int32_t)));
    if (!plan.fwd_queue.empty())
      std::memcpy(out + fq_at, plan.fwd_queue.data(), plan.fwd_queue.size() * sizeof(int32_t));
    const int64_t vq_at = fq_at + align16(int64_t(plan.fwd_queue.size() * sizeof(int32_t)));
    if (!plan.vjp_queue.empty())
      std::memcpy(out + vq_at, plan.vjp_queue.data(), plan.vjp_queue.size() * sizeof(int32_t));
    if (!plan.global_list.empty())
      std::memcpy(out + vq_at + align16(int64_t(plan.vjp_queue.size() * sizeof(int32_t))),
                  plan.global_list.data(), plan.global_list.size() * sizeof(int32_t));
    return LOITS_HIP_OK;
  } catch (const std::exception& e) {
    return fail(LOITS_HIP_ERUNTIME, std::string("batched_encode: ") + e.what());
  }
}

int loits_hip_batched_forward(const loits_hip_batched_args* args, void* stream) {
  try {
    const int st = check_args(args);
    if (st != LOITS_HIP_OK) return st;
    const loits_hip_layout& L = *args->layout;
    const BatchedParams P = params_of(args);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const bool f32 = L.dtype == LOITS_HIP_F32;
    const bool seq = L.scan_mode == 1;
    if (L.global_instances > 0) {  // global-table instances are built first
      const int tthreads = std::min(L.threads, 1024);
      const int rt = check_hip(
          f32 ? launch_tables<float>(P, tthreads, seq, L.lds_tables, true, s)
              : launch_tables<double>(P, tthreads, seq, L.lds_tables, true, s),
          "batched tables launch");
      if (rt != LOITS_HIP_OK) return rt;
    }
    if (L.persistent) {  // one launch for every chunk, from the queue
      if (L.n_chunks == 0) return check_hip(launch_finalize(P, s), "batched finalize launch");
      return check_hip(
          f32 ? launch_forward_persistent<float>(P, L.threads, seq, L.lds_persistent_forward, s)
              : launch_forward_persistent<double>(P, L.threads, seq, L.lds_persistent_forward, s),
          "batched persistent forward launch");
    }
    int rc = check_hip(f32 ? launch_sample<float>(P, L.threads, seq, L.lds_forward, s)
                           : launch_sample<double>(P, L.threads, seq, L.lds_forward, s),
                       "batched sample launch");
    if (rc != LOITS_HIP_OK) return rc;
    // the last sampling work-group finalizes (row offsets, totals, dense flag)
    return L.n_chunks > 0 ? LOITS_HIP_OK : check_hip(launch_finalize(P, s), "batched finalize launch");
  } catch (const std::exception& e) {
    return fail(LOITS_HIP_ERUNTIME, std::string("batched_forward: ") + e.what());
  }
}

int loits_hip_batched_pack(const loits_hip_batched_args* args, void* packed, int32_t* packed_bins,
                           void* stream) {
  try {
    const int st = check_args(args);
    if (st != LOITS_HIP_OK) return st;
    if (!packed) return fail(LOITS_HIP_EINVAL, "batched_pack: packed is NULL");
    const BatchedParams P = params_of(args);
    hipStream_t s = static_cast<hipStream_t>(stream);
    return check_hip(args->layout->dtype == LOITS_HIP_F32
                         ? launch_pack<float>(P, args->layout->threads, packed, packed_bins, s)
                         : launch_pack<double>(P, args->layout->threads, packed, packed_bins, s),
                     "batched pack launch");
  } catch (const std::exception& e) {
    return fail(LOITS_HIP_ERUNTIME, std::string("batched_pack: ") + e.what());
  }
}

int loits_hip_batched_backward(const loits_hip_batched_args* args, const void* grad_events,
                               void* grad_density, int32_t recompute_tables, void* stream) {
  try {
    const int st = check_args(args);
    if (st != LOITS_HIP_OK) return st;
    if (!grad_events && args->layout->capacity > 0)
      return fail(LOITS_HIP_EINVAL, "batched_backward: grad_events is NULL");
    if (!grad_density) return fail(LOITS_HIP_EINVAL, "batched_backward: grad_density is NULL");
    const loits_hip_layout& L = *args->layout;
    if (L.lds_backward > kMaxLds)
      return fail(LOITS_HIP_EUNSUPPORTED,
                  "batched_backward: curve tables of " + std::to_string(L.max_nodes) +
                      " nodes and their histograms exceed the LDS budget (" +
                      std::to_string(L.lds_backward) + " > " + std::to_string(kMaxLds) + " bytes)");
    const BatchedParams P = params_of(args);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const bool f32 = L.dtype == LOITS_HIP_F32;
    if (recompute_tables) {
      BatchedParams Q = P;
      Q.norm_out = nullptr;
      Q.cdf_out = nullptr;
      const int tthreads = std::min(L.threads, 1024);
      const int rc = check_hip(f32 ? launch_tables<float>(Q, tthreads, L.scan_mode == 1, L.lds_tables, false, s)
                                   : launch_tables<double>(Q, tthreads, L.scan_mode == 1, L.lds_tables, false, s),
                               "batched tables launch");
      if (rc != LOITS_HIP_OK) return rc;
    }
    if (L.persistent)
      return check_hip(f32 ? launch_vjp_persistent<float>(P, L.threads, L.lds_backward,
                                                          grad_events, grad_density, s)
                           : launch_vjp_persistent<double>(P, L.threads, L.lds_backward,
                                                           grad_events, grad_density, s),
                       "batched persistent vjp launch");
    return check_hip(f32 ? launch_vjp<float>(P, L.threads, L.lds_backward, grad_events,
                                             grad_density, s)
                         : launch_vjp<double>(P, L.threads, L.lds_backward, grad_events,
                                              grad_density, s),
                     "batched vjp launch");
  } catch (const std::exception& e) {
    return fail(LOITS_HIP_ERUNTIME, std::string("batched_backward: ") + e.what());
  }
}

}  // extern "C"
