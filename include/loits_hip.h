/*
 * loits_hip.h -- C ABI of the MI355X (gfx950) LOITS sampler.
 *
 * Plain pointers and sizes only.  Every device pointer is memory the caller
 * owns (PyTorch tensors in the Python package); the library allocates
 * nothing, frees nothing, and enqueues all work on the hipStream_t passed as
 * `stream` (NULL = the legacy default stream).  Every entry point returns a
 * loits_hip_status; on failure loits_hip_last_error() holds a message
 * (thread-local).  No C++ exception crosses this boundary.
 *
 * Two families of entry points:
 *
 *  1. The reference drop-in (2-axis, one call = one reference `forward` /
 *     `backward`), replacing the native cores behind cpp/bindings.cpp:24-215
 *     (cpp/loits_core.hpp:31-137, openmp/loits_core.hpp:352-459):
 *       loits_hip_allocate, loits_hip_forward, loits_hip_backward.
 *  2. The batched 1-D path: many independent x-curves ("instances") of
 *     uneven node and event counts sampled in single launches, with the
 *     reverse VJP recomputing the Philox stream instead of storing it:
 *       loits_hip_batched_layout, loits_hip_batched_encode,
 *       loits_hip_batched_forward, loits_hip_batched_pack,
 *       loits_hip_batched_backward.
 *     It has no single reference counterpart; under the 1-D mapping of the
 *     reference (nx = ny = 1, weights 1, counter base b*E, stream 2*sequence)
 *     its events, bin indices and gradients equal the x column of the
 *     reference forward/backward (cpp/bindings.cpp:24-208).
 *
 * Plus the RNG probe loits_hip_philox_fill / loits_hip_philox_words
 * (rng/philox.hpp:17-61, cpp/loits_core.cpp:39-53).
 */
#ifndef QUANTOM_LOITS_HIP_H
#define QUANTOM_LOITS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum loits_hip_status {
  LOITS_HIP_OK = 0,
  LOITS_HIP_EINVAL = 1,      /* bad argument (shape, dtype, size, pointer) */
  LOITS_HIP_ERUNTIME = 2,    /* a HIP runtime call failed */
  LOITS_HIP_EUNSUPPORTED = 3 /* a size the kernels do not handle */
} loits_hip_status;

typedef enum loits_hip_dtype { LOITS_HIP_F64 = 0, LOITS_HIP_F32 = 1 } loits_hip_dtype;

/* Thread-local message of the last failed call on this thread ("" if none). */
const char* loits_hip_last_error(void);

/* Library version string, e.g. "quantom-loits-hip 0.1 gfx950". */
const char* loits_hip_version(void);

/* Device probe: writes the gfx architecture name of `device` (e.g. "gfx950")
 * and its compute-unit count.  Replaces the availability probes of the
 * reference registry (loits.py:13-46). */
int loits_hip_device_info(int device, char* arch, int arch_len, int* cu_count);

/* Blocks until the work queued on `stream` (NULL: the default stream) is done.
 * The profiling regions of the bindings drain the stream at both ends with it,
 * as the reference's GPU binding waits on its queue when a region opens
 * (sycl/bindings.cpp:13-27). */
int loits_hip_synchronize(void* stream);

/* ---------------------------------------------------------------------- */
/* RNG (rng/philox.hpp:17-61)                                              */

/* values[p] = uniform_float(word p % 4 of philox4x32_10(p / 4, stream, seed))
 * for p < count: cpp/loits_core.cpp:39-53. */
int loits_hip_philox_fill(float* values, int64_t count, uint64_t seed, uint64_t stream_id,
                          void* stream);

/* words[4 i + w] = word w of philox4x32_10(first_block + i, stream_id, seed). */
int loits_hip_philox_words(uint32_t* words, int64_t first_block, int64_t blocks,
                           uint64_t stream_id, uint64_t seed, void* stream);

/* ---------------------------------------------------------------------- */
/* Reference drop-in (2-axis)                                               */

/* Shape of one reference call: cpp/loits_core.hpp:17-24 (cells = nx * ny). */
typedef struct loits_hip_shape {
  int64_t batch, nx, ny, kx, kq, cells;
} loits_hip_shape;

/* counts[i] = (int64)|weights[i] * n_events| for i < n, and
 * nmax_active[0] = max count, nmax_active[1] = sum; all device memory.
 * Replaces allocate_counts, cpp/loits_core.cpp:22-37 (the reference returns
 * nmax/active to the host; the caller reads nmax_active after the stream).
 * NaN weights give count 0; counts saturate at 2^62. */
int loits_hip_allocate(const void* weights, int dtype, int64_t n, int64_t n_events,
                       int64_t* counts, int64_t* nmax_active, void* stream);

typedef struct loits_hip_forward_args {
  loits_hip_shape shape;
  int dtype;             /* of every floating tensor below */
  int64_t nmax, active;  /* from loits_hip_allocate */
  uint64_t seed, sequence;
  /* inputs (device, contiguous, reference layouts cpp/loits_core.cpp:62-66, 90-94) */
  const void* x_bins;     /* [B, nx, kx] */
  const void* xsec_x;     /* [B, ny, nx, kx] */
  const void* q_bins;     /* [B, ny, kq] */
  const void* xsec_q;     /* [B, nx, ny, kq] */
  const uint8_t* acceptance; /* [B, nx, ny], 0 or 1 */
  const int64_t* counts;  /* [B, cells] */
  /* outputs (device) */
  double* norm_x;         /* [B, ny, nx]  (float64 whatever dtype is) */
  double* norm_q;         /* [B, nx, ny] */
  double* cdf_x;          /* like xsec_x, float64 */
  double* cdf_q;          /* like xsec_q, float64 */
  float* u_x;             /* [B, cells, nmax], every padded slot (u_x, u_q 16-byte aligned) */
  float* u_q;
  int16_t* interval_x;    /* [B*cells*nmax]; 0 in inactive slots (8-byte aligned) */
  int16_t* interval_q;
  void* events;           /* [active, 2] */
  int64_t* packed;        /* [active] */
  int64_t* row_offsets;   /* [B*cells + 1] or NULL */
  void* workspace;        /* loits_hip_forward_workspace_size() bytes */
  int64_t* valid;         /* device out [2]: kept events, error flags (0 = ok; bit 0: a
                             look-back wait ran out, bit 1: bounded capacities exceeded,
                             bit 2: a slot of a drop-free curve was dropped) */
} loits_hip_forward_args;

/* Bytes of device workspace loits_hip_forward needs for this call. */
int64_t loits_hip_forward_workspace_size(const loits_hip_shape* shape, int64_t nmax);

/* The reference forward (cpp/bindings.cpp:24-130): densities, CDFs, both
 * Philox streams (2*sequence, 2*sequence+1), intervals, interpolation and the
 * stable compaction of cpp/loits_core.cpp:239-266 (single pass, ordered by a
 * decoupled look-back over tiles, or, when every curve is drop-free, by the
 * active slots before each tile).  The CDF chain follows the reference's
 * sequential order, so norm, cdf, u, intervals, events and packed equal the
 * reference's.  No host synchronisation: valid[0] holds the kept count. */
int loits_hip_forward(const loits_hip_forward_args* args, void* stream);

/* Sync-free form of loits_hip_forward, for callers that cannot read nmax and
 * active on the host between allocation and sampling (HIP graph capture, a
 * loop that keeps buffers across calls).  args->nmax and args->active are
 * CAPACITIES: they size u_x/u_q/interval_x/interval_q (B*cells*nmax_cap),
 * events/packed (active_cap) and the workspace; the true nmax and active are
 * read on the device from nmax_active (written by loits_hip_allocate earlier
 * on the stream).  The reference's arrays are then prefixes of the capacity
 * buffers (u_x: the first B*cells*nmax floats, [B, cells, nmax]; events: the
 * first valid[0] rows).  If nmax or active exceeds its capacity nothing is
 * sampled and valid[1] gets bit 1 (value 2).  No reference counterpart: the
 * reference reads the sizes on the host (cpp/bindings.cpp:80-92); this entry
 * point is for the callers that cannot. */
int loits_hip_forward_bounded(const loits_hip_forward_args* args, const int64_t* nmax_active,
                              void* stream);

/* Bytes of the device scratch array behind loits_hip_forward_noalloc (one
 * copy per GPU the library runs on, held until the process ends). */
#define LOITS_HIP_NOALLOC_SCRATCH_BYTES (16 * 1024 * 1024)

/* Allocation-free form of loits_hip_forward: identical outputs, but the
 * workspace is the library's own static device scratch
 * (LOITS_HIP_NOALLOC_SCRATCH_BYTES per GPU), so the call makes no
 * hipMalloc/hipFree and no allocating library call.  args->workspace must be
 * NULL; a shape whose workspace (loits_hip_forward_workspace_size) exceeds
 * the scratch is rejected with LOITS_HIP_EINVAL.  Every call on one GPU
 * shares that GPU's scratch and clears what must start at zero itself, so
 * calls that overlap on one GPU (concurrent streams or threads) race and must
 * be serialised by the caller.  No reference counterpart: it is for callers
 * that cannot allocate (HIP graph capture, tight serving loops, harnesses that
 * forbid allocation). */
int loits_hip_forward_noalloc(const loits_hip_forward_args* args, void* stream);

typedef struct loits_hip_backward_args {
  loits_hip_shape shape;
  int dtype;
  int64_t nmax, nvalid;
  const void* grad_events;  /* [nvalid, 2] */
  const void* x_bins;
  const void* xsec_x;
  const void* q_bins;
  const void* xsec_q;
  const uint8_t* acceptance;
  const double* norm_x;
  const double* norm_q;
  const double* cdf_x;
  const double* cdf_q;
  const float* u_x;
  const float* u_q;
  const int16_t* interval_x;
  const int16_t* interval_q;
  const int64_t* packed;
  void* grad_xsec_x;        /* like xsec_x */
  void* grad_xsec_q;        /* like xsec_q */
  void* workspace;          /* loits_hip_backward_workspace_size() bytes */
  /* Optional, NULL to find them from `packed`: the rows of cell g are
   * [row_offsets[g], row_offsets[g + 1]) (batch * cells + 1 entries), as the
   * forward writes them and the reference's VJP reads them
   * (openmp/loits_core.cpp:307, :341).  Ranges are clamped to [0, nvalid). */
  const int64_t* row_offsets;
} loits_hip_backward_args;

int64_t loits_hip_backward_workspace_size(const loits_hip_shape* shape, int64_t nmax);

/* The reference backward (cpp/bindings.cpp:132-208): interpolation VJP
 * (cpp/loits_core.cpp:268-338, rows of one cell in row order as
 * openmp/loits_core.cpp:285-352), CDF VJP (:340-398), density VJP (:400-472). */
int loits_hip_backward(const loits_hip_backward_args* args, void* stream);

/* Sync-free form of loits_hip_backward after loits_hip_forward_bounded:
 * nmax is read from nmax_active[0] and the row count from valid[0] on the
 * device (args->nmax, args->nvalid are the capacities).  No reference
 * counterpart (see loits_hip_forward_bounded). */
int loits_hip_backward_bounded(const loits_hip_backward_args* args, const int64_t* nmax_active,
                               const int64_t* valid, void* stream);

/* ---------------------------------------------------------------------- */
/* Batched 1-D path                                                         */

/* One x-curve.  Its nodes are x[node_off .. node_off + n_nodes) with the
 * density at the same offsets; slot s (0 <= s < events) draws the uniform of
 * Philox counter p = counter_base + s on stream `stream` (word p % 4 of block
 * p / 4), exactly as the reference's slot p does (cpp/bindings.cpp:79-92).
 * Capacity row out_off + s belongs to slot s: the forward writes each slot's
 * value there, kept or not, with a keep mask; loits_hip_batched_pack gathers
 * the kept rows.  When every slot is kept the capacity layout is already the
 * packed output. */
typedef struct loits_hip_instance {
  int64_t node_off;
  int64_t n_nodes;      /* >= 2 */
  int64_t events;       /* >= 0 */
  int64_t out_off;      /* sum of `events` of the instances before it */
  uint64_t counter_base;
  uint64_t stream;
  int32_t accept;       /* acceptance flag: 0 multiplies the CDF by 0 */
  int32_t reserved;
} loits_hip_instance;

typedef struct loits_hip_tuning {
  int32_t threads;          /* work-group size: 256, 512 or 1024 */
  int32_t guide_shift;      /* guide cells = next_pow2(n_nodes) << guide_shift (0..3), clamped
                               to [2^12, 2^14] and lowered until the tables fit in LDS */
  int64_t chunk_events;     /* events per work-group (rounded up to 4*threads) */
  int32_t scan_mode;        /* 0 = parallel CDF scan, 1 = the reference's sequential chain */
  int32_t persistent;       /* 0 = one launch per kind of work: cell-table chunks, guide-table
                               chunks (forward); cell, guide and bin-range items (VJP).
                               1 = one persistent launch for the forward and one for the VJP:
                               resident work-groups take every chunk (VJP item) from a device
                               queue, largest first, until none is left.  Either way curves
                               whose per-bin tables exceed LDS ("global tables") are built by
                               a launch of their own before the forward's sampling. */
  int32_t table_format;     /* 0 = cell tables for float32 curves of <= 4096 nodes when
                               their chunks are <= 2^18 events (guide tables otherwise),
                               1 = guide tables for every curve */
  int32_t reserved;         /* must be 0 */
  /* Per grid-size bucket b = ceil(log2(n_nodes)) (curves of 2^(b-1) + 1 .. 2^b nodes,
   * b in 1..15); 0 keeps the values above for that bucket. */
  int32_t bucket_chunk_log2[16];   /* events per chunk = 2^v, v in 10..30 */
  int32_t bucket_guide_shift[16];  /* guide_shift + 1, in 1..4 */
} loits_hip_tuning;

typedef struct loits_hip_layout {
  int64_t n_instances, n_chunks, capacity, total_nodes, max_nodes;
  int64_t desc_bytes;       /* descriptor blob (host-encoded, copied to device) */
  int64_t table_bytes;      /* per-call tables (also the saved state) */
  int64_t scratch_bytes;    /* per-call scratch (counters, partial VJP sums) */
  int64_t lds_forward;      /* dynamic LDS of the sampling kernel */
  int64_t lds_backward;     /* dynamic LDS of the VJP kernel */
  int64_t lds_tables;       /* dynamic LDS of the standalone table build (VJP recompute) */
  int64_t n_segments;       /* chunks * segments per chunk (4 per wave of a work-group) */
  int64_t mask_words;       /* 64-bit keep-mask words in scratch */
  int64_t partial_doubles;  /* partial VJP sums in scratch (multi-chunk instances) */
  int32_t dtype, threads, scan_mode, persistent;
  int32_t grid_forward, grid_backward;
  int32_t global_instances; /* instances whose per-bin tables stay in global memory
                               (too large for the forward's LDS build; built by a
                               pre-pass, slower) */
  int32_t range_instances;  /* instances whose VJP splits the bins into ranges of at
                               most 7168 (float32) or 5120 (float64) bins (tables and
                               histograms too large for LDS); grid_backward counts one
                               work item per chunk and range */
  int32_t cell_instances;   /* instances on cell tables (table_format 0) */
  int32_t reserved;
  int64_t backward_plain;   /* work items of the instances without ranges (the cell-table
                               items, then the guide-table items; the ranged items follow) */
  int64_t cell_chunks;      /* forward work-groups of the cell-table instances (their own
                               launch; the other chunks take a second one) */
  int64_t backward_cell;    /* VJP work items of the cell-table instances */
  int64_t forward_items;    /* persistent forward queue length (its chunks) */
  int64_t lds_persistent_forward;  /* dynamic LDS of the persistent forward (every kind of chunk) */
} loits_hip_layout;

/* Host only: validate `instances` and compute sizes for `dtype`. */
int loits_hip_batched_layout(const loits_hip_instance* instances, int64_t n, int dtype,
                             const loits_hip_tuning* tuning, loits_hip_layout* layout);

/* Host only: write the descriptor blob (layout->desc_bytes bytes) to `host_desc`;
 * the caller copies it to device memory once and reuses it across calls. */
int loits_hip_batched_encode(const loits_hip_instance* instances, int64_t n,
                             const loits_hip_tuning* tuning, const loits_hip_layout* layout,
                             void* host_desc);

typedef struct loits_hip_batched_args {
  const loits_hip_layout* layout;
  const void* desc;        /* device copy of the descriptor blob */
  const void* x;           /* flat nodes, layout->dtype */
  const void* density;     /* flat densities, layout->dtype */
  uint64_t seed;
  int32_t use_window;      /* keep only window_lo <= value <= window_hi */
  int32_t reserved;
  double window_lo, window_hi;
  void* tables;            /* layout->table_bytes */
  void* scratch;           /* layout->scratch_bytes, zero-filled before its first call;
                              the kernels return their counters to zero */
  void* events;            /* [capacity] values, layout->dtype */
  int64_t* kept;           /* [n_instances] kept events per instance */
  int64_t* total;          /* device scalar: total kept */
  double* norm;            /* [n_instances] or NULL */
  double* cdf;             /* [total_nodes] or NULL */
  int32_t* bins;           /* [capacity] bin index per kept event, or NULL */
} loits_hip_batched_args;

/* Forward: curve tables, Philox uniforms, interval search, interpolation and
 * per-instance stable compaction into the capacity layout.  No host sync. */
int loits_hip_batched_forward(const loits_hip_batched_args* args, void* stream);

/* Gather the kept rows of the capacity layout into a contiguous [total] array
 * (stable order); does nothing when every slot was kept (the capacity layout
 * is then already packed).
 * `packed` has capacity rows; only the first *total are written. */
int loits_hip_batched_pack(const loits_hip_batched_args* args, void* packed, int32_t* packed_bins,
                           void* stream);

/* Reverse VJP w.r.t. the densities.  grad_events holds one cotangent per kept
 * event in packed (contiguous) order.  Tables come from the forward call
 * (layout->scan_mode decides how they were built); the Philox stream is
 * recomputed.  recompute_tables != 0 rebuilds them from x/density first. */
int loits_hip_batched_backward(const loits_hip_batched_args* args, const void* grad_events,
                               void* grad_density, int32_t recompute_tables, void* stream);

/* ---------------------------------------------------------------------- */
/* Soft-binning histograms of packed events, the discriminator input of the
 * GAN-style training harness.  No reference counterpart (the reference's
 * discriminator scores events directly, pytorch/gan.py:11-23).  For instance i
 * with rows [row_start[i], row_start[i+1]) of `values` (float32):
 *   hist[i][c] = sum_r exp(-(t_r - c - 1/2)^2 / (2 width^2)),  t = (v - lo) bins / (hi - lo),
 * over the seven bins nearest t_r; `hist` [n_instances, bins] is overwritten.
 * The backward writes d(sum_ic ghist[i][c] hist[i][c]) / d v_r to grad_values.
 * row_start is a device array of n_instances + 1 ascending offsets; bins <= 128.
 * NaN values take no part (no histogram weight, gradient 0); values outside
 * [lo, hi] by more than three bins reach no bin; rows before row_start[0] or
 * from row_start[n_instances] on take no part either (gradient 0).
 * Float atomics: sums may differ from run to run at float rounding. */
int loits_hip_soft_histogram(const float* values, const int64_t* row_start, int64_t n_instances,
                             int64_t n_rows, float lo, float hi, int bins, float width,
                             float* hist, void* stream);
int loits_hip_soft_histogram_backward(const float* values, const int64_t* row_start,
                                      int64_t n_instances, int64_t n_rows, float lo, float hi,
                                      int bins, float width, const float* grad_hist,
                                      float* grad_values, void* stream);

#ifdef __cplusplus
}
#endif
#endif
