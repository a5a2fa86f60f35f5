// bindings.cpp -- pybind11 front end over the C ABI (include/loits_hip.h).
//
// `forward` / `backward` keep the signatures and return values of the
// reference bindings (cpp/bindings.cpp:24-33, 132-147; state tuple
// cpp/bindings.cpp:126-129) with one extra trailing argument, the HIP stream
// handle (None = torch's current stream).  PyTorch allocates every tensor; the
// C ABI receives raw device pointers.  BatchedPlan wraps the batched 1-D path.
#include <torch/extension.h>
#include <ATen/record_function.h>

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "loits_hip.h"

namespace py = pybind11;

namespace {

// A profiler range (at::RecordFunction) over one stage of a binding.  With
// profile_regions set, the caller's stream is drained when the range opens and
// again before it closes, so the range covers the stage's GPU work and not
// only its launches; otherwise nothing is recorded and nothing waits.
class StageRange {
 public:
  StageRange(bool on, const char* name, void* stream) {
    if (!on) return;
    stream_ = stream;
    drain();
    range_ = std::make_unique<at::RecordFunction>(at::RecordScope::USER_SCOPE);
    if (range_->isActive()) range_->before(name);
  }
  ~StageRange() {
    if (range_) drain();  // the range itself ends when range_ is destroyed
  }
  StageRange(const StageRange&) = delete;
  StageRange& operator=(const StageRange&) = delete;

 private:
  void drain() { (void)loits_hip_synchronize(stream_); }
  void* stream_ = nullptr;
  std::unique_ptr<at::RecordFunction> range_;
};

void check(int status, const char* what) {
  if (status != LOITS_HIP_OK)
    throw std::runtime_error(std::string(what) + ": " + loits_hip_last_error());
}

void require_device(const at::Tensor& t, const char* name) {
  if (!t.is_cuda()) throw std::invalid_argument(std::string(name) + " must be a GPU tensor");
  if (!t.is_contiguous()) throw std::invalid_argument(std::string(name) + " must be contiguous");
}

void* stream_of(const py::object& stream, const at::Tensor& like) {
  if (!stream.is_none()) return reinterpret_cast<void*>(stream.cast<uintptr_t>());
  py::object cuda = py::module_::import("torch").attr("cuda");
  py::object dev = py::cast(like.device().index());
  return reinterpret_cast<void*>(cuda.attr("current_stream")(dev).attr("cuda_stream").cast<uintptr_t>());
}

// Makes an explicitly passed HIP stream torch's current stream for the
// lifetime of the object, so that the temporaries a binding allocates are
// ordered on that stream by the caching allocator and the host reads of
// sizes (.cpu()) wait for the kernels that write them.  No-op for None.
class StreamScope {
 public:
  StreamScope(const py::object& stream, const at::Tensor& like) {
    if (stream.is_none()) return;
    py::object cuda = py::module_::import("torch").attr("cuda");
    py::object ext = cuda.attr("ExternalStream")(stream.cast<uintptr_t>(),
                                                 py::arg("device") = like.device().index());
    ctx_ = cuda.attr("stream")(ext);
    ctx_.attr("__enter__")();
  }
  ~StreamScope() {
    if (ctx_) ctx_.attr("__exit__")(py::none(), py::none(), py::none());
  }
  StreamScope(const StreamScope&) = delete;
  StreamScope& operator=(const StreamScope&) = delete;

 private:
  py::object ctx_;
};

// Throws std::invalid_argument (ValueError in Python) unless `t` is a
// contiguous GPU tensor of dtype `dt` with `numel` elements (numel < 0: any).
void require_tensor(const at::Tensor& t, const char* name, at::ScalarType dt, int64_t numel) {
  require_device(t, name);
  if (t.scalar_type() != dt)
    throw std::invalid_argument(std::string(name) + " has dtype " +
                                std::string(c10::toString(t.scalar_type())) + ", expected " +
                                std::string(c10::toString(dt)));
  if (numel >= 0 && t.numel() != numel)
    throw std::invalid_argument(std::string(name) + " has " + std::to_string(t.numel()) +
                                " elements, expected " + std::to_string(numel));
}

void require_min(const at::Tensor& t, const char* name, int64_t numel) {
  if (t.numel() < numel)
    throw std::invalid_argument(std::string(name) + " has " + std::to_string(t.numel()) +
                                " elements, fewer than the " + std::to_string(numel) +
                                " required");
}

int dtype_code(const at::Tensor& t, const char* name) {
  if (t.scalar_type() == at::kDouble) return LOITS_HIP_F64;
  if (t.scalar_type() == at::kFloat) return LOITS_HIP_F32;
  throw std::invalid_argument(std::string(name) + " must be float64 or float32");
}

loits_hip_shape shape_of(const at::Tensor& x_bins, const at::Tensor& q_bins) {
  if (x_bins.dim() != 3 || q_bins.dim() != 3)
    throw std::invalid_argument("x_bins and q_bins must be [B, n, k]");
  return loits_hip_shape{x_bins.size(0), x_bins.size(1), q_bins.size(1),
                         x_bins.size(2), q_bins.size(2), x_bins.size(1) * q_bins.size(1)};
}

// cpp/bindings.cpp:24-130
std::vector<at::Tensor> forward(at::Tensor x_bins, at::Tensor xsec_x, at::Tensor q_bins,
                                at::Tensor xsec_q, at::Tensor weights, at::Tensor acceptance,
                                int64_t n_events, uint64_t seed, uint64_t sequence,
                                bool profile_regions, py::object stream_obj) {
  for (auto* p : {&x_bins, &xsec_x, &q_bins, &xsec_q, &weights, &acceptance})
    require_device(*p, "theory output");
  const int dt = dtype_code(xsec_x, "xsec_x");
  for (auto* p : {&x_bins, &q_bins, &xsec_q, &weights})
    if (p->scalar_type() != xsec_x.scalar_type())
      throw std::invalid_argument("all floating theory outputs must share xsec_x's dtype");
  if (acceptance.scalar_type() != at::kBool)
    throw std::invalid_argument("acceptance must be a bool tensor");
  const loits_hip_shape s = shape_of(x_bins, q_bins);
  if (xsec_x.numel() != s.batch * s.ny * s.nx * s.kx || xsec_q.numel() != s.batch * s.nx * s.ny * s.kq ||
      weights.numel() != s.batch * s.cells || acceptance.numel() != s.batch * s.cells)
    throw std::invalid_argument("theory output shapes do not match x_bins / q_bins");
  for (auto* p : {&x_bins, &q_bins, &xsec_q, &weights, &acceptance})
    if (p->device() != xsec_x.device())
      throw std::invalid_argument("all theory outputs must be on one GPU");
  StreamScope scope(stream_obj, xsec_x);
  void* stream = stream_of(stream_obj, xsec_x);
  StageRange total(profile_regions, "loits::forward", stream);
  auto i64 = xsec_x.options().dtype(at::kLong);
  auto f64 = xsec_x.options().dtype(at::kDouble);

  at::Tensor counts = at::empty({s.batch, s.cells}, i64);
  int64_t nmax = 0, active = 0;
  {
    StageRange region(profile_regions, "loits::forward::allocation", stream);
    at::Tensor na = at::empty({2}, i64);
    check(loits_hip_allocate(weights.data_ptr(), dt, s.batch * s.cells, n_events,
                             counts.data_ptr<int64_t>(), na.data_ptr<int64_t>(), stream),
          "loits_hip_allocate");
    at::Tensor host = na.cpu();  // the reference reads nmax/active on the host here too
    nmax = host[0].item<int64_t>();
    active = host[1].item<int64_t>();
  }
  const int64_t total_slots = s.batch * s.cells * nmax;
  at::Tensor norm_x = at::empty({s.batch, s.ny, s.nx}, f64);
  at::Tensor norm_q = at::empty({s.batch, s.nx, s.ny}, f64);
  at::Tensor cdf_x = at::empty(xsec_x.sizes(), f64);
  at::Tensor cdf_q = at::empty(xsec_q.sizes(), f64);
  at::Tensor u_x = at::empty({s.batch, s.cells, nmax}, xsec_x.options().dtype(at::kFloat));
  at::Tensor u_q = at::empty({s.batch, s.cells, nmax}, xsec_x.options().dtype(at::kFloat));
  at::Tensor interval_x = at::empty({total_slots}, xsec_x.options().dtype(at::kShort));
  at::Tensor interval_q = at::empty({total_slots}, xsec_x.options().dtype(at::kShort));
  at::Tensor events_storage = at::empty({active, 2}, xsec_x.options());
  at::Tensor packed_storage = at::empty({active}, i64);
  at::Tensor row_offsets = at::empty({s.batch * s.cells + 1}, i64);
  const int64_t ws_bytes = loits_hip_forward_workspace_size(&s, nmax);
  if (ws_bytes < 0) throw std::invalid_argument(loits_hip_last_error());
  at::Tensor workspace = at::empty({ws_bytes}, xsec_x.options().dtype(at::kByte));
  at::Tensor valid = at::empty({2}, i64);
  at::Tensor acc_u8 = acceptance.view(at::kByte);
  {
    // The HIP forward fuses the reference's stages into three kernels:
    // curves (rho + CDF, both axes), sampling (both Philox streams,
    // interpolation, compaction) and row offsets, all launched by the one
    // call below.  The reference's stage names are kept: the fused stages are
    // empty ranges here and their GPU time falls in stream_compaction.
    StageRange r1(profile_regions, "loits::forward::rho_x", stream);
    StageRange r2(profile_regions, "loits::forward::rho_q2", stream);
    StageRange r3(profile_regions, "loits::forward::cdf_x", stream);
    StageRange r4(profile_regions, "loits::forward::cdf_q2", stream);
  }
  {
    StageRange region(profile_regions, "loits::forward::stream_compaction", stream);
    {
      StageRange a(profile_regions, "loits::forward::random_x", stream);
      StageRange b(profile_regions, "loits::forward::random_q2", stream);
      StageRange c(profile_regions, "loits::forward::interpolation_x", stream);
      StageRange d(profile_regions, "loits::forward::interpolation_q2", stream);
    }
    loits_hip_forward_args a{};
    a.shape = s;
    a.dtype = dt;
    a.nmax = nmax;
    a.active = active;
    a.seed = seed;
    a.sequence = sequence;
    a.x_bins = x_bins.data_ptr();
    a.xsec_x = xsec_x.data_ptr();
    a.q_bins = q_bins.data_ptr();
    a.xsec_q = xsec_q.data_ptr();
    a.acceptance = acc_u8.data_ptr<uint8_t>();
    a.counts = counts.data_ptr<int64_t>();
    a.norm_x = norm_x.data_ptr<double>();
    a.norm_q = norm_q.data_ptr<double>();
    a.cdf_x = cdf_x.data_ptr<double>();
    a.cdf_q = cdf_q.data_ptr<double>();
    a.u_x = u_x.data_ptr<float>();
    a.u_q = u_q.data_ptr<float>();
    a.interval_x = interval_x.data_ptr<int16_t>();
    a.interval_q = interval_q.data_ptr<int16_t>();
    a.events = events_storage.data_ptr();
    a.packed = packed_storage.data_ptr<int64_t>();
    a.row_offsets = row_offsets.data_ptr<int64_t>();
    a.workspace = workspace.data_ptr();
    a.valid = valid.data_ptr<int64_t>();
    check(loits_hip_forward(&a, stream), "loits_hip_forward");
  }
  at::Tensor host_valid = valid.cpu();
  const int64_t flags = host_valid[1].item<int64_t>();
  if (flags & 4)
    throw std::runtime_error("loits_hip_forward: a slot of a drop-free curve was dropped");
  if (flags != 0)
    throw std::runtime_error("loits_hip_forward: compaction look-back did not complete");
  const int64_t nvalid = host_valid[0].item<int64_t>();
  StageRange region(profile_regions, "loits::forward::state_pack", stream);
  return {events_storage.narrow(0, 0, nvalid), norm_x, norm_q, cdf_x, cdf_q, u_x, u_q,
          interval_x, interval_q, packed_storage.narrow(0, 0, nvalid), row_offsets};
}

// cpp/bindings.cpp:132-208 (row_offsets accepted for the openmp-style state)
std::vector<at::Tensor> backward(at::Tensor grad_events, at::Tensor x_bins, at::Tensor xsec_x,
                                 at::Tensor q_bins, at::Tensor xsec_q, at::Tensor acceptance,
                                 at::Tensor norm_x, at::Tensor norm_q, at::Tensor cdf_x,
                                 at::Tensor cdf_q, at::Tensor u_x, at::Tensor u_q,
                                 at::Tensor interval_x, at::Tensor interval_q, at::Tensor packed,
                                 py::args rest) {
  // rest = ([row_offsets,] profile_regions [, stream])
  std::vector<py::object> tail;
  for (auto item : rest) tail.push_back(py::reinterpret_borrow<py::object>(item));
  at::Tensor row_offsets;
  if (!tail.empty() && py::hasattr(tail.front(), "dtype")) {
    row_offsets = tail.front().cast<at::Tensor>();
    tail.erase(tail.begin());
  }
  const bool profile_regions = !tail.empty() && tail[0].cast<bool>();
  py::object stream_obj = tail.size() > 1 ? tail[1] : py::none();
  grad_events = grad_events.contiguous();
  require_device(grad_events, "grad_events");
  const int dt = dtype_code(xsec_x, "xsec_x");
  if (grad_events.scalar_type() != xsec_x.scalar_type()) grad_events = grad_events.to(xsec_x.scalar_type());
  const loits_hip_shape s = shape_of(x_bins, q_bins);
  const int64_t gcells = s.batch * s.cells;
  const at::ScalarType rt = xsec_x.scalar_type();
  require_tensor(x_bins, "x_bins", rt, -1);
  require_tensor(q_bins, "q_bins", rt, -1);
  require_tensor(xsec_x, "xsec_x", rt, s.batch * s.ny * s.nx * s.kx);
  require_tensor(xsec_q, "xsec_q", rt, s.batch * s.nx * s.ny * s.kq);
  if (acceptance.scalar_type() != at::kBool || acceptance.numel() != gcells)
    throw std::invalid_argument("acceptance must be a bool tensor [B, nx, ny]");
  require_tensor(norm_x, "norm_x", at::kDouble, gcells);
  require_tensor(norm_q, "norm_q", at::kDouble, gcells);
  require_tensor(cdf_x, "cdf_x", at::kDouble, xsec_x.numel());
  require_tensor(cdf_q, "cdf_q", at::kDouble, xsec_q.numel());
  if (u_x.dim() != 3 || u_x.size(0) != s.batch || u_x.size(1) != s.cells)
    throw std::invalid_argument("u_x must be [B, cells, nmax]");
  const int64_t slots = u_x.numel();
  require_tensor(u_x, "u_x", at::kFloat, slots);
  require_tensor(u_q, "u_q", at::kFloat, slots);
  require_tensor(interval_x, "interval_x", at::kShort, slots);
  require_tensor(interval_q, "interval_q", at::kShort, slots);
  require_tensor(packed, "packed", at::kLong, -1);
  if (grad_events.dim() != 2 || grad_events.size(1) != 2 || grad_events.size(0) < packed.size(0))
    throw std::invalid_argument("grad_events must be [n, 2] with a row for every packed event");
  for (const at::Tensor* p : {&grad_events, &x_bins, &q_bins, &xsec_q, &norm_x, &norm_q, &cdf_x,
                              &cdf_q, &u_x, &u_q, &interval_x, &interval_q, &packed})
    if (p->device() != xsec_x.device())
      throw std::invalid_argument("backward: every tensor must be on xsec_x's GPU");
  StreamScope scope(stream_obj, xsec_x);
  void* stream = stream_of(stream_obj, xsec_x);
  StageRange total(profile_regions, "loits::backward", stream);
  at::Tensor grad_x = at::empty(xsec_x.sizes(), xsec_x.options());
  at::Tensor grad_q = at::empty(xsec_q.sizes(), xsec_q.options());
  const int64_t ws_bytes = loits_hip_backward_workspace_size(&s, u_x.size(2));
  if (ws_bytes < 0) throw std::invalid_argument(loits_hip_last_error());
  at::Tensor workspace = at::empty({ws_bytes}, xsec_x.options().dtype(at::kByte));
  at::Tensor acc_u8 = acceptance.contiguous().view(at::kByte);
  {
    StageRange a(profile_regions, "loits::backward::interpolation_x", stream);
    StageRange b(profile_regions, "loits::backward::interpolation_q2", stream);
  }
  {
    StageRange r1(profile_regions, "loits::backward::cdf_x", stream);
    StageRange r2(profile_regions, "loits::backward::cdf_q2", stream);
    StageRange r3(profile_regions, "loits::backward::rho_x", stream);
    StageRange r4(profile_regions, "loits::backward::rho_q2", stream);
    loits_hip_backward_args a{};
    a.shape = s;
    a.dtype = dt;
    a.nmax = u_x.size(2);
    a.nvalid = packed.size(0);
    a.grad_events = grad_events.data_ptr();
    a.x_bins = x_bins.data_ptr();
    a.xsec_x = xsec_x.data_ptr();
    a.q_bins = q_bins.data_ptr();
    a.xsec_q = xsec_q.data_ptr();
    a.acceptance = acc_u8.data_ptr<uint8_t>();
    a.norm_x = norm_x.data_ptr<double>();
    a.norm_q = norm_q.data_ptr<double>();
    a.cdf_x = cdf_x.data_ptr<double>();
    a.cdf_q = cdf_q.data_ptr<double>();
    a.u_x = u_x.data_ptr<float>();
    a.u_q = u_q.data_ptr<float>();
    a.interval_x = interval_x.data_ptr<int16_t>();
    a.interval_q = interval_q.data_ptr<int16_t>();
    a.packed = packed.data_ptr<int64_t>();
    a.grad_xsec_x = grad_x.data_ptr();
    a.grad_xsec_q = grad_q.data_ptr();
    a.workspace = workspace.data_ptr();
    // the forward's row offsets give each cell's rows without a search
    if (row_offsets.defined() && row_offsets.is_cuda() &&
        row_offsets.scalar_type() == at::kLong && row_offsets.is_contiguous() &&
        row_offsets.numel() == s.batch * s.cells + 1 && row_offsets.device() == xsec_x.device())
      a.row_offsets = row_offsets.data_ptr<int64_t>();
    check(loits_hip_backward(&a, stream), "loits_hip_backward");
  }
  StageRange region(profile_regions, "loits::backward::state_pack", stream);
  return {grad_x, grad_q};
}

// Sync-free reference forward/backward into caller-owned capacity buffers
// (loits_hip_forward_bounded / loits_hip_backward_bounded): nothing is read on
// the host, so a caller can capture the calls in a HIP graph.  `na` ([2],
// int64) receives nmax / active, `valid` ([2]) the kept count and error bits.
void forward_bounded(at::Tensor x_bins, at::Tensor xsec_x, at::Tensor q_bins, at::Tensor xsec_q,
                     at::Tensor weights, at::Tensor acceptance, int64_t n_events, uint64_t seed,
                     uint64_t sequence, at::Tensor counts, at::Tensor na, at::Tensor norm_x,
                     at::Tensor norm_q, at::Tensor cdf_x, at::Tensor cdf_q, at::Tensor u_x,
                     at::Tensor u_q, at::Tensor interval_x, at::Tensor interval_q,
                     at::Tensor events, at::Tensor packed, at::Tensor row_offsets,
                     at::Tensor workspace, at::Tensor valid, py::object stream_obj) {
  for (auto* p : {&x_bins, &xsec_x, &q_bins, &xsec_q, &weights, &acceptance, &counts, &na,
                  &norm_x, &norm_q, &cdf_x, &cdf_q, &u_x, &u_q, &interval_x, &interval_q,
                  &events, &packed, &row_offsets, &workspace, &valid})
    require_device(*p, "bounded forward tensor");
  const int dt = dtype_code(xsec_x, "xsec_x");
  const loits_hip_shape s = shape_of(x_bins, q_bins);
  const int64_t gcells = s.batch * s.cells;
  const at::ScalarType rt = xsec_x.scalar_type();
  require_tensor(x_bins, "x_bins", rt, -1);
  require_tensor(q_bins, "q_bins", rt, -1);
  require_tensor(xsec_x, "xsec_x", rt, s.batch * s.ny * s.nx * s.kx);
  require_tensor(xsec_q, "xsec_q", rt, s.batch * s.nx * s.ny * s.kq);
  require_tensor(weights, "weights", rt, gcells);
  require_tensor(events, "events", rt, -1);
  if (acceptance.scalar_type() != at::kBool || acceptance.numel() != gcells)
    throw std::invalid_argument("acceptance must be a bool tensor [B, nx, ny]");
  require_tensor(norm_x, "norm_x", at::kDouble, gcells);
  require_tensor(norm_q, "norm_q", at::kDouble, gcells);
  require_tensor(cdf_x, "cdf_x", at::kDouble, xsec_x.numel());
  require_tensor(cdf_q, "cdf_q", at::kDouble, xsec_q.numel());
  require_tensor(counts, "counts", at::kLong, -1);
  require_tensor(na, "na", at::kLong, -1);
  require_tensor(u_x, "u_x", at::kFloat, -1);
  require_tensor(u_q, "u_q", at::kFloat, -1);
  require_tensor(interval_x, "interval_x", at::kShort, -1);
  require_tensor(interval_q, "interval_q", at::kShort, -1);
  require_tensor(packed, "packed", at::kLong, -1);
  require_tensor(row_offsets, "row_offsets", at::kLong, -1);
  require_tensor(workspace, "workspace", at::kByte, -1);
  require_tensor(valid, "valid", at::kLong, -1);
  for (const at::Tensor* p : {&x_bins, &q_bins, &xsec_q, &weights, &acceptance, &counts, &na,
                              &norm_x, &norm_q, &cdf_x, &cdf_q, &u_x, &u_q, &interval_x,
                              &interval_q, &events, &packed, &row_offsets, &workspace, &valid})
    if (p->device() != xsec_x.device())
      throw std::invalid_argument("bounded forward: every tensor must be on xsec_x's GPU");
  const int64_t nmax_cap = gcells > 0 ? u_x.numel() / gcells : 0;
  const int64_t active_cap = packed.numel();
  if (u_q.numel() < gcells * nmax_cap || interval_x.numel() < gcells * nmax_cap ||
      interval_q.numel() < gcells * nmax_cap || events.numel() < 2 * active_cap ||
      row_offsets.numel() < gcells + 1 || na.numel() < 2 || valid.numel() < 2 ||
      counts.numel() < gcells)
    throw std::invalid_argument("bounded forward: a buffer is smaller than its capacity");
  const int64_t ws = loits_hip_forward_workspace_size(&s, nmax_cap);
  if (ws < 0 || workspace.numel() < ws)
    throw std::invalid_argument("bounded forward: workspace too small");
  void* stream = stream_of(stream_obj, xsec_x);
  check(loits_hip_allocate(weights.data_ptr(), dt, gcells, n_events, counts.data_ptr<int64_t>(),
                           na.data_ptr<int64_t>(), stream),
        "loits_hip_allocate");
  at::Tensor acc_u8 = acceptance.view(at::kByte);
  loits_hip_forward_args a{};
  a.shape = s;
  a.dtype = dt;
  a.nmax = nmax_cap;
  a.active = active_cap;
  a.seed = seed;
  a.sequence = sequence;
  a.x_bins = x_bins.data_ptr();
  a.xsec_x = xsec_x.data_ptr();
  a.q_bins = q_bins.data_ptr();
  a.xsec_q = xsec_q.data_ptr();
  a.acceptance = acc_u8.data_ptr<uint8_t>();
  a.counts = counts.data_ptr<int64_t>();
  a.norm_x = norm_x.data_ptr<double>();
  a.norm_q = norm_q.data_ptr<double>();
  a.cdf_x = cdf_x.data_ptr<double>();
  a.cdf_q = cdf_q.data_ptr<double>();
  a.u_x = u_x.data_ptr<float>();
  a.u_q = u_q.data_ptr<float>();
  a.interval_x = interval_x.data_ptr<int16_t>();
  a.interval_q = interval_q.data_ptr<int16_t>();
  a.events = events.data_ptr();
  a.packed = packed.data_ptr<int64_t>();
  a.row_offsets = row_offsets.data_ptr<int64_t>();
  a.workspace = workspace.data_ptr();
  a.valid = valid.data_ptr<int64_t>();
  check(loits_hip_forward_bounded(&a, na.data_ptr<int64_t>(), stream), "loits_hip_forward_bounded");
}

void backward_bounded(at::Tensor grad_events, at::Tensor x_bins, at::Tensor xsec_x,
                      at::Tensor q_bins, at::Tensor xsec_q, at::Tensor acceptance,
                      at::Tensor norm_x, at::Tensor norm_q, at::Tensor cdf_x, at::Tensor cdf_q,
                      at::Tensor u_x, at::Tensor u_q, at::Tensor interval_x,
                      at::Tensor interval_q, at::Tensor packed, at::Tensor na, at::Tensor valid,
                      at::Tensor grad_x, at::Tensor grad_q, at::Tensor workspace,
                      py::object stream_obj) {
  for (auto* p : {&grad_events, &x_bins, &xsec_x, &q_bins, &xsec_q, &acceptance, &norm_x,
                  &norm_q, &cdf_x, &cdf_q, &u_x, &u_q, &interval_x, &interval_q, &packed, &na,
                  &valid, &grad_x, &grad_q, &workspace})
    require_device(*p, "bounded backward tensor");
  const int dt = dtype_code(xsec_x, "xsec_x");
  if (grad_events.scalar_type() != xsec_x.scalar_type() ||
      grad_x.scalar_type() != xsec_x.scalar_type() || grad_q.scalar_type() != xsec_x.scalar_type())
    throw std::invalid_argument("bounded backward: gradients must have xsec_x's dtype");
  const loits_hip_shape s = shape_of(x_bins, q_bins);
  const int64_t gcells = s.batch * s.cells;
  const at::ScalarType rt = xsec_x.scalar_type();
  require_tensor(x_bins, "x_bins", rt, -1);
  require_tensor(q_bins, "q_bins", rt, -1);
  require_tensor(xsec_x, "xsec_x", rt, s.batch * s.ny * s.nx * s.kx);
  require_tensor(xsec_q, "xsec_q", rt, s.batch * s.nx * s.ny * s.kq);
  if (acceptance.scalar_type() != at::kBool || acceptance.numel() != gcells)
    throw std::invalid_argument("acceptance must be a bool tensor [B, nx, ny]");
  require_tensor(norm_x, "norm_x", at::kDouble, gcells);
  require_tensor(norm_q, "norm_q", at::kDouble, gcells);
  require_tensor(cdf_x, "cdf_x", at::kDouble, xsec_x.numel());
  require_tensor(cdf_q, "cdf_q", at::kDouble, xsec_q.numel());
  const int64_t nmax_cap = gcells > 0 ? u_x.numel() / gcells : 0;
  require_tensor(u_x, "u_x", at::kFloat, -1);
  require_tensor(u_q, "u_q", at::kFloat, -1);
  require_tensor(interval_x, "interval_x", at::kShort, -1);
  require_tensor(interval_q, "interval_q", at::kShort, -1);
  require_min(u_q, "u_q", gcells * nmax_cap);
  require_min(interval_x, "interval_x", gcells * nmax_cap);
  require_min(interval_q, "interval_q", gcells * nmax_cap);
  require_tensor(packed, "packed", at::kLong, -1);
  require_tensor(na, "na", at::kLong, -1);
  require_tensor(valid, "valid", at::kLong, -1);
  require_min(na, "na", 2);
  require_min(valid, "valid", 2);
  require_tensor(workspace, "workspace", at::kByte, -1);
  for (const at::Tensor* p : {&grad_events, &x_bins, &q_bins, &xsec_q, &acceptance, &norm_x,
                              &norm_q, &cdf_x, &cdf_q, &u_x, &u_q, &interval_x, &interval_q,
                              &packed, &na, &valid, &grad_x, &grad_q, &workspace})
    if (p->device() != xsec_x.device())
      throw std::invalid_argument("bounded backward: every tensor must be on xsec_x's GPU");
  if (grad_events.numel() < 2 * packed.numel() || grad_x.numel() != xsec_x.numel() ||
      grad_q.numel() != xsec_q.numel())
    throw std::invalid_argument("bounded backward: gradient buffers have the wrong size");
  const int64_t ws = loits_hip_backward_workspace_size(&s, 0);
  if (ws < 0 || workspace.numel() < ws)
    throw std::invalid_argument("bounded backward: workspace too small");
  at::Tensor acc_u8 = acceptance.view(at::kByte);
  loits_hip_backward_args a{};
  a.shape = s;
  a.dtype = dt;
  a.nmax = gcells > 0 ? u_x.numel() / gcells : 0;
  a.nvalid = packed.numel();
  a.grad_events = grad_events.data_ptr();
  a.x_bins = x_bins.data_ptr();
  a.xsec_x = xsec_x.data_ptr();
  a.q_bins = q_bins.data_ptr();
  a.xsec_q = xsec_q.data_ptr();
  a.acceptance = acc_u8.data_ptr<uint8_t>();
  a.norm_x = norm_x.data_ptr<double>();
  a.norm_q = norm_q.data_ptr<double>();
  a.cdf_x = cdf_x.data_ptr<double>();
  a.cdf_q = cdf_q.data_ptr<double>();
  a.u_x = u_x.data_ptr<float>();
  a.u_q = u_q.data_ptr<float>();
  a.interval_x = interval_x.data_ptr<int16_t>();
  a.interval_q = interval_q.data_ptr<int16_t>();
  a.packed = packed.data_ptr<int64_t>();
  a.grad_xsec_x = grad_x.data_ptr();
  a.grad_xsec_q = grad_q.data_ptr();
  a.workspace = workspace.data_ptr();
  check(loits_hip_backward_bounded(&a, na.data_ptr<int64_t>(), valid.data_ptr<int64_t>(),
                                   stream_of(stream_obj, xsec_x)),
        "loits_hip_backward_bounded");
}

at::Tensor philox_fill(int64_t count, uint64_t seed, uint64_t stream_id, int64_t device,
                       py::object stream_obj) {
  at::Tensor probe = at::empty({0}, at::TensorOptions().dtype(at::kFloat).device(at::kCUDA, device));
  StreamScope scope(stream_obj, probe);
  at::Tensor out = at::empty({count}, at::TensorOptions().dtype(at::kFloat).device(at::kCUDA, device));
  check(loits_hip_philox_fill(out.data_ptr<float>(), count, seed, stream_id,
                              stream_of(stream_obj, out)),
        "loits_hip_philox_fill");
  return out;
}

at::Tensor philox_words(int64_t first_block, int64_t blocks, uint64_t stream_id, uint64_t seed,
                        int64_t device, py::object stream_obj) {
  at::Tensor probe = at::empty({0}, at::TensorOptions().dtype(at::kInt).device(at::kCUDA, device));
  StreamScope scope(stream_obj, probe);
  at::Tensor out = at::empty({blocks * 4}, at::TensorOptions().dtype(at::kInt).device(at::kCUDA, device));
  check(loits_hip_philox_words(reinterpret_cast<uint32_t*>(out.data_ptr<int32_t>()), first_block,
                               blocks, stream_id, seed, stream_of(stream_obj, out)),
        "loits_hip_philox_words");
  return out;
}

// Soft-binning histograms (training harness).
at::Tensor soft_histogram(at::Tensor values, at::Tensor row_start, double lo, double hi,
                          int64_t bins, double width, py::object stream_obj) {
  require_device(values, "values");
  require_device(row_start, "row_start");
  if (values.scalar_type() != at::kFloat || row_start.scalar_type() != at::kLong)
    throw std::invalid_argument("values must be float32 and row_start int64");
  const int64_t n = row_start.numel() - 1;
  StreamScope scope(stream_obj, values);
  at::Tensor hist = at::empty({n > 0 ? n : 0, bins}, values.options());
  check(loits_hip_soft_histogram(values.data_ptr<float>(), row_start.data_ptr<int64_t>(), n,
                                 values.numel(), float(lo), float(hi), int(bins), float(width),
                                 hist.data_ptr<float>(), stream_of(stream_obj, values)),
        "loits_hip_soft_histogram");
  return hist;
}

at::Tensor soft_histogram_backward(at::Tensor values, at::Tensor row_start, double lo, double hi,
                                   int64_t bins, double width, at::Tensor grad_hist,
                                   py::object stream_obj) {
  require_device(values, "values");
  require_device(row_start, "row_start");
  require_device(grad_hist, "grad_hist");
  if (values.scalar_type() != at::kFloat || row_start.scalar_type() != at::kLong ||
      grad_hist.scalar_type() != at::kFloat)
    throw std::invalid_argument("values and grad_hist must be float32, row_start int64");
  const int64_t n = row_start.numel() - 1;
  if (grad_hist.numel() != n * bins) throw std::invalid_argument("grad_hist must be [n, bins]");
  StreamScope scope(stream_obj, values);
  at::Tensor grad = at::empty_like(values);
  check(loits_hip_soft_histogram_backward(values.data_ptr<float>(), row_start.data_ptr<int64_t>(),
                                          n, values.numel(), float(lo), float(hi), int(bins),
                                          float(width), grad_hist.data_ptr<float>(),
                                          grad.data_ptr<float>(), stream_of(stream_obj, values)),
        "loits_hip_soft_histogram_backward");
  return grad;
}

// ---------------------------------------------------------------------------
// Batched 1-D path.
class BatchedPlan {
 public:
  // instances: CPU int64 [n, 8] = node_off, n_nodes, events, out_off,
  // counter_base, stream, accept, 0 (counter/stream as two's-complement int64)
  // bucket_chunk_log2 / bucket_guide_shift: empty, or 16 per-bucket overrides
  // (loits_hip_tuning; 0 keeps the plan-wide value)
  BatchedPlan(at::Tensor instances, int64_t dtype, int64_t threads, int64_t guide_shift,
              int64_t chunk_events, int64_t scan_mode, int64_t persistent, int64_t table_format,
              std::vector<int64_t> bucket_chunk_log2, std::vector<int64_t> bucket_guide_shift) {
    instances = instances.contiguous();
    if (instances.device().is_cuda() || instances.scalar_type() != at::kLong ||
        instances.dim() != 2 || instances.size(1) != 8)
      throw std::invalid_argument("instances must be a CPU int64 tensor [n, 8]");
    const int64_t n = instances.size(0);
    inst_.resize(size_t(n));
    const int64_t* src = instances.data_ptr<int64_t>();
    for (int64_t i = 0; i < n; ++i) {
      loits_hip_instance& d = inst_[size_t(i)];
      d.node_off = src[8 * i + 0];
      d.n_nodes = src[8 * i + 1];
      d.events = src[8 * i + 2];
      d.out_off = src[8 * i + 3];
      d.counter_base = uint64_t(src[8 * i + 4]);
      d.stream = uint64_t(src[8 * i + 5]);
      d.accept = int32_t(src[8 * i + 6]);
      d.reserved = 0;
    }
    tuning_.threads = int32_t(threads);
    tuning_.guide_shift = int32_t(guide_shift);
    tuning_.chunk_events = chunk_events;
    tuning_.scan_mode = int32_t(scan_mode);
    tuning_.persistent = int32_t(persistent);
    tuning_.table_format = int32_t(table_format);
    tuning_.reserved = 0;
    for (auto* v : {&bucket_chunk_log2, &bucket_guide_shift})
      if (!v->empty() && v->size() != 16)
        throw std::invalid_argument("bucket overrides must have 16 entries (or none)");
    for (int b = 0; b < 16; ++b) {
      tuning_.bucket_chunk_log2[b] =
          bucket_chunk_log2.empty() ? 0 : int32_t(bucket_chunk_log2[size_t(b)]);
      tuning_.bucket_guide_shift[b] =
          bucket_guide_shift.empty() ? 0 : int32_t(bucket_guide_shift[size_t(b)]);
    }
    check(loits_hip_batched_layout(inst_.data(), n, int(dtype), &tuning_, &layout_),
          "loits_hip_batched_layout");
    desc_host_ = at::empty({layout_.desc_bytes}, at::TensorOptions().dtype(at::kByte));
    check(loits_hip_batched_encode(inst_.data(), n, &tuning_, &layout_, desc_host_.data_ptr()),
          "loits_hip_batched_encode");
  }

  py::dict layout() const {
    py::dict d;
    d["n_instances"] = layout_.n_instances;
    d["n_chunks"] = layout_.n_chunks;
    d["capacity"] = layout_.capacity;
    d["total_nodes"] = layout_.total_nodes;
    d["max_nodes"] = layout_.max_nodes;
    d["desc_bytes"] = layout_.desc_bytes;
    d["table_bytes"] = layout_.table_bytes;
    d["scratch_bytes"] = layout_.scratch_bytes;
    d["lds_forward"] = layout_.lds_forward;
    d["lds_backward"] = layout_.lds_backward;
    d["lds_tables"] = layout_.lds_tables;
    d["n_segments"] = layout_.n_segments;
    d["mask_words"] = layout_.mask_words;
    d["partial_doubles"] = layout_.partial_doubles;
    d["global_instances"] = layout_.global_instances;
    d["range_instances"] = layout_.range_instances;
    d["cell_instances"] = layout_.cell_instances;
    d["grid_backward"] = layout_.grid_backward;
    d["backward_plain"] = layout_.backward_plain;
    d["dtype"] = layout_.dtype;
    d["threads"] = layout_.threads;
    d["scan_mode"] = layout_.scan_mode;
    d["persistent"] = layout_.persistent;
    d["forward_items"] = layout_.forward_items;
    d["lds_persistent_forward"] = layout_.lds_persistent_forward;
    return d;
  }

  at::Tensor descriptors() const { return desc_host_; }

  // The descriptor blob with each instance's Philox stream replaced; the
  // layout does not depend on the streams, so the blob has the same size and
  // can be copied over a captured graph's device descriptors between replays.
  at::Tensor descriptors_for_streams(at::Tensor streams) const {
    streams = streams.contiguous();
    if (streams.device().is_cuda() || streams.scalar_type() != at::kLong || streams.dim() != 1 ||
        streams.size(0) != int64_t(inst_.size()))
      throw std::invalid_argument("streams must be a CPU int64 tensor [n_instances]");
    std::vector<loits_hip_instance> inst = inst_;
    const int64_t* src = streams.data_ptr<int64_t>();
    for (size_t i = 0; i < inst.size(); ++i) inst[i].stream = uint64_t(src[i]);
    at::Tensor out = at::empty({layout_.desc_bytes}, at::TensorOptions().dtype(at::kByte));
    check(loits_hip_batched_encode(inst.data(), int64_t(inst.size()), &tuning_, &layout_,
                                   out.data_ptr()),
          "loits_hip_batched_encode");
    return out;
  }

  // All tensors are device tensors the caller allocated with the layout's sizes.
  void forward(at::Tensor desc, at::Tensor x, at::Tensor density, uint64_t seed, bool use_window,
               double lo, double hi, at::Tensor tables, at::Tensor scratch, at::Tensor events,
               at::Tensor kept, at::Tensor total, py::object norm, py::object cdf, py::object bins,
               py::object stream_obj) {
    loits_hip_batched_args a = args(desc, x, density, seed, use_window, lo, hi, tables, scratch,
                                    events, kept, total, norm, cdf, bins);
    check(loits_hip_batched_forward(&a, stream_of(stream_obj, x)), "loits_hip_batched_forward");
  }

  void pack(at::Tensor desc, at::Tensor x, at::Tensor density, at::Tensor tables,
            at::Tensor scratch, at::Tensor events, at::Tensor kept, at::Tensor total,
            at::Tensor packed, py::object bins, py::object packed_bins, py::object stream_obj) {
    loits_hip_batched_args a = args(desc, x, density, 0, false, 0.0, 0.0, tables, scratch, events,
                                    kept, total, py::none(), py::none(), bins);
    int32_t* pb = packed_bins.is_none() ? nullptr : packed_bins.cast<at::Tensor>().data_ptr<int32_t>();
    check(loits_hip_batched_pack(&a, packed.data_ptr(), pb, stream_of(stream_obj, x)),
          "loits_hip_batched_pack");
  }

  void backward(at::Tensor desc, at::Tensor x, at::Tensor density, uint64_t seed, bool use_window,
                double lo, double hi, at::Tensor tables, at::Tensor scratch, at::Tensor events,
                at::Tensor kept, at::Tensor total, at::Tensor grad, at::Tensor grad_density,
                bool recompute_tables, py::object stream_obj) {
    loits_hip_batched_args a = args(desc, x, density, seed, use_window, lo, hi, tables, scratch,
                                    events, kept, total, py::none(), py::none(), py::none());
    check(loits_hip_batched_backward(&a, grad.data_ptr(), grad_density.data_ptr(),
                                     recompute_tables ? 1 : 0, stream_of(stream_obj, x)),
          "loits_hip_batched_backward");
  }

 private:
  loits_hip_batched_args args(at::Tensor& desc, at::Tensor& x, at::Tensor& density, uint64_t seed,
                              bool use_window, double lo, double hi, at::Tensor& tables,
                              at::Tensor& scratch, at::Tensor& events, at::Tensor& kept,
                              at::Tensor& total, const py::object& norm, const py::object& cdf,
                              const py::object& bins) {
    for (auto* t : {&desc, &x, &density, &tables, &scratch, &events, &kept, &total})
      require_device(*t, "batched tensor");
    const at::ScalarType want = layout_.dtype == LOITS_HIP_F32 ? at::kFloat : at::kDouble;
    if (x.scalar_type() != want || density.scalar_type() != want || events.scalar_type() != want)
      throw std::invalid_argument("x, density and events must have the plan's dtype");
    if (x.numel() < layout_.total_nodes || density.numel() < layout_.total_nodes ||
        events.numel() < layout_.capacity || tables.numel() < layout_.table_bytes ||
        scratch.numel() < layout_.scratch_bytes || desc.numel() < layout_.desc_bytes ||
        kept.numel() < layout_.n_instances || total.numel() < 1)
      throw std::invalid_argument("a batched tensor is smaller than the plan requires");
    loits_hip_batched_args a{};
    a.layout = &layout_;
    a.desc = desc.data_ptr();
    a.x = x.data_ptr();
    a.density = density.data_ptr();
    a.seed = seed;
    a.use_window = use_window ? 1 : 0;
    a.window_lo = lo;
    a.window_hi = hi;
    a.tables = tables.data_ptr();
    a.scratch = scratch.data_ptr();
    a.events = events.data_ptr();
    a.kept = kept.data_ptr<int64_t>();
    a.total = total.data_ptr<int64_t>();
    a.norm = norm.is_none() ? nullptr : norm.cast<at::Tensor>().data_ptr<double>();
    a.cdf = cdf.is_none() ? nullptr : cdf.cast<at::Tensor>().data_ptr<double>();
    a.bins = bins.is_none() ? nullptr : bins.cast<at::Tensor>().data_ptr<int32_t>();
    return a;
  }

  std::vector<loits_hip_instance> inst_;
  loits_hip_tuning tuning_{};
  loits_hip_layout layout_{};
  at::Tensor desc_host_;
};

py::tuple device_info(int device) {
  char arch[64] = {0};
  int cus = 0;
  check(loits_hip_device_info(device, arch, sizeof(arch), &cus), "loits_hip_device_info");
  return py::make_tuple(std::string(arch), cus);
}

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
  m.doc() = "MI355X LOITS sampler: pybind front end over the loits_hip C ABI";
  m.def("forward", &forward, py::arg("x_bins"), py::arg("xsec_x"), py::arg("q_bins"),
        py::arg("xsec_q"), py::arg("weights"), py::arg("acceptance"), py::arg("n_events"),
        py::arg("seed"), py::arg("sequence"), py::arg("profile_regions") = false,
        py::arg("stream") = py::none());
  m.def("backward", &backward);
  m.def("philox_fill", &philox_fill, py::arg("count"), py::arg("seed"), py::arg("stream_id"),
        py::arg("device") = 0, py::arg("stream") = py::none());
  m.def("philox_words", &philox_words, py::arg("first_block"), py::arg("blocks"),
        py::arg("stream_id"), py::arg("seed"), py::arg("device") = 0,
        py::arg("stream") = py::none());
  m.def("device_info", &device_info, py::arg("device") = 0);
  m.def("version", []() { return std::string(loits_hip_version()); });
  m.def("forward_bounded", &forward_bounded);
  m.def("backward_bounded", &backward_bounded);
  m.def("forward_workspace_size", [](at::Tensor x_bins, at::Tensor q_bins, int64_t nmax) {
    const loits_hip_shape s = shape_of(x_bins, q_bins);
    return loits_hip_forward_workspace_size(&s, nmax);
  });
  m.def("backward_workspace_size", [](at::Tensor x_bins, at::Tensor q_bins) {
    const loits_hip_shape s = shape_of(x_bins, q_bins);
    return loits_hip_backward_workspace_size(&s, 0);
  });
  m.def("soft_histogram", &soft_histogram, py::arg("values"), py::arg("row_start"),
        py::arg("lo"), py::arg("hi"), py::arg("bins"), py::arg("width"),
        py::arg("stream") = py::none());
  m.def("soft_histogram_backward", &soft_histogram_backward, py::arg("values"),
        py::arg("row_start"), py::arg("lo"), py::arg("hi"), py::arg("bins"), py::arg("width"),
        py::arg("grad_hist"), py::arg("stream") = py::none());
  py::class_<BatchedPlan>(m, "BatchedPlan")
      .def(py::init<at::Tensor, int64_t, int64_t, int64_t, int64_t, int64_t, int64_t, int64_t,
                    std::vector<int64_t>, std::vector<int64_t>>(),
           py::arg("instances"), py::arg("dtype"), py::arg("threads"), py::arg("guide_shift"),
           py::arg("chunk_events"), py::arg("scan_mode"), py::arg("persistent"),
           py::arg("table_format"), py::arg("bucket_chunk_log2") = std::vector<int64_t>(),
           py::arg("bucket_guide_shift") = std::vector<int64_t>())
      .def("layout", &BatchedPlan::layout)
      .def("descriptors", &BatchedPlan::descriptors)
      .def("descriptors_for_streams", &BatchedPlan::descriptors_for_streams)
      .def("forward", &BatchedPlan::forward)
      .def("pack", &BatchedPlan::pack)
      .def("backward", &BatchedPlan::backward);
}
