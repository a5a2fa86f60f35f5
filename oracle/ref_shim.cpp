// ref_shim.cpp -- extern "C" entry points around the reference's own compiled
// C++ cores (cpp/loits_core.cpp or openmp/loits_core.cpp, compiled in place
// from /root/reference by oracle/Makefile into oracle/_ref/).
//
// TEST INFRASTRUCTURE ONLY.  This file holds no reference code: it calls the
// reference functions through their own header so that Python (ctypes) can
// reach them.  REF_OPENMP selects the openmp core's namespace and signatures
// (openmp/loits_core.hpp:352-459) instead of the cpp core's
// (cpp/loits_core.hpp:31-137).
#include <cstdint>

#include "loits_core.hpp"
#include "philox.hpp"

#ifdef REF_OPENMP
namespace core = openmp_loits;
#else
namespace core = loits;
#endif

namespace {
core::Shape shape_of(const int64_t* s) {
  return core::Shape{s[0], s[1], s[2], s[3], s[4], s[5]};
}
const bool* as_bool(const uint8_t* a) { return reinterpret_cast<const bool*>(a); }
}  // namespace

extern "C" {

int ref_is_openmp() {
#ifdef REF_OPENMP
  return 1;
#else
  return 0;
#endif
}

void ref_philox(uint64_t block, uint64_t stream, uint64_t seed, uint32_t* out) {
  const auto r = quantom::rng::philox4x32_10(block, stream, seed);
  out[0] = r.x0;
  out[1] = r.x1;
  out[2] = r.x2;
  out[3] = r.x3;
}

void ref_fill_uniform(float* values, int64_t count, uint64_t seed, uint64_t stream) {
  core::fill_uniform(values, count, seed, stream);
}

void ref_allocate_counts(const double* weights, int64_t* counts, int64_t n_events,
                         const int64_t* shape, int64_t* out) {
  const auto a = core::allocate_counts(weights, counts, n_events, shape_of(shape));
  out[0] = a.nmax;
  out[1] = a.active;
}

void ref_density_x(const double* bins, const double* xsec, double* norm, double* rho,
                   const int64_t* shape) {
  core::density_x(bins, xsec, norm, rho, shape_of(shape));
}

void ref_density_q(const double* bins, const double* xsec, double* norm, double* rho,
                   const int64_t* shape) {
  core::density_q(bins, xsec, norm, rho, shape_of(shape));
}

void ref_cdf_x(const double* bins, const double* rho, const uint8_t* acc, double* cdf,
               const int64_t* shape) {
  core::cdf_x(bins, rho, as_bool(acc), cdf, shape_of(shape));
}

void ref_cdf_q(const double* bins, const double* rho, const uint8_t* acc, double* cdf,
               const int64_t* shape) {
  core::cdf_q(bins, rho, as_bool(acc), cdf, shape_of(shape));
}

void ref_interpolate_x(const double* bins, const double* cdf, const float* u,
                       const int64_t* counts, int64_t nmax, double* dense,
                       int16_t* indices, const int64_t* shape) {
  core::interpolate_x(bins, cdf, u, counts, nmax, dense, indices, shape_of(shape));
}

void ref_interpolate_q(const double* bins, const double* cdf, const float* u,
                       const int64_t* counts, int64_t nmax, double* dense,
                       int16_t* indices, const int64_t* shape) {
  core::interpolate_q(bins, cdf, u, counts, nmax, dense, indices, shape_of(shape));
}

int64_t ref_compact(const double* dense_x, const double* dense_q, const int64_t* counts,
                    int64_t nmax, double* events, int64_t* packed, int64_t* row_offsets,
                    const int64_t* shape) {
#ifdef REF_OPENMP
  return core::compact(dense_x, dense_q, counts, nmax, events, packed, row_offsets,
                       shape_of(shape));
#else
  (void)row_offsets;
  return core::compact(dense_x, dense_q, counts, nmax, events, packed, shape_of(shape));
#endif
}

void ref_interpolation_vjp_x(const double* grad_events, const int64_t* packed,
                             int64_t nvalid, const int64_t* row_offsets,
                             const double* bins, const double* cdf, const float* u,
                             const int16_t* indices, int64_t nmax, double* grad_cdf,
                             const int64_t* shape) {
#ifdef REF_OPENMP
  (void)nvalid;
  core::interpolation_vjp_x(grad_events, packed, row_offsets, bins, cdf, u, indices, nmax,
                            grad_cdf, shape_of(shape));
#else
  (void)row_offsets;
  core::interpolation_vjp_x(grad_events, packed, nvalid, bins, cdf, u, indices, nmax,
                            grad_cdf, shape_of(shape));
#endif
}

void ref_interpolation_vjp_q(const double* grad_events, const int64_t* packed,
                             int64_t nvalid, const int64_t* row_offsets,
                             const double* bins, const double* cdf, const float* u,
                             const int16_t* indices, int64_t nmax, double* grad_cdf,
                             const int64_t* shape) {
#ifdef REF_OPENMP
  (void)nvalid;
  core::interpolation_vjp_q(grad_events, packed, row_offsets, bins, cdf, u, indices, nmax,
                            grad_cdf, shape_of(shape));
#else
  (void)row_offsets;
  core::interpolation_vjp_q(grad_events, packed, nvalid, bins, cdf, u, indices, nmax,
                            grad_cdf, shape_of(shape));
#endif
}

void ref_cdf_vjp_x(const double* bins, const uint8_t* acc, const double* grad_cdf,
                   double* grad_rho, const int64_t* shape) {
  core::cdf_vjp_x(bins, as_bool(acc), grad_cdf, grad_rho, shape_of(shape));
}

void ref_cdf_vjp_q(const double* bins, const uint8_t* acc, const double* grad_cdf,
                   double* grad_rho, const int64_t* shape) {
  core::cdf_vjp_q(bins, as_bool(acc), grad_cdf, grad_rho, shape_of(shape));
}

void ref_density_vjp_x(const double* bins, const double* xsec, const double* norm,
                       const double* grad_rho, double* grad_xsec, const int64_t* shape) {
  core::density_vjp_x(bins, xsec, norm, grad_rho, grad_xsec, shape_of(shape));
}

void ref_density_vjp_q(const double* bins, const double* xsec, const double* norm,
                       const double* grad_rho, double* grad_xsec, const int64_t* shape) {
  core::density_vjp_q(bins, xsec, norm, grad_rho, grad_xsec, shape_of(shape));
}

}  // extern "C"
