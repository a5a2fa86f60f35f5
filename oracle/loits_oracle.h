/*
 * loits_oracle.h -- CPU restatement of the reference LOITS sampler (fp64).
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing under quantom_loits_amd/ links, loads or
 * calls this code; it is the checker that tests/, __graft_entry__.smoke() and
 * the cpu_baseline leg of bench.py compare the HIP path against.
 *
 * Every function restates one reference function (niteya-shah/quantom_loits)
 * and cites it.  Floating-point operation order follows the reference source;
 * the places where the reference build (g++ -O3 -march=native, whose default
 * -ffp-contract=fast fuses a multiply feeding an add) produces a fused
 * multiply-add are written here as explicit fma() calls, and this file is
 * compiled with -ffp-contract=off so that the explicit form is the only one.
 * tests/test_oracle_golden.py pins that choice against fixtures produced by
 * the compiled reference.
 *
 * Pinning status of each function (see tests/):
 *   philox4x32_10 / uniform_float : Random123 KATs + reference KAT + golden words
 *   allocate_counts               : golden fixtures + closed form (dyadic weights)
 *   density_* / cdf_*             : golden fixtures (bit-exact) + closed forms
 *   interpolate_*                 : golden fixtures + brute-force interval check
 *   compact                       : golden fixtures + brute-force filter
 *   *_vjp                         : golden fixtures + fp64 central FD (frozen
 *                                   interval) + adjoint dot-product identity
 * No function is "parity unpinned".
 */
#ifndef QUANTOM_LOITS_ORACLE_H
#define QUANTOM_LOITS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LOITS_ORACLE_EPSILON 1e-5 /* cpp/loits_core.hpp:15 */

typedef struct {
  int64_t batch, nx, ny, kx, kq, cells;
} loits_oracle_shape; /* cpp/loits_core.hpp:17-24 */

/* rng/philox.hpp:17-56 */
void loits_oracle_philox4x32_10(uint64_t block, uint64_t stream, uint64_t seed,
                                uint32_t out[4]);
/* rng/philox.hpp:59-61 */
float loits_oracle_uniform_float(uint32_t bits);
/* cpp/loits_core.cpp:39-53 (openmp/loits_core.cpp:53-68) */
void loits_oracle_fill_uniform(float* values, int64_t count, uint64_t seed,
                               uint64_t stream);
/* raw Philox words for `blocks` consecutive blocks (4 words each) */
void loits_oracle_philox_words(uint32_t* words, int64_t first_block,
                               int64_t blocks, uint64_t stream, uint64_t seed);

/* cpp/loits_core.cpp:22-37; returns nmax, writes active */
int64_t loits_oracle_allocate_counts(const double* weights, int64_t* counts,
                                     int64_t n_events,
                                     const loits_oracle_shape* s,
                                     int64_t* active);

/* cpp/loits_core.cpp:55-109 */
void loits_oracle_density_x(const double* bins, const double* xsec,
                            double* norm, double* rho,
                            const loits_oracle_shape* s);
void loits_oracle_density_q(const double* bins, const double* xsec,
                            double* norm, double* rho,
                            const loits_oracle_shape* s);
/* cpp/loits_core.cpp:111-165 */
void loits_oracle_cdf_x(const double* bins, const double* rho,
                        const uint8_t* acceptance, double* cdf,
                        const loits_oracle_shape* s);
void loits_oracle_cdf_q(const double* bins, const double* rho,
                        const uint8_t* acceptance, double* cdf,
                        const loits_oracle_shape* s);
/* cpp/loits_core.cpp:12-18 (the linear-scan interval), 167-237 */
void loits_oracle_interpolate_x(const double* bins, const double* cdf,
                                const float* u, const int64_t* counts,
                                int64_t nmax, double* dense, int16_t* indices,
                                const loits_oracle_shape* s);
void loits_oracle_interpolate_q(const double* bins, const double* cdf,
                                const float* u, const int64_t* counts,
                                int64_t nmax, double* dense, int16_t* indices,
                                const loits_oracle_shape* s);
/* openmp/loits_core.cpp:234-283 (two-pass, row_offsets); ordering equals
 * cpp/loits_core.cpp:239-266.  row_offsets may be NULL. */
int64_t loits_oracle_compact(const double* dense_x, const double* dense_q,
                             const int64_t* counts, int64_t nmax,
                             double* events, int64_t* packed,
                             int64_t* row_offsets,
                             const loits_oracle_shape* s);
/* cpp/loits_core.cpp:268-338 */
void loits_oracle_interpolation_vjp_x(const double* grad_events,
                                      const int64_t* packed, int64_t nvalid,
                                      const double* bins, const double* cdf,
                                      const float* u, const int16_t* indices,
                                      int64_t nmax, double* grad_cdf,
                                      const loits_oracle_shape* s);
void loits_oracle_interpolation_vjp_q(const double* grad_events,
                                      const int64_t* packed, int64_t nvalid,
                                      const double* bins, const double* cdf,
                                      const float* u, const int16_t* indices,
                                      int64_t nmax, double* grad_cdf,
                                      const loits_oracle_shape* s);
/* cpp/loits_core.cpp:340-398 */
void loits_oracle_cdf_vjp_x(const double* bins, const uint8_t* acceptance,
                            const double* grad_cdf, double* grad_rho,
                            const loits_oracle_shape* s);
void loits_oracle_cdf_vjp_q(const double* bins, const uint8_t* acceptance,
                            const double* grad_cdf, double* grad_rho,
                            const loits_oracle_shape* s);
/* cpp/loits_core.cpp:400-472 */
void loits_oracle_density_vjp_x(const double* bins, const double* xsec,
                                const double* norm, const double* grad_rho,
                                double* grad_xsec, const loits_oracle_shape* s);
void loits_oracle_density_vjp_q(const double* bins, const double* xsec,
                                const double* norm, const double* grad_rho,
                                double* grad_xsec, const loits_oracle_shape* s);

/*
 * One-dimensional batched restatement (the 1-D mapping nx = ny = 1 of the
 * reference, one curve per instance).  Instance i owns nodes
 * [node_off[i], node_off[i] + n_nodes[i]) of x / density, draws events[i]
 * uniforms from Philox stream streams[i] at counters counter_base[i] + slot,
 * and keeps an event when isfinite(value) && value != 0 and, if window != NULL,
 * window[0] <= value <= window[1].  Kept values are written in slot order,
 * instance after instance, into values_out; kept[i] receives the per-instance
 * count and the function returns the total.  With counter_base[i] = i*E and
 * streams[i] = 2*sequence this is exactly the x column of the reference
 * forward (cpp/bindings.cpp:24-130) under the 1-D mapping.
 */
int64_t loits_oracle_batched_forward(
    const double* x, const double* density, const int64_t* node_off,
    const int64_t* n_nodes, const int64_t* events, const uint64_t* counter_base,
    const uint64_t* streams, const uint8_t* accept, int64_t n_instances,
    uint64_t seed, const double* window, double* values_out, int32_t* bins_out,
    int64_t* kept, double* norm_out, double* cdf_out);

/* Reverse VJP of loits_oracle_batched_forward w.r.t. density (the x-axis
 * chain cpp/loits_core.cpp:268-302, 340-368, 400-435 under the 1-D mapping).
 * grad_values holds one cotangent per kept event, in output order. */
void loits_oracle_batched_backward(
    const double* x, const double* density, const int64_t* node_off,
    const int64_t* n_nodes, const int64_t* events, const uint64_t* counter_base,
    const uint64_t* streams, const uint8_t* accept, int64_t n_instances,
    uint64_t seed, const double* window, const double* grad_values,
    double* grad_density);

/* number of OpenMP threads the oracle uses (1 when built without OpenMP) */
int loits_oracle_threads(void);

#ifdef __cplusplus
}
#endif
#endif
