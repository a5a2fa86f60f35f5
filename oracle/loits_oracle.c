/*
 * loits_oracle.c -- CPU restatement of the reference LOITS sampler (fp64).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never part of it.
 * Built by oracle/Makefile with -ffp-contract=off; the fused multiply-adds
 * the reference build produces are explicit fma() calls below (each one is
 * marked "fused in the reference build").  See loits_oracle.h for the
 * pinning status of every function.
 */
#include "loits_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int loits_oracle_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* Philox4x32-10: rng/philox.hpp:17-56.  Counter (block lo, block hi, stream
 * lo, stream hi), key (seed lo, seed hi), ten rounds, key bumped by the Weyl
 * constants after every round but the last. */
void loits_oracle_philox4x32_10(uint64_t block, uint64_t stream, uint64_t seed,
                                uint32_t out[4]) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  uint32_t c0 = (uint32_t)block, c1 = (uint32_t)(block >> 32);
  uint32_t c2 = (uint32_t)stream, c3 = (uint32_t)(stream >> 32);
  uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)M0 * c0;
    uint64_t p1 = (uint64_t)M1 * c2;
    uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    uint32_t n1 = (uint32_t)p1;
    uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    uint32_t n3 = (uint32_t)p0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    if (r != 9) { k0 += W0; k1 += W1; }
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* rng/philox.hpp:59-61: top 24 bits scaled by 2^-24 (exact in float) */
float loits_oracle_uniform_float(uint32_t bits) {
  return (float)(bits >> 8) * 0x1.0p-24f;
}

/* cpp/loits_core.cpp:39-53: u[p] = word (p % 4) of philox(p / 4) */
void loits_oracle_fill_uniform(float* values, int64_t count, uint64_t seed,
                               uint64_t stream) {
  const int64_t blocks = (count + 3) / 4;
#pragma omp parallel for schedule(static)
  for (int64_t block = 0; block < blocks; ++block) {
    uint32_t w[4];
    loits_oracle_philox4x32_10((uint64_t)block, stream, seed, w);
    const int64_t base = block * 4;
    for (int e = 0; e < 4; ++e)
      if (base + e < count) values[base + e] = loits_oracle_uniform_float(w[e]);
  }
}

void loits_oracle_philox_words(uint32_t* words, int64_t first_block,
                               int64_t blocks, uint64_t stream, uint64_t seed) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < blocks; ++i)
    loits_oracle_philox4x32_10((uint64_t)(first_block + i), stream, seed,
                               words + 4 * i);
}

/* ------------------------------------------------------------------------ */
/* cpp/loits_core.cpp:22-37: count = (int64)|w * n_events| (truncation) */
int64_t loits_oracle_allocate_counts(const double* weights, int64_t* counts,
                                     int64_t n_events,
                                     const loits_oracle_shape* s,
                                     int64_t* active) {
  const int64_t n = s->batch * s->cells;
  const double scale = (double)n_events;
  int64_t nmax = 0, total = 0;
  for (int64_t i = 0; i < n; ++i) {
    const int64_t c = (int64_t)fabs(weights[i] * scale);
    counts[i] = c;
    total += c;
    if (c > nmax) nmax = c;
  }
  if (active) *active = total;
  return nmax;
}

/* ------------------------------------------------------------------------ */
/* One curve: trapezoid norm then rho = s * (1/n)  (cpp/loits_core.cpp:73-80) */
/* The reference build (g++ 11, -O3, x86-64 with AVX2/AVX-512) vectorizes the
 * norm sum and the density-VJP grad_norm sum as in-order ("fold-left")
 * reductions: 4-wide and 2-wide vector products added one by one without
 * fusion, then a scalar tail that is fused.  Term i of an nterms-long sum is
 * therefore fused iff it is that odd trailing term; the norm sum additionally
 * takes the all-fused scalar loop when it has at most three terms.  The
 * golden fixtures from the compiled reference pin this bit for bit. */
static inline int tail_fused(int64_t i, int64_t nterms) {
  return (nterms & 1) && i == nterms - 1;
}

static double curve_density(const double* xb, const double* xs, double* r,
                            int64_t k) {
  double n = 0.0;
  const int64_t nterms = k - 1;
  for (int64_t t = 1; t < k; ++t) {
    const double h = 0.5 * (xb[t] - xb[t - 1]);
    const double sum = xs[t - 1] + xs[t];
    if (nterms <= 3 || tail_fused(t - 1, nterms))
      n = fma(h, sum, n);
    else
      n = n + h * sum;
  }
  const double inv_n = 1.0 / n;
  if (r)
    for (int64_t t = 0; t < k; ++t) r[t] = xs[t] * inv_n;
  return n;
}

/* One curve: C_0 = 0, running trapezoid sum times acc (cpp/loits_core.cpp:130-135) */
static void curve_cdf(const double* xb, const double* r, double acc,
                      double* out, int64_t k) {
  double cumulative = 0.0;
  out[0] = 0.0;
  for (int64_t t = 1; t < k; ++t) {
    cumulative = fma(0.5 * (xb[t] - xb[t - 1]), r[t - 1] + r[t], cumulative); /* fused in the reference build */
    out[t] = cumulative * acc;
  }
}

/* cpp/loits_core.cpp:12-18: count-based interval by linear scan over all k */
static int64_t interval(const double* cdf, int64_t k, double u) {
  int64_t selected = -1;
  for (int64_t t = 0; t < k; ++t) selected += (int64_t)(u >= cdf[t]);
  if (selected < 0) selected = 0;
  if (selected > k - 2) selected = k - 2;
  return selected;
}

/* cpp/loits_core.cpp:190-195: inverse interpolation inside bin j */
static double interpolate_one(const double* bin, const double* curve,
                              int64_t j, double uv) {
  const double c0 = curve[j];
  const double inv_d = 1.0 / (curve[j + 1] - c0 + LOITS_ORACLE_EPSILON);
  const double m = (bin[j + 1] - bin[j]) * inv_d;
  return fma(m, uv - c0, bin[j]); /* fused in the reference build */
}

/* cpp/loits_core.cpp:294-301 (one row of the interpolation VJP) */
static void interpolation_vjp_one(const double* bin, const double* curve,
                                  double* gcurve, int64_t j, double uv,
                                  double gev) {
  const double c0 = curve[j];
  const double d = curve[j + 1] - c0 + LOITS_ORACLE_EPSILON;
  const double inv_d = 1.0 / d;
  const double t = uv - c0;
  const double db = bin[j + 1] - bin[j];
  const double upstream = gev * db;
  /* gcurve[j] += upstream * (-inv_d + t * inv_d * inv_d): both adds fused */
  gcurve[j] = fma(upstream, fma(t * inv_d, inv_d, -inv_d), gcurve[j]);
  /* gcurve[j + 1] -= upstream * t * inv_d * inv_d: fused negated multiply-add */
  gcurve[j + 1] = fma(-(upstream * t * inv_d), inv_d, gcurve[j + 1]);
}

/* cpp/loits_core.cpp:355-365 (one curve of the CDF VJP) */
static void curve_cdf_vjp(const double* xb, double acc, const double* gc,
                          double* gr, int64_t k) {
  memset(gr, 0, (size_t)k * sizeof(double));
  double suffix = 0.0;
  for (int64_t t = k - 1; t > 0; --t) {
    suffix += gc[t];
    const double h = 0.5 * (xb[t] - xb[t - 1]) * acc;
    gr[t - 1] = fma(h, suffix, gr[t - 1]); /* fused in the reference build */
    gr[t] = fma(h, suffix, gr[t]);         /* fused in the reference build */
  }
}

/* cpp/loits_core.cpp:416-431 (one curve of the density VJP) */
static void curve_density_vjp(const double* xb, const double* xs, double n,
                              const double* gr, double* gx, int64_t k) {
  const double inv_n = 1.0 / n;
  const double inv_n2 = inv_n * inv_n;
  double grad_norm = 0.0;
  for (int64_t t = 0; t < k; ++t) {
    gx[t] = gr[t] * inv_n;
    const double term = gr[t] * xs[t];
    if (tail_fused(t, k))
      grad_norm = fma(-term, inv_n2, grad_norm);
    else
      grad_norm = grad_norm - term * inv_n2;
  }
  for (int64_t t = 1; t < k; ++t) {
    const double h = 0.5 * (xb[t] - xb[t - 1]);
    gx[t - 1] = fma(h, grad_norm, gx[t - 1]); /* fused in the reference build */
    gx[t] = fma(h, grad_norm, gx[t]);         /* fused in the reference build */
  }
}

/* ------------------------------------------------------------------------ */
/* Layout helpers.  x-curve of cell (ix, iy) lives at [b, iy, ix, :]; its
 * nodes are x_bins[b, ix, :].  q-curve of cell (ix, iy) lives at
 * [b, ix, iy, :]; its nodes are q_bins[b, iy, :]
 * (cpp/loits_core.cpp:62-66, 90-94). */
static inline void decode_curve(int64_t curve, int64_t outer, int64_t inner,
                                int64_t* b, int64_t* o, int64_t* i) {
  *b = curve / (outer * inner);
  const int64_t r = curve - *b * outer * inner;
  *o = r / inner;
  *i = r - *o * inner;
}

/* cpp/loits_core.cpp:55-81 */
void loits_oracle_density_x(const double* bins, const double* xsec,
                            double* norm, double* rho,
                            const loits_oracle_shape* s) {
  const int64_t curves = s->batch * s->ny * s->nx;
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < curves; ++c) {
    int64_t b, iy, ix;
    decode_curve(c, s->ny, s->nx, &b, &iy, &ix);
    const double* xb = bins + (b * s->nx + ix) * s->kx;
    norm[c] = curve_density(xb, xsec + c * s->kx, rho + c * s->kx, s->kx);
  }
}

/* cpp/loits_core.cpp:83-109 */
void loits_oracle_density_q(const double* bins, const double* xsec,
                            double* norm, double* rho,
                            const loits_oracle_shape* s) {
  const int64_t curves = s->batch * s->nx * s->ny;
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < curves; ++c) {
    int64_t b, ix, iy;
    decode_curve(c, s->nx, s->ny, &b, &ix, &iy);
    const double* qb = bins + (b * s->ny + iy) * s->kq;
    norm[c] = curve_density(qb, xsec + c * s->kq, rho + c * s->kq, s->kq);
  }
}

/* cpp/loits_core.cpp:111-137 */
void loits_oracle_cdf_x(const double* bins, const double* rho,
                        const uint8_t* acceptance, double* cdf,
                        const loits_oracle_shape* s) {
  const int64_t curves = s->batch * s->ny * s->nx;
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < curves; ++c) {
    int64_t b, iy, ix;
    decode_curve(c, s->ny, s->nx, &b, &iy, &ix);
    const double acc = acceptance[(b * s->nx + ix) * s->ny + iy] ? 1.0 : 0.0;
    curve_cdf(bins + (b * s->nx + ix) * s->kx, rho + c * s->kx, acc,
              cdf + c * s->kx, s->kx);
  }
}

/* cpp/loits_core.cpp:139-165 */
void loits_oracle_cdf_q(const double* bins, const double* rho,
                        const uint8_t* acceptance, double* cdf,
                        const loits_oracle_shape* s) {
  const int64_t curves = s->batch * s->nx * s->ny;
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < curves; ++c) {
    int64_t b, ix, iy;
    decode_curve(c, s->nx, s->ny, &b, &ix, &iy);
    const double acc = acceptance[(b * s->nx + ix) * s->ny + iy] ? 1.0 : 0.0;
    curve_cdf(bins + (b * s->ny + iy) * s->kq, rho + c * s->kq, acc,
              cdf + c * s->kq, s->kq);
  }
}

/* cpp/loits_core.cpp:167-201 (openmp/loits_core.cpp:166-198) */
void loits_oracle_interpolate_x(const double* bins, const double* cdf,
                                const float* u, const int64_t* counts,
                                int64_t nmax, double* dense, int16_t* indices,
                                const loits_oracle_shape* s) {
  const int64_t gcells = s->batch * s->cells;
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t g = 0; g < gcells; ++g) {
    const int64_t b = g / s->cells, cell = g - b * s->cells;
    const int64_t ix = cell / s->ny, iy = cell - ix * s->ny;
    const double* bin = bins + (b * s->nx + ix) * s->kx;
    const double* curve = cdf + ((b * s->ny + iy) * s->nx + ix) * s->kx;
    const int64_t base = g * nmax;
    for (int64_t slot = 0; slot < counts[g]; ++slot) {
      const int64_t p = base + slot;
      const double uv = (double)u[p];
      const int64_t j = interval(curve, s->kx, uv);
      indices[p] = (int16_t)j;
      dense[p] = interpolate_one(bin, curve, j, uv);
    }
  }
}

/* cpp/loits_core.cpp:203-237 (openmp/loits_core.cpp:200-232) */
void loits_oracle_interpolate_q(const double* bins, const double* cdf,
                                const float* u, const int64_t* counts,
                                int64_t nmax, double* dense, int16_t* indices,
                                const loits_oracle_shape* s) {
  const int64_t gcells = s->batch * s->cells;
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t g = 0; g < gcells; ++g) {
    const int64_t b = g / s->cells, cell = g - b * s->cells;
    const int64_t ix = cell / s->ny, iy = cell - ix * s->ny;
    const double* bin = bins + (b * s->ny + iy) * s->kq;
    const double* curve = cdf + ((b * s->nx + ix) * s->ny + iy) * s->kq;
    const int64_t base = g * nmax;
    for (int64_t slot = 0; slot < counts[g]; ++slot) {
      const int64_t p = base + slot;
      const double uv = (double)u[p];
      const int64_t j = interval(curve, s->kq, uv);
      indices[p] = (int16_t)j;
      dense[p] = interpolate_one(bin, curve, j, uv);
    }
  }
}

/* keep iff both finite and x*q != 0 (cpp/loits_core.cpp:256) */
static inline int keep_pair(double xv, double qv) {
  return isfinite(xv) && isfinite(qv) && xv * qv != 0.0;
}

/* openmp/loits_core.cpp:234-283 */
int64_t loits_oracle_compact(const double* dense_x, const double* dense_q,
                             const int64_t* counts, int64_t nmax,
                             double* events, int64_t* packed,
                             int64_t* row_offsets,
                             const loits_oracle_shape* s) {
  const int64_t gcells = s->batch * s->cells;
  int64_t* offs = row_offsets ? row_offsets
                              : (int64_t*)malloc(sizeof(int64_t) * (size_t)(gcells + 1));
  offs[0] = 0;
#pragma omp parallel for schedule(static)
  for (int64_t g = 0; g < gcells; ++g) {
    int64_t valid = 0;
    for (int64_t slot = 0; slot < counts[g]; ++slot) {
      const int64_t p = g * nmax + slot;
      valid += keep_pair(dense_x[p], dense_q[p]);
    }
    offs[g + 1] = valid;
  }
  for (int64_t g = 0; g < gcells; ++g) offs[g + 1] += offs[g];
#pragma omp parallel for schedule(static)
  for (int64_t g = 0; g < gcells; ++g) {
    int64_t row = offs[g];
    for (int64_t slot = 0; slot < counts[g]; ++slot) {
      const int64_t p = g * nmax + slot;
      if (keep_pair(dense_x[p], dense_q[p])) {
        events[2 * row] = dense_x[p];
        events[2 * row + 1] = dense_q[p];
        packed[row] = p;
        ++row;
      }
    }
  }
  const int64_t total = offs[gcells];
  if (!row_offsets) free(offs);
  return total;
}

/* first row whose packed slot is >= p (rows are sorted by slot) */
static int64_t lower_row(const int64_t* packed, int64_t nvalid, int64_t p) {
  int64_t lo = 0, hi = nvalid;
  while (lo < hi) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (packed[mid] < p) lo = mid + 1; else hi = mid;
  }
  return lo;
}

/* cpp/loits_core.cpp:268-302.  Rows of one cell are contiguous and
 * accumulate in row order, as in openmp/loits_core.cpp:285-318, so the
 * per-cell parallel loop gives the serial loop's result bit for bit. */
void loits_oracle_interpolation_vjp_x(const double* grad_events,
                                      const int64_t* packed, int64_t nvalid,
                                      const double* bins, const double* cdf,
                                      const float* u, const int16_t* indices,
                                      int64_t nmax, double* grad_cdf,
                                      const loits_oracle_shape* s) {
  memset(grad_cdf, 0, sizeof(double) * (size_t)(s->batch * s->ny * s->nx * s->kx));
  const int64_t gcells = s->batch * s->cells;
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t g = 0; g < gcells; ++g) {
    const int64_t b = g / s->cells, cell = g - b * s->cells;
    const int64_t ix = cell / s->ny, iy = cell - ix * s->ny;
    const double* bin = bins + (b * s->nx + ix) * s->kx;
    const int64_t curve_off = ((b * s->ny + iy) * s->nx + ix) * s->kx;
    const int64_t r0 = lower_row(packed, nvalid, g * nmax);
    const int64_t r1 = lower_row(packed, nvalid, (g + 1) * nmax);
    for (int64_t row = r0; row < r1; ++row) {
      const int64_t p = packed[row];
      interpolation_vjp_one(bin, cdf + curve_off, grad_cdf + curve_off,
                            indices[p], (double)u[p], grad_events[2 * row]);
    }
  }
}

/* cpp/loits_core.cpp:304-338 */
void loits_oracle_interpolation_vjp_q(const double* grad_events,
                                      const int64_t* packed, int64_t nvalid,
                                      const double* bins, const double* cdf,
                                      const float* u, const int16_t* indices,
                                      int64_t nmax, double* grad_cdf,
                                      const loits_oracle_shape* s) {
  memset(grad_cdf, 0, sizeof(double) * (size_t)(s->batch * s->nx * s->ny * s->kq));
  const int64_t gcells = s->batch * s->cells;
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t g = 0; g < gcells; ++g) {
    const int64_t b = g / s->cells, cell = g - b * s->cells;
    const int64_t ix = cell / s->ny, iy = cell - ix * s->ny;
    const double* bin = bins + (b * s->ny + iy) * s->kq;
    const int64_t curve_off = ((b * s->nx + ix) * s->ny + iy) * s->kq;
    const int64_t r0 = lower_row(packed, nvalid, g * nmax);
    const int64_t r1 = lower_row(packed, nvalid, (g + 1) * nmax);
    for (int64_t row = r0; row < r1; ++row) {
      const int64_t p = packed[row];
      interpolation_vjp_one(bin, cdf + curve_off, grad_cdf + curve_off,
                            indices[p], (double)u[p], grad_events[2 * row + 1]);
    }
  }
}

/* cpp/loits_core.cpp:340-368 */
void loits_oracle_cdf_vjp_x(const double* bins, const uint8_t* acceptance,
                            const double* grad_cdf, double* grad_rho,
                            const loits_oracle_shape* s) {
  const int64_t curves = s->batch * s->ny * s->nx;
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < curves; ++c) {
    int64_t b, iy, ix;
    decode_curve(c, s->ny, s->nx, &b, &iy, &ix);
    const double acc = acceptance[(b * s->nx + ix) * s->ny + iy] ? 1.0 : 0.0;
    curve_cdf_vjp(bins + (b * s->nx + ix) * s->kx, acc, grad_cdf + c * s->kx,
                  grad_rho + c * s->kx, s->kx);
  }
}

/* cpp/loits_core.cpp:370-398 */
void loits_oracle_cdf_vjp_q(const double* bins, const uint8_t* acceptance,
                            const double* grad_cdf, double* grad_rho,
                            const loits_oracle_shape* s) {
  const int64_t curves = s->batch * s->nx * s->ny;
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < curves; ++c) {
    int64_t b, ix, iy;
    decode_curve(c, s->nx, s->ny, &b, &ix, &iy);
    const double acc = acceptance[(b * s->nx + ix) * s->ny + iy] ? 1.0 : 0.0;
    curve_cdf_vjp(bins + (b * s->ny + iy) * s->kq, acc, grad_cdf + c * s->kq,
                  grad_rho + c * s->kq, s->kq);
  }
}

/* cpp/loits_core.cpp:400-435 */
void loits_oracle_density_vjp_x(const double* bins, const double* xsec,
                                const double* norm, const double* grad_rho,
                                double* grad_xsec, const loits_oracle_shape* s) {
  const int64_t curves = s->batch * s->ny * s->nx;
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < curves; ++c) {
    int64_t b, iy, ix;
    decode_curve(c, s->ny, s->nx, &b, &iy, &ix);
    curve_density_vjp(bins + (b * s->nx + ix) * s->kx, xsec + c * s->kx, norm[c],
                      grad_rho + c * s->kx, grad_xsec + c * s->kx, s->kx);
  }
}

/* cpp/loits_core.cpp:437-472 */
void loits_oracle_density_vjp_q(const double* bins, const double* xsec,
                                const double* norm, const double* grad_rho,
                                double* grad_xsec, const loits_oracle_shape* s) {
  const int64_t curves = s->batch * s->nx * s->ny;
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < curves; ++c) {
    int64_t b, ix, iy;
    decode_curve(c, s->nx, s->ny, &b, &ix, &iy);
    curve_density_vjp(bins + (b * s->ny + iy) * s->kq, xsec + c * s->kq, norm[c],
                      grad_rho + c * s->kq, grad_xsec + c * s->kq, s->kq);
  }
}

/* ------------------------------------------------------------------------ */
/* 1-D batched restatement. */

static inline int keep_value(double v, const double* window) {
  if (!(isfinite(v) && v != 0.0)) return 0;
  if (window && !(v >= window[0] && v <= window[1])) return 0;
  return 1;
}

/* uniform for slot `slot` of an instance: counter p = base + slot */
static inline double slot_uniform(uint64_t base, int64_t slot, uint64_t stream,
                                  uint64_t seed) {
  const uint64_t p = base + (uint64_t)slot;
  uint32_t w[4];
  loits_oracle_philox4x32_10(p >> 2, stream, seed, w);
  return (double)loits_oracle_uniform_float(w[p & 3]);
}

int64_t loits_oracle_batched_forward(
    const double* x, const double* density, const int64_t* node_off,
    const int64_t* n_nodes, const int64_t* events, const uint64_t* counter_base,
    const uint64_t* streams, const uint8_t* accept, int64_t n_instances,
    uint64_t seed, const double* window, double* values_out, int32_t* bins_out,
    int64_t* kept, double* norm_out, double* cdf_out) {
  int64_t* cap = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n_instances + 1));
  cap[0] = 0;
  for (int64_t i = 0; i < n_instances; ++i) cap[i + 1] = cap[i] + events[i];
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t i = 0; i < n_instances; ++i) {
    const int64_t k = n_nodes[i];
    const double* xb = x + node_off[i];
    double* rho = (double*)malloc(sizeof(double) * (size_t)k);
    double* cdf = cdf_out ? cdf_out + node_off[i]
                          : (double*)malloc(sizeof(double) * (size_t)k);
    const double n = curve_density(xb, density + node_off[i], rho, k);
    if (norm_out) norm_out[i] = n;
    curve_cdf(xb, rho, (accept && !accept[i]) ? 0.0 : 1.0, cdf, k);
    int64_t row = cap[i];
    for (int64_t slot = 0; slot < events[i]; ++slot) {
      const double uv = slot_uniform(counter_base[i], slot, streams[i], seed);
      const int64_t j = interval(cdf, k, uv);
      const double v = interpolate_one(xb, cdf, j, uv);
      if (keep_value(v, window)) {
        values_out[row] = v;
        if (bins_out) bins_out[row] = (int32_t)j;
        ++row;
      }
    }
    kept[i] = row - cap[i];
    free(rho);
    if (!cdf_out) free(cdf);
  }
  /* stable left shift into the contiguous kept layout */
  int64_t out = 0;
  for (int64_t i = 0; i < n_instances; ++i) {
    if (out != cap[i]) {
      memmove(values_out + out, values_out + cap[i], sizeof(double) * (size_t)kept[i]);
      if (bins_out)
        memmove(bins_out + out, bins_out + cap[i], sizeof(int32_t) * (size_t)kept[i]);
    }
    out += kept[i];
  }
  free(cap);
  return out;
}

void loits_oracle_batched_backward(
    const double* x, const double* density, const int64_t* node_off,
    const int64_t* n_nodes, const int64_t* events, const uint64_t* counter_base,
    const uint64_t* streams, const uint8_t* accept, int64_t n_instances,
    uint64_t seed, const double* window, const double* grad_values,
    double* grad_density) {
  /* pass 1: kept count per instance (the forward's row numbering) */
  int64_t* rows = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n_instances + 1));
  double** cdfs = (double**)malloc(sizeof(double*) * (size_t)n_instances);
  double* norms = (double*)malloc(sizeof(double) * (size_t)n_instances);
  rows[0] = 0;
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t i = 0; i < n_instances; ++i) {
    const int64_t k = n_nodes[i];
    const double* xb = x + node_off[i];
    double* rho = (double*)malloc(sizeof(double) * (size_t)k);
    cdfs[i] = (double*)malloc(sizeof(double) * (size_t)k);
    norms[i] = curve_density(xb, density + node_off[i], rho, k);
    curve_cdf(xb, rho, (accept && !accept[i]) ? 0.0 : 1.0, cdfs[i], k);
    free(rho);
    int64_t kept = 0;
    for (int64_t slot = 0; slot < events[i]; ++slot) {
      const double uv = slot_uniform(counter_base[i], slot, streams[i], seed);
      const int64_t j = interval(cdfs[i], k, uv);
      kept += keep_value(interpolate_one(xb, cdfs[i], j, uv), window);
    }
    rows[i + 1] = kept;
  }
  for (int64_t i = 0; i < n_instances; ++i) rows[i + 1] += rows[i];
  /* pass 2: interpolation, CDF and density VJPs per instance */
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t i = 0; i < n_instances; ++i) {
    const int64_t k = n_nodes[i];
    const double* xb = x + node_off[i];
    const double* cdf = cdfs[i];
    double* gc = (double*)calloc((size_t)k, sizeof(double));
    double* gr = (double*)malloc(sizeof(double) * (size_t)k);
    int64_t row = rows[i];
    for (int64_t slot = 0; slot < events[i]; ++slot) {
      const double uv = slot_uniform(counter_base[i], slot, streams[i], seed);
      const int64_t j = interval(cdf, k, uv);
      if (!keep_value(interpolate_one(xb, cdf, j, uv), window)) continue;
      interpolation_vjp_one(xb, cdf, gc, j, uv, grad_values[row]);
      ++row;
    }
    curve_cdf_vjp(xb, (accept && !accept[i]) ? 0.0 : 1.0, gc, gr, k);
    curve_density_vjp(xb, density + node_off[i], norms[i], gr,
                      grad_density + node_off[i], k);
    free(gc);
    free(gr);
    free(cdfs[i]);
  }
  free(rows);
  free(cdfs);
  free(norms);
}
