/* ORACLE — test infrastructure only.  Imported by tests/ (and nothing in the
 * product path): a serial CPU restatement of the successor-table step of the
 * contact game, against which the GPU table builder (pe_car_stencil) is
 * checked.
 *
 * Follows local_game/grid_game.py:81-123 (VehicleGrid.stencil) and :38-79
 * (_nearest_supported), with the reference's float64 operations:
 *   - per axis, the lower grid index and the fraction toward the upper one:
 *     heading periodic, cell = ((h + pi) mod 2 pi) / (2 pi / n) with numpy's
 *     remainder (:96-99); other axes clipped, lower index =
 *     searchsorted(axis, z, 'right') - 1 kept in [0, n - 2], fraction
 *     (z - axis[j]) / (axis[j + 1] - axis[j]) (:100-106); validity = finite
 *     and within the axis ends +- 1e-9 (:94, :101);
 *   - vertex weight = product over axes 0..3 of (1 - frac) or frac, vertex c
 *     taking the upper index on axis k when bit 3 - k of c is set (BITS, :19);
 *   - illegal vertices weigh 0, the retained mass is the sum of the 16
 *     weights in numpy's pairwise order (:111-114), weights are divided by it
 *     when it exceeds 1e-9 and rounded to float32 (:115-120);
 *   - a valid successor with mass <= 1e-9 takes the fallback (:38-79): the
 *     min(64, n) nearest legal states in the features (x/dx, y/dx, cos h/dh,
 *     sin h/dh, v/dv) by squared distance summed in feature order, ties by
 *     index; the first 16 of them joined to the successor by a clear straight
 *     segment (pe_oracle_batch_connections); weights 1/max(d, 1e-8)^2 over
 *     their numpy-pairwise sum.
 * Pinned in tests/test_oracle_tables.py by multilinear exactness on linear
 * fields, renormalisation and periodicity known answers, numpy's own sum of
 * the returned raw weights, and scipy's cKDTree for the neighbour lists.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

void pe_oracle_batch_connections(const double *from, const double *to, int64_t n,
                                 const double *obs, int nc, double half,
                                 const double *param, uint8_t *out, int threads);

/* numpy's pairwise summation of count values (count <= 128 path). */
static double numpy_pairwise(const double *a, int count) {
  if (count < 8) {
    double s = 0.;
    for (int i = 0; i < count; i++) s += a[i];
    return s;
  }
  double r[8];
  memcpy(r, a, sizeof r);
  int i = 8;
  for (; i < count - (count % 8); i += 8)
    for (int j = 0; j < 8; j++) r[j] += a[i + j];
  double s = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
  for (; i < count; i++) s += a[i];
  return s;
}

static double numpy_remainder(double a, double b) {
  double m = fmod(a, b);
  if (m != 0. && ((b < 0) != (m < 0))) m += b;
  if (m == 0.) m = copysign(0., b);
  return m;
}

/* Cell of x on one axis: lower index, upper index, fraction; clears *ok
 * when x is outside the (non-periodic) axis. */
static void cell(const double *axis, int n, int periodic, double x, int *lo, int *hi,
                 double *frac, int *ok) {
  if (!isfinite(x) && (periodic || isnan(x))) { /* undefined in the reference */
    *ok = 0;
    *lo = periodic ? 0 : n - 2;
    *hi = periodic ? (n > 1) : n - 1;
    *frac = NAN;
    return;
  }
  if (periodic) {
    const double step = 2 * M_PI / n;
    const double c = numpy_remainder(x + M_PI, 2 * M_PI) / step;
    const double f = floor(c);
    long long j = (long long)f % n;
    if (j < 0) j += n;
    *lo = (int)j;
    *hi = (int)((j + 1) % n);
    *frac = c - f;
    return;
  }
  if (!(x >= axis[0] - 1e-9) || !(x <= axis[n - 1] + 1e-9)) *ok = 0;
  double z = x < axis[0] ? axis[0] : x;
  if (z > axis[n - 1]) z = axis[n - 1];
  int count = 0; /* axis values <= z */
  while (count < n && axis[count] <= z) count++;
  int j = count - 1;
  if (j < 0) j = 0;
  if (j > n - 2) j = n - 2;
  *lo = j;
  *hi = j + 1;
  *frac = (z - axis[j]) / (axis[j + 1] - axis[j]);
}

/* Scaled fallback features of one state. */
static void features(const double *s, double dx, double dh, double dv, double *f) {
  f[0] = s[0] / dx;
  f[1] = s[1] / dx;
  f[2] = cos(s[2]) / dh;
  f[3] = sin(s[2]) / dh;
  f[4] = s[3] / dv;
}

static double sq_distance(const double *a, const double *b) {
  double s = 0.;
  for (int e = 0; e < 5; e++) {
    const double t = a[e] - b[e];
    s = s + t * t;
  }
  return s;
}

/* The k nearest of n feature rows to q: ids and squared distances, ordered
 * by (distance, index).  Selection by repeated scans (k is at most 64). */
static void nearest(const double *feat, int n, const double *q, int k, int *ids,
                    double *d2) {
  double last_d = -1.;
  int last_i = -1;
  for (int r = 0; r < k; r++) {
    double best_d = INFINITY;
    int best_i = -1;
    for (int c = 0; c < n; c++) {
      const double d = sq_distance(q, feat + 5 * (size_t)c);
      const int after = d > last_d || (d == last_d && c > last_i);
      const int before = best_i < 0 || d < best_d || (d == best_d && c < best_i);
      if (after && before) {
        best_d = d;
        best_i = c;
      }
    }
    ids[r] = best_i;
    d2[r] = best_d;
    last_d = best_d;
    last_i = best_i;
  }
}

/* k nearest legal states (feature rows feat [n,5]) of m query states
 * [m,4] with the given scales: ids [m,k], distances (square roots) [m,k]. */
void pe_oracle_nearest(const double *feat, int n, const double *queries, int64_t m, int k,
                       double dx, double dh, double dv, int *ids, double *dist) {
  double *d2 = malloc(sizeof(double) * k);
  for (int64_t i = 0; i < m; i++) {
    double q[5];
    features(queries + 4 * i, dx, dh, dv, q);
    nearest(feat, n, q, k, ids + (size_t)k * i, d2);
    for (int r = 0; r < k; r++) dist[(size_t)k * i + r] = sqrt(d2[r]);
  }
  free(d2);
}

/* The whole stencil step for m successors; raw [m,16] receives the vertex
 * weights before renormalisation (0 for illegal vertices), may be NULL. */
int64_t pe_oracle_car_stencil(const int *n, const double *ax0, const double *ax1,
                              const double *ax2, const double *ax3, const int32_t *d2c,
                              const double *next, int64_t m, const double *obs, int nc,
                              double half, const double *param, int32_t *ids, float *w,
                              uint8_t *valid, double *mass, double *raw) {
  const double *axes[4] = {ax0, ax1, ax2, ax3};
  long long stride[4];
  stride[3] = 1;
  for (int k = 2; k >= 0; k--) stride[k] = stride[k + 1] * n[k + 1];
  const long long dense_n = stride[0] * n[0];
  int legal = 0;
  for (long long d = 0; d < dense_n; d++) legal += d2c[d] >= 0;
  double *lstates = malloc(sizeof(double) * 4 * (legal ? legal : 1));
  for (long long d = 0; d < dense_n; d++) {
    if (d2c[d] < 0) continue;
    long long r = d;
    for (int k = 0; k < 4; k++) {
      lstates[4 * (size_t)d2c[d] + k] = axes[k][r / stride[k]];
      r %= stride[k];
    }
  }
  const double dx = ax0[1] - ax0[0];
  const double dh = 2 * sin(M_PI / n[2]);
  double dv = INFINITY;
  for (int j = 1; j < n[3]; j++)
    if (ax3[j] - ax3[j - 1] < dv) dv = ax3[j] - ax3[j - 1];
  double *feat = malloc(sizeof(double) * 5 * (legal ? legal : 1));
  for (int c = 0; c < legal; c++) features(lstates + 4 * (size_t)c, dx, dh, dv, feat + 5 * (size_t)c);
  const int kk = legal < 64 ? legal : 64;
  int64_t fallbacks = 0;

  for (int64_t i = 0; i < m; i++) {
    const double *s = next + 4 * i;
    int ok = 1, lo[4], hi[4];
    double frac[4];
    for (int k = 0; k < 4; k++) {
      if (!isfinite(s[k])) ok = 0;
      cell(axes[k], n[k], k == 2, s[k], &lo[k], &hi[k], &frac[k], &ok);
    }
    double wv[16];
    int32_t iv[16];
    for (int b0 = 0; b0 < 2; b0++)
      for (int b1 = 0; b1 < 2; b1++)
        for (int b2 = 0; b2 < 2; b2++)
          for (int b3 = 0; b3 < 2; b3++) {
            const int bit[4] = {b0, b1, b2, b3};
            const int c = 8 * b0 + 4 * b1 + 2 * b2 + b3;
            double wc = 1.;
            long long dense = 0;
            for (int k = 0; k < 4; k++) {
              wc = wc * (bit[k] ? frac[k] : 1. - frac[k]);
              dense += (bit[k] ? hi[k] : lo[k]) * stride[k];
            }
            iv[c] = d2c[dense];
            wv[c] = iv[c] >= 0 ? wc : 0.;
          }
    const double mm = numpy_pairwise(wv, 16);
    if (raw) memcpy(raw + 16 * i, wv, sizeof wv);
    mass[i] = mm;
    for (int c = 0; c < 16; c++) {
      ids[16 * i + c] = iv[c] > 0 ? iv[c] : 0;
      w[16 * i + c] = mm > 1e-9 ? (float)(wv[c] / mm) : 0.f;
    }
    valid[i] = (uint8_t)(ok && mm > 1e-9);
    if (!(ok && mm <= 1e-9) || legal == 0) continue;
    /* fallback */
    fallbacks++;
    double q[5], d2[64];
    int cand[64];
    features(s, dx, dh, dv, q);
    nearest(feat, legal, q, kk, cand, d2);
    double from[4 * 64], to[4 * 64];
    uint8_t linked[64];
    for (int r = 0; r < kk; r++) {
      memcpy(from + 4 * r, s, 4 * sizeof(double));
      memcpy(to + 4 * r, lstates + 4 * (size_t)cand[r], 4 * sizeof(double));
    }
    pe_oracle_batch_connections(from, to, kk, obs, nc, half, param, linked, 1);
    double wf[16];
    int chosen[16], count = 0;
    for (int r = 0; r < kk && count < 16; r++) {
      if (!linked[r]) continue;
      double d = sqrt(d2[r]);
      if (d < 1e-8) d = 1e-8;
      wf[count] = 1. / (d * d);
      chosen[count++] = cand[r];
    }
    const double total = numpy_pairwise(wf, count);
    for (int c = 0; c < 16; c++) {
      ids[16 * i + c] = c < count ? chosen[c] : 0;
      w[16 * i + c] = c < count ? (float)(wf[c] / total) : 0.f;
    }
    valid[i] = (uint8_t)(count > 0);
  }
  free(lstates);
  free(feat);
  return fallbacks;
}

/* Multilinear stencil on a uniform grid of d <= 8 axes (the 5D reader's
 * stencil, reversible/relative_table.py:76-109, and the dense 8D game's
 * interpolation_stencil, reversible/world_table.py:13-45): per axis the
 * lower node and fraction (periodic axes: numpy remainder of q - lo by the
 * period, over the step; others: clipped to [lo, hi], lower node at most
 * n - 2, fraction clipped to [0, 1], validity within slack), then the 2^d
 * vertices in itertools.product order with weights multiplied from axis 0. */
void pe_oracle_grid_stencil(int d, const int *n, const double *lo, const double *hi,
                            const double *step, const double *period, unsigned periodic,
                            double slack, const double *states, int64_t m, int64_t *index,
                            double *weights, uint8_t *valid) {
  long long stride[8];
  stride[d - 1] = 1;
  for (int k = d - 2; k >= 0; k--) stride[k] = stride[k + 1] * n[k + 1];
  const int corners = 1 << d;
  for (int64_t i = 0; i < m; i++) {
    long long node[8][2];
    double fr[8];
    int ok = 1;
    for (int k = 0; k < d; k++) {
      double x = states[(size_t)d * i + k];
      if (!isfinite(x)) {
        ok = 0;
        x = 0.;
      }
      if ((periodic >> k) & 1u) {
        const double f = numpy_remainder(x - lo[k], period[k]) / step[k];
        const double fl = floor(f);
        long long j = (long long)fl % n[k];
        if (j < 0) j += n[k];
        node[k][0] = j;
        node[k][1] = (j + 1) % n[k];
        fr[k] = f - fl;
      } else {
        if (!(x >= lo[k] - slack && x <= hi[k] + slack)) ok = 0;
        double z = x < lo[k] ? lo[k] : x;
        if (z > hi[k]) z = hi[k];
        const double f = (z - lo[k]) / step[k];
        long long j = (long long)floor(f);
        if (j > n[k] - 2) j = n[k] - 2;
        double t = f - (double)j;
        t = t < 0. ? 0. : t > 1. ? 1. : t;
        node[k][0] = j;
        node[k][1] = j + 1;
        fr[k] = t;
      }
    }
    valid[i] = (uint8_t)ok;
    for (int c = 0; c < corners; c++) {
      long long at = 0;
      double w = 1.;
      for (int k = 0; k < d; k++) {
        const int up = (c >> (d - 1 - k)) & 1;
        at += node[k][up] * stride[k];
        w = w * (up ? fr[k] : 1. - fr[k]);
      }
      index[(size_t)corners * i + c] = at;
      weights[(size_t)corners * i + c] = w;
    }
  }
}
