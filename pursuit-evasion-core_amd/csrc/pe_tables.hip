// Successor tables of one car grid on the GPU: for each held-command
// successor state, the 16-vertex multilinear stencil over the car's legal
// grid states, with illegal vertices dropped and the rest renormalised, and
// a nearest-supported fallback for successors whose 16 vertices are all
// illegal.
//
// Replaces the host numpy stencil of local_game/grid_game.py:81-123
// (VehicleGrid.stencil) and the cKDTree fallback :38-79
// (_nearest_supported) that make_grid (:158-196) runs on every successor.
//
// Arithmetic follows the reference's float64 formulas operation by operation
// (this file is compiled without contraction), so the tables are the
// reference's bit for bit where its numpy arithmetic is defined:
//   heading cell  f = ((h + pi) mod 2 pi) / (2 pi / n), numpy's remainder;
//   other axes    j = searchsorted(axis, clip(x), 'right') - 1 clipped to
//                 [0, n - 2], frac = (z - axis[j]) / (axis[j + 1] - axis[j]);
//   weight of vertex c = product over the axes, axis 0 first, of frac or
//                 1 - frac (axis k is bit 3 - k of c);
//   retained mass = the 16 weights of the legal vertices summed in numpy's
//                 pairwise order for 16 terms, (w[i] + w[i + 8]) in pairs
//                 of pairs;
//   fallback      the k = min(64, n) legal states nearest in the scaled
//                 features (x / dx, y / dx, cos h / dh, sin h / dh, v / dv),
//                 squared distance summed over the features in order, ties
//                 by state index; the first 16 whose straight segment from
//                 the successor is clear; weights 1 / max(d, 1e-8)^2 over
//                 their numpy-pairwise sum.
// The fallback's tie order is this file's own: cKDTree leaves the order of
// equidistant neighbours unspecified.
//
// Layout: one thread per successor for the stencil; one 256-thread block
// per fallback successor (64 rounds of a block-wide lexicographic minimum,
// then the segment tests four at a time, one per wave).
#include <algorithm>
#include <cmath>
#include <vector>

#include "geometry_device.h"
#include "pe_common.h"

namespace pe {

struct TabAxes {
  int n[4];
  long long stride[4];  // dense strides, axis 0 slowest
  const double *axis[4];
  double heading_step;  // (2 pi) / n_heading
};

constexpr double kPi = 3.141592653589793;
constexpr double kTwoPi = 2 * kPi;

// numpy's pairwise sum of count <= 16 values (count < 8: left to right;
// otherwise eight partial sums, one more block of eight, then the rest).
__device__ __forceinline__ double numpy_sum16(const double *a, int count) {
  if (count < 8) {
    double r = 0;
    for (int i = 0; i < count; i++) r += a[i];
    return r;
  }
  double r[8];
  for (int j = 0; j < 8; j++) r[j] = a[j];
  int i = 8;
  if (count == 16) {
    for (int j = 0; j < 8; j++) r[j] += a[8 + j];
    i = 16;
  }
  double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
  for (; i < count; i++) res += a[i];
  return res;
}

// Lower vertex j and upper-vertex fraction of coordinate x on axis k.
__device__ __forceinline__ void axis_cell(const TabAxes &g, int k, double x, int *j_out,
                                          int *j1_out, double *frac, bool *valid) {
  const int n = g.n[k];
  if (k == 2) {
    if (!isfinite(x)) {
      *valid = false;
      *j_out = 0;
      *j1_out = n > 1 ? 1 : 0;
      *frac = NAN;
      return;
    }
    double r = fmod(x + kPi, kTwoPi);
    if (r < 0) r += kTwoPi;  // numpy: the remainder takes the divisor's sign
    const double f = r / g.heading_step;
    const double fl = floor(f);
    const int j = (int)((long long)fl % n);
    *j_out = j;
    *j1_out = (j + 1) % n;
    *frac = f - fl;
    return;
  }
  const double *ax = g.axis[k];
  if (isnan(x)) {
    *valid = false;
    *j_out = n - 2;
    *j1_out = n - 1;
    *frac = NAN;
    return;
  }
  if (!(x >= ax[0] - 1e-9 && x <= ax[n - 1] + 1e-9)) *valid = false;
  const double z = fmin(fmax(x, ax[0]), ax[n - 1]);
  int lo = 0, hi = n;  // count of axis values <= z
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (ax[mid] <= z) lo = mid + 1;
    else hi = mid;
  }
  int j = lo - 1;
  j = j < 0 ? 0 : j > n - 2 ? n - 2 : j;
  *j_out = j;
  *j1_out = j + 1;
  *frac = (z - ax[j]) / (ax[j + 1] - ax[j]);
}

__global__ __launch_bounds__(256) void tab_stencil(TabAxes g, const int *__restrict__ d2c,
                                                   const double *__restrict__ next, long long m,
                                                   int *__restrict__ ids, float *__restrict__ wts,
                                                   uint8_t *__restrict__ valid_out,
                                                   double *__restrict__ mass_out,
                                                   int *__restrict__ missing,
                                                   unsigned long long *__restrict__ n_missing) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const double *q = next + 4 * i;
  bool valid = true;
  bool finite = true;
  long long lower[4], upper[4];
  double frac[4];
  for (int k = 0; k < 4; k++) {
    finite = finite && isfinite(q[k]);
    int j, j1;
    axis_cell(g, k, q[k], &j, &j1, &frac[k], &valid);
    lower[k] = j * g.stride[k];
    upper[k] = j1 * g.stride[k];
  }
  valid = valid && finite;
  double w[16];
  int id[16];
  for (int c = 0; c < 16; c++) {
    long long dense = 0;
    double wc = 1.0;
    for (int k = 0; k < 4; k++) {
      const bool up = (c >> (3 - k)) & 1;
      dense += up ? upper[k] : lower[k];
      wc *= up ? frac[k] : 1.0 - frac[k];
    }
    id[c] = d2c[dense];
    w[c] = id[c] >= 0 ? wc : 0.0;
  }
  const double mass = numpy_sum16(w, 16);
  const bool kept = mass > 1e-9;
  const bool fallback = valid && mass <= 1e-9;
  int *row_ids = ids + 16 * i;
  float *row_w = wts + 16 * i;
  for (int c = 0; c < 16; c++) {
    row_ids[c] = id[c] > 0 ? id[c] : 0;
    row_w[c] = kept ? (float)(w[c] / mass) : 0.0f;
  }
  valid_out[i] = (valid && kept) ? 1 : 0;
  mass_out[i] = mass;
  if (fallback) missing[atomicAdd(n_missing, 1ull)] = (int)i;
}

// Scaled features of the n legal states (compact order): x / dx, y / dx,
// cos h / dh, sin h / dh, v / dv, from each state's dense grid index.
__global__ __launch_bounds__(256) void tab_features(TabAxes g, const long long *__restrict__ dense,
                                                    int n, double dx, double dh, double dv,
                                                    double *__restrict__ states,
                                                    double *__restrict__ feat) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  long long r = dense[c];
  double s[4];
  for (int k = 0; k < 4; k++) {
    const long long j = r / g.stride[k];
    r -= j * g.stride[k];
    s[k] = g.axis[k][j];
  }
  for (int k = 0; k < 4; k++) states[4 * (long long)c + k] = s[k];
  double *f = feat + 5 * (long long)c;
  f[0] = s[0] / dx;
  f[1] = s[1] / dx;
  f[2] = cos(s[2]) / dh;
  f[3] = sin(s[2]) / dh;
  f[4] = s[3] / dv;
}

// (d, i) < (e, j) lexicographically.
__device__ __forceinline__ bool key_less(double d, int i, double e, int j) {
  return d < e || (d == e && i < j);
}

__global__ __launch_bounds__(256) void tab_nearest(const int *__restrict__ missing, int count,
                                                   const double *__restrict__ next,
                                                   const double *__restrict__ states,
                                                   const double *__restrict__ feat, int n, int k,
                                                   double dx, double dh, double dv,
                                                   const double *__restrict__ obs, int nc,
                                                   double half, Car m, int *__restrict__ ids,
                                                   float *__restrict__ wts,
                                                   uint8_t *__restrict__ valid_out) {
  __shared__ double cand_d2[64];
  __shared__ int cand_i[64];
  __shared__ double red_d[4];
  __shared__ int red_i[4];
  __shared__ uint8_t linked[64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int item = blockIdx.x; item < count; item += gridDim.x) {
    const long long row = missing[item];
    const double *q = next + 4 * row;
    const double qf[5] = {q[0] / dx, q[1] / dx, cos(q[2]) / dh, sin(q[2]) / dh, q[3] / dv};
    double last_d = -1.0;
    int last_i = -1;
    for (int r = 0; r < k; r++) {
      double best_d = INFINITY;
      int best_i = 0x7fffffff;
      for (int c = threadIdx.x; c < n; c += blockDim.x) {
        const double *f = feat + 5 * (long long)c;
        double d2 = 0.0;
        for (int e = 0; e < 5; e++) {
          const double t = qf[e] - f[e];
          d2 = d2 + t * t;
        }
        if (key_less(last_d, last_i, d2, c) && key_less(d2, c, best_d, best_i)) {
          best_d = d2;
          best_i = c;
        }
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        const double od = __shfl_xor(best_d, off, 64);
        const int oi = __shfl_xor(best_i, off, 64);
        if (key_less(od, oi, best_d, best_i)) {
          best_d = od;
          best_i = oi;
        }
      }
      if (lane == 0) {
        red_d[wave] = best_d;
        red_i[wave] = best_i;
      }
      __syncthreads();
      double bd = red_d[0];
      int bi = red_i[0];
      for (int w = 1; w < 4; w++)
        if (key_less(red_d[w], red_i[w], bd, bi)) {
          bd = red_d[w];
          bi = red_i[w];
        }
      if (threadIdx.x == 0) {
        cand_d2[r] = bd;
        cand_i[r] = bi;
      }
      last_d = bd;
      last_i = bi;
      __syncthreads();
    }
    // Segment tests in candidate order, one candidate per wave, until the
    // tested prefix holds 16 connected states.
    int found = 0, tested = 0;
    for (int base = 0; base < k && found < 16; base += 4) {
      const int c = base + wave;
      if (c < k) {
        const bool ok = segment_clear(q, states + 4 * (long long)cand_i[c], obs, nc, half, m);
        if (lane == 0) linked[c] = ok ? 1 : 0;
      }
      __syncthreads();
      tested = std::min(base + 4, k);
      found = 0;
      for (int t = 0; t < tested; t++) found += linked[t];
    }
    if (threadIdx.x == 0) {
      double dist[16], w[16];
      int chosen[16], count_sel = 0;
      for (int t = 0; t < tested && count_sel < 16; t++) {
        if (linked[t]) {
          chosen[count_sel] = cand_i[t];
          dist[count_sel] = sqrt(cand_d2[t]);
          count_sel++;
        }
      }
      for (int s = 0; s < count_sel; s++) {
        const double d = fmax(dist[s], 1e-8);
        w[s] = 1.0 / (d * d);
      }
      const double total = numpy_sum16(w, count_sel);
      int *row_ids = ids + 16 * row;
      float *row_w = wts + 16 * row;
      for (int s = 0; s < 16; s++) {
        row_ids[s] = s < count_sel ? chosen[s] : 0;
        row_w[s] = s < count_sel ? (float)(w[s] / total) : 0.0f;
      }
      valid_out[row] = count_sel > 0 ? 1 : 0;
    }
    __syncthreads();
  }
}

}  // namespace pe

using namespace pe;

extern "C" int pe_car_stencil(const pe_car_axes *axes, const int32_t *dense_to_compact,
                              const double *next, int64_t m, const double *obstacles,
                              int32_t nc, double half, const double *param, int32_t *ids,
                              float *weights, uint8_t *valid, double *mass,
                              int64_t *fallbacks) {
  clear_error();
  if (!axes || !dense_to_compact || m < 0 || nc < 0 ||
      (m > 0 && (!next || !ids || !weights || !valid || !mass)) ||
      (param && nc > 0 && !obstacles))
    return fail(PE_ERR_INVALID, "invalid argument");
  long long dense_n = 1;
  for (int k = 0; k < 4; k++) {
    if (axes->n[k] < (k == 2 ? 1 : 2) || !axes->axis[k])
      return fail(PE_ERR_INVALID, "Expected four axes (x, y, heading, speed)");
    for (int j = 1; j < axes->n[k]; j++)
      if (!(axes->axis[k][j] > axes->axis[k][j - 1]))
        return fail(PE_ERR_INVALID, "axes must be strictly increasing");
    dense_n *= axes->n[k];
  }
  if (param) {
    for (int k = 0; k < 7; k++)
      if (!std::isfinite(param[k]))
        return fail(PE_ERR_INVALID, "non-finite vehicle parameters");
    if (!std::isfinite(half) || half <= 0)
      return fail(PE_ERR_INVALID, "invalid room half-width");
  }
  // Legal states in compact order, from the dense -> compact map.
  int n = 0;
  for (long long d = 0; d < dense_n; d++)
    if (dense_to_compact[d] >= 0) n++;
  std::vector<long long> dense_of(n, -1);
  for (long long d = 0; d < dense_n; d++) {
    const int c = dense_to_compact[d];
    if (c >= n) return fail(PE_ERR_INVALID, "dense_to_compact is not a compact numbering");
    if (c >= 0) {
      if (dense_of[c] >= 0) return fail(PE_ERR_INVALID, "dense_to_compact repeats an index");
      dense_of[c] = d;
    }
  }
  if (n == 0) return fail(PE_ERR_INVALID, "the grid has no legal state");
  if (fallbacks) *fallbacks = 0;
  int rc = require_device(nullptr);
  if (rc) return rc;
  if (m == 0) return PE_OK;

  std::vector<void *> owned;
  auto release = [&]() {
    for (void *p : owned) (void)hipFree(p);
  };
  auto alloc = [&](void **p, size_t bytes) -> int {
    PE_HIP(hipMalloc(p, bytes ? bytes : 8));
    owned.push_back(*p);
    return PE_OK;
  };
  auto body = [&]() -> int {
    TabAxes g{};
    long long stride = 1;
    for (int k = 3; k >= 0; k--) {
      g.n[k] = axes->n[k];
      g.stride[k] = stride;
      stride *= axes->n[k];
    }
    g.heading_step = kTwoPi / axes->n[2];
    void *d_axes, *d_d2c, *d_next, *d_ids, *d_w, *d_valid, *d_mass, *d_missing, *d_count;
    size_t axis_total = 0;
    for (int k = 0; k < 4; k++) axis_total += axes->n[k];
    int r;
    if ((r = alloc(&d_axes, 8 * axis_total)) || (r = alloc(&d_d2c, 4 * dense_n)) ||
        (r = alloc(&d_next, 32 * (size_t)m)) || (r = alloc(&d_ids, 64 * (size_t)m)) ||
        (r = alloc(&d_w, 64 * (size_t)m)) || (r = alloc(&d_valid, (size_t)m)) ||
        (r = alloc(&d_mass, 8 * (size_t)m)) || (r = alloc(&d_missing, 4 * (size_t)m)) ||
        (r = alloc(&d_count, 8)))
      return r;
    size_t off = 0;
    for (int k = 0; k < 4; k++) {
      PE_HIP(hipMemcpy((double *)d_axes + off, axes->axis[k], 8 * axes->n[k],
                       hipMemcpyHostToDevice));
      g.axis[k] = (const double *)d_axes + off;
      off += axes->n[k];
    }
    PE_HIP(hipMemcpy(d_d2c, dense_to_compact, 4 * dense_n, hipMemcpyHostToDevice));
    PE_HIP(hipMemcpy(d_next, next, 32 * (size_t)m, hipMemcpyHostToDevice));
    PE_HIP(hipMemset(d_count, 0, 8));
    hipLaunchKernelGGL(tab_stencil, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, 0, g,
                       (const int *)d_d2c, (const double *)d_next, (long long)m, (int *)d_ids,
                       (float *)d_w, (uint8_t *)d_valid, (double *)d_mass, (int *)d_missing,
                       (unsigned long long *)d_count);
    PE_LAUNCHED(tab_stencil);
    unsigned long long count = 0;
    PE_HIP(hipMemcpy(&count, d_count, 8, hipMemcpyDeviceToHost));
    if (count > 0 && param) {
      const double dx = axes->axis[0][1] - axes->axis[0][0];
      const double dh = 2 * std::sin(kPi / axes->n[2]);
      double dv = INFINITY;
      for (int j = 1; j < axes->n[3]; j++)
        dv = std::min(dv, axes->axis[3][j] - axes->axis[3][j - 1]);
      void *d_dense, *d_states, *d_feat, *d_obs;
      if ((r = alloc(&d_dense, 8 * (size_t)n)) || (r = alloc(&d_states, 32 * (size_t)n)) ||
          (r = alloc(&d_feat, 40 * (size_t)n)) || (r = alloc(&d_obs, 24 * (size_t)nc)))
        return r;
      PE_HIP(hipMemcpy(d_dense, dense_of.data(), 8 * (size_t)n, hipMemcpyHostToDevice));
      if (nc) PE_HIP(hipMemcpy(d_obs, obstacles, 24 * (size_t)nc, hipMemcpyHostToDevice));
      hipLaunchKernelGGL(tab_features, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, g,
                         (const long long *)d_dense, n, dx, dh, dv, (double *)d_states,
                         (double *)d_feat);
      PE_LAUNCHED(tab_features);
      const int k = n < 64 ? n : 64;
      const unsigned grid = (unsigned)std::min<unsigned long long>(count, 4096);
      Car car{param[0], param[1], param[2], param[3], param[4], param[5], param[6]};
      hipLaunchKernelGGL(tab_nearest, dim3(grid), dim3(256), 0, 0, (const int *)d_missing,
                         (int)count, (const double *)d_next, (const double *)d_states,
                         (const double *)d_feat, n, k, dx, dh, dv, (const double *)d_obs, nc,
                         half, car, (int *)d_ids, (float *)d_w, (uint8_t *)d_valid);
      PE_LAUNCHED(tab_nearest);
    }
    PE_HIP(hipMemcpy(ids, d_ids, 64 * (size_t)m, hipMemcpyDeviceToHost));
    PE_HIP(hipMemcpy(weights, d_w, 64 * (size_t)m, hipMemcpyDeviceToHost));
    PE_HIP(hipMemcpy(valid, d_valid, (size_t)m, hipMemcpyDeviceToHost));
    PE_HIP(hipMemcpy(mass, d_mass, 8 * (size_t)m, hipMemcpyDeviceToHost));
    if (fallbacks) *fallbacks = (int64_t)count;
    return PE_OK;
  };
  rc = body();
  release();
  return rc;
}

// ---------------------------------------------------------------------------
// Multilinear stencils on a uniform grid of up to 8 axes (the 5D reader's
// RelativeTable.stencil, reversible/relative_table.py:76-109, and the probe
// stencils of the dense 8D game, reversible/world_table.py:13-45).  Axis k:
// n[k] nodes lo[k] + j step[k].  Periodic axes: f = ((q - lo) mod period) /
// step, numpy's remainder, lower node floor(f) mod n, the upper node wraps.
// Other axes: valid within [lo, hi] +- slack, f = (clip(q, lo, hi) - lo) /
// step, lower node min(floor(f), n - 2), fraction clipped to [0, 1].
// Vertex c takes the upper node on axis k when bit d - 1 - k of c is set
// (itertools.product order); its weight is the product over the axes, axis
// 0 first.  Non-finite coordinates count as 0 and make the state invalid.
namespace pe {

struct UniformGrid {
  int d;
  int n[8];
  long long stride[8];
  double lo[8], hi[8], step[8], period[8];
  unsigned periodic;
  double slack;
};

__global__ __launch_bounds__(128) void grid_stencil(UniformGrid g, const double *__restrict__ q,
                                                    long long m, long long *__restrict__ index,
                                                    double *__restrict__ weight,
                                                    uint8_t *__restrict__ valid) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  long long low[8], high[8];
  double frac[8];
  bool ok = true;
  for (int k = 0; k < g.d; k++) {
    double x = q[(long long)g.d * i + k];
    if (!isfinite(x)) {
      ok = false;
      x = 0.0;
    }
    const int n = g.n[k];
    long long j, j1;
    double t;
    if ((g.periodic >> k) & 1u) {
      double r = fmod(x - g.lo[k], g.period[k]);
      if (r != 0.0 && r < 0.0) r += g.period[k];
      const double f = r / g.step[k];
      const double fl = floor(f);
      j = ((long long)fl % n + n) % n;
      j1 = (j + 1) % n;
      t = f - fl;
    } else {
      ok = ok && x >= g.lo[k] - g.slack && x <= g.hi[k] + g.slack;
      const double f = (fmin(fmax(x, g.lo[k]), g.hi[k]) - g.lo[k]) / g.step[k];
      j = (long long)floor(f);
      if (j > n - 2) j = n - 2;
      j1 = j + 1;
      t = fmin(fmax(f - (double)j, 0.0), 1.0);
    }
    low[k] = j * g.stride[k];
    high[k] = j1 * g.stride[k];
    frac[k] = t;
  }
  const int corners = 1 << g.d;
  long long *out_i = index + (long long)corners * i;
  double *out_w = weight + (long long)corners * i;
  for (int c = 0; c < corners; c++) {
    long long idx = 0;
    double w = 1.0;
    for (int k = 0; k < g.d; k++) {
      const bool up = (c >> (g.d - 1 - k)) & 1;
      idx += up ? high[k] : low[k];
      w *= up ? frac[k] : 1.0 - frac[k];
    }
    out_i[c] = idx;
    out_w[c] = w;
  }
  valid[i] = ok ? 1 : 0;
}

}  // namespace pe

extern "C" int pe_grid_stencil(int32_t d, const int32_t *n, const double *lo, const double *hi,
                               const double *step, const double *period, uint32_t periodic,
                               double slack, const double *states, int64_t m, int64_t *index,
                               double *weights, uint8_t *valid) {
  clear_error();
  if (d < 1 || d > 8 || !n || !lo || !hi || !step || m < 0 ||
      (periodic && !period) || (m > 0 && (!states || !index || !weights || !valid)) ||
      !(slack >= 0) || (periodic >> d) != 0)
    return fail(PE_ERR_INVALID, "invalid argument");
  UniformGrid g{};
  g.d = d;
  g.periodic = periodic;
  g.slack = slack;
  long long stride = 1;
  for (int k = d - 1; k >= 0; k--) {
    const bool wraps = (periodic >> k) & 1u;
    if (n[k] < (wraps ? 1 : 2) || !(step[k] > 0) || !std::isfinite(lo[k]) ||
        !std::isfinite(hi[k]) || !(hi[k] >= lo[k]) || (wraps && !(period[k] > 0)))
      return fail(PE_ERR_INVALID, "Expected uniformly spaced, nondegenerate axes");
    g.n[k] = n[k];
    g.stride[k] = stride;
    stride *= n[k];
    g.lo[k] = lo[k];
    g.hi[k] = hi[k];
    g.step[k] = step[k];
    g.period[k] = wraps ? period[k] : 0.0;
  }
  int rc = require_device(nullptr);
  if (rc) return rc;
  if (m == 0) return PE_OK;
  const long long corners = 1ll << d;
  void *d_q = nullptr, *d_i = nullptr, *d_w = nullptr, *d_v = nullptr;
  auto body = [&]() -> int {
    PE_HIP(hipMalloc(&d_q, 8 * (size_t)d * m));
    PE_HIP(hipMalloc(&d_i, 8 * (size_t)corners * m));
    PE_HIP(hipMalloc(&d_w, 8 * (size_t)corners * m));
    PE_HIP(hipMalloc(&d_v, (size_t)m));
    PE_HIP(hipMemcpy(d_q, states, 8 * (size_t)d * m, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(grid_stencil, dim3((unsigned)((m + 127) / 128)), dim3(128), 0, 0, g,
                       (const double *)d_q, (long long)m, (long long *)d_i, (double *)d_w,
                       (uint8_t *)d_v);
    PE_LAUNCHED(grid_stencil);
    PE_HIP(hipMemcpy(index, d_i, 8 * (size_t)corners * m, hipMemcpyDeviceToHost));
    PE_HIP(hipMemcpy(weights, d_w, 8 * (size_t)corners * m, hipMemcpyDeviceToHost));
    PE_HIP(hipMemcpy(valid, d_v, (size_t)m, hipMemcpyDeviceToHost));
    return PE_OK;
  };
  rc = body();
  for (void *p : {d_q, d_i, d_w, d_v})
    if (p) (void)hipFree(p);
  return rc;
}
