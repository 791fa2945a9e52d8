// Per-vehicle geometry of the contact game on the GPU, in double precision:
// clearance against square walls and circular obstacles, exact held-command
// arcs, certified swept-arc safety, braking-backup safety, batched action
// tables and configuration-space connections.
//
// Replaces the OpenMP C++ library local_game/native_geometry.cpp (functions
// batch_clearance :92-99, batch_backups :100-107, batch_actions :108-123,
// configuration_connections :237-252) that local_game/native.py binds.  The
// entry points keep its argument meaning (states [n,4] = x, y, heading,
// signed speed; obstacles [nc,3] = x, y, r; param = [low, high, a, k, L, W,
// dt]) and return a status instead of taking a thread count.
//
// Clearance queries run one thread per state.  The certified-motion
// queries (backups, actions, connections) run one wave per item: the
// reference's recursive Lipschitz bisection becomes a depth-first loop over
// an explicit stack, visiting the same intervals in the same order, and
// every clearance evaluation in it is split over the wave's 64 lanes (each
// lane a strided share of the circles, then a wave minimum).  The minimum
// does not depend on the order the circles are visited in, so every
// decision matches the serial walk.
#include <cmath>
#include <initializer_list>
#include <vector>

#include "pe_common.h"

namespace pe {

struct Car {
  double low, high, a, k, L, W, dt;
};

__device__ __forceinline__ void held_step(const double *s, double a, double k,
                                          double t, const Car &m, double *out) {
  const double bound = a >= 0 ? m.high : m.low;
  const double reach = fabs(a) > 1e-14 ? (bound - s[3]) / a : INFINITY;
  const double held = fmin(t, fmax(0., reach));
  const double v = fmax(m.low, fmin(m.high, s[3] + a * t));
  const double dist = s[3] * held + .5 * a * held * held + v * (t - held);
  const double turn = k * dist;
  const double half = turn * .5;
  const double sinc = fabs(half) > 1e-12 ? sin(half) / half : 1.;
  out[0] = s[0] + dist * sinc * cos(s[2] + half);
  out[1] = s[1] + dist * sinc * sin(s[2] + half);
  out[2] = remainder(s[2] + turn, 2 * M_PI);
  out[3] = v;
}

// WAVE: the calling wave's lanes share the circles and all receive the
// minimum.
template <bool WAVE = false>
__device__ double clearance(const double *s, const double *__restrict__ obs,
                            int nc, double half, const Car &m) {
  const double c = cos(s[2]), q = sin(s[2]);
  const double ex = .5 * (m.L * fabs(c) + m.W * fabs(q));
  const double ey = .5 * (m.L * fabs(q) + m.W * fabs(c));
  double best = fmin(half - fabs(s[0]) - ex, half - fabs(s[1]) - ey);
  const double reach = .5 * hypot(m.L, m.W);
  const int first = WAVE ? (int)(threadIdx.x & 63) : 0, stride = WAVE ? 64 : 1;
  for (int j = first; j < nc; j += stride) {
    const double dx = obs[3 * j] - s[0], dy = obs[3 * j + 1] - s[1];
    const double r = obs[3 * j + 2];
    const double far = reach + r + fmax(0., best);
    if (dx * dx + dy * dy > far * far) continue;  // cannot lower the minimum
    const double lx = fabs(dx * c + dy * q) - m.L * .5;
    const double ly = fabs(-dx * q + dy * c) - m.W * .5;
    best = fmin(best, hypot(fmax(0., lx), fmax(0., ly)) + fmin(fmax(lx, ly), 0.) - r);
  }
  if (WAVE) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) best = fmin(best, __shfl_xor(best, off, 64));
  }
  return best;
}

struct Piece {
  double lo, hi, cl, cr;
  int depth;
};

// All pieces of [0, t] along the held arc keep clearance > 1e-7 and are
// certified by the Lipschitz bound lip within bisection depth 18.
template <bool WAVE>
__device__ bool arc_clear(const double *s, double a, double k, double t,
                          double lip, double c0, double c1,
                          const double *__restrict__ obs, int nc, double half,
                          const Car &m) {
  Piece stack[20];
  int top = 0;
  stack[top++] = {0., t, c0, c1, 0};
  while (top > 0) {
    const Piece p = stack[--top];
    const double low = fmin(p.cl, p.cr);
    if (low <= 1e-7) return false;
    if (low > lip * (p.hi - p.lo) * .5 + 1e-7) continue;
    if (p.depth >= 18 || p.hi - p.lo < 1e-8) return false;
    const double mid = .5 * (p.lo + p.hi);
    double z[4];
    held_step(s, a, k, mid, m, z);
    const double cm = clearance<WAVE>(z, obs, nc, half, m);
    stack[top++] = {mid, p.hi, cm, p.cr, p.depth + 1};  // visited second
    stack[top++] = {p.lo, mid, p.cl, cm, p.depth + 1};  // visited first
  }
  return true;
}

template <bool WAVE>
__device__ bool motion_safe(const double *s, double a, double k, double t,
                            const double *__restrict__ obs, int nc, double half,
                            const Car &m) {
  double end[4];
  held_step(s, a, k, t, m, end);
  const double lip =
      fmax(fabs(s[3]), fabs(end[3])) * (1 + .5 * hypot(m.L, m.W) * fabs(k));
  return arc_clear<WAVE>(s, a, k, t, lip, clearance<WAVE>(s, obs, nc, half, m),
                         clearance<WAVE>(end, obs, nc, half, m), obs, nc, half, m);
}

// Held-command braking to rest along one of k in {0, -K, +K}
// (native_geometry.cpp:66-88): full-deceleration periods, then a final
// period with a = -v/dt; the same arc is swept by uniform deceleration.
template <bool WAVE>
__device__ bool backup_safe(const double *s, const double *__restrict__ obs,
                            int nc, double half, const Car &m) {
  if (clearance<WAVE>(s, obs, nc, half, m) <= 1e-7) return false;
  if (fabs(s[3]) < 1e-10) return true;
  const double speed = fabs(s[3]);
  int periods = (int)ceil(speed / (m.a * m.dt) - 1e-10);
  if (periods < 1) periods = 1;
  const double full = (periods - 1) * m.dt;
  const double last = speed - m.a * full;
  const double dist = speed * full - .5 * m.a * full * full + .5 * last * m.dt;
  const double t = 2 * dist / speed, a = -s[3] / t;
  const double curv[3] = {0., -m.k, m.k};
  for (int j = 0; j < 3; j++)
    if (motion_safe<WAVE>(s, a, curv[j], t, obs, nc, half, m)) return true;
  return false;
}

__global__ __launch_bounds__(128) void geom_clearance(
    const double *__restrict__ states, long long n,
    const double *__restrict__ obs, int nc, double half, Car m, double *out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s[4] = {states[4 * i], states[4 * i + 1], states[4 * i + 2],
                 states[4 * i + 3]};
  out[i] = clearance(s, obs, nc, half, m);
}

// Item of the calling wave (4 waves per 256-thread block).
__device__ __forceinline__ long long wave_item() {
  return (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
}

__global__ __launch_bounds__(256) void geom_backups(
    const double *__restrict__ states, long long n,
    const double *__restrict__ obs, int nc, double half, Car m, uint8_t *out) {
  const long long i = wave_item();
  if (i >= n) return;  // uniform per wave
  double s[4] = {states[4 * i], states[4 * i + 1], states[4 * i + 2],
                 states[4 * i + 3]};
  const bool good = backup_safe<true>(s, obs, nc, half, m);
  if ((threadIdx.x & 63) == 0) out[i] = good ? 1 : 0;
}

__global__ __launch_bounds__(256) void geom_actions(
    const double *__restrict__ states, long long n,
    const double *__restrict__ controls, int na, double dt,
    const double *__restrict__ obs, int nc, double half, Car m, uint8_t *ok,
    double *next) {
  const long long t = wave_item();
  if (t >= n * na) return;  // uniform per wave
  const long long i = t / na;
  const int u = (int)(t - i * na);
  double s[4] = {states[4 * i], states[4 * i + 1], states[4 * i + 2],
                 states[4 * i + 3]};
  const double a = controls[2 * u], k = controls[2 * u + 1];
  double z[4];
  held_step(s, a, k, dt, m, z);
  const bool good = motion_safe<true>(s, a, k, dt, obs, nc, half, m) &&
                    backup_safe<true>(z, obs, nc, half, m);
  if ((threadIdx.x & 63) == 0) {
    for (int c = 0; c < 4; c++) next[4 * t + c] = z[c];
    ok[t] = good ? 1 : 0;
  }
}

// The same two queries for the states of many maps at once (one launch for
// all of them): item i belongs to map map_of[i], whose circles are
// obs[obs_off[b] .. obs_off[b + 1]) and whose square has half-width
// halves[b].  Each wave runs the single-map code on its own map's data.
__global__ __launch_bounds__(256) void geom_backups_multi(
    const double *__restrict__ states, long long n, const int *__restrict__ map_of,
    const double *__restrict__ obs, const long long *__restrict__ obs_off,
    const double *__restrict__ halves, Car m, uint8_t *out) {
  const long long i = wave_item();
  if (i >= n) return;  // uniform per wave
  const int b = map_of[i];
  double s[4] = {states[4 * i], states[4 * i + 1], states[4 * i + 2],
                 states[4 * i + 3]};
  const bool good = backup_safe<true>(s, obs + 3 * obs_off[b], (int)(obs_off[b + 1] - obs_off[b]),
                                      halves[b], m);
  if ((threadIdx.x & 63) == 0) out[i] = good ? 1 : 0;
}

__global__ __launch_bounds__(256) void geom_actions_multi(
    const double *__restrict__ states, long long n, const int *__restrict__ map_of,
    const double *__restrict__ controls, int na, double dt,
    const double *__restrict__ obs_all, const long long *__restrict__ obs_off,
    const double *__restrict__ halves, Car m, uint8_t *ok, double *next) {
  const long long t = wave_item();
  if (t >= n * na) return;  // uniform per wave
  const long long i = t / na;
  const int u = (int)(t - i * na);
  const int b = map_of[i];
  const double *obs = obs_all + 3 * obs_off[b];
  const int nc = (int)(obs_off[b + 1] - obs_off[b]);
  const double half = halves[b];
  double s[4] = {states[4 * i], states[4 * i + 1], states[4 * i + 2],
                 states[4 * i + 3]};
  const double a = controls[2 * u], k = controls[2 * u + 1];
  double z[4];
  held_step(s, a, k, dt, m, z);
  const bool good = motion_safe<true>(s, a, k, dt, obs, nc, half, m) &&
                    backup_safe<true>(z, obs, nc, half, m);
  if ((threadIdx.x & 63) == 0) {
    for (int c = 0; c < 4; c++) next[4 * t + c] = z[c];
    ok[t] = good ? 1 : 0;
  }
}

// Straight configuration-space segment a -> b (heading by the short way)
// stays clear, certified by bisection to depth 20 (native_geometry.cpp:213-252).
__global__ __launch_bounds__(256) void geom_connections(
    const double *__restrict__ from, const double *__restrict__ to,
    long long n, const double *__restrict__ obs, int nc, double half, Car m,
    uint8_t *out) {
  const long long i = wave_item();
  if (i >= n) return;  // uniform per wave
  const double *a = from + 4 * i, *b = to + 4 * i;
  const double turn = remainder(b[2] - a[2], 2 * M_PI);
  const double lip =
      hypot(b[0] - a[0], b[1] - a[1]) + .5 * hypot(m.L, m.W) * fabs(turn);
  auto at = [&](double f) {
    const double s[4] = {a[0] + f * (b[0] - a[0]), a[1] + f * (b[1] - a[1]),
                         a[2] + f * turn, 0.};
    return clearance<true>(s, obs, nc, half, m);
  };
  Piece stack[22];
  int top = 0;
  stack[top++] = {0., 1., at(0.), at(1.), 0};
  bool good = true;
  while (top > 0) {
    const Piece p = stack[--top];
    const double low = fmin(p.cl, p.cr);
    if (low <= 1e-7) { good = false; break; }
    if (low > lip * (p.hi - p.lo) * .5 + 1e-7) continue;
    if (p.depth >= 20) { good = false; break; }
    const double mid = (p.lo + p.hi) * .5;
    const double cm = at(mid);
    stack[top++] = {mid, p.hi, cm, p.cr, p.depth + 1};
    stack[top++] = {p.lo, mid, p.cl, cm, p.depth + 1};
  }
  if ((threadIdx.x & 63) == 0) out[i] = good ? 1 : 0;
}

// Device buffers of one synchronous geometry call.
struct Scratch {
  std::vector<void *> ptrs;
  ~Scratch() {
    for (void *p : ptrs) hipFree(p);
  }
  int alloc(void **p, size_t bytes) {
    PE_HIP(hipMalloc(p, bytes ? bytes : 8));
    ptrs.push_back(*p);
    return PE_OK;
  }
  int upload(void **p, const void *host, size_t bytes) {
    if (bytes && !host) return fail(PE_ERR_INVALID, "null input buffer");
    int rc = alloc(p, bytes);
    if (rc) return rc;
    if (bytes) PE_HIP(hipMemcpy(*p, host, bytes, hipMemcpyHostToDevice));
    return PE_OK;
  }
};

static int check_args(const double *param, double half, int nc, long long n) {
  if (!param || n < 0 || nc < 0 || !std::isfinite(half) || half <= 0)
    return fail(PE_ERR_INVALID, "invalid geometry arguments");
  for (int k = 0; k < 7; k++)
    if (!std::isfinite(param[k]))
      return fail(PE_ERR_INVALID, "non-finite vehicle parameters");
  return PE_OK;
}

// Buffers the caller passes for n > 0 items must not be null.
static int check_buffers(long long n, std::initializer_list<const void *> ptrs) {
  if (n > 0)
    for (const void *p : ptrs)
      if (!p) return fail(PE_ERR_INVALID, "null buffer");
  return PE_OK;
}

static Car read_car(const double *p) {
  return Car{p[0], p[1], p[2], p[3], p[4], p[5], p[6]};
}

static unsigned blocks_for(long long n) {
  return (unsigned)((n + 127) / 128 > 0 ? (n + 127) / 128 : 1);
}

// 256-thread blocks of four one-wave items.
static unsigned wave_blocks_for(long long n) {
  return (unsigned)((n + 3) / 4 > 0 ? (n + 3) / 4 : 1);
}

}  // namespace pe

using namespace pe;

extern "C" int pe_geom_clearance(const double *states, int64_t n,
                                 const double *obstacles, int32_t nc,
                                 double half, const double *param,
                                 double *out) {
  clear_error();
  int rc = check_args(param, half, nc, n);
  if (!rc) rc = check_buffers(n, {states, out});
  if (rc || (rc = require_device(nullptr))) return rc;
  if (n == 0) return PE_OK;
  Scratch s;
  void *d_s, *d_o, *d_out;
  if ((rc = s.upload(&d_s, states, n * 32)) || (rc = s.upload(&d_o, obstacles, nc * 24)) ||
      (rc = s.alloc(&d_out, n * 8)))
    return rc;
  hipLaunchKernelGGL(geom_clearance, dim3(blocks_for(n)), dim3(128), 0, 0,
                     (const double *)d_s, (long long)n, (const double *)d_o, nc,
                     half, read_car(param), (double *)d_out);
  PE_LAUNCHED(geom_clearance);
  PE_HIP(hipMemcpy(out, d_out, n * 8, hipMemcpyDeviceToHost));
  return PE_OK;
}

extern "C" int pe_geom_backups(const double *states, int64_t n,
                               const double *obstacles, int32_t nc,
                               double half, const double *param,
                               uint8_t *out) {
  clear_error();
  int rc = check_args(param, half, nc, n);
  if (!rc) rc = check_buffers(n, {states, out});
  if (rc || (rc = require_device(nullptr))) return rc;
  if (n == 0) return PE_OK;
  Scratch s;
  void *d_s, *d_o, *d_out;
  if ((rc = s.upload(&d_s, states, n * 32)) || (rc = s.upload(&d_o, obstacles, nc * 24)) ||
      (rc = s.alloc(&d_out, n)))
    return rc;
  hipLaunchKernelGGL(geom_backups, dim3(wave_blocks_for(n)), dim3(256), 0, 0,
                     (const double *)d_s, (long long)n, (const double *)d_o, nc,
                     half, read_car(param), (uint8_t *)d_out);
  PE_LAUNCHED(geom_backups);
  PE_HIP(hipMemcpy(out, d_out, n, hipMemcpyDeviceToHost));
  return PE_OK;
}

extern "C" int pe_geom_actions(const double *states, int64_t n,
                               const double *controls, int32_t na, double dt,
                               const double *obstacles, int32_t nc,
                               double half, const double *param, uint8_t *ok,
                               double *next) {
  clear_error();
  int rc = check_args(param, half, nc, n);
  if (!rc) rc = check_buffers(n, {states, controls, ok, next});
  if (rc) return rc;
  if (na < 1 || !(dt >= 0) || !std::isfinite(dt))
    return fail(PE_ERR_INVALID, "invalid controls or duration");
  if ((rc = require_device(nullptr))) return rc;
  if (n == 0) return PE_OK;
  Scratch s;
  void *d_s, *d_u, *d_o, *d_ok, *d_next;
  const long long m = (long long)n * na;
  if ((rc = s.upload(&d_s, states, n * 32)) || (rc = s.upload(&d_u, controls, na * 16)) ||
      (rc = s.upload(&d_o, obstacles, nc * 24)) || (rc = s.alloc(&d_ok, m)) ||
      (rc = s.alloc(&d_next, m * 32)))
    return rc;
  hipLaunchKernelGGL(geom_actions, dim3(wave_blocks_for(m)), dim3(256), 0, 0,
                     (const double *)d_s, (long long)n, (const double *)d_u, na,
                     dt, (const double *)d_o, nc, half, read_car(param),
                     (uint8_t *)d_ok, (double *)d_next);
  PE_LAUNCHED(geom_actions);
  PE_HIP(hipMemcpy(ok, d_ok, m, hipMemcpyDeviceToHost));
  PE_HIP(hipMemcpy(next, d_next, m * 32, hipMemcpyDeviceToHost));
  return PE_OK;
}

// Map list of the multi-map queries: map_of in range, offsets ascending
// from 0, half-widths finite and positive.
static int check_maps(const double *param, const int32_t *map_of, long long n, int32_t n_maps,
                      const int64_t *obs_off, const double *halves) {
  if (!param || n < 0 || n_maps < 1 || (n && !map_of) || !obs_off || !halves)
    return fail(PE_ERR_INVALID, "invalid geometry arguments");
  if (obs_off[0] != 0) return fail(PE_ERR_INVALID, "obstacle offsets must start at 0");
  for (int b = 0; b < n_maps; b++) {
    if (obs_off[b + 1] < obs_off[b] || obs_off[b + 1] - obs_off[b] > INT32_MAX)
      return fail(PE_ERR_INVALID, "obstacle offsets must ascend");
    if (!std::isfinite(halves[b]) || halves[b] <= 0)
      return fail(PE_ERR_INVALID, "invalid half-width");
  }
  for (long long i = 0; i < n; i++)
    if (map_of[i] < 0 || map_of[i] >= n_maps) return fail(PE_ERR_INVALID, "map index out of range");
  return check_args(param, halves[0], 0, n);
}

extern "C" int pe_geom_backups_multi(const double *states, int64_t n, const int32_t *map_of,
                                     int32_t n_maps, const double *obstacles,
                                     const int64_t *obstacle_offsets, const double *halves,
                                     const double *param, uint8_t *out) {
  clear_error();
  int rc = check_maps(param, map_of, n, n_maps, obstacle_offsets, halves);
  if (!rc) rc = check_buffers(n, {states, out});
  if (rc || (rc = require_device(nullptr))) return rc;
  if (n == 0) return PE_OK;
  const long long nc = obstacle_offsets[n_maps];
  Scratch s;
  void *d_s, *d_m, *d_o, *d_off, *d_h, *d_out;
  if ((rc = s.upload(&d_s, states, n * 32)) || (rc = s.upload(&d_m, map_of, n * 4)) ||
      (rc = s.upload(&d_o, obstacles, nc * 24)) ||
      (rc = s.upload(&d_off, obstacle_offsets, (n_maps + 1) * 8ll)) ||
      (rc = s.upload(&d_h, halves, n_maps * 8ll)) || (rc = s.alloc(&d_out, n)))
    return rc;
  hipLaunchKernelGGL(geom_backups_multi, dim3(wave_blocks_for(n)), dim3(256), 0, 0,
                     (const double *)d_s, (long long)n, (const int *)d_m, (const double *)d_o,
                     (const long long *)d_off, (const double *)d_h, read_car(param),
                     (uint8_t *)d_out);
  PE_LAUNCHED(geom_backups_multi);
  PE_HIP(hipMemcpy(out, d_out, n, hipMemcpyDeviceToHost));
  return PE_OK;
}

extern "C" int pe_geom_actions_multi(const double *states, int64_t n, const int32_t *map_of,
                                     int32_t n_maps, const double *controls, int32_t na,
                                     double dt, const double *obstacles,
                                     const int64_t *obstacle_offsets, const double *halves,
                                     const double *param, uint8_t *ok, double *next) {
  clear_error();
  int rc = check_maps(param, map_of, n, n_maps, obstacle_offsets, halves);
  if (!rc) rc = check_buffers(n, {states, ok, next});
  if (rc) return rc;
  if (na < 1 || !controls || !(dt >= 0) || !std::isfinite(dt))
    return fail(PE_ERR_INVALID, "invalid controls or duration");
  if ((rc = require_device(nullptr))) return rc;
  if (n == 0) return PE_OK;
  const long long nc = obstacle_offsets[n_maps];
  const long long m = (long long)n * na;
  Scratch s;
  void *d_s, *d_m, *d_u, *d_o, *d_off, *d_h, *d_ok, *d_next;
  if ((rc = s.upload(&d_s, states, n * 32)) || (rc = s.upload(&d_m, map_of, n * 4)) ||
      (rc = s.upload(&d_u, controls, na * 16)) || (rc = s.upload(&d_o, obstacles, nc * 24)) ||
      (rc = s.upload(&d_off, obstacle_offsets, (n_maps + 1) * 8ll)) ||
      (rc = s.upload(&d_h, halves, n_maps * 8ll)) || (rc = s.alloc(&d_ok, m)) ||
      (rc = s.alloc(&d_next, m * 32)))
    return rc;
  hipLaunchKernelGGL(geom_actions_multi, dim3(wave_blocks_for(m)), dim3(256), 0, 0,
                     (const double *)d_s, (long long)n, (const int *)d_m, (const double *)d_u, na,
                     dt, (const double *)d_o, (const long long *)d_off, (const double *)d_h,
                     read_car(param), (uint8_t *)d_ok, (double *)d_next);
  PE_LAUNCHED(geom_actions_multi);
  PE_HIP(hipMemcpy(ok, d_ok, m, hipMemcpyDeviceToHost));
  PE_HIP(hipMemcpy(next, d_next, m * 32, hipMemcpyDeviceToHost));
  return PE_OK;
}

extern "C" int pe_geom_connections(const double *from, const double *to,
                                   int64_t n, const double *obstacles,
                                   int32_t nc, double half,
                                   const double *param, uint8_t *out) {
  clear_error();
  int rc = check_args(param, half, nc, n);
  if (!rc) rc = check_buffers(n, {from, to, out});
  if (rc || (rc = require_device(nullptr))) return rc;
  if (n == 0) return PE_OK;
  Scratch s;
  void *d_a, *d_b, *d_o, *d_out;
  if ((rc = s.upload(&d_a, from, n * 32)) || (rc = s.upload(&d_b, to, n * 32)) ||
      (rc = s.upload(&d_o, obstacles, nc * 24)) || (rc = s.alloc(&d_out, n)))
    return rc;
  hipLaunchKernelGGL(geom_connections, dim3(wave_blocks_for(n)), dim3(256), 0, 0,
                     (const double *)d_a, (const double *)d_b, (long long)n,
                     (const double *)d_o, nc, half, read_car(param),
                     (uint8_t *)d_out);
  PE_LAUNCHED(geom_connections);
  PE_HIP(hipMemcpy(out, d_out, n, hipMemcpyDeviceToHost));
  return PE_OK;
}
