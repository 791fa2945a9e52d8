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

#include "geometry_device.h"

namespace pe {

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

// One wave per pair: the segment test of segment_clear.
__global__ __launch_bounds__(256) void geom_connections(
    const double *__restrict__ from, const double *__restrict__ to,
    long long n, const double *__restrict__ obs, int nc, double half, Car m,
    uint8_t *out) {
  const long long i = wave_item();
  if (i >= n) return;  // uniform per wave
  const bool good = segment_clear(from + 4 * i, to + 4 * i, obs, nc, half, m);
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
