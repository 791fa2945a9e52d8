// Device-side vehicle geometry of the contact game, in double precision,
// shared by the geometry kernels (pe_geometry.hip) and the successor-table
// builder (pe_tables.hip): exact held-command arcs, clearance against square
// walls and circles, certified swept-arc safety, braking-backup safety and
// straight configuration-space connections (local_game/native_geometry.cpp:
// 8-88, 213-252).
#pragma once

#include <cmath>

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

// Straight configuration-space segment a -> b (heading by the short way)
// stays clear, certified by bisection to depth 20 (native_geometry.cpp:213-252).
// Called by a whole wave; every lane returns the same answer.
__device__ bool segment_clear(const double *a, const double *b, const double *__restrict__ obs,
                              int nc, double half, const Car &m) {
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
  while (top > 0) {
    const Piece p = stack[--top];
    const double low = fmin(p.cl, p.cr);
    if (low <= 1e-7) return false;
    if (low > lip * (p.hi - p.lo) * .5 + 1e-7) continue;
    if (p.depth >= 20) return false;
    const double mid = (p.lo + p.hi) * .5;
    const double cm = at(mid);
    stack[top++] = {mid, p.hi, cm, p.cr, p.depth + 1};
    stack[top++] = {p.lo, mid, p.cl, cm, p.depth + 1};
  }
  return true;
}

}  // namespace pe
