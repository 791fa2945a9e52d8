/*
 * ORACLE — test infrastructure only (never linked by the product).
 *
 * CPU restatement of the dense eight-state reversible-car game in the fixed
 * square (reversible/native_world.cpp:32-314): state
 * [xP, yP, thP, vP, xE, yE, thE, vE], row-major with vE fastest; float
 * fields and float local arithmetic in the reference's operation order
 * (this file is compiled without floating-point contraction); ENO1/ENO2
 * slopes, local Lax-Friedrichs dissipation, RK1/RK2; target = capture or
 * evader wall fault, avoid = pursuer wall fault, tube projection
 * max(min(value, previous), avoid) on the completed update; arrival times
 * by linear interpolation of the crossing.
 *
 * Pinned by golden vectors written by the reference executable itself
 * (tests/golden/make_golden.py), bit for bit.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#include "oracle.h"

/* std::min / std::max as the reference evaluates them. */
static inline float mn(float a, float b) { return b < a ? b : a; }
static inline float mx(float a, float b) { return a < b ? b : a; }

typedef struct {
  int n[8];
  int64_t stride[8];
  float h[8];
  int64_t N, inner;
} W8Grid;

static void w8_grid(const pe_oracle_world8d_cfg *c, W8Grid *g) {
  const float extent = c->extent;
  const double pi = acos(-1.);
  const float dx = 2 * extent / (c->positions - 1), dh = (float)(2 * pi / c->headings),
              dvp = (c->p_max - c->p_min) / (c->p_speeds - 1),
              dve = (c->e_max - c->e_min) / (c->e_speeds - 1);
  const int shape[8] = {c->positions, c->positions, c->headings, c->p_speeds,
                        c->positions, c->positions, c->headings, c->e_speeds};
  const float spacing[8] = {dx, dx, dh, dvp, dx, dx, dh, dve};
  g->stride[7] = 1;
  for (int k = 6; k >= 0; k--) g->stride[k] = g->stride[k + 1] * shape[k + 1];
  for (int k = 0; k < 8; k++) {
    g->n[k] = shape[k];
    g->h[k] = spacing[k];
  }
  g->N = g->stride[0] * shape[0];
  g->inner = (int64_t)c->headings * c->e_speeds;
}

/* Value at `off` cells from node j along one axis: periodic wrap, or the
 * signed linear extrapolation of the edge values past a non-periodic end. */
static float w8_at(const float *v, int64_t j, int64_t stride, int pos, int n, int off,
                   int periodic) {
  int q = pos + off;
  if (periodic) {
    q = ((q % n) + n) % n;
    return v[j + (int64_t)(q - pos) * stride];
  }
  if (q < 0) {
    const int64_t edge = j - (int64_t)pos * stride;
    const float a = v[edge], b = v[edge + stride];
    const float sign = (float)((a > 0) - (a < 0));
    return a + (float)(-q) * fabsf(b - a) * sign;
  }
  if (q >= n) {
    const int64_t edge = j + (int64_t)(n - 1 - pos) * stride;
    const float a = v[edge], b = v[edge - stride];
    const float sign = (float)((a > 0) - (a < 0));
    return a + (float)(q - n + 1) * fabsf(a - b) * sign;
  }
  return v[j + (int64_t)off * stride];
}

static void w8_slopes(const float *v, int64_t j, int64_t stride, int pos, int n, float h,
                      int periodic, int linear, int order, float *left, float *right) {
  if (linear && pos == 0) {
    *left = *right = (v[j + stride] - v[j]) / h;
    return;
  }
  if (linear && pos == n - 1) {
    *left = *right = (v[j] - v[j - stride]) / h;
    return;
  }
  const float f = v[j], fm = w8_at(v, j, stride, pos, n, -1, periodic),
              fp = w8_at(v, j, stride, pos, n, 1, periodic);
  float l = (f - fm) / h, r = (fp - f) / h;
  if (order == 2) {
    const float hh = h * h;
    const float dc = (fp - 2 * f + fm) / hh;
    const float dl =
        linear && pos < 2 ? dc : (f - 2 * fm + w8_at(v, j, stride, pos, n, -2, periodic)) / hh;
    const float dr =
        linear && pos + 2 >= n ? dc : (w8_at(v, j, stride, pos, n, 2, periodic) - 2 * fp + f) / hh;
    l += .5f * h * (fabsf(dl) <= fabsf(dc) ? dl : dc);
    r -= .5f * h * (fabsf(dc) <= fabsf(dr) ? dc : dr);
  }
  *left = l;
  *right = r;
}

typedef struct {
  float *x, *cs, *sn, *pv, *ev;
} W8Tables;

static void w8_tables(const pe_oracle_world8d_cfg *c, const W8Grid *g, W8Tables *t) {
  const double pi = acos(-1.);
  t->x = malloc(sizeof(float) * c->positions);
  t->cs = malloc(sizeof(float) * c->headings);
  t->sn = malloc(sizeof(float) * c->headings);
  t->pv = malloc(sizeof(float) * c->p_speeds);
  t->ev = malloc(sizeof(float) * c->e_speeds);
  for (int i = 0; i < c->positions; i++) t->x[i] = -c->extent + i * g->h[0];
  for (int i = 0; i < c->headings; i++) {
    t->cs[i] = (float)cos(-pi + i * 2 * pi / c->headings);
    t->sn[i] = (float)sin(-pi + i * 2 * pi / c->headings);
  }
  for (int i = 0; i < c->p_speeds; i++) t->pv[i] = c->p_min + i * g->h[3];
  for (int i = 0; i < c->e_speeds; i++) t->ev[i] = c->e_min + i * g->h[7];
}

static void w8_free(W8Tables *t) {
  free(t->x);
  free(t->cs);
  free(t->sn);
  free(t->pv);
  free(t->ev);
}

static void w8_decode(const W8Grid *g, int64_t j, int q[8]) {
  for (int k = 0; k < 8; k++) q[k] = (int)((j / g->stride[k]) % g->n[k]);
}

double pe_oracle_world8d_cfl_limit(const pe_oracle_world8d_cfg *c) {
  W8Grid g;
  w8_grid(c, &g);
  const float pabs = mx(fabsf(c->p_min), fabsf(c->p_max)),
              eabs = mx(fabsf(c->e_min), fabsf(c->e_max));
  return c->cfl / (2 * (double)(pabs + eabs) / g.h[0] +
                   (double)(pabs * c->p_curvature + eabs * c->e_curvature) / g.h[2] +
                   (double)c->p_accel / g.h[3] + (double)c->e_accel / g.h[7]);
}

int64_t pe_oracle_world8d_states(const pe_oracle_world8d_cfg *c) {
  W8Grid g;
  w8_grid(c, &g);
  return g.N;
}

/* Initial field max(target, avoid) and arrival 0 where it is <= 0. */
void pe_oracle_world8d_init(const pe_oracle_world8d_cfg *c, float *field, float *arrival) {
  W8Grid g;
  w8_grid(c, &g);
  W8Tables t;
  w8_tables(c, &g, &t);
  const float L = c->arena, radius = c->radius;
#pragma omp parallel for num_threads(c->threads) schedule(static)
  for (int64_t block = 0; block < g.N / g.inner; block++) {
    const int64_t begin = block * g.inner;
    int q[8];
    w8_decode(&g, begin, q);
    const float target = mn(hypotf(t.x[q[0]] - t.x[q[4]], t.x[q[1]] - t.x[q[5]]) - radius,
                            L - mx(fabsf(t.x[q[4]]), fabsf(t.x[q[5]])));
    const float avoid = mx(fabsf(t.x[q[0]]), fabsf(t.x[q[1]])) - L + 1e-6f;
    const float initial = mx(target, avoid);
    for (int64_t k = 0; k < g.inner; k++) {
      field[begin + k] = initial;
      arrival[begin + k] = initial <= 0 ? 0.f : NAN;
    }
  }
  w8_free(&t);
}

/* One RK stage src -> dst (base: the previous field of the blend stage). */
static void w8_advance(const pe_oracle_world8d_cfg *c, const W8Grid *g, const W8Tables *t,
                       const float *src, float *dst, float dt, const float *base, double tau,
                       float *arrival) {
  const int order = c->order, linear = c->linear_speed_boundary;
  const float L = c->arena, ap = c->p_accel, ae = c->e_accel, kp = c->p_curvature,
              ke = c->e_curvature;
  const int np = c->p_speeds, ne = c->e_speeds, nh = c->headings;
#pragma omp parallel for num_threads(c->threads) schedule(static)
  for (int64_t block = 0; block < g->N / g->inner; block++) {
    const int64_t begin = block * g->inner;
    int q[8];
    w8_decode(g, begin, q);
    const float vp = t->pv[q[3]], palo = q[3] == 0 ? 0 : -ap, pahi = q[3] == np - 1 ? 0 : ap;
    const float fpx = vp * t->cs[q[2]], fpy = vp * t->sn[q[2]], wp = fabsf(vp) * kp;
    const float avoid = mx(fabsf(t->x[q[0]]), fabsf(t->x[q[1]])) - L + 1e-6f;
    for (int eh = 0; eh < nh; eh++) {
      q[6] = eh;
      for (int iv = 0; iv < ne; iv++) {
        q[7] = iv;
        const int64_t j = begin + (int64_t)eh * ne + iv;
        const float ve = t->ev[iv], we = fabsf(ve) * ke, ealo = iv == 0 ? 0 : -ae,
                    eahi = iv == ne - 1 ? 0 : ae;
        const float f[8] = {fpx, fpy, 0, 0, ve * t->cs[eh], ve * t->sn[eh], 0, 0};
        float gk[8], H = 0, diss = 0;
        for (int k = 0; k < 8; k++) {
          float lm, rp;
          w8_slopes(src, j, g->stride[k], q[k], g->n[k], g->h[k], k == 2 || k == 6,
                    linear && (k == 3 || k == 7), order, &lm, &rp);
          gk[k] = .5f * (lm + rp);
          const float alpha = k == 2 ? wp : k == 3 ? ap : k == 6 ? we : k == 7 ? ae : fabsf(f[k]);
          H += f[k] * gk[k];
          diss += .5f * alpha * (rp - lm);
        }
        const float pa = palo * gk[3], pb = pahi * gk[3];
        const float ea = ealo * gk[7], eb = eahi * gk[7];
        H += -wp * fabsf(gk[2]) + we * fabsf(gk[6]) + mn(pa, pb) + mx(ea, eb);
        float value = src[j] + dt * (H + diss);
        if (base) value = .5f * (base[j] + value);
        if (order == 1 || base) {
          const float previous = base ? base[j] : src[j];
          value = mx(mn(value, previous), avoid);
          if (previous > 0 && value <= 0 && isnan(arrival[j]))
            arrival[j] = (float)tau + dt * previous / (previous - value);
        }
        dst[j] = value;
      }
    }
  }
}

/* Solve to `horizon` from the given field (time `start`), writing a frame
 * into `frames` (may be NULL) after each output_dt; returns steps taken.
 * stats[0] = the CFL dt limit. */
int64_t pe_oracle_world8d_solve(const pe_oracle_world8d_cfg *c, float *field, float *arrival,
                                double start, double horizon, double output_dt, float *frames,
                                double *stats) {
  W8Grid g;
  w8_grid(c, &g);
  W8Tables t;
  w8_tables(c, &g, &t);
  const double maxdt = pe_oracle_world8d_cfl_limit(c);
  float *b = malloc(sizeof(float) * g.N);
  float *a = field;
  double tau = start;
  int64_t steps = 0;
  const int count = (int)ceil((horizon - start) / output_dt - 1e-10);
  for (int frame = 1; frame <= count; frame++) {
    const double goal = fmin(horizon, start + frame * output_dt);
    while (tau < goal - 1e-10) {
      const double dt = fmin(maxdt, goal - tau);
      if (c->order == 1) {
        w8_advance(c, &g, &t, a, b, (float)dt, NULL, tau, arrival);
        float *s = a;
        a = b;
        b = s;
      } else {
        w8_advance(c, &g, &t, a, b, (float)dt, NULL, tau, arrival);
        w8_advance(c, &g, &t, b, a, (float)dt, a, tau, arrival);
      }
      tau += dt;
      steps++;
    }
    if (frames) memcpy(frames + (int64_t)(frame - 1) * g.N, a, sizeof(float) * g.N);
  }
  if (a != field) {
    memcpy(field, a, sizeof(float) * g.N);
    b = a;
  }
  free(b);
  w8_free(&t);
  if (stats) stats[0] = maxdt;
  return steps;
}
