/*
 * ORACLE — test infrastructure only (never linked by the product).
 *
 * CPU restatement, in double precision, of the reference's per-vehicle
 * geometry for the contact game (local_game/native_geometry.cpp):
 *
 *   step          exact held (a, k) arc with speed saturation    :8-21
 *   clearance     rectangle vs square walls and circles          :22-41
 *   interval      Lipschitz bisection of a swept arc, depth 18   :42-56
 *   motion_safe   the swept arc stays clear                      :57-65
 *   backup_safe   held-command braking on k in {0, -K, +K}       :66-88
 *   actions       successor + motion_safe && backup_safe(next)   :108-123
 *   connections   straight configuration-space segment is clear  :213-252
 *   contact       oriented-rectangle SAT payoff                  :124-135
 *
 * Pinned by the reference library itself (oracle/_ref/libnative_geometry.so,
 * compiled from native_geometry.cpp by oracle/ref.mk) on seeded inputs, and
 * by the closed forms of reversible/model.py:86-153 (exact arcs) and the
 * known answers of tests/test_contact_game.py:21-121.
 */
#include <math.h>
#include <stdint.h>
#include <omp.h>

#include "oracle.h"

typedef pe_oracle_model model_t;

void pe_oracle_step(const double *s, double a, double k, double t,
                    const model_t *m, double *out) {
  const double bound = a >= 0 ? m->high : m->low;
  const double reach = fabs(a) > 1e-14 ? (bound - s[3]) / a : INFINITY;
  const double held = fmin(t, fmax(0., reach));
  const double v = fmax(m->low, fmin(m->high, s[3] + a * t));
  const double dist = s[3] * held + .5 * a * held * held + v * (t - held);
  const double turn = k * dist;
  const double half = turn * .5;
  const double sinc = fabs(half) > 1e-12 ? sin(half) / half : 1.;
  out[0] = s[0] + dist * sinc * cos(s[2] + half);
  out[1] = s[1] + dist * sinc * sin(s[2] + half);
  out[2] = remainder(s[2] + turn, 2 * M_PI);
  out[3] = v;
}

double pe_oracle_clearance(const double *s, const double *obs, int nc,
                           double half, const model_t *m) {
  const double c = cos(s[2]), q = sin(s[2]);
  const double ex = .5 * (m->L * fabs(c) + m->W * fabs(q));
  const double ey = .5 * (m->L * fabs(q) + m->W * fabs(c));
  double best = fmin(half - fabs(s[0]) - ex, half - fabs(s[1]) - ey);
  const double reach = .5 * hypot(m->L, m->W);
  for (int j = 0; j < nc; j++) {
    const double dx = obs[3 * j] - s[0], dy = obs[3 * j + 1] - s[1];
    const double r = obs[3 * j + 2];
    const double far = reach + r + fmax(0., best);
    if (dx * dx + dy * dy > far * far)
      continue; /* cannot lower the minimum */
    const double lx = fabs(dx * c + dy * q) - m->L * .5;
    const double ly = fabs(-dx * q + dy * c) - m->W * .5;
    const double d = hypot(fmax(0., lx), fmax(0., ly)) + fmin(fmax(lx, ly), 0.) - r;
    best = fmin(best, d);
  }
  return best;
}

/* Bisection of [0, t] along the held arc: every piece must keep clearance
 * above 1e-7 and be certified by the Lipschitz bound, within depth 18. */
static int arc_clear(const double *s, double a, double k, double lip,
                     const double *obs, int nc, double half, const model_t *m,
                     double lo, double hi, double cl, double cr, int depth) {
  if (fmin(cl, cr) <= 1e-7)
    return 0;
  if (fmin(cl, cr) > lip * (hi - lo) * .5 + 1e-7)
    return 1;
  if (depth >= 18 || hi - lo < 1e-8)
    return 0;
  const double mid = .5 * (lo + hi);
  double z[4];
  pe_oracle_step(s, a, k, mid, m, z);
  const double cm = pe_oracle_clearance(z, obs, nc, half, m);
  return arc_clear(s, a, k, lip, obs, nc, half, m, lo, mid, cl, cm, depth + 1) &&
         arc_clear(s, a, k, lip, obs, nc, half, m, mid, hi, cm, cr, depth + 1);
}

int pe_oracle_motion_safe(const double *s, double a, double k, double t,
                          const double *obs, int nc, double half,
                          const model_t *m) {
  double end[4];
  pe_oracle_step(s, a, k, t, m, end);
  const double lip = fmax(fabs(s[3]), fabs(end[3])) *
                     (1 + .5 * hypot(m->L, m->W) * fabs(k));
  return arc_clear(s, a, k, lip, obs, nc, half, m, 0, t,
                   pe_oracle_clearance(s, obs, nc, half, m),
                   pe_oracle_clearance(end, obs, nc, half, m), 0);
}

int pe_oracle_backup_safe(const double *s, const double *obs, int nc,
                          double half, const model_t *m) {
  if (pe_oracle_clearance(s, obs, nc, half, m) <= 1e-7)
    return 0;
  if (fabs(s[3]) < 1e-10)
    return 1;
  /* Full-deceleration periods, then a final period with a = -v/dt. */
  const double speed = fabs(s[3]);
  int periods = (int)ceil(speed / (m->a * m->dt) - 1e-10);
  if (periods < 1)
    periods = 1;
  const double full = (periods - 1) * m->dt;
  const double last = speed - m->a * full;
  const double dist = speed * full - .5 * m->a * full * full + .5 * last * m->dt;
  /* Uniform deceleration over the same distance sweeps the same arc. */
  const double t = 2 * dist / speed, a = -s[3] / t;
  const double curv[3] = {0, -m->k, m->k};
  for (int j = 0; j < 3; j++)
    if (pe_oracle_motion_safe(s, a, curv[j], t, obs, nc, half, m))
      return 1;
  return 0;
}

void pe_oracle_batch_clearance(const double *states, int64_t n,
                               const double *obs, int nc, double half,
                               const double *param, double *out, int threads) {
  const model_t *m = (const model_t *)param;
#pragma omp parallel for num_threads(threads) schedule(static)
  for (int64_t i = 0; i < n; i++)
    out[i] = pe_oracle_clearance(states + 4 * i, obs, nc, half, m);
}

void pe_oracle_batch_backups(const double *states, int64_t n,
                             const double *obs, int nc, double half,
                             const double *param, uint8_t *out, int threads) {
  const model_t *m = (const model_t *)param;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 64)
  for (int64_t i = 0; i < n; i++)
    out[i] = (uint8_t)pe_oracle_backup_safe(states + 4 * i, obs, nc, half, m);
}

void pe_oracle_batch_actions(const double *states, int64_t n,
                             const double *controls, int na, double dt,
                             const double *obs, int nc, double half,
                             const double *param, uint8_t *ok, double *next,
                             int threads) {
  const model_t *m = (const model_t *)param;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 64)
  for (int64_t i = 0; i < n; i++)
    for (int u = 0; u < na; u++) {
      const double *s = states + 4 * i;
      const double a = controls[2 * u], k = controls[2 * u + 1];
      double *z = next + 4 * (i * na + u);
      pe_oracle_step(s, a, k, dt, m, z);
      ok[i * na + u] =
          (uint8_t)(pe_oracle_motion_safe(s, a, k, dt, obs, nc, half, m) &&
                    pe_oracle_backup_safe(z, obs, nc, half, m));
    }
}

/* Straight segment from a to b in (x, y, heading), heading by the short way. */
static double segment_clearance(const double *a, const double *b, double f,
                                double turn, const double *obs, int nc,
                                double half, const model_t *m) {
  const double s[4] = {a[0] + f * (b[0] - a[0]), a[1] + f * (b[1] - a[1]),
                       a[2] + f * turn, 0};
  return pe_oracle_clearance(s, obs, nc, half, m);
}

static int segment_clear(const double *a, const double *b, double turn,
                         double lip, const double *obs, int nc, double half,
                         const model_t *m, double lo, double hi, double cl,
                         double cr, int depth) {
  if (fmin(cl, cr) <= 1e-7)
    return 0;
  if (fmin(cl, cr) > lip * (hi - lo) * .5 + 1e-7)
    return 1;
  if (depth >= 20)
    return 0;
  const double mid = (lo + hi) * .5;
  const double cm = segment_clearance(a, b, mid, turn, obs, nc, half, m);
  return segment_clear(a, b, turn, lip, obs, nc, half, m, lo, mid, cl, cm, depth + 1) &&
         segment_clear(a, b, turn, lip, obs, nc, half, m, mid, hi, cm, cr, depth + 1);
}

void pe_oracle_batch_connections(const double *from, const double *to,
                                 int64_t n, const double *obs, int nc,
                                 double half, const double *param, uint8_t *out,
                                 int threads) {
  const model_t *m = (const model_t *)param;
#pragma omp parallel for num_threads(threads) schedule(static)
  for (int64_t i = 0; i < n; i++) {
    const double *a = from + 4 * i, *b = to + 4 * i;
    const double turn = remainder(b[2] - a[2], 2 * M_PI);
    const double lip = hypot(b[0] - a[0], b[1] - a[1]) +
                       .5 * hypot(m->L, m->W) * fabs(turn);
    out[i] = (uint8_t)segment_clear(a, b, turn, lip, obs, nc, half, m, 0, 1,
                                    pe_oracle_clearance(a, obs, nc, half, m),
                                    pe_oracle_clearance(b, obs, nc, half, m), 0);
  }
}

/* Rectangle-contact payoff of pursuer pose p and evader pose e. */
double pe_oracle_contact(const double *p, const double *e, double pl,
                         double pw, double el, double ew) {
  const double dx = e[0] - p[0], dy = e[1] - p[1], rel = e[2] - p[2];
  const double ac = fabs(cos(rel)), as = fabs(sin(rel));
  const double cp = cos(p[2]), sp = sin(p[2]), ce = cos(e[2]), se = sin(e[2]);
  const double g0 = fabs(dx * cp + dy * sp) - .5 * (pl + el * ac + ew * as);
  const double g1 = fabs(-dx * sp + dy * cp) - .5 * (pw + el * as + ew * ac);
  const double g2 = fabs(dx * ce + dy * se) - .5 * (el + pl * ac + pw * as);
  const double g3 = fabs(-dx * se + dy * ce) - .5 * (ew + pl * as + pw * ac);
  return fmax(fmax(g0, g1), fmax(g2, g3));
}
