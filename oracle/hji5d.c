/*
 * ORACLE — test infrastructure only. Nothing in the product links or calls
 * this file; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load it, as a checker.
 *
 * CPU restatement of the reference 5D relative ENO/Lax-Friedrichs HJI solve
 * (reversible/native_relative.cpp), written in C with OpenMP.
 *
 *   storage      [vP, vE, psi, y, x], x fastest      native_relative.cpp:74-75
 *   init         phi = hypot(x, y) - r, arrival 0/NaN native_relative.cpp:90-100
 *   slopes       ENO1/ENO2 one-sided differences      native_relative.cpp:38-67
 *   Hamiltonian  closed form + local LF dissipation   native_relative.cpp:144-157
 *   update       tube projection min(0, H), RK2 blend native_relative.cpp:158-165
 *   arrival      linear zero crossing                 native_relative.cpp:166-168
 *   time loop    CFL limit, outputs, RK1/RK2          native_relative.cpp:76-81,173-191
 *
 * Arithmetic is double precision with the operations in the same order as
 * the reference, so with float storage ("store_f64 = 0") the fields are
 * bit-identical to the reference binary's.  With "store_f64 = 1" the fields
 * are kept in double, which the tests use to state the float32 error of the
 * GPU path.
 *
 * Local exit (exit_radius > 0, the GPU variant gpu_relative.py:126-131 in its
 * float64 mode): a node with x*x + y*y above the decimal literal "%.9e" of
 * r^2 takes the literal "%.9e" of r and skips the update (and the step's
 * maxima); the initial plane takes float(r) where x*x + y*y > r*r in float64
 * (gpu_relative.py:167-169).
 *
 * Parity pinned by: golden vectors of the compiled reference binary
 * (tests/golden), the explicit control-enumeration Euler step of
 * tests/test_reversible_free_dp.py:16-94 restated in numpy, and invariants
 * (monotone fields, mirror symmetry, CFL closed form).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#include "oracle.h"

typedef struct {
  double dx, dphi, dvp, dve, limit;
  double exit_r2, exit_value; /* decimal literals of r^2 and r */
  size_t sy, sp, se, sv, n;
  double *sn, *cs;
} grid_t;

static void grid_setup(const pe_oracle_hji5d_cfg *c, grid_t *g) {
  const double pi = acos(-1.0);
  if (c->exit_radius > 0) {
    char text[64];
    snprintf(text, sizeof text, "%.9e", c->exit_radius * c->exit_radius);
    g->exit_r2 = strtod(text, NULL);
    snprintf(text, sizeof text, "%.9e", c->exit_radius);
    g->exit_value = strtod(text, NULL);
  }
  g->dx = 2 * c->extent / (c->nx - 1);
  g->dphi = 2 * pi / c->nh;
  g->dvp = (c->vpmax - c->vpmin) / (c->nvp - 1);
  g->dve = (c->vemax - c->vemin) / (c->nve - 1);
  g->sy = (size_t)c->nx;
  g->sp = (size_t)c->nx * c->nx;
  g->se = g->sp * c->nh;
  g->sv = g->se * c->nve;
  g->n = g->sv * c->nvp;
  {
    /* native_relative.cpp:76-81 */
    const double VP = fmax(fabs(c->vpmin), fabs(c->vpmax));
    const double VE = fmax(fabs(c->vemin), fabs(c->vemax));
    const double sum = (VP + VE + VP * c->kp * c->extent) / g->dx +
                       (VE + VP * c->kp * c->extent) / g->dx +
                       (VP * c->kp + VE * c->ke) / g->dphi + c->ap / g->dvp +
                       c->ae / g->dve;
    g->limit = c->cfl / sum;
  }
  g->sn = (double *)malloc(sizeof(double) * c->nh);
  g->cs = (double *)malloc(sizeof(double) * c->nh);
  for (int k = 0; k < c->nh; k++) {
    const double q = -pi + k * g->dphi;
    g->sn[k] = sin(q);
    g->cs[k] = cos(q);
  }
}

static void grid_free(grid_t *g) {
  free(g->sn);
  free(g->cs);
}

double pe_oracle_hji5d_cfl_limit(const pe_oracle_hji5d_cfg *c) {
  grid_t g;
  grid_setup(c, &g);
  const double v = g.limit;
  grid_free(&g);
  return v;
}

/* Values along one axis around the node: w[0..4] = offsets -2..+2.  A field
 * is either float (f32) or double (f64) storage. */
typedef struct {
  const float *f32;
  const double *f64;
} field_t;

static inline double fget(field_t s, ptrdiff_t j) {
  return s.f32 ? (double)s.f32[j] : s.f64[j];
}

/* One-sided ENO slopes along an axis (native_relative.cpp:38-67). */
static inline void eno(field_t s, size_t j, size_t stride, int pos, int n,
                       double h, int periodic, int order, double *dm,
                       double *dp) {
#define AT(off)                                                                \
  fget(s, (ptrdiff_t)j +                                                       \
              (ptrdiff_t)((periodic ? (((pos + (off)) % n + n) % n)            \
                                    : (pos + (off))) -                         \
                          pos) *                                               \
                  (ptrdiff_t)stride)
  const double v = fget(s, (ptrdiff_t)j);
  if (!periodic && pos == 0) {
    *dm = *dp = (AT(1) - v) / h;
    return;
  }
  if (!periodic && pos == n - 1) {
    *dm = *dp = (v - AT(-1)) / h;
    return;
  }
  const double vm = AT(-1), vp = AT(1);
  *dm = (v - vm) / h;
  *dp = (vp - v) / h;
  if (order == 1)
    return;
  const double hh = h * h;
  const double sc = (vp - 2 * v + vm) / hh;
  const double sl = (periodic || pos >= 2) ? (v - 2 * vm + AT(-2)) / hh : sc;
  const double sr = (periodic || pos + 2 < n) ? (AT(2) - 2 * vp + v) / hh : sc;
  /* strict comparisons, native_relative.cpp:65-66 */
  *dm += .5 * h * (fabs(sl) < fabs(sc) ? sl : sc);
  *dp -= .5 * h * (fabs(sc) < fabs(sr) ? sc : sr);
#undef AT
}

/* One RK stage src -> dst (native_relative.cpp:124-172).  When blend is set,
 * dst holds the value at the start of the step and receives
 * (previous + stage) / 2.  Returns max(0, max(dst - previous)); writes
 * max |dst - previous| into *change.  Returns NAN on a non-finite update. */
static double stage(const pe_oracle_hji5d_cfg *c, const grid_t *g,
                    field_t src, float *dst32, double *dst64, float *arrival,
                    double dt, double t, int blend, double *change) {
  double increase = 0, delta = 0;
  int bad = 0;
  const size_t rows = g->n / c->nx;
#pragma omp parallel for schedule(static) reduction(max : increase, delta)     \
    reduction(| : bad) num_threads(c->threads)
  for (size_t row = 0; row < rows; row++) {
    const int iy = (int)(row % c->nx);
    const int ih = (int)((row / c->nx) % c->nh);
    const int ie = (int)((row / ((size_t)c->nx * c->nh)) % c->nve);
    const int iv = (int)(row / ((size_t)c->nx * c->nh * c->nve));
    const double y = -c->extent + iy * g->dx, vp = c->vpmin + iv * g->dvp,
                 ve = c->vemin + ie * g->dve;
    const double wp = fabs(vp) * c->kp, we = fabs(ve) * c->ke;
    const double bx = ve * g->cs[ih] - vp, by = ve * g->sn[ih];
    const double plow = iv == 0 ? 0 : -c->ap;
    const double phigh = iv == c->nvp - 1 ? 0 : c->ap;
    const double elow = ie == 0 ? 0 : -c->ae;
    const double ehigh = ie == c->nve - 1 ? 0 : c->ae;
    for (int ix = 0; ix < c->nx; ix++) {
      const size_t j = row * c->nx + ix;
      const double x = -c->extent + ix * g->dx;
      if (c->exit_radius > 0 && x * x + y * y > g->exit_r2) {
        if (dst32)
          dst32[j] = (float)g->exit_value;
        else
          dst64[j] = (double)(float)g->exit_value;
        continue;
      }
      double xm, xp, ym, yp, hm, hp, vm, vpl, em, ep;
      eno(src, j, 1, ix, c->nx, g->dx, 0, c->order, &xm, &xp);
      eno(src, j, g->sy, iy, c->nx, g->dx, 0, c->order, &ym, &yp);
      eno(src, j, g->sp, ih, c->nh, g->dphi, 1, c->order, &hm, &hp);
      eno(src, j, g->sv, iv, c->nvp, g->dvp, 0, c->order, &vm, &vpl);
      eno(src, j, g->se, ie, c->nve, g->dve, 0, c->order, &em, &ep);
      const double gx = .5 * (xm + xp), gy = .5 * (ym + yp),
                   gh = .5 * (hm + hp), gv = .5 * (vm + vpl),
                   ge = .5 * (em + ep);
      /* std::min(a, b) is (b < a ? b : a); std::max(a, b) is (a < b ? b : a) */
      const double pa = plow * gv, pb = phigh * gv;
      const double ea = elow * ge, eb = ehigh * ge;
      double H = bx * gx + by * gy - wp * fabs(y * gx - x * gy - gh) +
                 we * fabs(gh) + (pb < pa ? pb : pa) + (ea < eb ? eb : ea);
      H += .5 * ((fabs(bx) + wp * fabs(y)) * (xp - xm) +
                 (fabs(by) + wp * fabs(x)) * (yp - ym) + (wp + we) * (hp - hm) +
                 c->ap * (vpl - vm) + c->ae * (ep - em));
      const double old = fget(src, (ptrdiff_t)j);
      const double previous =
          blend ? (dst32 ? (double)dst32[j] : dst64[j]) : old;
      double value = old + dt * (H < 0. ? H : 0.);
      if (blend)
        value = .5 * (previous + value);
      if (!isfinite(value))
        bad |= 1;
      double stored;
      if (dst32) {
        dst32[j] = (float)value;
        stored = (double)dst32[j];
      } else {
        dst64[j] = value;
        stored = value;
      }
      const double inc = stored - previous;
      if (inc > increase)
        increase = inc;
      if (fabs(inc) > delta)
        delta = fabs(inc);
      if ((blend || c->order == 1) && previous > 0 && stored <= 0 &&
          isnan(arrival[j]))
        arrival[j] = (float)(t + dt * previous / (previous - stored));
    }
  }
  if (change)
    *change = delta;
  return bad ? NAN : increase;
}

/* Initial field (native_relative.cpp:90-100). */
static void initialize(const pe_oracle_hji5d_cfg *c, const grid_t *g,
                       float *a32, double *a64, float *arrival) {
  const size_t rows = g->n / c->nx;
#pragma omp parallel for schedule(static) num_threads(c->threads)
  for (size_t row = 0; row < rows; row++) {
    const double y = -c->extent + (row % c->nx) * g->dx;
    for (int ix = 0; ix < c->nx; ix++) {
      const size_t j = row * c->nx + ix;
      const double x = -c->extent + ix * g->dx;
      double phi = hypot(x, y) - c->radius;
      if (c->exit_radius > 0 && x * x + y * y > c->exit_radius * c->exit_radius)
        phi = (double)(float)c->exit_radius;
      double stored;
      if (a32) {
        a32[j] = (float)phi;
        stored = a32[j];
      } else {
        a64[j] = phi;
        stored = phi;
      }
      arrival[j] = stored <= 0 ? 0.f : NAN;
    }
  }
}

/*
 * Whole solve.  frames (nullable) receives n_frames = outputs + 1 fields of n
 * values (float, or double when store_f64).  final_out / arrival_out receive
 * n values each.  stats[0..5] = cfl limit, internal steps, frames, maximum
 * increase, final time, last step's max |change|.  Optional trace receives
 * the max |change| of each step, up to trace_cap entries.
 * Returns 0, -1 on invalid options, -2 on a non-finite update.
 */
int pe_oracle_hji5d_solve(const pe_oracle_hji5d_cfg *c, double horizon,
                          double output_dt, void *frames, void *final_out,
                          float *arrival_out, double *stats, double *trace,
                          int64_t trace_cap) {
  if (c->nx < 3 || c->nh < 4 || c->nvp < 3 || c->nve < 3 || c->threads < 1 ||
      (c->order != 1 && c->order != 2) || c->extent <= c->radius ||
      c->radius <= 0 || horizon <= 0 || output_dt <= 0 || c->cfl <= 0 ||
      c->cfl > 1 || c->ap <= 0 || c->ae <= 0 || c->kp <= 0 || c->ke <= 0 ||
      c->vpmin >= c->vpmax || c->vemin >= c->vemax)
    return -1;
  grid_t g;
  grid_setup(c, &g);
  const size_t n = g.n;
  const size_t es = c->store_f64 ? sizeof(double) : sizeof(float);
  void *a = final_out;
  void *b = malloc(n * es);
  float *arrival = arrival_out;
#define F32(p) (c->store_f64 ? NULL : (float *)(p))
#define F64(p) (c->store_f64 ? (double *)(p) : NULL)
#define FIELD(p) ((field_t){F32(p), F64(p)})
  initialize(c, &g, F32(a), F64(a), arrival);
  int nframes = 0;
  if (frames)
    memcpy((char *)frames + (size_t)nframes * n * es, a, n * es);
  nframes++;
  double t = 0, maxinc = 0, last = 0;
  int64_t steps = 0;
  int rc = 0;
  /* native_relative.cpp:173-191 */
  const int outputs = (int)ceil(horizon / output_dt - 1e-10);
  for (int out = 1; out <= outputs && rc == 0; out++) {
    const double goal = fmin(horizon, out * output_dt);
    while (t < goal - 1e-12) {
      const double dt = g.limit < goal - t ? g.limit : goal - t;
      double inc, change = 0;
      if (c->order == 1) {
        inc = stage(c, &g, FIELD(a), F32(b), F64(b), arrival, dt, t, 0, &change);
        void *s = a;
        a = b;
        b = s;
      } else {
        stage(c, &g, FIELD(a), F32(b), F64(b), arrival, dt, t, 0, NULL);
        inc = stage(c, &g, FIELD(b), F32(a), F64(a), arrival, dt, t, 1, &change);
      }
      if (isnan(inc)) {
        rc = -2;
        break;
      }
      if (inc > maxinc)
        maxinc = inc;
      last = change;
      if (trace && steps < trace_cap)
        trace[steps] = change;
      t += dt;
      steps++;
    }
    if (frames)
      memcpy((char *)frames + (size_t)nframes * n * es, a, n * es);
    nframes++;
  }
  /* RK1 swaps buffers; the caller's final_out must hold the final field. */
  if (a != final_out) {
    memcpy(final_out, a, n * es);
    free(a);
  } else {
    free(b);
  }
  if (stats) {
    stats[0] = g.limit;
    stats[1] = (double)steps;
    stats[2] = nframes;
    stats[3] = maxinc;
    stats[4] = t;
    stats[5] = last;
  }
  grid_free(&g);
  return rc;
#undef F32
#undef F64
#undef FIELD
}

/*
 * Fixed-dt steps on a caller-owned float field, for timing the CPU baseline
 * on a bounded sample: `count` RK steps of size dt starting at time t.
 * Returns the wall seconds spent in the update loops.
 */
double pe_oracle_hji5d_steps(const pe_oracle_hji5d_cfg *c, float *field,
                             float *arrival, double dt, double t, int count) {
  grid_t g;
  grid_setup(c, &g);
  float *b = (float *)malloc(g.n * sizeof(float));
  const double start = omp_get_wtime();
  for (int k = 0; k < count; k++) {
    if (c->order == 1) {
      stage(c, &g, (field_t){field, NULL}, b, NULL, arrival, dt, t, 0, NULL);
      memcpy(field, b, g.n * sizeof(float));
    } else {
      stage(c, &g, (field_t){field, NULL}, b, NULL, arrival, dt, t, 0, NULL);
      stage(c, &g, (field_t){b, NULL}, field, NULL, arrival, dt, t, 1, NULL);
    }
    t += dt;
  }
  const double seconds = omp_get_wtime() - start;
  free(b);
  grid_free(&g);
  return seconds;
}

/* Initial field only, for tests of the init kernel. */
void pe_oracle_hji5d_init(const pe_oracle_hji5d_cfg *c, float *field,
                          float *arrival) {
  grid_t g;
  grid_setup(c, &g);
  initialize(c, &g, field, NULL, arrival);
  grid_free(&g);
}
