/* Test infrastructure only: drives the CPU oracle (oracle/*.c) on small
 * inputs so that tests/test_oracle_sanitizers.py can run it built with
 * AddressSanitizer and UndefinedBehaviorSanitizer.  Every oracle entry
 * point the tests bind is called at least once; the driver checks only that
 * the results are finite (the values are pinned by the other oracle tests). */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "../oracle/oracle.h"

double pe_oracle_hji5d_cfl_limit(const pe_oracle_hji5d_cfg *c);
void pe_oracle_hji5d_init(const pe_oracle_hji5d_cfg *c, float *field, float *arrival);
double pe_oracle_hji5d_steps(const pe_oracle_hji5d_cfg *c, float *field, float *arrival,
                             double dt, double t, int count);
int pe_oracle_hji5d_solve(const pe_oracle_hji5d_cfg *c, double horizon, double output_dt,
                          void *frames, void *final_out, float *arrival_out, double *stats,
                          double *trace, int64_t trace_cap);
void pe_oracle_batch_clearance(const double *states, int64_t n, const double *obs, int nc,
                               double half, const double *param, double *out, int threads);
void pe_oracle_batch_backups(const double *states, int64_t n, const double *obs, int nc,
                             double half, const double *param, uint8_t *out, int threads);
void pe_oracle_batch_actions(const double *states, int64_t n, const double *controls, int na,
                             double dt, const double *obs, int nc, double half,
                             const double *param, uint8_t *ok, double *next, int threads);
void pe_oracle_sl_init(int64_t np, int64_t ne, const double *pstates, const double *estates,
                       const uint8_t *inside, double exit_value, double pl, double pwid,
                       double el, double ewid, double *V, int threads);
double pe_oracle_sl_sweep(int64_t np, int64_t ne, int npa, int nea, const int32_t *pids,
                          const float *pw, const uint8_t *pallowed, const int32_t *eids,
                          const float *ew, const uint8_t *eallowed, const uint8_t *inside,
                          const double *V, double *out, int threads);
void pe_oracle_grid_stencil(int d, const int *n, const double *lo, const double *hi,
                            const double *step, const double *period, unsigned periodic,
                            double slack, const double *states, int64_t m, int64_t *index,
                            double *weights, uint8_t *valid);
int64_t pe_oracle_world8d_states(const pe_oracle_world8d_cfg *c);
void pe_oracle_world8d_init(const pe_oracle_world8d_cfg *c, float *field, float *arrival);
int64_t pe_oracle_world8d_solve(const pe_oracle_world8d_cfg *c, float *field, float *arrival,
                                double start, double horizon, double output_dt, float *frames,
                                double *stats);

static unsigned long long rng_state = 12345;
static double uniform(double lo, double hi) {
  rng_state = rng_state * 6364136223846793005ull + 1442695040888963407ull;
  return lo + (hi - lo) * (double)(rng_state >> 11) / 9007199254740992.0;
}

static int fail(const char *what) {
  fprintf(stderr, "oracle driver: %s\n", what);
  return 1;
}

static int hji5d(int order, int store_f64, double exit_radius) {
  pe_oracle_hji5d_cfg c = {9, 8, 5, 5, order, 2, store_f64, 2.0, 0.8, 0.3,
                           -1.0, 2.0, -0.5, 1.0, 3.0, 3.0, 1.0, 2.0, exit_radius};
  const size_t n = 9 * 9 * 8 * 5 * 5;
  float *f = malloc(n * sizeof(float)), *a = malloc(n * sizeof(float));
  pe_oracle_hji5d_init(&c, f, a);
  const double dt = pe_oracle_hji5d_cfl_limit(&c);
  pe_oracle_hji5d_steps(&c, f, a, dt, 0.0, 3);
  int bad = !(dt > 0);
  for (size_t j = 0; j < n; j++) bad |= !isfinite(f[j]);
  /* a whole solve with two output frames */
  const double horizon = 2 * dt * 3, output_dt = dt * 3;
  const size_t es = store_f64 ? sizeof(double) : sizeof(float);
  void *frames = malloc(3 * n * es), *final = malloc(n * es);
  double stats[16] = {0}, trace[64];
  bad |= pe_oracle_hji5d_solve(&c, horizon, output_dt, frames, final, a, stats, trace, 64) != 0;
  free(frames);
  free(final);
  free(f);
  free(a);
  return bad;
}

static int geometry(void) {
  const double param[7] = {-1.0, 2.0, 3.0, 1.0, 0.5, 0.3, 0.05};
  double obs[3 * 6], states[4 * 40], controls[2 * 9], next[4 * 40 * 9], clear[40];
  uint8_t ok[40 * 9], safe[40];
  for (int k = 0; k < 6; k++) {
    obs[3 * k] = uniform(-2.5, 2.5);
    obs[3 * k + 1] = uniform(-2.5, 2.5);
    obs[3 * k + 2] = uniform(0.05, 0.3);
  }
  for (int i = 0; i < 40; i++) {
    states[4 * i] = uniform(-2.5, 2.5);
    states[4 * i + 1] = uniform(-2.5, 2.5);
    states[4 * i + 2] = uniform(-3.14, 3.14);
    states[4 * i + 3] = uniform(-1.0, 2.0);
  }
  for (int u = 0; u < 9; u++) {
    controls[2 * u] = 3.0 * (u / 3 - 1);
    controls[2 * u + 1] = 1.0 * (u % 3 - 1);
  }
  pe_oracle_batch_clearance(states, 40, obs, 6, 3.0, param, clear, 2);
  pe_oracle_batch_backups(states, 40, obs, 6, 3.0, param, safe, 2);
  pe_oracle_batch_actions(states, 40, controls, 9, 0.05, obs, 6, 3.0, param, ok, next, 2);
  int bad = 0;
  for (int i = 0; i < 40; i++) bad |= !isfinite(clear[i]);
  for (int i = 0; i < 4 * 40 * 9; i++) bad |= !isfinite(next[i]);
  return bad;
}

static int sweep(void) {
  enum { NP = 30, NE = 34, NA = 9 };
  double ps[4 * NP], es[4 * NE], V[NP * NE], out[NP * NE];
  int32_t pid[NP * NA * 16], eid[NE * NA * 16];
  float pwt[NP * NA * 16], ewt[NE * NA * 16];
  uint8_t pal[NP * NA], eal[NE * NA], inside[NP * NE];
  for (int i = 0; i < 4 * NP; i++) ps[i] = uniform(-2, 2);
  for (int i = 0; i < 4 * NE; i++) es[i] = uniform(-2, 2);
  for (int r = 0; r < NP * NA; r++) {
    pal[r] = uniform(0, 1) < 0.8;
    for (int q = 0; q < 16; q++) {
      pid[16 * r + q] = (int32_t)uniform(0, NP - 0.5);
      pwt[16 * r + q] = (float)(1.0 / 16);
    }
  }
  for (int r = 0; r < NE * NA; r++) {
    eal[r] = uniform(0, 1) < 0.8;
    for (int q = 0; q < 16; q++) {
      eid[16 * r + q] = (int32_t)uniform(0, NE - 0.5);
      ewt[16 * r + q] = (float)(1.0 / 16);
    }
  }
  for (int k = 0; k < NP * NE; k++) inside[k] = uniform(0, 1) < 0.7;
  pe_oracle_sl_init(NP, NE, ps, es, inside, 3.0, 0.5, 0.3, 0.5, 0.3, V, 2);
  double change = 0;
  for (int s = 0; s < 4; s++) {
    change = pe_oracle_sl_sweep(NP, NE, NA, NA, pid, pwt, pal, eid, ewt, eal, inside, V, out, 2);
    for (int k = 0; k < NP * NE; k++) V[k] = out[k];
  }
  int bad = !isfinite(change);
  for (int k = 0; k < NP * NE; k++) bad |= !isfinite(V[k]);
  return bad;
}

static int stencil(void) {
  const int n[5] = {9, 9, 8, 5, 5};
  const double lo[5] = {-2, -2, -3.14159, -1, -0.5}, hi[5] = {2, 2, 3.14159, 2, 1};
  const double step[5] = {0.5, 0.5, 0.785398, 0.75, 0.375};
  const double period[5] = {0, 0, 6.283185, 0, 0};
  double q[5 * 50], w[50 * 32];
  int64_t idx[50 * 32];
  uint8_t valid[50];
  for (int i = 0; i < 50; i++)
    for (int k = 0; k < 5; k++) q[5 * i + k] = uniform(lo[k] - 0.3, hi[k] + 0.3);
  pe_oracle_grid_stencil(5, n, lo, hi, step, period, 1u << 2, 1e-10, q, 50, idx, w, valid);
  int bad = 0;
  for (int i = 0; i < 50 * 32; i++) bad |= !isfinite(w[i]);
  return bad;
}

static int world8d(void) {
  pe_oracle_world8d_cfg c = {4, 4, 3, 3, 2, 1, 2, 2.0f, 1.5f, 0.3f,
                             -1.0f, 2.0f, -0.5f, 1.0f, 3.0f, 3.0f, 1.0f, 2.0f, 0.8};
  const int64_t n = pe_oracle_world8d_states(&c);
  float *f = malloc((size_t)n * sizeof(float)), *a = malloc((size_t)n * sizeof(float));
  float *frames = malloc(2 * (size_t)n * sizeof(float));
  double stats[4] = {0};
  pe_oracle_world8d_init(&c, f, a);
  const int64_t steps = pe_oracle_world8d_solve(&c, f, a, 0.0, 0.1, 0.05, frames, stats);
  int bad = steps < 1;
  for (int64_t j = 0; j < n; j++) bad |= !isfinite(f[j]);
  free(f);
  free(a);
  free(frames);
  return bad;
}

int main(void) {
  if (hji5d(2, 0, 0.0) || hji5d(1, 0, 0.0) || hji5d(2, 1, 0.0) || hji5d(2, 0, 1.7))
    return fail("5D HJI");
  if (geometry()) return fail("geometry");
  if (sweep()) return fail("8D sweep");
  if (stencil()) return fail("grid stencil");
  if (world8d()) return fail("dense 8D");
  printf("oracle driver: ok\n");
  return 0;
}
