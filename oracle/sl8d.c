/*
 * ORACLE — test infrastructure only (never linked by the product).
 *
 * Double-precision CPU restatement of one Jacobi sweep of the compact
 * eight-state semi-Lagrangian min-max backup of the contact game
 * (local_game/grid_game.py:209-230 dense kernels, :399-421 packed kernels,
 * :286-333 / :511-581 drivers):
 *
 *   out = V
 *   for pairs (ip, ie) inside the exit radius:
 *     g_u(e')   = sum_k wP[ip,u,k] * V[idsP[ip,u,k], e']      (w > 0 only)
 *     Q(u, b)   = sum_q wE[ie,b,q] * g_u(idsE[ie,b,q])        (w > 0 only)
 *     out[ip,ie] = min(V[ip,ie], min_{u allowed} max_{b allowed} Q(u, b))
 *
 * V is the dense [np, ne] product of the two cars' legal states; pairs
 * outside the radius keep their value (the exit value).  The early exit of
 * the reference (worst >= best) never changes the result and is not used.
 *
 * Pinned by the joint-action brute force of tests/test_compact_dp.py:31-50
 * (restated in tests/test_oracle_sl8d.py) on seeded random tables, by
 * order-independence of the min-max, and by invariants (monotone sweeps,
 * captured pairs stay captured, exit values fixed).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <omp.h>

#include "oracle.h"

/* Returns max |out - V| over the sweep. inside[ip * ne + ie] != 0 marks the
 * pairs the backup updates. */
double pe_oracle_sl_sweep(int64_t np, int64_t ne, int npa, int nea,
                          const int32_t *pids, const float *pw,
                          const uint8_t *pallowed, const int32_t *eids,
                          const float *ew, const uint8_t *eallowed,
                          const uint8_t *inside, const double *V, double *out,
                          int threads) {
  double change = 0;
#pragma omp parallel num_threads(threads) reduction(max : change)
  {
    double *g = (double *)malloc(sizeof(double) * ne);
#pragma omp for schedule(dynamic, 1)
    for (int64_t ip = 0; ip < np; ip++) {
      for (int64_t ie = 0; ie < ne; ie++)
        out[ip * ne + ie] = V[ip * ne + ie];
      for (int u = 0; u < npa; u++) {
        if (!pallowed[ip * npa + u])
          continue;
        const int32_t *id = pids + (ip * npa + u) * 16;
        const float *w = pw + (ip * npa + u) * 16;
        for (int64_t e = 0; e < ne; e++) {
          double acc = 0;
          for (int k = 0; k < 16; k++)
            if (w[k] > 0)
              acc += (double)w[k] * V[(int64_t)id[k] * ne + e];
          g[e] = acc;
        }
        for (int64_t ie = 0; ie < ne; ie++) {
          if (!inside[ip * ne + ie])
            continue;
          double worst = -INFINITY;
          for (int b = 0; b < nea; b++) {
            if (!eallowed[ie * nea + b])
              continue;
            const int32_t *jd = eids + (ie * nea + b) * 16;
            const float *v = ew + (ie * nea + b) * 16;
            double q = 0;
            for (int j = 0; j < 16; j++)
              if (v[j] > 0)
                q += (double)v[j] * g[jd[j]];
            if (q > worst)
              worst = q;
          }
          if (worst < out[ip * ne + ie])
            out[ip * ne + ie] = worst;
        }
      }
      for (int64_t ie = 0; ie < ne; ie++) {
        const double d = fabs(out[ip * ne + ie] - V[ip * ne + ie]);
        if (d > change)
          change = d;
      }
    }
    free(g);
  }
  return change;
}

/* Initial field: rectangle-contact payoff inside the radius, exit value
 * outside (grid_game.py:200-208, :390-398; geometry.py:15-36). */
double pe_oracle_contact(const double *p, const double *e, double pl,
                         double pw, double el, double ew);

void pe_oracle_sl_init(int64_t np, int64_t ne, const double *pstates,
                       const double *estates, const uint8_t *inside,
                       double exit_value, double pl, double pwid, double el,
                       double ewid, double *V, int threads) {
#pragma omp parallel for num_threads(threads) schedule(static)
  for (int64_t ip = 0; ip < np; ip++)
    for (int64_t ie = 0; ie < ne; ie++)
      V[ip * ne + ie] =
          inside[ip * ne + ie]
              ? pe_oracle_contact(pstates + 4 * ip, estates + 4 * ie, pl, pwid,
                                  el, ewid)
              : exit_value;
}
