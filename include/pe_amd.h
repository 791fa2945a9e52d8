/*
 * pe_amd — MI355X-native game value-function solver, C ABI.
 *
 * Plain C types only: the caller owns every host buffer it passes; the
 * library owns its device memory and one HIP stream per handle.  Every call
 * is synchronous on return.  Functions return 0 on success or a negative
 * PE_ERR_* code; pe_last_error() gives the message of the calling thread's
 * last failure.  One call per handle at a time.
 */
#ifndef PE_AMD_H
#define PE_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PE_OK 0
#define PE_ERR_INVALID (-1)   /* invalid options or shapes (ValueError)      */
#define PE_ERR_MEMORY (-2)    /* device allocation failed (MemoryError)      */
#define PE_ERR_NONFINITE (-3) /* non-finite field update (RuntimeError)      */
#define PE_ERR_HIP (-4)       /* HIP runtime error (RuntimeError)            */
#define PE_ERR_NODEVICE (-5)  /* no usable gfx950 device (RuntimeError)      */

/* Message of the calling thread's last failure ("" if none). */
const char *pe_last_error(void);
/* ABI version of this header (PE_AMD_ABI). */
#define PE_AMD_ABI 1
int pe_abi_version(void);
/* sizeof of a struct of this header by name ("pe_sl_car", ...), or -1;
   lets bindings check their struct mirrors without a device. */
int pe_abi_struct_size(const char *name);
/* Device-to-device copy of `bytes` bytes by a plain streaming kernel,
   `repeats` times; *gbps receives (read + write bytes) / time.  The
   calibration point for HBM traffic counters and the achievable copy
   bandwidth.  No counterpart in the reference. */
int pe_copy_bandwidth(int64_t bytes, int32_t repeats, double *gbps);
/* Name, compute units and HBM bytes of the current device. */
int pe_device_info(char *name, int name_len, int32_t *compute_units,
                   int64_t *hbm_bytes);

/* ======================================================================
 * 5D relative HJI: ENO1/ENO2 + local Lax-Friedrichs + RK1/RK2 with tube
 * projection, storage [vP, vE, psi, y, x] float32, x fastest.
 *
 * Replaces the argv/process ABI of the reference executable
 * reversible/native_relative.cpp:236-271 (flags --positions ... --order,
 * run() at :68-235) that reversible/free_dp.py:199-261 spawns, and the
 * CuPy RawKernels advance/initialize of reversible/gpu_relative.py:19-46.
 * ====================================================================== */
typedef struct pe_hji5d_config {
  int32_t positions, headings, p_speeds, e_speeds; /* --positions ...     */
  int32_t order;     /* 1 or 2 (--order)                                   */
  int32_t precision; /* 0: float32 local arithmetic; 1: float64 local
                        arithmetic in the reference's operation order      */
  double extent;     /* --extent */
  double cfl;        /* --cfl, in (0, 1] */
  double radius;     /* --radius: capture radius (center_distance) */
  double p_min, p_max, e_min, e_max;           /* --p-min ... --e-max   */
  double p_acceleration, e_acceleration;       /* --p/e-acceleration    */
  double p_curvature, e_curvature;             /* --p/e-curvature       */
  double exit_radius; /* > 0: local exit clamp of gpu_relative.py:126-131 */
  int32_t capture;    /* 0: center distance; 1: rectangle contact
                         (gpu_relative.py:161-166)                          */
  int32_t vp_chunk;   /* vP values walked per GPU thread (0: automatic)    */
  double p_length, p_width, e_length, e_width; /* rectangle footprints */
  int32_t flags;      /* PE_HJI5D_* launch options (0: defaults)           */
  int32_t reserved;
} pe_hji5d_config;

/* pe_hji5d_config.flags: number blocks by psi-plane tiles (the slower,
   earlier block order), for comparison runs. */
#define PE_HJI5D_LAYOUT_PLANE 1
/* pe_hji5d_config.flags: use the generic stage template for FP32 too
   (comparison runs; the default FP32 kernel is instruction-lean). */
#define PE_HJI5D_GENERIC 2
/* pe_hji5d_config.flags: the single-plane FP32 kernel loads each node's
   neighbours just before computing it instead of one vP node ahead
   (comparison runs; the two-plane kernel always loads just before). */
#define PE_HJI5D_NO_PREFETCH 4
/* pe_hji5d_config.flags: FP32 kernel block order (psi, xy tile, vE), so that
   y-adjacent tiles run close together (comparison runs). */
#define PE_HJI5D_PSI_OUTER 8
/* pe_hji5d_config.flags: each FP32 kernel thread walks the column of one psi
   plane (comparison runs).  By default, with an even heading count, a thread
   walks the columns of two adjacent psi planes, which share their psi
   neighbour loads; the results are the same bits. */
#define PE_HJI5D_PSI_SINGLE 16

typedef struct pe_hji5d_stats {
  int64_t states;           /* grid nodes                                  */
  int64_t internal_steps;   /* RK steps taken                              */
  double time;              /* accumulated solve time (float64)            */
  double cfl_dt_limit;      /* native_relative.cpp:76-81                   */
  double maximum_increase;  /* max over steps of max(dst - previous)       */
  double last_residual;     /* max |V(t+dt) - V(t)| of the last step       */
  double update_seconds;    /* device time of the update kernels           */
  /* per-launch event timing, filled while profiling is on */
  double stage_ms[2];       /* summed duration: [stage 1, stage 2 (blend)] */
  int64_t stage_launches[2];
} pe_hji5d_stats;

typedef struct pe_hji5d pe_hji5d;

/* Allocate the fields on the current device and write the initial field
 * phi (native_relative.cpp:82-100). */
int pe_hji5d_create(const pe_hji5d_config *cfg, pe_hji5d **out);
/* CFL-limited RK steps until the solve time reaches goal
 * (native_relative.cpp:176-191).  max_steps <= 0: no limit. */
int pe_hji5d_advance_to(pe_hji5d *h, double goal, int64_t max_steps);
/* `count` RK steps of size dt (fixed-work stepping for benchmarks). */
int pe_hji5d_step(pe_hji5d *h, double dt, int64_t count);
/* CFL-limited steps past the current time until the per-step residual
 * max |dV| <= tol, checking every `check_every` steps, at most max_steps.
 * *converged receives 1 if the tolerance was met. */
int pe_hji5d_run_to_residual(pe_hji5d *h, double tol, int64_t max_steps,
                             int64_t check_every, int32_t *converged);
/* Copy the current field and/or arrival times (N floats each; NULL skips). */
int pe_hji5d_read(pe_hji5d *h, float *field, float *arrival);
/* Number of nodes with a finite arrival time. */
int pe_hji5d_resolved(pe_hji5d *h, int64_t *count);
/* Per-step residual trace recorded so far; *count receives its length. */
int pe_hji5d_trace(pe_hji5d *h, double *out, int64_t capacity, int64_t *count);
int pe_hji5d_get_stats(pe_hji5d *h, pe_hji5d_stats *out);
/* Per-launch hipEvent timing of the stage kernels on the handle's stream. */
int pe_hji5d_set_profiling(pe_hji5d *h, int32_t on);
void pe_hji5d_destroy(pe_hji5d *h);

/* ======================================================================
 * Contact-game geometry on the GPU (double precision).
 *
 * Same argument meaning as the reference's OpenMP library
 * local_game/native_geometry.cpp: states [n,4] = x, y, heading, signed
 * speed; obstacles [nc,3] = circle x, y, r; half = square room half-width;
 * param[7] = speed low, speed high, acceleration, curvature, body length,
 * body width, control period.  No thread count; a status is returned.
 * ====================================================================== */
/* batch_clearance, native_geometry.cpp:92-99 */
int pe_geom_clearance(const double *states, int64_t n, const double *obstacles,
                      int32_t nc, double half, const double *param,
                      double *out);
/* batch_backups, native_geometry.cpp:100-107 */
int pe_geom_backups(const double *states, int64_t n, const double *obstacles,
                    int32_t nc, double half, const double *param,
                    uint8_t *out);
/* batch_actions, native_geometry.cpp:108-123: ok [n,na], next [n,na,4] */
int pe_geom_actions(const double *states, int64_t n, const double *controls,
                    int32_t na, double dt, const double *obstacles, int32_t nc,
                    double half, const double *param, uint8_t *ok,
                    double *next);
/* pe_geom_backups and pe_geom_actions for the states of many maps in one
   launch (no counterpart in the reference, which builds one map per call):
   state i belongs to map map_of[i] in [0, n_maps), whose circles are
   obstacles[obstacle_offsets[b] .. obstacle_offsets[b + 1]) (offsets
   ascending from 0, n_maps + 1 entries) and whose room half-width is
   halves[b].  Results equal the single-map calls bit for bit. */
int pe_geom_backups_multi(const double *states, int64_t n, const int32_t *map_of,
                          int32_t n_maps, const double *obstacles,
                          const int64_t *obstacle_offsets, const double *halves,
                          const double *param, uint8_t *out);
int pe_geom_actions_multi(const double *states, int64_t n, const int32_t *map_of,
                          int32_t n_maps, const double *controls, int32_t na, double dt,
                          const double *obstacles, const int64_t *obstacle_offsets,
                          const double *halves, const double *param, uint8_t *ok,
                          double *next);
/* configuration_connections, native_geometry.cpp:237-252 */
int pe_geom_connections(const double *from, const double *to, int64_t n,
                        const double *obstacles, int32_t nc, double half,
                        const double *param, uint8_t *out);

/* ======================================================================
 * Successor tables of one car grid (the stencil step of make_grid).
 *
 * Replaces the host numpy stencil VehicleGrid.stencil
 * (local_game/grid_game.py:81-123) and its cKDTree fallback
 * _nearest_supported (:38-79) that make_grid (:158-196) applies to every
 * held-command successor.  Axes: x, y, heading, speed, each ascending;
 * heading is periodic with cells of 2 pi / n starting at -pi.  The dense
 * grid is row-major with axis 0 slowest; dense_to_compact [prod n] holds
 * each legal state's compact index or -1.
 * ====================================================================== */
typedef struct pe_car_axes {
  int32_t n[4];           /* axis lengths (heading >= 1, others >= 2)      */
  const double *axis[4];  /* axis values                                   */
} pe_car_axes;

/* For m successor states next [m,4]: the 16-vertex multilinear stencil
   (vertex c has the upper node along axis k when bit 3 - k of c is set)
   with illegal vertices dropped and the rest renormalised: ids [m,16]
   compact ids (0 where dropped), weights [m,16] float32, mass [m] the
   retained weight before renormalisation, valid [m] = finite, within the
   non-periodic axes (1e-9 slack) and supported.  A successor inside the
   axes whose vertices are all illegal (mass <= 1e-9) takes the fallback:
   among the min(64, n) legal states nearest in the scaled features
   (x/dx, y/dx, cos h/dh, sin h/dh, v/dv; dx = x spacing, dh = 2 sin(pi/nh),
   dv = smallest speed spacing; ties by compact index), the first 16 whose
   straight configuration segment from the successor is clear of the map
   (obstacles, half, param as pe_geom_connections), weighted by
   1 / max(d, 1e-8)^2 normalised; valid = 0 when none is.  param == NULL
   skips the fallback (such successors keep zero weights and valid = 0, as
   the reference's stencil of a grid without a geometry engine).
   *fallbacks (may be NULL) receives the number of successors whose
   vertices were all illegal. */
int pe_car_stencil(const pe_car_axes *axes, const int32_t *dense_to_compact,
                   const double *next, int64_t m, const double *obstacles,
                   int32_t nc, double half, const double *param, int32_t *ids,
                   float *weights, uint8_t *valid, double *mass,
                   int64_t *fallbacks);

/* Multilinear stencils of m states [m, d] on a uniform grid of d <= 8 axes
   (row-major, axis 0 slowest): node index [m, 2^d], weights [m, 2^d]
   (float64), valid [m].  Axis k has n[k] nodes lo[k] + j step[k] up to
   hi[k].  Axes in the `periodic` bit mask (bit k) wrap: f = ((q - lo) mod
   period[k]) / step (numpy's remainder), lower node floor(f) mod n.  Other
   axes: valid within [lo - slack, hi + slack], f = (clip(q, lo, hi) - lo) /
   step, lower node min(floor(f), n - 2), fraction clipped to [0, 1].
   Vertex c takes the upper node on axis k when bit d - 1 - k of c is set;
   its weight multiplies the axis factors from axis 0 up.  Non-finite
   coordinates make the state invalid.  Replaces the host numpy stencils
   RelativeTable.stencil (reversible/relative_table.py:76-109) and
   interpolation_stencil (reversible/world_table.py:13-45). */
int pe_grid_stencil(int32_t d, const int32_t *n, const double *lo,
                    const double *hi, const double *step, const double *period,
                    uint32_t periodic, double slack, const double *states,
                    int64_t m, int64_t *index, double *weights, uint8_t *valid);

/* ======================================================================
 * Compact 8D semi-Lagrangian min-max backup of the contact game.
 *
 * Value V over the dense product of the two cars' legal states, row-major
 * [p.n, e.n] float32.  One Jacobi sweep:
 *   V'(ip,ie) = min(V, min_{u allowed} max_{b allowed}
 *               sum_q wE[ie,b,q] sum_k wP[ip,u,k] V[idsP[ip,u,k], idsE[ie,b,q]])
 * for pairs inside the exit radius; other pairs keep their (exit) value.
 *
 * Replaces the CuPy RawKernels initialize/interpolate_p/reduce_e and
 * init_packed/p_packed/e_packed of local_game/grid_game.py:199-231,389-422
 * and their drivers backup_gpu / backup_packed (:286-333, :511-581), called
 * per sweep by local_game/reference.py:352-358.  Several maps of the same
 * action count solve together in one handle (one launch per kernel).
 * ====================================================================== */
typedef struct pe_sl_car {
  int32_t n;                  /* legal states (VehicleGrid.states)          */
  int32_t actions;            /* commands per state (9)                     */
  const double *states;       /* [n,4] x, y, heading, speed                 */
  const int32_t *successors;  /* [n,actions,16] compact stencil ids         */
  const float *weights;       /* [n,actions,16] stencil weights (>= 0)      */
  const uint8_t *allowed;     /* [n,actions] command admissible             */
  const int32_t *position;    /* [n] index of the state's (x, y) cell       */
  int32_t positions;          /* number of (x, y) cells                     */
  const double *clearance;    /* [n] obstacle clearance (map feature of the
                                 learned warm start); may be NULL           */
} pe_sl_car;

typedef struct pe_sl_game {
  pe_sl_car p, e;
  /* [p.positions, e.positions]: 1 where the centre distance <= exit radius
     (PairDomain.build, grid_game.py:350-367) */
  const uint8_t *inside;
  /* [p.positions, e.positions]: 1 where an inside pair's E successor can
     lie (distance <= exit radius + E successor reach); 0 skips work */
  const uint8_t *near;
  float exit_value;           /* value outside the radius (exit_radius)     */
  float p_length, p_width, e_length, e_width; /* rectangle footprints       */
} pe_sl_game;

#define PE_SL_EARLY_EXIT 1 /* stop a command once worst >= best (exact)     */
#define PE_SL_ORDERED 2    /* try last sweep's argmin/argmax commands first */
#define PE_SL_GRAPH 4      /* replay captured sweeps with a hipGraph        */
#define PE_SL_GLOBAL_GATHER 8 /* E side gathers G from global memory instead
                                 of LDS-staged rows (comparison runs)      */
#define PE_SL_ONE_PASS 32     /* with PE_SL_ORDERED: the one-pass E side
                                 (one read of each E stencil for all P
                                 commands); the default                    */
#define PE_SL_U_MAJOR 64      /* with PE_SL_ORDERED: the u-major E side
                                 (fewer registers; comparison runs)        */
#define PE_SL_FREEZE 128      /* pe_sl_run only: a map whose residual is at
                                 most tol at a check is no longer swept in
                                 that call (each map stops as if solved
                                 alone); all maps are swept again after   */
#define PE_SL_WIDE_IDS 16     /* E side reads the separate 32-bit id and
                                 weight tables instead of the packed
                                 96-byte records of 16-bit ids + weights
                                 (comparison runs)                         */
#define PE_SL_STENCIL_P_SIDE 256 /* FP32: P side sums each command's compacted
                                 stencil in vertex order instead of adding
                                 each of the P state's union rows to all
                                 commands' sums (comparison runs; equal to
                                 rounding)                                 */
#define PE_SL_SEPARATE_P_SIDE 512 /* FP32: the P side as its own launch writing
                                 the G buffer, which the E side stages,
                                 instead of each E-side block computing its
                                 P state's values into LDS (comparison
                                 runs; the same bits)                      */

typedef struct pe_sl_stats {
  int64_t sweeps;
  int32_t maps;
  int32_t lds_e_side;     /* threads per block of the E side that stages
                             G rows in LDS (256, 512 or 1024, chosen per
                             handle); 0: G is gathered from global memory */
  int64_t pairs;          /* stored pairs, all maps                         */
  int64_t active_pairs;   /* pairs inside the exit radius, all maps         */
  double update_seconds;  /* device time of the sweeps                      */
  double last_residual;   /* max over maps of the last sweep's max|dV|      */
  double p_ms, e_ms;      /* profiled kernel time (summed)                  */
  int64_t p_launches, e_launches;
} pe_sl_stats;

typedef struct pe_sl pe_sl;

int pe_sl_create(const pe_sl_game *maps, int32_t n_maps, pe_sl **out);
/* As pe_sl_create with the value precision: 0 = FP32 (the product path),
   1 = FP64, the oracle's arithmetic on the GPU (each stencil sum a separate
   multiply and add over the positive weights, left to right), for
   full-size precision checks.  FP64 handles sweep, run, read and
   initialise; queries and warm starts need FP32. */
int pe_sl_create_ex(const pe_sl_game *maps, int32_t n_maps, int32_t precision,
                    pe_sl **out);
/* v0 == NULL: the rectangle-contact payoff inside, exit value outside
   (init_packed, grid_game.py:390-398); else the concatenated [p.n*e.n]
   fields of all maps (a warm start). */
int pe_sl_initialize(pe_sl *h, const float *v0);
/* The same with a double field (converted to the handle's precision). */
int pe_sl_initialize_f64(pe_sl *h, const double *v0);
int pe_sl_sweep(pe_sl *h, int32_t sweeps, int32_t flags);
/* Sweeps until every map's residual max|dV| <= tol (checked every
   check_every sweeps) or max_sweeps. */
int pe_sl_run(pe_sl *h, double tol, int32_t max_sweeps, int32_t check_every,
              int32_t flags, int32_t *sweeps_done, int32_t *converged);
int pe_sl_read(pe_sl *h, int32_t map, float *field);
/* The same into doubles (exact for FP64 handles). */
int pe_sl_read_f64(pe_sl *h, int32_t map, double *field);
/* Field of `map` at n continuous pair states given as the P and E stencils
   (ids [n,16] into the cars' legal states, weights [n,16]):
   out[i] = sum_a pw[i,a] sum_b ew[i,b] V[pi[i,a], ei[i,b]]
   (query_field kernel, local_game/reference.py:127-135). */
int pe_sl_query(pe_sl *h, int32_t map, int64_t n, const int32_t *p_ids,
                const float *p_weights, const int32_t *e_ids,
                const float *e_weights, float *out);
/* Residual of every recorded sweep, [sweeps, maps] row-major. */
int pe_sl_trace(pe_sl *h, double *out, int64_t capacity, int64_t *count);
int pe_sl_get_stats(pe_sl *h, pe_sl_stats *out);
int pe_sl_set_profiling(pe_sl *h, int32_t on);
/* Threads per block of the LDS E side (256, 512 or 1024), replacing the
   size chosen at creation; no reference counterpart (a tuning and test
   knob: every size gives the same values). */
int pe_sl_set_lds_threads(pe_sl *h, int32_t threads);
void pe_sl_destroy(pe_sl *h);

/* ======================================================================
 * Queries of a saved field history (reader of a completed contact solve).
 *
 * Replaces the host gather of ContactReference.fields
 * (local_game/reference.py:423-439): for query i and frame t,
 *   out[t, i] = sum_k w[i,k] * (kept[i,k] ? float(hist[t, ids[i,k]]) : exit)
 * over `terms` product-stencil terms (256 for two 16-vertex stencils);
 * hist is the packed float16 history [frames, states] (fields.f16).
 * ====================================================================== */
typedef struct pe_history pe_history;
int pe_history_create(const uint16_t *f16, int64_t frames, int64_t states,
                      pe_history **out);
/* ids, w, kept: [n, terms]; out: [frames, n]. */
int pe_history_query(pe_history *h, int64_t n, int32_t terms, const int64_t *ids,
                     const float *w, const uint8_t *kept, float exit_value,
                     float *out);
void pe_history_destroy(pe_history *h);

/* ======================================================================
 * Batch queries of a saved 5D relative-game history (the reader of a
 * completed free_dp solve).
 *
 * Replaces the host numpy gathers of RelativeTable.stencil/field/value
 * (reversible/relative_table.py:76-161).  The history fields.f32
 * [frames, vP, vE, psi, y, x] is copied to the device once; each query is a
 * state [x, y, psi, vP, vE] (the reader's public order).  Axes are given in
 * storage order [vP, vE, psi, y, x]: first value, last value, spacing
 * (axis[1] - axis[0]).
 * ====================================================================== */
typedef struct pe_rtable pe_rtable;
int pe_rtable_create(const float *fields, int32_t frames, const int32_t *shape,
                     const double *axes_lo, const double *axes_hi,
                     const double *axes_step, const double *times,
                     pe_rtable **out);
/* Signed field at states [n, 5] and budgets [n]: multilinear in space,
   linear in time between saved frames; NaN outside the grid or the budget
   range (relative_table.py:118-131). */
int pe_rtable_field(pe_rtable *h, const double *states, const double *budgets,
                    int64_t n, double *out);
/* pe_rtable_field for callers that cannot allocate device memory (code under
   HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
   No counterpart in the reference.  The batch is staged through a fixed
   device array of the library (PE_RTABLE_NOALLOC_MAX queries, 3.5 MiB, held
   by each device it runs on until the process ends) instead of the handle's
   growable buffer; the call makes no hipMalloc / hipFree.  n above the limit
   is rejected with PE_ERR_INVALID.  All calls on one device share that array:
   calls that overlap on one device (any handles) race. */
#define PE_RTABLE_NOALLOC_MAX 65536
int pe_rtable_field_noalloc(pe_rtable *h, const double *states,
                            const double *budgets, int64_t n, double *out);
/* First crossing of zero of the interpolated history (relative_table.py:
   133-161); 0 for captured physical states; speed_limits = P min, P max,
   E min, E max. */
int pe_rtable_value(pe_rtable *h, const double *states, int64_t n,
                    double capture_radius, const double *speed_limits,
                    double *out);
void pe_rtable_destroy(pe_rtable *h);

/* ======================================================================
 * Learned value surrogate: MLP with BF16 MFMA kernels (FP32 accumulation).
 *
 * No counterpart in the reference for this network; the reference's own
 * learned value nets (local_game/learning.py, training.py) use PyTorch.
 * Input per point: 8 state coordinates [xP, yP, thP, vP, xE, yE, thE, vE]
 * then `extra` map features; features = normalised state, `rff` random
 * Fourier features (sin, cos), extras, E's offset in P's frame, cos and sin
 * of the relative heading, the distance, zero padding to a multiple of 16.
 * Loss: weighted MSE with weight 1 + 4 exp(-|y| / 0.2)
 * (local_game/training.py:351-352); optimiser AdamW (training.py:185-188).
 * ====================================================================== */
typedef struct pe_mlp_config {
  int32_t extra;        /* map features after the 8 state coordinates (<= 40) */
  int32_t rff;          /* random Fourier features (sin + cos), even        */
  int32_t hidden;       /* hidden width, multiple of 64                     */
  int32_t layers;       /* hidden SiLU layers                               */
  uint32_t seed;        /* initialisation and Fourier projection seed      */
  float rff_scale;      /* std of the Fourier projection                    */
  float state_shift[8]; /* normalisation: (x - shift) * scale               */
  float state_scale[8];
  float learning_rate, weight_decay, beta1, beta2, eps; /* AdamW */
  float under_weight;   /* loss weight factor where yhat < y (>= 1; 1 = the
                           symmetric weighted MSE)                          */
} pe_mlp_config;

typedef struct pe_mlp pe_mlp;
int pe_mlp_create(const pe_mlp_config *cfg, pe_mlp **out);
int pe_mlp_param_count(pe_mlp *h, int64_t *count);
/* Flattened FP32 parameters: per hidden layer W [out, in] then b [out];
   then the head w [hidden], b [1]. */
int pe_mlp_get_params(pe_mlp *h, float *out, int64_t count);
int pe_mlp_set_params(pe_mlp *h, const float *in, int64_t count);
/* Predictions for n points of raw input [n, 8 + extra]. */
int pe_mlp_forward(pe_mlp *h, const float *raw, int64_t n, float *out);
/* The bf16 feature rows [n, width] the first layer sees, as float. */
int pe_mlp_features(pe_mlp *h, const float *raw, int64_t n, float *features,
                    int32_t *width);
/* Loss and parameter gradients (flattened like the parameters) of one
   batch, computed by the forward and backward MFMA kernels. */
int pe_mlp_gradients(pe_mlp *h, const float *raw, const float *target,
                     int64_t n, float *grads, float *loss);
/* Device-resident training set, then `steps` AdamW steps on batches drawn
   on the device; losses[steps] (may be NULL). */
int pe_mlp_set_dataset(pe_mlp *h, const float *raw, const float *target,
                       int64_t n);
int pe_mlp_train(pe_mlp *h, int32_t steps, int32_t batch, uint32_t seed,
                 float *losses);
/* The GEMM kernel alone on bf16 bit patterns: C = A [M,K] x B, with B given
   [K,N] (transposed_b = 0) or [N,K] (transposed_b = 1).  For tests. */
int pe_mlp_gemm_check(const uint16_t *a, const uint16_t *b, int32_t M,
                      int32_t N, int32_t K, int32_t transposed_b, float *c);
void pe_mlp_destroy(pe_mlp *h);

/* Warm start of every map of a solver from the network:
   V0 = min(phi, mlp(state, phi, clearance_P, clearance_E) + delta) on the
   pairs inside the exit radius (the network needs extra = 3); the control
   order is reset so that the first sweep records it from V0.  No
   counterpart in the reference. */
int pe_sl_warm_start(pe_sl *sl, pe_mlp *mlp, float delta);
/* Command order of every inside pair from the network: the argmin P command
   and the argmax E command of one backup evaluated on min(phi, mlp) are
   stored as the pairs' control order, which the next sweep with
   PE_SL_ORDERED tries first (the values are those of any order; the
   handle's field is not changed).  Call after pe_sl_initialize or
   pe_sl_warm_start, which reset the order.  No counterpart in the
   reference. */
int pe_sl_order_from_mlp(pe_sl *sl, pe_mlp *mlp);
/* Raw network inputs [count, 11] of pairs [first, first + count) of a map
   (row-major pair index ip * e.n + ie), for building training sets. */
int pe_sl_pair_inputs(pe_sl *sl, int32_t map, int64_t first, int64_t count,
                      float *raw);

/* ======================================================================
 * Dense 8D square game: state [xP, yP, thP, vP, xE, yE, thE, vE] float32,
 * row-major with vE fastest; ENO1/ENO2 + local Lax-Friedrichs + RK1/RK2,
 * tube projection max(min(value, previous), avoid) on completed updates,
 * float local arithmetic in the reference's operation order.
 *
 * Replaces the argv/process ABI of the reference executable
 * reversible/native_world.cpp:76-114 (flags --positions ... --linear-speed-
 * boundary, run() at :115-366) that reversible/world_dp.py:329-413 spawns.
 * ====================================================================== */
typedef struct pe_world8d_config {
  int32_t positions, headings, p_speeds, e_speeds; /* --positions ...      */
  int32_t order;                 /* 1 or 2 (--order)                        */
  int32_t linear_speed_boundary; /* 1: inward one-sided slopes at the speed
                                    ends; 0: signed extrapolation (legacy) */
  int32_t flags;                 /* PE_WORLD8D_* launch options (0: default) */
  int32_t reserved;
  /* Physical values; rounded to float as the executable parses them. */
  double extent, arena, radius;  /* --extent --arena --radius               */
  double p_min, p_max, e_min, e_max;   /* --p-min ... --e-max              */
  double p_accel, e_accel;             /* --p-accel --e-accel              */
  double p_curvature, e_curvature;     /* --p-curvature --e-curvature      */
  double cfl;                          /* --cfl, in (0, 1] (double)        */
} pe_world8d_config;

/* pe_world8d_config.flags (comparison runs; results are identical).  The
   default stages each pursuer slab of evader states in LDS, one block of
   512 threads per slab (or per range of xE rows of a slab above 36 KiB).
   GENERIC: one thread per node; SLAB_256 / SLAB_1024: that many threads
   per slab block. */
#define PE_WORLD8D_GENERIC 1
#define PE_WORLD8D_SLAB_256 2
#define PE_WORLD8D_SLAB_1024 4

typedef struct pe_world8d_stats {
  int64_t states;           /* grid nodes                                  */
  int64_t internal_steps;   /* RK steps taken                              */
  double time;              /* current solve time                          */
  double max_internal_dt;   /* CFL step limit (native_world.cpp:133-135)   */
  double update_seconds;    /* device time of the stepping                 */
  int64_t probes;           /* probe stencils set                          */
  int32_t kernel;           /* 0: thread per node; else threads per LDS
                               slab block                                  */
  int32_t fast_division;    /* 1: the spacings allow the reciprocal division
                               (pe_world8d_check_division)                 */
} pe_world8d_stats;

typedef struct pe_world8d pe_world8d;

/* Allocate the fields on the current device and write the initial field
 * max(target, avoid) and arrival (0 where it is <= 0, else NaN)
 * (native_world.cpp:136-170). */
int pe_world8d_create(const pe_world8d_config *cfg, pe_world8d **out);
/* Continuation: replace the field (N floats) and the arrival times (NULL:
 * `time` where the field is <= 0, else NaN), and set the solve time
 * (native_world.cpp:171-190). */
int pe_world8d_set_field(pe_world8d *h, const float *field, const float *arrival,
                         double time);
/* CFL-limited RK steps until the solve time reaches goal
 * (native_world.cpp:317-329).  max_steps <= 0: no limit. */
int pe_world8d_advance_to(pe_world8d *h, double goal, int64_t max_steps);
/* Copy the current field and/or arrival times (N floats each; NULL skips). */
int pe_world8d_read(pe_world8d *h, float *field, float *arrival);
/* Number of non-finite values of the current field
 * (native_world.cpp:336-344). */
int pe_world8d_count_nonfinite(pe_world8d *h, int64_t *count);
/* Probe stencils, 256 node indices and weights per probe, validated as in
 * native_world.cpp:205-227 ("Bad probe stencil", "Bad probe weights"). */
int pe_world8d_set_probes(pe_world8d *h, const uint64_t *indices,
                          const double *weights, int64_t count);
/* values[q] = sum_k double(field[index[q,k]]) * weight[q,k], summed in k
 * order in float64 over the current field (native_world.cpp:241-248). */
int pe_world8d_probe(pe_world8d *h, double *values);
int pe_world8d_get_stats(pe_world8d *h, pe_world8d_stats *out);
void pe_world8d_destroy(pe_world8d *h);
/* Footprint-map mode of the same game (replaces the CuPy RawKernels
 * init/advance of reversible/gpu_map_game.py:14-75): clearance grids
 * [positions, positions, headings] over (x, y, heading) of each car's
 * footprint against the map; target = min(capture, E clearance), avoid =
 * -P clearance + 1e-6, heading spacing float 2 pi / headings; the field is
 * re-initialised at time 0.  Needs the linear speed boundary. */
int pe_world8d_set_clearance(pe_world8d *h, const float *p_clearance,
                             const float *e_clearance);
/* Replace the CFL step limit (the map pilot takes cfl / global_rate_bound,
 * gpu_map_game.py:239). */
int pe_world8d_set_step_limit(pe_world8d *h, double max_dt);
/* Self-check of the kernels' division a / b = correction steps on RN(1/b)
 * (used while every field value is +0 or of magnitude in [2^-60, 2^60])
 * against the IEEE division, for every float a with exponent in [lo, hi)
 * of either sign and +0: mismatches[k] for divisors[k] in [2^-20, 2^20].
 * No counterpart in the reference. */
int pe_world8d_check_division(const float *divisors, int32_t count, int32_t lo,
                              int32_t hi, int64_t *mismatches);

#ifdef __cplusplus
}
#endif
#endif /* PE_AMD_H */
