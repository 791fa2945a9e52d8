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
  int32_t reserved;
  double p_length, p_width, e_length, e_width; /* rectangle footprints */
} pe_hji5d_config;

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

#ifdef __cplusplus
}
#endif
#endif /* PE_AMD_H */
