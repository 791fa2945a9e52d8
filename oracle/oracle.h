/* ORACLE — test infrastructure only (see the header of each .c file). */
#ifndef PE_ORACLE_H
#define PE_ORACLE_H
#include <stddef.h>
#include <stdint.h>

/* 5D relative HJI options; same meaning as the relative5d argv flags of
 * reversible/native_relative.cpp:236-259. */
typedef struct {
  int nx, nh, nvp, nve, order, threads, store_f64;
  double extent, cfl, radius;
  double vpmin, vpmax, vemin, vemax, ap, ae, kp, ke;
  /* > 0: the local exit clamp of the reference's GPU variant in its float64
   * mode (reversible/gpu_relative.py:126-131, 167-169) */
  double exit_radius;
} pe_oracle_hji5d_cfg;

/* Dense 8D world game options; the argv flags of
 * reversible/native_world.cpp:77-104 (float parameters are floats there). */
typedef struct {
  int positions, headings, p_speeds, e_speeds, order, linear_speed_boundary, threads;
  float extent, arena, radius, p_min, p_max, e_min, e_max, p_accel, e_accel;
  float p_curvature, e_curvature;
  double cfl;
} pe_oracle_world8d_cfg;

/* Vehicle model parameters [low, high, a, k, L, W, dt]
 * (local_game/native_geometry.cpp:5-7,89-91). */
typedef struct {
  double low, high, a, k, L, W, dt;
} pe_oracle_model;

#endif
