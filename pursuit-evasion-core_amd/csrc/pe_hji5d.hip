// 5D relative HJI solve on one GPU: handle, initial field, time loop,
// residual trace and the FP32 stage kernel instantiation.
//
// Host time loop mirrors reversible/native_relative.cpp:173-191: CFL step
// dt = min(limit, goal - t), RK2 = stage(a -> b) then blend stage(b -> a),
// RK1 = stage(a -> b) and swap.  The solve time is accumulated in float64 on
// the host.  Each final stage max-reduces, on the device, the step's
// max(dst - previous) ("maximum_increase", native_relative.cpp:165) and
// max |dst - previous| (the residual) into a per-step slot of a ring that the
// host reads after each call, so stepping never waits on the device.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <vector>

#include "hji5d_fast.h"

namespace pe {

void hji5d_launch_f32(const Hji5dGrid &g, const StageArgs<float> &a,
                      bool blend, bool final_stage, hipStream_t stream) {
  const dim3 grid((unsigned)(g.per_xcd * 8)), block(256);
  if (g.fast) {
#define PE_FAST(B, O, P, PAIR)                                                              \
  hipLaunchKernelGGL((hji5d_stage_fast<B, O, P, PAIR>), grid, block, 0, stream, a, g, \
                     (int)final_stage)
    if (g.psi_pair) {
      if (g.order == 2) {
        if (blend) PE_FAST(true, 2, false, true);
        else PE_FAST(false, 2, false, true);
      } else {
        PE_FAST(false, 1, false, true);
      }
    } else if (g.order == 2) {
      if (blend) {
        if (g.prefetch) PE_FAST(true, 2, true, false);
        else PE_FAST(true, 2, false, false);
      } else {
        if (g.prefetch) PE_FAST(false, 2, true, false);
        else PE_FAST(false, 2, false, false);
      }
    } else {
      if (g.prefetch) PE_FAST(false, 1, true, false);
      else PE_FAST(false, 1, false, false);
    }
#undef PE_FAST
    return;
  }
  if (blend)
    hipLaunchKernelGGL((hji5d_stage<float, false, true>), grid, block, 0,
                       stream, a, g, (int)final_stage);
  else
    hipLaunchKernelGGL((hji5d_stage<float, false, false>), grid, block, 0,
                       stream, a, g, (int)final_stage);
}

// field[j] = plane[j % plane_size]; arrival = 0 where captured, NaN elsewhere
// (native_relative.cpp:90-100, gpu_relative.py:42-45).
__global__ __launch_bounds__(256) void hji5d_initialize(
    const float *__restrict__ plane, long long plane_size,
    float *__restrict__ field, float *__restrict__ arrival, long long n) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < n;
       j += stride) {
    const float v = plane[j % plane_size];
    field[j] = v;
    arrival[j] = v <= 0.f ? 0.f : __builtin_nanf("");
  }
}

__global__ __launch_bounds__(256) void count_finite(
    const float *__restrict__ x, long long n, unsigned long long *count) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  unsigned long long c = 0;
  for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < n;
       j += stride)
    c += isfinite(x[j]) ? 1ull : 0ull;
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
  if ((threadIdx.x & 63) == 0 && c) atomicAdd(count, c);
}

__global__ void fill_u64(unsigned long long *p, long long n,
                         unsigned long long v) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < n;
       j += stride)
    p[j] = v;
}

// Rectangle-contact payoff of E relative to P at the origin with zero
// heading (local_game/geometry.py:15-36 with p = 0).
static double rectangle_separation_origin(double x, double y, double psi,
                                          double pl, double pw, double el,
                                          double ew) {
  const double ca = std::fabs(std::cos(psi)), sa = std::fabs(std::sin(psi));
  double best = -std::numeric_limits<double>::infinity();
  // P frame (heading 0): d = (x, y).
  best = std::fmax(best, std::fabs(x) - pl / 2 - (el * ca + ew * sa) / 2);
  best = std::fmax(best, std::fabs(y) - pw / 2 - (el * sa + ew * ca) / 2);
  // E frame (heading psi).
  const double c = std::cos(psi), s = std::sin(psi);
  const double ex = x * c + y * s, ey = -x * s + y * c;
  best = std::fmax(best, std::fabs(ex) - el / 2 - (pl * ca + pw * sa) / 2);
  best = std::fmax(best, std::fabs(ey) - ew / 2 - (pl * sa + pw * ca) / 2);
  return best;
}

}  // namespace pe

using namespace pe;

struct pe_hji5d {
  pe_hji5d_config cfg{};
  int device = 0;
  hipStream_t stream = nullptr;
  Hji5dGrid g{};
  double limit = 0;
  long long n = 0;
  float *a = nullptr, *b = nullptr, *arrival = nullptr;
  float *a_alloc = nullptr, *b_alloc = nullptr;  // a, b with guard margins
  long long margin = 0;                          // floats before a and b
  void *sn = nullptr, *cs = nullptr;
  float *vpk = nullptr;  // [nvp][2] vP, |vP| kP in float
  // Reduction ring: slot k holds [increase, |change|, nonfinite, pad].
  unsigned long long *ring = nullptr;
  int ring_slots = 4096, ring_used = 0;
  unsigned long long *counter = nullptr;
  std::vector<double> trace;
  double t = 0, maxinc = 0, last = 0, update_seconds = 0;
  long long steps = 0;
  bool profiling = false;
  std::vector<hipEvent_t> events;  // start/stop pairs of profiled launches
  std::vector<int> event_kind;
  double stage_ms[2] = {0, 0};
  long long stage_launches[2] = {0, 0};
  hipEvent_t t0 = nullptr, t1 = nullptr;
};

static int validate(const pe_hji5d_config &c) {
  const double v[] = {c.extent, c.cfl, c.radius, c.p_min, c.p_max,
                      c.e_min, c.e_max, c.p_acceleration, c.e_acceleration,
                      c.p_curvature, c.e_curvature};
  for (double x : v)
    if (!std::isfinite(x)) return fail(PE_ERR_INVALID, "Invalid solver options");
  // native_relative.cpp:260-265
  if (c.positions < 3 || c.headings < 4 || c.p_speeds < 3 || c.e_speeds < 3 ||
      (c.order != 1 && c.order != 2) || c.extent <= c.radius ||
      c.radius <= 0 || c.cfl <= 0 || c.cfl > 1 || c.p_acceleration <= 0 ||
      c.e_acceleration <= 0 || c.p_curvature <= 0 || c.e_curvature <= 0 ||
      c.p_min >= c.p_max || c.e_min >= c.e_max ||
      (c.precision != 0 && c.precision != 1) ||
      (c.capture != 0 && c.capture != 1) || c.vp_chunk < 0 ||
      (c.flags & ~(PE_HJI5D_LAYOUT_PLANE | PE_HJI5D_GENERIC | PE_HJI5D_NO_PREFETCH |
                   PE_HJI5D_PSI_OUTER | PE_HJI5D_PSI_SINGLE)) != 0)
    return fail(PE_ERR_INVALID, "Invalid solver options");
  if (c.exit_radius > 0 && !(c.exit_radius < c.extent))
    return fail(PE_ERR_INVALID,
                "Local exit radius must lie inside the numerical grid");
  if (c.capture == 1 && !(c.p_length > 0 && c.p_width > 0 && c.e_length > 0 &&
                          c.e_width > 0))
    return fail(PE_ERR_INVALID, "Rectangle capture requires positive footprints");
  const double n = double(c.positions) * c.positions * c.headings *
                   c.p_speeds * c.e_speeds;
  if (n > 4.0e9 || double(c.headings) * c.positions * c.positions > 2.0e9 ||
      double(c.headings) * c.positions * c.positions / 256.0 * c.e_speeds *
              c.p_speeds > 2.0e9)
    return fail(PE_ERR_INVALID, "Grid too large for 32-bit plane indexing");
  return PE_OK;
}

static int flush_ring(pe_hji5d *h) {
  if (h->ring_used == 0) return PE_OK;
  std::vector<unsigned long long> host((size_t)h->ring_used * 4);
  PE_HIP(hipMemcpyAsync(host.data(), h->ring, host.size() * 8,
                        hipMemcpyDeviceToHost, h->stream));
  PE_HIP(hipStreamSynchronize(h->stream));
  bool bad = false;
  for (int k = 0; k < h->ring_used; k++) {
    double inc, chg;
    std::memcpy(&inc, &host[4 * k + 0], 8);
    std::memcpy(&chg, &host[4 * k + 1], 8);
    bad |= host[4 * k + 2] != 0;
    if (inc > h->maxinc) h->maxinc = inc;
    h->last = chg;
    h->trace.push_back(chg);
  }
  hipLaunchKernelGGL(fill_u64, dim3(16), dim3(256), 0, h->stream, h->ring,
                     (long long)h->ring_used * 4, 0ull);
  PE_LAUNCHED(fill_u64);
  h->ring_used = 0;
  if (bad) return fail(PE_ERR_NONFINITE, "Nonfinite field update");
  return PE_OK;
}

static void record(pe_hji5d *h, int kind, bool start) {
  if (!h->profiling) return;
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return;
  hipEventRecord(e, h->stream);
  h->events.push_back(e);
  if (start) h->event_kind.push_back(kind);
}

static int collect_events(pe_hji5d *h) {
  for (size_t k = 0; k + 1 < h->events.size(); k += 2) {
    float ms = 0;
    PE_HIP(hipEventElapsedTime(&ms, h->events[k], h->events[k + 1]));
    const int kind = h->event_kind[k / 2];
    h->stage_ms[kind] += ms;
    h->stage_launches[kind] += 1;
  }
  for (hipEvent_t e : h->events) hipEventDestroy(e);
  h->events.clear();
  h->event_kind.clear();
  return PE_OK;
}

// Launches one RK step of size dt at time t (no host synchronisation).
static int one_step(pe_hji5d *h, double dt) {
  if (h->ring_used == h->ring_slots) {
    int rc = flush_ring(h);
    if (rc) return rc;
  }
  unsigned long long *slot = h->ring + 4 * h->ring_used;
  const bool f64 = h->cfg.precision == 1;
  auto stage = [&](const float *src, float *dst, bool blend, bool fin) {
    const int kind = blend ? 1 : 0;
    record(h, kind, true);
    if (f64) {
      StageArgs<double> a{src, dst, h->arrival, (const double *)h->sn,
                          (const double *)h->cs, dt, h->t, slot};
      hji5d_launch_f64(h->g, a, blend, fin, h->stream);
    } else {
      StageArgs<float> a{src, dst, h->arrival, (const float *)h->sn,
                         (const float *)h->cs, (float)dt, h->t, slot, h->vpk};
      hji5d_launch_f32(h->g, a, blend, fin, h->stream);
    }
    record(h, kind, false);
  };
  if (h->cfg.order == 1) {
    stage(h->a, h->b, false, true);
    std::swap(h->a, h->b);
  } else {
    stage(h->a, h->b, false, false);
    stage(h->b, h->a, true, true);
  }
  PE_LAUNCHED(hji5d_stage);
  h->ring_used++;
  h->t += dt;
  h->steps++;
  return PE_OK;
}

static int begin_updates(pe_hji5d *h) {
  PE_HIP(hipEventRecord(h->t0, h->stream));
  return PE_OK;
}

static int end_updates(pe_hji5d *h) {
  PE_HIP(hipEventRecord(h->t1, h->stream));
  int rc = flush_ring(h);  // synchronises the stream
  float ms = 0;
  PE_HIP(hipEventSynchronize(h->t1));
  PE_HIP(hipEventElapsedTime(&ms, h->t0, h->t1));
  h->update_seconds += ms * 1e-3;
  if (h->profiling) {
    int rc2 = collect_events(h);
    if (rc2) return rc2;
  }
  return rc;
}

extern "C" int pe_hji5d_create(const pe_hji5d_config *cfg, pe_hji5d **out) {
  clear_error();
  if (!cfg || !out) return fail(PE_ERR_INVALID, "null argument");
  *out = nullptr;
  int rc = validate(*cfg);
  if (rc) return rc;
  int dev = 0;
  rc = require_device(&dev);
  if (rc) return rc;
  pe_hji5d *h = new pe_hji5d();
  h->cfg = *cfg;
  h->device = dev;
  const pe_hji5d_config &c = *cfg;
  const double pi = std::acos(-1.0);
  Hji5dGrid &g = h->g;
  g.nx = c.positions;
  g.nh = c.headings;
  g.nvp = c.p_speeds;
  g.nve = c.e_speeds;
  g.nx2 = g.nx * g.nx;
  g.plane = g.nh * g.nx2;
  g.tiles = (g.plane + 255) / 256;
  g.order = c.order;
  g.sp = g.nx2;
  g.se = g.plane;
  g.sv = (long long)g.plane * g.nve;
  g.extent = c.extent;
  g.dx = 2 * c.extent / (c.positions - 1);
  g.dphi = 2 * pi / c.headings;
  g.dvp = (c.p_max - c.p_min) / (c.p_speeds - 1);
  g.dve = (c.e_max - c.e_min) / (c.e_speeds - 1);
  g.vpmin = c.p_min;
  g.vemin = c.e_min;
  g.ap = c.p_acceleration;
  g.ae = c.e_acceleration;
  g.kp = c.p_curvature;
  g.ke = c.e_curvature;
  g.exit_on = c.exit_radius > 0;
  {
    // The reference formats r^2 and r with "%.9e" into its kernel source.
    char r2[64], r1[64];
    std::snprintf(r2, sizeof r2, "%.9e", c.exit_radius * c.exit_radius);
    std::snprintf(r1, sizeof r1, "%.9e", c.exit_radius);
    g.exit_r2 = std::strtod(r2, nullptr);
    g.exit_value = std::strtod(r1, nullptr);
    g.exit_r2f = std::strtof(r2, nullptr);
    g.exit_valuef = std::strtof(r1, nullptr);
    g.extentf = float(c.extent);
    g.dxf = (2.0f * g.extentf) / float(c.positions - 1);
  }
  // vP values per thread: by default the whole axis (fastest on C1 of the
  // chunk lengths 1, 2, 3, 5, 8 and 15).
  g.vpc = c.vp_chunk > 0 ? std::min(c.vp_chunk, g.nvp) : g.nvp;
  g.nvpc = (g.nvp + g.vpc - 1) / g.vpc;
  g.xytiles = (g.nx2 + 255) / 256;
  g.layout = (c.flags & PE_HJI5D_LAYOUT_PLANE) ? 0 : 1;
  g.blocks = (g.layout == 1 ? g.xytiles * g.nh : g.tiles) * g.nve * g.nvpc;
  g.per_xcd = (g.blocks + 7) / 8;
  h->n = g.sv * g.nvp;
  // The instruction-lean FP32 kernel (whole vP axis per thread, layout 1,
  // 32-bit byte offsets).
  g.fast = c.precision == 0 && g.layout == 1 && g.nvpc == 1 &&
           !(c.flags & PE_HJI5D_GENERIC) && double(h->n) * 4.0 < 4.0e9;
  g.prefetch = !(c.flags & PE_HJI5D_NO_PREFETCH);
  g.psi_outer = (c.flags & PE_HJI5D_PSI_OUTER) != 0;
  // Two psi planes per thread where the heading count is even (C1 blend
  // stage 178 -> 165 us, identical bits).
  g.psi_pair = g.fast && g.nh % 2 == 0 && !(c.flags & PE_HJI5D_PSI_SINGLE);
  if (g.psi_pair) {
    g.blocks = g.xytiles * (g.nh / 2) * g.nve;
    g.per_xcd = (g.blocks + 7) / 8;
  }
  {  // native_relative.cpp:76-81
    const double VP = std::max(std::fabs(c.p_min), std::fabs(c.p_max));
    const double VE = std::max(std::fabs(c.e_min), std::fabs(c.e_max));
    const double sum = (VP + VE + VP * c.p_curvature * c.extent) / g.dx +
                       (VE + VP * c.p_curvature * c.extent) / g.dx +
                       (VP * c.p_curvature + VE * c.e_curvature) / g.dphi +
                       c.p_acceleration / g.dvp + c.e_acceleration / g.dve;
    h->limit = c.cfl / sum;
  }
  auto cleanup = [&](int code) {
    pe_hji5d_destroy(h);
    return code;
  };
#define TRY(call)            \
  do {                       \
    int rc_ = (call);        \
    if (rc_) return cleanup(rc_); \
  } while (0)
  auto alloc = [&](void **p, size_t bytes) -> int {
    PE_HIP(hipMalloc(p, bytes));
    return PE_OK;
  };
  auto mk_stream = [&]() -> int {
    PE_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    PE_HIP(hipEventCreate(&h->t0));
    PE_HIP(hipEventCreate(&h->t1));
    return PE_OK;
  };
  TRY(mk_stream());
  // Fields with zeroed guard margins of 2 vE strides (+ alignment) on both
  // sides: the FP32 kernel loads x, y and vE neighbours unconditionally.
  h->margin = 2 * g.se + 64;
  const size_t padded = (h->n + 2 * h->margin) * sizeof(float);
  TRY(alloc((void **)&h->a_alloc, padded));
  TRY(alloc((void **)&h->b_alloc, padded));
  h->a = h->a_alloc + h->margin;
  h->b = h->b_alloc + h->margin;
  {
    auto zero = [&]() -> int {
      PE_HIP(hipMemset(h->a_alloc, 0, padded));
      PE_HIP(hipMemset(h->b_alloc, 0, padded));
      return PE_OK;
    };
    TRY(zero());
  }
  TRY(alloc((void **)&h->arrival, h->n * sizeof(float)));
  TRY(alloc((void **)&h->ring, (size_t)h->ring_slots * 4 * 8));
  TRY(alloc((void **)&h->counter, 8));
  // Heading tables (native_relative.cpp:84-89), in the local precision.
  {
    std::vector<double> sn(g.nh), cs(g.nh);
    for (int k = 0; k < g.nh; k++) {
      const double q = -pi + k * g.dphi;
      sn[k] = std::sin(q);
      cs[k] = std::cos(q);
    }
    const size_t es = c.precision == 1 ? 8 : 4;
    TRY(alloc(&h->sn, g.nh * es));
    TRY(alloc(&h->cs, g.nh * es));
    auto upload = [&](void *dst, const std::vector<double> &v) -> int {
      if (es == 8) {
        PE_HIP(hipMemcpy(dst, v.data(), v.size() * 8, hipMemcpyHostToDevice));
      } else {
        std::vector<float> f(v.begin(), v.end());
        PE_HIP(hipMemcpy(dst, f.data(), f.size() * 4, hipMemcpyHostToDevice));
      }
      return PE_OK;
    };
    TRY(upload(h->sn, sn));
    TRY(upload(h->cs, cs));
    // Pursuer speeds and their turn-rate bounds, rounded once from double.
    std::vector<float> vpk(2 * (size_t)g.nvp);
    for (int iv = 0; iv < g.nvp; iv++) {
      const double vp = g.vpmin + iv * g.dvp;
      vpk[2 * iv] = float(vp);
      vpk[2 * iv + 1] = float(std::fabs(vp) * g.kp);
    }
    TRY(alloc((void **)&h->vpk, vpk.size() * sizeof(float)));
    PE_HIP(hipMemcpy(h->vpk, vpk.data(), vpk.size() * sizeof(float), hipMemcpyHostToDevice));
  }
  // Initial capture plane on the host, broadcast on the device
  // (native_relative.cpp:93-96; gpu_relative.py:156-179).
  {
    const int nx = g.nx;
    const bool rect = c.capture == 1;
    const long long plane_size = rect ? (long long)g.plane : (long long)g.nx2;
    std::vector<float> plane((size_t)plane_size);
    for (int ih = 0; ih < (rect ? g.nh : 1); ih++)
      for (int iy = 0; iy < nx; iy++)
        for (int ix = 0; ix < nx; ix++) {
          const double x = -c.extent + ix * g.dx, y = -c.extent + iy * g.dx;
          double v;
          if (rect)
            v = rectangle_separation_origin(x, y, -pi + ih * g.dphi,
                                            c.p_length, c.p_width, c.e_length,
                                            c.e_width);
          else
            v = std::hypot(x, y) - c.radius;
          float f = float(v);
          // numpy float64 coordinates and r * r (gpu_relative.py:167-169)
          if (g.exit_on && x * x + y * y > c.exit_radius * c.exit_radius)
            f = float(c.exit_radius);
          plane[((size_t)ih * nx + iy) * nx + ix] = f;
        }
    float *d_plane = nullptr;
    TRY(alloc((void **)&d_plane, plane.size() * 4));
    auto init = [&]() -> int {
      PE_HIP(hipMemcpy(d_plane, plane.data(), plane.size() * 4,
                       hipMemcpyHostToDevice));
      hipLaunchKernelGGL(hji5d_initialize, dim3(4096), dim3(256), 0, h->stream,
                         d_plane, plane_size, h->a, h->arrival, h->n);
      PE_LAUNCHED(hji5d_initialize);
      hipLaunchKernelGGL(fill_u64, dim3(16), dim3(256), 0, h->stream, h->ring,
                         (long long)h->ring_slots * 4, 0ull);
      PE_LAUNCHED(fill_u64);
      PE_HIP(hipStreamSynchronize(h->stream));
      return PE_OK;
    };
    int rc_init = init();
    hipFree(d_plane);
    if (rc_init) return cleanup(rc_init);
  }
#undef TRY
  *out = h;
  return PE_OK;
}

extern "C" int pe_hji5d_advance_to(pe_hji5d *h, double goal,
                                   int64_t max_steps) {
  clear_error();
  if (!h || !std::isfinite(goal)) return fail(PE_ERR_INVALID, "invalid argument");
  PE_HIP(hipSetDevice(h->device));
  int rc = begin_updates(h);
  if (rc) return rc;
  int64_t done = 0;
  while (h->t < goal - 1e-12 && (max_steps <= 0 || done < max_steps)) {
    const double dt = std::min(h->limit, goal - h->t);
    rc = one_step(h, dt);
    if (rc) return rc;
    done++;
  }
  return end_updates(h);
}

extern "C" int pe_hji5d_step(pe_hji5d *h, double dt, int64_t count) {
  clear_error();
  if (!h || !(dt > 0) || !std::isfinite(dt) || count < 0)
    return fail(PE_ERR_INVALID, "invalid argument");
  PE_HIP(hipSetDevice(h->device));
  int rc = begin_updates(h);
  if (rc) return rc;
  for (int64_t k = 0; k < count; k++) {
    rc = one_step(h, dt);
    if (rc) return rc;
  }
  return end_updates(h);
}

extern "C" int pe_hji5d_run_to_residual(pe_hji5d *h, double tol,
                                        int64_t max_steps, int64_t check_every,
                                        int32_t *converged) {
  clear_error();
  if (!h || !(tol > 0) || max_steps < 1 || check_every < 1)
    return fail(PE_ERR_INVALID, "invalid argument");
  PE_HIP(hipSetDevice(h->device));
  if (converged) *converged = 0;
  int64_t done = 0;
  while (done < max_steps) {
    const int64_t chunk = std::min(check_every, max_steps - done);
    int rc = pe_hji5d_step(h, h->limit, chunk);
    if (rc) return rc;
    done += chunk;
    if (h->last <= tol) {
      if (converged) *converged = 1;
      break;
    }
  }
  return PE_OK;
}

extern "C" int pe_hji5d_read(pe_hji5d *h, float *field, float *arrival) {
  clear_error();
  if (!h) return fail(PE_ERR_INVALID, "null handle");
  PE_HIP(hipSetDevice(h->device));
  if (field)
    PE_HIP(hipMemcpyAsync(field, h->a, h->n * 4, hipMemcpyDeviceToHost,
                          h->stream));
  if (arrival)
    PE_HIP(hipMemcpyAsync(arrival, h->arrival, h->n * 4,
                          hipMemcpyDeviceToHost, h->stream));
  PE_HIP(hipStreamSynchronize(h->stream));
  return PE_OK;
}

extern "C" int pe_hji5d_resolved(pe_hji5d *h, int64_t *count) {
  clear_error();
  if (!h || !count) return fail(PE_ERR_INVALID, "invalid argument");
  PE_HIP(hipSetDevice(h->device));
  hipLaunchKernelGGL(fill_u64, dim3(1), dim3(64), 0, h->stream, h->counter,
                     1ll, 0ull);
  PE_LAUNCHED(fill_u64);
  hipLaunchKernelGGL(count_finite, dim3(2048), dim3(256), 0, h->stream,
                     h->arrival, h->n, h->counter);
  PE_LAUNCHED(count_finite);
  unsigned long long c = 0;
  PE_HIP(hipMemcpyAsync(&c, h->counter, 8, hipMemcpyDeviceToHost, h->stream));
  PE_HIP(hipStreamSynchronize(h->stream));
  *count = (int64_t)c;
  return PE_OK;
}

extern "C" int pe_hji5d_trace(pe_hji5d *h, double *out, int64_t capacity,
                              int64_t *count) {
  clear_error();
  if (!h) return fail(PE_ERR_INVALID, "null handle");
  const int64_t n = (int64_t)h->trace.size();
  if (out)
    for (int64_t k = 0; k < std::min(n, capacity); k++) out[k] = h->trace[k];
  if (count) *count = n;
  return PE_OK;
}

extern "C" int pe_hji5d_get_stats(pe_hji5d *h, pe_hji5d_stats *s) {
  clear_error();
  if (!h || !s) return fail(PE_ERR_INVALID, "invalid argument");
  s->states = h->n;
  s->internal_steps = h->steps;
  s->time = h->t;
  s->cfl_dt_limit = h->limit;
  s->maximum_increase = h->maxinc;
  s->last_residual = h->last;
  s->update_seconds = h->update_seconds;
  for (int k = 0; k < 2; k++) {
    s->stage_ms[k] = h->stage_ms[k];
    s->stage_launches[k] = h->stage_launches[k];
  }
  return PE_OK;
}

extern "C" int pe_hji5d_set_profiling(pe_hji5d *h, int32_t on) {
  clear_error();
  if (!h) return fail(PE_ERR_INVALID, "null handle");
  h->profiling = on != 0;
  for (int k = 0; k < 2; k++) {
    h->stage_ms[k] = 0;
    h->stage_launches[k] = 0;
  }
  return PE_OK;
}

extern "C" void pe_hji5d_destroy(pe_hji5d *h) {
  if (!h) return;
  hipSetDevice(h->device);
  if (h->stream) hipStreamSynchronize(h->stream);
  for (hipEvent_t e : h->events) hipEventDestroy(e);
  for (void *p : {(void *)h->a_alloc, (void *)h->b_alloc, (void *)h->arrival, h->sn, h->cs,
                  (void *)h->vpk, (void *)h->ring, (void *)h->counter})
    if (p) hipFree(p);
  if (h->t0) hipEventDestroy(h->t0);
  if (h->t1) hipEventDestroy(h->t1);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
}
