// Batch queries of a saved 5D relative-game history on the GPU (the reader
// reversible/relative_table.py:76-161: stencil, field, value).
//
// The history [frames, vP, vE, psi, y, x] float32 is resident in device
// memory.  One thread per query: it forms the 32-vertex multilinear stencil
// in float64 with the reader's formulas (storage-order coordinates
// [vP, vE, psi, y, x]; psi periodic as ((q + pi) mod 2 pi) / step; other axes
// clipped, last cell [n-2, n-1], weights multiplied axis by axis), then
//   field: linear interpolation in time between the saved frames around the
//          budget (relative_table.py:118-131);
//   value: bisection over the frames for the first interpolated value <= 0
//          and the linear crossing inside that interval, 0 for captured
//          physical states (relative_table.py:133-161).
// Vertex sums are accumulated in vertex order in float64.
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "pe_common.h"

namespace pe {

struct RTGrid {
  int n[5];               // storage order vP, vE, psi, y, x
  long long stride[5];
  double lo[5], hi[5], step[5];
  long long states;       // nodes per frame
  int frames;
};

struct RTStencil {
  long long base;         // index of the lower corner
  long long off[5];       // offset to the upper node along each axis
  double w[5];            // upper-node weight along each axis
  bool valid;
};

// relative_table.py:82-109 for one query (public order x, y, psi, vP, vE).
__device__ __forceinline__ RTStencil rt_stencil(const RTGrid &g, const double *s) {
  const double q[5] = {s[3], s[4], s[2], s[1], s[0]};
  RTStencil t;
  t.valid = true;
  t.base = 0;
  const double pi = 3.141592653589793;
  for (int k = 0; k < 5; k++) {
    double v = q[k];
    if (!isfinite(v)) {
      t.valid = false;
      v = 0;
    }
    const int n = g.n[k];
    long long j, j1;
    double w;
    if (k == 2) {
      // numpy remainder: fmod, moved into [0, 2 pi) when the signs differ
      double r = fmod(v + pi, 2 * pi);
      if (r != 0 && r < 0) r += 2 * pi;
      const double f = r / g.step[k];
      const double fl = floor(f);
      j = ((long long)fl % n + n) % n;
      w = f - fl;
      j1 = (j + 1) % n;
    } else {
      t.valid = t.valid && v >= g.lo[k] - 1e-10 && v <= g.hi[k] + 1e-10;
      const double c = fmin(fmax(v, g.lo[k]), g.hi[k]);
      const double f = (c - g.lo[k]) / g.step[k];
      j = (long long)floor(f);
      if (j > n - 2) j = n - 2;
      w = f - (double)j;
      j1 = j + 1;
    }
    t.base += j * g.stride[k];
    t.off[k] = (j1 - j) * g.stride[k];
    t.w[k] = w;
  }
  return t;
}

// Sum over the 32 vertices of frame t, vertex c having the upper node on
// axis k when bit (4 - k) of c is set (itertools.product order).
__device__ __forceinline__ double rt_sample(const float *__restrict__ fields, const RTGrid &g,
                                            const RTStencil &t, int frame) {
  const float *f = fields + (long long)frame * g.states;
  double sum = 0;
  for (int c = 0; c < 32; c++) {
    long long idx = t.base;
    double w = 1;
    for (int k = 0; k < 5; k++) {
      const bool up = (c >> (4 - k)) & 1;
      idx += up ? t.off[k] : 0;
      w *= up ? t.w[k] : 1 - t.w[k];
    }
    sum += (double)f[idx] * w;
  }
  return sum;
}

__global__ __launch_bounds__(256) void rt_field(const float *__restrict__ fields,
                                                const double *__restrict__ times, RTGrid g,
                                                const double *__restrict__ states,
                                                const double *__restrict__ budgets,
                                                long long count, double *__restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  RTStencil t = rt_stencil(g, states + 5 * i);
  double tau = budgets[i];
  const double horizon = times[g.frames - 1];
  const bool ok = t.valid && isfinite(tau) && tau >= 0 && tau <= horizon + 1e-10;
  if (!ok) {
    out[i] = NAN;
    return;
  }
  // k = clip(searchsorted(times, tau, 'right') - 1, 0, frames - 2)
  int lo = 0, hi = g.frames;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (times[mid] <= tau) lo = mid + 1;
    else hi = mid;
  }
  int k = lo - 1;
  k = k < 0 ? 0 : k > g.frames - 2 ? g.frames - 2 : k;
  const double fraction = (tau - times[k]) / (times[k + 1] - times[k]);
  const double a = rt_sample(fields, g, t, k), b = rt_sample(fields, g, t, k + 1);
  out[i] = a + fraction * (b - a);
}

__global__ __launch_bounds__(256) void rt_value(const float *__restrict__ fields,
                                                const double *__restrict__ times, RTGrid g,
                                                const double *__restrict__ states,
                                                long long count, double radius, double pmin,
                                                double pmax, double emin, double emax,
                                                double *__restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const double *s = states + 5 * i;
  bool physical = true;
  for (int k = 0; k < 5; k++) physical = physical && isfinite(s[k]);
  physical = physical && s[3] >= pmin - 1e-10 && s[3] <= pmax + 1e-10 &&
             s[4] >= emin - 1e-10 && s[4] <= emax + 1e-10;
  const bool captured = sqrt(s[0] * s[0] + s[1] * s[1]) <= radius;
  if (captured && physical) {
    out[i] = 0;
    return;
  }
  RTStencil t = rt_stencil(g, s);
  int lo = 0, hi = g.frames - 1;
  if (!t.valid || !(rt_sample(fields, g, t, hi) <= 0)) {
    out[i] = NAN;
    return;
  }
  while (hi - lo > 1) {
    const int mid = (lo + hi) / 2;
    if (rt_sample(fields, g, t, mid) <= 0) hi = mid;
    else lo = mid;
  }
  const double a = rt_sample(fields, g, t, lo), b = rt_sample(fields, g, t, hi);
  const double fraction = fabs(a - b) > 1e-14 ? a / (a - b) : 0.0;
  out[i] = times[lo] + fmin(fmax(fraction, 0.0), 1.0) * (times[hi] - times[lo]);
}

}  // namespace pe

using namespace pe;

struct pe_rtable {
  int device = 0;
  hipStream_t stream = nullptr;
  RTGrid g{};
  float *fields = nullptr;
  double *times = nullptr;
  double *buf = nullptr;   // states | budgets | out of the current batch
  long long capacity = 0;  // queries the buffer holds
};

extern "C" int pe_rtable_create(const float *fields, int32_t frames, const int32_t *shape,
                                const double *axes_lo, const double *axes_hi,
                                const double *axes_step, const double *times,
                                pe_rtable **out) {
  clear_error();
  if (!fields || !shape || !axes_lo || !axes_hi || !axes_step || !times || !out || frames < 2)
    return fail(PE_ERR_INVALID, "invalid argument");
  *out = nullptr;
  for (int k = 0; k < 5; k++)
    if (shape[k] < 2 || !(axes_step[k] > 0) || !std::isfinite(axes_lo[k]) ||
        !(axes_hi[k] > axes_lo[k]))
      return fail(PE_ERR_INVALID, "Expected five nondegenerate coordinate axes");
  for (int t = 1; t < frames; t++)
    if (!(times[t] > times[t - 1])) return fail(PE_ERR_INVALID, "Invalid output times");
  int dev = 0;
  int rc = require_device(&dev);
  if (rc) return rc;
  pe_rtable *h = new pe_rtable();
  h->device = dev;
  RTGrid &g = h->g;
  g.stride[4] = 1;
  for (int k = 3; k >= 0; k--) g.stride[k] = g.stride[k + 1] * shape[k + 1];
  for (int k = 0; k < 5; k++) {
    g.n[k] = shape[k];
    g.lo[k] = axes_lo[k];
    g.step[k] = axes_step[k];
    g.hi[k] = axes_hi[k];
  }
  g.states = g.stride[0] * shape[0];
  g.frames = frames;
  auto body = [&]() -> int {
    PE_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    const size_t bytes = sizeof(float) * (size_t)g.states * frames;
    if (hipMalloc(&h->fields, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return fail(PE_ERR_MEMORY, "cannot allocate the history");
    }
    PE_HIP(hipMemcpy(h->fields, fields, bytes, hipMemcpyHostToDevice));
    PE_HIP(hipMalloc(&h->times, sizeof(double) * frames));
    PE_HIP(hipMemcpy(h->times, times, sizeof(double) * frames, hipMemcpyHostToDevice));
    return PE_OK;
  };
  rc = body();
  if (rc) {
    pe_rtable_destroy(h);
    return rc;
  }
  *out = h;
  return PE_OK;
}

static int rt_reserve(pe_rtable *h, long long n) {
  if (n <= h->capacity) return PE_OK;
  if (h->buf) PE_HIP(hipFree(h->buf));
  h->buf = nullptr;
  h->capacity = 0;
  PE_HIP(hipMalloc(&h->buf, sizeof(double) * 7 * (size_t)n));
  h->capacity = n;
  return PE_OK;
}

extern "C" int pe_rtable_field(pe_rtable *h, const double *states, const double *budgets,
                               int64_t n, double *out) {
  clear_error();
  if (!h || n < 0 || (n > 0 && (!states || !budgets || !out)))
    return fail(PE_ERR_INVALID, "invalid argument");
  if (n == 0) return PE_OK;
  PE_HIP(hipSetDevice(h->device));
  int rc = rt_reserve(h, n);
  if (rc) return rc;
  double *s = h->buf, *b = s + 5 * n, *o = b + n;
  PE_HIP(hipMemcpyAsync(s, states, sizeof(double) * 5 * n, hipMemcpyHostToDevice, h->stream));
  PE_HIP(hipMemcpyAsync(b, budgets, sizeof(double) * n, hipMemcpyHostToDevice, h->stream));
  hipLaunchKernelGGL(rt_field, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream,
                     h->fields, h->times, h->g, s, b, (long long)n, o);
  PE_LAUNCHED(rt_field);
  PE_HIP(hipMemcpyAsync(out, o, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
  PE_HIP(hipStreamSynchronize(h->stream));
  return PE_OK;
}

extern "C" int pe_rtable_value(pe_rtable *h, const double *states, int64_t n,
                               double capture_radius, const double *speed_limits,
                               double *out) {
  clear_error();
  if (!h || n < 0 || !speed_limits || (n > 0 && (!states || !out)))
    return fail(PE_ERR_INVALID, "invalid argument");
  if (n == 0) return PE_OK;
  PE_HIP(hipSetDevice(h->device));
  int rc = rt_reserve(h, n);
  if (rc) return rc;
  double *s = h->buf, *o = s + 6 * n;
  PE_HIP(hipMemcpyAsync(s, states, sizeof(double) * 5 * n, hipMemcpyHostToDevice, h->stream));
  hipLaunchKernelGGL(rt_value, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream,
                     h->fields, h->times, h->g, s, (long long)n, capture_radius,
                     speed_limits[0], speed_limits[1], speed_limits[2], speed_limits[3], o);
  PE_LAUNCHED(rt_value);
  PE_HIP(hipMemcpyAsync(out, o, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
  PE_HIP(hipStreamSynchronize(h->stream));
  return PE_OK;
}

// Staging of pe_rtable_field_noalloc: states [n, 5] | budgets [n] | out [n]
// doubles for at most PE_RTABLE_NOALLOC_MAX queries (3.5 MiB), one copy per
// device for the life of the process.  Host code never passes `scratch`
// itself to a kernel or copy (in host code it is a stand-in, not a device
// address): it uses the device address hipGetSymbolAddress returns, looked
// up once per device.
static __device__ __align__(256) double scratch[7 * PE_RTABLE_NOALLOC_MAX];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double *g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// Device address of `scratch` on the current device `dev`.
static int rt_scratch(int dev, double **out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(PE_ERR_INVALID, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return fail(PE_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                  hipGetErrorString(e));
    }
    g_scratch_ptr[dev] = static_cast<double *>(p);
  }
  *out = g_scratch_ptr[dev];
  return PE_OK;
}

extern "C" int pe_rtable_field_noalloc(pe_rtable *h, const double *states,
                                       const double *budgets, int64_t n, double *out) {
  clear_error();
  if (!h || n < 0 || (n > 0 && (!states || !budgets || !out)))
    return fail(PE_ERR_INVALID, "invalid argument");
  if (n > PE_RTABLE_NOALLOC_MAX)
    return fail(PE_ERR_INVALID, "pe_rtable_field_noalloc: " + std::to_string(n) +
                                    " queries exceed the limit of " +
                                    std::to_string(PE_RTABLE_NOALLOC_MAX));
  if (n == 0) return PE_OK;
  PE_HIP(hipSetDevice(h->device));
  double *buf = nullptr;
  int rc = rt_scratch(h->device, &buf);
  if (rc) return rc;
  double *s = buf, *b = s + 5 * n, *o = b + n;
  PE_HIP(hipMemcpyAsync(s, states, sizeof(double) * 5 * n, hipMemcpyHostToDevice, h->stream));
  PE_HIP(hipMemcpyAsync(b, budgets, sizeof(double) * n, hipMemcpyHostToDevice, h->stream));
  hipLaunchKernelGGL(rt_field, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream,
                     h->fields, h->times, h->g, s, b, (long long)n, o);
  PE_LAUNCHED(rt_field);
  PE_HIP(hipMemcpyAsync(out, o, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
  PE_HIP(hipStreamSynchronize(h->stream));
  return PE_OK;
}

extern "C" void pe_rtable_destroy(pe_rtable *h) {
  if (!h) return;
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (void *p : {(void *)h->fields, (void *)h->times, (void *)h->buf})
    if (p) (void)hipFree(p);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}
