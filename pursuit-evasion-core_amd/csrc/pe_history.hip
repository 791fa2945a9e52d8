// Saved field history of a completed contact solve, queried on the GPU.
//
// ContactReference.fields (local_game/reference.py:423-439) evaluates, for
// every query state and every saved frame, the product of the two cars'
// 16-vertex stencils over the packed float16 history: 256 terms per query,
// each a packed pair index, a weight, and whether the pair is stored (pairs
// dropped by the exit radius contribute the exit value).  The reference
// gathers this on the host; here the history stays in HBM and one thread
// per (query, frame) sums its 256 terms, so a batch of queries costs one
// launch.  The host keeps the stencil construction and the final min with
// the payoff, which are cheap.
#include <hip/hip_fp16.h>

#include <cstring>
#include <vector>

#include "pe_common.h"

struct pe_history {
  int device = 0;
  hipStream_t stream = nullptr;
  long long frames = 0, states = 0;
  __half *hist = nullptr;  // [frames][states]
  // query staging, grown on demand
  long long cap = 0;
  int terms = 0;
  long long *ids = nullptr;
  float *w = nullptr;
  uint8_t *kept = nullptr;
  float *out = nullptr;
};

namespace pe {

__global__ __launch_bounds__(256) void history_query(const __half *__restrict__ hist,
                                                     long long states, long long n, int terms,
                                                     const long long *__restrict__ ids,
                                                     const float *__restrict__ w,
                                                     const uint8_t *__restrict__ kept,
                                                     float exit_value, float *__restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long t = blockIdx.y;
  if (i >= n) return;
  const __half *__restrict__ frame = hist + t * states;
  float acc = 0.f;
  for (int k = 0; k < terms; k++) {
    const long long q = i * terms + k;
    const float v = kept[q] ? __half2float(frame[ids[q]]) : exit_value;
    acc = fmaf(v, w[q], acc);
  }
  out[t * n + i] = acc;
}

}  // namespace pe

using namespace pe;

extern "C" int pe_history_create(const uint16_t *f16, int64_t frames, int64_t states,
                                 pe_history **out) {
  clear_error();
  if (!f16 || !out || frames < 1 || frames > 65535 || states < 1)
    return fail(PE_ERR_INVALID, "history [1..65535 frames, states >= 1] required");
  *out = nullptr;
  int dev = 0;
  int rc = require_device(&dev);
  if (rc) return rc;
  pe_history *h = new pe_history();
  h->device = dev;
  h->frames = frames;
  h->states = states;
  auto body = [&]() -> int {
    PE_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    const size_t bytes = sizeof(__half) * (size_t)frames * (size_t)states;
    if (hipMalloc(&h->hist, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return fail(PE_ERR_MEMORY, "cannot allocate the history on the device");
    }
    PE_HIP(hipMemcpy(h->hist, f16, bytes, hipMemcpyHostToDevice));
    return PE_OK;
  };
  rc = body();
  if (rc) {
    pe_history_destroy(h);
    return rc;
  }
  *out = h;
  return PE_OK;
}

extern "C" int pe_history_query(pe_history *h, int64_t n, int32_t terms, const int64_t *ids,
                                const float *weights, const uint8_t *kept, float exit_value,
                                float *out) {
  clear_error();
  if (!h || n < 0 || terms < 1 || terms > 4096 || (n > 0 && (!ids || !weights || !kept || !out)))
    return fail(PE_ERR_INVALID, "invalid query arguments");
  if (n == 0) return PE_OK;
  const long long m = n * (long long)terms;
  for (long long q = 0; q < m; q++)
    if (kept[q] && (ids[q] < 0 || ids[q] >= h->states))
      return fail(PE_ERR_INVALID, "pair index out of range");
  PE_HIP(hipSetDevice(h->device));
  if (n > h->cap || terms != h->terms) {
    for (void *p : {(void *)h->ids, (void *)h->w, (void *)h->kept, (void *)h->out})
      if (p) hipFree(p);
    h->ids = nullptr;
    h->w = nullptr;
    h->kept = nullptr;
    h->out = nullptr;
    h->cap = 0;
    if (hipMalloc(&h->ids, sizeof(long long) * m) != hipSuccess ||
        hipMalloc(&h->w, sizeof(float) * m) != hipSuccess ||
        hipMalloc(&h->kept, m) != hipSuccess ||
        hipMalloc(&h->out, sizeof(float) * n * h->frames) != hipSuccess) {
      (void)hipGetLastError();
      return fail(PE_ERR_MEMORY, "cannot allocate query buffers");
    }
    h->cap = n;
    h->terms = terms;
  }
  // Terms of dropped pairs keep whatever index the caller passed; they are
  // never dereferenced.
  PE_HIP(hipMemcpyAsync(h->ids, ids, sizeof(long long) * m, hipMemcpyHostToDevice, h->stream));
  PE_HIP(hipMemcpyAsync(h->w, weights, sizeof(float) * m, hipMemcpyHostToDevice, h->stream));
  PE_HIP(hipMemcpyAsync(h->kept, kept, m, hipMemcpyHostToDevice, h->stream));
  const dim3 grid((unsigned)((n + 255) / 256), (unsigned)h->frames);
  hipLaunchKernelGGL(history_query, grid, dim3(256), 0, h->stream, h->hist, h->states,
                     (long long)n, (int)terms, h->ids, h->w, h->kept, exit_value, h->out);
  PE_LAUNCHED(history_query);
  PE_HIP(hipMemcpyAsync(out, h->out, sizeof(float) * n * h->frames, hipMemcpyDeviceToHost,
                        h->stream));
  PE_HIP(hipStreamSynchronize(h->stream));
  return PE_OK;
}

extern "C" void pe_history_destroy(pe_history *h) {
  if (!h) return;
  hipSetDevice(h->device);
  if (h->stream) hipStreamSynchronize(h->stream);
  for (void *p : {(void *)h->hist, (void *)h->ids, (void *)h->w, (void *)h->kept, (void *)h->out})
    if (p) hipFree(p);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
}
