// Learned warm start of the contact-game backup: network inputs of pairs
// built on the device from the solver's car tables, and
// V0 = min(phi, mlp + delta) written into the solver's field.
//
// A value iteration V <- min(V, T V) started from any V0 with
// V* <= V0 <= phi converges to the same limit V* as the cold start from phi
// (the sweep is monotone and V* is its fixed point); delta lifts the network
// above V* where its error is within delta.  The plain sweep is unchanged.
#include "pe_internal.h"

namespace pe {

__device__ __forceinline__ float contact_payoff(const double *a, const double *b,
                                                const SlMap &m) {
  const double dx = b[0] - a[0], dy = b[1] - a[1], rel = b[2] - a[2];
  const double ac = fabs(cos(rel)), as = fabs(sin(rel));
  const double cp = cos(a[2]), sp = sin(a[2]), ce = cos(b[2]), se = sin(b[2]);
  const double g0 = fabs(dx * cp + dy * sp) - .5 * (m.pl + m.el * ac + m.ew * as);
  const double g1 = fabs(-dx * sp + dy * cp) - .5 * (m.pw + m.el * as + m.ew * ac);
  const double g2 = fabs(dx * ce + dy * se) - .5 * (m.el + m.pl * ac + m.pw * as);
  const double g3 = fabs(-dx * se + dy * ce) - .5 * (m.ew + m.pl * as + m.pw * ac);
  return float(fmax(fmax(g0, g1), fmax(g2, g3)));
}

// raw[i] = [P state (4), E state (4), phi, clearance P, clearance E]
__global__ __launch_bounds__(256) void pair_inputs(
    const SlMap *__restrict__ maps, int map, long long first, long long count,
    const double *__restrict__ pstates, const double *__restrict__ estates,
    const float *__restrict__ pclear, const float *__restrict__ eclear,
    float *__restrict__ raw) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const SlMap m = maps[map];
  const long long pair = first + i;
  const int ip = (int)(pair / m.ne), ie = (int)(pair - (long long)ip * m.ne);
  const double *a = pstates + 4ll * (m.p_row + ip);
  const double *b = estates + 4ll * (m.e_row + ie);
  float *r = raw + 11 * i;
  for (int d = 0; d < 4; d++) {
    r[d] = (float)a[d];
    r[4 + d] = (float)b[d];
  }
  r[8] = contact_payoff(a, b, m);
  r[9] = pclear[m.p_row + ip];
  r[10] = eclear[m.e_row + ie];
}

__global__ __launch_bounds__(256) void warm_combine(
    const SlMap *__restrict__ maps, int map, long long first, long long count,
    const float *__restrict__ raw, const float *__restrict__ yhat, float delta,
    const int *__restrict__ ppos, const int *__restrict__ epos,
    const uint8_t *__restrict__ inside, float *__restrict__ V,
    uint8_t *__restrict__ order) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const SlMap m = maps[map];
  const long long pair = first + i;
  const int ip = (int)(pair / m.ne), ie = (int)(pair - (long long)ip * m.ne);
  const long long at = m.v_off + pair;
  order[at] = 0xFF;
  if (!inside[m.tab_off + (long long)ppos[m.p_row + ip] * m.npos_e + epos[m.e_row + ie]]) {
    V[at] = m.exit_value;
    return;
  }
  const float phi = raw[11 * i + 8];
  V[at] = fminf(phi, yhat[i] + delta);
}

}  // namespace pe

using namespace pe;

extern "C" int pe_sl_pair_inputs(pe_sl *sl, int32_t map, int64_t first, int64_t count,
                                 float *raw) {
  clear_error();
  if (!sl || map < 0 || map >= sl->B || first < 0 || count < 0 || (count && !raw))
    return fail(PE_ERR_INVALID, "invalid argument");
  const SlMap &m = sl->maps[map];
  if (first + count > (long long)m.np * m.ne) return fail(PE_ERR_INVALID, "pair range");
  if (count == 0) return PE_OK;
  PE_HIP(hipSetDevice(sl->device));
  float *d = nullptr;
  PE_HIP(hipMalloc(&d, sizeof(float) * 11 * count));
  hipLaunchKernelGGL(pair_inputs, dim3((unsigned)((count + 255) / 256)), dim3(256), 0,
                     sl->stream, sl->d_maps, map, (long long)first, (long long)count,
                     sl->pstates, sl->estates, sl->pclear, sl->eclear, d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(raw, d, sizeof(float) * 11 * count,
                                          hipMemcpyDeviceToHost, sl->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(sl->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(PE_ERR_HIP, hipGetErrorString(e));
  return PE_OK;
}

extern "C" int pe_sl_order_from_mlp(pe_sl *sl, pe_mlp *mlp) {
  clear_error();
  if (!sl || !mlp) return fail(PE_ERR_INVALID, "invalid argument");
  if (mlp->shape.raw != 11) return fail(PE_ERR_INVALID, "the network needs extra = 3 inputs");
  if (sl->f64) return fail(PE_ERR_INVALID, "command orders need an FP32 solver handle");
  if (sl->device != mlp->device) return fail(PE_ERR_INVALID, "handles on different devices");
  PE_HIP(hipSetDevice(sl->device));
  PE_HIP(hipStreamSynchronize(sl->stream));
  const long long chunk = 1 << 18;
  int rc = mlp_reserve(mlp, chunk);
  if (rc) return rc;
  float *field = nullptr;  // [pairs] network field, then [pairs] probe output
  PE_HIP(hipMalloc(&field, sizeof(float) * 2 * (size_t)sl->pairs));
  auto body = [&]() -> int {
    for (int b = 0; b < sl->B; b++) {
      const SlMap &m = sl->maps[b];
      const long long n = (long long)m.np * m.ne;
      for (long long first = 0; first < n; first += chunk) {
        const long long count = std::min(chunk, n - first);
        const dim3 grid((unsigned)((count + 255) / 256));
        hipLaunchKernelGGL(pair_inputs, grid, dim3(256), 0, mlp->stream, sl->d_maps, b,
                           first, count, sl->pstates, sl->estates, sl->pclear, sl->eclear,
                           mlp->raw);
        PE_LAUNCHED(pair_inputs);
        int r = mlp_forward_loaded(mlp, count);
        if (r) return r;
        // min(phi, network) into the scratch field; orders reset to none
        hipLaunchKernelGGL(warm_combine, grid, dim3(256), 0, mlp->stream, sl->d_maps, b,
                           first, count, mlp->raw, mlp->yhat, 0.f, sl->ppos, sl->epos,
                           sl->inside, field, sl->order);
        PE_LAUNCHED(warm_combine);
      }
    }
    PE_HIP(hipStreamSynchronize(mlp->stream));
    return sl_probe_orders(sl, field, field + sl->pairs);
  };
  rc = body();
  (void)hipFree(field);
  return rc;
}

extern "C" int pe_sl_warm_start(pe_sl *sl, pe_mlp *mlp, float delta) {
  clear_error();
  if (!sl || !mlp || !std::isfinite(delta)) return fail(PE_ERR_INVALID, "invalid argument");
  if (mlp->shape.raw != 11) return fail(PE_ERR_INVALID, "the network needs extra = 3 inputs");
  if (sl->f64) return fail(PE_ERR_INVALID, "warm starts need an FP32 solver handle");
  if (sl->device != mlp->device) return fail(PE_ERR_INVALID, "handles on different devices");
  PE_HIP(hipSetDevice(sl->device));
  PE_HIP(hipStreamSynchronize(sl->stream));
  const long long chunk = 1 << 18;
  int rc = mlp_reserve(mlp, chunk);
  if (rc) return rc;
  for (int b = 0; b < sl->B; b++) {
    const SlMap &m = sl->maps[b];
    const long long n = (long long)m.np * m.ne;
    for (long long first = 0; first < n; first += chunk) {
      const long long count = std::min(chunk, n - first);
      const dim3 grid((unsigned)((count + 255) / 256));
      hipLaunchKernelGGL(pair_inputs, grid, dim3(256), 0, mlp->stream, sl->d_maps, b, first,
                         count, sl->pstates, sl->estates, sl->pclear, sl->eclear, mlp->raw);
      PE_LAUNCHED(pair_inputs);
      if ((rc = mlp_forward_loaded(mlp, count))) return rc;
      hipLaunchKernelGGL(warm_combine, grid, dim3(256), 0, mlp->stream, sl->d_maps, b, first,
                         count, mlp->raw, mlp->yhat, delta, sl->ppos, sl->epos, sl->inside,
                         sl->va, sl->order);
      PE_LAUNCHED(warm_combine);
    }
  }
  PE_HIP(hipStreamSynchronize(mlp->stream));
  // The sweep's second buffer must agree on the pairs it does not write.
  PE_HIP(hipMemcpyAsync(sl->vb, sl->va, sizeof(float) * sl->pairs, hipMemcpyDeviceToDevice,
                        sl->stream));
  PE_HIP(hipStreamSynchronize(sl->stream));
  sl->trace.clear();
  sl->sweeps = 0;
  sl->last = 0;
  return PE_OK;
}
