// Compact 8D semi-Lagrangian min-max backup of the contact game on one GPU.
//
// Semantics (local_game/grid_game.py:199-231 and :389-421, drivers
// :286-333 and :511-581): a Jacobi sweep maps V to
//   V'(ip,ie) = min(V(ip,ie), min_{u in A_P(ip)} max_{b in A_E(ie)} Q(u,b))
//   Q(u,b)    = sum_q wE[ie,b,q] * G_u(ip, idsE[ie,b,q])
//   G_u(ip,e) = sum_k wP[ip,u,k] * V(idsP[ip,u,k], e)
// for pairs whose centre distance is within the exit radius; the others
// keep the exit value.  The reference launches two kernels per P command
// and per chunk of 256 P states and stages every chunk through host memory;
// here one sweep is two launches over all maps of a batch, entirely in HBM:
//
//   sl_p_side  G_u(ip, e) for all u at once, one thread per (ip, e), e
//              fastest so the 16 V rows it gathers are read coalesced; the
//              P tables of (ip, u) are wave-uniform (scalar loads).  Entries
//              no inside pair can reach ("near" table) are skipped.
//   sl_e_side  one thread per pair (ip, ie), ie fastest: the min over P
//              commands of the max over E commands of the 16-point E
//              interpolation of G_u, with the reference's exact early exit
//              (a command whose running max reaches the running min cannot
//              lower it) and, optionally, the previous sweep's argmin /
//              argmax commands tried first so that exit comes sooner.  It
//              also max-reduces |V' - V| per map (the residual).
//   sl_e_side_lds  the same per pair, but one 1024-thread block per P state
//              first copies that state's G_u rows into LDS ([na][ne] floats,
//              94 KB at 2,612 E states), then walks the whole E row, so the
//              16-point E gathers hit LDS instead of one cache line per lane.
//              Used whenever the rows fit; bit-identical to sl_e_side.
//
// G holds [maps][u][ip][e] float32 (9 x the field) and is zero-filled once:
// entries skipped by sl_p_side are never read with a non-zero weight, and
// zero weights multiply finite values only, so no per-term branch is needed.
#include <cmath>
#include <cstring>
#include <vector>

#include "pe_internal.h"

namespace pe {




__global__ __launch_bounds__(256) void sl_initialize(
    const SlMap *__restrict__ maps, const double *__restrict__ pstates,
    const double *__restrict__ estates, const int *__restrict__ ppos,
    const int *__restrict__ epos, const uint8_t *__restrict__ inside,
    float *__restrict__ V, uint8_t *__restrict__ order) {
  const SlMap m = maps[blockIdx.z];
  const int ie = blockIdx.x * 256 + threadIdx.x;
  if (ie >= m.ne) return;
  for (int ip = blockIdx.y; ip < m.np; ip += gridDim.y) {
    const long long pair = m.v_off + (long long)ip * m.ne + ie;
    order[pair] = 0xFF;
    const int pp = ppos[m.p_row + ip], ep = epos[m.e_row + ie];
    if (!inside[m.tab_off + (long long)pp * m.npos_e + ep]) {
      V[pair] = m.exit_value;
      continue;
    }
    // Rectangle-contact payoff (local_game/geometry.py:15-36), in double.
    const double *a = pstates + 4ll * (m.p_row + ip);
    const double *b = estates + 4ll * (m.e_row + ie);
    const double dx = b[0] - a[0], dy = b[1] - a[1], rel = b[2] - a[2];
    const double ac = fabs(cos(rel)), as = fabs(sin(rel));
    const double cp = cos(a[2]), sp = sin(a[2]), ce = cos(b[2]), se = sin(b[2]);
    const double g0 = fabs(dx * cp + dy * sp) - .5 * (m.pl + m.el * ac + m.ew * as);
    const double g1 = fabs(-dx * sp + dy * cp) - .5 * (m.pw + m.el * as + m.ew * ac);
    const double g2 = fabs(dx * ce + dy * se) - .5 * (m.el + m.pl * ac + m.pw * as);
    const double g3 = fabs(-dx * se + dy * ce) - .5 * (m.ew + m.pl * as + m.pw * ac);
    V[pair] = float(fmax(fmax(g0, g1), fmax(g2, g3)));
  }
}

__global__ __launch_bounds__(256) void sl_p_side(
    const SlMap *__restrict__ maps, const float *__restrict__ V,
    float *__restrict__ G, const int *__restrict__ pids,
    const float *__restrict__ pw, const uint8_t *__restrict__ pallowed,
    const int *__restrict__ ppos, const int *__restrict__ epos,
    const uint8_t *__restrict__ near) {
  const SlMap m = maps[blockIdx.z];
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= m.ne) return;
  const float *__restrict__ Vm = V + m.v_off;
  const int ep = epos[m.e_row + e];
  for (int ip = blockIdx.y; ip < m.np; ip += gridDim.y) {
    const int row = m.p_row + ip;
    if (!near[m.tab_off + (long long)ppos[row] * m.npos_e + ep]) continue;
    for (int u = 0; u < m.na; u++) {
      const long long cmd = (long long)row * m.na + u;
      if (!pallowed[cmd]) continue;
      const int *id = pids + cmd * kStencil;
      const float *w = pw + cmd * kStencil;
      float acc = 0.f;
#pragma unroll
      for (int k = 0; k < kStencil; k++)
        acc = fmaf(w[k], Vm[(long long)id[k] * m.ne + e], acc);
      G[m.g_off + ((long long)u * m.np + ip) * m.ne + e] = acc;
    }
  }
}

// i-th command of the order that puts `first` in front (others ascending).
__device__ __forceinline__ int nth_command(int i, int first) {
  return i == 0 ? first : (i <= first ? i - 1 : i);
}

// Min over allowed P commands of the max over allowed E commands of one
// inside pair (ip, ie), starting from its current value; row(u) is the
// row G_u(ip, .).  Early exit: a P command whose running max over E
// commands reaches the running min cannot lower it (grid_game.py:409-421).
// Updates the stored command order of the pair.
template <bool EARLY, bool ORDERED, typename Row>
__device__ __forceinline__ float pair_backup(const SlMap &m, int prow, int erow, float start,
                                             long long pair, Row row,
                                             const int *__restrict__ eids,
                                             const float *__restrict__ ew,
                                             const uint8_t *__restrict__ eallowed,
                                             const uint8_t *__restrict__ pallowed,
                                             uint8_t *__restrict__ order) {
  float best = start;
  int ufirst = 0, bfirst = 0;
  if (ORDERED) {
    const int o = order[pair];
    if (o != 0xFF) {
      ufirst = o & 15;
      bfirst = o >> 4;
    }
  }
  int best_u = -1, best_b = -1;
  for (int i = 0; i < m.na; i++) {
    const int u = ORDERED ? nth_command(i, ufirst) : i;
    if (!pallowed[(long long)prow * m.na + u]) continue;
    const float *Gu = row(u);
    float worst = -1e30f;
    int arg = -1;
    bool cut = false;
    for (int j = 0; j < m.na; j++) {
      const int b = ORDERED ? nth_command(j, bfirst) : j;
      const long long cmd = (long long)erow * m.na + b;
      if (!eallowed[cmd]) continue;
      const int4 *id4 = reinterpret_cast<const int4 *>(eids + cmd * kStencil);
      const float4 *w4 = reinterpret_cast<const float4 *>(ew + cmd * kStencil);
      float q = 0.f;
#pragma unroll
      for (int c = 0; c < kStencil / 4; c++) {
        const int4 id = id4[c];
        const float4 w = w4[c];
        q = fmaf(w.x, Gu[id.x], q);
        q = fmaf(w.y, Gu[id.y], q);
        q = fmaf(w.z, Gu[id.z], q);
        q = fmaf(w.w, Gu[id.w], q);
      }
      if (q > worst) {
        worst = q;
        arg = b;
      }
      if (EARLY && worst >= best) {
        cut = true;  // this P command cannot lower the value
        break;
      }
    }
    if (!cut && worst < best) {
      best = worst;
      best_u = u;
      best_b = arg;
    }
  }
  if (ORDERED && best_u >= 0) order[pair] = (uint8_t)(best_u | (best_b << 4));
  return best;
}

// Block max of `change` into the map's residual slot (blockDim.x <= 1024).
__device__ __forceinline__ void block_residual(float change, unsigned long long *slot) {
  __shared__ float red[16];
  const float w = wave_max(change);
  const int waves = (int)(blockDim.x + 63) >> 6;
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = w;
  __syncthreads();
  if (threadIdx.x == 0) {
    float b = red[0];
    for (int k = 1; k < waves; k++) b = fmaxf(b, red[k]);
    raise_max(slot, (double)b);
  }
}

// E side with G gathered from global memory: one thread per pair.
template <bool EARLY, bool ORDERED>
__global__ __launch_bounds__(256) void sl_e_side(
    const SlMap *__restrict__ maps, const float *__restrict__ V,
    float *__restrict__ Vn, const float *__restrict__ G,
    const int *__restrict__ eids, const float *__restrict__ ew,
    const uint8_t *__restrict__ eallowed, const uint8_t *__restrict__ pallowed,
    const int *__restrict__ ppos, const int *__restrict__ epos,
    const uint8_t *__restrict__ inside, uint8_t *__restrict__ order,
    unsigned long long *__restrict__ slot) {
  const SlMap m = maps[blockIdx.z];
  const int ie = blockIdx.x * 256 + threadIdx.x;
  float change = 0.f;
  if (ie < m.ne) {
    const int erow = m.e_row + ie;
    const int ep = epos[erow];
    for (int ip = blockIdx.y; ip < m.np; ip += gridDim.y) {
      const long long pair = m.v_off + (long long)ip * m.ne + ie;
      const float start = V[pair];
      float best = start;
      const int prow = m.p_row + ip;
      if (inside[m.tab_off + (long long)ppos[prow] * m.npos_e + ep]) {
        auto row = [&](int u) { return G + m.g_off + ((long long)u * m.np + ip) * m.ne; };
        best = pair_backup<EARLY, ORDERED>(m, prow, erow, start, pair, row, eids, ew,
                                           eallowed, pallowed, order);
      }
      Vn[pair] = best;
      change = fmaxf(change, fabsf(best - start));
    }
  }
  block_residual(change, slot + blockIdx.z);
}

// E side with the G rows of one P state staged in LDS: one block per
// (P state, map) walks the whole E row, so the 16-point E gathers of every
// pair read LDS instead of scattered cache lines.  Dynamic LDS holds
// [na][ne] floats; only the rows of allowed P commands are filled, and only
// those are read.  Same arithmetic as sl_e_side, value for value.
template <bool EARLY, bool ORDERED>
__global__ __launch_bounds__(1024) void sl_e_side_lds(
    const SlMap *__restrict__ maps, const float *__restrict__ V,
    float *__restrict__ Vn, const float *__restrict__ G,
    const int *__restrict__ eids, const float *__restrict__ ew,
    const uint8_t *__restrict__ eallowed, const uint8_t *__restrict__ pallowed,
    const int *__restrict__ ppos, const int *__restrict__ epos,
    const uint8_t *__restrict__ inside, uint8_t *__restrict__ order,
    unsigned long long *__restrict__ slot) {
  extern __shared__ float tile[];
  const SlMap m = maps[blockIdx.z];
  const int ip = blockIdx.x;  // uniform per block
  const bool active = ip < m.np;
  const int prow = m.p_row + ip;
  if (active) {
    for (int u = 0; u < m.na; u++) {
      if (!pallowed[(long long)prow * m.na + u]) continue;
      const float *__restrict__ src = G + m.g_off + ((long long)u * m.np + ip) * m.ne;
      float *dst = tile + u * m.ne;
      for (int e = threadIdx.x; e < m.ne; e += blockDim.x) dst[e] = src[e];
    }
  }
  __syncthreads();
  float change = 0.f;
  if (active) {
    const int pp = ppos[prow];
    for (int ie = threadIdx.x; ie < m.ne; ie += blockDim.x) {
      const int erow = m.e_row + ie;
      const long long pair = m.v_off + (long long)ip * m.ne + ie;
      const float start = V[pair];
      float best = start;
      if (inside[m.tab_off + (long long)pp * m.npos_e + epos[erow]]) {
        auto row = [&](int u) { return tile + u * m.ne; };
        best = pair_backup<EARLY, ORDERED>(m, prow, erow, start, pair, row, eids, ew,
                                           eallowed, pallowed, order);
      }
      Vn[pair] = best;
      change = fmaxf(change, fabsf(best - start));
    }
  }
  block_residual(change, slot + blockIdx.z);
}

// Field sampled at continuous pair states through the product of the two
// cars' 16-vertex stencils (reference.py:127-135, query_field).  Pairs
// beyond the exit radius hold the exit value in the dense layout, which is
// what the packed reference returns for them.
__global__ __launch_bounds__(256) void sl_query(
    const float *__restrict__ V, int ne, long long n,
    const int *__restrict__ pi, const float *__restrict__ pw,
    const int *__restrict__ ei, const float *__restrict__ ew,
    float *__restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float total = 0.f;
  for (int a = 0; a < kStencil; a++) {
    const float wa = pw[kStencil * i + a];
    if (!(wa > 0.f)) continue;
    const float *row = V + (long long)pi[kStencil * i + a] * ne;
    float value = 0.f;
    for (int b = 0; b < kStencil; b++) {
      const float wb = ew[kStencil * i + b];
      if (wb > 0.f) value = fmaf(wb, row[ei[kStencil * i + b]], value);
    }
    total = fmaf(wa, value, total);
  }
  out[i] = total;
}

__global__ void sl_fill_u64(unsigned long long *p, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

__global__ void sl_fill_f32(float *p, long long n) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += stride)
    p[i] = 0.f;
}

}  // namespace pe

using namespace pe;

static dim3 pair_grid(const pe_sl *h) {
  const unsigned gy = (unsigned)std::min(h->np_max, 65535);
  return dim3((unsigned)((h->ne_max + 255) / 256), gy, (unsigned)h->B);
}

static void sl_record(pe_sl *h, int kind, bool start) {
  if (!h->profiling) return;
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return;
  hipEventRecord(e, h->stream);
  h->events.push_back(e);
  if (start) h->kinds.push_back(kind);
}

// Enqueues one sweep va -> vb writing its residuals to slot[0..B).
static void enqueue_sweep(pe_sl *h, const float *src, float *dst, int flags,
                          unsigned long long *slot, bool record) {
  const dim3 grid = pair_grid(h), block(256);
  if (record) sl_record(h, 0, true);
  hipLaunchKernelGGL(sl_p_side, grid, block, 0, h->stream, h->d_maps, src,
                     h->G, h->pids, h->pw, h->pallowed, h->ppos, h->epos,
                     h->near);
  if (record) sl_record(h, 0, false);
  if (record) sl_record(h, 1, true);
  const bool early = flags & PE_SL_EARLY_EXIT, ordered = flags & PE_SL_ORDERED;
  const bool lds = h->lds_bytes > 0 && !(flags & PE_SL_GLOBAL_GATHER);
  const dim3 lds_grid((unsigned)h->np_max, 1, (unsigned)h->B), lds_block(1024);
#define PE_E_SIDE(E, O)                                                           \
  if (lds)                                                                        \
    hipLaunchKernelGGL((sl_e_side_lds<E, O>), lds_grid, lds_block, h->lds_bytes,  \
                       h->stream, h->d_maps, src, dst, h->G, h->eids, h->ew,      \
                       h->eallowed, h->pallowed, h->ppos, h->epos, h->inside,     \
                       h->order, slot);                                           \
  else                                                                            \
    hipLaunchKernelGGL((sl_e_side<E, O>), grid, block, 0, h->stream, h->d_maps,   \
                       src, dst, h->G, h->eids, h->ew, h->eallowed, h->pallowed,  \
                       h->ppos, h->epos, h->inside, h->order, slot)
  if (early && ordered) PE_E_SIDE(true, true);
  else if (early) PE_E_SIDE(true, false);
  else if (ordered) PE_E_SIDE(false, true);
  else PE_E_SIDE(false, false);
#undef PE_E_SIDE
  if (record) sl_record(h, 1, false);
}

static int sl_flush(pe_sl *h) {
  if (h->ring_used == 0) return PE_OK;
  std::vector<unsigned long long> host((size_t)h->ring_used * h->B);
  PE_HIP(hipMemcpyAsync(host.data(), h->ring, host.size() * 8,
                        hipMemcpyDeviceToHost, h->stream));
  PE_HIP(hipStreamSynchronize(h->stream));
  for (int s = 0; s < h->ring_used; s++) {
    double worst = 0;
    for (int b = 0; b < h->B; b++) {
      double v;
      std::memcpy(&v, &host[(size_t)s * h->B + b], 8);
      h->trace.push_back(v);
      worst = std::max(worst, v);
    }
    h->last = worst;
  }
  const long long n = (long long)h->ring_used * h->B;
  hipLaunchKernelGGL(sl_fill_u64, dim3((unsigned)((n + 255) / 256)), dim3(256),
                     0, h->stream, h->ring, n);
  PE_LAUNCHED(sl_fill_u64);
  h->ring_used = 0;
  if (h->profiling) {
    for (size_t k = 0; k + 1 < h->events.size(); k += 2) {
      float ms = 0;
      PE_HIP(hipEventElapsedTime(&ms, h->events[k], h->events[k + 1]));
      h->ms[h->kinds[k / 2]] += ms;
      h->launches[h->kinds[k / 2]] += 1;
    }
    for (hipEvent_t e : h->events) hipEventDestroy(e);
    h->events.clear();
    h->kinds.clear();
  }
  return PE_OK;
}

static int build_graph(pe_sl *h, int flags) {
  if (h->graph && h->graph_flags == flags && h->graph_src == h->va) return PE_OK;
  if (h->graph) {
    hipGraphExecDestroy(h->graph);
    h->graph = nullptr;
  }
  hipGraph_t g;
  PE_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  // Zero the graph's residual slots with a kernel (no memset node).
  const long long n = (long long)h->graph_sweeps * h->B;
  hipLaunchKernelGGL(sl_fill_u64, dim3((unsigned)((n + 255) / 256)), dim3(256),
                     0, h->stream, h->gslots, n);
  for (int s = 0; s < h->graph_sweeps; s++) {
    const bool even = (s % 2) == 0;
    enqueue_sweep(h, even ? h->va : h->vb, even ? h->vb : h->va, flags,
                  h->gslots + (long long)s * h->B, false);
  }
  hipError_t end = hipStreamEndCapture(h->stream, &g);
  if (end != hipSuccess) {
    (void)hipGetLastError();
    return fail(PE_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(end));
  }
  hipError_t inst = hipGraphInstantiate(&h->graph, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (inst != hipSuccess) {
    (void)hipGetLastError();
    h->graph = nullptr;
    return fail(PE_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(inst));
  }
  h->graph_flags = flags;
  h->graph_src = h->va;
  return PE_OK;
}

static int do_sweeps(pe_sl *h, int count, int flags) {
  PE_HIP(hipEventRecord(h->t0, h->stream));
  int done = 0;
  if ((flags & PE_SL_GRAPH) && !h->profiling) {
    int rc = build_graph(h, flags & ~PE_SL_GRAPH);
    if (rc) return rc;
    while (count - done >= h->graph_sweeps) {
      if (h->ring_used + h->graph_sweeps > h->ring_slots) {
        rc = sl_flush(h);
        if (rc) return rc;
      }
      PE_HIP(hipGraphLaunch(h->graph, h->stream));
      PE_HIP(hipMemcpyAsync(h->ring + (long long)h->ring_used * h->B, h->gslots,
                            (size_t)h->graph_sweeps * h->B * 8,
                            hipMemcpyDeviceToDevice, h->stream));
      h->ring_used += h->graph_sweeps;
      done += h->graph_sweeps;
    }
  }
  for (; done < count; done++) {
    if (h->ring_used == h->ring_slots) {
      int rc = sl_flush(h);
      if (rc) return rc;
    }
    enqueue_sweep(h, h->va, h->vb, flags & ~PE_SL_GRAPH,
                  h->ring + (long long)h->ring_used * h->B, true);
    PE_LAUNCHED(sl_sweep);
    std::swap(h->va, h->vb);
    h->ring_used++;
  }
  h->sweeps += count;
  PE_HIP(hipEventRecord(h->t1, h->stream));
  int rc = sl_flush(h);
  if (rc) return rc;
  float ms = 0;
  PE_HIP(hipEventSynchronize(h->t1));
  PE_HIP(hipEventElapsedTime(&ms, h->t0, h->t1));
  h->update_seconds += ms * 1e-3;
  return PE_OK;
}

extern "C" int pe_sl_create(const pe_sl_game *games, int32_t n_maps,
                            pe_sl **out) {
  clear_error();
  if (!games || !out || n_maps < 1 || n_maps > 65535)
    return fail(PE_ERR_INVALID, "invalid map list");
  *out = nullptr;
  const int na = games[0].p.actions;
  if (na < 1 || na > kMaxActions)
    return fail(PE_ERR_INVALID, "1..15 commands per state required");
  for (int b = 0; b < n_maps; b++) {
    const pe_sl_game &g = games[b];
    for (const pe_sl_car *c : {&g.p, &g.e}) {
      if (c->n < 1 || c->actions != na || c->positions < 1 || !c->states ||
          !c->successors || !c->weights || !c->allowed || !c->position)
        return fail(PE_ERR_INVALID, "invalid car tables (map " + std::to_string(b) + ")");
      for (int i = 0; i < c->n; i++)
        if (c->position[i] < 0 || c->position[i] >= c->positions)
          return fail(PE_ERR_INVALID, "position index out of range");
      const long long m = (long long)c->n * na * kStencil;
      for (long long k = 0; k < m; k++)
        if (c->successors[k] < 0 || c->successors[k] >= c->n ||
            !(c->weights[k] >= 0.f) || !std::isfinite(c->weights[k]))
          return fail(PE_ERR_INVALID, "successor id or weight out of range");
    }
    if (!g.inside || !g.near || !std::isfinite(g.exit_value))
      return fail(PE_ERR_INVALID, "inside / near tables and exit value required");
  }
  int dev = 0;
  int rc = require_device(&dev);
  if (rc) return rc;
  pe_sl *h = new pe_sl();
  h->device = dev;
  h->B = n_maps;
  h->na = na;
  long long prow = 0, erow = 0, voff = 0, goff = 0, toff = 0;
  for (int b = 0; b < n_maps; b++) {
    const pe_sl_game &g = games[b];
    SlMap m{};
    m.np = g.p.n;
    m.ne = g.e.n;
    m.na = na;
    m.npos_e = g.e.positions;
    m.v_off = voff;
    m.g_off = goff;
    m.tab_off = toff;
    m.p_row = (int)prow;
    m.e_row = (int)erow;
    m.exit_value = g.exit_value;
    m.pl = g.p_length;
    m.pw = g.p_width;
    m.el = g.e_length;
    m.ew = g.e_width;
    h->maps.push_back(m);
    voff += (long long)m.np * m.ne;
    goff += (long long)na * m.np * m.ne;
    toff += (long long)g.p.positions * g.e.positions;
    prow += m.np;
    erow += m.ne;
    h->np_max = std::max(h->np_max, m.np);
    h->ne_max = std::max(h->ne_max, m.ne);
    for (int i = 0; i < m.np; i++)
      for (int j = 0; j < m.ne; j++)
        h->active += g.inside[(long long)g.p.position[i] * g.e.positions +
                              g.e.position[j]] != 0;
  }
  if (prow > INT32_MAX || erow > INT32_MAX) {
    delete h;
    return fail(PE_ERR_INVALID, "too many states");
  }
  h->pairs = voff;
  h->g_size = goff;
  h->graph_sweeps = 2;
  auto cleanup = [&](int code) {
    pe_sl_destroy(h);
    return code;
  };
  auto body = [&]() -> int {
    PE_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    // LDS E side when the G rows of one P state fit a workgroup's LDS.
    {
      int limit = 0;
      PE_HIP(hipDeviceGetAttribute(&limit, hipDeviceAttributeMaxSharedMemoryPerBlock, h->device));
      const size_t bytes = sizeof(float) * (size_t)na * (size_t)h->ne_max;
      if (bytes + 256 <= (size_t)limit) {
        // If the runtime refuses the LDS size for any variant, the global-gather
        // E side is used instead.
        bool ok = true;
        for (const void *f : {(const void *)sl_e_side_lds<true, true>,
                              (const void *)sl_e_side_lds<true, false>,
                              (const void *)sl_e_side_lds<false, true>,
                              (const void *)sl_e_side_lds<false, false>})
          ok = ok && hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)bytes) == hipSuccess;
        if (ok) h->lds_bytes = bytes;
        else (void)hipGetLastError();
      }
    }
    PE_HIP(hipEventCreate(&h->t0));
    PE_HIP(hipEventCreate(&h->t1));
    PE_HIP(hipMalloc(&h->d_maps, sizeof(SlMap) * n_maps));
    PE_HIP(hipMemcpy(h->d_maps, h->maps.data(), sizeof(SlMap) * n_maps,
                     hipMemcpyHostToDevice));
    auto cat = [&](auto **dst, auto getter, long long per_row, bool is_p) -> int {
      using T = std::remove_const_t<std::remove_pointer_t<decltype(getter(games[0].p))>>;
      const long long rows = is_p ? prow : erow;
      PE_HIP(hipMalloc((void **)dst, sizeof(T) * rows * per_row));
      long long at = 0;
      for (int b = 0; b < n_maps; b++) {
        const pe_sl_car &c = is_p ? games[b].p : games[b].e;
        PE_HIP(hipMemcpy((T *)*dst + at * per_row, getter(c),
                         sizeof(T) * c.n * per_row, hipMemcpyHostToDevice));
        at += c.n;
      }
      return PE_OK;
    };
    int r;
    for (int side = 0; side < 2; side++) {
      const bool P = side == 0;
      if ((r = cat(P ? &h->pids : &h->eids, [](const pe_sl_car &c) { return c.successors; },
                   (long long)na * kStencil, P)) ||
          (r = cat(P ? &h->pw : &h->ew, [](const pe_sl_car &c) { return c.weights; },
                   (long long)na * kStencil, P)) ||
          (r = cat(P ? &h->pallowed : &h->eallowed, [](const pe_sl_car &c) { return c.allowed; },
                   (long long)na, P)) ||
          (r = cat(P ? &h->pstates : &h->estates, [](const pe_sl_car &c) { return c.states; },
                   4, P)) ||
          (r = cat(P ? &h->ppos : &h->epos, [](const pe_sl_car &c) { return c.position; },
                   1, P)))
        return r;
      // Per-state clearance as float (zeros where not given).
      std::vector<float> clear;
      for (int b = 0; b < n_maps; b++) {
        const pe_sl_car &c = P ? games[b].p : games[b].e;
        for (int i = 0; i < c.n; i++) clear.push_back(c.clearance ? (float)c.clearance[i] : 0.f);
      }
      float **dst = P ? &h->pclear : &h->eclear;
      PE_HIP(hipMalloc((void **)dst, sizeof(float) * std::max<size_t>(clear.size(), 1)));
      PE_HIP(hipMemcpy(*dst, clear.data(), sizeof(float) * clear.size(), hipMemcpyHostToDevice));
    }
    PE_HIP(hipMalloc((void **)&h->inside, std::max(toff, 1ll)));
    PE_HIP(hipMalloc((void **)&h->near, std::max(toff, 1ll)));
    long long at = 0;
    for (int b = 0; b < n_maps; b++) {
      const long long sz = (long long)games[b].p.positions * games[b].e.positions;
      PE_HIP(hipMemcpy(h->inside + at, games[b].inside, sz, hipMemcpyHostToDevice));
      PE_HIP(hipMemcpy(h->near + at, games[b].near, sz, hipMemcpyHostToDevice));
      at += sz;
    }
    PE_HIP(hipMalloc((void **)&h->va, sizeof(float) * h->pairs));
    PE_HIP(hipMalloc((void **)&h->vb, sizeof(float) * h->pairs));
    PE_HIP(hipMalloc((void **)&h->order, h->pairs));
    PE_HIP(hipMalloc((void **)&h->G, sizeof(float) * h->g_size));
    PE_HIP(hipMalloc((void **)&h->ring, 8ll * h->ring_slots * n_maps));
    PE_HIP(hipMalloc((void **)&h->gslots, 8ll * h->graph_sweeps * n_maps));
    hipLaunchKernelGGL(sl_fill_f32, dim3(4096), dim3(256), 0, h->stream, h->G,
                       h->g_size);
    PE_LAUNCHED(sl_fill_f32);
    const long long nr = 1ll * h->ring_slots * n_maps;
    hipLaunchKernelGGL(sl_fill_u64, dim3((unsigned)((nr + 255) / 256)), dim3(256),
                       0, h->stream, h->ring, nr);
    PE_LAUNCHED(sl_fill_u64);
    PE_HIP(hipStreamSynchronize(h->stream));
    return PE_OK;
  };
  rc = body();
  if (rc) return cleanup(rc);
  rc = pe_sl_initialize(h, nullptr);
  if (rc) return cleanup(rc);
  *out = h;
  return PE_OK;
}

extern "C" int pe_sl_initialize(pe_sl *h, const float *v0) {
  clear_error();
  if (!h) return fail(PE_ERR_INVALID, "null handle");
  PE_HIP(hipSetDevice(h->device));
  hipLaunchKernelGGL(sl_initialize, pair_grid(h), dim3(256), 0, h->stream,
                     h->d_maps, h->pstates, h->estates, h->ppos, h->epos,
                     h->inside, h->va, h->order);
  PE_LAUNCHED(sl_initialize);
  if (v0) {
    for (long long k = 0; k < h->pairs; k++)
      if (!std::isfinite(v0[k])) return fail(PE_ERR_INVALID, "non-finite initial field");
    PE_HIP(hipMemcpyAsync(h->va, v0, sizeof(float) * h->pairs,
                          hipMemcpyHostToDevice, h->stream));
  }
  PE_HIP(hipStreamSynchronize(h->stream));
  h->trace.clear();
  h->sweeps = 0;
  h->last = 0;
  return PE_OK;
}

extern "C" int pe_sl_sweep(pe_sl *h, int32_t sweeps, int32_t flags) {
  clear_error();
  if (!h || sweeps < 0) return fail(PE_ERR_INVALID, "invalid argument");
  PE_HIP(hipSetDevice(h->device));
  return do_sweeps(h, sweeps, flags);
}

extern "C" int pe_sl_run(pe_sl *h, double tol, int32_t max_sweeps,
                         int32_t check_every, int32_t flags,
                         int32_t *sweeps_done, int32_t *converged) {
  clear_error();
  if (!h || !(tol >= 0) || max_sweeps < 0 || check_every < 1)
    return fail(PE_ERR_INVALID, "invalid argument");
  PE_HIP(hipSetDevice(h->device));
  int done = 0, ok = 0;
  while (done < max_sweeps) {
    const int chunk = std::min(check_every, max_sweeps - done);
    int rc = do_sweeps(h, chunk, flags);
    if (rc) return rc;
    done += chunk;
    if (h->last <= tol) {
      ok = 1;
      break;
    }
  }
  if (sweeps_done) *sweeps_done = done;
  if (converged) *converged = ok;
  return PE_OK;
}

extern "C" int pe_sl_read(pe_sl *h, int32_t map, float *field) {
  clear_error();
  if (!h || !field || map < 0 || map >= h->B)
    return fail(PE_ERR_INVALID, "invalid argument");
  PE_HIP(hipSetDevice(h->device));
  const SlMap &m = h->maps[map];
  PE_HIP(hipMemcpyAsync(field, h->va + m.v_off, sizeof(float) * m.np * m.ne,
                        hipMemcpyDeviceToHost, h->stream));
  PE_HIP(hipStreamSynchronize(h->stream));
  return PE_OK;
}

extern "C" int pe_sl_query(pe_sl *h, int32_t map, int64_t n,
                           const int32_t *p_ids, const float *p_weights,
                           const int32_t *e_ids, const float *e_weights,
                           float *out) {
  clear_error();
  if (!h || map < 0 || map >= h->B || n < 0 || (n && (!p_ids || !p_weights ||
                                                      !e_ids || !e_weights || !out)))
    return fail(PE_ERR_INVALID, "invalid argument");
  const SlMap &m = h->maps[map];
  for (long long k = 0; k < n * kStencil; k++)
    if (p_ids[k] < 0 || p_ids[k] >= m.np || e_ids[k] < 0 || e_ids[k] >= m.ne)
      return fail(PE_ERR_INVALID, "query stencil id out of range");
  if (n == 0) return PE_OK;
  PE_HIP(hipSetDevice(h->device));
  void *buf = nullptr;
  const size_t ib = sizeof(int) * n * kStencil, fb = sizeof(float) * n * kStencil;
  PE_HIP(hipMalloc(&buf, 2 * ib + 2 * fb + sizeof(float) * n));
  char *c = (char *)buf;
  int *d_pi = (int *)c, *d_ei = (int *)(c + ib);
  float *d_pw = (float *)(c + 2 * ib), *d_ew = (float *)(c + 2 * ib + fb);
  float *d_out = (float *)(c + 2 * ib + 2 * fb);
  auto body = [&]() -> int {
    PE_HIP(hipMemcpyAsync(d_pi, p_ids, ib, hipMemcpyHostToDevice, h->stream));
    PE_HIP(hipMemcpyAsync(d_ei, e_ids, ib, hipMemcpyHostToDevice, h->stream));
    PE_HIP(hipMemcpyAsync(d_pw, p_weights, fb, hipMemcpyHostToDevice, h->stream));
    PE_HIP(hipMemcpyAsync(d_ew, e_weights, fb, hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(sl_query, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       h->stream, (const float *)(h->va + m.v_off), m.ne,
                       (long long)n, d_pi, d_pw, d_ei, d_ew, d_out);
    PE_LAUNCHED(sl_query);
    PE_HIP(hipMemcpyAsync(out, d_out, sizeof(float) * n, hipMemcpyDeviceToHost,
                          h->stream));
    PE_HIP(hipStreamSynchronize(h->stream));
    return PE_OK;
  };
  const int rc = body();
  (void)hipFree(buf);
  return rc;
}

extern "C" int pe_sl_trace(pe_sl *h, double *out, int64_t capacity,
                           int64_t *count) {
  clear_error();
  if (!h) return fail(PE_ERR_INVALID, "null handle");
  const int64_t n = (int64_t)h->trace.size();
  if (out)
    for (int64_t k = 0; k < std::min(n, capacity); k++) out[k] = h->trace[k];
  if (count) *count = n;
  return PE_OK;
}

extern "C" int pe_sl_get_stats(pe_sl *h, pe_sl_stats *s) {
  clear_error();
  if (!h || !s) return fail(PE_ERR_INVALID, "invalid argument");
  std::memset(s, 0, sizeof(*s));
  s->sweeps = h->sweeps;
  s->maps = h->B;
  s->lds_e_side = h->lds_bytes > 0;
  s->pairs = h->pairs;
  s->active_pairs = h->active;
  s->update_seconds = h->update_seconds;
  s->last_residual = h->last;
  s->p_ms = h->ms[0];
  s->e_ms = h->ms[1];
  s->p_launches = h->launches[0];
  s->e_launches = h->launches[1];
  return PE_OK;
}

extern "C" int pe_sl_set_profiling(pe_sl *h, int32_t on) {
  clear_error();
  if (!h) return fail(PE_ERR_INVALID, "null handle");
  h->profiling = on != 0;
  h->ms[0] = h->ms[1] = 0;
  h->launches[0] = h->launches[1] = 0;
  return PE_OK;
}

extern "C" void pe_sl_destroy(pe_sl *h) {
  if (!h) return;
  hipSetDevice(h->device);
  if (h->stream) hipStreamSynchronize(h->stream);
  if (h->graph) hipGraphExecDestroy(h->graph);
  for (hipEvent_t e : h->events) hipEventDestroy(e);
  for (void *p : {(void *)h->d_maps, (void *)h->pids, (void *)h->eids,
                  (void *)h->ppos, (void *)h->epos, (void *)h->pw, (void *)h->ew,
                  (void *)h->pallowed, (void *)h->eallowed, (void *)h->inside,
                  (void *)h->near, (void *)h->order, (void *)h->pstates,
                  (void *)h->estates, (void *)h->pclear, (void *)h->eclear,
                  (void *)h->va, (void *)h->vb, (void *)h->G,
                  (void *)h->ring, (void *)h->gslots})
    if (p) hipFree(p);
  if (h->t0) hipEventDestroy(h->t0);
  if (h->t1) hipEventDestroy(h->t1);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
}
