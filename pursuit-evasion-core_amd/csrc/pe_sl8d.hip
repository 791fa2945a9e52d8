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
//   sl_p_side  G_u(ip, e) for all u, blocks of (ip, strided chunk of the
//              list of near E states of ip's position): only entries an
//              inside pair can reach are computed, e ascending so the V rows
//              are read nearly coalesced; the P tables of (ip, u) are
//              block-uniform (scalar loads) and compacted to their non-zero
//              weights (4-, 8- or 16-term sums).
//   sl_e_side  one thread per pair (ip, ie), ie fastest: the min over P
//              commands of the max over E commands of the 16-point E
//              interpolation of G_u, with the reference's exact early exit
//              (a command whose running max reaches the running min cannot
//              lower it) and, optionally, the previous sweep's argmin /
//              argmax commands tried first so that exit comes sooner.  It
//              also max-reduces |V' - V| per map (the residual).
//   sl_e_side_lds  the same per pair, but one 1024-thread block per P state
//              first copies that state's G_u rows into LDS ([na][ne] floats,
//              94 KB at 2,612 E states), then walks the list of inside E
//              states of the P state's position, so the 16-point E gathers
//              hit LDS instead of one cache line per lane.  Used whenever
//              the rows fit; bit-identical to sl_e_side.
//
// Both E sides read each (E state, command) stencil as one packed 96-byte
// record (16-bit ids, then the weights) when every map has fewer than
// 65,536 E states, and visit the commands in one of two orders (one pass
// over the E commands for all P commands, or u-major), chosen per handle
// from the LDS occupancy; all choices give the same values.
//
// G holds [maps][u][ip][e] float32 (9 x the field) and is zero-filled once:
// entries skipped by sl_p_side are never read with a non-zero weight, and
// zero weights multiply finite values only, so no per-term branch is needed.
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "sl8d_kernels.h"

namespace pe {




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

// One sweep va -> vb in the handle's precision.
static void enqueue_sweep(pe_sl *h, const float *src, float *dst, int flags,
                          unsigned long long *slot, bool record) {
  if (h->f64)
    sl_enqueue_f64(h, src, dst, flags, slot, record);
  else
    enqueue_sweep_t<float>(h, src, dst, flags, slot, record);
}

namespace pe {
// Records in h->order, for every inside pair of every active map, the
// argmin P command and the argmax E command of one backup evaluated on
// `field` (device memory in the handle's dense layout, FP32 handles): the
// E side starts each pair from +3e38, so every pair records its commands.
// The backup's values go to `scratch` and are dropped; the handle's value
// buffers, trace and sweep count are not touched.
int sl_probe_orders(pe_sl *h, const float *field, float *scratch) {
  h->probe_orders = true;
  enqueue_sweep(h, field, scratch, PE_SL_EARLY_EXIT | PE_SL_ORDERED, h->gslots, false);
  h->probe_orders = false;
  PE_HIP(hipGetLastError());
  PE_HIP(hipStreamSynchronize(h->stream));
  return PE_OK;
}
}  // namespace pe

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

static int build_graph(pe_sl *h, int flags, int gi) {
  hipGraphExec_t &exec = h->graph[gi];
  const int sweeps = h->graph_sweeps[gi];
  if (exec && h->graph_flags[gi] == flags && h->graph_src[gi] == h->va) return PE_OK;
  if (exec) {
    hipGraphExecDestroy(exec);
    exec = nullptr;
  }
  hipGraph_t g;
  PE_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
  // Zero the graph's residual slots with a kernel (no memset node).
  const long long n = (long long)sweeps * h->B;
  hipLaunchKernelGGL(sl_fill_u64, dim3((unsigned)((n + 255) / 256)), dim3(256),
                     0, h->stream, h->gslots, n);
  for (int s = 0; s < sweeps; s++) {
    const bool even = (s % 2) == 0;
    enqueue_sweep(h, even ? h->va : h->vb, even ? h->vb : h->va, flags,
                  h->gslots + (long long)s * h->B, false);
  }
  hipError_t end = hipStreamEndCapture(h->stream, &g);
  if (end != hipSuccess) {
    (void)hipGetLastError();
    return fail(PE_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(end));
  }
  hipError_t inst = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (inst != hipSuccess) {
    (void)hipGetLastError();
    exec = nullptr;
    return fail(PE_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(inst));
  }
  h->graph_flags[gi] = flags;
  h->graph_src[gi] = h->va;
  return PE_OK;
}

static void drop_graphs(pe_sl *h) {
  for (hipGraphExec_t &g : h->graph)
    if (g) {
      hipGraphExecDestroy(g);
      g = nullptr;
    }
}

static int do_sweeps(pe_sl *h, int count, int flags) {
  PE_HIP(hipEventRecord(h->t0, h->stream));
  int done = 0;
  if ((flags & PE_SL_GRAPH) && !h->profiling) {
    for (int gi = 0; gi < pe_sl::kGraphs; gi++) {
      const int sweeps = h->graph_sweeps[gi];
      if (count - done < sweeps) continue;
      int rc = build_graph(h, flags & ~PE_SL_GRAPH, gi);
      if (rc) return rc;
      while (count - done >= sweeps) {
        if (h->ring_used + sweeps > h->ring_slots) {
          rc = sl_flush(h);
          if (rc) return rc;
        }
        PE_HIP(hipGraphLaunch(h->graph[gi], h->stream));
        PE_HIP(hipMemcpyAsync(h->ring + (long long)h->ring_used * h->B, h->gslots,
                              (size_t)sweeps * h->B * 8, hipMemcpyDeviceToDevice, h->stream));
        h->ring_used += sweeps;
        done += sweeps;
      }
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

// Packed E records (16-bit byte offsets into a G row) when every map's row
// of E values spans at most 64 KiB.
// Packed E stencil records of every map, for the LDS E side: per (state,
// command) the 16 vertices as signed 16-bit byte offsets in the interleaved
// tile ([e][na] values of `esize` bytes) relative to the state's own
// values, then the 16 weights.  A vertex of zero weight, and every vertex of
// a command the state does not allow (never evaluated), gets offset 0.
// False when some offset does not fit 16 bits: then the E side reads the
// 32-bit id tables.
static bool pack_erecs(const pe_sl_game *games, int n_maps, int na, size_t esize,
                       std::vector<ERec> &out) {
  out.clear();
  for (int b = 0; b < n_maps; b++) {
    const pe_sl_car &c = games[b].e;
    for (long long r = 0; r < (long long)c.n * na; r++) {
      const long long self = r / na;
      const bool allowed = c.allowed[r] != 0;
      ERec rec;
      unsigned short off[kStencil];
      float *w = reinterpret_cast<float *>(&rec.w[0]);
      for (int k = 0; k < kStencil; k++) {
        const float wk = c.weights[r * kStencil + k];
        w[k] = wk;
        long long d = 0;
        if (allowed && wk != 0.f) d = ((long long)c.successors[r * kStencil + k] - self) * na *
                                      (long long)esize;
        if (d < -32768 || d > 32767) {
          out.clear();
          return false;
        }
        off[k] = (unsigned short)(short)d;
      }
      unsigned *ids = reinterpret_cast<unsigned *>(&rec.ids[0]);
      for (int k = 0; k < kStencil / 2; k++)
        ids[k] = (unsigned)off[2 * k] | ((unsigned)off[2 * k + 1] << 16);
      out.push_back(rec);
    }
  }
  return true;
}

extern "C" int pe_sl_create(const pe_sl_game *games, int32_t n_maps,
                            pe_sl **out) {
  return pe_sl_create_ex(games, n_maps, 0, out);
}

extern "C" int pe_sl_create_ex(const pe_sl_game *games, int32_t n_maps, int32_t precision,
                               pe_sl **out) {
  clear_error();
  if (!games || !out || n_maps < 1 || n_maps > 65535)
    return fail(PE_ERR_INVALID, "invalid map list");
  if (precision != 0 && precision != 1)
    return fail(PE_ERR_INVALID, "precision must be 0 (FP32) or 1 (FP64)");
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
  h->f64 = precision == 1;
  h->esize = h->f64 ? sizeof(double) : sizeof(float);
  h->B = n_maps;
  h->na = na;
  long long prow = 0, erow = 0, voff = 0, goff = 0, toff = 0;
  std::vector<int> ilist, ioff, nlist, noff;
  long long lds_slots[3] = {0, 0, 0};
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
    m.ipos_off = (int)ioff.size();
    m.active = 1;
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
    std::vector<int> count(g.p.positions, 0);
    for (int pp = 0; pp < g.p.positions; pp++) {
      ioff.push_back((int)std::min<size_t>(ilist.size(), INT32_MAX));
      noff.push_back((int)std::min<size_t>(nlist.size(), INT32_MAX));
      for (int j = 0; j < m.ne; j++) {
        const long long t = (long long)pp * g.e.positions + g.e.position[j];
        if (g.inside[t]) {
          ilist.push_back(j);
          count[pp]++;
        }
        if (g.near[t]) nlist.push_back(j);
      }
    }
    for (int i = 0; i < m.np; i++) {
      const int c = count[g.p.position[i]];
      h->active += c;
      for (int k = 0; k < 3; k++) {  // lane slots of E-side blocks of 256 << k threads
        const long long t = 256ll << k;
        lds_slots[k] += (c + t - 1) / t * t;
      }
    }
  }
  ioff.push_back((int)std::min<size_t>(ilist.size(), INT32_MAX));
  noff.push_back((int)std::min<size_t>(nlist.size(), INT32_MAX));
  if (prow > INT32_MAX || erow > INT32_MAX || ilist.size() >= (size_t)INT32_MAX ||
      nlist.size() >= (size_t)INT32_MAX) {
    delete h;
    return fail(PE_ERR_INVALID, "too many states");
  }
  h->pairs = voff;
  h->g_size = goff;
  auto cleanup = [&](int code) {
    pe_sl_destroy(h);
    return code;
  };
  std::vector<ERec> erecs;
  const bool packed = pack_erecs(games, n_maps, na, h->esize, erecs);
  auto body = [&]() -> int {
    PE_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    // LDS E side when the G rows of one P state fit a workgroup's LDS.
    {
      int limit = 0;
      PE_HIP(hipDeviceGetAttribute(&limit, hipDeviceAttributeMaxSharedMemoryPerBlock, h->device));
      const size_t bytes = h->esize * (size_t)na * (size_t)h->ne_max;
      // If the runtime refuses the LDS size, the global-gather E side is used.
      if (bytes + 256 <= (size_t)limit &&
          (h->f64 ? sl_allow_lds_f64(bytes) : sl_allow_lds<float>(bytes))) {
        h->lds_bytes = bytes;
        if (!h->f64) {
          double fill[3];
          for (int k = 0; k < 3; k++)
            fill[k] = lds_slots[k] ? (double)h->active / (double)lds_slots[k] : 1.0;
          h->lds_threads = sl_lds_threads(bytes, fill, packed, h->na <= 9);
        }
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
      // Allowed commands of each state as one bit mask.
      std::vector<uint16_t> mask;
      for (int b = 0; b < n_maps; b++) {
        const pe_sl_car &c = P ? games[b].p : games[b].e;
        for (int i = 0; i < c.n; i++) {
          unsigned bits = 0;
          for (int u = 0; u < na; u++)
            if (c.allowed[(long long)i * na + u]) bits |= 1u << u;
          mask.push_back((uint16_t)bits);
        }
      }
      uint16_t **mdst = P ? &h->pmask : &h->emask;
      PE_HIP(hipMalloc((void **)mdst, sizeof(uint16_t) * std::max<size_t>(mask.size(), 1)));
      PE_HIP(hipMemcpy(*mdst, mask.data(), sizeof(uint16_t) * mask.size(),
                       hipMemcpyHostToDevice));
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
    {
      // Compacted P stencils (see sl_p_side).
      const size_t n = (size_t)prow * na;
      std::vector<int> ids(n * kStencil);
      std::vector<float> wts(n * kStencil);
      std::vector<uint8_t> cnt(n);
      size_t at = 0;
      for (int b = 0; b < n_maps; b++) {
        const pe_sl_car &c = games[b].p;
        for (long long r = 0; r < (long long)c.n * na; r++, at++) {
          const int *si = c.successors + r * kStencil;
          const float *sw = c.weights + r * kStencil;
          int k = 0;
          for (int q = 0; q < kStencil; q++)
            if (sw[q] != 0.f) {
              ids[at * kStencil + k] = si[q];
              wts[at * kStencil + k] = sw[q];
              k++;
            }
          cnt[at] = (uint8_t)k;
          for (int q = k; q < kStencil; q++) {  // same row as the last term
            ids[at * kStencil + q] = k ? ids[at * kStencil + k - 1] : si[0];
            wts[at * kStencil + q] = 0.f;
          }
        }
      }
      PE_HIP(hipMemcpy(h->pids, ids.data(), sizeof(int) * ids.size(), hipMemcpyHostToDevice));
      PE_HIP(hipMemcpy(h->pw, wts.data(), sizeof(float) * wts.size(), hipMemcpyHostToDevice));
      PE_HIP(hipMalloc((void **)&h->pcount, std::max<size_t>(n, 1)));
      PE_HIP(hipMemcpy(h->pcount, cnt.data(), n, hipMemcpyHostToDevice));
      if (!h->f64 && na <= kUnionCommands) {
        // Union rows of each P state and their weight per command.
        std::vector<int> uoff(1, 0), urows;
        std::vector<float> uw;
        std::vector<int> rows;
        for (int b = 0; b < n_maps; b++) {
          const pe_sl_car &c = games[b].p;
          for (int i = 0; i < c.n; i++) {
            rows.clear();
            for (int u = 0; u < na; u++) {
              const long long cmd = (long long)i * na + u;
              if (!c.allowed[cmd]) continue;
              for (int q = 0; q < kStencil; q++)
                if (c.weights[cmd * kStencil + q] != 0.f)
                  rows.push_back(c.successors[cmd * kStencil + q]);
            }
            std::sort(rows.begin(), rows.end());
            rows.erase(std::unique(rows.begin(), rows.end()), rows.end());
            const size_t real = rows.size();
            while (rows.size() % kUnionBatch) rows.push_back(rows.back());  // zero weights
            const size_t at = uw.size();
            uw.resize(at + rows.size() * kUnionCommands, 0.f);
            for (int u = 0; u < na; u++) {
              const long long cmd = (long long)i * na + u;
              if (!c.allowed[cmd]) continue;
              for (int q = 0; q < kStencil; q++) {
                const float w = c.weights[cmd * kStencil + q];
                if (w == 0.f) continue;
                const size_t t = std::lower_bound(rows.begin(), rows.begin() + real,
                                                  c.successors[cmd * kStencil + q]) -
                                 rows.begin();
                uw[at + t * kUnionCommands + u] += w;
              }
            }
            urows.insert(urows.end(), rows.begin(), rows.end());
            uoff.push_back((int)urows.size());
          }
        }
        PE_HIP(hipMalloc((void **)&h->uoff, sizeof(int) * uoff.size()));
        PE_HIP(hipMemcpy(h->uoff, uoff.data(), sizeof(int) * uoff.size(),
                         hipMemcpyHostToDevice));
        PE_HIP(hipMalloc((void **)&h->urows, sizeof(int) * std::max<size_t>(urows.size(), 1)));
        PE_HIP(hipMemcpy(h->urows, urows.data(), sizeof(int) * urows.size(),
                         hipMemcpyHostToDevice));
        PE_HIP(hipMalloc((void **)&h->uw, sizeof(float) * std::max<size_t>(uw.size(), 1)));
        PE_HIP(hipMemcpy(h->uw, uw.data(), sizeof(float) * uw.size(), hipMemcpyHostToDevice));
      }
    }
    if (h->lds_bytes && packed) {  // only the LDS E side reads them
      PE_HIP(hipMalloc(&h->erec, sizeof(ERec) * std::max<size_t>(erecs.size(), 1)));
      PE_HIP(hipMemcpy(h->erec, erecs.data(), sizeof(ERec) * erecs.size(),
                       hipMemcpyHostToDevice));
    }
    PE_HIP(hipMalloc((void **)&h->ilist, sizeof(int) * std::max<size_t>(ilist.size(), 1)));
    PE_HIP(hipMalloc((void **)&h->ioff, sizeof(int) * ioff.size()));
    PE_HIP(hipMemcpy(h->ilist, ilist.data(), sizeof(int) * ilist.size(), hipMemcpyHostToDevice));
    PE_HIP(hipMemcpy(h->ioff, ioff.data(), sizeof(int) * ioff.size(), hipMemcpyHostToDevice));
    PE_HIP(hipMalloc((void **)&h->nlist, sizeof(int) * std::max<size_t>(nlist.size(), 1)));
    PE_HIP(hipMalloc((void **)&h->noff, sizeof(int) * noff.size()));
    PE_HIP(hipMemcpy(h->nlist, nlist.data(), sizeof(int) * nlist.size(), hipMemcpyHostToDevice));
    PE_HIP(hipMemcpy(h->noff, noff.data(), sizeof(int) * noff.size(), hipMemcpyHostToDevice));
    PE_HIP(hipMalloc((void **)&h->inside, std::max(toff, 1ll)));
    PE_HIP(hipMalloc((void **)&h->near, std::max(toff, 1ll)));
    long long at = 0;
    for (int b = 0; b < n_maps; b++) {
      const long long sz = (long long)games[b].p.positions * games[b].e.positions;
      PE_HIP(hipMemcpy(h->inside + at, games[b].inside, sz, hipMemcpyHostToDevice));
      PE_HIP(hipMemcpy(h->near + at, games[b].near, sz, hipMemcpyHostToDevice));
      at += sz;
    }
    PE_HIP(hipMalloc((void **)&h->va, h->esize * h->pairs));
    PE_HIP(hipMalloc((void **)&h->vb, h->esize * h->pairs));
    PE_HIP(hipMalloc((void **)&h->order, h->pairs));
    PE_HIP(hipMalloc((void **)&h->G, h->esize * h->g_size));
    PE_HIP(hipMalloc((void **)&h->ring, 8ll * h->ring_slots * n_maps));
    PE_HIP(hipMalloc((void **)&h->gslots, 8ll * h->graph_sweeps[0] * n_maps));
    hipLaunchKernelGGL(sl_fill_f32, dim3(4096), dim3(256), 0, h->stream, h->G,
                       h->g_size * (long long)(h->esize / sizeof(float)));
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

// Payoff initialisation plus an optional given field (float or double
// values, converted to the handle's precision on the host).
template <typename S>
static int initialize_from(pe_sl *h, const S *v0) {
  clear_error();
  if (!h) return fail(PE_ERR_INVALID, "null handle");
  PE_HIP(hipSetDevice(h->device));
  if (h->f64) {
    sl_initialize_f64(h);
  } else {
    hipLaunchKernelGGL(sl_initialize<float>, pair_grid(h), dim3(256), 0, h->stream,
                       h->d_maps, h->pstates, h->estates, h->ppos, h->epos, h->inside, h->va,
                       h->order);
  }
  PE_LAUNCHED(sl_initialize);
  if (v0) {
    for (long long k = 0; k < h->pairs; k++)
      if (!std::isfinite(v0[k])) return fail(PE_ERR_INVALID, "non-finite initial field");
    PE_HIP(hipStreamSynchronize(h->stream));
    if (h->f64) {
      std::vector<double> v(v0, v0 + h->pairs);
      PE_HIP(hipMemcpy(h->va, v.data(), sizeof(double) * h->pairs, hipMemcpyHostToDevice));
    } else {
      std::vector<float> v(v0, v0 + h->pairs);
      PE_HIP(hipMemcpy(h->va, v.data(), sizeof(float) * h->pairs, hipMemcpyHostToDevice));
    }
  }
  // Both value buffers hold the same field: the LDS E side writes inside
  // pairs only, so the others must already agree in the buffer it writes.
  PE_HIP(hipMemcpyAsync(h->vb, h->va, h->esize * h->pairs, hipMemcpyDeviceToDevice, h->stream));
  PE_HIP(hipStreamSynchronize(h->stream));
  h->trace.clear();
  h->sweeps = 0;
  h->last = 0;
  return PE_OK;
}

extern "C" int pe_sl_initialize(pe_sl *h, const float *v0) { return initialize_from(h, v0); }

extern "C" int pe_sl_initialize_f64(pe_sl *h, const double *v0) {
  return initialize_from(h, v0);
}

// Freezes every still-swept map whose residual in the last sweep is at most
// tol: its current field is copied into the other value buffer (so the
// ping-pong keeps it) and the sweep kernels skip it from then on.
static int freeze_converged(pe_sl *h, double tol, bool *frozen) {
  const double *last = h->trace.data() + h->trace.size() - h->B;
  bool changed = false;
  for (int b = 0; b < h->B; b++) {
    SlMap &m = h->maps[b];
    if (!m.active || !(last[b] <= tol)) continue;
    m.active = 0;
    changed = true;
    const size_t off = h->esize * (size_t)m.v_off, bytes = h->esize * (size_t)m.np * m.ne;
    PE_HIP(hipMemcpyAsync(reinterpret_cast<char *>(h->vb) + off,
                          reinterpret_cast<const char *>(h->va) + off, bytes,
                          hipMemcpyDeviceToDevice, h->stream));
  }
  if (!changed) return PE_OK;
  // Stream-ordered after the copies above and before the next sweep; the
  // synchronisation keeps the host array stable until the copy has read it.
  PE_HIP(hipMemcpyAsync(h->d_maps, h->maps.data(), sizeof(SlMap) * h->B,
                        hipMemcpyHostToDevice, h->stream));
  PE_HIP(hipStreamSynchronize(h->stream));
  *frozen = true;
  return PE_OK;
}

static bool valid_flags(int32_t flags) {
  const int32_t known = PE_SL_EARLY_EXIT | PE_SL_ORDERED | PE_SL_GRAPH | PE_SL_GLOBAL_GATHER |
                        PE_SL_WIDE_IDS | PE_SL_ONE_PASS | PE_SL_U_MAJOR | PE_SL_FREEZE |
                        PE_SL_STENCIL_P_SIDE | PE_SL_SEPARATE_P_SIDE;
  return !(flags & ~known) && !((flags & PE_SL_ONE_PASS) && (flags & PE_SL_U_MAJOR));
}

extern "C" int pe_sl_sweep(pe_sl *h, int32_t sweeps, int32_t flags) {
  clear_error();
  if (!h || sweeps < 0) return fail(PE_ERR_INVALID, "invalid argument");
  if (!valid_flags(flags)) return fail(PE_ERR_INVALID, "unknown or conflicting sweep flags");
  PE_HIP(hipSetDevice(h->device));
  return do_sweeps(h, sweeps, flags);
}

extern "C" int pe_sl_run(pe_sl *h, double tol, int32_t max_sweeps,
                         int32_t check_every, int32_t flags,
                         int32_t *sweeps_done, int32_t *converged) {
  clear_error();
  if (!h || !(tol >= 0) || max_sweeps < 0 || check_every < 1)
    return fail(PE_ERR_INVALID, "invalid argument");
  if (!valid_flags(flags)) return fail(PE_ERR_INVALID, "unknown or conflicting sweep flags");
  PE_HIP(hipSetDevice(h->device));
  const bool freeze = flags & PE_SL_FREEZE;
  flags &= ~PE_SL_FREEZE;
  int done = 0, ok = 0, rc = PE_OK;
  bool frozen = false;
  while (done < max_sweeps) {
    const int chunk = std::min(check_every, max_sweeps - done);
    if ((rc = do_sweeps(h, chunk, flags))) break;
    done += chunk;
    if (h->last <= tol) {
      ok = 1;
      break;
    }
    if (freeze && (rc = freeze_converged(h, tol, &frozen))) break;
  }
  if (frozen) {  // every map is swept again after this call
    for (SlMap &m : h->maps) m.active = 1;
    hipError_t e = hipMemcpyAsync(h->d_maps, h->maps.data(), sizeof(SlMap) * h->B,
                                  hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (!rc && e != hipSuccess) rc = fail(PE_ERR_HIP, hipGetErrorString(e));
  }
  if (rc) return rc;
  if (sweeps_done) *sweeps_done = done;
  if (converged) *converged = ok;
  return PE_OK;
}

// Copies map `map` of the current field out, converting between the
// handle's precision and the caller's.
template <typename S>
static int read_into(pe_sl *h, int32_t map, S *field) {
  clear_error();
  if (!h || !field || map < 0 || map >= h->B)
    return fail(PE_ERR_INVALID, "invalid argument");
  PE_HIP(hipSetDevice(h->device));
  const SlMap &m = h->maps[map];
  const long long n = (long long)m.np * m.ne;
  PE_HIP(hipStreamSynchronize(h->stream));
  if (h->f64) {
    std::vector<double> v((size_t)n);
    PE_HIP(hipMemcpy(v.data(), reinterpret_cast<double *>(h->va) + m.v_off, sizeof(double) * n,
                     hipMemcpyDeviceToHost));
    for (long long k = 0; k < n; k++) field[k] = (S)v[k];
  } else if (std::is_same<S, float>::value) {
    PE_HIP(hipMemcpy(field, h->va + m.v_off, sizeof(float) * n, hipMemcpyDeviceToHost));
  } else {
    std::vector<float> v((size_t)n);
    PE_HIP(hipMemcpy(v.data(), h->va + m.v_off, sizeof(float) * n, hipMemcpyDeviceToHost));
    for (long long k = 0; k < n; k++) field[k] = (S)v[k];
  }
  return PE_OK;
}

extern "C" int pe_sl_read(pe_sl *h, int32_t map, float *field) {
  return read_into(h, map, field);
}

extern "C" int pe_sl_read_f64(pe_sl *h, int32_t map, double *field) {
  return read_into(h, map, field);
}

extern "C" int pe_sl_query(pe_sl *h, int32_t map, int64_t n,
                           const int32_t *p_ids, const float *p_weights,
                           const int32_t *e_ids, const float *e_weights,
                           float *out) {
  clear_error();
  if (!h || map < 0 || map >= h->B || n < 0 || (n && (!p_ids || !p_weights ||
                                                      !e_ids || !e_weights || !out)))
    return fail(PE_ERR_INVALID, "invalid argument");
  if (h->f64) return fail(PE_ERR_INVALID, "field queries need an FP32 handle");
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
  s->lds_e_side = h->lds_bytes > 0 ? h->lds_threads : 0;
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

extern "C" int pe_sl_set_lds_threads(pe_sl *h, int32_t threads) {
  clear_error();
  if (!h) return fail(PE_ERR_INVALID, "null handle");
  if (threads != 256 && threads != 512 && threads != 1024)
    return fail(PE_ERR_INVALID, "LDS E-side blocks have 256, 512 or 1024 threads");
  if (h->graph[0] || h->graph[1]) {  // captured launches carry the old block size
    PE_HIP(hipStreamSynchronize(h->stream));
    drop_graphs(h);
  }
  h->lds_threads = threads;
  return PE_OK;
}

extern "C" void pe_sl_destroy(pe_sl *h) {
  if (!h) return;
  hipSetDevice(h->device);
  if (h->stream) hipStreamSynchronize(h->stream);
  drop_graphs(h);
  for (hipEvent_t e : h->events) hipEventDestroy(e);
  for (void *p : {(void *)h->d_maps, (void *)h->pids, (void *)h->eids, (void *)h->erec,
                  (void *)h->ilist, (void *)h->ioff, (void *)h->nlist, (void *)h->noff,
                  (void *)h->pcount, (void *)h->uoff, (void *)h->urows, (void *)h->uw,
                  (void *)h->ppos, (void *)h->epos, (void *)h->pw, (void *)h->ew,
                  (void *)h->pallowed, (void *)h->eallowed, (void *)h->pmask,
                  (void *)h->emask, (void *)h->inside,
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
