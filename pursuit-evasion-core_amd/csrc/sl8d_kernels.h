// Kernels of the compact 8D semi-Lagrangian backup, templated on the value
// type.  pe_sl8d.hip instantiates them in FP32 (the product path) and
// pe_sl8d_f64.hip in FP64, compiled without contraction so that every
// stencil sum is the oracle's: zero weights skipped, then, left to right,
// a multiply and an add (oracle/sl8d.c follows local_game/grid_game.py).
#pragma once

#include <algorithm>
#include <cmath>
#include <type_traits>

#include "pe_internal.h"

namespace pe {

// acc + w * v: one FP32 fma, or the FP64 multiply-then-add of positive
// weights only.
__device__ __forceinline__ float madd(float w, float v, float acc) { return fmaf(w, v, acc); }
__device__ __forceinline__ double madd(float w, double v, double acc) {
  return w > 0.f ? acc + (double)w * v : acc;
}

template <typename T>
__global__ __launch_bounds__(256) void sl_initialize(
    const SlMap *__restrict__ maps, const double *__restrict__ pstates,
    const double *__restrict__ estates, const int *__restrict__ ppos,
    const int *__restrict__ epos, const uint8_t *__restrict__ inside,
    T *__restrict__ V, uint8_t *__restrict__ order) {
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
    V[pair] = T(fmax(fmax(g0, g1), fmax(g2, g3)));
  }
}

// First N terms of a compacted P stencil, in order, all loads issued
// together.
template <int N, typename T>
__device__ __forceinline__ T p_sum(const int *__restrict__ id, const float *__restrict__ w,
                                   const T *__restrict__ Vm, int ne, int e) {
  T v[N];
#pragma unroll
  for (int q = 0; q < N; q++) v[q] = Vm[(long long)id[q] * ne + e];
  T acc = 0;
#pragma unroll
  for (int q = 0; q < N; q++) acc = madd(w[q], v[q], acc);
  return acc;
}

// P side: G_u(ip, e) for the near E states e of each P state ip: blocks
// (ip, chunk) walk the list of near E states of ip's position (ascending,
// so the V rows are still read nearly coalesced) in gridDim.y strided
// chunks, loading the P stencils once per block; entries of G that no
// inside pair can reach are not computed.  Few chunks when there are many
// P states (each block loops longer on its stencils), more when there are
// few (enough blocks to fill the GPU).  The P stencils are
// compacted: the vertices of non-zero weight first, in vertex order, then
// zero-weight padding, with their count per (state, command).  About half
// the weights are zero (illegal vertices, or a coordinate exactly on a grid
// node), so 4, 8 or 16 terms are summed; adding 0 * V to a sum that starts
// at +0 leaves it unchanged, so the result is the full 16-term sum.
constexpr int kPThreads = 256;             // block size of sl_p_side
constexpr long long kPSideBlocks = 16384;  // P-side blocks to aim for

template <typename T>
__global__ __launch_bounds__(kPThreads) void sl_p_side(
    const SlMap *__restrict__ maps, const T *__restrict__ V,
    T *__restrict__ G, const int *__restrict__ pids,
    const float *__restrict__ pw, const uint8_t *__restrict__ pcount,
    const uint8_t *__restrict__ pallowed,
    const int *__restrict__ ppos, const int *__restrict__ nlist,
    const int *__restrict__ noff) {
  const SlMap m = maps[blockIdx.z];
  // Consecutive blocks go to the 8 XCDs in turn; block x takes P state
  // start(x % 8) + x / 8, so each XCD walks a contiguous range of P states
  // and the V rows of their (nearby) successors stay in its L2.  A
  // bijection of [0, gridDim.x).
  const int gx = (int)gridDim.x, q = gx >> 3, r = gx & 7;
  const int xcd = (int)blockIdx.x & 7;
  const int ip = xcd * q + min(xcd, r) + ((int)blockIdx.x >> 3);
  if (ip >= m.np || !m.active) return;  // blockIdx.y: strided chunk of the list
  const T *__restrict__ Vm = V + m.v_off;
  const int row = m.p_row + ip;
  const int k0 = noff[m.ipos_off + ppos[row]];
  const int count = noff[m.ipos_off + ppos[row] + 1] - k0;
  for (int u = 0; u < m.na; u++) {
    const long long cmd = (long long)row * m.na + u;
    if (!pallowed[cmd]) continue;
    const int *id = pids + cmd * kStencil;
    const float *w = pw + cmd * kStencil;
    const int n = pcount[cmd];
    T *__restrict__ Gu = G + m.g_off + ((long long)u * m.np + ip) * m.ne;
    for (int k = blockIdx.y * kPThreads + threadIdx.x; k < count; k += gridDim.y * kPThreads) {
      const int e = nlist[k0 + k];
      Gu[e] = n <= 4   ? p_sum<4>(id, w, Vm, m.ne, e)
              : n <= 8 ? p_sum<8>(id, w, Vm, m.ne, e)
                       : p_sum<kStencil>(id, w, Vm, m.ne, e);
    }
  }
}

// i-th command of the order that puts `first` in front (others ascending).
__device__ __forceinline__ int nth_command(int i, int first) {
  return i == 0 ? first : (i <= first ? i - 1 : i);
}

// E stencil ids of one (state, command): 16 values of type ID, in vertex
// order: int E-state ids, or, in the packed records, signed 16-bit byte
// offsets in the interleaved LDS tile relative to the E state's own values
// (zero for a vertex of zero weight).
template <typename ID>
struct EIds;
template <>
struct EIds<int> {
  int4 v[4];
  __device__ __forceinline__ void load(const int *p) {
    const int4 *q = reinterpret_cast<const int4 *>(p);
#pragma unroll
    for (int c = 0; c < 4; c++) v[c] = q[c];
  }
  __device__ __forceinline__ int operator[](int k) const {
    const int4 &a = v[k >> 2];
    const int r = k & 3;
    return r == 0 ? a.x : r == 1 ? a.y : r == 2 ? a.z : a.w;
  }
};
template <>
struct EIds<unsigned short> {
  uint4 v[2];
  // offset of vertex k, sign-extended (one SDWA add with the base)
  __device__ __forceinline__ int operator[](int k) const {
    const uint4 &a = v[k >> 3];
    const int r = (k >> 1) & 3;
    const unsigned w = r == 0 ? a.x : r == 1 ? a.y : r == 2 ? a.z : a.w;
    return (int)(short)((k & 1) ? (w >> 16) : w);
  }
};

// Interleaved 16-bit offsets and weights of one (state, command): 96 bytes.
struct ERec {
  uint4 ids[2];
  float4 w[4];
};

// One E stencil in registers: from the packed records (ID = ERec) or from
// the separate 32-bit id and float weight tables (ID = int).
template <typename ID>
struct EStencil {
  EIds<typename std::conditional<std::is_same<ID, ERec>::value, unsigned short, int>::type> id;
  float4 w[4];
  __device__ __forceinline__ float weight(int k) const {
    const float4 &a = w[k >> 2];
    const int r = k & 3;
    return r == 0 ? a.x : r == 1 ? a.y : r == 2 ? a.z : a.w;
  }
  __device__ __forceinline__ void load(const ID *__restrict__ eids,
                                       const float *__restrict__ ew, long long cmd) {
    if constexpr (std::is_same<ID, ERec>::value) {
      const ERec *rec = eids + cmd;
      id.v[0] = rec->ids[0];
      id.v[1] = rec->ids[1];
#pragma unroll
      for (int c = 0; c < 4; c++) w[c] = rec->w[c];
    } else {
      id.load(eids + cmd * kStencil);
      const float4 *w4 = reinterpret_cast<const float4 *>(ew + cmd * kStencil);
#pragma unroll
      for (int c = 0; c < 4; c++) w[c] = w4[c];
    }
  }
};

template <typename T>
__device__ __forceinline__ unsigned lds_address(const T *p) {
  return (unsigned)(uintptr_t)(const __attribute__((address_space(3))) T *)p;
}
template <typename T>
__device__ __forceinline__ T lds_load(unsigned address) {
  return *(const __attribute__((address_space(3))) T *)(uintptr_t)address;
}

// The G values of one P state staged in LDS, interleaved: the value of E
// state e under P command u at element e * na + u.  The values of one
// vertex under all P commands are adjacent, so one vertex address serves
// every command (the command becomes the immediate offset of the LDS read).
template <typename T>
struct LdsTile {
  // LDS byte address of element 0 (int ids), or of the values of the
  // pair's own E state (the packed records' offsets are relative to it)
  unsigned base;
  int na;
  // byte address of vertex k of s under command 0, plus `add`
  template <typename ID>
  __device__ __forceinline__ unsigned vertex(const EStencil<ID> &s, int k, unsigned add) const {
    if constexpr (std::is_same<ID, ERec>::value)
      return add + (unsigned)s.id[k];
    else
      return add + __umul24((unsigned)s.id[k], (unsigned)na * (unsigned)sizeof(T));
  }
  // sum_k w_k G_u[id_k], vertex order (grid_game.py:406-408)
  template <typename ID>
  __device__ __forceinline__ T eval(const EStencil<ID> &s, int u) const {
    const unsigned b = base + (unsigned)u * (unsigned)sizeof(T);
    T q = 0;
#pragma unroll
    for (int k = 0; k < kStencil; k++) q = madd(s.weight(k), lds_load<T>(vertex(s, k, b)), q);
    return q;
  }
  // Two stencils under the same command at once (two independent chains).
  template <typename ID>
  __device__ __forceinline__ void eval2(const EStencil<ID> &sa, const EStencil<ID> &sb, int u,
                                        T &qa, T &qb) const {
    const unsigned b = base + (unsigned)u * (unsigned)sizeof(T);
    qa = 0;
    qb = 0;
#pragma unroll
    for (int k = 0; k < kStencil; k++) {
      qa = madd(sa.weight(k), lds_load<T>(vertex(sa, k, b)), qa);
      qb = madd(sb.weight(k), lds_load<T>(vertex(sb, k, b)), qb);
    }
  }
  // The vertex addresses of one stencil, for evaluating it under several
  // commands.
  struct Verts {
    unsigned a[kStencil];
  };
  template <typename ID>
  __device__ __forceinline__ Verts verts(const EStencil<ID> &s) const {
    Verts v;
#pragma unroll
    for (int k = 0; k < kStencil; k++) v.a[k] = vertex(s, k, base);
    return v;
  }
  // eval(s, u) from the vertex addresses; u a compile-time constant here.
  template <typename ID>
  __device__ __forceinline__ T eval_at(const Verts &v, const EStencil<ID> &s, int u) const {
    T q = 0;
#pragma unroll
    for (int k = 0; k < kStencil; k++)
      q = madd(s.weight(k), lds_load<T>(v.a[k] + (unsigned)u * (unsigned)sizeof(T)), q);
    return q;
  }
};

// The G rows of one P state in global memory: row u at p + u * ustride
// (G is [na][np][ne] per map), gathered with the 32-bit ids.
template <typename T>
struct GlobalRows {
  const T *__restrict__ p;
  long long ustride;
  __device__ __forceinline__ T eval(const EStencil<int> &s, int u) const {
    const T *__restrict__ Gu = p + u * ustride;
    T q = 0;
#pragma unroll
    for (int k = 0; k < kStencil; k++) q = madd(s.weight(k), Gu[s.id[k]], q);
    return q;
  }
  __device__ __forceinline__ void eval2(const EStencil<int> &sa, const EStencil<int> &sb, int u,
                                        T &qa, T &qb) const {
    const T *__restrict__ Gu = p + u * ustride;
    qa = 0;
    qb = 0;
#pragma unroll
    for (int k = 0; k < kStencil; k++) {
      qa = madd(sa.weight(k), Gu[sa.id[k]], qa);
      qb = madd(sb.weight(k), Gu[sb.id[k]], qb);
    }
  }
  struct Verts {};
  __device__ __forceinline__ Verts verts(const EStencil<int> &) const { return {}; }
  __device__ __forceinline__ T eval_at(const Verts &, const EStencil<int> &s, int u) const {
    return eval(s, u);
  }
};

// Min over allowed P commands of the max over allowed E commands of one
// inside pair (ip, ie), starting from its current value; `tile` holds the
// values G_u(ip, .) of the P state.  Early exit: a P command whose running
// max over E commands reaches the running min cannot lower it
// (grid_game.py:409-421).  Updates the stored command order of the pair.
//
// ORDER selects the visiting order:
//   kPlainOrder   commands in index order, u-major, as in the reference;
//   kUMajor       the stored best commands (u*, b*) first, then the others
//                 in index order, u-major;
//   kOnePass      one pass over the E commands (b* first) carrying every
//                 allowed P command along, so each E stencil is read once
//                 for all of them instead of once per P command (more
//                 registers); u* is judged first at the end;
//   kOnePass9     the same with room for 9 commands per state instead of
//                 15 (the {-A, 0, A} x {-K, 0, K} set): the running maxima
//                 take 6 registers fewer.
// The value and the stored commands are the same in all three: max and min
// are exact, each (u, b) combination is summed in the same order, ties go
// to the first command in visiting order, and a command cut against a
// bound could not have been chosen later (its max already reaches a value
// at or above the final min).
constexpr int kPlainOrder = 0, kUMajor = 1, kOnePass = 2, kOnePass9 = 3;

template <bool EARLY, int ORDER, typename T, typename Tile, typename ID>
__device__ __forceinline__ T pair_backup(const SlMap &m, int prow, int erow, T start,
                                         long long pair, const Tile &tile,
                                         const ID *__restrict__ eids,
                                         const float *__restrict__ ew,
                                         const uint16_t *__restrict__ emask,
                                         const uint16_t *__restrict__ pmask,
                                         uint8_t *__restrict__ order) {
  constexpr bool ORDERED = ORDER != kPlainOrder;
  constexpr bool ONE_PASS = ORDER == kOnePass || ORDER == kOnePass9;
  constexpr int MAXA = ORDER == kOnePass9 ? 9 : kMaxActions;  // commands per state
  T best = start;
  int ufirst = 0, bfirst = 0;
  if (ORDERED) {
    const int o = order[pair];
    if (o != 0xFF) {
      ufirst = o & 15;
      bfirst = o >> 4;
    }
  }
  const long long cmd0 = (long long)erow * m.na;
  // allowed commands, bit u: of the P state and of the E state
  const unsigned pm = pmask[prow], em = emask[erow];
  int best_u = -1, best_b = -1;
  // the E state's allowed commands in visiting order (bit j = j-th visited:
  // bfirst, then the others ascending)
  const unsigned ea_visit = ((em >> bfirst) & 1u) | ((em & ((1u << bfirst) - 1u)) << 1) |
                            (em & ~((2u << bfirst) - 1u));
  for (int i = 0; i < (ONE_PASS ? 0 : m.na); i++) {
    const int u = ORDERED ? nth_command(i, ufirst) : i;
    if (!((pm >> u) & 1u)) continue;
    T worst = T(-1e30);
    int arg = -1;
    bool cut = false;
    for (int j = 0; j < m.na; j++) {
      const int b = ORDERED ? nth_command(j, bfirst) : j;
      if (!((em >> b) & 1u)) continue;
      EStencil<ID> st;
      st.load(eids, ew, cmd0 + b);
      const T q = tile.eval(st, u);
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
  if constexpr (ONE_PASS) {
    // One pass over the E commands in visiting order carries every allowed
    // P command, u* included: each E stencil record is read once per pair.
    // A command is cut once its running max reaches the pair's start value
    // (it cannot lower the value).  u*, if it survives, then lowers the
    // bound, and the other survivors are judged against that: the same
    // outcome as visiting u* first, since the start value is at or above
    // every later bound and the running maxima only grow.
    T worst[MAXA];                // running max of each P command
    unsigned long long args = 0;  // its argmax, 4 bits per P command
#pragma unroll
    for (int u = 0; u < MAXA; u++) worst[u] = T(-1e30);
    unsigned alive = pm;
    for (unsigned todo = ea_visit; todo && alive; todo &= todo - 1) {
      const int b = nth_command(__builtin_ctz(todo), bfirst);
      EStencil<ID> st;
      st.load(eids, ew, cmd0 + b);
      const auto vt = tile.verts(st);
#pragma unroll
      for (int u = 0; u < MAXA; u++) {
        if (!((alive >> u) & 1u)) continue;
        const T q = tile.eval_at(vt, st, u);
        if (q > worst[u]) {
          worst[u] = q;
          args = (args & ~(15ull << (4 * u))) | ((unsigned long long)b << (4 * u));
        }
        if (EARLY && worst[u] >= best) alive &= ~(1u << u);
      }
    }
    // Survivors in visiting order: u*, then ascending u.
    if (((alive >> ufirst) & 1u) && worst[ufirst] < best) {
      best = worst[ufirst];
      best_u = ufirst;
      best_b = ea_visit ? (int)((args >> (4 * ufirst)) & 15) : -1;
    }
    alive &= ~(1u << ufirst);
#pragma unroll
    for (int u = 0; u < MAXA; u++) {
      if (((alive >> u) & 1u) && worst[u] < best) {
        best = worst[u];
        best_u = u;
        best_b = (int)((args >> (4 * u)) & 15);
      }
    }
  }
  if (ORDERED && best_u >= 0) order[pair] = (uint8_t)(best_u | (best_b << 4));
  return best;
}

// P side in FP32 from the union of a P state's stencil rows: the V rows that
// any allowed command of the P state reads with non-zero weight (22 per P
// state at C2 against 75 padded stencil terms over its 9 commands).  Each
// thread loads a row's value at its E state once and adds it, with that
// row's weight for every command, to all the commands' sums: G_u = sum_r
// W[r][u] V[row_r], rows ascending (then zero-weight padding).  The order
// of the terms differs from the stencils' vertex order and zero weights are
// added, so the sums agree with sl_p_side to rounding (the FP64 mode keeps
// sl_p_side, the oracle's order).  Blocks as in sl_p_side; rows of disallowed commands are not
// written.  MAXA: room for the commands of a state.
constexpr int kUnionCommands = 9;  // commands per state of sl_p_side_union
constexpr int kUnionBatch = 8;     // its union rows per batch of loads

// The P side's sums of one P state at E states e0 and e1 from its union rows
// r0 .. r0 + nr (padded to kUnionBatch): acc[u] = sum_r W[r][u] V[row_r],
// each row's value loaded once for both E states and all commands.
template <int MAXA>
__device__ __forceinline__ void union_sums(const float *__restrict__ Vm, int ne,
                                           const int *__restrict__ urows,
                                           const float *__restrict__ uw, int r0, int nr,
                                           int e0, int e1, float (&acc0)[MAXA],
                                           float (&acc1)[MAXA]) {
#pragma unroll
  for (int u = 0; u < MAXA; u++) acc0[u] = acc1[u] = 0.f;
  for (int t0 = 0; t0 < nr; t0 += kUnionBatch) {
    float v0[kUnionBatch], v1[kUnionBatch];
#pragma unroll
    for (int i = 0; i < kUnionBatch; i++) {
      const float *__restrict__ vrow = Vm + (long long)urows[r0 + t0 + i] * ne;
      v0[i] = vrow[e0];
      v1[i] = vrow[e1];
    }
#pragma unroll
    for (int i = 0; i < kUnionBatch; i++) {
      const float *__restrict__ w = uw + (long long)(r0 + t0 + i) * MAXA;
#pragma unroll
      for (int u = 0; u < MAXA; u++) {
        acc0[u] = fmaf(w[u], v0[i], acc0[u]);
        acc1[u] = fmaf(w[u], v1[i], acc1[u]);
      }
    }
  }
}

template <int MAXA>
__global__ __launch_bounds__(kPThreads) void sl_p_side_union(
    const SlMap *__restrict__ maps, const float *__restrict__ V, float *__restrict__ G,
    const int *__restrict__ uoff, const int *__restrict__ urows,
    const float *__restrict__ uw, const uint16_t *__restrict__ pmask,
    const int *__restrict__ ppos, const int *__restrict__ nlist,
    const int *__restrict__ noff) {
  const SlMap m = maps[blockIdx.z];
  const int gx = (int)gridDim.x, q = gx >> 3, r = gx & 7;
  const int xcd = (int)blockIdx.x & 7;
  const int ip = xcd * q + min(xcd, r) + ((int)blockIdx.x >> 3);
  if (ip >= m.np || !m.active) return;  // blockIdx.y: strided chunk of the list
  const float *__restrict__ Vm = V + m.v_off;
  const int row = m.p_row + ip;
  const int k0 = noff[m.ipos_off + ppos[row]];
  const int count = noff[m.ipos_off + ppos[row] + 1] - k0;
  const int r0 = uoff[row], nr = uoff[row + 1] - r0;
  const unsigned pm = pmask[row];
  const long long ustride = (long long)m.np * m.ne;
  float *__restrict__ Gp = G + m.g_off + (long long)ip * m.ne;
  // Each thread takes two E states of the list (k and k + stride) with the
  // same rows and weights: twice the loads in flight per weight fetch.
  const int stride = gridDim.y * kPThreads;
  for (int k = blockIdx.y * kPThreads + threadIdx.x; k < count; k += 2 * stride) {
    const bool two = k + stride < count;
    const int e0 = nlist[k0 + k], e1 = two ? nlist[k0 + k + stride] : e0;
    float acc0[MAXA], acc1[MAXA];
    union_sums<MAXA>(Vm, m.ne, urows, uw, r0, nr, e0, e1, acc0, acc1);
#pragma unroll
    for (int u = 0; u < MAXA; u++)
      if ((pm >> u) & 1u) {
        Gp[u * ustride + e0] = acc0[u];
        if (two) Gp[u * ustride + e1] = acc1[u];
      }
  }
}

// Where the fused E side finds its P side: the union rows and the near
// lists (FP32 handles with packed records).
struct FusedP {
  const int *__restrict__ uoff;
  const int *__restrict__ urows;
  const float *__restrict__ uw;
  const int *__restrict__ nlist;
  const int *__restrict__ noff;
};

// Block max of `change` into the map's residual slot (blockDim.x <= 1024).
template <typename T>
__device__ __forceinline__ void block_residual(T change, unsigned long long *slot) {
  __shared__ T red[16];
  const T w = wave_max(change);
  const int waves = (int)(blockDim.x + 63) >> 6;
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = w;
  __syncthreads();
  if (threadIdx.x == 0) {
    T b = red[0];
    for (int k = 1; k < waves; k++) b = b > red[k] ? b : red[k];
    raise_max(slot, (double)b);
  }
}

// E side with G gathered from global memory: one thread per pair, with the
// 32-bit id tables.
template <typename T, bool EARLY, int ORDER>
__global__ __launch_bounds__(256) void sl_e_side(
    const SlMap *__restrict__ maps, const T *__restrict__ V,
    T *__restrict__ Vn, const T *__restrict__ G,
    const int *__restrict__ eids, const float *__restrict__ ew,
    const uint16_t *__restrict__ emask, const uint16_t *__restrict__ pmask,
    const int *__restrict__ ppos, const int *__restrict__ epos,
    const uint8_t *__restrict__ inside, uint8_t *__restrict__ order,
    unsigned long long *__restrict__ slot, int probe) {
  const SlMap m = maps[blockIdx.z];
  if (!m.active) return;  // the whole block: no residual for a frozen map
  const int ie = blockIdx.x * 256 + threadIdx.x;
  T change = 0;
  if (ie < m.ne) {
    const int erow = m.e_row + ie;
    const int ep = epos[erow];
    for (int ip = blockIdx.y; ip < m.np; ip += gridDim.y) {
      const long long pair = m.v_off + (long long)ip * m.ne + ie;
      const T start = probe ? T(3.0e38f) : V[pair];
      T best = start;
      const int prow = m.p_row + ip;
      if (inside[m.tab_off + (long long)ppos[prow] * m.npos_e + ep]) {
        const GlobalRows<T> rows{G + m.g_off + (long long)ip * m.ne, (long long)m.np * m.ne};
        best = pair_backup<EARLY, ORDER>(m, prow, erow, start, pair, rows, eids, ew,
                                           emask, pmask, order);
      }
      Vn[pair] = best;
      const T d = best > start ? best - start : start - best;
      change = change > d ? change : d;
    }
  }
  block_residual(change, slot + blockIdx.z);
}

// E side with the G rows of one P state staged in LDS: one block per
// (P state, map) walks the inside pairs of that P state, so the 16-point E
// gathers of every pair read LDS instead of scattered cache lines.  Dynamic
// LDS holds the P state's G values as [ne][na] (see LdsTile); only the
// values of allowed P commands are read.  The pairs come from the list of inside E states
// of the P state's position, so every lane has a pair to update; the pairs
// outside the exit radius are not written (both value buffers already hold
// the same exit value there).  Same arithmetic as sl_e_side, value for value.
// sl_e_side_lds runs with 256, 512 or 1024 threads per block (chosen per
// handle from the occupancy and the pairs per P state).
constexpr int kLdsMaxThreads = 1024;

//
// FUSED (FP32, packed records): no G buffer and no P-side launch; the block
// computes its P state's values straight into the tile from V with the
// union rows (the sums of sl_p_side_union, bit for bit), for the near E
// states only: the non-zero-weight vertices of every inside pair's stencils
// are near, and the records point every zero-weight vertex at the pair's
// own (inside, so near) state.
template <typename T, bool EARLY, int ORDER, typename ID, bool FUSED = false>
__global__ __launch_bounds__(kLdsMaxThreads) void sl_e_side_lds(
    const SlMap *__restrict__ maps, const T *__restrict__ V,
    T *__restrict__ Vn, const T *__restrict__ G,
    const ID *__restrict__ eids, const float *__restrict__ ew,
    const uint16_t *__restrict__ emask, const uint16_t *__restrict__ pmask,
    const int *__restrict__ ppos, const int *__restrict__ ilist,
    const int *__restrict__ ioff, uint8_t *__restrict__ order,
    unsigned long long *__restrict__ slot, int probe, FusedP fp) {
  static_assert(!FUSED || (std::is_same<T, float>::value && std::is_same<ID, ERec>::value),
                "the fused P side is FP32 with packed records");
  extern __shared__ __align__(16) unsigned char tile_bytes[];
  T *tile = reinterpret_cast<T *>(tile_bytes);
  const SlMap m = maps[blockIdx.z];
  if (!m.active) return;  // the whole block: no residual for a frozen map
  const int ip = blockIdx.x;  // uniform per block
  const bool active = ip < m.np;
  const int prow = m.p_row + ip;
  if constexpr (FUSED) {
    if (active) {
      const int n0 = fp.noff[m.ipos_off + ppos[prow]];
      const int count = fp.noff[m.ipos_off + ppos[prow] + 1] - n0;
      const int r0 = fp.uoff[prow], nr = fp.uoff[prow + 1] - r0;
      const int nt = (int)blockDim.x;
      const float *__restrict__ Vm = V + m.v_off;
      for (int k = threadIdx.x; k < count; k += 2 * nt) {
        const bool two = k + nt < count;
        const int e0 = fp.nlist[n0 + k], e1 = two ? fp.nlist[n0 + k + nt] : e0;
        float acc0[kUnionCommands], acc1[kUnionCommands];
        union_sums<kUnionCommands>(Vm, m.ne, fp.urows, fp.uw, r0, nr, e0, e1, acc0, acc1);
#pragma unroll
        for (int u = 0; u < kUnionCommands; u++)
          if (u < m.na) {
            tile[e0 * m.na + u] = acc0[u];
            if (two) tile[e1 * m.na + u] = acc1[u];
          }
      }
    }
  } else if (active) {
    // All na rows (rows of disallowed commands are read but never used),
    // read row by row (t = u * ne + e, coalesced) with eight loads in
    // flight per thread and written to tile[e * na + u].
    const int total = m.na * m.ne;
    const T *__restrict__ src = G + m.g_off + (long long)ip * m.ne;
    const long long ustride = (long long)m.np * m.ne;
    constexpr int kBatch = 8;
    const int nt = (int)blockDim.x;
    // Element t0 = u * ne + e is read from src[u * ustride + e] (index g)
    // and written to tile[e * na + u] (index at); both advance by adds.
    const int u0 = threadIdx.x / m.ne;
    int e = threadIdx.x - u0 * m.ne;
    long long g = u0 * ustride + e;
    int at = e * m.na + u0;
    const int at_step = nt * m.na, at_wrap = 1 - m.ne * m.na;
    const long long g_wrap = ustride - m.ne;
    for (int t0 = threadIdx.x; t0 < total; t0 += kBatch * nt) {
      T v[kBatch];
      int to[kBatch];
#pragma unroll
      for (int b = 0; b < kBatch; b++) {
        if (t0 + b * nt < total) v[b] = src[g];
        to[b] = at;
        e += nt;
        g += nt;
        at += at_step;
        while (e >= m.ne) {
          e -= m.ne;
          g += g_wrap;
          at += at_wrap;
        }
      }
#pragma unroll
      for (int b = 0; b < kBatch; b++)
        if (t0 + b * nt < total) tile[to[b]] = v[b];
    }
  }
  __syncthreads();
  T change = 0;
  if (active) {
    const int k0 = ioff[m.ipos_off + ppos[prow]];
    const int count = ioff[m.ipos_off + ppos[prow] + 1] - k0;
    const unsigned tile_at = lds_address(tile);
    for (int k = threadIdx.x; k < count; k += blockDim.x) {
      const int ie = ilist[k0 + k];
      const LdsTile<T> lt{
          std::is_same<ID, ERec>::value
              ? tile_at + __umul24((unsigned)ie, (unsigned)m.na * (unsigned)sizeof(T))
              : tile_at,
          m.na};
      const int erow = m.e_row + ie;
      const long long pair = m.v_off + (long long)ip * m.ne + ie;
      const T start = probe ? T(3.0e38f) : V[pair];
      const T best = pair_backup<EARLY, ORDER>(m, prow, erow, start, pair, lt, eids, ew,
                                                 emask, pmask, order);
      Vn[pair] = best;
      const T d = best > start ? best - start : start - best;
      change = change > d ? change : d;
    }
  }
  block_residual(change, slot + blockIdx.z);
}


inline dim3 pair_grid(const pe_sl *h) {
  const unsigned gy = (unsigned)std::min(h->np_max, 65535);
  return dim3((unsigned)((h->ne_max + 255) / 256), gy, (unsigned)h->B);
}

inline void sl_record(pe_sl *h, int kind, bool start) {
  if (!h->profiling) return;
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return;
  hipEventRecord(e, h->stream);
  h->events.push_back(e);
  if (start) h->kinds.push_back(kind);
}

// Enqueues the separate P side of one sweep (G from src): the union-row
// kernel for FP32 handles that have union rows, the stencil kernel
// otherwise or with PE_SL_STENCIL_P_SIDE.
template <typename T>
void enqueue_p_side(pe_sl *h, const T *src, T *G, int flags, bool record) {
  if (record) sl_record(h, 0, true);
  const long long p_blocks = (long long)h->np_max * h->B;
  const int chunks = (int)std::max(
      1ll, std::min({(long long)(h->ne_max + kPThreads - 1) / kPThreads,
                     (kPSideBlocks + p_blocks - 1) / p_blocks, 65535ll}));
  const dim3 p_grid((unsigned)h->np_max, (unsigned)chunks, (unsigned)h->B);
  bool union_rows = false;
  if constexpr (std::is_same<T, float>::value) {
    if (h->uoff && !(flags & PE_SL_STENCIL_P_SIDE)) {
      // half the chunks: each thread takes two E states
      const dim3 u_grid(p_grid.x, (unsigned)std::max(1, (chunks + 1) / 2), p_grid.z);
      hipLaunchKernelGGL(sl_p_side_union<kUnionCommands>, u_grid, dim3(kPThreads), 0,
                         h->stream, h->d_maps, src, G, h->uoff, h->urows, h->uw, h->pmask,
                         h->ppos, h->nlist, h->noff);
      union_rows = true;
    }
  }
  if (!union_rows)
    hipLaunchKernelGGL(sl_p_side<T>, p_grid, dim3(kPThreads), 0, h->stream, h->d_maps, src, G,
                       h->pids, h->pw, h->pcount, h->pallowed, h->ppos, h->nlist, h->noff);
  if (record) sl_record(h, 0, false);
}

// Enqueues one sweep src -> dst (T values) writing its residuals to
// slot[0..B).
template <typename T>
void enqueue_sweep_t(pe_sl *h, const void *src_, void *dst_, int flags,
                     unsigned long long *slot, bool record) {
  const T *src = static_cast<const T *>(src_);
  T *dst = static_cast<T *>(dst_);
  T *G = reinterpret_cast<T *>(h->G);
  const dim3 grid = pair_grid(h), block(256);
  const bool lds = h->lds_bytes > 0 && !(flags & PE_SL_GLOBAL_GATHER);
  const bool packed = h->erec && !(flags & PE_SL_WIDE_IDS);
  // The E side computes its own P side from the union rows (no G buffer)
  // unless a separate P side is asked for.
  const bool fused = std::is_same<T, float>::value && lds && packed && h->uoff &&
                     !(flags & (PE_SL_STENCIL_P_SIDE | PE_SL_SEPARATE_P_SIDE));
  const FusedP fp{h->uoff, h->urows, h->uw, h->nlist, h->noff};
  if (!fused) enqueue_p_side<T>(h, src, G, flags, record);
  if (record) sl_record(h, 1, true);
  const bool early = flags & PE_SL_EARLY_EXIT, ordered = flags & PE_SL_ORDERED;
  const int probe = h->probe_orders ? 1 : 0;
  const dim3 lds_grid((unsigned)h->np_max, 1, (unsigned)h->B), lds_block(h->lds_threads);
  // One pass over the E commands unless u-major is asked for.
  const bool one_pass = !(flags & PE_SL_U_MAJOR);
#define PE_E_SIDE_LDS(E, O, ID, IDS, F)                                                \
  hipLaunchKernelGGL((sl_e_side_lds<T, E, O, ID, F>), lds_grid, lds_block, h->lds_bytes, \
                     h->stream, h->d_maps, src, dst, G, IDS, h->ew, h->emask, h->pmask,  \
                     h->ppos, h->ilist, h->ioff, h->order, slot, probe, fp)
#define PE_E_SIDE(E, O)                                                                \
  do {                                                                                 \
    if (fused) {                                                                       \
      if constexpr (std::is_same<T, float>::value)                                     \
        PE_E_SIDE_LDS(E, O, ERec, (const ERec *)h->erec, true);                        \
    } else if (lds && packed)                                                          \
      PE_E_SIDE_LDS(E, O, ERec, (const ERec *)h->erec, false);                         \
    else if (lds)                                                                      \
      PE_E_SIDE_LDS(E, O, int, h->eids, false);                                        \
    else                                                                               \
      hipLaunchKernelGGL((sl_e_side<T, E, O>), grid, block, 0, h->stream, h->d_maps,   \
                         src, dst, G, h->eids, h->ew, h->emask, h->pmask, h->ppos,     \
                         h->epos, h->inside, h->order, slot, probe);                   \
  } while (0)
  const bool nine = h->na <= 9;
  if (early && ordered && one_pass && nine) PE_E_SIDE(true, kOnePass9);
  else if (early && ordered && one_pass) PE_E_SIDE(true, kOnePass);
  else if (early && ordered) PE_E_SIDE(true, kUMajor);
  else if (early) PE_E_SIDE(true, kPlainOrder);
  else if (ordered && one_pass && nine) PE_E_SIDE(false, kOnePass9);
  else if (ordered && one_pass) PE_E_SIDE(false, kOnePass);
  else if (ordered) PE_E_SIDE(false, kUMajor);
  else PE_E_SIDE(false, kPlainOrder);
#undef PE_E_SIDE
#undef PE_E_SIDE_LDS
  if (record) sl_record(h, 1, false);
}

// Lets the LDS E side use `bytes` of dynamic LDS for every flag variant;
// false if the runtime refuses.
template <typename T>
bool sl_allow_lds(size_t bytes) {
  bool ok = true;
  for (const void *f : {(const void *)sl_e_side_lds<T, true, kOnePass, int>,
                        (const void *)sl_e_side_lds<T, true, kOnePass9, int>,
                        (const void *)sl_e_side_lds<T, false, kOnePass9, int>,
                        (const void *)sl_e_side_lds<T, true, kOnePass9, ERec>,
                        (const void *)sl_e_side_lds<T, false, kOnePass9, ERec>,
                        (const void *)sl_e_side_lds<T, true, kUMajor, int>,
                        (const void *)sl_e_side_lds<T, true, kPlainOrder, int>,
                        (const void *)sl_e_side_lds<T, false, kOnePass, int>,
                        (const void *)sl_e_side_lds<T, false, kUMajor, int>,
                        (const void *)sl_e_side_lds<T, false, kPlainOrder, int>,
                        (const void *)sl_e_side_lds<T, true, kOnePass, ERec>,
                        (const void *)sl_e_side_lds<T, true, kUMajor, ERec>,
                        (const void *)sl_e_side_lds<T, true, kPlainOrder, ERec>,
                        (const void *)sl_e_side_lds<T, false, kOnePass, ERec>,
                        (const void *)sl_e_side_lds<T, false, kUMajor, ERec>,
                        (const void *)sl_e_side_lds<T, false, kPlainOrder, ERec>})
    ok = ok && hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) ==
                   hipSuccess;
  if constexpr (std::is_same<T, float>::value) {
    for (const void *f : {(const void *)sl_e_side_lds<T, true, kOnePass9, ERec, true>,
                          (const void *)sl_e_side_lds<T, false, kOnePass9, ERec, true>,
                          (const void *)sl_e_side_lds<T, true, kOnePass, ERec, true>,
                          (const void *)sl_e_side_lds<T, false, kOnePass, ERec, true>,
                          (const void *)sl_e_side_lds<T, true, kUMajor, ERec, true>,
                          (const void *)sl_e_side_lds<T, false, kUMajor, ERec, true>,
                          (const void *)sl_e_side_lds<T, true, kPlainOrder, ERec, true>,
                          (const void *)sl_e_side_lds<T, false, kPlainOrder, ERec, true>})
      ok = ok && hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                     (int)bytes) == hipSuccess;
  }
  if (!ok) (void)hipGetLastError();
  return ok;
}

// Threads per block of the LDS E side for FP32 values in the default
// (one-pass) order, of 256, 512 and 1024.  Score: the blocks one CU holds
// at this kernel's registers and `bytes` of LDS, times the threads per
// block, times the fraction of lane slots that the inside pairs of the P
// states fill (fill[k] for blocks of 256 << k threads): the busy lanes per
// CU.  The smallest size within 80 % of the best score is taken: more,
// smaller blocks measured faster than a few percent more busy lanes in
// larger ones (C4 maps: 3.85 ms with 256 threads against 4.17 ms with 512,
// whose score is 20 % higher).
inline int sl_lds_threads(size_t bytes, const double fill[3], bool packed, bool nine) {
  const void *f = nine ? (packed ? (const void *)sl_e_side_lds<float, true, kOnePass9, ERec>
                                 : (const void *)sl_e_side_lds<float, true, kOnePass9, int>)
                       : (packed ? (const void *)sl_e_side_lds<float, true, kOnePass, ERec>
                                 : (const void *)sl_e_side_lds<float, true, kOnePass, int>);
  double score[3] = {-1, -1, -1}, best = -1;
  for (int k = 0; k < 3; k++) {
    const int t = 256 << k;
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, f, t, bytes) != hipSuccess) {
      (void)hipGetLastError();
      continue;
    }
    score[k] = blocks * t * fill[k];
    best = std::max(best, score[k]);
  }
  for (int k = 0; k < 3; k++)
    if (best > 0 && score[k] >= 0.8 * best) return 256 << k;
  return kLdsMaxThreads;
}

// FP64 instances (pe_sl8d_f64.hip, built without contraction).
void sl_enqueue_f64(pe_sl *h, const void *src, void *dst, int flags, unsigned long long *slot,
                    bool record);
void sl_initialize_f64(pe_sl *h);
bool sl_allow_lds_f64(size_t bytes);

}  // namespace pe
