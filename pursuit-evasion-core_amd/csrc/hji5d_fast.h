// FP32 RK stage of the 5D relative HJI update, written for instruction
// economy on gfx950: the cost of a node is the vector instructions it
// issues, its neighbour loads first, then its VALU (packed FP32 is not
// used).
//
// Measured on C1 (34.2M nodes), blend / stage-1 launch:
//   generic template hji5d_stage<float, false, *>            232 / 195 us
//   vP-walking threads + guard margins + selects              178 / 151 us
//     (about 200 VALU per node; wave-uniform branches around the boundary
//     cases instead of selects were slower, 196 / 176 us)
//   raw differences, folded scale factors, mask-free interior
//     vE blocks and vP iterations (~130 VALU per node)        176 / 146 us
//   + next vP node's neighbours requested one node early      169 / 141 us
//   two adjacent psi planes per thread (shared psi loads, 16
//     instead of 18 loads per blend node), loads not early    165 / 142 us
// So the stage is not VALU bound.  Nor is it HBM bound: the (psi, tile, vE)
// block order fetches 2.2x the bytes from memory in the same time.  What
// scales is the number of vector memory instructions per node (about 19):
// variants that omit neighbour loads run faster in proportion.  Taking the x
// neighbours from adjacent lanes with DPP wave shifts instead of loads was
// bit-identical but slower (173 / 148 us): the shifts, their wait states
// and the halo lanes cost more than the four loads.  Staging each vP node's
// xy tile with its halo rows in LDS (eight x/y neighbours from LDS, about two
// global loads per node for the tile) was bit-identical and much slower
// (228 / 202 us): the block barrier per vP node stops the waves from running
// ahead of one another, which is what hides the load latency here.  Staging
// a block's whole vP range of a few y rows once, behind a single barrier,
// was also bit-identical and slower (233 / 201 us with 512 threads, 234 / 207
// with 256).  A load-only kernel with these 17 loads per node takes 129 us on
// C1, and dwordx2 loads of the same bytes (half the instructions) 158 us:
// the load path is bound by bytes, and the ~130 VALU per node add ~113 us
// of issue time, so the stage sits near both floors.
//
// Same scheme as hji5d_stage<float, false, *> node by node
// (reversible/native_relative.cpp:38-67 slopes, :124-172 update), with the
// arithmetic reorganised (FP32 results agree to rounding, not bitwise):
//  * ENO differences are formed in value units and the 1/h factors are
//    folded into per-thread coefficients of the Hamiltonian: with
//    S = em + ep and D = ep - em in value units, the gradient is
//    g = S / (2h) and the Lax-Friedrichs jump is ep - em = D / h.
//  * Boundary ends of x, y, vE: setting dm = dp at the first node (dp = dm
//    at the last) makes the centre second difference exactly zero, which the
//    ENO selection then picks, so em = ep = that one-sided difference
//    (native_relative.cpp:47-54) with two selects instead of four.
//  * The fields carry zeroed guard margins of at least 2 * (vE stride)
//    floats on both sides, so every x, y, vE and psi neighbour load is
//    issued unconditionally from a launch-uniform base pointer plus the
//    node's 32-bit byte offset; values read past a boundary are discarded
//    by the selects.  psi is periodic and uniform per block, so its wrapped
//    neighbours are uniform base pointers too.
//  * Blocks whose vE index is at least 2 from both ends run a copy of the
//    column without vE masks; along vP the thread walks the axis with a
//    five-value window and sliding differences, and the interior vP nodes
//    run without vP masks.
//  * The rare arrival-time update stays in a (divergent) branch.
#pragma once

#include <type_traits>

#include "hji5d_stage.h"

namespace pe {

struct Axis {
  // per-thread boundary masks of a non-periodic axis
  bool first, last, has_m2, has_p2;
};

// Load / store at a launch-uniform base plus a 32-bit byte offset (the
// SGPR-base + VGPR-offset addressing form of global memory instructions).
__device__ __forceinline__ float ld(const float *__restrict__ base, unsigned byte_off) {
  return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(base) + byte_off);
}
__device__ __forceinline__ void st(float *__restrict__ base, unsigned byte_off, float v) {
  *reinterpret_cast<float *>(reinterpret_cast<char *>(base) + byte_off) = v;
}

__device__ __forceinline__ Axis axis_masks(int pos, int n) {
  return Axis{pos == 0, pos == n - 1, pos >= 2, pos + 2 < n};
}

// ENO2 one-sided differences in value units from the five values; returns
// s = em + ep and d = ep - em.  MASKED = false for an interior node of a
// non-periodic axis or any node of the periodic one.
template <bool MASKED>
__device__ __forceinline__ void eno2_raw(float v, float m2, float m1, float p1, float p2,
                                         const Axis &a, float &s, float &d) {
  float dm = v - m1, dp = p1 - v;
  if (MASKED) {
    dm = a.first ? dp : dm;
    dp = a.last ? dm : dp;
  }
  const float sc = dp - dm;
  float sl = dm - (m1 - m2), sr = (p2 - p1) - dp;
  if (MASKED) {
    sl = a.has_m2 ? sl : sc;
    sr = a.has_p2 ? sr : sc;
  }
  const float em = fmaf(.5f, fabsf(sl) < fabsf(sc) ? sl : sc, dm);
  const float ep = fmaf(-.5f, fabsf(sc) < fabsf(sr) ? sc : sr, dp);
  s = em + ep;
  d = ep - em;
}

// ENO1 in value units.
template <bool MASKED>
__device__ __forceinline__ void eno1_raw(float v, float m1, float p1, const Axis &a, float &s,
                                         float &d) {
  float dm = v - m1, dp = p1 - v;
  if (MASKED) {
    dm = a.first ? dp : dm;
    dp = a.last ? dm : dp;
  }
  s = dm + dp;
  d = dp - dm;
}

// Per-thread constants of one (vE, psi, y, x) column.
struct Column {
  unsigned j;            // element index of the node at vP index 0
  float ve_cs;           // ve cos(psi)
  float cx, cp;          // 1/(2 dx), 1/(2 dphi)
  float cxy, cxx;        // cx y, cx x
  float cby, caby;       // cx by, cx |by|
  float cay, cax;        // cx |y|, cx |x|
  float wecp;            // we cp
  float e_lo, e_hi;      // elow ce, ehigh ce (control bounds at the vE ends)
  float aev;             // ae ce
  Axis X, Y, E;
};

// The 16 neighbour values of one node (m2, m1, p1, p2 along x, y, vE, psi)
// and, in the blend stage, the previous value.
struct Neighbours {
  float x[4], y[4], e[4], h[4], prev;
};

// Sliding vP window of one column: w2, w3, w4 = V(iv), V(iv + 1), V(iv + 2);
// dpc = the forward difference at iv, dpp = the forward difference at
// iv - 1, c2p / r2p = the second differences at iv - 1 / iv.
struct VpWindow {
  float w2, w3, w4, dpc, dpp, c2p, r2p;

  __device__ __forceinline__ void start(const float *__restrict__ s0, unsigned j, unsigned sv) {
    w2 = ld(s0, j << 2);
    w3 = ld(s0, (j + sv) << 2);
    w4 = ld(s0, (j + 2 * sv) << 2);
    dpc = w3 - w2;
    dpp = c2p = r2p = 0.f;
  }

  // vP slopes of node iv in value units (s = em + ep, d = ep - em); moves
  // the differences on by one node.
  template <int ORDER, bool VINT>
  __device__ __forceinline__ void slopes(int iv, int nvp, float &s, float &d) {
    const float dn = w4 - w3;
    const float r2 = dn - dpc, c2 = r2p, l2 = c2p;
    float dm = dpp, dp = dpc;
    if (ORDER == 1) {
      if (!VINT) {
        dm = iv == 0 ? dp : dm;
        dp = iv == nvp - 1 ? dm : dp;
      }
      s = dm + dp;
      d = dp - dm;
    } else {
      const float sl2 = (VINT || iv >= 2) ? l2 : c2;
      const float sr2 = (VINT || iv + 2 < nvp) ? r2 : c2;
      float em = fmaf(.5f, fabsf(sl2) < fabsf(c2) ? sl2 : c2, dm);
      float ep = fmaf(-.5f, fabsf(c2) < fabsf(sr2) ? c2 : sr2, dp);
      if (!VINT) {
        const bool first = iv == 0, last = iv == nvp - 1;
        em = first ? dp : (last ? dm : em);
        ep = first ? dp : (last ? dm : ep);
      }
      s = em + ep;
      d = ep - em;
    }
    dpp = dpc;
    dpc = dn;
    c2p = c2;
    r2p = r2;
  }

  __device__ __forceinline__ void advance(float w5) {
    w2 = w3;
    w3 = w4;
    w4 = w5;
  }
};

// The slopes of one node along x, y, psi, vE (S = em + ep, D = ep - em in
// value units) and vP (sv, dv).
struct Slopes {
  float Sx, Dx, Sy, Dy, Sp, Dp, Se, De, sv, dv;
};

// Hamiltonian (native_relative.cpp:124-165) of one node in terms of its
// slopes, the RK step and the store; in the final stage also the step's
// maxima and the arrival time.  ve_cs, cby and caby are the psi-dependent
// column constants (those of `c`, or of the paired column).
//   bx gx + by gy - wp |y gx - x gy - gp| + we |gp|
//   + min(plow gv, phigh gv) + max(elow ge, ehigh ge)
//   + 1/2 [(|bx| + wp |y|)(xp - xm) + (|by| + wp |x|)(yp - ym)
//          + (wp + we)(pp - pm) + ap (vp+ - vp-) + ae (ve+ - ve-)]
template <bool BLEND, bool VINT, bool EINT>
__device__ __forceinline__ void update_node(const StageArgs<float> &a, const Column &c,
                                            float ve_cs, float cby, float caby, int iv,
                                            int nvp, float apv, float v, float prev,
                                            const Slopes &k, unsigned jb, int final_stage,
                                            float &inc_max, float &chg_max, bool &nonfinite) {
  const float vp = a.vpk[2 * iv], wp = a.vpk[2 * iv + 1];  // uniform: scalar loads
  const float bxc = (ve_cs - vp) * c.cx;
  float H = fmaf(bxc, k.Sx, cby * k.Sy);
  const float inner = fmaf(c.cxy, k.Sx, fmaf(-c.cxx, k.Sy, -c.cp * k.Sp));
  H = fmaf(-wp, fabsf(inner), H);
  H = fmaf(c.wecp, fabsf(k.Sp), H);
  // Control terms: at interior speeds min(-a g, a g) = -a |g| and
  // max(-a g, a g) = a |g| exactly.
  if (VINT) {
    H = fmaf(-apv, fabsf(k.sv), H);
  } else {
    const float p_lo = iv == 0 ? 0.f : -apv, p_hi = iv == nvp - 1 ? 0.f : apv;
    const float pa = p_lo * k.sv, pb = p_hi * k.sv;
    H += pb < pa ? pb : pa;
  }
  if (EINT) {
    H = fmaf(c.aev, fabsf(k.Se), H);
  } else {
    const float ea = c.e_lo * k.Se, eb = c.e_hi * k.Se;
    H += ea < eb ? eb : ea;
  }
  H = fmaf(fmaf(wp, c.cay, fabsf(bxc)), k.Dx, H);
  H = fmaf(fmaf(wp, c.cax, caby), k.Dy, H);
  H = fmaf(fmaf(wp, c.cp, c.wecp), k.Dp, H);
  H = fmaf(apv, k.dv, H);
  H = fmaf(c.aev, k.De, H);
  const float previous = BLEND ? prev : v;
  float value = fmaf(a.dt, H < 0.f ? H : 0.f, v);
  if (BLEND) value = .5f * (previous + value);
  nonfinite |= !isfinite(value);
  st(a.dst, jb, value);
  if (final_stage) {
    // Maxima on the bit patterns: inc_max >= 0, so signed integer order
    // agrees with float order against any inc; |inc| >= 0 likewise.
    const float inc = value - previous;
    inc_max = __int_as_float(max(__float_as_int(inc_max), __float_as_int(inc)));
    chg_max = __int_as_float(max(__float_as_int(chg_max), __float_as_int(fabsf(inc))));
    if (previous > 0.f && value <= 0.f) {
      const float old_arrival = ld(a.arrival, jb);
      if (isnan(old_arrival))
        st(a.arrival, jb,
           float(a.t + double(a.dt) * double(previous) / (double(previous) - double(value))));
    }
  }
}

// ENO slopes along x, y and vE of one node from its neighbours.
template <int ORDER, bool EINT>
__device__ __forceinline__ void xye_slopes(float v, const float x[4], const float y[4],
                                           const float e[4], const Column &c, Slopes &k) {
  if (ORDER == 1) {
    eno1_raw<true>(v, x[1], x[2], c.X, k.Sx, k.Dx);
    eno1_raw<true>(v, y[1], y[2], c.Y, k.Sy, k.Dy);
    eno1_raw<!EINT>(v, e[1], e[2], c.E, k.Se, k.De);
  } else {
    eno2_raw<true>(v, x[0], x[1], x[2], x[3], c.X, k.Sx, k.Dx);
    eno2_raw<true>(v, y[0], y[1], y[2], y[3], c.Y, k.Sy, k.Dy);
    eno2_raw<!EINT>(v, e[0], e[1], e[2], e[3], c.E, k.Se, k.De);
  }
}

// Pointers to the x, y and vE neighbours at -2, -1, +1, +2 of a node whose
// byte offset from `s0` is added at the load.
struct NeighbourBases {
  const float *__restrict__ x[4];
  const float *__restrict__ y[4];
  const float *__restrict__ e[4];

  __device__ __forceinline__ NeighbourBases(const float *__restrict__ s0, int nx, int se)
      : x{s0 - 2, s0 - 1, s0 + 1, s0 + 2},
        y{s0 - 2 * nx, s0 - nx, s0 + nx, s0 + 2 * nx},
        e{s0 - 2 * se, s0 - se, s0 + se, s0 + 2 * se} {}

  template <int ORDER>
  __device__ __forceinline__ void load(unsigned jb, float xv[4], float yv[4], float ev[4]) const {
    if (ORDER == 2) {
      xv[0] = ld(x[0], jb);
      xv[3] = ld(x[3], jb);
      yv[0] = ld(y[0], jb);
      yv[3] = ld(y[3], jb);
      ev[0] = ld(e[0], jb);
      ev[3] = ld(e[3], jb);
    }
    xv[1] = ld(x[1], jb);
    xv[2] = ld(x[2], jb);
    yv[1] = ld(y[1], jb);
    yv[2] = ld(y[2], jb);
    ev[1] = ld(e[1], jb);
    ev[2] = ld(e[2], jb);
  }
};

template <bool BLEND, int ORDER, bool EINT, bool PREFETCH>
__device__ __forceinline__ void hji5d_column(const StageArgs<float> &a, const Hji5dGrid &g,
                                             const Column &c, const float *__restrict__ hb[4],
                                             int final_stage, float &inc_max, float &chg_max,
                                             bool &nonfinite) {
  const float *__restrict__ s0 = a.src;
  const NeighbourBases nbase(s0, g.nx, (int)g.se);
  const float cv = .5f / float(g.dvp), apv = float(g.ap) * cv;
  const int nvp = g.nvp;
  const unsigned sv = (unsigned)g.sv;
  unsigned j = c.j;
  VpWindow win;
  win.start(s0, j, sv);
  auto load = [&](unsigned jb, Neighbours &n) {
    nbase.load<ORDER>(jb, n.x, n.y, n.e);
    if (ORDER == 2) {
      n.h[0] = ld(hb[0], jb);
      n.h[3] = ld(hb[3], jb);
    }
    n.h[1] = ld(hb[1], jb);
    n.h[2] = ld(hb[2], jb);
    if (BLEND) n.prev = ld(a.dst, jb);
  };
  // With PREFETCH the neighbours of the next vP node are requested before
  // the current node is computed, so each wave overlaps its own load
  // latency with its arithmetic.
  Neighbours cur;
  if (PREFETCH) load(j << 2, cur);

  auto node = [&](int iv, auto interior) {
    constexpr bool VINT = decltype(interior)::value;
    const float w5 = iv + 3 < nvp ? ld(s0, (j + 3 * sv) << 2) : 0.f;
    Neighbours nb;
    if (PREFETCH) {
      nb = cur;
      if (VINT || iv + 1 < nvp) load((j + sv) << 2, cur);
    } else {
      load(j << 2, nb);
    }
    Slopes k;
    win.slopes<ORDER, VINT>(iv, nvp, k.sv, k.dv);
    const float v = win.w2;
    xye_slopes<ORDER, EINT>(v, nb.x, nb.y, nb.e, c, k);
    if (ORDER == 1)
      eno1_raw<false>(v, nb.h[1], nb.h[2], c.E, k.Sp, k.Dp);
    else
      eno2_raw<false>(v, nb.h[0], nb.h[1], nb.h[2], nb.h[3], c.E, k.Sp, k.Dp);
    update_node<BLEND, VINT, EINT>(a, c, c.ve_cs, c.cby, c.caby, iv, nvp, apv, v, nb.prev, k,
                                   j << 2, final_stage, inc_max, chg_max, nonfinite);
    win.advance(w5);
    j += sv;
  };

  int iv = 0;
  for (; iv < 2 && iv < nvp; iv++) node(iv, std::false_type{});
#pragma unroll 2
  for (; iv + 2 < nvp; iv++) node(iv, std::true_type{});
  for (; iv < nvp; iv++) node(iv, std::false_type{});
}

// The psi-dependent constants of the second column of a pair.
struct PsiConstants {
  float ve_cs, cby, caby;
};

// Neighbour values of a node pair at psi planes ih and ih + 1 (ih even):
// x, y, vE neighbours of each node, the psi planes ih - 2, ih - 1, ih + 2,
// ih + 3 shared by both, and in the blend stage the previous values.
struct PairNeighbours {
  float xa[4], ya[4], ea[4], xb[4], yb[4], eb[4], h[4], preva, prevb;
};

// Two columns at the adjacent psi planes ih (even) and ih + 1, walked
// together along vP.  Their psi neighbours overlap: the pair loads planes
// ih - 2, ih - 1, ih + 2 and ih + 3, and each node takes the other's value
// as its psi neighbour at +1 or -1, so a node pair issues 4 psi loads
// instead of 8 (16 loads per node in the blend stage instead of 18).
// Every node is computed exactly as by hji5d_column.  The neighbours are
// loaded just before each node pair: requesting the next pair's early, as
// hji5d_column does, needs 155 VGPRs (3 waves per SIMD) and was slower.
template <bool BLEND, int ORDER, bool EINT>
__device__ __forceinline__ void hji5d_column_pair(const StageArgs<float> &a, const Hji5dGrid &g,
                                                  const Column &c, const PsiConstants &cb,
                                                  const float *__restrict__ hb[4],
                                                  int final_stage, float &inc_max,
                                                  float &chg_max, bool &nonfinite) {
  const float *__restrict__ s0 = a.src;
  const NeighbourBases nbase(s0, g.nx, (int)g.se);
  const float cv = .5f / float(g.dvp), apv = float(g.ap) * cv;
  const int nvp = g.nvp;
  const unsigned sv = (unsigned)g.sv, pb = (unsigned)g.nx2 << 2;  // column b: + pb bytes
  unsigned j = c.j;
  VpWindow wa, wb;
  wa.start(s0, j, sv);
  wb.start(s0, j + (unsigned)g.nx2, sv);
  auto load = [&](unsigned jb, PairNeighbours &n) {
    nbase.load<ORDER>(jb, n.xa, n.ya, n.ea);
    nbase.load<ORDER>(jb + pb, n.xb, n.yb, n.eb);
    if (ORDER == 2) {
      n.h[0] = ld(hb[0], jb);
      n.h[3] = ld(hb[3], jb);
    }
    n.h[1] = ld(hb[1], jb);
    n.h[2] = ld(hb[2], jb);
    if (BLEND) {
      n.preva = ld(a.dst, jb);
      n.prevb = ld(a.dst, jb + pb);
    }
  };
  auto node = [&](int iv, auto interior) {
    constexpr bool VINT = decltype(interior)::value;
    const bool more = iv + 3 < nvp;
    const float w5a = more ? ld(s0, (j + 3 * sv) << 2) : 0.f;
    const float w5b = more ? ld(s0, ((j + 3 * sv) << 2) + pb) : 0.f;
    PairNeighbours nb;
    load(j << 2, nb);
    Slopes ka, kb;
    wa.slopes<ORDER, VINT>(iv, nvp, ka.sv, ka.dv);
    wb.slopes<ORDER, VINT>(iv, nvp, kb.sv, kb.dv);
    const float va = wa.w2, vb = wb.w2;
    xye_slopes<ORDER, EINT>(va, nb.xa, nb.ya, nb.ea, c, ka);
    xye_slopes<ORDER, EINT>(vb, nb.xb, nb.yb, nb.eb, c, kb);
    if (ORDER == 1) {
      eno1_raw<false>(va, nb.h[1], vb, c.E, ka.Sp, ka.Dp);
      eno1_raw<false>(vb, va, nb.h[2], c.E, kb.Sp, kb.Dp);
    } else {
      eno2_raw<false>(va, nb.h[0], nb.h[1], vb, nb.h[2], c.E, ka.Sp, ka.Dp);
      eno2_raw<false>(vb, nb.h[1], va, nb.h[2], nb.h[3], c.E, kb.Sp, kb.Dp);
    }
    const unsigned jb = j << 2;
    update_node<BLEND, VINT, EINT>(a, c, c.ve_cs, c.cby, c.caby, iv, nvp, apv, va, nb.preva, ka,
                                   jb, final_stage, inc_max, chg_max, nonfinite);
    update_node<BLEND, VINT, EINT>(a, c, cb.ve_cs, cb.cby, cb.caby, iv, nvp, apv, vb, nb.prevb,
                                   kb, jb + pb, final_stage, inc_max, chg_max, nonfinite);
    wa.advance(w5a);
    wb.advance(w5b);
    j += sv;
  };

  int iv = 0;
  for (; iv < 2 && iv < nvp; iv++) node(iv, std::false_type{});
#pragma unroll 2
  for (; iv + 2 < nvp; iv++) node(iv, std::true_type{});
  for (; iv < nvp; iv++) node(iv, std::false_type{});
}

// PAIR: each thread walks the columns of two adjacent psi planes (ih even
// and ih + 1; nh even), the block index counts plane pairs (PREFETCH is
// then unused).
template <bool BLEND, int ORDER, bool PREFETCH, bool PAIR>
__global__ __launch_bounds__(256) void hji5d_stage_fast(StageArgs<float> a, Hji5dGrid g,
                                                        int final_stage) {
  const int hw = blockIdx.x;
  const int block = (hw & 7) * g.per_xcd + (hw >> 3);
  float inc_max = 0.f, chg_max = 0.f;
  bool nonfinite = false;
  if (block < g.blocks) {
    // Block order, fastest first: vE, then psi (or psi pair) and the xy
    // tile (psi before the tile by default; the tile before psi with
    // psi_outer).
    const int nhb = PAIR ? g.nh / 2 : g.nh;
    const int ie = block % g.nve;
    const int rest = block / g.nve;
    const int ihb = g.psi_outer ? rest / g.xytiles : rest % nhb;
    const int tile = g.psi_outer ? rest % g.xytiles : rest / nhb;
    const int ih = PAIR ? 2 * ihb : ihb;
    const int r = tile * 256 + (int)threadIdx.x;  // (y, x) index
    if (r < g.nx2) {
      const int iy = r / g.nx;
      const int ix = r - iy * g.nx;
      const float dx = float(g.dx), dphi = float(g.dphi), dve = float(g.dve);
      const float x = float(-g.extent) + float(ix) * dx;
      const float y = float(-g.extent) + float(iy) * dx;
      const unsigned j0 = (unsigned)(ie * (int)g.se + ih * g.nx2 + r);
      if (g.exit_on && exit_outside_f32(g, ix, iy)) {
        // Outside the local exit radius every vP node takes the exit value
        // (gpu_relative.py:126-131).
        const float ev = g.exit_valuef;
        for (int iv = 0; iv < g.nvp; iv++) {
          a.dst[j0 + (unsigned)iv * (unsigned)g.sv] = ev;
          if (PAIR) a.dst[j0 + (unsigned)g.nx2 + (unsigned)iv * (unsigned)g.sv] = ev;
        }
      } else {
        const float ve = float(g.vemin) + float(ie) * dve;
        const float cs = a.cs[ih], sn = a.sn[ih];
        const float by = ve * sn;
        Column c;
        c.j = j0;
        c.ve_cs = ve * cs;
        c.cx = .5f / dx;
        c.cp = .5f / dphi;
        const float ce = .5f / dve;
        c.cxy = c.cx * y;
        c.cxx = c.cx * x;
        c.cby = c.cx * by;
        c.caby = c.cx * fabsf(by);
        c.cay = c.cx * fabsf(y);
        c.cax = c.cx * fabsf(x);
        const float we = fabsf(ve) * float(g.ke), ae = float(g.ae);
        c.wecp = we * c.cp;
        c.e_lo = (ie == 0 ? 0.f : -ae) * ce;
        c.e_hi = (ie == g.nve - 1 ? 0.f : ae) * ce;
        c.aev = ae * ce;
        c.X = axis_masks(ix, g.nx);
        c.Y = axis_masks(iy, g.nx);
        c.E = axis_masks(ie, g.nve);
        // psi neighbours (periodic; ih is uniform per block): planes ih - 2,
        // ih - 1, ih + 1, ih + 2, or for a pair ih - 2, ih - 1, ih + 2, ih + 3
        const int nh = g.nh, nx2 = g.nx2;
        const int hm1 = ih == 0 ? nh - 1 : ih - 1;
        const int hm2 = ih >= 2 ? ih - 2 : ih - 2 + nh;
        const int hp1 = ih == nh - 1 ? 0 : ih + 1;
        const int hp2 = ih + 2 < nh ? ih + 2 : ih + 2 - nh;
        const int hp3 = ih + 3 < nh ? ih + 3 : ih + 3 - nh;
        const float *hb[4] = {a.src + (hm2 - ih) * nx2, a.src + (hm1 - ih) * nx2,
                              a.src + ((PAIR ? hp2 : hp1) - ih) * nx2,
                              a.src + ((PAIR ? hp3 : hp2) - ih) * nx2};
        const bool eint = ie >= 2 && ie + 2 < g.nve;
        if (PAIR) {
          const float by1 = ve * a.sn[ih + 1];
          const PsiConstants cb{ve * a.cs[ih + 1], c.cx * by1, c.cx * fabsf(by1)};
          if (eint)
            hji5d_column_pair<BLEND, ORDER, true>(a, g, c, cb, hb, final_stage, inc_max,
                                                  chg_max, nonfinite);
          else
            hji5d_column_pair<BLEND, ORDER, false>(a, g, c, cb, hb, final_stage, inc_max,
                                                   chg_max, nonfinite);
        } else if (eint) {
          hji5d_column<BLEND, ORDER, true, PREFETCH>(a, g, c, hb, final_stage, inc_max,
                                                     chg_max, nonfinite);
        } else {
          hji5d_column<BLEND, ORDER, false, PREFETCH>(a, g, c, hb, final_stage, inc_max,
                                                      chg_max, nonfinite);
        }
      }
    }
  }
  if (!final_stage) return;
  if (nonfinite) atomicMax(a.slot + 2, 1ull);
  __shared__ float red[2][4];
  const float i_w = wave_max(inc_max), c_w = wave_max(chg_max);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][wave] = i_w;
    red[1][wave] = c_w;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float i_b = red[0][0], c_b = red[1][0];
    for (int k = 1; k < 4; k++) {
      i_b = fmaxf(i_b, red[0][k]);
      c_b = fmaxf(c_b, red[1][k]);
    }
    raise_max(a.slot + 0, double(i_b));
    raise_max(a.slot + 1, double(c_b));
  }
}

}  // namespace pe
