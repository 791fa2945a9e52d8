// FP32 RK stage of the 5D relative HJI update, written for instruction
// economy on gfx950 (the stage is VALU-issue bound at about 200
// instructions per node in the generic template of hji5d_stage.h).
// Measured on C1: 151 us (stage 1) and 178 us (blend) per launch against
// 195 / 232 us for the generic template.  Wave-uniform branches around the
// y / vE / vP boundary cases instead of selects were slower (176 / 196 us).
//
// Same arithmetic as hji5d_stage<float, false, *> node by node
// (reversible/native_relative.cpp:38-67 slopes, :124-172 update), reorganised:
//  * The fields carry zeroed guard margins of at least 2 * (vE stride)
//    floats on both sides, so every x, y and vE neighbour load is issued
//    unconditionally from a launch-uniform base pointer plus the node's
//    32-bit index; values read past a boundary are discarded by selects.
//    psi is periodic and uses per-thread wrapped offsets.
//  * ENO boundary cases are per-thread masks computed once, applied with
//    selects instead of branches.
//  * Along vP the thread walks the axis with a five-value window, and the
//    first and second differences slide with it: one new difference of
//    each kind per node instead of two and three.
//  * The rare arrival-time update stays in a (divergent) branch.
#pragma once

#include "hji5d_stage.h"

namespace pe {

struct Axis {
  // per-thread boundary masks of a non-periodic axis
  bool first, last, has_m2, has_p2;
};

// Load at a launch-uniform base plus a 32-bit byte offset (the SGPR-base +
// VGPR-offset addressing form of global loads).
__device__ __forceinline__ float ld(const float *__restrict__ base, unsigned byte_off) {
  return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(base) + byte_off);
}

__device__ __forceinline__ Axis axis_masks(int pos, int n) {
  return Axis{pos == 0, pos == n - 1, pos >= 2, pos + 2 < n};
}

// ENO2 slopes from the five values (non-periodic axis, masks per thread).
__device__ __forceinline__ void eno2(float v, float m2, float m1, float p1, float p2,
                                     const Axis &a, float inv_h, float inv_hh,
                                     float half_h, float &dm, float &dp) {
  const float Dm = (v - m1) * inv_h, Dp = (p1 - v) * inv_h;
  const float sc = (p1 - 2.f * v + m1) * inv_hh;
  float sl = (v - 2.f * m1 + m2) * inv_hh;
  float sr = (p2 - 2.f * p1 + v) * inv_hh;
  sl = a.has_m2 ? sl : sc;
  sr = a.has_p2 ? sr : sc;
  float em = Dm + half_h * (fabsf(sl) < fabsf(sc) ? sl : sc);
  float ep = Dp - half_h * (fabsf(sc) < fabsf(sr) ? sc : sr);
  // one-sided raw differences at the ends (native_relative.cpp:47-54)
  em = a.first ? Dp : (a.last ? Dm : em);
  ep = a.first ? Dp : (a.last ? Dm : ep);
  dm = em;
  dp = ep;
}

__device__ __forceinline__ void eno1(float v, float m1, float p1, const Axis &a,
                                     float inv_h, float &dm, float &dp) {
  const float Dm = (v - m1) * inv_h, Dp = (p1 - v) * inv_h;
  dm = a.first ? Dp : (a.last ? Dm : Dm);
  dp = a.first ? Dp : (a.last ? Dm : Dp);
}

template <bool BLEND, int ORDER>
__global__ __launch_bounds__(256) void hji5d_stage_fast(StageArgs<float> a, Hji5dGrid g,
                                                        int final_stage) {
  const int hw = blockIdx.x;
  const int block = (hw & 7) * g.per_xcd + (hw >> 3);
  float inc_max = 0.f, chg_max = 0.f;
  bool nonfinite = false;
  if (block < g.blocks) {
    const int ie = block % g.nve;
    const int rest = block / g.nve;
    const int ih = rest % g.nh;
    const int r = (rest / g.nh) * 256 + (int)threadIdx.x;  // (y, x) index
    if (r < g.nx2) {
      const int iy = r / g.nx;
      const int ix = r - iy * g.nx;
      const float dx = float(g.dx), dphi = float(g.dphi), dvp = float(g.dvp),
                  dve = float(g.dve);
      const float ix_h = 1.f / dx, ip_h = 1.f / dphi, iv_h = 1.f / dvp, ie_h = 1.f / dve;
      const float ix_hh = ix_h * ix_h, ip_hh = ip_h * ip_h, iv_hh = iv_h * iv_h,
                  ie_hh = ie_h * ie_h;
      const float hx = .5f * dx, hp = .5f * dphi, hv = .5f * dvp, hev = .5f * dve;
      const float x = float(-g.extent) + float(ix) * dx;
      const float y = float(-g.extent) + float(iy) * dx;
      const float ve = float(g.vemin) + float(ie) * dve;
      const float ap = float(g.ap), ae = float(g.ae), kp = float(g.kp);
      const float we = fabsf(ve) * float(g.ke);
      const float cs = a.cs[ih], sn = a.sn[ih];
      const float by = ve * sn, aby = fabsf(by), ax = fabsf(x), ay = fabsf(y);
      const float elow = ie == 0 ? 0.f : -ae, ehigh = ie == g.nve - 1 ? 0.f : ae;
      const bool outside = g.exit_on && double(x * x + y * y) > g.exit_r2;
      const Axis X = axis_masks(ix, g.nx), Y = axis_masks(iy, g.nx),
                 E = axis_masks(ie, g.nve);
      // psi neighbours (periodic), per-thread wrapped offsets.
      const int hm1 = ih == 0 ? g.nh - 1 : ih - 1;
      const int hm2 = ih >= 2 ? ih - 2 : ih - 2 + g.nh;
      const int hp1 = ih == g.nh - 1 ? 0 : ih + 1;
      const int hp2 = ih + 2 < g.nh ? ih + 2 : ih + 2 - g.nh;
      const unsigned ohm1_b = (unsigned)((hm1 - ih) * g.nx2) << 2,
                     ohm2_b = (unsigned)((hm2 - ih) * g.nx2) << 2,
                     ohp1_b = (unsigned)((hp1 - ih) * g.nx2) << 2,
                     ohp2_b = (unsigned)((hp2 - ih) * g.nx2) << 2;
      // Launch-uniform neighbour bases (guard margins keep them in range).
      const float *__restrict__ s0 = a.src;
      const int nx = g.nx, se = (int)g.se;
      const float *__restrict__ sxm2 = s0 - 2, *__restrict__ sxm1 = s0 - 1,
                                *__restrict__ sxp1 = s0 + 1, *__restrict__ sxp2 = s0 + 2;
      const float *__restrict__ sym2 = s0 - 2 * nx, *__restrict__ sym1 = s0 - nx,
                                *__restrict__ syp1 = s0 + nx, *__restrict__ syp2 = s0 + 2 * nx;
      const float *__restrict__ sem2 = s0 - 2 * se, *__restrict__ sem1 = s0 - se,
                                *__restrict__ sep1 = s0 + se, *__restrict__ sep2 = s0 + 2 * se;
      const unsigned sv = (unsigned)g.sv;
      unsigned j = (unsigned)(ie * se + ih * g.nx2 + r);  // node at vP index 0
      // vP window w1..w4 = V(iv-1 .. iv+2) and its sliding differences:
      // at node iv, Dm = D(iv-1/2), Dp = D(iv+1/2), l2 = d2(iv-1),
      // c2 = d2(iv), r2 = d2(iv+1); each step computes only Dp and r2.
      float w1 = 0.f, w2 = ld(s0, j << 2);
      float w3 = ld(s0, (j + sv) << 2);
      float w4 = ld(s0, (j + 2 * sv) << 2);
      float Dp_prev = 0.f, c2_prev = 0.f, r2_prev = 0.f;
      for (int iv = 0; iv < g.nvp; iv++, j += sv) {
        const float w5 = iv + 3 < g.nvp ? ld(s0, (j + 3 * sv) << 2) : 0.f;
        const float Dm = Dp_prev, Dp = (w3 - w2) * iv_h;
        const float l2 = c2_prev, c2 = r2_prev, r2 = (w4 - 2.f * w3 + w2) * iv_hh;
        Dp_prev = Dp;
        c2_prev = c2;
        r2_prev = r2;
        const bool first = iv == 0, last = iv == g.nvp - 1;
        float vm, vpl;
        if (ORDER == 1) {
          vm = first ? Dp : Dm;
          vpl = last ? Dm : Dp;
        } else {
          const float sl2 = iv >= 2 ? l2 : c2;
          const float sr2 = iv + 2 < g.nvp ? r2 : c2;
          const float em = Dm + hv * (fabsf(sl2) < fabsf(c2) ? sl2 : c2);
          const float ep = Dp - hv * (fabsf(c2) < fabsf(sr2) ? c2 : sr2);
          vm = first ? Dp : (last ? Dm : em);
          vpl = first ? Dp : (last ? Dm : ep);
        }
        const unsigned jb = j << 2;
        if (outside) {
          a.dst[j] = float(g.exit_value);
        } else {
          const float v = w2;
          float xm, xp, ym, yp, pm, pp, em, ep;
          if (ORDER == 1) {
            eno1(v, ld(sxm1, jb), ld(sxp1, jb), X, ix_h, xm, xp);
            eno1(v, ld(sym1, jb), ld(syp1, jb), Y, ix_h, ym, yp);
            eno1(v, ld(sem1, jb), ld(sep1, jb), E, ie_h, em, ep);
            const float hm = ld(s0, jb + ohm1_b), hpv = ld(s0, jb + ohp1_b);
            pm = (v - hm) * ip_h;
            pp = (hpv - v) * ip_h;
          } else {
            eno2(v, ld(sxm2, jb), ld(sxm1, jb), ld(sxp1, jb), ld(sxp2, jb), X, ix_h, ix_hh, hx, xm, xp);
            eno2(v, ld(sym2, jb), ld(sym1, jb), ld(syp1, jb), ld(syp2, jb), Y, ix_h, ix_hh, hx, ym, yp);
            eno2(v, ld(sem2, jb), ld(sem1, jb), ld(sep1, jb), ld(sep2, jb), E, ie_h, ie_hh, hev, em, ep);
            const float hm2v = ld(s0, jb + ohm2_b), hm1v = ld(s0, jb + ohm1_b), hp1v = ld(s0, jb + ohp1_b),
                        hp2v = ld(s0, jb + ohp2_b);
            const Axis P{false, false, true, true};
            eno2(v, hm2v, hm1v, hp1v, hp2v, P, ip_h, ip_hh, hp, pm, pp);
          }
          const float vp = float(g.vpmin) + float(iv) * dvp;
          const float wp = fabsf(vp) * kp;
          const float bx = ve * cs - vp;
          const float gx = .5f * (xm + xp), gy = .5f * (ym + yp), gp = .5f * (pm + pp),
                      gv = .5f * (vm + vpl), ge = .5f * (em + ep);
          const float plow = iv == 0 ? 0.f : -ap, phigh = iv == g.nvp - 1 ? 0.f : ap;
          const float pa = plow * gv, pb = phigh * gv, ea = elow * ge, eb = ehigh * ge;
          float H = bx * gx + by * gy - wp * fabsf(y * gx - x * gy - gp) + we * fabsf(gp) +
                    (pb < pa ? pb : pa) + (ea < eb ? eb : ea);
          H += .5f * ((fabsf(bx) + wp * ay) * (xp - xm) + (aby + wp * ax) * (yp - ym) +
                      (wp + we) * (pp - pm) + ap * (vpl - vm) + ae * (ep - em));
          const float old = v;
          const float previous = BLEND ? a.dst[j] : old;
          float value = old + a.dt * (H < 0.f ? H : 0.f);
          if (BLEND) value = .5f * (previous + value);
          nonfinite |= !isfinite(value);
          a.dst[j] = value;
          if (final_stage) {
            const float inc = value - previous;
            inc_max = fmaxf(inc_max, inc);
            chg_max = fmaxf(chg_max, fabsf(inc));
            if (previous > 0.f && value <= 0.f) {
              const float old_arrival = a.arrival[j];
              if (isnan(old_arrival))
                a.arrival[j] = float(a.t + double(a.dt) * double(previous) /
                                              (double(previous) - double(value)));
            }
          }
        }
        w1 = w2;
        w2 = w3;
        w3 = w4;
        w4 = w5;
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
