// Dense eight-state reversible-car game in the fixed square on the GPU
// (the native_world8d backend, reversible/native_world.cpp:32-314, behind
// reversible/world_dp.py:177-426).
//
// State [xP, yP, thP, vP, xE, yE, thE, vE], row-major with vE fastest, float
// fields and float local arithmetic.  This file is built without
// floating-point contraction and keeps the reference's operation order, so
// every node update is the reference's, bit for bit:
//   * ENO1/ENO2 slopes per axis; the ENO2 selection uses <= (not <); the
//     position axes extrapolate past their ends with the signed edge slope,
//     headings are periodic, speeds use inward one-sided slopes at their
//     ends (linear speed boundary) or the signed extrapolation (legacy);
//   * Hamiltonian with local Lax-Friedrichs dissipation; RK1 or RK2;
//   * on the completed update, the tube projection max(min(value, previous),
//     avoid) with avoid the pursuer wall fault, and arrival times at the
//     crossing of zero.
// The initial field (capture or evader wall fault, max with the avoid
// surface) is computed on the host with the same float operations.
//
// Layout of a launch: one block per (xP, yP, thP, vP) slab, or per range of
// xE rows of it when the slab is above 36 KiB, whose inner (xE, yE, thE, vE)
// values (plus two xE rows on either side) are staged in LDS, so the pursuer
// coordinates and everything derived from them are uniform per block and
// the evader-axis neighbours are LDS reads.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "pe_common.h"

namespace pe {

// Axes in spacing classes: positions (0, 1, 4, 5), headings (2, 6), P speed
// (3), E speed (7).  Per-class constants keep the kernels' uniform values
// (and their scalar registers) few.
__host__ __device__ constexpr int w8_cls(int k) {
  return (k == 2 || k == 6) ? 1 : k == 3 ? 2 : k == 7 ? 3 : 0;
}

struct W8 {
  int n[8];              // axis lengths
  long long stride[8];   // element strides
  int s32[8];            // the same as 32-bit values (fields below 2^31 nodes)
  // per spacing class: h, h * h (float, as the scheme forms it), their
  // reciprocals rounded to nearest and the remainders (r + rl = 1/h to about
  // 2^-48)
  float h[4], hh[4], rh[4], rhh[4], rhl[4], rhhl[4];
  long long N;           // nodes
  int inner;             // nodes of one pursuer slab: n * n * nh * ne
  int inner_blocks;      // ceil(inner / 256)
  int order, linear;     // ENO order, linear speed boundary
  float L, ap, ae, kp, ke;
  // Inner-index decoding by multiply-high: r / n[k] = umulhi(r, magic[k])
  // for every r < inner when fast_decode (inner * n[k] < 2^32).
  unsigned magic[8];
  int fast_decode;
  int fast_division;     // divisors within [2^-20, 2^20]: see w8_div
  int part_rows;         // xE rows per block of the LDS kernel
  int slab_parts;        // blocks per slab: ceil(n / part_rows)
};

__device__ __forceinline__ float w8_mn(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float w8_mx(float a, float b) { return a < b ? b : a; }

// a / b, correctly rounded.  FAST: b > 0 is one of the grid spacings or
// their squares, y = RN(1/b) and y + ylo = 1/b to about 2^-48 relative
// (both host-computed).  q0 = RN(a y + RN(a ylo)) lies within half an ulp
// of a/b plus a term of relative size 2^-47, so within one ulp; with y
// within half an ulp of 1/b and q0 within one ulp of a/b, the remainder
// a - b q0 is exact and RN(q0 + (a - b q0) y) is the correctly rounded
// quotient (Markstein's theorem, P. Markstein, IBM J. Res. Dev. 34(1),
// 1990).  The theorem needs a normal quotient and an exact remainder: the
// kernels take this path only while every value of the field they read is
// +0 or finite with magnitude in [2^-60, 2^60], so that every non-zero
// numerator (a difference, second difference or extrapolation of such
// values) lies in [2^-83, 2^64]; a numerator +0 gives +0 as a / b does and
// -0 never occurs (the flag counts -0 as out of range).  Otherwise the IEEE
// division is used.
template <bool FAST>
__device__ __forceinline__ float w8_div(float a, float b, float y, float ylo) {
  if (!FAST) return a / b;
  const float q0 = fmaf(a, y, a * ylo);
  return fmaf(fmaf(-q0, b, a), y, q0);
}

// Whether a value keeps the next stage on the fast division (see w8_div).
__device__ __forceinline__ bool w8_in_range(float v) {
  const float a = fabsf(v);
  return (a >= 0x1p-60f && a <= 0x1p60f) || __float_as_uint(v) == 0u;
}

// Field accessors: global memory by node index (32-bit byte offsets from the
// field's base when the field is below 4 GiB), or a slab staged in LDS by
// index within the slab.  Same values, so the same arithmetic either way.
template <typename I>
struct W8Global {
  const float *__restrict__ p;
  __device__ __forceinline__ float operator()(I i) const {
    if constexpr (sizeof(I) == 4)
      return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(p) +
                                              ((unsigned)i << 2));
    else
      return p[i];
  }
};
struct W8Lds {
  const float *p;  // LDS values by inner index (offset by the staged rows)
  __device__ __forceinline__ float operator()(int i) const { return p[i]; }
};

// Value `off` cells from node j along one axis: periodic wrap, or the signed
// linear extrapolation of the edge values past a non-periodic end.
template <typename I, typename V>
__device__ __forceinline__ float w8_at(const V &v, I j, I stride, int pos, int n, int off,
                                       bool periodic) {
  int q = pos + off;
  if (periodic) {  // |off| <= 2 < n
    q = q < 0 ? q + n : q >= n ? q - n : q;
    return v(j + (I)(q - pos) * stride);
  }
  if (q < 0) {
    const I edge = j - (I)pos * stride;
    const float a = v(edge), b = v(edge + stride);
    const float sign = (float)((a > 0) - (a < 0));
    return a + (float)(-q) * fabsf(b - a) * sign;
  }
  if (q >= n) {
    const I edge = j + (I)(n - 1 - pos) * stride;
    const float a = v(edge), b = v(edge - stride);
    const float sign = (float)((a > 0) - (a < 0));
    return a + (float)(q - n + 1) * fabsf(a - b) * sign;
  }
  return v(j + (I)off * stride);
}

// Left and right ENO slopes of the node j (value f) along axis k.
template <bool FAST, typename I, typename V>
__device__ __forceinline__ void w8_slopes(const V &v, I j, float f, int k, int pos,
                                          const W8 &g, bool periodic, bool linear, float &left,
                                          float &right) {
  const I stride = sizeof(I) == 4 ? (I)g.s32[k] : (I)g.stride[k];
  const int n = g.n[k];
  const int c = w8_cls(k);
  const float h = g.h[c], rh = g.rh[c], rhl = g.rhl[c];
  if (linear && pos == 0) {
    left = right = w8_div<FAST>(v(j + stride) - f, h, rh, rhl);
    return;
  }
  if (linear && pos == n - 1) {
    left = right = w8_div<FAST>(f - v(j - stride), h, rh, rhl);
    return;
  }
  const float fm = w8_at(v, j, stride, pos, n, -1, periodic),
              fp = w8_at(v, j, stride, pos, n, 1, periodic);
  float l = w8_div<FAST>(f - fm, h, rh, rhl), r = w8_div<FAST>(fp - f, h, rh, rhl);
  if (g.order == 2) {
    const float hh = g.hh[c], rhh = g.rhh[c], rhhl = g.rhhl[c];
    const float dc = w8_div<FAST>(fp - 2 * f + fm, hh, rhh, rhhl);
    const float dl = linear && pos < 2
                         ? dc
                         : w8_div<FAST>(f - 2 * fm + w8_at(v, j, stride, pos, n, -2, periodic),
                                        hh, rhh, rhhl);
    const float dr = linear && pos + 2 >= n
                         ? dc
                         : w8_div<FAST>(w8_at(v, j, stride, pos, n, 2, periodic) - 2 * fp + f,
                                        hh, rhh, rhhl);
    l += .5f * h * (fabsf(dl) <= fabsf(dc) ? dl : dc);
    r -= .5f * h * (fabsf(dc) <= fabsf(dr) ? dc : dr);
  }
  left = l;
  right = r;
}

// w8_slopes for a speed axis with the linear speed boundary, without
// branches (its end cases differ from lane to lane): every lane loads
// in-range neighbours (its own value in place of a missing one), forms the
// ENO quantities and selects.  The selected values are the expressions of
// w8_slopes: at pos 0 both slopes are (v[j+s] - f) / h, the uncorrected right
// slope; at n-1 both are (f - v[j-s]) / h; at pos 1 (n-2) dl (dr) is dc.
template <bool FAST, typename I, typename V>
__device__ __forceinline__ void w8_slopes_speed(const V &v, I j, float f, int k, int pos,
                                                const W8 &g, float &left, float &right) {
  const I s = sizeof(I) == 4 ? (I)g.s32[k] : (I)g.stride[k];
  const int n = g.n[k];
  const int c = w8_cls(k);
  const float h = g.h[c], rh = g.rh[c], rhl = g.rhl[c];
  const float fm = pos > 0 ? v(j - s) : f, fp = pos < n - 1 ? v(j + s) : f;
  const float l0 = w8_div<FAST>(f - fm, h, rh, rhl), r0 = w8_div<FAST>(fp - f, h, rh, rhl);
  float l = l0, r = r0;
  if (g.order == 2) {
    const float hh = g.hh[c], rhh = g.rhh[c], rhhl = g.rhhl[c];
    const float fm2 = pos >= 2 ? v(j - 2 * s) : f, fp2 = pos + 2 < n ? v(j + 2 * s) : f;
    const float dc = w8_div<FAST>(fp - 2 * f + fm, hh, rhh, rhhl);
    const float dl = pos < 2 ? dc : w8_div<FAST>(f - 2 * fm + fm2, hh, rhh, rhhl);
    const float dr = pos + 2 >= n ? dc : w8_div<FAST>(fp2 - 2 * fp + f, hh, rhh, rhhl);
    l += .5f * h * (fabsf(dl) <= fabsf(dc) ? dl : dc);
    r -= .5f * h * (fabsf(dc) <= fabsf(dr) ? dc : dr);
  }
  left = pos == 0 ? r0 : pos == n - 1 ? l0 : l;
  right = pos == 0 ? r0 : pos == n - 1 ? l0 : r;
}

// Pursuer-side quantities of a slab (uniform over its nodes).
struct W8Slab {
  int q[4];
  float palo, pahi, fpx, fpy, wp, avoid;
};

__device__ __forceinline__ W8Slab w8_slab(long long s, const W8 &g, const float *__restrict__ x,
                                          const float *__restrict__ cs,
                                          const float *__restrict__ sn,
                                          const float *__restrict__ pv,
                                          const float *__restrict__ p_clearance) {
  W8Slab p;
  p.q[3] = (int)(s % g.n[3]);
  s /= g.n[3];
  p.q[2] = (int)(s % g.n[2]);
  s /= g.n[2];
  p.q[1] = (int)(s % g.n[1]);
  p.q[0] = (int)(s / g.n[1]);
  const float vp = pv[p.q[3]];
  p.palo = p.q[3] == 0 ? 0 : -g.ap;
  p.pahi = p.q[3] == g.n[3] - 1 ? 0 : g.ap;
  p.fpx = vp * cs[p.q[2]];
  p.fpy = vp * sn[p.q[2]];
  p.wp = fabsf(vp) * g.kp;
  p.avoid = p_clearance
                ? -p_clearance[((long long)p.q[0] * g.n[1] + p.q[1]) * g.n[2] + p.q[2]] + 1e-6f
                : w8_mx(fabsf(x[p.q[0]]), fabsf(x[p.q[1]])) - g.L + 1e-6f;
  return p;
}

// Per-launch arguments beside the fields.
struct W8Launch {
  const float *x, *cs, *sn, *pv, *ev;  // tables
  // Footprint-map mode: pursuer clearance [n, n, nh] over (xP, yP, thP); the
  // avoid value is -clearance + 1e-6 (gpu_map_game.py:72-73).  Null: the
  // pursuer wall fault of the square.
  const float *p_clearance;
  float dt;
  double tau;
  // Range flags (see w8_div): the launch that wrote src recorded its id in
  // *src_flag if it wrote a value out of range; this launch records `id` in
  // *dst_flag if it does.
  const unsigned long long *src_flag;
  unsigned long long src_id, *dst_flag, id;
};

// The update of node j = slab * inner + i (native_world.cpp:261-314): axes
// 0-3 read `outer` at j, axes 4-7 read `inner_v` at ii (i, or j when both
// are the global field).  Returns whether the written value is in range.
template <bool FAST, typename I, typename II, typename VO, typename VI>
__device__ __forceinline__ bool w8_node(const VO &outer, const VI &inner_v, I j, II ii,
                                        float center, const W8Slab &p, int i, float *dst,
                                        const float *base, float *__restrict__ arrival,
                                        const W8 &g, const W8Launch &L) {
  int q[8];
  q[0] = p.q[0];
  q[1] = p.q[1];
  q[2] = p.q[2];
  q[3] = p.q[3];
  if (g.fast_decode) {
    const int r6 = (int)__umulhi((unsigned)i, g.magic[7]);
    const int r5 = (int)__umulhi((unsigned)r6, g.magic[6]);
    q[4] = (int)__umulhi((unsigned)r5, g.magic[5]);
    q[7] = i - r6 * g.n[7];
    q[6] = r6 - r5 * g.n[6];
    q[5] = r5 - q[4] * g.n[5];
  } else {
    int r = i;
    q[7] = r % g.n[7];
    r /= g.n[7];
    q[6] = r % g.n[6];
    r /= g.n[6];
    q[5] = r % g.n[5];
    q[4] = r / g.n[5];
  }
  const int eh = q[6], iv = q[7];
  const float ve = L.ev[iv], we = fabsf(ve) * g.ke, ealo = iv == 0 ? 0 : -g.ae,
              eahi = iv == g.n[7] - 1 ? 0 : g.ae;
  const float f[8] = {p.fpx, p.fpy, 0, 0, ve * L.cs[eh], ve * L.sn[eh], 0, 0};
  float gk[8], H = 0, diss = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    float lm, rp;
    if (k < 4)
      w8_slopes<FAST>(outer, j, center, k, q[k], g, k == 2, g.linear && k == 3, lm, rp);
    else if (k == 7 && g.linear)
      w8_slopes_speed<FAST>(inner_v, ii, center, k, q[k], g, lm, rp);
    else
      w8_slopes<FAST>(inner_v, ii, center, k, q[k], g, k == 6, false, lm, rp);
    gk[k] = .5f * (lm + rp);
    const float alpha = k == 2 ? p.wp : k == 3 ? g.ap : k == 6 ? we : k == 7 ? g.ae : fabsf(f[k]);
    H += f[k] * gk[k];
    diss += .5f * alpha * (rp - lm);
  }
  H += -p.wp * fabsf(gk[2]) + we * fabsf(gk[6]) + w8_mn(p.palo * gk[3], p.pahi * gk[3]) +
       w8_mx(ealo * gk[7], eahi * gk[7]);
  float value = center + L.dt * (H + diss);
  if (base) value = .5f * (base[j] + value);
  if (g.order == 1 || base) {
    const float previous = base ? base[j] : center;
    value = w8_mx(w8_mn(value, previous), p.avoid);
    if (previous > 0 && value <= 0 && isnan(arrival[j]))
      arrival[j] = (float)L.tau + L.dt * previous / (previous - value);
  }
  dst[j] = value;
  return w8_in_range(value);
}

// Whether this launch may divide fast: the divisors allow it and the launch
// that wrote src wrote no value out of range.
__device__ __forceinline__ bool w8_fast(const W8 &g, const W8Launch &L) {
  return g.fast_division && *L.src_flag != L.src_id;
}

// One RK stage src -> dst, one thread per node; base = the previous field in
// the blend stage (may alias dst: each node reads base[j] before writing
// dst[j]).  Blocks cover 256 nodes of one slab.  I: index type (int when the
// field is below 4 GiB).
template <typename I>
__global__ __launch_bounds__(256) void world8d_stage(const float *__restrict__ src, float *dst,
                                                     const float *base,
                                                     float *__restrict__ arrival, W8 g,
                                                     W8Launch L) {
  const long long b = blockIdx.x;
  const long long slab = b / g.inner_blocks;  // (xP, yP, thP, vP), uniform
  const int i = (int)(b - slab * g.inner_blocks) * 256 + (int)threadIdx.x;
  if (i >= g.inner) return;
  const W8Slab p = w8_slab(slab, g, L.x, L.cs, L.sn, L.pv, L.p_clearance);
  const I j = (I)(slab * g.inner + i);
  const W8Global<I> v{src};
  const bool ok = w8_fast(g, L)
                      ? w8_node<true>(v, v, j, j, v(j), p, i, dst, base, arrival, g, L)
                      : w8_node<false>(v, v, j, j, v(j), p, i, dst, base, arrival, g, L);
  if (!ok) atomicMax(L.dst_flag, L.id);
}

// The same stage with one block per part of a slab: the block's evader
// states (a range of xE rows of the slab) and the two xE rows on either side
// are staged in LDS, so the evader-axis neighbours are LDS reads and only
// the pursuer-axis neighbours (block-uniform boundary cases) are global
// loads.  A node's xE neighbours, and the rows 0-1 / n-2, n-1 its xE
// extrapolation reads, lie in the staged rows.
template <int THREADS, typename I, bool SPLIT>
__global__ __launch_bounds__(THREADS) void world8d_stage_slab(const float *__restrict__ src,
                                                              float *dst, const float *base,
                                                              float *__restrict__ arrival, W8 g,
                                                              W8Launch L) {
  extern __shared__ float slab_values[];
  const long long slab = SPLIT ? blockIdx.x / g.slab_parts : blockIdx.x;
  const int part = SPLIT ? (int)(blockIdx.x - slab * g.slab_parts) : 0;
  const int row = (int)g.stride[4];  // evader states of one xE row
  const int x0 = SPLIT ? part * g.part_rows : 0;
  const int x1 = SPLIT ? min(g.n[4], x0 + g.part_rows) : g.n[4];
  const int w0 = SPLIT ? max(0, x0 - 2) : 0, w1 = SPLIT ? min(g.n[4], x1 + 2) : g.n[4];
  const int lo = w0 * row, hi = w1 * row;  // staged inner indices
  const float *__restrict__ vals = src + slab * g.inner;
  for (int i = lo + (int)threadIdx.x; i < hi; i += THREADS) slab_values[i - lo] = vals[i];
  __syncthreads();
  const W8Slab p = w8_slab(slab, g, L.x, L.cs, L.sn, L.pv, L.p_clearance);
  const W8Global<I> outer{src};
  // Indexed by inner index: the first staged value is inner index lo.
  const float *staged = slab_values - lo;
  const W8Lds inner_v{staged};
  const bool fast = w8_fast(g, L);
  bool ok = true;
  const int end = x1 * row;
#pragma unroll 1
  for (int i = x0 * row + (int)threadIdx.x; i < end; i += THREADS) {
    const I j = (I)(slab * g.inner + i);
    ok &= fast ? w8_node<true>(outer, inner_v, j, i, staged[i], p, i, dst, base, arrival, g,
                               L)
               : w8_node<false>(outer, inner_v, j, i, staged[i], p, i, dst, base, arrival, g,
                                L);
  }
  if (!ok) atomicMax(L.dst_flag, L.id);
}

// Count of non-finite values (grid-stride; one atomic per non-zero wave).
__global__ __launch_bounds__(256) void world8d_nonfinite(const float *__restrict__ v, long long n,
                                                         unsigned long long *__restrict__ count) {
  unsigned long long bad = 0;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    bad += !isfinite(v[i]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) bad += __shfl_xor(bad, off, 64);
  if ((threadIdx.x & 63) == 0 && bad) atomicAdd(count, bad);
}

// Probe values: one thread per probe, the 256 products summed in order in
// float64 (native_world.cpp:241-248).
__global__ __launch_bounds__(256) void world8d_probe(const float *__restrict__ v,
                                                     const unsigned long long *__restrict__ index,
                                                     const double *__restrict__ weight,
                                                     long long count, double *__restrict__ out) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= count) return;
  double sum = 0;
  for (int k = 0; k < 256; k++) sum += (double)v[index[q * 256 + k]] * weight[q * 256 + k];
  out[q] = sum;
}

// Self-check of w8_div<true> against the IEEE division: every float a of
// either sign with exponent in [lo, hi) (and +0) divided by b.
__global__ __launch_bounds__(256) void world8d_division_check(float b, float y, float ylo, int lo,
                                                              int hi,
                                                              unsigned long long *mismatches) {
  const unsigned long long span = (unsigned long long)(hi - lo) << 23;
  const unsigned long long total = 2 * span + 1;
  unsigned long long bad = 0;
  for (unsigned long long t = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
       t < total; t += (unsigned long long)gridDim.x * blockDim.x) {
    unsigned bits = 0;
    if (t < 2 * span) {
      const unsigned long long u = t % span;
      bits = (unsigned)(((unsigned long long)(lo + 127) << 23) + u) | (t >= span ? 0x80000000u : 0u);
    }
    const float a = __uint_as_float(bits);
    bad += __float_as_uint(w8_div<true>(a, b, y, ylo)) != __float_as_uint(a / b);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) bad += __shfl_xor(bad, off, 64);
  if ((threadIdx.x & 63) == 0 && bad) atomicAdd(mismatches, bad);
}

}  // namespace pe

using namespace pe;

// RN(1/b) for b > 0: the float nearest to 1/b among the double quotient's
// float rounding and its neighbours (b y is exact in double).
static float recip_nearest(float b) {
  const float y0 = (float)(1.0 / (double)b);
  float best = y0;
  double err = std::fabs(1.0 - (double)b * (double)y0);
  for (float c : {std::nextafter(y0, 0.f), std::nextafter(y0, INFINITY)}) {
    const double e = std::fabs(1.0 - (double)b * (double)c);
    if (e < err) {
      err = e;
      best = c;
    }
  }
  return best;
}

// 1/b - y rounded to float (y = RN(1/b)): 1 - b y is exact in double, so
// the remainder carries 1/b to about 2^-48 relative.
static float recip_remainder(float b, float y) {
  return (float)((1.0 - (double)b * (double)y) / (double)b);
}

// Per-class spacing constants (class c takes the spacing of its first axis).
static void world8d_set_spacings(W8 &g, const float spacing[8]) {
  g.fast_division = 1;
  for (int k : {0, 2, 3, 7}) {
    const int c = w8_cls(k);
    g.h[c] = spacing[k];
    g.hh[c] = spacing[k] * spacing[k];
    g.rh[c] = recip_nearest(g.h[c]);
    g.rhh[c] = recip_nearest(g.hh[c]);
    g.rhl[c] = recip_remainder(g.h[c], g.rh[c]);
    g.rhhl[c] = recip_remainder(g.hh[c], g.rhh[c]);
    if (!(g.h[c] >= 0x1p-20f && g.h[c] <= 0x1p20f && g.hh[c] >= 0x1p-20f && g.hh[c] <= 0x1p20f))
      g.fast_division = 0;
  }
}

// Whether a field keeps the stages on the fast division (w8_in_range).
static bool field_in_range(const float *v, long long n) {
  for (long long k = 0; k < n; k++) {
    const float a = std::fabs(v[k]);
    uint32_t bits;
    std::memcpy(&bits, &v[k], 4);
    if (!((a >= 0x1p-60f && a <= 0x1p60f) || bits == 0u)) return false;
  }
  return true;
}

struct pe_world8d {
  int device = 0;
  hipStream_t stream = nullptr;
  pe_world8d_config cfg{};
  W8 g{};
  float *a = nullptr, *b = nullptr, *arrival = nullptr;
  float *tables = nullptr;  // x | cs | sn | pv | ev
  float *p_clearance = nullptr;  // footprint-map mode (pe_world8d_set_clearance)
  unsigned long long *count = nullptr;   // non-finite counter
  unsigned long long *probe_index = nullptr;
  double *probe_weight = nullptr, *probe_out = nullptr;
  long long probes = 0;
  int variant = 0;  // 0: thread per node; else threads per LDS slab block
  size_t lds_bytes = 0;  // LDS of one slab block
  // Range flags of the two field buffers and the id of the launch (or
  // upload) that last wrote each; ids increase.
  unsigned long long *flags = nullptr, *a_flag = nullptr, *b_flag = nullptr;
  unsigned long long a_id = 0, b_id = 0, next_id = 1;
  double maxdt = 0, tau = 0, update_seconds = 0;
  long long steps = 0;
  hipEvent_t t0 = nullptr, t1 = nullptr;
  std::vector<float> x, cs, sn, pv, ev;
};

// Host tables and grid, in the reference's float / double operations
// (native_world.cpp:114-153).
static void world8d_setup(pe_world8d *h) {
  const pe_world8d_config &c = h->cfg;
  const int n = c.positions, nh = c.headings, np = c.p_speeds, ne = c.e_speeds;
  const float extent = (float)c.extent, pmin = (float)c.p_min, pmax = (float)c.p_max,
              emin = (float)c.e_min, emax = (float)c.e_max;
  const double pi = std::acos(-1.);
  const float dx = 2 * extent / (n - 1), dh = float(2 * pi / nh), dvp = (pmax - pmin) / (np - 1),
              dve = (emax - emin) / (ne - 1);
  W8 &g = h->g;
  const int shape[8] = {n, n, nh, np, n, n, nh, ne};
  const float spacing[8] = {dx, dx, dh, dvp, dx, dx, dh, dve};
  g.stride[7] = 1;
  for (int k = 6; k >= 0; k--) g.stride[k] = g.stride[k + 1] * shape[k + 1];
  for (int k = 0; k < 8; k++) {
    g.n[k] = shape[k];
    g.s32[k] = (int)std::min<long long>(g.stride[k], INT32_MAX);
  }
  world8d_set_spacings(g, spacing);
  g.N = g.stride[0] * n;
  g.inner = n * n * nh * ne;
  g.inner_blocks = (g.inner + 255) / 256;
  // floor(2^32 / d) + 1 = 2^32 / d + e with 0 < e <= 1, so
  // floor(r * magic / 2^32) = floor(r / d + r e / 2^32) = floor(r / d)
  // whenever r d < 2^32 (the fractional part of r / d is <= 1 - 1 / d).
  g.fast_decode = 1;
  for (int k = 0; k < 8; k++) {
    g.magic[k] = (unsigned)((1ull << 32) / (unsigned)shape[k] + 1);
    if (k >= 5 && (unsigned long long)g.inner * (unsigned)shape[k] >= (1ull << 32))
      g.fast_decode = 0;
  }
  g.order = c.order;
  g.linear = c.linear_speed_boundary;
  g.L = (float)c.arena;
  g.ap = (float)c.p_accel;
  g.ae = (float)c.e_accel;
  g.kp = (float)c.p_curvature;
  g.ke = (float)c.e_curvature;
  h->x.resize(n);
  h->cs.resize(nh);
  h->sn.resize(nh);
  h->pv.resize(np);
  h->ev.resize(ne);
  for (int i = 0; i < n; i++) h->x[i] = -extent + i * dx;
  for (int i = 0; i < nh; i++) {
    h->cs[i] = (float)std::cos(-pi + i * 2 * pi / nh);
    h->sn[i] = (float)std::sin(-pi + i * 2 * pi / nh);
  }
  for (int i = 0; i < np; i++) h->pv[i] = pmin + i * dvp;
  for (int i = 0; i < ne; i++) h->ev[i] = emin + i * dve;
  const float pabs = std::max(std::fabs(pmin), std::fabs(pmax)),
              eabs = std::max(std::fabs(emin), std::fabs(emax));
  h->maxdt = c.cfl / (2 * double(pabs + eabs) / dx + double(pabs * g.kp + eabs * g.ke) / dh +
                      double(g.ap) / dvp + double(g.ae) / dve);
}

// Initial field max(target, avoid) on the host (native_world.cpp:154-178).
static void world8d_initial(const pe_world8d *h, std::vector<float> &field,
                            std::vector<float> &arrival) {
  const W8 &g = h->g;
  const float L = (float)h->cfg.arena, radius = (float)h->cfg.radius;
  field.resize(g.N);
  arrival.resize(g.N);
  const long long inner = (long long)g.n[6] * g.n[7];
  const float *x = h->x.data();
  for (long long block = 0; block < g.N / inner; block++) {
    const long long begin = block * inner;
    int q[8];
    for (int k = 0; k < 8; k++) q[k] = (int)((begin / g.stride[k]) % g.n[k]);
    const float t1 = std::hypot(x[q[0]] - x[q[4]], x[q[1]] - x[q[5]]) - radius;
    const float t2 = L - std::max(std::fabs(x[q[4]]), std::fabs(x[q[5]]));
    const float target = t2 < t1 ? t2 : t1;
    const float fa = std::fabs(x[q[0]]), fb = std::fabs(x[q[1]]);
    const float avoid = (fa < fb ? fb : fa) - L + 1e-6f;
    const float initial = target < avoid ? avoid : target;
    for (long long k = 0; k < inner; k++) {
      field[begin + k] = initial;
      arrival[begin + k] = initial <= 0 ? 0.f : std::numeric_limits<float>::quiet_NaN();
    }
  }
}

// Range flag of a field just uploaded into h->a.
static int world8d_uploaded(pe_world8d *h, const float *field) {
  h->a_id = h->next_id++;
  const unsigned long long flag = field_in_range(field, h->g.N) ? 0 : h->a_id;
  PE_HIP(hipMemcpy(h->a_flag, &flag, 8, hipMemcpyHostToDevice));
  return PE_OK;
}

extern "C" int pe_world8d_create(const pe_world8d_config *cfg, pe_world8d **out) {
  clear_error();
  if (!cfg || !out) return fail(PE_ERR_INVALID, "null argument");
  *out = nullptr;
  const pe_world8d_config &c = *cfg;
  // native_world.cpp:104-111
  if (c.positions < 5 || c.headings < 4 || c.p_speeds < 3 || c.e_speeds < 3 ||
      !(c.extent > c.arena) || c.order < 1 || c.order > 2 || !(c.cfl > 0) || c.cfl > 1 ||
      !(c.p_min < c.p_max) || !(c.e_min < c.e_max) || !(c.p_accel > 0) || !(c.e_accel > 0) ||
      !(c.p_curvature > 0) || !(c.e_curvature > 0) || !(c.radius >= 0) ||
      (c.linear_speed_boundary != 0 && c.linear_speed_boundary != 1) ||
      (c.flags & ~(PE_WORLD8D_GENERIC | PE_WORLD8D_SLAB_256 | PE_WORLD8D_SLAB_1024)) != 0)
    return fail(PE_ERR_INVALID, "Invalid arguments");
  const double nodes = std::pow((double)c.positions, 4) * c.headings * c.headings *
                       c.p_speeds * c.e_speeds;
  if (nodes > 4.0e9) return fail(PE_ERR_INVALID, "grid too large");
  int dev = 0;
  int rc = require_device(&dev);
  if (rc) return rc;
  pe_world8d *h = new pe_world8d();
  h->device = dev;
  h->cfg = c;
  world8d_setup(h);
  // The LDS kernel stages a range of xE rows of a slab plus two rows on
  // either side: whole slabs up to 36 KiB, else the most rows whose region
  // fits 36 KiB (at least four blocks per CU).  512 threads per block; one
  // thread per node when not even one row fits 64 KiB.
  {
    W8 &g = h->g;
    const int n = g.n[4];
    const size_t row_bytes = sizeof(float) * (size_t)g.stride[4];
    int rows = n;
    if (sizeof(float) * (size_t)g.inner > 36864)
      rows = std::max(1, (int)(36864 / row_bytes) - 4);
    // balance the parts: the fewest parts of at most `rows` rows, evened out
    const int parts = (n + rows - 1) / rows;
    g.part_rows = (n + parts - 1) / parts;
    g.slab_parts = (n + g.part_rows - 1) / g.part_rows;
    const int staged = std::min(n, g.part_rows + 4);
    h->lds_bytes = row_bytes * (size_t)staged;
    if (!(c.flags & PE_WORLD8D_GENERIC) && h->lds_bytes <= 65536)
      h->variant = (c.flags & PE_WORLD8D_SLAB_256)    ? 256
                   : (c.flags & PE_WORLD8D_SLAB_1024) ? 1024
                                                      : 512;
  }
  auto body = [&]() -> int {
    PE_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    PE_HIP(hipEventCreate(&h->t0));
    PE_HIP(hipEventCreate(&h->t1));
    const size_t bytes = sizeof(float) * (size_t)h->g.N;
    if (hipMalloc(&h->a, bytes) != hipSuccess || hipMalloc(&h->b, bytes) != hipSuccess ||
        hipMalloc(&h->arrival, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return fail(PE_ERR_MEMORY, "cannot allocate the 8D fields");
    }
    std::vector<float> tab;
    for (auto *v : {&h->x, &h->cs, &h->sn, &h->pv, &h->ev}) tab.insert(tab.end(), v->begin(), v->end());
    PE_HIP(hipMalloc(&h->tables, sizeof(float) * tab.size()));
    PE_HIP(hipMemcpy(h->tables, tab.data(), sizeof(float) * tab.size(), hipMemcpyHostToDevice));
    PE_HIP(hipMalloc(&h->count, 8));
    PE_HIP(hipMalloc(&h->flags, 16));
    PE_HIP(hipMemset(h->flags, 0, 16));
    h->a_flag = h->flags;
    h->b_flag = h->flags + 1;
    std::vector<float> field, arrival;
    world8d_initial(h, field, arrival);
    PE_HIP(hipMemcpy(h->a, field.data(), bytes, hipMemcpyHostToDevice));
    PE_HIP(hipMemcpy(h->arrival, arrival.data(), bytes, hipMemcpyHostToDevice));
    return world8d_uploaded(h, field.data());
  };
  rc = body();
  if (rc) {
    pe_world8d_destroy(h);
    return rc;
  }
  *out = h;
  return PE_OK;
}

extern "C" int pe_world8d_set_field(pe_world8d *h, const float *field, const float *arrival,
                                    double time) {
  clear_error();
  if (!h || !field || !(time >= 0)) return fail(PE_ERR_INVALID, "invalid argument");
  PE_HIP(hipSetDevice(h->device));
  const long long N = h->g.N;
  PE_HIP(hipStreamSynchronize(h->stream));
  const size_t bytes = sizeof(float) * (size_t)N;
  PE_HIP(hipMemcpy(h->a, field, bytes, hipMemcpyHostToDevice));
  if (arrival) {
    PE_HIP(hipMemcpy(h->arrival, arrival, bytes, hipMemcpyHostToDevice));
  } else {
    // Captured nodes of a continuation arrive at its start (native_world.cpp:185-189).
    std::vector<float> arr((size_t)N);
    for (long long k = 0; k < N; k++)
      arr[k] = field[k] <= 0 ? (float)time : std::numeric_limits<float>::quiet_NaN();
    PE_HIP(hipMemcpy(h->arrival, arr.data(), bytes, hipMemcpyHostToDevice));
  }
  h->tau = time;
  return world8d_uploaded(h, field);
}

// One RK stage src -> dst (buffers a or b with their range flags and ids).
static int world8d_launch(pe_world8d *h, const float *src, float *dst, const float *base,
                          double dt) {
  const W8 &g = h->g;
  W8Launch L;
  L.x = h->tables;
  L.cs = L.x + g.n[0];
  L.sn = L.cs + g.n[2];
  L.pv = L.sn + g.n[2];
  L.ev = L.pv + g.n[3];
  L.p_clearance = h->p_clearance;
  L.dt = (float)dt;
  L.tau = h->tau;
  const bool from_a = src == h->a;
  L.src_flag = from_a ? h->a_flag : h->b_flag;
  L.src_id = from_a ? h->a_id : h->b_id;
  L.dst_flag = from_a ? h->b_flag : h->a_flag;
  L.id = h->next_id++;
  (from_a ? h->b_id : h->a_id) = L.id;
  const unsigned slabs = (unsigned)(g.N / g.inner);
  const size_t lds = h->lds_bytes;
  const unsigned parts = slabs * (unsigned)g.slab_parts;
  // 32-bit byte offsets from the field base while the field is below 4 GiB.
  const bool small = (unsigned long long)g.N * 4 < (1ull << 32);
  if (h->variant == 0) {
    if (small)
      hipLaunchKernelGGL(world8d_stage<int>, dim3(slabs * g.inner_blocks), dim3(256), 0,
                         h->stream, src, dst, base, h->arrival, g, L);
    else
      hipLaunchKernelGGL(world8d_stage<long long>, dim3(slabs * g.inner_blocks), dim3(256), 0,
                         h->stream, src, dst, base, h->arrival, g, L);
    PE_LAUNCHED(world8d_stage);
    return PE_OK;
  }
#define PE_W8_SLAB_S(T, S)                                                                 \
  if (small)                                                                               \
    hipLaunchKernelGGL((world8d_stage_slab<T, int, S>), dim3(parts), dim3(T), lds,         \
                       h->stream, src, dst, base, h->arrival, g, L);                       \
  else                                                                                     \
    hipLaunchKernelGGL((world8d_stage_slab<T, long long, S>), dim3(parts), dim3(T), lds,   \
                       h->stream, src, dst, base, h->arrival, g, L);
#define PE_W8_SLAB(T)              \
  if (g.slab_parts > 1) {          \
    PE_W8_SLAB_S(T, true)          \
  } else {                         \
    PE_W8_SLAB_S(T, false)         \
  }
  if (h->variant == 256) {
    PE_W8_SLAB(256)
  } else if (h->variant == 512) {
    PE_W8_SLAB(512)
  } else {
    PE_W8_SLAB(1024)
  }
#undef PE_W8_SLAB
#undef PE_W8_SLAB_S
  PE_LAUNCHED(world8d_stage_slab);
  return PE_OK;
}

// One RK1 or RK2 step of size dt from the current time.
static int world8d_step(pe_world8d *h, double dt) {
  int rc;
  if (h->cfg.order == 1) {
    if ((rc = world8d_launch(h, h->a, h->b, nullptr, dt))) return rc;
    std::swap(h->a, h->b);
    std::swap(h->a_flag, h->b_flag);
    std::swap(h->a_id, h->b_id);
  } else {
    if ((rc = world8d_launch(h, h->a, h->b, nullptr, dt))) return rc;
    if ((rc = world8d_launch(h, h->b, h->a, h->a, dt))) return rc;
  }
  h->tau += dt;
  h->steps++;
  return PE_OK;
}

extern "C" int pe_world8d_advance_to(pe_world8d *h, double goal, int64_t max_steps) {
  clear_error();
  if (!h || !std::isfinite(goal)) return fail(PE_ERR_INVALID, "invalid argument");
  PE_HIP(hipSetDevice(h->device));
  PE_HIP(hipEventRecord(h->t0, h->stream));
  int64_t taken = 0;
  // native_world.cpp:324-329
  while (h->tau < goal - 1e-10 && (max_steps <= 0 || taken < max_steps)) {
    const double dt = std::min(h->maxdt, goal - h->tau);
    int rc = world8d_step(h, dt);
    if (rc) return rc;
    taken++;
  }
  PE_HIP(hipEventRecord(h->t1, h->stream));
  PE_HIP(hipEventSynchronize(h->t1));
  float ms = 0;
  PE_HIP(hipEventElapsedTime(&ms, h->t0, h->t1));
  h->update_seconds += ms * 1e-3;
  return PE_OK;
}

extern "C" int pe_world8d_count_nonfinite(pe_world8d *h, int64_t *count) {
  clear_error();
  if (!h || !count) return fail(PE_ERR_INVALID, "invalid argument");
  PE_HIP(hipSetDevice(h->device));
  PE_HIP(hipMemsetAsync(h->count, 0, 8, h->stream));
  hipLaunchKernelGGL(world8d_nonfinite, dim3(2048), dim3(256), 0, h->stream, h->a, h->g.N,
                     h->count);
  PE_LAUNCHED(world8d_nonfinite);
  unsigned long long bad = 0;
  PE_HIP(hipMemcpyAsync(&bad, h->count, 8, hipMemcpyDeviceToHost, h->stream));
  PE_HIP(hipStreamSynchronize(h->stream));
  *count = (int64_t)bad;
  return PE_OK;
}

extern "C" int pe_world8d_set_probes(pe_world8d *h, const uint64_t *indices,
                                     const double *weights, int64_t count) {
  clear_error();
  if (!h || count < 0 || (count > 0 && (!indices || !weights)))
    return fail(PE_ERR_INVALID, "invalid argument");
  // native_world.cpp:215-226
  for (int64_t q = 0; q < count; q++) {
    double sum = 0;
    for (int k = 0; k < 256; k++) {
      const int64_t j = q * 256 + k;
      if (indices[j] >= (uint64_t)h->g.N || !std::isfinite(weights[j]) || weights[j] < 0)
        return fail(PE_ERR_INVALID, "Bad probe stencil");
      sum += weights[j];
    }
    if (std::fabs(sum - 1) > 1e-8) return fail(PE_ERR_INVALID, "Bad probe weights");
  }
  PE_HIP(hipSetDevice(h->device));
  PE_HIP(hipStreamSynchronize(h->stream));
  for (void *p : {(void *)h->probe_index, (void *)h->probe_weight, (void *)h->probe_out})
    if (p) PE_HIP(hipFree(p));
  h->probe_index = nullptr;
  h->probe_weight = h->probe_out = nullptr;
  h->probes = 0;
  if (count == 0) return PE_OK;
  const size_t terms = (size_t)count * 256;
  PE_HIP(hipMalloc(&h->probe_index, terms * 8));
  PE_HIP(hipMalloc(&h->probe_weight, terms * 8));
  PE_HIP(hipMalloc(&h->probe_out, (size_t)count * 8));
  PE_HIP(hipMemcpy(h->probe_index, indices, terms * 8, hipMemcpyHostToDevice));
  PE_HIP(hipMemcpy(h->probe_weight, weights, terms * 8, hipMemcpyHostToDevice));
  h->probes = count;
  return PE_OK;
}

extern "C" int pe_world8d_probe(pe_world8d *h, double *values) {
  clear_error();
  if (!h || (h->probes && !values)) return fail(PE_ERR_INVALID, "invalid argument");
  if (!h->probes) return PE_OK;
  PE_HIP(hipSetDevice(h->device));
  hipLaunchKernelGGL(world8d_probe, dim3((unsigned)((h->probes + 255) / 256)), dim3(256), 0,
                     h->stream, h->a, h->probe_index, h->probe_weight, h->probes, h->probe_out);
  PE_LAUNCHED(world8d_probe);
  PE_HIP(hipMemcpyAsync(values, h->probe_out, (size_t)h->probes * 8, hipMemcpyDeviceToHost,
                        h->stream));
  PE_HIP(hipStreamSynchronize(h->stream));
  return PE_OK;
}

extern "C" int pe_world8d_read(pe_world8d *h, float *field, float *arrival) {
  clear_error();
  if (!h) return fail(PE_ERR_INVALID, "null handle");
  PE_HIP(hipSetDevice(h->device));
  PE_HIP(hipStreamSynchronize(h->stream));
  const size_t bytes = sizeof(float) * (size_t)h->g.N;
  if (field) PE_HIP(hipMemcpy(field, h->a, bytes, hipMemcpyDeviceToHost));
  if (arrival) PE_HIP(hipMemcpy(arrival, h->arrival, bytes, hipMemcpyDeviceToHost));
  return PE_OK;
}

extern "C" int pe_world8d_set_clearance(pe_world8d *h, const float *p_clearance,
                                        const float *e_clearance) {
  clear_error();
  if (!h || !p_clearance || !e_clearance) return fail(PE_ERR_INVALID, "invalid argument");
  if (!h->cfg.linear_speed_boundary)
    return fail(PE_ERR_INVALID, "the footprint-map game uses the linear speed boundary");
  W8 &g = h->g;
  const int n = g.n[0], nh = g.n[2];
  const size_t poses = (size_t)n * n * nh;
  for (size_t k = 0; k < poses; k++)
    if (!std::isfinite(p_clearance[k]) || !std::isfinite(e_clearance[k]))
      return fail(PE_ERR_INVALID, "clearance grids must be finite");
  PE_HIP(hipSetDevice(h->device));
  PE_HIP(hipStreamSynchronize(h->stream));
  // Heading spacing as the map kernel forms it: float 2 pi / nh
  // (gpu_map_game.py:58).
  const float dh = 6.2831853071795864769f / nh;
  const float spacing[8] = {g.h[0], g.h[0], dh, g.h[2], g.h[0], g.h[0], dh, g.h[3]};
  world8d_set_spacings(g, spacing);
  if (!h->p_clearance) PE_HIP(hipMalloc(&h->p_clearance, sizeof(float) * poses));
  PE_HIP(hipMemcpy(h->p_clearance, p_clearance, sizeof(float) * poses, hipMemcpyHostToDevice));
  // Initial field max(min(capture, e), -p + 1e-6) (gpu_map_game.py:43-49),
  // capture = hypot((iP - iE) dx, (jP - jE) dx) - radius.
  const float extent = (float)h->cfg.extent, radius = (float)h->cfg.radius;
  const float dx = 2 * extent / (n - 1);
  std::vector<float> field((size_t)g.N), arrival((size_t)g.N);
  const long long inner = (long long)g.n[6] * g.n[7];
  for (long long block = 0; block < g.N / inner; block++) {
    const long long begin = block * inner;
    int q[8];
    for (int k = 0; k < 6; k++) q[k] = (int)((begin / g.stride[k]) % g.n[k]);
    const float cap = std::hypot((float)(q[0] - q[4]) * dx, (float)(q[1] - q[5]) * dx) - radius;
    const float pc = p_clearance[((size_t)q[0] * n + q[1]) * nh + q[2]];
    for (int th = 0; th < g.n[6]; th++) {
      const float ec = e_clearance[((size_t)q[4] * n + q[5]) * nh + th];
      const float v = std::fmax(std::fmin(cap, ec), -pc + 1e-6f);
      for (int iv = 0; iv < g.n[7]; iv++) {
        field[begin + (long long)th * g.n[7] + iv] = v;
        arrival[begin + (long long)th * g.n[7] + iv] =
            v <= 0 ? 0.f : std::numeric_limits<float>::quiet_NaN();
      }
    }
  }
  const size_t bytes = sizeof(float) * (size_t)g.N;
  PE_HIP(hipMemcpy(h->a, field.data(), bytes, hipMemcpyHostToDevice));
  PE_HIP(hipMemcpy(h->arrival, arrival.data(), bytes, hipMemcpyHostToDevice));
  h->tau = 0;
  return world8d_uploaded(h, field.data());
}

extern "C" int pe_world8d_set_step_limit(pe_world8d *h, double max_dt) {
  clear_error();
  if (!h || !(max_dt > 0) || !std::isfinite(max_dt))
    return fail(PE_ERR_INVALID, "invalid argument");
  h->maxdt = max_dt;
  return PE_OK;
}

extern "C" int pe_world8d_check_division(const float *divisors, int32_t count, int32_t lo,
                                         int32_t hi, int64_t *mismatches) {
  clear_error();
  if (!divisors || count < 0 || !mismatches || lo < -126 || hi > 128 || lo >= hi)
    return fail(PE_ERR_INVALID, "invalid argument");
  for (int32_t k = 0; k < count; k++)
    if (!(divisors[k] >= 0x1p-20f && divisors[k] <= 0x1p20f))
      return fail(PE_ERR_INVALID, "divisor outside [2^-20, 2^20]");
  int dev = 0;
  int rc = require_device(&dev);
  if (rc) return rc;
  unsigned long long *d = nullptr;
  PE_HIP(hipMalloc(&d, 8));
  for (int32_t k = 0; k < count; k++) {
    unsigned long long bad = 0;
    if (hipMemset(d, 0, 8) != hipSuccess) break;
    const float y = recip_nearest(divisors[k]);
    hipLaunchKernelGGL(world8d_division_check, dim3(4096), dim3(256), 0, 0, divisors[k], y,
                       recip_remainder(divisors[k], y), (int)lo, (int)hi, d);
    if (hipMemcpy(&bad, d, 8, hipMemcpyDeviceToHost) != hipSuccess) break;
    mismatches[k] = (int64_t)bad;
  }
  const hipError_t e = hipGetLastError();
  (void)hipFree(d);
  if (e != hipSuccess) return fail(PE_ERR_HIP, hipGetErrorString(e));
  return PE_OK;
}

extern "C" int pe_world8d_get_stats(pe_world8d *h, pe_world8d_stats *s) {
  clear_error();
  if (!h || !s) return fail(PE_ERR_INVALID, "invalid argument");
  std::memset(s, 0, sizeof(*s));
  s->states = h->g.N;
  s->internal_steps = h->steps;
  s->time = h->tau;
  s->max_internal_dt = h->maxdt;
  s->update_seconds = h->update_seconds;
  s->probes = h->probes;
  s->kernel = h->variant;
  s->fast_division = h->g.fast_division;
  return PE_OK;
}

extern "C" void pe_world8d_destroy(pe_world8d *h) {
  if (!h) return;
  hipSetDevice(h->device);
  if (h->stream) hipStreamSynchronize(h->stream);
  for (void *p : {(void *)h->a, (void *)h->b, (void *)h->arrival, (void *)h->tables,
                  (void *)h->count, (void *)h->probe_index, (void *)h->probe_weight,
                  (void *)h->probe_out, (void *)h->flags, (void *)h->p_clearance})
    if (p) hipFree(p);
  if (h->t0) hipEventDestroy(h->t0);
  if (h->t1) hipEventDestroy(h->t1);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
}
