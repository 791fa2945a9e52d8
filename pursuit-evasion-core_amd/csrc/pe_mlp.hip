// Learned value surrogate: a small MLP trained and evaluated with
// hand-written BF16 MFMA kernels (FP32 accumulation) on gfx950.
//
// Network (input -> 4 x 256 SiLU -> 1):
//   features f(x) = [xn (8), sin(B xn) (16), cos(B xn) (16), extra (3),
//                    E offset in P's frame (2), cos/sin relative heading (2),
//                    distance (1), 0 pad to a multiple of 16]
//   with xn the normalised 8D pair state [xP, yP, thP, vP, xE, yE, thE, vE],
//   B a fixed Gaussian matrix (random Fourier features) and `extra` the
//   per-node map features (contact payoff phi, P clearance, E clearance);
//   h1 = silu(W0 f + b0), h_{l+1} = silu(W_l h_l + b_l), y = w4 . h4 + b4.
// Weights are kept as FP32 masters (AdamW) with BF16 copies [out, in] and
// transposed BF16 copies [in, out] for the MFMA operands.
//
// GEMM: C[m, n] = sum_k A(m, k) B(k, n) with arbitrary strides, one
// BM x BN output tile per 256-thread block (128 x 128, 128 x 64 or 64 x 64:
// the largest that still gives 256 blocks); the four waves split it 2 x 2
// and each owns its quarter as 16 x 16 tiles of v_mfma_f32_16x16x32_bf16.
// K advances 64 at a time through double-buffered LDS: the next stage's
// global loads are issued before the current stage's MFMAs and written to
// the other buffer after them, one barrier per stage.  Both operands sit
// k-contiguous in LDS, so each lane's MFMA operand (8 consecutive k) is one
// 16-byte LDS read.  The epilogues (bias + SiLU, SiLU derivative, plain or
// accumulated store) are fused and store 4 consecutive outputs per lane.
// Weight gradients are split over k (at most 64 chunks of at least 256
// rows of the batch, each storing a partial tile) and the partials summed in
// a fixed order, so the training step is deterministic.
//
// MFMA 16x16x32 bf16 operand layout (wave64), lane l:
//   A fragment: row (l % 16), k = 8 * (l / 16) .. +7
//   B fragment: col (l % 16), k = 8 * (l / 16) .. +7
//   D fragment: col (l % 16), rows 4 * (l / 16) .. +3
// The kernel passes the B fragment as the first MFMA operand, which
// transposes D: lane l then holds row (l % 16), cols 4 * (l / 16) .. +3.
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "pe_internal.h"


namespace pe {

constexpr int kRelative = 5;  // relative-frame features appended by mlp_features
constexpr int kMaxExtra = 40;  // map features per point after the 8 state coordinates

__device__ __forceinline__ __bf16 to_bf16(float x) { return (__bf16)x; }
__device__ __forceinline__ float silu(float z) { return z / (1.f + __expf(-z)); }
__device__ __forceinline__ float silu_grad(float z) {
  const float s = 1.f / (1.f + __expf(-z));
  return s * (1.f + z * (1.f - s));
}

enum Epilogue : int {
  kStoreF32 = 0,     // C (float) = acc (+ beta * C)
  kBiasSilu = 1,     // Z = acc + bias (float); H = silu(Z) (bf16)
  kSiluGrad = 2,     // G (bf16) = acc * silu'(Z[m, n])
};

struct GemmArgs {
  const __bf16 *A;
  long long sam, sak;  // A(m, k) = A[m * sam + k * sak]
  const __bf16 *B;
  long long sbk, sbn;  // B(k, n) = B[k * sbk + n * sbn]
  int M, N, K;
  int epilogue;
  float beta;          // kStoreF32: C = acc + beta * C
  float *C;            // [M, N] float (kStoreF32 output)
  long long ldc;
  const float *bias;   // kBiasSilu
  __bf16 *H;           // kBiasSilu / kSiluGrad bf16 output [M, N]
  long long ldh;
  __bf16 *Z;           // pre-activations [M, N] in bf16: kBiasSilu writes
                       // them (if not null), kSiluGrad reads them
  long long ldz;
  __bf16 *HT;          // kBiasSilu / kSiluGrad: the bf16 output also stored
                       // transposed, HT[n * ldht + m] (if not null)
  long long ldht;
  int kchunk;          // > 0: block z sums k in [z * kchunk, (z + 1) * kchunk)
                       // and stores (kStoreF32, beta 0) at C + z * cstride
  long long cstride;
};

// One operand tile of ROWS rows x kK k staged through registers: each
// thread loads ROWS / 32 vectors of 8 bf16 from global memory (issued before
// the current tile's MFMAs) and writes them k-contiguous into LDS (after).
// X(row, k) = X[row * srow + k * sk].  k-contiguous operands load 8
// consecutive k of one row; row-contiguous ones 8 consecutive rows of one k,
// transposed on the way into LDS.  Out-of-range rows and k read as zero.
constexpr int kK = 64;      // K per LDS stage (two MFMA k-steps)
constexpr int kLdsPad = 8;  // LDS row padding (bf16) against bank conflicts

template <int ROWS>
struct TileLoad {
  static constexpr int V = ROWS / 32;
  bf16x8 v[V];

  __device__ __forceinline__ void load(const __bf16 *__restrict__ X, long long srow,
                                       long long sk, int row0, int rows, int k0, int kend,
                                       int tid) {
#pragma unroll
    for (int i = 0; i < V; i++) {
      const int q = tid + 256 * i;
      if (sk == 1) {
        const int row = row0 + q / (kK / 8), k = k0 + (q % (kK / 8)) * 8;
        if (row < rows && k + 8 <= kend && (srow % 8) == 0) {
          v[i] = *reinterpret_cast<const bf16x8 *>(X + row * srow + k);
        } else {
#pragma unroll
          for (int t = 0; t < 8; t++)
            v[i][t] = (row < rows && k + t < kend) ? X[row * srow + k + t] : to_bf16(0.f);
        }
      } else if (srow == 1) {
        const int k = k0 + q / (ROWS / 8), row = row0 + (q % (ROWS / 8)) * 8;
        if (k < kend && row + 8 <= rows && (sk % 8) == 0) {
          v[i] = *reinterpret_cast<const bf16x8 *>(X + k * sk + row);
        } else {
#pragma unroll
          for (int t = 0; t < 8; t++)
            v[i][t] = (k < kend && row + t < rows) ? X[k * sk + row + t] : to_bf16(0.f);
        }
      } else {
        const int row = row0 + q / (kK / 8), k = k0 + (q % (kK / 8)) * 8;
#pragma unroll
        for (int t = 0; t < 8; t++)
          v[i][t] = (row < rows && k + t < kend) ? X[row * srow + (k + t) * sk] : to_bf16(0.f);
      }
    }
  }

  __device__ __forceinline__ void store(__bf16 (*T)[kK + kLdsPad], long long srow,
                                        long long sk, int tid) const {
#pragma unroll
    for (int i = 0; i < V; i++) {
      const int q = tid + 256 * i;
      if (sk != 1 && srow == 1) {
        const int k = q / (ROWS / 8), r = (q % (ROWS / 8)) * 8;
#pragma unroll
        for (int t = 0; t < 8; t++) T[r + t][k] = v[i][t];
      } else {
        const int r = q / (kK / 8), c = (q % (kK / 8)) * 8;
        *reinterpret_cast<bf16x8 *>(&T[r][c]) = v[i];
      }
    }
  }
};

// C[m, n] = sum_k A(m, k) B(k, n), one BM x BN output tile per 256-thread
// block; the four waves split it 2 x 2 and each owns (BM/2) x (BN/2) as
// 16 x 16 tiles of v_mfma_f32_16x16x32_bf16.  K advances kK = 64 per stage
// through double-buffered LDS (one barrier per stage); both operands sit
// k-contiguous in LDS so each lane's MFMA operand (8 consecutive k) is one
// 16-byte LDS read.
template <int BM, int BN>
__global__ __launch_bounds__(256) void mfma_gemm(GemmArgs g) {
  constexpr int TM = BM / 32, TN = BN / 32;  // MFMA tiles per wave
  __shared__ __attribute__((aligned(16))) __bf16 As[2][BM][kK + kLdsPad];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[2][BN][kK + kLdsPad];
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * (BM / 2), wn = (wave & 1) * (BN / 2);
  int k_begin = 0, k_end = g.K;
  float *C = g.C;
  if (g.kchunk > 0) {
    k_begin = blockIdx.z * g.kchunk;
    k_end = min(g.K, k_begin + g.kchunk);
    C += blockIdx.z * g.cstride;
  }
  f32x4 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  TileLoad<BM> la;
  TileLoad<BN> lb;
  la.load(g.A, g.sam, g.sak, m0, g.M, k_begin, k_end, tid);
  lb.load(g.B, g.sbn, g.sbk, n0, g.N, k_begin, k_end, tid);
  la.store(As[0], g.sam, g.sak, tid);
  lb.store(Bs[0], g.sbn, g.sbk, tid);
  __syncthreads();
  int buf = 0;
  const int fr = lane & 15, fk = (lane >> 4) * 8;
  for (int k0 = k_begin; k0 < k_end; k0 += kK) {
    const bool more = k0 + kK < k_end;
    if (more) {
      la.load(g.A, g.sam, g.sak, m0, g.M, k0 + kK, k_end, tid);
      lb.load(g.B, g.sbn, g.sbk, n0, g.N, k0 + kK, k_end, tid);
    }
#pragma unroll
    for (int kk = 0; kk < kK; kk += 32) {
      bf16x8 a[TM], b[TN];
#pragma unroll
      for (int i = 0; i < TM; i++)
        a[i] = *reinterpret_cast<const bf16x8 *>(&As[buf][wm + 16 * i + fr][kk + fk]);
#pragma unroll
      for (int j = 0; j < TN; j++)
        b[j] = *reinterpret_cast<const bf16x8 *>(&Bs[buf][wn + 16 * j + fr][kk + fk]);
      // B first: the tile comes out transposed, so each lane holds four
      // consecutive n of one m (vector stores in the epilogue).
#pragma unroll
      for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[j], a[i], acc[i][j], 0, 0, 0);
    }
    if (more) {
      la.store(As[buf ^ 1], g.sam, g.sak, tid);
      lb.store(Bs[buf ^ 1], g.sbn, g.sbk, tid);
    }
    __syncthreads();
    buf ^= 1;
  }
  // The transposed output goes through LDS, which the K loop has finished
  // with (its last barrier): each wave's (BN/2) x (BM/2) block of HT is
  // staged m-contiguous in its own region, then stored 16 bytes per lane.
  constexpr int kTPad = BM / 2 + 8;
  static_assert(4 * (BN / 2) * kTPad <= 2 * BM * (kK + kLdsPad), "staging fits in As");
  __bf16(*Tt)[kTPad] = reinterpret_cast<__bf16(*)[kTPad]>(&As[0][0][0]) + wave * (BN / 2);
  // Epilogue: lane holds C[m][n .. n + 3] of each tile, m = lane % 16 and
  // n = 4 * (lane / 16) within the tile.
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++) {
      const int m = m0 + wm + 16 * i + (lane & 15);
      const int n = n0 + wn + 16 * j + 4 * (lane >> 4);
      if (m >= g.M) continue;
      const f32x4 x = acc[i][j];
      const bool whole = n + 3 < g.N;
      if (g.epilogue == kStoreF32) {
        float *c = C + m * g.ldc + n;
        if (whole && (g.ldc % 4) == 0) {
          f32x4 y = x;
          if (g.beta != 0.f) y += g.beta * *reinterpret_cast<const f32x4 *>(c);
          *reinterpret_cast<f32x4 *>(c) = y;
        } else {
#pragma unroll
          for (int v = 0; v < 4 && n + v < g.N; v++)
            c[v] = g.beta != 0.f ? x[v] + g.beta * c[v] : x[v];
        }
      } else {
        // bias + SiLU (Z kept if asked) or the SiLU derivative (Z read)
        const bool vec = whole && (g.ldz % 4) == 0 && (g.ldh % 4) == 0;
        bf16x4 hv, zv;
        if (g.epilogue == kBiasSilu) {
#pragma unroll
          for (int v = 0; v < 4; v++) {
            const float z = n + v < g.N ? x[v] + g.bias[n + v] : 0.f;
            hv[v] = to_bf16(silu(z));
            zv[v] = to_bf16(z);
          }
          if (g.Z) {
            if (vec) {
              *reinterpret_cast<bf16x4 *>(g.Z + m * g.ldz + n) = zv;
            } else {
#pragma unroll
              for (int v = 0; v < 4; v++)
                if (n + v < g.N) g.Z[m * g.ldz + n + v] = zv[v];
            }
          }
        } else {
          if (vec) {
            zv = *reinterpret_cast<const bf16x4 *>(g.Z + m * g.ldz + n);
          } else {
#pragma unroll
            for (int v = 0; v < 4; v++)
              zv[v] = n + v < g.N ? g.Z[m * g.ldz + n + v] : to_bf16(0.f);
          }
#pragma unroll
          for (int v = 0; v < 4; v++) hv[v] = to_bf16(x[v] * silu_grad((float)zv[v]));
        }
        if (vec) {
          *reinterpret_cast<bf16x4 *>(g.H + m * g.ldh + n) = hv;
        } else {
#pragma unroll
          for (int v = 0; v < 4; v++)
            if (n + v < g.N) g.H[m * g.ldh + n + v] = hv[v];
        }
        if (g.HT) {
#pragma unroll
          for (int v = 0; v < 4; v++)
            Tt[16 * j + 4 * (lane >> 4) + v][16 * i + (lane & 15)] = hv[v];
        }
      }
    }
  if (g.HT && g.epilogue != kStoreF32) {  // uniform over the block
    __syncthreads();
    // 8 lanes per row of BM/2 values (8 each), 64 / (BM/16) rows per pass
    constexpr int kLanesPerRow = BM / 16, kRowsPerPass = 64 / kLanesPerRow;
    const int r0 = lane / kLanesPerRow, c = (lane % kLanesPerRow) * 8;
    const int m = m0 + wm + c;
#pragma unroll
    for (int r = r0; r < BN / 2; r += kRowsPerPass) {
      const int n = n0 + wn + r;
      if (n >= g.N) continue;
      __bf16 *dst = g.HT + (long long)n * g.ldht + m;
      if (m + 8 <= g.M && (g.ldht % 8) == 0) {
        *reinterpret_cast<bf16x8 *>(dst) = *reinterpret_cast<const bf16x8 *>(&Tt[r][c]);
      } else {
#pragma unroll
        for (int t = 0; t < 8; t++)
          if (m + t < g.M) dst[t] = Tt[r][c + t];
      }
    }
  }
}

// ---------------------------------------------------------------- features



// xn = (x - shift) * scale for the 8 state coordinates; features into F
// and, if FT is not null, transposed into FT[c * ldt + i].
__global__ __launch_bounds__(256) void mlp_features(
    const float *__restrict__ raw, long long n, MlpShape s,
    const float *__restrict__ shift, const float *__restrict__ scale,
    const float *__restrict__ proj /* [rff/2, 8] */, __bf16 *__restrict__ F,
    __bf16 *__restrict__ FT, long long ldt) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float *x = raw + i * s.raw;
  float xn[8];
  for (int d = 0; d < 8; d++) xn[d] = (x[d] - shift[d]) * scale[d];
  __bf16 *f = F + i * s.width;
  int c = 0;
  for (int d = 0; d < 8; d++) f[c++] = to_bf16(xn[d]);
  const int half = s.rff / 2;
  for (int q = 0; q < half; q++) {
    float t = 0.f;
    for (int d = 0; d < 8; d++) t = fmaf(proj[q * 8 + d], xn[d], t);
    f[c + q] = to_bf16(__sinf(t));
    f[c + half + q] = to_bf16(__cosf(t));
  }
  c += s.rff;
  for (int d = 0; d < s.extra; d++) f[c++] = to_bf16(x[8 + d]);
  // Relative frame: E's offset in P's frame, relative heading, distance.
  const float cp = __cosf(x[2]), sp = __sinf(x[2]);
  const float dx = x[4] - x[0], dy = x[5] - x[1];
  f[c++] = to_bf16(cp * dx + sp * dy);
  f[c++] = to_bf16(-sp * dx + cp * dy);
  f[c++] = to_bf16(__cosf(x[6] - x[2]));
  f[c++] = to_bf16(__sinf(x[6] - x[2]));
  f[c++] = to_bf16(sqrtf(dx * dx + dy * dy));
  for (; c < s.width; c++) f[c] = to_bf16(0.f);
  if (FT)  // consecutive threads write consecutive i of each feature row
    for (c = 0; c < s.width; c++) FT[c * ldt + i] = f[c];
}

// y[i] = w . h[i] + b  (final 256 -> 1 layer); one wave per point.
__global__ __launch_bounds__(256) void mlp_head(const __bf16 *__restrict__ H,
                                                long long n, int width,
                                                const float *__restrict__ w,
                                                const float *__restrict__ b,
                                                float *__restrict__ y) {
  const long long i = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (i >= n) return;
  float t = 0.f;
  for (int k = lane; k < width; k += 64) t = fmaf(w[k], (float)H[i * width + k], t);
  for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
  if (lane == 0) y[i] = t + b[0];
}

// Weighted MSE on the head output: weight 1 + 4 exp(-|y| / 0.2) (the
// reference's training weighting, local_game/training.py:351-352).
// Underestimates (yhat < y) are weighted `under` times more, which biases
// the fit towards an upper envelope of the values.
// dy = 2 w (yhat - y) / n; loss accumulated (sum of w (yhat - y)^2 / n).
__global__ __launch_bounds__(256) void mlp_loss(const float *__restrict__ yhat,
                                                const float *__restrict__ y,
                                                long long n, float under,
                                                float *__restrict__ dy,
                                                float *__restrict__ loss) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (i < n) {
    const float r = yhat[i] - y[i];
    const float w = (1.f + 4.f * __expf(-fabsf(y[i]) / 0.2f)) * (r < 0.f ? under : 1.f);
    dy[i] = 2.f * w * r / (float)n;
    l = w * r * r / (float)n;
  }
  for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off, 64);
  if ((threadIdx.x & 63) == 0 && l != 0.f) atomicAdd(loss, l);
}

// Head backward: dZ3[i, k] = dy_i w4[k] silu'(Z3[i, k]) (bf16), stored row
// by row into dZ and transposed into dZT[k * ldt + i].  One 256-thread block
// per 64 x 64 tile: rows of the tile are written from registers, the
// transposed rows through a padded LDS tile, both coalesced.  The head's
// weight and bias gradients are column sums (column_partials with weights
// dy).
__global__ __launch_bounds__(256) void mlp_head_back(const __bf16 *__restrict__ Z, long long n,
                                                     int width, const float *__restrict__ w,
                                                     const float *__restrict__ dy,
                                                     __bf16 *__restrict__ dZ,
                                                     __bf16 *__restrict__ dZT, long long ldt) {
  __shared__ __bf16 tile[64][65];
  const long long i0 = (long long)blockIdx.x * 64;
  const int k0 = blockIdx.y * 64;
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int r = e >> 6, c = e & 63;
    const long long i = i0 + r;
    const int k = k0 + c;
    __bf16 v = to_bf16(0.f);
    if (i < n && k < width) {
      v = to_bf16(dy[i] * w[k] * silu_grad((float)Z[i * width + k]));
      dZ[i * width + k] = v;
    }
    tile[c][r] = v;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < 64 * 64; e += 256) {
    const int c = e >> 6, r = e & 63;
    const long long i = i0 + r;
    const int k = k0 + c;
    if (i < n && k < width) dZT[k * ldt + i] = tile[c][r];
  }
}

// Column sums of a row-major [n, width] matrix over chunks of `rows` rows:
// partial[c, k] = sum over rows i of chunk c of a_i X[i, k] (a_i = wt[i], or
// 1 without weights); with X null, partial[c, 0] = sum of wt over the
// chunk.  Threads run along k, so every row read is coalesced.
__global__ __launch_bounds__(256) void column_partials(const __bf16 *__restrict__ X,
                                                       const float *__restrict__ wt,
                                                       long long n, int width, int rows,
                                                       float *__restrict__ partial) {
  const long long r0 = (long long)blockIdx.x * rows;
  const long long r1 = min(n, r0 + rows);
  if (!X) {
    if (threadIdx.x == 0) {
      float t = 0.f;
      for (long long i = r0; i < r1; i++) t += wt[i];
      partial[blockIdx.x] = t;
    }
    return;
  }
  for (int k = threadIdx.x; k < width; k += blockDim.x) {
    float t = 0.f;
#pragma unroll 16
    for (long long i = r0; i < r1; i++)
      t = fmaf(wt ? wt[i] : 1.f, (float)X[i * width + k], t);
    partial[(long long)blockIdx.x * width + k] = t;
  }
}

// out[j] = sum over c of partial[c * count + j], in c order: one thread per
// output, reads coalesced across threads (many outputs, few chunks).
__global__ __launch_bounds__(256) void sum_partials(const float *__restrict__ partial,
                                                    int chunks, long long count,
                                                    float *__restrict__ out) {
  const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= count) return;
  float t = 0.f;
#pragma unroll 8
  for (int c = 0; c < chunks; c++) t += partial[(long long)c * count + j];
  out[j] = t;
}

// The same sum with one wave per output: lane l adds chunks l, l + 64, ...
// and the wave combines its lanes in a fixed butterfly order (few outputs,
// many chunks).
__global__ __launch_bounds__(256) void sum_partials_wide(const float *__restrict__ partial,
                                                         int chunks, long long count,
                                                         float *__restrict__ out) {
  const long long j = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= count) return;  // uniform per wave
  const int lane = threadIdx.x & 63;
  float t = 0.f;
  for (int c = lane; c < chunks; c += 64) t += partial[(long long)c * count + j];
  for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
  if (lane == 0) out[j] = t;
}

// AdamW over every parameter tensor of the network in one launch: segment
// k covers elements [first[k], first[k + 1]) of the concatenation.
constexpr int kMaxSegments = 2 * 8 + 2;
struct AdamSegment {
  float *p;
  const float *g;
  float *m1, *m2;
  __bf16 *copy, *copy_t;
  int rows, cols;
  float wd;
};
struct AdamSegments {
  AdamSegment s[kMaxSegments];
  long long first[kMaxSegments + 1];
  int count;
};

__global__ __launch_bounds__(256) void adamw_all(AdamSegments segs, float lr, float b1,
                                                 float b2, float eps, float c1, float c2) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= segs.first[segs.count]) return;
  int k = 0;
  while (t >= segs.first[k + 1]) k++;
  const AdamSegment &a = segs.s[k];
  const long long i = t - segs.first[k];
  const float gi = a.g[i];
  const float m1 = b1 * a.m1[i] + (1.f - b1) * gi;
  const float m2 = b2 * a.m2[i] + (1.f - b2) * gi * gi;
  a.m1[i] = m1;
  a.m2[i] = m2;
  const float x = a.p[i] * (1.f - lr * a.wd) - lr * (m1 / c1) / (sqrtf(m2 / c2) + eps);
  a.p[i] = x;
  if (a.copy) a.copy[i] = to_bf16(x);
  if (a.copy_t) {
    const int r = (int)(i / a.cols), c = (int)(i % a.cols);
    a.copy_t[(long long)c * a.rows + r] = to_bf16(x);
  }
}

__global__ __launch_bounds__(256) void gather_rows(const float *__restrict__ src,
                                                   const float *__restrict__ ysrc,
                                                   int width, long long n_src,
                                                   long long n, unsigned seed,
                                                   unsigned step,
                                                   float *__restrict__ dst,
                                                   float *__restrict__ ydst) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // splitmix-style hash of (seed, step, i) -> sample index
  unsigned long long z = ((unsigned long long)seed << 32) ^ ((unsigned long long)step * 0x9E3779B97F4A7C15ull) ^ (unsigned long long)i;
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  const long long j = (long long)(z % (unsigned long long)n_src);
  for (int c = 0; c < width; c++) dst[i * width + c] = src[j * width + c];
  ydst[i] = ysrc[j];
}

__global__ void fill_f32(float *p, long long n, float v) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    p[i] = v;
}

}  // namespace pe

using namespace pe;



// Launch with the largest output tile that still gives at least 256 blocks
// (one per CU), else 64 x 64.
static int launch_gemm_on(hipStream_t stream, const GemmArgs &g, int splits) {
  auto blocks = [&](int bm, int bn) {
    return (long long)((g.M + bm - 1) / bm) * ((g.N + bn - 1) / bn) * splits;
  };
  auto go = [&](auto kernel, int bm, int bn) {
    const dim3 grid((unsigned)((g.M + bm - 1) / bm), (unsigned)((g.N + bn - 1) / bn),
                    (unsigned)splits);
    hipLaunchKernelGGL(kernel, grid, dim3(256), 0, stream, g);
  };
  if (blocks(128, 128) >= 256) go(mfma_gemm<128, 128>, 128, 128);
  else if (blocks(128, 64) >= 256) go(mfma_gemm<128, 64>, 128, 64);
  else go(mfma_gemm<64, 64>, 64, 64);
  PE_LAUNCHED(mfma_gemm);
  return PE_OK;
}

static int launch_gemm(pe_mlp *h, const GemmArgs &g, int splits = 1) {
  return launch_gemm_on(h->stream, g, splits);
}

// Column sums: at least kSumRows rows per chunk and at most kSumChunks
// chunks.  Weight-gradient GEMMs: the batch is split into at most
// kMaxSplits chunks of at least kSplitK rows (multiples of the LDS stage),
// enough blocks to fill the chip for a batch of 8192 without moving more
// partial products than needed for larger ones.
constexpr int kSumRows = 16;
constexpr int kSumChunks = 512;
constexpr int kSplitK = 256;
constexpr int kMaxSplits = 64;

static int sum_rows(long long n) {
  return (int)std::max<long long>(kSumRows, (n + kSumChunks - 1) / kSumChunks);
}
static int split_chunk(long long n) {
  long long chunk = std::max<long long>(kSplitK, (n + kMaxSplits - 1) / kMaxSplits);
  return (int)((chunk + kK - 1) / kK * kK);
}

// out[width] = column sums of X [n, width] weighted by wt (see
// column_partials), through the handle's partial buffer.
static int column_sums(pe_mlp *h, const __bf16 *X, const float *wt, long long n, int width,
                       float *out) {
  const int rows = sum_rows(n);
  const int chunks = (int)((n + rows - 1) / rows);
  hipLaunchKernelGGL(column_partials, dim3((unsigned)chunks), dim3(256), 0, h->stream, X, wt,
                     n, width, rows, h->partial);
  PE_LAUNCHED(column_partials);
  const long long count = X ? width : 1;
  hipLaunchKernelGGL(sum_partials_wide, dim3((unsigned)((count + 3) / 4)), dim3(256), 0,
                     h->stream, h->partial, chunks, count, out);
  PE_LAUNCHED(sum_partials_wide);
  return PE_OK;
}

// Forward of n points already in h->raw; fills F, Hs, Zs (if keep) and yhat.
static int forward(pe_mlp *h, long long n, bool keep) {
  const MlpShape &s = h->shape;
  hipLaunchKernelGGL(mlp_features, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     h->stream, h->raw, n, s, h->shift, h->scale, h->proj, h->F,
                     keep ? h->FT : nullptr, h->ldt);
  PE_LAUNCHED(mlp_features);
  const __bf16 *x = h->F;
  int width = s.width;
  for (int l = 0; l < h->layers; l++) {
    GemmArgs g{};
    g.A = x;
    g.sam = width;
    g.sak = 1;
    g.B = h->Wh[l];         // B(k, n) = W[n, k]
    g.sbk = 1;
    g.sbn = h->fan_in[l];
    g.M = (int)n;
    g.N = h->fan_out[l];
    g.K = h->fan_in[l];
    g.epilogue = kBiasSilu;
    g.Z = keep ? h->Zs[l] : nullptr;
    g.ldz = h->fan_out[l];
    g.bias = h->b[l];
    g.H = h->Hs[l];
    g.ldh = h->fan_out[l];
    // the next layer's input, transposed for its weight gradient
    g.HT = keep && l + 1 < h->layers ? h->HT[l] : nullptr;
    g.ldht = h->ldt;
    int rc = launch_gemm(h, g);
    if (rc) return rc;
    x = h->Hs[l];
    width = h->fan_out[l];
  }
  hipLaunchKernelGGL(mlp_head, dim3((unsigned)((n * 64 + 255) / 256)), dim3(256), 0,
                     h->stream, x, n, width, h->w_head, h->b_head, h->yhat);
  PE_LAUNCHED(mlp_head);
  return PE_OK;
}

// Gradients of the weighted MSE of n points in raw / yt into gW, gb, ...;
// the batch loss is added to *loss (device memory, zeroed by the caller).
static int backward(pe_mlp *h, long long n, float *loss) {
  int rc = forward(h, n, true);
  if (rc) return rc;
  hipLaunchKernelGGL(mlp_loss, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream,
                     h->yhat, h->yt, n, h->cfg.under_weight, h->dy, loss);
  PE_LAUNCHED(mlp_loss);
  const int L = h->layers;
  const int top = h->fan_out[L - 1];
  if ((rc = column_sums(h, h->Hs[L - 1], h->dy, n, top, h->gw_head))) return rc;
  if ((rc = column_sums(h, nullptr, h->dy, n, 1, h->gb_head))) return rc;
  hipLaunchKernelGGL(mlp_head_back, dim3((unsigned)((n + 63) / 64), (unsigned)((top + 63) / 64)),
                     dim3(256), 0, h->stream, h->Zs[L - 1], n, top, h->w_head, h->dy,
                     h->dZ[L - 1], h->dZT[L - 1], h->ldt);
  PE_LAUNCHED(mlp_head_back);
  for (int l = L - 1; l >= 0; l--) {
    // gW[o, i] = sum_m dZ[m, o] X[m, i] over the transposed copies:
    // A(o, m) = dZT[o * ldt + m], B(m, i) = XT[i * ldt + m] (both k-contiguous)
    GemmArgs g{};
    g.A = h->dZT[l];
    g.sam = h->ldt;
    g.sak = 1;
    g.B = l == 0 ? h->FT : h->HT[l - 1];
    g.sbk = 1;
    g.sbn = h->ldt;
    g.M = h->fan_out[l];
    g.N = h->fan_in[l];
    g.K = (int)n;
    g.epilogue = kStoreF32;
    // split over the batch: partial products in h->partial, then summed
    const int kchunk = split_chunk(n);
    const int splits = (int)((n + kchunk - 1) / kchunk);
    const long long count = (long long)h->fan_out[l] * h->fan_in[l];
    g.C = h->partial;
    g.ldc = h->fan_in[l];
    g.kchunk = kchunk;
    g.cstride = count;
    if ((rc = launch_gemm(h, g, splits))) return rc;
    hipLaunchKernelGGL(sum_partials, dim3((unsigned)((count + 255) / 256)), dim3(256), 0,
                       h->stream, h->partial, splits, count, h->gW[l]);
    PE_LAUNCHED(sum_partials);
    if ((rc = column_sums(h, h->dZ[l], nullptr, n, h->fan_out[l], h->gb[l]))) return rc;
    if (l > 0) {
      // dZ_{l-1}[m, i] = (sum_o dZ_l[m, o] W_l[o, i]) * silu'(Z_{l-1}[m, i])
      GemmArgs d{};
      d.A = h->dZ[l];
      d.sam = h->fan_out[l];
      d.sak = 1;
      d.B = h->Wt[l];       // B(k = o, n = i) = Wt[i, o] (k-contiguous)
      d.sbk = 1;
      d.sbn = h->fan_out[l];
      d.M = (int)n;
      d.N = h->fan_in[l];
      d.K = h->fan_out[l];
      d.epilogue = kSiluGrad;
      d.H = h->dZ[l - 1];
      d.ldh = h->fan_in[l];
      d.Z = h->Zs[l - 1];
      d.ldz = h->fan_in[l];
      d.HT = h->dZT[l - 1];
      d.ldht = h->ldt;
      if ((rc = launch_gemm(h, d))) return rc;
    }
  }
  return PE_OK;
}

static int optimizer_step(pe_mlp *h) {
  h->step++;
  const pe_mlp_config &c = h->cfg;
  const float c1 = 1.f - std::pow(c.beta1, (float)h->step);
  const float c2 = 1.f - std::pow(c.beta2, (float)h->step);
  AdamSegments segs{};
  auto add = [&](float *p, float *g, float *m, float *v, long long n, __bf16 *cp, __bf16 *ct,
                 int rows, int cols, float wd) {
    segs.s[segs.count] = AdamSegment{p, g, m, v, cp, ct, rows, cols, wd};
    segs.first[segs.count + 1] = segs.first[segs.count] + n;
    segs.count++;
  };
  for (int l = 0; l < h->layers; l++) {
    add(h->W[l], h->gW[l], h->mW[l], h->vW[l], (long long)h->fan_out[l] * h->fan_in[l],
        h->Wh[l], h->Wt[l], h->fan_out[l], h->fan_in[l], c.weight_decay);
    add(h->b[l], h->gb[l], h->mb[l], h->vb[l], h->fan_out[l], nullptr, nullptr, 1,
        h->fan_out[l], 0.f);
  }
  const int top = h->fan_out[h->layers - 1];
  add(h->w_head, h->gw_head, h->mw_head, h->vw_head, top, nullptr, nullptr, 1, top,
      c.weight_decay);
  add(h->b_head, h->gb_head, h->mb_head, h->vb_head, 1, nullptr, nullptr, 1, 1, 0.f);
  const long long total = segs.first[segs.count];
  hipLaunchKernelGGL(adamw_all, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     h->stream, segs, c.learning_rate, c.beta1, c.beta2, c.eps, c1, c2);
  PE_LAUNCHED(adamw_all);
  return PE_OK;
}

static int refresh_copies(pe_mlp *h) {
  for (int l = 0; l < h->layers; l++) {
    const int rows = h->fan_out[l], cols = h->fan_in[l];
    std::vector<float> w((size_t)rows * cols);
    PE_HIP(hipMemcpy(w.data(), h->W[l], w.size() * 4, hipMemcpyDeviceToHost));
    std::vector<__bf16> a(w.size()), t(w.size());
    for (int r = 0; r < rows; r++)
      for (int c = 0; c < cols; c++) {
        a[(size_t)r * cols + c] = (__bf16)w[(size_t)r * cols + c];
        t[(size_t)c * rows + r] = (__bf16)w[(size_t)r * cols + c];
      }
    PE_HIP(hipMemcpy(h->Wh[l], a.data(), a.size() * 2, hipMemcpyHostToDevice));
    PE_HIP(hipMemcpy(h->Wt[l], t.data(), t.size() * 2, hipMemcpyHostToDevice));
  }
  return PE_OK;
}

// Flattened parameter order: for each hidden layer W[out,in], b[out]; then
// head w[top], b[1].
static long long param_count(const pe_mlp *h) {
  long long n = 0;
  for (int l = 0; l < h->layers; l++) n += (long long)h->fan_out[l] * (h->fan_in[l] + 1);
  return n + h->fan_out[h->layers - 1] + 1;
}

template <typename F>
static int for_each_param(pe_mlp *h, F f) {
  long long at = 0;
  int rc;
  for (int l = 0; l < h->layers; l++) {
    const long long nw = (long long)h->fan_out[l] * h->fan_in[l];
    if ((rc = f(h->W[l], h->gW[l], at, nw))) return rc;
    at += nw;
    if ((rc = f(h->b[l], h->gb[l], at, (long long)h->fan_out[l]))) return rc;
    at += h->fan_out[l];
  }
  const int top = h->fan_out[h->layers - 1];
  if ((rc = f(h->w_head, h->gw_head, at, (long long)top))) return rc;
  at += top;
  return f(h->b_head, h->gb_head, at, 1ll);
}

static int ensure_capacity(pe_mlp *h, long long n) {
  if (n <= h->cap) return PE_OK;
  auto free_all = [&]() {
    for (void *p : {(void *)h->raw, (void *)h->yt, (void *)h->yhat, (void *)h->dy, (void *)h->F,
                    (void *)h->partial, (void *)h->FT})
      if (p) (void)hipFree(p);
    for (auto *v : {&h->Hs, &h->dZ, &h->Zs, &h->HT, &h->dZT})
      for (auto *&p : *v) {
        if (p) (void)hipFree(p);
        p = nullptr;
      }
    h->raw = h->yt = h->yhat = h->dy = h->partial = nullptr;
    h->F = h->FT = nullptr;
  };
  free_all();
  h->cap = 0;
  const long long cap = std::max(n, 1024ll);
  PE_HIP(hipMalloc(&h->raw, sizeof(float) * cap * h->shape.raw));
  PE_HIP(hipMalloc(&h->yt, sizeof(float) * cap));
  PE_HIP(hipMalloc(&h->yhat, sizeof(float) * cap));
  PE_HIP(hipMalloc(&h->dy, sizeof(float) * cap));
  PE_HIP(hipMalloc(&h->F, sizeof(__bf16) * cap * h->shape.width));
  const long long ldt = (cap + 7) / 8 * 8;
  PE_HIP(hipMalloc(&h->FT, sizeof(__bf16) * ldt * h->shape.width));
  {
    long long widest = 1, largest = 1;
    for (int l = 0; l < h->layers; l++) {
      widest = std::max<long long>(widest, h->fan_out[l]);
      largest = std::max<long long>(largest, (long long)h->fan_out[l] * h->fan_in[l]);
    }
    const long long splits = (cap + split_chunk(cap) - 1) / split_chunk(cap);
    const long long chunks = (cap + sum_rows(cap) - 1) / sum_rows(cap);
    PE_HIP(hipMalloc(&h->partial,
                     sizeof(float) * std::max(splits * largest, chunks * widest)));
  }
  for (int l = 0; l < h->layers; l++) {
    PE_HIP(hipMalloc(&h->Hs[l], sizeof(__bf16) * cap * h->fan_out[l]));
    PE_HIP(hipMalloc(&h->dZ[l], sizeof(__bf16) * cap * h->fan_out[l]));
    PE_HIP(hipMalloc(&h->Zs[l], sizeof(__bf16) * cap * h->fan_out[l]));
    PE_HIP(hipMalloc(&h->HT[l], sizeof(__bf16) * ldt * h->fan_out[l]));
    PE_HIP(hipMalloc(&h->dZT[l], sizeof(__bf16) * ldt * h->fan_out[l]));
  }
  h->ldt = ldt;
  h->cap = cap;
  return PE_OK;
}

namespace pe {
int mlp_reserve(pe_mlp *h, long long n) { return ensure_capacity(h, n); }
int mlp_forward_loaded(pe_mlp *h, long long n) { return forward(h, n, false); }
}  // namespace pe

extern "C" int pe_mlp_create(const pe_mlp_config *cfg, pe_mlp **out) {
  clear_error();
  if (!cfg || !out) return fail(PE_ERR_INVALID, "null argument");
  *out = nullptr;
  const pe_mlp_config &c = *cfg;
  if (c.extra < 0 || c.extra > kMaxExtra || c.rff < 0 || c.rff % 2 || c.rff > 64 || c.hidden < 16 ||
      c.hidden % 64 || c.layers < 1 || c.layers > 8 || !(c.learning_rate > 0) ||
      !(c.under_weight >= 1.f && c.under_weight <= 1e4f))
    return fail(PE_ERR_INVALID, "invalid MLP configuration");
  int dev = 0;
  int rc = require_device(&dev);
  if (rc) return rc;
  pe_mlp *h = new pe_mlp();
  h->device = dev;
  h->cfg = c;
  h->shape.raw = 8 + c.extra;
  h->shape.extra = c.extra;
  h->shape.rff = c.rff;
  h->shape.width = ((8 + c.rff + c.extra + kRelative + 15) / 16) * 16;
  h->shape.hidden = c.hidden;
  h->layers = c.layers;
  for (int l = 0; l < c.layers; l++) {
    h->fan_in.push_back(l == 0 ? h->shape.width : c.hidden);
    h->fan_out.push_back(c.hidden);
  }
  h->W.resize(c.layers); h->b.resize(c.layers); h->gW.resize(c.layers); h->gb.resize(c.layers);
  h->mW.resize(c.layers); h->vW.resize(c.layers); h->mb.resize(c.layers); h->vb.resize(c.layers);
  h->Wh.resize(c.layers); h->Wt.resize(c.layers);
  h->Hs.resize(c.layers); h->dZ.resize(c.layers); h->Zs.resize(c.layers);
  h->HT.resize(c.layers); h->dZT.resize(c.layers);
  auto body = [&]() -> int {
    PE_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    std::mt19937_64 rng((unsigned long long)c.seed);
    std::normal_distribution<float> normal(0.f, 1.f);
    for (int l = 0; l < c.layers; l++) {
      const long long nw = (long long)h->fan_out[l] * h->fan_in[l];
      for (float **p : {&h->W[l], &h->gW[l], &h->mW[l], &h->vW[l]})
        PE_HIP(hipMalloc(p, sizeof(float) * nw));
      for (float **p : {&h->b[l], &h->gb[l], &h->mb[l], &h->vb[l]})
        PE_HIP(hipMalloc(p, sizeof(float) * h->fan_out[l]));
      PE_HIP(hipMalloc(&h->Wh[l], sizeof(__bf16) * nw));
      PE_HIP(hipMalloc(&h->Wt[l], sizeof(__bf16) * nw));
      // He-style init for SiLU; features beyond the real inputs are zero.
      const int real_in = l == 0 ? 8 + c.rff + c.extra + kRelative : h->fan_in[l];
      std::vector<float> w(nw, 0.f);
      const float sd = std::sqrt(2.f / real_in);
      for (int o = 0; o < h->fan_out[l]; o++)
        for (int i = 0; i < real_in; i++) w[(size_t)o * h->fan_in[l] + i] = sd * normal(rng);
      PE_HIP(hipMemcpy(h->W[l], w.data(), nw * 4, hipMemcpyHostToDevice));
      std::vector<float> zero(std::max<long long>(nw, h->fan_out[l]), 0.f);
      PE_HIP(hipMemcpy(h->b[l], zero.data(), h->fan_out[l] * 4, hipMemcpyHostToDevice));
      for (float *p : {h->mW[l], h->vW[l], h->gW[l]})
        PE_HIP(hipMemcpy(p, zero.data(), nw * 4, hipMemcpyHostToDevice));
      for (float *p : {h->mb[l], h->vb[l], h->gb[l]})
        PE_HIP(hipMemcpy(p, zero.data(), h->fan_out[l] * 4, hipMemcpyHostToDevice));
    }
    const int top = c.hidden;
    for (float **p : {&h->w_head, &h->gw_head, &h->mw_head, &h->vw_head})
      PE_HIP(hipMalloc(p, sizeof(float) * top));
    for (float **p : {&h->b_head, &h->gb_head, &h->mb_head, &h->vb_head, &h->loss})
      PE_HIP(hipMalloc(p, sizeof(float) * 1));
    std::vector<float> w(top), zero(top, 0.f);
    for (int i = 0; i < top; i++) w[i] = std::sqrt(1.f / top) * normal(rng);
    PE_HIP(hipMemcpy(h->w_head, w.data(), top * 4, hipMemcpyHostToDevice));
    for (float *p : {h->mw_head, h->vw_head, h->gw_head})
      PE_HIP(hipMemcpy(p, zero.data(), top * 4, hipMemcpyHostToDevice));
    for (float *p : {h->b_head, h->mb_head, h->vb_head, h->gb_head})
      PE_HIP(hipMemcpy(p, zero.data(), 4, hipMemcpyHostToDevice));
    // Normalisation and random Fourier projection.
    PE_HIP(hipMalloc(&h->shift, 32));
    PE_HIP(hipMalloc(&h->scale, 32));
    PE_HIP(hipMemcpy(h->shift, c.state_shift, 32, hipMemcpyHostToDevice));
    PE_HIP(hipMemcpy(h->scale, c.state_scale, 32, hipMemcpyHostToDevice));
    const int half = std::max(1, c.rff / 2);
    std::vector<float> proj((size_t)half * 8);
    for (auto &x : proj) x = c.rff_scale * normal(rng);
    PE_HIP(hipMalloc(&h->proj, proj.size() * 4));
    PE_HIP(hipMemcpy(h->proj, proj.data(), proj.size() * 4, hipMemcpyHostToDevice));
    int r = refresh_copies(h);
    if (r) return r;
    return ensure_capacity(h, 4096);
  };
  rc = body();
  if (rc) {
    pe_mlp_destroy(h);
    return rc;
  }
  *out = h;
  return PE_OK;
}

extern "C" int pe_mlp_param_count(pe_mlp *h, int64_t *count) {
  clear_error();
  if (!h || !count) return fail(PE_ERR_INVALID, "invalid argument");
  *count = param_count(h);
  return PE_OK;
}

extern "C" int pe_mlp_get_params(pe_mlp *h, float *out, int64_t count) {
  clear_error();
  if (!h || !out || count != param_count(h)) return fail(PE_ERR_INVALID, "size mismatch");
  PE_HIP(hipSetDevice(h->device));
  return for_each_param(h, [&](float *p, float *, long long at, long long n) -> int {
    PE_HIP(hipMemcpy(out + at, p, n * 4, hipMemcpyDeviceToHost));
    return PE_OK;
  });
}

extern "C" int pe_mlp_set_params(pe_mlp *h, const float *in, int64_t count) {
  clear_error();
  if (!h || !in || count != param_count(h)) return fail(PE_ERR_INVALID, "size mismatch");
  for (long long k = 0; k < count; k++)
    if (!std::isfinite(in[k])) return fail(PE_ERR_INVALID, "non-finite parameter");
  PE_HIP(hipSetDevice(h->device));
  int rc = for_each_param(h, [&](float *p, float *, long long at, long long n) -> int {
    PE_HIP(hipMemcpy(p, in + at, n * 4, hipMemcpyHostToDevice));
    return PE_OK;
  });
  if (rc) return rc;
  return refresh_copies(h);
}

extern "C" int pe_mlp_forward(pe_mlp *h, const float *raw, int64_t n, float *out) {
  clear_error();
  if (!h || n < 0 || (n && (!raw || !out))) return fail(PE_ERR_INVALID, "invalid argument");
  PE_HIP(hipSetDevice(h->device));
  const long long chunk = 1 << 18;
  for (long long at = 0; at < n; at += chunk) {
    const long long m = std::min(chunk, (long long)n - at);
    int rc = ensure_capacity(h, m);
    if (rc) return rc;
    PE_HIP(hipMemcpyAsync(h->raw, raw + at * h->shape.raw, sizeof(float) * m * h->shape.raw,
                          hipMemcpyHostToDevice, h->stream));
    rc = forward(h, m, false);
    if (rc) return rc;
    PE_HIP(hipMemcpyAsync(out + at, h->yhat, sizeof(float) * m, hipMemcpyDeviceToHost,
                          h->stream));
  }
  PE_HIP(hipStreamSynchronize(h->stream));
  return PE_OK;
}

extern "C" int pe_mlp_features(pe_mlp *h, const float *raw, int64_t n, float *features,
                               int32_t *width) {
  clear_error();
  if (!h || n < 0 || (n && (!raw || !features))) return fail(PE_ERR_INVALID, "invalid argument");
  if (width) *width = h->shape.width;
  if (n == 0) return PE_OK;
  PE_HIP(hipSetDevice(h->device));
  int rc = ensure_capacity(h, n);
  if (rc) return rc;
  PE_HIP(hipMemcpyAsync(h->raw, raw, sizeof(float) * n * h->shape.raw, hipMemcpyHostToDevice,
                        h->stream));
  hipLaunchKernelGGL(mlp_features, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, h->stream,
                     h->raw, (long long)n, h->shape, h->shift, h->scale, h->proj, h->F,
                     nullptr, 0ll);
  PE_LAUNCHED(mlp_features);
  std::vector<__bf16> f((size_t)n * h->shape.width);
  PE_HIP(hipMemcpyAsync(f.data(), h->F, f.size() * 2, hipMemcpyDeviceToHost, h->stream));
  PE_HIP(hipStreamSynchronize(h->stream));
  for (size_t k = 0; k < f.size(); k++) features[k] = (float)f[k];
  return PE_OK;
}

extern "C" int pe_mlp_gradients(pe_mlp *h, const float *raw, const float *target,
                                int64_t n, float *grads, float *loss) {
  clear_error();
  if (!h || n < 1 || !raw || !target || !grads) return fail(PE_ERR_INVALID, "invalid argument");
  PE_HIP(hipSetDevice(h->device));
  int rc = ensure_capacity(h, n);
  if (rc) return rc;
  PE_HIP(hipMemcpyAsync(h->raw, raw, sizeof(float) * n * h->shape.raw, hipMemcpyHostToDevice,
                        h->stream));
  PE_HIP(hipMemcpyAsync(h->yt, target, sizeof(float) * n, hipMemcpyHostToDevice, h->stream));
  hipLaunchKernelGGL(fill_f32, dim3(1), dim3(64), 0, h->stream, h->loss, 1ll, 0.f);
  if ((rc = backward(h, n, h->loss))) return rc;
  float l = 0;
  PE_HIP(hipMemcpyAsync(&l, h->loss, 4, hipMemcpyDeviceToHost, h->stream));
  PE_HIP(hipStreamSynchronize(h->stream));
  if (loss) *loss = l;
  return for_each_param(h, [&](float *, float *g, long long at, long long m) -> int {
    PE_HIP(hipMemcpy(grads + at, g, m * 4, hipMemcpyDeviceToHost));
    return PE_OK;
  });
}

extern "C" int pe_mlp_set_dataset(pe_mlp *h, const float *raw, const float *target, int64_t n) {
  clear_error();
  if (!h || n < 1 || !raw || !target) return fail(PE_ERR_INVALID, "invalid argument");
  PE_HIP(hipSetDevice(h->device));
  if (h->data_x) (void)hipFree(h->data_x);
  if (h->data_y) (void)hipFree(h->data_y);
  h->data_x = h->data_y = nullptr;
  PE_HIP(hipMalloc(&h->data_x, sizeof(float) * n * h->shape.raw));
  PE_HIP(hipMalloc(&h->data_y, sizeof(float) * n));
  PE_HIP(hipMemcpy(h->data_x, raw, sizeof(float) * n * h->shape.raw, hipMemcpyHostToDevice));
  PE_HIP(hipMemcpy(h->data_y, target, sizeof(float) * n, hipMemcpyHostToDevice));
  h->data_n = n;
  return PE_OK;
}

extern "C" int pe_mlp_train(pe_mlp *h, int32_t steps, int32_t batch, uint32_t seed,
                            float *losses) {
  clear_error();
  if (!h || steps < 0 || batch < 1 || !h->data_n)
    return fail(PE_ERR_INVALID, "invalid argument or no dataset");
  PE_HIP(hipSetDevice(h->device));
  int rc = ensure_capacity(h, batch);
  if (rc) return rc;
  // Per-step losses stay on the device until the end: no step waits on the
  // host.
  float *history = nullptr;
  PE_HIP(hipMalloc(&history, sizeof(float) * (steps > 0 ? steps : 1)));
  auto body = [&]() -> int {
    hipLaunchKernelGGL(fill_f32, dim3(64), dim3(256), 0, h->stream, history,
                       (long long)(steps > 0 ? steps : 1), 0.f);
    PE_LAUNCHED(fill_f32);
    const unsigned first = (unsigned)h->step;
    for (int s = 0; s < steps; s++) {
      hipLaunchKernelGGL(gather_rows, dim3((unsigned)((batch + 255) / 256)), dim3(256), 0,
                         h->stream, h->data_x, h->data_y, h->shape.raw, h->data_n,
                         (long long)batch, seed, first + (unsigned)s, h->raw, h->yt);
      PE_LAUNCHED(gather_rows);
      int r = backward(h, batch, history + s);
      if (r) return r;
      if ((r = optimizer_step(h))) return r;
    }
    if (losses && steps > 0)
      PE_HIP(hipMemcpyAsync(losses, history, sizeof(float) * steps, hipMemcpyDeviceToHost,
                            h->stream));
    PE_HIP(hipStreamSynchronize(h->stream));
    return PE_OK;
  };
  rc = body();
  (void)hipFree(history);
  return rc;
}

extern "C" int pe_mlp_gemm_check(const uint16_t *a, const uint16_t *b, int32_t M,
                                 int32_t N, int32_t K, int32_t transposed_b, float *c) {
  // C[m, n] = sum_k A[m, k] B(k, n); B given [K, N] (transposed_b = 0) or
  // [N, K] (transposed_b = 1); inputs are bf16 bit patterns.
  clear_error();
  if (!a || !b || !c || M < 1 || N < 1 || K < 1) return fail(PE_ERR_INVALID, "invalid argument");
  int rc = require_device(nullptr);
  if (rc) return rc;
  void *da = nullptr, *db = nullptr, *dc = nullptr;
  auto body = [&]() -> int {
    PE_HIP(hipMalloc(&da, 2ll * M * K));
    PE_HIP(hipMalloc(&db, 2ll * K * N));
    PE_HIP(hipMalloc(&dc, 4ll * M * N));
    PE_HIP(hipMemcpy(da, a, 2ll * M * K, hipMemcpyHostToDevice));
    PE_HIP(hipMemcpy(db, b, 2ll * K * N, hipMemcpyHostToDevice));
    GemmArgs g{};
    g.A = (const __bf16 *)da;
    g.sam = K;
    g.sak = 1;
    g.B = (const __bf16 *)db;
    if (transposed_b) {
      g.sbk = 1;
      g.sbn = K;
    } else {
      g.sbk = N;
      g.sbn = 1;
    }
    g.M = M;
    g.N = N;
    g.K = K;
    g.epilogue = kStoreF32;
    g.C = (float *)dc;
    g.ldc = N;
    int r = launch_gemm_on(0, g, 1);
    if (r) return r;
    PE_HIP(hipMemcpy(c, dc, 4ll * M * N, hipMemcpyDeviceToHost));
    return PE_OK;
  };
  rc = body();
  for (void *p : {da, db, dc})
    if (p) (void)hipFree(p);
  return rc;
}

extern "C" void pe_mlp_destroy(pe_mlp *h) {
  if (!h) return;
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  auto f = [](void *p) {
    if (p) (void)hipFree(p);
  };
  for (int l = 0; l < h->layers; l++) {
    for (void *p : {(void *)h->W[l], (void *)h->b[l], (void *)h->gW[l], (void *)h->gb[l],
                    (void *)h->mW[l], (void *)h->vW[l], (void *)h->mb[l], (void *)h->vb[l],
                    (void *)h->Wh[l], (void *)h->Wt[l], (void *)h->Hs[l], (void *)h->dZ[l],
                    (void *)h->Zs[l], (void *)h->HT[l], (void *)h->dZT[l]})
      f(p);
  }
  for (void *p : {(void *)h->w_head, (void *)h->b_head, (void *)h->gw_head, (void *)h->gb_head,
                  (void *)h->mw_head, (void *)h->vw_head, (void *)h->mb_head, (void *)h->vb_head,
                  (void *)h->shift, (void *)h->scale, (void *)h->proj, (void *)h->raw,
                  (void *)h->yt, (void *)h->yhat, (void *)h->dy, (void *)h->loss, (void *)h->F,
                  (void *)h->FT, (void *)h->partial,
                  (void *)h->data_x, (void *)h->data_y})
    f(p);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}
