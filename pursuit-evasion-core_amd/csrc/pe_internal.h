// Internal state of the pe_amd handles, shared by the translation units that
// operate on more than one handle kind (the learned warm start).
#pragma once

#include <vector>

#include "pe_common.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace pe {

constexpr int kStencil = 16;
constexpr int kMaxActions = 15;  // order byte packs two 4-bit command indices

struct SlMap {
  int np, ne, na, npos_e;
  long long v_off;      // pair offset of this map in the value arrays
  long long g_off;      // offset in G
  long long tab_off;    // offset in the inside / near tables
  int p_row, e_row;     // first state row in the concatenated car tables
  int ipos_off;         // first entry of this map's P positions in the inside lists
  int active;           // swept (1), or frozen by a converged pe_sl_run (0)
  float exit_value, pl, pw, el, ew;
};

struct MlpShape {
  int raw;      // raw inputs per point: 8 state + extra
  int extra;    // extra map features
  int rff;      // sin/cos pairs count * 2
  int width;    // padded feature width (multiple of 32)
  int hidden;
};

// Forward pass of n points whose raw inputs are already in h->raw; the
// predictions land in h->yhat (both device buffers of the handle).
int mlp_forward_loaded(struct ::pe_mlp *h, long long n);
int mlp_reserve(struct ::pe_mlp *h, long long n);

// One order probe of a solver handle (pe_sl8d.hip).
int sl_probe_orders(pe_sl *h, const float *field, float *scratch);

}  // namespace pe

struct pe_sl {
  int device = 0;
  hipStream_t stream = nullptr;
  int B = 0, na = 0;
  std::vector<pe::SlMap> maps;
  pe::SlMap *d_maps = nullptr;
  int np_max = 0, ne_max = 0;
  bool f64 = false;      // FP64 values (va, vb, G hold doubles)
  size_t esize = 4;      // bytes per value
  size_t lds_bytes = 0;  // G rows of one P state, [na][ne_max] floats (0: no LDS E side)
  int lds_threads = 1024;  // threads per block of the LDS E side
  long long pairs = 0, g_size = 0, active = 0;
  // car tables
  int *pids = nullptr, *eids = nullptr, *ppos = nullptr, *epos = nullptr;
  // E stencils as records of 16-bit relative LDS-tile offsets + weights
  // (LDS E side, when every offset fits)
  void *erec = nullptr;
  // E indices of the inside pairs of each (map, P position), ascending:
  // ilist[ioff[k] .. ioff[k + 1]) for k = SlMap::ipos_off + position.
  int *ilist = nullptr, *ioff = nullptr;
  int *nlist = nullptr, *noff = nullptr;  // the same for the near pairs
  // P stencils compacted (non-zero weights first, in vertex order) and the
  // count of non-zero weights per (state, command); pids / pw hold them.
  uint8_t *pcount = nullptr;
  // FP32 P side from the union of each P state's stencil rows (na <= 9):
  // P state row i reads urows[uoff[i] .. uoff[i + 1]) (ascending, within
  // its map) with weights uw[row][9] per command (0 where the command does
  // not use the row); null when not built
  int *uoff = nullptr, *urows = nullptr;
  float *uw = nullptr;
  float *pw = nullptr, *ew = nullptr;
  // allowed commands of each state as a bit mask (bit u), for the E side
  uint16_t *pmask = nullptr, *emask = nullptr;
  uint8_t *pallowed = nullptr, *eallowed = nullptr, *inside = nullptr,
          *near = nullptr, *order = nullptr;
  double *pstates = nullptr, *estates = nullptr;
  float *pclear = nullptr, *eclear = nullptr;  // per-state clearance (may be zero)
  float *va = nullptr, *vb = nullptr, *G = nullptr;
  // while set, the E side starts every pair from +3e38 instead of its value,
  // so each inside pair records its argmin / argmax commands (an order
  // probe into scratch buffers; the handle's fields are not touched)
  bool probe_orders = false;
  // residual ring [slots][B] and graph slots [graph_sweeps[0]][B]
  unsigned long long *ring = nullptr, *gslots = nullptr;
  int ring_slots = 256, ring_used = 0;
  std::vector<double> trace;  // [sweeps][B]
  double last = 0, update_seconds = 0;
  long long sweeps = 0;
  bool profiling = false;
  std::vector<hipEvent_t> events;
  std::vector<int> kinds;
  double ms[2] = {0, 0};
  long long launches[2] = {0, 0};
  hipEvent_t t0 = nullptr, t1 = nullptr;
  // captured sweeps: a long graph replayed while enough sweeps remain, then
  // a short one (both even, so the value buffers end where they started)
  static constexpr int kGraphs = 2;
  hipGraphExec_t graph[kGraphs] = {nullptr, nullptr};
  int graph_sweeps[kGraphs] = {16, 2};
  int graph_flags[kGraphs] = {-1, -1};
  const float *graph_src[kGraphs] = {nullptr, nullptr};  // field the sweeps start from
};

struct pe_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  pe_mlp_config cfg{};
  pe::MlpShape shape{};
  int layers = 0;                   // hidden layers (MFMA GEMMs)
  std::vector<int> fan_in, fan_out; // per GEMM layer
  // parameters: per layer W [out, in] fp32 master + bf16 + bf16^T, b [out]
  std::vector<float *> W, b, gW, gb, mW, vW, mb, vb;
  std::vector<__bf16 *> Wh, Wt;
  float *w_head = nullptr, *b_head = nullptr, *gw_head = nullptr, *gb_head = nullptr;
  float *mw_head = nullptr, *vw_head = nullptr, *mb_head = nullptr, *vb_head = nullptr;
  float *shift = nullptr, *scale = nullptr, *proj = nullptr;
  // activations for batches of up to cap points
  long long cap = 0;
  float *raw = nullptr, *yt = nullptr, *yhat = nullptr, *dy = nullptr, *loss = nullptr;
  float *partial = nullptr;  // split-K products and column-sum partials of a batch
  __bf16 *F = nullptr;
  std::vector<__bf16 *> Hs, dZ;  // per hidden layer
  std::vector<__bf16 *> Zs;  // pre-activations (bf16), kept for the SiLU derivative
  // transposed copies [width, ldt] of F, Hs and dZ, written by the kernels
  // that produce them during training: the operands of the weight-gradient
  // GEMMs, which sum over the batch, then sit k-contiguous
  __bf16 *FT = nullptr;
  std::vector<__bf16 *> HT, dZT;
  long long ldt = 0;         // cap rounded up to a multiple of 8
  // dataset
  float *data_x = nullptr, *data_y = nullptr;
  long long data_n = 0;
  long long step = 0;
};
