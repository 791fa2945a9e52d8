// FP64 instances of the 8D backup kernels.  Built with -ffp-contract=off so
// that each stencil sum is a separate multiply and add, left to right over
// the positive weights: the arithmetic of the FP64 oracle sweep, for
// full-size FP32-against-FP64 checks on the GPU.
#include "sl8d_kernels.h"

namespace pe {

void sl_enqueue_f64(pe_sl *h, const void *src, void *dst, int flags, unsigned long long *slot,
                    bool record) {
  enqueue_sweep_t<double>(h, src, dst, flags, slot, record);
}

void sl_initialize_f64(pe_sl *h) {
  hipLaunchKernelGGL(sl_initialize<double>, pair_grid(h), dim3(256), 0, h->stream, h->d_maps,
                     h->pstates, h->estates, h->ppos, h->epos, h->inside,
                     reinterpret_cast<double *>(h->va), h->order);
}

bool sl_allow_lds_f64(size_t bytes) { return sl_allow_lds<double>(bytes); }

}  // namespace pe
