"""Value-surrogate MLP on the GPU (BF16 MFMA kernels, FP32 accumulation).

    net = ValueMLP.for_game(p_grid, e_grid)      # 8 + 3 inputs, 32 RFF, 4 x 256
    net.fit(inputs, values, steps=2000, batch=8192)
    v = net(inputs)                              # predictions, float32
    solver.warm_start(net, delta)                # V0 = min(phi, net + delta)

Training follows the reference's supervised value-net conventions
(local_game/training.py): weighted MSE with weight 1 + 4 exp(-|y| / 0.2)
(:351-352) and AdamW (:185-188).  `under_weight` > 1 additionally weights
underestimates (prediction below the label), so the fit leans towards an
upper envelope: a warm start is only exact where it is not below the
converged value.  The network itself is new: the reference has no surrogate
on this path.
"""

from __future__ import annotations

import ctypes

import numpy as np

from .. import _native
from .._native import check, lib, ptr

EXTRA = 3  # phi, clearance P, clearance E


class ValueMLP:
    def __init__(self, *, shift, scale, extra=EXTRA, rff=32, hidden=256, layers=4, seed=0,
                 rff_scale=1.0, learning_rate=1e-3, weight_decay=1e-5, beta1=0.9,
                 beta2=0.999, eps=1e-8, under_weight=1.0):
        c = _native.MLPConfig()
        c.extra, c.rff, c.hidden, c.layers, c.seed = extra, rff, hidden, layers, seed
        c.rff_scale = rff_scale
        for k in range(8):
            c.state_shift[k] = float(shift[k])
            c.state_scale[k] = float(scale[k])
        c.learning_rate, c.weight_decay = learning_rate, weight_decay
        c.beta1, c.beta2, c.eps = beta1, beta2, eps
        c.under_weight = under_weight
        self.config = c
        self.inputs = 8 + extra
        self._h = ctypes.c_void_p()
        check(lib().pe_mlp_create(ctypes.byref(c), ctypes.byref(self._h)))

    @classmethod
    def for_game(cls, p, e, **kw):
        """Normalisation from the two cars' grids (positions, headings, speeds)."""
        lo = np.concatenate([p.states.min(0), e.states.min(0)])
        hi = np.concatenate([p.states.max(0), e.states.max(0)])
        shift = (hi + lo) / 2
        scale = 2.0 / np.maximum(hi - lo, 1e-6)
        return cls(shift=shift, scale=scale, **kw)

    def close(self):
        if self._h:
            lib().pe_mlp_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _raw(self, x):
        x = np.ascontiguousarray(x, np.float32)
        if x.ndim != 2 or x.shape[1] != self.inputs:
            raise ValueError(f"inputs [n, {self.inputs}] required")
        return x

    def __call__(self, x):
        x = self._raw(x)
        out = np.empty(len(x), np.float32)
        check(lib().pe_mlp_forward(self._h, ptr(x), len(x), ptr(out)))
        return out

    def features(self, x):
        x = self._raw(x)
        width = ctypes.c_int32()
        check(lib().pe_mlp_features(self._h, ptr(x), 0, None, ctypes.byref(width)))
        out = np.empty((len(x), width.value), np.float32)
        check(lib().pe_mlp_features(self._h, ptr(x), len(x), ptr(out), ctypes.byref(width)))
        return out

    def param_count(self):
        n = ctypes.c_int64()
        check(lib().pe_mlp_param_count(self._h, ctypes.byref(n)))
        return n.value

    def get_params(self):
        out = np.empty(self.param_count(), np.float32)
        check(lib().pe_mlp_get_params(self._h, ptr(out), len(out)))
        return out

    def set_params(self, flat):
        flat = np.ascontiguousarray(flat, np.float32)
        check(lib().pe_mlp_set_params(self._h, ptr(flat), len(flat)))

    def unflatten(self, flat):
        """[(W [out, in], b [out]) per hidden layer] + [(w [hidden], b [1])]."""
        c = self.config
        width = ((8 + c.rff + c.extra + 5 + 15) // 16) * 16
        shapes, at, out = [], 0, []
        fan_in = width
        for _ in range(c.layers):
            shapes.append(((c.hidden, fan_in), (c.hidden,)))
            fan_in = c.hidden
        for ws, bs in shapes:
            nw = ws[0] * ws[1]
            out.append((flat[at: at + nw].reshape(ws), flat[at + nw: at + nw + bs[0]]))
            at += nw + bs[0]
        out.append((flat[at: at + c.hidden], flat[at + c.hidden: at + c.hidden + 1]))
        return out

    def gradients(self, x, y):
        x = self._raw(x)
        y = np.ascontiguousarray(y, np.float32)
        grads = np.empty(self.param_count(), np.float32)
        loss = ctypes.c_float()
        check(lib().pe_mlp_gradients(self._h, ptr(x), ptr(y), len(x), ptr(grads),
                                     ctypes.byref(loss)))
        return loss.value, grads

    def fit(self, x, y, *, steps, batch=8192, seed=0):
        """`steps` AdamW steps on device-drawn batches; returns the losses."""
        x = self._raw(x)
        y = np.ascontiguousarray(y, np.float32)
        check(lib().pe_mlp_set_dataset(self._h, ptr(x), ptr(y), len(x)))
        losses = np.empty(steps, np.float32)
        check(lib().pe_mlp_train(self._h, int(steps), int(batch), int(seed), ptr(losses)))
        return losses


def gemm_check(a_bits, b_bits, transposed_b):
    """The MFMA GEMM kernel alone on bf16 bit patterns (uint16 arrays)."""
    a = np.ascontiguousarray(a_bits, np.uint16)
    b = np.ascontiguousarray(b_bits, np.uint16)
    M, K = a.shape
    N = b.shape[0] if transposed_b else b.shape[1]
    c = np.empty((M, N), np.float32)
    check(lib().pe_mlp_gemm_check(ptr(a), ptr(b), M, N, K, int(bool(transposed_b)), ptr(c)))
    return c


def to_bf16_bits(x):
    """Round float32 to bf16 (nearest even) bit patterns and back-values."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    return r, (r.astype(np.uint32) << 16).view(np.float32)


def training_flops(net, batch):
    """Algorithmic FLOPs of one training step of `batch` points: forward
    (2 per multiply-add of every layer, head included) plus backward (the
    weight gradients of every layer and the input gradients of all but the
    first)."""
    c = net.config
    width = ((8 + c.rff + c.extra + 5 + 15) // 16) * 16
    dims = [width] + [c.hidden] * c.layers
    layer = [dims[i] * dims[i + 1] for i in range(c.layers)]
    forward = 2 * batch * (sum(layer) + c.hidden)
    backward = 2 * batch * (sum(layer) + sum(layer[1:]) + 2 * c.hidden)
    return forward + backward


def main(argv=None):
    """Training throughput on synthetic data: `steps` AdamW steps at
    `batch` points, printed as one JSON line with the algorithmic TFLOP/s."""
    import argparse
    import json
    import time

    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--points", type=int, default=1 << 20)
    a = ap.parse_args(argv)
    rng = np.random.default_rng(0)
    x = rng.uniform(-1, 1, (a.points, 8 + EXTRA)).astype(np.float32)
    y = (np.sin(3 * x[:, 0]) * np.cos(2 * x[:, 4]) + 0.1 * x[:, 8]).astype(np.float32)
    with ValueMLP(shift=np.zeros(8), scale=np.ones(8)) as net:
        net.fit(x, y, steps=5, batch=a.batch)
        t0 = time.perf_counter()
        losses = net.fit(x, y, steps=a.steps, batch=a.batch, seed=1)
        seconds = time.perf_counter() - t0
        flops = training_flops(net, a.batch) * a.steps
    print(json.dumps(dict(steps=a.steps, batch=a.batch, seconds=seconds,
                          ms_per_step=seconds / a.steps * 1e3,
                          tflops=flops / seconds / 1e12, loss_first=float(losses[0]),
                          loss_last=float(losses[-1]))), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
