"""GPU checks of the BF16 MFMA MLP kernels.

* The GEMM kernel against a float64 matmul of the same bf16 values
  (exactly representable inputs, FP32 accumulation): relative error 1e-5.
  Odd sizes exercise the tile edges; both operand layouts are covered.
* Forward and backward against PyTorch autograd in float64 on the CPU
  (the checker), run on exactly the bf16 feature rows the kernels see and
  the bf16-rounded weights.  The kernels keep activations and gradients in
  bf16 between layers, so agreement is to bf16 precision: relative error of
  each gradient tensor <= 3e-2, cosine >= 0.999.  Both the symmetric loss
  and the one that weights underestimates 4x are checked.
* Training reduces the loss on a smooth target.
"""

import numpy as np
import pytest

from pe_amd.learned.mlp import ValueMLP, gemm_check, to_bf16_bits

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("M,N,K,tb", [(64, 64, 32, 1), (100, 70, 48, 1), (129, 256, 256, 0),
                                      (33, 17, 1000, 0)])
def test_mfma_gemm_matches_float64(gpu, M, N, K, tb):
    rng = np.random.default_rng(M * N + K)
    a_bits, a = to_bf16_bits(rng.normal(size=(M, K)).astype(np.float32))
    shape = (N, K) if tb else (K, N)
    b_bits, b = to_bf16_bits(rng.normal(size=shape).astype(np.float32))
    print(f"[gpu] gemm {M}x{N}x{K} tb={tb}", flush=True)
    c = gemm_check(a_bits, b_bits, tb)
    ref = a.astype(np.float64) @ (b.T if tb else b).astype(np.float64)
    err = np.abs(c - ref).max() / np.abs(ref).max()
    print(f"gemm {M}x{N}x{K}: rel err {err:.2e}")
    assert err < 1e-5


def make_net(seed=0, under_weight=1.0):
    return ValueMLP(shift=np.zeros(8), scale=np.full(8, 0.5), seed=seed, rff_scale=1.0,
                    under_weight=under_weight)


def random_inputs(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1.5, 1.5, (n, 11)).astype(np.float32)
    y = (np.sin(x[:, 0]) * np.cos(x[:, 5]) + 0.3 * x[:, 8]).astype(np.float32)
    return x, y


def torch_reference(net, feats, y, under_weight=1.0):
    torch = pytest.importorskip("torch")
    params = net.unflatten(net.get_params())
    ts = []
    h = torch.tensor(feats, dtype=torch.float64)
    for W, b in params[:-1]:
        _, Wb = to_bf16_bits(W)
        Wt = torch.tensor(Wb, dtype=torch.float64, requires_grad=True)
        bt = torch.tensor(b, dtype=torch.float64, requires_grad=True)
        ts += [Wt, bt]
        z = h @ Wt.T + bt
        h = z * torch.sigmoid(z)
    w4 = torch.tensor(params[-1][0], dtype=torch.float64, requires_grad=True)
    b4 = torch.tensor(params[-1][1], dtype=torch.float64, requires_grad=True)
    ts += [w4, b4]
    out = h @ w4 + b4
    yt = torch.tensor(y, dtype=torch.float64)
    w = (1 + 4 * torch.exp(-yt.abs() / 0.2)) * torch.where(out < yt, under_weight, 1.0)
    loss = (w * (out - yt) ** 2).mean()
    loss.backward()
    return out.detach().numpy(), loss.item(), [t.grad.numpy().ravel() for t in ts]


@pytest.mark.parametrize("under_weight,n", [(1.0, 512), (4.0, 512), (1.0, 333)])
def test_forward_and_gradients_match_float64_autograd(gpu, under_weight, n):
    """n = 333: a batch that is no multiple of 8 or of a GEMM tile, so the
    transposed copies of the weight-gradient GEMMs end in partial rows."""
    net = make_net(under_weight=under_weight)
    x, y = random_inputs(n, 1)
    print(f"[gpu] mlp forward/backward under_weight={under_weight} n={n}", flush=True)
    feats = net.features(x)
    pred = net(x)
    ref_out, ref_loss, ref_grads = torch_reference(net, feats, y, under_weight)
    rel = np.abs(pred - ref_out).max() / np.abs(ref_out).max()
    print(f"forward rel err {rel:.2e}")
    assert rel < 2e-2
    loss, grads = net.gradients(x, y)
    assert loss == pytest.approx(ref_loss, rel=3e-2)
    at = 0
    for k, g_ref in enumerate(ref_grads):
        g = grads[at: at + len(g_ref)]
        at += len(g_ref)
        rel = np.linalg.norm(g - g_ref) / max(np.linalg.norm(g_ref), 1e-12)
        cos = float(g @ g_ref / max(np.linalg.norm(g) * np.linalg.norm(g_ref), 1e-30))
        print(f"param tensor {k}: rel {rel:.2e} cos {cos:.5f}")
        assert rel < 3e-2 and cos > 0.999
    assert at == len(grads)


def test_training_reduces_loss(gpu):
    net = make_net(3)
    x, y = random_inputs(20000, 2)
    print("[gpu] mlp training", flush=True)
    losses = net.fit(x, y, steps=300, batch=1024)
    print(f"loss {losses[:10].mean():.4f} -> {losses[-10:].mean():.4f}")
    assert losses[-10:].mean() < 0.2 * losses[:10].mean()
    pred = net(x[:2000])
    assert np.sqrt(np.mean((pred - y[:2000]) ** 2)) < 0.1
