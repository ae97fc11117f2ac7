"""GPU: the band-proposal network kernels (BF16 MFMA GEMMs, im2col/col2im,
GELU, head, AdamW) against the numpy oracle with the same bf16 rounding.
FP32 accumulation order differs from the oracle's float64, and a flipped
bf16 rounding propagates, so the bar is 1e-2 relative on every gradient
tensor (norm-wise) and on the logits."""
import numpy as np
import pytest

import fbls_amd as F
from oracle import cnn as C

pytestmark = pytest.mark.gpu


def _data(B, L, cin, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1, (B, L, cin)).astype(np.float32)
    y = (x[:, :, 0] > 1.5).astype(np.float32)
    return x, y


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("shape", [(2, 64, 3, 16, 2, 3), (4, 256, 8, 64, 6, 7)])
def test_forward_and_gradients_match_oracle(gpu, shape):
    B, L, cin, width, layers, kw = shape
    print(f"[cnn] shape {shape}", flush=True)
    net = F.Cnn(gpu, in_channels=cin, width=width, layers=layers, kernel=kw,
                pos_weight=5.0, seed=3)
    orc = C.Net(cin, width, layers, kw, 5.0)
    p = net.params().astype(np.float64)
    assert p.size == orc.count
    x, y = _data(B, L, cin, 7)
    z = net.forward(x)
    zo = orc.forward(p, x, bf=True)
    assert _rel(z, zo) < 1e-2, _rel(z, zo)
    loss, g = net.loss_grad(x, y)
    lo, go = orc.loss_grad(p, x, y, bf=True)
    assert abs(loss - lo) <= 1e-3 * abs(lo)
    for l, (off, co, k, boff) in enumerate(orc.shapes):
        assert _rel(g[off:off + co * k], go[off:off + co * k]) < 1e-2, ("W", l)
        assert _rel(g[boff:boff + co], go[boff:boff + co]) < 1e-2, ("b", l)
    hw = slice(orc.head_w, orc.head_w + width)
    assert _rel(g[hw], go[hw]) < 1e-2
    assert abs(g[orc.head_b] - go[orc.head_b]) <= 1e-2 * abs(go[orc.head_b])


def test_adamw_step_matches_oracle_and_is_deterministic(gpu):
    x, y = _data(2, 128, 8, 11)
    a = F.Cnn(gpu, in_channels=8, width=32, layers=3, kernel=7, seed=5)
    b = F.Cnn(gpu, in_channels=8, width=32, layers=3, kernel=7, seed=5)
    orc = C.Net(8, 32, 3, 7, 20.0)
    p0 = a.params().astype(np.float64)
    _, g = a.loss_grad(x, y)
    a.set_params(p0)  # resets the optimiser
    a.train_step(x, y, lr=1e-3, weight_decay=0.1)
    mask = np.ones(orc.count, dtype=bool)
    for off, co, k, boff in orc.shapes:
        mask[boff:boff + co] = False
    mask[orc.head_b] = False
    want, _, _ = C.adamw(p0, g.astype(np.float64), np.zeros(orc.count), np.zeros(orc.count),
                         1, lr=1e-3, wd=0.1, decay_mask=mask)
    np.testing.assert_allclose(a.params(), want, rtol=1e-5, atol=1e-7)
    for _ in range(3):
        a.train_step(x, y)
    b.train_step(x, y, lr=1e-3, weight_decay=0.1)
    for _ in range(3):
        b.train_step(x, y)
    np.testing.assert_array_equal(a.params(), b.params())


def test_training_reduces_loss(gpu):
    x, y = _data(8, 512, 8, 13)
    net = F.Cnn(gpu, in_channels=8, width=64, layers=6, kernel=7, pos_weight=5.0, seed=2)
    first = net.train_step(x, y, lr=2e-3)
    for _ in range(60):
        last = net.train_step(x, y, lr=2e-3)
    print("[cnn] loss", first, "->", last, flush=True)
    assert last < 0.3 * first


@pytest.mark.parametrize("L", [2048, 200])
def test_fused_inference_is_bit_identical(gpu, L):
    # the proposal network's shape: 8 inputs, 64 channels, 6 layers, kernel 7;
    # L = 200 leaves a partial 128-bin tile at each sample's end
    print(f"[cnn] fused inference L={L}", flush=True)
    net = F.Cnn(gpu, in_channels=8, width=64, layers=6, kernel=7, seed=9)
    x, _ = _data(3, L, 8, 13)
    fused = net.forward(x)
    net.set_fused(False)
    plain = net.forward(x)
    net.set_fused(True)
    np.testing.assert_array_equal(fused, plain)
    orc = C.Net(8, 64, 6, 7, 20.0)
    assert _rel(fused, orc.forward(net.params().astype(np.float64), x, bf=True)) < 1e-2
