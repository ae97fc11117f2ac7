"""Pins the CPU oracle of the band-proposal network (oracle/cnn.py): its
gradient against central finite differences in float64, im2col / col2im as
adjoints, the bf16 rounding rule, and the first AdamW step's closed form."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from oracle import cnn as C  # noqa: E402


def _tiny(seed=0):
    rng = np.random.default_rng(seed)
    net = C.Net(in_channels=3, width=4, layers=3, kernel=5, pos_weight=7.0)
    p = rng.normal(0, 0.4, net.count)
    x = rng.normal(0, 1, (2, 11, 3))
    y = (rng.uniform(size=(2, 11)) < 0.2).astype(float)
    return net, p, x, y


def test_gradient_matches_finite_differences():
    net, p, x, y = _tiny()
    _, g = net.loss_grad(p, x, y, bf=False)
    h = 1e-6
    rng = np.random.default_rng(1)
    for i in rng.choice(net.count, 60, replace=False):
        e = np.zeros(net.count)
        e[i] = h
        lp, _ = net.loss_grad(p + e, x, y, bf=False)
        lm, _ = net.loss_grad(p - e, x, y, bf=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_im2col_col2im_adjoint():
    rng = np.random.default_rng(2)
    X = rng.normal(size=(3, 9, 4))
    D = rng.normal(size=(27, 7 * 4))
    lhs = float((C.im2col(X, 7) * D).sum())
    rhs = float((X.reshape(27, 4) * C.col2im(D, 3, 9, 4, 7)).sum())
    assert abs(lhs - rhs) < 1e-12 * max(1.0, abs(lhs))


def test_bf16_rounding():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.14159265, 1e-30],
                    dtype=np.float32)
    r = C.bf16(vals)
    assert r[0] == 1.0 and r[3] == -2.5
    assert r[1] == 1.0  # tie to even (mantissa 0)
    assert r[2] == 1.0 + 2 ** -7  # above the tie rounds up
    assert abs(r[4] - 3.140625) < 1e-12
    bits = r.astype(np.float32).view(np.uint32)
    assert np.all(bits & 0xFFFF == 0)


def test_adamw_first_step():
    p = np.array([0.5, -1.0, 2.0])
    g = np.array([0.1, -0.2, 0.0])
    q, m, v = C.adamw(p, g, np.zeros(3), np.zeros(3), 1, lr=0.01, wd=0.1,
                      decay_mask=np.array([True, True, False]))
    # step 1: m_hat = g, v_hat = g^2
    want = p - 0.01 * (g / (np.abs(g) + 1e-8) + np.array([0.1, 0.1, 0.0]) * p)
    np.testing.assert_allclose(q, want, rtol=1e-12)
