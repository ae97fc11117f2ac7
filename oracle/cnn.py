"""CPU oracle of the band-proposal network (numpy, float64 arithmetic).

TEST INFRASTRUCTURE ONLY: the checker of the library's BF16 MFMA network
kernels (fbls_cnn_*).  The network has no counterpart in the reference; this
restatement follows the library's definition: `layers` Conv1d layers
(channels-last, zero padding, odd kernel) with GELU, a per-bin linear head,
binary cross-entropy with a positive weight, mean over bins, AdamW.

With bf16=True the values the GPU rounds to bfloat16 (the network input,
every weight and activation entering a GEMM, and the pre-activation
gradients entering the weight/input-gradient GEMMs) are rounded the same way
(round to nearest even); everything else stays exact.  With bf16=False the
arithmetic is plain float64, which the finite-difference test pins.
"""
import math

import numpy as np
from scipy.special import erf


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def im2col(X, kw):
    """X (B, L, C) -> (B*L, kw*C) with column k*C + c = X[b, l + k - kw//2, c]."""
    B, L, C = X.shape
    h = kw // 2
    P = np.zeros((B, L + 2 * h, C))
    P[:, h:h + L] = X
    cols = np.concatenate([P[:, k:k + L] for k in range(kw)], axis=2)
    return cols.reshape(B * L, kw * C)


def col2im(D, B, L, C, kw):
    """Adjoint of im2col."""
    h = kw // 2
    D = D.reshape(B, L, kw, C)
    P = np.zeros((B, L + 2 * h, C))
    for k in range(kw):
        P[:, k:k + L] += D[:, :, k]
    return P[:, h:h + L].reshape(B * L, C)


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / math.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x / math.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / math.sqrt(2 * math.pi)


def softplus(x):
    return np.maximum(x, 0.0) + np.log1p(np.exp(-np.abs(x)))


def sigmoid(x):
    return 1.0 / (1.0 + np.exp(-x))


class Net:
    def __init__(self, in_channels, width, layers, kernel, pos_weight):
        self.cin, self.width, self.layers, self.kw = in_channels, width, layers, kernel
        self.pw = pos_weight
        self.shapes = []
        off = 0
        for l in range(layers):
            c = in_channels if l == 0 else width
            self.shapes.append((off, width, kernel * c, off + width * kernel * c))
            off += width * kernel * c + width
        self.head_w, self.head_b = off, off + width
        self.count = off + width + 1

    def unpack(self, p):
        Ws, bs = [], []
        for off, co, k, boff in self.shapes:
            Ws.append(p[off:off + co * k].reshape(co, k))
            bs.append(p[boff:boff + co])
        return Ws, bs, p[self.head_w:self.head_w + self.width], p[self.head_b]

    def loss_grad(self, p, x, y, bf=True):
        """Mean loss and flat gradient for x (B, L, Cin), y (B, L)."""
        p = np.asarray(p, dtype=np.float64)
        r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
        B, L, _ = x.shape
        Ws, bs, wh, bh = self.unpack(p)
        X = r(x)
        cols, Ys, Hs = [], [], []
        for l in range(self.layers):
            Xc = im2col(X, self.kw)
            Y = (Xc @ r(Ws[l]).T).astype(np.float32 if bf else np.float64) + bs[l]
            H = gelu(Y)
            cols.append(Xc)
            Ys.append(Y)
            Hs.append(r(H))
            X = r(H).reshape(B, L, self.width)
        Hl = Hs[-1]
        z = Hl @ wh + bh
        yy = y.reshape(-1)
        M = z.size
        terms = self.pw * yy * softplus(-z) + (1 - yy) * softplus(z)
        dz = (-self.pw * yy * sigmoid(-z) + (1 - yy) * sigmoid(z)) / M
        g = np.zeros(self.count)
        g[self.head_w:self.head_w + self.width] = Hl.T @ dz
        g[self.head_b] = dz.sum()
        dH = np.outer(dz, wh)
        for l in reversed(range(self.layers)):
            off, co, k, boff = self.shapes[l]
            dY = dH * gelu_grad(Ys[l])
            g[boff:boff + co] = dY.sum(axis=0)
            dYb = r(dY)
            g[off:off + co * k] = (dYb.T @ cols[l]).reshape(-1)
            if l > 0:
                dH = col2im(dYb @ r(Ws[l]), B, L, self.width, self.kw)
        return float(terms.sum() / M), g

    def forward(self, p, x, bf=True):
        p = np.asarray(p, dtype=np.float64)
        r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
        B, L, _ = x.shape
        Ws, bs, wh, bh = self.unpack(p)
        X = r(x)
        for l in range(self.layers):
            Y = (im2col(X, self.kw) @ r(Ws[l]).T).astype(np.float32 if bf else np.float64) + bs[l]
            X = r(gelu(Y)).reshape(B, L, self.width)
        return (X.reshape(B * L, -1) @ wh + bh).reshape(B, L)


def adamw(p, g, m, v, step, lr, wd, decay_mask, b1=0.9, b2=0.999, eps=1e-8):
    """One AdamW step (decoupled decay where decay_mask is True)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    p = p - lr * (mh / (np.sqrt(vh) + eps) + np.where(decay_mask, wd, 0.0) * p)
    return p, m, v
