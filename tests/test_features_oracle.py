"""Pins the oracle's coarse periodogram (oracle/proposal.c) by a direct
numpy restatement: for every period the bin of each point is
floor(t * nb / P) mod nb, every circular window of k bins is summed
point by point, windows with 3 .. n - 3 points and a positive precision
q = w - w^2 / C qualify, and the value is the largest a / sqrt(q)
(-inf when none qualifies).  The oracle picks its phase by the key a|a|/q
and recomputes the SNR there, so the two agree to rounding."""
import math
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402


def _direct(t, v, w, C, periods, durations, nb_min, nb_max):
    d0 = float(np.min(durations))
    out = np.full((len(periods), len(durations)), -np.inf)
    for i, P in enumerate(periods):
        nb = int(min(nb_max, max(nb_min, math.ceil(2.0 * P / d0))))
        scale = nb / P
        b = (np.floor(t * scale).astype(np.int64)) % nb
        for g, D in enumerate(durations):
            k = min(nb, max(1, int(math.floor(D * scale + 0.5))))
            best = -np.inf
            for j in range(nb):
                inside = ((b - j) % nb) < k
                c = int(inside.sum())
                a, ww = float(v[inside].sum()), float(w[inside].sum())
                q = ww - ww * ww / C
                if 3 <= c <= len(t) - 3 and q > 0:
                    best = max(best, a / math.sqrt(q))
            out[i, g] = best
    return out


def test_coarse_features_direct_restatement():
    rng = np.random.default_rng(4)
    n = 600
    t = np.sort(rng.uniform(0.0, 9.0, n))
    w = rng.uniform(0.5, 2.0, n) * 1e6
    y = rng.normal(0, 1e-3, n)
    P0, ep, dur = 2.3, 0.7, 0.15
    y[np.abs(((t - ep + P0 / 2) % P0) - P0 / 2) < dur / 2] += 4e-3  # deficit
    C = float(w.sum())
    v = w * y - w * float((w * y).sum()) / C
    periods = np.array([0.61, 1.15, 2.3, 3.9, 4.6])
    durations = np.array([0.04, 0.09, 0.2])
    got = O.coarse_features(t, v, w, C, periods, durations, 16, 64)
    want = _direct(t, v, w, C, periods, durations, 16, 64)
    fin = np.isfinite(want)
    assert np.array_equal(fin, np.isfinite(got))
    np.testing.assert_allclose(got[fin], want[fin], rtol=1e-12, atol=0)
    # the injected period stands out at its duration
    assert got[2, 1] == got[:, 1].max() and got[2, 1] > 5
