"""GPU: the evidence call replayed through a hipGraph, and the launch-shape
autotuner.

* Graph replays (captured per call, updated into one executable graph)
  interleaved with eager work and other shapes give bit-identical rows and
  tile parts to eager launches and to the oracle, on every replay.
* The autotuner measures every candidate shape, checks that each returns
  the same rows, and leaves the fastest on the context; the search after it
  is bit-identical to the default shape's.
"""
import math

import numpy as np
import pytest

import fbls_amd as F
import oracle as O

pytestmark = pytest.mark.gpu


def _same_rows(a, b):
    assert len(a) == len(b)
    for r, o in zip(a, b):
        for k in r:
            assert r[k] == o[k] or (isinstance(r[k], float) and math.isnan(r[k])
                                    and math.isnan(o[k])), k


def test_evidence_graph_replays(gpu):
    cfg = F.SearchConfig(period_min=0.5, period_max=13.7)
    t, f, e = F.synth_star(seed=1, sectors=1, sigma=1e-3, period=3.7, depth=2e-3,
                           duration=2.5 / 24)
    d = cfg.durations_days()
    s = F.family_state(t, f - 1.0, e, d, cfg.correlated_noise)
    dev = F.DeviceReady.upload(gpu, s)
    vq, wq, _, _ = dev.fields()
    periods = [3.7, 7.4, 3.7, 1.85, 3.7, 11.0, 3.7]
    gpu.set_graphs(False)
    eager = [F.compare_search(gpu, dev, p, d, s.rho, s.residual_energy, cfg, debug=True)
             for p in periods]
    n0 = gpu.set_graphs(True)
    for rep in range(3):
        for p, (er, edbg) in zip(periods, eager):
            print(f"[graph] replay {rep} P={p}", flush=True)
            gr, gdbg = F.compare_search(gpu, dev, p, d, s.rho, s.residual_energy, cfg,
                                        debug=True)
            _same_rows(gr, er)
            np.testing.assert_array_equal(gdbg["screen"], edbg["screen"])
            if er:
                np.testing.assert_array_equal(gdbg["parts"], edbg["parts"])
            # eager work between replays
            F.signal_search(gpu, dev, d, s.rho, F.SearchConfig(period_min=3.0,
                                                               period_max=3.2))
    assert gpu.set_graphs() - n0 == 3 * len(periods)
    orows = O.compare_search(s.t, vq, wq, s.baseline_norm, s.origin, s.residual_energy,
                             d, s.rho, 3.7, cfg)
    _same_rows(eager[0][0], orows)


def test_autotune(gpu):
    cfg = F.SearchConfig(period_min=0.5, period_max=13.7)
    t, f, e = F.synth_star(seed=3, sectors=2, sigma=1e-3, period=5.1, depth=2e-3,
                           duration=3.0 / 24)
    d = cfg.durations_days()
    s = F.family_state(t, f - 1.0, e, d, cfg.correlated_noise)
    dev = F.DeviceReady.upload(gpu, s)
    g2 = F.Gpu(0)
    rows0, _, _ = F.signal_search(g2, dev, d, s.rho, cfg)
    best, table = F.autotune(g2, dev, d, s.rho, cfg, repeats=2)
    print("[tune]", best, table, flush=True)
    assert len(table) >= 10
    assert best["seconds"] == min(e["seconds"] for e in table)
    rows1, _, _ = F.signal_search(g2, dev, d, s.rho, cfg)
    np.testing.assert_array_equal(rows0, rows1)
    g2.close()
