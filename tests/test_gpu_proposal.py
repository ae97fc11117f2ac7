"""GPU: the pieces of the learned band proposal.

* fbls_coarse_features against the oracle (bit-exact: exact fixed-point
  sums, the same FP64 expressions).
* A keep-masked batch search returns exactly the full search's rows at the
  kept periods, also when nearly every period passes the prescan (the
  restricted search's box-first prescan).
* With the committed trained network: on fresh held-out stars (seeds never
  used in training), the pruned search keeps >= 95 % of the full scan's
  recovered injections at the saved operating point, and the network
  proposes better than ranking bins by the raw coarse SNR at the same
  top-k (the method without the network).
"""
import os

import numpy as np
import pytest

import fbls_amd as F
import oracle as O
from fbls_amd import proposal as PR

pytestmark = pytest.mark.gpu


def test_coarse_features_bit_exact(gpu):
    pc = PR.ProposalConfig()
    cfg = PR.search_config(pc)
    res = PR.make_states(range(3), 4_900_000, cfg, workers=1)
    states = [r[0] for r in res]
    devs = F.DeviceReady.upload_packed(gpu, states)
    periods = PR.bin_centers(pc)[::37]
    gd = PR.group_durations(cfg, pc.groups)
    got = F.coarse_features(gpu, devs, periods, gd, pc.nb_min, pc.nb_max)
    for k, (s, dev) in enumerate(zip(states, devs)):
        vq, wq, _, _ = dev.fields()
        want = O.coarse_features(s.t, vq, wq, s.baseline_norm, periods, gd, pc.nb_min,
                                 pc.nb_max)
        np.testing.assert_array_equal(got[k], want)
    # the network input computed on the device is the host transform of them
    inputs = F.coarse_inputs(gpu, devs, periods, gd, pc.nb_min, pc.nb_max)
    np.testing.assert_array_equal(inputs, PR.transform(got))
    # the device-resident proposal equals the network run on those inputs
    net = F.Cnn(gpu, in_channels=pc.groups, width=pc.width, layers=pc.layers,
                kernel=pc.kernel, seed=3)
    logits = F.propose(gpu, net, devs, periods, gd, pc.nb_min, pc.nb_max)
    np.testing.assert_array_equal(logits, net.forward(inputs))


def test_keep_mask_rows_are_full_rows(gpu):
    pc = PR.ProposalConfig()
    cfg = PR.search_config(pc)
    res = PR.make_states(range(4), 4_900_100, cfg, workers=1)
    states = [r[0] for r in res]
    devs = F.DeviceReady.upload_packed(gpu, states)
    d = cfg.durations_days()
    rows, _ = F.signal_search_batch(gpu, devs, d, [s.rho for s in states], cfg)
    grids = PR.grids_of(devs, cfg)
    rng = np.random.default_rng(0)
    keeps = [rng.uniform(size=len(g)) < 0.3 for g in grids]
    prow, pstats = F.signal_search_batch(gpu, devs, d, [s.rho for s in states], cfg,
                                         keep=keeps)
    for k in range(len(states)):
        kept = set(grids[k][keeps[k]].tolist())
        want = np.array([r for r in rows[k] if r[0] in kept]).reshape(-1, 5)
        np.testing.assert_array_equal(prow[k], want)
        assert pstats[k]["periods"] == int(keeps[k].sum())
    # bin restriction: the rows of the grid periods inside the kept bins
    edges = PR.bin_edges(pc)
    bins = rng.uniform(size=(len(states), pc.bins)) < 0.05
    bins[1] = True  # one star keeps every bin: its full search
    brow, bstats = F.signal_search_bins(gpu, devs, d, [s.rho for s in states], cfg, bins,
                                        edges)
    for k in range(len(states)):
        b = np.searchsorted(edges, grids[k], side="right") - 1
        inside = bins[k][b]
        kept = set(grids[k][inside].tolist())
        want = np.array([r for r in rows[k] if r[0] in kept]).reshape(-1, 5)
        np.testing.assert_array_equal(brow[k], want)
        assert bstats[k]["periods"] == int(inside.sum())
    np.testing.assert_array_equal(brow[1], rows[1])


@pytest.mark.skipif(not os.path.exists(PR.DEFAULT_WEIGHTS), reason="no trained proposal")
def test_box_first_prescan_keeps_full_rows(gpu):
    """At a low threshold nearly every period passes the prescan.  A search
    restricted to every period (boxes first in the prescan; the staircase
    templates only for periods whose best box stays below the cut) returns
    the full search's rows and prescan survivor counts."""
    pc = PR.ProposalConfig()
    cfg = F.SearchConfig(period_min=pc.period_min, period_max=pc.period_max, threshold=3.0)
    res = PR.make_states(range(3), 4_900_200, cfg, workers=1)
    states = [r[0] for r in res]
    devs = F.DeviceReady.upload_packed(gpu, states)
    d = cfg.durations_days()
    rho = [s.rho for s in states]
    rows, st = F.signal_search_batch(gpu, devs, d, rho, cfg)
    keeps = [np.ones(len(g), bool) for g in PR.grids_of(devs, cfg)]
    krow, kst = F.signal_search_batch(gpu, devs, d, rho, cfg, keep=keeps)
    for k in range(len(states)):
        print(f"[box-first] star {k}: rows {len(rows[k])}, periods {st[k]['periods']}, "
              f"survivors {st[k]['prescan_survivors']}", flush=True)
        np.testing.assert_array_equal(krow[k], rows[k])
        assert kst[k]["prescan_survivors"] == st[k]["prescan_survivors"]
        assert st[k]["prescan_survivors"] >= 0.5 * st[k]["periods"]
    for x in devs:
        x.close()


def test_trained_proposal_recall_and_gain(gpu):
    net, pc, extra = PR.load(gpu)
    cfg = PR.search_config(pc)
    res = PR.make_states(range(1024), 4_950_000, cfg, workers=8)
    ev = PR.EvalSet(gpu, [r[0] for r in res], [r[1] for r in res], pc)
    topk, tau = int(extra["topk"]), float(extra["tau"])
    fb = tuple(float(v) for v in np.atleast_1d(extra["fallback"]))
    sw = ev.sweep(net, points=[(tau, topk, fb), (tau, topk, (0.0, 0.0))])
    ev.close()
    at = sw["learned"][0]
    print("[proposal]", at, flush=True)
    # 20,000 held-out stars: 99.9 % recall; ~430 recoveries here
    assert at["full_recovered"] >= 300
    assert at["recall"] >= 0.99
    assert at["stars_per_hour_pruned"] >= 5 * at["stars_per_hour_full"]
    learned0, raw0 = sw["learned"][1], sw["raw_snr_topk"][1]
    print("[proposal] learned", learned0["recall"], "raw", raw0["recall"], flush=True)
    # the network against the raw periodogram's top bins at the same kept share
    assert learned0["recall"] >= raw0["recall"] + 0.1


def test_coarse_features_unequal_weights_bit_exact(gpu):
    """The periodogram of stars whose weights differ point to point (the
    general three-field fold) equals the oracle, like that of stars with
    equal weights."""
    import dataclasses

    pc = PR.ProposalConfig()
    cfg = PR.search_config(pc)
    res = PR.make_states(range(2), 4_900_200, cfg, workers=1)
    rng = np.random.default_rng(5)
    states = [r[0] for r in res]
    states.append(dataclasses.replace(states[0], w=states[0].w * rng.uniform(0.6, 1.5, len(states[0].w))))
    devs = F.DeviceReady.upload_packed(gpu, states)
    periods = PR.bin_centers(pc)[::53]
    gd = PR.group_durations(cfg, pc.groups)
    got = F.coarse_features(gpu, devs, periods, gd, pc.nb_min, pc.nb_max)
    for k, (s, dev) in enumerate(zip(states, devs)):
        vq, wq, _, _ = dev.fields()
        assert (len(np.unique(wq)) == 1) == (k < 2)
        want = O.coarse_features(s.t, vq, wq, s.baseline_norm, periods, gd, pc.nb_min,
                                 pc.nb_max)
        np.testing.assert_array_equal(got[k], want)
