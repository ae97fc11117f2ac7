"""GPU: the end-to-end batch pipeline (fbls_amd.endtoend.run: preparation on
the library's threads, uploads on their own stream, searches on two device
contexts with a chunk ramp) returns, star for star, the best rows of the
same search run directly on the prepared stars -- the full scan and the
learned-pruned search at the committed operating point."""
import numpy as np
import pytest

import fbls_amd as F
from fbls_amd import endtoend as E
from fbls_amd import proposal as PR

pytestmark = pytest.mark.gpu


def _curves(n):
    out = []
    for i in range(n):
        kw = PR.star_params(i, 4_950_000)
        out.append(F.synth_star(**kw))
    return out


def _same_best(got, want):
    assert len(got) == len(want)
    for (p1, s1), (p2, s2) in zip(got, want):
        assert (np.isnan(p1) and np.isnan(p2)) or p1 == p2, (p1, p2)
        assert s1 == s2, (s1, s2)


def test_end_to_end_full_and_pruned_equal_direct_search(gpu):
    pc = PR.ProposalConfig()
    cfg = PR.search_config(pc)
    d = cfg.durations_days()
    curves = _curves(22)
    states = [F.family_state(t, f - 1.0, e, d, cfg.correlated_noise) for t, f, e in curves]
    devs = F.DeviceReady.upload_packed(gpu, states)
    rho = [s.rho for s in states]
    rows, _ = F.signal_search_batch(gpu, devs, d, rho, cfg)
    want_full = PR.best_rows(rows, cfg.threshold)
    net, pc, extra = PR.load(gpu)
    op = (float(extra["tau"]), int(extra["topk"]),
          tuple(float(v) for v in np.atleast_1d(extra["fallback"])))
    z = F.propose(gpu, net, devs, PR.bin_centers(pc), PR.group_durations(cfg, pc.groups),
                  pc.nb_min, pc.nb_max)
    with np.errstate(over="ignore"):
        probs = 1.0 / (1.0 + np.exp(-z))
    bins = PR.keep_bins(probs, pc, op[0], op[1])
    conf = probs.max(axis=1)
    bins[(conf >= op[2][0]) & (conf < op[2][1])] = True
    prow, _ = F.signal_search_bins(gpu, devs, d, rho, cfg, bins, PR.bin_edges(pc))
    want_pruned = PR.best_rows(prow, cfg.threshold)
    for x in devs:
        x.close()
    print(f"[e2e] stars {len(curves)}, full rows {sum(len(r) for r in rows)}", flush=True)
    full = E.run(gpu, curves, cfg, "full", chunk=8)
    _same_best(full["best"], want_full)
    assert full["stars"] == len(curves) and full["device_workers"] == 2
    assert full["stage_seconds"]["prepare"] > 0.0
    print("[e2e] pruned", flush=True)
    pruned = E.run(gpu, curves, cfg, "pruned", net=net, pc=pc, op=op, chunk=8)
    _same_best(pruned["best"], want_pruned)
    assert pruned["trial_evals"] < full["trial_evals"]
