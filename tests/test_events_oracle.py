"""Pins the single-event half of the oracle (oracle/events.c).

* single_event_search: a brute force over every (phase bin, duration,
  shape) template that sums the explicit staircase template h(t) = m(t) / L
  (m = number of nested layer windows holding the point) against the
  fields, with circular windows taken modulo nb: same survivors in the same
  order, scores and depths to 1e-12.  This pins the doubled prefix, the
  nested-interval staircase formula (kernels.rs:84-113), the usability
  rules and the record decoding (events.rs:137-151).
* The statistics: templates_scanned = nb * nd * ns, the bin count formula
  and the overflow rule.
* select_events: a direct Python restatement of peaks.rs:490-550 on random
  tables with prior masks.
* The search: a deep single dip beside a periodic planet is reported as a
  single event at its time and duration while the planet stays a peak.
"""
import math
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import fbls_amd as F  # noqa: E402
import oracle as O  # noqa: E402


def _quantised(state):
    return O.quantized(state["v"]), O.quantized(state["w"])


def _family(seed=3, n=400, span=4.0, dip=(2.1, 0.12, 6e-3), sigma=1e-3):
    rng = np.random.default_rng(seed)
    t = np.sort(rng.uniform(0.0, span, n)) + 2458000.0
    f = rng.normal(0.0, sigma, n)
    t0, d, depth = dip
    f[np.abs(t - 2458000.0 - t0) <= d / 2] -= depth
    e = np.full(n, sigma)
    return t, f, e


def _brute(t, v, w, C, origin, durations, rho, ingress, layers, thr, po):
    n = len(t)
    trial = t[-1] + 2 * durations[-1] + 1e-6
    nb = max(8, math.ceil(po * trial / durations[0]))
    scale = nb / trial
    b = np.floor(t * scale).astype(np.int64) % nb
    out = []
    L = float(layers)
    for j in range(nb):
        for di, D in enumerate(durations):
            raw_k = max(1, int(math.floor(D * scale + 0.5)))
            k = min(raw_k, nb // 2)
            inbox = ((b - j) % nb) < k
            cnt = int(inbox.sum())
            if not (raw_k < nb // 2 and 3 <= cnt <= n - 3):
                continue
            for ing in ingress:
                if ing == 0.0:
                    h = inbox.astype(float)
                else:
                    m = np.zeros(n)
                    for layer in range(layers):
                        fac = 1 - 2 * ing * (layer / (layers - 1) if layers > 1 else 0.0)
                        width = max(1, int(math.floor(k * fac + 0.5)))
                        start = j + (k - width) // 2
                        m += ((b - start) % nb) < width
                    h = m / L
                a = float(np.sum(h * v))
                hp = float(np.sum(h * w))
                norm = float(np.sum(h * h * w))
                q = norm - hp * hp / C
                if not (a > 0 and q > 0):
                    continue
                s = (a / rho[di]) / math.sqrt(q)
                if s >= thr:
                    out.append(((j + k / 2) / scale + origin, k / scale, s, a / q))
    return out, nb


def test_single_event_scan_matches_brute_force():
    t, f, e = _family()
    durations = np.array([0.05, 0.08, 0.12])
    st = O.family_state(t, f, e, durations, correlated=False)
    v, w = _quantised(st)
    cfg = F.SearchConfig(threshold=3.0, ingress_fractions=[0.0, 0.2, 0.45], shape_layers=4,
                         phase_oversample=4)
    rho = np.array([1.0, 1.1, 1.3])
    rows, stats = O.single_event_search(st["t"], v, w, st["baseline_norm"], st["origin"],
                                        durations, rho, cfg)
    want, nb = _brute(st["t"], v, w, st["baseline_norm"], st["origin"], durations, rho,
                      cfg.ingress_fractions, 4, 3.0, 4)
    assert stats["phase_bins"] == nb
    assert stats["templates_scanned"] == nb * 3 * 3
    assert stats["threshold_survivors"] == len(want) == len(rows) > 10
    for r, (ep, du, s, dep) in zip(rows, want):
        assert r["period"] == math.inf
        assert r["epoch"] == ep and r["duration"] == du
        assert abs(r["score"] - s) <= 1e-12 * abs(s)
        assert abs(r["depth"] - dep) <= 1e-12 * abs(dep)
    # the strongest survivor sits on the dip
    best = max(rows, key=lambda r: r["score"])
    assert abs(best["epoch"] - (2458000.0 + 2.1)) < 0.06
    # unsorted durations give the same rows (sorted with their rho)
    perm = [2, 0, 1]
    rows2, _ = O.single_event_search(st["t"], v, w, st["baseline_norm"], st["origin"],
                                     durations[perm], rho[perm], cfg)
    assert rows2 == rows


def test_single_event_capacity_overflow():
    t, f, e = _family(seed=4)
    d = np.array([0.05, 0.1])
    st = O.family_state(t, f, e, d, correlated=False)
    v, w = _quantised(st)
    cfg = F.SearchConfig(threshold=1.0)
    rows, stats = O.single_event_search(st["t"], v, w, st["baseline_norm"], st["origin"], d,
                                        np.ones(2), cfg, capacity=5)
    assert stats["overflow"] and stats["threshold_survivors"] > 5 and rows == []


def _select_py(table, time, thr, limit, taken):
    masks = [np.asarray(m, bool) for m in taken]
    rem = [r["score"] if math.isfinite(r["score"]) and r["score"] >= thr else -math.inf
           for r in table]
    out, it = [], 0
    time = np.asarray(time)
    while len(out) < limit and it < 5000 and table:
        it += 1
        i = max(range(len(table)), key=lambda a: (rem[a], a))  # last of equal maxima
        if not math.isfinite(rem[i]):
            break
        row = table[i]
        inside = np.abs(time - row["epoch"]) <= row["duration"] / 2
        for j, o in enumerate(table):
            if abs(o["epoch"] - row["epoch"]) <= (o["duration"] + row["duration"]) / 2:
                rem[j] = -math.inf
        pts = int(inside.sum())
        if pts < 3 or any((inside & m).sum() / pts >= 0.5 for m in masks):
            continue
        masks.append(np.abs(time - row["epoch"]) <= 0.75 * row["duration"])
        out.append(dict(epoch_bjd=row["epoch"], duration_hours=24 * row["duration"],
                        snr=row["score"], depth=row["depth"], depth_ppm=1e6 * row["depth"],
                        points=pts))
    return out


def test_select_events_restatement():
    rng = np.random.default_rng(5)
    time = np.sort(rng.uniform(0, 20, 3000))
    for trial in range(6):
        m = int(rng.integers(1, 400))
        table = [dict(epoch=float(rng.uniform(0, 20)), duration=float(rng.uniform(0.02, 0.4)),
                      score=float(rng.choice([rng.uniform(0, 20), 9.0])),
                      depth=float(rng.uniform(1e-4, 1e-2))) for _ in range(m)]
        taken = [np.abs(time - rng.uniform(0, 20)) < rng.uniform(0.1, 1.0) for _ in range(2)]
        got = O.select_events(table, time, 7.1, 20, taken)
        want = _select_py(table, time, 7.1, 20, taken)
        assert got == want, trial


def test_search_reports_single_dip_beside_planet():
    t, f, e = F.synth_star(seed=21, sectors=1, sigma=8e-4, period=3.7, depth=2.5e-3,
                           duration=2.5 / 24)
    tc, dd = t[0] + 11.37, 5.0 / 24
    f = f - 4e-3 * (np.abs(t - tc) <= dd / 2)
    cfg = F.SearchConfig(period_min=1.0, period_max=8.0, passes=1, max_candidates=3,
                         evidence_candidates=1)
    peaks, passes, ev = O.search_prepared(t, f - 1.0, e, cfg, with_events=True)
    assert any(abs(p["period_days"] / 3.7 - 1) < 1e-3 for p in peaks)
    assert len(ev["event_stats"]) == 1 and not ev["event_stats"][0]["overflow"]
    assert ev["events"], ev
    top = ev["events"][0]
    assert abs(top["epoch_bjd"] - tc) < dd / 2 and abs(top["duration_hours"] - 5.0) < 1.5
    assert top["points"] >= 3
    # events are ordered by the greedy selection: non-increasing SNR
    s = [x["snr"] for x in ev["events"]]
    assert s == sorted(s, reverse=True)
    cfg.single_events = False
    _, _, ev0 = O.search_prepared(t, f - 1.0, e, cfg, with_events=True)
    assert ev0["events"] == [] and ev0["event_stats"] == []
