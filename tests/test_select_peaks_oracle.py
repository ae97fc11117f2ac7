"""The oracle's greedy peak selection against an independent restatement.

`select_peaks` (isakruas/fbls src/select/peaks.rs:308-487) is restated here
in Python from the reference's definitions -- the BinaryHeap order (larger
score, then smaller row), lazy invalidation, the 8192-row refresh blocks of
stale rows, the rescoring of `Rescorer::evaluate` (peaks.rs:148-240, with
`count_rows` / `rescore_rows`, src/search/kernels.rs:934-1042), the overlap
and `min_events` rules, `without_strongest_event` (peaks.rs:243-299) and the
0.5-margin exclusion masks -- and compared with the oracle's C selection
(`orc_select_peaks`) on constructed tables: the accepted rows, their order,
isolated and leftover statistics, event counts and the final exclusion flags
must be identical.  The fields are fixed-point values, so every sum is exact
in any order and the two sides agree bit for bit.  The tables include exact
score ties at the top, prior masks and exclusions from an earlier pass, and
a cluster of more than 8192 rows around an accepted ephemeris, so that one
acceptance needs a second refresh block before a fresh row reaches the top.
"""
import heapq

import numpy as np
import pytest

import oracle as O

BLOCK = 8192          # peaks.rs:383
MAX_ITERATIONS = 20000
MARGIN = 0.5


# ------------------------------------------------------------ restatement

def rem_euclid(a, b):
    r = np.fmod(a, b)
    return np.where(r < 0.0, r + abs(b), r)


def round_half_away(x):
    """f64::round: halves away from zero (x - floor(x) is exact)."""
    ax = np.abs(x)
    f = np.floor(ax)
    return (np.sign(x) * (f + (ax - f >= 0.5))).astype(np.int64)


def transit_mask(time, period, epoch, duration, margin=0.0):
    phase = rem_euclid(time - epoch + period / 2.0, period) - period / 2.0
    return np.abs(phase) <= (1.0 + margin) * duration / 2.0


def overlap(a, b):
    total = int(a.sum())
    return 0.0 if total == 0 else int((a & b).sum()) / total


def event_count(time, period, epoch, duration, minimum_points=3):
    inside = transit_mask(time, period, epoch, duration)
    cycles = round_half_away((time[inside] - epoch) / period)
    _, counts = np.unique(cycles, return_counts=True)
    return int((counts >= minimum_points).sum())


def nearest_rho(fam, duration):
    best = 0
    for i in range(1, len(fam["dur"])):
        if abs(fam["dur"][i] - duration) < abs(fam["dur"][best] - duration):
            best = i
    return fam["rho"][best]


def box_phase(time, period, epoch, duration):
    """The rescoring kernels' in-transit test (kernels.rs:952-955)."""
    x = (time - epoch) * (1.0 / period) + 0.5
    return np.abs((x - np.floor(x) - 0.5) * period) <= duration / 2.0


def evaluate(idx, rows, family, fams, pos, time_all, excluded, threshold):
    out = []
    for i in idx:
        r = rows[i]
        P, E, D = r["period"], r["epoch"], r["duration"]
        inside = box_phase(time_all, P, E, D)
        total = max(int(inside.sum()), 1)
        unshared = int((inside & ~excluded).sum()) / total
        required = max(threshold * np.sqrt(unshared), 0.5 * threshold)
        f = fams[family[i]]
        x = (f["t"] + f["origin"] - E) * (1.0 / P) + 0.5
        keep = ~excluded[pos[family[i]]] & (np.abs((x - np.floor(x) - 0.5) * P) <= D / 2.0)
        a, hp = f["v"][keep].sum(), f["w"][keep].sum()
        norm = hp                      # the diagonal model: norm sums w as hp does
        q = norm - hp * hp / f["C"]
        score = (a / nearest_rho(f, D)) / np.sqrt(max(q, 1e-300)) if (q > 0 and a > 0) else -np.inf
        out.append((score, required))
    return out


def without_strongest_event(f, pos, r, excluded):
    ta = f["t"] + f["origin"]
    P, E, D = r["period"], r["epoch"], r["duration"]
    phase = rem_euclid(ta - E + P / 2.0, P) - P / 2.0
    member = ~excluded[pos] & (np.abs(phase) <= D / 2.0)
    if not member.any():
        return -np.inf
    cycles = round_half_away((ta[member] - E) / P)
    v, w = f["v"][member], f["w"][member]
    rho = nearest_rho(f, D)
    strongest, best = 0, -np.inf
    for c in np.unique(cycles):            # ascending, as the BTreeMap
        m = cycles == c
        a, hp = v[m].sum(), w[m].sum()
        s = a / rho / np.sqrt(max(hp - hp * hp / f["C"], 1e-300))
        if s > best:
            strongest, best = c, s
    rest = cycles != strongest
    a, hp = v[rest].sum(), w[rest].sum()
    q = hp - hp * hp / f["C"]
    return a / rho / np.sqrt(max(q, 1e-300)) if (q > 0 and a > 0) else -np.inf


def select_peaks_restated(fams, rows, family, time_all, threshold, limit, min_events,
                          masks=(), excluded=None):
    n_all = len(time_all)
    excluded = np.zeros(n_all, bool) if excluded is None else np.array(excluded, bool)
    masks = [np.array(m, bool) for m in masks]
    pos = [np.minimum(np.searchsorted(time_all, f["t"] + f["origin"], "left"), n_all - 1)
           for f in fams]
    remaining = [r["score"] if (np.isfinite(r["score"]) and r["score"] >= threshold)
                 else -np.inf for r in rows]
    fresh = [not excluded.any()] * len(rows)
    heap = [(-v, i) for i, v in enumerate(remaining) if np.isfinite(v)]
    heapq.heapify(heap)
    trace = {"refreshes": 0, "full_blocks": 0, "back_to_back": 0}
    last_was_refresh = [False]

    def refresh():
        trace["refreshes"] += 1
        stale, kept = [], []
        while heap:
            nv, i = heapq.heappop(heap)
            if -nv != remaining[i]:
                continue
            if fresh[i]:
                kept.append((nv, i))
            else:
                stale.append(i)
                if len(stale) == BLOCK:
                    trace["full_blocks"] += 1
                    break
        for e in kept:
            heapq.heappush(heap, e)
        for i, (score, required) in zip(stale, evaluate(stale, rows, family, fams, pos,
                                                        time_all, excluded, threshold)):
            remaining[i] = score if score >= required else -np.inf
            fresh[i] = True
            if np.isfinite(remaining[i]):
                heapq.heappush(heap, (-remaining[i], i))

    accepted = []
    iterations = 0
    while len(accepted) < limit and iterations < MAX_ITERATIONS:
        iterations += 1
        while heap and -heap[0][0] != remaining[heap[0][1]]:
            heapq.heappop(heap)
        if not heap:
            break
        i = heap[0][1]
        if not fresh[i]:
            if last_was_refresh[0]:
                trace["back_to_back"] += 1
            refresh()
            last_was_refresh[0] = True
            continue
        last_was_refresh[0] = False
        heapq.heappop(heap)
        isolated = remaining[i]
        remaining[i] = -np.inf
        r = rows[i]
        mask = transit_mask(time_all, r["period"], r["epoch"], r["duration"])
        if any(overlap(mask, m) >= 0.5 for m in masks):
            continue
        events = event_count(time_all, r["period"], r["epoch"], r["duration"], 3)
        if events < min_events:
            continue
        k = family[i]
        leftover = without_strongest_event(fams[k], pos[k], r, excluded)
        if min_events > 1 and leftover < 0.5 * threshold:
            continue
        wide = transit_mask(time_all, r["period"], r["epoch"], r["duration"], MARGIN)
        excluded |= wide
        masks.append(wide)
        accepted.append({"row": i, "snr": isolated, "snr_folded": r["score"],
                         "snr_without_strongest_event": leftover, "events": events,
                         "family": k, "period_days": r["period"], "epoch_bjd": r["epoch"],
                         "duration_hours": 24.0 * r["duration"], "depth": r["depth"]})
        fresh = [False] * len(rows)
        refresh()
        last_was_refresh[0] = True
    return accepted, excluded, trace


# ------------------------------------------------------------------ data

EPHEMERIDES = ((3.1, 0.7, 0.12), (5.7, 2.2, 0.2), (9.3, 4.1, 0.16))  # P, epoch offset, D


def problem(seed, n_rows, cluster, ties=3):
    """Two families over one 30-min cadence curve (the second a subset), with
    fixed-point fields carrying three boxes, and a table whose top rows are
    a tight cluster around the first box (with exact score ties)."""
    rng = np.random.default_rng(seed)
    t0 = 2458000.0
    time_all = t0 + np.arange(0.0, 27.4, 1.0 / 48.0)
    time_all = time_all[(time_all - t0 < 13.2) | (time_all - t0 > 14.2)]
    n_all = len(time_all)
    subset = np.sort(rng.choice(n_all, size=int(0.85 * n_all), replace=False))
    fams = []
    for k, idx in enumerate((np.arange(n_all), subset)):
        ta = time_all[idx]
        # precision p and deficit y = noise + boxes; fields v = p y and w = p
        # in fixed point (power-of-two scales, as the library uploads them)
        p = rng.uniform(0.25, 0.5, size=len(ta))
        y = rng.normal(0.0, 1.0, size=len(ta))
        for (P, off, D), depth in zip(EPHEMERIDES, (3.0, 2.5, 3.5)):
            y[transit_mask(ta, P, t0 + off, D)] += depth
        w = np.ldexp(np.rint(np.ldexp(p, 30)), -30)
        v = np.ldexp(np.rint(np.ldexp(p * y, 26)), -26)
        fams.append({"t": ta - ta[0], "origin": ta[0], "v": v, "w": w, "C": float(w.sum()),
                     "dur": np.array([0.06, 0.12, 0.2]),
                     "rho": np.array([1.0, 1.05, 1.2]) * (1.0 + 0.1 * k)})
    rows, family = [], []

    def add(P, E, D, score, fam):
        rows.append({"period": float(P), "epoch": float(E), "duration": float(D),
                     "score": float(score), "depth": float(rng.uniform(1e-4, 1e-2))})
        family.append(fam)

    P1, off1, D1 = EPHEMERIDES[0]
    top = 40.0
    for m in range(cluster):
        score = top if m < ties else rng.uniform(20.0, 39.0)
        add(P1 * (1 + rng.normal(0, 2e-5)), t0 + off1 + rng.normal(0, 0.01),
            D1 * rng.uniform(0.8, 1.2), score, int(rng.integers(0, 2)))
    for P, off, D in EPHEMERIDES[1:]:
        for _ in range(60):
            add(P * (1 + rng.normal(0, 1e-5)), t0 + off + rng.normal(0, 0.005),
                D * rng.uniform(0.9, 1.1), rng.uniform(10.0, 19.0), int(rng.integers(0, 2)))
    while len(rows) < n_rows:
        P = rng.uniform(0.5, 12.0)
        add(P, t0 + rng.uniform(0, P), rng.choice([0.06, 0.12, 0.2]),
            rng.uniform(6.0, 25.0), int(rng.integers(0, 2)))
    rows[-1]["score"] = np.nan                       # a non-finite row never enters
    perm = rng.permutation(len(rows))                # rows in no particular order
    return (fams, [rows[i] for i in perm], [family[i] for i in perm], time_all)


def check(fams, rows, family, time_all, threshold, limit, min_events, masks=(), excluded=None):
    got, ex = O.select_peaks(fams, rows, family, time_all, threshold, limit, min_events,
                             masks, excluded)
    want, wex, trace = select_peaks_restated(fams, rows, family, time_all, threshold, limit,
                                             min_events, masks, excluded)
    assert [p["row"] for p in got] == [p["row"] for p in want]
    for g, w in zip(got, want):
        for key, value in w.items():
            assert g[key] == value, (key, g[key], value)
        assert g["depth_ppm"] == 1e6 * w["depth"]
    np.testing.assert_array_equal(ex, wex)
    return want, trace


@pytest.mark.parametrize("min_events", [1, 2, 3])
def test_select_peaks_matches_restatement(min_events):
    fams, rows, family, time_all = problem(11 + min_events, 1500, 400)
    want, trace = check(fams, rows, family, time_all, 7.1, 6, min_events)
    assert len(want) >= 1
    # exact ties at the top: the smallest row index among them is taken first
    top_rows = sorted(i for i, r in enumerate(rows) if r["score"] == 40.0)
    assert want[0]["row"] == top_rows[0]


def test_select_peaks_with_prior_masks_and_exclusions():
    """A later pass: masks and exclusion flags from earlier acceptances make
    every row stale from the start."""
    fams, rows, family, time_all = problem(5, 1200, 300)
    P, off, D = EPHEMERIDES[2]
    prior = transit_mask(time_all, P, 2458000.0 + off, D, MARGIN)
    want, trace = check(fams, rows, family, time_all, 7.1, 5, 2, masks=[prior], excluded=prior)
    assert trace["refreshes"] >= 1
    assert all(abs(p["period_days"] - P) > 0.01 for p in want)


def test_select_peaks_second_refresh_block():
    """More than 8192 cluster rows above every other row: after the first
    acceptance the first refresh block rescores only cluster rows, so the top
    of the heap is still stale and a second block is rescored before the
    next acceptance."""
    fams, rows, family, time_all = problem(3, 14000, 11000)
    want, trace = check(fams, rows, family, time_all, 7.1, 3, 1)
    assert trace["full_blocks"] >= 1 and trace["back_to_back"] >= 1, trace
    assert len(want) >= 2
