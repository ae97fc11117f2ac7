"""Long-baseline batch: learned-pruned search with a full-scan fallback.

Stars of 1-26 sectors in two seasons around a one-year gap, periods
0.5-100 d (synth.long_baseline_star), searched three ways on one GPU:

* the full scan: fbls_signal_search_batch over every grid period of every
  star (the reference's search, prescan factor 2);
* the learned-pruned scan: the band-proposal network reads a coarse
  periodogram of each star and the bin-restricted batch search
  (fbls_signal_search_bins) visits only the grid periods of the bins it
  proposes; stars whose largest bin probability falls in a fallback band
  [lo, hi) are not trusted to the network and get the full scan.

The coarse periodogram keeps the coherence of the single-sector one the
network was trained on: its bins are log-spaced with a relative width that
shrinks as 1 / baseline (the grid's own period step, grid.rs:12-13, scales
the same way), so that the phase drift across a bin over the whole
baseline is the same as over one sector, and it is computed on the star
binned to 10 minutes (sums of v and w per bin, the bin's mean time), which
leaves the box statistic of every window of >= 30 minutes unchanged up to
the phase smear of one bin.  Stars with the same sector count share their
baseline, hence their bins: they go through the proposal as one group;
the pruned search takes every star with its own bins in one call
(fbls_signal_search_bins_ragged).

Recall is the share of the full scan's recovered injections (transits and
eclipsing-binary pairs, alias rule of tools/benchmark.py:26-27) that the
pruned search also recovers; stars/hour count every second of the pruned
path (binning, upload of the binned stars, periodogram, network, bin
selection, planning and scans).
"""
import concurrent.futures as cf
import math
import os
import time

import numpy as np

from . import (DeviceReady, FamilyState, SearchConfig, coarse_inputs, family_state,
               propose, signal_search_batch, signal_search_bins, signal_search_bins_ragged,
               synth)
from . import proposal as PR

PERIOD_MIN, PERIOD_MAX = 0.5, 100.0
BIN_MINUTES = 10.0


def search_config():
    return SearchConfig(period_min=PERIOD_MIN, period_max=PERIOD_MAX)


def _make_state(args):
    i, seed0, durations, correlated, mix = args
    (t, f, e), kw = synth.long_baseline_star(i, seed0, *mix)
    s = family_state(t, f - 1.0, e, durations, correlated)
    return s, kw.get("period", 0.0), kw["kind"], kw["sectors"]


def _jobs(indices, seed0, mix):
    cfg = search_config()
    return [(i, seed0, cfg.durations_days(), cfg.correlated_noise, mix) for i in indices]


def make_states(indices, seed0, workers=16, mix=(0.30, 0.05), pool=None):
    """[(state, injected period or 0, kind, sectors)] of the given stars
    (mix: the shares of transits and eclipsing binaries)."""
    jobs = _jobs(indices, seed0, mix)
    if pool is not None:
        return list(pool.map(_make_state, jobs, chunksize=2))
    if workers <= 1:
        return [_make_state(j) for j in jobs]
    import multiprocessing as mp

    with cf.ProcessPoolExecutor(workers, mp_context=mp.get_context("spawn")) as ex:
        return list(ex.map(_make_state, jobs, chunksize=2))


def binned_state(s: FamilyState, minutes=BIN_MINUTES):
    """The star summed into consecutive time bins of the given width: v and
    w add (the box statistic reads only their window sums), the bin's time
    is the mean of its points' times."""
    key = np.floor(s.t / (minutes / 1440.0)).astype(np.int64)
    starts = np.flatnonzero(np.r_[True, key[1:] != key[:-1]])
    count = np.diff(np.r_[starts, len(key)])
    t = np.add.reduceat(s.t, starts) / count
    v = np.add.reduceat(s.v, starts)
    w = np.add.reduceat(s.w, starts)
    return FamilyState(t, v, w, s.baseline_norm, s.origin, s.durations, s.rho,
                       s.inflation, s.segments)


def feature_bins(t_last, pc: PR.ProposalConfig, period_min=PERIOD_MIN,
                 period_max=PERIOD_MAX):
    """(centers, edges) of a star's coarse periodogram: log-spaced bins over
    [period_min, min(period_max, baseline)] whose relative width is the
    single-sector network's (pc: bins over [pc.period_min, pc.period_max]
    for a 27.4-d sector) scaled by 27.4 d / baseline.  edges[0] = 0 and
    edges[-1] = inf, as every grid period belongs to some bin."""
    hi = min(period_max, t_last)
    ref = math.log(pc.period_max / pc.period_min) / pc.bins
    step = ref * 27.4 / max(t_last, 27.4)
    lo_l, hi_l = math.log(period_min), math.log(hi)
    nbins = max(1, int(math.ceil((hi_l - lo_l) / step)))
    step = (hi_l - lo_l) / nbins
    centers = np.exp(lo_l + (np.arange(nbins) + 0.5) * step)
    edges = np.exp(lo_l + step * np.arange(nbins + 1))
    edges[0], edges[-1] = 0.0, np.inf
    return centers, edges


def groups_of(states):
    """Star indices grouped by baseline (equal sector counts share it)."""
    out = {}
    for k, s in enumerate(states):
        out.setdefault(float(s.t[-1]), []).append(k)
    return out


def _dilate(sel, dilate):
    out = sel.copy()
    for o in range(1, dilate + 1):
        out[o:] |= sel[:-o]
        out[:-o] |= sel[o:]
    return out


def _top(x, k, sel):
    k = min(k, len(x))
    if k > 0:
        sel[np.argpartition(-x, k - 1)[:k]] = True


def keep_from(probs, dilate, tau, topk=0):
    """Bins with p >= tau plus the topk most probable, dilated by `dilate`
    bins on each side."""
    sel = probs >= tau
    _top(probs, topk, sel)
    return _dilate(sel, dilate)


def operating_point(extra):
    """(tau, top-k, fallback band, raw top-k) of a saved proposal's extras
    (raw top-k 0 when the file has none)."""
    return (float(extra["tau"]), int(extra["topk"]),
            tuple(float(v) for v in np.atleast_1d(extra["fallback"])),
            int(extra["raw_topk"]) if "raw_topk" in extra else 0)


def keep_mask(probs, peaks, dilate, op):
    """The bins the pruned search scans at operating point op = (tau, top-k,
    (lo, hi), raw top-k): the network's (p >= tau and its top-k most
    probable) and the coarse periodogram's raw top-k highest-SNR bins, all
    dilated by `dilate` bins; every bin when the network's largest
    probability lies in the fallback band [lo, hi).  The raw bins catch
    strong signals the network ranks low (a short-period eclipsing binary
    whose subharmonic it prefers); they cost a few bins per star."""
    tau, topk, (lo, hi) = op[:3]
    raw = op[3] if len(op) > 3 else 0
    if lo <= float(probs.max()) < hi:
        return np.ones(len(probs), dtype=bool)
    sel = probs >= tau
    _top(probs, topk, sel)
    if raw and peaks is not None:
        _top(np.where(np.isfinite(peaks), peaks, -np.inf), raw, sel)
    return _dilate(sel, dilate)


class LongBatch:
    """A set of long-baseline stars resident on the GPU with their full-scan
    results; evaluates the learned-pruned search at operating points."""

    def __init__(self, gpu, res, net, pc, scan_chunk=32, log=None):
        self.gpu, self.net, self.pc = gpu, net, pc
        self.cfg = search_config()
        self.states = [r[0] for r in res]
        self.injected = [r[1] for r in res]
        self.kinds = [r[2] for r in res]
        self.sectors = [r[3] for r in res]
        self.rhos = [s.rho for s in self.states]
        self.d = self.cfg.durations_days()
        self.devs = DeviceReady.upload_packed(gpu, self.states)
        self.groups = groups_of(self.states)
        # full scan (warm the staging buffers on the first chunk first)
        signal_search_batch(gpu, self.devs[:2], self.d, self.rhos[:2], self.cfg)
        gpu.launch_totals(reset=True)
        t0 = time.perf_counter()
        best, templates, periods = [], 0, 0
        for c0 in range(0, len(self.devs), scan_chunk):
            rows, stats = signal_search_batch(gpu, self.devs[c0:c0 + scan_chunk], self.d,
                                              self.rhos[c0:c0 + scan_chunk], self.cfg)
            best += PR.best_rows(rows, self.cfg.threshold)
            templates += sum(st["templates_scanned"] for st in stats)
            periods += sum(st["periods"] for st in stats)
            if log:
                log(f"full scan {min(len(self.devs), c0 + scan_chunk)} of {len(self.devs)} "
                    f"stars ({time.perf_counter() - t0:.1f} s)")
        self.t_full = time.perf_counter() - t0
        self.full_totals = gpu.launch_totals(reset=True)
        self.templates_full, self.periods_full = templates, periods
        self.best_full = best
        self.rf = PR.recovered(best, self.injected)

    def proposal(self):
        """Per star: bin probabilities, the coarse periodogram's peak SNRs
        and the bin edges; the seconds of binning, upload and the
        device-resident periodogram + network."""
        gd = PR.group_durations(self.cfg, self.pc.groups)
        t0 = time.perf_counter()
        binned = [binned_state(s) for s in self.states]
        bdevs = DeviceReady.upload_packed(self.gpu, binned)
        n = len(self.states)
        probs, peaks, edges = [None] * n, [None] * n, [None] * n
        for t_last, idx in self.groups.items():
            centers, e = feature_bins(t_last, self.pc)
            per = max(1, (1 << 24) // len(centers))  # network batch limit
            for c0 in range(0, len(idx), per):
                part = idx[c0:c0 + per]
                z, pk = propose(self.gpu, self.net, [bdevs[k] for k in part], centers, gd,
                                self.pc.nb_min, self.pc.nb_max, peaks=True)
                with np.errstate(over="ignore"):  # exp overflow is p = 0
                    p = 1.0 / (1.0 + np.exp(-z))
                for j, k in enumerate(part):
                    probs[k], peaks[k], edges[k] = p[j], pk[j], e
        dt = time.perf_counter() - t0
        for b in bdevs:
            b.close()
        return probs, peaks, edges, dt

    def pruned(self, probs, peaks, edges, t_prop, op):
        """The bin-restricted search of every star at operating point op
        (keep_mask), timed from bin selection to the rows."""
        lo, hi = op[2]
        t0 = time.perf_counter()
        keeps, nfall = [], 0
        for k in range(len(self.states)):
            keeps.append(keep_mask(probs[k], peaks[k], self.pc.keep_dilate, op))
            nfall += int(lo <= float(probs[k].max()) < hi)
        # every star in one call (its own bins and edges): the same merged
        # launches for the whole batch
        rows, stats = signal_search_bins_ragged(self.gpu, self.devs, self.d, self.rhos,
                                                self.cfg, keeps, edges)
        best = PR.best_rows(rows, self.cfg.threshold)
        templates = sum(st["templates_scanned"] for st in stats)
        periods = sum(st["periods"] for st in stats)
        t_pruned = time.perf_counter() - t0
        rp = PR.recovered(best, self.injected)
        n = len(self.states)
        eb = np.array([k == "eb" for k in self.kinds])
        return {
            "stars": n, "tau": op[0], "topk": op[1], "fallback": list(op[2]),
            "raw_topk": op[3] if len(op) > 3 else 0,
            "full_recovered": int(self.rf.sum()),
            "pruned_recovered": int((self.rf & rp).sum()),
            "recall": float((self.rf & rp).sum() / max(self.rf.sum(), 1)),
            "eb_full_recovered": int((self.rf & eb).sum()),
            "eb_pruned_recovered": int((self.rf & rp & eb).sum()),
            "fallback_stars": nfall,
            "kept_period_fraction": periods / max(self.periods_full, 1),
            "periods_full": self.periods_full,
            "templates_full": self.templates_full, "templates_pruned": templates,
            "seconds_full_scan": self.t_full, "seconds_proposal": t_prop,
            "seconds_pruned_scan": t_pruned,
            "stars_per_hour_full": n * 3600.0 / self.t_full,
            "stars_per_hour_pruned": n * 3600.0 / (t_prop + t_pruned),
            "speedup": self.t_full / (t_prop + t_pruned),
        }

    def sweep(self, points, log=None):
        probs, peaks, edges, t_prop = self.proposal()
        out = []
        for op in points:
            out.append(self.pruned(probs, peaks, edges, t_prop, op))
            if log:
                r = out[-1]
                log(f"operating point {op}: recall {r['recall']:.4f} "
                    f"speedup {r['speedup']:.2f}")
        return out

    def summary(self):
        inj = np.array([p > 0 for p in self.injected])
        return {
            "stars": len(self.states),
            "sectors_min": int(min(self.sectors)), "sectors_max": int(max(self.sectors)),
            "points_max": int(max(len(s.t) for s in self.states)),
            "transits": int(sum(k == "transit" for k in self.kinds)),
            "eclipsing_binaries": int(sum(k == "eb" for k in self.kinds)),
            "full_recovered_of_injected": [int(self.rf.sum()), int(inj.sum())],
            "full_scan_seconds": self.t_full,
            "full_scan_stars_per_hour": len(self.states) * 3600.0 / self.t_full,
            "full_scan_trial_evals_per_s": self.templates_full / self.t_full,
            "full_scan_periods": self.periods_full,
            "full_scan_kernel_seconds": self.full_totals["seconds"],
            "full_scan_launches": self.full_totals["launches"],
            "full_scan_fold_pairs_per_kernel_s": (self.full_totals["fold_pairs"]
                                                  / self.full_totals["seconds"]),
        }

    def close(self):
        for d in self.devs:
            d.close()


# operating points: (tau, top-k, fallback band, raw top-k); the validation
# sweep tries those without a fallback band (a star with one is rescanned in
# full, which costs 2-4x the pruned search), the held-out report adds one
# with it
POINTS = [(tau, k, (0.0, 0.0), raw) for tau in (0.02, 0.05, 0.1, 0.2) for k in (0, 8)
          for raw in (0, 8)]
FALLBACK_POINT = (0.05, 8, (0.005, 0.1), 8)
# the coarse periodogram's raw top-k bins alone (the network keeps no bin):
# what the network adds over the method without it
RAW_ONLY = [(1.01, 0, (0.0, 0.0), 8), (1.01, 0, (0.0, 0.0), 32)]
WEIGHTS = os.path.join(PR.HERE, "data", "proposal_long.npz")
REPORT = os.path.join(PR.HERE, "data", "longbase_report.json")
ALIASES = (1 / 3, 0.5, 1.0, 2.0, 3.0)


def merge_sweeps(parts):
    """Per operating point, the sums over disjoint sets of stars."""
    out = []
    for j in range(len(parts[0])):
        rs = [p[j] for p in parts]
        m = {k: rs[0][k] for k in ("tau", "topk", "fallback", "raw_topk")}
        assert all({k: r[k] for k in m} == m for r in rs), "different operating points"
        for k in ("stars", "full_recovered", "pruned_recovered", "eb_full_recovered",
                  "eb_pruned_recovered", "fallback_stars", "templates_full",
                  "templates_pruned"):
            m[k] = int(sum(r[k] for r in rs))
        for k in ("seconds_full_scan", "seconds_proposal", "seconds_pruned_scan"):
            m[k] = float(sum(r[k] for r in rs))
        m["recall"] = m["pruned_recovered"] / max(m["full_recovered"], 1)
        m["periods_full"] = int(sum(r["periods_full"] for r in rs))
        m["kept_period_fraction"] = float(sum(r["kept_period_fraction"] * r["periods_full"]
                                              for r in rs) / m["periods_full"])
        m["stars_per_hour_full"] = m["stars"] * 3600.0 / m["seconds_full_scan"]
        m["stars_per_hour_pruned"] = m["stars"] * 3600.0 / (m["seconds_proposal"]
                                                            + m["seconds_pruned_scan"])
        m["speedup"] = m["stars_per_hour_pruned"] / m["stars_per_hour_full"]
        out.append(m)
    return out


def merge_summaries(parts):
    """The full-scan summary of disjoint sets of stars."""
    m = {"stars": sum(p["stars"] for p in parts),
         "sectors_min": min(p["sectors_min"] for p in parts),
         "sectors_max": max(p["sectors_max"] for p in parts),
         "points_max": max(p["points_max"] for p in parts)}
    for k in ("transits", "eclipsing_binaries", "full_scan_periods", "full_scan_launches"):
        m[k] = sum(p[k] for p in parts)
    m["full_recovered_of_injected"] = [sum(p["full_recovered_of_injected"][i] for p in parts)
                                       for i in (0, 1)]
    for k in ("full_scan_seconds", "full_scan_kernel_seconds"):
        m[k] = sum(p[k] for p in parts)
    m["full_scan_stars_per_hour"] = m["stars"] * 3600.0 / m["full_scan_seconds"]
    m["full_scan_trial_evals_per_s"] = sum(
        p["full_scan_trial_evals_per_s"] * p["full_scan_seconds"] for p in parts
    ) / m["full_scan_seconds"]
    m["full_scan_fold_pairs_per_kernel_s"] = sum(
        p["full_scan_fold_pairs_per_kernel_s"] * p["full_scan_kernel_seconds"] for p in parts
    ) / m["full_scan_kernel_seconds"]
    return m


def training_set(gpu, pc, indices, seed0, pool, log, crop=2048, negatives=2,
                 mix=(0.6, 0.1), chunk=96, seed=0):
    """Network inputs and labels for fine-tuning on long-baseline stars:
    crops of `crop` bins of each star's coarse periodogram (one around its
    label, `negatives` at random offsets).  Labels come from the GPU scan
    itself, restricted to the bins around the injected period and its
    aliases (P/3 ... 3P): the best row >= threshold there that is an alias
    of the injection marks its bin +- label_dilate (the single-sector
    labels mark the full scan's best row the same way)."""
    cfg = search_config()
    d = cfg.durations_days()
    gd = PR.group_durations(cfg, pc.groups)
    rng = np.random.default_rng(seed)
    xs, ys = [], []
    spans = [indices[c:c + chunk] for c in range(0, len(indices), chunk)]

    def submit(span):
        return pool.map(_make_state, _jobs(span, seed0, mix), chunksize=2)

    pending = submit(spans[0])
    for k, span in enumerate(spans):
        res = list(pending)
        pending = submit(spans[k + 1]) if k + 1 < len(spans) else None
        states = [r[0] for r in res]
        devs = DeviceReady.upload_packed(gpu, states)
        bdevs = DeviceReady.upload_packed(gpu, [binned_state(st) for st in states])
        for t_last, idx in groups_of(states).items():
            centers, edges = feature_bins(t_last, pc)
            x = np.empty((len(idx), len(centers), pc.groups), dtype=np.float32)
            per = max(1, (1 << 24) // len(centers))
            for c0 in range(0, len(idx), per):
                part = idx[c0:c0 + per]
                x[c0:c0 + len(part)] = coarse_inputs(gpu, [bdevs[j] for j in part], centers,
                                                     gd, pc.nb_min, pc.nb_max)
            nbin = len(centers)
            keep = np.zeros((len(idx), nbin), dtype=bool)
            for j, i in enumerate(idx):
                p = res[i][1]
                if p > 0:
                    for m in ALIASES:
                        b = int(np.searchsorted(edges, p * m, side="right")) - 1
                        if 0 <= b < nbin:
                            keep[j, max(0, b - 3):b + 4] = True
            rows, _ = signal_search_bins(gpu, [devs[i] for i in idx], d,
                                         [states[i].rho for i in idx], cfg, keep, edges)
            best = PR.best_rows(rows, cfg.threshold)
            for j, i in enumerate(idx):
                y = np.zeros(nbin, dtype=np.float32)
                p_inj, (p_best, _) = res[i][1], best[j]
                lab = -1
                if p_inj > 0 and p_best == p_best and PR.alias(p_best, p_inj) < 1e-3:
                    lab = int(np.searchsorted(edges, p_best, side="right")) - 1
                    y[max(0, lab - pc.label_dilate):lab + pc.label_dilate + 1] = 1.0
                starts = [int(rng.integers(0, nbin - crop + 1)) for _ in range(negatives)]
                if lab >= 0:  # a crop holding the label bin and its dilation
                    lo = min(nbin - crop, max(0, lab - crop + 1 + pc.label_dilate))
                    hi = max(lo, min(nbin - crop, lab - pc.label_dilate))
                    starts.append(int(rng.integers(lo, hi + 1)))
                for s0 in starts:
                    xs.append(x[j, s0:s0 + crop].copy())
                    ys.append(y[s0:s0 + crop].copy())
        for dv in devs + bdevs:
            dv.close()
        log(f"training stars {span[-1] + 1 - indices[0]} of {len(indices)}, crops {len(xs)}, "
            f"positive crops {int(sum(v.max() > 0 for v in ys))}")
    return np.stack(xs), np.stack(ys)


def main():
    import argparse
    import json
    import multiprocessing as mp

    from . import Gpu

    ap = argparse.ArgumentParser(description="long-baseline learned-pruned batch")
    ap.add_argument("--train", type=int, default=0,
                    help="fine-tune the single-sector network on this many stars first")
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--validate", type=int, default=320)
    ap.add_argument("--heldout", type=int, default=480)
    ap.add_argument("--weights", default=WEIGHTS,
                    help="the fine-tuned network to evaluate when not training")
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--report", default=None)
    ap.add_argument("--workers", type=int, default=16)
    ap.add_argument("--skip-fallback", action="store_true",
                    help="leave the fallback-band point out of the held-out sweep")
    ap.add_argument("--heldout-only", action="store_true",
                    help="keep the operating point of --weights and the validation figures "
                         "of --prior; measure the held-out stars again")
    ap.add_argument("--prior", default=REPORT,
                    help="the report whose training and validation figures are kept")
    ap.add_argument("--heldout-start", type=int, default=0,
                    help="first held-out star (seed 5,000,000 + index)")
    ap.add_argument("--short-sweep", action="store_true",
                    help="held-out sweep of the operating point, the raw-only points and "
                         "the fallback point only")
    ap.add_argument("--merge", nargs="+", default=None,
                    help="combine held-out reports of consecutive star ranges into --report")
    args = ap.parse_args()

    def log(msg):
        print(f"[longbase] {msg}", flush=True)

    if args.merge:
        parts = [json.load(open(p)) for p in args.merge]
        report = dict(parts[0])
        report["heldout"] = merge_summaries([r["heldout"] for r in parts])
        report["heldout_sweep"] = merge_sweeps([r["heldout_sweep"] for r in parts])
        report["heldout_at_operating_point"] = report["heldout_sweep"][0]
        report["heldout_stars"] = report["heldout"]["stars"]
        report["heldout_parts"] = [r.get("heldout_range") for r in parts]
        with open(args.report, "w") as f:
            json.dump(report, f, indent=1)
        log(json.dumps(report["heldout_at_operating_point"]))
        return

    pool = cf.ProcessPoolExecutor(args.workers, mp_context=mp.get_context("spawn"))
    gpu = Gpu(0)
    t0 = time.perf_counter()
    train_info = {}
    if args.train:
        net, pc, _ = PR.load(gpu)  # start from the single-sector network
        x, y = training_set(gpu, pc, list(range(args.train)), 5_300_000, pool, log)
        log(f"training crops {len(x)} ({time.perf_counter() - t0:.1f} s)")
        t1 = time.perf_counter()
        losses = PR.train(net, x, y, epochs=args.epochs, lr=args.lr, seed=1, log=log)
        train_info = {"train_stars": args.train, "train_crops": int(len(x)),
                      "positive_crops": int((y.max(axis=1) > 0).sum()),
                      "epochs": args.epochs, "lr": args.lr,
                      "train_seconds": time.perf_counter() - t1,
                      "final_loss": float(np.mean(losses[-50:]))}
        del x, y
        PR.save(net, pc, args.out)
        log(f"fine-tuned {train_info}")
    else:
        net, pc, _ = PR.load(gpu, args.weights)
        if os.path.exists(args.prior):  # the training of the evaluated network
            train_info = json.load(open(args.prior)).get("training", {})
    prior = json.load(open(args.prior)) if os.path.exists(args.prior) else {}
    if args.heldout_only:  # the committed operating point, chosen on validation
        _, _, extra = PR.load(gpu, args.weights)
        op = operating_point(extra)
        args.validate = 0
    vres = make_states(range(args.validate), 5_100_000, pool=pool) if args.validate else []
    span = range(args.heldout_start, args.heldout_start + args.heldout)
    t_prep = time.perf_counter()
    hres = make_states(span, 5_000_000, pool=pool) if args.heldout else []
    pool.shutdown()
    log(f"prepared {len(vres)} + {len(hres)} stars ({time.perf_counter() - t_prep:.1f} s "
        "for the held-out ones)")
    val_summary = prior.get("validation_full_scan")
    if args.heldout_only:
        point = prior.get("validation_at_operating_point")
        vs = prior.get("validation_sweep")
    else:
        val = LongBatch(gpu, vres, net, pc, log=log)
        val_summary = val.summary()
        log(f"validation full scan {val_summary}")
        vs = val.sweep(POINTS, log=lambda m: log("validation " + m))
        val.close()
        point = PR.fastest(vs, 0.99)
        if point is None:  # none reaches 99 %: the most cautious of the best-recall points
            top = max(r["recall"] for r in vs)
            point = max((r for r in vs if r["recall"] == top),
                        key=lambda r: r["kept_period_fraction"])
        op = (point["tau"], point["topk"], tuple(point["fallback"]), point["raw_topk"])
    summary, hs = None, [None]
    if hres:  # --heldout 0: select on validation only
        held = LongBatch(gpu, hres, net, pc, log=log)
        summary = held.summary()
        log(f"held-out full scan {summary}")
        extra = ([] if args.skip_fallback else [FALLBACK_POINT]) + RAW_ONLY
        rest = extra if args.short_sweep else POINTS + extra
        hs = held.sweep([op] + [p for p in rest if p != op],
                        log=lambda m: log("held-out " + m))
        held.close()
    report = {"training": train_info,
              "selection_rule": ("the fastest point whose validation recall reaches 99 %; "
                                 "when none does, the most cautious (largest kept fraction) "
                                 "of the points with the best validation recall"),
              "operating_point": {"tau": op[0], "topk": op[1], "fallback": list(op[2]),
                                  "raw_topk": op[3]},
              "validation_stars": prior.get("validation_stars") if args.heldout_only
              else len(vres),
              "validation_at_operating_point": point,
              "validation_full_scan": val_summary,
              "validation_sweep": vs,
              "heldout": summary, "heldout_at_operating_point": hs[0],
              "heldout_sweep": hs,
              "heldout_range": [span.start, span.stop] if hres else None}
    log(json.dumps(hs[0] if hres else point))
    PR.save(net, pc, args.out, tau=op[0], topk=op[1], fallback=np.array(op[2], dtype=float),
            raw_topk=op[3])
    if args.report:
        with open(args.report, "w") as f:
            json.dump(report, f, indent=1)


if __name__ == "__main__":
    main()
