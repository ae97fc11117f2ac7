"""Learned band proposal: which period bins get the full-resolution scan.

A small 1-D CNN (fbls_cnn_*) reads the coarse periodogram of a star
(fbls_coarse_features: 2048 log-spaced period bins x 8 representative
durations, the largest box SNR over phase) and outputs, per bin, the
probability that the full scan's best detection lies there.  The pruned
search scans only the grid periods of the proposed bins, each exactly as the
full search does (fbls_signal_search_batch with a keep mask), so its rows
are the full search's rows at those periods; stars whose proposal is not
confident fall back to the full scan.  Labels come from the full GPU scan
itself on seeded injection stars; recall is measured against the full
scan's recovered injections with the alias rule of the reference's
benchmark (min over m in {1/3, 1/2, 1, 2, 3} of |P / P_inj - m| / m < 1e-3,
tools/benchmark.py:26-27, src/validate.rs:41-49).

Usage (on a GPU):
    python -m fbls_amd.proposal --train 200000 --skip-eval --out w.npz
    python -m fbls_amd.proposal --reselect --validate 2000 --heldout 20000 \\
        --out fbls_amd/data/proposal.npz --report report.json
(the second reads the weights and training figures from fbls_amd/data)
"""
import argparse
import concurrent.futures as cf
import dataclasses
import json
import math
import os
import time

import numpy as np

from . import (SearchConfig, Cnn, DeviceReady, Gpu, coarse_inputs, family_state, propose,
               signal_search_batch, signal_search_bins, synth_star)

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "data", "proposal.npz")


@dataclasses.dataclass
class ProposalConfig:
    period_min: float = 0.5
    period_max: float = 13.7
    bins: int = 2048
    groups: int = 8
    width: int = 64
    layers: int = 6
    kernel: int = 7
    pos_weight: float = 20.0
    label_dilate: int = 2     # bins on each side of the detection's bin
    keep_dilate: int = 2      # bins added on each side of a proposed bin
    nb_min: int = 16
    nb_max: int = 1024


def search_config(pc: ProposalConfig):
    return SearchConfig(period_min=pc.period_min, period_max=pc.period_max)


def group_durations(cfg: SearchConfig, k):
    d = cfg.durations_days()
    return np.geomspace(d.min(), d.max(), k)


def bin_centers(pc: ProposalConfig):
    lo, hi = math.log(pc.period_min), math.log(pc.period_max)
    step = (hi - lo) / pc.bins
    return np.exp(lo + (np.arange(pc.bins) + 0.5) * step)


def bin_of(periods, pc: ProposalConfig):
    lo, hi = math.log(pc.period_min), math.log(pc.period_max)
    b = np.floor((np.log(np.asarray(periods)) - lo) / (hi - lo) * pc.bins)
    return np.clip(b, 0, pc.bins - 1).astype(np.int64)


def bin_edges(pc: ProposalConfig):
    """Period edges of the bins (bin k: edges[k] <= P < edges[k + 1]); the
    outer edges are open, as bin_of clips periods outside the range."""
    lo, hi = math.log(pc.period_min), math.log(pc.period_max)
    e = np.exp(lo + (hi - lo) * np.arange(pc.bins + 1) / pc.bins)
    e[0], e[-1] = 0.0, np.inf
    return e


def keep_bins(probs, pc: ProposalConfig, tau, topk):
    """Per star: bins with p >= tau plus the topk most probable, dilated by
    keep_dilate bins on each side (vectorised over stars)."""
    sel = probs >= tau
    if topk:
        idx = np.argpartition(-probs, topk - 1, axis=1)[:, :topk]
        np.put_along_axis(sel, idx, True, axis=1)
    out = sel.copy()
    for o in range(1, pc.keep_dilate + 1):
        out[:, o:] |= sel[:, :-o]
        out[:, :-o] |= sel[:, o:]
    return out


def transform(raw):
    """Network input from the raw coarse SNRs (no qualifying phase -> 0)."""
    x = np.where(np.isfinite(raw), raw, 0.0)
    return (np.clip(x, -5.0, 40.0) / 8.0).astype(np.float32)


def alias(p, cat):
    return min(abs(p / cat - m) / m for m in (1 / 3, 0.5, 1.0, 2.0, 3.0))


# ---------------------------------------------------------------- stars

def star_params(i, seed0):
    """Single-sector injection star i (the batched configuration's laws with
    one sector): sigma log-uniform 200-3000 ppm, half with a box of period
    log-uniform 0.5-13.7 d, depth log-uniform 200-10000 ppm, duration
    1 h + 7 h (P^(1/3) - 0.5^(1/3)) / (20^(1/3) - 0.5^(1/3))."""
    rng = np.random.default_rng(seed0 + i)
    sigma = float(np.exp(rng.uniform(np.log(200e-6), np.log(3000e-6))))
    kw = dict(seed=seed0 + i, sectors=1, sector_days=27.4, gap_days=1.0,
              cadence_days=1.0 / 720.0, t0=2458000.0, sigma=sigma)
    if rng.uniform() < 0.5:
        p = float(np.exp(rng.uniform(np.log(0.5), np.log(13.7))))
        depth = float(np.exp(rng.uniform(np.log(200e-6), np.log(1e-2))))
        dur = (1.0 + 7.0 * (p ** (1 / 3) - 0.5 ** (1 / 3))
               / (20 ** (1 / 3) - 0.5 ** (1 / 3))) / 24.0
        kw.update(period=p, depth=depth, duration=dur)
    return kw


def _make_state(args):
    i, seed0, durations, correlated = args
    kw = star_params(i, seed0)
    t, f, e = synth_star(**kw)
    s = family_state(t, f - 1.0, e, durations, correlated)
    return s, kw.get("period", 0.0)


def make_states(indices, seed0, cfg: SearchConfig, workers=16):
    d = cfg.durations_days()
    jobs = [(i, seed0, d, cfg.correlated_noise) for i in indices]
    if workers <= 1:
        return [_make_state(j) for j in jobs]
    import multiprocessing as mp

    with cf.ProcessPoolExecutor(workers, mp_context=mp.get_context("spawn")) as ex:
        return list(ex.map(_make_state, jobs, chunksize=32))


# ------------------------------------------------------------ labels, data

def best_rows(rows_per_star, threshold):
    """(period, score) of each star's best row, or (nan, -inf)."""
    out = []
    for r in rows_per_star:
        if len(r) and np.max(r[:, 3]) >= threshold:
            k = int(np.argmax(r[:, 3]))
            out.append((float(r[k, 0]), float(r[k, 3])))
        else:
            out.append((float("nan"), float("-inf")))
    return out


def labels_from(best, pc: ProposalConfig):
    y = np.zeros((len(best), pc.bins), dtype=np.float32)
    for s, (p, _) in enumerate(best):
        if p == p:
            b = int(bin_of([p], pc)[0])
            y[s, max(0, b - pc.label_dilate):b + pc.label_dilate + 1] = 1.0
    return y


def grids_of(devs, cfg: SearchConfig):
    from . import lib

    d = cfg.durations_days()
    dmin = float(d.min())
    out = []
    for dev in devs:
        hi = min(cfg.period_max, dev.t_last)
        lo = max(cfg.period_min, 4.0 * dmin * (1.0 + 1e-9))
        n = lib().fbls_period_grid_count(lo, hi, dev.t_last, dmin, cfg.period_oversample)
        g = np.empty(n)
        from . import _check, _dp
        _check(lib().fbls_period_grid(lo, hi, dev.t_last, dmin, cfg.period_oversample,
                                      _dp(g), n))
        out.append(g)
    return out


def chunk_data(gpu, states, pc: ProposalConfig, full=True, timed=False):
    """Features (and the full scan's rows when full) of a chunk of stars.
    timed: run each once untimed first, so that the times exclude the
    one-off growth of the library's staging buffers."""
    cfg = search_config(pc)
    devs = DeviceReady.upload_packed(gpu, states)
    gd = group_durations(cfg, pc.groups)
    centers = bin_centers(pc)
    if timed:
        coarse_inputs(gpu, devs, centers, gd, pc.nb_min, pc.nb_max)
    t0 = time.perf_counter()
    x = coarse_inputs(gpu, devs, centers, gd, pc.nb_min, pc.nb_max)
    t_feat = time.perf_counter() - t0
    rows = stats = None
    t_full = 0.0
    if full:
        rhos = [s.rho for s in states]
        if timed:
            signal_search_batch(gpu, devs, cfg.durations_days(), rhos, cfg)
        t0 = time.perf_counter()
        rows, stats = signal_search_batch(gpu, devs, cfg.durations_days(), rhos, cfg)
        t_full = time.perf_counter() - t0
    return devs, x, rows, stats, t_feat, t_full


# ---------------------------------------------------------------- proposal

def keep_masks(probs, grids, pc: ProposalConfig, tau, topk):
    """Per star: proposed bins = {p >= tau} plus the topk most probable,
    dilated by keep_dilate; keep = grid periods inside proposed bins."""
    keeps = []
    for s, g in enumerate(grids):
        p = probs[s]
        sel = p >= tau
        if topk:
            sel[np.argpartition(-p, topk - 1)[:topk]] = True
        if pc.keep_dilate:
            d = pc.keep_dilate
            idx = np.nonzero(sel)[0]
            grow = np.zeros_like(sel)
            for o in range(-d, d + 1):
                grow[np.clip(idx + o, 0, pc.bins - 1)] = True
            sel = grow
        keeps.append(sel[bin_of(g, pc)])
    return keeps


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def predict(net, x, batch=512):
    """Bin probabilities (float32) of the network for inputs x."""
    out = np.empty(x.shape[:2], dtype=np.float32)
    with np.errstate(over="ignore"):  # exp overflow is p = 0
        _predict_into(net, x, batch, out)
    return out


def _predict_into(net, x, batch, out):
    for i in range(0, len(x), batch):
        z = net.forward(x[i:i + batch])
        np.negative(z, out=z)
        np.exp(z, out=z)
        z += 1.0
        np.reciprocal(z, out=out[i:i + batch])


def train(net, x, y, epochs=8, batch=32, lr=2e-3, weight_decay=1e-4, seed=0, log=None):
    rng = np.random.default_rng(seed)
    n = len(x)
    steps = max(1, epochs * (n // batch))
    losses = []
    for step in range(steps):
        idx = rng.choice(n, batch, replace=False)
        # cosine decay of the learning rate
        rate = lr * 0.5 * (1 + math.cos(math.pi * step / steps))
        losses.append(net.train_step(x[idx], y[idx], lr=rate, weight_decay=weight_decay))
        if log and (step % 200 == 0 or step == steps - 1):
            log(f"step {step}/{steps} loss {np.mean(losses[-50:]):.5f}")
    return losses


def recovered(best, injected):
    return np.array([(p == p and inj > 0 and alias(p, inj) < 1e-3)
                     for (p, _), inj in zip(best, injected)])


# operating points: (tau, top-k bins, fallback band): a star keeps its bins
# with p >= tau plus its top-k bins; a star whose largest bin probability
# lies in [lo, hi) is uncertain and gets the full scan
POINTS = [(tau, k, fb) for tau in (0.05, 0.1, 0.15, 0.2) for k in (0, 4, 8)
          for fb in ((0.0, 0.0), (0.02, 0.1))]


class EvalSet:
    """A set of stars with their full-scan results, features and grids."""

    def __init__(self, gpu, states, injected, pc: ProposalConfig):
        self.gpu, self.states, self.pc = gpu, states, pc
        self.cfg = search_config(pc)
        self.injected = injected
        (self.devs, self.x, rows, stats, self.t_feat,
         self.t_full) = chunk_data(gpu, states, pc, full=True, timed=True)
        self.best_full = best_rows(rows, self.cfg.threshold)
        self.templates_full = int(sum(s["templates_scanned"] for s in stats))
        self.grids = grids_of(self.devs, self.cfg)
        self.edges = bin_edges(pc)
        self.rf = recovered(self.best_full, injected)

    def proposal(self, net):
        """Bin probabilities of the network with the periodogram computed
        and consumed on the device (fbls_propose), and the seconds taken."""
        gd = group_durations(self.cfg, self.pc.groups)
        centers = bin_centers(self.pc)
        propose(self.gpu, net, self.devs[:2], centers, gd, self.pc.nb_min, self.pc.nb_max)
        t0 = time.perf_counter()
        z = propose(self.gpu, net, self.devs, centers, gd, self.pc.nb_min, self.pc.nb_max)
        with np.errstate(over="ignore"):  # exp overflow is p = 0
            probs = 1.0 / (1.0 + np.exp(-z))
        return probs, time.perf_counter() - t0

    def pruned(self, probs, t_net, tau, topk, fallback):
        """The pruned search from the probabilities: bin selection (with the
        fallback stars searched in full) and the bin-restricted batch search,
        both timed."""
        t0 = time.perf_counter()
        conf = probs.max(axis=1)
        lo, hi = fallback
        fall = (conf >= lo) & (conf < hi)
        bins = keep_bins(probs, self.pc, tau, topk)
        bins[fall] = True
        prow, pstats = signal_search_bins(self.gpu, self.devs, self.cfg.durations_days(),
                                          [s.rho for s in self.states], self.cfg, bins,
                                          self.edges)
        t_pruned = time.perf_counter() - t0
        rp = recovered(best_rows(prow, self.cfg.threshold), self.injected)
        kept = sum(int(st["periods"]) for st in pstats)
        total = sum(len(g) for g in self.grids)
        n = len(self.states)
        return {
            "stars": n, "topk": topk, "fallback": fallback, "tau": tau,
            "full_recovered": int(self.rf.sum()),
            "pruned_recovered": int((self.rf & rp).sum()),
            "recall": float((self.rf & rp).sum() / max(self.rf.sum(), 1)),
            "fallback_stars": int(fall.sum()),
            "kept_period_fraction": kept / max(total, 1),
            "templates_full": self.templates_full,
            "templates_pruned": int(sum(s["templates_scanned"] for s in pstats)),
            "seconds_full_scan": self.t_full,
            "seconds_proposal": t_net,
            "seconds_pruned_scan": t_pruned,
            "stars_per_hour_full": n * 3600.0 / self.t_full,
            "stars_per_hour_pruned": n * 3600.0 / (t_net + t_pruned),
        }

    def sweep(self, net, points=POINTS):
        """Every operating point for the learned proposal, and beside each
        the method without the network: every star's top bins by raw coarse
        SNR, as many per star as the learned rule keeps on average."""
        probs, t_net = self.proposal(net)
        raw = self.x.max(axis=2).astype(np.float64)
        raw = raw / max(float(raw.max()), 1e-9)
        out = {"learned": [], "raw_snr_topk": []}
        raw_at = {}
        for tau, topk, fb in points:
            r = self.pruned(probs, t_net, tau, topk, fb)
            out["learned"].append(r)
            sel = (probs >= tau).sum(axis=1)
            k_raw = max(1, int(round(float(np.maximum(sel, topk).mean()))))
            if k_raw not in raw_at:  # no threshold and no fallback
                raw_at[k_raw] = self.pruned(raw, self.t_feat, 2.0, k_raw, (0.0, 0.0))
            out["raw_snr_topk"].append(raw_at[k_raw])
        return out

    def close(self):
        for d in self.devs:
            d.close()


def fastest(results, recall=0.99):
    """The operating point with the most stars/hour among those at or above
    the recall bar (None when none reaches it)."""
    ok = [r for r in results if r["recall"] >= recall]
    return max(ok, key=lambda r: r["stars_per_hour_pruned"]) if ok else None


def save(net, pc: ProposalConfig, path, **extra):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez(path, params=net.params(), config=json.dumps(dataclasses.asdict(pc)),
             **{k: np.asarray(v) for k, v in extra.items()})


def load(gpu, path=DEFAULT_WEIGHTS):
    """(network, ProposalConfig, extras) from a saved proposal."""
    z = np.load(path, allow_pickle=False)
    pc = ProposalConfig(**json.loads(str(z["config"])))
    net = Cnn(gpu, in_channels=pc.groups, width=pc.width, layers=pc.layers,
              kernel=pc.kernel, pos_weight=pc.pos_weight)
    net.set_params(z["params"])
    extras = {k: z[k] for k in z.files if k not in ("params", "config")}
    return net, pc, extras


def _noop(_):
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=200_000)
    ap.add_argument("--validate", type=int, default=2000)
    ap.add_argument("--heldout", type=int, default=20_000)
    ap.add_argument("--chunk", type=int, default=2000)
    ap.add_argument("--epochs", type=int, default=2)
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--report", default=None)
    ap.add_argument("--workers", type=int, default=16)
    ap.add_argument("--skip-eval", action="store_true",
                    help="train and save only; the operating point and the held-out "
                         "report come from a later --reselect run")
    ap.add_argument("--reselect", action="store_true",
                    help="keep the saved weights; redo the operating point and report")
    args = ap.parse_args()
    pc = ProposalConfig()
    cfg = search_config(pc)

    def log(msg):
        print(f"[proposal] {msg}", flush=True)

    import multiprocessing as mp

    pool = cf.ProcessPoolExecutor(args.workers, mp_context=mp.get_context("spawn"))
    list(pool.map(_noop, range(args.workers)))
    gpu = Gpu(0)
    d = cfg.durations_days()

    def chunks(total, seed0):
        """(states, injected periods) chunk by chunk; the host prepares the
        next chunk in the worker processes while the GPU works on this one."""
        spans = [range(c0, min(total, c0 + args.chunk)) for c0 in range(0, total, args.chunk)]

        def submit(idx):
            return pool.map(_make_state, [(i, seed0, d, cfg.correlated_noise) for i in idx],
                            chunksize=32)

        pending = submit(spans[0]) if spans else None
        for k, idx in enumerate(spans):
            res = list(pending)
            pending = submit(spans[k + 1]) if k + 1 < len(spans) else None
            yield idx, [r[0] for r in res], [r[1] for r in res]

    # disjoint seed ranges: training 4,000,000 + i (i < 200,000), held-out
    # 4,200,000 + i (i < 20,000), validation 4,250,000 + i
    seed_train, seed_val, seed_test = 4_000_000, 4_250_000, 4_200_000
    assert args.train <= seed_test - seed_train and args.heldout <= seed_val - seed_test
    t0 = time.perf_counter()
    prior_path = os.path.join(HERE, "data", "proposal_report.json")
    if args.reselect:
        net, _, _ = load(gpu)
        prior = json.load(open(prior_path)) if os.path.exists(prior_path) else {}
        t_train = float(prior.get("train_seconds", 0.0))
        losses = [float(prior.get("final_loss", 0.0))]
        positives = int(prior.get("train_positives", 0))
        args.train = int(prior.get("train_stars", args.train))
        args.epochs = int(prior.get("epochs", args.epochs))
    else:
        # the whole training set's inputs (float32, 64 KiB per star) and labels
        x = np.empty((args.train, pc.bins, pc.groups), dtype=np.float32)
        y = np.empty((args.train, pc.bins), dtype=np.float32)
        for idx, states, _ in chunks(args.train, seed_train):
            devs, xc, rows, _, _, _ = chunk_data(gpu, states, pc)
            x[idx.start:idx.stop] = xc
            y[idx.start:idx.stop] = labels_from(best_rows(rows, cfg.threshold), pc)
            for dv in devs:
                dv.close()
            log(f"train stars {idx.stop} ({time.perf_counter() - t0:.1f} s)")
        positives = int((y.sum(axis=1) > 0).sum())
        log(f"positives {positives} of {len(y)} stars")
        net = Cnn(gpu, in_channels=pc.groups, width=pc.width, layers=pc.layers,
                  kernel=pc.kernel, pos_weight=pc.pos_weight, seed=7)
        t1 = time.perf_counter()
        losses = train(net, x, y, epochs=args.epochs, log=log)
        t_train = time.perf_counter() - t1
        del x, y
        log(f"trained in {t_train:.1f} s")
    if args.skip_eval:
        pool.shutdown()
        save(net, pc, args.out, tau=0.1, topk=0, fallback=np.zeros(2))
        report = {"train_stars": args.train, "epochs": args.epochs,
                  "train_positives": positives, "train_seconds": t_train,
                  "final_loss": float(np.mean(losses[-50:])),
                  "config": dataclasses.asdict(pc)}
        if args.report:
            with open(args.report, "w") as f:
                json.dump(report, f, indent=1)
        return
    # operating point: the fastest reaching 99 % recall on validation stars
    vsweeps = []
    for _, vs, vi in chunks(args.validate, seed_val):
        val = EvalSet(gpu, vs, vi, pc)
        vsweeps.append(val.sweep(net)["learned"])
        val.close()
    vmerged = _merge(vsweeps)
    point = fastest(vmerged, 0.99) or max(vmerged, key=lambda r: r["recall"])
    topk, fallback, tau = point["topk"], point["fallback"], point["tau"]
    log(f"operating point tau {tau} top-k {topk} fallback {fallback}: validation {point}")
    learned, raw = [], []
    op = (tau, topk, tuple(fallback))
    for idx, hs, hi in chunks(args.heldout, seed_test):
        ev = EvalSet(gpu, hs, hi, pc)
        sw = ev.sweep(net, points=[op] + [p for p in POINTS if p != op])
        learned.append(sw["learned"])
        raw.append(sw["raw_snr_topk"])
        ev.close()
        log(f"held-out {idx.stop} stars ({time.perf_counter() - t0:.1f} s)")
    pool.shutdown()
    ml, mr = _merge(learned), _merge(raw)
    report = {
        "train_stars": args.train, "train_positives": positives,
        "validation_stars": args.validate,
        "heldout_stars": args.heldout, "epochs": args.epochs,
        "train_seconds": t_train, "final_loss": float(np.mean(losses[-50:])),
        "operating_point": {"tau": tau, "topk": topk, "fallback": fallback},
        "validation_at_operating_point": point,
        "heldout_at_operating_point": ml[0],
        "heldout_learned_sweep": ml, "heldout_raw_snr_sweep": mr,
        "config": dataclasses.asdict(pc),
    }
    save(net, pc, args.out, tau=tau, topk=topk, fallback=np.array(fallback, dtype=float))
    log(json.dumps(report["heldout_at_operating_point"]))
    if args.report:
        with open(args.report, "w") as f:
            json.dump(report, f, indent=1)


def _merge(chunks):
    """Per operating point, the sums over chunks of stars (recall, speed and
    kept fraction of the whole set)."""
    out = []
    for j in range(len(chunks[0])):
        rs = [c[j] for c in chunks]
        m = {k: rs[0][k] for k in ("topk", "fallback", "tau")}
        for k in ("stars", "full_recovered", "pruned_recovered", "fallback_stars",
                  "templates_full", "templates_pruned"):
            m[k] = int(sum(r[k] for r in rs))
        for k in ("seconds_full_scan", "seconds_proposal", "seconds_pruned_scan"):
            m[k] = float(sum(r[k] for r in rs))
        m["recall"] = m["pruned_recovered"] / max(m["full_recovered"], 1)
        m["kept_period_fraction"] = float(
            sum(r["kept_period_fraction"] * r["stars"] for r in rs) / m["stars"])
        m["stars_per_hour_full"] = m["stars"] * 3600.0 / m["seconds_full_scan"]
        m["stars_per_hour_pruned"] = m["stars"] * 3600.0 / (
            m["seconds_proposal"] + m["seconds_pruned_scan"])
        m["speedup"] = m["stars_per_hour_pruned"] / m["stars_per_hour_full"]
        m["evals_ratio"] = m["templates_full"] / max(m["templates_pruned"], 1)
        out.append(m)
    return out


if __name__ == "__main__":
    main()
