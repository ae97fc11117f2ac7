"""End-to-end batch throughput, as `fbls batch` measures it (isakruas/fbls
src/batch.rs:74-106: every second from a star's light curve to its result
counts): from light curves in host memory to each star's best row.

Host preparation (`family_state`, curve.rs:541-569) runs on worker threads
-- the library's calls release the interpreter lock -- for chunk k + 1 while
the GPU uploads, searches and selects chunk k.  The search of a chunk is one
of:

  full     the packed upload and the merged-launch full scan of every star
           (fbls_signal_search_batch);
  pruned   the packed upload, the learned proposal on the device
           (fbls_propose) and the bin-restricted search (fbls_signal_search_bins)
           at an operating point (tau, top-k, fallback band);
  long     the long-baseline variant: the binned proposal per baseline group
           and the ragged bin-restricted search (fbls_signal_search_bins_ragged).

Returns the stars' best (period, SNR), the wall seconds of the whole run and
the seconds of each stage (preparation waiting time included in `wait`).
"""
import concurrent.futures as cf
import time

import numpy as np

from . import (DeviceReady, family_state, propose, signal_search_batch, signal_search_bins,
               signal_search_bins_ragged)
from . import proposal as PR


def _prepare(curve, durations, correlated):
    t, f, e = curve
    return family_state(t, f - 1.0, e, durations, correlated)


class Stages:
    def __init__(self):
        self.seconds = {}

    def add(self, name, dt):
        self.seconds[name] = self.seconds.get(name, 0.0) + dt


def _search_full(gpu, states, cfg, st):
    t0 = time.perf_counter()
    devs = DeviceReady.upload_packed(gpu, states)
    t1 = time.perf_counter()
    rows, stats = signal_search_batch(gpu, devs, cfg.durations_days(),
                                      [s.rho for s in states], cfg)
    t2 = time.perf_counter()
    best = PR.best_rows(rows, cfg.threshold)
    for d in devs:
        d.close()
    t3 = time.perf_counter()
    st.add("upload", t1 - t0)
    st.add("search", t2 - t1)
    st.add("select", t3 - t2)
    return best, sum(s["templates_scanned"] for s in stats)


def _search_pruned(gpu, states, cfg, net, pc, op, st):
    tau, topk, (lo, hi) = op
    gd = PR.group_durations(cfg, pc.groups)
    centers, edges = PR.bin_centers(pc), PR.bin_edges(pc)
    t0 = time.perf_counter()
    devs = DeviceReady.upload_packed(gpu, states)
    t1 = time.perf_counter()
    z = propose(gpu, net, devs, centers, gd, pc.nb_min, pc.nb_max)
    with np.errstate(over="ignore"):  # exp overflow is p = 0
        probs = 1.0 / (1.0 + np.exp(-z))
    bins = PR.keep_bins(probs, pc, tau, topk)
    conf = probs.max(axis=1)
    bins[(conf >= lo) & (conf < hi)] = True
    t2 = time.perf_counter()
    rows, stats = signal_search_bins(gpu, devs, cfg.durations_days(), [s.rho for s in states],
                                     cfg, bins, edges)
    t3 = time.perf_counter()
    best = PR.best_rows(rows, cfg.threshold)
    for d in devs:
        d.close()
    t4 = time.perf_counter()
    st.add("upload", t1 - t0)
    st.add("proposal", t2 - t1)
    st.add("search", t3 - t2)
    st.add("select", t4 - t3)
    return best, sum(s["templates_scanned"] for s in stats)


def _search_long(gpu, states, cfg, net, pc, op, st):
    from . import longbase as LB

    tau, topk, (lo, hi) = op
    gd = PR.group_durations(cfg, pc.groups)
    t0 = time.perf_counter()
    devs = DeviceReady.upload_packed(gpu, states)
    binned = [LB.binned_state(s) for s in states]
    bdevs = DeviceReady.upload_packed(gpu, binned)
    t1 = time.perf_counter()
    keeps, edges = [None] * len(states), [None] * len(states)
    for t_last, idx in LB.groups_of(states).items():
        centers, e = LB.feature_bins(t_last, pc)
        per = max(1, (1 << 24) // len(centers))  # network batch limit
        for c0 in range(0, len(idx), per):
            part = idx[c0:c0 + per]
            z = propose(gpu, net, [bdevs[k] for k in part], centers, gd, pc.nb_min, pc.nb_max)
            with np.errstate(over="ignore"):
                p = 1.0 / (1.0 + np.exp(-z))
            for j, k in enumerate(part):
                keep = LB.keep_from(p[j], pc.keep_dilate, tau, topk)
                if lo <= float(p[j].max()) < hi:
                    keep[:] = True
                keeps[k], edges[k] = keep, e
    for b in bdevs:
        b.close()
    t2 = time.perf_counter()
    rows, stats = signal_search_bins_ragged(gpu, devs, cfg.durations_days(),
                                            [s.rho for s in states], cfg, keeps, edges)
    t3 = time.perf_counter()
    best = PR.best_rows(rows, cfg.threshold)
    for d in devs:
        d.close()
    t4 = time.perf_counter()
    st.add("upload", t1 - t0)
    st.add("proposal", t2 - t1)
    st.add("search", t3 - t2)
    st.add("select", t4 - t3)
    return best, sum(s["templates_scanned"] for s in stats)


def run(gpu, curves, cfg, mode="full", net=None, pc=None, op=None, chunk=128, workers=16):
    """One end-to-end batch run over `curves` ((time, flux, flux_err) per
    star, flux normalised to 1).  Returns a dict with the stars' best rows,
    the wall seconds, stars/hour, trial evals, the stage seconds and the
    worker count."""
    durations = cfg.durations_days()
    st = Stages()
    best, templates = [], 0
    with cf.ThreadPoolExecutor(workers) as pool:
        def submit(a, b):
            return [pool.submit(_prepare, c, durations, cfg.correlated_noise)
                    for c in curves[a:b]]

        t_start = time.perf_counter()
        pending = submit(0, chunk)
        for c0 in range(0, len(curves), chunk):
            t0 = time.perf_counter()
            states = [f.result() for f in pending]
            st.add("wait", time.perf_counter() - t0)
            if c0 + chunk < len(curves):
                pending = submit(c0 + chunk, c0 + 2 * chunk)
            if mode == "full":
                b, n = _search_full(gpu, states, cfg, st)
            elif mode == "pruned":
                b, n = _search_pruned(gpu, states, cfg, net, pc, op, st)
            elif mode == "long":
                b, n = _search_long(gpu, states, cfg, net, pc, op, st)
            else:
                raise ValueError(mode)
            best += b
            templates += n
        wall = time.perf_counter() - t_start
    return {"best": best, "seconds": wall, "stars": len(curves),
            "stars_per_hour": len(curves) * 3600.0 / wall, "trial_evals": templates,
            "stage_seconds": st.seconds, "host_workers": workers, "chunk": chunk}


def preparation_rate(curves, cfg, workers=1):
    """Stars/hour of the host preparation alone on `workers` threads."""
    durations = cfg.durations_days()
    t0 = time.perf_counter()
    if workers <= 1:
        for c in curves:
            _prepare(c, durations, cfg.correlated_noise)
    else:
        with cf.ThreadPoolExecutor(workers) as pool:
            list(pool.map(lambda c: _prepare(c, durations, cfg.correlated_noise), curves))
    return len(curves) * 3600.0 / (time.perf_counter() - t0)
