"""End-to-end batch throughput, as `fbls batch` measures it (isakruas/fbls
src/batch.rs:74-106: every second from a star's light curve to its result
counts): from light curves in host memory to each star's best row.

Host preparation (`family_state`, curve.rs:541-569) and upload of the
chunks ahead run on the library's own preparation threads, one
fbls_family_prepare_upload_many call per chunk (each star quantised and
written to pinned staging as it is prepared, then one copy on the library's
upload stream), in chunk order from one Python thread, while the GPU
searches and selects the chunks before.  The search of a chunk is one of:

  full     the packed upload and the merged-launch full scan of every star
           (fbls_signal_search_batch);
  pruned   the packed upload, the learned proposal on the device
           (fbls_propose) and the bin-restricted search (fbls_signal_search_bins)
           at an operating point (tau, top-k, fallback band);
  long     the long-baseline variant: the binned proposal per baseline group
           (network bins and the periodogram's raw top bins) and the ragged
           bin-restricted search (fbls_signal_search_bins_ragged).

Returns the stars' best (period, SNR), the wall seconds of the whole run and
the seconds of each stage (the time the search waited for a prepared and
uploaded chunk in `wait`).
"""
import concurrent.futures as cf
import os
import threading
import time

import numpy as np

from . import (DeviceReady, family_state, family_states, propose, signal_search_batch,
               signal_search_bins, signal_search_bins_ragged)
from . import proposal as PR

# threads that run the library's preparation (fbls_family_state_many): its
# pool, four short of the cores (at most 16), and the calling thread
_CORES = min(os.cpu_count() or 1, 16)
PREP_THREADS = max(1, _CORES - 4 if _CORES > 4 else 1) + 1


def _prepare(curve, durations, correlated):
    t, f, e = curve
    return family_state(t, f - 1.0, e, durations, correlated)


def _prepare_many(curves, durations, correlated):
    return family_states(curves, durations, correlated, flux_offset=-1.0)


class Stages:
    def __init__(self):
        self.seconds = {}
        self._lock = threading.Lock()

    def add(self, name, dt):
        with self._lock:
            self.seconds[name] = self.seconds.get(name, 0.0) + dt


def _upload(gpu, states, mode):
    """The chunk's device families (and, for the long mode, its binned
    families for the periodogram)."""
    devs = DeviceReady.upload_packed(gpu, states)
    if mode != "long":
        return devs, None
    from . import longbase as LB

    return devs, DeviceReady.upload_packed(gpu, [LB.binned_state(s) for s in states])


def _search_full(gpu, rho, devs, cfg, st):
    t1 = time.perf_counter()
    rows, stats = signal_search_batch(gpu, devs, cfg.durations_days(), rho, cfg)
    t2 = time.perf_counter()
    best = PR.best_rows(rows, cfg.threshold)
    t3 = time.perf_counter()
    st.add("search", t2 - t1)
    st.add("select", t3 - t2)
    return best, sum(s["templates_scanned"] for s in stats)


def _search_pruned(gpu, rho, devs, cfg, net, pc, op, st):
    tau, topk, (lo, hi) = op[:3]
    gd = PR.group_durations(cfg, pc.groups)
    centers, edges = PR.bin_centers(pc), PR.bin_edges(pc)
    t1 = time.perf_counter()
    z = propose(gpu, net, devs, centers, gd, pc.nb_min, pc.nb_max)
    with np.errstate(over="ignore"):  # exp overflow is p = 0
        probs = 1.0 / (1.0 + np.exp(-z))
    bins = PR.keep_bins(probs, pc, tau, topk)
    conf = probs.max(axis=1)
    bins[(conf >= lo) & (conf < hi)] = True
    t2 = time.perf_counter()
    rows, stats = signal_search_bins(gpu, devs, cfg.durations_days(), rho, cfg, bins, edges)
    t3 = time.perf_counter()
    best = PR.best_rows(rows, cfg.threshold)
    t4 = time.perf_counter()
    st.add("proposal", t2 - t1)
    st.add("search", t3 - t2)
    st.add("select", t4 - t3)
    return best, sum(s["templates_scanned"] for s in stats)


def _search_long(gpu, states, devs, bdevs, cfg, net, pc, op, st):
    from . import longbase as LB

    gd = PR.group_durations(cfg, pc.groups)
    t1 = time.perf_counter()
    keeps, edges = [None] * len(states), [None] * len(states)
    for t_last, idx in LB.groups_of(states).items():
        centers, e = LB.feature_bins(t_last, pc)
        per = max(1, (1 << 24) // len(centers))  # network batch limit
        for c0 in range(0, len(idx), per):
            part = idx[c0:c0 + per]
            z, pk = propose(gpu, net, [bdevs[k] for k in part], centers, gd, pc.nb_min,
                            pc.nb_max, peaks=True)
            with np.errstate(over="ignore"):
                p = 1.0 / (1.0 + np.exp(-z))
            for j, k in enumerate(part):
                keeps[k] = LB.keep_mask(p[j], pk[j], pc.keep_dilate, op)
                edges[k] = e
    t2 = time.perf_counter()
    rows, stats = signal_search_bins_ragged(gpu, devs, cfg.durations_days(),
                                            [s.rho for s in states], cfg, keeps, edges)
    t3 = time.perf_counter()
    best = PR.best_rows(rows, cfg.threshold)
    t4 = time.perf_counter()
    st.add("proposal", t2 - t1)
    st.add("search", t3 - t2)
    st.add("select", t4 - t3)
    return best, sum(s["templates_scanned"] for s in stats)


def _ramp(chunk):
    """Chunk sizes: chunk / 4, chunk / 2, then chunk for ever."""
    yield max(1, chunk // 4)
    yield max(1, chunk // 2)
    while True:
        yield chunk


def _peers(gpu, k):
    """k further device contexts on gpu's device, kept on gpu and reused by
    later runs (their staging and device arenas stay warm, as in a batch
    service that keeps its workers)."""
    from . import Gpu

    peers = gpu.__dict__.setdefault("_e2e_peers", [])
    while len(peers) < k:
        peers.append(Gpu(gpu.device))
    return peers[:k]


def _net_copy(net, gpu):
    """A copy of the proposal network on another device context, made once
    per (network, context) and refreshed from the network's current
    parameters on each run."""
    from . import Cnn

    copies = gpu.__dict__.setdefault("_e2e_nets", {})
    copy = copies.get(id(net))
    if copy is None or copy[0] is not net:
        c = net.cfg
        copy = (net, Cnn(gpu, in_channels=c.in_channels, width=c.width, layers=c.layers,
                         kernel=c.kernel, pos_weight=c.pos_weight))
        copies[id(net)] = copy
    copy[1].set_params(net.params())
    return copy[1]


def run(gpu, curves, cfg, mode="full", net=None, pc=None, op=None, chunk=96, device_workers=2):
    """One end-to-end batch run over `curves` ((time, flux, flux_err) per
    star, flux normalised to 1).  The stages overlap: preparation and
    upload on the library's preparation threads and upload stream (chunk
    after chunk, in order; at most device_workers + 1 chunks on the device
    and not yet searched), and the
    searches on `device_workers` threads, each with a device context of its
    own (and, for the learned modes, its own copy of the network), taking
    chunks in turn, so that one chunk's search overlaps the next one's
    proposal and host planning on the GPU -- the reference's `batch
    --workers` (src/batch.rs:47-68).  The first two chunks are a quarter and
    a half of `chunk` stars.  Returns a dict with the stars' best rows, the
    wall seconds, stars/hour, trial evals, the stage seconds summed over the
    device threads (`wait`: time a device thread waited for its chunk;
    `prepare`, `upload`: the feeding thread's busy time) and the thread
    counts."""
    if mode not in ("full", "pruned", "long"):
        raise ValueError(mode)
    durations = cfg.durations_days()
    st = Stages()
    # the first chunks are small, so the device starts while the host still
    # prepares the rest (chunk / 4, chunk / 2, then chunk)
    spans, a = [], 0
    for size in _ramp(chunk):
        if a >= len(curves):
            break
        spans.append((a, min(len(curves), a + size)))
        a += size
    nw = max(1, min(device_workers, len(spans)))
    ctxs = [gpu] + _peers(gpu, nw - 1)
    before = [g.launch_totals() for g in ctxs]
    nets = [net] + [_net_copy(net, g) for g in ctxs[1:]] if net is not None else [None] * nw
    ahead = threading.Semaphore(nw + 1)
    stop = threading.Event()  # a device thread failed: pending uploads give up
    results = [None] * len(spans)
    with cf.ThreadPoolExecutor(1) as feed, cf.ThreadPoolExecutor(nw) as dev:
        def prepare_upload(a, b):
            # at most nw + 1 chunks on the device and not yet searched
            while not ahead.acquire(timeout=0.1):
                if stop.is_set():
                    raise RuntimeError("the search of an earlier chunk failed")
            try:
                t0 = time.perf_counter()
                if mode == "long":  # the binned families come from the prepared arrays
                    states = _prepare_many(curves[a:b], durations, cfg.correlated_noise)
                    t1 = time.perf_counter()
                    devs, bdevs = _upload(gpu, states, mode)
                    rho = [s.rho for s in states]
                else:  # each star's fixed point and image written as it is prepared
                    devs, info = DeviceReady.prepare_upload(gpu, curves[a:b], durations,
                                                            cfg.correlated_noise,
                                                            flux_offset=-1.0)
                    states, bdevs, rho = None, None, list(info["rho"])
                    t1 = time.perf_counter()
                st.add("prepare", t1 - t0)
                st.add("upload", time.perf_counter() - t1)
            except BaseException:
                ahead.release()
                raise
            return states, rho, devs, bdevs

        def device_work(j):
            try:
                search_chunks(j)
            except BaseException:
                stop.set()
                raise

        def search_chunks(j):
            g, nt = ctxs[j], nets[j]
            for k in range(j, len(spans), nw):
                t0 = time.perf_counter()
                states, rho, devs, bdevs = ups[k].result()
                st.add("wait", time.perf_counter() - t0)
                try:
                    if mode == "full":
                        results[k] = _search_full(g, rho, devs, cfg, st)
                    elif mode == "pruned":
                        results[k] = _search_pruned(g, rho, devs, cfg, nt, pc, op, st)
                    else:
                        results[k] = _search_long(g, states, devs, bdevs, cfg, nt, pc, op, st)
                finally:
                    for d in devs + (bdevs or []):
                        d.close()
                    ahead.release()

        t_start = time.perf_counter()
        ups = [feed.submit(prepare_upload, a, b) for a, b in spans]
        for f in [dev.submit(device_work, j) for j in range(nw)]:
            f.result()
        wall = time.perf_counter() - t_start
    best, templates = [], 0
    for b, n in results:
        best += b
        templates += n
    # scan launches of this run on every context (the contexts' kernels
    # overlap in time, so their seconds add up to more than the wall time)
    totals = {k: 0 for k in before[0]}
    for g, b0 in zip(ctxs, before):
        for k, v in g.launch_totals().items():
            totals[k] += v - b0[k]
    return {"best": best, "seconds": wall, "stars": len(curves),
            "stars_per_hour": len(curves) * 3600.0 / wall, "trial_evals": templates,
            "stage_seconds": st.seconds, "host_workers": PREP_THREADS, "device_workers": nw,
            "chunk": chunk, "launch_totals": totals}


def preparation_rate(curves, cfg, threads=1):
    """Stars/hour of the host preparation alone: one star at a time on the
    calling thread (threads=1), or in chunks of 128 stars on the library's
    preparation threads."""
    durations = cfg.durations_days()
    t0 = time.perf_counter()
    if threads <= 1:
        for c in curves:
            _prepare(c, durations, cfg.correlated_noise)
    else:
        for a in range(0, len(curves), 128):
            _prepare_many(curves[a:a + 128], durations, cfg.correlated_noise)
    return len(curves) * 3600.0 / (time.perf_counter() - t0)
