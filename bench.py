#!/usr/bin/env python3
"""Benchmark of the FBLS full period scan on one MI355X.

Workload (one step): the full signal search of one multi-sector star -- four
contiguous 27.4-d sectors at 2-min cadence (~76k points after the 1-d
mid-sector gaps), AR(1) phi = 0.9 red noise plus white noise with total
sigma 600 ppm, an injected box (P = 11.3 d, depth 500 ppm, duration 4 h),
periods 0.5-20 d, the 13 default durations, prescan factor 2 (the reference
default).  Prepared-curve semantics: one family, no binning/detrending.

metric value: trial evals/s, "templates_scanned" of the reference
(periods x durations x shapes x phase bins, coarse + full pass) per second of
the whole step (host band loop included; inputs already resident in HBM).

roofline: the fused scan kernel against VALU issue, the unit that binds it.
The star (16 B per observation) stays in L2 and each workgroup folds it into
LDS, so HBM carries little more than the period records; VALU issue and LDS
are both about 70 % busy on the coarse launch.  The VALU instruction count
and the HBM bytes per launch come from the committed rocprofv3 counter
passes (profiles/), the launch duration from HIP events in this run.

cpu_baseline: the FP64 C oracle (same algorithm, OpenMP over periods) on the
host cores, over the same star, and on one core.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E datasheet


def workload():
    import fbls_amd as F
    from fbls_amd import synth

    t, f, e = synth.four_sector_red(seed=2)
    cfg = F.SearchConfig(period_min=0.5, period_max=20.0)
    d = cfg.durations_days()
    state = F.family_state(t, f - 1.0, e, d, cfg.correlated_noise)
    return cfg, d, state


def calibrate_copy(mib=256, reps=4):
    """Device-to-device copies of a known size: the HIP runtime's plain copy
    kernel, whose FETCH_SIZE / WRITE_SIZE scale the counters of the scan."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    nbytes = mib << 20
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # D2D
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"calibration_copy_bytes": nbytes, "reps": reps}))


COUNTER_PROFILE = os.path.join(ROOT, "profiles", "fused_scan_four_sector_star.md")

# VALU issue ceiling of the chip: each CU issues one wave64 VALU instruction
# per cycle (four 16-lane SIMDs, four cycles per wave64 instruction), 256 CUs
# at the 2.4 GHz peak engine clock
VALU_PEAK_G = 256 * 2.4


def measured_counters():
    """Per-launch counters of the scan kernel on this workload, from the
    committed rocprofv3 passes (HBM bytes from FETCH_SIZE / WRITE_SIZE scaled
    by the copy-kernel calibration; VALU wave-instructions from SQ_INSTS_VALU;
    the coarse launch's VALU and LDS busy fractions).  The counters need
    their own profiler runs, so the figures are read from the summary those
    runs produced."""
    import re

    try:
        text = open(COUNTER_PROFILE).read()
    except OSError:
        return None
    out = {"source": os.path.relpath(COUNTER_PROFILE, ROOT)}
    for key in ("traffic_bytes_per_launch", "valu_insts_per_launch", "valu_busy_coarse",
                "lds_busy_coarse", "valu_lane_insts_per_fold_pair", "fold_pairs_per_step"):
        k = re.search(key + r":\s*([0-9.eE+]+)", text)
        out[key] = float(k.group(1)) if k else None
    return out


def run_gpu(cfg, d, state, steps, warmup, launch=(0, 0, 0)):
    import fbls_amd as F

    gpu = F.Gpu(0)
    gpu.set_launch(*launch)
    dev = F.DeviceReady.upload(gpu, state)
    for _ in range(warmup):
        F.signal_search(gpu, dev, d, state.rho, cfg)
    per_step = []
    gpu.launch_totals(reset=True)
    t0 = time.perf_counter()
    for _ in range(steps):
        rows, grid, bands = F.signal_search(gpu, dev, d, state.rho, cfg)
        per_step.append(bands)
    elapsed = time.perf_counter() - t0
    totals = gpu.launch_totals(reset=True)
    return gpu, dev, rows, grid, per_step, elapsed, totals


def log(msg):
    """A progress line on stderr (the JSON line alone goes to stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _threads(fn, items, workers=16):
    """fn over items on host threads (the library's calls release the
    interpreter lock)."""
    import concurrent.futures as cf

    with cf.ThreadPoolExecutor(workers) as ex:
        return list(ex.map(fn, items))


def single_sector_curves(n_stars):
    """Fresh single-sector injection stars (seeds never used in training):
    [((time, flux, flux_err), injected period or 0)]."""
    from fbls_amd import proposal as PR
    from fbls_amd import synth_star

    def make(i):
        kw = PR.star_params(i, 4_300_000)
        return synth_star(**kw), kw.get("period", 0.0)

    return _threads(make, range(n_stars))


def batch_curves(n_stars):
    """The 512-star batch of uneven length (1-13 sectors)."""
    from fbls_amd import synth

    def make(i):
        curve, kw = synth.batch_star(i)
        return curve, kw.get("period", 0.0)

    return _threads(make, range(n_stars))


def long_curves(n_stars):
    """Fresh long-baseline stars (1-26 sectors around a one-year gap,
    periods 0.5-100 d; seeds never used in training or selection):
    [(curve, injected period or 0, kind, sectors)]."""
    from fbls_amd import synth

    def make(i):
        curve, kw = synth.long_baseline_star(i, 5_400_000)
        return curve, kw.get("period", 0.0), kw["kind"], kw["sectors"]

    return _threads(make, range(n_stars))


def prepare(curves, cfg):
    from fbls_amd import family_states

    d = cfg.durations_days()
    out = []
    for a in range(0, len(curves), 128):
        out += family_states(curves[a:a + 128], d, cfg.correlated_noise, flux_offset=-1.0)
    return out


def end_to_end(gpu, curves, injected, cfg, net=None, pc=None, op=None, mode="pruned"):
    """The full scan and the learned-pruned search of the same stars, each
    timed end to end from light curves in host memory (preparation on the
    library's preparation threads pipelined with the GPU, upload, proposal,
    search, selection; the median of three runs), with the recall of the
    pruned search against the full scan and the host preparation rate one
    star at a time and on the library's threads."""
    from fbls_amd import endtoend as E
    from fbls_amd import proposal as PR

    def median_run(**kw):
        """Three runs of the whole batch; the median by wall time, with
        every run's seconds."""
        runs = sorted((E.run(gpu, curves, cfg, **kw) for _ in range(3)),
                      key=lambda r: r["seconds"])
        r = runs[1]
        return r, {"seconds": r["seconds"], "stars_per_hour": r["stars_per_hour"],
                   "stage_seconds": r["stage_seconds"],
                   "seconds_of_runs": [x["seconds"] for x in runs]}

    E.run(gpu, curves[:4], cfg, "full", chunk=4)  # staging growth outside timing
    full, out_full = median_run(mode="full")
    out = {"stars": len(curves), "host_workers": full["host_workers"],
           "device_workers": full["device_workers"],
           "chunk_stars": full["chunk"], "runs": "median of 3", "full": out_full}
    rf = PR.recovered(full["best"], injected)
    out["full_scan_recovered"] = int(rf.sum())
    if net is not None:
        E.run(gpu, curves[:4], cfg, mode, net=net, pc=pc, op=op, chunk=4)
        pr, out["pruned"] = median_run(mode=mode, net=net, pc=pc, op=op)
        rp = PR.recovered(pr["best"], injected)
        out["recall_vs_full_scan"] = float((rf & rp).sum() / max(rf.sum(), 1))
        out["speedup"] = pr["stars_per_hour"] / full["stars_per_hour"]
    sample = curves[:min(64, len(curves))]
    out["host_preparation_stars_per_hour"] = {
        "1_thread": E.preparation_rate(sample, cfg, 1),
        f"{E.PREP_THREADS}_threads": E.preparation_rate(curves, cfg, E.PREP_THREADS)}
    return out


def batched(gpu, data):
    """The full scan of every star of the batch, end to end: preparation on
    the library's preparation threads pipelined with the packed upload and the merged-launch
    search (fbls_signal_search_batch) of chunks of 96 stars on two device
    contexts, and selection.  Rates are per wall second; the kernel seconds
    add up both contexts' launches (which overlap in time)."""
    import fbls_amd as F
    from fbls_amd import endtoend as E
    from fbls_amd import proposal as PR

    cfg = F.SearchConfig(period_min=0.5, period_max=20.0)
    curves = [c for c, _ in data]
    inj = [p for _, p in data]
    E.run(gpu, curves[:4], cfg, "full", chunk=4)  # staging growth outside timing
    r = E.run(gpu, curves, cfg, "full")
    tot = r["launch_totals"]
    found = PR.recovered(r["best"], inj)
    sizes = [len(c[0]) for c in curves]
    # the kernels alone, one context, no other work on the device: the scan's
    # own fold rate on these stars (the end-to-end run's two contexts overlap
    # in time, so the sum of their kernel seconds overstates the device time)
    sub = curves[:min(64, len(curves))]
    devs, info = F.DeviceReady.prepare_upload(gpu, sub, cfg.durations_days(),
                                              cfg.correlated_noise, flux_offset=-1.0)
    gpu.launch_totals(reset=True)
    F.signal_search_batch(gpu, devs, cfg.durations_days(), list(info["rho"]), cfg)
    one = gpu.launch_totals(reset=True)
    for x in devs:
        x.close()
    return {
        "workload": f"{len(curves)} stars, 1-13 sectors ({min(sizes)}-{max(sizes)} points), "
                    "sigma 200-3000 ppm, half with a box P 0.5-20 d; periods 0.5-20 d",
        "stars_per_hour": r["stars_per_hour"],
        "seconds": r["seconds"],
        "stage_seconds": r["stage_seconds"],
        "host_workers": r["host_workers"],
        "device_workers": r["device_workers"],
        "trial_evals_per_s": r["trial_evals"] / r["seconds"],
        "kernel_seconds_all_contexts": tot["seconds"],
        "launches": tot["launches"],
        "periods": int(tot["periods"]),
        "fold_pairs": int(tot["fold_pairs"]),
        "fold_pairs_per_kernel_s_all_contexts": tot["fold_pairs"] / tot["seconds"],
        "one_context": {"stars": len(sub), "kernel_seconds": one["seconds"],
                        "launches": one["launches"], "fold_pairs": int(one["fold_pairs"]),
                        "fold_pairs_per_kernel_s": one["fold_pairs"] / one["seconds"]},
        "injections_recovered": int(found.sum()),
        "injections": sum(1 for c in inj if c > 0),
        "host_preparation_stars_per_hour": {
            "1_thread": E.preparation_rate(curves[:16], cfg, 1),
            f"{E.PREP_THREADS}_threads": E.preparation_rate(curves, cfg, E.PREP_THREADS)},
    }


def learned_pruned(gpu, data):
    """The learned-pruned search against the full scan at the committed
    network's operating point: device-side (stars prepared and uploaded
    beforehand) and end to end."""
    from fbls_amd import proposal as PR

    n_stars = len(data)
    net, pc, extra = PR.load(gpu)
    cfg = PR.search_config(pc)
    curves = [c for c, _ in data]
    inj = [p for _, p in data]
    states = prepare(curves, cfg)
    ev = PR.EvalSet(gpu, states, inj, pc)
    topk, tau = int(extra["topk"]), float(extra["tau"])
    fb = tuple(float(v) for v in np.atleast_1d(extra["fallback"]))
    sw = ev.sweep(net, points=[(tau, topk, fb)])
    ev.close()
    r = sw["learned"][0]
    raw = sw["raw_snr_topk"][0]
    e2e = end_to_end(gpu, curves, inj, cfg, net, pc, (tau, topk, fb), "pruned")
    return {
        "workload": f"{n_stars} single-sector injection stars (sigma 200-3000 ppm, "
                    "half with a box), periods 0.5-13.7 d",
        "operating_point": {"top_k_bins": topk, "tau": tau, "fallback_band": list(fb)},
        "recall_vs_full_scan": r["recall"],
        "full_scan_recovered": r["full_recovered"],
        "stars_per_hour_full": r["stars_per_hour_full"],
        "stars_per_hour_pruned": r["stars_per_hour_pruned"],
        "speedup": r["stars_per_hour_pruned"] / r["stars_per_hour_full"],
        "trial_evals_full": r["templates_full"],
        "trial_evals_pruned": r["templates_pruned"],
        "kept_period_fraction": r["kept_period_fraction"],
        "fallback_stars": r["fallback_stars"],
        "raw_snr_ranking_recall_same_kept_bins": raw["recall"],
        "raw_snr_ranking_kept_period_fraction": raw["kept_period_fraction"],
        "end_to_end": e2e,
    }


def long_baseline(gpu, data):
    """The learned-pruned search of long-baseline stars (the network
    fine-tuned on them, at its committed operating point) against their
    full scan."""
    from fbls_amd import longbase as LB
    from fbls_amd import proposal as PR

    net, pc, extra = PR.load(gpu, LB.WEIGHTS)
    op = LB.operating_point(extra)
    cfg = LB.search_config()
    curves = [d[0] for d in data]
    states = prepare(curves, cfg)
    res = [(s, d[1], d[2], d[3]) for s, d in zip(states, data)]
    lb = LB.LongBatch(gpu, res, net, pc)
    summary = lb.summary()
    r = lb.sweep([op])[0]
    lb.close()
    return {
        "workload": f"{len(res)} stars, {summary['sectors_min']}-{summary['sectors_max']} "
                    "sectors in two seasons around a 365-d gap (up to "
                    f"{summary['points_max']} points), 30% transits + 5% eclipsing "
                    "binaries; periods 0.5-100 d",
        "operating_point": {"tau": op[0], "top_k_bins": op[1], "fallback_band": list(op[2]),
                            "raw_top_k_bins": op[3]},
        "recall_vs_full_scan": r["recall"],
        "full_scan_recovered": r["full_recovered"],
        "injections": summary["transits"] + summary["eclipsing_binaries"],
        "stars_per_hour_full": r["stars_per_hour_full"],
        "stars_per_hour_pruned": r["stars_per_hour_pruned"],
        "speedup": r["speedup"],
        "full_scan_trial_evals_per_s": summary["full_scan_trial_evals_per_s"],
        "full_scan_kernel_seconds": summary["full_scan_kernel_seconds"],
        "full_scan_launches": summary["full_scan_launches"],
        "full_scan_fold_pairs_per_kernel_s": summary["full_scan_fold_pairs_per_kernel_s"],
        "trial_evals_full": r["templates_full"],
        "trial_evals_pruned": r["templates_pruned"],
        "kept_period_fraction": r["kept_period_fraction"],
        "fallback_stars": r["fallback_stars"],
        "end_to_end": end_to_end(gpu, curves, [d[1] for d in data], cfg, net, pc, op, "long"),
    }


def long_star_stages(gpu, curves):
    """Stage attribution of each star's full search, one star at a time
    (FBLS_PROFILE: each band's kernel seconds split by the workgroups'
    shader cycles in the fold, the prefix sums and the template scan): per
    star its points, periods, prescan survivors, device seconds and fold
    pairs per device second; every band of the largest and the smallest."""
    import fbls_amd as F
    from fbls_amd import longbase as LB

    os.environ["FBLS_PROFILE"] = "1"
    cfg = LB.search_config()
    d = cfg.durations_days()
    keys = ("seconds_fold", "seconds_cumulative", "seconds_scan", "seconds_device")
    by_size = sorted((c[0] for c in curves), key=lambda c: len(c[0]))
    stars, detail = [], []
    for k, curve in enumerate(by_size):
        state = prepare([curve], cfg)[0]
        dev = F.DeviceReady.upload(gpu, state)
        for _ in range(2):  # the second search is reported
            _, _, bands = F.signal_search(gpu, dev, d, state.rho, cfg)
        dev.close()
        n = len(state.t)
        periods = sum(b["periods"] for b in bands)
        survivors = sum(b["prescan_survivors"] for b in bands)
        tot = {k2: sum(b[k2] for b in bands) for k2 in keys}
        stars.append({"points": n, "periods": periods, "survivors": survivors,
                      "fold_pairs": (periods + survivors) * n, **tot})
        if k in (0, len(by_size) - 1):
            detail.append([{k2: b[k2] for k2 in ("periods", "phase_bins", "prescan_survivors")
                            + keys} for b in bands])

    def rate(group):
        sec = sum(x["seconds_device"] for x in group)
        return sum(x["fold_pairs"] for x in group) / sec if sec else 0.0

    few = [x for x in stars if x["survivors"] < 0.1 * x["periods"]]
    return {"workload": f"full search of each of {len(stars)} long-baseline stars, one at a "
                        "time, periods 0.5-100 d",
            "stars": stars,
            "fold_pairs_per_device_s": rate(stars),
            "stars_with_few_survivors": len(few),
            "fold_pairs_per_device_s_few_survivors": rate(few),
            "bands_smallest": detail[0], "bands_largest": detail[-1]}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(cfg, d, state):
    """The FP64 C oracle's signal search of the same star on the host: on
    every OpenMP thread (the reported value) and on one thread."""
    import oracle as O

    def search(threads):
        O.set_threads(threads)
        t0 = time.perf_counter()
        rows, bands, g, st = O.signal_search(
            state.t, state.v, state.w, state.baseline_norm, state.origin, d,
            state.rho, cfg.ingress_fractions, cfg, cfg.period_min, cfg.period_max,
            cfg.period_oversample)
        dt = time.perf_counter() - t0
        return rows, sum(b["templates_scanned"] for b in bands), dt

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    rows, templates, dt = search(threads)
    _, templates1, dt1 = search(1)
    O.set_threads(0)
    return {"value": templates / dt, "unit": "trial evals/s", "cores": threads,
            "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"full star, one signal search ({templates:.3e} "
                      f"templates, {dt:.2f} s, FP64 C oracle, OpenMP)",
            "one_core": {"value": templates1 / dt1, "seconds": dt1,
                         "sample": "the same search on one thread"}}, rows


def oracle_errors(gpu, dev, state, d, cfg, rows, orows):
    """Largest relative error of each reported statistic of the GPU search
    against the FP64 oracle run on the unquantised (FP64) fields of the same
    star: SR (the reference's GLS amplitude SNR, `score`), dBIC and the
    evidence of the best row.  The reference defines no dBIC; here it is
    score^2 - ln(n): the whitened chi-square gain of the box depth at the
    found ephemeris minus the BIC charge of that one amplitude.  The GPU is
    bit-exact with the oracle on the fixed-point fields it scans; these
    errors are those of the 32-bit fixed point itself."""
    import oracle as O
    from fbls_amd import compare_search

    n = len(state.t)
    g = {float(r[0]): r for r in rows}
    o = {float(r[0]): r for r in orows}
    common = sorted(set(g) & set(o))

    def rel(a, b):
        return abs(a - b) / max(abs(b), 1e-300)

    sr = [rel(g[p][3], o[p][3]) for p in common]
    bic = [rel(g[p][3] ** 2 - np.log(n), o[p][3] ** 2 - np.log(n)) for p in common]
    best = float(rows[np.argmax(rows[:, 3])][0])
    ge = compare_search(gpu, dev, best, d, state.rho, state.residual_energy, cfg)
    oe = O.compare_search(state.t, state.v, state.w, state.baseline_norm, state.origin,
                          state.residual_energy, d, state.rho, best, cfg)
    keys = ("snr", "log_bf_planet_noise", "log_bf_binary_noise", "planet_log_odds")
    ev = {k: max(rel(a[k], b[k]) for a, b in zip(ge, oe)) for k in keys}
    return {"reference": "FP64 oracle on the unquantised fields",
            "rows_compared": len(common), "rows_gpu": len(rows), "rows_oracle": len(orows),
            "sr_max_rel": max(sr) if sr else None,
            "delta_bic_max_rel": max(bic) if bic else None,
            "evidence_best_row_max_rel": ev, "evidence_period": best}


def roofline(ctr, launch_s, launches, tot, steps, stream_bytes, fold_pairs):
    """The scan kernel against the unit that binds it.  The star (16 B per
    observation) stays in L2 and every workgroup re-reads it from there, so
    HBM carries only the records: the kernel is bound on chip, by VALU issue
    (with LDS close behind).  `achieved` is the VALU wave-instructions of one
    launch (SQ_INSTS_VALU, mean of the coarse and fine launch, from the
    committed profile) over the launch's mean duration measured here with HIP
    events on the library's stream; `peak` is one wave64 VALU instruction per
    CU per cycle at 2.4 GHz.  The HBM fraction is the measured FETCH + WRITE
    bytes per launch over the same duration."""
    valu = None if ctr is None else ctr["valu_insts_per_launch"]
    traffic = None if ctr is None else ctr["traffic_bytes_per_launch"]
    achieved = None if valu is None else valu / launch_s / 1e9
    hbm = None if traffic is None else traffic / launch_s / 1e9
    return {
        "bound": "valu",
        "achieved": achieved,
        "peak": VALU_PEAK_G,
        "unit": "G VALU wave-instructions/s",
        "frac": None if achieved is None else achieved / VALU_PEAK_G,
        "traffic": traffic,
        "kernel": "fused_scan<256> (coarse launch), fused_scan<512> (fine launch)",
        "launches_per_step": launches,
        "launch_ms": launch_s * 1e3,
        "kernel_ms_per_step": tot["seconds"] / steps * 1e3,
        "valu_insts_per_launch": valu,
        "valu_lane_insts_per_fold_pair": None if ctr is None
        else ctr["valu_lane_insts_per_fold_pair"],
        "fold_pairs_per_step": fold_pairs,
        "fold_pairs_per_kernel_s": fold_pairs / (tot["seconds"] / steps),
        "valu_busy_coarse": None if ctr is None else ctr["valu_busy_coarse"],
        "lds_busy_coarse": None if ctr is None else ctr["lds_busy_coarse"],
        "hbm": {"achieved": hbm, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": None if hbm is None else hbm / HBM_PEAK_GBS,
                "traffic_bytes_per_launch": traffic},
        # what a kernel that streamed the star from HBM for every period
        # would move (16 B per fold pair + 88 B per period record)
        "streaming_equivalent_bytes_per_launch": stream_bytes,
        "counter_source": None if ctr is None else ctr["source"],
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--lds", type=int, default=0,
                    help="LDS bytes per workgroup (0 = library default)")
    ap.add_argument("--block", type=int, default=0, help="256, 512 or 1024")
    ap.add_argument("--fold-mode", type=int, default=0,
                    help="0 default (equal weights fold no counts), 1 run-length "
                         "with the counts folded, 2 one atomic per pair")
    ap.add_argument("--proposal-stars", type=int, default=512,
                    help="stars of the learned-pruned comparison (0 = skip)")
    ap.add_argument("--batch-stars", type=int, default=512,
                    help="stars of the batched full scan of uneven stars (0 = skip)")
    ap.add_argument("--long-stars", type=int, default=32,
                    help="long-baseline stars of the learned-pruned comparison (0 = skip)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run the copy kernel used to scale counters")
    ap.add_argument("--only", choices=("learned_pruned", "batched", "long_baseline",
                                       "long_star_stages"),
                    help="run one section alone (its star count from the options above) and "
                         "print its JSON: the workload of a profile of that section's launches")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.only:
        import fbls_amd as F

        gpu = F.Gpu(0)
        if args.only == "learned_pruned":
            sec = learned_pruned(gpu, single_sector_curves(args.proposal_stars))
        elif args.only == "batched":
            sec = batched(gpu, batch_curves(args.batch_stars))
        elif args.only == "long_baseline":
            sec = long_baseline(gpu, long_curves(args.long_stars))
        else:
            sec = long_star_stages(gpu, long_curves(args.long_stars))
        print(json.dumps({args.only: sec}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: stars are independent and the "
              "multi-GPU layout (one process per GPU over a shared star list) "
              "is not run here; use --gpus 1", file=sys.stderr)
        sys.exit(2)

    pstars = single_sector_curves(args.proposal_stars) if args.proposal_stars > 0 else None
    bstars = batch_curves(args.batch_stars) if args.batch_stars > 0 else None
    lstars = long_curves(args.long_stars) if args.long_stars > 0 else None
    log("stars generated")
    cfg, d, state = workload()
    n = len(state.t)
    gpu, dev, rows, grid, per_step, elapsed, tot = run_gpu(
        cfg, d, state, args.steps, args.warmup,
        (args.lds, args.block, args.fold_mode))
    bands = per_step[-1]
    templates = sum(b["templates_scanned"] for b in bands)
    launch_periods = sum(b["periods"] + b["prescan_survivors"] for b in bands)
    # per launch of the scan kernel, from the library's HIP events on its
    # own stream over the timed steps
    launches = tot["launches"] / args.steps
    launch_s = tot["seconds"] / tot["launches"]
    stream_bytes = (16.0 * tot["fold_pairs"] + 88.0 * tot["periods"]) / tot["launches"]
    ctr = measured_counters()
    ms = elapsed / args.steps * 1e3
    value = templates / (elapsed / args.steps)
    best = rows[np.argmax(rows[:, 3])] if len(rows) else None

    out = {
        "metric": "BLS trial evals/s & stars/hr, 1 GPU: learned-pruned vs full scan @ \u226599% recall",
        "value_is": "trial evals/s of the full scan of the four-sector star (stars/hr in config; learned-pruned vs full scan in learned_pruned)",
        "value": value,
        "unit": "trial evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 phase + exact fixed-point sums + f64 merit",
        "data": "synthetic",
        "config": {
            "workload": "4-sector star, AR(1) red + white 600 ppm, box "
                        "P=11.3 d 500 ppm 4 h; periods 0.5-20 d; 13 durations;"
                        " prescan 2; prepared semantics",
            "observations": n,
            "periods": len(grid),
            "templates_per_step": templates,
            "fold_pairs_per_step": launch_periods * n,
            "stars_per_hour": 3600.0 / (elapsed / args.steps),
            "stars_per_hour_basis": "the star resident in HBM; the end_to_end sections "
                                    "include preparation and upload",
            "best_period": None if best is None else float(best[0]),
            "best_snr": None if best is None else float(best[3]),
        },
        "roofline": roofline(ctr, launch_s, launches, tot, args.steps, stream_bytes,
                             launch_periods * n),
    }
    log(f"four-sector star: {ms:.2f} ms per step")
    if pstars:
        out["learned_pruned"] = learned_pruned(gpu, pstars)
        log("single-sector learned-pruned vs full scan done")
    if bstars:
        out["batched"] = batched(gpu, bstars)
        log("batch of uneven stars done")
    if lstars:
        out["long_baseline"] = long_baseline(gpu, lstars)
        log("long-baseline learned-pruned vs full scan done")
    if not args.no_cpu:
        cb, orows = cpu_baseline(cfg, d, state)
        out["cpu_baseline"] = cb
        try:
            out["error_vs_fp64_oracle"] = oracle_errors(gpu, dev, state, d, cfg, rows, orows)
        except Exception as exc:  # a report field: never lose the bench line over it
            out["error_vs_fp64_oracle"] = {"failed": repr(exc)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
