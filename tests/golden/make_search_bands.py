"""Extracts the per-band scan statistics of the reference's one committed
search report (public/2026-11-08/refine/72668830/search.json in
isakruas/fbls) into tests/golden/search_bands_72668830.json.

Run in a container that mounts the reference at /root/reference; the output
is data only (counts per band), used to pin the period grid, band split,
phase bins, batches and templates_scanned.
"""
import json
import os

SRC = "/root/reference/public/2026-11-08/refine/72668830/search.json"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                   "search_bands_72668830.json")

d = json.load(open(SRC))
cfg = d["config"]
keep_cfg = {k: cfg[k] for k in (
    "durations_hours", "period_min", "period_max", "period_oversample",
    "phase_oversample", "prescan_factor", "prescan_fraction", "period_batch",
    "max_fold_mib", "ingress_fractions", "shape_layers", "threshold")}
passes = []
for p in d["statistics"]["passes"]:
    fams = []
    for f in p["families"]:
        fams.append({
            "family": f["family"],
            "durations_hours": f["durations_hours"],
            "periods": f["periods"],
            "observations": f["observations"],
            "bands": [{k: b[k] for k in ("periods", "prescan_survivors",
                                         "phase_bins", "batches",
                                         "templates_scanned",
                                         "exact_tile_retries")}
                      for b in f["bands"]],
        })
    passes.append({"pass": p["pass"], "families": fams})
json.dump({"source": "public/2026-11-08/refine/72668830/search.json",
           "config": keep_cfg,
           "periods_tested": d["statistics"]["periods_tested"],
           "passes": passes}, open(OUT, "w"), indent=1)
print("wrote", OUT)
