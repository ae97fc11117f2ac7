"""Synthetic workloads of the benchmark and tests (seeded, deterministic).

Each function returns (time, flux, flux_err) from the C generator
(fbls_synth_star: xorshift64 + Box-Muller as in survey/numeric.rs:345-364)
with normalised flux (1 + noise - transits) and constant flux_err.
"""
import numpy as np

from . import synth_star

CADENCE = 1.0 / 720.0  # 2 min
T0 = 2458000.0


def single_sector(seed=1):
    """One 27.4-d sector, 1-d mid-sector gap, white sigma 1000 ppm, box
    P = 3.7 d, depth 2000 ppm, duration 2.5 h, random epoch."""
    return synth_star(seed=seed, sectors=1, sector_days=27.4, gap_days=1.0,
                      cadence_days=CADENCE, t0=T0, sigma=1e-3, period=3.7,
                      depth=2e-3, duration=2.5 / 24.0)


def four_sector_red(seed=2):
    """Four contiguous sectors (~79k points, 1-d gap per sector), AR(1)
    phi = 0.9 red noise holding half the variance, total sigma 600 ppm, box
    P = 11.3 d, depth 500 ppm, duration 4 h, random epoch."""
    return synth_star(seed=seed, sectors=4, sector_days=27.4, gap_days=1.0,
                      cadence_days=CADENCE, t0=T0, sigma=600e-6,
                      red_fraction=0.5, ar_phi=0.9, period=11.3, depth=500e-6,
                      duration=4.0 / 24.0)


def batch_star(i, seed0=3000):
    """Star i of the 512-star batch: 1-13 sectors, sigma log-uniform
    200-3000 ppm, half with a box (P log-uniform 0.5-20 d, depth log-uniform
    200-10000 ppm, duration 1 h + 7 h (P^(1/3) - 0.5^(1/3)) /
    (20^(1/3) - 0.5^(1/3)))."""
    rng = np.random.default_rng(seed0 + i)
    sectors = int(rng.integers(1, 14))
    sigma = float(np.exp(rng.uniform(np.log(200e-6), np.log(3000e-6))))
    kw = dict(seed=seed0 + i, sectors=sectors, sector_days=27.4, gap_days=1.0,
              cadence_days=CADENCE, t0=T0, sigma=sigma)
    if rng.uniform() < 0.5:
        p = float(np.exp(rng.uniform(np.log(0.5), np.log(20.0))))
        depth = float(np.exp(rng.uniform(np.log(200e-6), np.log(1e-2))))
        dur = (1.0 + 7.0 * (p ** (1 / 3) - 0.5 ** (1 / 3))
               / (20 ** (1 / 3) - 0.5 ** (1 / 3))) / 24.0
        kw.update(period=p, depth=depth, duration=dur)
    return synth_star(**kw), kw


def _duration_days(p, p_max):
    """1 h + 7 h (P^(1/3) - 0.5^(1/3)) / (p_max^(1/3) - 0.5^(1/3)), in days."""
    return (1.0 + 7.0 * (p ** (1 / 3) - 0.5 ** (1 / 3))
            / (p_max ** (1 / 3) - 0.5 ** (1 / 3))) / 24.0


def long_baseline_star(i, seed0=5_000_000, transit=0.30, eb=0.05):
    """Star i of the long-baseline batch: 1-26 sectors in two seasons of
    ceil(k/2) and floor(k/2) contiguous 27.4-d sectors separated by a
    365-d gap (up to ~510k points over ~1080 d); sigma log-uniform
    200-3000 ppm of which an AR(1) (phi = 0.9) share uniform in [0, 0.5] of
    the variance; a share `transit` (30 %) carries a box (P log-uniform
    0.5-100 d, depth log-uniform 200-10000 ppm, duration 1-8 h growing as
    P^(1/3)), a share `eb` (5 %) an eclipsing-binary-like pair (primary
    depth log-uniform 2000-30000 ppm, secondary at phase 0.5 with a depth
    ratio uniform in [0.1, 0.9], the same period and duration laws).
    Returns ((time, flux, flux_err), kw) with kw["kind"] in
    {"none", "transit", "eb"}."""
    rng = np.random.default_rng(seed0 + i)
    sectors = int(rng.integers(1, 27))
    sigma = float(np.exp(rng.uniform(np.log(200e-6), np.log(3000e-6))))
    kw = dict(seed=seed0 + i, sectors=sectors, sector_days=27.4, gap_days=1.0,
              cadence_days=CADENCE, t0=T0, sigma=sigma,
              red_fraction=float(rng.uniform(0.0, 0.5)), ar_phi=0.9,
              baseline_gap_days=365.0 if sectors > 1 else 0.0)
    u = rng.uniform()
    p = float(np.exp(rng.uniform(np.log(0.5), np.log(100.0))))
    if u < transit:
        depth = float(np.exp(rng.uniform(np.log(200e-6), np.log(1e-2))))
        kw.update(period=p, depth=depth, duration=_duration_days(p, 100.0))
    elif u < transit + eb:
        depth = float(np.exp(rng.uniform(np.log(2e-3), np.log(3e-2))))
        kw.update(period=p, depth=depth, duration=_duration_days(p, 100.0),
                  secondary_depth=depth * float(rng.uniform(0.1, 0.9)))
    kind = "none" if u >= transit + eb else ("transit" if u < transit else "eb")
    return synth_star(**kw), dict(kw, kind=kind)


def write_npz(path, workload="four_sector_red", seed=None, raw=True):
    """Writes a workload as an NPZ light curve (time_bjd, flux, flux_err,
    sector).  raw: the flux in counts (x 12000) with a slow 6-day trend of
    0.4 %, as the raw search expects; otherwise the normalised flux."""
    fn = {"single_sector": single_sector, "four_sector_red": four_sector_red}[workload]
    t, f, e = fn() if seed is None else fn(seed=seed)
    sector = (np.floor((t - t[0]) / 28.4)).astype(np.int64)
    if raw:
        slow = 1.0 + 0.004 * np.sin(2.0 * np.pi * (t - t[0]) / 6.0)
        f, e = 12000.0 * f * slow, 12000.0 * e
    np.savez(path, time_bjd=t, flux=f, flux_err=e, sector=sector)


if __name__ == "__main__":
    import argparse

    ap = argparse.ArgumentParser(description="write a synthetic workload as NPZ")
    ap.add_argument("workload", choices=["single_sector", "four_sector_red"])
    ap.add_argument("out")
    ap.add_argument("--normalised", action="store_true")
    a = ap.parse_args()
    write_npz(a.out, a.workload, raw=not a.normalised)
