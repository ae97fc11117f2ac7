"""GPU: the batched multi-star search (packed ragged upload, every star's
bands merged into shared launches) against per-star searches and the
oracle.  Bit-exact: the same kernels see the same fixed-point fields."""
import numpy as np
import pytest

import fbls_amd as F
import oracle as O
from fbls_amd import synth

pytestmark = pytest.mark.gpu


def _states(indices, cfg, sectors=None):
    d = cfg.durations_days()
    out = []
    for i in indices:
        (t, f, e), kw = synth.batch_star(i)
        if sectors is not None:
            kw = dict(kw, sectors=sectors)
            t, f, e = F.synth_star(**kw)
        out.append(F.family_state(t, f - 1.0, e, d, cfg.correlated_noise))
    return d, out


@pytest.fixture(scope="module")
def six_star_batch(gpu):
    """Stars 0-5 of the 512-star batch (5-13 sectors, 95k-247k points)
    searched together in the merged launches."""
    cfg = F.SearchConfig(period_min=0.5, period_max=20.0)
    d, states = _states(range(6), cfg, sectors=None)
    print("[batch] n per star", [len(s.t) for s in states], flush=True)
    devs = F.DeviceReady.upload_packed(gpu, states)
    rows, stats = F.signal_search_batch(gpu, devs, d, [s.rho for s in states], cfg)
    yield cfg, d, states, devs, rows, stats
    for dv in devs:
        dv.close()


def test_batch_equals_single_star_searches(gpu, six_star_batch):
    cfg, d, states, devs, rows, stats = six_star_batch
    for k, (s, dev) in enumerate(zip(states, devs)):
        r1, grid, bands = F.signal_search(gpu, dev, d, s.rho, cfg)
        np.testing.assert_array_equal(rows[k], r1)
        assert stats[k]["templates_scanned"] == sum(b["templates_scanned"] for b in bands)
        assert stats[k]["periods"] == len(grid)


@pytest.mark.parametrize("k", range(6))
def test_batch_star_matches_oracle(gpu, six_star_batch, k):
    """Each star of the merged batch against the oracle's signal_search of
    that star alone (the oracle needs up to ~20 s for the 13-sector star)."""
    cfg, d, states, devs, rows, stats = six_star_batch
    s, dev = states[k], devs[k]
    print(f"[batch] oracle star {k} n={len(s.t)}", flush=True)
    vq, wq, _, _ = dev.fields()
    orows, obands, og, ost = O.signal_search(
        s.t, vq, wq, s.baseline_norm, s.origin, d, s.rho,
        cfg.ingress_fractions, cfg, cfg.period_min, cfg.period_max,
        cfg.period_oversample)
    assert ost == 0
    assert stats[k]["periods"] == og
    assert stats[k]["templates_scanned"] == sum(b["templates_scanned"] for b in obands)
    np.testing.assert_array_equal(rows[k], orows)


def test_batch_chunks_and_repeat(gpu):
    # long stars: more than 2^21 grid periods in total, so several chunks
    cfg = F.SearchConfig(period_min=0.5, period_max=20.0)
    d, states = _states(range(100, 112), cfg, sectors=13)
    print("[batch] chunks, n per star", [len(s.t) for s in states], flush=True)
    devs = F.DeviceReady.upload_packed(gpu, states)
    rows, stats = F.signal_search_batch(gpu, devs, d, [s.rho for s in states], cfg)
    assert sum(st["periods"] for st in stats) > (1 << 21)
    rows2, _ = F.signal_search_batch(gpu, devs, d, [s.rho for s in states], cfg)
    for a, b in zip(rows, rows2):
        np.testing.assert_array_equal(a, b)
    for k in (0, 7, 11):
        r1, _, _ = F.signal_search(gpu, devs[k], d, states[k].rho, cfg)
        np.testing.assert_array_equal(rows[k], r1)


def test_packed_and_list_uploads_agree(gpu, six_star_batch):
    """fbls_family_upload_packed (one concatenated array per field) and
    fbls_family_upload_list (one array per star, what DeviceReady.upload_packed
    calls) give the same fixed-point fields and the same search rows; a
    released upload's device buffer is reused by the next upload."""
    import ctypes

    cfg, d, states, devs, rows, stats = six_star_batch
    print("[batch] packed C entry vs list upload", flush=True)
    L = F.lib()
    k = len(states)
    offs = np.zeros(k + 1, dtype=np.uintp)
    offs[1:] = np.cumsum([len(s.t) for s in states])
    t = np.concatenate([s.t for s in states])
    v = np.concatenate([s.v for s in states])
    w = np.concatenate([s.w for s in states])
    c = np.array([s.baseline_norm for s in states])
    o = np.array([s.origin for s in states])
    P = ctypes.POINTER(ctypes.c_double)
    hs = (ctypes.c_void_p * k)()
    assert L.fbls_family_upload_packed(
        gpu._h, k, offs.ctypes.data_as(ctypes.POINTER(ctypes.c_size_t)), t.ctypes.data_as(P),
        v.ctypes.data_as(P), w.ctypes.data_as(P), c.ctypes.data_as(P), o.ctypes.data_as(P),
        hs) == 0
    packed = []
    for i, st in enumerate(states):
        dv = F.DeviceReady.__new__(F.DeviceReady)
        dv.gpu, dv.n, dv.baseline_norm, dv.origin = gpu, len(st.t), st.baseline_norm, st.origin
        dv.t_last = float(st.t[-1])
        dv._h = ctypes.c_void_p(hs[i])
        packed.append(dv)
    for a, b in zip(devs, packed):
        fa, fb = a.fields(), b.fields()
        for x, y in zip(fa, fb):
            np.testing.assert_array_equal(x, y)
    rows2, _ = F.signal_search_batch(gpu, packed, d, [s.rho for s in states], cfg)
    for a, b in zip(rows, rows2):
        np.testing.assert_array_equal(a, b)
    for dv in packed:
        dv.close()
    # the released buffer serves the next upload of the same stars
    again = F.DeviceReady.upload_packed(gpu, states)
    rows3, _ = F.signal_search_batch(gpu, again, d, [s.rho for s in states], cfg)
    for a, b in zip(rows, rows3):
        np.testing.assert_array_equal(a, b)
    for dv in again:
        dv.close()


def test_prepare_upload_equals_prepare_then_upload(gpu, six_star_batch):
    """fbls_family_prepare_upload_many (each star quantised and staged by its
    preparation thread) uploads the same fixed-point fields, times and
    correlated-noise factors as family_state followed by the list upload,
    and the batch search over its families returns the same rows."""
    cfg, d, states, devs, rows, stats = six_star_batch
    curves = [synth.batch_star(i)[0] for i in range(6)]
    print("[batch] prepare + upload in one call", flush=True)
    got, info = F.DeviceReady.prepare_upload(gpu, curves, d, cfg.correlated_noise,
                                             flux_offset=-1.0)
    for k, (g, s, dv) in enumerate(zip(got, states, devs)):
        assert g.n == len(s.t) and g.t_last == dv.t_last and g.origin == s.origin
        assert g.baseline_norm == s.baseline_norm
        np.testing.assert_array_equal(info["rho"][k], s.rho)
        for a, b in zip(g.fields(), dv.fields()):
            np.testing.assert_array_equal(a, b)
    rows2, _ = F.signal_search_batch(gpu, got, d, list(info["rho"]), cfg)
    for a, b in zip(rows2, rows):
        np.testing.assert_array_equal(a, b)
    for g in got:
        g.close()
    bad = (np.arange(4.0), np.ones(4), np.ones(4))
    with pytest.raises(F.FblsError) as ei:
        F.DeviceReady.prepare_upload(gpu, [curves[0], bad], d, True, flux_offset=-1.0)
    assert "star 1" in str(ei.value)
