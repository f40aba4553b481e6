"""Shared helpers for full-run tests: workloads, oracle CPU runs and the
statistical comparisons (yields within 3 sigma, chi-square on histograms)."""
from __future__ import annotations

import ctypes as C

import numpy as np

import oracle_lib as ol
from nebula_amd import _lib as L


def pencil_primaries(n, energy, pos=(0.0, 0.0, 5.0), direction=(0.0, 0.0, -1.0)):
    p = np.zeros(n, L.PARTICLE_DTYPE)
    p["pos"] = pos
    p["dir"] = direction
    p["kin_energy"] = energy
    return p


def oracle_run(og, om, primaries, seed=0x78C7E39B, threads=8, rng_kind=0, min_energy=0.0, time_limit=0.0):
    """cpu_mt_main-style run of the oracle; returns (stats dict, hist)."""
    n = len(primaries)
    tags = np.arange(n, dtype=np.uint32)
    hist = np.zeros(L.HIST_BINS, np.uint64)
    st = ol.RunStats()
    rc = ol.orc.orc_cpu_mt_run(og.h, om.h, ol.p(primaries), ol.p(tags), n, min_energy, seed, threads, rng_kind,
                              time_limit, None, None, 0, ol.p(hist), C.byref(st))
    assert rc == 0
    return {f: getattr(st, f) for f, _ in ol.RunStats._fields_}, hist


def yield_z(k1, n1, k2, n2):
    """z-score of the difference of two yields (detections per primary).
    Detections per primary are counts with variance estimated by the Poisson
    bound k/n^2 (cascades give over-dispersion; the per-primary variance is
    estimated from k as well, conservative for yields below ~1)."""
    y1, y2 = k1 / n1, k2 / n2
    var = k1 / n1 ** 2 + k2 / n2 ** 2
    return (y1 - y2) / np.sqrt(max(var, 1e-300))


def chi2_hist(h1, h2, min_count=20):
    """Two-sample chi-square on histograms (bins with enough counts merged
    from the top), returns p-value."""
    from scipy import stats
    h1 = np.asarray(h1, float); h2 = np.asarray(h2, float)
    # merge bins until each merged bin has >= min_count combined
    edges, acc1, acc2, m1, m2 = [], 0.0, 0.0, [], []
    for a, b in zip(h1, h2):
        acc1 += a; acc2 += b
        if acc1 + acc2 >= min_count:
            m1.append(acc1); m2.append(acc2); acc1 = acc2 = 0.0
    if acc1 + acc2 > 0 and m1:
        m1[-1] += acc1; m2[-1] += acc2
    m1, m2 = np.array(m1), np.array(m2)
    n1, n2 = m1.sum(), m2.sum()
    k1, k2 = np.sqrt(n2 / n1), np.sqrt(n1 / n2)
    chi = np.sum((k1 * m1 - k2 * m2) ** 2 / (m1 + m2))
    dof = len(m1) - 1
    return stats.chi2.sf(chi, dof), chi, dof


def ks_binned(h1, h2):
    """Kolmogorov-Smirnov distance between two binned distributions and the
    asymptotic p-value (1-eV bins: exact enough for tie-free comparison)."""
    from scipy import stats
    c1 = np.cumsum(h1) / max(np.sum(h1), 1); c2 = np.cumsum(h2) / max(np.sum(h2), 1)
    d = np.max(np.abs(c1 - c2))
    n1, n2 = np.sum(h1), np.sum(h2)
    en = np.sqrt(n1 * n2 / (n1 + n2))
    return d, stats.kstwobign.sf(en * d)
