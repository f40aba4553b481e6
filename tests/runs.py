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


def yield_z(k1, n1, k2, n2, dispersion=1.0):
    """z-score of the difference of two yields (detections per primary) from
    counts alone, with the per-primary variance taken as dispersion x the
    yield (1 = Poisson).  A primary's cascade can emit several secondaries,
    so the true variance is larger (about 1.6 x the SE yield at a line
    edge): pass the measured ratio, or use per_primary() and z_from() where
    detection records with primary tags are at hand."""
    y1, y2 = k1 / n1, k2 / n2
    var = dispersion * (k1 / n1 ** 2 + k2 / n2 ** 2)
    return (y1 - y2) / np.sqrt(max(var, 1e-300))


def per_primary(energies, tags, n):
    """SE (E < 50 eV) and BSE detections per primary from detection records
    (tags = primary index 0..n-1): {"se": (mean, variance), "bse": ...}."""
    e = np.asarray(energies, np.float32)
    t = np.asarray(tags, np.int64)
    se = np.bincount(t[e < 50], minlength=n)[:n]
    bse = np.bincount(t[e >= 50], minlength=n)[:n]
    return {"se": (se.mean(), se.var(ddof=1)), "bse": (bse.mean(), bse.var(ddof=1)), "n": n}


def z_from(a, b, key):
    """z-score of the difference of two per_primary() yields, each with its
    own measured per-primary variance."""
    (m1, v1), (m2, v2) = a[key], b[key]
    return (m1 - m2) / np.sqrt(max(v1 / a["n"] + v2 / b["n"], 1e-300))


def oracle_records(og, om, primaries, seed=0x78C7E39B, threads=16, rng_kind=0, per_primary_cap=8):
    """cpu_mt-style oracle run that keeps its detection records: (stats,
    histogram, detected energies, detected tags), tags = primary index."""
    n = len(primaries)
    tags = np.arange(n, dtype=np.uint32)
    cap = per_primary_cap * n
    det = np.zeros(cap, L.PARTICLE_DTYPE); dtags = np.zeros(cap, np.uint32)
    hist = np.zeros(L.HIST_BINS, np.uint64)
    st = ol.RunStats()
    rc = ol.orc.orc_cpu_mt_run(og.h, om.h, ol.p(np.ascontiguousarray(primaries)), ol.p(tags), n, 0.0, seed, threads,
                              rng_kind, 0.0, ol.p(det), ol.p(dtags), cap, ol.p(hist), C.byref(st))
    assert rc == 0 and st.detected <= cap
    k = st.detected
    return {f: getattr(st, f) for f, _ in ol.RunStats._fields_}, hist, det["kin_energy"][:k], dtags[:k]


def dispersion(energies, tags, n):
    """Variance-to-mean ratio of detections per primary (>= 1): one primary's
    detections are not independent draws, so counting statistics over
    detections understate the variance by this factor."""
    c = np.bincount(np.asarray(tags, np.int64), minlength=n)[:n]
    return max(1.0, float(c.var(ddof=1) / max(c.mean(), 1e-300)))


def spectra_agree(h1, h2, label, disp=1.0, p_min=0.0027):
    """Chi-square (merged bins) and KS on two detected-energy histograms, the
    statistics deflated by the detections' dispersion (histogram counts of
    cascades are over-dispersed by about that factor); asserts both p-values
    exceed p_min (3 sigma, two-sided) and returns them."""
    from scipy import stats
    _, chi, dof = chi2_hist(h1, h2)
    p_chi = stats.chi2.sf(chi / disp, dof)
    d, _ = ks_binned(h1, h2)
    n1, n2 = np.sum(h1), np.sum(h2)
    p_ks = stats.kstwobign.sf(np.sqrt(n1 * n2 / (n1 + n2) / disp) * d)
    print(f"  {label}: chi2 {chi:.1f}/{dof} (dispersion {disp:.2f}) p {p_chi:.3g}, KS {d:.4f} p {p_ks:.3g}")
    assert p_chi > p_min and p_ks > p_min, (label, p_chi, p_ks)
    return p_chi, p_ks


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
