"""GPU tests of the scan workloads: per-pixel tallies of a raster over the
line/trench sample against the oracle at selected pixels, the batched uneven
3-material scan in one persistent launch, and the autotuner."""
import numpy as np
import pytest

import oracle_lib as ol
import runs
import nebula_amd as na
from nebula_amd import _lib as L
from nebula_amd import autotune, scenes, synth, workloads
from nebula_amd.driver import Driver, raster_scan

pytestmark = pytest.mark.gpu


def _oracle_pixel(og, om, x, y, z, energy, n, sigma, seed):
    rng = np.random.default_rng(seed)
    p = runs.pencil_primaries(n, energy, pos=(x, y, z))
    p["pos"][:, 0] += rng.normal(0, sigma, n)
    p["pos"][:, 1] += rng.normal(0, sigma, n)
    st, _ = runs.oracle_run(og, om, p, threads=16)
    return st


def test_line_scan_pixels_match_oracle():
    """Raster across one line (trench, edge, top): per-pixel SE and BSE
    yields of the persistent kernel within 3 sigma of the oracle CPU driver."""
    print("[gpu] line scan pixels", flush=True)
    tris, specs, _ = workloads.lines()
    g = na.Geometry(tris); m = na.Materials(specs)
    og = ol.OrcGeometry(tris); om = ol.OrcMaterials([p.desc for p in m.prepared])
    xs = np.array([-32.0, -17.0, -16.0, 0.0], np.float32)      # trench, edge-1, edge, line top (line at 0)
    per = 20_000
    beam = raster_scan(xs, np.zeros(1, np.float32), scenes.LINE_HEIGHT + 5.0, 500.0, per, 1.0, seed=4)
    with Driver(g, m, capacity=4096, seed=8) as d:
        t = d.run_persistent(beam, per_pixel=True)
    assert t["stopped"] == 0
    for k, x in enumerate(xs):
        # as many oracle primaries as the GPU's, so that the oracle's own
        # sampling error does not dominate z at the edge pixels
        st = _oracle_pixel(og, om, float(x), 0.0, scenes.LINE_HEIGHT + 5.0, 500.0, per, 1.0, k)
        for gk, ck in (("pixel_se", "detected_se"), ("pixel_bse", "detected_bse")):
            z = runs.yield_z(int(t[gk][k]), per, st[ck], st["primaries"])
            print(f"  x={x:+.0f} {ck}: gpu {t[gk][k] / per:.4f} cpu {st[ck] / st['primaries']:.4f} z {z:+.2f}")
            assert abs(z) < 3.5, (x, ck, z)


def test_batched_uneven_scan_tallies():
    """4096 pixels, Poisson(200) primaries each, 200..2000 eV, 3 materials:
    one launch; per-pixel tallies add up to the totals; yields vs the oracle
    on a sample of the same pixels."""
    print("[gpu] batched uneven scan", flush=True)
    tris, specs, beam = workloads.scan3()
    g = na.Geometry(tris); m = na.Materials(specs)
    with Driver(g, m, capacity=4096, seed=9) as d:
        t = d.run_persistent(beam, per_pixel=True, hist=True)
    px = beam["pixels"]
    total = int(px["n_primaries"].sum())
    assert t["primaries"] == total and t["stopped"] == 0
    assert int(t["pixel_se"].sum()) == t["detected_se"] and int(t["pixel_bse"].sum()) == t["detected_bse"]
    # oracle on 2% of the primaries, drawn pixel by pixel from the same batch
    og = ol.OrcGeometry(tris); om = ol.OrcMaterials([p.desc for p in m.prepared])
    rng = np.random.default_rng(21)
    sel = rng.choice(len(px), 80, replace=False)
    prims = []
    for i in sel:
        n = int(px["n_primaries"][i])
        p = runs.pencil_primaries(n, float(px["energy"][i]), pos=(float(px["x"][i]), float(px["y"][i]), float(px["z"][i])))
        p["pos"][:, 0] += rng.normal(0, 1.0, n); p["pos"][:, 1] += rng.normal(0, 1.0, n)
        prims.append(p)
    prims = np.concatenate(prims)
    st, _ = runs.oracle_run(og, om, prims, threads=16)
    n_sel = int(px["n_primaries"][sel].sum())
    for gk, ck in (("pixel_se", "detected_se"), ("pixel_bse", "detected_bse")):
        k_gpu = int(t[gk][sel].sum())
        z = runs.yield_z(k_gpu, n_sel, st[ck], st["primaries"])
        print(f"  {ck}: gpu {k_gpu / n_sel:.4f} cpu {st[ck] / st['primaries']:.4f} z {z:+.2f}")
        assert abs(z) < 4, (ck, z)


def test_autotuner_returns_a_valid_config():
    print("[gpu] autotune", flush=True)
    tris, specs, beam = workloads.cube(n=20_000)
    g = na.Geometry(tris); m = na.Materials(specs)
    space = {"block_size": (128, 256), "blocks_per_cu": (2, 4), "stack_depth": (64,), "refill_threshold": (1, 16)}
    with Driver(g, m, capacity=1024, seed=3) as d:
        best, results = autotune.tune(d, beam, sample=20_000, space=space, use_cache=False)
        assert autotune.valid(best) and len(results) == 8
        t = d.run_persistent(beam, **best)
    assert t["primaries"] == 20_000 and t["stopped"] == 0
