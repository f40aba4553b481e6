"""GPU tests of the scan workloads: the batched uneven 3-material scan in
one persistent launch (its per-pixel tallies add up to the totals) and the
autotuner.  Yields and spectra of the line scan and of the batched scan
against the oracle are in tests/test_gpu_configs.py."""
import numpy as np
import pytest

import oracle_lib as ol
import runs
import nebula_amd as na
from nebula_amd import _lib as L
from nebula_amd import autotune, scenes, synth, workloads
from nebula_amd.driver import Driver, raster_scan

pytestmark = pytest.mark.gpu


def test_batched_uneven_scan_tallies():
    """4096 pixels, Poisson(200) primaries each, 200..2000 eV, 3 materials:
    one launch; per-pixel tallies add up to the totals."""
    print("[gpu] batched uneven scan", flush=True)
    tris, specs, beam = workloads.scan3()
    g = na.Geometry(tris); m = na.Materials(specs)
    with Driver(g, m, capacity=4096, seed=9) as d:
        t = d.run_persistent(beam, per_pixel=True, hist=True)
    px = beam["pixels"]
    total = int(px["n_primaries"].sum())
    assert t["primaries"] == total and t["stopped"] == 0
    assert int(t["pixel_se"].sum()) == t["detected_se"] and int(t["pixel_bse"].sum()) == t["detected_bse"]
    # yields and spectra of a sample of these pixels against the oracle:
    # tests/test_gpu_configs.py::test_batched_scan_sample_spectrum_matches_oracle


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
