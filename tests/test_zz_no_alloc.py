"""GPU tests of nbl_driver_run_persistent_list_noalloc, the persistent
transport that allocates nothing (primaries, tags and detection records in
the caller's device memory, the kernel's scratch in one fixed array per GPU),
against nbl_driver_run_persistent_list through the C ABI, each on two calls
in a row, the second with fewer primaries.

* Upward primaries in the vacuum gap under the detector plane: every one is
  detected unchanged, so the two entry points must give identical records
  (bit for bit, compared by tag).
* A 300 eV pencil beam into Si: per-primary SE and BSE yields with their own
  measured variances, compared at 3 sigma for the family of four z-scores
  (Bonferroni: |z| < 3.4 each), and the detected-energy spectra by
  chi-square and KS.
* The limits: an overflow depth beyond the scratch's stack space and an
  invalid primary are refused with an error, after which the driver runs.

Device memory for the caller's arrays comes from the HIP runtime the library
itself uses (hipMalloc through ctypes), not from the library."""
import ctypes as C

import numpy as np
import pytest

import runs
import nebula_amd as na
from nebula_amd import _lib as L
from nebula_amd import scenes, synth
from nebula_amd.driver import Driver

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    hip = C.CDLL("libamdhip64.so.7")             # the runtime libnbl_hip.so is linked against
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipDeviceSynchronize.argtypes = []
    return hip


class DevArray:
    """A caller-owned device copy of a numpy array (or zeroed space)."""

    def __init__(self, hip, like, upload=True):
        self.hip, self.nbytes, self.like = hip, like.nbytes, like
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), max(self.nbytes, 1)) == 0
        src = like if upload else np.zeros_like(like)
        assert hip.hipMemcpy(self.p, src.ctypes.data, self.nbytes, H2D) == 0

    def download(self):
        out = np.empty_like(self.like)
        assert self.hip.hipMemcpy(out.ctypes.data, self.p, self.nbytes, D2H) == 0
        return out

    def free(self):
        if self.p:
            self.hip.hipFree(self.p)
            self.p = C.c_void_p()


def _noalloc(drv, hip, prim, cap, **kw):
    """One call of the allocation-free entry point; returns (tallies, record particles, record tags)."""
    tags = np.arange(len(prim), dtype=np.uint32)
    dp, dt = DevArray(hip, prim), DevArray(hip, tags)
    rp = DevArray(hip, np.zeros(cap, L.PARTICLE_DTYPE), upload=False)
    rt = DevArray(hip, np.zeros(cap, np.uint32), upload=False)
    try:
        t = drv.run_persistent_list_noalloc(dp.p.value, dt.p.value, len(prim), rp.p.value, rt.p.value, cap, **kw)
        k = min(t["detected_se"] + t["detected_bse"], cap)
        return t, rp.download()[:k], rt.download()[:k]
    finally:
        for a in (dp, dt, rp, rt):
            a.free()


def _sorted_records(parts, tags):
    order = np.lexsort((parts["kin_energy"], parts["pos"][:, 0], tags))
    return parts[order], tags[order]


def _upward(n, seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(n, L.PARTICLE_DTYPE)
    p["pos"][:, 0] = rng.uniform(-900, 900, n)
    p["pos"][:, 1] = rng.uniform(-900, 900, n)
    p["pos"][:, 2] = rng.uniform(1, 9, n)                   # vacuum gap between the cube (z<0) and the detector (z=10)
    d = np.stack([rng.uniform(-0.3, 0.3, n), rng.uniform(-0.3, 0.3, n), np.ones(n)], 1)
    p["dir"] = (d / np.linalg.norm(d, axis=1)[:, None]).astype(np.float32)
    p["kin_energy"] = rng.uniform(5, 2000, n).astype(np.float32)
    return p


def test_noalloc_records_equal_list_records():
    """Deterministic transport (vacuum to the detector): identical records."""
    print("[gpu] noalloc: vacuum records, two calls", flush=True)
    hip = _hip()
    m = na.Materials([synth.SI])
    g = na.Geometry(scenes.cube_slab())
    with Driver(g, m, capacity=1024, seed=5) as da, Driver(g, m, capacity=1024, seed=5) as db:
        for n, seed in ((60_000, 1), (7_000, 2)):          # the second call is smaller
            prim = _upward(n, seed)
            ta = da.run_persistent_list(prim, detected=n + 16, count_work=True)
            tb, pb, gb = _noalloc(db, hip, prim, n + 16, count_work=True)
            assert ta["primaries"] == tb["primaries"] == n
            assert ta["detected_count"] == n and tb["detected_se"] + tb["detected_bse"] == n
            assert (ta["detected_se"], ta["detected_bse"]) == (tb["detected_se"], tb["detected_bse"])
            assert ta["events"] == tb["events"] > 0 and ta["stopped"] == tb["stopped"] == 0
            pa, ga = _sorted_records(*ta["detected"])
            pb, gb = _sorted_records(pb, gb)
            assert np.array_equal(ga, gb) and np.array_equal(ga, np.arange(n, dtype=np.uint32))
            assert np.array_equal(pa.view(np.uint8), pb.view(np.uint8))
            assert np.array_equal(pb["kin_energy"], prim["kin_energy"][gb])   # detected unchanged


def test_noalloc_yields_and_spectra_match_list():
    """300 eV pencil beam into Si (the energy of the rough-surface workload,
    whose cascades fit the scratch's stacks at the default launch shape):
    the two entry points agree statistically."""
    print("[gpu] noalloc: Si pencil beam, two calls", flush=True)
    hip = _hip()
    m = na.Materials([synth.SI])
    g = na.Geometry(scenes.cube_slab())
    z_bar = 3.4                                             # 3 sigma for the family of four (Bonferroni)
    with Driver(g, m, capacity=1024, seed=11) as da, Driver(g, m, capacity=1024, seed=12) as db:
        for n in (40_000, 12_000):                          # the second call is smaller
            prim = runs.pencil_primaries(n, 300.0)
            cap = 8 * n
            ta = da.run_persistent_list(prim, detected=cap, hist=True)
            tb, pb, gb = _noalloc(db, hip, prim, cap, hist=True)
            assert ta["primaries"] == tb["primaries"] == n
            assert ta["stopped"] == tb["stopped"] == 0
            k = tb["detected_se"] + tb["detected_bse"]
            assert len(gb) == k <= cap and (gb < n).all()
            assert int(np.sum(tb["hist"])) == int(np.sum(pb["kin_energy"] < L.HIST_BINS))
            pa, ga = ta["detected"]
            ya = runs.per_primary(pa["kin_energy"], ga, n)
            yb = runs.per_primary(pb["kin_energy"], gb, n)
            for key in ("se", "bse"):
                z = runs.z_from(ya, yb, key)
                print(f"  n={n} {key}: list {ya[key][0]:.4f} noalloc {yb[key][0]:.4f} z {z:+.2f}", flush=True)
                assert abs(z) < z_bar, (n, key, z)
            if n == 40_000:
                runs.spectra_agree(ta["hist"], tb["hist"], "list vs noalloc",
                                   disp=runs.dispersion(pb["kin_energy"], gb, n))


def test_noalloc_refuses_what_does_not_fit():
    print("[gpu] noalloc: limits", flush=True)
    hip = _hip()
    m = na.Materials([synth.SI])
    g = na.Geometry(scenes.cube_slab())
    with Driver(g, m, capacity=1024, seed=3) as d:
        prim = runs.pencil_primaries(1000, 300.0)
        with pytest.raises(L.NblError, match="exceeds the scratch"):
            _noalloc(d, hip, prim, 0, overflow_depth=1 << 20)
        bad = prim.copy()
        bad["dir"][17] = 0.0
        with pytest.raises(L.NblError, match="zero direction"):
            _noalloc(d, hip, bad, 0)
        # after the refusals the same driver still runs
        t, parts, tags = _noalloc(d, hip, prim, 8000)
        assert t["primaries"] == 1000 and len(tags) == t["detected_se"] + t["detected_bse"]
