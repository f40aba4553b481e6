"""GPU tests of the driver: in-house status sort, XORWOW streams, one
do_iteration at a time against the oracle's GPU-driver semantics, the
buffered frame loop of gpu_main.cu, and full-run statistics (yields within
3 sigma, chi-square and KS on detected-energy spectra) of the iteration
driver and the persistent kernel against the oracle's CPU driver."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as ol
import runs
import nebula_amd as na
from nebula_amd import _lib as L
from nebula_amd import scenes, synth
from nebula_amd.driver import Driver, WorkPool, raster_scan

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 63, 4096, 4097, 1_000_003])
def test_status_sort_is_stable_counting_sort(n):
    print(f"[gpu] sort n={n}", flush=True)
    rng = np.random.default_rng(n)
    codes = np.array([L.ST_PENDING, L.ST_INELASTIC, L.ST_ELASTIC, L.ST_INTERSECT, L.ST_NO_EVENT,
                      L.ST_DETECTED, L.ST_NEW_SECONDARY, L.ST_TERMINATED], np.uint8)
    keys = rng.choice(codes, n).astype(np.uint8)
    out = np.zeros(n, np.uint32)
    L.call("nbl_sort_status_batch", L.ptr(keys), n, L.ptr(out))
    assert np.array_equal(out, np.argsort(keys & 3, kind="stable").astype(np.uint32))
    same = np.full(n, L.ST_TERMINATED, np.uint8)
    L.call("nbl_sort_status_batch", L.ptr(same), n, L.ptr(out))
    assert np.array_equal(out, np.arange(n, dtype=np.uint32))


def test_xorwow_streams_match_oracle():
    print("[gpu] rng", flush=True)
    n, draws = 10_000, 64
    a = np.zeros(6 * n, np.uint32); b = np.zeros(6 * n, np.uint32)
    L.call("nbl_rng_seed", 1234, 77, n, L.ptr(a))
    ol.orc.orc_rng_seed(1234, 77, n, ol.p(b))
    assert np.array_equal(a, b)
    ga = np.zeros(n * draws, np.float32); oa = np.zeros(n * draws, np.float32)
    L.call("nbl_rng_uniform_batch", L.ptr(a), n, draws, L.ptr(ga))
    ol.orc.orc_rng_uniform_batch(ol.p(b), n, draws, ol.p(oa))
    assert np.array_equal(ga, oa)


def _scene_lines():
    tris = scenes.line_trench()
    specs = [synth.SI, synth.PMMA]
    return tris, specs


def test_iterations_match_oracle_stepwise():
    """do_iteration on the GPU vs the oracle's restatement of
    gpu_driver.inl:312-567 from the same slot state, one iteration at a time
    (the oracle is re-synchronised to the GPU state after every step).
    Capacity is small so that secondaries run out of free slots and the
    PENDING path is taken."""
    print("[gpu] stepwise iterations", flush=True)
    tris, specs = _scene_lines()
    g = na.Geometry(tris); m = na.Materials(specs)
    og = ol.OrcGeometry(tris); om = ol.OrcMaterials([p.desc for p in m.prepared])
    cap = 3000
    drv = Driver(g, m, capacity=cap, seed=42)
    osl = ol.orc.orc_slots_create(cap, 42)
    try:
        rng = np.random.default_rng(1)
        n = 2000
        prim = np.zeros(n, L.PARTICLE_DTYPE)
        prim["pos"] = np.stack([rng.uniform(-200, 200, n), rng.uniform(-200, 200, n), np.full(n, 55.0)], 1)
        prim["dir"] = (0, 0, -1)
        prim["kin_energy"] = rng.uniform(100, 1500, n)
        tags = np.arange(n, dtype=np.uint32)
        assert drv.push(prim, tags) == n
        assert ol.orc.orc_slots_push(osl, ol.p(prim), ol.p(tags), n) == n
        max_e = m.max_energy()
        seen_pending = False
        for it in range(40):
            s0 = drv.read_slots()
            view = L.SlotView(**{k: ol.p(v) for k, v in s0.items()})
            ol.orc.orc_slots_write(osl, C.byref(view))
            drv.do_iteration()
            ol.orc.orc_slots_iteration(osl, om.h, og.h, 0.0, max_e)
            a = drv.read_slots()
            b = {k: np.zeros_like(v) for k, v in s0.items()}
            ol.orc.orc_slots_read(osl, C.byref(L.SlotView(**{k: ol.p(v) for k, v in b.items()})))
            assert np.array_equal(a["particle_idx"], b["particle_idx"]), f"iteration {it}: sort differs"
            st_eq = (a["status"] == b["status"]).mean()
            assert st_eq > 0.995, f"iteration {it}: status agreement {st_eq}"
            for k in ("tags", "material", "last_tri"):
                assert (a[k] == b[k]).mean() > 0.995, f"iteration {it}: {k}"
            active = (a["status"] != L.ST_TERMINATED) & (b["status"] != L.ST_TERMINATED)
            pa, pb = a["particles"][active], b["particles"][active]
            ok = np.logical_and.reduce([np.all(np.isclose(pa[f], pb[f], rtol=1e-4, atol=1e-4).reshape(len(pa), -1), axis=1)
                                        for f in ("pos", "dir", "kin_energy")])
            assert ok.mean() > 0.99, f"iteration {it}: particle agreement {ok.mean()}"
            # an event's RNG use is identical: the per-slot streams stay in lockstep
            assert (a["rng"].reshape(-1, 6) == b["rng"].reshape(-1, 6)).all(axis=1).mean() > 0.995
            seen_pending |= bool((a["status"] == L.ST_PENDING).any())
        assert seen_pending, "the PENDING path was not exercised"
    finally:
        ol.orc.orc_slots_destroy(osl)
        drv.close()


def test_buffered_frame_loop_runs_all_primaries():
    """gpu_main.cu:178-213: input buffer refills, detection snapshot, frames."""
    print("[gpu] frame loop", flush=True)
    g = na.Geometry(scenes.cube_slab()); m = na.Materials([synth.SI])
    n = 3000
    pool = WorkPool(runs.pencil_primaries(n, 1000.0), np.arange(n, dtype=np.uint32))
    detected = []
    with Driver(g, m, capacity=200_000, seed=7) as drv:
        drv.allocate_input_buffers(1500)
        drv.push_to_buffer(pool)
        for frame in range(100_000):
            drv.push_to_simulation()
            drv.buffer_detected()
            for _ in range(20):
                drv.do_iteration()
            drv.push_to_buffer(pool)
            running, det = drv.flush_buffered()
            detected += det
            if running == 0 and pool.done():
                break
        drv.buffer_detected()
        running, det = drv.flush_buffered()
        detected += det
        assert running == 0 and pool.done()
    tags = np.array([t for _, t in detected])
    assert tags.max() < n
    e = np.array([p["kin_energy"] for p, _ in detected])
    # BSE yield of 1 keV on the synthetic Si against the oracle, within 3
    # sigma with the per-primary variance of the frame loop's own records
    og = ol.OrcGeometry(scenes.cube_slab()); om = ol.OrcMaterials([m.prepared[0].desc])
    st, _ = runs.oracle_run(og, om, runs.pencil_primaries(4 * n, 1000.0), threads=8)
    mean, var = runs.per_primary(e, tags, n)["bse"]
    z = runs.yield_z((e >= 50).sum(), n, st["detected_bse"], st["primaries"], dispersion=max(1.0, var / mean))
    assert abs(z) < 3, z


def _per_primary_counts(det_tags, n):
    return np.bincount(det_tags, minlength=n)


@pytest.mark.parametrize("energy", [1000.0, 300.0])
def test_persistent_yields_and_spectra_match_oracle(energy):
    """Cube slab, pencil beam: SE/BSE yields within 3 sigma, chi-square and KS
    on the detected-energy histogram, persistent kernel vs oracle CPU driver."""
    print(f"[gpu] persistent yields {energy} eV", flush=True)
    tris = scenes.cube_slab()
    g = na.Geometry(tris); m = na.Materials([synth.SI])
    og = ol.OrcGeometry(tris); om = ol.OrcMaterials([m.prepared[0].desc])
    n_cpu = 12_000 if energy > 500 else 30_000
    st, h_cpu = runs.oracle_run(og, om, runs.pencil_primaries(n_cpu, energy), threads=16)
    # GPU: 400 pixels at the beam point, so per-pixel yields give the variance
    n_pix, per = 400, 500
    beam = raster_scan(np.zeros(20), np.zeros(20), 5.0, energy, per, 0.0, seed=3)
    with Driver(g, m, capacity=1 << 16, seed=11) as drv:
        t = drv.run_persistent(beam, per_pixel=True, hist=True)
    n_gpu = n_pix * per
    assert t["primaries"] == n_gpu and t["stopped"] == 0
    for key, pix in (("detected_se", "pixel_se"), ("detected_bse", "pixel_bse")):
        y_gpu = t[pix].astype(float) / per
        se_gpu = y_gpu.std(ddof=1) / np.sqrt(n_pix)
        y_cpu = st[key] / st["primaries"]
        # CPU standard error from the GPU's per-primary dispersion (same physics)
        se_cpu = y_gpu.std(ddof=1) * np.sqrt(per) / np.sqrt(st["primaries"])
        z = (y_gpu.mean() - y_cpu) / np.hypot(se_gpu, se_cpu)
        print(f"  {key}: gpu {y_gpu.mean():.4f} cpu {y_cpu:.4f} z {z:+.2f}")
        assert abs(z) < 3, (key, z)
    p, chi, dof = runs.chi2_hist(t["hist"], h_cpu)
    d, pks = runs.ks_binned(t["hist"], h_cpu)
    print(f"  spectrum chi2 p={p:.3g} (chi {chi:.1f}/{dof}), KS D={d:.4f} p={pks:.3g}")
    assert p > 1e-3 and pks > 1e-3


def test_iteration_driver_and_persistent_agree():
    """The two GPU transport forms (reference-semantics iterations and the
    persistent kernel) give the same yields within 3 sigma."""
    print("[gpu] iteration vs persistent", flush=True)
    tris = scenes.cube_slab()
    g = na.Geometry(tris); m = na.Materials([synth.SI])
    n = 20_000
    with Driver(g, m, capacity=1 << 20, seed=5) as drv:
        assert drv.push(runs.pencil_primaries(n, 1000.0), np.arange(n, dtype=np.uint32)) == n
        det = []
        for _ in range(100_000):
            drv.do_iteration()
            if drv.get_running_count() == 0:
                break
        det = drv.flush_detected()
        assert drv.get_running_count() == 0
    e = np.array([p["kin_energy"] for p, _ in det])
    a = runs.per_primary(e, np.array([tg for _, tg in det]), n)
    with Driver(g, m, capacity=1 << 16, seed=6) as drv:
        t = drv.run_persistent_list(runs.pencil_primaries(200_000, 1000.0), detected=1_600_000)
    b = runs.per_primary(t["detected"][0]["kin_energy"], t["detected"][1], 200_000)
    for key in ("se", "bse"):
        z = runs.z_from(a, b, key)
        print(f"  {key}: iteration {a[key][0]:.4f} persistent {b[key][0]:.4f} z {z:+.2f}")
        assert abs(z) < 3


def test_persistent_list_mode_and_detection_records():
    """Explicit primaries through the persistent kernel, detected electrons
    returned as records: the records agree with the tallies of the same
    launch (count, SE/BSE split, 1-eV histogram, detector plane, tags), the
    yields with the oracle's within 3 sigma; a too-small buffer keeps the
    count; bad primaries and per-pixel tallies are refused."""
    print("[gpu] persistent list mode + detection records", flush=True)
    tris = scenes.cube_slab()
    g = na.Geometry(tris); m = na.Materials([synth.SI])
    og = ol.OrcGeometry(tris); om = ol.OrcMaterials([m.prepared[0].desc])
    n_cpu = 40_000
    st, _, oe, ot = runs.oracle_records(og, om, runs.pencil_primaries(n_cpu, 1000.0), threads=16)
    opp = runs.per_primary(oe, ot, n_cpu)
    n = 200_000
    prim = runs.pencil_primaries(n, 1000.0)
    tags = (np.arange(n) % 1000).astype(np.uint32)
    with Driver(g, m, capacity=1 << 16, seed=17) as drv:
        t = drv.run_persistent_list(prim, tags, hist=True, detected=4 * n)
        parts, dtags = t["detected"]
        assert t["primaries"] == n and t["stopped"] == 0
        assert t["detected_count"] == len(parts) == t["detected_se"] + t["detected_bse"]
        e = parts["kin_energy"]
        assert (e < 50).sum() == t["detected_se"] and (e >= 50).sum() == t["detected_bse"]
        h = np.bincount(np.floor(e[e < L.HIST_BINS]).astype(np.int64), minlength=L.HIST_BINS)
        assert np.array_equal(h, t["hist"].astype(np.int64))
        assert np.all(dtags < 1000) and np.all(np.abs(parts["pos"][:, 2] - 10.0) < 1e-3)
        for key, k in (("detected_se", "se"), ("detected_bse", "bse")):
            mean, var = opp[k]
            z = runs.yield_z(t[key], n, st[key], st["primaries"], dispersion=max(1.0, var / mean))
            print(f"  {key}: list {t[key] / n:.4f} oracle {st[key] / st['primaries']:.4f} z {z:+.2f}")
            assert abs(z) < 3
        small = drv.run_persistent_list(prim[:20_000], detected=100)
        assert small["detected_count"] > 100 and len(small["detected"][0]) == 100
        bad = prim[:10].copy()
        bad["dir"][3] = 0
        with pytest.raises(na.NblError):
            drv.run_persistent_list(bad)


def test_inelastic_batching_keeps_yields_and_histograms():
    """Inelastic events wait in batches of 8 by default (nbl_persistent_opts
    inelastic_batch / inelastic_wait); against no batching (1), the default
    and a long wait give the same yields and detected-energy histogram
    within statistics, for every primary."""
    print("[gpu] inelastic batching", flush=True)
    tris = scenes.rough_surface(n=128)
    g = na.Geometry(tris); m = na.Materials([synth.SI])
    n = 400_000
    prim = runs.pencil_primaries(n, 500.0, pos=(0.0, 0.0, 20.0))
    res, pp = {}, {}
    with Driver(g, m, capacity=1024, seed=8) as drv:
        for name, kw in (("default", {}), ("off", dict(inelastic_batch=1)),
                         ("long", dict(inelastic_batch=32, inelastic_wait=64))):
            res[name] = drv.run_persistent_list(prim, hist=True, detected=8 * n, **kw)
            pp[name] = runs.per_primary(res[name]["detected"][0]["kin_energy"], res[name]["detected"][1], n)
    base = res["off"]
    for name in ("default", "long"):
        t = res[name]
        assert t["primaries"] == base["primaries"] and t["stopped"] == 0
        for k in ("se", "bse"):
            z = runs.z_from(pp[name], pp["off"], k)
            print(f"  {name} {k}: {pp[name][k][0]:.4f} vs {pp['off'][k][0]:.4f}, z {z:+.2f}")
            assert abs(z) < 3
        a, b = base["hist"][:600].astype(float), t["hist"][:600].astype(float)
        keep = (a + b) > 20
        chi2 = float(np.mean((a - b)[keep] ** 2 / (a + b)[keep]))
        print(f"  {name} histogram chi2/bin {chi2:.3f} over {keep.sum()} bins")
        assert chi2 < 1.5, chi2


def test_tiny_stack_waits_for_room_and_counts_drops():
    """With the smallest stack (64 LDS entries, 1 in HBM) 2 keV cascades run
    out of room: inelastic events wait for it (pending_retries), a wavefront
    whose 64 lanes all wait drops the secondaries that do not fit and counts
    them as stopped, and every primary still runs to the end.  Dropping is
    an error (NBL_ERR_RUNTIME naming the count) unless the caller allows it."""
    print("[gpu] tiny stack", flush=True)
    tris = scenes.cube_slab()
    g = na.Geometry(tris); m = na.Materials([synth.SI])
    beam = raster_scan(np.zeros(1), np.zeros(1), 5.0, 2000.0, 20_000, 0.0, seed=31)
    with Driver(g, m, capacity=1024, seed=12) as drv:
        with pytest.raises(L.NblError) as e:
            drv.run_persistent(beam, stack_depth=64, overflow_depth=1)
        assert e.value.code == L.NBL_ERR_RUNTIME and "electrons stopped" in str(e.value)
        t = drv.run_persistent(beam, stack_depth=64, overflow_depth=1, allow_drops=True, count_work=True)
        ref = drv.run_persistent(beam, count_work=True)
    print(f"  tiny: pending {t['pending_retries']} stopped {t['stopped']} secondaries {t['secondaries']}; "
          f"default: pending {ref['pending_retries']} stopped {ref['stopped']}")
    assert t["primaries"] == ref["primaries"] == 20_000
    assert t["pending_retries"] > 0 and ref["stopped"] == 0
    assert 0 < t["secondaries"] and t["stopped"] < t["secondaries"]
    assert t["detected_bse"] > 0 and t["detected_se"] > 0


def test_work_counters_on_request():
    """nbl_persistent_opts.count_work: the default launch leaves the work
    counters at 0 and still counts primaries, detections and stopped
    electrons; with count_work every work counter is filled, consistently
    (a step per event, two free-path lookups per step in a material)."""
    print("[gpu] work counters on request", flush=True)
    g = na.Geometry(scenes.cube_slab()); m = na.Materials([synth.SI])
    beam = raster_scan(np.zeros(1), np.zeros(1), 5.0, 1000.0, 50_000, 0.0, seed=3)
    work = ("events", "lookups_1d", "lookups_2d", "rounddowns", "node_steps", "tri_tests", "boundary_events",
            "secondaries")
    with Driver(g, m, capacity=1024, seed=8) as drv:
        a = drv.run_persistent(beam)
        b = drv.run_persistent(beam, count_work=True)
    for t in (a, b):
        assert t["primaries"] == 50_000 and t["stopped"] == 0
        assert t["detected_se"] > 0 and t["detected_bse"] > 0
    for k in work:
        assert a[k] == 0 and b[k] > 0, (k, a[k], b[k])
    assert b["events"] > b["boundary_events"] and b["lookups_1d"] <= 2 * b["events"]
    assert b["lookups_2d"] >= b["rounddowns"]
