"""Full Penn inelastic model (physics/full_penn.h) on the GPU against the
reference's table_3D (golden fixture) and the CPU oracle: 3D table lookups
bit-exact; single Penn events within the physics tolerance; whole runs
(SE/BSE yields, detected spectrum) statistically."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle_lib as ol
import runs
import nebula_amd as na
from nebula_amd import _lib as L
from nebula_amd import scenes, synth
from nebula_amd.driver import Driver, raster_scan
from test_gpu_unit import PHYS_RTOL, PHYS_ATOL, close_particles, oracle_physics, random_particles
from test_penn import PENN_SPECS, _bits, _dummy_desc

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_table3d_matches_reference_fixture_and_oracle():
    print("[gpu] penn 3D tables", flush=True)
    f = np.load(os.path.join(GOLDEN, "tables3d.npz"), allow_pickle=False)
    d3 = np.ascontiguousarray(f["d3"])
    desc, keep = _dummy_desc(d3, f["scale"])
    arr = (L.MaterialDesc * 1)(desc)
    h = C.c_void_p()
    L.call("nbl_materials_create", arr, 1, C.byref(h))
    try:
        x, y, z = f["x"], f["y"], f["z"]
        for tab, rd, key in ((L.TAB_PENN_Q_ICDF, 0, "o_get"), (L.TAB_PENN_IONISATION, 1, "o_rounddown"),
                             (L.TAB_PENN_IONISATION, 0, "o_get")):
            o = np.zeros(len(x), np.float32)
            L.call("nbl_table3d_get_batch", h, 0, tab, rd, L.ptr(x), L.ptr(y), L.ptr(z), len(x), L.ptr(o))
            assert np.array_equal(_bits(o), _bits(f[key])), key
    finally:
        L.lib.nbl_materials_destroy(h)
    m = na.Materials(PENN_SPECS)
    om = ol.OrcMaterials([p.desc for p in m.prepared])
    rng = np.random.default_rng(9)
    n = 100_000
    x = rng.uniform(-0.5, 10, n).astype(np.float32)
    y = rng.uniform(-10, 1.1, n).astype(np.float32)
    z = rng.uniform(-0.1, 1.1, n).astype(np.float32)
    for mat in range(len(PENN_SPECS)):
        for tab in (L.TAB_PENN_Q_ICDF, L.TAB_PENN_IONISATION):
            for rd in (0, 1):
                a = m.table3d(mat, tab, x, y, z, rounddown=rd)
                b = np.zeros(n, np.float32)
                ol.orc.orc_table3d_get_batch(om.h, mat, tab, rd, ol.p(x), ol.p(y), ol.p(z), n, ol.p(b))
                assert np.array_equal(_bits(a), _bits(b)), (mat, tab, rd)
    k = na.Materials([synth.SI])                 # 3D lookups on a Kieft material are rejected
    with pytest.raises(L.NblError):
        k.table3d(0, L.TAB_PENN_Q_ICDF, x[:4], y[:4], z[:4])


def test_penn_physics_matches_oracle():
    """nbl_physics_batch(INELASTIC) on full Penn materials vs the oracle with
    the same injected uniforms; the bar is the one of the Kieft events."""
    print("[gpu] penn physics", flush=True)
    m = na.Materials(PENN_SPECS)
    om = ol.OrcMaterials([p.desc for p in m.prepared])
    rng = np.random.default_rng(321)
    n = 200_000
    parts = random_particles(n, rng)
    mat = rng.integers(0, len(PENN_SPECS), n).astype(np.int32)
    u = np.maximum(rng.uniform(0, 1, (n, L.NUM_UNIFORMS)), 2.0 ** -24).astype(np.float32)
    g = m.physics(L.EVENT_INELASTIC, parts, mat, u)
    o = oracle_physics(om, None, L.EVENT_INELASTIC, parts, mat, u)
    p2 = parts.copy()
    p2["kin_energy"] = np.nextafter(p2["kin_energy"], np.float32(np.inf))
    o2 = oracle_physics(om, None, L.EVENT_INELASTIC, p2, mat, u)
    ok = close_particles(g["particles"], o["particles"])
    ok_ulp = close_particles(o2["particles"], o["particles"])
    print(f"  primary within tol: gpu {ok.mean():.6f}, oracle+1ulp {ok_ulp.mean():.6f}")
    assert ok.mean() > 0.9995 or (ok.mean() > 0.995 and ok.sum() >= ok_ulp.sum())
    agree = (g["has_secondary"] == o["has_secondary"]).mean()
    print(f"  secondary created: gpu {g['has_secondary'].mean():.4f} oracle {o['has_secondary'].mean():.4f} "
          f"agreement {agree:.6f}")
    assert agree > 0.9995
    both = (g["has_secondary"] == 1) & (o["has_secondary"] == 1)
    assert both.sum() > n // 4
    sok = close_particles(g["secondary"], o["secondary"], both)
    both2 = (o2["has_secondary"] == 1) & (o["has_secondary"] == 1)
    sok_ulp = close_particles(o2["secondary"], o["secondary"], both2)
    print(f"  secondary within tol: gpu {sok.mean():.6f}, oracle+1ulp {sok_ulp.mean():.6f}")
    assert sok.mean() > 0.9995 or (sok.mean() > 0.99 and sok.mean() >= sok_ulp.mean())


@pytest.mark.parametrize("energy", [1000.0, 300.0])
def test_penn_persistent_yields_and_spectra_match_oracle(energy):
    """Si (full Penn) slab under a pencil beam: SE/BSE yields within 3 sigma,
    chi-square and KS on the detected spectrum, persistent kernel vs oracle."""
    print(f"[gpu] penn persistent yields {energy} eV", flush=True)
    tris = scenes.cube_slab()
    g = na.Geometry(tris); m = na.Materials([synth.SI_PENN])
    og = ol.OrcGeometry(tris); om = ol.OrcMaterials([m.prepared[0].desc])
    n_cpu = 12_000 if energy > 500 else 30_000
    st, h_cpu = runs.oracle_run(og, om, runs.pencil_primaries(n_cpu, energy), threads=16)
    n_pix, per = 400, 500
    beam = raster_scan(np.zeros(20), np.zeros(20), 5.0, energy, per, 0.0, seed=4)
    with Driver(g, m, capacity=1 << 16, seed=12) as drv:
        t = drv.run_persistent(beam, per_pixel=True, hist=True)
    n_gpu = n_pix * per
    assert t["primaries"] == n_gpu and t["stopped"] == 0
    for key, pix in (("detected_se", "pixel_se"), ("detected_bse", "pixel_bse")):
        y_gpu = t[pix].astype(float) / per
        se_gpu = y_gpu.std(ddof=1) / np.sqrt(n_pix)
        y_cpu = st[key] / st["primaries"]
        se_cpu = y_gpu.std(ddof=1) * np.sqrt(per) / np.sqrt(st["primaries"])
        z = (y_gpu.mean() - y_cpu) / np.hypot(se_gpu, se_cpu)
        print(f"  {key}: gpu {y_gpu.mean():.4f} cpu {y_cpu:.4f} z {z:+.2f}")
        assert abs(z) < 3, (key, z)
    p, chi, dof = runs.chi2_hist(t["hist"], h_cpu)
    d, pks = runs.ks_binned(t["hist"], h_cpu)
    print(f"  spectrum chi2 p={p:.3g} (chi {chi:.1f}/{dof}), KS D={d:.4f} p={pks:.3g}")
    assert p > 1e-3 and pks > 1e-3


def test_penn_iteration_driver_matches_persistent():
    """The reference-semantics iteration driver runs full Penn too; its
    yields agree with the persistent kernel's."""
    print("[gpu] penn iteration vs persistent", flush=True)
    tris = scenes.cube_slab()
    g = na.Geometry(tris); m = na.Materials([synth.SI_PENN])
    n = 20_000
    with Driver(g, m, capacity=1 << 20, seed=5) as drv:
        assert drv.push(runs.pencil_primaries(n, 1000.0), np.arange(n, dtype=np.uint32)) == n
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
