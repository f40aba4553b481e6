"""The nbl_gpu command (the reference's nebula_gpu front end: .tri + .pri +
materials in, 36-byte detection records out, prescan + frame loop through
the iteration driver) end to end on the GPU, against the persistent kernel
on the same geometry, materials and beam (yields within 4 sigma)."""
import os
import subprocess

import numpy as np
import pytest

import runs
import nebula_amd as na
from nebula_amd import _lib as L
from nebula_amd import files, scenes, synth
from nebula_amd.driver import Driver, raster_scan

pytestmark = pytest.mark.gpu

CLI = os.path.join(os.path.dirname(L.LIB_PATH), "nbl_gpu")


def test_cli_runs_and_matches_persistent_yields(tmp_path):
    print("[gpu] nbl_gpu end to end", flush=True)
    tris = scenes.cube_slab()
    a = tris["r0"].astype(np.float64)
    files.write_tri_file(tmp_path / "g.tri", tris["material_in"], tris["material_out"],
                         a, a + tris["e1"], a + tris["e2"], comment="flat Si slab + detector")
    n = 20_000
    prim = runs.pencil_primaries(n, 1000.0)
    pix = np.stack([np.arange(n) % 100, np.arange(n) // 100], axis=1).astype(np.int32)
    files.write_pri_file(tmp_path / "p.pri", prim, pix)
    synth.write_material_file(synth.SI, tmp_path / "si.nblmat")
    r = subprocess.run([CLI, "--capacity", "200000", "--prescan-size=1000", "--seed", "7",
                        str(tmp_path / "g.tri"), str(tmp_path / "p.pri"), str(tmp_path / "si.nblmat")],
                       capture_output=True, timeout=300)
    err = r.stderr.decode(errors="replace")
    print(err[-600:])
    assert r.returncode == 0, err
    rec = files.read_records(r.stdout)
    assert len(r.stdout) == 36 * len(rec) and len(rec) > 0
    assert np.all((rec["pixel"][:, 0] >= 0) & (rec["pixel"][:, 0] < 100))
    assert np.all((rec["pixel"][:, 1] >= 0) & (rec["pixel"][:, 1] < n // 100))
    assert np.all(np.abs(rec["pos"][:, 2] - 10.0) < 1e-3)      # all detected on the detector plane
    se, bse = int((rec["kin_energy"] < 50).sum()), int((rec["kin_energy"] >= 50).sum())
    g = na.Geometry(tris); m = na.Materials([synth.SI])
    with Driver(g, m, capacity=1 << 16, seed=3) as drv:
        t = drv.run_persistent(raster_scan(np.zeros(1), np.zeros(1), 5.0, 1000.0, 200_000, 0.0, seed=5))
    # per-primary variances from the records (pixel x + 100 y is the primary's index)
    pp = runs.per_primary(rec["kin_energy"], rec["pixel"][:, 0] + 100 * rec["pixel"][:, 1], n)
    for k_cli, k_p, key in ((se, t["detected_se"], "se"), (bse, t["detected_bse"], "bse")):
        z = runs.yield_z(k_cli, n, k_p, t["primaries"], dispersion=max(1.0, pp[key][1] / pp[key][0]))
        print(f"  cli {k_cli / n:.4f} persistent {k_p / t['primaries']:.4f} z {z:+.2f}")
        assert abs(z) < 3


def test_cli_usage_and_errors(tmp_path):
    r = subprocess.run([CLI], capture_output=True, timeout=60)
    assert r.returncode == 1 and b"Usage" in r.stderr
    r = subprocess.run([CLI, "--bogus", "1", "a", "b", "c"], capture_output=True, timeout=60)
    assert r.returncode == 1 and b"Unexpected key 'bogus'" in r.stderr
    (tmp_path / "g.tri").write_text("0 -123 0 0 0 1 0 0 0 1 0\n2 -123 0 0 0 1 0 0 0 1 0\n")
    (tmp_path / "p.pri").write_bytes(b"")
    synth.write_material_file(synth.SI, tmp_path / "si.nblmat")
    r = subprocess.run([CLI, str(tmp_path / "g.tri"), str(tmp_path / "p.pri"), str(tmp_path / "si.nblmat")],
                       capture_output=True, timeout=120)
    assert r.returncode == 1 and b"not enough materials" in r.stderr


def _cli(*args, timeout=300):
    r = subprocess.run([CLI, *map(str, args)], capture_output=True, timeout=timeout)
    err = r.stderr.decode(errors="replace")
    assert r.returncode == 0, err
    return r.stdout, err


def test_cli_reads_hdf5_materials(tmp_path):
    """The reference's HDF5 material layout (written by nbl_mat_write) in place
    of .nblmat: with --inelastic kieft the detected records are the same as
    from the .nblmat file (same tables bit for bit, same seed); by default the
    file's full Penn tables are used, and the yields match the persistent
    kernel's full Penn run."""
    print("[gpu] nbl_gpu with HDF5 materials", flush=True)
    pytest.importorskip("nebula_amd.matfile")
    from nebula_amd import matfile
    if not os.path.exists(os.path.join(os.path.dirname(L.LIB_PATH), "libnbl_mat.so")):
        pytest.skip("libnbl_mat.so not built")
    tris = scenes.cube_slab()
    a = tris["r0"].astype(np.float64)
    files.write_tri_file(tmp_path / "g.tri", tris["material_in"], tris["material_out"], a, a + tris["e1"],
                         a + tris["e2"])
    n = 20_000
    prim = runs.pencil_primaries(n, 1000.0)
    pix = np.stack([np.arange(n) % 100, np.arange(n) // 100], axis=1).astype(np.int32)
    files.write_pri_file(tmp_path / "p.pri", prim, pix)
    synth.write_material_file(synth.SI, tmp_path / "si.nblmat")
    matfile.write_material(tmp_path / "si.mat", synth.SI)
    common = ["--capacity", "200000", "--seed", "7", tmp_path / "g.tri", tmp_path / "p.pri"]
    out_nbl, _ = _cli(*common, tmp_path / "si.nblmat")
    out_h5, err = _cli("--inelastic", "kieft", *common, tmp_path / "si.mat")
    assert "Kieft-Bosch inelastic" in err
    ra, rb = files.read_records(out_nbl), files.read_records(out_h5)
    key = lambda r: np.lexsort([r["kin_energy"], r["pos"][:, 0], r["pos"][:, 1], r["pixel"][:, 0], r["pixel"][:, 1]])
    assert len(ra) == len(rb) and len(ra) > 0
    assert ra[key(ra)].tobytes() == rb[key(rb)].tobytes()
    out_p, err = _cli(*common, tmp_path / "si.mat")
    assert "full Penn inelastic" in err
    rec = files.read_records(out_p)
    se, bse = int((rec["kin_energy"] < 50).sum()), int((rec["kin_energy"] >= 50).sum())
    g = na.Geometry(tris); m = na.Materials([synth.SI_PENN])
    with Driver(g, m, capacity=1 << 16, seed=3) as drv:
        t = drv.run_persistent(raster_scan(np.zeros(1), np.zeros(1), 5.0, 1000.0, 200_000, 0.0, seed=5))
    pp = runs.per_primary(rec["kin_energy"], rec["pixel"][:, 0] + 100 * rec["pixel"][:, 1], n)
    for k_cli, k_p, key in ((se, t["detected_se"], "se"), (bse, t["detected_bse"], "bse")):
        z = runs.yield_z(k_cli, n, k_p, t["primaries"], dispersion=max(1.0, pp[key][1] / pp[key][0]))
        print(f"  cli (Penn, HDF5) {k_cli / n:.4f} persistent {k_p / t['primaries']:.4f} z {z:+.2f}")
        assert abs(z) < 3
    # a .nblmat file has no full Penn tables
    r = subprocess.run([CLI, "--inelastic", "penn", *map(str, common), str(tmp_path / "si.nblmat")],
                       capture_output=True, timeout=120)
    assert r.returncode == 1 and b"Full Penn model not found" in r.stderr


def test_cli_persistent_mode(tmp_path):
    """--persistent 1: the same files through the persistent kernel; the
    records sit on the detector with valid pixels and the yields agree with
    the frame loop's (same fresh seed, different random streams) within 3 sigma,
    each side's per-primary variance from its records."""
    print("[gpu] nbl_gpu --persistent", flush=True)
    tris = scenes.cube_slab()
    a = tris["r0"].astype(np.float64)
    files.write_tri_file(tmp_path / "g.tri", tris["material_in"], tris["material_out"], a, a + tris["e1"],
                         a + tris["e2"])
    n = 50_000
    prim = runs.pencil_primaries(n, 1000.0)
    pix = np.stack([np.arange(n) % 100, np.arange(n) // 100], axis=1).astype(np.int32)
    files.write_pri_file(tmp_path / "p.pri", prim, pix)
    synth.write_material_file(synth.SI, tmp_path / "si.nblmat")
    # The frame loop's outcome is a fixed function of its seed; a fixed seed
    # would compare every run against the same frame-loop sample, so that a
    # sample one sigma off the mean fails the 3-sigma bar far more often than
    # 0.27 %.  A fresh seed per run (printed) keeps the test a two-sample test.
    seed = int(np.random.SeedSequence().entropy % (1 << 31))
    print(f"  seed {seed}", flush=True)
    common = ["--capacity", "200000", "--seed", str(seed), tmp_path / "g.tri", tmp_path / "p.pri",
              tmp_path / "si.nblmat"]
    out_p, err = _cli("--persistent", "1", *common)
    assert "persistent" in err
    out_f, _ = _cli(*common)
    rp, rf = files.read_records(out_p), files.read_records(out_f)
    assert len(rp) > 0 and np.all(np.abs(rp["pos"][:, 2] - 10.0) < 1e-3)
    assert np.all((rp["pixel"][:, 0] >= 0) & (rp["pixel"][:, 0] < 100) & (rp["pixel"][:, 1] >= 0) &
                  (rp["pixel"][:, 1] < n // 100))
    idx = lambda r: r["pixel"][:, 0] + 100 * r["pixel"][:, 1]
    a, b = runs.per_primary(rp["kin_energy"], idx(rp), n), runs.per_primary(rf["kin_energy"], idx(rf), n)
    for key in ("se", "bse"):
        z = runs.z_from(a, b, key)
        print(f"  {key}: persistent {a[key][0]:.4f} frame loop {b[key][0]:.4f} z {z:+.2f}")
        assert abs(z) < 3
