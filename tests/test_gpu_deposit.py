"""GPU energy-deposition tally (the counterpart of the reference's
energydeposit_cpu_driver) against the oracle's restatement of the same
rule (a scatter event that changes the electron's energy deposits the
change at the event point), on a flat Si slab under a 1 keV pencil beam:
total deposited energy, deposit count and the depth profile, statistically;
plus invariants of the tally itself."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as ol
import runs
import nebula_amd as na
from nebula_amd import scenes, synth
from nebula_amd.driver import Driver, raster_scan

pytestmark = pytest.mark.gpu

SHAPE = (1, 1, 40)                         # nx, ny, nz: a depth profile in 5-nm slices
LO, HI = (-500.0, -500.0, -200.0), (500.0, 500.0, 0.0)


def _oracle(og, om, n, seed):
    prim = runs.pencil_primaries(n, 1000.0)
    tags = np.arange(n, dtype=np.uint32)
    grid = np.zeros(SHAPE[::-1], np.float64)
    tot = C.c_double(); cnt = C.c_uint64(); st = ol.RunStats()
    lo = np.asarray(LO, np.float32); hi = np.asarray(HI, np.float32)
    rc = ol.orc.orc_cpu_mt_run_deposit(og.h, om.h, ol.p(prim), ol.p(tags), n, 0.0, seed, 16, None, C.byref(st),
                                       *SHAPE, ol.p(lo), ol.p(hi), ol.p(grid), C.byref(tot), C.byref(cnt))
    assert rc == 0
    return grid, tot.value, cnt.value, {f: getattr(st, f) for f, _ in ol.RunStats._fields_}


@pytest.mark.parametrize("spec", [synth.SI, synth.SI_PENN], ids=lambda s: s.inelastic)
def test_deposit_tally_matches_oracle(spec):
    """Per-primary means over independent runs: 10 oracle runs of 3000
    primaries, 20 GPU launches of 20000; every depth bin, the total and the
    deposit count compared by Welch z at a family-wise 3 sigma (two-sided
    0.27 %, Bonferroni over the k quantities: |z| < 3.96 for k = 15).
    Kieft-Bosch and full Penn Si (the kernel's deposit variant with each
    inelastic model)."""
    print(f"[gpu] energy deposition vs oracle ({spec.inelastic})", flush=True)
    tris = scenes.cube_slab()
    g = na.Geometry(tris); m = na.Materials([spec])
    og = ol.OrcGeometry(tris); om = ol.OrcMaterials([m.prepared[0].desc])
    n_cpu, n_gpu = 3000, 20_000
    cpu = []
    for k in range(10):
        cgrid, ctot, ccnt, cst = _oracle(og, om, n_cpu, 1234 + k)
        # invariants of the oracle's tally: one deposit per energy-changing
        # elastic or inelastic event (nearly all of them), all deposits positive
        assert 0.99 * cst["lookups_2d"] <= ccnt <= cst["lookups_2d"]
        assert np.all(cgrid >= 0) and cgrid.sum() <= ctot * (1 + 1e-12)
        cpu.append(np.concatenate([cgrid.ravel(), [ctot, ccnt]]) / n_cpu)
    gpu = []
    with Driver(g, m, capacity=1 << 16, seed=99) as drv:
        for k in range(20):
            t = drv.run_persistent(raster_scan(np.zeros(1), np.zeros(1), 5.0, 1000.0, n_gpu, 0.0, seed=100 + k),
                                   deposit=dict(shape=SHAPE, lo=LO, hi=HI))
            assert t["primaries"] == n_gpu and t["stopped"] == 0
            assert 0.99 * t["lookups_2d"] <= t["deposit_count"] <= t["lookups_2d"]
            assert np.all(t["deposit"] >= 0) and t["deposit"].sum() <= t["deposit_total"] * (1 + 1e-9)
            gpu.append(np.concatenate([t["deposit"].ravel(), [t["deposit_total"], t["deposit_count"]]]) / n_gpu)
    cpu, gpu = np.array(cpu), np.array(gpu)
    se = np.sqrt(cpu.var(axis=0, ddof=1) / len(cpu) + gpu.var(axis=0, ddof=1) / len(gpu))
    # depth bins that hold at least 1 % of the largest bin's energy (the deep,
    # rarely reached bins have too few deposits per run for a variance estimate)
    prof = gpu.mean(axis=0)[:-2]
    ok = (se > 0) & np.concatenate([prof >= 0.01 * prof.max(), [True, True]])
    z = (cpu.mean(axis=0) - gpu.mean(axis=0))[ok] / se[ok]
    print(f"  total eV/primary: gpu {gpu[:, -2].mean():.3f} cpu {cpu[:, -2].mean():.3f}; deposits/primary: "
          f"gpu {gpu[:, -1].mean():.2f} cpu {cpu[:, -1].mean():.2f}")
    from scipy import stats
    z_crit = stats.norm.isf(0.0027 / 2 / len(z))
    print(f"  {int(ok.sum())} quantities (|z| < {z_crit:.2f}), z = " + " ".join(f"{v:+.2f}" for v in z))
    assert np.all(np.abs(z) < z_crit)


def test_deposit_grid_errors():
    tris = scenes.cube_slab()
    g = na.Geometry(tris); m = na.Materials([synth.SI])
    with Driver(g, m, capacity=1024, seed=1) as drv:
        beam = raster_scan(np.zeros(1), np.zeros(1), 5.0, 1000.0, 10, 0.0)
        with pytest.raises(na.NblError):
            drv.run_persistent(beam, deposit=dict(shape=(0, 1, 1), lo=LO, hi=HI))
        with pytest.raises(na.NblError):
            drv.run_persistent(beam, deposit=dict(shape=(1, 1, 1), lo=HI, hi=LO))


def test_deposit_records_match_the_grid_of_the_same_launch():
    """Per-event records (cpu_energydep.cpp:184-192: position after the
    event, energy before, energy lost, tag) from the launch that
    also fills the grid: binned with the kernel's own float voxel rule they
    give the grid (sums in double, order aside), the record count is the
    deposit count and their sum the total; every record loses a positive
    amount of at most its energy.  A second, count-only call and a call
    with a small buffer give the counts of their own launches."""
    print("[gpu] energy-deposition records vs grid", flush=True)
    tris = scenes.cube_slab()
    g = na.Geometry(tris); m = na.Materials([synth.SI])
    n = 2000
    with Driver(g, m, capacity=1024, seed=21) as drv:
        beam = raster_scan(np.zeros(1), np.zeros(1), 5.0, 1000.0, n, 0.0, seed=7)
        t = drv.run_persistent(beam, deposit=dict(shape=SHAPE, lo=LO, hi=HI), deposit_records=6_000_000)
        assert t["primaries"] == n and t["stopped"] == 0
        rec, tags = t["deposit_records"]
        k = t["deposit_records_count"]
        assert k == t["deposit_count"] == len(rec) > 100 * n
        assert np.all(tags == 0)                   # a pixel batch tags records with the pixel index
        assert np.all(rec["deposit"] > 0) and np.all(rec["deposit"] <= rec["energy"])
        # no electron has more than the beam energy plus the barrier it gains entering the solid
        barrier = m.prepared[0].desc.barrier
        assert np.all(rec["energy"] <= np.float32(1000.0 + barrier + 1e-3))
        assert abs(rec["deposit"].astype(np.float64).sum() - t["deposit_total"]) <= 1e-9 * t["deposit_total"]
        # the kernel's voxel rule, in float32: floor((p - lo) * (n / (hi - lo)))
        lo = np.asarray(LO, np.float32); hi = np.asarray(HI, np.float32)
        inv = np.asarray(SHAPE, np.float32) / (hi - lo)
        idx = np.floor((rec["pos"] - lo) * inv).astype(np.int64)
        inside = np.all((idx >= 0) & (idx < np.asarray(SHAPE)), axis=1)
        flat = (idx[inside, 2] * SHAPE[1] + idx[inside, 1]) * SHAPE[0] + idx[inside, 0]
        grid = np.bincount(flat, weights=rec["deposit"][inside].astype(np.float64),
                           minlength=int(np.prod(SHAPE))).reshape(SHAPE[::-1])
        assert np.allclose(grid, t["deposit"], rtol=1e-9, atol=1e-9 * t["deposit_total"])
        # records only (no grid), count only, and a buffer smaller than the count
        t2 = drv.run_persistent(beam, deposit_records=0)
        assert t2["deposit_records_count"] > 100 * n and len(t2["deposit_records"][0]) == 0
        t3 = drv.run_persistent(beam, deposit_records=1000)
        assert t3["deposit_records_count"] > 1000 and len(t3["deposit_records"][0]) == 1000
        assert np.all(t3["deposit_records"][0]["deposit"] > 0)
