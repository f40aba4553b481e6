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
