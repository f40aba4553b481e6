"""Benchmark: primary electrons/s of the GPU transport on one MI355X, beside
the CPU oracle on the same host.

Workload (rough surface): a 512 x 512 Gaussian height field (sigma 2 nm,
correlation length 10 nm, seed 1) over 1 um^2 of synthetic Si, triangulated
to 522,242 surface triangles (526,334 with walls, bottom and detector); 1e7
primaries of 300 eV uniform over the 1 um^2 area.  One step = transport of
all 1e7 primaries to the end of their cascades in one persistent launch.

Second part of the metric (learned vs table sampler, unless --no-learned):
the rough surface in 8 material stripes, 1e7 primaries of 500 eV, once with
the ICDF tables and once with the learned BF16 sampler trained on 1e8
samples of the same tables; primaries/s of both, yields, and the sampler's
KS distance per energy bin.

Third part (unless --no-penn): the rough-surface workload with Si under the
full Penn inelastic model (the reference's default physics) instead of
Kieft-Bosch: primaries/s and yields of one launch after a warm-up.

Output: one JSON line (metric, value, roofline, cpu_baseline, counters,
learned_vs_table, full_penn).
`--gpus N` with N > 1 exits: this build measures one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0   # MI355X HBM3E, AMD public spec
N_SIMDS = 1024           # 256 CUs x 4 SIMDs
# algorithmic bytes per unit of work (one event = one init + event step of one
# electron); the kernel's counters count the work it does: node_steps = child
# words, leaf records and rope records it loads, tri_tests = triangle records
# it tests (leaves a cull drops cost none)
BYTES = dict(event=122, lookup_1d=8, lookup_2d=16, rounddown=4, node_step=4, tri_test=44,
             boundary=52, secondary=65)
# the tallies a launch fills only when asked to count its work (count_work)
WORK_KEYS = ("events", "lookups_1d", "lookups_2d", "rounddowns", "node_steps", "tri_tests", "boundary_events",
             "secondaries", "pending_retries")


WORKLOAD_TEXT = {
    "rough": "rough surface: 512x512 Gaussian height field (sigma 2 nm, corr 10 nm, seed 1), 526334 triangles, "
             "synthetic Si, 300 eV uniform over 1 um^2, 1e7 primaries",
    "stripes8": "rough surface in 8 material stripes, 533502 triangles, 500 eV uniform over 1 um^2, 1e7 primaries",
    "rough500": "rough surface, one stripe material, 500 eV uniform over 1 um^2, 1e7 primaries",
    "rough_penn": "rough surface as 'rough' with Si under the full Penn inelastic model, 300 eV, 1e7 primaries",
    "lines": "Si + 5 PMMA lines (32 nm wide, 64 nm pitch, 50 nm tall), 10004 triangles, 500 eV Gaussian sigma 1 nm, "
             "256x64 pixels x 1000 primaries",
    "scan3": "lines geometry with 3 materials, 4096 pixels x Poisson(200) primaries, 200-2000 eV",
    "cube": "flat Si slab (1 um cube), 1 keV pencil beam, 1e5 primaries",
}


def profiled(workload, primaries, kernel_ms):
    """The committed rocprofv3 figures of this workload's kernel (HBM bytes
    per launch from FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated; issue
    counters), or None unless the workload and launch size match and this
    run's kernel time is within 15 % of the profiled one (same kernel build)."""
    summary = {"rough": "k_persistent_rough_surface_300eV"}.get(workload)
    if summary is None:
        return None, None
    path = os.path.join(ROOT, "profiles", summary + ".traffic.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        prof = json.load(f)
    if prof.get("workload") != workload or prof.get("primaries_per_launch") != primaries:
        return None, None
    if not abs(kernel_ms / prof["avg_kernel_ms"] - 1) < 0.15:
        return None, None
    return prof, os.path.relpath(path, ROOT)


def host_cpu():
    """Model name and logical CPU count of this host, and the threads this
    process may run on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}


def algorithmic_bytes(t):
    return (BYTES["event"] * t["events"] + BYTES["lookup_1d"] * t["lookups_1d"] +
            BYTES["lookup_2d"] * t["lookups_2d"] + BYTES["rounddown"] * t["rounddowns"] +
            BYTES["node_step"] * t["node_steps"] + BYTES["tri_test"] * t["tri_tests"] +
            BYTES["boundary"] * t["boundary_events"] + BYTES["secondary"] * t["secondaries"])


def learned_vs_table(primaries, seed, train_samples):
    """The 8-material stripes workload with the table path and with the
    learned sampler (trained here on samples of the same tables): primaries/s
    of each, yields, and the learned sampler's KS distance per energy bin."""
    import nebula_amd as na
    from nebula_amd import workloads
    from nebula_amd.driver import Driver
    tris, specs, beam = workloads.stripes8()
    beam = workloads.scaled(beam, primaries)
    e_beam = float(beam["pixels"]["energy"].max())
    g = na.Geometry(tris)
    m = na.Materials(specs)
    t0 = time.time()
    # train over every energy an electron can have in the materials: from the
    # tables' lower end (1 eV, below every barrier) to above the beam energy
    smp = na.Sampler(m, samples=train_samples, e_min=0.0, e_max=1.2 * e_beam, seed=seed)
    train_s = time.time() - t0
    # the elastic network alone (inelastic draws from the tables)
    smp_el = na.Sampler(m, samples=train_samples, e_min=0.0, e_max=1.2 * e_beam, seed=seed, kinds="elastic")
    rep = smp.report()
    edges = np.geomspace(5.0, 0.95 * e_beam, 9)
    ks = {k: [smp.ks_per_bin(kind, mat, edges, n=20000, seed=mat) for mat in range(len(specs))]
          for k, kind in (("elastic", 0), ("inelastic", 1))}
    out = {}
    for name, sampler in (("table", None), ("learned", smp), ("learned_elastic", smp_el)):
        with Driver(g, m, capacity=1024, seed=77) as drv:
            drv.set_sampler(sampler)
            work = drv.run_persistent(dict(beam, seed=1), count_work=True)   # warm-up, counts the work
            t1 = time.perf_counter()
            t = drv.run_persistent(dict(beam, seed=2))
            wall = time.perf_counter() - t1
        n = t["primaries"]
        out[name] = {"primaries_per_s": n / wall, "kernel_ms": t["kernel_ms"], "events_per_s": work["events"] / wall,
                     "se": t["detected_se"] / n, "bse": t["detected_bse"] / n, "stopped": t["stopped"],
                     "counts": (t["detected_se"], t["detected_bse"], n)}
        print(f"{name}: {n / wall:.4g} primaries/s, SE {out[name]['se']:.4f} BSE {out[name]['bse']:.4f}",
              file=sys.stderr, flush=True)
    def yield_z(name):
        z = {}
        for i, key in enumerate(("se", "bse")):
            (a, na_), (b, nb_) = (out["table"]["counts"][i], out["table"]["counts"][2]), \
                                 (out[name]["counts"][i], out[name]["counts"][2])
            # Poisson counting error of each yield
            z[key] = (b / nb_ - a / na_) / np.hypot(np.sqrt(max(a, 1)) / na_, np.sqrt(max(b, 1)) / nb_)
        return z
    z, z_el = yield_z("learned"), yield_z("learned_elastic")
    for v in out.values():
        v.pop("counts")
    return {"workload": WORKLOAD_TEXT["stripes8"].replace("1e7", f"{primaries:.0e}"),
            "table": out["table"], "learned": out["learned"],
            "speedup_learned_vs_table": out["learned"]["primaries_per_s"] / out["table"]["primaries_per_s"],
            "yield_z": z, "yields_equal_3sigma": bool(max(abs(z["se"]), abs(z["bse"])) < 3),
            "learned_elastic_only": {**out["learned_elastic"],
                                     "speedup_vs_table": out["learned_elastic"]["primaries_per_s"] /
                                     out["table"]["primaries_per_s"],
                                     "yield_z": z_el,
                                     "yields_equal_3sigma": bool(max(abs(z_el["se"]), abs(z_el["bse"])) < 3),
                                     "note": "elastic cos(theta) from the network, inelastic omega' from the tables"},
            "ks_per_bin": {"energy_edges_eV": [float(e) for e in edges],
                           "note": "per material (rows) and energy bin; table cells below the Fermi level excluded",
                           **{k: [[None if not np.isfinite(x) else round(float(x), 5) for x in row] for row in v]
                              for k, v in ks.items()},
                           "max": {k: float(np.nanmax(v)) for k, v in ks.items()}},
            "training": {"samples_per_network": rep["samples"], "seconds": train_s,
                         "mse_standardised": {"elastic": rep["loss_elastic"], "inelastic": rep["loss_inelastic"]},
                         "energy_range_eV": [rep["e_min"], rep["e_max"]]}}


def full_penn(tris, specs, beam, steps):
    """The same geometry and beam with every material under full Penn."""
    import nebula_amd as na
    from nebula_amd import synth
    from nebula_amd.driver import Driver
    m = na.Materials([synth.penn(s) if s.inelastic != "penn" else s for s in specs])
    g = na.Geometry(tris)
    with Driver(g, m, capacity=1024, seed=31) as drv:
        work = drv.run_persistent(dict(beam, seed=900), count_work=True)  # warm-up, counts the work
        ms, tot = [], None
        t1 = time.perf_counter()
        for s in range(steps):
            t = drv.run_persistent(dict(beam, seed=901 + s))
            ms.append(t["kernel_ms"])
            tot = t if tot is None else {k: tot[k] + t[k] for k in t}
        wall = time.perf_counter() - t1
    n = tot["primaries"]
    print(f"full Penn: {n / wall:.4g} primaries/s, SE {tot['detected_se'] / n:.4f} BSE {tot['detected_bse'] / n:.4f}",
          file=sys.stderr, flush=True)
    return {"workload": "same geometry and beam, materials under the full Penn inelastic model",
            "primaries_per_s": n / wall, "avg_kernel_ms": float(np.mean(ms)),
            "events_per_s": work["events"] * steps / wall,
            "se": tot["detected_se"] / n, "bse": tot["detected_bse"] / n, "stopped": tot["stopped"], "steps": steps}


def cpu_baseline(tris, specs, beam, n_sample, time_limit, threads):
    """The oracle's cpu_mt restatement on a bounded sample of the workload."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    import runs
    import nebula_amd as na
    from nebula_amd import workloads
    og = ol.OrcGeometry(tris)
    om = ol.OrcMaterials([na.prepared(s).desc for s in specs])
    prim, _ = workloads.sample_primaries(beam, n_sample, seed=12345)
    st, _ = runs.oracle_run(og, om, prim, threads=threads, time_limit=time_limit)
    return st


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="rough",
                    help="rough (default: the headline config), stripes8, rough500, lines, scan3, cube")
    ap.add_argument("--primaries", type=int, default=0, help="primaries per step (0 = the workload's own)")
    ap.add_argument("--cpu-sample", type=int, default=0,
                    help="primaries in the CPU baseline's sample (0 = 40,000 per host thread, at least 600,000)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="CPU baseline threads (0 = every thread this process may run on, "
                         "as cpu_mt_main.cpp:197 takes hardware_concurrency())")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-vs-table sampler comparison")
    ap.add_argument("--no-penn", action="store_true", help="skip the full Penn run of the same workload")
    ap.add_argument("--learned-primaries", type=int, default=10_000_000)
    ap.add_argument("--train-samples", type=int, default=100_000_000)
    ap.add_argument("--block-size", type=int, default=0)
    ap.add_argument("--blocks-per-cu", type=int, default=0)
    ap.add_argument("--stack-depth", type=int, default=0)
    ap.add_argument("--refill-threshold", type=int, default=0)
    ap.add_argument("--inelastic-batch", type=int, default=0)
    ap.add_argument("--inelastic-wait", type=int, default=0)
    ap.add_argument("--no-autotune", action="store_true",
                    help="launch with the library defaults (or the options given) instead of the autotuner's choice")
    ap.add_argument("--autotune-sample", type=int, default=1_000_000)
    ap.add_argument("--copy-bytes", type=int, default=1 << 30,
                    help="size of the plain copy kernel run first (counter scale, bandwidth yardstick); 0 = skip")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the transport shards by primaries with no "
              "cross-GPU traffic, and multi-GPU runs are not part of this build.", file=sys.stderr)
        sys.exit(2)

    import nebula_amd as na
    from nebula_amd import workloads
    from nebula_amd.driver import Driver

    t0 = time.time()
    tris, specs, base_beam = workloads.BY_NAME[args.workload]()
    base_beam = workloads.scaled(base_beam, args.primaries)
    n_step = int(base_beam["pixels"]["n_primaries"].sum())
    g = na.Geometry(tris)
    m = na.Materials(specs)
    info = g.info()
    name, cus = na.device_info()
    print(f"setup: {len(tris)} triangles, {info['n_nodes']} nodes, {info['n_leaves']} leaves, "
          f"device {name} ({cus} CUs), {time.time() - t0:.1f}s", file=sys.stderr, flush=True)

    copy_gbps = None
    if args.copy_bytes:
        import ctypes as C
        ms = C.c_double()
        na._lib.call("nbl_bench_copy", args.copy_bytes, 3, C.byref(ms))
        copy_gbps = 2 * args.copy_bytes / (ms.value * 1e-3) / 1e9
        print(f"copy kernel: {args.copy_bytes} B in {ms.value:.3f} ms = {copy_gbps:.0f} GB/s (read + write)",
              file=sys.stderr, flush=True)

    opts = dict(block_size=args.block_size, blocks_per_cu=args.blocks_per_cu,
                stack_depth=args.stack_depth, refill_threshold=args.refill_threshold,
                inelastic_batch=args.inelastic_batch, inelastic_wait=args.inelastic_wait)
    drv = Driver(g, m, capacity=1024, seed=2024)
    beam = lambda s: dict(base_beam, seed=s)
    tuned = None
    if not args.no_autotune and not any(opts.values()):
        # launch shape and cadence from the autotuner (cached per device,
        # geometry and beam), which replaces the reference's prescan heuristic
        from nebula_amd import autotune
        t_tune = time.time()
        best, results = autotune.tune(drv, base_beam, sample=args.autotune_sample, repeats=2)
        default_ms = next((r["kernel_ms"] for r in results
                           if all(r[k] == v for k, v in autotune.LIBRARY_DEFAULT.items())), None)
        best_ms = min(r["kernel_ms"] for r in results)
        opts.update(best)
        tuned = {"sample_primaries": args.autotune_sample, "candidates": len(results), "chosen": best,
                 "chosen_sample_ms": best_ms, "library_default_sample_ms": default_ms,
                 "seconds": round(time.time() - t_tune, 1)}
        print(f"autotune: {best} ({best_ms:.2f} ms per sample of {args.autotune_sample}; library default "
              f"{default_ms} ms; {len(results)} candidates)", file=sys.stderr, flush=True)
    # The work a launch does (events, table lookups, node steps, triangle
    # tests, ...) comes from a counting launch of the same workload (the first
    # warm-up); the timed launches run without the work counters, which cost
    # the kernel about 4 %.
    work = None
    for w in range(args.warmup):
        t = drv.run_persistent(beam(1000 + w), count_work=(w == 0), **opts)
        if w == 0:
            work = t
        print(f"warmup {w + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    totals = None
    kernel_ms = []
    t_start = time.perf_counter()
    for s in range(args.steps):
        t = drv.run_persistent(beam(s), **opts)
        kernel_ms.append(t["kernel_ms"])
        if totals is None:
            totals = {k: 0 for k in t}
        for k, v in t.items():
            totals[k] += v
        print(f"step {s + 1}/{args.steps}: {t['kernel_ms']:.1f} ms, SE {t['detected_se'] / t['primaries']:.4f} "
              f"BSE {t['detected_bse'] / t['primaries']:.4f}", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t_start
    assert totals["primaries"] == n_step * args.steps and totals["stopped"] == 0, totals
    if work is None:
        work = drv.run_persistent(beam(1000), count_work=True, **opts)
    ms_per_step = wall * 1000.0 / args.steps
    value = totals["primaries"] / wall
    avg_kernel_ms = float(np.mean(kernel_ms))
    per_launch = {k: v / args.steps for k, v in totals.items()}
    per_launch.update({k: float(work[k]) for k in WORK_KEYS})
    ach_gbps = algorithmic_bytes(per_launch) / (avg_kernel_ms * 1e-3) / 1e9
    prof, prof_src = profiled(args.workload, n_step, avg_kernel_ms)
    traffic = float(prof["traffic_bytes"]) if prof else None
    issue = None
    if prof and prof.get("valu_insts") and prof.get("gpu_cycles"):
        # vector instructions issued per SIMD per cycle of the launch (a SIMD
        # issues at most one per cycle; at 4 cycles per wave64 instruction
        # the profiled launch would not fit its own duration, so gfx950
        # retires many in fewer, and this issue rate is what is measured)
        per_simd = prof["valu_insts"] / N_SIMDS
        issue = {"what": "vector ALU instruction issue per SIMD per cycle", "valu_insts_per_launch": prof["valu_insts"],
                 "valu_insts_per_simd": per_simd, "launch_cycles": prof["gpu_cycles"],
                 "frac": per_simd / prof["gpu_cycles"], "cycles_per_valu_inst": prof["gpu_cycles"] / per_simd,
                 "valu_insts_per_event": prof["valu_insts"] / per_launch["events"], "source": prof_src}

    cpu = None
    if not args.no_cpu:
        hc = host_cpu()
        threads = args.cpu_threads or hc["affinity"]
        sample = args.cpu_sample or max(600_000, 40_000 * threads)
        st = cpu_baseline(tris, specs, base_beam, sample, args.cpu_seconds, threads)
        cpu = {"value": st["primaries"] / st["seconds"], "unit": "primaries/s", "cores": int(st["threads"]),
               "kind": "port",
               "sample": f"oracle cpu_mt restatement on {st['threads']} threads (every thread of this process's "
                         f"affinity, as cpu_mt_main.cpp:197), {st['primaries']} primaries drawn from the same pixel "
                         f"batch (sample of {sample}, stopped at {args.cpu_seconds:.0f} s), "
                         f"{st['seconds']:.1f} s, {st['events'] / st['seconds']:.3g} events/s",
               "host_nproc": hc["nproc"], "host_affinity_threads": hc["affinity"], "cpu_model": hc["model"],
               "events_per_s": st["events"] / st["seconds"],
               "se_yield": st["detected_se"] / max(st["primaries"], 1),
               "bse_yield": st["detected_bse"] / max(st["primaries"], 1)}
        print(f"cpu baseline: {cpu['value']:.0f} primaries/s on {cpu['cores']} threads "
              f"({hc['model']}, nproc {hc['nproc']})", file=sys.stderr, flush=True)

    penn = None
    if not args.no_penn and all(s.inelastic == "kieft" for s in specs):
        penn = full_penn(tris, specs, base_beam, max(1, min(args.steps, 2)))

    learned = None
    if not args.no_learned:
        learned = learned_vs_table(args.learned_primaries, 5, args.train_samples)

    out = {
        "metric": "primary electrons/s (1 MI355X) vs host CPU; learned vs table sampler at equal yield",
        "value": value,
        "unit": "primaries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": WORKLOAD_TEXT.get(args.workload, args.workload),
                   "name": args.workload, "primaries_per_step": n_step, "triangles": int(len(tris)),
                   "materials": len(specs), "pixels": int(len(base_beam["pixels"])),
                   "sampler": "table", "kernel": "persistent",
                   "launch": {k: v for k, v in opts.items()}, "autotune": tuned},
        # HBM roofline of the persistent kernel: algorithmic bytes of the work
        # the launch did (counters x bytes per unit) per second of launch; the
        # HBM bytes it actually moved (profiled) beside it; and the bound the
        # profile measures (vector issue), with its own fraction
        "roofline": {"bound": "hbm", "achieved": ach_gbps, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": ach_gbps / HBM_PEAK_GBPS, "traffic": traffic, "traffic_source": prof_src,
                     "kernel": "k_persistent", "avg_kernel_ms": avg_kernel_ms, "copy_kernel_GBps": copy_gbps,
                     "algorithmic_bytes_per_launch": algorithmic_bytes(per_launch),
                     "bytes_per_unit": BYTES,
                     "traffic_GBps": traffic / (avg_kernel_ms * 1e-3) / 1e9 if traffic else None,
                     "traffic_frac": traffic / (avg_kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBPS if traffic else None,
                     "measured_bound": issue},
        "cpu_baseline": cpu,
        "speedup_vs_cpu": (value / cpu["value"]) if cpu else None,
        "events_per_s": per_launch["events"] * args.steps / wall,
        "counters_per_step": per_launch,
        "yields": {"se": totals["detected_se"] / totals["primaries"], "bse": totals["detected_bse"] / totals["primaries"]},
        "learned_vs_table": learned,
        "full_penn": penn,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
