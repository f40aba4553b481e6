"""Autotuner for the persistent transport kernel.

Replaces the reference's prescan heuristic (gpu_main.cu:108-169, which picks
frame_size and batch_size from one pass over `prescan-size` primaries) with
a measured choice of the launch shape and scheduling cadence:

  block_size       threads per workgroup (wavefronts x 64)
  blocks_per_cu    resident workgroups per compute unit
  stack_depth      LDS entries of each wavefront's secondary stack
  refill_threshold idle lanes needed before a wavefront refills (compaction cadence)
  inelastic_batch  lanes that must want an inelastic event before a wavefront runs them

Each candidate runs a short sample of the caller's own beam (same geometry,
materials and energies); the fastest by device time wins.  Results are
cached per (device, geometry size, beam signature) in a JSON file.
"""
from __future__ import annotations

import hashlib
import itertools
import json
import os

import numpy as np

from . import _lib as L

CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_build", "autotune.json")

DEFAULT_SPACE = {
    "block_size": (64, 128, 256),
    "blocks_per_cu": (4, 5, 8, 10),
    "stack_depth": (64, 96, 112),
    "refill_threshold": (2, 4, 8),
    "inelastic_batch": (1, 8),
}
PATH_MEMO_WORDS = 15         # per-lane octree path memo in LDS (lds_path_cache::kWords)
MAX_WAVES_PER_CU = 20        # the table kernel is built for 5 wavefronts per SIMD


def _signature(driver, beam, sample, space):
    info = driver.geometry.info()
    px = np.asarray(beam["pixels"])
    h = hashlib.sha1()
    h.update(px.tobytes())
    h.update(str((info["n_triangles_in"], info["n_nodes"], info["n_leaves"], sample)).encode())
    # the library build and the search space: a rebuilt kernel is tuned afresh
    h.update(str(sorted(space.items())).encode())
    h.update(str(os.path.getmtime(L.LIB_PATH)).encode())
    name, cus = _device()
    return f"{name}:{cus}:{h.hexdigest()[:16]}"


def _device():
    from .api import device_info
    return device_info()


def _sample_beam(beam, sample):
    px = np.array(beam["pixels"], copy=True)
    total = int(px["n_primaries"].sum())
    if total > sample:
        scale = sample / total
        px["n_primaries"] = np.maximum(1, np.round(px["n_primaries"] * scale)).astype(np.uint32)
    return dict(beam, pixels=px)


def valid(cfg):
    waves = cfg["block_size"] // 64
    lds = waves * (9 * cfg["stack_depth"] + PATH_MEMO_WORDS * 64) * 4
    return (lds * cfg["blocks_per_cu"] <= 160 * 1024 and
            waves * cfg["blocks_per_cu"] <= MAX_WAVES_PER_CU)


LIBRARY_DEFAULT = {"block_size": 128, "blocks_per_cu": 10, "stack_depth": 96, "refill_threshold": 4,
                   "inelastic_batch": 8}


def tune(driver, beam, sample=200_000, space=None, repeats=1, use_cache=True, log=None):
    """Returns (best_config, results) for this driver and beam.  Every
    candidate of the space that fits a compute unit runs the sample `repeats`
    times (its best device time counts); results list them all."""
    space = space or DEFAULT_SPACE
    key = _signature(driver, beam, sample, space)
    cache = {}
    if use_cache and os.path.exists(CACHE):
        try:
            with open(CACHE) as f:
                cache = json.load(f)
        except (OSError, ValueError):
            cache = {}
        if key in cache:
            return cache[key]["best"], cache[key]["results"]
    sb = _sample_beam(beam, sample)
    names = list(space)
    results = []
    driver.run_persistent(sb)      # warm caches and code objects
    for values in itertools.product(*(space[k] for k in names)):
        cfg = dict(zip(names, values))
        if not valid(cfg):
            continue
        ms = min(driver.run_persistent(sb, **cfg)["kernel_ms"] for _ in range(repeats))
        results.append(dict(cfg, kernel_ms=ms))
        if log:
            log(f"autotune {cfg} -> {ms:.2f} ms")
    best = min(results, key=lambda r: r["kernel_ms"])
    best_cfg = {k: best[k] for k in names}
    if use_cache:
        cache[key] = {"best": best_cfg, "results": results}
        os.makedirs(os.path.dirname(CACHE), exist_ok=True)
        with open(CACHE, "w") as f:
            json.dump(cache, f, indent=1)
    return best_cfg, results
