"""The benchmark workloads as (triangles, material specs, pixel batch).

  cube      flat Si slab, 1 keV pencil beam
  lines     Si substrate + 5 PMMA lines, 500 eV Gaussian beam (sigma 1 nm),
            256 x 64 raster, 1000 primaries per pixel
  rough     512 x 512 rough Si surface, 300 eV uniform over 1 um^2, 1e7 primaries
  scan3     the lines geometry with 3 materials (lines alternate PMMA / W-like),
            4096 pixels, Poisson(200) primaries per pixel (seed 2),
            energies uniform in 200..2000 eV, one persistent launch
  stripes8  the rough surface in 8 material stripes, 500 eV, 1e7 primaries
"""
from __future__ import annotations

import numpy as np

from . import scenes, synth
from .driver import pencil_beam, raster_scan, uniform_beam
from . import _lib as L


def cube(n=100_000):
    return scenes.cube_slab(), [synth.SI], pencil_beam(1000.0, n, pos=(0.0, 0.0, 5.0))


def lines(nx=256, ny=64, per_pixel=1000, scale=1.0):
    tris = scenes.line_trench()
    xs = np.linspace(-160.0, 160.0, nx)
    ys = np.linspace(-40.0, 40.0, ny)
    beam = raster_scan(xs, ys, scenes.LINE_HEIGHT + 5.0, 500.0, max(1, int(per_pixel * scale)), 1.0, seed=1)
    return tris, [synth.SI, synth.PMMA], beam


def _top(tris):
    s = tris[tris["material_in"] != L.MAT_DETECTOR]
    return float(np.max(np.concatenate([s["r0"][:, 2], s["r0"][:, 2] + s["e1"][:, 2], s["r0"][:, 2] + s["e2"][:, 2]])))


def rough(n=10_000_000):
    tris = scenes.rough_surface()
    return tris, [synth.SI], uniform_beam(300.0, n, (500.0, 500.0), _top(tris) + 5.0, seed=0)


def scan3(n_pixels=4096, mean=200, seed=2):
    tris = scenes.line_trench(line_materials=(1, 2, 1, 2, 1))
    rng = np.random.default_rng(seed)
    side = int(np.sqrt(n_pixels))
    xs = np.linspace(-160.0, 160.0, side)
    ys = np.linspace(-40.0, 40.0, n_pixels // side)
    beam = raster_scan(xs, ys, scenes.LINE_HEIGHT + 5.0, 0.0, 0, 1.0, seed=3)
    px = beam["pixels"]
    px["n_primaries"] = rng.poisson(mean, len(px)).astype(np.uint32)
    px["energy"] = rng.uniform(200.0, 2000.0, len(px)).astype(np.float32)
    return tris, [synth.SI, synth.PMMA, synth.W_LIKE], beam


def stripes8(n=10_000_000):
    tris = scenes.rough_surface_stripes()
    return tris, synth.stripe_materials(8), uniform_beam(500.0, n, (500.0, 500.0), _top(tris) + 5.0, seed=0)


def rough500(n=10_000_000):
    """The rough surface in one material at 500 eV (single-material control for stripes8)."""
    tris = scenes.rough_surface()
    return tris, [synth.stripe_materials(8)[0]], uniform_beam(500.0, n, (500.0, 500.0), _top(tris) + 5.0, seed=0)


def rough_penn(n=10_000_000):
    """The rough-surface workload with Si under the full Penn inelastic model."""
    tris, _, beam = rough(n)
    return tris, [synth.SI_PENN], beam


BY_NAME = {"cube": cube, "lines": lines, "rough": rough, "scan3": scan3, "stripes8": stripes8, "rough500": rough500,
           "rough_penn": rough_penn}


def sample_primaries(beam, n, seed=0):
    """n primaries drawn on the host from a pixel batch (pixels in proportion
    to their primaries, positions by the pixel's profile), as particles with
    the pixel index as tag; used to run the same workload on the CPU."""
    rng = np.random.default_rng(seed)
    px = np.asarray(beam["pixels"])
    w = px["n_primaries"].astype(float)
    idx = rng.choice(len(px), n, p=w / w.sum())
    p = np.zeros(n, L.PARTICLE_DTYPE)
    x = px["x"][idx].astype(np.float64); y = px["y"][idx].astype(np.float64)
    gauss = px["profile"][idx] == 0
    x = np.where(gauss, x + rng.normal(0, 1, n) * px["sigma"][idx], x + px["half_x"][idx] * rng.uniform(-1, 1, n))
    y = np.where(gauss, y + rng.normal(0, 1, n) * px["sigma"][idx], y + px["half_y"][idx] * rng.uniform(-1, 1, n))
    p["pos"][:, 0], p["pos"][:, 1], p["pos"][:, 2] = x, y, px["z"][idx]
    d = np.asarray(beam.get("dir", (0.0, 0.0, -1.0)), np.float64)
    p["dir"] = d / np.linalg.norm(d)
    p["kin_energy"] = px["energy"][idx]
    return p, idx.astype(np.uint32)


def scaled(beam, primaries):
    """The same pixel batch with its primaries scaled to about `primaries` in total."""
    px = np.array(beam["pixels"], copy=True)
    total = px["n_primaries"].sum()
    if primaries and total != primaries:
        px["n_primaries"] = np.maximum(0, np.round(px["n_primaries"] * (primaries / total))).astype(np.uint32)
    return dict(beam, pixels=px)
