"""The persistent kernel's octree descent (nbl_device.h: descend with its
per-lane path memo and the leaves' ropes) against the reference-order
descent (descend without a memo: the location code's digits from the root,
then the position, octree.inl:47-84), on the CPU, call for call: leaf word,
location code, centre and half size must be identical, and the memo's
descent may load no node the reference-order descent does not (its node-step
counter counts the child words and rope records it actually loads).
The reference-order descent is the one the scalar propagate kernel uses,
which the propagate parity tests hold bit for bit to the oracle.  Scenes:
the rough surface (9 levels), a cluster of tiny triangles in a large box
(21 levels, deeper than the memo's LDS levels) and the same cluster in an
off-centre box of odd size, whose cell centres round.  A planted mistake (the
reference side sees positions on a node's centre plane one ulp higher) must
be caught, which shows that exact ties are exercised.  The neighbour-code
step (next_cell, in closed form) is checked against the reference's
digit-by-digit loop on every walk step and on random codes.  A new walk
inside the memo's leaf by more than the geometry's margin takes that leaf
without descending; positions within a few hundred ulp of the last leaf's
faces exercise the margin, and the shortcut without it must be caught.""" 
import os
import shutil
import subprocess

import numpy as np
import pytest

from nebula_amd import _lib as L
from nebula_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or ("/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else None)
CLUSTER = (0.1234, 0.2345, 0.3456)


def deep_cluster_scene():
    """400 triangles of ~1e-6 inside a 2e-5 cube, plus two triangles spanning
    a 200-wide box: the octree splits down to its depth limit of 21."""
    rng = np.random.default_rng(5)
    n = 400
    t = np.zeros(n + 2, L.TRIANGLE_DTYPE)
    base = np.array(CLUSTER, np.float32)
    t["r0"][:n] = base + rng.uniform(-1e-5, 1e-5, (n, 3)).astype(np.float32)
    t["e1"][:n] = rng.uniform(-1e-6, 1e-6, (n, 3))
    t["e2"][:n] = rng.uniform(-1e-6, 1e-6, (n, 3))
    t["r0"][n], t["e1"][n], t["e2"][n] = (-100, -100, -100), (200, 0, 0), (0, 200, 0)
    t["r0"][n + 1], t["e1"][n + 1], t["e2"][n + 1] = (100, 100, 100), (-200, 0, 0), (0, 0, -200)
    t["material_in"] = 0
    t["material_out"] = -123
    return t


SKEW_CLUSTER = (37.3217, 12.70713, 5.12341)


def skewed_cluster_scene():
    """The deep cluster with the root box off-centre and of an odd size, so
    that the cell centres of the descent round (their float sums need more
    than 24 bits below the first levels)."""
    rng = np.random.default_rng(6)
    n = 400
    t = np.zeros(n + 2, L.TRIANGLE_DTYPE)
    base = np.array(SKEW_CLUSTER, np.float32)
    t["r0"][:n] = base + rng.uniform(-1e-4, 1e-4, (n, 3)).astype(np.float32)
    t["e1"][:n] = rng.uniform(-1e-5, 1e-5, (n, 3))
    t["e2"][:n] = rng.uniform(-1e-5, 1e-5, (n, 3))
    lo, hi = np.array([-63.713, -48.131, -71.917]), np.array([139.373, 157.771, 120.311])
    t["r0"][n], t["e1"][n], t["e2"][n] = lo, (hi[0] - lo[0], 0, 0), (0, hi[1] - lo[1], 0)
    t["r0"][n + 1], t["e1"][n + 1], t["e2"][n + 1] = hi, (lo[0] - hi[0], 0, 0), (0, 0, lo[2] - hi[2])
    t["material_in"] = 0
    t["material_out"] = -123
    return t


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if HIPCC is None:
        pytest.skip("hipcc not available")
    d = tmp_path_factory.mktemp("descent")
    exe = d / "descent_memo_check"
    csrc = os.path.join(ROOT, "nebula_amd", "csrc")
    r = subprocess.run([HIPCC, "-O2", "-std=c++17", "-ffp-contract=off", "-x", "hip", "--offload-arch=gfx950",
                        "-I", csrc, "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "descent_memo_check.cpp"), os.path.join(csrc, "octree_build.cpp"),
                        "-o", str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    scenes_on_disk = {}
    for name, make in (("rough", scenes.rough_surface), ("deep", deep_cluster_scene),
                       ("skew", skewed_cluster_scene)):
        path = d / f"{name}.tri"
        np.ascontiguousarray(make()).tofile(path)
        scenes_on_disk[name] = str(path)
    return exe, scenes_on_disk


def _run(checker, scene, calls, seed, nudge=False, nomargin=False):
    exe, paths = checker
    args = [str(exe), paths[scene], str(calls), str(seed)]
    if scene == "deep":
        args += ["cluster", *(repr(c) for c in CLUSTER), "2e-5"]
    if scene == "skew":
        args += ["cluster", *(repr(c) for c in SKEW_CLUSTER), "2e-4"]
    if nudge:
        args.append("nudge")
    if nomargin:
        args.append("nomargin")
    return subprocess.run(args, capture_output=True, text=True, timeout=300)


@pytest.mark.parametrize("scene,depth", [("rough", 9), ("deep", 21), ("skew", 21)])
def test_memo_descent_matches_reference_order(checker, scene, depth):
    for seed in (1, 2):
        r = _run(checker, scene, 300_000, seed)
        assert r.returncode == 0, r.stdout + r.stderr
        f = r.stdout.split()
        stats = dict(zip(f[1::2], (int(x) for x in f[2::2])))
        assert stats["tree_depth"] == depth and stats["max_depth"] == depth
        assert stats["node_steps"] < stats["ref_node_steps"]
        assert min(stats["new"], stats["rope"], stats["code"]) > 50_000
        assert stats["on_plane"] > 30_000
        assert stats["leaf_box"] > 10_000 and stats["face"] > 10_000


@pytest.mark.parametrize("scene", ["rough", "deep"])
def test_planted_tie_mistake_is_caught(checker, scene):
    r = _run(checker, scene, 300_000, 3, nudge=True)
    assert r.returncode == 1 and "MISMATCH" in r.stdout, r.stdout + r.stderr


def test_leaf_box_without_margin_is_caught(checker):
    """The leaf-box shortcut of a new walk without its rounding margin
    returns a wrong leaf next to some face where the cell centres round
    (the skewed box; the other scenes' centres happen to be exact sums):
    the check must see it."""
    r = _run(checker, "skew", 300_000, 4, nomargin=True)
    assert r.returncode == 1 and "MISMATCH" in r.stdout, r.stdout + r.stderr
