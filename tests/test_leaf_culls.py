"""The octree's leaf culls (nebula_amd/csrc/octree_build.cpp: leaf_bounds;
nbl_device.h: leaf_dropped), which let the device walk drop leaves
whose triangles all face the ignored material (side cones) or lie beyond
the walk's segment (slabs), are checked on the CPU against the reference's
own float tests (octree.inl:128-138, triangle.inl:7-28) for directions and
segments at the cull thresholds, on three scenes; so are the leaves' ropes
(the node and centre a walk's next cell starts from) against a descent with
the reference's own centre formula; planted mistakes must be caught.  The device side is covered bit for bit by the
propagate and transport parity tests."""
import os
import shutil
import subprocess

import pytest

from nebula_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or ("/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else None)
SCENES = {"rough": scenes.rough_surface, "stripes": scenes.rough_surface_stripes, "lines": scenes.line_trench}


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if HIPCC is None:
        pytest.skip("hipcc not available")
    d = tmp_path_factory.mktemp("culls")
    exe = d / "leaf_cull_check"
    csrc = os.path.join(ROOT, "nebula_amd", "csrc")
    r = subprocess.run([HIPCC, "-O2", "-std=c++17", "-ffp-contract=off", "-x", "hip", "--offload-arch=gfx950",
                        "-I", csrc, "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "leaf_cull_check.cpp"), os.path.join(csrc, "octree_build.cpp"),
                        "-o", str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return d, exe


def _run(checker, name, draws, cone_scale=1.0, slab_scale=1.0, swap=False):
    d, exe = checker
    path = d / f"{name}.bin"
    if not path.exists():
        SCENES[name]().tofile(path)
    r = subprocess.run([str(exe), str(path), str(draws), str(cone_scale), str(slab_scale)] + (["swap"] if swap else []),
                       capture_output=True, text=True, timeout=300)
    vals = dict(zip(*[iter(r.stdout.split())] * 2))
    return r.returncode, {k: int(v) for k, v in vals.items()}


@pytest.mark.parametrize("name", ["rough", "stripes", "lines"])
def test_leaf_culls_agree_with_the_reference_tests(checker, name):
    rc, v = _run(checker, name, 24 if name != "lines" else 200)
    assert rc == 0 and v["violations"] == 0 and v["slab_violations"] == 0 and v["rope_violations"] == 0, v
    assert v["ropes"] > 30_000, v
    assert v["directions"] > 100_000 and v["segments"] > 30_000, v
    # the cones and slabs cover most candidates of these scenes (the rest never cull)
    assert v["in_cone_leaves"] > 0.8 * v["candidates"] and v["in_slab_leaves"] > 0.8 * v["candidates"], v


def test_planted_threshold_mistakes_are_caught(checker):
    rc, v = _run(checker, "rough", 8, cone_scale=0.9)
    assert rc == 1 and v["violations"] > 1000, v
    rc, v = _run(checker, "rough", 8, slab_scale=0.5)
    assert rc == 1 and v["slab_violations"] > 1000, v
    rc, v = _run(checker, "lines", 1, swap=True)
    assert rc == 1 and v["rope_violations"] == v["ropes"] > 0, v
