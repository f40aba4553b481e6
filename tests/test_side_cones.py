"""The octree's leaf side cones (nebula_amd/csrc/octree_build.cpp:
side_cone), which let the device walk skip leaves whose triangles all face
the ignored material, are checked on the CPU against the reference's float
side test itself (octree.inl:128-138) for directions at the cull threshold,
on three scenes; a planted mistake in the threshold must be caught.  The
device side is covered bit for bit by the propagate parity tests."""
import os
import shutil
import subprocess

import pytest

from nebula_amd import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or ("/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else None)


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if HIPCC is None:
        pytest.skip("hipcc not available")
    d = tmp_path_factory.mktemp("cones")
    exe = d / "side_cone_check"
    csrc = os.path.join(ROOT, "nebula_amd", "csrc")
    r = subprocess.run([HIPCC, "-O2", "-std=c++17", "-ffp-contract=off", "-x", "hip", "--offload-arch=gfx950",
                        "-I", csrc, "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "side_cone_check.cpp"), os.path.join(csrc, "octree_build.cpp"),
                        "-o", str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return d, exe


def _run(checker, name, tris, per_leaf, scale):
    d, exe = checker
    path = d / f"{name}.bin"
    if not path.exists():
        tris.tofile(path)
    r = subprocess.run([str(exe), str(path), str(per_leaf), str(scale)], capture_output=True, text=True, timeout=300)
    vals = dict(zip(*[iter(r.stdout.split())] * 2))
    return r.returncode, {k: int(v) for k, v in vals.items()}


@pytest.mark.parametrize("name", ["rough", "stripes", "lines"])
def test_side_cones_agree_with_every_side_test(checker, name):
    tris = {"rough": scenes.rough_surface, "stripes": scenes.rough_surface_stripes, "lines": scenes.line_trench}[name]()
    rc, v = _run(checker, name, tris, 24 if name != "lines" else 200, 1.0)
    assert rc == 0 and v["violations"] == 0, v
    assert v["directions"] > 100_000
    # the cones cover most candidates of these scenes (the rest never cull)
    assert v["in_cone_leaves"] > 0.8 * v["candidates"], v


def test_planted_threshold_mistake_is_caught(checker):
    rc, v = _run(checker, "rough", scenes.rough_surface(), 8, 0.9)
    assert rc == 1 and v["violations"] > 1000, v
