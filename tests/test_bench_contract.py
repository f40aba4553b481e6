"""bench.py's output contract: one JSON line with the metric, the whole-job
value, the step bookkeeping, the workload, a roofline object for the
persistent kernel and the CPU baseline; requests for more than one GPU end
with a message and status 2 before touching a device."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120, cwd=ROOT)
    assert r.returncode == 2 and "one GPU" in r.stderr


@pytest.mark.gpu
def test_bench_line_has_every_field():
    print("[gpu] bench contract", flush=True)
    r = subprocess.run([sys.executable, "-u", os.path.join(ROOT, "bench.py"), "--primaries", "200000", "--steps", "1",
                        "--warmup", "1", "--no-learned", "--no-penn", "--cpu-sample", "3000", "--cpu-seconds", "3",
                        "--copy-bytes", str(64 << 20)], capture_output=True, text=True, timeout=280, cwd=ROOT)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [l for l in r.stdout.strip().splitlines() if l.startswith("{")]
    assert len(lines) == 1
    d = json.loads(lines[0])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
              "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert k in d, k
    assert d["n_gpus"] == 1 and d["steps"] == 1 and d["warmup"] == 1 and d["higher_is_better"] is True
    assert d["value"] > 0 and d["ms_per_step"] > 0 and d["vs_baseline"] is None
    assert "workload" in d["config"] and d["config"]["primaries_per_step"] == 200000
    rl = d["roofline"]
    for k in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert k in rl, k
    assert rl["bound"] == "hbm" and rl["unit"] == "GB/s" and rl["peak"] == 8000.0
    assert rl["frac"] == pytest.approx(rl["achieved"] / rl["peak"])
    # the kernel's own work per second cannot exceed the peak it is held against
    assert 0 < rl["frac"] <= 1
    # a profile covers the full-size workload only: no traffic figure here
    assert rl["traffic"] is None and rl["traffic_GBps"] is None
    cb = d["cpu_baseline"]
    for k in ("value", "unit", "cores", "kind", "sample"):
        assert k in cb, k
    assert cb["value"] > 0 and cb["kind"] == "port" and cb["cores"] >= 1
    # every thread the process may run on (cpu_mt_main.cpp:197), with the host stated
    assert cb["cores"] == cb["host_affinity_threads"] and cb["host_nproc"] >= cb["cores"] and cb["cpu_model"]
