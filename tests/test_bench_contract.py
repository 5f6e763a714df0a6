"""bench.py's output contract: one JSON line with the metric, the whole-job value, the run shape,
the roofline of the dominant kernel and the CPU baseline; one GPU only."""

import json
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run_bench(*args, timeout=300):
    return subprocess.run([sys.executable, "-u", os.path.join(REPO, "bench.py"), *args], cwd=REPO,
                          stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=timeout)


def test_bench_refuses_several_gpus():
    """More than one GPU exits with status 2 and says why, before touching any GPU."""
    r = run_bench("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


@pytest.mark.gpu
def test_bench_json_line():
    """A small run prints one JSON line with the contract's keys and consistent values."""
    print("\n[bench contract] 64x4096, 3 steps", flush=True)
    r = run_bench("--ndet", "64", "--nt", "4096", "--steps", "3", "--warmup", "1", "--settle", "0",
                  "--no-cpu", "--batch-ntod", "0", "--no-cg", "--err-ndet", "8")
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [ln for ln in r.stdout.splitlines() if ln.strip()]
    assert len(lines) == 1
    d = json.loads(lines[0])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better",
              "scaling", "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert k in d, k
    assert d["n_gpus"] == 1 and d["steps"] == 3 and d["warmup"] == 1
    assert d["unit"] == "Gsamp/s" and d["higher_is_better"] is True and d["scaling"] == "weak"
    assert d["value"] > 0 and d["ms_per_step"] > 0
    # value is the whole job's throughput: samples per step over the time of one step.
    nsamp = d["config"]["nsamp"]
    assert nsamp == 64 * 4096
    assert abs(d["value"] - nsamp / (d["ms_per_step"] * 1e-3) / 1e9) <= 1e-2 * d["value"]   # ms rounded to 1e-4
    assert "64x4096" in d["config"]["workload"]
    rl = d["roofline"]
    assert rl["bound"] == "hbm" and rl["unit"] == "GB/s" and rl["peak"] == 8000.0
    assert abs(rl["frac"] - rl["achieved"] / rl["peak"]) < 1e-3
    assert d["cpu_baseline"] is None             # --no-cpu
    assert d["config"]["chosen_form"] in d["forms"]
