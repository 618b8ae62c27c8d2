"""bench.py's output line on a small workload: exactly one JSON object on
stdout with the keys of the measurement record (metric, value, unit,
n_gpus, steps, warmup, ms_per_step, higher_is_better, scaling, vs_baseline,
dtype, data, config), a roofline block for the dominant kernel and a CPU
baseline block; and the one-GPU refusal of --gpus 2."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OFF = ["--kln-events", "0", "--kln-direct-events", "0", "--config1-events", "0", "--adjoint-events", "0",
       "--p3d-events", "0", "--mixed-events", "0", "--avg-events", "0", "--jet-events", "0",
       "--default-events", "0", "--no-graph", "--no-autotune"]


def _bench(*args, timeout=240):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=timeout, cwd=ROOT)


def test_bench_prints_one_record_line():
    r = _bench("--events", "2000", "--steps", "2", "--warmup", "1", "--cpu-seconds", "1", "--parity-events", "8",
               *OFF)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [l for l in r.stdout.splitlines() if l.strip()]
    assert len(lines) == 1, r.stdout[-2000:]
    d = json.loads(lines[0])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
              "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert k in d, k
    assert d["n_gpus"] == 1 and d["steps"] == 2 and d["warmup"] == 1
    assert d["higher_is_better"] is True and d["data"].startswith("synthetic") and d["unit"] == "events/s"
    assert d["value"] > 0 and d["ms_per_step"] > 0
    assert d["config"]["events_per_step"] == 2000 and d["config"]["failed_events"] == 0
    roof = d["roofline"]
    for k in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert k in roof, k
    assert roof["achieved"] > 0 and roof["frac"] == pytest.approx(roof["achieved"] / roof["peak"])
    cpu = d["cpu_baseline"]
    for k in ("value", "unit", "cores", "kind", "sample"):
        assert k in cpu, k
    assert cpu["value"] > 0 and cpu["kind"] in ("port", "reference") and 1 <= cpu["cores"] <= 16
    eps = d["config"]["eps_err_vs_fp64"]
    assert eps["events"] == 8 and eps["eps_r2"] <= 1e-6 and eps["eps_rn"] <= 1e-6


def test_bench_refuses_more_than_one_gpu():
    r = _bench("--gpus", "2", timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr
