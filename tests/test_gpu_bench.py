"""GPU: bench.py's one-line JSON contract, on a short run (40 plain steps, a
101-step kernel-timing pass and one CPU-baseline run per thread count):
every field the contract names, with the types and the invariants the
measurement promises (the roofline's fraction is its achieved over its peak,
at least 100 timed launches, one GPU, the CPU baseline's thread counts and
host model)."""

from __future__ import annotations

import json
import subprocess
import sys
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]


def test_bench_prints_the_contract_line():
    cmd = [sys.executable, "-u", str(ROOT / "bench.py"), "--steps", "40", "--warmup", "2", "--no-batched",
           "--no-serving", "--no-ingest", "--no-plain-torch", "--cpu-runs", "1", "--no-fp64-check", "--kernel-steps",
           "101"]
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=400, cwd=ROOT)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith("{")]
    assert len(lines) == 1, r.stdout[-2000:]
    d = json.loads(lines[0])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better",
              "scaling", "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert k in d, k
    assert d["n_gpus"] == 1 and d["steps"] == 40 and d["warmup"] == 2
    assert d["value"] > 0 and d["ms_per_step"] > 0 and d["higher_is_better"] is True
    assert d["unit"] == "particle-steps/s" and d["dtype"] == "f32" and d["vs_baseline"] is None
    assert d["config"]["workload"] == "plain-gpu-1M-24h" and d["config"]["footprint_accumulation"] == "fixed"
    rf = d["roofline"]
    assert rf["bound"] == "hbm" and rf["unit"] == "GB/s" and rf["peak"] == 8000.0
    assert rf["achieved"] > 0 and rf["frac"] == pytest.approx(rf["achieved"] / rf["peak"], rel=1e-3)
    assert rf["bytes_per_particle_step"] == 90 and (rf["traffic"] is None or rf["traffic"] > 0)
    assert rf["kernel_launches_timed"] >= 100 and rf["particles_per_launch"] == 1 << 20
    cb = d["cpu_baseline"]
    assert cb["value"] > 0 and cb["unit"] == "particle-steps/s" and cb["kind"] == "port"
    assert cb["cores"] in [r["threads"] for r in cb["by_threads"]] and 16 in [r["threads"] for r in cb["by_threads"]]
    assert cb["cpu_model"] and cb["os_cpu_count"] >= 1 and "plain-cpu-10k-3h" in cb["sample"]
    assert all(r["particle_steps"] == 1_790_000 for r in cb["by_threads"])
    assert "error_vs_fp64" not in d
