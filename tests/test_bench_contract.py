"""CPU: the parts of bench.py's contract that need no GPU: more than one GPU
exits with a message (this engine runs on one GPU), and the flags the
measurement documents exist with their defaults (fixed-point accumulation,
re-binning every 32 steps, a kernel-timing pass of at least 288 launches,
90 algorithmic bytes per particle-step, the serving leg's 120 s of arrivals,
one batch release checked against the float64 oracle)."""

from __future__ import annotations

import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def test_documented_defaults():
    sys.path.insert(0, str(ROOT))
    import bench

    old = sys.argv
    try:
        sys.argv = ["bench.py"]
        a = bench.parse()
    finally:
        sys.argv = old
    assert a.gpus == 1 and a.steps == 1440 and a.accumulate == "fixed"
    assert a.reorder_every == 32 and a.kernel_steps == 288 and a.step_block == 128
    assert a.serve_seconds == 60.0 and a.check_releases == 1 and a.torch_steps == 60
    assert bench.BYTES_PER_PARTICLE_STEP == 90
    assert bench.BATCHED == "batched-gpu-2688r" and bench.CPU_WORKLOAD == "plain-cpu-10k-3h"
