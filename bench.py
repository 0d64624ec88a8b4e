"""Throughput benchmark of the backward-LPDM step engine on one MI355X.

Headline workload ("plain-gpu-1M-24h"): one release of 2^20 particles, 24 h
backward at dt = 60 s (1,440 steps), full physics (Hanna three-regime
turbulence with the Richardson closure aloft, meander on, per-column Emanuel
convection once per met hour), synthetic met on a 240 x 160 grid at 0.25 deg
with 60 AGL levels and 25 hourly snapshots, footprint 24 time-ago bins x 1
layer x 160 x 240 in float32.

A "step" is one model time step of that run over all 2^20 particles: the fused
step kernel (which, by default, also adds every particle's residence to the
footprint's exact fixed-point sums), the re-binning of the particles by met
cell every few steps and, on a bracket change, the hourly resample, window
build and convection. The float32 footprint is produced from the sums at the
end of the timed region, inside it. The met source hours are resident in HBM
before the timed region starts. ``value`` counts active particle-steps per
second of wall time over the timed steps.

Beside the headline the same JSON line carries:

* ``plain_torch``: the plain FP32 method as PyTorch ops on the same GPU and
  workload (one met hour, convection included), eager and captured as a graph;
* ``batched`` / ``batched_bf16``: the whole 2,688-release network batch
  (56 sites x 48 hourly releases, 2k-20k particles, 24-120 h back, 168 met
  hours) in one engine, float32 met and the bfloat16 precision layout;
* ``serving``: Poisson footprint requests on a bfloat16 met ring;
* ``ingest``: the headline workload with its met streamed from the host;
* ``cpu_baseline``: the same plain method as PyTorch ops on the host's CPU
  cores (the 10k-particle, 3 h workload, the reference's dynamic CPU path);
* ``error_vs_fp64``: errors of the engine's results against the float64
  oracle on the same noise (per-step positions over the first met hour;
  full-run footprint and residence total of sampled releases of the batch).

Prints one JSON line. ``--gpus N`` with N > 1 exits: this engine runs on one GPU.
"""

from __future__ import annotations

import argparse
import json
import os
import statistics
import subprocess
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

METRIC = "particle-steps/s, batched vs plain at stated error; p50/p99 footprint latency"
WORKLOAD = "plain-gpu-1M-24h"
CPU_WORKLOAD = "plain-cpu-10k-3h"
BATCHED = "batched-gpu-2688r"
BATCHED_BF16 = "batched-gpu-2688r-bf16"
SERVING = "serving-gpu-10d"
PEAK_HBM_GBS = 8000.0
# VALU issue model of MI355X: 256 CUs x 4 SIMDs, 2.4 GHz peak engine clock, a
# wave64 VALU instruction occupies its SIMD for 4 cycles
N_SIMD, CLOCK_HZ, VALU_CYCLES = 1024, 2.4e9, 4
# Algorithmic HBM bytes per active particle-step of the fused step kernel
# (float32 structure-of-arrays state): read lon lat z u' v' w' u_m v_m,
# weight, r_off, we_off, rel_idx, alive (49 B); write lon lat z u' v' w' u_m
# v_m, alive (33 B); the footprint emission, a 4 B cell key and a 4 B value
# (8 B). Noise is generated in the kernel (0 B). Total 90 B. (The fixed-point
# accumulation replaces the emission with one 8 B integer atomic per run of
# equal cells in a wavefront, at most 8 B per particle-step.)
BYTES_PER_PARTICLE_STEP = 90
# Positions as an int32 grid point plus float32 lon/lat offsets and z (16 B
# instead of 12 B), read and written: 8 B more per particle-step.
CELL_EXTRA_BYTES = 8


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1440)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--met", choices=("f32", "bf16"), default="f32")
    ap.add_argument("--reorder-every", type=int, default=32,
                    help="re-bin the particles by met cell every N steps (0 = never; 32 is the autotuner's "
                         "choice on this workload with the fixed-point accumulation)")
    ap.add_argument("--step-block", type=int, default=128,
                    help="threads per workgroup of the step launch (64..512; 128 is the autotuner's choice on "
                         "this workload)")
    ap.add_argument("--graph", action="store_true",
                    help="time the captured step body (graph replay) as the headline run instead of "
                         "eager launches (both are always measured)")
    ap.add_argument("--kernel-steps", type=int, default=288,
                    help="steps of the separate kernel-timing pass (HIP events around every step-kernel "
                         "launch; at least --steps)")
    ap.add_argument("--accumulate", choices=("fixed", "sort"), default="fixed",
                    help="footprint accumulation: exact fixed-point sums inside the step kernel, or keys and "
                         "values sorted by cell and reduced in a fixed order")
    ap.add_argument("--no-other-accumulation", action="store_true",
                    help="skip timing the headline steps once more with the other footprint accumulation")
    ap.add_argument("--no-overlap", action="store_true",
                    help="sort accumulation: accumulate each step's footprint on the step's own stream")
    ap.add_argument("--autotune", action="store_true", help="tune the headline run's launch parameters too")
    ap.add_argument("--no-plain-torch", action="store_true", help="skip the PyTorch-ops plain method leg")
    ap.add_argument("--torch-steps", type=int, default=60,
                    help="timed steps of the PyTorch-ops plain method (60 = one met hour, convection included)")
    ap.add_argument("--no-batched", action="store_true", help="skip the 2,688-release network batch legs")
    ap.add_argument("--no-batched-bf16", action="store_true", help="skip the bfloat16 network batch leg")
    ap.add_argument("--batched-steps", type=int, default=0, help="time only the first N steps of the batch")
    ap.add_argument("--no-autotune", action="store_true", help="batched legs with the default launch parameters")
    ap.add_argument("--batched-block", type=int, default=0, help="batched legs: step workgroup size (no tuning)")
    ap.add_argument("--batched-reorder", type=int, default=-1, help="batched legs: re-binning interval (no tuning)")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving leg")
    ap.add_argument("--serve-seconds", type=float, default=60.0,
                    help="seconds of Poisson request arrivals (20 requests/s) in the serving leg (600 = the "
                         "full 10-minute stream)")
    ap.add_argument("--serve-out", default="", help="keep the serving leg's Zarr stores in this directory")
    ap.add_argument("--serve-no-convection", action="store_true", help="serving leg without convection")
    ap.add_argument("--no-ingest", action="store_true", help="skip the streamed-met leg")
    ap.add_argument("--ingest-workers", type=int, default=8, help="host decode threads of the streamed-met leg")
    ap.add_argument("--no-cpu-baseline", action="store_true", help="skip the CPU baseline and the float64 checks")
    ap.add_argument("--cpu-runs", type=int, default=3, help="timed CPU baseline runs per thread count (median)")
    ap.add_argument("--no-fp64-check", action="store_true", help="skip the float64 error checks")
    ap.add_argument("--check-releases", type=int, default=1,
                    help="releases of the batch run alone through the float64 oracle (1-3: the smallest 24 h, "
                         "48 h and 72 h releases)")
    ap.add_argument("--check-particles", type=int, default=16384,
                    help="particles of the per-step position check over the first met hour")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only run a float32 copy kernel of this many MiB (reads and writes that "
                         "many bytes) 5 times, to calibrate the HBM byte counters; prints one JSON line")
    return ap.parse_args()


def copy_calibration(mib: int) -> int:
    import torch

    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"calibration": "copy", "bytes_read": n * 4, "bytes_written": n * 4, "launches": 5}))
    return 0


# ---------------------------------------------------------------------------
# timing helpers
# ---------------------------------------------------------------------------

def _timed_steps(eng, K: int, sptr: int) -> float:
    """Wall seconds of the next K engine steps (the last accumulations
    included), synchronised on both sides."""

    import torch

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.advance(K, sptr)
    eng.join(sptr)
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def _kernel_pass(eng, n_steps: int, stream, sptr: int) -> dict:
    """From the release state, run ``n_steps`` eager steps, each step-kernel
    launch recording its own start and end (HIP events of the launch, on the
    launch stream: glide_step_timing_arm); the average duration and the active
    particles of the launches that had active particles."""

    import ctypes

    import numpy as np
    import torch

    from glide_amd import _lib as L

    eng.reset()
    eng.use_graph = False
    torch.cuda.synchronize()
    n = min(n_steps, len(eng.sched))
    L.check(L.lib().glide_step_timing_arm(n))
    for _ in range(n):
        eng.run_step(sptr)
    eng.join(sptr)
    torch.cuda.synchronize()
    ms = np.zeros(n, np.float32)
    got = L.lib().glide_step_timing_read(ms.ctypes.data_as(ctypes.c_void_p), n)
    if got != n:
        raise RuntimeError(f"{got} of {n} step launches timed")
    act = eng.active_steps.cpu().numpy()[:n].astype(np.float64)
    timed = [(i, float(ms[i])) for i in range(n) if act[i] > 0]
    t_ms = np.array([t for _, t in timed])
    t_act = np.array([act[i] for i, _ in timed])
    return {"launches": int(len(t_ms)), "avg_us": float(t_ms.mean() * 1e3) if len(t_ms) else float("nan"),
            "particles_per_launch": float(t_act.mean()) if len(t_act) else 0.0,
            "bytes_per_s": (BYTES_PER_PARTICLE_STEP * float(t_act.sum()) / (float(t_ms.sum()) * 1e-3)
                            if len(t_ms) else float("nan"))}


def _counters(workload: str, met: str) -> dict:
    """The step kernel's committed rocprofv3 counter summary for this
    workload (HBM bytes per launch from FETCH_SIZE / WRITE_SIZE scaled by a
    copy kernel's reading; VALU instructions per launch), if present."""

    p = ROOT / "profiles" / f"step_kernel_counters_{workload}_{met}.json"
    if not p.exists():
        return {}
    try:
        return json.loads(p.read_text())
    except (OSError, ValueError):
        return {}


def _roofline(kp: dict, kernel: str, counters: dict) -> dict:
    achieved = kp["bytes_per_s"] / 1e9
    out = {
        "bound": "hbm",
        "achieved": round(achieved, 2),
        "peak": PEAK_HBM_GBS,
        "unit": "GB/s",
        "frac": round(achieved / PEAK_HBM_GBS, 5),
        "traffic": counters.get("traffic_bytes_per_launch"),
        "kernel": kernel,
        "kernel_avg_us": round(kp["avg_us"], 3),
        "kernel_launches_timed": kp["launches"],
        "particles_per_launch": round(kp["particles_per_launch"]),
        "bytes_per_particle_step": BYTES_PER_PARTICLE_STEP,
        "bytes_per_launch": round(BYTES_PER_PARTICLE_STEP * kp["particles_per_launch"]),
    }
    valu = counters.get("valu_insts_per_launch")
    if valu:
        # issue-limited time of the launch's VALU instructions on every SIMD
        floor_us = valu * VALU_CYCLES / (N_SIMD * CLOCK_HZ) * 1e6
        out["valu"] = {"insts_per_launch": valu, "issue_floor_us": round(floor_us, 2),
                       "issue_frac": round(floor_us / kp["avg_us"], 4),
                       "model": f"{VALU_CYCLES} cycles per wave64 instruction, {N_SIMD} SIMDs at "
                                f"{CLOCK_HZ / 1e9:.1f} GHz"}
    if counters:
        out["counters_source"] = counters.get("source")
    return out


# ---------------------------------------------------------------------------
# the headline: plain FP32 method, 2^20 particles, on the HIP engine
# ---------------------------------------------------------------------------

def headline_leg(args, stream, sptr: int) -> tuple[dict, dict]:
    import numpy as np
    import torch

    from glide_amd import _lib as L
    from glide_amd.autotune import tune
    from glide_amd.runner import BatchEngine
    from glide_amd.synthetic import SyntheticMet, workload_config

    cfg = workload_config(WORKLOAD)
    met = SyntheticMet(cfg["spec"])
    met_dtype = L.GLIDE_MET_BF16 if args.met == "bf16" else L.GLIDE_MET_F32
    kw = dict(met_dtype=met_dtype, overlap=not args.no_overlap, accumulate=args.accumulate)
    if args.autotune:
        tuned = tune(cfg, met, use_cache=False, **kw)
    else:
        tuned = {"best": {"reorder_every": args.reorder_every, "step_block": args.step_block}, "trials": []}
    eng = BatchEngine(cfg, device="cuda", met=met, use_graph=True, **kw, **tuned["best"])
    n_sched = len(eng.sched)
    K = max(1, min(args.steps, n_sched))
    W = max(0, min(args.warmup, n_sched))

    def timed_run(graph: bool) -> float:
        """Warm up (captures the step body in graph mode), reset, time K steps."""
        eng.use_graph = graph
        eng.step_events = None
        eng.run(W)
        torch.cuda.synchronize()
        eng.reset()
        return _timed_steps(eng, K, sptr)

    other_elapsed = timed_run(not args.graph)  # the other launch mode (replay vs eager)
    elapsed = timed_run(args.graph)  # the headline run
    act = eng.active_steps.cpu().numpy()[:K].astype(np.float64)  # counted by the step kernel
    psteps = int(act.sum())
    eng.use_graph = False
    kp = _kernel_pass(eng, max(K, args.kernel_steps), stream, sptr)
    kernel = (f"glide::step_kernel<{str(met_dtype == L.GLIDE_MET_BF16).lower()}, 0, "
              f"{str(args.accumulate == 'fixed').lower()}, false>")
    roofline = _roofline(kp, kernel, _counters(WORKLOAD, args.met) if args.accumulate == "fixed" else {})
    roofline["whole_step_achieved"] = round(BYTES_PER_PARTICLE_STEP * psteps / elapsed / 1e9, 2)
    out = {
        "metric": METRIC,
        "value": psteps / elapsed,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": elapsed * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if met_dtype == L.GLIDE_MET_F32 else "bf16-met/f32",
        "data": "synthetic (seeded met cube and release; no network)",
        "config": {
            "workload": WORKLOAD,
            "particles": eng.n,
            "dt_s": int(cfg["dt"]),
            "steps_in_run": n_sched,
            "grid": f"{cfg['spec'].nx}x{cfg['spec'].ny}x{cfg['spec'].za}",
            "met_hours": len(eng.met.src),  # resident in HBM before the timed region
            "met_storage": args.met,
            "convection": "per-column emanuel",
            "meander": True,
            "max_substeps": eng.h.max_substeps,
            "reorder_every": eng.reorder_every,
            "step_block": eng.step_block or 256,
            "autotune_trials": tuned["trials"],
            "step_graph": bool(args.graph),
            "footprint_accumulation": args.accumulate,
            "footprint": f"1x{cfg['grid']['n_t']}x1x{cfg['grid']['n_y']}x{cfg['grid']['n_x']} f32",
        },
        "active_particle_steps": psteps,
        ("eager" if args.graph else "graph_replay"): {
            "ms_per_step": other_elapsed * 1e3 / K,
            "value": psteps / other_elapsed,
        },
        "roofline": roofline,
    }
    if not args.no_other_accumulation:
        # the same K steps with the other footprint accumulation, on the same met
        other = "sort" if args.accumulate == "fixed" else "fixed"
        oe = BatchEngine(cfg, device="cuda", met=met, device_met=eng.met, accumulate=other, overlap=True,
                         met_dtype=met_dtype, **tuned["best"])
        oe.run(W)
        torch.cuda.synchronize()
        oe.reset()
        oel = _timed_steps(oe, K, sptr)
        out[f"accumulate_{other}"] = {"ms_per_step": oel * 1e3 / K, "value": psteps / oel,
                                      "method": ("keys and values written by the step, stable radix sort by cell, "
                                                 "fixed-order float32 reduction on a second stream" if other == "sort"
                                                 else "exact int64 fixed-point sums inside the step kernel")}
        oe.close()
        del oe
    # the same K steps with positions held as a grid point plus float32 offsets
    ce = BatchEngine(cfg, device="cuda", met=met, device_met=eng.met, met_dtype=met_dtype, cell_positions=True,
                     accumulate=args.accumulate, **tuned["best"])
    ce.run(W)
    torch.cuda.synchronize()
    ce.reset()
    cel = _timed_steps(ce, K, sptr)
    out["positions_cell"] = {"ms_per_step": cel * 1e3 / K, "value": psteps / cel,
                             "layout": "int32 grid point + float32 lon/lat offsets (16 B) instead of float32 "
                                       "lon/lat/z (12 B)"}
    ce.close()
    del ce
    evidence = {"positions": {}}
    if not args.no_fp64_check and not args.no_cpu_baseline:
        for name, mdt, cell in (("f32", L.GLIDE_MET_F32, False), ("bf16", L.GLIDE_MET_BF16, False),
                                ("f32_cell", L.GLIDE_MET_F32, True), ("bf16_cell", L.GLIDE_MET_BF16, True)):
            evidence["positions"][name] = position_snapshots(cfg, met, mdt, args.check_particles, cell=cell)
    eng.close()
    del eng
    torch.cuda.empty_cache()
    return out, evidence


def position_snapshots(cfg, met, met_dtype: int, n_particles: int, n_steps: int = 60, cell: bool = False) -> dict:
    """The headline release with ``n_particles`` particles stepped by the
    engine over the first met hour; the particle state (in particle-id order,
    positions exactly as the layout holds them) before the first step and
    after each step, for the per-step check against the float64 oracle."""

    import dataclasses

    import torch

    from glide_amd.runner import BatchEngine

    rel = dataclasses.replace(cfg["releases"][0], n_particles=int(n_particles))
    c = dict(cfg, releases=[rel])
    eng = BatchEngine(c, device="cuda", met=met, met_dtype=met_dtype, cell_positions=cell)
    snaps = [eng.state_numpy()]
    steps = eng.sched[:n_steps]
    for _ in steps:
        eng.run_step()
        snaps.append(eng.state_numpy())
    torch.cuda.synchronize()
    out = {"cfg": c, "sched": steps, "snaps": snaps, "length_s": eng.length_s, "seed": eng.seed,
           "r_off": eng.ra.r_off.copy(), "lo_off": eng.ra.lo_off.copy(), "uid": eng.ra.uid.copy()}
    eng.close()
    return out


# ---------------------------------------------------------------------------
# the plain method as PyTorch ops on the GPU (the comparison baseline)
# ---------------------------------------------------------------------------

def plain_torch_leg(args) -> dict:
    """The plain FP32 method as PyTorch ops on the same GPU and workload
    (grid_sample, searchsorted, randn, the fixed substep loop, materialised
    convection rows, scatter_add_): the first --torch-steps steps of the
    schedule (one met hour by default, its first step convecting), eager and
    with the step body captured as one graph, each after a 3-step warm-up and
    a reset."""

    import torch

    from glide_amd.runner import BatchEngine
    from glide_amd.synthetic import SyntheticMet, workload_config
    from glide_amd.torch_plain import TorchPlain

    cfg = workload_config(WORKLOAD)
    met = SyntheticMet(cfg["spec"])
    eng = BatchEngine(cfg, device="cuda", met=met)
    K = max(1, min(args.torch_steps, len(eng.sched)))
    res = {}
    for mode in ("eager", "graph"):
        tp = TorchPlain.from_engine(eng, seed=2, convection=True, graph=mode == "graph")
        tp.run(3)
        torch.cuda.synchronize()
        tp.reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        tp.run(K)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        psteps = tp.active_steps
        res[mode] = {"value": psteps / elapsed, "ms_per_step": elapsed * 1e3 / K, "active_particle_steps": psteps}
        del tp
        torch.cuda.empty_cache()
    eng.close()
    del eng
    torch.cuda.empty_cache()
    return {"workload": WORKLOAD, "value": res["graph"]["value"], "unit": "particle-steps/s", "steps": K,
            "ms_per_step": res["graph"]["ms_per_step"], "eager": res["eager"],
            "convection": "per column, once per met hour (the first timed step)",
            "method": "plain FP32 as PyTorch ops (grid_sample, randn, scatter_add_), step body captured as one "
                      "graph (value) and eager"}


# ---------------------------------------------------------------------------
# the network batch: 2,688 releases in one engine
# ---------------------------------------------------------------------------

BATCH_CANDIDATES = [dict(step_block=256, reorder_every=r) for r in (0, 16, 64)] + [
    dict(step_block=128, reorder_every=16), dict(step_block=512, reorder_every=16)]


def sampled_releases(cfg, n: int = 1) -> list[int]:
    """Releases of the batch checked against the float64 oracle: for the
    first ``n`` of the backward lengths 24, 48, 72 h, the release of that
    length with the fewest particles (the oracle runs one 24 h release of
    about 2k particles in about 20 s of one core, in float64 and again in
    float32)."""

    out = []
    for hours in (24, 48, 72)[:n]:
        rs = [r for r in cfg["releases"] if r.length_s == hours * 3600]
        if rs:
            out.append(min(rs, key=lambda r: (r.n_particles, r.release_idx)).release_idx)
    return out


def batched_leg(args, name: str, stream, sptr: int, shared: dict, evidence: dict) -> dict:
    """The whole network batch in one engine on one GPU: releases of 2k-20k
    particles and 24-120 h backward lengths, each stepped while the cursor is
    above its own terminus, footprints in per-release slabs of exact
    fixed-point sums (64-bit cells), float64 release totals, the live set
    compacted every met hour. Times the whole schedule."""

    import numpy as np
    import torch

    from glide_amd import _lib as L
    from glide_amd.autotune import tune
    from glide_amd.runner import BatchEngine, DeviceMet
    from glide_amd.synthetic import SyntheticMet, workload_config

    cfg = workload_config(name)
    bf16 = name.endswith("-bf16")
    # the bfloat16 leg is the precision layout: positions as an int32 grid
    # point plus float32 offsets as well (4 B more state read and written)
    cell = bf16
    bpp = BYTES_PER_PARTICLE_STEP + (CELL_EXTRA_BYTES if cell else 0)
    met_dtype = L.GLIDE_MET_BF16 if bf16 else L.GLIDE_MET_F32
    if "met" not in shared:
        shared["met"] = SyntheticMet(cfg["spec"])
    met = shared["met"]
    from glide_amd.schedule import cursor_schedule

    sched = cursor_schedule(cfg["releases"], int(cfg["dt"]))
    hours = sorted({s.bracket_hour for s in sched} | {s.bracket_hour + 1 for s in sched})
    t_setup = time.perf_counter()
    dm = DeviceMet(met, hours, device="cuda", met_dtype=met_dtype, sources=shared.get("sources"))
    shared["sources"] = dm.src
    kw = dict(met_dtype=met_dtype, compact=True, release_terminus=True, accumulate="fixed")
    if args.batched_block or args.batched_reorder >= 0:
        best = {}
        if args.batched_block:
            best["step_block"] = args.batched_block
        if args.batched_reorder >= 0:
            best["reorder_every"] = args.batched_reorder
        tuned = {"best": best, "trials": []}
    elif args.no_autotune:
        tuned = {"best": {}, "trials": []}
    elif "tuned" in shared:
        tuned = {"best": shared["tuned"], "trials": [], "from": BATCHED}
    else:
        tuned = tune(cfg, met, candidates=BATCH_CANDIDATES, use_cache=False, stream=stream, device_met=dm, **kw)
        shared["tuned"] = tuned["best"]
    eng = BatchEngine(cfg, device="cuda", met=met, device_met=dm, cell_positions=cell, **kw, **tuned["best"])
    n_sched = len(eng.sched)
    K = n_sched if args.batched_steps <= 0 else min(args.batched_steps, n_sched)
    eng.run(min(20, n_sched))  # warm-up, then back to the release state
    torch.cuda.synchronize()
    eng.reset()
    setup_s = time.perf_counter() - t_setup
    import ctypes

    L.check(L.lib().glide_step_timing_arm(K))  # every step launch times itself (its own events)
    elapsed = _timed_steps(eng, K, sptr)
    act = eng.active_steps.cpu().numpy()[:K].astype(np.float64)
    psteps = int(act.sum())
    ms = np.zeros(K, np.float32)
    got = L.lib().glide_step_timing_read(ms.ctypes.data_as(ctypes.c_void_p), K)
    if got == K:  # one launch per step
        timed = [(i, float(ms[i])) for i in range(K) if act[i] > 0]
        t_ms = np.array([t for _, t in timed])
        t_act = np.array([act[i] for i, _ in timed])
    else:  # steps over an empty live set launch nothing: pair the launches with the non-empty steps
        t_act = act[act > 0][:got]
        t_ms = ms[:t_act.size].astype(np.float64)
    kp = {"launches": int(len(t_ms)), "avg_us": float(t_ms.mean() * 1e3) if len(t_ms) else float("nan"),
          "particles_per_launch": float(t_act.mean()) if len(t_act) else 0.0,
          "bytes_per_s": (bpp * float(t_act.sum()) / (float(t_ms.sum()) * 1e-3)
                          if len(t_ms) else float("nan"))}
    roofline = _roofline(kp, f"glide::step_kernel<{str(bf16).lower()}, 0, true, {str(cell).lower()}>", {})
    roofline["bytes_per_particle_step"] = bpp
    roofline["bytes_per_launch"] = round(bpp * kp["particles_per_launch"])
    roofline["whole_step_achieved"] = round(bpp * psteps / elapsed / 1e9, 2)
    g = cfg["grid"]
    tot = eng.release_total.cpu().numpy()
    if K == n_sched:
        for r in sampled_releases(cfg, args.check_releases):
            evidence.setdefault(r, {})[name] = {
                "footprint": eng.release_footprint(r).cpu().numpy().astype(np.float64),
                "total": float(tot[r])}
    out = {
        "workload": name,
        "value": psteps / elapsed,
        "unit": "particle-steps/s",
        "ms_per_step": elapsed * 1e3 / K,
        "seconds": elapsed,
        "steps": K,
        "steps_in_run": n_sched,
        "active_particle_steps": psteps,
        "particles": eng.n,
        "releases": len(cfg["releases"]),
        "met_hours": len(hours),
        "met_storage": "bf16" if bf16 else "f32",
        "footprint": (f"{g['n_r']} per-release slabs x up to {g['n_t']} time-ago bins x 1 x {g['n_y']} x {g['n_x']} "
                      f"f32 from exact int64 fixed-point sums ({eng.acc_cells} cells)"),
        "release_totals": "f64",
        "setup_s": setup_s,
        "device_gb_allocated": round(torch.cuda.max_memory_allocated() / 1e9, 1),
        "autotune": {"chosen": tuned["best"], "trials": tuned["trials"], **({"from": tuned["from"]}
                                                                            if "from" in tuned else {})},
        "positions": "int32 grid point + float32 lon/lat offsets" if cell else "float32 lon/lat/z",
        "roofline": roofline,
    }
    eng.close()
    del eng
    torch.cuda.empty_cache()
    # the other position layout (int32 grid point + float32 offsets): the first
    # steps of the batch timed, and the sampled releases alone against the
    # float64 oracle (a release alone equals it in the batch)
    kc = min(2000, n_sched)
    times = {}
    for cell in (False, True):
        e2 = BatchEngine(cfg, device="cuda", met=met, device_met=dm, cell_positions=cell, **kw, **tuned["best"])
        times[cell] = _timed_steps(e2, kc, sptr)
        e2.close()
        del e2
        torch.cuda.empty_cache()
    out["positions_cell"] = {"steps": kc, "ms_per_step": times[True] * 1e3 / kc,
                             "float32_positions_ms_per_step": times[False] * 1e3 / kc,
                             "speed_vs_float32_positions": times[False] / times[True]}
    if K == n_sched:
        import dataclasses

        for r in sampled_releases(cfg, args.check_releases):
            rel = cfg["releases"][r]
            one = dataclasses.replace(rel, release_idx=0, noise_id=rel.release_idx)
            c1 = dict(cfg, releases=[one], grid=dict(cfg["grid"], n_r=1))
            e1 = BatchEngine(c1, device="cuda", met=met, device_met=dm, release_terminus=True, met_dtype=met_dtype,
                             cell_positions=True)
            e1.run()
            evidence[r][f"{name} alone, cell positions"] = {
                "footprint": e1.release_footprint(0).cpu().numpy().astype(np.float64),
                "total": float(e1.release_total.cpu().numpy()[0])}
            e1.close()
            del e1
    del dm
    torch.cuda.empty_cache()
    return out


def serving_leg(args, plain_value: float) -> dict:
    """Continuous-batching serving: Poisson requests (20/s) on a 240x160x60
    met ring of 240 distinct hourly brackets in bfloat16, each request a
    release of 1k-50k particles, 6-240 h back, at one of 100 sites, with
    per-column convection on each request's first step of every met hour.
    Requests are admitted as slots, particle-pool blocks and footprint-arena
    cells are free (pool and arena grow within free HBM), and are stepped
    together on their own cursors; finished footprints are written to one zlib
    Zarr store per request on host threads, and latency is arrival to
    footprint store on disk (the stores go to a temporary directory, removed
    afterwards, unless --serve-out names one)."""

    import shutil
    import tempfile

    import numpy as np
    import torch

    from glide_amd import _lib as L
    from glide_amd.runner import ConvectionOptions
    from glide_amd.serving import MetRing, ResultWriter, Server, latency_percentiles
    from glide_amd.synthetic import MetSpec, SyntheticMet, serving_requests, workload_config

    t0 = time.perf_counter()
    spec = MetSpec(nx=240, ny=160, za=60, n_hours=241, seed=20240105, lon0=2.0, lat0=72.0)
    met = SyntheticMet(spec)
    ring = MetRing(met, first_hour=0, n_hours=240, met_dtype=L.GLIDE_MET_BF16, convection=not args.serve_no_convection)
    ring_s = time.perf_counter() - t0
    base = workload_config("plain-gpu-1M-24h")
    reqs = serving_requests(spec, rate_per_s=20.0, seconds=args.serve_seconds)
    out_dir = args.serve_out or tempfile.mkdtemp(prefix="glide_serve_")
    writer = ResultWriter(out_dir, workers=4)
    srv = Server(ring, base["grid"], base["domain"], max_slots=1024, n_t_max=240, seed=17, writer=writer,
                 convection=None if args.serve_no_convection else ConvectionOptions())
    # warm-up: one request through, not timed
    warm = serving_requests(spec, rate_per_s=20.0, seconds=0.2, seed=99)[:1]
    for r in warm:
        srv.admit(r)
    srv.drain()
    stats = srv.serve(reqs, progress_every=500)
    res = stats.results
    lat = latency_percentiles(res)
    day = [r for r in res if 18 * 3600 <= r.req.length_s <= 30 * 3600]
    lat_day = latency_percentiles(day)
    tenk = [r for r in day if 5_000 <= r.req.n_particles <= 20_000]
    lat_tenk = latency_percentiles(tenk)
    value = stats.particle_steps / stats.busy_s if stats.busy_s > 0 else float("nan")
    out = {
        "workload": SERVING,
        "value": value,
        "unit": "particle-steps/s",
        "requests": len(reqs),
        "completed": len(res),
        "arrival_rate_per_s": 20.0,
        "arrival_seconds": args.serve_seconds,
        "busy_seconds": stats.busy_s,
        "engine_steps": stats.steps,
        "active_particle_steps": stats.particle_steps,
        "latency_s": {"p50": lat["p50"], "p99": lat["p99"]},
        "latency_18_30h_s": {"p50": lat_day["p50"], "p99": lat_day["p99"], "n": len(day)},
        "latency_5k_20k_particles_18_30h_s": {"p50": lat_tenk["p50"], "p99": lat_tenk["p99"], "n": len(tenk)},
        "met_storage": "bf16 ring, 240 distinct hourly brackets",
        "ring_gb": round(ring.nbytes / 1e9, 2),
        "ring_build_s": round(ring_s, 1),
        "convection": "per column, each request's first step of every met hour" if srv.conv else "off",
        "convect_launches": srv.convect_launches,
        "latency_to": "footprint Zarr store (zlib) on disk",
        "zarr_stores": len(writer.paths),
        "zarr_bytes": sum(f.stat().st_size for p in writer.paths for f in Path(p).rglob("*") if f.is_file()),
        "admission": {"max_slots": srv.max_slots, "pool_particles": srv.cap, "arena_cells": srv.arena.size,
                      "device_gb": round(srv.device_bytes() / 1e9, 2),
                      "rule": "slot, pool blocks and arena cells free; pool and arena grow within free HBM"},
        "vs_plain": value / plain_value if plain_value > 0 else None,
        "mean_particles": float(np.mean([r.n_particles for r in reqs])) if reqs else 0.0,
    }
    writer.close()
    if not args.serve_out:
        shutil.rmtree(out_dir, ignore_errors=True)
    del srv, ring
    torch.cuda.empty_cache()
    return out


def ingest_leg(args, sptr: int, plain_value: float) -> dict:
    """The plain workload with its met streamed instead of resident: source
    hours decoded on host worker threads into pinned memory and uploaded on a
    copy stream ahead of the cursor (glide_amd.ingest.MetIngest). Times the
    whole schedule from the first hour's decode. The synthetic decode (≈0.15 s
    of one core per 240x160 hour) is slower than the GPU's hour (≈10 ms), so
    this rate is the host decode's, and the reported wait is the time the
    caller was blocked on it."""

    import torch

    from glide_amd.ingest import MetIngest
    from glide_amd.runner import BatchEngine
    from glide_amd.synthetic import SyntheticMet, workload_config

    cfg = workload_config(WORKLOAD)
    met = SyntheticMet(cfg["spec"])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    with MetIngest(met, workers=args.ingest_workers, lookahead=2 * args.ingest_workers, max_resident=6) as ing:
        eng = BatchEngine(cfg, device="cuda", met=met, ingest=ing)
        for _ in range(len(eng.sched)):
            eng.run_step(sptr)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        st = dict(ing.stats)
    psteps = int(eng.active_steps.cpu().numpy().sum())
    eng.close()
    del eng
    torch.cuda.empty_cache()
    return {
        "workload": WORKLOAD,
        "value": psteps / elapsed,
        "unit": "particle-steps/s",
        "seconds": elapsed,
        "vs_resident": psteps / elapsed / plain_value if plain_value > 0 else None,
        "workers": args.ingest_workers,
        "hours_decoded": st["decoded"],
        "decode_cpu_s": st["decode_s"],
        "caller_wait_s": st["wait_s"],
        "upload_gb": st["upload_bytes"] / 1e9,
        "hours_per_s": st["uploads"] / elapsed,
    }



# ---------------------------------------------------------------------------
# the CPU baseline and the float64 checks (the oracle is the checker here)
# ---------------------------------------------------------------------------

def _cpu_model() -> str:
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
    except (OSError, subprocess.SubprocessError):
        txt = ""
    for line in txt.splitlines():
        if line.startswith("Model name:"):
            return line.split(":", 1)[1].strip()
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline_leg(args) -> dict:
    """The plain method as PyTorch ops on the host's CPU cores: the
    reference's default CPU path (the dynamic step over the active particles,
    max(k) substeps) restated in glide_amd.torch_plain, float32, on the whole
    10k-particle, 3 h workload (64x48x40 met, convection off), at
    torch.set_num_threads(n) for n = os.cpu_count() and n = 16. Per thread
    count: one warm-up run, then the median of --cpu-runs runs of the step loop
    (windows prepared; one run when the warm-up took over 20 s), and one run
    of the whole run (window building from the synthetic met included)."""

    import numpy as np
    import torch

    from glide_amd.synthetic import SyntheticMet, workload_config
    from glide_amd.torch_plain import TorchPlain
    from oracle import run as ORUN

    cfg = workload_config(CPU_WORKLOAD)
    met = SyntheticMet(cfg["spec"])
    geo = dict(lon0=float(met.lon[0]), lon1=float(met.lon[-1]), lat0=float(met.lat[0]), lat1=float(met.lat[-1]))
    levels = cfg["spec"].agl_levels
    prepared = {}
    fresh = ORUN.torch_windows(met)

    def cached(bh):
        if bh not in prepared:
            prepared[bh] = fresh(bh)
        return prepared[bh]

    def one(windows):
        t0 = time.perf_counter()
        tp = TorchPlain(cfg, windows, device="cpu", levels=levels, geometry=geo, seed=3, dynamic=True)
        t1 = time.perf_counter()
        tp.run(len(tp.sched))
        t2 = time.perf_counter()
        fp = tp.footprint.numpy()
        total = float(fp.sum())
        return tp.active_steps, t2 - t1, time.perf_counter() - t0, total

    n_cpu = os.cpu_count() or 1
    threads = sorted({min(n_cpu, 256), 16})
    prev = torch.get_num_threads()
    runs = {}
    try:
        for n in threads:
            torch.set_num_threads(n)
            _, warm_s, _, _ = one(cached)  # warm-up (also prepares the windows)
            # bounded: a thread count far above the cores this process may use
            # runs slowly; then the median is over fewer runs
            n_runs = max(1, args.cpu_runs) if warm_s < 20.0 else 1
            step_s = []
            for _ in range(n_runs):
                ps, s1, _, _ = one(cached)
                step_s.append(s1)
            _, _, e2e, _ = one(ORUN.torch_windows(met))  # windows built from the synthetic met again
            runs[n] = {"threads": n, "step_loop_s": statistics.median(step_s), "end_to_end_s": e2e,
                       "particle_steps": ps, "value": ps / statistics.median(step_s), "end_to_end_value": ps / e2e,
                       "runs": len(step_s), "warmup_s": warm_s}
    finally:
        torch.set_num_threads(prev)
    best = max(runs.values(), key=lambda r: r["value"])
    return {
        "value": best["value"],
        "unit": "particle-steps/s",
        "cores": best["threads"],
        "kind": "port",
        "sample": (f"{CPU_WORKLOAD}: 10,000 particles x 180 steps ({best['particle_steps']} particle-steps), "
                   f"the reference's dynamic CPU step as float32 PyTorch ops (glide_amd.torch_plain, "
                   f"dynamic=True); median of {best['runs']} step-loop runs after 1 warm-up at "
                   f"{best['threads']} threads"),
        "by_threads": [runs[n] for n in threads],
        "os_cpu_count": n_cpu,
        "affinity_cpus": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None,
        "cpu_model": _cpu_model(),
        "end_to_end_value": best["end_to_end_value"],
    }


def position_errors(ev: dict) -> dict:
    """Per-step position errors of the engine over the first met hour: each
    engine step from its own float32 state against one float64 oracle step
    from the same state, on the same Philox noise (convection included on the
    hour's first step). Relative errors |d lon| / |lon|, |d lat| / |lat| and
    |d z| / max(|z|, |z_pre|, 10 m) of the active particles."""

    import numpy as np

    from oracle import lpdm as O
    from oracle import run as ORUN

    from glide_amd.synthetic import SyntheticMet

    cfg = ev["cfg"]
    met = SyntheticMet(cfg["spec"])
    levels = cfg["spec"].agl_levels
    hours = {}

    def hour(h):
        if h not in hours:
            sh = met.hour(h)
            from oracle import resample as ors

            hours[h] = ors.resample_source_hour(sh, met, levels)
        return hours[h]

    conv = ORUN.column_convection(met, hours, ev["seed"]) if cfg.get("convection") else None
    p = O.HannaParams()
    windows = {}
    order = np.argsort(ev["uid"], kind="stable")  # snapshots are in particle-id order
    r_off, lo_off = ev["r_off"][order], ev["lo_off"][order]
    worst = {"lon": 0.0, "lat": 0.0, "z": 0.0}
    sums = {"lon": 0.0, "lat": 0.0, "z": 0.0}
    n_tot = 0
    frac_ok = []
    last_bracket = None
    for st, pre, post in zip(ev["sched"], ev["snaps"][:-1], ev["snaps"][1:]):
        bh = st.bracket_hour
        if bh not in windows:
            windows[bh] = ORUN.build_window(hour(bh), hour(bh + 1), levels, met.lon, met.lat, np.float64)
        co = np.float32(st.cursor_off)
        active = (r_off >= co) & (lo_off <= co) & pre["alive"]
        S = {k: pre[k].astype(np.float64) for k in ("lon", "lat", "z", "up", "vp", "wp", "um", "vm")}
        S["uid"] = pre["uid"]
        ref = O.step(S, active, windows[bh], st.alpha, float(st.delta_s), p, seed=ev["seed"],
                     step_index=st.noise_step)
        if bh != last_bracket:
            if conv is not None:
                ref = conv(ref, active, bh, st.alpha, st.noise_step)
            last_bracket = bh
        if not active.any():
            continue
        e = {"lon": np.abs(post["lon"][active] - ref["lon"][active]) / np.abs(ref["lon"][active]),
             "lat": np.abs(post["lat"][active] - ref["lat"][active]) / np.abs(ref["lat"][active]),
             "z": np.abs(post["z"][active] - ref["z"][active])
             / np.maximum(np.maximum(np.abs(ref["z"][active]), np.abs(S["z"][active])), 10.0)}
        for k in e:
            worst[k] = max(worst[k], float(e[k].max()))
            sums[k] += float(e[k].sum())
        n_tot += int(active.sum())
        frac_ok.append(float((np.maximum(np.maximum(e["lon"], e["lat"]), e["z"]) <= 1e-5).mean()))
    return {"steps": len(frac_ok), "particles": int(ev["snaps"][0]["lon"].size),
            "mean_rel": {k: sums[k] / max(n_tot, 1) for k in sums}, "max_rel": worst,
            "min_frac_within_1e-5": min(frac_ok) if frac_ok else None,
            "mean_frac_within_1e-5": float(np.mean(frac_ok)) if frac_ok else None}


def release_errors(r: int, results: dict) -> dict:
    """One release of the network batch run alone through the float64 oracle
    (its own terminus, its particle ids, the batch's seed and per-column
    convection): relative L2 of the engine's footprint slab and the relative
    error of its residence total, per engine run."""

    import dataclasses

    import numpy as np

    from oracle import run as ORUN

    from glide_amd.synthetic import SyntheticMet, workload_config

    cfg = workload_config(BATCHED)
    rel = cfg["releases"][r]
    one = dataclasses.replace(rel, release_idx=0, noise_id=rel.release_idx)
    c1 = dict(cfg, releases=[one], grid=dict(cfg["grid"], n_r=1))
    met = SyntheticMet(cfg["spec"])
    hours = {}
    t0 = time.perf_counter()
    ref = ORUN.run(c1, np.float64, met=met, hours_cache=hours, convection=ORUN.column_convection(met, hours, c1["seed"]),
                   release_terminus=True)
    g = cfg["grid"]
    rf = ref.footprint.reshape(g["n_t"], g["n_z"] * g["n_y"] * g["n_x"])
    out = {"release": r, "particles": one.n_particles, "length_h": one.length_s // 3600,
           "oracle_seconds": round(time.perf_counter() - t0, 1), "oracle_particle_steps": ref.particle_steps}
    rtot = float(ref.footprint.sum())
    # the float32 oracle on the same noise: how far float32 arithmetic alone
    # carries the full run from the float64 one (the envelope of a float32 engine)
    h32 = {}
    f32 = ORUN.run(c1, np.float32, met=met, hours_cache=h32, convection=ORUN.column_convection(met, h32, c1["seed"]),
                   release_terminus=True)
    ff = f32.footprint.astype(np.float64)
    out["float32_oracle"] = {"footprint_rel_l2": float(np.linalg.norm(ff - ref.footprint) / np.linalg.norm(ref.footprint)),
                             "total_rel_err": abs(float(ff.sum()) - rtot) / rtot}
    for name, res in results.items():
        fp = res["footprint"].reshape(-1, g["n_z"] * g["n_y"] * g["n_x"])
        nt = fp.shape[0]
        beyond = float(np.abs(rf[nt:]).sum())  # residence in bins past the slab (none expected)
        l2 = float(np.linalg.norm(fp - rf[:nt]) / np.linalg.norm(rf[:nt]))
        out[name] = {"footprint_rel_l2": l2, "total_rel_err": abs(res["total"] - rtot) / rtot,
                     "oracle_residence_beyond_slab": beyond}
    return out


def fp64_checks(evidence: dict, note=lambda msg: None) -> dict:
    out = {}
    for name, ev in evidence.get("positions", {}).items():
        out[f"positions_first_met_hour_{name}_met"] = position_errors(ev)
        note(f"positions {name}: checked")
    rel = []
    for r, res in sorted(evidence.get("releases", {}).items()):
        rel.append(release_errors(r, res))
        note(f"release {r}: checked")
    if rel:
        out["batched_releases"] = rel
    out["method"] = ("float64 oracle (oracle/) on the same Philox noise; positions: one oracle step from each "
                     "engine step's float32 start state; releases: the release alone, full run")
    return out


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(the path would shard by release with no collective in the step)", file=sys.stderr)
        return 2
    if args.copy_calibration_mib:
        return copy_calibration(args.copy_calibration_mib)
    import torch

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        return 2
    torch.cuda.set_device(0)
    stream = torch.cuda.Stream()  # default priority; graph capture needs a non-default stream
    sptr = int(stream.cuda_stream)
    torch.cuda.set_stream(stream)

    def note(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    out, evidence = headline_leg(args, stream, sptr)
    note(f"headline {out['value']:.4g} particle-steps/s")
    if not args.no_plain_torch:
        out["plain_torch"] = plain_torch_leg(args)
        pt = out["plain_torch"]["value"]
        out["vs_plain_torch"] = out["value"] / pt if pt > 0 else None
        note(f"plain torch {pt:.4g}")
    if not args.no_batched:
        shared: dict = {}
        evidence["releases"] = {}
        names = [BATCHED] + ([] if args.no_batched_bf16 else [BATCHED_BF16])
        for name in names:
            key = "batched_bf16" if name.endswith("-bf16") else "batched"
            out[key] = batched_leg(args, name, stream, sptr, shared, evidence["releases"])
            out[key]["vs_plain"] = out[key]["value"] / out["value"]
            if "plain_torch" in out:
                out[key]["vs_plain_torch"] = out[key]["value"] / out["plain_torch"]["value"]
            note(f"{name} {out[key]['value']:.4g}")
        shared.clear()
        torch.cuda.empty_cache()
    if not args.no_serving:
        out["serving"] = serving_leg(args, out["value"])
        note("serving done")
    if not args.no_ingest:
        out["ingest"] = ingest_leg(args, sptr, out["value"])
        note("ingest done")
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_leg(args)
        note(f"cpu baseline {out['cpu_baseline']['value']:.4g}")
        if not args.no_fp64_check:
            out["error_vs_fp64"] = fp64_checks(evidence, note)
            note("fp64 checks done")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
