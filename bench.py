"""Throughput benchmark of the backward-LPDM step engine on one MI355X.

Workload ("plain-gpu-1M-24h"): one release of 2^20 particles, 24 h backward at
dt = 60 s (1,440 steps), full physics (Hanna three-regime turbulence with the
Richardson closure aloft, meander on, per-column Emanuel convection once per
met hour), synthetic met on a 240 x 160 grid at 0.25 deg with 60 AGL levels and
25 hourly snapshots, footprint 24 time-ago bins x 1 layer x 160 x 240 in float32.

A "step" is one model time step of that run over all 2^20 particles: the fused
step kernel (which, by default, also adds every particle's residence to the
footprint's exact fixed-point sums), the re-binning of the particles by met
cell every few steps and, on a bracket change, the hourly resample, window
build and convection. The float32 footprint is produced from the sums at the
end of the timed region, inside it. The met source hours are
resident in HBM before the timed region starts. ``value`` counts active
particle-steps per second of wall time over the timed steps.

Prints one JSON line. ``--gpus N`` with N > 1 exits: this engine runs on one GPU.
"""

from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

METRIC = "particle-steps/s, batched vs plain at stated error; p50/p99 footprint latency"
WORKLOAD = "plain-gpu-1M-24h"
PEAK_HBM_GBS = 8000.0
# Algorithmic HBM bytes per active particle-step of the fused step kernel
# (float32 structure-of-arrays state): read lon lat z u' v' w' u_m v_m, weight,
# r_off, we_off, rel_idx, alive (49 B); write lon lat z u' v' w' u_m v_m, alive
# (33 B). Noise is generated in the kernel (0 B). With the sort accumulation
# the kernel also writes a 4 B cell key and a 4 B value (8 B; total 90 B); with
# the fixed-point accumulation it adds to the footprint sums instead, one 8 B
# integer atomic per run of equal cells in a wavefront, not per particle
# (total 82 B, the atomics not counted).
BYTES_PER_PARTICLE_STEP = {"fixed": 82, "sort": 90}


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1440)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--met", choices=("f32", "bf16"), default="f32")
    ap.add_argument("--reorder-every", type=int, default=32,
                    help="re-bin the particles by met cell every N steps (0 = never; 32 is the autotuner's "
                         "choice on this workload with the fixed-point accumulation: binned particles gather "
                         "from fewer cache lines and sum longer runs per atomic, and the gain decays as they "
                         "move)")
    ap.add_argument("--step-block", type=int, default=128,
                    help="threads per workgroup of the step launch (64..512; 128 is the autotuner's choice on "
                         "this workload)")
    ap.add_argument("--graph", action="store_true",
                    help="time the captured step body (graph replay) as the headline run instead of "
                         "eager launches (both are always measured)")
    ap.add_argument("--time-every", type=int, default=7,
                    help="time the step kernel with HIP events on every Nth timed step (in graph "
                         "mode those steps launch eagerly); 7 is coprime with the re-binning "
                         "intervals, so the samples cover every phase of the binning cycle")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the network-batch leg (batched-gpu-336r, bfloat16 met, compensated footprint)")
    ap.add_argument("--batched-steps", type=int, default=0,
                    help="time only the first N steps of the network batch (0 = all)")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving leg")
    ap.add_argument("--no-ingest", action="store_true", help="skip the streamed-met leg")
    ap.add_argument("--no-autotune", action="store_true", help="batched leg with the default launch parameters")
    ap.add_argument("--batched-block", type=int, default=0, help="batched leg: step workgroup size (no tuning)")
    ap.add_argument("--batched-reorder", type=int, default=-1, help="batched leg: re-binning interval (no tuning)")
    ap.add_argument("--autotune", action="store_true",
                    help="tune the headline run's launch parameters too (its trials then show in a kernel trace)")
    ap.add_argument("--accumulate", choices=("fixed", "sort"), default="fixed",
                    help="footprint accumulation: exact fixed-point sums inside the step kernel, or keys and "
                         "values sorted by cell and reduced in a fixed order")
    ap.add_argument("--no-other-accumulation", action="store_true",
                    help="skip timing the headline steps once more with the other footprint accumulation")
    ap.add_argument("--no-overlap", action="store_true",
                    help="sort accumulation: accumulate each step's footprint on the step's own stream "
                         "(no second stream)")
    ap.add_argument("--serve-out", default="", help="keep the serving leg's Zarr stores in this directory")
    ap.add_argument("--no-plain-torch", action="store_true", help="skip the PyTorch-ops plain method leg")
    ap.add_argument("--torch-steps", type=int, default=60, help="timed steps of the PyTorch-ops plain method")
    ap.add_argument("--ingest-workers", type=int, default=8, help="host decode threads of the streamed-met leg")
    ap.add_argument("--serve-seconds", type=float, default=30.0,
                    help="seconds of Poisson request arrivals (20 requests/s) in the serving leg")
    ap.add_argument("--cpu-particles", type=int, default=262144)
    ap.add_argument("--cpu-steps", type=int, default=40)
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


def cpu_baseline(cfg, n_particles: int, n_steps: int) -> dict:
    """The oracle's float32 plain method (the reference's CPU path, restated)
    on a bounded sample of the same workload: the first ``n_steps`` steps of
    the run with ``n_particles`` particles of the same release on the same
    grid. Times the step loop only."""

    import numpy as np

    from oracle import lpdm as O
    from oracle import run as ORUN

    rel = cfg["releases"][0]
    sample = dict(cfg)
    sample["releases"] = [type(rel)(**{**rel.__dict__, "n_particles": n_particles})]
    sample["convection"] = False
    res = ORUN.run(sample, dty=np.float32, params=O.HannaParams(), max_steps=n_steps)
    return {
        "value": res.particle_steps / res.step_seconds,
        "unit": "particle-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"first {n_steps} steps of {WORKLOAD} with {n_particles} particles "
                   f"(float32 numpy restatement of the reference's plain method, "
                   f"{res.particle_steps} particle-steps in {res.step_seconds:.2f} s, one core)"),
    }


BATCHED = "batched-gpu-336r"
SERVING = "serving-gpu-10d"


def serving_leg(args, plain_value: float) -> dict:
    """Continuous-batching serving: Poisson requests (20/s) on a 240x160x60
    met ring of 240 hourly brackets in bfloat16 (a 24 h synthetic cycle
    repeated over the 10-day span), each request a release of 1k-50k particles,
    6-240 h back, at one of 100 sites. Requests are admitted as slots and pool
    blocks free up and are stepped together on their own cursors; finished
    footprints are written to one zlib Zarr store per request on host threads,
    and latency is arrival to footprint store on disk (the stores go to a
    temporary directory, removed afterwards, unless --serve-out names one)."""

    import numpy as np
    import torch

    from glide_amd import _lib as L
    import shutil
    import tempfile

    from glide_amd.serving import MetRing, ResultWriter, Server, latency_percentiles
    from glide_amd.synthetic import MetSpec, SyntheticMet, serving_requests, workload_config

    spec = MetSpec(nx=240, ny=160, za=60, n_hours=25, seed=20240105, lon0=2.0, lat0=72.0)
    met = SyntheticMet(spec)
    ring = MetRing(met, first_hour=0, n_hours=240, met_dtype=L.GLIDE_MET_BF16, source_hour=lambda h: h % 24)
    base = workload_config("plain-gpu-1M-24h")
    reqs = serving_requests(spec, rate_per_s=20.0, seconds=args.serve_seconds)
    out_dir = args.serve_out or tempfile.mkdtemp(prefix="glide_serve_")
    writer = ResultWriter(out_dir, workers=4)
    srv = Server(ring, base["grid"], base["domain"], max_slots=160, pool_particles=6 << 20, n_t_max=240,
                 seed=17, writer=writer)
    # warm-up: one request through, not timed
    warm = serving_requests(spec, rate_per_s=20.0, seconds=0.2, seed=99)[:1]
    for r in warm:
        srv.admit(r)
    srv.drain()
    stats = srv.serve(reqs)
    res = stats.results
    lat = latency_percentiles(res)
    day = [r for r in res if 18 * 3600 <= r.req.length_s <= 30 * 3600]
    lat_day = latency_percentiles(day)
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
        "met_storage": "bf16 ring, 240 brackets",
        "latency_to": "footprint Zarr store (zlib) on disk",
        "zarr_stores": len(writer.paths),
        "zarr_bytes": sum(f.stat().st_size for p in writer.paths for f in Path(p).rglob("*") if f.is_file()),
        "slots": srv.max_slots,
        "vs_plain": value / plain_value if plain_value > 0 else None,
        "mean_particles": float(np.mean([r.n_particles for r in reqs])) if reqs else 0.0,
    }
    writer.close()
    if not args.serve_out:
        shutil.rmtree(out_dir, ignore_errors=True)
    del srv, ring
    torch.cuda.empty_cache()
    return out


def batched_leg(args, sptr: int, plain_value: float) -> dict:
    """One GPU's share of the network batch: 336 releases (7 sites x 48 hourly
    releases, 2k-20k particles, 24-120 h back) in one engine, with the
    precision layout of the batched method (bfloat16 met records; float32
    footprint from exact fixed-point sums, or with Neumaier compensation under
    the sort accumulation; float64 release totals) and the live set compacted
    every met hour. Times the whole schedule."""

    import numpy as np
    import torch

    from glide_amd import _lib as L
    from glide_amd.runner import BatchEngine
    from glide_amd.synthetic import SyntheticMet, workload_config

    from glide_amd.autotune import tune

    cfg = workload_config(BATCHED)
    met = SyntheticMet(cfg["spec"])
    kw = dict(met_dtype=L.GLIDE_MET_BF16, compensated=True, compact=True, overlap=not args.no_overlap,
              accumulate=args.accumulate)
    bpp = BYTES_PER_PARTICLE_STEP[args.accumulate]
    if args.batched_block or args.batched_reorder >= 0:
        best = {}
        if args.batched_block:
            best["step_block"] = args.batched_block
        if args.batched_reorder >= 0:
            best["reorder_every"] = args.batched_reorder
        tuned = {"best": best, "trials": []}
    elif args.no_autotune:
        tuned = {"best": {}, "trials": []}
    else:
        tuned = tune(cfg, met, use_cache=False, **kw)
    eng = BatchEngine(cfg, device="cuda", met=met, **kw, **tuned["best"])
    n_sched = len(eng.sched)
    K = n_sched if args.batched_steps <= 0 else min(args.batched_steps, n_sched)
    eng.run(min(20, n_sched))  # warm-up, then back to the release state
    torch.cuda.synchronize()
    eng.reset()
    torch.cuda.synchronize()
    ev = eng.time_step_kernel(K, torch.cuda.current_stream(), every=max(1, args.time_every))
    t0 = time.perf_counter()
    for _ in range(K):
        eng.run_step(sptr)
    eng.join(sptr)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    act = eng.active_steps.cpu().numpy()[:K].astype(np.float64)
    psteps = int(act.sum())
    g = cfg["grid"]
    value = psteps / elapsed
    timed = [(i, a.elapsed_time(b)) for i, (a, b) in sorted(ev.items()) if act[i] > 0]
    t_ms = np.array([t for _, t in timed])
    moved = float(sum(act[i] for i, _ in timed)) * bpp
    achieved = moved / (float(t_ms.sum()) * 1e-3) / 1e9 if len(t_ms) else float("nan")
    out = {
        "workload": BATCHED,
        "value": value,
        "unit": "particle-steps/s",
        "ms_per_step": elapsed * 1e3 / K,
        "steps": K,
        "steps_in_run": n_sched,
        "active_particle_steps": psteps,
        "particles": eng.n,
        "releases": len(cfg["releases"]),
        "met_storage": "bf16",
        "footprint": (f"{g['n_r']}x{g['n_t']}x1x{g['n_y']}x{g['n_x']} f32 "
                      + ("from exact int64 fixed-point sums" if args.accumulate == "fixed"
                         else "+ Neumaier compensation")),
        "release_totals": "f64",
        "vs_plain": value / plain_value if plain_value > 0 else None,
        "autotune": {"chosen": tuned["best"], "trials": tuned["trials"]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 5),
                     "kernel": f"glide::step_kernel<true, 0, {str(args.accumulate == 'fixed').lower()}>",
                     "kernel_launches_timed": len(t_ms),
                     "kernel_avg_us": round(float(t_ms.mean()) * 1e3, 3) if len(t_ms) else None,
                     "bytes_per_particle_step": bpp,
                     "whole_step_achieved": round(psteps * bpp / elapsed / 1e9, 2)},
        "error_bar": ("footprint within 1 % relative L2 per release of the float64 oracle on the same noise "
                      "(tests/test_gpu_batch.py, 6-release network batch on the 64x48 grid: measured <= 1.4e-3 "
                      "with bfloat16 met, <= 3e-4 with float32 met; residence totals within 8.5e-4 / 1.9e-4)"),
    }
    eng.close()
    del eng
    torch.cuda.empty_cache()
    return out


def plain_torch_leg(args, sptr: int) -> dict:
    """The plain FP32 method as PyTorch ops on the same GPU (the reference's
    GPU path restated: grid_sample, searchsorted, randn, the fixed substep
    loop, scatter_add_), on the same workload: the first --torch-steps steps
    of the schedule, after 3 warm-up steps on a separate instance."""

    import torch

    from glide_amd.runner import BatchEngine
    from glide_amd.synthetic import SyntheticMet, workload_config
    from glide_amd.torch_plain import TorchPlain

    cfg = workload_config(WORKLOAD)
    met = SyntheticMet(cfg["spec"])
    eng = BatchEngine(cfg, device="cuda", met=met)
    TorchPlain(eng, seed=1).run(3, sptr)
    torch.cuda.synchronize()
    tp = TorchPlain(eng, seed=2)
    K = max(1, min(args.torch_steps, len(eng.sched)))
    t0 = time.perf_counter()
    tp.run(K, sptr)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    psteps = int(tp.active_steps.item())
    out = {"workload": WORKLOAD, "value": psteps / elapsed, "unit": "particle-steps/s", "steps": K,
           "ms_per_step": elapsed * 1e3 / K, "active_particle_steps": psteps,
           "method": "plain FP32 as PyTorch ops (grid_sample, randn, scatter_add_), eager"}
    del tp, eng
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


def main() -> int:
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(the path would shard by release with no collective in the step)", file=sys.stderr)
        return 2
    if args.copy_calibration_mib:
        return copy_calibration(args.copy_calibration_mib)
    import numpy as np
    import torch

    from glide_amd import _lib as L
    from glide_amd.runner import BatchEngine
    from glide_amd.synthetic import SyntheticMet, workload_config

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        return 2
    torch.cuda.set_device(0)
    cfg = workload_config(WORKLOAD)
    met = SyntheticMet(cfg["spec"])
    met_dtype = L.GLIDE_MET_BF16 if args.met == "bf16" else L.GLIDE_MET_F32
    stream = torch.cuda.Stream()  # default priority; graph capture needs a non-default stream
    sptr = int(stream.cuda_stream)
    torch.cuda.set_stream(stream)
    from glide_amd.autotune import tune

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
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(K):
            eng.run_step(sptr)
        eng.join(sptr)  # the last steps' accumulations are part of the timed work
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    # the other launch mode first (replay vs eager), then the headline run
    other_elapsed = timed_run(not args.graph)
    eng.use_graph = args.graph
    eng.run(W)
    torch.cuda.synchronize()
    eng.reset()
    torch.cuda.synchronize()

    every = max(1, args.time_every)
    ev = eng.time_step_kernel(K, stream, every=every)  # HIP events around sampled step-kernel launches
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(K):
        eng.run_step(sptr)
    eng.join(sptr)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    act = eng.active_steps.cpu().numpy()[:K].astype(np.float64)  # counted by the step kernel
    psteps = int(act.sum())
    timed = [(i, a.elapsed_time(b)) for i, (a, b) in sorted(ev.items()) if act[i] > 0]
    t_ms = np.array([t for _, t in timed])
    t_act = np.array([act[i] for i, _ in timed])
    avg_ms = float(t_ms.mean()) if len(t_ms) else float("nan")
    bpp = BYTES_PER_PARTICLE_STEP[args.accumulate]
    bytes_launch = bpp * float(t_act.mean()) if len(t_act) else 0.0
    achieved = (bpp * float(t_act.sum()) / (float(t_ms.sum()) * 1e-3) / 1e9
                if len(t_ms) else float("nan"))
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 2),
        "peak": PEAK_HBM_GBS,
        "unit": "GB/s",
        "frac": round(achieved / PEAK_HBM_GBS, 5),
        "traffic": None,
        "kernel": (f"glide::step_kernel<{str(met_dtype == L.GLIDE_MET_BF16).lower()}, 0, "
                   f"{str(args.accumulate == 'fixed').lower()}>"),
        "kernel_avg_us": round(avg_ms * 1e3, 3),
        "kernel_launches_timed": len(t_ms),
        "bytes_per_launch": round(bytes_launch),
        "bytes_per_particle_step": bpp,
    }
    traffic_file = ROOT / "profiles" / f"traffic_{WORKLOAD}_{args.met}_{args.accumulate}.json"
    if traffic_file.exists():
        try:
            roofline["traffic"] = json.loads(traffic_file.read_text()).get("bytes_per_launch")
        except (OSError, ValueError):
            roofline["traffic"] = None
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
            "accumulate_on_second_stream": args.accumulate == "sort" and not args.no_overlap,
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
                         **{k: v for k, v in kw.items() if k not in ("accumulate", "overlap")}, **tuned["best"])
        oe.run(W)
        torch.cuda.synchronize()
        oe.reset()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(K):
            oe.run_step(sptr)
        oe.join(sptr)
        torch.cuda.synchronize()
        oel = time.perf_counter() - t0
        out[f"accumulate_{other}"] = {"ms_per_step": oel * 1e3 / K, "value": psteps / oel,
                                      "method": ("keys and values written by the step, stable radix sort by cell, "
                                                 "fixed-order float32 reduction on a second stream" if other == "sort"
                                                 else "exact int64 fixed-point sums inside the step kernel")}
        oe.close()
        del oe
    eng.close()
    del eng
    torch.cuda.empty_cache()
    if not args.no_plain_torch:
        out["plain_torch"] = plain_torch_leg(args, sptr)
        pt = out["plain_torch"]["value"]
        out["vs_plain_torch"] = out["value"] / pt if pt > 0 else None
    if not args.no_batched:
        out["batched"] = batched_leg(args, sptr, out["value"])
        if "plain_torch" in out:
            out["batched"]["vs_plain_torch"] = out["batched"]["value"] / pt if pt > 0 else None
    if not args.no_serving:
        out["serving"] = serving_leg(args, out["value"])
    if not args.no_ingest:
        out["ingest"] = ingest_leg(args, sptr, out["value"])
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(cfg, args.cpu_particles, args.cpu_steps)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
