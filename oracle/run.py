"""Whole-run oracle driver: the reference's static cursor loop for one batch.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows reference main.py:1293-1477 (S2 order): active mask, met bracket and
alpha, the scheme step, convection on a bracket change, time-ago bin from the
pre-step cursor, gridding of post-step positions, escape kill.
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from oracle import gridder as og
from oracle import lpdm
from oracle import resample as ors
from oracle import window as ow


def release_offsets(rel_seed: int, n: int, duration_s: float) -> np.ndarray:
    """Within-release release-time offsets, U[0, duration) from a CPU torch
    generator seeded with the release seed (reference release_generator.py:286-293)."""

    import torch

    g = torch.Generator(device="cpu")
    g.manual_seed(int(rel_seed))
    return torch.empty(n, dtype=torch.float32).uniform_(0.0, float(duration_s), generator=g).numpy()


def build_window(hs, he, levels, lon, lat, dty, meander_coeff=0.16, meander_r=1):
    """Window for the bracket (hs, he) of resampled hours; each h is
    (out3 [5, Za, Y, X], src2 [4, Y, X], p_level_hpa [Za])."""

    s3, s2, p_hpa = hs
    e3, e2, _ = he
    s3 = s3.astype(dty)
    e3 = e3.astype(dty)
    surf_s = np.stack([s2[0], s2[1], s2[2], s2[3], s3[3, 0]]).astype(dty)
    surf_e = np.stack([e2[0], e2[1], e2[2], e2[3], e3[3, 0]]).astype(dty)
    Za, Y, X = s3.shape[1:]
    height = np.broadcast_to(np.asarray(levels, dty)[:, None, None], (Za, Y, X))
    p_pa = np.asarray(p_hpa, dty) * 100.0
    t_mid = ow.window_mid(s3[3], e3[3])
    u_mid = ow.window_mid(s3[0], e3[0])
    v_mid = ow.window_mid(s3[1], e3[1])
    ft = ow.free_trop_fields(t_mid, u_mid, v_mid, height, p_pa).astype(dty)
    dens = ow.density_fields(t_mid, height, p_pa).astype(dty)
    mean = ow.meander_fields(u_mid, v_mid, meander_coeff, meander_r).astype(dty)
    return lpdm.Window(
        uvw_s=s3[0:3], uvw_e=e3[0:3], surf_s=surf_s, surf_e=surf_e, ft=ft, dens=dens,
        meander=mean, lon0=float(lon[0]), lon1=float(lon[-1]), lat0=float(lat[0]),
        lat1=float(lat[-1]), levels=np.asarray(levels, dty),
    )


@dataclass
class RunResult:
    footprint: np.ndarray
    particles: dict
    alive: np.ndarray
    escaped_per_step: list = field(default_factory=list)
    active_per_step: list = field(default_factory=list)
    particle_steps: int = 0
    step_seconds: float = 0.0
    total_seconds: float = 0.0


def schedule(releases, dt: int):
    """Cursor schedule of one batch (main.py:1338-1389): list of
    (cursor_s, delta_s) with cursor in seconds after the met epoch."""

    ends = [r.release_time_s + r.duration_s for r in releases]
    start = max(ends)
    term = min(e - r.length_s for e, r in zip(ends, releases))
    out = []
    t = start
    while t > term:
        d = min(dt, t - term)
        out.append((t, d))
        t -= d
    return out


def run(cfg, dty=np.float64, params: lpdm.HannaParams | None = None, met=None,
        max_steps: int | None = None, hours_cache=None, convection=None,
        release_terminus: bool = False) -> RunResult:
    """Run the plain method over the whole single-batch workload ``cfg``
    (from ``glide_amd.synthetic.workload_config``). ``release_terminus``
    steps each release only while the cursor is above its own terminus
    (window end - length), as when it runs alone; the default is the
    reference's mask (main.py:1396-1401) under the batch terminus."""

    from glide_amd.synthetic import SyntheticMet

    t_all = time.perf_counter()
    p = params or lpdm.HannaParams()
    spec = cfg["spec"]
    met = met or SyntheticMet(spec)
    levels = spec.agl_levels
    rels = cfg["releases"]
    grid, dom = cfg["grid"], cfg["domain"]
    dt = cfg["dt"]
    seed = cfg["seed"]
    batch_start = min(r.release_time_s for r in rels)
    P = {k: [] for k in ("lon", "lat", "z", "up", "vp", "wp", "um", "vm", "uid")}
    w_l, roff_l, weoff_l, ridx_l, lo_l, te_l = [], [], [], [], [], []
    for r in rels:
        n = r.n_particles
        P["lon"].append(np.full(n, r.lon, np.float32))
        P["lat"].append(np.full(n, r.lat, np.float32))
        P["z"].append(np.full(n, r.alt_agl_m, np.float32))
        for k in ("up", "vp", "wp", "um", "vm"):
            P[k].append(np.zeros(n, np.float32))
        # particle ids: (noise id % 4096) << 20 | k (one release of > 2^20: k alone)
        nid = r.release_idx if getattr(r, "noise_id", None) is None else r.noise_id
        base = 0 if n > (1 << 20) else (int(nid) % 4096) << 20
        P["uid"].append((np.uint32(base) + np.arange(n, dtype=np.uint32)).astype(np.uint32))
        w_l.append(np.full(n, 1.0 / n, np.float32))
        off = float(r.release_time_s - batch_start)
        ro = release_offsets(r.seed, n, r.duration_s) + np.float32(off)
        roff_l.append(ro)
        lo_l.append(ro - np.float32(r.length_s))  # each release's own backward length
        te_l.append(np.full(n, np.float32(off + r.duration_s) - np.float32(r.length_s), np.float32))
        weoff_l.append(np.full(n, off + r.duration_s, np.float32))
        ridx_l.append(np.full(n, r.release_idx - rels[0].release_idx, np.int64))  # batch-local (main.py:1299-1302)
    P = {k: np.concatenate(v) for k, v in P.items()}
    for k in ("lon", "lat", "z", "up", "vp", "wp", "um", "vm"):
        P[k] = P[k].astype(dty)
    weights = np.concatenate(w_l)
    roff = np.concatenate(roff_l)
    weoff = np.concatenate(weoff_l)
    ridx = np.concatenate(ridx_l)
    lo = np.concatenate(lo_l)
    te = np.concatenate(te_l)
    N = weights.size
    alive = np.ones(N, bool)
    fp = np.zeros(grid["n_r"] * grid["n_t"] * grid["n_z"] * grid["n_y"] * grid["n_x"], dty)

    hours = hours_cache if hours_cache is not None else {}

    def hour(h):
        if h not in hours:
            sh = met.hour(h)
            hours[h] = ors.resample_source_hour(sh, met, levels)
        return hours[h]

    windows = {}
    res = RunResult(footprint=fp, particles=P, alive=alive)
    sched = schedule(rels, dt)
    if max_steps is not None:
        sched = sched[:max_steps]
    last_bracket = None
    t_step = 0.0
    for si, (cur, delta) in enumerate(sched):
        cursor_off = float(cur - batch_start)
        if release_terminus:
            active = (roff >= np.float32(cursor_off)) & (np.float32(cursor_off) > te) & alive
        else:
            active = (roff >= np.float32(cursor_off)) & (lo <= np.float32(cursor_off)) & alive
        bh = cur // 3600
        if bh not in windows:
            windows.clear()
            windows[bh] = build_window(hour(bh), hour(bh + 1), levels, met.lon, met.lat, dty)
        win = windows[bh]
        alpha = max(0.0, min(1.0, ((cur - 0.5 * delta) - 3600.0 * bh) / 3600.0))
        t0 = time.perf_counter()
        noise_step = cur // dt  # Philox counter word 1: absolute time in steps
        P = lpdm.step(P, active, win, alpha, float(delta), p, seed=seed, step_index=noise_step)
        if bh != last_bracket:
            if convection is not None:
                P = convection(P, active, bh, alpha, noise_step)
            last_bracket = bh
        t_idx = og.time_bins(weoff, cursor_off, grid["n_t"])
        og.accumulate(fp, P["lon"].astype(np.float32), P["lat"].astype(np.float32),
                      P["z"].astype(np.float32), active, weights, t_idx, ridx, float(delta), grid)
        inside = og.within_met_domain(P["lon"], P["lat"], P["z"], dom)
        newly = active & ~inside
        alive &= ~newly
        t_step += time.perf_counter() - t0
        res.escaped_per_step.append(int(newly.sum()))
        res.active_per_step.append(int(active.sum()))
    res.particles = P
    res.alive = alive
    res.particle_steps = int(sum(res.active_per_step))
    res.step_seconds = t_step
    res.total_seconds = time.perf_counter() - t_all
    return res


def column_convection(met, hours: dict, seed: int, levels=None):
    """Per-column convection for ``run(convection=...)``: on the first step of
    each met hour (after the step, before the gridding), every active particle
    is redistributed through its nearest column's matrix with the engine's
    trigger, closure and Philox stream (reference emanuel.py:570-805 applied
    per column). ``hours`` must be the dict ``run`` fills (``hours_cache``)."""

    from oracle import convection as oc

    levels = met.spec.agl_levels if levels is None else levels

    def convect(P, active, bh, alpha, noise_step):
        s3, _, p_hpa = hours[bh]
        e3, _, _ = hours[bh + 1]
        conv = met.hour(bh).conv
        lcl, lnb, mb = oc.column_descriptors(conv, s3[3].astype(np.float32), e3[3].astype(np.float32), alpha,
                                             p_hpa, np.asarray(levels, np.float32))
        ny, nx = lcl.shape
        col = oc.nearest_column(P["lon"].astype(np.float32), P["lat"].astype(np.float32), met.lon[0],
                                met.lon[-1], met.lat[0], met.lat[-1], nx, ny)
        out = dict(P)
        out["z"] = oc.convect_columns(P["z"], active, col, lcl, lnb, mb.astype(np.float64), p_hpa, levels, seed,
                                      P["uid"], noise_step).astype(P["z"].dtype)
        return out

    return convect


def torch_windows(met, levels=None, hours=None):
    """Window provider for ``glide_amd.torch_plain.TorchPlain`` on the CPU:
    bracket hour ``bh``'s fields from this module's float32 window builder
    (resample, support stacks, meander), as CPU torch tensors in the layout
    TorchPlain reads. ``hours`` caches resampled hours across calls."""

    import torch

    levels = met.spec.agl_levels if levels is None else levels
    hours = {} if hours is None else hours

    def hour(h):
        if h not in hours:
            sh = met.hour(h)
            hours[h] = ors.resample_source_hour(sh, met, levels)
        return hours[h]

    def windows(bh: int) -> dict:
        hs, he = hour(bh), hour(bh + 1)
        w = build_window(hs, he, levels, met.lon, met.lat, np.float32)
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32))  # noqa: E731
        return dict(uvw=t(np.concatenate([w.uvw_s, w.uvw_e])), ft=t(w.ft), dens=t(w.dens), mea=t(w.meander),
                    surf_s=t(w.surf_s), surf_e=t(w.surf_e), t_s=t(hs[0][3]), t_e=t(he[0][3]),
                    conv=t(met.hour(bh).conv), p_hpa=torch.from_numpy(np.asarray(hs[2], np.float64)))

    return windows
