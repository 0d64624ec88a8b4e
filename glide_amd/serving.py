"""Serving: footprint requests stepped together by continuous batching.

A request is one point release (site, inlet height, release time, particle
count, release duration, backward length). Each admitted request takes a slot
with its own cursor; every engine step advances all running slots by one step
of their own schedules (``glide_slots_advance``), runs the fused step kernel
over the shared particle pool with each wavefront reading its slot's scalars
and its bracket's window from a device-resident met ring
(``glide_step_slots``), which adds every footprint to the exact fixed-point
sums of the slot's slab (converted to float when the request finishes,
``glide_fixed_finalize_slabs``). With convection on, a slot's first step of
each met hour leaves its emission to a convection launch over that slot's
particles (``glide_convect`` with the slot's scalars), after the step as in
the reference. A request is stepped exactly as it would be alone (its own
terminus, offsets relative to its own release time, noise keyed by its id and
absolute time), so its result does not depend on what else is running.

Admission is by free HBM: a request needs a free slot, a run of free
64-particle blocks in the particle pool and a footprint slab (its own hourly
time-ago bins) in the fixed-point arena; when the pool or the arena has no
room, they grow by as much as the device's free memory allows, and a request
that still does not fit waits. Finished slots are found by an asynchronous
copy of the slot table; their footprints are copied to pinned host memory on
the same stream, so the slot and its slab can be reused at once.

No counterpart in the reference, which runs one batch on one cursor
(main.py:1383-1477); the per-request schedule is the reference's schedule for
a batch holding that release alone (main.py:1338-1389, 892-938).
"""

from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from glide_amd import _lib as L
from glide_amd.runner import ConvectionOptions, DeviceMet, HannaOptions, fixed_exponent
from glide_amd.schedule import release_offsets
from glide_amd.synthetic import SyntheticMet

BLOCK = 64  # particles per wavefront; a slot owns whole blocks


def served_noise_hi(req_id: int, n_particles: int) -> int:
    """Philox counter word 3 of a served request's particles. A request of at
    most 2^20 particles has ids (req_id % 4096) << 20 | k, so word 3 carries
    the rest of its id (req_id >> 12); a larger request has ids k and word 3
    is req_id with the top bit set. Two requests served together therefore
    never share a (particle id, word 3) pair, and a request with req_id <
    4096 draws the noise of the same release in a batch."""

    if req_id < 0 or req_id >= (1 << 31):
        raise ValueError("req_id must be in [0, 2^31)")
    if n_particles <= (1 << 20):
        return int(req_id) >> 12
    return (1 << 31) | int(req_id)


@dataclass
class Request:
    """One footprint request (a point release)."""

    req_id: int
    lon: float
    lat: float
    alt_agl_m: float
    n_particles: int
    release_time_s: int  # seconds after the met epoch
    length_s: int  # backward length
    duration_s: int = 60  # release window
    seed: int = 0
    arrival_s: float = 0.0  # wall-clock offset of the arrival (request loop)


@dataclass
class Result:
    req: Request
    footprint: np.ndarray  # [n_t_used, n_z, n_y, n_x] float32 (the slab's first bins)
    total: float  # residence total (float64 segmented sums)
    escaped: int
    active_steps: int
    steps: int
    latency_s: float = float("nan")
    state: dict | None = None  # final particle state (Server(keep_state=True))


class MetRing:
    """Device-resident windows of consecutive met brackets.

    Bracket h (ring index h - first_hour) holds the packed window of hours h
    and h + 1 as ``glide_build_window`` writes it. ``source_hour`` maps a ring
    hour to the hour of the source met (identity, or a repeated cycle). With
    ``convection`` the ring also keeps, per source hour, what the convection
    trigger and closure read: the resampled temperature [nz][ny][nx], the
    level pressures and the column descriptors (CAPE, LCL and cloud-top
    heights, T_v excess). Source hours are uploaded and resampled
    ``upload_batch`` at a time, so the build holds only those (and the ring)."""

    def __init__(self, met: SyntheticMet, first_hour: int, n_hours: int, *, met_dtype: int = L.GLIDE_MET_BF16,
                 source_hour=None, device="cuda", meander_coeff: float = 0.16, meander_radius: int = 1,
                 stream: int | None = None, convection: bool = False, upload_batch: int = 8):
        self.first_hour = int(first_hour)
        self.n_hours = int(n_hours)
        self.met_dtype = met_dtype
        self.source_hour = source_hour or (lambda h: h)
        self.convection = bool(convection)
        dm = DeviceMet(met, [], device=device, met_dtype=met_dtype, meander_coeff=meander_coeff,
                       meander_radius=meander_radius, defer=True)
        self.dm = dm
        self.nx, self.ny, self.nz = dm.nx, dm.ny, dm.nz
        ncell = self.nz * self.ny * self.nx
        self.wind_stride = L.window_bytes(ncell, met_dtype)
        self.aux_stride = L.window_bytes(ncell, met_dtype)
        self.surf_stride = self.ny * self.nx * 12 * 4
        self.wind = torch.empty(n_hours * self.wind_stride, dtype=torch.uint8, device=device)
        self.aux = torch.empty(n_hours * self.aux_stride, dtype=torch.uint8, device=device)
        self.surf = torch.empty(n_hours * self.surf_stride, dtype=torch.uint8, device=device)
        # convection inputs per source hour: temperature, pressures, descriptors
        self.conv_t: dict[int, torch.Tensor] = {}
        self.conv_p: dict[int, torch.Tensor] = {}
        self.conv_d: dict[int, torch.Tensor] = {}
        st = L.stream_ptr() if stream is None else stream
        lib = L.lib()
        step = max(1, int(upload_batch))
        for j0 in range(0, n_hours, step):
            js = range(j0, min(n_hours, j0 + step))
            need = sorted({self.source_hour(first_hour + j + e) for j in js for e in (0, 1)})
            for h in [h for h in list(dm.src) if h not in need]:  # keep only this batch's hours
                dm.src.pop(h)
                dm.ring.pop(h, None)
            dm.upload(need)
            for j in js:
                h = first_hour + j
                s3, s2, p = dm.resampled(self.source_hour(h), st)
                e3, e2, _ = dm.resampled(self.source_hour(h + 1), st)
                L.check(lib.glide_build_window(
                    s3.data_ptr(), e3.data_ptr(), s2.data_ptr(), e2.data_ptr(), p.data_ptr(),
                    dm.levels_f.data_ptr(), self.nz, self.ny, self.nx, meander_coeff, meander_radius, met_dtype,
                    self.wind.data_ptr() + j * self.wind_stride, self.aux.data_ptr() + j * self.aux_stride,
                    self.surf.data_ptr() + j * self.surf_stride, st))
            if self.convection:
                for sh in need:
                    if sh not in self.conv_t:
                        r3, _, rp = dm.resampled(sh, st)
                        self.conv_t[sh] = r3[3].clone()
                        self.conv_p[sh] = rp.clone()
                        self.conv_d[sh] = dm.source(sh, st)["conv"].clone()
        torch.cuda.synchronize()
        dm.ring.clear()
        dm.src.clear()  # the ring holds everything the steps read
        self.grid = dm.grid
        self.levels_f = dm.levels_f
        self.desc = L.GlideMetRing(
            wind=self.wind.data_ptr(), aux=self.aux.data_ptr(), surf=self.surf.data_ptr(),
            wind_stride=self.wind_stride, aux_stride=self.aux_stride, surf_stride=self.surf_stride,
            first_hour=self.first_hour, n_hours=self.n_hours, dtype=met_dtype, reserved=0)

    @property
    def nbytes(self) -> int:
        """Device bytes the ring holds."""

        extra = sum(t.numel() * t.element_size() for d in (self.conv_t, self.conv_p, self.conv_d)
                    for t in d.values())
        return self.wind.numel() + self.aux.numel() + self.surf.numel() + extra

    def covers(self, t0_s: float, t1_s: float) -> bool:
        """True when every bracket of cursors in (t0, t1] is in the ring."""
        return (math.floor(t0_s / 3600.0) >= self.first_hour
                and math.floor(t1_s / 3600.0) < self.first_hour + self.n_hours)


@dataclass
class _Running:
    req: Request
    slot: int
    block0: int
    n_blocks: int
    admitted_s: float
    cell0: int  # the slab's first cell in the footprint arena
    n_t: int  # hourly time-ago bins of the slab
    # the slot's schedule as glide_slots_advance steps it (replayed on the host)
    t: float = 0.0
    term: float = 0.0
    noise_hi: int = 0
    k: int = 0
    last_d: float = 0.0
    bracket: int | None = None
    done: bool = False


@dataclass
class _Pending:
    res: Result
    event: torch.cuda.Event
    host: torch.Tensor
    total: torch.Tensor
    n_t: int


class _Staging:
    """Pinned host buffers for host-to-device copies that must not stall the
    stream: a ring of buffers, each reused after its last copy completed."""

    def __init__(self, nbytes: int, count: int = 4):
        self.bufs = [torch.empty(nbytes, dtype=torch.uint8).pin_memory() for _ in range(count)]
        self.events = [None] * count
        self.i = 0

    def take(self) -> tuple[int, torch.Tensor]:
        j = self.i
        self.i = (self.i + 1) % len(self.bufs)
        if self.events[j] is not None:
            self.events[j].synchronize()
        return j, self.bufs[j]

    def done(self, j: int) -> None:
        ev = torch.cuda.Event()
        ev.record()
        self.events[j] = ev


@dataclass
class ServeStats:
    results: list = field(default_factory=list)
    particle_steps: int = 0
    steps: int = 0
    busy_s: float = 0.0


class ResultWriter:
    """Writes finished requests' footprints to Zarr stores on host threads, off
    the serving loop (one zlib-deflated one-release store per request; the
    deflate releases the interpreter lock). A result is handed back, with its
    latency measured to the moment its store is on disk, once written; a write
    error re-raises in the serving loop."""

    def __init__(self, out_dir, *, workers: int = 4, compressor: str | None = "zlib"):
        from concurrent.futures import ThreadPoolExecutor

        self.out_dir = out_dir
        self.compressor = compressor
        self.pool = ThreadPoolExecutor(max_workers=int(workers), thread_name_prefix="glide-zarr")
        self.inflight: list = []
        self.bytes_written = 0
        self.paths: list[str] = []

    def submit(self, server: "Server", res: Result) -> None:
        self.inflight.append((res, self.pool.submit(server.write_result, res, self.out_dir,
                                                    compressor=self.compressor)))

    def take_done(self, now_fn, block: bool = False) -> list[Result]:
        done, keep = [], []
        for res, fut in self.inflight:
            if block or fut.done():
                path = fut.result()  # re-raises a write error here
                res.latency_s = now_fn() - res.req.arrival_s
                self.paths.append(path)
                done.append(res)
            else:
                keep.append((res, fut))
        self.inflight = keep
        return done

    @property
    def busy(self) -> bool:
        return bool(self.inflight)

    def close(self) -> None:
        self.pool.shutdown(wait=True)


class _Ranges:
    """First-fit allocator of [start, start + count) ranges of a growable span."""

    def __init__(self, size: int):
        self.size = int(size)
        self.free = [(0, self.size)] if size > 0 else []

    def take(self, n: int) -> int | None:
        for j, (b0, cnt) in enumerate(self.free):
            if cnt >= n:
                if cnt == n:
                    self.free.pop(j)
                else:
                    self.free[j] = (b0 + n, cnt - n)
                return b0
        return None

    def give(self, b0: int, n: int) -> None:
        merged = []
        for st, c in sorted(self.free + [(b0, n)]):
            if merged and merged[-1][0] + merged[-1][1] == st:
                merged[-1] = (merged[-1][0], merged[-1][1] + c)
            else:
                merged.append((st, c))
        self.free = merged

    def grow(self, new_size: int) -> None:
        if new_size > self.size:
            self.give(self.size, new_size - self.size)
            self.size = int(new_size)

    def tail_free(self) -> int:
        """Free units at the end of the span (a growth extends them)."""

        if self.free and self.free[-1][0] + self.free[-1][1] == self.size:
            return self.free[-1][1]
        return 0

    def high_water(self) -> int:
        """One past the last unit in use."""

        return self.size - self.tail_free()


class Server:
    """Continuous-batching footprint server on one GPU. All of its device work
    goes to the stream that is current when it is constructed; use it under
    that same stream.

    Capacity follows free HBM: the particle pool (64-particle blocks) and the
    footprint arena (one int64 fixed-point slab per running request, as many
    hourly time-ago bins as its backward length needs, placed by a 64-bit
    offset) start at ``pool_particles`` and ``arena_cells`` and grow on
    admission when a request does not fit, by as much as the device's free
    memory (torch's free cached blocks included) allows beyond
    ``hbm_reserve_bytes`` and, when set, within ``hbm_limit_bytes`` for the
    two together. A request that cannot be placed waits. ``max_slots`` bounds
    the requests running at once (the slot table)."""

    PARTICLE_ARRAYS = ("lon", "lat", "z", "up", "vp", "wp", "um", "vm", "weight", "r_off", "we_off", "rel_idx",
                       "uid", "alive", "lo_off")

    def __init__(self, ring: MetRing, grid: dict, domain: dict, *, n_t_max: int, seed: int, dt: int = 60,
                 hanna: HannaOptions | None = None, convection: ConvectionOptions | None = None,
                 max_slots: int = 1024, pool_particles: int = 1 << 16, arena_cells: int | None = None,
                 hbm_reserve_bytes: int = 2 << 30, hbm_limit_bytes: int | None = None, device="cuda",
                 stream: int | None = None, poll_every: int = 8, pinned_slabs: int = 8, keep_state: bool = False,
                 writer: ResultWriter | None = None):
        self.ring = ring
        if convection is not None and not ring.convection:
            raise ValueError("convection needs a MetRing built with convection=True")
        self.conv = convection
        self.grid = dict(grid)
        self.writer = writer  # finished footprints go to Zarr stores before they count as done
        self.dt = int(dt)
        self.seed = int(seed)
        self.h = hanna or HannaOptions()
        self.device = device
        self.stream = L.stream_ptr() if stream is None else stream
        self.max_slots = int(max_slots)
        self.n_t_max = int(n_t_max)
        self.poll_every = int(poll_every)
        self.keep_state = bool(keep_state)
        self.hbm_reserve = int(hbm_reserve_bytes)
        self.hbm_limit = None if hbm_limit_bytes is None else int(hbm_limit_bytes)
        d = device
        self.z_edges = torch.tensor(grid["z_edges"], dtype=torch.float32, device=d)
        self.n_z = len(grid["z_edges"]) - 1
        self.cpb = self.n_z * grid["n_y"] * grid["n_x"]  # cells per hourly bin
        # per-slot slab placement in the arena, scale exponent, totals
        self.slab_off = torch.zeros(self.max_slots, dtype=torch.int64, device=d)
        self.slab_nt = torch.zeros(self.max_slots, dtype=torch.int32, device=d)
        self.fixed_exp = torch.zeros(self.max_slots, dtype=torch.int32, device=d)
        self.release_total = torch.zeros(self.max_slots, dtype=torch.float64, device=d)
        self.fixed_ws = torch.empty(max(8, L.lib().glide_fixed_workspace_bytes(self.max_slots)), dtype=torch.uint8,
                                    device=d)
        self.fp_scratch = torch.empty(self.n_t_max * self.cpb, dtype=torch.float32, device=d)
        self.fg = L.GlideFootprintGrid(
            lon_min=grid["lon_min"], lon_max=grid["lon_max"], lat_min=grid["lat_min"], lat_max=grid["lat_max"],
            z_edges=self.z_edges.data_ptr(), n_r=self.max_slots, n_t=self.n_t_max, n_z=self.n_z,
            n_y=grid["n_y"], n_x=grid["n_x"], reserved=0,
            dom_lon_min=domain["lon_min"], dom_lon_max=domain["lon_max"], dom_lat_min=domain["lat_min"],
            dom_lat_max=domain["lat_max"], dom_alt_max=domain["alt_max"], t_bin=None,
            slab_off=self.slab_off.data_ptr(), slab_nt=self.slab_nt.data_ptr(), slab_cells=0)
        # growable particle pool (structure of arrays) and footprint arena
        self.P = {}
        self.n_blocks = 0
        self.wave_slot = torch.empty(0, dtype=torch.int32, device=d)
        self.wave_slot_h = np.zeros(0, np.int32)
        self.blocks = _Ranges(0)
        self.fixed = torch.empty(0, dtype=torch.int64, device=d)
        self.arena = _Ranges(0)
        self._grow_pool((int(pool_particles) + BLOCK - 1) // BLOCK)
        self._grow_arena(int(arena_cells) if arena_cells is not None else 8 * self.n_t_max * self.cpb)
        # slot table (device) and its pinned host snapshot
        self.slot_bytes = C.sizeof(L.GlideSlot)
        self.escaped_off = L.GlideSlot.escaped.offset
        self.slots = torch.zeros(self.max_slots * self.slot_bytes, dtype=torch.uint8, device=d)
        self.slots_host = torch.zeros_like(self.slots, device="cpu").pin_memory()
        self.slots_event = None
        self.free_slots = list(range(self.max_slots))[::-1]
        self.running: dict[int, _Running] = {}
        self.pending: list[_Pending] = []
        self.pinned = [torch.empty(self.n_t_max * self.cpb, dtype=torch.float32).pin_memory()
                       for _ in range(pinned_slabs)]
        self.dummy_escaped = torch.zeros(1, dtype=torch.int64, device=d)  # per-slot counters are in the slots
        self.staging = None
        self.tmpl = L.GlideStepParams(
            dt=float(self.dt), alpha=0.0, cursor_off=0.0, length_s=0.0, substep_c=self.h.substep_c,
            z_ubl_m=self.h.z_ubl_m, meander_timescale_s=self.h.meander_timescale_s, seed=self.seed,
            step_index=0, max_substeps=self.h.max_substeps,
            flags=self.h.flags() | L.F_EMIT | L.F_RELEASE_TERMINUS | L.F_FIXED)
        # convection: descriptors per (bracket hour, alpha), a few kept
        self._desc: dict[tuple, tuple] = {}
        if self.conv is not None:
            ncol = ring.ny * ring.nx
            self.conv_cloud = torch.empty(4 * ncol, dtype=torch.float32, device=d)
        self.convect_launches = 0
        self.step_count = 0
        self.particle_steps = 0
        self.completed: list[Result] = []  # finished requests not yet handed out
        self.inflight: list[torch.cuda.Event] = []  # one event per poll_every steps (backpressure)
        self.max_inflight = 4
        self.t_origin = time.perf_counter()

    # -- capacity ------------------------------------------------------------
    @property
    def hi_blocks(self) -> int:
        return self.blocks.high_water()

    @property
    def cap(self) -> int:
        return self.n_blocks * BLOCK

    def bytes_per_particle(self) -> int:
        return sum(4 if k not in ("alive",) else 1 for k in self.PARTICLE_ARRAYS) + 4 * 1 / BLOCK

    def device_bytes(self) -> int:
        """Bytes of the pool and the arena (what admission grows)."""

        pool = sum(t.numel() * t.element_size() for t in self.P.values()) + self.wave_slot.numel() * 4
        return pool + self.fixed.numel() * 8

    def hbm_available(self) -> int:
        """Bytes the server may still add to its pool and arena."""

        free, _ = torch.cuda.mem_get_info()
        cached = torch.cuda.memory_reserved() - torch.cuda.memory_allocated()
        avail = free + cached - self.hbm_reserve
        if self.hbm_limit is not None:
            avail = min(avail, self.hbm_limit - self.device_bytes())
        return int(avail)

    def _grow_pool(self, min_blocks: int) -> None:
        if min_blocks <= self.n_blocks:
            return
        old = self.n_blocks * BLOCK
        n = min_blocks * BLOCK
        types = dict(rel_idx=torch.int32, uid=torch.int32, alive=torch.uint8)
        newP = {}
        for k in self.PARTICLE_ARRAYS:
            t = torch.zeros(n, dtype=types.get(k, torch.float32), device=self.device)
            if old:
                t[:old].copy_(self.P[k][:old])
            newP[k] = t
        self.P = newP
        ws = torch.full((min_blocks,), -1, dtype=torch.int32, device=self.device)
        if self.n_blocks:
            ws[:self.n_blocks].copy_(self.wave_slot)
        self.wave_slot = ws
        self.wave_slot_h = np.concatenate([self.wave_slot_h, np.full(min_blocks - self.n_blocks, -1, np.int32)])
        self.blocks.grow(min_blocks)
        self.n_blocks = min_blocks

    def _grow_arena(self, min_cells: int) -> None:
        if min_cells <= self.arena.size:
            return
        a = torch.zeros(min_cells, dtype=torch.int64, device=self.device)
        if self.arena.size:
            a[:self.arena.size].copy_(self.fixed)
        self.fixed = a
        self.arena.grow(min_cells)
        self.fg.slab_cells = int(min_cells)

    def _fit(self, nb: int, cells: int, grow: bool) -> bool:
        """Whether a request of nb blocks and a slab of ``cells`` can be placed
        now, growing the pool and the arena within free HBM when ``grow``."""

        need_b = 0 if any(c >= nb for _, c in self.blocks.free) else nb - self.blocks.tail_free()
        need_c = 0 if any(c >= cells for _, c in self.arena.free) else cells - self.arena.tail_free()
        if need_b == 0 and need_c == 0:
            return True
        if not grow:
            return False
        per_block = BLOCK * self.bytes_per_particle()
        avail = self.hbm_available()
        # a growth reallocates: the old arrays are alive while they are copied
        cost = lambda b, c: (self.n_blocks + b) * per_block + (self.arena.size + c) * 8  # noqa: E731
        if cost(need_b, need_c) > avail:
            return False
        # grow by a margin (doubling where free HBM allows) so growths stay rare
        gb = max(need_b, self.n_blocks) if need_b else 0
        gc = max(need_c, self.arena.size) if need_c else 0
        if cost(gb, gc) > avail:
            gb, gc = need_b, need_c
        self._grow_pool(self.n_blocks + gb)
        self._grow_arena(self.arena.size + gc)
        return True

    def _slab_bins(self, req: Request) -> int:
        return int(math.ceil(req.length_s / 3600.0))

    def can_admit(self, req: Request, grow: bool = False) -> bool:
        end = req.release_time_s + req.duration_s
        if not self.ring.covers(end - req.length_s, end):
            raise ValueError(f"request {req.req_id}: met ring does not cover its window")
        if self._slab_bins(req) > self.n_t_max:
            raise ValueError(f"request {req.req_id}: backward length exceeds {self.n_t_max} h")
        if not self.free_slots:
            return False
        nb = (req.n_particles + BLOCK - 1) // BLOCK
        return self._fit(nb, self._slab_bins(req) * self.cpb, grow)

    def admit(self, req: Request, now_s: float | None = None) -> bool:
        """Place a request on the device if a slot is free and its particles
        and footprint slab fit (the pool and arena grow within free HBM).
        Its arrays go up through pinned staging, so admission does not wait
        for the steps already queued."""

        if not self.can_admit(req, grow=True):
            return False
        n = int(req.n_particles)
        nb = (n + BLOCK - 1) // BLOCK
        b0 = self.blocks.take(nb)
        n_t = self._slab_bins(req)
        c0 = self.arena.take(n_t * self.cpb)
        slot = self.free_slots.pop()
        i0, m = b0 * BLOCK, nb * BLOCK
        need = m * 4 * 11 + self.slot_bytes + nb * 4 + 512
        if self.staging is None or need > self.staging.bufs[0].numel():
            self.staging = _Staging(max(need, 4 << 20))
        j, buf = self.staging.take()
        off = 0

        def put(a: np.ndarray) -> torch.Tensor:
            nonlocal off
            raw = np.ascontiguousarray(a).view(np.uint8).reshape(-1)
            view = buf[off:off + raw.size]
            view.numpy()[:] = raw
            off += (raw.size + 15) & ~15
            return view

        # particles: offsets relative to the request's own release time
        ro = np.zeros(m, np.float32)
        ro[:n] = release_offsets(req.seed, n, req.duration_s)
        te = np.float32(req.duration_s) - np.float32(req.length_s)
        uid = np.zeros(m, np.uint32)
        uid[:n] = ((np.uint32(req.req_id % 4096) << np.uint32(20)) + np.arange(n, dtype=np.uint32)) \
            if n <= (1 << 20) else np.arange(n, dtype=np.uint32)
        noise_hi = served_noise_hi(req.req_id, n)
        alive = np.zeros(m, np.uint8)
        alive[:n] = 1
        host = dict(lon=np.full(m, req.lon, np.float32), lat=np.full(m, req.lat, np.float32),
                    z=np.full(m, req.alt_agl_m, np.float32), weight=np.full(m, 1.0 / n, np.float32),
                    r_off=ro, we_off=np.full(m, req.duration_s, np.float32), lo_off=np.full(m, te, np.float32),
                    rel_idx=np.full(m, slot, np.int32), uid=uid.view(np.int32), alive=alive)
        for k, v in host.items():
            self.P[k][i0:i0 + m].view(torch.uint8).copy_(put(v), non_blocking=True)
        for k in ("up", "vp", "wp", "um", "vm"):
            self.P[k][i0:i0 + m].zero_()
        # slot entry, slab placement and scale
        end = float(req.release_time_s + req.duration_s)
        sl = L.GlideSlot(t=end, term=end - float(req.length_s), rel_time=float(req.release_time_s),
                         state=L.GLIDE_SLOT_RUNNING, k=0, noise_hi=noise_hi)
        raw = np.frombuffer(bytes(sl), dtype=np.uint8)
        self.slots[slot * self.slot_bytes:(slot + 1) * self.slot_bytes].copy_(put(raw), non_blocking=True)
        self.wave_slot_h[b0:b0 + nb] = slot
        self.wave_slot[b0:b0 + nb].view(torch.uint8).copy_(put(self.wave_slot_h[b0:b0 + nb]), non_blocking=True)
        self.slab_off[slot:slot + 1].view(torch.uint8).copy_(put(np.array([c0], np.int64)), non_blocking=True)
        self.slab_nt[slot:slot + 1].view(torch.uint8).copy_(put(np.array([n_t], np.int32)), non_blocking=True)
        # the slot's scale: n particles of weight float32(1/n), whose sum n * w is exact in float64
        e = fixed_exponent(n * float(np.float32(1.0 / n)), req.length_s, self.dt).reshape(1)
        self.fixed_exp[slot:slot + 1].view(torch.uint8).copy_(put(e), non_blocking=True)
        self.staging.done(j)
        self.running[slot] = _Running(req, slot, b0, nb, now_s if now_s is not None else self.now(), c0, n_t,
                                      t=end, term=end - float(req.length_s), noise_hi=noise_hi)
        return True

    # -- stepping ----------------------------------------------------------
    def now(self) -> float:
        return time.perf_counter() - self.t_origin

    def _advance_host(self) -> list:
        """The slot schedule glide_slots_advance applies on the device, replayed
        on the host: the slots whose step is the first of a met hour for them,
        with that step's scalars (the reference's cursor loop for the release
        alone, main.py:1338-1389, 892-938)."""

        out = []
        for run in self.running.values():
            if run.done:
                continue
            if run.k > 0:
                run.t -= run.last_d
            if not run.t > run.term:
                run.done = True
                continue
            d = min(float(self.dt), run.t - run.term)
            bh = math.floor(run.t / 3600.0)
            alpha = max(0.0, min(1.0, ((run.t - 0.5 * d) - 3600.0 * bh) / 3600.0))
            if run.k == 0 or bh != run.bracket:
                out.append((run, bh, alpha, d))
            run.last_d = d
            run.bracket = bh
            run.k += 1
        return out

    def _descriptors(self, bh: int, alpha: float):
        """(i_lcl, i_lnb, m_b, layer_mass) of every column for bracket ``bh`` at
        ``alpha`` (glide_convect_columns on the ring's convection inputs)."""

        key = (bh, float(alpha))
        hit = self._desc.get(key)
        if hit is not None:
            return hit
        r = self.ring
        ncol = r.ny * r.nx
        hs, he = r.source_hour(bh), r.source_hour(bh + 1)
        d = self.device
        out = (torch.empty(ncol, dtype=torch.int32, device=d), torch.empty(ncol, dtype=torch.int32, device=d),
               torch.empty(ncol, dtype=torch.float32, device=d), torch.empty(r.nz, dtype=torch.float32, device=d))
        co = self.conv
        L.check(L.lib().glide_convect_columns(
            r.conv_d[hs].data_ptr(), r.conv_t[hs].data_ptr(), r.conv_t[he].data_ptr(), float(alpha),
            r.conv_p[hs].data_ptr(), r.levels_f.data_ptr(), r.nz, r.ny, r.nx, co.closure_c, co.trigger_dtv_k,
            co.min_cape_j_kg, co.min_cloud_depth_m, out[0].data_ptr(), out[1].data_ptr(), out[2].data_ptr(),
            out[3].data_ptr(), self.stream))
        if len(self._desc) >= 64:
            self._desc.pop(next(iter(self._desc)))
        self._desc[key] = out
        return out

    def _convect(self, run: "_Running", bh: int, alpha: float, d: float) -> None:
        """Convect one slot's particles after its step, then emit them (the
        slot's step left its emission and escape kill to this call)."""

        i_lcl, i_lnb, m_b, lm = self._descriptors(bh, alpha)
        i0, m = run.block0 * BLOCK, run.n_blocks * BLOCK
        P = self.P
        ptr = lambda k, sz=4: P[k].data_ptr() + i0 * sz  # noqa: E731
        parts = L.GlideParticles(
            n=m, pos_stride=1, lon=ptr("lon"), lat=ptr("lat"), z=ptr("z"), up=ptr("up"), vp=ptr("vp"),
            wp=ptr("wp"), um=ptr("um"), vm=ptr("vm"), weight=ptr("weight"), r_off=ptr("r_off"),
            we_off=ptr("we_off"), rel_idx=ptr("rel_idx"), uid=ptr("uid"), alive=ptr("alive", 1), active=None,
            lo_off=ptr("lo_off"))
        p = L.GlideStepParams(dt=d, alpha=alpha, cursor_off=run.t - run.req.release_time_s, length_s=0.0,
                              substep_c=self.h.substep_c, z_ubl_m=self.h.z_ubl_m,
                              meander_timescale_s=self.h.meander_timescale_s, seed=self.seed,
                              step_index=int(math.floor(run.t / self.dt)), max_substeps=self.h.max_substeps,
                              flags=L.F_EMIT | L.F_FIXED | L.F_RELEASE_TERMINUS, block=0, noise_hi=run.noise_hi)
        em = L.GlideEmit(keys=None, vals=None, escaped=self.slots.data_ptr() + run.slot * self.slot_bytes
                         + self.escaped_off, active=None, fixed=self.fixed.data_ptr(),
                         fixed_exp=self.fixed_exp.data_ptr(), fixed_copies=1)
        L.check(L.lib().glide_convect(C.byref(p), C.byref(parts), C.byref(self.ring.grid), i_lcl.data_ptr(),
                                      i_lnb.data_ptr(), m_b.data_ptr(), lm.data_ptr(), None,
                                      self.conv_cloud.data_ptr(), C.byref(self.fg), C.byref(em), self.stream))
        self.convect_launches += 1

    def step(self) -> None:
        """One engine step of every running slot."""

        lib = L.lib()
        st = self.stream
        conv = self.conv is not None
        L.check(lib.glide_slots_advance(self.slots.data_ptr(), self.max_slots, float(self.dt),
                                        L.GLIDE_SLOTS_CONVECT if conv else 0, st))
        firsts = self._advance_host()
        n = self.hi_blocks * BLOCK
        if n > 0:
            parts = L.GlideParticles(
                n=n, pos_stride=1, **{k: self.P[k].data_ptr() for k in self.PARTICLE_ARRAYS}, active=None)
            em = L.GlideEmit(keys=None, vals=None, escaped=self.dummy_escaped.data_ptr(), active=None,
                             fixed=self.fixed.data_ptr(), fixed_exp=self.fixed_exp.data_ptr(), fixed_copies=1)
            L.check(lib.glide_step_slots(C.byref(self.tmpl), self.slots.data_ptr(), self.wave_slot.data_ptr(),
                                         C.byref(parts), C.byref(self.ring.grid), C.byref(self.ring.desc),
                                         C.byref(self.fg), C.byref(em), st))
            if conv:
                for run, bh, alpha, d in firsts:
                    self._convect(run, bh, alpha, d)
        self.step_count += 1
        if self.step_count % self.poll_every == 0:
            # keep at most max_inflight * poll_every steps queued ahead of the GPU,
            # so admissions and finish checks see a current device
            ev = torch.cuda.Event()
            ev.record()
            self.inflight.append(ev)
            while len(self.inflight) > self.max_inflight:
                self.inflight.pop(0).synchronize()
            self.poll()

    def _slot_view(self) -> list:
        arr = (L.GlideSlot * self.max_slots).from_buffer_copy(self.slots_host.numpy().tobytes())
        return list(arr)

    def poll(self, block: bool = False) -> list[Result]:
        """Collect finished requests: read the slot table copied at the last
        poll (waiting for it when ``block``), start the footprint copies of
        finished slots, free them, and return results whose copies are done."""

        done = []
        if self.slots_event is not None and (block or self.slots_event.query()):
            if block:
                self.slots_event.synchronize()
            view = self._slot_view()
            for slot, run in list(self.running.items()):
                sv = view[slot]
                if sv.state == L.GLIDE_SLOT_FINISHED:
                    self._finish(slot, run, sv)
            self.slots_event = None
        if self.slots_event is None:
            self.slots_host.copy_(self.slots, non_blocking=True)
            self.slots_event = torch.cuda.Event()
            self.slots_event.record()
        keep = []
        for p in self.pending:
            if block:
                p.event.synchronize()
            if p.event.query():
                fp = p.host[: p.n_t * self.cpb].numpy().copy()
                p.res.footprint = fp.reshape(p.n_t, self.n_z, self.fg.n_y, self.fg.n_x)
                p.res.total = float(p.total[0])
                p.res.latency_s = self.now() - p.res.req.arrival_s
                self.pinned.append(p.host)
                done.append(p.res)
            else:
                keep.append(p)
        self.pending = keep
        if self.writer is not None:
            for res in done:
                self.writer.submit(self, res)
            done = self.writer.take_done(self.now, block=block and not self.pending)
        self.completed.extend(done)
        return done

    def _finish(self, slot: int, run: _Running, sv) -> None:
        req = run.req
        n_t = run.n_t
        cells = n_t * self.cpb
        host = self.pinned.pop() if self.pinned else torch.empty(self.n_t_max * self.cpb,
                                                                 dtype=torch.float32).pin_memory()
        # the slot's exact sums to float and its total, then clear its slab
        L.check(L.lib().glide_fixed_finalize_slabs(
            self.fixed.data_ptr(), 1, self.arena.size, self.slab_off.data_ptr(), self.slab_nt.data_ptr(), self.cpb,
            self.n_t_max, self.fixed_exp.data_ptr(), slot, slot + 1, self.fp_scratch.data_ptr(),
            self.release_total.data_ptr(), self.fixed_ws.data_ptr(), self.fixed_ws.numel(), self.stream))
        host[:cells].copy_(self.fp_scratch[:cells], non_blocking=True)
        total = torch.empty(1, dtype=torch.float64).pin_memory()
        total.copy_(self.release_total[slot:slot + 1], non_blocking=True)
        ev = torch.cuda.Event()
        ev.record()
        # reuse: the copies above precede these writes on the stream
        self.fixed[run.cell0:run.cell0 + cells].zero_()
        self.release_total[slot].zero_()
        self.slots[slot * self.slot_bytes:(slot + 1) * self.slot_bytes].zero_()
        self.wave_slot_h[run.block0:run.block0 + run.n_blocks] = -1
        self.wave_slot[run.block0:run.block0 + run.n_blocks].fill_(-1)
        self.blocks.give(run.block0, run.n_blocks)
        self.arena.give(run.cell0, cells)
        self.free_slots.append(slot)
        del self.running[slot]
        res = Result(req=req, footprint=None, total=float("nan"), escaped=int(sv.escaped),
                     active_steps=int(sv.active_steps), steps=int(sv.k))
        if self.keep_state:  # tests: the request's final particles (a synchronous copy)
            i0 = run.block0 * BLOCK
            res.state = {k: self.P[k][i0:i0 + req.n_particles].cpu().numpy().copy()
                         for k in ("lon", "lat", "z", "up", "vp", "wp", "um", "vm", "uid", "alive")}
        self.particle_steps += int(sv.active_steps)
        self.pending.append(_Pending(res, ev, host, total, n_t))

    def write_result(self, res: Result, out_dir, *, epoch_iso: str = "2024-01-01T00:00:00",
                     compressor: str | None = "zlib") -> str:
        """Write a finished request's footprint as a one-release footprint store
        (time-ago bins of one hour, as many as the request's length needs) and,
        when the server keeps final states, its endpoints as Parquet."""

        from pathlib import Path

        from .output import ZarrFootprintStore, footprint_coords, write_particles_parquet

        out = Path(out_dir) / f"request_{res.req.req_id:06d}"
        fp = res.footprint
        grid = dict(self.grid, n_t=fp.shape[0])
        store = ZarrFootprintStore(out / "footprint.zarr")
        store.create((1,) + fp.shape, coords=footprint_coords(grid, [res.req], length_s=res.req.length_s,
                                                              epoch_iso=epoch_iso),
                     attrs={"request_id": int(res.req.req_id), "residence_total": float(res.total),
                            "escaped": int(res.escaped), "steps": int(res.steps)}, compressor=compressor)
        store.write_region(0, fp[None])
        if res.state is not None:
            S = res.state
            w = np.full(S["lon"].shape, 1.0 / max(1, res.req.n_particles), np.float32)
            write_particles_parquet(out / "endpoints.parquet", S["lon"], S["lat"], S["z"], w)
        return str(out)

    def take_completed(self) -> list[Result]:
        out, self.completed = self.completed, []
        return out

    def drain(self) -> list[Result]:
        """Step until every admitted request has finished; return the results
        finished since the last hand-out."""

        while self.running or self.pending or (self.writer is not None and self.writer.busy):
            if self.running:
                self.step()
            self.poll(block=not self.running)
        return self.take_completed()

    # -- request loop ------------------------------------------------------
    def serve(self, requests: list[Request], max_wall_s: float | None = None, progress_every: int = 0,
              log=None) -> ServeStats:
        """Real-time request loop: requests arrive at ``arrival_s`` (seconds
        after the start), are admitted as slots and pool blocks free up, and
        are stepped together until all are done. With ``progress_every`` > 0,
        ``log`` (default: print to stderr) gets a line each time that many
        more requests have completed."""

        if progress_every and log is None:
            import sys

            def log(msg):
                print(msg, file=sys.stderr, flush=True)

        stats = ServeStats()
        queue = sorted(requests, key=lambda r: r.arrival_s)
        qi = 0
        waiting: list[Request] = []
        self.t_origin = time.perf_counter()
        t_busy0 = None
        while qi < len(queue) or waiting or self.running or self.pending or (
                self.writer is not None and self.writer.busy):
            now = self.now()
            while qi < len(queue) and queue[qi].arrival_s <= now:
                waiting.append(queue[qi])
                qi += 1
            while waiting and self.admit(waiting[0], now):
                waiting.pop(0)
            if self.running:
                if t_busy0 is None:
                    t_busy0 = now
                self.step()
            elif self.pending:
                self.poll(block=True)
            elif self.writer is not None and self.writer.busy:
                self.poll()  # writes finish on their threads; keep admitting arrivals meanwhile
                time.sleep(0.001)
            elif qi < len(queue):
                time.sleep(max(0.0, min(0.01, queue[qi].arrival_s - self.now())))
            done = self.take_completed()
            stats.results.extend(done)
            if progress_every and done and len(stats.results) // progress_every != \
                    (len(stats.results) - len(done)) // progress_every:
                log(f"[serve] {self.now():.0f} s: {len(stats.results)} of {len(queue)} requests done, "
                    f"{len(self.running)} running, {self.step_count} steps")
            if max_wall_s is not None and self.now() > max_wall_s:
                break
        torch.cuda.synchronize()
        stats.busy_s = self.now() - (t_busy0 or 0.0)
        stats.steps = self.step_count
        stats.particle_steps = self.particle_steps
        return stats


def latency_percentiles(results: list[Result], q=(50, 99)) -> dict:
    lat = np.array([r.latency_s for r in results], dtype=np.float64)
    if lat.size == 0:
        return {f"p{p}": float("nan") for p in q}
    return {f"p{p}": float(np.percentile(lat, p)) for p in q}
