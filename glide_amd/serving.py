"""Serving: footprint requests stepped together by continuous batching.

A request is one point release (site, inlet height, release time, particle
count, release duration, backward length). Each admitted request takes a slot
with its own cursor; every engine step advances all running slots by one step
of their own schedules (``glide_slots_advance``), runs the fused step kernel
over the shared particle pool with each wavefront reading its slot's scalars
and its bracket's window from a device-resident met ring
(``glide_step_slots``), which adds every footprint to its slot's exact
fixed-point sums (converted to float per slot when the request finishes,
``glide_fixed_finalize``; or, with ``accumulate="sort"``, keys and values
sorted into per-slot slabs by ``glide_footprint_accumulate_binned``). A request is stepped exactly
as it would be alone (its own terminus, offsets relative to its own release
time, noise keyed by its id and absolute time), so its result does not depend
on what else is running.

Admission is by free resources: a free slot (slabs are preallocated from free
HBM) and a run of free 64-particle blocks in the pool. Finished slots are found
by an asynchronous copy of the slot table; their footprints are copied to
pinned host memory on the same stream, so the slot can be reused at once.

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
from glide_amd.runner import DeviceMet, HannaOptions, fixed_exponent
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
    hour to the hour of the source met (identity, or a repeated cycle)."""

    def __init__(self, met: SyntheticMet, first_hour: int, n_hours: int, *, met_dtype: int = L.GLIDE_MET_BF16,
                 source_hour=None, device="cuda", meander_coeff: float = 0.16, meander_radius: int = 1,
                 stream: int | None = None):
        self.first_hour = int(first_hour)
        self.n_hours = int(n_hours)
        self.met_dtype = met_dtype
        self.source_hour = source_hour or (lambda h: h)
        hours = sorted({self.source_hour(h) for h in range(first_hour, first_hour + n_hours + 1)})
        dm = DeviceMet(met, hours, device=device, met_dtype=met_dtype, meander_coeff=meander_coeff,
                       meander_radius=meander_radius)
        self.dm = dm
        self.nx, self.ny, self.nz = dm.nx, dm.ny, dm.nz
        ncell = self.nz * self.ny * self.nx
        self.wind_stride = L.window_bytes(ncell, met_dtype)
        self.aux_stride = L.window_bytes(ncell, met_dtype)
        self.surf_stride = self.ny * self.nx * 12 * 4
        self.wind = torch.empty(n_hours * self.wind_stride, dtype=torch.uint8, device=device)
        self.aux = torch.empty(n_hours * self.aux_stride, dtype=torch.uint8, device=device)
        self.surf = torch.empty(n_hours * self.surf_stride, dtype=torch.uint8, device=device)
        st = L.stream_ptr() if stream is None else stream
        lib = L.lib()
        for j in range(n_hours):
            h = first_hour + j
            s3, s2, p = dm.resampled(self.source_hour(h), st)
            e3, e2, _ = dm.resampled(self.source_hour(h + 1), st)
            L.check(lib.glide_build_window(
                s3.data_ptr(), e3.data_ptr(), s2.data_ptr(), e2.data_ptr(), p.data_ptr(),
                dm.levels_f.data_ptr(), self.nz, self.ny, self.nx, meander_coeff, meander_radius, met_dtype,
                self.wind.data_ptr() + j * self.wind_stride, self.aux.data_ptr() + j * self.aux_stride,
                self.surf.data_ptr() + j * self.surf_stride, st))
        torch.cuda.synchronize()
        dm.ring.clear()
        dm.src.clear()  # the ring holds everything the steps read
        self.grid = dm.grid
        self.desc = L.GlideMetRing(
            wind=self.wind.data_ptr(), aux=self.aux.data_ptr(), surf=self.surf.data_ptr(),
            wind_stride=self.wind_stride, aux_stride=self.aux_stride, surf_stride=self.surf_stride,
            first_hour=self.first_hour, n_hours=self.n_hours, dtype=met_dtype, reserved=0)

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


class Server:
    """Continuous-batching footprint server on one GPU. All of its device work
    goes to the stream that is current when it is constructed; use it under
    that same stream."""

    def __init__(self, ring: MetRing, grid: dict, domain: dict, *, max_slots: int, pool_particles: int,
                 n_t_max: int, seed: int, dt: int = 60, hanna: HannaOptions | None = None, device="cuda",
                 stream: int | None = None, poll_every: int = 8, pinned_slabs: int = 16,
                 keep_state: bool = False, writer: ResultWriter | None = None, accumulate: str = "fixed"):
        self.ring = ring
        # accumulate: "fixed" adds each step's residence to exact int64 sums in
        # the step kernel (a finished slot's slab is converted to float when it
        # is collected); "sort" writes keys and values and sorts them by cell
        if accumulate not in ("fixed", "sort"):
            raise ValueError(f"accumulate must be 'fixed' or 'sort', not {accumulate!r}")
        self.fixed_mode = accumulate == "fixed"
        self.grid = dict(grid)
        self.writer = writer  # finished footprints go to Zarr stores before they count as done
        self.dt = int(dt)
        self.seed = int(seed)
        self.h = hanna or HannaOptions()
        self.device = device
        self.stream = L.stream_ptr() if stream is None else stream
        self.max_slots = int(max_slots)
        self.n_blocks = int(pool_particles + BLOCK - 1) // BLOCK
        self.cap = self.n_blocks * BLOCK
        self.n_t_max = int(n_t_max)
        self.poll_every = int(poll_every)
        self.keep_state = bool(keep_state)
        d = device
        # footprint pool: one slab of n_t_max time-ago bins per slot
        self.z_edges = torch.tensor(grid["z_edges"], dtype=torch.float32, device=d)
        self.n_z = len(grid["z_edges"]) - 1
        self.fg = L.GlideFootprintGrid(
            lon_min=grid["lon_min"], lon_max=grid["lon_max"], lat_min=grid["lat_min"], lat_max=grid["lat_max"],
            z_edges=self.z_edges.data_ptr(), n_r=self.max_slots, n_t=self.n_t_max, n_z=self.n_z,
            n_y=grid["n_y"], n_x=grid["n_x"], reserved=0,
            dom_lon_min=domain["lon_min"], dom_lon_max=domain["lon_max"], dom_lat_min=domain["lat_min"],
            dom_lat_max=domain["lat_max"], dom_alt_max=domain["alt_max"])
        # release-binned keys (sort): a slot's time bin is one per step (its
        # particles share the request's window end), so keys are slot x cell
        self.t_bin = torch.full((self.max_slots,), -1, dtype=torch.int32, device=d)
        self.fg.t_bin = None if self.fixed_mode else self.t_bin.data_ptr()
        self.cpb = self.n_z * grid["n_y"] * grid["n_x"]
        self.cpr = self.n_t_max * self.cpb
        self.n_cells = self.max_slots * self.cpr
        if self.n_cells >= 0xFFFFFFFF:
            raise MemoryError("the slot slabs exceed 32-bit cell keys; lower max_slots or n_t_max")
        self.footprint = torch.zeros(self.n_cells, dtype=torch.float32, device=d)
        self.release_total = torch.zeros(self.max_slots, dtype=torch.float64, device=d)
        if self.fixed_mode:
            # one int64 accumulator per slot slab and the slot's scale exponent
            # (set at admission); copies = 1: served releases are small
            self.fixed = torch.zeros(self.n_cells, dtype=torch.int64, device=d)
            self.fixed_exp = torch.zeros(self.max_slots, dtype=torch.int32, device=d)
            self.fixed_ws = torch.empty(max(8, L.lib().glide_fixed_workspace_bytes(1)), dtype=torch.uint8, device=d)
        # particle pool (structure of arrays)
        f32 = torch.float32
        types = dict(rel_idx=torch.int32, uid=torch.int32, alive=torch.uint8)
        names = ("lon", "lat", "z", "up", "vp", "wp", "um", "vm", "weight", "r_off", "we_off", "rel_idx",
                 "uid", "alive", "lo_off")
        self.P = {k: torch.zeros(self.cap, dtype=types.get(k, f32), device=d) for k in names}
        self.keys = torch.empty(self.cap, dtype=torch.int32, device=d)
        self.vals = torch.empty(self.cap, dtype=torch.float32, device=d)
        self.workspace = torch.empty(L.lib().glide_footprint_workspace_bytes(self.cap), dtype=torch.uint8, device=d)
        self.wave_slot_h = np.full(self.n_blocks, -1, np.int32)
        self.wave_slot = torch.from_numpy(self.wave_slot_h.copy()).to(d)
        # slot table (device) and its pinned host snapshot
        self.slot_bytes = C.sizeof(L.GlideSlot)
        self.slots = torch.zeros(self.max_slots * self.slot_bytes, dtype=torch.uint8, device=d)
        self.slots_host = torch.zeros_like(self.slots, device="cpu").pin_memory()
        self.slots_event = None
        self.free_slots = list(range(self.max_slots))[::-1]
        self.free_ranges = [(0, self.n_blocks)]  # (first block, count), sorted
        self.hi_blocks = 0
        self.running: dict[int, _Running] = {}
        self.pending: list[_Pending] = []
        self.pinned = [torch.empty(self.cpr, dtype=torch.float32).pin_memory() for _ in range(pinned_slabs)]
        self.dummy_escaped = torch.zeros(1, dtype=torch.int64, device=d)  # per-slot counters are in the slots
        self.max_request = 0
        self.staging = None
        self.tmpl = L.GlideStepParams(
            dt=float(self.dt), alpha=0.0, cursor_off=0.0, length_s=0.0, substep_c=self.h.substep_c,
            z_ubl_m=self.h.z_ubl_m, meander_timescale_s=self.h.meander_timescale_s, seed=self.seed,
            step_index=0, max_substeps=self.h.max_substeps,
            flags=self.h.flags() | L.F_EMIT | L.F_RELEASE_TERMINUS | (L.F_FIXED if self.fixed_mode else 0))
        self.step_count = 0
        self.particle_steps = 0
        self.completed: list[Result] = []  # finished requests not yet handed out
        self.inflight: list[torch.cuda.Event] = []  # one event per poll_every steps (backpressure)
        self.max_inflight = 4
        self.t_origin = time.perf_counter()

    # -- resources ---------------------------------------------------------
    def _take_blocks(self, nb: int) -> int | None:
        for j, (b0, cnt) in enumerate(self.free_ranges):
            if cnt >= nb:
                if cnt == nb:
                    self.free_ranges.pop(j)
                else:
                    self.free_ranges[j] = (b0 + nb, cnt - nb)
                return b0
        return None

    def _give_blocks(self, b0: int, nb: int) -> None:
        fr = sorted(self.free_ranges + [(b0, nb)])
        merged = []
        for s, c in fr:
            if merged and merged[-1][0] + merged[-1][1] == s:
                merged[-1] = (merged[-1][0], merged[-1][1] + c)
            else:
                merged.append((s, c))
        self.free_ranges = merged
        tail = merged[-1]
        self.hi_blocks = tail[0] if tail[0] + tail[1] == self.n_blocks else self.n_blocks

    def can_admit(self, req: Request) -> bool:
        end = req.release_time_s + req.duration_s
        if not self.ring.covers(end - req.length_s, end):
            raise ValueError(f"request {req.req_id}: met ring does not cover its window")
        if int(math.ceil(req.length_s / 3600.0)) > self.n_t_max:
            raise ValueError(f"request {req.req_id}: backward length exceeds {self.n_t_max} h")
        nb = (req.n_particles + BLOCK - 1) // BLOCK
        return bool(self.free_slots) and any(c >= nb for _, c in self.free_ranges)

    def admit(self, req: Request, now_s: float | None = None) -> bool:
        """Place a request on the device if a slot and pool blocks are free.
        Its arrays go up through pinned staging, so admission does not wait
        for the steps already queued."""

        if not self.can_admit(req):
            return False
        n = int(req.n_particles)
        nb = (n + BLOCK - 1) // BLOCK
        b0 = self._take_blocks(nb)
        slot = self.free_slots.pop()
        i0, m = b0 * BLOCK, nb * BLOCK
        self.hi_blocks = max(self.hi_blocks, b0 + nb)
        need = m * 4 * 11 + self.slot_bytes + nb * 4 + 256
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
        # slot entry
        end = float(req.release_time_s + req.duration_s)
        sl = L.GlideSlot(t=end, term=end - float(req.length_s), rel_time=float(req.release_time_s),
                         state=L.GLIDE_SLOT_RUNNING, k=0, noise_hi=noise_hi)
        raw = np.frombuffer(bytes(sl), dtype=np.uint8)
        self.slots[slot * self.slot_bytes:(slot + 1) * self.slot_bytes].copy_(put(raw), non_blocking=True)
        self.wave_slot_h[b0:b0 + nb] = slot
        self.wave_slot[b0:b0 + nb].view(torch.uint8).copy_(put(self.wave_slot_h[b0:b0 + nb]), non_blocking=True)
        if self.fixed_mode:
            # the slot's scale: n particles of weight float32(1/n), whose sum n * w is exact in float64
            e = fixed_exponent(n * float(np.float32(1.0 / n)), req.length_s, self.dt).reshape(1)
            self.fixed_exp[slot:slot + 1].view(torch.uint8).copy_(put(e), non_blocking=True)
        self.staging.done(j)
        self.running[slot] = _Running(req, slot, b0, nb, now_s if now_s is not None else self.now())
        return True

    # -- stepping ----------------------------------------------------------
    def now(self) -> float:
        return time.perf_counter() - self.t_origin

    def step(self) -> None:
        """One engine step of every running slot."""

        lib = L.lib()
        st = self.stream
        L.check(lib.glide_slots_advance(self.slots.data_ptr(), self.max_slots, float(self.dt), st))
        n = self.hi_blocks * BLOCK
        if n > 0:
            parts = L.GlideParticles(
                n=n, pos_stride=1, **{k: self.P[k].data_ptr() for k in (
                    "lon", "lat", "z", "up", "vp", "wp", "um", "vm", "weight", "r_off", "we_off", "rel_idx",
                    "uid", "alive", "lo_off")}, active=None)
            fx = dict(fixed=self.fixed.data_ptr(), fixed_exp=self.fixed_exp.data_ptr(), fixed_copies=1) \
                if self.fixed_mode else {}
            em = L.GlideEmit(keys=self.keys.data_ptr(), vals=self.vals.data_ptr(),
                             escaped=self.dummy_escaped.data_ptr(), active=None, **fx)
            L.check(lib.glide_step_slots(C.byref(self.tmpl), self.slots.data_ptr(), self.wave_slot.data_ptr(),
                                         C.byref(parts), C.byref(self.ring.grid), C.byref(self.ring.desc),
                                         C.byref(self.fg), C.byref(em), st))
        if n > 0 and not self.fixed_mode:
            L.check(lib.glide_footprint_accumulate_binned(
                self.keys.data_ptr(), self.vals.data_ptr(), n, self.max_slots, self.n_t_max, self.cpb,
                self.t_bin.data_ptr(), self.footprint.data_ptr(), None, self.release_total.data_ptr(),
                self.workspace.data_ptr(), self.workspace.numel(), st))
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
                fp = p.host[: p.n_t * (self.cpr // self.n_t_max)].numpy().copy()
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
        n_t = int(math.ceil(req.length_s / 3600.0))
        host = self.pinned.pop() if self.pinned else torch.empty(self.cpr, dtype=torch.float32).pin_memory()
        lo = slot * self.cpr
        if self.fixed_mode:
            # the slot's exact sums to its float slab and total, then clear them
            L.check(L.lib().glide_fixed_finalize(
                self.fixed.data_ptr() + 8 * lo, 1, self.cpr, self.cpr, self.fixed_exp.data_ptr() + 4 * slot,
                self.footprint.data_ptr() + 4 * lo, self.release_total.data_ptr() + 8 * slot,
                self.fixed_ws.data_ptr(), self.fixed_ws.numel(), self.stream))
        host.copy_(self.footprint[lo:lo + self.cpr], non_blocking=True)
        total = torch.empty(1, dtype=torch.float64).pin_memory()
        total.copy_(self.release_total[slot:slot + 1], non_blocking=True)
        ev = torch.cuda.Event()
        ev.record()
        # reuse: the copies above precede these writes on the stream
        self.footprint[lo:lo + self.cpr].zero_()
        self.release_total[slot].zero_()
        if self.fixed_mode:
            self.fixed[lo:lo + self.cpr].zero_()
        self.slots[slot * self.slot_bytes:(slot + 1) * self.slot_bytes].zero_()
        self.wave_slot_h[run.block0:run.block0 + run.n_blocks] = -1
        self.wave_slot[run.block0:run.block0 + run.n_blocks].fill_(-1)
        self._give_blocks(run.block0, run.n_blocks)
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
    def serve(self, requests: list[Request], max_wall_s: float | None = None) -> ServeStats:
        """Real-time request loop: requests arrive at ``arrival_s`` (seconds
        after the start), are admitted as slots and pool blocks free up, and
        are stepped together until all are done."""

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
            stats.results.extend(self.take_completed())
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
