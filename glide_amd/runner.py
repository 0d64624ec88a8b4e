"""Device-resident driver for one batch of releases.

Mirrors the reference's static per-step body (main.py:1383-1477, S2 order)
with everything on the GPU: the met source hours stay in HBM, the hourly
resample and window build run as column kernels on a bracket change, and each
step is one fused step kernel, which also adds every particle's residence to
the footprint (exact fixed-point sums; or, with the sort accumulation, writes
keys and values that a deterministic sort-and-reduce then accumulates) and
applies the escape kill.
Convection runs once per bracket, between the step and the gridding, as in the
reference; on that step the convection kernel emits the residence instead.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from glide_amd import _lib as L
from glide_amd.met import upload_source_hour
from glide_amd.schedule import cursor_schedule, release_arrays
from glide_amd.synthetic import SyntheticMet


@dataclass
class HannaOptions:
    """HannaScheme constructor options (reference hanna.py:440-520); the
    shipped configs use max_substeps=6 and meander on."""

    meander_enabled: bool = True
    meander_coefficient: float = 0.16
    meander_stencil_radius: int = 1
    meander_timescale_s: float = 1800.0
    substep_c: float = 0.5
    max_substeps: int = 6
    z_ubl_m: float = 2.0
    flexpart_tl_floors: bool = True
    surface_layer_override: bool = False

    def flags(self) -> int:
        f = 0
        if self.meander_enabled:
            f |= L.F_MEANDER
        if self.flexpart_tl_floors:
            f |= L.F_TL_FLOORS
        if self.surface_layer_override:
            f |= L.F_SL_OVERRIDE
        return f


@dataclass
class ConvectionOptions:
    """EmanuelReducedConvection constructor options (emanuel.py:526-559)."""

    closure_c: float = 0.03
    trigger_dtv_k: float = 0.9
    min_cape_j_kg: float = 50.0
    min_cloud_depth_m: float = 500.0
    # parcel_lift: column CAPE / LCL / cloud top / T_v excess from a surface
    # parcel lifted through each column's own temperature and humidity
    # (glide_column_cape); otherwise the met supplies them per column
    parcel_lift: bool = False


# counters each step's active count is spread over (GlideEmit.active_copies)
ACTIVE_COPIES = 64


def fixed_exponent(weight_sum: np.ndarray | float, length_s: float, dt: float) -> np.ndarray:
    """Power-of-two scale of a release's fixed-point footprint: its residence
    is at most T = weight_sum * (length_s + 2 dt) (every particle is active over
    at most that span), and e = 61 - ceil(log2 T) keeps T * 2^e <= 2^61, inside
    int64 with room to spare; one contribution then rounds by 2^-e, below
    T * 2^-61. T = 0 gives 0."""

    bound = np.abs(np.asarray(weight_sum, np.float64)) * (float(length_s) + 2.0 * float(dt))
    e = np.zeros(bound.shape, np.int32)
    pos = bound > 0
    e[pos] = 61 - np.ceil(np.log2(bound[pos])).astype(np.int32)
    return e


def _dev(a, dtype=None, device="cuda"):
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.to(device)


class DeviceMet:
    """Met source hours resident in HBM, the resampled ring and one packed
    window, for a :class:`SyntheticMet` grid."""

    def __init__(self, met: SyntheticMet, hours, *, device="cuda", met_dtype: int = L.GLIDE_MET_F32,
                 meander_coeff: float = 0.16, meander_radius: int = 1, ingest=None, sources: dict | None = None,
                 defer: bool = False):
        self.met = met
        # ingest (glide_amd.ingest.MetIngest): source hours stream in ahead of
        # the cursor instead of being made resident up front
        self.ingest = ingest
        spec = met.spec
        self.device = device
        self.nx, self.ny, self.nz = spec.nx, spec.ny, spec.za
        self.met_dtype = met_dtype
        self.meander_coeff = meander_coeff
        self.meander_radius = meander_radius
        lev = spec.agl_levels
        self.levels_np = lev
        self.levels_d = _dev(lev, torch.float64, device)
        self.levels_f = _dev(lev.astype(np.float32), None, device)
        self.lon = met.lon
        self.lat = met.lat
        self.terrain = _dev(met.terrain, None, device)
        mlon = np.clip(111320.0 * np.cos(np.deg2rad(self.lat)), 1.0, None)
        self.mlon = _dev(mlon, torch.float64, device)
        # sources: another DeviceMet's resident source hours (the same met in
        # another window precision), shared instead of uploaded again
        self.src = sources if sources is not None else {}
        self.zp = None
        if ingest is None:
            if not defer:  # defer: the caller uploads hours as it needs them (upload)
                self.upload(hours)
        else:
            first = max(hours) if ingest.direction < 0 else min(hours)
            self.zp = ingest.get(first)["src3"].shape[1]
        ncell = self.nz * self.ny * self.nx
        self.wind = torch.empty(L.window_bytes(ncell, met_dtype), dtype=torch.uint8, device=device)
        self.aux = torch.empty(L.window_bytes(ncell, met_dtype), dtype=torch.uint8, device=device)
        self.surf = torch.empty(self.ny * self.nx * 12 * 4, dtype=torch.uint8, device=device)  # float in both
        self.ring = {}
        self.window_hour = None
        self.grid = L.GlideGrid(nx=self.nx, ny=self.ny, nz=self.nz, reserved=0,
                                lon0=float(self.lon[0]), lon1=float(self.lon[-1]),
                                lat0=float(self.lat[0]), lat1=float(self.lat[-1]),
                                levels=self.levels_f.data_ptr())
        self.window = L.GlideWindow(wind=self.wind.data_ptr(), aux=self.aux.data_ptr(),
                                    surf=self.surf.data_ptr(), dtype=met_dtype, reserved=0)

    def upload(self, hours) -> None:
        """Make source hours resident (those not already)."""

        for h in hours:
            if h not in self.src:
                self.src[h] = upload_source_hour(self.met.hour(h), self.device)
        if self.zp is None and self.src:
            self.zp = next(iter(self.src.values()))["src3"].shape[1]

    def resampled(self, h: int, stream: int):
        """Resampled hour ``h`` (out3 [5, Za, Y, X], surface [4, Y, X], p_hpa [Za])."""

        if h in self.ring:
            return self.ring[h]
        s = self.source(h, stream)
        out3 = torch.empty((5, self.nz, self.ny, self.nx), dtype=torch.float32, device=self.device)
        p_out = torch.empty(self.nz, dtype=torch.float64, device=self.device)
        dlat = float(self.lat[1] - self.lat[0])
        dlon = float(self.lon[1] - self.lon[0])
        L.check(L.lib().glide_resample_hour(
            s["src3"].data_ptr(), 5, self.zp, self.ny, self.nx, s["height"].data_ptr(),
            self.levels_d.data_ptr(), self.nz, self.terrain.data_ptr(), self.mlon.data_ptr(),
            dlat, dlon, s["p_hpa"].data_ptr(), 0, 1, 2, out3.data_ptr(), p_out.data_ptr(), stream))
        # the window walks backward in time: keep only the hours next to h
        for old in [k for k in self.ring if abs(k - h) > 2]:
            del self.ring[old]
        self.ring[h] = (out3, s["src2"], p_out)
        return self.ring[h]

    def source(self, h: int, stream: int) -> dict:
        """Device arrays of source hour ``h``, ordered before work on ``stream``."""

        if self.ingest is not None:
            return self.ingest.get(h, stream)
        if h not in self.src:
            raise ValueError(f"met hour {h} not resident (have {sorted(self.src)})")
        return self.src[h]

    def build_window(self, h: int, stream: int) -> None:
        if self.window_hour == h:
            return
        s3, s2, p = self.resampled(h, stream)
        e3, e2, _ = self.resampled(h + 1, stream)
        L.check(L.lib().glide_build_window(
            s3.data_ptr(), e3.data_ptr(), s2.data_ptr(), e2.data_ptr(), p.data_ptr(),
            self.levels_f.data_ptr(), self.nz, self.ny, self.nx, self.meander_coeff,
            self.meander_radius, self.met_dtype, self.wind.data_ptr(), self.aux.data_ptr(),
            self.surf.data_ptr(), stream))
        self.window_hour = h


class BatchEngine:
    """One batch of releases on the device, stepped in the reference's order."""

    def __init__(self, cfg: dict, *, device="cuda", met: SyntheticMet | None = None,
                 met_dtype: int = L.GLIDE_MET_F32, hanna: HannaOptions | None = None,
                 convection: ConvectionOptions | None = None, compensated: bool = False,
                 device_met: DeviceMet | None = None, reorder_every: int = 0, use_graph: bool = False,
                 compact: bool = False, release_terminus: bool = False, diagnostics: bool = False,
                 ingest=None, step_block: int = 0, overlap: bool = False, binned_keys: bool = True,
                 accumulate: str = "fixed", fixed_copies: int = 0, bin_by_release: bool | None = None,
                 slabs: bool | None = None, cell_positions: bool = False):
        self.cfg = cfg
        # accumulate: how each step's residence reaches the footprint.
        #   "fixed": inside the step kernel, as exact 64-bit fixed-point integer
        #     sums per cell (GLIDE_F_FIXED), converted to float32 once when the
        #     footprint is read; order-free, so bit-identical under any launch
        #     shape, particle order or re-binning;
        #   "sort": keys and values written by the step, then a stable radix
        #     sort by cell and a fixed-order segmented float reduction
        #     (glide_footprint_accumulate).
        if accumulate not in ("fixed", "sort"):
            raise ValueError(f"accumulate must be 'fixed' or 'sort', not {accumulate!r}")
        self.fixed_mode = accumulate == "fixed"
        # fixed_copies: copies of the fixed-point accumulator the step's
        # workgroups spread their atomics over (0: as many as fit in 512 MiB, at
        # most 16; a cell crowded by a point release then takes 1/copies of the
        # atomics per address); the cells are the sums over the copies. The
        # step sums each wavefront's lanes per distinct cell before its atomics,
        # so 16 copies already keep a point release's cell from queueing, and
        # every copy is read once to make the float footprint
        self.fixed_copies_req = int(fixed_copies)
        # bin_by_release: the re-binning keeps each release's particles
        # together (release-major keys), so a wavefront's particles share a
        # release and its fixed-point atomics sum long runs; None: with the
        # fixed-point accumulation and more than one release
        self.bin_by_release = bin_by_release
        # binned_keys (sort): footprint keys without the time bin (a per-release
        # table carries it): shorter keys, fewer sort passes, the same sums bit
        # for bit
        self.binned_keys = bool(binned_keys) and not self.fixed_mode
        # overlap (sort): each eager step's footprint accumulation runs on a
        # second stream, concurrently with the next step's kernel (keys and
        # values double-buffered; a buffer is reused only after its accumulation)
        self.overlap = bool(overlap) and not self.fixed_mode
        # threads per workgroup of the step kernel (0 = the library's default);
        # a tuning parameter (glide_amd.autotune), results do not depend on it
        self.step_block = int(step_block)
        # diagnostics: per-step mean particle position (inside the step body)
        # and hourly grid-mean winds, for the trajectory table
        self.diagnostics = bool(diagnostics)
        # cell_positions: the precision layout, each particle's horizontal
        # position as its nearest met grid point (GlideParticles.cell) plus
        # float32 offsets in degrees (lon/lat hold the offsets)
        self.cell_positions = bool(cell_positions)
        if self.cell_positions and self.diagnostics:
            raise ValueError("diagnostics take absolute positions; cell_positions=True has none on the device")
        self.reorder_every = int(reorder_every)
        # compact: on every new met hour, sort the particles by met cell with the
        # ones that can never be active again (finished releases, escaped) last,
        # and launch over the live prefix only
        self.compact = bool(compact)
        # release_terminus: each release is stepped only while the cursor is
        # above its own terminus, as when it runs alone (its result does not
        # depend on the batch); default: the reference's batch terminus
        self.release_terminus = bool(release_terminus)
        self.extra_flags = (L.F_RELEASE_TERMINUS if self.release_terminus else 0) | (
            L.F_FIXED if self.fixed_mode else 0)
        self.use_graph = bool(use_graph)
        self._graphs: dict[tuple[int, int, int], int] = {}
        self.graph_launches = 0
        self.device = device
        self.h = hanna or HannaOptions()
        self.conv_opts = convection if convection is not None else (
            ConvectionOptions() if cfg.get("convection") else None)
        rels = cfg["releases"]
        self.length_s = max(int(r.length_s) for r in rels)
        # releases of different backward lengths pass each particle's earliest
        # active offset (lo_off); one length keeps the scalar form
        self.per_particle_length = len({int(r.length_s) for r in rels}) > 1 or self.release_terminus
        self.sched = cursor_schedule(rels, int(cfg["dt"]))
        # index of the last step of each step's met hour (the hourly compaction's horizon)
        self._bracket_end = list(range(len(self.sched)))
        for i in range(len(self.sched) - 2, -1, -1):
            if self.sched[i + 1].bracket_hour == self.sched[i].bracket_hour:
                self._bracket_end[i] = self._bracket_end[i + 1]
        hours = sorted({s.bracket_hour for s in self.sched} | {s.bracket_hour + 1 for s in self.sched})
        self.met = device_met or DeviceMet(met or SyntheticMet(cfg["spec"]), hours, device=device,
                                           met_dtype=met_dtype, ingest=ingest,
                                           meander_coeff=self.h.meander_coefficient,
                                           meander_radius=self.h.meander_stencil_radius)
        self.seed = int(cfg["seed"])
        self.ra = release_arrays(rels)
        self.n = int(self.ra.lon.size)
        self.n_live = self.n
        g = cfg["grid"]
        dom = cfg["domain"]
        self.grid_cfg = g
        self.z_edges = _dev(np.asarray(g["z_edges"], np.float32), None, device)
        self.fg = L.GlideFootprintGrid(
            lon_min=g["lon_min"], lon_max=g["lon_max"], lat_min=g["lat_min"], lat_max=g["lat_max"],
            z_edges=self.z_edges.data_ptr(), n_r=g["n_r"], n_t=g["n_t"], n_z=g["n_z"],
            n_y=g["n_y"], n_x=g["n_x"], reserved=0,
            dom_lon_min=dom["lon_min"], dom_lon_max=dom["lon_max"], dom_lat_min=dom["lat_min"],
            dom_lat_max=dom["lat_max"], dom_alt_max=dom["alt_max"])
        self.n_cells = g["n_r"] * g["n_t"] * g["n_z"] * g["n_y"] * g["n_x"]
        self.cells_per_release = self.n_cells // g["n_r"]
        # slabs: each release's footprint slab holds only its own hourly time
        # bins (ceil(length / 3600)), addressed by 64-bit offsets, so batches
        # beyond 2^32 flat cells fit (fixed-point accumulation only); None:
        # on when the flat layout would not fit 32-bit keys
        if slabs is None:
            slabs = self.fixed_mode and self.n_cells >= 0xFFFFFFFF
        if slabs and not self.fixed_mode:
            raise ValueError("per-release slabs need the fixed-point accumulation (accumulate='fixed')")
        if not slabs and self.n_cells >= 0xFFFFFFFF:
            raise ValueError(f"{self.n_cells} footprint cells exceed 32-bit keys; use slabs=True "
                             "(fixed-point accumulation)")
        self.slabs = bool(slabs)
        self.cells_per_bin = g["n_z"] * g["n_y"] * g["n_x"]
        if self.slabs:
            # a particle is active until the cursor reaches its r_off - length
            # (under the batch terminus), so its time-ago reaches at most
            # length + duration: hourly bins 0 .. floor((length + duration) / 3600)
            nt = np.zeros(g["n_r"], np.int64)
            for r in rels:
                loc = int(r.release_idx) - int(rels[0].release_idx)  # batch-local slab (release_arrays)
                nt[loc] = max(nt[loc], (int(r.length_s) + int(r.duration_s)) // 3600 + 1)
            self.slab_nt_np = np.clip(nt, 1, g["n_t"]).astype(np.int32)
            sizes = self.slab_nt_np.astype(np.int64) * self.cells_per_bin
            self.slab_off_np = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
            self.acc_cells = int(sizes.sum())
            self.slab_off = _dev(self.slab_off_np, None, device)
            self.slab_nt = _dev(self.slab_nt_np, None, device)
            self.fg.slab_off = self.slab_off.data_ptr()
            self.fg.slab_nt = self.slab_nt.data_ptr()
            self.fg.slab_cells = self.acc_cells
        else:
            self.acc_cells = self.n_cells
        self.compensated = compensated
        self._alloc()
        self.reset()

    STATE = ("lon", "lat", "z", "up", "vp", "wp", "um", "vm", "weight", "r_off", "we_off", "rel_idx",
             "uid", "alive", "lo_off")

    def _alloc(self):
        d = self.device
        n = self.n
        f32 = torch.float32
        dtypes = dict(rel_idx=torch.int32, uid=torch.int32, alive=torch.uint8)
        self.bufs = []
        keys = self.STATE + (("cell",) if self.cell_positions else ())
        dtypes["cell"] = torch.int32
        for _ in range(2):  # double buffer for the reordering permutation
            self.bufs.append({k: torch.empty(n, dtype=dtypes.get(k, f32), device=d) for k in keys})
        self.cur = 0
        self._parts = [self._make_parts(b) for b in self.bufs]
        self.keys = torch.empty(n, dtype=torch.int32, device=d)
        self.vals = torch.empty(n, dtype=f32, device=d)
        self._ext: dict[int, torch.cuda.Stream] = {}  # torch handles of the raw streams the engine is driven on
        self.emit_bufs = [(self.keys, self.vals)]
        if self.overlap:
            self.emit_bufs.append((torch.empty(n, dtype=torch.int32, device=d), torch.empty(n, dtype=f32, device=d)))
        # per emit buffer: the release time-bin table of binned keys and the
        # footprint grid naming it
        g = self.grid_cfg
        self.t_bins = [torch.full((g["n_r"],), -1, dtype=torch.int32, device=d) for _ in self.emit_bufs]
        self.fgs = []
        for tb in self.t_bins:
            fg = L.GlideFootprintGrid(**{f: getattr(self.fg, f) for f, _ in L.GlideFootprintGrid._fields_})
            fg.t_bin = tb.data_ptr() if self.binned_keys else None
            self.fgs.append(fg)
        self.acc_stream = None
        if self.overlap:
            # default priority, on a hardware queue apart from the step's
            self.acc_stream = L.queue_stream(torch.device(d).index or 0)
            self.ev_step = [torch.cuda.Event(), torch.cuda.Event()]
            self.ev_acc = [torch.cuda.Event(), torch.cuda.Event()]
        self.acc_pending = [False, False]
        self.escaped = torch.zeros(max(len(self.sched), 1), dtype=torch.int64, device=d)
        # per-step active counts, each spread over ACTIVE_COPIES counters (one
        # atomic per workgroup on a single counter would queue); read as sums
        self._active = torch.zeros((max(len(self.sched), 1), ACTIVE_COPIES), dtype=torch.int64, device=d)
        self._footprint = torch.zeros(self.acc_cells, dtype=f32, device=d)
        # Neumaier terms of the sort accumulation; the fixed-point sums are exact
        self.comp = torch.zeros(self.n_cells, dtype=f32, device=d) if self.compensated and not self.fixed_mode else None
        self._release_total = torch.zeros(self.grid_cfg["n_r"], dtype=torch.float64, device=d)
        self._fixed_stale = False
        if self.fixed_mode:
            copies = self.fixed_copies_req
            if copies <= 0:
                copies = 1
                while copies < 16 and 2 * copies * 8 * self.acc_cells <= (512 << 20):
                    copies *= 2
            self.fixed_copies = copies
            self.fixed = torch.zeros(copies * self.acc_cells, dtype=torch.int64, device=d)
            self.fixed_exp = _dev(self.fixed_exponents(), None, d)
            self.fixed_ws = torch.empty(max(8, L.lib().glide_fixed_workspace_bytes(self.grid_cfg["n_r"])),
                                        dtype=torch.uint8, device=d)
        ws = L.lib().glide_footprint_workspace_bytes(n)
        self.workspace = torch.empty(ws, dtype=torch.uint8, device=d)
        self.sort_ws = torch.empty(L.lib().glide_sort_workspace_bytes(n), dtype=torch.uint8, device=d)
        self.ro_keys = torch.empty(n, dtype=torch.int32, device=d)
        self.ro_idx = torch.empty(n, dtype=torch.int32, device=d)
        self.ro_keys_s = torch.empty(n, dtype=torch.int32, device=d)
        self.ro_idx_s = torch.empty(n, dtype=torch.int32, device=d)
        ncell = self.met.nz * self.met.ny * self.met.nx
        by_rel = self.bin_by_release if self.bin_by_release is not None else (
            self.fixed_mode and self.grid_cfg["n_r"] > 1)
        self.bin_groups = int(self.grid_cfg["n_r"]) if by_rel else 1
        # cells (per release group), the later bucket's, the sentinel; the
        # library drops the level from the cell when that key would not fit 32 bits
        if 2 * self.bin_groups * ncell >= 0xFFFFFFFF:
            ncell = self.met.ny * self.met.nx
        self.ro_bits = max(1, int(2 * self.bin_groups * ncell).bit_length())
        nc = self.met.ny * self.met.nx
        self.i_lcl = torch.empty(nc, dtype=torch.int32, device=d)
        self.i_lnb = torch.empty(nc, dtype=torch.int32, device=d)
        self.m_b = torch.empty(nc, dtype=f32, device=d)
        self.layer_mass = torch.empty(self.met.nz, dtype=f32, device=d)
        self.colsum_err = torch.zeros(nc, dtype=torch.float64, device=d)
        self.conv_lift = torch.empty((4, self.met.ny, self.met.nx), dtype=f32, device=d)
        self.conv_cloud = torch.empty(4 * nc, dtype=f32, device=d)  # per-column cloud totals of glide_convect
        self.live = torch.zeros(1, dtype=torch.int32, device=d)
        # Step table for the captured step body: one GlideStepParams row per
        # schedule step (the flags of a plain emitting step), and the device
        # cursor that the graph advances.
        flags = self.h.flags() | L.F_EMIT
        if self.h.max_substeps < 1 or self.h.substep_c <= 0 or any(st.delta_s <= 0 for st in self.sched):
            raise ValueError("step table rows need max_substeps >= 1, substep_c > 0 and dt > 0")
        rows = (L.GlideStepParams * max(len(self.sched), 1))(*[self.step_params(st, flags) for st in self.sched])
        self.step_table = torch.frombuffer(bytearray(bytes(rows)), dtype=torch.uint8).to(d)
        self.step_tmpl = self.step_params(self.sched[0], flags) if self.sched else None
        self.row = torch.zeros(L.GLIDE_ROW_WORDS, dtype=torch.int32, device=d)  # cursor, finished-workgroup counts
        if self.diagnostics:
            self.pos_means = torch.zeros((max(len(self.sched), 1), 3), dtype=torch.float64, device=d)
            self.means_ws = torch.empty(L.lib().glide_means_workspace_bytes(), dtype=torch.uint8, device=d)
            self.hour_wind: dict[int, torch.Tensor] = {}

    def _make_parts(self, b) -> L.GlideParticles:
        return L.GlideParticles(
            n=self.n, pos_stride=1, lon=b["lon"].data_ptr(), lat=b["lat"].data_ptr(), z=b["z"].data_ptr(),
            up=b["up"].data_ptr(), vp=b["vp"].data_ptr(), wp=b["wp"].data_ptr(), um=b["um"].data_ptr(),
            vm=b["vm"].data_ptr(), weight=b["weight"].data_ptr(), r_off=b["r_off"].data_ptr(),
            we_off=b["we_off"].data_ptr(), rel_idx=b["rel_idx"].data_ptr(), uid=b["uid"].data_ptr(),
            alive=b["alive"].data_ptr(), active=None,
            lo_off=b["lo_off"].data_ptr() if self.per_particle_length else None,
            cell=b["cell"].data_ptr() if self.cell_positions else None)

    # -- the cell layout: grid point + offsets <-> absolute coordinates ----
    def _cell_consts(self):
        g = self.met.grid
        f = np.float32
        return (f(g.lon0), f(g.lat0), f((g.lon1 - g.lon0) / (g.nx - 1)), f((g.lat1 - g.lat0) / (g.ny - 1)),
                f((g.nx - 1) / (g.lon1 - g.lon0)), f((g.ny - 1) / (g.lat1 - g.lat0)), g.nx, g.ny)

    def _cell_points(self, cell):
        """The float32 grid point of each cell, as the kernels form it
        (fma(cx, dlon, lon0), one rounding of an exact double sum)."""

        lon0, lat0, dlon, dlat, _, _, _, _ = self._cell_consts()
        c = np.asarray(cell, np.int64)
        cy, cx = c >> 16, c & 0xFFFF
        rlon = (np.float64(lon0) + cx * np.float64(dlon)).astype(np.float32)
        rlat = (np.float64(lat0) + cy * np.float64(dlat)).astype(np.float32)
        return rlon, rlat

    def to_cell(self, lon, lat):
        """(cell, lon offset, lat offset) of absolute positions: the nearest
        grid point and the float32 offsets from it."""

        lon0, lat0, _, _, sx, sy, nx, ny = self._cell_consts()
        lon = np.asarray(lon, np.float64)
        lat = np.asarray(lat, np.float64)
        cx = np.clip(np.floor((lon - lon0) * np.float64(sx) + 0.5), 0, nx - 1).astype(np.int64)
        cy = np.clip(np.floor((lat - lat0) * np.float64(sy) + 0.5), 0, ny - 1).astype(np.int64)
        cell = ((cy << 16) | cx).astype(np.int32)
        rlon, rlat = self._cell_points(cell)
        return cell, (lon - rlon).astype(np.float32), (lat - rlat).astype(np.float32)

    def from_cell(self, cell, off_lon, off_lat):
        """Absolute float64 positions of a cell layout: grid point + offset, exactly."""

        rlon, rlat = self._cell_points(cell)
        return (rlon.astype(np.float64) + np.asarray(off_lon, np.float64),
                rlat.astype(np.float64) + np.asarray(off_lat, np.float64))

    def _put_positions(self, b, lon, lat) -> None:
        if self.cell_positions:
            cell, ol, oa = self.to_cell(lon, lat)
            b["cell"].copy_(torch.from_numpy(cell))
            b["lon"].copy_(torch.from_numpy(ol))
            b["lat"].copy_(torch.from_numpy(oa))
        else:
            b["lon"].copy_(torch.from_numpy(np.asarray(lon, np.float32)))
            b["lat"].copy_(torch.from_numpy(np.asarray(lat, np.float32)))

    def _get_positions(self, b, order):
        lon, lat = b["lon"].cpu().numpy()[order], b["lat"].cpu().numpy()[order]
        if self.cell_positions:
            return self.from_cell(b["cell"].cpu().numpy()[order], lon, lat)
        return lon, lat

    @property
    def parts(self) -> L.GlideParticles:
        """The current buffer's particles, limited to the live prefix."""

        p = self._parts[self.cur]
        if self.n_live == self.n:
            return p
        q = L.GlideParticles(**{f: getattr(p, f) for f, _ in L.GlideParticles._fields_})
        q.n = self.n_live
        return q

    def __getattr__(self, name):
        # current-buffer state arrays: eng.lon, eng.z, eng.alive, ...
        if name in BatchEngine.STATE and "bufs" in self.__dict__:
            return self.__dict__["bufs"][self.__dict__["cur"]][name]
        raise AttributeError(name)

    def fixed_exponents(self) -> np.ndarray:
        """Per-release power-of-two scale of the fixed-point footprint: release
        r's residence is at most T_r = (sum of its weights) * (longest backward
        length + 2 dt) (every particle is active over at most that span), and
        e_r = 61 - ceil(log2 T_r) keeps T_r * 2^e_r <= 2^61, inside int64 with
        room to spare; the rounding of one contribution, 2^-e_r, is then below
        T_r * 2^-61."""

        n_r = self.grid_cfg["n_r"]
        w = np.bincount(np.clip(self.ra.rel_idx.astype(np.int64), 0, n_r - 1),
                        weights=np.abs(self.ra.weight.astype(np.float64)), minlength=n_r)[:n_r]
        return fixed_exponent(w, self.length_s, self.cfg["dt"])

    def _finalize_fixed(self, stream: int) -> None:
        """Fixed-point sums to the float32 footprint and float64 totals."""

        if self.slabs:
            g = self.grid_cfg
            L.check(L.lib().glide_fixed_finalize_slabs(
                self.fixed.data_ptr(), self.fixed_copies, self.acc_cells, self.slab_off.data_ptr(),
                self.slab_nt.data_ptr(), self.cells_per_bin, int(self.slab_nt_np.max()), self.fixed_exp.data_ptr(),
                0, g["n_r"], self._footprint.data_ptr(), self._release_total.data_ptr(), self.fixed_ws.data_ptr(),
                self.fixed_ws.numel(), stream))
        else:
            L.check(L.lib().glide_fixed_finalize(
                self.fixed.data_ptr(), self.fixed_copies, self.n_cells, self.cells_per_release,
                self.fixed_exp.data_ptr(), self._footprint.data_ptr(), self._release_total.data_ptr(),
                self.fixed_ws.data_ptr(), self.fixed_ws.numel(), stream))
        self._fixed_stale = False

    @property
    def active_steps(self) -> torch.Tensor:
        """Active particles per schedule step (int64 [steps]): the step's
        particle-steps, counted by the step kernel."""

        return self._active.sum(dim=1)

    def fixed_cells(self) -> torch.Tensor:
        """The fixed-point sums per cell (int64 [n_cells], summed over the
        accumulator's copies): cell c holds its residence times 2^fixed_exp[r]."""

        return self.fixed.view(self.fixed_copies, self.acc_cells).sum(dim=0)

    @property
    def footprint(self) -> torch.Tensor:
        """The footprint, float32 on the device, holding every step run so far
        (ordered before later work on the current stream): [n_r * n_t * n_z *
        n_y * n_x], or with per-release slabs the releases' slabs [T_r][Z][Y][X]
        back to back (``release_footprint(r)``)."""

        if self._fixed_stale:
            self._finalize_fixed(L.stream_ptr())
        return self._footprint

    @property
    def release_total(self) -> torch.Tensor:
        """Per-release residence (float64 [n_r]) of the steps run so far."""

        if self._fixed_stale:
            self._finalize_fixed(L.stream_ptr())
        return self._release_total

    def reset(self):
        """Back to the release state (positions at the sources, state zero,
        release order)."""

        if getattr(self, "acc_pending", None) is not None:
            self.join()  # no accumulation may still be writing the footprint being cleared
        self.cur = 0
        b = self.bufs[0]
        ra = self.ra
        for k, v in (("z", ra.z), ("weight", ra.weight), ("r_off", ra.r_off),
                     ("we_off", ra.we_off), ("rel_idx", ra.rel_idx), ("uid", ra.uid.view(np.int32)),
                     ("lo_off", ra.te_off if self.release_terminus else ra.lo_off)):
            b[k].copy_(torch.from_numpy(np.ascontiguousarray(v)))
        self._put_positions(b, ra.lon, ra.lat)
        for k in ("up", "vp", "wp", "um", "vm"):
            b[k].zero_()
        b["alive"].fill_(1)
        self.escaped.zero_()
        self._active.zero_()
        self._footprint.zero_()
        if self.comp is not None:
            self.comp.zero_()
        self._release_total.zero_()
        if self.fixed_mode:
            self.fixed.zero_()
            self._fixed_stale = False
        if self.diagnostics:
            self.pos_means.zero_()
        self.step_i = 0
        self.n_live = self.n
        self.horizon = -1e30  # no later bucket until the first hourly compaction
        self.last_bracket = None
        self.step_events = None
        self.row.zero_()

    def reorder(self, st, stream: int, count_live: bool = False) -> None:
        """Sort the whole particle state by met cell (gather locality);
        particles that can never be active again go to the end. With
        ``count_live`` (the hourly compaction) the particles that cannot be
        active before the last step of this met hour (released further back)
        go after the ones that can, and the launches that follow cover only
        those (one device-to-host read of the count); a reorder within the hour
        keeps the same split."""

        lib = L.lib()
        p = self.step_params(st, 0)
        if count_live:
            self.live.zero_()
            self.horizon = float(np.float32(self.sched[self._bracket_end[st.index]].cursor_off))
        L.check(lib.glide_reorder_keys(C.byref(p), C.byref(self._parts[self.cur]), C.byref(self.met.grid),
                                       self.horizon, self.bin_groups, self.ro_keys.data_ptr(), self.ro_idx.data_ptr(),
                                       self.live.data_ptr() if count_live else None, stream))
        # the hourly compaction sorts everything; a re-binning within the hour
        # sorts only the live prefix (what the steps launch over) and carries
        # the rest over in place (the key kernel wrote idx[i] = i)
        n_sort = self.n if count_live else self.n_live
        L.check(lib.glide_sort_pairs(self.ro_keys.data_ptr(), self.ro_idx.data_ptr(), n_sort, self.ro_bits,
                                     self.ro_keys_s.data_ptr(), self.ro_idx_s.data_ptr(), self.sort_ws.data_ptr(),
                                     self.sort_ws.numel(), stream))
        if n_sort < self.n:
            with torch.cuda.stream(self._stream_obj(stream)):
                self.ro_idx_s[n_sort:].copy_(self.ro_idx[n_sort:])
        nxt = 1 - self.cur
        L.check(lib.glide_permute_particles(C.byref(self._parts[self.cur]), C.byref(self._parts[nxt]),
                                            self.ro_idx_s.data_ptr(), stream))
        self.cur = nxt
        if count_live:
            self.n_live = int(self.live.item())

    def step_params(self, st, flags: int) -> L.GlideStepParams:
        h = self.h
        return L.GlideStepParams(
            dt=float(st.delta_s), alpha=float(st.alpha), cursor_off=float(st.cursor_off),
            length_s=float(self.length_s), substep_c=h.substep_c, z_ubl_m=h.z_ubl_m,
            meander_timescale_s=h.meander_timescale_s, seed=self.seed, step_index=st.noise_step,
            max_substeps=h.max_substeps, flags=flags | self.extra_flags, block=self.step_block)

    def emit(self, st, buf: int = 0) -> L.GlideEmit:
        keys, vals = self.emit_bufs[buf]
        return L.GlideEmit(keys=keys.data_ptr(), vals=vals.data_ptr(),
                           escaped=self.escaped.data_ptr() + 8 * st.index,
                           active=self._active.data_ptr() + 8 * ACTIVE_COPIES * st.index,
                           active_copies=ACTIVE_COPIES, **self._fixed_emit())

    def _fixed_emit(self) -> dict:
        if not self.fixed_mode:
            return {}
        return dict(fixed=self.fixed.data_ptr(), fixed_exp=self.fixed_exp.data_ptr(), fixed_copies=self.fixed_copies)

    def _stream_obj(self, ptr: int):
        s = self._ext.get(ptr)
        if s is None:
            s = self._ext[ptr] = (torch.cuda.ExternalStream(ptr) if ptr else torch.cuda.default_stream())
        return s

    def join(self, stream: int | None = None) -> None:
        """Order the work that follows on ``stream`` (default: the current
        stream) after every footprint accumulation still running on the
        accumulation stream (overlap mode), or convert the fixed-point sums to
        the float footprint on ``stream`` (fixed mode)."""

        ptr = L.stream_ptr() if stream is None else stream
        if self.fixed_mode:
            if self._fixed_stale:
                self._finalize_fixed(ptr)
            return
        self._wait_acc(ptr)

    def _wait_acc(self, ptr: int) -> None:
        """Order ``ptr`` after the accumulations running on the accumulation stream."""

        if not self.overlap:
            return
        main = self._stream_obj(ptr)
        for b in (0, 1):
            if self.acc_pending[b]:
                main.wait_event(self.ev_acc[b])
                self.acc_pending[b] = False

    def _capture(self, stream: int, steps: int = 1) -> int:
        """Capture ``steps`` plain step bodies back to back (table-driven step,
        which also advances the device cursor, and the footprint accumulation)
        for the current particle buffer. (A body whose accumulation of the
        previous step runs on a parallel branch beside the step kernel
        replayed more than twice as slowly as this serial one on MI355X, 0.35
        against 0.15 ms per step: the step kernel fills every CU, so the
        branches hardly overlap.)"""

        lib = L.lib()
        if stream == 0:
            raise ValueError("graph capture needs a non-default stream")
        em = L.GlideEmit(keys=self.keys.data_ptr(), vals=self.vals.data_ptr(), escaped=self.escaped.data_ptr(),
                         active=self._active.data_ptr(), active_copies=ACTIVE_COPIES, **self._fixed_emit())
        L.check(lib.glide_graph_begin(stream))
        try:
            for _ in range(steps):
                L.check(lib.glide_step_table(C.byref(self.step_tmpl), self.step_table.data_ptr(), len(self.sched),
                                             self.row.data_ptr(),
                                             C.byref(self.parts), C.byref(self.met.grid), C.byref(self.met.window),
                                             C.byref(self.fgs[0]), C.byref(em), stream))
                self._accumulate(stream)
                if self.diagnostics:
                    self._position_means(stream, row=True)
        finally:
            h = C.c_void_p()
            rc = lib.glide_graph_end(stream, C.byref(h))
        L.check(rc)
        return h.value

    def _accumulate(self, stream: int, buf: int = 0) -> None:
        if self.fixed_mode:
            self._fixed_stale = True  # the step kernels added to the fixed-point sums
            return
        if self.n_live == 0:
            return
        keys, vals = self.emit_bufs[buf]
        if self.binned_keys:
            g = self.grid_cfg
            L.check(L.lib().glide_footprint_accumulate_binned(
                keys.data_ptr(), vals.data_ptr(), self.n_live, g["n_r"], g["n_t"], g["n_z"] * g["n_y"] * g["n_x"],
                self.t_bins[buf].data_ptr(), self._footprint.data_ptr(),
                self.comp.data_ptr() if self.comp is not None else None, self._release_total.data_ptr(),
                self.workspace.data_ptr(), self.workspace.numel(), stream))
            return
        L.check(L.lib().glide_footprint_accumulate(
            keys.data_ptr(), vals.data_ptr(), self.n_live, self.n_cells, self.cells_per_release,
            self._footprint.data_ptr(), self.comp.data_ptr() if self.comp is not None else None,
            self._release_total.data_ptr(), self.workspace.data_ptr(), self.workspace.numel(), stream))

    def _position_means(self, stream: int, row: bool = False) -> None:
        """Mean lon/lat/z over all particles of the batch into pos_means[step]
        (the step is read from the device cursor inside a captured body)."""

        n_rows = self.pos_means.shape[0]
        out = self.pos_means.data_ptr() if row else self.pos_means.data_ptr() + 24 * self.step_i
        L.check(L.lib().glide_particle_means(
            C.byref(self._parts[self.cur]), self.row.data_ptr() if row else None, n_rows if row else 0, out,
            self.means_ws.data_ptr(), self.means_ws.numel(), stream))

    def _wind_means(self, h: int, stream: int) -> None:
        if h in self.hour_wind:
            return
        out3, _, _ = self.met.resampled(h, stream)
        m = torch.empty(3, dtype=torch.float64, device=self.device)
        nlev = self.met.nz * self.met.ny * self.met.nx
        L.check(L.lib().glide_field_means(out3.data_ptr(), nlev, 3, m.data_ptr(), self.means_ws.data_ptr(),
                                          self.means_ws.numel(), stream))
        self.hour_wind[h] = m

    def trajectory_rows(self, *, batch_idx: int = 0, step_offset: int = 0,
                        epoch_iso: str = "2024-01-01T00:00:00") -> list[dict]:
        """Per-step diagnostics rows of the steps taken, with the reference's
        columns (main.py:1614-1643): one device-to-host read at the end."""

        if not self.diagnostics:
            raise ValueError("engine built without diagnostics=True")
        torch.cuda.synchronize()
        k = self.step_i
        pos = self.pos_means.cpu().numpy()[:k]
        act = self.active_steps.cpu().numpy()[:k]
        esc = self.escaped.cpu().numpy()[:k]
        alive = self.n - np.cumsum(esc)
        wind = {h: m.cpu().numpy() for h, m in self.hour_wind.items()}
        epoch = np.datetime64(epoch_iso, "s")
        rows = []
        for i in range(k):
            st = self.sched[i]
            a = float(st.alpha)
            w = wind[st.bracket_hour] * (1.0 - a) + wind[st.bracket_hour + 1] * a
            rows.append({
                "step": step_offset + i + 1, "batch_idx": batch_idx,
                "time_hour_end": str(epoch + np.timedelta64(int(st.cursor_s), "s")) + "+00:00",
                "mean_lon": float(pos[i, 0]), "mean_lat": float(pos[i, 1]), "mean_alt_agl_m": float(pos[i, 2]),
                "u_mean_m_s": float(w[0]), "v_mean_m_s": float(w[1]), "w_mean_m_s": float(w[2]),
                "active_particles": int(act[i]), "escaped_this_step": int(esc[i]),
                "alive_particles": int(alive[i]),
            })
        return rows

    def endpoints(self) -> dict:
        """Particle endpoints (lon, lat, alt, weight, release_idx) in particle-id
        order, for the endpoint table."""

        torch.cuda.synchronize()
        b = self.bufs[self.cur]
        uid = b["uid"].cpu().numpy().view(np.uint32)
        order = np.argsort(uid, kind="stable")
        lon, lat = self._get_positions(b, order)
        return {"lon": np.asarray(lon, np.float32), "lat": np.asarray(lat, np.float32),
                "alt": b["z"].cpu().numpy()[order], "weight": b["weight"].cpu().numpy()[order],
                "release_idx": b["rel_idx"].cpu().numpy()[order].astype(np.int64)}

    def footprint_numpy(self) -> np.ndarray:
        """The footprint as (R, T, Z, Y, X) float32 on the host (the
        compensation terms added when the footprint is compensated; with
        per-release slabs, each release's bins past its own length are 0)."""

        self.join()  # after every step's accumulation (fixed mode: the float footprint made)
        torch.cuda.synchronize()
        g = self.grid_cfg
        f = self.footprint if self.comp is None else self.footprint + self.comp
        if not self.slabs:
            return f.cpu().numpy().reshape(g["n_r"], g["n_t"], g["n_z"], g["n_y"], g["n_x"])
        out = np.zeros((g["n_r"], g["n_t"], g["n_z"], g["n_y"], g["n_x"]), np.float32)
        host = f.cpu().numpy()
        for r in range(g["n_r"]):
            out[r, : self.slab_nt_np[r]] = self._slab(host, r)
        return out

    def _slab(self, flat, r: int):
        g = self.grid_cfg
        o = int(self.slab_off_np[r]) if self.slabs else r * self.cells_per_release
        nt = int(self.slab_nt_np[r]) if self.slabs else g["n_t"]
        return flat[o: o + nt * self.cells_per_bin].reshape(nt, g["n_z"], g["n_y"], g["n_x"])

    def release_footprint(self, r: int) -> torch.Tensor:
        """Release ``r``'s footprint [T_r, Z, Y, X] float32 on the device (T_r its
        own hourly bins with per-release slabs, else n_t)."""

        return self._slab(self.footprint, r)

    def write_outputs(self, out_dir, *, epoch_iso: str = "2024-01-01T00:00:00", release_start: int = 0,
                      store=None) -> dict:
        """Write the batch's outputs as the reference does at batch end
        (main.py:1648-1692): its footprint slabs into the Zarr store (created
        here for the batch's releases unless ``store`` is given, then written at
        ``release_start``), endpoints and, with diagnostics, the trajectory
        table as Parquet. Returns the paths."""

        from pathlib import Path

        from .output import (ZarrFootprintStore, footprint_coords, write_particles_parquet,
                             write_trajectory_parquet)

        out = Path(out_dir)
        fp = self.footprint_numpy()
        if store is None:
            store = ZarrFootprintStore(out / "footprint.zarr")
            store.create(fp.shape, coords=footprint_coords(self.grid_cfg, self.cfg["releases"], length_s=self.length_s,
                                                           epoch_iso=epoch_iso))
        store.write_region(release_start, fp)
        ep = self.endpoints()
        write_particles_parquet(out / "endpoints.parquet", ep["lon"], ep["lat"], ep["alt"], ep["weight"],
                                ep["release_idx"] + release_start)
        paths = {"footprint": str(store.path), "endpoints": str(out / "endpoints.parquet")}
        if self.diagnostics:
            write_trajectory_parquet(out / "trajectory.parquet", int(self.cfg["dt"]),
                                     self.trajectory_rows(epoch_iso=epoch_iso))
            paths["trajectory"] = str(out / "trajectory.parquet")
        return paths

    def close(self) -> None:
        """Free the captured graphs."""

        for h in self._graphs.values():
            L.lib().glide_graph_destroy(h)
        self._graphs.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run_step(self, stream: int | None = None) -> None:
        """Advance one step of the schedule."""

        lib = L.lib()
        if stream is None:
            stream = L.stream_ptr()
        st = self.sched[self.step_i]
        if self.fixed_mode:
            self._fixed_stale = True  # this step adds to the fixed-point sums
        new_bracket = st.bracket_hour != self.last_bracket
        if self.compact and new_bracket:
            self.reorder(st, stream, count_live=True)
        elif self.reorder_every and self.step_i > 0 and self.step_i % self.reorder_every == 0:
            self.reorder(st, stream)
        self.met.build_window(st.bracket_hour, stream)
        if self.diagnostics and new_bracket:
            self._wind_means(st.bracket_hour, stream)
            self._wind_means(st.bracket_hour + 1, stream)
        convect = new_bracket and self.conv_opts is not None
        ev = self.step_events.get(self.step_i) if self.step_events is not None else None
        if self.use_graph and not convect and ev is None:
            self._wait_acc(stream)  # the captured body emits into buffer 0 and accumulates in line
            L.check(lib.glide_graph_launch(self._graph(1, stream), stream))
            self.graph_launches += 1
            self.last_bracket = st.bracket_hour
            self.step_i += 1
            return
        flags = self.h.flags() | (0 if convect else L.F_EMIT)
        p = self.step_params(st, flags)
        buf = self.step_i & 1 if self.overlap else 0
        if self.overlap and self.acc_pending[buf]:
            # the accumulation of two steps ago still reads this buffer
            self._stream_obj(stream).wait_event(self.ev_acc[buf])
            self.acc_pending[buf] = False
        em = self.emit(st, buf)
        if ev is not None:
            ev[0].record(self._torch_stream)
        L.check(lib.glide_step(C.byref(p), C.byref(self.parts), C.byref(self.met.grid),
                               C.byref(self.met.window), C.byref(self.fgs[buf]), C.byref(em), stream))
        if ev is not None:
            ev[1].record(self._torch_stream)
        if convect:
            s3, _, phpa = self.met.resampled(st.bracket_hour, stream)
            e3, _, _ = self.met.resampled(st.bracket_hour + 1, stream)
            co = self.conv_opts
            nlev = self.met.nz * self.met.ny * self.met.nx
            if co.parcel_lift:
                conv = self.conv_lift
                L.check(lib.glide_column_cape(
                    s3.data_ptr() + 3 * nlev * 4, e3.data_ptr() + 3 * nlev * 4, s3.data_ptr() + 4 * nlev * 4,
                    e3.data_ptr() + 4 * nlev * 4, float(st.alpha), phpa.data_ptr(), self.met.levels_f.data_ptr(),
                    self.met.nz, self.met.ny, self.met.nx, None, None, None, None, conv.data_ptr(), stream))
            else:
                conv = self.met.source(st.bracket_hour, stream)["conv"]
            L.check(lib.glide_convect_columns(
                conv.data_ptr(), s3.data_ptr() + 3 * nlev * 4, e3.data_ptr() + 3 * nlev * 4,
                float(st.alpha), phpa.data_ptr(), self.met.levels_f.data_ptr(), self.met.nz,
                self.met.ny, self.met.nx, co.closure_c, co.trigger_dtv_k, co.min_cape_j_kg,
                co.min_cloud_depth_m, self.i_lcl.data_ptr(), self.i_lnb.data_ptr(),
                self.m_b.data_ptr(), self.layer_mass.data_ptr(), stream))
            pc = self.step_params(st, self.h.flags() | L.F_EMIT)
            L.check(lib.glide_convect(C.byref(pc), C.byref(self.parts), C.byref(self.met.grid),
                                      self.i_lcl.data_ptr(), self.i_lnb.data_ptr(), self.m_b.data_ptr(),
                                      self.layer_mass.data_ptr(), self.colsum_err.data_ptr(),
                                      self.conv_cloud.data_ptr(), C.byref(self.fgs[buf]), C.byref(em), stream))
        if self.overlap:
            self.ev_step[buf].record(self._stream_obj(stream))
            self.acc_stream.torch.wait_event(self.ev_step[buf])
            self._accumulate(self.acc_stream.ptr, buf)
            self.ev_acc[buf].record(self.acc_stream.torch)
            self.acc_pending[buf] = True
        else:
            self._accumulate(stream)
        if self.diagnostics:
            self._position_means(stream)
        self.last_bracket = st.bracket_hour
        self.step_i += 1
        if self.use_graph:
            L.check(lib.glide_row_set(self.row.data_ptr(), self.step_i, stream))

    def time_step_kernel(self, n_steps: int, stream, every: int = 1) -> dict:
        """Record HIP events around every ``every``-th of the next ``n_steps``
        step-kernel launches on ``stream`` (a torch stream, the one the kernels
        are launched on). Those steps launch eagerly. Returns {offset from the
        current step: (start, end)}."""

        self._torch_stream = stream
        evs = {i: (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for i in range(0, n_steps, max(1, every))}
        self.step_events = {self.step_i + i: e for i, e in evs.items()}
        return evs

    # steps per captured graph: a run of replayable steps is launched as
    # graphs of these sizes, largest first (one launch per graph, not per step)
    GRAPH_STEPS = (16, 8, 4, 2, 1)

    def _graph(self, steps: int, stream: int) -> int:
        """The captured body of ``steps`` steps for the current buffer and live count."""

        key = (self.cur, self.n_live, steps)
        g = self._graphs.get(key)
        if g is None:
            for k in [k for k in self._graphs if k[1] != self.n_live]:  # a shrunken live set
                L.lib().glide_graph_destroy(self._graphs.pop(k))
            g = self._graphs[key] = self._capture(stream, steps)
        return g

    def _replayable_run(self, end: int) -> int:
        """Steps from step_i (before ``end``) that replay the captured body
        back to back with nothing between them: the bracket of the step
        before (no window build, compaction or convection), no re-binning,
        no timing events."""

        k = 0
        for j in range(self.step_i, end):
            if self.sched[j].bracket_hour != self.last_bracket:
                break
            if self.reorder_every and j > 0 and j % self.reorder_every == 0:
                break
            if self.step_events is not None and j in self.step_events:
                break
            k += 1
        return k

    def advance(self, n_steps: int, stream: int | None = None) -> None:
        """Run the next ``n_steps`` steps of the schedule (fewer at its end).
        With ``use_graph``, runs of replayable steps go to the device as
        multi-step graphs; the result is the same as stepping one by one."""

        if stream is None:
            stream = L.stream_ptr()
        end = min(len(self.sched), self.step_i + n_steps)
        lib = L.lib()
        while self.step_i < end:
            k = self._replayable_run(end) if self.use_graph else 0
            if k < 2:
                self.run_step(stream)
                continue
            self.met.build_window(self.last_bracket, stream)  # (built already: a shared DeviceMet is re-checked)
            if self.fixed_mode:
                self._fixed_stale = True  # these steps add to the fixed-point sums
            self._wait_acc(stream)
            for size in self.GRAPH_STEPS:
                while k >= size:
                    L.check(lib.glide_graph_launch(self._graph(size, stream), stream))
                    self.graph_launches += 1
                    self.step_i += size
                    k -= size

    def run(self, n_steps: int | None = None) -> None:
        self.advance(len(self.sched) if n_steps is None else n_steps)
        self.join()

    def active_count(self, st) -> int:
        """Particles in their release window at step ``st`` (before escapes)."""

        co = np.float32(st.cursor_off)
        if self.release_terminus:
            act = (self.ra.r_off >= co) & (co > self.ra.te_off)
        else:
            act = (self.ra.r_off >= co) & (self.ra.lo_off <= co)
        return int(act.sum())

    def state_numpy(self) -> dict:
        """Particle state on the host, in particle-id (release) order."""

        torch.cuda.synchronize()
        b = self.bufs[self.cur]
        uid = b["uid"].cpu().numpy().view(np.uint32)
        order = np.argsort(uid, kind="stable")
        out = {k: b[k].cpu().numpy()[order] for k in ("z", "up", "vp", "wp", "um", "vm")}
        # absolute positions (float64 exactly when the cell layout holds them)
        out["lon"], out["lat"] = self._get_positions(b, order)
        out["alive"] = b["alive"].cpu().numpy().astype(bool)[order]
        out["uid"] = uid[order]
        return out

    def set_state(self, S: dict) -> None:
        """Load per-particle state given in particle-id order (tests)."""

        self.cur = 0
        b = self.bufs[0]
        ra = self.ra
        self.n_live = self.n
        self.horizon = -1e30
        for k, v in (("weight", ra.weight), ("r_off", ra.r_off), ("we_off", ra.we_off), ("rel_idx", ra.rel_idx),
                     ("uid", ra.uid.view(np.int32)),
                     ("lo_off", ra.te_off if self.release_terminus else ra.lo_off)):
            b[k].copy_(torch.from_numpy(np.ascontiguousarray(v)))
        for k in ("z", "up", "vp", "wp", "um", "vm"):
            b[k].copy_(torch.from_numpy(np.asarray(S[k], np.float32)))
        self._put_positions(b, S["lon"], S["lat"])
        b["alive"].fill_(1)
