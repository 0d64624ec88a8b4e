"""FootprintGridder drop-in (reference src/lpdm/footprint_gridder.py:28-197).

Same constructor and ``accumulate`` signature and the same ``.tensor``
(``[n_releases, n_t, n_z, n_y, n_x]``). The index math runs in the HIP key
kernel (bit-exact with the reference's float32 ops) and the scatter is the
deterministic sort-by-cell segmented reduction, so the result is identical from
run to run (the reference's ``scatter_add_`` uses float atomics on the GPU).
"""

from __future__ import annotations

import ctypes as C
import os
from typing import Sequence

import numpy as np
import torch

from glide_amd import _lib as L


class FootprintGridder:
    MAX_FOOTPRINT_GIB: float = 32.0

    def __init__(self, lon_bounds: tuple[float, float], lat_bounds: tuple[float, float],
                 z_edges_m: Sequence[float], n_time_bins: int, n_y: int, n_x: int, *,
                 n_releases: int = 1, device: torch.device | str = "cuda",
                 dtype: torch.dtype = torch.float32, compensated: bool = False) -> None:
        ze = np.asarray(list(z_edges_m), dtype=float)
        if ze.ndim != 1 or ze.size < 2:
            raise ValueError("z_edges_m must be a 1D sequence with at least 2 values")
        if not np.all(np.diff(ze) > 0):
            raise ValueError("z_edges_m must be strictly ascending")
        n_z = int(ze.size - 1)
        if min(n_time_bins, n_y, n_x, n_releases) <= 0 or n_z <= 0:
            raise ValueError("All footprint dimensions must be > 0")
        if dtype != torch.float32:
            raise ValueError("the device gridder accumulates float32 residence time")
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise ValueError("FootprintGridder runs on the GPU (device must be cuda)")
        self.dtype = dtype
        n_el = int(n_releases) * int(n_time_bins) * n_z * int(n_y) * int(n_x)
        cap_gib = float(os.environ.get("LPDM_FOOTPRINT_MAX_GIB") or self.MAX_FOOTPRINT_GIB)
        if n_el * 4 > int(cap_gib * 1024**3):
            raise MemoryError(
                f"Footprint tensor would require {n_el * 4 / 1024**3:.2f} GiB (shape (n_releases="
                f"{n_releases}, n_time_bins={n_time_bins}, n_z={n_z}, n_y={n_y}, n_x={n_x}), float32). "
                f"Cap is {cap_gib:.2f} GiB; raise it via LPDM_FOOTPRINT_MAX_GIB or reduce the grid.")
        if n_el >= 0xFFFFFFFF:
            raise ValueError("footprint has too many cells for 32-bit cell keys")
        self.tensor = torch.zeros((n_releases, n_time_bins, n_z, n_y, n_x), device=self.device, dtype=dtype)
        self.comp = torch.zeros_like(self.tensor) if compensated else None
        self.lon_min, self.lon_max = (float(v) for v in lon_bounds)
        self.lat_min, self.lat_max = (float(v) for v in lat_bounds)
        self.z_edges = torch.tensor(ze, dtype=torch.float32, device=self.device)
        self.z_min, self.z_max = float(ze[0]), float(ze[-1])
        self.n_releases, self.n_t, self.n_z, self.n_y, self.n_x = int(n_releases), int(n_time_bins), n_z, int(n_y), int(n_x)
        self.n_cells = n_el
        self.release_total = torch.zeros(self.n_releases, dtype=torch.float64, device=self.device)
        self._grid = L.GlideFootprintGrid(
            lon_min=self.lon_min, lon_max=self.lon_max, lat_min=self.lat_min, lat_max=self.lat_max,
            z_edges=self.z_edges.data_ptr(), n_r=self.n_releases, n_t=self.n_t, n_z=n_z, n_y=self.n_y,
            n_x=self.n_x, reserved=0)
        self._scratch = {}

    def _buffers(self, n: int):
        b = self._scratch.get(n)
        if b is None:
            ws = L.lib().glide_footprint_workspace_bytes(n)
            b = (torch.empty(n, dtype=torch.int32, device=self.device),
                 torch.empty(n, dtype=torch.float32, device=self.device),
                 torch.empty(ws, dtype=torch.uint8, device=self.device))
            self._scratch = {n: b}
        return b

    def accumulate(self, particles: torch.Tensor, active_mask: torch.Tensor, weights: torch.Tensor,
                   t_idx: torch.Tensor, release_idx: torch.Tensor, dt_seconds: float) -> None:
        """Add ``weights * dt_seconds`` of every active, in-range particle to its
        cell (reference footprint_gridder.py:117-197)."""

        if particles.ndim != 2 or particles.shape[1] < 3:
            raise ValueError("particles must be [N, >=3] (lon, lat, alt)")
        n = int(particles.shape[0])
        if n == 0:
            return
        if particles.dtype != torch.float32 or particles.stride(1) != 1:
            particles = particles[:, :3].to(torch.float32).contiguous()
        for name, t in (("active_mask", active_mask), ("weights", weights), ("t_idx", t_idx),
                        ("release_idx", release_idx)):
            if t.shape[0] != n:
                raise ValueError(f"{name} must have shape (N,) with N = {n}")
        am = active_mask.to(device=self.device, dtype=torch.bool).contiguous()
        w = weights.to(device=self.device, dtype=torch.float32).contiguous()
        ti = t_idx.to(device=self.device, dtype=torch.int64).contiguous()
        ri = release_idx.to(device=self.device, dtype=torch.int64).contiguous()
        keys, vals, ws = self._buffers(n)
        base = particles.data_ptr()
        lib = L.lib()
        s = L.stream_ptr()
        L.check(lib.glide_footprint_keys(n, particles.stride(0), base, base + 4, base + 8, am.data_ptr(),
                                         w.data_ptr(), ti.data_ptr(), ri.data_ptr(), float(dt_seconds),
                                         C.byref(self._grid), keys.data_ptr(), vals.data_ptr(), s))
        L.check(lib.glide_footprint_accumulate(
            keys.data_ptr(), vals.data_ptr(), n, self.n_cells, self.n_cells // self.n_releases,
            self.tensor.data_ptr(), self.comp.data_ptr() if self.comp is not None else None,
            self.release_total.data_ptr(), ws.data_ptr(), ws.numel(), s))

    def compensated_tensor(self) -> torch.Tensor:
        """``tensor + comp`` when built with ``compensated=True`` (else ``tensor``)."""

        return self.tensor if self.comp is None else self.tensor + self.comp
