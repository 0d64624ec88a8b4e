"""Met containers and the hourly resample, mirroring the reference's reader
interface (src/lpdm/met_reader.py:85-177) for the terrain-following path.

``resample_hour_to_agl`` is the device replacement of
``ArcoEra5ZarrReader._resample_hour_to_agl`` (met_reader.py:641-693): one
column kernel does the bracketing, the lerp, the ``w`` slope correction and the
bbox-mean level pressure.
"""

from __future__ import annotations

from dataclasses import dataclass
from datetime import datetime
from typing import Mapping

import numpy as np
import torch

from glide_amd import _lib as L


@dataclass(frozen=True)
class MetFieldMetadata:
    """reference met_reader.py:85-98."""

    lon: np.ndarray
    lat: np.ndarray
    level: np.ndarray
    pressure_level_hpa: np.ndarray
    time_start: datetime
    time_end: datetime
    variable_units: Mapping[str, str]


@dataclass(frozen=True)
class HourlyMetTensors:
    """Hour-bracketing tensors ``[C, Z, Y, X]`` (reference met_reader.py:101-153)."""

    hour_start: torch.Tensor
    hour_end: torch.Tensor
    metadata: MetFieldMetadata
    channel_names: tuple[str, ...]
    height_agl_m: torch.Tensor | None = None

    def _channel_index(self, name: str) -> int:
        try:
            return self.channel_names.index(name)
        except ValueError as exc:
            available = ", ".join(self.channel_names) or "<none>"
            raise KeyError(f"Channel {name!r} not present in met window. Available: {available}") from exc

    def channel(self, name: str) -> tuple[torch.Tensor, torch.Tensor]:
        i = self._channel_index(name)
        return self.hour_start[i], self.hour_end[i]

    def channels(self, *names: str) -> tuple[torch.Tensor, torch.Tensor]:
        idx = [self._channel_index(n) for n in names]
        return self.hour_start[idx], self.hour_end[idx]


def resample_hour_to_agl(src3: torch.Tensor, height_agl_p: torch.Tensor, terrain_m: torch.Tensor,
                         lat: np.ndarray, lon: np.ndarray, agl_levels: np.ndarray,
                         p_level_hpa: np.ndarray, *, uvw_index=(0, 1, 2), stream=None):
    """Resample one hour of 3-D channels ``src3 [C, Zp, Y, X]`` (device float32)
    from per-column heights ``height_agl_p [Zp, Y, X]`` (negative below ground)
    onto the ascending ``agl_levels``. Returns ``(dst3 [C, Za, Y, X],
    pressure_level_hpa [Za])`` on the device.

    Grid spacing must be uniform (ERA5 grids are); raises ``ValueError`` otherwise.
    """

    if src3.ndim != 4:
        raise ValueError(f"src3 must be [C, Zp, Y, X]; got {tuple(src3.shape)}")
    C_, Zp, Y, X = src3.shape
    if tuple(height_agl_p.shape) != (Zp, Y, X):
        raise ValueError(f"height_agl_p {tuple(height_agl_p.shape)} must match src3[1:] {(Zp, Y, X)}")
    agl = np.asarray(agl_levels, np.float64)
    if agl.ndim != 1 or not np.all(np.diff(agl) > 0):
        raise ValueError("agl_levels must be 1-D strictly ascending")
    lat = np.asarray(lat, np.float64)
    lon = np.asarray(lon, np.float64)
    dl = np.diff(lat)
    dn = np.diff(lon)
    if not (np.all(dl == dl[0]) and np.all(dn == dn[0])):
        raise ValueError("resample needs a uniformly spaced lat/lon grid")
    dev = src3.device
    mlon = torch.as_tensor(np.clip(111320.0 * np.cos(np.deg2rad(lat)), 1.0, None), dtype=torch.float64,
                           device=dev)
    agl_d = torch.as_tensor(agl, dtype=torch.float64, device=dev)
    p_d = torch.as_tensor(np.asarray(p_level_hpa, np.float64), device=dev)
    src3 = src3.to(torch.float32).contiguous()
    h = height_agl_p.to(device=dev, dtype=torch.float32).contiguous()
    terr = terrain_m.to(device=dev, dtype=torch.float32).contiguous()
    out = torch.empty((C_, agl.size, Y, X), dtype=torch.float32, device=dev)
    p_out = torch.empty(agl.size, dtype=torch.float64, device=dev)
    iu, iv, iw = uvw_index
    L.check(L.lib().glide_resample_hour(
        src3.data_ptr(), C_, Zp, Y, X, h.data_ptr(), agl_d.data_ptr(), agl.size, terr.data_ptr(),
        mlon.data_ptr(), float(dl[0]), float(dn[0]), p_d.data_ptr(), iu, iv, iw, out.data_ptr(),
        p_out.data_ptr(), L.stream_ptr(stream)))
    return out, p_out


def convert_source_units(src3: torch.Tensor, src2: torch.Tensor, geopotential: torch.Tensor | None,
                         geopotential_sfc: torch.Tensor | None, p_level_pa: torch.Tensor, *, w_ch: int = 2,
                         t_ch: int = 3, shf_ch: int = 3, shf_units: int = L.GLIDE_SHF_ACCUMULATED_J_M2,
                         accumulation_s: float = 3600.0, convert_w: bool = True, stream=None) -> torch.Tensor | None:
    """The reader's per-hour conversions of one raw hour on the device, in
    place (``glide_source_units``, replacing met_reader.py:1052-1078,
    1268-1300, 1353-1385): omega in ``src3[w_ch]`` to w (when ``convert_w``),
    ``src2[shf_ch]`` to upward W/m^2, and, when geopotential is given, the
    per-column heights above ground, returned as a new [Zp, Y, X] tensor."""

    C3, Zp, Y, X = src3.shape
    for t, name in ((src3, "src3"), (src2, "src2")):
        if t.dtype != torch.float32 or not t.is_contiguous():
            raise ValueError(f"{name} must be contiguous float32")
    p = p_level_pa.to(device=src3.device, dtype=torch.float64).contiguous()
    height = None
    if geopotential is not None:
        geopotential = geopotential.to(torch.float32).contiguous()
        geopotential_sfc = geopotential_sfc.to(torch.float32).contiguous()
        height = torch.empty((Zp, Y, X), dtype=torch.float32, device=src3.device)
    shf = src2[shf_ch]
    L.check(L.lib().glide_source_units(
        src3.data_ptr(), C3, Zp, Y, X, w_ch if convert_w else -1, t_ch, p.data_ptr(),
        geopotential.data_ptr() if geopotential is not None else None,
        geopotential_sfc.data_ptr() if geopotential is not None else None,
        height.data_ptr() if height is not None else None, shf.data_ptr(), int(shf_units), float(accumulation_s),
        L.stream_ptr(stream)))
    return height


def upload_source_hour(sh, device) -> dict:
    """Device arrays (src3, src2, height, p_hpa, conv) of one source hour on
    the current stream; an hour in the reader's raw units is converted on the
    device."""

    def dev(a, dtype=None):
        t = torch.from_numpy(np.ascontiguousarray(a))
        return (t.to(dtype) if dtype is not None else t).to(device)

    out = dict(src3=dev(sh.src3), src2=dev(sh.src2), p_hpa=dev(np.asarray(sh.p_level_pa) / 100.0, torch.float64),
               conv=dev(sh.conv))
    if getattr(sh, "units", "final") == "source":
        out["height"] = convert_source_units(out["src3"], out["src2"], dev(sh.geopotential),
                                             dev(sh.geopotential_sfc), dev(np.asarray(sh.p_level_pa), torch.float64))
    else:
        out["height"] = dev(sh.height_agl)
    return out
