"""Footprint and particle outputs in the reference's on-disk formats.

* Footprints: a Zarr v2 store as the reference's OutputWriter writes it through
  xarray (output_writer.py:114-243): one ``footprint`` float32 array with dims
  (release, time_ago, z_bin, latitude, longitude), chunked one release per
  chunk so each finished release is written on its own, and the coordinates of
  main._build_footprint_dataset_metadata (main.py:394-445). The store is
  written directly (JSON metadata, little-endian C-order chunks, raw or
  zlib-deflated) in the encoding xarray reads back: ``_ARRAY_DIMENSIONS`` on
  every array, non-index coordinates named in the data variable's
  ``coordinates`` attribute, datetimes as int64 nanoseconds with CF units, NaN
  as the footprint's fill value (so a release never written reads as NaN, as
  with the reference's lazily created store) and no fill value on coordinates
  (a fill value is a ``_FillValue`` to xarray and would mask real zeros), plus
  consolidated metadata.
* Particle endpoints and trajectory diagnostics: Parquet tables with the
  reference's columns (output_writer.py:54-101).

``OutputWriter`` keeps the reference's method names and arguments; arrays may
be numpy arrays or torch tensors.
"""

from __future__ import annotations

import json
import os
import zlib
from pathlib import Path
from typing import Any, Mapping, Sequence

import numpy as np

FOOTPRINT_DIMS = ("release", "time_ago", "z_bin", "latitude", "longitude")


def _np(a) -> np.ndarray:
    if hasattr(a, "detach"):  # torch tensor
        return a.detach().to(device="cpu").numpy()
    return np.asarray(a)


def _write_json(path: Path, obj) -> None:
    path.write_text(json.dumps(obj, indent=2, sort_keys=True, allow_nan=False))


def _array_meta(shape, chunks, dtype: str, fill, compressor=None) -> dict:
    return {"zarr_format": 2, "shape": list(shape), "chunks": list(chunks), "dtype": dtype,
            "compressor": compressor, "fill_value": fill, "order": "C", "filters": None, "dimension_separator": "."}


def _compressor(name: str | None, level: int = 1):
    """Zarr v2 compressor metadata: None, or zlib (numcodecs' Zlib codec id)."""

    if name is None:
        return None
    if name != "zlib":
        raise ValueError(f"unsupported compressor {name!r} (None or 'zlib')")
    return {"id": "zlib", "level": int(level)}


def _coord_dims(name: str, value) -> tuple[tuple[str, ...], Any]:
    """(dims, values) of a coordinate given as (dim, values) or bare values (a
    dimension coordinate named like its dim)."""

    if isinstance(value, tuple):
        dims, vals = value
        return ((dims,) if isinstance(dims, str) else tuple(dims)), vals
    return (name,), value


def footprint_coords(grid: Mapping[str, Any], releases: Sequence, *, length_s: float,
                     epoch_iso: str = "2024-01-01T00:00:00") -> dict:
    """Coordinates of the footprint store (main.py:394-445): per-release time,
    duration and position along ``release`` (and ``site`` when any release
    carries a label); time-ago bins of max(1, length / 3600 / n_t) hours; z-bin
    centres and edges; cell-centre latitudes/longitudes and their edges.
    ``grid`` holds the gridder bounds (lon_min/max, lat_min/max, n_x, n_y, n_t,
    z_edges); release times are seconds after ``epoch_iso``."""

    lon_e = np.linspace(grid["lon_min"], grid["lon_max"], grid["n_x"] + 1, dtype=np.float64)
    lat_e = np.linspace(grid["lat_min"], grid["lat_max"], grid["n_y"] + 1, dtype=np.float64)
    z_e = np.asarray(grid["z_edges"], dtype=np.float64)
    n_t = int(grid["n_t"])
    per_bin = max(1.0, float(length_s) / 3600.0 / max(1, n_t))
    t0 = np.arange(n_t, dtype=np.float64) * per_bin
    epoch = np.datetime64(epoch_iso, "ns")
    times = np.array([epoch + np.timedelta64(int(r.release_time_s), "s") for r in releases], dtype="datetime64[ns]")
    coords = {
        "release_time": ("release", times),
        "release_duration_seconds": ("release", np.array([float(r.duration_s) for r in releases])),
        "release_lon": ("release", np.array([float(r.lon) for r in releases])),
        "release_lat": ("release", np.array([float(r.lat) for r in releases])),
        "release_alt_agl_m": ("release", np.array([float(r.alt_agl_m) for r in releases])),
        "time_ago": np.arange(n_t, dtype=np.int64),
        "time_ago_start_hours": ("time_ago", t0),
        "time_ago_end_hours": ("time_ago", t0 + per_bin),
        "z_bin": 0.5 * (z_e[:-1] + z_e[1:]),
        "z_bottom_m": ("z_bin", z_e[:-1]),
        "z_top_m": ("z_bin", z_e[1:]),
        "latitude": 0.5 * (lat_e[:-1] + lat_e[1:]),
        "longitude": 0.5 * (lon_e[:-1] + lon_e[1:]),
        "latitude_edge": ("latitude_edge", lat_e),
        "longitude_edge": ("longitude_edge", lon_e),
    }
    if any(getattr(r, "label", None) is not None for r in releases):
        coords["site"] = ("release", np.array([getattr(r, "label", None) or "" for r in releases]))
    return coords


class ZarrFootprintStore:
    """A 5-D footprint store, created for the whole schedule and filled one
    release (or a block of releases) at a time."""

    def __init__(self, path: str | os.PathLike):
        self.path = Path(path)

    def create(self, shape: tuple[int, ...], *, coords: Mapping[str, Any] | None = None,
               attrs: Mapping[str, Any] | None = None, dims: tuple[str, ...] = FOOTPRINT_DIMS,
               compressor: str | None = None, level: int = 1) -> None:
        """Write the store's metadata and coordinates. ``compressor="zlib"``
        deflates the footprint chunks (sparse footprints shrink a hundredfold);
        default: raw chunks."""
        if len(shape) != 5 or len(dims) != 5:
            raise ValueError("footprint store shape must be (R, T, Z, Y, X)")
        shape = tuple(int(s) for s in shape)
        self.path.mkdir(parents=True, exist_ok=True)
        for old in self.path.glob("footprint/[0-9]*"):  # mode="w": a new store
            old.unlink()
        meta: dict[str, Any] = {".zgroup": {"zarr_format": 2}, ".zattrs": dict(attrs or {})}
        aux = []  # non-index coordinates
        for name, value in (coords or {}).items():
            cdims, vals = _coord_dims(name, value)
            v = _np(vals)
            if v.ndim != len(cdims):
                raise ValueError(f"coordinate {name}: {v.ndim}-d values for dims {cdims}")
            for d, n in zip(cdims, v.shape):
                if d in dims and n != shape[dims.index(d)]:
                    raise ValueError(f"coordinate {name}: length {n} on {d}, the store has {shape[dims.index(d)]}")
            if cdims != (name,):
                aux.append(name)
            ca: dict[str, Any] = {"_ARRAY_DIMENSIONS": list(cdims)}
            if np.issubdtype(v.dtype, np.datetime64):
                v = v.astype("datetime64[ns]").astype(np.int64)
                ca.update(units="nanoseconds since 1970-01-01", calendar="proleptic_gregorian")
            elif v.dtype.kind == "S":
                v = v.astype(str)
            v = np.ascontiguousarray(v.astype(v.dtype.newbyteorder("<")) if v.dtype.byteorder == ">" else v)
            cm = _array_meta(v.shape, v.shape, v.dtype.str, None)
            cdir = self.path / name
            cdir.mkdir(exist_ok=True)
            _write_json(cdir / ".zarray", cm)
            _write_json(cdir / ".zattrs", ca)
            (cdir / ".".join(["0"] * max(1, v.ndim))).write_bytes(v.tobytes())
            meta[f"{name}/.zarray"], meta[f"{name}/.zattrs"] = cm, ca
        fp_dir = self.path / "footprint"
        fp_dir.mkdir(exist_ok=True)
        am = _array_meta(shape, (1,) + shape[1:], "<f4", "NaN", _compressor(compressor, level))
        aa: dict[str, Any] = {"_ARRAY_DIMENSIONS": list(dims)}
        if aux:
            aa["coordinates"] = " ".join(sorted(aux))
        _write_json(fp_dir / ".zarray", am)
        _write_json(fp_dir / ".zattrs", aa)
        meta["footprint/.zarray"], meta["footprint/.zattrs"] = am, aa
        _write_json(self.path / ".zgroup", meta[".zgroup"])
        _write_json(self.path / ".zattrs", meta[".zattrs"])
        _write_json(self.path / ".zmetadata", {"zarr_consolidated_format": 1, "metadata": meta})

    def _meta(self) -> dict:
        return json.loads((self.path / "footprint" / ".zarray").read_text())

    def shape(self) -> tuple[int, ...]:
        return tuple(self._meta()["shape"])

    def write_region(self, release_start: int, footprint) -> None:
        """Write ``footprint`` (n, T, Z, Y, X) into releases
        [release_start, release_start + n), one chunk file per release."""

        arr = _np(footprint)
        if arr.ndim != 5:
            raise ValueError("footprint must have shape (R, T, Z, Y, X)")
        arr = arr.astype("<f4", copy=False)
        meta = self._meta()
        shape = tuple(meta["shape"])
        comp = meta.get("compressor")
        if tuple(arr.shape[1:]) != tuple(shape[1:]):
            raise ValueError(f"footprint block {arr.shape[1:]} does not match the store's {tuple(shape[1:])}")
        if release_start < 0 or release_start + arr.shape[0] > shape[0]:
            raise ValueError(f"releases [{release_start}, {release_start + arr.shape[0]}) outside [0, {shape[0]})")
        for j in range(arr.shape[0]):
            name = ".".join([str(release_start + j)] + ["0"] * 4)
            tmp = self.path / "footprint" / (name + ".tmp")
            raw = np.ascontiguousarray(arr[j]).tobytes()
            tmp.write_bytes(zlib.compress(raw, comp["level"]) if comp else raw)
            os.replace(tmp, self.path / "footprint" / name)

    def read(self, release: int) -> np.ndarray:
        """One release's footprint (T, Z, Y, X); NaN if it was never written."""

        meta = self._meta()
        shape = meta["shape"]
        f = self.path / "footprint" / ".".join([str(release)] + ["0"] * 4)
        if not f.exists():
            return np.full(shape[1:], np.nan, np.float32)
        raw = f.read_bytes()
        if meta.get("compressor"):
            raw = zlib.decompress(raw)
        return np.frombuffer(raw, dtype="<f4").reshape(shape[1:]).copy()

    def read_coord(self, name: str) -> np.ndarray:
        m = json.loads((self.path / name / ".zarray").read_text())
        raw = (self.path / name / ".".join(["0"] * max(1, len(m["shape"])))).read_bytes()
        return np.frombuffer(raw, dtype=np.dtype(m["dtype"])).reshape(m["shape"]).copy()


def _write_parquet(df, path) -> None:
    Path(path).parent.mkdir(parents=True, exist_ok=True)
    tmp = Path(str(path) + ".tmp")
    df.to_parquet(tmp, index=False)
    os.replace(tmp, path)


def write_particles_parquet(path: str | os.PathLike, lon, lat, alt, weight, release_idx=None) -> None:
    """Particle endpoints, columns [lon, lat, alt, weight] (+ release_idx int64
    for multi-release runs), as output_writer.py:54-85."""

    import pandas as pd

    cols = {"lon": _np(lon).astype(np.float32), "lat": _np(lat).astype(np.float32),
            "alt": _np(alt).astype(np.float32), "weight": _np(weight).astype(np.float32)}
    n = cols["lon"].shape[0]
    if any(v.shape != (n,) for v in cols.values()):
        raise ValueError("particles must have shape (N, 4)")
    df = pd.DataFrame(cols)
    if release_idx is not None:
        ridx = _np(release_idx).astype(np.int64)
        if ridx.shape != (n,):
            raise ValueError(f"release_idx shape {ridx.shape} does not match particles ({n},)")
        df["release_idx"] = ridx
    _write_parquet(df, path)


def write_trajectory_parquet(path: str | os.PathLike, step_seconds: int, rows: Sequence[Mapping[str, Any]]) -> None:
    """Per-step diagnostics rows; adds elapsed_seconds = step * step_seconds
    when a ``step`` column is present (output_writer.py:87-101)."""

    import pandas as pd

    df = pd.DataFrame(list(rows))
    if "step" in df.columns and "elapsed_seconds" not in df.columns:
        df["elapsed_seconds"] = df["step"].astype(np.int64) * int(step_seconds)
    _write_parquet(df, path)


class OutputWriter:
    """The reference's writer interface (output_writer.py:51-243) over local
    paths."""

    FOOTPRINT_DIMS = FOOTPRINT_DIMS

    def write_particles_parquet(self, path: str, particles, *, release_idx=None) -> None:
        arr = _np(particles)
        if arr.ndim != 2 or arr.shape[1] != 4:
            raise ValueError("particles must have shape (N, 4)")
        write_particles_parquet(path, arr[:, 0], arr[:, 1], arr[:, 2], arr[:, 3], release_idx)

    def write_trajectory_parquet(self, path: str, step_seconds: int, rows) -> None:
        write_trajectory_parquet(path, step_seconds, rows)

    def write_metadata_json(self, path: str, payload: Mapping[str, Any]) -> None:
        Path(path).parent.mkdir(parents=True, exist_ok=True)
        Path(path).write_text(json.dumps(payload, indent=2, sort_keys=True))

    def write_footprint_zarr(self, path: str, footprint, *, dims=FOOTPRINT_DIMS, coords=None, attrs=None) -> None:
        arr = _np(footprint)
        if arr.ndim != 5:
            raise ValueError("footprint must have shape (R, T, Z, Y, X)")
        store = ZarrFootprintStore(path)
        store.create(arr.shape, coords=coords, attrs=attrs, dims=tuple(dims))
        store.write_region(0, arr)

    def create_footprint_store(self, path: str, *, shape, dims=FOOTPRINT_DIMS, coords=None, attrs=None) -> None:
        ZarrFootprintStore(path).create(tuple(shape), coords=coords, attrs=attrs, dims=tuple(dims))

    def write_footprint_region(self, path: str, footprint, *, release_start: int, release_stop: int,
                               dims=FOOTPRINT_DIMS) -> None:
        arr = _np(footprint)
        if arr.ndim != 5:
            raise ValueError("footprint must have shape (R, T, Z, Y, X)")
        if arr.shape[0] != release_stop - release_start:
            raise ValueError(f"footprint leading dim {arr.shape[0]} != region size {release_stop - release_start} "
                             f"(release_start={release_start}, release_stop={release_stop})")
        ZarrFootprintStore(path).write_region(release_start, arr)
