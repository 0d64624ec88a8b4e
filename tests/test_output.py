"""CPU: the output formats (reference tests/test_output_writer.py).

xarray and zarr are not installed here, so the footprint store is read back by
a Zarr v2 decoder written in this file from the format's rules (array metadata
-> chunk grid -> chunk keys joined by the dimension separator, C-order
little-endian chunk bytes, the fill value where a chunk is missing) together
with the xarray conventions a reader relies on (_ARRAY_DIMENSIONS, the
``coordinates`` attribute, CF datetime units, _FillValue masking). The
reference's own assertions on the store, endpoints and trajectory tables are
restated on what that decoder returns."""

from __future__ import annotations

import itertools
import json
from pathlib import Path

import numpy as np
import pandas as pd
import pytest

from glide_amd.output import (FOOTPRINT_DIMS, OutputWriter, ZarrFootprintStore, footprint_coords,
                              write_trajectory_parquet)
from glide_amd.synthetic import Release


def _decode_array(root: Path, name: str):
    """(values, attrs) of one array of a Zarr v2 group, xarray-style: the
    fill value masks to NaN in float arrays, CF nanosecond datetimes decode."""

    meta = json.loads((root / name / ".zarray").read_text())
    attrs = json.loads((root / name / ".zattrs").read_text())
    assert meta["zarr_format"] == 2 and meta["order"] == "C"
    comp = meta["compressor"]
    assert comp is None or comp["id"] == "zlib"  # numcodecs Zlib: a zlib stream per chunk
    dt = np.dtype(meta["dtype"])
    shape, chunks = meta["shape"], meta["chunks"]
    fill = meta["fill_value"]
    if fill == "NaN":
        fill = np.nan
    out = np.full(shape, np.nan if fill is None and dt.kind == "f" else (fill if fill is not None else 0), dtype=dt)
    grid = [range((s + c - 1) // c) for s, c in zip(shape, chunks)] or [range(1)]
    sep = meta.get("dimension_separator", ".")
    for idx in itertools.product(*grid):
        f = root / name / sep.join(str(i) for i in idx)
        if not f.exists():
            continue
        raw = f.read_bytes()
        if comp is not None:
            import zlib

            raw = zlib.decompress(raw)
        block = np.frombuffer(raw, dtype=dt).reshape(chunks)
        sl = tuple(slice(i * c, min((i + 1) * c, s)) for i, c, s in zip(idx, chunks, shape))
        out[sl] = block[tuple(slice(0, x.stop - x.start) for x in sl)]
    if fill is not None and dt.kind == "f" and not (isinstance(fill, float) and np.isnan(fill)):
        out = np.where(out == fill, np.nan, out)  # xarray: _FillValue -> NaN
    if attrs.get("units", "").startswith("nanoseconds since 1970-01-01"):
        out = out.astype("datetime64[ns]")
    return out, attrs


def _open(root: Path) -> dict:
    cons = json.loads((root / ".zmetadata").read_text())["metadata"]
    names = sorted({k.split("/")[0] for k in cons if "/" in k})
    for n in names:  # consolidated metadata agrees with the files
        assert cons[f"{n}/.zarray"] == json.loads((root / n / ".zarray").read_text())
        assert cons[f"{n}/.zattrs"] == json.loads((root / n / ".zattrs").read_text())
    return {n: _decode_array(root, n) for n in names}


REF_COORDS = {
    "release_time": ("release", np.array([np.datetime64("2024-01-01T00:00:00", "ns")])),
    "release_duration_seconds": ("release", np.array([3600.0])),
    "release_lon": ("release", np.array([-122.3])),
    "release_lat": ("release", np.array([37.9])),
    "release_alt_agl_m": ("release", np.array([500.0])),
    "site": ("release", np.array(["MHD"])),
    "time_ago": np.array([0, 1], dtype=np.int64),
    "time_ago_start_hours": ("time_ago", np.array([0.0, 1.0])),
    "time_ago_end_hours": ("time_ago", np.array([1.0, 2.0])),
    "z_bin": np.array([500.0, 1500.0, 2500.0]),
    "z_bottom_m": ("z_bin", np.array([0.0, 1000.0, 2000.0])),
    "z_top_m": ("z_bin", np.array([1000.0, 2000.0, 3000.0])),
    "latitude": np.array([35.25, 35.75, 36.25, 36.75]),
    "longitude": np.array([-122.5, -122.0, -121.5, -121.0, -120.5]),
    "latitude_edge": ("latitude_edge", np.array([35.0, 35.5, 36.0, 36.5, 37.0])),
    "longitude_edge": ("longitude_edge", np.array([-122.75, -122.25, -121.75, -121.25, -120.75, -120.25])),
}


def test_footprint_zarr_roundtrips_coords(tmp_path):
    """Reference test_output_writer_footprint_zarr on the decoded store."""
    fp = np.ones((1, 2, 3, 4, 5), np.float32)
    path = tmp_path / "footprint.zarr"
    OutputWriter().write_footprint_zarr(str(path), fp, coords=REF_COORDS)
    ds = _open(path)
    v, a = ds["footprint"]
    assert v.shape == (1, 2, 3, 4, 5) and v.dtype == np.float32 and np.all(v == 1.0)
    assert tuple(a["_ARRAY_DIMENSIONS"]) == FOOTPRINT_DIMS
    aux = set(a["coordinates"].split())
    assert aux == {"release_time", "release_duration_seconds", "release_lon", "release_lat", "release_alt_agl_m",
                   "site", "time_ago_start_hours", "time_ago_end_hours", "z_bottom_m", "z_top_m"}
    assert ds["release_time"][1]["_ARRAY_DIMENSIONS"] == ["release"]
    assert ds["release_time"][0][0] == np.datetime64("2024-01-01T00:00:00", "ns")
    assert ds["release_lon"][0][0] == -122.3 and ds["release_lat"][0][0] == 37.9
    assert ds["release_alt_agl_m"][0][0] == 500.0 and ds["site"][0][0] == "MHD"
    assert np.array_equal(ds["z_bottom_m"][0], [0.0, 1000.0, 2000.0])  # a real 0 is not masked
    assert np.array_equal(ds["time_ago"][0], [0, 1])
    assert np.array_equal(ds["longitude_edge"][0], REF_COORDS["longitude_edge"][1])
    for name in ("time_ago", "z_bin", "latitude", "longitude", "latitude_edge", "longitude_edge"):
        assert ds[name][1]["_ARRAY_DIMENSIONS"] == [name]


def test_footprint_store_regions_and_errors(tmp_path):
    """create_footprint_store + write_footprint_region: one chunk per release,
    unwritten releases read as NaN (the store's fill), zeros stay zeros, shape
    errors as in the reference."""
    w = OutputWriter()
    path = str(tmp_path / "fp.zarr")
    shape = (4, 2, 3, 4, 5)
    w.create_footprint_store(path, shape=shape)
    meta = json.loads((Path(path) / "footprint" / ".zarray").read_text())
    assert meta["chunks"] == [1, 2, 3, 4, 5] and meta["dtype"] == "<f4" and meta["fill_value"] == "NaN"
    rng = np.random.default_rng(0)
    blk = rng.random((2,) + shape[1:]).astype(np.float32)
    blk[0, 0] = 0.0
    w.write_footprint_region(path, blk, release_start=1, release_stop=3)
    v, _ = _open(Path(path))["footprint"]
    assert np.all(np.isnan(v[0])) and np.all(np.isnan(v[3]))
    assert np.array_equal(v[1:3], blk)
    st = ZarrFootprintStore(path)
    assert np.array_equal(st.read(2), blk[1]) and np.all(np.isnan(st.read(0)))
    with pytest.raises(ValueError, match="must have shape"):
        w.write_footprint_zarr(str(tmp_path / "bad.zarr"), np.ones((2, 3, 4, 5), np.float32))
    with pytest.raises(ValueError, match="leading dim"):
        w.write_footprint_region(path, blk, release_start=0, release_stop=3)
    with pytest.raises(ValueError, match="outside"):
        w.write_footprint_region(path, blk, release_start=3, release_stop=5)
    with pytest.raises(ValueError, match="does not match"):
        w.write_footprint_region(path, blk[:, :1], release_start=0, release_stop=2)
    with pytest.raises(ValueError, match="store shape"):
        w.create_footprint_store(path, shape=(2, 3, 4, 5))


def test_footprint_coords_closed_forms():
    """main.py:394-445: edges are linspace over the grid bounds, centres their
    midpoints, time bins max(1, length / 3600 / n_t) hours wide."""
    grid = dict(lon_min=-10.0, lon_max=5.0, lat_min=40.0, lat_max=52.0, n_x=30, n_y=24, n_t=4,
                z_edges=[0.0, 50.0, 200.0])
    rels = [Release(lon=1.5, lat=45.0, alt_agl_m=100.0, n_particles=10, release_time_s=7200 + 60 * k,
                    duration_s=120, length_s=12 * 3600, release_idx=k) for k in range(3)]
    c = footprint_coords(grid, rels, length_s=12 * 3600)
    assert np.allclose(np.diff(c["longitude_edge"][1]), 0.5) and c["longitude_edge"][1][[0, -1]].tolist() == [-10, 5]
    assert np.allclose(c["latitude"], 40.25 + 0.5 * np.arange(24))
    assert c["time_ago_start_hours"][1].tolist() == [0, 3, 6, 9] and c["time_ago_end_hours"][1][-1] == 12
    assert footprint_coords(grid, rels, length_s=1800)["time_ago_end_hours"][1].tolist() == [1, 2, 3, 4]
    assert c["z_bin"].tolist() == [25.0, 125.0] and c["z_top_m"][1].tolist() == [50.0, 200.0]
    assert c["release_time"][1][2] == np.datetime64("2024-01-01T02:02:00", "ns")
    assert c["release_duration_seconds"][1].tolist() == [120.0] * 3 and "site" not in c


def test_particles_and_trajectory_parquet(tmp_path):
    """Reference test_output_writer_local_files and the release_idx column."""
    w = OutputWriter()
    p = np.array([[10.0, 20.0, 500.0, 0.5], [11.0, 21.0, 550.0, 0.5], [12.0, 22.0, 600.0, 0.5]], np.float32)
    w.write_particles_parquet(str(tmp_path / "a.parquet"), p[:2])
    df = pd.read_parquet(tmp_path / "a.parquet")
    assert list(df.columns) == ["lon", "lat", "alt", "weight"] and len(df) == 2
    w.write_particles_parquet(str(tmp_path / "b.parquet"), p, release_idx=np.array([0, 1, 1]))
    df = pd.read_parquet(tmp_path / "b.parquet")
    assert list(df.columns) == ["lon", "lat", "alt", "weight", "release_idx"]
    assert df["release_idx"].tolist() == [0, 1, 1] and df["release_idx"].dtype == np.int64
    assert np.array_equal(df[["lon", "lat", "alt", "weight"]].to_numpy(), p)
    with pytest.raises(ValueError, match="does not match particles"):
        w.write_particles_parquet(str(tmp_path / "c.parquet"), p[:1], release_idx=np.array([0, 1, 2]))
    with pytest.raises(ValueError, match="must have shape"):
        w.write_particles_parquet(str(tmp_path / "d.parquet"), p[:, :3])
    rows = [{"step": 1, "mean_lon": 10.1}, {"step": 2, "mean_lon": 10.2}]
    write_trajectory_parquet(tmp_path / "t.parquet", 300, rows)
    assert list(pd.read_parquet(tmp_path / "t.parquet")["elapsed_seconds"]) == [300, 600]
    w.write_metadata_json(str(tmp_path / "m.json"), {"hello": "world", "n": 2})
    assert json.loads((tmp_path / "m.json").read_text()) == {"hello": "world", "n": 2}


def test_zlib_store_roundtrips_and_shrinks_sparse_footprints(tmp_path):
    st = ZarrFootprintStore(tmp_path / "z.zarr")
    st.create((2, 24, 1, 160, 240), compressor="zlib")
    fp = np.zeros((2, 24, 1, 160, 240), np.float32)
    rng = np.random.default_rng(1)
    idx = rng.integers(0, fp[0].size, 5000)
    fp[0].reshape(-1)[idx] = rng.random(5000).astype(np.float32)
    st.write_region(0, fp[:1])
    v, _ = _open(tmp_path / "z.zarr")["footprint"]
    assert np.array_equal(v[0], fp[0]) and np.all(np.isnan(v[1]))
    assert np.array_equal(st.read(0), fp[0])
    size = (tmp_path / "z.zarr" / "footprint" / "0.0.0.0.0").stat().st_size
    assert size < fp[0].nbytes / 20
    with pytest.raises(ValueError, match="unsupported compressor"):
        ZarrFootprintStore(tmp_path / "b.zarr").create((1, 1, 1, 1, 1), compressor="lz4")
