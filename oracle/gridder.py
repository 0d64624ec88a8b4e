"""Numpy restatement of ``FootprintGridder.accumulate`` and the driver's
time-bin / escape bookkeeping.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Index arithmetic is done in float32 with IEEE division, exactly as the reference
does on float32 particle tensors (footprint_gridder.py:160-167); this is the
bit-exact contract the HIP key kernel is held to.
"""

from __future__ import annotations

import numpy as np


def cell_indices(lon, lat, z, grid):
    """(x, y, z) bin indices, reference footprint_gridder.py:160-167.

    x = floor((lon - lon_min) / max(1e-6, span) * n_x) in float32;
    z = bucketize(z, edges, right=False) - 1, i.e. bins (e_b, e_{b+1}].
    """

    f32 = np.float32
    lon = np.asarray(lon, f32)
    lat = np.asarray(lat, f32)
    z = np.asarray(z, f32)
    lon_min = f32(grid["lon_min"])
    lat_min = f32(grid["lat_min"])
    span_x = f32(max(1e-6, grid["lon_max"] - grid["lon_min"]))
    span_y = f32(max(1e-6, grid["lat_max"] - grid["lat_min"]))
    with np.errstate(invalid="ignore"):
        xi = np.floor(((lon - lon_min) / span_x) * f32(grid["n_x"])).astype(np.int64)
        yi = np.floor(((lat - lat_min) / span_y) * f32(grid["n_y"])).astype(np.int64)
    edges = np.asarray(grid["z_edges"], f32)
    zi = np.searchsorted(edges, z, side="left").astype(np.int64) - 1
    return xi, yi, zi


def time_bins(we_off, cursor_off: float, n_t: int):
    """reference main.py:1459-1465: floor(clamp(we_off - cursor_off, 0) / 3600)."""

    el = np.maximum(np.asarray(we_off, np.float32) - np.float32(cursor_off), np.float32(0.0))
    t = np.floor(el / np.float32(3600.0)).astype(np.int64)
    if n_t == 1:
        t = np.minimum(t, 0)
    return t


def keys_and_values(lon, lat, z, active, weights, t_idx, rel_idx, dt, grid):
    """Flat cell key (int64, -1 when dropped) and float32 value w*dt per particle
    (reference footprint_gridder.py:169-196)."""

    xi, yi, zi = cell_indices(lon, lat, z, grid)
    R, T, Zo, Y, X = grid["n_r"], grid["n_t"], grid["n_z"], grid["n_y"], grid["n_x"]
    r = np.asarray(rel_idx, np.int64)
    t = np.asarray(t_idx, np.int64)
    valid = (
        np.asarray(active, bool)
        & (r >= 0) & (r < R)
        & (xi >= 0) & (xi < X)
        & (yi >= 0) & (yi < Y)
        & (zi >= 0) & (zi < Zo)
        & (t >= 0) & (t < T)
    )
    key = ((((r * T + t) * Zo + zi) * Y + yi) * X + xi)
    key = np.where(valid, key, -1)
    val = np.asarray(weights, np.float32) * np.float32(dt)
    return key, np.where(valid, val, np.float32(0.0))


def accumulate(fp, lon, lat, z, active, weights, t_idx, rel_idx, dt, grid):
    """Add this step's residence time into ``fp`` (flat float64 or float32 view),
    in particle order (a fixed order: the oracle is deterministic)."""

    key, val = keys_and_values(lon, lat, z, active, weights, t_idx, rel_idx, dt, grid)
    m = key >= 0
    np.add.at(fp, key[m], val[m].astype(fp.dtype))


def within_met_domain(lon, lat, z, dom):
    """reference main.py:690-711: inclusive lon/lat bounds and alt <= alt_max."""

    return (
        (lon >= dom["lon_min"]) & (lon <= dom["lon_max"])
        & (lat >= dom["lat_min"]) & (lat <= dom["lat_max"])
        & (z <= dom["alt_max"])
    )
