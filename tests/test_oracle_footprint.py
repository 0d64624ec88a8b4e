"""Pins the oracle gridder with the reference's known answers
(tests/test_footprint.py:58-462): exact cells, drops at every edge, inactive
and out-of-range time/release bins, non-uniform z edges, per-release slices."""

from __future__ import annotations

import numpy as np

from oracle import gridder as G


def grid(n_r=1, n_t=3, z_edges=(0.0, 500.0, 1000.0), ny=4, nx=4):
    return dict(lon_min=-1.0, lon_max=1.0, lat_min=-1.0, lat_max=1.0, n_x=nx, n_y=ny, n_t=n_t, n_r=n_r,
                z_edges=list(z_edges), n_z=len(z_edges) - 1)


def acc(g, P, active, w, t, r, dt):
    fp = np.zeros(g["n_r"] * g["n_t"] * g["n_z"] * g["n_y"] * g["n_x"])
    G.accumulate(fp, P[:, 0], P[:, 1], P[:, 2], active, w, t, r, dt, g)
    return fp.reshape(g["n_r"], g["n_t"], g["n_z"], g["n_y"], g["n_x"])


def test_single_particle_lands_in_expected_cell():
    g = grid()
    fp = acc(g, np.array([[-0.75, -0.75, 250.0]]), [True], [0.5], [1], [0], 10.0)
    assert fp[0, 1, 0, 0, 0] == 5.0 and fp.sum() == 5.0


def test_out_of_bounds_and_inactive_dropped():
    g = grid()
    P = np.array([[-2.0, 0, 100], [2.0, 0, 100], [0, -2.0, 100], [0, 2.0, 100], [0, 0, -100],
                  [0, 0, 2000], [0, 0, 100], [0.1, 0.1, 100]])
    fp = acc(g, P, [True] * 7 + [False], np.ones(8), np.zeros(8, int), np.zeros(8, int), 1.0)
    assert abs(fp.sum() - 1.0) < 1e-12


def test_edges_half_open_bins():
    """x/y bins [k, k+1): lon_max is dropped, lon_min kept; z bins (e_b, e_b+1]:
    z == e_0 dropped, z == e_1 in bin 0 (bucketize right=False)."""
    g = grid()
    xi, yi, zi = G.cell_indices([-1.0, 1.0, 0.0, 0.0], [0.0, 0.0, -1.0, 1.0], [0.0, 500.0, 500.0001, 1000.0], g)
    assert list(xi[:2]) == [0, 4] and list(yi[2:]) == [0, 4]
    assert list(zi) == [-1, 0, 1, 1]


def test_time_and_release_bins():
    g = grid(n_r=2)
    P = np.array([[-0.5, -0.5, 100.0]] * 5)
    fp = acc(g, P, [True] * 5, np.ones(5), [0, 1, 2, 3, -1], [0, 1, 1, 0, 0], 1.0)
    assert fp[0, 0].sum() == 1.0 and fp[1, 1].sum() == 1.0 and fp[1, 2].sum() == 1.0 and fp.sum() == 3.0
    fp = acc(g, P, [True] * 5, np.ones(5), [0] * 5, [0, 1, 2, -1, 1], 1.0)
    assert fp[0].sum() == 1.0 and fp[1].sum() == 2.0


def test_non_uniform_z_edges():
    g = grid(n_t=1, z_edges=(0.0, 40.0, 1000.0, 5000.0), ny=2, nx=2)
    P = np.array([[-0.5, -0.5, 20.0], [-0.5, -0.5, 500.0], [-0.5, -0.5, 3000.0]])
    fp = acc(g, P, [True] * 3, np.ones(3), [0] * 3, [0] * 3, 1.0)
    assert fp[0, 0, 0].sum() == 1.0 and fp[0, 0, 1].sum() == 1.0 and fp[0, 0, 2].sum() == 1.0


def test_mass_conservation_in_bounds():
    g = grid()
    rng = np.random.default_rng(42)
    n = 100
    P = np.stack([rng.uniform(-0.99, 0.99, n), rng.uniform(-0.99, 0.99, n), rng.uniform(0.1, 999, n)], 1)
    w = np.full(n, 1.0 / n, np.float32)
    fp = acc(g, P, np.ones(n, bool), w, np.zeros(n, int), np.zeros(n, int), 5.0)
    assert abs(fp.sum() - float((w * np.float32(5.0)).astype(np.float64).sum())) < 1e-12


def test_time_bin_and_escape_predicates():
    t = G.time_bins(np.float32([60.0, 60.0, 60.0]), -3600.0, 3)
    assert list(t) == [1, 1, 1]
    assert list(G.time_bins(np.float32([60.0, 7200.0, 0.0]), 60.0, 3)) == [0, 1, 0]
    assert list(G.time_bins(np.float32([60.0, 99999.0]), -99999.0, 1)) == [0, 0]
    dom = dict(lon_min=0.0, lon_max=1.0, lat_min=0.0, lat_max=1.0, alt_max=100.0)
    ins = G.within_met_domain(np.array([0.0, 1.0, 1.0001, 0.5, 0.5]), np.array([0.0, 1.0, 0.5, 0.5, 0.5]),
                              np.array([0.0, 100.0, 1.0, 100.01, -5.0]), dom)
    assert list(ins) == [True, True, False, False, True]
