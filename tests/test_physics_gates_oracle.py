"""CPU: the reference's physics acceptance gates run on the oracle (the float64
restatement of the step), pinning the composite step the oracle restates
line by line (reference tests/test_dispersion_analytic.py:38-243,
tests/test_diffusion_pde.py:46-169, tests/test_plume_footprint.py:57-190,
tests/test_main_runtime.py:573-921). The GPU versions
(tests/test_gpu_physics_gates.py) run the reference's full sizes through the
engine; here the ensembles are smaller and the runs shorter so the CPU suite
stays within minutes, with the same set-ups and assertions at the sizes that
fit (each reduction noted; statistical bounds kept as in the reference
wherever the smaller ensemble still sits inside them)."""

from __future__ import annotations

import math

import numpy as np
import pytest

import physics_gates as G
from oracle import gridder as OG
from oracle import lpdm as O

PRESCRIBED = O.HannaParams(meander_enabled=False, prescribed=True, max_substeps=6)
ADVECT_ONLY = None


def _homogeneous(sigma, t_l):
    return G.uniform_window(levels=[0.0, 20000.0], lon=np.linspace(0.0, 1.0, 4), lat=np.linspace(1.0, 0.0, 4),
                            sigma=sigma, t_l=t_l)


def test_oracle_ou_autocorrelation_and_stationarity():
    n, sigma_w2, t_l, dt = 6000, 0.7, 100.0, 5.0  # reference n = 100,000
    sigma = math.sqrt(sigma_w2)
    w0 = np.random.default_rng(4111).standard_normal(n) * sigma
    st = G.OracleStepper(_homogeneous(sigma, t_l), 0.5, 0.5, np.full(n, 10000.0), params=PRESCRIBED, seed=4111,
                         w0=w0)
    for _ in range(int(2 * t_l / dt)):
        st.step(dt)
    n_frames = int(3 * t_l / dt) + 1
    frames = np.empty((n_frames, n))
    for k in range(n_frames):
        frames[k] = st.S["wp"]
        st.step(dt)
    var = frames.var(ddof=1)
    assert 0.97 * sigma_w2 <= var <= 1.03 * sigma_w2, var
    centered = frames - frames.mean()
    for tau_over_tl in (0.5, 1.0, 2.0):
        lag = int(round(tau_over_tl * t_l / dt))
        r = (centered[:-lag] * centered[lag:]).mean() / var
        assert abs(r - math.exp(-tau_over_tl)) < 0.02, (tau_over_tl, r)


def test_oracle_solid_body_rotation_returns_to_start():
    omega = 2.0 * math.pi / 3600.0
    win = G.rotation_window(omega)
    lon0, lat0, xy0 = G.rotation_start()
    errors = []
    for dt in (60.0, 30.0, 15.0):
        lo, la, z = lon0.copy(), lat0.copy(), np.full(4, 500.0)
        for _ in range(int(round(3600.0 / dt))):
            lo, la, z = O.advect_rk2(lo, la, z, win, 0.5, dt)
        errors.append(float(np.max(np.linalg.norm(G.rotation_xy(lo, la) - xy0, axis=1))))
    for i in range(len(errors) - 1):
        assert errors[i] / errors[i + 1] > 3.5, errors
    assert errors[-1] < 1e-3 * float(np.max(np.linalg.norm(xy0, axis=1)))
    assert np.all(z == 500.0)


def test_oracle_taylor_dispersion_curve():
    n, t_l, dt = 6000, 100.0, 1.0  # reference n = 100,000 and checkpoints to 30 T_L
    w0 = np.random.default_rng(2201).standard_normal(n)
    st = G.OracleStepper(_homogeneous(1.0, t_l), 0.5, 0.5, np.full(n, 10000.0), params=PRESCRIBED, seed=2201,
                         w0=w0)
    step = 0
    for c in (0.1, 0.3, 1.0):
        while step < int(round(c * t_l / dt)):
            st.step(dt)
            step += 1
        got = float(np.std(st.S["z"], ddof=1))
        expected = math.sqrt(G.taylor_var(1.0, t_l, c * t_l))
        assert abs(got - expected) / expected < 0.05, (c, got, expected)
    assert abs(float(np.std(st.S["z"], ddof=1)) - math.sqrt(G.taylor_var(1.0, t_l, 100.0))) > 0.0


def test_oracle_taylor_coarse_dt_bias_is_bounded():
    n, t_l = 6000, 100.0  # reference n = 100,000; the fine-dt leg runs on the GPU only
    total_t = 5.0 * t_l
    w0 = np.random.default_rng(71).standard_normal(n)
    st = G.OracleStepper(_homogeneous(1.0, t_l), 0.5, 0.5, np.full(n, 10000.0), params=PRESCRIBED, seed=71, w0=w0)
    for _ in range(int(round(total_t / 20.0))):
        st.step(20.0)
    got = float(np.std(st.S["z"], ddof=1))
    expected = math.sqrt(G.taylor_var(1.0, t_l, total_t))
    assert abs(got - expected) / expected < 0.15, got


def test_oracle_diffusion_limit_and_teeth_first_checkpoint():
    n = 5000  # reference n = 100,000 and checkpoints 300, 900, 1800 s
    lev = G.pde_levels()
    win = G.uniform_window(levels=lev, lon=np.linspace(0.0, 1.0, 4), lat=np.linspace(1.0, 0.0, 4),
                           sigma=G.pde_sigma(lev), t_l=G.PDE_TL)
    rng = np.random.default_rng(3301)
    z0 = rng.random(n) * (G.PDE_Z0[1] - G.PDE_Z0[0]) + G.PDE_Z0[0]
    w0 = rng.standard_normal(n) * np.sqrt((G.PDE_K0 + G.PDE_A * np.minimum(z0, G.PDE_ZCAP)) / G.PDE_TL)
    params = O.HannaParams(meander_enabled=False, prescribed=True, z_ubl_m=0.0)
    st = G.OracleStepper(win, 0.5, 0.5, z0, params=params, seed=3301, w0=w0)
    ref, edges = G.pde_reference((300.0,))
    for _ in range(300):
        st.step(G.PDE_DT)
        above = st.S["z"] > G.PDE_HTOP
        st.S["z"] = np.where(above, 2.0 * G.PDE_HTOP - st.S["z"], st.S["z"])
        st.S["wp"] = np.where(above, -st.S["wp"], st.S["wp"])
    hist, _ = np.histogram(st.S["z"], bins=edges)
    got = hist / hist.sum()
    want = ref[300.0]["true"]
    assert np.abs(got - want).sum() < 0.08
    assert abs(got[:3].sum() - want[:3].sum()) / want[:3].sum() < 0.12
    assert np.abs(got - ref[300.0]["half"]).sum() > 0.25
    assert np.abs(got - ref[300.0]["quarter"]).sum() > 0.45


def test_oracle_gaussian_plume_first_bands():
    n, n_steps = 4000, 400  # reference n = 200,000, 800 steps (the bands to 10 T_L fit in 2000 s)
    sigma = math.sqrt(G.PL_SIGMA2)
    rng = np.random.default_rng(9021)
    u0, v0, w0 = (rng.standard_normal(n) * sigma for _ in range(3))
    st = G.OracleStepper(G.plume_window(), 0.0, 0.0, np.full(n, G.PL_ZR), params=PRESCRIBED, seed=9021, u0=u0,
                         v0=v0, w0=w0)
    b = G.plume_gridder_bounds()
    grid = dict(lon_min=b["lon_bounds"][0], lon_max=b["lon_bounds"][1], lat_min=b["lat_bounds"][0],
                lat_max=b["lat_bounds"][1], z_edges=list(b["z_edges_m"]), n_r=1, n_t=1, n_z=2, n_y=G.PL_NY,
                n_x=G.PL_NX)
    fp = np.zeros(2 * G.PL_NY * G.PL_NX)
    zeros = np.zeros(n, np.int64)
    active = np.ones(n, bool)
    w = np.full(n, np.float32(1.0 / n))
    for _ in range(n_steps):
        st.step(G.PL_DT)
        OG.accumulate(fp, st.S["lon"], st.S["lat"], st.S["z"], active, w, zeros, zeros, G.PL_DT, grid)
    raw = fp.reshape(2, G.PL_NY, G.PL_NX)[0]
    ref = G.plume_analytic(n_steps)
    t_l = G.PL_TL
    # Bounds of the reference at its size; the 50x smaller ensemble leaves the
    # per-column means within them, the per-column maxima are 2x wider.
    G.plume_checks(raw, ref, bands=[(2 * t_l, 5 * t_l, 0.10, 0.02), (5 * t_l, 10 * t_l, 0.15, 0.05)],
                   corr_min=0.98, ratio_tol=0.03, width_tol=0.05)


def _oracle_wmc(met, z0, blh, seed, n_steps):
    win = G.wmc_oracle_window(met)
    params = O.HannaParams(meander_enabled=False, max_substeps=50)
    st = G.OracleStepper(win, 0.0, 45.0, z0, params=params, seed=seed)
    for _ in range(n_steps):
        st.step(15.0)
        above = st.S["z"] > blh
        st.S["z"] = np.where(above, 2.0 * blh - st.S["z"], st.S["z"])
        st.S["wp"] = np.where(above, -st.S["wp"], st.S["wp"])
    return st.S["z"]


def test_oracle_v1_well_mixed_backward():
    blh = 3000.0
    met = G.wmc_met(blh_m=blh)
    z0 = np.random.default_rng(202605301).random(3000) * blh  # reference n = 6,000
    z = _oracle_wmc(met, z0, blh, 202605301, 600)  # reference 1,500 steps (~40 T_L here)
    in_bl = (z >= 0.0) & (z <= blh)
    assert in_bl.mean() > 0.99
    counts, _ = np.histogram(z[in_bl], bins=np.linspace(0.0, blh, 11))
    interior = counts[2:8]
    assert float(np.sqrt(np.mean(((interior - interior.mean()) / interior.mean()) ** 2))) < 0.15, counts.tolist()


def test_oracle_v1_density_weighted_well_mixed():
    blh, z_top, p_bot, p_top, t_k = 5000.0, 6000.0, 1000.0, 500.0, 280.0
    met = G.wmc_met(blh_m=blh, p_top_hpa=p_top, z_top=z_top)
    rng = np.random.default_rng(202605302)
    n = 3000  # reference n = 6,000
    rho_max = float(G.density_at(0.0, p_bot_hpa=p_bot, p_top_hpa=p_top, z_top_m=z_top, t_kelvin=t_k))
    z_init: list[float] = []
    while len(z_init) < n:
        z_try = rng.uniform(0.0, blh, size=n - len(z_init))
        rho_try = G.density_at(z_try, p_bot_hpa=p_bot, p_top_hpa=p_top, z_top_m=z_top, t_kelvin=t_k)
        z_init.extend(z_try[rng.uniform(0.0, rho_max, size=z_try.shape) < rho_try].tolist())
    z = _oracle_wmc(met, np.array(z_init[:n]), blh, 202605302, 600)  # reference 1,500 steps
    in_bl = (z >= 0.0) & (z <= blh)
    assert in_bl.mean() > 0.99
    edges = np.linspace(0.0, blh, 11)
    counts, _ = np.histogram(z[in_bl], bins=edges)
    rho = G.density_at(0.5 * (edges[:-1] + edges[1:]), p_bot_hpa=p_bot, p_top_hpa=p_top, z_top_m=z_top,
                       t_kelvin=t_k)
    expected = counts.sum() * rho / rho.sum()
    ic, ie = counts[3:7], expected[3:7]
    rel_rms = float(np.sqrt(np.mean(((ic - ie) / ie) ** 2)))
    flat_rms = float(np.sqrt(np.mean(((ic - ic.mean()) / ic.mean()) ** 2)))
    assert rel_rms < 0.15, (rel_rms, counts.tolist())
    assert flat_rms > 0.5 * rel_rms or flat_rms < 0.05
