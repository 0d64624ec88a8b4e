"""GPU: the per-column surface-parcel lift (glide_column_cape) against the
oracle's compute_lcl_lnb_cape on every column, through the C ABI.

Both sides compute in float64 from the same float32 fields; they differ only
in the last bits of exp/log/pow and in the CAPE summation order. Bars: LCL and
LNB indices identical, CAPE within 1e-9 relative (+1e-9 J/kg), T_v excess
within 1e-9 K. The descriptors then feed glide_convect_columns: indices
identical, cloud-base flux within 1e-6 (float32 closure)."""

from __future__ import annotations

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

from glide_amd import _lib as L  # noqa: E402
from glide_amd.runner import BatchEngine, ConvectionOptions  # noqa: E402
from glide_amd.synthetic import SyntheticMet, workload_config  # noqa: E402
from oracle import convection as OC  # noqa: E402


def _stream():
    return torch.cuda.current_stream().cuda_stream


def _lift(t_s, t_e, q_s, q_e, alpha, p_hpa, levels):
    nz, ny, nx = t_s.shape
    dev = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in
           dict(t_s=t_s, t_e=t_e, q_s=q_s, q_e=q_e).items()}
    p = torch.from_numpy(np.asarray(p_hpa, np.float64)).cuda()
    lev = torch.from_numpy(np.asarray(levels, np.float32)).cuda()
    lcl = torch.empty((ny, nx), dtype=torch.int32, device="cuda")
    lnb = torch.empty_like(lcl)
    cape = torch.empty((ny, nx), dtype=torch.float64, device="cuda")
    dtv = torch.empty_like(cape)
    conv = torch.empty((4, ny, nx), dtype=torch.float32, device="cuda")
    print(f"glide_column_cape nz={nz} ny={ny} nx={nx} alpha={alpha}", flush=True)
    L.check(L.lib().glide_column_cape(dev["t_s"].data_ptr(), dev["t_e"].data_ptr(), dev["q_s"].data_ptr(),
                                      dev["q_e"].data_ptr(), float(alpha), p.data_ptr(), lev.data_ptr(), nz, ny, nx,
                                      lcl.data_ptr(), lnb.data_ptr(), cape.data_ptr(), dtv.data_ptr(),
                                      conv.data_ptr(), _stream()))
    torch.cuda.synchronize()
    return lcl.cpu().numpy(), lnb.cpu().numpy(), cape.cpu().numpy(), dtv.cpu().numpy(), conv.cpu().numpy()


def _check(got, want):
    lcl, lnb, cape, dtv = got[:4]
    olcl, olnb, ocape, odtv = want
    assert np.array_equal(lcl, olcl)
    assert np.array_equal(lnb, olnb)
    np.testing.assert_allclose(cape, ocape, rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(dtv, odtv, rtol=0, atol=1e-9)


def _soundings(nz=12):
    """The reference's moist-unstable and dry-stable soundings
    (tests/test_convection.py:160-196) and a cold column with a 1 km surface
    inversion under a 6.5 K/km lapse (conditionally unstable aloft), side by
    side."""
    p_hpa = np.exp(np.linspace(np.log(1000.0), np.log(100.0), nz))
    z_km = -8.0 * np.log(p_hpa / 1000.0)
    t_u = np.maximum(300.0 - 6.5 * z_km, 200.0)
    q_u = np.maximum(0.014 * np.exp(-(1000.0 - p_hpa) / 400.0), 1e-6)
    t_s = np.minimum(280.0 + 2.0 * z_km, 290.0)
    q_s = np.full(nz, 1e-5)
    t_w = np.where(z_km < 1.0, 265.0 + 8.0 * z_km, 273.0 - 6.5 * (z_km - 1.0))
    q_w = 0.003 * np.exp(-z_km / 2.0)
    T = np.stack([t_u, t_s, t_w], -1)[:, None, :].astype(np.float32)
    Q = np.stack([q_u, q_s, q_w], -1)[:, None, :].astype(np.float32)
    return T, Q, p_hpa


def test_column_cape_reference_soundings():
    T, Q, p_hpa = _soundings()
    levels = (np.arange(T.shape[0]) * 1000.0).astype(np.float32)
    for alpha in (0.0, 0.5):
        got = _lift(T, T + np.float32(0.5), Q, Q, alpha, p_hpa, levels)
        want = OC.column_cape(T, T + np.float32(0.5), Q, Q, alpha, p_hpa)
        _check(got, want)
    lcl, lnb, cape, dtv, conv = got
    assert lcl[0, 0] > 0 and lnb[0, 0] > lcl[0, 0] and cape[0, 0] > 100.0 and dtv[0, 0] > 0.0
    assert cape[0, 1] < 50.0  # the dry stable sounding
    np.testing.assert_array_equal(conv, OC.conv_descriptors_from_lift(lcl, lnb, cape, dtv, levels))


def test_column_cape_on_synthetic_met_and_descriptors():
    cfg = workload_config("plain-cpu-10k-3h")
    met = SyntheticMet(cfg["spec"])
    eng = BatchEngine(cfg, device="cuda", met=met, convection=ConvectionOptions(parcel_lift=True))
    s3, _, ph = eng.met.resampled(1, _stream())
    e3, _, _ = eng.met.resampled(2, _stream())
    torch.cuda.synchronize()
    t_s, q_s = s3[3].cpu().numpy(), s3[4].cpu().numpy()
    t_e, q_e = e3[3].cpu().numpy(), e3[4].cpu().numpy()
    p_hpa = ph.cpu().numpy()
    levels = met.spec.agl_levels.astype(np.float32)
    alpha = 0.37
    got = _lift(t_s, t_e, q_s, q_e, alpha, p_hpa, levels)
    want = OC.column_cape(t_s, t_e, q_s, q_e, alpha, p_hpa)
    _check(got, want)
    lcl, lnb, cape, dtv, conv = got
    print("columns with an LCL", (lcl >= 0).mean(), "CAPE >= 50", (cape >= 50).mean(), "max CAPE", cape.max())
    assert (lcl >= 0).mean() > 0.5
    # descriptors -> trigger and closure, as the engine chains them
    nz, ny, nx = eng.met.nz, eng.met.ny, eng.met.nx
    conv_d = torch.from_numpy(conv).cuda()
    L.check(L.lib().glide_convect_columns(conv_d.data_ptr(), s3[3].data_ptr(), e3[3].data_ptr(), alpha,
                                          ph.data_ptr(), eng.met.levels_f.data_ptr(), nz, ny, nx, 0.03, 0.9, 50.0,
                                          500.0, eng.i_lcl.data_ptr(), eng.i_lnb.data_ptr(), eng.m_b.data_ptr(),
                                          eng.layer_mass.data_ptr(), _stream()))
    torch.cuda.synchronize()
    oconv = OC.conv_descriptors_from_lift(*want, levels)
    olcl, olnb, omb = OC.column_descriptors(oconv, t_s, t_e, alpha, p_hpa, levels)
    assert np.array_equal(eng.i_lcl.cpu().numpy().reshape(ny, nx), olcl)
    assert np.array_equal(eng.i_lnb.cpu().numpy().reshape(ny, nx), olnb)
    np.testing.assert_allclose(eng.m_b.cpu().numpy().reshape(ny, nx), omb, rtol=1e-6)


def test_engine_runs_with_parcel_lift():
    """The engine's convection event with the device lift: the descriptors it
    computed for the bracket equal a direct lift at that step's alpha, and a
    run with it matches between eager steps and graph replay."""
    cfg = workload_config("plain-cpu-10k-3h")
    met = SyntheticMet(cfg["spec"])
    opts = ConvectionOptions(parcel_lift=True)
    eng = BatchEngine(cfg, device="cuda", met=met, convection=opts)
    sched = eng.sched
    first = next(i for i in range(1, len(sched)) if sched[i].bracket_hour != sched[i - 1].bracket_hour)
    eng.run(first + 1)
    torch.cuda.synchronize()
    st = sched[first]
    s3 = eng.met.resampled(st.bracket_hour, _stream())[0]
    e3 = eng.met.resampled(st.bracket_hour + 1, _stream())[0]
    ph = eng.met.resampled(st.bracket_hour, _stream())[2]
    direct = _lift(s3[3].cpu().numpy(), e3[3].cpu().numpy(), s3[4].cpu().numpy(), e3[4].cpu().numpy(), st.alpha,
                   ph.cpu().numpy(), met.spec.agl_levels)
    np.testing.assert_array_equal(eng.conv_lift.cpu().numpy(), direct[4])
    eng.run()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        g = BatchEngine(cfg, device="cuda", met=met, convection=opts, use_graph=True)
        g.run()
    torch.cuda.synchronize()
    a, b = eng.state_numpy(), g.state_numpy()
    for k in ("lon", "lat", "z", "wp"):
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(eng.footprint.cpu().numpy(), g.footprint.cpu().numpy())
    g.close()


def test_column_sum_check_many_levels_bit_exact():
    """The per-column float64 mass-conservation check (glide_convect with
    colsum_err) on 128 levels and hand-picked clouds, against the oracle's
    float32 matrix summed in float64 in index order, exactly: cloud blocks
    small enough to be staged and too large for it (rows below base x cloud
    levels > 1,024), levels above 64 (the second lane register), clouds from
    the ground, one-layer clouds, the top level as cloud top, and columns that
    do not fire (no cloud, no flux)."""
    import ctypes as C

    nz, ny, nx = 128, 2, 8
    rng = np.random.default_rng(7)
    p_hpa = np.linspace(1000.0, 20.0, nz)
    lm = OC.layer_masses(p_hpa * 100.0).astype(np.float32)
    clouds = [(0, 5), (1, 2), (10, 60), (40, 100), (30, 127), (63, 64), (64, 127), (3, 127),
              (-1, -1), (5, 9), (2, 50), (20, 90), (100, 127), (0, 127), (12, 13), (7, 8)]
    lcl = np.array([c[0] for c in clouds], np.int32).reshape(ny, nx)
    lnb = np.array([c[1] for c in clouds], np.int32).reshape(ny, nx)
    mb = rng.uniform(0.5, 40.0, (ny, nx)).astype(np.float32)
    mb[0, 1] = 0.0  # a cloud with no cloud-base flux does not fire
    assert any(a >= 0 and a * (b - a + 1) > 1024 for a, b in clouds)  # the unstaged path
    assert any(0 <= a and a * (b - a + 1) <= 1024 for a, b in clouds)  # the staged path
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    lcl_d, lnb_d, mb_d, lm_d = dev(lcl), dev(lnb), dev(mb), dev(lm)
    levels = dev(np.linspace(0.0, 20000.0, nz).astype(np.float32))
    one = torch.zeros(1, dtype=torch.float32, device="cuda")
    alive = torch.zeros(1, dtype=torch.uint8, device="cuda")
    parts = L.GlideParticles(n=0, pos_stride=1, lon=one.data_ptr(), lat=one.data_ptr(), z=one.data_ptr(),
                             r_off=one.data_ptr(), alive=alive.data_ptr())
    grid = L.GlideGrid(nx=nx, ny=ny, nz=nz, reserved=0, lon0=0.0, lon1=1.75, lat0=0.0, lat1=0.25,
                       levels=levels.data_ptr())
    prm = L.GlideStepParams(dt=60.0, alpha=0.5, max_substeps=1, flags=0)
    err = torch.full((ny * nx,), -1.0, dtype=torch.float64, device="cuda")
    print(f"glide_convect colsum nz={nz} columns={ny * nx}", flush=True)
    L.check(L.lib().glide_convect(C.byref(prm), C.byref(parts), C.byref(grid), lcl_d.data_ptr(), lnb_d.data_ptr(),
                                  mb_d.data_ptr(), lm_d.data_ptr(), err.data_ptr(), None, None, None, _stream()))
    torch.cuda.synchronize()
    got = err.cpu().numpy().reshape(ny, nx)
    for iy in range(ny):
        for ix in range(nx):
            f32 = OC.mass_flux_matrix_f32(lm, int(lcl[iy, ix]), int(lnb[iy, ix]), float(mb[iy, ix]))
            want = OC.column_sum_error(f32, lm)
            assert got[iy, ix] == want, (iy, ix, lcl[iy, ix], lnb[iy, ix], got[iy, ix], want)
    assert got[0, 1] == 0.0 and got[1, 0] == 0.0  # no flux / no cloud
