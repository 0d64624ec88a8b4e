"""GPU tests of the drop-in classes behind the reference's interfaces:
FootprintGridder (footprint_gridder.py), HannaScheme.step
(turbulence/hanna.py) and the device resample (met_reader.py:641-693)."""

from __future__ import annotations

from datetime import datetime, timezone

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from glide_amd.footprint import FootprintGridder  # noqa: E402
from glide_amd.met import HourlyMetTensors, MetFieldMetadata, resample_hour_to_agl  # noqa: E402
from glide_amd.runner import BatchEngine  # noqa: E402
from glide_amd.synthetic import SyntheticMet, workload_config  # noqa: E402
from glide_amd.turbulence import GPUEngine, HannaScheme, get_scheme, list_schemes  # noqa: E402
from oracle import lpdm as O  # noqa: E402
from oracle import resample as OR  # noqa: E402
from oracle import run as ORUN  # noqa: E402


def gridder(n_releases=1):
    return FootprintGridder((-1.0, 1.0), (-1.0, 1.0), (0.0, 500.0, 1000.0), 3, 4, 4, n_releases=n_releases,
                            device="cuda")


def col(v, n, dt=torch.int64):
    return torch.full((n,), v, dtype=dt, device="cuda")


def test_gridder_single_particle_and_drops():
    g = gridder()
    g.accumulate(torch.tensor([[-0.75, -0.75, 250.0]], device="cuda"), torch.tensor([True], device="cuda"),
                 torch.tensor([0.5], device="cuda"), col(1, 1), col(0, 1), 10.0)
    assert g.tensor[0, 1, 0, 0, 0].item() == 5.0 and g.tensor.sum().item() == 5.0
    g = gridder()
    P = torch.tensor([[-2.0, 0, 100], [2.0, 0, 100], [0, -2.0, 100], [0, 2.0, 100], [0, 0, -100],
                      [0, 0, 2000], [0, 0, 100]], device="cuda")
    g.accumulate(P, torch.ones(7, dtype=torch.bool, device="cuda"), torch.ones(7, device="cuda"), col(0, 7),
                 col(0, 7), 1.0)
    assert abs(g.tensor.sum().item() - 1.0) < 1e-12
    g = gridder(n_releases=2)
    P = torch.tensor([[-0.5, -0.5, 100.0]] * 4, device="cuda")
    g.accumulate(P, torch.tensor([True, True, False, True], device="cuda"), torch.ones(4, device="cuda"),
                 torch.tensor([0, 2, 0, 7], device="cuda"), torch.tensor([0, 1, 1, 0], device="cuda"), 2.0)
    assert g.tensor[0, 0].sum().item() == 2.0 and g.tensor[1, 2].sum().item() == 2.0 and g.tensor.sum() == 4.0


def test_gridder_validation_errors():
    with pytest.raises(ValueError):
        FootprintGridder((-1, 1), (-1, 1), (0.0, 0.0), 1, 2, 2, device="cuda")
    with pytest.raises(ValueError):
        FootprintGridder((-1, 1), (-1, 1), (0.0,), 1, 2, 2, device="cuda")
    with pytest.raises(ValueError):
        FootprintGridder((-1, 1), (-1, 1), (0.0, 1.0), 1, 2, 2, n_releases=0, device="cuda")
    with pytest.raises(MemoryError):
        FootprintGridder((-1, 1), (-1, 1), (0.0, 1.0), 100000, 1000, 1000, device="cuda")


def _window(met, h3_s, h3_e, src_s, src_e, p_hpa, levels):
    Z, Y, X = h3_s.shape[1:]
    names = ("u", "v", "w", "t", "q", "blh", "sp", "ustar", "shf")

    def full(h3, src2):
        s2 = torch.as_tensor(src2, device="cuda")[:, None].expand(-1, Z, -1, -1)
        return torch.cat([h3, s2], 0).contiguous()

    md = MetFieldMetadata(lon=met.lon, lat=met.lat, level=levels, pressure_level_hpa=p_hpa,
                          time_start=datetime(2024, 1, 1, tzinfo=timezone.utc),
                          time_end=datetime(2024, 1, 1, 1, tzinfo=timezone.utc), variable_units={})
    hag = torch.as_tensor(np.broadcast_to(levels[:, None, None], (Z, Y, X)).copy(), device="cuda")
    return HourlyMetTensors(full(h3_s, src_s), full(h3_e, src_e), md, names, hag)


def test_resample_api_bit_exact_and_hanna_step_matches_engine_and_oracle():
    cfg = workload_config("plain-cpu-10k-3h")
    met = SyntheticMet(cfg["spec"])
    lev = met.spec.agl_levels
    hs = {}
    for h in (2, 3):
        sh = met.hour(h)
        d3, p = resample_hour_to_agl(torch.as_tensor(sh.src3, device="cuda"),
                                     torch.as_tensor(sh.height_agl, device="cuda"),
                                     torch.as_tensor(met.terrain, device="cuda"), met.lat, met.lon, lev,
                                     sh.p_level_pa / 100.0)
        o3, _, op = OR.resample_hour(sh.src3, sh.src2, sh.height_agl, met.terrain, met.lat, met.lon, lev,
                                     sh.p_level_pa)
        assert np.array_equal(d3.cpu().numpy(), o3)
        np.testing.assert_allclose(p.cpu().numpy(), op, rtol=1e-12)
        hs[h] = (d3, sh.src2, p.cpu().numpy())
    mw = _window(met, hs[2][0], hs[3][0], hs[2][1], hs[3][1], hs[2][2], lev)
    assert "hanna_1982" in list_schemes()
    scheme = get_scheme("hanna_1982", meander_enabled=True, max_substeps=6, seed=cfg["seed"])
    assert isinstance(scheme, HannaScheme) and scheme.step_includes_advection(None)
    eng = BatchEngine(cfg, device="cuda", met=met)
    st = eng.sched[1]
    n = eng.n
    ra = eng.ra
    parts = torch.as_tensor(np.stack([ra.lon, ra.lat, ra.z, ra.weight], 1), device="cuda").contiguous()
    state = scheme.initialize_state(n, device=torch.device("cuda"), dtype=torch.float32)
    active_np = (ra.r_off >= np.float32(st.cursor_off)) & (ra.r_off - np.float32(eng.length_s)
                                                           <= np.float32(st.cursor_off))
    active_np[::7] = False  # some inactive particles must stay untouched
    active = torch.as_tensor(active_np, device="cuda")
    before = parts.clone()
    scheme.step_index = st.noise_step
    out, state = scheme.step(parts, state, mw, st.alpha, float(st.delta_s), active, GPUEngine())
    got = out.cpu().numpy()
    assert np.array_equal(got[~active_np], before.cpu().numpy()[~active_np])
    assert np.all(state["w_prime"].cpu().numpy()[~active_np] == 0.0)
    # the engine path, same kernel and noise: identical
    eng.step_i = st.index
    eng.run_step()
    es = eng.state_numpy()
    assert np.array_equal(got[active_np, 0], es["lon"][active_np])
    assert np.array_equal(got[active_np, 2], es["z"][active_np])
    # and the float64 oracle within the per-step bar
    win = ORUN.build_window(OR.resample_hour(*_src(met, 2)), OR.resample_hour(*_src(met, 3)), lev, met.lon,
                            met.lat, np.float64)
    S = {k: getattr(ra, k).astype(np.float64) for k in ("lon", "lat", "z")}
    for k in ("up", "vp", "wp", "um", "vm"):
        S[k] = np.zeros(n)
    S["uid"] = ra.uid
    ref = O.step(S, active_np, win, st.alpha, float(st.delta_s), O.HannaParams(), seed=cfg["seed"],
                 step_index=st.noise_step)
    e = np.abs(got[active_np, 0] - ref["lon"][active_np]) / np.abs(ref["lon"][active_np])
    assert (e <= 1e-5).mean() >= 0.999


def _src(met, h):
    sh = met.hour(h)
    return (sh.src3, sh.src2, sh.height_agl, met.terrain, met.lat, met.lon, met.spec.agl_levels, sh.p_level_pa)


def test_hanna_step_noise_keys_and_particle_layouts():
    """HannaScheme.step keys its noise by model time (a fresh instance at t
    equals one first called later in time; a repeated call at one model time
    and a new backward sweep draw afresh), by ``state["particle_id"]`` when
    given (a row permutation permutes the result), and takes the (N, 4) block
    at any layout (a strided view and a transposed one are updated in place,
    bit-identical to the contiguous case)."""
    import physics_gates as G

    met = G.wmc_met(blh_m=1500.0)
    met = HourlyMetTensors(met.hour_start.cuda(), met.hour_end.cuda(), met.metadata, met.channel_names,
                           met.height_agl_m.cuda())
    n = 4096
    g = torch.Generator(device="cpu").manual_seed(5)
    base = torch.zeros((n, 4))
    base[:, 1] = 45.0
    base[:, 2] = torch.empty(n).uniform_(10.0, 1400.0, generator=g)
    base[:, 3] = 1.0 / n
    base = base.cuda()
    active = torch.ones(n, dtype=torch.bool, device="cuda")
    eng = GPUEngine()

    def run(scheme, alpha, parts=None, state=None):
        parts = base.clone() if parts is None else parts
        st = scheme.initialize_state(n, device=torch.device("cuda"), dtype=torch.float32)
        if state:
            st.update(state)
        out, _ = scheme.step(parts, st, met, alpha, 60.0, active, eng)
        torch.cuda.synchronize()
        return out

    mk = lambda: HannaScheme(seed=9)  # noqa: E731
    fresh = run(mk(), 0.25)
    s = mk()
    run(s, 0.75)
    assert torch.equal(run(s, 0.25), fresh)
    assert not torch.equal(run(s, 0.25), fresh)  # a repeat at the same model time
    assert not torch.equal(run(s, 0.5), run(mk(), 0.5))  # a new sweep
    perm = torch.randperm(n, generator=torch.Generator().manual_seed(4)).cuda()
    ids = torch.arange(n, dtype=torch.int32, device="cuda")
    a = run(mk(), 0.25, state={"particle_id": ids})
    b = run(mk(), 0.25, parts=base[perm].contiguous(), state={"particle_id": ids[perm].contiguous()})
    assert torch.equal(a[perm], b)
    wide = torch.zeros((n, 7), device="cuda")
    wide[:, 2:6] = base
    out = run(mk(), 0.25, parts=wide[:, 2:6])
    assert out.data_ptr() == wide[:, 2:6].data_ptr() and torch.equal(out, fresh)
    assert torch.all(wide[:, :2] == 0) and torch.all(wide[:, 6] == 0)
    tr = base.t().contiguous().t()  # column-major (N, 4)
    out = run(mk(), 0.25, parts=tr)
    assert out.data_ptr() == tr.data_ptr() and torch.equal(out, fresh)
