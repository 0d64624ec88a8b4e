"""CPU: the drop-in classes' registry, constants and constructor checks
(reference tests/test_hanna.py:38-45, 290-296, 367-378). No compute."""

from __future__ import annotations

import pytest

pytest.importorskip("torch")

from glide_amd import hanna  # noqa: E402
from glide_amd.turbulence import HannaScheme, get_scheme, list_schemes  # noqa: E402


def test_hanna_is_registered():
    assert "hanna_1982" in list_schemes()
    instance = get_scheme("hanna_1982")
    assert isinstance(instance, HannaScheme)
    assert instance.required_met_keys() == ("t", "ustar", "shf")


def test_hanna_above_bl_constants_used():
    assert hanna.ABOVE_BL_SIGMA_M_S == 0.1
    assert hanna.ABOVE_BL_T_L_S == 100.0


def test_meander_constructor_validates_params():
    with pytest.raises(ValueError):
        HannaScheme(meander_coefficient=0.0)
    with pytest.raises(ValueError):
        HannaScheme(meander_stencil_radius=0)
    with pytest.raises(ValueError):
        HannaScheme(meander_timescale_seconds=0.0)


def test_convection_registry_and_constructor_checks():
    """reference tests/test_convection.py:50-56, 594-606."""
    from glide_amd.convection import (ConvectionScheme, EmanuelReducedConvection, NoConvection, register_scheme)
    from glide_amd.convection import get_scheme as get_conv
    from glide_amd.convection import list_schemes as list_conv

    names = list_conv()
    assert "none" in names and "emanuel_reduced" in names
    assert isinstance(get_conv("none"), NoConvection)
    e = get_conv("emanuel_reduced", closure_c=0.05)
    assert isinstance(e, EmanuelReducedConvection) and e.closure_c == 0.05 and e.columns == "bbox"
    assert e.required_met_keys() == ("t", "q") and get_conv("none").required_met_keys() == ()
    assert get_conv("none").initialize_state(3, device="cpu", dtype=None) == {}
    with pytest.raises(KeyError, match="Unknown convection scheme"):
        get_conv("kain_fritsch")
    for kw, msg in ((dict(closure_c=0.0), "closure_c"), (dict(trigger_dtv_k=-1.0), "trigger_dtv_k"),
                    (dict(min_cape_j_kg=-10.0), "min_cape_j_kg"), (dict(min_cloud_depth_m=-5.0), "min_cloud_depth_m"),
                    (dict(columns="rows"), "columns")):
        with pytest.raises(ValueError, match=msg):
            EmanuelReducedConvection(**kw)

    class Nameless(ConvectionScheme):
        name = ""

        def required_met_keys(self):
            return ()

        def maybe_convect(self, *a, **k):
            return None

    with pytest.raises(TypeError):
        register_scheme(Nameless)

    class Clash(Nameless):
        name = "none"

    with pytest.raises(ValueError, match="already registered"):
        register_scheme(Clash)


def test_noise_clock_keys():
    """Model-time step counters; repeats and new sweeps move the key's high word."""
    from datetime import datetime, timedelta, timezone

    import numpy as np

    from glide_amd.met import MetFieldMetadata
    from glide_amd.noise import NoiseClock, model_time_step

    t0 = datetime(2024, 1, 1, 6, tzinfo=timezone.utc)
    md = MetFieldMetadata(lon=np.zeros(2), lat=np.zeros(2), level=np.zeros(2), pressure_level_hpa=np.zeros(2),
                          time_start=t0, time_end=t0 + timedelta(hours=1), variable_units={})
    base = int(t0.timestamp()) // 60
    assert model_time_step(md, 0.5, 60.0) == base + 30
    assert model_time_step(md, 1.0 - 30 / 3600, 60.0) == base + 59  # the midpoint of the hour's last step
    c = NoiseClock()
    seq = [c.next(md, a, 60.0, seed=7) for a in (1.0 - 30 / 3600, 1.0 - 90 / 3600, 1.0 - 90 / 3600, 0.5, 0.9)]
    steps = [s for s, _ in seq]
    keys = [k for _, k in seq]
    assert steps == [base + 59, base + 58, base + 58, base + 30, base + 54]
    assert keys[0] == keys[1] == keys[3] == 7  # one backward sweep: the seed itself
    assert keys[2] == 7 ^ (1 << 32)  # a repeat at the same model time
    assert keys[4] == 7 ^ (1 << 48)  # time went forward: the next sweep
