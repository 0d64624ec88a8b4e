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
