"""The Hanna (1982) / FLEXPART free functions, evaluated on the GPU.

Drop-in for the reference's module-level functions of
``lpdm.turbulence.hanna`` (src/lpdm/turbulence/hanna.py:142-189, 192-337,
340-426) and ``HannaScheme._windowed_std`` (hanna.py:1786-1830): the same
names and argument order, over torch tensors. Every function runs the float32
device code that the fused step kernel and the window build execute
(``glide_hanna_eval`` / ``glide_windowed_std``), so the reference's unit tests
of these functions exercise the engine's own arithmetic. Inputs are broadcast
to one shape and moved to the GPU as float32; results are float32 tensors on
that GPU.
"""

from __future__ import annotations

import ctypes as C

import torch

from glide_amd import _lib as L

# Module constants (reference hanna.py:47-134).
EARTH_ROTATION_RATE_S = 7.2921e-5
ABOVE_BL_SIGMA_M_S = 0.1
ABOVE_BL_T_L_S = 100.0
FT_KZ_FLOOR_M2_S = 0.1
FT_KZ_CEIL_M2_S = 50.0

_OPS = dict(coriolis=1, air_density=2, obukhov=3, convective_velocity=4, in_bl=5, surface_layer_sigma_w=6,
            potential_temperature=7, brunt_vaisala=8, richardson=9, ft_diffusivity=10, ft_sigma_tl=11)
_N_OUT = dict(in_bl=6, ft_sigma_tl=2)


def _device(args) -> torch.device:
    for a in args:
        if isinstance(a, torch.Tensor) and a.is_cuda:
            return a.device
    return torch.device("cuda", torch.cuda.current_device())


def _eval(name: str, *args):
    dev = _device(args)
    ts = [torch.as_tensor(a, device=dev) for a in args]
    shape = torch.broadcast_shapes(*(t.shape for t in ts))
    ins = [t.to(torch.float32).broadcast_to(shape).contiguous() for t in ts]
    n_out = _N_OUT.get(name, 1)
    outs = [torch.empty(shape, dtype=torch.float32, device=dev) for _ in range(n_out)]
    in_p = (C.c_void_p * 6)(*[t.data_ptr() for t in ins])
    out_p = (C.c_void_p * 6)(*[t.data_ptr() for t in outs])
    with torch.cuda.device(dev):
        L.check(L.lib().glide_hanna_eval(_OPS[name], ins[0].numel(), in_p, out_p, L.stream_ptr()))
    return outs[0] if n_out == 1 else tuple(outs)


def coriolis_parameter(lat_deg):
    """f = 2 Omega sin(lat) (hanna.py:142-145)."""
    return _eval("coriolis", lat_deg)


def air_density(sp_pa, t_k):
    """rho = p / (R_d T) (hanna.py:148-150)."""
    return _eval("air_density", sp_pa, t_k)


def obukhov_length(ustar, shf, t_k, sp_pa):
    """L = -u*^3 T rho c_p / (kappa g H), +inf when |kappa g H| <= 1e-10
    (hanna.py:153-172)."""
    return _eval("obukhov", ustar, shf, t_k, sp_pa)


def convective_velocity(blh, shf, t_k, sp_pa):
    """w* = (g max(h, 50) H / (T rho c_p))^(1/3) for H > 0, else 0
    (hanna.py:175-189)."""
    return _eval("convective_velocity", blh, shf, t_k, sp_pa)


def in_bl_sigma_TL(z, blh, ustar, w_star, h_over_L, lat):  # noqa: N802 (reference name)
    """(sigma_u, sigma_v, sigma_w, T_Lu, T_Lv, T_Lw) of the in-BL regime that
    h/L selects (hanna.py:192-337)."""
    return _eval("in_bl", z, blh, ustar, w_star, h_over_L, lat)


def surface_layer_sigma_w(z, ustar, L):  # noqa: N803 (reference name)
    """Surface-layer sigma_w (hanna.py:405-426)."""
    return _eval("surface_layer_sigma_w", z, ustar, L)


def potential_temperature(t_k, p_pa):
    """theta = T (p0 / p)^(R_d / c_p) (hanna.py:340-345)."""
    return _eval("potential_temperature", t_k, p_pa)


def brunt_vaisala_squared(theta, dtheta_dz):
    """N^2 = g / theta dtheta/dz (hanna.py:348-353)."""
    return _eval("brunt_vaisala", theta, dtheta_dz)


def gradient_richardson(n2, shear_sq):
    """Ri = N^2 / max(S^2, 1e-8) (hanna.py:356-361)."""
    return _eval("richardson", n2, shear_sq)


def free_trop_diffusivity(z, shear_mag, ri):
    """Richardson-closure K_z, clamped to [0.1, 50] m^2/s (hanna.py:364-385)."""
    return _eval("ft_diffusivity", z, shear_mag, ri)


def free_trop_sigma_TL(k_z, n2):  # noqa: N802 (reference name)
    """(sigma_w, T_Lw) with sigma_w^2 T_Lw = K_z (hanna.py:388-402)."""
    return _eval("ft_sigma_tl", k_z, n2)


def windowed_std(field, radius: int = 1):
    """Population std of each cell's edge-clipped (2r+1)^2 horizontal
    neighbourhood of a [Z, Y, X] field (HannaScheme._windowed_std,
    hanna.py:1786-1830)."""

    if field.ndim != 3:
        raise ValueError("field must be [Z, Y, X]")
    dev = _device((field,))
    f = torch.as_tensor(field, device=dev).to(torch.float32).contiguous()
    out = torch.empty_like(f)
    nz, ny, nx = (int(v) for v in f.shape)
    with torch.cuda.device(dev):
        L.check(L.lib().glide_windowed_std(f.data_ptr(), nz, ny, nx, int(radius), out.data_ptr(), L.stream_ptr()))
    return out
