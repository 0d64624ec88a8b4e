"""Convection plugin interface and the HIP-backed Emanuel scheme.

Drop-in for the reference's ``ConvectionScheme`` ABC and registry
(src/lpdm/convection/base.py:32-128), ``NoConvection``
(convection/no_convection.py:23-44) and ``EmanuelReducedConvection``
(convection/emanuel.py:515-805), which the driver calls once per met-window
advance (main.py:1446-1458). ``maybe_convect`` takes the same arguments and
returns the same ``(particles, state)``.

``EmanuelReducedConvection`` runs as three device launches:

* ``glide_column_cape``: the surface-parcel lift (LCL, LNB, CAPE and the T_v
  excess at LCL+1, emanuel.py:282-399) in float64;
* ``glide_convect_columns``: the trigger (LCL and LNB exist, LNB > LCL, CAPE,
  T_v excess, cloud depth) and the closure m_b = c rho_LCL min(sqrt(2 CAPE), 5)
  times the 3600 s met interval (emanuel.py:629-679), and the layer masses;
* ``glide_convect``: per particle, the non-divergent mass-flux matrix of its
  column (emanuel.py:429-506), the diagonal closure and the backward transpose
  sampling with movers placed uniformly in the destination layer
  (emanuel.py:720-805), never materialising a matrix.

``columns="bbox"`` (default) is the reference's scheme: one column, the
bounding-box mean of the time-interpolated temperature, humidity and height
profiles (emanuel.py:593-627), flipped to ascending height when the met's
levels descend. ``columns="per_column"`` applies the same scheme to every met
column on its own (the parcel lift per column, each particle through its
nearest column's matrix).

Random numbers come from the engine's counter-based Philox stream (not torch's
generator), keyed by the seed (``seed``, or the given generator's initial
seed) and the particle id (``state["particle_id"]`` when present, else the
row index) and counted by model time (glide_amd.noise). A particle's draw
therefore does not depend on which others are active or on how many calls
came before.
"""

from __future__ import annotations

import ctypes as C
from abc import ABC, abstractmethod
from typing import ClassVar, TypeAlias

import numpy as np
import torch

from glide_amd import _lib as L
from glide_amd.met import HourlyMetTensors
from glide_amd.noise import NoiseClock, model_time_step  # noqa: F401  (model_time_step re-exported)
from glide_amd.turbulence import GPUEngine

ConvectionState: TypeAlias = dict[str, torch.Tensor]

Q_MIN_KG_KG = 0.0
MAX_LEVELS = 128  # the device column kernels' level limit


class ConvectionScheme(ABC):
    """Deep-convection parameterisation (reference convection/base.py:32-96)."""

    name: ClassVar[str]

    @abstractmethod
    def required_met_keys(self) -> tuple[str, ...]:
        """Met channels the scheme reads from ``HourlyMetTensors``."""

    def initialize_state(self, n_particles: int, *, device: torch.device, dtype: torch.dtype) -> ConvectionState:
        """Persistent per-scheme state; empty for stateless schemes."""

        del n_particles, device, dtype
        return {}

    @abstractmethod
    def maybe_convect(self, particles: torch.Tensor, state: ConvectionState, met_window: HourlyMetTensors, *,
                      t_alpha: float, dt_seconds: float, active_mask: torch.Tensor, engine: GPUEngine,
                      generator: torch.Generator | None = None) -> tuple[torch.Tensor, ConvectionState]:
        """Possibly move active particles vertically through convective columns;
        inactive particles are unchanged. Returns ``(particles, state)``."""


_REGISTRY: dict[str, type[ConvectionScheme]] = {}


def register_scheme(cls: type[ConvectionScheme]) -> type[ConvectionScheme]:
    """Register a scheme class under its ``name`` (reference base.py:102-113)."""

    if not getattr(cls, "name", ""):
        raise TypeError(f"{cls.__name__} must define a non-empty class-level `name` attribute")
    existing = _REGISTRY.get(cls.name)
    if existing is not None and existing is not cls:
        raise ValueError(f"Convection scheme name {cls.name!r} already registered to {existing.__name__}")
    _REGISTRY[cls.name] = cls
    return cls


def get_scheme(name: str, **kwargs: object) -> ConvectionScheme:
    """Construct a registered scheme by name (reference base.py:116-122)."""

    if name not in _REGISTRY:
        available = ", ".join(sorted(_REGISTRY)) or "<none>"
        raise KeyError(f"Unknown convection scheme {name!r}. Registered: {available}")
    return _REGISTRY[name](**kwargs)


def list_schemes() -> tuple[str, ...]:
    return tuple(sorted(_REGISTRY))


@register_scheme
class NoConvection(ConvectionScheme):
    """Pass-through: particles unchanged (reference no_convection.py:23-44)."""

    name: ClassVar[str] = "none"

    def required_met_keys(self) -> tuple[str, ...]:
        return ()

    def maybe_convect(self, particles, state, met_window, *, t_alpha, dt_seconds, active_mask, engine,
                      generator=None):
        del met_window, t_alpha, dt_seconds, active_mask, engine, generator
        return particles, state


@register_scheme
class EmanuelReducedConvection(ConvectionScheme):
    """Emanuel / Forster et al. (2007) reduced convection on the GPU (see the
    module docstring)."""

    name: ClassVar[str] = "emanuel_reduced"

    def __init__(self, closure_c: float = 0.03, trigger_dtv_k: float = 0.9, min_cape_j_kg: float = 50.0,
                 min_cloud_depth_m: float = 500.0, *, columns: str = "bbox", seed: int = 0) -> None:
        if closure_c <= 0:
            raise ValueError("closure_c must be > 0")
        if trigger_dtv_k < 0:
            raise ValueError("trigger_dtv_k must be >= 0")
        if min_cape_j_kg < 0:
            raise ValueError("min_cape_j_kg must be >= 0")
        if min_cloud_depth_m < 0:
            raise ValueError("min_cloud_depth_m must be >= 0")
        if columns not in ("bbox", "per_column"):
            raise ValueError("columns must be 'bbox' or 'per_column'")
        self.closure_c = float(closure_c)
        self.trigger_dtv_k = float(trigger_dtv_k)
        self.min_cape_j_kg = float(min_cape_j_kg)
        self.min_cloud_depth_m = float(min_cloud_depth_m)
        self.columns = columns
        self.seed = int(seed)
        self.last: dict | None = None  # the last call's column descriptors, heights and pressures (device)
        self.clock = NoiseClock()

    def required_met_keys(self) -> tuple[str, ...]:
        return ("t", "q")

    # -- columns ----------------------------------------------------------
    def _profiles(self, mw: HourlyMetTensors, t_alpha: float, device):
        """(t_s, t_e, q_s, q_e [Z, ny, nx] float32, alpha for the kernels,
        heights [Z] or [Z, ny, nx] float32, pressure hPa [Z] float64), ascending."""

        if mw.height_agl_m is None:
            raise ValueError("EmanuelReducedConvection requires HourlyMetTensors.height_agl_m "
                             "(3D geopotential height). The met reader provides it; synthetic readers must too.")
        f32 = torch.float32
        t_s, t_e = (a.to(device=device, dtype=f32) for a in mw.channel("t"))
        q_s, q_e = (a.to(device=device, dtype=f32) for a in mw.channel("q"))
        h = mw.height_agl_m.to(device=device, dtype=f32)
        p_hpa = torch.as_tensor(np.asarray(mw.metadata.pressure_level_hpa, np.float64), device=device)
        if self.columns == "bbox":
            # the reference's order: lerp per cell, then the bbox mean (emanuel.py:597-609)
            a = float(t_alpha)
            t_col = (t_s * (1.0 - a) + t_e * a).mean(dim=(1, 2))
            q_col = (q_s * (1.0 - a) + q_e * a).mean(dim=(1, 2)).clamp(min=Q_MIN_KG_KG)
            h_col = h.mean(dim=(1, 2))
            if not bool(h_col[-1] > h_col[0]):
                t_col, q_col, h_col, p_hpa = t_col.flip(0), q_col.flip(0), h_col.flip(0), p_hpa.flip(0)
            # one column, replicated over a 2 x 2 grid (the particle kernel's
            # smallest); the profiles are already at t_alpha
            rep = lambda c: c[:, None, None].expand(-1, 2, 2).contiguous()  # noqa: E731
            t3, q3 = rep(t_col), rep(q_col)
            return t3, t3, q3, q3, 0.0, h_col.contiguous(), p_hpa.contiguous()
        heights = h[:, 0, 0]
        if not bool(torch.all(heights[1:] > heights[:-1])):
            raise ValueError("per-column convection expects ascending terrain-following levels")
        if not torch.allclose(h, heights[:, None, None].expand_as(h), rtol=1e-6, atol=1e-3):
            raise ValueError("per-column convection expects height_agl_m to be one ladder broadcast over columns")
        return (t_s.contiguous(), t_e.contiguous(), q_s.clamp(min=Q_MIN_KG_KG).contiguous(),
                q_e.clamp(min=Q_MIN_KG_KG).contiguous(), float(t_alpha), heights.contiguous(), p_hpa.contiguous())

    def maybe_convect(self, particles: torch.Tensor, state: ConvectionState, met_window: HourlyMetTensors, *,
                      t_alpha: float, dt_seconds: float, active_mask: torch.Tensor, engine: GPUEngine,
                      generator: torch.Generator | None = None) -> tuple[torch.Tensor, ConvectionState]:
        """Backward convective redistribution of the active particles (reference
        emanuel.py:570-716). Updates ``particles`` in place (its altitude
        column only) and returns it with ``state``."""

        if particles.ndim != 2 or particles.shape[1] != 4 or particles.dtype != torch.float32:
            raise ValueError("particles must be a float32 (N, 4) [lon, lat, alt, weight] tensor")
        if not particles.is_cuda:
            raise ValueError("particles must be a CUDA tensor")
        n = int(particles.shape[0])
        if active_mask.shape != (n,):
            raise ValueError("active_mask must have shape (N,)")
        if not (0.0 <= float(t_alpha) <= 1.0):
            raise ValueError("t_alpha must be in [0, 1]")
        if float(dt_seconds) <= 0:
            raise ValueError("dt_seconds must be > 0")
        if not bool(torch.any(active_mask)):
            return particles, state
        del engine  # the device and dtype are the particles'
        device = particles.device
        t_s, t_e, q_s, q_e, alpha, heights, p_hpa = self._profiles(met_window, t_alpha, device)
        Z, ny, nx = (int(v) for v in t_s.shape)
        if not 2 <= Z <= MAX_LEVELS:
            raise ValueError(f"convection supports 2 to {MAX_LEVELS} levels, got {Z}")
        st = L.stream_ptr()
        lib = L.lib()
        ncol = ny * nx
        conv = torch.empty((4, ny, nx), dtype=torch.float32, device=device)
        L.check(lib.glide_column_cape(t_s.data_ptr(), t_e.data_ptr(), q_s.data_ptr(), q_e.data_ptr(), alpha,
                                      p_hpa.data_ptr(), heights.data_ptr(), Z, ny, nx, None, None, None, None,
                                      conv.data_ptr(), st))
        i_lcl = torch.empty(ncol, dtype=torch.int32, device=device)
        i_lnb = torch.empty(ncol, dtype=torch.int32, device=device)
        m_b = torch.empty(ncol, dtype=torch.float32, device=device)
        layer_mass = torch.empty(Z, dtype=torch.float32, device=device)
        L.check(lib.glide_convect_columns(conv.data_ptr(), t_s.data_ptr(), t_e.data_ptr(), alpha, p_hpa.data_ptr(),
                                          heights.data_ptr(), Z, ny, nx, self.closure_c, self.trigger_dtv_k,
                                          self.min_cape_j_kg, self.min_cloud_depth_m, i_lcl.data_ptr(),
                                          i_lnb.data_ptr(), m_b.data_ptr(), layer_mass.data_ptr(), st))
        self.last = {"i_lcl": i_lcl, "i_lnb": i_lnb, "m_b": m_b, "layer_mass": layer_mass, "conv": conv,
                     "heights": heights, "p_hpa": p_hpa}
        # movers change only the altitude column: a strided tensor is stepped through a contiguous copy
        work = particles if particles.stride(1) == 1 else particles.contiguous()
        am = active_mask.to(device=device, dtype=torch.bool).contiguous()
        pid = state.get("particle_id") if isinstance(state, dict) else None
        if pid is None:
            pid = torch.arange(n, dtype=torch.int32, device=device)
        elif pid.shape != (n,) or pid.dtype != torch.int32 or pid.device != device:
            raise ValueError(f"state['particle_id'] must be an int32 (N,) tensor on {device}")
        pid = pid.contiguous()
        base = work.data_ptr()
        parts = L.GlideParticles(n=n, pos_stride=work.stride(0), lon=base, lat=base + 4, z=base + 8, up=None,
                                 vp=None, wp=None, um=None, vm=None, weight=base + 12, r_off=None, we_off=None,
                                 rel_idx=None, uid=pid.data_ptr(), alive=None, active=am.data_ptr(), lo_off=None)
        md = met_window.metadata
        lon = np.asarray(md.lon, np.float64)
        lat = np.asarray(md.lat, np.float64)
        if self.columns == "bbox":  # any span: every column of the 2 x 2 grid holds the same descriptors
            g = L.GlideGrid(nx=2, ny=2, nz=Z, reserved=0, lon0=0.0, lon1=1.0, lat0=0.0, lat1=1.0,
                            levels=heights.data_ptr())
        else:
            g = L.GlideGrid(nx=nx, ny=ny, nz=Z, reserved=0, lon0=float(lon[0]), lon1=float(lon[-1]),
                            lat0=float(lat[0]), lat1=float(lat[-1]), levels=heights.data_ptr())
        seed = self.seed if generator is None else int(generator.initial_seed())
        step_index, key = self.clock.next(md, t_alpha, dt_seconds, seed)
        p = L.GlideStepParams(dt=float(dt_seconds), alpha=float(t_alpha), cursor_off=0.0, length_s=0.0,
                              substep_c=0.5, z_ubl_m=0.0, meander_timescale_s=1.0, seed=key,
                              step_index=step_index, max_substeps=1, flags=L.F_MASK_INPUT, block=0)
        L.check(lib.glide_convect(C.byref(p), C.byref(parts), C.byref(g), i_lcl.data_ptr(), i_lnb.data_ptr(),
                                  m_b.data_ptr(), layer_mass.data_ptr(), None, None, None, None, st))
        if work is not particles:
            particles[:, 2].copy_(work[:, 2])
        return particles, state
