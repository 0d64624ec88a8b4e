"""Turbulence plugin interface and the HIP-backed Hanna scheme.

Drop-in for the reference's ``TurbulenceScheme`` ABC and registry
(src/lpdm/turbulence/base.py:26-122) and ``HannaScheme`` (turbulence/hanna.py
:435-1978). ``HannaScheme.step`` takes the same arguments and returns the same
``(particles, state)``; the whole static step core (RK2 advection through the
time-interpolated winds, the Hanna/FLEXPART sigma and T_L with the
free-troposphere closure, the well-mixed and density drifts, the substepped
exact-OU update with clipping and reflection, meander and the active-mask gate)
is one fused HIP kernel launch, ``glide_step``. Like the reference's static
path it folds advection into the step (``step_includes_advection`` is True).

Per window, the met is staged once into address-stable packed device records
(the reference's ``_to_static_buffer`` lesson, hanna.py:1652-1680): winds at the
two bracket ends, the surface fields, and the window-midpoint free-troposphere,
density and meander stacks built by ``glide_build_window``.

Noise is the engine's counter-based Philox stream keyed by (seed, particle id)
and counted by (step, substep), so a step's result does not depend on which
particles are active or on launch order. The particle id is
``state["particle_id"]`` (int32) when the state holds one, else the row index;
the step counter is model time (glide_amd.noise), so it does not depend on how
many calls came before.
"""

from __future__ import annotations

import ctypes as C
from abc import ABC, abstractmethod
from typing import ClassVar, TypeAlias

import numpy as np
import torch

from glide_amd import _lib as L
from glide_amd.met import HourlyMetTensors
from glide_amd.noise import NoiseClock

TurbulenceState: TypeAlias = dict[str, torch.Tensor]

MEANDER_COEFFICIENT_DEFAULT = 0.16
MEANDER_STENCIL_RADIUS_DEFAULT = 1
MEANDER_TIMESCALE_S_DEFAULT = 1800.0
SUBSTEP_C_DEFAULT = 0.5
MAX_SUBSTEPS_DEFAULT = 50
Z_UBL_DEFAULT_M = 2.0


class GPUEngine:
    """Device/dtype holder passed through the scheme API (reference
    gpu_engine.py:124-201). The step arithmetic lives in the HIP kernels."""

    def __init__(self, *, device: torch.device | str = "cuda", dtype: torch.dtype = torch.float32) -> None:
        self.device = torch.device(device)
        self.dtype = dtype


class TurbulenceScheme(ABC):
    """reference turbulence/base.py:26-90."""

    name: ClassVar[str]

    @abstractmethod
    def required_met_keys(self) -> tuple[str, ...]:
        ...

    @abstractmethod
    def initialize_state(self, n_particles: int, *, device: torch.device, dtype: torch.dtype) -> TurbulenceState:
        ...

    @abstractmethod
    def step(self, particles: torch.Tensor, state: TurbulenceState, met_window: HourlyMetTensors,
             t_alpha: float, dt_seconds: float, active_mask: torch.Tensor,
             engine: GPUEngine) -> tuple[torch.Tensor, TurbulenceState]:
        ...

    def step_includes_advection(self, engine: GPUEngine) -> bool:
        del engine
        return False


_REGISTRY: dict[str, type[TurbulenceScheme]] = {}


def register_scheme(cls: type[TurbulenceScheme]) -> type[TurbulenceScheme]:
    if not getattr(cls, "name", ""):
        raise TypeError(f"{cls.__name__} must define a non-empty class-level `name` attribute")
    existing = _REGISTRY.get(cls.name)
    if existing is not None and existing is not cls:
        raise ValueError(f"Turbulence scheme name {cls.name!r} already registered to {existing.__name__}")
    _REGISTRY[cls.name] = cls
    return cls


def get_scheme(name: str, **kwargs: object) -> TurbulenceScheme:
    if name not in _REGISTRY:
        available = ", ".join(sorted(_REGISTRY)) or "<none>"
        raise KeyError(f"Unknown turbulence scheme {name!r}. Registered schemes: {available}")
    return _REGISTRY[name](**kwargs)


def list_schemes() -> tuple[str, ...]:
    return tuple(sorted(_REGISTRY))


@register_scheme
class HannaScheme(TurbulenceScheme):
    """Hanna 1982 / FLEXPART turbulence on the GPU (see module docstring)."""

    name: ClassVar[str] = "hanna_1982"
    DRIFT_FD_DELTA_M: ClassVar[float] = 1.0

    def __init__(self, *, meander_enabled: bool = False,
                 meander_coefficient: float = MEANDER_COEFFICIENT_DEFAULT,
                 meander_stencil_radius: int = MEANDER_STENCIL_RADIUS_DEFAULT,
                 meander_timescale_seconds: float = MEANDER_TIMESCALE_S_DEFAULT,
                 substep_c: float = SUBSTEP_C_DEFAULT, max_substeps: int = MAX_SUBSTEPS_DEFAULT,
                 z_ubl_m: float = Z_UBL_DEFAULT_M, static_substeps: bool | None = None,
                 flexpart_tl_floors: bool = True, surface_layer_override: bool = False,
                 seed: int = 0, met_storage: str = "f32") -> None:
        if meander_coefficient <= 0:
            raise ValueError("meander_coefficient must be > 0")
        if meander_stencil_radius < 1:
            raise ValueError("meander_stencil_radius must be >= 1")
        if meander_timescale_seconds <= 0:
            raise ValueError("meander_timescale_seconds must be > 0")
        if substep_c <= 0:
            raise ValueError("substep_c must be > 0")
        if max_substeps < 1:
            raise ValueError("max_substeps must be >= 1")
        if z_ubl_m < 0:
            raise ValueError("z_ubl_m must be >= 0")
        if met_storage not in ("f32", "bf16"):
            raise ValueError("met_storage must be 'f32' or 'bf16'")
        # The reference's static path (a fixed max_substeps loop, finished
        # particles advanced by sub_dt = 0) and dynamic path (each particle's own
        # k substeps over the active subset) give the same result here: the fused
        # kernel runs each particle's own k substeps and the counter-based noise
        # does not depend on the loop shape. Kept for the interface.
        self.static_substeps = static_substeps
        self.meander_enabled = bool(meander_enabled)
        self.meander_coefficient = float(meander_coefficient)
        self.meander_stencil_radius = int(meander_stencil_radius)
        self.meander_timescale_seconds = float(meander_timescale_seconds)
        self.substep_c = float(substep_c)
        self.max_substeps = int(max_substeps)
        self.z_ubl_m = float(z_ubl_m)
        self.flexpart_tl_floors = bool(flexpart_tl_floors)
        self.surface_layer_override = bool(surface_layer_override)
        self.seed = int(seed)
        # Philox counter word 1. None (default): model time (glide_amd.noise);
        # an int set by the caller is used, with the seed as the key, and
        # advanced by one per step() instead.
        self.step_index: int | None = None
        self.clock = NoiseClock()
        self.met_dtype = L.GLIDE_MET_BF16 if met_storage == "bf16" else L.GLIDE_MET_F32
        self._win_key = None
        self._win = None
        self._uid = None

    @staticmethod
    def _windowed_std(field: torch.Tensor, radius: int) -> torch.Tensor:
        """The meander stencil std (hanna.py:1786-1830), on the GPU."""

        from glide_amd.hanna import windowed_std

        return windowed_std(field, radius)

    def required_met_keys(self) -> tuple[str, ...]:
        return ("t", "ustar", "shf")

    def initialize_state(self, n_particles: int, *, device: torch.device, dtype: torch.dtype) -> TurbulenceState:
        st = {k: torch.zeros(n_particles, device=device, dtype=dtype) for k in ("u_prime", "v_prime", "w_prime")}
        if self.meander_enabled:
            st["u_meander"] = torch.zeros(n_particles, device=device, dtype=dtype)
            st["v_meander"] = torch.zeros(n_particles, device=device, dtype=dtype)
        return st

    def step_includes_advection(self, engine: GPUEngine) -> bool:
        del engine
        return True

    def _flags(self) -> int:
        f = L.F_MASK_INPUT
        if self.meander_enabled:
            f |= L.F_MEANDER
        if self.flexpart_tl_floors:
            f |= L.F_TL_FLOORS
        if self.surface_layer_override:
            f |= L.F_SL_OVERRIDE
        return f

    def _window(self, mw: HourlyMetTensors, device: torch.device):
        """Packed device window for ``mw``, built once per window start time."""

        key = (mw.metadata.time_start, device)
        if self._win_key == key:
            return self._win
        md = mw.metadata
        level = np.asarray(md.level, np.float64)
        if level.ndim != 1 or level.size < 2 or not np.all(np.diff(level) > 0):
            raise ValueError("HannaScheme expects an ascending terrain-following level ladder")
        if mw.height_agl_m is None:
            raise ValueError("HannaScheme free-troposphere turbulence requires HourlyMetTensors.height_agl_m")
        Z, Y, X = (int(v) for v in mw.hour_start.shape[1:])
        if Z > 128:
            raise ValueError("at most 128 met levels supported")
        h = mw.height_agl_m.detach().to("cpu", torch.float64).numpy()
        if not np.allclose(h, level[:, None, None], rtol=1e-6, atol=1e-3):
            raise ValueError("height_agl_m must be the terrain-following ladder broadcast over columns")
        f32 = torch.float32

        def stack3(hour):
            names = ("u", "v", "w", "t")
            parts = [hour[mw._channel_index(n)] for n in names]
            q = hour[mw._channel_index("q")] if "q" in mw.channel_names else torch.zeros_like(parts[0])
            return torch.stack(parts + [q]).to(device=device, dtype=f32).contiguous()

        def stack2(hour):
            return torch.stack([hour[mw._channel_index(n)][0] for n in ("blh", "sp", "ustar", "shf")]).to(
                device=device, dtype=f32).contiguous()

        s3, e3 = stack3(mw.hour_start), stack3(mw.hour_end)
        s2, e2 = stack2(mw.hour_start), stack2(mw.hour_end)
        p = torch.as_tensor(np.asarray(md.pressure_level_hpa, np.float64), device=device)
        lev = torch.as_tensor(level.astype(np.float32), device=device)
        if self._win is not None and self._win["shape"] == (Z, Y, X) and self._win["device"] == device:
            wind, aux, surf = self._win["wind"], self._win["aux"], self._win["surf"]  # address-stable
        else:
            wind = torch.empty(L.window_bytes(Z * Y * X, self.met_dtype), dtype=torch.uint8, device=device)
            aux = torch.empty(L.window_bytes(Z * Y * X, self.met_dtype), dtype=torch.uint8, device=device)
            surf = torch.empty(Y * X * 12 * 4, dtype=torch.uint8, device=device)  # float in both layouts
        L.check(L.lib().glide_build_window(
            s3.data_ptr(), e3.data_ptr(), s2.data_ptr(), e2.data_ptr(), p.data_ptr(), lev.data_ptr(), Z, Y, X,
            self.meander_coefficient, self.meander_stencil_radius, self.met_dtype, wind.data_ptr(),
            aux.data_ptr(), surf.data_ptr(), L.stream_ptr()))
        lon = np.asarray(md.lon, np.float64)
        lat = np.asarray(md.lat, np.float64)
        grid = L.GlideGrid(nx=X, ny=Y, nz=Z, reserved=0, lon0=float(lon[0]), lon1=float(lon[-1]),
                           lat0=float(lat[0]), lat1=float(lat[-1]), levels=lev.data_ptr())
        win = L.GlideWindow(wind=wind.data_ptr(), aux=aux.data_ptr(), surf=surf.data_ptr(),
                            dtype=self.met_dtype, reserved=0)
        self._win = dict(shape=(Z, Y, X), device=device, wind=wind, aux=aux, surf=surf, lev=lev, grid=grid,
                         win=win, keep=(s3, e3, s2, e2, p))
        self._win_key = key
        return self._win

    def step(self, particles: torch.Tensor, state: TurbulenceState, met_window: HourlyMetTensors,
             t_alpha: float, dt_seconds: float, active_mask: torch.Tensor,
             engine: GPUEngine) -> tuple[torch.Tensor, TurbulenceState]:
        """One backward step of every active particle (reference hanna.py:580-633).
        Updates ``particles`` and ``state`` in place; inactive particles are unchanged."""

        if particles.ndim != 2 or particles.shape[1] != 4 or particles.dtype != torch.float32:
            raise ValueError("particles must be a float32 (N, 4) [lon, lat, alt, weight] tensor")
        if not particles.is_cuda:
            raise ValueError("particles must be a CUDA tensor")
        n = int(particles.shape[0])
        if active_mask.shape != (n,):
            raise ValueError("active_mask must have shape (N,)")
        device = particles.device
        win = self._window(met_window, device)
        keys = ["u_prime", "v_prime", "w_prime"] + (["u_meander", "v_meander"] if self.meander_enabled else [])
        for k in keys:
            t = state[k]
            if t.shape != (n,) or t.dtype != torch.float32 or not t.is_contiguous() or t.device != device:
                raise ValueError(f"state[{k!r}] must be a contiguous float32 (N,) tensor on {device}")
        am = active_mask.to(device=device, dtype=torch.bool).contiguous()
        pid = state.get("particle_id")
        if pid is not None:
            if pid.shape != (n,) or pid.dtype != torch.int32 or pid.device != device:
                raise ValueError(f"state['particle_id'] must be an int32 (N,) tensor on {device}")
            uid = pid.contiguous()
        else:  # the row index
            if self._uid is None or self._uid.numel() != n or self._uid.device != device:
                self._uid = torch.arange(n, dtype=torch.int32, device=device)
            uid = self._uid
        # rows of 4 adjacent floats at any row stride; another layout through a contiguous copy
        work = particles if particles.stride(1) == 1 else particles.contiguous()
        if self.step_index is None:
            step_index, key = self.clock.next(met_window.metadata, t_alpha, dt_seconds, self.seed)
        else:
            step_index, key = int(self.step_index), self.seed
        base = work.data_ptr()
        parts = L.GlideParticles(
            n=n, pos_stride=work.stride(0), lon=base, lat=base + 4, z=base + 8,
            up=state["u_prime"].data_ptr(), vp=state["v_prime"].data_ptr(), wp=state["w_prime"].data_ptr(),
            um=state["u_meander"].data_ptr() if self.meander_enabled else None,
            vm=state["v_meander"].data_ptr() if self.meander_enabled else None,
            weight=base + 12, r_off=None, we_off=None, rel_idx=None, uid=uid.data_ptr(), alive=None,
            active=am.data_ptr())
        p = L.GlideStepParams(dt=float(dt_seconds), alpha=float(t_alpha), cursor_off=0.0, length_s=0.0,
                              substep_c=self.substep_c, z_ubl_m=self.z_ubl_m,
                              meander_timescale_s=self.meander_timescale_seconds, seed=key,
                              step_index=step_index, max_substeps=self.max_substeps, flags=self._flags(),
                              block=0)
        L.check(L.lib().glide_step(C.byref(p), C.byref(parts), C.byref(win["grid"]), C.byref(win["win"]),
                                   None, None, L.stream_ptr()))
        if work is not particles:
            particles.copy_(work)
        if self.step_index is not None:
            self.step_index += 1
        return particles, state
