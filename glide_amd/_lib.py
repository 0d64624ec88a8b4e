"""ctypes binding of ``libglide_hip.so`` (the C ABI in ``include/glide.h``).

The library is mandatory: :func:`lib` raises ``RuntimeError`` when it is not
built or cannot be loaded. There is no fallback path.
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libglide_hip.so"

ABI_VERSION = 8  # GLIDE_ABI_VERSION of include/glide.h

GLIDE_OK = 0
GLIDE_ERR_INVALID = -1
GLIDE_ERR_MEMORY = -2
GLIDE_ERR_DEVICE = -3

GLIDE_MET_F32 = 0
GLIDE_MET_BF16 = 1

F_MEANDER = 0x1
F_TL_FLOORS = 0x2
F_SL_OVERRIDE = 0x4
F_EMIT = 0x8
F_MASK_INPUT = 0x10
F_NO_TURB = 0x20
F_RELEASE_TERMINUS = 0x40
F_FIXED = 0x80
F_PRESCRIBED = 0x100

GLIDE_SLOT_FREE = 0
GLIDE_SLOT_RUNNING = 1
GLIDE_SLOT_FINISHED = 2
GLIDE_SLOTS_CONVECT = 0x1
GLIDE_ROW_WORDS = 66  # int32 words of the step cursor buffer (glide_step_table)
GLIDE_SHF_UPWARD_W_M2, GLIDE_SHF_DOWNWARD_W_M2, GLIDE_SHF_ACCUMULATED_J_M2 = 0, 1, 2
GLIDE_NOALLOC_MAX_PAIRS = 1 << 20  # pairs glide_footprint_accumulate_noalloc serves

vp = C.c_void_p


class GlideGrid(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("reserved", C.c_int32),
        ("lon0", C.c_double), ("lon1", C.c_double), ("lat0", C.c_double), ("lat1", C.c_double),
        ("levels", vp),
    ]


class GlideWindow(C.Structure):
    _fields_ = [("wind", vp), ("aux", vp), ("surf", vp), ("dtype", C.c_int32), ("reserved", C.c_int32)]


class GlideParticles(C.Structure):
    _fields_ = [
        ("n", C.c_int64), ("pos_stride", C.c_int64),
        ("lon", vp), ("lat", vp), ("z", vp),
        ("up", vp), ("vp", vp), ("wp", vp), ("um", vp), ("vm", vp),
        ("weight", vp), ("r_off", vp), ("we_off", vp), ("rel_idx", vp), ("uid", vp),
        ("alive", vp), ("active", vp), ("lo_off", vp), ("cell", vp),
    ]


class GlideFootprintGrid(C.Structure):
    _fields_ = [
        ("lon_min", C.c_double), ("lon_max", C.c_double), ("lat_min", C.c_double),
        ("lat_max", C.c_double), ("z_edges", vp),
        ("n_r", C.c_int32), ("n_t", C.c_int32), ("n_z", C.c_int32), ("n_y", C.c_int32),
        ("n_x", C.c_int32), ("reserved", C.c_int32),
        ("dom_lon_min", C.c_double), ("dom_lon_max", C.c_double), ("dom_lat_min", C.c_double),
        ("dom_lat_max", C.c_double), ("dom_alt_max", C.c_double), ("t_bin", vp),
        ("slab_off", vp), ("slab_nt", vp), ("slab_cells", C.c_int64),
    ]


class GlideStepParams(C.Structure):
    _fields_ = [
        ("dt", C.c_double), ("alpha", C.c_double), ("cursor_off", C.c_double),
        ("length_s", C.c_double), ("substep_c", C.c_double), ("z_ubl_m", C.c_double),
        ("meander_timescale_s", C.c_double), ("seed", C.c_uint64),
        ("step_index", C.c_int32), ("max_substeps", C.c_int32),
        ("flags", C.c_int32), ("block", C.c_int32),
        ("noise_hi", C.c_uint32), ("reserved", C.c_int32),
    ]


class GlideSlot(C.Structure):
    _fields_ = [
        ("t", C.c_double), ("term", C.c_double), ("rel_time", C.c_double),
        ("state", C.c_int32), ("k", C.c_int32), ("noise_hi", C.c_uint32), ("convect", C.c_int32),
        ("last_d", C.c_double), ("cursor_off", C.c_float), ("alpha", C.c_float), ("dt", C.c_float),
        ("noise_step", C.c_int32), ("bracket", C.c_int32), ("stepping", C.c_int32),
        ("escaped", C.c_uint64), ("active_steps", C.c_uint64),
    ]


class GlideMetRing(C.Structure):
    _fields_ = [
        ("wind", vp), ("aux", vp), ("surf", vp),
        ("wind_stride", C.c_int64), ("aux_stride", C.c_int64), ("surf_stride", C.c_int64),
        ("first_hour", C.c_int32), ("n_hours", C.c_int32), ("dtype", C.c_int32), ("reserved", C.c_int32),
    ]


class GlideEmit(C.Structure):
    _fields_ = [("keys", vp), ("vals", vp), ("escaped", vp), ("active", vp), ("fixed", vp), ("fixed_exp", vp),
                ("fixed_copies", C.c_int32), ("active_copies", C.c_int32)]


P = C.POINTER

# name -> (restype, argtypes); every symbol include/glide.h declares.
SIGNATURES = {
    "glide_abi_version": (C.c_int, []),
    "glide_last_error": (C.c_char_p, []),
    "glide_step": (C.c_int, [P(GlideStepParams), P(GlideParticles), P(GlideGrid), P(GlideWindow),
                             P(GlideFootprintGrid), P(GlideEmit), vp]),
    "glide_step_table": (C.c_int, [P(GlideStepParams), vp, C.c_int32, vp, P(GlideParticles), P(GlideGrid), P(GlideWindow),
                                   P(GlideFootprintGrid), P(GlideEmit), vp]),
    "glide_row_set": (C.c_int, [vp, C.c_int32, vp]),
    "glide_column_cape": (C.c_int, [vp, vp, vp, vp, C.c_double, vp, vp, C.c_int32, C.c_int32, C.c_int32,
                                    vp, vp, vp, vp, vp, vp]),
    "glide_footprint_accumulate_binned": (C.c_int, [vp, vp, C.c_int64, C.c_uint32, C.c_uint32, C.c_uint32, vp, vp,
                                                    vp, vp, vp, C.c_size_t, vp]),
    "glide_means_workspace_bytes": (C.c_size_t, []),
    "glide_fixed_workspace_bytes": (C.c_size_t, [C.c_int32]),
    "glide_fixed_finalize": (C.c_int, [vp, C.c_int32, C.c_int64, C.c_uint32, vp, vp, vp, vp, C.c_size_t, vp]),
    "glide_fixed_finalize_slabs": (C.c_int, [vp, C.c_int32, C.c_int64, vp, vp, C.c_uint32, C.c_int32, vp, C.c_int32,
                                             C.c_int32, vp, vp, vp, C.c_size_t, vp]),
    "glide_particle_means": (C.c_int, [P(GlideParticles), vp, C.c_int32, vp, vp, C.c_size_t, vp]),
    "glide_field_means": (C.c_int, [vp, C.c_int64, C.c_int32, vp, vp, C.c_size_t, vp]),
    "glide_slots_advance": (C.c_int, [vp, C.c_int32, C.c_double, C.c_int32, vp]),
    "glide_step_timing_arm": (C.c_int, [C.c_int32]),
    "glide_step_timing_read": (C.c_int, [vp, C.c_int32]),
    "glide_step_slots": (C.c_int, [P(GlideStepParams), vp, vp, P(GlideParticles), P(GlideGrid), P(GlideMetRing),
                                   P(GlideFootprintGrid), P(GlideEmit), vp]),
    "glide_footprint_keys": (C.c_int, [C.c_int64, C.c_int64, vp, vp, vp, vp, vp, vp, vp, C.c_double,
                                       P(GlideFootprintGrid), vp, vp, vp]),
    "glide_footprint_workspace_bytes": (C.c_size_t, [C.c_int64]),
    "glide_footprint_accumulate": (C.c_int, [vp, vp, C.c_int64, C.c_uint32, C.c_uint32, vp, vp, vp, vp,
                                             C.c_size_t, vp]),
    "glide_footprint_accumulate_noalloc": (C.c_int, [vp, vp, C.c_int64, C.c_uint32, C.c_uint32, vp, vp, vp, vp]),
    "glide_sort_workspace_bytes": (C.c_size_t, [C.c_int64]),
    "glide_sort_pairs": (C.c_int, [vp, vp, C.c_int64, C.c_int32, vp, vp, vp, C.c_size_t, vp]),
    "glide_reorder_keys": (C.c_int, [P(GlideStepParams), P(GlideParticles), P(GlideGrid), C.c_double, C.c_int32,
                                     vp, vp, vp, vp]),
    "glide_permute_particles": (C.c_int, [P(GlideParticles), P(GlideParticles), vp, vp]),
    "glide_resample_hour": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, vp, vp, C.c_int32,
                                      vp, vp, C.c_double, C.c_double, vp, C.c_int32, C.c_int32, C.c_int32,
                                      vp, vp, vp]),
    "glide_source_units": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, vp,
                                     vp, vp, vp, vp, C.c_int32, C.c_double, vp]),
    "glide_build_window": (C.c_int, [vp, vp, vp, vp, vp, vp, C.c_int32, C.c_int32, C.c_int32, C.c_double,
                                     C.c_int32, C.c_int32, vp, vp, vp, vp]),
    "glide_convect_columns": (C.c_int, [vp, vp, vp, C.c_double, vp, vp, C.c_int32, C.c_int32, C.c_int32,
                                        C.c_double, C.c_double, C.c_double, C.c_double, vp, vp, vp, vp, vp]),
    "glide_convect": (C.c_int, [P(GlideStepParams), P(GlideParticles), P(GlideGrid), vp, vp, vp, vp, vp, vp,
                                P(GlideFootprintGrid), P(GlideEmit), vp]),
    "glide_hanna_eval": (C.c_int, [C.c_int32, C.c_int64, vp, vp, vp]),
    "glide_windowed_std": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, vp, vp]),
    "glide_graph_begin": (C.c_int, [vp]),
    "glide_graph_end": (C.c_int, [vp, P(vp)]),
    "glide_graph_launch": (C.c_int, [vp, vp]),
    "glide_graph_destroy": (C.c_int, [vp]),
}

def window_bytes(ncell: int, met_dtype: int) -> int:
    """Bytes of one packed wind (or support) field of ``ncell`` met cells:
    bfloat16 records of 16 B per cell, or float32 plane A (16 B per cell) then
    plane B (8 B per cell)."""

    return int(ncell) * (16 if met_dtype == GLIDE_MET_BF16 else 24)


def unpack_window(wind, aux, nz: int, ny: int, nx: int, met_dtype: int):
    """Packed window fields (uint8 torch tensors, column-major cells
    [ny][nx][nz]) as float32 [nz, ny, nx, 6] tensors: wind channels u_s v_s w_s
    u_e v_e w_e, support channels sigma_w_ft T_Lw_ft rho drho_dz sigma_mu
    sigma_mv."""

    import torch

    ncell = nz * ny * nx
    out = []
    for buf, wind_order in ((wind, True), (aux, False)):
        if met_dtype == GLIDE_MET_BF16:
            u16 = buf[: ncell * 16].view(torch.int16).view(ncell, 8).to(torch.int32) & 0xFFFF
            rec = (u16 << 16).view(torch.float32)[:, :6]
        else:
            a = buf[: ncell * 16].view(torch.float32).view(ncell, 4)
            b = buf[ncell * 16: ncell * 24].view(torch.float32).view(ncell, 2)
            if wind_order:  # plane A u_s v_s u_e v_e, plane B w_s w_e
                rec = torch.stack([a[:, 0], a[:, 1], b[:, 0], a[:, 2], a[:, 3], b[:, 1]], dim=1)
            else:  # plane A sigma_w T_Lw rho drho_dz, plane B sigma_mu sigma_mv
                rec = torch.cat([a, b], dim=1)
        out.append(rec.reshape(ny, nx, nz, 6).permute(2, 0, 1, 3))
    return out[0], out[1]


_LIB = None


def lib() -> C.CDLL:
    """Load the HIP library (once). Raises RuntimeError if it is missing."""

    global _LIB
    if _LIB is not None:
        return _LIB
    path = Path(os.environ.get("GLIDE_HIP_LIB", LIB_PATH))
    if not path.exists():
        raise RuntimeError(
            f"{path} is not built; run `python -m glide_amd.build` (the HIP library is required)"
        )
    h = C.CDLL(str(path))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(h, name)
        fn.restype = res
        fn.argtypes = args
    if h.glide_abi_version() != ABI_VERSION:
        raise RuntimeError("libglide_hip.so ABI version mismatch")
    _LIB = h
    return h


def check(rc: int) -> None:
    """Map a glide status to the reference's exception types."""

    if rc == GLIDE_OK:
        return
    msg = (lib().glide_last_error() or b"").decode(errors="replace")
    if rc == GLIDE_ERR_INVALID:
        raise ValueError(msg)
    if rc == GLIDE_ERR_MEMORY:
        raise MemoryError(msg)
    raise RuntimeError(msg)


def ptr(t) -> int | None:
    """Device pointer of a torch tensor (None for None)."""

    if t is None:
        return None
    return t.data_ptr()


def stream_ptr(stream=None) -> int:
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)


_QUEUE_STREAMS: dict[int, "QueueStream"] = {}


class QueueStream:
    """A HIP stream on a hardware queue of its own, for work that must overlap
    the stream driving the step.

    A process has few hardware queues (``GPU_MAX_HW_QUEUES``, 4 by default)
    and the runtime shares them among its streams; torch's pool streams are
    spread over them in creation order, so a pool stream taken for the
    accumulation can sit on the same queue as the step's stream, where the two
    run one after the other instead of together. The runtime creates a
    CU-masked stream on a queue of its own, and a mask of every CU leaves the
    work where it would run anyway. Use ``queue_stream(device)``: one such
    stream per device serves the whole process until the interpreter exits,
    so no engine holds a queue of its own. ``torch`` is
    a torch handle for events."""

    def __init__(self, device: int = 0):
        import torch

        hip = C.CDLL("libamdhip64.so.7")  # the runtime torch and the library already use
        hip.hipExtStreamCreateWithCUMask.argtypes = [C.POINTER(C.c_void_p), C.c_uint32, C.POINTER(C.c_uint32)]
        hip.hipStreamSynchronize.argtypes = [C.c_void_p]
        hip.hipStreamDestroy.argtypes = [C.c_void_p]
        ncu = torch.cuda.get_device_properties(device).multi_processor_count
        words = (ncu + 31) // 32
        mask = (C.c_uint32 * words)(*([0xFFFFFFFF] * words))
        if ncu % 32:
            mask[words - 1] = (1 << (ncu % 32)) - 1
        raw = C.c_void_p()
        with torch.cuda.device(device):
            rc = hip.hipExtStreamCreateWithCUMask(C.byref(raw), words, mask)
        if rc != 0 or not raw.value:
            raise RuntimeError(f"hipExtStreamCreateWithCUMask failed (hipError {rc})")
        self.ptr = int(raw.value)
        self.torch = torch.cuda.ExternalStream(self.ptr, device=device)
        self._hip = hip

    def destroy(self) -> None:
        if self.ptr:
            ptr, self.ptr = C.c_void_p(self.ptr), 0
            self._hip.hipStreamSynchronize(ptr)
            self._hip.hipStreamDestroy(ptr)


def queue_stream(device: int = 0) -> QueueStream:
    """The process's stream on a hardware queue of its own for ``device``
    (destroyed, after its work ends, when the interpreter exits)."""

    s = _QUEUE_STREAMS.get(device)
    if s is None:
        if not _QUEUE_STREAMS:
            import atexit

            atexit.register(_destroy_queue_streams)
        s = _QUEUE_STREAMS[device] = QueueStream(device)
    return s


def _destroy_queue_streams() -> None:
    for s in _QUEUE_STREAMS.values():
        s.destroy()
    _QUEUE_STREAMS.clear()
