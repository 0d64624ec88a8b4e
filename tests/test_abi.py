"""The C ABI: every symbol include/glide.h declares is exported by the built
library and bound by the ctypes layer, and the ctypes structs have the exact
C layout (checked against gcc's sizeof/offsetof of the header). No compute."""

from __future__ import annotations

import ctypes as C
import re
import shutil
import subprocess
from pathlib import Path

import pytest

from glide_amd import _lib as L

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "glide.h"

STRUCTS = {
    "GlideGrid": L.GlideGrid,
    "GlideWindow": L.GlideWindow,
    "GlideParticles": L.GlideParticles,
    "GlideFootprintGrid": L.GlideFootprintGrid,
    "GlideStepParams": L.GlideStepParams,
    "GlideEmit": L.GlideEmit,
    "GlideSlot": L.GlideSlot,
    "GlideMetRing": L.GlideMetRing,
}


def declared_functions() -> set[str]:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b(glide_[a-z_0-9]+)\s*\(", text))


def test_header_declarations_are_bound():
    decl = declared_functions()
    assert decl, "no declarations parsed"
    assert decl == set(L.SIGNATURES), decl.symmetric_difference(L.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert lib.glide_abi_version() == L.ABI_VERSION == 8


def test_library_missing_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(L, "_LIB", None)
    monkeypatch.setenv("GLIDE_HIP_LIB", str(tmp_path / "nope.so"))
    with pytest.raises(RuntimeError, match="not built"):
        L.lib()


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_struct_layout_matches_c(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    for sname, cls in STRUCTS.items():
        lines.append(f'printf("{sname} size %zu\\n", sizeof({sname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{sname}.{fname} %zu\\n", offsetof({sname}, {fname}));')
    lines.append("return 0;}")
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = dict(line.rsplit(" ", 1) for line in out if line)
    for sname, cls in STRUCTS.items():
        assert int(got[f"{sname} size"]) == C.sizeof(cls), sname
        for fname, _ in cls._fields_:
            assert int(got[f"{sname}.{fname}"]) == getattr(cls, fname).offset, (sname, fname)


def test_errors_map_to_reference_exceptions():
    """Invalid arguments return GLIDE_ERR_INVALID -> ValueError without touching
    a device (argument checks run before any launch)."""

    lib = L.lib()
    with pytest.raises(ValueError, match="NULL"):
        L.check(lib.glide_step(None, None, None, None, None, None, None))
    with pytest.raises(ValueError, match="workspace too small"):
        L.check(lib.glide_footprint_accumulate(1, 1, 10, 100, 100, 1, None, None, 1, 0, None))
    g = L.GlideFootprintGrid(n_r=1, n_t=1, n_z=0, n_y=1, n_x=1)
    with pytest.raises(ValueError, match="footprint dimensions"):
        L.check(lib.glide_footprint_keys(1, 1, 1, 1, 1, 1, 1, 1, 1, 60.0, C.byref(g), 1, 1, None))
    assert lib.glide_footprint_workspace_bytes(0) > 0
    assert lib.glide_footprint_workspace_bytes(1 << 20) >= 16 * (1 << 20)


def test_step_block_is_validated_before_any_launch():
    """GlideStepParams.block (threads per workgroup of the step launch) must be
    0 or a multiple of 64 in [64, 512]; a bad value is an argument error raised
    before anything reaches the device (the pointers here are never read)."""

    lib = L.lib()
    fake = 0x1000
    s = L.GlideParticles(n=64, pos_stride=1, lon=fake, lat=fake, z=fake, up=fake, vp=fake, wp=fake, um=fake,
                         vm=fake, r_off=fake, alive=fake)
    g = L.GlideGrid(nx=4, ny=4, nz=4, lon0=0.0, lon1=1.0, lat0=0.0, lat1=1.0, levels=fake)
    w = L.GlideWindow(wind=fake, aux=fake, surf=fake, dtype=L.GLIDE_MET_F32)
    for bad in (32, 100, 576, 1024, -64):
        p = L.GlideStepParams(dt=60.0, substep_c=0.5, max_substeps=6, block=bad)
        with pytest.raises(ValueError, match="step block"):
            L.check(lib.glide_step(C.byref(p), C.byref(s), C.byref(g), C.byref(w), None, None, None))


def test_step_table_row_count_is_validated_before_any_launch():
    """glide_step_table needs the table's row count (>= 1): the kernel steps
    nothing for a cursor outside [0, n_rows), so a table without rows is an
    argument error, raised before anything reaches the device (the pointers
    here are never read)."""

    lib = L.lib()
    fake = 0x1000
    s = L.GlideParticles(n=64, pos_stride=1, lon=fake, lat=fake, z=fake, up=fake, vp=fake, wp=fake, um=fake,
                         vm=fake, r_off=fake, alive=fake)
    g = L.GlideGrid(nx=4, ny=4, nz=4, lon0=0.0, lon1=1.0, lat0=0.0, lat1=1.0, levels=fake)
    w = L.GlideWindow(wind=fake, aux=fake, surf=fake, dtype=L.GLIDE_MET_F32)
    p = L.GlideStepParams(dt=60.0, substep_c=0.5, max_substeps=6)
    for bad in (0, -1):
        with pytest.raises(ValueError, match="n_rows"):
            L.check(lib.glide_step_table(C.byref(p), fake, bad, fake, C.byref(s), C.byref(g), C.byref(w), None,
                                         None, None))
