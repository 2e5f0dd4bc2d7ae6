"""CPU checks of the C ABI: the ctypes mirror matches include/sphgpu.h, the
built library loads and exports every declared entry point, and host-side
argument validation returns the documented codes without touching a GPU."""
from __future__ import annotations

import ctypes as C
import pathlib
import re
import subprocess

import pytest

from sphcode_amd import _abi

ROOT = pathlib.Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "sphgpu.h"
LIB = ROOT / "sphcode_amd" / "lib" / "libsphgpu.so"


def declared_functions() -> list[str]:
    text = HEADER.read_text()
    return sorted(set(re.findall(r"\b(sphgpu_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_expected_entry_points():
    names = declared_functions()
    for must in ("sphgpu_create", "sphgpu_upload", "sphgpu_download", "sphgpu_integrate",
                 "sphgpu_pre_interaction", "sphgpu_fluid_force", "sphgpu_gravity_force",
                 "sphgpu_timestep", "sphgpu_predict", "sphgpu_correct", "sphgpu_make_tree"):
        assert must in names


def test_library_exports_every_declared_symbol():
    assert LIB.exists(), "run python -m sphcode_amd.build"
    out = subprocess.run(["nm", "-D", "--defined-only", str(LIB)], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (sphgpu_[a-z_0-9]+)", out))
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing


def test_library_loads_and_binds():
    lib = C.CDLL(str(LIB))
    fns = _abi.bind(lib, "sphgpu_", _abi.GPU_ONLY)
    assert fns["version"]() >= 1


def _layout_program() -> str:
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "sphgpu.h"', 'int main(void){']
    for st, cls in (("sphgpu_params", _abi.Params), ("sphgpu_particles", _abi.Particles),
                    ("sphgpu_stats", _abi.Stats), ("sphgpu_timers", _abi.Timers)):
        lines.append(f'printf("{st} size %zu\\n", sizeof({st}));')
        for name, _ in cls._fields_:
            lines.append(f'printf("{st}.{name} %zu\\n", offsetof({st}, {name}));')
    lines.append("return 0;}")
    return "\n".join(lines)


def test_struct_layouts_match_header(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(_layout_program())
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    got = dict(line.rsplit(" ", 1) for line in out.strip().splitlines())
    for st, cls in (("sphgpu_params", _abi.Params), ("sphgpu_particles", _abi.Particles),
                    ("sphgpu_stats", _abi.Stats), ("sphgpu_timers", _abi.Timers)):
        assert int(got[f"{st} size"]) == C.sizeof(cls), st
        for name, _ in cls._fields_:
            assert int(got[f"{st}.{name}"]) == getattr(cls, name).offset, f"{st}.{name}"


@pytest.mark.parametrize("field,value,code", [
    ("dim", 4, _abi.E_INVALID),
    ("kernel", 7, _abi.E_INVALID),
    ("sph_type", 5, _abi.E_INVALID),
    ("tree_max_level", 30, _abi.E_INVALID),
    ("tree_leaf_particle_num", 0, _abi.E_INVALID),
])
def test_create_rejects_bad_parameters(field, value, code):
    """Mirrors the reference's parameter errors (src/solver.cpp:198-278)
    plus the key-width limit of the Morton tree; validation runs before any
    device call."""
    lib = C.CDLL(str(LIB))
    fns = _abi.bind(lib, "sphgpu_", _abi.GPU_ONLY)
    p = _abi.default_params(3)
    p.gamma = 5.0 / 3.0
    setattr(p, field, value)
    ctx = C.c_void_p()
    assert fns["create"](C.byref(p), C.byref(ctx)) == code
    assert fns["last_error"]()


def test_wendland_in_one_dimension_rejected():
    """The reference's C4Kernel asserts DIM != 1
    (include/kernel/wendland_kernel.hpp:25-28): its 1D normalisation is 0."""
    lib = C.CDLL(str(LIB))
    fns = _abi.bind(lib, "sphgpu_", _abi.GPU_ONLY)
    p = _abi.default_params(1)
    p.gamma = 1.4
    p.kernel = _abi.WENDLAND
    ctx = C.c_void_p()
    assert fns["create"](C.byref(p), C.byref(ctx)) == _abi.E_INVALID
    assert b"Wendland" in fns["last_error"]()


def test_time_dependent_av_alpha_order_rejected():
    """alphaMax < alphaMin is an error (src/solver.cpp:227-228)."""
    lib = C.CDLL(str(LIB))
    fns = _abi.bind(lib, "sphgpu_", _abi.GPU_ONLY)
    p = _abi.default_params(3)
    p.gamma = 1.4
    p.av_use_time_dependent = 1
    p.av_alpha_max = 0.1
    p.av_alpha_min = 1.0
    ctx = C.c_void_p()
    assert fns["create"](C.byref(p), C.byref(ctx)) == _abi.E_INVALID
    assert b"alphaMax" in fns["last_error"]()
