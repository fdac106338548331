"""C-ABI boundary checks that need no GPU: the built library loads and exports exactly the
entry points include/relight.h declares, and the Python bindings declare the same set."""
import ctypes
import os
import re
import subprocess

import pytest

from relight_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "relight.h")


def header_symbols():
    text = open(HEADER).read()
    return set(re.findall(r"^\s*(?:int32_t|const char\s*\*)\s*(rl_\w+)\s*\(", text, re.M))


def test_header_declares_entry_points():
    syms = header_symbols()
    assert {"rl_create", "rl_scene_end", "rl_bake_direct", "rl_bake_ambient_occlusion",
            "rl_bake_batch", "rl_trace_segments", "rl_lightmap_read_hdr"} <= syms


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_capi.LIB_PATH):
        pytest.skip("librelight_amd.so not built")
    out = subprocess.run(["nm", "-D", "--defined-only", _capi.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\sT\s+(rl_\w+)$", out, re.M))
    missing = header_symbols() - exported
    assert not missing, f"not exported: {sorted(missing)}"
    lib = ctypes.CDLL(_capi.LIB_PATH)
    for name in header_symbols():
        assert hasattr(lib, name)


def test_bindings_cover_the_header():
    assert {name for name, _, _ in _capi.SIGNATURES} == header_symbols()


def test_struct_layouts_match_header():
    # rl_light_desc: 2 ints, 3 float[3], intensity, has_attenuation, 6 floats, 2 pointers, 2 uints.
    assert ctypes.sizeof(_capi.rl_light_desc) == 4 * 2 + 36 + 4 + 4 + 24 + 4 + 16 + 8
    assert ctypes.sizeof(_capi.rl_ao_settings) == 24
    assert ctypes.sizeof(_capi.rl_hit) == 24
    assert ctypes.sizeof(_capi.rl_stats) == 40 + 16 * 8 + 4 * 16
