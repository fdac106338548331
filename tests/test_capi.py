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


# Every struct of the header with its ctypes mirror; the C compiler's offsetof decides the layout.
_STRUCTS = ["rl_light_desc", "rl_ao_settings", "rl_hit", "rl_indirect_settings", "rl_tuning", "rl_bake_iterator", "rl_progress",
            "rl_solve_info", "rl_stats", "rl_bake_item", "rl_server_config", "rl_job_desc",
            "rl_job_status", "rl_server_stats"]


def test_struct_layouts_match_header(tmp_path):
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "relight.h"', "int main(void) {"]
    for name in _STRUCTS:
        py = getattr(_capi, name)
        lines.append(f'printf("{name} %zu\\n", sizeof({name}));')
        for field, _ in py._fields_:
            lines.append(f'printf("{name}.{field} %zu\\n", offsetof({name}, {field}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)],
                   check=True)
    got = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                         check=True).stdout.splitlines())
    for name in _STRUCTS:
        py = getattr(_capi, name)
        assert int(got[name]) == ctypes.sizeof(py), name
        for field, _ in py._fields_:
            assert int(got[f"{name}.{field}"]) == getattr(py, field).offset, f"{name}.{field}"

def test_argument_errors_without_a_device():
    """Entry points reject NULL handles before any device work (callable without a GPU)."""
    if not os.path.exists(_capi.LIB_PATH):
        pytest.skip("librelight_amd.so not built")
    lib = _capi.load()
    cfg = _capi.rl_server_config(1 << 30, 4, 1 << 20)
    out = ctypes.c_void_p()
    assert lib.rl_server_create(None, ctypes.byref(cfg), ctypes.byref(out)) == _capi.RL_INVALID_ARGUMENT
    assert b"NULL" in lib.rl_last_error()
    assert lib.rl_bake_items(None, None, 0) == _capi.RL_INVALID_ARGUMENT
    assert lib.rl_server_destroy(None) == _capi.RL_SUCCESS
    st = _capi.rl_job_status()
    assert lib.rl_server_poll(None, 1, ctypes.byref(st)) == _capi.RL_INVALID_ARGUMENT


def test_pass_and_tuning_argument_errors_without_a_device():
    """Recorded-pass and tuning entry points reject NULL handles and bad values before any device
    work (callable without a GPU)."""
    if not os.path.exists(_capi.LIB_PATH):
        pytest.skip("librelight_amd.so not built")
    lib = _capi.load()
    out = ctypes.c_void_p()
    assert lib.rl_pass_create(None, None, None, 0, 1, None, ctypes.byref(out)) == _capi.RL_INVALID_ARGUMENT
    assert lib.rl_pass_launch(None, None) == _capi.RL_INVALID_ARGUMENT
    assert lib.rl_pass_destroy(None) == _capi.RL_SUCCESS
    t = _capi.rl_tuning(_capi.RL_AO_KERNEL_REFILL, 16, 8, 8, _capi.RL_FG_KERNEL_AUTO)
    assert lib.rl_set_tuning(None, ctypes.byref(t)) == _capi.RL_INVALID_ARGUMENT
    assert lib.rl_get_tuning(None, ctypes.byref(t)) == _capi.RL_INVALID_ARGUMENT
