"""CPU-side checks of the C ABI: the library loads and exports every entry
point that include/ear.h declares; the ctypes mirror matches the header; host
error paths work without a device (no compute call is made here)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from ear_amd import _ffi as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ear.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ear_[a-z_0-9]+)\s*\(", src)))


def test_header_declares_what_ffi_binds():
    assert declared_functions() == sorted(F.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(F.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name


STRUCTS = ["ear_material", "ear_bvh_opts", "ear_scene_info", "ear_recorder", "ear_source", "ear_render_desc", "ear_backward_item",
           "ear_hist", "ear_stats", "ear_tuning"]


def test_struct_layouts_match_the_c_compiler(tmp_path):
    """sizeof / offsetof of every ABI struct as gcc lays them out == the ctypes mirror."""
    import subprocess

    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "ear.h"', "int main(void) {"]
    for st in STRUCTS:
        lines.append(f'printf("{st} size %zu\\n", sizeof({st}));')
        for field, _ in getattr(F, st)._fields_:
            lines.append(f'printf("{st} {field} %zu\\n", offsetof({st}, {field}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    for line in filter(None, out):
        st, field, val = line.split()
        cls = getattr(F, st)
        got = C.sizeof(cls) if field == "size" else getattr(cls, field).offset
        assert got == int(val), line


def test_status_strings_and_abi_version():
    lib = F.load()
    assert lib.ear_abi_version() == 7
    assert lib.ear_status_string(0) == b"ok"
    assert lib.ear_status_string(2) == b"Invalid material settings"
    assert lib.ear_status_string(99) == b"unknown status"


def test_null_arguments_are_rejected_without_a_device():
    lib = F.load()
    assert lib.ear_ctx_create(0, None) == 1
    assert lib.ear_ctx_destroy(None) == 1
    assert lib.ear_render(None, None, None, None, None) == 1
    assert lib.ear_scene_destroy(None) == 1
    assert lib.ear_ctx_last_error(None) == b""


def _has_device():
    try:
        hip = C.CDLL("libamdhip64.so")
        n = C.c_int(0)
        return hip.hipGetDeviceCount(C.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


@pytest.mark.skipif(_has_device(), reason="checks the no-device path")
def test_context_without_device_reports_no_device():
    lib = F.load()
    h = C.c_void_p()
    assert lib.ear_ctx_create(0, C.byref(h)) == 4  # EAR_ERR_NO_DEVICE
    from ear_amd import api

    with pytest.raises(api.EarError) as e:
        api.Context(0)
    assert e.value.status_name == "EAR_ERR_NO_DEVICE"
