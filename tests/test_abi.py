"""The C-ABI library: it loads without a GPU, exports every symbol that
include/bmfr_hip.h declares, validates parameters the way the reference's
configuration constrains them, and derives the reference's sizes
(opencl/bmfr.cpp:104-118).  No compute call is made here."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import bmfr_amd
from bmfr_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "bmfr_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(bmfr_[a-z_0-9]+)\s*\(",
                                 text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    bound = {n for n, _, _ in _lib.SIGNATURES}
    assert set(names) == bound, set(names) ^ bound


def test_api_version_and_status_strings():
    lib = _lib.load()
    assert lib.bmfr_api_version() == 3
    for st in range(6):
        assert lib.bmfr_status_string(st)
    assert lib.bmfr_status_string(99) == b"unknown status"


@pytest.mark.parametrize("w,h,ws,hs,blocks", [
    (256, 256, (256, 256), (288, 288), 81),
    (1280, 720, (1280, 736), (1312, 768), 984),   # 41x24, not the 943 of the config text
    (1920, 1080, (1920, 1088), (1952, 1120), 2135),
    (640, 360, (640, 384), (672, 416), 273),
    (2560, 1440, (2560, 1440), (2592, 1472), 3726),
    (47, 47, (64, 64), (96, 96), 9),
])
def test_sizes_match_reference_formulas(w, h, ws, hs, blocks):
    p = bmfr_amd.default_params(w, h)
    s = bmfr_amd.compute_sizes(p)
    assert (s.workset_width, s.workset_height) == ws
    assert (s.margin_width, s.margin_height) == hs
    assert s.blocks == blocks == s.blocks_x * s.blocks_y
    assert s.pixels == w * h
    assert s.padded_pixels == hs[0] * hs[1]


def test_reference_defaults():
    p = bmfr_amd.default_params(1280, 720)
    assert p.blend_alpha == np.float32(0.2)
    assert p.second_blend_alpha == np.float32(0.1)
    assert p.taa_blend_alpha == np.float32(0.2)
    assert p.noise_amount == 1e-2
    assert p.tmp_storage == bmfr_amd.STORAGE_F16   # USE_HALF_PRECISION_IN_TMP_DATA 1
    assert p.storage == bmfr_amd.STORAGE_F32


@pytest.mark.parametrize("kw", [
    dict(width=32), dict(height=46), dict(width=16), dict(storage=7), dict(tmp_storage=-1), dict(qr_mode=9),
    dict(blend_alpha=1.5), dict(taa_blend_alpha=float("nan")),
])
def test_invalid_params_rejected(kw):
    w = kw.pop("width", 64)
    h = kw.pop("height", 64)
    with pytest.raises(bmfr_amd.BmfrError) as e:
        bmfr_amd.default_params(w, h, **kw)
    assert e.value.status == bmfr_amd.ERR_INVALID_ARGUMENT


def test_null_arguments_return_status():
    lib = _lib.load()
    assert lib.bmfr_default_params(64, 64, None) == bmfr_amd.ERR_INVALID_ARGUMENT
    assert lib.bmfr_fitter(None, 0) == bmfr_amd.ERR_INVALID_ARGUMENT
    assert lib.bmfr_stream_close(None) == bmfr_amd.OK
    assert b"null" in lib.bmfr_last_error()
    d = ctypes.c_double()
    assert lib.bmfr_measure_lds_read_bandwidth(0, ctypes.byref(d)) == bmfr_amd.ERR_INVALID_ARGUMENT
    assert lib.bmfr_measure_lds_read_bandwidth(16, None) == bmfr_amd.ERR_INVALID_ARGUMENT


def test_stream_open_without_gpu_reports_no_device():
    if bmfr_amd.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(bmfr_amd.BmfrError) as e:
        bmfr_amd.Stream(bmfr_amd.default_params(64, 64))
    assert e.value.status == bmfr_amd.ERR_NO_DEVICE


def test_device_footprint_720p():
    # 14 float3 buffers, spp x2, prev_pixel, accept and tmp_data (half):
    # the per-stream admission figure.
    p = bmfr_amd.default_params(1280, 720)
    s = bmfr_amd.compute_sizes(p)
    P, Q = s.pixels, s.padded_pixels
    f32 = s.device_bytes
    assert 0.95 * (14 * 12 * P + 11 * P + 26 * Q) < f32 < 1.05 * (14 * 12 * P + 11 * P + 26 * Q)
    p16 = bmfr_amd.default_params(1280, 720, storage=bmfr_amd.STORAGE_F16)
    assert bmfr_amd.compute_sizes(p16).device_bytes < 0.7 * f32


def test_header_compiles_as_c99_and_cxx(tmp_path):
    """include/bmfr_hip.h is plain C: C99 with every warning, and C++."""
    inc = os.path.join(ROOT, "include")
    src = tmp_path / "h.c"
    src.write_text('#include "bmfr_hip.h"\nint main(void) { return bmfr_api_version() == 3 ? 0 : 1; }\n')
    for cmd in (["gcc", "-x", "c", "-std=c99", "-pedantic", "-Wall", "-Wextra", "-Werror"],
                ["g++", "-x", "c++", "-std=c++11", "-pedantic", "-Wall", "-Wextra", "-Werror"]):
        r = subprocess.run(cmd + ["-fsyntax-only", "-I", inc, str(src)], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


def test_struct_layouts_match_the_python_mirror(tmp_path):
    """The ctypes structures have the C compiler's sizes and offsets."""
    inc = os.path.join(ROOT, "include")
    structs = {"bmfr_params": _lib.BmfrParams, "bmfr_sizes": _lib.BmfrSizes,
               "bmfr_frame_request": _lib.BmfrFrameRequest, "bmfr_server_event": _lib.BmfrServerEvent}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "bmfr_hip.h"', 'int main(void) {']
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    r = subprocess.run(["gcc", "-std=c99", "-I", inc, str(src), "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    got = {}
    for line in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines():
        c, f, v = line.split()
        got[(c, f)] = int(v)
    for cname, py in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)


def test_new_entry_points_validate_without_a_gpu():
    lib = _lib.load()
    assert lib.bmfr_bind_device_inputs(None, None, None, None, None) == bmfr_amd.ERR_INVALID_ARGUMENT
    assert lib.bmfr_bind_output_surface(None, None, 0, 0) == bmfr_amd.ERR_INVALID_ARGUMENT
    assert lib.bmfr_server_submit(None, None) == bmfr_amd.ERR_INVALID_ARGUMENT
    n = ctypes.c_int32()
    assert lib.bmfr_server_poll(None, None, 0, 0, ctypes.byref(n)) == bmfr_amd.ERR_INVALID_ARGUMENT
    assert lib.bmfr_server_step_query(None, 1, None, None) == bmfr_amd.ERR_INVALID_ARGUMENT
    t0 = lib.bmfr_clock_seconds()
    assert lib.bmfr_clock_seconds() >= t0 > 0
