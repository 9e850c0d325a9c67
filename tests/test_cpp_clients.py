"""The C ABI driven from C++ rather than ctypes:
  - bmfr_frame_loop, the reference's frame loop (opencl/bmfr.cpp:417-485)
    written against include/bmfr_hip.h, gives the bits of the same frames
    driven through the Python mirror (separate stage calls and every fuse
    level of bmfr_run_frame);
  - bmfr_serve_load, the request-loop load generator the benchmark runs,
    completes a short served run with every admitted frame done.
Both programs link libbmfr_hip.so from the package directory."""
import json
import os
import subprocess

import numpy as np
import pytest

import bmfr_amd
from tests.harness import gpu_params, scene

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FRAME_LOOP = os.path.join(ROOT, "bmfr_amd", "bmfr_frame_loop")
SERVE_LOAD = os.path.join(ROOT, "bmfr_amd", "bmfr_serve_load")


def test_clients_are_built_against_the_package_library():
    for exe in (FRAME_LOOP, SERVE_LOAD):
        assert os.access(exe, os.X_OK), exe
        deps = subprocess.run(["ldd", exe], capture_output=True, text=True).stdout
        line = [l for l in deps.splitlines() if "libbmfr_hip.so" in l]
        assert line and os.path.join(ROOT, "bmfr_amd", "libbmfr_hip.so") in line[0], deps


@pytest.mark.gpu
@pytest.mark.parametrize("fused", [-1, 0, 1, 2])
def test_cpp_frame_loop_equals_python_mirror(gpu, tmp_path, fused):
    w, h, frames, seed = 130, 70, 5, 0xB3F40000 + 70
    out = tmp_path / "result.f32"
    args = [FRAME_LOOP, str(w), str(h), str(frames), str(out), str(seed)] + ([str(fused)] if fused >= 0 else [])
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "total time in all kernels" in r.stdout
    got = np.fromfile(out, np.float32)
    sc = scene(w, h, seed=seed)
    st = bmfr_amd.Stream(gpu_params(w, h))
    try:
        if fused >= 0:
            st.set_fused(fused)
        for fr in range(frames):
            f = sc.render_host(fr)
            st.upload_frame(f["albedo"], f["normals"], f["positions"], f["noisy"])
            vp, off = sc.frame_args(fr)
            if fused >= 0:
                st.run_frame(fr, vp, off)
            else:
                st.accumulate_noisy_data(fr, vp, off)
                st.fitter(fr)
                st.weighted_sum(fr)
                st.accumulate_filtered_data(fr)
                st.taa(fr)
                st.end_frame()
        want = st.download_result()
    finally:
        st.close()
    assert np.array_equal(np.nan_to_num(got, nan=7), np.nan_to_num(want, nan=7))


@pytest.mark.gpu
def test_cpp_serve_load_short_run(gpu):
    r = subprocess.run([SERVE_LOAD, "--mode", "served", "--streams", "8", "--seconds", "2", "--budget-gib", "1.5",
                        "--frames-rendered", "4"], capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stderr
    d = json.loads(r.stdout.strip().splitlines()[-1])
    assert d["failed"] == 0 and d["frames"] > 0
    assert d["admitted"] + d["never_admitted"] == 8
    assert d["frames"] + d["dropped"] == d["frames_arrived"]
    assert 0 < d["latency_ms_p50"] <= d["latency_ms_p99"] <= d["latency_ms_max"]
