"""bmfr_fit_blocks, the fitter entry point that allocates nothing (its
scratch is a fixed device array per GPU): through the C ABI its weights and
mins_maxs are bit-identical to bmfr_fitter's on the same block matrices, on
two calls in a row with the second smaller, in every QR mode; out-of-range
sizes are refused; its source calls no allocating routine."""
import ctypes
import os
import re

import numpy as np
import pytest

import bmfr_amd
from bmfr_amd import _lib
from tests.harness import DeviceBuffer, gpu_params, scene

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_source_allocates_nothing():
    src = open(os.path.join(ROOT, "bmfr_amd", "csrc", "noalloc.hip")).read()
    code = re.sub(r"//[^\n]*", "", src)
    for name in ("hipMalloc", "hipFree", "hipHostMalloc", "hipMallocAsync", "hipMemset", "malloc(", "new "):
        assert name not in code, name


def _stream_fit(w, h, frame, tmp_f16, qr, thr):
    """A stream's own stage 1 and fitter: (tmp_data device pointer, weights,
    mins_maxs, stream) -- the stream is returned open."""
    st = bmfr_amd.Stream(gpu_params(w, h, tmp_f16=tmp_f16, qr_mode=qr, refactor_threshold=thr))
    sc = scene(w, h, seed=0xB3F40000 + 80 + w)
    for fr in range(frame + 1):
        sc.render_to_stream(fr, st)
        st.accumulate_noisy_data(fr, *sc.frame_args(fr))
        if fr < frame:
            st.end_frame()
    st.fitter(frame)
    return st, st.read_buffer("weights"), st.read_buffer("mins_maxs")


def _fit_blocks(st, frame, tmp_f16, qr, thr):
    lib = _lib.load()
    n = st.sizes.blocks
    w_out, mm_out = DeviceBuffer(30 * 4 * n, fill=0), DeviceBuffer(12 * 4 * n, fill=0)
    try:
        rc = lib.bmfr_fit_blocks(st.buffer_device_ptr("tmp_data"), 1 if tmp_f16 else 0, n, frame, 1e-2, qr,
                                 ctypes.c_float(thr), w_out.ptr, mm_out.ptr, None)
        assert rc == 0, lib.bmfr_last_error()
        return w_out.read(np.float32), mm_out.read(np.float32)
    finally:
        w_out.free()
        mm_out.free()


@pytest.mark.gpu
@pytest.mark.parametrize("qr,thr", [(bmfr_amd.QR_FP32, 1e-6), (bmfr_amd.QR_KAHAN, 1e-6),
                                    (bmfr_amd.QR_KAHAN_FP64, 0.5)])
def test_fit_blocks_equals_fitter_two_calls(gpu, qr, thr):
    for (w, h, tmp_f16, frame) in [(256, 256, True, 2), (96, 64, False, 3)]:   # second call smaller
        st, w_want, mm_want = _stream_fit(w, h, frame, tmp_f16, qr, thr)
        try:
            w_got, mm_got = _fit_blocks(st, frame, tmp_f16, qr, thr)
        finally:
            st.close()
        assert np.array_equal(mm_got, mm_want), (w, h)
        assert np.array_equal(w_got, w_want), (w, h)


@pytest.mark.gpu
def test_fit_blocks_refuses_sizes(gpu):
    lib = _lib.load()
    buf = DeviceBuffer(1 << 20)
    try:
        for blocks in (0, 16385):
            assert lib.bmfr_fit_blocks(buf.ptr, 1, blocks, 0, 1e-2, 0, ctypes.c_float(1e-6), buf.ptr, buf.ptr,
                                       None) == bmfr_amd.ERR_INVALID_ARGUMENT
            assert b"blocks" in lib.bmfr_last_error()
        assert lib.bmfr_fit_blocks(None, 1, 4, 0, 1e-2, 0, ctypes.c_float(1e-6), buf.ptr, buf.ptr,
                                   None) == bmfr_amd.ERR_INVALID_ARGUMENT
        assert lib.bmfr_fit_blocks(buf.ptr, 1, 4, 0, 1e-2, 7, ctypes.c_float(1e-6), buf.ptr, buf.ptr,
                                   None) == bmfr_amd.ERR_INVALID_ARGUMENT
    finally:
        buf.free()
