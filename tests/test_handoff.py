"""Renderer handoff (opencl/bmfr.cpp:415-416, 478-480: "in real use case there
would not be WriteBuffer and ReadBuffer ... the input data comes from the
path tracer and output goes to the screen").

- Zero-copy inputs: a stream whose frames read caller-owned device G-buffers
  (two renderer sets used in turn, as a renderer with double-buffered
  G-buffers would) gives the bits of the same stream fed through its own
  buffers, at every fuse level, through the separate stage calls, and in
  batch and server steps; and its final frame meets the end-to-end bars
  against the CPU oracle fed the same frames.
- Output surfaces: every format holds the frame's TAA result as
  download_result returns it (float formats exactly, half rounded to
  nearest even, 8-bit as round(clamp(v, 0, 1) * 255)), alpha set, row
  padding untouched, written by every fused and batched TAA kernel.
"""
import numpy as np
import pytest

import bmfr_amd
from oracle import bind
from tests.harness import DeviceBuffer, FrameInputs, HipStream, gpu_params, quantize_u8, rel_err, scene

pytestmark = pytest.mark.gpu

SIZES = [(96, 64, False), (130, 70, True)]   # (w, h, storage_f16); 130x70 is ragged


def _run_own(st, sc, frames, fuse):
    st.set_fused(fuse)
    for fr in range(frames):
        sc.render_to_stream(fr, st)
        vp, off = sc.frame_args(fr)
        st.run_frame(fr, vp, off)
    st.synchronize()


def _run_bound(st, sc, frames, fuse, sets, producer=None, separate=False):
    st.set_fused(fuse)
    for fr in range(frames):
        inp = sets[fr % 2]
        inp.render(sc, fr, producer.handle if producer else None)
        if producer:
            producer.synchronize()
        st.bind_device_inputs(*inp.ptrs)
        vp, off = sc.frame_args(fr)
        if separate:
            st.accumulate_noisy_data(fr, vp, off)
            st.fitter(fr)
            st.weighted_sum(fr)
            st.accumulate_filtered_data(fr)
            st.taa(fr)
            st.end_frame()
        else:
            st.run_frame(fr, vp, off)
    st.synchronize()


def _history(st):
    names = ["result_prev", "out_prev", "noisy_acc_prev", "spp_prev", "weights", "mins_maxs", "prev_pixel",
             "accept"]
    d = {n: st.read_buffer(n) for n in names}
    d["download"] = st.download_result()
    return d


def _same(a, b):
    return np.array_equal(np.nan_to_num(a, nan=7.0), np.nan_to_num(b, nan=7.0)) and np.array_equal(
        np.isnan(a), np.isnan(b))


@pytest.mark.parametrize("w,h,g16", SIZES)
@pytest.mark.parametrize("mode", ["fuse0", "fuse1", "fuse2", "separate"])
def test_bound_inputs_equal_own_buffers(gpu, w, h, g16, mode):
    fuse = {"fuse0": 0, "fuse1": 1, "fuse2": 2, "separate": 0}[mode]
    sc = scene(w, h, seed=0xB3F40000 + 20)
    own = bmfr_amd.Stream(gpu_params(w, h, storage_f16=g16))
    bound = bmfr_amd.Stream(gpu_params(w, h, storage_f16=g16))
    sets = [FrameInputs(w, h, g16), FrameInputs(w, h, g16)]
    prod = HipStream()
    try:
        _run_own(own, sc, 6, fuse)
        _run_bound(bound, sc, 6, fuse, sets, producer=prod, separate=mode == "separate")
        a, b = _history(own), _history(bound)
        for k in a:
            assert _same(a[k], b[k]), k
    finally:
        for s in sets:
            s.free()
        prod.destroy()
        own.close()
        bound.close()


def test_bound_inputs_against_the_oracle(gpu):
    """End to end with caller-owned G-buffers: the final frame against the
    CPU oracle fed the same frames, with the plain path's FP32-tmp bar."""
    w, h = 96, 64
    sc = scene(w, h, seed=0xB3F40000 + 21)
    g = bmfr_amd.Stream(gpu_params(w, h, tmp_f16=False))
    o = bind.OracleStream(bind.params(w, h, half_tmp=False))
    sets = [FrameInputs(w, h, False), FrameInputs(w, h, False)]
    try:
        for fr in range(8):
            f = sc.render_host(fr)
            o.upload(f["albedo"], f["normals"], f["positions"], f["noisy"])
            sets[fr % 2].render(sc, fr)
            g.bind_device_inputs(*sets[fr % 2].ptrs)
            vp, off = sc.frame_args(fr)
            g.run_frame(fr, vp, off)
            want = o.run_frame(fr, vp, off)
        got = g.download_result()
        fin = np.isfinite(want) & np.isfinite(got)
        assert np.array_equal(np.isfinite(want), np.isfinite(got))
        err = rel_err(got[fin], want[fin])
        assert err.max() < 1e-3, float(err.max())
    finally:
        for s in sets:
            s.free()
        g.close()


def test_binding_rules(gpu):
    w, h = 96, 64
    st = bmfr_amd.Stream(gpu_params(w, h))
    buf = DeviceBuffer(16 * w * h * 3 + 16)
    try:
        with pytest.raises(bmfr_amd.BmfrError):
            st.bind_device_inputs(buf.ptr, None, buf.ptr, buf.ptr)       # all four or none
        with pytest.raises(bmfr_amd.BmfrError):
            st.bind_device_inputs(buf.ptr + 4, buf.ptr, buf.ptr, buf.ptr)   # 16-byte alignment
        st.bind_device_inputs(None, None, None, None)
        with pytest.raises(bmfr_amd.BmfrError):
            st.bind_output_surface(buf.ptr, bmfr_amd.SURFACE_RGBA_F32, w * 16 - 4)   # pitch too small
        with pytest.raises(bmfr_amd.BmfrError):
            st.bind_output_surface(buf.ptr, bmfr_amd.SURFACE_RGBA_F32, w * 16 + 8)   # pitch misaligned
        with pytest.raises(bmfr_amd.BmfrError):
            st.bind_output_surface(buf.ptr + 4, bmfr_amd.SURFACE_RGBA_F16, w * 8)    # base misaligned
        with pytest.raises(bmfr_amd.BmfrError):
            st.bind_output_surface(buf.ptr, 9, w * 16)
        st.bind_output_surface(buf.ptr, bmfr_amd.SURFACE_RGBA_U8, w * 4)
        st.bind_output_surface(None)
        assert st.hip_stream
    finally:
        buf.free()
        st.close()


def _expected_surface(res, fmt):
    """The surface bytes of a W*H*3 float result in format fmt."""
    rgb = res.reshape(-1, 3)
    if fmt == bmfr_amd.SURFACE_RGB_F32:
        return rgb.astype(np.float32)
    if fmt == bmfr_amd.SURFACE_RGBA_F32:
        return np.concatenate([rgb, np.ones((len(rgb), 1), np.float32)], 1)
    if fmt == bmfr_amd.SURFACE_RGBA_F16:
        return np.concatenate([rgb.astype(np.float16), np.ones((len(rgb), 1), np.float16)], 1)
    return np.concatenate([quantize_u8(rgb), np.full((len(rgb), 1), 255, np.uint8)], 1)


def _check_surface(surf, res, fmt, w, h, pitch, sentinel):
    pix, dt, ch = bmfr_amd.SURFACE_LAYOUT[fmt]
    raw = surf.read(np.uint8).reshape(h, pitch)
    assert (raw[:, w * pix:] == sentinel).all(), "row padding written"
    got = raw[:, :w * pix].copy().view(dt).reshape(-1, ch)
    want = _expected_surface(res, fmt)
    assert got.dtype == want.dtype
    assert _same(got.astype(np.float64), want.astype(np.float64)), fmt


@pytest.mark.parametrize("fmt", [bmfr_amd.SURFACE_RGB_F32, bmfr_amd.SURFACE_RGBA_F32, bmfr_amd.SURFACE_RGBA_F16,
                                 bmfr_amd.SURFACE_RGBA_U8])
@pytest.mark.parametrize("fuse,g16", [(0, False), (1, True), (2, False), (2, True)])
def test_output_surface_holds_the_result(gpu, fmt, fuse, g16):
    w, h = 130, 70
    pix = bmfr_amd.SURFACE_LAYOUT[fmt][0]
    pitch = w * pix + 64          # padded rows
    sc = scene(w, h, seed=0xB3F40000 + 22)
    st = bmfr_amd.Stream(gpu_params(w, h, storage_f16=g16))
    surf = DeviceBuffer(pitch * h, fill=0xA5)
    try:
        st.set_fused(fuse)
        st.bind_output_surface(surf.ptr, fmt, pitch)
        for fr in range(4):
            sc.render_to_stream(fr, st)
            vp, off = sc.frame_args(fr)
            st.run_frame(fr, vp, off)
            _check_surface(surf, st.download_result(), fmt, w, h, pitch, 0xA5)
    finally:
        surf.free()
        st.close()


@pytest.mark.parametrize("fuse", [bmfr_amd.FUSE_34, bmfr_amd.FUSE_345])
def test_batch_with_bound_inputs_and_surfaces_equals_alone(gpu, fuse):
    mix = [(96, 64, False), (130, 70, True), (200, 120, False)]
    scs = [scene(w, h, seed=0xB3F40000 + 30 + k) for k, (w, h, _) in enumerate(mix)]
    alone = [bmfr_amd.Stream(gpu_params(w, h, storage_f16=g)) for (w, h, g) in mix]
    members = [bmfr_amd.Stream(gpu_params(w, h, storage_f16=g)) for (w, h, g) in mix]
    sets = [[FrameInputs(w, h, g), FrameInputs(w, h, g)] for (w, h, g) in mix]
    surfs = [DeviceBuffer(w * h * 4) for (w, h, _) in mix]
    b = bmfr_amd.Batch(members)
    try:
        b.set_fused(fuse)
        b.set_graph(True)
        for m, s, (w, h, _) in zip(members, surfs, mix):
            m.bind_output_surface(s.ptr, bmfr_amd.SURFACE_RGBA_U8)
        for fr in range(5):
            args = [sc.frame_args(fr) for sc in scs]
            for k in range(len(mix)):
                sets[k][fr % 2].render(scs[k], fr)
                members[k].bind_device_inputs(*sets[k][fr % 2].ptrs)
            b.run([fr] * len(mix), np.concatenate([a[0] for a in args]), np.concatenate([a[1] for a in args]))
            b.synchronize()
            for k, (w, h, _) in enumerate(mix):
                _check_surface(surfs[k], members[k].download_result(), bmfr_amd.SURFACE_RGBA_U8, w, h, w * 4, None)
        for k in range(len(mix)):
            _run_own(alone[k], scs[k], 5, fuse)
            a, m = _history(alone[k]), _history(members[k])
            for name in a:
                assert _same(a[name], m[name]), (k, name)
    finally:
        b.close()
        for st in alone + members:
            st.close()
        for ss in sets:
            for s in ss:
                s.free()
        for s in surfs:
            s.free()
