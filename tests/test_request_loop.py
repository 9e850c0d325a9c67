"""The server's request loop (bmfr_server_submit / bmfr_server_poll): frames of
many render streams are queued as they arrive and a dispatcher thread in the
library forms batched steps continuously, at most two on the GPU at a time,
one frame per stream per step in submission order.  Checks:
  - every result is the same, bit for bit, as running the stream alone,
    whatever the steps it shared (mixed sizes, storage precisions and QR
    modes, streams opening and closing mid-run, the admission FIFO);
  - the completion events: one per frame, in order per stream, timestamps
    ordered arrival <= launch <= done, a step's size and GPU time;
  - step ids are a monotonically increasing count, and a reused record is
    reported as expired rather than another step's;
  - the rules: a stream's own input buffers hold one queued frame, closing
    waits for queued frames, destroying the server invalidates its streams.
"""
import time

import numpy as np
import pytest

import bmfr_amd
from tests.harness import DeviceBuffer, FrameInputs, HipStream, gpu_params, quantize_u8, scene

pytestmark = pytest.mark.gpu


def _alone(w, h, g16, sc, frames, qr=bmfr_amd.QR_FP32, thr=1e-6):
    st = bmfr_amd.Stream(gpu_params(w, h, storage_f16=g16, qr_mode=qr, refactor_threshold=thr))
    st.set_fused(bmfr_amd.FUSE_345)
    for fr in range(frames):
        sc.render_to_stream(fr, st)
        st.run_frame(fr, *sc.frame_args(fr))
    st.synchronize()
    out = st.download_result(), st.read_buffer("weights"), st.refactor_count()
    st.close()
    return out


class Client:
    """A renderer feeding one server stream: frames rendered into a ring of
    four caller-owned G-buffer sets on a producer HIP stream, each with an
    RGBA8 output surface of its own."""

    def __init__(self, sv, stream, w, h, g16, sc, prod):
        self.sv, self.st, self.w, self.h, self.sc, self.prod = sv, stream, w, h, sc, prod
        self.ring = [FrameInputs(w, h, g16) for _ in range(4)]
        self.surf = [DeviceBuffer(w * h * 4) for _ in range(4)]
        self.next = 0          # next frame to submit
        self.done = 0          # frames completed (in order)

    def can_submit(self):
        # set k % 4 is free once frame k - 4 and its successor (which reads
        # its normals and positions) have completed
        return self.next - self.done < 3

    def submit(self, tag):
        fr = self.next
        inp = self.ring[fr % 4]
        inp.render(self.sc, fr, self.prod.handle)
        vp, off = self.sc.frame_args(fr)
        self.sv.submit(self.st, fr, vp, off, inputs=inp.ptrs, surface=self.surf[fr % 4].ptr,
                       surface_format=bmfr_amd.SURFACE_RGBA_U8, producer_stream=self.prod.handle, tag=tag)
        self.next += 1

    def free(self):
        for x in self.ring:
            x.free()
        for s in self.surf:
            s.free()


MIX = [  # (w, h, storage_f16, qr_mode)
    (96, 64, False, bmfr_amd.QR_FP32),
    (130, 70, True, bmfr_amd.QR_FP32),
    (200, 120, False, bmfr_amd.QR_KAHAN),
    (64, 48, True, bmfr_amd.QR_KAHAN_FP64),
]


def test_request_loop_results_equal_alone(gpu):
    frames = 6
    sv = bmfr_amd.Server(1 << 32)
    prod = HipStream()
    scs = [scene(w, h, seed=0xB3F40000 + 40 + k) for k, (w, h, _, _) in enumerate(MIX)]
    clients = []
    try:
        for k, (w, h, g16, qr) in enumerate(MIX):
            st = sv.open_stream(gpu_params(w, h, storage_f16=g16, qr_mode=qr, refactor_threshold=0.5))
            clients.append(Client(sv, st, w, h, g16, scs[k], prod))
        events, tag = [], 0
        deadline = time.time() + 60
        while any(c.done < frames for c in clients):
            assert time.time() < deadline, "request loop stalled"
            for k, c in enumerate(clients):
                # stream k submits at its own pace: step membership changes
                if c.next < frames and c.can_submit() and (k % 2 == 0 or tag % 3 != 0):
                    c.submit(tag=1000 * k + c.next)
                tag += 1
            for e in sv.poll(timeout_us=2000):
                assert e["kind"] == bmfr_amd.EVENT_FRAME_DONE, e
                k = e["tag"] // 1000
                assert e["frame_number"] == clients[k].done == e["tag"] % 1000
                assert e["t_arrival"] <= e["t_launch"] <= e["t_done"]
                assert 1 <= e["step_streams"] <= len(MIX) and e["step_gpu_ms"] > 0
                clients[k].done += 1
                events.append(e)
        ids = [e["step_id"] for e in events]
        assert ids == sorted(ids) and ids[0] >= 1
        assert len(set(ids)) < len(ids), "no step held two streams"
        for k, (w, h, g16, qr) in enumerate(MIX):
            c = clients[k]
            res, wts, nref = _alone(w, h, g16, scs[k], frames, qr=qr, thr=0.5)
            got = c.st.download_result()
            assert np.array_equal(np.nan_to_num(res, nan=7), np.nan_to_num(got, nan=7)), k
            assert np.array_equal(wts, c.st.read_buffer("weights")), k
            assert nref == c.st.refactor_count(), k
            if qr == bmfr_amd.QR_KAHAN_FP64:
                assert nref > 0
            # the last frame's surface: the 8-bit form of the result
            surf = c.surf[(frames - 1) % 4].read(np.uint8).reshape(-1, 4)
            assert np.array_equal(surf[:, :3], quantize_u8(got.reshape(-1, 3))), k
            assert (surf[:, 3] == 255).all()
    finally:
        sv.close()
        prod.destroy()
        for c in clients:
            c.free()


def test_own_input_buffers_hold_one_queued_frame(gpu):
    """With the loop paused (no step in flight allowed) a frame stays queued:
    a second frame that would also need the stream's own input buffers is
    refused, one with device inputs is queued; resuming runs both in order."""
    sv = bmfr_amd.Server(1 << 30)
    w, h = 96, 64
    sc = scene(w, h)
    inp = FrameInputs(w, h, False)
    try:
        st = sv.open_stream(gpu_params(w, h))
        sv.set_policy(0, 64)                      # paused
        sc.render_to_stream(0, st)
        sv.submit(st, 0, *sc.frame_args(0), tag=0)
        with pytest.raises(bmfr_amd.BmfrError) as e:
            sv.submit(st, 1, *sc.frame_args(1), tag=1)
        assert e.value.status == bmfr_amd.ERR_STATE
        inp.render(sc, 1)
        sv.submit(st, 1, *sc.frame_args(1), inputs=inp.ptrs, tag=1)
        assert sv.poll(timeout_us=20000) == []     # nothing launched while paused
        sv.set_policy(2, 64)
        got = []
        deadline = time.time() + 30
        while len(got) < 2:
            assert time.time() < deadline
            got += sv.poll(timeout_us=100000)
        assert [e["frame_number"] for e in got] == [0, 1]
        assert got[0]["step_id"] < got[1]["step_id"]
        other = bmfr_amd.Stream(gpu_params(w, h))
        with pytest.raises(bmfr_amd.BmfrError):
            sv.submit(other, 0, *sc.frame_args(0))   # not a server stream
        other.close()
        # the same two frames alone
        res, _, _ = _alone(w, h, False, sc, 2)
        assert np.array_equal(res, st.download_result())
    finally:
        sv.close()
        inp.free()


def test_step_ids_monotonic_and_expiry(gpu):
    sv = bmfr_amd.Server(1 << 30)
    try:
        w, h = 64, 48
        st = sv.open_stream(gpu_params(w, h))
        sc = scene(w, h)
        ids = []
        for fr in range(10):
            sc.render_to_stream(fr, st)
            vp, off = sc.frame_args(fr)
            ids.append(sv.step([st], [fr], vp, off))
        assert ids == list(range(ids[0], ids[0] + 10))
        sv.synchronize()
        done, ms = sv.query(ids[-1])
        assert done and ms > 0
        with pytest.raises(bmfr_amd.BmfrError) as e:
            sv.query(ids[0])                     # its record has been reused
        assert e.value.status == bmfr_amd.ERR_STATE
        with pytest.raises(bmfr_amd.BmfrError) as e:
            sv.query(ids[-1] + 1)                # never launched
        assert e.value.status == bmfr_amd.ERR_INVALID_ARGUMENT
    finally:
        sv.close()


def test_admission_fifo_and_close_invalidation(gpu):
    p = gpu_params(96, 64)
    need = bmfr_amd.compute_sizes(p).device_bytes
    sv = bmfr_amd.Server(2 * need + need // 2)
    a = sv.request_stream(p, tag=1)
    b = sv.request_stream(p, tag=2)
    assert a is not None and b is not None
    assert sv.request_stream(p, tag=3) is None       # queued: no room
    assert sv.request_stream(p, tag=4) is None       # queued behind it
    assert sv.cancel_request(4) and not sv.cancel_request(4)
    with pytest.raises(bmfr_amd.BmfrError) as e:
        sv.request_stream(bmfr_amd.default_params(1920, 1080), tag=5)   # larger than the budget
    assert e.value.status == bmfr_amd.ERR_OUT_OF_BUDGET
    sv.close_stream(a)
    ev = sv.poll(timeout_us=1000000)
    assert len(ev) == 1 and ev[0]["kind"] == bmfr_amd.EVENT_STREAM_ADMITTED and ev[0]["tag"] == 3
    c = ev[0]["stream"]
    assert c is not None and sv.usage()[2] == 2
    # a server's stream is closed through the server only
    assert bmfr_amd._lib.load().bmfr_stream_close(c.handle) == bmfr_amd.ERR_STATE
    sv.close()
    assert b.handle is None and c.handle is None     # the server closed them


def test_batch_member_cannot_be_closed(gpu):
    sts = [bmfr_amd.Stream(gpu_params(96, 64)) for _ in range(2)]
    b = bmfr_amd.Batch(sts)
    with pytest.raises(bmfr_amd.BmfrError) as e:
        sts[0].close()
    assert e.value.status == bmfr_amd.ERR_STATE
    with pytest.raises(bmfr_amd.BmfrError):
        bmfr_amd.Batch([sts[1]])                      # already in a batch
    b.close()
    for s in sts:
        s.close()
