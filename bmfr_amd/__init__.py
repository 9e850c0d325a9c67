"""Host-side mirror of the reference's frame API over the gfx950 C ABI.

The reference drives five OpenCL kernels per frame from C++
(opencl/bmfr.cpp:417-485): accumulate_noisy_data, fitter, weighted_sum,
accumulate_filtered_data, taa, then swaps its double buffers.  ``Stream``
exposes the same five calls (same names, order and argument meaning) on a
device-resident stream from ``libbmfr_hip.so``; every call goes to a
hand-written HIP kernel.  Failures raise ``BmfrError`` carrying the C
status code.
"""
import ctypes
import weakref

import numpy as np

from . import _lib
from ._lib import BmfrFrameRequest, BmfrParams, BmfrServerEvent, BmfrSizes

# enum bmfr_status
OK = 0
ERR_INVALID_ARGUMENT = 1
ERR_OUT_OF_BUDGET = 2
ERR_HIP = 3
ERR_NO_DEVICE = 4
ERR_STATE = 5

STORAGE_F32 = 0
STORAGE_F16 = 1
QR_FP32 = 0
QR_KAHAN = 1
QR_KAHAN_FP64 = 2
FUSE_NONE, FUSE_34, FUSE_345 = 0, 1, 2   # bmfr_fuse_level
# bmfr_surface_format: (id, bytes per pixel, numpy dtype, channels)
SURFACE_RGB_F32, SURFACE_RGBA_F32, SURFACE_RGBA_F16, SURFACE_RGBA_U8 = 0, 1, 2, 3
SURFACE_LAYOUT = {SURFACE_RGB_F32: (12, np.float32, 3), SURFACE_RGBA_F32: (16, np.float32, 4),
                  SURFACE_RGBA_F16: (8, np.float16, 4), SURFACE_RGBA_U8: (4, np.uint8, 4)}
# bmfr_event_kind
EVENT_FRAME_DONE, EVENT_FRAME_FAILED, EVENT_STREAM_ADMITTED = 1, 2, 3

# enum bmfr_buffer: id -> (name, element dtype key, elements per pixel kind)
BUFFERS = {
    "albedo": 0, "normals_cur": 1, "normals_prev": 2, "positions_cur": 3,
    "positions_prev": 4, "noisy_in": 5, "noisy_acc_cur": 6, "noisy_acc_prev": 7,
    "spp_cur": 8, "spp_prev": 9, "prev_pixel": 10, "accept": 11, "tmp_data": 12,
    "weights": 13, "mins_maxs": 14, "filtered": 15, "tone_mapped": 16,
    "out_cur": 17, "out_prev": 18, "result_cur": 19, "result_prev": 20,
}
_STORAGE_BUFFERS = {"normals_cur", "normals_prev", "positions_cur", "positions_prev",
                    "noisy_in", "noisy_acc_cur", "noisy_acc_prev", "out_cur", "out_prev",
                    "result_cur", "result_prev"}
_U8_BUFFERS = {"spp_cur", "spp_prev", "accept"}


class BmfrError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"bmfr status {status}: {message}")
        self.status = status


def _check(status):
    if status != OK:
        lib = _lib.load()
        raise BmfrError(status, lib.bmfr_last_error().decode(errors="replace"))


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _f32(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} floats, got {a.size}")
    return a


def device_count():
    return _lib.load().bmfr_device_count()


def clock_seconds():
    """Seconds on the request loop's clock (bmfr_clock_seconds)."""
    return _lib.load().bmfr_clock_seconds()


def lds_read_bandwidth(steps=4096):
    """Measured LDS read bandwidth of the current GPU, bytes/s."""
    v = ctypes.c_double()
    _check(_lib.load().bmfr_measure_lds_read_bandwidth(steps, ctypes.byref(v)))
    return v.value


def default_params(width, height, **overrides):
    """Reference defaults (opencl/bmfr.cpp:58-62, 88) with optional overrides."""
    lib = _lib.load()
    p = BmfrParams()
    st = lib.bmfr_default_params(width, height, ctypes.byref(p))
    for k, v in overrides.items():
        setattr(p, k, v)
    if overrides or st:
        s = BmfrSizes()
        st = lib.bmfr_compute_sizes(ctypes.byref(p), ctypes.byref(s))
    _check(st)
    return p


def parse_feature_buffers(params, not_scaled, scaled):
    """Sets params' feature set from the reference's two feature strings,
    NOT_SCALED_FEATURE_BUFFERS and SCALED_FEATURE_BUFFERS (opencl/bmfr.cpp:65-77),
    e.g. ("1.f,normal.x,normal.y,normal.z,", "world_position.x*world_position.y").
    The reference's own strings select its set (feature_count 0)."""
    lib = _lib.load()
    _check(lib.bmfr_parse_feature_buffers(not_scaled.encode(), scaled.encode(), ctypes.byref(params)))
    return params


def compute_sizes(params):
    lib = _lib.load()
    s = BmfrSizes()
    _check(lib.bmfr_compute_sizes(ctypes.byref(params), ctypes.byref(s)))
    return s


class Stream:
    """One render stream with device-resident buffers and history."""

    def __init__(self, params):
        self._lib = _lib.load()
        self.params = params
        h = ctypes.c_void_p()
        _check(self._lib.bmfr_stream_open(ctypes.byref(params), ctypes.byref(h)))
        self._h = h
        s = BmfrSizes()
        _check(self._lib.bmfr_stream_sizes(h, ctypes.byref(s)))
        self.sizes = s

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            owner = getattr(self, "_owner", None)
            if owner is not None:
                owner.close_stream(self)
            else:
                _check(self._lib.bmfr_stream_close(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- inputs ---------------------------------------------------------------
    def upload_frame(self, albedo, normals, positions, noisy):
        n = self.params.width * self.params.height * 3
        arrs = [_f32(x, n) for x in (albedo, normals, positions, noisy)]
        _check(self._lib.bmfr_upload_frame(self._h, *[_fp(a) for a in arrs]))

    def copy_frame_inputs_device(self, albedo, normals, positions, noisy):
        """Device pointers (ints) of the frame's inputs; see
        bmfr_copy_frame_inputs_device."""
        _check(self._lib.bmfr_copy_frame_inputs_device(
            self._h, *[ctypes.c_void_p(p) for p in (albedo, normals, positions, noisy)]))

    def bind_device_inputs(self, albedo, normals, positions, noisy):
        """Zero-copy inputs of the next frame: caller-owned device buffers
        (ints; all None unbinds); see bmfr_bind_device_inputs."""
        _check(self._lib.bmfr_bind_device_inputs(
            self._h, *[ctypes.c_void_p(p) for p in (albedo, normals, positions, noisy)]))

    def bind_output_surface(self, surface, fmt=SURFACE_RGB_F32, row_pitch=None):
        """Every later frame's TAA result also goes to this device surface
        (an int pointer; None unbinds); row_pitch defaults to a packed row."""
        if row_pitch is None:
            row_pitch = self.params.width * SURFACE_LAYOUT[fmt][0]
        _check(self._lib.bmfr_bind_output_surface(self._h, ctypes.c_void_p(surface), int(fmt), int(row_pitch)))

    @property
    def hip_stream(self):
        """The hipStream_t (int) the stream's work runs on."""
        p = ctypes.c_void_p()
        _check(self._lib.bmfr_stream_hip_stream(self._h, ctypes.byref(p)))
        return p.value

    # -- the five stages (opencl/bmfr.cpp:429-476) --------------------------------
    def accumulate_noisy_data(self, frame, prev_frame_camera_matrix, pixel_offset):
        m = _f32(prev_frame_camera_matrix, 16)
        o = _f32(pixel_offset, 2)
        _check(self._lib.bmfr_accumulate_noisy_data(self._h, frame, _fp(m), _fp(o)))

    def fitter(self, frame):
        _check(self._lib.bmfr_fitter(self._h, frame))

    def fitter_fp64_weights(self):
        """FP64 least-squares weights of every block of the last fitter call."""
        out = np.empty(30 * self.sizes.blocks, np.float32)
        _check(self._lib.bmfr_fitter_fp64_weights(self._h, _fp(out)))
        return out

    def refactor_count(self):
        c = ctypes.c_int32()
        _check(self._lib.bmfr_refactor_count(self._h, ctypes.byref(c)))
        return c.value

    def weighted_sum(self, frame):
        _check(self._lib.bmfr_weighted_sum(self._h, frame))

    def accumulate_filtered_data(self, frame):
        _check(self._lib.bmfr_accumulate_filtered_data(self._h, frame))

    def taa(self, frame):
        _check(self._lib.bmfr_taa(self._h, frame))

    def end_frame(self):
        _check(self._lib.bmfr_end_frame(self._h))

    def run_frame(self, frame, prev_frame_camera_matrix, pixel_offset):
        m = _f32(prev_frame_camera_matrix, 16)
        o = _f32(pixel_offset, 2)
        _check(self._lib.bmfr_run_frame(self._h, frame, _fp(m), _fp(o)))

    def set_fused(self, level=True):
        """How run_frame runs stages 3-5: FUSE_NONE (False), FUSE_34 or
        FUSE_345 (True, a new stream's level)."""
        if isinstance(level, bool):
            level = FUSE_345 if level else FUSE_NONE
        _check(self._lib.bmfr_set_fused(self._h, int(level)))

    def set_fitter_config(self, rows_per_lane, blocks_per_group):
        _check(self._lib.bmfr_set_fitter_config(self._h, rows_per_lane, blocks_per_group))

    def synchronize(self):
        _check(self._lib.bmfr_synchronize(self._h))

    def reset(self):
        _check(self._lib.bmfr_stream_reset(self._h))

    # -- outputs ----------------------------------------------------------------
    def download_result(self):
        out = np.empty(self.params.width * self.params.height * 3, np.float32)
        _check(self._lib.bmfr_download_result(self._h, _fp(out)))
        return out

    def buffer_dtype(self, name):
        if name in _U8_BUFFERS:
            return np.uint8
        if name == "tmp_data":
            return np.float16 if self.params.tmp_storage == STORAGE_F16 else np.float32
        if name in _STORAGE_BUFFERS:
            return np.float16 if self.params.storage == STORAGE_F16 else np.float32
        return np.float32

    def buffer_bytes(self, name):
        n = ctypes.c_size_t()
        _check(self._lib.bmfr_buffer_bytes(self._h, BUFFERS[name], ctypes.byref(n)))
        return n.value

    def read_buffer(self, name):
        dt = np.dtype(self.buffer_dtype(name))
        nbytes = self.buffer_bytes(name)
        out = np.empty(nbytes // dt.itemsize, dt)
        _check(self._lib.bmfr_read_buffer(self._h, BUFFERS[name], out.ctypes.data, nbytes))
        return out

    def write_buffer(self, name, array):
        dt = np.dtype(self.buffer_dtype(name))
        a = np.ascontiguousarray(array, dtype=dt)
        nbytes = self.buffer_bytes(name)
        if a.nbytes != nbytes:
            raise ValueError(f"{name} holds {nbytes} bytes, got {a.nbytes}")
        _check(self._lib.bmfr_write_buffer(self._h, BUFFERS[name], a.ctypes.data, nbytes))

    def buffer_device_ptr(self, name):
        p = ctypes.c_void_p()
        _check(self._lib.bmfr_buffer_device_ptr(self._h, BUFFERS[name], ctypes.byref(p)))
        return p.value

    # -- timing -------------------------------------------------------------------
    def set_timing(self, enable=True):
        """True/1: per-stage events; 2: first-start/last-end events only."""
        level = int(enable) if not isinstance(enable, bool) else (1 if enable else 0)
        _check(self._lib.bmfr_set_timing(self._h, level))

    def stage_times(self):
        """ms of the last frame's five stages plus the first-start to
        last-end total (opencl/bmfr.cpp:497-502)."""
        out = np.empty(6, np.float32)
        _check(self._lib.bmfr_stage_times(self._h, _fp(out)))
        return out


    def record_marker(self, marker):
        _check(self._lib.bmfr_record_marker(self._h, marker))

    def marker_elapsed(self, m0, m1):
        ms = ctypes.c_float()
        _check(self._lib.bmfr_marker_elapsed(self._h, m0, m1, ctypes.byref(ms)))
        return ms.value


class Batch:
    """Several streams stepped together: one launch per stage over all of
    them (see bmfr_batch_* in include/bmfr_hip.h)."""

    def __init__(self, streams):
        self._lib = _lib.load()
        self.streams = list(streams)
        arr = (ctypes.c_void_p * len(self.streams))(*[s.handle for s in self.streams])
        h = ctypes.c_void_p()
        _check(self._lib.bmfr_batch_create(arr, len(self.streams), ctypes.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            _check(self._lib.bmfr_batch_destroy(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, frames, matrices, offsets, timed=False):
        n = len(self.streams)
        fr = np.ascontiguousarray(frames, np.int32)
        m = _f32(matrices, 16 * n)
        o = _f32(offsets, 2 * n)
        if fr.size != n:
            raise ValueError("one frame number per stream")
        _check(self._lib.bmfr_batch_run(self._h, fr.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                        _fp(m), _fp(o), 1 if timed else 0))

    def synchronize(self):
        _check(self._lib.bmfr_batch_synchronize(self._h))

    def set_graph(self, enable=True):
        _check(self._lib.bmfr_batch_set_graph(self._h, 1 if enable else 0))

    def set_fused(self, level=2):
        _check(self._lib.bmfr_batch_set_fused(self._h, int(level)))

    def set_lanes(self, lanes=2):
        _check(self._lib.bmfr_batch_set_lanes(self._h, int(lanes)))

    def set_fitter_config(self, rows_per_lane, blocks_per_group, grid_cap=0):
        _check(self._lib.bmfr_batch_set_fitter_config(self._h, rows_per_lane, blocks_per_group,
                                                      grid_cap))

    def stage_times(self):
        out = np.empty(6, np.float32)
        _check(self._lib.bmfr_batch_stage_times(self._h, _fp(out)))
        return out

    def autotune(self, reps=3):
        buf = ctypes.create_string_buffer(4096)
        _check(self._lib.bmfr_batch_autotune(self._h, reps, buf, len(buf)))
        return buf.value.decode()

    def record_marker(self, marker):
        _check(self._lib.bmfr_batch_record_marker(self._h, marker))

    def marker_elapsed(self, m0, m1):
        ms = ctypes.c_float()
        _check(self._lib.bmfr_batch_marker_elapsed(self._h, m0, m1, ctypes.byref(ms)))
        return ms.value


class Server:
    """Streams admitted by an HBM budget and stepped in dynamic batches,
    explicitly (step) or by the library's request loop (submit / poll); see
    bmfr_server_* in include/bmfr_hip.h.  Closing the server closes every
    stream it issued: their Python handles are invalidated."""

    def __init__(self, budget_bytes):
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        _check(self._lib.bmfr_server_create(int(budget_bytes), ctypes.byref(h)))
        self._h = h
        self._issued = weakref.WeakValueDictionary()   # stream handle -> Stream

    def close(self):
        if self._h:
            _check(self._lib.bmfr_server_destroy(self._h))
            self._h = None
            for st in list(self._issued.values()):
                st._h = None
            self._issued.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def usage(self):
        u, b, n = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int32()
        _check(self._lib.bmfr_server_usage(self._h, ctypes.byref(u), ctypes.byref(b), ctypes.byref(n)))
        return u.value, b.value, n.value

    def _wrap(self, h, params):
        st = Stream.__new__(Stream)
        st._lib, st.params, st._h = self._lib, params, ctypes.c_void_p(h)
        sz = BmfrSizes()
        _check(self._lib.bmfr_stream_sizes(st._h, ctypes.byref(sz)))
        st.sizes = sz
        st._owner = self
        self._issued[h] = st
        return st

    def open_stream(self, params):
        """A server-owned Stream; raises BmfrError(ERR_OUT_OF_BUDGET) when it
        does not fit the budget."""
        h = ctypes.c_void_p()
        _check(self._lib.bmfr_server_open_stream(self._h, ctypes.byref(params), ctypes.byref(h)))
        return self._wrap(h.value, params)

    def request_stream(self, params, tag):
        """Admission FIFO: a Stream when admitted now, else None (the stream
        arrives later as an EVENT_STREAM_ADMITTED event with `tag`, which
        poll() turns into a Stream)."""
        h = ctypes.c_void_p()
        _check(self._lib.bmfr_server_request_stream(self._h, ctypes.byref(params), int(tag), ctypes.byref(h)))
        self._pending_params = getattr(self, "_pending_params", {})
        if h.value:
            return self._wrap(h.value, params)
        self._pending_params[int(tag)] = params
        return None

    def cancel_request(self, tag):
        f = ctypes.c_int32()
        _check(self._lib.bmfr_server_cancel_request(self._h, int(tag), ctypes.byref(f)))
        getattr(self, "_pending_params", {}).pop(int(tag), None)
        return bool(f.value)

    def close_stream(self, st):
        _check(self._lib.bmfr_server_close_stream(self._h, st._h))
        self._issued.pop(st._h.value if isinstance(st._h, ctypes.c_void_p) else st._h, None)
        st._h = None

    def step(self, streams, frames, matrices, offsets):
        n = len(streams)
        arr = (ctypes.c_void_p * n)(*[s.handle for s in streams])
        fr = np.ascontiguousarray(frames, np.int32)
        m = _f32(matrices, 16 * n)
        o = _f32(offsets, 2 * n)
        sid = ctypes.c_int64()
        _check(self._lib.bmfr_server_step(self._h, n, arr, fr.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                          _fp(m), _fp(o), ctypes.byref(sid)))
        return sid.value

    def query(self, step_id):
        """(done, gpu_ms) of a step id returned by step()."""
        done, ms = ctypes.c_int32(), ctypes.c_float()
        _check(self._lib.bmfr_server_step_query(self._h, int(step_id), ctypes.byref(done), ctypes.byref(ms)))
        return bool(done.value), ms.value

    def synchronize(self):
        _check(self._lib.bmfr_server_synchronize(self._h))

    def set_fused(self, level=2):
        _check(self._lib.bmfr_server_set_fused(self._h, int(level)))

    def set_policy(self, max_in_flight=2, max_streams_per_step=64):
        _check(self._lib.bmfr_server_set_policy(self._h, int(max_in_flight), int(max_streams_per_step)))

    def submit(self, stream, frame, prev_frame_camera_matrix, pixel_offset, inputs=None, surface=None,
               surface_format=SURFACE_RGB_F32, surface_row_pitch=None, producer_stream=None, t_arrival=None,
               tag=0):
        """Queues one frame of `stream` for the request loop.  `inputs`: four
        device pointers (albedo, normals, positions, noisy) or None for the
        stream's own input buffers."""
        r = BmfrFrameRequest()
        r.stream = stream.handle.value if isinstance(stream.handle, ctypes.c_void_p) else stream.handle
        r.frame_number = int(frame)
        r.prev_frame_camera_matrix[:] = [float(v) for v in _f32(prev_frame_camera_matrix, 16)]
        r.pixel_offset[:] = [float(v) for v in _f32(pixel_offset, 2)]
        if inputs is not None:
            r.albedo, r.normals, r.positions, r.noisy = inputs
        if surface is not None:
            r.surface = surface
            r.surface_format = int(surface_format)
            r.surface_row_pitch = int(surface_row_pitch if surface_row_pitch is not None
                                      else stream.params.width * SURFACE_LAYOUT[surface_format][0])
        r.producer_stream = producer_stream
        r.t_arrival = clock_seconds() if t_arrival is None else float(t_arrival)
        r.tag = int(tag)
        _check(self._lib.bmfr_server_submit(self._h, ctypes.byref(r)))

    def poll(self, max_events=256, timeout_us=0):
        """Events since the last poll, as dicts; an admitted stream's event
        carries its Stream under "stream"."""
        buf = (BmfrServerEvent * max_events)()
        n = ctypes.c_int32()
        _check(self._lib.bmfr_server_poll(self._h, buf, max_events, int(timeout_us), ctypes.byref(n)))
        out = []
        for e in buf[:n.value]:
            d = {k: getattr(e, k) for k, _ in BmfrServerEvent._fields_}
            if e.kind == EVENT_STREAM_ADMITTED:
                params = getattr(self, "_pending_params", {}).pop(int(e.tag), None)
                d["stream"] = self._wrap(e.stream, params) if e.status == OK else None
            else:
                d["stream"] = self._issued.get(e.stream)
            out.append(d)
        return out


class Scene:
    """Seeded procedural plane+spheres scene (the input producer)."""

    def __init__(self, seed, spheres, firefly_probability, width, height, static_camera=False):
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        _check(self._lib.bmfr_scene_create(seed, spheres, firefly_probability, width, height,
                                           1 if static_camera else 0, ctypes.byref(h)))
        self._h = h
        self.width, self.height = width, height

    def close(self):
        if self._h:
            self._lib.bmfr_scene_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def camera(self, frame):
        vp = np.empty(16, np.float32)
        off = np.empty(2, np.float32)
        _check(self._lib.bmfr_scene_camera(self._h, frame, _fp(vp), _fp(off)))
        return vp, off

    def render_host(self, frame):
        n = self.width * self.height * 3
        bufs = [np.empty(n, np.float32) for _ in range(4)]
        _check(self._lib.bmfr_scene_render_host(self._h, frame, *[_fp(b) for b in bufs]))
        return dict(zip(("albedo", "normals", "positions", "noisy"), bufs))

    def render_to_stream(self, frame, stream):
        _check(self._lib.bmfr_scene_render_to_stream(self._h, frame, stream.handle))

    def render_device(self, frame, storage, albedo, normals, positions, noisy, hip_stream=None):
        """Renders a frame into caller-owned device buffers (int pointers)."""
        _check(self._lib.bmfr_scene_render_device(self._h, frame, int(storage), *[
            ctypes.c_void_p(p) for p in (albedo, normals, positions, noisy, hip_stream)]))

    def frame_args(self, frame):
        """(prev_frame_camera_matrix, pixel_offset) of `frame` as the reference
        binds them: camera_matrices[max(frame-1, 0)], pixel_offsets[frame]
        (opencl/bmfr.cpp:440-444)."""
        vp, _ = self.camera(max(frame - 1, 0))
        _, off = self.camera(frame)
        return vp, off
