"""ctypes binding of the reference runner (oracle/_ref/libbmfr_ref.so).

TEST INFRASTRUCTURE: runs the reference's own OpenCL kernels (opencl/bmfr.cl
compiled unmodified for gfx950, see oracle/Makefile) on the GPU through the
ROCm OpenCL runtime, with the reference host's buffers, argument bindings and
swaps (opencl/bmfr.cpp:315-484).  Available only where oracle/_ref/ was built
and an OpenCL GPU device is visible.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")
LIB_PATH = os.path.join(REF_DIR, "libbmfr_ref.so")

# buffer ids (ref_runner.cpp)
IDS = {"normals_cur": 0, "normals_prev": 1, "positions_cur": 2, "positions_prev": 3,
       "noisy_cur": 4, "noisy_prev": 5, "tmp_data": 6, "filtered": 7, "out_cur": 8,
       "out_prev": 9, "result_cur": 10, "result_prev": 11, "prev_pixel": 12, "accept": 13,
       "albedo": 14, "tone_mapped": 15, "weights": 16, "mins_maxs": 17, "spp_cur": 18,
       "spp_prev": 19}
U8 = {"accept", "spp_cur", "spp_prev"}

_lib = None


def available():
    return os.path.exists(LIB_PATH)


def load():
    global _lib
    if _lib is None:
        lib = ctypes.CDLL(LIB_PATH)
        V = ctypes.c_void_p
        lib.ref_error.restype = ctypes.c_char_p
        lib.ref_device_count.restype = ctypes.c_int
        lib.ref_open.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, ctypes.c_int, ctypes.POINTER(V)]
        lib.ref_upload.argtypes = [V, V, V, V, V]
        lib.ref_stage.argtypes = [V, ctypes.c_int, ctypes.c_int, V, V]
        lib.ref_end_frame.argtypes = [V]
        lib.ref_stage_ms.argtypes = [V, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        lib.ref_buffer_bytes.argtypes = [V, ctypes.c_int, ctypes.POINTER(ctypes.c_size_t)]
        lib.ref_read.argtypes = [V, ctypes.c_int, V, ctypes.c_size_t]
        lib.ref_write.argtypes = [V, ctypes.c_int, V, ctypes.c_size_t]
        lib.ref_close.argtypes = [V]
        _lib = lib
    return _lib


# Feature sets that oracle/Makefile also compiles the reference for (the
# FEATURE_BUFFERS strings of its -D options): name -> (factor pairs,
# not_scaled) in the oracle's ids (1, normal.xyz, world_position.xyz).
FEATURE_SETS = {
    "cross5": ([(0, 0), (2, 0), (4, 5), (6, 0), (4, 6)], 2),
    "quad16": ([(0, 0), (1, 0), (2, 0), (3, 0), (4, 0), (5, 0), (6, 0), (4, 4), (5, 5), (6, 6),
                (4, 5), (5, 6), (4, 6), (1, 4), (2, 5), (3, 6)], 4),
}


def code_object(w, h, half, feature_set=None):
    fs = f"_{feature_set}" if feature_set else ""
    return os.path.join(REF_DIR, f"bmfr_{w}x{h}_h{1 if half else 0}{fs}.co")


class RefRun:
    """The reference's frame chain on the GPU, stage by stage."""

    def __init__(self, w, h, half=True, feature_set=None):
        lib = load()
        self.lib, self.w, self.h, self.half = lib, w, h, half
        n, ns = 10, 4
        if feature_set:
            pairs, ns = FEATURE_SETS[feature_set]
            n = len(pairs)
        hd = ctypes.c_void_p()
        rc = lib.ref_open(code_object(w, h, half, feature_set).encode(), w, h, 1 if half else 0, n, n - ns,
                          ctypes.byref(hd))
        if rc != 0:
            raise RuntimeError(f"ref_open: {lib.ref_error().decode()}")
        self.h_ = hd

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(self.lib.ref_error().decode())

    def upload(self, albedo, normals, positions, noisy):
        arrs = [np.ascontiguousarray(a, np.float32) for a in (albedo, normals, positions, noisy)]
        self._check(self.lib.ref_upload(self.h_, *[a.ctypes.data for a in arrs]))

    def stage(self, i, frame, matrix=None, offset=None):
        m = np.ascontiguousarray(matrix if matrix is not None else np.zeros(16), np.float32)
        o = np.ascontiguousarray(offset if offset is not None else np.zeros(2), np.float32)
        self._check(self.lib.ref_stage(self.h_, i, frame, m.ctypes.data, o.ctypes.data))

    def run_frame(self, frame, matrix, offset):
        for i in range(5):
            self.stage(i, frame, matrix, offset)
        self.end_frame()

    def end_frame(self):
        self._check(self.lib.ref_end_frame(self.h_))

    def stage_ms(self, i):
        ms = ctypes.c_double()
        self._check(self.lib.ref_stage_ms(self.h_, i, ctypes.byref(ms)))
        return ms.value

    def dtype(self, name):
        if name in U8:
            return np.uint8
        if name == "tmp_data":
            return np.float16 if self.half else np.float32
        return np.float32

    def read(self, name, count=None):
        """Buffer contents; `count` elements (default: the whole buffer)."""
        dt = np.dtype(self.dtype(name))
        n = ctypes.c_size_t()
        self._check(self.lib.ref_buffer_bytes(self.h_, IDS[name], ctypes.byref(n)))
        count = n.value // dt.itemsize if count is None else count
        out = np.empty(count, dt)
        self._check(self.lib.ref_read(self.h_, IDS[name], out.ctypes.data, out.nbytes))
        return out

    def write(self, name, array):
        a = np.ascontiguousarray(array, self.dtype(name))
        self._check(self.lib.ref_write(self.h_, IDS[name], a.ctypes.data, a.nbytes))

    def close(self):
        if self.h_:
            self.lib.ref_close(self.h_)
            self.h_ = None
