"""ctypes binding of the CPU oracle (oracle/_build/libbmfr_oracle.so).

TEST INFRASTRUCTURE: only the tests, the smoke check and the CPU-baseline
leg of bench.py use this; the product library never does.  Each wrapper
mirrors one stage of opencl/bmfr.cl (see bmfr_oracle.h for citations) and
works on numpy arrays in the reference's buffer layouts.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libbmfr_oracle.so")


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_int32),
        ("height", ctypes.c_int32),
        ("position_limit_squared", ctypes.c_float),
        ("normal_limit_squared", ctypes.c_float),
        ("blend_alpha", ctypes.c_float),
        ("second_blend_alpha", ctypes.c_float),
        ("taa_blend_alpha", ctypes.c_float),
        ("noise_amount", ctypes.c_double),
        ("half_tmp", ctypes.c_int32),
        ("feature_count", ctypes.c_int32),          # 0: the reference's set
        ("features_not_scaled", ctypes.c_int32),
        ("feature_factors", (ctypes.c_uint8 * 2) * 16),
    ]


class OrcSizes(ctypes.Structure):
    _fields_ = [
        ("workset_w", ctypes.c_int32), ("workset_h", ctypes.c_int32),
        ("margin_w", ctypes.c_int32), ("margin_h", ctypes.c_int32),
        ("blocks_x", ctypes.c_int32), ("blocks_y", ctypes.c_int32),
        ("blocks", ctypes.c_int32), ("pixels", ctypes.c_int64), ("padded", ctypes.c_int64),
    ]


_P = ctypes.c_void_p
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER(OrcParams)
    sig = {
        "orc_sizes_of": (None, [ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(OrcSizes)]),
        "orc_default_params": (None, [ctypes.c_int32, ctypes.c_int32, P]),
        "orc_block_offset": (None, [ctypes.c_int32, _P]),
        "orc_mirror": (ctypes.c_int32, [ctypes.c_int32, ctypes.c_int32]),
        "orc_hash": (ctypes.c_uint32, [ctypes.c_uint32]),
        "orc_random": (ctypes.c_float, [ctypes.c_uint32]),
        "orc_add_random": (ctypes.c_float, [ctypes.c_float, ctypes.c_int32, ctypes.c_int32,
                                            ctypes.c_int32, ctypes.c_int32, ctypes.c_double]),
        "orc_add_random_n": (ctypes.c_float, [ctypes.c_float, ctypes.c_int32, ctypes.c_int32,
                                              ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                                              ctypes.c_int32]),
        "orc_feature_count": (ctypes.c_int32, [P]),
        "orc_not_scaled": (ctypes.c_int32, [P]),
        "orc_feature": (ctypes.c_float, [P, ctypes.c_int32, _P, _P]),
        "orc_f32_to_f16": (ctypes.c_uint16, [ctypes.c_float]),
        "orc_f16_to_f32": (ctypes.c_float, [ctypes.c_uint16]),
        "orc_scale": (ctypes.c_float, [ctypes.c_float, ctypes.c_float, ctypes.c_float]),
        "orc_rgb_to_ycocg": (None, [_P, _P]),
        "orc_ycocg_to_rgb": (None, [_P, _P]),
        "orc_tree_sum": (ctypes.c_float, [_P]),
        "orc_tree_min": (ctypes.c_float, [_P]),
        "orc_tree_max": (ctypes.c_float, [_P]),
        "orc_accumulate_noisy_data": (ctypes.c_int32, [P] + [_P] * 12 + [_P, _P, ctypes.c_int32]),
        "orc_fitter": (None, [P, _P, _P, _P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
        "orc_fitter_inputs": (None, [P, _P, ctypes.c_int32, ctypes.c_int32, _P, _P, _P]),
        "orc_lstsq_fp64": (ctypes.c_double, [_P, _P, ctypes.c_int32, ctypes.c_int32,
                                             ctypes.c_int32, _P]),
        "orc_weighted_sum": (None, [P, _P, _P, _P, _P, _P, ctypes.c_int32]),
        "orc_accumulate_filtered_data": (None, [P] + [_P] * 8 + [ctypes.c_int32]),
        "orc_taa": (None, [P, _P, _P, _P, _P, ctypes.c_int32]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(a):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


# factor ids of a feature term: 1, normal.xyz, world_position.xyz
FACTORS = {"1": 0, "n.x": 1, "n.y": 2, "n.z": 3, "p.x": 4, "p.y": 5, "p.z": 6}


def params(width, height, half_tmp=True, features=None, not_scaled=None, **kw):
    """Oracle parameters; `features` (optional) is a list of terms such as
    "1", "n.y", "p.x*p.z" (at most 16), the first `not_scaled` unscaled."""
    p = OrcParams()
    load().orc_default_params(width, height, ctypes.byref(p))
    p.half_tmp = 1 if half_tmp else 0
    if features is not None:
        set_features(p, features, not_scaled)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def factor_pairs(features):
    """[(a, b)] factor ids of feature terms ("p.x*p.y" -> (4, 5), "n.z" -> (3, 0))."""
    pairs = []
    for t in features:
        f = t.split("*")
        assert 1 <= len(f) <= 2, t
        pairs.append((FACTORS[f[0]], FACTORS[f[1]] if len(f) == 2 else 0))
    return pairs


def set_features(p, features, not_scaled):
    pairs = factor_pairs(features)
    assert 1 <= len(pairs) <= 16 and 0 <= not_scaled <= len(pairs)
    p.feature_count = len(pairs)
    p.features_not_scaled = not_scaled
    for i, (a, b) in enumerate(pairs):
        p.feature_factors[i][0] = a
        p.feature_factors[i][1] = b


def feature_count(p):
    return p.feature_count or 10


def scaled_count(p):
    return p.feature_count - p.features_not_scaled if p.feature_count else 6


def sizes(width, height):
    s = OrcSizes()
    load().orc_sizes_of(width, height, ctypes.byref(s))
    return s


class OracleStream:
    """CPU restatement of one stream: the reference's buffer set with its
    double buffers (opencl/bmfr.cpp:315-347, 482-484), driven stage by stage
    like the frame loop of opencl/bmfr.cpp:417-485."""

    def __init__(self, p):
        self.p = p
        s = sizes(p.width, p.height)
        self.s = s
        P = p.width * p.height
        z3 = lambda: np.zeros(3 * P, np.float32)
        self.albedo = z3()
        self.normals = [z3(), z3()]
        self.positions = [z3(), z3()]
        self.noisy_in = z3()
        self.noisy_acc = [z3(), z3()]
        self.spp = [np.zeros(P, np.uint8), np.zeros(P, np.uint8)]
        self.prev_pixel = np.zeros(2 * P, np.float32)
        self.accept = np.zeros(P, np.uint8)
        nf, ns = feature_count(p), scaled_count(p)
        self.tmp = np.zeros((nf + 3) * s.padded, np.float32)
        self.weights = np.zeros(3 * nf * s.blocks, np.float32)
        self.mins_maxs = np.zeros(max(2 * ns, 1) * s.blocks, np.float32)
        self.filtered = z3()
        self.tone = z3()
        self.out = [z3(), z3()]
        self.result = [z3(), z3()]
        self.swapped = False

    def cur(self):
        return 0 if self.swapped else 1

    def prev(self):
        return 1 if self.swapped else 0

    def upload(self, albedo, normals, positions, noisy):
        c = self.cur()
        self.albedo[:] = albedo
        self.normals[c][:] = normals
        self.positions[c][:] = positions
        self.noisy_in[:] = noisy

    def accumulate_noisy_data(self, frame, matrix, offset):
        c, pv = self.cur(), self.prev()
        m = np.ascontiguousarray(matrix, np.float32)
        o = np.ascontiguousarray(offset, np.float32)
        rc = load().orc_accumulate_noisy_data(
            ctypes.byref(self.p), _ptr(self.normals[c]), _ptr(self.normals[pv]),
            _ptr(self.positions[c]), _ptr(self.positions[pv]), _ptr(self.noisy_in),
            _ptr(self.noisy_acc[pv]), _ptr(self.spp[pv]), _ptr(self.noisy_acc[c]),
            _ptr(self.prev_pixel), _ptr(self.accept), _ptr(self.spp[c]), _ptr(self.tmp),
            _ptr(m), _ptr(o), frame)
        if rc != 0:
            raise ValueError(f"image {self.p.width}x{self.p.height} below the 47-pixel minimum")

    def fitter(self, frame, first=0, count=None):
        if count is None:
            count = self.s.blocks
        load().orc_fitter(ctypes.byref(self.p), _ptr(self.tmp), _ptr(self.mins_maxs),
                          _ptr(self.weights), frame, first, count)

    def weighted_sum(self, frame):
        c = self.cur()
        load().orc_weighted_sum(ctypes.byref(self.p), _ptr(self.weights), _ptr(self.mins_maxs),
                                _ptr(self.filtered), _ptr(self.normals[c]),
                                _ptr(self.positions[c]), frame)

    def accumulate_filtered_data(self, frame):
        c, pv = self.cur(), self.prev()
        load().orc_accumulate_filtered_data(
            ctypes.byref(self.p), _ptr(self.filtered), _ptr(self.prev_pixel), _ptr(self.accept),
            _ptr(self.albedo), _ptr(self.tone), _ptr(self.spp[c]), _ptr(self.out[pv]),
            _ptr(self.out[c]), frame)

    def taa(self, frame):
        c, pv = self.cur(), self.prev()
        load().orc_taa(ctypes.byref(self.p), _ptr(self.prev_pixel), _ptr(self.tone),
                       _ptr(self.result[c]), _ptr(self.result[pv]), frame)

    def end_frame(self):
        self.swapped = not self.swapped

    def run_frame(self, frame, matrix, offset):
        self.accumulate_noisy_data(frame, matrix, offset)
        self.fitter(frame)
        self.weighted_sum(frame)
        self.accumulate_filtered_data(frame)
        self.taa(frame)
        res = self.result[self.cur()].copy()
        self.end_frame()
        return res


def fitter_inputs(p, tmp, block, frame):
    """(A[1024,N], b[1024,3], mins_maxs[2S]) as the QR sees them."""
    A = np.zeros((1024, feature_count(p)), np.float64)
    b = np.zeros((1024, 3), np.float64)
    mm = np.zeros(max(2 * scaled_count(p), 1), np.float32)
    load().orc_fitter_inputs(ctypes.byref(p), _ptr(tmp), block, frame, _ptr(A), _ptr(b), _ptr(mm))
    return A, b, mm


def lstsq_fp64(A, b):
    A = np.ascontiguousarray(A, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    m, n = A.shape
    k = b.shape[1]
    w = np.zeros((n, k), np.float64)
    rc = load().orc_lstsq_fp64(_ptr(A), _ptr(b), m, n, k, _ptr(w))
    return w, rc
