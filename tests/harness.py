"""Shared helpers: drive the HIP stream (through the C ABI) and the CPU oracle
on identical inputs, frame by frame, the way opencl/bmfr.cpp:417-485 drives
the reference kernels."""
import numpy as np

import bmfr_amd
from oracle import bind

SEED = 0xB3F40000


def scene(w, h, spheres=8, firefly=1e-3, static=False, seed=SEED):
    return bmfr_amd.Scene(seed, spheres, firefly, w, h, static_camera=static)


def f16_round(a):
    return np.asarray(a, np.float32).astype(np.float16).astype(np.float32)


def fma32(a, b, c):
    """Elementwise float32 fused multiply-add, RN(a*b + c) with one rounding.
    a*b is exact in float64 (24 + 24 bits); s = a*b + c rounds in float64,
    and TwoSum gives that rounding's exact error e.  RN32(s) is RN32 of the
    exact value unless s lies exactly halfway between two float32 values,
    where the sign of e decides the side."""
    a, b, c = (np.asarray(v, np.float32).astype(np.float64) for v in (a, b, c))
    p = a * b
    s = p + c
    bb = s - p
    e = (p - (s - bb)) + (c - bb)
    r = s.astype(np.float32)
    r64 = r.astype(np.float64)
    up = np.nextafter(r, np.float32(np.inf)).astype(np.float64)
    dn = np.nextafter(r, np.float32(-np.inf)).astype(np.float64)
    with np.errstate(invalid="ignore", over="ignore"):
        at_up_mid = (s == (r64 + up) / 2) & (e > 0)
        at_dn_mid = (s == (r64 + dn) / 2) & (e < 0)
    r = np.where(at_up_mid, up.astype(np.float32), r)
    r = np.where(at_dn_mid, dn.astype(np.float32), r)
    return r.astype(np.float32)


def gpu_params(w, h, tmp_f16=True, storage_f16=False, **kw):
    return bmfr_amd.default_params(
        w, h, tmp_storage=bmfr_amd.STORAGE_F16 if tmp_f16 else bmfr_amd.STORAGE_F32,
        storage=bmfr_amd.STORAGE_F16 if storage_f16 else bmfr_amd.STORAGE_F32, **kw)


# The reference's feature set as explicit factor pairs (bmfr.cpp:65-77):
# 1, normal.xyz unscaled; world_position.xyz and its squares scaled.
REFERENCE_SET = ([(0, 0), (1, 0), (2, 0), (3, 0), (4, 0), (5, 0), (6, 0), (4, 4), (5, 5), (6, 6)], 4)


def set_factor_pairs(p, pairs, not_scaled):
    """Writes a feature set into library or oracle params (same fields)."""
    p.feature_count = len(pairs)
    p.features_not_scaled = not_scaled
    for i in range(16):
        a, b = pairs[i] if i < len(pairs) else (0, 0)
        p.feature_factors[i][0] = a
        p.feature_factors[i][1] = b


def copy_features(dst, src):
    dst.feature_count = src.feature_count
    dst.features_not_scaled = src.features_not_scaled
    for i in range(16):
        dst.feature_factors[i][0] = src.feature_factors[i][0]
        dst.feature_factors[i][1] = src.feature_factors[i][1]


class Lockstep:
    """A GPU stream and an oracle stream fed the same frames.  `features` is
    the reference's pair of feature strings (NOT_SCALED, SCALED), parsed by
    the library; `pairs` an explicit (factor pairs, not_scaled)."""

    def __init__(self, w, h, tmp_f16=True, storage_f16=False, sc=None, fitter_cfg=None, features=None,
                 pairs=None):
        self.w, self.h = w, h
        self.storage_f16 = storage_f16
        self.sc = sc or scene(w, h)
        gp = gpu_params(w, h, tmp_f16, storage_f16)
        if features is not None:
            bmfr_amd.parse_feature_buffers(gp, *features)
        if pairs is not None:
            set_factor_pairs(gp, *pairs)
        self.g = bmfr_amd.Stream(gp)
        if fitter_cfg:
            self.g.set_fitter_config(*fitter_cfg)
        op = bind.params(w, h, half_tmp=tmp_f16)
        copy_features(op, gp)
        self.o = bind.OracleStream(op)
        self.features = bind.feature_count(op)

    def upload(self, frame):
        f = self.sc.render_host(frame)
        self.g.upload_frame(f["albedo"], f["normals"], f["positions"], f["noisy"])
        if self.storage_f16:
            # the GPU stores these as binary16: feed the oracle the same values
            f = {k: (v if k == "albedo" else f16_round(v)) for k, v in f.items()}
        self.o.upload(f["albedo"], f["normals"], f["positions"], f["noisy"])
        return f

    def read(self, name):
        a = self.g.read_buffer(name)
        return a.astype(np.float32) if a.dtype == np.float16 else a

    def close(self):
        self.g.close()


def rel_err(a, b, floor=1e-2):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) / np.maximum(np.abs(b), floor)


def psnr(img, ref):
    mse = np.mean((np.asarray(img, np.float64) - np.asarray(ref, np.float64)) ** 2)
    return float("inf") if mse == 0 else 10 * np.log10(1.0 / mse)


class Hip:
    """The few HIP runtime calls the tests make themselves (device buffers a
    renderer would own, and a producer HIP stream), through the runtime's C
    API; plumbing only."""
    _rt = None

    @classmethod
    def rt(cls):
        if cls._rt is None:
            import ctypes
            rt = ctypes.CDLL("libamdhip64.so")
            vp = ctypes.c_void_p
            rt.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
            rt.hipFree.argtypes = [vp]
            rt.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
            rt.hipMemset.argtypes = [vp, ctypes.c_int, ctypes.c_size_t]
            rt.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
            rt.hipStreamDestroy.argtypes = [vp]
            rt.hipStreamSynchronize.argtypes = [vp]
            rt.hipDeviceSynchronize.argtypes = []
            cls._rt = rt
        return cls._rt

    @classmethod
    def check(cls, err, what):
        assert err == 0, f"{what} failed: hipError {err}"


class DeviceBuffer:
    """A caller-owned device allocation (int pointer in .ptr)."""

    def __init__(self, nbytes, fill=None):
        import ctypes
        p = ctypes.c_void_p()
        Hip.check(Hip.rt().hipMalloc(ctypes.byref(p), int(nbytes)), "hipMalloc")
        self.ptr, self.nbytes = p.value, int(nbytes)
        if fill is not None:
            Hip.check(Hip.rt().hipMemset(self.ptr, int(fill), self.nbytes), "hipMemset")

    def read(self, dtype=np.uint8):
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize, dtype)
        Hip.check(Hip.rt().hipDeviceSynchronize(), "hipDeviceSynchronize")
        Hip.check(Hip.rt().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2), "hipMemcpy D2H")
        return out

    def free(self):
        if self.ptr:
            Hip.rt().hipFree(self.ptr)
            self.ptr = None


class HipStream:
    def __init__(self):
        import ctypes
        p = ctypes.c_void_p()
        Hip.check(Hip.rt().hipStreamCreate(ctypes.byref(p)), "hipStreamCreate")
        self.handle = p.value

    def synchronize(self):
        Hip.check(Hip.rt().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def destroy(self):
        if self.handle:
            Hip.rt().hipStreamDestroy(self.handle)
            self.handle = None


class FrameInputs:
    """One renderer-owned set of a frame's G-buffers: albedo float[3P], the
    others float or half [3P] per the stream's storage."""

    def __init__(self, w, h, storage_f16):
        n = w * h * 3
        s = 2 if storage_f16 else 4
        self.storage = 1 if storage_f16 else 0
        self.bufs = [DeviceBuffer(4 * n), DeviceBuffer(s * n), DeviceBuffer(s * n), DeviceBuffer(s * n)]

    @property
    def ptrs(self):
        return tuple(b.ptr for b in self.bufs)

    def render(self, sc, frame, hip_stream=None):
        sc.render_device(frame, self.storage, *self.ptrs, hip_stream=hip_stream)

    def free(self):
        for b in self.bufs:
            b.free()


def quantize_u8(v):
    """round(clamp(v, 0, 1) * 255) in float32 with NaN -> 0: the 8-bit
    conversion of the reference's PNG output (opencl/bmfr.cpp:519-547)."""
    v = np.asarray(v, np.float32)
    c = np.where(v > 0, np.where(v < 1, v, np.float32(1)), np.float32(0)).astype(np.float32)
    return (c * np.float32(255) + np.float32(0.5)).astype(np.float32).astype(np.uint8)
