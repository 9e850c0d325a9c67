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
