"""Dataset-format IO and the reference's frame loop over a dataset directory.

The reference reads, per frame NN, ``color``, ``shading_normal``,
``world_position`` and ``albedo`` ``NN.exr`` (opencl/bmfr.cpp:49-53,
145-172), takes the per-frame view-projection matrices, pixel offsets and
reprojection limits from the dataset's ``camera_matrices.h``
(opencl/bmfr.cpp:43-47, 226-227, 440-444), runs the five stages per frame
(opencl/bmfr.cpp:417-485) and writes ``outputs/outputNN.png``
(opencl/bmfr.cpp:519-547).  The readers and writers are the library's own
(no OpenImageIO): scanline OpenEXR with NONE/RLE/ZIPS/ZIP compression, PNG
and FP32 EXR out.
"""
import ctypes
import os

import numpy as np

from . import _check, _fp, _lib

# file name stems, opencl/bmfr.cpp:49-53
NOISY, NORMAL, POSITION, ALBEDO = "color", "shading_normal", "world_position", "albedo"
EXR_NONE, EXR_ZIP = 0, 3


def exr_info(path):
    lib = _lib.load()
    w, h, c = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    _check(lib.bmfr_exr_info(os.fsencode(path), ctypes.byref(w), ctypes.byref(h), ctypes.byref(c)))
    return w.value, h.value, c.value


def read_exr_rgb(path, width=None, height=None):
    """(H, W, 3) float32 RGB of an OpenEXR file."""
    if width is None or height is None:
        width, height, _ = exr_info(path)
    out = np.empty((height, width, 3), np.float32)
    _check(_lib.load().bmfr_exr_read_rgb(os.fsencode(path), width, height, _fp(out)))
    return out


def write_exr_rgb(path, rgb, compression=EXR_ZIP):
    a = np.ascontiguousarray(rgb, np.float32)
    h, w = a.shape[:2]
    _check(_lib.load().bmfr_exr_write_rgb(os.fsencode(path), w, h, _fp(a), w, compression))


def write_png_rgb(path, rgb):
    """8-bit RGB PNG of an (H, W, 3) float image in [0, 1]."""
    a = np.ascontiguousarray(rgb, np.float32)
    h, w = a.shape[:2]
    _check(_lib.load().bmfr_png_write_rgb(os.fsencode(path), w, h, _fp(a), w))


def load_camera(path, frames):
    """(matrices (frames, 16), offsets (frames, 2), position_limit_squared,
    normal_limit_squared) from a camera_matrices.h."""
    m = np.empty((frames, 16), np.float32)
    o = np.empty((frames, 2), np.float32)
    pl, nl = ctypes.c_float(), ctypes.c_float()
    _check(_lib.load().bmfr_camera_matrices_load(os.fsencode(path), frames, _fp(m), _fp(o),
                                                 ctypes.byref(pl), ctypes.byref(nl)))
    return m, o, pl.value, nl.value


def load_frame(data_dir, frame, width, height):
    """The four input images of `frame` as flat float32 arrays of W*H*3."""
    return {k: read_exr_rgb(os.path.join(data_dir, f"{stem}{frame}.exr"), width, height).ravel()
            for k, stem in (("albedo", ALBEDO), ("normals", NORMAL), ("positions", POSITION),
                            ("noisy", NOISY))}


def run(data_dir, frames, out_dir=None, width=None, height=None, features=None, **params):
    """The reference's tasks() loop on a dataset directory: every frame's
    inputs uploaded, the five stages run, the result written to
    out_dir/outputNN.png.  The previous frame's matrix is used for
    reprojection (matrix_index = max(frame - 1, 0), opencl/bmfr.cpp:440-442).
    `features` is an optional (NOT_SCALED_FEATURE_BUFFERS,
    SCALED_FEATURE_BUFFERS) pair of the reference's feature strings
    (opencl/bmfr.cpp:65-77).  Returns the list of (H, W, 3) results."""
    from . import Stream, default_params, parse_feature_buffers
    if width is None or height is None:
        width, height, _ = exr_info(os.path.join(data_dir, f"{NOISY}0.exr"))
    mats, offs, pl, nl = load_camera(os.path.join(data_dir, "camera_matrices.h"), frames)
    p = default_params(width, height, position_limit_squared=pl, normal_limit_squared=nl, **params)
    if features is not None:
        parse_feature_buffers(p, *features)
    results = []
    with Stream(p) as st:
        for f in range(frames):
            g = load_frame(data_dir, f, width, height)
            st.upload_frame(g["albedo"], g["normals"], g["positions"], g["noisy"])
            st.run_frame(f, mats[max(f - 1, 0)], offs[f])
            res = st.download_result().reshape(height, width, 3)
            results.append(res)
            if out_dir:
                os.makedirs(out_dir, exist_ok=True)
                write_png_rgb(os.path.join(out_dir, f"output{f}.png"), res)
    return results
