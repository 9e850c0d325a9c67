"""Configurable feature sets (opencl/bmfr.cpp:63-77, 193-214): the parser of
the reference's feature strings, parameter validation, and the oracle's
feature-set generalisation pinned against

  - the reference's own set written out explicitly (must give the bits of
    the pinned default path, feature_count 0);
  - numpy evaluations of the feature expressions (stage 1's block-matrix
    layout, the weighted sum);
  - FP64 least squares and exact-fit recovery (the fitter with N columns);
  - the noise seed's BUFFER_COUNT stride (bmfr.cl:179-181) from orc_random.

CPU only; the GPU path is compared with this oracle in
tests/test_feature_sets_gpu.py."""
import ctypes

import numpy as np
import pytest

import bmfr_amd
from bmfr_amd import _lib
from oracle import bind
from tests.harness import REFERENCE_SET, copy_features, fma32, set_factor_pairs

lib = bind.load()

REF_NOT_SCALED = "1.f,normal.x,normal.y,normal.z,"           # bmfr.cpp:65-69
REF_SCALED = ("world_position.x,world_position.y,world_position.z,"
              "world_position.x*world_position.x,world_position.y*world_position.y,"
              "world_position.z*world_position.z")              # bmfr.cpp:71-77

# factor id -> name in the reference's expressions
NAMES = {0: "1.f", 1: "normal.x", 2: "normal.y", 3: "normal.z",
         4: "world_position.x", 5: "world_position.y", 6: "world_position.z"}

# A few configured sets: (factor pairs, not_scaled)
SETS = {
    "one": ([(0, 0)], 1),
    "one_scaled": ([(4, 0)], 0),
    "cross5": ([(0, 0), (2, 0), (4, 5), (6, 0), (4, 6)], 2),
    "lin7": ([(0, 0), (1, 0), (2, 0), (3, 0), (4, 0), (5, 0), (6, 0)], 4),
    "quad16": ([(0, 0), (1, 0), (2, 0), (3, 0), (4, 0), (5, 0), (6, 0), (4, 4), (5, 5), (6, 6),
                (4, 5), (5, 6), (4, 6), (1, 4), (2, 5), (3, 6)], 4),
}


def _params(w, h, **kw):
    p = _lib.BmfrParams()
    assert _lib.load().bmfr_default_params(w, h, ctypes.byref(p)) == 0
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _status(p):
    s = _lib.BmfrSizes()
    return _lib.load().bmfr_compute_sizes(ctypes.byref(p), ctypes.byref(s)), s


# ---- parser and validation --------------------------------------------------
def test_reference_strings_select_the_reference_set():
    p = _params(1280, 720)
    bmfr_amd.parse_feature_buffers(p, REF_NOT_SCALED, REF_SCALED)
    assert p.feature_count == 0
    st, s = _status(p)
    assert st == 0
    assert s.device_bytes == _status(_params(1280, 720))[1].device_bytes


def test_configured_set_is_parsed_term_by_term():
    p = _params(256, 256)
    bmfr_amd.parse_feature_buffers(p, " 1.f, normal.y ,",
                                   "world_position.x*world_position.z,world_position.y,normal.z*1")
    assert (p.feature_count, p.features_not_scaled) == (5, 2)
    got = [(p.feature_factors[i][0], p.feature_factors[i][1]) for i in range(5)]
    assert got == [(0, 0), (2, 0), (4, 6), (5, 0), (3, 0)]
    # counted like bmfr.cpp:193-201: commas of the first string, commas + 1 of the second
    p2 = _params(256, 256)
    bmfr_amd.parse_feature_buffers(p2, "1.f,", "world_position.x")
    assert (p2.feature_count, p2.features_not_scaled) == (2, 1)


@pytest.mark.parametrize("ns,sc", [
    ("1.f,albedo.x,", "world_position.x"),                 # not in scope of the kernels
    ("1.f,", "normal.x*normal.y*normal.z"),                 # three factors
    ("1.f,", "2.f"),
    ("", ""),                                              # no feature
    (",".join(["normal.x"] * 9) + ",", ",".join(["world_position.y"] * 8)),   # 17 terms
])
def test_parse_errors(ns, sc):
    p = _params(256, 256)
    with pytest.raises(bmfr_amd.BmfrError) as e:
        bmfr_amd.parse_feature_buffers(p, ns, sc)
    assert e.value.status == bmfr_amd.ERR_INVALID_ARGUMENT


def test_validation_of_configured_sets():
    bad = []
    p = _params(96, 64); p.feature_count = 17; bad.append(p)
    p = _params(96, 64); set_factor_pairs(p, *SETS["cross5"]); p.features_not_scaled = 6; bad.append(p)
    p = _params(96, 64); set_factor_pairs(p, *SETS["cross5"]); p.feature_factors[3][1] = 7; bad.append(p)
    p = _params(96, 64, qr_mode=bmfr_amd.QR_KAHAN); set_factor_pairs(p, *SETS["cross5"]); bad.append(p)
    for p in bad:
        assert _status(p)[0] == bmfr_amd.ERR_INVALID_ARGUMENT
    p = _params(96, 64); set_factor_pairs(p, *SETS["cross5"])
    assert _status(p)[0] == 0


def test_device_footprint_follows_the_feature_count():
    """tmp_data (N+3 columns), weights (3N), mins_maxs (2S) and the noise
    table ((N-1) x 1024 doubles) scale with the set."""
    base = _status(_params(256, 256, tmp_storage=bmfr_amd.STORAGE_F32))[1]
    prev = None
    for name in ["one", "cross5", "lin7", "quad16"]:
        p = _params(256, 256, tmp_storage=bmfr_amd.STORAGE_F32)
        set_factor_pairs(p, *SETS[name])
        s = _status(p)[1]
        n = p.feature_count
        # tmp_data dominates: (n - 10) columns of 4-byte elements over Q texels
        d = s.device_bytes - base.device_bytes
        assert abs(d - (n - 10) * 4 * s.padded_pixels) < 64 * s.blocks + 16 * 1024 * 8
        if prev is not None:
            assert s.device_bytes > prev
        prev = s.device_bytes


# ---- the oracle with a feature set ------------------------------------------
def _scene(w, h, seed=0xB3F40000):
    return bmfr_amd.Scene(seed, 8, 1e-3, w, h)


def _oracle(w, h, half, pairs=None):
    p = bind.params(w, h, half_tmp=half)
    if pairs is not None:
        set_factor_pairs(p, *pairs)
    return p, bind.OracleStream(p)


def _run_stage12(o, sc, frames):
    for fr in range(frames):
        f = sc.render_host(fr)
        o.upload(f["albedo"], f["normals"], f["positions"], f["noisy"])
        vp, off = sc.frame_args(fr)
        o.accumulate_noisy_data(fr, vp, off)
        if fr < frames - 1:
            o.end_frame()


@pytest.mark.parametrize("half", [True, False])
def test_explicit_reference_set_gives_the_default_bits(half):
    w, h = 70, 48
    sc = _scene(w, h)
    _, o0 = _oracle(w, h, half)
    _, o1 = _oracle(w, h, half, REFERENCE_SET)
    for fr in range(3):
        f = sc.render_host(fr)
        vp, off = sc.frame_args(fr)
        for o in (o0, o1):
            o.upload(f["albedo"], f["normals"], f["positions"], f["noisy"])
            o.accumulate_noisy_data(fr, vp, off)
            o.fitter(fr)
            o.weighted_sum(fr)
        for name in ["tmp", "weights", "mins_maxs", "filtered"]:
            assert np.array_equal(getattr(o0, name), getattr(o1, name)), (fr, name)
        o0.end_frame()
        o1.end_frame()


def _factor_arrays(nrm, pos):
    one = np.ones_like(pos[..., 0])
    return [one, nrm[..., 0], nrm[..., 1], nrm[..., 2], pos[..., 0], pos[..., 1], pos[..., 2]]


def _offset(frame):
    o = (ctypes.c_int32 * 2)()
    lib.orc_block_offset(frame, o)
    return o[0], o[1]


@pytest.mark.parametrize("name", ["one", "cross5", "quad16"])
def test_stage1_writes_the_set_in_block_layout(name):
    """Frame 0: tmp_data is [by][bx][feature][y][x] of the mirrored,
    jittered grid with N + 3 columns, each feature the float product of its
    two factors (bmfr.cl:448-476)."""
    w, h = 70, 48
    sc = _scene(w, h)
    f = sc.render_host(0)
    pairs, _ = SETS[name]
    p, o = _oracle(w, h, False, SETS[name])
    o.upload(f["albedo"], f["normals"], f["positions"], f["noisy"])
    vp, off = sc.frame_args(0)
    o.accumulate_noisy_data(0, vp, off)
    s, n = o.s, len(pairs)
    t = o.tmp.reshape(s.blocks_y, s.blocks_x, n + 3, 32, 32)
    grid = t.transpose(2, 0, 3, 1, 4).reshape(n + 3, s.margin_h, s.margin_w)
    ox, oy = _offset(0)
    gy, gx = np.mgrid[0:s.margin_h, 0:s.margin_w]
    mx = np.where(gx - 16 + ox < 0, -(gx - 16 + ox) - 1, gx - 16 + ox)
    mx = np.where(mx >= w, 2 * w - mx - 1, mx)
    my = np.where(gy - 16 + oy < 0, -(gy - 16 + oy) - 1, gy - 16 + oy)
    my = np.where(my >= h, 2 * h - my - 1, my)
    pos = f["positions"].reshape(h, w, 3)[my, mx]
    nrm = f["normals"].reshape(h, w, 3)[my, mx]
    col = f["noisy"].reshape(h, w, 3)[my, mx]
    fa = _factor_arrays(nrm, pos)
    want = [(fa[a] * fa[b]).astype(np.float32) for a, b in pairs] + [col[..., c] for c in range(3)]
    assert np.array_equal(grid, np.stack(want).astype(np.float32))


def test_noise_seed_strides_by_the_buffer_count():
    """bmfr.cl:179-181: random(id + 256 sv + 1024 f + 1024 BUFFER_COUNT frame)."""
    for (row, fb, frame, B) in [(0, 1, 0, 4), (1023, 15, 7, 19), (17, 4, 59, 8), (300, 2, 3, 13)]:
        arg = row + 1024 * fb + 1024 * B * frame
        r = np.float32(lib.orc_random(arg)) - np.float32(0.5)
        want = np.float32(np.float64(np.float32(-1.5)) + 1e-2 * 2.0 * np.float64(r))
        got = lib.orc_add_random_n(-1.5, row % 256, row // 256, fb, frame, 1e-2, B)
        assert np.float32(got) == want
        if B == 13:
            assert got == lib.orc_add_random(-1.5, row % 256, row // 256, fb, frame, 1e-2)


@pytest.mark.parametrize("name", ["one", "one_scaled", "cross5", "lin7", "quad16"])
def test_configured_fitter_matches_fp64_least_squares(name):
    w, h = 70, 48
    pairs, ns = SETS[name]
    n = len(pairs)
    p, o = _oracle(w, h, False, SETS[name])
    _run_stage12(o, _scene(w, h), 3)
    tmp0 = o.tmp.copy()
    o.fitter(2)
    s = o.s
    W = o.weights.reshape(s.blocks, n, 3).astype(np.float64)
    worst, used = 0.0, 0
    for blk in range(s.blocks):
        A, b, mm = bind.fitter_inputs(p, tmp0, blk, 2)
        assert A.shape == (1024, n)
        if n - ns:
            assert np.array_equal(mm, o.mins_maxs[blk * 2 * (n - ns):(blk + 1) * 2 * (n - ns)])
        ref, rc = bind.lstsq_fp64(A, b)
        if rc < 1e-6:   # quad16 on flat blocks: degree-2 terms nearly collinear
            continue
        used += 1
        fit, fit_r = A @ W[blk], A @ ref
        worst = max(worst, np.abs(fit - fit_r).max() / (np.abs(fit_r).max() + 1e-3))
    assert used >= s.blocks // 2
    assert worst < 2e-4, worst


@pytest.mark.parametrize("name", ["cross5", "quad16"])
def test_configured_fitter_exact_fit_recovers_generating_weights(name):
    w, h = 70, 48
    pairs, ns = SETS[name]
    n = len(pairs)
    p, o = _oracle(w, h, False, SETS[name])
    _run_stage12(o, _scene(w, h), 3)
    s = o.s
    rng = np.random.default_rng(11)
    wtrue = rng.uniform(-1, 1, (s.blocks, n, 3))
    tmp = o.tmp.reshape(s.blocks, n + 3, 1024)
    conds = []
    for blk in range(s.blocks):
        A, _, _ = bind.fitter_inputs(p, o.tmp, blk, 2)
        tmp[blk, n:n + 3, :] = (A @ wtrue[blk]).T.astype(np.float32)
        conds.append(np.linalg.cond(A))
    o.fitter(2)
    W = o.weights.reshape(s.blocks, n, 3)
    checked = 0
    for blk in range(s.blocks):
        if conds[blk] > 1e4:
            continue
        checked += 1
        err = np.abs(W[blk] - wtrue[blk]).max()
        assert err < 1e-6 * conds[blk] * 30 + 1e-4, (blk, err, conds[blk])
    assert checked >= s.blocks // 2


@pytest.mark.parametrize("name", ["one_scaled", "cross5", "quad16"])
def test_configured_weighted_sum_closed_form(name):
    """filtered = max(0, sum_f w_f * scale(feature_f)) in float32, features
    in order, each term one FMA, scaled ones with their block's bounds
    (bmfr.cl:727-750)."""
    w, h = 70, 48
    pairs, ns = SETS[name]
    n = len(pairs)
    sc = _scene(w, h)
    f = sc.render_host(0)
    p, o = _oracle(w, h, False, SETS[name])
    o.upload(f["albedo"], f["normals"], f["positions"], f["noisy"])
    s = o.s
    rng = np.random.default_rng(12)
    o.weights[:] = rng.uniform(-1, 1, o.weights.size).astype(np.float32)
    mm = rng.uniform(-3, 3, (s.blocks, max(n - ns, 1), 2)).astype(np.float32)
    mm[:, :, 1] = mm[:, :, 0] + rng.choice([0.5, 4.0], mm[:, :, 1].shape).astype(np.float32)
    o.mins_maxs[:] = mm.ravel()[:o.mins_maxs.size]
    frame = 3
    o.weighted_sum(frame)
    ox, oy = _offset(frame)
    y, x = np.mgrid[0:h, 0:w]
    g = (x + 16 - ox) // 32 + ((y + 16 - oy) // 32) * s.blocks_x
    nrm = f["normals"].reshape(h, w, 3)
    pos = f["positions"].reshape(h, w, 3)
    fa = _factor_arrays(nrm, pos)
    Wt = o.weights.reshape(s.blocks, n, 3)
    col = np.zeros((h, w, 3), np.float32)
    for i, (a, b) in enumerate(pairs):
        v = (fa[a] * fa[b]).astype(np.float32)
        if i >= ns:
            lo, hi = mm[g, i - ns, 0], mm[g, i - ns, 1]
            v = np.where(np.abs(hi - lo) > 1, (v - lo) / (hi - lo), v - lo).astype(np.float32)
        for c in range(3):
            col[..., c] = fma32(Wt[g, i, c], v, col[..., c])   # bmfr.cl:745, contracted
    want = np.where(col < 0, np.float32(0), col)
    assert np.array_equal(o.filtered.reshape(h, w, 3), want)


def test_oracle_and_library_params_share_the_feature_fields():
    gp = _params(96, 64)
    set_factor_pairs(gp, *SETS["quad16"])
    op = bind.params(96, 64)
    copy_features(op, gp)
    assert bind.feature_count(op) == 16 and bind.scaled_count(op) == 12
    assert lib.orc_feature_count(ctypes.byref(op)) == 16
    assert lib.orc_not_scaled(ctypes.byref(op)) == 4
