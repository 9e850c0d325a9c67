"""Pins of the CPU oracle (oracle/bmfr_oracle.c), CPU only.

Each check is chosen so that a plausible mistake in the restatement (a
dropped term, a wrong sign or index, a transposed operand) fails it: closed
forms, invariants, special cases that reduce to a library routine
(numpy.float16 rounding, numpy.linalg.lstsq) and brute force on tiny
inputs.  The reference ships no tests or golden vectors (it has none), so
these, plus the fixtures of the reference kernels' own runs under
tests/golden/, are what pin the oracle.
"""
import ctypes

import numpy as np
import pytest

from tests.harness import fma32

from oracle import bind

lib = bind.load()


# ---- binary16 conversion (vstore_half / vload_half, bmfr.cl:255-258) --------
def test_f16_conversion_matches_numpy_rne():
    rng = np.random.default_rng(1)
    vals = np.concatenate([
        rng.standard_normal(20000).astype(np.float32) * 10.0 ** rng.integers(-9, 6, 20000),
        np.array([0.0, -0.0, 65504.0, 65519.996, 65520.0, -65520.0, 6.1035156e-05,
                  5.96e-08, 2.98e-08, 2.9802322e-08, 1e-10, np.inf, -np.inf,
                  1.0009765625, 1.00048828125, 1.00146484375], np.float32),
        # every exact half plus/minus half an ulp: the tie cases
        (np.arange(0, 0x7c00, dtype=np.uint16).view(np.float16).astype(np.float32)),
    ]).astype(np.float32)
    mids = []
    h = np.arange(1, 0x7bff, dtype=np.uint16).view(np.float16).astype(np.float64)
    h2 = np.arange(2, 0x7c00, dtype=np.uint16).view(np.float16).astype(np.float64)
    mids = ((h + h2) / 2).astype(np.float32)
    vals = np.concatenate([vals, mids, -mids])
    got = np.array([lib.orc_f32_to_f16(float(v)) for v in vals], np.uint16)
    with np.errstate(over="ignore"):
        want = vals.astype(np.float16).view(np.uint16)
    assert np.array_equal(got, want)
    back = np.array([lib.orc_f16_to_f32(int(b)) for b in want[::7]], np.float32)
    assert np.array_equal(back, want[::7].view(np.float16).astype(np.float32))


# ---- hash and feature noise (bmfr.cl:161-182) -------------------------------
def _hash_np(a):
    # Robert Jenkins' six-shift 32-bit integer hash, written from its
    # published description with explicit uint32 wrap-around.
    a = a.astype(np.uint64)
    m = np.uint64(0xFFFFFFFF)
    a = ((a + np.uint64(0x7ED55D16)) + (a << np.uint64(12))) & m
    a = ((a ^ np.uint64(0xC761C23C)) ^ (a >> np.uint64(19))) & m
    a = ((a + np.uint64(0x165667B1)) + (a << np.uint64(5))) & m
    a = ((a + np.uint64(0xD3A2646C)) ^ (a << np.uint64(9))) & m
    a = ((a + np.uint64(0xFD7046C5)) + (a << np.uint64(3))) & m
    a = ((a ^ np.uint64(0xB55A4F09)) ^ (a >> np.uint64(16))) & m
    return a.astype(np.uint32)


def test_hash_known_answers():
    xs = np.array([0, 1, 2, 3, 255, 1023, 13312, 0x7FFFFFFF, 0xFFFFFFFF, 123456789],
                  np.uint32)
    want = _hash_np(xs)
    got = np.array([lib.orc_hash(int(x)) for x in xs], np.uint32)
    assert np.array_equal(got, want)
    # bijectivity on a window: a wrong shift direction usually breaks it
    win = np.arange(0, 1 << 16, dtype=np.uint32)
    assert len(np.unique(_hash_np(win))) == len(win)


def test_random_is_hash_over_two_pow_32_and_noise_bounds():
    for a in [0, 7, 99999, 0xFFFFFFFF]:
        h = int(_hash_np(np.array([a], np.uint32))[0])
        assert lib.orc_random(a) == np.float32(np.float32(h) / np.float32(2.0 ** 32))
    # add_random: value + 0.02*(r-0.5) in double, rounded once to float
    for (row, f, frame) in [(0, 1, 0), (1023, 9, 5), (17, 4, 59)]:
        arg = row + 1024 * f + 13 * 1024 * frame
        r = np.float32(lib.orc_random(arg)) - np.float32(0.5)
        want = np.float32(np.float64(np.float32(3.25)) + 1e-2 * 2.0 * np.float64(r))
        got = lib.orc_add_random(3.25, row % 256, row // 256, f, frame, 1e-2)
        assert np.float32(got) == want
        assert abs(got - 3.25) <= 0.0100001


# ---- jitter, mirror, block indexing (bmfr.cl:207-222, 267-285, 310-317, 718-722)
def _offset(frame):
    o = (ctypes.c_int32 * 2)()
    lib.orc_block_offset(frame, o)
    return o[0], o[1]


def test_block_offsets_range_and_period():
    offs = [_offset(f) for f in range(16)]
    assert len(set(offs)) == 16
    assert all(-16 <= x <= 14 and -16 <= y <= 14 for x, y in offs)
    assert all(x % 2 == 0 and y % 2 == 0 for x, y in offs)
    assert [_offset(f + 16) for f in range(16)] == offs


def test_mirror_closed_form():
    for n in [33, 48, 100]:
        for i in range(-n + 1, 2 * n):
            want = i if 0 <= i < n else (-i - 1 if i < 0 else 2 * n - i - 1)
            assert lib.orc_mirror(i, n) == want
            assert 0 <= lib.orc_mirror(i, n) < n


@pytest.mark.parametrize("w,h", [(47, 50), (64, 64), (70, 48)])
def test_padded_grid_covers_each_pixel_once_and_stage3_maps_back(w, h):
    """Every image pixel is written by exactly one non-mirrored stage-1
    work-item, and stage 3's block index of that pixel is that work-item's
    block (bijection invariant of the jittered block grid)."""
    s = bind.sizes(w, h)
    for frame in range(16):
        ox, oy = _offset(frame)
        gy, gx = np.mgrid[0:s.margin_h, 0:s.margin_w]
        px, py = gx - 16 + ox, gy - 16 + oy
        own = (px >= 0) & (px < w) & (py >= 0) & (py < h)
        counts = np.zeros((h, w), int)
        np.add.at(counts, (py[own], px[own]), 1)
        assert (counts == 1).all()
        blk_of_item = (gy // 32) * s.blocks_x + gx // 32
        y, x = np.mgrid[0:h, 0:w]
        g = (x + 16 - ox) // 32 + ((y + 16 - oy) // 32) * s.blocks_x
        owner_blk = np.zeros((h, w), int)
        owner_blk[py[own], px[own]] = blk_of_item[own]
        assert np.array_equal(g, owner_blk)


# ---- reductions (bmfr.cl:26-87) --------------------------------------------
def test_tree_reductions_exact_on_integers():
    rng = np.random.default_rng(2)
    for _ in range(20):
        v = rng.integers(-1000, 1000, 256).astype(np.float32)
        a, b, c = v.copy(), v.copy(), v.copy()
        assert lib.orc_tree_sum(a.ctypes.data) == float(v.astype(np.int64).sum())
        assert lib.orc_tree_min(b.ctypes.data) == v.min()
        assert lib.orc_tree_max(c.ctypes.data) == v.max()


def test_tree_sum_association():
    # 2^24 + 1 + ... is sensitive to order: pins the 4-way then 8-way shape
    v = np.zeros(256, np.float32)
    v[0] = 2.0 ** 24
    v[64] = 1.0
    v[128] = 1.0
    got = lib.orc_tree_sum(v.ctypes.data)
    # s[0] = s[0] + ((s[64] + s[128]) + s[192]) = 2^24 + 2 (exact)
    assert got == 2.0 ** 24 + 2
    v2 = np.zeros(256, np.float32)
    v2[0] = 2.0 ** 24
    v2[1] = 1.0
    v2[2] = 1.0
    # level 3 is ((s0 + s1) + s2): 2^24 + 1 rounds to 2^24 twice
    assert lib.orc_tree_sum(v2.ctypes.data) == 2.0 ** 24


# ---- scale, YCoCg (bmfr.cl:184-205) ------------------------------------------
def test_scale_branches():
    assert lib.orc_scale(5.0, 1.0, 9.0) == pytest.approx(0.5)
    assert lib.orc_scale(1.5, 1.0, 1.9) == pytest.approx(0.5)     # |max-min| <= 1: shift only
    assert lib.orc_scale(3.0, 1.0, 2.0) == pytest.approx(2.0)     # exactly 1: shift only
    assert lib.orc_scale(-2.0, -4.0, 0.0) == pytest.approx(0.5)


def test_ycocg_round_trip_and_matrix():
    rng = np.random.default_rng(3)
    for _ in range(50):
        rgb = (rng.integers(0, 256, 3) / 256.0).astype(np.float32)
        y = np.zeros(3, np.float32)
        back = np.zeros(3, np.float32)
        lib.orc_rgb_to_ycocg(rgb.ctypes.data, y.ctypes.data)
        lib.orc_ycocg_to_rgb(y.ctypes.data, back.ctypes.data)
        assert np.array_equal(back, rgb)            # dyadic values: exact
        M = np.array([[1, 2, 1], [2, 0, -2], [-1, 2, -1]], np.float64)
        assert np.allclose(y, M @ rgb)


# ---- stage 1 (bmfr.cl:287-485) -------------------------------------------------
def _scene(w, h, static=False, seed=0xB3F40000, spheres=8, firefly=1e-3):
    import bmfr_amd
    return bmfr_amd.Scene(seed, spheres, firefly, w, h, static_camera=static)


def test_stage1_frame0_pass_through_and_layout():
    w, h = 70, 48
    sc = _scene(w, h)
    f = sc.render_host(0)
    p = bind.params(w, h, half_tmp=False)
    o = bind.OracleStream(p)
    o.upload(f["albedo"], f["normals"], f["positions"], f["noisy"])
    vp, off = sc.frame_args(0)
    o.accumulate_noisy_data(0, vp, off)
    c = o.cur()
    assert np.array_equal(o.noisy_acc[c], f["noisy"])
    assert (o.spp[c] == 1).all() and (o.accept == 0).all()
    y, x = np.mgrid[0:h, 0:w]
    pp = o.prev_pixel.reshape(h, w, 2)
    assert np.array_equal(pp[..., 0], x) and np.array_equal(pp[..., 1], y)
    # tmp_data: block-major [by][bx][feature][y][x] of the mirrored, jittered grid
    s = o.s
    t = o.tmp.reshape(s.blocks_y, s.blocks_x, 13, 32, 32)
    grid = t.transpose(2, 0, 3, 1, 4).reshape(13, s.margin_h, s.margin_w)
    ox, oy = _offset(0)
    gy, gx = np.mgrid[0:s.margin_h, 0:s.margin_w]
    mx = np.where(gx - 16 + ox < 0, -(gx - 16 + ox) - 1, gx - 16 + ox)
    mx = np.where(mx >= w, 2 * w - mx - 1, mx)
    my = np.where(gy - 16 + oy < 0, -(gy - 16 + oy) - 1, gy - 16 + oy)
    my = np.where(my >= h, 2 * h - my - 1, my)
    pos = f["positions"].reshape(h, w, 3)[my, mx]
    nrm = f["normals"].reshape(h, w, 3)[my, mx]
    col = f["noisy"].reshape(h, w, 3)[my, mx]
    want = np.stack([np.ones_like(pos[..., 0]), nrm[..., 0], nrm[..., 1], nrm[..., 2],
                     pos[..., 0], pos[..., 1], pos[..., 2], pos[..., 0] ** 2, pos[..., 1] ** 2,
                     pos[..., 2] ** 2, col[..., 0], col[..., 1], col[..., 2]])
    assert np.array_equal(grid, want.astype(np.float32))


def test_stage1_static_camera_closed_form():
    """Static camera and pixel offset: every pixel reprojects onto itself
    (to float rounding), hit pixels accumulate spp = min(f+1, 255) and
    alpha = max(1/(f+1), 0.2); background pixels never accept."""
    w, h = 64, 48
    sc = _scene(w, h, static=True)
    p = bind.params(w, h, half_tmp=False)
    o = bind.OracleStream(p)
    y, x = np.mgrid[0:h, 0:w]
    for frame in range(6):
        f = sc.render_host(frame)
        o.upload(f["albedo"], f["normals"], f["positions"], f["noisy"])
        vp, off = sc.frame_args(frame)
        o.accumulate_noisy_data(frame, vp, off)
        c = o.cur()
        hit = (np.abs(f["normals"].reshape(h, w, 3)).sum(-1) > 0)
        pp = o.prev_pixel.reshape(h, w, 2)
        if frame > 0:
            assert np.abs(pp[..., 0] - x)[hit].max() < 2e-3
            assert np.abs(pp[..., 1] - y)[hit].max() < 2e-3
            assert (o.accept.reshape(h, w)[~hit] == 0).all()
        spp = o.spp[c].reshape(h, w)
        assert (spp[hit] == min(frame + 1, 255)).all()
        assert (spp[~hit] == 1).all()
        if frame > 0:
            # accumulated colour is the running mean until alpha hits 0.2
            acc = o.noisy_acc[c].reshape(h, w, 3)[hit]
            assert np.isfinite(acc).all()
        o.end_frame()


# ---- stage 2: the fitter against FP64 least squares (bmfr.cl:487-700) --------
def _fitter_case(w=70, h=48, frame=3, half=False, seed=0xB3F40000):
    sc = _scene(w, h, seed=seed)
    p = bind.params(w, h, half_tmp=half)
    o = bind.OracleStream(p)
    for fr in range(frame + 1):
        f = sc.render_host(fr)
        o.upload(f["albedo"], f["normals"], f["positions"], f["noisy"])
        vp, off = sc.frame_args(fr)
        o.accumulate_noisy_data(fr, vp, off)
        if fr < frame:
            o.end_frame()
    return p, o


def test_fp64_lstsq_matches_numpy():
    rng = np.random.default_rng(4)
    A = rng.standard_normal((1024, 10))
    b = rng.standard_normal((1024, 3))
    w, rc = bind.lstsq_fp64(A, b)
    ref = np.linalg.lstsq(A, b, rcond=None)[0]
    assert np.allclose(w, ref, rtol=1e-10, atol=1e-12)
    assert 0 < rc <= 1


@pytest.mark.parametrize("half", [False, True])
def test_fitter_weights_match_fp64_least_squares(half):
    p, o = _fitter_case(half=half)
    tmp0 = o.tmp.copy()
    o.fitter(3)
    s = o.s
    W = o.weights.reshape(s.blocks, 10, 3)
    worst = 0.0
    for blk in range(s.blocks):
        A, b, mm = bind.fitter_inputs(p, tmp0, blk, 3)
        assert np.array_equal(mm, o.mins_maxs[blk * 12:(blk + 1) * 12])
        ref = np.linalg.lstsq(A, b, rcond=None)[0]
        # compare fitted colours (well conditioned) and weights
        fit_o = A @ W[blk].astype(np.float64)
        fit_r = A @ ref
        scale = np.abs(fit_r).max() + 1e-3
        err = np.abs(fit_o - fit_r).max() / scale
        worst = max(worst, err)
    # FP32 Householder with the reference's sign choice; half mode rounds
    # every update to binary16 (bmfr.cl:651).
    assert worst < (2e-2 if half else 2e-4), worst


def test_fitter_exact_fit_recovers_generating_weights():
    """Colour columns that are an exact linear combination of the fitter's
    (scaled, noised) features give back the generating weights."""
    p, o = _fitter_case(half=False)
    s = o.s
    rng = np.random.default_rng(5)
    wtrue = rng.uniform(-1, 1, (s.blocks, 10, 3))
    tmp = o.tmp.reshape(s.blocks, 13, 1024)
    conds = []
    for blk in range(s.blocks):
        A, _, _ = bind.fitter_inputs(p, o.tmp, blk, 3)
        tmp[blk, 10:13, :] = (A @ wtrue[blk]).T.astype(np.float32)
        conds.append(np.linalg.cond(A))
    o.fitter(3)
    W = o.weights.reshape(s.blocks, 10, 3)
    for blk in range(s.blocks):
        # FP32 QR: error grows with the block's condition number
        err = np.abs(W[blk] - wtrue[blk]).max()
        assert err < 1e-6 * conds[blk] * 30 + 1e-4, (blk, err, conds[blk])


def test_fitter_householder_sign_is_reference_choice():
    """R_00 = +||column 0|| = +32 (all-ones column), i.e. u_0 = 1 - 32: with
    a single block whose colour equals 2.5 everywhere, weight 0 is 2.5 and
    the rest vanish."""
    w, h = 47, 47
    p = bind.params(w, h, half_tmp=False)
    o = bind.OracleStream(p)
    s = o.s
    rng = np.random.default_rng(6)
    t = o.tmp.reshape(s.blocks, 13, 1024)
    t[:, 0] = 1.0
    t[:, 1:10] = rng.uniform(-1, 1, (s.blocks, 9, 1024))
    t[:, 10:13] = 2.5
    o.fitter(0)
    W = o.weights.reshape(s.blocks, 10, 3)
    # features 4..9 get scaled; weight on the constant term absorbs the mean
    A, b, _ = bind.fitter_inputs(p, o.tmp, 0, 0)
    fit = A @ W[0]
    assert np.allclose(fit, 2.5, atol=1e-4)


# ---- stages 3-5 (bmfr.cl:703-974) -----------------------------------------------
def _rn32_exact(q):
    """Round a Fraction to the nearest float32, ties to even (normal range)."""
    from fractions import Fraction
    import math
    if q == 0:
        return np.float32(0)
    sign = -1 if q < 0 else 1
    q = abs(q)
    e = math.floor(math.log2(q))
    while Fraction(2) ** e > q:
        e -= 1
    while Fraction(2) ** (e + 1) <= q:
        e += 1
    ulp = Fraction(2) ** (e - 23)
    t = q / ulp
    n = t.numerator // t.denominator
    rem = t - n
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and n % 2 == 1):
        n += 1
    return np.float32(sign * float(n * ulp))


def test_fma32_is_the_single_rounding():
    """The numpy fused multiply-add the tests use against exact rational
    arithmetic: random operands over several exponent ranges, and operands
    built so that a*b + c lies a hair off a float32 rounding midpoint (where
    rounding twice, through float64, goes the wrong way)."""
    from fractions import Fraction
    rng = np.random.default_rng(5)
    a = rng.uniform(-2, 2, 4000).astype(np.float32)
    b = (rng.uniform(-2, 2, 4000) * 10.0 ** rng.integers(-6, 6, 4000)).astype(np.float32)
    c = (rng.uniform(-2, 2, 4000) * 10.0 ** rng.integers(-6, 6, 4000)).astype(np.float32)
    # midpoint cases: c in [1, 2) (float32 spacing 2^-23), a*b = +-(2^-24 - 2^-70),
    # so a*b + c lies 2^-70 inside a float32 midpoint and rounds to c, while
    # float64 rounds it onto the midpoint
    k = rng.integers(1, 4000, 200).astype(np.float64)
    c_mid = (1.0 + k * 2.0 ** -23).astype(np.float32)
    a_mid = np.full(200, np.float32(1.0 + 2.0 ** -23))
    h = np.float32(2.0 ** -24 - 2.0 ** -47)
    b_mid = np.where(rng.random(200) < 0.5, h, -h).astype(np.float32)
    a = np.concatenate([a, a_mid]).astype(np.float32)
    b = np.concatenate([b, b_mid]).astype(np.float32)
    c = np.concatenate([c, c_mid]).astype(np.float32)
    got = fma32(a, b, c)
    for i in range(len(a)):
        q = Fraction(float(a[i])) * Fraction(float(b[i])) + Fraction(float(c[i]))
        want = _rn32_exact(q)
        assert got[i] == want, (i, a[i], b[i], c[i], got[i], want)


def test_weighted_sum_unit_weights_pick_block_and_feature():
    w, h = 70, 48
    sc = _scene(w, h)
    f = sc.render_host(0)
    p = bind.params(w, h)
    o = bind.OracleStream(p)
    o.upload(f["albedo"], f["normals"], f["positions"], f["noisy"])
    s = o.s
    frame = 5
    ox, oy = _offset(frame)
    # weight of feature 0 (the constant 1) = block index; others 0
    W = np.zeros((s.blocks, 10, 3), np.float32)
    W[:, 0, 0] = np.arange(s.blocks)
    # feature 2 (normal.y) with weight 1 in the green channel
    W[:, 2, 1] = 1.0
    o.weights[:] = W.ravel()
    o.mins_maxs[:] = 0.0
    o.weighted_sum(frame)
    out = o.filtered.reshape(h, w, 3)
    y, x = np.mgrid[0:h, 0:w]
    g = (x + 16 - ox) // 32 + ((y + 16 - oy) // 32) * s.blocks_x
    assert np.array_equal(out[..., 0], g.astype(np.float32))
    ny = f["normals"].reshape(h, w, 3)[..., 1]
    assert np.array_equal(out[..., 1], np.where(ny < 0, 0, ny))
    assert (out[..., 2] == 0).all()


def test_accumulate_filtered_frame0_and_static_history():
    w, h = 48, 48
    P = w * h
    p = bind.params(w, h)
    o = bind.OracleStream(p)
    rng = np.random.default_rng(7)
    o.filtered[:] = rng.uniform(0, 2, 3 * P)
    o.albedo[:] = rng.uniform(0.1, 0.9, 3 * P)
    o.accumulate_filtered_data(0)
    c = o.cur()
    assert np.array_equal(o.out[c], o.filtered)
    want = np.clip(np.power(np.maximum(0, o.albedo * o.filtered), np.float32(0.454545)), 0, 1)
    assert np.allclose(o.tone, want, rtol=2e-7, atol=0)
    # static history: tap 0 only, prev_pixel = pixel, spp = s
    y, x = np.mgrid[0:h, 0:w]
    o.prev_pixel[:] = np.stack([x, y], -1).ravel()
    o.accept[:] = 1
    for s in [1, 3, 20]:
        o.spp[c][:] = s
        hist = rng.uniform(0, 2, 3 * P).astype(np.float32)
        o.out[o.prev()][:] = hist
        o.accumulate_filtered_data(1)
        a = np.float32(max(1.0 / s, 0.1))
        want = a * o.filtered + (np.float32(1) - a) * hist
        assert np.allclose(o.out[c], want, rtol=1e-6)


def test_taa_pass_through_constant_image_and_edges():
    w, h = 48, 48
    P = w * h
    p = bind.params(w, h)
    o = bind.OracleStream(p)
    y, x = np.mgrid[0:h, 0:w]
    o.prev_pixel[:] = np.stack([x, y], -1).ravel().astype(np.float32)
    rng = np.random.default_rng(8)
    o.tone[:] = rng.uniform(0, 1, 3 * P)
    o.taa(0)
    assert np.array_equal(o.result[o.cur()], o.tone)
    # constant image and constant history: identity
    o.tone[:] = np.tile(np.array([0.25, 0.5, 0.75], np.float32), P)
    o.result[o.prev()][:] = o.tone
    o.taa(1)
    assert np.allclose(o.result[o.cur()], o.tone, atol=1e-7)
    # prev pixel out of range -> pass-through
    o.prev_pixel[:] = -5.0
    o.result[o.prev()][:] = 0.0
    o.taa(1)
    assert np.array_equal(o.result[o.cur()], o.tone)
    # history clamped into the neighbourhood box: black history, mid-grey image
    o.prev_pixel[:] = np.stack([x, y], -1).ravel().astype(np.float32) - 0.5
    o.tone[:] = 0.5
    o.result[o.prev()][:] = 0.0
    o.taa(1)
    assert np.allclose(o.result[o.cur()], 0.5, atol=1e-6)


def test_minimum_image_size_is_the_mirror_validity_bound():
    """47 is the smallest size whose padded grid (up to WORKSET+29 after the
    largest jitter) mirrors back inside the image; the oracle refuses smaller
    images, as the C ABI does."""
    def valid(n):
        ws = 32 * ((n + 31) // 32)
        lo, hi = 0 - 16 - 16, ws + 31 - 16 + 14
        return all(0 <= lib.orc_mirror(i, n) < n for i in (lo, hi))
    assert [n for n in range(33, 200) if not valid(n)] == list(range(33, 47))
    o = bind.OracleStream(bind.params(64, 46))
    with pytest.raises(ValueError):
        o.accumulate_noisy_data(0, np.zeros(16), np.zeros(2))
