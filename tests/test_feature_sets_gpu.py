"""Configured feature sets on the GPU, through the C ABI (the configured-set
kernels: stage 1 and the weighted sums evaluate the set's factor pairs, the
fitter runs N + 3 columns).

  explicit reference set   the configured kernels with the reference's own
                           set give the bits of the specialised kernels
                           (feature_count 0), every buffer, every fuse level
  stage 1                  bit-exact against the oracle (tmp_data with N + 3
                           columns, accumulated colour, reprojection, spp)
  fitter                   mins_maxs bit-exact; fitted colour within the
                           stated tolerance of the oracle and no worse than
                           twice the oracle's own distance to FP64
  end to end               final frame per-pixel relative error
  batches and servers      refuse configured sets (BMFR_ERR_INVALID_ARGUMENT)
The reference's own kernels with the same sets are compared in
tests/test_reference_run.py."""
import numpy as np
import pytest

import bmfr_amd
from oracle import bind
from tests.harness import REFERENCE_SET, Lockstep, gpu_params, rel_err, scene, set_factor_pairs
from tests.test_feature_sets import SETS
from tests.test_gpu_parity import _stage1_frames

pytestmark = pytest.mark.gpu

BUFS = ["noisy_acc_cur", "prev_pixel", "accept", "spp_cur", "tmp_data", "mins_maxs", "weights",
        "out_cur", "result_cur"]


@pytest.mark.parametrize("tmp_f16", [True, False])
@pytest.mark.parametrize("level", [bmfr_amd.FUSE_NONE, bmfr_amd.FUSE_34, bmfr_amd.FUSE_345])
def test_explicit_reference_set_gives_the_specialised_bits(gpu, tmp_f16, level):
    w, h = 96, 64
    sc = scene(w, h)
    p0 = gpu_params(w, h, tmp_f16)
    p1 = gpu_params(w, h, tmp_f16)
    set_factor_pairs(p1, *REFERENCE_SET)
    g0, g1 = bmfr_amd.Stream(p0), bmfr_amd.Stream(p1)
    try:
        for g in (g0, g1):
            g.set_fused(level)
        for fr in range(4):
            f = sc.render_host(fr)
            vp, off = sc.frame_args(fr)
            for g in (g0, g1):
                g.upload_frame(f["albedo"], f["normals"], f["positions"], f["noisy"])
                if level == bmfr_amd.FUSE_NONE:
                    g.accumulate_noisy_data(fr, vp, off)
                    g.fitter(fr)
                    g.weighted_sum(fr)
                    g.accumulate_filtered_data(fr)
                    g.taa(fr)
                else:
                    g.run_frame(fr, vp, off)
            names = BUFS + (["filtered", "tone_mapped"] if level == bmfr_amd.FUSE_NONE else [])
            for name in names:
                if level != bmfr_amd.FUSE_NONE:
                    name = {"out_cur": "out_prev", "result_cur": "result_prev", "noisy_acc_cur": "noisy_acc_prev",
                            "spp_cur": "spp_prev"}.get(name, name)   # run_frame swapped the double buffers
                a, b = g0.read_buffer(name), g1.read_buffer(name)
                same = (a == b) | (np.isnan(a) & np.isnan(b)) if a.dtype.kind == "f" else a == b
                assert same.all(), (fr, name, int((~same).sum()))
            if level == bmfr_amd.FUSE_NONE:
                for g in (g0, g1):
                    g.end_frame()
    finally:
        g0.close()
        g1.close()


@pytest.mark.parametrize("name", ["one", "cross5", "quad16"])
@pytest.mark.parametrize("tmp_f16", [True, False])
def test_configured_stage1_bit_exact(gpu, name, tmp_f16):
    ls = Lockstep(96, 64, tmp_f16=tmp_f16, pairs=SETS[name])
    try:
        _stage1_frames(ls, 4)
    finally:
        ls.close()


def _fitted(p, tmp, frame, blocks, W, rc_min=0.0):
    """Per block: (fitted colour A W, A) for blocks whose FP64 reciprocal
    condition estimate is at least rc_min."""
    out = {}
    for blk in range(blocks):
        A, b, _ = bind.fitter_inputs(p, tmp, blk, frame)
        w64, rc = bind.lstsq_fp64(A, b)
        if rc >= rc_min:
            out[blk] = (A, w64)
    return out


@pytest.mark.parametrize("name", ["one", "one_scaled", "cross5", "lin7", "quad16"])
@pytest.mark.parametrize("tmp_f16", [True, False])
def test_configured_fitter_matches_oracle_and_fp64(gpu, name, tmp_f16):
    w, h = 256, 256
    ls = Lockstep(w, h, tmp_f16=tmp_f16, pairs=SETS[name])
    n = ls.features
    try:
        _stage1_frames(ls, 3, check=False)
        fr = 2
        tmp0 = ls.o.tmp.copy()
        assert np.array_equal(ls.read("tmp_data"), tmp0)
        ls.g.fitter(fr)
        ls.o.fitter(fr)
        s = ls.o.s
        assert np.array_equal(ls.read("mins_maxs")[:ls.o.mins_maxs.size], ls.o.mins_maxs)
        Wg = ls.read("weights").reshape(s.blocks, n, 3).astype(np.float64)
        Wo = ls.o.weights.reshape(s.blocks, n, 3).astype(np.float64)
        # a column that is constant over a block scales to zeros and has no
        # Householder vector (the reference divides by zero there too)
        assert np.array_equal(np.isfinite(Wg), np.isfinite(Wo))
        # blocks with a well-defined least-squares solution (rc >= 1e-6);
        # the degree-2 set is nearly collinear on flat blocks
        blocks = _fitted(ls.o.p, tmp0, fr, s.blocks, None, 1e-6)
        assert len(blocks) >= s.blocks // 2
        fe = fe64 = fo64 = 0.0
        for blk, (A, w64) in blocks.items():
            fg, fo, fd = A @ Wg[blk], A @ Wo[blk], A @ w64
            sc_ = np.abs(fo).max() + 1e-3
            fe = max(fe, float(np.abs(fg - fo).max() / sc_))
            fe64 = max(fe64, float(np.abs(fg - fd).max() / sc_))
            fo64 = max(fo64, float(np.abs(fo - fd).max() / sc_))
        print(f"{name} tmp_f16={tmp_f16}: {len(blocks)}/{s.blocks} blocks, vs oracle {fe:.2e}, "
              f"vs FP64 {fe64:.2e} (oracle {fo64:.2e})")
        # the same bars as the reference's set (tests/test_gpu_parity.py)
        assert fe < (3e-2 if tmp_f16 else 1e-4), fe
        assert fe64 < max(2.0 * fo64, 1e-4), (fe64, fo64)
    finally:
        ls.close()


@pytest.mark.parametrize("name", ["one", "cross5", "lin7", "quad16"])
@pytest.mark.parametrize("tmp_f16", [True, False])
def test_configured_end_to_end(gpu, name, tmp_f16):
    w, h, frames = 96, 64, 6
    ls = Lockstep(w, h, tmp_f16=tmp_f16, pairs=SETS[name])
    try:
        for fr in range(frames):
            ls.upload(fr)
            vp, off = ls.sc.frame_args(fr)
            ls.g.run_frame(fr, vp, off)
            res_o = ls.o.run_frame(fr, vp, off)
        res_g = ls.g.download_result()
        fin = np.isfinite(res_o) & np.isfinite(res_g)
        assert np.array_equal(np.isfinite(res_o), np.isfinite(res_g))
        err = rel_err(res_g[fin], res_o[fin])
        print(f"{name} tmp_f16={tmp_f16}: median {np.median(err):.2e}, p99.9 {np.percentile(err, 99.9):.2e}, "
              f"max {err.max():.2e}")
        if tmp_f16:
            # binary16 updates: a few pixels move when one rounding goes the
            # other way (the reference's set has the same bar)
            assert np.percentile(err, 99.9) < 1e-2 and np.median(err) < 1e-4
        else:
            assert err.max() < 1e-3, float(err.max())
    finally:
        ls.close()


def test_batches_and_servers_refuse_configured_sets(gpu):
    p = gpu_params(96, 64)
    set_factor_pairs(p, *SETS["cross5"])
    s = bmfr_amd.Stream(p)
    try:
        with pytest.raises(bmfr_amd.BmfrError) as e:
            bmfr_amd.Batch([s])
        assert e.value.status == bmfr_amd.ERR_INVALID_ARGUMENT
        sv = bmfr_amd.Server(1 << 30)
        try:
            with pytest.raises(bmfr_amd.BmfrError) as e:
                sv.open_stream(p)
            assert e.value.status == bmfr_amd.ERR_INVALID_ARGUMENT
        finally:
            sv.close()
        with pytest.raises(bmfr_amd.BmfrError):
            s.fitter_fp64_weights()
    finally:
        s.close()


@pytest.mark.parametrize("tmp_f16", [False, True])
def test_configured_set_full_size_1080p(gpu, tmp_f16):
    """16 features at 1920x1080: stage 1 bit-exact over two frames, the
    fitter's mins_maxs bit-exact, and on a sample of well-conditioned blocks
    the fitted colour no further from FP64 least squares than twice the
    oracle's own distance."""
    w, h = 1920, 1080
    ls = Lockstep(w, h, tmp_f16=tmp_f16, sc=scene(w, h, spheres=48), pairs=SETS["quad16"])
    n = ls.features
    try:
        _stage1_frames(ls, 2)
        fr = 1
        tmp0 = ls.o.tmp.copy()
        ls.g.fitter(fr)
        s = ls.o.s
        rng = np.random.default_rng(9)
        sample = np.unique(np.concatenate([[0, s.blocks - 1], rng.choice(s.blocks, 62, replace=False)]))
        for b in sample:
            ls.o.fitter(fr, int(b), 1)
        ns = n - ls.o.p.features_not_scaled
        mm_g = ls.read("mins_maxs").reshape(s.blocks, 2 * ns)
        mm_o = ls.o.mins_maxs.reshape(s.blocks, 2 * ns)
        assert np.array_equal(mm_g[sample], mm_o[sample])
        Wg = ls.read("weights").reshape(s.blocks, n, 3).astype(np.float64)
        Wo = ls.o.weights.reshape(s.blocks, n, 3).astype(np.float64)
        used = 0
        for b in sample:
            A, bb, _ = bind.fitter_inputs(ls.o.p, tmp0, int(b), fr)
            w64, rc = bind.lstsq_fp64(A, bb)
            if rc < 1e-6:
                continue
            used += 1
            sc_ = np.abs(A @ w64).max() + 1e-3
            eg = np.abs(A @ Wg[b] - A @ w64).max() / sc_
            eo = np.abs(A @ Wo[b] - A @ w64).max() / sc_
            assert eg <= 2.0 * eo + (1e-6 if not tmp_f16 else 1e-4), (b, eg, eo)
        assert used >= len(sample) // 2
    finally:
        ls.close()
