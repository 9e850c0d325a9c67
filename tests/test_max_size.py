"""The largest frame the API accepts, 16384x16384: 268M pixels, so a
three-float image buffer holds 3 GiB (the kernels address image buffers
with 32-bit byte offsets from the buffer's base) and the block matrix
(263,169 blocks of 13 x 1024 binary16 elements) 7 GiB.  Every kernel's
addressing at the far end of its buffers, through the size-independent
properties the chain has there (the CPU oracle cannot run whole frames of
this size in a test's time):
  - static camera: every hit pixel reprojects onto itself (to within 0.02
    pixel, a few ulp of coordinates near 16384), counts frames in spp; the
    background never accepts (the closed form of opencl/bmfr.cl:336-442);
  - frame 0: accumulate_filtered_data and TAA pass their inputs through
    (bmfr.cl:778-790, 884-890), and the result is in [0, 1];
  - the fused 3+4+5 kernel gives the separate stages' bits on both frames;
  - the fitter's weights on sampled blocks (the first, the last and the
    last block row included) against FP64 least squares on the block
    matrices stage 1 wrote, no worse than the CPU oracle's FP32 Householder
    on the same block (within a factor of two)."""
import ctypes

import numpy as np
import pytest

import bmfr_amd
from oracle import bind
from tests.harness import gpu_params, scene

pytestmark = pytest.mark.gpu

W = H = 16384


def _hit_mask(g):
    n = g.read_buffer("normals_cur").reshape(H, W, 3)
    return np.any(n != 0, axis=2)


def test_max_size_frame(gpu):
    sc = scene(W, H, spheres=24, static=True)
    sep = bmfr_amd.Stream(gpu_params(W, H))
    fus = bmfr_amd.Stream(gpu_params(W, H))
    try:
        sep.set_fused(0)
        fus.set_fused(bmfr_amd.FUSE_345)
        xs = np.arange(W, dtype=np.float32)[None, :]
        ys = np.arange(H, dtype=np.float32)[:, None]
        for fr in range(2):
            print(f"max size: frame {fr}", flush=True)
            vp, off = sc.frame_args(fr)
            sc.render_to_stream(fr, sep)
            sc.render_to_stream(fr, fus)
            sep.accumulate_noisy_data(fr, vp, off)
            hit = _hit_mask(sep)
            assert hit.mean() > 0.1, float(hit.mean())
            spp = sep.read_buffer("spp_cur").reshape(H, W)
            assert (spp[hit] == fr + 1).all()
            assert (spp[~hit] == 1).all()
            del spp
            if fr > 0:
                pp = sep.read_buffer("prev_pixel").reshape(H, W, 2)
                ex = np.abs(pp[..., 0] - xs)[hit]
                ey = np.abs(pp[..., 1] - ys)[hit]
                del pp
                print(f"max size: reprojection error x {ex.max():.2e}, y {ey.max():.2e}", flush=True)
                assert ex.max() < 0.02 and ey.max() < 0.02
                del ex, ey
                acc = sep.read_buffer("accept").reshape(H, W)
                assert (acc[~hit] == 0).all()
                del acc
            sep.fitter(fr)
            if fr > 0:
                _fitter_sample(sep, fr)
            sep.weighted_sum(fr)
            sep.accumulate_filtered_data(fr)
            sep.taa(fr)
            fus.run_frame(fr, vp, off)
            sep.synchronize()
            fus.synchronize()
            if fr == 0:
                assert np.array_equal(sep.read_buffer("out_cur"), sep.read_buffer("filtered"))
                assert np.array_equal(sep.read_buffer("result_cur"), sep.read_buffer("tone_mapped"))
            r = sep.read_buffer("result_cur")
            # frame 0's result is the tone-mapped colour, in [0, 1]; later
            # frames blend with history clamped in YCoCg, whose RGB may leave
            # that range (bmfr.cl:963-973)
            assert np.isfinite(r).all() and (fr > 0 or (r.min() >= 0.0 and r.max() <= 1.0))
            assert np.array_equal(r, fus.download_result())
            del r
            # run_frame has ended the fused stream's frame: its accumulated
            # colour is now the previous frame's
            assert np.array_equal(sep.read_buffer("out_cur"), fus.read_buffer("out_prev"))
            sep.end_frame()
    finally:
        sep.close()
        fus.close()
        sc.close()


def _fitter_sample(g, fr):
    """Sampled blocks' weights against FP64 least squares, next to the
    oracle's FP32 Householder on the same block matrix."""
    p = bind.params(W, H, half_tmp=True)
    s = bind.sizes(W, H)
    nb, bx = s.blocks, s.blocks_x
    tmp = g.read_buffer("tmp_data")            # binary16, [block][13][1024]
    wg = g.read_buffer("weights").reshape(nb, 10, 3).astype(np.float64)
    mm_g = g.read_buffer("mins_maxs").reshape(nb, 12)
    assert tmp.size == 13 * 1024 * nb and np.isfinite(wg).all()
    rng = np.random.default_rng(11)
    sample = np.unique(np.concatenate([[0, 1, bx - 1, nb - bx, nb - 2, nb - 1],
                                       rng.choice(nb, 24, replace=False)]))
    lib = bind.load()
    worse, singular = [], 0
    for blk in sample:
        m = np.ascontiguousarray(tmp[blk * 13 * 1024:(blk + 1) * 13 * 1024], np.float32)
        A, b, mm = bind.fitter_inputs(p, m, 0, fr)
        assert np.array_equal(mm, mm_g[blk]), blk
        wo = np.zeros(30, np.float32)
        mo = np.zeros(12, np.float32)
        mq = m.copy()   # the oracle's QR works in place; the copy must outlive the call
        lib.orc_fitter(ctypes.byref(p), bind._ptr(mq), bind._ptr(mo), bind._ptr(wo), fr, 0, 1)
        w64, rc = bind.lstsq_fp64(A, b)
        if rc < 1e-10:
            singular += 1
            continue
        scale = np.abs(A @ w64).max() + 1e-3
        eg = np.abs(A @ wg[blk] - A @ w64).max() / scale
        eo = np.abs(A @ wo.reshape(10, 3).astype(np.float64) - A @ w64).max() / scale
        worse.append(eg / max(eo, 1e-7))
        assert eg <= 2.0 * eo + 1e-4, (int(blk), eg, eo)
    assert singular <= len(sample) // 4
    print(f"max size: {len(worse)} blocks ({singular} singular skipped), fitted-colour error vs FP64 "
          f"relative to the oracle's: median {np.median(worse):.2f}, max {np.max(worse):.2f}", flush=True)


def test_max_size_batched(gpu):
    """The batched kernels (one launch over the concatenated tiles and
    blocks of the batch's members) at the largest frame, beside a small
    member: every member's results bit-identical to the stream run alone."""
    sizes = [(W, H), (96, 64)]
    scs = [scene(w, h, spheres=24, static=(w == W)) for (w, h) in sizes]
    alone = [bmfr_amd.Stream(gpu_params(w, h)) for (w, h) in sizes]
    member = [bmfr_amd.Stream(gpu_params(w, h)) for (w, h) in sizes]
    b = bmfr_amd.Batch(member)
    try:
        for fr in range(2):
            print(f"max size batched: frame {fr}", flush=True)
            args = [sc.frame_args(fr) for sc in scs]
            for k, sc in enumerate(scs):
                sc.render_to_stream(fr, alone[k])
                sc.render_to_stream(fr, member[k])
                alone[k].run_frame(fr, *args[k])
            b.run([fr] * len(sizes), np.concatenate([a[0] for a in args]), np.concatenate([a[1] for a in args]))
            b.synchronize()
            for k in range(len(sizes)):
                assert np.array_equal(alone[k].download_result(), member[k].download_result()), (fr, k)
                assert np.array_equal(alone[k].read_buffer("weights"), member[k].read_buffer("weights")), (fr, k)
    finally:
        b.close()
        for s in alone + member:
            s.close()
        for sc in scs:
            sc.close()
