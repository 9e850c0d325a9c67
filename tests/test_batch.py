"""Batched execution (one launch per stage over many streams): each stream's
results must be those of running it alone, bit for bit -- the batched
kernels run the same per-pixel and per-block code on the same inputs.
Covers mixed resolutions, mixed storage precision (several launch groups),
edge blocks, hipGraph replay interleaved with eager work, and errors."""
import numpy as np
import pytest

import bmfr_amd
from tests.harness import gpu_params, scene

pytestmark = pytest.mark.gpu

MIX = [  # (w, h, storage_f16, tmp_f16, seed offset, qr mode)
    (96, 64, False, True, 0, bmfr_amd.QR_FP32),
    (130, 70, False, True, 1, bmfr_amd.QR_FP32),     # ragged: W, H not multiples of 32
    (64, 48, True, True, 2, bmfr_amd.QR_FP32),       # FP16 storage: a second launch group
    (200, 120, False, False, 3, bmfr_amd.QR_FP32),   # FP32 tmp_data: a third group
    (96, 64, False, False, 4, bmfr_amd.QR_KAHAN),    # compensated QR
    (130, 70, True, False, 5, bmfr_amd.QR_KAHAN_FP64),   # + FP64 re-factorisation
    (1000, 47, False, True, 6, bmfr_amd.QR_FP32),    # one block row of the padded grid
    (47, 600, True, True, 7, bmfr_amd.QR_FP32),      # one block column
]
# A threshold at which many blocks are re-factorised in FP64 (the default,
# 1e-6, flags almost none of these scenes' blocks).
REFACTOR_THRESHOLD = 0.5


def _streams(mix):
    return [bmfr_amd.Stream(gpu_params(w, h, tmp_f16=t16, storage_f16=g16, qr_mode=qr,
                                       refactor_threshold=REFACTOR_THRESHOLD))
            for (w, h, g16, t16, _, qr) in mix]


def _scenes(mix):
    return [scene(w, h, seed=0xB3F40000 + 3 + k) for (w, h, _, _, k, _) in mix]


def _buffers(st, fuse):
    names = ["result_prev", "out_prev", "noisy_acc_prev", "spp_prev", "weights", "mins_maxs",
             "prev_pixel", "accept"]
    if fuse == bmfr_amd.FUSE_34:  # the 3+4+5 kernel keeps tone_mapped on chip
        names.append("tone_mapped")
    d = {n: st.read_buffer(n) for n in names}
    d["refactored"] = np.array([st.refactor_count()])
    return d


def _run(mix, frames, batched, graph=False, eager_between=False, fuse=bmfr_amd.FUSE_345, lanes=2):
    sts, scs = _streams(mix), _scenes(mix)
    for st in sts:
        st.set_fused(fuse)
    b = bmfr_amd.Batch(sts) if batched else None
    if b:
        b.set_fused(fuse)
        b.set_lanes(lanes)
    if graph:
        b.set_graph(True)
    try:
        for fr in range(frames):
            for st, sc in zip(sts, scs):
                sc.render_to_stream(fr, st)
            args = [sc.frame_args(fr) for sc in scs]
            if batched:
                b.run([fr] * len(sts), np.concatenate([a[0] for a in args]),
                      np.concatenate([a[1] for a in args]))
                if eager_between:
                    # eager work and synchronisation between graph replays
                    b.synchronize()
                    for st in sts:
                        st.read_buffer("weights")
            else:
                for st, (vp, off) in zip(sts, args):
                    st.run_frame(fr, vp, off)
        out = []
        for st in sts:
            st.synchronize()
            d = _buffers(st, fuse)
            d["download"] = st.download_result()
            out.append(d)
        return out
    finally:
        if b:
            b.close()
        for st in sts:
            st.close()


def _same(a, b):
    return np.array_equal(a, b) or np.array_equal(np.isnan(a), np.isnan(b)) and np.array_equal(
        np.nan_to_num(a), np.nan_to_num(b))


@pytest.mark.parametrize("fuse", [bmfr_amd.FUSE_34, bmfr_amd.FUSE_345])
def test_batch_equals_each_stream_alone(gpu, fuse):
    alone = _run(MIX, 5, batched=False, fuse=fuse)
    mixed = _run(MIX, 5, batched=True, fuse=fuse)
    for k, (a, m) in enumerate(zip(alone, mixed)):
        for name in a:
            assert _same(a[name], m[name]), (k, name)
    # the compensated streams: only the FP64 one re-factorises, and it does
    assert [int(a["refactored"][0]) > 0 for a in alone] == [m[5] == bmfr_amd.QR_KAHAN_FP64 for m in MIX]


def test_batch_graph_replay_equals_eager(gpu):
    eager = _run(MIX[:3] + MIX[4:], 6, batched=True)
    graph = _run(MIX[:3] + MIX[4:], 6, batched=True, graph=True, eager_between=True)
    for k, (a, m) in enumerate(zip(eager, graph)):
        for name in a:
            assert _same(a[name], m[name]), (k, name)


@pytest.mark.parametrize("lanes", [1, 3, 4])
def test_batch_pipeline_lanes_equal_each_stream_alone(gpu, lanes):
    """The software pipeline's lane count changes only the launch order:
    graph-replayed steps with 1, 3 or 4 lanes give each stream's alone bits."""
    alone = _run(MIX, 4, batched=False)
    mixed = _run(MIX, 4, batched=True, graph=True, eager_between=True, lanes=lanes)
    for k, (a, m) in enumerate(zip(alone, mixed)):
        for name in a:
            assert _same(a[name], m[name]), (lanes, k, name)


def test_batch_lanes_change_between_steps(gpu):
    """Regrouping the members (set_lanes between steps, after graph replays)
    keeps every stream's bits."""
    sts, scs = _streams(MIX), _scenes(MIX)
    b = bmfr_amd.Batch(sts)
    try:
        b.set_graph(True)
        for fr, lanes in enumerate([2, 4, 4, 1, 3, 2]):
            b.set_lanes(lanes)
            for st, sc in zip(sts, scs):
                sc.render_to_stream(fr, st)
            args = [sc.frame_args(fr) for sc in scs]
            b.run([fr] * len(sts), np.concatenate([a[0] for a in args]), np.concatenate([a[1] for a in args]))
        b.synchronize()
        got = [_buffers(st, bmfr_amd.FUSE_345) for st in sts]
        with pytest.raises(bmfr_amd.BmfrError):
            b.set_lanes(5)
    finally:
        b.close()
        for st in sts:
            st.close()
    alone = _run(MIX, 6, batched=False)
    for k, (a, m) in enumerate(zip(alone, got)):
        for name in m:
            assert _same(a[name], m[name]), (k, name)


@pytest.mark.parametrize("cfg", [(4, 1, 0), (8, 2, 0), (16, 1, 64), (16, 4, 8)])
def test_batch_fitter_configs_and_persistent_grid(gpu, cfg):
    """Every fitter shape, with a capped (persistent, several blocks per
    work-group) grid, gives the stream-alone weights of the same shape."""
    mix = MIX[:2]
    sts, scs = _streams(mix), _scenes(mix)
    alone = _streams(mix)
    for st in alone:
        st.set_fitter_config(cfg[0], cfg[1])
    b = bmfr_amd.Batch(sts)
    b.set_fitter_config(*cfg)
    try:
        for fr in range(3):
            args = [sc.frame_args(fr) for sc in scs]
            for st, a_st, sc in zip(sts, alone, scs):
                sc.render_to_stream(fr, st)
                sc.render_to_stream(fr, a_st)
            b.run([fr] * 2, np.concatenate([a[0] for a in args]), np.concatenate([a[1] for a in args]))
            for a_st, (vp, off) in zip(alone, args):
                a_st.run_frame(fr, vp, off)
        for st, a_st in zip(sts, alone):
            assert np.array_equal(st.read_buffer("weights"), a_st.read_buffer("weights"))
            assert np.array_equal(st.download_result(), a_st.download_result())
    finally:
        b.close()
        for st in sts + alone:
            st.close()


def test_batch_timing_and_errors(gpu):
    sts = _streams(MIX[:2])
    try:
        with pytest.raises(bmfr_amd.BmfrError):
            bmfr_amd.Batch([sts[0], sts[0]])
        b = bmfr_amd.Batch(sts)
        with pytest.raises(bmfr_amd.BmfrError):
            b.set_fitter_config(3, 1)
        with pytest.raises(bmfr_amd.BmfrError):
            b.stage_times()
        scs = _scenes(MIX[:2])
        args = [sc.frame_args(0) for sc in scs]
        for st, sc in zip(sts, scs):
            sc.render_to_stream(0, st)
        b.run([0, 0], np.concatenate([a[0] for a in args]), np.concatenate([a[1] for a in args]),
              timed=True)
        t = b.stage_times()
        assert (t > 0).all() and t[5] >= t[:5].max()
        with pytest.raises(ValueError):
            b.run([0], np.zeros(32), np.zeros(4))
        b.close()
    finally:
        for st in sts:
            st.close()
