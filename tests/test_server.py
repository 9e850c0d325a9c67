"""The server: admission by HBM budget, stream lifecycle and dynamic batched
steps (continuous batching).  Each stream's results must equal running it
alone, bit for bit, whatever other streams share its steps; failure paths
are forced from the host."""
import numpy as np
import pytest

import bmfr_amd
from tests.harness import gpu_params, scene

pytestmark = pytest.mark.gpu

SPECS = [(96, 64, False), (130, 70, True), (64, 48, False)]  # (w, h, storage_f16)


def test_admission_by_budget(gpu):
    p = gpu_params(96, 64)
    need = bmfr_amd.compute_sizes(p).device_bytes
    sv = bmfr_amd.Server(2 * need + need // 2)
    try:
        a = sv.open_stream(p)
        b = sv.open_stream(p)
        with pytest.raises(bmfr_amd.BmfrError) as e:
            sv.open_stream(p)
        assert e.value.status == bmfr_amd.ERR_OUT_OF_BUDGET
        used, budget, n = sv.usage()
        assert used == 2 * need and n == 2
        sv.close_stream(a)
        c = sv.open_stream(p)            # memory freed: admitted now
        assert sv.usage()[2] == 2
        with pytest.raises(bmfr_amd.BmfrError):
            sv.close_stream(a)           # already closed
        c.close()
        b.close()
        assert sv.usage()[0] == 0
    finally:
        sv.close()


def test_step_rejects_foreign_and_duplicate_streams(gpu):
    sv = bmfr_amd.Server(1 << 30)
    other = bmfr_amd.Stream(gpu_params(64, 48))
    try:
        s = sv.open_stream(gpu_params(64, 48))
        with pytest.raises(bmfr_amd.BmfrError):
            sv.step([other], [0], np.zeros(16), np.zeros(2))
        with pytest.raises(bmfr_amd.BmfrError):
            sv.step([s, s], [0, 0], np.zeros(32), np.zeros(4))
        with pytest.raises(bmfr_amd.BmfrError):
            sv.step([s], [-1], np.zeros(16), np.zeros(2))
    finally:
        other.close()
        sv.close()


def test_each_result_same_alone_and_in_the_mix(gpu):
    """Stream 0 is in every step, stream 1 in every other step, stream 2
    joins at step 3 -- continuous batching with changing step membership."""
    sv = bmfr_amd.Server(1 << 31)
    scs = [scene(w, h, seed=0xB3F40000 + 10 + k) for k, (w, h, _) in enumerate(SPECS)]
    served = [None, None, None]
    alone = [bmfr_amd.Stream(gpu_params(w, h, storage_f16=g16)) for (w, h, g16) in SPECS]
    next_frame = [0, 0, 0]
    try:
        served[0] = sv.open_stream(gpu_params(*SPECS[0][:2], storage_f16=SPECS[0][2]))
        served[1] = sv.open_stream(gpu_params(*SPECS[1][:2], storage_f16=SPECS[1][2]))
        pending_ids = []
        for step in range(8):
            if step == 3:
                served[2] = sv.open_stream(gpu_params(*SPECS[2][:2], storage_f16=SPECS[2][2]))
            members = [k for k in range(3) if served[k] is not None and (k != 1 or step % 2 == 0)]
            mats, offs, frames = [], [], []
            for k in members:
                fr = next_frame[k]
                scs[k].render_to_stream(fr, served[k])
                vp, off = scs[k].frame_args(fr)
                mats.append(vp)
                offs.append(off)
                frames.append(fr)
            pending_ids.append(sv.step([served[k] for k in members], frames, np.concatenate(mats),
                                       np.concatenate(offs)))
            for k in members:
                next_frame[k] += 1
        sv.synchronize()
        done, ms = sv.query(pending_ids[-1])
        assert done and ms > 0
        for k in range(3):
            for fr in range(next_frame[k]):
                scs[k].render_to_stream(fr, alone[k])
                alone[k].run_frame(fr, *scs[k].frame_args(fr))
            a = alone[k].download_result()
            m = served[k].download_result()
            assert np.array_equal(a, m), k
            assert np.array_equal(alone[k].read_buffer("weights"), served[k].read_buffer("weights")), k
    finally:
        for st in alone:
            st.close()
        sv.close()
