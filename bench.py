#!/usr/bin/env python3
"""Benchmark of the gfx950 BMFR frame chain on one MI355X.

Workload (the headline metric's configuration): plain single-stream BMFR at
1280x720, the seeded procedural plane + 24 spheres scene with a moving
camera, FP32 G-buffers and FP16 tmp_data (the reference defaults).  Per
frame the renderer stand-in writes the frame's G-buffers straight into the
stream's device buffers (not timed, like the reference's uploads,
opencl/bmfr.cpp:415-416); the five stages are then timed with HIP events on
the stream's own HIP stream.  ms/frame is the reference's "Total time in all
kernels": first stage's start to last stage's end (opencl/bmfr.cpp:497-502),
averaged over the timed frames.

Prints one JSON line.  `--gpus N` with N > 1 exits: every GPU box here has a
single GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# 256 CUs x 4 SIMDs at 2.4 GHz, one wave64 VALU instruction per SIMD every 4 cycles
VALU_ISSUE_PEAK = 1024 * 2.4e9 / 4 / 1e9
# SQ_INSTS_VALU / SQ_WAVES of the fitter (k_fitter<16, 4, __half, true, false>),
# profiles/fitter_720p_counters.md
FITTER_VALU_PER_WAVE = 5275


def reference_fitter_lds_bytes():
    """LDS bytes one block of the reference's fitter moves (bmfr.cl:510-656,
    13 buffers, 1024 rows): the 4 KB u_vec store of each of the 13 columns,
    a 4-byte u_vec read in the dot product and another in the update for
    every row at or below the pivot of every trailing column (the colour
    columns pivot at row 10 and each has two trailing columns,
    bmfr.cl:552, 606), and the min and max reductions' 2 x 4 KB of each of
    the 6 scaled features."""
    total = 13 * 4096 + 6 * 2 * 4096
    for col in range(13):
        cl = min(col, 10)
        m = 1024 - cl
        total += 8 * m * (13 - (cl + 1))
    return total
W, H = 1280, 720
SPHERES = 24
SEED = 0xB3F40000 + 1


def stage_bytes(P, Q, nB, s=4, t=2):
    """Algorithmic HBM bytes per launch of each stage (unique bytes: every
    input texel read once, every output texel written once)."""
    return {
        "accumulate_noisy_data": P * (21 * s + 11) + 13 * t * Q,
        "fitter": 13 * t * Q + 168 * nB,
        "weighted_sum": P * (6 * s + 12) + 168 * nB,
        "accumulate_filtered_data": P * (46 + 6 * s),
        "taa": P * (20 + 6 * s),
    }


STAGES = ["accumulate_noisy_data", "fitter", "weighted_sum", "accumulate_filtered_data", "taa"]


def chain_kernels(fuse, P, Q, nB, s=4):
    """The kernels of the chained frame at a fuse level: name -> (index of
    the stage whose events bracket it, algorithmic bytes per launch).  A fused
    kernel keeps the intermediates it folds in on chip: `filtered` (12 B/px
    written and read), and at level 2 `tone_mapped` (the same) and the
    second read of prev_pixel (8 B/px)."""
    sb = stage_bytes(P, Q, nB, s)
    k = {"accumulate_noisy_data": (0, sb["accumulate_noisy_data"]), "fitter": (1, sb["fitter"])}
    if fuse == 0:
        k.update({n: (i, sb[n]) for i, n in enumerate(STAGES) if i >= 2})
    elif fuse == 1:
        k["weighted_sum+accumulate_filtered_data"] = (
            3, sb["weighted_sum"] + sb["accumulate_filtered_data"] - 24 * P)
        k["taa"] = (4, sb["taa"])
    else:
        k["filter_taa"] = (3, sb["weighted_sum"] + sb["accumulate_filtered_data"] + sb["taa"] - 56 * P)
    return k


def run_plain(args):
    """Timed frames record events only at the first stage's start and the
    last stage's end (nothing between kernels); a second pass with events
    around every stage gives the per-stage breakdown."""
    import bmfr_amd
    p = bmfr_amd.default_params(W, H)
    st = bmfr_amd.Stream(p)
    st.set_fused(args.fuse)
    if args.fitter_config:
        st.set_fitter_config(*[int(x) for x in args.fitter_config.split(",")])
    sc = bmfr_amd.Scene(SEED, SPHERES, 1e-3, W, H)
    n_frames = args.warmup + 2 * args.steps
    cams = [sc.frame_args(f) for f in range(n_frames)]
    totals, per_stage = [], []
    wall0, wall = None, None
    for f in range(n_frames):
        if f == args.warmup:
            st.synchronize()
            st.set_timing(2)
            wall0 = time.perf_counter()
        if f == args.warmup + args.steps:
            st.synchronize()
            wall = time.perf_counter() - wall0
            st.set_timing(1)
        sc.render_to_stream(f, st)
        vp, off = cams[f]
        st.run_frame(f, vp, off)
        if args.warmup <= f < args.warmup + args.steps:
            totals.append(float(st.stage_times()[5]))
        elif f >= args.warmup + args.steps:
            per_stage.append(st.stage_times())
    # The fitter's time inside the chain with no event between it and stage 1:
    # frames alternate a window from stage 1's start to its end (A) and one
    # from stage 1's start to the fitter's end (B), one event pair each, so
    # that the events' own cost cancels in mean(B) - mean(A), which keeps the
    # chain's stage 1 -> fitter dispatch gap.  The rest of such a frame runs
    # as separate stages, outside both windows.
    st.synchronize()
    st.set_timing(0)
    win = ([], [])
    for k in range(2 * args.steps):
        f = n_frames + k
        sc.render_to_stream(f, st)
        vp, off = sc.frame_args(f)
        st.record_marker(0)
        st.accumulate_noisy_data(f, vp, off)
        if k % 2 == 0:
            st.record_marker(1)
        st.fitter(f)
        if k % 2 == 1:
            st.record_marker(1)
        st.weighted_sum(f)
        st.accumulate_filtered_data(f)
        st.taa(f)
        st.end_frame()
        win[k % 2].append(st.marker_elapsed(0, 1))
    fitter_chain_ms = float(np.mean(win[1]) - np.mean(win[0]))
    n_frames += 2 * args.steps
    # the fitter's kernel time without dispatch gaps: the last frame's fitter
    # launched back to back, events only around the whole run
    st.synchronize()
    reps = 50
    st.record_marker(0)
    for _ in range(reps):
        st.fitter(n_frames - 1)
    st.record_marker(1)
    fitter_b2b_ms = st.marker_elapsed(0, 1) / reps
    s = st.sizes
    st.close()
    sc.close()
    return np.array(totals), np.array(per_stage, np.float64), wall, s, fitter_b2b_ms, fitter_chain_ms


BATCH_MIX = [(640, 360)] * 4 + [(1280, 720)] * 6 + [(1920, 1080)] * 4 + [(2560, 1440)] * 2


def run_batched(steps, warmup, lanes=2):
    """16 concurrent streams of mixed resolutions, one frame each per step:
    every stage is one launch over all of them (the fitter one launch over
    all 22,988 blocks, its shape chosen by the autotuner), replayed from a
    captured hipGraph.  The
    renderer stand-in writes each stream's G-buffers before a step (not
    timed); the step is timed with HIP events on the batch's HIP stream."""
    import bmfr_amd
    sts = [bmfr_amd.Stream(bmfr_amd.default_params(w, h)) for (w, h) in BATCH_MIX]
    scs = [bmfr_amd.Scene(0xB3F40000 + 3 + k, SPHERES, 1e-3, w, h) for k, (w, h) in enumerate(BATCH_MIX)]
    b = bmfr_amd.Batch(sts)
    b.set_lanes(lanes)
    n = len(sts)
    pixels = sum(w * h for (w, h) in BATCH_MIX)
    blocks = sum(st.sizes.blocks for st in sts)

    def step(f, timed, graph):
        for st, sc in zip(sts, scs):
            sc.render_to_stream(f, st)
        args = [sc.frame_args(f) for sc in scs]
        b.set_graph(graph)
        b.run([f] * n, np.concatenate([a[0] for a in args]), np.concatenate([a[1] for a in args]),
              timed=timed)

    step(0, False, False)
    b.synchronize()
    tune = b.autotune(3)
    stage = None
    totals = []
    for f in range(1, warmup + steps + 1):
        timed = f > warmup
        if f == warmup:  # one eager, per-stage timed step for the breakdown
            step(f, True, False)
            stage = b.stage_times()
            continue
        step(f, timed, True)
        if timed:
            totals.append(float(b.stage_times()[5]))
    b.synchronize()
    b.close()
    for st in sts:
        st.close()
    t = np.array(totals)
    step_ms = float(t.mean())
    p720 = 1280 * 720
    return {
        "streams": n, "mix": "4x640x360 6x1280x720 4x1920x1080 2x2560x1440",
        "pixels_per_step": pixels, "blocks_per_step": blocks, "steps": len(t),
        "step_ms": round(step_ms, 4), "step_ms_p99": round(float(np.percentile(t, 99)), 4),
        "ms_per_720p_frame_equiv": round(step_ms * p720 / pixels, 5),
        "aggregate_frames_per_s": round(n / (step_ms * 1e-3), 1),
        "mpixel_per_s": round(pixels / (step_ms * 1e-3) / 1e6, 1),
        "stage_ms": {k: round(float(stage[i]), 4) for k, (i, _) in chain_kernels(2, 1, 1, 1).items()},
        "graph": True, "pipeline_lanes": lanes, "autotune": tune,
    }


def run_precision(frames):
    """1920x1080, 48 spheres, fireflies 5e-3: FP16 normals/positions/noisy
    colour/histories, an FP32 working matrix (tmp_data) with compensated QR
    and FP64 re-factorisation of blocks with min |R_ii| < 1e-6 max |R_ii|;
    beside it the same scene with the plain-path settings (FP32 storage,
    FP16 tmp_data, plain FP32 QR, the reference's default).  Reports
    ms/frame, the weight error of both against FP64 least squares on the
    same block matrices, and the final frame's error against the plain
    path."""
    import bmfr_amd
    w, h = 1920, 1080
    sc = bmfr_amd.Scene(0xB3F40000 + 2, 48, 5e-3, w, h)
    a = bmfr_amd.Stream(bmfr_amd.default_params(w, h, storage=bmfr_amd.STORAGE_F16,
                                                tmp_storage=bmfr_amd.STORAGE_F32,
                                                qr_mode=bmfr_amd.QR_KAHAN_FP64))
    b = bmfr_amd.Stream(bmfr_amd.default_params(w, h))
    a.set_timing(2)
    t = []
    for f in range(frames):
        sc.render_to_stream(f, a)
        sc.render_to_stream(f, b)
        vp, off = sc.frame_args(f)
        a.run_frame(f, vp, off)
        b.run_frame(f, vp, off)
        if f >= 3:
            t.append(float(a.stage_times()[5]))

    def werr(st):
        w32 = st.read_buffer("weights").reshape(-1, 10, 3).astype(np.float64)
        w64 = st.fitter_fp64_weights().reshape(-1, 10, 3).astype(np.float64)
        e = np.abs(w32 - w64).max(axis=(1, 2)) / (np.abs(w64).max(axis=(1, 2)) + 1e-6)
        e = e[np.isfinite(e)]
        return {"max": float(e.max()), "median": float(np.median(e)), "p99": float(np.percentile(e, 99))}

    ea, eb = werr(a), werr(b)
    refactored = a.refactor_count()
    ra, rb = a.download_result(), b.download_result()
    fin = np.isfinite(ra) & np.isfinite(rb)
    mse = float(np.mean((np.clip(ra[fin], 0, 1) - np.clip(rb[fin], 0, 1)) ** 2))
    rel = np.abs(ra[fin] - rb[fin]) / np.maximum(np.abs(rb[fin]), 1e-2)
    a.close()
    b.close()
    sc.close()
    return {"workload": "1920x1080 (2135 blocks), 48 spheres, fireflies 5e-3, FP16 storage, "
                        "FP32 tmp_data, QR_KAHAN_FP64", "frames": frames,
            "ms_per_frame": round(float(np.mean(t)), 4),
            "weight_err_vs_fp64": {k: round(v, 8) for k, v in ea.items()},
            "weight_err_vs_fp64_plain_path": {k: round(v, 8) for k, v in eb.items()},
            "blocks_refactored_last_frame": refactored,
            "final_frame_vs_plain_path": {"psnr_db": round(10 * np.log10(1.0 / max(mse, 1e-30)), 2),
                                            "median_rel": float(np.median(rel)),
                                            "p99_rel": float(np.percentile(rel, 99))}}


# Configured feature sets (the reference's FEATURE_BUFFERS strings,
# opencl/bmfr.cpp:65-77) timed on the plain 720p chain.
FEATURE_SETS = [
    ("linear 7", "1.f,normal.x,normal.y,normal.z,", "world_position.x,world_position.y,world_position.z"),
    ("quadratic 16", "1.f,normal.x,normal.y,normal.z,",
     "world_position.x,world_position.y,world_position.z,world_position.x*world_position.x,"
     "world_position.y*world_position.y,world_position.z*world_position.z,world_position.x*world_position.y,"
     "world_position.y*world_position.z,world_position.x*world_position.z,normal.x*world_position.x,"
     "normal.y*world_position.y,normal.z*world_position.z"),
]


def run_feature_sets(frames, fuse):
    """The plain 720p chain with each configured set: ms/frame (first start
    to last end) and the fitter's share (per-stage events)."""
    import bmfr_amd
    out = []
    sc = bmfr_amd.Scene(SEED, SPHERES, 1e-3, W, H)
    for name, ns, scl in FEATURE_SETS:
        p = bmfr_amd.default_params(W, H)
        bmfr_amd.parse_feature_buffers(p, ns, scl)
        st = bmfr_amd.Stream(p)
        st.set_fused(fuse)
        tot, fit = [], []
        for f in range(5 + 2 * frames):
            if f == 5:
                st.synchronize()
                st.set_timing(2)
            if f == 5 + frames:
                st.synchronize()
                st.set_timing(1)
            sc.render_to_stream(f, st)
            vp, off = sc.frame_args(f)
            st.run_frame(f, vp, off)
            if 5 <= f < 5 + frames:
                tot.append(float(st.stage_times()[5]))
            elif f >= 5 + frames:
                fit.append(float(st.stage_times()[1]))
        st.close()
        out.append({"set": name, "features": p.feature_count, "not_scaled": p.features_not_scaled,
                    "ms_per_frame": round(float(np.mean(tot)), 4), "fitter_ms": round(float(np.mean(fit)), 4)})
        progress(f"feature set {name}: {out[-1]['ms_per_frame']} ms/frame, fitter {out[-1]['fitter_ms']} ms")
    sc.close()
    return out


SERVE_LOAD = os.path.join(ROOT, "bmfr_amd", "bmfr_serve_load")


def run_serve_load(*opts, timeout=600):
    """One run of bmfr_serve_load, the C++ client of the library's request
    loop (bmfr_server_submit / bmfr_server_poll): its JSON line as a dict."""
    import subprocess
    r = subprocess.run([SERVE_LOAD] + [str(o) for o in opts], capture_output=True, text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"bmfr_serve_load {' '.join(map(str, opts))} failed ({r.returncode}): {r.stderr[-2000:]}")
    return json.loads(r.stdout.strip().splitlines()[-1])


def run_served(duration, budget_gib):
    """64 render streams opened and closed over `duration` seconds: per stream
    640x360 / 1280x720 / 1920x1080 / 2560x1440 with probabilities 0.2 / 0.5 /
    0.25 / 0.05, Poisson frame arrivals at 30 or 60 Hz, FP16 or FP32
    storage; admitted by the HBM budget (admission FIFO); every frame handed
    over as device G-buffers with an RGBA8 output surface."""
    return run_serve_load("--mode", "served", "--streams", 64, "--seconds", duration, "--budget-gib", budget_gib)


def run_serve_sweep(counts, duration):
    """p50/p99 frame latency against the number of concurrent 1280x720 FP32
    streams at 60 Hz (latency of the frames arriving after the first second),
    up to and past the GPU's saturation: a stream holds at most 3 frames in
    the denoiser and drops one that arrives beyond that."""
    out = []
    for n in counts:
        budget = max(4.0, 0.2 * n + 4.0)   # 191 MB a stream
        r = run_serve_load("--mode", "sweep", "--streams", n, "--seconds", duration, "--settle", 1.0,
                           "--budget-gib", budget)
        out.append({k: r[k] for k in ("streams", "frames_arrived", "frames", "dropped", "aggregate_frames_per_s",
                                      "mpixel_per_s", "latency_ms_p50", "latency_ms_p99", "latency_ms_max",
                                      "mean_streams_per_step", "step_gpu_ms_mean", "step_time_per_s")})
        progress(f"serve sweep: {n} streams, p99 {r['latency_ms_p99']} ms, {r['mean_streams_per_step']} "
                 f"streams/step, {r['dropped']} of {r['frames_arrived']} frames dropped, step time per s {r['step_time_per_s']}")
    return out


def measured_traffic(kernel):
    """HBM bytes per dispatch of `kernel` from the calibrated rocprofv3 PMC
    runs summarised in profiles/ (null when that summary is absent)."""
    path = os.path.join(ROOT, "profiles", "traffic_plain_720p.json")
    try:
        d = json.load(open(path))["bytes_per_dispatch"][kernel]
        return int(d["read"] + d["write"])
    except (OSError, KeyError, ValueError):
        return None


def copy_peak(reps=20):
    """Measured HBM copy bandwidth of the library's 16-byte copy kernel
    (four 11 MB buffers per launch), the practical streaming ceiling."""
    import bmfr_amd
    p = bmfr_amd.default_params(W, H)
    a, b = bmfr_amd.Stream(p), bmfr_amd.Stream(p)
    src = [a.buffer_device_ptr(n) for n in ("albedo", "normals_cur", "positions_cur", "noisy_in")]
    nbytes = sum(a.buffer_bytes(n) for n in ("albedo", "normals_cur", "positions_cur", "noisy_in"))
    for _ in range(3):
        b.copy_frame_inputs_device(*src)
    b.record_marker(0)
    for _ in range(reps):
        b.copy_frame_inputs_device(*src)
    b.record_marker(1)
    ms = b.marker_elapsed(0, 1)
    a.close()
    b.close()
    return 2.0 * nbytes * reps / (ms * 1e-3) / 1e9


def cpu_baseline(frames=8):
    """The CPU oracle (a scalar C restatement, OpenMP over rows/blocks) on the
    host cores: 720p frames of the same scene; ms/frame over frames 1..N-1."""
    import bmfr_amd
    from oracle import bind
    sc = bmfr_amd.Scene(SEED, SPHERES, 1e-3, W, H)
    o = bind.OracleStream(bind.params(W, H, half_tmp=True))
    times = []
    for f in range(frames):
        g = sc.render_host(f)
        o.upload(g["albedo"], g["normals"], g["positions"], g["noisy"])
        vp, off = sc.frame_args(f)
        t0 = time.perf_counter()
        o.run_frame(f, vp, off)
        times.append(time.perf_counter() - t0)
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": 1000.0 * float(np.mean(times[1:])), "unit": "ms/frame", "cores": cores,
            "kind": "port",
            "sample": f"CPU oracle (oracle/bmfr_oracle.c, OpenMP) on {frames} frames of the same "
                      f"1280x720 scene; mean of frames 1..{frames - 1}"}


def reference_kernels_same_gpu(frames=10):
    """The reference's own opencl/bmfr.cl, compiled unmodified for gfx950
    (oracle/_ref, built where the reference tree is present), run by an
    OpenCL host with the reference's buffers and bindings on this GPU through
    the ROCm OpenCL runtime: the same 1280x720 frames (24 spheres, FP16
    tmp_data), the sum of its five kernels' event times averaged over frames
    1..N-1 (opencl/bmfr.cpp:488-517).  A baseline beside the CPU one, never
    the measured path."""
    from oracle import ref
    if not ref.available():
        return {"ms_per_frame": None, "note": "oracle/_ref not built (the reference tree was absent at build)"}
    try:
        if ref.load().ref_device_count() < 1:
            return {"ms_per_frame": None, "note": "no OpenCL GPU device visible"}
        import bmfr_amd
        sc = bmfr_amd.Scene(SEED, SPHERES, 1e-3, W, H)
        r = ref.RefRun(W, H, True)
        per = []
        for f in range(frames):
            g = sc.render_host(f)
            r.upload(g["albedo"], g["normals"], g["positions"], g["noisy"])
            vp, off = sc.frame_args(f)
            for i in range(5):
                r.stage(i, f, vp, off)
            if f > 0:
                per.append([r.stage_ms(i) for i in range(5)])
            r.end_frame()
        r.close()
        sc.close()
    except (OSError, RuntimeError) as e:
        return {"ms_per_frame": None, "note": f"reference run unavailable: {e}"}
    per = np.array(per).mean(axis=0)
    return {"ms_per_frame": round(float(per.sum()), 4), "unit": "ms/frame",
            "stage_ms": [round(float(x), 4) for x in per], "kind": "reference",
            "sample": f"opencl/bmfr.cl on this MI355X (ROCm OpenCL), {frames} frames 1280x720, "
                      "sum of its five kernels' event times, frames 1..N-1"}


def progress(msg):
    """One line per finished phase on stderr (the JSON line stays alone on stdout)."""
    print(f"bench.py: {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--fitter-config", default="")
    ap.add_argument("--fuse", type=int, default=2, help="stages 3-5: 0 separate, 1 3+4 fused, 2 3+4+5 fused")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batched-steps", type=int, default=40)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batch-lanes", type=int, default=2)
    ap.add_argument("--serve-seconds", type=float, default=30.0)
    ap.add_argument("--serve-budget-gib", type=float, default=4.0)
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--serve-sweep", default="1,16,64,128,192,256,512,1024",
                    help="stream counts of the latency-vs-streams sweep (empty: skip)")
    ap.add_argument("--serve-sweep-seconds", type=float, default=5.0)
    ap.add_argument("--precision-frames", type=int, default=60)
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--no-feature-sets", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU; --gpus {args.gpus} is not supported (the path shards by "
              "whole render streams: each GPU's process would run its own streams, with nothing to "
              "exchange)", file=sys.stderr)
        sys.exit(2)

    totals, t, wall, s, fitter_b2b_ms, fitter_chain_ms = run_plain(args)
    ms_frame = float(totals.mean())
    progress(f"plain 1280x720: {ms_frame:.4f} ms/frame over {len(totals)} frames")
    P, Q, nB = s.pixels, s.padded_pixels, s.blocks
    kern = chain_kernels(args.fuse, P, Q, nB)
    stage_ms = {k: float(t[:, i].mean()) for k, (i, _) in kern.items()}
    sb = {k: b for k, (_, b) in kern.items()}
    stage_gbs = {k: sb[k] / (stage_ms[k] * 1e-3) / 1e9 for k in sb}
    # the metric's kernel (and the longest one in the kernel trace); every
    # stage's fraction of the HBM peak is reported beside it
    dom = "fitter"
    # its duration in the chain: the difference of the two event windows
    # (run_plain), free of the per-stage events' own cost
    fit_ms = fitter_chain_ms
    dom_gbs = sb[dom] / (fit_ms * 1e-3) / 1e9
    roof = {"kernel": dom, "bound": "hbm", "achieved": round(dom_gbs, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(dom_gbs / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(dom), "algorithmic_bytes": int(sb[dom]),
            "kernel_ms": round(fit_ms, 5),
            "kernel_ms_method": "in the chain: events from stage 1's start to the fitter's end minus "
                                "events from stage 1's start to its end, alternate frames",
            "stage_frac": {k: round(v / HBM_PEAK_GBS, 4) for k, v in stage_gbs.items()},
            # What actually binds the fitter (one wave per 32x32 block, ~1 wave
            # per SIMD at 720p): its dependent VALU instruction chain, not HBM.
            "binding_resource": "valu issue / latency (see issue)",
            # the same bytes over the kernel's time launched back to back (no
            # dispatch gap), the figure comparable to a kernel trace's duration
            "kernel_ms_back_to_back": round(fitter_b2b_ms, 5),
            "frac_back_to_back": round(sb[dom] / (fitter_b2b_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    # The fitter's own limit is instruction issue: its code is straight-line,
    # FITTER_VALU_PER_WAVE instructions per wave (one wave per block, the
    # launch padded to whole four-block work-groups per XCD), against one
    # wave64 VALU instruction per 4 cycles per SIMD.
    waves = 8 * (((nB + 7) // 8 + 3) // 4) * 4
    issue = waves * FITTER_VALU_PER_WAVE / (fit_ms * 1e-3) / 1e9
    roof["issue"] = {"bound": "valu", "achieved": round(issue, 1), "peak": VALU_ISSUE_PEAK,
                     "unit": "G wave-instructions/s", "frac": round(issue / VALU_ISSUE_PEAK, 4),
                     "valu_per_wave": FITTER_VALU_PER_WAVE, "waves": waves}
    # The reference's fitter streams its block through LDS; this one keeps it
    # in registers.  The reference's LDS bytes per block over the fitter's
    # time, against the measured LDS read peak, is the fitter's share of the
    # LDS roofline its method implies.
    import bmfr_amd
    lds_peak = bmfr_amd.lds_read_bandwidth() / 1e9
    lds_b = reference_fitter_lds_bytes()
    lds_rate = nB * lds_b / (fit_ms * 1e-3) / 1e9
    roof["lds"] = {"bound": "lds", "algorithmic_bytes_per_block": lds_b, "achieved": round(lds_rate, 1),
                   "peak": round(lds_peak, 1), "peak_nominal": round(256 * 128 * 2.4, 1), "unit": "GB/s",
                   "frac": round(lds_rate / lds_peak, 4)}
    progress(f"LDS read peak {lds_peak:.0f} GB/s; fitter at {lds_rate:.0f} GB/s of reference LDS traffic")
    batched = None
    if not args.no_batched:
        batched = run_batched(args.batched_steps, 5, args.batch_lanes)
        batched["speedup_vs_plain_720p"] = round(ms_frame / batched["ms_per_720p_frame_equiv"], 2)
        fl = batched["blocks_per_step"] * lds_b / (batched["stage_ms"]["fitter"] * 1e-3) / 1e9
        batched["fitter_lds"] = {"achieved": round(fl, 1), "peak": round(lds_peak, 1), "unit": "GB/s",
                                 "frac": round(fl / lds_peak, 4)}
        progress(f"batched: {batched['step_ms']} ms/step, x{batched['speedup_vs_plain_720p']} vs plain")
    precision = None
    if not args.no_precision:
        precision = run_precision(args.precision_frames)
        progress(f"precision 1920x1080: {precision['ms_per_frame']} ms/frame")
    feature_sets = None
    if not args.no_feature_sets:
        feature_sets = run_feature_sets(30, args.fuse)
    served = None
    if not args.no_served:
        served = run_served(args.serve_seconds, args.serve_budget_gib)
        progress(f"served: p99 {served['latency_ms_p99']} ms, {served['mpixel_per_s']} Mpixel/s")
    sweep = None
    if not args.no_served and args.serve_sweep:
        sweep = run_serve_sweep([int(x) for x in args.serve_sweep.split(",")], args.serve_sweep_seconds)
    copy_gbs = copy_peak()
    roof["copy_peak_gbs"] = round(copy_gbs, 1)
    cpu, reference_gpu = None, None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline()
        progress(f"cpu baseline: {cpu['value']:.1f} ms/frame on {cpu['cores']} threads")
        reference_gpu = reference_kernels_same_gpu()
        if reference_gpu.get("ms_per_frame"):
            reference_gpu["ratio_to_this_chain"] = round(reference_gpu["ms_per_frame"] / ms_frame, 2)
        progress(f"reference bmfr.cl on this GPU: {reference_gpu}")
    out = {
        "metric": "720p ms/frame plain vs batched; p99 ms vs streams; fitter % LDS/HBM peak",
        "value": round(ms_frame, 4),
        "unit": "ms/frame",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_frame, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded procedural plane + 24 spheres, moving camera, rendered on the GPU",
        "config": {"workload": "plain BMFR single stream 1280x720 (984 blocks/frame), FP32 "
                               "G-buffers, FP16 tmp_data", "width": W, "height": H,
                   "blocks": nB, "timed_frames": args.steps, "fuse_level": args.fuse},
        "mpixel_per_s": round(P / (ms_frame * 1e-3) / 1e6, 1),
        "frames_per_s_wall": round(args.steps / wall, 1),
        "ms_per_frame_with_stage_events": round(float(t[:, 5].mean()), 4),
        "stage_ms": {k: round(v, 4) for k, v in stage_ms.items()},
        "stage_gbs": {k: round(v, 1) for k, v in stage_gbs.items()},
        "roofline": roof,
        "batched": batched,
        "served": served,
        "latency_vs_streams": sweep,
        "precision": precision,
        "feature_sets": feature_sets,
        "cpu_baseline": cpu,
        "reference_gpu": reference_gpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
