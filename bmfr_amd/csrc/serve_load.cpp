// bmfr_serve_load: a C++ client of the server's request loop (include/
// bmfr_hip.h, bmfr_server_submit / bmfr_server_poll) that plays the role of
// many renderers: render streams open and close over a run, frames arrive as
// Poisson processes, each frame is handed over zero-copy (device G-buffers)
// with an 8-bit output surface, and the completion events give each frame's
// latency.  Prints one JSON line.
//
//   served: 64 streams over `seconds`, resolution 640x360 / 1280x720 /
//           1920x1080 / 2560x1440 with probabilities 0.2 / 0.5 / 0.25 / 0.05,
//           open at U[0, 0.8 T] for U[0.15, 0.4] T, 30 or 60 Hz, FP16 or FP32
//           storage (even odds), admitted by the HBM budget through the
//           server's admission FIFO (frames arriving before admission are
//           dropped);
//   sweep:  `streams` 1280x720 FP32 streams open for the whole run at 60 Hz;
//           latency of the frames arriving after `settle` seconds.
// A renderer holds at most `max_queue` frames of a stream in the denoiser
// at once and drops a frame that arrives when that many are outstanding (a
// real-time renderer skips a frame rather than queueing stale ones).
//
// Inputs: the procedural scene's frames 0..F-1 of each (resolution, storage)
// are rendered once into device buffers before the run, and a stream's frame
// f reads set (f + phase) mod F, read-only and shared by the streams of that
// kind; its previous-frame matrix is the previous set's camera.  The load
// measured is the denoiser's alone (no per-frame render on the GPU).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <queue>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "bmfr_hip.h"

namespace {

int admitted_before = 0;   // streams opened before the run's clock starts

#define CHECK(expr)                                                                          \
    do {                                                                                     \
        const int32_t st_ = (expr);                                                          \
        if (st_ != BMFR_OK) {                                                                \
            std::fprintf(stderr, "bmfr_serve_load: %s failed (%d): %s\n", #expr, st_,        \
                         bmfr_last_error());                                                 \
            std::exit(2);                                                                    \
        }                                                                                    \
    } while (0)
#define HIP_CHECK(expr)                                                                      \
    do {                                                                                     \
        const hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess) {                                                              \
            std::fprintf(stderr, "bmfr_serve_load: %s failed: %s\n", #expr,                  \
                         hipGetErrorString(e_));                                             \
            std::exit(2);                                                                    \
        }                                                                                    \
    } while (0)

struct Args {
    std::string mode = "served";
    int streams = 64;
    double seconds = 30.0;
    double settle = 0.0;
    double budget_gib = 4.0;
    uint64_t seed = 0xB3F40000ull + 4;
    int frames_rendered = 16;
    int max_queue = 3;
    int max_in_flight = 2;
    int spheres = 24;
};

// The frames of one (resolution, storage) kind, rendered once.
struct InputKind {
    int w = 0, h = 0;
    bool f16 = false;
    std::vector<void *> albedo, normals, positions, noisy;   // [F]
    std::vector<float> vp, offset;                           // [F][16], [F][2]
};

struct StreamSpec {
    int kind = 0;
    double t_open = 0, t_close = 0, rate = 60;
    int phase = 0;
};

struct StreamState {
    bmfr_stream stream = nullptr;
    bool requested = false, admitted = false, closed = false;
    double t_requested = 0;
    int next_frame = 0;       // frames submitted
    int outstanding = 0;      // submitted, not yet done
    void *surface = nullptr;
    int64_t pitch = 0;
};

Args parse(int argc, char **argv) {
    Args a;
    for (int i = 1; i + 1 < argc; i += 2) {
        const std::string k = argv[i], v = argv[i + 1];
        if (k == "--mode") a.mode = v;
        else if (k == "--streams") a.streams = std::atoi(v.c_str());
        else if (k == "--seconds") a.seconds = std::atof(v.c_str());
        else if (k == "--settle") a.settle = std::atof(v.c_str());
        else if (k == "--budget-gib") a.budget_gib = std::atof(v.c_str());
        else if (k == "--seed") a.seed = std::strtoull(v.c_str(), nullptr, 0);
        else if (k == "--frames-rendered") a.frames_rendered = std::atoi(v.c_str());
        else if (k == "--max-queue") a.max_queue = std::atoi(v.c_str());
        else if (k == "--max-in-flight") a.max_in_flight = std::atoi(v.c_str());
        else {
            std::fprintf(stderr, "bmfr_serve_load: unknown option %s\n", k.c_str());
            std::exit(2);
        }
    }
    if ((a.mode != "served" && a.mode != "sweep") || a.streams < 1 || a.seconds <= 0 || a.frames_rendered < 2 ||
        a.max_queue < 1) {
        std::fprintf(stderr, "bmfr_serve_load: bad arguments\n");
        std::exit(2);
    }
    return a;
}

int kind_of(std::vector<InputKind> &kinds, int w, int h, bool f16) {
    for (size_t i = 0; i < kinds.size(); ++i)
        if (kinds[i].w == w && kinds[i].h == h && kinds[i].f16 == f16) return (int)i;
    InputKind k;
    k.w = w;
    k.h = h;
    k.f16 = f16;
    kinds.push_back(k);
    return (int)kinds.size() - 1;
}

void render_kind(InputKind &k, int frames, int spheres, uint64_t seed) {
    bmfr_scene sc;
    CHECK(bmfr_scene_create(seed, spheres, 1e-3f, k.w, k.h, 0, &sc));
    const size_t n = (size_t)k.w * k.h * 3, g = k.f16 ? 2 : 4;
    k.vp.resize(16 * (size_t)frames);
    k.offset.resize(2 * (size_t)frames);
    for (int f = 0; f < frames; ++f) {
        void *p[4];
        HIP_CHECK(hipMalloc(&p[0], n * 4));
        for (int i = 1; i < 4; ++i) HIP_CHECK(hipMalloc(&p[i], n * g));
        CHECK(bmfr_scene_render_device(sc, f, k.f16 ? BMFR_STORAGE_F16 : BMFR_STORAGE_F32, p[0], p[1], p[2], p[3],
                                       nullptr));
        k.albedo.push_back(p[0]);
        k.normals.push_back(p[1]);
        k.positions.push_back(p[2]);
        k.noisy.push_back(p[3]);
        CHECK(bmfr_scene_camera(sc, f, &k.vp[16 * (size_t)f], &k.offset[2 * (size_t)f]));
    }
    HIP_CHECK(hipDeviceSynchronize());
    CHECK(bmfr_scene_destroy(sc));
}

double percentile(std::vector<double> v, double q) {
    if (v.empty()) return -1.0;
    std::sort(v.begin(), v.end());
    const double pos = q / 100.0 * (double)(v.size() - 1);
    const size_t lo = (size_t)pos, hi = std::min(lo + 1, v.size() - 1);
    return v[lo] + (v[hi] - v[lo]) * (pos - (double)lo);
}

}  // namespace

int main(int argc, char **argv) {
    const Args a = parse(argc, argv);
    std::mt19937_64 rng(a.seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);

    // ---- the streams of the run
    std::vector<InputKind> kinds;
    std::vector<StreamSpec> specs(a.streams);
    const int res[4][2] = {{640, 360}, {1280, 720}, {1920, 1080}, {2560, 1440}};
    const double cum[4] = {0.2, 0.7, 0.95, 1.0};
    for (int k = 0; k < a.streams; ++k) {
        StreamSpec &s = specs[k];
        if (a.mode == "served") {
            const double r = U(rng);
            int ri = 0;
            while (ri < 3 && r >= cum[ri]) ++ri;
            s.t_open = U(rng) * 0.8 * a.seconds;
            s.t_close = std::min(a.seconds, s.t_open + (0.15 + 0.25 * U(rng)) * a.seconds);
            s.rate = U(rng) < 0.5 ? 30.0 : 60.0;
            const bool f16 = U(rng) < 0.5;
            s.kind = kind_of(kinds, res[ri][0], res[ri][1], f16);
        } else {
            s.t_open = 0.0;
            s.t_close = a.seconds;
            s.rate = 60.0;
            s.kind = kind_of(kinds, 1280, 720, false);
        }
        s.phase = (int)(rng() % (uint64_t)a.frames_rendered);
    }
    for (size_t i = 0; i < kinds.size(); ++i)
        render_kind(kinds[i], a.frames_rendered, a.spheres, a.seed + 100 + i);

    bmfr_server sv;
    CHECK(bmfr_server_create((int64_t)(a.budget_gib * (double)(1ull << 30)), &sv));
    CHECK(bmfr_server_set_policy(sv, a.max_in_flight, 64));
    std::vector<StreamState> st(a.streams);
    if (a.mode == "sweep") {
        // every stream is open before the clock starts (opening allocates)
        for (int k = 0; k < a.streams; ++k) {
            const InputKind &kd = kinds[specs[k].kind];
            bmfr_params p;
            CHECK(bmfr_default_params(kd.w, kd.h, &p));
            CHECK(bmfr_server_open_stream(sv, &p, &st[k].stream));
            st[k].requested = st[k].admitted = true;
            ++admitted_before;
        }
    }

    // arrivals: a min-heap of (time, stream), Poisson per stream
    using Arr = std::pair<double, int>;
    std::priority_queue<Arr, std::vector<Arr>, std::greater<Arr>> arrivals;
    std::vector<std::exponential_distribution<double>> gap;
    for (int k = 0; k < a.streams; ++k) {
        gap.emplace_back(specs[k].rate);
        const double t = specs[k].t_open + gap[k](rng);
        if (t < specs[k].t_close) arrivals.push({t, k});
    }

    std::vector<double> lat;
    std::map<int64_t, std::pair<int, float>> steps;   // step id -> (streams, gpu ms)
    int64_t frames_done = 0, pixels_done = 0, dropped = 0, arrived = 0, failed = 0;
    // throughput and step time over the arrival window [0, seconds] (not the
    // drain and the closes after it): frames done and steps completed in it.
    // step_time_per_s sums the steps' GPU time (first start to last end) per
    // second of the window; steps on the server's two lanes overlap, so it
    // can exceed 1.
    int64_t frames_in_window = 0, pixels_in_window = 0;
    double gpu_ms_in_window = 0;
    int admitted = admitted_before, admitted_after_waiting = 0, never_admitted = 0;
    const double t0 = bmfr_clock_seconds();
    std::vector<bmfr_server_event> ev(1024);
    int64_t outstanding_total = 0;
    for (;;) {
        const double now = bmfr_clock_seconds() - t0;
        // opens: request admission (the FIFO admits later when memory frees)
        for (int k = 0; k < a.streams; ++k) {
            StreamState &s = st[k];
            if (!s.requested && now >= specs[k].t_open) {
                s.requested = true;
                s.t_requested = now;
                const InputKind &kd = kinds[specs[k].kind];
                bmfr_params p;
                CHECK(bmfr_default_params(kd.w, kd.h, &p));
                p.storage = kd.f16 ? BMFR_STORAGE_F16 : BMFR_STORAGE_F32;
                CHECK(bmfr_server_request_stream(sv, &p, (uint64_t)k, &s.stream));
                if (s.stream) {
                    s.admitted = true;
                    ++admitted;
                }
            }
        }
        // arrivals that are due
        while (!arrivals.empty() && arrivals.top().first <= now) {
            const Arr x = arrivals.top();
            arrivals.pop();
            const int k = x.second;
            StreamState &s = st[k];
            const StreamSpec &sp = specs[k];
            const double t_next = x.first + gap[k](rng);
            if (t_next < sp.t_close) arrivals.push({t_next, k});
            ++arrived;
            if (!s.admitted || s.closed || s.outstanding >= a.max_queue) {
                ++dropped;
                continue;
            }
            const InputKind &kd = kinds[sp.kind];
            if (!s.surface) {
                s.pitch = (int64_t)kd.w * 4;
                HIP_CHECK(hipMalloc(&s.surface, (size_t)s.pitch * kd.h));
            }
            const int F = a.frames_rendered;
            const int idx = (s.next_frame + sp.phase) % F, pidx = (idx + F - 1) % F;
            bmfr_frame_request r{};
            r.stream = s.stream;
            r.frame_number = s.next_frame;
            std::memcpy(r.prev_frame_camera_matrix, &kd.vp[16 * (size_t)(s.next_frame ? pidx : idx)],
                        sizeof(r.prev_frame_camera_matrix));
            std::memcpy(r.pixel_offset, &kd.offset[2 * (size_t)idx], sizeof(r.pixel_offset));
            r.albedo = kd.albedo[idx];
            r.normals = kd.normals[idx];
            r.positions = kd.positions[idx];
            r.noisy = kd.noisy[idx];
            r.surface = s.surface;
            r.surface_format = BMFR_SURFACE_RGBA_U8;
            r.surface_row_pitch = s.pitch;
            r.t_arrival = t0 + x.first;
            r.tag = (uint64_t)k;
            CHECK(bmfr_server_submit(sv, &r));
            ++s.next_frame;
            ++s.outstanding;
            ++outstanding_total;
        }
        // events: completions and admissions
        const double next_due = arrivals.empty() ? now + 1e-3 : arrivals.top().first;
        const int wait_us = (int)std::max(0.0, std::min(1000.0, (next_due - now) * 1e6));
        int32_t n = 0;
        CHECK(bmfr_server_poll(sv, ev.data(), (int32_t)ev.size(), wait_us, &n));
        for (int i = 0; i < n; ++i) {
            const bmfr_server_event &e = ev[i];
            const int k = (int)e.tag;
            if (e.kind == BMFR_EVENT_STREAM_ADMITTED) {
                if (e.status != BMFR_OK) {
                    std::fprintf(stderr, "bmfr_serve_load: admission failed: %d\n", e.status);
                    std::exit(2);
                }
                st[k].stream = e.stream;
                st[k].admitted = true;
                ++admitted;
                if (e.t_done - (t0 + st[k].t_requested) > 0.005) ++admitted_after_waiting;
                continue;
            }
            --st[k].outstanding;
            --outstanding_total;
            if (e.kind != BMFR_EVENT_FRAME_DONE) {
                ++failed;
                continue;
            }
            const InputKind &kd = kinds[specs[k].kind];
            ++frames_done;
            pixels_done += (int64_t)kd.w * kd.h;
            if (e.t_arrival - t0 >= a.settle) lat.push_back((e.t_done - e.t_arrival) * 1e3);
            if (e.t_done - t0 <= a.seconds) {
                ++frames_in_window;
                pixels_in_window += (int64_t)kd.w * kd.h;
                if (!steps.count(e.step_id)) gpu_ms_in_window += e.step_gpu_ms;
            }
            steps[e.step_id] = {e.step_streams, e.step_gpu_ms};
        }
        // closes: after the stream's end, once its frames are done (a stream
        // still waiting for admission withdraws its request)
        const double t = bmfr_clock_seconds() - t0;
        bool all_closed = true;
        for (int k = 0; k < a.streams; ++k) {
            StreamState &s = st[k];
            if (s.closed) continue;
            if (t >= specs[k].t_close && s.outstanding == 0 && s.requested) {
                if (s.admitted) {
                    CHECK(bmfr_server_close_stream(sv, s.stream));
                } else {
                    int32_t found = 0;
                    CHECK(bmfr_server_cancel_request(sv, (uint64_t)k, &found));
                    if (!found) {
                        all_closed = false;   // its admission event is on its way
                        continue;
                    }
                    ++never_admitted;
                }
                s.closed = true;
                if (s.surface) HIP_CHECK(hipFree(s.surface));
                s.surface = nullptr;
                continue;
            }
            all_closed = false;
        }
        if (all_closed && arrivals.empty() && outstanding_total == 0) break;
        if (t > a.seconds + 60.0) {
            std::fprintf(stderr, "bmfr_serve_load: did not drain within 60 s of the end\n");
            std::exit(3);
        }
    }
    const double wall = bmfr_clock_seconds() - t0;
    CHECK(bmfr_server_destroy(sv));
    double gpu_ms = 0, step_streams = 0;
    for (auto &kv : steps) {
        gpu_ms += kv.second.second;
        step_streams += kv.second.first;
    }
    (void)pixels_done;
    const double nsteps = (double)steps.size();
    std::printf("{\"mode\": \"%s\", \"streams\": %d, \"duration_s\": %.2f, \"budget_gib\": %.1f, "
                "\"admitted\": %d, \"admitted_after_waiting\": %d, \"never_admitted\": %d, "
                "\"frames_arrived\": %lld, \"frames\": %lld, \"dropped\": %lld, \"failed\": %lld, "
                "\"aggregate_frames_per_s\": %.1f, \"mpixel_per_s\": %.1f, "
                "\"latency_ms_p50\": %.3f, \"latency_ms_p99\": %.3f, \"latency_ms_max\": %.3f, "
                "\"steps\": %lld, \"mean_streams_per_step\": %.2f, \"step_gpu_ms_mean\": %.4f, "
                "\"step_time_per_s\": %.3f, \"max_queue_per_stream\": %d, \"max_in_flight\": %d}\n",
                a.mode.c_str(), a.streams, wall, a.budget_gib, admitted, admitted_after_waiting, never_admitted,
                (long long)arrived, (long long)frames_done, (long long)dropped, (long long)failed,
                (double)frames_in_window / a.seconds, (double)pixels_in_window / a.seconds / 1e6, percentile(lat, 50),
                percentile(lat, 99), lat.empty() ? -1.0 : *std::max_element(lat.begin(), lat.end()),
                (long long)steps.size(), nsteps ? step_streams / nsteps : 0.0, nsteps ? gpu_ms / nsteps : 0.0,
                gpu_ms_in_window * 1e-3 / a.seconds, a.max_queue, a.max_in_flight);
    for (InputKind &k : kinds)
        for (int f = 0; f < (int)k.albedo.size(); ++f) {
            (void)hipFree(k.albedo[f]);
            (void)hipFree(k.normals[f]);
            (void)hipFree(k.positions[f]);
            (void)hipFree(k.noisy[f]);
        }
    return failed ? 1 : 0;
}
