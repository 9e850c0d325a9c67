// Reference runner: TEST INFRASTRUCTURE ONLY.
//
// Runs the reference's own OpenCL kernels -- opencl/bmfr.cl, compiled
// unmodified for gfx950 by oracle/Makefile (target `ref`) into
// oracle/_ref/bmfr_<W>x<H>_h<0|1>.co -- on the GPU through the ROCm OpenCL
// runtime, so that the HIP path can be compared with the reference itself
// on identical inputs.  This host is written for that purpose; it reproduces
// what the reference's host does for the kernels:
//   buffer set and sizes        opencl/bmfr.cpp:315-343
//   static argument bindings    opencl/bmfr.cpp:349-384
//   per-frame bindings, ranges  opencl/bmfr.cpp:245-249, 429-476
//   double-buffer swaps         opencl/bmfr.cpp:482-484
// The -D values (resolution, features, limits 0.01/0.1) are fixed when the
// code object is compiled (opencl/bmfr.cpp:205-232).
#define CL_TARGET_OPENCL_VERSION 120
#include <CL/cl.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(const char *what, cl_int e) {
    char b[256];
    snprintf(b, sizeof(b), "%s failed (%d)", what, (int)e);
    g_err = b;
    return e == 0 ? -1 : (int)e;
}

#define CLT(x)                          \
    do {                                \
        cl_int e_ = (x);                \
        if (e_ != CL_SUCCESS) return fail(#x, e_); \
    } while (0)

// buffer ids of ref_read/ref_write
enum {
    R_NORMALS_CUR, R_NORMALS_PREV, R_POSITIONS_CUR, R_POSITIONS_PREV, R_NOISY_CUR, R_NOISY_PREV,
    R_TMP, R_FILTERED, R_OUT_CUR, R_OUT_PREV, R_RESULT_CUR, R_RESULT_PREV, R_PREV_PIXELS,
    R_ACCEPT, R_ALBEDO, R_TONE, R_WEIGHTS, R_MINS_MAXS, R_SPP_CUR, R_SPP_PREV, R_COUNT
};

struct Dbl {
    cl_mem m[2] = {nullptr, nullptr};
    bool swapped = false;
    cl_mem cur() const { return swapped ? m[0] : m[1]; }  // Double_buffer::current
    cl_mem prev() const { return swapped ? m[1] : m[0]; }
};

}  // namespace

struct RefRun {
    int W = 0, H = 0, half = 1;
    int ws_w = 0, ws_h = 0, mg_w = 0, mg_h = 0, blocks = 0;
    cl_context ctx = nullptr;
    cl_command_queue q = nullptr;
    cl_program prog = nullptr;
    cl_kernel k_accum = nullptr, k_fit = nullptr, k_ws = nullptr, k_af = nullptr, k_taa = nullptr;
    Dbl normals, positions, noisy, out, result, spp;
    cl_mem in_buffer = nullptr, filtered = nullptr, prev_pixels = nullptr, accept = nullptr,
           albedo = nullptr, tone = nullptr, weights = nullptr, mins_maxs = nullptr;
    size_t bytes[R_COUNT] = {};
    cl_event ev[5] = {};

    cl_mem resolve(int id) const {
        switch (id) {
            case R_NORMALS_CUR: return normals.cur();
            case R_NORMALS_PREV: return normals.prev();
            case R_POSITIONS_CUR: return positions.cur();
            case R_POSITIONS_PREV: return positions.prev();
            case R_NOISY_CUR: return noisy.cur();
            case R_NOISY_PREV: return noisy.prev();
            case R_TMP: return in_buffer;
            case R_FILTERED: return filtered;
            case R_OUT_CUR: return out.cur();
            case R_OUT_PREV: return out.prev();
            case R_RESULT_CUR: return result.cur();
            case R_RESULT_PREV: return result.prev();
            case R_PREV_PIXELS: return prev_pixels;
            case R_ACCEPT: return accept;
            case R_ALBEDO: return albedo;
            case R_TONE: return tone;
            case R_WEIGHTS: return weights;
            case R_MINS_MAXS: return mins_maxs;
            case R_SPP_CUR: return spp.cur();
            case R_SPP_PREV: return spp.prev();
            default: return nullptr;
        }
    }
};

static int make_buf(RefRun *r, cl_mem *m, size_t bytes) {
    cl_int e;
    *m = clCreateBuffer(r->ctx, CL_MEM_READ_WRITE, bytes, nullptr, &e);
    if (e != CL_SUCCESS) return fail("clCreateBuffer", e);
    // zero it so that runs are reproducible (the reference leaves it unset)
    std::vector<char> z(bytes, 0);
    CLT(clEnqueueWriteBuffer(r->q, *m, CL_TRUE, 0, bytes, z.data(), 0, nullptr, nullptr));
    return 0;
}

extern "C" {

const char *ref_error(void) { return g_err.c_str(); }

int ref_device_count(void) {
    cl_uint np = 0;
    if (clGetPlatformIDs(0, nullptr, &np) != CL_SUCCESS || np == 0) return 0;
    std::vector<cl_platform_id> ps(np);
    clGetPlatformIDs(np, ps.data(), nullptr);
    cl_uint total = 0;
    for (auto p : ps) {
        cl_uint nd = 0;
        if (clGetDeviceIDs(p, CL_DEVICE_TYPE_GPU, 0, nullptr, &nd) == CL_SUCCESS) total += nd;
    }
    return (int)total;
}

// features / scaled: the code object's BUFFER_COUNT - 3 and FEATURES_SCALED
// (10 and 6 for the reference's set), which size tmp_data, weights, mins_maxs
// and the local R matrix (opencl/bmfr.cpp:321-343, 354-358).
int ref_open(const char *code_object, int W, int H, int half, int features, int scaled, void **out) {
    *out = nullptr;
    FILE *f = fopen(code_object, "rb");
    if (!f) {
        g_err = std::string("cannot open ") + code_object;
        return -1;
    }
    std::vector<unsigned char> bin;
    unsigned char buf[65536];
    size_t n;
    while ((n = fread(buf, 1, sizeof(buf), f)) > 0) bin.insert(bin.end(), buf, buf + n);
    fclose(f);

    RefRun *r = new RefRun;
    r->W = W;
    r->H = H;
    r->half = half;
    // opencl/bmfr.cpp:107-118
    r->ws_w = 32 * ((W + 31) / 32);
    r->ws_h = 32 * ((H + 31) / 32);
    r->mg_w = r->ws_w + 32;
    r->mg_h = r->ws_h + 32;
    r->blocks = (r->mg_w / 32) * (r->mg_h / 32);

    cl_uint np = 0;
    CLT(clGetPlatformIDs(0, nullptr, &np));
    std::vector<cl_platform_id> ps(np);
    CLT(clGetPlatformIDs(np, ps.data(), nullptr));
    cl_device_id dev = nullptr;
    for (auto p : ps) {
        cl_uint nd = 0;
        if (clGetDeviceIDs(p, CL_DEVICE_TYPE_GPU, 1, &dev, &nd) == CL_SUCCESS && nd > 0) break;
        dev = nullptr;
    }
    if (!dev) {
        g_err = "no OpenCL GPU device";
        delete r;
        return -2;
    }
    cl_int e;
    r->ctx = clCreateContext(nullptr, 1, &dev, nullptr, nullptr, &e);
    if (e != CL_SUCCESS) return fail("clCreateContext", e);
    r->q = clCreateCommandQueue(r->ctx, dev, CL_QUEUE_PROFILING_ENABLE, &e);  // bmfr.cpp:191
    if (e != CL_SUCCESS) return fail("clCreateCommandQueue", e);
    const unsigned char *bp = bin.data();
    size_t bl = bin.size();
    cl_int bst;
    r->prog = clCreateProgramWithBinary(r->ctx, 1, &dev, &bl, &bp, &bst, &e);
    if (e != CL_SUCCESS) return fail("clCreateProgramWithBinary", e);
    e = clBuildProgram(r->prog, 1, &dev, "", nullptr, nullptr);
    if (e != CL_SUCCESS) {
        char log[4096] = {0};
        clGetProgramBuildInfo(r->prog, dev, CL_PROGRAM_BUILD_LOG, sizeof(log) - 1, log, nullptr);
        g_err = std::string("clBuildProgram: ") + log;
        return (int)e;
    }
    r->k_accum = clCreateKernel(r->prog, "accumulate_noisy_data", &e);
    if (e != CL_SUCCESS) return fail("clCreateKernel accumulate_noisy_data", e);
    r->k_fit = clCreateKernel(r->prog, "fitter", &e);
    if (e != CL_SUCCESS) return fail("clCreateKernel fitter", e);
    r->k_ws = clCreateKernel(r->prog, "weighted_sum", &e);
    if (e != CL_SUCCESS) return fail("clCreateKernel weighted_sum", e);
    r->k_af = clCreateKernel(r->prog, "accumulate_filtered_data", &e);
    if (e != CL_SUCCESS) return fail("clCreateKernel accumulate_filtered_data", e);
    r->k_taa = clCreateKernel(r->prog, "taa", &e);
    if (e != CL_SUCCESS) return fail("clCreateKernel taa", e);

    // Buffers, opencl/bmfr.cpp:315-343 (OUTPUT_SIZE = workset W x H)
    const size_t OUT = (size_t)r->ws_w * r->ws_h, MG = (size_t)r->mg_w * r->mg_h;
    const size_t P = (size_t)W * H;
    struct {
        cl_mem *m;
        size_t b;
        int id;
    } bufs[] = {
        {&r->normals.m[0], OUT * 12, -1}, {&r->normals.m[1], OUT * 12, -1},
        {&r->positions.m[0], OUT * 12, -1}, {&r->positions.m[1], OUT * 12, -1},
        {&r->noisy.m[0], OUT * 12, -1}, {&r->noisy.m[1], OUT * 12, -1},
        {&r->in_buffer, MG * (features + 3) * (half ? 2 : 4), R_TMP}, {&r->filtered, OUT * 12, R_FILTERED},
        {&r->out.m[0], MG * 12, -1}, {&r->out.m[1], MG * 12, -1},
        {&r->result.m[0], OUT * 12, -1}, {&r->result.m[1], OUT * 12, -1},
        {&r->prev_pixels, OUT * 8, R_PREV_PIXELS}, {&r->accept, OUT, R_ACCEPT},
        {&r->albedo, P * 12, R_ALBEDO}, {&r->tone, P * 12, R_TONE},
        {&r->weights, (size_t)r->blocks * 3 * features * 4, R_WEIGHTS},
        {&r->mins_maxs, (size_t)r->blocks * 2 * (scaled ? scaled : 1) * 4, R_MINS_MAXS},
        {&r->spp.m[0], OUT, -1}, {&r->spp.m[1], OUT, -1},
    };
    for (auto &b : bufs)
        if (int rc = make_buf(r, b.m, b.b)) return rc;
    r->bytes[R_NORMALS_CUR] = r->bytes[R_NORMALS_PREV] = OUT * 12;
    r->bytes[R_POSITIONS_CUR] = r->bytes[R_POSITIONS_PREV] = OUT * 12;
    r->bytes[R_NOISY_CUR] = r->bytes[R_NOISY_PREV] = OUT * 12;
    r->bytes[R_OUT_CUR] = r->bytes[R_OUT_PREV] = MG * 12;
    r->bytes[R_RESULT_CUR] = r->bytes[R_RESULT_PREV] = OUT * 12;
    r->bytes[R_SPP_CUR] = r->bytes[R_SPP_PREV] = OUT;
    for (auto &b : bufs)
        if (b.id >= 0) r->bytes[b.id] = b.b;

    // Static arguments, opencl/bmfr.cpp:349-384
    CLT(clSetKernelArg(r->k_accum, 0, sizeof(cl_mem), &r->prev_pixels));
    CLT(clSetKernelArg(r->k_accum, 1, sizeof(cl_mem), &r->accept));
    const size_t r_size = ((features + 1) * (features + 2) / 2) * sizeof(cl_float3);  // COMPRESSED_R, bmfr.cpp:354-358
    CLT(clSetKernelArg(r->k_fit, 0, 256 * sizeof(float), nullptr));
    CLT(clSetKernelArg(r->k_fit, 1, 1024 * sizeof(float), nullptr));
    CLT(clSetKernelArg(r->k_fit, 2, r_size, nullptr));
    CLT(clSetKernelArg(r->k_fit, 3, sizeof(cl_mem), &r->weights));
    CLT(clSetKernelArg(r->k_fit, 4, sizeof(cl_mem), &r->mins_maxs));
    CLT(clSetKernelArg(r->k_ws, 0, sizeof(cl_mem), &r->weights));
    CLT(clSetKernelArg(r->k_ws, 1, sizeof(cl_mem), &r->mins_maxs));
    CLT(clSetKernelArg(r->k_ws, 2, sizeof(cl_mem), &r->filtered));
    CLT(clSetKernelArg(r->k_af, 0, sizeof(cl_mem), &r->filtered));
    CLT(clSetKernelArg(r->k_af, 1, sizeof(cl_mem), &r->prev_pixels));
    CLT(clSetKernelArg(r->k_af, 2, sizeof(cl_mem), &r->accept));
    CLT(clSetKernelArg(r->k_af, 3, sizeof(cl_mem), &r->albedo));
    CLT(clSetKernelArg(r->k_af, 4, sizeof(cl_mem), &r->tone));
    CLT(clSetKernelArg(r->k_taa, 0, sizeof(cl_mem), &r->prev_pixels));
    CLT(clSetKernelArg(r->k_taa, 1, sizeof(cl_mem), &r->tone));
    CLT(clFinish(r->q));
    *out = r;
    return 0;
}

// opencl/bmfr.cpp:420-427: blocking uploads of W*H*3 floats.
int ref_upload(void *h, const float *albedo, const float *normals, const float *positions,
               const float *noisy) {
    RefRun *r = (RefRun *)h;
    const size_t b = (size_t)r->W * r->H * 12;
    CLT(clEnqueueWriteBuffer(r->q, r->albedo, CL_TRUE, 0, b, albedo, 0, nullptr, nullptr));
    CLT(clEnqueueWriteBuffer(r->q, r->normals.cur(), CL_TRUE, 0, b, normals, 0, nullptr, nullptr));
    CLT(clEnqueueWriteBuffer(r->q, r->positions.cur(), CL_TRUE, 0, b, positions, 0, nullptr, nullptr));
    CLT(clEnqueueWriteBuffer(r->q, r->noisy.cur(), CL_TRUE, 0, b, noisy, 0, nullptr, nullptr));
    return 0;
}

// One stage with the per-frame bindings of opencl/bmfr.cpp:429-476.
int ref_stage(void *h, int stage, int frame, const float *matrix, const float *offset) {
    RefRun *r = (RefRun *)h;
    const size_t accum_g[2] = {(size_t)r->mg_w, (size_t)r->mg_h};
    const size_t out_g[2] = {(size_t)r->ws_w, (size_t)r->ws_h};
    const size_t loc[2] = {8, 8};
    const size_t fit_g = 256 * (size_t)r->blocks, fit_l = 256;
    cl_mem m;
    if (r->ev[stage]) {
        clReleaseEvent(r->ev[stage]);
        r->ev[stage] = nullptr;
    }
    switch (stage) {
        case 0: {
            m = r->normals.cur(); CLT(clSetKernelArg(r->k_accum, 2, sizeof(cl_mem), &m));
            m = r->normals.prev(); CLT(clSetKernelArg(r->k_accum, 3, sizeof(cl_mem), &m));
            m = r->positions.cur(); CLT(clSetKernelArg(r->k_accum, 4, sizeof(cl_mem), &m));
            m = r->positions.prev(); CLT(clSetKernelArg(r->k_accum, 5, sizeof(cl_mem), &m));
            m = r->noisy.cur(); CLT(clSetKernelArg(r->k_accum, 6, sizeof(cl_mem), &m));
            m = r->noisy.prev(); CLT(clSetKernelArg(r->k_accum, 7, sizeof(cl_mem), &m));
            m = r->spp.prev(); CLT(clSetKernelArg(r->k_accum, 8, sizeof(cl_mem), &m));
            m = r->spp.cur(); CLT(clSetKernelArg(r->k_accum, 9, sizeof(cl_mem), &m));
            CLT(clSetKernelArg(r->k_accum, 10, sizeof(cl_mem), &r->in_buffer));
            cl_float16 mat;
            memcpy(&mat, matrix, sizeof(mat));
            cl_float2 off;
            off.s[0] = offset[0];
            off.s[1] = offset[1];
            cl_int fr = frame;
            CLT(clSetKernelArg(r->k_accum, 11, sizeof(cl_float16), &mat));
            CLT(clSetKernelArg(r->k_accum, 12, sizeof(cl_float2), &off));
            CLT(clSetKernelArg(r->k_accum, 13, sizeof(cl_int), &fr));
            CLT(clEnqueueNDRangeKernel(r->q, r->k_accum, 2, nullptr, accum_g, loc, 0, nullptr, &r->ev[0]));
            break;
        }
        case 1: {
            cl_int fr = frame;
            CLT(clSetKernelArg(r->k_fit, 5, sizeof(cl_mem), &r->in_buffer));
            CLT(clSetKernelArg(r->k_fit, 6, sizeof(cl_int), &fr));
            CLT(clEnqueueNDRangeKernel(r->q, r->k_fit, 1, nullptr, &fit_g, &fit_l, 0, nullptr, &r->ev[1]));
            break;
        }
        case 2: {
            cl_int fr = frame;
            m = r->normals.cur(); CLT(clSetKernelArg(r->k_ws, 3, sizeof(cl_mem), &m));
            m = r->positions.cur(); CLT(clSetKernelArg(r->k_ws, 4, sizeof(cl_mem), &m));
            m = r->noisy.cur(); CLT(clSetKernelArg(r->k_ws, 5, sizeof(cl_mem), &m));
            CLT(clSetKernelArg(r->k_ws, 6, sizeof(cl_int), &fr));
            CLT(clEnqueueNDRangeKernel(r->q, r->k_ws, 2, nullptr, out_g, loc, 0, nullptr, &r->ev[2]));
            break;
        }
        case 3: {
            cl_int fr = frame;
            m = r->spp.cur(); CLT(clSetKernelArg(r->k_af, 5, sizeof(cl_mem), &m));
            m = r->out.prev(); CLT(clSetKernelArg(r->k_af, 6, sizeof(cl_mem), &m));
            m = r->out.cur(); CLT(clSetKernelArg(r->k_af, 7, sizeof(cl_mem), &m));
            CLT(clSetKernelArg(r->k_af, 8, sizeof(cl_int), &fr));
            CLT(clEnqueueNDRangeKernel(r->q, r->k_af, 2, nullptr, out_g, loc, 0, nullptr, &r->ev[3]));
            break;
        }
        case 4: {
            cl_int fr = frame;
            m = r->result.cur(); CLT(clSetKernelArg(r->k_taa, 2, sizeof(cl_mem), &m));
            m = r->result.prev(); CLT(clSetKernelArg(r->k_taa, 3, sizeof(cl_mem), &m));
            CLT(clSetKernelArg(r->k_taa, 4, sizeof(cl_int), &fr));
            CLT(clEnqueueNDRangeKernel(r->q, r->k_taa, 2, nullptr, out_g, loc, 0, nullptr, &r->ev[4]));
            break;
        }
        default:
            g_err = "bad stage";
            return -1;
    }
    CLT(clFinish(r->q));
    return 0;
}

// opencl/bmfr.cpp:482-484
int ref_end_frame(void *h) {
    RefRun *r = (RefRun *)h;
    for (Dbl *d : {&r->normals, &r->positions, &r->noisy, &r->out, &r->result, &r->spp})
        d->swapped = !d->swapped;
    return 0;
}

// Kernel time of the last launch of `stage` (ms), from the queue's events.
int ref_stage_ms(void *h, int stage, double *ms) {
    RefRun *r = (RefRun *)h;
    if (stage < 0 || stage > 4 || !r->ev[stage]) return -1;
    cl_ulong t0, t1;
    CLT(clGetEventProfilingInfo(r->ev[stage], CL_PROFILING_COMMAND_START, sizeof(t0), &t0, nullptr));
    CLT(clGetEventProfilingInfo(r->ev[stage], CL_PROFILING_COMMAND_END, sizeof(t1), &t1, nullptr));
    *ms = (double)(t1 - t0) * 1e-6;
    return 0;
}

int ref_buffer_bytes(void *h, int id, size_t *bytes) {
    RefRun *r = (RefRun *)h;
    if (id < 0 || id >= R_COUNT) return -1;
    *bytes = r->bytes[id];
    return 0;
}

int ref_read(void *h, int id, void *dst, size_t bytes) {
    RefRun *r = (RefRun *)h;
    cl_mem m = r->resolve(id);
    if (!m || bytes > r->bytes[id]) {
        g_err = "bad buffer id or size";
        return -1;
    }
    CLT(clEnqueueReadBuffer(r->q, m, CL_TRUE, 0, bytes, dst, 0, nullptr, nullptr));
    return 0;
}

int ref_write(void *h, int id, const void *src, size_t bytes) {
    RefRun *r = (RefRun *)h;
    cl_mem m = r->resolve(id);
    if (!m || bytes > r->bytes[id]) {
        g_err = "bad buffer id or size";
        return -1;
    }
    CLT(clEnqueueWriteBuffer(r->q, m, CL_TRUE, 0, bytes, src, 0, nullptr, nullptr));
    return 0;
}

int ref_close(void *h) {
    RefRun *r = (RefRun *)h;
    if (!r) return 0;
    if (r->q) clFinish(r->q);
    for (auto &e : r->ev)
        if (e) clReleaseEvent(e);
    for (Dbl *d : {&r->normals, &r->positions, &r->noisy, &r->out, &r->result, &r->spp})
        for (auto &m : d->m)
            if (m) clReleaseMemObject(m);
    for (cl_mem m : {r->in_buffer, r->filtered, r->prev_pixels, r->accept, r->albedo, r->tone,
                     r->weights, r->mins_maxs})
        if (m) clReleaseMemObject(m);
    for (cl_kernel k : {r->k_accum, r->k_fit, r->k_ws, r->k_af, r->k_taa})
        if (k) clReleaseKernel(k);
    if (r->prog) clReleaseProgram(r->prog);
    if (r->q) clReleaseCommandQueue(r->q);
    if (r->ctx) clReleaseContext(r->ctx);
    delete r;
    return 0;
}

}  // extern "C"
