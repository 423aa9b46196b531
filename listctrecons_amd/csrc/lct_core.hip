// lct_core.hip -- context, device memory, error reporting, planner and
// ramp taps of liblct.so (host code).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "lct_internal.hpp"

namespace lct {

static thread_local std::string g_last_error;

int set_error(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

int hip_error(hipError_t e, const char *what)
{
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    return set_error(e == hipErrorOutOfMemory ? LCT_ERR_NOMEM : LCT_ERR_HIP, m);
}

void clear_error() { g_last_error.clear(); }

int scratch(lct_ctx *ctx, int slot, size_t bytes, void **out)
{
    lct_ctx::Buf &b = ctx->scratch[slot];
    if (b.bytes < bytes) {
        if (b.ptr) {
            LCT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
            LCT_CHECK_HIP(hipFree(b.ptr));
            b.ptr = nullptr;
            b.bytes = 0;
        }
        size_t want = (bytes + 4095) & ~size_t(4095);
        hipError_t e = hipMalloc(&b.ptr, want);
        if (e != hipSuccess) {
            b.ptr = nullptr;
            (void)hipGetLastError();
            return set_error(LCT_ERR_NOMEM, "device scratch allocation of " +
                                                std::to_string(want) + " bytes failed");
        }
        b.bytes = want;
    }
    *out = b.ptr;
    return LCT_OK;
}

static int staging_init(lct_ctx *ctx)
{
    lct_ctx::Staging &s = ctx->staging;
    if (s.base || s.failed) return LCT_OK;
    void *p = nullptr;
    if (hipHostMalloc(&p, 2 * lct_ctx::kStagingHalf, hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        s.failed = true;
        return LCT_OK;
    }
    for (int k = 0; k < 2; k++) {
        hipError_t e = hipEventCreateWithFlags(&s.ev[k], hipEventDisableTiming);
        if (e != hipSuccess) {
            if (k == 1) (void)hipEventDestroy(s.ev[0]);
            (void)hipHostFree(p);
            return hip_error(e, "hipEventCreateWithFlags");
        }
    }
    s.base = static_cast<char *>(p);
    return LCT_OK;
}

void staging_release(lct_ctx *ctx)
{
    lct_ctx::Staging &s = ctx->staging;
    if (!s.base) return;
    for (int k = 0; k < 2; k++) {
        if (s.pending[k]) (void)hipEventSynchronize(s.ev[k]);
        (void)hipEventDestroy(s.ev[k]);
    }
    (void)hipHostFree(s.base);
    s = lct_ctx::Staging{};
}

int upload(lct_ctx *ctx, void *d, const void *h, size_t bytes)
{
    // Ordered after earlier work on the context stream (which may still read
    // the previous table in `d`); returns once the host buffer is consumed.
    if (bytes == 0) return LCT_OK;
    LCT_CHECK(staging_init(ctx));
    lct_ctx::Staging &s = ctx->staging;
    if (!s.base || bytes > lct_ctx::kStagingHalf) {
        // no staging ring, or a table larger than a half: copy and wait
        LCT_CHECK_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, ctx->stream));
        LCT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
        return LCT_OK;
    }
    size_t off = (s.used + 255) & ~size_t(255);
    if (off + bytes > lct_ctx::kStagingHalf) {
        // switch halves; wait for the copies still reading the other one
        s.half ^= 1;
        if (s.pending[s.half]) LCT_CHECK_HIP(hipEventSynchronize(s.ev[s.half]));
        s.pending[s.half] = false;
        off = 0;
    }
    char *p = s.base + size_t(s.half) * lct_ctx::kStagingHalf + off;
    std::memcpy(p, h, bytes);
    s.used = off + bytes;
    // the half's event must follow every copy out of it: a copy on another
    // stream than the half's previous one is ordered after that one first
    if (s.pending[s.half] && s.ev_stream[s.half] != ctx->stream)
        LCT_CHECK_HIP(hipStreamWaitEvent(ctx->stream, s.ev[s.half], 0));
    LCT_CHECK_HIP(hipMemcpyAsync(d, p, bytes, hipMemcpyHostToDevice, ctx->stream));
    LCT_CHECK_HIP(hipEventRecord(s.ev[s.half], ctx->stream));
    s.pending[s.half] = true;
    s.ev_stream[s.half] = ctx->stream;
    return LCT_OK;
}

// RampFilterGPU::calculate_filter, Src/Filter.cpp:10-19 (double arithmetic,
// stored as float), with the apodisation windows this build defines for the
// empty Include/SmoothWindow.h.  tau = ds.
static double ramlak(long n, double ds)
{
    if (n == 0) return 0.25 / ds;
    double nn = double(n);
    return (std::sin(nn * LCT_PI_D) / (nn * LCT_PI_D) +
            (std::cos(nn * LCT_PI_D) - 1.0) / (nn * nn * LCT_PI_D * LCT_PI_D)) / (2.0 * ds);
}

void host_taps(int window, int ns, float ds_f, float *h)
{
    const double ds = ds_f;
    for (int i = 0; i < 2 * ns; i++) {
        long n = long(i) - ns;
        double v;
        if (window == LCT_WINDOW_HANN) {
            // |f| (0.5 + 0.5 cos(2 pi f tau)): the ramp taps smoothed by (1/4, 1/2, 1/4).
            v = 0.5 * ramlak(n, ds) + 0.25 * (ramlak(n - 1, ds) + ramlak(n + 1, ds));
        } else if (window == LCT_WINDOW_SHEPP_LOGAN) {
            // |f| sinc(f tau)  ->  -2 / (pi^2 tau (4 n^2 - 1)).
            double nn = double(n);
            v = -2.0 / (LCT_PI_D * LCT_PI_D * ds * (4.0 * nn * nn - 1.0));
        } else {
            v = ramlak(n, ds);
        }
        h[i] = float(v);
    }
}

int aux_init(lct_ctx *ctx)
{
    if (ctx->aux_ready) return LCT_OK;
    for (auto &s : ctx->aux)
        if (!s) LCT_CHECK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    for (auto &e : ctx->sync_ev)
        if (!e) LCT_CHECK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    ctx->aux_ready = true;
    return LCT_OK;
}

}  // namespace lct

using namespace lct;

extern "C" {

const char *lct_last_error(void) { return g_last_error.c_str(); }

const char *lct_version(void) { return "lct 1.0 gfx950"; }

int lct_ctx_create(int device, lct_ctx **out)
{
    if (!out) return set_error(LCT_ERR_INVALID, "lct_ctx_create: out is NULL");
    *out = nullptr;
    int n = 0;
    LCT_CHECK_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n)
        return set_error(LCT_ERR_INVALID, "lct_ctx_create: device " + std::to_string(device) +
                                              " not present (" + std::to_string(n) + " devices)");
    LCT_CHECK_HIP(hipSetDevice(device));
    LCT_CHECK(noalloc_scratch_init(device));
    lct_ctx *c = new lct_ctx();
    c->device = device;
    c->stream = nullptr;
    for (auto &e : c->ev) {
        hipError_t err = hipEventCreate(&e);
        if (err != hipSuccess) {
            delete c;
            return hip_error(err, "hipEventCreate");
        }
    }
    c->events_ready = true;
    *out = c;
    return LCT_OK;
}

int lct_ctx_destroy(lct_ctx *ctx)
{
    if (!ctx) return LCT_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &s : ctx->aux)
        if (s) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
    for (auto &e : ctx->sync_ev)
        if (e) (void)hipEventDestroy(e);
    staging_release(ctx);
    for (auto &b : ctx->scratch)
        if (b.ptr) (void)hipFree(b.ptr);
    if (ctx->events_ready)
        for (auto &e : ctx->ev) (void)hipEventDestroy(e);
    delete ctx;
    return LCT_OK;
}

int lct_ctx_set_stream(lct_ctx *ctx, void *stream)
{
    if (!ctx) return set_error(LCT_ERR_INVALID, "lct_ctx_set_stream: ctx is NULL");
    ctx->stream = static_cast<hipStream_t>(stream);
    return LCT_OK;
}

int lct_synchronize(lct_ctx *ctx)
{
    if (!ctx) return set_error(LCT_ERR_INVALID, "lct_synchronize: ctx is NULL");
    LCT_CHECK_HIP(hipSetDevice(ctx->device));
    LCT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    return LCT_OK;
}

int lct_set_profiling(lct_ctx *ctx, int enable)
{
    if (!ctx) return set_error(LCT_ERR_INVALID, "lct_set_profiling: ctx is NULL");
    ctx->profiling = enable != 0;
    return LCT_OK;
}

int lct_get_timings(lct_ctx *ctx, lct_timings *out)
{
    if (!ctx || !out) return set_error(LCT_ERR_INVALID, "lct_get_timings: NULL argument");
    *out = ctx->timings;
    return LCT_OK;
}

// GPUMemoryCtrl<T>::allocate_memory / free_memory / copy_*, Include/GPUMemoryCtrl.h:22-59.
int lct_device_malloc(lct_ctx *ctx, void **d_ptr, size_t bytes)
{
    if (!ctx || !d_ptr) return set_error(LCT_ERR_INVALID, "lct_device_malloc: NULL argument");
    *d_ptr = nullptr;
    LCT_CHECK_HIP(hipSetDevice(ctx->device));
    hipError_t e = hipMalloc(d_ptr, bytes ? bytes : 1);
    if (e != hipSuccess) {
        *d_ptr = nullptr;
        (void)hipGetLastError();
        return set_error(LCT_ERR_NOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
    }
    return LCT_OK;
}

int lct_device_free(lct_ctx *ctx, void *d_ptr)
{
    if (!ctx) return set_error(LCT_ERR_INVALID, "lct_device_free: ctx is NULL");
    if (!d_ptr) return LCT_OK;
    LCT_CHECK_HIP(hipSetDevice(ctx->device));
    LCT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    LCT_CHECK_HIP(hipFree(d_ptr));
    return LCT_OK;
}

static int copy(lct_ctx *ctx, void *dst, const void *src, size_t bytes, hipMemcpyKind kind)
{
    if (!ctx) return set_error(LCT_ERR_INVALID, "lct_memcpy: ctx is NULL");
    if (bytes == 0) return LCT_OK;
    if (!dst || !src) return set_error(LCT_ERR_INVALID, "lct_memcpy: NULL pointer");
    LCT_CHECK_HIP(hipSetDevice(ctx->device));
    LCT_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, kind, ctx->stream));
    if (kind != hipMemcpyDeviceToDevice) LCT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    return LCT_OK;
}

int lct_memcpy_h2d(lct_ctx *ctx, void *d_dst, const void *h_src, size_t bytes)
{
    return copy(ctx, d_dst, h_src, bytes, hipMemcpyHostToDevice);
}

int lct_memcpy_d2h(lct_ctx *ctx, void *h_dst, const void *d_src, size_t bytes)
{
    return copy(ctx, h_dst, d_src, bytes, hipMemcpyDeviceToHost);
}

int lct_memcpy_d2d(lct_ctx *ctx, void *d_dst, const void *d_src, size_t bytes)
{
    return copy(ctx, d_dst, d_src, bytes, hipMemcpyDeviceToDevice);
}

int lct_memset_zero(lct_ctx *ctx, void *d_dst, size_t bytes)
{
    if (!ctx) return set_error(LCT_ERR_INVALID, "lct_memset_zero: ctx is NULL");
    if (bytes == 0) return LCT_OK;
    LCT_CHECK_HIP(hipSetDevice(ctx->device));
    LCT_CHECK_HIP(hipMemsetAsync(d_dst, 0, bytes, ctx->stream));
    return LCT_OK;
}

int lct_ramp_taps(int window, int ns, float ds, float *h_taps)
{
    if (ns <= 0 || !(ds > 0.0f) || !h_taps || window < 0 || window > 2)
        return set_error(LCT_ERR_INVALID, "lct_ramp_taps: bad arguments");
    host_taps(window, ns, ds, h_taps);
    return LCT_OK;
}

}  // extern "C"
