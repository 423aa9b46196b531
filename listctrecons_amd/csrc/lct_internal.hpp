// lct_internal.hpp -- shared internals of liblct.so (context, errors,
// device scratch, launch helpers).  Not installed; include/lct.h is the ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "lct.h"

#define LCT_PI_D 3.141592653589793238462643383  // Include/Common.h:3
#define LCT_BLOCK_SLICES 32                     // Include/Common.h:4

struct lct_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool profiling = false;
    lct_timings timings{};

    // Scratch owned by the context, grown on demand and reused across calls.
    struct Buf {
        void *ptr = nullptr;
        size_t bytes = 0;
    };
    std::vector<Buf> scratch = std::vector<Buf>(9);

    hipEvent_t ev[8] = {};
    bool events_ready = false;

    // Streams and events of the overlapped resident pipeline (later views
    // rebinned and filtered while the blocks of earlier ones are
    // backprojected): non-blocking, default priority, created on first use.
    hipStream_t aux[2] = {nullptr, nullptr};
    hipEvent_t sync_ev[12] = {};
    bool aux_ready = false;

    // Pinned staging ring for host tables (two halves of kStagingHalf bytes):
    // an upload copies the table into the current half and enqueues an
    // asynchronous copy from there, so the host never waits for the device.
    // ev[k] follows the last copy out of half k; a half is refilled only once
    // that event has completed.
    static constexpr size_t kStagingHalf = size_t(8) << 20;
    struct Staging {
        char *base = nullptr;
        size_t used = 0;
        int half = 0;
        bool pending[2] = {false, false};
        hipEvent_t ev[2] = {};
        hipStream_t ev_stream[2] = {};  // stream of the copy ev[k] follows
        bool failed = false;  // pinned allocation refused: uploads wait instead
    } staging;

    // Where the last reconstruct call left its last block's filtered views
    // (scratch slots A / B, untouched until the next reconstruct call):
    // layout 1: rows [view][nt][ns] f32; layout 2: Q [view][cp][rp] of
    // `storage`; 0: nothing.  The block's views start at view0.
    // What the resident Q buffer was last cleared for: its guard slots are
    // zero while buffer, layout and extent are unchanged.
    struct QZero {
        const void *ptr = nullptr;
        int cp = 0, rp = 0;
        size_t esz = 0, bytes = 0;
    } qzero;

    struct LastBlock {
        int layout = 0, storage = 0, n_angle = 0, nt = 0, ns = 0, cp = 0, rp = 0;
        long view0 = 0;
        const void *ptr = nullptr;
    } last;
};

namespace lct {

// Scratch slots of lct_ctx::scratch.
enum ScratchSlot {
    kScratchA = 0,  // rebinned views
    kScratchB,      // filtered views
    kScratchTables, // per-view / per-channel geometry tables
    kScratchTaps,   // filter taps / Toeplitz operands
    kScratchImg,    // one block image (per-block pipeline)
    kScratchFan,    // fan window (per-block pipeline)
    kScratchMisc0,
    kScratchMisc1,
    kScratchPart,   // second partial sums of the backprojector's split pair ranges
};

int set_error(int code, const std::string &msg);
int hip_error(hipError_t e, const char *what);
void clear_error();

// Returns a device buffer of at least `bytes` in slot `slot` (grown, never shrunk).
int scratch(lct_ctx *ctx, int slot, size_t bytes, void **out);
int noalloc_scratch_init(int device);
// Creates ctx->aux and ctx->sync_ev on first use.
int aux_init(lct_ctx *ctx);

// Host-to-device copy on the context stream, ordered after earlier work on it;
// returns once `h` has been read (into the pinned staging ring), without
// waiting for the device.
int upload(lct_ctx *ctx, void *d, const void *h, size_t bytes);
// Waits for the staging ring's copies and frees it (context teardown).
void staging_release(lct_ctx *ctx);

// Kernel-side launchers (defined in the .hip files).
int launch_rebin(lct_ctx *ctx, int kind, int weight, const lct_ct_geom *g, const float *d_fan,
                 long fan_view0, long fan_views, float *d_par, long n_out, long out_view0,
                 int transposed);
int launch_weighting(lct_ctx *ctx, const lct_ct_geom *g, float *d_par);
int launch_filter_direct(lct_ctx *ctx, int ns, const float *d_taps, float *d_rows, long rows);
// Toeplitz-GEMM ramp filter on BF16 MFMA (algo LCT_FILTER_TOEPLITZ_BF16X3 or
// _BF16X6) of rows [n_cols][ns] f32 into the backprojector layout
// Q[view][cp][rp] of the given storage type.
int launch_filter_toeplitz(lct_ctx *ctx, int algo, int storage, int ns, int nt, const float *h_taps,
                           const float *d_in, void *d_q, long n_cols, int cp, int rp);
// Literal per-voxel backprojection of one 32-slice block (tables uploaded here).
int backproject_block_public(lct_ctx *ctx, int kind, const lct_ct_geom &g,
                             const lct_img_geom &ig, const float *d_q, const float *h_betas,
                             float *d_img);

// Host geometry tables evaluated in double from the float32 geometry.
void host_taps(int window, int ns, float ds, float *h);
// a_min / a_max / a_middle of each slice of a block (float32, as the reference).
void bp_windows(const lct_ct_geom &g, const lct_img_geom &ig, int *amin, int *amax, int *amid);
// cos (first n) and sin (next n) of float32 angles evaluated in double.
void angle_tables(const float *betas, int n, std::vector<float> &cs);

}  // namespace lct

#define LCT_CHECK_HIP(call)                                   \
    do {                                                      \
        hipError_t e__ = (call);                              \
        if (e__ != hipSuccess) return lct::hip_error(e__, #call); \
    } while (0)

#define LCT_CHECK(call)               \
    do {                              \
        int rc__ = (call);            \
        if (rc__ != LCT_OK) return rc__; \
    } while (0)
