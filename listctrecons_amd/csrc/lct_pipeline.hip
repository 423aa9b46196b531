// lct_pipeline.hip -- FDKReconstructor::reconstruct (Src/Reconstructor.cpp:30-49)
// over the whole planned range, with the projection file resident in HBM.
//
// LCT_PIPE_BLOCKS keeps the reference's structure: every 32-slice block
// re-rebins and re-filters its own window of n_angle_add views
// (Src/DataStructs.cpp:222-240, :77-108) and is backprojected voxel by voxel.
// LCT_PIPE_RESIDENT rebins and filters every view once into an HBM-resident
// filtered-projection array and backprojects all blocks from it
// (lct_resident.hip).
#include <algorithm>
#include <cmath>
#include <vector>

#include "lct_internal.hpp"

namespace lct {

int reconstruct_resident(lct_ctx *ctx, const lct_plan &plan, const lct_recon_opts &o,
                         const float *h_angles, const float *h_couch, const float *d_fan,
                         long file_views, int b0, int b1, float *d_vol);

// Stage timer: records events on the context stream when profiling is on.
struct StageClock {
    lct_ctx *ctx;
    hipEvent_t a, b;
    bool on;
    explicit StageClock(lct_ctx *c) : ctx(c), a(c->ev[6]), b(c->ev[7]), on(c->profiling) {}
    int start()
    {
        if (on) LCT_CHECK_HIP(hipEventRecord(a, ctx->stream));
        return LCT_OK;
    }
    int stop(float *acc)
    {
        if (!on) return LCT_OK;
        LCT_CHECK_HIP(hipEventRecord(b, ctx->stream));
        LCT_CHECK_HIP(hipEventSynchronize(b));
        float ms = 0.0f;
        LCT_CHECK_HIP(hipEventElapsedTime(&ms, a, b));
        *acc += ms;
        return LCT_OK;
    }
};

long long count_updates(const lct_plan &plan, int bp_kind)
{
    if (bp_kind == LCT_BP_PI_ORI) {
        // half a turn of views per voxel (a_min + 1 .. a_min + npt / 2,
        // Backprojector.cu:200-245) when its PI window lies inside the block's views
        return (long long)(plan.ctg.n_proj_turn / 2) * LCT_BLOCK_SLICES * plan.ig.nx * plan.ig.ny *
               plan.n_blocks;
    }
    // one update per view in each voxel's window (a_max - a_min per slice)
    int amin[32], amax[32], amid[32];
    bp_windows(plan.ctg, plan.ig, amin, amax, amid);
    long long per_column = 0;
    for (int iz = 0; iz < 32; iz++) per_column += amax[iz] - amin[iz];
    return per_column * plan.ig.nx * plan.ig.ny * plan.n_blocks;
}

// Blocks [b0, b1); block b's slab lands at d_vol + (b - b0) slabs.
static int reconstruct_blocks(lct_ctx *ctx, const lct_plan &plan, const lct_recon_opts &o,
                              const float *h_angles, const float *h_couch, const float *d_fan,
                              long file_views, int b0, int b1, float *d_vol)
{
    const lct_ct_geom &g = plan.ctg;
    const lct_img_geom &ig = plan.ig;
    const size_t view = size_t(g.nt) * g.ns;
    const size_t slab = size_t(ig.nx) * ig.ny * ig.nz;

    std::vector<float> taps(2 * size_t(g.ns));
    host_taps(o.window, g.ns, g.ds, taps.data());
    void *d_taps = nullptr, *d_par = nullptr;
    LCT_CHECK(scratch(ctx, kScratchTaps, taps.size() * 4, &d_taps));
    LCT_CHECK(upload(ctx, d_taps, taps.data(), taps.size() * 4));
    LCT_CHECK(scratch(ctx, kScratchA, view * g.n_angle * 4, &d_par));
    float *par = static_cast<float *>(d_par);

    const int nb = g.n_angle + g.n_proj_turn / 2 + 1;
    std::vector<float> betas(nb);
    StageClock clk(ctx);
    lct_timings &t = ctx->timings;
    for (int b = b0; b < b1; b++) {
        int bps = 0;
        LCT_CHECK(lct_block_info(&plan, h_couch, b, &bps, nullptr, nullptr, nullptr, nullptr));
        const long first = long(bps) + plan.data_offset;
        if (bps < 0 || first < 0 || first + g.n_angle_add > file_views ||
            bps + g.add_projections + g.n_angle > plan.n_readings)
            return set_error(LCT_ERR_RANGE, "block " + std::to_string(b) +
                                                 " projection window lies outside the data");
        // RawDataReader (Src/DataReader.cpp:8-19) + Rebiner: the block window
        // starts at file view `first`; texels outside its n_angle_add views are 0.
        LCT_CHECK(clk.start());
        LCT_CHECK(launch_rebin(ctx, o.rebin_kind, o.weight_kind, &g, d_fan, first, g.n_angle_add,
                               par, g.n_angle, g.add_projections, 0));
        LCT_CHECK(clk.stop(&t.rebin_ms));
        LCT_CHECK(clk.start());
        LCT_CHECK(launch_filter_direct(ctx, g.ns, static_cast<const float *>(d_taps), par,
                                       long(g.n_angle) * g.nt));
        LCT_CHECK(clk.stop(&t.filter_ms));
        // betas = tubeAngles[bps + add ...] (Backprojector.cu:145)
        for (int i = 0; i < nb; i++) {
            const long src = long(bps) + g.add_projections + i;
            betas[i] = src < plan.n_readings ? h_angles[src] : 0.0f;
        }
        LCT_CHECK(clk.start());
        LCT_CHECK(backproject_block_public(ctx, o.bp_kind, g, ig, par, betas.data(),
                                           d_vol + size_t(b - b0) * slab));
        LCT_CHECK(clk.stop(&t.backproject_ms));
        t.backproject_kernel_ms = t.backproject_ms;
        t.n_backproject_launches++;
    }
    // the last block's filtered views stay in `par` (CTDataCPU::rebin of the
    // reference after Filter.cu:69)
    ctx->last = lct_ctx::LastBlock{1, LCT_F32, g.n_angle, g.nt, g.ns, 0, 0, 0, par};
    return LCT_OK;
}

}  // namespace lct

using namespace lct;

// Blocks [b0, b1) of the plan; d_fan[0] is projection-file view fan_view0.
static int reconstruct_range(lct_ctx *ctx, const lct_plan *plan, const lct_recon_opts *opts,
                             const float *h_angles, const float *h_couch, const float *d_fan,
                             long fan_view0, long fan_views, int b0, int b1, float *d_vol,
                             const char *who)
{
    if (!ctx || !plan || !opts || !h_angles || !h_couch || !d_fan || !d_vol || fan_views <= 0 ||
        fan_view0 < 0)
        return set_error(LCT_ERR_INVALID, std::string(who) + ": NULL argument");
    const lct_recon_opts &o = *opts;
    if (o.rebin_kind < 0 || o.rebin_kind > 2 || o.weight_kind < 0 || o.weight_kind > 1 ||
        o.window < 0 || o.window > 2 || o.bp_kind < 0 || o.bp_kind > 1 || o.storage < 0 ||
        o.storage > 2 || o.filter_algo < 0 || o.filter_algo > 2 || o.interp < 0 || o.interp > 1)
        return set_error(LCT_ERR_INVALID, std::string(who) + ": option out of range");
    if (plan->n_blocks <= 0 || plan->ig.nz != LCT_BLOCK_SLICES || plan->ctg.n_angle <= 0)
        return set_error(LCT_ERR_INVALID, std::string(who) + ": plan has no blocks");
    if (b0 < 0 || b1 <= b0 || b1 > plan->n_blocks)
        return set_error(LCT_ERR_INVALID, std::string(who) + ": block range outside the plan");
    // index the device buffer from its first view
    lct_plan p = *plan;
    p.data_offset = int(long(plan->data_offset) - fan_view0);
    LCT_CHECK_HIP(hipSetDevice(ctx->device));
    ctx->timings = lct_timings{};
    ctx->last = lct_ctx::LastBlock{};
    hipEvent_t t0 = ctx->ev[4], t1 = ctx->ev[5];
    if (ctx->profiling) LCT_CHECK_HIP(hipEventRecord(t0, ctx->stream));
    int rc;
    if (o.pipeline == LCT_PIPE_RESIDENT)
        rc = reconstruct_resident(ctx, p, o, h_angles, h_couch, d_fan, fan_views, b0, b1, d_vol);
    else if (o.pipeline == LCT_PIPE_BLOCKS)
        rc = reconstruct_blocks(ctx, p, o, h_angles, h_couch, d_fan, fan_views, b0, b1, d_vol);
    else
        rc = set_error(LCT_ERR_INVALID, std::string(who) + ": unknown pipeline");
    if (rc != LCT_OK) return rc;
    ctx->timings.backproject_updates = count_updates(p, o.bp_kind) / p.n_blocks * (b1 - b0);
    if (ctx->profiling) {
        LCT_CHECK_HIP(hipEventRecord(t1, ctx->stream));
        LCT_CHECK_HIP(hipEventSynchronize(t1));
        LCT_CHECK_HIP(hipEventElapsedTime(&ctx->timings.total_ms, t0, t1));
    }
    return LCT_OK;
}

extern "C" int lct_reconstruct(lct_ctx *ctx, const lct_plan *plan, const lct_recon_opts *opts,
                               const float *h_angles, const float *h_couch, const float *d_fan,
                               long file_views, float *d_vol)
{
    if (!plan) return set_error(LCT_ERR_INVALID, "lct_reconstruct: NULL argument");
    return reconstruct_range(ctx, plan, opts, h_angles, h_couch, d_fan, 0, file_views, 0,
                             plan->n_blocks, d_vol, "lct_reconstruct");
}

extern "C" int lct_reconstruct_slabs(lct_ctx *ctx, const lct_plan *plan,
                                     const lct_recon_opts *opts, const float *h_angles,
                                     const float *h_couch, const float *d_fan, long fan_view0,
                                     long fan_views, int block_begin, int block_end, float *d_vol)
{
    return reconstruct_range(ctx, plan, opts, h_angles, h_couch, d_fan, fan_view0, fan_views,
                             block_begin, block_end, d_vol, "lct_reconstruct_slabs");
}
