// lct_plan.cpp -- the helical scan planner of ReconData (host only).
//
// Restates ReconData::convert_from_file_params (Src/DataStructs.cpp:121-250)
// and BlockInfo::update (:77-108).  The planner's decisions (guard views,
// allowed range, slice and block counts, each block's projection window) are
// integer and float32 decisions in the reference; they are evaluated here in
// the same float32 operations and order so every block reads the same views.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "lct.h"

namespace lct {
int set_error(int code, const std::string &msg);
}

namespace {

constexpr double kPi = 3.141592653589793238462643383;  // Include/Common.h:3
constexpr int kBlockSlices = 32;                        // Include/Common.h:4

// array_search (Src/DataStructs.cpp:12-34); the index bound is tested before
// the element is read, so a key beyond the data stops at n instead of reading
// array[n].
int search_views(float key, const float *a, int n, int direction)
{
    int i = 0;
    if (direction > 0) {
        while (i < n && key > a[i]) ++i;
    } else if (direction < 0) {
        while (i < n && key < a[i]) ++i;
    } else {
        while (i < n && key != a[i]) ++i;
    }
    return i;
}

// int(fabs(d)/d) with the reference's fallback to +1 (one-slice request).
int sign_or_one(float from, float to)
{
    float d = to - from;
    if (!(d != 0.0f)) return 1;
    int r = int(std::fabs(d) / d);
    return (r == 1 || r == -1) ? r : 1;
}

}  // namespace

extern "C" int lct_plan_helical(const lct_prm *prm, const float *ang, const float *couch,
                                lct_plan *out)
{
    using lct::set_error;
    if (!prm || !ang || !couch || !out)
        return set_error(LCT_ERR_INVALID, "lct_plan_helical: NULL argument");
    std::memset(out, 0, sizeof(*out));
    const lct_prm &p = *prm;
    if (p.readings <= 100)
        return set_error(LCT_ERR_INVALID, "lct_plan_helical: Readings must exceed 100 "
                                          "(the table direction is read from view 100)");
    if (p.ns <= 0 || p.nt <= 0 || p.n_proj_turn <= 1 || !(p.dz > 0.0f) || !(p.ds > 0.0f) ||
        !(p.dt > 0.0f) || !(p.r_src_to_iso > 0.0f) || p.nx <= 0 || p.ny <= 0)
        return set_error(LCT_ERR_INVALID, "lct_plan_helical: non-positive geometry value");

    // Guard views (:123): half the fan in views, plus 10.
    const float half_fan = std::max(p.central_channel, float(p.ns) - p.central_channel);
    const float fan_views = (half_fan * p.ds) * float(p.n_proj_turn);
    const int add = int(double(fan_views) / (double(2.0f * p.r_src_to_iso) * kPi) + 10);

    // Table direction and per-view orbit step (:169-175).
    const float d100 = couch[100] - couch[0];
    if (d100 == 0.0f)
        return set_error(LCT_ERR_INVALID, "lct_plan_helical: couch positions do not move");
    const int z_dir = int(std::fabs(d100) / d100);
    const bool wraps = (ang[1] < 1 && ang[0] > 6) || (ang[1] > 6 && ang[0] < 1);
    const float orbit = wraps ? ang[2] - ang[1] : ang[1] - ang[0];

    // Reconstructable range (:179-195).
    const float data_begin = couch[0];
    const float data_end = couch[p.readings - 1];
    const float padding =
        p.pitch_value * float(p.n_proj_turn / 2 + add + 256) / float(p.n_proj_turn);
    const float allowed_begin = data_begin + float(z_dir) * padding;
    const float allowed_end = data_end - float(z_dir) * padding;
    float start_pos, end_pos;
    if (z_dir > 0) {
        start_pos = std::ceil(allowed_begin + float(z_dir));
        end_pos = std::floor(allowed_end - float(z_dir));
    } else {
        start_pos = std::floor(allowed_begin + float(z_dir));
        end_pos = std::ceil(allowed_end - float(z_dir));
    }
    const int rd = sign_or_one(start_pos, end_pos);

    // Slice and block counts (:200-208): whole range, rounded down to 32 slices.
    const float recon_start = start_pos - float(rd) * p.dz;
    float recon_end = end_pos + float(rd) * p.dz;
    const int n_requested = int(std::floor(std::fabs(recon_end - recon_start) / p.dz) + 1.0f);
    const int n_recon = (n_requested - 1) - (n_requested - 1) % kBlockSlices;
    recon_end = recon_start + float(rd * (n_recon - 1)) * p.dz;
    const int n_blocks = n_recon / kBlockSlices;

    // (:210-218) the reference prints and exits.
    auto outside = [&](float v) {
        return (v > allowed_begin && v > allowed_end) || (v < allowed_begin && v < allowed_end);
    };
    if (outside(start_pos) || outside(end_pos)) {
        char msg[160];
        std::snprintf(msg, sizeof(msg),
                      "Requested reconstruction is outside of allowed data range: %.2f to %.2f",
                      allowed_begin, allowed_end);
        return set_error(LCT_ERR_RANGE, msg);
    }
    if (n_blocks <= 0)
        return set_error(LCT_ERR_RANGE, "lct_plan_helical: fewer than 32 reconstructable slices");

    // First block's projection window (:222-240).
    const float block_start = recon_start;
    const float block_end = block_start + float(rd * (kBlockSlices - 1)) * p.dz;
    int first = search_views(block_start, couch, p.readings, z_dir);
    int last = search_views(block_end, couch, p.readings, z_dir);
    const int half_turn = p.n_proj_turn / 2;
    if (first > last) {
        first = last - half_turn - add;
        last = first + half_turn + add;
    } else {
        first = first - half_turn - add;
        last = last + half_turn + add;
    }
    const int n_angle_add = last - first + 100;
    const int n_angle = n_angle_add - 2 * add;

    lct_ct_geom &g = out->ctg;  // CTGeom(...) at :242
    g.orbit = orbit;
    g.orbit_start = ang[0];
    g.ns = p.ns;
    g.nt = p.nt;
    g.ws = p.central_channel;
    g.wt = float(p.nt - 1) / 2.0f;
    g.n_angle = n_angle;
    g.n_angle_add = n_angle_add;
    g.ds = p.ds;
    g.dt = p.dt;
    g.pitch = p.pitch_value;
    g.source_z0 = -p.pitch_value / 2.0f;
    g.dso = p.r_src_to_iso;
    g.n_proj_turn = p.n_proj_turn;
    g.z_dir = z_dir;
    g.add_projections = add;

    lct_img_geom &ig = out->ig;  // ImgGeom(...) at :243 (dy = Dx, wz = 0)
    ig.nx = p.nx;
    ig.ny = p.ny;
    ig.nz = kBlockSlices;
    ig.dx = p.dx;
    ig.dy = p.dx;
    ig.dz = p.dz;
    ig.wx = float(p.nx - 1) / 2.0f + p.offset_x;
    ig.wy = float(p.ny - 1) / 2.0f + p.offset_y;
    ig.wz = 0.0f;

    out->n_readings = p.readings;
    out->data_begin_pos = data_begin;
    out->data_end_pos = data_end;
    out->allow_begin = allowed_begin;
    out->allow_end = allowed_end;
    out->start_pos = start_pos;
    out->end_pos = end_pos;
    out->recon_start_pos = recon_start;
    out->recon_end_pos = recon_end;
    out->n_slices_recon = n_recon;
    out->n_slices_block = kBlockSlices;
    out->n_blocks = n_blocks;
    out->data_offset = p.data_offset;
    out->recon_direction = rd;
    return LCT_OK;
}

extern "C" int lct_block_info(const lct_plan *plan, const float *couch, int block,
                              int *proj_start, int *proj_end, float *block_start,
                              float *block_end, int *ref_view)
{
    using lct::set_error;
    if (!plan || !couch || block < 0)
        return set_error(LCT_ERR_INVALID, "lct_block_info: bad argument");
    const lct_ct_geom &g = plan->ctg;
    const int rd = sign_or_one(plan->start_pos, plan->end_pos);
    const float bs = plan->recon_start_pos +
                     float(rd * block) * plan->ig.dz * float(plan->n_slices_block);
    const float be = bs + float(rd * (plan->n_slices_block - 1)) * plan->ig.dz;
    const int s0 = search_views(bs, couch, plan->n_readings, g.z_dir);
    const int e0 = search_views(be, couch, plan->n_readings, g.z_dir);
    const int half_turn = g.n_proj_turn / 2;
    int s, e;
    if (block == 0) {
        if (s0 > e0) { s = e0 - half_turn - g.add_projections; e = s + half_turn + g.add_projections; }
        else { s = s0 - half_turn - g.add_projections; e = e0 + half_turn + g.add_projections; }
    } else {
        s = (s0 > e0 ? e0 : s0) - half_turn - g.add_projections;
        e = s + g.n_angle_add;
    }
    if (proj_start) *proj_start = s;
    if (proj_end) *proj_end = e;
    if (block_start) *block_start = bs;
    if (block_end) *block_end = be;
    if (ref_view) *ref_view = s + half_turn + g.add_projections;
    return LCT_OK;
}

extern "C" int lct_slab_views(const lct_plan *plan, const float *couch, int block_begin,
                              int block_end, long *view_begin, long *view_end)
{
    using lct::set_error;
    if (!plan || !couch || !view_begin || !view_end || block_begin < 0 ||
        block_end <= block_begin || block_end > plan->n_blocks)
        return set_error(LCT_ERR_INVALID, "lct_slab_views: bad argument");
    long lo = 0, hi = 0;
    for (int b = block_begin; b < block_end; b++) {
        int s = 0;
        const int rc = lct_block_info(plan, couch, b, &s, nullptr, nullptr, nullptr, nullptr);
        if (rc != LCT_OK) return rc;
        const long first = long(s) + plan->data_offset;
        const long last = first + plan->ctg.n_angle_add;
        lo = b == block_begin ? first : std::min(lo, first);
        hi = b == block_begin ? last : std::max(hi, last);
    }
    *view_begin = lo;
    *view_end = hi;
    return LCT_OK;
}
