// lct_forward.hip -- ray-driven forward projector of a voxel volume for the
// helical curved-detector geometry (the same rays as lct_project_ellipsoids).
// It produces reprojections of reconstructed volumes, the acceptance test of
// the learned correction (a corrected volume must reproject at least as close
// to the measured data as plain FDK).  Not part of the reference.
//
// Ray of view v, row it, channel is: source S = (-dso sin b, dso cos b, z_v),
// in-plane direction angle b - (is - ws) ds / dso, slope (it - wt) dt / dso.
// The volume [nz][ny][nx] has voxel centres x = (ix - wx) dx,
// y = (iy - wy) dy, z = z0 + iz dz, and is sampled trilinearly (zero outside)
// at the midpoints of n equal steps of at most `step` mm between the ray's
// entry into and exit from the voxel-centre box.
#include <cmath>
#include <vector>

#include "lct_internal.hpp"

namespace lct {

struct FwdArgs {
    const float *vol;
    int nx, ny, nz;
    float wx, wy, dx, dy, dz, z0;
    int ns, nt;
    float ws, wt, ds_dso, dt_dso, dso, step;
    const float *views;  // per view: sin b, cos b, z, b
    float *out;
    long n_views;
};

__device__ __forceinline__ float sample3(const FwdArgs &A, float fx, float fy, float fz)
{
    // fx, fy, fz in voxel-index units
    const float x0f = floorf(fx), y0f = floorf(fy), z0f = floorf(fz);
    const int x0 = int(x0f), y0 = int(y0f), z0 = int(z0f);
    const float ax = fx - x0f, ay = fy - y0f, az = fz - z0f;
    float acc = 0.0f;
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const int zz = z0 + k;
        if (zz < 0 || zz >= A.nz) continue;
        const float wz = k ? az : 1.0f - az;
#pragma unroll
        for (int jj = 0; jj < 2; jj++) {
            const int yy = y0 + jj;
            if (yy < 0 || yy >= A.ny) continue;
            const float wy = jj ? ay : 1.0f - ay;
            const float *row = A.vol + (size_t(zz) * A.ny + yy) * A.nx;
            const float v0 = (x0 >= 0 && x0 < A.nx) ? row[x0] : 0.0f;
            const float v1 = (x0 + 1 >= 0 && x0 + 1 < A.nx) ? row[x0 + 1] : 0.0f;
            acc += wz * wy * fmaf(ax, v1 - v0, v0);
        }
    }
    return acc;
}

__global__ __launch_bounds__(256) void forward_project_kernel(FwdArgs A)
{
    const int is = blockIdx.x * blockDim.x + threadIdx.x;
    const int it = blockIdx.y;
    const long v = blockIdx.z;
    if (is >= A.ns || v >= A.n_views) return;
    const float sb = A.views[4 * v], cb = A.views[4 * v + 1], sz = A.views[4 * v + 2];
    const float b = A.views[4 * v + 3];
    const float a = (float(is) - A.ws) * A.ds_dso;
    const float slope = (float(it) - A.wt) * A.dt_dso;
    const float inv = rsqrtf(1.0f + slope * slope);
    float d[3] = {sinf(b - a) * inv, -cosf(b - a) * inv, slope * inv};
    const float S[3] = {-A.dso * sb, A.dso * cb, sz};
    // box of voxel centres
    const float lo[3] = {-A.wx * A.dx, -A.wy * A.dy, A.z0};
    const float hi[3] = {(A.nx - 1 - A.wx) * A.dx, (A.ny - 1 - A.wy) * A.dy,
                         A.z0 + (A.nz - 1) * A.dz};
    float t0 = -1e30f, t1 = 1e30f;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        if (fabsf(d[k]) < 1e-12f) {
            if (S[k] < lo[k] || S[k] > hi[k]) t1 = -1e30f;
            continue;
        }
        float ta = (lo[k] - S[k]) / d[k], tb = (hi[k] - S[k]) / d[k];
        if (ta > tb) { const float tmp = ta; ta = tb; tb = tmp; }
        t0 = fmaxf(t0, ta);
        t1 = fminf(t1, tb);
    }
    float sum = 0.0f;
    if (t1 > t0) {
        const int n = int(ceilf((t1 - t0) / A.step));
        const float h = (t1 - t0) / float(n);
        for (int k = 0; k < n; k++) {
            const float t = t0 + (float(k) + 0.5f) * h;
            const float px = S[0] + t * d[0], py = S[1] + t * d[1], pz = S[2] + t * d[2];
            sum += sample3(A, px / A.dx + A.wx, py / A.dy + A.wy, (pz - A.z0) / A.dz);
        }
        sum *= h;
    }
    A.out[(size_t(v) * A.nt + it) * A.ns + is] = sum;
}

}  // namespace lct

using namespace lct;

extern "C" int lct_forward_project(lct_ctx *ctx, const lct_ct_geom *g, const float *h_angles,
                                   const float *h_couch, long view0, long n_views,
                                   const float *d_vol, const lct_img_geom *vg, float vol_z0,
                                   int vol_nz, float step, float *d_out)
{
    if (!ctx || !g || !h_angles || !h_couch || !d_vol || !vg || !d_out || n_views < 0 ||
        view0 < 0 || vol_nz <= 0 || vg->nx <= 0 || vg->ny <= 0 || !(step > 0.0f) ||
        !(g->dso > 0.0f))
        return set_error(LCT_ERR_INVALID, "lct_forward_project: bad argument");
    if (n_views == 0) return LCT_OK;
    LCT_CHECK_HIP(hipSetDevice(ctx->device));
    std::vector<float> vt(4 * size_t(n_views));
    for (long v = 0; v < n_views; v++) {
        const double b = h_angles[view0 + v];
        vt[4 * v] = float(std::sin(b));
        vt[4 * v + 1] = float(std::cos(b));
        vt[4 * v + 2] = h_couch[view0 + v];
        vt[4 * v + 3] = float(b);
    }
    void *dv = nullptr;
    LCT_CHECK(scratch(ctx, kScratchMisc0, vt.size() * 4, &dv));
    LCT_CHECK(upload(ctx, dv, vt.data(), vt.size() * 4));
    FwdArgs a;
    a.vol = d_vol;
    a.nx = vg->nx;
    a.ny = vg->ny;
    a.nz = vol_nz;
    a.wx = vg->wx;
    a.wy = vg->wy;
    a.dx = vg->dx;
    a.dy = vg->dy;
    a.dz = vg->dz;
    a.z0 = vol_z0;
    a.ns = g->ns;
    a.nt = g->nt;
    a.ws = g->ws;
    a.wt = g->wt;
    a.ds_dso = float(double(g->ds) / double(g->dso));
    a.dt_dso = float(double(g->dt) / double(g->dso));
    a.dso = g->dso;
    a.step = step;
    a.out = d_out;
    for (long v0 = 0; v0 < n_views; v0 += 65535) {
        a.views = static_cast<const float *>(dv) + 4 * v0;
        a.n_views = std::min<long>(65535, n_views - v0);
        a.out = d_out + size_t(v0) * g->nt * g->ns;
        const dim3 grid(unsigned((g->ns + 255) / 256), unsigned(g->nt), unsigned(a.n_views));
        hipLaunchKernelGGL(forward_project_kernel, grid, dim3(256), 0, ctx->stream, a);
        LCT_CHECK_HIP(hipGetLastError());
    }
    LCT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    return LCT_OK;
}
