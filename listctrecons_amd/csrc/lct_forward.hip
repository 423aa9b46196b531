// lct_forward.hip -- ray-driven forward projector of a voxel volume for the
// helical curved-detector geometry (the same rays as lct_project_ellipsoids),
// and its adjoint.  The projector produces reprojections of reconstructed
// volumes, the acceptance test of the learned correction (a corrected volume
// must reproject at least as close to the measured data as plain FDK).  Not
// part of the reference.
//
// Ray of view v, row it, channel is: source S = (-dso sin b, dso cos b, z_v),
// in-plane direction angle b - (is - ws) ds / dso, slope (it - wt) dt / dso.
// The volume [nz][ny][nx] has voxel centres x = (ix - wx) dx,
// y = (iy - wy) dy, z = z0 + iz dz, and is sampled trilinearly (zero outside)
// at the midpoints of n equal steps of at most `step` mm between the ray's
// entry into and exit from the voxel-centre box.
//
// The adjoint walks the same samples and scatters h * p with the same eight
// trilinear weights (float atomics), so <A x, y> = <x, A^T y> up to rounding.
#include <cmath>
#include <vector>

#include "lct_internal.hpp"

namespace lct {

struct FwdArgs {
    const float *vol;  // forward: the volume read
    float *vol_out;    // adjoint: the volume accumulated
    int nx, ny, nz;
    float wx, wy, dx, dy, dz, z0;
    int ns, nt;
    float ws, wt, ds_dso, dt_dso, dso, step;
    const float *views;  // per view: sin b, cos b, z, b
    const float *proj;   // adjoint: the sinogram read
    float *out;          // forward: the sinogram written
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

// Adjoint of sample3: val times each weight into the voxels sample3 reads.
__device__ __forceinline__ void scatter3(const FwdArgs &A, float fx, float fy, float fz, float val)
{
    const float x0f = floorf(fx), y0f = floorf(fy), z0f = floorf(fz);
    const int x0 = int(x0f), y0 = int(y0f), z0 = int(z0f);
    const float ax = fx - x0f, ay = fy - y0f, az = fz - z0f;
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const int zz = z0 + k;
        if (zz < 0 || zz >= A.nz) continue;
        const float wz = k ? az : 1.0f - az;
#pragma unroll
        for (int jj = 0; jj < 2; jj++) {
            const int yy = y0 + jj;
            if (yy < 0 || yy >= A.ny) continue;
            const float w = wz * (jj ? ay : 1.0f - ay) * val;
            float *row = A.vol_out + (size_t(zz) * A.ny + yy) * A.nx;
            if (x0 >= 0 && x0 < A.nx) atomicAdd(row + x0, w * (1.0f - ax));
            if (x0 + 1 >= 0 && x0 + 1 < A.nx) atomicAdd(row + x0 + 1, w * ax);
        }
    }
}

// Source S, unit direction d and chord [t0, t1] of ray (v, it, is) through the
// box of voxel centres; false when the ray misses the box.
__device__ __forceinline__ bool ray_chord(const FwdArgs &A, int is, int it, long v, float (&S)[3],
                                          float (&d)[3], float &t0, float &t1)
{
    const float sb = A.views[4 * v], cb = A.views[4 * v + 1], sz = A.views[4 * v + 2];
    const float b = A.views[4 * v + 3];
    const float a = (float(is) - A.ws) * A.ds_dso;
    const float slope = (float(it) - A.wt) * A.dt_dso;
    const float inv = rsqrtf(1.0f + slope * slope);
    d[0] = sinf(b - a) * inv;
    d[1] = -cosf(b - a) * inv;
    d[2] = slope * inv;
    S[0] = -A.dso * sb;
    S[1] = A.dso * cb;
    S[2] = sz;
    const float lo[3] = {-A.wx * A.dx, -A.wy * A.dy, A.z0};
    const float hi[3] = {(A.nx - 1 - A.wx) * A.dx, (A.ny - 1 - A.wy) * A.dy,
                         A.z0 + (A.nz - 1) * A.dz};
    t0 = -1e30f;
    t1 = 1e30f;
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
    return t1 > t0;
}

__global__ __launch_bounds__(256) void forward_project_kernel(FwdArgs A)
{
    const int is = blockIdx.x * blockDim.x + threadIdx.x;
    const int it = blockIdx.y;
    const long v = blockIdx.z;
    if (is >= A.ns || v >= A.n_views) return;
    float S[3], d[3], t0, t1;
    float sum = 0.0f;
    if (ray_chord(A, is, it, v, S, d, t0, t1)) {
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

__global__ __launch_bounds__(256) void forward_adjoint_kernel(FwdArgs A)
{
    const int is = blockIdx.x * blockDim.x + threadIdx.x;
    const int it = blockIdx.y;
    const long v = blockIdx.z;
    if (is >= A.ns || v >= A.n_views) return;
    float S[3], d[3], t0, t1;
    if (!ray_chord(A, is, it, v, S, d, t0, t1)) return;
    const int n = int(ceilf((t1 - t0) / A.step));
    const float h = (t1 - t0) / float(n);
    const float val = h * A.proj[(size_t(v) * A.nt + it) * A.ns + is];
    if (val == 0.0f) return;
    for (int k = 0; k < n; k++) {
        const float t = t0 + (float(k) + 0.5f) * h;
        const float px = S[0] + t * d[0], py = S[1] + t * d[1], pz = S[2] + t * d[2];
        scatter3(A, px / A.dx + A.wx, py / A.dy + A.wy, (pz - A.z0) / A.dz, val);
    }
}

// Shared set-up of both entry points: argument checks, the per-view table and
// the geometry fields of FwdArgs.
static int forward_setup(lct_ctx *ctx, const lct_ct_geom *g, const float *h_angles,
                         const float *h_couch, long view0, long n_views, const lct_img_geom *vg,
                         float vol_z0, int vol_nz, float step, const char *who, FwdArgs &a,
                         const float **d_views)
{
    if (!ctx || !g || !h_angles || !h_couch || !vg || n_views < 0 || view0 < 0 || vol_nz <= 0 ||
        vg->nx <= 0 || vg->ny <= 0 || !(step > 0.0f) || !(g->dso > 0.0f) || g->ns <= 0 ||
        g->nt <= 0)
        return set_error(LCT_ERR_INVALID, std::string(who) + ": bad argument");
    LCT_CHECK_HIP(hipSetDevice(ctx->device));
    *d_views = nullptr;
    if (n_views > 0) {
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
        *d_views = static_cast<const float *>(dv);
    }
    a = FwdArgs{};
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
    return LCT_OK;
}

}  // namespace lct

using namespace lct;

extern "C" int lct_forward_project(lct_ctx *ctx, const lct_ct_geom *g, const float *h_angles,
                                   const float *h_couch, long view0, long n_views,
                                   const float *d_vol, const lct_img_geom *vg, float vol_z0,
                                   int vol_nz, float step, float *d_out)
{
    if (!d_vol || !d_out) return set_error(LCT_ERR_INVALID, "lct_forward_project: bad argument");
    FwdArgs a;
    const float *dv = nullptr;
    LCT_CHECK(forward_setup(ctx, g, h_angles, h_couch, view0, n_views, vg, vol_z0, vol_nz, step,
                            "lct_forward_project", a, &dv));
    if (n_views == 0) return LCT_OK;
    a.vol = d_vol;
    for (long v0 = 0; v0 < n_views; v0 += 65535) {
        a.views = dv + 4 * v0;
        a.n_views = std::min<long>(65535, n_views - v0);
        a.out = d_out + size_t(v0) * g->nt * g->ns;
        const dim3 grid(unsigned((g->ns + 255) / 256), unsigned(g->nt), unsigned(a.n_views));
        hipLaunchKernelGGL(forward_project_kernel, grid, dim3(256), 0, ctx->stream, a);
        LCT_CHECK_HIP(hipGetLastError());
    }
    LCT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    return LCT_OK;
}

extern "C" int lct_forward_project_adjoint(lct_ctx *ctx, const lct_ct_geom *g,
                                           const float *h_angles, const float *h_couch,
                                           long view0, long n_views, const float *d_proj,
                                           const lct_img_geom *vg, float vol_z0, int vol_nz,
                                           float step, float *d_vol)
{
    if (!d_proj || !d_vol)
        return set_error(LCT_ERR_INVALID, "lct_forward_project_adjoint: bad argument");
    FwdArgs a;
    const float *dv = nullptr;
    LCT_CHECK(forward_setup(ctx, g, h_angles, h_couch, view0, n_views, vg, vol_z0, vol_nz, step,
                            "lct_forward_project_adjoint", a, &dv));
    LCT_CHECK_HIP(hipMemsetAsync(d_vol, 0, sizeof(float) * size_t(vol_nz) * vg->ny * vg->nx,
                                 ctx->stream));
    a.vol_out = d_vol;
    for (long v0 = 0; v0 < n_views; v0 += 65535) {
        a.views = dv + 4 * v0;
        a.n_views = std::min<long>(65535, n_views - v0);
        a.proj = d_proj + size_t(v0) * g->nt * g->ns;
        const dim3 grid(unsigned((g->ns + 255) / 256), unsigned(g->nt), unsigned(a.n_views));
        hipLaunchKernelGGL(forward_adjoint_kernel, grid, dim3(256), 0, ctx->stream, a);
        LCT_CHECK_HIP(hipGetLastError());
    }
    LCT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    return LCT_OK;
}
