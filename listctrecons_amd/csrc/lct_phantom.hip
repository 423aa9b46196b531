// lct_phantom.hip -- synthetic helical scans: exact ellipsoid line integrals
// on a curved detector and Poisson noise.  The reference ships no data; these
// kernels make every benchmark input of the listed workloads on the device.
#include <cmath>
#include <vector>

#include "lct_internal.hpp"

namespace lct {

constexpr int kMaxEll = 320;
constexpr int kEllStride = 16;

// Ellipsoids staged in LDS; one thread per ray (view, row, channel), FP64
// arithmetic so the integrals match the FP64 generator to rounding.
__global__ __launch_bounds__(256) void project_ellipsoids_kernel(
    const double *__restrict__ ell, int n_ell, const double *__restrict__ views, long n_views,
    int ns, int nt, double ws, double wt, double ds_over_dso, double dt_over_dso,
    float *__restrict__ out)
{
    __shared__ double se[kMaxEll * kEllStride];
    for (int i = threadIdx.x; i < n_ell * kEllStride; i += blockDim.x) se[i] = ell[i];
    __syncthreads();
    const int is = blockIdx.x * blockDim.x + threadIdx.x;
    const int it = blockIdx.y;
    const long v = blockIdx.z;
    if (is >= ns || v >= n_views) return;
    // views[v] = (sin b, cos b, z, b) with the source at (-dso sin b, dso cos b, z)
    const double sb = views[4 * v], cb = views[4 * v + 1], sz = views[4 * v + 2];
    const double b = views[4 * v + 3];
    const double a = (double(is) - ws) * ds_over_dso;
    const double dso = views[4 * n_views];
    const double ux = sin(b - a), uy = -cos(b - a);
    const double slope = (double(it) - wt) * dt_over_dso;
    const double inv = 1.0 / sqrt(1.0 + slope * slope);
    const double d0 = ux * inv, d1 = uy * inv, d2 = slope * inv;
    const double S0 = -dso * sb, S1 = dso * cb, S2 = sz;
    double total = 0.0;
    for (int e = 0; e < n_ell; e++) {
        const double *E = se + e * kEllStride;
        const double *R = E + 6;
        const double r0 = S0 - E[0], r1 = S1 - E[1], r2 = S2 - E[2];
        const double p0 = (R[0] * r0 + R[3] * r1 + R[6] * r2) / E[3];
        const double p1 = (R[1] * r0 + R[4] * r1 + R[7] * r2) / E[4];
        const double p2 = (R[2] * r0 + R[5] * r1 + R[8] * r2) / E[5];
        const double q0 = (R[0] * d0 + R[3] * d1 + R[6] * d2) / E[3];
        const double q1 = (R[1] * d0 + R[4] * d1 + R[7] * d2) / E[4];
        const double q2 = (R[2] * d0 + R[5] * d1 + R[8] * d2) / E[5];
        const double A = q0 * q0 + q1 * q1 + q2 * q2;
        const double B = p0 * q0 + p1 * q1 + p2 * q2;
        const double C = p0 * p0 + p1 * p1 + p2 * p2 - 1.0;
        const double disc = B * B - A * C;
        if (disc > 0.0) total += E[15] * 2.0 * sqrt(disc) / A;
    }
    out[(size_t(v) * nt + it) * ns + is] = float(total);
}

// SplitMix64 finaliser (Steele, Lea, Flood 2014) as a counter-based stream.
__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ double uniform01(uint64_t seed, uint64_t idx, uint32_t draw)
{
    const uint64_t x = mix64(seed ^ mix64(idx * 0x9E3779B97F4A7C15ull + draw));
    return (double(x >> 11) + 0.5) * (1.0 / 9007199254740992.0);  // (0,1)
}

// Poisson(lam): multiplication method below 10, otherwise Hormann's PTRS
// transformed rejection (W. Hormann, "The transformed rejection method for
// generating Poisson random variables", Insurance: Math. Econ. 12, 1993).
__device__ double poisson(double lam, uint64_t seed, uint64_t idx)
{
    uint32_t draw = 0;
    if (lam < 10.0) {
        const double L = exp(-lam);
        double p = 1.0;
        int k = 0;
        do {
            p *= uniform01(seed, idx, draw++);
            ++k;
        } while (p > L && k < 1000);
        return double(k - 1);
    }
    const double slam = sqrt(lam), loglam = log(lam);
    const double b = 0.931 + 2.53 * slam;
    const double a = -0.059 + 0.02483 * b;
    const double inv_alpha = 1.1239 + 1.1328 / (b - 3.4);
    const double vr = 0.9277 - 3.6224 / (b - 2.0);
    for (int iter = 0; iter < 256; iter++) {
        const double U = uniform01(seed, idx, draw++) - 0.5;
        const double V = uniform01(seed, idx, draw++);
        const double us = 0.5 - fabs(U);
        const double k = floor((2.0 * a / us + b) * U + lam + 0.43);
        if (us >= 0.07 && V <= vr) return k;
        if (k < 0.0 || (us < 0.013 && V > us)) continue;
        if (log(V) + log(inv_alpha) - log(a / (us * us) + b) <= -lam + k * loglam - lgamma(k + 1.0))
            return k;
    }
    return floor(lam + 0.5);  // not reached in practice (acceptance ~ 0.9 per try)
}

__global__ __launch_bounds__(256) void poisson_noise_kernel(float *__restrict__ p, size_t n,
                                                            double i0, uint64_t seed)
{
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double lam = i0 * exp(-double(p[i]));
    const double cnt = poisson(lam, seed, i);
    p[i] = float(-log(fmax(cnt, 1.0) / i0));
}

// Phantom density at voxel centres (ss^3 sub-samples per voxel averaged).
__global__ __launch_bounds__(256) void voxelize_kernel(const double *__restrict__ ell, int n_ell,
                                                       int nx, int ny, int nz, double wx,
                                                       double wy, double dx, double dy,
                                                       double z0, double dz, int ss,
                                                       float *__restrict__ vol)
{
    __shared__ double se[kMaxEll * kEllStride];
    for (int i = threadIdx.x; i < n_ell * kEllStride; i += blockDim.x) se[i] = ell[i];
    __syncthreads();
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = blockIdx.y, iz = blockIdx.z;
    if (ix >= nx) return;
    double acc = 0.0;
    for (int sz = 0; sz < ss; sz++)
        for (int sy = 0; sy < ss; sy++)
            for (int sx = 0; sx < ss; sx++) {
                const double x = (double(ix) + (sx + 0.5) / ss - 0.5 - wx) * dx;
                const double y = (double(iy) + (sy + 0.5) / ss - 0.5 - wy) * dy;
                const double z = z0 + (double(iz) + (sz + 0.5) / ss - 0.5) * dz;
                for (int e = 0; e < n_ell; e++) {
                    const double *E = se + e * kEllStride;
                    const double *R = E + 6;
                    const double r0 = x - E[0], r1 = y - E[1], r2 = z - E[2];
                    const double p0 = (R[0] * r0 + R[3] * r1 + R[6] * r2) / E[3];
                    const double p1 = (R[1] * r0 + R[4] * r1 + R[7] * r2) / E[4];
                    const double p2 = (R[2] * r0 + R[5] * r1 + R[8] * r2) / E[5];
                    if (p0 * p0 + p1 * p1 + p2 * p2 <= 1.0) acc += E[15];
                }
            }
    vol[(size_t(iz) * ny + iy) * nx + ix] = float(acc / double(ss * ss * ss));
}

}  // namespace lct

using namespace lct;

extern "C" int lct_voxelize_ellipsoids(lct_ctx *ctx, const lct_img_geom *ig, double z0, int nz,
                                       const double *h_ell, int n_ell, int supersample,
                                       float *d_vol)
{
    if (!ctx || !ig || !h_ell || !d_vol || nz <= 0 || ig->nx <= 0 || ig->ny <= 0 || n_ell < 0 ||
        n_ell > kMaxEll || supersample < 1 || supersample > 8)
        return set_error(LCT_ERR_INVALID, "lct_voxelize_ellipsoids: bad argument");
    LCT_CHECK_HIP(hipSetDevice(ctx->device));
    void *de = nullptr;
    LCT_CHECK(scratch(ctx, kScratchMisc1, size_t(n_ell) * kEllStride * 8 + 8, &de));
    LCT_CHECK(upload(ctx, de, h_ell, size_t(n_ell) * kEllStride * 8));
    const dim3 grid(unsigned((ig->nx + 255) / 256), unsigned(ig->ny), unsigned(nz));
    hipLaunchKernelGGL(voxelize_kernel, grid, dim3(256), 0, ctx->stream,
                       static_cast<const double *>(de), n_ell, ig->nx, ig->ny, nz, double(ig->wx),
                       double(ig->wy), double(ig->dx), double(ig->dy), z0, double(ig->dz),
                       supersample, d_vol);
    LCT_CHECK_HIP(hipGetLastError());
    LCT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    return LCT_OK;
}

extern "C" int lct_project_ellipsoids(lct_ctx *ctx, const lct_ct_geom *g, const float *h_angles,
                                      const float *h_couch, long view0, long n_views,
                                      const double *h_ell, int n_ell, float *d_out)
{
    if (!ctx || !g || !h_angles || !h_couch || !h_ell || !d_out || n_views < 0 || view0 < 0 ||
        n_ell < 0 || n_ell > kMaxEll || g->ns <= 0 || g->nt <= 0 || !(g->dso > 0.0f))
        return set_error(LCT_ERR_INVALID, "lct_project_ellipsoids: bad argument");
    if (n_views == 0) return LCT_OK;
    LCT_CHECK_HIP(hipSetDevice(ctx->device));
    std::vector<double> vt(4 * size_t(n_views) + 1);
    for (long v = 0; v < n_views; v++) {
        const double b = h_angles[view0 + v];
        vt[4 * v] = std::sin(b);
        vt[4 * v + 1] = std::cos(b);
        vt[4 * v + 2] = h_couch[view0 + v];
        vt[4 * v + 3] = b;
    }
    vt[4 * size_t(n_views)] = g->dso;
    void *dv = nullptr, *de = nullptr;
    LCT_CHECK(scratch(ctx, kScratchMisc0, vt.size() * 8, &dv));
    LCT_CHECK(scratch(ctx, kScratchMisc1, size_t(n_ell) * kEllStride * 8 + 8, &de));
    LCT_CHECK(upload(ctx, dv, vt.data(), vt.size() * 8));
    LCT_CHECK(upload(ctx, de, h_ell, size_t(n_ell) * kEllStride * 8));
    // one launch per <= 65535 views (grid.z limit)
    for (long v0 = 0; v0 < n_views; v0 += 65535) {
        const long nv = std::min<long>(65535, n_views - v0);
        const dim3 grid(unsigned((g->ns + 255) / 256), unsigned(g->nt), unsigned(nv));
        hipLaunchKernelGGL(project_ellipsoids_kernel, grid, dim3(256), 0, ctx->stream,
                           static_cast<const double *>(de), n_ell,
                           static_cast<const double *>(dv) + 4 * v0, nv, g->ns, g->nt,
                           double(g->ws), double(g->wt), double(g->ds) / double(g->dso),
                           double(g->dt) / double(g->dso),
                           d_out + size_t(v0) * g->nt * g->ns);
        LCT_CHECK_HIP(hipGetLastError());
    }
    // the view table is read by the kernels above; keep it alive until they finish
    LCT_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    return LCT_OK;
}

extern "C" int lct_poisson_noise(lct_ctx *ctx, float *d_proj, size_t n, float i0, uint64_t seed)
{
    if (!ctx || !d_proj || !(i0 > 0.0f))
        return set_error(LCT_ERR_INVALID, "lct_poisson_noise: bad argument");
    if (n == 0) return LCT_OK;
    LCT_CHECK_HIP(hipSetDevice(ctx->device));
    hipLaunchKernelGGL(poisson_noise_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0,
                       ctx->stream, d_proj, n, double(i0), seed);
    LCT_CHECK_HIP(hipGetLastError());
    return LCT_OK;
}
