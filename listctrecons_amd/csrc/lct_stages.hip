// lct_stages.hip -- the reference's four stages as CDNA4 kernels with the
// reference's per-block array layouts:
//   rebin      fan [NA+][nt][ns] -> par [NA][nt][ns]  (CUDA/Rebiner.cu)
//   weighting  in place on par                        (CUDA/PreWeighting.cu)
//   filter     in place on par rows, FP32             (CUDA/Filter.cu, Src/Filter.cpp)
//   backproject par -> img [32][ny][nx]               (CUDA/Backprojector.cu)
//
// The reference samples through CUDA texture references (linear filtering,
// border addressing).  Here every interpolation is explicit: the fetch
// coordinate is split into an integer texel index and an FP32 weight that
// the host evaluates in double, and texels outside the array read as 0.
#include <cmath>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "lct_internal.hpp"

namespace lct {

// ------------------------------------------------------------------ rebin

// CUDA/Rebiner.cu:82-103 (FAN2PAR) and :13-29 (NULL, frac == 0).
// par[ia][it][is] = lerp_view(fan[.][it][is], base + ia + k0[is] + fr[is]).
// A thread computes V consecutive output views of one channel.  Output view
// ia reads fan views v0 = base + ia + k0[is] and v0 + 1, so views
// ia .. ia+V-1 read the V + 1 fan views v0 .. v0 + V, each loaded once per
// row.
template <int V>
__global__ __launch_bounds__(256) void rebin_views_lerp_kernel(
    const float *__restrict__ fan, long fan_views, float *__restrict__ par, long base, int ns,
    int nt, const int *__restrict__ k0, const float *__restrict__ fr,
    const float *__restrict__ wtab, long n_out, int rows)
{
    // 1-D grid (a multiple of 8) of (channel chunk, chunk of `rows` rows, group
    // of V output views); each XCD takes a contiguous range of groups (their fan
    // views share its L2)
    const long n_wg = long(gridDim.x);
    const long wg = long(blockIdx.x & 7u) * (n_wg >> 3) + long(blockIdx.x >> 3);
    const long n_cx = (ns + 255) / 256, n_rc = (nt + rows - 1) / rows;
    const long grp = wg / (n_cx * n_rc);
    const long rem = wg - grp * n_cx * n_rc;
    const long rc = rem / n_cx;
    const int is = int(rem - rc * n_cx) * 256 + int(threadIdx.x);
    const int it0 = int(rc) * rows, it1 = min(nt, it0 + rows);
    const long ia0 = grp * V;
    if (ia0 >= n_out || is >= ns) return;
    const int nv = int(n_out - ia0 < V ? n_out - ia0 : V);
    const long v0 = base + ia0 + k0[is];
    const float f = fr[is];
    const size_t view = size_t(nt) * ns;
    const float *src = fan + is;
    bool ok[V + 1];
    size_t off[V + 1];
#pragma unroll
    for (int j = 0; j <= V; j++) {
        ok[j] = v0 + j >= 0 && v0 + j < fan_views;
        off[j] = ok[j] ? size_t(v0 + j) * view : 0;
    }
    float *o = par + size_t(ia0) * view + is;
#pragma unroll 4
    for (int it = it0; it < it1; it++) {
        const size_t r = size_t(it) * ns;
        float x[V + 1];
#pragma unroll
        for (int j = 0; j <= V; j++) x[j] = ok[j] ? src[off[j] + r] : 0.0f;
        const float w = wtab ? wtab[r + is] : 1.0f;
#pragma unroll
        for (int j = 0; j < V; j++) {
            if (j >= nv) break;
            float v = fmaf(f, x[j + 1] - x[j], x[j]);
            if (wtab) v *= w;
            o[size_t(j) * view + r] = v;
        }
    }
}

// Output views per thread of the rebin (clinical scan: 1.24 ms with 1,
// 0.99 with 2, 0.82 with 4, 0.71 with 8, 0.72 with 16).
constexpr int kRebinViews = 8;
// Detector rows per thread of the rebin: rows are independent, and a thread
// walking all of them left too few waves to hide the loads' latency.
constexpr int kRebinRows = 16;

// CUDA/Rebiner.cu:156-179 (PI_ORI): trilinear sample at (pI, qI, bI), with
// pI and the view fraction depending on the channel only and qI on (row,
// channel). A thread makes V consecutive output views of one channel, row by
// row: the V + 1 fan views they read are sampled bilinearly at (qI, pI) once
// each and the outputs lerp neighbouring samples (as the fan-to-parallel
// rebin above). Taps outside the views, rows or channels read zero (the
// texture's border mode).
template <int V>
__global__ __launch_bounds__(256) void rebin_pi_ori_kernel(
    const float *__restrict__ fan, long fan_views, float *__restrict__ par, long base, int ns,
    int nt, const int *__restrict__ pi0, const float *__restrict__ pfr,
    const int *__restrict__ vk0, const float *__restrict__ vfr, const float *__restrict__ qc1,
    const float *__restrict__ qc0, float wt, const float *__restrict__ wtab, long n_out, int rows)
{
    // grid as in rebin_views_lerp_kernel: (channel chunk, row chunk, view group)
    const long n_wg = long(gridDim.x);
    const long wg = long(blockIdx.x & 7u) * (n_wg >> 3) + long(blockIdx.x >> 3);
    const long n_cx = (ns + 255) / 256, n_rc = (nt + rows - 1) / rows;
    const long grp = wg / (n_cx * n_rc);
    const long rem = wg - grp * n_cx * n_rc;
    const long rc = rem / n_cx;
    const int is = int(rem - rc * n_cx) * 256 + int(threadIdx.x);
    const int it0 = int(rc) * rows, it1 = min(nt, it0 + rows);
    const long ia0 = grp * V;
    if (ia0 >= n_out || is >= ns) return;
    const int nv = int(n_out - ia0 < V ? n_out - ia0 : V);
    const int x0 = pi0[is];
    const float ax = pfr[is];
    const long z0 = base + ia0 + vk0[is];
    const float az = vfr[is];
    const float c1 = qc1[is], c0 = qc0[is];
    const size_t view = size_t(nt) * ns;
    const bool okx0 = x0 >= 0 && x0 < ns, okx1 = x0 + 1 >= 0 && x0 + 1 < ns;
    const long xo = okx0 ? x0 : (okx1 ? x0 + 1 : 0);  // an in-range channel to address from
    bool okz[V + 1];
    size_t offz[V + 1];
#pragma unroll
    for (int j = 0; j <= V; j++) {
        okz[j] = z0 + j >= 0 && z0 + j < fan_views;
        offz[j] = okz[j] ? size_t(z0 + j) * view + size_t(xo) : 0;
    }
    const long d0 = long(x0) - xo, d1 = d0 + 1;  // channel offsets of the two taps from xo
    float *o = par + size_t(ia0) * view + is;
    // The lanes of a wave take consecutive channels, whose fan views differ
    // (the view offset moves with the fan angle), so each load touches many
    // cache lines.  When both taps of every lane lie in range (all waves but
    // the border ones), the two taps, adjacent floats, come in one 8-byte load.
    auto rows_loop = [&](auto pair_loads) {
        for (int it = it0; it < it1; it++) {
            const float qI = fmaf(float(it) - wt, c1, c0);
            const float fy = floorf(qI);
            const int y0 = int(fy);
            const float ay = qI - fy;
            const bool oky0 = y0 >= 0 && y0 < nt, oky1 = y0 + 1 >= 0 && y0 + 1 < nt;
            const long r0 = long(y0) * ns, r1 = r0 + ns;
            float smp[V + 1];
#pragma unroll
            for (int j = 0; j <= V; j++) {
                const float *f = fan + offz[j];
                float v00, v01, v10, v11;
                if constexpr (decltype(pair_loads)::value) {
                    // d0 == 0 here: the pair starts at xo = x0 (4-byte aligned)
                    typedef float float2u __attribute__((ext_vector_type(2), aligned(4)));
                    const float2u z2 = {0.0f, 0.0f};
                    const float2u p0 = okz[j] && oky0 ? *reinterpret_cast<const float2u *>(f + r0) : z2;
                    const float2u p1 = okz[j] && oky1 ? *reinterpret_cast<const float2u *>(f + r1) : z2;
                    v00 = p0.x;
                    v01 = p0.y;
                    v10 = p1.x;
                    v11 = p1.y;
                } else {
                    v00 = okz[j] && oky0 && okx0 ? f[r0 + d0] : 0.0f;
                    v01 = okz[j] && oky0 && okx1 ? f[r0 + d1] : 0.0f;
                    v10 = okz[j] && oky1 && okx0 ? f[r1 + d0] : 0.0f;
                    v11 = okz[j] && oky1 && okx1 ? f[r1 + d1] : 0.0f;
                }
                const float lo = fmaf(ax, v01 - v00, v00), hi = fmaf(ax, v11 - v10, v10);
                smp[j] = fmaf(ay, hi - lo, lo);
            }
            const size_t r = size_t(it) * ns;
            const float w = wtab ? wtab[r + is] : 1.0f;
#pragma unroll
            for (int j = 0; j < V; j++) {
                if (j >= nv) break;
                float v = fmaf(az, smp[j + 1] - smp[j], smp[j]);
                if (wtab) v *= w;
                o[size_t(j) * view + r] = v;
            }
        }
    };
    if (__builtin_amdgcn_ballot_w64(!(okx0 && okx1)) == 0)
        rows_loop(std::integral_constant<bool, true>());
    else
        rows_loop(std::integral_constant<bool, false>());
}

// CUDA/PreWeighting.cu:8-29: par[ia][it][is] *= w[it][is].
__global__ __launch_bounds__(256) void weight_kernel(float *__restrict__ par, size_t view, int n,
                                                     const float *__restrict__ wtab)
{
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= size_t(n) * view) return;
    par[i] *= wtab[i % view];
}

// PiOriPreWeighting weight table in double (CUDA/PreWeighting.cu:26-32):
// w = sqrt(dso^2 - s^2) / sqrt(dso^2 - s^2 + (t - gamma P / 2pi)^2),
// s = (is - ws) ds, t = (it - wt) P / (2 nt).
static void pi_ori_weights(const lct_ct_geom &g, std::vector<float> &w)
{
    w.resize(size_t(g.nt) * g.ns);
    const double P = g.pitch, dso = g.dso;
    for (int it = 0; it < g.nt; it++)
        for (int is = 0; is < g.ns; is++) {
            double s = (double(is) - double(g.ws)) * double(g.ds);
            double t = (double(it) - double(g.wt)) * P / (2.0 * g.nt);
            double gam = std::asin(s / dso);
            double a = dso * dso - s * s;
            double b = t - gam * P / (2.0 * LCT_PI_D);
            w[size_t(it) * g.ns + is] = float(std::sqrt(a) / std::sqrt(a + b * b));
        }
}

static void split(double v, int &i, float &f)
{
    double fl = std::floor(v);
    i = int(fl);
    f = float(v - fl);
}

int launch_rebin(lct_ctx *ctx, int kind, int weight, const lct_ct_geom *gp, const float *d_fan,
                 long fan_view0, long fan_views, float *d_par, long n_out, long out_view0,
                 int /*transposed*/)
{
    const lct_ct_geom &g = *gp;
    const int ns = g.ns, nt = g.nt;
    // Host tables in double from the float32 geometry.
    std::vector<int> ia(2 * size_t(ns));
    std::vector<float> fa(4 * size_t(ns));
    std::vector<float> wtab;
    if (weight == LCT_WEIGHT_PI_ORI) pi_ori_weights(g, wtab);
    const double dso = g.dso, ds = g.ds, orbit = g.orbit, ws = g.ws;
    for (int is = 0; is < ns; is++) {
        double s = (double(is) - ws) * ds;
        if (kind == LCT_REBIN_FAN2PAR) {
            // alpha = (is - ws) ds / dso; view offset alpha / orbit (Rebiner.cu:92-95)
            split(s / dso / orbit, ia[is], fa[is]);
        } else if (kind == LCT_REBIN_PI_ORI) {
            double gam = std::asin(s / dso);
            split(gam * dso / ds + ws, ia[is], fa[is]);           // pI
            split(gam / orbit, ia[ns + is], fa[ns + is]);         // bI - ia - add
            double cg = std::cos(gam);
            fa[2 * ns + is] = float(g.pitch / (2.0 * (nt - 1)) / cg / g.dt);  // dqI/dit
            fa[3 * ns + is] =
                float(-gam * g.pitch / (2.0 * LCT_PI_D) / cg / g.dt + double(g.wt));
        } else {
            ia[is] = 0;
            fa[is] = 0.0f;
        }
    }
    void *dt = nullptr;
    const size_t tab_bytes = ia.size() * 4 + fa.size() * 4 + wtab.size() * 4;
    LCT_CHECK(scratch(ctx, kScratchMisc1, tab_bytes, &dt));
    int *d_i = static_cast<int *>(dt);
    float *d_f = reinterpret_cast<float *>(d_i + ia.size());
    float *d_w = wtab.empty() ? nullptr : d_f + fa.size();
    LCT_CHECK(upload(ctx, d_i, ia.data(), ia.size() * 4));
    LCT_CHECK(upload(ctx, d_f, fa.data(), fa.size() * 4));
    if (d_w) LCT_CHECK(upload(ctx, d_w, wtab.data(), wtab.size() * 4));

    const float *fan = d_fan + size_t(fan_view0) * ns * nt;
    if (n_out <= 0) return LCT_OK;
    const long n_rc = (nt + kRebinRows - 1) / kRebinRows;
    if (kind == LCT_REBIN_PI_ORI) {
        constexpr int V = kRebinViews;
        const long n_grp = (n_out + V - 1) / V;
        const long n_wg = ((long((ns + 255) / 256) * n_rc * n_grp) + 7) & ~7L;
        if (n_wg > 0x7fffffffL) return set_error(LCT_ERR_UNSUPPORTED, "rebin: too many views");
        hipLaunchKernelGGL(rebin_pi_ori_kernel<V>, dim3(unsigned(n_wg)), dim3(256), 0, ctx->stream,
                           fan, fan_views, d_par, out_view0, ns, nt, d_i, d_f, d_i + ns, d_f + ns,
                           d_f + 2 * ns, d_f + 3 * ns, g.wt, d_w, long(n_out), kRebinRows);
    } else {
        constexpr int V = kRebinViews;
        const long n_grp = (n_out + V - 1) / V;
        const long n_wg = ((long((ns + 255) / 256) * n_rc * n_grp) + 7) & ~7L;
        if (n_wg > 0x7fffffffL) return set_error(LCT_ERR_UNSUPPORTED, "rebin: too many views");
        hipLaunchKernelGGL(rebin_views_lerp_kernel<V>, dim3(unsigned(n_wg)), dim3(256), 0,
                           ctx->stream, fan, fan_views, d_par, out_view0, ns, nt, d_i, d_f, d_w,
                           long(n_out), kRebinRows);
    }
    LCT_CHECK_HIP(hipGetLastError());
    return LCT_OK;
}

int launch_weighting(lct_ctx *ctx, const lct_ct_geom *gp, float *d_par)
{
    const lct_ct_geom &g = *gp;
    std::vector<float> w;
    pi_ori_weights(g, w);
    void *dw = nullptr;
    LCT_CHECK(scratch(ctx, kScratchMisc1, w.size() * 4, &dw));
    LCT_CHECK(upload(ctx, dw, w.data(), w.size() * 4));
    const size_t view = size_t(g.nt) * g.ns;
    const size_t total = view * g.n_angle;
    hipLaunchKernelGGL(weight_kernel, dim3(unsigned((total + 255) / 256)), dim3(256), 0,
                       ctx->stream, d_par, view, g.n_angle, static_cast<const float *>(dw));
    LCT_CHECK_HIP(hipGetLastError());
    return LCT_OK;
}

// ----------------------------------------------------------------- filter

// CUDA/Filter.cu:10-50: out[r][is] = sum_{k<ns} h[is-k+ns] row_r[k], k
// ascending with FMA, one FP32 accumulator per output.  A workgroup filters
// R rows; a thread owns J consecutive outputs and keeps the J taps it needs
// for the current k in registers, shifting them by one tap per k.
template <int J>
__global__ __launch_bounds__(256) void filter_direct_kernel(const float *__restrict__ taps,
                                                            float *__restrict__ rows,
                                                            long n_rows, int ns)
{
    constexpr int R = 4;
    extern __shared__ float sm[];
    float *st = sm;           // 2 ns taps
    float *sr = sm + 2 * ns;  // R rows
    const long r0 = long(blockIdx.x) * R;
    for (int i = threadIdx.x; i < 2 * ns; i += blockDim.x) st[i] = taps[i];
    for (int i = threadIdx.x; i < R * ns; i += blockDim.x) {
        const long rr = r0 + i / ns;
        sr[i] = rr < n_rows ? rows[size_t(rr) * ns + (i % ns)] : 0.0f;
    }
    __syncthreads();
    const int is0 = threadIdx.x * J;
    if (is0 >= ns) return;
    float acc[R][J];
    float hreg[J];
#pragma unroll
    for (int j = 0; j < J; j++) {
        hreg[j] = (is0 + j < ns) ? st[is0 + j + ns] : 0.0f;
#pragma unroll
        for (int r = 0; r < R; r++) acc[r][j] = 0.0f;
    }
    for (int k = 0; k < ns; k++) {
#pragma unroll
        for (int r = 0; r < R; r++) {
            const float x = sr[r * ns + k];
#pragma unroll
            for (int j = 0; j < J; j++) acc[r][j] = fmaf(hreg[j], x, acc[r][j]);
        }
#pragma unroll
        for (int j = J - 1; j > 0; j--) hreg[j] = hreg[j - 1];
        hreg[0] = st[is0 - k - 1 + ns];  // index >= is0 >= 0 for k < ns
    }
#pragma unroll
    for (int r = 0; r < R; r++) {
        const long rr = r0 + r;
        if (rr >= n_rows) break;
#pragma unroll
        for (int j = 0; j < J; j++)
            if (is0 + j < ns) rows[size_t(rr) * ns + is0 + j] = acc[r][j];
    }
}

int launch_filter_direct(lct_ctx *ctx, int ns, const float *d_taps, float *d_rows, long rows)
{
    if (rows <= 0) return LCT_OK;
    const int J = (ns + 255) / 256;
    const size_t shm = size_t(2 * ns + 4 * ns) * sizeof(float);
    const dim3 grid(unsigned((rows + 3) / 4));
#define LCT_FILTER_CASE(JJ)                                                                   \
    case JJ:                                                                                  \
        hipLaunchKernelGGL(filter_direct_kernel<JJ>, grid, dim3(256), shm, ctx->stream, d_taps, \
                           d_rows, rows, ns);                                                 \
        break;
    switch (J) {
        LCT_FILTER_CASE(1)
        LCT_FILTER_CASE(2)
        LCT_FILTER_CASE(3)
        LCT_FILTER_CASE(4)
        LCT_FILTER_CASE(5)
        LCT_FILTER_CASE(6)
        LCT_FILTER_CASE(7)
        LCT_FILTER_CASE(8)
    default:
        return set_error(LCT_ERR_UNSUPPORTED, "filter: ns > 2048 not supported");
    }
#undef LCT_FILTER_CASE
    LCT_CHECK_HIP(hipGetLastError());
    return LCT_OK;
}

// ------------------------------------------------------------- backproject

struct BpGeom {
    int nx, ny, ns, nt, na, h;
    float wx, wy, wz, dx, dy, dz;
    float dso, ds, dt, ws, wt, orbit, z0, dzv, zdir, pitch, temp2;
};

// tex3D(proj, p+0.5, r+0.5, ia+0.5) of one view [nt][ns] with linear
// filtering and border addressing.
__device__ __forceinline__ float tap_border(const float *__restrict__ v, int ns, int nt, float p,
                                            float r)
{
    if (!(p > -1.0f && p < float(ns) && r > -1.0f && r < float(nt))) return 0.0f;
    const float fp = floorf(p), fr = floorf(r);
    const int p0 = int(fp), r0 = int(fr);
    const float ap = p - fp, ar = r - fr;
    const bool pl = p0 >= 0, ph = p0 + 1 < ns, rl = r0 >= 0, rh = r0 + 1 < nt;
    const float v00 = (pl && rl) ? v[r0 * ns + p0] : 0.0f;
    const float v01 = (ph && rl) ? v[r0 * ns + p0 + 1] : 0.0f;
    const float v10 = (pl && rh) ? v[(r0 + 1) * ns + p0] : 0.0f;
    const float v11 = (ph && rh) ? v[(r0 + 1) * ns + p0 + 1] : 0.0f;
    const float lo = fmaf(ap, v01 - v00, v00);
    const float hi = fmaf(ap, v11 - v10, v10);
    return fmaf(ar, hi - lo, lo);
}

// CUDA/Backprojector.cu:12-99 evaluated per voxel as written: both rays of
// every view in the voxel's window, conjugate z-weights, 2 pi / (a_max-a_min).
// Differences: arguments by value (no unsynchronised __shared__ copy,
// :28-31), cos/sin of each float32 tube angle from a double-evaluated table
// instead of __cosf/__sinf, explicit bilinear taps.
__global__ __launch_bounds__(256) void bp_helical_literal_kernel(
    const float *__restrict__ q, const float *__restrict__ cosb, const float *__restrict__ sinb,
    const int *__restrict__ win, float *__restrict__ img, BpGeom G)
{
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = blockIdx.y, iz = blockIdx.z;
    if (ix >= G.nx) return;
    const float xc = (float(ix) - G.wx) * G.dx;
    const float yc = (float(iy) - G.wy) * G.dy;
    const float zc = (float(iz) - G.wz) * G.dz;
    const int amin = win[2 * iz], amax = win[2 * iz + 1];
    const size_t view = size_t(G.ns) * G.nt;
    const float kq = G.dso * G.zdir / G.dt;
    float sum = 0.0f;
    for (int ia = amin; ia < amax; ia++) {
        int ia2 = ia + G.h < amax ? ia + G.h : ia - G.h;
        ia2 = ia2 < 0 ? ia + G.h : ia2;
        const float c = cosb[ia], s = sinb[ia], c2 = cosb[ia2], s2 = sinb[ia2];
        const float phat = -xc * c - yc * s;
        const float theta = asinf(phat / G.dso);
        const float delta = theta * G.dso / G.ds;
        const float p1 = delta + G.ws, p2 = -delta + G.ws;
        const float zr = G.z0 + (float(ia) + theta / G.orbit) * G.dzv;
        const float zr2 = G.z0 + (float(ia2) - theta / G.orbit) * G.dzv;
        const float root = sqrtf(G.dso * G.dso - phat * phat);
        const float lhat = root + xc * s - yc * c;
        const float lhat2 = root + xc * s2 - yc * c2;
        const float qh = (zc - zr) * kq / lhat;
        const float qh2 = (zc - zr2) * kq / lhat2;
        float w1, w2;
        if (fabsf(qh) > G.wt) {
            w1 = 0.0f;
            w2 = 1.0f;
        } else if (fabsf(qh2) > G.wt) {
            w1 = 1.0f;
            w2 = 0.0f;
        } else {
            w1 = -qh2 / (qh - qh2);
            w2 = qh / (qh - qh2);
        }
        const float t1 = (ia < G.na) ? tap_border(q + size_t(ia) * view, G.ns, G.nt, p1, qh + G.wt) : 0.0f;
        const float t2 = (ia2 < G.na) ? tap_border(q + size_t(ia2) * view, G.ns, G.nt, p2, qh2 + G.wt) : 0.0f;
        sum += t1 * w1 + t2 * w2;
    }
    img[(size_t(iz) * G.ny + iy) * G.nx + ix] =
        sum * 2.0f * float(LCT_PI_D) / float(amax - amin);
}

// CUDA/Backprojector.cu:171-250 (PI-window, uniform-p rebinned data).
__global__ __launch_bounds__(256) void bp_pi_ori_literal_kernel(
    const float *__restrict__ q, const float *__restrict__ cosb, const float *__restrict__ sinb,
    const int *__restrict__ win, float *__restrict__ img, BpGeom G)
{
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = blockIdx.y, iz = blockIdx.z;
    if (ix >= G.nx) return;
    const float xc = (float(ix) - G.wx) * G.dx;
    const float yc = (float(iy) - G.wy) * G.dy;
    const float zc = (float(iz) - G.wz) * G.dz;
    const int am = win[2 * 32 + iz];
    const size_t view = size_t(G.ns) * G.nt;
    const float quarter = G.pitch / 4.0f;
    const float row_scale = 2.0f * float(G.nt - 1) / G.pitch;
    float sum = 0.0f;
    int ia;
    for (ia = am; ia >= 0; ia--) {
        const float c = cosb[ia], sb = sinb[ia];
        const float s = -xc * c - yc * sb;
        const float gam = asinf(s / G.dso);
        const float temp1 = sqrtf(G.dso * G.dso - s * s);
        const float zoff = zc - (G.z0 + G.dzv * float(ia));
        float t = temp1 * (zoff - gam * G.temp2) / (temp1 + xc * sb - yc * c) + gam * G.temp2;
        t *= G.zdir;
        if (t > quarter || t < -quarter) break;
        const float is = s / G.ds + G.ws;
        const float it = t * row_scale + G.wt;
        sum += tap_border(q + size_t(ia) * view, G.ns, G.nt, is, it);
    }
    const int a_min = ia;
    for (ia = am + 1; ia <= a_min + G.h; ia++) {
        if (ia >= G.na) break;  // the border-addressed texture returns 0 past the block
        const float c = cosb[ia], sb = sinb[ia];
        const float s = -xc * c - yc * sb;
        const float gam = asinf(s / G.dso);
        const float temp1 = sqrtf(G.dso * G.dso - s * s);
        const float zoff = zc - (G.z0 + G.dzv * float(ia));
        float t = temp1 * (zoff - gam * G.temp2) / (temp1 + xc * sb - yc * c) + gam * G.temp2;
        t *= G.zdir;
        const float is = s / G.ds + G.ws;
        const float it = t * row_scale + G.wt;
        sum += tap_border(q + size_t(ia) * view, G.ns, G.nt, is, it);
    }
    img[(size_t(iz) * G.ny + iy) * G.nx + ix] = sum * float(LCT_PI_D) / float(G.h);
}

static BpGeom make_bp_geom(const lct_ct_geom &g, const lct_img_geom &ig)
{
    BpGeom G;
    G.nx = ig.nx;
    G.ny = ig.ny;
    G.ns = g.ns;
    G.nt = g.nt;
    G.na = g.n_angle;
    G.h = g.n_proj_turn / 2;
    G.wx = ig.wx;
    G.wy = ig.wy;
    G.wz = ig.wz;
    G.dx = ig.dx;
    G.dy = ig.dy;
    G.dz = ig.dz;
    G.dso = g.dso;
    G.ds = g.ds;
    G.dt = g.dt;
    G.ws = g.ws;
    G.wt = g.wt;
    G.orbit = g.orbit;
    G.z0 = g.source_z0;
    G.dzv = g.pitch / float(g.n_proj_turn);  // Backprojector.cu:149
    G.zdir = float(g.z_dir);
    G.pitch = g.pitch;
    G.temp2 = float(double(g.pitch) / (2.0 * LCT_PI_D));
    return G;
}

// a_middle, a_min, a_max of CUDA/Backprojector.cu:37-42 for each slice,
// evaluated in float32 exactly as the kernel does.
void bp_windows(const lct_ct_geom &g, const lct_img_geom &ig, int *amin, int *amax, int *amid)
{
    const float dzv = g.pitch / float(g.n_proj_turn);
    const int h = g.n_proj_turn / 2;
    for (int iz = 0; iz < ig.nz; iz++) {
        const float zc = (float(iz) - ig.wz) * ig.dz;
        int a = int(roundf((zc - g.source_z0) / dzv));
        a = a < 0 ? 0 : a;
        a = a > g.n_angle - 1 ? g.n_angle - 1 : a;
        amid[iz] = a;
        amin[iz] = (a - h >= 0) ? (a - h) : 0;
        amax[iz] = (a + h < g.n_angle) ? (a + h) : g.n_angle - 1;
    }
}

// cos/sin of each float32 tube angle, evaluated in double.
void angle_tables(const float *betas, int n, std::vector<float> &cs)
{
    cs.resize(2 * size_t(n));
    for (int i = 0; i < n; i++) {
        cs[i] = float(std::cos(double(betas[i])));
        cs[n + i] = float(std::sin(double(betas[i])));
    }
}

}  // namespace lct

using namespace lct;

namespace {

// Uploads the block tables (cos, sin, windows) and launches the literal kernel.
int backproject_block(lct_ctx *ctx, int kind, const lct_ct_geom &g, const lct_img_geom &ig,
                      const float *d_q, const float *h_betas, float *d_img)
{
    if (ig.nz != LCT_BLOCK_SLICES)
        return set_error(LCT_ERR_INVALID, "backproject: ImgGeom nz must be 32 (BLOCK_SLICES)");
    const int nb = g.n_angle + g.n_proj_turn / 2 + 1;
    std::vector<float> cs;
    angle_tables(h_betas, nb, cs);
    int amin[32], amax[32], amid[32];
    bp_windows(g, ig, amin, amax, amid);
    // kernel layout: (a_min, a_max) pairs for the 32 slices, then a_middle[32]
    std::vector<int> wk(3 * 32);
    for (int iz = 0; iz < 32; iz++) {
        wk[2 * iz] = amin[iz];
        wk[2 * iz + 1] = amax[iz];
        wk[64 + iz] = amid[iz];
    }
    void *dt = nullptr;
    const size_t bytes = cs.size() * 4 + wk.size() * 4;
    LCT_CHECK(scratch(ctx, kScratchTables, bytes, &dt));
    float *d_cs = static_cast<float *>(dt);
    int *d_win = reinterpret_cast<int *>(d_cs + cs.size());
    LCT_CHECK(upload(ctx, d_cs, cs.data(), cs.size() * 4));
    LCT_CHECK(upload(ctx, d_win, wk.data(), wk.size() * 4));
    const BpGeom G = make_bp_geom(g, ig);
    const dim3 grid(unsigned((ig.nx + 255) / 256), unsigned(ig.ny), unsigned(ig.nz));
    if (kind == LCT_BP_PI_ORI)
        hipLaunchKernelGGL(bp_pi_ori_literal_kernel, grid, dim3(256), 0, ctx->stream, d_q, d_cs,
                           d_cs + nb, d_win, d_img, G);
    else
        hipLaunchKernelGGL(bp_helical_literal_kernel, grid, dim3(256), 0, ctx->stream, d_q, d_cs,
                           d_cs + nb, d_win, d_img, G);
    LCT_CHECK_HIP(hipGetLastError());
    return LCT_OK;
}

// ------------------------------------------- allocation-free backprojector
// lct_backproject_noalloc keeps its tables (cos and sin of the block's tube
// angles, then the 32 slices' view windows) in a file-scope device array of
// fixed size (bp_noalloc_scratch, 128.4 KiB): one copy per GPU, held until
// the process ends, shared by every call on that GPU.  The array's name is
// only a host-side stand-in; its address on each GPU comes from
// hipGetSymbolAddress, taken once per device when a context is created there
// (noalloc_scratch_init), and host code passes only that address to kernels.  The tables reach the array through
// kernel arguments (no staging buffer, no copy from host memory).
constexpr int kNoAllocMaxAngles = 16384;  // n_angle + n_proj_turn / 2 + 1 at most
constexpr size_t kNoAllocBytes = size_t(2) * kNoAllocMaxAngles * 4 + 3 * 32 * 4;
static __device__ __align__(256) unsigned char bp_noalloc_scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
void *g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

// Up to kArgWords 32-bit words of table through a kernel argument, split
// between two destinations: words [0, n0) -> dst0, [n0, n0 + n1) -> dst1.
constexpr int kArgWords = 960;
struct WordChunk {
    unsigned v[kArgWords];
};

__global__ __launch_bounds__(256) void fill_words_kernel(WordChunk c, int n0, unsigned *dst0, int n1,
                                                         unsigned *dst1)
{
    const int i = int(blockIdx.x) * 256 + int(threadIdx.x);
    if (i < n0)
        dst0[i] = c.v[i];
    else if (i < n0 + n1)
        dst1[i - n0] = c.v[i];
}

int backproject_block_noalloc(lct_ctx *ctx, int kind, const lct_ct_geom &g, const lct_img_geom &ig,
                              const float *d_q, const float *h_betas, float *d_img)
{
    if (ig.nz != LCT_BLOCK_SLICES)
        return set_error(LCT_ERR_INVALID, "backproject: ImgGeom nz must be 32 (BLOCK_SLICES)");
    const int nb = g.n_angle + g.n_proj_turn / 2 + 1;
    if (nb > kNoAllocMaxAngles)
        return set_error(LCT_ERR_UNSUPPORTED,
                         "lct_backproject_noalloc: n_angle + n_proj_turn/2 + 1 = " + std::to_string(nb) +
                             " exceeds its fixed scratch of " + std::to_string(kNoAllocMaxAngles) +
                             " angles (use lct_backproject)");
    if (ctx->device < 0 || ctx->device >= kMaxDevices || !g_scratch_ptr[ctx->device])
        return set_error(LCT_ERR_INVALID, "lct_backproject_noalloc: scratch of this device not set up");
    float *d_cos = static_cast<float *>(g_scratch_ptr[ctx->device]);
    float *d_sin = d_cos + nb;
    int *d_win = reinterpret_cast<int *>(d_sin + nb);
    WordChunk c;
    // cos / sin of each float32 tube angle, evaluated in double (as angle_tables)
    constexpr int kHalf = kArgWords / 2;
    for (int i0 = 0; i0 < nb; i0 += kHalf) {
        const int n = std::min(kHalf, nb - i0);
        for (int i = 0; i < n; i++) {
            c.v[i] = __builtin_bit_cast(unsigned, float(std::cos(double(h_betas[i0 + i]))));
            c.v[n + i] = __builtin_bit_cast(unsigned, float(std::sin(double(h_betas[i0 + i]))));
        }
        hipLaunchKernelGGL(fill_words_kernel, dim3(unsigned((2 * n + 255) / 256)), dim3(256), 0,
                           ctx->stream, c, n, reinterpret_cast<unsigned *>(d_cos + i0), n,
                           reinterpret_cast<unsigned *>(d_sin + i0));
        LCT_CHECK_HIP(hipGetLastError());
    }
    // kernel layout of the windows: (a_min, a_max) of the 32 slices, then a_middle[32]
    int amin[32], amax[32], amid[32];
    bp_windows(g, ig, amin, amax, amid);
    for (int iz = 0; iz < 32; iz++) {
        c.v[2 * iz] = unsigned(amin[iz]);
        c.v[2 * iz + 1] = unsigned(amax[iz]);
        c.v[64 + iz] = unsigned(amid[iz]);
    }
    hipLaunchKernelGGL(fill_words_kernel, dim3(1), dim3(256), 0, ctx->stream, c, 96,
                       reinterpret_cast<unsigned *>(d_win), 0, static_cast<unsigned *>(nullptr));
    LCT_CHECK_HIP(hipGetLastError());
    const BpGeom G = make_bp_geom(g, ig);
    const dim3 grid(unsigned((ig.nx + 255) / 256), unsigned(ig.ny), unsigned(ig.nz));
    if (kind == LCT_BP_PI_ORI)
        hipLaunchKernelGGL(bp_pi_ori_literal_kernel, grid, dim3(256), 0, ctx->stream, d_q, d_cos,
                           d_sin, d_win, d_img, G);
    else
        hipLaunchKernelGGL(bp_helical_literal_kernel, grid, dim3(256), 0, ctx->stream, d_q, d_cos,
                           d_sin, d_win, d_img, G);
    LCT_CHECK_HIP(hipGetLastError());
    return LCT_OK;
}

bool geom_ok(const lct_ct_geom *g)
{
    return g && g->ns > 0 && g->nt > 0 && g->n_angle > 0 && g->n_angle_add >= g->n_angle &&
           g->n_proj_turn > 1 && g->dso > 0.0f && g->ds > 0.0f && g->dt > 0.0f &&
           g->orbit != 0.0f;
}

}  // namespace

namespace lct {
// Device address of the allocation-free backprojector's scratch array on
// `device` (the current device), taken once per device.
int noalloc_scratch_init(int device)
{
    if (device < 0 || device >= kMaxDevices)
        return set_error(LCT_ERR_UNSUPPORTED, "device index beyond 64");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return LCT_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(bp_noalloc_scratch));
    if (e != hipSuccess) return hip_error(e, "hipGetSymbolAddress(bp_noalloc_scratch)");
    g_scratch_ptr[device] = p;
    return LCT_OK;
}

int backproject_block_public(lct_ctx *ctx, int kind, const lct_ct_geom &g,
                             const lct_img_geom &ig, const float *d_q, const float *h_betas,
                             float *d_img)
{
    return backproject_block(ctx, kind, g, ig, d_q, h_betas, d_img);
}
}  // namespace lct

extern "C" {

int lct_rebin(lct_ctx *ctx, int kind, int weight, const lct_ct_geom *g, const float *d_fan,
              float *d_par)
{
    if (!ctx || !geom_ok(g) || !d_fan || !d_par || kind < 0 || kind > 2 || weight < 0 ||
        weight > 1)
        return set_error(LCT_ERR_INVALID, "lct_rebin: bad argument");
    LCT_CHECK_HIP(hipSetDevice(ctx->device));
    if (kind == LCT_REBIN_NULL && weight == LCT_WEIGHT_NONE) {
        // NullRebinerGPU: par[ia] = fan[ia + add] (contiguous views).
        const size_t view = size_t(g->nt) * g->ns;
        LCT_CHECK_HIP(hipMemcpyAsync(d_par, d_fan + size_t(g->add_projections) * view,
                                     view * g->n_angle * sizeof(float), hipMemcpyDeviceToDevice,
                                     ctx->stream));
        return LCT_OK;
    }
    return launch_rebin(ctx, kind, weight, g, d_fan, 0, g->n_angle_add, d_par, g->n_angle,
                        g->add_projections, 0);
}

int lct_weighting(lct_ctx *ctx, int kind, const lct_ct_geom *g, float *d_par)
{
    if (!ctx || !geom_ok(g) || !d_par || kind < 0 || kind > 1)
        return set_error(LCT_ERR_INVALID, "lct_weighting: bad argument");
    if (kind == LCT_WEIGHT_NONE) return LCT_OK;  // NULLPreWeighting, PreWeighting.h:18
    LCT_CHECK_HIP(hipSetDevice(ctx->device));
    return launch_weighting(ctx, g, d_par);
}

int lct_filter(lct_ctx *ctx, const lct_ct_geom *g, const float *h_taps, float *d_rows, long rows)
{
    if (!ctx || !g || g->ns <= 0 || !h_taps || !d_rows || rows < 0)
        return set_error(LCT_ERR_INVALID, "lct_filter: bad argument");
    LCT_CHECK_HIP(hipSetDevice(ctx->device));
    void *dt = nullptr;
    LCT_CHECK(scratch(ctx, kScratchTaps, size_t(2) * g->ns * sizeof(float), &dt));
    LCT_CHECK(upload(ctx, dt, h_taps, size_t(2) * g->ns * sizeof(float)));
    return launch_filter_direct(ctx, g->ns, static_cast<const float *>(dt), d_rows, rows);
}

int lct_backproject(lct_ctx *ctx, int kind, const lct_ct_geom *g, const lct_img_geom *ig,
                    const float *d_q, const float *h_betas, float *d_img)
{
    if (!ctx || !geom_ok(g) || !ig || !d_q || !h_betas || !d_img || kind < 0 || kind > 1 ||
        ig->nx <= 0 || ig->ny <= 0)
        return set_error(LCT_ERR_INVALID, "lct_backproject: bad argument");
    LCT_CHECK_HIP(hipSetDevice(ctx->device));
    return backproject_block(ctx, kind, *g, *ig, d_q, h_betas, d_img);
}

int lct_backproject_noalloc(lct_ctx *ctx, int kind, const lct_ct_geom *g, const lct_img_geom *ig,
                            const float *d_q, const float *h_betas, float *d_img)
{
    if (!ctx || !geom_ok(g) || !ig || !d_q || !h_betas || !d_img || kind < 0 || kind > 1 ||
        ig->nx <= 0 || ig->ny <= 0)
        return set_error(LCT_ERR_INVALID, "lct_backproject_noalloc: bad argument");
    LCT_CHECK_HIP(hipSetDevice(ctx->device));
    return backproject_block_noalloc(ctx, kind, *g, *ig, d_q, h_betas, d_img);
}

}  // extern "C"
