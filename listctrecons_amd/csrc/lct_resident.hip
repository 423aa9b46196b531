// lct_resident.hip -- the HBM-resident pipeline and its backprojector.
//
// Every parallel view the planned blocks need is rebinned once and filtered
// once into Q, an HBM-resident array in the backprojector's layout:
//     Q[view][slot_ch][slot_row],  slot_ch = channel + 1  (cp = ns + 3 slots),
//                                  slot_row = row + 4     (rp = nt + 8 slots, nt % 8 == 0;
//                                                          f32: rp = nt + 9, odd)
// with zero guard slots, stored as f32, f16 or bf16.  A view is then one
// contiguous block of cp*rp elements, and the channel window a tile of voxel
// columns needs is one contiguous range.
//
// Backprojector (CUDA/Backprojector.cu:12-99 restructured).  In a voxel's
// window [a_middle - h, a_middle + h), h = npt/2, every view pair (j, j+h)
// is visited twice by the reference: once as (ia=j, ia2=j+h) and once as
// (ia=j+h, ia2=j).  With tube angles beta_{j+h} = beta_j + pi the two visits
// sample the same two rays with swapped weights, so each pair is evaluated
// once and counted twice; when both rays leave the detector the two visits
// pick different rays, and the pair then counts each ray once.  Per
// (column, pair) the z-independent terms (asin, sqrt, channel taps, source
// z) are computed once and reused for every slice of the 32-slice block
// whose window contains the pair.  Geometries where the folding does not
// hold (odd npt, clipped windows, irregular angles) use the per-voxel
// kernel of lct_stages.hip on the same resident arrays.
//
// The PI-original backprojector (CUDA/Backprojector.cu:171-250) folds the
// same pairs, a_middle - h < j <= a_middle: the reference walks back from
// a_middle while a view's ray lies in the voxel's PI window (|t| <= P/4) and
// then forward to half a turn after the first view outside it, so of each
// pair the voxel takes view j when its ray is in the window and view j + h
// otherwise (exact whenever the window test does not change sign twice along
// the walk).  Its rows are linear in z as the helical ones, and the window
// table marks the slices where every column of a tile takes the same view.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "lct_internal.hpp"

namespace lct {

constexpr int kRowPad = 4;  // zero row slots before row 0
// Backprojector tile: 16 (x) x 32 (y) voxel columns per 512-thread workgroup.
// Against 16 x 16, the staged window per column shrinks by a quarter and a
// thread stages 2 chunks per view pair instead of 3 (clinical f16 launch
// 33.4 -> 31.8 ms).
constexpr int kTile = 16;
constexpr int kTileY = 32;
constexpr int kBpThreads = kTile * kTileY;
// most pair ranges a (block, tile) is split into
constexpr int kMaxParts = 4;

template <typename T> struct Cvt;
template <> struct Cvt<float> {
    static __device__ __forceinline__ float to(float x) { return x; }
    static __device__ __forceinline__ float from(float x) { return x; }
};
template <> struct Cvt<__half> {
    static __device__ __forceinline__ __half to(float x) { return __float2half_rn(x); }
    static __device__ __forceinline__ float from(__half x) { return __half2float(x); }
};
template <> struct Cvt<__hip_bfloat16> {
    static __device__ __forceinline__ __hip_bfloat16 to(float x) { return __float2bfloat16(x); }
    static __device__ __forceinline__ float from(__hip_bfloat16 x) { return __bfloat162float(x); }
};

// --------------------------------------------------------------- filter

// Direct FP32 ramp convolution (CUDA/Filter.cu:42-46 order: k ascending,
// FMA) of 4 consecutive rows of one view, written channel-major into Q.
template <int J, typename T>
__global__ __launch_bounds__(256) void filter_to_q_kernel(const float *__restrict__ taps,
                                                          const float *__restrict__ in,
                                                          T *__restrict__ q, long n_rows, int ns,
                                                          int nt, int cp, int rp)
{
    constexpr int R = 4;
    extern __shared__ float sm[];
    float *st = sm;
    float *sr = sm + 2 * ns;
    const long r0 = long(blockIdx.x) * R;
    for (int i = threadIdx.x; i < 2 * ns; i += blockDim.x) st[i] = taps[i];
    for (int i = threadIdx.x; i < R * ns; i += blockDim.x) {
        const long rr = r0 + i / ns;
        sr[i] = rr < n_rows ? in[size_t(rr) * ns + (i % ns)] : 0.0f;
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
        hreg[0] = st[is0 - k - 1 + ns];
    }
    if (r0 >= n_rows) return;
    const long v = r0 / nt;
    const int it0 = int(r0 - v * nt);  // multiple of 4 (nt % 8 == 0)
    T *dst = q + size_t(v) * cp * rp + it0 + kRowPad;
#pragma unroll
    for (int j = 0; j < J; j++) {
        if (is0 + j >= ns) break;
        T *d = dst + size_t(is0 + j + 1) * rp;
#pragma unroll
        for (int r = 0; r < R; r++) d[r] = Cvt<T>::to(acc[r][j]);
    }
}

template <typename T>
static int launch_filter_to_q(lct_ctx *ctx, int ns, int nt, const float *d_taps,
                              const float *d_in, T *d_q, long rows, int cp, int rp)
{
    const int J = (ns + 255) / 256;
    const size_t shm = size_t(2 * ns + 4 * ns) * sizeof(float);
    const dim3 grid(unsigned((rows + 3) / 4));
#define LCT_FQ_CASE(JJ)                                                                        \
    case JJ:                                                                                   \
        hipLaunchKernelGGL((filter_to_q_kernel<JJ, T>), grid, dim3(256), shm, ctx->stream,     \
                           d_taps, d_in, d_q, rows, ns, nt, cp, rp);                           \
        break;
    switch (J) {
        LCT_FQ_CASE(1)
        LCT_FQ_CASE(2)
        LCT_FQ_CASE(3)
        LCT_FQ_CASE(4)
        LCT_FQ_CASE(5)
        LCT_FQ_CASE(6)
        LCT_FQ_CASE(7)
        LCT_FQ_CASE(8)
    default:
        return set_error(LCT_ERR_UNSUPPORTED, "filter: ns > 2048 not supported");
    }
#undef LCT_FQ_CASE
    LCT_CHECK_HIP(hipGetLastError());
    return LCT_OK;
}

// ------------------------------------------------------- backprojector

struct FoldBpArgs {
    const void *q;
    size_t view_stride;  // cp * rp elements
    int rp, cp, W, H;    // channel window of W = 2H + 2 slots
    int lp;              // LDS row pitch of a staged channel (rp + 1, or rp for D kernels:
                         // odd, so the channels that neighbouring lanes read fall in
                         // different banks)
    int asin_poly;       // every voxel column has |p_hat| / dso <= 0.7: asin_poly is valid
    int units;           // D kernels: 16-byte units copied per view window (see bp_fold_kernel)
    int n_parts;         // pairs of a block in n_parts ranges [jlo + split[p], jlo + split[p + 1])
    int split[kMaxParts + 1];
    float *part;         // sums of ranges 1 .. n_parts - 1, one call volume each (range 0: vol)
    size_t part_stride;  // floats per call volume
    const int *slab_tab;       // per block: qoff, slice0, a_middle[32]
    int n_pairs;               // pairs j of a block: a_middle[0] - h <= j < a_middle[31]
                               // (PI-original: a_middle[0] - h < j <= a_middle[31])
    float *vol;
    int n_tiles_x, n_tiles_y, n_wg, xcd_remap;
    int nx, ny, h;
    float wx, wy, wz, dx, dy, dz;
    float inv_dso, dso_ds, dso2, ws, wt, inv_orbit, z0, dzv, kc, ns_f, nt_f, scale;
    float inv_ds;           // PI-original channels: p = s / ds + ws
    int pi;                 // bp_windows_kernel: PI-original geometry and masks
    unsigned *winb;         // PI-original: per (block, tile, pair) slices taking view j + h alone
};

template <typename T> struct Chunk {
    static constexpr int kElems = 16 / int(sizeof(T));  // elements per 16-byte load
};

// Converts a 16-byte chunk of T to floats at a 4-byte aligned LDS address
// (two 16-bit values per 32-bit word, low half first).
template <typename T>
__device__ __forceinline__ void chunk_to_lds_odd(float *dst, const uint4 &v)
{
    if constexpr (sizeof(T) == 4) {
        dst[0] = __uint_as_float(v.x);
        dst[1] = __uint_as_float(v.y);
        dst[2] = __uint_as_float(v.z);
        dst[3] = __uint_as_float(v.w);
    } else {
        auto lo = [](unsigned w) { return Cvt<T>::from(__builtin_bit_cast(T, (unsigned short)(w & 0xffffu))); };
        auto hi = [](unsigned w) { return Cvt<T>::from(__builtin_bit_cast(T, (unsigned short)(w >> 16))); };
        dst[0] = lo(v.x);
        dst[1] = hi(v.x);
        dst[2] = lo(v.y);
        dst[3] = hi(v.y);
        dst[4] = lo(v.z);
        dst[5] = hi(v.z);
        dst[6] = lo(v.w);
        dst[7] = hi(v.w);
    }
}

// asin(u) for |u| <= 0.7 as u + u^3 P(u^2), P of degree 6: coefficients from
// a weighted least-squares fit of (asin(u) - u) / u^3 on Chebyshev nodes in
// u^2, reweighted toward equal relative error.  Evaluated in FP32 it stays
// within 0.83 ulp of the exact asin over the whole interval.
__device__ __forceinline__ float asin_poly(float u)
{
    const float t = u * u;
    float p = 0.106134832f;
    p = fmaf(p, t, -0.0844023898f);
    p = fmaf(p, t, 0.0719100833f);
    p = fmaf(p, t, 0.0179117583f);
    p = fmaf(p, t, 0.0462728813f);
    p = fmaf(p, t, 0.0749005452f);
    p = fmaf(p, t, 0.166668698f);
    return fmaf(u * t, p, u);
}

// First channel slot of the W-slot window of a view, from the channel
// coordinate c of the tile centre (window covers every tap of the tile).
__device__ __forceinline__ int window_start(float c, int H, int cp, int W)
{
    int s = int(floorf(c)) + 1 - H;
    s = s < 0 ? 0 : s;
    return s > cp - W ? cp - W : s;
}

// The staged windows are held in LDS as floats with an odd channel pitch lp
// (A.lp = rp + 1): neighbouring voxel columns read neighbouring channels, and
// an odd pitch spreads those reads over all LDS banks (the HBM pitch rp is a
// multiple of 8, which would put them in a quarter of the banks).  LP > 0:
// lp known at compile time, so the two channels of a bilinear tap are read
// with immediate LDS offsets from one address.
//
// G (LCT_INTERP_GATHER): the same arithmetic with every tap gathered from Q
// in HBM through the L1 / L2 caches instead: no staging, no barriers; the
// zero guard slots of Q give the reference's border behaviour.
// cs: (cos, sin) of each Q view's tube angle; zmask: per pair j - jlo, bit z
// set when a_middle[z] - h <= j < a_middle[z] (the pair lies in slice z's
// window; the same for every block); win: per (block, tile, pair) the first
// channel slots of the two views' LDS windows (.x = s1 | s2 << 16) and the
// slices on which every column of the tile has both rays on the detector
// (.y, see bp_windows_kernel).  All three are read with wave-uniform
// indices, so they arrive in scalar registers.
// f32 storage staged through registers takes 4 chunks per thread and would
// use 84 VGPRs (2 workgroups per CU); capped at 80 it keeps 3 (4 spilled).
// 16-bit storage needs 71 and no cap.
//
// D (f32 storage only): the windows are copied from Q into LDS by LDS-DMA
// loads (global_load_lds_dwordx4), with no registers or conversion in
// between.  Q then carries the odd pitch itself (rp = lp = nt + 9), so a
// view's window is one contiguous range, copied in 16-byte units from the
// 16-byte boundary at or below its start: the window's first element lands
// `shift` = (start mod 4) floats into its LDS region.  The two regions are
// 4 Up floats apart (Up = units per view rounded up to 64), so each wave's
// 64 units belong to one view and fill 1 KB of LDS in lane order.
//
// PI (the PI-original backprojector, Backprojector.cu:171-250): channels
// s / ds + ws, rows from t = root (zc - zr) / lhat +- gamma P / 2 pi, one view
// of each pair per voxel (see the file header).  Its slices carry one tap, so
// latency rather than issue limits it: it is held to 64 VGPRs, 8 waves per
// SIMD (4 workgroups per CU instead of 3 at 67 VGPRs; the two 8-byte values
// spilled are stored before the pair loop and reloaded after it): 21.0 ->
// 18.6 ms on the clinical volume, 245.7 -> 219.5 ms on the large one.  The
// helical form would spill 15 registers at 64, and keeps 6 waves.
template <typename T, int NPRE, int LP, bool G = false, bool D = false, bool PI = false>
__global__ __launch_bounds__(kBpThreads)
__attribute__((amdgpu_waves_per_eu(PI && !G ? 8 : ((sizeof(T) == 4 && !D) ? 6 : 1)))) void bp_fold_kernel(FoldBpArgs A, const float2 *__restrict__ cs,
                                                      const unsigned *__restrict__ zmask,
                                                      const uint2 *__restrict__ win)
{
    extern __shared__ float lds[];
    constexpr int CE = Chunk<T>::kElems;
    const int W = A.W, rp = A.rp, lp = LP > 0 ? LP : A.lp;
    const int view_elems = W * rp;          // one view's window in Q
    const int lds_view = W * lp;            // and in LDS (views j, j+h: one buffer of 2 W lp)
    // 16-byte chunks: view j's at chunk indices [0, n_vc), view j+h's at
    // [pad, pad + n_vc) with pad a multiple of 64, so each wave's chunks of a
    // slot belong to one view and share a scalar base address
    const int n_vc = view_elems / CE;
    const int pad = (n_vc + 63) & ~63;

    // XCD-aware order: consecutive workgroups are dealt round-robin to the 8
    // XCDs; give each XCD a contiguous range of tiles so the views it stages
    // stay in its own L2.
    int wg = blockIdx.x;
    if (A.xcd_remap) wg = (wg & 7) * (A.n_wg >> 3) + (wg >> 3);
    const int tiles = A.n_tiles_x * A.n_tiles_y;
    const int blk = wg / (tiles * A.n_parts);
    const int rem = wg - blk * tiles * A.n_parts;
    const int part = rem / tiles;  // which range of the block's pairs
    const int t = rem - part * tiles;
    const int tx0 = (t % A.n_tiles_x) * kTile, ty0 = (t / A.n_tiles_x) * kTileY;
    const int ix = tx0 + int(threadIdx.x & (kTile - 1));
    const int iy = ty0 + int(threadIdx.x / kTile);
    const int *st = A.slab_tab + blk * 34;
    const int qoff = st[0], slice0 = st[1];

    const float xc = (float(ix) - A.wx) * A.dx;
    const float yc = (float(iy) - A.wy) * A.dy;
    const T *Q = static_cast<const T *>(A.q);

    float acc[32];
#pragma unroll
    for (int z = 0; z < 32; z++) acc[z] = 0.0f;

    // pairs of any slice's window: a_middle[0] - h <= j < a_middle[31]; this
    // workgroup takes [jb, je) of them (the tables are indexed from jlo)
    // PI-original pairs: a_middle[0] - h < j <= a_middle[31] (one later)
    const int jlo = st[2] - A.h + (PI ? 1 : 0), jhi = st[33] + (PI ? 1 : 0);
    const int jb = jlo + A.split[part];
    const int je = part == A.n_parts - 1 ? jhi : jlo + A.split[part + 1];

    uint4 pre[NPRE];
    // Chunk slot i of this thread: chunk index c = tid + kBpThreads i, view
    // (c >= pad, the same on every lane of a wave), element offset in the
    // view's window and LDS offset, all fixed for the whole loop.  Chunk m
    // of a view holds CE consecutive rows of one channel (rp is a multiple
    // of 8): channel m % W, row block m / W.  Consecutive lanes thus take
    // consecutive channels, whose LDS addresses differ by the odd pitch lp
    // and fall in different banks when the chunk is written.  Both offsets
    // fit 16 bits (a window holds at most 2 * W * lp < 65536 floats) and
    // share one register.
    unsigned offs[NPRE];
    bool live[NPRE];
#pragma unroll
    for (int i = 0; i < NPRE; i++) {
        const int c = int(threadIdx.x) + i * kBpThreads;
        const int v = c >= pad ? 1 : 0;
        const int m = c - v * pad;
        live[i] = m < n_vc;
        const int mm = live[i] ? m : 0;
        const int ch = mm % W, rb = (mm / W) * CE;
        offs[i] = unsigned(ch * rp + rb) | (unsigned(v * lds_view + ch * lp + rb) << 16);
    }
    // chunk slot i belongs to view j + h on every lane of a wave or on none
    bool vbw[NPRE];
#pragma unroll
    for (int i = 0; i < NPRE; i++)
        vbw[i] = __builtin_amdgcn_readfirstlane(int(threadIdx.x) + i * kBpThreads) >= pad;
    const size_t hstride = size_t(A.h) * A.view_stride;
    // v0: first element of view j in Q.  Its running pointer qj (advanced by
    // one view per pair) replaces a 64-bit multiply per view, which took
    // ~40 scalar instructions per pair (each costs VALU issue slots here:
    // backprojection 28.6 -> 28.2 ms on the clinical volume, same results)
    auto fetch = [&](const T *v0, int s1, int s2) {
        const T *va = v0 + unsigned(s1 * rp);
        const T *vb = v0 + hstride + unsigned(s2 * rp);
#pragma unroll
        for (int i = 0; i < NPRE; i++) {
            const T *base = vbw[i] ? vb : va;
            if (live[i]) pre[i] = *reinterpret_cast<const uint4 *>(base + (offs[i] & 0xffffu));
        }
    };
    const T *qj = Q + size_t(qoff + jb) * A.view_stride;
    auto commit = [&](float *buf) {
#pragma unroll
        for (int i = 0; i < NPRE; i++)
            if (live[i]) chunk_to_lds_odd<T>(buf + (offs[i] >> 16), pre[i]);
    };

    // D: LDS-DMA copy of pair j's two windows (views j and j + h)
    const int Up = (A.units + 63) & ~63;
    auto first_elem = [&](int v, int sl) -> size_t {
        return size_t(qoff + v) * A.view_stride + size_t(sl) * rp;
    };
    auto dma = [&](int j, int s1, int s2) {
        const float *Qf = static_cast<const float *>(A.q);
        const size_t ea = first_elem(j, s1) & ~size_t(3), eb = first_elem(j + A.h, s2) & ~size_t(3);
        const int lane = int(threadIdx.x & 63);
        const int u_wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x) & ~63);
        for (int u0 = u_wave; u0 < 2 * Up; u0 += kBpThreads) {
            const bool vb = u0 >= Up;  // the same on every lane of the wave
            const int lu = u0 - (vb ? Up : 0) + lane;
            if (lu < A.units)
                __builtin_amdgcn_global_load_lds(
                    (__attribute__((address_space(1))) void *)(Qf + (vb ? eb : ea) + 4 * size_t(lu)),
                    (__attribute__((address_space(3))) void *)(lds + 4 * u0), 16, 0, 0);
        }
    };

    const float2 *csq = cs + qoff;
    const uint2 *wt = win + size_t(blk * tiles + t) * A.n_pairs;  // index j - jlo
    const unsigned *wtb = PI ? A.winb + size_t(blk * tiles + t) * A.n_pairs : nullptr;
    if constexpr (D) {
        const unsigned w = wt[jb - jlo].x;
        dma(jb, int(w & 0xffffu), int(w >> 16));
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        __syncthreads();
    } else if constexpr (!G) {
        const unsigned w = wt[jb - jlo].x;
        fetch(qj, int(w & 0xffffu), int(w >> 16));
        commit(lds);
        __syncthreads();
    }
    for (int j = jb; j < je; j++) {
        const float2 cs1 = csq[j], cs2 = csq[j + A.h];
        const float c = cs1.x, s = cs1.y, c2 = cs2.x, s2v = cs2.y;
        const unsigned zm = zmask[j - jlo];
        const unsigned fm = wt[j - jlo].y;
        const unsigned fm2 = PI ? wtb[j - jlo] : 0u;
        int s1 = 0, s2 = 0;
        const float *buf = lds;
        int view_off = lds_view;  // LDS offset of view j + h's window from view j's
        if constexpr (!G) {
            const unsigned w = wt[j - jlo].x;
            s1 = int(w & 0xffffu);
            s2 = int(w >> 16);
            if constexpr (D) {
                const int sa = int(first_elem(j, s1) & 3), sb = int(first_elem(j + A.h, s2) & 3);
                buf = lds + sa;
                view_off = 4 * Up + sb - sa;
            } else {
                // prefetch pair j + 1 into registers (the last pair is fetched twice)
                const int jn = min(j + 1, je - 1);
                const unsigned wn = wt[jn - jlo].x;
                fetch(j + 1 < je ? qj + A.view_stride : qj, int(wn & 0xffffu), int(wn >> 16));
            }
        }
        // z-independent terms of the pair (Backprojector.cu:53-62)
        const float phat = -xc * c - yc * s;
        const float u = phat * A.inv_dso;
        const float theta = A.asin_poly ? asin_poly(u) : asinf(u);
        // channel offset: theta dso / ds, or s / ds for the PI-original data
        // (uniform in s, Backprojector.cu:215-216)
        const float delta = PI ? phat * A.inv_ds : theta * A.dso_ds;
        const float p1 = fminf(fmaxf(A.ws + delta, -1.0f), A.ns_f);
        const float p2 = fminf(fmaxf(A.ws - delta, -1.0f), A.ns_f);
        const float fp1 = floorf(p1), fp2 = floorf(p2);
        const float ap1 = p1 - fp1, ap2 = p2 - fp2;
        // local channel slots; the window covers every tap of the tile, the
        // clamp only keeps LDS addressing inside the stage buffer
        const int l1 = min(max(int(fp1) + 1 - s1, 0), W - 2);
        const int l2 = min(max(int(fp2) + 1 - s2, 0), W - 2);
        const float *b1 = buf + __mul24(l1, lp);
        const float *b2 = buf + view_off + __mul24(l2, lp);
        // gather: channel slot = channel + 1 (slots 0, ns + 1 and ns + 2 are zero)
        const T *g1 = Q + size_t(qoff + j) * A.view_stride + size_t(int(fp1) + 1) * rp;
        const T *g2 = Q + size_t(qoff + j + A.h) * A.view_stride + size_t(int(fp2) + 1) * rp;
        const float root = __builtin_amdgcn_sqrtf(A.dso2 - phat * phat);
        const float lhat = root + xc * s - yc * c;
        const float lhat2 = root + xc * s2v - yc * c2;
        const float tz = theta * A.inv_orbit;
        const float zr = A.z0 + (float(j) + tz) * A.dzv;
        const float zr2 = A.z0 + (float(j + A.h) - tz) * A.dzv;
        // PI-original rows (Backprojector.cu:207-216): t = root (zc - zr) / lhat +
        // gamma P / 2 pi, row t 2 (nt - 1) / P + wt, with kc = z_dir 2 (nt - 1) / P
        const float kr = PI ? A.kc * root : A.kc;
        const float k1 = kr * __builtin_amdgcn_rcpf(lhat), k2 = kr * __builtin_amdgcn_rcpf(lhat2);
        // detector row offsets qh(z) = (zc(z) - zr) k1 as z K1 + O1 with zc(z) = (z - wz) dz:
        // the slice index is a literal, so no per-slice coordinate is held in a register
        const float K1 = k1 * A.dz, K2 = k2 * A.dz;
        // and the row coordinates in slot units (row + kRowPad) as r(z) = z K1 + R1,
        // R1 = O1 + row0, row0 = wt + kRowPad the slot of qh = 0
        const float row0 = A.wt + float(kRowPad);
        float R1, R2;
        if constexpr (PI) {
            const float gp = A.kc * tz * A.dzv;  // gamma P / 2 pi in rows (the ray through j + h: -gamma)
            R1 = fmaf(-A.wz, K1, fmaf(-zr, k1, row0 + gp));
            R2 = fmaf(-A.wz, K2, fmaf(-zr2, k2, row0 - gp));
        } else {
            R1 = fmaf(-A.wz, K1, fmaf(-zr, k1, row0));
            R2 = fmaf(-A.wz, K2, fmaf(-zr2, k2, row0));
        }
        // off the detector, a row coordinate is clamped to the guard slots
        // [kRowPad - 1, nt + kRowPad]: positive, so truncation is floor
        const float lo_r = float(kRowPad) - 1.0f, hi_r = A.nt_f + float(kRowPad);
        if constexpr (PI) {
            // Backprojector.cu:200-245: a voxel takes view j when that ray lies in
            // its PI window (|t| <= P / 4: on the detector rows), else view j + h.
            // The window table marks the slices where every column of the tile
            // takes view j (fm, ray on the detector) or view j + h (fm2, ray j
            // outside and ray j + h on the detector): one tap, no clamps.
            auto tap = [&](const float *bl, const T *gl, float ap, float r) -> float {
                const int i = int(r);
                const float ar = __builtin_amdgcn_fractf(r);
                float v00, v01, v10, v11;
                if constexpr (G) {
                    const T *q = gl + i;
                    v00 = Cvt<T>::from(q[0]);
                    v01 = Cvt<T>::from(q[1]);
                    v10 = Cvt<T>::from(q[rp]);
                    v11 = Cvt<T>::from(q[rp + 1]);
                } else {
                    const float *q = bl + i;
                    v00 = q[0];
                    v01 = q[1];
                    v10 = q[lp];
                    v11 = q[lp + 1];
                }
                const float lo = fmaf(ap, v10 - v00, v00), hi = fmaf(ap, v11 - v01, v01);
                return fmaf(ar, hi - lo, lo);
            };
            // One loop per slice class, each with one scalar bit test per slice; a
            // slice is in one class, so each is still summed once, in pair order.
            // (Tested in one loop as an if / else-if chain, the three classes
            // shared one tail behind register moves and flow-control branches:
            // 22.3 -> 21.1 ms on the clinical volume, 279 -> 245 ms on the large one.)
            const unsigned m1 = zm & fm, m2 = zm & ~fm & fm2, mx = zm & ~fm & ~fm2;
#pragma unroll
            for (int z = 0; z < 32; z++)
                if (m1 & (1u << z)) acc[z] += tap(b1, g1, ap1, fmaf(float(z), K1, R1));
            if (m2) {
#pragma unroll
                for (int z = 0; z < 32; z++)
                    if (m2 & (1u << z)) acc[z] += tap(b2, g2, ap2, fmaf(float(z), K2, R2));
            }
            if (mx) {
#pragma unroll
                for (int z = 0; z < 32; z++) {
                    if (mx & (1u << z)) {
                        // per column: the view first, then its one tap (rows
                        // clamped to the zero guard slots: the border)
                        const float r1 = fmaf(float(z), K1, R1), r2 = fmaf(float(z), K2, R2);
                        const bool in1 = fabsf(r1 - row0) <= A.wt;
                        acc[z] += tap(in1 ? b1 : b2, in1 ? g1 : g2, in1 ? ap1 : ap2,
                                      __builtin_amdgcn_fmed3f(in1 ? r1 : r2, lo_r, hi_r));
                    }
                }
            }
        } else {
    #pragma unroll
            for (int z = 0; z < 32; z++) {
                if (zm & (1u << z)) {  // pair inside this slice's window
                    float r1 = fmaf(float(z), K1, R1);
                    float r2 = fmaf(float(z), K2, R2);
                    // conjugate weights (w1, 1 - w1), w1 = qh2 / (qh2 - qh) with qh = r - row0
                    float w1 = (row0 - r2) * __builtin_amdgcn_rcpf(r1 - r2);
                    // keeps w1 computed on every lane, so the selects below stay
                    // selects instead of divergent branches (emits no instruction)
                    asm volatile("" : "+v"(w1));
                    // fm bit z (wave-uniform, from the window table): every column
                    // of the tile has both rays on the detector, so the row
                    // coordinates lie in the detector's slots and the detector
                    // tests, clamps and selects are skipped
                    if (!(fm & (1u << z))) {
                        // keeps these branches: not speculated into selects (no instruction)
                        asm volatile("");
                        // a ray off the detector (|qh| > wt) takes weight 0, or 1/2
                        // each when both are; when the compare masks show every
                        // lane of the wave with both rays on the detector, the
                        // clamps and selects are skipped
                        const bool out1 = fabsf(r1 - row0) > A.wt, out2 = fabsf(r2 - row0) > A.wt;
                        if (__builtin_amdgcn_ballot_w64(out1 || out2) != 0) {
                            asm volatile("");
                            w1 = out1 ? (out2 ? 0.5f : 0.0f) : (out2 ? 1.0f : w1);
                            r1 = __builtin_amdgcn_fmed3f(r1, lo_r, hi_r);
                            r2 = __builtin_amdgcn_fmed3f(r2, lo_r, hi_r);
                        }
                    }
                    const int i1 = int(r1), i2 = int(r2);
                    const float ar1 = __builtin_amdgcn_fractf(r1), ar2 = __builtin_amdgcn_fractf(r2);
                    float a00, a01, a10, a11, c00, c01, c10, c11;
                    if constexpr (G) {
                        const T *p1 = g1 + i1;
                        const T *p2 = g2 + i2;
                        a00 = Cvt<T>::from(p1[0]);
                        a01 = Cvt<T>::from(p1[1]);
                        a10 = Cvt<T>::from(p1[rp]);
                        a11 = Cvt<T>::from(p1[rp + 1]);
                        c00 = Cvt<T>::from(p2[0]);
                        c01 = Cvt<T>::from(p2[1]);
                        c10 = Cvt<T>::from(p2[rp]);
                        c11 = Cvt<T>::from(p2[rp + 1]);
                    } else {
                        const float *p1 = b1 + i1;
                        const float *p2 = b2 + i2;
                        a00 = p1[0];
                        a01 = p1[1];
                        a10 = p1[lp];
                        a11 = p1[lp + 1];
                        c00 = p2[0];
                        c01 = p2[1];
                        c10 = p2[lp];
                        c11 = p2[lp + 1];
                    }
                    const float la = fmaf(ap1, a10 - a00, a00), ha = fmaf(ap1, a11 - a01, a01);
                    const float lc = fmaf(ap2, c10 - c00, c00), hc = fmaf(ap2, c11 - c01, c01);
                    const float t1 = fmaf(ar1, ha - la, la);
                    const float t2 = fmaf(ar2, hc - lc, lc);
                    acc[z] = fmaf(w1, t1 - t2, acc[z] + t2);
                }
            }
        }
        if constexpr (D) {
            __syncthreads();
            if (j + 1 < je) {
                const unsigned wn = wt[j + 1 - jlo].x;
                dma(j + 1, int(wn & 0xffffu), int(wn >> 16));
            }
            // the copies count as vector-memory loads: this wave waits for
            // its own, the barrier for every wave's
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            __syncthreads();
        } else if constexpr (!G) {
            // one buffer: every wave is done with pair j before pair j + 1
            // overwrites it (half the LDS of double buffering, so five
            // workgroups fit a CU instead of four)
            __syncthreads();
            commit(lds);
            __syncthreads();
        }
        qj += A.view_stride;
    }
    if (ix < A.nx && iy < A.ny) {
        // several ranges: unscaled sums, added and scaled by bp_sum_parts_kernel
        float *o = (part ? A.part + size_t(part - 1) * A.part_stride : A.vol) +
                   (size_t(slice0) * A.ny + iy) * A.nx + ix;
        const float sc = A.n_parts == 1 ? A.scale : 1.0f;
        const size_t plane = size_t(A.nx) * A.ny;
#pragma unroll
        for (int z = 0; z < 32; z++) o[size_t(z) * plane] = acc[z] * sc;
    }
}

// vol = (vol + part[0] + ... + part[n_more - 1]) * scale over n floats, left
// to right: the pair ranges' sums of the backprojector in a fixed order
__global__ __launch_bounds__(256) void bp_sum_parts_kernel(float *__restrict__ vol,
                                                           const float *__restrict__ part,
                                                           size_t stride, int n_more, size_t n,
                                                           float scale)
{
    const size_t i = size_t(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    float v = vol[i];
    for (int p = 0; p < n_more; p++) v += part[size_t(p) * stride + i];
    vol[i] = v * scale;
}

// Detector row offsets qh(z) = z K + O of the two rays of pair j through
// voxel column (x, y) (fractional voxel indices), from the backprojector's
// per-pair terms (Backprojector.cu:53-70): ray 1 through view j (a = cos,
// sin), ray 2 through view j + h (b).  Hardware reciprocal and square root:
// the result only places slices against a margin of a quarter row.
__device__ __forceinline__ void pair_row_lines(const FoldBpArgs &A, float x, float y, float2 a,
                                               float2 b, int j, float &K1, float &O1, float &K2,
                                               float &O2)
{
    const float xc = (x - A.wx) * A.dx, yc = (y - A.wy) * A.dy;
    const float phat = -xc * a.x - yc * a.y;
    const float u = phat * A.inv_dso;
    const float theta = A.asin_poly ? asin_poly(u) : asinf(u);
    const float root = __builtin_amdgcn_sqrtf(A.dso2 - phat * phat);
    const float lhat = root + xc * a.y - yc * a.x;
    const float lhat2 = root + xc * b.y - yc * b.x;
    const float tz = theta * A.inv_orbit;
    const float zr = A.z0 + (float(j) + tz) * A.dzv;
    const float zr2 = A.z0 + (float(j + A.h) - tz) * A.dzv;
    // PI-original rows: kc = z_dir 2 (nt - 1) / P, t = root (zc - zr) / lhat +- gamma P / 2 pi
    const float kr = A.pi ? A.kc * root : A.kc;
    const float gp = A.pi ? A.kc * tz * A.dzv : 0.0f;
    const float k1 = kr * __builtin_amdgcn_rcpf(lhat), k2 = kr * __builtin_amdgcn_rcpf(lhat2);
    K1 = k1 * A.dz;
    K2 = k2 * A.dz;
    O1 = -A.wz * K1 - zr * k1 + gp;
    O2 = -A.wz * K2 - zr2 * k2 - gp;
}

// Narrows [zlo, zhi] to the slices z with lo <= z K + O <= hi (K != 0).
__device__ __forceinline__ void clip_line(float K, float O, float lo, float hi, float &zlo, float &zhi)
{
    const float r = __builtin_amdgcn_rcpf(K);
    const float za = (lo - O) * r, zb = (hi - O) * r;
    zlo = fmaxf(zlo, fminf(za, zb));
    zhi = fminf(zhi, fmaxf(za, zb));
}

// The slices 0..31 inside [zlo, zhi] (one within 1e-3 of an end is left out).
__device__ __forceinline__ unsigned slice_bits(float zlo, float zhi)
{
    const int z0 = int(ceilf(fminf(fmaxf(zlo + 1e-3f, 0.0f), 32.0f)));
    const int z1 = int(floorf(fmaxf(fminf(zhi - 1e-3f, 31.0f), -1.0f)));
    return z1 < z0 ? 0u : (z1 - z0 == 31 ? 0xffffffffu : ((1u << (z1 - z0 + 1)) - 1u) << z0);
}

// Per (block, tile, pair):
//  .x  first channel slots of the LDS windows: the tile centre's channel
//      coordinate through view j (first window) and its mirror 2 ws - c
//      through view j + h (second window), s1 | s2 << 16;
//  .y  bit z set when every voxel column of the tile has both rays of the
//      pair on the detector in slice z, with a quarter row to spare.
// qh is linear in z for each column, and smooth over the tile: a quadratic
// c4 x (1 - x) + c5 y (1 - y) away from the bilinear interpolation of its
// corner values (x, y in [0, 1] across the tile), so at most |c4|/4 + |c5|/4
// = dx + dy away, dx and dy its departures at the midpoints of two edges.
// With e = 2 (dx + dy) for the higher-order terms, |qh| <= wt - 1/4 - e at
// the four corners holds |qh| <= wt - 1/4 at every column.
__global__ __launch_bounds__(256) void bp_windows_kernel(FoldBpArgs A, int n_blocks,
                                                         const float2 *__restrict__ cs,
                                                         uint2 *__restrict__ win)
{
    const int tiles = A.n_tiles_x * A.n_tiles_y;
    const long n = long(n_blocks) * tiles * A.n_pairs;
    const long i = long(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    const int jj = int(i % A.n_pairs);
    const long bt = i / A.n_pairs;
    const int blk = int(bt / tiles), t = int(bt % tiles);
    const int tx0 = (t % A.n_tiles_x) * kTile, ty0 = (t / A.n_tiles_x) * kTileY;
    const float xcc = (float(tx0) + 0.5f * (kTile - 1) - A.wx) * A.dx;
    const float ycc = (float(ty0) + 0.5f * (kTileY - 1) - A.wy) * A.dy;
    const int *st = A.slab_tab + blk * 34;
    const int j = st[2] - A.h + (A.pi ? 1 : 0) + jj;
    const float2 c = cs[st[0] + j];
    const float ph = -xcc * c.x - ycc * c.y;
    const float c1 = A.ws + (A.pi ? ph * A.inv_ds : asinf(ph * A.inv_dso) * A.dso_ds);
    const int s1 = window_start(c1, A.H, A.cp, A.W);
    const int s2 = window_start(2.0f * A.ws - c1, A.H, A.cp, A.W);

    // qh(z) = z K + O of both rays at the corners [a][b] and two edge midpoints
    const float2 c2 = cs[st[0] + j + A.h];
    const float x0 = float(tx0), x1 = float(tx0 + kTile - 1);
    const float y0 = float(ty0), y1 = float(ty0 + kTileY - 1);
    float K[2][2][2], O[2][2][2], Kx[2], Ox[2], Ky[2], Oy[2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
            pair_row_lines(A, a ? x1 : x0, b ? y1 : y0, c, c2, j, K[a][b][0], O[a][b][0], K[a][b][1],
                           O[a][b][1]);
    pair_row_lines(A, 0.5f * (x0 + x1), y0, c, c2, j, Kx[0], Ox[0], Kx[1], Ox[1]);
    pair_row_lines(A, x0, 0.5f * (y0 + y1), c, c2, j, Ky[0], Oy[0], Ky[1], Oy[1]);
    float e = 0.0f;
#pragma unroll
    for (int v = 0; v < 2; v++) {
        // departures at z = 0 and z = 31 (linear in z between them)
        const float dx0 = fabsf(Ox[v] - 0.5f * (O[0][0][v] + O[1][0][v]));
        const float dx1 = fabsf(fmaf(31.0f, Kx[v], Ox[v]) -
                                0.5f * (fmaf(31.0f, K[0][0][v], O[0][0][v]) + fmaf(31.0f, K[1][0][v], O[1][0][v])));
        const float dy0 = fabsf(Oy[v] - 0.5f * (O[0][0][v] + O[0][1][v]));
        const float dy1 = fabsf(fmaf(31.0f, Ky[v], Oy[v]) -
                                0.5f * (fmaf(31.0f, K[0][0][v], O[0][0][v]) + fmaf(31.0f, K[0][1][v], O[0][1][v])));
        e = fmaxf(e, 2.0f * (fmaxf(dx0, dx1) + fmaxf(dy0, dy1)));
    }
    const float L = A.wt - 0.25f - e;
    if (A.pi) {
        // PI-original: fast = every column's ray j inside its PI window (|qh| <= wt,
        // here <= L: on the detector); winb = every column's ray j outside it on
        // one side (|qh| >= wt + 1/4 + e) and ray j + h on the detector (|qh| <= L)
        const float U = A.wt + 0.25f + e;
        float in_lo = -1e30f, in_hi = 1e30f, up_lo = -1e30f, up_hi = 1e30f, dn_lo = -1e30f,
              dn_hi = 1e30f, r2_lo = -1e30f, r2_hi = 1e30f;
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int b = 0; b < 2; b++) {
                clip_line(K[a][b][0], O[a][b][0], -L, L, in_lo, in_hi);
                clip_line(K[a][b][0], O[a][b][0], U, 1e30f, up_lo, up_hi);
                clip_line(K[a][b][0], O[a][b][0], -1e30f, -U, dn_lo, dn_hi);
                clip_line(K[a][b][1], O[a][b][1], -L, L, r2_lo, r2_hi);
            }
        const unsigned fast = L > 0.0f ? slice_bits(in_lo, in_hi) : 0u;
        const unsigned other = L > 0.0f ? (slice_bits(up_lo, up_hi) | slice_bits(dn_lo, dn_hi)) &
                                              slice_bits(r2_lo, r2_hi)
                                        : 0u;
        win[i] = make_uint2(unsigned(s1) | (unsigned(s2) << 16), fast);
        A.winb[i] = other & ~fast;
        return;
    }
    float zlo = -1e30f, zhi = 1e30f;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int v = 0; v < 2; v++) {
                // slices with -L <= z K + O <= L (K != 0: dz > 0 and k != 0)
                const float r = __builtin_amdgcn_rcpf(K[a][b][v]), o = O[a][b][v];
                const float za = (-L - o) * r, zb = (L - o) * r;
                zlo = fmaxf(zlo, fminf(za, zb));
                zhi = fminf(zhi, fmaxf(za, zb));
            }
    unsigned fast = 0;
    if (L > 0.0f) {
        // a slice within 1e-3 of an end of the interval is left out
        const int z0 = int(ceilf(fminf(fmaxf(zlo + 1e-3f, 0.0f), 32.0f)));
        const int z1 = int(floorf(fmaxf(fminf(zhi - 1e-3f, 31.0f), -1.0f)));
        for (int z = z0; z <= z1; z++) fast |= 1u << z;
    }
    win[i] = make_uint2(unsigned(s1) | (unsigned(s2) << 16), fast);
}

// Device tables of the pair-folded backprojector (see bp_fold_kernel).
struct BpTables {
    const float2 *cs;
    const unsigned *zmask;
    const uint2 *win;     // channel windows (unused by the gather path) and fast slices
};

template <typename T>
static int launch_bp_fold_gather(lct_ctx *ctx, FoldBpArgs &a, int n_blocks, const BpTables &tb)
{
    a.n_wg = a.n_tiles_x * a.n_tiles_y * n_blocks * a.n_parts;
    a.xcd_remap = (a.n_wg % 8 == 0) ? 1 : 0;
    a.lp = a.rp + 1;
    a.W = 2;  // no LDS window
    hipLaunchKernelGGL((bp_fold_kernel<T, 1, 0, true>), dim3(unsigned(a.n_wg)), dim3(kBpThreads), 0,
                       ctx->stream, a, tb.cs, tb.zmask, tb.win);
    LCT_CHECK_HIP(hipGetLastError());
    return LCT_OK;
}

template <typename T, bool PI>
static int launch_bp_fold(lct_ctx *ctx, FoldBpArgs &a, int n_blocks, const BpTables &tb)
{
    a.n_wg = a.n_tiles_x * a.n_tiles_y * n_blocks * a.n_parts;
    a.xcd_remap = (a.n_wg % 8 == 0) ? 1 : 0;
    constexpr int CE = Chunk<T>::kElems;
    a.lp = a.rp + 1;
    // chunk slots: view j's chunks, padded to a multiple of 64, then view j+h's
    const int n_vc = a.W * a.rp / CE;
    const int chunks = ((n_vc + 63) & ~63) + n_vc;
    const size_t shm = size_t(2) * a.W * a.lp * sizeof(float);
    if (shm > 160 * 1024)
        return set_error(LCT_ERR_UNSUPPORTED, "backprojector: channel window exceeds LDS");
    const int npre = (chunks + kBpThreads - 1) / kBpThreads;
    if (npre < 1 || npre > 8)
        return set_error(LCT_ERR_UNSUPPORTED, "backprojector: channel window too large");
    // LDS pitches of the 16-, 64- and 96-row detectors get immediate offsets
    using Kern = void (*)(FoldBpArgs, const float2 *, const unsigned *, const uint2 *);
    static const Kern table[4][8] = {
#define LCT_ROW(LP)                                                                              \
    {bp_fold_kernel<T, 1, LP, false, false, PI>, bp_fold_kernel<T, 2, LP, false, false, PI>,     \
     bp_fold_kernel<T, 3, LP, false, false, PI>, bp_fold_kernel<T, 4, LP, false, false, PI>,     \
     bp_fold_kernel<T, 5, LP, false, false, PI>, bp_fold_kernel<T, 6, LP, false, false, PI>,     \
     bp_fold_kernel<T, 7, LP, false, false, PI>, bp_fold_kernel<T, 8, LP, false, false, PI>}
        LCT_ROW(0), LCT_ROW(25), LCT_ROW(73), LCT_ROW(105)
#undef LCT_ROW
    };
    const int row = a.rp == 24 ? 1 : a.rp == 72 ? 2 : a.rp == 104 ? 3 : 0;
    const Kern k = table[row][npre - 1];
    if (shm > 64 * 1024)
        LCT_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(k),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, int(shm)));
    hipLaunchKernelGGL(k, dim3(unsigned(a.n_wg)), dim3(kBpThreads), shm, ctx->stream, a, tb.cs, tb.zmask,
                       tb.win);
    LCT_CHECK_HIP(hipGetLastError());
    return LCT_OK;
}

// f32 storage: windows copied by LDS-DMA, Q with the odd pitch rp = lp.
template <bool PI>
static int launch_bp_fold_dma(lct_ctx *ctx, FoldBpArgs &a, int n_blocks, const BpTables &tb)
{
    a.n_wg = a.n_tiles_x * a.n_tiles_y * n_blocks * a.n_parts;
    a.xcd_remap = (a.n_wg % 8 == 0) ? 1 : 0;
    a.lp = a.rp;
    // a window of W rp floats starting up to 3 floats into its first unit
    a.units = (a.W * a.rp + 3 + 3) / 4;
    const size_t shm = size_t(2) * ((a.units + 63) & ~63) * 16;
    if (shm > 160 * 1024)
        return set_error(LCT_ERR_UNSUPPORTED, "backprojector: channel window exceeds LDS");
    using Kern = void (*)(FoldBpArgs, const float2 *, const unsigned *, const uint2 *);
    const Kern k = a.rp == 25    ? bp_fold_kernel<float, 1, 25, false, true, PI>
                   : a.rp == 73  ? bp_fold_kernel<float, 1, 73, false, true, PI>
                   : a.rp == 105 ? bp_fold_kernel<float, 1, 105, false, true, PI>
                                 : bp_fold_kernel<float, 1, 0, false, true, PI>;
    if (shm > 64 * 1024)
        LCT_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(k),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, int(shm)));
    hipLaunchKernelGGL(k, dim3(unsigned(a.n_wg)), dim3(kBpThreads), shm, ctx->stream, a, tb.cs, tb.zmask,
                       tb.win);
    LCT_CHECK_HIP(hipGetLastError());
    return LCT_OK;
}

// Bound on |p_hat| / dso = sin(theta) over the volume's voxel columns (one
// voxel of margin).
static double column_u_max(const lct_ct_geom &g, const lct_img_geom &ig)
{
    const double rx = (std::max(std::fabs(0.0 - ig.wx), std::fabs(ig.nx - 1.0 - ig.wx)) + 1.0) * ig.dx;
    const double ry = (std::max(std::fabs(0.0 - ig.wy), std::fabs(ig.ny - 1.0 - ig.wy)) + 1.0) * ig.dy;
    return std::sqrt(rx * rx + ry * ry) / g.dso;
}

// Half-width (in channels) of the channel footprint of a tile, with
// the largest in-plane radius of the volume bounding the asin slope.
static int tile_half_width(const lct_ct_geom &g, const lct_img_geom &ig)
{
    const double u = std::min(column_u_max(g, ig), 0.999);
    const double slope = 1.0 / std::sqrt(1.0 - u * u);
    // |dx cos b + dy sin b| <= sqrt(dx^2 + dy^2) for the column offsets of a tile.
    // Taps of a column at channel p use channels floor(p) and floor(p) + 1, and
    // |p - c| <= E for the tile centre c, so H >= ceil(E) covers them; one
    // more channel absorbs FP32 rounding of the two asin evaluations.
    const double ex = (kTile - 1) * double(ig.dx), ey = (kTileY - 1) * double(ig.dy);
    const double half_extent = 0.5 * std::sqrt(ex * ex + ey * ey);
    return int(std::ceil(half_extent * slope / g.ds)) + 1;
}

// Parallel views [g0, g0 + nviews) that blocks [b0, b1) backproject, and
// each block's blockProjStart (bps[b - b0]).
struct Resident {
    long g0 = 0, nviews = 0;
    std::vector<int> bps;
};

static int plan_resident(const lct_plan &plan, const float *h_couch, long file_views, int b0,
                         int b1, Resident &r)
{
    const lct_ct_geom &g = plan.ctg;
    r.bps.resize(size_t(b1 - b0));
    long lo = 1L << 40, hi = -(1L << 40);
    for (int b = b0; b < b1; b++) {
        int s = 0;
        LCT_CHECK(lct_block_info(&plan, h_couch, b, &s, nullptr, nullptr, nullptr, nullptr));
        r.bps[size_t(b - b0)] = s;
        if (s < 0 || long(s) + plan.data_offset < 0 ||
            long(s) + plan.data_offset + g.n_angle_add > file_views ||
            s + g.add_projections + g.n_angle > plan.n_readings)
            return set_error(LCT_ERR_RANGE, "block " + std::to_string(b) +
                                                 " projection window lies outside the data");
        lo = std::min(lo, long(s) + g.add_projections);
        hi = std::max(hi, long(s) + g.add_projections + g.n_angle);
    }
    r.g0 = lo;
    r.nviews = hi - lo;
    return LCT_OK;
}

int reconstruct_resident(lct_ctx *ctx, const lct_plan &plan, const lct_recon_opts &o,
                         const float *h_angles, const float *h_couch, const float *d_fan,
                         long file_views, int b0, int b1, float *d_vol)
{
    const lct_ct_geom &g = plan.ctg;
    const lct_img_geom &ig = plan.ig;
    const int nb = b1 - b0;
    Resident R;
    LCT_CHECK(plan_resident(plan, h_couch, file_views, b0, b1, R));
    const size_t view = size_t(g.nt) * g.ns;
    const int h = g.n_proj_turn / 2;

    // Can the pair-folded backprojector be used?
    int amin[32], amax[32], amid[32];
    bp_windows(g, ig, amin, amax, amid);
    // the PI-original backprojector folds the same pairs (one view of each
    // counts, Backprojector.cu:200-245) but has no cache-gather form
    const bool pi = o.bp_kind == LCT_BP_PI_ORI;
    bool fold = (g.n_proj_turn % 2 == 0) && (g.nt % 8 == 0) &&
                (o.bp_kind == LCT_BP_HELICAL || (pi && o.interp != LCT_INTERP_GATHER));
    for (int z = 0; z < 32 && fold; z++)
        fold = (amid[z] - h >= 0) && (amid[z] + h < g.n_angle) && (z == 0 || amid[z] >= amid[z - 1]);
    for (long v = R.g0; v + h < R.g0 + R.nviews && fold; v++) {
        double d = double(h_angles[v + h]) - double(h_angles[v]) - LCT_PI_D;
        d = std::remainder(d, 2.0 * LCT_PI_D);
        fold = std::fabs(d) < 2e-5;
    }
    if (!fold && o.storage != LCT_F32)
        return set_error(LCT_ERR_UNSUPPORTED,
                         "reduced-precision storage needs the pair-folded backprojector (even "
                         "npt, unclipped windows, half-turn-symmetric angles, nt % 8 == 0)");

    std::vector<float> taps(2 * size_t(g.ns));
    host_taps(o.window, g.ns, g.ds, taps.data());
    void *d_taps = nullptr, *d_p = nullptr, *d_q = nullptr, *d_tab = nullptr;
    LCT_CHECK(scratch(ctx, kScratchTaps, taps.size() * 4, &d_taps));
    LCT_CHECK(upload(ctx, d_taps, taps.data(), taps.size() * 4));
    LCT_CHECK(scratch(ctx, kScratchA, view * R.nviews * 4, &d_p));
    float *P = static_cast<float *>(d_p);

    lct_timings &t = ctx->timings;
    hipEvent_t e0 = ctx->ev[6], e1 = ctx->ev[7];
    auto tic = [&]() -> int {
        if (ctx->profiling) LCT_CHECK_HIP(hipEventRecord(e0, ctx->stream));
        return LCT_OK;
    };
    auto toc = [&](float *acc) -> int {
        if (!ctx->profiling) return LCT_OK;
        LCT_CHECK_HIP(hipEventRecord(e1, ctx->stream));
        LCT_CHECK_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        LCT_CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
        *acc += ms;
        return LCT_OK;
    };

    if (!fold) {
        // Rebin every needed view once (fused pre-weighting), filter in place
        // and backproject each block from its offset with the per-voxel kernel.
        LCT_CHECK(tic());
        LCT_CHECK(launch_rebin(ctx, o.rebin_kind, o.weight_kind, &g, d_fan, 0, file_views, P,
                               R.nviews, R.g0 + plan.data_offset, 0));
        LCT_CHECK(toc(&t.rebin_ms));
        LCT_CHECK(tic());
        LCT_CHECK(launch_filter_direct(ctx, g.ns, static_cast<const float *>(d_taps), P,
                                       long(R.nviews) * g.nt));
        LCT_CHECK(toc(&t.filter_ms));
        const size_t slab = size_t(ig.nx) * ig.ny * ig.nz;
        std::vector<float> betas(g.n_angle + h + 1);
        for (int k = 0; k < nb; k++) {
            const long off = long(R.bps[k]) + g.add_projections - R.g0;
            for (size_t i = 0; i < betas.size(); i++) {
                const long src = long(R.bps[k]) + g.add_projections + long(i);
                betas[i] = src < plan.n_readings ? h_angles[src] : 0.0f;
            }
            LCT_CHECK(tic());
            LCT_CHECK(backproject_block_public(ctx, o.bp_kind, g, ig, P + size_t(off) * view,
                                               betas.data(), d_vol + size_t(k) * slab));
            LCT_CHECK(toc(&t.backproject_ms));
            t.backproject_kernel_ms = t.backproject_ms;
            t.n_backproject_launches++;
        }
        ctx->last = lct_ctx::LastBlock{1, LCT_F32, g.n_angle, g.nt, g.ns, 0, 0,
                                       long(R.bps[nb - 1]) + g.add_projections - R.g0, P};
        return LCT_OK;
    }

    // Q: every view rebinned (fused pre-weighting) and filtered once, in the
    // backprojector's layout and storage precision.
    // f32 windows are copied into LDS as they lie in Q, so Q carries the
    // backprojector's odd LDS pitch (one more zero row slot); 16-bit windows
    // are converted on the way and keep rp = nt + 8.
    const bool dma = o.storage == LCT_F32 && o.interp != LCT_INTERP_GATHER;
    const int cp = g.ns + 3, rp = g.nt + (dma ? 9 : 8);
    const size_t esz = o.storage == LCT_F32 ? 4 : 2;
    const size_t q_view = size_t(cp) * rp * esz;  // bytes of one view in Q
    // the copies of the last view's window read up to 24 bytes past it
    const size_t q_bytes = q_view * R.nviews + 256;
    LCT_CHECK(scratch(ctx, kScratchB, q_bytes, &d_q));
    // The filters write every channel and row slot of the views they fill and
    // never the guard slots, so Q's guards stay zero from one call to the
    // next: Q is cleared only when its buffer, its layout or its extent
    // changed since it was last cleared.
    lct_ctx::QZero &qz = ctx->qzero;
    if (qz.ptr != d_q || qz.cp != cp || qz.rp != rp || qz.esz != esz || qz.bytes < q_bytes) {
        LCT_CHECK_HIP(hipMemsetAsync(d_q, 0, q_bytes, ctx->stream));
        qz = lct_ctx::QZero{d_q, cp, rp, esz, q_bytes};
    }
    // rebin and filter of the resident views [v0, v1) on the context stream
    auto prep = [&](long v0, long v1) -> int {
        LCT_CHECK(tic());
        LCT_CHECK(launch_rebin(ctx, o.rebin_kind, o.weight_kind, &g, d_fan, 0, file_views,
                               P + size_t(v0) * view, v1 - v0, R.g0 + v0 + plan.data_offset, 0));
        LCT_CHECK(toc(&t.rebin_ms));
        const float *in = P + size_t(v0) * view;
        void *q = static_cast<char *>(d_q) + size_t(v0) * q_view;
        const long rows = (v1 - v0) * g.nt;
        const float *dt = static_cast<const float *>(d_taps);
        LCT_CHECK(tic());
        if (o.filter_algo != LCT_FILTER_DIRECT_F32 && g.ns % 4 == 0)
            LCT_CHECK(launch_filter_toeplitz(ctx, o.filter_algo, o.storage, g.ns, g.nt, taps.data(),
                                             in, q, rows, cp, rp));
        else if (o.storage == LCT_F32)
            LCT_CHECK(launch_filter_to_q<float>(ctx, g.ns, g.nt, dt, in, static_cast<float *>(q),
                                                rows, cp, rp));
        else if (o.storage == LCT_F16)
            LCT_CHECK(launch_filter_to_q<__half>(ctx, g.ns, g.nt, dt, in, static_cast<__half *>(q),
                                                 rows, cp, rp));
        else
            LCT_CHECK(launch_filter_to_q<__hip_bfloat16>(ctx, g.ns, g.nt, dt, in,
                                                         static_cast<__hip_bfloat16 *>(q), rows,
                                                         cp, rp));
        LCT_CHECK(toc(&t.filter_ms));
        return LCT_OK;
    };

    // Per-view cos/sin (double) of the resident views, the block table and
    // the slices whose window holds pair j (a_middle is the same in every block).
    const int n_cs = int(R.nviews + h + 1);
    std::vector<float> cs_host(2 * size_t(n_cs));
    {
        std::vector<float> b(n_cs);
        for (int i = 0; i < n_cs; i++) {
            const long src = R.g0 + i;
            b[i] = src < plan.n_readings ? h_angles[src] : 0.0f;
        }
        std::vector<float> sep;
        angle_tables(b.data(), n_cs, sep);  // [cos...][sin...] -> (cos, sin) pairs
        for (int i = 0; i < n_cs; i++) {
            cs_host[2 * size_t(i)] = sep[size_t(i)];
            cs_host[2 * size_t(i) + 1] = sep[size_t(n_cs) + i];
        }
    }
    std::vector<int> tab(34 * size_t(nb));
    for (int k = 0; k < nb; k++) {
        tab[34 * k] = int(long(R.bps[k]) + g.add_projections - R.g0);
        tab[34 * k + 1] = k * LCT_BLOCK_SLICES;
        for (int z = 0; z < 32; z++) tab[34 * k + 2 + z] = amid[z];
    }
    // helical: pairs a_middle - h <= j < a_middle of each slice (Backprojector.cu:38-47);
    // PI-original: a_middle - h < j <= a_middle, the views a_middle - h + 1 ..
    // a_middle + h among which the PI window of every voxel lies (its walk,
    // Backprojector.cu:200-245, starts at a_min > a_middle - h)
    std::vector<unsigned> zmask(size_t(amid[31] - amid[0] + h));
    for (size_t jj = 0; jj < zmask.size(); jj++) {
        const int j = amid[0] - h + (pi ? 1 : 0) + int(jj);
        unsigned m = 0;
        for (int z = 0; z < 32; z++)
            if (pi ? (amid[z] - h < j && j <= amid[z]) : (amid[z] - h <= j && j < amid[z])) m |= 1u << z;
        zmask[jj] = m;
    }
    const size_t tab_bytes = cs_host.size() * 4 + tab.size() * 4 + zmask.size() * 4;
    LCT_CHECK(scratch(ctx, kScratchTables, tab_bytes, &d_tab));
    float *d_cs = static_cast<float *>(d_tab);  // 8-byte aligned: cs_host has 2 n_cs floats
    int *d_st = reinterpret_cast<int *>(d_cs + cs_host.size());
    unsigned *d_zm = reinterpret_cast<unsigned *>(d_st + tab.size());
    const int n_pairs = int(zmask.size());
    const int n_tiles = ((ig.nx + kTile - 1) / kTile) * ((ig.ny + kTileY - 1) / kTileY);
    const long n_win = long(nb) * n_tiles * n_pairs;
    void *d_win = nullptr;
    // window table (8 B per entry), then the PI-original view-(j + h) slices (4 B)
    LCT_CHECK(scratch(ctx, kScratchMisc0, size_t(n_win) * (pi ? 12 : 8), &d_win));
    const BpTables tb{reinterpret_cast<const float2 *>(d_cs), d_zm, static_cast<const uint2 *>(d_win)};

    FoldBpArgs a{};
    a.q = d_q;
    a.view_stride = size_t(cp) * rp;
    a.rp = rp;
    a.cp = cp;
    a.H = tile_half_width(g, ig);
    a.asin_poly = column_u_max(g, ig) <= 0.7 ? 1 : 0;
    a.W = std::min(2 * a.H + 2, cp);
    a.slab_tab = d_st;
    a.n_pairs = n_pairs;
    a.vol = d_vol;
    a.n_tiles_x = (ig.nx + kTile - 1) / kTile;
    a.n_tiles_y = (ig.ny + kTileY - 1) / kTileY;
    a.nx = ig.nx;
    a.ny = ig.ny;
    a.h = h;
    a.wx = ig.wx;
    a.wy = ig.wy;
    a.wz = ig.wz;
    a.dx = ig.dx;
    a.dy = ig.dy;
    a.dz = ig.dz;
    a.inv_dso = float(1.0 / double(g.dso));
    a.dso_ds = float(double(g.dso) / double(g.ds));
    a.dso2 = float(double(g.dso) * double(g.dso));
    a.ws = g.ws;
    a.wt = g.wt;
    a.inv_orbit = float(1.0 / double(g.orbit));
    a.z0 = g.source_z0;
    a.dzv = float(double(g.pitch) / double(g.n_proj_turn));
    // rows per mm along z: dso z_dir / dt for the helical data, z_dir 2 (nt - 1) / P
    // for the PI-original rebin (rows P / 2 (nt - 1) apart, Backprojector.cu:216)
    a.kc = pi ? float(double(g.z_dir) * 2.0 * double(g.nt - 1) / double(g.pitch))
              : float(double(g.dso) * double(g.z_dir) / double(g.dt));
    a.inv_ds = float(1.0 / double(g.ds));
    a.pi = pi ? 1 : 0;
    a.winb = pi ? reinterpret_cast<unsigned *>(static_cast<char *>(d_win) + size_t(n_win) * 8) : nullptr;
    a.ns_f = float(g.ns);
    a.nt_f = float(g.nt);
    // helical: 2 pi / (a_max - a_min) over 2h views, each pair counted once here
    // (twice there); PI-original: pi / (npt / 2) (Backprojector.cu:247)
    a.scale = pi ? float(LCT_PI_D / double(h)) : float(2.0 * 2.0 * LCT_PI_D / double(2 * h));
    // Each (block, tile) can be backprojected by kParts workgroups, one per
    // range of the block's pairs (ranges of about equal work: a pair costs
    // about 4 slices' worth of set-up and staging plus its active slices).
    // Shorter workgroups shorten the launch's tail, where slots empty one by
    // one behind the last workgroups to start (on the clinical volume the
    // whole GPU idled ~3 ms there out of 29). The ranges are the same for
    // every block, so a block's sums do not depend on which blocks share its
    // launch, and a slab equals the same blocks of the whole volume bit for
    // bit. Measured (whole volumes back to back, same box): 16-bit storage
    // 32.1 -> 31.2 ms with 4 ranges (2: 32.0, 3: 31.5, 6: 31.2, 8: 31.3);
    // the large volume, 42 rounds of workgroups, 410 -> 409 ms; f32 storage
    // (the LDS-DMA kernel, 4 workgroups per CU) 39.0 -> 41.4 ms, so it keeps
    // one range.
    const int kParts = 4;
    const bool want_parts = o.storage != LCT_F32 && o.interp != LCT_INTERP_GATHER;
    a.n_parts = 1;
    a.part = nullptr;
    a.part_stride = size_t(nb) * LCT_BLOCK_SLICES * ig.ny * ig.nx;
    for (int &x : a.split) x = 0;
    if (want_parts && n_pairs >= 2 * kParts) {
        long total = 0, run = 0;
        for (unsigned m : zmask) total += m ? __builtin_popcount(m) + 4 : 0;
        int p = 1;
        for (int jj = 0; jj < n_pairs && p < kParts; jj++) {
            run += zmask[jj] ? __builtin_popcount(zmask[jj]) + 4 : 0;
            if (run * kParts >= total * p) a.split[p++] = jj + 1;
        }
        bool ok = p == kParts;
        for (int q = 1; q < kParts && ok; q++) ok = a.split[q] > a.split[q - 1] && a.split[q] < n_pairs;
        if (ok) {
            void *d_part = nullptr;
            LCT_CHECK(scratch(ctx, kScratchPart, (kParts - 1) * a.part_stride * 4, &d_part));
            a.part = static_cast<float *>(d_part);
            a.n_parts = kParts;
        } else {
            for (int &x : a.split) x = 0;
        }
    }
    // backprojection of blocks [k0, k1) in one launch on the context stream
    auto bp = [&](int k0, int k1) -> int {
        FoldBpArgs b = a;
        b.slab_tab = d_st + 34 * k0;
        BpTables tk = tb;  // the launch numbers its blocks from 0
        tk.win += size_t(k0) * n_tiles * n_pairs;
        if (b.winb) b.winb += size_t(k0) * n_tiles * n_pairs;
        LCT_CHECK(tic());
        if (o.interp == LCT_INTERP_GATHER) {
            if (o.storage == LCT_F32)
                LCT_CHECK(launch_bp_fold_gather<float>(ctx, b, k1 - k0, tk));
            else if (o.storage == LCT_F16)
                LCT_CHECK(launch_bp_fold_gather<__half>(ctx, b, k1 - k0, tk));
            else
                LCT_CHECK(launch_bp_fold_gather<__hip_bfloat16>(ctx, b, k1 - k0, tk));
        } else if (o.storage == LCT_F32)
            LCT_CHECK(pi ? launch_bp_fold_dma<true>(ctx, b, k1 - k0, tk)
                         : launch_bp_fold_dma<false>(ctx, b, k1 - k0, tk));
        else if (o.storage == LCT_F16)
            LCT_CHECK((pi ? launch_bp_fold<__half, true>(ctx, b, k1 - k0, tk)
                          : launch_bp_fold<__half, false>(ctx, b, k1 - k0, tk)));
        else
            LCT_CHECK((pi ? launch_bp_fold<__hip_bfloat16, true>(ctx, b, k1 - k0, tk)
                          : launch_bp_fold<__hip_bfloat16, false>(ctx, b, k1 - k0, tk)));
        // the backprojector alone, for its launch time (the sum pass follows)
        if (ctx->profiling) LCT_CHECK_HIP(hipEventRecord(ctx->ev[3], ctx->stream));
        if (a.n_parts > 1) {
            const size_t slab = size_t(LCT_BLOCK_SLICES) * ig.ny * ig.nx;
            const size_t n = size_t(k1 - k0) * slab;
            hipLaunchKernelGGL(bp_sum_parts_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0,
                               ctx->stream, d_vol + size_t(k0) * slab, a.part + size_t(k0) * slab,
                               a.part_stride, a.n_parts - 1, n, a.scale);
            LCT_CHECK_HIP(hipGetLastError());
        }
        LCT_CHECK(toc(&t.backproject_ms));
        if (ctx->profiling) {
            float ms = 0.0f;
            LCT_CHECK_HIP(hipEventElapsedTime(&ms, e0, ctx->ev[3]));
            t.backproject_kernel_ms += ms;
        }
        t.n_backproject_launches++;
        return LCT_OK;
    };
    auto upload_tables = [&]() -> int {
        LCT_CHECK(upload(ctx, d_cs, cs_host.data(), cs_host.size() * 4));
        LCT_CHECK(upload(ctx, d_st, tab.data(), tab.size() * 4));
        LCT_CHECK(upload(ctx, d_zm, zmask.data(), zmask.size() * 4));
        if (n_win > 0x7fffffffL * 256L)
            return set_error(LCT_ERR_UNSUPPORTED, "backprojector: window table too large");
        hipLaunchKernelGGL(bp_windows_kernel, dim3(unsigned((n_win + 255) / 256)), dim3(256), 0,
                           ctx->stream, a, nb, tb.cs, static_cast<uint2 *>(d_win));
        LCT_CHECK_HIP(hipGetLastError());
        return LCT_OK;
    };
    ctx->last = lct_ctx::LastBlock{2, o.storage, g.n_angle, g.nt, g.ns, cp, rp,
                                   long(R.bps[nb - 1]) + g.add_projections - R.g0, d_q};

    // Overlapping the rebin and filter with the backprojection pays only when
    // the filter is the direct FP32 convolution (7.8 ms per clinical volume:
    // 40.0 -> 38.4 ms per volume). Against the 2 ms Toeplitz filter the split
    // backprojector launches cost more than they hide (32.1 -> 32.6-33.7 ms
    // with f16 storage, 414 -> 420 ms on the large volume).
    const bool toeplitz = o.filter_algo != LCT_FILTER_DIRECT_F32 && g.ns % 4 == 0;
    if (ctx->profiling || nb < 2 || toeplitz) {
        // one stage after the other on the context stream (each one timed)
        LCT_CHECK(prep(0, R.nviews));
        LCT_CHECK(upload_tables());
        return bp(0, nb);
    }
    // Overlapped: the blocks go in up to 4 groups. The views a group needs
    // are rebinned and filtered on aux[0] in order, and each group is
    // backprojected as soon as its views are ready, alternately on the
    // context stream and aux[1], so one group's launch fills the GPU while
    // another's last workgroups drain and the next views are prepared
    // beside them. Every kernel and its inputs are those of the sequential
    // order, so the volume is the same bit for bit.
    LCT_CHECK(aux_init(ctx));
    LCT_CHECK(upload_tables());
    const hipStream_t main = ctx->stream, side = ctx->aux[0], other = ctx->aux[1];
    struct Restore {  // ctx->stream is switched per launch below
        lct_ctx *c;
        hipStream_t s;
        ~Restore() { c->stream = s; }
    } restore{ctx, main};
    hipEvent_t *ev = ctx->sync_ev;
    LCT_CHECK_HIP(hipEventRecord(ev[0], main));  // tables, windows and the caller's inputs
    LCT_CHECK_HIP(hipStreamWaitEvent(side, ev[0], 0));
    LCT_CHECK_HIP(hipStreamWaitEvent(other, ev[0], 0));
    const int per = (nb + 3) / 4;
    long ready = 0;  // views [0, ready) are filtered in Q
    int gi = 0;
    for (int k0 = 0; k0 < nb; k0 += per, gi++) {
        const int k1 = std::min(nb, k0 + per);
        // block k reads Q views qoff + j and qoff + j + h, j < a_middle[31]
        long need = R.nviews;
        if (k1 < nb) {
            need = 0;
            for (int k = k0; k < k1; k++) need = std::max(need, long(tab[34 * k]) + amid[31] + h);
            need = std::min(need, long(R.nviews));
        }
        if (need > ready) {
            ctx->stream = side;
            LCT_CHECK(prep(ready, need));
            ready = need;
        }
        LCT_CHECK_HIP(hipEventRecord(ev[1 + gi], side));
        const hipStream_t s = (gi % 2 == 0) ? main : other;
        LCT_CHECK_HIP(hipStreamWaitEvent(s, ev[1 + gi], 0));
        ctx->stream = s;
        LCT_CHECK(bp(k0, k1));
    }
    // the context stream follows everything
    LCT_CHECK_HIP(hipEventRecord(ev[1 + gi], other));
    LCT_CHECK_HIP(hipStreamWaitEvent(main, ev[1 + gi], 0));
    return LCT_OK;
}

// Q[view0 + v][c + 1][r + kRowPad] -> rows[v][r][c] (f32).
template <typename T>
__global__ __launch_bounds__(256) void q_to_rows_kernel(const T *__restrict__ q, long view0,
                                                        int cp, int rp, int ns, int nt,
                                                        float *__restrict__ rows)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    const long v = blockIdx.z;
    if (c >= ns) return;
    rows[(size_t(v) * nt + r) * ns + c] =
        Cvt<T>::from(q[(size_t(view0 + v) * cp + c + 1) * rp + r + kRowPad]);
}

}  // namespace lct

using namespace lct;

extern "C" int lct_last_block_filtered(lct_ctx *ctx, float *d_rows)
{
    if (!ctx || !d_rows) return set_error(LCT_ERR_INVALID, "lct_last_block_filtered: NULL argument");
    const lct_ctx::LastBlock &L = ctx->last;
    if (L.layout == 0)
        return set_error(LCT_ERR_INVALID, "lct_last_block_filtered: no reconstruction on this context");
    LCT_CHECK_HIP(hipSetDevice(ctx->device));
    const size_t view = size_t(L.nt) * L.ns;
    if (L.layout == 1) {
        LCT_CHECK_HIP(hipMemcpyAsync(d_rows, static_cast<const float *>(L.ptr) + L.view0 * view,
                                     view * L.n_angle * sizeof(float), hipMemcpyDeviceToDevice,
                                     ctx->stream));
        return LCT_OK;
    }
    const dim3 grid(unsigned((L.ns + 255) / 256), unsigned(L.nt), unsigned(L.n_angle));
    if (L.storage == LCT_F32)
        hipLaunchKernelGGL(q_to_rows_kernel<float>, grid, dim3(256), 0, ctx->stream,
                           static_cast<const float *>(L.ptr), L.view0, L.cp, L.rp, L.ns, L.nt, d_rows);
    else if (L.storage == LCT_F16)
        hipLaunchKernelGGL(q_to_rows_kernel<__half>, grid, dim3(256), 0, ctx->stream,
                           static_cast<const __half *>(L.ptr), L.view0, L.cp, L.rp, L.ns, L.nt, d_rows);
    else
        hipLaunchKernelGGL(q_to_rows_kernel<__hip_bfloat16>, grid, dim3(256), 0, ctx->stream,
                           static_cast<const __hip_bfloat16 *>(L.ptr), L.view0, L.cp, L.rp, L.ns,
                           L.nt, d_rows);
    LCT_CHECK_HIP(hipGetLastError());
    return LCT_OK;
}

extern "C" int lct_filter_to_bp_layout(lct_ctx *ctx, int algo, int storage, int ns, int nt,
                                       const float *h_taps, const float *d_rows, long views,
                                       void *d_q)
{
    if (!ctx || !h_taps || !d_rows || !d_q || ns <= 0 || nt <= 0 || nt % 8 != 0 || views < 0 ||
        storage < 0 || storage > 2 || algo < 0 || algo > 2)
        return set_error(LCT_ERR_INVALID, "lct_filter_to_bp_layout: bad argument");
    if (algo != LCT_FILTER_DIRECT_F32 && ns % 4 != 0)
        return set_error(LCT_ERR_UNSUPPORTED, "lct_filter_to_bp_layout: Toeplitz needs ns % 4 == 0");
    LCT_CHECK_HIP(hipSetDevice(ctx->device));
    const int cp = ns + 3, rp = nt + 8;
    const long rows = views * nt;
    if (algo != LCT_FILTER_DIRECT_F32)
        return launch_filter_toeplitz(ctx, algo, storage, ns, nt, h_taps, d_rows, d_q, rows, cp, rp);
    void *d_taps = nullptr;
    LCT_CHECK(scratch(ctx, kScratchTaps, size_t(2) * ns * sizeof(float), &d_taps));
    LCT_CHECK(upload(ctx, d_taps, h_taps, size_t(2) * ns * sizeof(float)));
    const float *t = static_cast<const float *>(d_taps);
    if (storage == LCT_F32)
        return launch_filter_to_q<float>(ctx, ns, nt, t, d_rows, static_cast<float *>(d_q), rows, cp, rp);
    if (storage == LCT_F16)
        return launch_filter_to_q<__half>(ctx, ns, nt, t, d_rows, static_cast<__half *>(d_q), rows, cp, rp);
    return launch_filter_to_q<__hip_bfloat16>(ctx, ns, nt, t, d_rows,
                                              static_cast<__hip_bfloat16 *>(d_q), rows, cp, rp);
}
