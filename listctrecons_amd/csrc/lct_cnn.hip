// lct_cnn.hip -- the learned artifact correction: a 3D residual CNN of six
// 3x3x3 convolutions (1 -> 32 -> 32 -> 32 -> 32 -> 32 -> 1, ReLU after the
// first five), y = x + net(x), trained with Adam on FP32 master weights.
// Not part of the reference (it has no learned component).
//
// Layout: hidden activations are channels-last bf16 [b][z][y][x][32]; the
// single-channel input / output / target are f32 [b][z][y][x].  Convolutions
// are cross-correlations with zero padding ('same'):
//   out[v][co] = sum_tap sum_ci W[tap][co][ci] in[v + off(tap)][ci],
//   tap = (dz+1) 9 + (dy+1) 3 + (dx+1).
// The 32 -> 32 layers run as implicit GEMMs on v_mfma_f32_16x16x32_bf16
// (M = voxels, N = output channels, K = one tap's 32 input channels per
// step, 27 steps), bf16 operands, FP32 accumulation.  Their data gradient is
// the same kernel with the taps mirrored and W transposed per tap, and the
// 32 -> 1 output layer is the same kernel with one N fragment whose column 0
// holds the (bf16) output weights.  The 1 -> 32 layer is an MFMA GEMM over
// the 27 taps.  The 32 -> 32 weight gradients are implicit GEMMs over voxels
// on the same MFMA; the weight gradients of the 1-channel layers and the bias
// gradients run on FP32 VALU.  Every reduction over voxels goes through
// per-chunk partial sums added in a fixed order (no float atomics), so the
// loss and the gradients are bit-reproducible.
#include <hip/hip_bf16.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>

#include "lct_internal.hpp"

namespace lct {

constexpr int kC = 32;      // hidden channels
constexpr int kTaps = 27;

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef unsigned short bf16_t;  // bf16 bits

// parameter offsets (floats) in the packed parameter vector
constexpr long kW1 = 0, kB1 = kW1 + kTaps * kC;
constexpr long kMid = kB1 + kC;  // 4 layers of W[27][32][32] + b[32]
constexpr long kMidStride = long(kTaps) * kC * kC + kC;
constexpr long kW6 = kMid + 4 * kMidStride, kB6 = kW6 + kTaps * kC;
constexpr long kParams = kB6 + 1;

__device__ __forceinline__ bf16_t f2bf(float x)
{
    unsigned u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return bf16_t(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16_t b) { return __uint_as_float(unsigned(b) << 16); }
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));
// (a, b) -> two bf16 in one word, rounded to nearest even by the hardware
// conversion (v_cvt_pk_bf16_f32): the same bits as f2bf for finite values
__device__ __forceinline__ unsigned pack_bf16x2(float a, float b)
{
    return __builtin_bit_cast(unsigned, __builtin_convertvector(f32x2_t{a, b}, bf16x2_t));
}

struct Shape {
    int B, D, H, W;
};

__device__ __forceinline__ long vox(const Shape &s, int b, int z, int y, int x)
{
    return ((long(b) * s.D + z) * s.H + y) * s.W + x;
}

__device__ __forceinline__ bool inside(const Shape &s, int z, int y, int x)
{
    return z >= 0 && z < s.D && y >= 0 && y < s.H && x >= 0 && x < s.W;
}

// -------------------------------------------- 32 -> 32 layers on MFMA
// Workgroup tile: 4 (z) x 4 (y) x 16 (x) output voxels of one batch item;
// wave w owns plane z0 + w (4 M-fragments of 16 voxels) and all 32 output
// channels (2 N-fragments).  The (4+2) x (4+2) x (16+2) halo of input
// voxels (32 bf16 channels each) is staged in LDS.
// mode 0: out = relu(conv + bias); mode 1: out = conv * (mask > 0);
// mode 2: yout = xres + bias[0] + conv, one output channel (w is [27][16][32]
// with only column 0 non-zero; one N fragment).
constexpr int kTZ = 4, kTY = 4, kTX = 16;  // 2 x 16 and 8 x 16 y-x tiles ran slower
constexpr int kHZ = kTZ + 2, kHY = kTY + 2, kHX = kTX + 2;
constexpr int kHalo = kHZ * kHY * kHX;  // 648 voxels, 41472 B of LDS
// Largest plane (H W voxels) the conv kernels take: the offsets of a tile's
// voxels and halo chunks from the tile's first voxel then fit 32 bits (bytes
// of a 32-channel bf16 halo: 6 planes x 2^23 x 64 B < 2^32).
constexpr long kMaxPlane = 1L << 23;

// The conv kernels run on a 3D grid: x = tile column, y = tile row, z = tile
// plane + tiles_z x batch item, so a workgroup finds its tile without the
// 64-bit divisions of a flat index.
struct TileOrg {
    int b, z0, y0, x0;
};
__device__ __forceinline__ TileOrg tile_origin(const Shape &s)
{
    const int tiles_z = (s.D + kTZ - 1) / kTZ;
    const int b = int(blockIdx.z) / tiles_z;
    return TileOrg{b, (int(blockIdx.z) - b * tiles_z) * kTZ, int(blockIdx.y) * kTY, int(blockIdx.x) * kTX};
}
// the (kTZ+2) x (kTY+2) x (kTX+2) halo lies inside the volume (every tile but
// those at its faces): no clamps or padding, addresses from the tile base
__device__ __forceinline__ bool halo_inside(const Shape &s, const TileOrg &o)
{
    return o.z0 >= 1 && o.z0 + kTZ < s.D && o.y0 >= 1 && o.y0 + kTY < s.H && o.x0 >= 1 &&
           o.x0 + kTX < s.W;
}
// the tile's output voxels all lie inside the volume
__device__ __forceinline__ bool tile_inside(const Shape &s, const TileOrg &o)
{
    return o.z0 + kTZ <= s.D && o.y0 + kTY <= s.H && o.x0 + kTX <= s.W;
}
// offset of tile voxel (zl, yl, xl) from the tile's first voxel
__device__ __forceinline__ unsigned tile_off(const Shape &s, int zl, int yl, int xl)
{
    return (unsigned(zl) * unsigned(s.H) + unsigned(yl)) * unsigned(s.W) + unsigned(xl);
}

// ------------------------------------------- layer 1 (1 -> 32) on MFMA
// h[v][co] = relu(b[co] + sum_tap w[tap][co] x[v + off(tap)]) as a GEMM with
// K = the 27 taps (padded to 32): the f32 input is split into bf16 hi + lo
// (two products, ~2^-17 relative), the weights are bf16, FP32 accumulation.
// Same 4 x 4 x 16 tile as above; the (4+2) x (4+2) x (16+2) f32 halo and the
// f32 output tile share one LDS buffer.  w: [32 co][32 taps] bf16, taps >= 27 zero.
__global__ __launch_bounds__(256) void conv_in1_kernel(const float *__restrict__ x,
                                                       const bf16_t *__restrict__ w,
                                                       const float *__restrict__ bias,
                                                       bf16_t *__restrict__ h, Shape s)
{
    __shared__ __attribute__((aligned(16))) float lds[kTZ * kTY * kTX * kC];  // 32 KB
    float *hx_ = lds;  // halo, 648 floats, until the products are done
    const TileOrg o = tile_origin(s);
    const int b = o.b, z0 = o.z0, y0 = o.y0, x0 = o.x0;

    constexpr int kStage = (kHalo + 255) / 256;
    float val[kStage];
    if (halo_inside(s, o)) {
        const float *src = x + vox(s, b, z0 - 1, y0 - 1, x0 - 1);
#pragma unroll
        for (int i = 0; i < kStage; i++) {
            const int c = min(int(threadIdx.x) + 256 * i, kHalo - 1);
            const int hx = c % kHX, hy = (c / kHX) % kHY, hz = c / (kHX * kHY);
            val[i] = src[tile_off(s, hz, hy, hx)];
        }
    } else {
#pragma unroll
        for (int i = 0; i < kStage; i++) {
            const int c = min(int(threadIdx.x) + 256 * i, kHalo - 1);
            const int hx = c % kHX, hy = (c / kHX) % kHY, hz = c / (kHX * kHY);
            const int gz = z0 + hz - 1, gy = y0 + hy - 1, gx = x0 + hx - 1;
            const bool pad = !inside(s, gz, gy, gx);
            const int cz = min(max(gz, 0), s.D - 1), cy = min(max(gy, 0), s.H - 1);
            const int cx = min(max(gx, 0), s.W - 1);
            const float v = x[vox(s, b, cz, cy, cx)];
            val[i] = pad ? 0.0f : v;
        }
    }
#pragma unroll
    for (int i = 0; i < kStage; i++) {
        const int c = int(threadIdx.x) + 256 * i;
        if (c < kHalo) hx_[c] = val[i];
    }
    __syncthreads();

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = (lane >> 4) * 8;
    bf16x8_t bfrag[2];
#pragma unroll
    for (int f = 0; f < 2; f++)
        bfrag[f] = *reinterpret_cast<const bf16x8_t *>(w + (f * 16 + li) * kC + lk);
    f32x4_t acc[kTY][2];
#pragma unroll
    for (int m = 0; m < kTY; m++) {
        // A fragment: voxel (wave, m, li), taps lk .. lk + 7, split as
        // v = hi + lo by the packed hardware conversion (round to nearest even,
        // two values per v_cvt_pk_bf16_f32; v - hi is exact in FP32)
        float v[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int tap = lk + j;
            v[j] = 0.0f;
            if (tap < kTaps) {
                const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
                v[j] = hx_[((wave + 1 + dz) * kHY + m + 1 + dy) * kHX + li + 1 + dx];
            }
        }
        bf16x8_t ahi, alo;
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const f32x2_t pv = {v[2 * e], v[2 * e + 1]};
            const bf16x2_t h = __builtin_convertvector(pv, bf16x2_t);
            const bf16x2_t l = __builtin_convertvector(pv - __builtin_convertvector(h, f32x2_t), bf16x2_t);
            ahi[2 * e] = h[0];
            ahi[2 * e + 1] = h[1];
            alo[2 * e] = l[0];
            alo[2 * e + 1] = l[1];
        }
#pragma unroll
        for (int f = 0; f < 2; f++) {
            acc[m][f] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, bfrag[f],
                                                                 f32x4_t{0.0f, 0.0f, 0.0f, 0.0f}, 0, 0, 0);
            acc[m][f] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo, bfrag[f], acc[m][f], 0, 0, 0);
        }
    }
    __syncthreads();  // halo reads done: the buffer becomes the output tile
    float *tile = lds;  // [256 voxels][32] f32
#pragma unroll
    for (int m = 0; m < kTY; m++)
#pragma unroll
        for (int f = 0; f < 2; f++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int vl = (wave * kTY + m) * kTX + (lane >> 4) * 4 + j;
                tile[vl * kC + f * 16 + li] = acc[m][f][j];
            }
    __syncthreads();
    // every pass of a thread writes the same 8 channels (256 is a multiple of 4)
    const int part = threadIdx.x & 3;
    float bv[8];
#pragma unroll
    for (int e = 0; e < 8; e++) bv[e] = bias[part * 8 + e];
    const bool whole = tile_inside(s, o);
    uint4 *dst = reinterpret_cast<uint4 *>(h + vox(s, b, z0, y0, x0) * kC) + part;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int vl = (int(threadIdx.x) >> 2) + 64 * k;
        const int xl = vl % kTX, yl = (vl / kTX) % kTY, zl = vl / (kTX * kTY);
        if (!whole && !inside(s, z0 + zl, y0 + yl, x0 + xl)) continue;
        const float *src = tile + vl * kC + part * 8;
        unsigned w2[4];
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const float a0 = fmaxf(src[2 * e] + bv[2 * e], 0.0f);
            const float a1 = fmaxf(src[2 * e + 1] + bv[2 * e + 1], 0.0f);
            w2[e] = pack_bf16x2(a0, a1);
        }
        dst[size_t(tile_off(s, zl, yl, xl)) * 4] = make_uint4(w2[0], w2[1], w2[2], w2[3]);
    }
}

template <int MODE>
__global__ __launch_bounds__(256) void conv_mid_kernel(const bf16_t *__restrict__ in,
                                                       const bf16_t *__restrict__ w,  // [27][n][k]
                                                       const float *__restrict__ bias,
                                                       const bf16_t *__restrict__ mask,
                                                       bf16_t *__restrict__ out, Shape s,
                                                       const float *__restrict__ xres,
                                                       float *__restrict__ yout)
{
    constexpr int NF = MODE == 2 ? 1 : 2;  // N fragments of 16 output channels
    __shared__ __attribute__((aligned(16))) bf16_t halo[kHalo * kC];
    const TileOrg o = tile_origin(s);
    const int b = o.b, z0 = o.z0, y0 = o.y0, x0 = o.x0;

    // stage the halo: 648 voxels x 4 chunks of 16 bytes, every load of a
    // thread issued before its LDS writes (one memory latency, not eleven)
    constexpr int kStage = (kHalo * 4 + 255) / 256;
    uint4 val[kStage];
    bool pad[kStage];
    if (halo_inside(s, o)) {
        // chunk (hz, hy, hx, part) at a 32-bit byte offset from the halo's first voxel
        const char *src = reinterpret_cast<const char *>(in + vox(s, b, z0 - 1, y0 - 1, x0 - 1) * kC);
#pragma unroll
        for (int i = 0; i < kStage; i++) {
            const int c = min(int(threadIdx.x) + 256 * i, kHalo * 4 - 1);
            const int hv = c >> 2, part = c & 3;
            const int hx = hv % kHX, hy = (hv / kHX) % kHY, hz = hv / (kHX * kHY);
            pad[i] = false;
            val[i] = *reinterpret_cast<const uint4 *>(src + (tile_off(s, hz, hy, hx) * (kC * 2) + part * 16));
        }
    } else {
        // loads from clamped, always valid addresses; the zero padding is
        // applied when the chunk is written
#pragma unroll
        for (int i = 0; i < kStage; i++) {
            const int c = min(int(threadIdx.x) + 256 * i, kHalo * 4 - 1);
            const int hv = c >> 2, part = c & 3;
            const int hx = hv % kHX, hy = (hv / kHX) % kHY, hz = hv / (kHX * kHY);
            const int gz = z0 + hz - 1, gy = y0 + hy - 1, gx = x0 + hx - 1;
            pad[i] = !inside(s, gz, gy, gx);
            const int cz = min(max(gz, 0), s.D - 1), cy = min(max(gy, 0), s.H - 1);
            const int cx = min(max(gx, 0), s.W - 1);
            val[i] = reinterpret_cast<const uint4 *>(in + vox(s, b, cz, cy, cx) * kC)[part];
        }
    }
#pragma unroll
    for (int i = 0; i < kStage; i++) {
        const int c = int(threadIdx.x) + 256 * i;
        if (c < kHalo * 4)
            reinterpret_cast<uint4 *>(halo)[c] = pad[i] ? make_uint4(0u, 0u, 0u, 0u) : val[i];
    }
    __syncthreads();

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = (lane >> 4) * 8;
    f32x4_t acc[kTY][NF];
#pragma unroll
    for (int m = 0; m < kTY; m++)
#pragma unroll
        for (int f = 0; f < NF; f++) acc[m][f] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};

    // weight fragments of tap + 1 are loaded while tap's products run
    constexpr int wn = NF * 16;  // rows (output channels) per tap of w
    // byte offsets from the uniform weight base: one 32-bit add per tap
    const char *wbase = reinterpret_cast<const char *>(w);
    const unsigned wlane = unsigned(li * kC + lk) * 2;
    auto load_b = [&](int tap, bf16x8_t (&bf)[NF]) {
#pragma unroll
        for (int f = 0; f < NF; f++)
            bf[f] = *reinterpret_cast<const bf16x8_t *>(wbase + (wlane + unsigned((tap * wn + f * 16) * kC * 2)));
    };
    constexpr int kPD = 3;  // taps of weight fragments in flight
    bf16x8_t bq[kPD][NF];
#pragma unroll
    for (int p = 0; p < kPD; p++) load_b(p, bq[p]);
    // A fragments (voxels x input channels) of one tap from the halo
    const bf16_t *hb = halo + (wave * kHY * kHX + li) * kC + lk;  // this lane's halo base
    auto load_a = [&](int tap, bf16x8_t (&af)[kTY]) {
        const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
#pragma unroll
        for (int m = 0; m < kTY; m++) {
            af[m] = *reinterpret_cast<const bf16x8_t *>(hb + (((1 + dz) * kHY + m + 1 + dy) * kHX + 1 + dx) * kC);
        }
    };
    // software pipeline: the halo reads of tap + 1 and the weight loads of
    // tap + kPD are issued before tap's products, and the scheduler may not
    // move them past those products (sched_barrier)
    bf16x8_t a_cur[kTY], a_nxt[kTY];
    load_a(0, a_cur);
#pragma unroll
    for (int tap = 0; tap < kTaps; tap++) {
        bf16x8_t bfrag[NF];
#pragma unroll
        for (int f = 0; f < NF; f++) bfrag[f] = bq[tap % kPD][f];
        if (tap + kPD < kTaps) load_b(tap + kPD, bq[tap % kPD]);
        if (tap + 1 < kTaps) load_a(tap + 1, a_nxt);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int m = 0; m < kTY; m++)
#pragma unroll
            for (int f = 0; f < NF; f++)
                acc[m][f] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a_cur[m], bfrag[f], acc[m][f], 0, 0, 0);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int m = 0; m < kTY; m++) a_cur[m] = a_nxt[m];
    }
    __syncthreads();  // halo no longer needed: reuse it for the f32 output tile
    float *tile = reinterpret_cast<float *>(halo);  // [256 voxels][32] f32 = 32 KB
#pragma unroll
    for (int m = 0; m < kTY; m++)
#pragma unroll
        for (int f = 0; f < NF; f++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int xl = (lane >> 4) * 4 + j;                 // voxel row within the fragment
                const int vl = (wave * kTY + m) * kTX + xl;          // local voxel
                tile[vl * kC + f * 16 + li] = acc[m][f][j];
            }
    __syncthreads();
    const bool whole = tile_inside(s, o);
    const long tile0 = vox(s, b, z0, y0, x0);  // the tile's first voxel
    if constexpr (MODE == 2) {
        // channel 0 of the tile, one output voxel per thread and pass
        const float *xr = xres + tile0;
        float *yo = yout + tile0;
#pragma unroll
        for (int k = 0; k < 1; k++) {
            const int vl = int(threadIdx.x) + 256 * k;
            const int xl = vl % kTX, yl = (vl / kTX) % kTY, zl = vl / (kTX * kTY);
            if (whole || inside(s, z0 + zl, y0 + yl, x0 + xl)) {
                const unsigned v = tile_off(s, zl, yl, xl);
                yo[v] = xr[v] + (bias[0] + tile[vl * kC]);
            }
        }
        return;
    }
    // every pass of a thread handles the same 8 channels (256 is a multiple of 4)
    const int part = threadIdx.x & 3;
    float bv[8];
    if constexpr (MODE == 0) {
#pragma unroll
        for (int e = 0; e < 8; e++) bv[e] = bias[part * 8 + e];
    }
    uint4 *dst = reinterpret_cast<uint4 *>(out + tile0 * kC) + part;
    const uint4 *mk4 = MODE == 1 ? reinterpret_cast<const uint4 *>(mask + tile0 * kC) + part : nullptr;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int vl = (int(threadIdx.x) >> 2) + 64 * k;
        const int xl = vl % kTX, yl = (vl / kTX) % kTY, zl = vl / (kTX * kTY);
        if (!whole && !inside(s, z0 + zl, y0 + yl, x0 + xl)) continue;
        const size_t v4 = size_t(tile_off(s, zl, yl, xl)) * 4;
        const float *src = tile + vl * kC + part * 8;
        unsigned w2[4];
        if constexpr (MODE == 0) {
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const float a0 = fmaxf(src[2 * e] + bv[2 * e], 0.0f);
                const float a1 = fmaxf(src[2 * e + 1] + bv[2 * e + 1], 0.0f);
                w2[e] = pack_bf16x2(a0, a1);
            }
        } else {
            const uint4 mk = mk4[v4];
            const unsigned mw[4] = {mk.x, mk.y, mk.z, mk.w};
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const float a0 = bf2f(bf16_t(mw[e] & 0xffffu)) > 0.0f ? src[2 * e] : 0.0f;
                const float a1 = bf2f(bf16_t(mw[e] >> 16)) > 0.0f ? src[2 * e + 1] : 0.0f;
                w2[e] = pack_bf16x2(a0, a1);
            }
        }
        dst[v4] = make_uint4(w2[0], w2[1], w2[2], w2[3]);
    }
}

// ------------------------------------------------------------ backward
//
// Every reduction over voxels (loss, weight and bias gradients) is split into
// a fixed number of chunks that depends only on the shape: each chunk writes
// its partial sums to its own row of a [chunks][params] buffer, and
// reduce_partials_kernel adds the rows in chunk order.  No float atomics: a
// loss / gradient evaluation is bit-reproducible, run after run.

constexpr int kMaxChunks = 256;

__device__ __forceinline__ void chunk_range(long n, int chunk, int n_chunks, long &v0, long &v1)
{
    v0 = n * chunk / n_chunks;
    v1 = n * (chunk + 1) / n_chunks;
}

// Sum of a 256-thread block's values in a fixed tree order (result on thread 0).
__device__ __forceinline__ float block_sum(float x, float *red)
{
    red[threadIdx.x] = x;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (int(threadIdx.x) < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    return red[0];
}

// out[p] = scale * sum_c part[c][p] in a fixed order: a block of 256
// threads covers P consecutive p (P = 64, 32 or 1) and 256 / P chunk groups;
// group k sums the chunks c = k, k + 256/P, ... ascending, and the groups'
// sums are added in group order.
__global__ __launch_bounds__(256) void reduce_partials_kernel(const float *__restrict__ part,
                                                              int n_chunks, long np, int P,
                                                              float *__restrict__ out, float scale)
{
    __shared__ float red[256];
    const int pl = int(threadIdx.x) % P, grp = int(threadIdx.x) / P, n_grp = 256 / P;
    const long p = long(blockIdx.x) * P + pl;
    float acc = 0.0f;
    if (p < np) {
#pragma unroll 8
        for (int c = grp; c < n_chunks; c += n_grp) acc += part[size_t(c) * np + p];
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    if (int(threadIdx.x) < P && p < np) {
        float sum = 0.0f;
        for (int k = 0; k < n_grp; k++) sum += red[k * P + threadIdx.x];
        out[p] = sum * scale;
    }
}

// loss = mean((y - t)^2); dy = 2 (y - t) / N.  Chunk c (one block) writes
// its sum of squares to part[c].
__global__ __launch_bounds__(256) void mse_kernel(const float *__restrict__ y,
                                                  const float *__restrict__ t,
                                                  float *__restrict__ dy, long n, int n_chunks,
                                                  float *__restrict__ part)
{
    __shared__ float red[256];
    long v0, v1;
    chunk_range(n, blockIdx.x, n_chunks, v0, v1);
    float acc = 0.0f;
    for (long i = v0 + threadIdx.x; i < v1; i += 256) {
        const float d = y[i] - t[i];
        dy[i] = 2.0f * d / float(n);
        acc = fmaf(d, d, acc);
    }
    const float sum = block_sum(acc, red);
    if (threadIdx.x == 0) part[blockIdx.x] = sum;
}

// dh5[v][ci] = (h5[v][ci] > 0) * sum_tap w6[tap][ci] dy[v - off(tap)]
__global__ __launch_bounds__(256) void dgrad_out1_kernel(const float *__restrict__ dy,
                                                         const float *__restrict__ w,
                                                         const bf16_t *__restrict__ h,
                                                         bf16_t *__restrict__ dh, Shape s)
{
    __shared__ float sw[kTaps * kC];
    for (int i = threadIdx.x; i < kTaps * kC; i += blockDim.x) sw[i] = w[i];
    __syncthreads();
    const long n = long(s.B) * s.D * s.H * s.W;
    const long v = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const int xx = int(v % s.W), yy = int((v / s.W) % s.H), zz = int((v / (long(s.W) * s.H)) % s.D);
    const int b = int(v / (long(s.W) * s.H * s.D));
    float g[kTaps];
#pragma unroll
    for (int t = 0; t < kTaps; t++) {
        const int dz = t / 9 - 1, dy2 = (t / 3) % 3 - 1, dx = t % 3 - 1;
        g[t] = inside(s, zz - dz, yy - dy2, xx - dx) ? dy[vox(s, b, zz - dz, yy - dy2, xx - dx)] : 0.0f;
    }
    const uint4 *hm = reinterpret_cast<const uint4 *>(h + v * kC);
    uint4 *dst = reinterpret_cast<uint4 *>(dh + v * kC);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint4 q = hm[k];
        const unsigned mw[4] = {q.x, q.y, q.z, q.w};
        unsigned o[4];
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int ci = k * 8 + 2 * e;
            float a0 = 0.0f, a1 = 0.0f;
#pragma unroll
            for (int t = 0; t < kTaps; t++) {
                a0 = fmaf(sw[t * kC + ci], g[t], a0);
                a1 = fmaf(sw[t * kC + ci + 1], g[t], a1);
            }
            a0 = bf2f(bf16_t(mw[e] & 0xffffu)) > 0.0f ? a0 : 0.0f;
            a1 = bf2f(bf16_t(mw[e] >> 16)) > 0.0f ? a1 : 0.0f;
            o[e] = unsigned(f2bf(a0)) | (unsigned(f2bf(a1)) << 16);
        }
        dst[k] = make_uint4(o[0], o[1], o[2], o[3]);
    }
}

// Weight gradient of a 32 -> 32 layer on MFMA:
//   dw[tap][co][ci] = sum_v g[v][co] a[v + off(tap)][ci]
// is 27 GEMMs of 32 x 32 with K = voxels, on v_mfma_f32_16x16x32_bf16
// (A = g^T, B = the shifted a; bf16 operands, FP32 accumulation).  A
// workgroup owns one half of ci (blockIdx.y) and a contiguous range of the
// conv tiles (4 z x 4 y x 16 x voxels, K = 256 per tile, 8 k-steps of 32);
// its partial sums stay in registers across the tiles and go to row
// blockIdx.x of the partial buffer at the end.  K must be contiguous per
// lane in both operands, so the tiles are staged transposed in LDS:
//   gT[co][v]                  v = (z * 4 + y) * 16 + x of the tile
//   aC[dx + 1][ci][hz][hy][x]  = a(x0 + x + dx, y0 + hy - 1, z0 + hz - 1)
// one copy of the ci half's halo per x shift, so that every tap's B fragment
// (8 consecutive x) is one aligned 16-byte read.  Wave w owns the taps
// w, w + 4, ... (7 or 6) and both co fragments: 56 accumulator registers.
// Lane (q = lane / 16) covers k = 8 q .. 8 q + 7 of a k-step s: voxel row
// 2 s + q / 2 of the tile, x = 8 (q % 2) .. + 7.  The pitches are odd
// multiples of 16 bytes, so the 16 lanes of a fragment row read 16-byte
// words in different bank groups.
constexpr int kGP = kTZ * kTY * kTX + 8;  // 264 bf16 per co row of gT (528 B = 33 x 16 B)
constexpr int kAP = kHZ * kHY * kTX + 8;  // 584 bf16 per ci row of aC (1168 B = 73 x 16 B)
constexpr int kHaloItems = kHalo * 2;     // halo voxels x 2 chunks of 8 channels (one ci half)

__global__ __launch_bounds__(256) void wgrad_mid_kernel(const bf16_t *__restrict__ g,
                                                        const bf16_t *__restrict__ a,
                                                        float *__restrict__ part, Shape s,
                                                        int n_tiles, int n_chunks)
{
    __shared__ __attribute__((aligned(16))) bf16_t gT[kC * kGP];      // 16.5 KB
    __shared__ __attribute__((aligned(16))) bf16_t aC[3 * 16 * kAP];  // 54.8 KB
    const int chunk = blockIdx.x, half = blockIdx.y;
    const int t_begin = int(long(n_tiles) * chunk / n_chunks);
    const int t_end = int(long(n_tiles) * (chunk + 1) / n_chunks);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, q = lane >> 4;
    const int tiles_x = (s.W + kTX - 1) / kTX, tiles_y = (s.H + kTY - 1) / kTY;
    const int tiles_z = (s.D + kTZ - 1) / kTZ;

    f32x4_t acc[7][2];
#pragma unroll
    for (int i = 0; i < 7; i++)
#pragma unroll
        for (int f = 0; f < 2; f++) acc[i][f] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};

    for (int tile = t_begin; tile < t_end; tile++) {
        long t = tile;
        const int tx = int(t % tiles_x);
        t /= tiles_x;
        const int ty = int(t % tiles_y);
        t /= tiles_y;
        const int tz = int(t % tiles_z);
        const int b = int(t / tiles_z);
        const int x0 = tx * kTX, y0 = ty * kTY, z0 = tz * kTZ;

        // loads first (one memory latency), then the transposing LDS writes
        uint4 gv[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int c = int(threadIdx.x) + 256 * i;  // < 1024
            const int vl = c >> 2, part_ = c & 3;
            const int gz = z0 + (vl >> 6), gy = y0 + ((vl >> 4) & 3), gx = x0 + (vl & 15);
            gv[i] = inside(s, gz, gy, gx)
                        ? reinterpret_cast<const uint4 *>(g + vox(s, b, gz, gy, gx) * kC)[part_]
                        : make_uint4(0u, 0u, 0u, 0u);
        }
        constexpr int kAS = (kHaloItems + 255) / 256;
        uint4 av[kAS];
#pragma unroll
        for (int i = 0; i < kAS; i++) {
            const int c = min(int(threadIdx.x) + 256 * i, kHaloItems - 1);
            const int hv = c >> 1, part_ = c & 1;
            const int hx = hv % kHX, hy = (hv / kHX) % kHY, hz = hv / (kHX * kHY);
            const int gz = z0 + hz - 1, gy = y0 + hy - 1, gx = x0 + hx - 1;
            av[i] = inside(s, gz, gy, gx)
                        ? reinterpret_cast<const uint4 *>(a + vox(s, b, gz, gy, gx) * kC + half * 16)[part_]
                        : make_uint4(0u, 0u, 0u, 0u);
        }
        __syncthreads();  // the previous tile's fragment reads are done
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int c = int(threadIdx.x) + 256 * i;
            const int vl = c >> 2, part_ = c & 3;
            const unsigned w4[4] = {gv[i].x, gv[i].y, gv[i].z, gv[i].w};
#pragma unroll
            for (int e = 0; e < 8; e++)
                gT[(part_ * 8 + e) * kGP + vl] = bf16_t(e & 1 ? w4[e >> 1] >> 16 : w4[e >> 1] & 0xffffu);
        }
#pragma unroll
        for (int i = 0; i < kAS; i++) {
            const int c = int(threadIdx.x) + 256 * i;
            if (c >= kHaloItems) break;
            const int hv = c >> 1, part_ = c & 1;
            const int hx = hv % kHX, hyz = hv / kHX;  // hyz = hz * 6 + hy
            const unsigned w4[4] = {av[i].x, av[i].y, av[i].z, av[i].w};
#pragma unroll
            for (int dxi = 0; dxi < 3; dxi++) {
                const int x = hx - dxi;
                if (x < 0 || x >= kTX) continue;
                bf16_t *d = aC + (dxi * 16 + part_ * 8) * kAP + hyz * kTX + x;
#pragma unroll
                for (int e = 0; e < 8; e++)
                    d[e * kAP] = bf16_t(e & 1 ? w4[e >> 1] >> 16 : w4[e >> 1] & 0xffffu);
            }
        }
        __syncthreads();

#pragma unroll
        for (int ks = 0; ks < 8; ks++) {
            const int row = 2 * ks + (q >> 1), x8 = (q & 1) * 8;
            const int zl = row >> 2, yl = row & 3;
            bf16x8_t af[2];
#pragma unroll
            for (int f = 0; f < 2; f++)
                af[f] = *reinterpret_cast<const bf16x8_t *>(gT + (f * 16 + li) * kGP + row * kTX + x8);
#pragma unroll
            for (int i = 0; i < 7; i++) {
                const int tap = wave + 4 * i;
                if (tap >= kTaps) break;  // wave-uniform
                const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
                const bf16x8_t bfr = *reinterpret_cast<const bf16x8_t *>(
                    aC + ((dx + 1) * 16 + li) * kAP + ((zl + dz + 1) * kHY + yl + dy + 1) * kTX + x8);
#pragma unroll
                for (int f = 0; f < 2; f++)
                    acc[i][f] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[f], bfr, acc[i][f], 0, 0, 0);
            }
        }
    }
    // C fragment: lane holds rows (co) 4 q + r, column (ci) li
    float *dst = part + size_t(chunk) * kTaps * kC * kC;
#pragma unroll
    for (int i = 0; i < 7; i++) {
        const int tap = wave + 4 * i;
        if (tap >= kTaps) break;
#pragma unroll
        for (int f = 0; f < 2; f++)
#pragma unroll
            for (int r = 0; r < 4; r++)
                dst[(size_t(tap) * kC + f * 16 + 4 * q + r) * kC + half * 16 + li] = acc[i][f][r];
    }
}

// Weight gradient of layer 1 (ci = x, one channel) and of layer 6 (co = y,
// one channel), on FP32 VALU (1/32 of a 32 -> 32 layer's work):
//   layer 1: dw[tap][c] = sum_v g1[v][c] x[v + off(tap)]
//   layer 6: dw[tap][c] = sum_v dy[v] h5[v + off(tap)][c]
// Workgroup = a contiguous range of the conv tiles (4 z x 4 y x 16 x); per
// tile the unshifted field is staged in LDS at the tile's voxels and the
// shifted one on the (4+2) x (4+2) x (16+2) halo.  Thread t owns channel
// t % 32 and the taps t / 32 + 8 k; its sums stay in registers across the
// tiles and go to part[chunk][tap][c].
template <bool LAYER6>
__global__ __launch_bounds__(256) void wgrad_io_kernel(const bf16_t *__restrict__ hid,
                                                       const float *__restrict__ one,
                                                       float *__restrict__ part, Shape s,
                                                       int n_tiles, int n_chunks)
{
    // layer 1: tile of g1 as f32 [256][32] + halo of x [648]
    // layer 6: tile of dy [256] + halo of h5 as bf16 [648][32]
    constexpr int kTileVox = kTZ * kTY * kTX;
    constexpr int kBytes = LAYER6 ? kTileVox * 4 + kHalo * kC * 2 : kTileVox * kC * 4 + kHalo * 4;
    __shared__ __attribute__((aligned(16))) unsigned char lds[kBytes];
    float *tile_f = reinterpret_cast<float *>(lds);                       // [256][32] or [256]
    float *halo_f = reinterpret_cast<float *>(lds + kTileVox * kC * 4);    // layer 1: [648]
    bf16_t *halo_h = reinterpret_cast<bf16_t *>(lds + kTileVox * 4);       // layer 6: [648][32]
    const int chunk = blockIdx.x;
    const int t_begin = int(long(n_tiles) * chunk / n_chunks);
    const int t_end = int(long(n_tiles) * (chunk + 1) / n_chunks);
    const int c = threadIdx.x & 31, tg = threadIdx.x >> 5;
    const int tiles_x = (s.W + kTX - 1) / kTX, tiles_y = (s.H + kTY - 1) / kTY;
    const int tiles_z = (s.D + kTZ - 1) / kTZ;
    int toff[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int tap = min(tg + 8 * k, kTaps - 1);
        toff[k] = ((tap / 9 - 1) * kHY + (tap / 3) % 3 - 1) * kHX + tap % 3 - 1;
    }
    const int nk = tg + 24 < kTaps ? 4 : 3;
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int tile = t_begin; tile < t_end; tile++) {
        long t = tile;
        const int tx = int(t % tiles_x);
        t /= tiles_x;
        const int ty = int(t % tiles_y);
        t /= tiles_y;
        const int tz = int(t % tiles_z);
        const int b = int(t / tiles_z);
        const int x0 = tx * kTX, y0 = ty * kTY, z0 = tz * kTZ;
        __syncthreads();  // the previous tile's reads are done
        if constexpr (!LAYER6) {
            for (int i = threadIdx.x; i < kTileVox * 4; i += 256) {
                const int vl = i >> 2, pt = i & 3;
                const int gz = z0 + (vl >> 6), gy = y0 + ((vl >> 4) & 3), gx = x0 + (vl & 15);
                uint4 u = make_uint4(0u, 0u, 0u, 0u);
                if (inside(s, gz, gy, gx)) u = reinterpret_cast<const uint4 *>(hid + vox(s, b, gz, gy, gx) * kC)[pt];
                const unsigned w4[4] = {u.x, u.y, u.z, u.w};
                float *d = tile_f + vl * kC + pt * 8;
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    d[2 * e] = bf2f(bf16_t(w4[e] & 0xffffu));
                    d[2 * e + 1] = bf2f(bf16_t(w4[e] >> 16));
                }
            }
            for (int i = threadIdx.x; i < kHalo; i += 256) {
                const int hx = i % kHX, hy = (i / kHX) % kHY, hz = i / (kHX * kHY);
                const int gz = z0 + hz - 1, gy = y0 + hy - 1, gx = x0 + hx - 1;
                halo_f[i] = inside(s, gz, gy, gx) ? one[vox(s, b, gz, gy, gx)] : 0.0f;
            }
        } else {
            for (int i = threadIdx.x; i < kTileVox; i += 256) {
                const int gz = z0 + (i >> 6), gy = y0 + ((i >> 4) & 3), gx = x0 + (i & 15);
                tile_f[i] = inside(s, gz, gy, gx) ? one[vox(s, b, gz, gy, gx)] : 0.0f;
            }
            for (int i = threadIdx.x; i < kHalo * 4; i += 256) {
                const int hv = i >> 2, pt = i & 3;
                const int hx = hv % kHX, hy = (hv / kHX) % kHY, hz = hv / (kHX * kHY);
                const int gz = z0 + hz - 1, gy = y0 + hy - 1, gx = x0 + hx - 1;
                uint4 u = make_uint4(0u, 0u, 0u, 0u);
                if (inside(s, gz, gy, gx)) u = reinterpret_cast<const uint4 *>(hid + vox(s, b, gz, gy, gx) * kC)[pt];
                reinterpret_cast<uint4 *>(halo_h)[i] = u;
            }
        }
        __syncthreads();
        for (int vl = 0; vl < kTileVox; vl++) {
            const int h = (((vl >> 6) + 1) * kHY + ((vl >> 4) & 3) + 1) * kHX + (vl & 15) + 1;
            if constexpr (!LAYER6) {
                const float gv = tile_f[vl * kC + c];
#pragma unroll
                for (int k = 0; k < 4; k++)
                    if (k < nk) acc[k] = fmaf(gv, halo_f[h + toff[k]], acc[k]);
            } else {
                const float dv = tile_f[vl];
#pragma unroll
                for (int k = 0; k < 4; k++)
                    if (k < nk) acc[k] = fmaf(dv, bf2f(halo_h[(h + toff[k]) * kC + c]), acc[k]);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 4; k++)
        if (k < nk) part[(size_t(chunk) * kTaps + tg + 8 * k) * kC + c] = acc[k];
}

// Bias gradients: db[c] = sum_v g[v][c] (bf16 channels-last, C = 32; part
// [chunk][32]) or sum_v g1[v] (part[chunk]).  Channels-last rows are read as
// 16-byte chunks: thread t takes chunk t % 4 (channels 8 (t % 4) ..) of every
// 64th voxel of the chunk's range.
__global__ __launch_bounds__(256) void bgrad_kernel(const bf16_t *__restrict__ g,
                                                    const float *__restrict__ g1,
                                                    float *__restrict__ part, long n,
                                                    int n_chunks)
{
    __shared__ float red[256 * 8];
    long v0, v1;
    chunk_range(n, blockIdx.x, n_chunks, v0, v1);
    if (!g) {
        float acc = 0.0f;
        for (long v = v0 + threadIdx.x; v < v1; v += 256) acc += g1[v];
        const float sum = block_sum(acc, red);
        if (threadIdx.x == 0) part[blockIdx.x] = sum;
        return;
    }
    const int pt = threadIdx.x & 3, lane_v = threadIdx.x >> 2;
    float acc[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    for (long v = v0 + lane_v; v < v1; v += 64) {
        const uint4 u = reinterpret_cast<const uint4 *>(g + v * kC)[pt];
        const unsigned w4[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
        for (int e = 0; e < 4; e++) {
            acc[2 * e] += bf2f(bf16_t(w4[e] & 0xffffu));
            acc[2 * e + 1] += bf2f(bf16_t(w4[e] >> 16));
        }
    }
#pragma unroll
    for (int e = 0; e < 8; e++) red[threadIdx.x * 8 + e] = acc[e];
    __syncthreads();
    if (threadIdx.x < 32) {
        const int ch = threadIdx.x, p = ch >> 3, e = ch & 7;
        float t = 0.0f;
        for (int l = 0; l < 64; l++) t += red[(l * 4 + p) * 8 + e];
        part[size_t(blockIdx.x) * kC + ch] = t;
    }
}

// Adam (FP32 master weights) and the bf16 operand copies of the 32 -> 32
// layers: wb[l][tap][co][ci] (forward) and wt[l][tap'][ci][co] with mirrored
// taps (data gradient).  Layers 1 and 6 use bf16-rounded FP32 weights.
__global__ __launch_bounds__(256) void adam_kernel(float *__restrict__ w, float *__restrict__ m,
                                                   float *__restrict__ v,
                                                   const float *__restrict__ g, long n, float lr,
                                                   float b1, float b2, float eps, float c1,
                                                   float c2, int update)
{
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n || !update) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ __launch_bounds__(256) void weights_to_operands_kernel(const float *__restrict__ w,
                                                                  bf16_t *__restrict__ wb,
                                                                  bf16_t *__restrict__ wt,
                                                                  float *__restrict__ wio,
                                                                  bf16_t *__restrict__ w6op,
                                                                  bf16_t *__restrict__ w1op)
{
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    const long n_mid = 4L * kTaps * kC * kC;
    if (i < n_mid) {
        const int l = int(i / (kTaps * kC * kC));
        const int r = int(i % (kTaps * kC * kC));
        const int tap = r / (kC * kC), co = (r / kC) % kC, ci = r % kC;
        const float x = w[kMid + l * kMidStride + r];
        const bf16_t bx = f2bf(x);
        wb[i] = bx;
        const int mtap = kTaps - 1 - tap;  // off(mtap) = -off(tap)
        wt[size_t(l) * kTaps * kC * kC + (size_t(mtap) * kC + ci) * kC + co] = bx;
    }
    if (i < long(kTaps) * kC) {
        wio[i] = bf2f(f2bf(w[kW1 + i]));
        wio[kTaps * kC + i] = bf2f(f2bf(w[kW6 + i]));
    }
    // input layer as an MFMA operand [32 co][32 taps], taps >= 27 zero
    if (i < long(kC) * kC) {
        const int co = int(i / kC), tap = int(i % kC);
        w1op[i] = tap < kTaps ? f2bf(w[kW1 + tap * kC + co]) : bf16_t(0);
    }
    // output layer as an MFMA operand [27][16][32]: column 0 = W6[tap][ci]
    if (i < long(kTaps) * 16 * kC) {
        const int tap = int(i / (16 * kC)), n = int((i / kC) % 16), ci = int(i % kC);
        w6op[i] = n == 0 ? f2bf(w[kW6 + tap * kC + ci]) : bf16_t(0);
    }
}

}  // namespace lct

using namespace lct;

struct lct_cnn {
    lct_ctx *ctx = nullptr;
    float *w = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;  // FP32 [kParams]
    bf16_t *wb = nullptr, *wt = nullptr;                             // 4 x 27 x 32 x 32
    bf16_t *w6op = nullptr;                                          // 27 x 16 x 32
    bf16_t *w1op = nullptr;                                          // 32 x 32
    float *wio = nullptr;                                            // rounded W1, W6
    bf16_t *act[6] = {};   // h1..h5 (index 1..5)
    bf16_t *grad[2] = {};  // ping-pong gradient activations
    float *dy = nullptr, *y = nullptr, *loss = nullptr;
    float *part = nullptr;  // [kMaxChunks][27 x 32 x 32] partial sums of the reductions
    long cap_vox = 0;
    long step = 0;
};

namespace {

int cnn_alloc(lct_cnn *n, long nvox)
{
    if (nvox <= n->cap_vox) return LCT_OK;
    LCT_CHECK_HIP(hipStreamSynchronize(n->ctx->stream));
    for (auto &p : n->act) if (p) { LCT_CHECK_HIP(hipFree(p)); p = nullptr; }
    for (auto &p : n->grad) if (p) { LCT_CHECK_HIP(hipFree(p)); p = nullptr; }
    if (n->dy) LCT_CHECK_HIP(hipFree(n->dy));
    if (n->y) LCT_CHECK_HIP(hipFree(n->y));
    n->dy = n->y = nullptr;
    n->cap_vox = 0;
    for (int l = 1; l <= 5; l++) LCT_CHECK_HIP(hipMalloc(&n->act[l], size_t(nvox) * kC * 2));
    for (auto &p : n->grad) LCT_CHECK_HIP(hipMalloc(&p, size_t(nvox) * kC * 2));
    LCT_CHECK_HIP(hipMalloc(&n->dy, size_t(nvox) * 4));
    LCT_CHECK_HIP(hipMalloc(&n->y, size_t(nvox) * 4));
    n->cap_vox = nvox;
    return LCT_OK;
}

int refresh_operands(lct_cnn *n)
{
    const long cnt = 4L * kTaps * kC * kC;
    hipLaunchKernelGGL(weights_to_operands_kernel, dim3(unsigned((cnt + 255) / 256)), dim3(256), 0,
                       n->ctx->stream, n->w, n->wb, n->wt, n->wio, n->w6op, n->w1op);
    LCT_CHECK_HIP(hipGetLastError());
    return LCT_OK;
}

unsigned blocks_for(long n) { return unsigned((n + 255) / 256); }

// Grid of the conv kernels (see tile_origin); the shapes the kernels take
int conv_grid(Shape s, dim3 *g)
{
    const long tx = (s.W + kTX - 1) / kTX, ty = (s.H + kTY - 1) / kTY;
    const long tzb = long(s.B) * ((s.D + kTZ - 1) / kTZ);
    if (long(s.H) * s.W > kMaxPlane)
        return set_error(LCT_ERR_UNSUPPORTED, "cnn: a plane of more than 2^23 voxels");
    if (ty > 65535 || tzb > 65535)
        return set_error(LCT_ERR_UNSUPPORTED, "cnn: more than 65535 tile rows or batch x tile planes");
    *g = dim3(unsigned(tx), unsigned(ty), unsigned(tzb));
    return LCT_OK;
}

int forward(lct_cnn *n, const float *x, Shape s, float *y)
{
    hipStream_t st = n->ctx->stream;
    dim3 grid;
    LCT_CHECK(conv_grid(s, &grid));
    hipLaunchKernelGGL(conv_in1_kernel, grid, dim3(256), 0, st, x, n->w1op,
                       n->w + kB1, n->act[1], s);
    for (int l = 0; l < 4; l++)
        hipLaunchKernelGGL(conv_mid_kernel<0>, grid, dim3(256), 0, st,
                           n->act[l + 1], n->wb + size_t(l) * kTaps * kC * kC,
                           n->w + kMid + l * kMidStride + kTaps * kC * kC, nullptr, n->act[l + 2], s,
                           nullptr, nullptr);
    hipLaunchKernelGGL(conv_mid_kernel<2>, grid, dim3(256), 0, st, n->act[5],
                       n->w6op, n->w + kB6, nullptr, nullptr, s, x, y);
    LCT_CHECK_HIP(hipGetLastError());
    return LCT_OK;
}

// Voxel chunks of the reductions: a function of the shape only, so the
// summation order (and every bit of the result) is the same on every run.
int voxel_chunks(long nv) { return int(std::min<long>(kMaxChunks, std::max<long>(1, (nv + 255) / 256))); }
long conv_tiles(Shape s)
{
    return long(s.B) * ((s.D + kTZ - 1) / kTZ) * ((s.H + kTY - 1) / kTY) * ((s.W + kTX - 1) / kTX);
}

void reduce_into(lct_cnn *n, int n_chunks, long np, float *out, float scale = 1.0f)
{
    const int P = np >= 64 ? 64 : np >= 32 ? 32 : 1;
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(unsigned((np + P - 1) / P)), dim3(256), 0,
                       n->ctx->stream, n->part, n_chunks, np, P, out, scale);
}

int backward(lct_cnn *n, const float *x, Shape s)
{
    const long nv = long(s.B) * s.D * s.H * s.W;
    hipStream_t st = n->ctx->stream;
    const int vc = voxel_chunks(nv);
    const long tiles = conv_tiles(s);
    if (tiles > 0x7fffffffL) return set_error(LCT_ERR_UNSUPPORTED, "cnn: too many voxel tiles");
    dim3 grid;
    LCT_CHECK(conv_grid(s, &grid));
    const int tc = int(std::min<long>(kMaxChunks, tiles));
    // layer 6
    hipLaunchKernelGGL(wgrad_io_kernel<true>, dim3(tc), dim3(256), 0, st, n->act[5], n->dy,
                       n->part, s, int(tiles), tc);
    reduce_into(n, tc, kTaps * kC, n->g + kW6);
    hipLaunchKernelGGL(bgrad_kernel, dim3(vc), dim3(256), 0, st, nullptr, n->dy, n->part, nv, vc);
    reduce_into(n, vc, 1, n->g + kB6);
    hipLaunchKernelGGL(dgrad_out1_kernel, dim3(blocks_for(nv)), dim3(256), 0, st, n->dy,
                       n->wio + kTaps * kC, n->act[5], n->grad[0], s);
    int cur = 0;
    // layers 5..2: grad[cur] holds dL/d(pre-activation of layer l+2) (masked)
    for (int l = 3; l >= 0; l--) {
        bf16_t *gk = n->grad[cur];
        float *gw = n->g + kMid + l * kMidStride;
        hipLaunchKernelGGL(wgrad_mid_kernel, dim3(tc, 2), dim3(256), 0, st, gk, n->act[l + 1],
                           n->part, s, int(tiles), tc);
        reduce_into(n, tc, long(kTaps) * kC * kC, gw);
        hipLaunchKernelGGL(bgrad_kernel, dim3(vc), dim3(256), 0, st, gk, nullptr, n->part, nv, vc);
        reduce_into(n, vc, kC, gw + kTaps * kC * kC);
        hipLaunchKernelGGL(conv_mid_kernel<1>, grid, dim3(256), 0, st, gk,
                           n->wt + size_t(l) * kTaps * kC * kC, nullptr, n->act[l + 1],
                           n->grad[cur ^ 1], s, nullptr, nullptr);
        cur ^= 1;
    }
    // layer 1: grad[cur] = dL/d(pre-activation of layer 1)
    hipLaunchKernelGGL(wgrad_io_kernel<false>, dim3(tc), dim3(256), 0, st, n->grad[cur], x,
                       n->part, s, int(tiles), tc);
    reduce_into(n, tc, kTaps * kC, n->g + kW1);
    hipLaunchKernelGGL(bgrad_kernel, dim3(vc), dim3(256), 0, st, n->grad[cur], nullptr, n->part, nv,
                       vc);
    reduce_into(n, vc, kC, n->g + kB1);
    LCT_CHECK_HIP(hipGetLastError());
    return LCT_OK;
}

bool shape_ok(int B, int D, int H, int W)
{
    return B > 0 && D > 0 && H > 0 && W > 0;
}

}  // namespace

extern "C" {

long lct_cnn_param_count(void) { return kParams; }

int lct_cnn_create(lct_ctx *ctx, uint64_t seed, lct_cnn **out)
{
    if (!ctx || !out) return set_error(LCT_ERR_INVALID, "lct_cnn_create: NULL argument");
    *out = nullptr;
    LCT_CHECK_HIP(hipSetDevice(ctx->device));
    lct_cnn *n = new lct_cnn();
    n->ctx = ctx;
    auto fail = [&](hipError_t e, const char *w) {
        lct_cnn_destroy(n);
        return hip_error(e, w);
    };
    hipError_t e;
    if ((e = hipMalloc(&n->w, kParams * 4)) != hipSuccess) return fail(e, "hipMalloc(w)");
    if ((e = hipMalloc(&n->m, kParams * 4)) != hipSuccess) return fail(e, "hipMalloc(m)");
    if ((e = hipMalloc(&n->v, kParams * 4)) != hipSuccess) return fail(e, "hipMalloc(v)");
    if ((e = hipMalloc(&n->g, kParams * 4)) != hipSuccess) return fail(e, "hipMalloc(g)");
    if ((e = hipMalloc(&n->wb, 4L * kTaps * kC * kC * 2)) != hipSuccess) return fail(e, "hipMalloc(wb)");
    if ((e = hipMalloc(&n->wt, 4L * kTaps * kC * kC * 2)) != hipSuccess) return fail(e, "hipMalloc(wt)");
    if ((e = hipMalloc(&n->wio, 2L * kTaps * kC * 4)) != hipSuccess) return fail(e, "hipMalloc(wio)");
    if ((e = hipMalloc(&n->w6op, long(kTaps) * 16 * kC * 2)) != hipSuccess)
        return fail(e, "hipMalloc(w6op)");
    if ((e = hipMalloc(&n->w1op, long(kC) * kC * 2)) != hipSuccess) return fail(e, "hipMalloc(w1op)");
    if ((e = hipMalloc(&n->loss, 4)) != hipSuccess) return fail(e, "hipMalloc(loss)");
    if ((e = hipMalloc(&n->part, size_t(kMaxChunks) * kTaps * kC * kC * 4)) != hipSuccess)
        return fail(e, "hipMalloc(part)");
    // He-normal initialisation from the seed; the last layer starts at 1/100 so
    // the initial correction is close to zero; zero biases.
    std::vector<float> h(kParams, 0.0f);
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    const double s1 = std::sqrt(2.0 / kTaps), sm = std::sqrt(2.0 / (kTaps * kC));
    for (long i = 0; i < kTaps * kC; i++) h[kW1 + i] = float(nd(rng) * s1);
    for (int l = 0; l < 4; l++)
        for (long i = 0; i < long(kTaps) * kC * kC; i++) h[kMid + l * kMidStride + i] = float(nd(rng) * sm);
    for (long i = 0; i < kTaps * kC; i++) h[kW6 + i] = float(nd(rng) * sm * 0.01);
    if ((e = hipMemcpy(n->w, h.data(), kParams * 4, hipMemcpyHostToDevice)) != hipSuccess)
        return fail(e, "hipMemcpy(w)");
    if ((e = hipMemset(n->m, 0, kParams * 4)) != hipSuccess) return fail(e, "hipMemset(m)");
    if ((e = hipMemset(n->v, 0, kParams * 4)) != hipSuccess) return fail(e, "hipMemset(v)");
    int rc = refresh_operands(n);
    if (rc != LCT_OK) {
        lct_cnn_destroy(n);
        return rc;
    }
    *out = n;
    return LCT_OK;
}

int lct_cnn_destroy(lct_cnn *n)
{
    if (!n) return LCT_OK;
    // the context may already be destroyed: wait on the device, not its stream
    (void)hipDeviceSynchronize();
    for (void *p : {static_cast<void *>(n->w), static_cast<void *>(n->m), static_cast<void *>(n->v),
                    static_cast<void *>(n->g), static_cast<void *>(n->wb), static_cast<void *>(n->wt),
                    static_cast<void *>(n->wio), static_cast<void *>(n->w6op),
                    static_cast<void *>(n->w1op),
                    static_cast<void *>(n->loss), static_cast<void *>(n->part),
                    static_cast<void *>(n->dy), static_cast<void *>(n->y)})
        if (p) (void)hipFree(p);
    for (auto p : n->act) if (p) (void)hipFree(p);
    for (auto p : n->grad) if (p) (void)hipFree(p);
    delete n;
    return LCT_OK;
}

int lct_cnn_get_params(lct_cnn *n, float *h_params)
{
    if (!n || !h_params) return set_error(LCT_ERR_INVALID, "lct_cnn_get_params: NULL argument");
    LCT_CHECK_HIP(hipMemcpyAsync(h_params, n->w, kParams * 4, hipMemcpyDeviceToHost, n->ctx->stream));
    LCT_CHECK_HIP(hipStreamSynchronize(n->ctx->stream));
    return LCT_OK;
}

int lct_cnn_set_params(lct_cnn *n, const float *h_params)
{
    if (!n || !h_params) return set_error(LCT_ERR_INVALID, "lct_cnn_set_params: NULL argument");
    LCT_CHECK(upload(n->ctx, n->w, h_params, kParams * 4));
    return refresh_operands(n);
}

int lct_cnn_get_grads(lct_cnn *n, float *h_grads)
{
    if (!n || !h_grads) return set_error(LCT_ERR_INVALID, "lct_cnn_get_grads: NULL argument");
    LCT_CHECK_HIP(hipMemcpyAsync(h_grads, n->g, kParams * 4, hipMemcpyDeviceToHost, n->ctx->stream));
    LCT_CHECK_HIP(hipStreamSynchronize(n->ctx->stream));
    return LCT_OK;
}

int lct_cnn_forward(lct_cnn *n, const float *d_x, int B, int D, int H, int W, float *d_y)
{
    if (!n || !d_x || !d_y || !shape_ok(B, D, H, W))
        return set_error(LCT_ERR_INVALID, "lct_cnn_forward: bad argument");
    LCT_CHECK_HIP(hipSetDevice(n->ctx->device));
    LCT_CHECK(cnn_alloc(n, long(B) * D * H * W));
    return forward(n, d_x, Shape{B, D, H, W}, d_y);
}

int lct_cnn_loss_grad(lct_cnn *n, const float *d_x, const float *d_t, int B, int D, int H, int W,
                      float *h_loss)
{
    if (!n || !d_x || !d_t || !shape_ok(B, D, H, W))
        return set_error(LCT_ERR_INVALID, "lct_cnn_loss_grad: bad argument");
    LCT_CHECK_HIP(hipSetDevice(n->ctx->device));
    const long nv = long(B) * D * H * W;
    LCT_CHECK(cnn_alloc(n, nv));
    const Shape s{B, D, H, W};
    LCT_CHECK(forward(n, d_x, s, n->y));
    // every gradient entry and the loss are written (not accumulated) by the
    // fixed-order reductions below
    const int vc = voxel_chunks(nv);
    hipLaunchKernelGGL(mse_kernel, dim3(vc), dim3(256), 0, n->ctx->stream, n->y, d_t, n->dy, nv, vc,
                       n->part);
    reduce_into(n, vc, 1, n->loss, 1.0f / float(nv));
    LCT_CHECK(backward(n, d_x, s));
    if (h_loss) {
        LCT_CHECK_HIP(hipMemcpyAsync(h_loss, n->loss, 4, hipMemcpyDeviceToHost, n->ctx->stream));
        LCT_CHECK_HIP(hipStreamSynchronize(n->ctx->stream));
    }
    return LCT_OK;
}

int lct_cnn_adam_step(lct_cnn *n, float lr, float beta1, float beta2, float eps)
{
    if (!n) return set_error(LCT_ERR_INVALID, "lct_cnn_adam_step: NULL argument");
    LCT_CHECK_HIP(hipSetDevice(n->ctx->device));
    n->step++;
    const float c1 = float(1.0 - std::pow(double(beta1), double(n->step)));
    const float c2 = float(1.0 - std::pow(double(beta2), double(n->step)));
    hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(kParams)), dim3(256), 0, n->ctx->stream, n->w,
                       n->m, n->v, n->g, kParams, lr, beta1, beta2, eps, c1, c2, 1);
    LCT_CHECK_HIP(hipGetLastError());
    return refresh_operands(n);
}

}  // extern "C"
