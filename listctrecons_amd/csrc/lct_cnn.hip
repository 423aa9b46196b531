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
// the 27 taps; all weight gradients run on FP32 VALU.
#include <hip/hip_bf16.h>

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
    const int tiles_x = (s.W + kTX - 1) / kTX, tiles_y = (s.H + kTY - 1) / kTY;
    const int tiles_z = (s.D + kTZ - 1) / kTZ;
    long t = blockIdx.x;
    const int tx = int(t % tiles_x);
    t /= tiles_x;
    const int ty = int(t % tiles_y);
    t /= tiles_y;
    const int tz = int(t % tiles_z);
    const int b = int(t / tiles_z);
    const int x0 = tx * kTX, y0 = ty * kTY, z0 = tz * kTZ;

    constexpr int kStage = (kHalo + 255) / 256;
    float val[kStage];
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
        // A fragment: voxel (wave, m, li), taps lk .. lk + 7
        bf16x8_t ahi, alo;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int tap = lk + j;
            float v = 0.0f;
            if (tap < kTaps) {
                const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
                v = hx_[((wave + 1 + dz) * kHY + m + 1 + dy) * kHX + li + 1 + dx];
            }
            const bf16_t hi = f2bf(v);
            ahi[j] = __builtin_bit_cast(__bf16, hi);
            alo[j] = __builtin_bit_cast(__bf16, f2bf(v - bf2f(hi)));
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
    for (int c = threadIdx.x; c < kTZ * kTY * kTX * 4; c += 256) {
        const int vl = c >> 2, part = c & 3;
        const int xl = vl % kTX, yl = (vl / kTX) % kTY, zl = vl / (kTX * kTY);
        const int gz = z0 + zl, gy = y0 + yl, gx = x0 + xl;
        if (!inside(s, gz, gy, gx)) continue;
        const float *src = tile + vl * kC + part * 8;
        unsigned o[4];
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int ch = part * 8 + 2 * e;
            const float a0 = fmaxf(src[2 * e] + bias[ch], 0.0f);
            const float a1 = fmaxf(src[2 * e + 1] + bias[ch + 1], 0.0f);
            o[e] = unsigned(f2bf(a0)) | (unsigned(f2bf(a1)) << 16);
        }
        reinterpret_cast<uint4 *>(h + vox(s, b, gz, gy, gx) * kC)[part] = make_uint4(o[0], o[1], o[2], o[3]);
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
    const int tiles_x = (s.W + kTX - 1) / kTX, tiles_y = (s.H + kTY - 1) / kTY;
    const int tiles_z = (s.D + kTZ - 1) / kTZ;
    long t = blockIdx.x;
    const int tx = int(t % tiles_x);
    t /= tiles_x;
    const int ty = int(t % tiles_y);
    t /= tiles_y;
    const int tz = int(t % tiles_z);
    const int b = int(t / tiles_z);
    const int x0 = tx * kTX, y0 = ty * kTY, z0 = tz * kTZ;

    // stage the halo: 648 voxels x 4 chunks of 16 bytes, every load of a
    // thread issued before its LDS writes (one memory latency, not eleven)
    // (loads from clamped, always valid addresses; the zero padding is
    // applied when the chunk is written)
    constexpr int kStage = (kHalo * 4 + 255) / 256;
    uint4 val[kStage];
    bool pad[kStage];
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
    auto load_b = [&](int tap, bf16x8_t (&bf)[NF]) {
#pragma unroll
        for (int f = 0; f < NF; f++)
            bf[f] = *reinterpret_cast<const bf16x8_t *>(w + (size_t(tap) * wn + f * 16 + li) * kC + lk);
    };
    bf16x8_t bnext[NF];
    load_b(0, bnext);
    for (int tap = 0; tap < kTaps; tap++) {
        const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
        bf16x8_t bfrag[NF];
#pragma unroll
        for (int f = 0; f < NF; f++) bfrag[f] = bnext[f];
        if (tap + 1 < kTaps) load_b(tap + 1, bnext);
#pragma unroll
        for (int m = 0; m < kTY; m++) {
            const int hz = wave + 1 + dz, hy = m + 1 + dy, hx = li + 1 + dx;
            const bf16x8_t afrag =
                *reinterpret_cast<const bf16x8_t *>(halo + ((hz * kHY + hy) * kHX + hx) * kC + lk);
#pragma unroll
            for (int f = 0; f < NF; f++)
                acc[m][f] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(afrag, bfrag[f], acc[m][f], 0, 0, 0);
        }
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
    if constexpr (MODE == 2) {
        // channel 0 of the tile, one output voxel per thread and pass
        for (int vl = int(threadIdx.x); vl < kTZ * kTY * kTX; vl += 256) {
            const int xl = vl % kTX, yl = (vl / kTX) % kTY, zl = vl / (kTX * kTY);
            const int gz = z0 + zl, gy = y0 + yl, gx = x0 + xl;
            if (inside(s, gz, gy, gx)) {
                const long gv = vox(s, b, gz, gy, gx);
                yout[gv] = xres[gv] + (bias[0] + tile[vl * kC]);
            }
        }
        return;
    }
    for (int c = threadIdx.x; c < kTZ * kTY * kTX * 4; c += 256) {
        const int vl = c >> 2, part = c & 3;
        const int xl = vl % kTX, yl = (vl / kTX) % kTY, zl = vl / (kTX * kTY);
        const int gz = z0 + zl, gy = y0 + yl, gx = x0 + xl;
        if (!inside(s, gz, gy, gx)) continue;
        const long gv = vox(s, b, gz, gy, gx);
        const float *src = tile + vl * kC + part * 8;
        unsigned o[4];
        if (MODE == 0) {
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int ch = part * 8 + 2 * e;
                const float a0 = fmaxf(src[2 * e] + bias[ch], 0.0f);
                const float a1 = fmaxf(src[2 * e + 1] + bias[ch + 1], 0.0f);
                o[e] = unsigned(f2bf(a0)) | (unsigned(f2bf(a1)) << 16);
            }
        } else {
            const uint4 mk = reinterpret_cast<const uint4 *>(mask + gv * kC)[part];
            const unsigned mw[4] = {mk.x, mk.y, mk.z, mk.w};
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const float a0 = bf2f(bf16_t(mw[e] & 0xffffu)) > 0.0f ? src[2 * e] : 0.0f;
                const float a1 = bf2f(bf16_t(mw[e] >> 16)) > 0.0f ? src[2 * e + 1] : 0.0f;
                o[e] = unsigned(f2bf(a0)) | (unsigned(f2bf(a1)) << 16);
            }
        }
        reinterpret_cast<uint4 *>(out + gv * kC)[part] = make_uint4(o[0], o[1], o[2], o[3]);
    }
}

// ------------------------------------------------------------ backward

// loss = mean((y - t)^2); dy = 2 (y - t) / N
__global__ __launch_bounds__(256) void mse_kernel(const float *__restrict__ y,
                                                  const float *__restrict__ t,
                                                  float *__restrict__ dy, long n,
                                                  float *__restrict__ loss)
{
    __shared__ float red[256];
    const long i = long(blockIdx.x) * blockDim.x + threadIdx.x;
    float d = 0.0f;
    if (i < n) {
        d = y[i] - t[i];
        dy[i] = 2.0f * d / float(n);
    }
    red[threadIdx.x] = d * d;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (int(threadIdx.x) < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, red[0] / float(n));
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

// Weight gradient of a 32 -> 32 layer for one tap over a chunk of voxels:
// dw[tap][co][ci] += sum_v g[v][co] a[v + off(tap)][ci]  (FP32 accumulation).
// Thread t owns co = t / 8 and ci = 4 (t % 8) .. +3; voxels are staged in LDS
// 128 at a time.
constexpr int kWgChunk = 128;
__global__ __launch_bounds__(256) void wgrad_mid_kernel(const bf16_t *__restrict__ g,
                                                        const bf16_t *__restrict__ a,
                                                        float *__restrict__ dw, Shape s,
                                                        long vox_per_wg)
{
    __shared__ __attribute__((aligned(16))) bf16_t sg[kWgChunk * kC];
    __shared__ __attribute__((aligned(16))) bf16_t sa[kWgChunk * kC];
    const int tap = blockIdx.y;
    const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
    const long n = long(s.B) * s.D * s.H * s.W;
    const long v_begin = long(blockIdx.x) * vox_per_wg;
    const long v_end = (v_begin + vox_per_wg < n ? v_begin + vox_per_wg : n);
    const int co = threadIdx.x >> 3, ci0 = (threadIdx.x & 7) * 4;
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (long vb = v_begin; vb < v_end; vb += kWgChunk) {
        // stage: 128 voxels x 4 chunks of 16 B for g and for the shifted a
        for (int c = threadIdx.x; c < kWgChunk * 4; c += 256) {
            const long v = vb + (c >> 2);
            const int part = c & 3;
            uint4 gv = make_uint4(0u, 0u, 0u, 0u), av = make_uint4(0u, 0u, 0u, 0u);
            if (v < v_end) {
                gv = reinterpret_cast<const uint4 *>(g + v * kC)[part];
                const int xx = int(v % s.W), yy = int((v / s.W) % s.H),
                          zz = int((v / (long(s.W) * s.H)) % s.D);
                const int b = int(v / (long(s.W) * s.H * s.D));
                if (inside(s, zz + dz, yy + dy, xx + dx))
                    av = reinterpret_cast<const uint4 *>(a + vox(s, b, zz + dz, yy + dy, xx + dx) * kC)[part];
            }
            reinterpret_cast<uint4 *>(sg)[c] = gv;
            reinterpret_cast<uint4 *>(sa)[c] = av;
        }
        __syncthreads();
        const int nv = int((v_end - vb < kWgChunk ? v_end - vb : kWgChunk));
        for (int i = 0; i < nv; i++) {
            const float gg = bf2f(sg[i * kC + co]);
            const uint2 q = *reinterpret_cast<const uint2 *>(sa + i * kC + ci0);
            acc[0] = fmaf(gg, bf2f(bf16_t(q.x & 0xffffu)), acc[0]);
            acc[1] = fmaf(gg, bf2f(bf16_t(q.x >> 16)), acc[1]);
            acc[2] = fmaf(gg, bf2f(bf16_t(q.y & 0xffffu)), acc[2]);
            acc[3] = fmaf(gg, bf2f(bf16_t(q.y >> 16)), acc[3]);
        }
        __syncthreads();
    }
    float *d = dw + (size_t(tap) * kC + co) * kC + ci0;
#pragma unroll
    for (int k = 0; k < 4; k++) atomicAdd(d + k, acc[k]);
}

// Weight gradient of layer 1 (ci = x, one channel) and of layer 6 (co = y, one
// channel): dw[tap][c] += sum_v G[v][c] A[v + off(tap)] (layer 1, G = g1, A = x)
// or sum_v dy[v] h5[v + off(tap)][c] (layer 6).  Thread = (channel, voxel lane).
template <bool LAYER6>
__global__ __launch_bounds__(256) void wgrad_io_kernel(const bf16_t *__restrict__ hid,
                                                       const float *__restrict__ one,
                                                       float *__restrict__ dw, Shape s,
                                                       long vox_per_wg)
{
    __shared__ float red[256];
    const int tap = blockIdx.y;
    const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
    const long n = long(s.B) * s.D * s.H * s.W;
    const long v_begin = long(blockIdx.x) * vox_per_wg;
    const long v_end = (v_begin + vox_per_wg < n ? v_begin + vox_per_wg : n);
    const int c = threadIdx.x & 31, lane_v = threadIdx.x >> 5;  // 8 voxel lanes
    float acc = 0.0f;
    for (long v = v_begin + lane_v; v < v_end; v += 8) {
        const int xx = int(v % s.W), yy = int((v / s.W) % s.H), zz = int((v / (long(s.W) * s.H)) % s.D);
        const int b = int(v / (long(s.W) * s.H * s.D));
        if (!inside(s, zz + dz, yy + dy, xx + dx)) continue;
        const long vs = vox(s, b, zz + dz, yy + dy, xx + dx);
        if (LAYER6)
            acc = fmaf(one[v], bf2f(hid[vs * kC + c]), acc);
        else
            acc = fmaf(bf2f(hid[v * kC + c]), one[vs], acc);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x < 32) {
        float t = 0.0f;
        for (int k = 0; k < 8; k++) t += red[k * 32 + threadIdx.x];
        atomicAdd(dw + tap * kC + threadIdx.x, t);
    }
}

// Bias gradients: db[c] += sum_v g[v][c] (bf16 channels-last, C = 32) or sum_v dy[v].
__global__ __launch_bounds__(256) void bgrad_kernel(const bf16_t *__restrict__ g,
                                                    const float *__restrict__ g1,
                                                    float *__restrict__ db, long n)
{
    __shared__ float red[256];
    const int c = threadIdx.x & 31, lane_v = threadIdx.x >> 5;
    float acc = 0.0f;
    const long stride = long(gridDim.x) * 8;
    for (long v = long(blockIdx.x) * 8 + lane_v; v < n; v += stride)
        acc += g ? bf2f(g[v * kC + c]) : (c == 0 ? g1[v] : 0.0f);
    red[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x < 32) {
        float t = 0.0f;
        for (int k = 0; k < 8; k++) t += red[k * 32 + threadIdx.x];
        if (g || threadIdx.x == 0) atomicAdd(db + threadIdx.x, t);
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

int forward(lct_cnn *n, const float *x, Shape s, float *y)
{
    hipStream_t st = n->ctx->stream;
    const long tiles = long(s.B) * ((s.D + kTZ - 1) / kTZ) * ((s.H + kTY - 1) / kTY) *
                       ((s.W + kTX - 1) / kTX);
    hipLaunchKernelGGL(conv_in1_kernel, dim3(unsigned(tiles)), dim3(256), 0, st, x, n->w1op,
                       n->w + kB1, n->act[1], s);
    for (int l = 0; l < 4; l++)
        hipLaunchKernelGGL(conv_mid_kernel<0>, dim3(unsigned(tiles)), dim3(256), 0, st,
                           n->act[l + 1], n->wb + size_t(l) * kTaps * kC * kC,
                           n->w + kMid + l * kMidStride + kTaps * kC * kC, nullptr, n->act[l + 2], s,
                           nullptr, nullptr);
    hipLaunchKernelGGL(conv_mid_kernel<2>, dim3(unsigned(tiles)), dim3(256), 0, st, n->act[5],
                       n->w6op, n->w + kB6, nullptr, nullptr, s, x, y);
    LCT_CHECK_HIP(hipGetLastError());
    return LCT_OK;
}

int backward(lct_cnn *n, const float *x, Shape s)
{
    const long nv = long(s.B) * s.D * s.H * s.W;
    hipStream_t st = n->ctx->stream;
    const long vpw = 8192;
    const unsigned chunks = unsigned((nv + vpw - 1) / vpw);
    // layer 6
    hipLaunchKernelGGL(wgrad_io_kernel<true>, dim3(chunks, kTaps), dim3(256), 0, st, n->act[5],
                       n->dy, n->g + kW6, s, vpw);
    hipLaunchKernelGGL(bgrad_kernel, dim3(256), dim3(256), 0, st, nullptr, n->dy, n->g + kB6, nv);
    hipLaunchKernelGGL(dgrad_out1_kernel, dim3(blocks_for(nv)), dim3(256), 0, st, n->dy,
                       n->wio + kTaps * kC, n->act[5], n->grad[0], s);
    const long tiles = long(s.B) * ((s.D + kTZ - 1) / kTZ) * ((s.H + kTY - 1) / kTY) *
                       ((s.W + kTX - 1) / kTX);
    int cur = 0;
    // layers 5..2: grad[cur] holds dL/d(pre-activation of layer l+2) (masked)
    for (int l = 3; l >= 0; l--) {
        bf16_t *gk = n->grad[cur];
        hipLaunchKernelGGL(wgrad_mid_kernel, dim3(chunks, kTaps), dim3(256), 0, st, gk,
                           n->act[l + 1], n->g + kMid + l * kMidStride, s, vpw);
        hipLaunchKernelGGL(bgrad_kernel, dim3(256), dim3(256), 0, st, gk, nullptr,
                           n->g + kMid + l * kMidStride + kTaps * kC * kC, nv);
        hipLaunchKernelGGL(conv_mid_kernel<1>, dim3(unsigned(tiles)), dim3(256), 0, st, gk,
                           n->wt + size_t(l) * kTaps * kC * kC, nullptr, n->act[l + 1],
                           n->grad[cur ^ 1], s, nullptr, nullptr);
        cur ^= 1;
    }
    // layer 1: grad[cur] = dL/d(pre-activation of layer 1)
    hipLaunchKernelGGL(wgrad_io_kernel<false>, dim3(chunks, kTaps), dim3(256), 0, st,
                       n->grad[cur], x, n->g + kW1, s, vpw);
    hipLaunchKernelGGL(bgrad_kernel, dim3(256), dim3(256), 0, st, n->grad[cur], nullptr,
                       n->g + kB1, nv);
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
                    static_cast<void *>(n->loss),
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
    LCT_CHECK_HIP(hipMemsetAsync(n->loss, 0, 4, n->ctx->stream));
    LCT_CHECK_HIP(hipMemsetAsync(n->g, 0, kParams * 4, n->ctx->stream));
    hipLaunchKernelGGL(mse_kernel, dim3(blocks_for(nv)), dim3(256), 0, n->ctx->stream, n->y, d_t,
                       n->dy, nv, n->loss);
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
