// lct_toeplitz.hip -- the ramp convolution (CUDA/Filter.cu:42-46) as a
// Toeplitz GEMM on gfx950 BF16 matrix cores with split operands.
//
//   Q[(v,r)][ch] = sum_k T[ch][k] P[(v,r)][k],   T[ch][k] = h[ch - k + ns]
//
// C (M = channels) = A (M x K, the Toeplitz taps) * B (K x N, N = every row
// of every view).  The taps are split two ways, h = h1 + h2 (bf16 each), the
// data two ways, x = x1 + x2 (~2^-17), and
//   h1 x1 + h1 x2 + h2 x1
// is accumulated in FP32 with three v_mfma_f32_16x16x32_bf16 per fragment
// (LCT_FILTER_TOEPLITZ_BF16X3).  LCT_FILTER_TOEPLITZ_BF16X6 splits both three
// ways, h = h1 + h2 + h3 and x = x1 + x2 + x3 (each exact: 3 x 8 bits cover an
// FP32 mantissa), and accumulates the six products of order <= 2,
//   h1 x1 + h1 x2 + h2 x1 + h1 x3 + h2 x2 + h3 x1,
// dropping terms of order 2^-24 relative: FP32-level accuracy at twice the
// MFMA work of the three-product form.
// The tap split's error is the same for every row, so it would bias the
// filter's DC response: h2 is rounded with error diffusion along the tap
// index (each tap's residual carried into the next one's rounding), so the
// residuals' sum over any contiguous run of taps -- every output channel's
// DC gain is such a sum -- stays within one h2 rounding step instead of
// growing with the run.  The data split's error is random per sample and
// averages out in the backprojection.  The result is written in the
// backprojector's channel-major layout Q[view][slot_ch][slot_row] (see
// lct_resident.hip) as f32, f16 or bf16.
//
// A fragments.  An A fragment (16 channels x 32 k) depends only on the
// diagonal d = m0 - k0 (a multiple of 16), so every fragment the GEMM needs
// is precomputed on the host in the MFMA register layout (lane l holds
// T[m0 + l%16][k0 + 8(l/16) + j], j = 0..7) and read as one coalesced
// 16-byte load per lane from an L2-resident table.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "lct_internal.hpp"

namespace lct {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

constexpr int kTM = 128;     // channels per workgroup tile (8 fragments of 16)
constexpr int kTN = 64;      // (view,row) columns per tile (4 fragments of 16)
constexpr int kTK = 32;      // K per MFMA step
constexpr int kBRow = 40;    // LDS row of one B column: 32 k + 8 pad (bf16)
constexpr int kRowPadT = 4;  // zero row slots before row 0 in Q (as lct_resident.hip)

struct ToeplitzArgs {
    const unsigned short *frag;  // [2][n_diag][64 lanes][8] bf16 bits (h1, h2)
    long frag_stride;            // n_diag * 64 * 8
    const float *in;             // P [(v,r)][ns] f32
    void *q;                     // Q [v][cp][rp]
    long n_cols;                 // V * nt
    int ns, nt, cp, rp, kp, n_m_tiles, n_n_wg;
    int xcd_remap;  // grid is a multiple of 8: deal consecutive tiles to one XCD
};

template <typename T> __device__ __forceinline__ T to_store(float x);
template <> __device__ __forceinline__ float to_store<float>(float x) { return x; }
template <> __device__ __forceinline__ __half to_store<__half>(float x) { return __float2half_rn(x); }
template <> __device__ __forceinline__ __hip_bfloat16 to_store<__hip_bfloat16>(float x)
{
    return __float2bfloat16(x);
}

// NP operand parts (2: three products, 3: six products).  NP = 2 keeps 4
// waves per SIMD (at most 128 registers, the accumulators in VGPRs, no
// spills); NP = 3 holds a third A fragment pair and B part, at 3 waves.
template <typename T, int NP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NP == 2 ? 4 : 3))) void
toeplitz_filter_kernel(ToeplitzArgs A)
{
    static_assert(NP == 2 || NP == 3, "two or three operand parts");
    __shared__ __attribute__((aligned(16))) unsigned short bh[2 * kTN * kBRow];
    __shared__ __attribute__((aligned(16))) unsigned short bl[2 * kTN * kBRow];
    // third data part (NP = 3; a 16-byte placeholder, never touched, for NP = 2)
    __shared__ __attribute__((aligned(16))) unsigned short b3[NP == 3 ? 2 * kTN * kBRow : 8];

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    // workgroups are dealt round-robin to the 8 XCDs; the remap gives each XCD
    // a contiguous range, so the channel tiles that read the same columns share
    // that XCD's L2 instead of fetching them once per XCD
    int b = int(blockIdx.x);
    if (A.xcd_remap) b = (b & 7) * int(gridDim.x >> 3) + (b >> 3);
    const int mt = b % A.n_m_tiles;
    const int nw = b / A.n_m_tiles;
    const int m0 = mt * kTM;
    const int li = lane & 15;
    const int lk = (lane >> 4) * 8;
    const int n_ksteps = A.kp / kTK;

    // B staging: thread -> column tid / 4, k group (tid % 4) * 8
    const int bcol = tid >> 2, bk = (tid & 3) * 8;
    auto stage_b = [&](long n0, int ks, float (&x)[8]) {
        const long col = n0 + bcol;
        const int k = ks * kTK + bk;
        if (col < A.n_cols && k < A.ns) {
            const float *src = A.in + size_t(col) * A.ns + k;
            if (k + 8 <= A.ns) {
                const float4 a = *reinterpret_cast<const float4 *>(src);
                const float4 b = *reinterpret_cast<const float4 *>(src + 4);
                x[0] = a.x; x[1] = a.y; x[2] = a.z; x[3] = a.w;
                x[4] = b.x; x[5] = b.y; x[6] = b.z; x[7] = b.w;
            } else {
#pragma unroll
                for (int j = 0; j < 8; j++) x[j] = (k + j < A.ns) ? src[j] : 0.0f;
            }
        } else {
#pragma unroll
            for (int j = 0; j < 8; j++) x[j] = 0.0f;
        }
    };
    // x = hi + lo (+ third): each part rounded to nearest even by the packed
    // hardware conversion (two values per v_cvt_pk_bf16_f32); every
    // remainder is exact in FP32
    auto commit_b = [&](int buf, const float (&x)[8]) {
        unsigned hi[4], lo[4], th[4];
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const f32x2 v = {x[2 * e], x[2 * e + 1]};
            const bf16x2 h = __builtin_convertvector(v, bf16x2);
            const f32x2 r = v - __builtin_convertvector(h, f32x2);
            const bf16x2 l = __builtin_convertvector(r, bf16x2);
            hi[e] = __builtin_bit_cast(unsigned, h);
            lo[e] = __builtin_bit_cast(unsigned, l);
            if constexpr (NP == 3) {
                const f32x2 r2 = r - __builtin_convertvector(l, f32x2);
                th[e] = __builtin_bit_cast(unsigned, __builtin_convertvector(r2, bf16x2));
            }
        }
        *reinterpret_cast<uint4 *>(bh + (buf * kTN + bcol) * kBRow + bk) = make_uint4(hi[0], hi[1], hi[2], hi[3]);
        *reinterpret_cast<uint4 *>(bl + (buf * kTN + bcol) * kBRow + bk) = make_uint4(lo[0], lo[1], lo[2], lo[3]);
        if constexpr (NP == 3)
            *reinterpret_cast<uint4 *>(b3 + (buf * kTN + bcol) * kBRow + bk) = make_uint4(th[0], th[1], th[2], th[3]);
    };

    const long n_tiles = (A.n_cols + kTN - 1) / kTN;
    for (long nt_i = nw; nt_i < n_tiles; nt_i += A.n_n_wg) {
        const long n0 = nt_i * kTN;
        f32x4 acc[2][4];
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int f = 0; f < 4; f++) acc[a][f] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};

        // A fragments of this wave's two 16-channel slices at k-step ks: diagonal
        // index (m - 32 ks + kp - 32) / 16 of the precomputed table. Those of
        // k-step ks + 1 are loaded while ks runs, so the table's L2 latency is
        // not waited on at every step.
        bf16x8 a1[2], a2[2], a3[NP == 3 ? 2 : 1];
        auto load_a = [&](int ks, bf16x8 (&p1)[2], bf16x8 (&p2)[2], bf16x8 (&p3)[NP == 3 ? 2 : 1]) {
#pragma unroll
            for (int a = 0; a < 2; a++) {
                const int m = m0 + (2 * wave + a) * 16;
                const long di = (m - ks * kTK + A.kp - kTK) / 16;
                const unsigned short *f = A.frag + (di * 64 + lane) * 8;
                p1[a] = *reinterpret_cast<const bf16x8 *>(f);
                p2[a] = *reinterpret_cast<const bf16x8 *>(f + A.frag_stride);
                if constexpr (NP == 3) p3[a] = *reinterpret_cast<const bf16x8 *>(f + 2 * A.frag_stride);
            }
        };
        load_a(0, a1, a2, a3);
        float xb[8];
        stage_b(n0, 0, xb);
        commit_b(0, xb);
        __syncthreads();
        for (int ks = 0; ks < n_ksteps; ks++) {
            const int buf = ks & 1;
            if (ks + 1 < n_ksteps) stage_b(n0, ks + 1, xb);
            // the compiler's MFMA / memory interleaving strategy for the k-step (it
            // only orders instructions): 1.76 -> 1.63 ms (NP = 2, strategy 1) and
            // 2.93 -> 2.84 ms (NP = 3, strategy 0) on the clinical scan
            __builtin_amdgcn_iglp_opt(NP == 2 ? 1 : 0);
            bf16x8 n1[2], n2[2], n3[NP == 3 ? 2 : 1];
            load_a(ks + 1 < n_ksteps ? ks + 1 : ks, n1, n2, n3);
#pragma unroll
            for (int f = 0; f < 4; f++) {
                const int c = f * 16 + li;
                const bf16x8 b_hi = *reinterpret_cast<const bf16x8 *>(bh + (buf * kTN + c) * kBRow + lk);
                const bf16x8 b_lo = *reinterpret_cast<const bf16x8 *>(bl + (buf * kTN + c) * kBRow + lk);
                bf16x8 b_th;
                if constexpr (NP == 3) b_th = *reinterpret_cast<const bf16x8 *>(b3 + (buf * kTN + c) * kBRow + lk);
#pragma unroll
                for (int a = 0; a < 2; a++) {
                    // smallest products first
                    if constexpr (NP == 3) {
                        acc[a][f] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a3[a], b_hi, acc[a][f], 0, 0, 0);
                        acc[a][f] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2[a], b_lo, acc[a][f], 0, 0, 0);
                        acc[a][f] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1[a], b_th, acc[a][f], 0, 0, 0);
                    }
                    acc[a][f] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2[a], b_hi, acc[a][f], 0, 0, 0);
                    acc[a][f] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1[a], b_lo, acc[a][f], 0, 0, 0);
                    acc[a][f] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1[a], b_hi, acc[a][f], 0, 0, 0);
                }
            }
            if (ks + 1 < n_ksteps) commit_b(buf ^ 1, xb);
#pragma unroll
            for (int a = 0; a < 2; a++) {
                a1[a] = n1[a];
                a2[a] = n2[a];
                if constexpr (NP == 3) a3[a] = n3[a];
            }
            __syncthreads();
        }
        // epilogue: D[i][n] -> Q[v][ch + 1][r + 4]
        T *q = static_cast<T *>(A.q);
#pragma unroll
        for (int f = 0; f < 4; f++) {
            const long col = n0 + f * 16 + li;
            if (col >= A.n_cols) continue;
            const long v = col / A.nt;
            const int r = int(col - v * A.nt);
            T *base = q + size_t(v) * A.cp * A.rp + r + kRowPadT;
#pragma unroll
            for (int a = 0; a < 2; a++) {
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int ch = m0 + (2 * wave + a) * 16 + (lane >> 4) * 4 + j;
                    if (ch < A.ns) base[size_t(ch + 1) * A.rp] = to_store<T>(acc[a][f][j]);
                }
            }
        }
    }
}

static unsigned short host_bf16_rn(float x)
{
    unsigned u;
    std::memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

static double host_bf16_f(unsigned short b)
{
    unsigned u = unsigned(b) << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

template <typename T, int NP>
int launch_filter_toeplitz_t(lct_ctx *ctx, int ns, int nt, const float *h_taps, const float *d_in,
                             T *d_q, long n_cols, int cp, int rp)
{
    const int mp = (ns + kTM - 1) / kTM * kTM;
    const int kp = (ns + kTK - 1) / kTK * kTK;
    // diagonals d = m - k0 from -(kp - 32) to mp - 16 in steps of 16
    const int n_diag = (mp + kp - kTK) / 16;
    const long stride = long(n_diag) * 64 * 8;
    // NP = 2: h = h1 + h2 per tap index, h2 rounded with the residual of the
    // previous taps carried along (error diffusion).  NP = 3: h = h1 + h2 + h3,
    // each the nearest bf16 of what the previous parts leave (exact)
    std::vector<unsigned short> t1(2 * size_t(ns)), t2(2 * size_t(ns)), t3(2 * size_t(ns), 0);
    double carry = 0.0;
    for (int idx = 0; idx < 2 * ns; idx++) {
        const double h = double(h_taps[idx]);
        t1[idx] = host_bf16_rn(float(h));
        if (NP == 2) {
            const double want = (h - host_bf16_f(t1[idx])) + carry;
            t2[idx] = host_bf16_rn(float(want));
            carry = want - host_bf16_f(t2[idx]);
        } else {
            const double r1 = h - host_bf16_f(t1[idx]);
            t2[idx] = host_bf16_rn(float(r1));
            t3[idx] = host_bf16_rn(float(r1 - host_bf16_f(t2[idx])));
        }
    }
    std::vector<unsigned short> frag(NP * size_t(stride));
    for (int di = 0; di < n_diag; di++) {
        const int d = di * 16 - (kp - kTK);
        for (int lane = 0; lane < 64; lane++)
            for (int j = 0; j < 8; j++) {
                // T[m + i][k0 + kk] = h[d + i - kk + ns]
                const int i = lane & 15, kk = (lane >> 4) * 8 + j;
                const int idx = d + i - kk + ns;
                const bool in = idx >= 0 && idx < 2 * ns;
                const size_t o = (size_t(di) * 64 + lane) * 8 + j;
                frag[o] = in ? t1[idx] : 0;
                frag[stride + o] = in ? t2[idx] : 0;
                if (NP == 3) frag[2 * stride + o] = in ? t3[idx] : 0;
            }
    }
    void *dt = nullptr;
    LCT_CHECK(scratch(ctx, kScratchTaps, frag.size() * 2, &dt));
    LCT_CHECK(upload(ctx, dt, frag.data(), frag.size() * 2));

    ToeplitzArgs a;
    a.frag = static_cast<const unsigned short *>(dt);
    a.frag_stride = stride;
    a.in = d_in;
    a.q = d_q;
    a.n_cols = n_cols;
    a.ns = ns;
    a.nt = nt;
    a.cp = cp;
    a.rp = rp;
    a.kp = kp;
    a.n_m_tiles = mp / kTM;
    const long n_tiles = (n_cols + kTN - 1) / kTN;
    a.n_n_wg = int(std::min<long>(n_tiles, std::max<long>(1, 4096 / a.n_m_tiles)));
    // a grid that is a multiple of 8 takes the XCD-aware order (fewer strided
    // workgroups still cover every column tile)
    while (a.n_n_wg > 8 && (a.n_m_tiles * a.n_n_wg) % 8 != 0) a.n_n_wg--;
    a.xcd_remap = (a.n_m_tiles * a.n_n_wg) % 8 == 0 ? 1 : 0;
    hipLaunchKernelGGL((toeplitz_filter_kernel<T, NP>), dim3(unsigned(a.n_m_tiles * a.n_n_wg)),
                       dim3(256), 0, ctx->stream, a);
    LCT_CHECK_HIP(hipGetLastError());
    return LCT_OK;
}

template <int NP>
static int launch_filter_toeplitz_np(lct_ctx *ctx, int storage, int ns, int nt, const float *h_taps,
                                     const float *d_in, void *d_q, long n_cols, int cp, int rp)
{
    if (storage == LCT_F32)
        return launch_filter_toeplitz_t<float, NP>(ctx, ns, nt, h_taps, d_in, static_cast<float *>(d_q),
                                                   n_cols, cp, rp);
    if (storage == LCT_F16)
        return launch_filter_toeplitz_t<__half, NP>(ctx, ns, nt, h_taps, d_in,
                                                    static_cast<__half *>(d_q), n_cols, cp, rp);
    return launch_filter_toeplitz_t<__hip_bfloat16, NP>(ctx, ns, nt, h_taps, d_in,
                                                        static_cast<__hip_bfloat16 *>(d_q), n_cols, cp, rp);
}

int launch_filter_toeplitz(lct_ctx *ctx, int algo, int storage, int ns, int nt, const float *h_taps,
                           const float *d_in, void *d_q, long n_cols, int cp, int rp)
{
    if (ns % 4 != 0)
        return set_error(LCT_ERR_UNSUPPORTED, "toeplitz filter: ns must be a multiple of 4");
    if (algo == LCT_FILTER_TOEPLITZ_BF16X6)
        return launch_filter_toeplitz_np<3>(ctx, storage, ns, nt, h_taps, d_in, d_q, n_cols, cp, rp);
    return launch_filter_toeplitz_np<2>(ctx, storage, ns, nt, h_taps, d_in, d_q, n_cols, cp, rp);
}

}  // namespace lct
