// issue_rates -- issue cost of the instruction classes the backprojector uses,
// measured on the GPU it runs on (gfx950).
//
//     listctrecons_amd/bin/issue_rates
//
// Each kernel runs a fixed loop of one instruction in 8 independent chains
// (or 8 FMAs plus n scalar instructions) on 256 CUs x 4 SIMDs x 8 waves, so
// issue, not latency, sets its time.  Printed: cycles per wave64
// instruction per SIMD at 2.4 GHz (the engine clock the bench's VALU peak
// uses).  The LDS rows print CU-cycles per wave-level ds_read2_b32 with four
// reads in flight.  Every loop is bounded, every LDS address is masked into
// the workgroup's buffer, and the instructions written in asm are ALU
// instructions only.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

#define CHECK(x)                                                                    \
    do {                                                                            \
        hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess) {                                                     \
            std::fprintf(stderr, "HIP error %s at line %d\n", hipGetErrorString(e_), \
                         __LINE__);                                                 \
            std::exit(1);                                                           \
        }                                                                           \
    } while (0)

constexpr int kIters = 32768;

// %0..%7 are the chains, %8 and %9 the operands; each line of the macro
// names its own chain
#define OP8(OP)                                                                              \
    asm volatile(OP(0) "\n" OP(1) "\n" OP(2) "\n" OP(3) "\n" OP(4) "\n" OP(5) "\n" OP(6) "\n" OP(7) \
                 : "+v"(x0), "+v"(x1), "+v"(x2), "+v"(x3), "+v"(x4), "+v"(x5), "+v"(x6), "+v"(x7) \
                 : "v"(a), "v"(b));

#define FMA_(k) "v_fma_f32 %" #k ", %8, %9, %" #k
#define ADD_(k) "v_add_f32 %" #k ", %8, %" #k
#define SUB_(k) "v_sub_f32 %" #k ", %8, %" #k
#define MUL_(k) "v_mul_f32 %" #k ", %8, %" #k
#define ADDU_(k) "v_add_u32 %" #k ", %8, %" #k
#define AND_(k) "v_and_b32 %" #k ", %8, %" #k
#define CVTI_(k) "v_cvt_i32_f32 %" #k ", %" #k
#define FRACT_(k) "v_fract_f32 %" #k ", %" #k
#define LSHLADD_(k) "v_lshl_add_u32 %" #k ", %" #k ", 2, %8"
#define CVTH_(k) "v_cvt_f32_f16 %" #k ", %" #k
#define MAX_(k) "v_max_f32 %" #k ", %8, %" #k
#define MED3_(k) "v_med3_f32 %" #k ", %" #k ", %8, %9"
#define FMAMIX_(k) "v_fma_mix_f32 %" #k ", %8, %9, %" #k " op_sel_hi:[0,1,0]"
#define PKFMA16_(k) "v_pk_fma_f16 %" #k ", %8, %9, %" #k
#define MOV_(k) "v_mov_b32 %" #k ", %8"
#define XOR_(k) "v_xor_b32 %" #k ", %8, %" #k
#define SHL_(k) "v_lshlrev_b32 %" #k ", 2, %" #k
#define MIN_(k) "v_min_f32 %" #k ", %8, %" #k
#define FLOOR_(k) "v_floor_f32 %" #k ", %" #k
#define CVTFI_(k) "v_cvt_f32_i32 %" #k ", %" #k
#define MADU24_(k) "v_mad_u32_u24 %" #k ", %" #k ", %8, %9"
#define DOT2_(k) "v_dot2_f32_f16 %" #k ", %8, %9, %" #k
#define FMA16_(k) "v_fma_f16 %" #k ", %8, %9, %" #k
#define RCP_(k) "v_rcp_f32 %" #k ", %" #k
#define SQRT_(k) "v_sqrt_f32 %" #k ", %" #k

#define KERNEL(NAME, OP)                                                                     \
    __global__ void NAME(float *out, float a_, float b_)                                     \
    {                                                                                        \
        float a = a_ + float(threadIdx.x) * 1e-9f, b = b_;                                   \
        float x0 = float(threadIdx.x), x1 = x0 + 1, x2 = x0 + 2, x3 = x0 + 3, x4 = x0 + 4,   \
              x5 = x0 + 5, x6 = x0 + 6, x7 = x0 + 7;                                         \
        for (int i = 0; i < kIters; i++) {                                                   \
            OP8(OP)                                                                          \
        }                                                                                    \
        out[blockIdx.x * blockDim.x + threadIdx.x] = x0 + x1 + x2 + x3 + x4 + x5 + x6 + x7;  \
    }

KERNEL(k_fma, FMA_)
KERNEL(k_add, ADD_)
KERNEL(k_sub, SUB_)
KERNEL(k_mul, MUL_)
KERNEL(k_addu, ADDU_)
KERNEL(k_and, AND_)
KERNEL(k_cvti, CVTI_)
KERNEL(k_fract, FRACT_)
KERNEL(k_lshladd, LSHLADD_)
KERNEL(k_cvth, CVTH_)
KERNEL(k_max, MAX_)
KERNEL(k_med3, MED3_)
KERNEL(k_fmamix, FMAMIX_)
KERNEL(k_pkfma16, PKFMA16_)
KERNEL(k_mov, MOV_)
KERNEL(k_xor, XOR_)
KERNEL(k_shl, SHL_)
KERNEL(k_min, MIN_)
KERNEL(k_floor, FLOOR_)
KERNEL(k_cvtfi, CVTFI_)
KERNEL(k_madu24, MADU24_)
KERNEL(k_dot2, DOT2_)
KERNEL(k_fma16, FMA16_)
KERNEL(k_rcp, RCP_)
KERNEL(k_sqrt, SQRT_)

// 8 FMAs and N scalar adds per iteration, from the same wave
template <int N>
__global__ void k_fma_salu(float *out, float a_, float b_)
{
    float a = a_ + float(threadIdx.x) * 1e-9f, b = b_;
    float x0 = float(threadIdx.x), x1 = x0 + 1, x2 = x0 + 2, x3 = x0 + 3, x4 = x0 + 4, x5 = x0 + 5,
          x6 = x0 + 6, x7 = x0 + 7;
    unsigned s0 = 0, s1 = 0;
    for (int i = 0; i < kIters; i++) {
        OP8(FMA_)
        if constexpr (N == 2)
            asm volatile("s_add_u32 %0, %0, 1\n s_add_u32 %1, %1, 1" : "+s"(s0), "+s"(s1) : : "scc");
        if constexpr (N == 4)
            asm volatile("s_add_u32 %0, %0, 1\n s_add_u32 %1, %1, 1\n s_add_u32 %0, %0, 1\n s_add_u32 %1, %1, 1"
                         : "+s"(s0), "+s"(s1) : : "scc");
        if constexpr (N == 8)
            asm volatile("s_add_u32 %0, %0, 1\n s_add_u32 %1, %1, 1\n s_add_u32 %0, %0, 1\n s_add_u32 %1, %1, 1\n"
                         "s_add_u32 %0, %0, 1\n s_add_u32 %1, %1, 1\n s_add_u32 %0, %0, 1\n s_add_u32 %1, %1, 1"
                         : "+s"(s0), "+s"(s1) : : "scc");
    }
    out[blockIdx.x * blockDim.x + threadIdx.x] = x0 + x1 + x2 + x3 + x4 + x5 + x6 + x7 + float(s0 + s1);
}

// Four ds_read2_b32 per iteration, lanes 73 floats apart (the backprojector's
// odd pitch), with 16 FMAs beside them (WITH_FMA) or none.
template <bool WITH_FMA>
__global__ void k_lds(float *out, float a_, float b_)
{
    __shared__ float buf[4096 + 128];
    for (int i = int(threadIdx.x); i < 4096 + 128; i += int(blockDim.x)) buf[i] = float(i) * 0.5f;
    __syncthreads();
    float a = a_ + float(threadIdx.x) * 1e-9f, b = b_;
    float x0 = float(threadIdx.x), x1 = x0 + 1, x2 = x0 + 2, x3 = x0 + 3, x4 = x0 + 4, x5 = x0 + 5,
          x6 = x0 + 6, x7 = x0 + 7;
    const unsigned base = unsigned((int(threadIdx.x & 63) * 73) & 4095);
    float acc = 0.0f;
    for (int i = 0; i < kIters; i++) {
        // both indices stay below 4096, and +74 below 4096 + 128
        const unsigned ia = (base + unsigned(i & 31) * 64u) & 4095u;
        const unsigned ic = (ia + 1024u) & 4095u;
        const float *pa = buf + ia, *pc = buf + ic;
        const float a00 = pa[0], a01 = pa[1], a10 = pa[73], a11 = pa[74];
        const float c00 = pc[0], c01 = pc[1], c10 = pc[73], c11 = pc[74];
        if constexpr (WITH_FMA) {
            OP8(FMA_)
            OP8(FMA_)
        }
        acc += (a00 + a01) * (a10 + a11) + (c00 + c01) * (c10 + c11);
    }
    out[blockIdx.x * blockDim.x + threadIdx.x] = x0 + x1 + x2 + x3 + x4 + x5 + x6 + x7 + acc;
}

using Kern = void (*)(float *, float, float);

static float time_ms(Kern k, int waves_per_simd, float *out, hipEvent_t e0, hipEvent_t e1)
{
    const dim3 grid(256 * unsigned(waves_per_simd)), block(256);
    hipLaunchKernelGGL(k, grid, block, 0, 0, out, 1.0001f, 0.5f);
    CHECK(hipGetLastError());
    CHECK(hipDeviceSynchronize());
    CHECK(hipEventRecord(e0));
    for (int r = 0; r < 5; r++) hipLaunchKernelGGL(k, grid, block, 0, 0, out, 1.0001f, 0.5f);
    CHECK(hipEventRecord(e1));
    CHECK(hipEventSynchronize(e1));
    float ms = 0.0f;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    return ms / 5.0f;
}

int main()
{
    int n_cu = 0;
    CHECK(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, 0));
    float *out = nullptr;
    CHECK(hipMalloc(&out, size_t(256) * 8 * 256 * sizeof(float)));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    const double clk = 2.4e9;
    const int wps = 8;  // waves per SIMD (grids sized for 256 CUs)
    std::printf("issue cost on %d CUs, %d waves per SIMD, at %.1f GHz\n", n_cu, wps, clk / 1e9);
    struct Row {
        const char *name;
        Kern k;
    };
    const Row rows[] = {
        {"v_fma_f32", k_fma},         {"v_add_f32", k_add},         {"v_sub_f32", k_sub},
        {"v_mul_f32", k_mul},         {"v_add_u32", k_addu},        {"v_and_b32", k_and},
        {"v_cvt_i32_f32", k_cvti},    {"v_fract_f32", k_fract},     {"v_lshl_add_u32", k_lshladd},
        {"v_cvt_f32_f16", k_cvth},    {"v_max_f32", k_max},         {"v_med3_f32", k_med3},
        {"v_fma_mix_f32", k_fmamix},  {"v_pk_fma_f16", k_pkfma16},  {"v_mov_b32", k_mov},
        {"v_xor_b32", k_xor},         {"v_lshlrev_b32", k_shl},     {"v_min_f32", k_min},
        {"v_floor_f32", k_floor},     {"v_cvt_f32_i32", k_cvtfi},   {"v_mad_u32_u24", k_madu24},
        {"v_dot2_f32_f16", k_dot2},   {"v_fma_f16", k_fma16},       {"v_rcp_f32", k_rcp},
        {"v_sqrt_f32", k_sqrt},
    };
    for (const Row &r : rows) {
        const float ms = time_ms(r.k, wps, out, e0, e1);
        const double per = ms * 1e-3 * clk / (double(kIters) * 8 * wps);
        std::printf("%-22s %7.3f ms  %5.2f cycles per wave-instruction per SIMD\n", r.name, ms, per);
        std::fflush(stdout);
    }
    const Row salu[] = {{"8 v_fma_f32 + 0 SALU", k_fma},
                        {"8 v_fma_f32 + 2 SALU", k_fma_salu<2>},
                        {"8 v_fma_f32 + 4 SALU", k_fma_salu<4>},
                        {"8 v_fma_f32 + 8 SALU", k_fma_salu<8>}};
    for (const Row &r : salu) {
        const float ms = time_ms(r.k, wps, out, e0, e1);
        const double per = ms * 1e-3 * clk / (double(kIters) * wps);
        std::printf("%-22s %7.3f ms  %5.2f cycles per iteration per SIMD\n", r.name, ms, per);
        std::fflush(stdout);
    }
    const Row lds[] = {{"4 ds_read2_b32", k_lds<false>}, {"4 ds_read2_b32 + 16 fma", k_lds<true>}};
    for (const Row &r : lds) {
        const float ms = time_ms(r.k, wps, out, e0, e1);
        const double per = ms * 1e-3 * clk / (double(kIters) * 4 * 4 * wps);
        std::printf("%-22s %7.3f ms  %5.2f CU-cycles per wave ds_read2_b32 (%.0f B/clk/CU)\n", r.name, ms, per,
                    64.0 * 8.0 / per);
        std::fflush(stdout);
    }
    CHECK(hipFree(out));
    return 0;
}
