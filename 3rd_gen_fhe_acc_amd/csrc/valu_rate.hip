// valu_rate.hip -- issue rate of the integer VALU instructions the transforms
// and the multiply-accumulate are built from, on the GPU it runs on.
//
// Each kernel runs CH independent register chains per lane, one or two
// instruction kinds alternating, so the SIMDs are limited by issue, not by
// latency or memory; the only memory access is one in-bounds store per
// thread at the end.  The program prints, per kernel, the launch time and the
// wave instructions issued per SIMD per cycle at the clock the caller states
// (default 2.4 GHz): 0.25 is one instruction per SIMD every 4 cycles; above
// that the SIMD issued some instructions in pairs, which rocprofv3 counts as
// SQ_ACTIVE_INST_VALU2.
//
// Usage: valu_rate [clock_GHz]
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

namespace {

constexpr int CH = 8;         // independent chains per lane and kind
constexpr int NT = 256;       // threads per workgroup
constexpr int BLOCKS = 4096;  // 16 workgroups per CU on 256 CUs
constexpr int ITERS = 4000;

// One kernel per instruction mix: per chain step, instruction X on chain a[j]
// and instruction Y on chain b[j] (Y may be empty).
#define FHEACC_RATE_KERNEL(NAME, X, Y)                                                      \
    __global__ __launch_bounds__(NT) void NAME(uint64_t *out, uint32_t y, int iters) {      \
        uint32_t a[CH], b[CH];                                                              \
        for (int j = 0; j < CH; j++) {                                                      \
            a[j] = threadIdx.x * 3 + j + 0x80000000u;                                       \
            b[j] = threadIdx.x * 5 + j;                                                     \
        }                                                                                   \
        for (int it = 0; it < iters; it++) {                                                \
            _Pragma("unroll") for (int j = 0; j < CH; j++) {                                \
                X;                                                                          \
                Y;                                                                          \
            }                                                                               \
        }                                                                                   \
        uint64_t s = 0;                                                                     \
        for (int j = 0; j < CH; j++) s += a[j] + b[j];                                      \
        out[blockIdx.x * NT + threadIdx.x] = s;                                             \
    }

#define OP(INSN, R) asm volatile(INSN " %0, %0, %1" : "+v"(R[j]) : "v"(y))
FHEACC_RATE_KERNEL(k_add, OP("v_add_u32", a), )
FHEACC_RATE_KERNEL(k_add3, asm volatile("v_add3_u32 %0, %0, %1, %1" : "+v"(a[j]) : "v"(y)), )
FHEACC_RATE_KERNEL(k_lshl_add, asm volatile("v_lshl_add_u32 %0, %0, 1, %1" : "+v"(a[j]) : "v"(y)), )
FHEACC_RATE_KERNEL(k_min, OP("v_min_u32", a), )
FHEACC_RATE_KERNEL(k_mul_lo, OP("v_mul_lo_u32", a), )
FHEACC_RATE_KERNEL(k_mul_hi, OP("v_mul_hi_u32", a), )
FHEACC_RATE_KERNEL(k_add_sub, OP("v_add_u32", a), OP("v_sub_u32", b))
FHEACC_RATE_KERNEL(k_mul_lo_add, OP("v_mul_lo_u32", a), OP("v_add_u32", b))
#undef OP

// v_mad_u64_u32, as the compiler emits it for a 32x32+64-bit multiply-add
// (written in C: the compiler then needs no hazard padding between them)
__global__ __launch_bounds__(NT) void k_mad_u64(uint64_t *out, uint32_t y, int iters) {
    uint64_t acc[CH];
    for (int j = 0; j < CH; j++) acc[j] = threadIdx.x * 3 + j;
    for (int it = 0; it < iters; it++)
#pragma unroll
        for (int j = 0; j < CH; j++) acc[j] = (uint64_t)(uint32_t)acc[j] * y + acc[j];
    uint64_t s = 0;
    for (int j = 0; j < CH; j++) s += acc[j];
    out[blockIdx.x * NT + threadIdx.x] = s;
}

typedef void (*Kernel)(uint64_t *, uint32_t, int);

bool ok(hipError_t e, const char *what) {
    if (e == hipSuccess) return true;
    fprintf(stderr, "valu_rate: %s: %s\n", what, hipGetErrorString(e));
    return false;
}

}  // namespace

int main(int argc, char **argv) {
    const double ghz = argc > 1 ? atof(argv[1]) : 2.4;
    int cus = 0;
    if (!ok(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0), "device attribute")) return 1;
    uint64_t *out = nullptr;
    if (!ok(hipMalloc(&out, (size_t)BLOCKS * NT * sizeof(uint64_t)), "hipMalloc")) return 1;
    hipEvent_t e0, e1;
    if (!ok(hipEventCreate(&e0), "event") || !ok(hipEventCreate(&e1), "event")) return 1;
    const struct {
        const char *name;
        Kernel k;
        int per_step;  // instructions per chain step
    } kinds[] = {{"v_add_u32", k_add, 1},           {"v_add3_u32", k_add3, 1},
                 {"v_lshl_add_u32", k_lshl_add, 1}, {"v_min_u32", k_min, 1},
                 {"v_mul_lo_u32", k_mul_lo, 1},     {"v_mul_hi_u32", k_mul_hi, 1},
                 {"v_mad_u64_u32", k_mad_u64, 1},   {"v_add_u32 + v_sub_u32", k_add_sub, 2},
                 {"v_mul_lo_u32 + v_add_u32", k_mul_lo_add, 2}};
    const double simds = 4.0 * cus;
    for (const auto &kd : kinds) {
        float best = 1e30f;
        for (int rep = 0; rep < 3; rep++) {
            hipLaunchKernelGGL(kd.k, dim3(BLOCKS), dim3(NT), 0, 0, out, 0x12345u, ITERS);
            if (!ok(hipEventRecord(e0), "record")) return 1;
            hipLaunchKernelGGL(kd.k, dim3(BLOCKS), dim3(NT), 0, 0, out, 0x12345u, ITERS);
            if (!ok(hipEventRecord(e1), "record") || !ok(hipEventSynchronize(e1), kd.name)) return 1;
            float ms = 0;
            if (!ok(hipEventElapsedTime(&ms, e0, e1), "elapsed")) return 1;
            if (ms < best) best = ms;
        }
        const double wave_instr = (double)BLOCKS * (NT / 64) * ITERS * CH * kd.per_step;
        const double rate = wave_instr / (best * 1e-3) / (simds * ghz * 1e9);
        printf("{\"instructions\": \"%s\", \"ms\": %.4f, \"wave_instr_per_simd_cycle\": %.4f, \"clock_GHz\": %.2f}\n",
               kd.name, best, rate, ghz);
    }
    if (!ok(hipFree(out), "hipFree")) return 1;
    return 0;
}
