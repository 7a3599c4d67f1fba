// mm_probe.hip - issue-rate probe for the integer instructions the step kernel is built from.
//
// The step kernel is bound by VALU issue, not by HBM or by a multiplier count (no MFMA applies:
// the arithmetic is exact modular integer work). Its roofline therefore needs the issue rate of
// gfx950 for the instructions it uses, measured on the device rather than taken from a data
// sheet. Each probe kernel runs, per thread, 8 independent dependency chains of one operation
// for `iters` iterations over a grid that fills every SIMD with 8 waves, and folds the chains into
// one word per thread (so nothing is optimized away). The host times the launch with HIP
// events; a rocprofv3 counter run of the same command gives SQ_INSTS_VALU and the clock.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/montmul_hip.h"

namespace {

thread_local std::string g_err;

enum ProbeOp { OP_MAD_U64_U32 = 0, OP_ADD_U32 = 1, OP_MUL_HI_U32 = 2, OP_ADD_U64 = 3, OP_MAD_AND_ADD = 4 };

// One instruction per chain and iteration, written as inline assembly so that the compiler can
// neither fuse nor strength-reduce the chains (arithmetic only: no memory access in the asm).
template <int OP>
__global__ __launch_bounds__(256) void k_probe_issue(uint32_t* out, uint32_t iters, uint32_t seed) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t a = (t ^ seed) | 1u, b = a * 2654435761u + seed;
    uint64_t c[8];
    uint32_t d[8];  // the 32-bit chains (OP_ADD_U32, OP_MUL_HI_U32, OP_MAD_AND_ADD)
#pragma unroll
    for (int j = 0; j < 8; j++) {
        c[j] = (uint64_t)(b + 77u * (uint32_t)j) << 7 | (uint32_t)j;
        d[j] = b ^ (uint32_t)j;
    }
    for (uint32_t i = 0; i < iters; i++) {
#pragma unroll
        for (int j = 0; j < 8; j++) {
            if constexpr (OP == OP_MAD_U64_U32) {
                uint64_t carry;  // the instruction's carry-out pair, unused
                asm volatile("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(c[j]), "=s"(carry) : "v"(a), "v"(b));
            } else if constexpr (OP == OP_ADD_U32) {
                asm volatile("v_add_u32 %0, %0, %1" : "+v"(d[j]) : "v"(a));
            } else if constexpr (OP == OP_MUL_HI_U32) {
                asm volatile("v_mul_hi_u32 %0, %0, %1" : "+v"(d[j]) : "v"(a));
            } else if constexpr (OP == OP_ADD_U64) {
                const uint64_t k = (uint64_t)a << 3 | (uint32_t)j;
                asm volatile("v_lshl_add_u64 %0, %0, 0, %1" : "+v"(c[j]) : "v"(k));
            } else {
                // A multiply-add and an independent 32-bit add in turn, as the step's code mixes
                // them (back-to-back multiply-adds take a wait state between them on gfx950).
                uint64_t carry;
                asm volatile("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(c[j]), "=s"(carry) : "v"(a), "v"(b));
                asm volatile("v_add_u32 %0, %0, %1" : "+v"(d[j]) : "v"(a));
            }
        }
    }
    uint64_t r = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) r ^= c[j] + d[j];
    out[t] = (uint32_t)r ^ (uint32_t)(r >> 32);
}

}  // namespace

extern "C" {

// Runs probe `op` (0: v_mad_u64_u32, 1: v_add_u32, 2: v_mul_hi_u32, 3: 64-bit add
// v_lshl_add_u64, 4: v_mad_u64_u32 and v_add_u32 in turn) on device
// `device` with `blocks` workgroups of 256 threads and `iters` iterations of 8 chained operations
// per thread. Writes the launch's device time (HIP events) in ms and a checksum of the results.
// Returns 0, or nonzero with mm_probe_last_error(). No reference counterpart: measurement.
int mm_probe_issue(int device, int op, uint32_t blocks, uint32_t iters, double* ms, uint64_t* checksum) {
    if (op < 0 || op > 4 || blocks == 0) {
        g_err = "probe: op must be 0..4 and blocks > 0";
        return 1;
    }
    auto chk = [](hipError_t e, const char* what) {
        if (e != hipSuccess) g_err = std::string(what) + ": " + hipGetErrorString(e);
        return e != hipSuccess;
    };
    if (chk(hipSetDevice(device), "hipSetDevice")) return 1;
    const size_t n = (size_t)blocks * 256;
    uint32_t* out = nullptr;
    if (chk(hipMalloc((void**)&out, n * 4), "hipMalloc")) return 1;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 1;
    if (!chk(hipEventCreate(&e0), "hipEventCreate") && !chk(hipEventCreate(&e1), "hipEventCreate")) {
        void (*kern)(uint32_t*, uint32_t, uint32_t) = op == 0   ? k_probe_issue<0>
                                                     : op == 1 ? k_probe_issue<1>
                                                     : op == 2 ? k_probe_issue<2>
                                                     : op == 3 ? k_probe_issue<3>
                                                               : k_probe_issue<4>;
        hipEventRecord(e0, 0);
        hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), 0, 0, out, iters, 12345u);
        hipEventRecord(e1, 0);
        if (!chk(hipGetLastError(), "probe launch") && !chk(hipEventSynchronize(e1), "probe")) {
            float f = 0;
            hipEventElapsedTime(&f, e0, e1);
            *ms = f;
            std::vector<uint32_t> h(n);
            if (!chk(hipMemcpy(h.data(), out, n * 4, hipMemcpyDeviceToHost), "hipMemcpy")) {
                uint64_t s = 0;
                for (uint32_t v : h) s = s * 1000003u + v;
                *checksum = s;
                rc = 0;
            }
        }
    }
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    hipFree(out);
    return rc;
}

const char* mm_probe_last_error(void) { return g_err.c_str(); }

}  // extern "C"
