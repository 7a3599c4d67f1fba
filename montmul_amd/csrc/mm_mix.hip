// mm_mix.hip - the mixed-length step kernel: slots of several lengths in one persistent launch.
// Built once per workgroup size (MM_MIX_NT = 512 or 1,024), each into an object of its own so
// that the two large kernels compile in parallel with the rest of the library.

#include "mm_slot_steps.h"

#ifndef MM_MIX_NT
#error "MM_MIX_NT (512 or 1024) must be defined"
#endif

namespace mm {

template <uint32_t NT, uint32_t LOGN>
__device__ __forceinline__ void mix_v2(const v2::VArgs& A, const uint32_t* list, const MixItem& it, unsigned char* lds) {
    constexpr uint32_t G = mix_groups(NT, LOGN);
    if constexpr ((1u << LOGN) / 8 > NT)
        return;
    else if constexpr (G > 1)
        steps_v2_group<LOGN, G>(A, list, it.first, it.count, it.steps, lds);
    else
        steps_v2<LOGN, true, false, ((1u << LOGN) / 8 < NT), mix_tiled(LOGN)>(A, list[it.first], it.steps, lds);
}

template <uint32_t NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(MM_V2_WAVES_PER_EU))) void k_mixed(
    const MixClass* __restrict__ cls, const MixItem* __restrict__ items, const uint32_t* __restrict__ list, uint32_t count,
    uint32_t* head) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ uint32_t next;
    for (;;) {
        // The barrier first: every thread has finished the previous item (and read `next`)
        // before thread 0 takes the next descriptor. It also keeps the compiler from merging
        // this `thread 0` branch with the previous item's last one (slot 0 writes pos and err):
        // merged, the other threads would reach the read of `next` without passing the fetch,
        // and under divergent-loop lowering wave 0's other lanes would loop on a stale value
        // while lane 0 waited, without end.
        __syncthreads();
        if (threadIdx.x == 0) next = atomicAdd(head, 1u);
        __syncthreads();
        const uint32_t i = __builtin_amdgcn_readfirstlane(next);
        if (i >= count) break;
        const MixItem it = items[i];
        const MixClass& M = cls[it.cls];
        const uint32_t slot = list[it.first];
        switch (M.kind) {
            case 9: mix_v2<NT, 9>(M.v, list, it, lds); break;
            case 10: mix_v2<NT, 10>(M.v, list, it, lds); break;
            case 11: mix_v2<NT, 11>(M.v, list, it, lds); break;
            case 12: mix_v2<NT, 12>(M.v, list, it, lds); break;
            case 13: mix_v2<NT, 13>(M.v, list, it, lds); break;
            case MIX_V1: steps_v1<false, false, false>(M.k, slot, it.steps, lds); break;
            case MIX_V1 | MIX_V1_WRAP: steps_v1<false, true, false>(M.k, slot, it.steps, lds); break;
            case MIX_V1 | MIX_V1_GLOBAL: steps_v1<true, false, false>(M.k, slot, it.steps, lds); break;
            case MIX_V1 | MIX_V1_GLOBAL | MIX_V1_WRAP: steps_v1<true, true, false>(M.k, slot, it.steps, lds); break;
            case MIX_V1 | MIX_V1_WRAP | MIX_V1_GROUPS: steps_v1_group(M.k, list, it.first, it.count, it.steps, lds); break;
            default: break;
        }
    }
}

template <>
hipError_t mix_launch<MM_MIX_NT>(uint32_t grid, size_t lds, hipStream_t sm, const MixClass* cls, const MixItem* items,
                                 const uint32_t* list, uint32_t count, uint32_t* head) {
    hipLaunchKernelGGL(k_mixed<MM_MIX_NT>, dim3(grid), dim3(MM_MIX_NT), lds, sm, cls, items, list, count, head);
    return hipGetLastError();
}

template <> hipError_t mix_lds_limit<MM_MIX_NT>(size_t max_lds, size_t* dyn) {
    const void* k = (const void*)k_mixed<MM_MIX_NT>;
    hipFuncAttributes a;
    hipError_t e = hipFuncGetAttributes(&a, k);
    if (e != hipSuccess) return e;
    *dyn = max_lds > a.sharedSizeBytes ? max_lds - a.sharedSizeBytes : 0;
    if (hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)*dyn) != hipSuccess) {
        (void)hipGetLastError();
        *dyn = std::min<size_t>(*dyn, 65536);
    }
    return hipSuccess;
}

}  // namespace mm
