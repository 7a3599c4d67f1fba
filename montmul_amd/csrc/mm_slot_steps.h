// mm_slot_steps.h - the steps of one slot as device functions, shared by the per-length step
// kernels (mm_hip.hip) and the mixed-length kernel (mm_mix.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "mm_step.h"
#include "mm_v2_step.h"

namespace mm {

struct KArgs {
    Tab t;
    HalfTab ht;
    Slots s;
};

// The register-transform kernels are held to 128 VGPRs: four waves per SIMD.
#ifndef MM_V2_WAVES_PER_EU
#define MM_V2_WAVES_PER_EU 4
#endif

// Up to `bits` steps on one slot (reference: `steps`, kernels/src/lib.rs:883-927). BARE runs
// the squaring alone and leaves x unchanged (lib.rs:949-961).
template <bool GLOBAL, bool WRAP, bool BARE>
__device__ __forceinline__ void steps_v1(const KArgs& A, uint32_t slot, uint32_t bits, unsigned char* lds) {
    const Slots& S = A.s;
    const uint32_t n = A.t.n, k = S.k;
    Work w;
    if (GLOBAL) {
        w.a1 = S.work1 + (size_t)slot * n;
        w.a2 = S.work2 + (size_t)slot * n;
        w.lt = S.lim + (size_t)slot * 3 * k;
        w.ls = w.lt + 2 * k;
        w.sc = (uint64_t*)lds;
    } else {
        w.a2 = (F2*)lds;
        w.a1 = (F1*)(lds + 16 * (size_t)n);
        w.lt = (uint32_t*)(lds + 24 * (size_t)n);
        w.ls = w.lt + 2 * k;
        w.sc = (uint64_t*)(lds + ((24 * (size_t)n + 12 * (size_t)k + 7) & ~(size_t)7));
    }
    uint32_t* x = S.x + (size_t)slot * k;
    const uint32_t* ex = S.exp + (size_t)slot * S.ew;
    const uint32_t nb = S.nbits[slot];
    uint32_t pos = S.pos[slot];
    uint32_t bad = 0;
    for (uint32_t it = 0; it < bits && pos < nb; it++, pos++) {
        const uint32_t bit = (ex[pos >> 5] >> (31 - (pos & 31))) & 1;
        const uint32_t c = bit ? S.base : 1;
        bad |= square(w, x, k, c, A.t) != 0;
        if (!BARE)
            bad |= redc<WRAP>(w, x, k, A.t, A.ht, S.nphat1 + (size_t)slot * n, S.nphat2 + (size_t)slot * n,
                              S.nhat1 + (size_t)slot * n, S.nhat2 + (size_t)slot * n);
    }
    if (tid() == 0) {
        S.pos[slot] = pos;
        if (bad) S.err[slot] = 1;
    }
}

// The same steps on register-resident transforms (wraparound path, n = 512 .. 8,192): the
// vector lives in registers, exchanges go through LDS; in global mode T and x stay in HBM and
// the exchanges use an LDS tile of at most 2,048 elements (4,096 at n = 8,192). WIDE: the
// workgroup has more than n / 8 threads (see v2_square).
template <uint32_t LOGN, bool GLOBAL, bool BARE, bool WIDE = false>
__device__ __forceinline__ void steps_v2(const v2::VArgs& A, uint32_t slot, uint32_t bits, unsigned char* lds) {
    const Slots& S = A.s;
    const uint32_t n = A.t.P.n, k = S.k;
    v2::Xbuf X;
    uint32_t* Tl;
    uint64_t* sc;
    if (GLOBAL) {
        // Transform vectors in registers between exchanges, exchanges through an LDS tile,
        // T and x in HBM.
        constexpr uint32_t tile = v2::Tiling<LOGN, true>::tile;
        X.z2 = (v2::Z2*)lds;
        X.z1 = (v2::Z1*)(lds + 16 * tile);
        Tl = S.lim + (size_t)slot * 3 * k;
        sc = (uint64_t*)(lds + 24 * tile);
    } else {
        X.z2 = (v2::Z2*)lds;
        X.z1 = (v2::Z1*)(lds + 16 * (size_t)n);
        Tl = (uint32_t*)(lds + 24 * (size_t)n);
        sc = (uint64_t*)(lds + 24 * (size_t)n + 8 * (size_t)k);
    }
    uint32_t* x = S.x + (size_t)slot * k;
    const uint32_t* ex = S.exp + (size_t)slot * S.ew;
    const uint32_t nb = S.nbits[slot];
    uint32_t pos = S.pos[slot];
    uint32_t bad = 0;
    for (uint32_t it = 0; it < bits && pos < nb; it++, pos++) {
        __syncthreads();  // x written by every thread in the previous step
        const uint32_t bit = (ex[pos >> 5] >> (31 - (pos & 31))) & 1;
        const F1* iw1 = bit ? A.iwc1 : A.t.iw1;
        const v2::S2* iw2 = bit ? A.iwc2 : A.t.iw2;
        bad |= v2::v2_square<LOGN, GLOBAL, WIDE>(A.t, X, Tl, sc, x, iw1, iw2, A.crt_inv);
        if (!BARE)
            bad |= v2::v2_redc<LOGN, GLOBAL, WIDE>(A, X, Tl, sc, x, S.nphat1 + (size_t)slot * n, S.nphat2 + (size_t)slot * n,
                                                   S.nhat1 + (size_t)slot * n, S.nhat2 + (size_t)slot * n);
    }
    if (tid() == 0) {
        S.pos[slot] = pos;
        if (bad) S.err[slot] = 1;
    }
}

// ---- Slots of several lengths in one persistent launch -------------------------------------
// A group of batches, one per length, is stepped by one launch of NT-thread workgroups. The
// work queue is an array of per-slot descriptors (length class, slot) in device memory, sorted
// longest first; every workgroup takes the next descriptor with an atomic add on a head counter
// and runs that slot's steps, until the queue is empty. No reference counterpart: the
// reference's batches hold one length each (batch.rs:64-77).
constexpr uint32_t MIX_V1 = 0x100, MIX_V1_GLOBAL = 0x200, MIX_V1_WRAP = 0x400;
struct MixClass {
    v2::VArgs v;    // register-transform lengths: kind = log2 n (9 .. 13)
    KArgs k;        // every other length: kind = MIX_V1 | flags
    uint32_t kind;
};
struct MixItem {
    uint32_t cls, slot;
};

// The mixed-length kernel of NT threads per workgroup (512 or 1,024; one object file each):
// launches it on `sm`, and sets its dynamic LDS limit to what the device allows beside its
// static LDS (written to *dyn).
template <uint32_t NT>
hipError_t mix_launch(uint32_t grid, size_t lds, hipStream_t sm, const MixClass* cls, const MixItem* items, uint32_t count,
                      uint32_t* head, uint32_t bits);
template <uint32_t NT> hipError_t mix_lds_limit(size_t max_lds, size_t* dyn);

}  // namespace mm
