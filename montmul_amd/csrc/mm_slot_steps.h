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
    uint32_t tile_m;  // global mode: LDS tile of the radix-4 transforms, in elements (0: none)
};

// LDS of the radix-4 step in global mode: the carry scratch, then (16-byte aligned) a tile of
// tile_m elements of one field at a time.
__host__ __device__ inline uint32_t v1_tile_offset(uint32_t threads) { return (lds_scratch_bytes(threads) + 15) & ~15u; }

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
        w.tile = lds + v1_tile_offset(blockDim.x);
        w.tile_m = A.tile_m;
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

// The same steps on register-resident transforms (wraparound path, n = 512 .. 16,384): the
// vector lives in registers, exchanges go through LDS; in global mode T and x stay in HBM and
// the exchanges use an LDS tile of at most 2,048 elements (4,096 at n = 8,192 and 16,384, where
// the vector is two sub-transforms of n/2 in turn). WIDE: the workgroup has more than n / 8
// threads (see v2_square).
// TILE: the exchange tile in global mode (1: the default of v2::tile_max; else its size in
// elements, as the autotuner picks it). REG_T: T may stay in registers from the square to the
// reduction (the step kernel of one length; the mixed-length kernel has no registers to spare).
template <uint32_t LOGN, bool GLOBAL, bool BARE, bool WIDE = false, uint32_t TILE = 1, bool REG_T = false>
__device__ __forceinline__ void steps_v2(const v2::VArgs& A, uint32_t slot, uint32_t bits, unsigned char* lds) {
    constexpr uint32_t TILED = GLOBAL ? TILE : 0;
    const Slots& S = A.s;
    const uint32_t n = A.t.P.n, k = S.k;
    v2::Xbuf X;
    uint32_t* Tl;
    uint64_t* sc;
    if (GLOBAL) {
        // Transform vectors in registers between exchanges, exchanges through an LDS tile,
        // T and x in HBM.
        constexpr uint32_t tile = v2::Tiling<LOGN, TILED>::tile;
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
        if constexpr (LOGN == 14) {
            // Two sub-transforms of n/2 (v2_square_split): the slot's HBM work vector holds the
            // half waiting its turn, and coefficients are staged in HBM.
            static_assert(GLOBAL && !WIDE, "n = 2^14 runs in global mode only");
            v2::Z1* w1 = (v2::Z1*)(S.work1 + (size_t)slot * n);
            v2::Z2* w2 = (v2::Z2*)(S.work2 + (size_t)slot * n);
            uint32_t* coef = A.coef + (size_t)slot * 6 * n;
            bad |= v2::v2_square_split<LOGN - 1>(A, X, Tl, sc, x, w1, w2, coef, iw1, iw2);
            if (!BARE)
                bad |= v2::v2_redc_split<LOGN - 1>(A, X, Tl, sc, x, w1, w2, coef, S.nphat1 + (size_t)slot * n,
                                                   S.nphat2 + (size_t)slot * n, S.nhat1 + (size_t)slot * n,
                                                   S.nhat2 + (size_t)slot * n);
        } else {
            // In global mode with a two-chunk tile T stays in registers from the square to the
            // reduction (run_bare keeps writing it to HBM, where mm_read_product finds it).
            // Up to n = 4,096: at n = 8,192 the 16 registers of T spill 19 VGPRs, and T through
            // HBM measured 0.7 % faster there.
            constexpr bool KEEP = REG_T && GLOBAL && !BARE && !WIDE && v2::Tiling<LOGN, TILED>::chunks == 2 && LOGN <= 12;
            uint32_t keep[16];
            bad |= v2::v2_square<LOGN, TILED, WIDE, 0, KEEP>(A.t, X, Tl, sc, x, iw1, iw2, A.crt_inv, true, keep);
            if (!BARE)
                bad |= v2::v2_redc<LOGN, TILED, WIDE, 0, KEEP>(A, X, Tl, sc, x, S.nphat1 + (size_t)slot * n,
                                                               S.nphat2 + (size_t)slot * n, S.nhat1 + (size_t)slot * n,
                                                               S.nhat2 + (size_t)slot * n, true, keep);
        }
    }
    if (tid() == 0) {
        S.pos[slot] = pos;
        if (bad) S.err[slot] = 1;
    }
}

// ---- Slots of several lengths in one persistent launch -------------------------------------
// A group of batches, one per length, is stepped by one launch of NT-thread workgroups. The
// work queue is an array of descriptors in device memory, sorted longest first: a length class,
// a run of `count` slots in a slot list, and a step count. Every workgroup takes the next
// descriptor with an atomic add on a head counter and steps its slots, until the queue is
// empty. A register-transform length whose n / 8 threads fill a fraction of the workgroup runs
// several slots side by side, one per thread group. No reference counterpart: the reference's
// batches hold one length each (batch.rs:64-77).
constexpr uint32_t MIX_V1 = 0x100, MIX_V1_GLOBAL = 0x200, MIX_V1_WRAP = 0x400, MIX_V1_GROUPS = 0x800;
struct MixClass {
    v2::VArgs v;    // register-transform lengths: kind = log2 n (9 .. 13)
    KArgs k;        // every other length: kind = MIX_V1 | flags
    uint32_t kind;
};
struct MixItem {
    uint32_t cls, first, count, steps;  // slots list[first .. first + count), up to `steps` steps each
};

// The register-transform step's exchange tile in the mixed launch, as the TILED argument of its
// templates (v2::Tiling) and in elements: half the vector up to n = 2,048, as in the plain
// kernel's default plan there (twice the slot groups per workgroup in the same LDS), else the
// built-in tile.
__host__ __device__ constexpr uint32_t mix_tiled(uint32_t logn) { return logn <= 11 ? (1u << logn) / 2 : 1u; }
__host__ __device__ constexpr uint32_t mix_tile(uint32_t logn) {
    return logn <= 11 ? (1u << logn) / 2 : logn > 12 ? 4096u : 2048u;
}
// LDS of one slot group of the register-transform step at log2 n = LOGN in global mode: the
// exchange tile (24 bytes per element) and the carry scratch of its n / 8 threads, 16-aligned.
__host__ __device__ constexpr uint32_t mix_group_lds(uint32_t logn) {
    return ((24u * mix_tile(logn) + 8u * (2u * (((1u << logn) / 8) / 64) + 2u)) + 15u) & ~15u;
}
// Slots per workgroup of NT threads at log2 n = LOGN: a power of two, at most NT / (n / 8), and
// within the LDS of one workgroup per CU (NT = 1,024) or of two (NT = 512) on a 160 KiB CU.
__host__ __device__ constexpr uint32_t mix_groups(uint32_t nt, uint32_t logn) {
    const uint32_t t = (1u << logn) / 8;
    if (t >= nt) return 1;
    const uint32_t cap = nt >= 1024 ? 159u * 1024u : 79u * 1024u;
    uint32_t g = nt / t;
    while (g > 1 && g * mix_group_lds(logn) > cap) g /= 2;
    return g;
}

// G slots of one register-transform length in one workgroup of G n/8 threads: thread group g
// steps slot list[first + g] (g < count; the other groups join every barrier and touch no
// memory) for up to `steps` steps, each group with its own exchange tile and carry scratch.
template <uint32_t LOGN, uint32_t G>
__device__ __forceinline__ void steps_v2_group(const v2::VArgs& A, const uint32_t* list, uint32_t first, uint32_t count,
                                               uint32_t steps, unsigned char* lds) {
    constexpr uint32_t T = (1u << LOGN) / 8, TILED = mix_tiled(LOGN), tile = v2::Tiling<LOGN, TILED>::tile;
    static_assert(tile == mix_tile(LOGN), "the group's tile is the one its LDS was sized for");
    const Slots& S = A.s;
    const uint32_t n = 1u << LOGN, k = S.k;
    const uint32_t g = __builtin_amdgcn_readfirstlane(threadIdx.x / T);  // whole waves per group
    const bool on = g < count;
    const uint32_t slot = list[first + (on ? g : 0)];
    unsigned char* base = lds + g * mix_group_lds(LOGN);
    v2::Xbuf X;
    X.z2 = (v2::Z2*)base;
    X.z1 = (v2::Z1*)(base + 16 * tile);
    uint64_t* sc = (uint64_t*)(base + 24 * tile);
    uint32_t* Tl = S.lim + (size_t)slot * 3 * k;
    uint32_t* x = S.x + (size_t)slot * k;
    const uint32_t* ex = S.exp + (size_t)slot * S.ew;
    const uint32_t nb = S.nbits[slot];
    uint32_t pos = S.pos[slot];
    uint32_t bad = 0;
    for (uint32_t it = 0; it < steps; it++) {
        __syncthreads();  // x written by every thread in the previous step
        const bool act = on && pos < nb;
        const uint32_t bit = act ? (ex[pos >> 5] >> (31 - (pos & 31))) & 1 : 0;
        const F1* iw1 = bit ? A.iwc1 : A.t.iw1;
        const v2::S2* iw2 = bit ? A.iwc2 : A.t.iw2;
        bad |= v2::v2_square<LOGN, TILED, true, T>(A.t, X, Tl, sc, x, iw1, iw2, A.crt_inv, act);
        bad |= v2::v2_redc<LOGN, TILED, true, T>(A, X, Tl, sc, x, S.nphat1 + (size_t)slot * n, S.nphat2 + (size_t)slot * n,
                                                  S.nhat1 + (size_t)slot * n, S.nhat2 + (size_t)slot * n, act);
        if (act) pos++;
    }
    if (on && (threadIdx.x & (T - 1)) == 0) {
        S.pos[slot] = pos;
        if (bad) S.err[slot] = 1;
    }
}

// LDS of one slot group of the radix-4 step: its LDS-resident layout at the largest grouped
// length, n = k = 256 with 64 threads (lds_bytes_resident(256, 256, 64)).
constexpr uint32_t V1_GROUP_LDS = 24u * 256u + 12u * 256u + 8u * (2u * 1u + 2u);

// blockDim / 64 slots of one radix-4 length (wraparound, n <= 256) in one workgroup, one wave each:
// wave g steps slot list[first + g] (g < count) for up to `steps` steps in its own LDS-resident
// layout. A wave without a slot, or whose slot has finished, steps a scratch copy in its own LDS
// (w.ls, which the wraparound reduction does not use) and writes no slot's state.
__device__ __forceinline__ void steps_v1_group(const KArgs& A, const uint32_t* list, uint32_t first, uint32_t count,
                                               uint32_t steps, unsigned char* lds) {
    const Slots& S = A.s;
    const uint32_t n = A.t.n, k = S.k;
    const uint32_t g = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const bool on = g < count;
    const uint32_t slot = list[first + (on ? g : 0)];
    unsigned char* base = lds + g * V1_GROUP_LDS;
    Work w;
    w.a2 = (F2*)base;
    w.a1 = (F1*)(base + 16 * (size_t)n);
    w.lt = (uint32_t*)(base + 24 * (size_t)n);
    w.ls = w.lt + 2 * k;
    w.sc = (uint64_t*)(base + ((24 * (size_t)n + 12 * (size_t)k + 7) & ~(size_t)7));
    uint32_t* x = S.x + (size_t)slot * k;
    const uint32_t* ex = S.exp + (size_t)slot * S.ew;
    const uint32_t nb = S.nbits[slot];
    uint32_t pos = S.pos[slot];
    uint32_t bad = 0;
    for (uint32_t it = 0; it < steps; it++) {
        const bool act = on && pos < nb;
        uint32_t* xs = act ? x : w.ls;
        const uint32_t bit = act ? (ex[pos >> 5] >> (31 - (pos & 31))) & 1 : 0;
        const uint32_t c = bit ? S.base : 1;
        const uint32_t b1 = square<OneWave>(w, xs, k, c, A.t) != 0;
        const uint32_t b2 = redc<true, OneWave>(w, xs, k, A.t, A.ht, S.nphat1 + (size_t)slot * n, S.nphat2 + (size_t)slot * n,
                                                S.nhat1 + (size_t)slot * n, S.nhat2 + (size_t)slot * n);
        if (act) {
            bad |= b1 | b2;
            pos++;
        }
    }
    if (on && (threadIdx.x & 63) == 0) {
        S.pos[slot] = pos;
        if (bad) S.err[slot] = 1;
    }
}

// The mixed-length kernel of NT threads per workgroup (512 or 1,024; one object file each):
// launches it on `sm`, and sets its dynamic LDS limit to what the device allows beside its
// static LDS (written to *dyn).
template <uint32_t NT>
hipError_t mix_launch(uint32_t grid, size_t lds, hipStream_t sm, const MixClass* cls, const MixItem* items,
                      const uint32_t* list, uint32_t count, uint32_t* head);
template <uint32_t NT> hipError_t mix_lds_limit(size_t max_lds, size_t* dyn);

}  // namespace mm
