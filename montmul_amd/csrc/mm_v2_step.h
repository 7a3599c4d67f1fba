// mm_v2_step.h - the fused Montgomery step on register-resident transforms (wraparound path).
#pragma once

#include "mm_v2.h"

namespace mm {
namespace v2 {

// Device tables of one length for the register transforms.
struct VTab {
    Plan P;
    const F1 *w1, *iw1;    // zeta^j, zeta^-j / n (canonical)
    const S2 *w2, *iw2;    // the same in GF(M61^2), split at bit 31
    const F1 *tw1, *itw1;  // pass twiddles (layout of Plan)
    const S2 *tw2, *itw2;
    const F1 *sw1, *isw1;  // n = 2^14 only: omega^j and omega^-j, j < n / 2 (the split stage)
    const S2 *sw2, *isw2;
};

struct VArgs {
    VTab t;     // length n
    VTab ht;    // length h = n / 2 (weights unused); at n = 2^14 also the two sub-transforms
    Slots s;
    u32 crt_inv;
    const F1* iwc1;  // zeta^-j c / n for the batch's base c: a 1 bit's factor folded into the
    const S2* iwc2;  // unweighting (c IDFT(X) = IDFT(c X))
    u32* coef;       // n = 2^14 only: coefficient staging, 6 n words per slot (3 per coefficient)
};

// LDS bytes of the LDS-resident layout: exchange buffer (24 n), T (8 k), carry scratch.
__host__ __device__ inline uint64_t v2_lds_resident(uint32_t n, uint32_t k, uint32_t threads) {
    return 24ull * n + 8ull * k + 8ull * (2 * (threads / 64) + 2);
}

// WIDE: the workgroup may have more threads than the vector's n / 8; threads t >= n / 8 join
// every barrier and touch no memory (the mixed-length launch runs every length in workgroups
// of the largest length's size). Without WIDE, blockDim = n / 8 and no thread is idle.
// GT > 0: the workgroup holds several slots, one per group of GT = n / 8 threads; t is the
// index within the group, and a group whose `on` is false (no slot) touches no memory.
#define MM_ACT(t) (on && (!WIDE || (t) < T))

// One step's squaring: T = c x^2 into Tl (2k limbs); returns nonzero on a carry out.
// iw1, iw2: the unweighting, zeta^-j / n or, on a 1 bit, zeta^-j c / n.
// KEEP (exchange tile of n/2 elements, one slot per workgroup): T stays in registers instead,
// thread t's limbs 8t .. 8t+7 in keep[0..7] and n + 8t .. n + 8t+7 in keep[8..15], for v2_redc.
template <u32 LOGN, u32 TILED, bool WIDE = false, u32 GT = 0, bool KEEP = false>
__device__ __forceinline__ u32 v2_square(const VTab& V, const Xbuf& X, u32* Tl, u64* sc, const u32* x, const F1* iw1,
                                         const S2* iw2, u32 inv, bool on = true, u32* keep = nullptr) {
    typedef CPlan<LOGN> P;
    u32 t = slot_tid<GT>();  // refreshed per phase: no address held across phases
    const u32 T = P::T, n = P::n;
    constexpr u32 NT = WIDE ? P::T : 0;
    const u32 r0 = P::radix(0), ls0 = P::lstride(0);
    Z1 v1[8];
    Z2 v2[8];
    if (MM_ACT(t)) {
#pragma unroll
        for (u32 e = 0; e < 8; e++) {
            u32 pos = slot_pos(t, e, r0, ls0, T);
            u32 xv = x[pos];
            v1[e] = Op1::real_mul(xv, V.w1[pos]);
            v2[e] = Op2::real_mul(xv, V.w2[pos]);
            MM_GROUP_FENCE(e);
        }
    }
    dif_regs<LOGN, TILED, GT>(t, v1, v2, V.tw1, V.tw2, X, on);
#pragma unroll
    for (u32 e = 0; e < 8; e++) {
        v1[e] = Op1::sqr(v1[e]);
        v2[e] = Op2::sqr(v2[e]);
    }
    dit_regs<LOGN, TILED, GT, false>(t, v1, v2, V.itw1, V.itw2, X, on);  // after the DIF's own exchanges
    t = slot_tid<GT>();
    typedef Tiling<LOGN, TILED> G;
    u32 l[16];
    u64 top;
    static_assert(!KEEP || (G::chunks == 2 && !WIDE && GT == 0), "T in registers: two chunks, one slot per workgroup");
    if constexpr (G::chunks == 1) {
        const Cbuf C = cbuf(X, n);
        __syncthreads();
        if (MM_ACT(t)) {
#pragma unroll
            for (u32 e = 0; e < 8; e++) {
                u32 pos = slot_pos(t, e, r0, ls0, T);
                Z1 a = Op1::mul(v1[e], iw1[pos]);
                Z2 b = Op2::mul(v2[e], iw2[pos]);
                cput(C, pos, crt_lazy(a.re, b.re, inv));
                cput(C, pos + n, crt_lazy(a.im, b.im, inv));
                MM_GROUP_FENCE(e);
            }
        }
        __syncthreads();
        t = slot_tid<GT>();
        top = carry_regs<16, NT>(t, [&](int i) { return MM_ACT(t) ? cget(C, 16 * t + i) : (u128)0; }, l, sc, 0);
    } else {
        // The tile holds 2 * tile = n coefficients: the real parts (coefficients 0 .. n-1) go
        // through it and are carried first, then the imaginary parts (n .. 2n-1) with the
        // carry out of the low half; only the imaginary residues are held in between.
        static_assert(2 * G::tile == n, "two coefficient chunks");
        const Cbuf C = cbuf(X, G::tile);
        u32 im1[8];
        u64 im2[8];
        __syncthreads();
        if (MM_ACT(t)) {
#pragma unroll
            for (u32 e = 0; e < 8; e++) {
                u32 pos = slot_pos(t, e, r0, ls0, T);
                Z1 a = Op1::mul(v1[e], iw1[pos]);
                Z2 b = Op2::mul(v2[e], iw2[pos]);
                cput(C, pos, crt_lazy(a.re, b.re, inv));
                im1[e] = a.im;
                im2[e] = b.im;
                pin(im1[e]);
                pin(im2[e]);
                MM_GROUP_FENCE(e);
            }
        }
        __syncthreads();
        u32 lo[8];
        t = slot_tid<GT>();
        const u64 mid = carry_regs<8, NT>(t, [&](int i) { return MM_ACT(t) ? cget(C, 8 * t + i) : (u128)0; }, lo, sc, 0);
        if (MM_ACT(t)) {
#pragma unroll
            for (u32 i = 0; i < 8; i++) {
                if constexpr (KEEP)
                    keep[i] = lo[i];
                else
                    Tl[8 * t + i] = lo[i];
            }
#pragma unroll
            for (u32 e = 0; e < 8; e++) cput(C, slot_pos(t, e, r0, ls0, T), crt_lazy(im1[e], im2[e], inv));
        }
        __syncthreads();
        u32 hi[8];
        t = slot_tid<GT>();
        const u64 top = carry_regs<8, NT>(t, [&](int i) { return MM_ACT(t) ? cget(C, 8 * t + i) : (u128)0; }, hi, sc, mid);
        if (MM_ACT(t)) {
#pragma unroll
            for (u32 i = 0; i < 8; i++) {
                if constexpr (KEEP)
                    keep[8 + i] = hi[i];
                else
                    Tl[n + 8 * t + i] = hi[i];
            }
        }
        return top != 0;
    }
    if (MM_ACT(t)) {
#pragma unroll
        for (u32 i = 0; i < 16; i++) Tl[16 * t + i] = l[i];
    }
    return top != 0;
}

// Reduction: x = T R^-1 (lazy) from T in Tl, through m = (T mod R) N' mod R and the cyclic
// product m N mod (R - 1) (montmul/src/cpu.rs:452-484, 297-328). Returns nonzero on a carry
// that must vanish but does not.
// KEEP: T comes from v2_square's registers (keep[16]); its low half reaches the transform's
// positions through the exchange buffer instead of HBM.
template <u32 LOGN, u32 TILED, bool WIDE = false, u32 GT = 0, bool KEEP = false>
__device__ __forceinline__ u32 v2_redc(const VArgs& A, const Xbuf& X, u32* Tl, u64* sc, u32* x, const F1* np1, const F2* np2,
                                       const F1* ab1, const F2* ab2, bool on = true, const u32* keep = nullptr) {
    const VTab& V = A.t;
    const VTab& H = A.ht;
    typedef CPlan<LOGN> P;
    typedef CPlan<LOGN - 1> PH;
    u32 t = slot_tid<GT>();  // refreshed per phase: no address held across phases
    const u32 T = P::T, n = P::n, k = n, h = n / 2, Th = PH::T;
    constexpr u32 NT = WIDE ? P::T : 0;
    const u32 r0 = P::radix(0), ls0 = P::lstride(0);
    const u32 inv = A.crt_inv;
    // Staging for k = n coefficients: the exchange buffer (one tile) holds 2 * tile of them.
    static_assert(2 * Tiling<LOGN, TILED>::tile >= (1u << LOGN), "coefficients fit the tile");
    const Cbuf C = cbuf(X, Tiling<LOGN, TILED>::tile);
    Z1 v1[8];
    Z2 v2[8];
    const u32* Tsrc = Tl;
    if constexpr (KEEP) {
        // The low half of T through the exchange buffer (n words): written in limb order,
        // read in the transform's; dif_regs orders its first use of the buffer with a barrier.
        u32* Tx = (u32*)X.z2;
        __syncthreads();  // the square's last reads of the buffer
#pragma unroll
        for (u32 i = 0; i < 8; i++) Tx[8 * t + i] = keep[i];
        Tsrc = Tx;
    }
    __syncthreads();  // T's limbs written by every thread
    if (MM_ACT(t)) {
#pragma unroll
        for (u32 e = 0; e < 8; e++) {
            u32 pos = slot_pos(t, e, r0, ls0, T);
            u32 tv = Tsrc[pos];
            v1[e] = Op1::real_mul(tv, V.w1[pos]);
            v2[e] = Op2::real_mul(tv, V.w2[pos]);
            MM_GROUP_FENCE(e);
        }
    }
    dif_regs<LOGN, TILED, GT>(t, v1, v2, V.tw1, V.tw2, X, on);
    t = slot_tid<GT>();
    if (MM_ACT(t)) {
#pragma unroll
        for (u32 e = 0; e < 8; e++) {
            v1[e] = Op1::mul(v1[e], np1[8 * t + e]);
            v2[e] = Op2::mul(v2[e], np2[8 * t + e]);
            MM_GROUP_FENCE(e);
        }
    }
    dit_regs<LOGN, TILED, GT, false>(t, v1, v2, V.itw1, V.itw2, X, on);  // after the DIF's own exchanges
    __syncthreads();
    t = slot_tid<GT>();
    if (MM_ACT(t)) {
#pragma unroll
        for (u32 e = 0; e < 8; e++) {
            u32 pos = slot_pos(t, e, r0, ls0, T);
            Z1 a = Op1::mul(v1[e], V.iw1[pos]);
            Z2 b = Op2::mul(v2[e], V.iw2[pos]);
            cput(C, pos, crt_lazy(a.re, b.re, inv));
            MM_GROUP_FENCE(e);
        }
    }
    __syncthreads();
    t = slot_tid<GT>();
    u32 m[8];
    carry_regs<8, NT>(t, [&](int i) { return MM_ACT(t) ? cget(C, 8 * t + i) : (u128)0; }, m, sc, 0);
    t = slot_tid<GT>();
    // m packed as z_j = m_2j + i m_2j+1 at half-length positions j = 4t .. 4t+3.
    if (MM_ACT(t)) {
#pragma unroll
        for (u32 i = 0; i < 4; i++) {
            u32 q = sw(4 * t + i);
            X.z1[q] = Z1{mod1(m[2 * i]), mod1(m[2 * i + 1])};
            X.z2[q] = Z2{m[2 * i], m[2 * i + 1]};
        }
    }
    __syncthreads();
    t = slot_tid<GT>();
    const bool hact = on && t < Th;
    if (hact) get(X, v1, v2, t, PH::radix(0), PH::lstride(0), Th);
    dif_regs<LOGN - 1, TILED, GT>(t, v1, v2, H.tw1, H.tw2, X, on);
    t = slot_tid<GT>();
    // Pointwise with the partner position: out_p = (z_p + conj z_q) A_p + (z_p - conj z_q) B_p.
    __syncthreads();
    if (hact) {
#pragma unroll
        for (u32 e = 0; e < 8; e++) {
            u32 q = sw(8 * t + e);
            X.z1[q] = v1[e];
            X.z2[q] = v2[e];
        }
    }
    __syncthreads();
    t = slot_tid<GT>();
    if (hact) {
#pragma unroll
        for (u32 e = 0; e < 8; e++) {
            u32 p = 8 * t + e, q = sw(partner(p));
            Z1 zq1 = Op1::conj(X.z1[q]);
            Z2 zq2 = Op2::conj(X.z2[q]);
            Z1 e1 = zadd<Op1>(v1[e], zq1), d1 = zsub<Op1>(v1[e], zq1);
            Z2 e2 = zadd<Op2>(v2[e], zq2), d2 = zsub<Op2>(v2[e], zq2);
            v1[e] = zadd<Op1>(Op1::mul(e1, ab1[p]), Op1::mul(d1, ab1[h + p]));
            v2[e] = zadd<Op2>(Op2::mul(e2, ab2[p]), Op2::mul(d2, ab2[h + p]));
            MM_GROUP_FENCE(e);
        }
    }
    t = slot_tid<GT>();
    dit_regs<LOGN - 1, TILED, GT>(t, v1, v2, H.itw1, H.itw2, X, on);
    t = slot_tid<GT>();
    // Cyclic product element j carries coefficients 2j (re) and 2j + 1 (im).
    __syncthreads();
    if (hact) {
#pragma unroll
        for (u32 e = 0; e < 8; e++) {
            u32 j = slot_pos(t, e, PH::radix(0), PH::lstride(0), Th);
            cput(C, 2 * j, crt_lazy(v1[e].re, v2[e].re, inv));
            cput(C, 2 * j + 1, crt_lazy(v1[e].im, v2[e].im, inv));
            MM_GROUP_FENCE(e);
        }
    }
    __syncthreads();
    t = slot_tid<GT>();
    // G = (cyclic + T mod R) mod (R - 1), x = floor(T / R) + G.
    u32 xs[8];
    u64 cc = carry_regs<8, NT>(
        t, [&](int i) { return MM_ACT(t) ? cget(C, 8 * t + i) + (KEEP ? keep[i] : Tl[8 * t + i]) : (u128)0; }, xs, sc, 0);
    // With several slot groups in the workgroup the fold count must be the same in all of them
    // (the carry holds barriers): they fold while any group has a carry left; a further fold
    // with carry 0 leaves a group's limbs as they are.
#pragma nounroll
    for (int guard = 0; guard < 4 && (GT ? __syncthreads_or(cc != 0) != 0 : cc != 0); guard++) {
        u32 prev[8];
#pragma unroll
        for (u32 i = 0; i < 8; i++) prev[i] = xs[i];
        cc = carry_regs<8, NT>(t, [&](int i) { return (u128)prev[i]; }, xs, sc, cc);
    }
    t = slot_tid<GT>();
    u32 lo[8];
#pragma unroll
    for (u32 i = 0; i < 8; i++) lo[i] = xs[i];
    u64 top = carry_regs<8, NT>(
        t, [&](int i) { return MM_ACT(t) ? (u128)lo[i] + (KEEP ? keep[8 + i] : Tl[k + 8 * t + i]) : (u128)0; }, xs, sc, 0);
    t = slot_tid<GT>();
    if (MM_ACT(t)) {
#pragma unroll
        for (u32 i = 0; i < 8; i++) x[8 * t + i] = xs[i];
    }
    return (top != 0) | (cc != 0);
}

// ---- n = 2^14 (L = 32,768): two register-resident sub-transforms of n / 2 -------------------
// 8 elements per thread would take 2,048 threads, more than a workgroup holds. A length-n DIF
// is instead one radix-2 stage, A_j = a_j + a_{j+n/2} and B_j = (a_j - a_{j+n/2}) omega^j
// (j < n/2), followed by length-n/2 DIFs of A and of B, whose bit-reversed outputs are the
// first and the second half of the length-n spectrum in the same bit-reversed order. The
// inverse mirrors it. The 1,024 threads hold A in registers (the layout of the n/2 plan) while B
// waits in the slot's HBM work vector; each thread stores and reloads only its own elements
// there, in thread-major order (coalesced, and no barrier needed). The sub-transforms use the
// n/2 tables, whose roots are omega^2. Coefficients are staged in HBM (3 words each, 6 n words
// per slot) instead of the LDS tile, which keeps exchanging the sub-transforms.

// Thread-major index of element e of thread t among T threads.
__device__ __forceinline__ u32 tmaj(u32 t, u32 e, u32 T) { return e * T + t; }

// The radix-2 split of the weighted input (coefficient j = in[j], j < n, the upper half of the
// 2n coefficients zero): A in registers, B stored to w[0, n/2) thread-major.
template <u32 LOGS>
__device__ __forceinline__ void split_load(const VTab& V, const u32* in, Z1* v1, Z2* v2, Z1* w1, Z2* w2) {
    typedef CPlan<LOGS> PS;
    constexpr u32 hs = PS::n, T = PS::T;
    const u32 t = slot_tid<0>();
#pragma unroll
    for (u32 e = 0; e < 8; e++) {
        const u32 j = slot_pos(t, e, PS::radix(0), PS::lstride(0), T);
        const u32 lo = in[j], hi = in[j + hs];
        const Z1 lo1 = Op1::real_mul(lo, V.w1[j]), hi1 = Op1::real_mul(hi, V.w1[j + hs]);
        const Z2 lo2 = Op2::real_mul(lo, V.w2[j]), hi2 = Op2::real_mul(hi, V.w2[j + hs]);
        v1[e] = zadd<Op1>(lo1, hi1);
        v2[e] = zadd<Op2>(lo2, hi2);
        w1[tmaj(t, e, T)] = Op1::mul(zsub<Op1>(lo1, hi1), V.sw1[j]);
        w2[tmaj(t, e, T)] = Op2::mul(zsub<Op2>(lo2, hi2), V.sw2[j]);
        MM_GROUP_FENCE(e);
    }
}

// One half of the product's spectrum through its sub-transform: DIF, the pointwise step
// (square, or times the constant spectrum cs at half offset `half`), DIT. In and out in
// registers, pass-0 layout of the n/2 plan.
template <u32 LOGS, bool SQUARE>
__device__ __forceinline__ void split_half(const VTab& S, const Xbuf& X, Z1* v1, Z2* v2, const F1* cs1, const F2* cs2) {
    const u32 t0 = slot_tid<0>();
    dif_regs<LOGS, true>(t0, v1, v2, S.tw1, S.tw2, X);
    if constexpr (SQUARE) {
#pragma unroll
        for (u32 e = 0; e < 8; e++) {
            v1[e] = Op1::sqr(v1[e]);
            v2[e] = Op2::sqr(v2[e]);
        }
    } else {
        const u32 t = slot_tid<0>();
#pragma unroll
        for (u32 e = 0; e < 8; e++) {
            v1[e] = Op1::mul(v1[e], cs1[8 * t + e]);
            v2[e] = Op2::mul(v2[e], cs2[8 * t + e]);
            MM_GROUP_FENCE(e);
        }
    }
    dit_regs<LOGS, true, 0, false>(slot_tid<0>(), v1, v2, S.itw1, S.itw2, X);  // after the DIF's own exchanges
}

// The spectrum's product through both halves: A from registers (its result A' to w[n/2, n)),
// then B from w[0, n/2) (its result times omega^-j back to w[0, n/2)). Both thread-major.
template <u32 LOGS, bool SQUARE>
__device__ __forceinline__ void split_product(const VArgs& A, const Xbuf& X, Z1* v1, Z2* v2, Z1* w1, Z2* w2, const F1* cs1,
                                              const F2* cs2) {
    typedef CPlan<LOGS> PS;
    constexpr u32 hs = PS::n, T = PS::T;
    const VTab& S = A.ht;
    split_half<LOGS, SQUARE>(S, X, v1, v2, cs1, cs2);
    u32 t = slot_tid<0>();
#pragma unroll
    for (u32 e = 0; e < 8; e++) {
        w1[hs + tmaj(t, e, T)] = v1[e];
        w2[hs + tmaj(t, e, T)] = v2[e];
        v1[e] = w1[tmaj(t, e, T)];
        v2[e] = w2[tmaj(t, e, T)];
    }
    split_half<LOGS, SQUARE>(S, X, v1, v2, cs1 ? cs1 + hs : cs1, cs2 ? cs2 + hs : cs2);
    t = slot_tid<0>();
#pragma unroll
    for (u32 e = 0; e < 8; e++) {
        const u32 j = slot_pos(t, e, PS::radix(0), PS::lstride(0), T);
        w1[tmaj(t, e, T)] = Op1::mul(v1[e], A.t.isw1[j]);
        w2[tmaj(t, e, T)] = Op2::mul(v2[e], A.t.isw2[j]);
        MM_GROUP_FENCE(e);
    }
}

// The inverse's last stage, out_j = A'_j + t_j and out_{j+n/2} = A'_j - t_j, then the
// unweighting and CRT; element i's coefficients go to C at i (real part) and, with IM, at
// i + n (imaginary part).
template <u32 LOGS, bool IM>
__device__ __forceinline__ void split_combine(const Cbuf& C, const Z1* w1, const Z2* w2, const F1* iw1, const S2* iw2, u32 inv) {
    typedef CPlan<LOGS> PS;
    constexpr u32 hs = PS::n, n = 2 * hs, T = PS::T;
    const u32 t = slot_tid<0>();
#pragma unroll
    for (u32 e = 0; e < 8; e++) {
        const u32 j = slot_pos(t, e, PS::radix(0), PS::lstride(0), T);
        const Z1 a1 = w1[hs + tmaj(t, e, T)], b1 = w1[tmaj(t, e, T)];
        const Z2 a2 = w2[hs + tmaj(t, e, T)], b2 = w2[tmaj(t, e, T)];
        const Z1 p1 = Op1::mul(zadd<Op1>(a1, b1), iw1[j]), q1 = Op1::mul(zsub<Op1>(a1, b1), iw1[j + hs]);
        const Z2 p2 = Op2::mul(zadd<Op2>(a2, b2), iw2[j]), q2 = Op2::mul(zsub<Op2>(a2, b2), iw2[j + hs]);
        cput(C, j, crt_lazy(p1.re, p2.re, inv));
        cput(C, j + hs, crt_lazy(q1.re, q2.re, inv));
        if (IM) {
            cput(C, j + n, crt_lazy(p1.im, p2.im, inv));
            cput(C, j + hs + n, crt_lazy(q1.im, q2.im, inv));
        }
        MM_GROUP_FENCE(e);
    }
}

// Squaring at n = 2^(LOGS+1) = 2^14: T = c x^2 into Tl (2k = 2n limbs); nonzero on a carry out.
template <u32 LOGS>
__device__ __forceinline__ u32 v2_square_split(const VArgs& A, const Xbuf& X, u32* Tl, u64* sc, const u32* x, Z1* w1, Z2* w2,
                                               u32* coef, const F1* iw1, const S2* iw2) {
    constexpr u32 n = 2u << LOGS;
    Z1 v1[8];
    Z2 v2[8];
    split_load<LOGS>(A.t, x, v1, v2, w1, w2);
    split_product<LOGS, true>(A, X, v1, v2, w1, w2, nullptr, nullptr);
    const Cbuf C{coef, coef + 2 * n, coef + 4 * n};
    split_combine<LOGS, true>(C, w1, w2, iw1, iw2, A.crt_inv);
    __syncthreads();
    u32 t = slot_tid<0>();
    u32 l[16];
    const u64 mid = carry_regs<16>(t, [&](int i) { return cget(C, 16 * t + i); }, l, sc, 0);
    t = slot_tid<0>();
#pragma unroll
    for (u32 i = 0; i < 16; i++) Tl[16 * t + i] = l[i];
    const u64 top = carry_regs<16>(t, [&](int i) { return cget(C, n + 16 * t + i); }, l, sc, mid);
    t = slot_tid<0>();
#pragma unroll
    for (u32 i = 0; i < 16; i++) Tl[n + 16 * t + i] = l[i];
    return top != 0;
}

// Reduction at n = 2^14, as v2_redc: m = (T mod R) N' mod R through the split transforms,
// then the cyclic product m N mod (R - 1) at h = n/2 (one register-resident transform pair),
// with the packed m and the partner exchange through the HBM work vector.
template <u32 LOGS>
__device__ __forceinline__ u32 v2_redc_split(const VArgs& A, const Xbuf& X, u32* Tl, u64* sc, u32* x, Z1* w1, Z2* w2, u32* coef,
                                             const F1* np1, const F2* np2, const F1* ab1, const F2* ab2) {
    typedef CPlan<LOGS> PS;
    constexpr u32 hs = PS::n, n = 2 * hs, k = n, T = PS::T;
    const u32 inv = A.crt_inv;
    const VTab& H = A.ht;
    const Cbuf C{coef, coef + 2 * n, coef + 4 * n};
    Z1 v1[8];
    Z2 v2[8];
    __syncthreads();  // T's limbs written by every thread
    split_load<LOGS>(A.t, Tl, v1, v2, w1, w2);
    split_product<LOGS, false>(A, X, v1, v2, w1, w2, np1, np2);
    split_combine<LOGS, false>(C, w1, w2, A.t.iw1, A.t.iw2, inv);
    __syncthreads();
    u32 t = slot_tid<0>();
    u32 m[16];
    carry_regs<16>(t, [&](int i) { return cget(C, 16 * t + i); }, m, sc, 0);
    // m packed as z_j = m_2j + i m_2j+1 (j < h = n/2) at j = 8t .. 8t+7, natural order.
    t = slot_tid<0>();
#pragma unroll
    for (u32 i = 0; i < 8; i++) {
        w1[8 * t + i] = Z1{mod1(m[2 * i]), mod1(m[2 * i + 1])};
        w2[8 * t + i] = Z2{m[2 * i], m[2 * i + 1]};
    }
    __syncthreads();
    t = slot_tid<0>();
#pragma unroll
    for (u32 e = 0; e < 8; e++) {
        const u32 j = slot_pos(t, e, PS::radix(0), PS::lstride(0), T);
        v1[e] = w1[j];
        v2[e] = w2[j];
    }
    dif_regs<LOGS, true>(t, v1, v2, H.tw1, H.tw2, X);
    // Pointwise with the partner position: out_p = (z_p + conj z_q) A_p + (z_p - conj z_q) B_p,
    // z_q read from the HBM work vector's second half.
    t = slot_tid<0>();
#pragma unroll
    for (u32 e = 0; e < 8; e++) {
        w1[hs + 8 * t + e] = v1[e];
        w2[hs + 8 * t + e] = v2[e];
    }
    __syncthreads();
    t = slot_tid<0>();
#pragma unroll
    for (u32 e = 0; e < 8; e++) {
        const u32 p = 8 * t + e, q = hs + partner(p);
        const Z1 zq1 = Op1::conj(w1[q]);
        const Z2 zq2 = Op2::conj(w2[q]);
        const Z1 e1 = zadd<Op1>(v1[e], zq1), d1 = zsub<Op1>(v1[e], zq1);
        const Z2 e2 = zadd<Op2>(v2[e], zq2), d2 = zsub<Op2>(v2[e], zq2);
        v1[e] = zadd<Op1>(Op1::mul(e1, ab1[p]), Op1::mul(d1, ab1[hs + p]));
        v2[e] = zadd<Op2>(Op2::mul(e2, ab2[p]), Op2::mul(d2, ab2[hs + p]));
        MM_GROUP_FENCE(e);
    }
    dit_regs<LOGS, true>(slot_tid<0>(), v1, v2, H.itw1, H.itw2, X);
    // Cyclic product element j carries coefficients 2j (re) and 2j + 1 (im).
    t = slot_tid<0>();
#pragma unroll
    for (u32 e = 0; e < 8; e++) {
        const u32 j = slot_pos(t, e, PS::radix(0), PS::lstride(0), T);
        cput(C, 2 * j, crt_lazy(v1[e].re, v2[e].re, inv));
        cput(C, 2 * j + 1, crt_lazy(v1[e].im, v2[e].im, inv));
        MM_GROUP_FENCE(e);
    }
    __syncthreads();
    t = slot_tid<0>();
    // G = (cyclic + T mod R) mod (R - 1), x = floor(T / R) + G.
    u32 xs[16];
    u64 cc = carry_regs<16>(t, [&](int i) { return cget(C, 16 * t + i) + Tl[16 * t + i]; }, xs, sc, 0);
#pragma nounroll
    for (int guard = 0; guard < 4 && cc != 0; guard++) {
        u32 prev[16];
#pragma unroll
        for (u32 i = 0; i < 16; i++) prev[i] = xs[i];
        cc = carry_regs<16>(t, [&](int i) { return (u128)prev[i]; }, xs, sc, cc);
    }
    t = slot_tid<0>();
    u32 lo[16];
#pragma unroll
    for (u32 i = 0; i < 16; i++) lo[i] = xs[i];
    const u64 top = carry_regs<16>(t, [&](int i) { return (u128)lo[i] + Tl[k + 16 * t + i]; }, xs, sc, 0);
    t = slot_tid<0>();
#pragma unroll
    for (u32 i = 0; i < 16; i++) x[16 * t + i] = xs[i];
    return (top != 0) | (cc != 0);
}

#undef MM_ACT

}  // namespace v2
}  // namespace mm
