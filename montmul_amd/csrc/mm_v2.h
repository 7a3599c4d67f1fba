// mm_v2.h - register-resident transforms for the fused step (lengths n = 512 .. 8,192, and the
// two halves of n = 16,384).
//
// Each thread holds 8 elements of both fields in registers; a transform is a sequence of
// radix-8 (or one radix-2/4) passes, each a butterfly over the thread's elements followed (DIF)
// or preceded (DIT) by that pass's twiddles, with an exchange through a swizzled LDS buffer
// between passes (the whole vector, or in global mode a tile of it; an exchange whose partners
// lie in one wave goes through that wave's own part of the buffer). The pass layouts make
// the pointwise steps exchange-free: after the last DIF pass thread t holds positions
// 8t .. 8t+7, which is also where the first DIT pass reads.
//
// Arithmetic is lazy. GF(M61) values are 64-bit words <= 2^61 + 6 (a fold (x & P2) + (x >> 61)
// after every operation keeps them there); GF(M31) values are 32-bit words in [0, P1], P1
// standing for 0. Products go through 32 x 32 -> 64 multiply-adds (v_mad_u64_u32) on 31-bit
// halves and one Mersenne fold. On gfx950 a multiply-add issues at about the rate of a 64-bit
// add (0.93 against 0.96 per CU per cycle; 32-bit adds 1.37), so the count of 64-bit
// instructions, not of multiplies, is the cost. Constants read from memory (twiddles, weights,
// spectra) are canonical.
//
// The roots are those of the reference (field.rs:237-267): in both fields omega^(n/4) = i and
// omega^(n/8) = 2^h (1 + i) with h = 30 (M61) and 15 (M31), so the radix-8 butterflies need
// only rotations and adds internally. The host asserts both identities when it builds tables.
#pragma once

#include <type_traits>

#include "mm_kernels.h"

namespace mm {
namespace v2 {

typedef uint64_t u64;
typedef uint32_t u32;

// ---- Lazy GF(M61) ---------------------------------------------------------------------------
constexpr u64 K4 = 4 * P2;  // a multiple of P2 above every lazy value

MM_HD u64 f2(u64 x) { return (x & P2) + (x >> 61); }  // < 2^64 in, <= 2^61 + 6 out
MM_HD u64 ad2(u64 a, u64 b) { return f2(a + b); }
MM_HD u64 sb2(u64 a, u64 b) { return f2(a + (K4 - b)); }
// x * 2^30 for x < 2^62: x = xh 2^31 + xl, 2^61 = 1, so x 2^30 = xh + xl 2^30.
MM_HD u64 rot2(u64 x) {
    u64 xl = (u32)x & 0x7fffffffu;
    return (xl << 30) + (x >> 31);
}
// Canonical value of a lazy one.
MM_HD u64 canon2(u64 x) { return x >= P2 ? x - P2 : x; }

// ---- GF(M61) products on 31-bit halves ------------------------------------------------------
// A lazy word x <= 2^61 + 6 is x1 2^31 + x0 with x0 < 2^31 and x1 <= 2^30. With a second word
// y = y1 2^31 + y0 (y0 < 2^31, y1 <= 2^31), x y = x0 y0 + (x0 y1 + x1 y0) 2^31 + x1 y1 2^62, and
// 2^62 = 2 (mod P2). A product (or a sum of two) is therefore a low column x0 y0 + (2 x1) y1
// and a middle column m = x0 y1 + x1 y0, each a chain of 32 x 32 -> 64 multiply-adds with no
// carries between words, and the middle column enters as
//     m 2^31 = (m >> 30) 2^61 + (m mod 2^30) 2^31 = (m >> 30) + (m mod 2^30) 2^31  (mod P2).
struct Sp {
    u32 x0, x1, x1d;  // x0, x1 and 2 x1 of a lazy word
};
// A canonical constant with both parts split at bit 31 (re = r1 2^31 + r0, im likewise).
struct alignas(16) S2 {
    u32 r0, r1, i0, i1;
};

MM_HD Sp spl(u64 x) {
    u32 x1 = (u32)(x >> 31);
    return Sp{(u32)x & P1, x1, x1 << 1};
}
MM_HD S2 spl_c(F2 c) { return S2{(u32)c.re & P1, (u32)(c.re >> 31), (u32)c.im & P1, (u32)(c.im >> 31)}; }
MM_HD u64 mad(u32 a, u32 b, u64 c) { return (u64)a * b + c; }
// 2^31 held in a scalar register: a multiply-add by it is one instruction where a 64-bit shift,
// two masks and an add would be four.
MM_HD u32 k31() {
#if defined(__HIP_DEVICE_COMPILE__)
    u32 k;
    asm("s_mov_b32 %0, 0x80000000" : "=s"(k));
    return k;
#else
    return 0x80000000u;
#endif
}
// acc + m 2^31 (mod P2), folded, for acc <= 2^63 + 2^62 and m < 2^64 - 2^34.
MM_HD u64 comb(u64 acc, u64 m) { return f2(mad((u32)m & 0x3fffffffu, k31(), acc + (m >> 30))); }
// a c + b d for lazy a, b and canonical c = c1 2^31 + c0, d = d1 2^31 + d0 (c1, d1 < 2^30):
// low column < 2 (2^62 + 2^61), middle column < 2^63.
MM_HD u64 dot2(Sp a, Sp b, u32 c0, u32 c1, u32 d0, u32 d1) {
    u64 acc = mad(a.x0, c0, mad(a.x1d, c1, mad(b.x0, d0, mad(b.x1d, d1, 0))));
    u64 m = mad(a.x0, c1, mad(a.x1, c0, mad(b.x0, d1, mad(b.x1, d0, 0))));
    return comb(acc, m);
}
// x y for lazy x and y = y1 2^31 + y0 with y0 < 2^31, y1 <= 2^31 (y < 2^62 + 2^31):
// low column <= 2^62 + 2^62, middle column < 2^62 + 2^61.
MM_HD u64 prod2(Sp x, u32 y0, u32 y1) { return comb(mad(x.x0, y0, mad(x.x1d, y1, 0)), mad(x.x0, y1, mad(x.x1, y0, 0))); }

// ---- Lazy GF(M31): words in [0, P1] -------------------------------------------------------
MM_HD u32 umin(u32 a, u32 b) { return a < b ? a : b; }
MM_HD u32 ad1(u32 a, u32 b) { u32 s = a + b; return umin(s, s - P1); }
MM_HD u32 sb1(u32 a, u32 b) { u32 d = a - b; return umin(d, d + P1); }
MM_HD u32 rot1(u32 x) { return ((x << 15) & P1) | (x >> 16); }  // x * 2^15
// P mod P1 into [0, P1] for P <= 2 P1^2.
MM_HD u32 red1w(u64 p) {
    u32 a = (u32)p & P1, b = (u32)(p >> 31);
    u32 s = a + (b & P1) + (b >> 31);
    return umin(s, s - P1);
}
MM_HD u32 canon1(u32 x) { return x == P1 ? 0 : x; }

// ---- Complex elements and per-field operations ----------------------------------------------
struct Z1 {
    u32 re, im;
};
struct alignas(16) Z2 {
    u64 re, im;
};

struct Op1 {
    typedef u32 B;
    typedef Z1 Z;
    typedef F1 C;  // canonical constant
    static MM_HD B ad(B a, B b) { return ad1(a, b); }
    static MM_HD B sb(B a, B b) { return sb1(a, b); }
    static MM_HD B rot(B a) { return rot1(a); }
    // First-stage add/sub of a butterfly: GF(M31) words have no headroom, so these reduce.
    static MM_HD B adL(B a, B b) { return ad1(a, b); }
    static MM_HD B sbL(B a, B b) { return sb1(a, b); }
    // x * w, w canonical.
    static MM_HD Z mul(Z x, C w) {
        u64 re = (u64)x.re * w.re + (u64)x.im * (P1 - w.im);
        u64 im = (u64)x.re * w.im + (u64)x.im * w.re;
        return Z{red1w(re), red1w(im)};
    }
    static MM_HD Z sqr(Z x) {
        u64 re = (u64)x.re * x.re + (u64)x.im * (P1 - x.im);
        u64 im = 2 * ((u64)x.re * x.im);
        return Z{red1w(re), red1w(im)};
    }
    static MM_HD Z scale(Z x, u32 c) { return Z{red1w((u64)x.re * c), red1w((u64)x.im * c)}; }
    // A real 32-bit value times a canonical weight.
    static MM_HD Z real_mul(u32 v, C w) {
        u32 m = mod1(v);
        return Z{red1w((u64)m * w.re), red1w((u64)m * w.im)};
    }
    static MM_HD Z conj(Z x) { return Z{x.re, sb1(0, x.im)}; }
};

struct Op2 {
    typedef u64 B;
    typedef Z2 Z;
    typedef S2 C;  // canonical constant, split
    static MM_HD B ad(B a, B b) { return ad2(a, b); }
    static MM_HD B sb(B a, B b) { return sb2(a, b); }
    static MM_HD B rot(B a) { return rot2(a); }
    // First-stage add/sub of a butterfly, unreduced: from lazy inputs (<= 2^61 + 6) the results
    // stay below 3 * 2^61 + 6, which the folded second-stage ad2/sb2 (K4 = 4 P2) accept.
    static MM_HD B adL(B a, B b) { return a + b; }
    static MM_HD B sbL(B a, B b) { return a + (2 * P2 - b); }
    // x * w: re = x.re w.re + x.im (P2 - w.im), im = x.re w.im + x.im w.re.
    static MM_HD Z mul(Z x, C w) {
        Sp a = spl(x.re), b = spl(x.im);
        u32 n0 = P1 - w.i0, n1 = 0x3fffffffu - w.i1;  // P2 - w.im, split
        return Z{dot2(a, b, w.r0, w.r1, n0, n1), dot2(a, b, w.i0, w.i1, w.r0, w.r1)};
    }
    static MM_HD Z mul(Z x, F2 w) { return mul(x, spl_c(w)); }
    // (re + i im)^2 = (re + im)(re - im) + i re (2 im).
    static MM_HD Z sqr(Z x) {
        u64 s = ad2(x.re, x.im), d = sb2(x.re, x.im);
        u64 y = x.im << 1;  // < 2^62 + 13
        return Z{prod2(spl(s), (u32)d & P1, (u32)(d >> 31)), prod2(spl(x.re), (u32)y & P1, (u32)(y >> 31))};
    }
    static MM_HD Z scale(Z x, u32 c) {
        // c <= 4: x c < 2^63.1.
        return Z{f2((u64)(u32)x.re * c + (((u64)(u32)(x.re >> 32) * c) << 32)),
                 f2((u64)(u32)x.im * c + (((u64)(u32)(x.im >> 32) * c) << 32))};
    }
    // A real 32-bit value times a canonical weight: v w = v w0 + (v w1) 2^31.
    static MM_HD Z real_mul(u32 v, C w) { return Z{comb(mad(v, w.r0, 0), mad(v, w.r1, 0)), comb(mad(v, w.i0, 0), mad(v, w.i1, 0))}; }
    static MM_HD Z conj(Z x) { return Z{x.re, sb2(0, x.im)}; }
};

template <class O> MM_HD typename O::Z zadd(typename O::Z a, typename O::Z b) {
    return {O::ad(a.re, b.re), O::ad(a.im, b.im)};
}
template <class O> MM_HD typename O::Z zsub(typename O::Z a, typename O::Z b) {
    return {O::sb(a.re, b.re), O::sb(a.im, b.im)};
}
// The same for a butterfly's first stage (unreduced in GF(M61)).
template <class O> MM_HD typename O::Z zaddL(typename O::Z a, typename O::Z b) {
    return {O::adL(a.re, b.re), O::adL(a.im, b.im)};
}
template <class O> MM_HD typename O::Z zsubL(typename O::Z a, typename O::Z b) {
    return {O::sbL(a.re, b.re), O::sbL(a.im, b.im)};
}
template <class O> MM_HD typename O::Z zsub_iL(typename O::Z a, typename O::Z b) {
    return {O::sbL(b.im, a.im), O::sbL(a.re, b.re)};
}
// (a - b) * i
template <class O> MM_HD typename O::Z zsub_i(typename O::Z a, typename O::Z b) {
    return {O::sb(b.im, a.im), O::sb(a.re, b.re)};
}
// (a - b) * (-i)
template <class O> MM_HD typename O::Z zsub_mi(typename O::Z a, typename O::Z b) {
    return {O::sb(a.im, b.im), O::sb(b.re, a.re)};
}

// ---- Butterflies over a thread's slots ------------------------------------------------------
// DIF radix 8 on v[o], v[o+1], ..., v[o+7]: output slot m holds frequency rev3(m).
template <class O> MM_HD void dif8(typename O::Z* v) {
    typedef typename O::Z Z;
    Z a0 = zaddL<O>(v[0], v[4]), a1 = zaddL<O>(v[1], v[5]), a2 = zaddL<O>(v[2], v[6]), a3 = zaddL<O>(v[3], v[7]);
    Z a4 = zsubL<O>(v[0], v[4]);
    Z d1 = zsubL<O>(v[1], v[5]);
    Z a5{O::rot(O::sb(d1.re, d1.im)), O::rot(O::ad(d1.re, d1.im))};  // d1 2^h (1 + i)
    Z a6 = zsub_iL<O>(v[2], v[6]);                                   // (v2 - v6) i
    Z e3 = zsubL<O>(v[7], v[3]);                                     // -(v3 - v7)
    Z a7{O::rot(O::ad(e3.re, e3.im)), O::rot(O::sb(e3.im, e3.re))};  // (v3 - v7) 2^h (-1 + i)
    Z b0 = zadd<O>(a0, a2), b2 = zsub<O>(a0, a2), b1 = zadd<O>(a1, a3), b3 = zsub_i<O>(a1, a3);
    Z b4 = zadd<O>(a4, a6), b6 = zsub<O>(a4, a6), b5 = zadd<O>(a5, a7), b7 = zsub_i<O>(a5, a7);
    v[0] = zadd<O>(b0, b1);
    v[1] = zsub<O>(b0, b1);
    v[2] = zadd<O>(b2, b3);
    v[3] = zsub<O>(b2, b3);
    v[4] = zadd<O>(b4, b5);
    v[5] = zsub<O>(b4, b5);
    v[6] = zadd<O>(b6, b7);
    v[7] = zsub<O>(b6, b7);
}
template <class O> MM_HD void dif4(typename O::Z* v) {
    typedef typename O::Z Z;
    Z a0 = zaddL<O>(v[0], v[2]), a2 = zsubL<O>(v[0], v[2]), a1 = zaddL<O>(v[1], v[3]), a3 = zsub_iL<O>(v[1], v[3]);
    v[0] = zadd<O>(a0, a1);
    v[1] = zsub<O>(a0, a1);
    v[2] = zadd<O>(a2, a3);
    v[3] = zsub<O>(a2, a3);
}
template <class O> MM_HD void dif2(typename O::Z* v) {
    typename O::Z a = v[0], b = v[1];
    v[0] = zadd<O>(a, b);
    v[1] = zsub<O>(a, b);
}
// DIT radix 8 (inverse roots): input slot m holds frequency rev3(m), output natural.
template <class O> MM_HD void dit8(typename O::Z* v) {
    typedef typename O::Z Z;
    Z b0 = zaddL<O>(v[0], v[1]), b1 = zsubL<O>(v[0], v[1]), b2 = zaddL<O>(v[2], v[3]), b3 = zsubL<O>(v[2], v[3]);
    Z b4 = zaddL<O>(v[4], v[5]), b5 = zsubL<O>(v[4], v[5]), b6 = zaddL<O>(v[6], v[7]), b7 = zsubL<O>(v[6], v[7]);
    // span 2: second element of pairs (1,3), (5,7) times -i.
    Z c0 = zadd<O>(b0, b2), c2 = zsub<O>(b0, b2);
    Z c1{O::ad(b1.re, b3.im), O::sb(b1.im, b3.re)}, c3{O::sb(b1.re, b3.im), O::ad(b1.im, b3.re)};
    Z c4 = zadd<O>(b4, b6), c6 = zsub<O>(b4, b6);
    Z c5{O::ad(b5.re, b7.im), O::sb(b5.im, b7.re)}, c7{O::sb(b5.re, b7.im), O::ad(b5.im, b7.re)};
    // span 4: t_i = omega8^-i c_{i+4}: 1, 2^h (1 - i), -i, 2^h (-1 - i).
    v[0] = zadd<O>(c0, c4);
    v[4] = zsub<O>(c0, c4);
    Z t1{O::rot(O::ad(c5.re, c5.im)), O::rot(O::sb(c5.im, c5.re))};
    v[1] = zadd<O>(c1, t1);
    v[5] = zsub<O>(c1, t1);
    v[2] = Z{O::ad(c2.re, c6.im), O::sb(c2.im, c6.re)};
    v[6] = Z{O::sb(c2.re, c6.im), O::ad(c2.im, c6.re)};
    typename O::B u3 = O::rot(O::sb(c7.im, c7.re)), w3 = O::rot(O::ad(c7.re, c7.im));  // t3 = (u3, -w3)
    v[3] = Z{O::ad(c3.re, u3), O::sb(c3.im, w3)};
    v[7] = Z{O::sb(c3.re, u3), O::ad(c3.im, w3)};
}
template <class O> MM_HD void dit4(typename O::Z* v) {
    typedef typename O::Z Z;
    Z b0 = zaddL<O>(v[0], v[1]), b1 = zsubL<O>(v[0], v[1]), b2 = zaddL<O>(v[2], v[3]), b3 = zsubL<O>(v[2], v[3]);
    v[0] = zadd<O>(b0, b2);
    v[2] = zsub<O>(b0, b2);
    v[1] = Z{O::ad(b1.re, b3.im), O::sb(b1.im, b3.re)};
    v[3] = Z{O::sb(b1.re, b3.im), O::ad(b1.im, b3.re)};
}
template <class O> MM_HD void dit2(typename O::Z* v) { dif2<O>(v); }

// ---- Pass plan ---------------------------------------------------------------------------
// DIF passes in order: radix r_p (2, 4 or 8; only the first may be below 8) and stride s_p,
// with s_0 = n / r_0 and s_{p+1} = s_p / r_{p+1}; the last stride is 1. tw_off[p] indexes the
// pass's twiddles (s_p * (r_p - 1) entries: for j < s_p and slot m = 1 .. r_p - 1, the
// twiddle omega_(r s)^(j rev_r(m)); the DIT table holds the inverses).
constexpr int MAXPASS = 6;
struct Plan {
    u32 n, logn, threads, npass;
    u32 radix[MAXPASS], lstride[MAXPASS], tw_off[MAXPASS];
};

// Radix-r bit reversal of m.
__device__ __forceinline__ u32 revr(u32 m, u32 r) {
    return r == 8 ? ((m & 1) << 2) | (m & 2) | (m >> 2) : r == 4 ? ((m & 1) << 1) | (m >> 1) : m;
}

// Element position of thread t's slot e in pass p (r = radix, ls = log2 stride, T threads).
__device__ __forceinline__ u32 slot_pos(u32 t, u32 e, u32 r, u32 ls, u32 T) {
    u32 u = e / r, m = e % r;
    u32 g = u * T + t;
    u32 b = g >> ls, j = g & ((1u << ls) - 1);
    return b * r * (1u << ls) + j + (m << ls);
}
// Swizzled index in the exchange buffer (bank-conflict-free 8-lane groups).
__device__ __forceinline__ u32 sw(u32 p) { return p ^ ((p >> 3) & 7); }

// The exchange buffer: n elements of each field (and 3 words per coefficient when reused
// for the carry's transposition).
struct Xbuf {
    Z2* z2;
    Z1* z1;
};

// Butterfly over the thread's slots for a pass of radix r.
template <class O, bool DIF> MM_HD void bfly(typename O::Z* v, u32 r) {
    if (r == 8) {
        if (DIF) dif8<O>(v); else dit8<O>(v);
    } else if (r == 4) {
        if (DIF) { dif4<O>(v); dif4<O>(v + 4); } else { dit4<O>(v); dit4<O>(v + 4); }
    } else {
#pragma unroll
        for (int g = 0; g < 8; g += 2) dif2<O>(v + g);
    }
}

// Twiddles of pass p for thread t, slots e with m != 0 (DIF: after, DIT: before the butterfly).
template <class O>
__device__ __forceinline__ void twiddle(typename O::Z* v, const typename O::C* tw, u32 t, u32 r, u32 ls, u32 T) {
    if (ls == 0) return;  // stride 1: all twiddles are 1
#pragma unroll
    for (u32 e = 0; e < 8; e++) {
        u32 m = e % r;
        if (m == 0) continue;
        u32 g = (e / r) * T + t, j = g & ((1u << ls) - 1);
        v[e] = O::mul(v[e], tw[j * (r - 1) + (m - 1)]);
    }
}

// Writes the thread's slots of pass p to the exchange buffer at their positions.
__device__ __forceinline__ void put(const Xbuf& X, const Z1* v1, const Z2* v2, u32 t, u32 r, u32 ls, u32 T) {
#pragma unroll
    for (u32 e = 0; e < 8; e++) {
        u32 q = sw(slot_pos(t, e, r, ls, T));
        X.z2[q] = v2[e];
        X.z1[q] = v1[e];
    }
}
__device__ __forceinline__ void get(const Xbuf& X, Z1* v1, Z2* v2, u32 t, u32 r, u32 ls, u32 T) {
#pragma unroll
    for (u32 e = 0; e < 8; e++) {
        u32 q = sw(slot_pos(t, e, r, ls, T));
        v2[e] = X.z2[q];
        v1[e] = X.z1[q];
    }
}

// The thread index through an opaque move: address arithmetic derived from it cannot be
// hoisted out of the step loop (where, for every slot and table, it would exhaust registers).
__device__ __forceinline__ u32 opaque_tid() {
    u32 t;
    asm volatile("v_mov_b32 %0, %1" : "=v"(t) : "v"(threadIdx.x));
    return t;
}

// The thread's index within its slot's thread group: the whole workgroup (GT = 0), or groups of
// GT threads (a power of two, whole waves) when several slots share a workgroup.
template <u32 GT> __device__ __forceinline__ u32 slot_tid() {
    const u32 t = opaque_tid();
    return GT ? (t & (GT - 1)) : t;
}

// Pins a value in registers at this point: the compiler may not sink its computation past
// later barriers (where it would keep the larger inputs live instead).
__device__ __forceinline__ void pin(u32& v) { asm volatile("" : "+v"(v)); }
__device__ __forceinline__ void pin(u64& v) { asm volatile("" : "+v"(v)); }

// Scheduling fence: keeps the compiler from interleaving phases (and hoisting their loads),
// which would multiply live registers across the unrolled transform.
#define MM_FENCE() __builtin_amdgcn_sched_barrier(0)
// Loops over a thread's elements that read constants from memory fence after every
// MM_GROUP elements: otherwise the scheduler issues all eight elements' loads at once and
// holds their registers until use. With two workgroups per CU, other waves hide the latency.
#ifndef MM_GROUP
#define MM_GROUP 2
#endif
#define MM_GROUP_FENCE(e)                              \
    do {                                               \
        if (((e) + 1) % MM_GROUP == 0) MM_FENCE();     \
    } while (0)

// The pass plan of length 2^LOGN at compile time (the same plan as make_plan on the host).
template <u32 LOGN> struct CPlan {
    static constexpr u32 n = 1u << LOGN, T = n / 8, b = LOGN % 3;
    static constexpr u32 npass = (b ? 1 : 0) + LOGN / 3;
    static constexpr u32 radix(u32 p) { return (b && p == 0) ? (1u << b) : 8; }
    static constexpr u32 lstride(u32 p) { return p == 0 ? LOGN - (b ? b : 3) : lstride(p - 1) - 3; }
    static constexpr u32 tw_off(u32 p) { return p == 0 ? 0 : tw_off(p - 1) + (1u << lstride(p - 1)) * (radix(p - 1) - 1); }
};

// Largest exchange tile in elements: with TILED (global mode) the exchange buffer holds at
// most 2,048 elements of both fields (48 KiB; two workgroups per CU up to n = 4,096), and a
// longer vector is exchanged in chunks. At n = 8,192 the 1,024-thread workgroup has a CU to
// itself, so its tile is 4,096 elements (96 KiB): two chunks, like n = 4,096, where four would
// hold twice the registers in flight.
//
// TILED selects the exchange buffer: 0 holds the whole vector (LDS-resident layout), 1 the
// default tile above, and a larger power of two that many elements (the autotuner's choice:
// one 4,096-element tile at n = 4,096 exchanges in one chunk at one workgroup per CU).
constexpr u32 tile_max(u32 logn) { return logn > 12 ? 4096 : 2048; }
template <u32 LOGN, u32 TILED> struct Tiling {
    static constexpr u32 n = 1u << LOGN;
    static constexpr u32 want = TILED == 0 ? n : TILED == 1 ? tile_max(LOGN) : TILED;
    static constexpr u32 tile = n > want ? want : n;
    static constexpr u32 chunks = n / tile;
    static_assert(TILED <= 1 || (TILED & (TILED - 1)) == 0, "a tile is a power of two");
};

// Moves the thread's slots from pass (ro, lso)'s layout to pass (ri, lsi)'s through the
// exchange buffer, one tile-sized chunk of positions at a time. With several chunks the inputs
// land in separate registers until every chunk's outputs are written (a thread's input slot
// can lie in an earlier chunk than the output slot held in the same register). Starts and
// ends with the barriers that order it against the buffer's other uses.
template <u32 LOGN, u32 TILED>
__device__ __forceinline__ void exchange(u32 t, bool act, Z1* v1, Z2* v2, const Xbuf& X, u32 ro, u32 lso, u32 ri, u32 lsi) {
    typedef Tiling<LOGN, TILED> G;
    constexpr u32 T = CPlan<LOGN>::T;
    if constexpr (G::chunks == 1) {
        __syncthreads();
        if (act) put(X, v1, v2, t, ro, lso, T);
        __syncthreads();
        if (act) get(X, v1, v2, t, ri, lsi, T);
    } else {
        Z1 n1[8];
        Z2 n2[8];
#pragma unroll
        for (u32 c = 0; c < G::chunks; c++) {
            __syncthreads();
            if (act) {
#pragma unroll
                for (u32 e = 0; e < 8; e++) {
                    u32 pos = slot_pos(t, e, ro, lso, T);
                    if (pos / G::tile == c) {
                        u32 q = sw(pos % G::tile);
                        X.z2[q] = v2[e];
                        X.z1[q] = v1[e];
                    }
                }
            }
            __syncthreads();
            if (act) {
#pragma unroll
                for (u32 e = 0; e < 8; e++) {
                    u32 pos = slot_pos(t, e, ri, lsi, T);
                    if (pos / G::tile == c) {
                        u32 q = sw(pos % G::tile);
                        n2[e] = X.z2[q];
                        n1[e] = X.z1[q];
                    }
                }
            }
        }
#pragma unroll
        for (u32 e = 0; e < 8; e++) {
            v1[e] = n1[e];
            v2[e] = n2[e];
        }
    }
}

// ---- Exchanges inside one wave ------------------------------------------------------------
// Between two radix-8 passes of stride 64 or less, the 64 threads of a wave hold exactly the 512
// positions [512 w, 512 w + 512) of wave w before and after (slot_pos of lane t & 63 plus 512 w).
// Such an exchange needs no workgroup barrier: each wave moves its block through its own 4 KiB
// of the exchange buffer, in three rounds of one 8-byte word per element (the M61 real parts,
// the M61 imaginary parts, the M31 pairs), so that no input has to be held beside its output.
// The wave's LDS instructions execute in order; the compiler is kept from reordering them
// across a round by a wave-scope fence.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// Word index of block position p (< 512, base-8 digits a2 a1 a0) in the wave's region: a1 ^= a2,
// a0 ^= a1. Each of the block's three layouts (strides 64, 8 and 1) then puts the 32 lanes of a
// half-wave 8-byte access on distinct bank pairs.
__device__ __forceinline__ u32 wsw(u32 p) { return p ^ ((p >> 3) & 7) ^ (((p >> 6) & 7) << 3); }
// True when the exchange from pass layout (ro, lso) to (ri, lsi) stays inside every wave.
__host__ __device__ constexpr bool wave_local(u32 ro, u32 lso, u32 ri, u32 lsi) { return ro == 8 && ri == 8 && lso <= 6 && lsi <= 6; }

__device__ __forceinline__ void exchange_wave(u32 t, bool act, Z1* v1, Z2* v2, const Xbuf& X, u32 lso, u32 lsi) {
    if (!act) return;
    const u32 l = t & 63;
    u64* R = (u64*)X.z2 + 512 * (t >> 6);
    u32 qo[8], qi[8];
#pragma unroll
    for (u32 e = 0; e < 8; e++) {
        qo[e] = wsw(slot_pos(l, e, 8, lso, 64));
        qi[e] = wsw(slot_pos(l, e, 8, lsi, 64));
    }
#pragma unroll
    for (u32 e = 0; e < 8; e++) R[qo[e]] = v2[e].re;
    wave_sync();
#pragma unroll
    for (u32 e = 0; e < 8; e++) v2[e].re = R[qi[e]];
    wave_sync();
#pragma unroll
    for (u32 e = 0; e < 8; e++) R[qo[e]] = v2[e].im;
    wave_sync();
#pragma unroll
    for (u32 e = 0; e < 8; e++) v2[e].im = R[qi[e]];
    wave_sync();
#pragma unroll
    for (u32 e = 0; e < 8; e++) R[qo[e]] = ((u64)v1[e].im << 32) | v1[e].re;
    wave_sync();
#pragma unroll
    for (u32 e = 0; e < 8; e++) {
        const u64 z = R[qi[e]];
        v1[e] = Z1{(u32)z, (u32)(z >> 32)};
    }
    wave_sync();
}

// One exchange between passes: inside the waves where it can be, else through the tile. A
// wave-local exchange that follows a tile exchange (or starts a transform whose caller asks for
// it) waits at a workgroup barrier first: the tile's last reads alias the waves' regions.
template <u32 LOGN, u32 TILED>
__device__ __forceinline__ void exchange_any(u32 t, bool act, Z1* v1, Z2* v2, const Xbuf& X, u32 ro, u32 lso, u32 ri, u32 lsi,
                                             bool& synced) {
    if (wave_local(ro, lso, ri, lsi)) {
        if (!synced) __syncthreads();
        synced = true;
        exchange_wave(t, act, v1, v2, X, lso, lsi);
    } else {
        exchange<LOGN, TILED>(t, act, v1, v2, X, ro, lso, ri, lsi);
        synced = false;
    }
}

// Calls f(std::integral_constant<u32, I>) for I = B .. E-1 in order: the passes of a transform
// as compile-time constants at every length (an unrolled loop over 5 passes is too large for
// the compiler to unroll at n >= 8,192, which left radix and stride as run-time values).
template <u32 B, u32 E, class Fn> __device__ __forceinline__ void static_for(Fn&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<u32, B>{});
        static_for<B + 1, E>(f);
    }
}

// Full DIF on registers laid out for pass 0; ends laid out for the last pass (positions
// 8t .. 8t+7). Threads t >= n / 8 hold no elements but join every barrier, and so do all
// threads when `on` is false (an idle slot group). GT: see slot_tid. SYNC: a barrier before a
// first exchange that is wave-local (false when, since the last barrier, every wave has used
// only its own region of the exchange buffer).
template <u32 LOGN, u32 TILED, u32 GT = 0, bool SYNC = true>
__device__ __forceinline__ void dif_regs(u32 t, Z1* v1, Z2* v2, const F1* tw1, const S2* tw2, const Xbuf& X, bool on = true) {
    typedef CPlan<LOGN> P;
    static_assert(16 * Tiling<LOGN, TILED>::tile >= 64 * P::T, "the wave regions fit the exchange buffer");
    const bool act = on && t < P::T;
    bool synced = !SYNC;
    static_for<0, P::npass>([&](auto pc) {
        constexpr u32 p = decltype(pc)::value;
        constexpr u32 r = P::radix(p), ls = P::lstride(p);
        t = slot_tid<GT>();
        if (act) {
            // The two fields' butterflies are independent. Up to n = 4,096 there is no fence
            // between them, so the scheduler may interleave the GF(M31) work with the GF(M61)
            // butterfly (+0.3 % at n = 4,096); the 1,024-thread lengths keep it (their A/B was
            // -0.4 % at n = 8,192). Nor is there one between the GF(M61) butterfly and its
            // twiddles, whose loads then issue under the butterfly's arithmetic (+1.5 % at
            // n = 4,096, +1.4 % at 8,192).
            bfly<Op1, true>(v1, r);
            twiddle<Op1>(v1, tw1 + P::tw_off(p), t, r, ls, P::T);
            if constexpr (LOGN > 12) MM_FENCE();
            bfly<Op2, true>(v2, r);
            twiddle<Op2>(v2, tw2 + P::tw_off(p), t, r, ls, P::T);
            MM_FENCE();
        }
        if constexpr (p + 1 < P::npass)
            exchange_any<LOGN, TILED>(t, act, v1, v2, X, r, ls, P::radix(p + 1), P::lstride(p + 1), synced);
    });
}

// Full DIT from the last pass's layout back to pass 0's (natural positions). SYNC: as dif_regs.
template <u32 LOGN, u32 TILED, u32 GT = 0, bool SYNC = true>
__device__ __forceinline__ void dit_regs(u32 t, Z1* v1, Z2* v2, const F1* itw1, const S2* itw2, const Xbuf& X, bool on = true) {
    typedef CPlan<LOGN> P;
    const bool act = on && t < P::T;
    bool synced = !SYNC;
    static_for<0, P::npass>([&](auto qc) {
        constexpr u32 p = P::npass - 1 - decltype(qc)::value;
        constexpr u32 r = P::radix(p), ls = P::lstride(p);
        t = slot_tid<GT>();
        if (act) {
            twiddle<Op1>(v1, itw1 + P::tw_off(p), t, r, ls, P::T);
            bfly<Op1, false>(v1, r);
            if constexpr (LOGN > 12) MM_FENCE();  // as in dif_regs
            twiddle<Op2>(v2, itw2 + P::tw_off(p), t, r, ls, P::T);
            bfly<Op2, false>(v2, r);
            MM_FENCE();
        }
        if constexpr (p > 0)
            exchange_any<LOGN, TILED>(t, act, v1, v2, X, r, ls, P::radix(p - 1), P::lstride(p - 1), synced);
    });
}

// ---- Carry over registers ------------------------------------------------------------------
// Thread t holds coefficients c[0..E) of positions E t .. E t + E - 1 (all blockDim threads,
// count = blockDim * E). carry_local sums them into limbs l with the local carry and returns the
// high part (sum c_i 2^(32 i) >> 32 E); carry_fix adds the neighbours' high parts and resolves
// the remaining 0/1 carries across the block: generate/propagate ballots over lanes, then over
// waves. Together they write the limbs of sum c_q 2^(32q) + cin and return the carry out of
// the top. sc: 2 * waves + 2 words.
template <int E, class Coef> __device__ __forceinline__ u64 carry_local(Coef coef, u32* l) {
    u128 acc = 0;
#pragma unroll
    for (int e = 0; e < E; e++) {
        acc += coef(e);
        l[e] = (u32)acc;
        acc >>= 32;
        if ((e & 3) == 3) MM_FENCE();  // coefficient reads four at a time
    }
    return (u64)acc;
}
// NT = 0: the carry spans all blockDim threads. NT > 0 (a multiple of 64): only threads t < NT
// hold coefficients; the other waves join the barriers and write nothing.
template <int E, u32 NT = 0> __device__ __forceinline__ u64 carry_fix(u32 t, u32* l, u64 hi, u64* sc, u64 cin) {
    const u32 nt = NT ? NT : blockDim.x, lane = t & 63, wave = t >> 6, nw = nt >> 6;
    const bool mine = NT == 0 || t < NT;
    const u64 up = __shfl_up(hi, 1, 64);
    __syncthreads();
    if (lane == 63 && mine) sc[wave] = hi;
    __syncthreads();
    u64 k = t == 0 ? cin : (lane == 0 ? sc[wave - 1] : up);
    bool ones = true;
#pragma unroll
    for (int e = 0; e < E; e++) {
        u64 v = (u64)l[e] + k;
        l[e] = (u32)v;
        k = v >> 32;
        ones &= l[e] == 0xffffffffu;
    }
    const u32 owner = nt - 1;
    const bool live = t < owner;
    const u64 g = __ballot(live && k != 0);
    const u64 y = g | __ballot(live && ones);
    const u64 s0 = g + y;
    const bool wgen = s0 < g, wprop = y == ~0ull;
    if (lane == 0 && mine) sc[nw + wave] = (u64)wgen | ((u64)wprop << 1);
    __syncthreads();
    const u64 f = lane < nw ? sc[nw + lane] : 0;
    const u64 gw = __ballot(f & 1), yw = gw | __ballot((f & 2) != 0);
    const u64 cw = ((gw + yw) ^ gw ^ yw) >> wave & 1;
    const u64 s = g + y + cw;
    u64 bit = ((s ^ g ^ y) >> lane) & 1;
#pragma unroll
    for (int e = 0; e < E; e++) {
        u64 v = (u64)l[e] + bit;
        l[e] = (u32)v;
        bit = v >> 32;
    }
    if (t == owner) sc[2 * nw] = hi + k + bit;
    __syncthreads();
    return sc[2 * nw];
}
template <int E, u32 NT = 0, class Coef> __device__ __forceinline__ u64 carry_regs(u32 t, Coef coef, u32* l, u64* sc, u64 cin) {
    const u64 hi = carry_local<E>(coef, l);
    return carry_fix<E, NT>(t, l, hi, sc, cin);
}

// Coefficient staging for the transposition: 3 words per coefficient in the exchange
// buffer's space: two words in the M61 half (16n bytes) and one in the M31 half (8n bytes),
// 2n coefficients in all.
struct Cbuf {
    u32 *w0, *w1, *w2;
};
__device__ __forceinline__ Cbuf cbuf(const Xbuf& X, u32 n) {
    u32* b2 = (u32*)X.z2;
    return Cbuf{b2, b2 + 2 * n, (u32*)X.z1};
}
__device__ __forceinline__ void cput(const Cbuf& C, u32 q, u128 v) {
    C.w0[q] = (u32)v;
    C.w1[q] = (u32)(v >> 32);
    C.w2[q] = (u32)(v >> 64);
}
__device__ __forceinline__ u128 cget(const Cbuf& C, u32 q) {
    return (u128)C.w0[q] | ((u128)C.w1[q] << 32) | ((u128)C.w2[q] << 64);
}

__device__ __forceinline__ u128 crt_lazy(u32 c1, u64 c2, u32 inv) { return crt(canon1(c1), canon2(c2), inv); }

}  // namespace v2
}  // namespace mm
