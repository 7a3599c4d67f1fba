// mm_field.h - exact arithmetic in GF(M31^2) and GF(M61^2) (i^2 = -1), for the host (table
// generation) and the gfx950 device (transforms).
//
// The fields and their roots are fixed by the reference (montmul/src/field.rs:7-11, 111-120,
// 237-267): P1 = 2^31 - 1, P2 = 2^61 - 1, both = 3 mod 4, so GF(p^2) = GF(p)[i] and the 2-Sylow
// subgroup has order 2^32 (M31) and 2^62 (M61). The same roots give the same spectra, so the
// device spectra match the reference's CPU spectra element for element.
//
// Every function returns canonical residues (< p) from canonical inputs. On gfx950 every
// VALU op in these sequences (v_mad_u64_u32 included) issues at the same rate, so the
// functions minimise instruction count: products of canonical operands are summed before a
// single reduction wherever the sum stays below 2^63 (M31) or 2^125 (M61).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define MM_HD __host__ __device__ __forceinline__

namespace mm {

typedef unsigned __int128 u128;

constexpr uint32_t P1 = 0x7fffffffu;
constexpr uint64_t P2 = 0x1fffffffffffffffull;

struct F1 {
    uint32_t re, im;
};
struct alignas(16) F2 {
    uint64_t re, im;
};

// ---- GF(M31) ----------------------------------------------------------------------------
MM_HD uint32_t r1(uint32_t x) { return x >= P1 ? x - P1 : x; }  // x < 2 P1 + 1 (wraps below)
MM_HD uint32_t add1(uint32_t a, uint32_t b) { return r1(a + b); }
MM_HD uint32_t sub1(uint32_t a, uint32_t b) { return a >= b ? a - b : a - b + P1; }
MM_HD uint32_t neg1(uint32_t a) { return a ? P1 - a : 0; }
// t mod P1 for t < 2^63: t = hi * 2^31 + lo with 2^31 = 1, and hi < 2^32 folds once more.
MM_HD uint32_t red1(uint64_t t) {
    uint32_t lo = (uint32_t)t & P1, hi = (uint32_t)(t >> 31);
    uint32_t s = lo + (hi & P1) + (hi >> 31);  // <= 2^32 - 1
    return r1(r1(s));
}
MM_HD uint32_t mul1(uint32_t a, uint32_t b) { return red1((uint64_t)a * b); }
// x mod P1 for any 32-bit x.
MM_HD uint32_t mod1(uint32_t x) { return r1((x & P1) + (x >> 31)); }

// ---- GF(M61) ----------------------------------------------------------------------------
MM_HD uint64_t r2(uint64_t x) { return x >= P2 ? x - P2 : x; }
MM_HD uint64_t add2(uint64_t a, uint64_t b) { return r2(a + b); }
MM_HD uint64_t sub2(uint64_t a, uint64_t b) { return a >= b ? a - b : a - b + P2; }
MM_HD uint64_t neg2(uint64_t a) { return a ? P2 - a : 0; }
// t mod P2 for t < 2^125: t = hi * 2^61 + lo with 2^61 = 1; hi < 2^64 folds once more.
MM_HD uint64_t red2(u128 t) {
    uint64_t lo = (uint64_t)t & P2, hi = (uint64_t)(t >> 61);
    uint64_t s = lo + (hi & P2) + (hi >> 61);  // < 2^62 + 8
    return r2(r2(s));
}
MM_HD u128 wmul(uint64_t a, uint64_t b) { return (u128)a * b; }
MM_HD uint64_t mul2(uint64_t a, uint64_t b) { return red2(wmul(a, b)); }

// ---- GF(p^2) ----------------------------------------------------------------------------
MM_HD F1 cadd(F1 a, F1 b) { return F1{add1(a.re, b.re), add1(a.im, b.im)}; }
MM_HD F1 csub(F1 a, F1 b) { return F1{sub1(a.re, b.re), sub1(a.im, b.im)}; }
MM_HD F2 cadd(F2 a, F2 b) { return F2{add2(a.re, b.re), add2(a.im, b.im)}; }
MM_HD F2 csub(F2 a, F2 b) { return F2{sub2(a.re, b.re), sub2(a.im, b.im)}; }
// a * b: re = a.re b.re + a.im (p - b.im), im = a.re b.im + a.im b.re, one reduction each.
MM_HD F1 cmul(F1 a, F1 b) {
    uint64_t re = (uint64_t)a.re * b.re + (uint64_t)a.im * (P1 - b.im);
    uint64_t im = (uint64_t)a.re * b.im + (uint64_t)a.im * b.re;
    return F1{red1(re), red1(im)};
}
MM_HD F2 cmul(F2 a, F2 b) {
    u128 re = wmul(a.re, b.re) + wmul(a.im, P2 - b.im);
    u128 im = wmul(a.re, b.im) + wmul(a.im, b.re);
    return F2{red2(re), red2(im)};
}
MM_HD F1 csqr(F1 a) {
    uint64_t re = (uint64_t)a.re * a.re + (uint64_t)a.im * (P1 - a.im);
    uint64_t im = 2 * ((uint64_t)a.re * a.im);
    return F1{red1(re), red1(im)};
}
MM_HD F2 csqr(F2 a) {
    u128 re = wmul(a.re, a.re) + wmul(a.im, P2 - a.im);
    u128 im = 2 * wmul(a.re, a.im);
    return F2{red2(re), red2(im)};
}
// Multiplication by a small scalar c <= 4 (the base).
MM_HD F1 cscale_small(F1 a, uint32_t c) { return F1{red1((uint64_t)a.re * c), red1((uint64_t)a.im * c)}; }
MM_HD F2 cscale_small(F2 a, uint32_t c) { return F2{red2((u128)a.re * c), red2((u128)a.im * c)}; }
MM_HD F1 cconj(F1 a) { return F1{a.re, neg1(a.im)}; }
MM_HD F2 cconj(F2 a) { return F2{a.re, neg2(a.im)}; }
// a * i and a * (-i).
MM_HD F1 cmul_i(F1 a) { return F1{neg1(a.im), a.re}; }
MM_HD F2 cmul_i(F2 a) { return F2{neg2(a.im), a.re}; }
MM_HD F1 cmul_mi(F1 a) { return F1{a.im, neg1(a.re)}; }
MM_HD F2 cmul_mi(F2 a) { return F2{a.im, neg2(a.re)}; }
// a + i b and a - i b.
MM_HD F1 cadd_i(F1 a, F1 b) { return F1{sub1(a.re, b.im), add1(a.im, b.re)}; }
MM_HD F1 csub_i(F1 a, F1 b) { return F1{add1(a.re, b.im), sub1(a.im, b.re)}; }
MM_HD F2 cadd_i(F2 a, F2 b) { return F2{sub2(a.re, b.im), add2(a.im, b.re)}; }
MM_HD F2 csub_i(F2 a, F2 b) { return F2{add2(a.re, b.im), sub2(a.im, b.re)}; }

MM_HD bool ceq(F1 a, F1 b) { return a.re == b.re && a.im == b.im; }
MM_HD bool ceq(F2 a, F2 b) { return a.re == b.re && a.im == b.im; }

// ---- CRT --------------------------------------------------------------------------------
// The unique C < P1 * P2 with C = c1 (mod P1) and C = c2 (mod P2), for canonical c1, c2:
// C = c2 + P2 * ((c1 - c2) * P2^-1 mod P1). `inv` = (P2 mod P1)^-1 mod P1.
MM_HD u128 crt(uint32_t c1, uint64_t c2, uint32_t inv) {
    // c2 mod P1: 2^31 = 1, so fold the 61-bit value in 31-bit pieces.
    uint32_t m = ((uint32_t)c2 & P1) + ((uint32_t)(c2 >> 31) & P1) + (uint32_t)(c2 >> 62);
    m = r1(r1(m));
    uint32_t h = mul1(sub1(c1, m), inv);
    return (u128)c2 + (((u128)h << 61) - h);
}

}  // namespace mm
