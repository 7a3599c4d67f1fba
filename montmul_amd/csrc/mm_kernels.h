// mm_kernels.h - device code of the batched Montgomery step for gfx950.
//
// One workgroup owns one slot (one modulus) and runs whole steps
//     x <- x^2 * c^bit * R^-1 mod N   (lazy range [0, 2N), R = 2^(32k)),
// reading its exponent bits on the device. A step is (montmul/kernels/src/lib.rs:859-881,
// 790-857 for the reference's version):
//   1. T = x^2 c by a weighted ("right-angle") transform of length n = L/2 over
//      GF(M31^2) x GF(M61^2): coefficient j in the real part and j + n in the imaginary part
//      of element j, weights zeta^j with zeta^n = i, so the cyclic length-n transform computes
//      products modulo X^(2n) + 1. CRT gives exact coefficients < M31 * M61; a block-wide
//      carry turns them into the 2k limbs of T.
//   2. m = (T mod R) N' mod R: the same transform of T's low k limbs, times the spectrum of
//      N' (computed once per batch), inverse, carry of k coefficients.
//   3. With the wraparound correction (k = L/2): the cyclic product m N mod (R - 1) through
//      complex transforms of half length h = n/2 of the packed limbs z_j = m_2j + i m_2j+1,
//      with per-slot factors A | B of N, and x = floor(T / R) + ((cyclic + T mod R) mod (R-1)),
//      an end-around carry, and a check that the sum fits k limbs. Without it: S = m N by
//      full-length transforms and x = floor(T / R) + floor(S / R) + [T mod R != 0].
//
// Transforms are in-place radix-4 passes: DIF (natural order in, bit-reversed out) and DIT
// (bit-reversed in, natural out). A pass reads four elements at distance q, applies one
// radix-4 butterfly with the pass twiddles (W^j, W^2j, W^3j), and writes them back; q
// shrinks by 4 per DIF pass and grows by 4 per DIT pass. An odd log2 n adds one radix-2 stage
// with unit twiddles: the last DIF stage, the first DIT stage.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mm_field.h"

namespace mm {

// Per-length device tables of one field pair (built on the host, mm_tables.h).
struct Tab {
    const F1 *w1, *iw1;    // zeta^j and zeta^-j / n, j < n
    const F2 *w2, *iw2;
    const F1 *dtw1;        // DIF pass twiddle triples (W^j, W^2j, W^3j), passes in order
    const F2 *dtw2;
    const F1 *itw1;        // DIT pass twiddle triples with the inverse roots
    const F2 *itw2;
    uint32_t n, logn;
    uint32_t i_is_plus;    // omega^(n/4) == +i (1) or -i (0)
    uint32_t crt_inv;
};

// Tables of the half-length path (length h = n/2) plus its unpack roots.
struct HalfTab {
    Tab t;                 // transforms of length h (weights unused)
    const F1 *u1, *iu1;    // w^rev(p), w^-rev(p), w a primitive 2h-th root, p < h
    const F2 *u2, *iu2;
};

// Per-slot state in device memory, structure of arrays over slots.
struct Slots {
    uint32_t *x;           // k limbs per slot
    const F1 *nphat1;      // n elements per slot: spectrum of N'
    const F2 *nphat2;
    const F1 *nhat1;       // n elements per slot: factors A | B (wraparound) or spectrum of N
    const F2 *nhat2;
    F1 *work1;             // n elements per slot (global mode only)
    F2 *work2;
    uint32_t *lim;         // global mode: 3k limbs per slot (T, then m or S's top half)
    const uint32_t *exp;   // ew words per slot, most significant bit first
    const uint32_t *nbits;
    uint32_t *pos;
    uint32_t *err;         // set to 1 when a carry that must vanish does not
    uint32_t k, ew, base, slots;
};

// ---- Small helpers ------------------------------------------------------------------------
__device__ __forceinline__ uint32_t tid() { return threadIdx.x; }
__device__ __forceinline__ uint32_t nthreads() { return blockDim.x; }

// Thread geometry of the radix-4 step: the whole workgroup steps one slot, or (slot groups in
// the mixed-length launch) each wave of 64 threads steps its own slot. Barriers stay
// workgroup-wide, so every group runs the same sequence of them.
struct WholeBlock {
    static constexpr bool grouped = false;
    static __device__ __forceinline__ uint32_t t() { return threadIdx.x; }
    static __device__ __forceinline__ uint32_t n() { return blockDim.x; }
};
// The lane id (mbcnt over a full exec mask) equals threadIdx.x & 63 for whole-wave groups and
// needs no work-item ID. An earlier, heavily spilling build of the 1,024-thread mixed-length
// kernel stepped wrongly with threadIdx.x & 63 here and exactly with the lane id; builds of the
// present source are exact with either.
struct OneWave {
    static constexpr bool grouped = true;
    static __device__ __forceinline__ uint32_t t() { return __lane_id(); }
    static __device__ __forceinline__ uint32_t n() { return 64; }
};

template <class F> struct FieldOf;
template <> struct FieldOf<F1> { typedef uint32_t B; };
template <> struct FieldOf<F2> { typedef uint64_t B; };

// Multiplication by the fourth root of unity omega^(n/4) (= +i or -i) and its inverse.
template <class F> __device__ __forceinline__ F mul_i4(F a, bool plus) { return plus ? cmul_i(a) : cmul_mi(a); }

// ---- Radix-4 passes -----------------------------------------------------------------------
// One DIF radix-4 butterfly group: elements i0 + {0, q, 2q, 3q}, twiddles W^j, W^2j, W^3j.
template <class F>
__device__ __forceinline__ void dif4(F* a, uint32_t i0, uint32_t q, const F* tw, bool unit, bool plus) {
    F x0 = a[i0], x1 = a[i0 + q], x2 = a[i0 + 2 * q], x3 = a[i0 + 3 * q];
    F y0 = cadd(x0, x2), y1 = cadd(x1, x3), y2 = csub(x0, x2), y3 = mul_i4(csub(x1, x3), plus);
    F b0 = cadd(y0, y1), b1 = csub(y0, y1), b2 = cadd(y2, y3), b3 = csub(y2, y3);
    if (!unit) {
        b1 = cmul(b1, tw[1]);
        b2 = cmul(b2, tw[0]);
        b3 = cmul(b3, tw[2]);
    }
    a[i0] = b0;
    a[i0 + q] = b1;
    a[i0 + 2 * q] = b2;
    a[i0 + 3 * q] = b3;
}

// One DIT radix-4 butterfly group: x1 *= V^2j, x2 *= V^j, x3 *= V^3j, then two stages.
template <class F>
__device__ __forceinline__ void dit4(F* a, uint32_t i0, uint32_t q, const F* tw, bool unit, bool plus) {
    F x0 = a[i0], x1 = a[i0 + q], x2 = a[i0 + 2 * q], x3 = a[i0 + 3 * q];
    if (!unit) {
        x1 = cmul(x1, tw[1]);
        x2 = cmul(x2, tw[0]);
        x3 = cmul(x3, tw[2]);
    }
    F y0 = cadd(x0, x1), y1 = csub(x0, x1), y2 = cadd(x2, x3);
    // J = V^q = omega^(-n/4): -i when omega^(n/4) = +i.
    F y3 = mul_i4(csub(x2, x3), !plus);
    a[i0] = cadd(y0, y2);
    a[i0 + 2 * q] = csub(y0, y2);
    a[i0 + q] = cadd(y1, y3);
    a[i0 + 3 * q] = csub(y1, y3);
}

// Number of radix-4 passes and the offset (in triples) of pass p's twiddles.
__host__ __device__ __forceinline__ uint32_t dif_pass_q(uint32_t logn, uint32_t p) { return 1u << (logn - 2 - 2 * p); }
__host__ __device__ __forceinline__ uint32_t dit_pass_q(uint32_t logn, uint32_t p) { return 1u << ((logn & 1) + 2 * p); }

// One radix-4 DIF pass of stride q over a[0, len) (len a multiple of 4q), twiddles from offset
// `off` (in triples).
template <class F, class Geo>
__device__ __forceinline__ void dif_pass(F* a, uint32_t len, uint32_t q, uint32_t lq, const F* tw, bool plus) {
    for (uint32_t b = Geo::t(); b < len / 4; b += Geo::n()) {
        uint32_t j = b & (q - 1);
        uint32_t i0 = ((b >> lq) << (lq + 2)) + j;
        dif4(a, i0, q, tw + 3 * j, q == 1, plus);
    }
}
template <class F, class Geo>
__device__ __forceinline__ void dit_pass(F* a, uint32_t len, uint32_t q, uint32_t lq, const F* tw, bool plus) {
    for (uint32_t b = Geo::t(); b < len / 4; b += Geo::n()) {
        uint32_t j = b & (q - 1);
        uint32_t i0 = ((b >> lq) << (lq + 2)) + j;
        dit4(a, i0, q, tw + 3 * j, q == 1, plus);
    }
}
template <class F, class Geo>
__device__ __forceinline__ void radix2_pass(F* a, uint32_t len) {
    for (uint32_t b = Geo::t(); b < len / 2; b += Geo::n()) {
        F u = a[2 * b], v = a[2 * b + 1];
        a[2 * b] = cadd(u, v);
        a[2 * b + 1] = csub(u, v);
    }
}

// Full DIF over a (length t.n). Ends with a barrier.
//
// With a tile (m elements of LDS, a power of two, a in global memory: lib.rs:455-511), the passes
// whose butterfly groups span more than m elements run in place, and the rest, which act within
// aligned blocks of at most m, run block after block in the tile: one global round trip for
// them instead of one per pass. Each thread copies back exactly the elements it copied in, so
// the next block's copy-in needs no barrier. A pass's twiddles depend only on the position in
// its butterfly group, so a block uses the same twiddle rows as the whole vector.
template <class F, class Geo = WholeBlock>
__device__ void dif(F* a, const Tab& t, const F* dtw, F* tile = nullptr, uint32_t m = 0) {
    if (m == 0) tile = nullptr;  // no tile
    const uint32_t n = t.n, logn = t.logn, npass = logn / 2;
    const bool plus = t.i_is_plus;
    uint32_t off = 0, p = 0;
    for (; p < npass; p++) {
        const uint32_t q = dif_pass_q(logn, p), lq = logn - 2 - 2 * p;
        if (tile && 4 * q <= m) break;
        __syncthreads();
        dif_pass<F, Geo>(a, n, q, lq, dtw + 3 * off, plus);
        off += q;
    }
    if (tile && (p < npass || (logn & 1))) {
        const uint32_t len = m < n ? m : n;
        for (uint32_t c = 0; c < n; c += len) {
            if (c == 0) __syncthreads();  // the in-place passes are done
            for (uint32_t j = Geo::t(); j < len; j += Geo::n()) tile[j] = a[c + j];
            uint32_t o = off;
            for (uint32_t pp = p; pp < npass; pp++) {
                const uint32_t q = dif_pass_q(logn, pp), lq = logn - 2 - 2 * pp;
                __syncthreads();
                dif_pass<F, Geo>(tile, len, q, lq, dtw + 3 * o, plus);
                o += q;
            }
            if (logn & 1) {
                __syncthreads();
                radix2_pass<F, Geo>(tile, len);
            }
            __syncthreads();
            for (uint32_t j = Geo::t(); j < len; j += Geo::n()) a[c + j] = tile[j];
        }
    } else if (logn & 1) {
        __syncthreads();
        radix2_pass<F, Geo>(a, n);
    }
    __syncthreads();
}

// Full DIT over a (length t.n), unscaled. Ends with a barrier. With a tile: the mirror image
// of dif's, the passes within blocks of at most m first, block after block in the tile, then
// the wider passes in place.
template <class F, class Geo = WholeBlock>
__device__ void dit(F* a, const Tab& t, const F* itw, F* tile = nullptr, uint32_t m = 0) {
    if (m == 0) tile = nullptr;  // no tile
    const uint32_t n = t.n, logn = t.logn, npass = logn / 2;
    const bool plus = t.i_is_plus;
    uint32_t off = 0, p = 0;
    if (tile) {
        const uint32_t len = m < n ? m : n;
        uint32_t np = 0;  // passes that act within blocks of len
        while (np < npass && 4 * dit_pass_q(logn, np) <= len) np++;
        for (uint32_t c = 0; c < n; c += len) {
            if (c == 0) __syncthreads();  // the vector's last writes (by any thread) are done
            for (uint32_t j = Geo::t(); j < len; j += Geo::n()) tile[j] = a[c + j];
            if (logn & 1) {
                __syncthreads();
                radix2_pass<F, Geo>(tile, len);
            }
            uint32_t o = 0;
            for (uint32_t pp = 0; pp < np; pp++) {
                const uint32_t q = dit_pass_q(logn, pp), lq = (logn & 1) + 2 * pp;
                __syncthreads();
                dit_pass<F, Geo>(tile, len, q, lq, itw + 3 * o, plus);
                o += q;
            }
            __syncthreads();
            for (uint32_t j = Geo::t(); j < len; j += Geo::n()) a[c + j] = tile[j];
        }
        p = np;
        for (uint32_t pp = 0; pp < np; pp++) off += dit_pass_q(logn, pp);
    } else if (logn & 1) {
        __syncthreads();
        radix2_pass<F, Geo>(a, n);
    }
    for (; p < npass; p++) {
        const uint32_t q = dit_pass_q(logn, p), lq = (logn & 1) + 2 * p;
        __syncthreads();
        dit_pass<F, Geo>(a, n, q, lq, itw + 3 * off, plus);
        off += q;
    }
    __syncthreads();
}

// ---- Whole-vector element passes ------------------------------------------------------------
// a_j = (lim_j, lim_{j+n}) * zeta^j for the first cnt coefficients (zero beyond), both fields.
template <class Geo = WholeBlock>
__device__ __forceinline__ void load_weighted(F1* a1, F2* a2, const uint32_t* lim, uint32_t cnt, const Tab& t) {
    const uint32_t n = t.n;
    for (uint32_t j = Geo::t(); j < n; j += Geo::n()) {
        uint32_t re = j < cnt ? lim[j] : 0, im = j + n < cnt ? lim[j + n] : 0;
        a1[j] = cmul(F1{mod1(re), mod1(im)}, t.w1[j]);
        a2[j] = cmul(F2{(uint64_t)re, (uint64_t)im}, t.w2[j]);
    }
}

template <class Geo = WholeBlock>
__device__ __forceinline__ void unweight(F1* a1, F2* a2, const Tab& t) {
    for (uint32_t j = Geo::t(); j < t.n; j += Geo::n()) {
        a1[j] = cmul(a1[j], t.iw1[j]);
        a2[j] = cmul(a2[j], t.iw2[j]);
    }
}

// ---- Block-wide carry ---------------------------------------------------------------------
// Writes the limbs of sum_q coef(q) 2^(32q), q < count, through store(q, limb) and returns the
// carry out of coefficient count - 1 (plus cin, added at q = 0). coef(q) < 2^96.
//
// Each thread converts E consecutive coefficients of a chunk of blockDim * E into E limbs and
// a high part H < 2^63, then adds its left neighbour's H (wave shuffle, or LDS across waves).
// What remains between threads is a 0/1 carry per thread, resolved by carry lookahead:
// generate/propagate ballots over the wave (64 lanes in one 64-bit add), then over the waves.
// `sc` is LDS scratch of 2 * (blockDim / 64) + 2 words. E >= 2 keeps the neighbour carry's
// effect to 0 or 1 for every thread except the one owning coefficient count - 1.
template <int E, class Coef, class Store, class Geo = WholeBlock>
__device__ uint64_t block_carry(Coef coef, uint32_t count, Store store, uint64_t* sc, uint64_t cin) {
    const uint32_t t = Geo::t(), nt = Geo::n(), lane = t & 63, wave = t >> 6, nw = nt >> 6;
    uint64_t chunk_in = cin;
    uint32_t parity = 0;
    for (uint32_t base = 0; base < count; base += nt * E) {
        const uint32_t q0 = base + t * E;
        const uint32_t owner = min((count - 1 - base) / E, nt - 1);
        uint32_t l[E];
        u128 acc = 0;
#pragma unroll
        for (int e = 0; e < E; e++) {
            if (q0 + e < count) {
                acc += coef(q0 + e);
                l[e] = (uint32_t)acc;
                acc >>= 32;
            } else {
                l[e] = 0;
            }
        }
        uint64_t hi = (uint64_t)acc;
        // Neighbour carry: lane - 1 of this wave, or lane 63 of the previous wave.
        uint64_t up = __shfl_up(hi, 1, 64);
        if (lane == 63) sc[wave] = hi;
        __syncthreads();
        uint64_t c = t == 0 ? chunk_in : (lane == 0 ? sc[wave - 1] : up);
        bool ones = true;
#pragma unroll
        for (int e = 0; e < E; e++) {
            if (q0 + e < count) {
                uint64_t v = (uint64_t)l[e] + c;
                l[e] = (uint32_t)v;
                c = v >> 32;
                ones &= l[e] == 0xffffffffu;
            }
        }
        // c is now this thread's 0/1 carry out (any value for the owner).
        const bool live = t < owner;
        const uint64_t g = __ballot(live && c != 0);
        const uint64_t y = g | __ballot(live && ones);
        const uint64_t s0 = g + y;
        const bool wgen = s0 < g;            // carry out of lane 63 with no carry in
        const bool wprop = y == ~0ull;       // a carry into lane 0 would leave lane 63
        if (lane == 0) sc[nw + wave] = (uint64_t)wgen | ((uint64_t)wprop << 1);
        __syncthreads();
        const uint64_t f = lane < nw ? sc[nw + lane] : 0;
        const uint64_t gw = __ballot(f & 1), yw = gw | __ballot((f & 2) != 0);
        const uint64_t cw = ((gw + yw) ^ gw ^ yw) >> wave & 1;  // carry into this wave
        const uint64_t s = g + y + cw;
        uint64_t bit = ((s ^ g ^ y) >> lane) & 1;                // carry into this lane
        if (t > owner) bit = 0;
#pragma unroll
        for (int e = 0; e < E; e++) {
            if (q0 + e < count) {
                uint64_t v = (uint64_t)l[e] + bit;
                l[e] = (uint32_t)v;
                bit = v >> 32;
                store(q0 + e, l[e]);
            }
        }
        if (t == owner) sc[2 * nw + parity] = hi + c + bit;
        __syncthreads();
        chunk_in = sc[2 * nw + parity];
        parity ^= 1;
    }
    return chunk_in;
}

// Coefficient q of a product held in (a1, a2): re of element q for q < n, im of q - n.
struct ProductCoef {
    const F1* a1;
    const F2* a2;
    uint32_t n, inv;
    __device__ __forceinline__ u128 operator()(uint32_t q) const {
        bool hi = q >= n;
        uint32_t j = hi ? q - n : q;
        F1 e1 = a1[j];
        F2 e2 = a2[j];
        return hi ? crt(e1.im, e2.im, inv) : crt(e1.re, e2.re, inv);
    }
};

struct StoreLimbs {
    uint32_t* out;
    __device__ __forceinline__ void operator()(uint32_t q, uint32_t v) const { out[q] = v; }
};

// ---- Wraparound pieces --------------------------------------------------------------------
// In a bit-reversed half spectrum of length h, the slot holding the conjugate-symmetric index
// of the one at slot p: rev((h - rev(p)) mod h). It stays inside p's octave [2^t, 2^(t+1)) and
// mirrors p there, which is 3 * 2^t - 1 - p, independent of h (lib.rs:530-541).
__host__ __device__ __forceinline__ uint32_t partner(uint32_t p) {
    if (p == 0) return 0;
    uint32_t top = 31 - __builtin_clz(p);
    return (3u << top) - 1 - p;
}

}  // namespace mm
