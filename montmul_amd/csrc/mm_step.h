// mm_step.h - the fused Montgomery step and its setup kernels (device side).
#pragma once

#include "mm_kernels.h"

namespace mm {

// Carry-lookahead chunk width: coefficients per thread.
constexpr int CARRY_E = 8;

// Vector storage of one workgroup: a1/a2 are n elements in LDS (LDS mode) or in the slot's
// HBM work buffer (global mode). `lt` holds T's 2k limbs, `ls` k limbs of m or of S's top
// half; `sc` is the carry scratch (LDS in both modes).
struct Work {
    F1* a1;
    F2* a2;
    uint32_t* lt;
    uint32_t* ls;
    uint64_t* sc;
    // Global mode: an LDS tile of tile_m elements (16 tile_m bytes, one field at a time) for the
    // transforms' narrow passes (dif/dit); none when tile_m is 0.
    unsigned char* tile = nullptr;
    uint32_t tile_m = 0;
    __device__ __forceinline__ F1* tile1() const { return (F1*)tile; }
    __device__ __forceinline__ F2* tile2() const { return (F2*)tile; }
};

// LDS layout of LDS mode: F2[n] | F1[n] | T: u32[2k] | u32[k] | scratch. Global mode keeps
// only the scratch in LDS.
__host__ __device__ inline uint32_t lds_scratch_bytes(uint32_t threads) { return 8 * (2 * (threads / 64) + 2); }
__host__ __device__ inline uint64_t lds_bytes_resident(uint32_t n, uint32_t k, uint32_t threads) {
    return ((24ull * n + 12ull * k + 7) & ~7ull) + lds_scratch_bytes(threads);
}

// Squaring: T = c x^2 into w.lt (2k limbs); returns the carry out of 2k limbs.
template <class Geo = WholeBlock>
__device__ uint64_t square(const Work& w, const uint32_t* x, uint32_t k, uint32_t c, const Tab& t) {
    load_weighted<Geo>(w.a1, w.a2, x, k, t);
    dif<F1, Geo>(w.a1, t, t.dtw1, w.tile1(), w.tile_m);
    dif<F2, Geo>(w.a2, t, t.dtw2, w.tile2(), w.tile_m);
    for (uint32_t j = Geo::t(); j < t.n; j += Geo::n()) {
        F1 v1 = csqr(w.a1[j]);
        F2 v2 = csqr(w.a2[j]);
        if (c != 1) {
            v1 = cscale_small(v1, c);
            v2 = cscale_small(v2, c);
        }
        w.a1[j] = v1;
        w.a2[j] = v2;
    }
    dit<F1, Geo>(w.a1, t, t.itw1, w.tile1(), w.tile_m);
    dit<F2, Geo>(w.a2, t, t.itw2, w.tile2(), w.tile_m);
    unweight<Geo>(w.a1, w.a2, t);
    __syncthreads();
    return block_carry<CARRY_E, ProductCoef, StoreLimbs, Geo>(ProductCoef{w.a1, w.a2, t.n, t.crt_inv}, 2 * k, StoreLimbs{w.lt}, w.sc, 0);
}

// a <- a (.) b elementwise, b in device memory.
template <class Geo = WholeBlock>
__device__ __forceinline__ void pointwise_const(const Work& w, const F1* b1, const F2* b2, uint32_t n) {
    for (uint32_t j = Geo::t(); j < n; j += Geo::n()) {
        w.a1[j] = cmul(w.a1[j], b1[j]);
        w.a2[j] = cmul(w.a2[j], b2[j]);
    }
    __syncthreads();
}

// Stores limb q of m as the packed half-length element z_{q/2} (re for even q, im for odd).
struct StorePacked {
    F1* a1;
    F2* a2;
    __device__ __forceinline__ void operator()(uint32_t q, uint32_t v) const {
        uint32_t j = q >> 1;
        if (q & 1) {
            a1[j].im = mod1(v);
            a2[j].im = v;
        } else {
            a1[j].re = mod1(v);
            a2[j].re = v;
        }
    }
};

// Coefficient q of the packed cyclic product plus T's limb q.
struct CyclicCoef {
    const F1* a1;
    const F2* a2;
    const uint32_t* lt;
    uint32_t inv;
    __device__ __forceinline__ u128 operator()(uint32_t q) const {
        F1 e1 = a1[q >> 1];
        F2 e2 = a2[q >> 1];
        u128 c = (q & 1) ? crt(e1.im, e2.im, inv) : crt(e1.re, e2.re, inv);
        return c + lt[q];
    }
};

struct LimbsCoef {
    const uint32_t* a;
    __device__ __forceinline__ u128 operator()(uint32_t q) const { return a[q]; }
};

struct SumCoef {  // a[q] + b[q]
    const uint32_t* a;
    const uint32_t* b;
    __device__ __forceinline__ u128 operator()(uint32_t q) const { return (u128)a[q] + b[q]; }
};

struct Sum3Coef {  // a[q] + b[q] + (q == 0 ? flag : 0)
    const uint32_t* a;
    const uint32_t* b;
    uint32_t flag;
    __device__ __forceinline__ u128 operator()(uint32_t q) const { return (u128)a[q] + b[q] + (q == 0 ? flag : 0); }
};

// Stores only limbs q >= k, at q - k.
struct StoreTop {
    uint32_t* out;
    uint32_t k;
    __device__ __forceinline__ void operator()(uint32_t q, uint32_t v) const {
        if (q >= k) out[q - k] = v;
    }
};

// half_pointwise (out of place): out_p = (z_p + conj z_q) A_p + (z_p - conj z_q) B_p with
// q = partner(p), A at ab[p] and B at ab[h + p]. Ends with a barrier.
template <class F, class Geo = WholeBlock>
__device__ __forceinline__ void half_pointwise(const F* z, F* out, const F* ab, uint32_t h) {
    for (uint32_t p = Geo::t(); p < h; p += Geo::n()) {
        F zp = z[p], zq = cconj(z[partner(p)]);
        F e = cadd(zp, zq), d = csub(zp, zq);
        out[p] = cadd(cmul(e, ab[p]), cmul(d, ab[h + p]));
    }
    __syncthreads();
}

// Reduction (steps 2 and 3): from T in w.lt, writes T R^-1 (lazy) to x. Returns nonzero when
// a carry that must vanish did not.
template <bool WRAP, class Geo = WholeBlock>
__device__ uint32_t redc(const Work& w, uint32_t* x, uint32_t k, const Tab& t, const HalfTab& ht,
                         const F1* nphat1, const F2* nphat2, const F1* nhat1, const F2* nhat2) {
    // Slot groups share the workgroup's barriers: the reduction without wraparound ORs a flag
    // over the whole workgroup, so it runs on whole workgroups only.
    static_assert(WRAP || !Geo::grouped, "slot groups need the wraparound reduction");
    // m = (T mod R) N' mod R
    load_weighted<Geo>(w.a1, w.a2, w.lt, k, t);
    dif<F1, Geo>(w.a1, t, t.dtw1, w.tile1(), w.tile_m);
    dif<F2, Geo>(w.a2, t, t.dtw2, w.tile2(), w.tile_m);
    pointwise_const<Geo>(w, nphat1, nphat2, t.n);
    dit<F1, Geo>(w.a1, t, t.itw1, w.tile1(), w.tile_m);
    dit<F2, Geo>(w.a2, t, t.itw2, w.tile2(), w.tile_m);
    unweight<Geo>(w.a1, w.a2, t);
    __syncthreads();
    if (WRAP) {
        // m goes straight into the packed half-length vector (every read of the product
        // precedes the carry's first barrier; every packed write follows its last).
        block_carry<CARRY_E, ProductCoef, StorePacked, Geo>(ProductCoef{w.a1, w.a2, t.n, t.crt_inv}, k, StorePacked{w.a1, w.a2},
                                                            w.sc, 0);
        // The packed spectrum occupies elements [0, h), the product goes to [h, 2h = n).
        const uint32_t h = ht.t.n;
        dif<F1, Geo>(w.a1, ht.t, ht.t.dtw1, w.tile1(), w.tile_m);
        dif<F2, Geo>(w.a2, ht.t, ht.t.dtw2, w.tile2(), w.tile_m);
        F1* b1 = w.a1 + h;
        F2* b2 = w.a2 + h;
        half_pointwise<F1, Geo>(w.a1, b1, nhat1, h);
        half_pointwise<F2, Geo>(w.a2, b2, nhat2, h);
        dit<F1, Geo>(b1, ht.t, ht.t.itw1, w.tile1(), w.tile_m);
        dit<F2, Geo>(b2, ht.t, ht.t.itw2, w.tile2(), w.tile_m);
        // G = (cyclic + T mod R) mod (R - 1), then x = floor(T / R) + G.
        uint64_t c = block_carry<CARRY_E, CyclicCoef, StoreLimbs, Geo>(CyclicCoef{b1, b2, w.lt, t.crt_inv}, k, StoreLimbs{x}, w.sc, 0);
        // End-around carry: 2^(32k) = 1 mod R - 1. One fold in practice; bounded by the
        // value's size in principle. Slot groups fold while any group has a carry (the carry
        // holds barriers); a fold with carry 0 leaves the limbs as they are.
        for (int guard = 0; guard < 4 && (Geo::grouped ? __syncthreads_or(c != 0) != 0 : c != 0); guard++)
            c = block_carry<CARRY_E, LimbsCoef, StoreLimbs, Geo>(LimbsCoef{x}, k, StoreLimbs{x}, w.sc, c);
        uint64_t top = block_carry<CARRY_E, SumCoef, StoreLimbs, Geo>(SumCoef{x, w.lt + k}, k, StoreLimbs{x}, w.sc, 0);
        return (top != 0) | (c != 0);
    } else {
        bool nz = false;
        for (uint32_t j = tid(); j < k; j += nthreads()) nz |= w.lt[j] != 0;
        const uint32_t flag = __syncthreads_or(nz) ? 1 : 0;
        block_carry<CARRY_E>(ProductCoef{w.a1, w.a2, t.n, t.crt_inv}, k, StoreLimbs{w.ls}, w.sc, 0);
        // S = m N
        load_weighted(w.a1, w.a2, w.ls, k, t);
        dif(w.a1, t, t.dtw1, w.tile1(), w.tile_m);
        dif(w.a2, t, t.dtw2, w.tile2(), w.tile_m);
        pointwise_const(w, nhat1, nhat2, t.n);
        dit(w.a1, t, t.itw1, w.tile1(), w.tile_m);
        dit(w.a2, t, t.itw2, w.tile2(), w.tile_m);
        unweight(w.a1, w.a2, t);
        __syncthreads();
        uint64_t top_s = block_carry<CARRY_E>(ProductCoef{w.a1, w.a2, t.n, t.crt_inv}, 2 * k, StoreTop{w.ls, k}, w.sc, 0);
        // x = floor(T / R) + floor(S / R) + [T mod R != 0]
        uint64_t top = block_carry<CARRY_E>(Sum3Coef{w.lt + k, w.ls, flag}, k, StoreLimbs{x}, w.sc, 0);
        return (top_s != 0) | (top != 0);
    }
}

}  // namespace mm
