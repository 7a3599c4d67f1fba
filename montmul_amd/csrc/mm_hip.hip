// mm_hip.hip - the HIP backend: kernels, per-length tables, batch state and the C ABI
// declared in include/montmul_hip.h.
//
// Replaces the reference's `Gpu` backend (montmul/src/gpu/mod.rs, gpu/device.rs, gpu/plan.rs)
// and its kernels (montmul/kernels/src/lib.rs). One workgroup owns one slot and runs whole
// steps on the device; the host uploads loaded values and exponents, launches, and checks the
// per-slot error words.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "../../include/montmul_hip.h"
#include "mm_slot_steps.h"

// Device scratch of the allocation-free product (mm_mul_noalloc): the spectrum of one factor and
// the working vector (n elements of both fields each, 24 n bytes) and 3k limbs, carved from this
// fixed array. Every GPU the library runs on holds its own copy until the process ends. In host
// code the name stands for a host-side shadow, not a device address: host code uses only the
// per-device address that hipGetSymbolAddress returns (Ctx::scratch).
static constexpr size_t MUL_SCRATCH_BYTES = 4u << 20;
static __device__ __align__(256) unsigned char scratch[MUL_SCRATCH_BYTES];

namespace mm {

// ---- Kernels ---------------------------------------------------------------------------------

// Up to `bits` steps on slot first + blockIdx.x.
template <bool GLOBAL, bool WRAP, bool BARE>
__global__ __launch_bounds__(1024) void k_steps(KArgs A, uint32_t first, uint32_t bits) {
    extern __shared__ __align__(16) unsigned char lds[];
    steps_v1<GLOBAL, WRAP, BARE>(A, first + blockIdx.x, bits, lds);
}

// The kernels are held to 128 VGPRs (four waves per SIMD), so two 512-thread workgroups share a
// CU wherever LDS allows.
template <uint32_t LOGN, bool GLOBAL, bool BARE, uint32_t TILE = 1>
__global__ __launch_bounds__(LOGN > 12 ? 1024 : 512) __attribute__((amdgpu_waves_per_eu(MM_V2_WAVES_PER_EU))) void k_steps_v2(v2::VArgs A, uint32_t first, uint32_t bits) {
    extern __shared__ __align__(16) unsigned char lds[];
    steps_v2<LOGN, GLOBAL, BARE, false, TILE, true>(A, first + blockIdx.x, bits, lds);
}

// ModBatch::is_one (batch.rs:179-183) for a list of slots, one workgroup per listed slot:
// out[b] = 1 if slot list[b]'s raw value (k limbs of x) equals either of its two kept values
// (R mod N, then R mod N + N: 2k limbs per slot in `ones`), else 0.
__global__ void k_is_one(const uint32_t* __restrict__ x, const uint32_t* __restrict__ ones, const uint32_t* __restrict__ list,
                         uint32_t k, int32_t* __restrict__ out) {
    const uint32_t slot = list[blockIdx.x];
    const uint32_t* xs = x + (size_t)slot * k;
    const uint32_t* a = ones + (size_t)slot * 2 * k;
    bool ne0 = false, ne1 = false;
    for (uint32_t j = threadIdx.x; j < k; j += blockDim.x) {
        const uint32_t v = xs[j];
        ne0 |= v != a[j];
        ne1 |= v != a[k + j];
    }
    const int d0 = __syncthreads_or(ne0);
    const int d1 = __syncthreads_or(ne1);
    if (threadIdx.x == 0) out[blockIdx.x] = (!d0 || !d1) ? 1 : 0;
}

__global__ void k_fill_u32(uint32_t* p, uint32_t count, uint32_t v) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < count) p[j] = v;
}

// The unweighting tables times c (a batch's base) for the register-transform step: a 1 bit's
// factor c enters there instead of as a pointwise product (c IDFT(X) = IDFT(c X)).
__global__ void k_scale_tab(const F1* iw1, const v2::S2* iw2, uint32_t c, uint32_t n, F1* o1, v2::S2* o2) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const F1 a = iw1[j];
    o1[j] = F1{mul1(a.re, c), mul1(a.im, c)};
    const v2::S2 b = iw2[j];
    const uint64_t re = ((uint64_t)b.r1 << 31) | b.r0, im = ((uint64_t)b.i1 << 31) | b.i0;
    o2[j] = v2::spl_c(F2{mul2(re, c), mul2(im, c)});
}

// ---- Transform layer (reference: trait Transform, transform.rs:22-46) ----------------------
// `count` vectors of n = L/2 elements, vector v at f1 + v n and f2 + v n. Element j of a vector
// holds coefficients j (real part) and j + n (imaginary part). One workgroup per vector for the
// in-place transforms; grid-stride loops over all elements for the element-wise kernels.

// Coefficients (L per vector, row-major) into elements (Cpu::upload, cpu.rs:353-364).
__global__ void k_t_upload(const uint32_t* c, uint32_t n, uint64_t total, F1* f1, F2* f2) {
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < total; e += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t v = e / n, j = e % n;
        const uint32_t re = c[2 * n * v + j], im = c[2 * n * v + n + j];
        f1[e] = F1{mod1(re), mod1(im)};
        f2[e] = F2{re, im};
    }
}

// Weighted DIF in place: natural order in, bit-reversed out (forward1, cpu.rs:86-95).
__global__ __launch_bounds__(1024) void k_t_forward(Tab t, F1* f1, F2* f2) {
    F1* a1 = f1 + (size_t)blockIdx.x * t.n;
    F2* a2 = f2 + (size_t)blockIdx.x * t.n;
    for (uint32_t j = tid(); j < t.n; j += nthreads()) {
        a1[j] = cmul(a1[j], t.w1[j]);
        a2[j] = cmul(a2[j], t.w2[j]);
    }
    dif(a1, t, t.dtw1);
    dif(a2, t, t.dtw2);
}

// DIT in place and the unweighting zeta^-j / n (inverse1, cpu.rs:96-105).
__global__ __launch_bounds__(1024) void k_t_inverse(Tab t, F1* f1, F2* f2) {
    F1* a1 = f1 + (size_t)blockIdx.x * t.n;
    F2* a2 = f2 + (size_t)blockIdx.x * t.n;
    dit(a1, t, t.itw1);
    dit(a2, t, t.itw2);
    unweight(a1, a2, t);
}

// out = scale[v] a (.) b (Cpu::pointwise, cpu.rs:397-406 with pointwise1, 131-140); no scales: 1.
__global__ void k_t_pointwise(F1* o1, F2* o2, const F1* a1, const F2* a2, const F1* b1, const F2* b2, const uint32_t* scales,
                              uint32_t n, uint64_t total) {
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < total; e += (uint64_t)gridDim.x * blockDim.x) {
        F1 r1 = cmul(a1[e], b1[e]);
        F2 r2 = cmul(a2[e], b2[e]);
        if (scales) {
            const uint32_t c = scales[e / n];
            r1 = cscale_small(r1, c);
            r2 = cscale_small(r2, c);
        }
        o1[e] = r1;
        o2[e] = r2;
    }
}

// Exact coefficients in [0, Q) by CRT, two words (low, high) each (Cpu::download, cpu.rs:365-377).
__global__ void k_t_download(const F1* f1, const F2* f2, uint32_t n, uint64_t total, uint32_t inv, uint64_t* out) {
    for (uint64_t e = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; e < total; e += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t v = e / n, j = e % n;
        const u128 lo = crt(f1[e].re, f2[e].re, inv), hi = crt(f1[e].im, f2[e].im, inv);
        const uint64_t q0 = 2 * n * v + j, q1 = q0 + n;
        out[2 * q0] = (uint64_t)lo;
        out[2 * q0 + 1] = (uint64_t)(lo >> 64);
        out[2 * q1] = (uint64_t)hi;
        out[2 * q1 + 1] = (uint64_t)(hi >> 64);
    }
}

// Weighted forward transform of k-limb strings into out (n elements per vector), in place in
// the output (reference: `t_load_forward`, lib.rs:1040-1049; used for N' and, without the
// wraparound correction, N).
__global__ __launch_bounds__(1024) void k_load_forward(Tab t, const uint32_t* limbs, uint32_t k, F1* out1, F2* out2) {
    const uint32_t v = blockIdx.x, n = t.n;
    F1* a1 = out1 + (size_t)v * n;
    F2* a2 = out2 + (size_t)v * n;
    load_weighted(a1, a2, limbs + (size_t)v * k, k, t);
    __syncthreads();
    dif(a1, t, t.dtw1);
    dif(a2, t, t.dtw2);
}

// Per-slot factors A | B of the half path for N (k = 2h limbs): N-hat as in the reference's
// `half_nhat` (cpu.rs:239-249) and the factors of `half_factors` (cpu.rs:261-277), computed on
// the device. tmp holds 2h elements per slot; ab receives A at [0, h) and B at [h, 2h).
template <class F>
__device__ void half_factors_dev(F* z, F* s, F* ab, const F* u, const F* iu, uint32_t h, F c) {
    // s_p = half_nhat at p: (x, y) = (e + wo, e - wo), e = z_p + conj z_q, wo = (z_p - conj z_q)(-i) u_p.
    for (uint32_t p = tid(); p < h; p += nthreads()) {
        F zp = z[p], zq = cconj(z[partner(p)]);
        F e = cadd(zp, zq), wo = cmul(cmul_mi(csub(zp, zq)), u[p]);
        F xx = cadd(e, wo), yy = csub(e, wo);
        s[p] = p == 0 ? cadd(cmul(xx, c), cmul_i(cmul(yy, c))) : cmul(xx, c);
    }
    __syncthreads();
    for (uint32_t p = tid(); p < h; p += nthreads()) {
        F nx, ny;
        if (p == 0) {
            nx = s[0];
            nx.im = 0;
            ny = s[0];
            ny.re = s[0].im;
            ny.im = 0;
        } else {
            nx = s[p];
            ny = cconj(s[partner(p)]);
        }
        F one{};
        one.re = 1;
        F i_iu = cmul_i(iu[p]);
        F X = cmul(nx, cadd(one, i_iu)), Y = cmul(ny, csub(one, i_iu));
        ab[p] = cadd(X, Y);
        ab[h + p] = cmul(cmul_mi(csub(X, Y)), u[p]);
    }
}

__global__ __launch_bounds__(1024) void k_half_factors(HalfTab ht, const uint32_t* nlimbs, uint32_t k, F1* tmp1, F2* tmp2,
                                                       F1* ab1, F2* ab2, F1 c1, F2 c2) {
    const uint32_t v = blockIdx.x, h = ht.t.n;
    F1* z1 = tmp1 + (size_t)v * 2 * h;
    F2* z2 = tmp2 + (size_t)v * 2 * h;
    const uint32_t* nl = nlimbs + (size_t)v * k;
    for (uint32_t j = tid(); j < h; j += nthreads()) {
        uint32_t re = nl[2 * j], im = nl[2 * j + 1];
        z1[j] = F1{mod1(re), mod1(im)};
        z2[j] = F2{(uint64_t)re, (uint64_t)im};
    }
    __syncthreads();
    dif(z1, ht.t, ht.t.dtw1);
    dif(z2, ht.t, ht.t.dtw2);
    half_factors_dev(z1, z1 + h, ab1 + (size_t)v * 2 * h, ht.u1, ht.iu1, h, c1);
    half_factors_dev(z2, z2 + h, ab2 + (size_t)v * 2 * h, ht.u2, ht.iu2, h, c2);
}

// x[out] <- x[a] x[b] R^-1 with out's modulus (reference: `mul_slots`, lib.rs:963-988). One
// workgroup; vectors in device memory (mb: F(x_b), w: the working vector), limbs in lim.
template <bool WRAP>
__global__ __launch_bounds__(1024) void k_mul(KArgs A, uint32_t out, uint32_t ia, uint32_t ib, F1* mb1, F2* mb2, F1* w1,
                                              F2* w2, uint32_t* lim) {
    extern __shared__ __align__(16) unsigned char lds[];
    const Slots& S = A.s;
    const uint32_t n = A.t.n, k = S.k;
    Work w{w1, w2, lim, lim + 2 * k, (uint64_t*)lds};
    load_weighted(mb1, mb2, S.x + (size_t)ib * k, k, A.t);
    __syncthreads();
    dif(mb1, A.t, A.t.dtw1);
    dif(mb2, A.t, A.t.dtw2);
    load_weighted(w.a1, w.a2, S.x + (size_t)ia * k, k, A.t);
    __syncthreads();
    dif(w.a1, A.t, A.t.dtw1);
    dif(w.a2, A.t, A.t.dtw2);
    pointwise_const(w, mb1, mb2, n);
    dit(w.a1, A.t, A.t.itw1);
    dit(w.a2, A.t, A.t.itw2);
    unweight(w.a1, w.a2, A.t);
    __syncthreads();
    uint32_t bad = block_carry<CARRY_E>(ProductCoef{w.a1, w.a2, n, A.t.crt_inv}, 2 * k, StoreLimbs{w.lt}, w.sc, 0) != 0;
    bad |= redc<WRAP>(w, S.x + (size_t)out * k, k, A.t, A.ht, S.nphat1 + (size_t)out * n, S.nphat2 + (size_t)out * n,
                      S.nhat1 + (size_t)out * n, S.nhat2 + (size_t)out * n);
    if (tid() == 0 && bad) S.err[out] = 1;
}

// ---- Host-side field helpers (table generation) ----------------------------------------------
template <class F> F cone() { F r{}; r.re = 1; return r; }
template <class F> F ci() { F r{}; r.im = 1; return r; }
template <class F> F cpow(F b, u128 e) {
    F r = cone<F>();
    while (e) {
        if (e & 1) r = cmul(r, b);
        b = cmul(b, b);
        e >>= 1;
    }
    return r;
}
inline F1 cinv(F1 a) {
    uint32_t nrm = add1(mul1(a.re, a.re), mul1(a.im, a.im));
    uint32_t r = 1, b = nrm;
    for (uint64_t e = P1 - 2; e; e >>= 1) { if (e & 1) r = mul1(r, b); b = mul1(b, b); }
    return F1{mul1(a.re, r), mul1(neg1(a.im), r)};
}
inline F2 cinv(F2 a) {
    uint64_t nrm = add2(mul2(a.re, a.re), mul2(a.im, a.im));
    uint64_t r = 1, b = nrm;
    for (uint64_t e = P2 - 2; e; e >>= 1) { if (e & 1) r = mul2(r, b); b = mul2(b, b); }
    return F2{mul2(a.re, r), mul2(neg2(a.im), r)};
}
template <class F> struct Ord;
template <> struct Ord<F1> { static constexpr uint32_t E = 32; static constexpr uint64_t COF = (1ull << 30) - 1; };
template <> struct Ord<F2> { static constexpr uint32_t E = 62; static constexpr uint64_t COF = (1ull << 60) - 1; };
// The generator of the 2-Sylow subgroup that the reference uses (field.rs:111-120): the first
// (a + i)^cofactor, a = 1, 2, ..., whose order is the full 2^E.
template <class F> F sylow2() {
    for (uint32_t a = 1;; a++) {
        F z{};
        z.re = a;
        z.im = 1;
        F g = cpow(z, (u128)Ord<F>::COF);
        if (!ceq(cpow(g, (u128)1 << (Ord<F>::E - 1)), cone<F>())) return g;
    }
}
template <class F> const F& gen() { static const F g = sylow2<F>(); return g; }

inline uint32_t ilog2(uint64_t n) { uint32_t l = 0; while ((1ull << l) < n) l++; return l; }

// Host copies of one field's tables at length n.
template <class F> struct HostTab {
    std::vector<F> w, iw, dtw, itw, u, iu;
    bool plus = true;
};

// Roots as the reference picks them (field.rs:237-267): omega = g^(2^(E - log n)) and
// zeta = g^(2^(E - log n - 2)), inverted unless zeta^n = i.
template <class F> HostTab<F> make_tab(uint32_t n, bool weights, bool unpack) {
    HostTab<F> T;
    const uint32_t logn = ilog2(n), E = Ord<F>::E;
    const F g = gen<F>();
    const F omega = cpow(g, (u128)1 << (E - logn));
    if (n >= 4) {
        F i4 = cpow(omega, (u128)(n / 4));
        F mi{};
        mi.im = std::is_same<F, F1>::value ? (uint64_t)P1 - 1 : P2 - 1;
        if (ceq(i4, ci<F>())) T.plus = true;
        else if (ceq(i4, mi)) T.plus = false;
        else { fprintf(stderr, "montmul_hip: omega^(n/4) is not +-i\n"); abort(); }
    }
    if (weights) {
        F zeta = cpow(g, (u128)1 << (E - logn - 2));
        if (!ceq(cpow(zeta, (u128)n), ci<F>())) zeta = cinv(zeta);
        F nn{};
        nn.re = n;
        F ninv = cinv(nn);
        F zi = cinv(zeta);
        T.w.resize(n);
        T.iw.resize(n);
        F a = cone<F>(), b = ninv;
        for (uint32_t j = 0; j < n; j++) {
            T.w[j] = a;
            T.iw[j] = b;
            a = cmul(a, zeta);
            b = cmul(b, zi);
        }
    }
    const uint32_t npass = logn / 2;
    for (uint32_t p = 0; p < npass; p++) {
        uint32_t q = dif_pass_q(logn, p);
        F W = cpow(omega, (u128)1 << (2 * p)), x = cone<F>();
        for (uint32_t j = 0; j < q; j++) {
            F x2 = cmul(x, x);
            T.dtw.push_back(x);
            T.dtw.push_back(x2);
            T.dtw.push_back(cmul(x2, x));
            x = cmul(x, W);
        }
    }
    const F oinv = cinv(omega);
    for (uint32_t p = 0; p < npass; p++) {
        uint32_t q = dit_pass_q(logn, p);
        F V = cpow(oinv, (u128)(n / (4 * q))), x = cone<F>();
        for (uint32_t j = 0; j < q; j++) {
            F x2 = cmul(x, x);
            T.itw.push_back(x);
            T.itw.push_back(x2);
            T.itw.push_back(cmul(x2, x));
            x = cmul(x, V);
        }
    }
    if (unpack) {
        // u[p] = w^rev(p), iu[p] = w^-rev(p), w a primitive 2n-th root (reference: Unpack,
        // cpu.rs:165-179, for the half length n = h).
        const F w = cpow(g, (u128)1 << (E - (logn + 1))), wi = cinv(w);
        std::vector<F> pw(n), pwi(n);
        F a = cone<F>(), b = cone<F>();
        for (uint32_t j = 0; j < n; j++) { pw[j] = a; pwi[j] = b; a = cmul(a, w); b = cmul(b, wi); }
        T.u.resize(n);
        T.iu.resize(n);
        for (uint32_t p = 0; p < n; p++) {
            uint32_t r = 0;
            for (uint32_t bit = 0; bit < logn; bit++) if (p >> bit & 1) r |= 1u << (logn - 1 - bit);
            T.u[p] = pw[r];
            T.iu[p] = pwi[r];
        }
    }
    return T;
}

// Pass plan of the register transforms at length n (n >= 64): one radix-2 or radix-4 pass
// first when log2 n is not a multiple of 3, then radix-8 passes down to stride 1.
inline v2::Plan make_plan(uint32_t n) {
    v2::Plan P{};
    P.n = n;
    P.logn = ilog2(n);
    P.threads = n / 8;
    uint32_t b = P.logn % 3, s = n, np = 0;
    if (b) {
        P.radix[np] = 1u << b;
        s /= P.radix[np];
        P.lstride[np++] = ilog2(s);
    }
    while (s > 1) {
        P.radix[np] = 8;
        s /= 8;
        P.lstride[np++] = ilog2(s);
    }
    P.npass = np;
    uint32_t off = 0;
    for (uint32_t p = 0; p < np; p++) {
        P.tw_off[p] = off;
        off += (1u << P.lstride[p]) * (P.radix[p] - 1);
    }
    return P;
}

inline uint32_t host_revr(uint32_t m, uint32_t r) {
    return r == 8 ? ((m & 1) << 2) | (m & 2) | (m >> 2) : r == 4 ? ((m & 1) << 1) | (m >> 1) : m;
}

// Pass twiddles of the plan: for pass p (radix r, stride s), j < s and slot m = 1 .. r-1, the
// element omega_(r s)^(j rev_r(m)) (inverse roots for the DIT table).
template <class F> std::vector<F> make_v2_twiddles(const v2::Plan& P, bool inverse) {
    const uint32_t E = Ord<F>::E;
    const F omega = cpow(gen<F>(), (u128)1 << (E - P.logn));
    const F base = inverse ? cinv(omega) : omega;
    std::vector<F> out;
    for (uint32_t p = 0; p < P.npass; p++) {
        const uint32_t r = P.radix[p], s = 1u << P.lstride[p];
        const F root = cpow(base, (u128)(P.n / (r * s)));
        F rj = cone<F>();
        for (uint32_t j = 0; j < s; j++) {
            F pw[8];
            pw[0] = cone<F>();
            for (uint32_t q = 1; q < r; q++) pw[q] = cmul(pw[q - 1], rj);
            for (uint32_t m = 1; m < r; m++) out.push_back(pw[host_revr(m, r)]);
            rj = cmul(rj, root);
        }
    }
    return out;
}

// omega^j (or omega^-j), j < n/2, with omega the n-th root of the transforms of length n: the
// twiddles of the radix-2 stage that splits a length-n transform into two of n/2.
template <class F> std::vector<F> split_roots(uint32_t n, bool inverse) {
    const F omega = cpow(gen<F>(), (u128)1 << (Ord<F>::E - ilog2(n)));
    const F base = inverse ? cinv(omega) : omega;
    std::vector<F> out(n / 2);
    F r = cone<F>();
    for (uint32_t j = 0; j < n / 2; j++) {
        out[j] = r;
        r = cmul(r, base);
    }
    return out;
}

// GF(M61^2) constants in the split form of the register transforms' products.
inline std::vector<v2::S2> split_all(const std::vector<F2>& v) {
    std::vector<v2::S2> out(v.size());
    for (size_t i = 0; i < v.size(); i++) out[i] = v2::spl_c(v[i]);
    return out;
}

// omega^(n/4) = i and omega^(n/8) = 2^h (1 + i) (h = 15 in M31, 30 in M61), which the radix-8
// butterflies assume.
inline bool check_small_roots() {
    F1 w8a = cpow(gen<F1>(), (u128)1 << (Ord<F1>::E - 3));
    F2 w8b = cpow(gen<F2>(), (u128)1 << (Ord<F2>::E - 3));
    F1 w4a = cmul(w8a, w8a);
    F2 w4b = cmul(w8b, w8b);
    return w8a.re == (1u << 15) && w8a.im == (1u << 15) && w8b.re == (1ull << 30) && w8b.im == (1ull << 30) &&
           w4a.re == 0 && w4a.im == 1 && w4b.re == 0 && w4b.im == 1;
}

}  // namespace mm

using namespace mm;

// ---- Backend state ----------------------------------------------------------------------------
namespace {

thread_local std::string g_err;
int fail(const std::string& m) { g_err = m; return 1; }
int hip_fail(const char* what, hipError_t e) { return fail(std::string(what) + ": " + hipGetErrorString(e)); }
#define HIPCHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) return hip_fail(#x, e_); } while (0)

struct DevTables {
    uint32_t n = 0;
    Tab t{};
    v2::VTab vt{};  // register-transform tables (n >= 64)
    HalfTab ht{};  // tables at length n (the half path of length 2n uses these)
    std::vector<void*> bufs;
    int device = 0;
    ~DevTables() {
        hipSetDevice(device);
        for (void* b : bufs) hipFree(b);
    }
};

// A launch plan the autotuner can choose: kernel (register radix-8 or radix-4), mode (global or
// LDS-resident), the register kernel's exchange tile in global mode (1: its default) and slots
// per CU (0: as many as fit; 1: one, by padding the workgroup's LDS past half a CU's).
struct PlanChoice {
    bool v2 = false, global = false;
    uint32_t tile = 1, slots_per_cu = 0;
};

struct Ctx {
    mm_options opts{};
    int device = 0;
    hipStream_t stream = nullptr;
    int cus = 0;
    size_t max_lds = 0;
    size_t max_dyn_lds = 0;     // dynamic LDS every step kernel can take
    size_t max_dyn_lds_v2 = 0;  // the same for the register-transform kernels
    size_t max_dyn_lds_mix[2] = {0, 0};  // the mixed-length kernel, 512 and 1,024 threads
    // Mixed-length launches: head counter, length classes and work items in one device
    // buffer (grown on demand), and the latest launch's milliseconds per step.
    unsigned char* mix_buf = nullptr;
    size_t mix_cap = 0;
    double mix_ms_per_step = 0;
    // Autotuner choices: (n, wraparound, slot-count octave) -> (register kernel, global mode).
    std::map<std::tuple<uint32_t, bool, uint32_t>, PlanChoice> tuned;
    std::string name;
    std::mutex mu;
    std::map<uint32_t, std::shared_ptr<DevTables>> tabs;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // This device's copy of the fixed product scratch (null, with scratch_err, if its address
    // could not be taken).
    unsigned char* scratch = nullptr;
    std::string scratch_err;
};

// Device address of the fixed product scratch on `device`, taken once per device (with that
// device current) and kept: each GPU holds its own copy of the array at its own address.
int scratch_address(int device, unsigned char** out) {
    static std::mutex mu;
    static std::map<int, unsigned char*> addr;
    std::lock_guard<std::mutex> g(mu);
    auto it = addr.find(device);
    if (it != addr.end()) {
        *out = it->second;
        return 0;
    }
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return hip_fail("hipGetSymbolAddress(scratch)", e != hipSuccess ? e : hipErrorInvalidSymbol);
    addr[device] = (unsigned char*)p;
    *out = (unsigned char*)p;
    return 0;
}

template <class T> int upload_vec(DevTables& D, const std::vector<T>& v, const T** out) {
    *out = nullptr;
    if (v.empty()) return 0;
    void* p = nullptr;
    HIPCHK(hipMalloc(&p, v.size() * sizeof(T)));
    D.bufs.push_back(p);
    HIPCHK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    *out = (const T*)p;
    return 0;
}

// Tables at complex length n, built and uploaded on first use (reference: Gpu::tables,
// gpu/mod.rs:186-190). The same object serves as the half tables of length-2n transforms.
int get_tables(Ctx* C, uint32_t n, std::shared_ptr<DevTables>* out) {
    std::lock_guard<std::mutex> g(C->mu);
    auto it = C->tabs.find(n);
    if (it != C->tabs.end()) { *out = it->second; return 0; }
    auto D = std::make_shared<DevTables>();
    D->device = C->device;
    D->n = n;
    HostTab<F1> h1 = make_tab<F1>(n, true, true);
    HostTab<F2> h2 = make_tab<F2>(n, true, true);
    if (h1.plus != h2.plus && n >= 4) return fail("fourth roots differ between fields");
    Tab& t = D->t;
    t.n = n;
    t.logn = ilog2(n);
    t.i_is_plus = h1.plus;
    t.crt_inv = 0;
    {
        // (P2 mod P1)^-1 mod P1 (field.rs:283-286).
        uint32_t b = (uint32_t)(P2 % P1), r = 1;
        for (uint64_t e = P1 - 2; e; e >>= 1) { if (e & 1) r = mul1(r, b); b = mul1(b, b); }
        t.crt_inv = r;
    }
    if (upload_vec(*D, h1.w, &t.w1) || upload_vec(*D, h1.iw, &t.iw1) || upload_vec(*D, h2.w, &t.w2) ||
        upload_vec(*D, h2.iw, &t.iw2) || upload_vec(*D, h1.dtw, &t.dtw1) || upload_vec(*D, h2.dtw, &t.dtw2) ||
        upload_vec(*D, h1.itw, &t.itw1) || upload_vec(*D, h2.itw, &t.itw2))
        return 1;
    if (n >= 64) {
        if (!check_small_roots()) return fail("roots of unity differ from the radix-8 butterflies' assumption");
        v2::VTab& V = D->vt;
        V.P = make_plan(n);
        V.w1 = t.w1; V.iw1 = t.iw1;
        if (upload_vec(*D, split_all(h2.w), &V.w2) || upload_vec(*D, split_all(h2.iw), &V.iw2) ||
            upload_vec(*D, make_v2_twiddles<F1>(V.P, false), &V.tw1) || upload_vec(*D, split_all(make_v2_twiddles<F2>(V.P, false)), &V.tw2) ||
            upload_vec(*D, make_v2_twiddles<F1>(V.P, true), &V.itw1) || upload_vec(*D, split_all(make_v2_twiddles<F2>(V.P, true)), &V.itw2))
            return 1;
        // n = 2^14: the radix-2 stage in front of the two sub-transforms of n/2.
        if (n == 16384 && (upload_vec(*D, split_roots<F1>(n, false), &V.sw1) ||
                           upload_vec(*D, split_all(split_roots<F2>(n, false)), &V.sw2) ||
                           upload_vec(*D, split_roots<F1>(n, true), &V.isw1) ||
                           upload_vec(*D, split_all(split_roots<F2>(n, true)), &V.isw2)))
            return 1;
    }
    D->ht.t = t;
    if (upload_vec(*D, h1.u, &D->ht.u1) || upload_vec(*D, h1.iu, &D->ht.iu1) || upload_vec(*D, h2.u, &D->ht.u2) ||
        upload_vec(*D, h2.iu, &D->ht.iu2))
        return 1;
    C->tabs[n] = D;
    *out = D;
    return 0;
}

struct State {
    Ctx* C = nullptr;
    int device = 0;  // C's device, kept so that a state can be freed after its context is closed
    uint32_t k = 0, len = 0, n = 0, base = 0, slots = 0;
    bool global = false, wrap = true, v2 = false;
    uint32_t threads = 0, tile = 1, slots_per_cu = 0;
    uint32_t v1_tile = 0;  // radix-4 kernel in global mode: its transforms' LDS tile (elements)
    size_t lds = 0;
    std::shared_ptr<DevTables> tab, half;
    // device buffers
    uint32_t *d_x = nullptr, *d_exp = nullptr, *d_nbits = nullptr, *d_pos = nullptr, *d_err = nullptr, *d_lim = nullptr;
    F1 *d_nphat1 = nullptr, *d_nhat1 = nullptr, *d_work1 = nullptr;
    F2 *d_nphat2 = nullptr, *d_nhat2 = nullptr, *d_work2 = nullptr;
    F1* d_iwc1 = nullptr;  // unweighting times base (register-transform step)
    uint32_t* d_coef = nullptr;  // n = 2^14 register step: coefficient staging, 6 n words per slot
    v2::S2* d_iwc2 = nullptr;
    // set_modulus staging (N' and N limbs, the half factors' work vectors), allocated on the
    // first admission and kept: no allocation or free (which can wait on the whole device)
    // per admitted slot.
    uint32_t* d_stage_lim = nullptr;
    F1* d_stage1 = nullptr;
    F2* d_stage2 = nullptr;
    // is_one on the device: the two raw values that stand for 1 (2k limbs per slot), a slot list
    // and a flag per listed slot.
    uint32_t* d_ones = nullptr;
    uint32_t* d_ones_list = nullptr;
    int32_t* d_ones_out = nullptr;
    std::vector<uint8_t> ones_set;
    uint32_t ew = 0;                  // exponent row stride in words
    size_t exp_cap = 0;               // words allocated for the exponent rows
    // host mirrors
    std::vector<uint32_t> h_x;
    bool x_host_valid = true;         // every row of h_x equals the device's x
    std::vector<uint8_t> x_dirty;     // rows written on the host since the last upload
    std::vector<uint32_t> x_dirty_list;
    std::vector<std::vector<uint32_t>> exps;
    std::vector<uint32_t> nbits, pos;
    std::vector<uint8_t> e_dirty;     // exponents set on the host since the last upload
    std::vector<uint32_t> e_dirty_list;
    bool exp_all = true;              // the device exponent buffer needs a full upload
    double ms_per_slot_step[2] = {0, 0};
    bool tuned = false;
    double last_ms = 0;
    uint32_t last_launches = 0;
    bool product_in_hbm = false;  // the latest run was run_bare: d_lim holds its last product
    std::vector<uint32_t> failed;     // every slot whose carry check failed in the latest run or mul
    ~State() {
        hipSetDevice(device);
        for (void* p : {(void*)d_x, (void*)d_exp, (void*)d_nbits, (void*)d_pos, (void*)d_err, (void*)d_lim,
                        (void*)d_nphat1, (void*)d_nhat1, (void*)d_work1, (void*)d_nphat2, (void*)d_nhat2, (void*)d_work2,
                        (void*)d_iwc1, (void*)d_iwc2, (void*)d_stage_lim, (void*)d_stage1, (void*)d_stage2, (void*)d_coef,
                        (void*)d_ones, (void*)d_ones_list, (void*)d_ones_out})
            if (p) hipFree(p);
    }
    v2::VArgs vargs() const {
        v2::VArgs A{};
        A.t = tab->vt;
        A.ht = half->vt;
        A.s = args().s;
        A.crt_inv = tab->t.crt_inv;
        A.iwc1 = d_iwc1;
        A.iwc2 = d_iwc2;
        A.coef = d_coef;
        return A;
    }
    // The step kernel's arguments: args() plus the radix-4 kernel's LDS tile in global mode.
    KArgs step_args() const {
        KArgs A = args();
        A.tile_m = !v2 && global ? v1_tile : 0;
        return A;
    }
    KArgs args() const {
        KArgs A{};
        A.t = tab->t;
        if (half) A.ht = half->ht;
        Slots& s = A.s;
        s.x = d_x;
        s.nphat1 = d_nphat1; s.nphat2 = d_nphat2;
        s.nhat1 = d_nhat1; s.nhat2 = d_nhat2;
        s.work1 = d_work1; s.work2 = d_work2;
        s.lim = d_lim;
        s.exp = d_exp; s.nbits = d_nbits; s.pos = d_pos; s.err = d_err;
        s.k = k; s.ew = ew ? ew : 1; s.base = base; s.slots = slots;
        return A;
    }
};

// Threads per workgroup: one radix-4 group per thread, 64 to 1024.
uint32_t plan_threads(uint32_t n) { return std::min<uint32_t>(1024, std::max<uint32_t>(64, n / 4)); }

template <class T> int dalloc(T** p, size_t count) {
    *p = nullptr;
    if (count == 0) return 0;
    hipError_t e = hipMalloc((void**)p, count * sizeof(T));
    if (e != hipSuccess) return hip_fail("hipMalloc", e);
    return 0;
}

typedef void (*StepKernel)(KArgs, uint32_t, uint32_t);
typedef void (*StepKernelV2)(v2::VArgs, uint32_t, uint32_t);
template <uint32_t LOGN> StepKernelV2 pick_v2(bool global, bool bare) {
    if (global) return bare ? k_steps_v2<LOGN, true, true> : k_steps_v2<LOGN, true, false>;
    return bare ? k_steps_v2<LOGN, false, true> : k_steps_v2<LOGN, false, false>;
}
// The exchange tiles the autotuner may pick besides the default (v2::tile_max), in global mode:
// one tile of the whole vector at n = 4,096 (one chunk per exchange, one workgroup per CU), and
// half the vector at n = 512 .. 2,048 (two chunks per exchange, half the LDS per workgroup).
constexpr uint32_t V2_WIDE_TILE = 4096;
bool v2_tile_ok(uint32_t n, bool global, uint32_t tile) {
    return tile == 1 || (global && ((tile == V2_WIDE_TILE && n == 4096) || (tile == n / 2 && n >= 512 && n <= 2048)));
}
// A tile asked for by its size: the size of the kernel's built-in tile (v2::tile_max, or the
// whole vector below it) is that tile, 1.
uint32_t v2_tile_norm(uint32_t n, uint32_t tile) { return tile == std::min<uint32_t>(n, n > 4096 ? 4096 : 2048) ? 1 : tile; }
// The register kernel's exchange tile when none is asked for: in global mode at n = 512 .. 2,048
// half the vector, whose halved LDS lets a CU hold as many workgroups as its registers allow
// (four waves per SIMD rather than three) and keeps T in registers; the two chunks per exchange
// cost less than that gains (at 4,096 slots: L = 1,024 +22 %, 2,048 +18 %, 4,096 +8 %; at L =
// 2,048 with 2,048 slots, which then all fit at once, +35 %). Else the built-in tile.
uint32_t v2_default_tile(uint32_t n, bool global) { return global && n >= 512 && n <= 2048 ? n / 2 : 1; }
template <uint32_t LOGN> StepKernelV2 pick_half_v2(bool bare) {
    return bare ? k_steps_v2<LOGN, true, true, (1u << LOGN) / 2> : k_steps_v2<LOGN, true, false, (1u << LOGN) / 2>;
}
StepKernelV2 pick_step_v2(uint32_t logn, bool global, bool bare, uint32_t tile = 1) {
    if (tile != 1) {
        if (!global || !v2_tile_ok(1u << logn, global, tile)) return nullptr;
        if (logn == 12) return bare ? k_steps_v2<12, true, true, V2_WIDE_TILE> : k_steps_v2<12, true, false, V2_WIDE_TILE>;
        if (logn == 11) return pick_half_v2<11>(bare);
        if (logn == 10) return pick_half_v2<10>(bare);
        if (logn == 9) return pick_half_v2<9>(bare);
        return nullptr;
    }
    switch (logn) {
        case 9: return pick_v2<9>(global, bare);
        case 10: return pick_v2<10>(global, bare);
        case 11: return pick_v2<11>(global, bare);
        case 12: return pick_v2<12>(global, bare);
        case 13: return pick_v2<13>(global, bare);
        // n = 2^14 runs two sub-transforms of n/2 in turn, in global mode only.
        case 14: return global ? (bare ? k_steps_v2<14, true, true> : k_steps_v2<14, true, false>) : nullptr;
        default: return nullptr;
    }
}
StepKernel pick_step(bool global, bool wrap, bool bare) {
    if (global) {
        if (bare) return wrap ? k_steps<true, true, true> : k_steps<true, false, true>;
        return wrap ? k_steps<true, true, false> : k_steps<true, false, false>;
    }
    if (bare) return wrap ? k_steps<false, true, true> : k_steps<false, false, true>;
    return wrap ? k_steps<false, true, false> : k_steps<false, false, false>;
}

// Raises each step kernel's dynamic LDS limit to what the device allows beside the kernel's
// static LDS, and records the smallest such limit for planning.
int set_lds_limit(Ctx* C) {
    const StepKernel ks[] = {k_steps<true, true, true>,   k_steps<true, false, true>,  k_steps<true, true, false>,
                             k_steps<true, false, false>, k_steps<false, true, true>,  k_steps<false, false, true>,
                             k_steps<false, true, false>, k_steps<false, false, false>};
    size_t dyn_max = C->max_lds;
    for (StepKernel k : ks) {
        hipFuncAttributes a;
        HIPCHK(hipFuncGetAttributes(&a, (const void*)k));
        size_t dyn = C->max_lds > a.sharedSizeBytes ? C->max_lds - a.sharedSizeBytes : 0;
        if (hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn) != hipSuccess) {
            (void)hipGetLastError();
            dyn = std::min<size_t>(dyn, 65536);
        }
        dyn_max = std::min(dyn_max, dyn);
    }
    for (uint32_t i = 0; i < 32; i++) {
        // Every length and mode with the default tile, then the other tiles in global mode.
        const uint32_t logn = i < 24 ? 9 + i / 4 : 9 + (i - 24) / 2;
        StepKernelV2 k = i < 24 ? pick_step_v2(logn, (i >> 1) & 1, i & 1)
                                : pick_step_v2(logn, true, i & 1, logn == 12 ? V2_WIDE_TILE : (1u << logn) / 2);
        if (!k) continue;
        hipFuncAttributes a;
        HIPCHK(hipFuncGetAttributes(&a, (const void*)k));
        size_t dyn = C->max_lds > a.sharedSizeBytes ? C->max_lds - a.sharedSizeBytes : 0;
        if (hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn) != hipSuccess) {
            (void)hipGetLastError();
            dyn = std::min<size_t>(dyn, 65536);
        }
        C->max_dyn_lds_v2 = std::min(C->max_dyn_lds_v2 ? C->max_dyn_lds_v2 : dyn, dyn);
    }
    HIPCHK(mix_lds_limit<512>(C->max_lds, &C->max_dyn_lds_mix[0]));
    HIPCHK(mix_lds_limit<1024>(C->max_lds, &C->max_dyn_lds_mix[1]));
    C->max_dyn_lds = dyn_max;
    return 0;
}

// Uploads the rows of x and the exponents written on the host since the last upload: runs of
// consecutive dirty slots go in one copy each, so that admitting a few slots into a large batch
// moves only their rows (reference: Gpu::push, gpu/mod.rs:161-176, which uploads everything).
// Calls copy(first, count) for each run of consecutive slots in `list` (sorted in place).
template <class Copy> int slot_runs(std::vector<uint32_t>& list, Copy&& copy) {
    std::sort(list.begin(), list.end());
    for (size_t a = 0; a < list.size();) {
        size_t b = a + 1;
        while (b < list.size() && list[b] == list[b - 1] + 1) b++;
        if (copy(list[a], list[b - 1] + 1 - list[a])) return 1;
        a = b;
    }
    return 0;
}

// The x part of push: rows of x written on the host since the last upload. Allocates nothing.
int push_x(State* S) {
    if (S->x_dirty_list.empty()) return 0;
    hipStream_t sm = S->C->stream;
    const uint32_t k = S->k;
    if (slot_runs(S->x_dirty_list, [&](uint32_t first, uint32_t cnt) -> int {
            HIPCHK(hipMemcpyAsync(S->d_x + (size_t)first * k, S->h_x.data() + (size_t)first * k, (size_t)cnt * k * 4,
                                  hipMemcpyHostToDevice, sm));
            return 0;
        }))
        return 1;
    for (uint32_t i : S->x_dirty_list) S->x_dirty[i] = 0;
    S->x_dirty_list.clear();
    return 0;
}

int push(State* S) {
    hipStream_t sm = S->C->stream;
    auto runs = [&](std::vector<uint32_t>& list, auto&& copy) -> int { return slot_runs(list, copy); };
    if (push_x(S)) return 1;
    if (!S->e_dirty_list.empty() || S->exp_all) {
        // Rows are as wide as the longest exponent set so far, so that re-arming a batch with
        // short exponents uploads only those words. The buffer holds at least k words per slot,
        // so that exponents below the modulus (N - 1 for Fermat tests) widen the rows without
        // reallocating; a wider row uploads every row again.
        uint32_t ew = S->ew;
        for (uint32_t i : S->e_dirty_list) ew = std::max<uint32_t>(ew, (uint32_t)S->exps[i].size());
        if (ew > S->ew || !S->d_exp) {
            const size_t need = (size_t)S->slots * ew;
            if (need > S->exp_cap || !S->d_exp) {
                const size_t cap = (size_t)S->slots * std::max(ew, S->k);
                HIPCHK(hipStreamSynchronize(sm));
                if (S->d_exp) hipFree(S->d_exp);
                S->d_exp = nullptr;
                S->exp_cap = 0;
                if (dalloc(&S->d_exp, cap)) return 1;
                S->exp_cap = cap;
            }
            S->ew = ew;
            S->exp_all = true;
        }
        std::vector<uint32_t> all;
        if (S->exp_all) {
            all.resize(S->slots);
            for (uint32_t i = 0; i < S->slots; i++) all[i] = i;
        }
        std::vector<uint32_t>& list = S->exp_all ? all : S->e_dirty_list;
        std::vector<uint32_t> words;
        if (runs(list, [&](uint32_t first, uint32_t cnt) -> int {
                words.assign((size_t)cnt * S->ew, 0);
                for (uint32_t i = 0; i < cnt; i++)
                    std::copy(S->exps[first + i].begin(), S->exps[first + i].end(), words.begin() + (size_t)i * S->ew);
                HIPCHK(hipMemcpyAsync(S->d_exp + (size_t)first * S->ew, words.data(), words.size() * 4, hipMemcpyHostToDevice, sm));
                // the staging vector is reused: finish this copy first
                HIPCHK(hipStreamSynchronize(sm));
                return 0;
            }))
            return 1;
        HIPCHK(hipMemcpyAsync(S->d_nbits, S->nbits.data(), S->slots * 4, hipMemcpyHostToDevice, sm));
        HIPCHK(hipMemcpyAsync(S->d_pos, S->pos.data(), S->slots * 4, hipMemcpyHostToDevice, sm));
        HIPCHK(hipStreamSynchronize(sm));
        for (uint32_t i : S->e_dirty_list) S->e_dirty[i] = 0;
        S->e_dirty_list.clear();
        S->exp_all = false;
    }
    return 0;
}

int pull_x(State* S) {
    if (S->x_host_valid) return 0;
    if (push(S)) return 1;  // rows written on the host go to the device first
    HIPCHK(hipMemcpyAsync(S->h_x.data(), S->d_x, S->h_x.size() * 4, hipMemcpyDeviceToHost, S->C->stream));
    HIPCHK(hipStreamSynchronize(S->C->stream));
    S->x_host_valid = true;
    return 0;
}

// First slot whose error word is set, -1 if none, -2 on a device error. Every flagged slot is
// appended to S->failed (mm_failed_slots) and all error words are cleared: a failure is reported
// once, for every slot it hit.
int64_t check_err(State* S, const char* what) {
    std::vector<uint32_t> err(S->slots);
    hipError_t e = hipMemcpyAsync(err.data(), S->d_err, S->slots * 4, hipMemcpyDeviceToHost, S->C->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(S->C->stream);
    if (e != hipSuccess) { hip_fail(what, e); return MM_EFAULT; }
    const size_t before = S->failed.size();
    for (uint32_t i = 0; i < S->slots; i++)
        if (err[i]) S->failed.push_back(i);
    if (S->failed.size() == before) return MM_OK;
    if (hipMemsetAsync(S->d_err, 0, S->slots * 4, S->C->stream) != hipSuccess || hipStreamSynchronize(S->C->stream) != hipSuccess) {
        fail(std::string(what) + ": clearing error words failed");
        return MM_EFAULT;
    }
    const uint32_t first = S->failed[before];
    const size_t more = S->failed.size() - before - 1;
    g_err = std::string(what) + ": carry out of a product that must fit, slot " + std::to_string(first) +
            (more ? " (and " + std::to_string(more) + " more slot" + (more > 1 ? "s" : "") + ")" : "");
    return (int64_t)first;
}

// Runs `reps` launches of one chunk on the context's stream and waits for them; writes their
// device time. Two or more repeats are captured once into a HIP graph and replayed back to back,
// with no host round trip between them (reference: the per-launch loop of Gpu::steps,
// gpu/mod.rs:113-158, synchronizes after every launch). `launch` enqueues kernels only (no
// memset nodes), and returns hipSuccess or the launch error.
template <class Launch> int run_chunks(Ctx* C, uint64_t reps, Launch&& launch, float* ms) {
    hipStream_t sm = C->stream;
    *ms = 0;
    if (reps <= 1) {
        hipEventRecord(C->ev0, sm);
        hipError_t le = launch();
        if (le != hipSuccess) return hip_fail("step kernel launch", le);
        hipEventRecord(C->ev1, sm);
    } else {
        hipGraph_t g = nullptr;
        hipGraphExec_t ge = nullptr;
        HIPCHK(hipStreamBeginCapture(sm, hipStreamCaptureModeThreadLocal));
        hipError_t le = launch();
        hipError_t ce = hipStreamEndCapture(sm, &g);
        if (le != hipSuccess || ce != hipSuccess) {
            if (g) hipGraphDestroy(g);
            return hip_fail("step chunk capture", le != hipSuccess ? le : ce);
        }
        hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (ie != hipSuccess) return hip_fail("step chunk graph", ie);
        hipEventRecord(C->ev0, sm);
        hipError_t ge_err = hipSuccess;
        for (uint64_t r = 0; r < reps && ge_err == hipSuccess; r++) ge_err = hipGraphLaunch(ge, sm);
        hipEventRecord(C->ev1, sm);
        hipError_t se = hipEventSynchronize(C->ev1);
        hipGraphExecDestroy(ge);
        if (ge_err != hipSuccess) return hip_fail("step chunk graph launch", ge_err);
        if (se != hipSuccess) return hip_fail("step kernel", se);
    }
    hipError_t se = hipEventSynchronize(C->ev1);
    if (se != hipSuccess) return hip_fail("step kernel", se);
    hipEventElapsedTime(ms, C->ev0, C->ev1);
    return 0;
}

// Runs up to `bits` steps on every slot with bits left (reference: Gpu::steps, gpu/mod.rs:113-158).
int64_t steps(State* S, uint64_t bits, bool bare) {
    Ctx* C = S->C;
    if (hipSetDevice(C->device) != hipSuccess) { fail("hipSetDevice failed"); return MM_EFAULT; }
    uint64_t most = 0;
    for (uint32_t i = 0; i < S->slots; i++) most = std::max<uint64_t>(most, S->nbits[i] - S->pos[i]);
    uint64_t left = std::min(bits, most);
    S->last_ms = 0;
    S->last_launches = 0;
    S->failed.clear();
    if (left == 0) return MM_OK;
    if (push(S)) return MM_EFAULT;
    // Only run_bare leaves each step's product in HBM: a full step of the register kernel keeps
    // it on chip where it can.
    S->product_in_hbm = bare;
    KArgs A = S->step_args();
    StepKernel kern = pick_step(S->global, S->wrap, bare);
    v2::VArgs A2{};
    StepKernelV2 kern2 = nullptr;
    if (S->v2) {
        A2 = S->vargs();
        kern2 = pick_step_v2(ilog2(S->n), S->global, bare, S->tile);
    }
    while (left > 0) {
        // Steps per launch: all of them, or about max_launch_ms worth from the latest rate.
        uint64_t s = left;
        if (C->opts.max_launch_ms > 0) {
            double e = S->ms_per_slot_step[bare];
            s = e > 0 ? (uint64_t)std::max(1.0, C->opts.max_launch_ms / (e * std::max<uint32_t>(1, S->slots))) : 1;
            s = std::min(s, left);
        }
        s = std::min<uint64_t>(s, 0xffffffffu);
        // Once the rate is known, the remaining whole chunks go as replays of one graph; a carry
        // failure is then reported after them (the failing slot's error word stays set).
        const uint64_t reps = C->opts.max_launch_ms > 0 && S->ms_per_slot_step[bare] > 0 ? std::max<uint64_t>(1, left / s) : 1;
        float ms = 0;
        if (run_chunks(C, reps,
                       [&]() {
                           if (kern2)
                               hipLaunchKernelGGL(kern2, dim3(S->slots), dim3(S->threads), S->lds, C->stream, A2, 0u, (uint32_t)s);
                           else
                               hipLaunchKernelGGL(kern, dim3(S->slots), dim3(S->threads), S->lds, C->stream, A, 0u, (uint32_t)s);
                           return hipGetLastError();
                       },
                       &ms))
            return MM_EFAULT;
        S->last_ms += ms;
        S->last_launches += (uint32_t)reps;
        S->ms_per_slot_step[bare] = std::max(1e-6, (double)ms) / ((double)(s * reps) * std::max<uint32_t>(1, S->slots));
        int64_t r = check_err(S, bare ? "run_bare" : "run");
        for (uint32_t i = 0; i < S->slots; i++) S->pos[i] += (uint32_t)std::min<uint64_t>(s * reps, S->nbits[i] - S->pos[i]);
        if (!bare) S->x_host_valid = false;
        if (r != MM_OK) return r;
        left -= s * reps;
    }
    return MM_OK;
}

// ---- Plans and the autotuner ------------------------------------------------------------------
// The radix-4 kernel's LDS tile in global mode, in elements of one field (32 KiB of F2): its
// transforms run the passes that act within 2,048 elements there (lib.rs:455-511). By default
// from n = 32,768 (L = 65,536) on, where the vector's in-place passes miss L2: +10 % at
// L = 65,536 and +19 % at 131,072, but -16 % at L = 8,192, where they hit it and the tile's
// power-of-two strides conflict in LDS banks. mm_options.tile = 2,048 asks for it at any length.
constexpr uint32_t V1_TILE = 2048, V1_TILE_MIN_N = 32768;
uint32_t v1_tile_for(uint32_t n, bool global, uint32_t tile) {
    if (!global) return 0;
    return tile == V1_TILE || n >= V1_TILE_MIN_N ? std::min<uint32_t>(n, V1_TILE) : 0;
}

// Dynamic LDS at which a second workgroup no longer fits a CU beside the first.
uint64_t one_per_cu_lds(const Ctx* C) { return C->max_lds / 2 + 2048; }

// The launch shape of a plan: the register kernel (n/8 threads; an exchange tile in global mode,
// the whole vector and T resident otherwise) or the radix-4 kernel (one radix-4 group per
// thread; the vector resident in LDS, or in HBM with the carry scratch and a tile for the
// transforms' narrow passes in LDS). With one slot per CU the workgroup's dynamic LDS is padded
// past half of a CU's, so no second one fits.
void set_plan(State* S, const PlanChoice& p) {
    const uint32_t n = S->n, k = S->k;
    S->v2 = p.v2;
    S->global = p.global;
    S->tile = p.global ? (p.v2 ? v2_tile_norm(S->n, p.tile) : p.tile) : 1;
    S->slots_per_cu = p.slots_per_cu;
    S->v1_tile = 0;
    if (p.v2) {
        S->threads = std::min<uint32_t>(n / 8, 1024);  // n = 2^14: two halves of n/2, n/16 threads
        const uint64_t want = S->tile != 1 ? S->tile : (n > 4096 ? 4096 : 2048);  // v2::Tiling
        const uint64_t tile = std::min<uint64_t>(n, want);
        S->lds = p.global ? 24 * tile + lds_scratch_bytes(S->threads) : v2::v2_lds_resident(n, k, S->threads);
    } else {
        S->threads = plan_threads(n);
        S->v1_tile = v1_tile_for(n, p.global, p.tile);
        S->lds = p.global ? v1_tile_offset(S->threads) + 16ull * S->v1_tile : lds_bytes_resident(n, k, S->threads);
    }
    if (p.slots_per_cu == 1) S->lds = std::max<size_t>(S->lds, one_per_cu_lds(S->C));
}
void set_plan(State* S, bool v2, bool global) {
    PlanChoice p;
    p.v2 = v2;
    p.global = global;
    p.tile = v2 ? v2_default_tile(S->n, global) : 1;
    set_plan(S, p);
}

// Whether a plan can run the batch: the register kernel needs the wraparound path and
// n = 512 .. 16,384 (16,384 in global mode only); an LDS-resident plan, or a padded one, needs
// its bytes within the kernel's dynamic LDS; a wider tile is built for n = 4,096 only.
bool plan_fits(const State* S, const PlanChoice& p) {
    const Ctx* C = S->C;
    if (p.v2 && !(S->wrap && S->n >= 512 && S->n <= 16384)) return false;
    if (p.v2 && S->n > 8192 && !p.global) return false;  // n = 2^14: global mode only
    const uint32_t tile = p.v2 ? v2_tile_norm(S->n, p.tile) : p.tile;
    if (p.v2 && !v2_tile_ok(S->n, p.global, tile)) return false;
    if (!p.v2 && p.tile != 1 && !(p.global && p.tile == V1_TILE)) return false;
    uint64_t lds;
    if (p.v2) {
        const uint64_t want = tile != 1 ? tile : (S->n > 4096 ? 4096 : 2048);
        lds = p.global ? 24 * std::min<uint64_t>(S->n, want) + lds_scratch_bytes(std::min<uint32_t>(S->n / 8, 1024))
                       : v2::v2_lds_resident(S->n, S->k, S->n / 8);
    } else {
        lds = p.global ? v1_tile_offset(plan_threads(S->n)) + 16ull * v1_tile_for(S->n, true, p.tile)
                       : lds_bytes_resident(S->n, S->k, plan_threads(S->n));
    }
    if (p.slots_per_cu == 1) lds = std::max<uint64_t>(lds, one_per_cu_lds(C));
    return lds <= (p.v2 ? C->max_dyn_lds_v2 : C->max_dyn_lds);
}
bool plan_fits(const State* S, bool v2, bool global) {
    PlanChoice p;
    p.v2 = v2;
    p.global = global;
    p.tile = v2 ? v2_default_tile(S->n, global) : 1;
    return plan_fits(S, p);
}

// Times every plan that fits for 16 steps on the batch's own slots and keeps the fastest:
// kernel × mode × (register kernel in global mode) exchange tile × slots per CU. The slots hold
// x = 0 and a 16-bit zero exponent for the timing (every kernel's work is the same for any data,
// and 0 stays 0), and are cleared afterwards. The choice is cached per (n, wraparound,
// slot-count octave), so later batches of that shape take it without timing.
int autotune(State* S, bool v2ok) {
    Ctx* C = S->C;
    const auto key = std::make_tuple(S->n, S->wrap, ilog2(S->slots));
    bool have = false;
    PlanChoice best;
    best.v2 = S->v2;
    best.global = S->global;
    {
        std::lock_guard<std::mutex> g(C->mu);
        auto it = C->tuned.find(key);
        if (it != C->tuned.end()) {
            have = true;
            best = it->second;
        }
    }
    if (!have) {
        const uint32_t T = 16;
        hipStream_t sm = C->stream;
        std::vector<uint32_t> nb(S->slots, T);
        HIPCHK(hipMemcpyAsync(S->d_nbits, nb.data(), S->slots * 4ull, hipMemcpyHostToDevice, sm));
        double best_ms = 1e300;
        std::vector<PlanChoice> cands;
        for (int v = v2ok ? 1 : 0; v >= 0; v--)
            for (int g = 1; g >= 0; g--)
                for (uint32_t tile : {1u, V2_WIDE_TILE, V1_TILE, S->n / 2})
                    for (uint32_t spc : {0u, 1u}) {
                        PlanChoice p;
                        p.v2 = v;
                        p.global = g;
                        p.tile = v && g ? v2_tile_norm(S->n, tile) : tile;
                        p.slots_per_cu = spc;
                        bool seen = false;
                        for (const PlanChoice& q : cands)
                            seen |= q.v2 == p.v2 && q.global == p.global && q.tile == p.tile && q.slots_per_cu == p.slots_per_cu;
                        if (!seen && plan_fits(S, p)) cands.push_back(p);
                    }
        for (const PlanChoice& p : cands) {
            set_plan(S, p);
            double t = 1e300;
            for (int rep = 0; rep < 2; rep++) {  // the first launch also loads the code
                HIPCHK(hipMemsetAsync(S->d_pos, 0, S->slots * 4ull, sm));
                HIPCHK(hipEventRecord(C->ev0, sm));
                if (p.v2)
                    hipLaunchKernelGGL(pick_step_v2(ilog2(S->n), p.global, false, S->tile), dim3(S->slots), dim3(S->threads),
                                       S->lds, sm, S->vargs(), 0u, T);
                else
                    hipLaunchKernelGGL(pick_step(p.global, S->wrap, false), dim3(S->slots), dim3(S->threads), S->lds, sm,
                                       S->step_args(), 0u, T);
                HIPCHK(hipGetLastError());
                HIPCHK(hipEventRecord(C->ev1, sm));
                HIPCHK(hipEventSynchronize(C->ev1));
                float ms = 0;
                hipEventElapsedTime(&ms, C->ev0, C->ev1);
                t = std::min(t, (double)ms);
            }
            if (t < best_ms) {
                best_ms = t;
                best = p;
            }
        }
        HIPCHK(hipMemsetAsync(S->d_nbits, 0, S->slots * 4ull, sm));
        HIPCHK(hipMemsetAsync(S->d_pos, 0, S->slots * 4ull, sm));
        HIPCHK(hipMemsetAsync(S->d_err, 0, S->slots * 4ull, sm));
        HIPCHK(hipStreamSynchronize(sm));
        std::lock_guard<std::mutex> g(C->mu);
        C->tuned[key] = best;
    }
    set_plan(S, best);
    S->tuned = true;
    return 0;
}

// ---- Several lengths in one launch ----------------------------------------------------------
// The length class of a batch in the mixed launch of nt threads: the register-transform step in
// global mode (kind log2 n), or the radix-4 step, LDS-resident when its vector fits in 64 KiB.
// Allocates the HBM buffers that mode needs and the batch's own plan did not. Writes the LDS
// bytes a workgroup needs.
// LDS that one workgroup of the mixed launch may take: a CU's 160 KiB for one workgroup of 1,024
// threads, half of it for two of 512.
inline uint64_t mix_lds_cap(uint32_t nt) { return nt >= 1024 ? 159u * 1024u : 79u * 1024u; }

int mix_class(State* S, uint32_t nt, MixClass* out, uint64_t* lds, uint32_t* groups) {
    const uint32_t n = S->n, k = S->k;
    const size_t sl = S->slots;
    memset(out, 0, sizeof(*out));
    *groups = 1;
    // The register-transform step joins the mixed launch up to n = 8,192; n = 2^14 (two halves
    // in turn) runs the radix-4 step there, like longer lengths.
    if (S->v2 && n <= 8192) {
        if (!S->d_lim && dalloc(&S->d_lim, sl * 3 * k)) return 1;
        out->kind = ilog2(n);
        out->v = S->vargs();
        *groups = mix_groups(nt, out->kind);
        *lds = *groups > 1 ? (uint64_t)*groups * mix_group_lds(out->kind) : 24ull * mix_tile(out->kind) + lds_scratch_bytes(nt);
        return 0;
    }
    if (S->wrap && n <= 256 && (uint64_t)(nt / 64) * V1_GROUP_LDS <= mix_lds_cap(nt)) {
        // One wave per slot, NT / 64 slots per workgroup, each LDS-resident.
        out->kind = MIX_V1 | MIX_V1_WRAP | MIX_V1_GROUPS;
        out->k = S->args();
        *groups = nt / 64;
        *lds = (uint64_t)*groups * V1_GROUP_LDS;
        return 0;
    }
    const uint64_t resident = lds_bytes_resident(n, k, nt);
    const bool global = resident > 65536;
    if (global && ((!S->d_work1 && dalloc(&S->d_work1, sl * n)) || (!S->d_work2 && dalloc(&S->d_work2, sl * n)) ||
                   (!S->d_lim && dalloc(&S->d_lim, sl * 3 * k))))
        return 1;
    out->kind = MIX_V1 | (global ? MIX_V1_GLOBAL : 0) | (S->wrap ? MIX_V1_WRAP : 0);
    out->k = S->args();
    out->k.tile_m = v1_tile_for(n, global, 1);  // its transforms' LDS tile, where it pays
    *lds = global ? v1_tile_offset(nt) + 16ull * out->k.tile_m : resident;
    return 0;
}

// Up to `bits` steps on every slot with bits left in every batch of the group, in launches of
// the mixed-length kernel. Returns MM_OK, the first slot whose carry-out check failed (its
// batch's index in *which), or MM_EFAULT.
int64_t group_steps(State* const* G, size_t count, uint64_t bits, size_t* which) {
    if (count == 0) return MM_OK;
    Ctx* C = G[0]->C;
    for (size_t g = 0; g < count; g++)
        if (G[g]->C != C) { fail("a group's batches must share one backend"); return MM_EFAULT; }
    if (hipSetDevice(C->device) != hipSuccess) { fail("hipSetDevice failed"); return MM_EFAULT; }
    uint64_t most = 0;
    uint32_t nt = 512;
    for (size_t g = 0; g < count; g++) {
        const State* S = G[g];
        for (uint32_t i = 0; i < S->slots; i++) most = std::max<uint64_t>(most, S->nbits[i] - S->pos[i]);
        if (S->v2 && S->n <= 8192 && S->n / 8 > nt) nt = S->n / 8;
        G[g]->last_ms = 0;
        G[g]->last_launches = 0;
        G[g]->failed.clear();
    }
    uint64_t left = std::min(bits, most);
    if (left == 0) return MM_OK;
    std::vector<MixClass> cls(count);
    std::vector<uint32_t> groups(count, 1);
    uint64_t lds = 0;
    for (size_t g = 0; g < count; g++) {
        uint64_t l = 0;
        if (push(G[g]) || mix_class(G[g], nt, &cls[g], &l, &groups[g])) return MM_EFAULT;
        lds = std::max(lds, l);
    }
    const size_t lim = C->max_dyn_lds_mix[nt == 1024];
    if (lds > lim) {
        fail("mixed launch needs " + std::to_string(lds) + " bytes of LDS, " + std::to_string(lim) + " available");
        return MM_EFAULT;
    }
    // Workgroups resident per CU: four waves per SIMD of 128 VGPRs, so 1,024 threads take a CU
    // and 512-thread workgroups pair up when their LDS allows.
    const uint32_t per_cu = nt == 1024 ? 1 : (2 * lds <= C->max_lds ? 2 : 1);
    hipStream_t sm = C->stream;
    std::vector<MixItem> items;
    std::vector<uint32_t> list;
    std::vector<double> cost;
    while (left > 0) {
        uint64_t s = left;
        if (C->opts.max_launch_ms > 0)
            s = C->mix_ms_per_step > 0 ? (uint64_t)std::max(1.0, C->opts.max_launch_ms / C->mix_ms_per_step) : 16;
        s = std::min<uint64_t>(std::min(s, left), 0xffffffffu);
        // Work items: the slots with bits left, in runs of up to `groups` slots of one length
        // (slots with the most steps together), the most work first (longest-processing-time
        // order: the longest chains start at once, short ones fill in behind them).
        items.clear();
        list.clear();
        cost.clear();
        for (size_t g = 0; g < count; g++) {
            const State* S = G[g];
            const double per_step = (double)S->n * ilog2(S->n) * (S->v2 ? 1.0 : 3.0);
            std::vector<std::pair<uint64_t, uint32_t>> todo;  // (steps, slot)
            for (uint32_t i = 0; i < S->slots; i++) {
                const uint64_t r = std::min<uint64_t>(s, S->nbits[i] - S->pos[i]);
                if (r > 0) todo.push_back({r, i});
            }
            std::stable_sort(todo.begin(), todo.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
            for (size_t a = 0; a < todo.size(); a += groups[g]) {
                const size_t b = std::min(todo.size(), a + groups[g]);
                items.push_back(MixItem{(uint32_t)g, (uint32_t)list.size(), (uint32_t)(b - a), (uint32_t)todo[a].first});
                for (size_t j = a; j < b; j++) list.push_back(todo[j].second);
                cost.push_back((double)todo[a].first * per_step);
            }
        }
        std::vector<uint32_t> order(items.size());
        for (uint32_t i = 0; i < order.size(); i++) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return cost[a] > cost[b]; });
        std::vector<MixItem> queue(items.size());
        for (size_t i = 0; i < order.size(); i++) queue[i] = items[order[i]];
        const size_t cls_off = 256, items_off = cls_off + ((count * sizeof(MixClass) + 255) & ~(size_t)255);
        const size_t list_off = items_off + ((queue.size() * sizeof(MixItem) + 255) & ~(size_t)255);
        const size_t need = list_off + list.size() * sizeof(uint32_t);
        if (need > C->mix_cap) {
            if (hipStreamSynchronize(sm) != hipSuccess) { fail("stream synchronize failed"); return MM_EFAULT; }
            if (C->mix_buf) hipFree(C->mix_buf);
            C->mix_buf = nullptr;
            C->mix_cap = 0;
            if (dalloc(&C->mix_buf, need * 2)) return MM_EFAULT;
            C->mix_cap = need * 2;
        }
        uint32_t* head = (uint32_t*)C->mix_buf;
        const MixClass* dcls = (const MixClass*)(C->mix_buf + cls_off);
        const MixItem* ditems = (const MixItem*)(C->mix_buf + items_off);
        const uint32_t* dlist = (const uint32_t*)(C->mix_buf + list_off);
        hipError_t ce = hipMemcpyAsync(C->mix_buf + cls_off, cls.data(), count * sizeof(MixClass), hipMemcpyHostToDevice, sm);
        if (ce == hipSuccess)
            ce = hipMemcpyAsync(C->mix_buf + items_off, queue.data(), queue.size() * sizeof(MixItem), hipMemcpyHostToDevice, sm);
        if (ce == hipSuccess)
            ce = hipMemcpyAsync(C->mix_buf + list_off, list.data(), list.size() * sizeof(uint32_t), hipMemcpyHostToDevice, sm);
        if (ce != hipSuccess) { hip_fail("mixed launch upload", ce); return MM_EFAULT; }
        // The upload must land before a replayed graph reads it.
        if (hipStreamSynchronize(sm) != hipSuccess) { fail("mixed launch upload failed"); return MM_EFAULT; }
        const uint32_t grid = (uint32_t)std::min<size_t>(queue.size(), (size_t)C->cus * per_cu);
        // Whole chunks after the first go as replays of one graph (queue reset + launch); slots
        // that finish early take their descriptor and return at once.
        const uint64_t reps = C->opts.max_launch_ms > 0 && C->mix_ms_per_step > 0 ? std::max<uint64_t>(1, left / s) : 1;
        float ms = 0;
        if (run_chunks(C, reps,
                       [&]() {
                           hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, sm, head, 1u, 0u);
                           hipError_t fe = hipGetLastError();
                           if (fe != hipSuccess) return fe;
                           return nt == 1024 ? mix_launch<1024>(grid, lds, sm, dcls, ditems, dlist, (uint32_t)queue.size(), head)
                                             : mix_launch<512>(grid, lds, sm, dcls, ditems, dlist, (uint32_t)queue.size(), head);
                       },
                       &ms))
            return MM_EFAULT;
        C->mix_ms_per_step = std::max(1e-6, (double)ms) / (double)(s * reps);
        int64_t first = MM_OK;
        for (size_t g = 0; g < count; g++) {
            State* S = G[g];
            S->last_ms += ms;
            S->last_launches += (uint32_t)reps;
            const int64_t r = check_err(S, "run");
            if (r == MM_EFAULT) return r;
            if (r != MM_OK && first == MM_OK) {
                first = r;
                if (which) *which = g;
            }
            for (uint32_t i = 0; i < S->slots; i++) S->pos[i] += (uint32_t)std::min<uint64_t>(s * reps, S->nbits[i] - S->pos[i]);
            S->x_host_valid = false;
        }
        if (first != MM_OK) return first;
        left -= s * reps;
    }
    return MM_OK;
}

}  // namespace

// ---- C ABI -------------------------------------------------------------------------------------
extern "C" {

const char* mm_last_error(void) { return g_err.c_str(); }

int mm_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) { fail("hipGetDeviceCount failed"); return -1; }
    return n;
}

int mm_hip_open(const mm_options* opts, mm_backend* out) {
    memset(out, 0, sizeof(*out));
    auto C = new Ctx();
    C->opts = opts ? *opts : mm_options{0, 0, 1, 0, 0, 0};
    C->device = C->opts.device;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || C->device < 0 || C->device >= count) {
        delete C;
        return fail("no HIP device " + std::to_string(opts ? opts->device : 0));
    }
    hipDeviceProp_t p;
    if (hipSetDevice(C->device) != hipSuccess || hipGetDeviceProperties(&p, C->device) != hipSuccess) {
        delete C;
        return fail("cannot query HIP device");
    }
    C->cus = p.multiProcessorCount;
    C->max_lds = p.sharedMemPerBlock;
    // Marketing name with the architecture (e.g. "AMD Instinct MI355X (gfx950)").
    C->name = p.name[0] ? std::string(p.name) + " (" + std::string(p.gcnArchName).substr(0, 6) + ")" : std::string(p.gcnArchName);
    if (hipStreamCreateWithFlags(&C->stream, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&C->ev0) != hipSuccess ||
        hipEventCreate(&C->ev1) != hipSuccess || set_lds_limit(C)) {
        delete C;
        return 1;
    }
    // The fixed product scratch's address on this device (taken on the device's first open).
    // Only mm_mul_noalloc needs it: a failure is reported there.
    if (scratch_address(C->device, &C->scratch)) C->scratch_err = g_err;
    out->self = C;
    out->set_modulus = mm_set_modulus;
    out->run_group = mm_run_group;
    out->create = mm_create;
    out->destroy = mm_destroy;
    out->load = mm_load;
    out->set_exponent = mm_set_exponent;
    out->run = mm_run;
    out->mul = mm_mul;
    out->read = mm_read;
    out->remaining = mm_remaining;
    out->failed_slots = mm_failed_slots;
    out->last_error = mm_last_error;
    out->set_ones = mm_set_ones;
    out->is_one_many = mm_is_one_many;
    // Limits (gpu/mod.rs:272-283): bound Q = P1 * P2, lengths 2^3..2^24, full_k = wraparound.
    u128 Q = (u128)P1 * P2;
    out->bound_lo = (uint64_t)Q;
    out->bound_hi = (uint64_t)(Q >> 64);
    out->min_len = 8;
    out->max_len = 1u << 24;
    out->full_k = C->opts.wraparound ? 1 : 0;
    return 0;
}

void mm_hip_close(mm_backend* be) {
    Ctx* C = (Ctx*)be->self;
    if (!C) return;
    hipSetDevice(C->device);
    C->tabs.clear();
    if (C->mix_buf) hipFree(C->mix_buf);
    if (C->ev0) hipEventDestroy(C->ev0);
    if (C->ev1) hipEventDestroy(C->ev1);
    if (C->stream) hipStreamDestroy(C->stream);
    delete C;
    be->self = nullptr;
}

// Backend::create (reference: Gpu::create, gpu/mod.rs:325-370).
int mm_create(void* self, const mm_batch_setup* st, void** out) {
    Ctx* C = (Ctx*)self;
    *out = nullptr;
    if (hipSetDevice(C->device) != hipSuccess) return fail("hipSetDevice failed");
    if (st->len < 8 || (st->len & (st->len - 1)) || st->len > (1u << 24)) return fail("unsupported transform length");
    if (st->slots == 0) return fail("empty batch");
    const uint32_t n = st->len / 2, k = st->k;
    if (k > n) return fail("k exceeds L/2");
    const bool wrap = C->opts.wraparound != 0;
    if (wrap && k != n) return fail("wraparound needs k = L/2");
    std::unique_ptr<State> S(new State());
    S->C = C;
    S->device = C->device;
    S->k = k; S->len = st->len; S->n = n; S->base = st->base; S->slots = st->slots; S->wrap = wrap;
    const char* kenv = getenv("MONTMUL_KERNEL");
    const bool v2ok = plan_fits(S.get(), true, true) && !(kenv && strcmp(kenv, "v1") == 0);
    if (v2ok) {
        // Global mode unless LDS is preferred: its smaller LDS footprint (T in HBM, exchange
        // tiles of at most 2,048 elements) fits more workgroups per CU, and that wins at every
        // length of this kernel (squarings x bits per second, 4,096 slots: L = 1,024: 3.73e11
        // against 3.21e11 resident; 2,048: 3.41 against 2.96; 4,096: 3.74 against 3.22;
        // 8,192: 3.3 M against 2.3 M squarings/s).
        set_plan(S.get(), true, C->opts.force_global || !plan_fits(S.get(), true, false) || !C->opts.prefer_lds);
    } else {
        set_plan(S.get(), false, C->opts.force_global || !plan_fits(S.get(), false, false));
    }
    const bool tune = C->opts.autotune && !C->opts.force_global;
    if (!tune && (C->opts.tile != 0 || C->opts.slots_per_cu != 0)) {
        PlanChoice p;
        p.v2 = S->v2;
        p.global = S->global;
        p.tile = C->opts.tile != 0 ? C->opts.tile : S->v2 ? v2_default_tile(n, S->global) : 1;
        p.slots_per_cu = C->opts.slots_per_cu;
        if (p.slots_per_cu > 1 || !plan_fits(S.get(), p))
            return fail("tile " + std::to_string(C->opts.tile) + " / slots per CU " + std::to_string(C->opts.slots_per_cu) +
                        " is not a plan for L = " + std::to_string(st->len));
        set_plan(S.get(), p);
    }
    // Device memory: x, N'-hat, N-hat or A|B, work vectors and limbs in global mode, exponents
    // of up to k words, per-slot words, the two raw values of 1 with a list word and a flag,
    // tables (about 6n elements per field, twice), a margin.
    const uint64_t per_slot = 4ull * k + 48ull * n + (S->global || tune ? 24ull * n + 12ull * k : 0) + 16 + 4ull * k +
                              (v2ok && n > 8192 ? 24ull * n : 0) + 8ull * k + 8;
    const uint64_t need = (uint64_t)st->slots * per_slot + 300ull * n + 48ull * n * 2 + (64ull << 20);
    size_t freeb = 0, total = 0;
    if (hipMemGetInfo(&freeb, &total) != hipSuccess) return fail("hipMemGetInfo failed");
    if (need > freeb) return fail("batch needs " + std::to_string(need) + " bytes, " + std::to_string(freeb) + " free");
    if (get_tables(C, n, &S->tab)) return 1;
    if (wrap && get_tables(C, n / 2, &S->half)) return 1;
    const size_t sl = st->slots;
    if (dalloc(&S->d_ones, sl * 2 * k) || dalloc(&S->d_ones_list, sl) || dalloc(&S->d_ones_out, sl)) return 1;
    S->ones_set.assign(sl, 0);
    if (dalloc(&S->d_x, sl * k) || dalloc(&S->d_nbits, sl) || dalloc(&S->d_pos, sl) || dalloc(&S->d_err, sl) ||
        dalloc(&S->d_nphat1, sl * n) || dalloc(&S->d_nphat2, sl * n) || dalloc(&S->d_nhat1, sl * n) ||
        dalloc(&S->d_nhat2, sl * n) || dalloc(&S->d_exp, sl))
        return 1;
    S->ew = 1;
    S->exp_cap = sl;
    if ((S->global || tune) && (dalloc(&S->d_work1, sl * n) || dalloc(&S->d_work2, sl * n) || dalloc(&S->d_lim, sl * 3 * k)))
        return 1;
    if (v2ok) {
        if (dalloc(&S->d_iwc1, n) || dalloc(&S->d_iwc2, n)) return 1;
        if (n > 8192 && dalloc(&S->d_coef, sl * 6 * n)) return 1;
        hipLaunchKernelGGL(k_scale_tab, dim3((n + 255) / 256), dim3(256), 0, C->stream, S->tab->vt.iw1, S->tab->vt.iw2,
                           st->base, n, S->d_iwc1, S->d_iwc2);
        HIPCHK(hipGetLastError());
    }
    hipStream_t sm = C->stream;
    HIPCHK(hipMemsetAsync(S->d_x, 0, sl * k * 4, sm));
    HIPCHK(hipMemsetAsync(S->d_err, 0, sl * 4, sm));
    HIPCHK(hipMemsetAsync(S->d_nbits, 0, sl * 4, sm));
    HIPCHK(hipMemsetAsync(S->d_pos, 0, sl * 4, sm));
    HIPCHK(hipMemsetAsync(S->d_exp, 0, sl * 4, sm));
    // Constants: N'-hat by a weighted forward transform; A|B (or N-hat) from N.
    {
        uint32_t* limbs = nullptr;
        if (dalloc(&limbs, sl * k)) return 1;
        std::unique_ptr<uint32_t, void (*)(uint32_t*)> guard(limbs, [](uint32_t* p) { hipFree(p); });
        const uint32_t tt = plan_threads(n);
        HIPCHK(hipMemcpyAsync(limbs, st->nprime, sl * k * 4, hipMemcpyHostToDevice, sm));
        hipLaunchKernelGGL(k_load_forward, dim3(st->slots), dim3(tt), 0, sm, S->tab->t, limbs, k, S->d_nphat1, S->d_nphat2);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(limbs, st->n, sl * k * 4, hipMemcpyHostToDevice, sm));
        if (wrap) {
            const uint32_t h = n / 2;
            F1 c1 = cinv(F1{mod1(8 * h), 0});
            F2 c2 = cinv(F2{(uint64_t)(8 * h), 0});
            // The work of the half factors lives in the N'-hat-sized N-hat buffer's twin: use a
            // temporary of 2h = n elements per slot.
            F1* t1 = nullptr;
            F2* t2 = nullptr;
            if (dalloc(&t1, sl * n) || dalloc(&t2, sl * n)) { hipFree(t1); return 1; }
            hipLaunchKernelGGL(k_half_factors, dim3(st->slots), dim3(plan_threads(h)), 0, sm, S->half->ht, limbs, k, t1, t2,
                               S->d_nhat1, S->d_nhat2, c1, c2);
            hipError_t le = hipGetLastError();
            hipError_t se = hipStreamSynchronize(sm);
            hipFree(t1);
            hipFree(t2);
            if (le != hipSuccess) return hip_fail("half factors launch", le);
            if (se != hipSuccess) return hip_fail("half factors", se);
        } else {
            hipLaunchKernelGGL(k_load_forward, dim3(st->slots), dim3(tt), 0, sm, S->tab->t, limbs, k, S->d_nhat1, S->d_nhat2);
            HIPCHK(hipGetLastError());
        }
        HIPCHK(hipStreamSynchronize(sm));
    }
    if (tune && autotune(S.get(), v2ok)) return 1;
    S->h_x.assign(sl * k, 0);
    S->x_dirty.assign(sl, 0);
    S->exps.assign(sl, {});
    S->e_dirty.assign(sl, 0);
    S->nbits.assign(sl, 0);
    S->pos.assign(sl, 0);
    *out = S.release();
    return 0;
}

void mm_destroy(void* st) { delete (State*)st; }

int mm_load(void* st, size_t i, const uint32_t* limbs) {  // gpu/mod.rs:371-375
    State* S = (State*)st;
    if (i >= S->slots) return fail("slot out of range");
    std::copy(limbs, limbs + S->k, S->h_x.begin() + i * S->k);
    if (!S->x_dirty[i]) {
        S->x_dirty[i] = 1;
        S->x_dirty_list.push_back((uint32_t)i);
    }
    return 0;
}

int mm_set_exponent(void* st, size_t i, const uint32_t* words, size_t nwords, uint64_t nbits) {  // gpu/mod.rs:376-382
    State* S = (State*)st;
    if (i >= S->slots) return fail("slot out of range");
    if (nbits > 32ull * nwords || nbits > 0xffffffffull) return fail("nbits exceeds the exponent words");
    S->exps[i].assign(words, words + nwords);
    S->nbits[i] = (uint32_t)nbits;
    S->pos[i] = 0;
    if (!S->e_dirty[i]) {
        S->e_dirty[i] = 1;
        S->e_dirty_list.push_back((uint32_t)i);
    }
    return 0;
}

// Device constants of one slot for a new modulus: N'-hat, and A|B (or N-hat), as mm_create
// computes them for every slot. The exponent is cleared.
int mm_set_modulus(void* st, size_t i, const uint32_t* nl, const uint32_t* npl, const uint32_t* two_n) {
    State* S = (State*)st;
    Ctx* C = S->C;
    (void)two_n;  // the device path has no use for 2N
    if (i >= S->slots) return fail("slot out of range");
    if (hipSetDevice(C->device) != hipSuccess) return fail("hipSetDevice failed");
    S->ones_set[i] = 0;  // the values of 1 belong to the old modulus until set_ones
    const uint32_t k = S->k, n = S->n;
    hipStream_t sm = C->stream;
    if (!S->d_stage_lim && dalloc(&S->d_stage_lim, 2ull * k)) return 1;
    if (S->wrap && ((!S->d_stage1 && dalloc(&S->d_stage1, n)) || (!S->d_stage2 && dalloc(&S->d_stage2, n)))) return 1;
    uint32_t* lim = S->d_stage_lim;
    F1* t1 = S->d_stage1;
    F2* t2 = S->d_stage2;
    HIPCHK(hipMemcpyAsync(lim, npl, (size_t)k * 4, hipMemcpyHostToDevice, sm));
    HIPCHK(hipMemcpyAsync(lim + k, nl, (size_t)k * 4, hipMemcpyHostToDevice, sm));
    const uint32_t tt = plan_threads(n);
    hipLaunchKernelGGL(k_load_forward, dim3(1), dim3(tt), 0, sm, S->tab->t, lim, k, S->d_nphat1 + i * n, S->d_nphat2 + i * n);
    HIPCHK(hipGetLastError());
    if (S->wrap) {
        const uint32_t h = n / 2;
        F1 c1 = cinv(F1{mod1(8 * h), 0});
        F2 c2 = cinv(F2{(uint64_t)(8 * h), 0});
        hipLaunchKernelGGL(k_half_factors, dim3(1), dim3(plan_threads(h)), 0, sm, S->half->ht, lim + k, k, t1, t2,
                           S->d_nhat1 + i * n, S->d_nhat2 + i * n, c1, c2);
    } else {
        hipLaunchKernelGGL(k_load_forward, dim3(1), dim3(tt), 0, sm, S->tab->t, lim + k, k, S->d_nhat1 + i * n,
                           S->d_nhat2 + i * n);
    }
    hipError_t le = hipGetLastError();
    hipError_t se = hipStreamSynchronize(sm);  // the staging buffers are reused by the next call
    if (le != hipSuccess) return hip_fail("set_modulus launch", le);
    if (se != hipSuccess) return hip_fail("set_modulus", se);
    S->exps[i].clear();
    S->nbits[i] = 0;
    S->pos[i] = 0;
    if (!S->e_dirty[i]) {
        S->e_dirty[i] = 1;
        S->e_dirty_list.push_back((uint32_t)i);
    }
    return 0;
}

int64_t mm_run(void* st, uint64_t bits) { return steps((State*)st, bits, false); }
int64_t mm_run_bare(void* st, uint64_t bits) { return steps((State*)st, bits, true); }

// Batches of several lengths stepped together by the mixed-length kernel (no reference
// counterpart; see mm_backend.run_group).
int64_t mm_run_group(void* const* states, size_t count, uint64_t bits, size_t* which) {
    std::vector<State*> G(count);
    for (size_t g = 0; g < count; g++) {
        G[g] = (State*)states[g];
        if (!G[g]) { fail("null batch in a group"); return MM_EFAULT; }
    }
    return group_steps(G.data(), count, bits, which);
}

int64_t mm_mul(void* st, size_t out, size_t a, size_t b) {  // gpu/mod.rs:386-391
    State* S = (State*)st;
    Ctx* C = S->C;
    if (out >= S->slots || a >= S->slots || b >= S->slots) { fail("slot out of range"); return MM_EFAULT; }
    if (hipSetDevice(C->device) != hipSuccess || push(S)) return MM_EFAULT;
    S->failed.clear();
    const uint32_t n = S->n, k = S->k;
    F1 *mb1 = nullptr, *w1 = nullptr;
    F2 *mb2 = nullptr, *w2 = nullptr;
    uint32_t* lim = nullptr;
    int64_t r = MM_EFAULT;
    if (!dalloc(&mb1, n) && !dalloc(&mb2, n) && !dalloc(&w1, n) && !dalloc(&w2, n) && !dalloc(&lim, 3ull * k)) {
        auto kern = S->wrap ? k_mul<true> : k_mul<false>;
        hipLaunchKernelGGL(kern, dim3(1), dim3(S->threads), lds_scratch_bytes(S->threads), C->stream, S->args(), (uint32_t)out,
                           (uint32_t)a, (uint32_t)b, mb1, mb2, w1, w2, lim);
        hipError_t le = hipGetLastError();
        hipError_t se = hipStreamSynchronize(C->stream);
        if (le != hipSuccess) hip_fail("mul launch", le);
        else if (se != hipSuccess) hip_fail("mul", se);
        else r = check_err(S, "mul");
    }
    for (void* p : {(void*)mb1, (void*)mb2, (void*)w1, (void*)w2, (void*)lim}) if (p) hipFree(p);
    S->x_host_valid = false;
    return r;
}

// mm_mul with its scratch (the spectrum of x_b, the working vector, 3k limbs) carved from this
// device's fixed scratch array instead of allocated per call. Allocates nothing: x rows loaded
// on the host go up by copies into the batch's existing buffer (exponents wait for the next
// run), and only the output slot's error word is read back. The kernel writes every scratch
// word before reading it, so nothing needs clearing between calls.
int64_t mm_mul_noalloc(void* st, size_t out, size_t a, size_t b) {
    State* S = (State*)st;
    Ctx* C = S->C;
    if (out >= S->slots || a >= S->slots || b >= S->slots) { fail("slot out of range"); return MM_EFAULT; }
    if (!C->scratch) { fail("mul_noalloc: no fixed scratch on this device: " + C->scratch_err); return MM_EFAULT; }
    const uint32_t n = S->n, k = S->k;
    auto up = [](size_t v) { return (v + 255) & ~(size_t)255; };
    const size_t o_mb2 = up(n * sizeof(F1)), o_w1 = o_mb2 + up(n * sizeof(F2)), o_w2 = o_w1 + up(n * sizeof(F1));
    const size_t o_lim = o_w2 + up(n * sizeof(F2)), need = o_lim + up(12ull * k);
    if (need > MUL_SCRATCH_BYTES) {
        fail("mul_noalloc: length " + std::to_string(S->len) + " needs " + std::to_string(need) +
             " bytes of scratch, the fixed scratch holds " + std::to_string(MUL_SCRATCH_BYTES) + " (lengths up to 131072)");
        return MM_EFAULT;
    }
    if (hipSetDevice(C->device) != hipSuccess) { fail("hipSetDevice failed"); return MM_EFAULT; }
    if (push_x(S)) return MM_EFAULT;
    S->failed.clear();
    unsigned char* base = C->scratch;
    auto kern = S->wrap ? k_mul<true> : k_mul<false>;
    hipLaunchKernelGGL(kern, dim3(1), dim3(S->threads), lds_scratch_bytes(S->threads), C->stream, S->args(), (uint32_t)out,
                       (uint32_t)a, (uint32_t)b, (F1*)base, (F2*)(base + o_mb2), (F1*)(base + o_w1), (F2*)(base + o_w2),
                       (uint32_t*)(base + o_lim));
    hipError_t le = hipGetLastError();
    if (le != hipSuccess) { hip_fail("mul_noalloc launch", le); return MM_EFAULT; }
    S->x_host_valid = false;
    uint32_t err = 0;
    hipError_t e = hipMemcpyAsync(&err, S->d_err + out, 4, hipMemcpyDeviceToHost, C->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(C->stream);
    if (e != hipSuccess) { hip_fail("mul_noalloc", e); return MM_EFAULT; }
    if (err) {
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, C->stream, S->d_err + out, 1u, 0u);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(C->stream);
        if (e != hipSuccess) { hip_fail("mul_noalloc: clearing the error word", e); return MM_EFAULT; }
        fail("mul_noalloc: carry out of a product that must fit, slot " + std::to_string(out));
        S->failed.push_back((uint32_t)out);
        return (int64_t)out;
    }
    return MM_OK;
}

// The raw values that stand for 1, for is_one_many (no reference counterpart).
int mm_set_ones(void* st, size_t first, size_t count, const uint32_t* targets) {
    State* S = (State*)st;
    if (first > S->slots || count > S->slots - first) return fail("set_ones: slot range past the batch");
    if (count == 0) return 0;
    if (hipSetDevice(S->C->device) != hipSuccess) return fail("hipSetDevice failed");
    const size_t w = 2ull * S->k;
    HIPCHK(hipMemcpyAsync(S->d_ones + first * w, targets, count * w * 4, hipMemcpyHostToDevice, S->C->stream));
    HIPCHK(hipStreamSynchronize(S->C->stream));
    std::fill(S->ones_set.begin() + first, S->ones_set.begin() + first + count, 1);
    return 0;
}

// ModBatch::is_one (batch.rs:179-183) for many slots on the device: the rows written on the host
// go up first, then one workgroup compares each listed slot's x with its two kept values, and
// one copy brings the flags back.
int mm_is_one_many(void* st, const uint32_t* slots, size_t count, int32_t* out) {
    State* S = (State*)st;
    for (size_t i = 0; i < count; i++) {
        if (slots[i] >= S->slots) return fail("is_one_many: slot " + std::to_string(slots[i]) + " out of range");
        if (!S->ones_set[slots[i]]) return fail("is_one_many: slot " + std::to_string(slots[i]) + " has no value of 1 kept");
    }
    if (count == 0) return 0;
    if (hipSetDevice(S->C->device) != hipSuccess) return fail("hipSetDevice failed");
    if (push_x(S)) return 1;
    hipStream_t sm = S->C->stream;
    // The list and flag buffers hold one word per slot of the batch: longer lists (a slot may
    // be listed more than once) go in parts of that size.
    for (size_t a = 0; a < count; a += S->slots) {
        const size_t c = std::min<size_t>(S->slots, count - a);
        HIPCHK(hipMemcpyAsync(S->d_ones_list, slots + a, c * 4, hipMemcpyHostToDevice, sm));
        hipLaunchKernelGGL(k_is_one, dim3((uint32_t)c), dim3(256), 0, sm, S->d_x, S->d_ones, S->d_ones_list, S->k, S->d_ones_out);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(out + a, S->d_ones_out, c * 4, hipMemcpyDeviceToHost, sm));
        HIPCHK(hipStreamSynchronize(sm));
    }
    return 0;
}

int mm_read(void* st, size_t i, uint32_t* out) {  // gpu/mod.rs:392-394
    State* S = (State*)st;
    if (i >= S->slots) return fail("slot out of range");
    if (S->x_host_valid || S->x_dirty[i]) {
        std::copy(S->h_x.begin() + i * S->k, S->h_x.begin() + (i + 1) * S->k, out);
        return 0;
    }
    // The device row alone (the reference downloads every slot here, gpu/mod.rs:247-269).
    if (hipSetDevice(S->C->device) != hipSuccess) return fail("hipSetDevice failed");
    HIPCHK(hipMemcpyAsync(out, S->d_x + i * S->k, (size_t)S->k * 4, hipMemcpyDeviceToHost, S->C->stream));
    HIPCHK(hipStreamSynchronize(S->C->stream));
    return 0;
}

uint64_t mm_remaining(void* st, size_t i) {  // gpu/mod.rs:395-397
    State* S = (State*)st;
    if (i >= S->slots) return 0;
    return S->nbits[i] - S->pos[i];
}

int mm_push(void* st) {
    State* S = (State*)st;
    if (hipSetDevice(S->C->device) != hipSuccess) return fail("hipSetDevice failed");
    if (push(S)) return 1;
    HIPCHK(hipStreamSynchronize(S->C->stream));
    return 0;
}

int mm_plan(void* st, mm_plan_info* out) {
    State* S = (State*)st;
    out->global = S->global;
    out->threads = S->threads;
    out->lds_bytes = (uint32_t)S->lds;
    out->slots_per_wg = 1;
    out->kernel = S->v2 ? 8 : 4;
    out->tuned = S->tuned ? 1 : 0;
    out->tile = !S->global ? 0
                : S->v2       ? (S->tile != 1 ? S->tile : std::min<uint32_t>(S->n, S->n > 4096 ? 4096 : 2048))
                              : S->v1_tile;
    out->slots_per_cu = S->slots_per_cu;
    return 0;
}

uint64_t mm_free_memory(void* self) {
    Ctx* C = (Ctx*)self;
    size_t f = 0, t = 0;
    hipSetDevice(C->device);
    hipMemGetInfo(&f, &t);
    return f;
}
int mm_cu_count(void* self) { return ((Ctx*)self)->cus; }
uint64_t mm_max_lds(void* self) { return ((Ctx*)self)->max_lds; }
int mm_device_name(void* self, char* buf, size_t len) {
    Ctx* C = (Ctx*)self;
    if (len) { strncpy(buf, C->name.c_str(), len - 1); buf[len - 1] = 0; }
    return 0;
}
// T = c x^2 (2k limbs) of slot i's latest step, where global mode keeps it in HBM.
int mm_read_product(void* st, size_t i, uint32_t* out) {
    State* S = (State*)st;
    if (i >= S->slots) return fail("slot out of range");
    if (!S->global || !S->d_lim) return fail("the product stays in HBM only in global mode");
    if (!S->product_in_hbm) return fail("the product is left in HBM by run_bare only");
    if (hipSetDevice(S->C->device) != hipSuccess) return fail("hipSetDevice failed");
    HIPCHK(hipMemcpyAsync(out, S->d_lim + i * 3 * (size_t)S->k, 2ull * S->k * 4, hipMemcpyDeviceToHost, S->C->stream));
    HIPCHK(hipStreamSynchronize(S->C->stream));
    return 0;
}

size_t mm_failed_slots(void* st, uint32_t* out, size_t max) {
    const State* S = (const State*)st;
    for (size_t i = 0; i < S->failed.size() && i < max; i++) out[i] = S->failed[i];
    return S->failed.size();
}

int mm_last_run_kernel_ms(void* st, double* ms, uint32_t* launches) {
    State* S = (State*)st;
    *ms = S->last_ms;
    *launches = S->last_launches;
    return 0;
}

}  // extern "C"

// ---- Transform layer (trait Transform for Gpu, gpu/mod.rs:286-319; load_forward and
// download_raw, gpu/mod.rs:212-225) ------------------------------------------------------------
struct mm_tbuf {
    Ctx* C;
    uint32_t len, count;
    F1* f1;
    F2* f2;
    int device;  // C's device: a buffer can be freed after its context is closed
};
struct mm_tscales {
    Ctx* C;
    uint32_t count;
    uint32_t* s;
    int device;
};

namespace {
// Grid of a grid-stride element kernel: a few workgroups per CU.
dim3 elem_grid(const Ctx* C, uint64_t total) {
    const uint64_t want = (total + 255) / 256, cap = 8ull * (uint64_t)std::max(1, C->cus);
    return dim3((uint32_t)std::max<uint64_t>(1, std::min(want, cap)));
}
int check_len(Ctx* C, uint32_t len) {
    if (len < 8 || (len & (len - 1)) || len > (1u << 24)) return fail("unsupported transform length");
    (void)C;
    return 0;
}
int new_tbuf(Ctx* C, uint32_t len, uint32_t count, mm_tbuf** out) {
    *out = nullptr;
    if (check_len(C, len)) return 1;
    if (count == 0) return fail("empty transform batch");
    std::unique_ptr<mm_tbuf> b(new mm_tbuf{C, len, count, nullptr, nullptr, C->device});
    const size_t el = (size_t)len / 2 * count;
    if (dalloc(&b->f1, el) || dalloc(&b->f2, el)) {
        hipFree(b->f1);
        return 1;
    }
    *out = b.release();
    return 0;
}
int match(const mm_tbuf* b, uint32_t len, uint32_t count) {
    if (!b) return fail("null transform buffer");
    if (b->len != len || b->count != count) return fail("transform buffer holds another length or count");
    return 0;
}
}  // namespace

extern "C" {

int mm_t_upload(void* self, const uint32_t* coeffs, uint32_t len, uint32_t count, mm_tbuf** out) {
    Ctx* C = (Ctx*)self;
    if (hipSetDevice(C->device) != hipSuccess) return fail("hipSetDevice failed");
    if (new_tbuf(C, len, count, out)) return 1;
    mm_tbuf* b = *out;
    uint32_t* d = nullptr;
    const size_t words = (size_t)len * count;
    if (dalloc(&d, words)) { mm_t_free(b); *out = nullptr; return 1; }
    hipError_t e = hipMemcpyAsync(d, coeffs, words * 4, hipMemcpyHostToDevice, C->stream);
    const uint64_t total = (uint64_t)len / 2 * count;
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_t_upload, elem_grid(C, total), dim3(256), 0, C->stream, d, len / 2, total, b->f1, b->f2);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(C->stream);
    hipFree(d);
    if (e != hipSuccess) { mm_t_free(b); *out = nullptr; return hip_fail("upload", e); }
    return 0;
}

int mm_t_load_forward(void* self, const uint32_t* limbs, uint32_t k, uint32_t len, uint32_t count, mm_tbuf** out) {
    Ctx* C = (Ctx*)self;
    if (hipSetDevice(C->device) != hipSuccess) return fail("hipSetDevice failed");
    if (check_len(C, len)) return 1;
    if (k > len / 2) return fail("k exceeds L/2");
    std::shared_ptr<DevTables> T;
    if (get_tables(C, len / 2, &T)) return 1;
    if (new_tbuf(C, len, count, out)) return 1;
    mm_tbuf* b = *out;
    uint32_t* d = nullptr;
    const size_t words = (size_t)k * count;
    if (dalloc(&d, std::max<size_t>(words, 1))) { mm_t_free(b); *out = nullptr; return 1; }
    hipError_t e = words ? hipMemcpyAsync(d, limbs, words * 4, hipMemcpyHostToDevice, C->stream) : hipSuccess;
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_load_forward, dim3(count), dim3(plan_threads(len / 2)), 0, C->stream, T->t, d, k, b->f1, b->f2);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(C->stream);
    hipFree(d);
    if (e != hipSuccess) { mm_t_free(b); *out = nullptr; return hip_fail("load_forward", e); }
    return 0;
}

int mm_t_download(const mm_tbuf* b, uint64_t* out) {
    if (!b) return fail("null transform buffer");
    Ctx* C = b->C;
    if (hipSetDevice(C->device) != hipSuccess) return fail("hipSetDevice failed");
    std::shared_ptr<DevTables> T;
    if (get_tables(C, b->len / 2, &T)) return 1;
    uint64_t* d = nullptr;
    const size_t words = 2ull * b->len * b->count;
    if (dalloc(&d, words)) return 1;
    const uint64_t total = (uint64_t)b->len / 2 * b->count;
    hipLaunchKernelGGL(k_t_download, elem_grid(C, total), dim3(256), 0, C->stream, b->f1, b->f2, b->len / 2, total,
                       T->t.crt_inv, d);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, d, words * 8, hipMemcpyDeviceToHost, C->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(C->stream);
    hipFree(d);
    if (e != hipSuccess) return hip_fail("download", e);
    return 0;
}

int mm_t_download_raw(const mm_tbuf* b, uint32_t* f1, uint64_t* f2) {
    if (!b) return fail("null transform buffer");
    Ctx* C = b->C;
    if (hipSetDevice(C->device) != hipSuccess) return fail("hipSetDevice failed");
    const size_t el = (size_t)b->len / 2 * b->count;
    HIPCHK(hipMemcpyAsync(f1, b->f1, el * sizeof(F1), hipMemcpyDeviceToHost, C->stream));
    HIPCHK(hipMemcpyAsync(f2, b->f2, el * sizeof(F2), hipMemcpyDeviceToHost, C->stream));
    HIPCHK(hipStreamSynchronize(C->stream));
    return 0;
}

int mm_t_upload_scales(void* self, const uint32_t* scales, uint32_t count, mm_tscales** out) {
    Ctx* C = (Ctx*)self;
    *out = nullptr;
    if (hipSetDevice(C->device) != hipSuccess) return fail("hipSetDevice failed");
    std::unique_ptr<mm_tscales> s(new mm_tscales{C, count, nullptr, C->device});
    if (dalloc(&s->s, std::max<uint32_t>(count, 1))) return 1;
    if (count) {
        hipError_t e = hipMemcpy(s->s, scales, (size_t)count * 4, hipMemcpyHostToDevice);
        if (e != hipSuccess) { hipFree(s->s); return hip_fail("upload_scales", e); }
    }
    *out = s.release();
    return 0;
}

static int transform_pass(mm_tbuf* b, uint32_t len, uint32_t count, bool fwd) {
    if (match(b, len, count)) return 1;
    Ctx* C = b->C;
    if (hipSetDevice(C->device) != hipSuccess) return fail("hipSetDevice failed");
    std::shared_ptr<DevTables> T;
    if (get_tables(C, len / 2, &T)) return 1;
    if (fwd)
        hipLaunchKernelGGL(k_t_forward, dim3(count), dim3(plan_threads(len / 2)), 0, C->stream, T->t, b->f1, b->f2);
    else
        hipLaunchKernelGGL(k_t_inverse, dim3(count), dim3(plan_threads(len / 2)), 0, C->stream, T->t, b->f1, b->f2);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(C->stream));
    return 0;
}
int mm_t_forward(mm_tbuf* b, uint32_t len, uint32_t count) { return transform_pass(b, len, count, true); }
int mm_t_inverse(mm_tbuf* b, uint32_t len, uint32_t count) { return transform_pass(b, len, count, false); }

int mm_t_pointwise(mm_tbuf* out, const mm_tbuf* a, const mm_tbuf* b, const mm_tscales* s, uint32_t len, uint32_t count) {
    if (match(out, len, count) || match(a, len, count) || match(b, len, count)) return 1;
    if (s && s->count < count) return fail("fewer scales than vectors");
    Ctx* C = out->C;
    if (hipSetDevice(C->device) != hipSuccess) return fail("hipSetDevice failed");
    const uint64_t total = (uint64_t)len / 2 * count;
    hipLaunchKernelGGL(k_t_pointwise, elem_grid(C, total), dim3(256), 0, C->stream, out->f1, out->f2, a->f1, a->f2, b->f1,
                       b->f2, s ? s->s : nullptr, len / 2, total);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(C->stream));
    return 0;
}

void mm_t_free(mm_tbuf* b) {
    if (!b) return;
    hipSetDevice(b->device);
    hipFree(b->f1);
    hipFree(b->f2);
    delete b;
}
void mm_t_free_scales(mm_tscales* s) {
    if (!s) return;
    hipSetDevice(s->device);
    hipFree(s->s);
    delete s;
}

}  // extern "C"
