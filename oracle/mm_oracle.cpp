// mm_oracle.cpp - TEST INFRASTRUCTURE ONLY. Not part of the product.
//
// A CPU restatement of the reference's arithmetic for the batched Montgomery step, used by
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker. Nothing in
// montmul_amd/ links, loads or calls it.
//
// What it restates (every function cites the reference line it follows):
//  - field.rs        GF(M31^2) and GF(M61^2) arithmetic, the 2-Sylow generator search, the
//                    per-length tables (tw, itw, w, iw) and CRT.
//  - cpu.rs          the `Cpu` backend: radix-2 DIF/DIT, weighted forward/inverse, carry, the
//                    wraparound half path (Half, half_nhat, half_pointwise, half_factors,
//                    wrap_combine), REDC, square_step, the Transform layer, and the Backend
//                    methods, threaded over slots (rayon in the reference, std::thread here).
//  - gmp.rs          the `Gmp` backend: closed-form REDC on GMP integers.
//
// One addition beyond the reference: each Cpu slot carries an error word with the meaning of
// the GPU's (kernels/src/lib.rs:753-754): set when the carry out of T's 2k limbs, or of the
// reduced value's k limbs, is nonzero. The reference's CPU path only debug-asserts these.
//
// Pinning (tests/test_oracle.py): golden vectors written by tests/golden/make_golden.py from
// the closed form with Python integers, GMP's mpz_powm, the exactness closed form
// (tests/common/mod.rs:46-51), known tribonacci primes (OEIS A092835), a brute-force
// O(n^2) DFT over GF(p^2) at small n, CRT round trips and the generator orders.

#include <gmp.h>

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../include/montmul_hip.h"

namespace {

typedef unsigned __int128 u128;

// ---- field.rs:7-42 --------------------------------------------------------------------
constexpr uint32_t P1 = 0x7fffffffu;
constexpr uint64_t P2 = 0x1fffffffffffffffull;
constexpr u128 Q = (u128)P1 * (u128)P2;  // field.rs:11

inline uint32_t add1(uint32_t a, uint32_t b) { uint32_t s = a + b; return s >= P1 ? s - P1 : s; }
inline uint32_t sub1(uint32_t a, uint32_t b) { return a >= b ? a - b : a + P1 - b; }
inline uint32_t mul1(uint32_t a, uint32_t b) {
    uint64_t t = (uint64_t)a * b;
    uint32_t s = (uint32_t)((t & P1) + (t >> 31));
    return s >= P1 ? s - P1 : s;
}
inline uint64_t add2(uint64_t a, uint64_t b) { uint64_t s = a + b; return s >= P2 ? s - P2 : s; }
inline uint64_t sub2(uint64_t a, uint64_t b) { return a >= b ? a - b : a + P2 - b; }
inline uint64_t mul2(uint64_t a, uint64_t b) {
    u128 t = (u128)a * b;
    uint64_t s = ((uint64_t)t & P2) + (uint64_t)(t >> 61);
    return s >= P2 ? s - P2 : s;
}

// ---- field.rs:74-86: elements -----------------------------------------------------------
struct F1 { uint32_t re, im; };
struct F2 { uint64_t re, im; };
inline bool operator==(F1 a, F1 b) { return a.re == b.re && a.im == b.im; }
inline bool operator==(F2 a, F2 b) { return a.re == b.re && a.im == b.im; }
inline bool operator!=(F1 a, F1 b) { return !(a == b); }
inline bool operator!=(F2 a, F2 b) { return !(a == b); }

// field.rs:88-109
uint32_t base_pow1(uint32_t b, uint64_t e) {
    uint32_t r = 1;
    while (e > 0) { if (e & 1) r = mul1(r, b); b = mul1(b, b); e >>= 1; }
    return r;
}
uint64_t base_pow2(uint64_t b, uint64_t e) {
    uint64_t r = 1;
    while (e > 0) { if (e & 1) r = mul2(r, b); b = mul2(b, b); e >>= 1; }
    return r;
}

// Elem for F1 (field.rs:131-176) and F2 (field.rs:178-222), as overloads.
inline F1 from_u32(F1*, uint32_t v) { return F1{v >= P1 ? v - P1 : v, 0}; }
inline F2 from_u32(F2*, uint32_t v) { return F2{(uint64_t)v, 0}; }
inline F1 add(F1 a, F1 o) { return F1{add1(a.re, o.re), add1(a.im, o.im)}; }
inline F2 add(F2 a, F2 o) { return F2{add2(a.re, o.re), add2(a.im, o.im)}; }
inline F1 sub(F1 a, F1 o) { return F1{sub1(a.re, o.re), sub1(a.im, o.im)}; }
inline F2 sub(F2 a, F2 o) { return F2{sub2(a.re, o.re), sub2(a.im, o.im)}; }
inline F1 mul(F1 a, F1 o) {
    return F1{sub1(mul1(a.re, o.re), mul1(a.im, o.im)), add1(mul1(a.re, o.im), mul1(a.im, o.re))};
}
inline F2 mul(F2 a, F2 o) {
    return F2{sub2(mul2(a.re, o.re), mul2(a.im, o.im)), add2(mul2(a.re, o.im), mul2(a.im, o.re))};
}
inline F1 scale(F1 a, uint32_t c) { c = c >= P1 ? c - P1 : c; return F1{mul1(a.re, c), mul1(a.im, c)}; }
inline F2 scale(F2 a, uint32_t c) { return F2{mul2(a.re, (uint64_t)c), mul2(a.im, (uint64_t)c)}; }
inline F1 inv(F1 a) {
    uint32_t n = add1(mul1(a.re, a.re), mul1(a.im, a.im));
    uint32_t ni = base_pow1(n, (uint64_t)P1 - 2);
    return F1{mul1(a.re, ni), mul1(sub1(0, a.im), ni)};
}
inline F2 inv(F2 a) {
    uint64_t n = add2(mul2(a.re, a.re), mul2(a.im, a.im));
    uint64_t ni = base_pow2(n, P2 - 2);
    return F2{mul2(a.re, ni), mul2(sub2(0, a.im), ni)};
}
template <class F> F one() { F r{}; r.re = 1; return r; }
template <class F> F unit_i() { F r{}; r.im = 1; return r; }
template <class F> F zero() { return F{}; }
template <class F> struct Order;
template <> struct Order<F1> { static constexpr uint32_t LOG2 = 32; static constexpr u128 COFACTOR = ((u128)1 << 30) - 1; };
template <> struct Order<F2> { static constexpr uint32_t LOG2 = 62; static constexpr u128 COFACTOR = ((u128)1 << 60) - 1; };

// Elem::pow (field.rs:60-71)
template <class F> F fpow(F b, u128 e) {
    F r = one<F>();
    while (e > 0) { if (e & 1) r = mul(r, b); b = mul(b, b); e >>= 1; }
    return r;
}

// find_sylow2 (field.rs:111-120) with new_from_parts (field.rs:122-129): the smallest a >= 1
// such that (a + i)^cofactor has order exactly 2^ORDER_LOG2.
template <class F> F sylow2() {
    for (uint32_t a = 1;; a++) {
        F part = add(from_u32((F*)nullptr, a), scale(unit_i<F>(), 1));
        F g = fpow(part, Order<F>::COFACTOR);
        if (fpow(g, (u128)1 << (Order<F>::LOG2 - 1)) != one<F>()) return g;
    }
}
template <class F> const F& sylow2_cached() { static const F g = sylow2<F>(); return g; }

// FieldTables::new (field.rs:237-267)
template <class F> struct FieldTables {
    std::vector<F> tw, itw, w, iw;
    explicit FieldTables(size_t n) {
        uint32_t log = 0; while (((size_t)1 << log) < n) log++;
        const F g = sylow2_cached<F>();
        F omega = fpow(g, (u128)1 << (Order<F>::LOG2 - log));
        F zeta = fpow(g, (u128)1 << (Order<F>::LOG2 - log - 2));
        if (fpow(zeta, (u128)n) != unit_i<F>()) zeta = inv(zeta);
        auto powers = [](F r, size_t m) {
            std::vector<F> v(m); F x = one<F>();
            for (size_t i = 0; i < m; i++) { v[i] = x; x = mul(x, r); }
            return v;
        };
        F ninv = inv(from_u32((F*)nullptr, (uint32_t)n));
        iw = powers(inv(zeta), n);
        for (auto& x : iw) x = mul(x, ninv);
        tw = powers(omega, n / 2);
        itw = powers(inv(omega), n / 2);
        w = powers(zeta, n);
    }
};

// Tables (field.rs:269-281)
struct Tables {
    size_t n;
    FieldTables<F1> t1;
    FieldTables<F2> t2;
    explicit Tables(size_t n_) : n(n_), t1(n_), t2(n_) {}
};
const FieldTables<F1>& ft(const Tables& t, F1*) { return t.t1; }
const FieldTables<F2>& ft(const Tables& t, F2*) { return t.t2; }

// crt_inv, crt (field.rs:283-294)
uint32_t crt_inv() { return base_pow1((uint32_t)(P2 % P1), (uint64_t)P1 - 2); }
inline u128 crt(uint32_t c1, uint64_t c2, uint32_t inv) {
    uint32_t c2m = (uint32_t)(c2 % P1);
    uint32_t h = mul1(sub1(c1, c2m), inv);
    return (u128)c2 + (u128)P2 * h;
}

// ---- cpu.rs:49-83: radix-2 transforms --------------------------------------------------
template <class F> void dif(F* a, size_t n, const std::vector<F>& tw) {
    size_t h = n / 2, s = 0;
    while (h >= 1) {
        for (size_t g = 0; g < n; g += 2 * h)
            for (size_t j = 0; j < h; j++) {
                F u = a[g + j], v = a[g + j + h];
                a[g + j] = add(u, v);
                a[g + j + h] = mul(sub(u, v), tw[j << s]);
            }
        h /= 2; s += 1;
    }
}
template <class F> void dit(F* a, size_t n, const std::vector<F>& itw) {
    size_t h = 1, s = 0;
    while (((size_t)1 << s) < n) s++;
    while (h < n) {
        s -= 1;
        for (size_t g = 0; g < n; g += 2 * h)
            for (size_t j = 0; j < h; j++) {
                F u = a[g + j];
                F v = mul(a[g + j + h], itw[j << s]);
                a[g + j] = add(u, v);
                a[g + j + h] = sub(u, v);
            }
        h *= 2;
    }
}

// forward1, inverse1 (cpu.rs:85-103)
void forward1(F1* a1, F2* a2, const Tables& t) {
    for (size_t j = 0; j < t.n; j++) { a1[j] = mul(a1[j], t.t1.w[j]); a2[j] = mul(a2[j], t.t2.w[j]); }
    dif(a1, t.n, t.t1.tw);
    dif(a2, t.n, t.t2.tw);
}
void inverse1(F1* a1, F2* a2, const Tables& t) {
    dit(a1, t.n, t.t1.itw);
    dit(a2, t.n, t.t2.itw);
    for (size_t j = 0; j < t.n; j++) { a1[j] = mul(a1[j], t.t1.iw[j]); a2[j] = mul(a2[j], t.t2.iw[j]); }
}

// forward_limbs (cpu.rs:105-115)
void forward_limbs(const uint32_t* c, size_t clen, F1* a1, F2* a2, const Tables& t) {
    size_t n = t.n;
    auto get = [&](size_t q) { return q < clen ? c[q] : 0u; };
    for (size_t j = 0; j < n; j++) {
        uint32_t re = get(j), im = get(j + n);
        a1[j] = F1{re % P1, im % P1};
        a2[j] = F2{(uint64_t)re, (uint64_t)im};
    }
    forward1(a1, a2, t);
}

// carry (cpu.rs:117-129): 2n limbs of sum c_q 2^(32q); returns the carry out of the top.
u128 carry(const F1* a1, const F2* a2, size_t n, uint32_t* out, uint32_t inv) {
    u128 c = 0;
    for (size_t q = 0; q < 2 * n; q++) {
        size_t j = q % n;
        u128 v = q < n ? crt(a1[j].re, a2[j].re, inv) : crt(a1[j].im, a2[j].im, inv);
        c += v;
        out[q] = (uint32_t)c;
        c >>= 32;
    }
    return c;
}

// pointwise1 (cpu.rs:131-140)
void pointwise1(F1* a1, F2* a2, const F1* b1, const F2* b2, size_t n, uint32_t c) {
    for (size_t j = 0; j < n; j++) {
        a1[j] = mul(a1[j], b1[j]);
        a2[j] = mul(a2[j], b2[j]);
        if (c != 1) { a1[j] = scale(a1[j], c); a2[j] = scale(a2[j], c); }
    }
}

// Conj (cpu.rs:142-163)
inline F1 conj(F1 a) { return F1{a.re, sub1(0, a.im)}; }
inline F2 conj(F2 a) { return F2{a.re, sub2(0, a.im)}; }
inline void parts(F1 a, F1& x, F1& y) { x = F1{a.re, 0}; y = F1{a.im, 0}; }
inline void parts(F2 a, F2& x, F2& y) { x = F2{a.re, 0}; y = F2{a.im, 0}; }

// Unpack (cpu.rs:165-179): u[p] = w^rev(p), iu[p] = w^-rev(p), w a primitive 2h-th root.
template <class F> struct Unpack {
    std::vector<F> u, iu;
    Unpack(size_t h, const std::vector<size_t>& revs, const F* omega) {
        uint32_t lg = 0; while (((size_t)1 << lg) < 2 * h) lg++;
        F w = fpow(sylow2_cached<F>(), (u128)1 << (Order<F>::LOG2 - lg));
        if (omega && mul(w, w) != *omega) { fprintf(stderr, "oracle: unpack root mismatch\n"); abort(); }
        u.resize(h); iu.resize(h);
        F wi = inv(w);
        for (size_t p = 0; p < h; p++) { u[p] = fpow(w, (u128)revs[p]); iu[p] = fpow(wi, (u128)revs[p]); }
    }
};

// Half (cpu.rs:181-202)
struct Half {
    std::shared_ptr<Tables> tab;  // tables at length h
    std::vector<size_t> partner;
    std::unique_ptr<Unpack<F1>> w1;
    std::unique_ptr<Unpack<F2>> w2;
    explicit Half(std::shared_ptr<Tables> t) : tab(t) {
        size_t h = t->n;
        uint32_t bits = 0; while (((size_t)1 << bits) < h) bits++;
        auto rev = [bits](size_t p) {
            size_t r = 0;
            for (uint32_t b = 0; b < bits; b++) if (p >> b & 1) r |= (size_t)1 << (bits - 1 - b);
            return r;
        };
        std::vector<size_t> revs(h);
        for (size_t p = 0; p < h; p++) revs[p] = rev(p);
        partner.resize(h);
        for (size_t p = 0; p < h; p++) partner[p] = rev((h - revs[p]) % h);
        const F1* o1 = h >= 4 ? &t->t1.tw[1] : nullptr;
        const F2* o2 = h >= 4 ? &t->t2.tw[1] : nullptr;
        w1.reset(new Unpack<F1>(h, revs, o1));
        w2.reset(new Unpack<F2>(h, revs, o2));
    }
};
const Unpack<F1>& hw(const Half& hf, F1*) { return *hf.w1; }
const Unpack<F2>& hw(const Half& hf, F2*) { return *hf.w2; }

// half_spectrum (cpu.rs:212-219)
template <class F> inline void half_spectrum(F zk, F zj, F u, F& x, F& y) {
    F mi = sub(zero<F>(), unit_i<F>());
    F e = add(zk, conj(zj));
    F wo = mul(mul(sub(zk, conj(zj)), mi), u);
    x = add(e, wo);
    y = sub(e, wo);
}
// half_one (cpu.rs:204-210)
template <class F> inline F half_one(F zk, F zj, F nx, F ny, F u, F iu) {
    F x, y;
    half_spectrum(zk, zj, u, x, y);
    F px = mul(x, nx), py = mul(y, ny);
    return add(add(px, py), mul(mul(sub(px, py), iu), unit_i<F>()));
}
// half_pointwise (cpu.rs:221-237)
template <class F> void half_pointwise(F* z, size_t h, const F* s, const std::vector<size_t>& partner, const Unpack<F>& w) {
    F x0, y0;
    parts(s[0], x0, y0);
    for (size_t p = 0; p < h; p++) {
        size_t q = partner[p];
        if (q < p) continue;
        F zp = z[p], zq = z[q];
        F nx = p == 0 ? x0 : s[p];
        F ny = p == 0 ? y0 : conj(s[q]);
        z[p] = half_one(zp, zq, nx, ny, w.u[p], w.iu[p]);
        if (q != p) z[q] = half_one(zq, zp, s[q], conj(s[p]), w.u[q], w.iu[q]);
    }
}
// half_nhat (cpu.rs:239-249)
template <class F> void half_nhat(F* z, size_t h, const std::vector<size_t>& partner, const Unpack<F>& w) {
    F c = inv(from_u32((F*)nullptr, (uint32_t)(8 * h)));
    std::vector<F> src(z, z + h);
    for (size_t p = 0; p < h; p++) {
        F x, y;
        half_spectrum(src[p], src[partner[p]], w.u[p], x, y);
        z[p] = p == 0 ? add(mul(x, c), mul(mul(y, c), unit_i<F>())) : mul(x, c);
    }
}
// half_forward (cpu.rs:286-295)
void half_forward(const uint32_t* c, F1* a1, F2* a2, const Half& hf) {
    for (size_t j = 0; j < hf.tab->n; j++) {
        uint32_t re = c[2 * j], im = c[2 * j + 1];
        a1[j] = F1{re % P1, im % P1};
        a2[j] = F2{(uint64_t)re, (uint64_t)im};
    }
    dif(a1, hf.tab->n, hf.tab->t1.tw);
    dif(a2, hf.tab->n, hf.tab->t2.tw);
}
// half_nhat_of (cpu.rs:251-259)
void half_nhat_of(const uint32_t* n, const Half& hf, std::vector<F1>& a1, std::vector<F2>& a2) {
    size_t h = hf.tab->n;
    a1.assign(h, F1{}); a2.assign(h, F2{});
    half_forward(n, a1.data(), a2.data(), hf);
    half_nhat(a1.data(), h, hf.partner, *hf.w1);
    half_nhat(a2.data(), h, hf.partner, *hf.w2);
}
// half_factors (cpu.rs:261-277): A at p, B at h + p.
template <class F> std::vector<F> half_factors(const std::vector<F>& s, const std::vector<size_t>& partner, const Unpack<F>& w) {
    size_t h = s.size();
    F i = unit_i<F>(), mi = sub(zero<F>(), i);
    std::vector<F> out(2 * h);
    for (size_t p = 0; p < h; p++) {
        F nx, ny;
        if (p == 0) parts(s[0], nx, ny); else { nx = s[p]; ny = conj(s[partner[p]]); }
        F x = mul(nx, add(one<F>(), mul(i, w.iu[p])));
        F y = mul(ny, sub(one<F>(), mul(i, w.iu[p])));
        out[p] = add(x, y);
        out[h + p] = mul(mul(sub(x, y), mi), w.u[p]);
    }
    return out;
}

// wrap_combine (cpu.rs:297-328). Returns nonzero if a carry that must vanish did not.
int wrap_combine(const F1* a1, const F2* a2, const uint32_t* t, uint32_t* x, size_t k, uint32_t inv) {
    u128 c = 0;
    for (size_t q = 0; q < k; q++) {
        F1 e1 = a1[q / 2]; F2 e2 = a2[q / 2];
        c += q % 2 == 0 ? crt(e1.re, e2.re, inv) : crt(e1.im, e2.im, inv);
        c += t[q];
        x[q] = (uint32_t)c;
        c >>= 32;
    }
    // End-around carry: 2^(32k) = 1 mod R - 1.
    while (c != 0) {
        for (size_t q = 0; q < k; q++) {
            c += x[q];
            x[q] = (uint32_t)c;
            c >>= 32;
            if (c == 0) break;
        }
    }
    for (size_t q = 0; q < k; q++) {
        c += (u128)x[q] + t[k + q];
        x[q] = (uint32_t)c;
        c >>= 32;
    }
    return c != 0;
}

// ---- Tables cache (Cpu::tables, cpu.rs:38-41) ------------------------------------------
std::mutex g_tab_mu;
std::map<size_t, std::shared_ptr<Tables>> g_tabs;
std::shared_ptr<Tables> tables_for_n(size_t n) {
    std::lock_guard<std::mutex> g(g_tab_mu);
    auto it = g_tabs.find(n);
    if (it != g_tabs.end()) return it->second;
    auto t = std::make_shared<Tables>(n);
    g_tabs[n] = t;
    return t;
}

thread_local std::string g_err;
int fail(const std::string& m) { g_err = m; return 1; }

// Parallel for over [0, count) on `threads` threads (rayon's par_iter in the reference).
template <class Fn> void par_for(size_t count, unsigned threads, Fn fn) {
    if (threads <= 1 || count <= 1) { for (size_t i = 0; i < count; i++) fn(i); return; }
    std::atomic<size_t> next(0);
    std::vector<std::thread> ts;
    unsigned nt = (unsigned)std::min<size_t>(threads, count);
    for (unsigned t = 0; t < nt; t++)
        ts.emplace_back([&] { for (size_t i; (i = next.fetch_add(1)) < count;) fn(i); });
    for (auto& t : ts) t.join();
}
unsigned default_threads() { unsigned h = std::thread::hardware_concurrency(); return h ? h : 1; }

// ---- Cpu backend (cpu.rs:409-578) --------------------------------------------------------
struct CpuConfig { int32_t wraparound; int32_t full_k; int32_t threads; };

struct Slot {  // cpu.rs:409-419
    std::vector<F1> nhat1, nphat1;
    std::vector<F2> nhat2, nphat2;
    std::vector<uint32_t> two_n, x, exp;
    uint64_t nbits = 0, pos = 0;
    uint32_t err = 0;
};

struct CpuState {  // cpu.rs:421-427
    size_t k;
    uint32_t base;
    unsigned threads;
    std::shared_ptr<Tables> tab;
    std::unique_ptr<Half> half;
    std::vector<Slot> slots;
    std::vector<uint32_t> failed;  // every slot whose carry check failed in the latest run or mul
};

struct Scratch {  // cpu.rs:429-440
    std::vector<F1> a1; std::vector<F2> a2; std::vector<uint32_t> t, s;
    explicit Scratch(size_t n) : a1(n), a2(n), t(2 * n), s(2 * n) {}
};

// redc (cpu.rs:452-484). Returns nonzero if a carry that must vanish did not.
int redc(size_t k, const Tables& tab, const Half* half, const Slot& s, Scratch& sc, uint32_t* x, uint32_t inv) {
    bool flag = false;
    for (size_t q = 0; q < k; q++) flag |= sc.t[q] != 0;
    forward_limbs(sc.t.data(), k, sc.a1.data(), sc.a2.data(), tab);
    pointwise1(sc.a1.data(), sc.a2.data(), s.nphat1.data(), s.nphat2.data(), tab.n, 1);
    inverse1(sc.a1.data(), sc.a2.data(), tab);
    carry(sc.a1.data(), sc.a2.data(), tab.n, sc.s.data(), inv);
    std::vector<uint32_t> m(sc.s.begin(), sc.s.begin() + k);
    if (half) {
        half_forward(m.data(), sc.a1.data(), sc.a2.data(), *half);
        size_t h = half->tab->n;
        half_pointwise(sc.a1.data(), h, s.nhat1.data(), half->partner, *half->w1);
        half_pointwise(sc.a2.data(), h, s.nhat2.data(), half->partner, *half->w2);
        dit(sc.a1.data(), h, half->tab->t1.itw);
        dit(sc.a2.data(), h, half->tab->t2.itw);
        return wrap_combine(sc.a1.data(), sc.a2.data(), sc.t.data(), x, k, inv);
    }
    forward_limbs(m.data(), k, sc.a1.data(), sc.a2.data(), tab);
    pointwise1(sc.a1.data(), sc.a2.data(), s.nhat1.data(), s.nhat2.data(), tab.n, 1);
    inverse1(sc.a1.data(), sc.a2.data(), tab);
    u128 top = carry(sc.a1.data(), sc.a2.data(), tab.n, sc.s.data(), inv);
    int bad = top != 0;
    for (size_t q = 2 * k; q < 2 * tab.n; q++) bad |= sc.s[q] != 0;
    uint64_t c = flag;
    for (size_t j = 0; j < k; j++) {
        c += (uint64_t)sc.t[k + j] + sc.s[k + j];
        x[j] = (uint32_t)c;
        c >>= 32;
    }
    return bad || c != 0;
}

// square_step (cpu.rs:486-497), with the squaring's carry check of the GPU path
// (kernels/src/lib.rs:908-912).
void square_step(size_t k, const Tables& tab, const Half* half, Slot& s, Scratch& sc, uint32_t c, uint32_t inv) {
    forward_limbs(s.x.data(), k, sc.a1.data(), sc.a2.data(), tab);
    for (size_t j = 0; j < tab.n; j++) {
        sc.a1[j] = scale(mul(sc.a1[j], sc.a1[j]), c);
        sc.a2[j] = scale(mul(sc.a2[j], sc.a2[j]), c);
    }
    inverse1(sc.a1.data(), sc.a2.data(), tab);
    u128 top = carry(sc.a1.data(), sc.a2.data(), tab.n, sc.t.data(), inv);
    int bad = top != 0;
    for (size_t q = 2 * k; q < 2 * tab.n; q++) bad |= sc.t[q] != 0;
    std::vector<uint32_t> x(k);
    bad |= redc(k, tab, half, s, sc, x.data(), inv);
    s.x = x;
    if (bad) s.err = 1;
}

}  // namespace

// ========================================================================================
// C entry points. Prefix oc_: the Cpu restatement; og_: the Gmp restatement.
// ========================================================================================
extern "C" {

const char* oc_last_error(void) { return g_err.c_str(); }

// Backend::create for Cpu (cpu.rs:502-533).
int oc_create(void* self, const mm_batch_setup* st, void** out) {
    const CpuConfig* cfg = (const CpuConfig*)self;
    if (st->len < 8 || (st->len & (st->len - 1))) return fail("unsupported length");
    auto S = new CpuState();
    S->k = st->k; S->base = st->base;
    S->threads = cfg->threads > 0 ? (unsigned)cfg->threads : default_threads();
    S->tab = tables_for_n(st->len / 2);
    size_t n = S->tab->n;
    if (cfg->wraparound) {
        if (st->k != n) { delete S; return fail("wraparound needs k = L/2"); }
        S->half.reset(new Half(tables_for_n(st->len / 4)));
    }
    S->slots.resize(st->slots);
    par_for(st->slots, S->threads, [&](size_t i) {
        Slot& s = S->slots[i];
        const uint32_t* ni = st->n + i * st->k;
        const uint32_t* npi = st->nprime + i * st->k;
        Scratch sc(n);
        if (S->half) {
            half_nhat_of(ni, *S->half, s.nhat1, s.nhat2);
        } else {
            forward_limbs(ni, st->k, sc.a1.data(), sc.a2.data(), *S->tab);
            s.nhat1 = sc.a1; s.nhat2 = sc.a2;
        }
        forward_limbs(npi, st->k, sc.a1.data(), sc.a2.data(), *S->tab);
        s.nphat1 = sc.a1; s.nphat2 = sc.a2;
        s.two_n.assign(st->two_n + i * st->k, st->two_n + (i + 1) * st->k);
        s.x.assign(st->k, 0);
    });
    *out = S;
    return 0;
}
void oc_destroy(void* st) { delete (CpuState*)st; }
// The table's set_modulus (an extension of this library's backend table, no counterpart in
// cpu.rs): the slot's spectra as oc_create computes them, exponent cleared.
int oc_set_modulus(void* st, size_t i, const uint32_t* n, const uint32_t* nprime, const uint32_t* two_n) {
    CpuState* S = (CpuState*)st;
    if (i >= S->slots.size()) return fail("slot out of range");
    Slot& s = S->slots[i];
    size_t nn = S->tab->n;
    Scratch sc(nn);
    if (S->half) {
        half_nhat_of(n, *S->half, s.nhat1, s.nhat2);
    } else {
        forward_limbs(n, S->k, sc.a1.data(), sc.a2.data(), *S->tab);
        s.nhat1 = sc.a1; s.nhat2 = sc.a2;
    }
    forward_limbs(nprime, S->k, sc.a1.data(), sc.a2.data(), *S->tab);
    s.nphat1 = sc.a1; s.nphat2 = sc.a2;
    s.two_n.assign(two_n, two_n + S->k);
    s.exp.clear();
    s.nbits = s.pos = 0;
    return 0;
}
int oc_load(void* st, size_t i, const uint32_t* limbs) {  // cpu.rs:534-536
    CpuState* S = (CpuState*)st;
    if (i >= S->slots.size()) return fail("slot out of range");
    S->slots[i].x.assign(limbs, limbs + S->k);
    return 0;
}
int oc_set_exponent(void* st, size_t i, const uint32_t* words, size_t nwords, uint64_t nbits) {  // cpu.rs:537-542
    CpuState* S = (CpuState*)st;
    if (i >= S->slots.size()) return fail("slot out of range");
    if (nbits > 32 * (uint64_t)nwords) return fail("nbits exceeds the exponent words");
    Slot& s = S->slots[i];
    s.exp.assign(words, words + nwords);
    s.nbits = nbits;
    s.pos = 0;
    return 0;
}
int64_t oc_run(void* st, uint64_t bits) {  // cpu.rs:543-557
    CpuState* S = (CpuState*)st;
    uint32_t inv = crt_inv();
    const Tables& tab = *S->tab;
    const Half* half = S->half.get();
    par_for(S->slots.size(), S->threads, [&](size_t i) {
        Slot& s = S->slots[i];
        Scratch sc(tab.n);
        for (uint64_t b = 0; b < bits; b++) {
            if (s.pos >= s.nbits) break;
            uint32_t bit = (s.exp[s.pos / 32] >> (31 - s.pos % 32)) & 1;
            square_step(S->k, tab, half, s, sc, bit ? S->base : 1, inv);
            s.pos += 1;
        }
    });
    // The reference checks these carries with debug assertions (cpu.rs:474-483) and its GPU
    // backend panics naming the first failing slot (gpu/mod.rs:179-184): run returns that slot,
    // and every failing slot is kept for oc_failed_slots.
    S->failed.clear();
    for (size_t i = 0; i < S->slots.size(); i++)
        if (S->slots[i].err) S->failed.push_back((uint32_t)i);
    for (auto& s : S->slots) s.err = 0;
    if (!S->failed.empty()) {
        g_err = "run: carry out of a product that must fit, slot " + std::to_string(S->failed[0]);
        return (int64_t)S->failed[0];
    }
    return MM_OK;
}
size_t oc_failed_slots(void* st, uint32_t* out, size_t max) {
    const CpuState* S = (const CpuState*)st;
    for (size_t i = 0; i < S->failed.size() && i < max; i++) out[i] = S->failed[i];
    return S->failed.size();
}
int64_t oc_mul(void* st, size_t out, size_t a, size_t b) {  // cpu.rs:558-571
    CpuState* S = (CpuState*)st;
    if (out >= S->slots.size() || a >= S->slots.size() || b >= S->slots.size()) { fail("slot out of range"); return MM_EFAULT; }
    const Tables& tab = *S->tab;
    uint32_t inv = crt_inv();
    Scratch sc(tab.n);
    std::vector<F1> b1(tab.n); std::vector<F2> b2(tab.n);
    forward_limbs(S->slots[b].x.data(), S->k, b1.data(), b2.data(), tab);
    forward_limbs(S->slots[a].x.data(), S->k, sc.a1.data(), sc.a2.data(), tab);
    pointwise1(sc.a1.data(), sc.a2.data(), b1.data(), b2.data(), tab.n, 1);
    inverse1(sc.a1.data(), sc.a2.data(), tab);
    u128 top = carry(sc.a1.data(), sc.a2.data(), tab.n, sc.t.data(), inv);
    int bad = top != 0;
    for (size_t q = 2 * S->k; q < 2 * tab.n; q++) bad |= sc.t[q] != 0;
    std::vector<uint32_t> x(S->k);
    bad |= redc(S->k, tab, S->half.get(), S->slots[out], sc, x.data(), inv);
    S->slots[out].x = x;
    S->failed.clear();
    if (bad) {
        S->failed.push_back((uint32_t)out);
        g_err = "mul: carry out of a product that must fit, slot " + std::to_string(out);
        return (int64_t)out;
    }
    return MM_OK;
}
int oc_read(void* st, size_t i, uint32_t* out) {  // cpu.rs:572-574
    CpuState* S = (CpuState*)st;
    if (i >= S->slots.size()) return fail("slot out of range");
    memcpy(out, S->slots[i].x.data(), S->k * 4);
    return 0;
}
uint64_t oc_remaining(void* st, size_t i) {  // cpu.rs:575-577
    CpuState* S = (CpuState*)st;
    return S->slots[i].nbits - S->slots[i].pos;
}

// Limits for Cpu (cpu.rs:337-347): bound Q, lengths 2^3..2^24, full_k = full_k || wraparound.
void oc_limits(const void* self, uint64_t* bound_lo, uint64_t* bound_hi, uint32_t* min_len, uint32_t* max_len, int32_t* full_k) {
    const CpuConfig* cfg = (const CpuConfig*)self;
    *bound_lo = (uint64_t)Q; *bound_hi = (uint64_t)(Q >> 64);
    *min_len = 8; *max_len = 1u << 24;
    *full_k = cfg->full_k || cfg->wraparound;
}

// ---- Transform layer (cpu.rs:349-407); f1 holds (re, im) u32 pairs, f2 (re, im) u64 pairs.
void oc_t_upload(const uint32_t* coeffs, size_t len, size_t count, uint32_t* f1, uint64_t* f2) {
    size_t n = len / 2;
    for (size_t v = 0; v < count; v++) {
        const uint32_t* c = coeffs + v * len;
        for (size_t j = 0; j < n; j++) {
            f1[2 * (v * n + j)] = c[j] % P1; f1[2 * (v * n + j) + 1] = c[j + n] % P1;
            f2[2 * (v * n + j)] = c[j]; f2[2 * (v * n + j) + 1] = c[j + n];
        }
    }
}
void oc_t_download(const uint32_t* f1, const uint64_t* f2, size_t len, size_t count, uint64_t* out) {
    size_t n = len / 2;
    uint32_t inv = crt_inv();
    for (size_t v = 0; v < count; v++)
        for (size_t j = 0; j < n; j++) {
            size_t e = v * n + j;
            u128 lo = crt(f1[2 * e], f2[2 * e], inv), hi = crt(f1[2 * e + 1], f2[2 * e + 1], inv);
            size_t q0 = v * len + j, q1 = v * len + j + n;
            out[2 * q0] = (uint64_t)lo; out[2 * q0 + 1] = (uint64_t)(lo >> 64);
            out[2 * q1] = (uint64_t)hi; out[2 * q1 + 1] = (uint64_t)(hi >> 64);
        }
}
void oc_t_forward(uint32_t* f1, uint64_t* f2, size_t len, size_t count) {
    auto t = tables_for_n(len / 2);
    size_t n = len / 2;
    par_for(count, default_threads(), [&](size_t v) { forward1((F1*)f1 + v * n, (F2*)f2 + v * n, *t); });
}
void oc_t_inverse(uint32_t* f1, uint64_t* f2, size_t len, size_t count) {
    auto t = tables_for_n(len / 2);
    size_t n = len / 2;
    par_for(count, default_threads(), [&](size_t v) { inverse1((F1*)f1 + v * n, (F2*)f2 + v * n, *t); });
}
void oc_t_pointwise(uint32_t* o1, uint64_t* o2, const uint32_t* a1, const uint64_t* a2, const uint32_t* b1,
                    const uint64_t* b2, const uint32_t* scales, size_t len, size_t count) {
    size_t n = len / 2;
    for (size_t v = 0; v < count; v++) {
        uint32_t c = scales ? scales[v] : 1;
        memcpy(o1 + 2 * v * n, a1 + 2 * v * n, n * 8);
        memcpy(o2 + 2 * v * n, a2 + 2 * v * n, n * 16);
        pointwise1((F1*)o1 + v * n, (F2*)o2 + v * n, (const F1*)b1 + v * n, (const F2*)b2 + v * n, n, c);
    }
}

// ---- Pieces exposed for pinning tests --------------------------------------------------
void oc_sylow2(uint32_t* g1, uint64_t* g2) {
    F1 a = sylow2_cached<F1>(); F2 b = sylow2_cached<F2>();
    g1[0] = a.re; g1[1] = a.im; g2[0] = b.re; g2[1] = b.im;
}
// Tables at complex length n: tw, itw (n/2 each), w, iw (n each), for both fields.
void oc_tables(size_t n, uint32_t* tw1, uint32_t* itw1, uint32_t* w1, uint32_t* iw1, uint64_t* tw2, uint64_t* itw2,
               uint64_t* w2, uint64_t* iw2) {
    auto t = tables_for_n(n);
    memcpy(tw1, t->t1.tw.data(), n / 2 * 8); memcpy(itw1, t->t1.itw.data(), n / 2 * 8);
    memcpy(w1, t->t1.w.data(), n * 8); memcpy(iw1, t->t1.iw.data(), n * 8);
    memcpy(tw2, t->t2.tw.data(), n / 2 * 16); memcpy(itw2, t->t2.itw.data(), n / 2 * 16);
    memcpy(w2, t->t2.w.data(), n * 16); memcpy(iw2, t->t2.iw.data(), n * 16);
}
uint32_t oc_crt_inv(void) { return crt_inv(); }
void oc_crt(uint32_t c1, uint64_t c2, uint64_t* out) { u128 v = crt(c1, c2, crt_inv()); out[0] = (uint64_t)v; out[1] = (uint64_t)(v >> 64); }
void oc_mul1(uint32_t a, uint32_t b, uint32_t* r) { *r = mul1(a, b); }
void oc_mul2(uint64_t a, uint64_t b, uint64_t* r) { *r = mul2(a, b); }
void oc_cmul1(const uint32_t* a, const uint32_t* b, uint32_t* r) { F1 x = mul(F1{a[0], a[1]}, F1{b[0], b[1]}); r[0] = x.re; r[1] = x.im; }
void oc_cmul2(const uint64_t* a, const uint64_t* b, uint64_t* r) { F2 x = mul(F2{a[0], a[1]}, F2{b[0], b[1]}); r[0] = x.re; r[1] = x.im; }
// The half path's pieces for k = 2h limbs of N: partner positions, N-hat as stored by
// half_nhat, and the A|B factors of half_factors (2h elements per field).
void oc_half_partner(size_t h, uint64_t* out) {
    Half hf(tables_for_n(h));
    for (size_t p = 0; p < h; p++) out[p] = hf.partner[p];
}
void oc_half_factors(const uint32_t* n_limbs, size_t h, uint32_t* f1, uint64_t* f2) {
    Half hf(tables_for_n(h));
    std::vector<F1> s1; std::vector<F2> s2;
    half_nhat_of(n_limbs, hf, s1, s2);
    auto a = half_factors(s1, hf.partner, *hf.w1);
    auto b = half_factors(s2, hf.partner, *hf.w2);
    memcpy(f1, a.data(), 2 * h * 8);
    memcpy(f2, b.data(), 2 * h * 16);
}
// Unweighted DIF of length n (the half path's forward transform) on one vector per field.
void oc_dif(uint32_t* f1, uint64_t* f2, size_t n) {
    auto t = tables_for_n(n);
    dif((F1*)f1, n, t->t1.tw);
    dif((F2*)f2, n, t->t2.tw);
}

// ---- Gmp backend (gmp.rs:1-122) ------------------------------------------------------------
struct GmpSlot {  // gmp.rs:304-316
    mpz_t n, nprime, two_n, x, t, m;
    std::vector<uint32_t> exp;
    uint64_t nbits = 0, pos = 0;
    GmpSlot() { mpz_inits(n, nprime, two_n, x, t, m, nullptr); }
    ~GmpSlot() { mpz_clears(n, nprime, two_n, x, t, m, nullptr); }
    // Slot::redc (gmp.rs:45-55): m = (T mod R) N' mod R, x = (T + m N) / R.
    void redc(mp_bitcnt_t rbits) {
        mpz_fdiv_r_2exp(m, t, rbits);
        mpz_mul(m, m, nprime);
        mpz_fdiv_r_2exp(m, m, rbits);
        mpz_mul(m, m, n);
        mpz_add(t, t, m);
        mpz_fdiv_q_2exp(t, t, rbits);
        mpz_swap(x, t);
    }
    // Slot::square_step (gmp.rs:57-61)
    void square_step(uint32_t c, mp_bitcnt_t rbits) {
        mpz_mul(t, x, x);
        mpz_mul_ui(t, t, c);
        redc(rbits);
    }
};
struct GmpState { size_t k; uint32_t base; unsigned threads; std::vector<std::unique_ptr<GmpSlot>> slots; };

static void limbs_to_mpz(mpz_t z, const uint32_t* l, size_t k) { mpz_import(z, k, -1, 4, 0, 0, l); }
// to_limbs (batch.rs:34-39): fails when the value does not fit k limbs.
static int mpz_to_limbs(const mpz_t z, uint32_t* l, size_t k) {
    size_t cnt = 0;
    memset(l, 0, k * 4);
    if (mpz_sgn(z) < 0 || mpz_sizeinbase(z, 2) > 32 * k) return fail("value does not fit " + std::to_string(k) + " limbs");
    mpz_export(l, &cnt, -1, 4, 0, 0, z);
    return 0;
}

struct GmpConfig { int32_t threads; int32_t full_k; };

int og_create(void* self, const mm_batch_setup* st, void** out) {  // gmp.rs:73-88
    const GmpConfig* cfg = (const GmpConfig*)self;
    auto S = new GmpState();
    S->k = st->k; S->base = st->base;
    S->threads = cfg && cfg->threads > 0 ? (unsigned)cfg->threads : default_threads();
    for (size_t i = 0; i < st->slots; i++) {
        S->slots.emplace_back(new GmpSlot());
        GmpSlot& s = *S->slots.back();
        limbs_to_mpz(s.n, st->n + i * st->k, st->k);
        limbs_to_mpz(s.nprime, st->nprime + i * st->k, st->k);
        limbs_to_mpz(s.two_n, st->two_n + i * st->k, st->k);
    }
    *out = S;
    return 0;
}
void og_destroy(void* st) { delete (GmpState*)st; }
// set_modulus for the Gmp restatement (extension, as oc_set_modulus).
int og_set_modulus(void* st, size_t i, const uint32_t* n, const uint32_t* nprime, const uint32_t* two_n) {
    GmpState* S = (GmpState*)st;
    if (i >= S->slots.size()) return fail("slot out of range");
    GmpSlot& s = *S->slots[i];
    limbs_to_mpz(s.n, n, S->k);
    limbs_to_mpz(s.nprime, nprime, S->k);
    limbs_to_mpz(s.two_n, two_n, S->k);
    s.exp.clear();
    s.nbits = s.pos = 0;
    return 0;
}
int og_load(void* st, size_t i, const uint32_t* limbs) {  // gmp.rs:89-91
    GmpState* S = (GmpState*)st;
    if (i >= S->slots.size()) return fail("slot out of range");
    limbs_to_mpz(S->slots[i]->x, limbs, S->k);
    return 0;
}
int og_set_exponent(void* st, size_t i, const uint32_t* words, size_t nwords, uint64_t nbits) {  // gmp.rs:92-97
    GmpState* S = (GmpState*)st;
    if (i >= S->slots.size()) return fail("slot out of range");
    GmpSlot& s = *S->slots[i];
    s.exp.assign(words, words + nwords);
    s.nbits = nbits; s.pos = 0;
    return 0;
}
int64_t og_run(void* st, uint64_t bits) {  // gmp.rs:98-109
    GmpState* S = (GmpState*)st;
    mp_bitcnt_t rbits = 32 * S->k;
    par_for(S->slots.size(), S->threads, [&](size_t i) {
        GmpSlot& s = *S->slots[i];
        for (uint64_t b = 0; b < bits; b++) {
            if (s.pos >= s.nbits) break;
            uint32_t bit = (s.exp[s.pos / 32] >> (31 - s.pos % 32)) & 1;
            s.square_step(bit ? S->base : 1, rbits);
            s.pos += 1;
        }
    });
    return MM_OK;
}
int64_t og_mul(void* st, size_t out, size_t a, size_t b) {  // gmp.rs:110-116
    GmpState* S = (GmpState*)st;
    mpz_t t; mpz_init(t);
    mpz_mul(t, S->slots[a]->x, S->slots[b]->x);
    GmpSlot& s = *S->slots[out];
    mpz_swap(s.t, t);
    mpz_clear(t);
    s.redc(32 * S->k);
    return MM_OK;
}
int og_read(void* st, size_t i, uint32_t* out) {  // gmp.rs:117-119
    GmpState* S = (GmpState*)st;
    if (i >= S->slots.size()) return fail("slot out of range");
    return mpz_to_limbs(S->slots[i]->x, out, S->k);
}
uint64_t og_remaining(void* st, size_t i) {  // gmp.rs:120-122
    GmpState* S = (GmpState*)st;
    return S->slots[i]->nbits - S->slots[i]->pos;
}

// One closed-form step on raw limbs: x' = (x^2 c + m N) / R with m = ((x^2 c) mod R) N' mod R,
// R = 2^(32k). All arrays hold k limbs. Used to pin individual steps.
void og_step(const uint32_t* x, const uint32_t* n, const uint32_t* nprime, size_t k, uint32_t c, uint32_t* out) {
    GmpSlot s;
    limbs_to_mpz(s.x, x, k); limbs_to_mpz(s.n, n, k); limbs_to_mpz(s.nprime, nprime, k);
    s.square_step(c, 32 * k);
    mpz_to_limbs(s.x, out, k);
}

// GMP's powm, product mod m and invert on little-endian u32 limbs: the tests' expected values at sizes where
// Python integers are slow (the reference's tests take them from rug's pow_mod and invert).
// out holds nm limbs. og_invert returns 0 when a has no inverse mod m.
void og_powm(const uint32_t* b, size_t nb, const uint32_t* e, size_t ne, const uint32_t* m, size_t nm, uint32_t* out) {
    mpz_t B, E, M, R;
    mpz_inits(B, E, M, R, nullptr);
    limbs_to_mpz(B, b, nb); limbs_to_mpz(E, e, ne); limbs_to_mpz(M, m, nm);
    mpz_powm(R, B, E, M);
    mpz_to_limbs(R, out, nm);
    mpz_clears(B, E, M, R, nullptr);
}

void og_mulmod(const uint32_t* a, size_t na, const uint32_t* b, size_t nb, const uint32_t* m, size_t nm, uint32_t* out) {
    mpz_t A, B, M;
    mpz_inits(A, B, M, nullptr);
    limbs_to_mpz(A, a, na); limbs_to_mpz(B, b, nb); limbs_to_mpz(M, m, nm);
    mpz_mul(A, A, B);
    mpz_mod(A, A, M);
    mpz_to_limbs(A, out, nm);
    mpz_clears(A, B, M, nullptr);
}

int og_invert(const uint32_t* a, size_t na, const uint32_t* m, size_t nm, uint32_t* out) {
    mpz_t A, M, R;
    mpz_inits(A, M, R, nullptr);
    limbs_to_mpz(A, a, na); limbs_to_mpz(M, m, nm);
    const int ok = mpz_invert(R, A, M) != 0;
    if (ok) mpz_to_limbs(R, out, nm);
    mpz_clears(A, M, R, nullptr);
    return ok;
}

}  // extern "C"
