// mm_host.cpp - the host API over any backend table: ModBatch, fermat, Params and the score
// workload (montmul/src/batch.rs, montmul/src/score.rs:45-61).
//
// The reference's `ModBatch<B: Backend>` is generic over its backend; here the backend is an
// `mm_backend` function table, so this one host drives the HIP backend of this library or any
// other implementation of the table. Big-integer setup (N', R mod N, R^-1 mod N) uses GMP, as
// the reference does through `rug` (batch.rs:84-94), one thread per modulus group.

#include <gmp.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/montmul_hip.h"

namespace {

thread_local std::string g_err;
int fail(const std::string& m) { g_err = m; return 1; }

struct Mpz {
    mpz_t v;
    Mpz() { mpz_init(v); }
    Mpz(const Mpz& o) { mpz_init_set(v, o.v); }
    Mpz& operator=(const Mpz& o) { mpz_set(v, o.v); return *this; }
    ~Mpz() { mpz_clear(v); }
};

void from_signed_limbs(mpz_t z, int sign, const uint32_t* l, size_t nl) {
    mpz_import(z, nl, -1, 4, 0, 0, l);
    if (sign < 0) mpz_neg(z, z);
}

// to_limbs (batch.rs:34-39): k little-endian limbs; fails when the value does not fit.
bool to_limbs(const mpz_t z, uint32_t* out, size_t k) {
    memset(out, 0, k * 4);
    if (mpz_sgn(z) < 0 || mpz_sizeinbase(z, 2) > 32 * k) return false;
    if (mpz_sgn(z) == 0) return true;
    size_t cnt = 0;
    mpz_export(out, &cnt, -1, 4, 0, 0, z);
    return true;
}

size_t significant_bits(const mpz_t z) { return mpz_sgn(z) == 0 ? 0 : mpz_sizeinbase(z, 2); }

template <class Fn> void par_for(size_t count, Fn fn) {
    unsigned h = std::thread::hardware_concurrency();
    unsigned nt = (unsigned)std::min<size_t>(h ? h : 1, count);
    if (nt <= 1) { for (size_t i = 0; i < count; i++) fn(i); return; }
    std::atomic<size_t> next(0);
    std::vector<std::thread> ts;
    for (unsigned t = 0; t < nt; t++)
        ts.emplace_back([&] { for (size_t i; (i = next.fetch_add(1)) < count;) fn(i); });
    for (auto& t : ts) t.join();
}

typedef unsigned __int128 u128;

// The Montgomery constants of an odd modulus N > 1 for R = 2^(32k) (batch.rs:84-91):
// N' = -N^-1 mod R, R mod N and R^-1 mod N. The reference inverts twice (N mod R, then R mod N);
// here neither needs an extended gcd:
// - N^-1 mod 2^p comes from Newton's iteration x <- x (2 - N x) mod 2^p, which doubles the
//   count of correct low bits (Hensel lifting): from 64 bits (five rounds on one machine word)
//   up a ladder of precisions ceil(32k / 2^j), so the last round works at exactly 32k bits.
// - With N N^-1 = 1 + q R (0 <= q < N), 1 + N N' = (N - q) R, so R^-1 mod N = (1 + N N') / R,
//   one product and a shift (this is REDC(1)).
// Returns false when N is even or not below R.
bool montgomery_constants(const mpz_t n, uint32_t k, mpz_t nprime, mpz_t r_mod_n, mpz_t rinv) {
    const mp_bitcnt_t bits = 32 * (mp_bitcnt_t)k;
    if (mpz_sgn(n) <= 0 || mpz_even_p(n) || mpz_sizeinbase(n, 2) > bits) return false;
    const uint64_t n0 = (uint64_t)mpz_getlimbn(n, 0);
    uint64_t x0 = n0;  // N N = 1 mod 8: three bits to start
    for (int i = 0; i < 5; i++) x0 *= 2 - n0 * x0;
    std::vector<mp_bitcnt_t> ladder;
    for (mp_bitcnt_t p = bits; p > 64; p = (p + 1) / 2) ladder.push_back(p);
    Mpz x, t, nlow;
    mpz_set_ui(x.v, (unsigned long)x0);
    if (bits < 64) mpz_fdiv_r_2exp(x.v, x.v, bits);
    for (size_t j = ladder.size(); j-- > 0;) {
        const mp_bitcnt_t p = ladder[j];
        mpz_fdiv_r_2exp(nlow.v, n, p);
        mpz_mul(t.v, nlow.v, x.v);
        mpz_fdiv_r_2exp(t.v, t.v, p);
        mpz_ui_sub(t.v, 2, t.v);
        mpz_mul(x.v, x.v, t.v);
        mpz_fdiv_r_2exp(x.v, x.v, p);
    }
    // nprime = R - x (x is odd, so nonzero)
    mpz_set_ui(t.v, 0);
    mpz_setbit(t.v, bits);
    mpz_sub(nprime, t.v, x.v);
    mpz_mod(r_mod_n, t.v, n);
    mpz_mul(rinv, n, nprime);
    mpz_add_ui(rinv, rinv, 1);
    mpz_fdiv_q_2exp(rinv, rinv, bits);
    return true;
}

}  // namespace

// Modulus (batch.rs:44-50).
struct MmModulus { Mpz n, r, rinv; };

struct mm_modbatch {  // batch.rs:53-58
    mm_backend be;
    void* state = nullptr;
    uint32_t k = 0, len = 0;
    std::vector<MmModulus> moduli;
    std::vector<uint32_t> failed;  // the latest run's or mul's failing slot, for backends without failed_slots
    ~mm_modbatch() { if (state) be.destroy(state); }
};

namespace {

// The two raw values that stand for 1 in slot i (ModBatch::is_one, batch.rs:179-183): R mod N
// and R mod N + N, k limbs each, into out[0 .. 2k).
bool ones_of(const MmModulus& m, uint32_t k, uint32_t* out) {
    Mpz r2;
    mpz_add(r2.v, m.r.v, m.n.v);
    return to_limbs(m.r.v, out, k) && to_limbs(r2.v, out + k, k);
}

// Hands slots [first, first + count)'s values of 1 to a backend that tests is_one on the device,
// 256 slots per call.
int keep_ones(mm_modbatch* b, size_t first, size_t count) {
    if (!b->be.set_ones) return 0;
    const size_t w = 2ull * b->k;
    for (size_t a = first; a < first + count; a += 256) {
        const size_t c = std::min<size_t>(256, first + count - a);
        std::vector<uint32_t> t(c * w);
        std::atomic<int> bad(0);
        par_for(c, [&](size_t j) {
            if (!ones_of(b->moduli[a + j], b->k, &t[j * w])) bad = 1;
        });
        if (bad) return fail("R mod N + N does not fit " + std::to_string(b->k) + " limbs");
        if (b->be.set_ones(b->state, a, c, t.data())) return fail(b->be.last_error());
    }
    return 0;
}

}  // namespace

extern "C" {

const char* mm_host_last_error(void) { return g_err.c_str(); }

// Params::for_bits (batch.rs:20-25).
int mm_params_for_bits(uint32_t bits, uint32_t min_len, uint32_t max_len, uint32_t* k, uint32_t* len) {
    uint64_t kk = ((uint64_t)bits + 4 + 31) / 32;
    for (uint64_t l = min_len; l && l <= max_len; l *= 2)
        if (l >= 2 * kk) { *k = (uint32_t)kk; *len = (uint32_t)l; return 0; }
    return fail("no transform length for " + std::to_string(bits) + " bits");
}

// Params::check_exact (batch.rs:27-31): max(c, 1) * k * (2^32 - 1)^2 < bound, in exact
// arithmetic (the left side can exceed 128 bits only for k far beyond any supported length).
int mm_params_check_exact(uint64_t k, uint32_t c, uint64_t bound_lo, uint64_t bound_hi) {
    Mpz lhs, bound, t;
    mpz_set_ui(lhs.v, c ? c : 1);
    mpz_mul_ui(lhs.v, lhs.v, k);
    mpz_set_ui(t.v, 0xffffffffu);
    mpz_mul(t.v, t.v, t.v);
    mpz_mul(lhs.v, lhs.v, t.v);
    mpz_set_ui(bound.v, bound_hi);
    mpz_mul_2exp(bound.v, bound.v, 64);
    mpz_add_ui(bound.v, bound.v, bound_lo);
    if (mpz_cmp(lhs.v, bound.v) < 0) return 0;
    char* a = mpz_get_str(nullptr, 10, lhs.v);
    char* b = mpz_get_str(nullptr, 10, bound.v);
    std::string m = std::string("c * k * (2^w - 1)^2 = ") + a + " exceeds the coefficient bound " + b;
    void (*freefn)(void*, size_t);
    mp_get_memory_functions(nullptr, nullptr, &freefn);
    freefn(a, strlen(a) + 1);
    freefn(b, strlen(b) + 1);
    return fail(m);
}

// ModBatch::new (batch.rs:62-105).
int mm_modbatch_new(const mm_backend* be, size_t count, const uint32_t* limbs, const size_t* offsets,
                    const int8_t* signs, uint32_t base, mm_modbatch** out) {
    *out = nullptr;
    if (count == 0) return fail("empty batch");
    if (base < 1 || base > 4)
        return fail("base " + std::to_string(base) + " does not fit the lazy range (c <= 4)");
    std::vector<Mpz> ns(count);
    uint32_t k = 0, len = 0;
    for (size_t i = 0; i < count; i++) {
        from_signed_limbs(ns[i].v, signs[i], limbs + offsets[i], offsets[i + 1] - offsets[i]);
        if (mpz_sgn(ns[i].v) <= 0 || mpz_even_p(ns[i].v) || mpz_cmp_ui(ns[i].v, 1) <= 0)
            return fail("modulus must be odd and > 1");
        size_t bits = significant_bits(ns[i].v);
        if (bits > 0xffffffffu) return fail("modulus too large");
        uint32_t pk, pl;
        if (mm_params_for_bits((uint32_t)bits, be->min_len, be->max_len, &pk, &pl)) return 1;
        if (i == 0) { k = pk; len = pl; continue; }
        if (pl != len)
            return fail("moduli need different transform lengths (" + std::to_string(len) + " vs " +
                        std::to_string(pl) + ")");
        k = std::max(k, pk);
    }
    if (be->full_k) k = len / 2;
    if (mm_params_check_exact(k, base, be->bound_lo, be->bound_hi)) return 1;

    auto B = new mm_modbatch();
    B->be = *be; B->k = k; B->len = len;
    B->moduli.resize(count);
    std::vector<uint32_t> nl(count * (size_t)k), npl(count * (size_t)k), tnl(count * (size_t)k);
    std::atomic<int> bad(0);
    // Two modular inversions per modulus dominate the setup of large batches.
    par_for(count, [&](size_t i) {
        Mpz np, two;
        const mpz_t& n = ns[i].v;
        MmModulus& m = B->moduli[i];
        mpz_set(m.n.v, n);
        if (!montgomery_constants(n, k, np.v, m.r.v, m.rinv.v)) { bad = 1; return; }
        mpz_mul_2exp(two.v, n, 1);
        if (!to_limbs(n, &nl[i * k], k) || !to_limbs(np.v, &npl[i * k], k) || !to_limbs(two.v, &tnl[i * k], k))
            bad = 1;
    });
    if (bad) { delete B; return fail("value does not fit " + std::to_string(k) + " limbs"); }
    mm_batch_setup st{k, len, base, (uint32_t)count, nl.data(), npl.data(), tnl.data()};
    if (be->create(be->self, &st, &B->state)) {
        g_err = be->last_error();
        B->state = nullptr;
        delete B;
        return 1;
    }
    if (keep_ones(B, 0, count)) {
        delete B;
        return 1;
    }
    *out = B;
    return 0;
}

// Admission of another modulus into slot i (no reference counterpart; the reference builds a
// new ModBatch): the setup values of mm_modbatch_new for one modulus, then the backend's
// set_modulus.
int mm_modbatch_set_modulus(mm_modbatch* b, size_t i, int sign, const uint32_t* limbs, size_t nl) {
    if (i >= b->moduli.size()) return fail("slot " + std::to_string(i) + " out of range");
    if (!b->be.set_modulus) return fail("this backend cannot replace a slot's modulus");
    Mpz n, np, two;
    from_signed_limbs(n.v, sign, limbs, nl);
    if (mpz_sgn(n.v) <= 0 || mpz_even_p(n.v) || mpz_cmp_ui(n.v, 1) <= 0) return fail("modulus must be odd and > 1");
    uint32_t pk, pl;
    size_t bits = significant_bits(n.v);
    if (bits > 0xffffffffu || mm_params_for_bits((uint32_t)bits, b->be.min_len, b->be.max_len, &pk, &pl)) return 1;
    if (pl != b->len || pk > b->k)
        return fail("modulus needs transform length " + std::to_string(pl) + ", the batch has " + std::to_string(b->len));
    const uint32_t k = b->k;
    MmModulus m;
    mpz_set(m.n.v, n.v);
    if (!montgomery_constants(n.v, k, np.v, m.r.v, m.rinv.v)) return fail("modulus not invertible mod R");
    mpz_mul_2exp(two.v, n.v, 1);
    std::vector<uint32_t> nl_(k), npl(k), tnl(k);
    if (!to_limbs(n.v, nl_.data(), k) || !to_limbs(np.v, npl.data(), k) || !to_limbs(two.v, tnl.data(), k))
        return fail("value does not fit " + std::to_string(k) + " limbs");
    if (b->be.set_modulus(b->state, i, nl_.data(), npl.data(), tnl.data())) return fail(b->be.last_error());
    mpz_set(b->moduli[i].n.v, m.n.v);
    mpz_set(b->moduli[i].r.v, m.r.v);
    mpz_set(b->moduli[i].rinv.v, m.rinv.v);
    return keep_ones(b, i, 1);
}

void mm_modbatch_free(mm_modbatch* b) { delete b; }

void mm_modbatch_params(const mm_modbatch* b, uint32_t* k, uint32_t* len) { *k = b->k; *len = b->len; }
size_t mm_modbatch_slots(const mm_modbatch* b) { return b->moduli.size(); }
void* mm_modbatch_state(mm_modbatch* b) { return b->state; }

static int check_slot(const mm_modbatch* b, size_t i) {
    if (i >= b->moduli.size()) return fail("slot " + std::to_string(i) + " out of range");
    return 0;
}

// ModBatch::load (batch.rs:125-130).
int mm_modbatch_load(mm_modbatch* b, size_t i, int sign, const uint32_t* limbs, size_t nl) {
    if (check_slot(b, i)) return 1;
    const MmModulus& m = b->moduli[i];
    Mpz x;
    from_signed_limbs(x.v, sign, limbs, nl);
    mpz_mul(x.v, x.v, m.r.v);
    mpz_mod(x.v, x.v, m.n.v);
    std::vector<uint32_t> l(b->k);
    to_limbs(x.v, l.data(), b->k);
    if (b->be.load(b->state, i, l.data())) return fail(b->be.last_error());
    return 0;
}

// ModBatch::load_raw (batch.rs:133-137).
int mm_modbatch_load_raw(mm_modbatch* b, size_t i, int sign, const uint32_t* limbs, size_t nl) {
    if (check_slot(b, i)) return 1;
    Mpz x, two;
    from_signed_limbs(x.v, sign, limbs, nl);
    mpz_mul_2exp(two.v, b->moduli[i].n.v, 1);
    if (mpz_sgn(x.v) < 0 || mpz_cmp(x.v, two.v) >= 0) return fail("raw value outside [0, 2N)");
    std::vector<uint32_t> l(b->k);
    to_limbs(x.v, l.data(), b->k);
    if (b->be.load(b->state, i, l.data())) return fail(b->be.last_error());
    return 0;
}

// ModBatch::set_exponent (batch.rs:140-149): bit p of the words is bit nbits - 1 - p of e.
int mm_modbatch_set_exponent(mm_modbatch* b, size_t i, int sign, const uint32_t* limbs, size_t nl) {
    if (check_slot(b, i)) return 1;
    if (sign < 0) return fail("exponent must be non-negative");
    Mpz e;
    from_signed_limbs(e.v, sign, limbs, nl);
    size_t nbits = significant_bits(e.v);
    std::vector<uint32_t> words((nbits + 31) / 32, 0);
    for (size_t p = 0; p < nbits; p++)
        if (mpz_tstbit(e.v, nbits - 1 - p)) words[p / 32] |= 1u << (31 - p % 32);
    if (b->be.set_exponent(b->state, i, words.data(), words.size(), nbits)) return fail(b->be.last_error());
    return 0;
}

int64_t mm_modbatch_run(mm_modbatch* b, uint64_t bits) {  // batch.rs:152-154
    int64_t r = b->be.run(b->state, bits);
    b->failed.clear();
    if (r >= 0) b->failed.push_back((uint32_t)r);
    if (r != MM_OK) g_err = b->be.last_error();
    return r;
}

// Every slot whose carry check failed in the latest run or mul: the backend's own list when it
// keeps one, else the one slot that run returned.
size_t mm_modbatch_failed_slots(mm_modbatch* b, uint32_t* out, size_t max) {
    if (b->be.failed_slots) return b->be.failed_slots(b->state, out, max);
    for (size_t i = 0; i < b->failed.size() && i < max; i++) out[i] = b->failed[i];
    return b->failed.size();
}

uint64_t mm_modbatch_remaining(mm_modbatch* b, size_t i) {  // batch.rs:157-159
    if (check_slot(b, i)) return 0;
    return b->be.remaining(b->state, i);
}

int64_t mm_modbatch_mul(mm_modbatch* b, size_t out, size_t a, size_t c) {  // batch.rs:162-166
    if (check_slot(b, out) || check_slot(b, a) || check_slot(b, c)) return MM_EFAULT;
    if (mpz_cmp(b->moduli[out].n.v, b->moduli[a].n.v) || mpz_cmp(b->moduli[a].n.v, b->moduli[c].n.v)) {
        fail("mul: slots must share a modulus");
        return MM_EFAULT;
    }
    int64_t r = b->be.mul(b->state, out, a, c);
    b->failed.clear();
    if (r >= 0) b->failed.push_back((uint32_t)r);
    if (r != MM_OK) g_err = b->be.last_error();
    return r;
}

int mm_modbatch_read_raw(mm_modbatch* b, size_t i, uint32_t* out) {  // batch.rs:168-170
    if (check_slot(b, i)) return 1;
    if (b->be.read(b->state, i, out)) return fail(b->be.last_error());
    return 0;
}

int mm_modbatch_read(mm_modbatch* b, size_t i, uint32_t* out) {  // batch.rs:173-176
    if (mm_modbatch_read_raw(b, i, out)) return 1;
    const MmModulus& m = b->moduli[i];
    Mpz x;
    mpz_import(x.v, b->k, -1, 4, 0, 0, out);
    mpz_mul(x.v, x.v, m.rinv.v);
    mpz_mod(x.v, x.v, m.n.v);
    to_limbs(x.v, out, b->k);
    return 0;
}

int mm_modbatch_is_one(mm_modbatch* b, size_t i) {  // batch.rs:179-183
    std::vector<uint32_t> l(b->k);
    if (mm_modbatch_read_raw(b, i, l.data())) return -1;
    const MmModulus& m = b->moduli[i];
    Mpz v, r2;
    mpz_import(v.v, b->k, -1, 4, 0, 0, l.data());
    mpz_add(r2.v, m.r.v, m.n.v);
    return mpz_cmp(v.v, m.r.v) == 0 || mpz_cmp(v.v, r2.v) == 0;
}

int mm_modbatch_is_one_many(mm_modbatch* b, const uint32_t* slots, size_t count, int32_t* out) {
    for (size_t j = 0; j < count; j++)
        if (check_slot(b, slots[j])) return 1;
    if (b->be.is_one_many) return b->be.is_one_many(b->state, slots, count, out) ? fail(b->be.last_error()) : 0;
    for (size_t j = 0; j < count; j++) {
        const int v = mm_modbatch_is_one(b, slots[j]);
        if (v < 0) return 1;
        out[j] = v;
    }
    return 0;
}

int mm_modbatch_modulus(mm_modbatch* b, size_t i, uint32_t* out) {
    if (check_slot(b, i)) return 1;
    to_limbs(b->moduli[i].n.v, out, b->k);
    return 0;
}

// ---- Mixed lengths ----------------------------------------------------------------------------
// A batch whose moduli need different transform lengths: one ModBatch per length (Params of
// each modulus alone, batch.rs:20-25), stepped together by the backend's run_group. Slot i is
// the i-th modulus given. No reference counterpart: ModBatch::new rejects mixed lengths
// (batch.rs:64-77).
struct mm_mixbatch {
    std::vector<mm_modbatch*> groups;                  // ascending L
    std::vector<std::vector<size_t>> members;          // group -> slots
    std::vector<std::pair<uint32_t, uint32_t>> where;  // slot -> (group, index in the group)
    ~mm_mixbatch() {
        for (mm_modbatch* g : groups) delete g;
    }
};

int mm_mixbatch_new(const mm_backend* be, size_t count, const uint32_t* limbs, const size_t* offsets, const int8_t* signs,
                    uint32_t base, mm_mixbatch** out) {
    *out = nullptr;
    if (count == 0) return fail("empty batch");
    std::map<uint32_t, std::vector<size_t>> by_len;
    for (size_t i = 0; i < count; i++) {
        Mpz n;
        from_signed_limbs(n.v, signs[i], limbs + offsets[i], offsets[i + 1] - offsets[i]);
        const size_t bits = significant_bits(n.v);
        if (bits > 0xffffffffu) return fail("modulus too large");
        uint32_t k, len;
        if (mm_params_for_bits((uint32_t)bits, be->min_len, be->max_len, &k, &len)) return 1;
        by_len[len].push_back(i);
    }
    std::unique_ptr<mm_mixbatch> M(new mm_mixbatch());
    M->where.resize(count);
    for (auto& kv : by_len) {
        const std::vector<size_t>& idx = kv.second;
        std::vector<uint32_t> l;
        std::vector<size_t> off{0};
        std::vector<int8_t> sg;
        for (size_t i : idx) {
            l.insert(l.end(), limbs + offsets[i], limbs + offsets[i + 1]);
            off.push_back(l.size());
            sg.push_back(signs[i]);
        }
        mm_modbatch* b;
        if (mm_modbatch_new(be, idx.size(), l.data(), off.data(), sg.data(), base, &b)) return 1;
        const uint32_t g = (uint32_t)M->groups.size();
        M->groups.push_back(b);
        M->members.push_back(idx);
        for (size_t j = 0; j < idx.size(); j++) M->where[idx[j]] = {g, (uint32_t)j};
    }
    *out = M.release();
    return 0;
}

void mm_mixbatch_free(mm_mixbatch* b) { delete b; }
size_t mm_mixbatch_slots(const mm_mixbatch* b) { return b->where.size(); }
size_t mm_mixbatch_groups(const mm_mixbatch* b) { return b->groups.size(); }

static mm_modbatch* mix_slot(const mm_mixbatch* b, size_t i, size_t* j) {
    if (i >= b->where.size()) {
        fail("slot " + std::to_string(i) + " out of range");
        return nullptr;
    }
    *j = b->where[i].second;
    return b->groups[b->where[i].first];
}

int mm_mixbatch_params(const mm_mixbatch* b, size_t i, uint32_t* k, uint32_t* len) {
    size_t j;
    mm_modbatch* g = mix_slot(b, i, &j);
    if (!g) return 1;
    mm_modbatch_params(g, k, len);
    return 0;
}

int mm_mixbatch_load(mm_mixbatch* b, size_t i, int sign, const uint32_t* limbs, size_t nl) {
    size_t j;
    mm_modbatch* g = mix_slot(b, i, &j);
    return g ? mm_modbatch_load(g, j, sign, limbs, nl) : 1;
}

int mm_mixbatch_load_raw(mm_mixbatch* b, size_t i, int sign, const uint32_t* limbs, size_t nl) {
    size_t j;
    mm_modbatch* g = mix_slot(b, i, &j);
    return g ? mm_modbatch_load_raw(g, j, sign, limbs, nl) : 1;
}

int mm_mixbatch_set_exponent(mm_mixbatch* b, size_t i, int sign, const uint32_t* limbs, size_t nl) {
    size_t j;
    mm_modbatch* g = mix_slot(b, i, &j);
    return g ? mm_modbatch_set_exponent(g, j, sign, limbs, nl) : 1;
}

// Every group through one run_group call when the backend has it, else group after group.
// Returns MM_OK, the first failing slot (its index in this batch), or MM_EFAULT.
int64_t mm_mixbatch_run(mm_mixbatch* b, uint64_t bits) {
    const mm_backend& be = b->groups[0]->be;
    for (mm_modbatch* g : b->groups) g->failed.clear();
    if (be.run_group && b->groups.size() > 1) {
        std::vector<void*> st;
        for (mm_modbatch* g : b->groups) st.push_back(g->state);
        size_t which = 0;
        const int64_t r = be.run_group(st.data(), st.size(), bits, &which);
        if (r == MM_OK) return r;
        if (r < 0 || which >= b->groups.size() || (size_t)r >= b->members[which].size()) {
            g_err = be.last_error();
            return r < 0 ? r : MM_EFAULT;
        }
        const size_t slot = b->members[which][(size_t)r];
        g_err = std::string(be.last_error()) + " (slot " + std::to_string(slot) + " of the mixed batch)";
        return (int64_t)slot;
    }
    // Group after group; every group takes its steps, and the first failing slot is returned.
    int64_t first = MM_OK;
    std::string why;
    for (size_t g = 0; g < b->groups.size(); g++) {
        const int64_t r = mm_modbatch_run(b->groups[g], bits);
        if (r == MM_OK) continue;
        if (r < 0 || (size_t)r >= b->members[g].size()) return r < 0 ? r : MM_EFAULT;
        if (first == MM_OK) {
            first = (int64_t)b->members[g][(size_t)r];
            why = g_err + " (slot " + std::to_string(first) + " of the mixed batch)";
        }
    }
    if (first != MM_OK) g_err = why;
    return first;
}

// Every slot of the mixed batch whose carry check failed in the latest run, ascending.
size_t mm_mixbatch_failed_slots(mm_mixbatch* b, uint32_t* out, size_t max) {
    std::vector<uint32_t> all, part;
    for (size_t g = 0; g < b->groups.size(); g++) {
        const size_t cnt = mm_modbatch_failed_slots(b->groups[g], nullptr, 0);
        part.assign(cnt, 0);
        mm_modbatch_failed_slots(b->groups[g], part.data(), cnt);
        for (uint32_t j : part)
            if (j < b->members[g].size()) all.push_back((uint32_t)b->members[g][j]);
    }
    std::sort(all.begin(), all.end());
    for (size_t i = 0; i < all.size() && i < max; i++) out[i] = all[i];
    return all.size();
}

uint64_t mm_mixbatch_remaining(mm_mixbatch* b, size_t i) {
    size_t j;
    mm_modbatch* g = mix_slot(b, i, &j);
    return g ? mm_modbatch_remaining(g, j) : 0;
}

int mm_mixbatch_read_raw(mm_mixbatch* b, size_t i, uint32_t* out) {
    size_t j;
    mm_modbatch* g = mix_slot(b, i, &j);
    return g ? mm_modbatch_read_raw(g, j, out) : 1;
}

int mm_mixbatch_read(mm_mixbatch* b, size_t i, uint32_t* out) {
    size_t j;
    mm_modbatch* g = mix_slot(b, i, &j);
    return g ? mm_modbatch_read(g, j, out) : 1;
}

int mm_mixbatch_is_one(mm_mixbatch* b, size_t i) {
    size_t j;
    mm_modbatch* g = mix_slot(b, i, &j);
    return g ? mm_modbatch_is_one(g, j) : -1;
}

int mm_mixbatch_is_one_many(mm_mixbatch* b, const uint32_t* slots, size_t count, int32_t* out) {
    // Per length: the listed slots of each group go to that group's is_one_many together.
    std::vector<std::vector<uint32_t>> idx(b->groups.size()), pos(b->groups.size());
    for (size_t j = 0; j < count; j++) {
        if (slots[j] >= b->where.size()) return fail("slot " + std::to_string(slots[j]) + " out of range");
        const auto& w = b->where[slots[j]];
        idx[w.first].push_back(w.second);
        pos[w.first].push_back((uint32_t)j);
    }
    for (size_t g = 0; g < b->groups.size(); g++) {
        if (idx[g].empty()) continue;
        std::vector<int32_t> r(idx[g].size());
        if (mm_modbatch_is_one_many(b->groups[g], idx[g].data(), idx[g].size(), r.data())) return 1;
        for (size_t j = 0; j < r.size(); j++) out[pos[g][j]] = r[j];
    }
    return 0;
}

int mm_mixbatch_check_pow(mm_mixbatch* b, size_t i, uint32_t base, const uint32_t* e, size_t ne) {
    size_t j;
    mm_modbatch* g = mix_slot(b, i, &j);
    return g ? mm_modbatch_check_pow(g, j, base, e, ne) : -1;
}

// The backend state of group `group` (groups in ascending L; NULL if out of range): for plan
// queries and kernel timing.
void* mm_mixbatch_group_state(mm_mixbatch* b, size_t group) {
    return group < b->groups.size() ? b->groups[group]->state : nullptr;
}

// fermat (batch.rs:187-198).
int mm_fermat(const mm_backend* be, size_t count, const uint32_t* limbs, const size_t* offsets,
              const int8_t* signs, uint32_t base, uint64_t chunk, uint8_t* prp) {
    if (chunk == 0) return fail("chunk must be positive");
    mm_modbatch* b;
    if (mm_modbatch_new(be, count, limbs, offsets, signs, base, &b)) return 1;
    uint32_t one = 1;
    for (size_t i = 0; i < count; i++) {
        Mpz e;
        mpz_set(e.v, b->moduli[i].n.v);
        mpz_sub_ui(e.v, e.v, 1);
        std::vector<uint32_t> el((significant_bits(e.v) + 31) / 32 + 1, 0);
        size_t cnt = 0;
        if (mpz_sgn(e.v)) mpz_export(el.data(), &cnt, -1, 4, 0, 0, e.v);
        if (mm_modbatch_load(b, i, 1, &one, 1) || mm_modbatch_set_exponent(b, i, 1, el.data(), el.size())) {
            mm_modbatch_free(b);
            return 1;
        }
    }
    for (;;) {
        bool any = false;
        for (size_t i = 0; i < count && !any; i++) any = b->be.remaining(b->state, i) > 0;
        if (!any) break;
        int64_t r = mm_modbatch_run(b, chunk);
        if (r != MM_OK) { mm_modbatch_free(b); return 1; }
    }
    std::vector<uint32_t> all(count);
    std::vector<int32_t> ones(count);
    for (size_t i = 0; i < count; i++) all[i] = (uint32_t)i;
    if (mm_modbatch_is_one_many(b, all.data(), count, ones.data())) { mm_modbatch_free(b); return 1; }
    for (size_t i = 0; i < count; i++) prp[i] = (uint8_t)ones[i];
    mm_modbatch_free(b);
    return 0;
}

// The score's check of one slot (score.rs:107-115): no exponent bits left and
// read(slot) == base^e mod N, with e given as little-endian limbs. 1 if it holds, 0 if not,
// negative on error.
int mm_modbatch_check_pow(mm_modbatch* b, size_t i, uint32_t base, const uint32_t* e, size_t ne) {
    if (check_slot(b, i)) return -1;
    if (b->be.remaining(b->state, i) != 0) return 0;
    std::vector<uint32_t> l(b->k);
    if (mm_modbatch_read(b, i, l.data())) return -1;
    Mpz got, want, ez, bz;
    mpz_import(got.v, b->k, -1, 4, 0, 0, l.data());
    mpz_import(ez.v, ne, -1, 4, 0, 0, e);
    mpz_set_ui(bz.v, base);
    mpz_powm(want.v, bz.v, ez.v, b->moduli[i].n.v);
    return mpz_cmp(got.v, want.v) == 0;
}

// check_pow on every slot: the raw limbs of all slots are read on the calling thread (the
// backend is driven from one thread), then the reductions and GMP's powm run on `threads`
// threads. Returns -1 when every slot has no bits left and reads base^e mod N, the first slot
// that does not, or -2 on an error. exps: slot i's exponent is the little-endian limbs
// exps[offsets[i] .. offsets[i+1]).
int64_t mm_modbatch_check_pow_range(mm_modbatch* b, uint32_t base, size_t first_slot, size_t count,
                                    const uint32_t* exps, const size_t* offsets, uint32_t threads) {
    const size_t k = b->k;
    if (first_slot > b->moduli.size() || count > b->moduli.size() - first_slot) {
        fail("check_pow_range: slot range past the batch");
        return -2;
    }
    std::vector<uint32_t> raw(count * k);
    for (size_t i = 0; i < count; i++) {
        if (b->be.remaining(b->state, first_slot + i) != 0) return (int64_t)(first_slot + i);
        if (b->be.read(b->state, first_slot + i, raw.data() + i * k)) {
            fail(b->be.last_error());
            return -2;
        }
    }
    const size_t slots = count;
    std::atomic<size_t> next{0};
    std::atomic<int64_t> first{-1};
    auto work = [&] {
        Mpz got, want, ez, bz;
        mpz_set_ui(bz.v, base);
        for (size_t i; (i = next.fetch_add(1)) < slots;) {
            const MmModulus& m = b->moduli[first_slot + i];
            mpz_import(got.v, k, -1, 4, 0, 0, raw.data() + i * k);
            mpz_mul(got.v, got.v, m.rinv.v);
            mpz_mod(got.v, got.v, m.n.v);
            mpz_import(ez.v, offsets[i + 1] - offsets[i], -1, 4, 0, 0, exps + offsets[i]);
            mpz_powm(want.v, bz.v, ez.v, m.n.v);
            if (mpz_cmp(got.v, want.v) != 0) {
                int64_t cur = first.load();
                while ((cur < 0 || (int64_t)i < cur) && !first.compare_exchange_weak(cur, (int64_t)i)) {
                }
            }
        }
    };
    std::vector<std::thread> pool;
    for (uint32_t t = 1; t < std::max<uint32_t>(1, threads); t++) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
    return first.load() < 0 ? -1 : (int64_t)first_slot + first.load();
}

// The score workload (score.rs:45-61): `slots` moduli of `bits` bits from GMP's default
// random state seeded with `seed` (rug's RandState::new + seed + Integer::random_bits), top bit
// and bit 0 set, and per modulus the top `ebits` bits of N - 1, i.e. (N - 1) >> (bits - ebits).
// Writes ceil(bits/32) limbs per modulus and ceil(ebits/32) limbs per exponent, little-endian.
int mm_score_workload(uint32_t bits, size_t slots, uint32_t ebits, uint64_t seed, uint32_t* moduli, uint32_t* exps) {
    if (slots == 0 || bits < 2 || ebits == 0 || ebits > bits) return fail("slots > 0 and 0 < ebits <= bits required");
    gmp_randstate_t rs;
    gmp_randinit_default(rs);
    Mpz s, n, e;
    mpz_set_ui(s.v, (unsigned long)seed);
    gmp_randseed(rs, s.v);
    size_t nl = (bits + 31) / 32, el = (ebits + 31) / 32;
    for (size_t i = 0; i < slots; i++) {
        mpz_urandomb(n.v, rs, bits);
        mpz_setbit(n.v, bits - 1);
        mpz_setbit(n.v, 0);
        to_limbs(n.v, moduli + i * nl, nl);
        mpz_sub_ui(e.v, n.v, 1);
        mpz_fdiv_q_2exp(e.v, e.v, bits - ebits);
        to_limbs(e.v, exps + i * el, el);
    }
    gmp_randclear(rs);
    return 0;
}

}  // extern "C"
