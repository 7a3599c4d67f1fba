// mm_multi.cpp - several backends as one (reference: Multi, montmul/src/multi.rs:11-156).
//
// A batch's slots are split into contiguous shards, one per part, whose sizes differ by at most
// one (multi.rs:39-42). Creating and running the shards happens on one host thread per shard;
// every per-slot call goes to the slot's shard. Slots are independent, so only a `mul` whose
// operands live on different shards needs the host: it reads both, multiplies with GMP and loads
// the result (multi.rs:128-138). A failure on a shard is reported as the reference relabels its
// panics (multi.rs:58-70): a message ending in "slot i" becomes "... shard s, slot start + i",
// any other "shard s (first slot start): message".
//
// With one HIP context per part (mm_hip_open with the device of each part), the shards run on
// separate streams, of one device or of several.

#include <gmp.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/montmul_hip.h"

namespace {

thread_local std::string g_err;
int fail(const std::string& m) {
    g_err = m;
    return 1;
}

struct Multi {
    std::vector<mm_backend> parts;
};

struct MultiState {
    std::vector<mm_backend> parts;   // copies of the parts' tables: a state outlives mm_multi_close
    std::vector<void*> shards;
    std::vector<size_t> starts;      // first slot of each shard
    size_t slots = 0;
    uint32_t k = 0;
    std::vector<uint32_t> n;         // N per slot (k limbs each), for mul across shards
    std::vector<uint32_t> failed;    // every failing slot of the latest run or mul (global numbers)
    ~MultiState() {
        for (size_t s = 0; s < shards.size(); s++)
            if (shards[s]) parts[s].destroy(shards[s]);
    }
    // (shard, index within the shard) of slot i (multi.rs:80-85).
    size_t locate(size_t i, size_t* j) const {
        const size_t s = std::upper_bound(starts.begin(), starts.end(), i) - starts.begin() - 1;
        *j = i - starts[s];
        return s;
    }
};

// Contiguous shard ranges over `slots`, sizes differing by at most one, none empty (multi.rs:39-42).
std::vector<std::pair<size_t, size_t>> split(size_t slots, size_t parts) {
    const size_t shards = std::min(parts, slots);
    std::vector<std::pair<size_t, size_t>> out;
    for (size_t s = 0; s < shards; s++) out.push_back({s * slots / shards, (s + 1) * slots / shards});
    return out;
}

// Runs f(s) for every shard, each on its own thread (directly for one shard) (multi.rs:44-55).
template <class F> void each(size_t count, F&& f) {
    if (count == 1) {
        f(0);
        return;
    }
    std::vector<std::thread> th;
    for (size_t s = 0; s < count; s++) th.emplace_back([&f, s] { f(s); });
    for (auto& t : th) t.join();
}

// The reference's relabelling of a shard's failure (multi.rs:58-70): "... slot i" becomes
// "... shard s, slot start + i" (whatever follows the number, such as a count of further
// failing slots, is kept), any other message "shard s (first slot start): message".
std::string relabel(size_t s, size_t start, const std::string& msg) {
    for (size_t at = msg.rfind("slot "); at != std::string::npos; at = at ? msg.rfind("slot ", at - 1) : std::string::npos) {
        size_t e = at + 5;
        unsigned long long i = 0;
        while (e < msg.size() && isdigit((unsigned char)msg[e]) && e < at + 5 + 18) i = 10 * i + (unsigned)(msg[e++] - '0');
        if (e > at + 5)
            return msg.substr(0, at) + "shard " + std::to_string(s) + ", slot " + std::to_string(start + i) + msg.substr(e);
    }
    return "shard " + std::to_string(s) + " (first slot " + std::to_string(start) + "): " + msg;
}

int mt_create(void* self, const mm_batch_setup* st, void** out) {
    Multi* M = (Multi*)self;
    *out = nullptr;
    if (st->slots == 0) return fail("empty batch");
    std::unique_ptr<MultiState> S(new MultiState());
    S->parts = M->parts;
    S->slots = st->slots;
    S->k = st->k;
    S->n.assign(st->n, st->n + (size_t)st->slots * st->k);
    const auto ranges = split(st->slots, M->parts.size());
    S->shards.assign(ranges.size(), nullptr);
    for (auto& r : ranges) S->starts.push_back(r.first);
    std::vector<std::string> errs(ranges.size());
    std::vector<int> rc(ranges.size(), 0);
    each(ranges.size(), [&](size_t s) {
        const size_t a = ranges[s].first, cnt = ranges[s].second - a, k = st->k;
        mm_batch_setup sub{st->k, st->len, st->base, (uint32_t)cnt, st->n + a * k, st->nprime + a * k, st->two_n + a * k};
        rc[s] = M->parts[s].create(M->parts[s].self, &sub, &S->shards[s]);
        if (rc[s]) errs[s] = M->parts[s].last_error();
    });
    for (size_t s = 0; s < ranges.size(); s++)
        if (rc[s]) return fail(relabel(s, ranges[s].first, errs[s]));
    *out = S.release();
    return 0;
}

void mt_destroy(void* st) { delete (MultiState*)st; }

int mt_load(void* st, size_t i, const uint32_t* limbs) {
    MultiState* S = (MultiState*)st;
    if (i >= S->slots) return fail("slot " + std::to_string(i) + " out of range");
    size_t j;
    const size_t s = S->locate(i, &j);
    if (S->parts[s].load(S->shards[s], j, limbs)) return fail(relabel(s, S->starts[s], S->parts[s].last_error()));
    return 0;
}

int mt_set_exponent(void* st, size_t i, const uint32_t* words, size_t nwords, uint64_t nbits) {
    MultiState* S = (MultiState*)st;
    if (i >= S->slots) return fail("slot " + std::to_string(i) + " out of range");
    size_t j;
    const size_t s = S->locate(i, &j);
    if (S->parts[s].set_exponent(S->shards[s], j, words, nwords, nbits))
        return fail(relabel(s, S->starts[s], S->parts[s].last_error()));
    return 0;
}

// Every shard's failing slots in global numbers, ascending, after a run or mul that returned r.
void collect_failed(MultiState* S, size_t s, int64_t r, std::vector<uint32_t>& out) {
    const mm_backend& P = S->parts[s];
    if (P.failed_slots) {
        std::vector<uint32_t> v(P.failed_slots(S->shards[s], nullptr, 0));
        P.failed_slots(S->shards[s], v.data(), v.size());
        for (uint32_t j : v) out.push_back((uint32_t)(S->starts[s] + j));
    } else if (r >= 0) {
        out.push_back((uint32_t)(S->starts[s] + r));
    }
}

int64_t mt_run(void* st, uint64_t bits) {
    MultiState* S = (MultiState*)st;
    const size_t cnt = S->shards.size();
    std::vector<int64_t> rc(cnt, MM_OK);
    std::vector<std::string> errs(cnt);
    each(cnt, [&](size_t s) {
        rc[s] = S->parts[s].run(S->shards[s], bits);
        if (rc[s] != MM_OK) errs[s] = S->parts[s].last_error();
    });
    S->failed.clear();
    int64_t first = MM_OK;
    std::string why;
    for (size_t s = 0; s < cnt; s++) {
        if (rc[s] == MM_OK) continue;
        if (rc[s] < 0) {
            fail(relabel(s, S->starts[s], errs[s]));
            return MM_EFAULT;
        }
        collect_failed(S, s, rc[s], S->failed);
        if (first == MM_OK) {
            first = (int64_t)(S->starts[s] + rc[s]);
            why = relabel(s, S->starts[s], errs[s]);
        }
    }
    std::sort(S->failed.begin(), S->failed.end());
    if (first != MM_OK) g_err = why;
    return first;
}

int mt_read(void* st, size_t i, uint32_t* out) {
    MultiState* S = (MultiState*)st;
    if (i >= S->slots) return fail("slot " + std::to_string(i) + " out of range");
    size_t j;
    const size_t s = S->locate(i, &j);
    if (S->parts[s].read(S->shards[s], j, out)) return fail(relabel(s, S->starts[s], S->parts[s].last_error()));
    return 0;
}

// Within one shard its part computes the product; across shards the host does, with the
// result in [0, N) (multi.rs:128-138).
int64_t mt_mul(void* st, size_t out, size_t a, size_t b) {
    MultiState* S = (MultiState*)st;
    if (out >= S->slots || a >= S->slots || b >= S->slots) {
        fail("slot out of range");
        return MM_EFAULT;
    }
    size_t jo, ja, jb;
    const size_t so = S->locate(out, &jo), sa = S->locate(a, &ja), sb = S->locate(b, &jb);
    S->failed.clear();
    if (so == sa && sa == sb) {
        const int64_t r = S->parts[so].mul(S->shards[so], jo, ja, jb);
        if (r == MM_OK) return r;
        const std::string msg = relabel(so, S->starts[so], S->parts[so].last_error());
        if (r < 0) {
            fail(msg);
            return MM_EFAULT;
        }
        collect_failed(S, so, r, S->failed);
        g_err = msg;
        return (int64_t)(S->starts[so] + r);
    }
    const size_t k = S->k;
    std::vector<uint32_t> la(k), lb(k);
    if (mt_read(st, a, la.data()) || mt_read(st, b, lb.data())) return MM_EFAULT;
    mpz_t n, x, y, r;
    mpz_inits(n, x, y, r, nullptr);
    mpz_import(n, k, -1, 4, 0, 0, S->n.data() + out * k);
    mpz_import(x, k, -1, 4, 0, 0, la.data());
    mpz_import(y, k, -1, 4, 0, 0, lb.data());
    mpz_setbit(r, 32 * k);  // R = 2^(32k)
    int64_t res = MM_OK;
    if (!mpz_invert(r, r, n)) {
        fail("mul: R is not invertible modulo N");
        res = MM_EFAULT;
    } else {
        mpz_mul(x, x, y);
        mpz_mul(x, x, r);
        mpz_mod(x, x, n);
        std::vector<uint32_t> lo(k, 0);
        size_t cnt = 0;
        mpz_export(lo.data(), &cnt, -1, 4, 0, 0, x);
        if (mt_load(st, out, lo.data())) res = MM_EFAULT;
    }
    mpz_clears(n, x, y, r, nullptr);
    return res;
}

uint64_t mt_remaining(void* st, size_t i) {
    MultiState* S = (MultiState*)st;
    if (i >= S->slots) return 0;
    size_t j;
    const size_t s = S->locate(i, &j);
    return S->parts[s].remaining(S->shards[s], j);
}

int mt_set_modulus(void* st, size_t i, const uint32_t* n, const uint32_t* nprime, const uint32_t* two_n) {
    MultiState* S = (MultiState*)st;
    if (i >= S->slots) return fail("slot " + std::to_string(i) + " out of range");
    size_t j;
    const size_t s = S->locate(i, &j);
    const mm_backend& P = S->parts[s];
    if (!P.set_modulus) return fail("shard " + std::to_string(s) + ": the part has no set_modulus");
    if (P.set_modulus(S->shards[s], j, n, nprime, two_n)) return fail(relabel(s, S->starts[s], P.last_error()));
    std::copy(n, n + S->k, S->n.begin() + i * S->k);
    return 0;
}

size_t mt_failed_slots(void* st, uint32_t* out, size_t max) {
    const MultiState* S = (const MultiState*)st;
    for (size_t i = 0; i < S->failed.size() && i < max; i++) out[i] = S->failed[i];
    return S->failed.size();
}

const char* mt_last_error(void) { return g_err.c_str(); }

}  // namespace

extern "C" {

int mm_multi_open(const mm_backend* parts, size_t count, mm_backend* out) {
    memset(out, 0, sizeof(*out));
    if (count == 0) return fail("Multi: no parts");
    for (size_t i = 1; i < count; i++) {
        const mm_backend& p = parts[i];
        const mm_backend& q = parts[0];
        if (p.bound_lo != q.bound_lo || p.bound_hi != q.bound_hi || p.min_len != q.min_len || p.max_len != q.max_len ||
            p.full_k != q.full_k)
            return fail("Multi: part " + std::to_string(i) + " has other limits than part 0");
    }
    Multi* M = new Multi();
    M->parts.assign(parts, parts + count);
    out->self = M;
    out->create = mt_create;
    out->destroy = mt_destroy;
    out->load = mt_load;
    out->set_exponent = mt_set_exponent;
    out->run = mt_run;
    out->mul = mt_mul;
    out->read = mt_read;
    out->remaining = mt_remaining;
    bool setmod = true;
    for (size_t i = 0; i < count; i++) setmod = setmod && parts[i].set_modulus;
    out->set_modulus = setmod ? mt_set_modulus : nullptr;
    out->run_group = nullptr;
    out->failed_slots = mt_failed_slots;
    out->last_error = mt_last_error;
    out->bound_lo = parts[0].bound_lo;
    out->bound_hi = parts[0].bound_hi;
    out->min_len = parts[0].min_len;
    out->max_len = parts[0].max_len;
    out->full_k = parts[0].full_k;
    return 0;
}

void mm_multi_close(mm_backend* be) {
    delete (Multi*)be->self;
    be->self = nullptr;
}

size_t mm_multi_parts(const mm_backend* be) { return ((const Multi*)be->self)->parts.size(); }

const char* mm_multi_last_error(void) { return g_err.c_str(); }

}  // extern "C"
