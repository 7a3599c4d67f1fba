// mm_serve.cpp - a request server for base-b Fermat PRP tests over length pools.
//
// No reference counterpart: the reference runs fixed batches (ModBatch, fermat()). Here requests
// (one modulus each, exponent N - 1) arrive at any time. Each transform length L gets a pool: a
// ModBatch of `pool_slots` slots on a backend table of its own (for the HIP backend, its own
// stream), driven by one host thread. A pool's round is:
//   1. admit waiting requests into free slots (mm_modbatch_set_modulus, load 1, exponent N - 1);
//   2. run every slot with bits left for a chunk of steps sized to `target_ms` from the pool's
//      measured time per step (slots with no bits left are skipped by the kernel);
//   3. retire finished slots: is_one on their raw values, all at once (on the device when the
//      backend tests it there: one kernel and a copy of one word per slot) and a completion
//      record each.
// Pools run concurrently, so requests of different lengths share the GPU through streams.
//
// Admission by device memory: with hbm_bytes set, a pool reserves slots x mm_server_slot_bytes(L)
// from the budget when it creates its batch, cut to what is left. A length that finds no room
// waits; a pool with no active slot and no waiting request frees its batch while another length
// waits, and waiting pools retry on every release.

#include <gmp.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/montmul_hip.h"

namespace {

thread_local std::string g_err;
int fail(const std::string& m) { g_err = m; return 1; }

using Clock = std::chrono::steady_clock;

struct Request {
    int64_t id;
    std::vector<uint32_t> n;       // little-endian limbs of N
    std::vector<uint32_t> e;       // N - 1 (a Fermat test), or the exponent given at submit
    uint64_t bits;                 // bits of e: the steps it takes
    double t_submit;
};

struct Pool;

}  // namespace

struct mm_server {
    mm_server_options opt{};
    mm_open_fn open = nullptr;
    mm_close_fn close = nullptr;
    void* user = nullptr;
    mm_backend probe{};  // limits for routing
    bool probe_open = false;
    Clock::time_point t0;

    std::mutex mu;                       // guards everything below
    std::condition_variable done_cv;     // completions
    bool stop = false;
    int64_t next_id = 0;
    std::map<uint32_t, std::unique_ptr<Pool>> pools;
    std::deque<mm_server_result> done;
    std::unordered_map<int64_t, std::vector<uint32_t>> raws;
    mm_server_stats stats{};
    std::string last_error_text;
    void* owned_user = nullptr;  // freed with the server (the HIP factory's options)
    uint64_t hbm_used = 0;       // bytes reserved by live pool batches
    uint64_t hbm_epoch = 0;      // counts releases, so that waiting pools retry
    uint32_t deferred = 0;       // pools waiting for device memory

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

struct Pool {
    mm_server* S;
    uint32_t len;
    std::deque<Request> waiting;         // guarded by S->mu
    std::condition_variable cv;          // new requests or stop
    std::thread th;

    // owned by the pool thread
    mm_backend be{};
    bool be_open = false;
    mm_modbatch* batch = nullptr;
    std::vector<int64_t> slot_req;
    std::vector<double> slot_submit, slot_admit;
    std::vector<uint64_t> slot_bits;
    std::vector<uint32_t> free_slots;
    double ms_per_step = 0;
    uint64_t bytes = 0;          // reserved for this pool's batch (guarded by S->mu)
    bool deferred = false;       // waiting for device memory (guarded by S->mu)
    uint64_t seen_epoch = 0;     // S->hbm_epoch at the last failed reservation

    void finish(size_t slot, int prp, const char* err) {
        mm_server_result r{};
        r.id = slot_req[slot];
        r.prp = err ? -1 : prp;
        r.len = len;
        r.bits = slot_bits[slot];
        r.submit_s = slot_submit[slot];
        r.admit_s = slot_admit[slot];
        std::vector<uint32_t> raw;
        if (S->opt.keep_raw && !err) {
            uint32_t k, l;
            mm_modbatch_params(batch, &k, &l);
            raw.resize(k);
            if (mm_modbatch_read_raw(batch, slot, raw.data())) raw.clear();
        }
        std::lock_guard<std::mutex> g(S->mu);
        r.done_s = S->now();
        if (!raw.empty()) S->raws[r.id] = std::move(raw);
        S->done.push_back(r);
        S->stats.completed++;
        if (err) S->stats.failed++;
        S->done_cv.notify_all();
        slot_req[slot] = -1;
        free_slots.push_back((uint32_t)slot);
    }

    // Creates the batch on first use, every slot on the first request's modulus. Returns 0, 1 on
    // error, or 2 when the memory budget has no room for a slot yet.
    int ensure_batch(const Request& r) {
        if (batch) return 0;
        if (!be_open) {
            if (S->open(S->user, &be)) return fail("cannot open a backend for length " + std::to_string(len));
            be_open = true;
        }
        uint32_t cnt = std::max<uint32_t>(1, S->opt.pool_slots);
        const uint64_t per = mm_server_slot_bytes(len);
        {
            std::lock_guard<std::mutex> g(S->mu);
            if (S->opt.hbm_bytes) {
                if (per > S->opt.hbm_bytes)
                    return fail("a slot of length " + std::to_string(len) + " needs " + std::to_string(per) +
                                " bytes, more than the budget of " + std::to_string(S->opt.hbm_bytes));
                const uint64_t room = S->opt.hbm_bytes > S->hbm_used ? S->opt.hbm_bytes - S->hbm_used : 0;
                cnt = (uint32_t)std::min<uint64_t>(cnt, room / per);
                if (cnt == 0) {
                    if (!deferred) {
                        deferred = true;
                        S->deferred++;
                        // idle pools check S->deferred when woken
                        for (auto& p : S->pools)
                            if (p.second.get() != this) p.second->cv.notify_all();
                    }
                    seen_epoch = S->hbm_epoch;
                    return 2;
                }
            }
            bytes = (uint64_t)cnt * per;
            S->hbm_used += bytes;
            if (deferred) {
                deferred = false;
                S->deferred--;
            }
        }
        std::vector<uint32_t> limbs;
        std::vector<size_t> offsets(1, 0);
        std::vector<int8_t> signs(cnt, 1);
        for (uint32_t i = 0; i < cnt; i++) {
            limbs.insert(limbs.end(), r.n.begin(), r.n.end());
            offsets.push_back(limbs.size());
        }
        if (mm_modbatch_new(&be, cnt, limbs.data(), offsets.data(), signs.data(), S->opt.base, &batch)) {
            batch = nullptr;
            const std::string why = mm_host_last_error();
            give_back();
            return fail(why);
        }
        slot_req.assign(cnt, -1);
        slot_submit.assign(cnt, 0);
        slot_admit.assign(cnt, 0);
        slot_bits.assign(cnt, 0);
        free_slots.clear();
        for (uint32_t i = cnt; i-- > 0;) free_slots.push_back(i);
        return 0;
    }

    // Returns this pool's reservation to the budget and wakes every pool (guarded by S->mu
    // inside).
    void give_back() {
        std::lock_guard<std::mutex> g(S->mu);
        S->hbm_used -= bytes;
        bytes = 0;
        S->hbm_epoch++;
        for (auto& p : S->pools) p.second->cv.notify_all();
    }

    // Frees an idle batch so that a waiting length can use its memory.
    void release() {
        mm_modbatch_free(batch);
        batch = nullptr;
        slot_req.clear();
        free_slots.clear();
        {
            std::lock_guard<std::mutex> g(S->mu);
            S->stats.releases++;
        }
        give_back();
    }

    // 0: admitted; 1: failed; 2: no memory for this length's batch yet; 3: no free slot (the
    // batch was cut to fewer slots than the requests taken).
    int admit(Request& r) {
        const int e = ensure_batch(r);
        if (e) return e;
        if (free_slots.empty()) return 3;
        const uint32_t slot = free_slots.back();
        const uint32_t one = 1;
        if (mm_modbatch_set_modulus(batch, slot, 1, r.n.data(), r.n.size()) ||
            mm_modbatch_load(batch, slot, 1, &one, 1) ||
            mm_modbatch_set_exponent(batch, slot, 1, r.e.data(), r.e.size()))
            return 1;
        free_slots.pop_back();
        slot_req[slot] = r.id;
        slot_submit[slot] = r.t_submit;
        slot_bits[slot] = r.bits;
        std::lock_guard<std::mutex> g(S->mu);
        slot_admit[slot] = S->now();
        S->stats.admitted++;
        return 0;
    }

    void reject(const Request& r, const std::string& why) {
        std::lock_guard<std::mutex> g(S->mu);
        mm_server_result res{};
        res.id = r.id;
        res.prp = -1;
        res.len = len;
        res.bits = r.bits;
        res.submit_s = r.t_submit;
        res.admit_s = res.done_s = S->now();
        S->done.push_back(res);
        S->stats.completed++;
        S->stats.failed++;
        S->last_error_text = why;
        S->done_cv.notify_all();
    }

    size_t active() const { return slot_req.size() - free_slots.size(); }

    void loop() {
        for (;;) {
            std::vector<Request> take;
            bool idle_release = false;
            {
                std::unique_lock<std::mutex> lk(S->mu);
                cv.wait(lk, [&] {
                    return S->stop || (!waiting.empty() && batch && !free_slots.empty()) ||
                           (!waiting.empty() && !batch && (!deferred || S->hbm_epoch != seen_epoch)) || active() > 0 ||
                           (batch && active() == 0 && waiting.empty() && S->deferred > 0);
                });
                if (S->stop) break;
                if (batch && active() == 0 && waiting.empty() && S->deferred > 0) {
                    idle_release = true;
                } else {
                    size_t room = batch ? free_slots.size() : std::max<uint32_t>(1, S->opt.pool_slots);
                    while (!waiting.empty() && take.size() < room) {
                        take.push_back(std::move(waiting.front()));
                        waiting.pop_front();
                    }
                }
            }
            if (idle_release) {
                release();
                continue;
            }
            for (size_t i = 0; i < take.size(); i++) {
                const int a = admit(take[i]);
                if (a == 2 || a == 3) {
                    // No memory for this length yet, or no slot left: the rest wait, in order.
                    std::lock_guard<std::mutex> g(S->mu);
                    for (size_t j = take.size(); j-- > i;) waiting.push_front(std::move(take[j]));
                    break;
                }
                if (a == 1) reject(take[i], g_err.empty() ? std::string(mm_host_last_error()) : g_err);
            }
            if (!batch || active() == 0) continue;
            // One chunk of steps, sized from the measured time per step.
            uint64_t chunk = S->opt.max_chunk ? S->opt.max_chunk : 4096;
            if (ms_per_step > 0 && S->opt.target_ms > 0)
                chunk = std::min<uint64_t>(chunk, std::max<uint64_t>(1, (uint64_t)(S->opt.target_ms / ms_per_step)));
            else if (ms_per_step == 0)
                chunk = std::min<uint64_t>(chunk, 16);
            const auto t = Clock::now();
            const int64_t rr = mm_modbatch_run(batch, chunk);
            const double ms = std::chrono::duration<double, std::milli>(Clock::now() - t).count();
            if (rr == MM_OK) ms_per_step = ms / (double)chunk;
            {
                std::lock_guard<std::mutex> g(S->mu);
                S->stats.rounds++;
                S->stats.busy_ms += ms;
            }
            if (rr != MM_OK) {
                // A carry error fails every slot it hit (the others run on); any other failure
                // ends every active request.
                const std::string why = mm_host_last_error();
                std::vector<uint8_t> bad(slot_req.size(), rr < 0 ? 1 : 0);
                if (rr >= 0) {
                    std::vector<uint32_t> hit(mm_modbatch_failed_slots(batch, nullptr, 0));
                    mm_modbatch_failed_slots(batch, hit.data(), hit.size());
                    hit.push_back((uint32_t)rr);
                    for (uint32_t i : hit)
                        if (i < bad.size()) bad[i] = 1;
                }
                for (size_t i = 0; i < slot_req.size(); i++)
                    if (slot_req[i] >= 0 && bad[i]) finish(i, 0, why.c_str());
                continue;
            }
            // Finished slots: is_one for all of them at once (on the device, one small copy).
            std::vector<uint32_t> fin;
            for (size_t i = 0; i < slot_req.size(); i++)
                if (slot_req[i] >= 0 && mm_modbatch_remaining(batch, i) == 0) fin.push_back((uint32_t)i);
            if (fin.empty()) continue;
            std::vector<int32_t> one(fin.size(), -1);
            const bool ok = mm_modbatch_is_one_many(batch, fin.data(), fin.size(), one.data()) == 0;
            const std::string why = ok ? std::string() : "is_one failed: " + std::string(mm_host_last_error());
            for (size_t j = 0; j < fin.size(); j++) {
                const size_t i = fin[j];
                {
                    std::lock_guard<std::mutex> g(S->mu);
                    S->stats.squarings += slot_bits[i] ? slot_bits[i] - 1 : 0;
                }
                finish(i, ok ? one[j] : -1, ok ? nullptr : why.c_str());
            }
        }
        if (batch) mm_modbatch_free(batch);
        batch = nullptr;
        {
            std::lock_guard<std::mutex> g(S->mu);
            S->hbm_used -= bytes;
            bytes = 0;
            if (deferred) {
                deferred = false;
                S->deferred--;
            }
        }
        if (be_open && S->close) S->close(S->user, &be);
        be_open = false;
    }
};

// The HIP backend as a pool's table: one context (stream) per pool.
struct HipUser {
    mm_options o;
};
int hip_open(void* user, mm_backend* out) { return mm_hip_open(&((HipUser*)user)->o, out); }
void hip_close(void* user, mm_backend* be) {
    (void)user;
    mm_hip_close(be);
}

}  // namespace

extern "C" {

// At L = 32,768 the register-transform step also stages its coefficients in HBM (24 n = 12 L).
uint64_t mm_server_slot_bytes(uint32_t len) { return 50ull * len + 24 + (len == 32768 ? 12ull * len : 0); }

int mm_server_new(const mm_server_options* o, mm_open_fn open, mm_close_fn close, void* user, mm_server** out) {
    *out = nullptr;
    if (!open) return fail("no backend factory");
    std::unique_ptr<mm_server> S(new mm_server());
    S->opt = *o;
    if (S->opt.base < 1 || S->opt.base > 4) return fail("base must be 1 .. 4");
    if (S->opt.pool_slots == 0) S->opt.pool_slots = 256;
    S->open = open;
    S->close = close;
    S->user = user;
    if (open(user, &S->probe)) return fail("cannot open a backend");
    S->probe_open = true;
    S->t0 = Clock::now();
    *out = S.release();
    return 0;
}

int mm_server_new_hip(const mm_server_options* o, const mm_options* hip, mm_server** out) {
    auto* u = new HipUser{*hip};
    if (mm_server_new(o, hip_open, hip_close, u, out)) {
        delete u;
        return 1;
    }
    (*out)->owned_user = u;
    return 0;
}

// Queues base^e mod N: e = N - 1 (a Fermat test) when exp is null, else the given exponent.
static int64_t submit(mm_server* S, const uint32_t* limbs, size_t nl, const uint32_t* exp, size_t ne) {
    mpz_t n, e;
    mpz_inits(n, e, nullptr);
    mpz_import(n, nl, -1, 4, 0, 0, limbs);
    if (mpz_even_p(n) || mpz_cmp_ui(n, 1) <= 0) {
        mpz_clears(n, e, nullptr);
        fail("modulus must be odd and > 1");
        return -1;
    }
    const size_t bits = mpz_sizeinbase(n, 2);
    uint32_t k, len;
    if (mm_params_for_bits((uint32_t)bits, S->probe.min_len, S->probe.max_len, &k, &len)) {
        mpz_clears(n, e, nullptr);
        fail(mm_host_last_error());
        return -1;
    }
    Request r;
    r.n.assign((bits + 31) / 32, 0);
    size_t cnt = 0;
    mpz_export(r.n.data(), &cnt, -1, 4, 0, 0, n);
    if (exp)
        mpz_import(e, ne, -1, 4, 0, 0, exp);
    else
        mpz_sub_ui(e, n, 1);
    r.bits = mpz_sgn(e) ? mpz_sizeinbase(e, 2) : 0;
    r.e.assign((r.bits + 31) / 32 + 1, 0);
    mpz_export(r.e.data(), &cnt, -1, 4, 0, 0, e);
    mpz_clears(n, e, nullptr);
    std::lock_guard<std::mutex> g(S->mu);
    if (S->stop) {
        fail("server stopped");
        return -1;
    }
    r.id = S->next_id++;
    r.t_submit = S->now();
    auto it = S->pools.find(len);
    if (it == S->pools.end()) {
        std::unique_ptr<Pool> p(new Pool());
        p->S = S;
        p->len = len;
        Pool* raw = p.get();
        it = S->pools.emplace(len, std::move(p)).first;
        raw->th = std::thread([raw] { raw->loop(); });
    }
    const int64_t id = r.id;
    it->second->waiting.push_back(std::move(r));
    it->second->cv.notify_one();
    S->stats.submitted++;
    return id;
}

int64_t mm_server_submit(mm_server* S, const uint32_t* limbs, size_t nl) { return submit(S, limbs, nl, nullptr, 0); }

int64_t mm_server_submit_exp(mm_server* S, const uint32_t* limbs, size_t nl, const uint32_t* e, size_t ne) {
    if (!e && ne) {
        fail("null exponent");
        return -1;
    }
    static const uint32_t zero = 0;
    return submit(S, limbs, nl, e ? e : &zero, e ? ne : 1);
}

size_t mm_server_poll(mm_server* S, mm_server_result* out, size_t max, double wait_s) {
    std::unique_lock<std::mutex> lk(S->mu);
    if (S->done.empty() && wait_s > 0)
        S->done_cv.wait_for(lk, std::chrono::duration<double>(wait_s), [&] { return !S->done.empty(); });
    size_t got = 0;
    while (got < max && !S->done.empty()) {
        out[got++] = S->done.front();
        S->done.pop_front();
    }
    return got;
}

int mm_server_raw(mm_server* S, int64_t id, uint32_t* out, size_t max, size_t* nl) {
    std::lock_guard<std::mutex> g(S->mu);
    auto it = S->raws.find(id);
    if (it == S->raws.end()) return fail("no raw value kept for request " + std::to_string(id));
    *nl = it->second.size();
    std::copy(it->second.begin(), it->second.begin() + std::min(max, it->second.size()), out);
    S->raws.erase(it);
    return 0;
}

int mm_server_stats_get(mm_server* S, mm_server_stats* out) {
    std::lock_guard<std::mutex> g(S->mu);
    *out = S->stats;
    out->pools = (uint32_t)S->pools.size();
    out->hbm_used = S->hbm_used;
    out->waiting = 0;
    for (auto& p : S->pools) out->waiting += p.second->waiting.size();
    return 0;
}

const char* mm_server_last_error(mm_server* S) {
    if (S) {
        std::lock_guard<std::mutex> g(S->mu);
        if (!S->last_error_text.empty()) {
            g_err = S->last_error_text;
        }
    }
    return g_err.c_str();
}

void mm_server_free(mm_server* S) {
    if (!S) return;
    {
        std::lock_guard<std::mutex> g(S->mu);
        S->stop = true;
        for (auto& p : S->pools) p.second->cv.notify_all();
    }
    for (auto& p : S->pools)
        if (p.second->th.joinable()) p.second->th.join();
    S->pools.clear();
    if (S->probe_open && S->close) S->close(S->user, &S->probe);
    delete (HipUser*)S->owned_user;
    delete S;
}

}  // extern "C"
