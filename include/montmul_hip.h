/*
 * montmul_hip.h - C ABI of the MI355X (gfx950) backend for batched Montgomery squaring.
 *
 * Two layers, both plain C types, no exceptions across the boundary:
 *
 *  1. The backend (one batch of slots on one device). These entry points are what a
 *     foreign-language `impl Backend for Hip` binds; each replaces one method of the
 *     reference's `Backend`/`Limits` traits (montmul/src/transform.rs:11-20, 61-77), whose
 *     GPU implementation is `Gpu` (montmul/src/gpu/mod.rs:272-398).
 *
 *  2. The host API (`ModBatch`, `fermat`, `Params`; montmul/src/batch.rs:9-198), written in
 *     C++ over any backend given as an `mm_backend` table, so that the same host code drives
 *     this library's `Hip` backend or any other implementation of the table.
 *
 * Conventions
 *  - Status codes: 0 = success, nonzero = error; the message of the calling thread's last
 *    error is returned by mm_last_error(). `run` and `mul` return MM_OK (-1) on success, the
 *    index (>= 0) of the first slot whose carry check failed (the reference panics with
 *    "... slot i", gpu/mod.rs:179-184), or MM_EFAULT (-2) on any other failure.
 *  - Every call that touches the device makes the batch's device current on the calling
 *    thread first (HIP's current device is per thread).
 *  - A backend context (mm_hip_open) has one stream, shared by all of its batches: drive a
 *    context from one thread at a time. Separate contexts, on one device or several, may be
 *    used from separate threads concurrently.
 *  - Limbs are 32-bit little-endian: limb q holds bits 32q .. 32q+31.
 *  - Exponents are most significant bit first: bit p is (words[p / 32] >> (31 - p % 32)) & 1
 *    (transform.rs:67).
 */
#ifndef MONTMUL_HIP_H
#define MONTMUL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MM_OK ((int64_t)-1)
#define MM_EFAULT ((int64_t)-2)

/* Static data of one batch. Replaces `BatchSetup` (transform.rs:49-59). */
typedef struct mm_batch_setup {
    uint32_t k;              /* limbs per residue */
    uint32_t len;            /* transform length L (real coefficients; L/2 complex elements) */
    uint32_t base;           /* c, the multiplier applied on exponent bits equal to 1 */
    uint32_t slots;          /* number of moduli */
    const uint32_t *n;       /* slots * k limbs: N */
    const uint32_t *nprime;  /* slots * k limbs: N' = -N^-1 mod R, R = 2^(32k) */
    const uint32_t *two_n;   /* slots * k limbs: 2N */
} mm_batch_setup;

/*
 * A backend as a table of functions over opaque state. Replaces the generic parameter
 * `B: Backend + Limits` of `ModBatch<B>` (batch.rs:53-58, transform.rs:11-20, 61-77).
 */
typedef struct mm_backend {
    void *self;
    /* Backend::create (transform.rs:64). Writes an opaque state to *state. */
    int (*create)(void *self, const mm_batch_setup *setup, void **state);
    void (*destroy)(void *state);
    /* Backend::load (transform.rs:66): k limbs of a Montgomery-form value in [0, 2N). */
    int (*load)(void *state, size_t slot, const uint32_t *limbs);
    /* Backend::set_exponent (transform.rs:68); resets the slot's position to 0. */
    int (*set_exponent)(void *state, size_t slot, const uint32_t *words, size_t nwords, uint64_t nbits);
    /* Backend::run (transform.rs:70): up to `bits` steps on every slot with bits left. */
    int64_t (*run)(void *state, uint64_t bits);
    /* Backend::mul (transform.rs:72): slot out <- a * b * R^-1 (shared modulus). */
    int64_t (*mul)(void *state, size_t out, size_t a, size_t b);
    /* Backend::read (transform.rs:74): k limbs of slot `slot`. */
    int (*read)(void *state, size_t slot, uint32_t *out);
    /* Backend::remaining (transform.rs:76). */
    uint64_t (*remaining)(void *state, size_t slot);
    /* Reprograms slot `slot` for another modulus of the batch's k and L (N, N' and 2N as k
     * limbs each, as in mm_batch_setup) and clears its exponent; x is left as is. NULL when the
     * backend cannot. No reference counterpart: it admits new work into a running batch
     * (serving) without rebuilding the batch. */
    int (*set_modulus)(void *state, size_t slot, const uint32_t *n, const uint32_t *nprime, const uint32_t *two_n);
    /* Runs up to `bits` steps on every slot with bits left in `count` states of this backend
     * together (typically one per length), as `run` on each would; returns MM_OK, the first
     * failing slot with its state's index in *which, or MM_EFAULT. NULL when the backend has
     * no joint run (the host then runs the states one after another). No reference
     * counterpart: the reference's batches hold one length each (batch.rs:64-77). */
    int64_t (*run_group)(void *const *states, size_t count, uint64_t bits, size_t *which);
    /* Every slot whose carry check failed in the state's latest run (or mul), ascending: writes
     * up to `max` of them and returns their count (0 after a clean run). run returns the first
     * of them; the reference panics on the first (gpu/mod.rs:179-184). NULL when the backend
     * reports only the first. */
    size_t (*failed_slots)(void *state, uint32_t *out, size_t max);
    /* Message of the calling thread's last error of this backend. */
    const char *(*last_error)(void);
    /* Limits::coefficient_bound (transform.rs:13), as a 128-bit value. */
    uint64_t bound_lo, bound_hi;
    /* Limits::supported_lengths (transform.rs:15): every power of two in [min_len, max_len]. */
    uint32_t min_len, max_len;
    /* Limits::full_k (transform.rs:17): batches use k = L/2. */
    int32_t full_k;
    /* Keeps on the device, for `count` slots from `first`, the two raw values that stand for 1:
     * R mod N and R mod N + N, k limbs each (2k limbs per slot, slot after slot), for
     * is_one_many. NULL when the backend has no device-side test (ModBatch then reads each
     * slot). No reference counterpart, like is_one_many. */
    int (*set_ones)(void *state, size_t first, size_t count, const uint32_t *targets);
    /* ModBatch::is_one (batch.rs:179-183) on the device for `count` slots: out[i] = 1 if slot
     * slots[i]'s raw value equals one of the two values set_ones kept for it, else 0. One small
     * copy back for all of them instead of a read per slot (serving retires finished slots
     * with it). Returns 0, or nonzero (a slot out of range or without kept values). NULL as
     * set_ones. */
    int (*is_one_many)(void *state, const uint32_t *slots, size_t count, int32_t *out);
} mm_backend;

/* ------------------------------------------------------------------------------------ */
/* Layer 1: the Hip backend.                                                             */
/* ------------------------------------------------------------------------------------ */

/* Replaces `GpuOptions` (gpu/mod.rs:28-45). */
typedef struct mm_options {
    int32_t device;          /* HIP device ordinal */
    int32_t force_global;    /* keep transform vectors in HBM even when LDS would hold them */
    int32_t wraparound;      /* reduction through the half-length cyclic product; k = L/2 */
    uint32_t max_launch_ms;  /* target duration of one kernel launch; 0 = no splitting */
    int32_t prefer_lds;      /* keep vectors in LDS whenever they fit (the reference planner's
                              * rule, plan.rs:26-42); by default the register-transform kernel
                              * (wraparound, L = 1,024 .. 8,192) runs in global mode, whose smaller
                              * LDS footprint fits more workgroups per CU and is faster there */
    int32_t autotune;        /* time every plan that can run a new batch's length (radix-8
                              * register kernel or radix-4 kernel, global mode with LDS tiles or
                              * LDS-resident, the register kernel's exchange tile, one or as many
                              * slots per CU as fit) for 16 steps on the batch's own slots and keep
                              * the fastest; the choice is cached per length and slot-count octave */
    uint32_t tile;           /* without autotune, global mode: the register kernel's exchange tile
                              * in elements (0: the default, half the vector at L = 1,024 .. 4,096,
                              * 2,048 at L = 8,192 and 4,096 at 16,384; also built: the whole
                              * vector at L = 1,024 .. 4,096 and 4,096 at L = 8,192), or for the
                              * radix-4 kernel 2,048 to run its narrow passes in an LDS tile at
                              * any length (by default from L = 65,536) */
    uint32_t slots_per_cu;   /* without autotune: 1 pads each workgroup's LDS so that one slot runs
                              * per CU; 0: as many as fit */
} mm_options;

/* Launch plan of a batch. Replaces `LaunchPlan`/`Mode` (gpu/plan.rs:5-22). */
typedef struct mm_plan_info {
    int32_t global;          /* 0: vector resident in LDS; 1: vector in HBM, LDS tiles */
    uint32_t threads;        /* threads per workgroup */
    uint32_t lds_bytes;      /* dynamic LDS per workgroup */
    uint32_t slots_per_wg;   /* slots handled by one workgroup */
    uint32_t kernel;         /* 4: radix-4 step (LDS or global); 8: register-resident radix-8 step */
    int32_t tuned;           /* 1 when the autotuner chose this plan */
    uint32_t tile;           /* global mode: the LDS tile in elements (the register kernel's exchange
                              * tile, or the radix-4 kernel's tile for its narrow passes); else 0 */
    uint32_t slots_per_cu;   /* 0: as many workgroups per CU as fit; 1: LDS padded to one per CU */
} mm_plan_info;

/* Replaces Gpu::device_count (gpu/mod.rs:70-72). Returns -1 on error. */
int mm_device_count(void);
/* Replaces Gpu::with_options (gpu/mod.rs:62-67): fills `out` with this backend's table. */
int mm_hip_open(const mm_options *opts, mm_backend *out);
/* Releases what mm_hip_open allocated; every state created from it must be destroyed first. */
void mm_hip_close(mm_backend *be);

/* The table's functions, also callable directly with the `self`/state of a Hip table. */
int mm_create(void *self, const mm_batch_setup *setup, void **state);    /* gpu/mod.rs:325-370 */
void mm_destroy(void *state);                                          /* state drop */
int mm_load(void *state, size_t slot, const uint32_t *limbs);          /* gpu/mod.rs:371-375 */
int mm_set_exponent(void *state, size_t slot, const uint32_t *words, size_t nwords,
                    uint64_t nbits);                                   /* gpu/mod.rs:376-382 */
int64_t mm_run(void *state, uint64_t bits);                            /* gpu/mod.rs:383-385 */
int64_t mm_mul(void *state, size_t out, size_t a, size_t b);           /* gpu/mod.rs:386-391 */
int mm_read(void *state, size_t slot, uint32_t *out);                  /* gpu/mod.rs:392-394 */
uint64_t mm_remaining(void *state, size_t slot);                       /* gpu/mod.rs:395-397 */
const char *mm_last_error(void);

/* GPU-only extras (gpu/mod.rs:79-111). */
/* Gpu::run_bare (gpu/mod.rs:109-111): each step is the squaring alone; x is left unchanged. */
int64_t mm_run_bare(void *state, uint64_t bits);
/* Uploads loaded values and exponents now instead of at the next run (gpu/mod.rs:161-176). */
int mm_push(void *state);
/* Gpu::plan_of (gpu/mod.rs:102-104). */
int mm_plan(void *state, mm_plan_info *out);
/* Gpu::free_memory, sm_count, max_dyn_smem, device_name (gpu/mod.rs:83-99). */
uint64_t mm_free_memory(void *self);
int mm_cu_count(void *self);
uint64_t mm_max_lds(void *self);
int mm_device_name(void *self, char *buf, size_t len);
/* The Hip table's set_modulus (see mm_backend): device constants of one slot rebuilt. */
int mm_set_modulus(void *state, size_t slot, const uint32_t *n, const uint32_t *nprime, const uint32_t *two_n);
/* The Hip table's run_group (see mm_backend): the states' slots are stepped by one persistent
 * launch of the mixed-length kernel per chunk. Its workgroups take per-slot descriptors (length
 * class, slot) from a queue in device memory, longest work first, and step each slot at its own
 * length. Every state must come from the same mm_hip_open. */
int64_t mm_run_group(void *const *states, size_t count, uint64_t bits, size_t *which);
/* Device time of the step kernel launches of the latest run, in milliseconds (HIP events on
 * the batch's stream), and the number of launches. No reference counterpart: measurement. */
int mm_last_run_kernel_ms(void *state, double *ms, uint32_t *launches);
/* After mm_run_bare in global mode, the squaring's output of slot `slot`'s latest step: T = c x^2
 * as 2k limbs (c = base on a 1 bit, else 1); x is the step's input. Fails in LDS-resident mode
 * and after a full run (whose steps keep T on chip where they can). No reference counterpart:
 * it lets tests check the squaring half alone. */
int mm_read_product(void *state, size_t slot, uint32_t *out);
/* The Hip table's failed_slots (see mm_backend). */
size_t mm_failed_slots(void *state, uint32_t *out, size_t max);
/* The Hip table's set_ones and is_one_many (see mm_backend): the two values of 1 per slot are
 * kept beside x (8k bytes per slot); one workgroup per listed slot compares its x with them.
 * A slot's kept values are dropped when mm_set_modulus replaces its modulus. */
int mm_set_ones(void *state, size_t first, size_t count, const uint32_t *targets);
int mm_is_one_many(void *state, const uint32_t *slots, size_t count, int32_t *out);
/* mm_mul that allocates nothing. No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same result and return values as mm_mul; its device scratch (48 n + 12 k bytes
 * for L = 2n) comes from a fixed 4 MiB array that each GPU holds for the life of the process,
 * so lengths above 131,072 are rejected with an error (MM_EFAULT). Loaded rows of x are copied
 * up first; exponents set on the host wait for the next run. Every call on one GPU shares that
 * GPU's array: calls that overlap on one GPU (from contexts driven by different threads) race. */
int64_t mm_mul_noalloc(void *state, size_t out, size_t a, size_t b);

/* ------------------------------------------------------------------------------------ */
/* The transform layer of a Hip table (trait Transform, transform.rs:22-46; Gpu's impl,      */
/* gpu/mod.rs:286-319). A buffer holds `count` vectors of length L = len: L/2 elements of  */
/* GF(M31^2) x GF(M61^2) each, element j carrying coefficient j (real part) and j + L/2     */
/* (imaginary part); forward, pointwise and inverse give products modulo X^L + 1, exact      */
/* while coefficients stay below Q = M31 * M61. Spectra are in the reference's order.       */
/* Functions return 0 or nonzero with mm_last_error().                                      */
/* ------------------------------------------------------------------------------------ */

typedef struct mm_tbuf mm_tbuf;
typedef struct mm_tscales mm_tscales;

/* Transform::upload (transform.rs:28): len * count coefficients, row-major, untransformed. */
int mm_t_upload(void *self, const uint32_t *coeffs, uint32_t len, uint32_t count, mm_tbuf **out);
/* Gpu::load_forward (gpu/mod.rs:218-225): weights and forward-transforms `count` strings of
 * k <= len/2 limbs each (row-major), zero above. */
int mm_t_load_forward(void *self, const uint32_t *limbs, uint32_t k, uint32_t len, uint32_t count, mm_tbuf **out);
/* Transform::download (transform.rs:30): len * count exact coefficients in [0, Q), written as
 * two 64-bit words (low, high) each. */
int mm_t_download(const mm_tbuf *b, uint64_t *out);
/* Gpu::download_raw (gpu/mod.rs:212-215): the elements without CRT, as (re, im) pairs:
 * len/2 * count pairs of 32-bit words (M31) and of 64-bit words (M61). */
int mm_t_download_raw(const mm_tbuf *b, uint32_t *f1, uint64_t *f2);
/* Transform::upload_scales (transform.rs:31). */
int mm_t_upload_scales(void *self, const uint32_t *scales, uint32_t count, mm_tscales **out);
/* Transform::forward, inverse (transform.rs:33-34), in place; len and count must match. */
int mm_t_forward(mm_tbuf *b, uint32_t len, uint32_t count);
int mm_t_inverse(mm_tbuf *b, uint32_t len, uint32_t count);
/* Transform::pointwise (transform.rs:35-45): out[i] = scale[i] a[i] (.) b[i]; scales may be NULL. */
int mm_t_pointwise(mm_tbuf *out, const mm_tbuf *a, const mm_tbuf *b, const mm_tscales *scales, uint32_t len,
                   uint32_t count);
/* Drops of a buffer or a scale vector (device memory freed). */
void mm_t_free(mm_tbuf *b);
void mm_t_free_scales(mm_tscales *s);

/* Issue-rate probe (no reference counterpart: measurement). Runs `blocks` workgroups of 256
 * threads, each thread `iters` iterations of 8 independent chains of one instruction: op 0
 * v_mad_u64_u32, 1 v_add_u32, 2 v_mul_hi_u32, 3 v_lshl_add_u64 (64-bit add), 4 v_mad_u64_u32 and
 * v_add_u32 in turn. Writes the launch's device time in ms and a checksum of the results. */
int mm_probe_issue(int device, int op, uint32_t blocks, uint32_t iters, double *ms, uint64_t *checksum);
const char *mm_probe_last_error(void);

/* ------------------------------------------------------------------------------------ */
/* Several backends as one (reference: Multi, montmul/src/multi.rs:11-156).               */
/* ------------------------------------------------------------------------------------ */

/* Fills `out` with a table whose batches split their slots into contiguous shards, one per
 * part (sizes differing by at most one, multi.rs:39-42), created and run on one host thread
 * per shard. Per-slot calls go to the slot's shard; a mul across shards runs on the host with
 * GMP (multi.rs:128-138). A shard's failure is relabelled as the reference's panics are
 * (multi.rs:58-70): "... slot i" becomes "... shard s, slot g" with g the batch's slot, any
 * other message "shard s (first slot f): ...". The parts (for instance one mm_hip_open per
 * device, or several on one device) must outlive the table and have equal limits; each part
 * is driven from one thread at a time. */
int mm_multi_open(const mm_backend *parts, size_t count, mm_backend *out);
void mm_multi_close(mm_backend *be);
size_t mm_multi_parts(const mm_backend *be);
const char *mm_multi_last_error(void);

/* ------------------------------------------------------------------------------------ */
/* Layer 2: the host API over any backend table.                                          */
/* ------------------------------------------------------------------------------------ */

typedef struct mm_modbatch mm_modbatch;

/* Params::for_bits (batch.rs:20-25): k = ceil((bits + 4) / 32), L = the smallest supported
 * length >= 2k. Returns nonzero when no supported length is large enough. */
int mm_params_for_bits(uint32_t bits, uint32_t min_len, uint32_t max_len, uint32_t *k, uint32_t *len);
/* Params::check_exact (batch.rs:27-31): max(c, 1) * k * (2^32 - 1)^2 < bound. */
int mm_params_check_exact(uint64_t k, uint32_t c, uint64_t bound_lo, uint64_t bound_hi);

/*
 * ModBatch::new (batch.rs:62-105). Moduli are given as `count` signed integers: modulus i
 * has limbs limbs[offsets[i] .. offsets[i+1]) (little-endian magnitude) and sign signs[i]
 * (-1, 0 or 1). The backend table must outlive the batch.
 */
int mm_modbatch_new(const mm_backend *be, size_t count, const uint32_t *limbs, const size_t *offsets,
                    const int8_t *signs, uint32_t base, mm_modbatch **out);
void mm_modbatch_free(mm_modbatch *b);
/* ModBatch::params (batch.rs:107-109) and slots (110-112). */
void mm_modbatch_params(const mm_modbatch *b, uint32_t *k, uint32_t *len);
size_t mm_modbatch_slots(const mm_modbatch *b);
/* The backend state, for backend-specific extras (ModBatch::backend_and_state, batch.rs:120-122). */
void *mm_modbatch_state(mm_modbatch *b);
/* ModBatch::load (batch.rs:125-130): any integer x, stored as x * R mod N. */
int mm_modbatch_load(mm_modbatch *b, size_t slot, int sign, const uint32_t *limbs, size_t nlimbs);
/* ModBatch::load_raw (batch.rs:133-137): fails unless 0 <= xm < 2N. */
int mm_modbatch_load_raw(mm_modbatch *b, size_t slot, int sign, const uint32_t *limbs, size_t nlimbs);
/* ModBatch::set_exponent (batch.rs:140-149); e >= 0. */
int mm_modbatch_set_exponent(mm_modbatch *b, size_t slot, int sign, const uint32_t *limbs, size_t nlimbs);
/* ModBatch::run (batch.rs:152-154). Returns as mm_run. */
int64_t mm_modbatch_run(mm_modbatch *b, uint64_t bits);
/* ModBatch::remaining (batch.rs:157-159). */
uint64_t mm_modbatch_remaining(mm_modbatch *b, size_t slot);
/* ModBatch::mul (batch.rs:162-166); slots must share a modulus. Returns as mm_mul. */
int64_t mm_modbatch_mul(mm_modbatch *b, size_t out, size_t a, size_t c);
/* Every slot whose carry check failed in the latest run or mul, ascending (the backend's
 * failed_slots, or the one slot run returned when the table has none): writes up to `max` and
 * returns the count. No reference counterpart (the reference panics on the first slot). */
size_t mm_modbatch_failed_slots(mm_modbatch *b, uint32_t *out, size_t max);
/* ModBatch::read_raw (batch.rs:168-170): k limbs in [0, 2N). */
int mm_modbatch_read_raw(mm_modbatch *b, size_t slot, uint32_t *out);
/* ModBatch::read (batch.rs:173-176): k limbs of raw * R^-1 mod N, in [0, N). */
int mm_modbatch_read(mm_modbatch *b, size_t slot, uint32_t *out);
/* ModBatch::is_one (batch.rs:179-183): 1 if the raw value is R mod N or R mod N + N, 0 if
 * not, negative on error. */
int mm_modbatch_is_one(mm_modbatch *b, size_t slot);
/* is_one for `count` slots at once: out[i] = 1 or 0 for slot slots[i]. On a backend with
 * is_one_many the test runs on the device; otherwise each slot is read. Returns 0 or nonzero
 * on error. No reference counterpart (the reference calls is_one per slot, batch.rs:179-183). */
int mm_modbatch_is_one_many(mm_modbatch *b, const uint32_t *slots, size_t count, int32_t *out);
/* k limbs of slot i's modulus (ModBatch::modulus, batch.rs:113-115). */
int mm_modbatch_modulus(mm_modbatch *b, size_t slot, uint32_t *out);
/* Replaces slot i's modulus by another odd N > 1 needing the batch's L (k limbs at most): N', R
 * mod N, R^-1 mod N and 2N are computed as in mm_modbatch_new and the backend's set_modulus
 * rebuilds the slot; its exponent is cleared. Fails when the backend has no set_modulus. No
 * reference counterpart (the reference rebuilds a ModBatch): admission for serving. */
int mm_modbatch_set_modulus(mm_modbatch *b, size_t slot, int sign, const uint32_t *limbs, size_t nlimbs);

/* A batch of moduli needing different transform lengths (no reference counterpart: ModBatch::new
 * rejects mixed lengths, batch.rs:64-77). Each modulus gets the Params of its own size
 * (batch.rs:20-25); the moduli of one length form one ModBatch, and run() steps all of them
 * together through the backend's run_group (one after another if it has none). Slot i is the
 * i-th modulus given; the per-slot functions behave as the mm_modbatch ones with that slot's
 * k limbs, and run returns MM_OK, the first failing slot of this batch, or MM_EFAULT. */
typedef struct mm_mixbatch mm_mixbatch;
int mm_mixbatch_new(const mm_backend *be, size_t count, const uint32_t *limbs, const size_t *offsets,
                    const int8_t *signs, uint32_t base, mm_mixbatch **out);
void mm_mixbatch_free(mm_mixbatch *b);
size_t mm_mixbatch_slots(const mm_mixbatch *b);
size_t mm_mixbatch_groups(const mm_mixbatch *b);
int mm_mixbatch_params(const mm_mixbatch *b, size_t slot, uint32_t *k, uint32_t *len);
int mm_mixbatch_load(mm_mixbatch *b, size_t slot, int sign, const uint32_t *limbs, size_t nlimbs);
int mm_mixbatch_load_raw(mm_mixbatch *b, size_t slot, int sign, const uint32_t *limbs, size_t nlimbs);
int mm_mixbatch_set_exponent(mm_mixbatch *b, size_t slot, int sign, const uint32_t *limbs, size_t nlimbs);
int64_t mm_mixbatch_run(mm_mixbatch *b, uint64_t bits);
size_t mm_mixbatch_failed_slots(mm_mixbatch *b, uint32_t *out, size_t max);
uint64_t mm_mixbatch_remaining(mm_mixbatch *b, size_t slot);
int mm_mixbatch_read_raw(mm_mixbatch *b, size_t slot, uint32_t *out);
int mm_mixbatch_read(mm_mixbatch *b, size_t slot, uint32_t *out);
int mm_mixbatch_is_one(mm_mixbatch *b, size_t slot);
/* mm_modbatch_is_one_many over a mixed batch: each length's listed slots are tested together. */
int mm_mixbatch_is_one_many(mm_mixbatch *b, const uint32_t *slots, size_t count, int32_t *out);
int mm_mixbatch_check_pow(mm_mixbatch *b, size_t slot, uint32_t base, const uint32_t *e, size_t ne);
/* The backend state of group `group` (groups in ascending L), NULL if out of range. */
void *mm_mixbatch_group_state(mm_mixbatch *b, size_t group);

/* fermat (batch.rs:187-198): base-`base` Fermat PRP test of every modulus; prp[i] = 1 or 0. */
int mm_fermat(const mm_backend *be, size_t count, const uint32_t *limbs, const size_t *offsets,
              const int8_t *signs, uint32_t base, uint64_t chunk, uint8_t *prp);

/* The score's check of one slot (montmul/src/score.rs:107-115): no exponent bits left and
 * read(slot) == base^e mod N, e given as little-endian limbs. 1 if so, 0 if not, < 0 on error. */
int mm_modbatch_check_pow(mm_modbatch *b, size_t slot, uint32_t base, const uint32_t *e, size_t ne);
/* The same check on slots first_slot .. first_slot+count-1 (the j-th one's exponent: limbs
 * exps[offsets[j] .. offsets[j+1])). The slots are read on the calling thread and GMP's powm
 * runs on `threads` threads. Returns -1 when every slot matches, else the lowest slot that does
 * not (or has bits left), or -2 on an error. No reference counterpart: the score checks a few
 * slots; this checks a whole batch in parts a caller can report progress between. */
int64_t mm_modbatch_check_pow_range(mm_modbatch *b, uint32_t base, size_t first_slot, size_t count,
                                    const uint32_t *exps, const size_t *offsets, uint32_t threads);

/* The score workload (montmul/src/score.rs:45-61): `slots` moduli of `bits` bits drawn from
 * GMP's default random state seeded with `seed`, top bit and bit 0 set, and per modulus the top
 * `ebits` bits of N - 1. Writes ceil(bits/32) limbs per modulus and ceil(ebits/32) per exponent. */
int mm_score_workload(uint32_t bits, size_t slots, uint32_t ebits, uint64_t seed, uint32_t *moduli,
                      uint32_t *exps);

/* Message of the calling thread's last host-layer error. */
const char *mm_host_last_error(void);

/* ------------------------------------------------------------------------------------ */
/* Serving: base-c Fermat PRP requests over length pools (no reference counterpart; the  */
/* reference runs fixed batches). Each transform length gets a pool, a batch of           */
/* `pool_slots` slots on a backend table of its own driven by one host thread: a round     */
/* admits waiting requests into free slots (mm_modbatch_set_modulus), runs a chunk of      */
/* steps sized to `target_ms`, and retires finished slots (is_one on that slot's raw).     */
/* Pools run concurrently (on the HIP backend: one stream each).                           */
/* ------------------------------------------------------------------------------------ */

typedef struct mm_server mm_server;
/* Fills `out` with a fresh backend table (each pool opens its own) / releases it. */
typedef int (*mm_open_fn)(void *user, mm_backend *out);
typedef void (*mm_close_fn)(void *user, mm_backend *be);

typedef struct mm_server_options {
    uint32_t base;        /* Fermat base c, 1 .. 4 */
    uint32_t pool_slots;  /* slots per length pool (0: 256) */
    double target_ms;     /* wall time of one pool round; steps per round follow from it */
    uint32_t max_chunk;   /* most steps per round (0: 4096) */
    int32_t keep_raw;     /* keep each result's final raw limbs for mm_server_raw */
    uint64_t hbm_bytes;   /* device memory all pools together may take (0: no limit). A pool's
                           * slot count is cut to what fits; a length that does not fit waits,
                           * and idle pools give their memory back while a length is waiting */
} mm_server_options;

typedef struct mm_server_result {
    int64_t id;
    int32_t prp;          /* 1: c^(N-1) = 1 mod N; 0: not; -1: failed */
    uint32_t len;         /* transform length of its pool */
    uint64_t bits;        /* bits of N - 1 (the squarings it took) */
    double submit_s, admit_s, done_s;  /* seconds since the server started */
} mm_server_result;

typedef struct mm_server_stats {
    uint64_t submitted, admitted, completed, failed, rounds, squarings, waiting;
    double busy_ms;       /* wall time of the pools' runs, summed over pools */
    uint32_t pools;
    uint64_t hbm_used;    /* device memory the live pools hold, by mm_server_slot_bytes */
    uint64_t releases;    /* idle pools that gave their memory back to a waiting length */
} mm_server_stats;

/* Device memory of one pool slot at transform length L in the server's accounting: on the HIP
 * backend in global mode, x and the exponent (k words each), N'-hat and A|B (24 L bytes), the
 * work vectors (12 L) and the limb staging (6 L), with k = L/2: 46 L + 16 bytes, plus 12 L at
 * L = 32,768 for the coefficient staging of its register-transform step. */
uint64_t mm_server_slot_bytes(uint32_t len);

int mm_server_new(const mm_server_options *opts, mm_open_fn open, mm_close_fn close, void *user, mm_server **out);
/* A server whose pools are HIP backends opened with `hip` (one context and stream per pool). */
int mm_server_new_hip(const mm_server_options *opts, const mm_options *hip, mm_server **out);
/* Queues a Fermat test of the odd N > 1 given as little-endian limbs. Returns its id (>= 0),
 * or -1 with mm_server_last_error(). */
int64_t mm_server_submit(mm_server *s, const uint32_t *limbs, size_t nlimbs);
/* Queues base^e mod N for a given exponent e >= 0 (little-endian limbs) instead of N - 1; its
 * result's prp is 1 when base^e = 1 mod N, and `bits` is the bit length of e. */
int64_t mm_server_submit_exp(mm_server *s, const uint32_t *limbs, size_t nlimbs, const uint32_t *e, size_t ne);
/* Up to `max` completions, waiting up to `wait_s` seconds for the first. Returns the count. */
size_t mm_server_poll(mm_server *s, mm_server_result *out, size_t max, double wait_s);
/* The final raw Montgomery limbs of a completed request (when keep_raw), once. */
int mm_server_raw(mm_server *s, int64_t id, uint32_t *out, size_t max_limbs, size_t *nlimbs);
int mm_server_stats_get(mm_server *s, mm_server_stats *out);
const char *mm_server_last_error(mm_server *s);
/* Stops every pool (requests in flight are dropped) and frees the server. */
void mm_server_free(mm_server *s);

#ifdef __cplusplus
}
#endif

#endif /* MONTMUL_HIP_H */
