#!/usr/bin/env python3
"""Headline benchmark: modular squarings per second at 100,000-bit moduli on one MI355X.

Workload (the reference's score workload, montmul/src/score.rs:26-28, 45-61, 93-123): random
odd 100,000-bit moduli from GMP's generator seeded with 2026 (top bit and bit 0 set),
k = 4,096 limbs, L = 8,192, base 3, exponent = the top 2,000 bits of N - 1 in every slot.
A bench "step" is one pass of the hot path over the batch: re-arm every slot (load 1, set the
exponent) and run(2,000), i.e. slots x 2,000 modular squarings. Warm-up passes are the same
run(2,000), so every step-kernel launch of a run with a fixed slot count does the same work.

The slot count is the best of 256 x {1, 2, 4, 8, 16} (CUs x {1, 2, 4, 8, 16}, as the
reference's candidates, main.rs:199-216) at 200-bit exponents, unless --slots is given. The
headline is global mode (transform vectors in HBM, as the workload specifies); LDS mode
(vectors resident in LDS) is measured and reported beside it. Slot 0, the last slot and two
seeded random slots must equal 3^e mod N (GMP powm) or no number is reported
(score.rs:105-115); after the timed passes every headline slot is checked the same way, on
the host's threads (--no-check-all skips it).

Prints one JSON line (the last line of stdout). Progress goes to stderr.
"""

import argparse
import json
import os
import random
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "squarings/s at 100k-bit, 1 MI355X"
HBM_PEAK = 8.0e12          # MI355X HBM3E, bytes/s
# Attainable VALU issue of the step kernel's own instruction mix, in wave instructions per CU
# per cycle: the issue rates measured by mm_probe_issue (32-bit operations 1.37, 64-bit 0.95)
# weighted by the kernel's INT32 / INT64 / other counts (profiles/valu_issue_probe.md), over
# 256 CUs at 2.4 GHz.
VALU_MIX_PER_CU_CYCLE = {8192: 1.22}
VALU_PROBE = "profiles/valu_issue_probe.md"
# Per slot-step of the global-mode step kernel, by transform length L, from the rocprofv3 summary
# named in PROFILE: HBM bytes moved (FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000, scales from a
# 1 GiB device copy in the same run) and VALU wave instructions (SQ_INSTS_VALU). Counters cannot
# be read from inside this process; the summaries are re-measured whenever the kernel changes.
PROFILE = {8192: "profiles/k_steps_v2_global_100kbit_4096slots.md",
           32768: "profiles/k_steps_v2_split_300kbit_1024slots.md"}
TRAFFIC_PER_SLOT_STEP = {8192: 324_955, 32768: 9_757_633}
VALU_PER_SLOT_STEP = {8192: 170_223, 32768: 832_892}
REF_5090 = 1_536_441.0     # reference, RTX 5090, 170 slots (README.md:14)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def arm(batch, exps):
    for i, e in enumerate(exps):
        batch.load(i, 1)
        batch.set_exponent(i, e)


def median(v):
    return statistics.median(v)


def mad(v):
    m = median(v)
    return median([abs(x - m) for x in v])


def measure(mm, hip, moduli, exps, slots, ebits, repeats, warmup):
    """`warmup` untimed and `repeats` timed re-armed run(ebits) passes."""
    b = mm.ModBatch(hip, moduli[:slots], 3)
    ex = exps[:slots]
    for _ in range(warmup):
        arm(b, ex)
        b.run(ebits)
    rates, secs, kms, ups = [], [], [], []
    for _ in range(repeats):
        arm(b, ex)
        t_up = time.perf_counter()
        mm.Hip.push(b)  # inputs resident in HBM before the clock starts
        t0 = time.perf_counter()
        b.run(ebits)
        dt = time.perf_counter() - t0
        ms, launches = mm.Hip.last_run_kernel_ms(b)
        rates.append(slots * ebits / dt)
        secs.append(dt)
        ups.append(t0 - t_up)
        kms.append(ms / max(1, launches))
    return b, rates, secs, kms, ups


def check(b, exps, seed):
    slots = b.slots()
    rng = random.Random(seed + 1)
    picked = [0, slots - 1] + [rng.randrange(slots) for _ in range(2)]
    for i in picked:
        if not b.check_pow(i, 3, exps[i]):
            raise SystemExit(f"slot {i} disagrees with pow_mod; no score reported")
    return picked


def sweep(mm, hip, moduli, exps, candidates, sweep_ebits):
    best, res = None, {}
    for s in candidates:
        b, rates, _, _, _ = measure(mm, hip, moduli, exps, s, sweep_ebits, 1, 1)
        b.free()
        res[s] = rates[0]
        log(f"  sweep slots {s}: {rates[0]:.4g} sq/s")
        if best is None or rates[0] > res[best]:
            best = s
    return best, res


def cpu_baseline(bits, threads, gmp_steps=8_000, ntt_steps=1_000, single_steps=2_000):
    """The oracle's restatements of the reference's CPU backends on this host's cores, on the
    score workload (seed 2026, the top bits of N - 1 as exponents), as main.rs:214 scores them:
    slots = threads x {1, 2}, the better rate kept. Gmp (closed-form REDC on GMP integers,
    gmp.rs) runs `gmp_steps` steps per slot count (2 to 5 s each on 16 threads), the Cpu restatement (NTT with
    wraparound, cpu.rs) `ntt_steps` (1.5 to 3 s each), and Gmp once more with one slot on one
    thread for `single_steps` steps: seconds per squaring beside the reference's 2.709e-4 s on an
    Apple M5 core (score.rs:11-13)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import montmul_amd as mm
    from oracle_backends import Cpu, Gmp

    steps = max(gmp_steps, ntt_steps, single_steps)
    moduli, exps = mm.score_workload(bits, 2 * threads, steps, 2026)

    def rate(backend, slots, nsteps):
        b = mm.ModBatch(backend, moduli[:slots], 3)
        arm(b, exps[:slots])
        b.run(16)
        arm(b, exps[:slots])
        t0 = time.perf_counter()
        b.run(nsteps)
        dt = time.perf_counter() - t0
        b.free()
        return slots * nsteps / dt, dt

    out = {"gmp": (0, 0, 0), "cpu_ntt": (0, 0, 0)}
    for key, make, nsteps in (("gmp", lambda: Gmp(threads=threads), gmp_steps),
                              ("cpu_ntt", lambda: Cpu(wraparound=True, threads=threads), ntt_steps)):
        for slots in (threads, 2 * threads):
            r, dt = rate(make(), slots, nsteps)
            log(f"  cpu baseline {key}: {slots} slots on {threads} threads, {r:.1f} sq/s ({dt:.1f} s)")
            if r > out[key][0]:
                out[key] = (r, slots, dt)
    r1, dt1 = rate(Gmp(threads=1), 1, single_steps)
    out["gmp_single"] = (r1, 1, dt1)
    out["steps"] = {"gmp": gmp_steps, "cpu_ntt": ntt_steps, "gmp_single": single_steps}
    return out


def valu_roof(L, slot_steps_per_s):
    """The kernel's VALU issue rate (instructions per slot-step from the profile, rate from this
    run's kernel time) against the rate its instruction mix can issue, where measured."""
    achieved = VALU_PER_SLOT_STEP[L] * slot_steps_per_s
    out = {"bound": "valu", "achieved": round(achieved / 1e9, 2), "unit": "G wave-instr/s",
           "instr_per_slot_step": VALU_PER_SLOT_STEP[L], "source": PROFILE[L],
           "per_cu_per_cycle_at_2.4GHz": round(achieved / (256 * 2.4e9), 4)}
    if L in VALU_MIX_PER_CU_CYCLE:
        peak = VALU_MIX_PER_CU_CYCLE[L] * 256 * 2.4e9
        out.update(peak=round(peak / 1e9, 1), frac=round(achieved / peak, 4), peak_source=VALU_PROBE)
    return out


def host_threads():
    """Logical cores this process may run on, up to the 16 host threads of a one-GPU share."""
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return max(1, min(16, avail))


def copy_calibration(nbytes):
    """Four device-to-device copies of `nbytes` (the HIP runtime's blit kernel,
    __amd_rocclr_copyBuffer): a plain copy of known size whose FETCH_SIZE / WRITE_SIZE counters
    calibrate the step kernel's in a profiled run."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) != 0:
            raise SystemExit("copy calibration: hipMalloc failed")
    hip.hipMemset(src, 1, ctypes.c_size_t(nbytes))
    for _ in range(4):
        if hip.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3) != 0:  # hipMemcpyDeviceToDevice
            raise SystemExit("copy calibration: hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    return {"bytes_read": nbytes, "bytes_written": nbytes, "launches": 4, "kernel": "__amd_rocclr_copyBuffer"}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5, help="timed passes (repeats of run(ebits))")
    ap.add_argument("--warmup", type=int, default=1, help="untimed passes before the timed ones")
    ap.add_argument("--slots", type=int, default=0, help="slot count (0: best of 256 x {1,2,4,8,16})")
    ap.add_argument("--bits", type=int, default=100_000)
    ap.add_argument("--ebits", type=int, default=2000)
    ap.add_argument("--sweep-ebits", type=int, default=200)
    ap.add_argument("--modes", default="global,lds")
    ap.add_argument("--tile", type=int, default=0,
                    help="global mode: the exchange tile in elements (0: the default plan's)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-check-all", action="store_true",
                    help="skip the check of every headline slot against GMP's powm (about 3 minutes on 16 threads)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="also run a plain device copy of MIB MiB (4 times): the reference kernel for scaling "
                         "the FETCH_SIZE / WRITE_SIZE counters of a profiled run")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (slots would shard "
              "across devices with no exchange per step). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import montmul_amd as mm

    t_setup = time.perf_counter()
    maxslots = args.slots or 256 * 16
    moduli, exps = mm.score_workload(args.bits, maxslots, args.ebits, 2026)
    log(f"workload: {maxslots} moduli of {args.bits} bits in {time.perf_counter() - t_setup:.1f} s")
    modes = [m for m in args.modes.split(",") if m]
    results = {}
    for mode in modes:
        hip = mm.Hip(force_global=(mode == "global"), prefer_lds=(mode == "lds"), tile=args.tile if mode == "global" else 0)
        cus = hip.cu_count()
        if args.slots:
            best, sw = args.slots, {}
        else:
            cands = [cus * f for f in (1, 2, 4, 8, 16)]
            log(f"mode {mode}: sweep {cands}")
            best, sw = sweep(mm, hip, moduli, exps, cands, args.sweep_ebits)
        log(f"mode {mode}: {best} slots, {args.warmup} warm-up + {args.steps} timed passes of run({args.ebits})")
        b, rates, secs, kms, ups = measure(mm, hip, moduli, exps, best, args.ebits, args.steps, args.warmup)
        picked = check(b, exps, 2026)
        all_slots = None
        if mode == "global" and not args.no_check_all:
            # Every slot of the last timed pass against GMP's powm on the host share's threads.
            t_chk, part = time.perf_counter(), 512
            for first in range(0, best, part):
                bad = b.check_pow_range(3, first, exps[first:min(best, first + part)], host_threads())
                if bad is not None:
                    raise SystemExit(f"slot {bad} disagrees with pow_mod; no score reported")
                log(f"  slots {first}..{min(best, first + part) - 1} equal 3^e mod N")
            all_slots = {"slots": best, "equal_to_powm": True, "seconds": round(time.perf_counter() - t_chk, 1),
                         "threads": host_threads()}
            log(f"mode {mode}: all {best} slots equal 3^e mod N ({all_slots['seconds']} s)")
        plan = mm.Hip.plan_of(b)
        L, k = b.params().len, b.params().k
        # run over run_bare (score.rs:138-160) from the step kernel's own device time: the whole
        # step against the squaring alone, on the same slots and exponents.
        arm(b, exps[:best])
        mm.Hip.run_bare(b, args.ebits)
        bare_ms, bare_launches = mm.Hip.last_run_kernel_ms(b)
        overhead = statistics.mean(kms) / (bare_ms / max(1, bare_launches))
        results[mode] = dict(slots=best, rates=rates, secs=secs, kernel_ms=kms, plan=plan, L=L, k=k, ups=ups,
                             sweep={str(s): round(r) for s, r in sw.items()}, checked=picked, overhead=overhead,
                             all_slots=all_slots)
        log(f"mode {mode}: median {median(rates):.4g} sq/s, kernel {statistics.mean(kms):.1f} ms/launch")
        b.free()
        hip.close()

    head = results.get("global") or results[modes[0]]
    slots, ebits, L = head["slots"], args.ebits, head["L"]
    total = slots * ebits * len(head["secs"])
    value = total / sum(head["secs"])
    kms = statistics.mean(head["kernel_ms"])
    # Algorithmic HBM bytes per slot-step in global mode: the per-slot constant spectra N'-hat
    # and A|B (12 L bytes each, read once per step) plus x read and written (2 L each): 28 L.
    alg_bytes = 28 * L * slots * ebits
    achieved = alg_bytes / (kms * 1e-3)
    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "squarings/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * statistics.mean(head["secs"]), 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": round(value / REF_5090, 4) if args.bits == 100_000 else None,
        "vs_baseline_ref": "reference Gpu backend, RTX 5090, 170 slots, 1,536,441 squarings/s (README.md:14)",
        "dtype": "u32",
        "data": "synthetic (GMP generator seeded 2026, as the reference's score workload)",
        "config": {"workload": f"{args.bits:,}-bit moduli, L={L}, k={head['k']}, base 3, top {ebits:,} bits of N-1, " +
                               ("global mode (T, x and constants in HBM; transform exchanges through LDS tiles)"
                                if "global" in results else f"{modes[0]} mode"),
                   "bits": args.bits, "L": L, "k": head["k"], "slots": slots, "ebits": ebits,
                   "mode": "global" if "global" in results else modes[0],
                   "threads_per_wg": head["plan"]["threads"], "lds_bytes": head["plan"]["lds_bytes"],
                   "tile": head["plan"]["tile"]},
        # The same passes with the upload of x and the exponents (host to device) inside the
        # clock, as the reference's score times run() (score.rs:98-102).
        "value_incl_upload": round(total / (sum(head["secs"]) + sum(head["ups"])), 1),
        "upload_s_per_pass": round(statistics.mean(head["ups"]), 4),
        "overhead": {"run_over_run_bare": round(head["overhead"], 4),
                     "source": "step-kernel device time of run(ebits) over run_bare(ebits), same slots (score.rs:138-160)"},
        "median": round(median(head["rates"]), 1),
        "mad": round(mad(head["rates"]), 1),
        "fermat_tests_per_hour": round(value * 3600 / args.bits, 1),
        "roofline": {"bound": "hbm", "achieved": round(achieved / 1e9, 2), "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK, 5),
                     "traffic": (TRAFFIC_PER_SLOT_STEP[L] * slots * ebits) if L in TRAFFIC_PER_SLOT_STEP else None,
                     "traffic_per_slot_step": TRAFFIC_PER_SLOT_STEP.get(L), "traffic_source": PROFILE.get(L),
                     "kernel": "k_steps_v2 (global mode)", "kernel_ms_per_launch": round(kms, 3),
                     "alg_bytes_per_slot_step": 28 * L},
        "checked_slots": head["checked"],
        "checked_all_slots": head.get("all_slots"),
        # The bound that binds: VALU issue (integer work, no MFMA). Instructions per slot-step
        # from the profile, rate from this run's kernel time.
        "valu": (valu_roof(L, slots * ebits / (kms * 1e-3)) if L in VALU_PER_SLOT_STEP else None),
        "sweep": head["sweep"],
    }
    if "lds" in results and "global" in results:
        r = results["lds"]
        line["lds_mode"] = {"value": round(r["slots"] * ebits * len(r["secs"]) / sum(r["secs"]), 1),
                            "median": round(median(r["rates"]), 1), "mad": round(mad(r["rates"]), 1),
                            "slots": r["slots"], "kernel_ms_per_launch": round(statistics.mean(r["kernel_ms"]), 3),
                            "lds_bytes": r["plan"]["lds_bytes"], "sweep": r["sweep"]}
    if not args.no_cpu_baseline:
        # Every logical core this process may run on, up to the 16 host threads that a one-GPU
        # share of the GPU host allows (its worker-pool limit); both counts are reported.
        avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        threads = host_threads()
        log(f"cpu baseline on {threads} threads ({avail} logical cores available, {os.cpu_count()} in the host)")
        cb = cpu_baseline(args.bits, threads)
        g, c, g1 = cb["gmp"], cb["cpu_ntt"], cb["gmp_single"]
        line["cpu_baseline"] = {
            "value": round(g[0], 1), "unit": "squarings/s", "cores": threads, "kind": "port",
            "sample": f"oracle Gmp restatement (gmp.rs closed-form REDC), score workload at {args.bits} bits: best of "
                      f"{threads} x {{1, 2}} slots (main.rs:214) = {g[1]} slots x {cb['steps']['gmp']} steps, {g[2]:.2f} s wall",
            "logical_cores_available": avail, "host_cpu_count": os.cpu_count(),
            "cpu_ntt_value": round(c[0], 1),
            "cpu_ntt_sample": f"oracle Cpu restatement (cpu.rs NTT, wraparound): {c[1]} slots x {cb['steps']['cpu_ntt']} "
                              f"steps on {threads} threads, {c[2]:.2f} s wall",
            "gmp_single_core": {"seconds_per_squaring": g1[2] / cb["steps"]["gmp_single"],
                                "reference_m5_seconds_per_squaring": 2.709e-4,
                                "sample": f"1 slot x {cb['steps']['gmp_single']} steps on 1 thread, {g1[2]:.2f} s wall"}}
    if args.copy_calibration:
        line["copy_calibration"] = copy_calibration(args.copy_calibration << 20)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
