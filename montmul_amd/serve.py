"""Serving benchmark: Fermat PRP requests arriving as a Poisson process, served over length pools.

    python -m montmul_amd.serve [--rates 1,2,5,10,20,50,100,200] [--duration 20] \
        [--min-bits 4096] [--max-bits 262144] [--planted 0.1] [--seed 3] [--burst N]

Requests are base-3 Fermat tests (exponent N - 1) of random odd moduli whose bit sizes are
log-uniform in [min-bits, max-bits]. About `planted` of them are known primes in that range:
Mersenne primes 2^p - 1 and prime tribonacci numbers (OEIS A092835). Every result is checked:
planted primes must come back prime, and a few sampled requests are run again alone (one
request in a batch of its own) and must give the same raw Montgomery limbs.

For each rate, prints one JSON line with:
- p50 and p99 latency (submit to result) and p50 of the queueing part (submit to admit);
- aggregate squarings/s over the rate's wall time;
- the same requests run afterwards as plain per-length batches (one fermat() batch per L,
  lengths one after another), for comparison.

With --burst N, N requests arrive at once instead (one line, "rate_per_s": null). Both forms then
start with all the work in hand, so their squarings/s compare the throughput of continuous
batching across lengths with that of per-length batches.
"""

import argparse
import json
import math
import os
import random
import sys
import time

if __package__ in (None, ""):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

# Each length pool drives its own stream. With fewer hardware queues than pools (HIP's default
# is 4 per process), two pools share a queue, and one pool's admission launches wait behind the
# other's rounds: p99 queueing in a 300-request burst over five pools was 1.0 s with 4 queues
# and 0.12 s with 8. The runtime reads this before its first call, so it is set here, before
# anything touches the device, unless the caller chose a value.
os.environ.setdefault("GPU_MAX_HW_QUEUES", "8")

import montmul_amd as mm

MERSENNE_EXPONENTS = [4253, 4423, 9689, 9941, 11213, 19937, 21701, 23209, 44497, 86243, 110503, 132049, 216091]
TRIBONACCI_PRIME_INDICES = [4202, 18699, 96879]  # OEIS A092835: T(n) is prime


def tribonacci(n):
    a, b, c = 0, 0, 1
    for _ in range(n):
        a, b, c = b, c, a + b + c
    return a


def planted_primes(min_bits, max_bits):
    out = [(1 << p) - 1 for p in MERSENNE_EXPONENTS if min_bits <= p <= max_bits]
    for n in TRIBONACCI_PRIME_INDICES:
        t = tribonacci(n)
        if min_bits <= t.bit_length() <= max_bits:
            out.append(t)
    return out


def workload(rng, rate, duration, min_bits, max_bits, planted_frac, primes):
    """(arrival time, modulus, planted) for a Poisson process of `rate` per second."""
    t, out = 0.0, []
    while True:
        t += rng.expovariate(rate)
        if t >= duration:
            return out
        if primes and rng.random() < planted_frac:
            out.append((t, rng.choice(primes), True))
            continue
        bits = int(math.exp(rng.uniform(math.log(min_bits), math.log(max_bits))))
        n = rng.getrandbits(bits) | (1 << (bits - 1)) | 1
        out.append((t, n, False))


def burst_workload(rng, count, min_bits, max_bits, planted_frac, primes):
    """`count` requests arriving at time 0, drawn as in workload()."""
    out = []
    for _ in range(count):
        if primes and rng.random() < planted_frac:
            out.append((0.0, rng.choice(primes), True))
            continue
        bits = int(math.exp(rng.uniform(math.log(min_bits), math.log(max_bits))))
        out.append((0.0, rng.getrandbits(bits) | (1 << (bits - 1)) | 1, False))
    return out


def pct(v, q):
    v = sorted(v)
    return v[min(len(v) - 1, int(q * (len(v) - 1) + 0.5))] if v else None


def run_rate(args, rate, rng, primes):
    if rate is None:
        reqs = burst_workload(rng, args.burst, args.min_bits, args.max_bits, args.planted, primes)
    else:
        reqs = workload(rng, rate, args.duration, args.min_bits, args.max_bits, args.planted, primes)
    srv = mm.Server(base=3, pool_slots=args.pool_slots, target_ms=args.target_ms, keep_raw=True,
                    device=args.device, hbm_bytes=int(args.hbm_gib * (1 << 30)))
    ids, res = {}, {}
    t0 = time.perf_counter()
    for at, n, planted in reqs:
        while True:
            now = time.perf_counter() - t0
            if now >= at:
                break
            for r in srv.poll(256, min(0.01, at - now)):
                res[r["id"]] = r
        ids[srv.submit(n)] = (n, planted)
    deadline = time.perf_counter() + args.drain
    shown = time.perf_counter()
    while len(res) < len(ids) and time.perf_counter() < deadline:
        for r in srv.poll(256, 0.1):
            res[r["id"]] = r
        if time.perf_counter() - shown > 20:
            shown = time.perf_counter()
            print(f"rate {rate}: {len(res)} of {len(ids)} requests done", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    done = [r for r in res.values() if r["prp"] >= 0]
    lat = [r["done_s"] - r["submit_s"] for r in done]
    queue = [r["admit_s"] - r["submit_s"] for r in done]
    squarings = sum(r["bits"] for r in done)
    bad_planted = [i for i, (n, planted) in ids.items() if planted and i in res and res[i]["prp"] != 1]
    # alone vs mix: a few finished requests again in a batch of their own
    sample = random.Random(rate if rate is not None else args.seed).sample(sorted(res), min(args.check_alone, len(res)))
    hip = mm.Hip(device=args.device)
    mismatch = []
    for i in sample:
        n = ids[i][0]
        b = mm.ModBatch(hip, [n], 3)
        b.load(0, 1)
        b.set_exponent(0, n - 1)
        b.run(1 << 31)
        if srv.raw(i) != b.read_raw(0):
            mismatch.append(i)
        b.free()
    st = srv.stats()
    srv.close()
    line = {
        "metric": "serving: Fermat PRP requests, Poisson arrivals", "rate_per_s": rate,
        "requests": len(ids), "completed": len(done), "failed": len(res) - len(done),
        "unfinished": len(ids) - len(res), "bits": [args.min_bits, args.max_bits],
        "latency_p50_s": pct(lat, 0.5), "latency_p99_s": pct(lat, 0.99),
        "queue_p50_s": pct(queue, 0.5), "queue_p99_s": pct(queue, 0.99),
        "squarings": squarings, "wall_s": round(wall, 3), "sq_per_s": round(squarings / wall, 1),
        "pools": st["pools"], "rounds": st["rounds"], "pool_releases": st["releases"],
        "planted": sum(1 for v in ids.values() if v[1]),
        "planted_not_prime": len(bad_planted), "alone_checked": len(sample), "alone_mismatch": len(mismatch),
    }
    if not args.no_baseline and done:
        # the same moduli as plain per-length batches, one length after another
        by_len = {}
        for i in res:
            by_len.setdefault(res[i]["len"], []).append(ids[i][0])
        t1 = time.perf_counter()
        for length in sorted(by_len):
            mm.fermat(hip, by_len[length], 3, 1 << 16)
        tb = time.perf_counter() - t1
        line["batch_sq_per_s"] = round(squarings / tb, 1)
        line["batch_wall_s"] = round(tb, 3)
    hip.close()
    return line


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--rates", default="1,2,5,10,20,50,100,200")
    ap.add_argument("--duration", type=float, default=20.0, help="seconds of arrivals per rate")
    ap.add_argument("--drain", type=float, default=600.0, help="most seconds to wait for the last results")
    ap.add_argument("--min-bits", type=int, default=4096)
    ap.add_argument("--max-bits", type=int, default=262144)
    ap.add_argument("--planted", type=float, default=0.1)
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--pool-slots", type=int, default=256)
    ap.add_argument("--target-ms", type=float, default=20.0)
    ap.add_argument("--check-alone", type=int, default=4)
    ap.add_argument("--device", type=int, default=0)
    ap.add_argument("--no-baseline", action="store_true")
    ap.add_argument("--burst", type=int, default=0, help="N requests at once instead of the rate sweep")
    ap.add_argument("--hbm-gib", type=float, default=0.0,
                    help="device memory the pools may take, GiB (0: no limit; admission by footprint)")
    args = ap.parse_args()
    rng = random.Random(args.seed)
    primes = planted_primes(args.min_bits, args.max_bits)
    ok = True
    rates = [None] if args.burst else [float(r) for r in args.rates.split(",") if r]
    for rate in rates:
        line = run_rate(args, rate, rng, primes)
        print(json.dumps(line), flush=True)
        ok &= line["planted_not_prime"] == 0 and line["alone_mismatch"] == 0 and line["failed"] == 0
    sys.exit(0 if ok else 1)


if __name__ == "__main__":
    main()
