"""Mixed-length batch benchmark: moduli of many sizes stepped by one persistent launch against
the same moduli run as plain per-length batches, one length after another.

    python -m montmul_amd.mixed [--count 1000] [--seed 11] [--min-bits 2048] [--max-bits 200000] \
        [--ebits 4000] [--check 16] [--parts]

Moduli: `count` random odd numbers whose bit sizes are log-uniform in [min-bits, max-bits]
(Python `random` seeded `seed`), each at the transform length of its own size (L = 256 ..
16,384 for the defaults). Exponent: the top min(ebits, bits) bits of N - 1. Base 3, every slot
loaded with 1.

Both forms run after one untimed pass; the timed region is one run(ebits) of the mixed batch,
and one run(ebits) of every per-length ModBatch in turn. Checks: every slot's raw limbs equal in
both forms, and read == pow(3, e, N) (GMP) on `check` slots spread over the lengths.

Prints one JSON line: squarings/s of both forms (squarings = sum of exponent bits), the mixed
launch's kernel milliseconds (HIP events), slots per length, and the check counts.

`--parts` also times the mixed launch on parts of the workload, to show what bounds it: the
longest length's slots (plus one shortest-length slot of 16 steps, so that they run in the mixed
kernel rather than their own), every other slot alone, and every other slot plus one slot of the
longest length with a 16-bit exponent (which puts the others in that length's workgroup size
while adding almost no work). Their kernel milliseconds go into the line as "parts_ms".
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

import montmul_amd as mm


def workload(count, seed, min_bits, max_bits, ebits):
    rng = random.Random(seed)
    moduli, exps = [], []
    for _ in range(count):
        bits = int(math.exp(rng.uniform(math.log(min_bits), math.log(max_bits))))
        n = rng.getrandbits(bits) | (1 << (bits - 1)) | 1
        moduli.append(n)
        exps.append((n - 1) >> max(0, bits - ebits))
    return moduli, exps


def arm(batch, exps):
    for i, e in enumerate(exps):
        batch.load(i, 1)
        batch.set_exponent(i, e)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--count", type=int, default=1000)
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--min-bits", type=int, default=2048)
    ap.add_argument("--max-bits", type=int, default=200000)
    ap.add_argument("--ebits", type=int, default=4000)
    ap.add_argument("--check", type=int, default=16)
    ap.add_argument("--device", type=int, default=0)
    ap.add_argument("--parts", action="store_true", help="also time the longest length and the rest apart")
    args = ap.parse_args()
    moduli, exps = workload(args.count, args.seed, args.min_bits, args.max_bits, args.ebits)
    squarings = sum(e.bit_length() for e in exps)
    hip = mm.Hip(device=args.device)
    t0 = time.perf_counter()
    mix = mm.MixBatch(hip, moduli, 3)
    setup_s = time.perf_counter() - t0
    by_len = {}
    for i in range(len(moduli)):
        by_len.setdefault(mix.params(i).len, []).append(i)
    print(f"{len(moduli)} moduli, {squarings} squarings, lengths "
          + ", ".join(f"{l}: {len(v)}" for l, v in sorted(by_len.items())) + f"; setup {setup_s:.2f} s",
          file=sys.stderr, flush=True)

    arm(mix, exps)
    mix.run(args.ebits)  # untimed pass
    arm(mix, exps)
    t = time.perf_counter()
    mix.run(args.ebits)
    mixed_s = time.perf_counter() - t
    kernel_ms = mm.Hip.last_run_kernel_ms(_State(mix.group_state(0)))[0]
    print(f"mixed: {mixed_s:.3f} s, kernel {kernel_ms:.1f} ms", file=sys.stderr, flush=True)

    plain = {}
    for length, idx in sorted(by_len.items()):
        b = mm.ModBatch(hip, [moduli[i] for i in idx], 3)
        arm(b, [exps[i] for i in idx])
        b.run(args.ebits)
        plain[length] = b
    plain_s, plain_kernel_ms, per_len = 0.0, 0.0, {}
    for length, idx in sorted(by_len.items()):
        b = plain[length]
        arm(b, [exps[i] for i in idx])
        t = time.perf_counter()
        b.run(args.ebits)
        dt = time.perf_counter() - t
        plain_s += dt
        plain_kernel_ms += mm.Hip.last_run_kernel_ms(b)[0]
        per_len[length] = round(dt * 1000, 1)
    print(f"plain per-length: {plain_s:.3f} s", file=sys.stderr, flush=True)

    raw_mismatch = 0
    for length, idx in sorted(by_len.items()):
        for j, i in enumerate(idx):
            if plain[length].read_raw(j) != mix.read_raw(i) or mix.remaining(i) != 0:
                raw_mismatch += 1
    rng = random.Random(args.seed + 1)
    sample = set()
    for idx in by_len.values():
        sample.add(idx[0])
    while len(sample) < min(args.check, len(moduli)):
        sample.add(rng.randrange(len(moduli)))
    powm_mismatch = sum(0 if mix.check_pow(i, 3, exps[i]) else 1 for i in sorted(sample))
    line = {
        "metric": "mixed-length squarings/s: one persistent launch vs per-length batches",
        "workload": f"{len(moduli)} moduli, bits log-uniform in [{args.min_bits}, {args.max_bits}] (seed {args.seed}), "
                    f"base 3, e = top min({args.ebits}, bits) bits of N-1",
        "squarings": squarings,
        "mixed_s": round(mixed_s, 4), "mixed_kernel_ms": round(kernel_ms, 2),
        "mixed_sq_per_s": round(squarings / mixed_s, 1),
        "plain_s": round(plain_s, 4), "plain_kernel_ms": round(plain_kernel_ms, 2),
        "plain_sq_per_s": round(squarings / plain_s, 1),
        "speedup": round(plain_s / mixed_s, 3),
        "plain_ms_per_length": per_len,
        "slots_per_length": {l: len(v) for l, v in sorted(by_len.items())},
        "setup_s": round(setup_s, 2),
        "raw_mismatch": raw_mismatch, "powm_checked": len(sample), "powm_mismatch": powm_mismatch,
        "device": hip.device_name(),
    }
    if args.parts:
        line["parts_ms"] = parts(hip, moduli, exps, by_len, args.ebits)
    print(json.dumps(line), flush=True)
    for b in plain.values():
        b.free()
    mix.free()
    hip.close()
    sys.exit(0 if raw_mismatch == 0 and powm_mismatch == 0 else 1)


def parts(hip, moduli, exps, by_len, ebits):
    """Kernel milliseconds of the mixed launch on parts of the workload (one untimed pass each)."""
    longest, shortest = max(by_len), min(by_len)
    big = by_len[longest]
    rest = [i for length, idx in sorted(by_len.items()) if length != longest for i in idx]
    one, tiny = big[0], by_len[shortest][0]

    def short(i):  # the top 16 bits of slot i's exponent: 16 steps
        return exps[i] >> max(0, exps[i].bit_length() - 16)

    # One length alone would run its own kernel, so the longest length gets one slot of the
    # shortest beside it, which keeps it in the mixed launch at the cost of 16 short steps.
    cases = {
        f"L={longest} ({len(big)} slots) + one L={shortest} slot of 16 steps": (
            [moduli[i] for i in big] + [moduli[tiny]], [exps[i] for i in big] + [short(tiny)]),
        f"the rest alone ({len(rest)} slots)": ([moduli[i] for i in rest], [exps[i] for i in rest]),
        f"the rest + one L={longest} slot of 16 steps": (
            [moduli[i] for i in rest] + [moduli[one]], [exps[i] for i in rest] + [short(one)]),
    }
    out = {}
    for label, (ms, es) in cases.items():
        mix = mm.MixBatch(hip, ms, 3)
        arm(mix, es)
        mix.run(ebits)
        arm(mix, es)
        mix.run(ebits)
        out[label] = round(mm.Hip.last_run_kernel_ms(_State(mix.group_state(0)))[0], 1)
        print(f"part: {label}: kernel {out[label]} ms", file=sys.stderr, flush=True)
        mix.free()
    return out


class _State:
    """A stand-in with the `state` attribute Hip.last_run_kernel_ms reads."""

    def __init__(self, state):
        self.state = state


if __name__ == "__main__":
    main()
