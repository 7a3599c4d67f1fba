"""Command-line front end (reference: montmul/src/main.rs):

    python -m montmul_amd bench --bits B --slots S --steps K [--bare] [--global | --lds] [--csv PATH]
    python -m montmul_amd prp --index N
    python -m montmul_amd score [--slots S] [--label TEXT] [--rocprof] [--csv PATH]
    python -m montmul_amd fermat --bits B --slots S [--seed N] [--planted FILE] [--planted-slots P]
                                 [--check K] [--global] [--csv PATH]

Options are `--key value` pairs and bare `--flag`s; unknown or repeated keys are errors
(main.rs:21-66). Results print as one `key=value` line, appended to a CSV file (header first)
with --csv. The backend is the MI355X (`--backend gpu`, the default and only value here: the
reference's cpu and gmp backends are test oracles of this repository, not part of the library).
"""

import csv
import os
import random
import shutil
import subprocess
import sys
import tempfile
import time

import montmul_amd as mm
from montmul_amd import score as sc

ALLOWED = {  # main.rs:21-30, plus fermat (the full-test workload) and --global
    "bench": (["backend", "bits", "slots", "steps", "csv"], ["bare", "global", "lds"]),
    "prp": (["index", "backend"], []),
    "score": (["backend", "slots", "label", "csv"], ["rocprof"]),
    "fermat": (["backend", "bits", "slots", "seed", "planted", "planted-slots", "check", "csv"], ["global", "lds"]),
    "rocprof-launch": (["slots"], []),
}


class Args:
    """main.rs:32-66: rejects unknown or repeated keys and a value that is itself an option."""

    def __init__(self, cmd, argv):
        keys, flag_keys = ALLOWED.get(cmd, ([], []))
        self.opts, self.flags = {}, []
        it = iter(argv)
        for a in it:
            if not a.startswith("--"):
                raise SystemExit(f"unexpected argument {a}")
            key = a[2:]
            if key in self.opts or key in self.flags:
                raise SystemExit(f"--{key} given twice")
            if key in flag_keys:
                self.flags.append(key)
            elif key in keys:
                v = next(it, None)
                if v is None or v.startswith("--"):
                    raise SystemExit(f"--{key} needs a value")
                self.opts[key] = v
            else:
                raise SystemExit(f"{cmd} does not take --{key}")

    def get(self, key, default=None):
        return self.opts.get(key, default)

    def num(self, key, default=None):
        v = self.opts.get(key)
        if v is None:
            if default is None:
                raise SystemExit(f"missing --{key}")
            return default
        try:
            return int(v)
        except ValueError:
            raise SystemExit(f"--{key}: cannot parse {v}")

    def flag(self, key):
        return key in self.flags


def backend_of(args):
    b = args.get("backend", "gpu")
    if b != "gpu":
        raise SystemExit(f"unknown or unavailable backend {b} (this library's backend is gpu)")
    return b


def device_name(hip):
    """main.rs:84-97: whitespace to '_' so that it fits a key=value field."""
    return "_".join(hip.device_name().split()).replace(",", "")


def emit(fields, path):
    """Prints `key=value` fields and appends them to a CSV file (main.rs:111-122)."""
    print(" ".join(f"{k}={v}" for k, v in fields), flush=True)
    if path:
        d = os.path.dirname(path)
        if d:
            os.makedirs(d, exist_ok=True)
        fresh = not os.path.exists(path) or os.path.getsize(path) == 0
        with open(path, "a", newline="") as f:
            w = csv.writer(f)
            if fresh:
                w.writerow([k for k, _ in fields])
            w.writerow([v for _, v in fields])


def cmd_bench(args):
    """main.rs:124-180: random odd moduli and `steps`-bit exponents, one warm-up run(16), then one
    timed run(steps) (run_bare with --bare)."""
    backend_of(args)
    bits, slots, steps = args.num("bits"), args.num("slots"), args.num("steps")
    if bits < 2 or slots <= 0 or steps <= 0 or steps > bits:
        raise SystemExit("bits >= 2, slots > 0 and 0 < steps <= bits required")
    hip = mm.Hip(force_global=args.flag("global"), prefer_lds=args.flag("lds"))
    moduli, exps = mm.score_workload(bits, slots, steps, 0)
    b = mm.ModBatch(hip, moduli, 3)
    run = (lambda k: mm.Hip.run_bare(b, k)) if args.flag("bare") else b.run
    sc.arm(b, exps)
    run(16)
    sc.arm(b, exps)
    t = time.perf_counter()
    run(steps)
    secs = time.perf_counter() - t
    emit([("backend", "gpu"), ("device", device_name(hip)), ("bits", bits), ("L", b.params().len), ("slots", slots),
          ("steps", steps), ("bare", str(args.flag("bare")).lower()), ("mode", "global" if mm.Hip.plan_of(b)["global"] else "lds"),
          ("us_per_step_slot", f"{secs * 1e6 / (steps * slots):.3f}"),
          ("bitsteps_per_s", f"{bits * slots * steps / secs:.4e}")], args.get("csv"))


def tribonacci(n):
    a, b, c = 0, 0, 1
    for _ in range(n):
        a, b, c = b, c, a + b + c
    return a


def cmd_prp(args):
    """main.rs:182-189: Fermat base 3 on T(index)."""
    backend_of(args)
    i = args.num("index")
    t = tribonacci(i)
    start = time.perf_counter()
    got = mm.fermat(mm.Hip(), [t], 3, 1 << 16)
    print(f"T({i}) bits={t.bit_length()} prp={str(got[0]).lower()} seconds={time.perf_counter() - start:.3f}")


def candidates(hip):
    """main.rs:200-216: CUs x {1, 2, 4, 8, 16}, keeping the batches that fit in device memory."""
    moduli, _ = mm.score_workload(100_000, 1, 2000, 2026)
    out = []
    for f in (1, 2, 4, 8, 16):
        s = f * hip.cu_count()
        try:
            mm.ModBatch(hip, moduli * s, 3).free()
            out.append(s)
        except mm.MontmulError:
            pass
    return out


def rocprof_metrics(slots):
    """The --ncu analogue (main.rs:218-250): one 16-step launch of the score workload under
    rocprofv3, giving the step kernel's VALU issue (SQ_INSTS_VALU over the CUs' cycles,
    GRBM_GUI_ACTIVE summed over 8 XCDs, in percent of one wave instruction per CU per cycle; it
    can pass 100, as gfx950 issues simple 32-bit operations at up to about 1.37 per cycle) and
    its HBM traffic rate (FETCH_SIZE x 2, the scale of a plain copy measured on this GPU, plus
    WRITE_SIZE, over 8 TB/s)."""
    prof = shutil.which("rocprofv3")
    if not prof:
        raise RuntimeError("rocprofv3 not found")
    me = [sys.executable, "-m", "montmul_amd", "rocprof-launch", "--slots", str(slots)]
    vals = {}
    with tempfile.TemporaryDirectory() as d:
        for i, counters in enumerate((["SQ_INSTS_VALU", "GRBM_GUI_ACTIVE"], ["FETCH_SIZE"], ["WRITE_SIZE"])):
            out = os.path.join(d, str(i))
            subprocess.run([prof, "--pmc", *counters, "--output-format", "csv", "-d", out, "-o", "p", "--", *me],
                           check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, timeout=600)
            for root, _, files in os.walk(out):
                for fn in files:
                    if fn.endswith("counter_collection.csv"):
                        for r in csv.DictReader(open(os.path.join(root, fn))):
                            if "k_steps" in r["Kernel_Name"]:
                                vals[r["Counter_Name"]] = float(r["Counter_Value"])
                                vals["ns"] = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
    cus = mm.Hip().cu_count()
    cycles = vals["GRBM_GUI_ACTIVE"] / 8  # per XCD
    valu_pct = 100.0 * vals["SQ_INSTS_VALU"] / (cus * cycles)
    hbm = (2.0 * vals["FETCH_SIZE"] + vals["WRITE_SIZE"]) * 1024
    mem_pct = 100.0 * hbm / (vals["ns"] * 1e-9) / 8.0e12
    return valu_pct, mem_pct


def cmd_score(args):
    """main.rs:252-299: the standard score at --slots or at the best candidate, the run/run_bare
    overhead, optional counters, and the GMP core equivalent."""
    backend_of(args)
    hip = mm.Hip()
    slots = args.num("slots", 0) or sc.best_slots(lambda: hip, 100_000, candidates(hip))
    s = sc.score(lambda: hip, sc.ScoreConfig.standard(slots))
    s.overhead = sc.gpu_overhead(lambda: hip, sc.ScoreConfig.standard(slots))
    valu_pct = mem_pct = None
    if args.flag("rocprof"):
        try:
            valu_pct, mem_pct = rocprof_metrics(slots)
        except Exception as e:  # a failed profile leaves the columns empty (main.rs:271-276)
            print(f"warning: rocprof profile failed, valu_pct and mem_pct left empty: {e}", file=sys.stderr)
    opt = lambda v: "" if v is None else f"{v:.4f}"
    label = args.get("label", "").replace('"', '""')
    emit([("timestamp", int(time.time())), ("label", f'"{label}"'), ("device", device_name(hip)), ("backend", "gpu"),
          ("slots", s.slots), ("modsq_per_s", f"{s.modsq_per_s:.2f}"), ("mad", f"{s.mad:.2f}"),
          ("fermat_tests_per_hour", f"{s.fermat_tests_per_hour:.4f}"), ("overhead", opt(s.overhead)),
          ("valu_pct", opt(valu_pct)), ("mem_pct", opt(mem_pct)),
          ("gmp_core_equiv", opt(s.modsq_per_s * sc.GMP_M5_SECONDS_PER_MODSQ_100K))], args.get("csv"))


def cmd_rocprof_launch(args):
    """Internal (main.rs:334-343): one 16-step launch of the score workload, for rocprof."""
    cfg = sc.ScoreConfig.standard(args.num("slots"))
    moduli, exps = sc.workload(cfg)
    b = mm.ModBatch(mm.Hip(), moduli, 3)
    sc.arm(b, exps)
    b.run(16)


def cmd_fermat(args):
    """Full Fermat tests of a batch (e = N - 1, run until no bits remain): random odd moduli from
    GMP's generator seeded --seed, the first --planted-slots slots cycling through the known primes
    of --planted (one hexadecimal prime per line). Planted primes must come out prime; every slot
    reported prime and --check seeded random others are checked against GMP powm."""
    backend_of(args)
    bits, slots = args.num("bits"), args.num("slots")
    seed, check = args.num("seed", 7), args.num("check", 8)
    primes = []
    if args.get("planted"):
        primes = [int(l, 16) for l in open(args.get("planted")) if l.strip()]
        primes = [p for p in primes if p.bit_length() == bits]
    planted = min(args.num("planted-slots", len(primes) and slots // 2), slots) if primes else 0
    rand, _ = mm.score_workload(bits, slots - planted, bits, seed) if slots > planted else ([], [])
    moduli = [primes[i % len(primes)] for i in range(planted)] + rand
    hip = mm.Hip(force_global=args.flag("global"), prefer_lds=args.flag("lds"))
    b = mm.ModBatch(hip, moduli, 3)
    for i, n in enumerate(moduli):
        b.load(i, 1)
        b.set_exponent(i, n - 1)
    t = time.perf_counter()
    while any(b.remaining(i) for i in range(slots)):
        b.run(1 << 16)
    secs = time.perf_counter() - t
    prp = [b.is_one(i) for i in range(slots)]
    bad_planted = sum(1 for i in range(planted) if not prp[i])
    rng = random.Random(seed)
    others = [i for i in range(planted, slots)]
    to_check = sorted(set(i for i in range(slots) if prp[i] and i >= planted) |
                      set(rng.sample(others, min(check, len(others)))))
    wrong = [i for i in to_check if not b.check_pow(i, 3, moduli[i] - 1)]
    squarings = sum(n.bit_length() for n in moduli)  # N - 1 has the bits of N for N > 2
    emit([("backend", "gpu"), ("device", device_name(hip)), ("bits", bits), ("L", b.params().len), ("slots", slots),
          ("mode", "global" if mm.Hip.plan_of(b)["global"] else "lds"), ("planted", planted),
          ("distinct_planted", len(primes)), ("planted_not_prime", bad_planted),
          ("random_prp", sum(prp[planted:])), ("checked_powm", len(to_check)), ("powm_mismatch", len(wrong)),
          ("seconds", f"{secs:.3f}"), ("modsq_per_s", f"{squarings / secs:.1f}")], args.get("csv"))
    if bad_planted or wrong:
        raise SystemExit(1)


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if not argv:
        raise SystemExit("usage: python -m montmul_amd bench|prp|score|fermat [options]")
    cmd, rest = argv[0], argv[1:]
    if cmd not in ALLOWED:
        raise SystemExit(f"unknown command {cmd}")
    args = Args(cmd, rest)
    {"bench": cmd_bench, "prp": cmd_prp, "score": cmd_score, "fermat": cmd_fermat,
     "rocprof-launch": cmd_rocprof_launch}[cmd](args)


if __name__ == "__main__":
    main()
