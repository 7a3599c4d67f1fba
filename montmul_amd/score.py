"""The score: how many Montgomery squarings per second a backend sustains on a seeded batch of
moduli. Every repeat re-arms all slots and times one run; the median rate and its median absolute
deviation are the result, and none is given unless sampled slots equal 3^e mod N
(reference: montmul/src/score.rs).
"""

import random
import statistics
import time

from . import Hip, ModBatch, MontmulError, score_workload

# Median seconds per modular squaring of the reference's Gmp backend, one slot on one M5
# performance core at 100,000 bits (score.rs:11-13).
GMP_M5_SECONDS_PER_MODSQ_100K = 2.709e-4


class ScoreConfig:
    """score.rs:15-30: 100,000-bit moduli (L = 8,192), 2,000-bit exponents, 5 repeats, seed 2026."""

    def __init__(self, bits=100_000, slots=1, ebits=2_000, repeats=5, seed=2026):
        self.bits, self.slots, self.ebits, self.repeats, self.seed = bits, slots, ebits, repeats, seed

    @staticmethod
    def standard(slots):
        return ScoreConfig(slots=slots)


class Score:
    """score.rs:32-43."""

    def __init__(self, slots, modsq_per_s, mad, bits, overhead=None):
        self.slots = slots
        self.modsq_per_s = modsq_per_s
        self.mad = mad
        self.fermat_tests_per_hour = modsq_per_s * 3600.0 / bits
        self.overhead = overhead

    def __repr__(self):
        return f"Score(slots={self.slots}, modsq_per_s={self.modsq_per_s:.1f}, mad={self.mad:.1f})"


def workload(cfg):
    """score.rs:45-61: GMP's generator seeded `seed`, top bit and bit 0 set; the top `ebits` bits
    of N - 1 as exponents."""
    if cfg.slots <= 0 or cfg.repeats <= 0:
        raise MontmulError("slots and repeats must be positive")
    if not 0 < cfg.ebits <= cfg.bits:
        raise MontmulError("ebits must be in 1..=bits")
    return score_workload(cfg.bits, cfg.slots, cfg.ebits, cfg.seed)


def arm(batch, exps):
    """score.rs:63-69: loads 1 and the exponent into every slot."""
    for i, e in enumerate(exps):
        batch.load(i, 1)
        batch.set_exponent(i, e)


def median(v):
    return statistics.median(v)


def mad(v):
    m = median(v)
    return median([abs(x - m) for x in v])


def accept(old, new):
    """score.rs:86-89: `new` is kept only if its median beats `old` by at least 2 % and by more
    than twice the larger MAD."""
    gain = new.modsq_per_s - old.modsq_per_s
    return gain >= 0.02 * old.modsq_per_s and gain > 2.0 * max(old.mad, new.mad)


def score(make, cfg):
    """score.rs:91-123: one warm-up run(64), then `repeats` timed run(ebits) on freshly loaded
    slots; fails if slot 0, the last slot or two seeded random slots disagree with pow_mod."""
    moduli, exps = workload(cfg)
    batch = ModBatch(make(), moduli, 3)
    arm(batch, exps)
    batch.run(64)
    rates = []
    for _ in range(cfg.repeats):
        arm(batch, exps)
        t = time.perf_counter()
        batch.run(cfg.ebits)
        rates.append(cfg.slots * cfg.ebits / (time.perf_counter() - t))
    rng = random.Random(cfg.seed + 1)
    for i in [0, cfg.slots - 1] + [rng.randrange(cfg.slots) for _ in range(2)]:
        if batch.remaining(i) != 0 or batch.read(i) != pow(3, exps[i], moduli[i]):
            raise MontmulError(f"slot {i} disagrees with pow_mod")
    batch.free()
    return Score(cfg.slots, median(rates), mad(rates), cfg.bits)


def best_slots(make, bits, candidates):
    """score.rs:125-136: the candidate with the highest score at ebits = 200 and one repeat."""
    if not candidates:
        raise MontmulError("no candidate slot counts")
    best = (candidates[0], 0.0)
    for s in candidates:
        r = score(make, ScoreConfig(bits=bits, slots=s, ebits=200, repeats=1))
        if r.modsq_per_s > best[1]:
            best = (s, r.modsq_per_s)
    return best[0]


def gpu_overhead(make, cfg):
    """score.rs:138-160: median time of run over that of run_bare on the score workload (one
    untimed pass of each first), i.e. the whole step over the squaring alone."""
    moduli, exps = workload(cfg)
    batch = ModBatch(make(), moduli, 3)

    def timed(bare):
        arm(batch, exps)
        t = time.perf_counter()
        if bare:
            Hip.run_bare(batch, cfg.ebits)
        else:
            batch.run(cfg.ebits)
        return time.perf_counter() - t

    timed(False)
    timed(True)
    full = [timed(False) for _ in range(cfg.repeats)]
    bare = [timed(True) for _ in range(cfg.repeats)]
    batch.free()
    return median(full) / median(bare)
