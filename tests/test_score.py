"""The score front end on CPU (montmul/src/score.rs and montmul/tests/score.rs): median and MAD,
the acceptance rule, a checked score on the oracle's Cpu backend, a backend whose reads are
corrupted in one slot (no score), and a backend whose runs sleep on a schedule (the median of
re-armed repeats)."""

import ctypes
import time

import pytest

import montmul_amd as mm
from montmul_amd import MontmulError, score
from oracle_backends import Cpu


def test_median_and_mad():
    """score.rs:168-176."""
    assert score.median([5.0, 1.0, 3.0]) == 3.0
    assert score.median([4.0, 1.0, 3.0, 10.0]) == 3.5
    # deviations from 4: 3, 1, 1, 0, 96 -> median 1 (the mean deviation would be 20.2)
    assert score.mad([1.0, 5.0, 3.0, 4.0, 100.0]) == 1.0
    # deviations from 3.5: 2.5, 0.5, 0.5, 6.5 -> median 1.5
    assert score.mad([1.0, 4.0, 3.0, 10.0]) == 1.5


def _at(rate, mad):
    return score.Score(1, rate, mad, 1000)


def test_acceptance_rule():
    """score.rs:178-185: at least 2 % faster and by more than twice the larger MAD."""
    assert score.accept(_at(100.0, 0.5), _at(103.0, 0.5))
    assert not score.accept(_at(100.0, 0.1), _at(101.9, 0.1)), "below 2 %"
    assert not score.accept(_at(100.0, 2.0), _at(103.0, 0.5)), "within 2 MAD_old"
    assert not score.accept(_at(100.0, 0.5), _at(103.0, 1.5)), "within 2 MAD_new"
    assert not score.accept(_at(100.0, 0.5), _at(97.0, 0.5)), "slower"


class _Wrapped(mm.Backend):
    """The oracle's Cpu backend with its run and/or read replaced by Python functions that get
    the inner table (a BackendTable), the state pointer and the call's arguments."""

    def __init__(self, run=None, read=None):
        self.inner = Cpu()
        t = mm.BackendTable()
        for name, _ in mm.BackendTable._fields_:
            setattr(t, name, getattr(self.inner.table, name))
        self._cbs = []
        if run:
            cb = mm._RUN(lambda st, bits: run(self.inner.table, ctypes.c_void_p(st), bits))
            self._cbs.append(cb)
            t.run = cb
        if read:
            cb = mm._READ(lambda st, i, out: read(self.inner.table, ctypes.c_void_p(st), i, out))
            self._cbs.append(cb)
            t.read = cb
        self.table = t


def test_score_is_positive_and_checked():
    """tests/score.rs:54-60 on the oracle's Cpu."""
    cfg = score.ScoreConfig(bits=1000, slots=4, ebits=100, repeats=3, seed=1)
    s = score.score(Cpu, cfg)
    assert s.modsq_per_s > 0 and s.mad >= 0
    assert abs(s.fermat_tests_per_hour - s.modsq_per_s * 3600 / 1000) < 1e-6 * s.modsq_per_s


def test_wrong_backend_gets_no_score():
    """tests/score.rs:62-68: reads of the last slot have bit 0 of limb 0 flipped; the check
    against pow_mod must reject the score."""

    def read(inner, st, i, out):
        r = inner.read(st, i, out)
        if i == 3:
            out[0] ^= 1
        return r

    cfg = score.ScoreConfig(bits=1000, slots=4, ebits=100, repeats=1, seed=1)
    with pytest.raises(MontmulError, match="slot 3 disagrees with pow_mod"):
        score.score(lambda: _Wrapped(read=read), cfg)


def test_score_is_median_rate_over_reloaded_repeats():
    """tests/score.rs:107-121: a warm-up, then timed repeats that sleep 200, 200 and 20 ms. The
    median rate is slots ebits / 200 ms with a MAD near 0 (the mean rate would be about 4 times
    higher), and there is one run with work per repeat plus the warm-up: each repeat re-arms."""
    schedule, calls = [0, 200, 200, 20], [0]

    def run(inner, st, bits):
        if inner.remaining(st, 0) > 0:
            n = calls[0]
            calls[0] += 1
            time.sleep(schedule[min(n, len(schedule) - 1)] / 1000)
        return inner.run(st, bits)

    cfg = score.ScoreConfig(bits=200, slots=4, ebits=100, repeats=3, seed=1)
    s = score.score(lambda: _Wrapped(run=run), cfg)
    want = cfg.slots * cfg.ebits / 0.200
    assert abs(s.modsq_per_s - want) < 0.3 * want, (s.modsq_per_s, want)
    assert s.mad < 0.1 * want
    assert calls[0] == cfg.repeats + 1
