"""The committed golden vectors (tests/golden/golden.npz, written by
tests/golden/make_golden.py from the CPU oracle) against the oracle as it is
now, bit for bit, plus the parts of them that closed forms fix without the
oracle: the colex index (src/disorder.cu:159-171), comb frequencies
(src/mc.cu:449-483), permutations and counter sums of the exchange traces.
"""
import math
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

GOLDEN = os.path.join(HERE, "golden", "golden.npz")


@pytest.fixture(scope="module")
def golden():
    with np.load(GOLDEN, allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


def test_oracle_reproduces_golden(oracle, golden):
    import make_golden

    now = make_golden.generate()
    assert sorted(now) == sorted(golden)
    for k, v in golden.items():
        w = now[k]
        assert w.dtype == v.dtype and w.shape == v.shape, k
        # bit-for-bit, FP64 included
        assert v.tobytes() == w.tobytes(), k


def test_golden_colex_closed_form(golden):
    for (i, j, k, l), q in zip(golden["colex_quads"], golden["colex_q"]):
        assert q == math.comb(int(l), 4) + math.comb(int(k), 3) + math.comb(int(j), 2) + int(i)


def test_golden_comb_and_mask_structure(golden):
    import make_golden

    for N, fmc, gamma, seed in make_golden.MASK_CASES:
        w = golden[make_golden.case_key("omega", N, fmc, seed)]
        m = golden[make_golden.case_key("mask", N, fmc, gamma, seed)]
        assert len(m) == math.comb(N, 4) and set(np.unique(m)) <= {0, 1, 2, 4}
        if fmc == 1:
            assert np.array_equal(w, np.arange(N, dtype=float))
            if gamma == 0.0:  # only i+l = j+k survives, in channel 1
                assert set(np.unique(m)) <= {0, 2}
        else:
            assert np.all(np.diff(w) >= 0) and w.min() >= 0 and w.max() < 1


def test_golden_exchange_traces_are_permutations(golden):
    import make_golden

    for NT, nrep, rounds, _, _, seed48 in make_golden.EXCH_CASES:
        k = make_golden.case_key("exch", NT, nrep, seed48)
        perm, betas, beta = golden[k + "_perm"], golden[k + "_betas"], golden[k + "_beta"]
        assert golden[k + "_prop"].tolist() == [rounds * nrep] * (NT - 1)
        assert np.all(golden[k + "_acc"] <= golden[k + "_prop"])
        for r in range(rounds):
            assert sorted(perm[r].tolist()) == list(range(NT * nrep))
            for s in range(NT * nrep):
                assert betas[r][perm[r][s]] == beta[s // nrep]
