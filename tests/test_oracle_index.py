"""Pins the oracle's integer / index functions (CPU only).

- splitmix64: published test vector of Steele et al.'s SplitMix64 (seed 0).
- Philox4x32-10: Random123 known-answer vectors.
- colex encode/decode: exhaustive bijection (reference test src/test_quartet_index.cu:28-137).
- 3-type enumeration: set equality with brute force (src/test_enum.cpp:15-145) and
  with the reference's own three_type() compiled from src/test_enum.cpp in oracle/_ref.
"""
import itertools

import numpy as np
import pytest


def test_splitmix64_vector(oracle):
    got = [int(x) for x in oracle.splitmix64_seq(0, 4)]
    assert got == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F, 0xF88BB8A8724C81EC]


@pytest.mark.parametrize("ctr,key,want", [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
])
def test_philox_kat(oracle, ctr, key, want):
    assert [int(x) for x in oracle.philox(ctr, key)] == want


@pytest.mark.parametrize("N", range(4, 21))
def test_colex_bijection(oracle, N):
    # colex order: enumerate by l, then k, j, i ascending
    q = 0
    for l in range(3, N):
        for k in range(2, l):
            for j in range(1, k):
                for i in range(j):
                    assert oracle.encode(i, j, k, l) == q
                    assert oracle.decode(q) == (i, j, k, l)
                    q += 1
    assert q == oracle.n_quartets(N)


def _brute(N, i0, j0):
    return {q for q in itertools.combinations(range(N), 4) if i0 in q or j0 in q}


@pytest.mark.parametrize("N", [4, 5, 6, 7, 9, 12])
def test_three_type_vs_brute(oracle, N):
    for i0 in range(N):
        for j0 in range(i0 + 1, N):
            got = oracle.three_type(N, i0, j0)
            s = {tuple(int(x) for x in r) for r in got}
            assert len(s) == len(got), "duplicate quartet in 3-type enumeration"
            assert s == _brute(N, i0, j0)
            assert len(got) == oracle.n_inv(N)


@pytest.mark.parametrize("N,pairs", [(10, [(0, 1), (3, 7), (8, 9)]), (16, [(0, 15), (5, 6)]),
                                     (20, [(2, 11)])])
def test_three_type_vs_reference_build(oracle, N, pairs):
    if oracle.ref_lib() is None:
        pytest.skip("reference sources not mounted; oracle/_ref not built")
    for i0, j0 in pairs:
        mine = {tuple(int(x) for x in r) for r in oracle.three_type(N, i0, j0)}
        ref = {tuple(int(x) for x in r) for r in oracle.ref_three_type(N, i0, j0)}
        assert mine == ref


def test_reference_enum_program_passes(oracle):
    if oracle.ref_lib() is None:
        pytest.skip("reference sources not mounted; oracle/_ref not built")
    assert oracle.ref_lib().ref_test_enum_run() == 0


def test_three_type_order_matches_reference_t_index(oracle):
    # The reference walks t = 0..n_h4-1 with type 1 first; our restatement
    # yields the same order: first C(N-2,3) rows contain i0, next contain j0.
    N, i0, j0 = 11, 2, 6
    got = oracle.three_type(N, i0, j0)
    n12 = (N - 2) * (N - 3) * (N - 4) // 6
    assert all(i0 in r and j0 not in r for r in got[:n12])
    assert all(j0 in r and i0 not in r for r in got[n12:2 * n12])
    assert all(i0 in r and j0 in r for r in got[2 * n12:])


def test_sweep_draws_do_not_reuse_spin_init_blocks(oracle):
    """Spin initialisation draws Philox blocks (pair, 0, 0, 1) under the
    replica key; the sweep's per-step blocks must never coincide with them."""
    seed = 0x243F6A8885A308D3
    key = [seed & 0xFFFFFFFF, seed >> 32]
    for s in range(16):
        init = np.array([(int(x) + 0.5) / 2**32 for x in oracle.philox([s, 0, 0, 1], key)])
        u = oracle.step_uniforms(seed, s, 0)
        assert not np.allclose(u[:4], init)
        assert not np.allclose(u[4:6], init[:2])
