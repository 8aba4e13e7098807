"""The HIP path (libpspin.so) against the committed golden vectors
(tests/golden/golden.npz): replica seeds, FMC frequencies, g4 masks and the
g2 survivor pattern, sparse quartet selection and PT exchange traces, all bit
for bit.  The same vectors are checked against the oracle on the CPU
(tests/test_golden.py).
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden  # noqa: E402  (case tables only; nothing is generated here)

GOLDEN = os.path.join(HERE, "golden", "golden.npz")


@pytest.fixture(scope="module")
def golden():
    with np.load(GOLDEN, allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


def test_replica_seeds(pspin, golden):
    for seed, n in ((42, 8), (77777, 5), (2024, 3)):
        st = pspin.PSpin(pspin.SimConfig.default(N=8, nrep=n, seed=seed), pspin.DENSE)
        seeds, _ = st.rng_info()
        st.close()
        assert np.array_equal(seeds, golden[make_golden.case_key("seeds", seed, n)])


@pytest.mark.parametrize("case", make_golden.MASK_CASES)
def test_disorder_masks(pspin, golden, case):
    N, fmc, gamma, seed = case
    print(f"[golden masks] N={N} fmc={fmc} gamma={gamma} seed={seed}", flush=True)
    st = pspin.PSpin(pspin.SimConfig.default(N=N, nrep=2, seed=seed, fmc_mode=fmc, gamma=gamma), pspin.DENSE)
    w = st.omega()
    _, mask = st.g4()
    g2 = st.g2()
    st.close()
    assert w.tobytes() == golden[make_golden.case_key("omega", N, fmc, seed)].tobytes()
    assert np.array_equal(mask, golden[make_golden.case_key("mask", N, fmc, gamma, seed)])
    if fmc == 2:
        assert np.array_equal(g2 != 0, golden[make_golden.case_key("g2keep", N, gamma, seed)].astype(bool))


@pytest.mark.parametrize("case", make_golden.SPARSE_CASES)
def test_sparse_selection(pspin, golden, case):
    N, fmc, gamma, seed, J4, J40 = case
    print(f"[golden sparse] N={N} fmc={fmc} seed={seed}", flush=True)
    cfg = pspin.SimConfig.default(N=N, nrep=2, seed=seed, fmc_mode=fmc, gamma=gamma, alpha0=0.5, J0=0.0)
    assert cfg.alpha * cfg.J == J4 and cfg.alpha0 * cfg.J0 == J40
    st = pspin.PSpin(cfg, pspin.SPARSE)
    q, g = st.sparse_quartets()
    nnz = st.info.n_quart_active
    st.close()
    assert np.array_equal(q, golden[make_golden.case_key("sparse_q", N, fmc, seed)])
    assert nnz == int(golden[make_golden.case_key("sparse_nnz", N, fmc, seed)][0])
    # FP64 couplings: device Box-Muller (log, sincos) vs glibc, within 1e-12
    np.testing.assert_allclose(g, golden[make_golden.case_key("sparse_g", N, fmc, seed)], rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("case", make_golden.EXCH_CASES)
def test_exchange_traces(pspin, golden, case):
    NT, nrep, rounds, _, _, seed48 = case
    print(f"[golden exchange] NT={NT} nrep={nrep}", flush=True)
    k = make_golden.case_key("exch", NT, nrep, seed48)
    st = pspin.PSpin(pspin.SimConfig.default(N=8, nrep=NT * nrep, seed=5), pspin.DENSE)
    st.pt_init(NT, nrep, golden[k + "_beta"], seed48)
    for r in range(rounds):
        st.set_energies(golden[k + "_E"][r])
        st.pt_exchange()
        assert np.array_equal(st.pt_perm(), golden[k + "_perm"][r]), f"round {r}"
    _, _, acc, prop = st.pt_exchange_stats()
    x = st.pt_rng48()
    st.close()
    assert np.array_equal(acc, golden[k + "_acc"]) and np.array_equal(prop, golden[k + "_prop"])
    nxt = (0x5DEECE66D * x + 0xB) & ((1 << 48) - 1)
    assert nxt / 2.0 ** 48 == golden[k + "_next"][0]
