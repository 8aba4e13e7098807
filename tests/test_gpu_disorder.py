"""GPU parity of the disorder / initial-state path (libpspin.so vs the oracle).

Bit-exact: omega, FMC g4 masks, g2 survivor pattern, active counts, sparse
quartet selection (indices and channels).  FP64 values (couplings, initial
spins, initial energies) within 1e-12 relative: the GPU's Box-Muller uses the
device log/sincos, the oracle glibc's.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _mk(P, variant=0, **kw):
    cfg = P.SimConfig.default(**kw)
    return P.PSpin(cfg, variant), cfg


def _normals(O, seed, tag, idx):
    return np.array([O.normal(seed, tag, int(k)) for k in idx])


@pytest.mark.parametrize("N,fmc,gamma", [(10, 0, 0.0), (16, 2, 0.05), (12, 1, 1.0), (32, 2, 0.05)])
def test_dense_disorder(pspin, oracle, N, fmc, gamma):
    print(f"[dense disorder] N={N} fmc={fmc} gamma={gamma}", flush=True)
    seed = 77777
    st, cfg = _mk(pspin, N=N, nrep=3, seed=seed, fmc_mode=fmc, gamma=gamma, J=2.0, alpha=0.5, alpha0=0.0)
    w = st.omega()
    w_or = oracle.omega(N, fmc, seed) if fmc else np.zeros(N)
    assert np.array_equal(w, w_or)
    g4, mask = st.g4()
    mask_or = oracle.g4_mask(N, fmc, w_or, gamma)
    assert np.array_equal(mask, mask_or), "FMC mask not bit-exact"
    n_quart = int(mask_or.astype(bool).sum()) if fmc else len(mask_or)
    assert st.info.n_quart_active == n_quart
    g2 = st.g2()
    if fmc == 1:
        assert st.info.h2_active == 0 and st.info.n_pairs_active == 0 and not g2.any()
    else:
        keep = oracle.g2_keep(N, w_or, gamma) if fmc else np.ones((N, N), np.uint8)
        assert np.array_equal(g2 != 0, keep.astype(bool))
        n_pairs = int(np.triu(keep).sum())
        assert st.info.n_pairs_active == n_pairs
        sig2 = 1.0 * math.sqrt(N / n_pairs)
        iu = np.triu_indices(N)
        z = _normals(oracle, seed + 1000, 2, iu[0] * N + iu[1])
        want = np.where(keep[iu] != 0, z * sig2, 0.0)
        np.testing.assert_allclose(g2[iu], want, rtol=1e-12, atol=1e-14)
        assert np.array_equal(g2, g2.T)
    sig4 = 1.0 * math.sqrt(N / n_quart)
    sel = np.arange(0, len(g4), max(1, len(g4) // 400))
    z4 = _normals(oracle, seed + 2000 + 12345, 4, sel)
    np.testing.assert_allclose(g4[sel], z4 * sig4, rtol=1e-12, atol=1e-14)
    st.close()


def test_mean_shift(pspin, oracle):
    # src/test_mean_shift.cu: shift = J_p0 * N / n_surv on non-zero entries.
    N, J0, a0 = 12, 0.5, 0.5
    st, cfg = _mk(pspin, N=N, nrep=1, seed=4321, fmc_mode=0, J0=J0, alpha0=a0)
    st0, _ = _mk(pspin, N=N, nrep=1, seed=4321, fmc_mode=0, J0=0.0, alpha0=a0)
    g4, _ = st.g4()
    g40, _ = st0.g4()
    np.testing.assert_allclose(g4 - g40, a0 * J0 * N / math.comb(N, 4), rtol=0, atol=1e-12)
    g2, g20 = st.g2(), st0.g2()
    nz = g20 != 0
    np.testing.assert_allclose((g2 - g20)[nz], (1 - a0) * J0 * N / (N * (N + 1) / 2), rtol=0, atol=1e-12)
    st.close()
    st0.close()


@pytest.mark.parametrize("N,fmc,gamma", [(16, 2, 0.05), (24, 2, 0.05), (12, 0, 0.0), (20, 1, 1.0)])
def test_sparse_selection_bit_exact(pspin, oracle, N, fmc, gamma):
    print(f"[sparse selection] N={N} fmc={fmc}", flush=True)
    seed = 42
    st, cfg = _mk(pspin, 1, N=N, nrep=2, seed=seed, fmc_mode=fmc, gamma=gamma, alpha0=0.5, J0=0.0)
    q, g = st.sparse_quartets()
    w = oracle.omega(N, fmc, seed) if fmc else np.zeros(N)
    mask = oracle.g4_mask(N, fmc, w, gamma)
    g_unit = _normals(oracle, seed + 2000 + 12345, 4, range(len(mask)))
    q_or, g_or, nnz = oracle.sparse_select(N, mask, g_unit, N, seed, cfg.alpha * cfg.J, cfg.alpha0 * cfg.J0)
    assert st.info.n_quart_active == nnz
    assert np.array_equal(q, q_or), "sparse quartet selection not bit-exact"
    np.testing.assert_allclose(g, g_or, rtol=1e-12, atol=1e-14)
    st.close()


@pytest.mark.parametrize("variant", [0, 1])
def test_initial_spins_and_energy(pspin, oracle, variant):
    N, nrep, seed = 14, 5, 2024
    st, cfg = _mk(pspin, variant, N=N, nrep=nrep, seed=seed, fmc_mode=2, gamma=0.2)
    s = st.spins()
    seeds = oracle.replica_seeds(seed, nrep)
    for r in range(nrep):
        z = _normals(oracle, int(seeds[r]), 1, range(2 * N))
        raw = z[0::2] + 1j * z[1::2]
        if variant == 0:
            want = raw * math.sqrt(N / np.sum(np.abs(raw) ** 2))
            assert abs(np.sum(np.abs(s[r]) ** 2) - N) < 1e-10
        else:
            want = raw * (N / np.sum(np.abs(raw) ** 4)) ** 0.25
            assert abs(np.sum(np.abs(s[r]) ** 4) - N) < 1e-10
        np.testing.assert_allclose(s[r], want, rtol=1e-12, atol=1e-13)
    E, acc, prop = st.results()
    e2g, e4g = st.split_energies()
    g2 = st.g2()
    for r in range(nrep):
        if variant == 0:
            g4, mask = st.g4()
            e2, e4 = oracle.energy(N, s[r], g2, st.info.h2_active, g4, mask)
        else:
            q, gq = st.sparse_quartets()
            e2, e4 = oracle.energy_sparse(N, s[r], g2, st.info.h2_active, q, gq)
        assert abs(e2g[r] - e2) <= 1e-12 * max(1, abs(e2))
        assert abs(e4g[r] - e4) <= 1e-12 * max(1, abs(e4))
        assert abs(E[r] - (e2 + e4)) <= 1e-12 * max(1, abs(e2 + e4))
    assert not acc.any() and not prop.any()
    st.close()


@pytest.mark.parametrize("N,nrep,fmc,gamma", [(33, 10, 0, 0.0), (64, 6, 2, 0.05), (40, 9, 1, 1.0)])
def test_dense_energies_match_oracle(pspin, oracle, N, nrep, fmc, gamma):
    """FP64 full energies (workgroups of 4 replicas over balanced chunks of
    colex runs; 10 and 9 replicas leave a partial group) against the oracle's
    FP64 energy, after a few sweeps away from the initial state."""
    print(f"[dense energies] N={N} nrep={nrep} fmc={fmc}", flush=True)
    st, cfg = _mk(pspin, 0, N=N, nrep=nrep, seed=77, fmc_mode=fmc, gamma=gamma)
    st.set_uniform_beta(1.0)
    st.sweep(3)
    z = st.spins()
    e2g, e4g = st.split_energies()
    g2 = st.g2()
    g4, mask = st.g4()
    for r in range(nrep):
        e2, e4 = oracle.energy(N, z[r], g2, st.info.h2_active, g4, mask)
        assert abs(e2g[r] - e2) <= 1e-12 * max(1, abs(e2)), (r, e2g[r], e2)
        assert abs(e4g[r] - e4) <= 1e-12 * max(1, abs(e4)), (r, e4g[r], e4)
    st.close()
