"""Batched sparse PT (pspin_batch_*: one workgroup per disorder sample, 4 or
8 lanes per chain) against single PSPIN_SPARSE states with the same seeds and
against the FP64 oracle.

A batch sample starts from the same disorder, spins and RNG seeds as the
single state (same set-up path), runs the same Philox stream and the same
exchange stream (srand48(seed+5000)), so chains agree with the single-state
kernel up to FP32 summation order of Delta E: spins to 1e-9, running
energies to 1e-5 N, identical permutations and counters.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _single(pspin, N, seed, NT, nr, beta, fmc, gamma):
    cfg = pspin.SimConfig.default(N=N, nrep=NT * nr, seed=seed, fmc_mode=fmc, gamma=gamma)
    st = pspin.PSpin(cfg, pspin.SPARSE)
    st.pt_init(NT, nr, beta, seed + 5000)
    return st


@pytest.mark.parametrize("pt_freq", [0, 1, 3])
def test_batch_matches_single_states(pspin, oracle, pt_freq):
    Ns, seeds = [12, 24, 16, 20], [101, 202, 303, 404]
    NT, nr = 3, 2
    T, beta = oracle.temperatures(NT, 0.5, 1.5, 1)
    base = pspin.SimConfig.default(fmc_mode=2, gamma=0.2)
    b = pspin.PSpinBatch(base, Ns, seeds, NT, nr, beta)
    b.set_resync_interval(0)
    assert b.n_samples == 4 and b.chains == 4 * NT * nr
    assert b.pair_updates_per_sweep == sum(NT * nr * (n // 2) for n in Ns)
    sweeps = 24
    b.run(sweeps, pt_freq)
    for s, (N, seed) in enumerate(zip(Ns, seeds)):
        print(f"[batch vs single] N={N} seed={seed} pt_freq={pt_freq}", flush=True)
        st = _single(pspin, N, seed, NT, nr, beta, 2, 0.2)
        st.set_resync_interval(0)
        if pt_freq:
            st.pt_run(sweeps, pt_freq)
            np.testing.assert_array_equal(b.perm(s), st.pt_perm())
            _, _, acc_s, prop_s = st.pt_exchange_stats()
            acc_b, prop_b = b.exchange_stats(s)
            np.testing.assert_array_equal(acc_b, acc_s)
            np.testing.assert_array_equal(prop_b, prop_s)
        else:
            st.sweep(sweeps)
        Eb, ab, pb = b.results(s)
        Es, as_, ps_ = st.results()
        np.testing.assert_allclose(b.spins(s), st.spins(), atol=1e-9)
        # running energies accumulate FP32 Delta E in a different order
        np.testing.assert_allclose(Eb, Es, rtol=0, atol=1e-5 * N)
        np.testing.assert_array_equal(ab, as_)
        np.testing.assert_array_equal(pb, ps_)
        st.close()
    b.close()


def test_batch_matches_single_states_at_cfg3_shape(pspin, oracle):
    """The benchmark's shape per sample (24 temperatures in [0.2, 1.6] x 4
    replicas = 96 chains, uniform FMC gamma 0.05, N sampled quartets) at its
    two largest sizes, N = 48 and 64, with an exchange after every sweep and
    the FP64 resync at sweep 8."""
    Ns, seeds = [48, 64, 64, 48], [42, 43, 44, 45]
    NT, nr = 24, 4
    T, beta = oracle.temperatures(NT, 0.2, 1.6, 1)
    base = pspin.SimConfig.default(fmc_mode=2, gamma=0.05)
    b = pspin.PSpinBatch(base, Ns, seeds, NT, nr, beta)
    b.set_resync_interval(8)
    sweeps = 12
    b.run(sweeps, 1)
    for s, (N, seed) in enumerate(zip(Ns, seeds)):
        print(f"[batch cfg3 shape] N={N} seed={seed}", flush=True)
        st = _single(pspin, N, seed, NT, nr, beta, 2, 0.05)
        st.set_resync_interval(8)
        st.pt_run(sweeps, 1)
        np.testing.assert_array_equal(b.perm(s), st.pt_perm())
        _, _, acc_s, prop_s = st.pt_exchange_stats()
        acc_b, prop_b = b.exchange_stats(s)
        np.testing.assert_array_equal(acc_b, acc_s)
        np.testing.assert_array_equal(prop_b, prop_s)
        Eb, ab, pb = b.results(s)
        Es, as_, ps_ = st.results()
        np.testing.assert_allclose(b.spins(s), st.spins(), atol=1e-9)
        np.testing.assert_allclose(Eb, Es, rtol=0, atol=1e-5 * N)
        np.testing.assert_array_equal(ab, as_)
        np.testing.assert_array_equal(pb, ps_)
        st.close()
    b.close()


@pytest.mark.parametrize("NT,nr", [(4, 24), (4, 32), (5, 32), (2, 96)])
def test_batch_workgroup_shapes_match_single_states(pspin, oracle, NT, nr):
    """96, 128, 160 and 192 chains per sample: 8 lanes per chain in 768 and
    1024 threads, then 4 lanes per chain (more than 128 chains); 96 replicas
    per temperature take the exchange's path for more than 64 per level."""
    Ns, seeds = [12, 16], [505, 606]
    T, beta = oracle.temperatures(NT, 0.4, 1.5, 1)
    base = pspin.SimConfig.default(fmc_mode=2, gamma=0.2)
    b = pspin.PSpinBatch(base, Ns, seeds, NT, nr, beta)
    b.set_resync_interval(0)
    sweeps = 12
    b.run(sweeps, 1)
    for s, (N, seed) in enumerate(zip(Ns, seeds)):
        print(f"[batch shapes] chains={NT * nr} N={N} seed={seed}", flush=True)
        st = _single(pspin, N, seed, NT, nr, beta, 2, 0.2)
        st.set_resync_interval(0)
        st.pt_run(sweeps, 1)
        np.testing.assert_array_equal(b.perm(s), st.pt_perm())
        _, _, acc_s, prop_s = st.pt_exchange_stats()
        acc_b, prop_b = b.exchange_stats(s)
        np.testing.assert_array_equal(acc_b, acc_s)
        np.testing.assert_array_equal(prop_b, prop_s)
        Eb, ab, pb = b.results(s)
        Es, as_, ps_ = st.results()
        np.testing.assert_allclose(b.spins(s), st.spins(), atol=1e-9)
        np.testing.assert_allclose(Eb, Es, rtol=0, atol=1e-5 * N)
        np.testing.assert_array_equal(ab, as_)
        np.testing.assert_array_equal(pb, ps_)
        st.close()
    b.close()


def test_batch_energy_constraint_and_resync(pspin, oracle):
    Ns, seeds = [24, 32, 48, 64, 32, 24], [1, 2, 3, 4, 5, 6]
    NT, nr = 4, 2
    T, beta = oracle.temperatures(NT, 0.2, 1.6, 1)
    base = pspin.SimConfig.default(fmc_mode=2, gamma=0.05)
    b = pspin.PSpinBatch(base, Ns, seeds, NT, nr, beta)
    b.set_resync_interval(32)
    b.run(96, 1)  # resync after sweeps 32, 64, 96
    for s, (N, seed) in enumerate(zip(Ns, seeds)):
        st = _single(pspin, N, seed, NT, nr, beta, 2, 0.05)
        g2 = st.g2()
        q, gq = st.sparse_quartets()
        z = b.spins(s)
        E, _, prop = b.results(s)
        assert np.all(prop == 96 * (N // 2))
        assert np.max(np.abs(np.sum(np.abs(z) ** 4, axis=1) - N)) < 1e-9
        for c in range(NT * nr):
            e2, e4 = oracle.energy_sparse(N, z[c], g2, st.info.h2_active, q, gq)
            assert abs(E[c] - (e2 + e4)) <= 1e-9 * max(1.0, abs(e2 + e4))
        st.close()
    b.close()


def _batch_state(b, n):
    out = []
    for s in range(n):
        E, a, p = b.results(s)
        out.append((b.spins(s), E, a, p, b.perm(s), *b.exchange_stats(s)))
    return out


def _same(x, y):
    for u, v in zip(x, y):
        for p, q in zip(u, v):
            np.testing.assert_array_equal(p, q)


def test_batch_result_independent_of_launch_split_and_autotune(pspin, oracle):
    """Sweeps per launch is a pure launch-shape choice: every chain's state is
    device-resident, so any split gives the same bits; the autotuner restores
    the batch after its trials."""
    Ns, seeds = [16, 32, 24, 48], [7, 8, 9, 10]
    NT, nr = 4, 2
    T, beta = oracle.temperatures(NT, 0.3, 1.6, 1)
    base = pspin.SimConfig.default(fmc_mode=2, gamma=0.1)
    runs = []
    for chunk in (16, 64, 256):
        b = pspin.PSpinBatch(base, Ns, seeds, NT, nr, beta)
        b.set_resync_interval(100)
        b.set_chunk(chunk)
        b.run(700, 1)  # graph replays plus a remainder launch, resyncs at 100, 200, ...
        runs.append(_batch_state(b, len(Ns)))
        b.close()
    _same(runs[0], runs[1])
    _same(runs[0], runs[2])
    b = pspin.PSpinBatch(base, Ns, seeds, NT, nr, beta)
    b.set_resync_interval(100)
    b.run(300, 1)
    before = _batch_state(b, len(Ns))
    chunk, ms = b.autotune(256, 1)
    print(f"  autotuned sweeps/launch {chunk}: {ms:.4f} ms/sweep")
    assert chunk in (16, 32, 64, 128, 256) and ms > 0
    _same(before, _batch_state(b, len(Ns)))
    b.run(400, 1)
    _same(runs[0], _batch_state(b, len(Ns)))  # 300 + 400 = 700 sweeps, same bits
    b.close()
