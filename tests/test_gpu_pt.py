"""GPU parallel tempering (libpspin.so) vs the reference's host exchange.

Bit-exact: given the same energies and srand48 seed, the GPU exchange makes
the same decisions as src/parallel_tempering.cu:316-352 run with glibc
drand48 (the oracle), leaves the same permutation, counters and betas, and
the same 48-bit stream state.  Then the reference's test_replica_exchange
(src/test_replica_exchange.cu "8 6 65536 5", rates in [0,1] hard, soft
monotonicity/ergodicity reported), and the T_min energy series.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("NT,nrep", [(6, 1), (32, 8), (5, 100), (3, 64)])
def test_exchange_bit_exact(pspin, oracle, NT, nrep):
    print(f"[exchange bit-exact] NT={NT} nrep={nrep}", flush=True)
    cfg = pspin.SimConfig.default(N=8, nrep=NT * nrep, seed=5)
    st = pspin.PSpin(cfg, 0)
    T, beta = oracle.temperatures(NT, 0.2, 1.6, 1)
    seed48 = 42 + 5000
    st.pt_init(NT, nrep, beta, seed48)
    oracle.srand48(seed48)
    perm = np.arange(NT * nrep, dtype=np.int32)
    acc = np.zeros(NT - 1, dtype=np.int64)
    prop = np.zeros(NT - 1, dtype=np.int64)
    rng = np.random.default_rng(NT * 1000 + nrep)
    for rnd in range(25):
        # energies with a trend along the ladder plus noise, so both branches occur
        E = rng.normal(size=NT * nrep) * 3.0 + np.linspace(0, -5, NT * nrep)
        st.set_energies(E)
        st.pt_exchange()
        betas = oracle.pt_exchange(NT, nrep, beta, E, perm, acc, prop)
        gperm = st.pt_perm()
        assert np.array_equal(gperm, perm), f"round {rnd}"
    acc_w, prop_w, acc_t, prop_t = st.pt_exchange_stats()
    assert np.array_equal(acc_t, acc) and np.array_equal(prop_t, prop)
    assert np.array_equal(acc_w, acc) and np.array_equal(prop_w, prop)
    # stream state: advance glibc by one more draw on both sides and compare
    x = st.pt_rng48()
    nxt = (0x5DEECE66D * x + 0xB) & ((1 << 48) - 1)
    assert oracle.drand48() == nxt / 2.0 ** 48
    st.pt_reset_window()
    acc_w, prop_w, acc_t2, _ = st.pt_exchange_stats()
    assert not acc_w.any() and not prop_w.any() and np.array_equal(acc_t2, acc)
    st.close()


def _band_energies(NT, nrep, beta, perm, x, rng):
    """Energies by replica for which every exchange pair of the next round
    takes a draw and its delta lies within 2e-5 of log u of the very draw it
    takes (or below -700): every decision of the round goes to the FP64 test
    on the draw itself.  x is the 48-bit drand48 state before the round; the
    construction follows the reference's sequential rule (levels in order,
    ladders in order within a level, a draw only for delta < 0)."""
    mask = (1 << 48) - 1
    E = np.zeros(NT * nrep)
    ec = rng.normal(size=nrep)
    E[perm[:nrep]] = ec
    band = 0
    for t in range(NT - 1):
        db = beta[t] - beta[t + 1]
        for r in range(nrep):
            x = (0x5DEECE66D * x + 0xB) & mask
            u = x / 2.0 ** 48
            if rng.random() < 0.05:
                d = -800.0
            else:
                # (not exactly log u: at an exact tie the FP64 exponential and
                # glibc's may round a last bit apart, which decides the tie)
                d = (np.log(u) if u > 0 else -1.0) + rng.choice([2e-5, -2e-5, 1e-6, -1e-6, 1e-9, -1e-9])
                if d >= 0:
                    d = -1.0
            eb = ec[r] - d / db
            E[perm[(t + 1) * nrep + r]] = eb
            delta = db * (ec[r] - eb)
            band += (delta < -700.0) or abs(delta - np.log(u)) <= 3e-5
            if not (delta >= 0 or u < np.exp(delta)):
                ec[r] = eb
    return E, band


@pytest.mark.parametrize("NT,nrep", [(8, 4), (32, 8), (5, 100)])
def test_exchange_band_decisions_bit_exact(pspin, oracle, NT, nrep):
    # The exchange decides u < exp(delta) from logf of the draw outside a
    # 3e-5 band and by the FP64 test on the draw itself inside it (and below
    # delta = -700); random energies reach the band for about 6e-5 of the
    # draws.  Here the energies are built so that (nearly) every pair of
    # every round is a band case, and the GPU must still match the oracle
    # bit for bit (permutation, counters, drand48 state).  100 ladders run
    # the many-ladder kernel.
    print(f"[exchange band] NT={NT} nrep={nrep}", flush=True)
    st = pspin.PSpin(pspin.SimConfig.default(N=8, nrep=NT * nrep, seed=5), 0)
    T, beta = oracle.temperatures(NT, 0.2, 1.6, 1)
    seed48 = 1234 + NT
    st.pt_init(NT, nrep, beta, seed48)
    oracle.srand48(seed48)
    perm = np.arange(NT * nrep, dtype=np.int32)
    acc = np.zeros(NT - 1, dtype=np.int64)
    prop = np.zeros(NT - 1, dtype=np.int64)
    rng = np.random.default_rng(NT * 7 + nrep)
    x = st.pt_rng48()
    band = pairs = 0
    for rnd in range(10):
        E, b = _band_energies(NT, nrep, beta, perm, x, rng)
        band += b
        pairs += (NT - 1) * nrep
        st.set_energies(E)
        st.pt_exchange()
        oracle.pt_exchange(NT, nrep, beta, E, perm, acc, prop)
        np.testing.assert_array_equal(st.pt_perm(), perm, err_msg=f"round {rnd}")
        x = st.pt_rng48()
    acc_w, prop_w, acc_t, prop_t = st.pt_exchange_stats()
    np.testing.assert_array_equal(acc_t, acc)
    np.testing.assert_array_equal(prop_t, prop)
    nxt = (0x5DEECE66D * x + 0xB) & ((1 << 48) - 1)
    assert oracle.drand48() == nxt / 2.0 ** 48
    print(f"band cases {band} of {pairs} pairs", flush=True)
    assert band >= pairs // 2
    st.close()


def test_reference_replica_exchange(pspin, oracle):
    # src/test_replica_exchange.cu with run_tests.sh "8 6 65536 5" (seed 65432,
    # linear T in [0.5, 2], srand48(seed+7777)); 16380 sweeps here.
    N, NT, sweeps, pt_freq = 8, 6, 16380, 5
    cfg = pspin.SimConfig.default(N=N, nrep=NT, T=2.0, J=2.0, J0=0.0, alpha=0.5, alpha0=0.0, seed=65432)
    st = pspin.PSpin(cfg, 0)
    T, beta = oracle.temperatures(NT, 0.5, 2.0, 0)
    st.pt_init(NT, 1, beta, 65432 + 7777)
    st.pt_enable_series(sweeps // pt_freq + 8)
    st.pt_run(sweeps, pt_freq)
    _, _, acc, prop = st.pt_exchange_stats()
    assert np.all(prop == sweeps // pt_freq)
    rates = acc / prop
    assert np.all((rates >= 0) & (rates <= 1))
    print("  rates:", np.round(rates, 4))
    viol = int(np.sum(rates[1:] > rates[:-1] + 0.15))
    print(f"  monotonicity violations (soft): {viol}")
    # T_min series = energy of the replica at the coldest slot after each round
    ser = st.pt_series()
    assert ser.shape == (sweeps // pt_freq, 1)
    E, _, _ = st.results()
    perm = st.pt_perm()
    assert ser[-1, 0] == E[perm[NT - 1]]
    st.close()


def test_pt_run_counts_and_betas(pspin, oracle):
    NT, nr = 4, 3
    cfg = pspin.SimConfig.default(N=12, nrep=NT * nr, seed=17)
    st = pspin.PSpin(cfg, 0)
    T, beta = oracle.temperatures(NT, 0.3, 1.5, 1)
    st.pt_init(NT, nr, beta, 17 + 5000)
    st.pt_run(7, 3)   # exchanges after sweeps 3 and 6
    st.pt_run(5, 3)   # after 9 and 12
    _, _, acc, prop = st.pt_exchange_stats()
    assert np.all(prop == 4 * nr)
    _, _, p = st.results()
    assert np.all(p == 12 * (12 // 2))
    perm = st.pt_perm()
    assert sorted(perm.tolist()) == list(range(NT * nr))
    st.close()


# (the 64- and 65-temperature cases sit on either side of the largest ladder
# that cooperative launches hand over by tagged words; 2 temperatures is the
# shortest ladder)
@pytest.mark.parametrize("N,NT,nr,kernel", [(16, 4, 2, 0), (64, 32, 8, 2), (32, 8, 4, 3), (16, 64, 2, 0),
                                            (16, 65, 1, 0), (16, 2, 16, 0)])
def test_launch_modes_give_identical_chains(pspin, oracle, N, NT, nr, kernel):
    # Whole PT rounds run as one cooperative launch with a grid barrier (auto),
    # as captured graphs, or as direct launches (with the exchange in the
    # sweep launch, or, when profiled, in a launch of its own).  The kernels
    # keep their RNG counters and PT state on the device, so every way must
    # give the same chains bit for bit.
    T, beta = oracle.temperatures(NT, 0.4, 1.6, 1)
    out = []
    for mode, prof in ((pspin.LAUNCH_AUTO, False), (pspin.LAUNCH_GRAPHS, False), (pspin.LAUNCH_DIRECT, False),
                       (pspin.LAUNCH_DIRECT, True), (pspin.LAUNCH_AUTO, True)):
        print(f"[launch mode] {mode} profiling={prof}", flush=True)
        cfg = pspin.SimConfig.default(N=N, nrep=NT * nr, seed=321)
        st = pspin.PSpin(cfg, 0)
        if kernel:
            st.set_kernel(kernel, 0)
            st.set_pair_schedule(8)
        st.pt_init(NT, nr, beta, 321 + 5000)
        st.set_launch_mode(mode)
        st.set_profiling(prof)
        st.pt_run(96, 1)
        st.pt_run(40, 3)
        E, acc, prop = st.results()
        out.append((st.spins(), E, acc, st.pt_perm(), st.pt_rng48(), st.rng_info()[1], st.pt_exchange_stats()))
        if prof:
            p = st.profile()
            assert p["n_sweep"] == 136
        st.close()
    a = out[0]
    for b in out[1:]:
        np.testing.assert_array_equal(a[0], b[0])
        np.testing.assert_array_equal(a[1], b[1])
        np.testing.assert_array_equal(a[2], b[2])
        np.testing.assert_array_equal(a[3], b[3])
        for x, y in zip(a[6], b[6]):
            np.testing.assert_array_equal(x, y)
        assert a[4] == b[4] and a[5] == b[5] == 136


def test_cooperative_rounds_match_kernel_boundary_exchange(pspin, oracle):
    # 1024 PT rounds at the N=64 benchmark shape (fully connected, 32 T x 8
    # ladders, pair-row kernel, ladder-shared pair streams): the cooperative
    # launch, whose rounds meet at an in-kernel grid barrier, against direct
    # launches with the exchange in a kernel of its own (profiling on), where
    # every handoff is ordered by a kernel boundary.  The chains must agree
    # bit for bit after every block of 64 rounds; the first block that
    # differs names the round range.
    N, NT, nr = 64, 32, 8
    T, beta = oracle.temperatures(NT, 0.2, 1.6, 1)
    sts = []
    for mode, prof in ((pspin.LAUNCH_AUTO, False), (pspin.LAUNCH_DIRECT, True)):
        cfg = pspin.SimConfig.default(N=N, nrep=NT * nr, seed=42)
        st = pspin.PSpin(cfg, 0)
        st.set_kernel(pspin.KERNEL_PAIRS, 0)
        st.set_pair_schedule(8)
        st.pt_init(NT, nr, beta, 42 + 5000)
        st.set_launch_mode(mode)
        st.set_profiling(prof)
        sts.append(st)
    block, nblocks = 64, 16
    for b in range(nblocks):
        snap = []
        for st in sts:
            st.pt_run(block, 1)
            E, acc, prop = st.results()
            snap.append((st.spins(), E, acc, st.pt_perm(), st.pt_rng48(), st.pt_exchange_stats()))
        a, c = snap
        where = f"rounds {b * block}..{(b + 1) * block - 1}"
        np.testing.assert_array_equal(a[3], c[3], err_msg=where)
        np.testing.assert_array_equal(a[1], c[1], err_msg=where)
        np.testing.assert_array_equal(a[0], c[0], err_msg=where)
        np.testing.assert_array_equal(a[2], c[2], err_msg=where)
        assert a[4] == c[4], where
        for x, y in zip(a[5], c[5]):
            np.testing.assert_array_equal(x, y, err_msg=where)
    assert sts[1].profile()["n_sweep"] == block * nblocks
    for st in sts:
        st.close()


@pytest.mark.parametrize("NT,nr", [(16, 4), (80, 1)])
def test_cross_workgroup_wait_bound_is_reported(pspin, oracle, NT, nr):
    # Cooperative launches whose cross-workgroup waits may poll only twice:
    # the rounds cannot meet in time, every wait ends at its bound, and
    # pt_run reports the expired wait instead of returning normally.  The
    # default bound gives a normal run.  16 temperatures use the tagged
    # handoffs (workgroup 0 exchanges), 80 the arrival count (the last
    # workgroup to arrive exchanges).
    T, beta = oracle.temperatures(NT, 0.5, 1.5, 1)
    for limit in (1, 0):
        st = pspin.PSpin(pspin.SimConfig.default(N=16, nrep=NT * nr, seed=5), 0)
        st.pt_init(NT, nr, beta, 9)
        st.set_wait_limit(limit)
        if limit:
            with pytest.raises(pspin.PSpinError, match="expired"):
                st.pt_run(64, 1)
        else:
            st.pt_run(64, 1)
            E, acc, prop = st.results()
            assert np.all(np.isfinite(E)) and prop.sum() > 0
        st.close()


def test_periodic_resync(pspin, oracle):
    cfg = pspin.SimConfig.default(N=14, nrep=4, seed=77)
    st = pspin.PSpin(cfg, 0)
    st.set_resync_interval(16)
    st.set_uniform_beta(1.3)
    st.sweep(48)  # resyncs after sweeps 16, 32, 48
    E, _, _ = st.results()
    e2, e4 = st.split_energies()
    np.testing.assert_array_equal(E, e2 + e4)
    st.set_resync_interval(0)
    st.sweep(5)
    E, _, _ = st.results()
    e2, e4 = st.split_energies()
    assert np.all(np.abs(E - (e2 + e4)) < 1e-4) and np.any(E != e2 + e4)
    st.close()
