"""Size-independent properties at the benchmark configurations' full sizes.

The oracle replays and statistics run at sizes it finishes in seconds; here
the GPU runs the configurations the benchmark quotes and checks what must
hold at any size:
  * the constraint sum|a|^2 = N of every replica to FP64 rounding;
  * running energies equal to the FP64 full energies right after a resync,
    and within 1e-5 N after further FP32-Delta-E sweeps;
  * counters: N/2 proposals per replica per sweep, acceptances bounded by
    proposals, one exchange proposal per ladder and temperature pair per round;
  * PT keeps every replica in its ladder (slot t*nrep+k holds a replica r with
    r mod nrep == k), and the permutation stays a permutation;
  * Metropolis acceptance falls from T_max to T_min.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _check(st, NT, nr, sweeps, tag):
    N = st.N
    z = st.spins()
    assert np.max(np.abs(np.sum(np.abs(z) ** 2, axis=1) - N)) < 1e-9 * N, tag
    perm = st.pt_perm()
    assert np.array_equal(np.sort(perm), np.arange(NT * nr)), tag
    assert np.all(perm.reshape(NT, nr) % nr == np.arange(nr)[None, :]), tag
    E, acc, prop = st.results()
    assert np.all(prop == sweeps * (N // 2)), tag
    assert np.all((acc >= 0) & (acc <= prop)), tag
    _, _, _, xp = st.pt_exchange_stats()
    assert np.all(xp == sweeps * nr), tag
    rate = (acc[perm.reshape(NT, nr)].sum(axis=1) / prop[perm.reshape(NT, nr)].sum(axis=1))
    assert rate[0] > rate[-1], (tag, rate)
    return E


def _energies(st):
    e2, e4 = st.split_energies()
    return e2 + e4


@pytest.mark.parametrize("name,kw,NT,nr,Tmin,groups,fmt,sweeps", [
    ("config 2: N=64 FC, 32 x 8, 8 pair groups", dict(N=64, fmc_mode=0), 32, 8, 0.2, 8, 0, 512),
    ("config 1: N=32 uniform FMC, 20 x 4", dict(N=32, fmc_mode=2, gamma=0.05), 20, 4, 0.2, 0, 0, 1024),
    ("config 4: N=128 comb, BF16 hi/lo, 16 x 4", dict(N=128, fmc_mode=1, gamma=1.0), 16, 4, 0.3, 0, 1, 256),
])
def test_full_size_invariants(pspin, oracle, name, kw, NT, nr, Tmin, groups, fmt, sweeps):
    print(f"[full size] {name}", flush=True)
    cfg = pspin.SimConfig.default(nrep=NT * nr, seed=42, **kw)
    st = pspin.PSpin(cfg, pspin.DENSE)
    st.set_coupling_format(fmt)
    st.set_pair_schedule(groups)
    T, beta = oracle.temperatures(NT, Tmin, 1.6, 1)
    st.pt_init(NT, nr, beta, 42 + 5000)
    st.set_resync_interval(sweeps // 2)
    st.pt_run(sweeps, 1)  # resyncs after sweeps/2 and sweeps
    E = _check(st, NT, nr, sweeps, name)
    E64 = _energies(st)
    assert np.max(np.abs(E - E64) / np.maximum(1.0, np.abs(E64))) < 1e-12, name
    st.set_resync_interval(0)
    st.pt_run(64, 1)
    E = _check(st, NT, nr, sweeps + 64, name)
    drift = np.max(np.abs(E - _energies(st)))
    print(f"  running-energy drift after 64 unresynced sweeps: {drift:.2e} (N = {st.N})", flush=True)
    assert drift <= 1e-5 * st.N, name
    st.close()


def test_full_size_learned_mixture(pspin, oracle):
    """Config 5 system (N=64 uniform FMC, 32 x 8) with one learned move per 4
    sweeps: the constraint holds and every move leaves the FP64 energy."""
    NT, nr = 32, 8
    cfg = pspin.SimConfig.default(N=64, nrep=NT * nr, seed=42, fmc_mode=2, gamma=0.05)
    st = pspin.PSpin(cfg, pspin.DENSE)
    T, beta = oracle.temperatures(NT, 0.2, 1.6, 1)
    st.pt_init(NT, nr, beta, 42 + 5000)
    st.pt_run(128, 1)
    lp = pspin.LearnedMove(st, pspin.LPConfig.default(hidden=512, batch_cap=1024, data_cap=8, seed=1,
                                                      log_sigma0=np.log(0.02)))
    lp.run(256, 1, 4)
    _check(st, NT, nr, 128 + 256, "config 5 mixture")
    lp.move()
    E, _, _ = st.results()
    _, dbg, _, _ = lp.last()
    assert np.array_equal(E, np.where(dbg[:, 5] > 0, dbg[:, 3], dbg[:, 2]))
    assert np.max(np.abs(E - _energies(st)) / np.maximum(1.0, np.abs(E))) < 1e-11
    a, p = lp.stats()
    print(f"  learned-move acceptance T_max {a[0] / p[0]:.3f}, T_min {a[-1] / p[-1]:.3f}", flush=True)
    assert np.all(p == (256 // 4 + 1) * nr)
    lp.close()
    st.close()
