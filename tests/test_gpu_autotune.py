"""The sweep autotuner times candidates on the live state and leaves it untouched."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,fmc", [(16, 0), (32, 2)])
def test_autotune_restores_state_and_picks_a_candidate(pspin, oracle, N, fmc):
    cfg = pspin.SimConfig.default(N=N, nrep=32, seed=5, fmc_mode=fmc, gamma=0.1)
    st = pspin.PSpin(cfg, pspin.DENSE)
    T, beta = oracle.temperatures(4, 0.5, 2.0, 1)
    st.pt_init(4, 8, beta, 11)
    st.pt_run(20, 1)
    s0 = st.spins()
    E0, a0, p0 = st.results()
    _, c0 = st.rng_info()
    kernel, block, ms = st.autotune(4)
    print(f"  N={N} fmc={fmc}: kernel {kernel} block {block} {ms:.4f} ms/sweep")
    assert ms > 0
    assert st.info.kernel == kernel
    assert kernel in (pspin.KERNEL_LISTS, pspin.KERNEL_PAIRS, pspin.KERNEL_GROUPS)
    if fmc != 0:
        assert kernel in (pspin.KERNEL_LISTS, pspin.KERNEL_GROUPS)
    assert block % 64 == 0 and 64 <= block <= 1024
    assert np.array_equal(st.spins(), s0)
    E1, a1, p1 = st.results()
    assert np.array_equal(E1, E0) and np.array_equal(a1, a0) and np.array_equal(p1, p0)
    assert st.rng_info()[1] == c0
    # the chain continues with the chosen kernel
    st.pt_run(20, 1)
    _, _, p2 = st.results()
    assert np.all(p2 == p0 + 20 * (N // 2))
    st.close()
