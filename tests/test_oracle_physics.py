"""Pins the oracle's FP64 physics (CPU only), each against something other
than its own formula:

- FMC masks: hand-derived comb cases, brute-force recount, the reference's own
  test case (src/test_fmc_mask.cu:70-111: omega = drand48()*N after
  srand48(42)), uniform-omega survival rate vs a Monte Carlo estimate.
- Energy: global phase invariance, homogeneity (E2 ~ lambda^2, E4 ~ lambda^4),
  per-channel trigonometric closed form on unit-modulus spins, all-ones
  configuration = minus the sum of active couplings.
- Delta E: brute-force E(new) - E(old).
- Proposals: |a_i|^2+|a_j|^2 and |a_i|^4+|a_j|^4 invariants, closed forms.
- Exchange: glibc drand48 itself plus a hand-traced 3-temperature case.
- Sparse sampling: permutation property and uniform selection.
"""
import itertools
import math

import numpy as np
import pytest


# ----------------------------------------------------------------- FMC
def test_comb_mask_gamma0_only_channel1(oracle):
    # Comb omega_k = k with sorted i<j<k<l: ch0 (i+j-k-l) and ch2 (i+k-j-l) are
    # always <= -2, so only ch1 (i+l = j+k) can pass at gamma = 0.
    N = 12
    w = oracle.omega(N, 1, 42)
    assert np.array_equal(w, np.arange(N, dtype=float))
    m = oracle.g4_mask(N, 1, w, 0.0)
    q = 0
    for l in range(3, N):
        for k in range(2, l):
            for j in range(1, k):
                for i in range(j):
                    want = 2 if i + l == j + k else 0
                    assert m[q] == want, (i, j, k, l)
                    q += 1


def test_comb_mask_gamma1_counts(oracle):
    # gamma = 1 keeps ch1 with |i+l-j-k| <= 1; count by brute force.
    N = 14
    m = oracle.g4_mask(N, 1, np.arange(N, dtype=float), 1.0)
    want = sum(1 for i, j, k, l in itertools.combinations(range(N), 4) if abs(i + l - j - k) <= 1)
    assert int((m == 2).sum()) == want and int((m != 0).sum()) == want


def test_reference_fmc_mask_case(oracle):
    # src/test_fmc_mask.cu:70-111 with its run_tests.sh arguments "10 1.0 1"
    # (comb) and the uniform branch: omega = drand48()*N after srand48(42).
    N, gamma = 10, 1.0
    oracle.srand48(42)
    w = np.array([oracle.drand48() * N for _ in range(N)])
    m = oracle.g4_mask(N, 2, w, gamma)
    q = 0
    for l in range(3, N):
        for k in range(2, l):
            for j in range(1, k):
                for i in range(j):
                    c0 = abs(w[i] + w[j] - w[k] - w[l]) <= gamma
                    c1 = abs(w[i] + w[l] - w[j] - w[k]) <= gamma
                    c2 = abs(w[i] + w[k] - w[j] - w[l]) <= gamma
                    want = 1 if c0 else 2 if c1 else 4 if c2 else 0
                    assert m[q] == want
                    q += 1
    keep = oracle.g2_keep(N, w, gamma)
    for i in range(N):
        for j in range(i, N):
            assert bool(keep[i, j] == 0) == ((i != j) and abs(w[i] - w[j]) > gamma)


def test_uniform_omega_sorted_and_splitmix(oracle):
    N, seed = 32, 42
    w = oracle.omega(N, 2, seed)
    raw = np.array([(int(z) >> 11) / 2.0 ** 53 for z in oracle.splitmix64_seq(seed + 3000, N)])
    assert np.array_equal(w, np.sort(raw))


def test_uniform_fmc_survival_rate(oracle):
    # Survival of at least one channel for U[0,1) frequencies: Monte Carlo
    # estimate of P(min over 3 channels |sum| <= gamma) with sorted indices.
    N, gamma = 40, 0.05
    w = oracle.omega(N, 2, 7)
    m = oracle.g4_mask(N, 2, w, gamma)
    rate = float((m != 0).mean())
    rng = np.random.default_rng(0)
    x = np.sort(rng.random((200000, 4)), axis=1)
    c0 = np.abs(x[:, 0] + x[:, 1] - x[:, 2] - x[:, 3]) <= gamma
    c1 = np.abs(x[:, 0] + x[:, 3] - x[:, 1] - x[:, 2]) <= gamma
    c2 = np.abs(x[:, 0] + x[:, 2] - x[:, 1] - x[:, 3]) <= gamma
    p = float((c0 | c1 | c2).mean())
    assert abs(rate - p) < 0.03, (rate, p)


# ----------------------------------------------------------------- energy
def _rand_system(N, seed=1, fmc=0):
    rng = np.random.default_rng(seed)
    g2 = rng.normal(size=(N, N))
    g2 = np.triu(g2) + np.triu(g2, 1).T
    nq = N * (N - 1) * (N - 2) * (N - 3) // 24
    g4 = rng.normal(size=nq) * 0.3
    mask = rng.choice(np.array([0, 1, 2, 4], dtype=np.uint8), size=nq) if fmc else np.ones(nq, np.uint8)
    s = rng.normal(size=N) + 1j * rng.normal(size=N)
    s *= math.sqrt(N / np.sum(np.abs(s) ** 2))
    return g2, g4, mask, s


def test_energy_phase_invariance_and_homogeneity(oracle):
    N = 9
    g2, g4, mask, s = _rand_system(N, fmc=1)
    e2, e4 = oracle.energy(N, s, g2, 1, g4, mask)
    e2p, e4p = oracle.energy(N, s * np.exp(0.7j), g2, 1, g4, mask)
    assert abs(e2 - e2p) < 1e-12 and abs(e4 - e4p) < 1e-12
    lam = 1.3
    e2l, e4l = oracle.energy(N, s * lam, g2, 1, g4, mask)
    assert abs(e2l - lam ** 2 * e2) < 1e-11 and abs(e4l - lam ** 4 * e4) < 1e-10


@pytest.mark.parametrize("ch", [0, 1, 2])
def test_energy_channel_trig_closed_form(oracle, ch):
    # Unit-modulus spins a_k = exp(i t_k): the single quartet (0,1,2,3) of N=4
    # contributes -g cos(+-t0 +- t1 +- t2 +- t3) with the channel's conjugation
    # signs: ch0 (+,+,-,-), ch1 (+,-,-,+), ch2 (+,-,+,-) (src/mc.cu:385-394).
    t = np.array([0.3, 1.1, -0.4, 2.0])
    s = np.exp(1j * t)
    g = 0.77
    sign = {0: [1, 1, -1, -1], 1: [1, -1, -1, 1], 2: [1, -1, 1, -1]}[ch]
    _, e4 = oracle.energy(4, s, np.zeros((4, 4)), 0, np.array([g]), np.array([1 << ch], dtype=np.uint8))
    assert abs(e4 - (-g * math.cos(np.dot(sign, t)))) < 1e-14


def test_energy_h2_trig_and_all_ones(oracle):
    N = 6
    g2, g4, mask, _ = _rand_system(N, seed=3, fmc=1)
    t = np.linspace(0.1, 2.9, N)
    e2, _ = oracle.energy(N, np.exp(1j * t), g2, 1, g4, mask)
    want = -sum(g2[i, j] * math.cos(t[i] - t[j]) for i in range(N) for j in range(i + 1, N))
    assert abs(e2 - want) < 1e-12
    e2, e4 = oracle.energy(N, np.ones(N, complex), g2, 1, g4, mask)
    assert abs(e2 + np.triu(g2, 1).sum()) < 1e-12
    assert abs(e4 + g4[mask != 0].sum()) < 1e-12


@pytest.mark.parametrize("N,fmc", [(8, 0), (9, 1), (11, 1)])
def test_delta_e_equals_energy_difference(oracle, N, fmc):
    g2, g4, mask, s = _rand_system(N, seed=N, fmc=fmc)
    rng = np.random.default_rng(5)
    for _ in range(20):
        i0, j0 = sorted(rng.choice(N, 2, replace=False))
        ni, nj = oracle.propose_sphere(s[i0], s[j0], *rng.random(3))
        s2 = s.copy()
        s2[i0] = ni[0] + 1j * ni[1]
        s2[j0] = nj[0] + 1j * nj[1]
        dE, scale = oracle.delta_e(N, s, g2, 1, g4, mask, i0, j0, ni, nj)
        e_old = sum(oracle.energy(N, s, g2, 1, g4, mask))
        e_new = sum(oracle.energy(N, s2, g2, 1, g4, mask))
        assert abs(dE - (e_new - e_old)) <= 1e-12 * max(1.0, scale)
        assert scale >= abs(dE)


# ----------------------------------------------------------------- proposals
def test_sphere_proposal_closed_form(oracle):
    ai, aj = np.array([0.6, -0.8]), np.array([1.2, 0.5])
    ua, u1, u2 = 0.13, 0.71, 0.29
    ni, nj = oracle.propose_sphere(ai, aj, ua, u1, u2)
    R = ai @ ai + aj @ aj
    zi, zj = complex(*ai), complex(*aj)
    wi = zi * np.exp(-2j * np.pi * u1) * math.sqrt(R / abs(zi) ** 2) * math.cos(2 * np.pi * ua)
    wj = zj * np.exp(-2j * np.pi * u2) * math.sqrt(R / abs(zj) ** 2) * math.sin(2 * np.pi * ua)
    assert abs(complex(*ni) - wi) < 1e-14 and abs(complex(*nj) - wj) < 1e-14
    assert abs(ni @ ni + nj @ nj - R) < 1e-14


def test_cube_proposal_invariant(oracle):
    rng = np.random.default_rng(2)
    for _ in range(50):
        ai, aj = rng.normal(size=2), rng.normal(size=2)
        u, u1, u2 = rng.random(3)
        ni, nj = oracle.propose_cube(ai, aj, u, u1, u2)
        s4 = (ai @ ai) ** 2 + (aj @ aj) ** 2
        assert abs((ni @ ni) ** 2 + (nj @ nj) ** 2 - s4) < 1e-13 * s4
        assert abs((nj @ nj) ** 2 / s4 - u) < 1e-12
        assert abs(math.atan2(ni[1], ni[0]) % (2 * np.pi) - 2 * np.pi * u1) < 1e-9


def test_pair_pick(oracle):
    N = 7
    seen = set()
    for u1 in np.linspace(0.01, 1.0, 40):
        for u2 in np.linspace(0.01, 1.0, 40):
            i0, j0 = oracle.pair_pick(N, u1, u2)
            assert 0 <= i0 < j0 < N
            seen.add((i0, j0))
    assert len(seen) == N * (N - 1) // 2
    assert oracle.pair_pick(N, 1.0, 1.0) == (N - 2, N - 1)  # clamps of src/mc.cu:131-136


# ----------------------------------------------------------------- exchange
def test_pt_exchange_hand_trace(oracle):
    # NT=3, nrep=1; beta = 1, 2, 4.  E = [5, 1, 3] by physical replica.
    # t=0: delta=(1-2)(5-1)=-4 -> draw u0: accept iff u0 < e^-4.
    # t=1: pair (perm[1], perm[2]); computed after t=0's outcome.
    beta = np.array([1.0, 2.0, 4.0])
    E = np.array([5.0, 1.0, 3.0])
    seed = 99
    oracle.srand48(seed)
    u = [oracle.drand48() for _ in range(2)]
    perm = np.array([0, 1, 2], dtype=np.int32)
    acc = np.zeros(2, dtype=np.int64)
    prop = np.zeros(2, dtype=np.int64)
    oracle.srand48(seed)
    betas = oracle.pt_exchange(3, 1, beta, E, perm, acc, prop)
    want = [0, 1, 2]
    k = 0
    a0 = u[k] < math.exp(-4.0); k += 1
    if a0:
        want[0], want[1] = want[1], want[0]
    d1 = (2.0 - 4.0) * (E[want[1]] - E[want[2]])
    a1 = d1 >= 0 or (u[k] < math.exp(d1))
    if a1:
        want[1], want[2] = want[2], want[1]
    assert list(perm) == want
    assert list(acc) == [int(a0), int(a1)] and list(prop) == [1, 1]
    assert all(betas[want[t]] == beta[t] for t in range(3))


def test_drand48_lcg_matches_glibc(oracle):
    # The GPU exchange steps a 48-bit LCG; check the POSIX recurrence against glibc.
    s = 12345 + 5000
    x = ((s & 0xFFFFFFFF) << 16) | 0x330E
    oracle.srand48(s)
    for _ in range(100):
        x = (0x5DEECE66D * x + 0xB) & ((1 << 48) - 1)
        assert oracle.drand48() == x / 2.0 ** 48


# ----------------------------------------------------------------- sparse sampling
def test_partial_shuffle_properties(oracle):
    import ctypes
    arr = np.arange(100, dtype=np.int64)
    oracle.lib().or_partial_shuffle(arr, 100, 10, 4042)
    assert sorted(arr.tolist()) == list(range(100))
    counts = np.zeros(50)
    for seed in range(4000):
        a = np.arange(50, dtype=np.int64)
        oracle.lib().or_partial_shuffle(a, 50, 5, seed)
        counts[a[:5]] += 1
    p = counts / 4000.0
    assert np.all(np.abs(p - 0.1) < 0.025)


def test_sparse_select_structure(oracle):
    N = 16
    w = oracle.omega(N, 2, 42)
    gamma = 0.05
    mask = oracle.g4_mask(N, 2, w, gamma)
    g_unit = np.array([oracle.normal(42 + 2000 + 12345, 4, q) for q in range(len(mask))])
    q, g, nnz = oracle.sparse_select(N, mask, g_unit, N, 42, 1.0, 0.0)
    assert nnz == int(np.count_nonzero(mask))
    assert len(q) == min(N, nnz)
    for row, gv in zip(q, g):
        i, j, k, l, ch = row
        assert i < j < k < l < N
        qq = oracle.encode(i, j, k, l)
        assert mask[qq] == (1 << ch)
        assert abs(gv - g_unit[qq] * math.sqrt(N / len(q))) < 1e-15


def test_cpu_pt_overlap_histogram_brackets_q2(oracle):
    """The |q| histogram of the CPU PT loop counts every measurement once per
    temperature, and its bin edges bracket the accumulated |q|^2: binning q^2,
    or an off-by-one bin, breaks one of the two bounds."""
    N, NT, nr, Q = 6, 3, 2, 8
    rng = np.random.default_rng(12)
    g2 = rng.normal(size=(N, N))
    g2 = (g2 + g2.T) / 2
    g4 = rng.normal(size=oracle.n_quartets(N))
    mask = np.ones(oracle.n_quartets(N), dtype=np.uint8)
    a0 = rng.normal(size=(NT * nr, 2 * N))
    a0 *= math.sqrt(N) / np.linalg.norm(a0, axis=1, keepdims=True)
    beta = np.array([0.2, 0.6, 1.8])
    out = oracle.cpu_pt(N, NT, nr, a0.reshape(-1), g2, 1, g4, mask, beta, oracle.replica_seeds(31, NT * nr),
                        3000, 5, 500, qbins=Q)
    n, qh, q2 = out["n"], out["qh"], out["q2"]
    assert n == 2500
    np.testing.assert_array_equal(qh.sum(axis=1), np.full(NT, float(n)))
    lo = (np.arange(Q) / Q) ** 2
    hi = (np.arange(1, Q + 1) / Q) ** 2
    assert np.all(qh @ lo <= q2 + 1e-9) and np.all(q2 <= qh @ hi + 1e-9), (qh @ lo, q2, qh @ hi)
    assert np.all(qh[:, 1:].sum(axis=1) > 0)  # the histogram is not degenerate


@pytest.mark.parametrize("N,fmc,gamma", [(9, 0, 0.0), (12, 2, 0.1), (11, 1, 1.0)])
def test_cpu_pt_site_lists_follow_the_enumeration(oracle, N, fmc, gamma):
    """The list sweep visits the same active quartets as the reference's
    3-type enumeration with the same per-term arithmetic, so on the same
    inputs the two CPU PT loops take the same decisions: identical counters
    and spins and energies equal to rounding.  A list that drops the
    quartets holding both sites, or counts them twice, changes Delta E by
    O(1) and the chains part within a few sweeps."""
    rng = np.random.default_rng(N)
    NT, nr = 3, 2
    w = oracle.omega(N, fmc, 5) if fmc else np.zeros(N)
    mask = oracle.g4_mask(N, fmc, w, gamma)
    g4 = rng.normal(size=len(mask))
    g4[rng.random(len(mask)) < 0.05] = 0.0  # zero couplings are skipped by both
    g2 = rng.normal(size=(N, N))
    g2 = (g2 + g2.T) / 2
    a0 = rng.normal(size=(NT * nr, 2 * N))
    a0 *= math.sqrt(N) / np.linalg.norm(a0, axis=1, keepdims=True)
    beta = np.array([0.3, 0.9, 2.5])
    seeds = oracle.replica_seeds(77, NT * nr)
    outs = [oracle.cpu_pt(N, NT, nr, a0.reshape(-1), g2, 1, g4, mask, beta, seeds, 60, 9, 10, lists=lst)
            for lst in (False, True)]
    a, b = outs
    for k in ("acc", "prop", "ex_acc", "ex_prop"):
        np.testing.assert_array_equal(a[k], b[k])
    np.testing.assert_allclose(a["spins"], b["spins"], rtol=0, atol=1e-12)
    np.testing.assert_allclose(a["e"], b["e"], rtol=1e-12)
    assert a["acc"].sum() > 0


def test_sparse_select_hand_trace(oracle):
    """A selection worked out by hand (src/mc_sparse.cu:340-349,489-537).
    N = 5: quartets q = 0..4 are (0,1,2,3), (0,1,2,4), (0,1,3,4), (0,2,3,4),
    (1,2,3,4).  mask = [0x1, 0, 0x6, 0x4, 0x2], unit couplings
    [0.5, 1.0, -2.0, 0.0, 3.0]: q1 has no channel, q3 a zero coupling, q2 two
    channels, so the candidates q*3+ch in ascending order are [0, 7, 8, 13].
    The seed is 2^64 - 4000, so the shuffle's splitmix64 state seed + 4000
    wraps to 0 and its draws are the published SplitMix64 sequence for seed 0:
    r0 = 0xE220A8397B1DCDAF, r1 = 0x6E789E6AA1B965F4.
      i = 0: j = 0 + r0 mod 4 = 3 -> swap positions 0, 3: [13, 7, 8, 0]
      i = 1: j = 1 + r1 mod 3 = 1 -> no swap
    so the two selected terms are 13 = (q4, ch 1) and 7 = (q2, ch 1), with
    g = g_unit * J4 sqrt(N / 2) + J4_0 N / 2."""
    assert 0xE220A8397B1DCDAF % 4 == 3 and 0x6E789E6AA1B965F4 % 3 == 0
    mask = np.array([1, 0, 6, 4, 2], dtype=np.uint8)
    g_unit = np.array([0.5, 1.0, -2.0, 0.0, 3.0])
    J4, J40 = 1.25, 0.3
    q, g, nnz = oracle.sparse_select(5, mask, g_unit, 2, 2 ** 64 - 4000, J4, J40)
    assert nnz == 4
    assert q.tolist() == [[1, 2, 3, 4, 1], [0, 1, 3, 4, 1]]
    s, m = J4 * math.sqrt(5 / 2), J40 * 5 / 2
    assert g.tolist() == [3.0 * s + m, -2.0 * s + m]


def _sparse_system(rng, N, nq):
    """N sites, nq random quartets (distinct sites, channel 0..2), Gaussian
    couplings, symmetric g2, replicas on the cube sum |a|^4 = N."""
    q = np.zeros((nq, 5), dtype=np.int32)
    for t in range(nq):
        q[t, :4] = np.sort(rng.choice(N, 4, replace=False))
        q[t, 4] = rng.integers(0, 3)
    gq = rng.normal(size=nq)
    g2 = rng.normal(size=(N, N))
    g2 = (g2 + g2.T) / 2
    np.fill_diagonal(g2, 0.0)
    return q, gq, g2


def _cube_spins(rng, R, N):
    a = rng.normal(size=(R, N)) + 1j * rng.normal(size=(R, N))
    a *= (N / np.sum(np.abs(a) ** 4, axis=1, keepdims=True)) ** 0.25
    return a


def test_cpu_pt_sparse_replays_step_by_step(oracle):
    """The sparse CPU chain is the composition of the pinned primitives:
    step uniforms, pair pick (src/mc.cu:129-141), cube move
    (include/spins_sparse.h:26-56), sparse Delta E (src/mc_sparse.cu:146-210)
    and the Metropolis rule, in that order, every step.  A replay in Python
    from those primitives must end on the same spins and counters (no
    exchange: one temperature)."""
    rng = np.random.default_rng(4)
    N, nr, sweeps = 10, 3, 6
    q, gq, g2 = _sparse_system(rng, N, N)
    a0 = _cube_spins(rng, nr, N)
    beta = np.array([1.3])
    seeds = oracle.replica_seeds(19, nr)
    out = oracle.cpu_pt_sparse(N, 1, nr, a0, g2, 1, q, gq, beta, seeds, sweeps, 3, 0)
    got = out["spins"].reshape(nr, N, 2)
    acc = 0
    for r in range(nr):
        s = a0[r].copy()
        for sw in range(sweeps):
            for k in range(N // 2):
                u = oracle.step_uniforms(int(seeds[r]), sw, k)
                i0, j0 = oracle.pair_pick(N, u[0], u[1])
                ni, nj = oracle.propose_cube(s[i0], s[j0], u[2], u[3], u[4])
                ni, nj = complex(*ni), complex(*nj)
                dE, _ = oracle.delta_e_sparse(N, s, g2, 1, q, gq, i0, j0, ni, nj)
                if dE <= 0 or u[5] < math.exp(-beta[0] * dE):
                    s[i0], s[j0] = ni, nj
                    acc += 1
        np.testing.assert_allclose(got[r, :, 0] + 1j * got[r, :, 1], s, rtol=0, atol=1e-12)
    assert out["acc"].sum() == acc and 0 < acc < nr * sweeps * (N // 2)
    assert out["prop"].sum() == nr * sweeps * (N // 2)


def test_cpu_pt_sparse_energy_and_constraint(oracle):
    """Over a PT run the sparse chains keep sum |a|^4 = N, and the energies
    they carry (sums of per-step Delta E, reported per slot) equal the FP64
    energies of the final spins: a dropped, doubled or sign-flipped term of
    Delta E breaks the second.  At beta = 0 every move is accepted."""
    rng = np.random.default_rng(8)
    N, NT, nr, sweeps = 14, 3, 2, 300
    q, gq, g2 = _sparse_system(rng, N, N)
    a0 = _cube_spins(rng, NT * nr, N)
    beta = np.array([0.4, 1.0, 2.5])
    out = oracle.cpu_pt_sparse(N, NT, nr, a0, g2, 1, q, gq, beta, oracle.replica_seeds(5, NT * nr), sweeps, 11,
                               sweeps - 1)
    sp = out["spins"].reshape(NT * nr, N, 2)
    a = sp[..., 0] + 1j * sp[..., 1]
    np.testing.assert_allclose(np.sum(np.abs(a) ** 4, axis=1), N, rtol=1e-10)
    e_full = sum(sum(oracle.energy_sparse(N, a[r], g2, 1, q, gq)) for r in range(NT * nr))
    assert abs(out["e"].sum() * N - e_full) < 1e-9 * max(1.0, abs(e_full)), (out["e"].sum() * N, e_full)
    assert out["ex_prop"].sum() == (NT - 1) * nr and out["n"] == 1
    hot = oracle.cpu_pt_sparse(N, 1, 2, a0[:2], g2, 1, q, gq, np.array([0.0]), oracle.replica_seeds(5, 2), 20, 11, 0)
    assert hot["acc"].sum() == hot["prop"].sum() == 2 * 20 * (N // 2)
