"""Learned global move on the GPU against the oracle.

* The BF16-MFMA MLP: forward outputs and the loss gradient against the numpy
  oracle with the device's BF16 rounding points (oracle/mlp.py), which is
  itself pinned by finite differences and torch autograd.
* One move, replayed replica by replica with the FP64 oracle
  (or_lp_propose / or_lp_accept / or_energy): proposal amplitudes, both
  proposal densities, both energies, the MH log-ratio and the decision.
* The stationary law: learned moves alone at beta = 0 give Dirichlet(1/2)
  intensities; the PT mixture with learned moves matches plain PT.
* Training on PT snapshots lowers the loss, lengthens accepted moves and
  raises the energy ESS per move of chains of learned moves alone.
"""
import math

import numpy as np
import pytest

from oracle import mlp

pytestmark = pytest.mark.gpu

LSMIN, LSMAX = -9.0, 1.0


def _state(pspin, N=16, nrep=64, seed=7, fmc=0):
    cfg = pspin.SimConfig.default(N=N, nrep=nrep, seed=seed, fmc_mode=fmc)
    return pspin.PSpin(cfg, pspin.DENSE)


def _lp(pspin, st, hidden=128, **kw):
    cfg = pspin.LPConfig.default(hidden=hidden, batch_cap=512, data_cap=64, **kw)
    return pspin.LearnedMove(st, cfg)


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, float) - np.asarray(b, float)) / max(np.linalg.norm(b), 1e-30))


# (N, hidden, rows): a small network, and the benchmark's (N = 64, hidden 512)
# at its training batch of 4096 rows (K = 512 products, weight gradients in 8
# K slices of 512 rows).
SHAPES = [(16, 128, 200), (64, 512, 4096)]


def _lp_shape(pspin, st, hidden, rows):
    if rows <= 512:
        return _lp(pspin, st, hidden=hidden)
    return pspin.LearnedMove(st, pspin.LPConfig.default(hidden=hidden, batch_cap=rows, data_cap=4))


@pytest.mark.parametrize("N,hidden,rows", SHAPES)
def test_mlp_forward_matches_oracle(pspin, N, hidden, rows):
    st = _state(pspin, N=N)
    lp = _lp_shape(pspin, st, hidden, rows)
    d = mlp.Dims(N=st.N, hid=lp.hidden)
    assert lp.n_params == d.n_params and lp.in_pad == d.in_pad
    rng = np.random.default_rng(1)
    p = mlp.init_params(d, rng, gain=1.0)
    lp.set_params(p)
    assert np.array_equal(lp.params(), p)
    x = rng.normal(size=(rows, d.in_dim)).astype(np.float32)
    got = lp.forward(x)
    want, _ = mlp.forward(p, d, x, emulate=True)
    plain, _ = mlp.forward(p, d, x, emulate=False)
    err, err_plain = _rel(got, want), _rel(got, plain)
    print(f"  forward: rel err vs BF16-emulating oracle {err:.2e}, vs plain FP64 {err_plain:.2e}")
    assert err < 1e-3
    assert np.max(np.abs(got - want)) < 5e-3 * np.max(np.abs(want))
    assert err < err_plain / 4  # the device really rounds where the oracle says
    lp.close()
    st.close()


@pytest.mark.parametrize("N,hidden,rows", SHAPES)
def test_mlp_loss_gradient_matches_oracle(pspin, N, hidden, rows):
    st = _state(pspin, N=N)
    lp = _lp_shape(pspin, st, hidden, rows)
    d = mlp.Dims(N=st.N, hid=lp.hidden)
    rng = np.random.default_rng(2)
    p = mlp.init_params(d, rng, gain=0.5)
    lp.set_params(p)
    M = rows
    x = rng.normal(size=(M, d.in_dim)).astype(np.float32)
    U = (x[:, :d.dim] + 0.05 * rng.normal(size=(M, d.dim))).astype(np.float32)
    Y = (U + 0.1 * rng.normal(size=(M, d.dim))).astype(np.float32)
    cfg = lp.cfg
    loss, g = lp.loss_grad(x, U, Y)
    want_loss, want_g = mlp.loss_grad(p, d, x, U, Y, cfg.log_sigma0, cfg.log_sigma_min, cfg.log_sigma_max, True)
    _, plain_g = mlp.loss_grad(p, d, x, U, Y, cfg.log_sigma0, cfg.log_sigma_min, cfg.log_sigma_max, False)
    print(f"  loss gpu {loss:.8f} oracle {want_loss:.8f}")
    assert abs(loss - want_loss) < 1e-5 * abs(want_loss)
    o = d.offsets()
    for name, a, b in [("W1", "w1", "b1"), ("b1", "b1", "w2"), ("W2", "w2", "b2"), ("b2", "b2", "w3"),
                       ("W3", "w3", "b3"), ("b3", "b3", "n")]:
        e = _rel(g[o[a]:o[b]], want_g[o[a]:o[b]])
        ep = _rel(g[o[a]:o[b]], plain_g[o[a]:o[b]])
        print(f"  grad {name}: rel err {e:.2e} (vs plain FP64 {ep:.2e})")
        assert e < 2e-3, name
    lp.close()
    st.close()


def _slot_betas(st, beta_sched):
    perm = st.pt_perm()
    b = np.zeros(st.nrep)
    for pos, r in enumerate(perm):
        b[r] = beta_sched[pos // st.nrep_per_temp]
    return b


# (N, NT, replicas per temperature, FMC mode, hidden width, moves): two small
# systems (2: uniform FMC, quartets in channels 0 and 1) and the learned-move
# benchmark's shape (N = 64 uniform FMC, 32 x 8 replicas, hidden 512).
REPLAY_CASES = [(16, 8, 8, 0, 128, 3), (16, 8, 8, 2, 128, 3), (64, 32, 8, 2, 512, 2)]


@pytest.mark.parametrize("N,NT,nr,fmc,hidden,moves", REPLAY_CASES)
def test_learned_move_replays_on_oracle(pspin, oracle, N, NT, nr, fmc, hidden, moves):
    st = _state(pspin, N=N, nrep=NT * nr, seed=11, fmc=fmc)
    R = st.nrep
    T, beta = oracle.temperatures(NT, 0.4, 2.0, 1)
    st.pt_init(NT, nr, beta, 99)
    st.pt_run(40, 1)
    lp = _lp(pspin, st, hidden=hidden, init_gain=0.3, log_sigma0=math.log(0.08))
    d = mlp.Dims(N=N, hid=lp.hidden)
    g2 = st.g2()
    g4, mask = st.g4()
    seeds, _ = st.rng_info()
    cfg = lp.cfg
    ls_scale = math.log(cfg.scale)
    p = lp.params()
    n_acc = n_close = 0
    for mv in range(moves):
        s0 = st.spins()
        b_r = _slot_betas(st, beta)
        lp.move()
        prop, dbg, Of, Orv = lp.last()
        s1 = st.spins()
        E1, _, _ = st.results()
        for r in range(R):
            a = np.stack([s0[r].real, s0[r].imag], -1).reshape(-1)
            _, _, u = oracle.lp_gauge(N, a)
            want_of, _ = mlp.forward(p, d, np.concatenate([u, [b_r[r]]])[None].astype(np.float32), emulate=True)
            assert np.max(np.abs(Of[r] - want_of[0])) < 5e-3 * max(1.0, np.max(np.abs(want_of)))
            pr, lqf = oracle.lp_propose(N, a, Of[r], cfg.log_sigma0, LSMIN, LSMAX, ls_scale, int(seeds[r]), mv)
            pg = np.stack([prop[r].real, prop[r].imag], -1).reshape(-1)
            assert np.max(np.abs(pr - pg)) < 1e-11, (mv, r)
            assert abs(lqf - dbg[r, 0]) < 1e-8 * max(1.0, abs(lqf))
            e_cur = sum(oracle.energy(N, a, g2, 1, g4, mask))
            e_new = sum(oracle.energy(N, pr, g2, 1, g4, mask))
            assert abs(e_cur - dbg[r, 2]) < 1e-11 * max(1.0, abs(e_cur))
            assert abs(e_new - dbg[r, 3]) < 1e-11 * max(1.0, abs(e_new))
            out = oracle.lp_accept(N, b_r[r], a, pr, e_cur, e_new, Orv[r], cfg.log_sigma0, LSMIN, LSMAX, ls_scale,
                                   lqf, int(seeds[r]), mv)
            assert abs(out[1] - dbg[r, 1]) < 1e-8 * max(1.0, abs(out[1]))
            assert abs(out[4] - dbg[r, 4]) < 1e-7 * max(1.0, abs(out[4]))
            assert np.allclose(out[6:8], dbg[r, 6:8], rtol=1e-12, atol=1e-12)
            _, u_acc = oracle.lp_draws(int(seeds[r]), mv, N)
            if abs(math.log(u_acc) - out[4]) < 1e-6:
                n_close += 1
                continue
            assert out[5] == dbg[r, 5], (mv, r, out, dbg[r])
            acc = bool(dbg[r, 5])
            n_acc += acc
            want_s = prop[r] if acc else s0[r]
            assert np.array_equal(s1[r], want_s)
            assert E1[r] == (dbg[r, 3] if acc else dbg[r, 2])
    print(f"  replayed {moves * R} moves, {n_acc} accepted, {n_close} decided within 1e-6 of the threshold")
    assert 0 < n_acc < moves * R
    assert np.all(lp.move_counters() == moves)
    lp.close()
    st.close()


def test_learned_moves_alone_sample_dirichlet_half_at_beta0(pspin):
    """Intensities x_k/N ~ Beta(1/2, (N-1)/2) under prod|a_k|^-1 on the sphere."""
    N, R = 16, 256
    st = _state(pspin, N=N, nrep=R, seed=3)
    st.set_uniform_beta(0.0)
    lp = _lp(pspin, st, init_gain=0.1, log_sigma0=math.log(0.3))
    d = mlp.Dims(N=N, hid=lp.hidden)
    p = lp.params()
    p[d.offsets()["b3"]] += 0.3  # drift Re a_0 upwards: a wrong density ratio would bias <x_0>
    lp.set_params(p)
    for _ in range(300):
        lp.move()
    sum0 = np.zeros(R)
    sum2 = np.zeros(R)
    n = 0
    for _ in range(1500):
        lp.move()
        if _ % 5 == 0:
            x = np.abs(st.spins()) ** 2 / N
            sum0 += x[:, 0]
            sum2 += (x ** 2).mean(axis=1)
            n += 1
    a, pr = lp.stats()
    acc = a.sum() / pr.sum()
    m0, se0 = (sum0 / n).mean(), (sum0 / n).std(ddof=1) / math.sqrt(R)
    m2, se2 = (sum2 / n).mean(), (sum2 / n).std(ddof=1) / math.sqrt(R)
    want2 = 0.5 * 1.5 / ((N / 2) * (N / 2 + 1))
    flat2 = 2.0 / (N * (N + 1))
    print(f"  acceptance {acc:.3f}; <x0/N> {m0:.5f} +- {se0:.5f} (want {1 / N:.5f}); "
          f"<(x/N)^2> {m2:.6f} +- {se2:.6f} (want {want2:.6f}, flat measure {flat2:.6f})")
    assert 0.03 < acc < 0.97
    assert abs(m0 - 1.0 / N) < 5 * se0
    assert abs(m2 - want2) < 5 * se2
    assert abs(m2 - flat2) > 5 * se2
    lp.close()
    st.close()


def _ladder_means(st, NT, nr, sweeps, every, step):
    """Per-ladder, per-temperature mean energy per site over `sweeps` (step(n) advances n sweeps)."""
    acc = np.zeros((nr, NT))
    n = 0
    for _ in range(sweeps // every):
        step(every)
        E, _, _ = st.results()
        perm = st.pt_perm().reshape(NT, nr)
        acc += (E[perm] / st.N).T
        n += 1
    return acc / n


def test_pt_with_learned_moves_matches_plain_pt(pspin, oracle):
    """Mixing learned moves into PT keeps every temperature's <E/N> (independent ladders)."""
    NT, nr, N = 4, 16, 16
    T, beta = oracle.temperatures(NT, 0.8, 2.5, 1)
    res = []
    for mode in ("plain", "mixed"):
        st = _state(pspin, N=N, nrep=NT * nr, seed=5)
        st.pt_init(NT, nr, beta, 17 if mode == "plain" else 18)
        lp = _lp(pspin, st, init_gain=0.1, log_sigma0=math.log(0.05), seed=41)
        if mode == "plain":
            st.pt_run(500, 1)
            m = _ladder_means(st, NT, nr, 6000, 4, lambda k: st.pt_run(k, 1))
        else:
            lp.run(500, 1, 2)
            m = _ladder_means(st, NT, nr, 6000, 4, lambda k: lp.run(k, 1, 2))
            a, p = lp.stats()
            print("  learned-move acceptance per T:", np.round(a / np.maximum(p, 1), 3))
            assert np.all(a > 0)
        res.append(m)
        lp.close()
        st.close()
    diff = res[0].mean(axis=0) - res[1].mean(axis=0)
    se = np.sqrt(res[0].var(axis=0, ddof=1) / nr + res[1].var(axis=0, ddof=1) / nr)
    z = diff / se
    for t in range(NT):
        print(f"  T={T[t]:.3f}: plain {res[0][:, t].mean():.5f} mixed {res[1][:, t].mean():.5f} z {z[t]:.2f}")
    assert np.all(np.abs(z) < 4.0), z


def test_training_lowers_loss_and_lengthens_moves(pspin, oracle):
    """A network trained on PT snapshots proposes longer accepted moves than the initial one."""
    NT, nr, N = 4, 16, 16
    T, beta = oracle.temperatures(NT, 0.8, 2.5, 1)
    st = _state(pspin, N=N, nrep=NT * nr, seed=9)
    st.pt_init(NT, nr, beta, 5)
    st.pt_run(300, 1)
    lp = _lp(pspin, st, seed=3)
    lp.collect(64, 1, 1)
    assert lp.data_size() == 64

    def esjd(n_moves):
        """Mean squared displacement per move (gauge-invariant |a' a^* |-based)."""
        tot = 0.0
        for _ in range(n_moves):
            s0 = st.spins()
            lp.move()
            s1 = st.spins()
            ov = np.abs(np.sum(s1 * np.conj(s0), axis=1))
            tot += np.mean(2 * N - 2 * ov)
        return tot / n_moves

    before = esjd(40)
    hist = lp.train(300, batch=512, lag=1, lr=1e-3)
    print(f"  loss {hist[0]:.3f} -> {hist[-10:].mean():.3f}")
    assert hist[-10:].mean() < hist[:10].mean() - 1.0
    after = esjd(40)
    print(f"  squared jump per move: initial network {before:.4f}, trained {after:.4f}")
    assert after > before
    lp.close()
    st.close()


def test_training_speeds_energy_decorrelation(pspin, oracle):
    """Energy ESS of chains of learned moves alone (no sweeps, no exchange),
    the trained network against the initial one, on the same state, at the
    benchmark's proposal scale (s = 0.3, mean scale s^2).

    For a stationary energy series, 1 - rho_1 = E[(E' - E)^2] / (2 Var E), and
    the ESS per sample (1 - rho_1) / (1 + rho_1) is close to (1 - rho_1) / 2
    for rho_1 near 1.  Var E is the same equilibrium variance for both
    networks (pooled over both runs), so the ratio of the mean squared energy
    change per move is the ratio of the energy ESS per move.  Training on PT
    snapshots must raise it at every temperature."""
    NT, nr, N = 4, 16, 16
    T, beta = oracle.temperatures(NT, 0.8, 2.5, 1)
    st = _state(pspin, N=N, nrep=NT * nr, seed=13)
    st.pt_init(NT, nr, beta, 9)
    st.pt_run(300, 1)
    lp = _lp(pspin, st, seed=5)
    lp.set_scale(0.3)
    lp.set_mean_scale(0.09)
    lp.collect(64, 1, 1)
    perm = st.pt_perm().reshape(NT, nr)  # moves alone do not exchange: slots stay fixed

    def series(n_moves):
        E = np.zeros((n_moves + 1, NT * nr))
        E[0] = st.results()[0]
        lp.reset_stats()
        for k in range(n_moves):
            lp.move()
            E[k + 1] = st.results()[0]
        a, p = lp.stats()
        return E, a / np.maximum(p, 1)

    E0, acc0 = series(600)
    hist = lp.train(300, batch=512, lag=1, lr=1e-3)
    assert hist[-10:].mean() < hist[:10].mean()
    E1, acc1 = series(600)
    ratio = np.zeros(NT)
    for t in range(NT):
        e0, e1 = E0[:, perm[t]], E1[:, perm[t]]
        var = np.concatenate([e0 - e0.mean(axis=0), e1 - e1.mean(axis=0)]).var()
        j0 = np.mean(np.diff(e0, axis=0) ** 2)
        j1 = np.mean(np.diff(e1, axis=0) ** 2)
        r0, r1 = 1 - j0 / (2 * var), 1 - j1 / (2 * var)
        ratio[t] = ((1 - r1) / (1 + r1)) / ((1 - r0) / (1 + r0))
        print(f"  T={T[t]:.3f}: rho1 initial {r0:.4f} (acc {acc0[t]:.3f}) trained {r1:.4f} "
              f"(acc {acc1[t]:.3f}); energy ESS per move x{ratio[t]:.2f}")
    assert np.all(ratio > 1.3), ratio
    assert ratio.mean() > 1.5, ratio
    lp.close()
    st.close()
