"""Pins of the learned-move oracle (oracle/pspin_oracle.cpp or_lp_*, oracle/mlp.py).

The reference has no learned proposal, so nothing of it can be compared with
reference outputs; every function is pinned by mathematics instead:

* log J_n(alpha) -- the closed form at alpha = 0, the three-term recurrence
  for alpha >= 0 and scipy adaptive quadrature for alpha < 0;
* the projected-Gaussian density -- the uniform special case (mu = 0, equal
  sigma gives 1/|S^(d-1)|), normalisation over the sphere by Monte Carlo, and
  the distribution of radially projected Gaussian draws;
* the MLP backward -- central finite differences and torch autograd (FP64);
* the MH move as a whole -- its CPU chain must sample the stationary law of
  the pair-rotation sweep: Dirichlet(1/2) intensities at beta = 0 and the
  pair-rotation chain's mean energy at beta > 0.
"""
import math

import numpy as np
import pytest

from oracle import mlp

LS0, LSMIN, LSMAX = math.log(0.05), -9.0, 1.0


def _jn_scaled_recurrence(n, alpha):
    """K_m = J_m exp(-alpha^2/2) by the forward recurrence (stable for alpha >= 0)."""
    k0 = math.sqrt(math.pi / 2) * math.erfc(-alpha / math.sqrt(2))
    k1 = math.exp(-alpha * alpha / 2) + alpha * k0
    if n == 0:
        return k0
    for m in range(2, n + 1):
        k0, k1 = k1, alpha * k1 + (m - 1) * k0
    return k1


@pytest.mark.parametrize("n", [1, 2, 7, 31, 127, 255])
def test_log_jn_closed_form_at_zero(oracle, n):
    want = (n - 1) / 2 * math.log(2) + math.lgamma((n + 1) / 2)
    assert abs(oracle.log_jn(n, 0.0) - want) < 1e-12 * max(1, abs(want))


@pytest.mark.parametrize("n", [1, 5, 63, 127])
@pytest.mark.parametrize("alpha", [0.3, 2.0, 11.0, 37.0])
def test_log_jn_recurrence(oracle, n, alpha):
    want = math.log(_jn_scaled_recurrence(n, alpha)) + alpha * alpha / 2
    assert abs(oracle.log_jn(n, alpha) - want) < 1e-11 * max(1, abs(want))


@pytest.mark.parametrize("n", [3, 127])
@pytest.mark.parametrize("alpha", [-60.0, -7.5, -0.4, 90.0, 400.0])
def test_log_jn_quadrature(oracle, n, alpha):
    from scipy import integrate
    sq = math.sqrt(alpha * alpha + 4 * n)
    ss = 0.5 * (alpha + sq) if alpha >= 0 else 2 * n / (sq - alpha)
    fs = n * math.log(ss) - ss * ss / 2 + alpha * ss
    f = lambda s: math.exp(n * math.log(s) - s * s / 2 + alpha * s - fs) if s > 0 else 0.0
    w = 1 / math.sqrt(n / ss ** 2 + 1)
    val, err = integrate.quad(f, max(0.0, ss - 60 * w), ss + 60 * (1 + w), points=[ss], limit=400,
                              epsabs=0, epsrel=1e-13)
    assert abs(oracle.log_jn(n, alpha) - (fs + math.log(val))) < 1e-10


@pytest.mark.parametrize("d", [2, 8, 128])
def test_density_uniform_special_case(oracle, d):
    rng = np.random.default_rng(d)
    y = rng.normal(size=d)
    y /= np.linalg.norm(y)
    want = math.lgamma(d / 2) - math.log(2) - d / 2 * math.log(math.pi)
    for s in (-3.0, 0.0, 1.5):
        got = oracle.lp_log_density(y, np.zeros(d), np.full(d, s))
        assert abs(got - want) < 1e-10


@pytest.mark.parametrize("d", [4, 6])
def test_density_normalises_on_sphere(oracle, d):
    rng = np.random.default_rng(7 + d)
    mu = rng.normal(size=d)
    ls = np.log(rng.uniform(0.6, 1.4, size=d))
    n = 60000
    y = rng.normal(size=(n, d))
    y /= np.linalg.norm(y, axis=1, keepdims=True)
    p = np.exp([oracle.lp_log_density(v, mu, ls) for v in y])
    area = 2 * math.pi ** (d / 2) / math.gamma(d / 2)
    est = p.mean() * area
    se = p.std() * area / math.sqrt(n)
    assert abs(est - 1.0) < 5 * se, (est, se)


def test_density_matches_projected_draws(oracle):
    """Mean of a test function under projected draws equals its density integral."""
    d = 4
    rng = np.random.default_rng(3)
    mu = np.array([1.2, -0.3, 0.4, 0.9])
    ls = np.log(np.array([0.7, 1.1, 0.5, 0.9]))
    z = mu + np.exp(ls) * rng.normal(size=(200000, d))
    y = z / np.linalg.norm(z, axis=1, keepdims=True)
    f = lambda v: v[..., 0] ** 2 + 0.5 * v[..., 1] * v[..., 3]
    mc = f(y).mean()
    mc_se = f(y).std() / math.sqrt(len(y))
    u = rng.normal(size=(200000, d))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    p = np.exp([oracle.lp_log_density(v, mu, ls) for v in u])
    area = 2 * math.pi ** (d / 2) / math.gamma(d / 2)
    iw = f(u) * p * area
    assert abs(iw.mean() - mc) < 5 * math.hypot(mc_se, iw.std() / math.sqrt(len(u)))


def _batch(seed, N=8, M=5, hid=64, gain=1.0):
    rng = np.random.default_rng(seed)
    d = mlp.Dims(N=N, hid=hid)
    p = mlp.init_params(d, rng, gain=gain).astype(np.float64)
    x = rng.normal(size=(M, d.in_dim))
    U = x[:, :d.dim] + 0.1 * rng.normal(size=(M, d.dim))
    Y = U + 0.3 * rng.normal(size=(M, d.dim))
    return d, p, x, U, Y


def test_mlp_backward_finite_differences():
    d, p, x, U, Y = _batch(11)
    # keep log-scales away from the clamp so the loss is smooth
    loss, g = mlp.loss_grad(p, d, x, U, Y, 0.0, -50.0, 50.0, emulate=False)
    rng = np.random.default_rng(5)
    nz = [i for i in range(len(p)) if i % (d.in_pad) < d.in_dim or i >= d.offsets()["b1"]]
    idx = rng.choice(nz, size=40, replace=False)
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        lp_, _ = mlp.loss_grad(pp, d, x, U, Y, 0.0, -50.0, 50.0, emulate=False)
        lm_, _ = mlp.loss_grad(pm, d, x, U, Y, 0.0, -50.0, 50.0, emulate=False)
        fd = (lp_ - lm_) / (2 * h)
        assert abs(fd - g[i]) < 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_mlp_matches_torch_autograd():
    torch = pytest.importorskip("torch")
    d, p, x, U, Y = _batch(12)
    loss, g = mlp.loss_grad(p, d, x, U, Y, LS0, -3.0, 3.0, emulate=False)
    P = mlp.unpack(p, d)
    T = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in P.items()}
    X0 = torch.zeros((x.shape[0], d.in_pad), dtype=torch.float64)
    X0[:, :d.in_dim] = torch.tensor(x)
    h = torch.nn.functional.silu(X0 @ T["W1"].T + T["b1"])
    h = torch.nn.functional.silu(h @ T["W2"].T + T["b2"])
    O = h @ T["W3"].T + T["b3"]
    D = d.dim
    mu = torch.tensor(U) + O[:, :D]
    ls = torch.clamp(O[:, D:] + LS0, -3.0, 3.0)
    lt = ((0.5 * ((torch.tensor(Y) - mu) * torch.exp(-ls)) ** 2 + ls).sum()) / x.shape[0]
    lt.backward()
    assert abs(lt.item() - loss) < 1e-12 * abs(loss)
    gt = mlp.pack({k: T[k].grad.numpy() for k in T}, d)
    assert np.allclose(g, gt, rtol=1e-10, atol=1e-13)


def test_bf16_rounding_is_nearest_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.0e-3], dtype=np.float32)
    got = mlp.bf16(vals)
    # 1 + 2^-8 is a tie -> even (1.0); 1 + 3*2^-8 is a tie -> even (1 + 2^-6)
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == 1.0 + 2 ** -6 and got[3] == 1.0
    assert got[4] == -2.5
    assert abs(got[5] - 3.0e-3) <= 3.0e-3 * 2 ** -8


# ---------------------------------------------------------------------------
# CPU chains of the learned move: its stationary law
# ---------------------------------------------------------------------------
def _lp_chain(oracle, N, beta, a0, p, d, n_moves, seed, energy, scale, ls0):
    a = np.array(a0, dtype=np.float64)
    E = energy(a)
    xs, es, acc = [], [], 0
    ls_ = math.log(scale)
    for mv in range(n_moves):
        _, _, u = oracle.lp_gauge(N, a)
        Of, _ = mlp.forward(p, d, np.concatenate([u, [beta]])[None], emulate=False)
        prop, lqf = oracle.lp_propose(N, a, Of[0].astype(np.float32), ls0, LSMIN, LSMAX, ls_, seed, mv)
        _, _, up = oracle.lp_gauge(N, prop)
        Or, _ = mlp.forward(p, d, np.concatenate([up, [beta]])[None], emulate=False)
        En = energy(prop)
        out = oracle.lp_accept(N, beta, a, prop, E, En, Or[0].astype(np.float32), ls0, LSMIN, LSMAX, ls_, lqf,
                               seed, mv)
        if out[5]:
            a, E = prop, En
            acc += 1
        xs.append(a[0::2] ** 2 + a[1::2] ** 2)
        es.append(E)
    return np.array(xs), np.array(es), acc / n_moves


def _batch_mean_se(v, nb=40):
    b = np.array_split(np.asarray(v), nb)
    m = np.array([x.mean() for x in b])
    return m.mean(), m.std(ddof=1) / math.sqrt(nb)


def test_learned_chain_beta0_dirichlet_half(oracle):
    """At beta = 0 the target is prod |a_k|^-1 on the sphere: x_k/N ~ Beta(1/2, (N-1)/2).

    The network is random and strongly state dependent, and its mean output
    pushes Re a_0 up, so a wrong proposal-density ratio biases <x_0> and a
    wrong stationary weight biases <(x/N)^2>."""
    N = 4
    rng = np.random.default_rng(21)
    d = mlp.Dims(N=N, hid=32)
    p = mlp.init_params(d, rng, gain=1.0).astype(np.float64)
    p[d.offsets()["b3"]] += 1.0
    a0 = rng.normal(size=2 * N)
    a0 *= math.sqrt(N) / np.linalg.norm(a0)
    xs, _, acc = _lp_chain(oracle, N, 0.0, a0, p, d, 30000, 99, lambda a: 0.0, 1.0, math.log(0.6))
    assert 0.05 < acc < 0.95, acc
    xs = xs[2000:] / N
    m0, se0 = _batch_mean_se(xs[:, 0])
    assert abs(m0 - 1.0 / N) < 4.5 * se0, (m0, se0)
    m2, se = _batch_mean_se((xs ** 2).mean(axis=1))
    want = 0.5 * 1.5 / (2.0 * 3.0)  # Beta(1/2, 3/2) second moment
    flat = 1.0 * 2.0 / (4.0 * 5.0)  # Beta(1, 3): the flat measure
    assert abs(m2 - want) < 4.5 * se, (m2, want, se)
    assert abs(m2 - flat) > 4 * se


def test_learned_chain_matches_pair_rotation_chain(oracle):
    """At beta > 0 the learned-only chain and the pair-rotation chain agree on <E>."""
    N = 4
    rng = np.random.default_rng(4)
    g2 = rng.normal(size=(N, N))
    g2 = (g2 + g2.T) / 2
    g4 = rng.normal(size=oracle.n_quartets(N))
    mask = np.ones(oracle.n_quartets(N), dtype=np.uint8)
    energy = lambda a: sum(oracle.energy(N, a, g2, 1, g4, mask))
    beta = 1.3
    a0 = rng.normal(size=2 * N)
    a0 *= math.sqrt(N) / np.linalg.norm(a0)
    d = mlp.Dims(N=N, hid=32)
    p = mlp.init_params(d, np.random.default_rng(8), gain=0.5).astype(np.float64)
    _, es, acc = _lp_chain(oracle, N, beta, a0, p, d, 24000, 7, energy, 1.0, math.log(0.35))
    assert 0.05 < acc < 0.95, acc
    m_l, se_l = _batch_mean_se(es[2000:])
    # pair-rotation reference chain: independent runs at the same beta
    runs = []
    for k in range(8):
        nr = 4
        out = oracle.cpu_pt(N, 1, nr, np.tile(a0, nr), g2, 1, g4, mask, np.array([beta]),
                            oracle.replica_seeds(700 + k, nr), 20000, 1, 1000)
        runs.append(out["e"][0] / (out["n"] * nr) * N)
    m_p, se_p = np.mean(runs), np.std(runs, ddof=1) / math.sqrt(len(runs))
    z = (m_l - m_p) / math.hypot(se_l, se_p)
    assert abs(z) < 4.5, (m_l, se_l, m_p, se_p, z)
