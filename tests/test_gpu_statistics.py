"""Statistical parity of GPU parallel tempering with the reference's logic.

For each of 12 disorder samples the GPU (FP32 midpoint Delta E, GPU exchange)
and the oracle's CPU PT (FP64 sweeps by the reference's 3-type enumeration,
host drand48 exchange, src/mc.cu:66-318 and src/parallel_tempering.cu:273-352)
run on the same disorder from the same initial spins with independent RNG
streams.  Per sample the difference of each observable is formed; over the
samples the mean difference and its jackknife error give z-scores (the
reference's analysis averages over samples with jackknife,
src/analysis_pt.cpp:846-900).  Observables per temperature: <E/N>, the
specific heat Cv = N beta^2 (<e^2> - <e>^2) (src/analysis_pt.cpp:846-900),
<|q|^2> and the histogram P(|q|) in QB bins over [0, 1] (q = sum_k a_k^1
conj(a_k^2) / N between the two replicas of a slot, the overlap of
notes/quantities/observables.tex); the Metropolis acceptance and the
exchange rate per temperature pair.

Systems: N=10 fully connected (the pair-row kernel; per-replica and
ladder-shared pair streams) and N=32 uniform FMC gamma=0.05 (the group
kernel that carries the FMC and comb configurations) with FP32 and with
BF16 hi/lo couplings.  At N=32 the CPU chains visit each site's active
quartets from a table (oracle.cpu_pt(lists=True): the same terms and
arithmetic as the 3-type enumeration, pinned against it in
tests/test_oracle_physics.py).
"""
import concurrent.futures as cf

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NT, NR = 5, 2
BURN, MEAS = 500, 2000
QB = 4  # |q| histogram bins


# name: (N, fmc mode, gamma, sweep kernel, coupling format, pair groups)
SYSTEMS = {
    "fc10": (10, 0, 0.0, 0, 0, 0),
    "fc10_ladder_pairs": (10, 0, 0.0, 0, 0, NR),
    "fmc32_groups_f32": (32, 2, 0.05, 3, 0, NR),
    "fmc32_groups_bf16": (32, 2, 0.05, 3, 1, NR),
}


def _gpu_run(pspin, seed, beta, system):
    N, fmc, gamma, kernel, fmt, groups = SYSTEMS[system]
    beta = np.asarray(beta)
    cfg = pspin.SimConfig.default(N=N, nrep=NT * NR, seed=seed, fmc_mode=fmc, gamma=gamma)
    st = pspin.PSpin(cfg, pspin.DENSE)
    if fmt:
        st.set_coupling_format(fmt)
    if kernel:
        st.set_kernel(kernel, 0)
        assert st.info.kernel == kernel
    st.set_pair_schedule(groups)
    st.pt_init(NT, NR, beta, seed + 5000)
    s0 = st.spins()
    g2 = st.g2()
    g4, mask = st.g4()
    rs, _ = st.rng_info()
    st.pt_run(BURN, 1)
    st.reset_counters()
    st.pt_reset_window()
    e = np.zeros(NT)
    e2 = np.zeros(NT)
    q2 = np.zeros(NT)
    qh = np.zeros((NT, QB))
    n = 0
    for _ in range(MEAS // 2):
        st.pt_run(2, 1)
        E, _, _ = st.results()
        perm = st.pt_perm().reshape(NT, NR)
        z = st.spins()
        e += (E[perm] / N).mean(axis=1)
        e2 += ((E[perm] / N) ** 2).mean(axis=1)
        q = np.sum(z[perm[:, 0]] * np.conj(z[perm[:, 1]]), axis=1) / N
        q2 += np.abs(q) ** 2
        qh[np.arange(NT), np.minimum(QB - 1, (np.abs(q) * QB).astype(int))] += 1
        n += 1
    _, a, p = st.results()
    xa, xp, _, _ = st.pt_exchange_stats()
    st.close()
    cv = N * beta ** 2 * (e2 / n - (e / n) ** 2)
    obs = np.concatenate([e / n, cv, q2 / n, (qh / n).reshape(-1), [a.sum() / p.sum()], xa / xp])
    return obs, dict(s0=s0, g2=g2, g4=g4, mask=mask, rs=rs, h2=st.info.h2_active)


def _cpu_run(oracle, seed, beta, d, system):
    N = SYSTEMS[system][0]
    beta = np.asarray(beta)
    rs = d["rs"] ^ np.uint64(0x9E3779B97F4A7C15)
    out = oracle.cpu_pt(N, NT, NR, d["s0"], d["g2"], d["h2"], d["g4"], d["mask"], beta, rs, BURN + MEAS,
                        seed + 5000 + 17, BURN, qbins=QB, lists=N > 16)
    n = out["n"]
    e = out["e"] / (n * NR)
    cv = N * beta ** 2 * (out["e2"] / (n * NR) - e ** 2)
    q2 = out["q2"] / n
    acc = out["acc"].sum() / out["prop"].sum()
    return np.concatenate([e, cv, q2, (out["qh"] / n).reshape(-1), [acc], out["ex_acc"] / out["ex_prop"]])


@pytest.mark.parametrize("system", list(SYSTEMS))
def test_gpu_pt_matches_reference_logic_statistically(pspin, oracle, system):
    """Pair groups = NR: each ladder's replicas share one pair stream (the
    layout the benchmark uses), the ladders keep independent streams."""
    print(f"[statistics] {system}", flush=True)
    N = SYSTEMS[system][0]
    T, beta = oracle.temperatures(NT, 0.5, 2.0, 1)
    seeds = [1000 + 7 * s for s in range(12)]
    gpu, data = [], []
    for sd in seeds:
        o, d = _gpu_run(pspin, sd, beta, system)
        gpu.append(o)
        data.append(d)
    with cf.ThreadPoolExecutor(max_workers=12) as ex:
        cpu = list(ex.map(lambda a: _cpu_run(oracle, a[0], beta, a[1], system), zip(seeds, data)))
    gpu, cpu = np.array(gpu), np.array(cpu)
    diff = gpu - cpu
    S = len(seeds)
    jk = np.array([np.delete(diff, i, axis=0).mean(axis=0) for i in range(S)])
    se = np.sqrt((S - 1) / S * np.sum((jk - jk.mean(axis=0)) ** 2, axis=0))
    z = diff.mean(axis=0) / np.maximum(se, 1e-12)
    names = ([f"e[T={t:.3f}]" for t in T] + [f"Cv[T={t:.3f}]" for t in T] + [f"q2[T={t:.3f}]" for t in T] +
             [f"P|q|{b}[T={t:.3f}]" for t in T for b in range(QB)] + ["acc"] + [f"ex[{t}]" for t in range(NT - 1)])
    for nm, g, c, zz in zip(names, gpu.mean(axis=0), cpu.mean(axis=0), z):
        print(f"  {nm:14s} gpu {g: .5f}  cpu {c: .5f}  z {zz: .2f}")
    assert np.all(np.abs(z) < 3.5), z
    assert np.mean(np.abs(z) > 2.0) <= 0.25, z


# Sparse (smoothed-cube) systems: N = 24 with N quartets sampled from the
# uniform-FMC survivors (gamma = 0.2), the single-state kernel (LISTS layout,
# cube move) and the batched kernel (all 12 samples in one launch), against
# the oracle's CPU PT of the same sparse systems (src/mc_sparse.cu:63-253
# chains in the reference's PT loop).
SPARSE_N, SPARSE_GAMMA = 24, 0.2


def _sparse_state(pspin, seed, beta):
    cfg = pspin.SimConfig.default(N=SPARSE_N, nrep=NT * NR, seed=seed, fmc_mode=2, gamma=SPARSE_GAMMA)
    st = pspin.PSpin(cfg, pspin.SPARSE)
    st.pt_init(NT, NR, np.asarray(beta), seed + 5000)
    return st


def _observe(N, E, perm, z, acc):
    """One measurement of a PT state: per slot E/N, (E/N)^2, |q|^2 and the |q| bin."""
    perm = perm.reshape(NT, NR)
    e = (E[perm] / N).mean(axis=1)
    e2 = ((E[perm] / N) ** 2).mean(axis=1)
    q = np.sum(z[perm[:, 0]] * np.conj(z[perm[:, 1]]), axis=1) / N
    acc[0] += e
    acc[1] += e2
    acc[2] += np.abs(q) ** 2
    acc[3][np.arange(NT), np.minimum(QB - 1, (np.abs(q) * QB).astype(int))] += 1


def _finish(N, beta, acc, n, a, p, xa, xp):
    e, e2, q2, qh = acc
    cv = N * beta ** 2 * (e2 / n - (e / n) ** 2)
    return np.concatenate([e / n, cv, q2 / n, (qh / n).reshape(-1), [a.sum() / p.sum()], xa / xp])


@pytest.mark.parametrize("kernel", ["single", "batch"])
def test_gpu_sparse_pt_matches_reference_logic_statistically(pspin, oracle, kernel):
    print(f"[statistics] sparse cube N={SPARSE_N} {kernel}", flush=True)
    N = SPARSE_N
    T, beta = oracle.temperatures(NT, 0.5, 2.0, 1)
    seeds = [2000 + 11 * s for s in range(12)]
    data, gpu = [], []
    for sd in seeds:  # disorder, spins and RNG seeds of each sample (the batch builds the same)
        st = _sparse_state(pspin, sd, beta)
        q, gq = st.sparse_quartets()
        rs, _ = st.rng_info()
        data.append(dict(s0=st.spins(), g2=st.g2(), q=q, gq=gq, rs=rs, h2=st.info.h2_active))
        if kernel == "single":
            st.pt_run(BURN, 1)
            st.reset_counters()
            st.pt_reset_window()
            acc = [np.zeros(NT), np.zeros(NT), np.zeros(NT), np.zeros((NT, QB))]
            for _ in range(MEAS // 2):
                st.pt_run(2, 1)
                _observe(N, st.results()[0], st.pt_perm(), st.spins(), acc)
            _, a, p = st.results()
            xa, xp, _, _ = st.pt_exchange_stats()
            gpu.append(_finish(N, beta, acc, MEAS // 2, a, p, xa, xp))
        st.close()
    if kernel == "batch":
        base = pspin.SimConfig.default(fmc_mode=2, gamma=SPARSE_GAMMA)
        b = pspin.PSpinBatch(base, [N] * len(seeds), seeds, NT, NR, beta)
        b.set_resync_interval(0)
        b.run(BURN, 1)
        a0 = [b.results(s)[1:] for s in range(len(seeds))]
        x0 = [b.exchange_stats(s) for s in range(len(seeds))]
        accs = [[np.zeros(NT), np.zeros(NT), np.zeros(NT), np.zeros((NT, QB))] for _ in seeds]
        for _ in range(MEAS // 2):
            b.run(2, 1)
            for s in range(len(seeds)):
                _observe(N, b.results(s)[0], b.perm(s), b.spins(s), accs[s])
        for s in range(len(seeds)):
            _, a, p = b.results(s)
            xa, xp = b.exchange_stats(s)
            gpu.append(_finish(N, beta, accs[s], MEAS // 2, a - a0[s][0], p - a0[s][1], xa - x0[s][0],
                               xp - x0[s][1]))
        b.close()

    def cpu_run(args):
        sd, d = args
        rs = d["rs"] ^ np.uint64(0x9E3779B97F4A7C15)
        out = oracle.cpu_pt_sparse(N, NT, NR, d["s0"], d["g2"], d["h2"], d["q"], d["gq"], beta, rs, BURN + MEAS,
                                   sd + 5000 + 17, BURN, qbins=QB)
        n = out["n"]
        e = out["e"] / (n * NR)
        cv = N * beta ** 2 * (out["e2"] / (n * NR) - e ** 2)
        return np.concatenate([e, cv, out["q2"] / n, (out["qh"] / n).reshape(-1),
                               [out["acc"].sum() / out["prop"].sum()], out["ex_acc"] / out["ex_prop"]])

    with cf.ThreadPoolExecutor(max_workers=12) as ex:
        cpu = list(ex.map(cpu_run, zip(seeds, data)))
    gpu, cpu = np.array(gpu), np.array(cpu)
    diff = gpu - cpu
    S = len(seeds)
    jk = np.array([np.delete(diff, i, axis=0).mean(axis=0) for i in range(S)])
    se = np.sqrt((S - 1) / S * np.sum((jk - jk.mean(axis=0)) ** 2, axis=0))
    z = diff.mean(axis=0) / np.maximum(se, 1e-12)
    names = ([f"e[T={t:.3f}]" for t in T] + [f"Cv[T={t:.3f}]" for t in T] + [f"q2[T={t:.3f}]" for t in T] +
             [f"P|q|{b}[T={t:.3f}]" for t in T for b in range(QB)] + ["acc"] + [f"ex[{t}]" for t in range(NT - 1)])
    for nm, g, c, zz in zip(names, gpu.mean(axis=0), cpu.mean(axis=0), z):
        print(f"  {nm:14s} gpu {g: .5f}  cpu {c: .5f}  z {zz: .2f}")
    assert np.all(np.abs(z) < 3.5), z
    assert np.mean(np.abs(z) > 2.0) <= 0.25, z
