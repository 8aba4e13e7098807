"""Throughput benchmark: dense spherical parallel tempering on one MI355X.

Headline workload (the reference's `parallel_tempering -N 64 -NT 32 -nrep 8
-Tmin 0.2 -Tmax 1.6 -log_temp -pt_freq 1 -fmc 0 -seed 42`): N=64 fully
connected (C(64,4) = 635,376 quartets, real couplings, one channel each), 32
geometric temperatures x 8 ladders = 256 replicas, an exchange after every
sweep.  A step is one production launch: a cooperative launch of 32 PT rounds
(a sweep of all 256 replicas, N/2 = 32 pair proposals each, then the
exchange), all on the GPU.

Reports one JSON line: pair-updates/s (whole job, wall clock around exactly K
steps after W warm-up steps, inputs resident in HBM, no profiling), the sweep
kernel's roofline (HIP events around every launch on the library's stream, in
a separate pass of the same launch path), energy ESS/s at the coldest slot
from a 2^16-sweep series, a CPU baseline (the oracle's OpenMP transcription of
the reference sweep, FP64) on a bounded sample, the reference's per-replica
pair schedule, the other configurations (N=32 uniform FMC, N=128 comb with
BF16 hi/lo couplings and its Delta E error against FP64, the batched sparse
workload) and the learned-vs-plain comparison.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
FP32_PEAK = 157.3e12  # MI355X FP32, flop/s
LDS_PEAK = 128 * 256 * 2.4e9  # 128 B/clk per CU, 256 CUs, 2.4 GHz
ROUNDS = 32  # PT rounds per cooperative launch (the library's production launch)


def n_inv(N):
    return 2 * (N - 2) * (N - 3) * (N - 4) // 6 + (N - 2) * (N - 3) // 2


def bytes_per_pair_update(N, h2_active, b_q=4, b_2=4):
    """Algorithmic bytes of one pair-update: each touched quartet's coupling
    (b_q bytes, real FP32) plus a 1-byte channel code, plus two g2 rows."""
    return n_inv(N) * (b_q + 1) + (2 * N * b_2 if h2_active else 0)


def tau_int_sokal(x, c=5.0):
    """Integrated autocorrelation time with Sokal's automatic window."""
    x = np.asarray(x, dtype=np.float64)
    n = len(x)
    if n < 8:
        return float("nan")
    x = x - x.mean()
    f = np.fft.rfft(x, 2 * n)
    acf = np.fft.irfft(f * np.conj(f))[:n]
    if acf[0] <= 0:
        return float("nan")
    rho = acf / acf[0]
    tau = 1.0
    for M in range(1, n):
        tau += 2.0 * rho[M]
        if M >= c * tau:
            break
    return max(tau, 1.0)


def ess_blocks(series):
    """series [rounds, ladders] of an equilibrated block (no burn-in dropped):
    ESS summed over ladders and the mean tau_int."""
    ess, taus = 0.0, []
    for r in range(series.shape[1]):
        t = tau_int_sokal(series[:, r])
        taus.append(t)
        if np.isfinite(t):
            ess += series.shape[0] / t
    return ess, float(np.nanmean(taus))


def tau_e_crossing(series):
    """Reference-style tau_E: first lag where the normalised autocorrelation
    falls below 1/e (src/analysis_pt.cpp:2694-2760), averaged over ladders."""
    out = []
    for r in range(series.shape[1]):
        x = series[:, r] - series[:, r].mean()
        n = len(x)
        f = np.fft.rfft(x, 2 * n)
        acf = np.fft.irfft(f * np.conj(f))[:n]
        if acf[0] <= 0:
            continue
        rho = acf / acf[0]
        below = np.nonzero(rho < math.exp(-1.0))[0]
        if len(below):
            out.append(float(below[0]))
    return float(np.mean(out)) if out else None


def learned_vs_plain(args):
    """Energy ESS/s at the coldest slot with and without the learned global
    move, on the N=64 uniform-FMC (gamma 0.05) system, 32 geometric
    temperatures in [0.2,1.6] x 8 ladders, exchange every sweep.  The
    network (2N+1 -> 512 -> 512 -> 4N) is trained on the GPU on 8192
    snapshots of all 256 replicas (2^16 configurations per temperature) from
    the plain run; the learned run mixes one move per `every` sweeps.  Plain
    and learned production blocks alternate on the same equilibrated state, so
    each rate has a block-to-block error."""
    import cuda_pspin_amd as P

    N, NT, nr = 64, 32, 8
    R = NT * nr
    T = np.array([args.Tmax * (args.Tmin / args.Tmax) ** (t / (NT - 1)) for t in range(NT)])
    st = P.PSpin(P.SimConfig.default(N=N, nrep=R, seed=args.seed, fmc_mode=2, gamma=0.05, dev=0), P.DENSE)
    if args.no_graphs:
        st.set_profiling(True)  # per-launch events instead of captured graphs
    if args.direct_launches:
        st.set_launch_mode(P.LAUNCH_DIRECT)  # PT rounds as ordinary launches (same chains)
    st.pt_init(NT, nr, 1.0 / T, args.seed + 5000)
    st.pt_run(args.lp_burn, 1)
    tuned = st.autotune(8)  # the sweep kernel both runs use; restores the state
    lcfg = P.LPConfig.default(hidden=512, batch_cap=4096, data_cap=args.lp_snapshots, seed=args.seed)
    lp = P.LearnedMove(st, lcfg)
    t0 = time.perf_counter()
    lp.collect(args.lp_snapshots, 1, 1)
    t_collect = time.perf_counter() - t0
    t0 = time.perf_counter()
    hist = lp.train(args.lp_train_steps, batch=4096, lag=args.lp_lag, lr=1e-3)
    t_train = time.perf_counter() - t0
    lp.set_scale(args.lp_scale)
    lp.set_mean_scale(args.lp_scale ** 2)
    S = args.lp_block
    rec = {"plain": [], "learned": []}
    series_all = {"plain": [], "learned": []}
    lp.reset_stats()
    for b in range(args.lp_blocks):
        for mode in ("plain", "learned"):
            st.pt_enable_series(S)
            t0 = time.perf_counter()
            if mode == "plain":
                st.pt_run(S, 1)
            else:
                lp.run(S, 1, args.lp_every)
            w = time.perf_counter() - t0
            ser = st.pt_series()
            ess, tau = ess_blocks(ser)
            rec[mode].append((w, ess, tau))
            series_all[mode].append(ser)
    acc, prop = lp.stats()
    lp.close()
    st.close()
    out = {}
    for mode in ("plain", "learned"):
        w = np.array([x[0] for x in rec[mode]])
        e = np.array([x[1] for x in rec[mode]])
        rate = e / w
        out[mode] = {"pair_updates_per_s": float(S * R * (N // 2) / w.mean()), "ms_per_sweep": float(w.mean() / S * 1e3),
                     "energy_ess_per_s_tmin": float(rate.mean()),
                     "energy_ess_per_s_tmin_stderr": float(rate.std(ddof=1) / math.sqrt(len(rate))) if len(rate) > 1 else None,
                     "tau_int_sweeps": float(np.mean([x[2] for x in rec[mode]])),
                     "tau_e_crossing_sweeps": tau_e_crossing(np.concatenate(series_all[mode], axis=0))}
    prod_learned = float(np.sum([x[0] for x in rec["learned"]]))
    ess_learned = float(np.sum([x[1] for x in rec["learned"]]))
    out["learned"]["energy_ess_per_s_tmin_incl_training"] = ess_learned / (prod_learned + t_collect + t_train)
    out["learned"]["move_acceptance_by_T"] = {f"{T[t]:.3f}": float(acc[t] / max(prop[t], 1)) for t in (0, NT // 2, NT - 1)}
    flops_step = 2 * 4096 * ((2 * N + 1 + 63) // 64 * 64 * 512 + 512 * 512 + 512 * 4 * N) * 3
    out["training"] = {"snapshots": args.lp_snapshots, "configs_per_T": args.lp_snapshots * nr,
                       "collect_s": t_collect, "train_s": t_train, "steps": args.lp_train_steps, "batch": 4096,
                       "loss_first": float(np.mean(hist[:20])), "loss_last": float(np.mean(hist[-20:])),
                       "step_ms": t_train / args.lp_train_steps * 1e3,
                       "gemm_tflops": flops_step / (t_train / args.lp_train_steps) / 1e12,
                       "mfma_frac_bf16_dense": flops_step / (t_train / args.lp_train_steps) / 2.5e15}
    r = np.array([x[1] / x[0] for x in rec["learned"]]) / np.array([x[1] / x[0] for x in rec["plain"]])
    out["ess_ratio_learned_over_plain"] = float(r.mean())
    out["ess_ratio_stderr"] = float(r.std(ddof=1) / math.sqrt(len(r))) if len(r) > 1 else None
    out["config"] = {"workload": "dense spherical PT N=64 fmc uniform gamma=0.05, NT=32 x nrep=8, T in "
                                 f"[{args.Tmin},{args.Tmax}] geometric, exchange every sweep",
                     "learned": f"one move per {args.lp_every} sweeps, scale {args.lp_scale}, mean scale "
                                f"{args.lp_scale ** 2:.3g}, training lag {args.lp_lag}", "blocks": args.lp_blocks,
                     "sweeps_per_block": S,
                     "burn_in_sweeps": args.lp_burn,
                     "sweep_kernel": {"kernel": tuned[0], "block": tuned[1], "trial_ms_per_sweep": tuned[2]}}
    return out


def cpu_baseline(st, N, NT, nr, beta_sched, n_warm=16, n_sweeps=256):
    """The oracle's OpenMP transcription of the reference sweep (FP64, the
    reference's 3-type enumeration with its index decode), one replica chain
    per thread, on a bounded sample: as many replicas of the benchmark state
    as there are cores, spread over the temperature slots, n_warm untimed
    sweeps and then n_sweeps timed ones."""
    import oracle as O

    O.build()
    L = O.lib()
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    if os.environ.get("OMP_NUM_THREADS"):
        cores = min(cores, int(os.environ["OMP_NUM_THREADS"]))
    R = NT * nr
    n = max(1, min(cores, R))
    perm = st.pt_perm()
    slots = np.linspace(0, R - 1, n).round().astype(int)
    reps = perm[slots]
    betas = np.ascontiguousarray([beta_sched[s // nr] for s in slots], dtype=np.float64)
    g2 = np.ascontiguousarray(st.g2(), dtype=np.float64).reshape(-1)
    g4, mask = st.g4()
    s = st.spins()[reps]
    spins = np.ascontiguousarray(np.stack([s.real, s.imag], -1), dtype=np.float64).reshape(-1)
    seeds = np.ascontiguousarray(st.rng_info()[0][reps])
    E = np.zeros(n)
    acc = np.zeros(n, dtype=np.int64)
    h2 = int(st.info.h2_active)
    L.or_cpu_baseline_sweeps(N, n, spins, g2, h2, g4, mask, betas, seeds, 1 << 40, n_warm, E, acc, cores)
    secs = L.or_cpu_baseline_sweeps(N, n, spins, g2, h2, g4, mask, betas, seeds, (1 << 40) + n_warm, n_sweeps, E,
                                    acc, cores)
    pu = n_sweeps * n * (N // 2)
    return {"value": pu / secs, "unit": "pair-updates/s", "cores": cores, "kind": "port",
            "sample": f"{n} replicas of the benchmark state (one per thread, spread over the {NT} temperatures), "
                      f"{n_warm} warm-up sweeps then {n_sweeps} timed sweeps: {pu} pair-updates in {secs:.1f} s, "
                      "FP64, reference 3-type enumeration, no exchange"}


def copy_calibration(nbytes):
    """A plain device-to-device copy of known size (hipMemcpy, HIP's own blit
    kernel), run under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE to calibrate
    the counters against bytes actually moved."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(src), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(src, 1, C.c_size_t(nbytes)) == 0
    for _ in range(4):
        assert hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"copy_calibration_bytes": nbytes, "copies": 4}))


def load_traffic(key):
    """HBM bytes per sweep launch of the named workload, measured with
    rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE and scaled by the copy calibration
    (profiles/traffic.json), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        rec = json.load(open(path))
    except (OSError, ValueError):
        return None
    return rec.get(key, {}).get("hbm_bytes_per_launch")


def splitmix64(state):
    z = (state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return z ^ (z >> 31)


def flops_per_pair_update(N, n_quart_active, h2_active):
    """FP32 flops of one pair-update in the reference's algorithm: 24 per
    active quartet holding i0 or j0 (three complex products of the
    differential and the Re(g .) multiply-add, src/mc.cu:182-275), i.e.
    n_inv(N) 24 when every quartet is active, plus 8 per g2 entry of the two
    rows (src/mc.cu:160-180)."""
    c4 = N * (N - 1) * (N - 2) * (N - 3) // 24
    return 24.0 * n_quart_active * n_inv(N) / c4 + (2 * N * 4 if h2_active else 0)


def geometric_betas(NT, Tmin, Tmax):
    A = (Tmin / Tmax) ** (1.0 / (NT - 1))
    T = np.array([Tmax * A ** t for t in range(NT)])
    return T, 1.0 / T


def dense_pt(P, N, NT, nr, Tmin, Tmax, fmc, gamma, coupling, pair_groups, seed, steps, warmup, prof_launches):
    """Dense PT on one state: autotuned sweep kernel, W warm-up steps, then
    EXACTLY K timed steps, a step being one production launch: a cooperative
    launch of ROUNDS PT rounds (a sweep of every replica, then the exchange),
    with no profiling.  A separate pass of prof_launches launches records HIP
    events around every launch on the library's stream (the same launch
    path) for the kernel's duration.  Returns the state and the figures."""
    R = NT * nr
    T, beta = geometric_betas(NT, Tmin, Tmax)
    cfg = P.SimConfig.default(N=N, nrep=R, seed=seed, fmc_mode=fmc, gamma=gamma, dev=0)
    t0 = time.time()
    st = P.PSpin(cfg, P.DENSE)
    if coupling == "bf16":
        st.set_coupling_format(1)
    st.set_pair_schedule(pair_groups)
    tuned = st.autotune(8)  # times every kernel / tile candidate on this state and restores it
    st.pt_init(NT, nr, beta, seed + 5000)
    t_init = time.time() - t0
    st.pt_run(warmup * ROUNDS, 1)
    t0 = time.perf_counter()
    st.pt_run(steps * ROUNDS, 1)  # synchronises before returning
    wall = time.perf_counter() - t0
    st.set_profiling(True)
    st.pt_run(prof_launches * ROUNDS, 1)
    prof = st.profile()
    st.set_profiling(False)
    launch_ms = prof["sweep_ms"] / max(prof["n_launch"], 1)
    per_launch = prof["n_sweep"] / max(prof["n_launch"], 1)
    pu_sweep = R * (N // 2)
    fpu = flops_per_pair_update(N, st.info.n_quart_active, st.info.h2_active)
    achieved = pu_sweep * per_launch * fpu / (launch_ms * 1e-3)
    out = {"value": steps * ROUNDS * pu_sweep / wall, "unit": "pair-updates/s", "steps": steps, "warmup": warmup,
           "ms_per_step": wall / steps * 1e3, "ms_per_sweep": wall / (steps * ROUNDS) * 1e3,
           "config": {"N": N, "NT": NT, "nrep": nr, "replicas": R, "T": [Tmin, Tmax], "pt_freq": 1, "fmc": fmc,
                      "gamma": gamma, "coupling": coupling, "pair_groups": pair_groups, "seed": seed,
                      "rounds_per_step": ROUNDS, "sweeps_timed": steps * ROUNDS,
                      "quartets_active": st.info.n_quart_active, "h2_active": st.info.h2_active,
                      "autotuned": {"kernel": tuned[0], "block": tuned[1], "trial_ms_per_sweep": tuned[2]}},
           "init_s": t_init,
           "sweep_launch": {"path": "cooperative" if per_launch == ROUNDS else "direct", "sweeps_per_launch": per_launch,
                            "launch_ms": launch_ms, "launches_timed": prof["n_launch"]},
           "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": FP32_PEAK / 1e12, "unit": "TFLOP/s",
                        "frac": achieved / FP32_PEAK,
                        "traffic": load_traffic(f"N{N}_fmc{fmc}_k{st.info.kernel}_{coupling}"),
                        "flops_per_pair_update": fpu, "pair_updates_per_launch": int(pu_sweep * per_launch),
                        "launch_ms": launch_ms,
                        "hbm_algorithmic_bytes_per_pair_update": bytes_per_pair_update(N, st.info.h2_active)
                        if fmc == 0 else None}}
    return st, out


def energy_ess(st, nr, burn, prod):
    """Energy series of the coldest slot (the replica holding T_min after each
    exchange), `burn` sweeps dropped, `prod` recorded; per ladder
    ESS = prod / tau_int (Sokal window, c = 5), summed over ladders, over the
    wall time of the recorded sweeps.  Null unless every ladder's series spans
    at least 50 tau_int."""
    st.pt_run(burn, 1)
    st.pt_enable_series(prod)
    t0 = time.perf_counter()
    st.pt_run(prod, 1)
    w = time.perf_counter() - t0
    ser = st.pt_series()
    taus = np.array([tau_int_sokal(ser[:, r]) for r in range(nr)])
    per = ser.shape[0] / taus
    ok = bool(np.all(np.isfinite(taus)) and per.min() >= 50)
    ess = float(per.sum())
    return {"energy_ess_per_s_tmin": ess / w if ok else None,
            "energy_ess_per_s_tmin_stderr": float(math.sqrt(nr) * per.std(ddof=1) / w) if ok and nr > 1 else None,
            "tau_int_sweeps_by_ladder": [float(t) for t in taus],
            "series_over_tau_min": float(per.min()), "burn_in_sweeps": burn, "recorded_sweeps": int(ser.shape[0]),
            "recorded_wall_s": w, "tau_e_crossing_sweeps": tau_e_crossing(ser)}


def delta_e_error(st, n_sweeps=2):
    """FP32 Delta E of every proposal of n_sweeps logged sweeps against the
    library's FP64 full energies of the configurations before and after the
    proposal (E64(after) - E64(before); the FP64 difference is exact to
    ~1e-13 here).  The replay applies the logged accepted amplitudes, and
    must end on the kernel's own final spins."""
    R, N = st.nrep, st.N
    z = st.spins().copy()
    st.enable_step_log(n_sweeps)
    st.sweep(n_sweeps)
    log, n = st.step_log()
    z_end = st.spins()
    steps = log.reshape(R, n * (N // 2))
    rows = np.arange(R)

    def energies(zz):
        st.set_spins(zz)
        e2, e4 = st.split_energies()
        return e2 + e4

    e_cur = energies(z)
    err, d64 = [], []
    for k in range(steps.shape[1]):
        rec = steps[:, k]
        zp = z.copy()
        zp[rows, rec["i0"]] = rec["ai"][:, 0] + 1j * rec["ai"][:, 1]
        zp[rows, rec["j0"]] = rec["aj"][:, 0] + 1j * rec["aj"][:, 1]
        e_p = energies(zp)
        d = e_p - e_cur
        err.append(rec["dE"].astype(np.float64) - d)
        d64.append(d)
        acc = rec["acc"].astype(bool)
        z[acc] = zp[acc]
        e_cur = np.where(acc, e_p, e_cur)
    err, d64 = np.abs(np.concatenate(err)), np.concatenate(d64)
    st.set_spins(z_end)
    rms64 = float(np.sqrt(np.mean(d64 ** 2)))
    return {"proposals": int(err.size), "max_abs": float(err.max()), "rms_abs": float(np.sqrt(np.mean(err ** 2))),
            "rms_dE64": rms64, "max_abs_over_rms_dE64": float(err.max()) / rms64,
            "rms_abs_over_rms_dE64": float(np.sqrt(np.mean(err ** 2))) / rms64,
            "replay_ends_on_kernel_spins": bool(np.array_equal(z, z_end))}


def run_batch(P, seed, steps, warmup, n_samples=256, chunk=0):
    """Batched sparse PT: 256 disorder samples, N_s uniform on {24,32,48,64}
    (splitmix64(seed+s) mod 4), cube constraint, uniform FMC gamma=0.05, N
    sampled quartets each, 24 geometric temperatures in [0.2,1.6] x 4 replicas,
    FP64 resync every 1024 sweeps, sweeps per launch autotuned, graph-replayed
    launches.  A step is one sweep of the whole batch plus its exchanges."""
    NT, nr = 24, 4
    Ns = [(24, 32, 48, 64)[splitmix64(seed + s) % 4] for s in range(n_samples)]
    seeds = [seed + s for s in range(n_samples)]
    T, beta = geometric_betas(NT, 0.2, 1.6)
    base = P.SimConfig.default(fmc_mode=2, gamma=0.05, dev=0)
    t0 = time.time()
    b = P.PSpinBatch(base, Ns, seeds, NT, nr, beta)
    t_init = time.time() - t0
    b.set_resync_interval(1024)
    b.run(warmup, 1)
    if chunk:
        b.set_chunk(chunk)
        trial_ms = None
    else:
        chunk, trial_ms = b.autotune(512, 1)  # sweeps per launch; restores the batch afterwards
    t0 = time.perf_counter()
    b.run(steps, 1)
    wall = time.perf_counter() - t0
    pu = b.pair_updates_per_sweep * steps
    # LDS bytes a pair-update must read: the two sites' H4 entries (4 per site
    # on average: N quartets x 4 sites / N; 8 B each) with three FP64 complex
    # amplitudes each (48 B), and their H2 entries (2 gamma (N-1) partners on
    # average; 8 B entry + one 16 B amplitude)
    bpu = float(np.mean([2 * (4 * (8 + 48) + 2 * 0.05 * (n - 1) * (8 + 16)) for n in Ns]))
    achieved = pu / wall * bpu
    traffic = load_traffic("batch_256")
    out = {"value": pu / wall, "unit": "pair-updates/s", "steps": steps, "warmup": warmup,
           "ms_per_step": wall / steps * 1e3,
           "config": {"workload": "batched sparse cube PT: 256 samples N in {24,32,48,64}, fmc uniform 0.05, "
                                  "NT=24 x nrep=4, exchange every sweep, FP64 resync every 1024 sweeps",
                      "samples": n_samples, "chains": b.chains, "N_hist": {str(n): Ns.count(n) for n in (24, 32, 48, 64)},
                      "autotuned_sweeps_per_launch": chunk, "autotune_trial_ms_per_sweep": trial_ms},
           "init_s": t_init,
           "roofline": {"bound": "lds", "achieved": achieved / 1e9, "peak": LDS_PEAK / 1e9, "unit": "GB/s",
                        "frac": achieved / LDS_PEAK, "bytes_per_pair_update": bpu,
                        "traffic": None if traffic is None else traffic / load_traffic_sweeps("batch_256") * chunk,
                        "note": "terms and amplitudes are LDS-resident; the chain steps are VALU-issue-bound"}}
    b.close()
    return out


def load_traffic_sweeps(key):
    try:
        return json.load(open(os.path.join(ROOT, "profiles", "traffic.json")))[key]["sweeps_per_launch"]
    except (OSError, ValueError, KeyError):
        return 1


HEADLINE = "pair-updates/s + energy ESS/s at T_min, N=64 PT: learned vs plain, 1 MI355X"

# The configurations besides the headline (cfg 2): N, NT, nrep, Tmin, Tmax,
# fmc, gamma, coupling, iteration count of the reference invocation.
CONFIGS = {
    "cfg1": dict(N=32, NT=20, nr=4, Tmin=0.2, Tmax=1.6, fmc=2, gamma=0.05, coupling="f32", sweeps=1 << 14,
                 workload="dense spherical PT N=32 fmc uniform gamma=0.05, NT=20 x nrep=4, T in [0.2,1.6] "
                          "geometric, exchange every sweep, 2^14 sweeps"),
    "cfg4": dict(N=128, NT=16, nr=4, Tmin=0.3, Tmax=1.6, fmc=1, gamma=1.0, coupling="bf16", sweeps=1 << 14,
                 workload="dense spherical PT N=128 comb fmc gamma=1 (H2 off), BF16 hi/lo couplings, NT=16 x "
                          "nrep=4, T in [0.3,1.6] geometric, exchange every sweep, 2^14 sweeps"),
}


def config_line(P, name, seed):
    c = CONFIGS[name]
    steps = c["sweeps"] // ROUNDS
    st, out = dense_pt(P, c["N"], c["NT"], c["nr"], c["Tmin"], c["Tmax"], c["fmc"], c["gamma"], c["coupling"],
                       c["nr"], seed, steps, 8, 16)
    out["config"]["workload"] = c["workload"]
    if name == "cfg4":
        out["delta_e_vs_fp64"] = delta_e_error(st, 2)
    st.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=512, help="timed steps (one cooperative launch of 32 PT rounds each)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--Tmin", type=float, default=0.2)
    ap.add_argument("--Tmax", type=float, default=1.6)
    ap.add_argument("--ess-sweeps", type=int, default=1 << 16,
                    help="sweeps of the energy-ESS run (the first half is burn-in)")
    ap.add_argument("--cpu-sweeps", type=int, default=256)
    ap.add_argument("--cpu-warmup", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-vs-plain comparison")
    ap.add_argument("--no-configs", action="store_true",
                    help="headline only: skip the per-replica pair schedule, cfg 1 / 3 / 4 and the learned comparison")
    ap.add_argument("--copy-calib", type=int, default=0, help="only run the counter-calibration copy of this size")
    ap.add_argument("--workload", choices=["dense", "batch", "learned", "cfg1", "cfg4"], default="dense")
    ap.add_argument("--no-graphs", action="store_true", help="learned workload: launch kernels one by one")
    ap.add_argument("--direct-launches", action="store_true",
                    help="learned workload: PT rounds as ordinary launches, not cooperative ones (identical chains; "
                         "for profiler runs, which end cleanly only without cooperative launches)")
    ap.add_argument("--batch-chunk", type=int, default=0,
                    help="batch workload: sweeps per launch (0: autotuned); a fixed value gives every launch the "
                         "same size, for counter runs")
    ap.add_argument("--lp-burn", type=int, default=8192)
    ap.add_argument("--lp-snapshots", type=int, default=8192)
    ap.add_argument("--lp-train-steps", type=int, default=1500)
    ap.add_argument("--lp-scale", type=float, default=0.2)
    ap.add_argument("--lp-every", type=int, default=8)
    ap.add_argument("--lp-lag", type=int, default=1,
                    help="training pairs (snapshot t, replica r) -> (t + lag, r): the move learns to predict lag sweeps ahead")
    ap.add_argument("--lp-blocks", type=int, default=4)
    ap.add_argument("--lp-block", type=int, default=16384)
    args = ap.parse_args()
    if args.copy_calib:
        copy_calibration(args.copy_calib)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(independent ladders would shard with no communication). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import cuda_pspin_amd as P

    if args.workload == "batch":
        print(json.dumps(run_batch(P, args.seed, args.steps * ROUNDS, args.warmup * ROUNDS, chunk=args.batch_chunk)))
        return
    if args.workload == "learned":
        print(json.dumps({"learned_vs_plain": learned_vs_plain(args)}))
        return
    if args.workload in CONFIGS:
        print(json.dumps(config_line(P, args.workload, args.seed)))
        return

    # cfg 2: N=64 fully connected, 32 x 8, exchange every sweep, ladder-shared pair streams
    N, NT, nr = 64, 32, 8
    T, beta = geometric_betas(NT, args.Tmin, args.Tmax)
    st, head = dense_pt(P, N, NT, nr, args.Tmin, args.Tmax, 0, 0.0, "f32", nr, args.seed, args.steps, args.warmup,
                        min(args.steps, 64))
    cfg = head["config"]
    cfg["workload"] = (f"dense spherical PT N={N} fully connected (fmc 0), NT={NT} x nrep={nr}, T in "
                       f"[{args.Tmin},{args.Tmax}] geometric, exchange every sweep; a step is one cooperative "
                       f"launch of {ROUNDS} PT rounds")
    out = {"metric": HEADLINE, "value": head["value"], "unit": "pair-updates/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded disorder and spins)", "config": cfg,
           "ms_per_sweep": head["ms_per_sweep"], "init_s": head["init_s"], "sweep_launch": head["sweep_launch"],
           "roofline": head["roofline"]}
    if args.ess_sweeps > 0:
        ess = energy_ess(st, nr, args.ess_sweeps // 2, args.ess_sweeps - args.ess_sweeps // 2)
        out["energy_ess_per_s_tmin"] = ess.pop("energy_ess_per_s_tmin")
        out["energy_ess"] = ess
    else:
        out["energy_ess_per_s_tmin"] = None
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(st, N, NT, nr, beta, args.cpu_warmup, args.cpu_sweeps)
    st.close()
    # the reference's schedule: every replica draws its own pairs (src/mc.cu:129-141)
    if args.no_configs:
        print(json.dumps(out))
        return
    st, ref = dense_pt(P, N, NT, nr, args.Tmin, args.Tmax, 0, 0.0, "f32", 0, args.seed, args.steps, args.warmup,
                       min(args.steps, 16))
    st.close()
    out["per_replica_pairs"] = {k: ref[k] for k in ("value", "ms_per_step", "ms_per_sweep", "sweep_launch")}
    out["per_replica_pairs"]["roofline_frac"] = ref["roofline"]["frac"]
    out["other_configs"] = {name: config_line(P, name, args.seed) for name in CONFIGS}
    out["other_configs"]["cfg3"] = run_batch(P, args.seed, 1 << 15, 256)
    if not args.no_learned:
        out["learned_vs_plain"] = learned_vs_plain(args)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
