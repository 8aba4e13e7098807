"""Throughput benchmark: dense spherical parallel tempering on one MI355X.

Workload (the reference's `parallel_tempering -N 64 -NT 32 -nrep 8 -Tmin 0.2
-Tmax 1.6 -log_temp -pt_freq 1 -fmc 0 -seed 42`): N=64 fully connected
(C(64,4) = 635,376 quartets, real couplings, one channel each), 32 geometric
temperatures x 8 ladders = 256 replicas, an exchange after every sweep.
A step is one sweep of all 256 replicas (N/2 = 32 pair proposals each) plus
its exchange round, all on the GPU.

Reports one JSON line: pair-updates/s (whole job, timed by wall clock around
K steps after W warm-up steps, inputs resident in HBM), energy ESS/s at the
coldest slot, the sweep kernel's roofline figure (HIP events around every
sweep launch on the library's stream) and a CPU baseline (the oracle's
OpenMP transcription of the reference sweep, FP64) on a bounded sample.
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


def ess_of_series(series):
    """series [rounds, ladders]: drop the first half, sum n/tau over ladders."""
    half = series.shape[0] // 2
    prod = series[half:]
    ess = 0.0
    taus = []
    for r in range(prod.shape[1]):
        t = tau_int_sokal(prod[:, r])
        taus.append(t)
        if np.isfinite(t):
            ess += prod.shape[0] / t
    return ess, taus


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
    st.pt_init(NT, nr, 1.0 / T, args.seed + 5000)
    st.pt_run(args.lp_burn, 1)
    tuned = st.autotune(8)  # the sweep kernel both runs use; restores the state
    lcfg = P.LPConfig.default(hidden=512, batch_cap=4096, data_cap=args.lp_snapshots, seed=args.seed)
    lp = P.LearnedMove(st, lcfg)
    t0 = time.perf_counter()
    lp.collect(args.lp_snapshots, 1, 1)
    t_collect = time.perf_counter() - t0
    t0 = time.perf_counter()
    hist = lp.train(args.lp_train_steps, batch=4096, lag=1, lr=1e-3)
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
                                f"{args.lp_scale ** 2:.3g}", "blocks": args.lp_blocks, "sweeps_per_block": S,
                     "burn_in_sweeps": args.lp_burn,
                     "sweep_kernel": {"kernel": tuned[0], "block": tuned[1], "trial_ms_per_sweep": tuned[2]}}
    return out


def cpu_baseline(st, N, nrep, betas, seed, n_sweeps):
    import oracle as O

    O.build()
    L = O.lib()
    g2 = np.ascontiguousarray(st.g2(), dtype=np.float64).reshape(-1)
    g4, mask = st.g4()
    s = st.spins()
    spins = np.ascontiguousarray(np.stack([s.real, s.imag], -1), dtype=np.float64).reshape(-1)
    seeds, _ = st.rng_info()
    E = np.zeros(nrep)
    acc = np.zeros(nrep, dtype=np.int64)
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    if os.environ.get("OMP_NUM_THREADS"):
        cores = min(cores, int(os.environ["OMP_NUM_THREADS"]))
    secs = L.or_cpu_baseline_sweeps(N, nrep, spins, g2, int(st.info.h2_active), g4, mask,
                                    np.ascontiguousarray(betas), seeds, 1 << 40, n_sweeps, E, acc, cores)
    pu = n_sweeps * nrep * (N // 2)
    return {"value": pu / secs, "unit": "pair-updates/s", "cores": cores, "kind": "port",
            "sample": f"{n_sweeps} sweep(s) x {nrep} replicas of the same N={N} system "
                      f"({pu} pair-updates, {secs:.1f} s, FP64, reference 3-type enumeration)"}


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


def load_traffic(N, fmc, kernel):
    """HBM bytes per sweep launch measured with rocprofv3 --pmc (profiles/), or None."""
    return load_traffic_key(f"N{N}_fmc{fmc}_k{kernel}")


def load_traffic_key(key):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        rec = json.load(open(path))
    except ValueError:
        return None
    return rec.get(key, {}).get("hbm_bytes_per_launch")


def batch_traffic(sweeps_per_launch):
    """HBM bytes of one batch launch of sweeps_per_launch sweeps, from the
    measured bytes per sweep (the per-sweep draw scratch dominates)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        per = json.load(open(path)).get("batch_256", {}).get("hbm_bytes_per_sweep")
    except (OSError, ValueError):
        return None
    return None if per is None else per * sweeps_per_launch


def splitmix64(state):
    z = (state + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return z ^ (z >> 31)


def run_batch(args):
    """Batched sparse PT: 256 disorder samples, N_s uniform on {24,32,48,64}
    (splitmix64(seed+s) mod 4), cube constraint, uniform FMC gamma=0.05, N
    sampled quartets each, 24 geometric temperatures in [0.2,1.6] x 4 replicas,
    FP64 resync every 1024 sweeps, one launch per 64 sweeps (graph-replayed)."""
    import cuda_pspin_amd as P

    n_samples, NT, nr = args.samples, 24, 4
    Ns = [(24, 32, 48, 64)[splitmix64(args.seed + s) % 4] for s in range(n_samples)]
    seeds = [args.seed + s for s in range(n_samples)]
    T = np.array([1.6 * (0.2 / 1.6) ** (t / (NT - 1)) for t in range(NT)])
    base = P.SimConfig.default(fmc_mode=2, gamma=0.05, dev=0)
    t0 = time.time()
    b = P.PSpinBatch(base, Ns, seeds, NT, nr, 1.0 / T)
    t_init = time.time() - t0
    b.set_resync_interval(1024)
    b.run(args.warmup, 1)
    chunk, trial_ms = b.autotune(512, 1)  # sweeps per launch; restores the batch afterwards
    t0 = time.perf_counter()
    b.run(args.steps, 1)
    wall = time.perf_counter() - t0
    pu = b.pair_updates_per_sweep * args.steps
    # algorithmic bytes of a sparse pair-update: the N stored quartet terms
    # scanned (16 B each: indices + FP64 coupling) plus two g2 rows (8 B each)
    bpu = np.mean([16 * n + 2 * n * 8 for n in Ns])
    achieved = pu / wall * bpu
    out = {"metric": "pair-updates/s (batched sparse PT, 256 samples, 1 MI355X)", "value": pu / wall,
           "unit": "pair-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded disorder and spins)",
           "config": {"workload": "batched sparse cube PT: 256 samples N in {24,32,48,64}, fmc uniform 0.05, "
                                  "NT=24 x nrep=4, exchange every sweep, resync 1024",
                      "samples": n_samples, "chains": b.chains, "N_hist": {str(n): Ns.count(n) for n in (24, 32, 48, 64)},
                      "autotuned_sweeps_per_launch": chunk, "autotune_trial_ms_per_sweep": trial_ms},
           "init_s": t_init,
           "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK, "traffic": batch_traffic(chunk),
                        "note": "terms and amplitudes are LDS-resident; latency-bound per lane"}}
    b.close()
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16384)
    ap.add_argument("--warmup", type=int, default=64)
    ap.add_argument("--N", type=int, default=64)
    ap.add_argument("--NT", type=int, default=32)
    ap.add_argument("--nrep", type=int, default=8)
    ap.add_argument("--fmc", type=int, default=0)
    ap.add_argument("--gamma", type=float, default=0.0)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--kernel", type=int, default=0)
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--coupling", choices=["f32", "bf16"], default="f32")
    ap.add_argument("--Tmin", type=float, default=0.2)
    ap.add_argument("--Tmax", type=float, default=1.6)
    ap.add_argument("--cpu-sweeps", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calib", type=int, default=0, help="only run the counter-calibration copy of this size")
    ap.add_argument("--workload", choices=["dense", "batch", "learned"], default="dense")
    ap.add_argument("--samples", type=int, default=256)
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-vs-plain comparison")
    ap.add_argument("--no-graphs", action="store_true", help="launch kernels one by one (for tracing tools)")
    ap.add_argument("--pair-groups", type=int, default=8,
                    help="replicas r = g (mod G) share a pair stream, one group per XCD and per ladder "
                         "(0: per-replica pairs, as the reference)")
    ap.add_argument("--lp-burn", type=int, default=8192)
    ap.add_argument("--lp-snapshots", type=int, default=8192)
    ap.add_argument("--lp-train-steps", type=int, default=1500)
    ap.add_argument("--lp-scale", type=float, default=0.2)
    ap.add_argument("--lp-every", type=int, default=8)
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

    if args.workload == "batch":
        run_batch(args)
        return
    if args.workload == "learned":
        print(json.dumps({"learned_vs_plain": learned_vs_plain(args)}))
        return

    import cuda_pspin_amd as P

    N, NT, nr = args.N, args.NT, args.nrep
    nrep = NT * nr
    Tmin, Tmax = args.Tmin, args.Tmax
    A = (Tmin / Tmax) ** (1.0 / (NT - 1))
    T = np.array([Tmax * A ** t for t in range(NT)])
    beta = 1.0 / T
    cfg = P.SimConfig.default(N=N, nrep=nrep, seed=args.seed, fmc_mode=args.fmc, gamma=args.gamma, dev=0)
    t0 = time.time()
    st = P.PSpin(cfg, P.DENSE)
    if args.coupling == "bf16":
        st.set_coupling_format(1)
    st.set_pair_schedule(args.pair_groups)
    tuned = None
    if args.kernel == 0 and args.block == 0:
        tuned = st.autotune(8)  # times every kernel / tile candidate on this state
    else:
        st.set_kernel(args.kernel, args.block)
    t_init = time.time() - t0
    st.pt_init(NT, nr, beta, args.seed + 5000)
    st.pt_run(args.warmup, 1)
    st.pt_enable_series(args.steps)
    st.set_profiling(True)
    t0 = time.perf_counter()
    st.pt_run(args.steps, 1)  # synchronises before returning
    wall = time.perf_counter() - t0
    prof = st.profile()
    st.set_profiling(False)
    series = st.pt_series()
    ess, taus = ess_of_series(series)

    pu_per_step = nrep * (N // 2)
    value = args.steps * pu_per_step / wall
    kern_ms = prof["sweep_ms"] / max(prof["n_sweep"], 1)  # per sweep (exchange included when fused)
    sweeps_per_launch = prof["n_sweep"] / max(prof["n_launch"], 1)
    bpu = bytes_per_pair_update(N, st.info.h2_active)
    if args.fmc != 0:
        # FMC systems are stored sparsely: each surviving (site, term) entry is a
        # 4-byte coupling plus 4 bytes of partner indices
        bpu = 2.0 * st.info.n_site_entries / N * 8 + (2 * N * 4 if st.info.h2_active else 0)
    achieved = pu_per_step * bpu / (kern_ms * 1e-3)
    out = {
        "metric": "pair-updates/s + energy ESS/s at T_min, N=64 PT: learned vs plain, 1 MI355X",
        "value": value,
        "unit": "pair-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded disorder and spins)",
        "config": {"workload": f"dense spherical PT N={N} fmc={args.fmc} NT={NT} x nrep={nr} "
                               f"T in [{Tmin},{Tmax}] geometric, exchange every sweep",
                   "N": N, "NT": NT, "nrep": nr, "replicas": nrep, "pt_freq": 1, "fmc": args.fmc,
                   "sweeps_timed": args.steps, "kernel": st.info.kernel, "coupling": args.coupling,
                   "pair_groups": args.pair_groups,
                   "autotuned": None if tuned is None else {"kernel": tuned[0], "block": tuned[1],
                                                             "trial_ms_per_sweep": tuned[2]}},
        "energy_ess_per_s_tmin": ess / wall,
        "energy_tau_int_tmin_sweeps": float(np.nanmean(taus)) if taus else None,
        "sweep_kernel_ms": kern_ms,
        "sweeps_per_launch": sweeps_per_launch,
        "exchange_kernel_ms": (prof["exch_ms"] / prof["n_exch"]) if prof["n_exch"] else None,
        "init_s": t_init,
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": load_traffic(N, args.fmc, st.info.kernel),
                     "bytes_per_pair_update": bpu,
                     "pair_updates_per_launch": int(round(pu_per_step * sweeps_per_launch)),
                     "launch_ms": kern_ms * sweeps_per_launch},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(st, N, nrep, np.repeat(beta, nr), args.seed, args.cpu_sweeps)
    st.close()
    if not args.no_learned:
        out["learned_vs_plain"] = learned_vs_plain(args)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
