// pt_main.cpp -- drop-in `bin/parallel_tempering` (and, built with
// -DPSPIN_PT_SPARSE, `bin/parallel_tempering_sparse`): the reference's PT
// drivers (src/parallel_tempering.cu:33-529, src/parallel_tempering_sparse.cu)
// with the same flags, exit codes, run directory and output files.
//
// The sweeps and the replica exchange both run on the GPU (pspin_pt_run);
// the host only wakes at save points.  Exchange decisions are the
// reference's (same order, same srand48(seed+5000) drand48 stream).
#include <sys/prctl.h>
#include <sys/stat.h>

#include <cstring>
#include <ctime>
#include <string>

#include "../../../include/pspin_mc.hpp"

#ifdef PSPIN_PT_SPARSE
#define PT_PREFIX "PTS"
#define PT_TITLE "    p-Spin 2+4 :: Sparse Parallel Tempering       "
#define PT_NAME "Sparse PT"
#define PT_PNAME "PTS_N%d_R%d_S%d"
#else
#define PT_PREFIX "PT"
#define PT_TITLE "         p-Spin 2+4 :: Parallel Tempering         "
#define PT_NAME "Parallel Tempering"
#define PT_PNAME "PT_N%d_NR%d_S%d"
#endif

static void box_top() { printf("+--------------------------------------------------+\n"); }
static void box_title(const char* s) { printf("| %s |\n", s); }
static void box_sec(const char* s) { printf("-- %s ------\n", s); }

static double now_s() {
    timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec + t.tv_nsec * 1e-9;
}

int main(int argc, char** argv) {
    // PT flags first; everything else goes to parse_args (:35-56).
    double Tmax = -1.0, Tmin = -1.0;
    int NT = -1, pt_freq = -1, log_temp = 0;
    std::vector<char*> rest{argv[0]};
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        bool has = i + 1 < argc;
        if (a == "-Tmax" && has) Tmax = atof(argv[++i]);
        else if (a == "-Tmin" && has) Tmin = atof(argv[++i]);
        else if (a == "-NT" && has) NT = atoi(argv[++i]);
        else if (a == "-pt_freq" && has) pt_freq = atoi(argv[++i]);
        else if (a == "-log_temp") log_temp = 1;
        else rest.push_back(argv[i]);
    }
    if (Tmax < 0 || Tmin < 0 || NT < 2 || pt_freq < 1) {
        fprintf(stderr, "%s requires: -Tmax t -Tmin t -NT nt -pt_freq p\n", PT_NAME);
        fprintf(stderr,
                "Usage: %s -N n -Tmax t -Tmin t -NT nt -pt_freq p -iter k [-nrep n] [-J j] [-seed s] "
                "[-save_freq f] [-label l] [-dev d] [-fmc 0|1|2] [-gamma g] [-verbose [0|1|2]]\n",
                argv[0]);
        return 1;
    }
    if (Tmin > Tmax) {
        fprintf(stderr, "Error: Tmin (%.4f) > Tmax (%.4f)\n", Tmin, Tmax);
        return 1;
    }
    SimConfig cfg = parse_args((int)rest.size(), rest.data());
    const int nrep_per_temp = cfg.nrep;
    const int total = NT * nrep_per_temp;
    cfg.nrep = total;
    cfg.T = Tmax;
    {
        char pname[16];
        snprintf(pname, sizeof(pname), PT_PNAME, cfg.N, nrep_per_temp, cfg.label >= 0 ? cfg.label : 0);
        prctl(PR_SET_NAME, pname);
    }

    // Temperature schedule, t = 0 hottest (:93-109).
    std::vector<double> T(NT), beta(NT);
    const double A = (NT > 1) ? pow(Tmin / Tmax, 1.0 / (NT - 1)) : 1.0;
    const double dT = (NT > 1) ? (Tmax - Tmin) / (NT - 1) : 0.0;
    for (int t = 0; t < NT; t++) {
        T[t] = log_temp ? Tmax * pow(A, t) : Tmax - t * dT;
        beta[t] = 1.0 / T[t];
    }

    const int label = cfg.label >= 0 ? cfg.label : 0;
    char dir[256];
    make_run_dir(dir, sizeof(dir), "data", PT_PREFIX, cfg.N, cfg.alpha, cfg.J, cfg.J0, cfg.alpha0, NT, nrep_per_temp,
                 label);
    mkdir("data", 0755);
    mkdir(dir, 0755);
    auto path = [&](const char* f) { return std::string(dir) + "/" + f; };

    if (FILE* fp = fopen(path("params.txt").c_str(), "w")) {
        fprintf(fp, "N %d\nJ %.10g\nJ0 %.10g\nalpha %.10g\nalpha0 %.10g\nNT %d\nnrep %d\nlabel %d\nseed %lu\n", cfg.N,
                cfg.J, cfg.J0, cfg.alpha, cfg.alpha0, NT, nrep_per_temp, label, (unsigned long)cfg.seed);
        fprintf(fp, "nsweeps %d\nTmin %.10g\nTmax %.10g\n", cfg.mc_iterations, Tmin, Tmax);
        fclose(fp);
    }
    FILE* fout = fopen(path("energy_accept.txt").c_str(), "w");
    FILE* fex = fopen(path("exchanges.txt").c_str(), "w");
    FILE* fes = fopen(path("energy_split.txt").c_str(), "w");
    if (!fout || !fex || !fes) {
        fprintf(stderr, "Cannot open output files\n");
        return 2;
    }
    FILE* fconf = fopen(path("configs.bin").c_str(), "wb");
    if (!fconf) {
        fprintf(stderr, "Cannot open configs.bin\n");
        return 2;
    }
    {
        int hdr[3] = {cfg.N, NT, nrep_per_temp};
        fwrite(hdr, sizeof(int), 3, fconf);
    }
    fprintf(fout, "# sweep\tTidx\tT");
    for (int r = 0; r < nrep_per_temp; r++) fprintf(fout, "\tE%d/N\tacc%d", r, r);
    fprintf(fout, "\n");
    fprintf(fex, "# sweep\tTidx\tT_high\tT_low\tn_acc\tn_prop\trate\n");
    fprintf(fes, "# sweep\tTidx\tT");
    for (int r = 0; r < nrep_per_temp; r++) fprintf(fes, "\tE2_%d/N\tE4_%d/N", r, r);
    fprintf(fes, "\n");

    // Memory pre-check with this library's storage (FP64 disorder for the
    // energy kernels plus the FP32 sweep tables).
    const long long nq = n_quartets(cfg.N);
    long long mem = (long long)cfg.N * cfg.N * 12 + (long long)total * (cfg.N * 16 + 48);
#ifdef PSPIN_PT_SPARSE
    mem += nq * 9 + (long long)cfg.N * 48;
#else
    mem += nq * 9 + nq * 16;
#endif
    size_t free_mem = 0, total_mem = 0;
    char gpu[256] = {0};
    pspin_die_on(pspin_device_info(cfg.dev, &free_mem, &total_mem, gpu, sizeof(gpu)), "device query");
    if (mem > (long long)(free_mem * 0.9)) {
        fprintf(stderr, "ERROR: Not enough GPU memory! Need %.1fMB, have %.1fMB free\n", mem / 1e6, free_mem / 1e6);
        return 1;
    }

    printf("\n");
    box_top();
    box_title(PT_TITLE);
    box_top();
    if (cfg.verbose >= 1) {
        printf("  %-22s %s\n", "GPU", gpu);
        printf("  %-22s %d\n", "N", cfg.N);
#ifdef PSPIN_PT_SPARSE
        printf("  %-22s smoothed cube (sum|a|^4 = N)\n", "constraint");
#endif
        printf("  %-22s %d\n  %-22s %d\n  %-22s %d\n", "NT", NT, "nrep/T", nrep_per_temp, "total replicas", total);
        printf("  %-22s %.6f\n  %-22s %.6f\n", "Tmax", Tmax, "Tmin", Tmin);
        if (log_temp) printf("  %-22s geometric (A=%.8f)\n", "T schedule", A);
        else printf("  %-22s linear (dT=%.8f)\n", "T schedule", dT);
        printf("  %-22s %d\n  %-22s %d\n  %-22s %d\n", "MC sweeps", cfg.mc_iterations, "pt_freq", pt_freq, "save_freq",
               cfg.save_freq);
        printf("  %-22s %llu\n", "seed", (unsigned long long)cfg.seed);
        printf("  %-22s %.4f\n  %-22s %.4f\n  %-22s %.4f\n  %-22s %.4f\n", "J", cfg.J, "J0", cfg.J0, "alpha",
               cfg.alpha, "alpha0", cfg.alpha0);
        printf("  %-22s %lld\n  %-22s %lld\n", "pairs", n_pairs(cfg.N), "quartets", nq);
        printf("  %-22s %.1f MB (%.0f MB free / %.0f MB)\n\n", "memory", mem / 1e6, free_mem / 1e6, total_mem / 1e6);
    }

#ifdef PSPIN_PT_SPARSE
    MCStateSparse st = mc_sparse_init(cfg);
#else
    MCState st = mc_init(cfg);
#endif
    if (cfg.fmc_mode > 0) {
        const char* names[] = {"FC", "comb", "uniform"};
        printf("  %-22s %s (gamma=%.6f)  pairs=%lld/%lld  quartets=%lld/%lld\n", "FMC active", names[cfg.fmc_mode],
               cfg.gamma, st.n_pairs_active, n_pairs(cfg.N), st.n_quart_active, n_g4_total(cfg.N));
        if (FILE* ff = fopen(path("frequencies.txt").c_str(), "w")) {
            fprintf(ff, "# i omega_i\n");
            for (int i = 0; i < cfg.N; i++) fprintf(ff, "%d\t%.12f\n", i, st.h_omega[i]);
            fclose(ff);
        }
    }
    if (FILE* ft = fopen(path("temperatures.txt").c_str(), "w")) {
        fprintf(ft, "# Tidx T beta\n");
        for (int t = 0; t < NT; t++) fprintf(ft, "%d\t%.12f\t%.12f\n", t, T[t], beta[t]);
        fclose(ft);
    }
#ifdef PSPIN_PT_SPARSE
    {  // quartets.txt (src/parallel_tempering_sparse.cu:277-296)
        std::vector<int> q((size_t)5 * st.n_quartets);
        std::vector<double> g(st.n_quartets);
        pspin_die_on(pspin_get_sparse_quartets(st.h, q.data(), g.data()), "quartets");
        if (FILE* fq = fopen(path("quartets.txt").c_str(), "w")) {
            fprintf(fq, "# idx i j k l ch Re(g)\n");
            for (int s = 0; s < st.n_quartets; s++)
                fprintf(fq, "%d\t%d\t%d\t%d\t%d\t%d\t%.12e\n", s, q[5 * s], q[5 * s + 1], q[5 * s + 2], q[5 * s + 3],
                        q[5 * s + 4], g[s]);
            fclose(fq);
        }
    }
#endif

    pspin_die_on(pspin_pt_init(st.h, NT, nrep_per_temp, beta.data(), (long)cfg.seed + 5000), "pt_init");
    std::vector<double> E(total), e2(total), e4(total), reim(2 * cfg.N);
    std::vector<long long> acc(total), prop(total);
    std::vector<int> perm(total);
    std::vector<pspin_complex> spins((size_t)total * cfg.N);
    std::vector<long long> xa_w(NT - 1), xp_w(NT - 1), xa_t(NT - 1), xp_t(NT - 1);

    if (cfg.verbose >= 1) box_sec("Running");
    const double t0 = now_s();
    int s = 0;  // sweeps completed
    while (s < cfg.mc_iterations) {
        // Run on the GPU up to the next save point (:355).
        int next = ((s / cfg.save_freq) + 1) * cfg.save_freq;
        if (next > cfg.mc_iterations) next = cfg.mc_iterations;
        pspin_die_on(pspin_pt_run(st.h, next - s, pt_freq), "pt_run");
        s = next;
        const int sw = s;  // the reference's s + 1 at its save point
        mc_get_results(st, E.data(), acc.data(), prop.data());
        pspin_die_on(pspin_pt_get_perm(st.h, perm.data()), "perm");
        pspin_die_on(pspin_pt_get_exchange_stats(st.h, xa_w.data(), xp_w.data(), xa_t.data(), xp_t.data()), "stats");
        for (int t = 0; t < NT; t++) {
            fprintf(fout, "%d\t%d\t%.8f", sw, t, T[t]);
            for (int r = 0; r < nrep_per_temp; r++) {
                const int ph = perm[t * nrep_per_temp + r];
                const double a = prop[ph] > 0 ? (double)acc[ph] / prop[ph] : 0.0;
                fprintf(fout, "\t%.8f\t%.5f", E[ph] / cfg.N, a);
            }
            fprintf(fout, "\n");
        }
        fflush(fout);
        for (int t = 0; t < NT - 1; t++) {
            const double rate = xp_w[t] > 0 ? (double)xa_w[t] / xp_w[t] : 0.0;
            fprintf(fex, "%d\t%d\t%.8f\t%.8f\t%lld\t%lld\t%.5f\n", sw, t, T[t], T[t + 1], xa_w[t], xp_w[t], rate);
        }
        fflush(fex);
        pspin_die_on(pspin_pt_reset_window(st.h), "reset window");
        mc_compute_split_energies(st, e2.data(), e4.data());
        for (int t = 0; t < NT; t++) {
            fprintf(fes, "%d\t%d\t%.8f", sw, t, T[t]);
            for (int r = 0; r < nrep_per_temp; r++) {
                const int ph = perm[t * nrep_per_temp + r];
                fprintf(fes, "\t%.8f\t%.8f", e2[ph] / cfg.N, e4[ph] / cfg.N);
            }
            fprintf(fes, "\n");
        }
        fflush(fes);
        pspin_die_on(pspin_reset_counters(st.h), "reset counters");
        if (sw - 1 >= cfg.mc_iterations / 2) {  // second half (:399-416)
            mc_get_spins(st, spins.data());
            fwrite(&sw, sizeof(int), 1, fconf);
            for (int t = 0; t < NT; t++)
                for (int r = 0; r < nrep_per_temp; r++) {
                    const pspin_complex* rp = &spins[(size_t)perm[t * nrep_per_temp + r] * cfg.N];
                    for (int i = 0; i < cfg.N; i++) {
                        reim[2 * i] = rp[i].x;
                        reim[2 * i + 1] = rp[i].y;
                    }
                    fwrite(reim.data(), sizeof(double), 2 * cfg.N, fconf);
                }
            fflush(fconf);
        }
        if (cfg.verbose >= 2) {
            const int tc = NT - 1, tm = NT / 2, th = 0;
            auto exr = [&](int t) { return (t >= 0 && t < NT - 1 && xp_t[t] > 0) ? (double)xa_t[t] / xp_t[t] : 0.0; };
            auto mcr = [&](int t) {
                double sum = 0;
                for (int r = 0; r < nrep_per_temp; r++) {
                    const int ph = perm[t * nrep_per_temp + r];
                    sum += prop[ph] > 0 ? (double)acc[ph] / prop[ph] : 0.0;
                }
                return sum / nrep_per_temp;
            };
            printf("  %6d  E/N[%d]=% .3e ex=%.2e mc=%.2e  E/N[%d]=% .3e ex=%.2e mc=%.2e  E/N[%d]=% .3e ex=%.2e mc=%.2e"
                   "  [%.2e s/it]\n",
                   sw, tc, E[perm[tc * nrep_per_temp]] / cfg.N, exr(tc - 1), mcr(tc), tm,
                   E[perm[tm * nrep_per_temp]] / cfg.N, exr(tm), mcr(tm), th, E[perm[th * nrep_per_temp]] / cfg.N,
                   exr(th), mcr(th), (now_s() - t0) / sw);
        } else if (cfg.verbose == 1) {
            printf("  sweep %d/%d:  E_hot/N=% .3e  E_cold/N=% .3e\n", sw, cfg.mc_iterations, E[perm[0]] / cfg.N,
                   E[perm[(NT - 1) * nrep_per_temp]] / cfg.N);
        }
    }
    const double elapsed = now_s() - t0;
    printf("\n");
    box_sec("Summary");
    printf("  %-22s %.3f s\n", "total time", elapsed);
    printf("  %-22s %.4f ms\n", "time/sweep", elapsed / cfg.mc_iterations * 1e3);
    printf("  %-22s %d\n  %-22s %d\n\n", "sweeps", cfg.mc_iterations, "total replicas", total);
    box_sec("Exchange Rates");
    for (int t = 0; t < NT - 1; t++) {
        const double rate = xp_t[t] > 0 ? (double)xa_t[t] / xp_t[t] : 0.0;
        printf("  T[%d]-T[%d]  (%.4f-%.4f)  %lld/%lld = %.4f\n", t, t + 1, T[t], T[t + 1], xa_t[t], xp_t[t], rate);
    }
    printf("\n");
    if (FILE* ft = fopen(path("time.txt").c_str(), "w")) {
#ifdef PSPIN_PT_SPARSE
        // the sparse driver adds the number of sampled quartets
        // (src/parallel_tempering_sparse.cu:505-517)
        fprintf(ft,
                "# N NT nrep_per_temp total_replicas mc_iterations pt_freq save_freq Tmax Tmin n_sparse total_time_s "
                "ms_per_sweep\n");
        fprintf(ft, "%d %d %d %d %d %d %d %.6f %.6f %d %.6f %.6f\n", cfg.N, NT, nrep_per_temp, total,
                cfg.mc_iterations, pt_freq, cfg.save_freq, Tmax, Tmin, st.n_quartets, elapsed,
                elapsed / cfg.mc_iterations * 1e3);
#else
        fprintf(ft,
                "# N NT nrep_per_temp total_replicas mc_iterations pt_freq save_freq Tmax Tmin total_time_s "
                "ms_per_sweep\n");
        fprintf(ft, "%d %d %d %d %d %d %d %.6f %.6f %.6f %.6f\n", cfg.N, NT, nrep_per_temp, total, cfg.mc_iterations,
                pt_freq, cfg.save_freq, Tmax, Tmin, elapsed, elapsed / cfg.mc_iterations * 1e3);
#endif
        fclose(ft);
    }
    fclose(fout);
    fclose(fex);
    fclose(fes);
    fclose(fconf);
    mc_free(st);
    return 0;
}
