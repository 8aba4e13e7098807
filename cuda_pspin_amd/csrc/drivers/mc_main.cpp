// mc_main.cpp -- drop-in `bin/montecarlo` (fixed-temperature MC, the
// reference's src/montecarlo.cu:23-231): same flags, data/N{N}_NR{nrep}[_S{l}]
// directory, energy_accept.txt with cumulative acceptance, per-replica
// configs/conf_r{r}_iter{s}.bin in the second half, time_per_iter.txt.
#include <sys/prctl.h>
#include <sys/stat.h>

#include <ctime>
#include <string>

#include "../../../include/pspin_mc.hpp"

static double now_s() {
    timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec + t.tv_nsec * 1e-9;
}

int main(int argc, char** argv) {
    SimConfig cfg = parse_args(argc, argv);
    {
        char pname[16];
        snprintf(pname, sizeof(pname), "MC_N%d_NR%d_S%d", cfg.N, cfg.nrep, cfg.label >= 0 ? cfg.label : 0);
        prctl(PR_SET_NAME, pname);
    }
    char dir[128];
    if (cfg.label >= 0) snprintf(dir, sizeof(dir), "data/N%d_NR%d_S%d", cfg.N, cfg.nrep, cfg.label);
    else snprintf(dir, sizeof(dir), "data/N%d_NR%d", cfg.N, cfg.nrep);
    mkdir("data", 0755);
    mkdir(dir, 0755);
    const std::string confdir = std::string(dir) + "/configs";
    mkdir(confdir.c_str(), 0755);
    const std::string efile = std::string(dir) + "/energy_accept.txt";
    FILE* fout = fopen(efile.c_str(), "w");
    if (!fout) {
        fprintf(stderr, "Cannot open %s\n", efile.c_str());
        return 2;
    }
    fprintf(fout, "# iter");
    for (int r = 0; r < cfg.nrep; r++) fprintf(fout, "\tE%d/N\tacc%d", r, r);
    fprintf(fout, "\n");

    const long long nq = n_quartets(cfg.N);
    const long long mem = (long long)cfg.N * cfg.N * 12 + nq * 25 + (long long)cfg.nrep * (cfg.N * 16 + 48);
    size_t free_mem = 0, total_mem = 0;
    char gpu[256] = {0};
    pspin_die_on(pspin_device_info(cfg.dev, &free_mem, &total_mem, gpu, sizeof(gpu)), "device query");
    if (mem > (long long)(free_mem * 0.9)) {
        fprintf(stderr, "ERROR: Not enough GPU memory! Need %.1fMB, have %.1fMB free\n", mem / 1e6, free_mem / 1e6);
        return 1;
    }
    if (cfg.verbose) {
        printf("\n+--------------------------------------------------+\n");
        printf("|             p-Spin 2+4 :: Monte Carlo            |\n");
        printf("+--------------------------------------------------+\n");
        printf("  %-22s %s\n  %-22s %d\n  %-22s %d\n  %-22s %.6f\n  %-22s %.4f\n", "GPU", gpu, "N", cfg.N, "nrep",
               cfg.nrep, "T", cfg.T, "J", cfg.J);
        printf("  %-22s %d\n  %-22s %d\n  %-22s %llu\n", "MC sweeps", cfg.mc_iterations, "save_freq", cfg.save_freq,
               "seed", (unsigned long long)cfg.seed);
        printf("  %-22s %lld\n  %-22s %lld\n\n", "pairs", n_pairs(cfg.N), "quartets", nq);
    }

    MCState st = mc_init(cfg);
    if (cfg.fmc_mode > 0) {
        const char* names[] = {"FC", "comb", "uniform"};
        printf("  %-22s %s (gamma=%.4f)  pairs=%lld/%lld  quartets=%lld/%lld\n", "FMC active", names[cfg.fmc_mode],
               cfg.gamma, st.n_pairs_active, n_pairs(cfg.N), st.n_quart_active, n_g4_total(cfg.N));
        if (FILE* ff = fopen((std::string(dir) + "/frequencies.txt").c_str(), "w")) {
            fprintf(ff, "# i omega_i\n");
            for (int i = 0; i < cfg.N; i++) fprintf(ff, "%d\t%.12f\n", i, st.h_omega[i]);
            fclose(ff);
        }
    }
    std::vector<double> E(cfg.nrep), reim(2 * cfg.N);
    std::vector<long long> acc(cfg.nrep), prop(cfg.nrep);
    std::vector<pspin_complex> spins((size_t)cfg.nrep * cfg.N);
    pspin_die_on(pspin_set_uniform_beta(st.h, 1.0 / cfg.T), "beta");
    const double t0 = now_s();
    int s = 0;
    while (s < cfg.mc_iterations) {
        int next = ((s / cfg.save_freq) + 1) * cfg.save_freq;
        if (next > cfg.mc_iterations) next = cfg.mc_iterations;
        pspin_die_on(pspin_sweep(st.h, next - s), "sweep");
        s = next;
        mc_get_results(st, E.data(), acc.data(), prop.data());
        fprintf(fout, "%d", s);
        for (int r = 0; r < cfg.nrep; r++)
            fprintf(fout, "\t%.8f\t%.5f", E[r] / cfg.N, prop[r] > 0 ? (double)acc[r] / prop[r] : 0.0);
        fprintf(fout, "\n");
        fflush(fout);
        if (s - 1 >= cfg.mc_iterations / 2) {
            mc_get_spins(st, spins.data());
            for (int r = 0; r < cfg.nrep; r++) {
                char cf[512];
                snprintf(cf, sizeof(cf), "%s/conf_r%d_iter%d.bin", confdir.c_str(), r, s);
                if (FILE* fc = fopen(cf, "wb")) {
                    for (int i = 0; i < cfg.N; i++) {
                        reim[2 * i] = spins[(size_t)r * cfg.N + i].x;
                        reim[2 * i + 1] = spins[(size_t)r * cfg.N + i].y;
                    }
                    fwrite(reim.data(), sizeof(double), 2 * cfg.N, fc);
                    fclose(fc);
                }
            }
        }
        if (cfg.verbose) {
            printf("  iter %d/%d:", s, cfg.mc_iterations);
            for (int r = 0; r < cfg.nrep; r++)
                printf("  [%d] E/N=% .3e acc=% .3e", r, E[r] / cfg.N, prop[r] > 0 ? (double)acc[r] / prop[r] : 0.0);
            printf("\n");
        }
    }
    const double elapsed = now_s() - t0, per = elapsed / cfg.mc_iterations;
    printf("\n-- Summary ------\n  %-22s %.3f s\n  %-22s %.4f ms\n  %-22s %d\n  %-22s %d\n\n", "total time", elapsed,
           "time/sweep", per * 1e3, "N", cfg.N, "nrep", cfg.nrep);
    if (FILE* ft = fopen((std::string(dir) + "/time_per_iter.txt").c_str(), "w")) {
        fprintf(ft, "# N nrep mc_iterations total_time_s ms_per_iter\n");
        fprintf(ft, "%d %d %d %.6f %.6f\n", cfg.N, cfg.nrep, cfg.mc_iterations, elapsed, per * 1e3);
        fclose(ft);
    }
    if (cfg.verbose) {
        mc_get_results(st, E.data(), acc.data(), prop.data());
        printf("-- Results ------\n");
        for (int r = 0; r < cfg.nrep; r++)
            printf("  Replica %-12d E/N = % .3e   acc = % .3e\n", r, E[r] / cfg.N,
                   prop[r] > 0 ? (double)acc[r] / prop[r] : 0.0);
        printf("\n");
    }
    fclose(fout);
    mc_free(st);
    return 0;
}
