// sa_main.cpp -- drop-in `bin/simulated_annealing` (the reference's
// src/simulated_annealing.cu:30-314): -Tmax -Tmin -ntemp plus the common
// flags; -iter is an exponent (2^iter sweeps per temperature); geometric
// cooling; acceptance windows reset per temperature and per save;
// data/SA_N{N}_NR{nrep}_S{l}/ with energy_accept.txt, configs/conf_r{r}_T{T}.bin
// for the second half of the temperatures, time.txt.
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
    double Tmax = -1.0, Tmin = -1.0;
    int ntemp = -1;
    std::vector<char*> rest{argv[0]};
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        bool has = i + 1 < argc;
        if (a == "-Tmax" && has) Tmax = atof(argv[++i]);
        else if (a == "-Tmin" && has) Tmin = atof(argv[++i]);
        else if (a == "-ntemp" && has) ntemp = atoi(argv[++i]);
        else rest.push_back(argv[i]);
    }
    if (Tmax < 0 || Tmin < 0 || ntemp < 1) {
        fprintf(stderr, "Simulated Annealing requires: -Tmax tmax -Tmin tmin -ntemp nt\n");
        fprintf(stderr,
                "Usage: %s -N n -Tmax tmax -Tmin tmin -ntemp nt -iter k [-nrep n] [-J j] [-seed s] "
                "[-save_freq f] [-label l] [-dev d] [-fmc 0|1|2] [-gamma g] [-verbose [0|1|2]]\n",
                argv[0]);
        return 1;
    }
    if (Tmin > Tmax) {
        fprintf(stderr, "Error: Tmin (%.4f) > Tmax (%.4f)\n", Tmin, Tmax);
        return 1;
    }
    SimConfig cfg = parse_args((int)rest.size(), rest.data());
    const int per_temp = 1 << cfg.mc_iterations;
    {
        char pname[16];
        snprintf(pname, sizeof(pname), "SA_N%d_NR%d_S%d", cfg.N, cfg.nrep, cfg.label >= 0 ? cfg.label : 0);
        prctl(PR_SET_NAME, pname);
    }
    const int label = cfg.label >= 0 ? cfg.label : 0;
    char dir[256];
    snprintf(dir, sizeof(dir), "data/SA_N%d_NR%d_S%d", cfg.N, cfg.nrep, label);
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
    fprintf(fout, "# T\tsweep");
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
    printf("\n+--------------------------------------------------+\n");
    printf("|        p-Spin 2+4 :: Simulated Annealing         |\n");
    printf("+--------------------------------------------------+\n");
    if (cfg.verbose >= 2) printf("  %-22s %s\n  %-22s %d\n  %-22s %d\n", "GPU", gpu, "N", cfg.N, "nrep", cfg.nrep);

    cfg.T = Tmax;
    MCState st = mc_init(cfg);
    if (cfg.fmc_mode > 0) {
        const char* names[] = {"FC", "comb", "uniform"};
        printf("  %-22s %s (gamma=%.6f)  pairs=%lld/%lld  quartets=%lld/%lld\n", "FMC active", names[cfg.fmc_mode],
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
    const double A = (ntemp > 1) ? pow(Tmin / Tmax, 1.0 / (ntemp - 1)) : 1.0;
    printf("  %-22s %.8f\n", "schedule (A)", A);
    printf("-- Running ------\n");
    const double t0 = now_s();
    for (int step = 0; step < ntemp; step++) {
        const double Tt = Tmax * pow(A, step);
        pspin_die_on(pspin_set_uniform_beta(st.h, 1.0 / Tt), "beta");
        pspin_die_on(pspin_reset_counters(st.h), "reset");
        int s = 0;
        while (s < per_temp) {
            int next = ((s / cfg.save_freq) + 1) * cfg.save_freq;
            if (next > per_temp) next = per_temp;
            pspin_die_on(pspin_sweep(st.h, next - s), "sweep");
            s = next;
            mc_get_results(st, E.data(), acc.data(), prop.data());
            fprintf(fout, "%.8f\t%d", Tt, s);
            for (int r = 0; r < cfg.nrep; r++)
                fprintf(fout, "\t%.8f\t%.5f", E[r] / cfg.N, prop[r] > 0 ? (double)acc[r] / prop[r] : 0.0);
            fprintf(fout, "\n");
            fflush(fout);
            if (cfg.verbose >= 2) {
                printf("  %6d", s);
                for (int r = 0; r < cfg.nrep; r++)
                    printf("  % .3e  % .3e", E[r] / cfg.N, prop[r] > 0 ? (double)acc[r] / prop[r] : 0.0);
                printf("  [%.2e s/it]\n", (now_s() - t0) / ((long long)step * per_temp + s));
            }
            pspin_die_on(pspin_reset_counters(st.h), "reset");
        }
        if (step >= ntemp / 2) {
            mc_get_spins(st, spins.data());
            for (int r = 0; r < cfg.nrep; r++) {
                char cf[512];
                snprintf(cf, sizeof(cf), "%s/conf_r%d_T%.6f.bin", confdir.c_str(), r, Tt);
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
        if (cfg.verbose == 1) {
            printf("  [%d/%d] T=%.6f", step + 1, ntemp, Tt);
            for (int r = 0; r < cfg.nrep; r++) printf("  E/N=%.4f", E[r] / cfg.N);
            printf("\n");
        }
    }
    const double elapsed = now_s() - t0;
    printf("\n-- Summary ------\n  %-22s %.3f s\n  %-22s %d x %d = %lld\n\n", "total time", elapsed, "total sweeps",
           ntemp, per_temp, (long long)ntemp * per_temp);
    if (FILE* ft = fopen((std::string(dir) + "/time.txt").c_str(), "w")) {
        fprintf(ft, "# N nrep ntemp sweeps_per_temp Tmax Tmin total_time_s\n");
        fprintf(ft, "%d %d %d %d %.6f %.6f %.6f\n", cfg.N, cfg.nrep, ntemp, per_temp, Tmax, Tmin, elapsed);
        fclose(ft);
    }
    fclose(fout);
    mc_free(st);
    return 0;
}
