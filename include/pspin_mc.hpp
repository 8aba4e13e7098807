// pspin_mc.hpp -- C++ mirror of the reference's host API over the C ABI in
// pspin.h, with the reference's names and error behaviour (print, exit(1)):
//   SimConfig / default_config / parse_args / fmt_param / make_run_dir
//       (include/config.h:9-85, src/config.cu:6-70)
//   MCState + mc_init / mc_free / mc_sweep / mc_sweep_pt / mc_set_betas /
//       mc_get_results / mc_get_spins / mc_compute_split_energies
//       (include/mc.h:9-56)
//   MCStateSparse + mc_sparse_* (include/mc_sparse.h:15-61)
//   n_quartets / n_g4_total / n_pairs (include/disorder.h:32-44)
//   generate_g2 / generate_g4 / rescale_* / shift_mean_couplings / apply_fmc_* /
//       init_g4_mask / count_active_g4 (include/disorder.h:23-69),
//       compute_energy (include/hamiltonian.h:11-17), init_spins_uniform
//       (include/spins.h:9), init_spins_cube (include/spins_sparse.h:10)
// Spins are exchanged as pspin_complex {x, y}, the layout of cuDoubleComplex.
// Device buffers are owned by the opaque handle; the driver-side reads and
// resets the reference does directly on d_accepted / d_proposed / d_quartets
// are pspin_reset_counters / pspin_get_sparse_quartets here.
#pragma once

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "pspin.h"

struct pspin_complex {
    double x, y;
};

using SimConfig = pspin_sim_config;

inline void pspin_die_on(int rc, const char* what) {
    if (rc != PSPIN_OK) {
        fprintf(stderr, "%s failed: %s\n", what, pspin_last_error());
        exit(1);
    }
}

inline SimConfig default_config() {
    SimConfig c;
    pspin_default_config(&c);
    return c;
}

inline SimConfig parse_args(int argc, char** argv) {
    SimConfig c;
    if (pspin_parse_args(argc, argv, &c) != PSPIN_OK) {
        fprintf(stderr, "%s\n", pspin_last_error());
        if (!strncmp(pspin_last_error(), "Unknown", 7))
            fprintf(stderr,
                    "Usage: %s [-N n] [-nrep n] [-T temp] [-J coupling] [-J0 mean] [-alpha a] [-alpha0 a0] "
                    "[-iter n] [-seed s] [-save_freq k] [-label l] [-dev d] [-fmc 0|1|2] [-gamma g] "
                    "[-verbose [0|1|2]]\n",
                    argv[0]);
        exit(1);
    }
    return c;
}

inline long long n_quartets(int N) { return (long long)N * (N - 1) * (N - 2) * (N - 3) / 24; }
inline long long n_g4_total(int N) { return n_quartets(N); }
inline long long n_pairs(int N) { return (long long)N * (N + 1) / 2; }

// "%.6g" with a decimal point guaranteed (include/config.h:52-59).
inline void fmt_param(char* buf, int sz, double v) {
    snprintf(buf, sz, "%.6g", v);
    if (!strchr(buf, '.') && !strchr(buf, 'e')) {
        int n = (int)strlen(buf);
        if (n + 3 < sz) { buf[n] = '.'; buf[n + 1] = '0'; buf[n + 2] = '\0'; }
    }
}

// {base}/{prefix}_N{N}_a{alpha}_R{J/J0|inf}_a0{alpha0}_NT{NT}_NR{nrep}[_S{label}]
// (include/config.h:61-85).
inline void make_run_dir(char* buf, int sz, const char* base, const char* prefix, int N, double alpha, double J,
                         double J0, double alpha0, int NT, int nrep, int label) {
    char sa[32], sR[32], sa0[32];
    fmt_param(sa, sizeof(sa), alpha);
    fmt_param(sa0, sizeof(sa0), alpha0);
    if (J0 == 0.0) snprintf(sR, sizeof(sR), "inf");
    else fmt_param(sR, sizeof(sR), J / J0);
    if (label >= 0)
        snprintf(buf, sz, "%s/%s_N%d_a%s_R%s_a0%s_NT%d_NR%d_S%d", base, prefix, N, sa, sR, sa0, NT, nrep, label);
    else
        snprintf(buf, sz, "%s/%s_N%d_a%s_R%s_a0%s_NT%d_NR%d", base, prefix, N, sa, sR, sa0, NT, nrep);
}

struct MCStateBase {
    pspin_state* h = nullptr;
    int N = 0, nrep = 0;
    std::vector<double> omega;  // h_omega (empty when FC)
    double* h_omega = nullptr;
    long long n_pairs_active = 0, n_quart_active = 0;
    int h2_active = 0;
};

struct MCState : MCStateBase {};
struct MCStateSparse : MCStateBase {
    int n_quartets = 0;
};

inline void mc_fill_common(MCStateBase& s, const SimConfig& cfg) {
    pspin_info inf;
    pspin_die_on(pspin_get_info(s.h, &inf), "pspin_get_info");
    s.N = inf.N;
    s.nrep = inf.nrep;
    s.n_pairs_active = inf.n_pairs_active;
    s.n_quart_active = inf.n_quart_active;
    s.h2_active = inf.h2_active;
    if (cfg.fmc_mode > 0) {
        s.omega.resize(s.N);
        pspin_die_on(pspin_get_omega(s.h, s.omega.data()), "pspin_get_omega");
        s.h_omega = s.omega.data();
    }
}

inline MCState mc_init(const SimConfig& cfg) {
    MCState s;
    pspin_die_on(pspin_create(&cfg, PSPIN_DENSE, &s.h), "mc_init");
    mc_fill_common(s, cfg);
    return s;
}

inline MCStateSparse mc_sparse_init(const SimConfig& cfg) {
    MCStateSparse s;
    pspin_die_on(pspin_create(&cfg, PSPIN_SPARSE, &s.h), "mc_sparse_init");
    mc_fill_common(s, cfg);
    pspin_info inf;
    pspin_die_on(pspin_get_info(s.h, &inf), "pspin_get_info");
    s.n_quartets = inf.n_quartets;
    return s;
}

inline void mc_free(MCStateBase& s) {
    pspin_destroy(s.h);
    s.h = nullptr;
}
inline void mc_sparse_free(MCStateSparse& s) { mc_free(s); }

inline void mc_sweep(MCStateBase& s, const SimConfig& cfg) {
    pspin_die_on(pspin_set_uniform_beta(s.h, 1.0 / cfg.T), "mc_sweep");
    pspin_die_on(pspin_sweep(s.h, 1), "mc_sweep");
}
inline void mc_sweep_pt(MCStateBase& s) { pspin_die_on(pspin_sweep(s.h, 1), "mc_sweep_pt"); }
inline void mc_sparse_sweep_pt(MCStateSparse& s) { mc_sweep_pt(s); }

inline void mc_set_betas(MCStateBase& s, const double* b) { pspin_die_on(pspin_set_betas(s.h, b), "mc_set_betas"); }
inline void mc_sparse_set_betas(MCStateSparse& s, const double* b) { mc_set_betas(s, b); }

inline void mc_get_results(const MCStateBase& s, double* E, long long* acc, long long* prop) {
    pspin_die_on(pspin_get_results(s.h, E, acc, prop), "mc_get_results");
}
inline void mc_sparse_get_results(const MCStateSparse& s, double* E, long long* acc, long long* prop) {
    mc_get_results(s, E, acc, prop);
}

inline void mc_get_spins(const MCStateBase& s, pspin_complex* out) {
    pspin_die_on(pspin_get_spins(s.h, reinterpret_cast<double*>(out)), "mc_get_spins");
}
inline void mc_sparse_get_spins(const MCStateSparse& s, pspin_complex* out) { mc_get_spins(s, out); }

inline void mc_compute_split_energies(const MCStateBase& s, double* e2, double* e4) {
    pspin_die_on(pspin_split_energies(s.h, e2, e4), "mc_compute_split_energies");
}
inline void mc_sparse_compute_split_energies(const MCStateSparse& s, double* e2, double* e4) {
    mc_compute_split_energies(s, e2, e4);
}

// ---------------------------------------------------------------------------
// Free functions of include/disorder.h:23-69, include/hamiltonian.h:11-17,
// include/spins.h:9 and include/spins_sparse.h:10 on device buffers of
// pspin_complex (the layout of cuDoubleComplex).
// ---------------------------------------------------------------------------
inline void generate_g2(pspin_complex* d_g2, int N, double J, unsigned long long seed) {
    pspin_die_on(pspin_generate_g2(reinterpret_cast<double*>(d_g2), N, J, seed), "generate_g2");
}
inline void generate_g4(pspin_complex* d_g4, int N, double J, unsigned long long seed) {
    pspin_die_on(pspin_generate_g4(reinterpret_cast<double*>(d_g4), N, J, seed), "generate_g4");
}
inline void rescale_g2(pspin_complex* d_g2, int N, double J, long long n_surviving) {
    pspin_die_on(pspin_rescale_g2(reinterpret_cast<double*>(d_g2), N, J, n_surviving), "rescale_g2");
}
inline void rescale_g4(pspin_complex* d_g4, int N, double J, long long n_surviving) {
    pspin_die_on(pspin_rescale_g4(reinterpret_cast<double*>(d_g4), N, J, n_surviving), "rescale_g4");
}
inline void shift_mean_couplings(pspin_complex* d_c, long long n, double mean) {
    pspin_die_on(pspin_shift_mean_couplings(reinterpret_cast<double*>(d_c), n, mean), "shift_mean_couplings");
}
inline void apply_fmc_g2(pspin_complex* d_g2, int N, const double* d_omega, double gamma) {
    pspin_die_on(pspin_apply_fmc_g2(reinterpret_cast<double*>(d_g2), N, d_omega, gamma), "apply_fmc_g2");
}
inline void init_g4_mask(uint8_t* d_mask, int N) { pspin_die_on(pspin_init_g4_mask(d_mask, N), "init_g4_mask"); }
inline void apply_fmc_g4(uint8_t* d_mask, int N, const double* d_omega, double gamma) {
    pspin_die_on(pspin_apply_fmc_g4(d_mask, N, d_omega, gamma), "apply_fmc_g4");
}
inline long long count_active_g4(const uint8_t* d_mask, int N) {
    long long n = 0;
    pspin_die_on(pspin_count_active_g4(d_mask, N, &n), "count_active_g4");
    return n;
}
inline double compute_energy(const pspin_complex* d_spins, const pspin_complex* d_g2, const pspin_complex* d_g4,
                             const uint8_t* d_g4_mask, int N) {
    double e = 0.0;
    pspin_die_on(pspin_compute_energy(reinterpret_cast<const double*>(d_spins), reinterpret_cast<const double*>(d_g2),
                                      reinterpret_cast<const double*>(d_g4), d_g4_mask, N, &e),
                 "compute_energy");
    return e;
}
inline void init_spins_uniform(pspin_complex* d_spins, int N, unsigned long long seed) {
    pspin_die_on(pspin_init_spins_uniform(reinterpret_cast<double*>(d_spins), N, seed), "init_spins_uniform");
}
inline void init_spins_cube(pspin_complex* d_spins, int N, unsigned long long seed) {
    pspin_die_on(pspin_init_spins_cube(reinterpret_cast<double*>(d_spins), N, seed), "init_spins_cube");
}
