/* pspin.h -- C ABI of the MI355X-native 2+4 p-spin Monte Carlo library
 * (libpspin.so, hand-written gfx950 HIP kernels).
 *
 * The reference exposes a C++ header API (no FFI): MCState / MCStateSparse
 * plus free functions.  Every entry point below replaces one of them; the
 * citation names the reference declaration (file:line).  A C++ mirror with
 * the reference's own names lives in include/pspin_mc.hpp.
 *
 * Conventions
 *  - Every call returns PSPIN_OK (0) or a negative error code; the message is
 *    available from pspin_last_error().  (The reference prints and exit(1)s
 *    on any CUDA error, src/mc.cu:12-20; the C++ mirror restores that.)
 *  - Calls are synchronous on return, like the reference's API, which ends
 *    every call with cudaDeviceSynchronize (src/mc.cu:765,799).
 *  - One handle is single-threaded.  Host buffers are caller-allocated.
 *  - Spins cross the boundary as interleaved (re, im) doubles, replica-major
 *    [nrep][N][2], the byte layout of cuDoubleComplex arrays in the reference.
 */
#ifndef PSPIN_H
#define PSPIN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PSPIN_OK 0
#define PSPIN_ERR_ARG (-1)
#define PSPIN_ERR_HIP (-2)
#define PSPIN_ERR_STATE (-3)
#define PSPIN_ERR_NOMEM (-4)

/* Constraint / disorder variants. */
#define PSPIN_DENSE 0   /* spherical sum|a|^2=N, dense C(N,4) H4 (mc.h)            */
#define PSPIN_SPARSE 1  /* smoothed cube sum|a|^4=N, N sampled quartets (mc_sparse.h) */

/* Same fields and meaning as SimConfig (include/config.h:9-25). */
typedef struct pspin_sim_config {
    int save_freq;
    int verbose;
    int N;
    int nrep;
    double T;
    double J;
    double J0;
    double alpha;
    double alpha0;
    int mc_iterations;
    uint64_t seed;
    int label;
    int dev;
    int fmc_mode; /* 0 fully connected, 1 comb, 2 uniform */
    double gamma;
} pspin_sim_config;

typedef struct pspin_info {
    int N;
    int nrep;
    int variant;
    int h2_active;
    long long n_pairs_active;  /* MCState::n_pairs_active (include/mc.h:24) */
    long long n_quart_active;  /* MCState::n_quart_active (include/mc.h:25) */
    int n_quartets;            /* MCStateSparse::n_quartets (include/mc_sparse.h:19); C(N,4) if dense */
    long long n_site_entries;  /* stored (site, quartet) H4 entries used by the sweep */
    int coupling_format;       /* PSPIN_COUPLING_* */
    int kernel;                /* PSPIN_KERNEL_* in use */
} pspin_info;

typedef struct pspin_state pspin_state;

const char* pspin_last_error(void);

/* Default SimConfig values (include/config.h:28-46). */
void pspin_default_config(pspin_sim_config* cfg);

/* parse_args (src/config.cu:6-70): same flags, same auto-gamma.  Returns
 * PSPIN_ERR_ARG on an unknown flag or N<4 (the reference exits 1). */
int pspin_parse_args(int argc, char** argv, pspin_sim_config* cfg);

/* mc_init (include/mc.h:33) / mc_sparse_init (include/mc_sparse.h:42):
 * disorder (g2, g4, FMC mask), spins, RNG keys and FP64 initial energies.
 * variant: PSPIN_DENSE or PSPIN_SPARSE. */
int pspin_create(const pspin_sim_config* cfg, int variant, pspin_state** out);

/* mc_free (include/mc.h:36) / mc_sparse_free (include/mc_sparse.h:45). */
int pspin_destroy(pspin_state* st);

int pspin_get_info(const pspin_state* st, pspin_info* info);

/* mc_set_betas (include/mc.h:52) / mc_sparse_set_betas (mc_sparse.h:51):
 * betas[nrep] by physical replica. */
int pspin_set_betas(pspin_state* st, const double* betas);

/* mc_sweep_pt (include/mc.h:49) / mc_sparse_sweep_pt (mc_sparse.h:48),
 * n_sweeps times with the current betas.  mc_sweep (include/mc.h:39) is
 * pspin_set_uniform_beta(1/T) followed by this call. */
int pspin_sweep(pspin_state* st, int n_sweeps);
int pspin_set_uniform_beta(pspin_state* st, double beta);

/* mc_get_results (include/mc.h:42-43). */
int pspin_get_results(const pspin_state* st, double* E, long long* accepted, long long* proposed);

/* The drivers' cudaMemset of d_accepted / d_proposed
 * (src/parallel_tempering.cu:396-397, src/simulated_annealing.cu:205-206). */
int pspin_reset_counters(pspin_state* st);

/* mc_get_spins (include/mc.h:46) / mc_sparse_get_spins (mc_sparse.h:57). */
int pspin_get_spins(const pspin_state* st, double* re_im);

/* Direct writes of d_spins in the reference tests (src/test_delta_e.cu:138);
 * recomputes the FP64 energies. */
int pspin_set_spins(pspin_state* st, const double* re_im);

/* mc_compute_split_energies (include/mc.h:56) / sparse (mc_sparse.h:60):
 * FP64 E2, E4 per replica, i<j H2 convention. */
int pspin_split_energies(const pspin_state* st, double* e2, double* e4);

/* compute_energy (include/hamiltonian.h:11-17) for all replicas, written
 * into the running energies: the FP64 resync of the running accumulator. */
int pspin_resync_energy(pspin_state* st);

/* Periodic FP64 resync inside pspin_sweep / pspin_pt_run: every `sweeps`
 * sweeps (default 1024; 0 = never, as in the reference, which never resyncs
 * its FP64 running energy, src/mc.cu:301). */
int pspin_set_resync_interval(pspin_state* st, int sweeps);
int pspin_set_energies(pspin_state* st, const double* E); /* test hook */

/* Disorder read-back, for tests and the drivers' output files
 * (frequencies.txt src/parallel_tempering.cu:249-256; quartets.txt
 * src/parallel_tempering_sparse.cu:277-296; d_g2 / d_g4 / d_g4_mask reads in
 * src/test_delta_e.cu:138-145).  omega[N] (zeros in FC mode), g2[N*N],
 * g4[C(N,4)] and mask[C(N,4)] (dense only), quartets[n_quartets*5] as
 * (i,j,k,l,ch) and gq[n_quartets] (sparse only). */
int pspin_get_omega(const pspin_state* st, double* omega);
int pspin_get_g2(const pspin_state* st, double* g2);
int pspin_get_g4(const pspin_state* st, double* g4, uint8_t* mask);
int pspin_get_sparse_quartets(const pspin_state* st, int* quartets, double* gq);

/* ---------------- free functions on caller-owned device buffers ----------
 * The reference's disorder.h, hamiltonian.h, spins.h and spins_sparse.h API.
 * Pointers named d_* are device pointers (hipMalloc, or a framework's
 * tensor storage).  Couplings and spins are interleaved FP64 complex (the
 * byte layout of cuDoubleComplex): g2 [N*N][2], g4 [C(N,4)][2], spins [N][2];
 * masks are one byte per colex quartet q = C(l,4)+C(k,3)+C(j,2)+i; omega is
 * FP64 [N].  Kernels run on the null stream and the call waits for them.
 * The Gaussian values come from this library's Philox4x32-10 streams, not
 * from cuRAND, so they match the reference in distribution only; applied in
 * the reference's order (generate, FMC, count, rescale, shift) they equal
 * what pspin_create builds for the same seed, bit for bit. */

/* generate_g2 (include/disorder.h:23): unit real normals, symmetric, diagonal
 * kept, imaginary parts zero.  J is unused, as in the reference. */
int pspin_generate_g2(double* d_g2, int N, double J, unsigned long long seed);
/* generate_g4 (include/disorder.h:29): one unit real normal per quartet
 * (stream key seed + 12345, as src/disorder.cu:98). */
int pspin_generate_g4(double* d_g4, int N, double J, unsigned long long seed);
/* rescale_g2 / rescale_g4 (include/disorder.h:47,51): both parts times
 * J sqrt(N / n_surviving); nothing when n_surviving <= 0. */
int pspin_rescale_g2(double* d_g2, int N, double J, long long n_surviving);
int pspin_rescale_g4(double* d_g4, int N, double J, long long n_surviving);
/* shift_mean_couplings (include/disorder.h:55): real part + mean on every
 * entry with a non-zero part; nothing when n <= 0 or mean == 0. */
int pspin_shift_mean_couplings(double* d_c, long long n, double mean);
/* apply_fmc_g2 (include/disorder.h:58): zero g2[i][j] and g2[j][i], i<j, when
 * |omega_i - omega_j| > gamma; the diagonal is kept. */
int pspin_apply_fmc_g2(double* d_g2, int N, const double* d_omega, double gamma);
/* init_g4_mask (include/disorder.h:62): every quartet channel 0 (0x1). */
int pspin_init_g4_mask(uint8_t* d_mask, int N);
/* apply_fmc_g4 (include/disorder.h:66): overwrites the mask with the first
 * channel whose balance holds, |((w_i+w_j)-w_k)-w_l| <= gamma (0x1), then
 * |((w_i+w_l)-w_j)-w_k| (0x2), then |((w_i+w_k)-w_j)-w_l| (0x4), else 0,
 * FP64 sums left to right: bit-exact against the reference's kernel. */
int pspin_apply_fmc_g4(uint8_t* d_mask, int N, const double* d_omega, double gamma);
/* count_active_g4 (include/disorder.h:69): set bits of the mask, counted on
 * the GPU (the reference copies the mask to the host). */
int pspin_count_active_g4(const uint8_t* d_mask, int N, long long* n_active);
/* compute_energy (include/hamiltonian.h:11-17): FP64 H2 + H4 of one
 * configuration, H4 = -sum_q sum_{ch in mask[q]} Re[g4_q f_ch(a)] with the
 * channel products of src/hamiltonian.cu:101-108 (complex couplings
 * allowed).  H2 sums over i<j, the convention of the reference's sweep and
 * CPU oracle (src/mc.cu:164-179, src/test_delta_e.cu:38-43), not the i<=j of
 * src/hamiltonian.cu:30-38.  d_g2 may be NULL (no two-body term).  The sums
 * run in a fixed order, so the result is deterministic.  Allocates a
 * workspace of one double per workgroup (at most 1025) per call, as the
 * reference allocates its partial sums (src/hamiltonian.cu:168-196). */
int pspin_compute_energy(const double* d_spins, const double* d_g2, const double* d_g4, const uint8_t* d_mask, int N,
                         double* energy);
/* pspin_compute_energy without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Its workspace (1025 doubles: the H4 partials and the result) is a static
 * device array of the library, one per GPU, looked up once per device; the
 * call makes no hipMalloc, hipFree or other allocating call.  Limit: N <= 256
 * (PSPIN_ERR_ARG above).  Calls on one GPU share the array, so calls that
 * overlap on one GPU (from several host threads) race; every partial is
 * rewritten by each call, so nothing needs clearing between calls. */
int pspin_compute_energy_noalloc(const double* d_spins, const double* d_g2, const double* d_g4, const uint8_t* d_mask,
                                 int N, double* energy);
/* init_spins_uniform (include/spins.h:9): 2N normals rescaled to
 * sum |a|^2 = N.  init_spins_cube (include/spins_sparse.h:10): rescaled to
 * sum |a|^4 = N. */
int pspin_init_spins_uniform(double* d_spins, int N, unsigned long long seed);
int pspin_init_spins_cube(double* d_spins, int N, unsigned long long seed);
/* Colex decode of n quartet indices into sorted (i, j, k, l) int quadruples
 * [n][4]: the device decoder every dense kernel uses (the round trip of
 * src/test_quartet_index.cu:56-74). */
int pspin_decode_quartets(const long long* d_q, int* d_ijkl, long long n);

/* ---------------- parallel tempering (src/parallel_tempering.cu) -------- */
/* Set up NT x nrep_per_temp replicas (nrep must equal their product),
 * betas from beta_sched[NT] (logical index t*nrep_per_temp+r, :273-278), the
 * identity permutation (:283-284) and the exchange stream srand48(seed48)
 * (:304; pass (long)seed + 5000 to match the reference). */
int pspin_pt_init(pspin_state* st, int NT, int nrep_per_temp, const double* beta_sched, long seed48);

/* The driver's main loop (:312-352) on the GPU: n_sweeps sweeps, with a
 * replica exchange after every sweep s where (sweep_count+1) % pt_freq == 0.
 * Exchange decisions follow the reference exactly: hot-to-cold sequential
 * order, delta = (b_t - b_{t+1})(E_a - E_b), one drand48 draw per delta<0. */
int pspin_pt_run(pspin_state* st, int n_sweeps, int pt_freq);

/* One exchange round using the current running energies (test hook). */
int pspin_pt_exchange(pspin_state* st);

/* perm[NT*nrep_per_temp]: logical -> physical replica (:280-284). */
int pspin_pt_get_perm(const pspin_state* st, int* perm);

/* Exchange counters per temperature pair [NT-1]: window (since the last
 * pspin_pt_reset_window) and cumulative (:297-301). */
int pspin_pt_get_exchange_stats(const pspin_state* st, long long* acc_win, long long* prop_win,
                                long long* acc_tot, long long* prop_tot);
int pspin_pt_reset_window(pspin_state* st);

/* drand48 state of the exchange stream (48 bits), for checking against libc. */
int pspin_pt_get_rng48(const pspin_state* st, uint64_t* x48);

/* Energy at the coldest slot (t = NT-1) of each ladder r, recorded after
 * every exchange round into a device ring of `capacity` rounds (0 disables).
 * pspin_pt_get_series copies min(n_recorded, capacity) rounds oldest first,
 * layout [round][nrep_per_temp]; *n_out receives the number of rounds. */
int pspin_pt_enable_series(pspin_state* st, int capacity);
int pspin_pt_get_series(const pspin_state* st, double* out, int* n_out);

/* ---------------- batches of sparse PT systems --------------------------
 * The reference runs one disorder sample per process (scan_samples_pt loops
 * over labels).  A batch holds n_samples independent PSPIN_SPARSE systems,
 * sample s with N = Ns[s] and seed = seeds[s] (other fields from `base`), each
 * with NT x nrep_per_temp chains on the geometric/linear schedule beta_sched,
 * and runs them all in one launch: one workgroup per sample, 8 lanes per
 * chain (4 above 128 chains per sample; NT * nrep_per_temp <= 256), the
 * exchange of src/parallel_tempering.cu:316-352 per sample with
 * srand48(seed + 5000).  Sample s starts from the same disorder, spins and
 * RNG seeds as pspin_create(.., PSPIN_SPARSE) with that seed. */
typedef struct pspin_batch pspin_batch;
int pspin_batch_create(const pspin_sim_config* base, int n_samples, const int* Ns, const uint64_t* seeds, int NT,
                       int nrep_per_temp, const double* beta_sched, pspin_batch** out);
int pspin_batch_destroy(pspin_batch* b);
int pspin_batch_set_resync_interval(pspin_batch* b, int sweeps);
/* n_sweeps sweeps; exchange after each sweep completing a multiple of pt_freq (0: none). */
int pspin_batch_run(pspin_batch* b, int n_sweeps, int pt_freq);
int pspin_batch_info(const pspin_batch* b, int* n_samples, long long* chains, long long* pair_updates_per_sweep);
/* Per-sample read-back; chains are ordered t * nrep_per_temp + r (physical). */
int pspin_batch_get_results(const pspin_batch* b, int sample, double* E, long long* acc, long long* prop);
int pspin_batch_get_spins(const pspin_batch* b, int sample, double* re_im);
int pspin_batch_get_perm(const pspin_batch* b, int sample, int* perm);
int pspin_batch_get_exchange_stats(const pspin_batch* b, int sample, long long* acc_tot, long long* prop_tot);
/* Sweeps per kernel launch (default 64; results do not depend on it), and a
 * timing autotuner over 16..256 that restores the batch after each trial. */
int pspin_batch_set_chunk(pspin_batch* b, int sweeps_per_launch);
int pspin_batch_autotune(pspin_batch* b, int trial_sweeps, int pt_freq, int verbose, int* chunk,
                         double* ms_per_sweep);

/* ---------------- test and profiling hooks (no reference counterpart) --- */
/* Per-step log of the next sweeps: capacity_sweeps sweeps per replica.
 * Record layout (48 bytes): int i0, j0, accepted; float dE32;
 * double ai_new[2], aj_new[2].  Layout [rep][sweep][step]. */
int pspin_enable_step_log(pspin_state* st, int capacity_sweeps);
int pspin_get_step_log(const pspin_state* st, void* out, int* n_sweeps_logged);

/* Per-replica RNG seeds (splitmix64 chain, src/mc.cu:418-423) and the global
 * sweep counter that keys each sweep's Philox stream. */
int pspin_get_rng_info(const pspin_state* st, uint64_t* seeds, uint64_t* sweep_counter);

/* Sweep-kernel timing: when on, every sweep launch (one or more sweeps, and
 * the exchanges of a multi-round PT launch) is bracketed by HIP events on the
 * library's stream; totals and the sweeps they cover accumulate until reset.
 * Separate exchange launches are timed on their own.  Graph replays are not
 * timed, so profiling never takes the graph path. */
int pspin_set_profiling(pspin_state* st, int on);
int pspin_get_profile(pspin_state* st, double* sweep_ms, long long* n_sweeps, double* exch_ms,
                      long long* n_exch_launches);
/* Sweep launches behind pspin_get_profile's totals (a multi-round PT launch
 * counts once).  No counterpart in the reference. */
int pspin_get_profile_launches(pspin_state* st, long long* n_sweep_launches);

/* How pspin_pt_run launches whole PT rounds (pt_freq sweeps, one exchange).
 * No counterpart in the reference (its host loop launches each sweep):
 *   PSPIN_LAUNCH_AUTO    one cooperative launch of many rounds, when every
 *                        workgroup fits on the device at once: workgroup 0
 *                        exchanges on energies published as tagged words and
 *                        publishes each replica's new level the same way (up
 *                        to 64 temperatures; with more, the last workgroup to
 *                        arrive exchanges); otherwise as PSPIN_LAUNCH_GRAPHS;
 *   PSPIN_LAUNCH_GRAPHS  captured graphs of one sweep launch per round;
 *   PSPIN_LAUNCH_DIRECT  one sweep launch per round from the host.
 * The exchange runs in the last workgroup of the round's sweep launch when up
 * to 64 ladders fit its LDS, else as a launch of its own.  Every mode gives
 * the same chains bit for bit. */
enum { PSPIN_LAUNCH_AUTO = 0, PSPIN_LAUNCH_GRAPHS = 1, PSPIN_LAUNCH_DIRECT = 2 };
int pspin_set_launch_mode(pspin_state* st, int mode);

/* Bound on the polls of every cross-workgroup wait in a cooperative launch
 * (0: the default, about a second of polling).  A wait that reaches it is
 * flagged, the launch goes on with what it has, and pspin_pt_run returns
 * PSPIN_ERR_HIP.  For tests of that error path; no counterpart in the
 * reference, whose rounds meet at kernel boundaries. */
int pspin_set_wait_limit(pspin_state* st, unsigned polls);

/* cudaMemGetInfo / cudaGetDeviceProperties counterpart for the drivers'
 * memory pre-check and banner (src/parallel_tempering.cu:178-202).  dev < 0
 * means the current device.  name may be NULL. */
int pspin_device_info(int dev, size_t* free_bytes, size_t* total_bytes, char* name, int name_len);

/* Kernel selection and coupling storage (for the autotuner and tests). */
#define PSPIN_KERNEL_AUTO 0
#define PSPIN_KERNEL_LISTS 1  /* per-site (partner, conj, conj, g) lists       */
#define PSPIN_KERNEL_PAIRS 2  /* dense FC per-site pair-row layout             */
#define PSPIN_KERNEL_GROUPS 3 /* dense: per-site chunks of up to 4 partners
                               * sharing one conjugated pair (any channels)   */
#define PSPIN_COUPLING_F32 0
#define PSPIN_COUPLING_BF16X2 1 /* BF16 hi/lo split, FP32 accumulation       */
int pspin_set_kernel(pspin_state* st, int kernel, int block_size);
/* Pair schedule.  groups = 0 (default): every replica draws its own pair,
 * as the reference does (src/mc.cu:129-141, per-replica curand states).
 * groups = G > 0: replicas r with equal r mod G draw their pairs from one
 * shared stream (key splitmix64 of seed + 7000 + g); angles and acceptance
 * stay per replica.  Each replica's transition kernel is unchanged (the pair
 * sequence does not depend on the state), so every stationary law is too.
 * With G = 8 the replicas of one group run on one XCD and read the same
 * coupling slices at each step from that XCD's L2.  pspin_get_pair_keys
 * gives each replica's group key (0 with per-replica pairs). */
int pspin_set_pair_schedule(pspin_state* st, int groups);
/* Tile shape of the pair-row kernel: waves per workgroup (1..16) and rows per
 * batch (4 or 8; loads run one batch ahead per lane).  Fails with
 * PSPIN_ERR_ARG when a wave would get more than 64 batches.  pspin_autotune
 * picks it. */
int pspin_set_pair_tiles(pspin_state* st, int nwaves, int rows);
/* Coupling entries in flight per thread in the list kernel (8 or 16; 16 for
 * the dense variant only).  The group kernel keeps 4 (batch 8) or 6 (batch
 * 16) chunks of up to 4 entries in flight.  pspin_autotune picks it. */
int pspin_set_list_batch(pspin_state* st, int batch);
int pspin_get_pair_keys(const pspin_state* st, uint64_t* keys);
/* Timing autotuner for the sweep (replaces optimal_block_size_for_sweep,
 * src/mc.cu:709-744, which only queries occupancy): times each candidate
 * kernel / tile shape for trial_sweeps sweeps on the live state, restores the
 * state after every trial, keeps the fastest.  verbose prints one
 * "[auto-tune]" line per candidate.  Outputs may be NULL. */
int pspin_autotune(pspin_state* st, int trial_sweeps, int verbose, int* kernel, int* block, double* ms_per_sweep);
/* Coupling storage of the list and group kernels: FP32 (default) or BF16
 * hi/lo pairs (g ~ hi + lo, ~16 mantissa bits) with FP32 accumulation. */
int pspin_set_coupling_format(pspin_state* st, int format);

/* ------------------------------------------------------------------------
 * Learned global-move proposal.  It has no counterpart in the reference,
 * which has only the pair-rotation move (include/spins.h:17-52).
 *
 * An MLP (2N+1 inputs: Re/Im of the amplitudes in a gauge-fixed frame plus
 * beta; two SiLU hidden layers; 4N outputs) gives the mean and log-scale of
 * a diagonal Gaussian over the 2N real coordinates.  A draw is projected
 * radially onto sum|a|^2 = N and accepted by Metropolis-Hastings against
 * pi ~ exp(-beta H) prod_k |a_k|^-1, the stationary law of the pair-rotation
 * sweep, with the exact projected-Gaussian density and FP64 full energies.
 * Dense (spherical) states only; N a multiple of 16, at most 128.
 * ---------------------------------------------------------------------- */
typedef struct pspin_learned pspin_learned;
typedef struct pspin_lp_config {
    int hidden;           /* hidden width, multiple of 64 (512)                    */
    int batch_cap;        /* largest training batch, multiple of 64 (4096)         */
    long long data_cap;   /* snapshot capacity: data_cap x nrep configurations     */
    double log_sigma0;    /* added to the log-scale output (log 0.05)              */
    double log_sigma_min; /* log-scale clamp (-9, 1)                               */
    double log_sigma_max;
    double scale;         /* proposal scale factor on sigma (1)                    */
    double init_gain;     /* scale of the output layer's initial weights (0.01)    */
    uint64_t seed;        /* weight initialisation and batch sampling              */
    double mean_scale;    /* factor on the mean offset: mean = u + mean_scale*out (1) */
} pspin_lp_config;

int pspin_lp_default_config(pspin_lp_config* cfg);
int pspin_lp_create(pspin_state* st, const pspin_lp_config* cfg, pspin_learned** out);
int pspin_lp_destroy(pspin_learned* lp);
/* Flat FP32 parameters [W1 (hidden x in_pad) b1 W2 b2 W3 (4N x hidden) b3];
 * in_pad = 2N+1 rounded up to 64, padding columns of W1 are zero. */
long long pspin_lp_n_params(const pspin_learned* lp);
int pspin_lp_dims(const pspin_learned* lp, int* in_dim, int* in_pad, int* hidden, int* out_dim);
int pspin_lp_get_params(const pspin_learned* lp, float* out);
int pspin_lp_set_params(pspin_learned* lp, const float* in); /* also resets Adam */
int pspin_lp_set_scale(pspin_learned* lp, double scale);
int pspin_lp_set_mean_scale(pspin_learned* lp, double mean_scale);
/* Network evaluation on host rows x [M][in_dim] (M <= batch_cap): FP32
 * outputs [M][4N]. */
int pspin_lp_forward(pspin_learned* lp, const float* x, int M, float* out);
/* Gaussian negative log-likelihood of targets y [M][2N] given inputs x
 * [M][in_dim] and gauge-fixed amplitudes u [M][2N] (mean = u + output):
 * mean loss per row and the gradient of the mean loss [n_params]. */
int pspin_lp_loss_grad(pspin_learned* lp, const float* x, const float* u, const float* y, int M, double* loss,
                       float* grad);
/* Training data: record every replica's amplitudes and beta (one snapshot),
 * or run PT (pspin_pt_run) recording a snapshot every `every` sweeps. */
int pspin_lp_record(pspin_learned* lp);
int pspin_lp_collect(pspin_learned* lp, long long n_snapshots, int every, int pt_freq);
long long pspin_lp_data_size(const pspin_learned* lp);
/* n_steps Adam steps on random (snapshot t, replica r) -> (t + lag, r)
 * pairs; loss_hist [n_steps] (may be NULL) receives the batch losses. */
int pspin_lp_train(pspin_learned* lp, long long n_steps, int batch, int lag, double lr, double* loss_hist);
/* One learned move for every replica (exact MH); energies are resynced to
 * the FP64 full energy as a side effect. */
int pspin_lp_move(pspin_learned* lp);
/* PT with learned moves: n_sweeps sweeps, exchange every pt_freq sweeps,
 * one learned move after every `move_every` sweeps (move_every a multiple
 * of pt_freq). */
int pspin_lp_run(pspin_learned* lp, long long n_sweeps, int pt_freq, int move_every);
/* Accepted / proposed learned moves per temperature slot [NT] (one slot
 * before pspin_pt_init); reset with pspin_lp_reset_stats. */
int pspin_lp_get_stats(const pspin_learned* lp, long long* acc, long long* prop);
int pspin_lp_reset_stats(pspin_learned* lp);
/* Last move, per replica: proposal amplitudes [R][N][2] (may be NULL) and
 * dbg [R][8] = log q(a'|a), log q(a|a'), E(a), E(a'), log MH ratio,
 * accepted, log w(a), log w(a') with w = prod_k |a_k|^-1; net outputs at
 * the current state and at the proposal [R][4N] (may be NULL). */
int pspin_lp_get_last(const pspin_learned* lp, double* prop, double* dbg, float* out_fwd, float* out_rev);
int pspin_lp_get_move_counters(const pspin_learned* lp, uint64_t* ctr);

#ifdef __cplusplus
}
#endif
#endif /* PSPIN_H */
