// pspin_internal.h -- shared device helpers and the library state behind
// the C ABI in include/pspin.h.  gfx950 (CDNA4, wave64) only.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "../../include/pspin.h"

#define PS_HD __host__ __device__ __forceinline__

namespace ps {

constexpr int kMaxN = 256;         // site indices are packed in 8 bits
constexpr int kMaxPairsN = 128;    // largest N served by the dense pair-row kernel
constexpr int kWave = 64;          // CDNA wavefront width
constexpr uint32_t kTagSpins = 1;  // Philox stream tags (second key word is the seed)
constexpr uint32_t kTagG2 = 2;
constexpr uint32_t kTagG4 = 4;
// Fourth Philox counter word of the per-step sweep draws (two blocks per
// step).  Distinct from the tags above, which share the replica key with
// spin initialisation, so no sweep draw reuses an initialisation block.
constexpr uint32_t kCtrStep0 = 16;
constexpr uint32_t kCtrStep1 = 17;
constexpr uint32_t kCtrPair = 18;    // pair draws of a shared pair schedule (group key)
constexpr uint64_t kPairKeyOffset = 7000;  // group g's key: splitmix64 of (seed + 7000 + g)


// ---------------------------------------------------------------------------
// splitmix64 (Steele, Lea, Flood 2014) -- the reference's seed chain
// (src/mc.cu:42-47).
// ---------------------------------------------------------------------------
PS_HD uint64_t splitmix64(uint64_t* s) {
    uint64_t z = (*s += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

// Philox key of pair-schedule group g (replicas r with r % groups == g).
PS_HD uint64_t pair_group_key(uint64_t seed, int g) {
    uint64_t s = seed + kPairKeyOffset + (uint64_t)g;
    return splitmix64(&s);
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): counter-based, so every
// (replica, sweep, step) draw is addressable without stored RNG state.
// ---------------------------------------------------------------------------
struct U4 {
    uint32_t x, y, z, w;
};

PS_HD U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        U4 n;
        n.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
        n.y = (uint32_t)p1;
        n.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
        n.w = (uint32_t)p0;
        c = n;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

PS_HD double u32_to_unit(uint32_t x) { return ((double)x + 0.5) * (1.0 / 4294967296.0); }
PS_HD double u64_to_unit(uint32_t hi, uint32_t lo) {
    uint64_t m = ((uint64_t)(hi >> 5) << 26) | (uint64_t)(lo >> 6);
    return ((double)m + 0.5) * (1.0 / 9007199254740992.0);
}

// Standard normal number k of stream (seed, tag): Box-Muller on two 53-bit
// uniforms from one Philox block per pair of normals.
__device__ __forceinline__ double normal_k(uint64_t seed, uint32_t tag, uint64_t k) {
    uint64_t pair = k >> 1;
    U4 x = philox4x32_10(U4{(uint32_t)pair, (uint32_t)(pair >> 32), 0u, tag}, (uint32_t)seed,
                         (uint32_t)(seed >> 32));
    double u1 = u64_to_unit(x.x, x.y), u2 = u64_to_unit(x.z, x.w);
    double r = sqrt(-2.0 * log(u1));
    double s, c;
    sincos(2.0 * M_PI * u2, &s, &c);
    return (k & 1) ? r * s : r * c;
}

// ---------------------------------------------------------------------------
// drand48 (POSIX): X' = (0x5DEECE66D X + 0xB) mod 2^48, value X'/2^48;
// srand48(s) sets X = (low 32 bits of s) << 16 | 0x330E.  lcg48_skip applies
// k steps at once so a wavefront can serve several draws in parallel.
// ---------------------------------------------------------------------------
constexpr uint64_t kLcgA = 0x5DEECE66DULL;
constexpr uint64_t kLcgC = 0xBULL;
constexpr uint64_t kMask48 = (1ULL << 48) - 1;

PS_HD uint64_t srand48_state(long s) { return (((uint64_t)(uint32_t)s) << 16) | 0x330EULL; }

PS_HD uint64_t lcg48_skip(uint64_t x, uint64_t k) {
    uint64_t am = 1, ap = 0, cm = kLcgA, cp = kLcgC;
    while (k) {
        if (k & 1) {
            am = am * cm;
            ap = ap * cm + cp;
        }
        cp = (cm + 1) * cp;
        cm = cm * cm;
        k >>= 1;
    }
    return (am * x + ap) & kMask48;
}

// Affine coefficients (A, C) of k steps: X_{n+k} = A X_n + C (mod 2^48).
PS_HD void lcg48_coeffs(uint64_t k, uint64_t* A, uint64_t* C) {
    uint64_t am = 1, ap = 0, cm = kLcgA, cp = kLcgC;
    while (k) {
        if (k & 1) {
            am = am * cm;
            ap = ap * cm + cp;
        }
        cp = (cm + 1) * cp;
        cm = cm * cm;
        k >>= 1;
    }
    *A = am & kMask48;
    *C = ap & kMask48;
}

PS_HD double lcg48_unit(uint64_t x) { return (double)x * (1.0 / 281474976710656.0); }

PS_HD long long n_quartets(int N) { return (long long)N * (N - 1) * (N - 2) * (N - 3) / 24; }
PS_HD long long n_pairs(int N) { return (long long)N * (N + 1) / 2; }

// Colex decode q -> (i<j<k<l), q = C(l,4)+C(k,3)+C(j,2)+i (src/disorder.cu:159-171).
// Branch-light: l from a float estimate corrected by integer compares.
__device__ __forceinline__ void decode_quartet(long long q, int& i, int& j, int& k, int& l) {
    int L = (int)floor(pow(24.0 * (double)q, 0.25)) + 2;
    while ((long long)L * (L - 1) * (L - 2) * (L - 3) / 24 > q) L--;
    while ((long long)(L + 1) * L * (L - 1) * (L - 2) / 24 <= q) L++;
    long long rem = q - (long long)L * (L - 1) * (L - 2) * (L - 3) / 24;
    int K = (int)floor(cbrt(6.0 * (double)rem)) + 1;
    if (K < 2) K = 2;
    while ((long long)K * (K - 1) * (K - 2) / 6 > rem) K--;
    while ((long long)(K + 1) * K * (K - 1) / 6 <= rem) K++;
    rem -= (long long)K * (K - 1) * (K - 2) / 6;
    int J = (int)floor(sqrt(2.0 * (double)rem)) + 1;
    if (J < 1) J = 1;
    while ((long long)J * (J - 1) / 2 > rem) J--;
    while ((long long)(J + 1) * J / 2 <= rem) J++;
    rem -= (long long)J * (J - 1) / 2;
    i = (int)rem; j = J; k = K; l = L;
}

// Conjugation masks per channel (bit p set = sorted position p conjugated),
// src/mc.cu:119-121.
PS_HD int conj_mask(int ch) { return ch == 0 ? 0xC : (ch == 1 ? 0x6 : 0xA); }

// One step of a sweep, precomputed for all steps at sweep start.
struct StepPre {
    int i0, j0;
    double c[6];  // sphere: cos a, sin a, cos p1, sin p1, cos p2, sin p2; cube: sqrt(1-u), sqrt(u), phases
    double thr;   // accept iff dE < thr (thr = -log(u)/beta; +inf when beta == 0)
};

// Per-step log record (48 bytes), see pspin_enable_step_log.
struct StepRec {
    int i0, j0, acc;
    float dE;
    double ai[2], aj[2];
};

}  // namespace ps

// ---------------------------------------------------------------------------
// Library state (the MCState / MCStateSparse of include/mc.h:9-30 and
// include/mc_sparse.h:15-36, extended with the PT state that the reference
// keeps in its driver).
// ---------------------------------------------------------------------------
struct pspin_state {
    pspin_sim_config cfg{};
    int variant = PSPIN_DENSE;
    int N = 0, nrep = 0;
    int device = 0;
    hipStream_t stream = nullptr;

    int h2_active = 0;
    long long n_pairs_active = 0, n_quart_active = 0;
    std::vector<double> h_omega;

    // disorder
    double* d_g2_64 = nullptr;  // [N*N] real, symmetric
    float* d_g2_32 = nullptr;   // [N*N]
    double* d_g4_64 = nullptr;  // [C(N,4)] dense only
    uint8_t* d_mask = nullptr;  // [C(N,4)] dense only
    int n_sq = 0;               // sparse quartet count
    int* d_sq = nullptr;        // [n_sq*5] (i,j,k,l,ch)
    double* d_sqg = nullptr;    // [n_sq]
    std::vector<int> h_sq;
    std::vector<double> h_sqg;

    // per-site H4 lists used by the sweep: entry = {p | m<<8 | n<<16, g (f32 bits)}
    int* d_site_off = nullptr;      // [N+1]
    uint2* d_site_ent = nullptr;    // [n_ent]
    long long n_ent = 0;
    int max_site_len = 0;
    std::vector<uint2> h_site_ent;  // FP32 entries, kept to re-encode the coupling words
    int coupling_format = PSPIN_COUPLING_F32;
    // dense: per-site chunks of H4 entries sharing the conjugated pair (m, n)
    // (PSPIN_KERNEL_GROUPS): header {m | n << 8, p0 | p1 << 8 | p2 << 16 | p3 << 24}
    // and four coupling words (FP32 bits or BF16 hi/lo), zero in unused slots
    int* d_grp_off = nullptr;       // [N+1] chunk offsets per site
    uint2* d_grp_hdr = nullptr;     // [n_grp + 1], the last chunk all zero
    uint4* d_grp_g = nullptr;       // [n_grp + 1]
    long long n_grp = 0;
    int max_site_grp = 0;
    std::vector<uint4> h_grp_g;     // FP32 words, kept to re-encode the coupling words

    // dense fully connected per-site pair tiles (PSPIN_KERNEL_PAIRS)
    float* d_ptab = nullptr;       // [N][pair_slice_len]
    uint32_t* d_pinfo = nullptr;   // [N][pair_npad] 8 o(b) | 8 o(c) << 16 | partner-a flag << 31
    long long pair_slice_len = 0;  // floats per site
    int pair_npairs = 0, pair_npad = 0;
    std::vector<int> h_tile_rows;  // rows of pair block k's tile (multiple of 4)
    std::vector<int> h_tile_off;   // offset of pair block k's tile in a slice (floats)
    std::vector<uint32_t> h_pair_bc;  // pairs b | c << 16 (others relabelled), b descending
    int4* d_tasks = nullptr;
    int* d_wave_tasks = nullptr;
    int pair_nwaves = 0;
    int pair_rows = 16; // rows in flight per batch in the pair-row kernel (8 or 16)
    int list_batch = 8; // entries in flight per thread in the list kernel (8 or 16; dense only)

    // replicas
    double* d_spins = nullptr;  // [nrep][N][2]
    double* d_E = nullptr;
    double* d_betas = nullptr;
    long long* d_acc = nullptr;
    long long* d_prop = nullptr;
    uint64_t* d_seeds = nullptr;
    std::vector<uint64_t> h_seeds;
    uint64_t sweep_counter = 0;     // host mirror of d_ctr (all replicas advance together)
    uint64_t* d_ctr = nullptr;      // [nrep] per-replica sweep counters
    int resync_every = 1024;        // FP64 energy resync interval in sweeps (0 = never)
    long long since_resync = 0;
    // captured sweep+exchange rounds
    hipGraphExec_t graph = nullptr;
    int graph_pt_freq = 0, graph_rounds = 0;
    uint64_t graph_gen = 0;  // bumped whenever captured sweep graphs become stale
    double* d_e2 = nullptr;
    double* d_e4 = nullptr;
    int* d_echunk = nullptr;    // dense energy: pair-chunk bounds [n_echunk + 1]
    double* d_epart = nullptr;  // dense energy: H4 partial sums [n_echunk][nrep]
    int n_echunk = 0;

    // parallel tempering
    int pt_NT = 0, pt_nrep = 0;
    double* d_beta_sched = nullptr;
    int* d_perm = nullptr;
    long long* d_ex = nullptr;  // [4][NT-1]: acc_win, prop_win, acc_tot, prop_tot
    uint64_t* d_rng48 = nullptr;
    long long pt_sweeps = 0;
    double* d_series = nullptr;
    int series_cap = 0;
    long long* d_series_n = nullptr;
    unsigned int* d_done = nullptr;  // [3]: workgroups arrived, unused, barrier wait expired
    // tagged handoff words of cooperative launches: [2 nrep] energy halves, [nrep] new levels
    unsigned long long* d_hand = nullptr;
    unsigned hand_tag = 0;  // tag base of the next cooperative launch (grows by its rounds)
    unsigned wait_limit = 0;  // polls of a cross-workgroup wait before it is flagged (0: default)

    // step log
    ps::StepRec* d_log = nullptr;
    int log_cap = 0;
    int log_n = 0;

    // kernel choice and profiling
    int pair_groups = 0;  // 0: every replica draws its own pairs; G: replicas r = g mod G share
    int kernel = PSPIN_KERNEL_LISTS;
    int block_size = 0;
    bool prof = false;
    std::vector<hipEvent_t> ev_free;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_sweep, ev_exch;
    int launch_mode = PSPIN_LAUNCH_AUTO;
    double prof_sweep_ms = 0.0, prof_exch_ms = 0.0;
    long long prof_n_sweep = 0, prof_n_exch = 0, prof_n_launch = 0;
};

// ---------------------------------------------------------------------------
// Internal host launchers (one translation unit per kernel family).
// All launch on st->stream and return hipError_t.
// ---------------------------------------------------------------------------
namespace ps {
void set_error(const std::string& msg);

hipError_t launch_g2_build(double* g2, int N, uint64_t seed, int fmc_on, const double* d_omega, double gamma,
                           hipStream_t s);
hipError_t launch_g4_build(double* g4, uint8_t* mask, int N, uint64_t seed, int fmc_mode, const double* d_omega,
                           double gamma, hipStream_t s);
hipError_t launch_scale_shift(double* g, long long n, double sigma, double mean, hipStream_t s);
hipError_t launch_to_f32(const double* src, float* dst, long long n, hipStream_t s);
hipError_t launch_init_spins(double* spins, int N, int nrep, const uint64_t* seeds, int cube, hipStream_t s);

int energy_chunk_count(int N, int nrep);              // pair chunks of the dense energy kernel
std::vector<int> energy_chunks(int N, int nchunk);    // their bounds, about equal quartet counts
hipError_t launch_energy_dense(const pspin_state* st, double* e2, double* e4);
hipError_t launch_energy_sparse(const pspin_state* st, double* e2, double* e4);
hipError_t launch_add(const double* a, const double* b, double* out, int n, hipStream_t s);

// Replica-exchange arguments (exchange.hip), also carried by a sweep launch
// that runs the exchange in its last workgroup.
struct ExArgs {
    int NT, nrep;
    const double* beta_sched;
    const double* E;
    int* perm;
    double* betas;
    long long* ex;  // [4][NT-1]: acc_win, prop_win, acc_tot, prop_tot
    uint64_t* rng48;
    double* series;
    long long* series_n;
    int series_cap;
};
ExArgs exchange_args(const pspin_state* st);
// LDS bytes of the one-wave exchange for NT x nrep slots (layout in ExLds,
// exchange_dev.h): energies by slot and by replica, draw logs, and the
// permutation per slot with a spare row on each side; betas, accept counts and
// the four exchange counters per level; the 48-bit state, the series counter
// and the number of draws taken.
inline size_t exchange_lds_bytes(int NT, int nrep) {
    return (size_t)(NT + 2) * nrep * sizeof(int) + (size_t)NT * nrep * (2 * sizeof(double) + sizeof(float)) +
           (size_t)NT * 6 * sizeof(double) + 3 * sizeof(double);
}
// LDS bytes of the many-ladder exchange kernel (more than 64 ladders).
inline size_t exchange_wide_lds_bytes(int NT, int nrep) {
    return (size_t)NT * nrep * (sizeof(double) + sizeof(int)) + (size_t)NT * (sizeof(double) + 8);
}
// LDS bytes of the exchange kernel that serves NT x nrep slots.
inline size_t exchange_kernel_lds_bytes(int NT, int nrep) {
    return nrep <= 64 ? exchange_lds_bytes(NT, nrep) : exchange_wide_lds_bytes(NT, nrep);
}
// Whether a sweep launch can run the exchange in its last workgroup.
bool exchange_fusable(const pspin_state* st);

// Sweeps; with exchange = true the replica exchange follows (fused into the
// same launch when exchange_fusable, else a separate launch).
hipError_t launch_sweep(pspin_state* st, int n_sweeps, bool exchange = false);
// `rounds` PT rounds (pt_freq sweeps, then the exchange) in one cooperative
// launch with a grid barrier between rounds; pt_rounds_ok says whether every
// workgroup can be resident at once (and the exchange can be fused).
bool pt_rounds_ok(const pspin_state* st);
hipError_t launch_pt_rounds(pspin_state* st, int pt_freq, int rounds);

hipError_t launch_exchange(pspin_state* st);

// pspin_pt_run without the final stream synchronisation (api.cpp).
// launch_rounds: PT rounds per cooperative launch or captured graph (0: the
// default, about 32 sweeps).
int pt_advance(pspin_state* st, int n_sweeps, int pt_freq, int launch_rounds = 0);
// After the stream has drained: PSPIN_ERR_HIP if a grid-barrier wait of a
// cooperative PT launch expired (rounds then ran without their exchange).
int pt_check_barrier(pspin_state* st);
}  // namespace ps
