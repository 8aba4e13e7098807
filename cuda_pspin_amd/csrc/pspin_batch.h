// pspin_batch.h -- device-side description of a batch of sparse PT systems
// (one workgroup per disorder sample, 4 or 8 lanes per chain; see batch.hip).
#pragma once

#include "pspin_internal.h"

namespace ps {

struct BatchSample {
    int N, chains, chain_off;  // chains = NT * nrep_per_temp, first global chain index
    long long spin_off;        // offset of the sample's spins in doubles
    int e4_off, n_e4;          // per-site H4 entries (concatenated lists)
    int e2_off, n_e2;          // per-site H2 entries {k, g f32}
    int o_off;                 // offsets o4/o2 [N+1] start
    int q_off, n_q;            // FP64 quartets (resync)
    int h2_off, n_h2;          // FP64 i<j pairs (resync)
};

// One step's draws of one chain: the pair and everything of the cube proposal
// (include/spins_sparse.h:26-56) and of the Metropolis threshold that does
// not depend on the amplitudes.  They live in registers (batch.hip).
struct BatchDraw {
    int ij;              // i0 | j0 << 16
    int pad;
    double c1, s1, c2, s2;  // cos and sin of the two phases
    double q1, q2;          // sqrt(1 - u), sqrt(u) of the intensity split u
    double lu;              // -log(u_accept); the threshold is lu / beta
};

struct BatchArgs {
    const BatchSample* samples;  // in launch order
    int n_sweeps, pt_freq, NT, resync_every;
    const uint2* e4;
    const uint2* e2;
    const int* o4;
    const int* o2;
    const int* quartets;  // [*][5]
    const double* qg;
    const int2* h2ij;
    const double* h2g;
    double* spins;
    double* E;
    long long* acc;
    long long* prop;
    const uint64_t* seeds;
    uint64_t* ctr;
    int* since;
    double* betas;
    int* perm;  // per sample [chains], logical t*nr+r -> chain within the sample
    const double* beta_sched;
    long long* ex;      // per launch slot [4][NT-1]
    uint64_t* rng48;    // per launch slot
    long long* done;    // per launch slot: sweeps done (pt_freq alignment)
};

// Byte offsets of the batch kernel's LDS for one sample (C chains, NT levels,
// nr = C / NT per level): FP64 amplitudes [N][C | 1], H4 and H2 entries,
// list offsets, then the exchange's arrays.
struct BatchLdsLayout {
    size_t amp, e4, e2, o4, o2;
    size_t sE;     // [C] energies by chain
    size_t sES;    // [C] energies by slot
    size_t sB;     // [C] betas by chain
    size_t sperm;  // [C + 1] chain by slot, and a spare entry for idle lanes' writes
    size_t jA, jC; // [65] drand48 jump-ahead coefficients of 0..64 steps
    size_t sacc;   // [NT] accepted swaps per level over the launch
    size_t sBs;    // [NT] beta schedule
    size_t sDb;    // [NT] beta_t - beta_{t+1}
    size_t sAB;    // accept bits, one 64-bit word per lane and 64 levels
    size_t sLU;    // [C] logf of each draw the round can take
    size_t end;
};

__host__ __device__ inline BatchLdsLayout batch_lds_layout(const BatchSample& s, int NT) {
    auto al8 = [](size_t b) { return (b + 7) / 8 * 8; };
    const size_t C = (size_t)s.chains;
    BatchLdsLayout L;
    L.amp = 0;
    L.e4 = (size_t)s.N * (C | 1) * 16;
    L.e2 = L.e4 + (size_t)s.n_e4 * 8;
    L.o4 = L.e2 + (size_t)s.n_e2 * 8;
    L.o2 = L.o4 + (size_t)(s.N + 1) * 4;
    L.sE = al8(L.o2 + (size_t)(s.N + 1) * 4);
    L.sES = L.sE + C * 8;
    L.sB = L.sES + C * 8;
    L.sperm = L.sB + C * 8;
    L.jA = al8(L.sperm + (C + 1) * 4);
    L.jC = L.jA + 65 * 8;
    L.sacc = L.jC + 65 * 8;
    L.sBs = L.sacc + (size_t)NT * 8;
    L.sDb = L.sBs + (size_t)NT * 8;
    L.sAB = L.sDb + (size_t)NT * 8;
    L.sLU = L.sAB + (size_t)((NT - 1 + 63) / 64 * 64) * 8;
    L.end = L.sLU + C * 4;
    return L;
}

constexpr int kBatchThreads = 1024;  // largest workgroup of the batch kernel (draws: all threads; chains: 4 or 8 lanes each)

int batch_lanes(int chains);    // lanes per chain: 8 up to 128 chains per sample, else 4
int batch_threads(int chains);  // workgroup size: 512, 768 or 1024
size_t batch_lds_bytes(const BatchSample& s, int NT);
hipError_t batch_prepare(int chains, size_t lds);
hipError_t launch_batch(const BatchArgs& a, int n_samples, int chains, size_t lds, hipStream_t s);

}  // namespace ps
