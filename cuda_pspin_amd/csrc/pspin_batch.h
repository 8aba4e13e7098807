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

// One step's draws of one chain, made for a whole sweep before the chains run:
// the pair and everything of the cube proposal (include/spins_sparse.h:26-56)
// and of the Metropolis threshold that does not depend on the amplitudes.
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
    BatchDraw* draws;   // [chains * max_steps] scratch: sample s at chain_off * max_steps, [step][chain]
    int max_steps;      // largest N / 2 of the batch
};

constexpr int kBatchThreads = 1024;  // largest workgroup of the batch kernel (draws: all threads; chains: 4 or 8 lanes each)

int batch_lanes(int chains);    // lanes per chain: 8 up to 128 chains per sample, else 4
int batch_threads(int chains);  // workgroup size: 512, 768 or 1024
size_t batch_lds_bytes(const BatchSample& s, int NT);
hipError_t batch_prepare(int chains, size_t lds);
hipError_t launch_batch(const BatchArgs& a, int n_samples, int chains, size_t lds, hipStream_t s);

}  // namespace ps
