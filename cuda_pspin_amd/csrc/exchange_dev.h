// exchange_dev.h -- device side of the replica exchange, shared by the
// exchange kernels and the sweep launch that runs the exchange in its last
// workgroup (decision rule of src/parallel_tempering.cu:316-352).
#pragma once
#include "pspin_internal.h"

namespace ps {

// Accept test u < exp(delta) for delta < 0, decided exactly as in FP64: an
// FP32 exponential (relative error well under 1e-5 over the whole range,
// flushing to 0 below about -87) settles every u outside a 2e-5 relative
// band around it; inside the band the FP64 exp decides.
__device__ __forceinline__ bool accept_below(double u, double delta) {
    const double e32 = (double)__expf((float)delta);
    if (u < e32 * (1.0 - 2e-5)) return true;
    if (u > e32 * (1.0 + 2e-5)) return false;
    return u < exp(delta);
}

// One wave, up to 64 ladders.  LDS: sE[n] energies by slot, sB[NT] betas,
// sacc[NT] accepted swaps per level, sP[n] replicas by slot (n = NT nrep).
// exchange_stage fills it with nthr threads (loads first, then LDS writes);
// the caller then orders LDS (barrier or wave fence) before exchange_wave.
__device__ inline void exchange_stage(const ExArgs& a, void* lds, int tid, int nthr) {
    const int NT = a.NT, n = NT * a.nrep;
    double* sE = (double*)lds;
    double* sB = sE + n;
    long long* sacc = (long long*)(sB + NT);
    int* sP = (int*)(sacc + NT);
    for (int i0 = tid; i0 < n; i0 += 4 * nthr) {
        int p[4];
        double e[4];
#pragma unroll
        for (int k = 0; k < 4; k++) p[k] = i0 + k * nthr < n ? a.perm[i0 + k * nthr] : 0;
#pragma unroll
        for (int k = 0; k < 4; k++) e[k] = i0 + k * nthr < n ? a.E[p[k]] : 0.0;
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (i0 + k * nthr < n) {
                sP[i0 + k * nthr] = p[k];
                sE[i0 + k * nthr] = e[k];
            }
    }
    for (int i = tid; i < NT; i += nthr) sB[i] = a.beta_sched[i];
}

__device__ inline void exchange_wave(const ExArgs& a, void* lds, int lane) {
    const int NT = a.NT, nr = a.nrep, n = NT * nr;
    const bool valid = lane < nr;
    double* sE = (double*)lds;
    double* sB = sE + n;
    long long* sacc = (long long*)(sB + NT);
    int* sP = (int*)(sacc + NT);
    // Draw of rank `lane` in a level: the lcg48 state X advanced lane + 1 steps.
    uint64_t Ak, Ck;
    lcg48_coeffs((uint64_t)lane + 1, &Ak, &Ck);
    uint64_t X = *a.rng48;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    int pc = valid ? sP[lane] : 0;         // replica carried down the column
    double ec = valid ? sE[lane] : 0.0;    // its energy
    double bt = sB[0];
    int pn = valid ? sP[nr + lane] : 0;    // slot t+1 before level t
    double en = valid ? sE[nr + lane] : 0.0;
    double bn = sB[1];
    for (int t = 0; t < NT - 1; t++) {
        const int pb = pn;
        const double eb = en, b1 = bn;
        if (t + 2 < NT) {  // next level's untouched slot, off the chain
            pn = valid ? sP[(t + 2) * nr + lane] : 0;
            en = valid ? sE[(t + 2) * nr + lane] : 0.0;
            bn = sB[t + 2];
        }
        const double delta = (bt - b1) * (ec - eb);
        const bool need = valid && !(delta >= 0.0);
        const unsigned long long m = __ballot(need);
        const uint64_t x = (Ak * X + Ck) & kMask48;
        const int rank = __popcll(m & lt);
        const uint32_t xlo = (uint32_t)__shfl((int)(uint32_t)x, rank, 64);
        const uint32_t xhi = (uint32_t)__shfl((int)(uint32_t)(x >> 32), rank, 64);
        const int tot = __popcll(m);
        if (tot > 0) {
            const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, tot - 1);
            const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), tot - 1);
            X = ((uint64_t)hi << 32) | lo;
        }
        bool accept = valid;
        if (need) accept = accept_below(lcg48_unit(((uint64_t)xhi << 32) | xlo), delta);
        // slot t keeps the carried replica unless the pair swaps
        const int pf = accept ? pb : pc;
        if (valid) {
            a.perm[t * nr + lane] = pf;
            a.betas[pf] = bt;
        }
        if (!accept) {  // otherwise the carried replica moves down to slot t+1
            pc = pb;
            ec = eb;
        }
        bt = b1;
        const int nacc = __popcll(__ballot(valid && accept));
        if (lane == 0) sacc[t] = nacc;
    }
    if (valid) {
        a.perm[(NT - 1) * nr + lane] = pc;
        a.betas[pc] = bt;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (int t = lane; t < NT - 1; t += 64) {  // all four loads before any store
        const long long c0 = a.ex[0 * (NT - 1) + t], c1 = a.ex[1 * (NT - 1) + t];
        const long long c2 = a.ex[2 * (NT - 1) + t], c3 = a.ex[3 * (NT - 1) + t];
        const long long k = sacc[t];
        a.ex[0 * (NT - 1) + t] = c0 + k;
        a.ex[1 * (NT - 1) + t] = c1 + nr;
        a.ex[2 * (NT - 1) + t] = c2 + k;
        a.ex[3 * (NT - 1) + t] = c3 + nr;
    }
    if (a.series) {
        const long long round = *a.series_n;
        if (valid) a.series[(size_t)(round % a.series_cap) * nr + lane] = ec;
    }
    if (lane == 0) {
        *a.rng48 = X;
        if (a.series) *a.series_n = *a.series_n + 1;
    }
}

}  // namespace ps
