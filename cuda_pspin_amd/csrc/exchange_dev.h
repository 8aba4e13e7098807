// exchange_dev.h -- device side of the replica exchange, shared by the
// exchange kernels and the sweep launch that runs the exchange in its last
// workgroup (decision rule of src/parallel_tempering.cu:316-352).
#pragma once
#include "pspin_internal.h"

namespace ps {

// exp(x) for x <= 0 in FP64: Cody and Waite's reduction x = k ln2 + r (ln2
// in two parts), then the Taylor series of e^r, |r| <= ln2/2, to 18 terms
// (truncation below 1e-19 relative) by Horner's rule; a few ulp from the
// correctly rounded value.  Only draws inside the accept band use it, so its
// speed does not matter; it holds almost no constants, so none are hoisted
// into the exchange loop (where they would be spilled).
__device__ inline double exp_band(double x) {
    if (x < -745.0) return 0.0;
    const double k = rint(x * 1.4426950408889634);
    const double r = fma(-k, 1.90821492927058770002e-10, fma(-k, 6.93147180369123816490e-01, x));
    double p = 1.0;
    for (int i = 18; i >= 1; i--) p = fma(p, r / (double)i, 1.0);
    return ldexp(p, (int)k);
}

// Accept test u < exp(delta) for delta < 0, decided as in FP64: an FP32
// exponential (relative error well under 1e-5 over the whole range, flushing
// to 0 below about -87) settles every u outside a 2e-5 relative band around
// it; inside the band the FP64 exponential decides.
__device__ __forceinline__ bool accept_below(double u, double delta) {
    const double e32 = (double)__expf((float)delta);
    if (u < e32 * (1.0 - 2e-5)) return true;
    if (u > e32 * (1.0 + 2e-5)) return false;
    return u < exp_band(delta);
}

// Exchange data crosses workgroups inside one launch (energies written by
// every workgroup, the permutation and betas by the exchanging one), so it
// moves with agent-scope atomic loads and stores, which no XCD serves from a
// stale cache line.  Ordering comes from agent-scope release RMWs (arrival,
// generation bump) and acquire fences (last arriver, released waiters) in
// the sweep launch's grid barrier.
template <typename T>
__device__ __forceinline__ T dev_load(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void dev_store(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// LDS of the one-wave exchange (n = NT nrep slots): sE[n] energies by slot,
// sEr[n] energies by replica (as loaded), sB[NT] betas, sacc[NT] accepted
// swaps per level, sctr[4 NT] the exchange counters as loaded, sX[n] the
// lcg48 states of every draw the round can take, sX0 the 48-bit state and
// sser the series counter as loaded, sP[n] replicas by slot.
struct ExLds {
    double* sE;
    double* sEr;
    double* sB;
    long long* sacc;
    long long* sctr;
    uint64_t* sX;
    uint64_t* sX0;
    long long* sser;
    int* sP;
    __device__ ExLds(void* lds, int NT, int n) {
        sE = (double*)lds;
        sEr = sE + n;
        sB = sEr + n;
        sacc = (long long*)(sB + NT);
        sctr = sacc + NT;
        sX = (uint64_t*)(sctr + 4 * NT);
        sX0 = sX + n;
        sser = (long long*)(sX0 + 1);
        sP = (int*)(sser + 1);
    }
};

// exchange_stage fills the slots with nthr threads.  Everything the round
// reads from global memory is requested at once, in one round trip: the
// permutation, every replica's energy (by replica, so no load waits for the
// permutation), the betas, the exchange counters, the 48-bit state and the
// series counter.  The energies are then put in slot order in LDS.  It
// contains a workgroup barrier, so all nthr threads call it; the caller then
// orders LDS (barrier or wave fence) before exchange_wave.
__device__ inline void exchange_stage(const ExArgs& a, void* lds, int tid, int nthr) {
    asm volatile("" : "+v"(tid));  // nothing thread-derived hoisted out of a caller's loop
    const int NT = a.NT, n = NT * a.nrep;
    ExLds L(lds, NT, n);
    for (int i0 = tid; i0 < n; i0 += 4 * nthr) {
        int p[4];
        double e[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int i = i0 + k * nthr;
            p[k] = i < n ? dev_load(a.perm + i) : 0;
            e[k] = i < n ? dev_load(a.E + i) : 0.0;
        }
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (i0 + k * nthr < n) {
                L.sP[i0 + k * nthr] = p[k];
                L.sEr[i0 + k * nthr] = e[k];
            }
    }
    for (int i = tid; i < NT; i += nthr) L.sB[i] = a.beta_sched[i];
    for (int i = tid; i < 4 * (NT - 1); i += nthr) L.sctr[i] = dev_load(a.ex + i);
    if (tid == 0) {
        *L.sX0 = dev_load(a.rng48);
        *L.sser = a.series ? dev_load(a.series_n) : 0;
    }
    __syncthreads();
    for (int i = tid; i < n; i += nthr) L.sE[i] = L.sEr[L.sP[i]];
}

__device__ __forceinline__ uint64_t readlane64(uint64_t v, int l) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l);
    return ((uint64_t)hi << 32) | lo;
}

// The level chain: lane r carries ladder r's moving energy and replica; a
// level's draws are read by rank from the precomputed states, so the chain
// holds no permute, no 64-bit product and no global access.  The result goes
// out after the last level.
__device__ inline void exchange_wave(const ExArgs& a, void* lds, int lane) {
    // An opaque copy of the lane index: in a kernel that runs the exchange
    // between sweeps, nothing lane-derived here may be hoisted out of that loop
    // (it would stay live, and spilled, through every sweep).
    asm volatile("" : "+v"(lane));
    const int NT = a.NT, nr = a.nrep, n = NT * nr;
    const bool valid = lane < nr;
    ExLds L(lds, NT, n);
    const uint64_t X0 = *L.sX0;
    {
        // X_{k+1} for k = 0 .. nr (NT - 1) - 1: lane l jumps l + 1 steps, then
        // 64 steps at a time (lane 63's coefficients).
        uint64_t A, C;
        lcg48_coeffs((uint64_t)lane + 1, &A, &C);
        const uint64_t A64 = readlane64(A, 63), C64 = readlane64(C, 63);
        uint64_t x = (A * X0 + C) & kMask48;
        for (int k = lane; k < nr * (NT - 1); k += 64) {
            L.sX[k] = x;
            x = (A64 * x + C64) & kMask48;
        }
    }
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    int pc = valid ? L.sP[lane] : 0;         // replica carried down the column
    double ec = valid ? L.sE[lane] : 0.0;    // its energy
    double bt = L.sB[0];
    int pn = valid ? L.sP[nr + lane] : 0;    // slot t+1 before level t
    double en = valid ? L.sE[nr + lane] : 0.0;
    double bn = L.sB[1];
    int base = 0;  // draws taken so far
    for (int t = 0; t < NT - 1; t++) {
        const int pb = pn;
        const double eb = en, b1 = bn;
        if (t + 2 < NT) {  // next level's untouched slot, off the chain
            pn = valid ? L.sP[(t + 2) * nr + lane] : 0;
            en = valid ? L.sE[(t + 2) * nr + lane] : 0.0;
            bn = L.sB[t + 2];
        }
        const double delta = (bt - b1) * (ec - eb);
        const bool need = valid && !(delta >= 0.0);
        const unsigned long long m = __ballot(need);
        bool accept = valid;
        if (need) accept = accept_below(lcg48_unit(L.sX[base + __popcll(m & lt)]), delta);
        base += __popcll(m);
        // slot t keeps the carried replica unless the pair swaps (slot t is not
        // read again: the prefetch is two slots ahead)
        if (valid) L.sP[t * nr + lane] = accept ? pb : pc;
        if (!accept) {  // otherwise the carried replica moves down to slot t+1
            pc = pb;
            ec = eb;
        }
        bt = b1;
        const int nacc = __popcll(__ballot(valid && accept));
        if (lane == 0) L.sacc[t] = nacc;
    }
    if (valid) L.sP[(NT - 1) * nr + lane] = pc;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (int i = lane; i < n; i += 64) {
        const int p = L.sP[i];
        dev_store(a.perm + i, p);
        dev_store(a.betas + p, L.sB[i / nr]);
    }
    for (int t = lane; t < NT - 1; t += 64) {  // the counters as staged, plus this round
        long long* ex = a.ex;
        const long long k = L.sacc[t];
        dev_store(ex + t, L.sctr[t] + k);
        dev_store(ex + (NT - 1) + t, L.sctr[(NT - 1) + t] + nr);
        dev_store(ex + 2 * (NT - 1) + t, L.sctr[2 * (NT - 1) + t] + k);
        dev_store(ex + 3 * (NT - 1) + t, L.sctr[3 * (NT - 1) + t] + nr);
    }
    if (a.series) {
        const long long round = *L.sser;
        if (valid) dev_store(a.series + (size_t)(round % a.series_cap) * nr + lane, ec);
    }
    if (lane == 0) {
        dev_store(a.rng48, base > 0 ? L.sX[base - 1] : X0);
        if (a.series) dev_store(a.series_n, *L.sser + 1);
    }
}

}  // namespace ps
