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
// swaps per level, sctr[4 NT] the exchange counters as loaded, sX0 the
// 48-bit state and sser the series counter as loaded, sAB the accept bits of
// each lane (64 levels per word, 64 words per group of 64 levels), sLU[n]
// logf of every draw the round can take, sP[n + 1] replicas by slot (the
// last entry takes the writes of lanes past the ladders).
__host__ __device__ inline int exchange_ab_words(int NT) { return (NT - 1 + 63) / 64 * 64; }
struct ExLds {
    double* sE;
    double* sEr;
    double* sB;
    long long* sacc;
    long long* sctr;
    uint64_t* sX0;
    long long* sser;
    unsigned long long* sAB;
    float* sLU;
    int* sP;
    __device__ ExLds(void* lds, int NT, int n) {
        sE = (double*)lds;
        sEr = sE + n;
        sB = sEr + n;
        sacc = (long long*)(sB + NT);
        sctr = sacc + NT;
        sX0 = (uint64_t*)(sctr + 4 * NT);
        sser = (long long*)(sX0 + 1);
        sAB = (unsigned long long*)(sser + 1);
        sLU = (float*)(sAB + exchange_ab_words(NT));
        sP = (int*)(sLU + n);
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
// level's draws are read by rank from the precomputed logs, so the chain
// holds no permute, no 64-bit product, no exponential and no global access.
// The result goes out after the last level.
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
        // draw k (k = 0 .. nr (NT - 1) - 1) is X_{k+1}: lane l jumps l + 1
        // steps, then 64 steps at a time (lane 63's coefficients)
        uint64_t A, C;
        lcg48_coeffs((uint64_t)lane + 1, &A, &C);
        const uint64_t A64 = readlane64(A, 63), C64 = readlane64(C, 63);
        uint64_t x = (A * X0 + C) & kMask48;
        for (int k = lane; k < nr * (NT - 1); k += 64) {
            L.sLU[k] = logf((float)lcg48_unit(x));
            x = (A64 * x + C64) & kMask48;
        }
    }
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // The level loop is kept free of divergent branches and of stores that
    // wait on a lane mask (each costs the lone wave issue cycles on every
    // level): lanes past the ladders read ladder nr - 1's slots, unused, and
    // write their slot to the spare entry sP[n]; every lane reads the draw
    // log of its rank; accepts are kept as bits and counted after the loop.
    const int lr = valid ? lane : nr - 1;
    int pc = L.sP[lr];          // replica carried down the column
    double ec = L.sE[lr];       // its energy
    double bt = L.sB[0];
    int pn = L.sP[nr + lr];     // slot t+1 before level t
    double en = L.sE[nr + lr];
    double bn = L.sB[1];
    int base = 0;  // draws taken so far
    unsigned long long abits = 0;  // this lane's accepts, bit t % 64
    for (int t = 0; t < NT - 1; t++) {
        const int pb = pn;
        const double eb = en, b1 = bn;
        const int t2 = t + 2 < NT ? t + 2 : NT - 1;  // next level's untouched slot, off the chain
        pn = L.sP[t2 * nr + lr];
        en = L.sE[t2 * nr + lr];
        bn = L.sB[t2];
        // keep these reads issued here: the end of the level copies them into
        // the next level's registers, and waits for them there
        __builtin_amdgcn_sched_barrier(0);
        const double delta = (bt - b1) * (ec - eb);
        const bool need = valid && !(delta >= 0.0);
        const unsigned long long m = __ballot(need);
        // rank k of this lane's draw (k < n: base + popcount stays within the
        // nr (NT - 1) draws of the round)
        const int k = base + __popcll(m & lt);
        // u < exp(delta) decided from lu = logf((float)u), |lu - log u| < 1e-5
        // (u rounded to FP32, then logf within a few ulp of |lu| <= 34; u = 0
        // gives -inf): outside a 3e-5 band around delta the order of lu and
        // delta is that of u and exp(delta).  Below delta = -700 every u > 0
        // (at least 2^-48) exceeds exp(delta); that case goes to FP64 too.
        const double dd = delta - (double)L.sLU[k];
        // (bitwise: no short-circuit branches)
        bool accept = valid & (!need | (dd > 3e-5));
        const bool band = need & ((delta < -700.0) | (fabs(dd) <= 3e-5));
        // about 6e-5 of draws: the decision in FP64 on the draw itself (kept
        // off the fall-through path, which every level takes)
        if (__builtin_expect(__ballot(band) != 0, 0)) {
            if (band) {
                uint64_t A, C;
                lcg48_coeffs((uint64_t)k + 1, &A, &C);
                accept = lcg48_unit((A * X0 + C) & kMask48) < exp_band(delta);
            }
        }
        base += __popcll(m);
        // slot t keeps the carried replica unless the pair swaps (slot t is not
        // read again: the prefetch is two slots ahead)
        L.sP[valid ? t * nr + lane : n] = accept ? pb : pc;
        // a swap moves the carried replica down to slot t+1; otherwise slot
        // t+1's replica is carried on
        pc = accept ? pc : pb;
        ec = accept ? ec : eb;
        bt = b1;
        abits |= (unsigned long long)accept << (t & 63);
        if (__builtin_expect((t & 63) == 63 || t == NT - 2, 0)) {  // uniform: every 64 levels and at the end
            L.sAB[(t >> 6) * 64 + lane] = abits;
            abits = 0;
        }
    }
    if (valid) L.sP[(NT - 1) * nr + lane] = pc;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // the betas first: in a multi-round launch every other workgroup is
    // waiting for its own
    for (int i = lane; i < n; i += 64) dev_store(a.betas + L.sP[i], L.sB[i / nr]);
    for (int i = lane; i < n; i += 64) dev_store(a.perm + i, L.sP[i]);
    for (int t = lane; t < NT - 1; t += 64) {  // accepted swaps per level
        const unsigned long long* w = L.sAB + (t >> 6) * 64;
        long long c = 0;
        for (int r = 0; r < nr; r++) c += (w[r] >> (t & 63)) & 1ull;
        L.sacc[t] = c;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
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
        uint64_t A, C;  // the state after the round's base draws
        lcg48_coeffs((uint64_t)base, &A, &C);
        dev_store(a.rng48, (A * X0 + C) & kMask48);
        if (a.series) dev_store(a.series_n, *L.sser + 1);
    }
}

}  // namespace ps
