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
// sEr[n] energies by replica as loaded (after the chain: the energy that ends
// in slot NT - 1 of each ladder), sB[NT] betas, sacc[NT] accepted swaps per
// level, sctr[4 NT] the exchange counters as loaded, sX0 the 48-bit state and
// sser the series counter as loaded, sDraws the draws the round took, sLU[n]
// logf of every draw the round can take, sP[n] replicas by slot with a row of
// nr entries before slot 0 and one after slot NT - 1 (the chain's writes of
// "slot -1", its writes from lanes past the ladders, its read past the last
// slot).
struct ExLds {
    double* sE;
    double* sEr;
    double* sB;
    long long* sacc;
    long long* sctr;
    uint64_t* sX0;
    long long* sser;
    long long* sDraws;
    float* sLU;
    int* sP;
    __device__ ExLds(void* lds, int NT, int nr) {
        const int n = NT * nr;
        sE = (double*)lds;
        sEr = sE + n;
        sB = sEr + n;
        sacc = (long long*)(sB + NT);
        sctr = sacc + NT;
        sX0 = (uint64_t*)(sctr + 4 * NT);
        sser = (long long*)(sX0 + 1);
        sDraws = sser + 1;
        sLU = (float*)(sDraws + 1);
        sP = (int*)(sLU + n) + nr;
    }
};

// Replica r's energy from its two tagged words (see exchange_stage); at most
// `limit` polls.
__device__ inline double tagged_energy(const unsigned long long* ew, int r, unsigned tag, unsigned* expired,
                                       unsigned limit) {
    unsigned long long lo = 0, hi = 0;
    for (unsigned n = 0;; n++) {
        lo = dev_load(ew + 2 * r);
        hi = dev_load(ew + 2 * r + 1);
        if ((unsigned)(lo >> 32) == tag && (unsigned)(hi >> 32) == tag) break;
        if (n == limit) {
            __hip_atomic_fetch_or(expired, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
        }
        __builtin_amdgcn_s_sleep(2);
    }
    return __longlong_as_double((long long)((hi << 32) | (lo & 0xffffffffull)));
}

// exchange_stage fills the slots with nthr threads.  Everything the round
// reads from global memory is requested at once, in one round trip: the
// permutation, every replica's energy (by replica, so no load waits for the
// permutation), the betas, the exchange counters, the 48-bit state and the
// series counter.  Meanwhile the threads compute logf of every draw the round
// can take.  The energies are then put in slot order in LDS.  It contains a
// workgroup barrier, so all nthr threads call it; the caller then orders LDS
// (barrier) before exchange_wave.
//
// With ew, the energies come as tagged words instead of from a.E: replica r
// publishes {tag, low half} in ew[2r] and {tag, high half} in ew[2r + 1]; a
// thread polls a replica's two words until both carry this round's tag (a
// bounded wait: past `limit` polls it sets *expired and takes what it read).
__device__ inline void exchange_stage(const ExArgs& a, void* lds, int tid, int nthr,
                                      const unsigned long long* ew = nullptr, unsigned tag = 0,
                                      unsigned* expired = nullptr, unsigned limit = 0) {
    asm volatile("" : "+v"(tid));  // nothing thread-derived hoisted out of a caller's loop
    const int NT = a.NT, nr = a.nrep, n = NT * nr, nd = nr * (NT - 1), nc = 4 * (NT - 1);
    ExLds L(lds, NT, nr);
    // every load of the round is issued before any result is used
    const uint64_t X0 = dev_load(a.rng48);
    int p[4];
    double e[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int i = tid + k * nthr;
        p[k] = i < n ? dev_load(a.perm + i) : 0;
        e[k] = i < n && !ew ? dev_load(a.E + i) : 0.0;
    }
    const double bs = tid < NT ? a.beta_sched[tid] : 0.0;
    long long c[2];
#pragma unroll
    for (int k = 0; k < 2; k++) c[k] = tid + k * nthr < nc ? dev_load(a.ex + tid + k * nthr) : 0;
    const long long ser = tid == 0 && a.series ? dev_load(a.series_n) : 0;
    // logf of every draw the round can take, while the loads are in flight:
    // draw k (k = 0 .. nd - 1) is X_{k+1}; thread tid jumps tid + 1 steps,
    // then nthr steps at a time
    if (tid < nd) {
        uint64_t A, C, An = 1, Cn = 0;
        lcg48_coeffs((uint64_t)tid + 1, &A, &C);
        if (nd > nthr) lcg48_coeffs((uint64_t)nthr, &An, &Cn);
        uint64_t x = (A * X0 + C) & kMask48;
        for (int k = tid; k < nd; k += nthr) {
            L.sLU[k] = logf((float)lcg48_unit(x));
            x = (An * x + Cn) & kMask48;
        }
    }
    if (ew)
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (tid + k * nthr < n) e[k] = tagged_energy(ew, tid + k * nthr, tag, expired, limit);
#pragma unroll
    for (int k = 0; k < 4; k++)
        if (tid + k * nthr < n) {
            L.sP[tid + k * nthr] = p[k];
            L.sEr[tid + k * nthr] = e[k];
        }
    if (tid < NT) L.sB[tid] = bs;
#pragma unroll
    for (int k = 0; k < 2; k++)
        if (tid + k * nthr < nc) L.sctr[tid + k * nthr] = c[k];
    if (tid == 0) {
        *L.sX0 = X0;
        *L.sser = ser;
    }
    // what the first batches did not cover (small workgroups only)
    for (int i = tid + 4 * nthr; i < n; i += nthr) {
        L.sP[i] = dev_load(a.perm + i);
        L.sEr[i] = ew ? tagged_energy(ew, i, tag, expired, limit) : dev_load(a.E + i);
    }
    for (int i = tid + nthr; i < NT; i += nthr) L.sB[i] = a.beta_sched[i];
    for (int i = tid + 2 * nthr; i < nc; i += nthr) L.sctr[i] = dev_load(a.ex + i);
    __syncthreads();
    for (int i = tid; i < n; i += nthr) L.sE[i] = L.sEr[L.sP[i]];
}

// The level chain (wave 0; lane r carries ladder r).  Level t's decision
// needs the draws' ranks, so the chain from one level to the next runs through
// the ballot of who needs a draw, the read of that draw's log and the compare.
// Everything else is taken off that path: each level computes both candidate
// deltas of the next one (the carried energy continues if this level swaps,
// slot t+1's original is carried if not) and their need masks, and only picks
// between them once this level's decision is known.  Accepted swaps are
// counted per level by popcount; the slots, the counts, the draws taken and
// the energy that ends in slot NT - 1 of each ladder go to LDS for
// exchange_commit.
__device__ inline void exchange_wave(const ExArgs& a, void* lds, int lane) {
    // An opaque copy of the lane index: in a kernel that runs the exchange
    // between sweeps, nothing lane-derived here may be hoisted out of that loop
    // (it would stay live, and spilled, through every sweep).
    asm volatile("" : "+v"(lane));
    const int NT = a.NT, nr = a.nrep, n = NT * nr;
    const bool valid = lane < nr;
    ExLds L(lds, NT, nr);
    const uint64_t X0 = *L.sX0;
    // The one wave that runs the chain takes issue priority over the
    // workgroup's other waves (which draw the next sweep meanwhile).
    __builtin_amdgcn_s_setprio(3);
    // Lanes past the ladders read ladder nr - 1's slots, need no draw, never
    // swap, and write their slots to the row after the last slot.  Lane sets
    // (who needs a draw, who swaps) are kept as 64-bit masks, combined by
    // scalar instructions.
    const int lr = valid ? lane : nr - 1;
    const unsigned long long vmask = __ballot(valid);
    // Level t's inputs come as two candidates chosen by the previous decision
    // (mask accm): A if level t-1 swapped (the replica carried into t-1 goes
    // on), R if not (slot t's original is carried).  Level 0 has one candidate.
    int pcA = L.sP[lr];
    double ecA = L.sE[lr];
    int pb = L.sP[nr + lr];  // slot t+1's original, the partner at level t
    double eb = L.sE[nr + lr];
    double b1 = L.sB[1];
    double dA = (L.sB[0] - b1) * (ecA - eb);
    unsigned long long mA = __ballot(!(dA >= 0.0)) & vmask;  // lanes needing a draw
    int pcR = pcA;
    double ecR = ecA, dR = dA;
    unsigned long long mR = mA;
    unsigned long long accm = ~0ull;
    int base = 0;  // draws taken before this level
    // Running LDS pointers: the slot written at level t is t - 1 (level 0
    // writes the row before slot 0); the slot read ahead is t + 2 (the last
    // level reads past slot NT - 1: the row after it, sEr and sacc, unused).
    int* pw = valid ? L.sP + lane - nr : L.sP + n;
    const int sw = valid ? nr : 0;
    const int* pp = L.sP + 2 * nr + lr;
    const double* pe = L.sE + 2 * nr + lr;
    const double* pbt = L.sB + 2;
    for (int t0 = 0; t0 < NT - 1; t0 += 64) {
        const int t1 = NT - 1 < t0 + 64 ? NT - 1 : t0 + 64;
        int cnt = 0;  // accepted swaps of level t0 + l, in lane l
        for (int t = t0; t < t1; t++) {
            // slot t+2's original, for the next level's candidates (read first,
            // so that the wait for this level's draw covers it)
            const int pn = *pp;
            const double en = *pe;
            const double b2 = *pbt;
            pp += nr;
            pe += nr;
            pbt++;
            __builtin_amdgcn_sched_barrier(0);
            const unsigned long long m = (accm & mA) | (~accm & mR);
            // rank k of this lane's draw (k < n: base + popcount stays within
            // the nr (NT - 1) draws of the round)
            const int k = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                                         __builtin_amdgcn_mbcnt_lo((unsigned)m, (unsigned)base));
            // the previous level's slot goes out before the draw is read: LDS
            // completes in order, so the wait for the draw covers the store and
            // the next level starts with nothing outstanding
            const bool acc = __builtin_amdgcn_inverse_ballot_w64(accm);
            *pw = acc ? pcR : pcA;
            pw += sw;
            __builtin_amdgcn_sched_barrier(0);
            const float lu = L.sLU[k];
            __builtin_amdgcn_sched_barrier(0);
            // off the chain: the previous decision's picks and both candidates
            // of the next level
            const int pc = acc ? pcA : pcR;
            const double ec = acc ? ecA : ecR;
            const double delta = acc ? dA : dR;
            const double nA = (b1 - b2) * (ec - en);
            const double nR = (b1 - b2) * (eb - en);
            const unsigned long long nmA = __ballot(!(nA >= 0.0)) & vmask;
            const unsigned long long nmR = __ballot(!(nR >= 0.0)) & vmask;
            // u < exp(delta) decided from lu = logf((float)u), |lu - log u| <
            // 1e-5 (u rounded to FP32, then logf within a few ulp of |lu| <= 34;
            // u = 0 gives -inf): outside a 3e-5 band around delta the order of
            // lu and delta is that of u and exp(delta).  Below delta = -700
            // every u > 0 (at least 2^-48) exceeds exp(delta); that case goes to
            // FP64 too.
            const double dd = delta - (double)lu;
            unsigned long long am = (~m | __ballot(dd > 3e-5)) & vmask;
            const unsigned long long bm = m & (__ballot(delta < -700.0) | __ballot(fabs(dd) <= 3e-5));
            // about 6e-5 of draws: the decision in FP64 on the draw itself
            if (__builtin_expect(bm != 0, 0)) {
                bool accept = __builtin_amdgcn_inverse_ballot_w64(am);
                if (__builtin_amdgcn_inverse_ballot_w64(bm)) {
                    uint64_t A, C;
                    lcg48_coeffs((uint64_t)k + 1, &A, &C);
                    accept = lcg48_unit((A * X0 + C) & kMask48) < exp_band(delta);
                }
                am = __ballot(accept) & vmask;
            }
            base += __popcll(m);
            cnt = lane == t - t0 ? __popcll(am) : cnt;
            // this level's outcomes become the next level's candidates
            pcA = pc;
            ecA = ec;
            pcR = pb;
            ecR = eb;
            dA = nA;
            dR = nR;
            mA = nmA;
            mR = nmR;
            pb = pn;
            eb = en;
            b1 = b2;
            accm = am;
        }
        if (t0 + lane < t1) L.sacc[t0 + lane] = cnt;
    }
    const bool acc = __builtin_amdgcn_inverse_ballot_w64(accm);
    *pw = acc ? pcR : pcA;  // slot NT - 2
    if (valid) {
        L.sP[(NT - 1) * nr + lane] = acc ? pcA : pcR;
        L.sEr[lane] = acc ? ecA : ecR;  // the energy that ends in slot NT - 1
    }
    if (lane == 0) *L.sDraws = base;
    __builtin_amdgcn_s_setprio(0);
}

// exchange_commit writes the round out with nthr threads, after a barrier
// that follows exchange_wave: every replica's new beta first (in a
// multi-round launch the other workgroups wait for their own), then the
// permutation, the exchange counters, the T_min series and the 48-bit state
// after the draws taken.  The thread that writes replica `self`'s beta also
// puts it in *self_beta (self < 0: none).  With lvw, each replica's new level
// goes out first, as one word {tag, level} in lvw[replica].
__device__ inline void exchange_commit(const ExArgs& a, void* lds, int tid, int nthr, int self,
                                       double* self_beta, unsigned long long* lvw = nullptr,
                                       unsigned tag = 0) {
    asm volatile("" : "+v"(tid));  // nothing thread-derived hoisted out of a caller's loop
    const int NT = a.NT, nr = a.nrep, n = NT * nr;
    ExLds L(lds, NT, nr);
    for (int i = tid; i < n; i += nthr) {
        const int p = L.sP[i];
        const int t = i / nr;
        if (lvw) dev_store(lvw + p, ((unsigned long long)tag << 32) | (unsigned)t);
        const double b = L.sB[t];
        dev_store(a.betas + p, b);
        if (p == self) *self_beta = b;
    }
    for (int i = tid; i < n; i += nthr) dev_store(a.perm + i, L.sP[i]);
    for (int t = tid; t < NT - 1; t += nthr) {  // the counters as staged, plus this round
        long long* ex = a.ex;
        const long long k = L.sacc[t];
        dev_store(ex + t, L.sctr[t] + k);
        dev_store(ex + (NT - 1) + t, L.sctr[(NT - 1) + t] + nr);
        dev_store(ex + 2 * (NT - 1) + t, L.sctr[2 * (NT - 1) + t] + k);
        dev_store(ex + 3 * (NT - 1) + t, L.sctr[3 * (NT - 1) + t] + nr);
    }
    if (a.series)
        for (int r = tid; r < nr; r += nthr)
            dev_store(a.series + (size_t)(*L.sser % a.series_cap) * nr + r, L.sEr[r]);
    if (tid == 0) {
        uint64_t A, C;  // the state after the round's draws
        lcg48_coeffs((uint64_t)*L.sDraws, &A, &C);
        dev_store(a.rng48, (A * *L.sX0 + C) & kMask48);
        if (a.series) dev_store(a.series_n, *L.sser + 1);
    }
}

}  // namespace ps
