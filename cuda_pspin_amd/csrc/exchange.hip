// exchange.hip -- parallel-tempering replica exchange on gfx950 (the
// reference does it on the host between sweeps, src/parallel_tempering.cu:316-352).
//
// Decisions are identical to the reference's for the same energies and the
// same drand48 stream: temperature pairs t = 0..NT-2 in order (hot to cold),
// within each t the ladders r = 0..nrep-1 in order, delta =
// (beta_t - beta_{t+1}) (E_a - E_b), accept when delta >= 0, otherwise draw
// u = drand48() and accept when u < exp(delta).  One wavefront serves 64
// ladders per pass: a ballot over "needs a draw" gives each lane its rank in
// the reference's draw order, and lcg48_skip jumps the 48-bit state to it.
//
// Up to 64 ladders (exchange_wave, run by k_pt_exchange or by the last
// workgroup of a sweep launch): slots are staged in LDS once, lane r carries
// ladder r's moving energy and replica down the levels in registers (level t
// only swaps slots t and t+1, so slot t is final after level t and the values
// of slot t+1 before it are the untouched ones).  While staging, the threads
// compute in parallel every draw the round can take (X_{k+1} = A_{k+1} X + C_{k+1})
// and store logf of each; a level reads its draws by rank and compares their
// logs with delta, so the chain from one level to the next holds a ballot,
// one LDS read and a compare.  A draw within 3e-5 of the decision boundary in
// log space is decided by the FP64 test on the draw itself.  The results go
// out with every thread (exchange_commit).
// More ladders (k_pt_exchange_wide) walk column blocks of 64 in LDS.
#include "exchange_dev.h"

namespace ps {


__global__ __launch_bounds__(64) void k_pt_exchange_wide(ExArgs a) {
    // LDS: the permutation and the energies by slot (sEp[s] = E[perm[s]],
    // swapped together), betas, per-pair accept counts.  Lane r owns ladder
    // column r at every level, so levels need no barrier.
    extern __shared__ double lds[];
    const int lane = threadIdx.x;
    const int NT = a.NT, nr = a.nrep, n = NT * nr;
    double* sEp = lds;
    double* sB = lds + n;
    long long* sacc = (long long*)(lds + n + NT);  // [NT-1] accepted swaps this round
    int* sperm = (int*)(sacc + NT);
    for (int i = lane; i < n; i += 64) {
        const int p = a.perm[i];
        sperm[i] = p;
        sEp[i] = a.E[p];
    }
    for (int i = lane; i < NT; i += 64) sB[i] = a.beta_sched[i];
    // Jump table: k steps of the 48-bit LCG, k = 0..64, as one multiply-add.
    __shared__ uint64_t jA[65], jC[65];
    lcg48_coeffs((uint64_t)lane + 1, &jA[lane + 1], &jC[lane + 1]);
    if (lane == 0) { jA[0] = 1; jC[0] = 0; }
    __syncthreads();
    uint64_t X = *a.rng48;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int t = 0; t < NT - 1; t++) {
        const double db = sB[t] - sB[t + 1];
        long long nacc = 0;
        for (int c0 = 0; c0 < nr; c0 += 64) {
            const int r = c0 + lane;
            const bool valid = r < nr;
            const int sa = t * nr + r, sb = sa + nr;
            double ea = 0.0, eb = 0.0, delta = 0.0;
            if (valid) {
                ea = sEp[sa];
                eb = sEp[sb];
                delta = db * (ea - eb);
            }
            const bool need = valid && !(delta >= 0.0);
            const unsigned long long m = __ballot(need);
            bool accept = valid;
            if (need) {
                const int k = __popcll(m & lt) + 1;
                const uint64_t x = (jA[k] * X + jC[k]) & kMask48;
                accept = accept_below(lcg48_unit(x), delta);
            }
            const int tot = __popcll(m);
            X = (jA[tot] * X + jC[tot]) & kMask48;
            if (valid && accept) {
                const int tmp = sperm[sa];
                sperm[sa] = sperm[sb];
                sperm[sb] = tmp;
                sEp[sa] = eb;
                sEp[sb] = ea;
            }
            nacc += __popcll(__ballot(valid && accept));
        }
        if (lane == 0) sacc[t] = nacc;
    }
    __syncthreads();
    // One independent read-modify-write per counter, all lanes at once.
    for (int t = lane; t < NT - 1; t += 64) {
        a.ex[0 * (NT - 1) + t] += sacc[t];
        a.ex[1 * (NT - 1) + t] += nr;
        a.ex[2 * (NT - 1) + t] += sacc[t];
        a.ex[3 * (NT - 1) + t] += nr;
    }
    for (int i = lane; i < n; i += 64) {
        a.perm[i] = sperm[i];
        a.betas[sperm[i]] = sB[i / nr];
    }
    if (a.series) {
        const long long round = *a.series_n;
        double* row = a.series + (size_t)(round % a.series_cap) * nr;
        for (int r = lane; r < nr; r += 64) row[r] = sEp[(NT - 1) * nr + r];
    }
    __syncthreads();
    if (lane == 0) {
        *a.rng48 = X;
        if (a.series) *a.series_n += 1;
    }
}

__global__ __launch_bounds__(64) void k_pt_exchange(ExArgs a) {
    extern __shared__ double lds_ex[];
    exchange_stage(a, lds_ex, threadIdx.x, 64);
    __syncthreads();
    exchange_wave(a, lds_ex, threadIdx.x);
    __syncthreads();
    exchange_commit(a, lds_ex, threadIdx.x, 64, -1, nullptr);
}

ExArgs exchange_args(const pspin_state* st) {
    ExArgs a;
    a.NT = st->pt_NT;
    a.nrep = st->pt_nrep;
    a.beta_sched = st->d_beta_sched;
    a.E = st->d_E;
    a.perm = st->d_perm;
    a.betas = st->d_betas;
    a.ex = st->d_ex;
    a.rng48 = st->d_rng48;
    a.series = st->series_cap > 0 ? st->d_series : nullptr;
    a.series_n = st->d_series_n;
    a.series_cap = st->series_cap;
    return a;
}

hipError_t launch_exchange(pspin_state* st) {
    const ExArgs a = exchange_args(st);
    if (a.nrep <= 64) {
        hipLaunchKernelGGL(k_pt_exchange, dim3(1), dim3(64), exchange_lds_bytes(a.NT, a.nrep), st->stream, a);
    } else {
        hipLaunchKernelGGL(k_pt_exchange_wide, dim3(1), dim3(64), exchange_wide_lds_bytes(a.NT, a.nrep), st->stream, a);
    }
    return hipGetLastError();
}

}  // namespace ps
