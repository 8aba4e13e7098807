// batch.hip -- many independent sparse (smoothed-cube) PT systems of uneven N
// in one launch on gfx950.
//
// The reference runs one disorder sample per process
// (src/parallel_tempering_sparse.cu, launched per sample by scan_samples_pt).
// Each sparse step touches only ~4N/N = 4 quartet terms per site plus the
// few FMC-surviving g2 neighbours, so a workgroup per replica (as in
// k_sweep) spends its time in barriers.  Here one workgroup owns one disorder
// sample and one LANE owns one chain (replica x temperature) of it: a chain
// steps through its sweep with no barriers, reading its own FP64 amplitudes
// from LDS and the sample's per-site term lists from LDS.  After every
// pt_freq sweeps the workgroup meets once, wave 0 runs the replica exchange
// with the reference's decision order and drand48 stream for that sample,
// and every chain picks up its new beta.
//
// Per chain the arithmetic is the dense kernel's: FP64 cube proposal
// (include/spins_sparse.h:26-56), FP32 midpoint local-field Delta E
// (exact for the multilinear H, see sweep.hip), threshold Metropolis, FP64
// running energy, FP64 resync every resync_every sweeps from FP64 terms.
// Randomness is the same Philox stream as k_sweep (replica seed, sweep
// counter, step), so a sample here follows the same chains as a single
// PSPIN_SPARSE state with the same seed, up to FP32 summation order.
#include "pspin_internal.h"
#include "pspin_batch.h"

namespace ps {

__device__ __forceinline__ float2 f2(double2 v) { return make_float2((float)v.x, (float)v.y); }

// LDS layout (dynamic): FP64 amplitudes [N][C] (site-major so lanes of one
// site are adjacent), H4 entries, H2 entries, offsets, exchange scratch.
__global__ __launch_bounds__(256) void k_batch(BatchArgs a) {
    extern __shared__ double lds[];
    const BatchSample S = a.samples[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63;
    const int N = S.N, C = S.chains, NT = a.NT, nr = C / NT;
    const bool live = tid < C;
    double2* amp = (double2*)lds;                       // [N][C]
    uint2* e4 = (uint2*)(amp + (size_t)N * C);          // H4 entries
    uint2* e2 = e4 + S.n_e4;                            // H2 entries {k, g}
    int* o4 = (int*)(e2 + S.n_e2);                      // [N+1]
    int* o2 = o4 + N + 1;                               // [N+1]
    double* sE = (double*)(((uintptr_t)(o2 + N + 1) + 7) & ~(uintptr_t)7);  // [C]
    double* sB = sE + C;                                // [C] beta by chain
    int* sperm = (int*)(sB + C);                        // [C]
    uint64_t* jA = (uint64_t*)(((uintptr_t)(sperm + C) + 7) & ~(uintptr_t)7);  // [65]
    uint64_t* jC = jA + 65;
    long long* sacc = (long long*)(jC + 65);            // [NT]

    // ---- stage the sample's terms and the chains' amplitudes
    for (int i = tid; i < S.n_e4; i += blockDim.x) e4[i] = a.e4[S.e4_off + i];
    for (int i = tid; i < S.n_e2; i += blockDim.x) e2[i] = a.e2[S.e2_off + i];
    for (int i = tid; i <= N; i += blockDim.x) {
        o4[i] = a.o4[S.o_off + i];
        o2[i] = a.o2[S.o_off + i];
    }
    if (live)
        for (int x = 0; x < N; x++) {
            const double* g = a.spins + S.spin_off + ((size_t)tid * N + x) * 2;
            amp[x * C + tid] = make_double2(g[0], g[1]);
        }
    if (tid < C) sperm[tid] = a.perm[S.chain_off + tid];
    if (tid < 64) lcg48_coeffs((uint64_t)tid + 1, &jA[tid + 1], &jC[tid + 1]);
    if (tid == 0) {
        jA[0] = 1;
        jC[0] = 0;
    }
    const int ch = S.chain_off + tid;
    double beta = live ? a.betas[ch] : 0.0;
    double E = live ? a.E[ch] : 0.0;
    const uint64_t seed = live ? a.seeds[ch] : 0;
    uint64_t sweep = live ? a.ctr[ch] : 0;
    long long nacc = 0, nprop = 0;
    int since = live ? a.since[ch] : 0;
    uint64_t X = (tid < 64) ? a.rng48[blockIdx.x] : 0;  // wave 0 keeps the sample's drand48 state
    const long long done0 = a.done[blockIdx.x];
    __syncthreads();

    const int nsteps = N / 2;
    for (int sw = 0; sw < a.n_sweeps; sw++) {
        if (live) {
            const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
            for (int step = 0; step < nsteps; step++) {
                const U4 x = philox4x32_10(U4{(uint32_t)sweep, (uint32_t)(sweep >> 32), (uint32_t)step, kCtrStep0}, k0, k1);
                const U4 y = philox4x32_10(U4{(uint32_t)sweep, (uint32_t)(sweep >> 32), (uint32_t)step, kCtrStep1}, k0, k1);
                // pair pick (src/mc_sparse.cu:118-126)
                int ri = (int)(u32_to_unit(x.x) * N);
                if (ri >= N) ri = N - 1;
                int rj = (int)(u32_to_unit(x.y) * (N - 1));
                if (rj >= N - 1) rj = N - 2;
                if (rj >= ri) rj++;
                const int i0 = ri < rj ? ri : rj, j0 = ri < rj ? rj : ri;
                // cube proposal (include/spins_sparse.h:26-56)
                const double2 ai = amp[i0 * C + tid], aj = amp[j0 * C + tid];
                const double u = u32_to_unit(x.z);
                double s1, c1, s2, c2;
                sincos(2.0 * M_PI * u32_to_unit(x.w), &s1, &c1);
                sincos(2.0 * M_PI * u32_to_unit(y.x), &s2, &c2);
                const double r1 = ai.x * ai.x + ai.y * ai.y, r2 = aj.x * aj.x + aj.y * aj.y;
                const double sq = sqrt(r1 * r1 + r2 * r2);
                const double m1 = sqrt(sq * sqrt(1.0 - u)), m2 = sqrt(sq * sqrt(u));
                const double2 ni = make_double2(m1 * c1, m1 * s1), nj = make_double2(m2 * c2, m2 * s2);
                const float2 di = make_float2((float)(ni.x - ai.x), (float)(ni.y - ai.y));
                const float2 dj = make_float2((float)(nj.x - aj.x), (float)(nj.y - aj.y));
                const float2 mi = make_float2((float)(0.5 * (ai.x + ni.x)), (float)(0.5 * (ai.y + ni.y)));
                const float2 mj = make_float2((float)(0.5 * (aj.x + nj.x)), (float)(0.5 * (aj.y + nj.y)));
                // local fields at the midpoint configuration
                float dE = 0.f;
                for (int pass = 0; pass < 2; pass++) {
                    const int site = pass ? j0 : i0, other = pass ? i0 : j0;
                    const float2 mo = pass ? mi : mj;
                    float2 F = make_float2(0.f, 0.f);
                    for (int e = o2[site]; e < o2[site + 1]; e++) {
                        const uint2 en = e2[e];
                        const int k = (int)en.x;
                        const float g = __uint_as_float(en.y);
                        const float2 t = (k == other) ? mo : f2(amp[k * C + tid]);
                        F.x += g * t.x;
                        F.y -= g * t.y;
                    }
                    for (int e = o4[site]; e < o4[site + 1]; e++) {
                        const uint2 en = e4[e];
                        const int p = en.x & 0xff, m = (en.x >> 8) & 0xff, n = (en.x >> 16) & 0xff;
                        const float g = __uint_as_float(en.y);
                        const float2 tp = (p == other) ? mo : f2(amp[p * C + tid]);
                        const float2 tm = (m == other) ? mo : f2(amp[m * C + tid]);
                        const float2 tn = (n == other) ? mo : f2(amp[n * C + tid]);
                        const float2 mn = make_float2(tm.x * tn.x - tm.y * tn.y, tm.x * tn.y + tm.y * tn.x);
                        F.x += g * (tp.x * mn.x + tp.y * mn.y);
                        F.y += g * (tp.y * mn.x - tp.x * mn.y);
                    }
                    const float2 d = pass ? dj : di;
                    dE -= d.x * F.x - d.y * F.y;
                }
                const double thr = (beta == 0.0) ? INFINITY : -log(u32_to_unit(y.y)) / beta;
                nprop++;
                if (beta == 0.0 || (double)dE < thr) {
                    amp[i0 * C + tid] = ni;
                    amp[j0 * C + tid] = nj;
                    E += (double)dE;
                    nacc++;
                }
            }
            sweep++;
            since++;
            if (a.resync_every > 0 && since >= a.resync_every) {
                // FP64 resync from the sample's FP64 terms (src/mc_sparse.cu:258-315, i<j H2)
                double e = 0.0;
                for (int t = 0; t < S.n_q; t++) {
                    const int* v = a.quartets + (size_t)5 * (S.q_off + t);
                    const double2 A0 = amp[v[0] * C + tid], A1 = amp[v[1] * C + tid];
                    const double2 A2 = amp[v[2] * C + tid], A3 = amp[v[3] * C + tid];
                    const int cm = conj_mask(v[4]);
                    const double2 s1_ = (cm & 2) ? make_double2(A1.x, -A1.y) : A1;
                    const double2 s2_ = (cm & 4) ? make_double2(A2.x, -A2.y) : A2;
                    const double2 s3_ = (cm & 8) ? make_double2(A3.x, -A3.y) : A3;
                    const double2 p01 = make_double2(A0.x * s1_.x - A0.y * s1_.y, A0.x * s1_.y + A0.y * s1_.x);
                    const double2 p23 = make_double2(s2_.x * s3_.x - s2_.y * s3_.y, s2_.x * s3_.y + s2_.y * s3_.x);
                    e += -a.qg[S.q_off + t] * (p01.x * p23.x - p01.y * p23.y);
                }
                for (int t = 0; t < S.n_h2; t++) {
                    const int2 ij = a.h2ij[S.h2_off + t];
                    const double2 A = amp[ij.x * C + tid], B = amp[ij.y * C + tid];
                    e += -a.h2g[S.h2_off + t] * (A.x * B.x + A.y * B.y);
                }
                E = e;
                since = 0;
            }
        }
        const bool exch = a.pt_freq > 0 && (done0 + sw + 1) % a.pt_freq == 0;
        if (exch) {
            // Exchange (src/parallel_tempering.cu:316-352) for this sample.
            if (live) sE[tid] = E;
            __syncthreads();
            if (tid < 64) {
                const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
                for (int t = 0; t < NT - 1; t++) {
                    const double db = a.beta_sched[t] - a.beta_sched[t + 1];
                    long long na = 0;
                    for (int c0 = 0; c0 < nr; c0 += 64) {
                        const int r = c0 + lane;
                        const bool valid = r < nr;
                        double delta = 0.0;
                        if (valid) delta = db * (sE[sperm[t * nr + r]] - sE[sperm[(t + 1) * nr + r]]);
                        const bool need = valid && !(delta >= 0.0);
                        const unsigned long long m = __ballot(need);
                        bool accept = valid;
                        if (need) {
                            const int k = __popcll(m & lt) + 1;
                            accept = lcg48_unit((jA[k] * X + jC[k]) & kMask48) < exp(delta);
                        }
                        const int tot = __popcll(m);
                        X = (jA[tot] * X + jC[tot]) & kMask48;
                        if (valid && accept) {
                            const int tmp = sperm[t * nr + r];
                            sperm[t * nr + r] = sperm[(t + 1) * nr + r];
                            sperm[(t + 1) * nr + r] = tmp;
                        }
                        na += __popcll(__ballot(valid && accept));
                    }
                    if (lane == 0) sacc[t] = na;
                }
                for (int t = lane; t < NT - 1; t += 64) {
                    long long* ex = a.ex + (size_t)blockIdx.x * 4 * (NT - 1);
                    ex[t] += sacc[t];
                    ex[(NT - 1) + t] += nr;
                    ex[2 * (NT - 1) + t] += sacc[t];
                    ex[3 * (NT - 1) + t] += nr;
                }
                for (int i = lane; i < C; i += 64) sB[sperm[i]] = a.beta_sched[i / nr];
            }
            __syncthreads();
            if (live) beta = sB[tid];
        }
    }

    // ---- write back
    if (live) {
        for (int x = 0; x < N; x++) {
            double* g = a.spins + S.spin_off + ((size_t)tid * N + x) * 2;
            const double2 v = amp[x * C + tid];
            g[0] = v.x;
            g[1] = v.y;
        }
        a.E[ch] = E;
        a.acc[ch] += nacc;
        a.prop[ch] += nprop;
        a.ctr[ch] = sweep;
        a.since[ch] = since;
        a.betas[ch] = beta;
        a.perm[S.chain_off + tid] = sperm[tid];
    }
    if (tid == 0) {
        a.rng48[blockIdx.x] = X;
        a.done[blockIdx.x] = done0 + a.n_sweeps;
    }
}

size_t batch_lds_bytes(const BatchSample& s, int NT) {
    size_t b = (size_t)s.N * s.chains * 16 + (size_t)(s.n_e4 + s.n_e2) * 8 + (size_t)2 * (s.N + 1) * 4;
    b = (b + 7) / 8 * 8 + (size_t)s.chains * 16 + (size_t)s.chains * 4;
    b = (b + 7) / 8 * 8 + 2 * 65 * 8 + (size_t)NT * 8;
    return b;
}

hipError_t batch_prepare(size_t lds) {
    // Dynamic LDS above 64 KB must be allowed explicitly for the kernel.
    return hipFuncSetAttribute((const void*)k_batch, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

hipError_t launch_batch(const BatchArgs& a, int n_samples, int threads, size_t lds, hipStream_t s) {
    hipLaunchKernelGGL(k_batch, dim3(n_samples), dim3(threads), lds, s, a);
    return hipGetLastError();
}

}  // namespace ps
