// batch.hip -- many independent sparse (smoothed-cube) PT systems of uneven N
// in one launch on gfx950.
//
// The reference runs one disorder sample per process
// (src/parallel_tempering_sparse.cu, launched per sample by scan_samples_pt).
// Each sparse step touches only ~4N/N = 4 quartet terms per site plus the
// few FMC-surviving g2 neighbours, so a workgroup per replica (as in
// k_sweep) spends its time in barriers.  Here one workgroup owns one disorder
// sample and a group of 8 (or 4) consecutive lanes owns one chain (replica x
// temperature) of it: a chain steps through its sweep with no barriers,
// reading its FP64 amplitudes and the sample's per-site term lists from LDS.
// The first half of the chain's lanes sums the local field of site i, the
// second half that of site j, each lane every (L/2)-th entry of the lists,
// and a DPP sum combines them.  After every pt_freq sweeps the workgroup
// meets once, wave 0 runs the replica exchange with the reference's decision
// order and drand48 stream for that sample, and every chain picks up its new
// beta.
//
// A sweep runs in two phases.  First the workgroup makes the state-independent
// draws of all (step, chain) pairs of the sample (Philox blocks, pair, phases,
// intensity split, -log u of the threshold) into a global scratch array,
// [step][chain]; these FP64 transcendentals are most of a step's
// instructions, and this way they run with the whole workgroup's waves hiding
// their latency.  Then each chain's lanes walk its steps, prefetching the next
// step's draws, with only the amplitude-dependent part left: three square
// roots of the cube proposal, the local fields and the decision.  The draws
// do not depend on beta, so after a sweep that ends in an exchange, wave 0
// runs the (serial) exchange while the other waves make the next sweep's
// draws.
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

// The draws of one step of one chain (k_sweep's make_step with the cube
// proposal's amplitude-independent factors; src/mc_sparse.cu:118-126,
// include/spins_sparse.h:34-41).
__device__ __noinline__ BatchDraw batch_draw(uint64_t seed, uint64_t sweep, int step, int N) {
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    const U4 x = philox4x32_10(U4{(uint32_t)sweep, (uint32_t)(sweep >> 32), (uint32_t)step, kCtrStep0}, k0, k1);
    const U4 y = philox4x32_10(U4{(uint32_t)sweep, (uint32_t)(sweep >> 32), (uint32_t)step, kCtrStep1}, k0, k1);
    int ri = (int)(u32_to_unit(x.x) * N);
    if (ri >= N) ri = N - 1;
    int rj = (int)(u32_to_unit(x.y) * (N - 1));
    if (rj >= N - 1) rj = N - 2;
    if (rj >= ri) rj++;
    BatchDraw d;
    d.ij = (ri < rj ? ri : rj) | ((ri < rj ? rj : ri) << 16);
    d.pad = 0;
    const double u = u32_to_unit(x.z);
    sincospi(2.0 * u32_to_unit(x.w), &d.s1, &d.c1);  // 2 pi u, u in (0, 1): no large-argument reduction
    sincospi(2.0 * u32_to_unit(y.x), &d.s2, &d.c2);
    d.q1 = sqrt(1.0 - u);
    d.q2 = sqrt(u);
    d.lu = -log(u32_to_unit(y.y));
    return d;
}

// Sum over the L lanes of one chain (L = 4 or 8 consecutive lanes) with DPP
// moves.  Each stage adds a lane's value to its partner's, and a + b == b + a
// exactly, so every lane of the chain ends with the same bits.
template <int L>
__device__ __forceinline__ float chain_sum(float v) {
    static_assert(L == 4 || L == 8, "4 or 8 lanes per chain");
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xb1, 0xf, 0xf, false));  // quad_perm 1,0,3,2
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4e, 0xf, 0xf, false));  // quad_perm 2,3,0,1
    if (L == 8)
        v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x141, 0xf, 0xf, false));  // row_half_mirror
    return v;
}

// LDS layout (dynamic): FP64 amplitudes [N][P] with P = C | 1 (site-major so
// the chains of one site are adjacent; the odd pitch puts the different sites
// that the lanes of one chain read in different banks), H4 entries, H2
// entries, offsets, exchange scratch.
//
// Lanes: L consecutive lanes own one chain c = tid / L.  All of them make the
// chain's proposal and decision (same inputs, same bits); lanes [0, L/2) sum
// the fields of site i over every (L/2)-th entry of its lists, lanes [L/2, L)
// those of site j, and chain_sum adds the L partial Delta E.  Lane sub = 0
// writes the accepted amplitudes and owns the chain's counters and energy.
// The design point is one workgroup per CU (a sample at N = 64 with 96 chains
// takes about 104 KB of LDS), so a SIMD holds TB / 256 of its waves and the
// compiler may use the registers that leaves.
template <int L, int TB>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(1, TB / 256))) void k_batch(BatchArgs a) {
    extern __shared__ double lds[];
    const BatchSample S = a.samples[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63;
    const int N = S.N, C = S.chains, NT = a.NT, nr = C / NT, P = C | 1;
    constexpr int H = L / 2;
    const int c = tid / L, sub = tid % L, half = sub / H, hs = sub % H;
    const bool live = c < C, lead = live && sub == 0;
    double2* amp = (double2*)lds;                       // [N][P]
    uint2* e4 = (uint2*)(amp + (size_t)N * P);          // H4 entries
    uint2* e2 = e4 + S.n_e4;                            // H2 entries {k, g}
    int* o4 = (int*)(e2 + S.n_e2);                      // [N+1]
    int* o2 = o4 + N + 1;                               // [N+1]
    double* sE = (double*)(((uintptr_t)(o2 + N + 1) + 7) & ~(uintptr_t)7);  // [C]
    double* sB = sE + C;                                // [C] beta by chain
    int* sperm = (int*)(sB + C);                        // [C]
    uint64_t* jA = (uint64_t*)(((uintptr_t)(sperm + C) + 7) & ~(uintptr_t)7);  // [65]
    uint64_t* jC = jA + 65;
    long long* sacc = (long long*)(jC + 65);            // [NT] accepted swaps per level over the launch
    double* sBs = (double*)(sacc + NT);                 // [NT] beta schedule
    double* sDb = sBs + NT;                             // [NT] beta_t - beta_{t+1}

    // ---- stage the sample's terms and the chains' amplitudes
    for (int i = tid; i < S.n_e4; i += blockDim.x) e4[i] = a.e4[S.e4_off + i];
    for (int i = tid; i < S.n_e2; i += blockDim.x) e2[i] = a.e2[S.e2_off + i];
    for (int i = tid; i <= N; i += blockDim.x) {
        o4[i] = a.o4[S.o_off + i];
        o2[i] = a.o2[S.o_off + i];
    }
    for (int x = tid; x < N * C; x += blockDim.x) {
        const int cc = x / N, site = x - cc * N;
        const double* g = a.spins + S.spin_off + (size_t)x * 2;
        amp[site * P + cc] = make_double2(g[0], g[1]);
    }
    if (tid < C) sperm[tid] = a.perm[S.chain_off + tid];
    if (tid < 64) lcg48_coeffs((uint64_t)tid + 1, &jA[tid + 1], &jC[tid + 1]);
    if (tid == 0) {
        jA[0] = 1;
        jC[0] = 0;
    }
    if (tid < NT) {
        sBs[tid] = a.beta_sched[tid];
        sacc[tid] = 0;
        if (tid < NT - 1) sDb[tid] = a.beta_sched[tid] - a.beta_sched[tid + 1];
    }
    const int ch = S.chain_off + c;
    double beta = live ? a.betas[ch] : 0.0;
    double E = live ? a.E[ch] : 0.0;
    uint64_t sweep = live ? a.ctr[ch] : 0;
    long long nacc = 0, nprop = 0, nex = 0;
    int since = live ? a.since[ch] : 0;
    uint64_t X = (tid < 64) ? a.rng48[blockIdx.x] : 0;  // wave 0 keeps the sample's drand48 state
    const long long done0 = a.done[blockIdx.x];
    __syncthreads();

    const int nsteps = N / 2;
    BatchDraw* dr = a.draws + (size_t)S.chain_off * a.max_steps;  // this sample's [step][chain] draws
    // Phase 1 of sweep sw: every (step, chain) draw, by threads [first, blockDim.x).
    auto make_draws = [&](int sw, int first) {
        for (int x = tid - first; x < nsteps * C; x += (int)blockDim.x - first) {
            const int st = x / C, cx = x - st * C;
            const int gc = S.chain_off + cx;
            dr[x] = batch_draw(a.seeds[gc], a.ctr[gc] + (uint64_t)sw, st, N);
        }
    };
    if (a.n_sweeps > 0) make_draws(0, 0);
    __syncthreads();
    for (int sw = 0; sw < a.n_sweeps; sw++) {
        // Phase 2: each chain's steps, the next step's draws loaded one ahead.
        if (live) {
            BatchDraw nd = dr[c];
            for (int step = 0; step < nsteps; step++) {
                const BatchDraw d = nd;
                if (step + 1 < nsteps) nd = dr[(size_t)(step + 1) * C + c];
                const int i0 = d.ij & 0xffff, j0 = d.ij >> 16;
                // this lane's site (i for the first half of the chain's lanes, j for the second);
                // its list bounds and first entries are read before the proposal's arithmetic
                const int site = half ? j0 : i0, other = half ? i0 : j0;
                const int b2 = o2[site] + hs, n2 = o2[site + 1], b4 = o4[site] + hs, n4 = o4[site + 1];
                const double2 ai = amp[i0 * P + c], aj = amp[j0 * P + c];
                // an index inside the sample's entry arrays (the value is unused when the list is empty)
                uint2 en2 = e2[b2 < n2 ? b2 : (n2 > 0 ? n2 - 1 : 0)];
                uint2 en4 = e4[b4 < n4 ? b4 : (n4 > 0 ? n4 - 1 : 0)];
                // cube proposal (include/spins_sparse.h:26-56)
                const double r1 = ai.x * ai.x + ai.y * ai.y, r2 = aj.x * aj.x + aj.y * aj.y;
                const double sq = sqrt(r1 * r1 + r2 * r2);
                const double m1 = sqrt(sq * d.q1), m2 = sqrt(sq * d.q2);
                const double2 ni = make_double2(m1 * d.c1, m1 * d.s1), nj = make_double2(m2 * d.c2, m2 * d.s2);
                const double2 as = half ? aj : ai, ns = half ? nj : ni;
                const double2 ao = half ? ai : aj, no = half ? ni : nj;
                const double thr = (beta == 0.0) ? INFINITY : d.lu / beta;
                const float2 dd = make_float2((float)(ns.x - as.x), (float)(ns.y - as.y));
                const float2 mo = make_float2((float)(0.5 * (ao.x + no.x)), (float)(0.5 * (ao.y + no.y)));
                // local field at the midpoint configuration, this lane's share of the entries
                // (every H-th entry in list order; each iteration reads the next entry ahead)
                float2 F = make_float2(0.f, 0.f);
                for (int e = b2; e < n2; e += H) {
                    const uint2 en = en2;
                    const int k = (int)en.x;
                    const float g = __uint_as_float(en.y);
                    const double2 Ak = amp[k * P + c];  // unconditional: no branch around the LDS read
                    en2 = e2[e + H < n2 ? e + H : e];
                    __builtin_amdgcn_sched_barrier(0);  // both reads in flight together
                    const float2 tk = f2(Ak);
                    const float2 t = (k == other) ? mo : tk;
                    F.x += g * t.x;
                    F.y -= g * t.y;
                }
                for (int e = b4; e < n4; e += H) {
                    const uint2 en = en4;
                    en4 = e4[e + H < n4 ? e + H : e];
                    const int p = en.x & 0xff, m = (en.x >> 8) & 0xff, n = (en.x >> 16) & 0xff;
                    const float g = __uint_as_float(en.y);
                    // the three reads go out together; the midpoint replaces the other site's value
                    const double2 Ap = amp[p * P + c], Am = amp[m * P + c], An = amp[n * P + c];
                    __builtin_amdgcn_sched_barrier(0);  // keep the three reads in flight together
                    const float2 ap = f2(Ap), am = f2(Am), an = f2(An);
                    const float2 tp = (p == other) ? mo : ap;
                    const float2 tm = (m == other) ? mo : am;
                    const float2 tn = (n == other) ? mo : an;
                    const float2 mn = make_float2(tm.x * tn.x - tm.y * tn.y, tm.x * tn.y + tm.y * tn.x);
                    F.x += g * (tp.x * mn.x + tp.y * mn.y);
                    F.y += g * (tp.y * mn.x - tp.x * mn.y);
                }
                const float dE = chain_sum<L>(-(dd.x * F.x - dd.y * F.y));
                nprop++;
                if (beta == 0.0 || (double)dE < thr) {
                    if (sub == 0) {
                        amp[i0 * P + c] = ni;
                        amp[j0 * P + c] = nj;
                    }
                    // the chain's other lanes read these next step from the same
                    // wave: keep every later LDS access after the stores
                    __builtin_amdgcn_wave_barrier();
                    E += (double)dE;
                    nacc++;
                }
            }
            sweep++;
            since++;
            if (a.resync_every > 0 && since >= a.resync_every) {
                // FP64 resync from the sample's FP64 terms (src/mc_sparse.cu:258-315, i<j H2)
                if (sub == 0) {
                    double e = 0.0;
                    for (int t = 0; t < S.n_q; t++) {
                        const int* v = a.quartets + (size_t)5 * (S.q_off + t);
                        const double2 A0 = amp[v[0] * P + c], A1 = amp[v[1] * P + c];
                        const double2 A2 = amp[v[2] * P + c], A3 = amp[v[3] * P + c];
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
                        const double2 A = amp[ij.x * P + c], B = amp[ij.y * P + c];
                        e += -a.h2g[S.h2_off + t] * (A.x * B.x + A.y * B.y);
                    }
                    E = e;
                }
                since = 0;
            }
        }
        const bool exch = a.pt_freq > 0 && (done0 + sw + 1) % a.pt_freq == 0;
        if (exch && lead) sE[c] = E;
        __syncthreads();  // the chains are done with this sweep's draws and have published their energies
        // Wave 0 runs the exchange while the other waves make the next sweep's draws.
        if (sw + 1 < a.n_sweeps && (!exch || tid >= 64)) make_draws(sw + 1, exch ? 64 : 0);
        if (exch) {
            // Exchange (src/parallel_tempering.cu:316-352) for this sample.
            if (tid < 64) {
                const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
                if (nr <= 64) {
                    // Lane r carries the chain that climbs column r: at level t it
                    // holds the chain that levels < t left at (t, r).  The chain at
                    // (t + 1, r) is not written before level t, so it and its
                    // energy are read a level ahead, off the serial path.
                    const bool valid = lane < nr;
                    int cur = valid ? sperm[lane] : 0;
                    double Ecur = valid ? sE[cur] : 0.0;
                    int nx = valid ? sperm[nr + lane] : 0;
                    double Enx = valid ? sE[nx] : 0.0;
                    for (int t = 0; t < NT - 1; t++) {
                        const int n1 = nx;
                        const double E1 = Enx;
                        if (valid && t + 2 < NT) {
                            nx = sperm[(t + 2) * nr + lane];
                            Enx = sE[nx];
                        }
                        const double delta = sDb[t] * (Ecur - E1);
                        const bool need = valid && !(delta >= 0.0);
                        const unsigned long long m = __ballot(need);
                        bool accept = valid;
                        if (need) {
                            const int k = __popcll(m & lt) + 1;
                            accept = lcg48_unit((jA[k] * X + jC[k]) & kMask48) < exp(delta);
                        }
                        const int tot = __popcll(m);
                        X = (jA[tot] * X + jC[tot]) & kMask48;
                        if (valid) {
                            sperm[t * nr + lane] = accept ? n1 : cur;
                            if (!accept) {
                                cur = n1;
                                Ecur = E1;
                            }
                        }
                        const unsigned long long am = __ballot(valid && accept);
                        if (lane == 0) sacc[t] += __popcll(am);
                    }
                    if (valid) sperm[(NT - 1) * nr + lane] = cur;
                } else {
                    for (int t = 0; t < NT - 1; t++) {
                        long long na = 0;
                        for (int c0 = 0; c0 < nr; c0 += 64) {
                            const int r = c0 + lane;
                            const bool valid = r < nr;
                            double delta = 0.0;
                            if (valid) delta = sDb[t] * (sE[sperm[t * nr + r]] - sE[sperm[(t + 1) * nr + r]]);
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
                        if (lane == 0) sacc[t] += na;
                    }
                }
                for (int i = lane; i < C; i += 64) sB[sperm[i]] = sBs[i / nr];
            }
            nex++;
        }
        __syncthreads();  // the next sweep's draws are made and the new betas published
        if (exch && live) beta = sB[c];
    }

    // ---- exchange counters of the launch, once
    if (tid < 64 && nex > 0) {
        long long* ex = a.ex + (size_t)blockIdx.x * 4 * (NT - 1);
        for (int t = lane; t < NT - 1; t += 64) {
            ex[t] += sacc[t];
            ex[(NT - 1) + t] += nr * nex;
            ex[2 * (NT - 1) + t] += sacc[t];
            ex[3 * (NT - 1) + t] += nr * nex;
        }
    }

    // ---- write back (amplitudes were last written by other waves' lead lanes)
    __syncthreads();
    for (int x = tid; x < N * C; x += blockDim.x) {
        const int cc = x / N, site = x - cc * N;
        double* g = a.spins + S.spin_off + (size_t)x * 2;
        const double2 v = amp[site * P + cc];
        g[0] = v.x;
        g[1] = v.y;
    }
    if (lead) {
        a.E[ch] = E;
        a.acc[ch] += nacc;
        a.prop[ch] += nprop;
        a.ctr[ch] = sweep;
        a.since[ch] = since;
        a.betas[ch] = beta;
    }
    if (tid < C) a.perm[S.chain_off + tid] = sperm[tid];
    if (tid == 0) {
        a.rng48[blockIdx.x] = X;
        a.done[blockIdx.x] = done0 + a.n_sweeps;
    }
}

// Lanes per chain and workgroup size: 8 lanes in the smallest of 512, 768 or
// 1024 threads that holds them (fewer threads leave each more registers), 4
// lanes in 1024 threads above 128 chains.
int batch_lanes(int chains) { return chains <= kBatchThreads / 8 ? 8 : 4; }

int batch_threads(int chains) {
    const int need = chains * batch_lanes(chains);
    return need <= 512 ? 512 : (need <= 768 ? 768 : 1024);
}

namespace {
const void* batch_kernel(int chains) {
    if (batch_lanes(chains) == 4) return (const void*)k_batch<4, 1024>;
    switch (batch_threads(chains)) {
        case 512: return (const void*)k_batch<8, 512>;
        case 768: return (const void*)k_batch<8, 768>;
        default: return (const void*)k_batch<8, 1024>;
    }
}
}  // namespace

size_t batch_lds_bytes(const BatchSample& s, int NT) {
    size_t b = (size_t)s.N * (s.chains | 1) * 16 + (size_t)(s.n_e4 + s.n_e2) * 8 + (size_t)2 * (s.N + 1) * 4;
    b = (b + 7) / 8 * 8 + (size_t)s.chains * 16 + (size_t)s.chains * 4;
    b = (b + 7) / 8 * 8 + 2 * 65 * 8 + (size_t)NT * 8 * 3;
    return b;
}

hipError_t batch_prepare(int chains, size_t lds) {
    // Dynamic LDS above 64 KB must be allowed explicitly for the kernel.
    return hipFuncSetAttribute(batch_kernel(chains), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

hipError_t launch_batch(const BatchArgs& a, int n_samples, int chains, size_t lds, hipStream_t s) {
    void* args[] = {(void*)&a};
    return hipLaunchKernel(batch_kernel(chains), dim3(n_samples), dim3(batch_threads(chains)), args, lds, s);
}

}  // namespace ps
