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
// Draws.  The state-independent part of a step (Philox blocks, pair,
// phases, intensity split, -log u of the threshold) is mostly FP64
// transcendentals.  Each lane of a chain makes one step's draws in L: lane
// `sub` holds, in registers, the draws of step b L + sub of the current
// block b, and the chain's lanes take step t's draws from lane t mod L of
// their group by ds_bpermute.  When a block's last step starts, every lane
// makes its step of the next block (or of the next sweep's first block: the
// draws do not depend on beta), so the draw work is spread over the L lanes
// and interleaved with the chains' LDS-latency-bound term loops, and no draw
// leaves the registers.
//
// Exchange.  After a sweep that ends in an exchange, wave 0 runs it with the
// reference's decision order and drand48 stream for that sample
// (src/parallel_tempering.cu:316-352): the logs of every draw the round can
// take are made in parallel first (the stream's jump-ahead), so the serial
// level chain holds no exponential and no 48-bit product; a decision within
// a narrow band of its draw's log is settled in FP64 on the draw itself.
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

// The kernel's LDS for one sample (offsets from batch_lds_layout).
struct BatchLds {
    double2* amp;  // [N][C | 1]
    uint2* e4;
    uint2* e2;
    int* o4;
    int* o2;
    double* sE;
    double* sES;
    double* sB;
    int* sperm;
    uint64_t* jA;
    uint64_t* jC;
    long long* sacc;
    double* sBs;
    double* sDb;
    unsigned long long* sAB;
    float* sLU;
    __device__ BatchLds(void* lds, const BatchSample& s, int NT) {
        const BatchLdsLayout L = batch_lds_layout(s, NT);
        char* b = (char*)lds;
        amp = (double2*)(b + L.amp);
        e4 = (uint2*)(b + L.e4);
        e2 = (uint2*)(b + L.e2);
        o4 = (int*)(b + L.o4);
        o2 = (int*)(b + L.o2);
        sE = (double*)(b + L.sE);
        sES = (double*)(b + L.sES);
        sB = (double*)(b + L.sB);
        sperm = (int*)(b + L.sperm);
        jA = (uint64_t*)(b + L.jA);
        jC = (uint64_t*)(b + L.jC);
        sacc = (long long*)(b + L.sacc);
        sBs = (double*)(b + L.sBs);
        sDb = (double*)(b + L.sDb);
        sAB = (unsigned long long*)(b + L.sAB);
        sLU = (float*)(b + L.sLU);
    }
};

__device__ __forceinline__ void wave_sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One exchange round of a sample with nr <= 64 replicas per level, on one
// wave (src/parallel_tempering.cu:316-352: levels from the hottest, columns
// in order, a draw only for delta < 0).  Lane r carries the chain that climbs
// column r: at level t it holds the chain that levels < t left at (t, r).
// Slot t+2's chain and energy are read ahead, off the serial path; a level's
// draw is found by its rank among the round's draws.  Returns the drand48
// state after the round.
__device__ __forceinline__ uint64_t batch_exchange_chain(const BatchLds& B, uint64_t X0, int lane, int NT, int nr) {
    const int n = NT * nr, ndraw = nr * (NT - 1);
    const bool valid = lane < nr;
    // energies in slot order, and logf of every draw the round can take
    // (draw k is X_{k+1}: lane l jumps l + 1 steps, then 64 at a time)
    for (int i = lane; i < n; i += 64) B.sES[i] = B.sE[B.sperm[i]];
    {
        uint64_t x = (B.jA[lane + 1] * X0 + B.jC[lane + 1]) & kMask48;
        const uint64_t A64 = B.jA[64], C64 = B.jC[64];
        for (int k = lane; k < ndraw; k += 64) {
            B.sLU[k] = logf((float)lcg48_unit(x));
            x = (A64 * x + C64) & kMask48;
        }
    }
    wave_sync_lds();
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int lr = valid ? lane : nr - 1;  // lanes past the ladders follow the last column, unused
    int pc = B.sperm[lr];                  // chain carried down the column
    double ec = B.sES[lr];                 // its energy
    int pn = B.sperm[nr + lr];             // slot t+1 before level t
    double en = B.sES[nr + lr];
    double dn = B.sDb[0];
    int base = 0;                  // draws taken so far
    unsigned long long abits = 0;  // this lane's accepts, bit t % 64
    for (int t = 0; t < NT - 1; t++) {
        const int pb = pn;
        const double eb = en, db = dn;
        const int t2 = t + 2 < NT ? t + 2 : NT - 1;
        pn = B.sperm[t2 * nr + lr];
        en = B.sES[t2 * nr + lr];
        dn = B.sDb[t + 1 < NT - 1 ? t + 1 : t];
        __builtin_amdgcn_sched_barrier(0);  // the next level's reads go out here
        const double delta = db * (ec - eb);
        const bool need = valid && !(delta >= 0.0);
        const unsigned long long m = __ballot(need);
        // rank of this lane's draw (below ndraw for every lane that needs one;
        // at most ndraw, an entry of the array, for the others)
        const int k = base + __popcll(m & lt);
        // u < exp(delta) from lu = logf((float)u), |lu - log u| < 1e-5: outside a
        // 3e-5 band around delta the order of lu and delta is that of u and
        // exp(delta).  Below delta = -700 (and for u = 0) the FP64 test decides.
        const double dd = delta - (double)B.sLU[k];
        bool accept = valid & (!need | (dd > 3e-5));
        const bool band = need & ((delta < -700.0) | (fabs(dd) <= 3e-5));
        if (__builtin_expect(__ballot(band) != 0, 0)) {
            if (band) {
                uint64_t A, Cc;
                lcg48_coeffs((uint64_t)k + 1, &A, &Cc);
                accept = lcg48_unit((A * X0 + Cc) & kMask48) < exp(delta);
            }
        }
        base += __popcll(m);
        // slot t keeps the carried chain unless the pair swaps; a swap carries it on
        B.sperm[valid ? t * nr + lane : n] = accept ? pb : pc;
        pc = accept ? pc : pb;
        ec = accept ? ec : eb;
        abits |= (unsigned long long)accept << (t & 63);
        if (__builtin_expect((t & 63) == 63 || t == NT - 2, 0)) {
            B.sAB[(t >> 6) * 64 + lane] = abits;
            abits = 0;
        }
    }
    if (valid) B.sperm[(NT - 1) * nr + lane] = pc;
    wave_sync_lds();
    for (int t = lane; t < NT - 1; t += 64) {  // accepted swaps per level
        const unsigned long long* w = B.sAB + (t >> 6) * 64;
        long long cnt = 0;
        for (int r = 0; r < nr; r++) cnt += (long long)((w[r] >> (t & 63)) & 1ull);
        B.sacc[t] += cnt;
    }
    uint64_t A, Cc;
    lcg48_coeffs((uint64_t)base, &A, &Cc);
    return (A * X0 + Cc) & kMask48;
}

// The same round for more than 64 replicas per level: columns in blocks of
// 64 lanes, permutation entries swapped in LDS.
__device__ __forceinline__ uint64_t batch_exchange_wide(const BatchLds& B, uint64_t X, int lane, int NT, int nr) {
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int t = 0; t < NT - 1; t++) {
        long long na = 0;
        for (int c0 = 0; c0 < nr; c0 += 64) {
            const int r = c0 + lane;
            const bool valid = r < nr;
            double delta = 0.0;
            if (valid) delta = B.sDb[t] * (B.sE[B.sperm[t * nr + r]] - B.sE[B.sperm[(t + 1) * nr + r]]);
            const bool need = valid && !(delta >= 0.0);
            const unsigned long long m = __ballot(need);
            bool accept = valid;
            if (need) {
                const int k = __popcll(m & lt) + 1;
                accept = lcg48_unit((B.jA[k] * X + B.jC[k]) & kMask48) < exp(delta);
            }
            const int tot = __popcll(m);
            X = (B.jA[tot] * X + B.jC[tot]) & kMask48;
            if (valid && accept) {
                const int tmp = B.sperm[t * nr + r];
                B.sperm[t * nr + r] = B.sperm[(t + 1) * nr + r];
                B.sperm[(t + 1) * nr + r] = tmp;
            }
            na += __popcll(__ballot(valid && accept));
        }
        if (lane == 0) B.sacc[t] += na;
    }
    wave_sync_lds();
    return X;
}

// Lane addr / 4's value of v (ds_bpermute; the source lane is active).
__device__ __forceinline__ double bperm_f64(int addr, double v) {
    const int lo = __builtin_amdgcn_ds_bpermute(addr, __double2loint(v));
    const int hi = __builtin_amdgcn_ds_bpermute(addr, __double2hiint(v));
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ BatchDraw bperm_draw(int addr, const BatchDraw& o) {
    BatchDraw d;
    d.ij = __builtin_amdgcn_ds_bpermute(addr, o.ij);
    d.pad = 0;
    d.c1 = bperm_f64(addr, o.c1);
    d.s1 = bperm_f64(addr, o.s1);
    d.c2 = bperm_f64(addr, o.c2);
    d.s2 = bperm_f64(addr, o.s2);
    d.q1 = bperm_f64(addr, o.q1);
    d.q2 = bperm_f64(addr, o.q2);
    d.lu = bperm_f64(addr, o.lu);
    return d;
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
// takes about 106 KB of LDS), so a SIMD holds TB / 256 of its waves and the
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
    const BatchLds B(lds, S, NT);
    double2* amp = B.amp;
    uint2* e4 = B.e4;
    uint2* e2 = B.e2;
    int* o4 = B.o4;
    int* o2 = B.o2;
    double* sE = B.sE;
    double* sB = B.sB;
    int* sperm = B.sperm;
    uint64_t* jA = B.jA;
    uint64_t* jC = B.jC;
    long long* sacc = B.sacc;
    double* sBs = B.sBs;
    double* sDb = B.sDb;

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
    const uint64_t cseed = live ? a.seeds[ch] : 0;
    long long nacc = 0, nprop = 0, nex = 0;
    int since = live ? a.since[ch] : 0;
    uint64_t X = (tid < 64) ? a.rng48[blockIdx.x] : 0;  // wave 0 keeps the sample's drand48 state
    const long long done0 = a.done[blockIdx.x];

    const int nsteps = N / 2;
    // This lane's draws: step blk * L + sub of the current block.  The first
    // block of the launch's first sweep is made before the staging barrier.
    BatchDraw own{};
    if (live && a.n_sweeps > 0) own = batch_draw(cseed, sweep, sub, N);
    const int gaddr = (lane & ~(L - 1)) * 4;  // ds_bpermute address of lane 0 of this chain's group
    __syncthreads();

    for (int sw = 0; sw < a.n_sweeps; sw++) {
        if (live) {
            for (int step = 0; step < nsteps; step++) {
                const int q = step & (L - 1);
                const bool blk_end = q == L - 1 || step == nsteps - 1;  // uniform
                const BatchDraw d = bperm_draw(gaddr + q * 4, own);
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
                if (blk_end) {
                    // every lane makes its step of the next block: the next L steps of
                    // this sweep, or the first L of the next sweep (steps past the
                    // sweep's end are made and never read; nothing that is live
                    // across the call is left but the chain's state)
                    const bool next_sw = step + 1 >= nsteps;
                    if (!next_sw || sw + 1 < a.n_sweeps)
                        own = batch_draw(cseed, next_sw ? sweep + 1 : sweep, (next_sw ? 0 : step + 1) + sub, N);
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
        if (!exch) continue;
        if (lead) sE[c] = E;
        __syncthreads();  // every chain has published its energy
        // Exchange (src/parallel_tempering.cu:316-352) for this sample, on wave 0.
        if (tid < 64) {
            if (nr <= 64)
                X = batch_exchange_chain(B, X, lane, NT, nr);
            else
                X = batch_exchange_wide(B, X, lane, NT, nr);
            for (int i = lane; i < C; i += 64) sB[sperm[i]] = sBs[i / nr];
        }
        nex++;
        __syncthreads();  // the new betas are published
        if (live) beta = sB[c];
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

size_t batch_lds_bytes(const BatchSample& s, int NT) { return batch_lds_layout(s, NT).end; }

hipError_t batch_prepare(int chains, size_t lds) {
    // Dynamic LDS above 64 KB must be allowed explicitly for the kernel.
    return hipFuncSetAttribute(batch_kernel(chains), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

hipError_t launch_batch(const BatchArgs& a, int n_samples, int chains, size_t lds, hipStream_t s) {
    void* args[] = {(void*)&a};
    return hipLaunchKernel(batch_kernel(chains), dim3(n_samples), dim3(batch_threads(chains)), args, lds, s);
}

}  // namespace ps
