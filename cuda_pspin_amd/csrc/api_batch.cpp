// api_batch.cpp -- C ABI for batches of sparse PT systems (pspin_batch_*),
// run by the one-launch kernel in batch.hip.  Each sample's disorder, spins,
// seeds and initial energies come from the same set-up as a single
// PSPIN_SPARSE state with that sample's seed (mc_sparse_init,
// src/mc_sparse.cu:390-604), so a batch sample and the single state start
// from identical data.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "pspin_batch.h"

#define PB_TRY_HIP(call)                                                                              \
    do {                                                                                              \
        hipError_t _e = (call);                                                                       \
        if (_e != hipSuccess) {                                                                       \
            ps::set_error(std::string("HIP error ") + hipGetErrorString(_e) + " (" #call ")");         \
            return PSPIN_ERR_HIP;                                                                     \
        }                                                                                             \
    } while (0)
#define PB_REQUIRE(cond, code, msg) \
    do {                            \
        if (!(cond)) {              \
            ps::set_error(msg);     \
            return code;            \
        }                           \
    } while (0)

struct pspin_batch {
    int n = 0, NT = 0, nr = 0;
    long long chains = 0;
    hipStream_t stream = nullptr;
    std::vector<ps::BatchSample> h_samples;  // launch order (largest N first)
    std::vector<int> slot_of;                // sample index -> launch slot
    ps::BatchSample* d_samples = nullptr;
    uint2 *d_e4 = nullptr, *d_e2 = nullptr;
    int *d_o4 = nullptr, *d_o2 = nullptr, *d_q = nullptr;
    double *d_qg = nullptr, *d_h2g = nullptr;
    int2* d_h2ij = nullptr;
    double *d_spins = nullptr, *d_E = nullptr, *d_betas = nullptr, *d_sched = nullptr;
    long long *d_acc = nullptr, *d_prop = nullptr, *d_ex = nullptr, *d_done = nullptr;
    uint64_t *d_seeds = nullptr, *d_ctr = nullptr, *d_rng48 = nullptr;
    int *d_since = nullptr, *d_perm = nullptr;
    int resync_every = 1024;
    int chains_per_sample = 1;  // NT * nrep_per_temp (picks the lanes per chain)
    size_t lds = 0;
    long long pair_updates_per_sweep = 0;
    hipGraphExec_t graph = nullptr;
    int graph_sweeps = 0, graph_freq = 0;
    int chunk = 64;  // sweeps per launch (autotuned; results do not depend on it)
    size_t n_spin = 0, n_perm = 0, n_ex = 0;
};

namespace {

template <class T>
hipError_t dalloc(T** p, size_t n, const void* src, hipStream_t s) {
    *p = nullptr;
    hipError_t e = hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T));
    if (e != hipSuccess || !src || n == 0) return e;
    e = hipMemcpyAsync(*p, src, n * sizeof(T), hipMemcpyHostToDevice, s);
    return e != hipSuccess ? e : hipStreamSynchronize(s);
}

template <class T>
void dfree(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

void batch_free(pspin_batch* b) {
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    dfree(b->d_samples); dfree(b->d_e4); dfree(b->d_e2); dfree(b->d_o4); dfree(b->d_o2); dfree(b->d_q);
    dfree(b->d_qg); dfree(b->d_h2g); dfree(b->d_h2ij); dfree(b->d_spins); dfree(b->d_E); dfree(b->d_betas);
    dfree(b->d_sched); dfree(b->d_acc); dfree(b->d_prop); dfree(b->d_ex); dfree(b->d_done); dfree(b->d_seeds);
    dfree(b->d_ctr); dfree(b->d_rng48); dfree(b->d_since); dfree(b->d_perm);
    if (b->stream) (void)hipStreamDestroy(b->stream);
}

// Same packing as the single-state lists: the term -g Re[a_x a_p conj(a_m) conj(a_n)].
uint32_t pack_term(const int* v, int ch, int s) {
    const int cmask = ps::conj_mask(ch), bit = (cmask >> s) & 1;
    int p = -1, mn[2], c = 0;
    for (int t = 0; t < 4; t++) {
        if (t == s) continue;
        if (((cmask >> t) & 1) == bit && p < 0) p = t;
        else mn[c++] = t;
    }
    return (uint32_t)v[p] | ((uint32_t)v[mn[0]] << 8) | ((uint32_t)v[mn[1]] << 16);
}

int batch_args(pspin_batch* b, ps::BatchArgs& a, int n_sweeps, int pt_freq) {
    a.samples = b->d_samples;
    a.n_sweeps = n_sweeps;
    a.pt_freq = pt_freq;
    a.NT = b->NT;
    a.resync_every = b->resync_every;
    a.e4 = b->d_e4; a.e2 = b->d_e2; a.o4 = b->d_o4; a.o2 = b->d_o2;
    a.quartets = b->d_q; a.qg = b->d_qg; a.h2ij = b->d_h2ij; a.h2g = b->d_h2g;
    a.spins = b->d_spins; a.E = b->d_E; a.acc = b->d_acc; a.prop = b->d_prop; a.seeds = b->d_seeds;
    a.ctr = b->d_ctr; a.since = b->d_since; a.betas = b->d_betas; a.perm = b->d_perm;
    a.beta_sched = b->d_sched; a.ex = b->d_ex; a.rng48 = b->d_rng48; a.done = b->d_done;
    return PSPIN_OK;
}

}  // namespace

extern "C" {

int pspin_batch_create(const pspin_sim_config* base, int n_samples, const int* Ns, const uint64_t* seeds, int NT,
                       int nrep_per_temp, const double* beta_sched, pspin_batch** out) {
    PB_REQUIRE(base && Ns && seeds && beta_sched && out, PSPIN_ERR_ARG, "null argument");
    PB_REQUIRE(n_samples >= 1 && NT >= 2 && nrep_per_temp >= 1, PSPIN_ERR_ARG, "need n_samples >= 1, NT >= 2");
    const int C = NT * nrep_per_temp;
    PB_REQUIRE(C <= 256, PSPIN_ERR_ARG, "NT * nrep_per_temp must be <= 256 (4 lanes per chain in a 1024-thread workgroup)");
    *out = nullptr;
    pspin_batch* b = new pspin_batch();
    b->n = n_samples;
    b->NT = NT;
    b->nr = nrep_per_temp;
    auto fail = [&](int rc) {
        batch_free(b);
        delete b;
        return rc;
    };
    if (base->dev >= 0 && hipSetDevice(base->dev) != hipSuccess) return fail(PSPIN_ERR_HIP);
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return fail(PSPIN_ERR_HIP);

    // Launch order: largest N first, so the longest samples start first.
    std::vector<int> order(n_samples);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return Ns[x] > Ns[y]; });
    b->slot_of.assign(n_samples, 0);

    std::vector<uint2> e4, e2;
    std::vector<int> o4, o2, q;
    std::vector<double> qg, h2g, spins, E, betas;
    std::vector<int2> h2ij;
    std::vector<uint64_t> cseeds, rng48;
    std::vector<int> perm;
    size_t lds_max = 0;
    for (int slot = 0; slot < n_samples; slot++) {
        const int s = order[slot];
        b->slot_of[s] = slot;
        pspin_sim_config c = *base;
        c.N = Ns[s];
        c.seed = seeds[s];
        c.nrep = C;
        c.dev = -1;
        pspin_state* st = nullptr;
        int rc = pspin_create(&c, PSPIN_SPARSE, &st);
        if (rc) return fail(rc);
        const int N = c.N;
        ps::BatchSample S{};
        S.N = N;
        S.chains = C;
        S.chain_off = (int)b->chains;
        S.spin_off = (long long)spins.size();
        // spins, energies, seeds
        std::vector<double> sp((size_t)2 * N * C), en(C);
        rc = pspin_get_spins(st, sp.data());
        if (!rc) rc = pspin_get_results(st, en.data(), nullptr, nullptr);
        std::vector<double> g2((size_t)N * N);
        if (!rc) rc = pspin_get_g2(st, g2.data());
        pspin_info inf{};
        if (!rc) rc = pspin_get_info(st, &inf);
        std::vector<int> sq((size_t)5 * std::max(inf.n_quartets, 1));
        std::vector<double> sg(std::max(inf.n_quartets, 1));
        if (!rc) rc = pspin_get_sparse_quartets(st, sq.data(), sg.data());
        std::vector<uint64_t> rs(C);
        if (!rc) rc = pspin_get_rng_info(st, rs.data(), nullptr);
        pspin_destroy(st);
        if (rc) return fail(rc);
        spins.insert(spins.end(), sp.begin(), sp.end());
        E.insert(E.end(), en.begin(), en.end());
        cseeds.insert(cseeds.end(), rs.begin(), rs.end());
        for (int i = 0; i < C; i++) {
            betas.push_back(beta_sched[i / nrep_per_temp]);
            perm.push_back(i);
        }
        rng48.push_back(ps::srand48_state((long)seeds[s] + 5000));
        // H4 per-site lists and FP64 quartets
        S.q_off = (int)qg.size();
        S.n_q = inf.n_quartets;
        std::vector<std::vector<uint2>> per(N);
        for (int t = 0; t < inf.n_quartets; t++) {
            const int* v = &sq[5 * t];
            float gf = (float)sg[t];
            uint32_t gb;
            std::memcpy(&gb, &gf, 4);
            for (int p = 0; p < 4; p++) per[v[p]].push_back(make_uint2(pack_term(v, v[4], p), gb));
            q.insert(q.end(), v, v + 5);
            qg.push_back(sg[t]);
        }
        S.o_off = (int)o4.size();
        S.e4_off = (int)e4.size();
        S.e2_off = (int)e2.size();
        int acc4 = 0, acc2 = 0;
        S.h2_off = (int)h2g.size();
        for (int x = 0; x < N; x++) {
            o4.push_back(acc4);
            o2.push_back(acc2);
            for (const uint2& en2 : per[x]) e4.push_back(en2);
            acc4 += (int)per[x].size();
            if (inf.h2_active)
                for (int k = 0; k < N; k++) {
                    const double g = g2[(size_t)x * N + k];
                    if (k == x || g == 0.0) continue;
                    float gf = (float)g;
                    uint32_t gb;
                    std::memcpy(&gb, &gf, 4);
                    e2.push_back(make_uint2((uint32_t)k, gb));
                    acc2++;
                    if (k > x) {
                        h2ij.push_back(make_int2(x, k));
                        h2g.push_back(g);
                    }
                }
        }
        o4.push_back(acc4);
        o2.push_back(acc2);
        S.n_e4 = acc4;
        S.n_e2 = acc2;
        S.n_h2 = (int)h2g.size() - S.h2_off;
        b->chains += C;
        b->pair_updates_per_sweep += (long long)C * (N / 2);
        lds_max = std::max(lds_max, ps::batch_lds_bytes(S, NT));
        b->h_samples.push_back(S);
    }
    PB_REQUIRE(lds_max <= 160 * 1024, PSPIN_ERR_ARG, "a sample needs more than 160 KB of LDS (N * chains too large)");
    b->lds = lds_max;
    // each chain runs on 4 or 8 lanes
    b->chains_per_sample = C;
    hipStream_t s = b->stream;
    const size_t nch = (size_t)b->chains;
    std::vector<uint64_t> zero64(nch, 0);
    std::vector<int> zero32(nch, 0);
    std::vector<long long> zll(std::max<size_t>(nch, (size_t)n_samples * 4 * (NT - 1)), 0);
#define PB_ALLOC(p, n, src) \
    if (dalloc(&(p), (n), (src), s) != hipSuccess) return fail(PSPIN_ERR_HIP)
    PB_ALLOC(b->d_samples, b->h_samples.size(), b->h_samples.data());
    PB_ALLOC(b->d_e4, e4.size(), e4.data());
    PB_ALLOC(b->d_e2, e2.size(), e2.data());
    PB_ALLOC(b->d_o4, o4.size(), o4.data());
    PB_ALLOC(b->d_o2, o2.size(), o2.data());
    PB_ALLOC(b->d_q, q.size(), q.data());
    PB_ALLOC(b->d_qg, qg.size(), qg.data());
    PB_ALLOC(b->d_h2ij, h2ij.size(), h2ij.data());
    PB_ALLOC(b->d_h2g, h2g.size(), h2g.data());
    PB_ALLOC(b->d_spins, spins.size(), spins.data());
    PB_ALLOC(b->d_E, E.size(), E.data());
    PB_ALLOC(b->d_betas, betas.size(), betas.data());
    PB_ALLOC(b->d_sched, (size_t)NT, beta_sched);
    PB_ALLOC(b->d_acc, nch, zll.data());
    PB_ALLOC(b->d_prop, nch, zll.data());
    PB_ALLOC(b->d_ex, (size_t)n_samples * 4 * (NT - 1), zll.data());
    PB_ALLOC(b->d_done, (size_t)n_samples, zll.data());
    PB_ALLOC(b->d_seeds, cseeds.size(), cseeds.data());
    PB_ALLOC(b->d_ctr, nch, zero64.data());
    PB_ALLOC(b->d_rng48, rng48.size(), rng48.data());
    PB_ALLOC(b->d_since, nch, zero32.data());
    PB_ALLOC(b->d_perm, perm.size(), perm.data());
#undef PB_ALLOC
    b->n_spin = spins.size();
    b->n_perm = perm.size();
    b->n_ex = (size_t)n_samples * 4 * (NT - 1);
    if (ps::batch_prepare(b->chains_per_sample, b->lds) != hipSuccess) {
        ps::set_error("the batch kernel cannot get " + std::to_string(b->lds) + " bytes of LDS");
        return fail(PSPIN_ERR_HIP);
    }
    *out = b;
    return PSPIN_OK;
}

int pspin_batch_destroy(pspin_batch* b) {
    if (!b) return PSPIN_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    batch_free(b);
    delete b;
    return PSPIN_OK;
}

int pspin_batch_set_resync_interval(pspin_batch* b, int sweeps) {
    PB_REQUIRE(b && sweeps >= 0, PSPIN_ERR_ARG, "bad argument");
    b->resync_every = sweeps;
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
    return PSPIN_OK;
}

}  // extern "C"

namespace {

// n_sweeps sweeps of every chain; an exchange for every sample after each
// sweep that completes a multiple of pt_freq (0 = no exchange).  Launches of
// b->chunk sweeps are replayed from a captured graph of ~512 sweeps.  All
// chain state (RNG and resync counters, permutation, drand48 state) lives in
// device memory, so the result does not depend on the launch split.
int batch_advance(pspin_batch* b, int n_sweeps, int pt_freq) {
    const int chunk = b->chunk, launches = std::max(1, 512 / chunk);
    ps::BatchArgs a;
    int done = 0;
    while (done < n_sweeps) {
        const int left = n_sweeps - done;
        if (left >= chunk * launches) {
            if (!b->graph || b->graph_freq != pt_freq || b->graph_sweeps != chunk) {
                if (b->graph) (void)hipGraphExecDestroy(b->graph);
                b->graph = nullptr;
                batch_args(b, a, chunk, pt_freq);
                hipGraph_t g = nullptr;
                PB_TRY_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
                hipError_t e = hipSuccess;
                for (int k = 0; k < launches && e == hipSuccess; k++)
                    e = ps::launch_batch(a, b->n, b->chains_per_sample, b->lds, b->stream);
                hipError_t e2 = hipStreamEndCapture(b->stream, &g);
                PB_TRY_HIP(e);
                PB_TRY_HIP(e2);
                hipError_t e3 = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
                PB_TRY_HIP(e3);
                b->graph_freq = pt_freq;
                b->graph_sweeps = chunk;
            }
            PB_TRY_HIP(hipGraphLaunch(b->graph, b->stream));
            done += chunk * launches;
            continue;
        }
        const int c = std::min(left, chunk);
        batch_args(b, a, c, pt_freq);
        PB_TRY_HIP(ps::launch_batch(a, b->n, b->chains_per_sample, b->lds, b->stream));
        done += c;
    }
    return PSPIN_OK;
}

}  // namespace

extern "C" {

int pspin_batch_run(pspin_batch* b, int n_sweeps, int pt_freq) {
    PB_REQUIRE(b && n_sweeps >= 0 && pt_freq >= 0, PSPIN_ERR_ARG, "bad argument");
    const int rc = batch_advance(b, n_sweeps, pt_freq);
    if (rc) return rc;
    PB_TRY_HIP(hipStreamSynchronize(b->stream));
    return PSPIN_OK;
}

int pspin_batch_set_chunk(pspin_batch* b, int sweeps_per_launch) {
    PB_REQUIRE(b && sweeps_per_launch >= 1 && sweeps_per_launch <= 4096, PSPIN_ERR_ARG, "bad argument");
    b->chunk = sweeps_per_launch;
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
    return PSPIN_OK;
}

// Times sweeps-per-launch candidates on the live batch and keeps the fastest;
// every mutable device array is saved first and restored after each trial.
int pspin_batch_autotune(pspin_batch* b, int trial_sweeps, int pt_freq, int verbose, int* chunk_out,
                         double* ms_per_sweep) {
    PB_REQUIRE(b && trial_sweeps >= 1 && pt_freq >= 0, PSPIN_ERR_ARG, "bad argument");
    const long long nch = b->chains;
    struct Buf { void* dev; size_t bytes; void* saved; };
    std::vector<Buf> bufs = {{b->d_spins, b->n_spin * 8, nullptr}, {b->d_E, (size_t)nch * 8, nullptr},
                             {b->d_betas, (size_t)nch * 8, nullptr}, {b->d_acc, (size_t)nch * 8, nullptr},
                             {b->d_prop, (size_t)nch * 8, nullptr}, {b->d_ex, b->n_ex * 8, nullptr},
                             {b->d_done, (size_t)b->n * 8, nullptr}, {b->d_ctr, (size_t)nch * 8, nullptr},
                             {b->d_rng48, (size_t)b->n * 8, nullptr}, {b->d_since, (size_t)nch * 4, nullptr},
                             {b->d_perm, b->n_perm * 4, nullptr}};
    int rc = PSPIN_OK;
    auto fail = [&](hipError_t e) {
        ps::set_error(std::string("HIP error in pspin_batch_autotune: ") + hipGetErrorString(e));
        rc = PSPIN_ERR_HIP;
    };
    for (Buf& x : bufs) {
        hipError_t e = hipMalloc(&x.saved, std::max<size_t>(x.bytes, 1));
        if (!e) e = hipMemcpyAsync(x.saved, x.dev, x.bytes, hipMemcpyDeviceToDevice, b->stream);
        if (e) { fail(e); break; }
    }
    const int old = b->chunk;
    int best_chunk = old;
    double best = 1e300;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    for (int c : {16, 32, 64, 128, 256}) {
        if (rc) break;
        b->chunk = c;
        if (b->graph) (void)hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
        hipError_t e = hipEventRecord(e0, b->stream);
        if (!e && (rc = batch_advance(b, trial_sweeps, pt_freq))) break;
        if (!e) e = hipEventRecord(e1, b->stream);
        if (!e) e = hipEventSynchronize(e1);
        float ms = 0.0f;
        if (!e) e = hipEventElapsedTime(&ms, e0, e1);
        for (Buf& x : bufs)
            if (!e) e = hipMemcpyAsync(x.dev, x.saved, x.bytes, hipMemcpyDeviceToDevice, b->stream);
        if (e) { fail(e); break; }
        const double per = ms / trial_sweeps;
        if (verbose) std::printf("[auto-tune] batch %d samples: %d sweeps/launch: %.4f ms/sweep\n", b->n, c, per);
        if (per < best) {
            best = per;
            best_chunk = c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(b->stream);
    for (Buf& x : bufs)
        if (x.saved) (void)hipFree(x.saved);
    b->chunk = rc ? old : best_chunk;
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
    if (rc) return rc;
    if (verbose) std::printf("[auto-tune] batch -> %d sweeps/launch (%.4f ms/sweep)\n", best_chunk, best);
    if (chunk_out) *chunk_out = best_chunk;
    if (ms_per_sweep) *ms_per_sweep = best;
    return PSPIN_OK;
}

int pspin_batch_info(const pspin_batch* b, int* n_samples, long long* chains, long long* pair_updates_per_sweep) {
    PB_REQUIRE(b, PSPIN_ERR_ARG, "null argument");
    if (n_samples) *n_samples = b->n;
    if (chains) *chains = b->chains;
    if (pair_updates_per_sweep) *pair_updates_per_sweep = b->pair_updates_per_sweep;
    return PSPIN_OK;
}

int pspin_batch_get_results(const pspin_batch* b, int sample, double* E, long long* acc, long long* prop) {
    PB_REQUIRE(b && sample >= 0 && sample < b->n, PSPIN_ERR_ARG, "bad sample");
    const ps::BatchSample& S = b->h_samples[b->slot_of[sample]];
    if (E) PB_TRY_HIP(hipMemcpyAsync(E, b->d_E + S.chain_off, S.chains * sizeof(double), hipMemcpyDeviceToHost, b->stream));
    if (acc)
        PB_TRY_HIP(hipMemcpyAsync(acc, b->d_acc + S.chain_off, S.chains * sizeof(long long), hipMemcpyDeviceToHost, b->stream));
    if (prop)
        PB_TRY_HIP(hipMemcpyAsync(prop, b->d_prop + S.chain_off, S.chains * sizeof(long long), hipMemcpyDeviceToHost, b->stream));
    PB_TRY_HIP(hipStreamSynchronize(b->stream));
    return PSPIN_OK;
}

int pspin_batch_get_spins(const pspin_batch* b, int sample, double* re_im) {
    PB_REQUIRE(b && re_im && sample >= 0 && sample < b->n, PSPIN_ERR_ARG, "bad sample");
    const ps::BatchSample& S = b->h_samples[b->slot_of[sample]];
    PB_TRY_HIP(hipMemcpyAsync(re_im, b->d_spins + S.spin_off, (size_t)2 * S.N * S.chains * sizeof(double),
                              hipMemcpyDeviceToHost, b->stream));
    PB_TRY_HIP(hipStreamSynchronize(b->stream));
    return PSPIN_OK;
}

int pspin_batch_get_perm(const pspin_batch* b, int sample, int* perm) {
    PB_REQUIRE(b && perm && sample >= 0 && sample < b->n, PSPIN_ERR_ARG, "bad sample");
    const ps::BatchSample& S = b->h_samples[b->slot_of[sample]];
    PB_TRY_HIP(hipMemcpyAsync(perm, b->d_perm + S.chain_off, S.chains * sizeof(int), hipMemcpyDeviceToHost, b->stream));
    PB_TRY_HIP(hipStreamSynchronize(b->stream));
    return PSPIN_OK;
}

int pspin_batch_get_exchange_stats(const pspin_batch* b, int sample, long long* acc_tot, long long* prop_tot) {
    PB_REQUIRE(b && sample >= 0 && sample < b->n, PSPIN_ERR_ARG, "bad sample");
    const int n = b->NT - 1;
    std::vector<long long> h((size_t)4 * n);
    PB_TRY_HIP(hipMemcpyAsync(h.data(), b->d_ex + (size_t)b->slot_of[sample] * 4 * n, h.size() * sizeof(long long),
                              hipMemcpyDeviceToHost, b->stream));
    PB_TRY_HIP(hipStreamSynchronize(b->stream));
    if (acc_tot) std::memcpy(acc_tot, &h[2 * n], n * sizeof(long long));
    if (prop_tot) std::memcpy(prop_tot, &h[3 * n], n * sizeof(long long));
    return PSPIN_OK;
}

}  // extern "C"
