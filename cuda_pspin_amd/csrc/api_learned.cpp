// api_learned.cpp -- C ABI of the learned global-move proposal
// (include/pspin.h, pspin_lp_*): parameter set-up, the training loop
// (gather -> forward -> NLL -> backward -> Adam, all on the GPU), the move
// sequence (gauge -> forward -> propose -> forward -> FP64 energies -> MH)
// and its mixture with PT sweeps.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "pspin_learned.h"

#define LP_TRY_HIP(call)                                                                               \
    do {                                                                                               \
        hipError_t _e = (call);                                                                        \
        if (_e != hipSuccess) {                                                                        \
            ps::set_error(std::string("HIP error ") + hipGetErrorString(_e) + " at " + __FILE__ + ":" + \
                          std::to_string(__LINE__) + " (" #call ")");                                  \
            return PSPIN_ERR_HIP;                                                                      \
        }                                                                                              \
    } while (0)

#define LP_REQUIRE(cond, code, msg) \
    do {                            \
        if (!(cond)) {              \
            ps::set_error(msg);     \
            return code;            \
        }                           \
    } while (0)

namespace {

template <class T>
hipError_t dmalloc(T** p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    return hipMalloc((void**)p, n * sizeof(T));
}

template <class T>
void dfree(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

uint16_t host_bf16(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

int round64(long long x) { return (int)((x + 63) / 64 * 64); }

hipError_t alloc_act(pspin_mlp_act& a, int cap, const pspin_learned* lp, bool train) {
    a.cap = cap;
    hipError_t e;
    if ((e = dmalloc(&a.X0, (size_t)cap * lp->in_pad))) return e;
    if ((e = dmalloc(&a.Z1, (size_t)cap * lp->hid))) return e;
    if ((e = dmalloc(&a.H1, (size_t)cap * lp->hid))) return e;
    if ((e = dmalloc(&a.Z2, (size_t)cap * lp->hid))) return e;
    if ((e = dmalloc(&a.H2, (size_t)cap * lp->hid))) return e;
    if ((e = dmalloc(&a.O, (size_t)cap * lp->out_dim))) return e;
    if ((e = hipMemsetAsync(a.X0, 0, (size_t)cap * lp->in_pad * sizeof(uint16_t), lp->st->stream))) return e;
    if (!train) return hipSuccess;
    if ((e = dmalloc(&a.U, (size_t)cap * lp->dim))) return e;
    if ((e = dmalloc(&a.Y, (size_t)cap * lp->dim))) return e;
    if ((e = dmalloc(&a.dO, (size_t)cap * lp->out_dim))) return e;
    if ((e = dmalloc(&a.dZ2, (size_t)cap * lp->hid))) return e;
    if ((e = dmalloc(&a.dZ1, (size_t)cap * lp->hid))) return e;
    if ((e = dmalloc(&a.rowloss, (size_t)cap))) return e;
    return hipSuccess;
}

void free_act(pspin_mlp_act& a) {
    dfree(a.X0); dfree(a.Z1); dfree(a.H1); dfree(a.Z2); dfree(a.H2); dfree(a.O);
    dfree(a.U); dfree(a.Y); dfree(a.dO); dfree(a.dZ2); dfree(a.dZ1); dfree(a.rowloss);
}

void free_lp(pspin_learned* lp) {
    free_act(lp->tr); free_act(lp->fw); free_act(lp->rv);
    dfree(lp->d_p); dfree(lp->d_m); dfree(lp->d_v); dfree(lp->d_g); dfree(lp->d_pb);
    dfree(lp->d_gpart);
    dfree(lp->d_data); dfree(lp->d_dbeta);
    dfree(lp->d_u); dfree(lp->d_up); dfree(lp->d_gc); dfree(lp->d_gcp); dfree(lp->d_prop); dfree(lp->d_dbg);
    dfree(lp->d_part); dfree(lp->d_mctr); dfree(lp->d_lacc); dfree(lp->d_lprop); dfree(lp->d_terms);
    if (lp->graph) (void)hipGraphExecDestroy(lp->graph);
    lp->graph = nullptr;
}

hipError_t sync(const pspin_learned* lp) { return hipStreamSynchronize(lp->st->stream); }

constexpr int kMaxSplits = 8;

// Forward pass over M (multiple of 64) rows of act.X0.
hipError_t forward(pspin_learned* lp, pspin_mlp_act& a, int M) {
    const hipStream_t s = lp->st->stream;
    ps::GemmArgs g{};
    g.A = a.X0; g.a_m = lp->in_pad; g.a_k = 1;
    g.B = lp->d_pb + lp->off_w1; g.b_n = lp->in_pad; g.b_k = 1;
    g.M = M; g.N = lp->hid; g.K = lp->in_pad;
    g.bias = lp->d_p + lp->off_b1; g.Cf = a.Z1; g.ldc = lp->hid; g.Cb = a.H1; g.ldcb = lp->hid;
    hipError_t e = ps::launch_gemm(g, ps::kEpiBiasSilu, 1, s);
    if (e) return e;
    g.A = a.H1; g.a_m = lp->hid;
    g.B = lp->d_pb + lp->off_w2; g.b_n = lp->hid;
    g.K = lp->hid;
    g.bias = lp->d_p + lp->off_b2; g.Cf = a.Z2; g.Cb = a.H2;
    if ((e = ps::launch_gemm(g, ps::kEpiBiasSilu, 1, s))) return e;
    g.A = a.H2;
    g.B = lp->d_pb + lp->off_w3;
    g.N = lp->out_dim;
    g.bias = lp->d_p + lp->off_b3; g.Cf = a.O; g.ldc = lp->out_dim; g.Cb = nullptr;
    return ps::launch_gemm(g, ps::kEpiBiasF32, 1, s);
}

// K splits of a weight-gradient product over M batch rows: the largest
// power of two <= kMaxSplits with M a multiple of 64 x splits.
int grad_splits(int M) {
    int sp = kMaxSplits;
    while (sp > 1 && M % (64 * sp)) sp >>= 1;
    return sp;
}

// The gradient of a backward pass over M rows: with split K it is left as
// grad_splits(M) slices of the flat parameter vector in d_gpart (slice z at
// d_gpart + z n_par), to be summed in slice order; otherwise it is d_g.
const float* grad_slices(const pspin_learned* lp, int M) { return grad_splits(M) > 1 ? lp->d_gpart : lp->d_g; }

// dW [rows][cols] = dZ^T X and db [rows] = column sums of dZ, dZ [M][rows]
// and X [M][cols] BF16, at (off_w, off_b) of the gradient's slices.
hipError_t weight_grad(pspin_learned* lp, const uint16_t* dZ, int rows, const uint16_t* X, int cols, int M,
                       long long off_w, long long off_b) {
    const hipStream_t s = lp->st->stream;
    const int sp = grad_splits(M);
    float* G = (float*)grad_slices(lp, M);
    ps::GemmArgs g{};
    g.A = dZ; g.a_m = 1; g.a_k = rows;
    g.B = X; g.b_n = 1; g.b_k = cols;
    g.M = rows; g.N = cols; g.K = M;
    g.Cf = G + off_w; g.ldc = cols; g.cz = lp->n_par;
    g.rowsum = G + off_b; g.rz = lp->n_par;
    return ps::launch_gemm(g, ps::kEpiF32, sp, s);
}

// dX = (dZ W) * silu'(Z) into BF16 out, dZ [M][n_out], W [n_out][n_in].
hipError_t data_grad(pspin_learned* lp, const uint16_t* dZ, int n_out, long long off_w, int n_in, const float* Z,
                     uint16_t* out, int M) {
    ps::GemmArgs g{};
    g.A = dZ; g.a_m = n_out; g.a_k = 1;
    g.B = lp->d_pb + off_w; g.b_n = 1; g.b_k = n_in;
    g.M = M; g.N = n_in; g.K = n_out;
    g.Z = Z; g.ldz = n_in; g.Cb = out; g.ldcb = n_in;
    return ps::launch_gemm(g, ps::kEpiDSilu, 1, lp->st->stream);
}

// Backward pass from act.dO (BF16) into lp->d_g.
hipError_t backward(pspin_learned* lp, pspin_mlp_act& a, int M) {
    hipError_t e;
    if ((e = weight_grad(lp, a.dO, lp->out_dim, a.H2, lp->hid, M, lp->off_w3, lp->off_b3))) return e;
    if ((e = data_grad(lp, a.dO, lp->out_dim, lp->off_w3, lp->hid, a.Z2, a.dZ2, M))) return e;
    if ((e = weight_grad(lp, a.dZ2, lp->hid, a.H1, lp->hid, M, lp->off_w2, lp->off_b2))) return e;
    if ((e = data_grad(lp, a.dZ2, lp->hid, lp->off_w2, lp->hid, a.Z1, a.dZ1, M))) return e;
    return weight_grad(lp, a.dZ1, lp->hid, a.X0, lp->in_pad, M, lp->off_w1, lp->off_b1);
}

ps::MoveArgs move_args(pspin_learned* lp) {
    pspin_state* st = lp->st;
    ps::MoveArgs a{};
    a.N = lp->N; a.R = lp->R; a.in_pad = lp->in_pad; a.out_ld = lp->out_dim;
    a.ls0 = lp->cfg.log_sigma0; a.ls_min = lp->cfg.log_sigma_min; a.ls_max = lp->cfg.log_sigma_max;
    a.log_scale = std::log(lp->cfg.scale);
    a.mean_scale = lp->cfg.mean_scale;
    a.spins = st->d_spins; a.spins_w = st->d_spins; a.betas = st->d_betas; a.E = st->d_E;
    a.seeds = st->d_seeds; a.mctr = lp->d_mctr;
    a.X0 = lp->fw.X0; a.X0p = lp->rv.X0;
    a.U = lp->d_u; a.Up = lp->d_up; a.gc = lp->d_gc; a.gcp = lp->d_gcp;
    a.Of = lp->fw.O; a.Orv = lp->rv.O;
    a.prop = lp->d_prop; a.dbg = lp->d_dbg; a.part = lp->d_part; a.n_split = lp->n_split;
    a.perm = st->pt_NT >= 2 ? st->d_perm : nullptr;
    a.nrep_per_temp = st->pt_NT >= 2 ? st->pt_nrep : 1;
    a.lacc = lp->d_lacc; a.lprop = lp->d_lprop;
    return a;
}

hipError_t move_launch(pspin_learned* lp, const ps::MoveArgs& a) {
    const hipStream_t s = lp->st->stream;
    hipError_t e;
    if ((e = ps::launch_lp_prep(a, s))) return e;
    if ((e = forward(lp, lp->fw, lp->Rpad))) return e;
    if ((e = ps::launch_lp_propose(a, s))) return e;
    if ((e = forward(lp, lp->rv, lp->Rpad))) return e;
    if ((e = ps::launch_lp_energy(a, lp->d_terms, lp->n_terms, lp->d_part, s))) return e;
    return ps::launch_lp_accept(a, s);
}

// Capture (every / pt_freq rounds of pt_freq sweeps + one exchange, then one
// learned move) into lp->graph unless the current capture matches.  All
// kernels read their counters and PT state from device memory, so replays
// continue the chains.
int ensure_graph(pspin_learned* lp, int pt_freq, int every) {
    const ps::MoveArgs a = move_args(lp);
    if (lp->graph && lp->graph_pt_freq == pt_freq && lp->graph_every == every &&
        lp->graph_gen == lp->st->graph_gen && std::memcmp(&a, &lp->graph_args, sizeof(a)) == 0)
        return PSPIN_OK;
    if (lp->graph) (void)hipGraphExecDestroy(lp->graph);
    lp->graph = nullptr;
    pspin_state* st = lp->st;
    hipGraph_t g = nullptr;
    LP_TRY_HIP(hipStreamBeginCapture(st->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (int k = 0; every > 0 && k < every / pt_freq && e == hipSuccess; k++) {
        e = ps::launch_sweep(st, pt_freq, true);
    }
    if (e == hipSuccess) e = move_launch(lp, a);
    hipError_t e2 = hipStreamEndCapture(st->stream, &g);
    LP_TRY_HIP(e);
    LP_TRY_HIP(e2);
    hipError_t e3 = hipGraphInstantiate(&lp->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    LP_TRY_HIP(e3);
    lp->graph_pt_freq = pt_freq;
    lp->graph_every = every;
    lp->graph_args = a;
    lp->graph_gen = st->graph_gen;
    return PSPIN_OK;
}

// FP64 Hamiltonian terms of the state: H4 quartets (colex walk of the dense
// coupling table and mask) and, when H2 is active, the pairs i < j.  A term
// is {i | j << 7 | k << 14 | l << 21 | ch << 28, 0, g} (ch 3: pair), and
// the terms of one (j, k, l, ch) are consecutive: the energy kernel sums
// g a_i over such a run before multiplying by the product of the others.
int build_terms(pspin_learned* lp) {
    pspin_state* st = lp->st;
    const int N = lp->N;
    std::vector<uint4> terms;
    auto push = [&](uint32_t i, uint32_t j, uint32_t k, uint32_t l, uint32_t ch, double g) {
        uint64_t b;
        std::memcpy(&b, &g, 8);
        terms.push_back(make_uint4(i | j << 7 | k << 14 | l << 21 | ch << 28, 0u, (uint32_t)b, (uint32_t)(b >> 32)));
    };
    const long long nq = ps::n_quartets(N);
    std::vector<double> g4((size_t)nq);
    std::vector<uint8_t> mask((size_t)nq);
    LP_TRY_HIP(hipMemcpyAsync(g4.data(), st->d_g4_64, (size_t)nq * 8, hipMemcpyDeviceToHost, st->stream));
    LP_TRY_HIP(hipMemcpyAsync(mask.data(), st->d_mask, (size_t)nq, hipMemcpyDeviceToHost, st->stream));
    LP_TRY_HIP(hipStreamSynchronize(st->stream));
    long long q = 0;  // colex index of (0, j, k, l)
    for (int l = 3; l < N; l++)
        for (int k = 2; k < l; k++)
            for (int j = 1; j < k; q += j, j++)
                for (uint32_t ch = 0; ch < 3; ch++)
                    for (int i = 0; i < j; i++) {
                        const uint8_t m = mask[(size_t)(q + i)];
                        const double g = g4[(size_t)(q + i)];
                        if (m == 0 || g == 0.0 || ch != ((m & 1) ? 0u : ((m & 2) ? 1u : 2u))) continue;
                        push(i, j, k, l, ch, g);
                    }
    if (q != nq) {
        ps::set_error("learned move: quartet walk does not cover the coupling table");
        return PSPIN_ERR_ARG;
    }
    if (st->h2_active) {
        std::vector<double> g2((size_t)N * N);
        LP_TRY_HIP(hipMemcpyAsync(g2.data(), st->d_g2_64, (size_t)N * N * 8, hipMemcpyDeviceToHost, st->stream));
        LP_TRY_HIP(hipStreamSynchronize(st->stream));
        for (int j = 1; j < N; j++)
            for (int i = 0; i < j; i++)
                if (g2[(size_t)i * N + j] != 0.0) push(i, j, 0, 0, 3, g2[(size_t)i * N + j]);
    }
    lp->n_terms = (long long)terms.size();
    LP_TRY_HIP(dmalloc(&lp->d_terms, terms.size()));
    if (!terms.empty())
        LP_TRY_HIP(hipMemcpy(lp->d_terms, terms.data(), terms.size() * sizeof(uint4), hipMemcpyHostToDevice));
    lp->n_split = ps::lp_energy_splits(lp->n_terms);
    LP_TRY_HIP(ps::lp_energy_prepare());
    LP_TRY_HIP(dmalloc(&lp->d_part, (size_t)lp->n_split * 2 * lp->R));
    return PSPIN_OK;
}

// Host-side initial parameters: uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)),
// output layer scaled by init_gain, padding columns of W1 zero.
std::vector<float> init_params(const pspin_learned* lp) {
    std::vector<float> p((size_t)lp->n_par, 0.0f);
    uint64_t s = lp->cfg.seed ^ 0x6C65617264ULL;
    auto uni = [&](double bound) {
        const uint64_t x = ps::splitmix64(&s);
        return (float)(((double)(x >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0) * bound);
    };
    const double b1 = 1.0 / std::sqrt((double)lp->in_dim), b2 = 1.0 / std::sqrt((double)lp->hid);
    for (int o = 0; o < lp->hid; o++)
        for (int k = 0; k < lp->in_dim; k++) p[(size_t)(lp->off_w1 + (long long)o * lp->in_pad + k)] = uni(b1);
    for (int o = 0; o < lp->hid; o++) p[(size_t)(lp->off_b1 + o)] = uni(b1);
    for (long long i = 0; i < (long long)lp->hid * lp->hid; i++) p[(size_t)(lp->off_w2 + i)] = uni(b2);
    for (int o = 0; o < lp->hid; o++) p[(size_t)(lp->off_b2 + o)] = uni(b2);
    const double g3 = b2 * lp->cfg.init_gain;
    for (long long i = 0; i < (long long)lp->out_dim * lp->hid; i++) p[(size_t)(lp->off_w3 + i)] = uni(g3);
    for (int o = 0; o < lp->out_dim; o++) p[(size_t)(lp->off_b3 + o)] = uni(g3);
    return p;
}

int upload_params(pspin_learned* lp, const float* p) {
    const hipStream_t s = lp->st->stream;
    const size_t n = (size_t)lp->n_par;
    LP_TRY_HIP(hipMemcpyAsync(lp->d_p, p, n * 4, hipMemcpyHostToDevice, s));
    LP_TRY_HIP(hipMemsetAsync(lp->d_m, 0, n * 4, s));
    LP_TRY_HIP(hipMemsetAsync(lp->d_v, 0, n * 4, s));
    LP_TRY_HIP(ps::launch_f32_to_bf16(lp->d_p, lp->d_pb, (long long)n, s));
    LP_TRY_HIP(hipStreamSynchronize(s));
    lp->adam_t = 0;
    return PSPIN_OK;
}

// Host rows x [M][in_dim] -> BF16 act.X0 [Mpad][in_pad] (zero padding).
int upload_inputs(pspin_learned* lp, const float* x, int M, int Mpad) {
    std::vector<uint16_t> h((size_t)Mpad * lp->in_pad, 0);
    for (int m = 0; m < M; m++)
        for (int k = 0; k < lp->in_dim; k++)
            h[(size_t)m * lp->in_pad + k] = host_bf16(x[(size_t)m * lp->in_dim + k]);
    LP_TRY_HIP(hipMemcpyAsync(lp->tr.X0, h.data(), h.size() * 2, hipMemcpyHostToDevice, lp->st->stream));
    LP_TRY_HIP(hipStreamSynchronize(lp->st->stream));
    return PSPIN_OK;
}

}  // namespace

extern "C" {

int pspin_lp_default_config(pspin_lp_config* cfg) {
    LP_REQUIRE(cfg, PSPIN_ERR_ARG, "null config");
    cfg->hidden = 512;
    cfg->batch_cap = 4096;
    cfg->data_cap = 1024;
    cfg->log_sigma0 = std::log(0.05);
    cfg->log_sigma_min = -9.0;
    cfg->log_sigma_max = 1.0;
    cfg->scale = 1.0;
    cfg->init_gain = 0.01;
    cfg->seed = 12345;
    cfg->mean_scale = 1.0;
    return PSPIN_OK;
}

int pspin_lp_create(pspin_state* st, const pspin_lp_config* cfg, pspin_learned** out) {
    LP_REQUIRE(st && cfg && out, PSPIN_ERR_ARG, "null argument");
    *out = nullptr;
    LP_REQUIRE(st->variant == PSPIN_DENSE, PSPIN_ERR_ARG, "the learned move needs the spherical (dense) variant");
    LP_REQUIRE(st->N % 16 == 0 && st->N <= ps::kLearnedMaxN, PSPIN_ERR_ARG,
               "the learned move needs N a multiple of 16, at most 128");
    LP_REQUIRE(cfg->hidden >= 64 && cfg->hidden % 64 == 0, PSPIN_ERR_ARG, "hidden must be a positive multiple of 64");
    LP_REQUIRE(cfg->batch_cap >= 64 && cfg->batch_cap % 64 == 0, PSPIN_ERR_ARG,
               "batch_cap must be a positive multiple of 64");
    LP_REQUIRE(cfg->data_cap >= 2, PSPIN_ERR_ARG, "data_cap must be >= 2");
    LP_REQUIRE(cfg->scale > 0.0 && cfg->log_sigma_min < cfg->log_sigma_max, PSPIN_ERR_ARG, "bad scale parameters");
    LP_TRY_HIP(hipSetDevice(st->device));
    pspin_learned* lp = new pspin_learned();
    lp->st = st;
    lp->cfg = *cfg;
    lp->N = st->N;
    lp->R = st->nrep;
    lp->Rpad = round64(lp->R);
    lp->dim = 2 * lp->N;
    lp->in_dim = lp->dim + 1;
    lp->in_pad = round64(lp->in_dim);
    lp->hid = cfg->hidden;
    lp->out_dim = 2 * lp->dim;
    lp->off_w1 = 0;
    lp->off_b1 = lp->off_w1 + (long long)lp->hid * lp->in_pad;
    lp->off_w2 = lp->off_b1 + lp->hid;
    lp->off_b2 = lp->off_w2 + (long long)lp->hid * lp->hid;
    lp->off_w3 = lp->off_b2 + lp->hid;
    lp->off_b3 = lp->off_w3 + (long long)lp->out_dim * lp->hid;
    lp->n_par = lp->off_b3 + lp->out_dim;
    lp->data_cap = cfg->data_cap;
    auto fail = [&](int rc) {
        free_lp(lp);
        delete lp;
        return rc;
    };
    hipError_t e = hipSuccess;
    const size_t np = (size_t)lp->n_par;
    if (!e) e = dmalloc(&lp->d_p, np);
    if (!e) e = dmalloc(&lp->d_m, np);
    if (!e) e = dmalloc(&lp->d_v, np);
    if (!e) e = dmalloc(&lp->d_g, np);
    if (!e) e = dmalloc(&lp->d_pb, np);
    if (!e) e = dmalloc(&lp->d_gpart, (size_t)kMaxSplits * np);
    if (!e) e = alloc_act(lp->tr, cfg->batch_cap, lp, true);
    if (!e) e = alloc_act(lp->fw, lp->Rpad, lp, false);
    if (!e) e = alloc_act(lp->rv, lp->Rpad, lp, false);
    if (!e) e = dmalloc(&lp->d_data, (size_t)lp->data_cap * lp->R * lp->dim);
    if (!e) e = dmalloc(&lp->d_dbeta, (size_t)lp->data_cap * lp->R);
    if (!e) e = dmalloc(&lp->d_u, (size_t)lp->R * lp->dim);
    if (!e) e = dmalloc(&lp->d_up, (size_t)lp->R * lp->dim);
    if (!e) e = dmalloc(&lp->d_gc, (size_t)lp->R * 2);
    if (!e) e = dmalloc(&lp->d_gcp, (size_t)lp->R * 2);
    if (!e) e = dmalloc(&lp->d_prop, (size_t)lp->R * lp->dim);
    if (!e) e = dmalloc(&lp->d_dbg, (size_t)lp->R * 8);
    if (!e) e = dmalloc(&lp->d_mctr, (size_t)lp->R);
    if (!e) e = dmalloc(&lp->d_lacc, (size_t)lp->R);
    if (!e) e = dmalloc(&lp->d_lprop, (size_t)lp->R);
    if (!e) e = hipMemsetAsync(lp->d_mctr, 0, (size_t)lp->R * 8, st->stream);
    if (!e) e = hipMemsetAsync(lp->d_lacc, 0, (size_t)lp->R * 8, st->stream);
    if (!e) e = hipMemsetAsync(lp->d_lprop, 0, (size_t)lp->R * 8, st->stream);
    if (!e) e = hipMemsetAsync(lp->d_dbg, 0, (size_t)lp->R * 8 * 8, st->stream);
    if (!e) e = hipMemsetAsync(lp->d_g, 0, np * 4, st->stream);
    if (e) {
        ps::set_error(std::string("HIP error in pspin_lp_create: ") + hipGetErrorString(e));
        return fail(e == hipErrorOutOfMemory ? PSPIN_ERR_NOMEM : PSPIN_ERR_HIP);
    }
    int rc = build_terms(lp);
    if (rc) return fail(rc);
    const std::vector<float> p = init_params(lp);
    if ((rc = upload_params(lp, p.data()))) return fail(rc);
    *out = lp;
    return PSPIN_OK;
}

int pspin_lp_destroy(pspin_learned* lp) {
    if (!lp) return PSPIN_OK;
    (void)hipStreamSynchronize(lp->st->stream);
    free_lp(lp);
    delete lp;
    return PSPIN_OK;
}

long long pspin_lp_n_params(const pspin_learned* lp) { return lp ? lp->n_par : -1; }

int pspin_lp_dims(const pspin_learned* lp, int* in_dim, int* in_pad, int* hidden, int* out_dim) {
    LP_REQUIRE(lp, PSPIN_ERR_ARG, "null handle");
    if (in_dim) *in_dim = lp->in_dim;
    if (in_pad) *in_pad = lp->in_pad;
    if (hidden) *hidden = lp->hid;
    if (out_dim) *out_dim = lp->out_dim;
    return PSPIN_OK;
}

int pspin_lp_get_params(const pspin_learned* lp, float* out) {
    LP_REQUIRE(lp && out, PSPIN_ERR_ARG, "null argument");
    LP_TRY_HIP(hipMemcpyAsync(out, lp->d_p, (size_t)lp->n_par * 4, hipMemcpyDeviceToHost, lp->st->stream));
    LP_TRY_HIP(sync(lp));
    return PSPIN_OK;
}

int pspin_lp_set_params(pspin_learned* lp, const float* in) {
    LP_REQUIRE(lp && in, PSPIN_ERR_ARG, "null argument");
    return upload_params(lp, in);
}

int pspin_lp_set_scale(pspin_learned* lp, double scale) {
    LP_REQUIRE(lp && scale > 0.0, PSPIN_ERR_ARG, "scale must be > 0");
    lp->cfg.scale = scale;
    return PSPIN_OK;
}

int pspin_lp_set_mean_scale(pspin_learned* lp, double mean_scale) {
    LP_REQUIRE(lp && mean_scale >= 0.0, PSPIN_ERR_ARG, "mean scale must be >= 0");
    lp->cfg.mean_scale = mean_scale;
    return PSPIN_OK;
}

int pspin_lp_forward(pspin_learned* lp, const float* x, int M, float* out) {
    LP_REQUIRE(lp && x && out && M > 0 && M <= lp->cfg.batch_cap, PSPIN_ERR_ARG, "bad argument");
    const int Mpad = round64(M);
    int rc = upload_inputs(lp, x, M, Mpad);
    if (rc) return rc;
    LP_TRY_HIP(forward(lp, lp->tr, Mpad));
    LP_TRY_HIP(hipMemcpyAsync(out, lp->tr.O, (size_t)M * lp->out_dim * 4, hipMemcpyDeviceToHost, lp->st->stream));
    LP_TRY_HIP(sync(lp));
    return PSPIN_OK;
}

int pspin_lp_loss_grad(pspin_learned* lp, const float* x, const float* u, const float* y, int M, double* loss,
                       float* grad) {
    LP_REQUIRE(lp && x && u && y && M > 0 && M <= lp->cfg.batch_cap, PSPIN_ERR_ARG, "bad argument");
    const int Mpad = round64(M);
    const hipStream_t s = lp->st->stream;
    int rc = upload_inputs(lp, x, M, Mpad);
    if (rc) return rc;
    LP_TRY_HIP(hipMemcpyAsync(lp->tr.U, u, (size_t)M * lp->dim * 4, hipMemcpyHostToDevice, s));
    LP_TRY_HIP(hipMemcpyAsync(lp->tr.Y, y, (size_t)M * lp->dim * 4, hipMemcpyHostToDevice, s));
    LP_TRY_HIP(forward(lp, lp->tr, Mpad));
    LP_TRY_HIP(ps::launch_lp_nll(lp->tr.O, lp->out_dim, lp->tr.U, lp->tr.Y, lp->dim, Mpad, M, lp->cfg.log_sigma0,
                                 lp->cfg.log_sigma_min, lp->cfg.log_sigma_max, lp->tr.dO, lp->tr.rowloss, s));
    LP_TRY_HIP(backward(lp, lp->tr, Mpad));
    if (grad_splits(Mpad) > 1) LP_TRY_HIP(ps::launch_sum_slices(lp->d_gpart, grad_splits(Mpad), lp->n_par, lp->d_g, s));
    std::vector<double> rl((size_t)M);
    LP_TRY_HIP(hipMemcpyAsync(rl.data(), lp->tr.rowloss, (size_t)M * 8, hipMemcpyDeviceToHost, s));
    if (grad) LP_TRY_HIP(hipMemcpyAsync(grad, lp->d_g, (size_t)lp->n_par * 4, hipMemcpyDeviceToHost, s));
    LP_TRY_HIP(hipStreamSynchronize(s));
    if (loss) {
        double t = 0.0;
        for (double v : rl) t += v;
        *loss = t / M;
    }
    return PSPIN_OK;
}

int pspin_lp_record(pspin_learned* lp) {
    LP_REQUIRE(lp, PSPIN_ERR_ARG, "null handle");
    LP_REQUIRE(lp->data_n < lp->data_cap, PSPIN_ERR_STATE, "training data buffer is full");
    pspin_state* st = lp->st;
    LP_TRY_HIP(ps::launch_lp_record(st->d_spins, st->d_betas, lp->R, lp->dim,
                                    lp->d_data + (size_t)lp->data_n * lp->R * lp->dim,
                                    lp->d_dbeta + (size_t)lp->data_n * lp->R, st->stream));
    lp->data_n++;
    return PSPIN_OK;
}

int pspin_lp_collect(pspin_learned* lp, long long n_snapshots, int every, int pt_freq) {
    LP_REQUIRE(lp && n_snapshots >= 0 && every >= 1 && pt_freq >= 1, PSPIN_ERR_ARG, "bad argument");
    LP_REQUIRE(lp->data_n + n_snapshots <= lp->data_cap, PSPIN_ERR_ARG, "snapshots exceed data_cap");
    pspin_state* st = lp->st;
    for (long long i = 0; i < n_snapshots; i++) {
        const int rc = st->pt_NT >= 2 ? pspin_pt_run(st, every, pt_freq) : pspin_sweep(st, every);
        if (rc) return rc;
        if (const int r2 = pspin_lp_record(lp)) return r2;
    }
    LP_TRY_HIP(sync(lp));
    return PSPIN_OK;
}

long long pspin_lp_data_size(const pspin_learned* lp) { return lp ? lp->data_n : -1; }

int pspin_lp_train(pspin_learned* lp, long long n_steps, int batch, int lag, double lr, double* loss_hist) {
    LP_REQUIRE(lp && n_steps >= 0 && lr > 0.0, PSPIN_ERR_ARG, "bad argument");
    LP_REQUIRE(batch >= 64 && batch % 64 == 0 && batch <= lp->cfg.batch_cap, PSPIN_ERR_ARG,
               "batch must be a multiple of 64 within batch_cap");
    LP_REQUIRE(lag >= 1 && lp->data_n > lag, PSPIN_ERR_STATE, "not enough snapshots for this lag");
    const hipStream_t s = lp->st->stream;
    double* d_hist = nullptr;
    LP_TRY_HIP(dmalloc(&d_hist, (size_t)std::max(1LL, n_steps)));
    int rc = PSPIN_OK;
    for (long long it = 0; it < n_steps && rc == PSPIN_OK; it++) {
        hipError_t e = ps::launch_lp_gather(lp->d_data, lp->d_dbeta, lp->data_n, lp->R, lp->N, lag, lp->cfg.seed,
                                            lp->train_ctr++, batch, lp->in_pad, lp->tr.X0, lp->tr.U, lp->tr.Y, s);
        if (!e) e = forward(lp, lp->tr, batch);
        if (!e)
            e = ps::launch_lp_nll(lp->tr.O, lp->out_dim, lp->tr.U, lp->tr.Y, lp->dim, batch, batch,
                                  lp->cfg.log_sigma0, lp->cfg.log_sigma_min, lp->cfg.log_sigma_max, lp->tr.dO,
                                  lp->tr.rowloss, s);
        if (!e) e = backward(lp, lp->tr, batch);
        if (!e) e = ps::launch_adam(lp->d_p, lp->d_m, lp->d_v, grad_slices(lp, batch), grad_splits(batch), lp->d_pb,
                                    lp->n_par, (float)lr, 0.9f, 0.999f, 1e-8f, ++lp->adam_t, s);
        if (!e) e = ps::launch_sum_rows(lp->tr.rowloss, batch, 1.0 / batch, d_hist + it, s);
        if (e) {
            ps::set_error(std::string("HIP error in pspin_lp_train: ") + hipGetErrorString(e));
            rc = PSPIN_ERR_HIP;
        }
    }
    if (rc == PSPIN_OK && loss_hist && n_steps > 0) {
        hipError_t e = hipMemcpyAsync(loss_hist, d_hist, (size_t)n_steps * 8, hipMemcpyDeviceToHost, s);
        if (e) rc = PSPIN_ERR_HIP;
    }
    hipError_t e = hipStreamSynchronize(s);
    if (e && rc == PSPIN_OK) {
        ps::set_error(std::string("HIP error in pspin_lp_train: ") + hipGetErrorString(e));
        rc = PSPIN_ERR_HIP;
    }
    dfree(d_hist);
    return rc;
}

int pspin_lp_move(pspin_learned* lp) {
    LP_REQUIRE(lp, PSPIN_ERR_ARG, "null handle");
    if (lp->st->prof || lp->st->d_log) {  // per-launch recording: no graph
        LP_TRY_HIP(move_launch(lp, move_args(lp)));
    } else {
        int rc = ensure_graph(lp, 1, 0);
        if (rc) return rc;
        LP_TRY_HIP(hipGraphLaunch(lp->graph, lp->st->stream));
    }
    LP_TRY_HIP(sync(lp));
    lp->st->since_resync = 0;  // the move leaves every running energy at its FP64 value
    return PSPIN_OK;
}

int pspin_lp_run(pspin_learned* lp, long long n_sweeps, int pt_freq, int move_every) {
    LP_REQUIRE(lp && n_sweeps >= 0 && pt_freq >= 1 && move_every >= 1, PSPIN_ERR_ARG, "bad argument");
    LP_REQUIRE(move_every % pt_freq == 0, PSPIN_ERR_ARG, "move_every must be a multiple of pt_freq");
    pspin_state* st = lp->st;
    LP_REQUIRE(st->pt_NT >= 2, PSPIN_ERR_STATE, "pspin_pt_init has not been called");
    const bool can_graph = !st->prof && !st->d_log;
    // The sweeps between two moves as one cooperative launch of their rounds
    // (plain PT's launch path), then the move's captured graph; or, where
    // cooperative launches are off, one graph of rounds and move together.
    const bool coop = can_graph && st->launch_mode == PSPIN_LAUNCH_AUTO && ps::pt_rounds_ok(st);
    long long done = 0;
    while (done < n_sweeps) {
        const int chunk = (int)std::min<long long>(move_every, n_sweeps - done);
        if (chunk == move_every && coop && st->pt_sweeps % pt_freq == 0) {
            int rc = ps::pt_advance(st, chunk, pt_freq, chunk / pt_freq);
            if (rc) return rc;
            if ((rc = ensure_graph(lp, 1, 0))) return rc;
            LP_TRY_HIP(hipGraphLaunch(lp->graph, st->stream));
            st->since_resync = 0;
        } else if (chunk == move_every && can_graph && st->pt_sweeps % pt_freq == 0) {
            int rc = ensure_graph(lp, pt_freq, move_every);
            if (rc) return rc;
            LP_TRY_HIP(hipGraphLaunch(lp->graph, st->stream));
            st->sweep_counter += (uint64_t)chunk;
            st->pt_sweeps += chunk;
            st->since_resync = 0;
        } else {
            int rc = ps::pt_advance(st, chunk, pt_freq);
            if (rc) return rc;
            if (chunk == move_every) {
                LP_TRY_HIP(move_launch(lp, move_args(lp)));
                st->since_resync = 0;
            }
        }
        done += chunk;
    }
    LP_TRY_HIP(sync(lp));
    return ps::pt_check_barrier(st);
}

int pspin_lp_get_stats(const pspin_learned* lp, long long* acc, long long* prop) {
    LP_REQUIRE(lp && acc && prop, PSPIN_ERR_ARG, "null argument");
    const int n = lp->st->pt_NT >= 2 ? lp->st->pt_NT : 1;
    LP_TRY_HIP(hipMemcpyAsync(acc, lp->d_lacc, (size_t)n * 8, hipMemcpyDeviceToHost, lp->st->stream));
    LP_TRY_HIP(hipMemcpyAsync(prop, lp->d_lprop, (size_t)n * 8, hipMemcpyDeviceToHost, lp->st->stream));
    LP_TRY_HIP(sync(lp));
    return PSPIN_OK;
}

int pspin_lp_reset_stats(pspin_learned* lp) {
    LP_REQUIRE(lp, PSPIN_ERR_ARG, "null handle");
    LP_TRY_HIP(hipMemsetAsync(lp->d_lacc, 0, (size_t)lp->R * 8, lp->st->stream));
    LP_TRY_HIP(hipMemsetAsync(lp->d_lprop, 0, (size_t)lp->R * 8, lp->st->stream));
    LP_TRY_HIP(sync(lp));
    return PSPIN_OK;
}

int pspin_lp_get_last(const pspin_learned* lp, double* prop, double* dbg, float* out_fwd, float* out_rev) {
    LP_REQUIRE(lp, PSPIN_ERR_ARG, "null handle");
    const hipStream_t s = lp->st->stream;
    if (prop) LP_TRY_HIP(hipMemcpyAsync(prop, lp->d_prop, (size_t)lp->R * lp->dim * 8, hipMemcpyDeviceToHost, s));
    if (dbg) LP_TRY_HIP(hipMemcpyAsync(dbg, lp->d_dbg, (size_t)lp->R * 8 * 8, hipMemcpyDeviceToHost, s));
    if (out_fwd)
        LP_TRY_HIP(hipMemcpyAsync(out_fwd, lp->fw.O, (size_t)lp->R * lp->out_dim * 4, hipMemcpyDeviceToHost, s));
    if (out_rev)
        LP_TRY_HIP(hipMemcpyAsync(out_rev, lp->rv.O, (size_t)lp->R * lp->out_dim * 4, hipMemcpyDeviceToHost, s));
    LP_TRY_HIP(sync(lp));
    return PSPIN_OK;
}

int pspin_lp_get_move_counters(const pspin_learned* lp, uint64_t* ctr) {
    LP_REQUIRE(lp && ctr, PSPIN_ERR_ARG, "null argument");
    LP_TRY_HIP(hipMemcpyAsync(ctr, lp->d_mctr, (size_t)lp->R * 8, hipMemcpyDeviceToHost, lp->st->stream));
    LP_TRY_HIP(sync(lp));
    return PSPIN_OK;
}

}  // extern "C"
