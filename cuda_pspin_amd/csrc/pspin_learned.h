// pspin_learned.h -- state and launchers of the learned global-move proposal.
//
// The proposal is a multilayer perceptron (input 2N+1, two SiLU hidden
// layers, output 4N) whose output gives the mean and log-scale of a
// diagonal Gaussian over the 2N real coordinates of the amplitudes in a
// gauge-fixed frame.  A draw is projected radially onto the sphere
// sum |a|^2 = N and accepted by Metropolis-Hastings with the exact
// projected-Gaussian density, so the chain keeps the stationary law of the
// pair-rotation sweep (pi ~ exp(-beta H) prod_k |a_k|^-1 on the sphere).
//
// Training runs on the GPU: BF16 MFMA GEMMs with FP32 accumulation for the
// forward and backward passes, FP32 master weights and Adam.
#pragma once

#include "pspin_internal.h"

namespace ps {

constexpr uint32_t kCtrLearned = 32;  // fourth Philox counter word of learned-move draws

// Epilogues of the GEMM kernel.
enum GemmEpi { kEpiF32 = 0, kEpiBiasF32 = 1, kEpiBiasSilu = 2, kEpiDSilu = 3 };

// C(m, n) = sum_k A(m, k) B(n, k), BF16 operands, FP32 accumulation.
// A(m, k) = A[m * a_m + k * a_k]; exactly one of a_m, a_k is 1 (likewise B).
// M, N multiples of 64; K a multiple of 64 x splits; strides multiples of 8.
struct GemmArgs {
    const uint16_t* A;
    long long a_m, a_k;
    const uint16_t* B;
    long long b_n, b_k;
    int M, N, K;
    const float* bias;  // [N]          (kEpiBiasF32, kEpiBiasSilu)
    float* Cf;          // FP32 out     (kEpiF32, kEpiBiasF32; pre-activation for kEpiBiasSilu)
    long long ldc;
    uint16_t* Cb;       // BF16 out     (kEpiBiasSilu: activation; kEpiDSilu: gradient)
    long long ldcb;
    const float* Z;     // pre-activation read by kEpiDSilu
    long long ldz;
    float* rowsum;      // [splits][M] row sums of A over each K slice, or null
    long long cz;       // split-K slice strides of Cf and rowsum (0: M * ldc and M)
    long long rz;
};

// splits > 1 (kEpiF32 only): Cf receives `splits` partial products, slice z
// at Cf + z * cz, to be summed in slice order (launch_sum_slices, launch_adam).
hipError_t launch_gemm(const GemmArgs& g, int epi, int splits, hipStream_t s);
hipError_t launch_sum_slices(const float* part, int nz, long long n, float* out, hipStream_t s);
// Adam on FP32 masters with the gradient g[i] = sum over z < nz of g[z n + i]
// (split-K slices summed in order, as launch_sum_slices does).
hipError_t launch_adam(float* p, float* m, float* v, const float* g, int nz, uint16_t* pb, long long n, float lr,
                       float b1, float b2, float eps, long long t, hipStream_t s);
hipError_t launch_f32_to_bf16(const float* src, uint16_t* dst, long long n, hipStream_t s);

// Kernel arguments of one learned move over R replicas (one wave each).
struct MoveArgs {
    int N, R, in_pad, out_ld;
    double ls0, ls_min, ls_max, log_scale, mean_scale;
    const double* spins;  // [R][N][2] current amplitudes
    double* spins_w;      // same buffer, written on acceptance
    const double* betas;
    double* E;            // running energies, set to the FP64 full energy
    const uint64_t* seeds;
    uint64_t* mctr;       // [R] move counters (Philox counter)
    uint16_t* X0;         // network input rows at the current state [Rpad][in_pad]
    uint16_t* X0p;        // network input rows at the proposal
    double* U;            // [R][2N] gauge-fixed current state
    double* Up;           // [R][2N] gauge-fixed proposal
    double* gc;           // [R][2]  gauge phase of the current state
    double* gcp;          // [R][2]  gauge phase of the proposal
    const float* Of;      // [Rpad][out_ld] network output at the current state
    const float* Orv;     // [Rpad][out_ld] network output at the proposal
    double* prop;         // [R][N][2]
    double* dbg;          // [R][8]
    const double* part;   // [n_split][2R] energy partial sums
    int n_split;
    const int* perm;      // [R] slot -> replica (PT), or null
    int nrep_per_temp;
    long long* lacc;      // per temperature slot
    long long* lprop;
};

hipError_t launch_lp_prep(const MoveArgs& a, hipStream_t s);
hipError_t launch_lp_propose(const MoveArgs& a, hipStream_t s);
hipError_t launch_lp_energy(const MoveArgs& a, const uint4* terms, long long nt, double* part, hipStream_t s);
hipError_t lp_energy_prepare();  // once per process and device, before the first launch
int lp_energy_blocks(int R);     // workgroups per term split (64 configurations each)
int lp_energy_splits(long long n_terms);  // term ranges (partial sums per configuration)
constexpr int kLearnedMaxN = 128;
hipError_t launch_lp_accept(const MoveArgs& a, hipStream_t s);
hipError_t launch_lp_record(const double* spins, const double* betas, int R, int dim, float* out, float* obeta,
                            hipStream_t s);
hipError_t launch_lp_gather(const float* data, const float* dbeta, long long n_snap, int R, int N, int lag,
                            uint64_t seed, uint64_t step, int M, int in_pad, uint16_t* X0, float* U, float* Y,
                            hipStream_t s);
hipError_t launch_lp_nll(const float* O, int out_ld, const float* U, const float* Y, int dim, int M_pad, int M_valid,
                         double ls0, double ls_min, double ls_max, uint16_t* dO, double* rowloss, hipStream_t s);
hipError_t launch_sum_rows(const double* x, int M, double scale, double* out, hipStream_t s);

}  // namespace ps

// Activations of one MLP pass over `cap` rows (a multiple of 64).
struct pspin_mlp_act {
    int cap = 0;
    uint16_t* X0 = nullptr;  // [cap][in_pad]   BF16 input
    float* Z1 = nullptr;     // [cap][hid]      pre-activations
    uint16_t* H1 = nullptr;  // [cap][hid]
    float* Z2 = nullptr;
    uint16_t* H2 = nullptr;
    float* O = nullptr;      // [cap][out_dim]  linear output
    // training only
    float* U = nullptr;      // [cap][2N]  gauge-fixed input, FP32
    float* Y = nullptr;      // [cap][2N]  gauge-fixed target
    uint16_t* dO = nullptr;  // [cap][out_dim]
    uint16_t* dZ2 = nullptr; // [cap][hid]
    uint16_t* dZ1 = nullptr;
    double* rowloss = nullptr;  // [cap]
};

struct pspin_learned {
    pspin_state* st = nullptr;
    pspin_lp_config cfg{};
    int N = 0, R = 0, Rpad = 0;
    int dim = 0, in_dim = 0, in_pad = 0, hid = 0, out_dim = 0;

    // parameters: flat FP32 masters [W1 b1 W2 b2 W3 b3], Adam moments,
    // gradient, BF16 copy used by the GEMMs
    long long n_par = 0;
    long long off_w1 = 0, off_b1 = 0, off_w2 = 0, off_b2 = 0, off_w3 = 0, off_b3 = 0;
    float* d_p = nullptr;
    float* d_m = nullptr;
    float* d_v = nullptr;
    float* d_g = nullptr;
    uint16_t* d_pb = nullptr;
    long long adam_t = 0;
    float* d_gpart = nullptr;  // split-K slices of the gradient [kSplits][n_par]

    pspin_mlp_act tr;  // training batch
    pspin_mlp_act fw;  // move: forward pass at the current state
    pspin_mlp_act rv;  // move: forward pass at the proposal

    // training data: snapshots of every replica, FP32 amplitudes and beta
    float* d_data = nullptr;   // [data_cap][R][dim]
    float* d_dbeta = nullptr;  // [data_cap][R]
    long long data_cap = 0, data_n = 0;
    uint64_t train_ctr = 0;

    // move buffers
    double* d_u = nullptr;      // [R][dim] gauge-fixed current state
    double* d_up = nullptr;     // [R][dim] gauge-fixed proposal
    double* d_gc = nullptr;     // [R][2]   gauge phase of the current state
    double* d_gcp = nullptr;    // [R][2]   gauge phase of the proposal
    double* d_prop = nullptr;   // [R][N][2] proposal amplitudes
    double* d_dbg = nullptr;    // [R][8]   logq_f, logq_r, E_cur, E_new, log r, accepted, logw_cur, logw_new
    double* d_part = nullptr;   // [n_split][2R] energy partial sums
    int n_split = 0;
    uint64_t* d_mctr = nullptr; // [R] per-replica move counters
    long long* d_lacc = nullptr;   // [n_slots]
    long long* d_lprop = nullptr;  // [n_slots]
    int n_slots = 1;

    // FP64 Hamiltonian terms for full energies: {i | j<<8 | k<<16 | l<<24, ch, g}
    uint4* d_terms = nullptr;
    long long n_terms = 0;

    // captured (every/pt_freq PT rounds + one move); every == 0: the move alone
    hipGraphExec_t graph = nullptr;
    int graph_pt_freq = -1, graph_every = -1;
    uint64_t graph_gen = 0;
    ps::MoveArgs graph_args{};
};
