// disorder.hip -- coupling generation, FMC filters and spin initialisation on
// gfx950.  Semantics follow src/disorder.cu and src/spins*.cu of the
// reference; the Gaussian source is this library's Philox/Box-Muller stream
// (the reference's cuRAND streams are not reproducible).
#include "pspin_internal.h"

namespace ps {

// g2: unit normals on i<=j mirrored to the lower triangle, the diagonal
// kept (src/disorder.cu:46-55,57-80); FMC zeroes i!=j pairs with
// |w_i - w_j| > gamma (src/disorder.cu:130-143).
__global__ void k_g2_build(double* __restrict__ g2, int N, uint64_t seed, int fmc_on,
                           const double* __restrict__ omega, double gamma) {
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= N * N) return;
    int i = idx / N, j = idx - i * N;
    int a = i < j ? i : j, b = i < j ? j : i;
    double v = normal_k(seed, kTagG2, (uint64_t)a * N + b);
    if (fmc_on && a != b && fabs(omega[a] - omega[b]) > gamma) v = 0.0;
    g2[idx] = v;
}

// Single-channel FMC mask of quartet q: the first channel whose frequency
// balance holds wins (src/disorder.cu:153-183).  The FP64 sums are evaluated
// left to right with no multiply, so nothing is contracted and the result is
// bit-exact against any IEEE FP64 evaluation in the same order.
__device__ __forceinline__ uint8_t fmc_mask(long long q, const double* __restrict__ omega, double gamma) {
    int i, j, k, l;
    decode_quartet(q, i, j, k, l);
    const double wi = omega[i], wj = omega[j], wk = omega[k], wl = omega[l];
    if (fabs(wi + wj - wk - wl) <= gamma) return 1u;
    if (fabs(wi + wl - wj - wk) <= gamma) return 2u;
    if (fabs(wi + wk - wj - wl) <= gamma) return 4u;
    return 0u;
}

// g4: one unit normal per quartet (src/disorder.cu:93-113) and the
// single-channel FMC mask, first passing channel wins, FP64 sums evaluated
// left to right (src/disorder.cu:153-189).  fmc_mode 0 marks channel 0.
__global__ void k_g4_build(double* __restrict__ g4, uint8_t* __restrict__ mask, int N, long long nq,
                           uint64_t seed, int fmc_mode, const double* __restrict__ omega, double gamma) {
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    g4[q] = normal_k(seed, kTagG4, (uint64_t)q);
    mask[q] = fmc_mode != 0 ? fmc_mask(q, omega, gamma) : (uint8_t)1u;
}

// Rescale by sigma, then shift non-zero entries by mean
// (src/disorder.cu:28-35,82-124,199-216).  The product is rounded on its own
// (no fused multiply-add), as in the reference's two separate kernels.
__global__ void k_scale_shift(double* __restrict__ g, long long n, double sigma, double mean) {
#pragma clang fp contract(off)
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v = g[i] * sigma;
    if (v != 0.0) v += mean;
    g[i] = v;
}

__global__ void k_to_f32(const double* __restrict__ s, float* __restrict__ d, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = (float)s[i];
}

// The block's threads fill one configuration: 2N normals of the replica's
// stream, then a fixed-order LDS tree sum of |a|^2 (or |a|^4) and the
// rescale to sum = N (src/spins.cu:27-113, src/spins_sparse.cu:28-80).
__device__ __forceinline__ void init_spins_block(double* __restrict__ s, int N, uint64_t seed, int cube) {
    __shared__ double red[256];
    double acc = 0.0;
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        double re = normal_k(seed, kTagSpins, 2 * (uint64_t)i);
        double im = normal_k(seed, kTagSpins, 2 * (uint64_t)i + 1);
        s[2 * i] = re;
        s[2 * i + 1] = im;
        double r2 = re * re + im * im;
        acc += cube ? r2 * r2 : r2;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    double tot = red[0];
    double scale = cube ? sqrt(sqrt((double)N / tot)) : sqrt((double)N / tot);
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        s[2 * i] *= scale;
        s[2 * i + 1] *= scale;
    }
}

// Spins of every replica, one 256-thread block each, sum|a|^2 = N (sphere)
// or sum|a|^4 = N (cube).
__global__ __launch_bounds__(256) void k_init_spins(double* __restrict__ spins, int N,
                                                    const uint64_t* __restrict__ seeds, int cube) {
    init_spins_block(spins + (size_t)2 * N * blockIdx.x, N, seeds[blockIdx.x], cube);
}

static inline unsigned grid_for(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

hipError_t launch_g2_build(double* g2, int N, uint64_t seed, int fmc_on, const double* d_omega, double gamma,
                           hipStream_t s) {
    hipLaunchKernelGGL(k_g2_build, dim3(grid_for((long long)N * N, 256)), dim3(256), 0, s, g2, N, seed, fmc_on,
                       d_omega, gamma);
    return hipGetLastError();
}

hipError_t launch_g4_build(double* g4, uint8_t* mask, int N, uint64_t seed, int fmc_mode, const double* d_omega,
                           double gamma, hipStream_t s) {
    long long nq = n_quartets(N);
    hipLaunchKernelGGL(k_g4_build, dim3(grid_for(nq, 256)), dim3(256), 0, s, g4, mask, N, nq, seed, fmc_mode,
                       d_omega, gamma);
    return hipGetLastError();
}

hipError_t launch_scale_shift(double* g, long long n, double sigma, double mean, hipStream_t s) {
    hipLaunchKernelGGL(k_scale_shift, dim3(grid_for(n, 256)), dim3(256), 0, s, g, n, sigma, mean);
    return hipGetLastError();
}

hipError_t launch_to_f32(const double* src, float* dst, long long n, hipStream_t s) {
    hipLaunchKernelGGL(k_to_f32, dim3(grid_for(n, 256)), dim3(256), 0, s, src, dst, n);
    return hipGetLastError();
}

hipError_t launch_init_spins(double* spins, int N, int nrep, const uint64_t* seeds, int cube, hipStream_t s) {
    hipLaunchKernelGGL(k_init_spins, dim3(nrep), dim3(256), 0, s, spins, N, seeds, cube);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// The reference's free functions of include/disorder.h, include/spins.h and
// include/spins_sparse.h on caller-owned device buffers: couplings and spins
// are interleaved FP64 complex (the byte layout of cuDoubleComplex), masks one
// byte per colex quartet, omega FP64.  The steps and their order are those of
// the state's own disorder set-up above, so generate, FMC, rescale and shift
// on these buffers reproduce pspin_create's couplings bit for bit.
// ---------------------------------------------------------------------------

// g2 as the reference generates it (src/disorder.cu:57-80): unit normals,
// imaginary part zero, upper triangle mirrored, diagonal kept.
__global__ void k_gen_g2_c(double2* __restrict__ g2, int N, uint64_t seed) {
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= N * N) return;
    int i = idx / N, j = idx - i * N;
    int a = i < j ? i : j, b = i < j ? j : i;
    g2[idx] = make_double2(normal_k(seed, kTagG2, (uint64_t)a * N + b), 0.0);
}

// g4: one real unit normal per quartet (src/disorder.cu:93-113).
__global__ void k_gen_g4_c(double2* __restrict__ g4, long long nq, uint64_t seed) {
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nq) g4[q] = make_double2(normal_k(seed, kTagG4, (uint64_t)q), 0.0);
}

// Both parts times sigma (src/disorder.cu:28-35).
__global__ void k_scale_c(double2* __restrict__ g, long long n, double sigma) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) g[i] = make_double2(g[i].x * sigma, g[i].y * sigma);
}

// Real part plus mean where either part is non-zero (src/disorder.cu:199-216).
__global__ void k_shift_c(double2* __restrict__ g, long long n, double mean) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double2 v = g[i];
    if (v.x != 0.0 || v.y != 0.0) g[i] = make_double2(v.x + mean, v.y);
}

// Zero (i,j) and (j,i) for i<j with |w_i - w_j| > gamma (src/disorder.cu:130-143).
__global__ void k_fmc_g2_c(double2* __restrict__ g2, int N, const double* __restrict__ omega, double gamma) {
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= N * N) return;
    int i = idx / N, j = idx - i * N;
    if (i < j && fabs(omega[i] - omega[j]) > gamma) {
        g2[idx] = make_double2(0.0, 0.0);
        g2[(size_t)j * N + i] = make_double2(0.0, 0.0);
    }
}

__global__ void k_mask_fill(uint8_t* __restrict__ mask, long long nq) {
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nq) mask[q] = 1u;
}

__global__ void k_mask_fmc(uint8_t* __restrict__ mask, long long nq, const double* __restrict__ omega, double gamma) {
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nq) mask[q] = fmc_mask(q, omega, gamma);
}

// Set bits of the mask: a grid-stride count per thread (8 bytes at a time
// where the buffer allows), a wave sum by shuffles, one atomic per wave.
// Integer sums, so the order does not matter.
__global__ void k_mask_popcount(const uint8_t* __restrict__ mask, long long nq, unsigned long long* __restrict__ out) {
    const long long n8 = nq >> 3;
    const long long stride = (long long)gridDim.x * blockDim.x;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long c = 0;
    for (long long w = t; w < n8; w += stride) c += __popcll(((const unsigned long long*)mask)[w]);
    for (long long q = (n8 << 3) + t; q < nq; q += stride) c += __popc(mask[q]);
    for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(out, c);
}

// Colex decode of caller-given indices (the round trip of src/test_quartet_index.cu:56-74).
__global__ void k_decode(const long long* __restrict__ q, int4* __restrict__ out, long long n) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int i, j, k, l;
    decode_quartet(q[t], i, j, k, l);
    out[t] = make_int4(i, j, k, l);
}

// One configuration of N amplitudes.
__global__ __launch_bounds__(256) void k_init_spins_one(double* __restrict__ spins, int N, uint64_t seed, int cube) {
    init_spins_block(spins, N, seed, cube);
}

}  // namespace ps

namespace {

#define FF_REQUIRE(cond, code, msg) \
    do {                            \
        if (!(cond)) {              \
            ps::set_error(msg);     \
            return code;            \
        }                           \
    } while (0)

// The free functions launch on the null stream and wait, as the reference's
// end with cudaDeviceSynchronize: results are visible on return.
int ff_finish(const char* what) {
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(0);
    if (e != hipSuccess) {
        ps::set_error(std::string("HIP error ") + hipGetErrorString(e) + " in " + what);
        return PSPIN_ERR_HIP;
    }
    return PSPIN_OK;
}

inline unsigned blocks(long long n) { return (unsigned)((n + 255) / 256); }

// Quartet buffers hold C(N,4) entries; site indices fit 8 bits.
bool quartet_n_ok(int N) { return N >= 4 && N <= ps::kMaxN; }

}  // namespace

extern "C" {

int pspin_generate_g2(double* d_g2, int N, double J, unsigned long long seed) {
    (void)J;  // unit variance, as the reference; rescale_g2 applies J
    FF_REQUIRE(d_g2 && N >= 2 && N <= ps::kMaxN, PSPIN_ERR_ARG, "pspin_generate_g2: bad argument");
    hipLaunchKernelGGL(ps::k_gen_g2_c, dim3(blocks((long long)N * N)), dim3(256), 0, 0, (double2*)d_g2, N,
                       (uint64_t)seed);
    return ff_finish("pspin_generate_g2");
}

int pspin_generate_g4(double* d_g4, int N, double J, unsigned long long seed) {
    (void)J;
    FF_REQUIRE(d_g4 && quartet_n_ok(N), PSPIN_ERR_ARG, "pspin_generate_g4: bad argument");
    const long long nq = ps::n_quartets(N);
    // the reference offsets the generator seed by 12345 (src/disorder.cu:98)
    hipLaunchKernelGGL(ps::k_gen_g4_c, dim3(blocks(nq)), dim3(256), 0, 0, (double2*)d_g4, nq,
                       (uint64_t)seed + 12345u);
    return ff_finish("pspin_generate_g4");
}

int pspin_rescale_g2(double* d_g2, int N, double J, long long n_surviving) {
    FF_REQUIRE(d_g2 && N >= 1 && N <= ps::kMaxN, PSPIN_ERR_ARG, "pspin_rescale_g2: bad argument");
    if (n_surviving <= 0) return PSPIN_OK;  // everything filtered: nothing to rescale
    const double sigma = J * sqrt((double)N / (double)n_surviving);
    hipLaunchKernelGGL(ps::k_scale_c, dim3(blocks((long long)N * N)), dim3(256), 0, 0, (double2*)d_g2,
                       (long long)N * N, sigma);
    return ff_finish("pspin_rescale_g2");
}

int pspin_rescale_g4(double* d_g4, int N, double J, long long n_surviving) {
    FF_REQUIRE(d_g4 && quartet_n_ok(N), PSPIN_ERR_ARG, "pspin_rescale_g4: bad argument");
    if (n_surviving <= 0) return PSPIN_OK;
    const double sigma = J * sqrt((double)N / (double)n_surviving);
    const long long nq = ps::n_quartets(N);
    hipLaunchKernelGGL(ps::k_scale_c, dim3(blocks(nq)), dim3(256), 0, 0, (double2*)d_g4, nq, sigma);
    return ff_finish("pspin_rescale_g4");
}

int pspin_shift_mean_couplings(double* d_c, long long n, double mean) {
    if (n <= 0 || mean == 0.0) return PSPIN_OK;
    FF_REQUIRE(d_c, PSPIN_ERR_ARG, "pspin_shift_mean_couplings: null buffer");
    hipLaunchKernelGGL(ps::k_shift_c, dim3(blocks(n)), dim3(256), 0, 0, (double2*)d_c, n, mean);
    return ff_finish("pspin_shift_mean_couplings");
}

int pspin_apply_fmc_g2(double* d_g2, int N, const double* d_omega, double gamma) {
    FF_REQUIRE(d_g2 && d_omega && N >= 1 && N <= ps::kMaxN, PSPIN_ERR_ARG, "pspin_apply_fmc_g2: bad argument");
    hipLaunchKernelGGL(ps::k_fmc_g2_c, dim3(blocks((long long)N * N)), dim3(256), 0, 0, (double2*)d_g2, N, d_omega,
                       gamma);
    return ff_finish("pspin_apply_fmc_g2");
}

int pspin_init_g4_mask(uint8_t* d_mask, int N) {
    FF_REQUIRE(d_mask && quartet_n_ok(N), PSPIN_ERR_ARG, "pspin_init_g4_mask: bad argument");
    const long long nq = ps::n_quartets(N);
    hipLaunchKernelGGL(ps::k_mask_fill, dim3(blocks(nq)), dim3(256), 0, 0, d_mask, nq);
    return ff_finish("pspin_init_g4_mask");
}

int pspin_apply_fmc_g4(uint8_t* d_mask, int N, const double* d_omega, double gamma) {
    FF_REQUIRE(d_mask && d_omega && quartet_n_ok(N), PSPIN_ERR_ARG, "pspin_apply_fmc_g4: bad argument");
    const long long nq = ps::n_quartets(N);
    hipLaunchKernelGGL(ps::k_mask_fmc, dim3(blocks(nq)), dim3(256), 0, 0, d_mask, nq, d_omega, gamma);
    return ff_finish("pspin_apply_fmc_g4");
}

int pspin_count_active_g4(const uint8_t* d_mask, int N, long long* n_active) {
    FF_REQUIRE(d_mask && n_active && quartet_n_ok(N), PSPIN_ERR_ARG, "pspin_count_active_g4: bad argument");
    const long long nq = ps::n_quartets(N);
    unsigned long long* d_cnt = nullptr;
    hipError_t e = hipMalloc((void**)&d_cnt, sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemsetAsync(d_cnt, 0, sizeof(unsigned long long), 0);
    if (e == hipSuccess) {
        const unsigned g = std::max(1u, std::min(2048u, blocks((nq + 7) / 8)));
        hipLaunchKernelGGL(ps::k_mask_popcount, dim3(g), dim3(256), 0, 0, d_mask, nq, d_cnt);
        e = hipGetLastError();
    }
    unsigned long long h = 0;
    if (e == hipSuccess) e = hipMemcpy(&h, d_cnt, sizeof(h), hipMemcpyDeviceToHost);
    if (d_cnt) (void)hipFree(d_cnt);
    if (e != hipSuccess) {
        ps::set_error(std::string("HIP error ") + hipGetErrorString(e) + " in pspin_count_active_g4");
        return PSPIN_ERR_HIP;
    }
    *n_active = (long long)h;
    return PSPIN_OK;
}

int pspin_decode_quartets(const long long* d_q, int* d_ijkl, long long n) {
    if (n <= 0) return PSPIN_OK;
    FF_REQUIRE(d_q && d_ijkl, PSPIN_ERR_ARG, "pspin_decode_quartets: null buffer");
    hipLaunchKernelGGL(ps::k_decode, dim3(blocks(n)), dim3(256), 0, 0, d_q, (int4*)d_ijkl, n);
    return ff_finish("pspin_decode_quartets");
}

int pspin_init_spins_uniform(double* d_spins, int N, unsigned long long seed) {
    FF_REQUIRE(d_spins && N >= 1, PSPIN_ERR_ARG, "pspin_init_spins_uniform: bad argument");
    hipLaunchKernelGGL(ps::k_init_spins_one, dim3(1), dim3(256), 0, 0, d_spins, N, (uint64_t)seed, 0);
    return ff_finish("pspin_init_spins_uniform");
}

int pspin_init_spins_cube(double* d_spins, int N, unsigned long long seed) {
    FF_REQUIRE(d_spins && N >= 1, PSPIN_ERR_ARG, "pspin_init_spins_cube: bad argument");
    hipLaunchKernelGGL(ps::k_init_spins_one, dim3(1), dim3(256), 0, 0, d_spins, N, (uint64_t)seed, 1);
    return ff_finish("pspin_init_spins_cube");
}

}  // extern "C"
