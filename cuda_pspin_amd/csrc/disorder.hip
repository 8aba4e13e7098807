// disorder.hip -- coupling generation, FMC filters and spin initialisation on
// gfx950.  Semantics follow src/disorder.cu and src/spins*.cu of the
// reference; the Gaussian source is this library's Philox/Box-Muller stream
// (the reference's cuRAND streams are not reproducible).
#include "pspin_internal.h"

namespace ps {

// Colex decode q -> (i<j<k<l), q = C(l,4)+C(k,3)+C(j,2)+i (src/disorder.cu:159-171).
// Branch-light: l from a float estimate corrected by integer compares.
__device__ __forceinline__ void decode_quartet(long long q, int& i, int& j, int& k, int& l) {
    int L = (int)floor(pow(24.0 * (double)q, 0.25)) + 2;
    while ((long long)L * (L - 1) * (L - 2) * (L - 3) / 24 > q) L--;
    while ((long long)(L + 1) * L * (L - 1) * (L - 2) / 24 <= q) L++;
    long long rem = q - (long long)L * (L - 1) * (L - 2) * (L - 3) / 24;
    int K = (int)floor(cbrt(6.0 * (double)rem)) + 1;
    if (K < 2) K = 2;
    while ((long long)K * (K - 1) * (K - 2) / 6 > rem) K--;
    while ((long long)(K + 1) * K * (K - 1) / 6 <= rem) K++;
    rem -= (long long)K * (K - 1) * (K - 2) / 6;
    int J = (int)floor(sqrt(2.0 * (double)rem)) + 1;
    if (J < 1) J = 1;
    while ((long long)J * (J - 1) / 2 > rem) J--;
    while ((long long)(J + 1) * J / 2 <= rem) J++;
    rem -= (long long)J * (J - 1) / 2;
    i = (int)rem; j = J; k = K; l = L;
}

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

// g4: one unit normal per quartet (src/disorder.cu:93-113) and the
// single-channel FMC mask, first passing channel wins, FP64 sums evaluated
// left to right (src/disorder.cu:153-189).  fmc_mode 0 marks channel 0.
__global__ void k_g4_build(double* __restrict__ g4, uint8_t* __restrict__ mask, int N, long long nq,
                           uint64_t seed, int fmc_mode, const double* __restrict__ omega, double gamma) {
    long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    g4[q] = normal_k(seed, kTagG4, (uint64_t)q);
    uint8_t m = 1u;
    if (fmc_mode != 0) {
        int i, j, k, l;
        decode_quartet(q, i, j, k, l);
        double wi = omega[i], wj = omega[j], wk = omega[k], wl = omega[l];
        m = 0u;
        if (fabs(wi + wj - wk - wl) <= gamma) m = 1u;
        else if (fabs(wi + wl - wj - wk) <= gamma) m = 2u;
        else if (fabs(wi + wk - wj - wl) <= gamma) m = 4u;
    }
    mask[q] = m;
}

// Rescale by sigma, then shift non-zero entries by mean
// (src/disorder.cu:28-35,82-124,199-216).
__global__ void k_scale_shift(double* __restrict__ g, long long n, double sigma, double mean) {
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

// Spins: 2N normals per replica, then rescaled to sum|a|^2 = N
// (src/spins.cu:27-113) or sum|a|^4 = N (src/spins_sparse.cu:28-80).
// One 256-thread block per replica; fixed-order LDS tree reduction.
__global__ __launch_bounds__(256) void k_init_spins(double* __restrict__ spins, int N,
                                                    const uint64_t* __restrict__ seeds, int cube) {
    __shared__ double red[256];
    int rep = blockIdx.x;
    double* s = spins + (size_t)2 * N * rep;
    uint64_t seed = seeds[rep];
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

}  // namespace ps
