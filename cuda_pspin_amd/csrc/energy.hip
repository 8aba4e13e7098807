// energy.hip -- FP64 full and split energies on gfx950 (the reference's
// init_energy / split_energy kernels, src/mc.cu:323-405,567-683, and their
// sparse counterparts, src/mc_sparse.cu:258-315,663-756).
//
// H2 uses the i<j convention of the sweep and the CPU gold oracle
// (src/mc.cu:344-353, src/test_delta_e.cu:38-43).  The reductions run in a
// fixed order, so results are deterministic for identical inputs (the
// reference's atomicAdd is not).
//
// Dense layout: one block per replica; thread t walks the (k,l) pairs
// t, t+B, ... and, for each, the contiguous colex run q = C(l,4)+C(k,3)+C(j,2)+i
// over j<k, i<j -- so the coupling and mask reads are unit-stride and the
// quartet indices need no decode.  Spins of the replica sit in LDS.
#include "pspin_internal.h"

namespace ps {

struct c64 {
    double x, y;
};
__device__ __forceinline__ c64 cm(c64 a, c64 b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ c64 cc(c64 a) { return {a.x, -a.y}; }

template <int BS>
__device__ __forceinline__ double block_sum(double v, double* red) {
    red[threadIdx.x] = v;
    __syncthreads();
#pragma unroll
    for (int w = BS / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    double r = red[0];
    __syncthreads();
    return r;
}

__device__ __forceinline__ double h2_part(const c64* s, const double* __restrict__ g2, int N) {
    double e = 0.0;
    long long np = (long long)N * (N - 1) / 2;
    for (long long p = threadIdx.x; p < np; p += blockDim.x) {
        int j = (int)(0.5 + sqrt(0.25 + 2.0 * (double)p));
        while ((long long)j * (j - 1) / 2 > p) j--;
        while ((long long)(j + 1) * j / 2 <= p) j++;
        int i = (int)(p - (long long)j * (j - 1) / 2);
        c64 pr = cm(s[i], cc(s[j]));
        e += -g2[i * N + j] * pr.x;
    }
    return e;
}

__device__ __forceinline__ double quartet_re(c64 a0, c64 a1, c64 a2, c64 a3, int ch) {
    // ch0: a0 a1 a2* a3*, ch1: a0 a1* a2* a3, ch2: a0 a1* a2 a3* (src/mc.cu:385-394)
    c64 s1 = ch == 0 ? a1 : cc(a1);
    c64 s2 = ch == 2 ? a2 : cc(a2);
    c64 s3 = ch == 1 ? a3 : cc(a3);
    return cm(cm(a0, s1), cm(s2, s3)).x;
}

constexpr int kEnergyBS = 256;

__global__ __launch_bounds__(kEnergyBS) void k_energy_dense(const double* __restrict__ spins,
                                                            const double* __restrict__ g2, int h2_active,
                                                            const double* __restrict__ g4,
                                                            const uint8_t* __restrict__ mask, int N,
                                                            double* __restrict__ e2, double* __restrict__ e4) {
    __shared__ c64 s[kMaxN];
    __shared__ double red[kEnergyBS];
    int rep = blockIdx.x;
    const double* sp = spins + (size_t)2 * N * rep;
    for (int i = threadIdx.x; i < N; i += blockDim.x) s[i] = {sp[2 * i], sp[2 * i + 1]};
    __syncthreads();
    double eh2 = h2_active ? h2_part(s, g2, N) : 0.0;
    double eh4 = 0.0;
    long long npl = (long long)N * (N - 1) / 2;
    for (long long p = threadIdx.x; p < npl; p += blockDim.x) {
        // pair (k<l) in colex order: p = C(l,2) + k
        int l = (int)(0.5 + sqrt(0.25 + 2.0 * (double)p));
        while ((long long)l * (l - 1) / 2 > p) l--;
        while ((long long)(l + 1) * l / 2 <= p) l++;
        int k = (int)(p - (long long)l * (l - 1) / 2);
        if (k < 2) continue;
        long long base = (long long)l * (l - 1) * (l - 2) * (l - 3) / 24 + (long long)k * (k - 1) * (k - 2) / 6;
        c64 ak = s[k], al = s[l];
        long long q = base;
        for (int j = 1; j < k; j++) {
            c64 aj = s[j];
            for (int i = 0; i < j; i++, q++) {
                uint8_t m = mask[q];
                double g = g4[q];
                if (m == 0 || g == 0.0) continue;
                int ch = (m & 1) ? 0 : ((m & 2) ? 1 : 2);
                eh4 += -g * quartet_re(s[i], aj, ak, al, ch);
            }
        }
    }
    eh2 = block_sum<kEnergyBS>(eh2, red);
    eh4 = block_sum<kEnergyBS>(eh4, red);
    if (threadIdx.x == 0) {
        e2[rep] = eh2;
        e4[rep] = eh4;
    }
}

__global__ __launch_bounds__(kEnergyBS) void k_energy_sparse(const double* __restrict__ spins,
                                                             const double* __restrict__ g2, int h2_active,
                                                             const int* __restrict__ sq,
                                                             const double* __restrict__ sqg, int nq, int N,
                                                             double* __restrict__ e2, double* __restrict__ e4) {
    __shared__ c64 s[kMaxN];
    __shared__ double red[kEnergyBS];
    int rep = blockIdx.x;
    const double* sp = spins + (size_t)2 * N * rep;
    for (int i = threadIdx.x; i < N; i += blockDim.x) s[i] = {sp[2 * i], sp[2 * i + 1]};
    __syncthreads();
    double eh2 = h2_active ? h2_part(s, g2, N) : 0.0;
    double eh4 = 0.0;
    for (int t = threadIdx.x; t < nq; t += blockDim.x) {
        const int* v = sq + 5 * t;
        eh4 += -sqg[t] * quartet_re(s[v[0]], s[v[1]], s[v[2]], s[v[3]], v[4]);
    }
    eh2 = block_sum<kEnergyBS>(eh2, red);
    eh4 = block_sum<kEnergyBS>(eh4, red);
    if (threadIdx.x == 0) {
        e2[rep] = eh2;
        e4[rep] = eh4;
    }
}

__global__ void k_add(const double* a, const double* b, double* out, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = a[i] + b[i];
}

hipError_t launch_energy_dense(const pspin_state* st, double* e2, double* e4) {
    hipLaunchKernelGGL(k_energy_dense, dim3(st->nrep), dim3(kEnergyBS), 0, st->stream, st->d_spins, st->d_g2_64,
                       st->h2_active, st->d_g4_64, st->d_mask, st->N, e2, e4);
    return hipGetLastError();
}

hipError_t launch_energy_sparse(const pspin_state* st, double* e2, double* e4) {
    hipLaunchKernelGGL(k_energy_sparse, dim3(st->nrep), dim3(kEnergyBS), 0, st->stream, st->d_spins, st->d_g2_64,
                       st->h2_active, st->d_sq, st->d_sqg, st->n_sq, st->N, e2, e4);
    return hipGetLastError();
}

hipError_t launch_add(const double* a, const double* b, double* out, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_add, dim3((n + 255) / 256), dim3(256), 0, s, a, b, out, n);
    return hipGetLastError();
}

}  // namespace ps
