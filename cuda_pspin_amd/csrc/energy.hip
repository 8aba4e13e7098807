// energy.hip -- FP64 full and split energies on gfx950 (the reference's
// init_energy / split_energy kernels, src/mc.cu:323-405,567-683, and their
// sparse counterparts, src/mc_sparse.cu:258-315,663-756).
//
// H2 uses the i<j convention of the sweep and the CPU gold oracle
// (src/mc.cu:344-353, src/test_delta_e.cu:38-43).  The reductions run in a
// fixed order, so results are deterministic for identical inputs (the
// reference's atomicAdd is not).
//
// Dense layout: the quartets of a (k,l) pair form the contiguous colex run
// q = C(l,4) + C(k,3) + r, r = C(j,2) + i over i < j < k.  A workgroup takes
// kEnRep replicas (amplitudes in LDS) and a chunk of pairs (the chunks hold
// equal quartet counts); each wave takes every 4th pair of the chunk and its
// lanes take every 64th quartet of the run, so the coupling and mask reads
// are coalesced and each is used for kEnRep replicas.  The products of a_k
// and a_l are formed once per run and replica:
//   ch0  a_i a_j a_k* a_l* = a_i [a_j (a_k a_l)*]
//   ch1  a_i a_j* a_k* a_l = a_i [a_j* (a_k* a_l)]
//   ch2  a_i a_j* a_k a_l* = a_i [a_j* (a_k a_l*)]
// Each workgroup writes one partial sum per replica; a second kernel adds the
// H2 part and the partials in chunk order.
#include "pspin_internal.h"

#include <algorithm>
#include <vector>

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
constexpr int kEnRep = 4;  // replicas per workgroup of the dense H4 kernel

// H4 partial sums of kEnRep replicas over one chunk of (k,l) pairs.
__global__ __launch_bounds__(kEnergyBS) void k_energy_dense_part(const double* __restrict__ spins, int nrep,
                                                                 const double* __restrict__ g4,
                                                                 const uint8_t* __restrict__ mask, int N,
                                                                 const int* __restrict__ chunks,
                                                                 double* __restrict__ part) {
    __shared__ c64 s[kEnRep][kMaxN];
    __shared__ double red[kEnergyBS / 64][kEnRep];
    const int r0 = blockIdx.y * kEnRep, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int e = threadIdx.x; e < kEnRep * N; e += blockDim.x) {
        const int q = e / N, k = e - q * N;
        c64 v = {0.0, 0.0};
        if (r0 + q < nrep) v = {spins[((size_t)(r0 + q) * N + k) * 2], spins[((size_t)(r0 + q) * N + k) * 2 + 1]};
        s[q][k] = v;
    }
    __syncthreads();
    double en[kEnRep];
#pragma unroll
    for (int q = 0; q < kEnRep; q++) en[q] = 0.0;
    const int p1 = chunks[blockIdx.x + 1];
    for (int p = chunks[blockIdx.x] + w; p < p1; p += kEnergyBS / 64) {
        // pair (k < l) in colex order: p = C(l,2) + k (wave-uniform)
        int l = (int)(0.5 + sqrt(0.25 + 2.0 * (double)p));
        while (l * (l - 1) / 2 > p) l--;
        while ((l + 1) * l / 2 <= p) l++;
        const int k = p - l * (l - 1) / 2;
        const int n = k * (k - 1) / 2;  // quartets i < j < k of the run
        if (n == 0) continue;
        const long long base = (long long)l * (l - 1) * (l - 2) * (l - 3) / 24 + (long long)k * (k - 1) * (k - 2) / 6;
        c64 qkl[kEnRep], rkl[kEnRep], skl[kEnRep];
#pragma unroll
        for (int q = 0; q < kEnRep; q++) {
            const c64 ak = s[q][k], al = s[q][l];
            qkl[q] = cc(cm(ak, al));
            rkl[q] = cm(cc(ak), al);
            skl[q] = cm(ak, cc(al));
        }
        for (int r = lane; r < n; r += 64) {
            int j = (int)((1.0f + sqrtf(1.0f + 8.0f * (float)r)) * 0.5f);
            if (j * (j - 1) / 2 > r) j--;
            else if ((j + 1) * j / 2 <= r) j++;
            const int i = r - j * (j - 1) / 2;
            const uint8_t m = mask[base + r];
            const double g = g4[base + r];
            if (m == 0 || g == 0.0) continue;
            const int ch = (m & 1) ? 0 : ((m & 2) ? 1 : 2);
#pragma unroll
            for (int q = 0; q < kEnRep; q++) {
                const c64 aj = s[q][j], ai = s[q][i];
                const c64 u = ch == 0 ? cm(aj, qkl[q]) : cm(cc(aj), ch == 1 ? rkl[q] : skl[q]);
                en[q] -= g * (ai.x * u.x - ai.y * u.y);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < kEnRep; q++) {
        double v = en[q];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if (lane == 0) red[w][q] = v;
    }
    __syncthreads();
    if (threadIdx.x < kEnRep && r0 + (int)threadIdx.x < nrep) {
        double v = 0.0;
        for (int x = 0; x < kEnergyBS / 64; x++) v += red[x][threadIdx.x];
        part[(size_t)blockIdx.x * nrep + r0 + threadIdx.x] = v;
    }
}

// Per replica: the H2 part and the H4 partials in chunk order.
__global__ __launch_bounds__(kEnergyBS) void k_energy_dense_fin(const double* __restrict__ spins,
                                                                const double* __restrict__ g2, int h2_active, int N,
                                                                const double* __restrict__ part, int nchunk, int nrep,
                                                                double* __restrict__ e2, double* __restrict__ e4) {
    __shared__ c64 s[kMaxN];
    __shared__ double red[kEnergyBS];
    const int rep = blockIdx.x;
    const double* sp = spins + (size_t)2 * N * rep;
    for (int i = threadIdx.x; i < N; i += blockDim.x) s[i] = {sp[2 * i], sp[2 * i + 1]};
    __syncthreads();
    double eh2 = h2_active ? h2_part(s, g2, N) : 0.0;
    eh2 = block_sum<kEnergyBS>(eh2, red);
    if (threadIdx.x == 0) {
        double eh4 = 0.0;
        for (int c = 0; c < nchunk; c++) eh4 += part[(size_t)c * nrep + rep];
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

int energy_chunk_count(int N, int nrep) {
    const int groups = (nrep + kEnRep - 1) / kEnRep, npairs = N * (N - 1) / 2;
    return std::max(1, std::min(npairs, std::min(256, (2048 + groups - 1) / groups)));
}

std::vector<int> energy_chunks(int N, int nchunk) {
    // pair ranges [c[x], c[x+1]) with about equal quartet counts C(k,2)
    const int npairs = N * (N - 1) / 2;
    long long total = 0;
    for (int l = 1; l < N; l++)
        for (int k = 0; k < l; k++) total += (long long)k * (k - 1) / 2;
    std::vector<int> c(nchunk + 1, npairs);
    c[0] = 0;
    long long acc = 0;
    int x = 1, p = 0;
    for (int l = 1; l < N; l++)
        for (int k = 0; k < l; k++, p++) {
            while (x < nchunk && acc >= total * x / nchunk) c[x++] = p;
            acc += (long long)k * (k - 1) / 2;
        }
    return c;
}

hipError_t launch_energy_dense(const pspin_state* st, double* e2, double* e4) {
    const int groups = (st->nrep + kEnRep - 1) / kEnRep;
    hipLaunchKernelGGL(k_energy_dense_part, dim3(st->n_echunk, groups), dim3(kEnergyBS), 0, st->stream, st->d_spins,
                       st->nrep, st->d_g4_64, st->d_mask, st->N, st->d_echunk, st->d_epart);
    hipLaunchKernelGGL(k_energy_dense_fin, dim3(st->nrep), dim3(kEnergyBS), 0, st->stream, st->d_spins, st->d_g2_64,
                       st->h2_active, st->N, st->d_epart, st->n_echunk, st->nrep, e2, e4);
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
