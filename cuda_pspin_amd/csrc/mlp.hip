// mlp.hip -- BF16 MFMA GEMMs and the small kernels of MLP training on
// gfx950 (CDNA4).
//
// One GEMM kernel serves the three products of a dense layer:
//   forward        Z  = X  W^T      A = X  [M][K]    (k-contiguous)   B = W [N][K] (k-contiguous)
//   data gradient  dX = dZ W        A = dZ [M][Nout] (k-contiguous)   B = W        (n-contiguous)
//   weight grad.   dW = dZ^T X      A = dZ           (m-contiguous)   B = X        (n-contiguous)
// A 256-thread workgroup computes a 64x64 tile with four waves, each a
// 32x32 block accumulated by v_mfma_f32_32x32x16_bf16 over K steps of 32;
// operand tiles are staged in LDS as [row][k] with a padded pitch so each
// lane's 8-element fragment is one 16-byte LDS read.  Epilogues fuse the
// bias, the SiLU activation (storing the FP32 pre-activation for the
// backward pass and the BF16 activation for the next layer) and the SiLU
// derivative of the data gradient.
#include "pspin_learned.h"

namespace ps {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// FP32 -> BF16, round to nearest even (inputs are finite).
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
    float s = 1.0f / (1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

constexpr int kTile = 64;
constexpr int kTK = 32;
constexpr int kPitch = kTK + 8;  // 80-byte rows: 16-byte aligned, staggered banks

// Stage a 64 x 32 tile (rows r0.., k0..) of P into L[row][k].
// KC: P is contiguous along k (s = row stride); else contiguous along rows
// (s = k stride).  Each thread moves one 16-byte vector.
template <bool KC>
__device__ __forceinline__ void stage(const uint16_t* __restrict__ P, long long s, int r0, int k0,
                                      uint16_t (*L)[kPitch]) {
    const int t = threadIdx.x;
    if (KC) {
        const int r = t >> 2, c = (t & 3) * 8;
        *(uint4*)&L[r][c] = *(const uint4*)(P + (long long)(r0 + r) * s + (k0 + c));
    } else {
        const int k = t >> 3, r = (t & 7) * 8;
        const uint4 v = *(const uint4*)(P + (long long)(k0 + k) * s + (r0 + r));
        const uint16_t* e = (const uint16_t*)&v;
#pragma unroll
        for (int i = 0; i < 8; i++) L[r + i][k] = e[i];
    }
}

template <bool AKC, bool BKC, int EPI>
__global__ __launch_bounds__(256) void k_gemm(GemmArgs a) {
    __shared__ __attribute__((aligned(16))) uint16_t As[kTile][kPitch];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[kTile][kPitch];
    const int m0 = blockIdx.y * kTile, n0 = blockIdx.x * kTile;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
    const int fr = lane & 31, fk = (lane >> 5) * 8;
    const long long sa = AKC ? a.a_m : a.a_k;
    const long long sb = BKC ? a.b_n : a.b_k;
    f32x16 acc = {};
    for (int k0 = 0; k0 < a.K; k0 += kTK) {
        stage<AKC>(a.A, sa, m0, k0, As);
        stage<BKC>(a.B, sb, n0, k0, Bs);
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kTK; kk += 16) {
            const bf16x8 fa = *(const bf16x8*)&As[wm + fr][kk + fk];
            const bf16x8 fb = *(const bf16x8*)&Bs[wn + fr][kk + fk];
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa, fb, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    // accumulator element r of lane l: row 8(r/4) + 4(l/32) + r%4, column l%32
    const int n = n0 + wn + fr;
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int m = m0 + wm + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const float v = acc[r];
        if (EPI == kEpiF32) {
            a.Cf[(long long)m * a.ldc + n] = v;
        } else if (EPI == kEpiBiasF32) {
            a.Cf[(long long)m * a.ldc + n] = v + a.bias[n];
        } else if (EPI == kEpiBiasSilu) {
            const float z = v + a.bias[n];
            a.Cf[(long long)m * a.ldc + n] = z;
            a.Cb[(long long)m * a.ldcb + n] = f2bf(silu(z));
        } else {
            const float z = a.Z[(long long)m * a.ldz + n];
            a.Cb[(long long)m * a.ldcb + n] = f2bf(v * dsilu(z));
        }
    }
}

template <bool AKC, bool BKC>
static hipError_t launch_epi(const GemmArgs& g, int epi, hipStream_t s) {
    dim3 grid(g.N / kTile, g.M / kTile), block(256);
    switch (epi) {
    case kEpiF32: hipLaunchKernelGGL((k_gemm<AKC, BKC, kEpiF32>), grid, block, 0, s, g); break;
    case kEpiBiasF32: hipLaunchKernelGGL((k_gemm<AKC, BKC, kEpiBiasF32>), grid, block, 0, s, g); break;
    case kEpiBiasSilu: hipLaunchKernelGGL((k_gemm<AKC, BKC, kEpiBiasSilu>), grid, block, 0, s, g); break;
    case kEpiDSilu: hipLaunchKernelGGL((k_gemm<AKC, BKC, kEpiDSilu>), grid, block, 0, s, g); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_gemm(const GemmArgs& g, int epi, hipStream_t s) {
    const bool akc = g.a_k == 1, bkc = g.b_k == 1;
    if ((akc ? g.a_m : g.a_k) % 8 || (bkc ? g.b_n : g.b_k) % 8) return hipErrorInvalidValue;
    if (!akc && g.a_m != 1) return hipErrorInvalidValue;
    if (!bkc && g.b_n != 1) return hipErrorInvalidValue;
    if (g.M % kTile || g.N % kTile || g.K % kTK || g.M <= 0 || g.N <= 0 || g.K <= 0) return hipErrorInvalidValue;
    if (akc && bkc) return launch_epi<true, true>(g, epi, s);
    if (akc && !bkc) return launch_epi<true, false>(g, epi, s);
    if (!akc && !bkc) return launch_epi<false, false>(g, epi, s);
    return launch_epi<false, true>(g, epi, s);
}

// Column sums of a BF16 [M][N] matrix (bias gradients), fixed order.
__global__ __launch_bounds__(256) void k_colsum_bf16(const uint16_t* __restrict__ X, int M, long long ld,
                                                     float* __restrict__ out) {
    __shared__ float red[4][64];
    const int c = blockIdx.x * 64 + (threadIdx.x & 63), g = threadIdx.x >> 6;
    float s = 0.0f;
    for (int r = g; r < M; r += 4) s += __uint_as_float((uint32_t)X[(long long)r * ld + c] << 16);
    red[g][threadIdx.x & 63] = s;
    __syncthreads();
    if (g == 0) out[c] = (red[0][c & 63] + red[1][c & 63]) + (red[2][c & 63] + red[3][c & 63]);
}

hipError_t launch_colsum_bf16(const uint16_t* X, int M, int N, long long ld, float* out, hipStream_t s) {
    if (N % 64) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_colsum_bf16, dim3(N / 64), dim3(256), 0, s, X, M, ld, out);
    return hipGetLastError();
}

// Adam (Kingma & Ba 2015) on FP32 masters; refreshes the BF16 copy.
__global__ void k_adam(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ g, uint16_t* __restrict__ pb, long long n, float lr, float b1,
                       float b2, float eps, float bc1, float bc2) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    p[i] = pi;
    pb[i] = f2bf(pi);
}

hipError_t launch_adam(float* p, float* m, float* v, const float* g, uint16_t* pb, long long n, float lr,
                       float b1, float b2, float eps, long long t, hipStream_t s) {
    const float bc1 = 1.0f - powf(b1, (float)t), bc2 = 1.0f - powf(b2, (float)t);
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, m, v, g, pb, n, lr, b1, b2,
                       eps, bc1, bc2);
    return hipGetLastError();
}

__global__ void k_f32_to_bf16(const float* __restrict__ src, uint16_t* __restrict__ dst, long long n) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = f2bf(src[i]);
}

hipError_t launch_f32_to_bf16(const float* src, uint16_t* dst, long long n, hipStream_t s) {
    hipLaunchKernelGGL(k_f32_to_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, dst, n);
    return hipGetLastError();
}

}  // namespace ps
