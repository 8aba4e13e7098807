// mlp.hip -- BF16 MFMA GEMMs and the small kernels of MLP training on
// gfx950 (CDNA4).
//
// One GEMM kernel serves the three products of a dense layer:
//   forward        Z  = X  W^T      A = X  [M][K]    (k-contiguous)   B = W [N][K] (k-contiguous)
//   data gradient  dX = dZ W        A = dZ [M][Nout] (k-contiguous)   B = W        (n-contiguous)
//   weight grad.   dW = dZ^T X      A = dZ           (m-contiguous)   B = X        (n-contiguous)
// A 256-thread workgroup computes a 64x64 tile with four waves, each a
// 32x32 block accumulated by v_mfma_f32_32x32x16_bf16 over K steps of 64;
// operand tiles are staged in LDS as [row][k] with a padded pitch so each
// lane's 8-element fragment is one 16-byte LDS read.  Epilogues fuse the
// bias, the SiLU activation (storing the FP32 pre-activation for the
// backward pass and the BF16 activation for the next layer) and the SiLU
// derivative of the data gradient; weight-gradient products (K = batch
// rows) split K over gridDim.z and also produce the bias gradient.
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
constexpr int kTK = 64;
constexpr int kPitch = kTK + 8;  // 144-byte rows: 16-byte aligned, staggered banks

// Global -> registers: a 64 x 64 tile (rows r0.., k0..) of P, 16 elements
// per thread.  KC: P contiguous along k (s = row stride); else contiguous
// along rows (s = k stride).
template <bool KC>
__device__ __forceinline__ void fetch(const uint16_t* __restrict__ P, long long s, int r0, int k0, uint4& v0,
                                      uint4& v1) {
    const int t = threadIdx.x;
    const uint16_t* q = KC ? P + (long long)(r0 + (t >> 2)) * s + (k0 + (t & 3) * 16)
                           : P + (long long)(k0 + (t >> 2)) * s + (r0 + (t & 3) * 16);
    v0 = *(const uint4*)q;
    v1 = *(const uint4*)(q + 8);
}

__device__ __forceinline__ uint16_t half_of(uint4 v, int i) {
    const uint32_t w = i < 2 ? v.x : (i < 4 ? v.y : (i < 6 ? v.z : v.w));
    return (uint16_t)((i & 1) ? (w >> 16) : (w & 0xFFFFu));
}

// Registers -> LDS tile L[row][k].
template <bool KC>
__device__ __forceinline__ void put(uint4 v0, uint4 v1, uint16_t (*L)[kPitch]) {
    const int t = threadIdx.x;
    if (KC) {
        const int r = t >> 2, c = (t & 3) * 16;
        *(uint4*)&L[r][c] = v0;
        *(uint4*)&L[r][c + 8] = v1;
    } else {
        const int k = t >> 2, r = (t & 3) * 16;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            L[r + i][k] = half_of(v0, i);
            L[r + 8 + i][k] = half_of(v1, i);
        }
    }
}

// Epilogue of a 32 x 32 accumulator block at rows mb.., column n of the lane:
// accumulator element r of lane l is row 8(r/4) + 4(l/32) + r%4, column l%32.
template <int EPI>
__device__ __forceinline__ void epilogue(const GemmArgs& a, const f32x16& acc, float* Cf, int mb, int n, int lane) {
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int m = mb + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const float v = acc[r];
        if (EPI == kEpiF32) {
            Cf[(long long)m * a.ldc + n] = v;
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

// Block tile 64 x 64, four waves of 32 x 32, K steps of 64 with the next
// step's operands fetched into registers while the current step's MFMAs run
// (double-buffered LDS, one barrier per step).  gridDim.z > 1 splits K:
// slice z covers [z K/Z, (z+1) K/Z) and writes FP32 partials (kEpiF32 only).
// When rowsum is set, workgroups of the first column block also sum the
// rows of A over their K slice (the bias gradient of a weight-gradient
// product, whose A is the layer's output gradient transposed).
template <bool AKC, bool BKC, int EPI>
__global__ __launch_bounds__(256) void k_gemm(GemmArgs a) {
    __shared__ __attribute__((aligned(16))) uint16_t As[2][kTile][kPitch];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[2][kTile][kPitch];
    const int m0 = blockIdx.y * kTile, n0 = blockIdx.x * kTile;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
    const int fr = lane & 31, fk = (lane >> 5) * 8;
    const long long sa = AKC ? a.a_m : a.a_k;
    const long long sb = BKC ? a.b_n : a.b_k;
    const int kslice = a.K / (int)gridDim.z;
    const int kb = blockIdx.z * kslice, ke = kb + kslice;
    const bool rows = a.rowsum && blockIdx.x == 0 && threadIdx.x < kTile;
    float rs = 0.0f;
    f32x16 acc = {};
    uint4 ra0, ra1, rb0, rb1;
    fetch<AKC>(a.A, sa, m0, kb, ra0, ra1);
    fetch<BKC>(a.B, sb, n0, kb, rb0, rb1);
    put<AKC>(ra0, ra1, As[0]);
    put<BKC>(rb0, rb1, Bs[0]);
    __syncthreads();
    int cur = 0;
    for (int k0 = kb; k0 < ke; k0 += kTK) {
        const bool more = k0 + kTK < ke;
        if (more) {
            fetch<AKC>(a.A, sa, m0, k0 + kTK, ra0, ra1);
            fetch<BKC>(a.B, sb, n0, k0 + kTK, rb0, rb1);
        }
#pragma unroll
        for (int kk = 0; kk < kTK; kk += 16) {
            const bf16x8 fa = *(const bf16x8*)&As[cur][wm + fr][kk + fk];
            const bf16x8 fb = *(const bf16x8*)&Bs[cur][wn + fr][kk + fk];
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa, fb, acc, 0, 0, 0);
        }
        if (rows) {
#pragma unroll 8
            for (int kk = 0; kk < kTK; kk++) rs += __uint_as_float((uint32_t)As[cur][threadIdx.x][kk] << 16);
        }
        if (more) {
            put<AKC>(ra0, ra1, As[cur ^ 1]);
            put<BKC>(rb0, rb1, Bs[cur ^ 1]);
        }
        __syncthreads();
        cur ^= 1;
    }
    if (rows) a.rowsum[(long long)blockIdx.z * (a.rz ? a.rz : a.M) + m0 + threadIdx.x] = rs;
    epilogue<EPI>(a, acc, a.Cf + (long long)blockIdx.z * (a.cz ? a.cz : (long long)a.M * a.ldc), m0 + wm,
                  n0 + wn + fr, lane);
}

// Forward product of a few rows (the learned move's M = R rows): one wave
// per 32 x 32 output tile, so M = 256 rows and N = 512 give 128 workgroups
// instead of 32.  Both operands are k-contiguous (X [M][K], W [N][K]) and are
// loaded straight from global memory into the MFMA's operand layout (lane l
// takes row l % 32 and the 8 elements at k + 8 (l / 32) of each 16-wide k
// step): no LDS, no barrier.  K = 64 KS is fixed at compile time, so the
// loads of up to four 64-wide k blocks are in flight at once (all of them
// for K <= 256) and the wave waits for memory about once, not once a block.
template <int EPI, int KS>
__global__ __launch_bounds__(64) void k_gemm_rows(GemmArgs a) {
    constexpr int D = KS < 4 ? KS : 4;  // k blocks in flight
    const int lane = threadIdx.x;
    const int m0 = blockIdx.y * 32, n0 = blockIdx.x * 32;
    const int fr = lane & 31, fk = (lane >> 5) * 8;
    const uint16_t* Ap = a.A + (long long)(m0 + fr) * a.a_m + fk;
    const uint16_t* Bp = a.B + (long long)(n0 + fr) * a.b_n + fk;
    bf16x8 fa[D][4], fb[D][4];
#pragma unroll
    for (int p = 0; p < D; p++)
#pragma unroll
        for (int u = 0; u < 4; u++) {
            fa[p][u] = *(const bf16x8*)(Ap + 64 * p + 16 * u);
            fb[p][u] = *(const bf16x8*)(Bp + 64 * p + 16 * u);
        }
    // keep the loads above issued before any MFMA (the scheduler would
    // otherwise interleave them with the MFMAs, a few loads in flight)
    __builtin_amdgcn_sched_barrier(0);
    f32x16 acc = {};
#pragma unroll
    for (int kb = 0; kb < KS; kb++) {
        const int p = kb % D;
#pragma unroll
        for (int u = 0; u < 4; u++) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(fa[p][u], fb[p][u], acc, 0, 0, 0);
        if (kb + D < KS) {
#pragma unroll
            for (int u = 0; u < 4; u++) {
                fa[p][u] = *(const bf16x8*)(Ap + 64 * (kb + D) + 16 * u);
                fb[p][u] = *(const bf16x8*)(Bp + 64 * (kb + D) + 16 * u);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    epilogue<EPI>(a, acc, a.Cf, m0, n0 + fr, lane);
}

template <int EPI>
static hipError_t launch_rows(const GemmArgs& g, hipStream_t s) {
    dim3 grid(g.N / 32, g.M / 32), block(64);
    switch (g.K / 64) {
    case 1: hipLaunchKernelGGL((k_gemm_rows<EPI, 1>), grid, block, 0, s, g); break;
    case 2: hipLaunchKernelGGL((k_gemm_rows<EPI, 2>), grid, block, 0, s, g); break;
    case 3: hipLaunchKernelGGL((k_gemm_rows<EPI, 3>), grid, block, 0, s, g); break;
    case 4: hipLaunchKernelGGL((k_gemm_rows<EPI, 4>), grid, block, 0, s, g); break;
    case 5: hipLaunchKernelGGL((k_gemm_rows<EPI, 5>), grid, block, 0, s, g); break;
    case 6: hipLaunchKernelGGL((k_gemm_rows<EPI, 6>), grid, block, 0, s, g); break;
    case 7: hipLaunchKernelGGL((k_gemm_rows<EPI, 7>), grid, block, 0, s, g); break;
    case 8: hipLaunchKernelGGL((k_gemm_rows<EPI, 8>), grid, block, 0, s, g); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <bool AKC, bool BKC>
static hipError_t launch_epi(const GemmArgs& g, int epi, int splits, hipStream_t s) {
    dim3 grid(g.N / kTile, g.M / kTile, splits), block(256);
    switch (epi) {
    case kEpiF32: hipLaunchKernelGGL((k_gemm<AKC, BKC, kEpiF32>), grid, block, 0, s, g); break;
    case kEpiBiasF32: hipLaunchKernelGGL((k_gemm<AKC, BKC, kEpiBiasF32>), grid, block, 0, s, g); break;
    case kEpiBiasSilu: hipLaunchKernelGGL((k_gemm<AKC, BKC, kEpiBiasSilu>), grid, block, 0, s, g); break;
    case kEpiDSilu: hipLaunchKernelGGL((k_gemm<AKC, BKC, kEpiDSilu>), grid, block, 0, s, g); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Rows at or below which a forward product takes k_gemm_rows, and its
// largest K.
constexpr int kRowsKernelMaxM = 512;
constexpr int kRowsKernelMaxK = 512;

hipError_t launch_gemm(const GemmArgs& g, int epi, int splits, hipStream_t s) {
    const bool akc = g.a_k == 1, bkc = g.b_k == 1;
    if (akc && bkc && splits == 1 && g.M <= kRowsKernelMaxM && (epi == kEpiBiasSilu || epi == kEpiBiasF32) &&
        g.M > 0 && g.M % 32 == 0 && g.N > 0 && g.N % 32 == 0 && g.K > 0 && g.K % 64 == 0 &&
        g.K <= kRowsKernelMaxK && g.a_m % 8 == 0 && g.b_n % 8 == 0)
        return epi == kEpiBiasSilu ? launch_rows<kEpiBiasSilu>(g, s) : launch_rows<kEpiBiasF32>(g, s);
    if ((akc ? g.a_m : g.a_k) % 8 || (bkc ? g.b_n : g.b_k) % 8) return hipErrorInvalidValue;
    if (!akc && g.a_m != 1) return hipErrorInvalidValue;
    if (!bkc && g.b_n != 1) return hipErrorInvalidValue;
    if (splits < 1 || (splits > 1 && epi != kEpiF32)) return hipErrorInvalidValue;
    if (g.M % kTile || g.N % kTile || g.M <= 0 || g.N <= 0 || g.K <= 0 || g.K % (kTK * splits))
        return hipErrorInvalidValue;
    if (akc && bkc) return launch_epi<true, true>(g, epi, splits, s);
    if (akc && !bkc) return launch_epi<true, false>(g, epi, splits, s);
    if (!akc && !bkc) return launch_epi<false, false>(g, epi, splits, s);
    return launch_epi<false, true>(g, epi, splits, s);
}

// out[i] = sum_z part[z * n + i], z in order (split-K reduction).
__global__ void k_sum_slices(const float* __restrict__ part, int nz, long long n, float* __restrict__ out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (int z = 0; z < nz; z++) s += part[(long long)z * n + i];
    out[i] = s;
}

hipError_t launch_sum_slices(const float* part, int nz, long long n, float* out, hipStream_t s) {
    hipLaunchKernelGGL(k_sum_slices, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, part, nz, n, out);
    return hipGetLastError();
}

// Adam (Kingma & Ba 2015) on FP32 masters; refreshes the BF16 copy.
__global__ void k_adam(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ g, int nz, uint16_t* __restrict__ pb, long long n, float lr,
                       float b1, float b2, float eps, float bc1, float bc2) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gi = 0.0f;  // split-K slices in order, as k_sum_slices adds them
    for (int z = 0; z < nz; z++) gi += g[(long long)z * n + i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    p[i] = pi;
    pb[i] = f2bf(pi);
}

hipError_t launch_adam(float* p, float* m, float* v, const float* g, int nz, uint16_t* pb, long long n, float lr,
                       float b1, float b2, float eps, long long t, hipStream_t s) {
    const float bc1 = 1.0f - powf(b1, (float)t), bc2 = 1.0f - powf(b2, (float)t);
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, m, v, g, nz, pb, n, lr, b1,
                       b2, eps, bc1, bc2);
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
