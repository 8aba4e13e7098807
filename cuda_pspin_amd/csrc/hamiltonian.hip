// hamiltonian.hip -- compute_energy (include/hamiltonian.h:11-17,
// src/hamiltonian.cu:23-200) on caller-owned device buffers, FP64.
//
// One configuration of N amplitudes, complex couplings allowed.  The
// amplitudes are staged in LDS; workgroup b walks quartets b*256 + t in
// grid-stride order, decodes each with the shared colex decoder and adds
// -Re[g_q f_ch(a)] for every channel set in the mask, then reduces its
// threads in a fixed LDS tree to one partial.  One last workgroup adds the
// two-body term (i<j) and the partials in index order, so the result does
// not depend on scheduling.
#include <mutex>

#include "pspin_internal.h"

namespace ps {

constexpr int kEnThreads = 256;
constexpr int kEnMaxParts = 1024;  // H4 workgroups at most (grid-stride beyond)

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }

// Fixed-order tree sum over the workgroup's 256 threads; thread 0 gets it.
__device__ __forceinline__ double block_sum(double v, double* red) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int w = kEnThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    return red[0];
}

// H4 partial sums (src/hamiltonian.cu:61-117): channel products
//   ch 0: a_i a_j conj(a_k) conj(a_l)
//   ch 1: a_i conj(a_j) conj(a_k) a_l
//   ch 2: a_i conj(a_j) a_k conj(a_l)
// each multiplied as g ((s0 s1)(s2 s3)).
__global__ __launch_bounds__(kEnThreads) void k_energy_h4_one(const double2* __restrict__ spins,
                                                              const double2* __restrict__ g4,
                                                              const uint8_t* __restrict__ mask, int N, long long nq,
                                                              double* __restrict__ part) {
    __shared__ double2 a[kMaxN];
    __shared__ double red[kEnThreads];
    for (int i = threadIdx.x; i < N; i += kEnThreads) a[i] = spins[i];
    __syncthreads();
    double acc = 0.0;
    for (long long q = (long long)blockIdx.x * kEnThreads + threadIdx.x; q < nq;
         q += (long long)gridDim.x * kEnThreads) {
        const uint32_t m = mask[q];
        if (!m) continue;
        int i, j, k, l;
        decode_quartet(q, i, j, k, l);
        const double2 g = g4[q], ai = a[i], aj = a[j], ak = a[k], al = a[l];
        if (m & 1u) acc -= cmul(g, cmul(cmul(ai, aj), cmul(cconj(ak), cconj(al)))).x;
        if (m & 2u) acc -= cmul(g, cmul(cmul(ai, cconj(aj)), cmul(cconj(ak), al))).x;
        if (m & 4u) acc -= cmul(g, cmul(cmul(ai, cconj(aj)), cmul(ak, cconj(al)))).x;
    }
    const double s = block_sum(acc, red);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// H2 over i<j, -Re[g_ij a_i conj(a_j)] (src/hamiltonian.cu:23-56 with the
// sweep's i<j convention), plus the H4 partials in order.  out[0] = E.
__global__ __launch_bounds__(kEnThreads) void k_energy_fin_one(const double2* __restrict__ spins,
                                                               const double2* __restrict__ g2, int N,
                                                               const double* __restrict__ part, int nparts,
                                                               double* __restrict__ out) {
    __shared__ double red[kEnThreads];
    double h2 = 0.0;
    if (g2)
        for (int idx = threadIdx.x; idx < N * N; idx += kEnThreads) {
            const int i = idx / N, j = idx - i * N;
            if (i < j) h2 -= cmul(g2[idx], cmul(spins[i], cconj(spins[j]))).x;
        }
    const double e2 = block_sum(h2, red);
    if (threadIdx.x == 0) {
        double e4 = 0.0;
        for (int b = 0; b < nparts; b++) e4 += part[b];
        out[0] = e2 + e4;
    }
}

// Workgroups of the H4 pass for N sites (one partial each).
inline int energy_parts(int N) {
    const long long nq = n_quartets(N);
    return (int)std::max(1LL, std::min<long long>(kEnMaxParts, (nq + kEnThreads - 1) / kEnThreads));
}

// Both passes on the null stream; ws holds energy_parts(N) partials and the
// result after them.
hipError_t launch_energy_one(const double* spins, const double* g2, const double* g4, const uint8_t* mask, int N,
                             double* ws) {
    const int parts = energy_parts(N);
    hipLaunchKernelGGL(k_energy_h4_one, dim3(parts), dim3(kEnThreads), 0, 0, (const double2*)spins,
                       (const double2*)g4, mask, N, n_quartets(N), ws);
    hipLaunchKernelGGL(k_energy_fin_one, dim3(1), dim3(kEnThreads), 0, 0, (const double2*)spins,
                       (const double2*)g2, N, ws, parts, ws + parts);
    return hipGetLastError();
}

// Workspace of pspin_compute_energy_noalloc: the H4 partials and the
// result, one copy per GPU for the life of the process.  Host code must not
// pass `scratch` itself to a kernel (the name is a host-side stand-in); it
// uses the device address that hipGetSymbolAddress returns for the current
// device, kept in scratch_dev.
constexpr int kScratchDoubles = kEnMaxParts + 1;
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
static std::mutex scratch_mu;
static double* scratch_dev[kMaxDevices];

// The current device's address of `scratch`, looked up once per device.
hipError_t scratch_address(double** out) {
    int d = 0;
    hipError_t e = hipGetDevice(&d);
    if (e != hipSuccess) return e;
    if (d < 0 || d >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (!scratch_dev[d]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        scratch_dev[d] = (double*)p;
    }
    *out = scratch_dev[d];
    return hipSuccess;
}

}  // namespace ps

extern "C" {

int pspin_compute_energy(const double* d_spins, const double* d_g2, const double* d_g4, const uint8_t* d_mask, int N,
                         double* energy) {
    if (!(d_spins && d_g4 && d_mask && energy && N >= 4 && N <= ps::kMaxN)) {
        ps::set_error("pspin_compute_energy: bad argument (device spins, g4, mask, a host result, 4 <= N <= 256)");
        return PSPIN_ERR_ARG;
    }
    const int parts = ps::energy_parts(N);
    double* ws = nullptr;
    hipError_t e = hipMalloc((void**)&ws, (size_t)(parts + 1) * sizeof(double));
    if (e == hipSuccess) e = ps::launch_energy_one(d_spins, d_g2, d_g4, d_mask, N, ws);
    if (e == hipSuccess) e = hipMemcpy(energy, ws + parts, sizeof(double), hipMemcpyDeviceToHost);
    if (ws) (void)hipFree(ws);
    if (e != hipSuccess) {
        ps::set_error(std::string("HIP error ") + hipGetErrorString(e) + " in pspin_compute_energy");
        return PSPIN_ERR_HIP;
    }
    return PSPIN_OK;
}

int pspin_compute_energy_noalloc(const double* d_spins, const double* d_g2, const double* d_g4, const uint8_t* d_mask,
                                 int N, double* energy) {
    if (!(d_spins && d_g4 && d_mask && energy && N >= 4)) {
        ps::set_error("pspin_compute_energy_noalloc: bad argument (device spins, g4, mask and a host result)");
        return PSPIN_ERR_ARG;
    }
    if (N > ps::kMaxN || ps::energy_parts(N) + 1 > ps::kScratchDoubles) {
        ps::set_error("pspin_compute_energy_noalloc: N = " + std::to_string(N) + " is above the maximum of " +
                      std::to_string(ps::kMaxN) + " sites");
        return PSPIN_ERR_ARG;
    }
    double* ws = nullptr;
    hipError_t e = ps::scratch_address(&ws);
    if (e == hipSuccess) e = ps::launch_energy_one(d_spins, d_g2, d_g4, d_mask, N, ws);
    if (e == hipSuccess) e = hipMemcpy(energy, ws + ps::energy_parts(N), sizeof(double), hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        ps::set_error(std::string("HIP error ") + hipGetErrorString(e) + " in pspin_compute_energy_noalloc");
        return PSPIN_ERR_HIP;
    }
    return PSPIN_OK;
}

}  // extern "C"
