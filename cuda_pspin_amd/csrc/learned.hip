// learned.hip -- the learned global move on gfx950: gauge fixing, the
// projected-Gaussian proposal with its exact FP64 density, FP64 full
// energies of current and proposed states, Metropolis-Hastings, and the
// data path of training (snapshots, pair gathering, Gaussian NLL).
//
// Target.  The pair-rotation sweep samples, on the sphere sum|a|^2 = N with
// its surface measure, pi(a) ~ exp(-beta H(a)) prod_k |a_k|^-1 (its
// intensity split is arcsine-distributed, include/spins.h:25-51, and no
// Hastings factor is applied, src/mc.cu:288-296).  The learned move targets
// the same pi, so mixing it with sweeps keeps the stationary law.
//
// Gauge.  H is invariant under a global phase.  The network sees the state
// rotated so that its largest-intensity amplitude (smallest index on ties)
// is real positive: u = a c with c = conj(a_k*) / |a_k*|.  A proposal drawn
// in that frame is rotated back with conj(c).  Rotations are isometries of
// the sphere, so q(a'|a) is the frame density evaluated at a' c.
//
// Projected Gaussian.  In the frame, z ~ N(mu, diag sigma^2) over the d = 2N
// real coordinates and y = z / |z| on the unit sphere; a' = sqrt(N) y conj(c).
// mu = u + m O[0:2N] and log sigma = clamp(O[2N:4N] + ls0) + log s, with the
// proposal scale s and the mean-offset scale m (m = s^2 keeps a learned
// Langevin-like drift consistent with the shrunken step).
// The density of y w.r.t. the surface measure of the unit sphere is
//   p(y) = (2 pi)^(-d/2) prod sigma_i^-1 exp(-C/2) A^(-d/2) J_{d-1}(B / sqrt A),
//   A = sum y_i^2/sigma_i^2, B = sum y_i mu_i/sigma_i^2, C = sum mu_i^2/sigma_i^2,
//   J_n(alpha) = int_0^inf s^n exp(-s^2/2 + alpha s) ds,
// (integrate the Gaussian along the ray r y, r > 0, with Jacobian r^(d-1)).
// log J_n is computed by the trapezoid rule on a grid centred on the
// maximum s* of the log-concave integrand with step w/4 (w the Laplace
// width), extended on each side until the integrand falls below 1e-40 of
// its maximum: for an integrand analytic around the real axis and
// negligible at the ends the trapezoid error is exponentially small.
// The common factor sqrt(N)^-(d-1) of the densities on the radius-sqrt(N)
// sphere cancels in the MH ratio.
//
// Randomness: Philox4x32-10 keyed by the replica seed, counter
// (move, block, 0, kCtrLearned): blocks 0..N-1 give the 2N normals
// (Box-Muller, two per block), block 0xFFFFFFFF the accept uniform.
#include "pspin_learned.h"

namespace ps {

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ uint16_t f2bf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// Gauge phase c = conj(a_k*) / |a_k*| of a [N][2] (whole wave, result in
// every lane).
__device__ __forceinline__ void gauge_of(const double* __restrict__ a, int N, double& cr, double& ci) {
    const int lane = threadIdx.x & 63;
    double best = -1.0;
    int bk = 0x7FFFFFFF;
    for (int k = lane; k < N; k += 64) {
        const double x = a[2 * k] * a[2 * k] + a[2 * k + 1] * a[2 * k + 1];
        if (x > best) {
            best = x;
            bk = k;
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double ob = __shfl_xor(best, o, 64);
        const int ok = __shfl_xor(bk, o, 64);
        if (ob > best || (ob == best && ok < bk)) {
            best = ob;
            bk = ok;
        }
    }
    const double r = sqrt(best);
    cr = a[2 * bk] / r;
    ci = -a[2 * bk + 1] / r;
}

// log J_n(alpha), whole wave.
__device__ double log_jn(int n, double alpha) {
    const int lane = threadIdx.x & 63;
    const double dn = (double)n;
    const double sq = sqrt(alpha * alpha + 4.0 * dn);
    const double ss = alpha >= 0.0 ? 0.5 * (alpha + sq) : 2.0 * dn / (sq - alpha);
    const double fs = dn * log(ss) - 0.5 * ss * ss + alpha * ss;
    const double w = 1.0 / sqrt(dn / (ss * ss) + 1.0);
    const double h = 0.25 * w;
    double sum = 0.0;
    for (int base = 0, it = 0; it < 256; base += 64, it++) {  // peak and right side
        const double s = ss + (double)(base + lane) * h;
        const double e = exp(dn * log(s) - 0.5 * s * s + alpha * s - fs);
        sum += e;
        if (!(__shfl(e, 63, 64) >= 1e-40)) break;
    }
    for (int base = 1, it = 0; it < 256; base += 64, it++) {  // left side
        const double s = ss - (double)(base + lane) * h;
        const double e = s > 0.0 ? exp(dn * log(s) - 0.5 * s * s + alpha * s - fs) : 0.0;
        sum += e;
        if (!(__shfl(e, 63, 64) >= 1e-40)) break;
    }
    return fs + log(h * wsum(sum));
}

// coordinates per lane (d = lane + 64 t) at the largest N; loops run over
// all kMaxPer slots, unrolled, so the per-lane arrays stay in registers
constexpr int kMaxPer = (2 * kLearnedMaxN + 63) / 64;

// log density of unit vector y (this lane's coordinates d = lane + 64 t)
// under the projected Gaussian with means mu and log-scales ls.
__device__ __forceinline__ double log_proj_density(int dim, const double* y, const double* mu, const double* ls) {
    double A = 0.0, B = 0.0, C = 0.0, L = 0.0;
#pragma unroll
    for (int t = 0; t < kMaxPer; t++) {
        const double iv = exp(-2.0 * ls[t]);
        A += y[t] * y[t] * iv;
        B += y[t] * mu[t] * iv;
        C += mu[t] * mu[t] * iv;
        L += ls[t];
    }
    A = wsum(A);
    B = wsum(B);
    C = wsum(C);
    L = wsum(L);
    const double hd = 0.5 * (double)dim;
    return -hd * log(2.0 * M_PI) - L - 0.5 * C - hd * log(A) + log_jn(dim - 1, B / sqrt(A));
}



// Gauge-fix amplitudes a [N][2] with c and write the FP64 frame coordinates
// and the BF16 network input row (beta appended, zero padding).
__device__ __forceinline__ void write_frame(const double* __restrict__ a, int N, double cr, double ci, double beta,
                                            double* __restrict__ U, uint16_t* __restrict__ X, int in_pad) {
    const int lane = threadIdx.x & 63;
    for (int k = lane; k < N; k += 64) {
        const double re = a[2 * k] * cr - a[2 * k + 1] * ci;
        const double im = a[2 * k] * ci + a[2 * k + 1] * cr;
        U[2 * k] = re;
        U[2 * k + 1] = im;
        X[2 * k] = f2bf16((float)re);
        X[2 * k + 1] = f2bf16((float)im);
    }
    for (int c = 2 * N + lane; c < in_pad; c += 64) X[c] = c == 2 * N ? f2bf16((float)beta) : (uint16_t)0;
}

__global__ __launch_bounds__(64) void k_lp_prep(MoveArgs a) {
    const int r = blockIdx.x;
    const double* s = a.spins + (size_t)2 * a.N * r;
    double cr, ci;
    gauge_of(s, a.N, cr, ci);
    write_frame(s, a.N, cr, ci, a.betas[r], a.U + (size_t)2 * a.N * r, a.X0 + (size_t)a.in_pad * r, a.in_pad);
    if (threadIdx.x == 0) {
        a.gc[2 * r] = cr;
        a.gc[2 * r + 1] = ci;
    }
}

__device__ __forceinline__ double clampd(double x, double lo, double hi) { return fmin(fmax(x, lo), hi); }

__global__ __launch_bounds__(64) void k_lp_propose(MoveArgs a) {
    const int r = blockIdx.x, lane = threadIdx.x;
    const int N = a.N, dim = 2 * N;
    const uint64_t seed = a.seeds[r], mv = a.mctr[r];
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    const double* U = a.U + (size_t)dim * r;
    const float* O = a.Of + (size_t)a.out_ld * r;
    double mu[kMaxPer], ls[kMaxPer], z[kMaxPer];
    double z2 = 0.0;
#pragma unroll
    for (int t = 0; t < kMaxPer; t++) {
        const int d = lane + 64 * t;
        mu[t] = 0.0;
        ls[t] = 0.0;
        z[t] = 0.0;
        if (d < dim) {
            mu[t] = U[d] + a.mean_scale * (double)O[d];
            ls[t] = clampd((double)O[dim + d] + a.ls0, a.ls_min, a.ls_max) + a.log_scale;
            const uint32_t b = (uint32_t)(d >> 1);
            const U4 x = philox4x32_10(U4{(uint32_t)mv, (uint32_t)(mv >> 32), b, kCtrLearned}, k0, k1);
            const double u1 = u64_to_unit(x.x, x.y), u2 = u64_to_unit(x.z, x.w);
            const double rad = sqrt(-2.0 * log(u1));
            double sn, cs;
            sincos(2.0 * M_PI * u2, &sn, &cs);
            const double xi = (d & 1) ? rad * sn : rad * cs;
            z[t] = mu[t] + exp(ls[t]) * xi;
            z2 += z[t] * z[t];
        }
    }
    z2 = wsum(z2);
    const double inv = 1.0 / sqrt(z2);
#pragma unroll
    for (int t = 0; t < kMaxPer; t++) z[t] *= inv;  // unit vector y (zero beyond dim)
#pragma unroll
    for (int t = 0; t < kMaxPer; t++)
        if (lane + 64 * t >= dim) mu[t] = 0.0, ls[t] = 0.0;
    const double lqf = log_proj_density(dim, z, mu, ls);
    // back to the physical frame: a' = sqrt(N) y conj(c); (re, im) of a
    // site sit in neighbouring lanes
    const double cr = a.gc[2 * r], ci = a.gc[2 * r + 1], rn = sqrt((double)N);
    double* P = a.prop + (size_t)dim * r;
    double pa[kMaxPer];
#pragma unroll
    for (int t = 0; t < kMaxPer; t++) {
        const double y = z[t] * rn;
        const double other = __shfl_xor(y, 1, 64);
        const int d = lane + 64 * t;
        const double re = (d & 1) ? other : y;
        const double im = (d & 1) ? y : other;
        pa[t] = (d & 1) ? (im * cr - re * ci) : (re * cr + im * ci);
        if (d < dim) P[d] = pa[t];
    }
    // gauge of the proposal, from registers
    double best = -1.0;
    int bk = 0x7FFFFFFF;
#pragma unroll
    for (int t = 0; t < kMaxPer; t++) {
        const double o = __shfl_xor(pa[t], 1, 64);
        const int d = lane + 64 * t;
        if (d < dim && !(d & 1)) {
            const double x = pa[t] * pa[t] + o * o;
            if (x > best) {
                best = x;
                bk = d >> 1;
            }
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double ob = __shfl_xor(best, o, 64);
        const int ok = __shfl_xor(bk, o, 64);
        if (ob > best || (ob == best && ok < bk)) {
            best = ob;
            bk = ok;
        }
    }
    const int ts = (2 * bk) >> 6, ls_ = (2 * bk) & 63;
    double sel = 0.0;
#pragma unroll
    for (int t = 0; t < kMaxPer; t++)
        if (t == ts) sel = pa[t];
    const double kre = __shfl(sel, ls_, 64), kim = __shfl(sel, ls_ + 1, 64);
    const double rk = sqrt(best);
    const double cpr = kre / rk, cpi = -kim / rk;
    double* Up = a.Up + (size_t)dim * r;
    uint16_t* Xp = a.X0p + (size_t)a.in_pad * r;
#pragma unroll
    for (int t = 0; t < kMaxPer; t++) {
        const double o = __shfl_xor(pa[t], 1, 64);
        const int d = lane + 64 * t;
        const double re = (d & 1) ? o : pa[t];
        const double im = (d & 1) ? pa[t] : o;
        const double u = (d & 1) ? (re * cpi + im * cpr) : (re * cpr - im * cpi);
        if (d < dim) {
            Up[d] = u;
            Xp[d] = f2bf16((float)u);
        }
    }
    for (int c = dim + lane; c < a.in_pad; c += 64) Xp[c] = c == dim ? f2bf16((float)a.betas[r]) : (uint16_t)0;
    if (lane == 0) {
        a.gcp[2 * r] = cpr;
        a.gcp[2 * r + 1] = cpi;
        a.dbg[8 * r + 0] = lqf;
    }
}

// 16-byte aligned, so LDS reads and writes of one are single ds_read_b128 /
// ds_write_b128 (a wave's 64 consecutive entries then touch every bank
// equally; two 8-byte halves conflict two-way)
struct alignas(16) c64d {
    double x, y;
};
__device__ __forceinline__ c64d cmul(c64d p, c64d q) { return {p.x * q.x - p.y * q.y, p.x * q.y + p.y * q.x}; }
__device__ __forceinline__ c64d ccj(c64d p) { return {p.x, -p.y}; }

// Full FP64 energies of 2R configurations (current states, then proposals).
// A workgroup holds 64 configurations, one per lane, with their amplitudes
// in LDS as [site][configuration] (16 B per entry: a wave's read of one site
// is 1 KB of consecutive addresses, free of bank conflicts), and takes one
// of n_split ranges of the term list; its 16 waves take a sixteenth of the
// range each.  A wave loads its terms 64 at a time with one vector load and
// broadcasts each with readlane, so a term's indices are scalars and its
// branches uniform.
//
// A term is {x, 0, g}: x = i | j << 7 | k << 14 | l << 21 | ch << 28.  The
// list is grouped by key = x >> 7 = (j, k, l, ch): a group's terms differ
// only in i, and its energy is -Re[(sum_i g a_i) u] with u fixed by the key:
//   ch0  a_i a_j a_k* a_l*   u = a_j (a_k a_l)*
//   ch1  a_i a_j* a_k* a_l   u = a_j* (a_k* a_l)
//   ch2  a_i a_j* a_k a_l*   u = a_j* (a_k a_l*)
//   ch3  a_i a_j*  (pairs)   u = a_j*
// so a term costs one amplitude read and two FMAs per configuration, and a
// group three reads and two complex products more.  The waves' sums are
// added in wave order, and the accept kernel adds the n_split partials in a
// fixed order.
constexpr int kEnLanes = 64;
constexpr int kEnWaves = 16;
constexpr int kEnThreads = kEnLanes * kEnWaves;
constexpr int kEnMaxSplit = 64;
constexpr int kEnTermsPerSplit = 2048;

// A run of terms: the open group's key, its sum S = sum_i g a_i and its u.
struct EnRun {
    c64d S, u;
    uint32_t key;
};

// Close the open group into en and open `key` when it differs.
__device__ __forceinline__ void en_group(EnRun& r, uint32_t key, const c64d* A, double& en) {
    if (key == r.key) return;
    en -= r.S.x * r.u.x - r.S.y * r.u.y;
    r.S = {0.0, 0.0};
    const uint32_t ch = key >> 21;
    const c64d aj = A[(key & 127u) * kEnLanes];
    if (ch == 3u) {
        r.u = ccj(aj);
    } else {
        const c64d ak = A[((key >> 7) & 127u) * kEnLanes], al = A[((key >> 14) & 127u) * kEnLanes];
        if (ch == 0u) r.u = cmul(aj, ccj(cmul(ak, al)));
        else if (ch == 1u) r.u = cmul(ccj(aj), cmul(ccj(ak), al));
        else r.u = cmul(ccj(aj), cmul(ak, ccj(al)));
    }
    r.key = key;
}

__global__ __launch_bounds__(kEnThreads, 8) void k_lp_energy(const double* __restrict__ cur,
                                                             const double* __restrict__ prop, int R, int N,
                                                             const uint4* __restrict__ terms, long long nt,
                                                             double* __restrict__ part) {
    extern __shared__ c64d amp[];  // [N][kEnLanes]
    __shared__ double red[kEnWaves][kEnLanes];
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int c0 = blockIdx.x * kEnLanes;
    for (int e = threadIdx.x; e < N * kEnLanes; e += kEnThreads) {
        const int k = e >> 6, c = c0 + (e & 63);
        c64d v = {0.0, 0.0};
        if (c < 2 * R) {
            const double* src = c < R ? cur + (size_t)2 * N * c : prop + (size_t)2 * N * (c - R);
            const double2 d = *(const double2*)(src + 2 * k);  // rows of 2N doubles: 16-byte aligned
            v = {d.x, d.y};
        }
        amp[e] = v;
    }
    __syncthreads();
    const long long nq = (long long)gridDim.y * kEnWaves;
    const long long wq = (long long)blockIdx.y * kEnWaves + wv;
    const long long t0 = nt * wq / nq, t1 = nt * (wq + 1) / nq;
    const c64d* A = amp + lane;
    double en = 0.0;
    EnRun run{{0.0, 0.0}, {0.0, 0.0}, 0xFFFFFFFFu};
    for (long long b = t0; b < t1; b += kEnLanes) {
        // one load per 64 terms (other waves cover its latency)
        const uint4 w = b + lane < t1 ? terms[b + lane] : make_uint4(0u, 0u, 0u, 0u);
        const int n = (int)min((long long)kEnLanes, t1 - b);
        for (int e = 0; e < n; e++) {
            const uint32_t x = __builtin_amdgcn_readlane(w.x, e);
            const double g = __hiloint2double(__builtin_amdgcn_readlane((int)w.w, e),
                                              __builtin_amdgcn_readlane((int)w.z, e));
            const c64d ai = A[(x & 127u) * kEnLanes];
            en_group(run, x >> 7, A, en);
            run.S.x += g * ai.x;
            run.S.y += g * ai.y;
        }
    }
    en -= run.S.x * run.u.x - run.S.y * run.u.y;
    red[wv][lane] = en;
    __syncthreads();
    if (threadIdx.x < kEnLanes) {
        const int c = c0 + threadIdx.x;
        if (c < 2 * R) {
            double v = 0.0;
            for (int x = 0; x < kEnWaves; x++) v += red[x][threadIdx.x];
            part[(size_t)blockIdx.y * 2 * R + c] = v;
        }
    }
}

__global__ __launch_bounds__(64) void k_lp_accept(MoveArgs a) {
    const int r = blockIdx.x, lane = threadIdx.x;
    const int N = a.N, dim = 2 * N;
    double ec = 0.0, en = 0.0;
    for (int s = lane; s < a.n_split; s += 64) {
        ec += a.part[(size_t)s * 2 * a.R + r];
        en += a.part[(size_t)s * 2 * a.R + a.R + r];
    }
    ec = wsum(ec);
    en = wsum(en);
    // reverse density: current state in the proposal's frame, under the
    // network output at the proposal
    const double cr = a.gcp[2 * r], ci = a.gcp[2 * r + 1];
    const double* S = a.spins + (size_t)dim * r;
    const double* Up = a.Up + (size_t)dim * r;
    const float* O = a.Orv + (size_t)a.out_ld * r;
    double v[kMaxPer], mu[kMaxPer], ls[kMaxPer];
    double v2 = 0.0, lw = 0.0, lwp = 0.0;
    const double* P = a.prop + (size_t)dim * r;
#pragma unroll
    for (int t = 0; t < kMaxPer; t++) {
        const int d = lane + 64 * t;
        v[t] = mu[t] = ls[t] = 0.0;
        if (d < dim) {
            const int k = d >> 1;
            const double re = S[2 * k], im = S[2 * k + 1];
            v[t] = (d & 1) ? re * ci + im * cr : re * cr - im * ci;
            v2 += v[t] * v[t];
            mu[t] = Up[d] + a.mean_scale * (double)O[d];
            ls[t] = clampd((double)O[dim + d] + a.ls0, a.ls_min, a.ls_max) + a.log_scale;
            if (!(d & 1)) {
                lw -= 0.5 * log(re * re + im * im);
                lwp -= 0.5 * log(P[2 * k] * P[2 * k] + P[2 * k + 1] * P[2 * k + 1]);
            }
        }
    }
    v2 = wsum(v2);
    lw = wsum(lw);
    lwp = wsum(lwp);
    const double inv = 1.0 / sqrt(v2);
#pragma unroll
    for (int t = 0; t < kMaxPer; t++) v[t] *= inv;
    const double lqr = log_proj_density(dim, v, mu, ls);
    const double lqf = a.dbg[8 * r + 0];
    const double beta = a.betas[r];
    const double logr = -beta * (en - ec) + (lwp - lw) + (lqr - lqf);
    const uint64_t seed = a.seeds[r], mv = a.mctr[r];
    const U4 x = philox4x32_10(U4{(uint32_t)mv, (uint32_t)(mv >> 32), 0xFFFFFFFFu, kCtrLearned}, (uint32_t)seed,
                               (uint32_t)(seed >> 32));
    const double u = u64_to_unit(x.x, x.y);
    const bool acc = log(u) < logr;
    int slot = 0;
    if (a.perm) {
        int found = -1;
        for (int j = lane; j < a.R; j += 64)
            if (a.perm[j] == r) found = j;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) found = max(found, __shfl_xor(found, o, 64));
        slot = found >= 0 ? found / a.nrep_per_temp : 0;
    }
    __syncthreads();  // all lanes have read the current state
    if (acc) {
        double* W = a.spins_w + (size_t)dim * r;
        for (int d = lane; d < dim; d += 64) W[d] = P[d];
    }
    if (lane == 0) {
        a.E[r] = acc ? en : ec;
        double* D = a.dbg + 8 * r;
        D[1] = lqr;
        D[2] = ec;
        D[3] = en;
        D[4] = logr;
        D[5] = acc ? 1.0 : 0.0;
        D[6] = lw;
        D[7] = lwp;
        atomicAdd((unsigned long long*)&a.lprop[slot], 1ULL);
        if (acc) atomicAdd((unsigned long long*)&a.lacc[slot], 1ULL);
        a.mctr[r] = mv + 1;
    }
}

hipError_t lp_energy_prepare() {
    // dynamic LDS of [N][64] amplitudes: 128 KB at the largest N
    return hipFuncSetAttribute((const void*)k_lp_energy, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)(kLearnedMaxN * kEnLanes * sizeof(c64d)));
}

int lp_energy_blocks(int R) { return (2 * R + kEnLanes - 1) / kEnLanes; }
int lp_energy_splits(long long n_terms) {
    // about kEnTermsPerSplit terms per workgroup (128 per wave), at most kEnMaxSplit ranges
    return (int)std::max(1LL, std::min<long long>(kEnMaxSplit, (n_terms + kEnTermsPerSplit - 1) / kEnTermsPerSplit));
}

hipError_t launch_lp_prep(const MoveArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_lp_prep, dim3(a.R), dim3(64), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_lp_propose(const MoveArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_lp_propose, dim3(a.R), dim3(64), 0, s, a);
    return hipGetLastError();
}
hipError_t launch_lp_energy(const MoveArgs& a, const uint4* terms, long long nt, double* part, hipStream_t s) {
    if (a.N > kLearnedMaxN) return hipErrorInvalidValue;
    dim3 grid(lp_energy_blocks(a.R), a.n_split);
    hipLaunchKernelGGL(k_lp_energy, grid, dim3(kEnThreads), (size_t)a.N * kEnLanes * sizeof(c64d), s, a.spins,
                       (const double*)a.prop, a.R, a.N, terms, nt, part);
    return hipGetLastError();
}
hipError_t launch_lp_accept(const MoveArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_lp_accept, dim3(a.R), dim3(64), 0, s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Training data path
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_lp_record(const double* __restrict__ spins, const double* __restrict__ betas,
                                                  int dim, float* __restrict__ out, float* __restrict__ obeta) {
    const int r = blockIdx.x;
    for (int d = threadIdx.x; d < dim; d += 64) out[(size_t)dim * r + d] = (float)spins[(size_t)dim * r + d];
    if (threadIdx.x == 0) obeta[r] = (float)betas[r];
}

hipError_t launch_lp_record(const double* spins, const double* betas, int R, int dim, float* out, float* obeta,
                            hipStream_t s) {
    hipLaunchKernelGGL(k_lp_record, dim3(R), dim3(64), 0, s, spins, betas, dim, out, obeta);
    return hipGetLastError();
}

// Row m of a training batch: a random (snapshot t, replica r) with
// t + lag < n_snap; input = gauge-fixed snapshot t, target = snapshot
// t + lag of the same replica in the same frame.
__global__ __launch_bounds__(64) void k_lp_gather(const float* __restrict__ data, const float* __restrict__ dbeta,
                                                  long long n_snap, int R, int N, int lag, uint64_t seed,
                                                  uint64_t step, int in_pad, uint16_t* __restrict__ X0,
                                                  float* __restrict__ U, float* __restrict__ Y) {
    __shared__ double a[2 * kMaxN];
    const int m = blockIdx.x, lane = threadIdx.x, dim = 2 * N;
    const U4 x = philox4x32_10(U4{(uint32_t)step, (uint32_t)(step >> 32), (uint32_t)m, kCtrLearned + 1},
                               (uint32_t)seed, (uint32_t)(seed >> 32));
    const long long nt = n_snap - lag;
    long long t = (long long)(u32_to_unit(x.x) * (double)nt);
    int r = (int)(u32_to_unit(x.y) * (double)R);
    t = t < nt ? t : nt - 1;
    r = r < R ? r : R - 1;
    const float* A = data + ((size_t)t * R + r) * dim;
    const float* B = data + ((size_t)(t + lag) * R + r) * dim;
    for (int d = lane; d < dim; d += 64) a[d] = (double)A[d];
    __syncthreads();
    double cr, ci;
    gauge_of(a, N, cr, ci);
    for (int k = lane; k < N; k += 64) {
        const double ur = a[2 * k] * cr - a[2 * k + 1] * ci, ui = a[2 * k] * ci + a[2 * k + 1] * cr;
        const double br = B[2 * k], bi = B[2 * k + 1];
        U[(size_t)m * dim + 2 * k] = (float)ur;
        U[(size_t)m * dim + 2 * k + 1] = (float)ui;
        Y[(size_t)m * dim + 2 * k] = (float)(br * cr - bi * ci);
        Y[(size_t)m * dim + 2 * k + 1] = (float)(br * ci + bi * cr);
        X0[(size_t)m * in_pad + 2 * k] = f2bf16((float)ur);
        X0[(size_t)m * in_pad + 2 * k + 1] = f2bf16((float)ui);
    }
    const float beta = dbeta[(size_t)t * R + r];
    for (int c = dim + lane; c < in_pad; c += 64) X0[(size_t)m * in_pad + c] = c == dim ? f2bf16(beta) : (uint16_t)0;
}

hipError_t launch_lp_gather(const float* data, const float* dbeta, long long n_snap, int R, int N, int lag,
                            uint64_t seed, uint64_t step, int M, int in_pad, uint16_t* X0, float* U, float* Y,
                            hipStream_t s) {
    hipLaunchKernelGGL(k_lp_gather, dim3(M), dim3(64), 0, s, data, dbeta, n_snap, R, N, lag, seed, step, in_pad, X0,
                       U, Y);
    return hipGetLastError();
}

// Gaussian negative log-likelihood per row (mean = U + O[:dim], log-scale =
// clamp(O[dim:] + ls0)), its gradient w.r.t. O scaled by inv_m, as BF16.
// Rows m >= M_valid get zero loss and gradient.
__global__ __launch_bounds__(64) void k_lp_nll(const float* __restrict__ O, int out_ld, const float* __restrict__ U,
                                               const float* __restrict__ Y, int dim, int M_valid, double ls0,
                                               double ls_min, double ls_max, double inv_m,
                                               uint16_t* __restrict__ dO, double* __restrict__ rowloss) {
    const int m = blockIdx.x, lane = threadIdx.x;
    double loss = 0.0;
    for (int d = lane; d < dim; d += 64) {
        float gm = 0.0f, gl = 0.0f;
        if (m < M_valid) {
            const double mu = (double)U[(size_t)m * dim + d] + (double)O[(size_t)m * out_ld + d];
            const double raw = (double)O[(size_t)m * out_ld + dim + d] + ls0;
            const double ls = clampd(raw, ls_min, ls_max);
            const double is = exp(-ls);
            const double rr = ((double)Y[(size_t)m * dim + d] - mu) * is;
            loss += 0.5 * rr * rr + ls;
            gm = (float)(-rr * is * inv_m);
            gl = (raw > ls_min && raw < ls_max) ? (float)((1.0 - rr * rr) * inv_m) : 0.0f;
        }
        dO[(size_t)m * out_ld + d] = f2bf16(gm);
        dO[(size_t)m * out_ld + dim + d] = f2bf16(gl);
    }
    loss = wsum(loss);
    if (lane == 0) rowloss[m] = loss;
}

hipError_t launch_lp_nll(const float* O, int out_ld, const float* U, const float* Y, int dim, int M_pad, int M_valid,
                         double ls0, double ls_min, double ls_max, uint16_t* dO, double* rowloss, hipStream_t s) {
    hipLaunchKernelGGL(k_lp_nll, dim3(M_pad), dim3(64), 0, s, O, out_ld, U, Y, dim, M_valid, ls0, ls_min, ls_max,
                       1.0 / (double)M_valid, dO, rowloss);
    return hipGetLastError();
}

// Fixed-order sum of rows [0, M) into out[0] (one workgroup).
__global__ __launch_bounds__(256) void k_sum_rows(const double* __restrict__ x, int M, double scale,
                                                  double* __restrict__ out) {
    __shared__ double red[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < M; i += 256) s += x[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0] * scale;
}

hipError_t launch_sum_rows(const double* x, int M, double scale, double* out, hipStream_t s) {
    hipLaunchKernelGGL(k_sum_rows, dim3(1), dim3(256), 0, s, x, M, scale, out);
    return hipGetLastError();
}

}  // namespace ps
