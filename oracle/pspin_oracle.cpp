// pspin_oracle.cpp -- CPU FP64 restatement of the reference's 2+4 p-spin
// Monte Carlo path.
//
// TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
// product path: only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it.  Nothing in cuda_pspin_amd/ links or calls it.
//
// Every function cites the reference file:line whose behaviour it restates
// (paths relative to the reference repository root).  Functions that belong to
// this build's own design (Philox streams, Box-Muller normals) are marked so:
// they exist here so that the GPU's use of them can be checked bit-for-bit.
//
// Parity pinning (what checks each function, see tests/test_oracle_*.py):
//   splitmix64            published test vector (seed 0)
//   omega / masks         hand-enumerated tiny cases, comb closed forms,
//                         brute-force recount, uniform-FMC survival law
//   colex encode/decode   exhaustive bijection N<=20
//   three_type            set equality with brute force and with the
//                         reference's own test_enum.cpp built in oracle/_ref
//   energy                gauge/phase invariance, homogeneity, per-channel
//                         trigonometric closed form, all-ones sum of couplings
//   delta_e               equals E(new)-E(old) by brute force
//   proposals             invariants (|a|^2 / |a|^4 sums), closed forms
//   pt_exchange           glibc drand48 itself; hand-traced 3-temperature case
//   partial_shuffle       permutation property + uniform-selection statistics
//   philox4x32_10         Random123 known-answer vectors
//   log_jn / lp density   closed form at alpha = 0, recurrence, quadrature,
//                         sphere normalisation, uniform special case
//   learned move (MH)     stationary law on the CPU: Dirichlet(1/2) intensities
//                         at beta = 0 and agreement with the pair-rotation
//                         chain at beta > 0
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <array>
#include <memory>
#include <vector>
#include <chrono>
#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {

// ---------------------------------------------------------------------------
// splitmix64 (src/mc.cu:42-47; src/mc_sparse.cu:38-43)
// ---------------------------------------------------------------------------
uint64_t or_splitmix64(uint64_t* state) {
    uint64_t z = (*state += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

// Per-replica seeds: chain of splitmix64 from the master seed
// (src/mc.cu:418-423).
void or_replica_seeds(uint64_t seed, int nrep, uint64_t* out) {
    uint64_t st = seed;
    for (int r = 0; r < nrep; r++) out[r] = or_splitmix64(&st);
}

// ---------------------------------------------------------------------------
// Counting helpers (include/disorder.h:32-44)
// ---------------------------------------------------------------------------
long long or_n_quartets(int N) { return (long long)N * (N - 1) * (N - 2) * (N - 3) / 24; }
long long or_n_pairs(int N) { return (long long)N * (N + 1) / 2; }

// ---------------------------------------------------------------------------
// Colex quartet index (src/mc.cu:241-243 encode; src/disorder.cu:159-171 decode)
// q = C(l,4) + C(k,3) + C(j,2) + i  for i<j<k<l
// ---------------------------------------------------------------------------
long long or_encode(int i, int j, int k, int l) {
    return (long long)l * (l - 1) * (l - 2) * (l - 3) / 24 + (long long)k * (k - 1) * (k - 2) / 6 +
           (long long)j * (j - 1) / 2 + i;
}

void or_decode(long long q, int* out4) {
    // Largest l with C(l,4) <= q, then k, j, and the remainder is i.
    int l = 3;
    while ((long long)(l + 1) * l * (l - 1) * (l - 2) / 24 <= q) l++;
    long long rem = q - (long long)l * (l - 1) * (l - 2) * (l - 3) / 24;
    int k = 2;
    while ((long long)(k + 1) * k * (k - 1) / 6 <= rem) k++;
    rem -= (long long)k * (k - 1) * (k - 2) / 6;
    int j = 1;
    while ((long long)(j + 1) * j / 2 <= rem) j++;
    rem -= (long long)j * (j - 1) / 2;
    out4[0] = (int)rem; out4[1] = j; out4[2] = k; out4[3] = l;
}

// ---------------------------------------------------------------------------
// Frequencies (src/mc.cu:449-460): comb omega_k = k, uniform omega from
// splitmix64(seed+3000) >> 11 scaled by 2^-53, then sorted ascending.
// fmc_mode 0 writes nothing.
// ---------------------------------------------------------------------------
void or_omega(int N, int fmc_mode, uint64_t seed, double* omega) {
    if (fmc_mode == 1) {
        for (int i = 0; i < N; i++) omega[i] = (double)i;
    } else if (fmc_mode == 2) {
        uint64_t st = seed + 3000;
        for (int i = 0; i < N; i++) {
            uint64_t z = or_splitmix64(&st);
            omega[i] = (double)(z >> 11) / (double)(1ULL << 53);
        }
        std::sort(omega, omega + N);
    }
}

// ---------------------------------------------------------------------------
// FMC filters (src/disorder.cu:130-143 for g2, 153-183 for the g4 mask).
// g2: keep[i*N+j] = 0 where i != j and |w_i - w_j| > gamma (both triangles);
//     the diagonal is never filtered (the filter only visits i<j).
// g4: first channel whose condition holds wins, evaluated left to right in
//     FP64 exactly as written: ((w_a + w_b) - w_c) - w_d.
// ---------------------------------------------------------------------------
void or_g2_keep(int N, const double* w, double gamma, uint8_t* keep) {
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) keep[i * N + j] = 1;
    for (int i = 0; i < N; i++)
        for (int j = i + 1; j < N; j++)
            if (std::fabs(w[i] - w[j]) > gamma) { keep[i * N + j] = 0; keep[j * N + i] = 0; }
}

uint8_t or_mask_one(const double* w, int i, int j, int k, int l, double gamma) {
    if (std::fabs(w[i] + w[j] - w[k] - w[l]) <= gamma) return 1u;
    if (std::fabs(w[i] + w[l] - w[j] - w[k]) <= gamma) return 2u;
    if (std::fabs(w[i] + w[k] - w[j] - w[l]) <= gamma) return 4u;
    return 0u;
}

// fmc_mode 0: every quartet is channel 0 (src/disorder.cu:186-189).
void or_g4_mask(int N, int fmc_mode, const double* w, double gamma, uint8_t* mask) {
    long long q = 0;
    for (int l = 3; l < N; l++)
        for (int k = 2; k < l; k++)
            for (int j = 1; j < k; j++)
                for (int i = 0; i < j; i++, q++)
                    mask[q] = (fmc_mode == 0) ? 1u : or_mask_one(w, i, j, k, l, gamma);
}

long long or_popcount_mask(const uint8_t* mask, long long n) {
    long long c = 0;
    for (long long q = 0; q < n; q++) c += __builtin_popcount(mask[q]);
    return c;
}

// ---------------------------------------------------------------------------
// 3-type enumeration of the quartets touched by a pair move (src/mc.cu:111-243).
// Writes sorted quartets (4 ints each) in the reference's t order: type 1
// {i0,a,b,c}, type 2 {j0,a,b,c}, type 3 {i0,j0,a,b}; returns the count.
// The (a,b,c) decode here walks colex order directly instead of the
// reference's cbrt/sqrt inversion; order and content are the same.
// ---------------------------------------------------------------------------
static inline int skip2(int x, int i0, int j0) {
    if (x >= i0) x++;
    if (x >= j0) x++;
    return x;
}
static inline void sort4(int* v) { std::sort(v, v + 4); }

long long or_three_type(int N, int i0, int j0, int* out) {
    long long n = 0;
    int M = N - 2;
    for (int type = 0; type < 2; type++) {
        int site = type == 0 ? i0 : j0;
        for (int cc = 2; cc < M; cc++)
            for (int bb = 1; bb < cc; bb++)
                for (int aa = 0; aa < bb; aa++) {
                    int v[4] = {site, skip2(aa, i0, j0), skip2(bb, i0, j0), skip2(cc, i0, j0)};
                    sort4(v);
                    std::memcpy(out + 4 * n, v, sizeof(v));
                    n++;
                }
    }
    for (int bb = 1; bb < M; bb++)
        for (int aa = 0; aa < bb; aa++) {
            int v[4] = {i0, j0, skip2(aa, i0, j0), skip2(bb, i0, j0)};
            sort4(v);
            std::memcpy(out + 4 * n, v, sizeof(v));
            n++;
        }
    return n;
}

// ---------------------------------------------------------------------------
// Hamiltonian as computed by the reference (src/mc.cu:344-394,
// src/test_delta_e.cu:32-82):
//   H2 = -sum_{i<j} Re[g_ij a_i conj(a_j)]      (diagonal never contributes)
//   H4 = -sum_{q active} Re[g_q f_ch(a)]  with
//        ch0: a_i a_j conj(a_k) conj(a_l)   (conj mask 0xC)
//        ch1: a_i conj(a_j) conj(a_k) a_l   (conj mask 0x6)
//        ch2: a_i conj(a_j) a_k conj(a_l)   (conj mask 0xA)
// Couplings are real (src/disorder.cu:73-74,110-111).  Spins are interleaved
// (re, im) doubles.  A quartet is active when mask != 0 and g != 0.
// ---------------------------------------------------------------------------
struct cplx { double re, im; };
static inline cplx cmul(cplx a, cplx b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
static inline cplx cconj(cplx a) { return {a.re, -a.im}; }
static inline cplx ld(const double* s, int k) { return {s[2 * k], s[2 * k + 1]}; }

static const int kConjMask[3] = {0xC, 0x6, 0xA};

static inline cplx quartet_product(const double* s, const int* idx, int ch) {
    int cm = kConjMask[ch];
    cplx f[4];
    for (int p = 0; p < 4; p++) {
        cplx v = ld(s, idx[p]);
        f[p] = ((cm >> p) & 1) ? cconj(v) : v;
    }
    return cmul(cmul(f[0], f[1]), cmul(f[2], f[3]));
}

static inline int mask_channel(uint8_t m) { return (m & 1) ? 0 : (m & 2) ? 1 : (m & 4) ? 2 : -1; }

double or_energy_h2(int N, const double* s, const double* g2) {
    double e = 0.0;
    for (int i = 0; i < N; i++)
        for (int j = i + 1; j < N; j++) {
            double g = g2[i * N + j];
            cplx p = cmul(ld(s, i), cconj(ld(s, j)));
            e += -g * p.re;
        }
    return e;
}

double or_energy_h4(int N, const double* s, const double* g4, const uint8_t* mask) {
    double e = 0.0;
    long long q = 0;
    for (int l = 3; l < N; l++)
        for (int k = 2; k < l; k++)
            for (int j = 1; j < k; j++)
                for (int i = 0; i < j; i++, q++) {
                    uint8_t m = mask[q];
                    if (!m || g4[q] == 0.0) continue;
                    int idx[4] = {i, j, k, l};
                    for (int ch = 0; ch < 3; ch++)
                        if (m & (1 << ch)) e += -g4[q] * quartet_product(s, idx, ch).re;
                }
    return e;
}

void or_energy(int N, const double* s, const double* g2, int h2_active, const double* g4,
               const uint8_t* mask, double* e2, double* e4) {
    *e2 = h2_active ? or_energy_h2(N, s, g2) : 0.0;
    *e4 = or_energy_h4(N, s, g4, mask);
}

// Sparse H4 (src/mc_sparse.cu:292-306; src/test_sparse_dense.cu:43-75).
// quartets: nq rows of (i,j,k,l,ch); g: nq real couplings.
double or_energy_h4_sparse(const double* s, int nq, const int* quartets, const double* g) {
    double e = 0.0;
    for (int t = 0; t < nq; t++) e += -g[t] * quartet_product(s, quartets + 5 * t, quartets[5 * t + 4]).re;
    return e;
}

// ---------------------------------------------------------------------------
// Per-proposal Delta E in FP64 following the reference sweep's enumeration
// (src/mc.cu:160-275): factorised H2 rows for k != i0,j0 plus the (i0,j0)
// pair term, and full new-minus-old products over the 3-type quartets.
// scale receives sum over touched terms of |g| (|f_old| + |f_new|), the
// magnitude that bounds the FP32 rounding of the same sum.
// ---------------------------------------------------------------------------
void or_delta_e(int N, const double* s, const double* g2, int h2_active, const double* g4,
                const uint8_t* mask, int i0, int j0, const double* ai_new, const double* aj_new,
                double* dE_out, double* scale_out) {
    cplx ao_i = ld(s, i0), ao_j = ld(s, j0);
    cplx an_i = {ai_new[0], ai_new[1]}, an_j = {aj_new[0], aj_new[1]};
    cplx di = {an_i.re - ao_i.re, an_i.im - ao_i.im};
    cplx dj = {an_j.re - ao_j.re, an_j.im - ao_j.im};
    double dE = 0.0, scale = 0.0;
    if (h2_active) {
        cplx si = {0, 0}, sj = {0, 0};
        for (int k = 0; k < N; k++) {
            if (k == i0 || k == j0) continue;
            cplx akc = cconj(ld(s, k));
            double gi = g2[i0 * N + k], gj = g2[j0 * N + k];
            si.re += gi * akc.re; si.im += gi * akc.im;
            sj.re += gj * akc.re; sj.im += gj * akc.im;
            double ak = std::hypot(akc.re, akc.im);
            scale += (std::fabs(gi) * (std::hypot(ao_i.re, ao_i.im) + std::hypot(an_i.re, an_i.im)) +
                      std::fabs(gj) * (std::hypot(ao_j.re, ao_j.im) + std::hypot(an_j.re, an_j.im))) * ak;
        }
        dE -= cmul(di, si).re;
        dE -= cmul(dj, sj).re;
        double gij = g2[i0 * N + j0];
        cplx op = cmul(ao_i, cconj(ao_j)), np = cmul(an_i, cconj(an_j));
        dE -= gij * (np.re - op.re);
        scale += std::fabs(gij) * (std::hypot(op.re, op.im) + std::hypot(np.re, np.im));
    }
    std::vector<int> qs((size_t)4 * (2 * (N - 2) * (N - 3) * (N - 4) / 6 + (N - 2) * (N - 3) / 2 + 1));
    long long n = or_three_type(N, i0, j0, qs.data());
    std::vector<double> snew(s, s + 2 * N);
    snew[2 * i0] = an_i.re; snew[2 * i0 + 1] = an_i.im;
    snew[2 * j0] = an_j.re; snew[2 * j0 + 1] = an_j.im;
    for (long long t = 0; t < n; t++) {
        const int* v = &qs[4 * t];
        long long q = or_encode(v[0], v[1], v[2], v[3]);
        uint8_t m = mask[q];
        double g = g4[q];
        if (g == 0.0 || m == 0) continue;
        for (int ch = 0; ch < 3; ch++) {
            if (!(m & (1 << ch))) continue;
            cplx fo = quartet_product(s, v, ch), fn = quartet_product(snew.data(), v, ch);
            dE -= g * (fn.re - fo.re);
            scale += std::fabs(g) * (std::hypot(fo.re, fo.im) + std::hypot(fn.re, fn.im));
        }
    }
    *dE_out = dE;
    if (scale_out) *scale_out = scale;
}

// Sparse-quartet Delta E (src/mc_sparse.cu:146-210), same scale convention.
void or_delta_e_sparse(int N, const double* s, const double* g2, int h2_active, int nq,
                       const int* quartets, const double* g, int i0, int j0, const double* ai_new,
                       const double* aj_new, double* dE_out, double* scale_out) {
    std::vector<double> snew(s, s + 2 * N);
    snew[2 * i0] = ai_new[0]; snew[2 * i0 + 1] = ai_new[1];
    snew[2 * j0] = aj_new[0]; snew[2 * j0 + 1] = aj_new[1];
    double dE = 0.0, scale = 0.0;
    if (h2_active) {
        // Same rows as the dense path: pairs touching i0 or j0.
        for (int i = 0; i < N; i++)
            for (int j = i + 1; j < N; j++) {
                if (i != i0 && i != j0 && j != i0 && j != j0) continue;
                double gg = g2[i * N + j];
                cplx po = cmul(ld(s, i), cconj(ld(s, j)));
                cplx pn = cmul(ld(snew.data(), i), cconj(ld(snew.data(), j)));
                dE -= gg * (pn.re - po.re);
                scale += std::fabs(gg) * (std::hypot(po.re, po.im) + std::hypot(pn.re, pn.im));
            }
    }
    for (int t = 0; t < nq; t++) {
        const int* v = quartets + 5 * t;
        bool inv = false;
        for (int p = 0; p < 4; p++) inv |= (v[p] == i0 || v[p] == j0);
        if (!inv) continue;
        cplx fo = quartet_product(s, v, v[4]), fn = quartet_product(snew.data(), v, v[4]);
        dE -= g[t] * (fn.re - fo.re);
        scale += std::fabs(g[t]) * (std::hypot(fo.re, fo.im) + std::hypot(fn.re, fn.im));
    }
    *dE_out = dE;
    if (scale_out) *scale_out = scale;
}

// ---------------------------------------------------------------------------
// Proposals given explicit uniforms.
// Pair pick (src/mc.cu:129-141): u in (0,1].
// ---------------------------------------------------------------------------
void or_pair_pick(int N, double u1, double u2, int* i0, int* j0) {
    int ri = (int)(u1 * N);
    if (ri >= N) ri = N - 1;
    int rj = (int)(u2 * (N - 1));
    if (rj >= N - 1) rj = N - 2;
    if (rj >= ri) rj++;
    *i0 = ri < rj ? ri : rj;
    *j0 = ri < rj ? rj : ri;
}

// Spherical pair rotation (include/spins.h:17-52).
void or_propose_sphere(const double* ai, const double* aj, double u_alpha, double u_phi1,
                       double u_phi2, double* out_i, double* out_j) {
    double alpha = 2.0 * M_PI * u_alpha, phi1 = 2.0 * M_PI * u_phi1, phi2 = 2.0 * M_PI * u_phi2;
    double sa = std::sin(alpha), ca = std::cos(alpha);
    double s1 = std::sin(phi1), c1 = std::cos(phi1), s2 = std::sin(phi2), c2 = std::cos(phi2);
    double r1 = ai[0] * ai[0] + ai[1] * ai[1], r2 = aj[0] * aj[0] + aj[1] * aj[1];
    double nx1 = ai[0] * c1 + ai[1] * s1, ny1 = -ai[0] * s1 + ai[1] * c1;
    double nx2 = aj[0] * c2 + aj[1] * s2, ny2 = -aj[0] * s2 + aj[1] * c2;
    double f1 = std::sqrt((r1 + r2) / r1) * ca, f2 = std::sqrt((r1 + r2) / r2) * sa;
    out_i[0] = nx1 * f1; out_i[1] = ny1 * f1;
    out_j[0] = nx2 * f2; out_j[1] = ny2 * f2;
}

// Cube move (include/spins_sparse.h:26-56): absolute new phases, S4 split by u.
void or_propose_cube(const double* ai, const double* aj, double u, double u_phi1, double u_phi2,
                     double* out_i, double* out_j) {
    double phi1 = u_phi1 * 2.0 * M_PI, phi2 = u_phi2 * 2.0 * M_PI;
    double r1 = ai[0] * ai[0] + ai[1] * ai[1], r2 = aj[0] * aj[0] + aj[1] * aj[1];
    double S4 = r1 * r1 + r2 * r2, sq = std::sqrt(S4);
    double m1 = std::sqrt(sq * std::sqrt(1.0 - u)), m2 = std::sqrt(sq * std::sqrt(u));
    out_i[0] = m1 * std::cos(phi1); out_i[1] = m1 * std::sin(phi1);
    out_j[0] = m2 * std::cos(phi2); out_j[1] = m2 * std::sin(phi2);
}

// Metropolis rule (src/mc.cu:285-304): beta==0 accepts; dE<=0 accepts; else
// accept when u < exp(-beta dE).
int or_metropolis(double beta, double dE, double u) {
    if (beta == 0.0) return 1;
    if (dE <= 0.0) return 1;
    return u < std::exp(-beta * dE) ? 1 : 0;
}

// ---------------------------------------------------------------------------
// Temperature schedule (src/parallel_tempering.cu:93-109).
// ---------------------------------------------------------------------------
void or_temperatures(int NT, double Tmin, double Tmax, int log_temp, double* T, double* beta) {
    if (log_temp) {
        double A = (NT > 1) ? std::pow(Tmin / Tmax, 1.0 / (NT - 1)) : 1.0;
        for (int t = 0; t < NT; t++) { T[t] = Tmax * std::pow(A, t); beta[t] = 1.0 / T[t]; }
    } else {
        double dT = (NT > 1) ? (Tmax - Tmin) / (NT - 1) : 0.0;
        for (int t = 0; t < NT; t++) { T[t] = Tmax - t * dT; beta[t] = 1.0 / T[t]; }
    }
}

// ---------------------------------------------------------------------------
// PT replica exchange (src/parallel_tempering.cu:316-350), sequential hot to
// cold, glibc drand48 drawn only when delta < 0.  The caller seeds with
// or_srand48((long)seed + 5000) as the reference does (:304).
// acc/prop have NT-1 entries and are incremented.  betas is rebuilt by
// physical replica index (:348-350).
// ---------------------------------------------------------------------------
void or_srand48(long s) { srand48(s); }
double or_drand48(void) { return drand48(); }

void or_pt_exchange(int NT, int nrep, const double* beta_sched, const double* E, int* perm,
                    long long* acc, long long* prop, double* betas) {
    for (int t = 0; t < NT - 1; t++)
        for (int r = 0; r < nrep; r++) {
            int a = perm[t * nrep + r], b = perm[(t + 1) * nrep + r];
            double delta = (beta_sched[t] - beta_sched[t + 1]) * (E[a] - E[b]);
            prop[t]++;
            bool accept = (delta >= 0.0);
            if (!accept) accept = (drand48() < std::exp(delta));
            if (accept) {
                acc[t]++;
                int tmp = perm[t * nrep + r];
                perm[t * nrep + r] = perm[(t + 1) * nrep + r];
                perm[(t + 1) * nrep + r] = tmp;
            }
        }
    if (betas)
        for (int t = 0; t < NT; t++)
            for (int r = 0; r < nrep; r++) betas[perm[t * nrep + r]] = beta_sched[t];
}

// ---------------------------------------------------------------------------
// Sparse quartet sampling (src/mc_sparse.cu:340-349, 489-537).
// Candidates: q*3+ch in ascending q for every set mask bit where g4[q] != 0.
// Partial Fisher-Yates with a splitmix64 stream seeded seed+4000.
// Returns n_sel; fills quartets (5 ints each) and couplings
// g = g4_unit[q]*J4*sqrt(N/n_sel) + J4_0*N/n_sel.  n_nonzero_out receives
// the candidate count.
// ---------------------------------------------------------------------------
void or_partial_shuffle(long long* arr, long long count, int n_select, uint64_t seed) {
    uint64_t st = seed;
    for (int i = 0; i < n_select && i < (int)count; i++) {
        uint64_t r = or_splitmix64(&st);
        long long j = i + (long long)(r % (uint64_t)(count - i));
        std::swap(arr[i], arr[j]);
    }
}

int or_sparse_select(int N, const uint8_t* mask, const double* g4_unit, int n_req, uint64_t seed,
                     double J4, double J4_0, int* quartets, double* g, long long* n_nonzero_out) {
    long long nq = or_n_quartets(N);
    std::vector<long long> cand;
    for (long long q = 0; q < nq; q++) {
        if (g4_unit[q] == 0.0) continue;
        for (int ch = 0; ch < 3; ch++)
            if (mask[q] & (1 << ch)) cand.push_back(q * 3 + ch);
    }
    long long n_nonzero = (long long)cand.size();
    if (n_nonzero_out) *n_nonzero_out = n_nonzero;
    int n_sel = n_req;
    if (n_nonzero < n_sel) n_sel = (int)n_nonzero;
    or_partial_shuffle(cand.data(), n_nonzero, n_sel, seed + 4000);
    double sigma = J4 * std::sqrt((double)N / (double)n_sel);
    double mean = (J4_0 != 0.0) ? J4_0 * (double)N / (double)n_sel : 0.0;
    for (int s = 0; s < n_sel; s++) {
        long long enc = cand[s];
        int ch = (int)(enc % 3);
        long long q = enc / 3;
        int v[4];
        or_decode(q, v);
        for (int p = 0; p < 4; p++) quartets[5 * s + p] = v[p];
        quartets[5 * s + 4] = ch;
        g[s] = g4_unit[q] * sigma + mean;
    }
    return n_sel;
}

// ---------------------------------------------------------------------------
// Build's own design, restated here as the checker of the GPU's use of it:
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), Box-Muller normals.
// ---------------------------------------------------------------------------
static inline void philox_round(uint32_t* c, const uint32_t* k) {
    uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c[1] ^ k[0];
    uint32_t n2 = hi0 ^ c[3] ^ k[1];
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
}

void or_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    uint32_t k[2] = {key[0], key[1]};
    for (int r = 0; r < 10; r++) {
        philox_round(c, k);
        if (r < 9) { k[0] += 0x9E3779B9u; k[1] += 0xBB67AE85u; }
    }
    std::memcpy(out, c, sizeof(c));
}

// Uniform in (0,1) from 32 bits: (x + 0.5) 2^-32.
double or_u32_to_unit(uint32_t x) { return ((double)x + 0.5) * (1.0 / 4294967296.0); }
// Uniform in (0,1) with 53 bits from two words.
double or_u64_to_unit(uint32_t hi, uint32_t lo) {
    uint64_t m = ((uint64_t)(hi >> 5) << 26) | (uint64_t)(lo >> 6);
    return ((double)m + 0.5) * (1.0 / 9007199254740992.0);
}

// Standard normal number k of the stream (seed, tag): one Philox block per
// pair, Box-Muller on two 53-bit uniforms.
double or_normal(uint64_t seed, uint32_t tag, uint64_t k) {
    uint64_t pair = k >> 1;
    uint32_t ctr[4] = {(uint32_t)pair, (uint32_t)(pair >> 32), 0u, tag};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t x[4];
    or_philox4x32_10(ctr, key, x);
    double u1 = or_u64_to_unit(x[0], x[1]), u2 = or_u64_to_unit(x[2], x[3]);
    double r = std::sqrt(-2.0 * std::log(u1));
    return (k & 1) ? r * std::sin(2.0 * M_PI * u2) : r * std::cos(2.0 * M_PI * u2);
}

// Uniforms of one sweep step of replica with seed rseed at global sweep
// counter `sweep`: u[0..5] = pair u1, pair u2, alpha (or cube u), phi1, phi2,
// Metropolis u.
void or_step_uniforms(uint64_t rseed, uint64_t sweep, int step, double* u) {
    uint32_t key[2] = {(uint32_t)rseed, (uint32_t)(rseed >> 32)};
    uint32_t c0[4] = {(uint32_t)sweep, (uint32_t)(sweep >> 32), (uint32_t)step, 16u};
    uint32_t c1[4] = {(uint32_t)sweep, (uint32_t)(sweep >> 32), (uint32_t)step, 17u};
    uint32_t x[4], y[4];
    or_philox4x32_10(c0, key, x);
    or_philox4x32_10(c1, key, y);
    u[0] = or_u32_to_unit(x[0]);
    u[1] = or_u32_to_unit(x[1]);
    u[2] = or_u32_to_unit(x[2]);
    u[3] = or_u32_to_unit(x[3]);
    u[4] = or_u32_to_unit(y[0]);
    u[5] = or_u32_to_unit(y[1]);
}

// Shared pair schedule (this build's option): the two pair uniforms of the
// step come from the group stream keyed by gkey, counter word 18; the rest
// as or_step_uniforms.  gkey = 0 means per-replica pairs.
void or_step_uniforms_grouped(uint64_t rseed, uint64_t gkey, uint64_t sweep, int step, double* u) {
    or_step_uniforms(rseed, sweep, step, u);
    if (gkey) {
        uint32_t key[2] = {(uint32_t)gkey, (uint32_t)(gkey >> 32)};
        uint32_t c[4] = {(uint32_t)sweep, (uint32_t)(sweep >> 32), (uint32_t)step, 18u}, z[4];
        or_philox4x32_10(c, key, z);
        u[0] = or_u32_to_unit(z[0]);
        u[1] = or_u32_to_unit(z[1]);
    }
}

// Group key of pair-schedule group g: splitmix64 of (seed + 7000 + g).
uint64_t or_pair_group_key(uint64_t seed, int g) {
    uint64_t s = seed + 7000ULL + (uint64_t)g;
    return or_splitmix64(&s);
}

// ---------------------------------------------------------------------------
// CPU baseline: OpenMP transcription of the reference dense sweep
// (src/mc.cu:66-318): one replica chain per thread, FP64 throughout, the
// reference's 3-type enumeration with cbrt/sqrt decode and colex gather,
// factorised H2, Metropolis with exp.  Uniforms come from or_step_uniforms.
// Returns wall seconds of the timed sweeps; energies/acc updated in place.
// ---------------------------------------------------------------------------
static void ref_sweep_one(int N, double* s, const double* g2, int h2_active, const double* g4,
                          const uint8_t* mask, double beta, uint64_t rseed, uint64_t sweep, double* E,
                          long long* acc) {
    long long n12 = (long long)(N - 2) * (N - 3) * (N - 4) / 6;
    long long n3 = (long long)(N - 2) * (N - 3) / 2;
    long long nh4 = 2 * n12 + n3;
    for (int step = 0; step < N / 2; step++) {
        double u[6];
        or_step_uniforms(rseed, sweep, step, u);
        int i0, j0;
        or_pair_pick(N, u[0], u[1], &i0, &j0);
        double ai[2] = {s[2 * i0], s[2 * i0 + 1]}, aj[2] = {s[2 * j0], s[2 * j0 + 1]};
        double ni[2], nj[2];
        or_propose_sphere(ai, aj, u[2], u[3], u[4], ni, nj);
        cplx di = {ni[0] - ai[0], ni[1] - ai[1]}, dj = {nj[0] - aj[0], nj[1] - aj[1]};
        double dE = 0.0;
        if (h2_active) {
            cplx si = {0, 0}, sj = {0, 0};
            for (int k = 0; k < N; k++) {
                if (k == i0 || k == j0) continue;
                cplx akc = cconj(ld(s, k));
                si.re += g2[i0 * N + k] * akc.re; si.im += g2[i0 * N + k] * akc.im;
                sj.re += g2[j0 * N + k] * akc.re; sj.im += g2[j0 * N + k] * akc.im;
            }
            dE -= cmul(di, si).re + cmul(dj, sj).re;
            double gij = g2[i0 * N + j0];
            cplx op = cmul({ai[0], ai[1]}, cconj({aj[0], aj[1]}));
            cplx np = cmul({ni[0], ni[1]}, cconj({nj[0], nj[1]}));
            dE -= gij * (np.re - op.re);
        }
        for (long long t = 0; t < nh4; t++) {
            int idx[4];
            if (t < 2 * n12) {
                long long tt = t < n12 ? t : t - n12;
                int cc = (int)std::cbrt(6.0 * (double)tt);
                if (cc < 2) cc = 2;
                while ((long long)cc * (cc - 1) * (cc - 2) / 6 > tt) cc--;
                while ((long long)(cc + 1) * cc * (cc - 1) / 6 <= tt) cc++;
                long long rem = tt - (long long)cc * (cc - 1) * (cc - 2) / 6;
                int bb = (int)std::sqrt(2.0 * (double)rem);
                if (bb < 1) bb = 1;
                while ((long long)bb * (bb - 1) / 2 > rem) bb--;
                while ((long long)(bb + 1) * bb / 2 <= rem) bb++;
                int aa = (int)(rem - (long long)bb * (bb - 1) / 2);
                idx[0] = t < n12 ? i0 : j0;
                idx[1] = skip2(aa, i0, j0); idx[2] = skip2(bb, i0, j0); idx[3] = skip2(cc, i0, j0);
            } else {
                long long tt = t - 2 * n12;
                int bb = (int)std::sqrt(2.0 * (double)tt);
                if (bb < 1) bb = 1;
                while ((long long)bb * (bb - 1) / 2 > tt) bb--;
                while ((long long)(bb + 1) * bb / 2 <= tt) bb++;
                int aa = (int)(tt - (long long)bb * (bb - 1) / 2);
                idx[0] = i0; idx[1] = j0; idx[2] = skip2(aa, i0, j0); idx[3] = skip2(bb, i0, j0);
            }
            sort4(idx);
            long long q = or_encode(idx[0], idx[1], idx[2], idx[3]);
            double g = g4[q];
            uint8_t m = mask[q];
            if (g == 0.0 || m == 0) continue;
            int ch = mask_channel(m);
            cplx fo = quartet_product(s, idx, ch);
            double sv[8];
            for (int p = 0; p < 4; p++) {
                int x = idx[p];
                const double* src = (x == i0) ? ni : (x == j0) ? nj : &s[2 * x];
                sv[2 * p] = src[0]; sv[2 * p + 1] = src[1];
            }
            int loc[4] = {0, 1, 2, 3};
            cplx fn = quartet_product(sv, loc, ch);
            dE -= g * (fn.re - fo.re);
        }
        if (or_metropolis(beta, dE, u[5])) {
            s[2 * i0] = ni[0]; s[2 * i0 + 1] = ni[1];
            s[2 * j0] = nj[0]; s[2 * j0 + 1] = nj[1];
            *E += dE;
            (*acc)++;
        }
    }
}

double or_cpu_baseline_sweeps(int N, int nrep, double* spins, const double* g2, int h2_active,
                              const double* g4, const uint8_t* mask, const double* betas,
                              const uint64_t* rseeds, uint64_t sweep0, int n_sweeps, double* E,
                              long long* acc, int n_threads) {
#ifdef _OPENMP
    if (n_threads > 0) omp_set_num_threads(n_threads);
#endif
    auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel for schedule(dynamic, 1)
    for (int r = 0; r < nrep; r++)
        for (int sw = 0; sw < n_sweeps; sw++)
            ref_sweep_one(N, spins + (size_t)2 * N * r, g2, h2_active, g4, mask, betas[r], rseeds[r],
                          sweep0 + sw, &E[r], &acc[r]);
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

// ---------------------------------------------------------------------------
// CPU parallel tempering with the reference's driver logic
// (src/parallel_tempering.cu:273-352): FP64 sweeps (ref_sweep_one, the
// reference's 3-type enumeration), exchange after every sweep in hot-to-cold
// order with a POSIX drand48 stream seeded (seed48 << 16 | 0x330E) kept
// locally (thread-safe, the same recurrence glibc uses).
// Measures, from sweep `measure_from` on, per temperature slot t:
//   e_sum[t], e2_sum[t]  (E/N, (E/N)^2 summed over replicas r of slot t)
//   q2_sum[t]            (|q|^2, q = sum_k a_k^(r0) conj(a_k^(r1)) / N, replicas 0,1 of the slot)
//   acc[t], prop[t]      (Metropolis acceptances / proposals made at slot t)
// and ex_acc[t], ex_prop[t] for t < NT-1.  *n_meas = measured sweeps.
// ---------------------------------------------------------------------------
static double lcg48_next(uint64_t* x) {
    *x = (0x5DEECE66DULL * *x + 0xBULL) & ((1ULL << 48) - 1);
    return (double)*x / 281474976710656.0;
}

// The same sweep with the quartet terms of each site listed once up front:
// for the pair (i0, j0) it visits the active quartets holding i0, then those
// holding j0 but not i0 -- the set the 3-type enumeration visits, minus the
// inactive ones it skips -- with the same per-term FP64 arithmetic
// (g (f_new - f_old)).  Only the order of the sum differs, so Delta E agrees
// with ref_sweep_one to rounding.  It lets the statistical comparisons run at
// N = 32 in seconds; ref_sweep_one stays the timed CPU baseline.
struct SiteTerms {
    std::vector<std::array<int, 4>> idx;
    std::vector<int> ch;
    std::vector<double> g;
    std::vector<std::vector<int>> of_site;
    SiteTerms(int N, const double* g4, const uint8_t* mask) : of_site(N) {
        long long q = 0;
        for (int l = 3; l < N; l++)
            for (int k = 2; k < l; k++)
                for (int j = 1; j < k; j++)
                    for (int i = 0; i < j; i++, q++) {
                        if (g4[q] == 0.0 || mask[q] == 0) continue;
                        const int t = (int)g.size();
                        idx.push_back({i, j, k, l});
                        ch.push_back(mask_channel(mask[q]));
                        g.push_back(g4[q]);
                        for (int x : {i, j, k, l}) of_site[x].push_back(t);
                    }
    }
};

static void lists_sweep_one(int N, double* s, const double* g2, int h2_active, const SiteTerms& L, double beta,
                            uint64_t rseed, uint64_t sweep, double* E, long long* acc) {
    for (int step = 0; step < N / 2; step++) {
        double u[6];
        or_step_uniforms(rseed, sweep, step, u);
        int i0, j0;
        or_pair_pick(N, u[0], u[1], &i0, &j0);
        double ai[2] = {s[2 * i0], s[2 * i0 + 1]}, aj[2] = {s[2 * j0], s[2 * j0 + 1]};
        double ni[2], nj[2];
        or_propose_sphere(ai, aj, u[2], u[3], u[4], ni, nj);
        cplx di = {ni[0] - ai[0], ni[1] - ai[1]}, dj = {nj[0] - aj[0], nj[1] - aj[1]};
        double dE = 0.0;
        if (h2_active) {
            cplx si = {0, 0}, sj = {0, 0};
            for (int k = 0; k < N; k++) {
                if (k == i0 || k == j0) continue;
                cplx akc = cconj(ld(s, k));
                si.re += g2[i0 * N + k] * akc.re; si.im += g2[i0 * N + k] * akc.im;
                sj.re += g2[j0 * N + k] * akc.re; sj.im += g2[j0 * N + k] * akc.im;
            }
            dE -= cmul(di, si).re + cmul(dj, sj).re;
            double gij = g2[i0 * N + j0];
            cplx op = cmul({ai[0], ai[1]}, cconj({aj[0], aj[1]}));
            cplx np = cmul({ni[0], ni[1]}, cconj({nj[0], nj[1]}));
            dE -= gij * (np.re - op.re);
        }
        for (int pass = 0; pass < 2; pass++)
            for (int t : L.of_site[pass == 0 ? i0 : j0]) {
                const std::array<int, 4>& q4 = L.idx[t];
                if (pass == 1 && (q4[0] == i0 || q4[1] == i0 || q4[2] == i0 || q4[3] == i0)) continue;
                int idx[4] = {q4[0], q4[1], q4[2], q4[3]};
                cplx fo = quartet_product(s, idx, L.ch[t]);
                double sv[8];
                for (int p = 0; p < 4; p++) {
                    int x = idx[p];
                    const double* src = (x == i0) ? ni : (x == j0) ? nj : &s[2 * x];
                    sv[2 * p] = src[0]; sv[2 * p + 1] = src[1];
                }
                int loc[4] = {0, 1, 2, 3};
                cplx fn = quartet_product(sv, loc, L.ch[t]);
                dE -= L.g[t] * (fn.re - fo.re);
            }
        if (or_metropolis(beta, dE, u[5])) {
            s[2 * i0] = ni[0]; s[2 * i0 + 1] = ni[1];
            s[2 * j0] = nj[0]; s[2 * j0 + 1] = nj[1];
            *E += dE;
            (*acc)++;
        }
    }
}

// The PT driver loop shared by the dense and sparse CPU chains: sweep(r,
// beta, sweep index, &E[r], &acc) advances replica r by one sweep; E holds
// the replicas' initial energies.
extern "C++" {
template <typename Sweep>
static void cpu_pt_loop(Sweep sweep, std::vector<double>& E, int N, int NT, int nr, double* spins,
                        const double* beta_sched, int n_sweeps, long seed48, int measure_from, double* e_sum,
                        double* e2_sum, double* q2_sum, long long* acc, long long* prop, long long* ex_acc,
                        long long* ex_prop, int* n_meas, int qbins, double* qh_sum);
}

static void cpu_pt_impl(bool lists, int N, int NT, int nr, double* spins, const double* g2, int h2_active,
                        const double* g4, const uint8_t* mask, const double* beta_sched, const uint64_t* rseeds,
                        int n_sweeps, long seed48, int measure_from, double* e_sum, double* e2_sum, double* q2_sum,
                        long long* acc, long long* prop, long long* ex_acc, long long* ex_prop, int* n_meas, int qbins,
                        double* qh_sum);

void or_cpu_pt(int N, int NT, int nr, double* spins, const double* g2, int h2_active, const double* g4,
               const uint8_t* mask, const double* beta_sched, const uint64_t* rseeds, int n_sweeps, long seed48,
               int measure_from, double* e_sum, double* e2_sum, double* q2_sum, long long* acc, long long* prop,
               long long* ex_acc, long long* ex_prop, int* n_meas, int qbins, double* qh_sum) {
    cpu_pt_impl(false, N, NT, nr, spins, g2, h2_active, g4, mask, beta_sched, rseeds, n_sweeps, seed48, measure_from,
                e_sum, e2_sum, q2_sum, acc, prop, ex_acc, ex_prop, n_meas, qbins, qh_sum);
}

// or_cpu_pt with lists_sweep_one in place of ref_sweep_one.
void or_cpu_pt_lists(int N, int NT, int nr, double* spins, const double* g2, int h2_active, const double* g4,
                     const uint8_t* mask, const double* beta_sched, const uint64_t* rseeds, int n_sweeps, long seed48,
                     int measure_from, double* e_sum, double* e2_sum, double* q2_sum, long long* acc, long long* prop,
                     long long* ex_acc, long long* ex_prop, int* n_meas, int qbins, double* qh_sum) {
    cpu_pt_impl(true, N, NT, nr, spins, g2, h2_active, g4, mask, beta_sched, rseeds, n_sweeps, seed48, measure_from,
                e_sum, e2_sum, q2_sum, acc, prop, ex_acc, ex_prop, n_meas, qbins, qh_sum);
}

static void cpu_pt_impl(bool lists, int N, int NT, int nr, double* spins, const double* g2, int h2_active,
                        const double* g4, const uint8_t* mask, const double* beta_sched, const uint64_t* rseeds,
                        int n_sweeps, long seed48, int measure_from, double* e_sum, double* e2_sum, double* q2_sum,
                        long long* acc, long long* prop, long long* ex_acc, long long* ex_prop, int* n_meas, int qbins,
                        double* qh_sum) {
    const int R = NT * nr;
    std::unique_ptr<SiteTerms> L;
    if (lists) L.reset(new SiteTerms(N, g4, mask));
    std::vector<double> E(R);
    for (int r = 0; r < R; r++) {
        double e2, e4;
        or_energy(N, spins + (size_t)2 * N * r, g2, h2_active, g4, mask, &e2, &e4);
        E[r] = e2 + e4;
    }
    auto sweep = [&](int r, double beta, int sw, double* Er, long long* a) {
        if (lists)
            lists_sweep_one(N, spins + (size_t)2 * N * r, g2, h2_active, *L, beta, rseeds[r], (uint64_t)sw, Er, a);
        else
            ref_sweep_one(N, spins + (size_t)2 * N * r, g2, h2_active, g4, mask, beta, rseeds[r], (uint64_t)sw, Er,
                          a);
    };
    cpu_pt_loop(sweep, E, N, NT, nr, spins, beta_sched, n_sweeps, seed48, measure_from, e_sum, e2_sum, q2_sum, acc,
                prop, ex_acc, ex_prop, n_meas, qbins, qh_sum);
}

extern "C++" {
template <typename Sweep>
static void cpu_pt_loop(Sweep sweep, std::vector<double>& E, int N, int NT, int nr, double* spins,
                        const double* beta_sched, int n_sweeps, long seed48, int measure_from, double* e_sum,
                        double* e2_sum, double* q2_sum, long long* acc, long long* prop, long long* ex_acc,
                        long long* ex_prop, int* n_meas, int qbins, double* qh_sum) {
    const int R = NT * nr;
    std::vector<int> perm(R), slot(R);
    for (int i = 0; i < R; i++) perm[i] = i;
    uint64_t X = (((uint64_t)(uint32_t)seed48) << 16) | 0x330EULL;
    *n_meas = 0;
    for (int sw = 0; sw < n_sweeps; sw++) {
        for (int t = 0; t < NT; t++)
            for (int k = 0; k < nr; k++) slot[perm[t * nr + k]] = t;
        for (int r = 0; r < R; r++) {
            long long a = 0;
            const int t = slot[r];
            sweep(r, beta_sched[t], sw, &E[r], &a);
            if (sw >= measure_from) {
                acc[t] += a;
                prop[t] += N / 2;
            }
        }
        for (int t = 0; t < NT - 1; t++)
            for (int k = 0; k < nr; k++) {
                const int pa = perm[t * nr + k], pb = perm[(t + 1) * nr + k];
                const double d = (beta_sched[t] - beta_sched[t + 1]) * (E[pa] - E[pb]);
                bool ok = d >= 0.0;
                if (!ok) ok = lcg48_next(&X) < std::exp(d);
                if (sw >= measure_from) ex_prop[t]++;
                if (ok) {
                    if (sw >= measure_from) ex_acc[t]++;
                    std::swap(perm[t * nr + k], perm[(t + 1) * nr + k]);
                }
            }
        if (sw >= measure_from) {
            for (int t = 0; t < NT; t++) {
                for (int k = 0; k < nr; k++) {
                    const double e = E[perm[t * nr + k]] / N;
                    e_sum[t] += e;
                    e2_sum[t] += e * e;
                }
                if (nr >= 2) {
                    const double* a0 = spins + (size_t)2 * N * perm[t * nr];
                    const double* a1 = spins + (size_t)2 * N * perm[t * nr + 1];
                    double qr = 0, qi = 0;
                    for (int i = 0; i < N; i++) {
                        qr += a0[2 * i] * a1[2 * i] + a0[2 * i + 1] * a1[2 * i + 1];
                        qi += a0[2 * i + 1] * a1[2 * i] - a0[2 * i] * a1[2 * i + 1];
                    }
                    const double q2 = (qr * qr + qi * qi) / ((double)N * N);
                    q2_sum[t] += q2;
                    if (qbins > 0) {  // histogram of |q| over [0, 1] (|q| <= 1 on the sphere)
                        const int b = std::min(qbins - 1, (int)(std::sqrt(q2) * qbins));
                        qh_sum[(size_t)t * qbins + b] += 1.0;
                    }
                }
            }
            (*n_meas)++;
        }
    }
}
}  // extern "C++"

// ---------------------------------------------------------------------------
// Sparse (smoothed-cube) chains (src/mc_sparse.cu:63-253) in the same PT
// driver loop (src/parallel_tempering_sparse.cu runs the dense driver's
// exchange): the cube move (include/spins_sparse.h:26-56) from the step's
// uniforms, H2 as in the dense sweep, H4 over the selected quartets holding
// i0 or j0 (i0's list, then j0's without those holding i0) as g (f_new -
// f_old) in FP64 -- the reference's full and differential products, equal
// up to rounding -- and Metropolis with exp.
// ---------------------------------------------------------------------------
struct SparseTerms {
    const int* q;
    const double* g;
    std::vector<std::vector<int>> of_site;
    SparseTerms(int N, int nq, const int* quartets, const double* gq) : q(quartets), g(gq), of_site(N) {
        for (int t = 0; t < nq; t++)
            for (int p = 0; p < 4; p++) {
                std::vector<int>& v = of_site[quartets[5 * t + p]];
                if (v.empty() || v.back() != t) v.push_back(t);
            }
    }
};

static void sparse_sweep_one(int N, double* s, const double* g2, int h2_active, const SparseTerms& L, double beta,
                             uint64_t rseed, uint64_t sweep, double* E, long long* acc) {
    for (int step = 0; step < N / 2; step++) {
        double u[6];
        or_step_uniforms(rseed, sweep, step, u);
        int i0, j0;
        or_pair_pick(N, u[0], u[1], &i0, &j0);
        double ai[2] = {s[2 * i0], s[2 * i0 + 1]}, aj[2] = {s[2 * j0], s[2 * j0 + 1]};
        double ni[2], nj[2];
        or_propose_cube(ai, aj, u[2], u[3], u[4], ni, nj);
        cplx di = {ni[0] - ai[0], ni[1] - ai[1]}, dj = {nj[0] - aj[0], nj[1] - aj[1]};
        double dE = 0.0;
        if (h2_active) {
            cplx si = {0, 0}, sj = {0, 0};
            for (int k = 0; k < N; k++) {
                if (k == i0 || k == j0) continue;
                cplx akc = cconj(ld(s, k));
                si.re += g2[i0 * N + k] * akc.re; si.im += g2[i0 * N + k] * akc.im;
                sj.re += g2[j0 * N + k] * akc.re; sj.im += g2[j0 * N + k] * akc.im;
            }
            dE -= cmul(di, si).re + cmul(dj, sj).re;
            double gij = g2[i0 * N + j0];
            cplx op = cmul({ai[0], ai[1]}, cconj({aj[0], aj[1]}));
            cplx np = cmul({ni[0], ni[1]}, cconj({nj[0], nj[1]}));
            dE -= gij * (np.re - op.re);
        }
        for (int pass = 0; pass < 2; pass++)
            for (int t : L.of_site[pass == 0 ? i0 : j0]) {
                const int* v = L.q + 5 * t;
                if (pass == 1 && (v[0] == i0 || v[1] == i0 || v[2] == i0 || v[3] == i0)) continue;
                int idx[4] = {v[0], v[1], v[2], v[3]};
                cplx fo = quartet_product(s, idx, v[4]);
                double sv[8];
                for (int p = 0; p < 4; p++) {
                    int x = idx[p];
                    const double* src = (x == i0) ? ni : (x == j0) ? nj : &s[2 * x];
                    sv[2 * p] = src[0]; sv[2 * p + 1] = src[1];
                }
                int loc[4] = {0, 1, 2, 3};
                cplx fn = quartet_product(sv, loc, v[4]);
                dE -= L.g[t] * (fn.re - fo.re);
            }
        if (or_metropolis(beta, dE, u[5])) {
            s[2 * i0] = ni[0]; s[2 * i0 + 1] = ni[1];
            s[2 * j0] = nj[0]; s[2 * j0 + 1] = nj[1];
            *E += dE;
            (*acc)++;
        }
    }
}

// CPU PT of a sparse system: the measurements of or_cpu_pt.
void or_cpu_pt_sparse(int N, int NT, int nr, double* spins, const double* g2, int h2_active, int nq,
                      const int* quartets, const double* gq, const double* beta_sched, const uint64_t* rseeds,
                      int n_sweeps, long seed48, int measure_from, double* e_sum, double* e2_sum, double* q2_sum,
                      long long* acc, long long* prop, long long* ex_acc, long long* ex_prop, int* n_meas, int qbins,
                      double* qh_sum) {
    const int R = NT * nr;
    const SparseTerms L(N, nq, quartets, gq);
    std::vector<double> E(R);
    for (int r = 0; r < R; r++) {
        const double* sr = spins + (size_t)2 * N * r;
        E[r] = (h2_active ? or_energy_h2(N, sr, g2) : 0.0) + or_energy_h4_sparse(sr, nq, quartets, gq);
    }
    auto sweep = [&](int r, double beta, int sw, double* Er, long long* a) {
        sparse_sweep_one(N, spins + (size_t)2 * N * r, g2, h2_active, L, beta, rseeds[r], (uint64_t)sw, Er, a);
    };
    cpu_pt_loop(sweep, E, N, NT, nr, spins, beta_sched, n_sweeps, seed48, measure_from, e_sum, e2_sum, q2_sum, acc,
                prop, ex_acc, ex_prop, n_meas, qbins, qh_sum);
}

// ---------------------------------------------------------------------------
// Learned global move (this build's own design; the reference has only the
// pair rotation).  Restated here so the GPU move can be replayed and its
// stationary law checked on the CPU.
//
//  gauge      u = a c, c = conj(a_k*)/|a_k*|, k* = argmax |a_k|^2 (first)
//  proposal   z = mu + sigma xi in the frame, y = z/|z|, a' = sqrt(N) y conj(c)
//             mu = u + m O[0:2N], log sigma = clamp(O[2N:4N] + ls0) + log_scale
//  density    projected Gaussian on the unit sphere (w.r.t. surface measure)
//             p(y) = (2pi)^(-d/2) prod sigma^-1 e^(-C/2) A^(-d/2) J_{d-1}(B/sqrt A)
//  MH         log r = -beta (E' - E) + log w(a') - log w(a) + log q(a|a') - log q(a'|a),
//             w = prod_k |a_k|^-1 (the pair-rotation chain's stationary density)
//
// log J_n(alpha) = log int_0^inf s^n exp(-s^2/2 + alpha s) ds by the
// trapezoid rule in t = log s (8 nodes per Laplace width in t) -- a
// different variable and window from the device code's trapezoid in s.  Pinned against the closed form at alpha = 0, the
// three-term recurrence J_n = alpha J_{n-1} + (n-1) J_{n-2} for alpha >= 0 and
// adaptive quadrature for alpha < 0 (tests/test_oracle_learned.py).
// ---------------------------------------------------------------------------
double or_log_jn(int n, double alpha) {
    // s = e^t: J_n = int exp(g(t)) dt, g(t) = (n+1) t - e^(2t)/2 + alpha e^t,
    // smooth and decaying on both sides, so the trapezoid rule converges
    // exponentially for every n >= 0.
    const double m = (double)n + 1.0;
    const double sq = std::sqrt(alpha * alpha + 4.0 * m);
    const double so = alpha >= 0.0 ? 0.5 * (alpha + sq) : 2.0 * m / (sq - alpha);  // e^t at the maximum
    const double to = std::log(so);
    const double go = m * to - 0.5 * so * so + alpha * so;
    const double wt = 1.0 / std::sqrt(so * (2.0 * so - alpha));
    const double h = wt / 8.0;
    const double left = std::max(40.0 * wt, 110.0 / m), right = 40.0 * wt;
    const long long jl = (long long)std::ceil(left / h), jr = (long long)std::ceil(right / h);
    double sum = 0.0;
    for (long long j = -jl; j <= jr; j++) {
        const double t = to + h * (double)j;
        const double e = std::exp(t);
        sum += std::exp(m * t - 0.5 * e * e + alpha * e - go);
    }
    return go + std::log(sum * h);
}

double or_lp_log_density(int d, const double* y, const double* mu, const double* ls) {
    double A = 0.0, B = 0.0, C = 0.0, L = 0.0;
    for (int i = 0; i < d; i++) {
        const double iv = std::exp(-2.0 * ls[i]);
        A += y[i] * y[i] * iv;
        B += y[i] * mu[i] * iv;
        C += mu[i] * mu[i] * iv;
        L += ls[i];
    }
    const double hd = 0.5 * (double)d;
    return -hd * std::log(2.0 * M_PI) - L - 0.5 * C - hd * std::log(A) + or_log_jn(d - 1, B / std::sqrt(A));
}

int or_lp_gauge(int N, const double* a, double* c, double* u) {
    int bk = 0;
    double best = -1.0;
    for (int k = 0; k < N; k++) {
        const double x = a[2 * k] * a[2 * k] + a[2 * k + 1] * a[2 * k + 1];
        if (x > best) {
            best = x;
            bk = k;
        }
    }
    const double r = std::sqrt(best);
    c[0] = a[2 * bk] / r;
    c[1] = -a[2 * bk + 1] / r;
    if (u)
        for (int k = 0; k < N; k++) {
            u[2 * k] = a[2 * k] * c[0] - a[2 * k + 1] * c[1];
            u[2 * k + 1] = a[2 * k] * c[1] + a[2 * k + 1] * c[0];
        }
    return bk;
}

// Normals xi[0..2N) and the accept uniform of move `mv` of a replica.
void or_lp_draws(uint64_t seed, uint64_t mv, int N, double* xi, double* u_acc) {
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int b = 0; b < N; b++) {
        uint32_t ctr[4] = {(uint32_t)mv, (uint32_t)(mv >> 32), (uint32_t)b, 32u}, x[4];
        or_philox4x32_10(ctr, key, x);
        const double u1 = or_u64_to_unit(x[0], x[1]), u2 = or_u64_to_unit(x[2], x[3]);
        const double rad = std::sqrt(-2.0 * std::log(u1));
        xi[2 * b] = rad * std::cos(2.0 * M_PI * u2);
        xi[2 * b + 1] = rad * std::sin(2.0 * M_PI * u2);
    }
    if (u_acc) {
        uint32_t ctr[4] = {(uint32_t)mv, (uint32_t)(mv >> 32), 0xFFFFFFFFu, 32u}, x[4];
        or_philox4x32_10(ctr, key, x);
        *u_acc = or_u64_to_unit(x[0], x[1]);
    }
}

static void lp_mu_ls(int N, const double* u, const float* O, double ls0, double ls_min, double ls_max,
                     double log_scale, double mean_scale, double* mu, double* ls) {
    const int d = 2 * N;
    for (int i = 0; i < d; i++) {
        mu[i] = u[i] + mean_scale * (double)O[i];
        ls[i] = std::min(std::max((double)O[d + i] + ls0, ls_min), ls_max) + log_scale;
    }
}

// Proposal of move `mv` from a [N][2] given the network output O_f [4N] at
// a: writes prop [N][2], returns log q(a'|a).
double or_lp_propose(int N, const double* a, const float* O_f, double ls0, double ls_min, double ls_max,
                     double log_scale, double mean_scale, uint64_t seed, uint64_t mv, double* prop) {
    const int d = 2 * N;
    std::vector<double> u(d), mu(d), ls(d), xi(d), z(d);
    double c[2];
    or_lp_gauge(N, a, c, u.data());
    lp_mu_ls(N, u.data(), O_f, ls0, ls_min, ls_max, log_scale, mean_scale, mu.data(), ls.data());
    or_lp_draws(seed, mv, N, xi.data(), nullptr);
    double z2 = 0.0;
    for (int i = 0; i < d; i++) {
        z[i] = mu[i] + std::exp(ls[i]) * xi[i];
        z2 += z[i] * z[i];
    }
    const double inv = 1.0 / std::sqrt(z2);
    for (int i = 0; i < d; i++) z[i] *= inv;
    const double lq = or_lp_log_density(d, z.data(), mu.data(), ls.data());
    const double rn = std::sqrt((double)N);
    for (int k = 0; k < N; k++) {
        const double re = z[2 * k] * rn, im = z[2 * k + 1] * rn;
        prop[2 * k] = re * c[0] + im * c[1];
        prop[2 * k + 1] = im * c[0] - re * c[1];
    }
    return lq;
}

// MH decision of move `mv`: out[8] = log q(a'|a), log q(a|a'), E, E', log r,
// accepted, log w(a), log w(a').  O_r is the network output at the proposal.
int or_lp_accept(int N, double beta, const double* a, const double* prop, double E, double E_new,
                 const float* O_r, double ls0, double ls_min, double ls_max, double log_scale,
                 double mean_scale, double lq_f,
                 uint64_t seed, uint64_t mv, double* out) {
    const int d = 2 * N;
    std::vector<double> up(d), mu(d), ls(d), v(d);
    double c[2];
    or_lp_gauge(N, prop, c, up.data());
    lp_mu_ls(N, up.data(), O_r, ls0, ls_min, ls_max, log_scale, mean_scale, mu.data(), ls.data());
    double v2 = 0.0, lw = 0.0, lwp = 0.0;
    for (int k = 0; k < N; k++) {
        v[2 * k] = a[2 * k] * c[0] - a[2 * k + 1] * c[1];
        v[2 * k + 1] = a[2 * k] * c[1] + a[2 * k + 1] * c[0];
        v2 += v[2 * k] * v[2 * k] + v[2 * k + 1] * v[2 * k + 1];
        lw -= 0.5 * std::log(a[2 * k] * a[2 * k] + a[2 * k + 1] * a[2 * k + 1]);
        lwp -= 0.5 * std::log(prop[2 * k] * prop[2 * k] + prop[2 * k + 1] * prop[2 * k + 1]);
    }
    const double inv = 1.0 / std::sqrt(v2);
    for (int i = 0; i < d; i++) v[i] *= inv;
    const double lq_r = or_lp_log_density(d, v.data(), mu.data(), ls.data());
    const double logr = -beta * (E_new - E) + (lwp - lw) + (lq_r - lq_f);
    double u;
    std::vector<double> xi(d);
    or_lp_draws(seed, mv, N, xi.data(), &u);
    const int acc = std::log(u) < logr;
    out[0] = lq_f;
    out[1] = lq_r;
    out[2] = E;
    out[3] = E_new;
    out[4] = logr;
    out[5] = acc;
    out[6] = lw;
    out[7] = lwp;
    return acc;
}

int or_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
