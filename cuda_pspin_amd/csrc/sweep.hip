// sweep.hip -- fused Monte Carlo sweep on gfx950 (the reference's
// mc_sweep_kernel, src/mc.cu:66-318, and mc_sparse_sweep_kernel,
// src/mc_sparse.cu:63-253, re-designed).
//
// One workgroup owns one replica and makes N/2 pair proposals per sweep.
//
// Delta E by the midpoint rule.  Every Hamiltonian term is a real
// multilinear form in the amplitudes of distinct sites, so for a move
// (a_i, a_j) -> (a_i', a_j') with d_x = a_x' - a_x and midpoints
// m_x = (a_x + a_x')/2 it is exact that
//     Delta E = -Re[d_i F_i(s~)] - Re[d_j F_j(s~)],
// where s~ is the configuration with a_i, a_j replaced by their midpoints and
// F_x is the local field of site x (the coefficient of a_x in -E):
//     F_x = sum_{k != x} g2_xk conj(s~_k) + sum_{q containing x} g_q X_{x,q}(s~).
// Writing each H4 term as -g Re[a_x X] with X = a_p conj(a_m) conj(a_n)
// (p = the position sharing x's conjugation in the channel mask of
// src/mc.cu:119-121, m, n the other two) covers the reference's three
// enumeration types (src/mc.cu:111-275) with no special case for quartets
// that contain both sites, and in differential form, so FP32 arithmetic does
// not cancel new-minus-old products.
//
// Three coupling layouts feed the same step machinery:
//  LISTS  site x owns a contiguous list of (p | m<<8 | n<<16, g) entries
//         (any channel pattern: FMC, comb, sparse quartets); coalesced 8-byte
//         loads, the three amplitudes from LDS.
//  PAIRS  fully connected (every quartet channel 0): site x owns dense FP32
//         tiles indexed by (row y, pair (b,c)) over the other sites, each
//         triple a < b < c of them once, at row y = the site of a.  Pairs are
//         ordered by b descending and cut into blocks of 256; block k's tile
//         holds rows 0..b_max(k) with zeros where a >= b.  A lane owns 4
//         consecutive pairs and walks rows with one 16-byte load per row,
//         accumulating W_bc = sum_y g_ybc s~_y against the broadcast row
//         amplitude s~_y (two rows per 16-byte LDS read).  No predicate or
//         index arithmetic per row: 4 bytes and 2 FMA per coupling.  The local
//         field is then sum_bc U_bc W_bc (partner a) or U_bc conj(W_bc)
//         (partner c) with U_bc a product of two amplitudes, from a per-site
//         table of LDS offsets and partner flags.  A static list of row
//         segments balances the waves.
//  GROUPS dense, any channel pattern (FMC, comb): site x's entries grouped by
//         their conjugated pair {m, n} in chunks of up to four partners, so a
//         chunk adds conj(s~_m s~_n) sum_k g_k s~_pk to the field of x: one
//         complex product per chunk, one LDS read and 2 FMA per entry.
//
// Randomness: Philox4x32-10 keyed by the replica seed (the reference's
// splitmix64 chain, src/mc.cu:418-423) and countered by (sweep, step), all N/2
// steps drawn in parallel at sweep start.  Proposal arithmetic and the master
// amplitudes are FP64 (the constraint is kept to FP64 rounding, as in
// the reference); Delta E is FP32; the running energy is FP64.
#include "exchange_dev.h"

namespace ps {

// Default bound on a cross-workgroup wait (s_sleep 2 polls, about a second).
constexpr unsigned kBarrierSpinLimit = 1u << 23;
// What a workgroup leaves in its beta slot before arriving at the grid
// barrier: a NaN bit pattern, never a beta (betas are finite and >= 0).
constexpr unsigned long long kBetaMarker = 0x7FF8BE7A0DDC0DE5ull;
// Temperatures a tagged launch keeps in LDS (a new level is looked up there).
constexpr int kTagMaxNT = 64;

struct SweepArgs {
    int N, nrep, h2_active, n_sweeps;
    int rounds;             // PT rounds of n_sweeps sweeps + exchange (> 1: cooperative launch, grid barrier)
    int g2_lds;             // the FP32 two-body matrix is staged in dynamic LDS
    int pair_groups;        // 0: pairs from the replica stream; G > 0: from group (rep % G)'s stream
    uint64_t pair_seed;     // run seed the group keys derive from
    double* spins;
    const float* g2;
    const int* site_off;
    const uint2* site_ent;
    const double* betas;
    double* E;
    long long* acc;
    long long* prop;
    const uint64_t* seeds;
    uint64_t* ctr;  // [nrep] per-replica sweep counter (Philox counter), advanced by the kernel
    StepRec* log;
    int log_cap;   // sweeps per replica in the log
    int log_base;  // sweeps already logged
    // PAIRS layout
    const float* ptab;        // [N][slice_len] per-site pair tiles
    const uint32_t* pinfo;    // [N][npad] per-site pair amplitude offsets and partner flags
    long long slice_len;
    int npad;
    // replica exchange by the last workgroup to finish (exchange != 0)
    int exchange;
    unsigned int* done;     // [0] workgroups arrived, [1] unused, [2] barrier wait expired
    ExArgs ex;
    // cooperative launches with NT <= kTagMaxNT: workgroup 0 exchanges, and
    // energies and new levels move as tagged words (tag of round rd: tag0 + rd + 1)
    int tagged;
    unsigned tag0;
    unsigned spin_limit;      // polls of a cross-workgroup wait before it is flagged
    unsigned long long* ew;   // [2 nrep] energy halves {tag, half}
    unsigned long long* lvw;  // [nrep] new level {tag, level}
    const int4* tasks;        // row batches {slot | first << 1 | last << 2 | rows << 8, first pair, first row, tile offset}
    const int* wave_tasks;    // [nwaves+1] batch ranges per wave (at most 64 each)
    // GROUPS layout
    const int* grp_off;       // [N+1] chunk offsets per site
    const uint2* grp_hdr;     // [chunks] {m | n << 8, p0 | p1 << 8 | p2 << 16 | p3 << 24}
    const uint4* grp_g;       // [chunks] four coupling words
    int grp_zero;             // index of the all-zero chunk after the last
};

// Coupling layouts of the sweep kernel.
constexpr int kLists = 0, kPairs = 1, kGroups = 2;

__device__ __forceinline__ float2 cmulf(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Sum over the 64 lanes of a wave with DPP moves (no LDS round trips): pairs
// and quads by quad_perm, rows of 16 by row_ror 4 and 8, then row_bcast:15
// into rows 1 and 3 and row_bcast:31 into rows 2 and 3.  Lane 63 holds the
// total; it is read back as a wave-uniform value.
__device__ __forceinline__ float wave_sum(float v) {
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xb1, 0xf, 0xf, false));  // quad_perm 1,0,3,2
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4e, 0xf, 0xf, false));  // quad_perm 2,3,0,1
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x124, 0xf, 0xf, false)); // row_ror 4
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x128, 0xf, 0xf, false)); // row_ror 8
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x142, 0xa, 0xf, false)); // row_bcast:15
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x143, 0xc, 0xf, false)); // row_bcast:31
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// The pair of a step from two uniform words (src/mc.cu:129-141).
__device__ __forceinline__ void pick_pair(uint32_t a, uint32_t b, int N, int& i0, int& j0) {
    const double u1 = u32_to_unit(a), u2 = u32_to_unit(b);
    int ri = (int)(u1 * N);
    if (ri >= N) ri = N - 1;
    int rj = (int)(u2 * (N - 1));
    if (rj >= N - 1) rj = N - 2;
    if (rj >= ri) rj++;
    i0 = ri < rj ? ri : rj;
    j0 = ri < rj ? rj : ri;
}

// Draws and trigonometry for one step (pair pick src/mc.cu:129-141; angles
// include/spins.h:25-32 or include/spins_sparse.h:34-41).
// With a shared pair schedule (gkey != 0) the two pair uniforms come from
// the group's stream; everything else stays per replica.
template <bool CUBE>
__device__ __forceinline__ StepPre make_step(uint64_t seed, uint64_t gkey, uint64_t sweep, int step, int N,
                                             double beta) {
    uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    U4 x = philox4x32_10(U4{(uint32_t)sweep, (uint32_t)(sweep >> 32), (uint32_t)step, kCtrStep0}, k0, k1);
    U4 y = philox4x32_10(U4{(uint32_t)sweep, (uint32_t)(sweep >> 32), (uint32_t)step, kCtrStep1}, k0, k1);
    if (gkey) {
        const U4 z = philox4x32_10(U4{(uint32_t)sweep, (uint32_t)(sweep >> 32), (uint32_t)step, kCtrPair},
                                   (uint32_t)gkey, (uint32_t)(gkey >> 32));
        x.x = z.x;
        x.y = z.y;
    }
    StepPre p;
    pick_pair(x.x, x.y, N, p.i0, p.j0);
    double ua = u32_to_unit(x.z), up1 = u32_to_unit(x.w), up2 = u32_to_unit(y.x), uacc = u32_to_unit(y.y);
    // angles 2 pi u with u in (0, 1): sincospi needs no large-argument reduction
    if (CUBE) {
        p.c[0] = sqrt(1.0 - ua);
        p.c[1] = sqrt(ua);
    } else {
        sincospi(2.0 * ua, &p.c[1], &p.c[0]);
    }
    sincospi(2.0 * up1, &p.c[3], &p.c[2]);
    sincospi(2.0 * up2, &p.c[5], &p.c[4]);
    // u < exp(-beta dE)  <=>  dE < -log(u)/beta  (src/mc.cu:289-296)
    p.thr = (beta == 0.0) ? INFINITY : -log(uacc) / beta;
    return p;
}

// make_step split over five waves (each wave's lanes take the steps): part
// 0 the pair from the replica stream (unless gkey != 0) and the amplitude
// split, parts 1 and 2 the two phases, part 3 the threshold, part 4 the
// pair from the group stream (gkey != 0).  Every field gets the value
// make_step computes, by the same operations; a wave's latency is one
// Philox block and one trigonometric function instead of all of them.
// With defer, part 3 stores -log(u) and the caller divides by beta later
// (thr_fixup), the same two operations.
template <bool CUBE>
__device__ __noinline__ void make_step_part(StepPre& p, int part, uint64_t seed, uint64_t gkey, uint64_t sweep,
                                               int step, int N, double beta, bool defer) {
    const U4 ctr0 = U4{(uint32_t)sweep, (uint32_t)(sweep >> 32), (uint32_t)step, kCtrStep0};
    const U4 ctr1 = U4{(uint32_t)sweep, (uint32_t)(sweep >> 32), (uint32_t)step, kCtrStep1};
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    if (part == 4) {
        if (gkey) {
            const U4 z = philox4x32_10(U4{(uint32_t)sweep, (uint32_t)(sweep >> 32), (uint32_t)step, kCtrPair},
                                       (uint32_t)gkey, (uint32_t)(gkey >> 32));
            pick_pair(z.x, z.y, N, p.i0, p.j0);
        }
    } else if (part <= 1) {
        const U4 x = philox4x32_10(ctr0, k0, k1);
        if (part == 0) {
            if (!gkey) pick_pair(x.x, x.y, N, p.i0, p.j0);
            const double ua = u32_to_unit(x.z);
            if (CUBE) {
                p.c[0] = sqrt(1.0 - ua);
                p.c[1] = sqrt(ua);
            } else {
                sincospi(2.0 * ua, &p.c[1], &p.c[0]);
            }
        } else {
            sincospi(2.0 * u32_to_unit(x.w), &p.c[3], &p.c[2]);
        }
    } else {
        const U4 y = philox4x32_10(ctr1, k0, k1);
        if (part == 2) sincospi(2.0 * u32_to_unit(y.x), &p.c[5], &p.c[4]);
        else if (defer) p.thr = -log(u32_to_unit(y.y));
        else p.thr = (beta == 0.0) ? INFINITY : -log(u32_to_unit(y.y)) / beta;
    }
}

// The threshold of a step drawn with defer, once beta is known.
__device__ __forceinline__ void thr_fixup(StepPre& p, double beta) {
    p.thr = (beta == 0.0) ? INFINITY : p.thr / beta;
}

// Local-field contribution of the thread for the current step, LISTS layout.
// The two site lists are walked as one range; each thread issues LB (8 or
// 16) independent 8-byte loads before using any, so enough bytes are in
// flight to cover L2/MALL latency.  Entries past the end read as g = 0.

// Coupling word of a list entry: FP32, or a BF16 hi/lo pair (hi in the low
// half) whose sum carries ~16 mantissa bits; accumulation stays FP32.
template <bool BF16>
__device__ __forceinline__ float coupling(uint32_t w) {
    if (!BF16) return __uint_as_float(w);
    return __uint_as_float(w << 16) + __uint_as_float(w & 0xffff0000u);
}

template <bool BF16, int kListBatch>
__device__ __forceinline__ float field_lists(const SweepArgs& a, const float2* s32, int i0, int j0, float2 di,
                                             float2 dj, int tid, int bdim) {
    float2 Fi = make_float2(0.f, 0.f), Fj = make_float2(0.f, 0.f);
    const int oi = a.site_off[i0], ni_ = a.site_off[i0 + 1] - oi;
    const int oj = a.site_off[j0], nj_ = a.site_off[j0 + 1] - oj;
    const int n_tot = ni_ + nj_;
    for (int e0 = tid; e0 < n_tot; e0 += kListBatch * bdim) {
        uint2 en[kListBatch];
#pragma unroll
        for (int u = 0; u < kListBatch; u++) {
            const int e = e0 + u * bdim;
            en[u] = (e < n_tot) ? a.site_ent[e < ni_ ? oi + e : oj + (e - ni_)] : make_uint2(0u, 0u);
        }
#pragma unroll
        for (int u = 0; u < kListBatch; u++) {
            const float2 tp = s32[en[u].x & 0xff], tm = s32[(en[u].x >> 8) & 0xff], tn = s32[(en[u].x >> 16) & 0xff];
            const float g = coupling<BF16>(en[u].y);
            const float2 mn = cmulf(tm, tn);
            // g X with X = tp * conj(tm tn)
            const float xr = g * (tp.x * mn.x + tp.y * mn.y), xi = g * (tp.y * mn.x - tp.x * mn.y);
            const bool isj = e0 + u * bdim >= ni_;
            Fi.x += isj ? 0.f : xr;
            Fi.y += isj ? 0.f : xi;
            Fj.x += isj ? xr : 0.f;
            Fj.y += isj ? xi : 0.f;
        }
    }
    return -(di.x * Fi.x - di.y * Fi.y) - (dj.x * Fj.x - dj.y * Fj.y);
}

// Local-field contribution of the thread for the current step, PAIRS layout
// (fully connected, channel 0: f = a_p0 a_p1 conj(a_p2) conj(a_p3)).
// For site x and others a<b<c the conjugation partner of x is a when
// o(b) > x, else c.  With z = conj(s_c) W the pair's term is conj(s_b z') with
// z' = conj(z) for partner a and z otherwise, so a segment accumulates
// Hc = sum s_b z' and contributes -Re(d conj(Hc)).
//
// A wave's work is a static list of batches of at most PB rows (a multiple
// of 4), descriptor
//   {slot | first << 1 | last << 2 | rows << 8 | lanes(quad q) << (12 + 7 q),
//    first pair, first row, tile offset},
// held one per lane.  lanes(q) is the number of lanes (a prefix) that hold a
// pair of the quad's rows; only they load and multiply it.  Loads run one batch ahead of the arithmetic in one
// register buffer: as soon as a quad of rows has been multiplied, its
// registers take the next batch's loads for the same quad.  The last batch of
// a step prefetches the next step's first batch (its sites are drawn at sweep
// start), so those loads are in flight across the barrier.
template <int PB>
struct PairBuf {
    float4 g[PB];
    uint4 inf;  // pair-info words of the buffered batch (first batch of a segment)
};

__device__ __forceinline__ int4 desc_at(int4 d, int b) {
    return make_int4(__builtin_amdgcn_readlane(d.x, b), __builtin_amdgcn_readlane(d.y, b),
                     __builtin_amdgcn_readlane(d.z, b), __builtin_amdgcn_readlane(d.w, b));
}

__device__ __forceinline__ int quad_lanes(int x, int q) { return (x >> (12 + 7 * q)) & 127; }

// Loads of rows [u0, u0 + 4) of batch bd for the given pair site (lanes
// holding pairs of those rows only; the others neither load nor multiply).
template <int PB>
__device__ __forceinline__ void pair_load_quad(const SweepArgs& a, PairBuf<PB>& B, int4 bd, int site, int lane, int u0) {
    if (lane < quad_lanes(bd.x, u0 / 4)) {
        const float4* gp = (const float4*)(a.ptab + (size_t)site * a.slice_len + bd.w) + lane;
#pragma unroll
        for (int v = 0; v < 4; v++) B.g[u0 + v] = gp[(u0 + v) * 64];
    }
}

template <int PB>
__device__ __forceinline__ void pair_issue(const SweepArgs& a, PairBuf<PB>& B, int4 bd, int i0, int j0, int lane) {
    const int nr = (bd.x >> 8) & 15;
    const int site = (bd.x & 1) ? j0 : i0;
    if (bd.x & 2) B.inf = *(const uint4*)(a.pinfo + (size_t)site * a.npad + bd.y + 4 * lane);
#pragma unroll
    for (int u = 0; u < PB; u += 4)
        if (u < nr) pair_load_quad<PB>(a, B, bd, site, lane, u);
}

// Multiplies buffered batch bd into acc while loading batch bn (sites ni, nj;
// bn.x == 0 loads nothing), and returns the segment's contribution when bd is
// its last batch.
template <int PB>
__device__ __forceinline__ float pair_step(const SweepArgs& a, PairBuf<PB>& B, int4 bd, int4 bn, int ni, int nj,
                                           const float2* s32, float2 di, float2 dj, float2 (&acc)[4], uint4& inf,
                                           int lane) {
    const int nr = (bd.x >> 8) & 15, nn = (bn.x >> 8) & 15;
    const int nsite = (bn.x & 1) ? nj : ni;
    if (__builtin_expect(bd.x & 2, 0)) {  // a segment's first batch (about a third of them)
        inf = B.inf;
#pragma unroll
        for (int j = 0; j < 4; j++) acc[j] = make_float2(0.f, 0.f);
    }
    if (__builtin_expect(bn.x & 2, 0)) B.inf = *(const uint4*)(a.pinfo + (size_t)nsite * a.npad + bn.y + 4 * lane);
    const float4* tp = (const float4*)(s32 + bd.z);
#pragma unroll
    for (int u = 0; u < PB; u += 4) {
        if (u < nr && lane < quad_lanes(bd.x, u / 4)) {
            const float4 t01 = tp[u / 2], t23 = tp[u / 2 + 1];
            const float ta[8] = {t01.x, t01.y, t01.z, t01.w, t23.x, t23.y, t23.z, t23.w};
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const float4 q = B.g[u + v];
                const float tx = ta[2 * v], ty = ta[2 * v + 1];
                acc[0].x += q.x * tx; acc[0].y += q.x * ty;
                acc[1].x += q.y * tx; acc[1].y += q.y * ty;
                acc[2].x += q.z * tx; acc[2].y += q.z * ty;
                acc[3].x += q.w * tx; acc[3].y += q.w * ty;
            }
        }
        if (u < nn) pair_load_quad<PB>(a, B, bn, nsite, lane, u);
    }
    if (__builtin_expect(!(bd.x & 4), 1)) return 0.f;  // not a segment's last batch
    const uint32_t e[4] = {inf.x, inf.y, inf.z, inf.w};
    float2 Hc = make_float2(0.f, 0.f);
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const char* sb = (const char*)s32;
        const float2 tb = *(const float2*)(sb + (e[j] & 0xffffu));
        const float2 tc = *(const float2*)(sb + ((e[j] >> 16) & 0x7fffu));
        const float zx = tc.x * acc[j].x + tc.y * acc[j].y;
        const float zy = __uint_as_float(__float_as_uint(tc.x * acc[j].y - tc.y * acc[j].x) ^ (e[j] & 0x80000000u));
        Hc.x += tb.x * zx - tb.y * zy;
        Hc.y += tb.x * zy + tb.y * zx;
    }
    const float2 d = (bd.x & 1) ? dj : di;
    return -(d.x * Hc.x + d.y * Hc.y);
}

// Local-field contribution of the thread, GROUPS layout (dense, any channel
// pattern).  Site x's H4 entries are grouped by their conjugated pair {m, n}
// and cut into chunks of up to four partners p, so a chunk adds
// conj(s~_m s~_n) * sum_k g_k s~_pk to the field of x: 3 LDS reads and 4 FMA
// per chunk plus 1 LDS read and 2 FMA per entry, against 3 reads and about
// 10 FMA per entry in the LISTS layout.  The chunk ranges of the step's two
// sites are laid end to end; thread t takes chunks (bdim - 1 - t) + u bdim,
// GB at a time.  Loads run one batch ahead of the arithmetic, and the last
// batch of a step loads the next step's first batch, so those loads cross the
// barrier.
template <int GB>
struct GrpBuf {
    uint2 h[GB];
    uint4 g[GB];
};

struct GrpRange {
    int oi, ci, oj, n_tot;  // chunk offset and count of site i0, offset of j0, total
};

// Wave-uniform (the sites are), so the batch loop runs on the scalar unit.
__device__ __forceinline__ GrpRange grp_range(const int* goff, int i0, int j0) {
    GrpRange r;
    r.oi = __builtin_amdgcn_readfirstlane(goff[i0]);
    r.ci = __builtin_amdgcn_readfirstlane(goff[i0 + 1]) - r.oi;
    r.oj = __builtin_amdgcn_readfirstlane(goff[j0]);
    r.n_tot = r.ci + __builtin_amdgcn_readfirstlane(goff[j0 + 1]) - r.oj;
    return r;
}

// Loads chunk e of range r into slot u of the buffer; past the range, the
// all-zero chunk (s~_0 with zero couplings).  No branch, so the buffer keeps
// its registers from one batch to the next.
template <int GB>
__device__ __forceinline__ void grp_load(const SweepArgs& a, GrpBuf<GB>& B, int u, const GrpRange& r, int e) {
    const int idx = e < r.ci ? r.oi + e : (e < r.n_tot ? r.oj + (e - r.ci) : a.grp_zero);
    B.h[u] = a.grp_hdr[idx];
    B.g[u] = a.grp_g[idx];
}

// The chunk in slot u: -Re[d conj(s~_m s~_n) sum_k g_k s~_pk], d of its site.
template <bool BF16, int GB>
__device__ __forceinline__ float grp_chunk(const GrpBuf<GB>& B, int u, const float2* s32, bool isj, float2 di,
                                           float2 dj) {
    const uint2 h = B.h[u];
    const float2 sm = s32[h.x & 0xff], sn = s32[(h.x >> 8) & 0xff];
    const float cx = sm.x * sn.x - sm.y * sn.y, cy = -(sm.x * sn.y + sm.y * sn.x);  // conj(s_m s_n)
    const uint32_t gw[4] = {B.g[u].x, B.g[u].y, B.g[u].z, B.g[u].w};
    float wx = 0.f, wy = 0.f;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const float2 sp = s32[(h.y >> (8 * k)) & 0xff];
        const float g = coupling<BF16>(gw[k]);
        wx += g * sp.x;
        wy += g * sp.y;
    }
    const float xr = wx * cx - wy * cy, xi = wx * cy + wy * cx;
    const float2 d = isj ? dj : di;
    return -(d.x * xr - d.y * xi);
}

// A proposal for one step under one hypothesis about the step before it.
struct Prop {
    double2 ni, nj;      // proposed FP64 amplitudes
    float2 oi, oj;       // FP32 of the current amplitudes
    float2 fi, fj;       // FP32 of the proposed amplitudes
    float2 mi, mj;       // FP32 midpoints
    float2 di, dj;       // FP32 differences new - old
};
// (oi, oj), (fi, fj), (mi, mj), (di, dj) are each read as one float4
static_assert(offsetof(Prop, oi) % 16 == 0 && offsetof(Prop, fi) % 16 == 0 && offsetof(Prop, mi) % 16 == 0 &&
                  offsetof(Prop, di) % 16 == 0 && sizeof(Prop) % 16 == 0,
              "Prop pairs must be 16-byte aligned");

// Constraint-preserving proposal (include/spins.h:17-52,
// include/spins_sparse.h:26-56), FP64, from the current amplitudes ai, aj.
template <bool CUBE>
__device__ __noinline__ Prop propose(const StepPre& P, double2 ai, double2 aj) {
    const double r1 = ai.x * ai.x + ai.y * ai.y, r2 = aj.x * aj.x + aj.y * aj.y;
    double2 ni, nj;
    if (CUBE) {
        const double sq = sqrt(r1 * r1 + r2 * r2);
        const double m1 = sqrt(sq * P.c[0]), m2 = sqrt(sq * P.c[1]);
        ni = make_double2(m1 * P.c[2], m1 * P.c[3]);
        nj = make_double2(m2 * P.c[4], m2 * P.c[5]);
    } else {
        const double f1 = sqrt((r1 + r2) / r1) * P.c[0], f2 = sqrt((r1 + r2) / r2) * P.c[1];
        ni = make_double2((ai.x * P.c[2] + ai.y * P.c[3]) * f1, (-ai.x * P.c[3] + ai.y * P.c[2]) * f1);
        nj = make_double2((aj.x * P.c[4] + aj.y * P.c[5]) * f2, (-aj.x * P.c[5] + aj.y * P.c[4]) * f2);
    }
    Prop q;
    q.ni = ni;
    q.nj = nj;
    q.oi = make_float2((float)ai.x, (float)ai.y);
    q.oj = make_float2((float)aj.x, (float)aj.y);
    q.fi = make_float2((float)ni.x, (float)ni.y);
    q.fj = make_float2((float)nj.x, (float)nj.y);
    q.mi = make_float2((float)(0.5 * (ai.x + ni.x)), (float)(0.5 * (ai.y + ni.y)));
    q.mj = make_float2((float)(0.5 * (aj.x + nj.x)), (float)(0.5 * (aj.y + nj.y)));
    q.di = make_float2((float)(ni.x - ai.x), (float)(ni.y - ai.y));
    q.dj = make_float2((float)(nj.x - aj.x), (float)(nj.y - aj.y));
    return q;
}

// The per-step values every wave reads at the head of a step.
struct StepShared {
    StepPre pre[kMaxN / 2];
    Prop prop[3][2];                 // [step mod 3][outcome of the step before]
    __align__(16) float part[2][16]; // [step mod 2][wave] partial Delta E; unused slots stay 0
};

// Step-head gather: three 4-byte LDS reads per lane, each from a wave-uniform
// base plus a lane offset, issued together with the resolved step's proposal
// reads (one round trip after the barrier):
//   part  lanes 0-15: the 16 partial Delta E of step sp
//   pres  lanes 0-15: step sp's StepPre words (i0, j0, ..., thr in 14-15)
//   pren  lanes 0-1:  i0, j0 of step sn
struct Head {
    uint32_t part, pres, pren;
};

__device__ __forceinline__ Head head_gather(const StepShared& ss, int lane, int sn, int sp) {
    const int l16 = lane & 15;
    Head h;
    h.part = ((const uint32_t*)ss.part[sp & 1])[l16];
    h.pres = ((const uint32_t*)&ss.pre[sp])[l16];
    h.pren = ((const uint32_t*)&ss.pre[sn])[lane & 1];
    return h;
}

// Sum of the 16 partials in lanes 0-15 (one DPP row) in a fixed order:
// pairs, quads, then rows of 8 and 16 by rotation; read from lane 0.
__device__ __forceinline__ float head_sum16(uint32_t hv) {
    float v = __uint_as_float(hv);
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xb1, 0xf, 0xf, false));  // quad_perm 1,0,3,2
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4e, 0xf, 0xf, false));  // quad_perm 2,3,0,1
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x124, 0xf, 0xf, false)); // row_ror 4
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x128, 0xf, 0xf, false)); // row_ror 8
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0));
}

// Step pipeline, one workgroup barrier per step.  During step t's field
// phase, lanes 0 and 1 of wave 0 compute step t+1's proposal under the two
// outcomes of step t (rejected, accepted): the pairs of consecutive steps may
// share a site, and then the proposal depends on the outcome.  After the
// barrier every wave sums the partial Delta E of step t, decides acceptance
// (the same FP32 sum and FP64 threshold in every wave), and updates its own
// FP32 copy of the amplitudes: the final values of step t's sites, then the
// midpoints of step t+1's.  Private copies are written with wave-uniform
// addresses and values by all lanes, so each wave's LDS program order is the
// only ordering needed.  Wave 0 owns the FP64 amplitudes.
template <bool CUBE, int LAYOUT, bool BF16, int PB = 8>
__global__ __launch_bounds__(1024) void k_sweep(SweepArgs a) {
    constexpr bool PAIRS = LAYOUT == kPairs, GROUPS = LAYOUT == kGroups;
    __shared__ double2 s64[kMaxN];
    __shared__ __align__(16) float2 view[16][kMaxN];
    __shared__ StepShared ss;
    StepPre* pre = ss.pre;
    Prop(*prop)[2] = ss.prop;
    float(*part)[16] = ss.part;
    __shared__ int last_wg;
    __shared__ double beta_new;  // the beta after an exchange, read by thread 0
    __shared__ double sbs[kTagMaxNT];  // tagged launches: the temperatures by level
    __shared__ int goff[kMaxN + 1];  // GROUPS: chunk offsets per site
    extern __shared__ float g2s[];  // [N][N] two-body couplings when g2_lds

    const int rep = blockIdx.x;
    const int tid = threadIdx.x, bdim = blockDim.x;
    // The wave index is uniform across the wave; saying so lets the compiler
    // keep the task and row tables on the scalar path.
    const int nw = bdim >> 6, w = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const int N = a.N, nsteps = N / 2;
    double* gs = a.spins + (size_t)2 * N * rep;
    float2* s32 = view[w];

    for (int i = tid; i < N; i += bdim) s64[i] = make_double2(gs[2 * i], gs[2 * i + 1]);
    if (tid < 32) part[tid >> 4][tid & 15] = 0.f;
    // pad entries are read by zero-coupling rows of the pair tiles
    for (int i = lane; i < ((N + 3) & ~3); i += 64)
        s32[i] = i < N ? make_float2((float)gs[2 * i], (float)gs[2 * i + 1]) : make_float2(0.f, 0.f);
    double beta = a.betas[rep];
    const uint64_t seed = a.seeds[rep];
    const uint64_t sweep0 = a.ctr[rep];
    const uint64_t gkey = a.pair_groups > 0 ? pair_group_key(a.pair_seed, rep % a.pair_groups) : 0;
    double E = 0.0;
    long long nacc = 0, nprop = 0;
    if (tid == 0) E = a.E[rep];
    if (a.h2_active && a.g2_lds)
        for (int i = tid; i < N * N; i += bdim) g2s[i] = a.g2[i];
    if (GROUPS)
        for (int i = tid; i <= N; i += bdim) goff[i] = a.grp_off[i];
    if (a.tagged)
        for (int i = tid; i < a.ex.NT; i += bdim) sbs[i] = a.ex.beta_sched[i];
    // GROUPS: chunks are dealt from the last thread down, so a partial last
    // batch leaves wave 0 (which also forms the next proposal) the least work
    const int gt = bdim - 1 - tid;
    GrpBuf<PB> gB;

    // PAIRS: this wave's batch descriptors, one per lane (an even count)
    int4 desc = make_int4(0, 0, 0, 0);
    int nb = 0;
    PairBuf<PB> bA;
    if (PAIRS) {
        const int t0 = a.wave_tasks[w];
        nb = a.wave_tasks[w + 1] - t0;
        if (lane < nb) desc = a.tasks[t0 + lane];
    }

    // A round's first sweep is drawn during the previous round's exchange
    // (waves 1.. while wave 0 runs it or waits for it), with the thresholds
    // left undivided until the new beta is known.
    bool drawn = false;
    for (int rd = 0; rd < a.rounds; rd++) {
    for (int sw = 0; sw < a.n_sweeps; sw++) {
        const uint64_t sweep = sweep0 + (uint64_t)rd * a.n_sweeps + (uint64_t)sw;
        if (drawn) {  // the steps were drawn during the exchange; only the thresholds need beta
            for (int k = tid; k < nsteps; k += bdim) thr_fixup(pre[k], beta);
        } else {
            for (int part = w; part < 5; part += nw)  // five waves draw the steps together
                for (int k = lane; k < nsteps; k += 64)
                    make_step_part<CUBE>(pre[k], part, seed, gkey, sweep, k, N, beta, false);
        }
        drawn = false;
        __syncthreads();
        if (PAIRS && nb > 0)
            pair_issue<PB>(a, bA, desc_at(desc, 0), __builtin_amdgcn_readfirstlane(pre[0].i0),
                           __builtin_amdgcn_readfirstlane(pre[0].j0), lane);
        if (GROUPS) {
            const GrpRange r0 = grp_range(goff, __builtin_amdgcn_readfirstlane(pre[0].i0),
                                          __builtin_amdgcn_readfirstlane(pre[0].j0));
#pragma unroll
            for (int u = 0; u < PB; u++) grp_load<PB>(a, gB, u, r0, gt + u * bdim);
        }
        StepRec* logp = nullptr;
        const int slog = a.log_base + rd * a.n_sweeps + sw;
        if (a.log && slog < a.log_cap) logp = a.log + ((size_t)rep * a.log_cap + slog) * nsteps;
        if (w == 0 && lane < 2) {
            const StepPre& P = pre[0];
            prop[0][lane] = propose<CUBE>(P, s64[P.i0], s64[P.j0]);
        }
        __syncthreads();

        int hyp = 0;  // outcome of the step before the current one (selects its proposal)
        // ring slots of prop[] (step mod 3) for the step resolved, this step and the next
        int r_sp = 0, r_st = 0, r_nx = 1;
        for (int step = 0; step <= nsteps; step++) {
            // Step head.  The values it needs before the decision (partial sums,
            // sites, threshold, both outcomes of the resolved step) are read in
            // one round trip after the barrier; the partials, sites and
            // threshold by lane gathers taken with readlane, the amplitudes
            // as uniform-address vector reads.  The proposal picked by the
            // decision is a second read (cheaper in issue slots, which the four
            // waves of a SIMD share, than gathering both hypotheses).  Every
            // wave sums the partials in the same fixed order and so takes the
            // same decision.
            const bool more = step < nsteps;
            const int sn = more ? step : 0;
            const int sp = step > 0 ? step - 1 : 0;  // the step resolved here (none at step 0)
            const Head hv = head_gather(ss, lane, sn, sp);
            const Prop& Q = prop[r_sp][hyp];
            const float4 o01 = *(const float4*)&Q.oi, f01 = *(const float4*)&Q.fi;
            if (step > 0) {
                // Resolve step-1 (src/mc.cu:284-304).
                const float dE = head_sum16(hv.part);
                const double thr = __hiloint2double((int)__builtin_amdgcn_readlane(hv.pres, 15),
                                                    (int)__builtin_amdgcn_readlane(hv.pres, 14));
                const int ri = __builtin_amdgcn_readlane(hv.pres, 0);
                const int rj = __builtin_amdgcn_readlane(hv.pres, 1);
                const bool accept = (beta == 0.0) | ((double)dE < thr);
                s32[ri] = accept ? make_float2(f01.x, f01.y) : make_float2(o01.x, o01.y);
                s32[rj] = accept ? make_float2(f01.z, f01.w) : make_float2(o01.z, o01.w);
                if (__builtin_expect(w == 0 && accept, 0)) {
                    const double2 ni = Q.ni, nj = Q.nj;
                    s64[ri] = ni;
                    s64[rj] = nj;
                }
                if (tid == 0) {
                    if (accept) {
                        E += (double)dE;
                        nacc++;
                    }
                    nprop++;
                    if (__builtin_expect(logp != nullptr, 0)) {
                        StepRec r;
                        r.i0 = ri; r.j0 = rj; r.acc = accept; r.dE = dE;
                        r.ai[0] = Q.ni.x; r.ai[1] = Q.ni.y;
                        r.aj[0] = Q.nj.x; r.aj[1] = Q.nj.y;
                        logp[sp] = r;
                    }
                }
                hyp = __builtin_amdgcn_readfirstlane((int)accept);
            }
            if (!more) break;

            // wave-uniform: the address arithmetic of the tiles stays scalar
            const int i0 = __builtin_amdgcn_readlane(hv.pren, 0), j0 = __builtin_amdgcn_readlane(hv.pren, 1);
            const float4 m01 = *(const float4*)&prop[r_st][hyp].mi;
            const float4 d01 = *(const float4*)&prop[r_st][hyp].di;
            const float2 di = make_float2(d01.x, d01.y), dj = make_float2(d01.z, d01.w);
            s32[i0] = make_float2(m01.x, m01.y);
            s32[j0] = make_float2(m01.z, m01.w);
            const Prop& Qs = prop[r_st][hyp];
            float local = 0.f;
            if (__builtin_expect(a.h2_active && w == nw - 1, 0)) {
                float2 Fi = make_float2(0.f, 0.f), Fj = make_float2(0.f, 0.f);
                if (a.g2_lds) {  // no global load here: it would wait behind the prefetched rows
                    for (int k = lane; k < N; k += 64) {
                        const float2 t = s32[k];
                        const float gi = k != i0 ? g2s[i0 * N + k] : 0.f, gj = k != j0 ? g2s[j0 * N + k] : 0.f;
                        Fi.x += gi * t.x; Fi.y -= gi * t.y;
                        Fj.x += gj * t.x; Fj.y -= gj * t.y;
                    }
                } else {
                    const float* ri = a.g2 + (size_t)i0 * N;
                    const float* rj = a.g2 + (size_t)j0 * N;
                    for (int k = lane; k < N; k += 64) {
                        const float2 t = s32[k];
                        if (k != i0) { const float g = ri[k]; Fi.x += g * t.x; Fi.y -= g * t.y; }
                        if (k != j0) { const float g = rj[k]; Fj.x += g * t.x; Fj.y -= g * t.y; }
                    }
                }
                local = -(di.x * Fi.x - di.y * Fi.y) - (dj.x * Fj.x - dj.y * Fj.y);
            }
            if (PAIRS) {
                float2 acc[4];
                uint4 inf = make_uint4(0u, 0u, 0u, 0u);
                const int ni = __builtin_amdgcn_readfirstlane(step + 1 < nsteps ? pre[step + 1].i0 : 0);
                const int nj = __builtin_amdgcn_readfirstlane(step + 1 < nsteps ? pre[step + 1].j0 : 0);
                int4 bd = nb > 0 ? desc_at(desc, 0) : make_int4(0, 0, 0, 0);
                for (int b = 0; b < nb; b++) {
                    int4 bn = make_int4(0, 0, 0, 0);
                    if (b + 1 < nb) bn = desc_at(desc, b + 1);
                    else if (step + 1 < nsteps) bn = desc_at(desc, 0);
                    const bool same = b + 1 < nb;
                    local += pair_step<PB>(a, bA, bd, bn, same ? i0 : ni, same ? j0 : nj, s32, di, dj, acc, inf, lane);
                    bd = bn;
                }
            } else if (GROUPS) {
                const GrpRange r = grp_range(goff, i0, j0);
                const bool more = step + 1 < nsteps;
                const int ni = __builtin_amdgcn_readfirstlane(more ? pre[step + 1].i0 : 0);
                const int nj = __builtin_amdgcn_readfirstlane(more ? pre[step + 1].j0 : 1);
                const GrpRange rn = grp_range(goff, ni, nj);
                const int per = PB * bdim;
                const int nbt = r.n_tot > per ? (r.n_tot + per - 1) / per : 1;
                // As soon as a chunk has been used, its registers take the
                // same slot of the next batch (of this step or the next).
                for (int b = 0; b < nbt; b++) {
                    const bool same = b + 1 < nbt;
                    const GrpRange& rl = same ? r : rn;
                    const int e1 = same ? gt + (b + 1) * per : gt;
#pragma unroll
                    for (int u = 0; u < PB; u++) {
                        local += grp_chunk<BF16, PB>(gB, u, s32, gt + b * per + u * bdim >= r.ci, di, dj);
                        grp_load<PB>(a, gB, u, rl, e1 + u * bdim);  // (after a sweep's last step: unused)
                    }
                }
            } else
                local += field_lists<BF16, PB>(a, s32, i0, j0, di, dj, tid, bdim);
            local = wave_sum(local);
            if (lane == 0) part[step & 1][w] = local;
            if (__builtin_expect(w == 0 && lane < 2 && step + 1 < nsteps, 0)) {
                // Next step's proposal if this step is rejected (lane 0) or accepted (lane 1).
                const StepPre& Pn = pre[step + 1];
                double2 ai = s64[Pn.i0], aj = s64[Pn.j0];
                if (lane == 1) {
                    if (Pn.i0 == i0) ai = Qs.ni; else if (Pn.i0 == j0) ai = Qs.nj;
                    if (Pn.j0 == i0) aj = Qs.ni; else if (Pn.j0 == j0) aj = Qs.nj;
                }
                prop[r_nx][lane] = propose<CUBE>(Pn, ai, aj);
            }
            r_sp = r_st;
            r_st = r_nx;
            r_nx = r_nx == 2 ? 0 : r_nx + 1;
            __syncthreads();
        }
        __syncthreads();  // pre[] and prop[] are rewritten by the next sweep
    }
    if (a.exchange) {
        // Without tagged handoffs (below), the last workgroup out runs the
        // replica exchange on every energy.  Between
        // rounds of a cooperative launch the others wait for their new beta
        // (a grid barrier: an arrival count, and each replica's beta slot).
        // Ordering across workgroups follows the HIP memory model at agent
        // scope: each arrival is a release RMW, so the workgroup's energy
        // store and the marker it leaves in its beta slot come before it; the
        // last arriver acquires before staging the energies, so its stores of
        // every replica's new beta come after every marker.  A waiter polls
        // its own slot until the marker is gone and takes the value it reads
        // as its beta: one read, and no other data from the exchange is used
        // by it (the next exchanging workgroup reads the permutation after its
        // own acquire, which the previous exchanger's next arrival releases).
        // The data moves by agent-scope atomics (dev_load, dev_store), so no
        // access is served from a stale cache line.
        const bool more = rd + 1 < a.rounds;
        // the next round's first sweep, drawn by waves 1.. (nothing when the
        // workgroup is one wave)
        auto draw_next = [&]() {
            const uint64_t sweep_n = sweep0 + (uint64_t)(rd + 1) * a.n_sweeps;
            for (int part = w - 1; part < 5; part += nw - 1)
                for (int k = lane; k < nsteps; k += 64)
                    make_step_part<CUBE>(pre[k], part, seed, gkey, sweep_n, k, N, beta, true);
        };
        if (a.tagged) {
            // Tagged handoffs (cooperative launches, every workgroup resident):
            // each workgroup publishes its energy as two words {tag, half};
            // workgroup 0 gathers them, runs the exchange and publishes each
            // replica's new level as one word {tag, level}.  A word carries
            // all that is read from it, so no ordering between words is
            // needed: a reader that sees this round's tag has this round's
            // value.  The exchange state (permutation, counters, LCG state,
            // series) is written and read by workgroup 0 alone.
            const unsigned tag = a.tag0 + (unsigned)rd + 1u;
            if (tid == 0) {
                const unsigned long long eb = (unsigned long long)__double_as_longlong(E);
                dev_store(a.ew + 2 * rep, ((unsigned long long)tag << 32) | (eb & 0xffffffffull));
                dev_store(a.ew + 2 * rep + 1, ((unsigned long long)tag << 32) | (eb >> 32));
            }
            if (rep == 0) {
                exchange_stage(a.ex, &view[0][0], tid, bdim, a.ew, tag, a.done + 2, a.spin_limit);
                __syncthreads();
                if (w == 0) exchange_wave(a.ex, &view[0][0], lane);
                else if (more) draw_next();
                __syncthreads();
                exchange_commit(a.ex, &view[0][0], tid, bdim, rep, &beta_new, a.lvw, tag);
                if (more) {
                    __syncthreads();  // the exchange's LDS is read: the amplitude copies go back over it
                    for (int i = lane; i < ((N + 3) & ~3); i += 64)
                        s32[i] = i < N ? make_float2((float)s64[i].x, (float)s64[i].y) : make_float2(0.f, 0.f);
                }
            } else if (more) {
                if (tid == 0) {
                    // Bounded wait, as below.
                    double b = beta;
                    for (unsigned n = 0;; n++) {
                        const unsigned long long v = dev_load(a.lvw + rep);
                        if ((unsigned)(v >> 32) == tag) {
                            b = sbs[(unsigned)v];
                            break;
                        }
                        if (n == a.spin_limit) {
                            __hip_atomic_fetch_or(a.done + 2, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            break;
                        }
                        __builtin_amdgcn_s_sleep(2);
                    }
                    beta_new = b;
                } else if (w > 0) {
                    draw_next();
                }
            }
        } else {
            if (tid == 0) {
                dev_store(a.E + rep, E);
                if (more) dev_store((unsigned long long*)(a.betas + rep), kBetaMarker);
                last_wg = __hip_atomic_fetch_add(a.done, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT) == gridDim.x - 1;
                if (last_wg) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            }
            __syncthreads();
            if (last_wg) {
                if (tid == 0) dev_store(a.done, 0u);
                exchange_stage(a.ex, &view[0][0], tid, bdim);
                __syncthreads();
                if (w == 0) exchange_wave(a.ex, &view[0][0], lane);
                else if (more) draw_next();
                __syncthreads();
                // every thread writes the round out; this replica's new beta comes back in beta_new
                exchange_commit(a.ex, &view[0][0], tid, bdim, rep, &beta_new);
                if (more) {
                    __syncthreads();  // the exchange's LDS is read: the amplitude copies go back over it
                    for (int i = lane; i < ((N + 3) & ~3); i += 64)
                        s32[i] = i < N ? make_float2((float)s64[i].x, (float)s64[i].y) : make_float2(0.f, 0.f);
                }
            } else if (more) {
                if (tid == 0) {
                    // Bounded wait: every workgroup is resident (cooperative launch), so
                    // the new beta comes; past the bound the kernel flags it and keeps
                    // the old beta.
                    double b = beta;
                    for (unsigned n = 0;; n++) {
                        const unsigned long long v = dev_load((const unsigned long long*)(a.betas + rep));
                        if (v != kBetaMarker) {
                            b = __longlong_as_double((long long)v);
                            break;
                        }
                        if (n == a.spin_limit) {
                            __hip_atomic_fetch_or(a.done + 2, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            break;
                        }
                        __builtin_amdgcn_s_sleep(2);
                    }
                    beta_new = b;
                } else if (w > 0) {
                    draw_next();
                }
            }
        }
        if (more) {
            __syncthreads();
            beta = beta_new;  // this replica's temperature after the exchange
            drawn = nw > 1;
        }
    }
    }

    for (int i = tid; i < N; i += bdim) {
        gs[2 * i] = s64[i].x;
        gs[2 * i + 1] = s64[i].y;
    }
    if (tid == 0) {
        a.E[rep] = E;
        a.acc[rep] += nacc;
        a.prop[rep] += nprop;
        a.ctr[rep] = sweep0 + (uint64_t)a.rounds * a.n_sweeps;
    }
}

constexpr size_t kG2LdsMax = 64 * 1024;  // two-body matrix staged in LDS up to N = 128

bool exchange_fusable(const pspin_state* st) {
    return st->pt_NT >= 2 && st->pt_nrep <= 64 && st->d_done &&
           exchange_lds_bytes(st->pt_NT, st->pt_nrep) <= sizeof(float2) * 16 * kMaxN;
}

using SweepKernel = void (*)(SweepArgs);

// The kernel instance for the state's variant, layout, tile and coupling format.
SweepKernel sweep_kernel(const pspin_state* st) {
    const bool bf16 = st->coupling_format == PSPIN_COUPLING_BF16X2;
    if (st->variant == PSPIN_SPARSE) return bf16 ? k_sweep<true, kLists, true> : k_sweep<true, kLists, false>;
    if (st->kernel == PSPIN_KERNEL_PAIRS)
        return st->pair_rows == 4 ? k_sweep<false, kPairs, false, 4> : k_sweep<false, kPairs, false, 8>;
    if (st->kernel == PSPIN_KERNEL_GROUPS) {
        if (st->list_batch == 16) return bf16 ? k_sweep<false, kGroups, true, 6> : k_sweep<false, kGroups, false, 6>;
        return bf16 ? k_sweep<false, kGroups, true, 4> : k_sweep<false, kGroups, false, 4>;
    }
    if (st->list_batch == 16) return bf16 ? k_sweep<false, kLists, true, 16> : k_sweep<false, kLists, false, 16>;
    return bf16 ? k_sweep<false, kLists, true> : k_sweep<false, kLists, false>;
}

int sweep_block(const pspin_state* st) {
    int bs = st->kernel == PSPIN_KERNEL_PAIRS ? 64 * st->pair_nwaves : st->block_size;
    if (bs <= 0 && st->kernel == PSPIN_KERNEL_GROUPS) {
        // A few chunks per thread per step.
        const long long per_step = 2LL * st->max_site_grp;
        bs = per_step >= 4096 ? 1024 : per_step >= 2048 ? 512 : per_step >= 512 ? 256 : 64;
    } else if (bs <= 0) {
        // Enough threads that each touches a few entries per step.
        long long per_step = 2LL * st->max_site_len;
        bs = per_step >= 8192 ? 1024 : per_step >= 1024 ? 256 : 64;
    }
    return bs;
}

SweepArgs sweep_args(pspin_state* st, int n_sweeps, bool exchange, int rounds) {
    SweepArgs a;
    a.exchange = exchange && exchange_fusable(st);
    a.done = st->d_done;
    a.tagged = 0;
    a.tag0 = 0;
    a.spin_limit = st->wait_limit ? st->wait_limit : kBarrierSpinLimit;
    a.ew = st->d_hand;
    a.lvw = st->d_hand ? st->d_hand + 2 * st->nrep : nullptr;
    if (a.exchange) a.ex = exchange_args(st);
    a.rounds = rounds;
    a.N = st->N;
    a.nrep = st->nrep;
    a.h2_active = st->h2_active;
    a.g2_lds = st->h2_active && (size_t)st->N * st->N * sizeof(float) <= kG2LdsMax;
    a.n_sweeps = n_sweeps;
    a.spins = st->d_spins;
    a.g2 = st->d_g2_32;
    a.site_off = st->d_site_off;
    a.site_ent = st->d_site_ent;
    a.betas = st->d_betas;
    a.E = st->d_E;
    a.acc = st->d_acc;
    a.prop = st->d_prop;
    a.seeds = st->d_seeds;
    a.ctr = st->d_ctr;
    a.log = st->d_log;
    a.log_cap = st->log_cap;
    a.log_base = st->log_n;
    a.ptab = st->d_ptab;
    a.pinfo = st->d_pinfo;
    a.slice_len = st->pair_slice_len;
    a.npad = st->pair_npad;
    a.tasks = st->d_tasks;
    a.wave_tasks = st->d_wave_tasks;
    a.grp_off = st->d_grp_off;
    a.grp_hdr = st->d_grp_hdr;
    a.grp_g = st->d_grp_g;
    a.grp_zero = (int)st->n_grp;
    a.pair_groups = st->pair_groups;
    a.pair_seed = st->cfg.seed;
    return a;
}

size_t sweep_shm(const SweepArgs& a) { return a.g2_lds ? (size_t)a.N * a.N * sizeof(float) : 0; }


hipError_t launch_sweep(pspin_state* st, int n_sweeps, bool exchange) {
    SweepArgs a = sweep_args(st, n_sweeps, exchange, 1);
    hipLaunchKernelGGL(sweep_kernel(st), dim3(st->nrep), dim3(sweep_block(st)), sweep_shm(a), st->stream, a);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && exchange && !a.exchange) e = launch_exchange(st);
    return e;
}

bool pt_rounds_ok(const pspin_state* st) {
    if (!exchange_fusable(st)) return false;
    int dev = 0, coop = 0, ncu = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return false;
    if (hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) != hipSuccess || !coop) return false;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return false;
    SweepArgs a = sweep_args(const_cast<pspin_state*>(st), 1, true, 1);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)sweep_kernel(st), sweep_block(st),
                                                     sweep_shm(a)) != hipSuccess)
        return false;
    return (long long)per_cu * ncu >= st->nrep;
}

hipError_t launch_pt_rounds(pspin_state* st, int pt_freq, int rounds) {
    SweepArgs a = sweep_args(st, pt_freq, true, rounds);
    // Tags only grow from one launch to the next (never replayed from a
    // graph), so a handoff word holds this round's tag or an older one.
    a.tagged = a.exchange && st->pt_NT <= kTagMaxNT && st->d_hand;
    if (a.tagged) {
        a.tag0 = st->hand_tag;
        st->hand_tag += (unsigned)rounds;
    }
    void* args[] = {&a};
    return hipLaunchCooperativeKernel((const void*)sweep_kernel(st), dim3(st->nrep), dim3(sweep_block(st)), args,
                                      (unsigned)sweep_shm(a), st->stream);
}

}  // namespace ps
