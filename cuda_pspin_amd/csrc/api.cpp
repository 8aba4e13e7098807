// api.cpp -- C ABI of libpspin.so (include/pspin.h): state set-up, disorder
// orchestration, sweep / PT loop control, read-back.  Host side of the
// reference's mc_init / mc_sparse_init (src/mc.cu:411-562,
// src/mc_sparse.cu:390-604) and of its PT driver loop
// (src/parallel_tempering.cu:273-352).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "pspin_internal.h"

namespace ps {
static thread_local std::string g_err;
void set_error(const std::string& m) { g_err = m; }
}  // namespace ps

#define PS_TRY_HIP(call)                                                                               \
    do {                                                                                               \
        hipError_t _e = (call);                                                                        \
        if (_e != hipSuccess) {                                                                        \
            ps::set_error(std::string("HIP error ") + hipGetErrorString(_e) + " at " + __FILE__ + ":" + \
                          std::to_string(__LINE__) + " (" #call ")");                                  \
            return PSPIN_ERR_HIP;                                                                      \
        }                                                                                              \
    } while (0)

#define PS_REQUIRE(cond, code, msg)   \
    do {                              \
        if (!(cond)) {                \
            ps::set_error(msg);       \
            return code;              \
        }                             \
    } while (0)

namespace {

template <class T>
hipError_t dmalloc(T** p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    return hipMalloc((void**)p, n * sizeof(T));
}

template <class T>
void dfree(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

// Copies and fills go through the library's stream and complete before return.
hipError_t cpy(const pspin_state* st, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, st->stream);
    return e != hipSuccess ? e : hipStreamSynchronize(st->stream);
}

hipError_t fill0(const pspin_state* st, void* dst, size_t bytes) {
    hipError_t e = hipMemsetAsync(dst, 0, bytes, st->stream);
    return e != hipSuccess ? e : hipStreamSynchronize(st->stream);
}

void free_all(pspin_state* st) {
    dfree(st->d_g2_64); dfree(st->d_g2_32); dfree(st->d_g4_64); dfree(st->d_mask);
    dfree(st->d_sq); dfree(st->d_sqg); dfree(st->d_site_off); dfree(st->d_site_ent);
    dfree(st->d_grp_off); dfree(st->d_grp_hdr); dfree(st->d_grp_g);
    dfree(st->d_spins); dfree(st->d_E); dfree(st->d_betas); dfree(st->d_acc); dfree(st->d_prop);
    dfree(st->d_seeds); dfree(st->d_e2); dfree(st->d_e4); dfree(st->d_ctr);
    dfree(st->d_echunk); dfree(st->d_epart);
    if (st->graph) (void)hipGraphExecDestroy(st->graph);
    st->graph = nullptr;
    dfree(st->d_beta_sched); dfree(st->d_perm); dfree(st->d_ex); dfree(st->d_rng48); dfree(st->d_done); dfree(st->d_hand);
    dfree(st->d_series); dfree(st->d_series_n); dfree(st->d_log);
    dfree(st->d_ptab); dfree(st->d_pinfo);
    dfree(st->d_tasks); dfree(st->d_wave_tasks);
    for (auto& pr : st->ev_sweep) { st->ev_free.push_back(pr.first); st->ev_free.push_back(pr.second); }
    for (auto& pr : st->ev_exch) { st->ev_free.push_back(pr.first); st->ev_free.push_back(pr.second); }
    st->ev_sweep.clear(); st->ev_exch.clear();
    for (hipEvent_t e : st->ev_free) (void)hipEventDestroy(e);
    st->ev_free.clear();
    if (st->stream) (void)hipStreamDestroy(st->stream);
    st->stream = nullptr;
}

// Per-site H4 entry of a quartet term for the site at sorted position s:
// partner p = the other position with the same conjugation bit, m, n the
// remaining two (the term is -g Re[a_site a_p conj(a_m) conj(a_n)]).
inline uint32_t pack_entry(const int* v, int ch, int s) {
    int cmask = ps::conj_mask(ch);
    int bit = (cmask >> s) & 1;
    int p = -1, mn[2], c = 0;
    for (int t = 0; t < 4; t++) {
        if (t == s) continue;
        if (((cmask >> t) & 1) == bit && p < 0) p = t;
        else mn[c++] = t;
    }
    return (uint32_t)v[p] | ((uint32_t)v[mn[0]] << 8) | ((uint32_t)v[mn[1]] << 16);
}

struct Term {
    int v[4];
    int ch;
    double g;
};

int build_site_lists(pspin_state* st, const std::vector<Term>& terms) {
    const int N = st->N;
    std::vector<int> cnt(N + 1, 0);
    for (const Term& t : terms)
        for (int s = 0; s < 4; s++) cnt[t.v[s] + 1]++;
    for (int i = 0; i < N; i++) cnt[i + 1] += cnt[i];
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    std::vector<uint2> ent((size_t)cnt[N]);
    for (const Term& t : terms) {
        float g = (float)t.g;
        uint32_t gb;
        std::memcpy(&gb, &g, 4);
        for (int s = 0; s < 4; s++) {
            uint2 e;
            e.x = pack_entry(t.v, t.ch, s);
            e.y = gb;
            ent[fill[t.v[s]]++] = e;
        }
    }
    st->n_ent = (long long)ent.size();
    st->h_site_ent = ent;
    st->max_site_len = 0;
    for (int i = 0; i < N; i++) st->max_site_len = std::max(st->max_site_len, cnt[i + 1] - cnt[i]);
    PS_TRY_HIP(dmalloc(&st->d_site_off, N + 1));
    PS_TRY_HIP(dmalloc(&st->d_site_ent, ent.size()));
    PS_TRY_HIP(cpy(st, st->d_site_off, cnt.data(), (N + 1) * sizeof(int), hipMemcpyHostToDevice));
    if (!ent.empty())
        PS_TRY_HIP(cpy(st, st->d_site_ent, ent.data(), ent.size() * sizeof(uint2), hipMemcpyHostToDevice));
    return PSPIN_OK;
}

// Per-site chunks of the group kernel.  Site x's entries (p, m, n, g) are
// grouped by the conjugated pair {m, n}: the field of x gets
// conj(a_m a_n) * sum_p g a_p from each group, one product per chunk of up
// to four partners.  Groups are emitted in (m, n) order and cut into chunks
// of 4; unused slots have g = 0 and p = m.
int build_site_groups(pspin_state* st, const std::vector<Term>& terms) {
    const int N = st->N;
    std::vector<std::vector<std::pair<uint32_t, std::pair<int, float>>>> per(N);  // key m | n << 8, (p, g)
    for (const Term& t : terms) {
        const float g = (float)t.g;
        for (int s = 0; s < 4; s++) {
            const uint32_t e = pack_entry(t.v, t.ch, s);
            const int p = e & 0xff, m = (e >> 8) & 0xff, n = (e >> 16) & 0xff;
            const uint32_t key = (uint32_t)std::min(m, n) | ((uint32_t)std::max(m, n) << 8);
            per[t.v[s]].push_back({key, {p, g}});
        }
    }
    std::vector<int> off(N + 1, 0);
    std::vector<uint2> hdr;
    std::vector<uint4> gw;
    for (int x = 0; x < N; x++) {
        auto& v = per[x];
        std::stable_sort(v.begin(), v.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
        for (size_t i = 0; i < v.size();) {
            size_t j = i;
            while (j < v.size() && v[j].first == v[i].first) j++;
            for (size_t c = i; c < j; c += 4) {
                const uint32_t m = v[i].first & 0xff;
                uint32_t ps_[4] = {m, m, m, m}, gb[4] = {0u, 0u, 0u, 0u};
                for (size_t k = c; k < std::min(j, c + 4); k++) {
                    ps_[k - c] = (uint32_t)v[k].second.first;
                    std::memcpy(&gb[k - c], &v[k].second.second, 4);
                }
                hdr.push_back(make_uint2(v[i].first, ps_[0] | ps_[1] << 8 | ps_[2] << 16 | ps_[3] << 24));
                gw.push_back(make_uint4(gb[0], gb[1], gb[2], gb[3]));
            }
            i = j;
        }
        off[x + 1] = (int)hdr.size();
        v.clear();
        v.shrink_to_fit();
    }
    st->n_grp = (long long)hdr.size();
    st->max_site_grp = 0;
    for (int x = 0; x < N; x++) st->max_site_grp = std::max(st->max_site_grp, off[x + 1] - off[x]);
    // one all-zero chunk after the last: the sweep's loads past a range read it
    hdr.push_back(make_uint2(0u, 0u));
    gw.push_back(make_uint4(0u, 0u, 0u, 0u));
    st->h_grp_g = gw;
    PS_TRY_HIP(dmalloc(&st->d_grp_off, N + 1));
    PS_TRY_HIP(dmalloc(&st->d_grp_hdr, hdr.size()));
    PS_TRY_HIP(dmalloc(&st->d_grp_g, gw.size()));
    PS_TRY_HIP(cpy(st, st->d_grp_off, off.data(), (N + 1) * sizeof(int), hipMemcpyHostToDevice));
    PS_TRY_HIP(cpy(st, st->d_grp_hdr, hdr.data(), hdr.size() * sizeof(uint2), hipMemcpyHostToDevice));
    PS_TRY_HIP(cpy(st, st->d_grp_g, gw.data(), gw.size() * sizeof(uint4), hipMemcpyHostToDevice));
    return PSPIN_OK;
}

// Work list of the pair-row kernel for at most `nw_req` waves.  The tiles
// (slot 0 = i0, slot 1 = j0) x (pair block k) are laid end to end and cut into
// contiguous runs, one per wave, at multiples of 4 rows; a run that crosses a
// tile boundary becomes several segments.  The largest run cost is minimised
// by bisection over the cost bound with a greedy fill, where a segment costs
// kSegCost plus kRowCost per row (instruction counts of the kernel's set-up
// and row loop).  Wave 0 also computes the next step's proposal in FP64 and
// starts with kPropCost.  Segments are then cut into batches of `rows` rows
// (4 or 8), the kernel's unit of loads in flight.  Each 4-row quad of a batch
// carries the number of lanes that hold a pair of its first row; the other
// lanes' entries are zeros and are neither loaded nor multiplied.
int build_pair_tasks(pspin_state* st, int nw_req, int rows) {
    constexpr long long kSegCost = 64, kRowCost = 9, kPropCost = 150;
    const int nblk = (int)st->h_tile_rows.size();
    struct Seg { int slot, k, y0, nr; };
    auto fill = [&](long long C, std::vector<std::vector<Seg>>* out) -> int {
        int waves = 1;
        long long c = std::min(kPropCost, C - kSegCost - 4 * kRowCost);
        if (out) out->assign(1, {});
        for (int slot = 0; slot < 2; slot++)
            for (int k = 0; k < nblk; k++) {
                int y = 0;
                const int ny = st->h_tile_rows[k];
                while (y < ny) {
                    long long avail = C - c - kSegCost;
                    int take = avail >= 4 * kRowCost ? (int)(avail / kRowCost) / 4 * 4 : 0;
                    take = std::min(take, ny - y);
                    if (take == 0) {
                        if (c == 0) return 1 << 30;  // bound below one segment
                        waves++;
                        c = 0;
                        if (out) out->push_back({});
                        continue;
                    }
                    if (out) out->back().push_back({slot, k, y, take});
                    c += kSegCost + kRowCost * take;
                    y += take;
                }
            }
        return waves;
    };
    long long total = 0;
    for (int k = 0; k < nblk; k++) total += 2 * (kSegCost + kRowCost * st->h_tile_rows[k]);
    const int nwr = std::max(1, std::min(nw_req, 16));
    long long lo = kSegCost + 4 * kRowCost, hi = total;
    while (lo < hi) {
        const long long mid = (lo + hi) / 2;
        if (fill(mid, nullptr) <= nwr) hi = mid;
        else lo = mid + 1;
    }
    std::vector<std::vector<Seg>> per;
    fill(lo, &per);
    const int nw = (int)per.size();
    const int pb = rows == 4 ? 4 : 8;
    // Lanes holding a pair of row y (ext > y): a prefix, since b falls with the lane.
    const std::vector<uint32_t>& pbc = st->h_pair_bc;
    const int npairs = (int)pbc.size();
    auto lanes_at = [&](int k, int y) {
        int L = 0;
        while (L < 64 && 256 * k + 4 * L < npairs && (int)(pbc[(size_t)256 * k + 4 * L] & 0xffff) + 1 > y) L++;
        return L;
    };
    std::vector<int4> flat;
    std::vector<int> wt(nw + 1, 0);
    for (int w = 0; w < nw; w++) {
        int nb = 0;
        for (const Seg& g : per[w])
            for (int r = 0; r < g.nr; r += pb) {
                const int n = std::min(pb, g.nr - r);
                int x = g.slot | (r == 0 ? 2 : 0) | (r + n == g.nr ? 4 : 0) | (n << 8);
                for (int q = 0; q < n / 4; q++) x |= lanes_at(g.k, g.y0 + r + 4 * q) << (12 + 7 * q);
                flat.push_back(make_int4(x, 256 * g.k, g.y0 + r, st->h_tile_off[g.k] + 256 * (g.y0 + r)));
                nb++;
            }
        PS_REQUIRE(nb <= 64, PSPIN_ERR_ARG, "pair-row work list: more than 64 batches for one wave (use more waves)");
        wt[w + 1] = (int)flat.size();
    }
    dfree(st->d_tasks);
    dfree(st->d_wave_tasks);
    PS_TRY_HIP(dmalloc(&st->d_tasks, flat.size()));
    PS_TRY_HIP(dmalloc(&st->d_wave_tasks, wt.size()));
    PS_TRY_HIP(cpy(st, st->d_tasks, flat.data(), flat.size() * sizeof(int4), hipMemcpyHostToDevice));
    PS_TRY_HIP(cpy(st, st->d_wave_tasks, wt.data(), wt.size() * sizeof(int), hipMemcpyHostToDevice));
    st->pair_nwaves = nw;
    st->pair_rows = pb;
    return PSPIN_OK;
}

// Dense per-site pair tiles for the fully connected case (every quartet
// channel 0).  The other sites of x are relabelled 0..M-1 (M = N-1,
// o(y) = y + (y >= x)) and each unordered triple a < b < c of them is stored
// once, at (row o(a), pair (b,c)).  Pairs (b<c) are ordered by b descending
// and cut into blocks of 256 (64 lanes x 4); block k holds rows
// y = 0 .. b_max(k) (a < b <= b_max means o(a) <= b_max) rounded up to 4, as a
// dense [rows][256] FP32 tile with zeros where a >= b, at y = x and past the
// last pair.  Rows are indexed by the site itself so the kernel reads the row
// amplitude s[y] at a fixed offset.  pinfo[x][P] = 8 o(b) | 8 o(c) << 16 |
// (o(b) > x) << 31: the LDS byte offsets of the pair's amplitudes and the
// partner flag (partner a when o(b) > x, else c).
int build_pair_layout(pspin_state* st, const std::vector<double>& g4, const std::vector<uint8_t>& mask) {
    const int N = st->N, M = N - 1;
    std::vector<uint32_t> pbc;
    for (int b = M - 2; b >= 1; b--)
        for (int c = b + 1; c < M; c++) pbc.push_back((uint32_t)b | ((uint32_t)c << 16));
    const int npairs = (int)pbc.size(), nblk = (npairs + 255) / 256, npad = 256 * nblk;
    const int n4 = (N + 3) / 4 * 4;
    std::vector<int> trows(nblk), toff(nblk);
    long long off = 0;
    for (int k = 0; k < nblk; k++) {
        const int bmax = (int)(pbc[(size_t)256 * k] & 0xffff);
        trows[k] = std::min(n4, (bmax + 1 + 3) / 4 * 4);
        toff[k] = (int)off;
        off += 256LL * trows[k];
    }
    const long long slice = off;
    std::vector<float> tab((size_t)N * slice, 0.0f);
    std::vector<uint32_t> info((size_t)N * npad, 0u);
    for (int x = 0; x < N; x++) {
        float* sl = &tab[(size_t)x * slice];
        for (int P = 0; P < npairs; P++) {
            const int b = pbc[P] & 0xffff, c = pbc[P] >> 16;
            const int ob = b + (b >= x), oc = c + (c >= x), k = P / 256, p = P % 256;
            info[(size_t)x * npad + P] = (uint32_t)(8 * ob) | ((uint32_t)(8 * oc) << 16) | ((uint32_t)(b >= x) << 31);
            for (int a = 0; a < b; a++) {
                const int y = a + (a >= x);
                int v[4] = {x, y, ob, oc};
                std::sort(v, v + 4);
                const long long q = (long long)v[3] * (v[3] - 1) * (v[3] - 2) * (v[3] - 3) / 24 +
                                    (long long)v[2] * (v[2] - 1) * (v[2] - 2) / 6 + (long long)v[1] * (v[1] - 1) / 2 + v[0];
                sl[toff[k] + 256LL * y + p] = (mask[q] == 1u) ? (float)g4[q] : 0.0f;
            }
        }
    }
    st->pair_slice_len = slice;
    st->pair_npairs = npairs;
    st->pair_npad = npad;
    st->h_tile_rows = trows;
    st->h_tile_off = toff;
    st->h_pair_bc = pbc;
    PS_TRY_HIP(dmalloc(&st->d_ptab, tab.size()));
    PS_TRY_HIP(dmalloc(&st->d_pinfo, info.size()));
    PS_TRY_HIP(cpy(st, st->d_ptab, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice));
    PS_TRY_HIP(cpy(st, st->d_pinfo, info.data(), info.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    return build_pair_tasks(st, 16, 8);
}

// splitmix64 frequencies (src/mc.cu:449-460).
void make_omega(const pspin_sim_config& c, std::vector<double>& w) {
    w.assign(c.N, 0.0);
    if (c.fmc_mode == 1) {
        for (int i = 0; i < c.N; i++) w[i] = (double)i;
    } else if (c.fmc_mode == 2) {
        uint64_t s = c.seed + 3000;
        for (int i = 0; i < c.N; i++) w[i] = (double)(ps::splitmix64(&s) >> 11) / (double)(1ULL << 53);
        std::sort(w.begin(), w.end());
    }
}

// Dense g2 set-up shared by both variants (src/mc.cu:436-498,
// src/mc_sparse.cu:404-475): generate, FMC-filter, count i<=j survivors,
// rescale, mean-shift.
int setup_g2(pspin_state* st, const double* d_omega, double J2, double J2_0) {
    const int N = st->N;
    const pspin_sim_config& c = st->cfg;
    PS_TRY_HIP(dmalloc(&st->d_g2_64, (size_t)N * N));
    PS_TRY_HIP(dmalloc(&st->d_g2_32, (size_t)N * N));
    PS_TRY_HIP(hipMemsetAsync(st->d_g2_64, 0, (size_t)N * N * sizeof(double), st->stream));
    if (c.fmc_mode != 1)
        PS_TRY_HIP(ps::launch_g2_build(st->d_g2_64, N, c.seed + 1000, c.fmc_mode > 0, d_omega, c.gamma, st->stream));
    if (c.fmc_mode == 0) {
        st->n_pairs_active = ps::n_pairs(N);
    } else if (c.fmc_mode == 1) {
        st->n_pairs_active = 0;
    } else {
        std::vector<double> h((size_t)N * N);
        PS_TRY_HIP(hipMemcpyAsync(h.data(), st->d_g2_64, h.size() * sizeof(double), hipMemcpyDeviceToHost, st->stream));
        PS_TRY_HIP(hipStreamSynchronize(st->stream));
        long long n = 0;
        for (int i = 0; i < N; i++)
            for (int j = i; j < N; j++) n += (h[(size_t)i * N + j] != 0.0);
        st->n_pairs_active = n;
    }
    if (st->n_pairs_active > 0) {
        double sigma = J2 * std::sqrt((double)N / (double)st->n_pairs_active);
        double mean = (J2_0 != 0.0) ? J2_0 * (double)N / (double)st->n_pairs_active : 0.0;
        PS_TRY_HIP(ps::launch_scale_shift(st->d_g2_64, (long long)N * N, sigma, mean, st->stream));
    }
    st->h2_active = st->n_pairs_active > 0 ? 1 : 0;
    PS_TRY_HIP(ps::launch_to_f32(st->d_g2_64, st->d_g2_32, (long long)N * N, st->stream));
    return PSPIN_OK;
}

int recompute_energies(pspin_state* st) {
    if (st->variant == PSPIN_DENSE)
        PS_TRY_HIP(ps::launch_energy_dense(st, st->d_e2, st->d_e4));
    else
        PS_TRY_HIP(ps::launch_energy_sparse(st, st->d_e2, st->d_e4));
    PS_TRY_HIP(ps::launch_add(st->d_e2, st->d_e4, st->d_E, st->nrep, st->stream));
    return PSPIN_OK;
}

hipEvent_t take_event(pspin_state* st) {
    if (!st->ev_free.empty()) {
        hipEvent_t e = st->ev_free.back();
        st->ev_free.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

// Fold completed event pairs into the totals (blocks on the last one).
int harvest_events(pspin_state* st) {
    for (int pass = 0; pass < 2; pass++) {
        auto& v = pass == 0 ? st->ev_sweep : st->ev_exch;
        double& tot = pass == 0 ? st->prof_sweep_ms : st->prof_exch_ms;
        for (auto& pr : v) {
            PS_TRY_HIP(hipEventSynchronize(pr.second));
            float ms = 0.f;
            PS_TRY_HIP(hipEventElapsedTime(&ms, pr.first, pr.second));
            tot += ms;
            st->ev_free.push_back(pr.first);
            st->ev_free.push_back(pr.second);
        }
        v.clear();
    }
    return PSPIN_OK;
}

int sweep_launch(pspin_state* st, int n, bool exchange = false) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (st->prof) {
        e0 = take_event(st);
        e1 = take_event(st);
        PS_TRY_HIP(hipEventRecord(e0, st->stream));
    }
    PS_TRY_HIP(ps::launch_sweep(st, n, exchange));
    st->sweep_counter += (uint64_t)n;
    if (st->d_log) st->log_n = std::min(st->log_cap, st->log_n + n);
    if (st->prof) {
        PS_TRY_HIP(hipEventRecord(e1, st->stream));
        st->ev_sweep.push_back({e0, e1});
        st->prof_n_sweep += n;
        st->prof_n_launch++;
        if (st->ev_sweep.size() >= 512) {
            int rc = harvest_events(st);
            if (rc) return rc;
        }
    }
    return PSPIN_OK;
}

// `rounds` PT rounds in one cooperative launch (timed like a sweep launch).
int rounds_launch(pspin_state* st, int pt_freq, int rounds) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (st->prof) {
        e0 = take_event(st);
        e1 = take_event(st);
        PS_TRY_HIP(hipEventRecord(e0, st->stream));
    }
    PS_TRY_HIP(ps::launch_pt_rounds(st, pt_freq, rounds));
    const int n = rounds * pt_freq;
    st->sweep_counter += (uint64_t)n;
    if (st->prof) {
        PS_TRY_HIP(hipEventRecord(e1, st->stream));
        st->ev_sweep.push_back({e0, e1});
        st->prof_n_sweep += n;
        st->prof_n_launch++;
        if (st->ev_sweep.size() >= 512) {
            int rc = harvest_events(st);
            if (rc) return rc;
        }
    }
    return PSPIN_OK;
}

int exchange_launch(pspin_state* st) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (st->prof) {
        e0 = take_event(st);
        e1 = take_event(st);
        PS_TRY_HIP(hipEventRecord(e0, st->stream));
    }
    PS_TRY_HIP(ps::launch_exchange(st));
    if (st->prof) {
        PS_TRY_HIP(hipEventRecord(e1, st->stream));
        st->ev_exch.push_back({e0, e1});
        st->prof_n_exch++;
        if (st->ev_exch.size() >= 512) {
            int rc = harvest_events(st);
            if (rc) return rc;
        }
    }
    return PSPIN_OK;
}

constexpr int kGraphSweeps = 32;  // sweeps per captured graph (rounded to whole PT rounds)

void drop_graph(pspin_state* st) {
    if (st->graph) (void)hipGraphExecDestroy(st->graph);
    st->graph = nullptr;
    st->graph_gen++;
    st->graph_pt_freq = st->graph_rounds = 0;
}

// Capture `rounds` rounds of (pt_freq sweeps, one exchange) into a graph.
// The kernels read their RNG counters and PT state from device memory, so
// replaying the same graph continues the chains correctly.
int ensure_graph(pspin_state* st, int pt_freq, int rounds) {
    if (st->graph && st->graph_pt_freq == pt_freq && st->graph_rounds == rounds) return PSPIN_OK;
    drop_graph(st);
    hipGraph_t g = nullptr;
    PS_TRY_HIP(hipStreamBeginCapture(st->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (int k = 0; k < rounds && e == hipSuccess; k++) {
        e = ps::launch_sweep(st, pt_freq, true);
    }
    hipError_t e2 = hipStreamEndCapture(st->stream, &g);
    PS_TRY_HIP(e);
    PS_TRY_HIP(e2);
    hipError_t e3 = hipGraphInstantiate(&st->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    PS_TRY_HIP(e3);
    st->graph_pt_freq = pt_freq;
    st->graph_rounds = rounds;
    return PSPIN_OK;
}

// FP64 recomputation of the running energies every resync_every sweeps.
int maybe_resync(pspin_state* st) {
    if (st->resync_every > 0 && st->since_resync >= st->resync_every) {
        int rc = recompute_energies(st);
        if (rc) return rc;
        st->since_resync = 0;
    }
    return PSPIN_OK;
}

}  // namespace

namespace ps {
// PT sweeps and exchanges queued on the library stream, no final sync (the
// body of pspin_pt_run; also used between learned moves).
int pt_advance(pspin_state* st, int n_sweeps, int pt_freq, int launch_rounds) {
    PS_REQUIRE(st && n_sweeps >= 0 && pt_freq >= 1, PSPIN_ERR_ARG, "bad argument");
    PS_REQUIRE(st->pt_NT >= 2, PSPIN_ERR_STATE, "pspin_pt_init has not been called");
    int done = 0;
    // Whole rounds (pt_freq sweeps + one exchange) go through one cooperative
    // launch or a captured graph of `rounds` rounds (see pspin_set_launch_mode);
    // nothing per sweep may be recorded, and graphs are not timed.
    const int rounds = launch_rounds > 0 ? launch_rounds : std::max(1, kGraphSweeps / pt_freq);
    while (done < n_sweeps) {
        const int left = n_sweeps - done;
        const bool whole = st->pt_sweeps % pt_freq == 0 && left >= rounds * pt_freq && !st->d_log;
        const bool coop = whole && st->launch_mode == PSPIN_LAUNCH_AUTO && ps::pt_rounds_ok(st);
        const bool graph = whole && !coop && !st->prof && st->launch_mode != PSPIN_LAUNCH_DIRECT;
        if (coop || graph) {
            int rc = PSPIN_OK;
            if (coop) {
                if ((rc = rounds_launch(st, pt_freq, rounds))) return rc;
            } else {
                if ((rc = ensure_graph(st, pt_freq, rounds))) return rc;
                PS_TRY_HIP(hipGraphLaunch(st->graph, st->stream));
                st->sweep_counter += (uint64_t)(rounds * pt_freq);
            }
            const int nsw = rounds * pt_freq;
            st->pt_sweeps += nsw;
            done += nsw;
            st->since_resync += nsw;
            if ((rc = maybe_resync(st))) return rc;
            continue;
        }
        // Sweeps until the next exchange point ((s+1) % pt_freq == 0).
        long long to_next = pt_freq - (st->pt_sweeps % pt_freq);
        int chunk = (int)std::min<long long>(to_next, left);
        // The exchange runs in the sweep launch unless the launch is timed on
        // its own or the energies are resynchronised in between.
        const bool at_exchange = (st->pt_sweeps + chunk) % pt_freq == 0;
        const bool resync_due = st->resync_every > 0 && st->since_resync + chunk >= st->resync_every;
        const bool fuse = at_exchange && !resync_due && !st->prof && ps::exchange_fusable(st);
        int rc = sweep_launch(st, chunk, fuse);
        if (rc) return rc;
        st->pt_sweeps += chunk;
        done += chunk;
        st->since_resync += chunk;
        if ((rc = maybe_resync(st))) return rc;
        if (at_exchange && !fuse) {
            rc = exchange_launch(st);
            if (rc) return rc;
        }
    }
    return PSPIN_OK;
}
int pt_check_barrier(pspin_state* st) {
    unsigned int expired = 0;
    PS_TRY_HIP(hipMemcpy(&expired, st->d_done + 2, sizeof(expired), hipMemcpyDeviceToHost));
    PS_REQUIRE(!expired, PSPIN_ERR_HIP, "PT grid barrier wait expired: rounds ran without their exchange");
    return PSPIN_OK;
}
}  // namespace ps

extern "C" {

const char* pspin_last_error(void) { return ps::g_err.c_str(); }

void pspin_default_config(pspin_sim_config* c) {
    std::memset(c, 0, sizeof(*c));
    c->N = 64;
    c->nrep = 1;
    c->T = 1.0;
    c->J = 2.0;
    c->J0 = 0.0;
    c->alpha = 0.5;
    c->alpha0 = 0.5;
    c->mc_iterations = 10000;
    c->seed = 42;
    c->save_freq = 100;
    c->verbose = 0;
    c->label = -1;
    c->dev = -1;
    c->fmc_mode = 0;
    c->gamma = 0.0;
}

int pspin_parse_args(int argc, char** argv, pspin_sim_config* cfg) {
    pspin_default_config(cfg);
    for (int i = 1; i < argc; i++) {
        const char* a = argv[i];
        bool has = i + 1 < argc;
        if (!strcmp(a, "-N") && has) cfg->N = atoi(argv[++i]);
        else if (!strcmp(a, "-nrep") && has) cfg->nrep = atoi(argv[++i]);
        else if (!strcmp(a, "-T") && has) cfg->T = atof(argv[++i]);
        else if (!strcmp(a, "-J") && has) cfg->J = atof(argv[++i]);
        else if (!strcmp(a, "-J0") && has) cfg->J0 = atof(argv[++i]);
        else if (!strcmp(a, "-alpha") && has) cfg->alpha = atof(argv[++i]);
        else if (!strcmp(a, "-alpha0") && has) cfg->alpha0 = atof(argv[++i]);
        else if (!strcmp(a, "-iter") && has) cfg->mc_iterations = atoi(argv[++i]);
        else if (!strcmp(a, "-seed") && has) cfg->seed = strtoull(argv[++i], nullptr, 10);
        else if (!strcmp(a, "-save_freq") && has) cfg->save_freq = atoi(argv[++i]);
        else if (!strcmp(a, "-label") && has) cfg->label = atoi(argv[++i]);
        else if (!strcmp(a, "-dev") && has) cfg->dev = atoi(argv[++i]);
        else if (!strcmp(a, "-fmc") && has) cfg->fmc_mode = atoi(argv[++i]);
        else if (!strcmp(a, "-gamma") && has) cfg->gamma = atof(argv[++i]);
        else if (!strcmp(a, "-verbose")) {
            if (i + 1 < argc && argv[i + 1][0] >= '0' && argv[i + 1][0] <= '9') cfg->verbose = atoi(argv[++i]);
            else cfg->verbose = 1;
        } else {
            ps::set_error(std::string("Unknown argument: ") + a);
            return PSPIN_ERR_ARG;
        }
    }
    if (cfg->N < 4) {
        ps::set_error("Error: N must be >= 4 for 4-body interactions.");
        return PSPIN_ERR_ARG;
    }
    if (cfg->fmc_mode > 0 && cfg->gamma <= 0.0) {
        if (cfg->fmc_mode == 1) cfg->gamma = 0.0;
        else if (cfg->fmc_mode == 2) cfg->gamma = 1.0 / (2.0 * (cfg->N - 1));
    }
    return PSPIN_OK;
}

int pspin_create(const pspin_sim_config* cfg, int variant, pspin_state** out) {
    PS_REQUIRE(cfg && out, PSPIN_ERR_ARG, "null argument");
    PS_REQUIRE(cfg->N >= 4 && cfg->N <= ps::kMaxN, PSPIN_ERR_ARG, "N must be in [4, 256]");
    PS_REQUIRE(cfg->nrep >= 1, PSPIN_ERR_ARG, "nrep must be >= 1");
    PS_REQUIRE(variant == PSPIN_DENSE || variant == PSPIN_SPARSE, PSPIN_ERR_ARG, "unknown variant");
    PS_REQUIRE(cfg->fmc_mode >= 0 && cfg->fmc_mode <= 2, PSPIN_ERR_ARG, "fmc_mode must be 0, 1 or 2");
    *out = nullptr;
    pspin_state* st = new pspin_state();
    st->cfg = *cfg;
    st->variant = variant;
    st->N = cfg->N;
    st->nrep = cfg->nrep;
    int rc = PSPIN_OK;
    auto fail = [&](int code) {
        free_all(st);
        delete st;
        return code;
    };
#define PS_CREATE_HIP(call)                                                                             \
    do {                                                                                                \
        hipError_t _e = (call);                                                                         \
        if (_e != hipSuccess) {                                                                         \
            ps::set_error(std::string("HIP error ") + hipGetErrorString(_e) + " (" #call ")");          \
            return fail(PSPIN_ERR_HIP);                                                                 \
        }                                                                                               \
    } while (0)
    if (cfg->dev >= 0) PS_CREATE_HIP(hipSetDevice(cfg->dev));
    PS_CREATE_HIP(hipGetDevice(&st->device));
    PS_CREATE_HIP(hipStreamCreateWithFlags(&st->stream, hipStreamNonBlocking));

    const int N = st->N, nrep = st->nrep;
    // Per-replica seeds (src/mc.cu:418-423).
    st->h_seeds.resize(nrep);
    {
        uint64_t m = cfg->seed;
        for (int r = 0; r < nrep; r++) st->h_seeds[r] = ps::splitmix64(&m);
    }
    const double J2 = (1.0 - cfg->alpha) * cfg->J, J4 = cfg->alpha * cfg->J;
    const double J2_0 = (1.0 - cfg->alpha0) * cfg->J0, J4_0 = cfg->alpha0 * cfg->J0;

    make_omega(*cfg, st->h_omega);
    double* d_omega = nullptr;
    PS_CREATE_HIP(dmalloc(&d_omega, N));
    PS_CREATE_HIP(cpy(st, d_omega, st->h_omega.data(), N * sizeof(double), hipMemcpyHostToDevice));

    if ((rc = setup_g2(st, d_omega, J2, J2_0))) { (void)hipFree(d_omega); return fail(rc); }

    const long long nq = ps::n_quartets(N);
    std::vector<Term> terms;
    if (variant == PSPIN_DENSE) {
        // Dense g4 + single-channel mask, rescale and shift (src/mc.cu:445-502).
        PS_CREATE_HIP(dmalloc(&st->d_g4_64, nq));
        PS_CREATE_HIP(dmalloc(&st->d_mask, nq));
        PS_CREATE_HIP(ps::launch_g4_build(st->d_g4_64, st->d_mask, N, cfg->seed + 2000 + 12345, cfg->fmc_mode, d_omega,
                                          cfg->gamma, st->stream));
        std::vector<uint8_t> hm((size_t)nq);
        PS_CREATE_HIP(hipMemcpyAsync(hm.data(), st->d_mask, nq, hipMemcpyDeviceToHost, st->stream));
        PS_CREATE_HIP(hipStreamSynchronize(st->stream));
        if (cfg->fmc_mode == 0) {
            st->n_quart_active = nq;
        } else {
            long long c = 0;
            for (long long q = 0; q < nq; q++) c += __builtin_popcount(hm[q]);
            st->n_quart_active = c;
        }
        if (st->n_quart_active > 0) {
            double sigma = J4 * std::sqrt((double)N / (double)st->n_quart_active);
            double mean = (J4_0 != 0.0) ? J4_0 * (double)N / (double)st->n_quart_active : 0.0;
            PS_CREATE_HIP(ps::launch_scale_shift(st->d_g4_64, nq, sigma, mean, st->stream));
        }
        std::vector<double> hg((size_t)nq);
        PS_CREATE_HIP(hipMemcpyAsync(hg.data(), st->d_g4_64, nq * sizeof(double), hipMemcpyDeviceToHost, st->stream));
        PS_CREATE_HIP(hipStreamSynchronize(st->stream));
        long long q = 0;
        for (int l = 3; l < N; l++)
            for (int k = 2; k < l; k++)
                for (int j = 1; j < k; j++)
                    for (int i = 0; i < j; i++, q++) {
                        uint8_t m = hm[q];
                        if (!m || hg[q] == 0.0) continue;
                        int ch = (m & 1) ? 0 : ((m & 2) ? 1 : 2);
                        terms.push_back(Term{{i, j, k, l}, ch, hg[q]});
                    }
        st->n_sq = (int)std::min<long long>(nq, 2147483647LL);
        if (cfg->fmc_mode == 0 && N >= 8 && N <= ps::kMaxPairsN) {
            if ((rc = build_pair_layout(st, hg, hm))) { (void)hipFree(d_omega); return fail(rc); }
            st->kernel = PSPIN_KERNEL_PAIRS;
        }
    } else {
        // Sparse: unit g4 + mask, survivors in ascending q*3+ch, partial
        // Fisher-Yates with splitmix64(seed+4000), N quartets
        // (src/mc_sparse.cu:408-549).
        double* d_g4 = nullptr;
        uint8_t* d_m = nullptr;
        PS_CREATE_HIP(dmalloc(&d_g4, nq));
        PS_CREATE_HIP(dmalloc(&d_m, nq));
        PS_CREATE_HIP(ps::launch_g4_build(d_g4, d_m, N, cfg->seed + 2000 + 12345, cfg->fmc_mode, d_omega, cfg->gamma,
                                          st->stream));
        std::vector<double> hg((size_t)nq);
        std::vector<uint8_t> hm((size_t)nq);
        PS_CREATE_HIP(hipMemcpyAsync(hg.data(), d_g4, nq * sizeof(double), hipMemcpyDeviceToHost, st->stream));
        PS_CREATE_HIP(hipMemcpyAsync(hm.data(), d_m, nq, hipMemcpyDeviceToHost, st->stream));
        PS_CREATE_HIP(hipStreamSynchronize(st->stream));
        (void)hipFree(d_g4);
        (void)hipFree(d_m);
        std::vector<long long> cand;
        for (long long q = 0; q < nq; q++) {
            if (hg[q] == 0.0) continue;
            for (int ch = 0; ch < 3; ch++)
                if (hm[q] & (1 << ch)) cand.push_back(q * 3 + ch);
        }
        st->n_quart_active = (long long)cand.size();
        int n_sel = N;
        if ((long long)cand.size() < n_sel) {
            fprintf(stderr, "WARNING: only %lld non-zero quartets after FMC, requested %d. Using all.\n",
                    (long long)cand.size(), n_sel);
            n_sel = (int)cand.size();
        }
        uint64_t rs = cfg->seed + 4000;
        for (int i = 0; i < n_sel && i < (int)cand.size(); i++) {
            uint64_t r = ps::splitmix64(&rs);
            long long j = i + (long long)(r % (uint64_t)(cand.size() - i));
            std::swap(cand[i], cand[j]);
        }
        st->n_sq = n_sel;
        const double sigma = n_sel > 0 ? J4 * std::sqrt((double)N / (double)n_sel) : 0.0;
        const double mean = (J4_0 != 0.0 && n_sel > 0) ? J4_0 * (double)N / (double)n_sel : 0.0;
        st->h_sq.resize((size_t)5 * n_sel);
        st->h_sqg.resize(n_sel);
        for (int s = 0; s < n_sel; s++) {
            long long enc = cand[s];
            int ch = (int)(enc % 3);
            long long q = enc / 3;
            // colex decode (src/mc_sparse.cu:320-335)
            int l = 3;
            while ((long long)(l + 1) * l * (l - 1) * (l - 2) / 24 <= q) l++;
            long long rem = q - (long long)l * (l - 1) * (l - 2) * (l - 3) / 24;
            int k = 2;
            while ((long long)(k + 1) * k * (k - 1) / 6 <= rem) k++;
            rem -= (long long)k * (k - 1) * (k - 2) / 6;
            int j = 1;
            while ((long long)(j + 1) * j / 2 <= rem) j++;
            rem -= (long long)j * (j - 1) / 2;
            int v[4] = {(int)rem, j, k, l};
            for (int p = 0; p < 4; p++) st->h_sq[5 * s + p] = v[p];
            st->h_sq[5 * s + 4] = ch;
            st->h_sqg[s] = hg[q] * sigma + mean;
            terms.push_back(Term{{v[0], v[1], v[2], v[3]}, ch, st->h_sqg[s]});
        }
        PS_CREATE_HIP(dmalloc(&st->d_sq, (size_t)5 * n_sel));
        PS_CREATE_HIP(dmalloc(&st->d_sqg, (size_t)n_sel));
        if (n_sel > 0) {
            PS_CREATE_HIP(cpy(st, st->d_sq, st->h_sq.data(), st->h_sq.size() * sizeof(int), hipMemcpyHostToDevice));
            PS_CREATE_HIP(cpy(st, st->d_sqg, st->h_sqg.data(), n_sel * sizeof(double), hipMemcpyHostToDevice));
        }
    }
    (void)hipFree(d_omega);
    if ((rc = build_site_lists(st, terms))) return fail(rc);
    if (variant == PSPIN_DENSE) {
        if ((rc = build_site_groups(st, terms))) return fail(rc);
        if (st->kernel != PSPIN_KERNEL_PAIRS) st->kernel = PSPIN_KERNEL_GROUPS;
    }
    terms.clear();
    terms.shrink_to_fit();

    // Replicas (src/mc.cu:506-544).
    PS_CREATE_HIP(dmalloc(&st->d_spins, (size_t)2 * N * nrep));
    PS_CREATE_HIP(dmalloc(&st->d_seeds, nrep));
    PS_CREATE_HIP(cpy(st, st->d_seeds, st->h_seeds.data(), nrep * sizeof(uint64_t), hipMemcpyHostToDevice));
    PS_CREATE_HIP(ps::launch_init_spins(st->d_spins, N, nrep, st->d_seeds, variant == PSPIN_SPARSE, st->stream));
    PS_CREATE_HIP(dmalloc(&st->d_ctr, nrep));
    PS_CREATE_HIP(hipMemsetAsync(st->d_ctr, 0, nrep * sizeof(uint64_t), st->stream));
    PS_CREATE_HIP(dmalloc(&st->d_E, nrep));
    PS_CREATE_HIP(dmalloc(&st->d_e2, nrep));
    PS_CREATE_HIP(dmalloc(&st->d_e4, nrep));
    if (variant == PSPIN_DENSE) {
        st->n_echunk = ps::energy_chunk_count(N, nrep);
        const std::vector<int> ch = ps::energy_chunks(N, st->n_echunk);
        PS_CREATE_HIP(dmalloc(&st->d_echunk, ch.size()));
        PS_CREATE_HIP(cpy(st, st->d_echunk, ch.data(), ch.size() * sizeof(int), hipMemcpyHostToDevice));
        PS_CREATE_HIP(dmalloc(&st->d_epart, (size_t)st->n_echunk * nrep));
    }
    PS_CREATE_HIP(dmalloc(&st->d_betas, nrep));
    PS_CREATE_HIP(dmalloc(&st->d_acc, nrep));
    PS_CREATE_HIP(dmalloc(&st->d_prop, nrep));
    PS_CREATE_HIP(hipMemsetAsync(st->d_acc, 0, nrep * sizeof(long long), st->stream));
    PS_CREATE_HIP(hipMemsetAsync(st->d_prop, 0, nrep * sizeof(long long), st->stream));
    {
        std::vector<double> b(nrep, cfg->T > 0 ? 1.0 / cfg->T : 0.0);
        PS_CREATE_HIP(cpy(st, st->d_betas, b.data(), nrep * sizeof(double), hipMemcpyHostToDevice));
    }
    if ((rc = recompute_energies(st))) return fail(rc);
    PS_CREATE_HIP(hipStreamSynchronize(st->stream));
#undef PS_CREATE_HIP
    *out = st;
    return PSPIN_OK;
}

int pspin_destroy(pspin_state* st) {
    if (!st) return PSPIN_OK;
    if (st->stream) (void)hipStreamSynchronize(st->stream);
    free_all(st);
    delete st;
    return PSPIN_OK;
}

int pspin_get_info(const pspin_state* st, pspin_info* info) {
    PS_REQUIRE(st && info, PSPIN_ERR_ARG, "null argument");
    info->N = st->N;
    info->nrep = st->nrep;
    info->variant = st->variant;
    info->h2_active = st->h2_active;
    info->n_pairs_active = st->n_pairs_active;
    info->n_quart_active = st->n_quart_active;
    info->n_quartets = st->n_sq;
    info->n_site_entries = st->n_ent;
    info->coupling_format = st->coupling_format;
    info->kernel = st->kernel;
    return PSPIN_OK;
}

int pspin_set_betas(pspin_state* st, const double* betas) {
    PS_REQUIRE(st && betas, PSPIN_ERR_ARG, "null argument");
    PS_TRY_HIP(hipMemcpyAsync(st->d_betas, betas, st->nrep * sizeof(double), hipMemcpyHostToDevice, st->stream));
    PS_TRY_HIP(hipStreamSynchronize(st->stream));
    return PSPIN_OK;
}

int pspin_set_uniform_beta(pspin_state* st, double beta) {
    PS_REQUIRE(st, PSPIN_ERR_ARG, "null argument");
    std::vector<double> b(st->nrep, beta);
    return pspin_set_betas(st, b.data());
}

int pspin_sweep(pspin_state* st, int n_sweeps) {
    PS_REQUIRE(st && n_sweeps >= 0, PSPIN_ERR_ARG, "bad argument");
    if (n_sweeps == 0) return PSPIN_OK;
    int done = 0;
    while (done < n_sweeps) {
        // Split at resync points so the FP64 resync lands every resync_every sweeps.
        int chunk = n_sweeps - done;
        if (st->resync_every > 0) chunk = (int)std::min<long long>(chunk, st->resync_every - st->since_resync);
        if (chunk <= 0) chunk = 1;
        int rc = sweep_launch(st, chunk);
        if (rc) return rc;
        done += chunk;
        st->since_resync += chunk;
        if ((rc = maybe_resync(st))) return rc;
    }
    PS_TRY_HIP(hipStreamSynchronize(st->stream));
    return PSPIN_OK;
}

int pspin_get_results(const pspin_state* st, double* E, long long* acc, long long* prop) {
    PS_REQUIRE(st, PSPIN_ERR_ARG, "null argument");
    if (E) PS_TRY_HIP(hipMemcpyAsync(E, st->d_E, st->nrep * sizeof(double), hipMemcpyDeviceToHost, st->stream));
    if (acc) PS_TRY_HIP(hipMemcpyAsync(acc, st->d_acc, st->nrep * sizeof(long long), hipMemcpyDeviceToHost, st->stream));
    if (prop)
        PS_TRY_HIP(hipMemcpyAsync(prop, st->d_prop, st->nrep * sizeof(long long), hipMemcpyDeviceToHost, st->stream));
    PS_TRY_HIP(hipStreamSynchronize(st->stream));
    return PSPIN_OK;
}

int pspin_reset_counters(pspin_state* st) {
    PS_REQUIRE(st, PSPIN_ERR_ARG, "null argument");
    PS_TRY_HIP(hipMemsetAsync(st->d_acc, 0, st->nrep * sizeof(long long), st->stream));
    PS_TRY_HIP(hipMemsetAsync(st->d_prop, 0, st->nrep * sizeof(long long), st->stream));
    PS_TRY_HIP(hipStreamSynchronize(st->stream));
    return PSPIN_OK;
}

int pspin_get_spins(const pspin_state* st, double* re_im) {
    PS_REQUIRE(st && re_im, PSPIN_ERR_ARG, "null argument");
    PS_TRY_HIP(hipMemcpyAsync(re_im, st->d_spins, (size_t)2 * st->N * st->nrep * sizeof(double),
                              hipMemcpyDeviceToHost, st->stream));
    PS_TRY_HIP(hipStreamSynchronize(st->stream));
    return PSPIN_OK;
}

int pspin_set_spins(pspin_state* st, const double* re_im) {
    PS_REQUIRE(st && re_im, PSPIN_ERR_ARG, "null argument");
    PS_TRY_HIP(hipMemcpyAsync(st->d_spins, re_im, (size_t)2 * st->N * st->nrep * sizeof(double),
                              hipMemcpyHostToDevice, st->stream));
    int rc = recompute_energies(st);
    if (rc) return rc;
    PS_TRY_HIP(hipStreamSynchronize(st->stream));
    return PSPIN_OK;
}

int pspin_split_energies(const pspin_state* st, double* e2, double* e4) {
    PS_REQUIRE(st, PSPIN_ERR_ARG, "null argument");
    pspin_state* s = const_cast<pspin_state*>(st);
    if (s->variant == PSPIN_DENSE)
        PS_TRY_HIP(ps::launch_energy_dense(s, s->d_e2, s->d_e4));
    else
        PS_TRY_HIP(ps::launch_energy_sparse(s, s->d_e2, s->d_e4));
    if (e2) PS_TRY_HIP(hipMemcpyAsync(e2, s->d_e2, s->nrep * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    if (e4) PS_TRY_HIP(hipMemcpyAsync(e4, s->d_e4, s->nrep * sizeof(double), hipMemcpyDeviceToHost, s->stream));
    PS_TRY_HIP(hipStreamSynchronize(s->stream));
    return PSPIN_OK;
}

int pspin_resync_energy(pspin_state* st) {
    PS_REQUIRE(st, PSPIN_ERR_ARG, "null argument");
    int rc = recompute_energies(st);
    if (rc) return rc;
    PS_TRY_HIP(hipStreamSynchronize(st->stream));
    return PSPIN_OK;
}

int pspin_set_energies(pspin_state* st, const double* E) {
    PS_REQUIRE(st && E, PSPIN_ERR_ARG, "null argument");
    PS_TRY_HIP(hipMemcpyAsync(st->d_E, E, st->nrep * sizeof(double), hipMemcpyHostToDevice, st->stream));
    PS_TRY_HIP(hipStreamSynchronize(st->stream));
    return PSPIN_OK;
}

int pspin_get_omega(const pspin_state* st, double* omega) {
    PS_REQUIRE(st && omega, PSPIN_ERR_ARG, "null argument");
    std::memcpy(omega, st->h_omega.data(), st->N * sizeof(double));
    return PSPIN_OK;
}

int pspin_get_g2(const pspin_state* st, double* g2) {
    PS_REQUIRE(st && g2, PSPIN_ERR_ARG, "null argument");
    PS_TRY_HIP(hipMemcpyAsync(g2, st->d_g2_64, (size_t)st->N * st->N * sizeof(double), hipMemcpyDeviceToHost,
                              st->stream));
    PS_TRY_HIP(hipStreamSynchronize(st->stream));
    return PSPIN_OK;
}

int pspin_get_g4(const pspin_state* st, double* g4, uint8_t* mask) {
    PS_REQUIRE(st, PSPIN_ERR_ARG, "null argument");
    PS_REQUIRE(st->variant == PSPIN_DENSE, PSPIN_ERR_STATE, "dense couplings exist only for PSPIN_DENSE");
    long long nq = ps::n_quartets(st->N);
    if (g4) PS_TRY_HIP(hipMemcpyAsync(g4, st->d_g4_64, nq * sizeof(double), hipMemcpyDeviceToHost, st->stream));
    if (mask) PS_TRY_HIP(hipMemcpyAsync(mask, st->d_mask, nq, hipMemcpyDeviceToHost, st->stream));
    PS_TRY_HIP(hipStreamSynchronize(st->stream));
    return PSPIN_OK;
}

int pspin_get_sparse_quartets(const pspin_state* st, int* quartets, double* gq) {
    PS_REQUIRE(st, PSPIN_ERR_ARG, "null argument");
    PS_REQUIRE(st->variant == PSPIN_SPARSE, PSPIN_ERR_STATE, "sparse quartets exist only for PSPIN_SPARSE");
    if (quartets) std::memcpy(quartets, st->h_sq.data(), st->h_sq.size() * sizeof(int));
    if (gq) std::memcpy(gq, st->h_sqg.data(), st->h_sqg.size() * sizeof(double));
    return PSPIN_OK;
}

// ----------------------------- parallel tempering -------------------------
int pspin_pt_init(pspin_state* st, int NT, int nrep_per_temp, const double* beta_sched, long seed48) {
    PS_REQUIRE(st && beta_sched, PSPIN_ERR_ARG, "null argument");
    PS_REQUIRE(NT >= 2 && nrep_per_temp >= 1 && (long long)NT * nrep_per_temp == st->nrep, PSPIN_ERR_ARG,
               "NT * nrep_per_temp must equal nrep and NT >= 2");
    PS_REQUIRE(ps::exchange_kernel_lds_bytes(NT, nrep_per_temp) <= 60 * 1024, PSPIN_ERR_ARG,
               "too many replicas for the exchange kernel (NT * nrep_per_temp <= about 2100 with up to 64 "
               "ladders, 5000 with more)");
    drop_graph(st);
    dfree(st->d_beta_sched); dfree(st->d_perm); dfree(st->d_ex); dfree(st->d_rng48); dfree(st->d_done); dfree(st->d_hand);
    st->pt_NT = NT;
    st->pt_nrep = nrep_per_temp;
    st->pt_sweeps = 0;
    PS_TRY_HIP(dmalloc(&st->d_beta_sched, NT));
    PS_TRY_HIP(dmalloc(&st->d_perm, (size_t)NT * nrep_per_temp));
    PS_TRY_HIP(dmalloc(&st->d_ex, (size_t)4 * (NT - 1)));
    PS_TRY_HIP(dmalloc(&st->d_rng48, 1));
    PS_TRY_HIP(dmalloc(&st->d_done, 3));
    PS_TRY_HIP(fill0(st, st->d_done, 3 * sizeof(unsigned int)));
    PS_TRY_HIP(dmalloc(&st->d_hand, (size_t)3 * NT * nrep_per_temp));
    PS_TRY_HIP(fill0(st, st->d_hand, (size_t)3 * NT * nrep_per_temp * sizeof(unsigned long long)));
    st->hand_tag = 0;
    std::vector<int> perm((size_t)NT * nrep_per_temp);
    std::vector<double> betas((size_t)NT * nrep_per_temp);
    for (int i = 0; i < NT * nrep_per_temp; i++) {
        perm[i] = i;
        betas[i] = beta_sched[i / nrep_per_temp];
    }
    uint64_t x = ps::srand48_state(seed48);
    PS_TRY_HIP(cpy(st, st->d_beta_sched, beta_sched, NT * sizeof(double), hipMemcpyHostToDevice));
    PS_TRY_HIP(cpy(st, st->d_perm, perm.data(), perm.size() * sizeof(int), hipMemcpyHostToDevice));
    PS_TRY_HIP(cpy(st, st->d_betas, betas.data(), betas.size() * sizeof(double), hipMemcpyHostToDevice));
    PS_TRY_HIP(cpy(st, st->d_rng48, &x, sizeof(uint64_t), hipMemcpyHostToDevice));
    PS_TRY_HIP(fill0(st, st->d_ex, 4 * (NT - 1) * sizeof(long long)));
    return PSPIN_OK;
}

int pspin_pt_run(pspin_state* st, int n_sweeps, int pt_freq) {
    const int rc = ps::pt_advance(st, n_sweeps, pt_freq);
    if (rc) return rc;
    PS_TRY_HIP(hipStreamSynchronize(st->stream));
    return ps::pt_check_barrier(st);
}

int pspin_pt_exchange(pspin_state* st) {
    PS_REQUIRE(st && st->pt_NT >= 2, PSPIN_ERR_STATE, "pspin_pt_init has not been called");
    int rc = exchange_launch(st);
    if (rc) return rc;
    PS_TRY_HIP(hipStreamSynchronize(st->stream));
    return PSPIN_OK;
}

int pspin_pt_get_perm(const pspin_state* st, int* perm) {
    PS_REQUIRE(st && perm && st->pt_NT >= 2, PSPIN_ERR_STATE, "pspin_pt_init has not been called");
    PS_TRY_HIP(hipMemcpyAsync(perm, st->d_perm, (size_t)st->pt_NT * st->pt_nrep * sizeof(int), hipMemcpyDeviceToHost,
                              st->stream));
    PS_TRY_HIP(hipStreamSynchronize(st->stream));
    return PSPIN_OK;
}

int pspin_pt_get_exchange_stats(const pspin_state* st, long long* acc_win, long long* prop_win, long long* acc_tot,
                                long long* prop_tot) {
    PS_REQUIRE(st && st->pt_NT >= 2, PSPIN_ERR_STATE, "pspin_pt_init has not been called");
    const int n = st->pt_NT - 1;
    std::vector<long long> h((size_t)4 * n);
    PS_TRY_HIP(hipMemcpyAsync(h.data(), st->d_ex, h.size() * sizeof(long long), hipMemcpyDeviceToHost, st->stream));
    PS_TRY_HIP(hipStreamSynchronize(st->stream));
    if (acc_win) std::memcpy(acc_win, &h[0], n * sizeof(long long));
    if (prop_win) std::memcpy(prop_win, &h[n], n * sizeof(long long));
    if (acc_tot) std::memcpy(acc_tot, &h[2 * n], n * sizeof(long long));
    if (prop_tot) std::memcpy(prop_tot, &h[3 * n], n * sizeof(long long));
    return PSPIN_OK;
}

int pspin_pt_reset_window(pspin_state* st) {
    PS_REQUIRE(st && st->pt_NT >= 2, PSPIN_ERR_STATE, "pspin_pt_init has not been called");
    PS_TRY_HIP(hipMemsetAsync(st->d_ex, 0, 2 * (st->pt_NT - 1) * sizeof(long long), st->stream));
    PS_TRY_HIP(hipStreamSynchronize(st->stream));
    return PSPIN_OK;
}

int pspin_pt_get_rng48(const pspin_state* st, uint64_t* x48) {
    PS_REQUIRE(st && x48 && st->pt_NT >= 2, PSPIN_ERR_STATE, "pspin_pt_init has not been called");
    PS_TRY_HIP(hipMemcpyAsync(x48, st->d_rng48, sizeof(uint64_t), hipMemcpyDeviceToHost, st->stream));
    PS_TRY_HIP(hipStreamSynchronize(st->stream));
    return PSPIN_OK;
}

int pspin_pt_enable_series(pspin_state* st, int capacity) {
    PS_REQUIRE(st && capacity >= 0 && st->pt_NT >= 2, PSPIN_ERR_STATE, "pspin_pt_init has not been called");
    drop_graph(st);
    dfree(st->d_series);
    dfree(st->d_series_n);
    st->series_cap = capacity;
    if (capacity > 0) {
        PS_TRY_HIP(dmalloc(&st->d_series, (size_t)capacity * st->pt_nrep));
        PS_TRY_HIP(dmalloc(&st->d_series_n, 1));
        PS_TRY_HIP(fill0(st, st->d_series_n, sizeof(long long)));
    }
    return PSPIN_OK;
}

int pspin_pt_get_series(const pspin_state* st, double* out, int* n_out) {
    PS_REQUIRE(st && n_out, PSPIN_ERR_ARG, "null argument");
    *n_out = 0;
    if (st->series_cap <= 0) return PSPIN_OK;
    long long n = 0;
    PS_TRY_HIP(hipMemcpyAsync(&n, st->d_series_n, sizeof(long long), hipMemcpyDeviceToHost, st->stream));
    PS_TRY_HIP(hipStreamSynchronize(st->stream));
    const int cap = st->series_cap, nr = st->pt_nrep;
    const long long have = std::min<long long>(n, cap);
    if (out && have > 0) {
        std::vector<double> h((size_t)cap * nr);
        PS_TRY_HIP(cpy(st, h.data(), st->d_series, h.size() * sizeof(double), hipMemcpyDeviceToHost));
        const long long first = n - have;
        for (long long k = 0; k < have; k++) {
            long long slot = (first + k) % cap;
            std::memcpy(out + k * nr, &h[(size_t)slot * nr], nr * sizeof(double));
        }
    }
    *n_out = (int)have;
    return PSPIN_OK;
}

// ----------------------------- hooks --------------------------------------
int pspin_enable_step_log(pspin_state* st, int capacity_sweeps) {
    PS_REQUIRE(st && capacity_sweeps >= 0, PSPIN_ERR_ARG, "bad argument");
    drop_graph(st);
    dfree(st->d_log);
    st->log_cap = capacity_sweeps;
    st->log_n = 0;
    if (capacity_sweeps > 0)
        PS_TRY_HIP(dmalloc(&st->d_log, (size_t)st->nrep * capacity_sweeps * (st->N / 2)));
    return PSPIN_OK;
}

int pspin_get_step_log(const pspin_state* st, void* out, int* n_sweeps_logged) {
    PS_REQUIRE(st, PSPIN_ERR_ARG, "null argument");
    if (n_sweeps_logged) *n_sweeps_logged = st->log_n;
    if (out && st->d_log)
        PS_TRY_HIP(cpy(st, out, st->d_log, (size_t)st->nrep * st->log_cap * (st->N / 2) * sizeof(ps::StepRec),
                             hipMemcpyDeviceToHost));
    return PSPIN_OK;
}

int pspin_get_rng_info(const pspin_state* st, uint64_t* seeds, uint64_t* sweep_counter) {
    PS_REQUIRE(st, PSPIN_ERR_ARG, "null argument");
    if (seeds) std::memcpy(seeds, st->h_seeds.data(), st->nrep * sizeof(uint64_t));
    if (sweep_counter) *sweep_counter = st->sweep_counter;
    return PSPIN_OK;
}

int pspin_set_profiling(pspin_state* st, int on) {
    PS_REQUIRE(st, PSPIN_ERR_ARG, "null argument");
    int rc = harvest_events(st);
    if (rc) return rc;
    st->prof = on != 0;
    st->prof_sweep_ms = st->prof_exch_ms = 0.0;
    st->prof_n_sweep = st->prof_n_exch = st->prof_n_launch = 0;
    return PSPIN_OK;
}

int pspin_get_profile_launches(pspin_state* st, long long* n_sweep_launches) {
    PS_REQUIRE(st && n_sweep_launches, PSPIN_ERR_ARG, "null argument");
    *n_sweep_launches = st->prof_n_launch;
    return PSPIN_OK;
}

int pspin_set_launch_mode(pspin_state* st, int mode) {
    PS_REQUIRE(st && mode >= PSPIN_LAUNCH_AUTO && mode <= PSPIN_LAUNCH_DIRECT, PSPIN_ERR_ARG,
               "launch mode must be PSPIN_LAUNCH_AUTO, _GRAPHS or _DIRECT");
    st->launch_mode = mode;
    return PSPIN_OK;
}

int pspin_set_wait_limit(pspin_state* st, unsigned polls) {
    PS_REQUIRE(st, PSPIN_ERR_ARG, "null argument");
    st->wait_limit = polls;
    return PSPIN_OK;
}

int pspin_get_profile(pspin_state* st, double* sweep_ms, long long* n_sweep, double* exch_ms, long long* n_exch) {
    PS_REQUIRE(st, PSPIN_ERR_ARG, "null argument");
    int rc = harvest_events(st);
    if (rc) return rc;
    if (sweep_ms) *sweep_ms = st->prof_sweep_ms;
    if (n_sweep) *n_sweep = st->prof_n_sweep;
    if (exch_ms) *exch_ms = st->prof_exch_ms;
    if (n_exch) *n_exch = st->prof_n_exch;
    return PSPIN_OK;
}

static const char* kernel_name(int k) {
    return k == PSPIN_KERNEL_PAIRS ? "pair-row" : k == PSPIN_KERNEL_GROUPS ? "group" : "list";
}

// BF16 hi/lo split of an FP32 coupling: hi = bf16(g) (round to nearest even),
// lo = bf16(g - hi); packed hi | lo << 16.
static uint32_t bf16_rne(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    const uint32_t lsb = (u >> 16) & 1u;
    return (u + 0x7fffu + lsb) >> 16;
}

int pspin_set_coupling_format(pspin_state* st, int format) {
    PS_REQUIRE(st, PSPIN_ERR_ARG, "null argument");
    PS_REQUIRE(format == PSPIN_COUPLING_F32 || format == PSPIN_COUPLING_BF16X2, PSPIN_ERR_ARG, "unknown format");
    PS_REQUIRE(format == PSPIN_COUPLING_F32 || st->kernel == PSPIN_KERNEL_LISTS || st->kernel == PSPIN_KERNEL_GROUPS,
               PSPIN_ERR_STATE,
               "BF16 hi/lo couplings are served by the list and group kernels (pspin_set_kernel(.., "
               "PSPIN_KERNEL_LISTS or PSPIN_KERNEL_GROUPS, ..))");
    auto enc = [format](uint32_t& w) {
        if (format != PSPIN_COUPLING_BF16X2) return;
        float g;
        std::memcpy(&g, &w, 4);
        const uint32_t hi = bf16_rne(g);
        float hf;
        const uint32_t hb = hi << 16;
        std::memcpy(&hf, &hb, 4);
        const uint32_t lo = bf16_rne(g - hf);
        w = hi | (lo << 16);
    };
    std::vector<uint2> ent = st->h_site_ent;
    for (uint2& e : ent) enc(e.y);
    if (!ent.empty())
        PS_TRY_HIP(cpy(st, st->d_site_ent, ent.data(), ent.size() * sizeof(uint2), hipMemcpyHostToDevice));
    std::vector<uint4> gw = st->h_grp_g;
    for (uint4& g : gw) {
        enc(g.x);
        enc(g.y);
        enc(g.z);
        enc(g.w);
    }
    if (!gw.empty())
        PS_TRY_HIP(cpy(st, st->d_grp_g, gw.data(), gw.size() * sizeof(uint4), hipMemcpyHostToDevice));
    drop_graph(st);
    st->coupling_format = format;
    return PSPIN_OK;
}

int pspin_set_resync_interval(pspin_state* st, int sweeps) {
    PS_REQUIRE(st && sweeps >= 0, PSPIN_ERR_ARG, "bad argument");
    st->resync_every = sweeps;
    st->since_resync = 0;
    return PSPIN_OK;
}

int pspin_device_info(int dev, size_t* free_bytes, size_t* total_bytes, char* name, int name_len) {
    int cur = 0;
    PS_TRY_HIP(hipGetDevice(&cur));
    if (dev >= 0 && dev != cur) PS_TRY_HIP(hipSetDevice(dev));
    size_t f = 0, t = 0;
    PS_TRY_HIP(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    if (name && name_len > 0) {
        hipDeviceProp_t prop;
        PS_TRY_HIP(hipGetDeviceProperties(&prop, dev >= 0 ? dev : cur));
        std::snprintf(name, (size_t)name_len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    return PSPIN_OK;
}

// Sweep autotuner (the reference picks a block size from the occupancy API
// alone, src/mc.cu:709-744, and prints an [auto-tune] line).  This one times
// every candidate on the live state -- the list kernel at 64..1024 threads and,
// for fully connected systems, the pair-row kernel at 4..16 waves -- and
// restores spins, energies, counters and RNG counters after each trial, so a
// call leaves the chain exactly as it found it.
int pspin_autotune(pspin_state* st, int trial_sweeps, int verbose, int* kernel, int* block, double* ms_per_sweep) {
    PS_REQUIRE(st && trial_sweeps >= 1, PSPIN_ERR_ARG, "bad argument");
    PS_REQUIRE(!st->d_log, PSPIN_ERR_STATE, "disable the step log before autotuning");
    const size_t ns = (size_t)st->nrep * 2 * st->N;
    double *sv = nullptr, *ev = nullptr;
    long long *av = nullptr, *pv = nullptr;
    uint64_t* cv = nullptr;
    auto cleanup = [&]() { dfree(sv); dfree(ev); dfree(av); dfree(pv); dfree(cv); };
    hipError_t e = dmalloc(&sv, ns);
    if (!e) e = dmalloc(&ev, st->nrep);
    if (!e) e = dmalloc(&av, st->nrep);
    if (!e) e = dmalloc(&pv, st->nrep);
    if (!e) e = dmalloc(&cv, st->nrep);
    auto save = [&](bool to_saved) {
        auto c = [&](void* d, const void* s2, size_t b) { return hipMemcpyAsync(d, s2, b, hipMemcpyDeviceToDevice, st->stream); };
        hipError_t r = hipSuccess;
        if (to_saved) {
            if (!r) r = c(sv, st->d_spins, ns * 8);
            if (!r) r = c(ev, st->d_E, st->nrep * 8);
            if (!r) r = c(av, st->d_acc, st->nrep * 8);
            if (!r) r = c(pv, st->d_prop, st->nrep * 8);
            if (!r) r = c(cv, st->d_ctr, st->nrep * 8);
        } else {
            if (!r) r = c(st->d_spins, sv, ns * 8);
            if (!r) r = c(st->d_E, ev, st->nrep * 8);
            if (!r) r = c(st->d_acc, av, st->nrep * 8);
            if (!r) r = c(st->d_prop, pv, st->nrep * 8);
            if (!r) r = c(st->d_ctr, cv, st->nrep * 8);
        }
        return r;
    };
    if (!e) e = save(true);
    if (e) {
        cleanup();
        ps::set_error(std::string("HIP error in pspin_autotune: ") + hipGetErrorString(e));
        return PSPIN_ERR_HIP;
    }
    struct Cand { int kernel, param, rows; };
    std::vector<Cand> cands;
    const bool bf16 = st->coupling_format == PSPIN_COUPLING_BF16X2;
    if (st->d_ptab && !bf16 && st->variant == PSPIN_DENSE)
        for (int rows : {4, 8})
            for (int nw : {4, 8, 12, 16}) cands.push_back({PSPIN_KERNEL_PAIRS, nw, rows});
    for (int lb : {8, 16}) {
        if (lb == 16 && st->variant != PSPIN_DENSE) continue;
        for (int bs : {64, 128, 256, 512, 1024}) cands.push_back({PSPIN_KERNEL_LISTS, bs, lb});
        if (st->d_grp_off)
            for (int bs : {128, 256, 512, 1024}) cands.push_back({PSPIN_KERNEL_GROUPS, bs, lb});
    }
    const int old_kernel = st->kernel, old_block = st->block_size, old_nw = st->pair_nwaves,
              old_rows = st->pair_rows, old_lb = st->list_batch;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    double best = 1e300;
    Cand pick = {old_kernel, old_kernel == PSPIN_KERNEL_PAIRS ? old_nw : old_block,
                 old_kernel == PSPIN_KERNEL_PAIRS ? old_rows : old_lb};
    int rc = PSPIN_OK;
    for (const Cand& c : cands) {
        if (c.kernel == PSPIN_KERNEL_PAIRS) {
            rc = build_pair_tasks(st, c.param, c.rows);
            if (rc == PSPIN_ERR_ARG) {  // work list does not fit this shape
                rc = PSPIN_OK;
                continue;
            }
            if (rc) break;
        } else {
            st->block_size = c.param;
            st->list_batch = c.rows;
        }
        st->kernel = c.kernel;
        hipError_t r = ps::launch_sweep(st, 1);  // warm-up
        if (!r) r = hipEventRecord(e0, st->stream);
        if (!r) r = ps::launch_sweep(st, trial_sweeps);
        if (!r) r = hipEventRecord(e1, st->stream);
        if (!r) r = hipEventSynchronize(e1);
        float ms = 0.0f;
        if (!r) r = hipEventElapsedTime(&ms, e0, e1);
        if (!r) r = save(false);
        if (r) {
            ps::set_error(std::string("HIP error in pspin_autotune: ") + hipGetErrorString(r));
            rc = PSPIN_ERR_HIP;
            break;
        }
        const double per = ms / trial_sweeps;
        if (verbose)
            std::printf("[auto-tune] sweep N=%d nrep=%d %s %s=%d depth=%d: %.4f ms/sweep\n", st->N, st->nrep,
                        kernel_name(c.kernel), c.kernel == PSPIN_KERNEL_PAIRS ? "waves" : "block",
                        c.param, c.rows, per);
        if (per < best) {
            best = per;
            pick = c;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc == PSPIN_OK) {
        st->kernel = pick.kernel;
        if (pick.kernel == PSPIN_KERNEL_PAIRS) {
            rc = build_pair_tasks(st, pick.param, pick.rows);
        } else {
            st->block_size = pick.param;
            st->list_batch = pick.rows;
            if (st->d_ptab && (st->pair_nwaves != old_nw || st->pair_rows != old_rows))
                rc = build_pair_tasks(st, old_nw, old_rows);
        }
    } else {
        st->kernel = old_kernel;
        st->block_size = old_block;
        st->list_batch = old_lb;
        if (st->d_ptab) (void)build_pair_tasks(st, old_nw, old_rows);
    }
    (void)hipStreamSynchronize(st->stream);
    cleanup();
    drop_graph(st);
    if (rc) return rc;
    if (verbose)
        std::printf("[auto-tune] sweep N=%d -> %s %s=%d depth=%d (%.4f ms/sweep)\n", st->N,
                    kernel_name(pick.kernel),
                    pick.kernel == PSPIN_KERNEL_PAIRS ? "waves" : "block", pick.param, pick.rows, best);
    if (kernel) *kernel = pick.kernel;
    if (block) *block = pick.kernel == PSPIN_KERNEL_PAIRS ? 64 * pick.param : pick.param;
    if (ms_per_sweep) *ms_per_sweep = best;
    return PSPIN_OK;
}

int pspin_set_pair_tiles(pspin_state* st, int nwaves, int rows) {
    PS_REQUIRE(st && st->d_ptab, PSPIN_ERR_STATE, "no pair-row layout (dense fully connected, 8 <= N <= 128)");
    PS_REQUIRE(nwaves >= 1 && nwaves <= 16 && (rows == 4 || rows == 8), PSPIN_ERR_ARG,
               "nwaves in [1, 16], rows 4 or 8");
    drop_graph(st);
    return build_pair_tasks(st, nwaves, rows);
}

int pspin_set_list_batch(pspin_state* st, int batch) {
    PS_REQUIRE(st && (batch == 8 || batch == 16), PSPIN_ERR_ARG, "batch must be 8 or 16");
    PS_REQUIRE(batch == 8 || st->variant == PSPIN_DENSE, PSPIN_ERR_ARG, "the sparse variant uses batch 8");
    drop_graph(st);
    st->list_batch = batch;
    return PSPIN_OK;
}

int pspin_set_pair_schedule(pspin_state* st, int groups) {
    PS_REQUIRE(st && groups >= 0 && groups <= st->nrep, PSPIN_ERR_ARG, "groups must be in [0, nrep]");
    drop_graph(st);
    st->pair_groups = groups;
    return PSPIN_OK;
}

int pspin_get_pair_keys(const pspin_state* st, uint64_t* keys) {
    PS_REQUIRE(st && keys, PSPIN_ERR_ARG, "null argument");
    for (int r = 0; r < st->nrep; r++)
        keys[r] = st->pair_groups > 0 ? ps::pair_group_key(st->cfg.seed, r % st->pair_groups) : 0;
    return PSPIN_OK;
}

int pspin_set_kernel(pspin_state* st, int kernel, int block_size) {
    PS_REQUIRE(st, PSPIN_ERR_ARG, "null argument");
    PS_REQUIRE(kernel == PSPIN_KERNEL_AUTO || kernel == PSPIN_KERNEL_LISTS || kernel == PSPIN_KERNEL_PAIRS ||
                   kernel == PSPIN_KERNEL_GROUPS,
               PSPIN_ERR_ARG, "unknown kernel");
    PS_REQUIRE(kernel != PSPIN_KERNEL_PAIRS || st->d_ptab, PSPIN_ERR_ARG,
               "PSPIN_KERNEL_PAIRS needs a dense fully connected system (fmc_mode 0, 8 <= N <= 128)");
    PS_REQUIRE(kernel != PSPIN_KERNEL_GROUPS || st->d_grp_off, PSPIN_ERR_ARG,
               "PSPIN_KERNEL_GROUPS needs the dense variant");
    PS_REQUIRE(block_size == 0 || (block_size % 64 == 0 && block_size >= 64 && block_size <= 1024), PSPIN_ERR_ARG,
               "block_size must be 0 or a multiple of 64 in [64, 1024]");
    const int k = kernel != PSPIN_KERNEL_AUTO ? kernel
                  : st->d_ptab                ? PSPIN_KERNEL_PAIRS
                  : st->d_grp_off             ? PSPIN_KERNEL_GROUPS
                                              : PSPIN_KERNEL_LISTS;
    PS_REQUIRE(k != PSPIN_KERNEL_PAIRS || st->coupling_format == PSPIN_COUPLING_F32, PSPIN_ERR_STATE,
               "the pair-row kernel stores FP32 couplings; set PSPIN_COUPLING_F32 first");
    drop_graph(st);
    st->kernel = k;
    st->block_size = block_size;
    return PSPIN_OK;
}

}  // extern "C"
