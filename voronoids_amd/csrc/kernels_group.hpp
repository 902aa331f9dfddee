// kernels_group.hpp -- lane-group versions of k_cavity / k_commit: G lanes of a wavefront
// cooperate on one window point, NG = 64 / G points per wavefront, one wavefront per block.
//
// Same algorithm and state as kernels.hpp (the reference path they replace is cited there);
// only the mapping of work to lanes changes, to shorten the dependent-load chain of a point:
//   walk   : the D+1 face tests of a step run on D+1 lanes at once;
//   cavity : breadth-first by levels; the (simplex, face) pairs of a level are spread over the
//            G lanes, the cavity and facet lists live in LDS, new members are deduplicated and
//            appended with group ballots and popcount prefixes;
//   commit : one lane per boundary facet builds the new simplex, one lane per (facet, ridge)
//            rotates around the ridge and finds the neighboring facet by an LDS scan.
#pragma once
#include "kernels.hpp"

namespace vo {

template <int G> struct Group {
    static_assert(G == 4 || G == 8 || G == 16 || G == 32 || G == 64, "group size divides 64");
    static constexpr int NG = 64 / G;
    int lane, gid, gl;
    __device__ Group() {
        lane = (int)__lane_id();
        gid = lane / G;
        gl = lane % G;
    }
    __device__ __forceinline__ unsigned long long ballot(bool p) const {
        const unsigned long long b = __ballot(p);
        if constexpr (G == 64) return b;
        else return (b >> (gid * G)) & ((1ull << G) - 1ull);
    }
    __device__ __forceinline__ int below(unsigned long long m) const {
        return __popcll(m & ((1ull << gl) - 1ull));
    }
    __device__ __forceinline__ int bcast(int v, int k) const { return __shfl(v, gid * G + k, 64); }
};

// LDS list capacities per point (a point whose cavity outgrows them is re-run alone) and the
// least waves per SIMD the cavity kernel is compiled for; uniform 3D data at n = 4M peaks at
// 116 killed / 106 facets
#ifndef VO_GROUP_KC3
#define VO_GROUP_KC3 192
#endif
#ifndef VO_GROUP_FC3
#define VO_GROUP_FC3 256
#endif
#ifndef VO_COMMIT_WAVES
#define VO_COMMIT_WAVES 4    // 3D
#endif
#ifndef VO_COMMIT_WAVES2
#define VO_COMMIT_WAVES2 4   // 2D
#endif
#ifndef VO_CAVITY_WAVES2
#define VO_CAVITY_WAVES2 3   // 2D (the kernel reaches 4 waves at its natural register count)
#endif
#ifndef VO_CAVITY_WAVES
#define VO_CAVITY_WAVES 3
#endif
#ifndef VO_GROUP_KC2
#define VO_GROUP_KC2 64
#endif
#ifndef VO_GROUP_FC2
#define VO_GROUP_FC2 96
#endif
template <int D> constexpr int kGroupKC = D == 3 ? VO_GROUP_KC3 : VO_GROUP_KC2;    // cavity entries per point
template <int D> constexpr int kGroupFC = D == 3 ? VO_GROUP_FC3 : VO_GROUP_FC2;    // boundary facets per point
// Cavity members whose adjacency the breadth-first search keeps in LDS: it arrives with the
// member's coordinates when the member is tested (one run of loads), so expanding the member
// costs no further request; members past this index read their adjacency from the record.
// About 6 KB (3D) / 3 KB (2D) per wavefront, so that LDS does not limit the waves per CU.
template <int D, int G> constexpr int kGroupAC =
    (D == 3 ? 384 : 192) / (64 / G) < kGroupKC<D> ? (D == 3 ? 384 : 192) / (64 / G) : kGroupKC<D>;


// LDS list scans, four entries per ds_read_b128 (entries at or past n are ignored: they may
// hold another point's values)
__device__ __forceinline__ bool lds_contains(const int* list, int n, int x) {
    const int4* q = reinterpret_cast<const int4*>(list);
    for (int j = 0; j < n; j += 4) {
        const int4 w = q[j >> 2];
        if ((w.x == x) | ((j + 1 < n) & (w.y == x)) | ((j + 2 < n) & (w.z == x)) | ((j + 3 < n) & (w.w == x))) return true;
    }
    return false;
}
// index of the first entry equal to x (shift = 0) or whose value >> 2 equals x (shift = 2), -1
template <int SHIFT>
__device__ __forceinline__ int lds_find(const int* list, int n, int x) {
    const int4* q = reinterpret_cast<const int4*>(list);
    for (int j = 0; j < n; j += 4) {
        const int4 w = q[j >> 2];
        if ((w.x >> SHIFT) == x) return j;
        if (j + 1 < n && (w.y >> SHIFT) == x) return j + 1;
        if (j + 2 < n && (w.z >> SHIFT) == x) return j + 2;
        if (j + 3 < n && (w.w >> SHIFT) == x) return j + 3;
    }
    return -1;
}

// LDS of one wavefront of the lane-group cavity search
template <int D, int G> struct alignas(16) CavityLds {
    int4 adj[Group<G>::NG][kGroupAC<D, G>];   // adjacency of the first members of each point's cavity
    int4 na[64];                               // adjacency of each lane's candidate
    int cav[Group<G>::NG][kGroupKC<D>];        // killed lists
    int tmp[64];
};

// One wavefront's share of a round's cavity searches: window slots it.off + kb + lane group for
// kb = it.base, it.base + it.stride, ... below it.lim.  Run by k_cavity_g (one wavefront per
// block) and by the one-workgroup early-round kernel (one share per wavefront).
template <int D, int G>
__device__ __forceinline__ void cavity_pass(const Ctx c, int nwin, unsigned long long key_hi, const SlotIter it,
                                            CavityLds<D, G>& L) {
    using V = VT<D>;
    using F = typename V::F;
    constexpr int M = D + 1;
    constexpr int KC = kGroupKC<D>;
    constexpr int FC = kGroupFC<D>;
    static_assert(KC % 4 == 0, "int4 scans");
    constexpr int AC = kGroupAC<D, G>;
    auto& s_cav = L.cav;
    auto& s_tmp = L.tmp;
    auto& s_adj = L.adj;
    auto& s_na = L.na;
    const Group<G> g;
    const bool lead = g.gl == 0;
    LocalCnt lc = {};
    const WinView wv = win_view(c);
    for (int kb = it.base; kb < it.lim; kb += it.stride) {
    const int slot = it.off + kb + g.gid;
    const bool valid = kb + g.gid < it.lim;
    int st = valid ? ST_OK : ST_IDLE;
    int p = 0, pv = 0, t = 0;
    F pf{};
    int2 tt = make_int2(0, 0);
    unsigned long long key = 0;
#ifdef VO_POINT_PROBE
    const unsigned long long pr_t0 = wall_clock64();
    unsigned pr_walk = 0, pr_hops = 0, pr_tw = 0, pr_src = 0;
    float pr_d0 = 0.f, pr_dh = 0.f;
#endif
    if (st == ST_OK) {
        p = win_point(wv, slot);
        pv = c.M + p;
        key = key_hi | (unsigned)p;
        pf = V::ld(c.X32, pv);
        t = c.cp[p];
#ifdef VO_POINT_PROBE
        if (t < 0) {
            const HintView hv = hint_view(c, p);
            int lv = 0;
            for (int l = 1; l <= hv.levels; ++l)
                if (hv.h[hint_cell<D>(hv, pf, l)] >= 0) lv = l;
            pr_src = 1 + lv;
        }
#endif
        if (t < 0) t = hint_lookup<D>(hint_view(c, p), pf);
#ifdef VO_POINT_PROBE
        {
            F a0[4];
            int4 aa0;
            ld_ax<D>(c, t, aa0, a0);
            const float dx = a0[0].x - pf.x, dy = a0[0].y - pf.y;
            float d2 = dx * dx + dy * dy;
            if constexpr (D == 3) d2 += (a0[0].z - pf.z) * (a0[0].z - pf.z);
            pr_dh = sqrtf(d2);
        }
#endif
        tt = ld_v01<D>(c, t);
        for (int hops = 0; tt.x < 0; ++hops) {
            if (hops > kChainLimit) { if (lead) atomicOr(&c.st->err, ERR_CHAIN); st = ST_BROKEN; break; }
            t = tt.y;
            tt = ld_v01<D>(c, t);
            if (lead) lc.chain++;
#ifdef VO_POINT_PROBE
            pr_hops++;
#endif
        }
    }
    // ---- walk: lane k of the group tests face (r + k) mod M.  A walk far longer than the round's
    // others would hold the whole round: past walk_budget moves the point stops where it is and
    // resumes from there next round (only in windows of 64 or more, whose other points progress)
    bool walking = st == ST_OK;
    bool paused = false;
    const int budget = (c.walk_budget > 0 && nwin >= 64) ? c.walk_budget : 0x7fffffff;
    int came = -1;
    uint32_t rng = mix32((uint32_t)p * 0x9E3779B9u + 0x632BE5ABu);
    F a[4];
    int4 aa = make_int4(-1, -1, -1, -1);
    if (walking) ld_ax<D>(c, t, aa, a);
#ifdef VO_POINT_PROBE
    if (walking) {
        const float dx = a[0].x - pf.x, dy = a[0].y - pf.y;
        float d2 = dx * dx + dy * dy;
        if constexpr (D == 3) d2 += (a[0].z - pf.z) * (a[0].z - pf.z);
        pr_d0 = sqrtf(d2);
    }
#endif
    for (int step = 0; __any(walking); ++step) {
        int s = 1, nb = -1;
        bool test = false;
        if (walking) {
            if (lead) s_adj[g.gid][0] = aa;    // the simplex the walk ends in is member 0
            rng = rng * 1664525u + 1013904223u;
            const int r = (int)((rng >> 16) % M);
            int k = r + g.gl;
            if (k >= M) k -= M;
            if (g.gl < M) {
                nb = comp(aa, k);
                test = !(nb >= 0 && (nb >> 2) == came);
                if (test) s = orient_at<D>(c, a, t, k, pv, pf, lc);
            }
        }
        const unsigned long long und = g.ballot(test && s == SIGN_UNDECIDED);
        const unsigned long long neg = g.ballot(test && s != SIGN_UNDECIDED && s < 0);
        const int first = neg ? __ffsll((unsigned long long)neg) - 1 : 0;
        const int nb_sel = g.bcast(nb, first);
        if (walking) {
            if (lead) lc.walk++;
#ifdef VO_POINT_PROBE
            pr_walk++;
#endif
            if (und) {
                st = ST_DEFER;
                walking = false;
            } else if (!neg) {
                walking = false;
            } else if (nb_sel < 0 || step > kWalkLimit) {
                if (lead) atomicOr(&c.st->err, nb_sel < 0 ? ERR_HULL : ERR_WALK);
                st = ST_BROKEN;
                walking = false;
            } else {
                came = t;
                t = nb_sel >> 2;
                ld_ax<D>(c, t, aa, a);
                if (step + 1 >= budget) {
                    st = ST_LOST;
                    paused = true;
                    walking = false;
                }
            }
        }
    }
    if (st == ST_OK || st == ST_DEFER) {
        if (lead) {
            c.cp[p] = t;
            hint_store<D>(hint_view(c, p), pf, a, t);
        }
    } else if (paused && lead) {
        c.cp[p] = t;          // no hint: t need not be near p
    }
#ifdef VO_POINT_PROBE
    pr_tw = (unsigned)(wall_clock64() - pr_t0);
#endif
    // ---- the containing simplex must conflict (else p repeats one of its vertices)
    int s0 = 0;
    if (st == ST_OK && lead) s0 = conflict_at<D>(c, a, t, pv, pf, lc);
    s0 = g.bcast(s0, 0);
    if (st == ST_OK) {
        if (s0 == SIGN_UNDECIDED) st = ST_DEFER;
        else if (s0 <= 0) { st = ST_DUP; if (lead) lc.dups++; }
    }
    // ---- cavity: level-synchronous breadth-first search
    // With abort_early every member is reserved when it is found; the value the reservation
    // replaced (kold) is examined one pass later, once that pass's loads have returned, so the
    // check adds no wait.  A value below the point's key is a higher-priority point of this round
    // (earlier rounds have larger keys): the point cannot win and stops searching.
    int nk = 0, nf = 0, lo = 0, hi = 0, ci = 0;
    bool bfs = st == ST_OK;
    unsigned long long kold = ~0ull;
    if (bfs) {
        if (lead) {
            s_cav[g.gid][0] = t;
            if (c.abort_early) kold = atomicMin(kres_p<D>(c, t), key);
        }
        nk = 1;
        hi = 1;
    }
    __builtin_amdgcn_wave_barrier();
    while (__any(bfs)) {
        const int items = bfs ? (hi - lo) * M : 0;
        const int i = ci + g.gl;
        const bool valid = bfs && i < items;
        int tet = 0, face = 0, n = -1, s = 0, nb = -1;
        bool hull = false, test = false;
        if (valid) {
            const int mi = lo + i / M;
            tet = s_cav[g.gid][mi];
            face = i % M;
            // the LDS entry, or past AC (rare: large cavities) the record's, read by a volatile
            // load: a plain one let the compiler merge the two addresses into one flat load,
            // slower than an LDS read on the common path
            if (mi < AC) nb = reinterpret_cast<const int*>(s_adj[g.gid])[mi * 4 + face];
            else nb = *reinterpret_cast<const volatile int*>(a_ptr<D>(c, tet, face));
            hull = nb < 0;
            if (!hull) {
                n = nb >> 2;
                test = !lds_contains(s_cav[g.gid], nk, n);
                if (test) {
                    F an[4];
                    int4 na;
                    ld_ax<D>(c, n, na, an);
                    s_na[g.lane] = na;
                    s = conflict_at<D>(c, an, n, pv, pf, lc);
                }
            }
        }
        const unsigned long long und = g.ballot(test && s == SIGN_UNDECIDED);
        const unsigned long long lostm = g.ballot(bfs && kold < key);
        bool isnew = test && s != SIGN_UNDECIDED && s > 0;
        s_tmp[g.lane] = isnew ? n : -1;
        __builtin_amdgcn_wave_barrier();
        if (isnew && lds_contains(s_tmp + g.gid * G, g.gl, n)) isnew = false;
        const bool isfac = valid && (hull || (test && s != SIGN_UNDECIDED && s <= 0));
        const unsigned long long newm = g.ballot(isnew), facm = g.ballot(isfac);
        const int cnew = __popcll(newm), cfac = __popcll(facm);
        if (bfs) {
            if (lostm) {
                st = ST_LOST;
                bfs = false;
            } else if (und) {
                st = ST_DEFER;
                bfs = false;
            } else if (nk + cnew > KC || nf + cfac > FC) {
                st = ST_OVERFLOW;          // re-run alone with global-memory lists
                bfs = false;
            } else {
                if (isnew) {
                    const int at = nk + g.below(newm);
                    s_cav[g.gid][at] = n;
                    if (at < AC) s_adj[g.gid][at] = s_na[g.lane];
                    if (c.abort_early) kold = atomicMin(kres_p<D>(c, n), key);
                }
                if (isfac) {
                    // facets go straight to the point's global list: only the killed list,
                    // which every candidate is checked against, occupies LDS
                    const size_t fi = (size_t)slot * c.fcap + nf + g.below(facm);
                    c.fac[fi] = (tet << 2) | face;
                    c.fnb[fi] = hull ? -1 : nb;
                    // (only the stricter commit rule reads outside reservations)
                    if (!hull && !shared_facets(c)) atomicMin(nres_p<D>(c, n), key);
                }
                nk += cnew;
                nf += cfac;
                ci += G;
                if (ci >= items) {
                    lo = hi;
                    hi = nk;
                    ci = 0;
                    if (lo == hi) bfs = false;
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    if (st == ST_OK) {
        for (int j = g.gl; j < nk; j += G) {
            const int x = s_cav[g.gid][j];
            if (!c.abort_early) atomicMin(kres_p<D>(c, x), key);
            c.cav[(size_t)slot * c.kcap + (j)] = x;
        }
    }
    if (valid && lead) {
        c.nk[slot] = nk;
        c.nf[slot] = nf;
        c.status[slot] = st;
#ifdef VO_POINT_PROBE
        if (c.probe) {
            const unsigned long long pr_t1 = wall_clock64();
            c.probe[3 * (size_t)slot] = make_uint4((unsigned)pr_t0, (unsigned)(pr_t0 >> 32), pr_tw, (unsigned)(pr_t1 - pr_t0));
            c.probe[3 * (size_t)slot + 1] = make_uint4(pr_walk, pr_hops, (unsigned)nk, (unsigned)st | (unsigned)nf << 8);
            c.probe[3 * (size_t)slot + 2] = make_uint4(pr_src, (unsigned)p, __float_as_uint(pr_d0), __float_as_uint(pr_dh));
        }
#endif
    }
    __builtin_amdgcn_wave_barrier();       // the LDS lists are reused by the next slots
    }
    flush_counters(c.cnt, lc);
}

template <int D, int G>
__global__ void __launch_bounds__(64, D == 3 ? VO_CAVITY_WAVES : VO_CAVITY_WAVES2) k_cavity_g(Ctx c, int nwin, unsigned long long key_hi) {
    __shared__ CavityLds<D, G> L;
    // the shared-facet list starts empty for this round's commit (the launch before this one
    // applied the previous round's)
    round_params(c, nwin, key_hi);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        c.st->n_fix = 0;
        // the slots whose outcomes this round's commit writes to pflag, for the next planning
        // launch; an empty round (after a planner halt) keeps the last compaction's count
        c.st->flag_n = c.pflag ? nwin : 0;
        if (nwin > 0) c.st->n_compact = 0;
    }
    cavity_pass<D, G>(c, nwin, key_hi, slot_iter(c, nwin, Group<G>::NG), L);
}

// Partner of ridge k of facet j of a winner's cavity boundary (fv: the new simplices' vertex
// tuples, p first, then the facet's D vertices): the ridge is the facet's vertices other than
// fv[j][k]; the boundary of a cavity is a closed manifold, so exactly one other facet holds all
// of the ridge.  Returns that facet and, in kk, the index (1..D) of its vertex off the ridge,
// which is the face of its new simplex that the two new simplices share; -1 if none is found.
// 3D: the tuple in LDS is the new simplex (p, f1, f2, f3); 2D: the facet (f1, f2) only
// The scan reads four tuples per step (indices clamped to the list, so every read is in it)
// before comparing, one LDS round trip per four facets.
__device__ __forceinline__ int match_ridge(const int4* fv, int nf, int j, int k, int& kk) {
    const int4 a = fv[j];
    const int u = k == 1 ? a.z : a.y;
    const int w = k == 3 ? a.z : a.w;
    for (int j0 = 0; j0 < nf; j0 += 4) {
        int4 o[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) o[r] = fv[j0 + r < nf ? j0 + r : nf - 1];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int jj = j0 + r;
            const bool hy = (o[r].y == u) | (o[r].y == w), hz = (o[r].z == u) | (o[r].z == w);
            const bool hw = (o[r].w == u) | (o[r].w == w);
            if (jj < nf && jj != j && (int)hy + (int)hz + (int)hw == 2) {
                kk = !hy ? 1 : (!hz ? 2 : 3);
                return jj;
            }
        }
    }
    return -1;
}
__device__ __forceinline__ int match_ridge(const int2* fv, int nf, int j, int k, int& kk) {
    const int2 a = fv[j];
    const int u = k == 1 ? a.y : a.x;
    for (int j0 = 0; j0 < nf; j0 += 4) {
        int2 o[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) o[r] = fv[j0 + r < nf ? j0 + r : nf - 1];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int jj = j0 + r;
            if (jj < nf && jj != j && ((o[r].x == u) | (o[r].y == u))) {
                kk = o[r].x == u ? 2 : 1;
                return jj;
            }
        }
    }
    return -1;
}
template <int D> struct FacetTuple { using T = int4; };
template <> struct FacetTuple<2> { using T = int2; };
__device__ __forceinline__ void facet_tuple(int4& out, int4 nv) { out = nv; }
__device__ __forceinline__ void facet_tuple(int2& out, int4 nv) { out = make_int2(nv.y, nv.z); }

// Facet tuples a group keeps in LDS for matching ridges (4 KB per block in 3D, 2 KB in 2D); a
// winner with more facets links its internal faces by rotating around each ridge through its
// killed simplices.
#ifndef VO_RIDGE_MATCH3
#define VO_RIDGE_MATCH3 1
#endif
#ifndef VO_MATCH_BYTES3
#define VO_MATCH_BYTES3 4096
#endif
#ifndef VO_MATCH_MIN3
#define VO_MATCH_MIN3 0      // least tuples per point in 3D (more LDS for the small lane groups)
#endif
template <int D, int G> constexpr int kGroupMC0 =
    (D == 3 ? VO_MATCH_BYTES3 / (16 * (64 / G)) : 2048 / (8 * (64 / G)));
template <int D, int G> constexpr int kGroupMC1 =
    (D == 3 && kGroupMC0<D, G> < VO_MATCH_MIN3) ? VO_MATCH_MIN3 : kGroupMC0<D, G>;
template <int D, int G> constexpr int kGroupMC =
    (D == 3 && !VO_RIDGE_MATCH3) ? 0 : (kGroupMC1<D, G> < kGroupFC<D> ? kGroupMC1<D, G> : kGroupFC<D>);

// the new simplex (p first, then the facet) from its LDS tuple
__device__ __forceinline__ int4 tuple_simplex(int4 fv, int pv) { return make_int4(pv, fv.y, fv.z, fv.w); }
__device__ __forceinline__ int4 tuple_simplex(int2 fv, int pv) { return make_int4(pv, fv.x, fv.y, -1); }

// Coordinates of new simplex N = (p, nv_1 .. nv_D): p's are pf; each facet vertex's come from
// the record of the simplex across the facet, which holds the same vertex and lies outside the
// cavity (no winner rewrites its vertices or coordinates), or, on the super-simplex hull, from
// the vertex table.  Selects only, no dynamically indexed private array.
template <int D>
__device__ __forceinline__ void write_coords_from_outside(const Ctx& c, int N, int4 nv, typename VT<D>::F pf,
                                                          int outside) {
    using F = typename VT<D>::F;
    if (outside < 0) {
        F x[4];
        x[0] = pf;
        x[1] = VT<D>::ld(c.X32, nv.y);
        x[2] = VT<D>::ld(c.X32, nv.z);
        x[3] = D == 3 ? VT<D>::ld(c.X32, nv.w) : pf;
        store_tx<D>(c, N, x);
        return;
    }
    const int O = outside >> 2;
    const int4 ov = ld_v<D>(c, O);
    // position in O of each facet vertex
    int m[3];
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const int vk = comp(nv, k + 1);
        m[k] = ov.x == vk ? 0 : (ov.y == vk ? 1 : (ov.z == vk ? 2 : 3));
    }
    if constexpr (D == 3) {
        // the packed floats of O, vertex m at X[3m .. 3m+2]
        const float4* q = reinterpret_cast<const float4*>(rp<D>(c, O) + Rec<D>::X);
        const float4 f0 = q[0], f1 = q[1], f2 = q[2];
        const float X[12] = {f0.x, f0.y, f0.z, f0.w, f1.x, f1.y, f1.z, f1.w, f2.x, f2.y, f2.z, f2.w};
        float o[9];
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int j = 0; j < 3; ++j)
                o[3 * k + j] = m[k] == 0 ? X[j] : (m[k] == 1 ? X[3 + j] : (m[k] == 2 ? X[6 + j] : X[9 + j]));
        float4* w = reinterpret_cast<float4*>(rp<D>(c, N) + Rec<D>::X);
        w[0] = make_float4(pf.x, pf.y, pf.z, o[0]);
        w[1] = make_float4(o[1], o[2], o[3], o[4]);
        w[2] = make_float4(o[5], o[6], o[7], o[8]);
    } else {
        F ox[4], x[4];
        load_tx<D>(c, O, ox);
        x[0] = pf;
        x[1] = m[0] == 0 ? ox[0] : (m[0] == 1 ? ox[1] : ox[2]);
        x[2] = m[1] == 0 ? ox[0] : (m[1] == 1 ? ox[1] : ox[2]);
        x[3] = pf;
        store_tx<D>(c, N, x);
    }
}

// Two winners of a round may have adjacent cavities (sharing a boundary facet) when the later
// one's point lies outside the sphere of the simplex the earlier one builds on that facet: then
// inserting the earlier point first leaves the later one's cavity unchanged, and committing both
// at once equals inserting them in priority order.  True iff vertex pv lies outside or on the
// sphere of (p0, facet i of simplex t); t is pv's killed simplex on the facet, so (pv, facet as
// replace_first orders it) is positively oriented and p0, across the facet, is positive with two
// facet vertices swapped.  Only the FP32 tier is tried here (the commit kernel's registers are
// few): undecided counts as inside, and the point waits for the next round.
template <int D>
__device__ __forceinline__ bool outside_next_star(const Ctx& c, int t, int i, int p0, int pv, typename VT<D>::F pf,
                                                  LocalCnt& lc) {
    using V = VT<D>;
    if (!(c.filters & 1)) return false;
    const int4 nv = replace_first<D>(ld_v<D>(c, t), i, pv);
    const typename V::F a = V::ld(c.X32, c.M + p0), b = V::ld(c.X32, nv.z), d = V::ld(c.X32, nv.y);
    int s;
    if constexpr (D == 3) s = t1_insphere(a, b, d, V::ld(c.X32, nv.w), pf);
    else s = t1_incircle(a, b, d, pf);
    if (s == SIGN_UNDECIDED) return false;
    lc.t1[1]++;
    return s <= 0;
}

// LDS of one wavefront of the lane-group commit
template <int D, int G> struct alignas(16) CommitLds {
    static constexpr int DW = (kGroupFC<D> * D + 31) / 32;   // one bit per (facet, ridge): internal face linked
    int fac[Group<G>::NG][kGroupFC<D>];                     // facet list, then the new simplices
    typename FacetTuple<D>::T fv[Group<G>::NG][kGroupMC<D, G> > 0 ? kGroupMC<D, G> : 1];   // facet tuples
    unsigned done[Group<G>::NG][DW];
};

// One wavefront's share of a round's commits (slots as in cavity_pass)
template <int D, int G>
__device__ __forceinline__ void commit_pass(const Ctx c, int nwin, unsigned long long key_hi, const SlotIter it,
                                            CommitLds<D, G>& L) {
    constexpr int M = D + 1;
    constexpr int FC = kGroupFC<D>;
    static_assert(FC % 4 == 0, "int4 scans");
    constexpr int DW = CommitLds<D, G>::DW;
    constexpr int MC = kGroupMC<D, G>;
    auto& s_fac = L.fac;
    auto& s_done = L.done;
    auto& s_fv = L.fv;
    const Group<G> g;
    const bool lead = g.gl == 0;
    // winners may share facets (see outside_next_star; linked by k_fix_shared)
    const bool refined = shared_facets(c);
    LocalCnt lc = {};
    const WinView wv = win_view(c);
    for (int kb = it.base; kb < it.lim; kb += it.stride) {
    const int slot = it.off + kb + g.gid;
    const bool valid = kb + g.gid < it.lim;
    int st = ST_IDLE, p = 0, nk = 0, nf = 0;
    unsigned long long key = 0;
    bool ok = true, adj = false;
#ifdef VO_POINT_PROBE
    const unsigned long long pr_t0 = wall_clock64();
#endif
    if (valid) {
        st = c.status[slot];
        // the list sizes load with the status (selected after, not loaded under the status
        // test, which costs a second round trip)
        const int nk0 = c.nk[slot], nf0 = c.nf[slot];
        p = win_point(wv, slot);
        key = key_hi | (unsigned)p;
        nk = st == ST_OK ? nk0 : 0;
        nf = st == ST_OK ? nf0 : 0;
        if (st == ST_OK) {
            // members hold their kill reservation; outside neighbors are killed by no lower key
            // (with reuse, also: members are nobody's outside neighbor with a lower key; without,
            // an outside neighbor killed by a lower key marks its facet (bit 31) for the sphere
            // test below).  Member and facet checks share one pass with unconditional loads
            // (indices clamped to the lists, never empty for an OK point): three dependent loads
            const int n_it = nk > nf ? nk : nf;
            // two entries per lane and pass, their loads interleaved (with few lanes per point,
            // as in large windows, the passes would otherwise run one after another)
            for (int j0 = g.gl; j0 < n_it; j0 += 2 * G) {
                const int j1 = j0 + G;
                const int jk0 = j0 < nk ? j0 : nk - 1, jf0 = j0 < nf ? j0 : nf - 1;
                const int jk1 = j1 < nk ? j1 : nk - 1, jf1 = j1 < nf ? j1 : nf - 1;
                const int x0 = c.cav[(size_t)slot * c.kcap + jk0], x1 = c.cav[(size_t)slot * c.kcap + jk1];
                const int fe0 = c.fac[(size_t)slot * c.fcap + jf0], fe1 = c.fac[(size_t)slot * c.fcap + jf1];
                const int nb0 = c.fnb[(size_t)slot * c.fcap + jf0], nb1 = c.fnb[(size_t)slot * c.fcap + jf1];
                const ulonglong2 r0 = ld_res<D>(c, x0), r1 = ld_res<D>(c, x1);
                const unsigned long long ko0 = *kres_p<D>(c, nb0 >= 0 ? nb0 >> 2 : x0);
                const unsigned long long ko1 = *kres_p<D>(c, nb1 >= 0 ? nb1 >> 2 : x1);
                if (refined) {
                    const bool a0 = j0 < nf && nb0 >= 0 && ko0 < key, a1 = j1 < nf && nb1 >= 0 && ko1 < key;
                    if (j0 < nf) s_fac[g.gid][j0] = a0 ? (fe0 | (int)0x80000000) : fe0;
                    if (j1 < nf) s_fac[g.gid][j1] = a1 ? (fe1 | (int)0x80000000) : fe1;
                    ok = ok & (j0 >= nk || r0.x == key) & (j1 >= nk || r1.x == key);
                    adj = adj | a0 | a1;
                } else {
                    if (j0 < nf) s_fac[g.gid][j0] = fe0;
                    if (j1 < nf) s_fac[g.gid][j1] = fe1;
                    ok = ok & (j0 >= nk || ((r0.x == key) & (r0.y > key))) & (j0 >= nf || nb0 < 0 || ko0 > key);
                    ok = ok & (j1 >= nk || ((r1.x == key) & (r1.y > key))) & (j1 >= nf || nb1 < 0 || ko1 > key);
                }
            }
        }
    }
    // a point that holds every kill but touches a higher-priority cavity: the sphere test on
    // each shared facet (the higher point is the least killer of the simplex across it, the only
    // one of its killers that can win); nsh counts the facets to link after the commit
    int nsh = 0;
    const bool need = refined && g.ballot(adj) != 0 && g.ballot(!ok) == 0 && st == ST_OK;
    __builtin_amdgcn_wave_barrier();
    if (need) {
        const typename VT<D>::F pf = VT<D>::ld(c.X32, c.M + p);
        bool pass = true;
        for (int j0 = 0; j0 < nf; j0 += G) {
            const int j = j0 + g.gl;
            bool sh = false;
            if (j < nf) {
                const int fe = s_fac[g.gid][j];
                if (fe < 0) {
                    sh = true;
                    const int ff = fe & 0x7fffffff;
                    const int nb = c.fnb[(size_t)slot * c.fcap + j];
                    const unsigned long long ko = *kres_p<D>(c, nb >> 2);
                    pass = pass && outside_next_star<D>(c, ff >> 2, ff & 3, (int)(uint32_t)ko, c.M + p, pf, lc);
                }
            }
            nsh += __popcll(g.ballot(sh));
        }
        ok = ok && pass;
    }
    bool win = st == ST_OK && g.ballot(!ok) == 0;
    // entries of the shared-facet list, reserved by the lead; a full list makes the point wait
    int fix_base = 0;
    if (refined) {
        int fb = 0;
        if (win && lead && nsh > 0) {
            fb = atomicAdd(&c.st->n_fix, nsh);
            if (fb + nsh > c.fix_cap) {
                for (int e2 = fb; e2 < c.fix_cap; ++e2) c.fix[e2] = make_int2(-1, -1);
                fb = -1;
            }
        }
        fb = g.bcast(fb, 0);
        win = win && fb >= 0;
        fix_base = fb;
    }
    // Slots.  With reuse (and the facet tuples within the LDS match list), facet j's new simplex
    // takes the winner's killed record j for j < nr = min(nk, nf) and a fresh slot past that, so
    // that records stay about as many as live simplices and new ones land on lines the round has
    // just read; only the fresh ones are allocated (one range per winning group, the lead asks).
    const bool reuse = c.reuse && nf <= MC;
    const int nr = reuse ? (nk < nf ? nk : nf) : 0;
    const int base = g.bcast(alloc_slots(c, win && lead, nf - nr, p), 0);
    const bool build = win && base >= 0;
    if (refined && win && !build && lead)
        for (int e2 = 0; e2 < nsh; ++e2) c.fix[fix_base + e2] = make_int2(-1, -1);
    // a winner's check loops ran to completion, so s_fac holds its whole facet list
    __builtin_amdgcn_wave_barrier();
    const int pv = c.M + p;
    if (build) {
        const typename VT<D>::F pf = VT<D>::ld(c.X32, pv);
        if (reuse) {
            // 1a: every read of the killed records before any of them is rewritten: the new
            // simplex's vertices (its tuple, kept for the ridge match) and the simplex across
            // facet j, which replaces facet j in s_fac
            for (int j = g.gl; j < nf; j += G) {
                const int fe = s_fac[g.gid][j];
                const int t = fe >> 2, i = fe & 3;
                facet_tuple(s_fv[g.gid][j], replace_first<D>(ld_v<D>(c, t), i, pv));
                s_fac[g.gid][j] = *a_ptr<D>(c, t, i);
            }
            __builtin_amdgcn_wave_barrier();
            // 1b: the new simplices, their outside links and the outside neighbors' patches;
            // s_fac[j] becomes facet j's slot
            for (int j = g.gl; j < nf; j += G) {
                const int N = j < nr ? c.cav[(size_t)slot * c.kcap + j] : base + (j - nr);
                const int outside = s_fac[g.gid][j];
                const int4 nv = tuple_simplex(s_fv[g.gid][j], pv);
                st_v<D>(c, N, nv);
                write_coords_from_outside<D>(c, N, nv, pf, outside);
                *a_ptr<D>(c, N, 0) = outside;
                if (outside >= 0) *a_ptr<D>(c, outside >> 2, outside & 3) = N << 2;
                s_fac[g.gid][j] = N;
            }
            __builtin_amdgcn_wave_barrier();
        }
        // phase 1: the new simplex of facet j, its outside link and the outside neighbor's patch;
        // a shared facet (marked) is left unlinked and listed for k_fix_shared instead, and its
        // mark is cleared for the lookups that follow
        int run = 0;
        for (int j0 = 0; !reuse && j0 < nf; j0 += G) {
            const int j = j0 + g.gl;
            const int fm = j < nf ? s_fac[g.gid][j] : 0;
            const bool sh = fm < 0;
            const unsigned long long shm = g.ballot(sh);
            if (j < nf) {
                const int fe = fm & 0x7fffffff;
                const int t = fe >> 2, i = fe & 3;
                // every load of killed record t ahead of the new record's stores, in one round trip
                const int4 tv = ld_v<D>(c, t);
                const int outside = *a_ptr<D>(c, t, i);
                const StarSrc<D> xs = ld_star_src<D>(c, t);
                const int4 nv = replace_first<D>(tv, i, pv);
                const int N = base + j;
                st_v<D>(c, N, nv);
                store_star_coords<D>(c, xs, i, pf, N);
                *a_ptr<D>(c, N, 0) = outside;
                if (sh) c.fix[fix_base + run + g.below(shm)] = make_int2(N, fe);
                else if (outside >= 0) *a_ptr<D>(c, outside >> 2, outside & 3) = N << 2;
                if (nf <= MC) facet_tuple(s_fv[g.gid][j], nv);
                if (sh) s_fac[g.gid][j] = fe;
            }
            run += __popcll(shm);
        }
        __builtin_amdgcn_wave_barrier();
        // phase 2: internal faces, one lane per (facet, ridge); each link is written from both
        // sides by the lane that finds it, and the other side's lane skips it when it comes later
        for (int w = g.gl; w < DW; w += G) s_done[g.gid][w] = 0u;
        __builtin_amdgcn_wave_barrier();
        // matched in LDS: the facet sharing the ridge is the partner (no memory traffic)
        for (int q = g.gl; nf <= MC && q < nf * D; q += G) {
            if (s_done[g.gid][q >> 5] & (1u << (q & 31))) continue;
            const int j = q / D, k = 1 + q % D;
            int kk = 1;
            const int jj = match_ridge(s_fv[g.gid], nf, j, k, kk);
            const int Nj = reuse ? s_fac[g.gid][j] : base + j;
            if (jj < 0) {
                atomicOr(&c.st->err, ERR_WALK);
                *a_ptr<D>(c, Nj, k) = -1;
                continue;
            }
            const int Njj = reuse ? s_fac[g.gid][jj] : base + jj;
            *a_ptr<D>(c, Nj, k) = (Njj << 2) | kk;
            *a_ptr<D>(c, Njj, kk) = (Nj << 2) | k;
            const int qq = jj * D + kk - 1;
            atomicOr(&s_done[g.gid][qq >> 5], 1u << (qq & 31));
        }
        // too many facets for the LDS tuples: rotate around the ridge through the killed simplices
        for (int q = g.gl; nf > MC && q < nf * D; q += G) {
            if (s_done[g.gid][q >> 5] & (1u << (q & 31))) continue;
            const int j = q / D, k = 1 + q % D;
            const int fe = s_fac[g.gid][j];
            const int t = fe >> 2, i = fe & 3;
            const int4 tvt = ld_v<D>(c, t);
            int4 TA = ld_a<D>(c, t);
            const int4 nv = replace_first<D>(tvt, i, pv);
            const int cv = comp(nv, k);
            int T = t;
            int4 TV = tvt;
            int e = index_of<D>(TV, cv);
            int y = comp(tvt, i);
            for (int guard = 0;; ++guard) {
                const int nbe = comp(TA, e);
                if (nbe < 0) break;
                // reservation, vertices and adjacency of the next simplex in one round trip
                const int T2 = nbe >> 2;
                const unsigned long long kr = *kres_p<D>(c, T2);
                const int4 TV2 = ld_v<D>(c, T2);
                const int4 TA2 = ld_a<D>(c, T2);
                if (kr != key) break;
                if (guard > 4096) { atomicOr(&c.st->err, ERR_WALK); break; }
                const int w = comp(TV2, nbe & 3);
                e = index_of<D>(TV2, y);
                y = w;
                T = T2;
                TV = TV2;
                TA = TA2;
            }
            const int target = (T << 2) | e;
            const int jj = lds_find<0>(s_fac[g.gid], nf, target);
            const int4 nv2 = replace_first<D>(TV, e, pv);
            int kk = 1;
            for (int m = 1; m < M; ++m) {
                const int x = comp(nv2, m);
                bool in_ridge = false;
                for (int r = 1; r < M; ++r)
                    if (r != k && comp(nv, r) == x) in_ridge = true;
                if (!in_ridge) { kk = m; break; }
            }
            if (jj < 0) {
                atomicOr(&c.st->err, ERR_WALK);
                *a_ptr<D>(c, base + j, k) = -1;
                continue;
            }
            *a_ptr<D>(c, base + j, k) = ((base + jj) << 2) | kk;
            *a_ptr<D>(c, base + jj, kk) = ((base + j) << 2) | k;
            const int qq = jj * D + kk - 1;
            atomicOr(&s_done[g.gid][qq >> 5], 1u << (qq & 31));
        }
        // phase 3: killed simplices die, forwarding to a new simplex on one of their facets; with
        // reuse only the records past the reused ones die (nk > nf), forwarding to facet 0's
        for (int j = g.gl; !reuse && j < nk; j += G) {
            const int t = c.cav[(size_t)slot * c.kcap + (j)];
            const int r = lds_find<2>(s_fac[g.gid], nf, t);
            const int fwd = base + (r < 0 ? 0 : r);
            kill<D>(c, t, fwd);
        }
        for (int j = nr + g.gl; reuse && j < nk; j += G) kill<D>(c, c.cav[(size_t)slot * c.kcap + (j)], s_fac[g.gid][0]);
        if (lead) {
            lc.killed += nk;
            lc.created += nf;
            lc.wins += 1;
            lc.maxk = nk > (int)lc.maxk ? nk : lc.maxk;
            lc.maxf = nf > (int)lc.maxf ? nf : lc.maxf;
        }
    }
#ifdef VO_POINT_PROBE
    if (c.probe && valid && lead) {
        const unsigned long long pr_t1 = wall_clock64();
        const unsigned flags = (unsigned)st | (unsigned)win << 4 | (unsigned)build << 5 | (unsigned)(nf > MC) << 6 |
                               (unsigned)need << 7 | (unsigned)nsh << 8;
        uint4* q = c.probe + 3 * (size_t)c.win_cap + 2 * (size_t)slot;
        q[0] = make_uint4((unsigned)pr_t0, (unsigned)(pr_t0 >> 32), (unsigned)(pr_t1 - pr_t0), flags);
        q[1] = make_uint4((unsigned)nk, (unsigned)nf, 0u, 0u);
    }
#endif
    // cavities too large for the LDS lists are re-run alone after the round
    const bool big = lead && st == ST_OVERFLOW;
    bool again = lead && ((st == ST_OK && !build) || st == ST_DEFER || st == ST_LOST);
    {
        const unsigned long long bm = __ballot(big);
        int bpos = 0;
        if (g.lane == 0 && bm) bpos = atomicAdd(&c.st->n_big, __popcll(bm));
        bpos = __shfl(bpos, 0, 64) + __popcll(bm & ((1ull << g.lane) - 1ull));
        if (big) {
            if (bpos < c.big_cap) c.big[bpos] = p;
            else again = true;             // list full: retry in the next window
        }
    }
    set_outcome(c, lead && st != ST_IDLE, p, build, nf - nk, !build && again, !build && !again && st == ST_DUP);
    if (again) lc.pend++;
    if (c.pflag) {
        // the outcome at the slot's own position: no returning atomic on one counter per wavefront
        if (lead && valid) c.pflag[slot] = again ? p : -1;
    } else {
        const unsigned long long mask = __ballot(again);
        const int cnt = __popcll(mask);
        int pos0 = 0;
        if (g.lane == 0 && cnt) pos0 = atomicAdd(&c.st->n_next_pending, cnt);
        pos0 = __shfl(pos0, 0, 64);
        if (again) {
            const int pos = pos0 + __popcll(mask & ((1ull << g.lane) - 1ull));
            if (pos < c.pend_cap) c.pend_next[pos] = p;
            else atomicOr(&c.st->err, ERR_PENDING);
        }
    }
    __builtin_amdgcn_wave_barrier();       // s_fac is reused by the next slots
    }
    flush_counters(c.cnt, lc);
}

template <int D, int G>
__global__ void __launch_bounds__(64, D == 3 ? VO_COMMIT_WAVES : VO_COMMIT_WAVES2) k_commit_g(Ctx c, int nwin, unsigned long long key_hi) {
    __shared__ CommitLds<D, G> L;
    round_params(c, nwin, key_hi);
    commit_pass<D, G>(c, nwin, key_hi, slot_iter(c, nwin, Group<G>::NG), L);
}

}  // namespace vo
