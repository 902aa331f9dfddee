// runtime.hip -- host runtime and C ABI of libvoronoids_hip.so (see include/voronoids.h).
//
// The round driver replaces add_points_to_tree / insert_points_parallel
// (src/delaunay_tree.rs:213-386) and the insert_point loop of voronoids.delaunay
// (src/lib.rs:104-125): every round takes a window of uninserted points (the previous
// round's losers first, then fresh points in priority order), runs k_cavity and k_commit,
// resolves deferred predicates exactly when any were queued, and adapts the window size to
// keep about half of the window winning.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/voronoids.h"
#include "host_common.hpp"
#include "kernels.hpp"
#include "kernels_group.hpp"
#include "compact.hpp"
#include "canon.hpp"

#pragma clang fp contract(off)

using namespace vo;

using namespace vo_host;

namespace vo_host {
int canonical_from_simplices(int M, const int32_t* s, int64_t nl, int32_t** T, int64_t* n_T, int32_t** E,
                             int64_t* n_E) {
    const int32_t first = 2 * M;
    std::vector<std::array<int32_t, 4>> tris;
    std::vector<std::array<int32_t, 2>> edges;
    tris.reserve(nl);
    edges.reserve((size_t)nl * (M == 4 ? 6 : 3));
    for (int64_t i = 0; i < nl; ++i) {
        std::array<int32_t, 4> r = {INT32_MAX, INT32_MAX, INT32_MAX, INT32_MAX};
        bool real = true;
        for (int k = 0; k < M; ++k) {
            const int32_t v = s[i * M + k];
            r[k] = v >= first ? v - first : -1;
            real &= v >= first;
        }
        for (int a = 0; a < M; ++a)
            for (int b = a + 1; b < M; ++b)
                if (r[a] >= 0 && r[b] >= 0) edges.push_back({std::min(r[a], r[b]), std::max(r[a], r[b])});
        if (real) {
            std::sort(r.begin(), r.begin() + M);
            tris.push_back(r);
        }
    }
    std::sort(tris.begin(), tris.end());
    std::sort(edges.begin(), edges.end());
    edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
    *T = (int32_t*)std::malloc(sizeof(int32_t) * std::max<size_t>(1, tris.size() * M));
    *E = (int32_t*)std::malloc(sizeof(int32_t) * std::max<size_t>(1, edges.size() * 2));
    if (!*T || !*E) return fail(VO_ERR_OOM, "host allocation");
    for (size_t i = 0; i < tris.size(); ++i)
        for (int k = 0; k < M; ++k) (*T)[i * M + k] = tris[i][k];
    for (size_t i = 0; i < edges.size(); ++i) { (*E)[2 * i] = edges[i][0]; (*E)[2 * i + 1] = edges[i][1]; }
    *n_T = (int64_t)tris.size();
    *n_E = (int64_t)edges.size();
    return 0;
}
}  // namespace vo_host


// Round planning on the device for one triangulation: after each round a one-thread kernel
// absorbs the round's outcome (the host's window rule, the same arithmetic) and sizes the next
// window, so that rounds run back to back while the host only watches snapshots.  A round that
// needs the host (points re-run alone, queued exact predicates, an error) or more simplex
// space halts the planner; the rounds after a halt are empty until the host takes over.
enum : int { HALT_HOST = 1, HALT_SPACE = 2, HALT_DONE = 4, HALT_STALL = 8, HALT_BOOT = 16 };
struct RunPlan {
    BatchCtl ctl;              // nwin / key_hi of the planned round (read by the round kernels)
    long long inserted, dups, windows, n_live, tet_cap;
    double gamma, big_frac_lo, big_frac_hi, small_frac_lo, small_frac_hi;
    int head, n_pts, npend, nfresh, nwin, rounds, stalled, cur;
    int launched;              // a planned round ran and is not absorbed yet
    int halt;                  // HALT_* bits
    int win_cap, big_window, per_point;
    int exact_dev;             // queued exact predicates are evaluated in the stream (k_exact after
                               // each commit) instead of halting for the host
    int nq_done;               // the halted round's queue count, evaluated in the stream
    int min_window;            // least window size
    int boot_limit;            // k_boot: the planner halts (HALT_BOOT) before a window larger than
    int boot_rounds_end;       // this, or once `rounds` reaches boot_rounds_end (0: never)
    int chunk_rate_milli;      // records reserved per window point x 1000 across the allocation
                               // stripes of a round (0: no stripes)
    int chunk;                 // the planned round's chunk per stripe
};

struct vo_tri {
    int dev = 0, D = 3, M = 4, S = 4;
    hipStream_t s = nullptr;
    double c[3] = {0, 0, 0}, r = 0, R = 0;
    double sv[4][3] = {};
    int tet0[4] = {0, 1, 2, 3};
    // points
    int64_t n_pts = 0, vcap = 0, head = 0;
    std::vector<int32_t> input_of;     // point index -> input index
    double* X64 = nullptr;
    float* X32 = nullptr;
    int* cp = nullptr;
    int* vmap = nullptr;               // vertex id -> reference id
    // simplices
    char* rec = nullptr;               // simplex records (128 B in 3D, 64 B in 2D)
    char* rec_spare = nullptr;         // compaction target, same capacity
    int* cmap = nullptr;               // compaction: old record -> new record
    int* chist = nullptr;              // compaction: bin counts / cursors
    int64_t retired = 0;               // records dropped by compactions (max_simplex_id counts them)
    double compact_ratio = 0.0;        // compact when records exceed this many per live simplex (0: never)
    int64_t compact_min = 1 << 20;     // ... and at least this many records are allocated
    int64_t n_compact = 0;
    double ms_compact = 0;
    int xcd = 0;                       // window sorted into 8 spatial regions, one per XCD
    int* reg_meta = nullptr;           // [0..7] first slot, [8..15] count, [16..23] scatter cursors
    int* win2 = nullptr;               // the window in region order
    unsigned char* regs = nullptr;     // region of each window slot
    int64_t tcap = 0;
    int n_tets = 1;
    int64_t n_live = 1;
    // window
    int win_cap = 0, kcap = 0, fcap = 0;
    int* win = nullptr;
    int* pend[2] = {nullptr, nullptr};
    int cur = 0, npend = 0;
    int* cav = nullptr;
    int* fac = nullptr;
#ifdef VO_POINT_PROBE
    uint4* probe = nullptr;            // per-point probe of the cavity kernel (diagnostic builds)
    FILE* probe_f = nullptr;
#endif
    int* fnb = nullptr;                // the outside link of each facet, as the cavity search saw it
    int* nk = nullptr;
    int* nf = nullptr;
    int* status = nullptr;
    // points re-run alone: their list, and large single-point cavity buffers
    int* big = nullptr;
    int* cav1 = nullptr;
    int* fac1 = nullptr;
    int* misc1 = nullptr;              // nk, nf, status of the single point
    int big_cap = 1 << 20;
    // hints
    int* hint = nullptr;
    int hlevels = 0;
    float glo[3] = {0, 0, 0};
    float ginv = 1;
    // control
    RoundState* st = nullptr;
    RoundState* st_host = nullptr;
    Counters* cnt = nullptr;
    QEntry* q = nullptr;
    int qcap = 0;
    CEntry* cache = nullptr;
    uint32_t cmask = 0;
    double* ws = nullptr;
    int ws_cap = 0, ws_threads = 0;
    hipEvent_t ev[6] = {};
    // device-planned rounds: planner state, pinned snapshots (one per chunk in flight) and
    // per-round timing events (cavity start, commit start, commit end)
    RunPlan* plan = nullptr;
    RunPlan* plan_host = nullptr;
    std::vector<hipEvent_t> rev;
    hipEvent_t chunk_ev[2] = {};
    int pipelined = 1;                 // 0: the host plans every round and waits for it
    int chunk_rounds = 4;              // rounds submitted per chunk
    int round_timing = 0;              // events around each round's kernels (cost: ~3% at n = 4M)
    int abort_early = 1;               // losing cavity searches stop at the first reservation they lose
    int walk_budget = 0;               // lane-group walk moves per round before a point pauses (0 = off)
    int alloc_stripes = 1;             // device-planned rounds allocate from per-round striped chunks
    AllocStripe* ast = nullptr;        // their cursors (kAllocStripes)
    int outcome_flags = 1;             // device-planned rounds: the commit writes each slot's outcome
                                       // and the planning launch compacts the pending list
    int* pflag = nullptr;              // those outcomes (win_cap)
    int boot_window = 0;               // first rounds in one workgroup (k_boot) while windows stay at
                                       // or below this many points (0 = off; set per dimension)
    int reuse = -1;                    // new simplices take their winner's killed records first: 1 on,
                                       // 0 off, -1 only when append-only records would pass 2^29
    int reuse_on = 0;                  // this run's choice
    int2* fix = nullptr;               // shared facets of the round's winners (k_fix_shared)
    int fix_cap = 0;
    int shared_facets = 1;             // 0: the stricter rule (winners' cavities never adjacent)
    int min_window = 64;               // least points per round (8 measured ~2% slower at 3D 4M / 2D 10M)
    int64_t small_run = 64;            // runs of at most this many points use the host loop
    int exact_dev = 0;                 // this run has queued exact predicates: evaluate them in the stream
    int64_t stalled = 0;
    double gamma = 0.25;
    int filters = 3;
    int64_t big_window = 0;            // windows at least this large target big_frac_lo..hi winners (0: off)
    double big_frac_lo = 0.55, big_frac_hi = 0.8;
    double small_frac_lo = 0.35, small_frac_hi = 0.6;   // the same for smaller windows
    int local_order = 0;               // priority order: random, Morton-sorted in chunks of position/this (0: off)
    int group = 0;                     // lanes per window point (0 = chosen per round, 1 = one thread)
    int64_t lanes_target = 0;          // auto group size: largest G with window * G <= this (0 = default)
    int64_t commit_lanes_target = 0;   // the same for the commit kernel (0 = lanes_target)
    int64_t inserted = 0, dups = 0, rounds = 0, windows_total = 0;
    int64_t rounds0 = 0;               // rounds before the last vo_tri_reset (keys keep decreasing)
    double ms_run = 0, ms_cav = 0, ms_com = 0, ms_ex = 0, ms_other = 0;
    int64_t n_cav = 0, n_com = 0, n_solo = 0;
    bool dirty_dup = false;
};

namespace {

Ctx make_ctx(vo_tri* t) {
    Ctx c;
    c.rec = t->rec;
    c.X64 = t->X64;
    c.X32 = t->X32;
    c.cp = t->cp;
    c.win = t->win;
    c.pend_next = t->pend[t->cur ^ 1];
    c.big = t->big;
    c.cav = t->cav;
    c.fac = t->fac;
    c.fnb = t->fnb;
    c.nk = t->nk;
    c.nf = t->nf;
    c.status = t->status;
    c.hint = t->hint;
    for (int i = 0; i < 3; ++i) c.glo[i] = t->glo[i];
    c.ginv = t->ginv;
    c.hlevels = t->hlevels;
    c.st = t->st;
    c.cnt = t->cnt;
    c.q = t->q;
    c.qcap = t->qcap;
    c.cache = t->cache;
    c.cmask = t->cmask;
    c.M = t->M;
    c.filters = t->filters;
    c.solo = 0;
    c.abort_early = t->abort_early;
    c.walk_budget = t->walk_budget;
    c.reuse = t->reuse_on;
    c.fix = t->shared_facets ? t->fix : nullptr;
    c.fix_cap = t->fix_cap;
    c.win_cap = t->win_cap;
    c.kcap = t->kcap;
    c.fcap = t->fcap;
    c.tet_cap = (int)t->tcap;
    c.sets = nullptr;
    c.pset = nullptr;
    c.ctl = nullptr;
    c.big_cap = t->win_cap;
    c.pend_cap = t->win_cap;
    c.region = nullptr;
    return c;
}

// next window = pending losers followed by fresh points head .. head + nfresh - 1
// staged points in priority order: slot i takes input point perm[i] (pts row-major, D columns);
// X64 / X32 rows have stride S with the unused components 0, X32 = fl32(x - c) (the double
// difference rounded once, as on the host), vmap = the point's reference vertex id
__global__ void k_stage_points(const double* pts, const int32_t* perm, int64_t n, int D, int S, double c0, double c1,
                               double c2, double* X64, float* X32, int* vmap, int64_t id0) {
    const double cc[3] = {c0, c1, c2};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t src = perm[i];
        for (int k = 0; k < S; ++k) {
            const double x = k < D ? pts[src * D + k] : 0.0;
            X64[i * S + k] = x;
            X32[i * S + k] = k < D ? (float)(x - cc[k]) : 0.0f;
        }
        vmap[i] = (int)(id0 + src);
    }
}

__global__ void k_window(int* win, const int* pend, int npend, int head, int nfresh) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < npend) win[i] = pend[i];
    else if (i < npend + nfresh) win[i] = head + (i - npend);
}

// region of point p: the top three bits of its Morton code over the hint grid's box (3D: the
// octant; 2D: y quarter and x half), the same bits that lead compaction's record order
template <int D> __device__ __forceinline__ int point_region(const Ctx& c, int p) {
    const typename VT<D>::F x = VT<D>::ld(c.X32, c.M + p);
    const float ux = (x.x - c.glo[0]) * c.ginv, uy = (x.y - c.glo[1]) * c.ginv;
    if constexpr (D == 3) {
        const float uz = (x.z - c.glo[2]) * c.ginv;
        return ((uz >= 0.5f) << 2) | ((uy >= 0.5f) << 1) | (ux >= 0.5f);
    } else {
        const int qy = (int)fminf(fmaxf(uy * 4.0f, 0.0f), 3.0f);
        return ((qy >> 1) << 2) | ((ux >= 0.5f) << 1) | (qy & 1);
    }
}

template <int D>
__global__ void __launch_bounds__(256) k_region_count(Ctx c, const int* win, int nwin, unsigned char* regs, int* meta) {
    __shared__ int s_n[8];
    if (threadIdx.x < 8) s_n[threadIdx.x] = 0;
    __syncthreads();
    for (int i = blockIdx.x * 256 + threadIdx.x; i < nwin; i += gridDim.x * 256) {
        const int r = point_region<D>(c, win[i]);
        regs[i] = (unsigned char)r;
        atomicAdd(&s_n[r], 1);
    }
    __syncthreads();
    if (threadIdx.x < 8 && s_n[threadIdx.x]) atomicAdd(&meta[8 + threadIdx.x], s_n[threadIdx.x]);
}

__global__ void k_region_scan(int* meta) {
    int acc = 0;
    for (int r = 0; r < 8; ++r) {
        meta[r] = acc;
        meta[16 + r] = acc;
        acc += meta[8 + r];
    }
}

__global__ void __launch_bounds__(256) k_region_scatter(const int* win, int nwin, const unsigned char* regs, int* meta,
                                                         int* win2) {
    for (int i = blockIdx.x * 256 + threadIdx.x; i < nwin; i += gridDim.x * 256)
        win2[atomicAdd(&meta[16 + regs[i]], 1)] = win[i];
}

template <int D>
__global__ void k_live_flags(Ctx c, int n, int* flag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[i] = *v_ptr<D>(c, i, 0) >= 0;
}

// exclusive scan of flags within each 256-block (block sums out), then add block offsets
__global__ void k_block_scan(const int* in, int* out, int* bsum, int n) {
    __shared__ int s[256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int v = i < n ? in[i] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        int add = threadIdx.x >= (unsigned)o ? s[threadIdx.x - o] : 0;
        __syncthreads();
        s[threadIdx.x] += add;
        __syncthreads();
    }
    if (i < n) out[i] = s[threadIdx.x] - v;
    if (threadIdx.x == 255) bsum[blockIdx.x] = s[255];
}
__global__ void k_add_offsets(int* out, const int* boff, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] += boff[blockIdx.x];
}


// the lists a device-planned chunk starts from are complete: no planning launch may apply the
// shared-facet list or compact the outcome flags of a round finished before it
__global__ void k_clear_fix(RoundState* st) {
    st->n_fix = 0;
    st->flag_n = 0;
    st->n_compact = 0;
}

__global__ void k_reset_round(RoundState* st) {
    st->n_next_pending = 0;
    st->n_winners = 0;
    st->n_q = 0;
    st->n_fix = 0;
    st->n_live_delta = 0;
    st->n_big = 0;
}

// Absorb the last planned round and plan the next one (one thread).  Absorbing follows the
// host's rule in finish_round: duplicates are the window points neither inserted nor sent
// on; the window factor gamma grows when more than the target winning fraction won and
// shrinks when fewer did; the next window is gamma x inserted points (at least 8, at most the
// window capacity), the pending points first.
__global__ void __launch_bounds__(kCntStripes) k_round_totals(const Counters* cnt, RoundState* st) {
    round_totals(cnt, st);
}

__device__ void plan_step(RunPlan& p, RoundState* st, bool absorb);

// Block 0 plans the next round; blocks 1.. link the facets the last round's winners share (the
// shared-facet list, as k_fix_shared) and compact its outcome flags into pend_dst, the pending
// list of the next window (with Ctx::pflag).  The planner reads neither: one launch for all.
// The list count is cleared by the next cavity kernel, which every planning launch precedes.
template <int D>
__global__ void __launch_bounds__(256) k_plan_one(RunPlan* P, Ctx c, const Counters* cnt, int* pend_dst) {
    if (blockIdx.x > 0) {
        if (!c.reuse && c.fix) fix_shared_pass<D>(c, (blockIdx.x - 1) * blockDim.x + threadIdx.x, (gridDim.x - 1) * blockDim.x);
        // the last round's outcome flags (st->flag_n slots) into the next window's pending list
        if (c.pflag)
            append_flagged(c.pflag, c.st->flag_n, pend_dst, c.pend_cap, &c.st->n_compact, &c.st->err, blockIdx.x - 1,
                           gridDim.x - 1);
        return;
    }
    RoundState* st = c.st;
    __shared__ int s_absorb, s_base, s_chunk;
    RunPlan& p = *P;
    if (threadIdx.x == 0) s_absorb = p.launched && !(st->n_big > 0 || (st->n_q > 0 && !p.exact_dev) || st->err);
    __syncthreads();
    // with outcome flags the count of points sent on comes from the statistics (the compaction
    // runs beside this block)
    if (s_absorb) round_totals(cnt, st, c.pflag != nullptr);
    if (threadIdx.x == 0) {
        plan_step(p, st, s_absorb);
        // halted with the last queue already evaluated in the stream: the empty rounds that
        // follow must not evaluate it again (the host reads the count from nq_done)
        if (p.halt && p.exact_dev && st->n_q > 0) {
            p.nq_done = st->n_q;
            st->n_q = 0;
        }
        // a planned round: its allocation chunks at the top of the records (the space check
        // of plan_step counted them)
        s_chunk = (c.ast && p.launched && p.ctl.nwin > 0) ? p.chunk : 0;
        s_base = st->n_tets;
        if (s_chunk) st->n_tets = s_base + kAllocStripes * s_chunk;
    }
    __syncthreads();
    if (s_chunk && threadIdx.x < kAllocStripes) {
        c.ast[threadIdx.x].cur = s_base + threadIdx.x * s_chunk;
        c.ast[threadIdx.x].end = s_base + (threadIdx.x + 1) * s_chunk;
    }
}

__device__ void plan_step(RunPlan& p, RoundState* st, bool s_absorb) {
    p.ctl.nwin = 0;
    if (p.launched) {
        if (!s_absorb) {
            p.halt |= HALT_HOST;
            return;
        }
        const int nwin = p.nwin, w = st->n_winners, nx = st->n_next_pending;
        const int consumed = nwin - nx - w;
        p.dups += consumed;
        p.inserted += w;
        p.n_live += st->n_live_delta;
        p.windows += nwin;
        const double frac = nwin ? (double)w / nwin : 1.0;
        const bool big = p.big_window > 0 && nwin >= p.big_window;
        const double hi = big ? p.big_frac_hi : p.small_frac_hi, lo = big ? p.big_frac_lo : p.small_frac_lo;
        if (frac > hi) p.gamma = fmin(p.gamma * 1.5, 4.0);
        else if (frac < lo) p.gamma = fmax(p.gamma * 0.7, 0.002);
        p.stalled = (w == 0 && consumed == 0 && st->n_q == 0) ? p.stalled + 1 : 0;
        p.npend = nx;
        p.cur ^= 1;
        p.head += p.nfresh;
        p.rounds++;
        p.launched = 0;
        if (p.stalled > 64) p.halt |= HALT_STALL;
    }
    if (p.halt) return;
    if (p.head >= p.n_pts && p.npend == 0) {
        p.halt |= HALT_DONE;
        return;
    }
    const double target = fmax((double)p.min_window, p.gamma * (double)(p.inserted > 1 ? p.inserted : 1));
    int nwin = (int)fmin(target, (double)p.win_cap);
    if (p.npend > nwin) nwin = p.npend;
    const int nfresh = min(p.n_pts - p.head, nwin - p.npend);
    nwin = p.npend + nfresh;
    if (p.boot_limit > 0 && (nwin > p.boot_limit || p.rounds >= p.boot_rounds_end)) {
        p.halt |= HALT_BOOT;
        return;
    }
    const long long chunk =
        p.chunk_rate_milli > 0 ? ((long long)nwin * p.chunk_rate_milli / 1000 + kAllocStripes - 1) / kAllocStripes + 16 : 0;
    if ((long long)st->n_tets + (long long)nwin * p.per_point + chunk * kAllocStripes + 1024 > p.tet_cap) {
        p.halt |= HALT_SPACE;
        return;
    }
    p.nfresh = nfresh;
    p.nwin = nwin;
    p.chunk = (int)chunk;
    p.ctl.nwin = nwin;
    p.ctl.npend = p.npend;
    p.ctl.virt = 1;            // pending points, then fresh points head .. (not written out)
    p.ctl.vhead = p.head;
    p.ctl.key_hi = (unsigned long long)(0xFFFFFFFFu - (uint32_t)p.rounds) << 32;
    st->n_next_pending = 0;
    st->n_winners = 0;
    st->n_q = 0;
    st->n_live_delta = 0;
    st->n_big = 0;
    p.launched = 1;
}

// ---------------------------------------------------------------- early rounds, one workgroup
// A run's first rounds have windows of tens to a few hundred points.  Such a round is a short
// chain of dependent memory accesses per point, and as launched rounds it costs four launches
// (plan, cavity, commit, shared-facet links), each starting with cold caches.  k_boot runs these
// rounds in one workgroup: its wavefronts take the window's slots as the blocks of k_cavity_g /
// k_commit_g do (the same pass functions), barriers take the place of launches, and the records
// stay in the CU's caches from round to round.  The planner is the launched rounds' planner
// (plan_step), with boot_limit set: it halts before a window above the limit or at
// boot_rounds_end, and the kernel clears that halt on exit, leaving no round planned, so the
// host carries on from the plan as after any halt.  A round the kernel cannot absorb (a cavity
// beyond the LDS lists, queued exact predicates, an error) halts the planner for the host, as in
// k_plan_one without in-stream exact evaluation.  The loop is bounded by max_rounds.
constexpr int kBootMaxRounds = 4096;
template <int D> constexpr int kBootWaves = D == 3 ? 8 : 12;
template <int D> constexpr int kBootLanes = 8;      // lanes per point in both passes

template <int D>
__global__ void __launch_bounds__(64 * kBootWaves<D>) k_boot(Ctx c, RunPlan* P, int* pend0, int* pend1, int max_rounds) {
    constexpr int G = kBootLanes<D>;
    constexpr int NW = kBootWaves<D>;
    __shared__ union {
        CavityLds<D, G> cav[NW];
        CommitLds<D, G> com[NW];
    } U;
    __shared__ int s_absorb, s_go;     // two flags: a wave may still read one while thread 0 sets the other
    const int w = threadIdx.x >> 6;
    RunPlan& p = *P;
    for (int r = 0; r <= max_rounds; ++r) {
        if (threadIdx.x == 0)
            s_absorb = p.launched && !(c.st->n_big > 0 || c.st->n_q > 0 || c.st->err);
        __syncthreads();
        const bool absorb = s_absorb != 0;
        if (absorb) round_totals(c.cnt, c.st);
        if (threadIdx.x == 0) {
            if (r == max_rounds && !p.halt) p.halt |= HALT_BOOT;    // bound: plan nothing more
            plan_step(p, c.st, absorb);
            s_go = !p.halt && p.ctl.nwin > 0;
        }
        __syncthreads();
        if (!s_go) break;
        const int nwin = p.ctl.nwin;
        const unsigned long long key_hi = p.ctl.key_hi;
        c.win = p.cur ? pend1 : pend0;
        c.pend_next = p.cur ? pend0 : pend1;
        cavity_pass<D, G>(c, nwin, key_hi, SlotIter{w * Group<G>::NG, NW * Group<G>::NG, nwin, 0}, U.cav[w]);
        __syncthreads();
        commit_pass<D, G>(c, nwin, key_hi, SlotIter{w * Group<G>::NG, NW * Group<G>::NG, nwin, 0}, U.com[w]);
        __syncthreads();
        if (!c.reuse && c.fix) {
            fix_shared_pass<D>(c, threadIdx.x, blockDim.x);
            __syncthreads();
        }
        if (threadIdx.x == 0) c.st->n_fix = 0;     // applied: the next round's commit starts a new list
    }
    if (threadIdx.x == 0) {
        p.halt &= ~HALT_BOOT;
        p.boot_limit = 0;
    }
}




// grow the simplex records to cap (copying the first n_tets); new records are all 0xFF bytes:
// dead, hull adjacency, reservations at the largest key
int grow_tets(vo_tri* t, int64_t cap) {
    if (cap <= t->tcap) return 0;
    if (cap > (int64_t)1 << 29) return fail(VO_ERR_OOM, "simplex capacity above 2^29 (adjacency packs index << 2)");
    const size_t B = (size_t)record_bytes(t->D);
    char* rec;
    int rc;
    if ((rc = dev_alloc(&rec, (size_t)cap * B))) return rc;
    const size_t keep = t->rec ? (size_t)t->n_tets * B : 0;
    VO_HIP(hipMemsetAsync(rec + keep, 0xFF, (size_t)cap * B - keep, t->s));
    if (t->rec) {
        VO_HIP(hipMemcpyAsync(rec, t->rec, keep, hipMemcpyDeviceToDevice, t->s));
        VO_HIP(hipStreamSynchronize(t->s));
        hipFree(t->rec);
    }
    t->rec = rec;
    t->tcap = cap;
    if (t->rec_spare) {                // reallocated at the next compaction
        hipFree(t->rec_spare);
        t->rec_spare = nullptr;
    }
    if (t->cmap) {
        hipFree(t->cmap);
        t->cmap = nullptr;
    }
    return 0;
}

int grow_points(vo_tri* t, int64_t cap) {
    if (cap <= t->vcap) return 0;
    if (t->M + cap > INT32_MAX / 2) return fail(VO_ERR_INVALID, "too many points");
    double* X64;
    float* X32;
    int *cp, *vmap;
    int rc;
    const int64_t nv = t->M + cap;
    if ((rc = dev_alloc(&X64, nv * t->S)) || (rc = dev_alloc(&X32, nv * t->S)) || (rc = dev_alloc(&cp, cap)) ||
        (rc = dev_alloc(&vmap, nv)))
        return rc;
    if (t->X64) {
        const int64_t old = t->M + t->n_pts;
        VO_HIP(hipMemcpyAsync(X64, t->X64, sizeof(double) * old * t->S, hipMemcpyDeviceToDevice, t->s));
        VO_HIP(hipMemcpyAsync(X32, t->X32, sizeof(float) * old * t->S, hipMemcpyDeviceToDevice, t->s));
        VO_HIP(hipMemcpyAsync(cp, t->cp, sizeof(int) * std::max<int64_t>(t->n_pts, 1), hipMemcpyDeviceToDevice, t->s));
        VO_HIP(hipMemcpyAsync(vmap, t->vmap, sizeof(int) * old, hipMemcpyDeviceToDevice, t->s));
        VO_HIP(hipStreamSynchronize(t->s));
        hipFree(t->X64); hipFree(t->X32); hipFree(t->cp); hipFree(t->vmap);
    }
    t->X64 = X64; t->X32 = X32; t->cp = cp; t->vmap = vmap;
    t->vcap = cap;
    return 0;
}

int upload_super(vo_tri* t) {
    std::vector<double> x64((size_t)t->M * t->S, 0.0);
    std::vector<float> x32((size_t)t->M * t->S, 0.0f);
    std::vector<int> vm(t->M);
    for (int v = 0; v < t->M; ++v) {
        for (int k = 0; k < t->D; ++k) {
            x64[v * t->S + k] = t->sv[v][k];
            x32[v * t->S + k] = (float)(t->sv[v][k] - t->c[k]);
        }
        vm[v] = v;
    }
    VO_HIP(hipMemcpyAsync(t->X64, x64.data(), sizeof(double) * x64.size(), hipMemcpyHostToDevice, t->s));
    VO_HIP(hipMemcpyAsync(t->X32, x32.data(), sizeof(float) * x32.size(), hipMemcpyHostToDevice, t->s));
    VO_HIP(hipMemcpyAsync(t->vmap, vm.data(), sizeof(int) * vm.size(), hipMemcpyHostToDevice, t->s));
    VO_HIP(hipStreamSynchronize(t->s));
    return 0;
}

void free_all(vo_tri* t) {
    void* ptrs[] = {t->X64, t->X32, t->cp, t->vmap, t->rec, t->rec_spare, t->cmap, t->chist, t->reg_meta, t->win2,
                    t->regs, t->win, t->pend[0], t->pend[1],
                    t->cav, t->fac, t->fnb, t->nk, t->nf, t->status, t->hint, t->st, t->cnt, t->q, t->cache, t->ws,
                    t->big, t->cav1, t->fac1, t->misc1, t->plan, t->fix, t->ast, t->pflag};
    for (void* p : ptrs)
        if (p) hipFree(p);
#ifdef VO_POINT_PROBE
    if (t->probe) hipFree(t->probe);
    if (t->probe_f) std::fclose(t->probe_f);
#endif
    if (t->st_host) hipHostFree(t->st_host);
    if (t->plan_host) hipHostFree(t->plan_host);
    for (auto& e : t->ev)
        if (e) hipEventDestroy(e);
    for (auto& e : t->rev)
        if (e) hipEventDestroy(e);
    for (auto& e : t->chunk_ev)
        if (e) hipEventDestroy(e);
    if (t->s) hipStreamDestroy(t->s);
}

int sync_state(vo_tri* t) {
    VO_HIP(hipMemcpyAsync(t->st_host, t->st, sizeof(RoundState), hipMemcpyDeviceToHost, t->s));
    VO_HIP(hipStreamSynchronize(t->s));
    return 0;
}

int check_err(vo_tri* t) {
    const unsigned e = t->st_host->err;
    if (!e) return 0;
    if (e & ERR_OVERFLOW) return fail(VO_ERR_OVERFLOW, "a cavity exceeded the per-point buffers");
    if (e & ERR_TETCAP) return fail(VO_ERR_OVERFLOW, "simplex capacity exceeded");
    if (e & ERR_PRECISION) return fail(VO_ERR_PRECISION, "exact predicate left the f64 range or workspace");
    if (e & ERR_CACHE) return fail(VO_ERR_OVERFLOW, "exact-result cache full");
    if (e & ERR_HULL) return fail(VO_ERR_INVALID, "a point lies outside the super simplex");
    return fail(VO_ERR_INTERNAL, "walk, chain or rotation bound exceeded (err=" + std::to_string(e) + ")");
}

int run_exact(vo_tri* t, int nq) {
    if (nq <= 0) return 0;
    nq = std::min(nq, t->qcap);
    const int threads = 64;
    const int blocks = std::max(1, std::min((nq + threads - 1) / threads, t->ws_threads / threads));
    VO_HIP(hipEventRecord(t->ev[4], t->s));
    if (t->D == 3)
        hipLaunchKernelGGL(k_exact<4>, dim3(blocks), dim3(threads), 0, t->s, t->q, nq, t->X64, t->ws, t->ws_cap, t->cache,
                           t->cmask, t->st, t->cnt, t->qcap);
    else
        hipLaunchKernelGGL(k_exact<2>, dim3(blocks), dim3(threads), 0, t->s, t->q, nq, t->X64, t->ws, t->ws_cap, t->cache,
                           t->cmask, t->st, t->cnt, t->qcap);
    VO_HIP(hipGetLastError());
    VO_HIP(hipEventRecord(t->ev[5], t->s));
    VO_HIP(hipEventSynchronize(t->ev[5]));
    float ms = 0;
    hipEventElapsedTime(&ms, t->ev[4], t->ev[5]);
    t->ms_ex += ms;
    return 0;
}

// Points whose cavity outgrew the group kernel's LDS lists, re-run one at a time with the
// thread kernels over single-point lists of big_cap entries.  A point alone in its round wins
// unless a predicate defers (then it joins the next window).  The solo runs add to the same
// device round state (winners, live delta, pending list, exact queue) as the round itself.
int run_solo(vo_tri* t, int n_big) {
    int rc;
    t->n_tets = t->st_host->n_tets;     // grow_tets below copies the first n_tets simplices
    for (int i = 0; i < n_big; ++i) {
        t->rounds++;
        const unsigned long long key_hi = (unsigned long long)(0xFFFFFFFFu - (uint32_t)t->rounds) << 32;
        Ctx c = make_ctx(t);
        c.win = t->big + i;
        c.cav = t->cav1;
        c.fac = t->fac1;
        c.nk = t->misc1;
        c.nf = t->misc1 + 1;
        c.status = t->misc1 + 2;
        c.win_cap = 1;
        c.kcap = t->big_cap;
        c.fcap = t->big_cap;
        c.solo = 1;
        const int need = t->n_tets + t->big_cap + 1024;
        if (need > t->tcap) {
            if ((rc = grow_tets(t, std::max<int64_t>(need, t->tcap * 3 / 2)))) return rc;
            c = make_ctx(t);
            c.win = t->big + i; c.cav = t->cav1; c.fac = t->fac1;
            c.nk = t->misc1; c.nf = t->misc1 + 1; c.status = t->misc1 + 2;
            c.win_cap = 1; c.kcap = t->big_cap; c.fcap = t->big_cap; c.solo = 1;
        }
        if (t->D == 3) {
            hipLaunchKernelGGL(k_cavity<3>, dim3(1), dim3(256), 0, t->s, c, 1, key_hi);
            hipLaunchKernelGGL(k_commit<3>, dim3(1), dim3(256), 0, t->s, c, 1, key_hi);
        } else {
            hipLaunchKernelGGL(k_cavity<2>, dim3(1), dim3(256), 0, t->s, c, 1, key_hi);
            hipLaunchKernelGGL(k_commit<2>, dim3(1), dim3(256), 0, t->s, c, 1, key_hi);
        }
        VO_HIP(hipGetLastError());
        if ((rc = sync_state(t)) || (rc = check_err(t))) return rc;
        t->n_tets = t->st_host->n_tets;
    }
    t->n_solo += n_big;
    return 0;
}

// Live records to a fresh array in spatial order (compact.hpp), between rounds: the pending
// points' conflict pointers and the walk hints are first moved to live simplices, then
// renumbered with the records.
int compact(vo_tri* t) {
    const int D = t->D;
    const int n = t->n_tets;
    const size_t B = (size_t)record_bytes(D);
    int rc;
    if (!t->rec_spare) {
        // a fresh spare starts as all 0xFF records (largest reservation keys); afterwards its
        // records keep keys of earlier rounds, which are larger than any later round's
        if ((rc = dev_alloc(&t->rec_spare, (size_t)t->tcap * B))) return rc;
        VO_HIP(hipMemsetAsync(t->rec_spare, 0xFF, (size_t)t->tcap * B, t->s));
    }
    if (!t->cmap && (rc = dev_alloc(&t->cmap, (size_t)t->tcap))) return rc;
    if (!t->chist && (rc = dev_alloc(&t->chist, kCompactBins))) return rc;
    VO_HIP(hipEventRecord(t->ev[4], t->s));
    const Ctx c = make_ctx(t);
    const int cells = (int)hint_cells(D, t->hlevels);
    const int blocks = std::max(1, std::min(4096, (n + 255) / 256));
    int* pend = t->pend[t->cur];
    if (D == 3) {
        if (t->npend) hipLaunchKernelGGL(k_chase_points<3>, dim3(256), dim3(256), 0, t->s, c, pend, t->npend);
        hipLaunchKernelGGL(k_chase_hints<3>, dim3(1024), dim3(256), 0, t->s, c, t->hint, cells);
    } else {
        if (t->npend) hipLaunchKernelGGL(k_chase_points<2>, dim3(256), dim3(256), 0, t->s, c, pend, t->npend);
        hipLaunchKernelGGL(k_chase_hints<2>, dim3(1024), dim3(256), 0, t->s, c, t->hint, cells);
    }
    VO_HIP(hipMemsetAsync(t->chist, 0, sizeof(int) * kCompactBins, t->s));
    if (D == 3)
        hipLaunchKernelGGL(k_bin_count<3>, dim3(blocks), dim3(256), 0, t->s, c, n, t->chist, t->glo[0], t->glo[1],
                           t->glo[2], t->ginv);
    else
        hipLaunchKernelGGL(k_bin_count<2>, dim3(blocks), dim3(256), 0, t->s, c, n, t->chist, t->glo[0], t->glo[1],
                           t->glo[2], t->ginv);
    std::vector<int> h(kCompactBins);
    VO_HIP(hipMemcpyAsync(h.data(), t->chist, sizeof(int) * kCompactBins, hipMemcpyDeviceToHost, t->s));
    VO_HIP(hipStreamSynchronize(t->s));
    int live = 0;
    for (int i = 0; i < kCompactBins; ++i) {
        const int k = h[i];
        h[i] = live;
        live += k;
    }
    VO_HIP(hipMemcpyAsync(t->chist, h.data(), sizeof(int) * kCompactBins, hipMemcpyHostToDevice, t->s));
    if (D == 3) {
        hipLaunchKernelGGL(k_bin_place<3>, dim3(blocks), dim3(256), 0, t->s, c, n, t->chist, t->cmap, t->glo[0],
                           t->glo[1], t->glo[2], t->ginv);
        hipLaunchKernelGGL(k_move<3>, dim3(blocks), dim3(256), 0, t->s, c, t->rec_spare, t->cmap, n);
    } else {
        hipLaunchKernelGGL(k_bin_place<2>, dim3(blocks), dim3(256), 0, t->s, c, n, t->chist, t->cmap, t->glo[0],
                           t->glo[1], t->glo[2], t->ginv);
        hipLaunchKernelGGL(k_move<2>, dim3(blocks), dim3(256), 0, t->s, c, t->rec_spare, t->cmap, n);
    }
    if (t->npend) hipLaunchKernelGGL(k_remap_points, dim3(256), dim3(256), 0, t->s, t->cp, pend, t->npend, t->cmap);
    hipLaunchKernelGGL(k_remap_hints, dim3(1024), dim3(256), 0, t->s, t->hint, cells, t->cmap);
    VO_HIP(hipGetLastError());
    VO_HIP(hipMemcpyAsync(&t->st->n_tets, &live, sizeof(int), hipMemcpyHostToDevice, t->s));
    VO_HIP(hipEventRecord(t->ev[5], t->s));
    if ((rc = sync_state(t)) || (rc = check_err(t))) return rc;
    float ms = 0;
    hipEventElapsedTime(&ms, t->ev[4], t->ev[5]);
    t->ms_compact += ms;
    std::swap(t->rec, t->rec_spare);
    // the records past the live ones are made fresh again (all 0xFF: dead): the allocation
    // stripes leave unused slots below the count, which must not look live
    {
        const size_t B = (size_t)record_bytes(t->D);
        VO_HIP(hipMemsetAsync(t->rec + (size_t)live * B, 0xFF, (size_t)(t->tcap - live) * B, t->s));
    }
    t->retired += n - live;
    t->n_tets = live;
    t->n_live = live;
    t->n_compact++;
    return 0;
}

#define VO_LAUNCH_G(DIM, GL)                                                                     \
    do {                                                                                         \
        const dim3 grid_(blocks(64 / GL));                                                       \
        if (which) hipLaunchKernelGGL((k_commit_g<DIM, GL>), grid_, dim3(64), 0, t->s, c, nwin, key_hi); \
        else hipLaunchKernelGGL((k_cavity_g<DIM, GL>), grid_, dim3(64), 0, t->s, c, nwin, key_hi);       \
    } while (0)

// which = 0: locate + cavity + reservation, 1: check + commit
// lanes per point for this round: small windows are latency-bound (more lanes per point
// shorten each point's dependent-load chain); large windows are throughput-bound, where four
// lanes per point (face tests in parallel, LDS lists, one run of loads per simplex) beat one
// thread per point (measured: 2D n = 10M 114 -> 108 ms), so the choice stops at 4 unless
// "group_lanes" = 1 asks for the thread kernels
// The 3D commit of small windows takes up to a whole wavefront per point (facets and ridge
// rotations over 64 lanes): measured faster than 32 lanes for the commit, slower for the cavity.
int auto_group(const vo_tri* t, int nwin, int which) {
    if (t->group) return t->group;
    const int64_t base = t->lanes_target ? t->lanes_target : (t->D == 3 ? (1 << 21) : (1 << 16));
    const int64_t target = which && t->commit_lanes_target ? t->commit_lanes_target : base;
    static const int sizes[] = {64, 32, 16, 8};
    const int gmax = t->D == 3 ? (which ? 64 : 32) : 16;
    for (int g : sizes)
        if (g <= gmax && (int64_t)nwin * g <= target) return g;
    return 4;
}

// nwin sizes the grid (the kernels read the exact window from c.ctl when it is set); G = 0:
// lanes per point chosen from nwin
int launch_round(vo_tri* t, const Ctx& c, int nwin, unsigned long long key_hi, int which, int G = 0, bool fix = true) {
    if (!G) G = auto_group(t, nwin, which);
    // with regions, a multiple of 8 blocks and slack for unequal region sizes
    auto blocks = [&](int per_block) {
        const int b = (nwin + per_block - 1) / per_block;
        return c.region ? ((b + b / 8 + 8 + 7) & ~7) : b;
    };
    if (G == 1) {
        const int grid = blocks(256);
        if (t->D == 3) {
            if (which) hipLaunchKernelGGL(k_commit<3>, dim3(grid), dim3(256), 0, t->s, c, nwin, key_hi);
            else hipLaunchKernelGGL(k_cavity<3>, dim3(grid), dim3(256), 0, t->s, c, nwin, key_hi);
        } else {
            if (which) hipLaunchKernelGGL(k_commit<2>, dim3(grid), dim3(256), 0, t->s, c, nwin, key_hi);
            else hipLaunchKernelGGL(k_cavity<2>, dim3(grid), dim3(256), 0, t->s, c, nwin, key_hi);
        }
    } else if (t->D == 3 && which && G == 64) {
        hipLaunchKernelGGL((k_commit_g<3, 64>), dim3(blocks(1)), dim3(64), 0, t->s, c, nwin, key_hi);
    } else if (t->D == 3) {
        switch (G) {
            case 4: VO_LAUNCH_G(3, 4); break;
            case 8: VO_LAUNCH_G(3, 8); break;
            case 16: VO_LAUNCH_G(3, 16); break;
            default: VO_LAUNCH_G(3, 32); break;
        }
    } else {
        switch (G) {
            case 4: VO_LAUNCH_G(2, 4); break;
            case 8: VO_LAUNCH_G(2, 8); break;
            case 16: VO_LAUNCH_G(2, 16); break;
            default: VO_LAUNCH_G(2, 32); break;
        }
    }
    VO_HIP(hipGetLastError());
    // links across the facets two winners share (only without reuse; the thread commit keeps to
    // disjoint, non-adjacent cavities)
    if (fix && which && G != 1 && !c.reuse && c.fix) {
        const int fb = std::max(1, std::min(1024, nwin / 64 + 1));
        if (t->D == 3) hipLaunchKernelGGL(k_fix_shared<3>, dim3(fb), dim3(256), 0, t->s, c);
        else hipLaunchKernelGGL(k_fix_shared<2>, dim3(fb), dim3(256), 0, t->s, c);
        VO_HIP(hipGetLastError());
    }
    return 0;
}
#undef VO_LAUNCH_G

}  // namespace

// ------------------------------------------------------------------ predicate tiers, one by one
// Sign of each predicate instance through the tier cascade of the construction: FP32 on
// coordinates stored as fl32(x - center), then the FP64 filter, then FP64 expansions; the tier
// that decided is reported.  Grid-stride over the instances, one expansion workspace slice per
// thread.
template <int D>
__global__ void __launch_bounds__(64) k_eval_predicates(int type, const double* pts, int n, double c0, double c1,
                                                        double c2, double* ws, int ws_cap, int32_t* sign,
                                                        int32_t* tier) {
    using F = typename VT<D>::F;
    const int A = pred_arity((uint32_t)type);
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const double cc[3] = {c0, c1, c2};
    for (int i = tid; i < n; i += gridDim.x * blockDim.x) {
        const double* P[5];
        F f[5];
        for (int k = 0; k < A; ++k) {
            P[k] = pts + ((size_t)i * A + k) * D;
            float q[3] = {0.f, 0.f, 0.f};
            for (int j = 0; j < D; ++j) q[j] = (float)(P[k][j] - cc[j]);
            if constexpr (D == 3) f[k] = make_float4(q[0], q[1], q[2], 0.f);
            else f[k] = make_float2(q[0], q[1]);
        }
        int s, t = 1;
        if constexpr (D == 3) {
            s = type == PRED_ORIENT3 ? t1_orient3d(f[0], f[1], f[2], f[3]) : t1_insphere(f[0], f[1], f[2], f[3], f[4]);
            if (s == SIGN_UNDECIDED) {
                t = 2;
                s = type == PRED_ORIENT3 ? t2_orient3d(P[0], P[1], P[2], P[3]) : t2_insphere(P[0], P[1], P[2], P[3], P[4]);
            }
        } else {
            s = type == PRED_ORIENT2 ? t1_orient2d(f[0], f[1], f[2]) : t1_incircle(f[0], f[1], f[2], f[3]);
            if (s == SIGN_UNDECIDED) {
                t = 2;
                s = type == PRED_ORIENT2 ? t2_orient2d(P[0], P[1], P[2]) : t2_incircle(P[0], P[1], P[2], P[3]);
            }
        }
        if (s == SIGN_UNDECIDED) {
            t = 3;
            Ws w{ws + (size_t)tid * ws_cap, ws_cap, 0, 0};
            switch (type) {
                case PRED_ORIENT2: s = ex_orient2d(w, P[0], P[1], P[2]); break;
                case PRED_INCIRCLE: s = ex_incircle(w, P[0], P[1], P[2], P[3]); break;
                case PRED_ORIENT3: s = ex_orient3d(w, P[0], P[1], P[2], P[3]); break;
                default: s = ex_insphere(w, P[0], P[1], P[2], P[3], P[4]); break;
            }
            if (w.ovf) s = SIGN_UNDECIDED;
        }
        sign[i] = s;
        tier[i] = t;
    }
}

namespace {

// simplex records for n_tets + need, growing by at least half (grow_tets copies the live part)
int ensure_records(vo_tri* t, int64_t need) {
    if (need <= t->tcap) return 0;
    const int64_t cap_max = (int64_t)1 << 29;
    return grow_tets(t, std::max(need, std::min(t->tcap * 3 / 2, cap_max)));
}

// records a window point may create at most, as the planners reserve them
int64_t window_records(int D) { return D == 3 ? 64 : 16; }

// Record reuse for this run: on request, or (auto) when append-only records for the points still
// to insert would not fit the 2^29 the adjacency packing allows.  Reuse costs time (a winner's new
// simplices scatter over its killed records' lines instead of filling fresh consecutive ones:
// measured 3D n = 4M 147 -> 154 ms, 2D n = 10M 99 -> 114 ms) and needs the stricter round rule.
void choose_reuse(vo_tri* t, int64_t to_insert) {
    // append-only growth per point: about 27 (3D) / 6 (2D) live simplices, with allocation
    // stripes the unused chunk tails (measured up to 39 / 8.5 records per point in all), and
    // the reservation of the run's last rounds (64 / 16 records per window point, plus their
    // chunks: about 5 per input point), which must fit below the cap as well
    const int64_t per = t->alloc_stripes ? (t->D == 3 ? 46 : 14) : (t->D == 3 ? 34 : 10);
    const int64_t append_only = (int64_t)t->n_tets + per * to_insert + 4096;
    t->reuse_on = t->reuse == 1 || (t->reuse < 0 && append_only > ((int64_t)1 << 29));
}

// records reserved per point to insert: a winner creates about 27 (3D) / 6 (2D) simplices; with
// reuse all but about 7 (3D) / 2 (2D) of them take the winner's killed records
int64_t records_per_point(const vo_tri* t) {
    if (t->reuse_on) return t->D == 3 ? 12 : 4;
    // allocation stripes leave the unused part of each round's chunks behind (measured: 35-39
    // records per point in 3D, uniform and clustered; 8.5 in 2D)
    if (t->alloc_stripes) return t->D == 3 ? 44 : 12;
    return t->D == 3 ? 30 : 8;
}

// records a planned round may claim: per_point for each window point, plus the stripes' chunks
// (the planner's space check, plan_step)
int chunk_rate_milli(const vo_tri* t) { return (t->alloc_stripes && !t->reuse_on) ? (t->D == 3 ? 25000 : 5400) : 0; }
int64_t round_records(const vo_tri* t, int64_t nwin) {
    const int64_t rate = chunk_rate_milli(t);
    const int64_t chunk = rate ? (nwin * rate / 1000 + kAllocStripes - 1) / kAllocStripes + 16 : 0;
    return nwin * window_records(t->D) + chunk * kAllocStripes + 1024;
}

// The host's side of a finished round whose state is in st_host: errors, points re-run alone,
// the exact queue, then the counters and the window factor (k_plan_one applies the same rule
// on the device).
int finish_round(vo_tri* t, int nwin, int nfresh, FILE* trace, float ms_cav, float ms_com, bool exact_done = false) {
    int rc;
    if ((rc = check_err(t))) return rc;
    const int n_big = t->st_host->n_big;
    if (n_big > 0 && (rc = run_solo(t, n_big))) return rc;
    const RoundState& r = *t->st_host;
    if (r.n_q > 0 && !exact_done) {      // (exact_done: evaluated in the stream after the commit)
        if ((rc = run_exact(t, r.n_q))) return rc;
        if ((rc = sync_state(t)) || (rc = check_err(t))) return rc;
    }
    // the round's winners (solo runs included) from the statistics
    hipLaunchKernelGGL(k_round_totals, dim3(1), dim3(kCntStripes), 0, t->s, t->cnt, t->st);
    if ((rc = sync_state(t))) return rc;
    t->n_tets = r.n_tets;
    t->n_live += r.n_live_delta;
    t->inserted += r.n_winners;
    t->windows_total += nwin;
    if (trace) {
        std::vector<Counters> hs(kCntStripes);
        hipMemcpy(hs.data(), t->cnt, sizeof(Counters) * kCntStripes, hipMemcpyDeviceToHost);
        const Counters h = sum_counters(hs.data());
        std::fprintf(trace, "%lld %d %d %d %d %.4f %.4f %d %lld %d %d %llu %llu %llu %llu\n", (long long)t->rounds, nwin,
                     t->npend, nfresh, r.n_winners, ms_cav, ms_com, r.n_q, (long long)t->inserted, n_big,
                     r.n_next_pending, h.walk_steps, h.chain_hops, h.killed, h.created);
    }
    const int64_t consumed = (int64_t)nwin - r.n_next_pending - r.n_winners;   // duplicates
    t->dups += consumed;
    const double frac = nwin ? (double)r.n_winners / nwin : 1.0;
    // keep about half of the window winning; windows past big_window are throughput-bound
    // (a loser's cavity is wasted work, a round costs little beyond its points), so there
    // the target winning fraction is higher
    const bool big = t->big_window > 0 && nwin >= t->big_window;
    const double hi = big ? t->big_frac_hi : t->small_frac_hi, lo = big ? t->big_frac_lo : t->small_frac_lo;
    if (frac > hi) t->gamma = std::min(t->gamma * 1.5, 4.0);
    else if (frac < lo) t->gamma = std::max(t->gamma * 0.7, 0.002);
    // every round either inserts, drops a duplicate, or resolves queued predicates
    t->stalled = (r.n_winners == 0 && consumed == 0 && r.n_q == 0) ? t->stalled + 1 : 0;
    if (t->stalled > 64) return fail(VO_ERR_INTERNAL, "no progress for 64 rounds");
    t->npend = r.n_next_pending;
    t->cur ^= 1;
    t->head += nfresh;
    t->rounds++;
    return 0;
}

// window of the next round by the host's rule: (nwin, nfresh)
std::pair<int, int> host_window(const vo_tri* t) {
    const double target = std::max((double)t->min_window, t->gamma * (double)std::max<int64_t>(t->inserted, 1));
    int nwin = (int)std::min<double>(target, t->win_cap);
    nwin = std::max(nwin, t->npend);
    const int nfresh = (int)std::min<int64_t>(t->n_pts - t->head, nwin - t->npend);
    return {t->npend + nfresh, nfresh};
}

// Rounds planned on the host, one at a time: plan, launch, wait for the round's state, finish.
int run_host_planned(vo_tri* t, FILE* trace) {
    const int D = t->D;
    int rc = 0;
    while (t->head < t->n_pts || t->npend > 0) {
        if (t->npend > t->win_cap) return fail(VO_ERR_INTERNAL, "pending list beyond the window capacity");
        const auto [nwin, nfresh] = host_window(t);
        if ((rc = ensure_records(t, (int64_t)t->n_tets + (int64_t)nwin * window_records(D) + 1024))) return rc;
        const int grid = (nwin + 255) / 256;
        hipLaunchKernelGGL(k_window, dim3(grid), dim3(256), 0, t->s, t->win, t->pend[t->cur], t->npend, (int)t->head,
                           nfresh);
        hipLaunchKernelGGL(k_reset_round, dim3(1), dim3(1), 0, t->s, t->st);
        const unsigned long long key_hi = (unsigned long long)(0xFFFFFFFFu - (uint32_t)t->rounds) << 32;
        Ctx c = make_ctx(t);
#ifdef VO_POINT_PROBE
        // VO_PROBE=<path>: per round a header (round, nwin, 0, 0), three uint4 per window slot from
        // the cavity kernel and two from the commit kernel
        if (const char* pp = std::getenv("VO_PROBE")) {
            if (!t->probe && (rc = dev_alloc(&t->probe, 5 * (size_t)t->win_cap))) return rc;
            if (!t->probe_f) t->probe_f = std::fopen(pp, "ab");
            VO_HIP(hipMemsetAsync(t->probe, 0, sizeof(uint4) * 5 * (size_t)t->win_cap, t->s));
            c.probe = t->probe;
        }
#endif
        if (t->xcd) {
            if (!t->reg_meta && ((rc = dev_alloc(&t->reg_meta, 24)) || (rc = dev_alloc(&t->win2, t->win_cap)) ||
                                 (rc = dev_alloc(&t->regs, t->win_cap))))
                return rc;
            VO_HIP(hipMemsetAsync(t->reg_meta, 0, sizeof(int) * 24, t->s));
            const int rb = std::max(1, std::min(1024, (nwin + 255) / 256));
            if (D == 3) hipLaunchKernelGGL(k_region_count<3>, dim3(rb), dim3(256), 0, t->s, c, t->win, nwin, t->regs, t->reg_meta);
            else hipLaunchKernelGGL(k_region_count<2>, dim3(rb), dim3(256), 0, t->s, c, t->win, nwin, t->regs, t->reg_meta);
            hipLaunchKernelGGL(k_region_scan, dim3(1), dim3(1), 0, t->s, t->reg_meta);
            hipLaunchKernelGGL(k_region_scatter, dim3(rb), dim3(256), 0, t->s, t->win, nwin, t->regs, t->reg_meta, t->win2);
            c.win = t->win2;
            c.region = t->reg_meta;
        }
        VO_HIP(hipEventRecord(t->ev[2], t->s));
        if ((rc = launch_round(t, c, nwin, key_hi, 0))) return rc;
        VO_HIP(hipEventRecord(t->ev[3], t->s));
        if ((rc = launch_round(t, c, nwin, key_hi, 1))) return rc;
        VO_HIP(hipEventRecord(t->ev[1], t->s));
        if ((rc = sync_state(t))) return rc;
        float a = 0, b = 0;
        hipEventElapsedTime(&a, t->ev[2], t->ev[3]);
        hipEventElapsedTime(&b, t->ev[3], t->ev[1]);
        t->ms_cav += a;
        t->ms_com += b;
        t->n_cav++;
        t->n_com++;
#ifdef VO_POINT_PROBE
        if (c.probe && t->probe_f) {
            std::vector<uint4> h(5 * (size_t)nwin + 1);
            h[0] = make_uint4((unsigned)t->rounds, (unsigned)nwin, 0u, 0u);
            VO_HIP(hipMemcpy(h.data() + 1, t->probe, sizeof(uint4) * 3 * (size_t)nwin, hipMemcpyDeviceToHost));
            VO_HIP(hipMemcpy(h.data() + 1 + 3 * (size_t)nwin, t->probe + 3 * (size_t)t->win_cap,
                             sizeof(uint4) * 2 * (size_t)nwin, hipMemcpyDeviceToHost));
            std::fwrite(h.data(), sizeof(uint4), h.size(), t->probe_f);
        }
#endif
        if ((rc = finish_round(t, nwin, nfresh, trace, a, b))) return rc;
        if (t->compact_ratio > 0 && t->n_tets >= t->compact_min && (double)t->n_tets > t->compact_ratio * (double)t->n_live &&
            (rc = compact(t)))
            return rc;
    }
    return 0;
}

// the device planner's state from the host's fields (nothing launched, no halt)
RunPlan plan_from_host(const vo_tri* t) {
    RunPlan p = {};
    p.inserted = t->inserted;
    p.dups = t->dups;
    p.windows = t->windows_total;
    p.n_live = t->n_live;
    p.tet_cap = t->tcap;
    p.gamma = t->gamma;
    p.big_frac_lo = t->big_frac_lo;
    p.big_frac_hi = t->big_frac_hi;
    p.small_frac_lo = t->small_frac_lo;
    p.small_frac_hi = t->small_frac_hi;
    p.head = (int)t->head;
    p.n_pts = (int)t->n_pts;
    p.npend = t->npend;
    p.rounds = (int)t->rounds;
    p.stalled = (int)t->stalled;
    p.cur = t->cur;
    p.win_cap = t->win_cap;
    p.big_window = (int)t->big_window;
    p.per_point = (int)window_records(t->D);
    p.exact_dev = t->exact_dev;
    p.min_window = t->min_window;
    p.chunk_rate_milli = chunk_rate_milli(t);
    return p;
}

void plan_to_host(vo_tri* t, const RunPlan& p) {
    t->inserted = p.inserted;
    t->dups = p.dups;
    t->windows_total = p.windows;
    t->n_live = p.n_live;
    t->gamma = p.gamma;
    t->head = p.head;
    t->npend = p.npend;
    t->rounds = p.rounds;
    t->stalled = p.stalled;
    t->cur = p.cur;
}

// cavity and commit kernel time of the R rounds of chunk slot `slot`
void add_chunk_times(vo_tri* t, int slot, int R) {
    if (!t->round_timing) return;
    for (int r = 0; r < R; ++r) {
        hipEvent_t* e = &t->rev[(size_t)(slot * R + r) * 3];
        float a = 0, b = 0;
        hipEventElapsedTime(&a, e[0], e[1]);
        hipEventElapsedTime(&b, e[1], e[2]);
        t->ms_cav += a;
        t->ms_com += b;
    }
}

// Rounds planned on the device (k_plan_one), submitted in chunks of R with two chunks in
// flight: the host reads a chunk's planner snapshot while the next chunk runs, and takes over
// only when the planner halts (points re-run alone, queued exact predicates, an error, no
// simplex space, or done).  Each round's grid is sized from the snapshot two chunks back,
// scaled by the growth the window rule allows over that many rounds; the kernels stop at the
// planned window.
// The host's view after the device planner halted with snapshot P (the stream drained): its
// counters and lists, and the planned round it could not absorb, if any, finished on the host
int absorb_halted_plan(vo_tri* t, const RunPlan& P, bool flags = false) {
    int rc;
    if ((rc = sync_state(t))) return rc;
    // a round the planner did not absorb, run with outcome flags: its pending list was compacted
    // by the halting launch (count in n_compact); the host's finish appends after it
    if (flags && P.launched) {
        t->st_host->n_next_pending = t->st_host->n_compact;
        VO_HIP(hipMemcpyAsync(&t->st->n_next_pending, &t->st_host->n_next_pending, sizeof(int), hipMemcpyHostToDevice,
                              t->s));
    }
    // launches counted for the rounds that ran (the empty rounds after a halt are not)
    t->n_cav += P.rounds - t->rounds + P.launched;
    t->n_com += P.rounds - t->rounds + P.launched;
    plan_to_host(t, P);
    t->n_tets = t->st_host->n_tets;
    // a round that queued exact predicates: the rest of this run evaluates them in the stream
    const bool exact_done = P.exact_dev != 0;
    if (exact_done && P.nq_done > 0) t->st_host->n_q = P.nq_done;
    if (P.launched && t->st_host->n_q > 0) t->exact_dev = 1;
    if (P.launched && (rc = finish_round(t, P.nwin, P.nfresh, nullptr, 0.f, 0.f, exact_done))) return rc;
    if ((rc = check_err(t))) return rc;
    if (P.halt & HALT_STALL) return fail(VO_ERR_INTERNAL, "no progress for 64 rounds");
    return 0;
}

int run_device_planned(vo_tri* t) {
    const int D = t->D, R = t->chunk_rounds;
    int rc;
    bool booted = false;
    if (!t->plan) {
        if ((rc = dev_alloc(&t->plan, 1))) return rc;
        if (hipHostMalloc((void**)&t->plan_host, 2 * sizeof(RunPlan)) != hipSuccess) return fail(VO_ERR_OOM, "pinned");
        for (auto& e : t->chunk_ev) VO_HIP(hipEventCreate(&e));
    }
    while ((int)t->rev.size() < 2 * R * 3) {
        hipEvent_t e;
        VO_HIP(hipEventCreate(&e));
        t->rev.push_back(e);
    }
    // one reservation for the staged points (about 27 records per point in 3D, 6 in 2D);
    // a planner halt grows it when a distribution needs more
    {
        const int64_t left = t->n_pts - t->head + t->npend;
        const int64_t want =
            std::min<int64_t>((int64_t)t->n_tets + left * records_per_point(t) + 4096, (int64_t)1 << 29);
        if ((rc = ensure_records(t, want))) return rc;
    }
    // the lane-group commit of device-planned rounds writes outcome flags (the planning launch's
    // extra blocks, which link shared facets, compact them)
    const bool flags = t->outcome_flags && !t->reuse_on && t->shared_facets && t->group != 1;
    while (t->head < t->n_pts || t->npend > 0) {
        if (t->npend > t->win_cap) return fail(VO_ERR_INTERNAL, "pending list beyond the window capacity");
        // the shared-facet list and outcome flags of a round finished before this point are
        // applied already; the first planning launch must not apply them again (the cavity
        // kernels clear the counts after that)
        hipLaunchKernelGGL(k_clear_fix, dim3(1), dim3(1), 0, t->s, t->st);
        VO_HIP(hipGetLastError());
        {
            const int nwin = host_window(t).first;
            if ((rc = ensure_records(t, (int64_t)t->n_tets + round_records(t, nwin)))) return rc;
        }
        t->plan_host[0] = plan_from_host(t);
        // small first windows: rounds in one workgroup (k_boot) until the window outgrows
        // boot_window, once per run; the host then plans again from the kernel's final plan
        const bool boot = !booted && t->boot_window > 0 && t->group != 1 && !t->exact_dev &&
                          host_window(t).first <= t->boot_window;
        if (boot) {
            t->plan_host[0].boot_limit = t->boot_window;
            t->plan_host[0].boot_rounds_end = (int)std::min<int64_t>(t->rounds + kBootMaxRounds, INT32_MAX);
        }
        VO_HIP(hipMemcpyAsync(t->plan, &t->plan_host[0], sizeof(RunPlan), hipMemcpyHostToDevice, t->s));
        if (boot) {
            booted = true;
            Ctx c = make_ctx(t);
            c.ctl = &t->plan->ctl;
            if (D == 3)
                hipLaunchKernelGGL(k_boot<3>, dim3(1), dim3(64 * kBootWaves<3>), 0, t->s, c, t->plan, t->pend[0],
                                   t->pend[1], kBootMaxRounds);
            else
                hipLaunchKernelGGL(k_boot<2>, dim3(1), dim3(64 * kBootWaves<2>), 0, t->s, c, t->plan, t->pend[0],
                                   t->pend[1], kBootMaxRounds);
            VO_HIP(hipGetLastError());
            VO_HIP(hipMemcpyAsync(&t->plan_host[0], t->plan, sizeof(RunPlan), hipMemcpyDeviceToHost, t->s));
            VO_HIP(hipStreamSynchronize(t->s));
            // the kernel's rounds are no launches of the round kernels, and their time is not
            // in the round kernels' timings (ms_run holds it)
            const int64_t n_cav = t->n_cav, n_com = t->n_com;
            if ((rc = absorb_halted_plan(t, t->plan_host[0]))) return rc;
            t->n_cav = n_cav;
            t->n_com = n_com;
            continue;
        }
        int cur = t->cur;
        int k = 0, inflight = 0;
        double base = host_window(t).first, gam = t->gamma;
        bool halted = false;
        while (!halted) {
            // window estimates for the rounds of this chunk: typical (lanes per point) and high (grid)
            const double grow = 1.0 + 0.5 * gam;
            const double lag = inflight * R;
            const int est_mid = (int)std::min<double>(t->win_cap, base * std::pow(grow, lag + 0.5 * R));
            const int est_hi = (int)std::min<double>(t->win_cap, base * std::pow(grow, lag + R) + 64);
            const int Gc = auto_group(t, est_mid, 0), Gm = auto_group(t, est_mid, 1);
            const int slot = k & 1;
            for (int r = 0; r < R; ++r) {
                // the window is virtual: the kernels read the pending list, then count fresh points
                Ctx c = make_ctx(t);
                c.ctl = &t->plan->ctl;
                c.win = t->pend[cur];
                c.pend_next = t->pend[cur ^ 1];
                if (chunk_rate_milli(t)) c.ast = t->ast;
                if (flags) c.pflag = t->pflag;
                // the plan, with the previous round's shared-facet links (the thread kernels of
                // group_lanes 1 leave none) and pending list
                const int fixb = (!c.reuse && c.fix && t->group != 1) ? std::max(1, std::min(1024, est_hi / 64 + 1)) : 0;
                if (D == 3)
                    hipLaunchKernelGGL(k_plan_one<3>, dim3(1 + fixb), dim3(256), 0, t->s, t->plan, c, t->cnt, t->pend[cur]);
                else
                    hipLaunchKernelGGL(k_plan_one<2>, dim3(1 + fixb), dim3(256), 0, t->s, t->plan, c, t->cnt, t->pend[cur]);
                hipEvent_t* e = &t->rev[(size_t)(slot * R + r) * 3];
                if (t->round_timing) VO_HIP(hipEventRecord(e[0], t->s));
                if ((rc = launch_round(t, c, est_hi, 0, 0, Gc))) return rc;
                if (t->round_timing) VO_HIP(hipEventRecord(e[1], t->s));
                if ((rc = launch_round(t, c, est_hi, 0, 1, Gm, false))) return rc;
                if (t->round_timing) VO_HIP(hipEventRecord(e[2], t->s));
                if (t->exact_dev) {
                    // the round's queued predicates (count read from the round state; none: the
                    // kernel returns at once), before the planner absorbs the round
                    if (D == 3)
                        hipLaunchKernelGGL(k_exact<4>, dim3(t->ws_threads / 64), dim3(64), 0, t->s, t->q, -1, t->X64, t->ws,
                                           t->ws_cap, t->cache, t->cmask, t->st, t->cnt, t->qcap);
                    else
                        hipLaunchKernelGGL(k_exact<2>, dim3(t->ws_threads / 64), dim3(64), 0, t->s, t->q, -1, t->X64, t->ws,
                                           t->ws_cap, t->cache, t->cmask, t->st, t->cnt, t->qcap);
                    VO_HIP(hipGetLastError());
                }
                cur ^= 1;
            }
            VO_HIP(hipMemcpyAsync(&t->plan_host[slot], t->plan, sizeof(RunPlan), hipMemcpyDeviceToHost, t->s));
            VO_HIP(hipEventRecord(t->chunk_ev[slot], t->s));
            ++k;
            if (++inflight < 2) continue;
            const int hs = (k - 2) & 1;
            VO_HIP(hipEventSynchronize(t->chunk_ev[hs]));
            add_chunk_times(t, hs, R);
            --inflight;
            const RunPlan& sp = t->plan_host[hs];
            if (sp.halt) {
                halted = true;
            } else {
                base = std::max<double>(sp.npend, std::min<double>(t->win_cap, std::max((double)t->min_window, sp.gamma * (double)std::max<long long>(sp.inserted, 1))));
                gam = sp.gamma;
            }
        }
        // the newer chunk ends with no-op rounds; its snapshot is the planner's final state
        VO_HIP(hipStreamSynchronize(t->s));
        add_chunk_times(t, (k - 1) & 1, R);
        if ((rc = absorb_halted_plan(t, t->plan_host[(k - 1) & 1], flags))) return rc;
    }
    return 0;
}

}  // namespace

namespace {

// RAII device buffer for the canonical export's temporaries
template <class T> struct DevBuf {
    T* p = nullptr;
    bool own = true;           // false: a view of one of the handle's idle buffers
    ~DevBuf() { if (p && own) hipFree(p); }
};

// n elements in an idle buffer of the handle when it holds them (the window's lists are free
// between runs, and every call of the handle waits for its stream before it returns), else a
// new allocation: large exports then allocate nothing
template <class T> int idle_or_alloc(DevBuf<T>& b, void* idle, size_t idle_bytes, size_t n) {
    if (idle && sizeof(T) * std::max<size_t>(n, 1) <= idle_bytes) {
        b.p = static_cast<T*>(idle);
        b.own = false;
        return 0;
    }
    return dev_alloc(&b.p, n);
}

// exclusive scan of n ints (device), total of the inputs to *total
int device_excl_scan(hipStream_t s, const int* in, int* out, int64_t n, int64_t* total) {
    if (n <= 0) {
        *total = 0;
        return 0;
    }
    const int64_t nb = (n + 255) / 256;
    DevBuf<int> bsum, boff;
    int rc;
    if ((rc = dev_alloc(&bsum.p, (size_t)nb)) || (rc = dev_alloc(&boff.p, (size_t)nb))) return rc;
    hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nb), dim3(256), 0, s, in, out, bsum.p, n);
    VO_HIP(hipGetLastError());
    int64_t t2 = 0;
    if (nb > 1) {
        if ((rc = device_excl_scan(s, bsum.p, boff.p, nb, &t2))) return rc;
        hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nb), dim3(256), 0, s, out, boff.p, n);
        VO_HIP(hipGetLastError());
    }
    int last_in = 0, last_out = 0;
    VO_HIP(hipMemcpyAsync(&last_out, out + n - 1, sizeof(int), hipMemcpyDeviceToHost, s));
    VO_HIP(hipMemcpyAsync(&last_in, in + n - 1, sizeof(int), hipMemcpyDeviceToHost, s));
    VO_HIP(hipStreamSynchronize(s));
    *total = (int64_t)last_out + last_in;
    return 0;
}

// stable LSD radix sort of n 64-bit keys (only the low `bits` bits are compared) with an
// optional 32-bit payload; keys/vals and the temporaries swap roles each pass, so the result is
// in keys/vals on return
int radix_sort(hipStream_t s, unsigned long long*& keys, uint32_t*& vals, unsigned long long*& ktmp, uint32_t*& vtmp,
               int64_t n, int bits) {
    if (n <= 1) return 0;
    const int64_t ntiles = (n + kRadixTile - 1) / kRadixTile;
    DevBuf<int> hist, off;
    int rc;
    if ((rc = dev_alloc(&hist.p, (size_t)(256 * ntiles))) || (rc = dev_alloc(&off.p, (size_t)(256 * ntiles)))) return rc;
    for (int shift = 0; shift < bits; shift += 8) {
        hipLaunchKernelGGL(k_radix_hist, dim3((unsigned)ntiles), dim3(kRadixThreads), 0, s, keys, n, shift, hist.p, (int)ntiles);
        VO_HIP(hipGetLastError());
        int64_t tot = 0;
        if ((rc = device_excl_scan(s, hist.p, off.p, 256 * ntiles, &tot))) return rc;
        if (tot != n) return fail(VO_ERR_INTERNAL, "radix sort: histogram total differs from the key count");
        hipLaunchKernelGGL(k_radix_scatter, dim3((unsigned)ntiles), dim3(kRadixThreads), 0, s, keys, vals, ktmp, vtmp, n,
                           shift, off.p, (int)ntiles);
        VO_HIP(hipGetLastError());
        std::swap(keys, ktmp);
        std::swap(vals, vtmp);
    }
    return 0;
}

int bit_length(int64_t x) {
    int b = 0;
    while (x > 0) {
        ++b;
        x >>= 1;
    }
    return b;
}

}  // namespace

// ------------------------------------------------------------------ allocation-free predicates
// vo_predicates_noalloc stages its instances, signs, tiers and expansion workspaces in one
// file-scope device array instead of hipMalloc'd buffers.  Each device the library runs on holds
// its own copy (about 16.5 MiB) until the process ends.  Layout (offsets 256 B aligned):
//   [0, kNaPts)                 instance coordinates, up to kNaMaxN x 5 x 3 f64
//   [kNaPts, + kNaWs)           expansion workspaces, kNaThreads x kNaWsCap f64 (as vo_predicates)
//   [.., + kNaMaxN x 4 B)       signs, then tiers
constexpr int64_t kNaMaxN = 1 << 16;
constexpr int kNaBlocks = 4, kNaThreadsPerBlock = 64, kNaThreads = kNaBlocks * kNaThreadsPerBlock;
constexpr int kNaWsCap = 1 << 12;
constexpr size_t kNaPts = (size_t)kNaMaxN * 5 * 3 * sizeof(double);
constexpr size_t kNaWs = (size_t)kNaThreads * kNaWsCap * sizeof(double);
constexpr size_t kNaOut = (size_t)kNaMaxN * sizeof(int32_t);
constexpr size_t kNaBytes = kNaPts + kNaWs + 2 * kNaOut;
static_assert(kNaPts % 256 == 0 && kNaWs % 256 == 0 && kNaOut % 256 == 0, "256 B aligned parts");

static __device__ __align__(256) unsigned char scratch[kNaBytes];

namespace {
// the device address of `scratch` on each device, looked up once with that device current (in
// host code the array's name is not a device address)
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch[kMaxDevices] = {};

int scratch_on(int dev, unsigned char** out) {
    if (dev < 0 || dev >= kMaxDevices) return fail(VO_ERR_INVALID, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(VO_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[dev] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch[dev];
    return 0;
}
}  // namespace

// =================================================================== C ABI
extern "C" {

int vo_predicates(int hip_device, int type, const double* pts, int64_t n, const double* center, int32_t* signs,
                  int32_t* tiers) {
    if (type < 1 || type > 4) return fail(VO_ERR_INVALID, "type must be 1 (orient2d), 2 (incircle), 3 (orient3d) or 4 (insphere)");
    if (n < 0 || (n > 0 && (!pts || !center || !signs || !tiers))) return fail(VO_ERR_INVALID, "null argument");
    if (n > (int64_t)1 << 28) return fail(VO_ERR_INVALID, "too many predicates for one call");
    if (n == 0) return 0;
    const int D = type <= 2 ? 2 : 3;
    const int A = pred_arity((uint32_t)type);
    for (int64_t i = 0; i < n * A * D; ++i)
        if (!std::isfinite(pts[i])) return fail(VO_ERR_INVALID, "non-finite coordinate");
    int ndev = 0;
    VO_HIP(hipGetDeviceCount(&ndev));
    if (hip_device < 0 || hip_device >= ndev) return fail(VO_ERR_INVALID, "no such HIP device");
    VO_HIP(hipSetDevice(hip_device));
    const int threads = 64, blocks = 256, ws_cap = 1 << 12;
    double *dp = nullptr, *ws = nullptr;
    int32_t *ds = nullptr, *dt = nullptr;
    int rc;
    auto release = [&] { hipFree(dp); hipFree(ws); hipFree(ds); hipFree(dt); };
    if ((rc = dev_alloc(&dp, (size_t)n * A * D)) || (rc = dev_alloc(&ws, (size_t)threads * blocks * ws_cap)) ||
        (rc = dev_alloc(&ds, (size_t)n)) || (rc = dev_alloc(&dt, (size_t)n))) {
        release();
        return rc;
    }
    hipError_t e = hipMemcpy(dp, pts, sizeof(double) * (size_t)n * A * D, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        if (D == 3)
            hipLaunchKernelGGL(k_eval_predicates<3>, dim3(blocks), dim3(threads), 0, 0, type, dp, (int)n, center[0],
                               center[1], center[2], ws, ws_cap, ds, dt);
        else
            hipLaunchKernelGGL(k_eval_predicates<2>, dim3(blocks), dim3(threads), 0, 0, type, dp, (int)n, center[0],
                               center[1], 0.0, ws, ws_cap, ds, dt);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(signs, ds, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(tiers, dt, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost);
    release();
    if (e != hipSuccess) return fail(VO_ERR_HIP, std::string("vo_predicates: ") + hipGetErrorString(e));
    return 0;
}

int vo_predicates_noalloc(int hip_device, int type, const double* pts, int64_t n, const double* center,
                          int32_t* signs, int32_t* tiers) {
    if (type < 1 || type > 4) return fail(VO_ERR_INVALID, "type must be 1 (orient2d), 2 (incircle), 3 (orient3d) or 4 (insphere)");
    if (n < 0 || (n > 0 && (!pts || !center || !signs || !tiers))) return fail(VO_ERR_INVALID, "null argument");
    if (n > kNaMaxN)
        return fail(VO_ERR_INVALID, "vo_predicates_noalloc: " + std::to_string(n) + " instances exceed the maximum of " +
                                        std::to_string(kNaMaxN) + " per call (split the batch, or use vo_predicates)");
    if (n == 0) return 0;
    const int D = type <= 2 ? 2 : 3;
    const int A = pred_arity((uint32_t)type);
    for (int64_t i = 0; i < n * A * D; ++i)
        if (!std::isfinite(pts[i])) return fail(VO_ERR_INVALID, "non-finite coordinate");
    int ndev = 0;
    VO_HIP(hipGetDeviceCount(&ndev));
    if (hip_device < 0 || hip_device >= ndev) return fail(VO_ERR_INVALID, "no such HIP device");
    VO_HIP(hipSetDevice(hip_device));
    unsigned char* base = nullptr;
    int rc;
    if ((rc = scratch_on(hip_device, &base))) return rc;
    double* dp = reinterpret_cast<double*>(base);
    double* ws = reinterpret_cast<double*>(base + kNaPts);
    int32_t* ds = reinterpret_cast<int32_t*>(base + kNaPts + kNaWs);
    int32_t* dt = reinterpret_cast<int32_t*>(base + kNaPts + kNaWs + kNaOut);
    // every sign, tier and workspace entry a call reads is written by that call first: nothing
    // left by an earlier call needs clearing
    VO_HIP(hipMemcpy(dp, pts, sizeof(double) * (size_t)n * A * D, hipMemcpyHostToDevice));
    if (D == 3)
        hipLaunchKernelGGL(k_eval_predicates<3>, dim3(kNaBlocks), dim3(kNaThreadsPerBlock), 0, 0, type, dp, (int)n,
                           center[0], center[1], center[2], ws, kNaWsCap, ds, dt);
    else
        hipLaunchKernelGGL(k_eval_predicates<2>, dim3(kNaBlocks), dim3(kNaThreadsPerBlock), 0, 0, type, dp, (int)n,
                           center[0], center[1], 0.0, ws, kNaWsCap, ds, dt);
    VO_HIP(hipGetLastError());
    VO_HIP(hipMemcpy(signs, ds, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost));
    VO_HIP(hipMemcpy(tiers, dt, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost));
    return 0;
}

const char* vo_last_error(void) { return vo_host::g_err.c_str(); }
int vo_version(void) { return 1; }

int vo_bounding_sphere(int dim, const double* pts, int64_t n, double* center, double* radius) {
    if (dim != 2 && dim != 3) return fail(VO_ERR_INVALID, "dim must be 2 or 3");
    if (n < 1 || !pts || !center || !radius) return fail(VO_ERR_INVALID, "bad argument");
    vo_host::bounding_sphere(dim, pts, n, center, radius);
    return 0;
}
void vo_free(void* p) { std::free(p); }

int vo_circumsphere(int dim, const double* verts, double* center, double* radius) {
    if (dim != 2 && dim != 3) return fail(VO_ERR_INVALID, "dim must be 2 or 3");
    if (!verts || !center || !radius) return fail(VO_ERR_INVALID, "null argument");
    if (!vo_host::ref_circumsphere(dim, verts, center, radius))
        return fail(VO_ERR_INVALID, "degenerate simplex: the LU solve has a zero pivot");
    return 0;
}

int vo_in_sphere(int dim, const double* p, const double* center, double radius) {
    if (dim != 2 && dim != 3) return fail(VO_ERR_INVALID, "dim must be 2 or 3");
    if (!p || !center) return fail(VO_ERR_INVALID, "null argument");
    return vo_host::ref_in_sphere(dim, p, center, radius) ? 1 : 0;
}

int vo_circumspheres(int dim, const double* coords, int64_t n_coords, const int32_t* simplices, int64_t n,
                     double* centers, double* radii) {
    if (dim != 2 && dim != 3) return fail(VO_ERR_INVALID, "dim must be 2 or 3");
    if (n < 0 || (n > 0 && (!coords || !simplices || !centers || !radii))) return fail(VO_ERR_INVALID, "null argument");
    const int M = dim + 1;
    for (int64_t i = 0; i < n; ++i) {
        double v[12];
        for (int k = 0; k < M; ++k) {
            const int32_t id = simplices[i * M + k];
            if (id < 0 || id >= n_coords) return fail(VO_ERR_INVALID, "vertex id out of range");
            for (int j = 0; j < dim; ++j) v[k * dim + j] = coords[(int64_t)id * dim + j];
        }
        vo_host::ref_circumsphere(dim, v, centers + i * dim, radii + i);   // degenerate: NaN
    }
    return 0;
}

int vo_tri_create(int hip_device, int dim, const double* pts, int64_t n, vo_tri** out) {
    if (!out) return fail(VO_ERR_INVALID, "out is null");
    *out = nullptr;
    if (dim != 2 && dim != 3) return fail(VO_ERR_INVALID, "dim must be 2 or 3");
    if (n < 0 || (n > 0 && !pts)) return fail(VO_ERR_INVALID, "bad point buffer");
    for (int64_t i = 0; i < n * dim; ++i)
        if (!std::isfinite(pts[i])) return fail(VO_ERR_INVALID, "non-finite coordinate");
    int ndev = 0;
    VO_HIP(hipGetDeviceCount(&ndev));
    if (hip_device < 0 || hip_device >= ndev) return fail(VO_ERR_INVALID, "no such HIP device");
    VO_HIP(hipSetDevice(hip_device));
    vo_tri* t = new vo_tri();
    t->dev = hip_device;
    t->D = dim;
    t->M = dim + 1;
    t->S = dim == 3 ? 4 : 2;
    int rc = 0;
    auto bail = [&](int code) {
        free_all(t);
        delete t;
        return code;
    };
    if (hipStreamCreateWithFlags(&t->s, hipStreamNonBlocking) != hipSuccess) return bail(fail(VO_ERR_HIP, "stream"));
    for (auto& e : t->ev)
        if (hipEventCreate(&e) != hipSuccess) return bail(fail(VO_ERR_HIP, "event"));
    set_frame(dim, pts, n, t->c, &t->r);
    t->R = t->r * 10.0;
    super_vertices(dim, t->c, t->R, t->sv);
    first_simplex_order(dim, t->sv, t->tet0);
    // windows: capacity and per-point cavity buffers
    // a window never holds more than the set's points
    {
        int wc = 4096;
        while (wc < n && wc < (dim == 3 ? (1 << 20) : (1 << 21))) wc <<= 1;
        t->win_cap = wc;
    }
    t->kcap = dim == 3 ? 256 : 96;
    t->fcap = dim == 3 ? 384 : 128;
    t->qcap = 1 << 20;
    t->cmask = (1u << 22) - 1;
    t->ws_threads = 1024;
    t->ws_cap = 1 << 14;
    t->gamma = dim == 3 ? 0.2 : 0.6;
    t->big_window = dim == 3 ? 16384 : 65536;
    // measured at 3D n = 4M: 95 -> 109 ms with 256 (one CU holds too few lanes for 3D cavities and
    // commits); 2D n = 10M: 35.4 -> 34.9 ms
    t->boot_window = dim == 3 ? 0 : 256;
    // large 3D windows aim at 65% winners (with shared facets: 3D n = 4M 115 -> 112 ms; 2D unchanged)
    t->big_frac_lo = dim == 3 ? 0.65 : 0.55;
    if ((rc = dev_alloc(&t->win, t->win_cap)) || (rc = dev_alloc(&t->pend[0], t->win_cap)) ||
        (rc = dev_alloc(&t->pend[1], t->win_cap)) || (rc = dev_alloc(&t->cav, (size_t)t->win_cap * t->kcap)) ||
        (rc = dev_alloc(&t->fac, (size_t)t->win_cap * t->fcap)) || (rc = dev_alloc(&t->fnb, (size_t)t->win_cap * t->fcap)) ||
        (rc = dev_alloc(&t->nk, t->win_cap)) ||
        (rc = dev_alloc(&t->nf, t->win_cap)) || (rc = dev_alloc(&t->status, t->win_cap)) ||
        (rc = dev_alloc(&t->st, 1)) || (rc = dev_alloc(&t->cnt, kCntStripes)) || (rc = dev_alloc(&t->q, t->qcap)) ||
        (rc = dev_alloc(&t->cache, (size_t)t->cmask + 1)) || (rc = dev_alloc(&t->ws, (size_t)t->ws_threads * t->ws_cap)) ||
        (rc = dev_alloc(&t->big, t->win_cap)) || (rc = dev_alloc(&t->cav1, t->big_cap)) ||
        (rc = dev_alloc(&t->fac1, t->big_cap)) || (rc = dev_alloc(&t->misc1, 4)) ||
        (rc = dev_alloc(&t->fix, (size_t)2 * t->win_cap)) || (rc = dev_alloc(&t->ast, kAllocStripes)) ||
        (rc = dev_alloc(&t->pflag, t->win_cap)))
        return bail(rc);
    t->fix_cap = 2 * t->win_cap;
    if (hipHostMalloc((void**)&t->st_host, sizeof(RoundState)) != hipSuccess) return bail(fail(VO_ERR_OOM, "pinned"));
    hipMemsetAsync(t->st, 0, sizeof(RoundState), t->s);
    hipMemsetAsync(t->cnt, 0, sizeof(Counters) * kCntStripes, t->s);
    hipMemsetAsync(t->cache, 0xFF, sizeof(CEntry) * ((size_t)t->cmask + 1), t->s);
    if ((rc = grow_points(t, std::max<int64_t>(n, 16)))) return bail(rc);
    if ((rc = upload_super(t))) return bail(rc);
    if ((rc = grow_tets(t, 1024))) return bail(rc);
    {
        unsigned char r0[128];
        root_record(dim, t->tet0, 0, t->sv, t->c, r0);
        if (hipMemcpyAsync(t->rec, r0, (size_t)record_bytes(dim), hipMemcpyHostToDevice, t->s) != hipSuccess)
            return bail(fail(VO_ERR_HIP, "init record"));
    }
    RoundState init = {};
    init.n_tets = 1;
    if (hipMemcpyAsync(t->st, &init, sizeof(init), hipMemcpyHostToDevice, t->s) != hipSuccess ||
        hipStreamSynchronize(t->s) != hipSuccess)
        return bail(fail(VO_ERR_HIP, "init"));
    *out = t;
    return 0;
}

void vo_tri_destroy(vo_tri* t) {
    if (!t) return;
    hipSetDevice(t->dev);
    free_all(t);
    delete t;
}

int vo_tri_reset(vo_tri* t, const double* pts, int64_t n) {
    if (!t) return fail(VO_ERR_INVALID, "null handle");
    if (n < 0 || (n > 0 && !pts)) return fail(VO_ERR_INVALID, "bad point buffer");
    const int dim = t->D;
    for (int64_t i = 0; i < n * dim; ++i)
        if (!std::isfinite(pts[i])) return fail(VO_ERR_INVALID, "non-finite coordinate");
    VO_HIP(hipSetDevice(t->dev));
    VO_HIP(hipStreamSynchronize(t->s));
    // the frame and super simplex of the new point set, as vo_tri_create
    set_frame(dim, pts, n, t->c, &t->r);
    t->R = t->r * 10.0;
    super_vertices(dim, t->c, t->R, t->sv);
    first_simplex_order(dim, t->sv, t->tet0);
    // host state of an empty tree; the round counter goes on, so that every reservation key
    // of the new set is below the keys left in the records of the old one
    t->n_pts = t->head = 0;
    t->input_of.clear();
    // records above the count stay fresh (all 0xFF: dead), which the allocation stripes rely on:
    // a chunk's unused slots lie below the count and must not look live
    if (t->n_tets > 1)
        VO_HIP(hipMemsetAsync(t->rec + record_bytes(dim), 0xFF, (size_t)(t->n_tets - 1) * record_bytes(dim), t->s));
    t->n_tets = 1;
    t->n_live = 1;
    t->retired = 0;
    t->cur = t->npend = 0;
    t->inserted = t->dups = t->windows_total = 0;
    t->rounds0 = t->rounds;
    t->stalled = 0;
    t->exact_dev = 0;
    t->gamma = dim == 3 ? 0.2 : 0.6;
    t->ms_run = t->ms_cav = t->ms_com = t->ms_ex = t->ms_other = t->ms_compact = 0;
    t->n_cav = t->n_com = t->n_solo = t->n_compact = 0;
    t->dirty_dup = false;
    // device state: super vertices, the first record, cleared round state and statistics, no
    // cached exact signs (they are keyed by vertex ids), and a new hint grid at the next stage
    if (t->hint) {
        hipFree(t->hint);
        t->hint = nullptr;
    }
    int rc;
    if ((rc = grow_points(t, std::max<int64_t>(n, 16))) || (rc = upload_super(t))) return rc;
    unsigned char r0[128];
    root_record(dim, t->tet0, 0, t->sv, t->c, r0);
    VO_HIP(hipMemcpyAsync(t->rec, r0, (size_t)record_bytes(dim), hipMemcpyHostToDevice, t->s));
    RoundState init = {};
    init.n_tets = 1;
    VO_HIP(hipMemcpyAsync(t->st, &init, sizeof(init), hipMemcpyHostToDevice, t->s));
    VO_HIP(hipMemsetAsync(t->cnt, 0, sizeof(Counters) * kCntStripes, t->s));
    VO_HIP(hipMemsetAsync(t->cache, 0xFF, sizeof(CEntry) * ((size_t)t->cmask + 1), t->s));
    VO_HIP(hipStreamSynchronize(t->s));
    return 0;
}

int vo_tri_stage(vo_tri* t, const double* pts, int64_t n, uint64_t seed) {
    if (!t) return fail(VO_ERR_INVALID, "null handle");
    if (n < 0 || (n > 0 && !pts)) return fail(VO_ERR_INVALID, "bad point buffer");
    if (n == 0) return 0;
    const int D = t->D, S = t->S, M = t->M;
    for (int64_t i = 0; i < n * D; ++i)
        if (!std::isfinite(pts[i])) return fail(VO_ERR_INVALID, "non-finite coordinate");
    VO_HIP(hipSetDevice(t->dev));
    int rc;
    if ((rc = grow_points(t, t->n_pts + n))) return rc;
    const std::vector<int32_t> perm =
        t->local_order ? priority_order_local(D, pts, n, seed, t->c, t->r, t->local_order) : priority_order(n, seed);
    // ids are committed (n_pts, input_of) only once every upload and allocation has succeeded,
    // so that a failed stage leaves the handle as it was
    const int64_t base_in = t->n_pts;
    std::vector<int32_t> in_of(n);
    for (int64_t i = 0; i < n; ++i) in_of[i] = (int32_t)(base_in + perm[i]);
    const int64_t v0 = M + t->n_pts;
    {
        // the points as given and the order go up; the device puts them in order (a gather in
        // random order costs the host ~190 ms at n = 4M, the device well under a millisecond)
        DevBuf<double> dpts;
        DevBuf<int32_t> dperm;
        if ((rc = dev_alloc(&dpts.p, (size_t)n * D)) || (rc = dev_alloc(&dperm.p, (size_t)n))) return rc;
        VO_HIP(hipMemcpyAsync(dpts.p, pts, sizeof(double) * (size_t)n * D, hipMemcpyHostToDevice, t->s));
        VO_HIP(hipMemcpyAsync(dperm.p, perm.data(), sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, t->s));
        const int grid = (int)std::min<int64_t>((n + 255) / 256, 4096);
        hipLaunchKernelGGL(k_stage_points, dim3(grid), dim3(256), 0, t->s, dpts.p, dperm.p, n, D, S, t->c[0], t->c[1],
                           D == 3 ? t->c[2] : 0.0, t->X64 + v0 * S, t->X32 + v0 * S, t->vmap + v0,
                           (int64_t)(2 * M + base_in));
        VO_HIP(hipGetLastError());
        VO_HIP(hipStreamSynchronize(t->s));     // before the staging buffers are freed
    }
    VO_HIP(hipMemsetAsync(t->cp + t->n_pts, 0xFF, sizeof(int) * n, t->s));
    const int64_t n_after = t->n_pts + n;
    // walk-start hints: grids of 2^l cells per axis over the bounding sphere's box, finest
    // level about 4 points a cell; level 0 holds the initial simplex
    if (!t->hint) {
        const int L = hint_levels(D, n_after);
        t->hlevels = L;
        const int64_t cells = hint_cells(D, L);
        for (int k = 0; k < 3; ++k) t->glo[k] = (float)(-t->r);
        t->ginv = (float)(1.0 / (2.0 * t->r));
        if ((rc = dev_alloc(&t->hint, cells))) return rc;
        VO_HIP(hipMemsetAsync(t->hint, 0xFF, sizeof(int) * cells, t->s));
        VO_HIP(hipMemsetAsync(t->hint, 0, sizeof(int), t->s));
    }
    // simplex capacity for everything staged (uniform data creates ~27 (3D) / 6 (2D) per point)
    choose_reuse(t, n_after - t->inserted - t->dups);
    const int64_t want = (int64_t)t->n_tets + records_per_point(t) * (n_after - t->inserted - t->dups) + 4096;
    if ((rc = grow_tets(t, want))) return rc;
    VO_HIP(hipStreamSynchronize(t->s));
    t->n_pts = n_after;
    t->input_of.insert(t->input_of.end(), in_of.begin(), in_of.end());
    return 0;
}

int vo_tri_run(vo_tri* t) {
    if (!t) return fail(VO_ERR_INVALID, "null handle");
    VO_HIP(hipSetDevice(t->dev));
    choose_reuse(t, t->n_pts - t->head + t->npend);
    t->ms_cav = t->ms_com = t->ms_ex = t->ms_other = 0;
    t->n_cav = t->n_com = 0;
    t->stalled = 0;
    t->exact_dev = 0;
    const int64_t dups0 = t->dups;
    VO_HIP(hipEventRecord(t->ev[0], t->s));
    // VO_TRACE=<path>: append one line per round (diagnostics only; rounds planned on the host)
    struct TraceFile {
        FILE* f = nullptr;
        ~TraceFile() { if (f) std::fclose(f); }
    } tf;
    if (const char* tp = std::getenv("VO_TRACE")) tf.f = std::fopen(tp, "a");
    // the per-round host decisions of compaction and XCD regions need the host loop; so does a
    // small run (a few points need a round or two, and the device planner submits at least two
    // chunks of rounds: one point through the host loop takes about a third of the time)
    const int64_t todo = t->n_pts - t->head + t->npend;
    const bool host_loop = !t->pipelined || tf.f || t->xcd || t->compact_ratio > 0 || todo <= t->small_run;
    int rc = host_loop ? run_host_planned(t, tf.f) : run_device_planned(t);
    if (rc) return rc;
    VO_HIP(hipEventRecord(t->ev[1], t->s));
    VO_HIP(hipEventSynchronize(t->ev[1]));
    float ms = 0;
    hipEventElapsedTime(&ms, t->ev[0], t->ev[1]);
    t->ms_run = ms;
    t->ms_other = ms - t->ms_cav - t->ms_com - t->ms_ex;
    if (t->dups > dups0) return fail(VO_ERR_DUPLICATE, std::to_string(t->dups - dups0) + " duplicate points skipped");
    return 0;
}

int vo_tri_insert(vo_tri* t, const double* pts, int64_t n, uint64_t seed) {
    int rc = vo_tri_stage(t, pts, n, seed);
    return rc ? rc : vo_tri_run(t);
}

int vo_tri_counts(const vo_tri* t, int64_t* n_vertices, int64_t* n_simplices, int64_t* max_simplex_id) {
    if (!t) return fail(VO_ERR_INVALID, "null handle");
    if (n_vertices) *n_vertices = t->M + t->inserted;
    if (n_simplices) *n_simplices = t->n_live;
    if (max_simplex_id) {
        // simplices ever created (records are reused, so they are counted, not the records)
        VO_HIP(hipSetDevice(t->dev));
        std::vector<Counters> hs(kCntStripes);
        VO_HIP(hipMemcpy(hs.data(), t->cnt, sizeof(Counters) * kCntStripes, hipMemcpyDeviceToHost));
        *max_simplex_id = (int64_t)sum_counters(hs.data()).created + t->M;
    }
    return 0;
}

// export position of every record (exclusive scan of the live flags) and the live count
static int live_positions(vo_tri* t, DevBuf<int>& pos, int64_t* n_live) {
    const int n = t->n_tets;
    const int nb = (n + 255) / 256;
    DevBuf<int> flag, bsum;
    int rc;
    if ((rc = idle_or_alloc(flag, t->cav, sizeof(int) * (size_t)t->win_cap * t->kcap, (size_t)n)) ||
        (rc = idle_or_alloc(pos, t->fac, sizeof(int) * (size_t)t->win_cap * t->fcap, (size_t)n)) ||
        (rc = idle_or_alloc(bsum, t->big, sizeof(int) * (size_t)t->win_cap, (size_t)nb)))
        return rc;
    const Ctx c = make_ctx(t);
    if (t->D == 3) hipLaunchKernelGGL(k_live_flags<3>, dim3(nb), dim3(256), 0, t->s, c, n, flag.p);
    else hipLaunchKernelGGL(k_live_flags<2>, dim3(nb), dim3(256), 0, t->s, c, n, flag.p);
    hipLaunchKernelGGL(k_block_scan, dim3(nb), dim3(256), 0, t->s, flag.p, pos.p, bsum.p, n);
    std::vector<int> hb(nb);
    VO_HIP(hipMemcpyAsync(hb.data(), bsum.p, sizeof(int) * nb, hipMemcpyDeviceToHost, t->s));
    VO_HIP(hipStreamSynchronize(t->s));
    int64_t acc = 0;
    for (int i = 0; i < nb; ++i) { int v = hb[i]; hb[i] = (int)acc; acc += v; }
    VO_HIP(hipMemcpyAsync(bsum.p, hb.data(), sizeof(int) * nb, hipMemcpyHostToDevice, t->s));
    hipLaunchKernelGGL(k_add_offsets, dim3(nb), dim3(256), 0, t->s, pos.p, bsum.p, n);
    VO_HIP(hipGetLastError());
    VO_HIP(hipStreamSynchronize(t->s));     // before bsum and flag are freed
    *n_live = acc;
    return 0;
}

static int export_device(const vo_tri* tc, DevBuf<int32_t>& simp, DevBuf<int32_t>& nbr, int64_t* n_live_out) {
    vo_tri* t = const_cast<vo_tri*>(tc);
    const int n = t->n_tets;
    const int nb = (n + 255) / 256;
    DevBuf<int> pos;
    int64_t acc = 0;
    int rc;
    if ((rc = live_positions(t, pos, &acc))) return rc;
    const int M = t->M;
    const size_t fnb_bytes = sizeof(int) * (size_t)t->win_cap * t->fcap;
    if (sizeof(int32_t) * 2 * (size_t)std::max<int64_t>(acc, 1) * M <= fnb_bytes) {
        simp.p = t->fnb;
        nbr.p = t->fnb + (size_t)std::max<int64_t>(acc, 1) * M;
        simp.own = nbr.own = false;
    } else if ((rc = dev_alloc(&simp.p, acc * M)) || (rc = dev_alloc(&nbr.p, acc * M))) {
        return rc;
    }
    const Ctx c = make_ctx(t);
    if (t->D == 3)
        hipLaunchKernelGGL(k_export<3>, dim3(nb), dim3(256), 0, t->s, c, pos.p, n, t->vmap, simp.p, nbr.p);
    else
        hipLaunchKernelGGL(k_export<2>, dim3(nb), dim3(256), 0, t->s, c, pos.p, n, t->vmap, simp.p, nbr.p);
    VO_HIP(hipGetLastError());
    VO_HIP(hipStreamSynchronize(t->s));
    *n_live_out = acc;
    return 0;
}

int vo_tri_export(const vo_tri* t, int32_t* simplices, int32_t* neighbors) {
    if (!t) return fail(VO_ERR_INVALID, "null handle");
    VO_HIP(hipSetDevice(t->dev));
    DevBuf<int32_t> ds, dn;
    int64_t nl;
    int rc = export_device(t, ds, dn, &nl);
    if (rc) return rc;
    // the caller's arrays hold n_live rows (vo_tri_counts): never write past them
    if (nl != t->n_live) return fail(VO_ERR_INTERNAL, "live records differ from the live simplex count");
    const size_t bytes = sizeof(int32_t) * (size_t)nl * t->M;
    if (simplices && nl) VO_HIP(hipMemcpy(simplices, ds.p, bytes, hipMemcpyDeviceToHost));
    if (neighbors && nl) VO_HIP(hipMemcpy(neighbors, dn.p, bytes, hipMemcpyDeviceToHost));
    return 0;
}

// Canonical form on the device (canon.hpp): real simplices as sorted tuples in lexicographic
// order (two stable radix sorts: columns 2-3, then columns 0-1), real edges sorted and unique.
// The same output as canonical_from_simplices on the host, which the engine's results use.
static int canonical_device(const vo_tri* t, int32_t** T, int64_t* n_T, int32_t** E, int64_t* n_E) {
    const int M = t->M;
    const hipStream_t s = t->s;
    DevBuf<int32_t> ds, dn;
    int64_t nl = 0;
    int rc = export_device(t, ds, dn, &nl);
    if (rc) return rc;
    if (dn.own) hipFree(dn.p);
    dn.p = nullptr;
    const int first = 2 * M;
    const int nb = std::max(1, bit_length((int64_t)t->input_of.size() - 1));
    const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(4096, (nl + 255) / 256));
    // ---- simplices
    int64_t nt = 0;
    DevBuf<int32_t> dT;
    {
        DevBuf<int4> rows, crow;
        DevBuf<int> flag, pos;
        if ((rc = dev_alloc(&rows.p, (size_t)std::max<int64_t>(nl, 1))) || (rc = dev_alloc(&flag.p, (size_t)std::max<int64_t>(nl, 1))) ||
            (rc = dev_alloc(&pos.p, (size_t)std::max<int64_t>(nl, 1))))
            return rc;
        if (M == 4) hipLaunchKernelGGL(k_canon_rows<4>, dim3(grid), dim3(256), 0, s, ds.p, nl, first, rows.p, flag.p);
        else hipLaunchKernelGGL(k_canon_rows<3>, dim3(grid), dim3(256), 0, s, ds.p, nl, first, rows.p, flag.p);
        VO_HIP(hipGetLastError());
        if ((rc = device_excl_scan(s, flag.p, pos.p, nl, &nt))) return rc;
        const size_t nt1 = (size_t)std::max<int64_t>(nt, 1);
        DevBuf<unsigned long long> klo, khi, ktmp;
        DevBuf<uint32_t> perm, ptmp;
        if ((rc = dev_alloc(&crow.p, nt1)) || (rc = dev_alloc(&klo.p, nt1)) || (rc = dev_alloc(&khi.p, nt1)) ||
            (rc = dev_alloc(&ktmp.p, nt1)) || (rc = dev_alloc(&perm.p, nt1)) || (rc = dev_alloc(&ptmp.p, nt1)) ||
            (rc = dev_alloc(&dT.p, nt1 * M)))
            return rc;
        if (M == 4) hipLaunchKernelGGL(k_canon_keys<4>, dim3(grid), dim3(256), 0, s, rows.p, flag.p, pos.p, nl, nb, crow.p, klo.p, khi.p);
        else hipLaunchKernelGGL(k_canon_keys<3>, dim3(grid), dim3(256), 0, s, rows.p, flag.p, pos.p, nl, nb, crow.p, klo.p, khi.p);
        hipLaunchKernelGGL(k_iota_u32, dim3(grid), dim3(256), 0, s, perm.p, nt);
        VO_HIP(hipGetLastError());
        // least significant columns first, then the leading two (stable): lexicographic order
        if ((rc = radix_sort(s, klo.p, perm.p, ktmp.p, ptmp.p, nt, M == 4 ? 2 * nb : nb))) return rc;
        hipLaunchKernelGGL(k_gather_u64, dim3(grid), dim3(256), 0, s, khi.p, perm.p, klo.p, nt);
        VO_HIP(hipGetLastError());
        if ((rc = radix_sort(s, klo.p, perm.p, ktmp.p, ptmp.p, nt, 2 * nb))) return rc;
        if (M == 4) hipLaunchKernelGGL(k_canon_out_rows<4>, dim3(grid), dim3(256), 0, s, crow.p, perm.p, nt, dT.p);
        else hipLaunchKernelGGL(k_canon_out_rows<3>, dim3(grid), dim3(256), 0, s, crow.p, perm.p, nt, dT.p);
        VO_HIP(hipGetLastError());
    }
    // ---- edges
    int64_t ne = 0;
    DevBuf<int32_t> dE;
    {
        const int P = M * (M - 1) / 2;
        const int64_t nk = nl * P;
        const size_t nk1 = (size_t)std::max<int64_t>(nk, 1);
        DevBuf<unsigned long long> keys, ktmp;
        DevBuf<int> flag, pos;
        uint32_t* nov = nullptr;
        uint32_t* novt = nullptr;
        if ((rc = dev_alloc(&keys.p, nk1)) || (rc = dev_alloc(&ktmp.p, nk1)) || (rc = dev_alloc(&flag.p, nk1)) ||
            (rc = dev_alloc(&pos.p, nk1)))
            return rc;
        if (M == 4) hipLaunchKernelGGL(k_canon_edge_keys<4>, dim3(grid), dim3(256), 0, s, ds.p, nl, first, nb, keys.p);
        else hipLaunchKernelGGL(k_canon_edge_keys<3>, dim3(grid), dim3(256), 0, s, ds.p, nl, first, nb, keys.p);
        VO_HIP(hipGetLastError());
        if ((rc = radix_sort(s, keys.p, nov, ktmp.p, novt, nk, 2 * nb))) return rc;
        const unsigned egrid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(8192, (nk + 255) / 256));
        hipLaunchKernelGGL(k_canon_edge_flags, dim3(egrid), dim3(256), 0, s, keys.p, nk, flag.p);
        VO_HIP(hipGetLastError());
        if ((rc = device_excl_scan(s, flag.p, pos.p, nk, &ne))) return rc;
        if ((rc = dev_alloc(&dE.p, (size_t)std::max<int64_t>(ne, 1) * 2))) return rc;
        hipLaunchKernelGGL(k_canon_edge_out, dim3(egrid), dim3(256), 0, s, keys.p, flag.p, pos.p, nk, nb, dE.p);
        VO_HIP(hipGetLastError());
    }
    *T = (int32_t*)std::malloc(sizeof(int32_t) * (size_t)std::max<int64_t>(1, nt * M));
    *E = (int32_t*)std::malloc(sizeof(int32_t) * (size_t)std::max<int64_t>(1, ne * 2));
    if (!*T || !*E) {
        std::free(*T);
        std::free(*E);
        *T = *E = nullptr;
        return fail(VO_ERR_OOM, "host allocation");
    }
    if (nt) VO_HIP(hipMemcpyAsync(*T, dT.p, sizeof(int32_t) * (size_t)nt * M, hipMemcpyDeviceToHost, s));
    if (ne) VO_HIP(hipMemcpyAsync(*E, dE.p, sizeof(int32_t) * (size_t)ne * 2, hipMemcpyDeviceToHost, s));
    VO_HIP(hipStreamSynchronize(s));
    *n_T = nt;
    *n_E = ne;
    return 0;
}

int vo_tri_canonical(const vo_tri* t, int32_t** T, int64_t* n_T, int32_t** E, int64_t* n_E) {
    if (!t || !T || !n_T || !E || !n_E) return fail(VO_ERR_INVALID, "null argument");
    VO_HIP(hipSetDevice(t->dev));
    return canonical_device(t, T, n_T, E, n_E);
}

}  // extern "C"

namespace {

// Violation counts of the two checks, with the exact tier's retry: a pass whose predicates
// needed the expansions queues them; they are evaluated into the cache and the pass is run
// again (each pass resolves what the last one queued).  pairs = 0: the local check (k_check),
// 1: the reference's all-pairs check (k_check_pairs).
int run_check(vo_tri* t, int pairs, int64_t* out) {
    DevBuf<unsigned long long> bad;
    DevBuf<int> unres;
    int rc;
    if ((rc = dev_alloc(&bad.p, 1)) || (rc = dev_alloc(&unres.p, 1))) return rc;
    const int n = t->n_tets;
    const int nv = (int)(t->M + t->n_pts);
    // super vertex w stands for itself and its ghost copies (3D: vertex 0 has ghosts 4, 5, 6
    // and vertex 1 has ghost 7; 2D: vertices 0, 1, 2 have ghosts 3, 4, 5)
    const int4 weight = t->D == 3 ? make_int4(4, 2, 1, 1) : make_int4(2, 2, 2, 0);
    for (int attempt = 0; attempt < 16; ++attempt) {
        VO_HIP(hipMemsetAsync(bad.p, 0, sizeof(unsigned long long), t->s));
        VO_HIP(hipMemsetAsync(unres.p, 0, sizeof(int), t->s));
        hipLaunchKernelGGL(k_reset_round, dim3(1), dim3(1), 0, t->s, t->st);
        Ctx c = make_ctx(t);
        const unsigned grid = (unsigned)std::max(1, std::min(n, 8192));
        if (pairs && t->D == 3)
            hipLaunchKernelGGL(k_check_pairs<3>, dim3(grid), dim3(256), 0, t->s, c, n, nv, weight, bad.p, unres.p);
        else if (pairs)
            hipLaunchKernelGGL(k_check_pairs<2>, dim3(grid), dim3(256), 0, t->s, c, n, nv, weight, bad.p, unres.p);
        else if (t->D == 3)
            hipLaunchKernelGGL(k_check<3>, dim3((n + 255) / 256), dim3(256), 0, t->s, c, n, bad.p, unres.p);
        else
            hipLaunchKernelGGL(k_check<2>, dim3((n + 255) / 256), dim3(256), 0, t->s, c, n, bad.p, unres.p);
        VO_HIP(hipGetLastError());
        unsigned long long hb = 0;
        int hu = 0;
        VO_HIP(hipMemcpyAsync(&hb, bad.p, sizeof(hb), hipMemcpyDeviceToHost, t->s));
        VO_HIP(hipMemcpyAsync(&hu, unres.p, sizeof(hu), hipMemcpyDeviceToHost, t->s));
        if ((rc = sync_state(t))) return rc;
        if (hu == 0) {
            *out = (int64_t)hb;
            return check_err(t);
        }
        if ((rc = run_exact(t, t->st_host->n_q))) return rc;
    }
    return fail(VO_ERR_INTERNAL, "check: predicates stayed unresolved");
}

// internal vertex index of reference vertex id `id` (an inserted or staged input point), -1
int vertex_of_id(const vo_tri* t, int64_t id) {
    const int64_t in = id - 2 * t->M;
    if (in < 0) return -1;
    for (size_t i = 0; i < t->input_of.size(); ++i)
        if (t->input_of[i] == in) return (int)(t->M + i);
    return -1;
}

}  // namespace

extern "C" {

int vo_tri_check(vo_tri* t, int64_t* n_violations) {
    if (!t || !n_violations) return fail(VO_ERR_INVALID, "null argument");
    VO_HIP(hipSetDevice(t->dev));
    return run_check(t, 0, n_violations);
}

int vo_tri_check_pairs(vo_tri* t, int64_t* n_pairs) {
    if (!t || !n_pairs) return fail(VO_ERR_INVALID, "null argument");
    const double work = (double)t->n_live * (double)(t->M + t->n_pts);
    if (work > (double)((int64_t)1 << 36))
        return fail(VO_ERR_INVALID, "vo_tri_check_pairs is O(simplices x vertices), limited to 2^36 pairs; "
                                    "use vo_tri_check for large trees");
    VO_HIP(hipSetDevice(t->dev));
    return run_check(t, 1, n_pairs);
}

int vo_tri_locate(vo_tri* t, const double* p, int32_t* killed, int64_t killed_cap, int64_t* n_killed,
                  int32_t* facets, int64_t facets_cap, int64_t* n_facets) {
    if (!t || !p || !n_killed || !n_facets) return fail(VO_ERR_INVALID, "null argument");
    if (killed_cap < 0 || facets_cap < 0 || (killed_cap > 0 && !killed) || (facets_cap > 0 && !facets))
        return fail(VO_ERR_INVALID, "bad output buffer");
    const int D = t->D, M = t->M;
    for (int k = 0; k < D; ++k)
        if (!std::isfinite(p[k])) return fail(VO_ERR_INVALID, "non-finite coordinate");
    VO_HIP(hipSetDevice(t->dev));
    const Ctx c = make_ctx(t);
    const int start = t->hint ? -1 : 0;          // no hint grid before the first stage: the first record
    float f[3] = {0, 0, 0};
    double x[3] = {0, 0, 0};
    for (int k = 0; k < D; ++k) {
        x[k] = p[k];
        f[k] = (float)(p[k] - t->c[k]);
    }
    if (D == 3) {
        QueryPoint<3> q{make_float4(f[0], f[1], f[2], 0.f), {x[0], x[1], x[2]}};
        hipLaunchKernelGGL(k_locate_query<3>, dim3(1), dim3(64), 0, t->s, c, q, start, t->ws, t->ws_cap, t->cav1, t->fac1,
                           t->big_cap, t->misc1);
    } else {
        QueryPoint<2> q{make_float2(f[0], f[1]), {x[0], x[1], 0.0}};
        hipLaunchKernelGGL(k_locate_query<2>, dim3(1), dim3(64), 0, t->s, c, q, start, t->ws, t->ws_cap, t->cav1, t->fac1,
                           t->big_cap, t->misc1);
    }
    VO_HIP(hipGetLastError());
    int res[3] = {0, 0, 0};
    VO_HIP(hipMemcpyAsync(res, t->misc1, sizeof(res), hipMemcpyDeviceToHost, t->s));
    VO_HIP(hipStreamSynchronize(t->s));
    if (res[2] == 1)
        return fail(VO_ERR_DUPLICATE, "the point coincides with a vertex: no simplex has it strictly inside its sphere");
    if (res[2] == 2) return fail(VO_ERR_OVERFLOW, "the conflict region exceeds the query lists");
    if (res[2] == 3) return fail(VO_ERR_INVALID, "the point lies outside the super simplex");
    if (res[2] == 4) return fail(VO_ERR_PRECISION, "an exact sign left the f64 range or the workspace");
    const int nk = res[0], nf = res[1];
    *n_killed = nk;
    *n_facets = nf;
    if (nk > killed_cap || nf > facets_cap)
        return fail(VO_ERR_OVERFLOW, "output buffers too small: n_killed / n_facets hold the sizes needed");
    DevBuf<int> pos;
    DevBuf<int32_t> dk, df;
    int64_t nl = 0;
    int rc;
    if ((rc = live_positions(t, pos, &nl)) || (rc = dev_alloc(&dk.p, (size_t)nk)) ||
        (rc = dev_alloc(&df.p, (size_t)nf * (M + 1))))
        return rc;
    const unsigned grid = (unsigned)((std::max(nk, nf) + 255) / 256);
    if (D == 3)
        hipLaunchKernelGGL(k_locate_rows<3>, dim3(grid), dim3(256), 0, t->s, c, pos.p, t->vmap, t->cav1, nk, t->fac1, nf,
                           dk.p, df.p);
    else
        hipLaunchKernelGGL(k_locate_rows<2>, dim3(grid), dim3(256), 0, t->s, c, pos.p, t->vmap, t->cav1, nk, t->fac1, nf,
                           dk.p, df.p);
    VO_HIP(hipGetLastError());
    VO_HIP(hipMemcpyAsync(killed, dk.p, sizeof(int32_t) * (size_t)nk, hipMemcpyDeviceToHost, t->s));
    VO_HIP(hipMemcpyAsync(facets, df.p, sizeof(int32_t) * (size_t)nf * (M + 1), hipMemcpyDeviceToHost, t->s));
    VO_HIP(hipStreamSynchronize(t->s));
    std::sort(killed, killed + nk);           // the reference sorts and dedups (src/delaunay_tree.rs:55-56)
    return 0;
}

int vo_tri_set_vertex(vo_tri* t, int64_t id, const double* coords, int records) {
    if (!t || !coords) return fail(VO_ERR_INVALID, "null argument");
    const int D = t->D, S = t->S;
    for (int k = 0; k < D; ++k)
        if (!std::isfinite(coords[k])) return fail(VO_ERR_INVALID, "non-finite coordinate");
    const int vi = vertex_of_id(t, id);
    if (vi < 0) return fail(VO_ERR_INVALID, "no input point has vertex id " + std::to_string(id));
    VO_HIP(hipSetDevice(t->dev));
    double x64[4] = {0, 0, 0, 0};
    float x32[4] = {0, 0, 0, 0};
    for (int k = 0; k < D; ++k) {
        x64[k] = coords[k];
        x32[k] = (float)(coords[k] - t->c[k]);     // as staged: the f64 difference rounded once
    }
    VO_HIP(hipMemcpyAsync(t->X64 + (size_t)vi * S, x64, sizeof(double) * S, hipMemcpyHostToDevice, t->s));
    VO_HIP(hipMemcpyAsync(t->X32 + (size_t)vi * S, x32, sizeof(float) * S, hipMemcpyHostToDevice, t->s));
    if (records && t->n_tets > 0) {
        const Ctx c = make_ctx(t);
        const unsigned grid = (unsigned)((t->n_tets + 255) / 256);
        if (D == 3)
            hipLaunchKernelGGL(k_patch_vertex<3>, dim3(grid), dim3(256), 0, t->s, c, t->n_tets, vi,
                               make_float4(x32[0], x32[1], x32[2], 0.f));
        else
            hipLaunchKernelGGL(k_patch_vertex<2>, dim3(grid), dim3(256), 0, t->s, c, t->n_tets, vi,
                               make_float2(x32[0], x32[1]));
        VO_HIP(hipGetLastError());
    }
    // exact signs cached by vertex ids no longer hold for the moved vertex
    VO_HIP(hipMemsetAsync(t->cache, 0xFF, sizeof(CEntry) * ((size_t)t->cmask + 1), t->s));
    VO_HIP(hipStreamSynchronize(t->s));
    return 0;
}

int vo_tri_swap_simplex_vertices(vo_tri* t, int64_t simplex, int i, int j) {
    if (!t) return fail(VO_ERR_INVALID, "null handle");
    if (i < 0 || j < 0 || i > t->D || j > t->D || i == j) return fail(VO_ERR_INVALID, "i, j must be distinct in 0..dim");
    if (simplex < 0 || simplex >= t->n_live) return fail(VO_ERR_INVALID, "no live simplex at that export index");
    VO_HIP(hipSetDevice(t->dev));
    DevBuf<int> pos;
    int64_t nl = 0;
    int rc;
    if ((rc = live_positions(t, pos, &nl))) return rc;
    const Ctx c = make_ctx(t);
    const unsigned grid = (unsigned)((t->n_tets + 255) / 256);
    if (t->D == 3)
        hipLaunchKernelGGL(k_swap_in_record<3>, dim3(grid), dim3(256), 0, t->s, c, t->n_tets, pos.p, (int)simplex, i, j);
    else
        hipLaunchKernelGGL(k_swap_in_record<2>, dim3(grid), dim3(256), 0, t->s, c, t->n_tets, pos.p, (int)simplex, i, j);
    VO_HIP(hipGetLastError());
    VO_HIP(hipStreamSynchronize(t->s));
    return 0;
}

int vo_tri_stats(const vo_tri* t, vo_stats* out) {
    if (!t || !out) return fail(VO_ERR_INVALID, "null argument");
    VO_HIP(hipSetDevice(t->dev));
    std::vector<Counters> hs(kCntStripes);
    VO_HIP(hipMemcpy(hs.data(), t->cnt, sizeof(Counters) * kCntStripes, hipMemcpyDeviceToHost));
    const Counters h = sum_counters(hs.data());
    std::memset(out, 0, sizeof(*out));
    out->rounds = t->rounds - t->rounds0;
    out->points_inserted = t->inserted;
    out->points_duplicate = t->dups;
    out->windows_total = t->windows_total;
    out->winners_total = (int64_t)h.winners;
    for (int i = 0; i < 2; ++i) {
        out->tier1[i] = (int64_t)h.t1[i];
        out->tier2[i] = (int64_t)h.t2[i];
        out->tier3[i] = (int64_t)h.t3[i];
        out->deferred[i] = (int64_t)h.def[i];
    }
    out->exact_evaluated = (int64_t)h.exact_evaluated;
    out->walk_steps = (int64_t)h.walk_steps;
    out->chain_hops = (int64_t)h.chain_hops;
    out->killed = (int64_t)h.killed;
    out->created = (int64_t)h.created;
    out->max_killed = (int64_t)h.max_killed;
    out->max_facets = (int64_t)h.max_facets;
    out->ms_run = t->ms_run;
    out->ms_locate_cavity = t->ms_cav;
    out->ms_commit = t->ms_com;
    out->ms_exact = t->ms_ex;
    out->ms_other = t->ms_other;
    out->launches_locate_cavity = t->n_cav;
    out->launches_commit = t->n_com;
    out->solo_runs = t->n_solo;
    out->compactions = t->n_compact;
    out->ms_compact = t->ms_compact;
    out->duplicates_detected = (int64_t)h.dups;
    out->records = t->n_tets;
    return 0;
}

int vo_tri_config(vo_tri* t, const char* key, int64_t value) {
    if (!t || !key) return fail(VO_ERR_INVALID, "null argument");
    const std::string k(key);
    if (k == "filters") {
        if (value < 0 || value > 3) return fail(VO_ERR_INVALID, "filters must be a mask of 1 (fp32) and 2 (fp64)");
        t->filters = (int)value;
    } else if (k == "group_lanes") {
        if (value != 0 && value != 1 && value != 4 && value != 8 && value != 16 && value != 32)
            return fail(VO_ERR_INVALID, "group_lanes must be 0 (per round), 1, 4, 8, 16 or 32");
        t->group = (int)value;
    } else if (k == "lanes_target") {
        if (value < 0) return fail(VO_ERR_INVALID, "lanes_target must be >= 0");
        t->lanes_target = value;
    } else if (k == "commit_lanes_target") {
        if (value < 0) return fail(VO_ERR_INVALID, "commit_lanes_target must be >= 0");
        t->commit_lanes_target = value;
    } else if (k == "window_gamma_milli") {
        if (value <= 0) return fail(VO_ERR_INVALID, "window_gamma_milli must be positive");
        t->gamma = (double)value / 1000.0;
    } else if (k == "big_window") {
        if (value < 0) return fail(VO_ERR_INVALID, "big_window must be >= 0");
        t->big_window = value;
    } else if (k == "big_frac_lo_milli" || k == "big_frac_hi_milli") {
        if (value <= 0 || value >= 1000) return fail(VO_ERR_INVALID, k + " must be in 1..999");
        (k == "big_frac_lo_milli" ? t->big_frac_lo : t->big_frac_hi) = (double)value / 1000.0;
    } else if (k == "outcome_flags") {
        if (value != 0 && value != 1) return fail(VO_ERR_INVALID, "outcome_flags must be 0 or 1");
        t->outcome_flags = (int)value;
    } else if (k == "xcd_regions") {
        if (value != 0 && value != 1) return fail(VO_ERR_INVALID, "xcd_regions must be 0 or 1");
        t->xcd = (int)value;
    } else if (k == "compact_min") {
        if (value < 16) return fail(VO_ERR_INVALID, "compact_min must be >= 16");
        t->compact_min = value;
    } else if (k == "compact_ratio_milli") {
        if (value != 0 && value < 1500) return fail(VO_ERR_INVALID, "compact_ratio_milli must be 0 or >= 1500");
        t->compact_ratio = (double)value / 1000.0;
    } else if (k == "pipelined") {
        if (value != 0 && value != 1) return fail(VO_ERR_INVALID, "pipelined must be 0 or 1");
        t->pipelined = (int)value;
    } else if (k == "round_timing") {
        if (value != 0 && value != 1) return fail(VO_ERR_INVALID, "round_timing must be 0 or 1");
        t->round_timing = (int)value;
    } else if (k == "small_frac_lo_milli" || k == "small_frac_hi_milli") {
        if (value < 1 || value > 999) return fail(VO_ERR_INVALID, k + " must be in 1..999");
        (k == "small_frac_lo_milli" ? t->small_frac_lo : t->small_frac_hi) = (double)value / 1000.0;
    } else if (k == "small_run") {
        if (value < 0) return fail(VO_ERR_INVALID, "small_run must be >= 0");
        t->small_run = value;
    } else if (k == "min_window") {
        if (value < 1 || value > (1 << 16)) return fail(VO_ERR_INVALID, "min_window must be in 1..65536");
        t->min_window = (int)value;
    } else if (k == "abort_early") {
        if (value != 0 && value != 1) return fail(VO_ERR_INVALID, "abort_early must be 0 or 1");
        t->abort_early = (int)value;
    } else if (k == "alloc_stripes") {
        if (value != 0 && value != 1) return fail(VO_ERR_INVALID, "alloc_stripes must be 0 or 1");
        t->alloc_stripes = (int)value;
    } else if (k == "boot_window") {
        if (value < 0 || value > 65536) return fail(VO_ERR_INVALID, "boot_window must be in 0..65536");
        t->boot_window = (int)value;
    } else if (k == "walk_budget") {
        if (value < 0 || value > (1 << 20)) return fail(VO_ERR_INVALID, "walk_budget must be in 0..2^20");
        t->walk_budget = (int)value;
    } else if (k == "shared_facets") {
        if (value != 0 && value != 1) return fail(VO_ERR_INVALID, "shared_facets must be 0 or 1");
        t->shared_facets = (int)value;
    } else if (k == "reuse") {
        if (value < -1 || value > 1) return fail(VO_ERR_INVALID, "reuse must be -1 (auto), 0 or 1");
        t->reuse = (int)value;
    } else if (k == "chunk_rounds") {
        if (value < 1 || value > 64) return fail(VO_ERR_INVALID, "chunk_rounds must be in 1..64");
        t->chunk_rounds = (int)value;
    } else if (k == "local_order") {
        if (value < 0 || value > (1 << 20)) return fail(VO_ERR_INVALID, "local_order must be in 0..2^20");
        t->local_order = (int)value;
    } else {
        return fail(VO_ERR_INVALID, "unknown option " + k);
    }
    return 0;
}

int vo_tri_vertices(const vo_tri* t, double* coords) {
    if (!t || !coords) return fail(VO_ERR_INVALID, "null argument");
    VO_HIP(hipSetDevice(t->dev));
    const int D = t->D, M = t->M, S = t->S;
    for (int v = 0; v < 2 * M; ++v) {
        const int src = v < M ? v : (D == 3 ? (v == 2 * M - 1 ? 1 : 0) : v - M);   // ghosts copy super vertices
        for (int k = 0; k < D; ++k) coords[v * D + k] = t->sv[src][k];
    }
    std::vector<double> x((size_t)t->n_pts * S);
    if (t->n_pts)
        VO_HIP(hipMemcpy(x.data(), t->X64 + (size_t)M * S, sizeof(double) * x.size(), hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < t->n_pts; ++i) {
        const int64_t in = t->input_of[i];
        for (int k = 0; k < D; ++k) coords[(2 * M + in) * D + k] = x[i * S + k];
    }
    return 0;
}

}  // extern "C"
