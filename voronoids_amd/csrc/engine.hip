// engine.hip -- several independent point sets triangulated together on one device: the
// batched and served modes of include/voronoids.h (vo_engine_*, vo_batch_triangulate,
// vo_server_*).
//
// Each set is the construction of vo_tri_create + vo_tri_insert (DelaunayTree::new and
// add_points_to_tree, src/delaunay_tree.rs:389-510 / :544-640 and :336-386) on its own points;
// the sets share the device arrays and every round's kernels:
//   vertex arena  : per set, its d+1 super vertices then its points in priority order, so
//                   vertex id = M + point index holds across sets;
//   simplex arena : per set, ranges handed out by the host (first-fit); the commit kernels
//                   take slots from the set's own cursor; a full range sends the winner back
//                   to the window and flags the set for another range;
//   hint arena    : per set, its multi-level walk-start grid.
// A round plans every resident set's window on the device (k_plan, k_fill), then runs the
// cavity / commit / exact kernels of the one-set path over all sets at once.  R rounds plus a
// completion scan are captured in one HIP graph per lane-group size and replayed; the host
// works only between replays: it admits queued sets whose footprint fits, re-runs oversized
// cavities alone, extends full ranges, and exports and frees completed sets.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "../../include/voronoids.h"
#include "host_common.hpp"
#include "kernels.hpp"
#include "kernels_group.hpp"

#pragma clang fp contract(off)

using namespace vo;
using namespace vo_host;

namespace {

// ------------------------------------------------------------------ device side
struct DoneRec {
    int slot, live, inserted, dups;
    int cursor, fail_lo, end;   // the last range's written part ends at min of the three
};

struct AdmitDesc {
    double c[3];
    double sv[4][3];
    float glo[3];
    float ginv;
    long long pt_off;      // first staged point of the set
    int slot, n, v0, root, range_end, hint_off, hlevels, hcells;
    int tet0[4];
    unsigned long long root_rec[16];   // the first simplex's record (128 B in 3D, 64 B in 2D)
};

struct ExportChunk {
    int start, len, idx;   // simplex range piece, index of its set among the exported sets
};

struct ExtendRec {
    int slot, start, len, pad;
};

constexpr int kPlanThreads = 1024;

// block-wide exclusive scan of v (1024 threads); returns the prefix, total in *tot
__device__ int block_excl_scan_1024(int v, int* s_w, int* tot) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_w[w] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int i = 0; i < kPlanThreads / 64; ++i) {
            const int t = s_w[i];
            s_w[i] = acc;
            acc += t;
        }
        s_w[kPlanThreads / 64] = acc;
    }
    __syncthreads();
    const int r = s_w[w] + x - v;
    *tot = s_w[kPlanThreads / 64];
    __syncthreads();
    return r;
}

// Window plan of one round: adapt the window ratio from the last round's winner fraction,
// advance the round key, and give every resident set max(min_window, gamma * inserted) window
// points, its pending points first; when the sum exceeds the window capacity every set's
// fresh share is scaled down.
__global__ void __launch_bounds__(kPlanThreads) k_plan(BatchCtl* ctl, SetRec* sets, int nslots, RoundState* st,
                                                       int win_cap, int win_limit, int pend_cap, const Counters* cnt) {
    __shared__ int s_w[kPlanThreads / 64 + 1];
    __shared__ int s_npend;
    __shared__ float s_gamma;
    __shared__ unsigned long long s_sum;
    round_totals(cnt, st);       // the last round's winners
    if (threadIdx.x == 0) {
        // (a count past the list's capacity has already raised ERR_PENDING: never read past it)
        const int npend = min(st->n_next_pending, pend_cap);
        const int prev = ctl->nwin;
        float g = ctl->gamma;
        if (prev > 0) {
            // as the one-set driver: about half of a window wins, more of a large one
            const float frac = (float)st->n_winners / (float)prev;
            const bool big = ctl->big_window > 0 && prev >= ctl->big_window;
            const float hi = big ? 0.8f : 0.6f, lo = big ? 0.55f : 0.35f;
            if (frac > hi) g = fminf(g * 1.5f, 4.0f);
            else if (frac < lo) g = fmaxf(g * 0.7f, 0.002f);
        }
        ctl->gamma = g;
        ctl->round += 1u;
        ctl->key_hi = (unsigned long long)(0xFFFFFFFFu - ctl->round) << 32;
        // at most win_cap pending points enter the window; the rest start the next list
        const int take = min(npend, win_cap);
        st->n_next_pending = npend - take;
        st->n_winners = 0;
        st->n_q = 0;
        st->n_fix = 0;
        st->n_live_delta = 0;
        ctl->tail = npend - take;
        s_npend = take;
        s_gamma = g;
        s_sum = 0;
    }
    __syncthreads();
    const int npend = s_npend;
    const float g = s_gamma;
    const int minw = ctl->min_window;
    nslots = min(nslots, ctl->slots);
    unsigned long long part = 0;
    for (int s = threadIdx.x; s < nslots; s += kPlanThreads) {
        SetRec& r = sets[s];
        int fresh = 0;
        if ((r.flags & (SET_ACTIVE | SET_NEED_SPACE)) == SET_ACTIVE) {
            const float want = fmaxf((float)minw, g * (float)r.inserted);
            const int w = want > 1.0e9f ? 1000000000 : (int)want;
            fresh = min(r.last - r.head, max(0, w - r.pend));
        }
        r.pend = 0;
        r.fresh = fresh;
        part += (unsigned long long)fresh;
    }
    atomicAdd(&s_sum, part);
    __syncthreads();
    const unsigned long long want_total = s_sum;
    const long long room = win_limit > npend ? (long long)win_limit - npend : 0;
    const bool scale = want_total > 0 && (long long)want_total > room;
    int carry = 0;
    for (int s0 = 0; s0 < nslots; s0 += kPlanThreads) {
        const int s = s0 + (int)threadIdx.x;
        int fresh = 0;
        if (s < nslots) {
            fresh = sets[s].fresh;
            if (scale) fresh = (int)((unsigned long long)fresh * (unsigned long long)room / want_total);
        }
        int tot;
        const int off = block_excl_scan_1024(fresh, s_w, &tot);
        if (s < nslots) {
            sets[s].fresh = fresh;
            sets[s].woff = carry + off;
        }
        carry += tot;
    }
    if (threadIdx.x == 0) {
        ctl->npend = npend;
        ctl->nwin = npend + carry;
    }
}

// window = the pending points, then each set's fresh points head .. head + fresh - 1; pending
// points beyond the window's capacity move to the head of the next pending list
__global__ void __launch_bounds__(256) k_fill(const BatchCtl* ctl, SetRec* sets, int nslots, const int* pend,
                                              int* pend_next, int* win) {
    const int npend = ctl->npend, tail = ctl->tail;
    nslots = min(nslots, ctl->slots);
    for (int i = blockIdx.x * 256 + threadIdx.x; i < npend; i += gridDim.x * 256) win[i] = pend[i];
    for (int i = blockIdx.x * 256 + threadIdx.x; i < tail; i += gridDim.x * 256) pend_next[i] = pend[npend + i];
    for (int s = blockIdx.x; s < nslots; s += gridDim.x) {
        const int fresh = sets[s].fresh;
        if (fresh == 0) continue;
        const int head = sets[s].head, off = npend + sets[s].woff;
        for (int j = threadIdx.x; j < fresh; j += 256) win[off + j] = head + j;
        __syncthreads();
        if (threadIdx.x == 0) sets[s].head = head + fresh;
    }
}

// the points a captured round's lane-group commit sent on (its outcome flags), appended to the
// round's pending list after the part k_fill carried over: one atomic per 1,024 window slots
// instead of one per commit wavefront on the list's counter
__global__ void __launch_bounds__(256) k_append_pending(Ctx c) {
    append_flagged(c.pflag, c.ctl->nwin, c.pend_next, c.pend_cap, &c.st->n_next_pending, &c.st->err, blockIdx.x,
                   gridDim.x);
}

// completed sets (every point inserted or dropped as a duplicate) leave the rounds
__global__ void k_collect(BatchCtl* ctl, SetRec* sets, int nslots, DoneRec* done) {
    nslots = min(nslots, ctl->slots);
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < nslots; s += gridDim.x * blockDim.x) {
        SetRec& r = sets[s];
        if ((r.flags & SET_ACTIVE) && r.inserted + r.dups == r.n) {
            r.flags = SET_DONE;
            const int k = atomicAdd(&ctl->n_done, 1);
            done[k] = DoneRec{s, r.live, r.inserted, r.dups, r.cursor, r.fail_lo, r.end};
        }
    }
}

__global__ void k_clear_queue(RoundState* st) { st->n_q = 0; }

// slots that may be in use, set before each replay (the planner, fill and collection scan only these)
__global__ void k_set_slots(BatchCtl* ctl, int slots) { ctl->slots = slots; }

__global__ void k_boundary_reset(BatchCtl* ctl, RoundState* st, unsigned int round) {
    ctl->n_done = 0;
    ctl->round = round;
    st->n_big = 0;
    st->n_nospace = 0;
}

// arrays touched by admission
struct Arena {
    double* X64;
    float* X32;
    int* cp;
    int* vmap;
    int* pset;
    char* rec;             // simplex records
    int* hint;
    SetRec* sets;
};


// one block per admitted set: hint grid cleared (level 0 = the first simplex), super
// vertices, first simplex and set record written.  The rest of the simplex range is not
// cleared: a record is read only after the commit wrote it, the export scans only the written
// part of each range, and reservation keys left by earlier sets are from earlier rounds, so
// larger than any current key (a never-used arena starts as all 0xFF bytes)
template <int D>
__global__ void __launch_bounds__(256) k_admit_sets(const AdmitDesc* d, Arena a) {
    constexpr int M = D + 1, S = D == 3 ? 4 : 2;
    const AdmitDesc& e = d[blockIdx.x];
    for (int i = threadIdx.x; i < e.hcells; i += 256) a.hint[e.hint_off + i] = -1;
    __syncthreads();
    if (threadIdx.x != 0) return;
    a.hint[e.hint_off] = e.root;
    for (int k = 0; k < M; ++k) {
        for (int j = 0; j < S; ++j) {
            const double x = j < D ? e.sv[k][j] : 0.0;
            a.X64[(size_t)(e.v0 + k) * S + j] = x;
            a.X32[(size_t)(e.v0 + k) * S + j] = j < D ? (float)(x - e.c[j]) : 0.0f;
        }
        a.vmap[e.v0 + k] = k;
    }
    unsigned long long* q = reinterpret_cast<unsigned long long*>(a.rec + (size_t)e.root * (D == 3 ? 128 : 64));
    for (int k = 0; k < (D == 3 ? 16 : 8); ++k) q[k] = e.root_rec[k];
    SetRec r = {};
    r.cursor = e.root + 1;
    r.end = e.range_end;
    r.head = e.v0;
    r.last = e.v0 + e.n;
    r.flags = SET_ACTIVE;
    r.hint_off = e.hint_off;
    r.hlevels = e.hlevels;
    r.glo[0] = e.glo[0]; r.glo[1] = e.glo[1]; r.glo[2] = e.glo[2];
    r.ginv = e.ginv;
    r.live = 1;
    r.n = e.n;
    r.fail_lo = 0x7FFFFFFF;
    a.sets[e.slot] = r;
}

// staged points (priority order per set) into the vertex arena
template <int D>
__global__ void __launch_bounds__(256) k_admit_points(const AdmitDesc* d, int nd, const double* pts, const int* inidx,
                                                      long long total, Arena a) {
    constexpr int M = D + 1, S = D == 3 ? 4 : 2;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < total; i += (long long)gridDim.x * 256) {
        int lo = 0, hi = nd - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (d[mid].pt_off <= i) lo = mid;
            else hi = mid - 1;
        }
        const AdmitDesc& e = d[lo];
        const int j = (int)(i - e.pt_off);
        const int v = e.v0 + M + j;
        const int p = v - M;
        for (int k = 0; k < S; ++k) {
            const double x = k < D ? pts[i * D + k] : 0.0;
            a.X64[(size_t)v * S + k] = x;
            a.X32[(size_t)v * S + k] = k < D ? (float)(x - e.c[k]) : 0.0f;
        }
        a.vmap[v] = 2 * M + inidx[i];
        a.pset[p] = e.slot;
        a.cp[p] = -1;
    }
}

// a fresh simplex range for a set whose range filled up
template <int D>
__global__ void __launch_bounds__(256) k_extend(const ExtendRec* x, char* rec, SetRec* sets) {
    const ExtendRec e = x[blockIdx.x];
    (void)rec;
    if (threadIdx.x == 0) {
        sets[e.slot].fail_lo = 0x7FFFFFFF;
        sets[e.slot].cursor = e.start;
        sets[e.slot].end = e.start + e.len;
        sets[e.slot].flags &= ~SET_NEED_SPACE;
    }
}

// live simplices of completed sets, vertex ids mapped to the reference id space, compacted
// per set (block scan + one atomicAdd per 256 simplices on the set's output cursor)
template <int M>
__global__ void __launch_bounds__(256) k_export_sets(const ExportChunk* ch, int nch, Ctx c, const int* vmap,
                                                     const long long* obase, const int* olen, int* ocur, int32_t* out) {
    __shared__ int s_w[4 + 1];
    __shared__ int s_base;
    for (int b = blockIdx.x; b < nch; b += gridDim.x) {
        const ExportChunk e = ch[b];
        for (int o = 0; o < e.len; o += 256) {
            const bool in = o + (int)threadIdx.x < e.len;
            const int4 v = in ? ld_v<M - 1>(c, e.start + o + threadIdx.x) : make_int4(-1, 0, 0, 0);
            const int live = in && v.x >= 0;
            const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
            int x = live;
            for (int k = 1; k < 64; k <<= 1) {
                const int y = __shfl_up(x, k, 64);
                if (lane >= k) x += y;
            }
            if (lane == 63) s_w[w] = x;
            __syncthreads();
            if (threadIdx.x == 0) {
                int acc = 0;
                for (int i = 0; i < 4; ++i) {
                    const int t = s_w[i];
                    s_w[i] = acc;
                    acc += t;
                }
                s_base = acc ? atomicAdd(&ocur[e.idx], acc) : 0;
            }
            __syncthreads();
            const int pos = s_base + s_w[w] + x - 1;
            if (live && pos < olen[e.idx]) {       // a count beyond the set's live count is reported by the host
                const long long row = obase[e.idx] + pos;
                out[row * M + 0] = vmap[v.x];
                out[row * M + 1] = vmap[v.y];
                out[row * M + 2] = vmap[v.z];
                if (M == 4) out[row * M + 3] = vmap[v.w];
            }
            __syncthreads();
        }
    }
}

__global__ void k_fill_u64(unsigned long long* p, long long n, unsigned long long v) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) p[i] = v;
}
__global__ void k_fill_i32(int* p, long long n, int v) {
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) p[i] = v;
}

// ------------------------------------------------------------------ host side
double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// first-fit allocator of index ranges, coalescing on release
struct RangeAlloc {
    std::map<int64_t, int64_t> free_;      // start -> length
    int64_t cap = 0, used = 0;
    void init(int64_t c) {
        free_.clear();
        cap = c;
        used = 0;
        if (c > 0) free_[0] = c;
    }
    int64_t alloc(int64_t len) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= len) {
                const int64_t st = it->first, rest = it->second - len;
                free_.erase(it);
                if (rest > 0) free_[st + len] = rest;
                used += len;
                return st;
            }
        }
        return -1;
    }
    void release(int64_t st, int64_t len) {
        if (len <= 0) return;
        used -= len;
        auto it = free_.emplace(st, len).first;
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
                it = pv;
            }
        }
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
    }
};

struct HSet {
    int64_t tag = 0;
    int64_t n = 0;
    uint64_t seed = 0;
    std::vector<double> pts;               // as submitted, until prepared
    const double* ext = nullptr;           // or the caller's buffer (vo_batch_triangulate), until prepared
    // prepared by the engine's worker threads: frame, super simplex, points in priority order
    std::atomic<bool> ready{false};
    double c[3] = {0, 0, 0}, r = 1.0;
    double sv[4][3] = {};
    int tet0[4] = {0, 1, 2, 3};
    std::vector<double> packed;            // n x dim, priority order
    std::vector<int32_t> inidx;            // input index of each packed point
    int slot = -1;
    int64_t v0 = 0, h0 = 0, hcells = 0;
    int hlevels = 0;
    // simplex ranges [start, start + len) and, once closed, how many records were written
    struct Rng { int64_t start, len, used; };
    std::vector<Rng> ranges;
    double t_submit = 0, t_admit = 0;
    unsigned round_admit = 0;
};

// host work of a set that needs no engine state: the reference's bounding sphere (its frame),
// the super simplex, and the points gathered in priority order
void prepare_set(HSet* h, int D) {
    const double* pts = h->ext ? h->ext : h->pts.data();
    set_frame(D, pts, h->n, h->c, &h->r);
    super_vertices(D, h->c, h->r * 10.0, h->sv);
    first_simplex_order(D, h->sv, h->tet0);
    // the one-set path's priority order, so that every set (duplicates included) matches it
    const std::vector<int32_t> perm = priority_order(h->n, h->seed);
    h->packed.resize((size_t)h->n * D);
    h->inidx.resize((size_t)h->n);
    for (int64_t i = 0; i < h->n; ++i) {
        const int32_t src = perm[i];
        for (int j = 0; j < D; ++j) h->packed[(size_t)i * D + j] = pts[(size_t)src * D + j];
        h->inidx[i] = src;
    }
    std::vector<double>().swap(h->pts);
    h->ext = nullptr;
    h->ready.store(true, std::memory_order_release);
}

}  // namespace

struct vo_engine {
    int dev = 0, D = 3, M = 4, S = 4;
    hipStream_t s = nullptr;
    int64_t vcap = 0, tcap = 0, hcap = 0;
    int nslots = 8192;
    int win_cap = 0, win_limit = 0, kcap = 0, fcap = 0, qcap = 0, big_cap = 0, solo_cap = 1 << 20;
    uint32_t cmask = 0;
    int ws_threads = 1024, ws_cap = 1 << 14;
    // device arrays
    Arena a{};
    int *win = nullptr, *pend[2] = {nullptr, nullptr}, *cav = nullptr, *fac = nullptr, *fnb = nullptr, *nk = nullptr, *nf = nullptr,
        *status = nullptr, *big = nullptr, *cav1 = nullptr, *fac1 = nullptr, *misc1 = nullptr;
    RoundState* st = nullptr;
    RoundState* st_host = nullptr;
    BatchCtl* ctl = nullptr;
    BatchCtl* ctl_host = nullptr;
    Counters* cnt_host = nullptr;          // device counters (kCntStripes copies) as of the last replay (pinned)
    Counters* cnt = nullptr;
    QEntry* q = nullptr;
    CEntry* cache = nullptr;
    double* ws = nullptr;
    DoneRec* done = nullptr;
    // staging (admission, export, extension)
    char* stage_h = nullptr;
    char* stage_d = nullptr;
    size_t stage_cap = 0;
    int32_t* out_d = nullptr;
    size_t out_cap = 0;                     // int32 elements
    // host bookkeeping
    RangeAlloc valloc, talloc, halloc;
    std::set<int> free_slots;          // the lowest free slot is taken first, so that slots in use stay dense
    std::set<int> used_slots;
    std::deque<std::unique_ptr<HSet>> queue;
    std::vector<std::unique_ptr<HSet>> resident;   // by slot
    int64_t n_resident = 0;
    std::deque<vo_set_result> ready;
    bool cache_stale = false;
    // graphs, one per lane-group size
    hipGraphExec_t graph[6] = {};
    int graph_grid_g[6] = {};
    int rounds_per_replay = 8;
    int group = 0;
    int64_t lanes_target = 0;
    int min_window = 8;
    int tpp_milli = 0;
    int filters = 3;
    int abort_early = 1;               // as vo_tri_config "abort_early"
    int walk_budget = 0;               // as vo_tri_config "walk_budget"
    int reuse = 0;                     // as vo_tri_config "reuse" (0 or 1 here)
    int2* fix = nullptr;               // shared facets of a round's winners (k_fix_shared)
    int outcome_flags = 1;             // as vo_tri_config "outcome_flags": lane-group commits write
    int* pflag = nullptr;              // per-slot outcomes (win_cap), appended by k_append_pending
    int fix_cap = 0;
    float gamma0 = 0.2f;
    bool launched = false;
    int last_nwin = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // host worker threads: prepare queued sets ahead of admission, copy exported results
    std::vector<std::thread> workers;
    std::mutex prep_mu;
    std::condition_variable prep_cv, ready_cv;
    std::deque<std::function<void()>> tasks;   // copies go to the front, preparations to the back
    bool prep_stop = false;
    int32_t* out_h = nullptr;                   // pinned bounce buffer of the export
    // exported results whose rows still sit in out_h: copied into their own buffers later, so
    // that the copies overlap the next replay
    std::vector<vo_set_result> copy_res;
    std::vector<int64_t> copy_off;              // first row in out_h of each
    // the export's device-to-host copies run on a stream of their own, so that the next replay
    // starts while they are in flight; copy_done follows them, and the host waits for it before
    // it reads out_h / ocur_h or the export reuses out_d
    hipStream_t cs = nullptr;
    hipEvent_t export_done = nullptr, copy_done = nullptr;
    bool copy_pending = false;
    int* ocur_h = nullptr;                      // pinned: rows the export wrote per set (checked)
    // stats
    vo_engine_stats stats{};
};

namespace {

void worker(vo_engine* e) {
    for (;;) {
        std::function<void()> task;
        {
            std::unique_lock<std::mutex> lk(e->prep_mu);
            e->prep_cv.wait(lk, [&] { return e->prep_stop || !e->tasks.empty(); });
            if (e->prep_stop) return;
            task = std::move(e->tasks.front());
            e->tasks.pop_front();
        }
        task();
    }
}

void stop_workers(vo_engine* e) {
    {
        std::lock_guard<std::mutex> lk(e->prep_mu);
        e->prep_stop = true;
        e->tasks.clear();
    }
    e->prep_cv.notify_all();
    for (auto& t : e->workers)
        if (t.joinable()) t.join();
    e->workers.clear();
}

void enqueue(vo_engine* e, std::unique_ptr<HSet> h) {
    HSet* raw = h.get();
    e->queue.push_back(std::move(h));
    e->stats.sets_submitted++;
    const int D = e->D;
    {
        std::lock_guard<std::mutex> lk(e->prep_mu);
        e->tasks.push_back([e, raw, D] {
            prepare_set(raw, D);
            { std::lock_guard<std::mutex> l2(e->prep_mu); }
            e->ready_cv.notify_all();
        });
    }
    e->prep_cv.notify_one();
}

// block until the queue's first set is prepared
void wait_ready(vo_engine* e, HSet* h) {
    if (h->ready.load(std::memory_order_acquire)) return;
    std::unique_lock<std::mutex> lk(e->prep_mu);
    e->ready_cv.wait(lk, [&] { return h->ready.load(std::memory_order_acquire); });
}

// run fn(0 .. n-1) on the workers (ahead of queued preparations) and on this thread.  The
// shared counters live in a shared_ptr that every helper holds, so a helper that starts after
// the loop is over finds no index left and touches nothing else; fn is called only for indices
// taken before the last one completes, i.e. while this call is still waiting.
void parallel_for(vo_engine* e, int n, const std::function<void(int)>& fn) {
    if (n <= 0) return;
    struct Shared {
        std::atomic<int> next{0}, done{0};
        std::mutex m;
        std::condition_variable cv;
    };
    auto sh = std::make_shared<Shared>();
    const std::function<void(int)>* f = &fn;
    auto body = [sh, f, n] {
        for (int i; (i = sh->next.fetch_add(1)) < n;) {
            (*f)(i);
            if (sh->done.fetch_add(1) + 1 == n) {
                std::lock_guard<std::mutex> lk(sh->m);
                sh->cv.notify_all();
            }
        }
    };
    const int helpers = std::min<int>((int)e->workers.size(), n - 1);
    {
        std::lock_guard<std::mutex> lk(e->prep_mu);
        for (int k = 0; k < helpers; ++k) e->tasks.push_front(body);
    }
    e->prep_cv.notify_all();
    body();
    std::unique_lock<std::mutex> lk(sh->m);
    sh->cv.wait(lk, [&] { return sh->done.load() == n; });
}

void engine_free(vo_engine* e) {
    stop_workers(e);
    if (e->cs) hipStreamSynchronize(e->cs);
    void* ptrs[] = {e->a.X64, e->a.X32, e->a.cp, e->a.vmap, e->a.pset, e->a.rec, e->a.hint, e->a.sets,
                    e->win, e->pend[0], e->pend[1], e->cav, e->fac, e->fnb, e->nk, e->nf, e->status, e->big, e->cav1,
                    e->fac1, e->misc1, e->st, e->ctl, e->cnt, e->q, e->cache, e->ws, e->done, e->stage_d, e->out_d,
                    e->fix, e->pflag};
    for (void* p : ptrs)
        if (p) hipFree(p);
    if (e->st_host) hipHostFree(e->st_host);
    if (e->ctl_host) hipHostFree(e->ctl_host);
    if (e->cnt_host) hipHostFree(e->cnt_host);
    if (e->stage_h) hipHostFree(e->stage_h);
    if (e->out_h) hipHostFree(e->out_h);
    for (auto& g : e->graph)
        if (g) hipGraphExecDestroy(g);
    if (e->ev0) hipEventDestroy(e->ev0);
    if (e->ev1) hipEventDestroy(e->ev1);
    if (e->export_done) hipEventDestroy(e->export_done);
    if (e->copy_done) hipEventDestroy(e->copy_done);
    if (e->ocur_h) hipHostFree(e->ocur_h);
    if (e->cs) hipStreamDestroy(e->cs);
    if (e->s) hipStreamDestroy(e->s);
    for (auto& r : e->ready) std::free(r.simplices);
}

Ctx engine_ctx(vo_engine* e) {
    Ctx c{};
    c.rec = e->a.rec;
    c.X64 = e->a.X64;
    c.X32 = e->a.X32;
    c.cp = e->a.cp;
    c.win = e->win;
    c.pend_next = e->pend[0];          // rounds alternate the two lists; solo runs append to list 0
    c.big = e->big;
    c.cav = e->cav;
    c.fac = e->fac;
    c.fnb = e->fnb;
    c.nk = e->nk;
    c.nf = e->nf;
    c.status = e->status;
    c.hint = e->a.hint;
    c.ginv = 1.0f;
    c.hlevels = 0;
    c.st = e->st;
    c.cnt = e->cnt;
    c.q = e->q;
    c.qcap = e->qcap;
    c.cache = e->cache;
    c.cmask = e->cmask;
    c.filters = e->filters;
    c.solo = 0;
    c.abort_early = e->abort_early;
    c.walk_budget = e->walk_budget;
    c.reuse = e->reuse;
    c.fix = e->fix;
    c.fix_cap = e->fix_cap;
    c.M = e->M;
    c.win_cap = e->win_cap;
    c.kcap = e->kcap;
    c.fcap = e->fcap;
    c.tet_cap = (int)e->tcap;
    c.sets = e->a.sets;
    c.pset = e->a.pset;
    c.ctl = e->ctl;
    c.big_cap = e->big_cap;
    c.pend_cap = e->win_cap + e->big_cap;
    return c;
}

int ensure_stage(vo_engine* e, size_t bytes) {
    if (bytes <= e->stage_cap) return 0;
    size_t cap = std::max<size_t>(bytes, e->stage_cap * 3 / 2);
    cap = (cap + 4095) & ~(size_t)4095;
    if (e->stage_h) hipHostFree(e->stage_h);
    if (e->stage_d) hipFree(e->stage_d);
    e->stage_h = nullptr;
    e->stage_d = nullptr;
    e->stage_cap = 0;
    if (hipHostMalloc((void**)&e->stage_h, cap) != hipSuccess) return fail(VO_ERR_OOM, "pinned staging buffer");
    int rc = dev_alloc(&e->stage_d, cap);
    if (rc) return rc;
    e->stage_cap = cap;
    return 0;
}

// as the one-set driver's auto_group: latency-bound small windows get more lanes per point,
// large ones four (the thread kernels only on request, "group_lanes" = 1).  3D without a
// lanes_target: 32 lanes per point up to 65,536 window points, 16 beyond.  A batch's window is
// the sum of many small sets' windows, and each set's rounds stay chains of dependent loads
// however many sets share a round, so the one-set rule (4 or 8 lanes for large windows) loses
// there: 4,096 3D sets of 64..65,536 points 1.04 -> 0.87 s; served p50/p99 unchanged, saturated
// rate +9%.  2D keeps the one-set rule (fixed 8 or 16 lanes: 0.24 -> 0.31 / 0.42 s on such sets).
int lanes_group(const vo_engine* e, int64_t nwin) {
    if (e->group) return e->group;
    if (e->D == 3 && !e->lanes_target) return nwin * 32 <= (1 << 21) ? 32 : 16;
    const int64_t target = e->lanes_target ? e->lanes_target : (e->D == 3 ? (1 << 21) : (1 << 16));
    static const int sizes[] = {32, 16, 8};
    const int gmax = e->D == 3 ? 32 : 16;
    for (int g : sizes)
        if (g <= gmax && nwin * g <= target) return g;
    return 4;
}

int gindex(int G) { return G == 1 ? 0 : G == 4 ? 1 : G == 8 ? 2 : G == 16 ? 3 : 4; }

#define VO_ENGINE_G(DIM, GL, which)                                                              \
    do {                                                                                         \
        if (which) hipLaunchKernelGGL((k_commit_g<DIM, GL>), dim3(gblocks), dim3(64), 0, e->s, c, 0, 0ull); \
        else hipLaunchKernelGGL((k_cavity_g<DIM, GL>), dim3(gblocks), dim3(64), 0, e->s, c, 0, 0ull);       \
    } while (0)

void launch_round_kernels(vo_engine* e, const Ctx& c, int G, int which) {
    const int tblocks = std::max(1, std::min(2048, (e->win_cap + 255) / 256));
    // after a lane-group commit without reuse: links across the facets two winners share
    struct FixAfter {
        vo_engine* e;
        const Ctx& c;
        int G, which, blocks;
        ~FixAfter() {
            if (!which || G == 1 || c.reuse || !c.fix) return;
            if (e->D == 3) hipLaunchKernelGGL(k_fix_shared<3>, dim3(blocks), dim3(256), 0, e->s, c);
            else hipLaunchKernelGGL(k_fix_shared<2>, dim3(blocks), dim3(256), 0, e->s, c);
        }
    } fix_after{e, c, G, which, std::max(1, std::min(1024, e->win_cap / 64 + 1))};
    const int gblocks = std::max(1, std::min(8192, (e->win_cap * G + 63) / 64));
    if (G == 1) {
        if (e->D == 3) {
            if (which) hipLaunchKernelGGL(k_commit<3>, dim3(tblocks), dim3(256), 0, e->s, c, 0, 0ull);
            else hipLaunchKernelGGL(k_cavity<3>, dim3(tblocks), dim3(256), 0, e->s, c, 0, 0ull);
        } else {
            if (which) hipLaunchKernelGGL(k_commit<2>, dim3(tblocks), dim3(256), 0, e->s, c, 0, 0ull);
            else hipLaunchKernelGGL(k_cavity<2>, dim3(tblocks), dim3(256), 0, e->s, c, 0, 0ull);
        }
    } else if (e->D == 3 && which && G == 32) {
        // as the one-set driver: a whole wavefront per point for the commit of small windows
        const int cblocks = std::max(1, std::min(8192, e->win_cap));
        hipLaunchKernelGGL((k_commit_g<3, 64>), dim3(cblocks), dim3(64), 0, e->s, c, 0, 0ull);
    } else if (e->D == 3) {
        switch (G) {
            case 4: VO_ENGINE_G(3, 4, which); break;
            case 8: VO_ENGINE_G(3, 8, which); break;
            case 16: VO_ENGINE_G(3, 16, which); break;
            default: VO_ENGINE_G(3, 32, which); break;
        }
    } else {
        switch (G) {
            case 4: VO_ENGINE_G(2, 4, which); break;
            case 8: VO_ENGINE_G(2, 8, which); break;
            default: VO_ENGINE_G(2, 16, which); break;
        }
    }
}
#undef VO_ENGINE_G

// R rounds (plan, fill, cavity, commit, exact) and a completion scan, captured once per
// lane-group size
int get_graph(vo_engine* e, int G, hipGraphExec_t* out) {
    const int gi = gindex(G);
    if (e->graph[gi]) {
        *out = e->graph[gi];
        return 0;
    }
    VO_HIP(hipStreamBeginCapture(e->s, hipStreamCaptureModeThreadLocal));
    const int exact_blocks = e->ws_threads / 64;
    // round r reads pending list r % 2 and writes list (r + 1) % 2; an even round count ends
    // every replay on list 0
    for (int r = 0; r < e->rounds_per_replay; ++r) {
        Ctx c = engine_ctx(e);
        int* src = e->pend[r & 1];
        c.pend_next = e->pend[(r + 1) & 1];
        if (G != 1 && e->outcome_flags) c.pflag = e->pflag;
        hipLaunchKernelGGL(k_plan, dim3(1), dim3(kPlanThreads), 0, e->s, e->ctl, e->a.sets, e->nslots, e->st,
                           e->win_cap, e->win_limit, e->win_cap + e->big_cap, e->cnt);
        hipLaunchKernelGGL(k_fill, dim3(2048), dim3(256), 0, e->s, e->ctl, e->a.sets, e->nslots, src, c.pend_next,
                           e->win);
        launch_round_kernels(e, c, G, 0);
        launch_round_kernels(e, c, G, 1);
        if (c.pflag)
            hipLaunchKernelGGL(k_append_pending, dim3(std::max(1, std::min(1024, (e->win_cap + 1023) / 1024))), dim3(256), 0,
                               e->s, c);
        if (e->D == 3)
            hipLaunchKernelGGL(k_exact<4>, dim3(exact_blocks), dim3(64), 0, e->s, e->q, -1, e->a.X64, e->ws, e->ws_cap,
                               e->cache, e->cmask, e->st, e->cnt, e->qcap);
        else
            hipLaunchKernelGGL(k_exact<2>, dim3(exact_blocks), dim3(64), 0, e->s, e->q, -1, e->a.X64, e->ws, e->ws_cap,
                               e->cache, e->cmask, e->st, e->cnt, e->qcap);
    }
    hipLaunchKernelGGL(k_collect, dim3(32), dim3(256), 0, e->s, e->ctl, e->a.sets, e->nslots, e->done);
    hipGraph_t g = nullptr;
    hipError_t er = hipStreamEndCapture(e->s, &g);
    if (er != hipSuccess) return fail(VO_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(er));
    hipGraphExec_t x = nullptr;
    er = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (er != hipSuccess) return fail(VO_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(er));
    e->graph[gi] = x;
    *out = x;
    return 0;
}

void drop_graphs(vo_engine* e) {
    for (auto& g : e->graph)
        if (g) {
            hipGraphExecDestroy(g);
            g = nullptr;
        }
}

int64_t first_range_len(const vo_engine* e, int64_t n) {
    const int64_t milli = e->tpp_milli ? e->tpp_milli : e->reuse ? (e->D == 3 ? 12000 : 4000) : (e->D == 3 ? 30000 : 7000);
    return n * milli / 1000 + 64;
}

// admit queued sets (first come, first served) while slots, vertices, hint cells and a first
// simplex range fit; stage them in one upload and scatter them with two kernels
int admit(vo_engine* e) {
    if (e->queue.empty()) return 0;
    std::vector<HSet*> adm;
    while (!e->queue.empty() && !e->free_slots.empty()) {
        HSet* h = e->queue.front().get();
        if (!h->ready.load(std::memory_order_acquire)) {
            if (e->n_resident > 0 || !adm.empty()) break;   // admitted at a later boundary
            wait_ready(e, h);
        }
        const int L = hint_levels(e->D, h->n);
        const int64_t hc = hint_cells(e->D, L);
        const int64_t tl = first_range_len(e, h->n);
        const int64_t v0 = e->valloc.alloc(e->M + h->n);
        if (v0 < 0) break;
        const int64_t h0 = e->halloc.alloc(hc);
        if (h0 < 0) { e->valloc.release(v0, e->M + h->n); break; }
        const int64_t t0 = e->talloc.alloc(tl);
        if (t0 < 0) { e->valloc.release(v0, e->M + h->n); e->halloc.release(h0, hc); break; }
        h->v0 = v0;
        h->h0 = h0;
        h->hcells = hc;
        h->hlevels = L;
        h->ranges.push_back({t0, tl, -1});
        h->slot = *e->free_slots.begin();
        e->free_slots.erase(e->free_slots.begin());
        e->used_slots.insert(h->slot);
        h->t_admit = now_s();
        h->round_admit = e->ctl_host->round;
        adm.push_back(h);
        e->resident[h->slot] = std::move(e->queue.front());
        e->queue.pop_front();
        e->n_resident++;
    }
    if (!e->queue.empty() && adm.empty() && e->n_resident == 0)
        return fail(VO_ERR_OOM, "a queued set does not fit the empty engine (raise max_points / max_simplices)");
    if (!e->queue.empty()) e->stats.admission_waits++;
    if (adm.empty()) return 0;
    if (e->cache_stale) {
        // exact results are keyed by vertex ids, which freed sets hand on to new ones
        hipLaunchKernelGGL(k_fill_u64, dim3(1024), dim3(256), 0, e->s, reinterpret_cast<unsigned long long*>(e->cache),
                           (long long)(sizeof(CEntry) / 8) * ((long long)e->cmask + 1), ~0ull);
        e->cache_stale = false;
    }
    const int D = e->D;
    int64_t npts = 0;
    for (HSet* h : adm) npts += h->n;
    const size_t desc_b = sizeof(AdmitDesc) * adm.size();
    const size_t pts_off = (desc_b + 255) & ~(size_t)255;
    const size_t pts_b = sizeof(double) * (size_t)npts * D;
    const size_t idx_off = (pts_off + pts_b + 255) & ~(size_t)255;
    const size_t total = idx_off + sizeof(int) * (size_t)std::max<int64_t>(npts, 1);
    int rc = ensure_stage(e, total);
    if (rc) return rc;
    AdmitDesc* desc = reinterpret_cast<AdmitDesc*>(e->stage_h);
    double* P = reinterpret_cast<double*>(e->stage_h + pts_off);
    int* I = reinterpret_cast<int*>(e->stage_h + idx_off);
    int64_t off = 0;
    for (size_t k = 0; k < adm.size(); ++k) {
        HSet* h = adm[k];
        AdmitDesc& d = desc[k];
        std::memset(&d, 0, sizeof(d));
        for (int j = 0; j < 3; ++j) d.c[j] = h->c[j];
        std::memcpy(d.sv, h->sv, sizeof(d.sv));
        for (int j = 0; j < 3; ++j) d.glo[j] = (float)(-h->r);
        d.ginv = (float)(1.0 / (2.0 * h->r));
        d.pt_off = off;
        d.slot = h->slot;
        d.n = (int)h->n;
        d.v0 = (int)h->v0;
        d.root = (int)h->ranges[0].start;
        d.range_end = (int)(h->ranges[0].start + h->ranges[0].len);
        d.hint_off = (int)h->h0;
        d.hlevels = h->hlevels;
        d.hcells = (int)h->hcells;
        for (int j = 0; j < 4; ++j) d.tet0[j] = h->tet0[j];
        root_record(D, h->tet0, (int)h->v0, h->sv, h->c, reinterpret_cast<unsigned char*>(d.root_rec));
        off += h->n;
    }
    // the prepared points into the pinned staging buffer, one set per worker task
    parallel_for(e, (int)adm.size(), [&](int k) {
        HSet* h = adm[k];
        const long long o = desc[k].pt_off;
        if (h->n) {
            std::memcpy(P + o * D, h->packed.data(), sizeof(double) * (size_t)h->n * D);
            std::memcpy(I + o, h->inidx.data(), sizeof(int32_t) * (size_t)h->n);
        }
        std::vector<double>().swap(h->packed);
        std::vector<int32_t>().swap(h->inidx);
    });
    VO_HIP(hipMemcpyAsync(e->stage_d, e->stage_h, total, hipMemcpyHostToDevice, e->s));
    const AdmitDesc* dd = reinterpret_cast<const AdmitDesc*>(e->stage_d);
    const double* Pd = reinterpret_cast<const double*>(e->stage_d + pts_off);
    const int* Id = reinterpret_cast<const int*>(e->stage_d + idx_off);
    const int nd = (int)adm.size();
    const int pblocks = (int)std::max<int64_t>(1, std::min<int64_t>(4096, (npts + 255) / 256));
    if (D == 3) {
        hipLaunchKernelGGL(k_admit_sets<3>, dim3(nd), dim3(256), 0, e->s, dd, e->a);
        if (npts) hipLaunchKernelGGL(k_admit_points<3>, dim3(pblocks), dim3(256), 0, e->s, dd, nd, Pd, Id, (long long)npts, e->a);
    } else {
        hipLaunchKernelGGL(k_admit_sets<2>, dim3(nd), dim3(256), 0, e->s, dd, e->a);
        if (npts) hipLaunchKernelGGL(k_admit_points<2>, dim3(pblocks), dim3(256), 0, e->s, dd, nd, Pd, Id, (long long)npts, e->a);
    }
    VO_HIP(hipGetLastError());
    // the staging buffer is reused by the next boundary: wait for the upload
    VO_HIP(hipStreamSynchronize(e->s));
    e->stats.sets_admitted += nd;
    e->stats.max_sets_resident = std::max<int64_t>(e->stats.max_sets_resident, e->n_resident);
    return 0;
}

int sync_ctl(vo_engine* e) {
    VO_HIP(hipMemcpyAsync(e->st_host, e->st, sizeof(RoundState), hipMemcpyDeviceToHost, e->s));
    VO_HIP(hipMemcpyAsync(e->ctl_host, e->ctl, sizeof(BatchCtl), hipMemcpyDeviceToHost, e->s));
    VO_HIP(hipStreamSynchronize(e->s));
    return 0;
}

int check_err(vo_engine* e) {
    const unsigned er = e->st_host->err;
    if (!er) return 0;
    if (er & ERR_OVERFLOW) return fail(VO_ERR_OVERFLOW, "a cavity exceeded the single-point buffers");
    if (er & ERR_PRECISION) return fail(VO_ERR_PRECISION, "exact predicate left the f64 range or workspace");
    if (er & ERR_CACHE) return fail(VO_ERR_OVERFLOW, "exact-result cache full");
    if (er & ERR_HULL) return fail(VO_ERR_INVALID, "a point lies outside its set's super simplex");
    return fail(VO_ERR_INTERNAL, "walk, chain or rotation bound exceeded (err=" + std::to_string(er) + ")");
}

// points whose cavity outgrew the round kernels' lists, re-run one at a time with the thread
// kernels over single-point lists.  They share the round state: a solo point that defers or
// finds its set's range full joins pending list 0, which the next replay's first round reads,
// and its exact predicates are resolved here so that the retry can use them.  A completion
// scan follows; the round counter moves past the solo keys.
int run_solo(vo_engine* e) {
    const int nb = std::min(e->st_host->n_big, e->big_cap);
    if (nb <= 0) return 0;
    unsigned round = e->ctl_host->round;
    // the replay's own exact pass already resolved its queue
    hipLaunchKernelGGL(k_clear_queue, dim3(1), dim3(1), 0, e->s, e->st);
    for (int i = 0; i < nb; ++i) {
        ++round;
        const unsigned long long key_hi = (unsigned long long)(0xFFFFFFFFu - round) << 32;
        Ctx c = engine_ctx(e);
        c.ctl = nullptr;
        c.win = e->big + i;
        c.cav = e->cav1;
        c.fac = e->fac1;
        c.nk = e->misc1;
        c.nf = e->misc1 + 1;
        c.status = e->misc1 + 2;
        c.win_cap = 1;
        c.kcap = e->solo_cap;
        c.fcap = e->solo_cap;
        c.solo = 1;
        if (e->D == 3) {
            hipLaunchKernelGGL(k_cavity<3>, dim3(1), dim3(256), 0, e->s, c, 1, key_hi);
            hipLaunchKernelGGL(k_commit<3>, dim3(1), dim3(256), 0, e->s, c, 1, key_hi);
        } else {
            hipLaunchKernelGGL(k_cavity<2>, dim3(1), dim3(256), 0, e->s, c, 1, key_hi);
            hipLaunchKernelGGL(k_commit<2>, dim3(1), dim3(256), 0, e->s, c, 1, key_hi);
        }
    }
    VO_HIP(hipGetLastError());
    if (e->D == 3)
        hipLaunchKernelGGL(k_exact<4>, dim3(e->ws_threads / 64), dim3(64), 0, e->s, e->q, -1, e->a.X64, e->ws, e->ws_cap,
                           e->cache, e->cmask, e->st, e->cnt, e->qcap);
    else
        hipLaunchKernelGGL(k_exact<2>, dim3(e->ws_threads / 64), dim3(64), 0, e->s, e->q, -1, e->a.X64, e->ws, e->ws_cap,
                           e->cache, e->cmask, e->st, e->cnt, e->qcap);
    hipLaunchKernelGGL(k_collect, dim3(32), dim3(256), 0, e->s, e->ctl, e->a.sets, e->nslots, e->done);
    VO_HIP(hipGetLastError());
    int rc;
    if ((rc = sync_ctl(e)) || (rc = check_err(e))) return rc;
    e->ctl_host->round = round;
    e->stats.solo_runs += nb;
    return 0;
}

// another simplex range for every set whose range filled up
int extend_ranges(vo_engine* e) {
    std::vector<SetRec> recs(e->nslots);
    VO_HIP(hipMemcpyAsync(recs.data(), e->a.sets, sizeof(SetRec) * e->nslots, hipMemcpyDeviceToHost, e->s));
    VO_HIP(hipStreamSynchronize(e->s));
    std::vector<ExtendRec> ext;
    for (int s = 0; s < e->nslots; ++s) {
        if ((recs[s].flags & (SET_ACTIVE | SET_NEED_SPACE)) != (SET_ACTIVE | SET_NEED_SPACE)) continue;
        HSet* h = e->resident[s].get();
        if (!h) continue;
        const int64_t left = std::max<int64_t>(h->n - recs[s].inserted - recs[s].dups, 1);
        int64_t len = std::max<int64_t>(first_range_len(e, left) + 4096, 16384);
        int64_t st = e->talloc.alloc(len);
        if (st < 0) {
            len = 16384;
            st = e->talloc.alloc(len);
        }
        if (st < 0) continue;                       // retried after other sets free their ranges
        // the full range is closed: its written part ends at the first failed request
        HSet::Rng& last = h->ranges.back();
        const int64_t used_end = std::min<int64_t>({(int64_t)recs[s].cursor, (int64_t)recs[s].fail_lo, last.start + last.len});
        last.used = std::max<int64_t>(0, used_end - last.start);
        h->ranges.push_back({st, len, -1});
        ext.push_back(ExtendRec{s, (int)st, (int)len, 0});
    }
    if (ext.empty()) return 0;
    int rc = ensure_stage(e, sizeof(ExtendRec) * ext.size());
    if (rc) return rc;
    std::memcpy(e->stage_h, ext.data(), sizeof(ExtendRec) * ext.size());
    VO_HIP(hipMemcpyAsync(e->stage_d, e->stage_h, sizeof(ExtendRec) * ext.size(), hipMemcpyHostToDevice, e->s));
    if (e->D == 3)
        hipLaunchKernelGGL(k_extend<3>, dim3((unsigned)ext.size()), dim3(256), 0, e->s,
                           reinterpret_cast<const ExtendRec*>(e->stage_d), e->a.rec, e->a.sets);
    else
        hipLaunchKernelGGL(k_extend<2>, dim3((unsigned)ext.size()), dim3(256), 0, e->s,
                           reinterpret_cast<const ExtendRec*>(e->stage_d), e->a.rec, e->a.sets);
    VO_HIP(hipGetLastError());
    VO_HIP(hipStreamSynchronize(e->s));
    e->stats.range_extensions += (int64_t)ext.size();
    return 0;
}

// the export's copies of the last collect have landed in out_h / ocur_h
int wait_copy(vo_engine* e) {
    if (!e->copy_pending) return 0;
    VO_HIP(hipEventSynchronize(e->copy_done));
    e->copy_pending = false;
    return 0;
}

// rows of exported sets from the pinned bounce buffer into each result's own buffer (worker
// threads); the results join the ready queue
int flush_copies(vo_engine* e) {
    int rc = wait_copy(e);
    if (rc) return rc;
    const int ns = (int)e->copy_res.size();
    if (ns == 0) return 0;
    for (int i = 0; i < ns; ++i)
        if (e->ocur_h[i] != (int)e->copy_res[i].n_simplices)
            return fail(VO_ERR_INTERNAL, "export count differs from the set's live count");
    const int M = e->M;
    const int32_t* src = e->out_h;
    std::atomic<int> oom{0};
    parallel_for(e, ns, [&](int i) {
        vo_set_result& r = e->copy_res[i];
        const size_t cnt = (size_t)r.n_simplices * M;
        int32_t* b = (int32_t*)std::malloc(sizeof(int32_t) * std::max<size_t>(1, cnt));
        if (!b) { oom.store(1); return; }
        if (cnt) std::memcpy(b, src + (size_t)e->copy_off[i] * M, sizeof(int32_t) * cnt);
        r.simplices = b;
    });
    const double t = now_s();
    for (auto& r : e->copy_res) {
        r.t_done = t;
        e->ready.push_back(r);
    }
    e->copy_res.clear();
    e->copy_off.clear();
    if (oom.load()) return fail(VO_ERR_OOM, "host result buffer");
    return 0;
}

// export and free completed sets: their live simplices go to the pinned bounce buffer (one
// device-to-host copy), their arena ranges and slot are released at once, and their result
// buffers are filled by flush_copies
int collect(vo_engine* e, int n_done) {
    if (n_done <= 0) return 0;
    int rc0 = flush_copies(e);
    if (rc0) return rc0;
    std::vector<DoneRec> dr(n_done);
    VO_HIP(hipMemcpyAsync(dr.data(), e->done, sizeof(DoneRec) * n_done, hipMemcpyDeviceToHost, e->s));
    VO_HIP(hipStreamSynchronize(e->s));
    const int M = e->M;
    // outputs in groups that fit the export buffer
    size_t k0 = 0;
    while (k0 < dr.size()) {
        std::vector<long long> obase;
        std::vector<ExportChunk> ch;
        long long rows = 0;
        size_t k1 = k0;
        while (k1 < dr.size()) {
            const long long lv = dr[k1].live;
            if (k1 > k0 && (size_t)(rows + lv) * M > e->out_cap) break;
            if ((size_t)lv * M > e->out_cap) {
                // a set larger than the buffer: grow it (device buffer and its pinned bounce)
                int rcw = wait_copy(e);
                if (rcw) return rcw;
                hipFree(e->out_d);
                hipHostFree(e->out_h);
                e->out_d = nullptr;
                e->out_h = nullptr;
                e->out_cap = 0;
                int rc = dev_alloc(&e->out_d, (size_t)lv * M);
                if (rc) return rc;
                if (hipHostMalloc((void**)&e->out_h, sizeof(int32_t) * (size_t)lv * M) != hipSuccess)
                    return fail(VO_ERR_OOM, "pinned export buffer");
                e->out_cap = (size_t)lv * M;
            }
            obase.push_back(rows);
            rows += lv;
            HSet* h = e->resident[dr[k1].slot].get();
            {
                const DoneRec& d = dr[k1];
                HSet::Rng& last = h->ranges.back();
                const int64_t used_end = std::min<int64_t>({(int64_t)d.cursor, (int64_t)d.fail_lo, last.start + last.len});
                last.used = std::max<int64_t>(0, used_end - last.start);
            }
            for (auto& rg : h->ranges)
                for (int64_t a = 0; a < rg.used; a += 4096)
                    ch.push_back(ExportChunk{(int)(rg.start + a), (int)std::min<int64_t>(4096, rg.used - a),
                                             (int)(k1 - k0)});
            ++k1;
        }
        const int ns = (int)(k1 - k0);
        const size_t ob_off = 0;
        const size_t ol_off = (sizeof(long long) * ns + 255) & ~(size_t)255;
        const size_t oc_off = ol_off + ((sizeof(int) * ns + 255) & ~(size_t)255);
        const size_t ch_off = oc_off + ((sizeof(int) * ns + 255) & ~(size_t)255);
        const size_t total = ch_off + sizeof(ExportChunk) * ch.size();
        int rc = flush_copies(e);              // out_h is about to be overwritten
        if (rc) return rc;
        rc = ensure_stage(e, total);
        if (rc) return rc;
        std::memcpy(e->stage_h + ob_off, obase.data(), sizeof(long long) * ns);
        for (int i = 0; i < ns; ++i) reinterpret_cast<int*>(e->stage_h + ol_off)[i] = dr[k0 + i].live;
        std::memset(e->stage_h + oc_off, 0, sizeof(int) * ns);
        std::memcpy(e->stage_h + ch_off, ch.data(), sizeof(ExportChunk) * ch.size());
        VO_HIP(hipMemcpyAsync(e->stage_d, e->stage_h, total, hipMemcpyHostToDevice, e->s));
        const int blocks = (int)std::min<size_t>(ch.size(), 4096);
        if (M == 4)
            hipLaunchKernelGGL(k_export_sets<4>, dim3(blocks), dim3(256), 0, e->s,
                               reinterpret_cast<const ExportChunk*>(e->stage_d + ch_off), (int)ch.size(), engine_ctx(e),
                               e->a.vmap, reinterpret_cast<const long long*>(e->stage_d + ob_off),
                               reinterpret_cast<const int*>(e->stage_d + ol_off),
                               reinterpret_cast<int*>(e->stage_d + oc_off), e->out_d);
        else
            hipLaunchKernelGGL(k_export_sets<3>, dim3(blocks), dim3(256), 0, e->s,
                               reinterpret_cast<const ExportChunk*>(e->stage_d + ch_off), (int)ch.size(), engine_ctx(e),
                               e->a.vmap, reinterpret_cast<const long long*>(e->stage_d + ob_off),
                               reinterpret_cast<const int*>(e->stage_d + ol_off),
                               reinterpret_cast<int*>(e->stage_d + oc_off), e->out_d);
        VO_HIP(hipGetLastError());
        // the row counts (from the staging buffer, which the next admission overwrites: on this
        // stream), then on the copy stream one device-to-host copy of the rows into the pinned
        // bounce buffer (every set's rows go to its own result buffer later, on the worker
        // threads).  The next replay need not wait for it: the export has read the finished
        // sets' records by then, their ranges are handed on only to later work of this stream,
        // and out_d is rewritten only after the host has waited for copy_done.
        VO_HIP(hipMemcpyAsync(e->ocur_h, e->stage_d + oc_off, sizeof(int) * ns, hipMemcpyDeviceToHost, e->s));
        VO_HIP(hipEventRecord(e->export_done, e->s));
        VO_HIP(hipStreamWaitEvent(e->cs, e->export_done, 0));
        if (rows > 0)
            VO_HIP(hipMemcpyAsync(e->out_h, e->out_d, sizeof(int32_t) * (size_t)rows * M, hipMemcpyDeviceToHost, e->cs));
        VO_HIP(hipEventRecord(e->copy_done, e->cs));
        e->copy_pending = true;
        for (int i = 0; i < ns; ++i) {
            const DoneRec& d = dr[k0 + i];
            std::unique_ptr<HSet> h = std::move(e->resident[d.slot]);
            vo_set_result r{};
            r.tag = h->tag;
            r.dim = e->D;
            r.status = d.dups > 0 ? VO_ERR_DUPLICATE : VO_OK;
            r.n_points = h->n;
            r.points_duplicate = d.dups;
            r.n_simplices = d.live;
            r.simplices = nullptr;
            r.t_submit = h->t_submit;
            r.t_admit = h->t_admit;
            r.rounds = (int64_t)(e->ctl_host->round - h->round_admit);
            e->copy_res.push_back(r);
            e->copy_off.push_back(obase[i]);
            e->valloc.release(h->v0, e->M + h->n);
            e->halloc.release(h->h0, h->hcells);
            for (auto& rg : h->ranges) e->talloc.release(rg.start, rg.len);
            e->free_slots.insert(d.slot);
            e->used_slots.erase(d.slot);
            e->n_resident--;
            e->stats.sets_done++;
        }
        e->cache_stale = true;
        k0 = k1;
    }
    return 0;
}

}  // namespace

extern "C" {

double vo_clock(void) { return now_s(); }

int vo_engine_create(int hip_device, int dim, int64_t max_points, int64_t max_simplices, vo_engine** out) {
    if (!out) return fail(VO_ERR_INVALID, "out is null");
    *out = nullptr;
    if (dim != 2 && dim != 3) return fail(VO_ERR_INVALID, "dim must be 2 or 3");
    if (max_points < 16 || max_points > ((int64_t)1 << 30)) return fail(VO_ERR_INVALID, "max_points out of range");
    if (max_simplices == 0) max_simplices = std::min<int64_t>(max_points * (dim == 3 ? 30 : 8), ((int64_t)1 << 29) - 1);
    if (max_simplices < 1024 || max_simplices >= ((int64_t)1 << 29))
        return fail(VO_ERR_INVALID, "max_simplices must be in [1024, 2^29)");
    int ndev = 0;
    VO_HIP(hipGetDeviceCount(&ndev));
    if (hip_device < 0 || hip_device >= ndev) return fail(VO_ERR_INVALID, "no such HIP device");
    VO_HIP(hipSetDevice(hip_device));
    std::unique_ptr<vo_engine> e(new vo_engine());
    vo_engine* E = e.get();
    E->dev = hip_device;
    E->D = dim;
    E->M = dim + 1;
    E->S = dim == 3 ? 4 : 2;
    E->vcap = max_points;
    E->tcap = max_simplices;
    E->hcap = std::max<int64_t>(max_points, 1 << 16);   // a set's grid holds at most 0.82 cells per point
    E->win_cap = dim == 3 ? (1 << 20) : (1 << 21);
    E->win_limit = E->win_cap;
    E->kcap = dim == 3 ? 256 : 96;
    E->fcap = dim == 3 ? 384 : 128;
    E->qcap = 1 << 20;
    E->big_cap = 1 << 16;
    E->cmask = (1u << 22) - 1;
    E->gamma0 = dim == 3 ? 0.2f : 0.6f;
    E->resident.resize(E->nslots);
    for (int s = 0; s < E->nslots; ++s) E->free_slots.insert(s);
    auto bail = [&](int rc) {
        engine_free(E);
        return rc;
    };
    if (hipStreamCreateWithFlags(&E->s, hipStreamNonBlocking) != hipSuccess) return bail(fail(VO_ERR_HIP, "stream"));
    if (hipStreamCreateWithFlags(&E->cs, hipStreamNonBlocking) != hipSuccess) return bail(fail(VO_ERR_HIP, "stream"));
    if (hipEventCreateWithFlags(&E->export_done, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&E->copy_done, hipEventDisableTiming) != hipSuccess)
        return bail(fail(VO_ERR_HIP, "event"));
    if (hipHostMalloc((void**)&E->ocur_h, sizeof(int) * (size_t)E->nslots) != hipSuccess)
        return bail(fail(VO_ERR_OOM, "pinned export counts"));
    if (hipEventCreate(&E->ev0) != hipSuccess || hipEventCreate(&E->ev1) != hipSuccess)
        return bail(fail(VO_ERR_HIP, "event"));
    const int S = E->S;
    int rc;
    if ((rc = dev_alloc(&E->a.X64, (size_t)E->vcap * S)) || (rc = dev_alloc(&E->a.X32, (size_t)E->vcap * S)) ||
        (rc = dev_alloc(&E->a.cp, E->vcap)) || (rc = dev_alloc(&E->a.vmap, E->vcap)) ||
        (rc = dev_alloc(&E->a.pset, E->vcap)) || (rc = dev_alloc(&E->a.rec, (size_t)E->tcap * record_bytes(dim))) ||
        (rc = dev_alloc(&E->a.hint, E->hcap)) || (rc = dev_alloc(&E->a.sets, E->nslots)) ||
        (rc = dev_alloc(&E->win, E->win_cap)) || (rc = dev_alloc(&E->pend[0], E->win_cap + E->big_cap)) ||
        (rc = dev_alloc(&E->pend[1], E->win_cap + E->big_cap)) ||
        (rc = dev_alloc(&E->cav, (size_t)E->win_cap * E->kcap)) || (rc = dev_alloc(&E->fac, (size_t)E->win_cap * E->fcap)) ||
        (rc = dev_alloc(&E->fnb, (size_t)E->win_cap * E->fcap)) ||
        (rc = dev_alloc(&E->nk, E->win_cap)) || (rc = dev_alloc(&E->nf, E->win_cap)) ||
        (rc = dev_alloc(&E->status, E->win_cap)) || (rc = dev_alloc(&E->big, E->big_cap)) ||
        (rc = dev_alloc(&E->cav1, E->solo_cap)) || (rc = dev_alloc(&E->fac1, E->solo_cap)) ||
        (rc = dev_alloc(&E->misc1, 4)) || (rc = dev_alloc(&E->st, 1)) || (rc = dev_alloc(&E->ctl, 1)) ||
        (rc = dev_alloc(&E->cnt, kCntStripes)) || (rc = dev_alloc(&E->q, E->qcap)) ||
        (rc = dev_alloc(&E->cache, (size_t)E->cmask + 1)) || (rc = dev_alloc(&E->ws, (size_t)E->ws_threads * E->ws_cap)) ||
        (rc = dev_alloc(&E->done, E->nslots)) || (rc = dev_alloc(&E->fix, (size_t)2 * E->win_cap)) ||
        (rc = dev_alloc(&E->pflag, E->win_cap)))
        return bail(rc);
    E->fix_cap = 2 * E->win_cap;
    E->out_cap = (size_t)1 << 26;
    if ((rc = dev_alloc(&E->out_d, E->out_cap))) return bail(rc);
    if (hipHostMalloc((void**)&E->out_h, sizeof(int32_t) * E->out_cap) != hipSuccess)
        return bail(fail(VO_ERR_OOM, "pinned export buffer"));
    if (hipHostMalloc((void**)&E->st_host, sizeof(RoundState)) != hipSuccess ||
        hipHostMalloc((void**)&E->ctl_host, sizeof(BatchCtl)) != hipSuccess ||
        hipHostMalloc((void**)&E->cnt_host, sizeof(Counters) * kCntStripes) != hipSuccess)
        return bail(fail(VO_ERR_OOM, "pinned"));
    // reservations start at the largest key; no set record is active
    hipLaunchKernelGGL(k_fill_u64, dim3(4096), dim3(256), 0, E->s, reinterpret_cast<unsigned long long*>(E->a.rec),
                       (long long)E->tcap * record_bytes(dim) / 8, ~0ull);
    hipLaunchKernelGGL(k_fill_u64, dim3(1024), dim3(256), 0, E->s, reinterpret_cast<unsigned long long*>(E->cache),
                       (long long)(sizeof(CEntry) / 8) * ((long long)E->cmask + 1), ~0ull);
    hipLaunchKernelGGL(k_fill_i32, dim3(64), dim3(256), 0, E->s, reinterpret_cast<int*>(E->a.sets),
                       (long long)(sizeof(SetRec) / 4) * E->nslots, 0);
    RoundState st0 = {};
    BatchCtl c0 = {};
    c0.gamma = E->gamma0;
    c0.min_window = E->min_window;
    c0.big_window = dim == 3 ? 16384 : 65536;
    *E->st_host = st0;
    *E->ctl_host = c0;
    if (hipMemcpyAsync(E->st, E->st_host, sizeof(RoundState), hipMemcpyHostToDevice, E->s) != hipSuccess ||
        hipMemcpyAsync(E->ctl, E->ctl_host, sizeof(BatchCtl), hipMemcpyHostToDevice, E->s) != hipSuccess ||
        hipMemsetAsync(E->cnt, 0, sizeof(Counters) * kCntStripes, E->s) != hipSuccess || hipStreamSynchronize(E->s) != hipSuccess)
        return bail(fail(VO_ERR_HIP, "engine init"));
    if (hipGetLastError() != hipSuccess) return bail(fail(VO_ERR_HIP, "engine init kernels"));
    {
        const unsigned hw = std::thread::hardware_concurrency();
        const int nw = (int)std::max(1u, std::min(8u, hw ? hw / 2 : 1u));
        for (int i = 0; i < nw; ++i) E->workers.emplace_back(worker, E);
    }
    E->valloc.init(E->vcap);
    E->talloc.init(E->tcap);
    E->halloc.init(E->hcap);
    E->stats.vertex_capacity = E->vcap;
    E->stats.simplex_capacity = E->tcap;
    *out = e.release();
    return 0;
}

void vo_engine_destroy(vo_engine* e) {
    if (!e) return;
    hipSetDevice(e->dev);
    if (e->s) hipStreamSynchronize(e->s);
    engine_free(e);
    delete e;
}

int vo_engine_config(vo_engine* e, const char* key, int64_t value) {
    if (!e || !key) return fail(VO_ERR_INVALID, "null argument");
    const std::string k(key);
    if (k == "rounds_per_replay") {
        if (value < 2 || value > 64 || (value & 1)) return fail(VO_ERR_INVALID, "rounds_per_replay must be even, 2..64");
        e->rounds_per_replay = (int)value;
        drop_graphs(e);
    } else if (k == "lanes_target") {
        if (value < 0) return fail(VO_ERR_INVALID, "lanes_target must be >= 0");
        e->lanes_target = value;
    } else if (k == "group_lanes") {
        const int gmax = e->D == 3 ? 32 : 16;
        if (value != 0 && value != 1 && value != 4 && value != 8 && value != 16 && !(value == 32 && gmax == 32))
            return fail(VO_ERR_INVALID, "group_lanes must be 0, 1, 4, 8, 16 (or 32 in 3D)");
        e->group = (int)value;
    } else if (k == "min_window") {
        if (value < 1 || value > (1 << 20)) return fail(VO_ERR_INVALID, "min_window must be in 1..2^20");
        e->min_window = (int)value;
        e->ctl_host->min_window = (int)value;
        VO_HIP(hipSetDevice(e->dev));
        VO_HIP(hipMemcpy(&e->ctl->min_window, &e->ctl_host->min_window, sizeof(int), hipMemcpyHostToDevice));
    } else if (k == "big_window") {
        if (value < 0 || value > (1 << 30)) return fail(VO_ERR_INVALID, "big_window must be in 0..2^30");
        e->ctl_host->big_window = (int)value;
        VO_HIP(hipSetDevice(e->dev));
        VO_HIP(hipMemcpy(&e->ctl->big_window, &e->ctl_host->big_window, sizeof(int), hipMemcpyHostToDevice));
    } else if (k == "window_cap") {
        if (value < 64 || value > e->win_cap) return fail(VO_ERR_INVALID, "window_cap must be in 64..allocated window");
        e->win_limit = (int)value;
        drop_graphs(e);
    } else if (k == "filters") {
        if (value < 0 || value > 3) return fail(VO_ERR_INVALID, "filters must be a mask of 1 (fp32) and 2 (fp64)");
        e->filters = (int)value;
        drop_graphs(e);
    } else if (k == "abort_early") {
        if (value != 0 && value != 1) return fail(VO_ERR_INVALID, "abort_early must be 0 or 1");
        e->abort_early = (int)value;
        drop_graphs(e);
    } else if (k == "outcome_flags") {
        if (value != 0 && value != 1) return fail(VO_ERR_INVALID, "outcome_flags must be 0 or 1");
        e->outcome_flags = (int)value;
        drop_graphs(e);
    } else if (k == "walk_budget") {
        if (value < 0 || value > (1 << 20)) return fail(VO_ERR_INVALID, "walk_budget must be in 0..2^20");
        e->walk_budget = (int)value;
        drop_graphs(e);
    } else if (k == "reuse") {
        if (value != 0 && value != 1) return fail(VO_ERR_INVALID, "reuse must be 0 or 1");
        e->reuse = (int)value;
        for (auto& g : e->graph)              // the captured rounds depend on it
            if (g) {
                hipGraphExecDestroy(g);
                g = nullptr;
            }
        drop_graphs(e);
    } else if (k == "tets_per_point_milli") {
        if (value < 1000 || value > 1000000) return fail(VO_ERR_INVALID, "tets_per_point_milli must be in 1000..1e6");
        e->tpp_milli = (int)value;
    } else if (k == "window_gamma_milli") {
        if (value <= 0) return fail(VO_ERR_INVALID, "window_gamma_milli must be positive");
        e->ctl_host->gamma = (float)value / 1000.0f;
        VO_HIP(hipSetDevice(e->dev));
        VO_HIP(hipMemcpy(&e->ctl->gamma, &e->ctl_host->gamma, sizeof(float), hipMemcpyHostToDevice));
    } else {
        return fail(VO_ERR_INVALID, "unknown option " + k);
    }
    return 0;
}

int vo_engine_submit(vo_engine* e, const double* pts, int64_t n, uint64_t seed, int64_t tag) {
    if (!e) return fail(VO_ERR_INVALID, "null handle");
    if (n < 0 || (n > 0 && !pts)) return fail(VO_ERR_INVALID, "bad point buffer");
    if (n + e->M > e->vcap) return fail(VO_ERR_INVALID, "set larger than the engine's vertex arena");
    for (int64_t i = 0; i < n * e->D; ++i)
        if (!std::isfinite(pts[i])) return fail(VO_ERR_INVALID, "non-finite coordinate");
    std::unique_ptr<HSet> h(new HSet());
    h->tag = tag;
    h->n = n;
    h->seed = seed;
    h->pts.assign(pts, pts + n * e->D);
    h->t_submit = now_s();
    enqueue(e, std::move(h));
    return 0;
}

int vo_engine_launch(vo_engine* e) {
    if (!e) return fail(VO_ERR_INVALID, "null handle");
    if (e->launched) return fail(VO_ERR_INVALID, "a replay is already in flight");
    VO_HIP(hipSetDevice(e->dev));
    const double t0 = now_s();
    int rc = admit(e);
    const double ms = (now_s() - t0) * 1e3;
    e->stats.ms_boundary += ms;
    e->stats.ms_admit += ms;
    if (rc) return rc;
    if (e->n_resident == 0) return flush_copies(e);
    const int64_t pred = std::max<int64_t>((int64_t)e->last_nwin * 3 / 2, e->n_resident * (int64_t)e->min_window);
    const int G = lanes_group(e, std::min<int64_t>(pred, e->win_limit));
    hipGraphExec_t g;
    if ((rc = get_graph(e, G, &g))) return rc;
    hipLaunchKernelGGL(k_set_slots, dim3(1), dim3(1), 0, e->s, e->ctl, e->used_slots.empty() ? 0 : *e->used_slots.rbegin() + 1);
    VO_HIP(hipGetLastError());
    VO_HIP(hipEventRecord(e->ev0, e->s));
    VO_HIP(hipGraphLaunch(g, e->s));
    VO_HIP(hipEventRecord(e->ev1, e->s));
    VO_HIP(hipMemcpyAsync(e->st_host, e->st, sizeof(RoundState), hipMemcpyDeviceToHost, e->s));
    VO_HIP(hipMemcpyAsync(e->ctl_host, e->ctl, sizeof(BatchCtl), hipMemcpyDeviceToHost, e->s));
    VO_HIP(hipMemcpyAsync(e->cnt_host, e->cnt, sizeof(Counters) * kCntStripes, hipMemcpyDeviceToHost, e->s));
    e->launched = true;
    // the previous boundary's result copies overlap this replay
    const double tc = now_s();
    rc = flush_copies(e);
    e->stats.ms_collect += (now_s() - tc) * 1e3;
    return rc;
}

int vo_engine_finish(vo_engine* e, int64_t* n_left) {
    if (!e) return fail(VO_ERR_INVALID, "null handle");
    VO_HIP(hipSetDevice(e->dev));
    int rc;
    if (e->launched) {
        e->launched = false;
        VO_HIP(hipStreamSynchronize(e->s));
        const double t0 = now_s();
        float ms = 0;
        hipEventElapsedTime(&ms, e->ev0, e->ev1);
        e->stats.ms_replays += ms;
        e->stats.replays++;
        e->stats.rounds += e->rounds_per_replay;
        e->last_nwin = e->ctl_host->nwin;
        // points inserted so far, all sets (device count): the progress measure of a replay
        e->stats.points_inserted = (int64_t)sum_counters(e->cnt_host).winners;
        if ((rc = check_err(e))) return rc;
        if (e->st_host->n_big > 0 && (rc = run_solo(e))) return rc;
        if (e->st_host->n_nospace > 0 && (rc = extend_ranges(e))) return rc;
        const double tc = now_s();
        if ((rc = collect(e, e->ctl_host->n_done))) return rc;
        e->stats.ms_collect += (now_s() - tc) * 1e3;
        hipLaunchKernelGGL(k_boundary_reset, dim3(1), dim3(1), 0, e->s, e->ctl, e->st, e->ctl_host->round);
        VO_HIP(hipGetLastError());
        e->ctl_host->n_done = 0;
        e->st_host->n_big = 0;
        e->st_host->n_nospace = 0;
        e->stats.ms_boundary += (now_s() - t0) * 1e3;
    }
    if (n_left) *n_left = (int64_t)e->queue.size() + e->n_resident;
    return 0;
}

int vo_engine_step(vo_engine* e, int64_t* n_left) {
    int rc = vo_engine_launch(e);
    if (rc) return rc;
    return vo_engine_finish(e, n_left);
}

int vo_engine_results(vo_engine* e, vo_set_result* out, int64_t max, int64_t* n_out) {
    if (!e || !n_out || (max > 0 && !out)) return fail(VO_ERR_INVALID, "null argument");
    if (!e->launched) {               // with a replay in flight, its launch already flushed
        int rc = flush_copies(e);
        if (rc) return rc;
    }
    int64_t k = 0;
    while (k < max && !e->ready.empty()) {
        out[k++] = e->ready.front();
        e->ready.pop_front();
    }
    *n_out = k;
    return 0;
}

int vo_engine_stats_get(const vo_engine* e, vo_engine_stats* out) {
    if (!e || !out) return fail(VO_ERR_INVALID, "null argument");
    *out = e->stats;
    out->sets_resident = e->n_resident;
    out->sets_queued = (int64_t)e->queue.size();
    out->simplex_in_use = e->talloc.used;
    return 0;
}

int vo_batch_triangulate(vo_engine* e, const double* pts, const int64_t* set_offsets, int64_t n_sets, uint64_t seed,
                         vo_set_result* results) {
    if (!e || !set_offsets || (n_sets > 0 && (!pts || !results))) return fail(VO_ERR_INVALID, "null argument");
    if (n_sets < 0) return fail(VO_ERR_INVALID, "n_sets < 0");
    std::memset(results, 0, sizeof(vo_set_result) * (size_t)n_sets);
    VO_HIP(hipSetDevice(e->dev));
    for (int64_t k = 0; k < n_sets; ++k) {
        const int64_t a = set_offsets[k], b = set_offsets[k + 1];
        if (a < 0 || b < a) return fail(VO_ERR_INVALID, "set offsets must be non-decreasing from 0");
        if (b - a + e->M > e->vcap) return fail(VO_ERR_INVALID, "set larger than the engine's vertex arena");
    }
    {
        std::atomic<int> bad{0};
        const int D = e->D;
        parallel_for(e, (int)n_sets, [&](int k) {
            for (int64_t i = set_offsets[k] * D; i < set_offsets[k + 1] * D; ++i)
                if (!std::isfinite(pts[i])) { bad.store(1); return; }
        });
        if (bad.load()) return fail(VO_ERR_INVALID, "non-finite coordinate");
    }
    // the sets read the caller's buffer until prepared: no copy; every queued set of this call
    // is waited for before an early return
    std::vector<HSet*> mine;
    mine.reserve((size_t)n_sets);
    const double t_sub = now_s();
    for (int64_t k = 0; k < n_sets; ++k) {
        std::unique_ptr<HSet> h(new HSet());
        h->tag = k;
        h->n = set_offsets[k + 1] - set_offsets[k];
        h->seed = seed + (uint64_t)k;
        h->ext = pts + set_offsets[k] * e->D;
        h->t_submit = t_sub;
        mine.push_back(h.get());
        enqueue(e, std::move(h));
    }
    auto settle = [&](int rc) {
        for (auto& q : e->queue)
            for (HSet* m : mine)
                if (q.get() == m) wait_ready(e, m);
        return rc;
    };
    std::vector<char> got((size_t)n_sets, 0);
    int64_t left = 1, have = 0;
    std::vector<vo_set_result> buf(256);
    int64_t last_progress = e->stats.points_inserted, stall = 0;
    while (left > 0 || have < n_sets) {
        // launch (which starts the previous boundary's result copies beside the replay),
        // collect what is ready, then finish the replay
        int rc = vo_engine_launch(e);
        if (rc) return settle(rc);
        int64_t n = 0;
        auto harvest = [&] {
            do {
                vo_engine_results(e, buf.data(), (int64_t)buf.size(), &n);
                for (int64_t i = 0; i < n; ++i) {
                    const int64_t k = buf[i].tag;
                    if (k >= 0 && k < n_sets && !got[k]) {
                        results[k] = buf[i];
                        got[k] = 1;
                        ++have;
                    } else {
                        std::free(buf[i].simplices);
                    }
                }
            } while (n == (int64_t)buf.size());
        };
        harvest();
        rc = vo_engine_finish(e, &left);
        if (rc) return settle(rc);
        if (left == 0) harvest();
        // guard: a set waiting for space that nothing can free
        vo_engine_stats sst;
        vo_engine_stats_get(e, &sst);
        const int64_t prog = sst.points_inserted + sst.sets_done + sst.range_extensions + sst.sets_admitted;
        stall = prog == last_progress ? stall + 1 : 0;
        last_progress = prog;
        if (left == 0 && have < n_sets) return settle(fail(VO_ERR_INTERNAL, "engine drained with results missing"));
        if (stall > 64) return settle(fail(VO_ERR_OOM, "no progress for 64 replays: simplex arena exhausted"));
    }
    return 0;
}

int vo_result_canonical(const vo_set_result* r, int32_t** T, int64_t* n_T, int32_t** E, int64_t* n_E) {
    if (!r || !T || !n_T || !E || !n_E) return fail(VO_ERR_INVALID, "null argument");
    return canonical_from_simplices(r->dim + 1, r->simplices, r->n_simplices, T, n_T, E, n_E);
}

}  // extern "C"

// ====================================================================== served mode
// A driver thread owns a 2D and a 3D engine.  Requests are queued from any thread; each pass
// of the driver moves new requests into their engine's queue (keeping the arrival time),
// launches a replay on every engine with work (two streams, so the 2D and 3D rounds overlap
// on the device), then finishes both: admission by footprint happens at the next launch,
// completed sets are exported and their results queued for the caller.
namespace {

struct ServerReq {
    int dim;
    std::vector<double> pts;
    uint64_t seed;
    int64_t tag;
    double t_submit;
};

}  // namespace

struct vo_server {
    int dev = 0;
    vo_engine* eng[2] = {nullptr, nullptr};
    std::thread driver;
    std::mutex mu;
    std::condition_variable cv_in, cv_out;
    std::deque<ServerReq> inbox;
    std::deque<vo_set_result> outbox;
    bool stop = false;
    int err = 0;
    std::string err_msg;
};

namespace {

void server_loop(vo_server* s) {
    hipSetDevice(s->dev);
    std::vector<vo_set_result> buf(256);
    for (;;) {
        std::deque<ServerReq> take;
        bool busy = false;
        for (vo_engine* e : s->eng) busy |= !e->queue.empty() || e->n_resident > 0;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            if (!busy) s->cv_in.wait(lk, [&] { return s->stop || !s->inbox.empty(); });
            if (s->stop && s->inbox.empty() && !busy) return;
            take.swap(s->inbox);
        }
        int rc = 0;
        for (ServerReq& r : take) {
            vo_engine* e = s->eng[r.dim == 3 ? 1 : 0];
            std::unique_ptr<HSet> h(new HSet());
            h->tag = r.tag;
            h->n = (int64_t)r.pts.size() / r.dim;
            h->seed = r.seed;
            h->pts.swap(r.pts);
            h->t_submit = r.t_submit;
            enqueue(e, std::move(h));
        }
        for (vo_engine* e : s->eng)
            if (!rc && (!e->queue.empty() || e->n_resident > 0)) rc = vo_engine_launch(e);
        // results of the previous pass: their copies ran beside the replays just launched
        std::vector<vo_set_result> got;
        auto harvest = [&] {
            for (vo_engine* e : s->eng) {
                int64_t n = 0;
                do {
                    if (vo_engine_results(e, buf.data(), (int64_t)buf.size(), &n) && !rc) rc = VO_ERR_OOM;
                    got.insert(got.end(), buf.begin(), buf.begin() + n);
                } while (n == (int64_t)buf.size());
            }
        };
        harvest();
        for (vo_engine* e : s->eng)
            if (!rc) rc = vo_engine_finish(e, nullptr);
        bool idle = true;
        for (vo_engine* e : s->eng) idle &= e->queue.empty() && e->n_resident == 0;
        if (idle || rc) harvest();          // nothing will launch next: copy now
        {
            std::lock_guard<std::mutex> lk(s->mu);
            for (auto& r : got) s->outbox.push_back(r);
            if (rc) {
                s->err = rc;
                s->err_msg = g_err;
                s->stop = true;
            }
        }
        s->cv_out.notify_all();
        if (rc) return;
    }
}

}  // namespace

extern "C" {

int vo_server_create(int hip_device, int64_t max_points_2d, int64_t max_points_3d, vo_server** out) {
    if (!out) return fail(VO_ERR_INVALID, "out is null");
    *out = nullptr;
    std::unique_ptr<vo_server> s(new vo_server());
    s->dev = hip_device;
    int rc = vo_engine_create(hip_device, 2, max_points_2d, 0, &s->eng[0]);
    if (rc) return rc;
    rc = vo_engine_create(hip_device, 3, max_points_3d, 0, &s->eng[1]);
    if (rc) {
        vo_engine_destroy(s->eng[0]);
        return rc;
    }
    vo_server* raw = s.get();
    raw->driver = std::thread(server_loop, raw);
    *out = s.release();
    return 0;
}

int vo_server_submit(vo_server* s, int dim, const double* pts, int64_t n, uint64_t seed, int64_t tag) {
    if (!s) return fail(VO_ERR_INVALID, "null handle");
    if (dim != 2 && dim != 3) return fail(VO_ERR_INVALID, "dim must be 2 or 3");
    if (n < 0 || (n > 0 && !pts)) return fail(VO_ERR_INVALID, "bad point buffer");
    if (n + dim + 1 > s->eng[dim == 3 ? 1 : 0]->vcap) return fail(VO_ERR_INVALID, "request larger than the engine's vertex arena");
    for (int64_t i = 0; i < n * dim; ++i)
        if (!std::isfinite(pts[i])) return fail(VO_ERR_INVALID, "non-finite coordinate");
    ServerReq r;
    r.dim = dim;
    r.pts.assign(pts, pts + n * dim);
    r.seed = seed;
    r.tag = tag;
    r.t_submit = now_s();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->err) return fail(s->err, "server stopped: " + s->err_msg);
        if (s->stop) return fail(VO_ERR_INVALID, "server stopping");
        s->inbox.push_back(std::move(r));
    }
    s->cv_in.notify_one();
    return 0;
}

int vo_server_results(vo_server* s, vo_set_result* out, int64_t max, int64_t* n_out, int timeout_ms) {
    if (!s || !n_out || (max > 0 && !out)) return fail(VO_ERR_INVALID, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->outbox.empty() && timeout_ms > 0)
        s->cv_out.wait_for(lk, std::chrono::milliseconds(timeout_ms), [&] { return !s->outbox.empty() || s->err; });
    int64_t k = 0;
    while (k < max && !s->outbox.empty()) {
        out[k++] = s->outbox.front();
        s->outbox.pop_front();
    }
    *n_out = k;
    if (k == 0 && s->err) return fail(s->err, "server stopped: " + s->err_msg);
    return 0;
}

int vo_server_stats_get(const vo_server* s, int dim, vo_engine_stats* out) {
    if (!s || !out || (dim != 2 && dim != 3)) return fail(VO_ERR_INVALID, "bad argument");
    // a snapshot read while the driver runs: counters may be mid-update
    return vo_engine_stats_get(s->eng[dim == 3 ? 1 : 0], out);
}

void vo_server_destroy(vo_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_in.notify_all();
    if (s->driver.joinable()) s->driver.join();
    for (auto& r : s->outbox) std::free(r.simplices);
    for (vo_engine* e : s->eng) vo_engine_destroy(e);
    delete s;
}

}  // extern "C"
