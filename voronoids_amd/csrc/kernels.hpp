// kernels.hpp -- the per-round device kernels of the parallel Bowyer-Watson construction.
//
// Reference path replaced (src/delaunay_tree.rs): locate/find_all_neighbors :33-75 become the
// walk + cavity BFS of k_cavity; get_new_simplices :77-123, pair_simplices :674-695 and
// insert_point / insert_points_parallel :125-334 become k_commit; the static level schedule
// of make_queue/find_placement (src/scheduler.rs:6-55) becomes per-round reservation.
//
// Device state (D = 2 or 3, M = D + 1 vertices per simplex).  Each simplex is one record of
// 128 B (3D) / 64 B (2D), so that a visit touches one cache line and one page:
//   v       vertex ids, positively oriented; dead: v[0] = -1, v[1] = forward simplex
//   a       (neighbor << 2) | face-in-neighbor across the face opposite vertex k, -1 on the
//           super-simplex hull
//   x       fp32 vertex coordinates (translated by the bounding-sphere center), vertex order
//   kres, nres: 64-bit reservation keys ((2^32-1-round) << 32 | priority): the least key of
//           the round's points that kill the simplex / that have it as an outside neighbor
//   X64     f64 coordinates, stride 4 (3D) or 2 (2D); vertex v = super (v < M) or point v-M
//   X32     f32 coordinates translated by the bounding-sphere center, same strides
//   cp[p]   conflict pointer of point p: a simplex containing it (or stale: dead -> forward)
// Simplices are append-only: a killed simplex keeps a forward pointer into its killer's
// star, so stale conflict pointers are re-bucketed lazily by a walk when the point is next
// in a window.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "exact.hpp"
#include "predicates.hpp"

namespace vo {

// ST_LOST: the point stopped its cavity search at a simplex that a higher-priority point of the
// round had already reserved, or its walk ran out of the round's step budget (it cannot win the
// round; it is retried in the next one)
enum : int { ST_OK = 0, ST_DEFER = 1, ST_DUP = 2, ST_OVERFLOW = 3, ST_BROKEN = 4, ST_IDLE = 5, ST_LOST = 6 };
enum : uint32_t { ERR_OVERFLOW = 1, ERR_WALK = 2, ERR_CHAIN = 4, ERR_TETCAP = 8, ERR_HULL = 16, ERR_CACHE = 32,
                  ERR_PRECISION = 64, ERR_PENDING = 128 };

constexpr int kWalkLimit = 1 << 16;
constexpr int kChainLimit = 1 << 20;

// Device-side statistics, 64-bit, reduced per wave.  kCntStripes copies, each on lines of its
// own: a wave adds to the copy of its block (block index mod kCntStripes), so that the
// per-wave atomics of a round spread over many addresses instead of queueing at one (measured:
// one shared copy cost the 3D n = 4M construction a fifth of its time).  Readers sum the
// copies (max_* take the maximum).
constexpr int kCntStripes = 128;
struct alignas(256) Counters {
    unsigned long long t1[2], t2[2], t3[2], def[2];
    unsigned long long walk_steps, chain_hops, killed, created, winners, dups, exact_evaluated;
    unsigned long long max_killed, max_facets;
    unsigned long long pend;       // points sent on to the next window
};

// Round totals.  The two that every wavefront of a round updates atomically sit on lines of
// their own, so that their atomics queue separately; the rarely touched rest share one line.
// n_winners and n_live_delta are not atomic targets: round_totals derives them from the
// striped statistics after the round.
struct RoundState {
    alignas(256) int n_tets;          // simplices allocated (append-only high-water mark; one set only)
    alignas(256) int n_next_pending;  // losers and deferred points appended for the next window (a
                                      // round with outcome flags: from the statistics, round_totals)
    alignas(256) int n_q;             // exact-queue fill (may exceed capacity: excess dropped)
    alignas(256) int n_fix;           // shared-facet links to repair after the commit (see Ctx::fix)
    unsigned int err;                 // sticky ERR_* bits
    int n_big;                        // points whose cavity outgrew the group kernel's LDS lists
    int n_nospace;                    // several sets: winners whose set had no simplex space left
    int n_winners;                    // winners of the round (round_totals)
    int n_live_delta;                 // sum over its winners of (created - killed) (round_totals)
    unsigned long long acc_winners;   // statistics totals at the last round_totals call
    long long acc_live;
    unsigned long long acc_pend;
    int flag_n;                       // window slots of the last round whose outcomes are in Ctx::pflag
    alignas(256) int n_compact;       // pending points the planning launch placed from Ctx::pflag
};

// Several independent point sets sharing one device arena (batched and served modes).
// Set s owns a contiguous vertex block (its d+1 super vertices, then its points in priority
// order), one or more simplex ranges handed out by the host, and its own hint grid.
enum : int { SET_ACTIVE = 1, SET_NEED_SPACE = 2, SET_DONE = 4 };
// The five fields that every commit wavefront of a set updates atomically (the allocation
// cursor and four outcome counters) sit 256 B apart, each alone on its line, and apart from the
// fields the kernels read: atomics to one line queue one behind another (loads of that line
// too), and a large set is served by thousands of wavefronts per round.
struct SetRec {
    alignas(256) int cursor; // simplex allocation in the current range: next free
    alignas(256) int end;  // one past the last slot of the current range
    int head, last;        // fresh (never windowed) points: head .. last-1 (point indices)
    int flags;             // SET_*
    int hint_off, hlevels; // hint grid of the set: offset in the hint arena, finest level
    float glo[3], ginv;    // grid origin and 1 / extent in the set's translated frame
    int n;                 // points of the set
    int fresh, woff;       // planner: fresh points windowed this round and their offset
    int fail_lo;           // least slot of a failed request in the current range (slots below it are all written)
    alignas(256) int inserted; // points inserted
    alignas(256) int dups;     // points dropped as duplicates
    alignas(256) int pend;     // points of the set sent to the pending list in the current round
    alignas(256) int live;     // live simplices of the set
};

// round control planned on the device when several sets share the arrays: the kernels read
// the window size and the reservation key of the round from here
struct BatchCtl {
    int nwin, npend;       // window size; pending points copied into the window
    int tail;              // pending points beyond the window's capacity, carried to the next list
    float gamma;
    unsigned int round;
    unsigned long long key_hi;
    int n_done;            // sets completed since the host last collected them
    int n_active;
    int min_window;        // least fresh window of a set per round
    int big_window;        // windows at least this large aim at a higher winning fraction
    int slots;             // set slots that may be in use: one past the highest (host, before each replay)
    int virt, vhead;       // virt = 1: the window is not written out: slots below npend are the
                           // pending list (Ctx::win), slot i >= npend is point vhead + i - npend
};

// Simplex slots of one set's device-planned rounds come from kAllocStripes chunks that the
// planner reserves for each round at the top of the record array (AllocStripe: next free, end);
// a commit wavefront takes its range from the chunk of stripe blockIdx % kAllocStripes, and from
// the shared tail (RoundState::n_tets) once that chunk is spent.  Returning atomics on one
// address serialize (about 11 ns each on MI355X), so a single counter for every wavefront of a
// large round was a chain as long as the round's wavefronts.
constexpr int kAllocStripes = 64;
struct alignas(256) AllocStripe {
    int cur, end;
};

struct Ctx {
    char* rec;             // simplex records (layout in Rec<D>)
    const double* X64;
    const float* X32;
    int* cp;
    const int* win;
    int* pend_next;
    int* big;              // points to re-run alone with large cavity buffers
    int* cav;              // killed lists, one run per window slot: cav[slot * kcap + k]
    int* fac;              // boundary facets (t << 2 | face), likewise: fac[slot * fcap + k]
    uint4* probe = nullptr;  // builds with VO_POINT_PROBE (diagnostics): three entries per window slot
                             // from the lane-group cavity kernel (times, walk, hops, sizes), then
                             // from 3 win_cap on two per slot from the lane-group commit kernel
    int* fnb;              // lane-group kernels: the adjacency entry across each facet as the cavity
                           // search saw it (the simplex outside << 2 | its face, -1 on the hull),
                           // parallel to fac; the commit reads it instead of the record, which
                           // winners of the same commit may be patching
    int* nk;
    int* nf;
    int* status;
    int* hint;             // multi-level grid: level l has 2^l cells per axis, -1 = unset
    float glo[3];          // grid origin (translated frame)
    float ginv;            // 1 / grid extent
    int hlevels;           // finest level
    RoundState* st;
    AllocStripe* ast = nullptr;   // one set, device-planned rounds: the round's allocation chunks
    int* pflag = nullptr;  // device-planned rounds (one set) and the engine's captured rounds: the
                           // lane-group commit writes each window slot's outcome here (its point if
                           // it goes on to the next window, else -1) instead of appending to
                           // pend_next through one counter; append_flagged compacts it later
    Counters* cnt;
    QEntry* q;
    int qcap;
    CEntry* cache;
    uint32_t cmask;
    int filters;           // bit 0: FP32 tier on, bit 1: FP64 filter on (both on in production)
    int solo;              // the window holds one point: cavity membership is stamped in kres
    int reuse;             // lane-group commit: new simplices take the winner's killed records first
    int2* fix;             // lane-group commit without reuse: a winner whose cavity touches a
    int fix_cap;           // higher-priority one lists (new simplex, killed record << 2 | face) for
                           // each shared facet; k_fix_shared links the two sides after the commit
    int abort_early;       // lane-group cavity search: reserve each member as it is found and stop
                           // once a reservation shows a higher-priority point of the round
    int walk_budget;       // lane-group cavity walk in windows of at least 64 points: a point that
                           // has moved this many times stops and resumes there next round (0 = off)
    int M;                 // vertex id of point p is M + p
    int win_cap, kcap, fcap;
    int tet_cap;
    // several sets (null / 0 with one set)
    SetRec* sets;
    const int* pset;       // set of point p
    const BatchCtl* ctl;   // window size and key of the round
    int big_cap;           // capacity of the big list
    int pend_cap;          // capacity of the pending list
    const int* region;     // XCD regions (null: off): [0..7] first slot, [8..15] slot count of region r
};

// Window slots served by this block.  With XCD regions the window is sorted by region (the
// octant of the point, the top three bits of the Morton order that compaction sorts records
// by) and block b serves region b % 8 only: blocks are dispatched to the eight XCDs in turn,
// so each XCD's L2 and translation caches see one region's records.  Iterate with
//   for (kb = it.base; kb < it.lim; kb += it.stride) { k = kb + part; valid = k < it.lim;
//                                                    slot = it.off + k; ... }
struct SlotIter {
    int base, stride, lim, off;
};
__device__ __forceinline__ SlotIter slot_iter(const Ctx& c, int nwin, int per_block) {
    if (c.region) {
        const int r = blockIdx.x & 7;
        return SlotIter{(int)(blockIdx.x >> 3) * per_block, (int)(gridDim.x >> 3) * per_block, c.region[8 + r],
                        c.region[r]};
    }
    return SlotIter{(int)blockIdx.x * per_block, (int)gridDim.x * per_block, nwin, 0};
}

// window size and reservation key of the round: kernel arguments with one set, planned on
// the device with several
__device__ __forceinline__ void round_params(const Ctx& c, int& nwin, unsigned long long& key_hi) {
    if (c.ctl) {
        nwin = c.ctl->nwin;
        key_hi = c.ctl->key_hi;
    }
}

// the round's window: point of slot i
struct WinView {
    const int* w;
    int np, head;
};
__device__ __forceinline__ WinView win_view(const Ctx& c) {
    if (c.ctl && c.ctl->virt) return WinView{c.win, c.ctl->npend, c.ctl->vhead};
    return WinView{c.win, 0x7fffffff, 0};
}
__device__ __forceinline__ int win_point(const WinView& v, int slot) {
    return slot < v.np ? v.w[slot] : v.head + (slot - v.np);
}

// hint grid of a point's set
struct HintView {
    int* h;
    float glo[3];
    float ginv;
    int levels;
};
__device__ __forceinline__ HintView hint_view(const Ctx& c, int p) {
    HintView v;
    if (c.sets) {
        const SetRec& r = c.sets[c.pset[p]];
        v.h = c.hint + r.hint_off;
        v.glo[0] = r.glo[0]; v.glo[1] = r.glo[1]; v.glo[2] = r.glo[2];
        v.ginv = r.ginv;
        v.levels = r.hlevels;
    } else {
        v.h = c.hint;
        v.glo[0] = c.glo[0]; v.glo[1] = c.glo[1]; v.glo[2] = c.glo[2];
        v.ginv = c.ginv;
        v.levels = c.hlevels;
    }
    return v;
}

template <int D> struct VT;
template <> struct VT<3> {
    using F = float4;
    static constexpr int S = 4;
    static __device__ __forceinline__ F ld(const float* X, int v) { return reinterpret_cast<const float4*>(X)[v]; }
};
template <> struct VT<2> {
    using F = float2;
    static constexpr int S = 2;
    static __device__ __forceinline__ F ld(const float* X, int v) { return reinterpret_cast<const float2*>(X)[v]; }
};

__device__ __forceinline__ int comp(int4 a, int k) {
    return k == 0 ? a.x : (k == 1 ? a.y : (k == 2 ? a.z : a.w));
}
__device__ __forceinline__ void set_comp(int4& a, int k, int v) {
    if (k == 0) a.x = v; else if (k == 1) a.y = v; else if (k == 2) a.z = v; else a.w = v;
}

// simplex record layout: byte offsets of v (int x M), a (int x M), x (fp32 coordinates, packed
// in vertex order), kres, nres; 3D: v int4 | a int4 | x 12 floats | kres | nres | 32 B pad;
// 2D: v int3 | a int3 | x 6 floats | kres | nres.  Adjacency and coordinates are contiguous and
// 16 B aligned (3D: bytes 16..79, 2D with the vertices: 0..47), so that a walk step or a cavity
// test reads them with four (3D) / three (2D) 16 B loads: several loads of one line by one lane
// are separate requests, and at full load the request rate, not the line rate, bounds the
// random gathers.  kres and nres share one 16 B load.
template <int D> struct Rec;
template <> struct Rec<3> { static constexpr int B = 128, V = 0, A = 16, X = 32, K = 80, N = 88; };
template <> struct Rec<2> { static constexpr int B = 64, V = 0, A = 12, X = 24, K = 48, N = 56; };

template <int D> __device__ __forceinline__ char* rp(const Ctx& c, int t) { return c.rec + (size_t)t * Rec<D>::B; }
template <int D> __device__ __forceinline__ int* v_ptr(const Ctx& c, int t, int k) {
    return reinterpret_cast<int*>(rp<D>(c, t) + Rec<D>::V) + k;
}
template <int D> __device__ __forceinline__ int* a_ptr(const Ctx& c, int t, int k) {
    return reinterpret_cast<int*>(rp<D>(c, t) + Rec<D>::A) + k;
}
template <int D> __device__ __forceinline__ unsigned long long* kres_p(const Ctx& c, int t) {
    return reinterpret_cast<unsigned long long*>(rp<D>(c, t) + Rec<D>::K);
}
template <int D> __device__ __forceinline__ unsigned long long* nres_p(const Ctx& c, int t) {
    return reinterpret_cast<unsigned long long*>(rp<D>(c, t) + Rec<D>::N);
}
// vertex ids (2D: w = -1)
template <int D> __device__ __forceinline__ int4 ld_v(const Ctx& c, int t) {
    if constexpr (D == 3) return *reinterpret_cast<const int4*>(rp<D>(c, t) + Rec<D>::V);
    else {
        const int* q = v_ptr<D>(c, t, 0);
        return make_int4(q[0], q[1], q[2], -1);
    }
}
// v[0] and v[1] of record t (a dead record's mark and forward pointer) in one 8-byte load.  A
// relaxed atomic load, which the compiler never splits: from an int4 it loaded v[0], waited,
// and only then loaded v[1] when v[0] < 0, two dependent round trips per forward hop.  Nothing
// writes these words while a cavity search runs.
template <int D> __device__ __forceinline__ int2 ld_v01(const Ctx& c, int t) {
    const unsigned long long w = __hip_atomic_load(reinterpret_cast<unsigned long long*>(v_ptr<D>(c, t, 0)),
                                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    return make_int2((int)(unsigned)w, (int)(unsigned)(w >> 32));
}
template <int D> __device__ __forceinline__ void st_v(const Ctx& c, int t, int4 v) {
    if constexpr (D == 3) *reinterpret_cast<int4*>(rp<D>(c, t) + Rec<D>::V) = v;
    else {
        int* q = v_ptr<D>(c, t, 0);
        q[0] = v.x; q[1] = v.y; q[2] = v.z;
    }
}
// adjacency (2D: w = -1)
template <int D> __device__ __forceinline__ int4 ld_a(const Ctx& c, int t) {
    if constexpr (D == 3) return *reinterpret_cast<const int4*>(rp<D>(c, t) + Rec<D>::A);
    else {
        const int* q = a_ptr<D>(c, t, 0);
        return make_int4(q[0], q[1], q[2], -1);
    }
}
// a dead simplex keeps a forward pointer to a live one of its killer's star
template <int D> __device__ __forceinline__ void kill(const Ctx& c, int t, int fwd) {
    int* q = v_ptr<D>(c, t, 0);
    q[0] = -1;
    q[1] = fwd;
}

template <int D> __device__ __forceinline__ int index_of(int4 a, int v) {
    if (a.x == v) return 0;
    if (a.y == v) return 1;
    if (a.z == v) return 2;
    return (D == 3 && a.w == v) ? 3 : -1;
}

// Simplex t with vertex i replaced by p, reordered by an even permutation so that p comes
// first: orientation is preserved, and faces 1..D of the result contain p.
template <int D> __device__ __forceinline__ int4 replace_first(int4 v, int i, int p) {
    if (D == 3) {
        switch (i) {
            case 0: return make_int4(p, v.y, v.z, v.w);
            case 1: return make_int4(p, v.x, v.w, v.z);
            case 2: return make_int4(p, v.x, v.y, v.w);
            default: return make_int4(p, v.x, v.z, v.y);
        }
    } else {
        switch (i) {
            case 0: return make_int4(p, v.y, v.z, -1);
            case 1: return make_int4(p, v.z, v.x, -1);
            default: return make_int4(p, v.x, v.y, -1);
        }
    }
}

struct LocalCnt {
    unsigned t1[2], t2[2], t3[2], def[2];
    unsigned walk, chain, killed, created, wins, dups, pend;
    unsigned maxk, maxf;
};

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ long long wave_sum_i64(long long v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ unsigned wave_max(unsigned v) {
    for (int o = 32; o > 0; o >>= 1) { unsigned w = __shfl_xor(v, o, 64); v = w > v ? w : v; }
    return v;
}

// one atomic per wave per counter; all lanes of the wave must call this
// sum of the striped copies (host side)
inline Counters sum_counters(const Counters* stripes) {
    Counters r = {};
    for (int i = 0; i < kCntStripes; ++i) {
        const Counters& x = stripes[i];
        for (int k = 0; k < 2; ++k) {
            r.t1[k] += x.t1[k];
            r.t2[k] += x.t2[k];
            r.t3[k] += x.t3[k];
            r.def[k] += x.def[k];
        }
        r.walk_steps += x.walk_steps;
        r.chain_hops += x.chain_hops;
        r.killed += x.killed;
        r.created += x.created;
        r.winners += x.winners;
        r.dups += x.dups;
        r.pend += x.pend;
        r.exact_evaluated += x.exact_evaluated;
        r.max_killed = x.max_killed > r.max_killed ? x.max_killed : r.max_killed;
        r.max_facets = x.max_facets > r.max_facets ? x.max_facets : r.max_facets;
    }
    return r;
}

__device__ inline void flush_counters(Counters* c, const LocalCnt& l) {
    c += blockIdx.x & (kCntStripes - 1);
    const bool lead = __lane_id() == 0;
#define VO_FLUSH(field, src)                                          \
    {                                                                 \
        unsigned long long s = wave_sum((unsigned long long)(src));   \
        if (lead && s) atomicAdd(&c->field, s);                       \
    }
    VO_FLUSH(t1[0], l.t1[0]) VO_FLUSH(t1[1], l.t1[1]) VO_FLUSH(t2[0], l.t2[0]) VO_FLUSH(t2[1], l.t2[1])
    VO_FLUSH(t3[0], l.t3[0]) VO_FLUSH(t3[1], l.t3[1]) VO_FLUSH(def[0], l.def[0]) VO_FLUSH(def[1], l.def[1])
    VO_FLUSH(walk_steps, l.walk) VO_FLUSH(chain_hops, l.chain) VO_FLUSH(killed, l.killed)
    VO_FLUSH(created, l.created) VO_FLUSH(winners, l.wins) VO_FLUSH(dups, l.dups) VO_FLUSH(pend, l.pend)
#undef VO_FLUSH
    unsigned mk = wave_max(l.maxk), mf = wave_max(l.maxf);
    if (lead) {
        if (mk) atomicMax(&c->max_killed, (unsigned long long)mk);
        if (mf) atomicMax(&c->max_facets, (unsigned long long)mf);
    }
}

// Winners and the change in live simplices since the last call, from the striped statistics
// (each commit wavefront adds its winners, created and killed simplices there), into
// st->n_winners / st->n_live_delta, and with pend_count the points sent on to the next window
// into st->n_next_pending (rounds whose commit wrote Ctx::pflag instead of appending).  Called
// by every thread of one block (block-uniform).
__device__ inline void round_totals(const Counters* cnt, RoundState* st, bool pend_count = false) {
    __shared__ unsigned long long s_w[16], s_p[16];
    __shared__ long long s_l[16];
    unsigned long long w = 0, pn = 0;
    long long l = 0;
    for (int k = threadIdx.x; k < kCntStripes; k += blockDim.x) {
        w += cnt[k].winners;
        pn += cnt[k].pend;
        l += (long long)cnt[k].created - (long long)cnt[k].killed;
    }
    w = wave_sum(w);
    pn = wave_sum(pn);
    l = wave_sum_i64(l);
    const int wv = threadIdx.x / 64, nw = (blockDim.x + 63) / 64;
    if (__lane_id() == 0 && wv < 16) {
        s_w[wv] = w;
        s_p[wv] = pn;
        s_l[wv] = l;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long W = 0, P = 0;
        long long L = 0;
        for (int i = 0; i < nw && i < 16; ++i) {
            W += s_w[i];
            P += s_p[i];
            L += s_l[i];
        }
        st->n_winners = (int)(W - st->acc_winners);
        st->n_live_delta = (int)(L - st->acc_live);
        if (pend_count) st->n_next_pending = (int)(P - st->acc_pend);
        st->acc_winners = W;
        st->acc_live = L;
        st->acc_pend = P;
    }
    __syncthreads();
}

// exclusive prefix sum across the 64 lanes; all lanes must call
__device__ __forceinline__ int wave_excl_scan(int v, int& total) {
    int lane = __lane_id();
    int x = v;
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    total = __shfl(x, 63, 64);
    return x - v;
}

// The points of a round's outcome flags (Ctx::pflag, slots 0 .. n-1; -1: nothing) appended to
// dst at positions taken from *counter: block b of nb handles 1,024 slots per step with one
// atomic per step (256 threads, every thread of the block calls this).  The list's order is
// free: the next window holds the same points either way.
__device__ inline void append_flagged(const int* flags, int n, int* dst, int cap, int* counter, unsigned* err, int b,
                                      int nb) {
    __shared__ int s_w[4], s_base;
    const int wv = threadIdx.x >> 6, lane = (int)__lane_id();
    for (int b0 = b * 1024; b0 < n; b0 += nb * 1024) {
        int v[4], k = 0;
        for (int j = 0; j < 4; ++j) {
            const int i = b0 + j * 256 + threadIdx.x;
            v[j] = i < n ? flags[i] : -1;
            k += v[j] >= 0;
        }
        int tot = 0;
        const int off = wave_excl_scan(k, tot);
        if (lane == 0) s_w[wv] = tot;
        __syncthreads();
        if (threadIdx.x == 0) {
            const int all = s_w[0] + s_w[1] + s_w[2] + s_w[3];
            s_base = all ? atomicAdd(counter, all) : 0;
        }
        __syncthreads();
        int pos = s_base + off;
        for (int w = 0; w < wv; ++w) pos += s_w[w];
        for (int j = 0; j < 4; ++j) {
            if (v[j] < 0) continue;
            if (pos < cap) dst[pos] = v[j];
            else atomicOr(err, ERR_PENDING);
            ++pos;
        }
        __syncthreads();                   // s_w and s_base are rewritten by the next step
    }
}

// Simplex slots for this lane's request of n (req false: none): the base slot, or -1 when
// there is no request or no room (one set: ERR_TETCAP; several sets: the set is flagged for
// another range and the point retries).  One atomicAdd per wavefront when all requests come
// from one set (always, with one set), else one per request.  Every lane of the wavefront
// must call this.
__device__ inline int alloc_slots(const Ctx& c, bool req, int n, int p) {
    if (!c.sets) {
        int total = 0;
        const int off = wave_excl_scan(req ? n : 0, total);
        int base0 = 0;
        if (__lane_id() == 0 && total > 0) {
            base0 = -1;
            if (c.ast) {
                AllocStripe& a = c.ast[blockIdx.x & (kAllocStripes - 1)];
                const int b = atomicAdd(&a.cur, total);
                if (b + total <= a.end) base0 = b;
            }
            if (base0 < 0) base0 = atomicAdd(&c.st->n_tets, total);
        }
        base0 = __shfl(base0, 0, 64);
        if (!req) return -1;
        if (base0 + off + n > c.tet_cap) {
            atomicOr(&c.st->err, ERR_TETCAP);
            return -1;
        }
        return base0 + off;
    }
    const int s = req ? c.pset[p] : -1;
    const unsigned long long rm = __ballot(req);
    if (!rm) return -1;
    const int lead = __ffsll((long long)rm) - 1;
    const int s0 = __shfl(s, lead, 64);
    int base = -1, set = s;
    if (__ballot(req && s != s0) == 0) {
        int total = 0;
        const int off = wave_excl_scan(req ? n : 0, total);
        int base0 = 0;
        if ((int)__lane_id() == lead) base0 = atomicAdd(&c.sets[s0].cursor, total);
        base0 = __shfl(base0, lead, 64);
        base = base0 + off;
        set = s0;
    } else if (req) {
        base = atomicAdd(&c.sets[s].cursor, n);
    }
    if (!req) return -1;
    if (base + n > c.sets[set].end) {
        atomicMin(&c.sets[set].fail_lo, base);
        atomicOr(&c.sets[set].flags, SET_NEED_SPACE);
        atomicAdd(&c.st->n_nospace, 1);
        return -1;
    }
    return base;
}

// a point's outcome in its set's counters (several sets only): inserted (with its live-count
// change), back to the pending list, or dropped as a duplicate.  One atomic per counter and
// wavefront when all reporting lanes share a set, else one per lane.  Every lane of the
// wavefront must call this.
__device__ inline void set_outcome(const Ctx& c, bool report, int p, bool ins, int dlive, bool pend, bool dup) {
    if (!c.sets) return;
    const int s = report ? c.pset[p] : -1;
    const unsigned long long rm = __ballot(report);
    if (!rm) return;
    const int lead = __ffsll((long long)rm) - 1;
    const int s0 = __shfl(s, lead, 64);
    if (__ballot(report && s != s0) == 0) {
        const long long ni = wave_sum_i64(report && ins ? 1 : 0), nl = wave_sum_i64(report && ins ? dlive : 0);
        const long long np = wave_sum_i64(report && pend ? 1 : 0), nd = wave_sum_i64(report && dup ? 1 : 0);
        if ((int)__lane_id() == lead) {
            SetRec& r = c.sets[s0];
            if (ni) atomicAdd(&r.inserted, (int)ni);
            if (nl) atomicAdd(&r.live, (int)nl);
            if (np) atomicAdd(&r.pend, (int)np);
            if (nd) atomicAdd(&r.dups, (int)nd);
        }
    } else if (report) {
        SetRec& r = c.sets[s];
        if (ins) {
            atomicAdd(&r.inserted, 1);
            atomicAdd(&r.live, dlive);
        } else if (pend) {
            atomicAdd(&r.pend, 1);
        } else if (dup) {
            atomicAdd(&r.dups, 1);
        }
    }
}

__device__ __forceinline__ void enqueue(const Ctx& c, uint32_t type, const uint32_t* ids) {
    int slot = atomicAdd(&c.st->n_q, 1);
    if (slot < c.qcap) {
        c.q[slot].type = type;
        for (int i = 0; i < 5; ++i) c.q[slot].id[i] = ids[i];
    }
}

// ---------------------------------------------------------------- predicates
// Tier 1 reads the simplex's own coordinate record tx[t] (the fp32 coordinates of its
// vertices, in vertex order); tiers 2 and 3 need vertex ids and load tv[t] only then.

// 3D coordinates: 12 packed floats in three float4 <-> four float4 with w = 0
__device__ __forceinline__ void unpack_x3(float4 f0, float4 f1, float4 f2, float4* a) {
    a[0] = make_float4(f0.x, f0.y, f0.z, 0.f);
    a[1] = make_float4(f0.w, f1.x, f1.y, 0.f);
    a[2] = make_float4(f1.z, f1.w, f2.x, 0.f);
    a[3] = make_float4(f2.y, f2.z, f2.w, 0.f);
}

template <int D> __device__ __forceinline__ void load_tx(const Ctx& c, int t, typename VT<D>::F* a) {
    if constexpr (D == 3) {
        const float4* q = reinterpret_cast<const float4*>(rp<D>(c, t) + Rec<D>::X);
        unpack_x3(q[0], q[1], q[2], a);
    } else {
        const float2* q = reinterpret_cast<const float2*>(rp<D>(c, t) + Rec<D>::X);
        a[0] = q[0]; a[1] = q[1]; a[2] = q[2];
    }
}

template <int D> __device__ __forceinline__ void store_tx(const Ctx& c, int t, const typename VT<D>::F* a) {
    if constexpr (D == 3) {
        float4* q = reinterpret_cast<float4*>(rp<D>(c, t) + Rec<D>::X);
        q[0] = make_float4(a[0].x, a[0].y, a[0].z, a[1].x);
        q[1] = make_float4(a[1].y, a[1].z, a[2].x, a[2].y);
        q[2] = make_float4(a[2].z, a[3].x, a[3].y, a[3].z);
    } else {
        float2* q = reinterpret_cast<float2*>(rp<D>(c, t) + Rec<D>::X);
        q[0] = a[0]; q[1] = a[1]; q[2] = a[2];
    }
}

// adjacency and coordinates of simplex t in one run of 16 B loads (2D: with the vertices)
template <int D> __device__ __forceinline__ void ld_ax(const Ctx& c, int t, int4& adj, typename VT<D>::F* a) {
    if constexpr (D == 3) {
        const float4* q = reinterpret_cast<const float4*>(rp<D>(c, t) + Rec<D>::A);
        const float4 f0 = q[0], f1 = q[1], f2 = q[2], f3 = q[3];
        adj = make_int4(__float_as_int(f0.x), __float_as_int(f0.y), __float_as_int(f0.z), __float_as_int(f0.w));
        unpack_x3(f1, f2, f3, a);
    } else {
        const int4* q = reinterpret_cast<const int4*>(rp<D>(c, t));
        const int4 w0 = q[0], w1 = q[1], w2 = q[2];
        adj = make_int4(w0.w, w1.x, w1.y, -1);
        a[0] = make_float2(__int_as_float(w1.z), __int_as_float(w1.w));
        a[1] = make_float2(__int_as_float(w2.x), __int_as_float(w2.y));
        a[2] = make_float2(__int_as_float(w2.z), __int_as_float(w2.w));
    }
}

// kres and nres of simplex t in one 16 B load
template <int D> __device__ __forceinline__ ulonglong2 ld_res(const Ctx& c, int t) {
    return *reinterpret_cast<const ulonglong2*>(rp<D>(c, t) + Rec<D>::K);
}

// tiers 2 and 3 of the in-sphere test of vertex pv against simplex v
template <int D> __device__ int conflict_slow(const Ctx& c, int4 v, int pv, LocalCnt& lc) {
    constexpr int S = VT<D>::S;
    const double* X = c.X64;
    int s = SIGN_UNDECIDED;
    if (c.filters & 2) {
        if (D == 3)
            s = t2_insphere(X + S * v.x, X + S * v.y, X + S * v.z, X + S * v.w, X + S * pv);
        else
            s = t2_incircle(X + S * v.x, X + S * v.y, X + S * v.z, X + S * pv);
        if (s != SIGN_UNDECIDED) { lc.t2[1]++; return s; }
    }
    uint32_t ids[5];
    const uint32_t type = D == 3 ? PRED_INSPHERE : PRED_INCIRCLE;
    if (D == 3) { ids[0] = v.x; ids[1] = v.y; ids[2] = v.z; ids[3] = v.w; ids[4] = pv; }
    else { ids[0] = v.x; ids[1] = v.y; ids[2] = v.z; ids[3] = pv; ids[4] = kEmpty; }
    s = cache_lookup(c.cache, c.cmask, type, ids);
    if (s != SIGN_UNDECIDED) { lc.t3[1]++; return s; }
    lc.def[1]++;
    enqueue(c, type, ids);
    return SIGN_UNDECIDED;
}

// tiers 2 and 3 of an orientation test on vertex ids
template <int D> __device__ int orient_slow(const Ctx& c, const int* ids, LocalCnt& lc) {
    constexpr int S = VT<D>::S;
    const double* X = c.X64;
    int s = SIGN_UNDECIDED;
    if (c.filters & 2) {
        if (D == 3) s = t2_orient3d(X + S * ids[0], X + S * ids[1], X + S * ids[2], X + S * ids[3]);
        else s = t2_orient2d(X + S * ids[0], X + S * ids[1], X + S * ids[2]);
        if (s != SIGN_UNDECIDED) { lc.t2[0]++; return s; }
    }
    uint32_t u[5] = {(uint32_t)ids[0], (uint32_t)ids[1], (uint32_t)ids[2], D == 3 ? (uint32_t)ids[3] : kEmpty, kEmpty};
    const uint32_t type = D == 3 ? PRED_ORIENT3 : PRED_ORIENT2;
    s = cache_lookup(c.cache, c.cmask, type, u);
    if (s != SIGN_UNDECIDED) { lc.t3[0]++; return s; }
    lc.def[0]++;
    enqueue(c, type, u);
    return SIGN_UNDECIDED;
}

// p (vertex pv, coordinates pf) strictly inside the circumsphere of simplex t whose vertex
// coordinates are a[]: > 0
template <int D> __device__ __forceinline__ int conflict_at(const Ctx& c, const typename VT<D>::F* a, int t, int pv,
                                                            typename VT<D>::F pf, LocalCnt& lc) {
    if (c.filters & 1) {
        int s;
        if constexpr (D == 3) s = t1_insphere(a[0], a[1], a[2], a[3], pf);
        else s = t1_incircle(a[0], a[1], a[2], pf);
        if (s != SIGN_UNDECIDED) { lc.t1[1]++; return s; }
    }
    return conflict_slow<D>(c, ld_v<D>(c, t), pv, lc);
}

// orientation of simplex t (coordinates a[]) with vertex k replaced by p: < 0 means p lies
// beyond face k
template <int D> __device__ __forceinline__ int orient_at(const Ctx& c, const typename VT<D>::F* a, int t, int k,
                                                          int pv, typename VT<D>::F pf, LocalCnt& lc) {
    if (c.filters & 1) {
        typename VT<D>::F b[4];
        for (int i = 0; i < D + 1; ++i) b[i] = i == k ? pf : a[i];
        int s;
        if constexpr (D == 3) s = t1_orient3d(b[0], b[1], b[2], b[3]);
        else s = t1_orient2d(b[0], b[1], b[2]);
        if (s != SIGN_UNDECIDED) { lc.t1[0]++; return s; }
    }
    const int4 v = ld_v<D>(c, t);
    int ids[4] = {v.x, v.y, v.z, v.w};
    ids[k] = pv;
    return orient_slow<D>(c, ids, lc);
}

// coordinate record of replace_first(v, i, p) from the old record ox: p first, then the old
// vertices in the order replace_first puts them (selects only: a dynamically indexed private
// array would live in scratch memory)
template <int D>
__device__ __forceinline__ void star_record(const typename VT<D>::F* ox, int i, typename VT<D>::F pf,
                                            typename VT<D>::F* nx) {
    nx[0] = pf;
    if constexpr (D == 3) {
        nx[1] = i == 0 ? ox[1] : ox[0];
        nx[2] = i == 0 ? ox[2] : (i == 1 ? ox[3] : (i == 2 ? ox[1] : ox[2]));
        nx[3] = i == 0 ? ox[3] : (i == 1 ? ox[2] : (i == 2 ? ox[3] : ox[1]));
    } else {
        nx[1] = i == 0 ? ox[1] : (i == 1 ? ox[2] : ox[0]);
        nx[2] = i == 0 ? ox[2] : (i == 1 ? ox[0] : ox[1]);
    }
}

// Coordinates of the new simplex replace_first(v(t), i, p) written to record N straight from
// t's packed record (3D: scalar selects on the packed floats, no private array)
// 3D: the packed coordinates (f0, f1, f2) of a killed simplex with vertex i replaced by pf, in
// replace_first's vertex order, stored as simplex N's packed coordinates
__device__ __forceinline__ void store_star_coords3(const Ctx& c, float4 f0, float4 f1, float4 f2, int i, float4 pf,
                                                   int N) {
    // old vertex k = (X[3k], X[3k+1], X[3k+2]) of the packed floats
    const float X[12] = {f0.x, f0.y, f0.z, f0.w, f1.x, f1.y, f1.z, f1.w, f2.x, f2.y, f2.z, f2.w};
    // new vertices 1..3 come from old vertices s1, s2, s3 (replace_first's even permutation)
    const int s1 = i == 0 ? 1 : 0;
    const int s2 = i == 0 ? 2 : (i == 1 ? 3 : (i == 2 ? 1 : 2));
    const int s3 = i == 0 ? 3 : (i == 1 ? 2 : (i == 2 ? 3 : 1));
    float o[9];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        o[j] = s1 == 0 ? X[j] : (s1 == 1 ? X[3 + j] : (s1 == 2 ? X[6 + j] : X[9 + j]));
        o[3 + j] = s2 == 0 ? X[j] : (s2 == 1 ? X[3 + j] : (s2 == 2 ? X[6 + j] : X[9 + j]));
        o[6 + j] = s3 == 0 ? X[j] : (s3 == 1 ? X[3 + j] : (s3 == 2 ? X[6 + j] : X[9 + j]));
    }
    float4* w = reinterpret_cast<float4*>(rp<3>(c, N) + Rec<3>::X);
    w[0] = make_float4(pf.x, pf.y, pf.z, o[0]);
    w[1] = make_float4(o[1], o[2], o[3], o[4]);
    w[2] = make_float4(o[5], o[6], o[7], o[8]);
}

// the packed coordinates of killed simplex t (3D: three float4, 2D: three float2), loaded
// apart from their use so that a caller can issue the load before its stores: a load placed
// after a store to another record waits for it (the compiler cannot tell the records apart)
template <int D> struct StarSrc;
template <> struct StarSrc<3> { float4 f0, f1, f2; };
template <> struct StarSrc<2> { float2 o0, o1, o2; };
template <int D> __device__ __forceinline__ StarSrc<D> ld_star_src(const Ctx& c, int t) {
    if constexpr (D == 3) {
        const float4* q = reinterpret_cast<const float4*>(rp<D>(c, t) + Rec<D>::X);
        return StarSrc<3>{q[0], q[1], q[2]};
    } else {
        const float2* q = reinterpret_cast<const float2*>(rp<D>(c, t) + Rec<D>::X);
        return StarSrc<2>{q[0], q[1], q[2]};
    }
}

// coordinates of replace_first(v(t), i, p) as simplex N's record, from t's (src)
template <int D>
__device__ __forceinline__ void store_star_coords(const Ctx& c, const StarSrc<D>& src, int i, typename VT<D>::F pf,
                                                  int N) {
    if constexpr (D == 3) {
        store_star_coords3(c, src.f0, src.f1, src.f2, i, pf, N);
    } else {
        // scalar selects on the three points (selecting elements of a private array by i
        // became an indexed array in scratch memory, a store and two reloads per facet)
        const float2 o0 = src.o0, o1 = src.o1, o2 = src.o2;
        const bool i0 = i == 0, i1 = i == 1;
        typename VT<D>::F nx[4];
        nx[0] = pf;
        nx[1] = make_float2(i0 ? o1.x : (i1 ? o2.x : o0.x), i0 ? o1.y : (i1 ? o2.y : o0.y));
        nx[2] = make_float2(i0 ? o2.x : (i1 ? o0.x : o1.x), i0 ? o2.y : (i1 ? o0.y : o1.y));
        nx[3] = pf;
        store_tx<D>(c, N, nx);
    }
}

// position in the old simplex of vertex k (k >= 1) of replace_first(v, i, p)
template <int D> __device__ __forceinline__ int first_src(int i, int k) {
    if (D == 3) {
        const int code = i == 0 ? 0x321 : (i == 1 ? 0x230 : (i == 2 ? 0x310 : 0x120));
        return (code >> (4 * (k - 1))) & 0xF;
    } else {
        const int code = i == 0 ? 0x21 : (i == 1 ? 0x02 : 0x10);
        return (code >> (4 * (k - 1))) & 0xF;
    }
}

// cell of pf at level l of the hint grid (levels stored coarse to fine, one after another)
template <int D> __device__ __forceinline__ int hint_cell(const HintView& v, typename VT<D>::F pf, int l) {
    const int g = 1 << l;
    const float fg = (float)g;
    const int off = D == 3 ? (((1 << (3 * l)) - 1) / 7) : (((1 << (2 * l)) - 1) / 3);
    int ix = (int)fminf(fmaxf((pf.x - v.glo[0]) * v.ginv * fg, 0.0f), fg - 1.0f);
    int iy = (int)fminf(fmaxf((pf.y - v.glo[1]) * v.ginv * fg, 0.0f), fg - 1.0f);
    int iz = 0;
    if constexpr (D == 3) iz = (int)fminf(fmaxf((pf.z - v.glo[2]) * v.ginv * fg, 0.0f), fg - 1.0f);
    return off + (iz * g + iy) * g + ix;
}

// walk start for a fresh point: the finest level whose cell has been visited (level 0 holds
// the set's first simplex); the levels' loads are issued together
template <int D> __device__ __forceinline__ int hint_lookup(const HintView& v, typename VT<D>::F pf) {
    constexpr int LMAX = D == 3 ? 8 : 12;
    int h[LMAX + 1];
#pragma unroll
    for (int l = 0; l <= LMAX; ++l) h[l] = l <= v.levels ? v.h[l == 0 ? 0 : hint_cell<D>(v, pf, l)] : -1;
    int best = h[0];
#pragma unroll
    for (int l = 1; l <= LMAX; ++l)
        if (h[l] >= 0) best = h[l];
    return best;
}

// A walk that ended in simplex t (vertex coordinates a) leaves t as the hint of pf's cells, at
// the levels whose cells are at least as wide as t (level 1 always).  A finer cell written while
// the simplices around it were much larger than the cell would hand a point that arrives there
// much later a stale simplex: the forward chain from it follows one branch of that simplex's
// history and can end far from the point, and the point's walk is then long.
template <int D>
__device__ __forceinline__ void hint_store(const HintView& v, typename VT<D>::F pf, const typename VT<D>::F* a, int t) {
    float e2 = 0.f;
#pragma unroll
    for (int k = 0; k <= D; ++k) {
        const float dx = a[k].x - pf.x, dy = a[k].y - pf.y;
        float d2 = dx * dx + dy * dy;
        if constexpr (D == 3) d2 += (a[k].z - pf.z) * (a[k].z - pf.z);
        e2 = fmaxf(e2, d2);
    }
    // the cell width of level l is 1 / (ginv 2^l)
    const float w = v.ginv * sqrtf(e2);
    const int lw = w > 0.f ? (int)floorf(-log2f(w)) : v.levels;
    const int lmax = min(max(lw, 1), v.levels);
    for (int l = 1; l <= lmax; ++l) v.h[hint_cell<D>(v, pf, l)] = t;
}

__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16;
    return x;
}

// ---------------------------------------------------------------- k_cavity
// One thread per window point: follow dead simplices forward, walk to the simplex that
// contains p (orientation tests), BFS the conflict region (in-sphere tests), then reserve
// the killed simplices (kres) and the outside neighbors (nres) with 64-bit atomicMin.
template <int D>
__global__ void __launch_bounds__(256) k_cavity(Ctx c, int nwin, unsigned long long key_hi) {
    using V = VT<D>;
    constexpr int M = D + 1;
    constexpr int KL = 32;                       // cavity entries mirrored in LDS
    __shared__ int s_cav[KL][256];
    const int lt = threadIdx.x;
    LocalCnt lc = {};
    round_params(c, nwin, key_hi);
    const WinView wv = win_view(c);
    const SlotIter it = slot_iter(c, nwin, blockDim.x);
    for (int kb = it.base; kb < it.lim; kb += it.stride) {
    const int slot = it.off + kb + threadIdx.x;
    if (kb + (int)threadIdx.x < it.lim) {
        const int p = win_point(wv, slot);
        const int pv = c.M + p;
        const typename V::F pf = V::ld(c.X32, pv);
        const HintView hv = hint_view(c, p);
        int st = ST_OK;
        int t = c.cp[p];
        if (t < 0) t = hint_lookup<D>(hv, pf);
        int2 tt = ld_v01<D>(c, t);
        for (int hops = 0; tt.x < 0; ++hops) {
            if (hops > kChainLimit) { atomicOr(&c.st->err, ERR_CHAIN); st = ST_BROKEN; break; }
            t = tt.y;
            tt = ld_v01<D>(c, t);
            lc.chain++;
        }
        // remembering stochastic visibility walk
        int came = -1;
        uint32_t rng = mix32((uint32_t)p * 0x9E3779B9u + 0x632BE5ABu);
        typename V::F a[4];
        load_tx<D>(c, t, a);
        for (int step = 0; st == ST_OK; ++step) {
            if (step > kWalkLimit) { atomicOr(&c.st->err, ERR_WALK); st = ST_BROKEN; break; }
            const int4 aa = ld_a<D>(c, t);
            rng = rng * 1664525u + 1013904223u;
            const int r = (int)((rng >> 16) % M);
            int moved = 0;
            for (int kk = 0; kk < M; ++kk) {
                int k = r + kk;
                if (k >= M) k -= M;
                const int nb = comp(aa, k);
                if (nb >= 0 && (nb >> 2) == came) continue;
                const int s = orient_at<D>(c, a, t, k, pv, pf, lc);
                if (s == SIGN_UNDECIDED) { st = ST_DEFER; break; }
                if (s < 0) {
                    if (nb < 0) { atomicOr(&c.st->err, ERR_HULL); st = ST_BROKEN; break; }
                    came = t;
                    t = nb >> 2;
                    load_tx<D>(c, t, a);
                    moved = 1;
                    break;
                }
            }
            lc.walk++;
            if (!moved) break;
        }
        c.cp[p] = t;
        hint_store<D>(hv, pf, a, t);
        int nk = 0, nf = 0;
        if (st == ST_OK) {
            const int s0 = conflict_at<D>(c, a, t, pv, pf, lc);
            if (s0 == SIGN_UNDECIDED) st = ST_DEFER;
            else if (s0 <= 0) st = ST_DUP;   // on the circumsphere of its own simplex: a vertex
        }
        const unsigned long long skey = key_hi | (unsigned)p;
        if (st == ST_OK) {
            c.cav[(size_t)slot * c.kcap] = t;
            s_cav[0][lt] = t;
            if (c.solo) *kres_p<D>(c, t) = skey;
            nk = 1;
            for (int qi = 0; qi < nk && st == ST_OK; ++qi) {
                const int ct = qi < KL ? s_cav[qi][lt] : c.cav[(size_t)slot * c.kcap + (qi)];
                const int4 aa = ld_a<D>(c, ct);
                for (int f = 0; f < M; ++f) {
                    const int nb = comp(aa, f);
                    if (nb >= 0) {
                        const int n = nb >> 2;
                        bool seen = false;
                        if (c.solo) {
                            seen = *kres_p<D>(c, n) == skey;
                        } else {
                            const int nl = nk < KL ? nk : KL;
                            for (int j = 0; j < nl && !seen; ++j) seen = s_cav[j][lt] == n;
                            for (int j = KL; j < nk && !seen; ++j) seen = c.cav[(size_t)slot * c.kcap + (j)] == n;
                        }
                        if (seen) continue;
                        typename V::F an[4];
                        load_tx<D>(c, n, an);
                        const int s = conflict_at<D>(c, an, n, pv, pf, lc);
                        if (s == SIGN_UNDECIDED) { st = ST_DEFER; break; }
                        if (s > 0) {
                            if (nk == c.kcap) { st = ST_OVERFLOW; break; }
                            c.cav[(size_t)slot * c.kcap + (nk)] = n;
                            if (nk < KL) s_cav[nk][lt] = n;
                            if (c.solo) *kres_p<D>(c, n) = skey;
                            nk++;
                            continue;
                        }
                    }
                    if (nf == c.fcap) { st = ST_OVERFLOW; break; }
                    c.fac[(size_t)slot * c.fcap + (nf)] = (ct << 2) | f;
                    nf++;
                }
            }
        }
        if (st == ST_OK) {
            const unsigned long long key = key_hi | (unsigned)p;
            for (int j = 0; j < nk; ++j) atomicMin(kres_p<D>(c, c.cav[(size_t)slot * c.kcap + (j)]), key);
            for (int j = 0; j < nf; ++j) {
                const int fe = c.fac[(size_t)slot * c.fcap + (j)];
                const int nb = *a_ptr<D>(c, fe >> 2, fe & 3);
                if (nb >= 0) atomicMin(nres_p<D>(c, nb >> 2), key);
            }
        }
        // a window point whose cavity outgrows the lists is re-run alone; alone, it is an error
        if (st == ST_OVERFLOW && c.solo) atomicOr(&c.st->err, ERR_OVERFLOW);
        if (st == ST_DUP) lc.dups++;
        c.nk[slot] = nk;
        c.nf[slot] = nf;
        c.status[slot] = st;
    }
    }
    flush_counters(c.cnt, lc);
}

// ---------------------------------------------------------------- k_commit
// One thread per window point.  A point wins when it holds the least key on every simplex it
// kills and no earlier point of the round kills one of its outside neighbors or has one of
// its killed simplices as an outside neighbor: the winners' cavities are then pairwise
// disjoint and non-adjacent, and committing them in parallel equals inserting them one by
// one.  Winners get their simplex slots from one atomicAdd per wave (exclusive wave scan),
// build one simplex per boundary facet, link internal faces by rotating around each ridge
// inside the cavity, patch the outside neighbors, and leave forward pointers in the killed
// simplices.  Losers and deferred points go to the next window (ballot + popcount).
template <int D>
__global__ void __launch_bounds__(256) k_commit(Ctx c, int nwin, unsigned long long key_hi) {
    constexpr int M = D + 1;
    LocalCnt lc = {};
    round_params(c, nwin, key_hi);
    const WinView wv = win_view(c);
    const SlotIter it = slot_iter(c, nwin, blockDim.x);
    for (int kb = it.base; kb < it.lim; kb += it.stride) {
    const int slot = it.off + kb + threadIdx.x;
    const bool valid = kb + (int)threadIdx.x < it.lim;
    int st = ST_IDLE, p = 0, nk = 0, nf = 0;
    bool win = false;
    unsigned long long key = 0;
    if (valid) {
        st = c.status[slot];
        p = win_point(wv, slot);
        key = key_hi | (unsigned)p;
        if (st == ST_OK) {
            nk = c.nk[slot];
            nf = c.nf[slot];
            win = true;
            for (int j = 0; j < nk && win; ++j) {
                const int t = c.cav[(size_t)slot * c.kcap + (j)];
                win = *kres_p<D>(c, t) == key && *nres_p<D>(c, t) > key;
            }
            for (int j = 0; j < nf && win; ++j) {
                const int fe = c.fac[(size_t)slot * c.fcap + (j)];
                const int nb = *a_ptr<D>(c, fe >> 2, fe & 3);
                if (nb >= 0) win = *kres_p<D>(c, nb >> 2) > key;
            }
        }
    }
    const int base = alloc_slots(c, win, nf, p);
    if (base < 0) win = false;
    if (win) {
        const int pv = c.M + p;
        const typename VT<D>::F pf = VT<D>::ld(c.X32, pv);
        // phase 1: the new simplex of facet j takes slot base + j; its face 0 faces the outside
        // neighbor, which is patched to point back; the cavity side of the facet is overwritten
        // with the marker -2 - (base + j) so that ridge rotations find it without a search
        for (int j = 0; j < nf; ++j) {
            const int fe = c.fac[(size_t)slot * c.fcap + (j)];
            const int t = fe >> 2, i = fe & 3;
            int* slot_ptr = a_ptr<D>(c, t, i);
            const int outside = *slot_ptr;
            const int N = base + j;
            *a_ptr<D>(c, N, 0) = outside;
            if (outside >= 0) *a_ptr<D>(c, outside >> 2, outside & 3) = N << 2;
            *slot_ptr = -2 - N;
        }
        // phase 2: internal faces.  Face k (k >= 1) of the new simplex contains p and the ridge
        // nv \ {p, nv[k]}; rotating around that ridge through the cavity reaches the other
        // boundary facet containing it, whose marker names the neighbor.
        for (int j = 0; j < nf; ++j) {
            const int fe = c.fac[(size_t)slot * c.fcap + (j)];
            const int t = fe >> 2, i = fe & 3;
            const int4 tvt = ld_v<D>(c, t);
            const int4 tat = ld_a<D>(c, t);
            const int4 nv = replace_first<D>(tvt, i, pv);
            const int N = base + j;
            for (int k = 1; k < M; ++k) {
                const int cv = comp(nv, k);
                int4 TV = tvt, TA = tat;
                int e = index_of<D>(TV, cv);
                int y = comp(tvt, i);
                int nbe = comp(TA, e);
                for (int guard = 0; nbe >= 0; ++guard) {
                    if (guard > 4096) { atomicOr(&c.st->err, ERR_WALK); break; }
                    const int T2 = nbe >> 2;
                    const int4 TV2 = ld_v<D>(c, T2);
                    const int w = comp(TV2, nbe & 3);
                    e = index_of<D>(TV2, y);
                    y = w;
                    TV = TV2;
                    TA = ld_a<D>(c, T2);
                    nbe = comp(TA, e);
                }
                if (nbe > -2) { atomicOr(&c.st->err, ERR_WALK); continue; }
                const int N2 = -2 - nbe;
                const int4 nv2 = replace_first<D>(TV, e, pv);
                int kk = 1;
                for (int m = 1; m < M; ++m) {
                    const int x = comp(nv2, m);
                    bool in_ridge = false;
                    for (int r = 1; r < M; ++r)
                        if (r != k && comp(nv, r) == x) in_ridge = true;
                    if (!in_ridge) { kk = m; break; }
                }
                *a_ptr<D>(c, N, k) = (N2 << 2) | kk;
            }
            st_v<D>(c, N, nv);
            typename VT<D>::F ox[4], nx[4];
            load_tx<D>(c, t, ox);
            star_record<D>(ox, i, pf, nx);
            store_tx<D>(c, N, nx);
        }
        // phase 3: killed simplices die with a forward pointer into the star
        for (int j = 0; j < nk; ++j) kill<D>(c, c.cav[(size_t)slot * c.kcap + (j)], base);
        for (int j = 0; j < nf; ++j) *v_ptr<D>(c, c.fac[(size_t)slot * c.fcap + (j)] >> 2, 1) = base + j;
        lc.killed += nk;
        lc.created += nf;
        lc.wins += 1;
        lc.maxk = nk > (int)lc.maxk ? nk : lc.maxk;
        lc.maxf = nf > (int)lc.maxf ? nf : lc.maxf;
    }
    // cavities too large for the per-point lists are re-run alone after the round
    const bool big = st == ST_OVERFLOW && !c.solo;
    bool again = (st == ST_OK && !win) || st == ST_DEFER || st == ST_LOST;
    {
        const unsigned long long bm = __ballot(big);
        int bpos = 0;
        if (__lane_id() == 0 && bm) bpos = atomicAdd(&c.st->n_big, __popcll(bm));
        bpos = __shfl(bpos, 0, 64) + __popcll(bm & ((1ull << __lane_id()) - 1ull));
        if (big) {
            if (bpos < c.big_cap) c.big[bpos] = p;
            else again = true;             // list full: retry in the next window
        }
    }
    set_outcome(c, st != ST_IDLE, p, win, nf - nk, !win && again, !win && !again && st == ST_DUP);
    // next window: losers (valid cavity, lost a reservation) and deferred points
    const unsigned long long mask = __ballot(again);
    const int cnt = __popcll(mask);
    if (again) lc.pend++;
    int pos0 = 0;
    if (__lane_id() == 0 && cnt) pos0 = atomicAdd(&c.st->n_next_pending, cnt);
    pos0 = __shfl(pos0, 0, 64);
    if (again) {
        const int pos = pos0 + __popcll(mask & ((1ull << __lane_id()) - 1ull));
        if (pos < c.pend_cap) c.pend_next[pos] = p;
        else atomicOr(&c.st->err, ERR_PENDING);
    }
    }
    flush_counters(c.cnt, lc);
}

// ---------------------------------------------------------------- k_exact (tier 3)
template <int S>
__global__ void __launch_bounds__(64) k_exact(const QEntry* q, int nq, const double* X64, double* ws, int ws_cap,
                                              CEntry* cache, uint32_t cmask, RoundState* st, Counters* cnt, int qcap) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int nth = gridDim.x * blockDim.x;
    if (nq < 0) nq = min(st->n_q, qcap);     // queue fill of the round, read on the device
    for (int i = tid; i < nq; i += nth) {
        Ws w{ws + (size_t)tid * ws_cap, ws_cap, 0, 0};
        const QEntry e = q[i];
        const double* P[5];
        for (int j = 0; j < pred_arity(e.type); ++j) P[j] = X64 + (size_t)S * e.id[j];
        int s;
        switch (e.type) {
            case PRED_ORIENT2: s = ex_orient2d(w, P[0], P[1], P[2]); break;
            case PRED_INCIRCLE: s = ex_incircle(w, P[0], P[1], P[2], P[3]); break;
            case PRED_ORIENT3: s = ex_orient3d(w, P[0], P[1], P[2], P[3]); break;
            default: s = ex_insphere(w, P[0], P[1], P[2], P[3], P[4]); break;
        }
        if (w.ovf || s == SIGN_UNDECIDED) { atomicOr(&st->err, ERR_PRECISION); continue; }
        cache_insert(cache, cmask, e.type, e.id, s, &st->err);
        atomicAdd(&cnt[blockIdx.x & (kCntStripes - 1)].exact_evaluated, 1ull);
    }
}

// ---------------------------------------------------------------- export / check
// compacted export: vertex ids mapped to the reference id space, neighbors to export indices
template <int D>
__global__ void k_export(Ctx c, const int* pos, int n, const int* vmap, int32_t* simp, int32_t* nbr) {
    constexpr int M = D + 1;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int4 v = ld_v<D>(c, i);
    if (v.x < 0) return;
    const int o = pos[i];
    const int4 a = ld_a<D>(c, i);
    for (int k = 0; k < M; ++k) {
        if (simp) simp[(size_t)o * M + k] = vmap[comp(v, k)];
        if (nbr) {
            const int nb = comp(a, k);
            nbr[(size_t)o * M + k] = nb < 0 ? -1 : pos[nb >> 2];
        }
    }
}

// exact local Delaunay + adjacency symmetry + orientation, one thread per simplex
template <int D>
__global__ void __launch_bounds__(256) k_check(Ctx c, int n, unsigned long long* bad, int* unresolved) {
    using V = VT<D>;
    constexpr int M = D + 1;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    LocalCnt lc = {};
    unsigned nbad = 0, nun = 0;
    if (t < n) {
        const int4 v = ld_v<D>(c, t);
        if (v.x >= 0) {
            const int4 a = ld_a<D>(c, t);
            typename V::F x[4], x2[4];
            load_tx<D>(c, t, x);
            bool coords_ok = true;                   // the record matches the vertex table
            for (int k = 0; k < M; ++k) {
                const typename V::F g = V::ld(c.X32, comp(v, k));
                coords_ok &= g.x == x[k].x && g.y == x[k].y;
            }
            if (!coords_ok) nbad++;
            int s = orient_at<D>(c, x, t, 0, v.x, x[0], lc);
            if (s == SIGN_UNDECIDED) nun++;
            else if (s <= 0) nbad++;
            for (int f = 0; f < M; ++f) {
                const int nb = comp(a, f);
                if (nb < 0) continue;
                const int n2 = nb >> 2, f2 = nb & 3;
                const int4 v2 = ld_v<D>(c, n2);
                if (v2.x < 0 || *a_ptr<D>(c, n2, f2) != ((t << 2) | f)) { nbad++; continue; }
                const int w = comp(v2, f2);
                load_tx<D>(c, n2, x2);
                const int cs = conflict_at<D>(c, x, t, w, x2[f2], lc);
                if (cs == SIGN_UNDECIDED) nun++;
                else if (cs > 0) nbad++;
            }
        }
    }
    const unsigned long long sb = wave_sum(nbad), su = wave_sum(nun);
    if (__lane_id() == 0) {
        if (sb) atomicAdd(bad, sb);
        if (su) atomicAdd(unresolved, (int)su);
    }
    flush_counters(c.cnt, lc);
}

// check_delaunay with the reference's own semantics (src/delaunay_tree.rs:512-541 / 642-671):
// for every live simplex made only of inserted points, the vertices strictly inside its
// circumsphere that are not its own, counted with exact signs (inside = in-sphere sign times
// orientation sign > 0, so a simplex whose orientation a moved vertex flipped still counts
// what lies inside its sphere).  Vertex w < M is a super vertex; w_weight[w] is how many
// reference vertices share its coordinates (itself plus its ghost copies).  One block per
// simplex at a time, the block's threads over the vertices: O(T * V), for small trees.
template <int D>
__global__ void __launch_bounds__(256) k_check_pairs(Ctx c, int n, int nv, int4 w_weight, unsigned long long* bad,
                                                     int* unresolved) {
    using V = VT<D>;
    constexpr int M = D + 1;
    __shared__ int s_o;
    LocalCnt lc = {};
    unsigned long long nbad = 0;
    unsigned nun = 0;
    for (int t = blockIdx.x; t < n; t += gridDim.x) {
        const int4 v = ld_v<D>(c, t);
        if (v.x < 0) continue;
        bool real = true;
        for (int k = 0; k < M; ++k) real &= comp(v, k) >= c.M;
        if (!real) continue;
        typename V::F x[4];
        load_tx<D>(c, t, x);
        // the orientation once per simplex (one queue entry when it needs the exact tier)
        if (threadIdx.x == 0) s_o = orient_at<D>(c, x, t, 0, v.x, x[0], lc);
        __syncthreads();
        const int o = s_o;
        __syncthreads();
        if (o == SIGN_UNDECIDED) {
            if (threadIdx.x == 0) nun++;
            continue;
        }
        if (o == 0) continue;                 // flat: no sphere, nothing inside
        for (int w = threadIdx.x; w < nv; w += blockDim.x) {
            bool own = false;
            for (int k = 0; k < M; ++k) own |= comp(v, k) == w;
            if (own) continue;
            const int s = conflict_at<D>(c, x, t, w, V::ld(c.X32, w), lc);
            if (s == SIGN_UNDECIDED) nun++;
            else if (s * o > 0) nbad += w < M ? (unsigned long long)comp(w_weight, w) : 1ull;
        }
    }
    const unsigned long long sb = wave_sum(nbad), su = wave_sum(nun);
    if (__lane_id() == 0) {
        if (sb) atomicAdd(bad, sb);
        if (su) atomicAdd(unresolved, (int)su);
    }
    flush_counters(c.cnt, lc);
}

// Round rule of the lane-group kernels.  Winners hold every kill; then either (shared facets)
// a winner's cavity may touch a higher-priority winner's when the sphere test of the commit
// admits it, the links across such facets being set by k_fix_shared, or (stricter, with record
// reuse, or with no shared-facet list) the cavities of winners are also pairwise non-adjacent,
// which the commit checks through outside reservations (nres).
__device__ __forceinline__ bool shared_facets(const Ctx& c) { return !c.reuse && c.fix != nullptr; }

// ---------------------------------------------------------------- k_fix_shared
// Links across the facets shared by two winners of a round (a lower-priority winner W next to a
// higher-priority cavity, admitted by the commit's sphere test).  W built its new simplex N on
// the facet without linking it, and listed (N, A << 2 | face) with A its killed simplex on the
// facet.  Whatever is across that face of A now is N's neighbor: the higher-priority winner's
// new simplex on the facet (its outside patch landed in A, which only it writes there), or, when
// that point did not win, the old simplex across the facet, which nobody patched.  One thread per
// entry; the two ends are set here, after every commit of the round has finished.
template <int D>
__device__ __forceinline__ void fix_shared_pass(const Ctx& c, int first, int stride) {
    const int n = min(c.st->n_fix, c.fix_cap);
    for (int e = first; e < n; e += stride) {
        const int2 f = c.fix[e];
        if (f.x < 0) continue;                  // reserved by a point that then did not win
        const int X = *a_ptr<D>(c, f.y >> 2, f.y & 3);
        *a_ptr<D>(c, f.x, 0) = X;
        if (X >= 0) *a_ptr<D>(c, X >> 2, X & 3) = f.x << 2;
    }
}
template <int D>
__global__ void __launch_bounds__(256) k_fix_shared(Ctx c) {
    fix_shared_pass<D>(c, blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x);
}

// ---------------------------------------------------------------- locate query
// locate + get_new_simplices (src/delaunay_tree.rs:33-123) for one point q that is not a
// vertex of the tree, against the current tree, changing nothing: walk from the hint grid (or
// from record `start` when it is >= 0) to
// the simplex containing q, then search the conflict region breadth-first.  Signs come from the
// same cascade as the construction, the exact tier evaluated inline on q's f64 coordinates
// (no queue: q has no vertex id), in the workspace ws.  One thread.
//   out_k[0..nk): killed records, out_f[0..nf): boundary facets (t << 2 | face), res[0..2]:
//   nk, nf, status (0 ok, 1 q on the sphere of its own simplex: a vertex, 2 lists full,
//   3 outside the super simplex or a walk bound, 4 an exact sign left the f64 range)
template <int D> struct QueryPoint {
    typename VT<D>::F f;   // fl32(q - center)
    double x[3];           // q
};
template <int D>
__device__ int q_sign(const Ctx& c, const typename VT<D>::F* a, const int* ids, int k, const QueryPoint<D>& q, bool sphere,
                      double* ws, int ws_cap, bool& ovf) {
    constexpr int S = VT<D>::S;
    if (c.filters & 1) {
        int s;
        if (sphere) {
            if constexpr (D == 3) s = t1_insphere(a[0], a[1], a[2], a[3], q.f);
            else s = t1_incircle(a[0], a[1], a[2], q.f);
        } else {
            typename VT<D>::F b[4];
            for (int i = 0; i < D + 1; ++i) b[i] = i == k ? q.f : a[i];
            if constexpr (D == 3) s = t1_orient3d(b[0], b[1], b[2], b[3]);
            else s = t1_orient2d(b[0], b[1], b[2]);
        }
        if (s != SIGN_UNDECIDED) return s;
    }
    const double* P[5];
    for (int i = 0; i < D + 1; ++i) P[i] = c.X64 + (size_t)S * ids[i];
    if (!sphere) P[k] = q.x;
    int s = SIGN_UNDECIDED;
    if (c.filters & 2) {
        if (sphere) s = D == 3 ? t2_insphere(P[0], P[1], P[2], P[3], q.x) : t2_incircle(P[0], P[1], P[2], q.x);
        else s = D == 3 ? t2_orient3d(P[0], P[1], P[2], P[3]) : t2_orient2d(P[0], P[1], P[2]);
        if (s != SIGN_UNDECIDED) return s;
    }
    Ws w{ws, ws_cap, 0, 0};
    if (sphere) s = D == 3 ? ex_insphere(w, P[0], P[1], P[2], P[3], q.x) : ex_incircle(w, P[0], P[1], P[2], q.x);
    else s = D == 3 ? ex_orient3d(w, P[0], P[1], P[2], P[3]) : ex_orient2d(w, P[0], P[1], P[2]);
    if (w.ovf) ovf = true;
    return s;
}

template <int D>
__global__ void k_locate_query(Ctx c, QueryPoint<D> q, int start, double* ws, int ws_cap, int* out_k, int* out_f,
                               int cap, int* res) {
    using V = VT<D>;
    constexpr int M = D + 1;
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    bool ovf = false;
    int status = 0, nk = 0, nf = 0;
    int t = start >= 0 ? start : hint_lookup<D>(hint_view(c, 0), q.f);
    int4 tt = ld_v<D>(c, t);
    for (int hops = 0; tt.x < 0 && status == 0; ++hops) {
        if (hops > kChainLimit) status = 3;
        t = tt.y;
        tt = ld_v<D>(c, t);
    }
    typename V::F a[4];
    int came = -1;
    for (int step = 0; status == 0; ++step) {
        if (step > kWalkLimit) { status = 3; break; }
        load_tx<D>(c, t, a);
        const int4 v = ld_v<D>(c, t), aa = ld_a<D>(c, t);
        const int ids[4] = {v.x, v.y, v.z, v.w};
        int next = -1;
        for (int k = 0; k < M && next < 0 && status == 0; ++k) {
            const int nb = comp(aa, k);
            if (nb >= 0 && (nb >> 2) == came) continue;
            const int s = q_sign<D>(c, a, ids, k, q, false, ws, ws_cap, ovf);
            if (ovf) status = 4;
            else if (s < 0) {
                if (nb < 0) status = 3;
                else next = nb >> 2;
            }
        }
        if (next < 0) break;
        came = t;
        t = next;
    }
    if (status == 0) {
        const int4 v = ld_v<D>(c, t);
        const int ids[4] = {v.x, v.y, v.z, v.w};
        const int s0 = q_sign<D>(c, a, ids, 0, q, true, ws, ws_cap, ovf);
        if (ovf) status = 4;
        else if (s0 <= 0) status = 1;
    }
    if (status == 0) {
        out_k[nk++] = t;
        for (int i = 0; i < nk && status == 0; ++i) {
            const int ct = out_k[i];
            const int4 aa = ld_a<D>(c, ct);
            for (int f = 0; f < M && status == 0; ++f) {
                const int nb = comp(aa, f);
                bool fac = nb < 0;
                if (!fac) {
                    const int n = nb >> 2;
                    bool seen = false;
                    for (int j = 0; j < nk && !seen; ++j) seen = out_k[j] == n;
                    if (seen) continue;
                    typename V::F an[4];
                    load_tx<D>(c, n, an);
                    const int4 vn = ld_v<D>(c, n);
                    const int idn[4] = {vn.x, vn.y, vn.z, vn.w};
                    const int s = q_sign<D>(c, an, idn, 0, q, true, ws, ws_cap, ovf);
                    if (ovf) { status = 4; break; }
                    if (s > 0) {
                        if (nk == cap) { status = 2; break; }
                        out_k[nk++] = n;
                        continue;
                    }
                    fac = true;
                }
                if (nf == cap) { status = 2; break; }
                out_f[nf++] = (ct << 2) | f;
            }
        }
    }
    res[0] = nk;
    res[1] = nf;
    res[2] = status;
}

// rows of a located cavity in the export's numbering: killed records -> export positions;
// facet j -> (killed position, outside position or -1, the facet's D vertex ids in the
// reference id space, in the order replace_first gives the new simplex after q)
template <int D>
__global__ void k_locate_rows(Ctx c, const int* pos, const int* vmap, const int* in_k, int nk, const int* in_f, int nf,
                              int32_t* out_k, int32_t* out_f) {
    constexpr int M = D + 1;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nk) out_k[i] = pos[in_k[i]];
    if (i < nf) {
        const int fe = in_f[i];
        const int t = fe >> 2, face = fe & 3;
        const int nb = *a_ptr<D>(c, t, face);
        const int4 nv = replace_first<D>(ld_v<D>(c, t), face, -1);
        int32_t* row = out_f + (size_t)i * (M + 1);
        row[0] = pos[t];
        row[1] = nb < 0 ? -1 : pos[nb >> 2];
        for (int k = 1; k < M; ++k) row[1 + k] = vmap[comp(nv, k)];
    }
}

// Vertex vi moved: rewrite its fp32 coordinates f in every live record that holds it.
template <int D>
__global__ void __launch_bounds__(256) k_patch_vertex(Ctx c, int n, int vi, typename VT<D>::F f) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int4 v = ld_v<D>(c, t);
    if (v.x < 0) return;
    const int k = index_of<D>(v, vi);
    if (k < 0) return;
    typename VT<D>::F x[4];
    load_tx<D>(c, t, x);
#pragma unroll
    for (int j = 0; j < D + 1; ++j)
        if (j == k) x[j] = f;
    store_tx<D>(c, t, x);
}

// The live record at export position `target` (pos: exclusive scan of the live flags) swaps
// its vertex entries i and j and their coordinates; its adjacency is left as it was.
template <int D>
__global__ void __launch_bounds__(256) k_swap_in_record(Ctx c, int n, const int* pos, int target, int i, int j) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n || pos[t] != target) return;
    int4 v = ld_v<D>(c, t);
    if (v.x < 0) return;
    typename VT<D>::F x[4], y[4];
    load_tx<D>(c, t, x);
    const int vi = comp(v, i), vj = comp(v, j);
    set_comp(v, i, vj);
    set_comp(v, j, vi);
    typename VT<D>::F xi = x[0], xj = x[0];
#pragma unroll
    for (int k = 0; k < D + 1; ++k) {
        if (k == i) xi = x[k];
        if (k == j) xj = x[k];
    }
#pragma unroll
    for (int k = 0; k < D + 1; ++k) y[k] = k == i ? xj : (k == j ? xi : x[k]);
    st_v<D>(c, t, v);
    store_tx<D>(c, t, y);
}

}  // namespace vo
