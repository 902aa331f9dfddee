/*
 * voronoids.h -- C ABI of libvoronoids_hip.so, the MI355X (gfx950) Delaunay builder.
 *
 * This is the boundary a foreign-language facade binds (ctypes, a Rust extern "C" block,
 * cgo).  Plain pointers and sizes only; every function returns 0 or a negative VO_ERR_*
 * code and never aborts; vo_last_error() gives a thread-local message for the last failure.
 * Ownership: the caller owns every host buffer (copied during the call, never retained);
 * the library owns all device memory behind a vo_tri handle.  A vo_tri is single-writer
 * (the reference mutates through &mut self); distinct handles may be used from different
 * threads.
 *
 * Ids in exported views follow the reference: super vertices 0..d, ghost vertices d+1..2d+1
 * (never referenced here: the hull is -1), input point i is vertex 2(d+1)+i.
 */
#ifndef VORONOIDS_H
#define VORONOIDS_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    VO_OK = 0,
    VO_ERR_INVALID = -1,    /* bad argument (dimension, null pointer, NaN/inf coordinate) */
    VO_ERR_OOM = -2,        /* device or host allocation failed */
    VO_ERR_HIP = -3,        /* a HIP runtime call failed */
    VO_ERR_DUPLICATE = -4,  /* some inserted points coincide with existing vertices (skipped) */
    VO_ERR_OVERFLOW = -5,   /* a cavity exceeded the per-point buffers */
    VO_ERR_PRECISION = -6,  /* an exact evaluation left the f64 exponent range */
    VO_ERR_INTERNAL = -7    /* a walk or chain exceeded its bound: a bug */
};

typedef struct vo_tri vo_tri;

typedef struct {
    /* rounds */
    int64_t rounds, points_inserted, points_duplicate, windows_total, winners_total;
    /* predicate calls resolved per tier, index 0 = orient, 1 = in-sphere/in-circle */
    int64_t tier1[2], tier2[2], tier3[2], deferred[2];
    int64_t exact_evaluated;
    /* events */
    int64_t walk_steps, chain_hops, killed, created, max_killed, max_facets;
    /* device time of the last vo_tri_run, ms (hipEvent) and per kernel kind */
    double ms_run, ms_locate_cavity, ms_commit, ms_exact, ms_other;
    int64_t launches_locate_cavity, launches_commit;
    /* points whose cavity outgrew the per-point LDS lists and were inserted alone */
    int64_t solo_runs;
    /* points found to coincide with an existing vertex, counted on the device */
    int64_t duplicates_detected;
    /* compactions of the simplex records (live ones moved to a dense array in spatial order)
     * and their device time, ms (included in ms_run) */
    int64_t compactions;
    double ms_compact;
    /* simplex records in use, live and dead: with record reuse on (vo_tri_config "reuse") about
     * the live count, since new simplices take their winner's killed records; append-only (the
     * default below 2^29 records) about 4x that in 3D */
    int64_t records;
} vo_stats;

/* DelaunayTree::<3,4>::new / ::<2,3>::new (src/delaunay_tree.rs:389-510, :544-640):
 * bounding sphere of pts (src/geometry.rs:99-141), radius x10, super simplex; no point is
 * inserted.  pts is row-major n x dim f64. */
int vo_tri_create(int hip_device, int dim, const double* pts, int64_t n, vo_tri** out);
void vo_tri_destroy(vo_tri* t);
/* the same handle made into DelaunayTree::new(pts) again (a new point set of the same dimension;
 * no counterpart in the reference, which builds a new tree): its device buffers are kept, so a
 * caller triangulating many sets one after another allocates nothing per set */
int vo_tri_reset(vo_tri* t, const double* pts, int64_t n);

/* add_points_to_tree (src/delaunay_tree.rs:336-386) / the insert_point loop of
 * voronoids.delaunay (src/lib.rs:104-125): inserts n points in a random priority order drawn
 * from seed.  Equivalent to vo_tri_stage followed by vo_tri_run. */
int vo_tri_insert(vo_tri* t, const double* pts, int64_t n, uint64_t seed);
/* stage: copy n points to the device in priority order (no insertion) */
int vo_tri_stage(vo_tri* t, const double* pts, int64_t n, uint64_t seed);
/* run: insert every staged point (device-resident input; timed by ms_run) */
int vo_tri_run(vo_tri* t);

/* n_vertices counts super vertices and inserted points; n_simplices counts live simplices
 * (no ghosts); max_simplex_id = simplices ever created - 1 + (dim + 1), the reference's
 * max_simplex_id convention (creation order differs from the reference's). */
int vo_tri_counts(const vo_tri* t, int64_t* n_vertices, int64_t* n_simplices, int64_t* max_simplex_id);

/* live simplices: simplices[n_s*(dim+1)] vertex ids (reference id space, positively
 * oriented), neighbors[n_s*(dim+1)] = index (in this export) of the simplex across the face
 * opposite each vertex, or -1 on the hull of the super simplex.  Either pointer may be NULL. */
int vo_tri_export(const vo_tri* t, int32_t* simplices, int32_t* neighbors);

/* canonical output: T = simplices made only of input points, as input indices, rows sorted,
 * lexicographic; E = input-point pairs co-occurring in any live simplex, sorted, unique.
 * Arrays are malloc'd by the library; release with vo_free. */
int vo_tri_canonical(const vo_tri* t, int32_t** T, int64_t* n_T, int32_t** E, int64_t* n_E);
void vo_free(void* p);

/* check_delaunay (src/delaunay_tree.rs:512-541, :642-671) as an exact O(T) device test:
 * every face's opposite vertex outside the circumsphere, adjacency symmetric, every simplex
 * positively oriented.  Writes the number of violations. */
int vo_tri_check(vo_tri* t, int64_t* n_violations);

/* locate + get_new_simplices (src/delaunay_tree.rs:33-75, :77-123) for one point p[dim] that is
 * not yet in the tree, against the tree as it is (nothing changes): the simplices whose
 * circumsphere strictly contains p -- the cavity, as positions in vo_tri_export's numbering,
 * sorted ascending as the reference sorts them -- and the cavity's boundary facets, one row of
 * dim + 2 ints per facet: (cavity simplex, the simplex across the facet or -1 on the super-simplex
 * hull, then the facet's dim vertex ids), so that inserting p as vertex id creates one simplex
 * [id, facet vertices] per row (positively oriented in that order).  Exact signs.  n_killed and
 * n_facets always receive the sizes; the rows are written only when both fit (else
 * VO_ERR_OVERFLOW).  VO_ERR_DUPLICATE when p coincides with a vertex (the reference panics there,
 * :47-54), VO_ERR_INVALID when p lies outside the super simplex. */
int vo_tri_locate(vo_tri* t, const double* p, int32_t* killed, int64_t killed_cap, int64_t* n_killed,
                  int32_t* facets, int64_t facets_cap, int64_t* n_facets);

/* check_delaunay with the reference's own semantics (src/delaunay_tree.rs:512-541, :642-671):
 * the number of (simplex made only of input points, vertex not in it) pairs whose vertex lies
 * strictly inside the simplex's circumsphere, super and ghost vertices included (a ghost counts
 * as the super vertex it copies); the reference returns true iff this is 0.  Exact signs.  It
 * is O(simplices x vertices): VO_ERR_INVALID beyond 2^36 pairs (use vo_tri_check there). */
int vo_tri_check_pairs(vo_tri* t, int64_t* n_pairs);

/* The reference's tree exposes its state as pub fields (src/delaunay_tree.rs:24-30), so a
 * caller can move a vertex (Vertex::coordinates, :18-22) or reorder a simplex's vertices
 * (Simplex::vertices, :10-16) and then ask check_delaunay about the result.  These two do the
 * same to the device state; nothing is re-triangulated.
 *   vo_tri_set_vertex: input point with reference vertex id `id` (2(dim+1) + input index) moves
 *     to coords[dim]; records = 1 also rewrites the fp32 copies held in its simplices' records
 *     (records = 0 leaves them stale, which vo_tri_check reports as record mismatches).
 *   vo_tri_swap_simplex_vertices: the live simplex at position `simplex` of vo_tri_export's
 *     order swaps its vertex entries i and j (with their coordinates; its neighbor entries stay). */
int vo_tri_set_vertex(vo_tri* t, int64_t id, const double* coords, int records);
int vo_tri_swap_simplex_vertices(vo_tri* t, int64_t simplex, int i, int j);

int vo_tri_stats(const vo_tri* t, vo_stats* out);

/* tuning and test knobs (no counterpart in the reference): "filters" = mask of the FP32 (1)
 * and FP64-filter (2) predicate tiers, 3 by default; 0 sends every predicate to the exact
 * expansion tier.  "group_lanes" = lanes cooperating on one window point in the round
 * kernels (1, 4, 8, 16, 32; 0 = chosen per round: the largest size with window x size <=
 * "lanes_target", default 2^21 in 3D and 2^16 in 2D, and at least 4; 1 = one thread per point;
 * "commit_lanes_target", default 0 = "lanes_target", is the commit kernel's own target).  "window_gamma_milli" = initial window
 * size per inserted point x 1000.  "local_order" = 0 (default: the seeded random priority order) or k > 0
 * (the same order with consecutive chunks of position/k points sorted along the Morton curve; set before
 * vo_tri_stage).  "big_window" (default 16384 in 3D, 65536 in 2D; 0 = off), "big_frac_lo_milli" /
 * "big_frac_hi_milli" (default 650 / 800 in 3D, 550 / 800 in 2D): rounds whose window is at least big_window steer the window
 * size so that this fraction of it wins (smaller windows aim at "small_frac_lo_milli" .. "small_frac_hi_milli", default 350 / 600).  "compact_ratio_milli" (default
 * 0 = never; measured no faster at n = 4M: every visit misses the caches either way): between rounds, when the record array holds more than this many records (x 1000) per live
 * simplex, the live ones move to a dense array in spatial order; "compact_min" (default 2^20): not before this
 * many records exist.  "xcd_regions" (0 default / 1): each round's window is sorted into 8 spatial regions and
 * block b of the round kernels serves region b % 8, so that each of the 8 XCDs (blocks are dispatched to them
 * in turn) caches one region's records; measured no faster at n = 4M, kept as an option.  "pipelined" (1
 * default / 0): rounds planned on the device and submitted in chunks, the host taking over only when the
 * planner halts; 0 plans and waits for every round on the host (compaction, XCD regions and VO_TRACE always
 * do); both give the same rounds and result.  "chunk_rounds" (default 4, 1..64): rounds per submitted chunk.
 * "round_timing" (0 default / 1): HIP events around every round's cavity and commit kernels, reported in
 * vo_stats ms_locate_cavity / ms_commit (the host loop always times; the events cost about 3% at n = 4M).
 * "abort_early" (1 default / 0): the lane-group cavity search reserves each conflicting simplex as it
 * finds it and stops as soon as a reservation shows a higher-priority point of the round (that point
 * cannot win; it retries next round).  Same triangulation either way; with 1 the set of winners of a
 * round depends on timing, so round counts and simplex ids may differ from run to run (0: the same
 * rounds, windows and ids on every run; measured 157 -> 147 ms at 3D n = 4M).  "min_window" (default 64,
 * 1..65536): least points per round (the window is max(min_window, gamma x inserted)).  "small_run" (default
 * 64): a vo_tri_run with at most this many points to insert plans its rounds on the host (one point: 0.16 ms
 * in 2D instead of 0.22 ms through the device planner, which submits at least two chunks of rounds).
 * "walk_budget" (default per the autotuner's table, else 0 = off): in windows of 64 or more points, a
 * point whose walk has moved this many times stops there and resumes next round, so that one long
 * walk does not hold a whole round.  Like abort_early, a budget makes the winners of a round depend
 * on timing (a walk's length depends on which concurrent write a hint cell kept); 0 with
 * abort_early 0 gives the same rounds on every run.
 * "boot_window" (default 0 in 3D, 256 in 2D; 0..65536, 0 = off): a vo_tri_run whose first windows hold at
 * most this many points runs those rounds in one workgroup (one launch, barriers between the
 * phases, the records in one CU's caches) until the window outgrows it; same rounds, windows and
 * result (measured 2D n = 10M 35.4 -> 34.9 ms; 3D n = 4M 95 -> 109 ms, so off there).  Those rounds
 * count in vo_stats rounds, not in the launch counts, and their time only in ms_run.
 * "shared_facets" (1 default / 0): a window point whose cavity touches (shares a facet with) a
 * higher-priority point's may still win when it lies outside the sphere of the simplex that point
 * builds on the shared facet (inserting that point first would leave its cavity unchanged); the links
 * across such facets are set after the commit.  0: winners' cavities are also pairwise non-adjacent
 * (same triangulation; measured 3D n = 4M 149 -> 117 ms and 2D n = 10M 99 -> 74 ms with 1).
 * "reuse" (-1 default = auto / 0 / 1): new simplices take their winner's killed records first, so that
 * records stay about as many as live simplices (append-only: about 4x in 3D); auto turns it on only when
 * the append-only records of a run would pass 2^29 (with the default allocation chunks: from about
 * 11.6 M points in 3D, 38 M in 2D), since it is slower (3D n = 4M +5%, 2D n = 10M +15%:
 * a winner's new simplices then scatter over its killed records' lines); with reuse the rounds keep
 * winners' cavities non-adjacent.
 * "alloc_stripes" (1 default / 0): device-planned rounds without reuse take new simplex records from 64
 * chunks the planning launch reserves per round (one add to the record count per round instead of a
 * returning atomic on it per commit wavefront, which serialize); the chunks' unused tails stay unused
 * records (about 36 per point in 3D, 8.5 in 2D, against 27 / 6 live).  Measured 3D n = 4M
 * 88.8 -> 83.1 ms, 2D n = 10M 33.2 -> 32.7 ms.
 * "outcome_flags" (1 default / 0): in device-planned rounds with shared facets the commit writes each
 * window slot's outcome at the slot (the point when it goes on to the next window) instead of
 * appending it to the pending list with a returning atomic per wavefront; the next planning launch
 * compacts the flags and takes the count from the statistics.  Same rounds and result; measured
 * 3D n = 4M 83.0 -> 80.0 ms, 2D n = 10M 33.2 -> 31.2 ms. */
int vo_tri_config(vo_tri* t, const char* key, int64_t value);

/* The precision tiers one predicate at a time (no counterpart in the reference, whose in_sphere
 * (src/geometry.rs:91-97) is a plain f64 distance test): signs of n instances of one predicate through
 * the construction's cascade -- FP32 on fl32(x - center), FP64 filter, FP64 expansions -- and the tier
 * that decided each (1, 2, 3).  type: 1 orient2d(a,b,c), 2 incircle(a,b,c,d), 3 orient3d(a,b,c,d),
 * 4 insphere(a,b,c,d,e); pts = n x arity x dim f64.  Signs: orient = sign det[b-a; c-a(; d-a)];
 * incircle / insphere > 0 iff the last point is inside the circle / sphere of the others when they are
 * positively oriented.  A sign of 2 means the expansions left the f64 range. */
int vo_predicates(int hip_device, int type, const double* pts, int64_t n, const double* center, int32_t* signs,
                  int32_t* tiers);

/* vo_predicates without any device allocation: no counterpart in the reference; for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments, signs and tiers as vo_predicates, but n is at most 65,536 per
 * call (VO_ERR_INVALID beyond): the instances, signs, tiers and expansion workspaces are staged in
 * one static device array of about 16.5 MiB that each device holds from the first call until the
 * process ends.  Calls that overlap on one device share that array and race: serialize them. */
int vo_predicates_noalloc(int hip_device, int type, const double* pts, int64_t n, const double* center,
                          int32_t* signs, int32_t* tiers);

/* input vertex coordinates in the reference id space (for PyVertex.point): (2(dim+1) + n) x dim */
int vo_tri_vertices(const vo_tri* t, double* coords);

const char* vo_last_error(void);
int vo_version(void);

/* geometry::bounding_sphere (src/geometry.rs:99-141), host only: center[dim] and radius of the sphere
 * DelaunayTree::new builds its super simplex around (per-axis min/max midpoint, then the radius x 1.5
 * for each point, in input order, not strictly inside).  n >= 1. */
int vo_bounding_sphere(int dim, const double* pts, int64_t n, double* center, double* radius);

/* geometry::circumsphere (src/geometry.rs:58-88, with circumsphere_2d :2-22 and circumsphere_3d
 * :24-56), host only, bit for bit: the 2D perpendicular-bisector slope formula (an axis-parallel
 * first or second edge gives the formula's IEEE result, e.g. NaN for a horizontal one) and the
 * 3D LU solve in nalgebra 0.32's order; verts = (dim + 1) x dim, vertex 0 first.  A 3D simplex
 * with a zero pivot returns VO_ERR_INVALID with NaN outputs (the reference panics there). */
int vo_circumsphere(int dim, const double* verts, double* center, double* radius);
/* the same for n simplices given as vertex ids into coords (n_coords x dim); degenerate ones get
 * NaN (used for PySimplex.center / radius) */
int vo_circumspheres(int dim, const double* coords, int64_t n_coords, const int32_t* simplices, int64_t n,
                     double* centers, double* radii);
/* geometry::in_sphere (src/geometry.rs:91-97), host only: 1 iff sum_k (c_k - p_k)^2 < r * r (the
 * sum in axis order, no fused multiply-add), else 0 */
int vo_in_sphere(int dim, const double* p, const double* center, double radius);

/* ------------------------------------------------------------------------------------------
 * Several independent point sets on one device (batched and served modes).
 *
 * No counterpart in the reference, whose unit of work is one DelaunayTree: each set here is
 * the same construction as vo_tri_create + vo_tri_insert(pts, n, seed) on its own points
 * (its own bounding sphere and super simplex, its own priority order from seed) and yields the
 * same simplices; the sets share the device arrays and every round's kernels.  A vo_engine
 * holds sets of one dimension; a vo_server runs a 2D and a 3D engine behind a request queue.
 * ------------------------------------------------------------------------------------------ */
typedef struct vo_engine vo_engine;

typedef struct {
    int64_t tag;            /* the caller's tag given at submission */
    int32_t dim;
    int32_t status;         /* VO_OK, or VO_ERR_DUPLICATE (coincident points skipped, rest inserted) */
    int64_t n_points;
    int64_t points_duplicate;
    int64_t n_simplices;
    int32_t* simplices;     /* n_simplices x (dim+1) live simplices, reference id space (input
                               point i = 2(dim+1) + i, super vertices 0..dim); release with vo_free */
    double t_submit, t_admit, t_done;   /* seconds on vo_clock(): queued, resident, result ready */
    int64_t rounds;         /* engine rounds the set was resident for */
} vo_set_result;

/* engine statistics; points_inserted counts every set's inserted points as of the last replay */
typedef struct {
    int64_t sets_submitted, sets_admitted, sets_done, sets_resident, sets_queued;
    int64_t points_inserted, rounds, replays, solo_runs, range_extensions, admission_waits;
    int64_t vertex_capacity, simplex_capacity, simplex_in_use, max_sets_resident;
    double ms_replays, ms_boundary;      /* device time inside graph replays; host time between them */
    double ms_admit, ms_collect;         /* parts of ms_boundary: admission (upload, scatter) and export */
} vo_engine_stats;

/* max_points: vertex-arena capacity (points resident at once, super vertices included);
 * max_simplices: simplex-arena capacity (< 2^29; 0 = 30 (3D) / 8 (2D) per point).  Device
 * memory is about 48 B per vertex plus 112 B (3D) / 80 B (2D) per simplex, allocated once. */
int vo_engine_create(int hip_device, int dim, int64_t max_points, int64_t max_simplices, vo_engine** out);
void vo_engine_destroy(vo_engine* e);
/* queue one set (copied); it becomes resident when its footprint fits */
int vo_engine_submit(vo_engine* e, const double* pts, int64_t n, uint64_t seed, int64_t tag);
/* admit what fits, replay the captured rounds once, re-run oversized cavities, extend full
 * simplex ranges, export and free completed sets.  n_left = sets queued or resident after. */
int vo_engine_step(vo_engine* e, int64_t* n_left);
/* the same step split in two: launch returns once the replay is queued on the engine's stream
 * (so that engines on other streams can overlap it); finish waits for it and does the rest */
int vo_engine_launch(vo_engine* e);
int vo_engine_finish(vo_engine* e, int64_t* n_left);
/* move up to max completed results into out (ownership of simplices passes to the caller) */
int vo_engine_results(vo_engine* e, vo_set_result* out, int64_t max, int64_t* n_out);
int vo_engine_stats_get(const vo_engine* e, vo_engine_stats* out);
/* knobs: "rounds_per_replay" (1..64, default 8), "lanes_target" (as vo_tri_config; unset in 3D,
 * 32 lanes per point up to 65,536 window points and 16 beyond), "group_lanes" (0 = per replay, or
 * 1/4/8/16/32), "min_window" (least fresh points per set
 * and round, default 8), "window_cap" (points per round, up to the allocated 2^20/2^21),
 * "tets_per_point_milli" (first simplex range per point x 1000, default 30000 (3D) /
 * 7000 (2D)), "window_gamma_milli" (initial window per inserted point x 1000), "filters",
 * "abort_early" and "big_window" (as vo_tri_config), "outcome_flags" (1 default / 0: as vo_tri_config;
 * the lane-group commit writes per-slot outcomes and a kernel after it appends them to the pending
 * list; measured 4,096 3D sets: replays 588 -> 576 ms, 2D: 122 -> 115 ms) */
int vo_engine_config(vo_engine* e, const char* key, int64_t value);
/* batched mode: n_sets sets, set k = points set_offsets[k] .. set_offsets[k+1]-1 of pts
 * (row-major, dim columns), seed of set k = seed + k; runs to completion and writes results[k]
 * for every set (tag = k).  Returns the first error; results of completed sets are valid. */
int vo_batch_triangulate(vo_engine* e, const double* pts, const int64_t* set_offsets, int64_t n_sets,
                         uint64_t seed, vo_set_result* results);
/* canonical T / E (as vo_tri_canonical) of a result's simplices; arrays released with vo_free */
int vo_result_canonical(const vo_set_result* r, int32_t** T, int64_t* n_T, int32_t** E, int64_t* n_E);
double vo_clock(void);

/* served mode: requests of either dimension arrive from any thread; a driver thread admits
 * them into a 2D and a 3D engine by device-memory footprint as slots free up, replays both
 * engines' rounds on two streams, and queues results as sets complete (continuous batching). */
typedef struct vo_server vo_server;
int vo_server_create(int hip_device, int64_t max_points_2d, int64_t max_points_3d, vo_server** out);
int vo_server_submit(vo_server* s, int dim, const double* pts, int64_t n, uint64_t seed, int64_t tag);
/* completed results, up to max; waits up to timeout_ms for at least one when none is ready */
int vo_server_results(vo_server* s, vo_set_result* out, int64_t max, int64_t* n_out, int timeout_ms);
int vo_server_stats_get(const vo_server* s, int dim, vo_engine_stats* out);
/* stops the driver after the work in hand; results not collected are freed */
void vo_server_destroy(vo_server* s);

#ifdef __cplusplus
}
#endif
#endif
