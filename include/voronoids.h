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
} vo_stats;

/* DelaunayTree::<3,4>::new / ::<2,3>::new (src/delaunay_tree.rs:389-510, :544-640):
 * bounding sphere of pts (src/geometry.rs:99-141), radius x10, super simplex; no point is
 * inserted.  pts is row-major n x dim f64. */
int vo_tri_create(int hip_device, int dim, const double* pts, int64_t n, vo_tri** out);
void vo_tri_destroy(vo_tri* t);

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

int vo_tri_stats(const vo_tri* t, vo_stats* out);

/* tuning and test knobs (no counterpart in the reference): "filters" = mask of the FP32 (1)
 * and FP64-filter (2) predicate tiers, 3 by default; 0 sends every predicate to the exact
 * expansion tier.  "group_lanes" = lanes cooperating on one window point in the round
 * kernels (1, 4, 8, 16, 32; 0 = chosen per round: the largest size with window x size <=
 * "lanes_target", default 2^20 in 3D and 2^16 in 2D, from the autotuner).  "window_gamma_milli" = initial window
 * size per inserted point x 1000. */
int vo_tri_config(vo_tri* t, const char* key, int64_t value);

/* input vertex coordinates in the reference id space (for PyVertex.point): (2(dim+1) + n) x dim */
int vo_tri_vertices(const vo_tri* t, double* coords);

const char* vo_last_error(void);
int vo_version(void);

#ifdef __cplusplus
}
#endif
#endif
