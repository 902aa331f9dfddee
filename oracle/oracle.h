/*
 * oracle.h -- CPU oracle for the Delaunay construction path of kazewong/Voronoids.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker and the CPU baseline:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
 * product (voronoids_amd, libvoronoids_hip.so) never links or calls it.
 *
 * It restates, in plain C, the reference's sequential insertion
 *   DelaunayTree::new          src/delaunay_tree.rs:389-510 (3D), :544-640 (2D)
 *   TreeUpdate::new            src/delaunay_tree.rs:709-740
 *   locate/find_all_neighbors  src/delaunay_tree.rs:33-75
 *   get_new_simplices          src/delaunay_tree.rs:77-123
 *   insert_point               src/delaunay_tree.rs:125-211
 *   pair_simplices             src/delaunay_tree.rs:674-695
 * and its parallel path
 *   add_points_to_tree         src/delaunay_tree.rs:336-386
 *   insert_points_parallel     src/delaunay_tree.rs:213-334
 *   make_queue/find_placement  src/scheduler.rs:6-55
 * with the same vertex ids (super 0..d, ghosts d+1..2d+1, real 2(d+1)+i), simplex ids,
 * vertex orders and neighbor-list orders.
 *
 * The conflict predicate is selectable:
 *   OR_MODE_EXACT : exact in-sphere sign (big-integer arithmetic, exact_pred.c); output is
 *                   DT(P u S), S = the reference's super simplex.
 *   OR_MODE_REF   : the reference's own arithmetic (stored circumsphere from the slope /
 *                   LU formulas of src/geometry.rs:2-56, strict distance test :91-97).
 */
#ifndef VORONOIDS_ORACLE_H
#define VORONOIDS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_MODE_EXACT = 0, OR_MODE_REF = 1 };

/* exact signs (+1/0/-1) of the predicates, big-integer arithmetic behind a double filter.
 *   orient2d(a,b,c)      = det[b-a; c-a]                       (>0: counter-clockwise)
 *   orient3d(a,b,c,d)    = det[b-a; c-a; d-a]
 *   incircle(a,b,c,d)    > 0 iff d inside circle(abc) when orient2d(a,b,c) > 0
 *   insphere(a,b,c,d,e)  > 0 iff e inside sphere(abcd) when orient3d(a,b,c,d) > 0      */
int or_orient2d(const double* a, const double* b, const double* c);
int or_orient3d(const double* a, const double* b, const double* c, const double* d);
int or_incircle(const double* a, const double* b, const double* c, const double* d);
int or_insphere(const double* a, const double* b, const double* c, const double* d, const double* e);
/* same, but always through the big-integer path (no filter); for tests of the filter */
int or_orient3d_bigint(const double* a, const double* b, const double* c, const double* d);
int or_insphere_bigint(const double* a, const double* b, const double* c, const double* d, const double* e);
int or_orient2d_bigint(const double* a, const double* b, const double* c);
int or_incircle_bigint(const double* a, const double* b, const double* c, const double* d);

/* reference geometry, same evaluation order as src/geometry.rs */
void or_circumsphere(int dim, const double* verts /* (dim+1) x dim */, double* center, double* radius);
int  or_in_sphere(int dim, const double* p, const double* center, double radius);
void or_bounding_sphere(int dim, const double* pts, int64_t n, double* center, double* radius);
/* the 2(dim+1) super + ghost vertex coordinates DelaunayTree::new builds (row-major) */
void or_super_vertices(int dim, const double* pts, int64_t n, double* out);
/* cos/sin of 2pi/3 and 4pi/3 as used by DelaunayTree::new */
void or_trig_constants(double* out4);

typedef struct or_tree or_tree;
/* DelaunayTree::new(points): bounding sphere + super/ghost setup; points are NOT inserted */
or_tree* or_tree_new(int dim, const double* pts, int64_t n, int mode);
void     or_tree_free(or_tree* t);
/* n sequential TreeUpdate::new + insert_point, vertex id = vertices.len() (src/lib.rs:117-120).
 * input_base: input index recorded for the first of these points.  Returns 0, or -1 when
 * locate finds no conflicting simplex (the reference panics there: duplicate point). */
int      or_tree_insert_seq(or_tree* t, const double* pts, int64_t n, int64_t input_base);
/* add_points_to_tree(points) with nthreads OpenMP threads.  Cavity overlaps inside a level,
 * which the reference does not detect, are serialized instead of racing; their count is
 * returned through n_overlap (nullable).  n_levels receives the number of levels. */
int      or_tree_add_points_parallel(or_tree* t, const double* pts, int64_t n, int64_t input_base,
                                     int nthreads, int64_t* n_levels, int64_t* n_overlap);
/* reference entry point voronoids.delaunay (src/lib.rs:104-125) for any dim:
 * n <= 1e5 -> sequential, else 1e5 sequential + add_points_to_tree(rest). */
or_tree* or_delaunay(int dim, const double* pts, int64_t n, int mode, int nthreads);

int64_t or_tree_max_simplex_id(const or_tree* t);
int64_t or_tree_n_vertices(const or_tree* t);
int64_t or_tree_n_simplices(const or_tree* t);   /* live, ghosts included */
/* live simplices in id order: ids[n_s], verts[n_s*M], nbrs[n_s*M] (-1 padded), centers[n_s*dim],
 * radii[n_s]; any output pointer may be NULL */
int      or_tree_export(const or_tree* t, int64_t* ids, int32_t* verts, int32_t* nbrs,
                        double* centers, double* radii);
/* vertex id -> input index (-1 for super/ghost), coords[nv*dim] (nullable) */
int      or_tree_vertices(const or_tree* t, int64_t* input_index, double* coords);
/* incidence list of one vertex (PyVertex.simplex); returns its length, writes min(len,cap) */
int64_t  or_tree_vertex_simplices(const or_tree* t, int64_t v, int64_t* out, int64_t cap);
/* check_delaunay (src/delaunay_tree.rs:512-541 / 642-671) with the tree's predicate: number of
 * (simplex, vertex) violations over simplices made only of real vertices (O(T*V)). */
int64_t  or_tree_check_delaunay(const or_tree* t);
/* vertex v moves to p[dim] without re-triangulation (the reference's pub Vertex::coordinates);
 * -1 if v is out of range */
int      or_tree_set_vertex(or_tree* t, int64_t v, const double* p);

/* find_placement (src/scheduler.rs:30-55) on an explicit queue: footprints given as CSR */
int      or_find_placement(int64_t n, const int64_t* fp_off, const int64_t* fp, int64_t* placement);

#ifdef __cplusplus
}
#endif
#endif
