/*
 * tree.c -- restatement of the reference's DelaunayTree bookkeeping (sequential insert_point
 * loop and the level-scheduled add_points_to_tree path).
 *
 * TEST INFRASTRUCTURE ONLY (oracle + CPU baseline).  Every function cites the reference
 * code it follows.  Data layout is plain arrays indexed by id instead of DashMaps; ids,
 * vertex orders and neighbor-list orders are the reference's.  The kd-tree nearest-neighbour
 * query (kiddo) is replaced by a uniform grid with the same answer (ties -> smallest id).
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define MAXM 4

typedef struct { int32_t* a; int32_t n, cap; } ivec;

static void iv_push(ivec* v, int32_t x) {
    if (v->n == v->cap) {
        v->cap = v->cap ? v->cap * 2 : 8;
        v->a = (int32_t*)realloc(v->a, sizeof(int32_t) * (size_t)v->cap);
    }
    v->a[v->n++] = x;
}
static int iv_contains(const ivec* v, int32_t x) {
    for (int32_t i = 0; i < v->n; ++i)
        if (v->a[i] == x) return 1;
    return 0;
}
static void iv_remove_all(ivec* v, int32_t x) { /* Vec::retain(|&y| y != x) */
    int32_t w = 0;
    for (int32_t i = 0; i < v->n; ++i)
        if (v->a[i] != x) v->a[w++] = v->a[i];
    v->n = w;
}
static int cmp_i32(const void* a, const void* b) {
    int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
    return (x > y) - (x < y);
}
static void iv_sort_dedup(ivec* v) {
    qsort(v->a, (size_t)v->n, sizeof(int32_t), cmp_i32);
    int32_t w = 0;
    for (int32_t i = 0; i < v->n; ++i)
        if (w == 0 || v->a[w - 1] != v->a[i]) v->a[w++] = v->a[i];
    v->n = w;
}

typedef struct {
    double lo[3], h;
    int gn[3];
    int64_t ncell;
    int32_t* head;
} grid_t;

struct or_tree {
    int dim, M, mode, nsv;
    int64_t nv, cap_v;
    double* X;
    int64_t* vin;
    ivec* inc;
    int* vlock;
    int32_t* gnext;
    int64_t max_id, cap_s;
    int32_t* sv;
    int32_t* sn;
    uint8_t* snn;
    uint8_t* alive;
    int8_t* sor;
    double* sc;
    double* sr;
    int64_t n_live;
    grid_t g;
};

/* ---------------- storage growth ---------------- */
static void grow_v(or_tree* t, int64_t need) {
    if (need <= t->cap_v) return;
    int64_t cap = t->cap_v ? t->cap_v : 64;
    while (cap < need) cap *= 2;
    t->X = (double*)realloc(t->X, sizeof(double) * (size_t)(cap * t->dim));
    t->vin = (int64_t*)realloc(t->vin, sizeof(int64_t) * (size_t)cap);
    t->inc = (ivec*)realloc(t->inc, sizeof(ivec) * (size_t)cap);
    t->vlock = (int*)realloc(t->vlock, sizeof(int) * (size_t)cap);
    t->gnext = (int32_t*)realloc(t->gnext, sizeof(int32_t) * (size_t)cap);
    for (int64_t i = t->cap_v; i < cap; ++i) {
        t->inc[i].a = NULL; t->inc[i].n = 0; t->inc[i].cap = 0;
        t->vlock[i] = 0;
        t->gnext[i] = -1;
        t->vin[i] = -1;
    }
    t->cap_v = cap;
}

static void grow_s(or_tree* t, int64_t need) {
    if (need <= t->cap_s) return;
    int64_t cap = t->cap_s ? t->cap_s : 64;
    while (cap < need) cap *= 2;
    int M = t->M;
    t->sv = (int32_t*)realloc(t->sv, sizeof(int32_t) * (size_t)(cap * M));
    t->sn = (int32_t*)realloc(t->sn, sizeof(int32_t) * (size_t)(cap * M));
    t->snn = (uint8_t*)realloc(t->snn, (size_t)cap);
    t->alive = (uint8_t*)realloc(t->alive, (size_t)cap);
    t->sor = (int8_t*)realloc(t->sor, (size_t)cap);
    t->sc = (double*)realloc(t->sc, sizeof(double) * (size_t)(cap * t->dim));
    t->sr = (double*)realloc(t->sr, sizeof(double) * (size_t)cap);
    memset(t->alive + t->cap_s, 0, (size_t)(cap - t->cap_s));
    t->cap_s = cap;
}

/* ---------------- nearest neighbour (kiddo nearest_one replacement) ---------------- */
static void grid_init(or_tree* t, const double* pts, int64_t n) {
    int d = t->dim;
    double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
    if (n > 0) {
        for (int i = 0; i < d; ++i) { lo[i] = INFINITY; hi[i] = -INFINITY; }
        for (int64_t k = 0; k < n; ++k)
            for (int i = 0; i < d; ++i) {
                double x = pts[k * d + i];
                if (x < lo[i]) lo[i] = x;
                if (x > hi[i]) hi[i] = x;
            }
    }
    double ext = 0;
    for (int i = 0; i < d; ++i) if (hi[i] - lo[i] > ext) ext = hi[i] - lo[i];
    if (!(ext > 0)) ext = 1.0;
    double vol = 1.0;
    for (int i = 0; i < d; ++i) vol *= fmax(hi[i] - lo[i], ext * 1e-3);
    double target = n > 16 ? (double)n / 2.0 : 8.0;
    if (target > 4.0e7) target = 4.0e7;
    double h = pow(vol / target, 1.0 / d);
    t->g.h = h;
    t->g.ncell = 1;
    for (int i = 0; i < d; ++i) {
        t->g.lo[i] = lo[i];
        double c = floor((hi[i] - lo[i]) / h) + 1;
        if (c > 4096) c = 4096;
        t->g.gn[i] = (int)c;
        t->g.ncell *= t->g.gn[i];
    }
    for (int i = d; i < 3; ++i) { t->g.gn[i] = 1; t->g.lo[i] = 0; }
    t->g.head = (int32_t*)malloc(sizeof(int32_t) * (size_t)t->g.ncell);
    for (int64_t c = 0; c < t->g.ncell; ++c) t->g.head[c] = -1;
}

static int cell_coord(const grid_t* g, int i, double x) {
    double f = floor((x - g->lo[i]) / g->h);
    if (!(f >= 0)) return 0;
    if (f >= g->gn[i]) return g->gn[i] - 1;
    return (int)f;
}

static int64_t cell_of(const or_tree* t, const double* p) {
    const grid_t* g = &t->g;
    int c[3] = {0, 0, 0};
    for (int i = 0; i < t->dim; ++i) c[i] = cell_coord(g, i, p[i]);
    return ((int64_t)c[2] * g->gn[1] + c[1]) * g->gn[0] + c[0];
}

static void grid_add(or_tree* t, int64_t v) {
    int64_t c = cell_of(t, t->X + v * t->dim);
    t->gnext[v] = t->g.head[c];
    t->g.head[c] = (int32_t)v;
}

static double dist2(int d, const double* a, const double* b) {
    double s = 0.0;
    for (int i = 0; i < d; ++i) s += (a[i] - b[i]) * (a[i] - b[i]);
    return s;
}

static void scan_cell(const or_tree* t, int64_t c, const double* p, double* best, int64_t* bid) {
    for (int32_t v = t->g.head[c]; v >= 0; v = t->gnext[v]) {
        double dd = dist2(t->dim, p, t->X + (int64_t)v * t->dim);
        if (dd < *best || (dd == *best && v < *bid)) { *best = dd; *bid = v; }
    }
}

static int64_t nn_query(const or_tree* t, const double* p) {
    double best = INFINITY;
    int64_t bid = -1;
    for (int v = 0; v < t->nsv; ++v) {
        double dd = dist2(t->dim, p, t->X + v * t->dim);
        if (dd < best) { best = dd; bid = v; }
    }
    const grid_t* g = &t->g;
    int c[3] = {0, 0, 0};
    for (int i = 0; i < t->dim; ++i) c[i] = cell_coord(g, i, p[i]);
    int maxr = 0;
    for (int i = 0; i < t->dim; ++i) {
        int a = c[i], b = g->gn[i] - 1 - c[i];
        if (a > maxr) maxr = a;
        if (b > maxr) maxr = b;
    }
    for (int r = 0; r <= maxr; ++r) {
        if (r >= 1) {
            double gap = (r - 1) * g->h;
            if (gap * gap > best * (1.0 + 1e-9)) break;
        }
        int z0 = t->dim == 3 ? c[2] - r : 0, z1 = t->dim == 3 ? c[2] + r : 0;
        for (int z = z0; z <= z1; ++z) {
            if (z < 0 || z >= g->gn[2]) continue;
            int zedge = t->dim == 3 && (z == z0 || z == z1);
            for (int y = c[1] - r; y <= c[1] + r; ++y) {
                if (y < 0 || y >= g->gn[1]) continue;
                int yedge = (y == c[1] - r || y == c[1] + r);
                if (zedge || yedge) {
                    for (int x = c[0] - r; x <= c[0] + r; ++x) {
                        if (x < 0 || x >= g->gn[0]) continue;
                        scan_cell(t, ((int64_t)z * g->gn[1] + y) * g->gn[0] + x, p, &best, &bid);
                    }
                } else {
                    int xs[2] = {c[0] - r, c[0] + r};
                    for (int k = 0; k < (r == 0 ? 1 : 2); ++k) {
                        int x = xs[k];
                        if (x < 0 || x >= g->gn[0]) continue;
                        scan_cell(t, ((int64_t)z * g->gn[1] + y) * g->gn[0] + x, p, &best, &bid);
                    }
                }
            }
        }
    }
    return bid;
}

/* ---------------- conflict predicate ---------------- */
static int is_ghost_vertex(const or_tree* t, int32_t v) { return v >= t->M && v < t->nsv; }

static int orientation_of(const or_tree* t, const int32_t* vs) {
    const double* X = t->X;
    int d = t->dim;
    for (int i = 0; i < t->M; ++i)
        if (is_ghost_vertex(t, vs[i])) return 0;
    if (d == 2) return or_orient2d(X + vs[0] * 2, X + vs[1] * 2, X + vs[2] * 2);
    return or_orient3d(X + vs[0] * 3, X + vs[1] * 3, X + vs[2] * 3, X + vs[3] * 3);
}

/* in_sphere(p, simplex.center, simplex.radius) of the reference (src/geometry.rs:91-97), or
 * its exact counterpart: p strictly inside the circumsphere; ghost simplices never. */
static int conflict(const or_tree* t, int64_t s, const double* p) {
    if (t->mode == OR_MODE_REF) return or_in_sphere(t->dim, p, t->sc + s * t->dim, t->sr[s]);
    const int32_t* vs = t->sv + s * t->M;
    int o = t->sor[s];
    if (o == 0) return 0;
    const double* X = t->X;
    int sg;
    if (t->dim == 2)
        sg = or_incircle(X + vs[0] * 2, X + vs[1] * 2, X + vs[2] * 2, p);
    else
        sg = or_insphere(X + vs[0] * 3, X + vs[1] * 3, X + vs[2] * 3, X + vs[3] * 3, p);
    return sg * o > 0;
}

/* ---------------- DelaunayTree::new ---------------- */
static void put_simplex(or_tree* t, int64_t id, const int32_t* vs, const int32_t* nb, int nnb,
                        const double* c, double r, int locked) {
    int M = t->M, d = t->dim;
    memcpy(t->sv + id * M, vs, sizeof(int32_t) * (size_t)M);
    for (int i = 0; i < M; ++i) t->sn[id * M + i] = i < nnb ? nb[i] : -1;
    t->snn[id] = (uint8_t)nnb;
    for (int i = 0; i < d; ++i) t->sc[id * d + i] = c[i];
    t->sr[id] = r;
    t->sor[id] = t->mode == OR_MODE_EXACT ? (int8_t)orientation_of(t, vs) : 0;
    if (!t->alive[id]) {
        if (locked) __atomic_fetch_add(&t->n_live, 1, __ATOMIC_RELAXED);
        else t->n_live++;
    }
    t->alive[id] = 1;
}

/* src/delaunay_tree.rs:389-510 (3D) and :544-640 (2D) */
or_tree* or_tree_new(int dim, const double* pts, int64_t n, int mode) {
    if (dim != 2 && dim != 3) return NULL;
    or_tree* t = (or_tree*)calloc(1, sizeof(or_tree));
    t->dim = dim; t->M = dim + 1; t->mode = mode; t->nsv = 2 * (dim + 1);
    grow_v(t, t->nsv + 64);
    grow_s(t, 64);
    double sv[8 * 3];
    or_super_vertices(dim, pts, n, sv);
    memcpy(t->X, sv, sizeof(double) * (size_t)(t->nsv * dim));
    t->nv = t->nsv;
    double c[3], R;
    or_bounding_sphere(dim, pts, n, c, &R);
    R *= 10.0;
    double zero[3] = {0, 0, 0};
    if (dim == 3) {
        static const int32_t inc0[8][4] = {{0, 1, 2, 3}, {0, 1, 3, 4}, {0, 1, 2, 4}, {0, 2, 3, 4},
                                           {1}, {2}, {3}, {4}};
        static const int ninc[8] = {4, 4, 4, 4, 1, 1, 1, 1};
        for (int v = 0; v < 8; ++v)
            for (int k = 0; k < ninc[v]; ++k) iv_push(&t->inc[v], inc0[v][k]);
        int32_t s0[4] = {0, 1, 2, 3}, n0[4] = {1, 2, 3, 4};
        int32_t g[4][4] = {{4, 0, 1, 2}, {5, 0, 2, 3}, {6, 0, 3, 1}, {7, 1, 2, 3}};
        int32_t gn[1] = {0};
        put_simplex(t, 0, s0, n0, 4, c, R, 0);
        for (int i = 0; i < 4; ++i) put_simplex(t, i + 1, g[i], gn, 1, zero, 0.0, 0);
        t->max_id = 4;
    } else {
        static const int32_t inc0[6][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1}, {2}, {3}};
        static const int ninc[6] = {3, 3, 3, 1, 1, 1};
        for (int v = 0; v < 6; ++v)
            for (int k = 0; k < ninc[v]; ++k) iv_push(&t->inc[v], inc0[v][k]);
        int32_t s0[3] = {0, 1, 2}, n0[3] = {1, 2, 3};
        int32_t g[3][3] = {{3, 0, 1}, {4, 0, 2}, {5, 1, 2}};
        int32_t gn[1] = {0};
        put_simplex(t, 0, s0, n0, 3, c, R, 0);
        for (int i = 0; i < 3; ++i) put_simplex(t, i + 1, g[i], gn, 1, zero, 0.0, 0);
        t->max_id = 3;
    }
    grid_init(t, pts, n);
    return t;
}

void or_tree_free(or_tree* t) {
    if (!t) return;
    for (int64_t i = 0; i < t->cap_v; ++i) free(t->inc[i].a);
    free(t->X); free(t->vin); free(t->inc); free(t->vlock); free(t->gnext);
    free(t->sv); free(t->sn); free(t->snn); free(t->alive); free(t->sor); free(t->sc); free(t->sr);
    free(t->g.head);
    free(t);
}

/* ---------------- TreeUpdate ---------------- */
typedef struct {
    double p[3];
    int64_t id;            /* vertex id */
    int64_t input;         /* input index */
    ivec killed;           /* sorted */
    ivec news;             /* C*M vertex ids */
    ivec nbk;              /* C pairs (outside neighbor, killed) */
    ivec pairs;            /* local 1-based id pairs from pair_simplices */
    double* cen;           /* C*dim */
    double* rad;           /* C */
    int ccap;
    int ok;
} upd_t;

static void upd_free(upd_t* u) {
    free(u->killed.a); free(u->news.a); free(u->nbk.a); free(u->pairs.a); free(u->cen); free(u->rad);
    memset(u, 0, sizeof(*u));
}

/* locate + find_all_neighbors: src/delaunay_tree.rs:33-75 */
static int locate(const or_tree* t, const double* p, ivec* out, ivec* stack) {
    out->n = 0;
    int64_t v = nn_query(t, p);
    if (v < 0) return -1;
    const ivec* list = &t->inc[v];
    for (int32_t k = 0; k < list->n; ++k) {
        int32_t s = list->a[k];
        if (!conflict(t, s, p)) continue;
        iv_push(out, s);
        stack->n = 0;
        iv_push(stack, s);
        while (stack->n) {
            int32_t q = stack->a[--stack->n];
            for (int j = 0; j < t->snn[q]; ++j) {
                int32_t nb = t->sn[(int64_t)q * t->M + j];
                if (!iv_contains(out, nb) && conflict(t, nb, p)) {
                    iv_push(out, nb);
                    iv_push(stack, nb);
                }
            }
        }
    }
    if (out->n == 0) return -1;
    iv_sort_dedup(out);
    return 0;
}

/* get_new_simplices (src/delaunay_tree.rs:77-123) for every killed simplex, then
 * pair_simplices (:674-695): TreeUpdate::new, src/delaunay_tree.rs:709-740 */
static int tree_update(const or_tree* t, upd_t* u, ivec* stack) {
    int M = t->M, d = t->dim;
    u->news.n = u->nbk.n = u->pairs.n = 0;
    if (locate(t, u->p, &u->killed, stack) != 0) { u->ok = 0; return -1; }
    for (int32_t kk = 0; kk < u->killed.n; ++kk) {
        int32_t k = u->killed.a[kk];
        const int32_t* kv = t->sv + (int64_t)k * M;
        for (int j = 0; j < t->snn[k]; ++j) {
            int32_t nb = t->sn[(int64_t)k * M + j];
            if (conflict(t, nb, u->p)) continue;
            const int32_t* nv = t->sv + (int64_t)nb * M;
            int32_t ns[MAXM];
            ns[0] = (int32_t)u->id;
            int cnt = 1;
            for (int i = 0; i < M; ++i) {
                int in = 0;
                for (int q = 0; q < M; ++q) in |= kv[q] == nv[i];
                if (in && cnt < M) ns[cnt++] = nv[i];
            }
            for (int i = 0; i < M; ++i) iv_push(&u->news, ns[i]);
            iv_push(&u->nbk, nb);
            iv_push(&u->nbk, k);
        }
    }
    int C = u->news.n / M;
    if (C > u->ccap) {
        u->ccap = C * 2;
        u->cen = (double*)realloc(u->cen, sizeof(double) * (size_t)(u->ccap * d));
        u->rad = (double*)realloc(u->rad, sizeof(double) * (size_t)u->ccap);
    }
    for (int i = 0; i < C; ++i) {
        double vx[4 * 3];
        const int32_t* ns = u->news.a + i * M;
        for (int j = 0; j < d; ++j) vx[j] = u->p[j];
        for (int q = 1; q < M; ++q)
            for (int j = 0; j < d; ++j) vx[q * d + j] = t->X[(int64_t)ns[q] * d + j];
        or_circumsphere(d, vx, u->cen + i * d, &u->rad[i]);
    }
    for (int i = 0; i < C; ++i)
        for (int j = i + 1; j < C; ++j) {
            int count = 0;
            for (int k = 0; k < M; ++k) {
                int32_t x = u->news.a[j * M + k];
                for (int q = 0; q < M; ++q)
                    if (u->news.a[i * M + q] == x) { count++; break; }
            }
            if (count == d) {
                iv_push(&u->pairs, i + 1); iv_push(&u->pairs, j + 1);
                iv_push(&u->pairs, j + 1); iv_push(&u->pairs, i + 1);
            }
        }
    u->ok = 1;
    return 0;
}

static void atomic_min_i32(int32_t* p, int32_t v) {
    int32_t cur = __atomic_load_n(p, __ATOMIC_RELAXED);
    while (v < cur && !__atomic_compare_exchange_n(p, &cur, v, 1, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {}
}

static void lock_v(or_tree* t, int64_t v) {
    while (__atomic_exchange_n(&t->vlock[v], 1, __ATOMIC_ACQUIRE)) {}
}
static void unlock_v(or_tree* t, int64_t v) { __atomic_store_n(&t->vlock[v], 0, __ATOMIC_RELEASE); }

/* Body shared by insert_point (src/delaunay_tree.rs:125-211) and one update of
 * insert_points_parallel (:233-331): new simplex ids base+1..base+C, neighbors [nb],
 * patch the outside neighbors, push internal adjacencies, vertex + incidence lists, remove
 * the killed simplices.  Vertex storage and the simplex arrays must already be grown. */
static void apply_update(or_tree* t, const upd_t* u, int64_t base, int locked) {
    int M = t->M, d = t->dim;
    int C = u->news.n / M;
    int64_t v = u->id;
    for (int j = 0; j < d; ++j) t->X[v * d + j] = u->p[j];
    for (int i = 0; i < C; ++i) {
        int32_t nb = u->nbk.a[2 * i];
        put_simplex(t, base + i + 1, u->news.a + i * M, &nb, 1, u->cen + i * d, u->rad[i], locked);
    }
    for (int i = 0; i < C; ++i) {
        int32_t nb = u->nbk.a[2 * i], k = u->nbk.a[2 * i + 1];
        for (int j = 0; j < t->snn[nb]; ++j)
            if (t->sn[(int64_t)nb * M + j] == k) t->sn[(int64_t)nb * M + j] = (int32_t)(base + i + 1);
    }
    for (int q = 0; q < u->pairs.n; q += 2) {
        int64_t a = base + u->pairs.a[q], b = base + u->pairs.a[q + 1];
        if (t->snn[a] < M) t->sn[a * M + t->snn[a]++] = (int32_t)b;
    }
    t->vin[v] = u->input;
    t->inc[v].n = 0;
    for (int i = 0; i < C; ++i)
        for (int j = 0; j < M; ++j) {
            int64_t w = u->news.a[i * M + j];
            if (locked) lock_v(t, w);
            iv_push(&t->inc[w], (int32_t)(base + i + 1));
            if (locked) unlock_v(t, w);
        }
    for (int32_t kk = 0; kk < u->killed.n; ++kk) {
        int32_t k = u->killed.a[kk];
        for (int i = 0; i < M; ++i) {
            int64_t w = t->sv[(int64_t)k * M + i];
            if (locked) lock_v(t, w);
            iv_remove_all(&t->inc[w], k);
            if (locked) unlock_v(t, w);
        }
    }
    for (int32_t kk = 0; kk < u->killed.n; ++kk) {
        int32_t k = u->killed.a[kk];
        if (t->alive[k]) {
            t->alive[k] = 0;
            if (locked) __atomic_fetch_sub(&t->n_live, 1, __ATOMIC_RELAXED);
            else t->n_live--;
        }
    }
}

/* sequential path: TreeUpdate::new(n_points + i) + insert_point, src/lib.rs:117-120 */
int or_tree_insert_seq(or_tree* t, const double* pts, int64_t n, int64_t input_base) {
    upd_t u;
    memset(&u, 0, sizeof(u));
    ivec stack = {0, 0, 0};
    int rc = 0;
    for (int64_t i = 0; i < n; ++i) {
        for (int j = 0; j < t->dim; ++j) u.p[j] = pts[i * t->dim + j];
        u.id = t->nv;
        u.input = input_base + i;
        if (tree_update(t, &u, &stack) != 0) { rc = -1; break; }
        int C = u.news.n / t->M;
        grow_v(t, t->nv + 1);
        grow_s(t, t->max_id + C + 1);
        apply_update(t, &u, t->max_id, 0);
        t->nv++;
        grid_add(t, u.id);
        t->max_id += C;
    }
    upd_free(&u);
    free(stack.a);
    return rc;
}

/* find_placement, src/scheduler.rs:30-55 -- serial greedy levels.  occupancy[s] keeps only
 * its last occupant: the reference reads occupancy[s][len-2], i.e. the previous one. */
int or_find_placement(int64_t n, const int64_t* fp_off, const int64_t* fp, int64_t* placement) {
    int64_t maxs = -1;
    for (int64_t i = 0; i < fp_off[n]; ++i) if (fp[i] > maxs) maxs = fp[i];
    int64_t* last = (int64_t*)malloc(sizeof(int64_t) * (size_t)(maxs + 1 > 0 ? maxs + 1 : 1));
    for (int64_t s = 0; s <= maxs; ++s) last[s] = -1;
    for (int64_t id = 0; id < n; ++id) {
        int64_t best = 0;
        for (int64_t q = fp_off[id]; q < fp_off[id + 1]; ++q) {
            int64_t s = fp[q];
            int64_t prev = last[s];
            last[s] = id;
            int64_t lv = prev < 0 ? 1 : placement[prev] + 1;
            if (lv > best) best = lv;
        }
        placement[id] = best;
    }
    free(last);
    return 0;
}

/* add_points_to_tree, src/delaunay_tree.rs:336-386 with make_queue (src/scheduler.rs:6-28)
 * and insert_points_parallel (:213-334).  Updates of one level whose killed simplices or
 * outside neighbors collide (the reference races there) are recomputed and applied one by
 * one after the level's conflict-free updates. */
int or_tree_add_points_parallel(or_tree* t, const double* pts, int64_t n, int64_t input_base,
                                int nthreads, int64_t* n_levels, int64_t* n_overlap) {
    int M = t->M, d = t->dim;
    if (nthreads > 0) omp_set_num_threads(nthreads);
    if (n_overlap) *n_overlap = 0;
    if (n_levels) *n_levels = 0;
    if (n == 0) return 0;
    /* make_queue: locate + 2-hop footprint, sorted and deduplicated */
    int64_t* fp_cnt = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
    ivec* fps = (ivec*)calloc((size_t)n, sizeof(ivec));
    int bad = 0;
#pragma omp parallel
    {
        ivec killed = {0, 0, 0}, stack = {0, 0, 0};
#pragma omp for schedule(dynamic, 256)
        for (int64_t i = 0; i < n; ++i) {
            if (locate(t, pts + i * d, &killed, &stack) != 0) { bad = 1; continue; }
            ivec* f = &fps[i];
            for (int32_t a = 0; a < killed.n; ++a) {
                int32_t s = killed.a[a];
                for (int j = 0; j < t->snn[s]; ++j) {
                    int32_t s1 = t->sn[(int64_t)s * M + j];
                    for (int q = 0; q < t->snn[s1]; ++q) iv_push(f, t->sn[(int64_t)s1 * M + q]);
                }
            }
            iv_sort_dedup(f);
            fp_cnt[i + 1] = f->n;
        }
        free(killed.a); free(stack.a);
    }
    if (bad) {
        for (int64_t i = 0; i < n; ++i) free(fps[i].a);
        free(fps); free(fp_cnt);
        return -1;
    }
    for (int64_t i = 0; i < n; ++i) fp_cnt[i + 1] += fp_cnt[i];
    int64_t* fp = (int64_t*)malloc(sizeof(int64_t) * (size_t)(fp_cnt[n] > 0 ? fp_cnt[n] : 1));
    for (int64_t i = 0; i < n; ++i) {
        for (int32_t q = 0; q < fps[i].n; ++q) fp[fp_cnt[i] + q] = fps[i].a[q];
        free(fps[i].a);
    }
    free(fps);
    int64_t* placement = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    or_find_placement(n, fp_cnt, fp, placement);
    free(fp); free(fp_cnt);
    int64_t L = 0;
    for (int64_t i = 0; i < n; ++i) if (placement[i] > L) L = placement[i];
    if (n_levels) *n_levels = L;
    /* batches[level] = queue entries with that level, in queue order */
    int64_t* boff = (int64_t*)calloc((size_t)L + 2, sizeof(int64_t));
    for (int64_t i = 0; i < n; ++i) boff[placement[i] + 1]++;
    for (int64_t l = 0; l <= L; ++l) boff[l + 1] += boff[l];
    int64_t* order = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)(L + 1));
    for (int64_t l = 0; l <= L; ++l) fill[l] = boff[l];
    for (int64_t i = 0; i < n; ++i) order[fill[placement[i]]++] = i;
    free(fill);
    free(placement);

    int64_t maxb = 0;
    for (int64_t l = 1; l <= L; ++l) if (boff[l + 1] - boff[l] > maxb) maxb = boff[l + 1] - boff[l];
    upd_t* ups = (upd_t*)calloc((size_t)maxb, sizeof(upd_t));
    int64_t* off = (int64_t*)malloc(sizeof(int64_t) * (size_t)(maxb + 1));
    uint8_t* deferred = (uint8_t*)malloc((size_t)maxb);
    int32_t* owner = NULL;
    int64_t owner_cap = 0;
    int rc = 0;
    for (int64_t l = 1; l <= L && rc == 0; ++l) {
        int64_t b0 = boff[l], nb = boff[l + 1] - boff[l];
        int64_t n_points = t->nv;
        /* par: TreeUpdate::new(n_points + id, p, self) against the current tree */
#pragma omp parallel
        {
            ivec stack = {0, 0, 0};
#pragma omp for schedule(dynamic, 64)
            for (int64_t u = 0; u < nb; ++u) {
                int64_t qi = order[b0 + u];
                for (int j = 0; j < d; ++j) ups[u].p[j] = pts[qi * d + j];
                ups[u].id = n_points + u;
                ups[u].input = input_base + qi;
                tree_update(t, &ups[u], &stack);
            }
            free(stack.a);
        }
        for (int64_t u = 0; u < nb; ++u) if (!ups[u].ok) rc = -1;
        if (rc) break;
        /* collision check: lowest batch index owns each killed simplex / outside neighbor */
        if (owner_cap < t->max_id + 1) {
            owner_cap = (t->max_id + 1) * 2;
            owner = (int32_t*)realloc(owner, sizeof(int32_t) * (size_t)owner_cap);
        }
#pragma omp parallel for
        for (int64_t u = 0; u < nb; ++u) {
            for (int32_t a = 0; a < ups[u].killed.n; ++a) owner[ups[u].killed.a[a]] = INT32_MAX;
            for (int32_t a = 0; a < ups[u].nbk.n; a += 2) owner[ups[u].nbk.a[a]] = INT32_MAX;
        }
#pragma omp parallel for
        for (int64_t u = 0; u < nb; ++u) {
            for (int32_t a = 0; a < ups[u].killed.n; ++a) atomic_min_i32(&owner[ups[u].killed.a[a]], (int32_t)u);
            for (int32_t a = 0; a < ups[u].nbk.n; a += 2) atomic_min_i32(&owner[ups[u].nbk.a[a]], (int32_t)u);
        }
        int64_t nd = 0;
        for (int64_t u = 0; u < nb; ++u) {
            int ok = 1;
            for (int32_t a = 0; a < ups[u].killed.n && ok; ++a) ok = owner[ups[u].killed.a[a]] == u;
            for (int32_t a = 0; a < ups[u].nbk.n && ok; a += 2) ok = owner[ups[u].nbk.a[a]] == u;
            deferred[u] = !ok;
            nd += !ok;
        }
        if (n_overlap) *n_overlap += nd;
        /* exclusive prefix sum of created-simplex counts (src/delaunay_tree.rs:214-225) */
        off[0] = 0;
        for (int64_t u = 0; u < nb; ++u) off[u + 1] = off[u] + (deferred[u] ? 0 : ups[u].news.n / M);
        grow_v(t, n_points + nb);
        grow_s(t, t->max_id + off[nb] + 1);
        for (int64_t u = 0; u < nb; ++u)
            if (!deferred[u]) {   /* serial kd adds (src/delaunay_tree.rs:228-231) */
                for (int j = 0; j < d; ++j) t->X[ups[u].id * d + j] = ups[u].p[j];
                grid_add(t, ups[u].id);
            }
        int64_t base = t->max_id;
#pragma omp parallel for schedule(dynamic, 16)
        for (int64_t u = 0; u < nb; ++u)
            if (!deferred[u]) apply_update(t, &ups[u], base + off[u], 1);
        t->max_id += off[nb];
        t->nv = n_points + nb;
        /* colliding updates: recompute against the updated tree, apply in batch order */
        if (nd) {
            ivec stack = {0, 0, 0};
            for (int64_t u = 0; u < nb; ++u) {
                if (!deferred[u]) continue;
                if (tree_update(t, &ups[u], &stack) != 0) { rc = -1; break; }
                int C = ups[u].news.n / M;
                grow_s(t, t->max_id + C + 1);
                apply_update(t, &ups[u], t->max_id, 0);
                grid_add(t, ups[u].id);
                t->max_id += C;
            }
            free(stack.a);
        }
    }
    for (int64_t u = 0; u < maxb; ++u) upd_free(&ups[u]);
    free(ups); free(off); free(deferred); free(owner); free(order); free(boff);
    return rc;
}

/* voronoids.delaunay, src/lib.rs:104-125 */
or_tree* or_delaunay(int dim, const double* pts, int64_t n, int mode, int nthreads) {
    or_tree* t = or_tree_new(dim, pts, n, mode);
    if (!t) return NULL;
    const int64_t seq = 100000;
    int rc;
    if (n > seq) {
        rc = or_tree_insert_seq(t, pts, seq, 0);
        if (rc == 0) rc = or_tree_add_points_parallel(t, pts + seq * dim, n - seq, seq, nthreads, NULL, NULL);
    } else {
        rc = or_tree_insert_seq(t, pts, n, 0);
    }
    if (rc != 0) { or_tree_free(t); return NULL; }
    return t;
}

int64_t or_tree_max_simplex_id(const or_tree* t) { return t->max_id; }
int64_t or_tree_n_vertices(const or_tree* t) { return t->nv; }
int64_t or_tree_n_simplices(const or_tree* t) { return t->n_live; }

int or_tree_export(const or_tree* t, int64_t* ids, int32_t* verts, int32_t* nbrs, double* centers, double* radii) {
    int M = t->M, d = t->dim;
    int64_t k = 0;
    for (int64_t s = 0; s <= t->max_id; ++s) {
        if (!t->alive[s]) continue;
        if (ids) ids[k] = s;
        if (verts) memcpy(verts + k * M, t->sv + s * M, sizeof(int32_t) * (size_t)M);
        if (nbrs) memcpy(nbrs + k * M, t->sn + s * M, sizeof(int32_t) * (size_t)M);
        if (centers) memcpy(centers + k * d, t->sc + s * d, sizeof(double) * (size_t)d);
        if (radii) radii[k] = t->sr[s];
        k++;
    }
    return 0;
}

int or_tree_vertices(const or_tree* t, int64_t* input_index, double* coords) {
    if (input_index) memcpy(input_index, t->vin, sizeof(int64_t) * (size_t)t->nv);
    if (coords) memcpy(coords, t->X, sizeof(double) * (size_t)(t->nv * t->dim));
    return 0;
}

int64_t or_tree_vertex_simplices(const or_tree* t, int64_t v, int64_t* out, int64_t cap) {
    if (v < 0 || v >= t->nv) return -1;
    const ivec* l = &t->inc[v];
    for (int32_t i = 0; i < l->n && i < cap; ++i) out[i] = l->a[i];
    return l->n;
}

/* check_delaunay, src/delaunay_tree.rs:512-541 / 642-671 */
int64_t or_tree_check_delaunay(const or_tree* t) {
    int M = t->M;
    int64_t bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(dynamic, 64)
    for (int64_t s = 0; s <= t->max_id; ++s) {
        if (!t->alive[s]) continue;
        const int32_t* vs = t->sv + s * M;
        int real = 1;
        for (int i = 0; i < M; ++i) real &= vs[i] >= t->nsv;
        if (!real) continue;
        for (int64_t v = 0; v < t->nv; ++v) {
            int in = 0;
            for (int i = 0; i < M; ++i) in |= vs[i] == v;
            if (in) continue;
            if (conflict(t, s, t->X + v * t->dim)) bad++;
        }
    }
    return bad;
}

/* Vertex v moves to p: the reference's Vertex::coordinates is a pub field
 * (src/delaunay_tree.rs:18-22), so a caller can do this before check_delaunay.  Nothing is
 * re-triangulated; exact mode re-derives the stored orientation of v's simplices (its conflict
 * test reads coordinates), reference mode keeps its stored spheres, as the reference does. */
int or_tree_set_vertex(or_tree* t, int64_t v, const double* p) {
    if (v < 0 || v >= t->nv) return -1;
    memcpy(t->X + v * t->dim, p, sizeof(double) * (size_t)t->dim);
    if (t->mode != OR_MODE_EXACT) return 0;
    for (int64_t s = 0; s <= t->max_id; ++s) {
        if (!t->alive[s]) continue;
        const int32_t* vs = t->sv + s * t->M;
        int has = 0;
        for (int i = 0; i < t->M; ++i) has |= vs[i] == v;
        if (has) t->sor[s] = (int8_t)orientation_of(t, vs);
    }
    return 0;
}
