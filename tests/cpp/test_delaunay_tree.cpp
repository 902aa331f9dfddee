// C++ tests of the DelaunayTree facade (include/voronoids.hpp), run on one GPU.
//
// test_delaunay_tree_3d / test_delaunay_tree_2d follow the reference's own tests
// (tests/test_delaunay_tree.rs:6-37 and :39-60) step by step: new, max_simplex_id, single
// insert_point calls with ids n_points + i, add_points_to_tree, check_delaunay.  The points are
// drawn as the reference draws them, StdRng::seed_from_u64(0) and Uniform::from(0.0..1.0) in the
// same order, through include/voronoids_rand.hpp (rand 0.8 restated: its ChaCha core is checked
// against RFC vectors, its seed expansion and float mapping are not checkable here).  On top of
// the reference's checks every result is compared with the CPU oracle (oracle/liboracle.so,
// exact predicates, sequential insertion of the same points): canonical simplices and edges, and
// per-vertex incidence counts of the vertices() view.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <random>
#include <set>
#include <stdexcept>
#include <vector>

#include "oracle.h"
#include "voronoids.hpp"
#include "voronoids_rand.hpp"

using voronoids::delaunay_tree::Canonical;
using voronoids::delaunay_tree::DelaunayTree;
using voronoids::delaunay_tree::TreeUpdate;

static int g_failures = 0;
#define CHECK(cond)                                                                  \
    do {                                                                             \
        if (!(cond)) {                                                               \
            std::fprintf(stderr, "%s:%d: CHECK failed: %s\n", __FILE__, __LINE__, #cond); \
            ++g_failures;                                                            \
        }                                                                            \
    } while (0)

template <std::size_t N>
static std::vector<std::array<double, N>> uniform_points(std::mt19937_64& rng, std::size_t n) {
    std::uniform_real_distribution<double> dist(0.0, 1.0);
    std::vector<std::array<double, N>> v(n);
    for (auto& p : v)
        for (auto& x : p) x = dist(rng);
    return v;
}

// the canonical form of an oracle tree: ghost simplices dropped, vertex ids mapped to input
// indices; T = all-input simplices (rows sorted, lexicographic), E = input pairs sharing a simplex
static Canonical oracle_canonical(const or_tree* t, int dim) {
    const int M = dim + 1, nsv = 2 * M;
    const int64_t ns = or_tree_n_simplices(t), nv = or_tree_n_vertices(t);
    std::vector<int32_t> verts((std::size_t)ns * M);
    std::vector<int64_t> vin((std::size_t)nv);
    or_tree_export(t, nullptr, verts.data(), nullptr, nullptr, nullptr);
    or_tree_vertices(t, vin.data(), nullptr);
    std::set<std::vector<int32_t>> T;
    std::set<std::pair<int32_t, int32_t>> E;
    for (int64_t s = 0; s < ns; ++s) {
        bool ghost = false, real = true;
        std::vector<int32_t> row(M);
        for (int k = 0; k < M; ++k) {
            const int32_t v = verts[s * M + k];
            ghost |= v >= dim + 1 && v < nsv;
            row[k] = (int32_t)vin[v];
            real &= row[k] >= 0;
        }
        if (ghost) continue;
        for (int a = 0; a < M; ++a)
            for (int b = a + 1; b < M; ++b)
                if (row[a] >= 0 && row[b] >= 0) E.insert({std::min(row[a], row[b]), std::max(row[a], row[b])});
        if (real) {
            std::sort(row.begin(), row.end());
            T.insert(row);
        }
    }
    Canonical c;
    c.arity = M;
    for (const auto& r : T) c.T.insert(c.T.end(), r.begin(), r.end());
    for (const auto& e : E) {
        c.E.push_back(e.first);
        c.E.push_back(e.second);
    }
    return c;
}

// per-vertex count of non-ghost simplices in the oracle tree
static std::vector<int64_t> oracle_incidence(const or_tree* t, int dim) {
    const int M = dim + 1, nsv = 2 * M;
    const int64_t ns = or_tree_n_simplices(t), nv = or_tree_n_vertices(t);
    std::vector<int32_t> verts((std::size_t)ns * M);
    or_tree_export(t, nullptr, verts.data(), nullptr, nullptr, nullptr);
    std::vector<int64_t> cnt((std::size_t)nv, 0);
    for (int64_t s = 0; s < ns; ++s) {
        bool ghost = false;
        for (int k = 0; k < M; ++k) ghost |= verts[s * M + k] >= dim + 1 && verts[s * M + k] < nsv;
        if (!ghost)
            for (int k = 0; k < M; ++k) ++cnt[verts[s * M + k]];
    }
    return cnt;
}

template <std::size_t N>
static std::vector<double> flat(const std::vector<std::array<double, N>>& v) {
    std::vector<double> f;
    for (const auto& p : v) f.insert(f.end(), p.begin(), p.end());
    return f;
}

// the snapshot views: symmetric adjacency, incidence lists that add up, circumspheres through
// every vertex of the simplices made of input points, incidence counts equal to the oracle's
template <std::size_t N, std::size_t M>
static void check_views(const DelaunayTree<N, M>& tree, const or_tree* ot) {
    const auto simplices = tree.simplices();
    const auto vertices = tree.vertices();
    CHECK(simplices.size() == tree.num_simplices());
    CHECK(vertices.size() == tree.num_vertices());
    std::size_t incid = 0, spheres = 0;
    for (std::size_t s = 0; s < simplices.size(); ++s) {
        for (std::size_t j : simplices[s].neighbors) {
            const auto& back = simplices[j].neighbors;
            CHECK(std::find(back.begin(), back.end(), s) != back.end());
        }
        bool real = true;
        for (std::size_t v : simplices[s].vertices) real &= v >= 2 * M;
        if (!real) continue;
        ++spheres;
        for (std::size_t v : simplices[s].vertices) {
            double d2 = 0.0;
            for (std::size_t k = 0; k < N; ++k) {
                const double d = vertices[v].coordinates[k] - simplices[s].center[k];
                d2 += d * d;
            }
            CHECK(std::fabs(std::sqrt(d2) - simplices[s].radius) <= 1e-6 * (1.0 + simplices[s].radius));
        }
    }
    CHECK(spheres > 0);
    const std::vector<int64_t> ref = oracle_incidence(ot, (int)N);
    CHECK(ref.size() == vertices.size());
    for (std::size_t v = 0; v < vertices.size(); ++v) {
        incid += vertices[v].simplex.size();
        for (std::size_t s : vertices[v].simplex) {
            const auto& vs = simplices[s].vertices;
            CHECK(std::find(vs.begin(), vs.end(), v) != vs.end());
        }
        if (v >= 2 * M && v < ref.size()) CHECK((int64_t)vertices[v].simplex.size() == ref[v]);
    }
    CHECK(incid == M * simplices.size());
}

static void check_canonical_equal(const Canonical& a, const Canonical& b, const char* what) {
    const bool eq = a.arity == b.arity && a.T == b.T && a.E == b.E;
    if (!eq)
        std::fprintf(stderr, "%s: canonical forms differ (T %zu vs %zu rows, E %zu vs %zu)\n", what, a.n_T(), b.n_T(),
                     a.n_E(), b.n_E());
    CHECK(eq);
}

// n points of N coordinates from the reference's generator, point by point, coordinate by coordinate
template <std::size_t N>
static std::vector<std::array<double, N>> reference_points(voronoids::rand_compat::StdRng& rng, std::size_t n) {
    const voronoids::rand_compat::UniformF64 dist(0.0, 1.0);
    std::vector<std::array<double, N>> v(n);
    for (auto& p : v)
        for (auto& x : p) x = dist.sample(rng);
    return v;
}

static void test_delaunay_tree_3d() {
    auto rng = voronoids::rand_compat::StdRng::seed_from_u64(0);
    const auto vertices = reference_points<3>(rng, 1000);
    DelaunayTree<3, 4> delaunay_tree(vertices);
    CHECK(delaunay_tree.max_simplex_id() == 4);
    const std::size_t n_points = delaunay_tree.num_vertices();
    CHECK(n_points == 8);
    for (std::size_t i = 0; i < 100; ++i) {
        const TreeUpdate<3, 4> update(n_points + i, vertices[i], delaunay_tree);
        delaunay_tree.insert_point(update);
    }
    const auto vertices2 = reference_points<3>(rng, 1000);
    delaunay_tree.add_points_to_tree(vertices2);
    CHECK(delaunay_tree.check_delaunay());
    CHECK(delaunay_tree.num_vertices() == 8 + 1100);

    const std::vector<double> a = flat(vertices), b = flat(vertices2);
    or_tree* ot = or_tree_new(3, a.data(), 1000, OR_MODE_EXACT);
    CHECK(or_tree_insert_seq(ot, a.data(), 100, 0) == 0);
    CHECK(or_tree_insert_seq(ot, b.data(), 1000, 100) == 0);
    CHECK(or_tree_check_delaunay(ot) == 0);
    check_canonical_equal(delaunay_tree.canonical(), oracle_canonical(ot, 3), "3d");
    check_views(delaunay_tree, ot);
    or_tree_free(ot);
    std::printf("test_delaunay_tree_3d: %zu simplices, max_simplex_id %zu\n", delaunay_tree.num_simplices(),
                delaunay_tree.max_simplex_id());
}

static void test_delaunay_tree_2d() {
    const std::vector<std::array<double, 2>> vertices = {{0.3, 0.1}, {1.0, 0.2}, {0.1, 1.0}, {0.5, 0.5}};
    DelaunayTree<2, 3> delaunay_tree(vertices);
    CHECK(delaunay_tree.max_simplex_id() == 3);
    auto rng = voronoids::rand_compat::StdRng::seed_from_u64(0);
    const voronoids::rand_compat::UniformF64 dist(0.0, 1.0);
    const std::size_t n_points = delaunay_tree.num_vertices();
    CHECK(n_points == 6);
    std::vector<double> inserted;
    double total_ms = 0.0, worst_ms = 0.0;
    for (std::size_t i = 0; i < 1000; ++i) {
        const auto start = std::chrono::steady_clock::now();
        const double x = dist.sample(rng), y = dist.sample(rng);   // the reference's evaluation order
        const std::array<double, 2> point = {x, y};
        const TreeUpdate<2, 3> update(n_points + i, point, delaunay_tree);
        delaunay_tree.insert_point(update);
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - start).count();
        total_ms += ms;
        worst_ms = std::max(worst_ms, ms);
        inserted.insert(inserted.end(), point.begin(), point.end());
    }
    std::printf("test_delaunay_tree_2d: 1000 insert_point calls, %.3f ms mean, %.3f ms worst\n", total_ms / 1000,
                worst_ms);
    CHECK(delaunay_tree.check_delaunay());

    const std::vector<double> a = flat(vertices);
    or_tree* ot = or_tree_new(2, a.data(), 4, OR_MODE_EXACT);
    CHECK(or_tree_insert_seq(ot, inserted.data(), 1000, 0) == 0);
    check_canonical_equal(delaunay_tree.canonical(), oracle_canonical(ot, 2), "2d");
    check_views(delaunay_tree, ot);
    or_tree_free(ot);
}

// voronoids.delaunay (src/lib.rs:104-125) against the oracle's restatement of it
static void test_delaunay_function() {
    std::mt19937_64 rng(7);
    const auto pts = uniform_points<3>(rng, 20000);
    const auto tree = voronoids::delaunay_tree::delaunay<3>(pts, 5);
    CHECK(tree.check_delaunay());
    const std::vector<double> a = flat(pts);
    or_tree* ot = or_delaunay(3, a.data(), 20000, OR_MODE_EXACT, 0);
    CHECK(ot != nullptr);
    check_canonical_equal(tree.canonical(), oracle_canonical(ot, 3), "delaunay<3>");
    or_tree_free(ot);
    const vo_stats st = tree.stats();
    CHECK(st.points_inserted == 20000);
    std::printf("test_delaunay_function: 20000 points in %lld rounds, %.2f ms on the device\n", (long long)st.rounds,
                st.ms_run);
}

// insert_points_parallel (src/delaunay_tree.rs:213-334) with the updates of 1,000 points, then the
// rest through add_points_to_tree; geometry::bounding_sphere bit for bit against the oracle's
// restatement of src/geometry.rs:99-141
static void test_insert_points_parallel_and_geometry() {
    auto rng = voronoids::rand_compat::StdRng::seed_from_u64(11);
    const auto pts = reference_points<3>(rng, 3000);
    const auto bs = voronoids::geometry::bounding_sphere<3>(pts);
    const std::vector<double> a = flat(pts);
    double oc[3], orad;
    or_bounding_sphere(3, a.data(), 3000, oc, &orad);
    CHECK(bs.first[0] == oc[0] && bs.first[1] == oc[1] && bs.first[2] == oc[2] && bs.second == orad);
    CHECK(voronoids::geometry::in_sphere<3>(bs.first, bs.first, bs.second));
    CHECK(!voronoids::geometry::in_sphere<3>({bs.first[0] + bs.second, bs.first[1], bs.first[2]}, bs.first, bs.second));

    DelaunayTree<3, 4> tree(pts);
    const std::size_t n0 = tree.num_vertices();
    std::vector<TreeUpdate<3, 4>> ups;
    for (std::size_t i = 0; i < 1000; ++i) ups.emplace_back(n0 + i, pts[i], tree);
    bool threw = false;
    try {
        std::vector<TreeUpdate<3, 4>> bad = {TreeUpdate<3, 4>(n0 + 1, pts[0], tree)};
        tree.insert_points_parallel(bad);
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    CHECK(threw);
    tree.insert_points_parallel(ups);
    CHECK(tree.num_vertices() == n0 + 1000);
    tree.add_points_to_tree(std::vector<std::array<double, 3>>(pts.begin() + 1000, pts.end()));
    CHECK(tree.check_delaunay());
    or_tree* ot = or_tree_new(3, a.data(), 3000, OR_MODE_EXACT);
    CHECK(or_tree_insert_seq(ot, a.data(), 3000, 0) == 0);
    check_canonical_equal(tree.canonical(), oracle_canonical(ot, 3), "insert_points_parallel");
    or_tree_free(ot);

    // the Simplex view's spheres are the geometry module's
    const auto simplices = tree.simplices();
    const auto x = tree.coordinates();
    for (std::size_t s = 0; s < simplices.size(); s += 97) {
        std::array<std::array<double, 3>, 4> v;
        for (int k = 0; k < 4; ++k) v[k] = x[simplices[s].vertices[k]];
        const auto cs = voronoids::geometry::circumsphere<3, 4>(v);
        CHECK(cs.first == simplices[s].center && cs.second == simplices[s].radius);
    }
    std::printf("test_insert_points_parallel_and_geometry: ok\n");
}

// error behaviour: ids, foreign updates, invalid coordinates and coincident points
static void test_errors() {
    std::mt19937_64 rng(3);
    const auto pts = uniform_points<2>(rng, 200);
    DelaunayTree<2, 3> tree(pts);
    tree.add_points_to_tree(pts);
    const std::size_t next = tree.num_vertices();
    CHECK(next == 6 + 200);

    bool threw = false;
    try {
        tree.insert_point(TreeUpdate<2, 3>(next + 1, {0.25, 0.75}, tree));
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    CHECK(threw);

    DelaunayTree<2, 3> other(pts);
    threw = false;
    try {
        tree.insert_point(TreeUpdate<2, 3>(next, {0.25, 0.75}, other));
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    CHECK(threw);

    int code = 0;
    try {
        tree.insert_point(TreeUpdate<2, 3>(next, {std::numeric_limits<double>::quiet_NaN(), 0.5}, tree));
    } catch (const voronoids::Error& e) {
        code = e.code();
    }
    CHECK(code == VO_ERR_INVALID);
    CHECK(tree.num_vertices() == next);   // a rejected point consumes no id

    code = 0;
    try {
        tree.insert_point(TreeUpdate<2, 3>(next, pts[17], tree));   // coincides with point 17
    } catch (const voronoids::Error& e) {
        code = e.code();
    }
    CHECK(code == VO_ERR_DUPLICATE);
    CHECK(tree.num_vertices() == next + 1);
    CHECK(tree.check_delaunay());

    tree.insert_point(TreeUpdate<2, 3>(next + 1, {0.25, 0.75}, tree));
    CHECK(tree.check_delaunay());
    const vo_stats st = tree.stats();
    CHECK(st.points_inserted == 201);
    std::printf("test_errors: ok\n");
}

int main() {
    const struct {
        const char* name;
        void (*fn)();
    } tests[] = {{"test_delaunay_tree_3d", test_delaunay_tree_3d},
                 {"test_delaunay_tree_2d", test_delaunay_tree_2d},
                 {"test_delaunay_function", test_delaunay_function},
                 {"test_insert_points_parallel_and_geometry", test_insert_points_parallel_and_geometry},
                 {"test_errors", test_errors}};
    for (const auto& t : tests) {
        std::printf("running %s\n", t.name);
        std::fflush(stdout);
        const int before = g_failures;
        try {
            t.fn();
        } catch (const std::exception& e) {
            std::fprintf(stderr, "%s: unexpected exception: %s\n", t.name, e.what());
            ++g_failures;
        }
        std::printf("%s %s\n", t.name, g_failures == before ? "passed" : "FAILED");
        std::fflush(stdout);
    }
    if (g_failures) std::fprintf(stderr, "%d check(s) failed\n", g_failures);
    return g_failures ? 1 : 0;
}
