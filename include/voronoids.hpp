/*
 * voronoids.hpp -- C++17 facade over the C ABI of libvoronoids_hip.so (voronoids.h), header only.
 *
 * It mirrors the reference's Rust API, voronoids::delaunay_tree (src/delaunay_tree.rs), for C++
 * callers: DelaunayTree<N, M> with new (the constructor), add_points_to_tree, insert_point,
 * insert_points_parallel, check_delaunay and max_simplex_id, TreeUpdate<N, M>, the Simplex / Vertex
 * records of its DashMaps as snapshot views, and the geometry module's circumsphere, in_sphere and
 * bounding_sphere.  Every call goes through the C ABI; the construction runs on the
 * GPU.  Link with -lvoronoids_hip.
 *
 * Differences from the reference, all forced by the construction living on the device:
 *   - TreeUpdate::new (src/delaunay_tree.rs:709-740) computes the cavity of a point on the host
 *     against the current tree; here it records the id and the point, and the cavity is found on
 *     the device when insert_point applies it.  An update therefore never goes stale.
 *   - The id of an inserted point is the next free one (vertices().size() before the insertion,
 *     as in the reference's tests and src/lib.rs:117-120); insert_point rejects any other id with
 *     std::invalid_argument instead of keying a DashMap entry by it.
 *   - Failures throw voronoids::Error (code = VO_ERR_*, what() = vo_last_error()) where the
 *     reference panics: a coincident point (the reference's locate unwraps nothing,
 *     src/delaunay_tree.rs:33-47) throws VO_ERR_DUPLICATE after every other point was inserted.
 *   - Simplex ids in the views are positions in an export of the live simplices, not creation
 *     ids; max_simplex_id() follows the reference's counter convention (voronoids.h).
 *   - check_delaunay() is the exact O(T) test of vo_tri_check (every face's opposite vertex
 *     outside the circumsphere) rather than the reference's O(T x V) distance loop with prints
 *     (src/delaunay_tree.rs:512-541); violations() gives the count.
 */
#ifndef VORONOIDS_HPP
#define VORONOIDS_HPP

#include <algorithm>
#include <array>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "voronoids.h"

namespace voronoids {

/* a failed C ABI call: code is the VO_ERR_* value, what() the library's message */
class Error : public std::runtime_error {
public:
    Error(int code, const std::string& message) : std::runtime_error(message), code_(code) {}
    int code() const noexcept { return code_; }

private:
    int code_;
};

namespace detail {
inline void check(int rc) {
    if (rc == VO_OK) return;
    const char* m = vo_last_error();
    throw Error(rc, (m && *m) ? std::string(m) : "voronoids: error " + std::to_string(rc));
}

}  // namespace detail

/* the reference's geometry module (src/geometry.rs) on the host */
namespace geometry {

/* circumsphere (src/geometry.rs:58-88): the reference's own formulas bit for bit, computed by the
 * library (vo_circumsphere): the slope formula in 2D (:2-22; an axis-parallel first or second edge
 * gives its IEEE result, NaN for a horizontal one) and nalgebra's LU solve in 3D (:24-56).  A 3D
 * simplex with a zero pivot throws voronoids::Error (the reference panics on the failed solve). */
template <std::size_t N, std::size_t M>
std::pair<std::array<double, N>, double> circumsphere(const std::array<std::array<double, N>, M>& vertices) {
    static_assert(N == 2 || N == 3, "2D and 3D only");
    static_assert(M == N + 1, "N + 1 points");
    double v[M * N];
    for (std::size_t k = 0; k < M; ++k)
        for (std::size_t j = 0; j < N; ++j) v[k * N + j] = vertices[k][j];
    std::pair<std::array<double, N>, double> out;
    detail::check(vo_circumsphere((int)N, v, out.first.data(), &out.second));
    return out;
}

/* in_sphere (src/geometry.rs:91-97): the reference's strict test, squared distance < radius^2,
 * evaluated by the library (compiled without fused multiply-adds, so that the caller's compiler
 * flags cannot change a boundary decision) */
template <std::size_t N>
bool in_sphere(const std::array<double, N>& vertex, const std::array<double, N>& center, double radius) {
    static_assert(N == 2 || N == 3, "2D and 3D only");
    const int r = vo_in_sphere((int)N, vertex.data(), center.data(), radius);
    detail::check(r < 0 ? r : 0);
    return r == 1;
}

/* bounding_sphere (src/geometry.rs:99-141), bit for bit (vo_bounding_sphere) */
template <std::size_t N>
std::pair<std::array<double, N>, double> bounding_sphere(const std::vector<std::array<double, N>>& vertices) {
    static_assert(N == 2 || N == 3, "2D and 3D only");
    std::pair<std::array<double, N>, double> out;
    detail::check(vo_bounding_sphere((int)N, vertices.empty() ? nullptr : vertices[0].data(), (int64_t)vertices.size(),
                                     out.first.data(), &out.second));
    return out;
}

}  // namespace geometry

namespace delaunay_tree {

/* struct Simplex (src/delaunay_tree.rs:10-16): vertex ids, circumsphere, adjacent simplices */
template <std::size_t N, std::size_t M>
struct Simplex {
    std::array<std::size_t, M> vertices;
    std::array<double, N> center;
    double radius;
    std::vector<std::size_t> neighbors;
};

/* struct Vertex (src/delaunay_tree.rs:18-22): coordinates and the simplices that use it */
template <std::size_t N>
struct Vertex {
    std::array<double, N> coordinates;
    std::vector<std::size_t> simplex;
};

/* T: simplices made only of input points, as sorted tuples of input indices (order of
 * insertion), rows in lexicographic order; E: the input-point pairs that share a live simplex,
 * sorted and unique.  The same form for any insertion order. */
struct Canonical {
    std::size_t arity = 0;
    std::vector<int32_t> T;
    std::vector<int32_t> E;
    std::size_t n_T() const { return arity ? T.size() / arity : 0; }
    std::size_t n_E() const { return E.size() / 2; }
};

template <std::size_t N, std::size_t M>
class DelaunayTree;

/* struct TreeUpdate (src/delaunay_tree.rs:700-740): one point to insert with its vertex id */
template <std::size_t N, std::size_t M>
class TreeUpdate {
public:
    TreeUpdate(std::size_t id, const std::array<double, N>& vertex, const DelaunayTree<N, M>& tree)
        : id_(id), vertex_(vertex), tree_(&tree) {}
    std::size_t id() const noexcept { return id_; }
    const std::array<double, N>& vertex() const noexcept { return vertex_; }
    const DelaunayTree<N, M>* tree() const noexcept { return tree_; }

private:
    std::size_t id_;
    std::array<double, N> vertex_;
    const DelaunayTree<N, M>* tree_;
};

/* struct DelaunayTree (src/delaunay_tree.rs:24-30) on one HIP device */
template <std::size_t N, std::size_t M>
class DelaunayTree {
    static_assert(N == 2 || N == 3, "DelaunayTree is implemented for N = 2 and N = 3");
    static_assert(M == N + 1, "M must be N + 1");

public:
    using Point = std::array<double, N>;

    /* DelaunayTree::new (src/delaunay_tree.rs:389-510 (3D), :544-640 (2D)): the super simplex
     * around the bounding sphere of vertices; no point is inserted */
    explicit DelaunayTree(const std::vector<Point>& vertices, int device = 0) {
        detail::check(vo_tri_create(device, (int)N, flat(vertices), (int64_t)vertices.size(), &h_));
    }
    DelaunayTree(const DelaunayTree&) = delete;
    DelaunayTree& operator=(const DelaunayTree&) = delete;
    DelaunayTree(DelaunayTree&& o) noexcept : h_(std::exchange(o.h_, nullptr)), n_input_(o.n_input_) {}
    DelaunayTree& operator=(DelaunayTree&& o) noexcept {
        if (this != &o) {
            if (h_) vo_tri_destroy(h_);
            h_ = std::exchange(o.h_, nullptr);
            n_input_ = o.n_input_;
        }
        return *this;
    }
    ~DelaunayTree() {
        if (h_) vo_tri_destroy(h_);
    }

    /* add_points_to_tree (src/delaunay_tree.rs:336-386): every point inserted (ids continue from
     * vertices().size()), in a random priority order drawn from seed; the result does not depend
     * on the order */
    void add_points_to_tree(const std::vector<Point>& vertices, uint64_t seed = 0) {
        if (vertices.empty()) return;
        stage(flat(vertices), vertices.size(), seed);
        detail::check(vo_tri_run(h_));
    }

    /* insert_point (src/delaunay_tree.rs:125-211) of an update made for this tree */
    void insert_point(const TreeUpdate<N, M>& update) {
        if (update.tree() != this) throw std::invalid_argument("insert_point: the update was made for another tree");
        if (update.id() != num_vertices())
            throw std::invalid_argument("insert_point: vertex id " + std::to_string(update.id()) +
                                        " is not the next free id " + std::to_string(num_vertices()));
        stage(update.vertex().data(), 1, 0);
        detail::check(vo_tri_run(h_));
    }

    /* insert_points_parallel (src/delaunay_tree.rs:213-334): the updates' points inserted together
     * (ids num_vertices() .. + updates.size() - 1, in order).  The reference expects updates whose
     * cavities do not overlap (one level of find_placement); here any set is valid. */
    void insert_points_parallel(const std::vector<TreeUpdate<N, M>>& updates, uint64_t seed = 0) {
        if (updates.empty()) return;
        std::vector<double> pts;
        pts.reserve(updates.size() * N);
        for (std::size_t i = 0; i < updates.size(); ++i) {
            if (updates[i].tree() != this)
                throw std::invalid_argument("insert_points_parallel: an update was made for another tree");
            if (updates[i].id() != num_vertices() + i)
                throw std::invalid_argument("insert_points_parallel: update " + std::to_string(i) + " has vertex id " +
                                            std::to_string(updates[i].id()) + ", not " +
                                            std::to_string(num_vertices() + i));
            pts.insert(pts.end(), updates[i].vertex().begin(), updates[i].vertex().end());
        }
        stage(pts.data(), updates.size(), seed);
        detail::check(vo_tri_run(h_));
    }

    /* check_delaunay (src/delaunay_tree.rs:512-541, :642-671) */
    bool check_delaunay() const { return violations() == 0; }
    int64_t violations() const {
        int64_t n = 0;
        detail::check(vo_tri_check(h_, &n));
        return n;
    }

    /* the max_simplex_id field (src/delaunay_tree.rs:29): dim + 1 after new */
    std::size_t max_simplex_id() const {
        int64_t m = 0;
        detail::check(vo_tri_counts(h_, nullptr, nullptr, &m));
        return (std::size_t)m;
    }
    /* vertices.len(): super and ghost vertices (2 (N + 1)) plus every point handed in */
    std::size_t num_vertices() const noexcept { return 2 * M + n_input_; }
    /* simplices.len() over live simplices (the hull of the super simplex has no ghost cells here) */
    std::size_t num_simplices() const {
        int64_t s = 0;
        detail::check(vo_tri_counts(h_, nullptr, &s, nullptr));
        return (std::size_t)s;
    }

    /* live simplices as vertex ids (simp) and the export index across each face (nbr, -1 on
     * the hull), row-major with M columns */
    void export_arrays(std::vector<int32_t>& simp, std::vector<int32_t>& nbr) const {
        const std::size_t n = num_simplices();
        simp.assign(n * M, 0);
        nbr.assign(n * M, 0);
        detail::check(vo_tri_export(h_, simp.data(), nbr.data()));
    }
    /* coordinates of every vertex id (super and ghost vertices included) */
    std::vector<Point> coordinates() const {
        std::vector<Point> x(num_vertices());
        detail::check(vo_tri_vertices(h_, x.empty() ? nullptr : x[0].data()));
        return x;
    }

    /* snapshot of the simplices map: one entry per live simplex, in export order */
    std::vector<Simplex<N, M>> simplices() const {
        std::vector<int32_t> s, nb;
        export_arrays(s, nb);
        const std::vector<Point> x = coordinates();
        std::vector<Simplex<N, M>> out(s.size() / M);
        std::vector<double> centers(out.size() * N), radii(out.size());
        detail::check(vo_circumspheres((int)N, x.empty() ? nullptr : x[0].data(), (int64_t)x.size(), s.data(),
                                       (int64_t)out.size(), centers.data(), radii.data()));
        for (std::size_t i = 0; i < out.size(); ++i) {
            for (std::size_t k = 0; k < M; ++k) {
                out[i].vertices[k] = (std::size_t)s[i * M + k];
                if (nb[i * M + k] >= 0) out[i].neighbors.push_back((std::size_t)nb[i * M + k]);
            }
            for (std::size_t j = 0; j < N; ++j) out[i].center[j] = centers[i * N + j];
            out[i].radius = radii[i];
        }
        return out;
    }
    /* snapshot of the vertices map, indexed by vertex id */
    std::vector<Vertex<N>> vertices() const {
        std::vector<int32_t> s, nb;
        export_arrays(s, nb);
        const std::vector<Point> x = coordinates();
        std::vector<Vertex<N>> out(x.size());
        for (std::size_t v = 0; v < x.size(); ++v) out[v].coordinates = x[v];
        for (std::size_t i = 0; i < s.size(); ++i) out[(std::size_t)s[i]].simplex.push_back(i / M);
        return out;
    }

    Canonical canonical() const {
        int32_t *T = nullptr, *E = nullptr;
        int64_t nT = 0, nE = 0;
        detail::check(vo_tri_canonical(h_, &T, &nT, &E, &nE));
        Canonical c;
        c.arity = M;
        c.T.assign(T, T + nT * (int64_t)M);
        c.E.assign(E, E + 2 * nE);
        vo_free(T);
        vo_free(E);
        return c;
    }

    /* tuning knobs of vo_tri_config */
    void config(const std::string& key, int64_t value) { detail::check(vo_tri_config(h_, key.c_str(), value)); }
    vo_stats stats() const {
        vo_stats s{};
        detail::check(vo_tri_stats(h_, &s));
        return s;
    }
    vo_tri* handle() const noexcept { return h_; }

private:
    static const double* flat(const std::vector<Point>& v) { return v.empty() ? nullptr : v[0].data(); }
    /* a failed stage consumes no ids; a staged point keeps its id even if the run then rejects
     * it as a duplicate */
    void stage(const double* pts, std::size_t n, uint64_t seed) {
        detail::check(vo_tri_stage(h_, pts, (int64_t)n, seed));
        n_input_ += n;
    }

    vo_tri* h_ = nullptr;
    std::size_t n_input_ = 0;
};

/* voronoids.delaunay (src/lib.rs:104-125): DelaunayTree::new(points), then every point inserted */
template <std::size_t N>
DelaunayTree<N, N + 1> delaunay(const std::vector<std::array<double, N>>& points, uint64_t seed = 0, int device = 0) {
    DelaunayTree<N, N + 1> t(points, device);
    t.add_points_to_tree(points, seed);
    return t;
}

}  // namespace delaunay_tree
}  // namespace voronoids

#endif
