"""The C++ facade (include/voronoids.hpp) over the C ABI.

CPU: the header compiles with every member of DelaunayTree<3, 4> and <2, 3> instantiated, and
the rand 0.8 streams of the reference's tests (include/voronoids_rand.hpp) pass their
known-answer tests.
GPU: tests/cpp/_bin/test_delaunay_tree (built by tests/cpp/Makefile from build()) runs the
reference's tree tests (tests/test_delaunay_tree.rs) through the facade and checks each result
against the oracle; it exits non-zero on any failed check.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "_bin", "test_delaunay_tree")

_INSTANTIATE = r"""
#include "voronoids.hpp"
template class voronoids::delaunay_tree::DelaunayTree<3, 4>;
template class voronoids::delaunay_tree::DelaunayTree<2, 3>;
template class voronoids::delaunay_tree::TreeUpdate<3, 4>;
template class voronoids::delaunay_tree::TreeUpdate<2, 3>;
template voronoids::delaunay_tree::DelaunayTree<3, 4> voronoids::delaunay_tree::delaunay<3>(
    const std::vector<std::array<double, 3>>&, uint64_t, int);
template voronoids::delaunay_tree::DelaunayTree<2, 3> voronoids::delaunay_tree::delaunay<2>(
    const std::vector<std::array<double, 2>>&, uint64_t, int);
template std::pair<std::array<double, 3>, double> voronoids::geometry::circumsphere<3, 4>(
    const std::array<std::array<double, 3>, 4>&);
template bool voronoids::geometry::in_sphere<2>(const std::array<double, 2>&, const std::array<double, 2>&, double);
template std::pair<std::array<double, 3>, double> voronoids::geometry::bounding_sphere<3>(
    const std::vector<std::array<double, 3>>&);
"""


def test_facade_header_compiles(tmp_path):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not found")
    src = tmp_path / "inst.cpp"
    src.write_text(_INSTANTIATE)
    r = subprocess.run([cxx, "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror",
                        "-I", os.path.join(ROOT, "include"), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_rand_compat_known_answers(tmp_path):
    # include/voronoids_rand.hpp: ChaCha against RFC 8439 / RFC 7539 vectors, buffer boundaries
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not found")
    exe = tmp_path / "rand_compat"
    r = subprocess.run([cxx, "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
                        "-o", str(exe), os.path.join(ROOT, "tests", "cpp", "test_rand_compat.cpp")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "rand_compat: ok" in r.stdout, r.stdout + r.stderr


_SPHERES = r"""
#include <cmath>
#include <cstdio>
#include "voronoids.hpp"
int main() {
    using voronoids::geometry::circumsphere;
    using voronoids::geometry::in_sphere;
    // tests/test_geometry.rs:5-15, bit for bit: the unit corner tetrahedron
    const auto k = circumsphere<3, 4>({{{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}}});
    bool ok = k.first[0] == 0.5 && k.first[1] == 0.5 && k.first[2] == 0.5 && k.second == 0.8660254037844386;
    // circumsphere_2d's slopes exact in binary: center (3, 1), radius fl(sqrt 10)
    const auto e = circumsphere<2, 3>({{{0, 0}, {2, 4}, {6, 2}}});
    ok = ok && e.first[0] == 3.0 && e.first[1] == 1.0 && e.second == std::sqrt(10.0);
    // a horizontal first edge: the reference's slope formula gives NaN
    const auto h = circumsphere<2, 3>({{{0, 0}, {4, 0}, {0, 3}}});
    ok = ok && std::isnan(h.first[0]) && std::isnan(h.second);
    // a translated regular tetrahedron: every vertex at the radius
    const std::array<std::array<double, 3>, 4> s = {{{1001, 1001, 1001}, {1001, 999, 999}, {999, 1001, 999}, {999, 999, 1001}}};
    const auto r = circumsphere<3, 4>(s);
    for (int i = 0; i < 4; ++i) {
        const double d = std::sqrt((s[i][0] - r.first[0]) * (s[i][0] - r.first[0]) + (s[i][1] - r.first[1]) * (s[i][1] - r.first[1]) +
                                   (s[i][2] - r.first[2]) * (s[i][2] - r.first[2]));
        ok = ok && std::fabs(d - r.second) < 1e-12 && std::fabs(r.second - std::sqrt(3.0)) < 1e-12;
    }
    // coplanar points: the reference's LU solve fails (it panics); the facade throws
    bool threw = false;
    try {
        circumsphere<3, 4>({{{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0}}});
    } catch (const voronoids::Error& err) {
        threw = err.code() == VO_ERR_INVALID;
    }
    ok = ok && threw;
    // in_sphere is strict: a point on the sphere is outside
    ok = ok && !in_sphere<3>({3, 4, 0}, {0, 0, 0}, 5.0) && in_sphere<3>({3, 4 - 0x1p-50, 0}, {0, 0, 0}, 5.0);
    std::printf(ok ? "spheres ok\n" : "spheres FAILED\n");
    return ok ? 0 : 1;
}
"""


def test_facade_circumsphere_known_answers(tmp_path):
    # the geometry module of the facade, which calls the library (vo_circumsphere / vo_in_sphere)
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not found")
    libdir = os.path.join(ROOT, "voronoids_amd")
    src, exe = tmp_path / "spheres.cpp", tmp_path / "spheres"
    src.write_text(_SPHERES)
    r = subprocess.run([cxx, "-std=c++17", "-O2", "-mfma", "-Wall", "-Wextra", "-Werror", "-I",
                        os.path.join(ROOT, "include"), "-o", str(exe), str(src), "-L", libdir, "-lvoronoids_hip",
                        "-Wl,-rpath," + libdir], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "spheres ok" in r.stdout, r.stdout + r.stderr


@pytest.mark.gpu
def test_cpp_delaunay_tree_facade():
    assert os.path.exists(BIN), "tests/cpp/_bin/test_delaunay_tree is missing: run __graft_entry__.build()"
    print(f"running {BIN}", flush=True)
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    for name in ("test_delaunay_tree_3d", "test_delaunay_tree_2d", "test_delaunay_function",
                 "test_insert_points_parallel_and_geometry", "test_errors"):
        assert f"{name} passed" in r.stdout
