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
    // the unit corner tetrahedron: center (1/2, 1/2, 1/2), radius sqrt(3)/2
    const double t[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    const double* pt[4] = {t[0], t[1], t[2], t[3]};
    double c3[3], r3;
    voronoids::detail::circumsphere<3>(pt, c3, &r3);
    // a right triangle: center at the hypotenuse's midpoint (2, 1.5), radius 2.5
    const double u[3][2] = {{0, 0}, {4, 0}, {0, 3}};
    const double* pu[3] = {u[0], u[1], u[2]};
    double c2[2], r2;
    voronoids::detail::circumsphere<2>(pu, c2, &r2);
    // a translated, rotated regular tetrahedron: every vertex at the radius
    const double s[4][3] = {{1001, 1001, 1001}, {1001, 999, 999}, {999, 1001, 999}, {999, 999, 1001}};
    const double* ps[4] = {s[0], s[1], s[2], s[3]};
    double cs[3], rs;
    voronoids::detail::circumsphere<3>(ps, cs, &rs);
    bool ok = std::fabs(c3[0] - 0.5) < 1e-15 && std::fabs(c3[1] - 0.5) < 1e-15 && std::fabs(c3[2] - 0.5) < 1e-15 &&
              std::fabs(r3 - std::sqrt(3.0) / 2) < 1e-15 && std::fabs(c2[0] - 2) < 1e-15 && std::fabs(c2[1] - 1.5) < 1e-15 &&
              std::fabs(r2 - 2.5) < 1e-15;
    for (int i = 0; i < 4; ++i) {
        const double d = std::sqrt((s[i][0] - cs[0]) * (s[i][0] - cs[0]) + (s[i][1] - cs[1]) * (s[i][1] - cs[1]) +
                                   (s[i][2] - cs[2]) * (s[i][2] - cs[2]));
        ok = ok && std::fabs(d - rs) < 1e-12 && std::fabs(rs - std::sqrt(3.0)) < 1e-12;
    }
    std::printf(ok ? "spheres ok\n" : "spheres FAILED\n");
    return ok ? 0 : 1;
}
"""


def test_facade_circumsphere_known_answers(tmp_path):
    # the Simplex view's center and radius (include/voronoids.hpp detail::circumsphere)
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not found")
    src, exe = tmp_path / "spheres.cpp", tmp_path / "spheres"
    src.write_text(_SPHERES)
    r = subprocess.run([cxx, "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
                        "-o", str(exe), str(src)], capture_output=True, text=True)
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
