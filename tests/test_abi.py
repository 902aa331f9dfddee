"""The C-ABI library builds, loads and exports every symbol include/voronoids.h declares
(no GPU calls)."""
import ctypes
import os
import re

from voronoids_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    with open(os.path.join(ROOT, "include", "voronoids.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(vo_[a-z_]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared_symbols() == sorted(_lib.EXPORTED)


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    for name in declared_symbols():
        assert hasattr(L, name), name
        assert ctypes.cast(getattr(L, name), ctypes.c_void_p).value


def test_version_and_error_without_gpu():
    L = _lib.lib()
    assert L.vo_version() == 1
    assert isinstance(L.vo_last_error(), bytes)
    out = ctypes.c_void_p()
    # invalid dimension is rejected before any device call
    rc = L.vo_tri_create(0, 4, None, 0, ctypes.byref(out))
    assert rc == _lib.VO_ERR_INVALID and not out.value
    assert b"dim" in L.vo_last_error()


def test_result_canonical_is_host_only_and_matches_the_oracle():
    # vo_result_canonical needs no GPU: canonical T / E of a simplex list in the reference id
    # space, checked against the oracle's own canonicalizer on random tuples
    import numpy as np
    from oracle import oracle as O
    from voronoids_amd import batch as B
    rng = np.random.default_rng(5)
    for dim in (2, 3):
        M = dim + 1
        S = rng.integers(0, 2 * M + 40, size=(200, M)).astype(np.int32)
        S = S[np.array([len(set(r)) == M for r in S])]
        S = np.unique(np.sort(S, axis=1), axis=0)            # distinct simplices, as a triangulation has
        T, E = B.canonical_from_simplices(S, dim)
        ref = O.canonical_from_real(np.where(S >= 2 * M, S - 2 * M, -1), dim)
        assert np.array_equal(T, ref.T) and np.array_equal(E, ref.E)


def test_bounding_sphere_is_the_reference_s_bit_for_bit():
    # host-only entry point: the 10k-point configuration's bounding sphere against the known-answer
    # bits the oracle is pinned to (tests/golden/kat.json), and the closed-form case
    import json
    import struct
    import numpy as np
    G = os.path.join(ROOT, "tests", "golden")
    with open(os.path.join(G, "kat.json")) as f:
        kat = json.load(f)
    P = np.ascontiguousarray(np.load(os.path.join(G, "c1_points.npy")))
    L = _lib.lib()
    dp = ctypes.POINTER(ctypes.c_double)
    c = np.zeros(3)
    r = ctypes.c_double()
    assert L.vo_bounding_sphere(3, P.ctypes.data_as(dp), P.shape[0], c.ctypes.data_as(dp), ctypes.byref(r)) == 0
    bits = lambda x: "0x%016x" % struct.unpack("<Q", struct.pack("<d", float(x)))[0]
    assert [bits(x) for x in c] == kat["c1_bounding_sphere_hex"]["center"]
    assert bits(r.value) == kat["c1_bounding_sphere_hex"]["radius"]
    Q = np.array([[0.0, 0.0, 0.5], [1.0, 2.0, 1.5], [0.5, 1.0, 0.0], [0.5, 1.0, 2.0]])
    assert L.vo_bounding_sphere(3, Q.ctypes.data_as(dp), 4, c.ctypes.data_as(dp), ctypes.byref(r)) == 0
    assert c.tolist() == [0.5, 1.0, 1.0] and r.value == 1.5
    assert L.vo_bounding_sphere(4, Q.ctypes.data_as(dp), 4, c.ctypes.data_as(dp), ctypes.byref(r)) == _lib.VO_ERR_INVALID


def test_bounding_sphere_equals_the_oracle_bitwise():
    # the product's host routine and the oracle's restatement of src/geometry.rs:99-141 on random
    # sets, on the reference test's 2D points (one of them on the first sphere: the strict test
    # decides by rounding) and on sets where the 1.5x growth fires
    import numpy as np
    from oracle import oracle as O
    from voronoids_amd import datasets
    L = _lib.lib()
    dp = ctypes.POINTER(ctypes.c_double)
    sets = [np.array([[0.3, 0.1], [1.0, 0.2], [0.1, 1.0], [0.5, 0.5]])]
    for dim in (2, 3):
        for seed in range(5):
            sets.append(datasets.uniform(1000 + 37 * seed, dim, seed=500 + seed))
        sets.append(np.vstack([datasets.uniform(50, dim, seed=9), np.full((1, dim), 7.0), -np.full((1, dim), 9.0)]))
    for P in sets:
        P = np.ascontiguousarray(P, dtype=np.float64)
        dim = P.shape[1]
        c = np.zeros(dim)
        r = ctypes.c_double()
        assert L.vo_bounding_sphere(dim, P.ctypes.data_as(dp), P.shape[0], c.ctypes.data_as(dp), ctypes.byref(r)) == 0
        oc, orad = O.bounding_sphere(P)
        assert c.tobytes() == np.asarray(oc, np.float64).tobytes() and r.value == float(orad), P.shape


_STATIC_MAIN = r"""
#include <stdio.h>
#include "voronoids.h"
int main(void) {
    const double pts[6] = {-1, 0, 0, 1, 0, 0};
    double c[3], r;
    if (vo_version() != 1 || vo_bounding_sphere(3, pts, 2, c, &r) != 0) return 1;
    printf("static ok %g %g %g %g\n", c[0], c[1], c[2], r);
    return r == 1.5 ? 0 : 2;
}
"""


def test_static_archive_links_and_runs(tmp_path):
    # libvoronoids_hip.a (built beside the shared library): a C program linked against the archive
    import shutil
    import subprocess
    gcc = shutil.which("gcc")
    archive = os.path.join(ROOT, "voronoids_amd", "libvoronoids_hip.a")
    assert os.path.exists(archive), "run __graft_entry__.build()"
    src, exe = tmp_path / "m.c", tmp_path / "m"
    src.write_text(_STATIC_MAIN)
    r = subprocess.run([gcc, "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src), archive,
                        "-L/opt/rocm/lib", "-lamdhip64", "-lstdc++", "-lm", "-lpthread", "-Wl,-rpath,/opt/rocm/lib"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "static ok 0 0 0 1.5" in r.stdout, r.stdout + r.stderr
