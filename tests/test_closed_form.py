"""Known answers derived by hand from the reference's formulas, written as literal bits, so that
the oracle and the product are pinned from outside themselves:

* DelaunayTree::new's super vertices (src/delaunay_tree.rs:391-412 (3D), :546-558 (2D)) for point
  sets whose bounding sphere (src/geometry.rs:99-141) is known in closed form;
* circumsphere_2d (src/geometry.rs:2-22) on triangles whose slope formula is exact in binary;
* the product's geometry module (vo_circumsphere / vo_in_sphere, host only) against the same
  answers and, bit for bit, against the oracle's restatement on random and axis-aligned inputs.
"""
import ctypes
import math
import struct

import numpy as np
import pytest

from oracle import oracle as O
from voronoids_amd import _lib

_dp = ctypes.POINTER(ctypes.c_double)


def bits(x):
    return "0x%016x" % struct.unpack("<Q", struct.pack("<d", float(x)))[0]


def f(h):
    return struct.unpack("<d", struct.pack("<Q", int(h, 16)))[0]


# glibc's cos/sin of 2pi/3 and 4pi/3 (the reference computes them with f64::cos / f64::sin):
#   cos(2pi/3) = -1/2 + 2^-52 (2 ulp of 1/2 above -1/2), cos(4pi/3) = -1/2 - 2^-51 (4 ulp below)
COS1, SIN1 = "0xbfdffffffffffffc", "0x3febb67ae8584cab"
COS2, SIN2 = "0xbfe0000000000004", "0xbfebb67ae8584ca8"

# 3D, points (-1,0,0), (1,0,0): box center 0, both corner distances 1, radius = 1 (strict '>' keeps
# the lower one); the first point is not strictly inside (1 < 1 fails) -> 1.5; R = 10 * 1.5 = 15.
#   v0 = (0, 0, 15), v1 = (15, 0, -15)
#   v2 = (15 cos1, 15 sin1, -15): 15 (-1/2 + 2^-52) = -7.5 + 3.75 ulp(7.5) -> -7.5 + 4 ulp = 0xc01dfffffffffffc
#   v3 = (15 cos2, 15 sin2, -15): -7.5 - 15 * 2^-51 = -7.5 - 7.5 ulp -> (tie, to even) -7.5 - 8 ulp
#        = 0xc01e000000000008
#   15 sin1 and 15 sin2 round to 0x4029fb1339d2c7e0 / 0xc029fb1339d2c7de (exact products rounded once)
SUPER_3D_UNIT = [
    ["0x0000000000000000", "0x0000000000000000", "0x402e000000000000"],
    ["0x402e000000000000", "0x0000000000000000", "0xc02e000000000000"],
    ["0xc01dfffffffffffc", "0x4029fb1339d2c7e0", "0xc02e000000000000"],
    ["0xc01e000000000008", "0xc029fb1339d2c7de", "0xc02e000000000000"],
]
# 3D, points (0,0,0), (2,4,4): center (1,2,2), corner distances 3, radius 3 -> 4.5 (first point on
# the sphere), R = 45: v0 = c + (0,0,45) = (1,2,47), v1 = (46, 2, -43), v2 = (1 + fl(45 cos1),
# 2 + fl(45 sin1), -43), v3 likewise with cos2 / sin2 (each product, then each sum, rounded once)
SUPER_3D_SHIFTED = [
    ["0x3ff0000000000000", "0x4000000000000000", "0x4047800000000000"],
    ["0x4047000000000000", "0x4000000000000000", "0xc045800000000000"],
    ["0xc0357ffffffffffd", "0x40447c4e6b5e15e8", "0xc045800000000000"],
    ["0xc035800000000006", "0xc0427c4e6b5e15e6", "0xc045800000000000"],
]
# 2D, points (-1,0), (1,0): c = 0, radius 1 -> 1.5, R = 15: v0 = (15, 0), v1 = (15 cos1, 15 sin1),
# v2 = (15 cos2, 15 sin2)
SUPER_2D_UNIT = [
    ["0x402e000000000000", "0x0000000000000000"],
    ["0xc01dfffffffffffc", "0x4029fb1339d2c7e0"],
    ["0xc01e000000000008", "0xc029fb1339d2c7de"],
]

CASES = [
    (np.array([[-1.0, 0, 0], [1.0, 0, 0]]), [0.0, 0.0, 0.0], 1.5, SUPER_3D_UNIT),
    (np.array([[0.0, 0, 0], [2.0, 4, 4]]), [1.0, 2.0, 2.0], 4.5, SUPER_3D_SHIFTED),
    (np.array([[-1.0, 0], [1.0, 0]]), [0.0, 0.0], 1.5, SUPER_2D_UNIT),
]


def expected_with_ghosts(sv):
    """super vertices followed by the ghost vertices of new(): 3D ghosts 4, 5, 6 copy vertex 0
    and 7 copies vertex 1 (src/delaunay_tree.rs:407-412); 2D ghosts 3, 4, 5 copy 0, 1, 2 (:559-566)"""
    return sv + ([sv[0], sv[0], sv[0], sv[1]] if len(sv) == 4 else [sv[0], sv[1], sv[2]])


def test_trig_constants_are_the_literals():
    assert [bits(x) for x in O.trig_constants()] == [COS1, SIN1, COS2, SIN2]
    assert f(COS1) == -0.5 + 2.0 ** -52 and f(COS2) == -0.5 - 2.0 ** -51


@pytest.mark.parametrize("case", range(len(CASES)))
def test_oracle_super_vertices_hand_derived(case):
    P, c, r, sv = CASES[case]
    oc, orad = O.bounding_sphere(P)
    assert oc.tolist() == c and orad == r
    S = O.super_vertices(P)
    assert [[bits(x) for x in row] for row in S] == expected_with_ghosts(sv)


@pytest.mark.parametrize("case", range(len(CASES)))
def test_product_bounding_sphere_hand_derived(case):
    P, c, r, _ = CASES[case]
    dim = P.shape[1]
    cc = np.zeros(dim)
    rr = ctypes.c_double()
    _lib.check(_lib.lib().vo_bounding_sphere(dim, P.ctypes.data_as(_dp), P.shape[0], cc.ctypes.data_as(_dp),
                                             ctypes.byref(rr)))
    assert cc.tolist() == c and rr.value == r


# circumsphere_2d (src/geometry.rs:2-22) on triangles whose slopes and their negative reciprocals
# are powers of two, so that every step of the formula is exact:
#   (0,0), (2,4), (6,2): m_ab = 2, m_bc = -1/2, m_d = -1/2, m_e = 2, d = (1,2), e = (4,3)
#     x = (-1/2*1 - 2*4 + 3 - 2) / (-1/2 - 2) = -7.5 / -2.5 = 3, y = -1/2 (3 - 1) + 2 = 1, r = fl(sqrt 10)
#   (0,0), (0,2), (2,4): vertical first edge, m_ab = +inf, m_d = -0.0, m_bc = 1, m_e = -1, d = (0,1),
#     e = (1,3): x = (-0.0*0 + 1*1 + 3 - 1) / (-0.0 + 1) = 3, y = -0.0 (3 - 0) + 1 = 1, r = fl(sqrt 10)
#   (0,0), (4,0), (0,3): horizontal first edge, m_ab = 0, m_d = -inf, x = -inf / -inf = NaN
SQRT10 = "0x40094c583ada5b53"
TRI_2D = [
    ([[0.0, 0.0], [2.0, 4.0], [6.0, 2.0]], [3.0, 1.0], SQRT10),
    ([[0.0, 0.0], [0.0, 2.0], [2.0, 4.0]], [3.0, 1.0], SQRT10),
]


def product_circumsphere(v):
    v = np.ascontiguousarray(v, dtype=np.float64)
    dim = v.shape[1]
    c = np.zeros(dim)
    r = ctypes.c_double()
    rc = _lib.lib().vo_circumsphere(dim, v.ctypes.data_as(_dp), c.ctypes.data_as(_dp), ctypes.byref(r))
    return rc, c, r.value


@pytest.mark.parametrize("k", range(len(TRI_2D)))
def test_circumsphere_2d_exact_cases(k):
    v, c, r = TRI_2D[k]
    oc, orad = O.circumsphere(np.array(v))
    assert oc.tolist() == c and bits(orad) == r
    rc, pc, prad = product_circumsphere(v)
    assert rc == 0 and pc.tolist() == c and bits(prad) == r
    assert math.sqrt(10.0) == f(r)


def test_circumsphere_2d_horizontal_edge_is_nan_like_the_reference():
    v = np.array([[0.0, 0.0], [4.0, 0.0], [0.0, 3.0]])
    oc, orad = O.circumsphere(v)
    assert np.isnan(oc).all() and np.isnan(orad)
    rc, pc, prad = product_circumsphere(v)
    assert rc == 0 and np.isnan(pc).all() and np.isnan(prad)


def test_circumsphere_3d_reference_kat_product():
    # tests/test_geometry.rs:5-15: the unit corner tetrahedron
    rc, c, r = product_circumsphere([[0.0, 0, 0], [1.0, 0, 0], [0.0, 1, 0], [0.0, 0, 1]])
    assert rc == 0 and c.tolist() == [0.5, 0.5, 0.5] and r == 0.8660254037844386


def test_circumsphere_3d_singular_is_an_error():
    # the reference unwraps a failed LU solve (src/geometry.rs:49): coplanar points have no sphere
    rc, c, r = product_circumsphere([[0.0, 0, 0], [1.0, 0, 0], [0.0, 1, 0], [1.0, 1, 0]])
    assert rc == _lib.VO_ERR_INVALID and np.isnan(c).all()


@pytest.mark.parametrize("dim", [2, 3])
def test_product_circumsphere_equals_oracle_bitwise(dim):
    rng = np.random.default_rng(40 + dim)
    cases = [rng.random((dim + 1, dim)) for _ in range(400)]
    cases += [rng.random((dim + 1, dim)) * 10.0 ** rng.integers(-8, 8) + rng.random(dim) * 100 for _ in range(200)]
    # axis-aligned edges and repeated coordinates: pivots swap, slopes become 0 or infinite
    for _ in range(200):
        v = rng.integers(-3, 4, size=(dim + 1, dim)).astype(np.float64) / 2.0
        cases.append(v)
    n_cmp = 0
    for v in cases:
        oc, orad = O.circumsphere(v)
        rc, pc, prad = product_circumsphere(v)
        if rc != 0:
            continue                      # zero pivot (3D): the reference panics
        assert [bits(x) for x in pc] == [bits(x) for x in oc] and bits(prad) == bits(orad), v
        n_cmp += 1
    assert n_cmp > 600


def test_in_sphere_strict_boundary():
    lib = _lib.lib()
    c = np.array([0.0, 0.0, 0.0])
    on = np.array([3.0, 4.0, 0.0])               # distance^2 = 25 = r * r: not inside
    inside = np.array([3.0, 4.0 - 2.0 ** -50, 0.0])
    assert lib.vo_in_sphere(3, on.ctypes.data_as(_dp), c.ctypes.data_as(_dp), 5.0) == 0
    assert lib.vo_in_sphere(3, inside.ctypes.data_as(_dp), c.ctypes.data_as(_dp), 5.0) == 1
    rng = np.random.default_rng(9)
    for _ in range(500):
        p, q = rng.random(3), rng.random(3)
        r = float(rng.random())
        assert lib.vo_in_sphere(3, p.ctypes.data_as(_dp), q.ctypes.data_as(_dp), r) == O.lib().or_in_sphere(
            3, p.ctypes.data_as(_dp), q.ctypes.data_as(_dp), ctypes.c_double(r))
