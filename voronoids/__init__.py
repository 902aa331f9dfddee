"""voronoids -- the reference's Python module name (src/lib.rs:127-134) over the MI355X build.

``import voronoids; voronoids.delaunay(pts)`` runs as in the reference's README: the module
exposes ``delaunay`` (src/lib.rs:104-125), ``PyDelauanyTree`` (:62-102, the typo is the
reference's), ``PyVertex`` (:12-29) and ``PySimplex`` (:31-60), all from ``voronoids_amd``,
whose construction runs on the GPU through libvoronoids_hip.so.  ``delaunay`` accepts 3D points
as the reference does, and 2D points as well.  The batched and served modes (``batch_triangulate``,
``Engine``, ``Server``) have no counterpart in the reference.
"""
from voronoids_amd import (DelaunayTree, Engine, PyDelauanyTree, PySimplex, PyVertex, Server, SetResult, VoError,
                           batch_triangulate, delaunay)

__all__ = ["delaunay", "PyDelauanyTree", "PyVertex", "PySimplex", "DelaunayTree", "VoError",
           "batch_triangulate", "Engine", "Server", "SetResult"]
