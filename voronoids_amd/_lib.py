"""ctypes binding of libvoronoids_hip.so (include/voronoids.h).

Loading fails loudly when the library has not been built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# VORONOIDS_LIB names another build of the same library (e.g. a tuning variant)
LIB_PATH = os.environ.get("VORONOIDS_LIB") or os.path.join(HERE, "libvoronoids_hip.so")

VO_OK = 0
VO_ERR_INVALID = -1
VO_ERR_OOM = -2
VO_ERR_HIP = -3
VO_ERR_DUPLICATE = -4
VO_ERR_OVERFLOW = -5
VO_ERR_PRECISION = -6
VO_ERR_INTERNAL = -7

EXPORTED = [
    "vo_tri_create", "vo_tri_destroy", "vo_tri_insert", "vo_tri_stage", "vo_tri_run", "vo_tri_counts",
    "vo_tri_export", "vo_tri_canonical", "vo_free", "vo_tri_check", "vo_tri_stats", "vo_tri_config", "vo_tri_vertices",
    "vo_last_error", "vo_version",
    "vo_engine_create", "vo_engine_destroy", "vo_engine_submit", "vo_engine_step", "vo_engine_launch",
    "vo_engine_finish", "vo_engine_results", "vo_engine_stats_get", "vo_engine_config", "vo_batch_triangulate",
    "vo_result_canonical", "vo_clock",
    "vo_server_create", "vo_server_submit", "vo_server_results", "vo_server_stats_get", "vo_server_destroy",
    "vo_predicates", "vo_bounding_sphere",
    "vo_tri_check_pairs", "vo_tri_set_vertex", "vo_tri_swap_simplex_vertices",
    "vo_circumsphere", "vo_circumspheres", "vo_in_sphere", "vo_predicates_noalloc", "vo_tri_locate",
    "vo_tri_reset",
]


class VoSetResult(ctypes.Structure):
    _fields_ = [
        ("tag", ctypes.c_int64), ("dim", ctypes.c_int32), ("status", ctypes.c_int32),
        ("n_points", ctypes.c_int64), ("points_duplicate", ctypes.c_int64), ("n_simplices", ctypes.c_int64),
        ("simplices", ctypes.POINTER(ctypes.c_int32)),
        ("t_submit", ctypes.c_double), ("t_admit", ctypes.c_double), ("t_done", ctypes.c_double),
        ("rounds", ctypes.c_int64),
    ]


class VoEngineStats(ctypes.Structure):
    _fields_ = [(name, ctypes.c_int64) for name in (
        "sets_submitted", "sets_admitted", "sets_done", "sets_resident", "sets_queued",
        "points_inserted", "rounds", "replays", "solo_runs", "range_extensions", "admission_waits",
        "vertex_capacity", "simplex_capacity", "simplex_in_use", "max_sets_resident")] + [
        ("ms_replays", ctypes.c_double), ("ms_boundary", ctypes.c_double), ("ms_admit", ctypes.c_double),
        ("ms_collect", ctypes.c_double)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class VoStats(ctypes.Structure):
    _fields_ = [
        ("rounds", ctypes.c_int64), ("points_inserted", ctypes.c_int64), ("points_duplicate", ctypes.c_int64),
        ("windows_total", ctypes.c_int64), ("winners_total", ctypes.c_int64),
        ("tier1", ctypes.c_int64 * 2), ("tier2", ctypes.c_int64 * 2), ("tier3", ctypes.c_int64 * 2),
        ("deferred", ctypes.c_int64 * 2), ("exact_evaluated", ctypes.c_int64),
        ("walk_steps", ctypes.c_int64), ("chain_hops", ctypes.c_int64), ("killed", ctypes.c_int64),
        ("created", ctypes.c_int64), ("max_killed", ctypes.c_int64), ("max_facets", ctypes.c_int64),
        ("ms_run", ctypes.c_double), ("ms_locate_cavity", ctypes.c_double), ("ms_commit", ctypes.c_double),
        ("ms_exact", ctypes.c_double), ("ms_other", ctypes.c_double),
        ("launches_locate_cavity", ctypes.c_int64), ("launches_commit", ctypes.c_int64),
        ("solo_runs", ctypes.c_int64), ("duplicates_detected", ctypes.c_int64),
        ("compactions", ctypes.c_int64), ("ms_compact", ctypes.c_double), ("records", ctypes.c_int64),
    ]

    def as_dict(self) -> dict:
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if isinstance(v, ctypes.Array) else v
        return out


class VoError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"voronoids error {code}: {msg}")
        self.code = code


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built: run `python -m voronoids_amd.build` (hipcc, gfx950)")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        dp = ctypes.POINTER(ctypes.c_double)
        i32p = ctypes.POINTER(ctypes.c_int32)
        i64p = ctypes.POINTER(ctypes.c_int64)
        L.vo_tri_create.argtypes = [ctypes.c_int, ctypes.c_int, dp, ctypes.c_int64, ctypes.POINTER(P)]
        L.vo_bounding_sphere.argtypes = [ctypes.c_int, dp, ctypes.c_int64, dp, dp]
        L.vo_tri_destroy.argtypes = [P]
        L.vo_tri_reset.argtypes = [P, dp, ctypes.c_int64]
        L.vo_tri_destroy.restype = None
        for name in ("vo_tri_insert", "vo_tri_stage"):
            getattr(L, name).argtypes = [P, dp, ctypes.c_int64, ctypes.c_uint64]
        L.vo_tri_run.argtypes = [P]
        L.vo_tri_counts.argtypes = [P, i64p, i64p, i64p]
        L.vo_tri_export.argtypes = [P, i32p, i32p]
        L.vo_tri_canonical.argtypes = [P, ctypes.POINTER(i32p), i64p, ctypes.POINTER(i32p), i64p]
        L.vo_free.argtypes = [ctypes.c_void_p]
        L.vo_free.restype = None
        L.vo_tri_check.argtypes = [P, i64p]
        L.vo_tri_check_pairs.argtypes = [P, i64p]
        L.vo_tri_locate.argtypes = [P, dp, i32p, ctypes.c_int64, i64p, i32p, ctypes.c_int64, i64p]
        L.vo_tri_set_vertex.argtypes = [P, ctypes.c_int64, dp, ctypes.c_int]
        L.vo_tri_swap_simplex_vertices.argtypes = [P, ctypes.c_int64, ctypes.c_int, ctypes.c_int]
        L.vo_circumsphere.argtypes = [ctypes.c_int, dp, dp, dp]
        L.vo_circumspheres.argtypes = [ctypes.c_int, dp, ctypes.c_int64, i32p, ctypes.c_int64, dp, dp]
        L.vo_in_sphere.argtypes = [ctypes.c_int, dp, dp, ctypes.c_double]
        L.vo_tri_stats.argtypes = [P, ctypes.POINTER(VoStats)]
        L.vo_tri_vertices.argtypes = [P, dp]
        L.vo_tri_config.argtypes = [P, ctypes.c_char_p, ctypes.c_int64]
        res_p = ctypes.POINTER(VoSetResult)
        L.vo_engine_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int64, ctypes.POINTER(P)]
        L.vo_engine_destroy.argtypes = [P]
        L.vo_engine_destroy.restype = None
        L.vo_engine_submit.argtypes = [P, dp, ctypes.c_int64, ctypes.c_uint64, ctypes.c_int64]
        L.vo_engine_step.argtypes = [P, i64p]
        L.vo_engine_launch.argtypes = [P]
        L.vo_engine_finish.argtypes = [P, i64p]
        L.vo_engine_results.argtypes = [P, res_p, ctypes.c_int64, i64p]
        L.vo_engine_stats_get.argtypes = [P, ctypes.POINTER(VoEngineStats)]
        L.vo_engine_config.argtypes = [P, ctypes.c_char_p, ctypes.c_int64]
        L.vo_batch_triangulate.argtypes = [P, dp, i64p, ctypes.c_int64, ctypes.c_uint64, res_p]
        L.vo_result_canonical.argtypes = [res_p, ctypes.POINTER(i32p), i64p, ctypes.POINTER(i32p), i64p]
        L.vo_clock.restype = ctypes.c_double
        L.vo_server_create.argtypes = [ctypes.c_int, ctypes.c_int64, ctypes.c_int64, ctypes.POINTER(P)]
        L.vo_server_submit.argtypes = [P, ctypes.c_int, dp, ctypes.c_int64, ctypes.c_uint64, ctypes.c_int64]
        L.vo_server_results.argtypes = [P, res_p, ctypes.c_int64, i64p, ctypes.c_int]
        L.vo_server_stats_get.argtypes = [P, ctypes.c_int, ctypes.POINTER(VoEngineStats)]
        L.vo_server_destroy.argtypes = [P]
        L.vo_server_destroy.restype = None
        L.vo_predicates.argtypes = [ctypes.c_int, ctypes.c_int, dp, ctypes.c_int64, dp, i32p, i32p]
        L.vo_predicates_noalloc.argtypes = [ctypes.c_int, ctypes.c_int, dp, ctypes.c_int64, dp, i32p, i32p]
        L.vo_last_error.restype = ctypes.c_char_p
        L.vo_version.restype = ctypes.c_int
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != VO_OK:
        msg = lib().vo_last_error()
        raise VoError(rc, msg.decode() if msg else "")
