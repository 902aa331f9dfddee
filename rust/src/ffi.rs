//! The C ABI of libvoronoids_hip (include/voronoids.h), the declarations this crate calls.
//!
//! Every function returns 0 or a negative `VO_ERR_*` code (vo_in_sphere: 1 / 0) and never
//! aborts; `vo_last_error` holds a thread-local message for the last failure.  Host buffers are
//! copied during the call and never retained; the library owns the device memory behind a
//! `vo_tri` handle.
#![allow(non_camel_case_types)]

use std::os::raw::{c_char, c_int, c_void};

/// Opaque handle: one triangulation resident on one HIP device.
#[repr(C)]
pub struct vo_tri {
    _private: [u8; 0],
}

/// `vo_stats` of include/voronoids.h, field for field (every field is 8 bytes, so the layout has
/// no padding).
#[repr(C)]
#[derive(Debug, Default, Clone, Copy)]
pub struct vo_stats {
    pub rounds: i64,
    pub points_inserted: i64,
    pub points_duplicate: i64,
    pub windows_total: i64,
    pub winners_total: i64,
    /// predicate calls resolved per tier; index 0 = orient, 1 = in-sphere / in-circle
    pub tier1: [i64; 2],
    pub tier2: [i64; 2],
    pub tier3: [i64; 2],
    pub deferred: [i64; 2],
    pub exact_evaluated: i64,
    pub walk_steps: i64,
    pub chain_hops: i64,
    pub killed: i64,
    pub created: i64,
    pub max_killed: i64,
    pub max_facets: i64,
    /// device time of the last run, ms
    pub ms_run: f64,
    pub ms_locate_cavity: f64,
    pub ms_commit: f64,
    pub ms_exact: f64,
    pub ms_other: f64,
    pub launches_locate_cavity: i64,
    pub launches_commit: i64,
    pub solo_runs: i64,
    pub duplicates_detected: i64,
    pub compactions: i64,
    pub ms_compact: f64,
    pub records: i64,
}

extern "C" {
    pub fn vo_tri_create(hip_device: c_int, dim: c_int, pts: *const f64, n: i64, out: *mut *mut vo_tri) -> c_int;
    pub fn vo_tri_destroy(t: *mut vo_tri);
    pub fn vo_tri_reset(t: *mut vo_tri, pts: *const f64, n: i64) -> c_int;
    pub fn vo_tri_insert(t: *mut vo_tri, pts: *const f64, n: i64, seed: u64) -> c_int;
    pub fn vo_tri_stage(t: *mut vo_tri, pts: *const f64, n: i64, seed: u64) -> c_int;
    pub fn vo_tri_run(t: *mut vo_tri) -> c_int;
    pub fn vo_tri_counts(t: *const vo_tri, n_vertices: *mut i64, n_simplices: *mut i64, max_simplex_id: *mut i64)
        -> c_int;
    pub fn vo_tri_export(t: *const vo_tri, simplices: *mut i32, neighbors: *mut i32) -> c_int;
    pub fn vo_tri_canonical(t: *const vo_tri, tris: *mut *mut i32, n_t: *mut i64, edges: *mut *mut i32, n_e: *mut i64)
        -> c_int;
    pub fn vo_free(p: *mut c_void);
    pub fn vo_tri_check(t: *mut vo_tri, n_violations: *mut i64) -> c_int;
    pub fn vo_tri_locate(
        t: *mut vo_tri,
        p: *const f64,
        killed: *mut i32,
        killed_cap: i64,
        n_killed: *mut i64,
        facets: *mut i32,
        facets_cap: i64,
        n_facets: *mut i64,
    ) -> c_int;
    pub fn vo_tri_check_pairs(t: *mut vo_tri, n_pairs: *mut i64) -> c_int;
    pub fn vo_tri_set_vertex(t: *mut vo_tri, id: i64, coords: *const f64, records: c_int) -> c_int;
    pub fn vo_tri_swap_simplex_vertices(t: *mut vo_tri, simplex: i64, i: c_int, j: c_int) -> c_int;
    pub fn vo_tri_stats(t: *const vo_tri, out: *mut vo_stats) -> c_int;
    pub fn vo_tri_config(t: *mut vo_tri, key: *const c_char, value: i64) -> c_int;
    pub fn vo_tri_vertices(t: *const vo_tri, coords: *mut f64) -> c_int;
    pub fn vo_last_error() -> *const c_char;
    pub fn vo_version() -> c_int;
    pub fn vo_bounding_sphere(dim: c_int, pts: *const f64, n: i64, center: *mut f64, radius: *mut f64) -> c_int;
    pub fn vo_circumsphere(dim: c_int, verts: *const f64, center: *mut f64, radius: *mut f64) -> c_int;
    pub fn vo_circumspheres(
        dim: c_int,
        coords: *const f64,
        n_coords: i64,
        simplices: *const i32,
        n: i64,
        centers: *mut f64,
        radii: *mut f64,
    ) -> c_int;
    pub fn vo_in_sphere(dim: c_int, p: *const f64, center: *const f64, radius: f64) -> c_int;
}
