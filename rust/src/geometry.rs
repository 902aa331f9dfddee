//! geometry (src/geometry.rs) through the library's host routines, which restate the reference's
//! formulas bit for bit: `circumsphere` (:58-88, the 2D slope formula :2-22 and the 3D LU solve in
//! nalgebra 0.32's order :24-56), `in_sphere` (:91-97) and `bounding_sphere` (:99-141).
use std::os::raw::c_int;

use crate::error::check;
use crate::ffi;

/// Center and radius of the sphere through the first N + 1 vertices (N = 2 or 3), as
/// src/geometry.rs:58-88 computes them.  Other sizes give the zero sphere, as there.  A 3D simplex
/// whose LU solve meets a zero pivot panics (the reference unwraps the solve).
pub fn circumsphere<const N: usize, const M: usize>(vertices: [[f64; N]; M]) -> ([f64; N], f64) {
    let mut center = [0.0; N];
    if !((N == 2 && M >= 3) || (N == 3 && M >= 4)) {
        return (center, 0.0);
    }
    let mut radius = 0.0;
    // SAFETY: [[f64; N]; M] is M * N contiguous f64 and the library reads the first (N + 1) * N;
    // center holds N f64.
    let rc = unsafe {
        ffi::vo_circumsphere(N as c_int, vertices.as_ptr().cast::<f64>(), center.as_mut_ptr(), &mut radius)
    };
    if let Err(e) = check(rc) {
        panic!("circumsphere of {:?}: {}", vertices, e);
    }
    (center, radius)
}

/// Strictly inside: sum_k (center_k - vertex_k)^2 < radius^2, summed in axis order with no fused
/// multiply-add (src/geometry.rs:91-97).
pub fn in_sphere<const N: usize>(vertex: [f64; N], center: [f64; N], radius: f64) -> bool {
    if N == 2 || N == 3 {
        // SAFETY: both arrays hold N f64.
        return unsafe { ffi::vo_in_sphere(N as c_int, vertex.as_ptr(), center.as_ptr(), radius) } == 1;
    }
    // other dimensions: the same sum here (Rust never contracts a * b + c into a fused operation)
    let mut distance: f64 = 0.0;
    for k in 0..N {
        let d = center[k] - vertex[k];
        distance += d * d;
    }
    distance < radius * radius
}

/// The sphere DelaunayTree::new builds its super simplex around (src/geometry.rs:99-141): the
/// per-axis min / max midpoint, then the radius grown to 1.5x the distance of each point, in
/// input order, that is not strictly inside.  N = 2 or 3; panics on an empty input.
pub fn bounding_sphere<const N: usize>(vertices: Vec<[f64; N]>) -> ([f64; N], f64) {
    assert!(N == 2 || N == 3, "bounding_sphere: N must be 2 or 3");
    let mut c = [0.0f64; 3];
    let mut radius = 0.0;
    // SAFETY: Vec<[f64; N]> is len * N contiguous f64; c holds 3 >= N f64.
    let rc = unsafe {
        ffi::vo_bounding_sphere(
            N as c_int,
            vertices.as_ptr().cast::<f64>(),
            vertices.len() as i64,
            c.as_mut_ptr(),
            &mut radius,
        )
    };
    if let Err(e) = check(rc) {
        panic!("bounding_sphere: {}", e);
    }
    let mut center = [0.0; N];
    center.copy_from_slice(&c[..N]);
    (center, radius)
}
