//! voronoids over libvoronoids_hip: the public Rust API of kazewong/Voronoids (src/lib.rs:3-5:
//! `delaunay_tree`, `geometry`, `scheduler`) with the triangulation built and held on an MI355X
//! (gfx950) by the library's HIP kernels.  The crate's library name is `voronoids`, so code written
//! against the reference keeps its `use voronoids::...` lines; what differs is listed in the
//! `delaunay_tree` module's documentation.
pub mod delaunay_tree;
pub mod error;
pub mod ffi;
pub mod geometry;
pub mod scheduler;

pub use delaunay_tree::{Cavity, DelaunayTree, Simplex, TreeUpdate, Vertex};
pub use error::VoError;

/// voronoids.delaunay (src/lib.rs:104-125) for Rust callers: DelaunayTree::new(points), then every
/// point inserted.
pub fn delaunay(points: Vec<[f64; 3]>) -> DelaunayTree<3, 4> {
    let mut tree = DelaunayTree::<3, 4>::new(points.clone());
    tree.add_points_to_tree(points);
    tree
}

/// the library's version number (vo_version)
pub fn library_version() -> i32 {
    // SAFETY: no arguments.
    unsafe { ffi::vo_version() }
}
