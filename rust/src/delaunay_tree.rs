//! delaunay_tree (src/delaunay_tree.rs) over libvoronoids_hip: the triangulation lives in the
//! memory of one HIP device and every operation is a call through the C ABI.
//!
//! What a caller of the reference finds here, and what differs:
//! - `Simplex`, `Vertex`, `TreeUpdate` with the reference's fields and `DelaunayTree::new`,
//!   `locate`, `get_new_simplices`, `insert_point`, `insert_points_parallel`,
//!   `add_points_to_tree`, `check_delaunay` and the `max_simplex_id` field, with the reference's
//!   signatures (src/delaunay_tree.rs:24-386, :389-541, :544-671, :698-740).
//! - The reference's `vertices` and `simplices` fields are DashMaps in host memory.  Here the
//!   state is on the device, so they are methods that return a snapshot (`HashMap`s keyed the same
//!   way); `num_vertices()` is `vertices.len()` without the snapshot.  There is no `kdtree` field:
//!   the device locates points by walking.
//! - Simplex ids: ghost simplices keep the reference's ids 1..=N+1, vertex order, zero sphere and
//!   single neighbor (:458-502, :606-632); the super simplex is id 0 until the first insertion;
//!   every other live simplex is N + 2 + its position in the library's export order.  The ids of a
//!   snapshot hold until the tree changes (the reference's ids survive insertions).
//! - `add_points_to_tree` inserts in a random priority order drawn from a seed (0 unless
//!   `try_add_points_to_tree` is given another); the triangulation does not depend on it.
//! - Vertex order within a new simplex is positively oriented, so `get_new_simplices` lists the
//!   same vertex sets as the reference in another order (new vertex first in both).
use std::collections::HashMap;
use std::ffi::CString;
use std::os::raw::{c_int, c_void};
use std::ptr;

use crate::error::{check, VoError, VO_ERR_INVALID, VO_ERR_OVERFLOW};
use crate::ffi;
use crate::geometry::circumsphere;

/// struct Simplex (src/delaunay_tree.rs:10-16)
#[derive(Debug, Clone, PartialEq)]
pub struct Simplex<const N: usize, const M: usize> {
    pub vertices: [usize; M],
    pub center: [f64; N],
    pub radius: f64,
    pub neighbors: Vec<usize>,
}

/// struct Vertex (src/delaunay_tree.rs:18-22)
#[derive(Debug, Clone, PartialEq)]
pub struct Vertex<const N: usize> {
    pub coordinates: [f64; N],
    pub simplex: Vec<usize>,
}

/// The cavity of a point against the tree as it is (locate + get_new_simplices,
/// src/delaunay_tree.rs:33-123, in one device call): the ids of the simplices whose circumsphere
/// strictly contains the point, sorted, and one entry per boundary facet: (cavity simplex id, id of
/// the simplex across the facet -- a ghost id on the super-simplex hull --, the facet's N vertex
/// ids, ordered so that [new vertex, facet...] is positively oriented).
#[derive(Debug, Clone, PartialEq)]
pub struct Cavity {
    pub killed: Vec<usize>,
    pub facets: Vec<(usize, usize, Vec<usize>)>,
}

/// Ghost simplices of DelaunayTree::new (src/delaunay_tree.rs:467-502 (3D), :606-632 (2D)): the
/// ghost vertex, then the super-hull face it guards; ghost k has simplex id k + 1.
fn ghost_simplices(dim: usize) -> Vec<Vec<usize>> {
    if dim == 3 {
        vec![vec![4, 0, 1, 2], vec![5, 0, 2, 3], vec![6, 0, 3, 1], vec![7, 1, 2, 3]]
    } else {
        vec![vec![3, 0, 1], vec![4, 0, 2], vec![5, 1, 2]]
    }
}

/// the id of the ghost simplex guarding a super-hull face
fn ghost_of(ghosts: &[Vec<usize>], face: &[usize]) -> usize {
    let mut f = face.to_vec();
    f.sort_unstable();
    for (k, g) in ghosts.iter().enumerate() {
        let mut h = g[1..].to_vec();
        h.sort_unstable();
        if h == f {
            return k + 1;
        }
    }
    panic!("hull facet {:?} is not a face of the super simplex", face);
}

/// New simplices that share a ridge (N common vertices): both orders of every such pair, pairs
/// in (i, j) order with i < j, as src/delaunay_tree.rs:674-695 lists them.  Each simplex's M faces
/// are hashed, so this is linear in the number of simplices.
fn pair_simplices<const M: usize>(simplices: &[[usize; M]], simplices_id: &[usize]) -> Vec<(usize, usize)> {
    let mut by_face: HashMap<Vec<usize>, Vec<usize>> = HashMap::new();
    for (i, s) in simplices.iter().enumerate() {
        for skip in 0..M {
            let mut face: Vec<usize> = (0..M).filter(|&k| k != skip).map(|k| s[k]).collect();
            face.sort_unstable();
            by_face.entry(face).or_default().push(i);
        }
    }
    let mut pairs: Vec<(usize, usize)> = Vec::new();
    for owners in by_face.values() {
        for a in 0..owners.len() {
            for b in (a + 1)..owners.len() {
                pairs.push((owners[a].min(owners[b]), owners[a].max(owners[b])));
            }
        }
    }
    pairs.sort_unstable();
    pairs.dedup();
    let mut out = Vec::with_capacity(2 * pairs.len());
    for (i, j) in pairs {
        out.push((simplices_id[i], simplices_id[j]));
        out.push((simplices_id[j], simplices_id[i]));
    }
    out
}

/// struct DelaunayTree (src/delaunay_tree.rs:24-30) on one HIP device.  M must be N + 1 and N 2 or
/// 3.  A tree may move between threads; it is used by one at a time (`&mut self` for every change,
/// and the handle is not shared).
pub struct DelaunayTree<const N: usize, const M: usize> {
    handle: *mut ffi::vo_tri,
    /// host copy of every vertex's coordinates by id: 2 (N + 1) super and ghost vertices, then each
    /// point handed in (the library keeps the same table on the device)
    coords: Vec<[f64; N]>,
    /// the sphere new() stores on simplex 0: (bounding-sphere center, 10 x radius)
    root_sphere: ([f64; N], f64),
    /// the max_simplex_id field (src/delaunay_tree.rs:29): N + 1 after new, then grows with every
    /// simplex created
    pub max_simplex_id: usize,
}

// SAFETY: the handle is owned by this value alone and the library allows a handle to be used from
// any thread, one call at a time; `&self` methods that reach the device are not Sync-safe, and the
// raw pointer keeps the type !Sync.
unsafe impl<const N: usize, const M: usize> Send for DelaunayTree<N, M> {}

impl<const N: usize, const M: usize> Drop for DelaunayTree<N, M> {
    fn drop(&mut self) {
        if !self.handle.is_null() {
            // SAFETY: the handle came from vo_tri_create and is destroyed once.
            unsafe { ffi::vo_tri_destroy(self.handle) };
            self.handle = ptr::null_mut();
        }
    }
}

impl DelaunayTree<3, 4> {
    /// DelaunayTree::<3,4>::new (src/delaunay_tree.rs:389-510): the super simplex around the
    /// bounding sphere of `vertices` (radius x 10) on HIP device 0; no point is inserted.
    pub fn new(vertices: Vec<[f64; 3]>) -> Self {
        Self::create(&vertices, 0).unwrap_or_else(|e| panic!("DelaunayTree::new: {}", e))
    }
    pub fn new_on_device(vertices: &[[f64; 3]], hip_device: i32) -> Result<Self, VoError> {
        Self::create(vertices, hip_device)
    }
}

impl DelaunayTree<2, 3> {
    /// DelaunayTree::<2,3>::new (src/delaunay_tree.rs:544-640)
    pub fn new(vertices: Vec<[f64; 2]>) -> Self {
        Self::create(&vertices, 0).unwrap_or_else(|e| panic!("DelaunayTree::new: {}", e))
    }
    pub fn new_on_device(vertices: &[[f64; 2]], hip_device: i32) -> Result<Self, VoError> {
        Self::create(vertices, hip_device)
    }
}

impl<const N: usize, const M: usize> DelaunayTree<N, M> {
    fn create(vertices: &[[f64; N]], hip_device: i32) -> Result<Self, VoError> {
        assert!((N == 2 || N == 3) && M == N + 1, "DelaunayTree is built for <2, 3> and <3, 4>");
        let mut handle: *mut ffi::vo_tri = ptr::null_mut();
        // SAFETY: vertices is len * N contiguous f64; handle receives the new tree.
        check(unsafe {
            ffi::vo_tri_create(
                hip_device as c_int,
                N as c_int,
                vertices.as_ptr().cast::<f64>(),
                vertices.len() as i64,
                &mut handle,
            )
        })?;
        let mut tree = Self { handle, coords: Vec::new(), root_sphere: ([0.0; N], 10.0), max_simplex_id: 0 };
        // from here on, an error drops `tree`, which destroys the handle
        let mut super_vertices = vec![[0.0f64; N]; 2 * M];
        // SAFETY: with no point staged the library writes the 2 (N + 1) super and ghost rows.
        check(unsafe { ffi::vo_tri_vertices(tree.handle, super_vertices.as_mut_ptr().cast::<f64>()) })?;
        tree.coords = super_vertices;
        if !vertices.is_empty() {
            let mut c = [0.0f64; 3];
            let mut r = 0.0f64;
            // SAFETY: as above; c holds 3 >= N f64.
            check(unsafe {
                ffi::vo_bounding_sphere(
                    N as c_int,
                    vertices.as_ptr().cast::<f64>(),
                    vertices.len() as i64,
                    c.as_mut_ptr(),
                    &mut r,
                )
            })?;
            let mut center = [0.0; N];
            center.copy_from_slice(&c[..N]);
            tree.root_sphere = (center, 10.0 * if r > 0.0 { r } else { 1.0 });
        }
        tree.refresh_max_simplex_id()?;
        Ok(tree)
    }

    /// (vertices incl. super vertices, live simplices, max_simplex_id) as the library counts them
    fn counts(&self) -> Result<(i64, i64, i64), VoError> {
        let (mut v, mut s, mut m) = (0i64, 0i64, 0i64);
        // SAFETY: three valid out-pointers.
        check(unsafe { ffi::vo_tri_counts(self.handle, &mut v, &mut s, &mut m) })?;
        Ok((v, s, m))
    }

    fn refresh_max_simplex_id(&mut self) -> Result<(), VoError> {
        self.max_simplex_id = self.counts()?.2 as usize;
        Ok(())
    }

    /// copy points to the device; each takes the next vertex id (also one the run later finds to
    /// coincide with a vertex)
    fn stage(&mut self, pts: &[[f64; N]], seed: u64) -> Result<(), VoError> {
        // SAFETY: pts is len * N contiguous f64.
        check(unsafe { ffi::vo_tri_stage(self.handle, pts.as_ptr().cast::<f64>(), pts.len() as i64, seed) })?;
        self.coords.extend_from_slice(pts);
        Ok(())
    }

    fn run(&mut self) -> Result<(), VoError> {
        // SAFETY: a live handle.
        let rc = check(unsafe { ffi::vo_tri_run(self.handle) });
        self.refresh_max_simplex_id()?;
        rc
    }

    /// vertices.len() of the reference: 2 (N + 1) super and ghost vertices plus every point
    /// handed in
    pub fn num_vertices(&self) -> usize {
        self.coords.len()
    }

    /// live simplices (the super simplex's ghosts not counted)
    pub fn num_simplices(&self) -> usize {
        self.counts().map(|c| c.1 as usize).unwrap_or_else(|e| panic!("num_simplices: {}", e))
    }

    // -- construction ----------------------------------------------------------------------

    /// add_points_to_tree (src/delaunay_tree.rs:336-386): every point inserted, with vertex ids
    /// num_vertices() .. + vertices.len() - 1 in input order.
    pub fn add_points_to_tree(&mut self, vertices: Vec<[f64; N]>) {
        self.try_add_points_to_tree(&vertices, 0).unwrap_or_else(|e| panic!("add_points_to_tree: {}", e));
    }

    /// add_points_to_tree with the seed of the random priority order (the result does not depend
    /// on it); a point that coincides with a vertex gives VO_ERR_DUPLICATE after the others are in.
    pub fn try_add_points_to_tree(&mut self, vertices: &[[f64; N]], seed: u64) -> Result<(), VoError> {
        if vertices.is_empty() {
            return Ok(());
        }
        self.stage(vertices, seed)?;
        self.run()
    }

    /// insert_point (src/delaunay_tree.rs:125-211) of an update made for this tree whose id is the
    /// next free vertex id
    pub fn insert_point(&mut self, update: &TreeUpdate<N, M>) {
        self.try_insert_point(update).unwrap_or_else(|e| panic!("insert_point: {}", e));
    }

    pub fn try_insert_point(&mut self, update: &TreeUpdate<N, M>) -> Result<(), VoError> {
        if update.tree != self.handle as usize {
            return Err(VoError::invalid("insert_point: the update was made for another tree".to_string()));
        }
        if update.id != self.num_vertices() {
            return Err(VoError::invalid(format!(
                "insert_point: vertex id {} is not the next free id {}",
                update.id,
                self.num_vertices()
            )));
        }
        self.stage(std::slice::from_ref(&update.vertex), 0)?;
        self.run()
    }

    /// insert_points_parallel (src/delaunay_tree.rs:213-334): the updates' points inserted together,
    /// with ids num_vertices() .. in order.  The reference expects updates whose cavities do not
    /// overlap (one level of find_placement); here any set of updates is valid.
    #[allow(clippy::ptr_arg)]
    pub fn insert_points_parallel(&mut self, updates: &Vec<TreeUpdate<N, M>>) {
        self.try_insert_points_parallel(updates, 0).unwrap_or_else(|e| panic!("insert_points_parallel: {}", e));
    }

    pub fn try_insert_points_parallel(&mut self, updates: &[TreeUpdate<N, M>], seed: u64) -> Result<(), VoError> {
        if updates.is_empty() {
            return Ok(());
        }
        let base = self.num_vertices();
        let mut pts: Vec<[f64; N]> = Vec::with_capacity(updates.len());
        for (i, u) in updates.iter().enumerate() {
            if u.tree != self.handle as usize {
                return Err(VoError::invalid(format!("insert_points_parallel: update {} was made for another tree", i)));
            }
            if u.id != base + i {
                return Err(VoError::invalid(format!(
                    "insert_points_parallel: update {} has vertex id {}, not {}",
                    i,
                    u.id,
                    base + i
                )));
            }
            pts.push(u.vertex);
        }
        self.stage(&pts, seed)?;
        self.run()
    }

    // -- queries ---------------------------------------------------------------------------

    /// locate (src/delaunay_tree.rs:33-58): the ids of the simplices whose circumsphere strictly
    /// contains `vertex`, sorted.  Panics, as the reference does, when `vertex` coincides with a
    /// vertex of the tree.
    pub fn locate(&self, vertex: [f64; N]) -> Vec<usize> {
        match self.cavity(&vertex) {
            Ok(c) => c.killed,
            Err(e) => panic!("No simplex found for vertex {:?}: {}", vertex, e),
        }
    }

    /// The cavity of `vertex` (not yet in the tree) with exact signs; nothing changes.
    pub fn cavity(&self, vertex: &[f64; N]) -> Result<Cavity, VoError> {
        let root_only = self.counts()?.1 == 1 && self.only_the_super_simplex()?;
        let mut cap_k: usize = 64;
        let mut cap_f: usize = 128;
        loop {
            let mut killed = vec![0i32; cap_k];
            let mut facets = vec![0i32; cap_f * (M + 1)];
            let (mut nk, mut nf) = (0i64, 0i64);
            // SAFETY: killed holds cap_k ints and facets cap_f rows of N + 2 ints; the library
            // writes rows only when both fit.
            let rc = unsafe {
                ffi::vo_tri_locate(
                    self.handle,
                    vertex.as_ptr(),
                    killed.as_mut_ptr(),
                    cap_k as i64,
                    &mut nk,
                    facets.as_mut_ptr(),
                    cap_f as i64,
                    &mut nf,
                )
            };
            if rc == VO_ERR_OVERFLOW && (nk as usize > cap_k || nf as usize > cap_f) {
                cap_k = cap_k.max(nk as usize);
                cap_f = cap_f.max(nf as usize);
                continue;
            }
            check(rc)?;
            let id = |pos: i32| -> usize {
                if root_only {
                    0
                } else {
                    N + 2 + pos as usize
                }
            };
            let ghosts = ghost_simplices(N);
            let killed_ids: Vec<usize> = killed[..nk as usize].iter().map(|&p| id(p)).collect();
            let mut rows: Vec<(usize, usize, Vec<usize>)> = Vec::with_capacity(nf as usize);
            for r in facets.chunks_exact(M + 1).take(nf as usize) {
                let face: Vec<usize> = r[2..].iter().map(|&v| v as usize).collect();
                let across = if r[1] >= 0 { id(r[1]) } else { ghost_of(&ghosts, &face) };
                rows.push((id(r[0]), across, face));
            }
            // grouped by killed simplex, as the reference walks the killed sites in order
            rows.sort_by_key(|row| row.0);
            return Ok(Cavity { killed: killed_ids, facets: rows });
        }
    }

    /// true while the only live simplex is the super simplex (no point inserted yet)
    fn only_the_super_simplex(&self) -> Result<bool, VoError> {
        let mut s = vec![0i32; M];
        // SAFETY: one live simplex: M vertex ids; neighbors not requested.
        check(unsafe { ffi::vo_tri_export(self.handle, s.as_mut_ptr(), ptr::null_mut()) })?;
        Ok(s.iter().all(|&v| (v as usize) < M))
    }

    /// the simplices inserting `vertex` as `vertex_id` creates on the facets of the cavity simplex
    /// `killed_site_id`, with their circumspheres and (simplex across the facet, killed_site_id)
    /// (get_new_simplices, src/delaunay_tree.rs:77-123).  Empty when `killed_site_id` is not in the
    /// cavity of `vertex`.
    #[allow(clippy::type_complexity)]
    pub fn get_new_simplices(
        &self,
        killed_site_id: usize,
        vertex: [f64; N],
        vertex_id: usize,
    ) -> (Vec<[usize; M]>, Vec<[f64; N]>, Vec<f64>, Vec<(usize, usize)>) {
        let cavity = self.cavity(&vertex).unwrap_or_else(|e| panic!("get_new_simplices: {}", e));
        self.new_simplices(&cavity, Some(killed_site_id), vertex, vertex_id)
    }

    #[allow(clippy::type_complexity)]
    fn new_simplices(
        &self,
        cavity: &Cavity,
        only: Option<usize>,
        vertex: [f64; N],
        vertex_id: usize,
    ) -> (Vec<[usize; M]>, Vec<[f64; N]>, Vec<f64>, Vec<(usize, usize)>) {
        let mut simplices: Vec<[usize; M]> = Vec::new();
        let mut centers: Vec<[f64; N]> = Vec::new();
        let mut radii: Vec<f64> = Vec::new();
        let mut neighbors: Vec<(usize, usize)> = Vec::new();
        for (killed, across, face) in cavity.facets.iter() {
            if only.map_or(false, |k| k != *killed) {
                continue;
            }
            let mut s = [0usize; M];
            s[0] = vertex_id;
            s[1..].copy_from_slice(face);
            let mut pts = [[0.0f64; N]; M];
            pts[0] = vertex;
            for i in 1..M {
                pts[i] = self.coords[s[i]];
            }
            let (center, radius) = circumsphere(pts);
            simplices.push(s);
            centers.push(center);
            radii.push(radius);
            neighbors.push((*across, *killed));
        }
        (simplices, centers, radii, neighbors)
    }

    /// check_delaunay (src/delaunay_tree.rs:512-541, :642-671): no vertex strictly inside the
    /// circumsphere of a simplex made only of input points.  Exact signs on the device; the
    /// reference's all-pairs count where it has at most 2^36 pairs, else the exact O(simplices)
    /// local test (empty spheres of all adjacent pairs, symmetric adjacency, positive orientation).
    pub fn check_delaunay(&self) -> bool {
        match self.violating_pairs() {
            Ok(n) => n == 0,
            Err(e) if e.code == VO_ERR_INVALID => {
                self.violations().unwrap_or_else(|e| panic!("check_delaunay: {}", e)) == 0
            }
            Err(e) => panic!("check_delaunay: {}", e),
        }
    }

    /// the reference's count behind check_delaunay: (all-input simplex, vertex not in it) pairs with
    /// the vertex strictly inside the simplex's circumsphere (vo_tri_check_pairs)
    pub fn violating_pairs(&self) -> Result<i64, VoError> {
        let mut n = 0i64;
        // SAFETY: a live handle and one out-pointer.
        check(unsafe { ffi::vo_tri_check_pairs(self.handle, &mut n) })?;
        Ok(n)
    }

    /// violations of the exact local test (vo_tri_check), O(simplices)
    pub fn violations(&self) -> Result<i64, VoError> {
        let mut n = 0i64;
        // SAFETY: as above.
        check(unsafe { ffi::vo_tri_check(self.handle, &mut n) })?;
        Ok(n)
    }

    // -- the reference's state fields as snapshots ----------------------------------------

    /// live simplex rows (vertex ids, positively oriented) and the export position across each
    /// face (-1 on the super-simplex hull)
    fn export(&self) -> Result<(Vec<i32>, Vec<i32>), VoError> {
        let n = self.counts()?.1 as usize;
        let mut s = vec![0i32; n * M];
        let mut nb = vec![0i32; n * M];
        // SAFETY: both buffers hold n rows of M ints.
        check(unsafe { ffi::vo_tri_export(self.handle, s.as_mut_ptr(), nb.as_mut_ptr()) })?;
        Ok((s, nb))
    }

    /// the simplices field (src/delaunay_tree.rs:28): live simplices with the reference's
    /// circumspheres, and the ghost simplices
    pub fn simplices(&self) -> HashMap<usize, Simplex<N, M>> {
        self.try_simplices().unwrap_or_else(|e| panic!("simplices: {}", e))
    }

    pub fn try_simplices(&self) -> Result<HashMap<usize, Simplex<N, M>>, VoError> {
        let (s, nb) = self.export()?;
        let n = s.len() / M;
        let root_only = n == 1 && s.iter().all(|&v| (v as usize) < M);
        let id = |pos: usize| -> usize {
            if root_only {
                0
            } else {
                N + 2 + pos
            }
        };
        let ghosts = ghost_simplices(N);
        let mut ghost_neighbors: Vec<Vec<usize>> = vec![Vec::new(); ghosts.len()];
        let mut centers = vec![0.0f64; n * N];
        let mut radii = vec![0.0f64; n];
        if n > 0 {
            // SAFETY: coords holds coords.len() rows of N f64, s holds n rows of M ids into it,
            // centers / radii n rows.
            check(unsafe {
                ffi::vo_circumspheres(
                    N as c_int,
                    self.coords.as_ptr().cast::<f64>(),
                    self.coords.len() as i64,
                    s.as_ptr(),
                    n as i64,
                    centers.as_mut_ptr(),
                    radii.as_mut_ptr(),
                )
            })?;
        }
        let mut out: HashMap<usize, Simplex<N, M>> = HashMap::with_capacity(n + ghosts.len());
        for i in 0..n {
            let mut vertices = [0usize; M];
            for k in 0..M {
                vertices[k] = s[i * M + k] as usize;
            }
            let mut neighbors: Vec<usize> = Vec::with_capacity(M);
            for k in 0..M {
                let x = nb[i * M + k];
                if x >= 0 {
                    neighbors.push(id(x as usize));
                } else {
                    let face: Vec<usize> = (0..M).filter(|&j| j != k).map(|j| vertices[j]).collect();
                    let g = ghost_of(&ghosts, &face);
                    ghost_neighbors[g - 1].push(id(i));
                    neighbors.push(g);
                }
            }
            let (center, radius) = if id(i) == 0 {
                self.root_sphere
            } else {
                let mut c = [0.0; N];
                c.copy_from_slice(&centers[i * N..(i + 1) * N]);
                (c, radii[i])
            };
            out.insert(id(i), Simplex { vertices, center, radius, neighbors });
        }
        for (k, g) in ghosts.iter().enumerate() {
            let mut vertices = [0usize; M];
            vertices.copy_from_slice(g);
            out.insert(k + 1, Simplex { vertices, center: [0.0; N], radius: 0.0, neighbors: ghost_neighbors[k].clone() });
        }
        Ok(out)
    }

    /// the vertices field (src/delaunay_tree.rs:27): every vertex id with its coordinates and the
    /// ids of the simplices (ghosts included) that use it, ascending
    pub fn vertices(&self) -> HashMap<usize, Vertex<N>> {
        self.try_vertices().unwrap_or_else(|e| panic!("vertices: {}", e))
    }

    pub fn try_vertices(&self) -> Result<HashMap<usize, Vertex<N>>, VoError> {
        let simplices = self.try_simplices()?;
        let mut incident: Vec<Vec<usize>> = vec![Vec::new(); self.coords.len()];
        let mut ids: Vec<usize> = simplices.keys().copied().collect();
        ids.sort_unstable();
        for k in ids {
            for &v in simplices[&k].vertices.iter() {
                incident[v].push(k);
            }
        }
        let mut out: HashMap<usize, Vertex<N>> = HashMap::with_capacity(self.coords.len());
        for (v, simplex) in incident.into_iter().enumerate() {
            out.insert(v, Vertex { coordinates: self.coords[v], simplex });
        }
        Ok(out)
    }

    /// simplices made only of input points as sorted tuples of input indices, rows sorted, and the
    /// input-point pairs that share a live simplex (the same for any insertion order)
    pub fn canonical(&self) -> Result<(Vec<[i32; M]>, Vec<[i32; 2]>), VoError> {
        let mut t: *mut i32 = ptr::null_mut();
        let mut e: *mut i32 = ptr::null_mut();
        let (mut nt, mut ne) = (0i64, 0i64);
        // SAFETY: out-pointers; on success the library hands over two malloc'd arrays of nt * M and
        // ne * 2 ints, released with vo_free below.
        check(unsafe { ffi::vo_tri_canonical(self.handle, &mut t, &mut nt, &mut e, &mut ne) })?;
        let rows = |p: *mut i32, n: i64, w: usize| -> Vec<i32> {
            if p.is_null() || n <= 0 {
                Vec::new()
            } else {
                // SAFETY: p holds n * w ints (see above).
                let held = unsafe { std::slice::from_raw_parts(p, n as usize * w) };
                held.to_vec()
            }
        };
        let tv = rows(t, nt, M);
        let ev = rows(e, ne, 2);
        // SAFETY: both arrays came from the library's allocator (null is allowed).
        unsafe {
            ffi::vo_free(t.cast::<c_void>());
            ffi::vo_free(e.cast::<c_void>());
        }
        let tris: Vec<[i32; M]> = tv
            .chunks_exact(M)
            .map(|c| {
                let mut a = [0i32; M];
                a.copy_from_slice(c);
                a
            })
            .collect();
        let edges: Vec<[i32; 2]> = ev.chunks_exact(2).map(|c| [c[0], c[1]]).collect();
        Ok((tris, edges))
    }

    // -- the reference's pub fields, written ----------------------------------------------

    /// Vertex::coordinates of an inserted point set anew (the reference's field is pub); nothing
    /// is re-triangulated, so check_delaunay may then fail.
    pub fn set_vertex(&mut self, id: usize, coordinates: [f64; N]) -> Result<(), VoError> {
        // SAFETY: coordinates holds N f64.
        check(unsafe { ffi::vo_tri_set_vertex(self.handle, id as i64, coordinates.as_ptr(), 1) })?;
        self.coords[id] = coordinates;
        Ok(())
    }

    /// Simplex::vertices entries i and j of the live simplex at export position `position` swapped
    pub fn swap_simplex_vertices(&mut self, position: usize, i: usize, j: usize) -> Result<(), VoError> {
        // SAFETY: a live handle; the library checks the indices.
        check(unsafe { ffi::vo_tri_swap_simplex_vertices(self.handle, position as i64, i as c_int, j as c_int) })
    }

    // -- knobs and statistics (no counterpart in the reference) ---------------------------

    /// vo_tri_config: tuning and test knobs of the round kernels
    pub fn config(&mut self, key: &str, value: i64) -> Result<(), VoError> {
        let k = CString::new(key).map_err(|_| VoError::invalid("config key holds a NUL byte".to_string()))?;
        // SAFETY: k is NUL-terminated and outlives the call.
        check(unsafe { ffi::vo_tri_config(self.handle, k.as_ptr(), value) })
    }

    /// rounds, predicate tiers, events and device times of the last run
    pub fn stats(&self) -> Result<ffi::vo_stats, VoError> {
        let mut s = ffi::vo_stats::default();
        // SAFETY: s is a vo_stats.
        check(unsafe { ffi::vo_tri_stats(self.handle, &mut s) })?;
        Ok(s)
    }

    /// the tree made into DelaunayTree::new(vertices) again, its device buffers kept
    pub fn reset(&mut self, vertices: &[[f64; N]]) -> Result<(), VoError> {
        // SAFETY: vertices is len * N contiguous f64.
        check(unsafe { ffi::vo_tri_reset(self.handle, vertices.as_ptr().cast::<f64>(), vertices.len() as i64) })?;
        if !vertices.is_empty() {
            let mut c = [0.0f64; 3];
            let mut r = 0.0f64;
            // SAFETY: as in create.
            check(unsafe {
                ffi::vo_bounding_sphere(
                    N as c_int,
                    vertices.as_ptr().cast::<f64>(),
                    vertices.len() as i64,
                    c.as_mut_ptr(),
                    &mut r,
                )
            })?;
            let mut center = [0.0; N];
            center.copy_from_slice(&c[..N]);
            self.root_sphere = (center, 10.0 * if r > 0.0 { r } else { 1.0 });
        } else {
            self.root_sphere = ([0.0; N], 10.0);
        }
        // the super vertices depend on the new bounding sphere
        let mut super_vertices = vec![[0.0f64; N]; 2 * M];
        // SAFETY: with no point staged the library writes the 2 (N + 1) super and ghost rows.
        check(unsafe { ffi::vo_tri_vertices(self.handle, super_vertices.as_mut_ptr().cast::<f64>()) })?;
        self.coords = super_vertices;
        self.refresh_max_simplex_id()
    }
}

/// struct TreeUpdate (src/delaunay_tree.rs:698-740): one point to insert with its vertex id, and
/// what inserting it into the tree as it was when the update was made kills and creates.
#[derive(Debug, Clone)]
pub struct TreeUpdate<const N: usize, const M: usize> {
    id: usize,
    vertex: [f64; N],
    /// identity of the tree the update was made for
    tree: usize,
    killed_sites: Vec<usize>,
    simplices: Vec<[usize; M]>,
    simplices_id: Vec<usize>,
    centers: Vec<[f64; N]>,
    radii: Vec<f64>,
    neighbors: Vec<(usize, usize)>,
    new_neighbors: Vec<(usize, usize)>,
}

impl<const N: usize, const M: usize> TreeUpdate<N, M> {
    /// TreeUpdate::new (src/delaunay_tree.rs:709-740): locate `vertex`, and the new simplices with
    /// their spheres, outside neighbors and pairings (simplices_id 1..=count, as there).  Panics,
    /// as the reference does, when `vertex` coincides with a vertex of the tree.
    pub fn new(id: usize, vertex: [f64; N], tree: &DelaunayTree<N, M>) -> Self {
        let cavity = tree.cavity(&vertex).unwrap_or_else(|e| panic!("No simplex found for vertex {:?}: {}", vertex, e));
        let (simplices, centers, radii, neighbors) = tree.new_simplices(&cavity, None, vertex, id);
        let simplices_id: Vec<usize> = (1..simplices.len() + 1).collect();
        let new_neighbors = pair_simplices::<M>(&simplices, &simplices_id);
        TreeUpdate {
            id,
            vertex,
            tree: tree.handle as usize,
            killed_sites: cavity.killed,
            simplices,
            simplices_id,
            centers,
            radii,
            neighbors,
            new_neighbors,
        }
    }

    pub fn id(&self) -> usize {
        self.id
    }
    pub fn vertex(&self) -> [f64; N] {
        self.vertex
    }
    pub fn killed_sites(&self) -> &[usize] {
        &self.killed_sites
    }
    pub fn simplices(&self) -> &[[usize; M]] {
        &self.simplices
    }
    pub fn simplices_id(&self) -> &[usize] {
        &self.simplices_id
    }
    pub fn centers(&self) -> &[[f64; N]] {
        &self.centers
    }
    pub fn radii(&self) -> &[f64] {
        &self.radii
    }
    /// (simplex across the facet, killed simplex) per new simplex
    pub fn neighbors(&self) -> &[(usize, usize)] {
        &self.neighbors
    }
    /// pairs of new simplices (by simplices_id) that share a ridge, both orders
    pub fn new_neighbors(&self) -> &[(usize, usize)] {
        &self.new_neighbors
    }
}
