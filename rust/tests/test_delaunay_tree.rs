// The reference's tests/test_delaunay_tree.rs, run against the device build, with the results
// asserted (the reference prints them): max_simplex_id after new, check_delaunay after the
// insertions, and the views.  `delaunay_tree.vertices.len()` of the reference reads
// `num_vertices()` here (the vertices field is a device-side table; `vertices()` snapshots it).
use rand::distributions::{Distribution, Uniform};
use rand::prelude::*;
use std::collections::HashSet;
use voronoids::delaunay_tree::{DelaunayTree, TreeUpdate};

fn sorted(v: &[usize]) -> Vec<usize> {
    let mut s = v.to_vec();
    s.sort_unstable();
    s
}

#[test]
fn test_delaunay_tree_3d() {
    let mut vertices = vec![];
    let mut rng = StdRng::seed_from_u64(0);
    let dist = Uniform::from(0.0..1.0);
    for _ in 0..1000 {
        let point = [dist.sample(&mut rng), dist.sample(&mut rng), dist.sample(&mut rng)];
        vertices.push(point);
    }
    let mut delaunay_tree = DelaunayTree::<3, 4>::new(vertices.clone());
    assert_eq!(delaunay_tree.max_simplex_id, 4);
    let n_points = delaunay_tree.num_vertices();
    assert_eq!(n_points, 8);

    for i in 0..100 {
        let update = TreeUpdate::new(n_points + i, vertices[i], &delaunay_tree);
        delaunay_tree.insert_point(&update);
    }
    let mut vertices2: Vec<[f64; 3]> = vec![];
    for _ in 0..1000 {
        let point = [dist.sample(&mut rng), dist.sample(&mut rng), dist.sample(&mut rng)];
        vertices2.push(point);
    }
    delaunay_tree.add_points_to_tree(vertices2);
    assert!(delaunay_tree.check_delaunay());
    assert_eq!(delaunay_tree.num_vertices(), 8 + 1100);
}

#[test]
fn test_delaunay_tree_2d() {
    let vertices = vec![[0.3, 0.1], [1.0, 0.2], [0.1, 1.0], [0.5, 0.5]];
    let mut delaunay_tree = DelaunayTree::<2, 3>::new(vertices.clone());
    assert_eq!(delaunay_tree.max_simplex_id, 3);
    let mut rng = StdRng::seed_from_u64(0);
    let dist = Uniform::from(0.0..1.0);
    let n_points = delaunay_tree.num_vertices();
    for i in 0..1000 {
        let point = [dist.sample(&mut rng), dist.sample(&mut rng)];
        let update = TreeUpdate::new(n_points + i, point, &delaunay_tree);
        delaunay_tree.insert_point(&update);
    }
    assert!(delaunay_tree.check_delaunay());
}

#[test]
fn test_update_is_the_cavity() {
    // TreeUpdate::new against the tree before the insertion: the killed simplices disappear and the
    // new simplices appear (as vertex sets) once the point is inserted
    let mut rng = StdRng::seed_from_u64(3);
    let dist = Uniform::from(0.0..1.0);
    let pts: Vec<[f64; 3]> = (0..500).map(|_| [dist.sample(&mut rng), dist.sample(&mut rng), dist.sample(&mut rng)]).collect();
    let mut tree = DelaunayTree::<3, 4>::new(pts.clone());
    tree.add_points_to_tree(pts[..400].to_vec());
    let before: HashSet<Vec<usize>> = tree.simplices().values().map(|s| sorted(&s.vertices)).collect();
    let id = tree.num_vertices();
    let update = TreeUpdate::new(id, pts[400], &tree);
    let snapshot = tree.simplices();
    let killed: HashSet<Vec<usize>> = update.killed_sites().iter().map(|k| sorted(&snapshot[k].vertices)).collect();
    tree.insert_point(&update);
    let after: HashSet<Vec<usize>> = tree.simplices().values().map(|s| sorted(&s.vertices)).collect();
    let created: HashSet<Vec<usize>> = update.simplices().iter().map(|s| sorted(s)).collect();
    assert_eq!(before.difference(&after).cloned().collect::<HashSet<_>>(), killed);
    assert_eq!(after.difference(&before).cloned().collect::<HashSet<_>>(), created);
    assert_eq!(update.new_neighbors().len() % 2, 0);
    assert!(tree.check_delaunay());
}

#[test]
fn test_views() {
    let mut rng = StdRng::seed_from_u64(5);
    let dist = Uniform::from(0.0..1.0);
    let pts: Vec<[f64; 3]> = (0..300).map(|_| [dist.sample(&mut rng), dist.sample(&mut rng), dist.sample(&mut rng)]).collect();
    let tree = voronoids::delaunay(pts.clone());
    let simplices = tree.simplices();
    let vertices = tree.vertices();
    assert_eq!(vertices.len(), 8 + 300);
    for (k, g) in [[4, 0, 1, 2], [5, 0, 2, 3], [6, 0, 3, 1], [7, 1, 2, 3]].iter().enumerate() {
        let s = &simplices[&(k + 1)];
        assert_eq!(&s.vertices, g);
        assert_eq!(s.radius, 0.0);
        assert_eq!(s.neighbors.len(), 1);
    }
    for (id, s) in simplices.iter() {
        for n in s.neighbors.iter() {
            assert!(simplices[n].neighbors.contains(id));
        }
        for v in s.vertices.iter() {
            assert!(vertices[v].simplex.contains(id));
        }
    }
    for (i, p) in pts.iter().enumerate() {
        assert_eq!(vertices[&(8 + i)].coordinates, *p);
    }
    assert!(tree.max_simplex_id >= *simplices.keys().max().unwrap());
}
