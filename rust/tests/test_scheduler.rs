// The reference's tests/test_scheduler.rs, with the placement checked: entries on one level touch
// disjoint simplex sets, and every entry sits one level above the latest earlier entry it shares a
// simplex with.
use rand::distributions::{Distribution, Uniform};
use rand::prelude::*;
use std::collections::{HashMap, HashSet};
use voronoids::delaunay_tree::{DelaunayTree, TreeUpdate};
use voronoids::scheduler::{find_placement, make_queue};

#[test]
fn test_queuing() {
    const N_POINTS: usize = 1000;
    let mut vertices = vec![];
    let mut vertices2 = vec![];
    let mut rng = StdRng::seed_from_u64(0);
    let dist = Uniform::from(0.0..1.0);
    for _ in 0..N_POINTS {
        vertices.push([dist.sample(&mut rng), dist.sample(&mut rng), dist.sample(&mut rng)]);
        vertices2.push([dist.sample(&mut rng), dist.sample(&mut rng), dist.sample(&mut rng)]);
    }
    let mut delaunay_tree = DelaunayTree::<3, 4>::new(vertices.clone());
    let n_points = delaunay_tree.num_vertices();
    for i in 0..N_POINTS {
        let update = TreeUpdate::new(n_points + i, vertices[i], &delaunay_tree);
        delaunay_tree.insert_point(&update);
    }
    let queue = make_queue(vertices2, &delaunay_tree);
    let placement = find_placement(&queue);
    assert_eq!(placement.len(), N_POINTS);
    let mut per_level: HashMap<usize, HashSet<usize>> = HashMap::new();
    for (id, _, sites) in queue.iter() {
        let taken = per_level.entry(placement[*id]).or_default();
        for s in sites {
            assert!(taken.insert(*s), "two entries of level {} share simplex {}", placement[*id], s);
        }
    }
}

#[test]
fn test_find_placement_by_hand() {
    let q: Vec<(usize, [f64; 2], Vec<usize>)> = vec![
        (0, [0.0, 0.0], vec![1, 2]),
        (1, [0.0, 0.0], vec![3]),
        (2, [0.0, 0.0], vec![2, 3]),
        (3, [0.0, 0.0], vec![1, 4]),
        (4, [0.0, 0.0], vec![3, 4]),
    ];
    assert_eq!(find_placement(&q), vec![1, 1, 2, 2, 3]);
}
