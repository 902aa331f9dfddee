//! scheduler (src/scheduler.rs): the reference's plan for inserting many points at once.  The
//! device build does not use it (window points reserve their cavities with 64-bit atomics and a
//! round's winners are the points that hold every reservation they need); it is kept for callers
//! that drive insert_points_parallel level by level themselves.
use std::collections::HashMap;

use crate::delaunay_tree::DelaunayTree;

/// make_queue (src/scheduler.rs:6-28): for each point (in input order, id = its index) the
/// simplices within two neighbor steps of its cavity, sorted and unique.
pub fn make_queue<const N: usize, const M: usize>(
    vertices: Vec<[f64; N]>,
    tree: &DelaunayTree<N, M>,
) -> Vec<(usize, [f64; N], Vec<usize>)> {
    let simplices = tree.simplices();
    vertices
        .into_iter()
        .enumerate()
        .map(|(id, vertex)| {
            let killed = tree.locate(vertex);
            let mut sites: Vec<usize> = Vec::new();
            for k in killed.iter() {
                for first in simplices[k].neighbors.iter() {
                    sites.extend(simplices[first].neighbors.iter().copied());
                }
            }
            sites.sort_unstable();
            sites.dedup();
            (id, vertex, sites)
        })
        .collect()
}

/// find_placement (src/scheduler.rs:30-55): the level of each queue entry, 1 + the level of the
/// latest earlier entry that touches one of its sites (1 when none does), maximised over its
/// sites.  Entries are taken in queue order; ids must be 0..queue.len() and sites unique, as
/// make_queue returns them.  Levels of the entries touching one site strictly increase, so the
/// latest toucher's level is all that is kept per site.  Panics on an entry with no site, as the
/// reference does.
#[allow(clippy::ptr_arg)]
pub fn find_placement<const N: usize>(queue: &Vec<(usize, [f64; N], Vec<usize>)>) -> Vec<usize> {
    let mut latest: HashMap<usize, usize> = HashMap::new();
    let mut placement: Vec<usize> = vec![0; queue.len()];
    for (id, _, sites) in queue.iter() {
        let level = sites
            .iter()
            .map(|s| latest.get(s).map_or(1, |l| l + 1))
            .max()
            .unwrap_or_else(|| panic!("find_placement: entry {} touches no simplex", id));
        for s in sites.iter() {
            latest.insert(*s, level);
        }
        placement[*id] = level;
    }
    placement
}
