// The reference's tests/test_geometry.rs with its known answer, plus the 2D formula and in_sphere.
use rand::distributions::{Distribution, Uniform};
use rand::prelude::*;
use voronoids::geometry::{bounding_sphere, circumsphere, in_sphere};

#[test]
fn test_circumsphere() {
    let vertices = [[1.0, 0.0, 0.0], [0.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]];
    let (center, radius) = circumsphere(vertices);
    assert_eq!(center, [0.5, 0.5, 0.5]);
    assert_eq!(radius, 0.8660254037844386);
}

#[test]
fn test_circumsphere_2d() {
    // the slope formula's result for a triangle with no axis-parallel first or second edge
    let (center, radius) = circumsphere([[0.0, 0.0], [2.0, 1.0], [1.0, 3.0]]);
    for v in [[0.0, 0.0], [2.0, 1.0], [1.0, 3.0]].iter() {
        let d = ((v[0] - center[0]).powi(2) + (v[1] - center[1]).powi(2)).sqrt();
        assert!((d - radius).abs() < 1e-12);
    }
}

#[test]
fn test_in_sphere_is_strict() {
    assert!(in_sphere([0.5, 0.5, 0.5], [0.0, 0.0, 0.0], 1.0));
    assert!(!in_sphere([1.0, 0.0, 0.0], [0.0, 0.0, 0.0], 1.0));
    assert!(!in_sphere([0.0, 2.0], [0.0, 0.0], 1.0));
}

#[test]
fn test_boundsphere() {
    const N_TEST: usize = 1000;
    let mut rng = StdRng::seed_from_u64(0);
    let dist = Uniform::from(-1.0..1.0);
    let mut point_test: [[f64; 3]; N_TEST] = [[0.0; 3]; N_TEST];
    for p in point_test.iter_mut() {
        *p = [dist.sample(&mut rng), dist.sample(&mut rng), dist.sample(&mut rng)];
    }
    let (center, radius) = bounding_sphere(point_test.to_vec());
    for p in point_test.iter() {
        let d = ((p[0] - center[0]).powi(2) + (p[1] - center[1]).powi(2) + (p[2] - center[2]).powi(2)).sqrt();
        assert!(d <= radius);
    }
}
