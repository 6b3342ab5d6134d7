//! The crate's own hand-checked examples, run against the GPU library (needs an MI355X).
//! Sources: src/blas.rs:711-723, :828-849, :875-887; src/lapack/mod.rs:483-495, :497-508,
//! :537-558, :560-607.
use algebra_linear_hip::{blas, lapack, lapack::LapackError};

fn sample4() -> Vec<f64> {
    vec![4.0, 1.0, 2.0, 1.0, 1.0, 3.0, 0.0, 1.0, 2.0, 0.0, 5.0, 2.0, 1.0, 1.0, 2.0, 6.0]
}

#[test]
fn dgemv_hand_values() {
    let a = vec![1.0, 4.0, 2.0, 5.0, 3.0, 6.0];
    let mut y = vec![0.0; 2];
    blas::dgemv(1.0, &a, &[1.0, 1.0, 1.0], 0.0, &mut y, 2, 3);
    assert_eq!(y, vec![6.0, 15.0]);
}

#[test]
fn dlange_norms() {
    let a = vec![1.0, 4.0, 2.0, 5.0, 3.0, 6.0];
    assert_eq!(blas::dlange('M', &a, 2, 3), 6.0);
    assert_eq!(blas::dlange('1', &a, 2, 3), 9.0);
    assert_eq!(blas::dlange('I', &a, 2, 3), 15.0);
    assert!((blas::dlange('F', &a, 2, 3) - 91.0f64.sqrt()).abs() < 1e-12);
}

#[test]
#[should_panic(expected = "dgemm: a must be m*k")]
fn dgemm_shape_panic_has_the_crates_text() {
    let mut c = vec![0.0; 4];
    blas::dgemm(1.0, &[0.0; 5], &[0.0; 6], 0.0, &mut c, 2, 2, 3);
}

#[test]
fn dgesv_recovers_a_known_solution() {
    let a0 = sample4();
    let x = [1.0, -2.0, 3.0, 5.0];
    let mut b = vec![0.0; 4];
    for i in 0..4 {
        for j in 0..4 {
            b[i] += a0[i + j * 4] * x[j];
        }
    }
    let mut a = a0.clone();
    lapack::dgesv(&mut a, &mut b).unwrap();
    for i in 0..4 {
        assert!((b[i] - x[i]).abs() < 1e-12);
    }
}

#[test]
fn dgetrf_singular_matrix() {
    let mut a = vec![1.0, 2.0, 3.0, 2.0, 4.0, 6.0, 1.0, 1.0, 1.0];
    assert_eq!(lapack::dgetrf(&mut a, 3).unwrap_err(), LapackError::Singular);
}

#[test]
fn dpotrf_two_by_two_and_not_spd() {
    let mut a = vec![4.0, 2.0, 2.0, 3.0];
    lapack::dpotrf(&mut a, 2).unwrap();
    assert!((a[0] - 2.0).abs() < 1e-15 && (a[1] - 1.0).abs() < 1e-15 && (a[3] - 2.0f64.sqrt()).abs() < 1e-15);
    let mut b = vec![1.0, 2.0, 2.0, 1.0];
    assert_eq!(lapack::dpotrf(&mut b, 2).unwrap_err(), LapackError::Singular);
}

#[test]
fn bad_shape_is_an_error_not_a_panic() {
    let mut a = vec![0.0; 8];
    let mut b = vec![0.0; 3];
    assert_eq!(lapack::dgesv(&mut a, &mut b).unwrap_err(), LapackError::BadShape);
}
