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

// ---- complex twins: src/lapack/complex.rs:422-451 (test matrices), :483-494, :500-508, :532-575

use algebra_linear_hip::Complex64;

fn cpx(re: f64, im: f64) -> Complex64 {
    Complex64::new(re, im)
}

fn general3() -> Vec<Complex64> {
    vec![cpx(4.0, 1.0), cpx(1.0, 3.0), cpx(2.0, -1.0), cpx(1.0, -2.0), cpx(3.0, 0.0), cpx(1.0, 1.0), cpx(2.0, 0.0),
         cpx(0.0, -1.0), cpx(5.0, 2.0)]
}

fn spd3() -> Vec<Complex64> {
    vec![cpx(5.0, 0.0), cpx(1.0, -1.0), cpx(0.0, 2.0), cpx(1.0, 1.0), cpx(6.0, 0.0), cpx(1.0, 2.0), cpx(0.0, -2.0),
         cpx(1.0, -2.0), cpx(7.0, 0.0)]
}

#[test]
fn zgesv_recovers_the_known_solution() {
    let a = general3();
    let x_true = [cpx(1.0, -2.0), cpx(3.0, 0.5), cpx(-1.0, 1.0)];
    let mut b = vec![cpx(0.0, 0.0); 3];
    for i in 0..3 {
        for j in 0..3 {
            b[i] += a[i + j * 3] * x_true[j];
        }
    }
    let mut aw = a.clone();
    lapack::zgesv(&mut aw, &mut b).unwrap();
    for i in 0..3 {
        assert!((b[i] - x_true[i]).norm() < 1e-12);
    }
}

#[test]
fn zgetrf_reports_the_singular_matrix() {
    let mut a = vec![cpx(1.0, 1.0), cpx(3.0, 0.0), cpx(2.0, 2.0), cpx(6.0, 0.0)];
    assert_eq!(lapack::zgetrf(&mut a, 2), Err(LapackError::Singular));
}

#[test]
fn zpotrf_factorises_both_triangles() {
    let a = spd3();
    let mut lower = a.clone();
    lapack::zpotrf(&mut lower, 3, false).unwrap();
    let mut upper = a.clone();
    lapack::zpotrf(&mut upper, 3, true).unwrap();
    for i in 0..3 {
        for j in 0..=i {
            let mut s = cpx(0.0, 0.0);
            for k in 0..=j {
                s += lower[i + k * 3] * lower[j + k * 3].conj();
            }
            assert!((s - a[i + j * 3]).norm() < 1e-12);
        }
        for j in i..3 {
            let mut s = cpx(0.0, 0.0);
            for k in 0..=i {
                s += upper[k + i * 3].conj() * upper[k + j * 3];
            }
            assert!((s - a[i + j * 3]).norm() < 1e-12);
        }
    }
}

#[test]
#[should_panic(expected = "zgemm_tn: a must be k*m")]
fn zgemm_shape_panic_has_the_crates_text() {
    let mut c = vec![cpx(0.0, 0.0); 4];
    blas::zgemm_tn(cpx(1.0, 0.0), &[cpx(0.0, 0.0); 5], &[cpx(0.0, 0.0); 6], cpx(0.0, 0.0), &mut c, 2, 2, 3);
}
