//! The C ABI of rustymath's MI355X dense linear-algebra library (`include/hip_sys.h`, `libhip_sys.so`): raw
//! declarations only. Every function returns 0, or -i for an invalid argument i, or a positive status for a device
//! failure (100000: a device-side bounded wait ran out; 100001: valid arguments beyond what the library handles);
//! `hs_last_error` gives the calling thread's message. Matrices are row-major with a leading dimension and 64-bit
//! sizes (CBLAS row-major, ILP64), as `core/src/linalg/accelerate.rs` already passes them to a vendor BLAS.
#![allow(non_camel_case_types)]
#![no_std]
use core::ffi::{c_char, c_int, c_void};

pub const HS_NO_TRANS: c_int = 111;
pub const HS_TRANS: c_int = 112;
pub const HS_CONJ_TRANS: c_int = 113;
pub const HS_UPPER: c_int = 121;
pub const HS_LOWER: c_int = 122;
pub const HS_NON_UNIT: c_int = 131;
pub const HS_UNIT: c_int = 132;
pub const HS_LEFT: c_int = 141;
pub const HS_RIGHT: c_int = 142;

#[link(name = "hip_sys")]
extern "C" {
    pub fn hs_init() -> c_int;
    pub fn hs_finalize() -> c_int;
    pub fn hs_set_exchange(mode: c_int, local: *mut c_int) -> c_int;
    pub fn hs_last_error() -> *const c_char;
    pub fn hs_dgemm(op_a: c_int, op_b: c_int, m: i64, n: i64, k: i64, alpha: f64, a: *const f64, lda: i64,
                    b: *const f64, ldb: i64, beta: f64, c: *mut f64, ldc: i64) -> c_int;
    pub fn hs_sgemm(op_a: c_int, op_b: c_int, m: i64, n: i64, k: i64, alpha: f32, a: *const f32, lda: i64,
                    b: *const f32, ldb: i64, beta: f32, c: *mut f32, ldc: i64) -> c_int;
    pub fn hs_dgemm_batched(op_a: c_int, op_b: c_int, m: i64, n: i64, k: i64, alpha: f64, a: *const f64, lda: i64,
                            sa: i64, b: *const f64, ldb: i64, sb: i64, beta: f64, c: *mut f64, ldc: i64, sc: i64,
                            batch: i64) -> c_int;
    pub fn hs_sgemm_batched(op_a: c_int, op_b: c_int, m: i64, n: i64, k: i64, alpha: f32, a: *const f32, lda: i64,
                            sa: i64, b: *const f32, ldb: i64, sb: i64, beta: f32, c: *mut f32, ldc: i64, sc: i64,
                            batch: i64) -> c_int;
    pub fn hs_dtrsm(side: c_int, uplo: c_int, trans: c_int, diag: c_int, m: i64, n: i64, a: *const f64, lda: i64,
                    x: *mut f64, ldx: i64) -> c_int;
    pub fn hs_strsm(side: c_int, uplo: c_int, trans: c_int, diag: c_int, m: i64, n: i64, a: *const f32, lda: i64,
                    x: *mut f32, ldx: i64) -> c_int;
    pub fn hs_dlu_factor(a: *const f64, n: i64, factors: *mut f64, perm: *mut i64, pivots: *mut i64, info: *mut i64) -> c_int;
    pub fn hs_slu_factor(a: *const f32, n: i64, factors: *mut f32, perm: *mut i64, pivots: *mut i64, info: *mut i64) -> c_int;
    pub fn hs_dlu_solve(factors: *const f64, perm: *const i64, n: i64, b: *const f64, k: i64, x: *mut f64) -> c_int;
    pub fn hs_slu_solve(factors: *const f32, perm: *const i64, n: i64, b: *const f32, k: i64, x: *mut f32) -> c_int;
    pub fn hs_dlu_inverse(factors: *const f64, perm: *const i64, n: i64, out: *mut f64) -> c_int;
    pub fn hs_slu_inverse(factors: *const f32, perm: *const i64, n: i64, out: *mut f32) -> c_int;
    pub fn hs_dinverse(a: *const f64, n: i64, out: *mut f64, det: *mut f64, regular: *mut c_int) -> c_int;
    pub fn hs_sinverse(a: *const f32, n: i64, out: *mut f32, det: *mut f32, regular: *mut c_int) -> c_int;
    pub fn hs_dsolve(a: *const f64, n: i64, b: *const f64, k: i64, x: *mut f64, regular: *mut c_int) -> c_int;
    pub fn hs_ssolve(a: *const f32, n: i64, b: *const f32, k: i64, x: *mut f32, regular: *mut c_int) -> c_int;
    pub fn hs_dcholesky(a: *const f64, n: i64, l: *mut f64, ok: *mut c_int) -> c_int;
    pub fn hs_scholesky(a: *const f32, n: i64, l: *mut f32, ok: *mut c_int) -> c_int;
    pub fn hs_dsolve_stack(count: i64, n: i64, k: i64, a: *const f64, b: *const f64, x: *mut f64, regular: *mut c_int) -> c_int;
    pub fn hs_ssolve_stack(count: i64, n: i64, k: i64, a: *const f32, b: *const f32, x: *mut f32, regular: *mut c_int) -> c_int;
    pub fn hs_dlu_stack(count: i64, n: i64, a: *const f64, factors: *mut f64, perm: *mut i64) -> c_int;
    pub fn hs_dlu_solve_stack(count: i64, n: i64, k: i64, factors: *const f64, perm: *const i64, b: *const f64,
                              x: *mut f64) -> c_int;
    pub fn hs_slu_solve_stack(count: i64, n: i64, k: i64, factors: *const f32, perm: *const i64, b: *const f32,
                              x: *mut f32) -> c_int;
    pub fn hs_slu_stack(count: i64, n: i64, a: *const f32, factors: *mut f32, perm: *mut i64) -> c_int;
    pub fn hs_dgesv_batched(count: i64, n: *const i64, k: *const i64, a: *const *const f64, b: *const *const f64,
                            x: *const *mut f64, regular: *mut c_int) -> c_int;
    pub fn hs_sgesv_batched(count: i64, n: *const i64, k: *const i64, a: *const *const f32, b: *const *const f32,
                            x: *const *mut f32, regular: *mut c_int) -> c_int;
    pub fn hs_dev_dgemm(op_a: c_int, op_b: c_int, m: i64, n: i64, k: i64, alpha: f64, a: *const f64, lda: i64,
                        b: *const f64, ldb: i64, beta: f64, c: *mut f64, ldc: i64, stream: *mut c_void) -> c_int;
    pub fn hs_dev_sgemm(op_a: c_int, op_b: c_int, m: i64, n: i64, k: i64, alpha: f32, a: *const f32, lda: i64,
                        b: *const f32, ldb: i64, beta: f32, c: *mut f32, ldc: i64, stream: *mut c_void) -> c_int;
    pub fn hs_dev_dtrsm(side: c_int, uplo: c_int, trans: c_int, diag: c_int, m: i64, n: i64, a: *const f64, lda: i64,
                        x: *mut f64, ldx: i64, stream: *mut c_void) -> c_int;
    pub fn hs_dev_strsm(side: c_int, uplo: c_int, trans: c_int, diag: c_int, m: i64, n: i64, a: *const f32, lda: i64,
                        x: *mut f32, ldx: i64, stream: *mut c_void) -> c_int;
    pub fn hs_dev_dgemm_batched(op_a: c_int, op_b: c_int, m: i64, n: i64, k: i64, alpha: f64, a: *const f64,
                                lda: i64, sa: i64, b: *const f64, ldb: i64, sb: i64, beta: f64, c: *mut f64,
                                ldc: i64, sc: i64, batch: i64, stream: *mut c_void) -> c_int;
    pub fn hs_dev_sgemm_batched(op_a: c_int, op_b: c_int, m: i64, n: i64, k: i64, alpha: f32, a: *const f32,
                                lda: i64, sa: i64, b: *const f32, ldb: i64, sb: i64, beta: f32, c: *mut f32,
                                ldc: i64, sc: i64, batch: i64, stream: *mut c_void) -> c_int;
    pub fn hs_dev_dgetrf(n: i64, a: *mut f64, lda: i64, perm: *mut i64, info: *mut i64, stream: *mut c_void) -> c_int;
    pub fn hs_dev_sgetrf(n: i64, a: *mut f32, lda: i64, perm: *mut i64, info: *mut i64, stream: *mut c_void) -> c_int;
    pub fn hs_dev_dgesv(n: i64, a: *mut f64, lda: i64, perm: *mut i64, b: *mut f64, k: i64, ldb: i64,
                        regular: *mut c_int, stream: *mut c_void) -> c_int;
    pub fn hs_dev_sgesv(n: i64, a: *mut f32, lda: i64, perm: *mut i64, b: *mut f32, k: i64, ldb: i64,
                        regular: *mut c_int, stream: *mut c_void) -> c_int;
    pub fn hs_dev_dgetrs(trans: c_int, n: i64, k: i64, factors: *const f64, ldf: i64, perm: *const i64, b: *mut f64,
                         ldb: i64, stream: *mut c_void) -> c_int;
    pub fn hs_dev_sgetrs(trans: c_int, n: i64, k: i64, factors: *const f32, ldf: i64, perm: *const i64, b: *mut f32,
                         ldb: i64, stream: *mut c_void) -> c_int;
    pub fn hs_dev_dgetrs_noalloc(trans: c_int, n: i64, k: i64, factors: *const f64, ldf: i64, perm: *const i64,
                                 b: *mut f64, ldb: i64, stream: *mut c_void) -> c_int;
    pub fn hs_dev_dgetri(n: i64, factors: *const f64, ldf: i64, perm: *const i64, out: *mut f64, ldo: i64,
                         stream: *mut c_void) -> c_int;
    pub fn hs_dev_sgetri(n: i64, factors: *const f32, ldf: i64, perm: *const i64, out: *mut f32, ldo: i64,
                         stream: *mut c_void) -> c_int;
    pub fn hs_dev_dpotrf(n: i64, a: *mut f64, lda: i64, ok: *mut c_int, stream: *mut c_void) -> c_int;
    pub fn hs_dev_spotrf(n: i64, a: *mut f32, lda: i64, ok: *mut c_int, stream: *mut c_void) -> c_int;
    pub fn hs_dev_dsgesv(n: i64, a: *const f64, lda: i64, b: *const f64, x: *mut f64, max_iter: c_int, tol: f64,
                         iters: *mut c_int, berr: *mut f64, stream: *mut c_void) -> c_int;
    pub fn hs_dev_sgemm_bf16_update(m: i64, n: i64, k: i64, a: *const f32, lda: i64, b: *const f32, ldb: i64,
                                    c: *mut f32, ldc: i64, terms: c_int, stream: *mut c_void) -> c_int;
    pub fn hs_dev_dbgesv(n: i64, a: *const f64, lda: i64, b: *const f64, x: *mut f64, max_iter: c_int,
                         restart: c_int, terms: c_int, tol: f64, iters: *mut c_int, inner: *mut c_int,
                         berr: *mut f64, stream: *mut c_void) -> c_int;
    pub fn hs_dev_dgesv_batched_plan(count: i64, n: *const i64, k: *const i64, a_off: *const i64, b_off: *const i64,
                                     p_off: *const i64, a: *mut f64, b: *mut f64, perm: *mut i64,
                                     regular: *mut c_int, mode: c_int, panel_width: c_int, wgs_per_cu: c_int,
                                     plan: *mut i64) -> c_int;
    pub fn hs_dev_sgesv_batched_plan(count: i64, n: *const i64, k: *const i64, a_off: *const i64, b_off: *const i64,
                                     p_off: *const i64, a: *mut f32, b: *mut f32, perm: *mut i64,
                                     regular: *mut c_int, mode: c_int, panel_width: c_int, wgs_per_cu: c_int,
                                     plan: *mut i64) -> c_int;
    pub fn hs_dev_dsgesv_batched_plan(count: i64, n: *const i64, k: *const i64, a_off: *const i64, b_off: *const i64,
                                      p_off: *const i64, a: *const f64, b: *const f64, x: *mut f64, a32: *mut f32,
                                      r32: *mut f32, perm: *mut i64, regular: *mut c_int, iters: *mut c_int,
                                      berr: *mut f64, max_iter: c_int, tol_per_n: f64, panel_width: c_int,
                                      wgs_per_cu: c_int, plan: *mut i64) -> c_int;
    pub fn hs_dev_gesv_batched_run(plan: i64, graph: c_int, stream: *mut c_void) -> c_int;
    pub fn hs_dev_gesv_batched_free(plan: i64) -> c_int;
    pub fn hs_profile_panel(on: c_int, out: *mut i64) -> c_int;
    pub fn hs_profile_gemm(on: c_int, ms: *mut f64, launches: *mut i64, flops: *mut f64, bytes: *mut f64) -> c_int;
    pub fn hs_profile_lu(on: c_int, out: *mut f64) -> c_int;
    pub fn hs_profile_batched(on: c_int, out: *mut i64) -> c_int;
}
