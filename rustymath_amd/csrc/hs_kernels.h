// Internal launch API of the hip-sys kernels (C++; the C ABI in hip_sys.h wraps it).
// Every matrix is row-major with a leading dimension; `ta`/`tb` read an operand as the transpose of
// what it stores. Launches go to `s` and return a status (hs_common.h).
#pragma once
#include "hs_common.h"

namespace hs {

// Tiled::gemm (core/src/linalg.rs:772, :799-807): c = alpha op(a) op(b) + beta c, with the reference's
// routing: nothing for m or n = 0, a scaling for k = 0 or alpha = 0, `by_vector` for m or n = 1
// (:1020-1073) and the MFMA kernel otherwise. Each result starts from 0 or beta c and takes alpha a b
// in order of p, one fused multiply-add per step (:1353-1426). in_order keeps every shape on the MFMA kernel (no
// `by_vector` orders for m or n = 1), for the library's own blocked substitutions.
template <typename T>
int gemm(bool ta, bool tb, int64_t m, int64_t n, int64_t k, T alpha, const T* a, int64_t lda,
         const T* b, int64_t ldb, T beta, T* c, int64_t ldc, hipStream_t s, bool in_order = false);

// `gemm` for `batch` cells at once (a stack's products, one launch): cell i takes a + i sa, b + i sb and c + i sc
// (a stride of 0 repeats an operand); every cell's routing and arithmetic are gemm's.
template <typename T>
int gemm_batched(bool ta, bool tb, int64_t m, int64_t n, int64_t k, T alpha, const T* a, int64_t lda, int64_t sa,
                 const T* b, int64_t ldb, int64_t sb, T beta, T* c, int64_t ldc, int64_t sc, int64_t batch,
                 hipStream_t s);

// BF16-input, FP32-accumulate product for the low-precision factorization: c (f32) = c - a b with a, b rounded to
// bfloat16 once (terms = 1) or split into bfloat16 pairs hi + lo with three products (terms = 3), on MFMA 16x16x32 bf16.
int gemm_bf16_update(int64_t m, int64_t n, int64_t k, const float* a, int64_t lda, const float* b, int64_t ldb,
                     float* c, int64_t ldc, hipStream_t s, int terms);

// LU with partial pivoting of the row-major n×n matrix `a` (ld `lda`) in place: L strictly below the
// diagonal (unit diagonal implicit), U on and above; pivots[c] = the row swapped with row c (absolute,
// 0-based, LAPACK ipiv style). The arithmetic is rustymath's lu_factor (core/src/linalg/factor.rs:617),
// item for item. A zero pivot leaves its column undivided (factor.rs:606); the C ABI's `info` reports the first one
// from the factors' diagonal afterwards. Above 32 x 512 rows the panels are factored in memory, column by
// column, without look-ahead (the same arithmetic).
// bf16_terms (FP32 only, 0 = none): the trailing products after each 256-column block take their operands as bfloat16
// (gemm_bf16_update with that many terms); panels, row moves and the rows of U stay FP32. Not the reference's
// arithmetic: the low-precision factors of the BF16 mixed solve.
// perm_out (device, n items, may be null): rustymath's `perm` (position -> original row), composed on the GPU from the
// panels' row moves. keep_rows leaves the factors row-major (row i holds L's row i left of the diagonal and U's right of
// it) instead of transposing them into the reference's column-major layout: for the mixed solves' own substitutions.
template <typename T>
int getrf(int64_t n, T* a, int64_t lda, int64_t* pivots, hipStream_t s, int bf16_terms = 0, int64_t* perm_out = nullptr,
          bool keep_rows = false);

// Solves with factors from getrf: x = U⁻¹ L⁻¹ P b in place on b (n×k, ld ldb, row-major), with the
// reference's substitution orders (factor.rs:633-641, :668-748). `trans` solves Aᵀ x = b instead
// (Uᵀ, then Lᵀ, then the inverse row permutation).
template <typename T>
int getrs(bool trans, int64_t n, int64_t k, const T* lu, int64_t lda, const int64_t* pivots, T* b, int64_t ldb,
          hipStream_t s);

// `Tiled::triangular` (core/src/linalg.rs:781-787) with rustymath's native orders (`triangular_native`,
// core/src/linalg/factor.rs:711-721): solves op(a) x = b (right = false, a m×m) or x op(a) = b (right = true, a n×n)
// in place on x (m×n, ld ldx). The stored triangle is the lower one (stored_lower) or the upper one; ta reads a
// transposed; unit takes the diagonal as ones. `contiguous` says whether the caller's x is one column with ldx = 1,
// which takes `single`'s orders (x may be a compact device copy of a strided column). Synchronises `s`.
template <typename T>
int trsm(bool right, bool stored_lower, bool ta, bool unit, int64_t m, int64_t n, const T* a, int64_t lda, T* x,
         int64_t ldx, bool contiguous, hipStream_t s);
// `single` (factor.rs:725-748) for a triangle read by columns (entry (i, j) = f[j ld + i]), one contiguous column x.
template <typename T>
int trsv1_columns(bool lower, bool unit, int64_t n, const T* f, int64_t ld, T* x, hipStream_t s);

// getrs for FP64 with no allocation: its workspace is a static device array of the library (4 MiB and 256 bytes per
// device), so n <= 65536, k <= 16 and n k <= 2^18; calls that overlap on one device race on that array.
int getrs_noalloc(bool trans, int64_t n, int64_t k, const double* lu, int64_t lda, const int64_t* perm, double* b,
                  int64_t ldb, hipStream_t s);

// The inverse from factors and perm as getrf leaves them (`lu_inverse`, factor.rs:645-664): out (n×n, row-major, ld
// ldo) = U⁻¹ L⁻¹ with its columns permuted back, in the reference's orders of operations.
template <typename T>
int getri(int64_t n, const T* lu, int64_t lda, const int64_t* perm, T* out, int64_t ldo, hipStream_t s);

// The Cholesky factor in place (`cholesky_factor`, factor.rs:777-797): a (row-major n×n, ld lda, lower triangle read)
// becomes L with zeros above the diagonal, in the reference's order of operations; *ok = 0 (a's contents then
// undefined) where a diagonal item is not > 0, as the reference gives None. Synchronises `s`.
template <typename T>
int potrf(int64_t n, T* a, int64_t lda, int* ok, hipStream_t s);

// Mixed-precision solve: FP32 factors on MFMA, FP64 iterative refinement on the GPU (ir.hip).
int gesv_ir(int64_t n, const double* a, int64_t lda, const double* b, double* x, int max_iter, double tol,
            int* iters, double* berr, hipStream_t s);
// Mixed-precision solve with BF16 trailing products: FP32 panels and BF16-input, FP32-accumulate products, then
// FP64 GMRES-based refinement (GMRES on A d = r, left-preconditioned by the low-precision LU factors, restarted every
// `restart` steps) until the backward error reaches tol (ir.hip). iters: refinement steps; inner: GMRES steps.
int gesv_gmres_ir(int64_t n, const double* a, int64_t lda, const double* b, double* x, int max_iter, int restart,
                  double tol, int* iters, int* inner, double* berr, hipStream_t s, int terms);

// Batched solves (batched.hip): a plan for `count` systems with orders n[i] <= 512 and k[i] <= 8 right-hand
// sides, A_i row-major at a + a_off[i], B_i row-major n×k at b + b_off[i] (overwritten by X_i where regular),
// perm_i at perm + p_off[i]; regular[i] = 1 where the LU solve ran, 0 where rustymath would take least
// squares. Host arrays describe device buffers. Mode 0 factors and solves; modes 1 and 2 solve A X = B and
// Aᵀ X = B with the factors and perm that a mode-0 run left. Tuning knobs: panel width (16 or 32) and
// workgroups per CU (1 of 512 threads, or 2 of 256). A run replays the plan's hipGraph (or launches directly) on `s`.
int batched_plan(int64_t count, const int64_t* n, const int64_t* k, const int64_t* a_off, const int64_t* b_off,
                 const int64_t* p_off, bool f64, void* a, void* b, int64_t* perm, int* regular, int mode,
                 int panel_width, int wgs_per_cu, long long* handle);
// The FP32-LU + FP64-refinement form of a plan over the same systems: a (f64) and b (f64) are read, never written;
// a32 (Σ n² FP32 items, a's offsets) gets float32(A_i)'s factors, r32 (Σ n k, b's offsets) is the FP32 solves'
// workspace, x (f64, b's offsets) the refined solutions; per system, perm, regular, the corrections applied (iters)
// and the final backward error (berr). A system stops once ||B - A X||_F / (||A||_F ||X||_F) <= tol_per_n n 2^-52 or
// after max_iter corrections; its test runs on the GPU. Runs and frees as batched_run / batched_free.
int batched_ir_plan(int64_t count, const int64_t* n, const int64_t* k, const int64_t* a_off, const int64_t* b_off,
                    const int64_t* p_off, const double* a, const double* b, double* x, float* a32, float* r32,
                    int64_t* perm, int* regular, int* iters, double* berr, int max_iter, double tol_per_n,
                    int panel_width, int wgs_per_cu, long long* handle);
int batched_run(long long handle, bool use_graph, hipStream_t s);
int batched_free(long long handle);
// The batched kernel's per-phase profile (workgroup 0, 100 MHz ticks): on = 1 clears and enables (runs then
// launch directly, not from the graph), 0 disables; out[8] gets the sums and, in out[7], the systems timed.
int batched_profile(int on, long long* out);

// The MFMA product's live profile: on = 1 clears and starts, 0 stops; gives the summed event time (ms), the
// number of launches, their algorithmic flops (2 m n k each) and bytes (a, b and c once each, c twice where it is
// read) since the last call.
int gemm_profile(int on, double* ms, int64_t* launches, double* flops, double* bytes);

// The panel kernel's per-phase profile: on = 1 clears and enables, 0 disables; out[8] gets tick sums.
int panel_profile(int on, long long* out);

// The LU's live profile: trail_rest and the panel launches bracketed by HIP events (see hs_profile_lu in hip_sys.h).
int lu_profile(int on, double* out);

// Converts between the device layouts: transpose (square or rectangular), f64 <-> f32 casts.
template <typename T>
int transpose(int64_t rows, int64_t cols, const T* src, int64_t lds, T* dst, int64_t ldd, hipStream_t s);
int cast_d2s(int64_t rows, int64_t cols, const double* src, int64_t lds, float* dst, int64_t ldd, hipStream_t s);
// dst = float(src) over a compact array of `count` items and part[b] = the sum of squares of src that block b of the
// `blocks`-block grid read (FP64): ||src||² is then the sum of the parts, the same on every run.
int cast_d2s_sumsq(int64_t count, const double* src, float* dst, double* part, int blocks, hipStream_t s);
int cast_s2d(int64_t rows, int64_t cols, const float* src, int64_t lds, double* dst, int64_t ldd, hipStream_t s);

}  // namespace hs
