/* hip_sys.h — the C ABI of rustymath's MI355X (gfx950) dense linear-algebra path.
 *
 * Conventions (those of the reference's vendor seam, core/src/linalg/accelerate.rs):
 *   - matrices are row-major with a leading dimension (CBLAS row-major, accelerate.rs:12), 64-bit sizes
 *     (ILP64, accelerate.rs:9); op codes 111 / 112 / 113 = Normal / Transpose / Adjoint (accelerate.rs:36-37),
 *     Adjoint being Transpose for real items;
 *   - LU factors are rustymath's: column-major n×n ("row j holds column j", src/linalg/stacked.rs:25),
 *     L strictly below the diagonal with a unit diagonal implied, U on and above; perm[i] is the source
 *     row of row i (core/src/linalg/factor.rs:132-134);
 *   - every entry point returns 0, or -i where argument i is invalid (LAPACK style), or > 0 for a HIP
 *     runtime failure (100000 = a device-side bounded wait ran out; 100001 = valid arguments beyond what this
 *     library handles, such as an LU of order 2^31 or more); hs_last_error() gives the text.
 *     A zero pivot is not an error (factor.rs:606): the factorizations report the first one in `info`; NaN propagates;
 *   - the caller owns every buffer it passes. Host-pointer entry points copy in and out through a
 *     device-memory pool the library keeps across calls (as LuWork keeps buffers, factor.rs:196-201).
 *     hs_dev_* entry points take device pointers and a hipStream_t (NULL = the null stream) and
 *     leave results in device memory;
 *   - thread safety: each host thread gets its own stream; calls from several threads may overlap.
 *
 * The arithmetic of every entry point is the reference's: a product takes its terms in order of p, one
 * fused multiply-add each; the factorization and solves follow lu_factor and lu_solve's orders of
 * operations (factor.rs:580-748).
 */
#ifndef HIP_SYS_H
#define HIP_SYS_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HS_NO_TRANS 111
#define HS_TRANS 112
#define HS_CONJ_TRANS 113
#define HS_UPPER 121
#define HS_LOWER 122
#define HS_NON_UNIT 131
#define HS_UNIT 132
#define HS_LEFT 141
#define HS_RIGHT 142

/* Library set-up and tear-down; hs_finalize frees the device-memory pool. hs_init also probes the current device's
 * workgroup placement (see hs_set_exchange). */
int hs_init(void);
int hs_finalize(void);
/* How the workgroups of an LU panel (above order 512) and of a single-column substitution exchange pivots and rows
 * (no counterpart in the reference, whose LU runs on CPU threads). mode 0, the default: through one XCD's L2 where a
 * probe of the current device (at hs_init, or at its first factorization or substitution) finds launch indices
 * congruent mod 8 on one XCD, as round-robin dispatch places them; otherwise, and always with mode 1, every word is
 * published at device scope, which is slower and independent of placement. Results never depend on the mode.
 * `local` (may be NULL) gets 1 where the current device now exchanges through one L2, 0 otherwise. */
int hs_set_exchange(int mode, int* local);
/* The calling thread's last error text. */
const char* hs_last_error(void);

/* c = alpha op(a) op(b) + beta c, row-major, m×k times k×n.
 * Replaces `Tiled::gemm` (core/src/linalg.rs:772, :799-807) and the CBLAS call of its vendor seam
 * `accelerate::product` (core/src/linalg/accelerate.rs:208-223); routes as the reference does:
 * m or n = 0: nothing; k = 0 or alpha = 0: c = beta c (`scale`, linalg.rs:1000); m or n = 1: `by_vector`
 * (linalg.rs:1020); otherwise the tiled product (linalg.rs:1353). */
int hs_dgemm(int op_a, int op_b, int64_t m, int64_t n, int64_t k, double alpha, const double* a, int64_t lda,
             const double* b, int64_t ldb, double beta, double* c, int64_t ldc);
int hs_sgemm(int op_a, int op_b, int64_t m, int64_t n, int64_t k, float alpha, const float* a, int64_t lda,
             const float* b, int64_t ldb, float beta, float* c, int64_t ldc);

/* `batch` products at once, c_i = alpha op(a_i) op(b_i) + beta c_i with cell i at a + i·sa, b + i·sb, c + i·sc (items;
 * a stride of 0 repeats an operand; cells of c may not overlap, else -16): a stack's `Inner::inner` over its frames
 * (src/linalg/stacked.rs over core/src/linalg.rs:118-122), each cell exactly as hs_dgemm computes it, in one launch.
 * Strides are arguments 9, 12 and 16, batch 17. */
int hs_dgemm_batched(int op_a, int op_b, int64_t m, int64_t n, int64_t k, double alpha, const double* a, int64_t lda,
                     int64_t sa, const double* b, int64_t ldb, int64_t sb, double beta, double* c, int64_t ldc,
                     int64_t sc, int64_t batch);
int hs_sgemm_batched(int op_a, int op_b, int64_t m, int64_t n, int64_t k, float alpha, const float* a, int64_t lda,
                     int64_t sa, const float* b, int64_t ldb, int64_t sb, float beta, float* c, int64_t ldc,
                     int64_t sc, int64_t batch);

/* Triangular solves in place on x (row-major m×n, ld ldx): op(A) X = B (side 141, A m×m) or X op(A) = B (side 142,
 * A n×n), A row-major with ld lda, of which only the triangle `uplo` names is read (121 upper, 122 lower; diag 131
 * non-unit, 132 unit diagonal taken as ones); trans 111 / 112 / 113 (113 is 112 for real items).
 * Replaces `Tiled::triangular` (core/src/linalg.rs:781-787), one of `dispatch!`'s three sites, and its vendor form
 * `accelerate::triangular` (core/src/linalg/accelerate.rs:67-70: the same codes, alpha = 1). Orders of operations are
 * `triangular_native`'s (core/src/linalg/factor.rs:711-721): `single` for one contiguous column (ldx = 1), the halving
 * `forward` / `backward` otherwise, a right-side solve as the left-side solve of the transposed view. The reference's
 * own callers pass (122, 131) or (121, 132); the other two pairs follow the same native orders. */
int hs_dtrsm(int side, int uplo, int trans, int diag, int64_t m, int64_t n, const double* a, int64_t lda, double* x,
             int64_t ldx);
int hs_strsm(int side, int uplo, int trans, int diag, int64_t m, int64_t n, const float* a, int64_t lda, float* x,
             int64_t ldx);

/* The LU factorization with partial pivoting of the row-major n×n matrix a: factors (n×n, column-major),
 * perm (n) and, where `pivots` is not NULL, the row swapped with row c at step c (n). Where `info` is not NULL it
 * gets 0, or 1 + the first column whose pivot is exactly zero (LAPACK's convention; the column is left undivided,
 * as factor.rs:606 leaves it, and the factorization goes on).
 * Replaces `Factor::matrix_lu_into` / `lu_factor` (core/src/linalg/factor.rs:85-87, :617-630). */
int hs_dlu_factor(const double* a, int64_t n, double* factors, int64_t* perm, int64_t* pivots, int64_t* info);
int hs_slu_factor(const float* a, int64_t n, float* factors, int64_t* perm, int64_t* pivots, int64_t* info);

/* x (n×k, row-major) with A x = b, from factors and perm as hs_*lu_factor gives them; an out-of-range
 * row position is argument -2. Replaces `Factor::matrix_lu_solve` / `lu_solve` (factor.rs:88, :633-641). */
int hs_dlu_solve(const double* factors, const int64_t* perm, int64_t n, const double* b, int64_t k, double* x);
int hs_slu_solve(const float* factors, const int64_t* perm, int64_t n, const float* b, int64_t k, float* x);

/* The inverse (row-major n×n) from factors and perm as hs_*lu_factor gives them; an out-of-range row position is
 * argument -2. Replaces `Factor::matrix_lu_inverse_into` / `lu_inverse` (factor.rs:91, :645-664) behind `Lu::inverse`
 * (:150-154), in the orders of rustymath's own kernels. */
int hs_dlu_inverse(const double* factors, const int64_t* perm, int64_t n, double* out);
int hs_slu_inverse(const float* factors, const int64_t* perm, int64_t n, float* out);

/* `LuWork::inverse_into` for a square a (factor.rs:215-219), which `pinv` takes for a regular matrix (:211-213):
 * out = a⁻¹ and *det = its determinant where every pivot exceeds eps·n·max|a| (*regular = 1); *regular = 0 and out
 * untouched where the reference would take the SVD's pseudoinverse, which this library leaves to the caller. */
int hs_dinverse(const double* a, int64_t n, double* out, double* det, int* regular);
int hs_sinverse(const float* a, int64_t n, float* out, float* det, int* regular);

/* `LuWork::solve` for a square a (factor.rs:207-226): x = a⁻¹ b through LU where every pivot exceeds
 * eps·n·max|a| (*regular = 1); *regular = 0 and x untouched where the reference would take least squares,
 * which this library leaves to the caller. */
int hs_dsolve(const double* a, int64_t n, const double* b, int64_t k, double* x, int* regular);
int hs_ssolve(const float* a, int64_t n, const float* b, int64_t k, float* x, int* regular);

/* The Cholesky factor L (row-major n×n, zeros above the diagonal) with L Lᵀ = a, reading only a's lower triangle;
 * *ok = 0 and l untouched where a is not positive definite (a diagonal item not > 0), where the reference gives None.
 * Replaces `Factor::matrix_cholesky` / `cholesky_factor` (factor.rs:93, :267, :777-797), in its order of operations. */
int hs_dcholesky(const double* a, int64_t n, double* l, int* ok);
int hs_scholesky(const float* a, int64_t n, float* l, int* ok);

/* Stacks of `count` cells of one order n (row-major cells back to back): `Stacked::solve` (src/linalg/stacked.rs:88-101)
 * gives x (count×n×k) and regular (count) as hs_*solve does for each cell (x holds b where regular = 0), and
 * `Stacked::lu` (:174-187) gives factors (count×n×n, each column-major as hs_*lu_factor gives them) and perm
 * (count×n). Orders up to 512 with at most 8 right-hand sides run as one batched launch (as the batched plan below,
 * bit for bit each cell's single-matrix result); larger cells run one by one. */
int hs_dsolve_stack(int64_t count, int64_t n, int64_t k, const double* a, const double* b, double* x, int* regular);
int hs_ssolve_stack(int64_t count, int64_t n, int64_t k, const float* a, const float* b, float* x, int* regular);
int hs_dlu_stack(int64_t count, int64_t n, const double* a, double* factors, int64_t* perm);
/* `Stacked::lu_solve` (src/linalg/stacked.rs:189-204): x (count×n×k) from reference-layout factors (count×n×n,
 * column-major cells) and perm (count×n, every item in [0, n), else -5), each cell as hs_*lu_solve gives it. Cells up
 * to order 512 with at most 8 right-hand sides and permutations for perm run as one batched launch. */
int hs_dlu_solve_stack(int64_t count, int64_t n, int64_t k, const double* factors, const int64_t* perm, const double* b,
                       double* x);
int hs_slu_solve_stack(int64_t count, int64_t n, int64_t k, const float* factors, const int64_t* perm, const float* b,
                       float* x);
int hs_slu_stack(int64_t count, int64_t n, const float* a, float* factors, int64_t* perm);

/* Uneven batched solves from host memory: `count` systems, system i being a[i] (row-major n[i]×n[i]) with b[i]
 * (row-major n[i]×k[i]), each solved as `LuWork::solve` (factor.rs:207-226) solves a cell of a stack over
 * `parallel::stack` (core/src/parallel.rs:118-192); the reference has no single call for systems of different orders
 * (a stack holds one order), so this is the per-cell loop a caller would write, as one persistent batched launch for
 * the systems up to order 512 with at most 8 right-hand sides (others one by one). x[i] (n[i]×k[i]) receives X_i where
 * regular[i] = 1 and b[i]'s copy where regular[i] = 0 (the reference would take least squares; an empty system counts
 * as not regular, factor.rs:222). regular may be NULL. Each X_i is bit for bit the single-matrix hs_*solve result. */
int hs_dgesv_batched(int64_t count, const int64_t* n, const int64_t* k, const double* const* a, const double* const* b,
                     double* const* x, int* regular);
int hs_sgesv_batched(int64_t count, const int64_t* n, const int64_t* k, const float* const* a, const float* const* b,
                     float* const* x, int* regular);

/* Device-resident forms (no counterpart in the reference, whose API takes host slices): pointers are
 * device memory, `stream` a hipStream_t (NULL = the null stream). */
int hs_dev_dgemm(int op_a, int op_b, int64_t m, int64_t n, int64_t k, double alpha, const double* a, int64_t lda,
                 const double* b, int64_t ldb, double beta, double* c, int64_t ldc, void* stream);
int hs_dev_sgemm(int op_a, int op_b, int64_t m, int64_t n, int64_t k, float alpha, const float* a, int64_t lda,
                 const float* b, int64_t ldb, float beta, float* c, int64_t ldc, void* stream);
/* hs_dtrsm on device memory (a, x) and `stream`; synchronises `stream` before returning. */
int hs_dev_dtrsm(int side, int uplo, int trans, int diag, int64_t m, int64_t n, const double* a, int64_t lda, double* x,
                 int64_t ldx, void* stream);
int hs_dev_strsm(int side, int uplo, int trans, int diag, int64_t m, int64_t n, const float* a, int64_t lda, float* x,
                 int64_t ldx, void* stream);
/* hs_dgemm_batched on device memory and `stream` (asynchronous, as hs_dev_dgemm). */
int hs_dev_dgemm_batched(int op_a, int op_b, int64_t m, int64_t n, int64_t k, double alpha, const double* a,
                         int64_t lda, int64_t sa, const double* b, int64_t ldb, int64_t sb, double beta, double* c,
                         int64_t ldc, int64_t sc, int64_t batch, void* stream);
int hs_dev_sgemm_batched(int op_a, int op_b, int64_t m, int64_t n, int64_t k, float alpha, const float* a,
                         int64_t lda, int64_t sa, const float* b, int64_t ldb, int64_t sb, float beta, float* c,
                         int64_t ldc, int64_t sc, int64_t batch, void* stream);
/* Factors the row-major a (ld lda) in place into column-major factors (same buffer, same ld) and writes
 * perm (device, n int64) and, where not NULL, *info (host) as hs_dlu_factor gives it. Synchronises `stream` before
 * returning. */
int hs_dev_dgetrf(int64_t n, double* a, int64_t lda, int64_t* perm, int64_t* info, void* stream);
int hs_dev_sgetrf(int64_t n, float* a, int64_t lda, int64_t* perm, int64_t* info, void* stream);
/* `LuWork::solve` (factor.rs:207-226) on device memory: factors a (ld lda) in place as hs_dev_dgetrf does, writes perm
 * (device), and sets *regular (host) to 1 where every pivot exceeds eps·n·max|a| (NaN counting as not negligible), then
 * solves b (n×k row-major, ld ldb, device) in place; *regular = 0 and b untouched where the reference would take least
 * squares, which this library leaves to the caller. Synchronises `stream`. */
int hs_dev_dgesv(int64_t n, double* a, int64_t lda, int64_t* perm, double* b, int64_t k, int64_t ldb, int* regular,
                 void* stream);
int hs_dev_sgesv(int64_t n, float* a, int64_t lda, int64_t* perm, float* b, int64_t k, int64_t ldb, int* regular,
                 void* stream);
/* Solves in place on b (n×k row-major, ld ldb >= k) with device factors (ld ldf >= n): A x = b (trans = 0) in the
 * reference's substitution orders, or Aᵀ x = b (trans = 1, for adjoints). */
int hs_dev_dgetrs(int trans, int64_t n, int64_t k, const double* factors, int64_t ldf, const int64_t* perm, double* b,
                  int64_t ldb, void* stream);
int hs_dev_sgetrs(int trans, int64_t n, int64_t k, const float* factors, int64_t ldf, const int64_t* perm, float* b,
                  int64_t ldb, void* stream);
/* hs_dev_dgetrs that allocates nothing (no counterpart in the reference): for callers that cannot allocate device
 * memory, such as code under HIP graph capture or a serving loop that forbids it. The same arguments, results and
 * bits as hs_dev_dgetrs; its workspace is a static array of the library in each device's memory (4 MiB + 256 B, held
 * from a device's first call until the process ends), which limits it to n <= 65536 (else -2), k <= 16 and n·k <=
 * 2^18 (else -3). Every call on one device shares that device's array: calls that overlap on one device race, so a
 * caller serialises them. Synchronises `stream` before returning. */
int hs_dev_dgetrs_noalloc(int trans, int64_t n, int64_t k, const double* factors, int64_t ldf, const int64_t* perm,
                          double* b, int64_t ldb, void* stream);
/* The inverse from device factors (ld ldf) and perm into out (row-major, ld ldo), as hs_dlu_inverse computes it. */
int hs_dev_dgetri(int64_t n, const double* factors, int64_t ldf, const int64_t* perm, double* out, int64_t ldo,
                  void* stream);
int hs_dev_sgetri(int64_t n, const float* factors, int64_t ldf, const int64_t* perm, float* out, int64_t ldo,
                  void* stream);

/* The Cholesky factor in place on device memory (ld lda); *ok (host) as hs_dcholesky gives it. Synchronises `stream`. */
int hs_dev_dpotrf(int64_t n, double* a, int64_t lda, int* ok, void* stream);
int hs_dev_spotrf(int64_t n, float* a, int64_t lda, int* ok, void* stream);

/* Mixed precision (no counterpart in the reference, which has no iterative refinement): solves A x = b for
 * the n×n f64 matrix a (row-major, lda = n) and one right-hand side b (device, n f64) by factoring a copy of A
 * in FP32 on MFMA and refining x in FP64 on the GPU until the normwise backward error
 * ||b - A x||_2 / (||A||_F ||x||_2) is at most tol or max_iter corrections have run. Writes x (device, n),
 * the number of corrections and the final backward error. A is read, never written. */
int hs_dev_dsgesv(int64_t n, const double* a, int64_t lda, const double* b, double* x, int max_iter, double tol,
                  int* iters, double* berr, void* stream);
/* The BF16 factorization's trailing product on its own: c -= a b for row-major FP32 operands (a m×k, b k×n, c m×n)
 * taken as bfloat16, products summed in FP32 on MFMA: terms = 1 rounds each item to bfloat16 (nearest even); terms = 3
 * splits each item into hi = bf16(x) and lo = bf16(x - hi) and adds a_lo b_hi, a_hi b_lo and a_hi b_hi. */
int hs_dev_sgemm_bf16_update(int64_t m, int64_t n, int64_t k, const float* a, int64_t lda, const float* b, int64_t ldb,
                             float* c, int64_t ldc, int terms, void* stream);
/* The BF16 form of hs_dev_dsgesv (no counterpart in the reference): the factorization's trailing products take their
 * operands as bfloat16 (`terms` as in hs_dev_sgemm_bf16_update: 1 = rounded once, 3 = split pairs) and sum in FP32
 * on MFMA (panels and rows of U in FP32), and each refinement step solves A d = r by GMRES in FP64,
 * left-preconditioned by those factors and restarted every `restart` (1..200) steps, since classical refinement cannot
 * converge from such factors at cond(A) near 1e6. iters = refinement steps, inner = GMRES steps (each one FP64
 * product with A and one pair of triangular solves). */
int hs_dev_dbgesv(int64_t n, const double* a, int64_t lda, const double* b, double* x, int max_iter, int restart,
                  int terms, double tol, int* iters, int* inner, double* berr, void* stream);

/* Batched uneven solves: the per-cell solves of a stack, which rustymath runs one `LuWork::solve` per Rayon
 * worker (src/linalg/stacked.rs:88-101 over core/src/parallel.rs:118-192 and core/src/linalg/factor.rs:207-226),
 * as one persistent launch over all cells. A plan covers `count` systems in device buffers: A_i (n[i]×n[i],
 * row-major, n[i] <= 512) at a + a_off[i], B_i (n[i]×k[i], row-major, k[i] <= 8) at b + b_off[i], and perm_i
 * (n[i] int64) at perm + p_off[i]; n, k and the offsets are host arrays.
 * mode 0: a run overwrites each A_i with its factors (row perm_i[j] holds position j, rows are not moved) and,
 *   where regular[i] = 1, B_i with X_i, bit for bit the reference's `solve` of that cell; regular[i] = 0 marks a
 *   cell the reference would solve by least squares, whose B_i is left as it was.
 * mode 1 / 2: solves A_i X = B_i (the reference's `lu_solve`, factor.rs:633-641) or A_iᵀ X = B_i (Uᵀ, Lᵀ, then
 *   the inverse permutation; no counterpart in the reference, for adjoints) in place on B_i with the factors
 *   and perm that a mode-0 run left; `regular` may be NULL.
 * panel_width (16 or 32) and wgs_per_cu (1: a workgroup of 512 threads per CU; 2: two of 256 threads, so two
 * systems per CU) are tuning knobs that leave the results unchanged. A run with
 * graph = 1 replays a hipGraph captured at the plan's first run. The plan keeps the pointers; runs that overlap
 * on one plan race. */
int hs_dev_dgesv_batched_plan(int64_t count, const int64_t* n, const int64_t* k, const int64_t* a_off,
                              const int64_t* b_off, const int64_t* p_off, double* a, double* b, int64_t* perm,
                              int* regular, int mode, int panel_width, int wgs_per_cu, int64_t* plan);
int hs_dev_sgesv_batched_plan(int64_t count, const int64_t* n, const int64_t* k, const int64_t* a_off,
                              const int64_t* b_off, const int64_t* p_off, float* a, float* b, int64_t* perm,
                              int* regular, int mode, int panel_width, int wgs_per_cu, int64_t* plan);
/* The FP32-LU + FP64-refinement form of the batched solves (no counterpart in the reference, which has no refinement):
 * the same systems and offsets; a and b (f64, device) are read, never written. Each system is factored in FP32 on
 * MFMA (a32: Σ n² FP32 items at a's offsets, left holding float32(A_i)'s factors as a mode-0 plan leaves them, with
 * perm and regular), solved, then refined on the GPU: X += the FP32 solve of the scaled residual, R = B - A X in
 * FP64, until the normwise backward error ||R||_F / (||A||_F ||X||_F) is at most tol_per_n · n_i · 2^-52 or
 * max_iter (0..100) corrections have run. r32 (Σ n k FP32 items at b's offsets) is workspace; x (f64, b's offsets)
 * gets X_i, iters[i] the corrections applied and berr[i] the final backward error (NaN where regular[i] = 0, whose x
 * is left as it was). One run is the whole sequence (one persistent launch per pass), replayed from a hipGraph with
 * graph = 1; run and free it as any plan. */
int hs_dev_dsgesv_batched_plan(int64_t count, const int64_t* n, const int64_t* k, const int64_t* a_off,
                               const int64_t* b_off, const int64_t* p_off, const double* a, const double* b, double* x,
                               float* a32, float* r32, int64_t* perm, int* regular, int* iters, double* berr,
                               int max_iter, double tol_per_n, int panel_width, int wgs_per_cu, int64_t* plan);
int hs_dev_gesv_batched_run(int64_t plan, int graph, void* stream);
int hs_dev_gesv_batched_free(int64_t plan);

/* Profiling (no counterpart in the reference): on = 1 clears and starts the panel kernel's per-phase timers
 * (100 MHz ticks summed over launches: candidate search, publish, exchange wait, pivot bookkeeping, update),
 * on = 0 stops them; `out` (8 values, may be NULL) receives the sums. */
int hs_profile_panel(int on, int64_t* out);
/* Profiling of the MFMA product: on = 1 clears and starts bracketing every product launch with HIP events on its
 * stream, on = 0 stops; gives the launches' summed time (ms), their count, their algorithmic flops (2 m n k) and
 * algorithmic bytes (a and b read once, c written once and, where beta is not 0, read once). Any output may be NULL. */
int hs_profile_gemm(int on, double* ms, int64_t* launches, double* flops, double* bytes);
/* Profiling of the LU (no counterpart in the reference): the FP32 and FP64 factorizations of order above 512 run each
 * block's trailing product partly on the panel launches' idle workgroups (their workers) and the remaining tiles in
 * trail_rest. on = 1 clears and starts bracketing every panel and trail_rest launch with HIP events on its stream;
 * on = 0 stops. out (8 values, may be NULL) gets, since the last call: trail_rest's summed time (ms), launches, the
 * flops of its tiles (2 rows cols k each) and their algorithmic bytes (the tiles' items of c read and written once,
 * the rows of a and columns of b they touch read once); then the panel launches' summed time (ms), launches, their
 * own factorization flops and the flops of the tiles their workers computed. */
int hs_profile_lu(int on, double* out);
/* Profiling of the batched solves: on = 1 clears and starts per-phase timers in workgroup 0 (100 MHz ticks: setup,
 * panel columns, trailing update, regularity check and L, U, then in out[5] the panels' rows of U), so that runs
 * launch directly rather than replay their graph; on = 0 stops. `out` (8 values, may be NULL) gets the sums, out[7]
 * the systems timed. */
int hs_profile_batched(int on, int64_t* out);

#ifdef __cplusplus
}
#endif
#endif
