/* oracle.c — CPU restatement of rustymath's dense linear-algebra hot path (matrix products, LU with
 * partial pivoting, LU solves, inverse and determinants) at f64 and f32.
 *
 * TEST INFRASTRUCTURE ONLY. Tests, the smoke check and the CPU-baseline leg of the benchmark load it as
 * the checker; the HIP library never links, loads or calls it, and no product path falls back to it.
 *
 * Reference (AnswerDotAI/rustymath, Rust, no Rust toolchain in this image, so restated here):
 *   core/src/linalg.rs          gemm :799-807, scale :1000-1008, by_vector :1020-1073 (dots :903,
 *                               narrow_dots :923, narrow_sums :941, scaled_runs :883), native_gemm
 *                               :1077-1081, product :1353-1426 (KC/MC/NC :389-395), pack :1429,
 *                               the x86-64 AVX2 tiles times64 6x8 and times32 6x16 (:412-413, :1282-1283)
 *   core/src/linalg/factor.rs   stride :72, LEAF :79, NS :82, Lu :132-194, LuWork::solve/regular
 *                               :207-226, largest :476, greatest :488, divide :506, swap_rows :572,
 *                               lu_panel :580-615, lu_factor :617-630, lu_solve :633-641,
 *                               lu_inverse :645-664, forward :668, backward :693, triangular_native
 *                               :711, single :725
 *   core/src/linalg/scalar.rs   mul_acc = fused multiply-add :52; core/src/float.rs:5 (FUSED on x86-64-v3)
 *
 * Pinned by tests/test_oracle.py against the reference's own known answers and invariants; see the
 * test file's docstrings for each check. Built by oracle/Makefile with -O3 -march=x86-64-v3, the
 * reference's target (.cargo/config.toml:5-7), and -ffp-contract=off so that only the explicit
 * fused multiply-adds below fuse. */
#include <immintrin.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_KC 256
#define OR_MC 128
#define OR_NC 2048

/* Threads for the products' bands of rows: the reference's `set_threads` (core/src/parallel.rs:15). */
void or_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : omp_get_num_procs());
#else
    (void)n;
#endif
}
int or_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* `times64` (core/src/linalg.rs:1282): a 6x8 tile of f64 in twelve 4-lane registers, one
 * `_mm256_fmadd_pd` per item per step. `load` starts from c, otherwise from zero. */
static void micro_d(int64_t kc, const double* ap, const double* bp, double* c, int64_t ldc, int rows, int cols, int load) {
    __m256d acc[6][2];
    double t[6][8];
    int whole = rows == 6 && cols == 8;
    for (int r = 0; r < 6; ++r) {
        if (!load) { acc[r][0] = _mm256_setzero_pd(); acc[r][1] = _mm256_setzero_pd(); continue; }
        if (whole) { acc[r][0] = _mm256_loadu_pd(c + r * ldc); acc[r][1] = _mm256_loadu_pd(c + r * ldc + 4); continue; }
        for (int s = 0; s < 8; ++s) t[r][s] = r < rows && s < cols ? c[r * ldc + s] : 0.0;
        acc[r][0] = _mm256_loadu_pd(t[r]); acc[r][1] = _mm256_loadu_pd(t[r] + 4);
    }
    for (int64_t p = 0; p < kc; ++p) {
        __m256d b0 = _mm256_loadu_pd(bp + p * 8), b1 = _mm256_loadu_pd(bp + p * 8 + 4);
        for (int r = 0; r < 6; ++r) {
            __m256d a = _mm256_set1_pd(ap[p * 6 + r]);
            acc[r][0] = _mm256_fmadd_pd(a, b0, acc[r][0]);
            acc[r][1] = _mm256_fmadd_pd(a, b1, acc[r][1]);
        }
    }
    for (int r = 0; r < 6; ++r) {
        if (whole) { _mm256_storeu_pd(c + r * ldc, acc[r][0]); _mm256_storeu_pd(c + r * ldc + 4, acc[r][1]); continue; }
        if (r >= rows) continue;
        _mm256_storeu_pd(t[r], acc[r][0]); _mm256_storeu_pd(t[r] + 4, acc[r][1]);
        for (int s = 0; s < cols; ++s) c[r * ldc + s] = t[r][s];
    }
}

/* `times32` (core/src/linalg.rs:1283): a 6x16 tile of f32 in twelve 8-lane registers. */
static void micro_s(int64_t kc, const float* ap, const float* bp, float* c, int64_t ldc, int rows, int cols, int load) {
    __m256 acc[6][2];
    float t[6][16];
    int whole = rows == 6 && cols == 16;
    for (int r = 0; r < 6; ++r) {
        if (!load) { acc[r][0] = _mm256_setzero_ps(); acc[r][1] = _mm256_setzero_ps(); continue; }
        if (whole) { acc[r][0] = _mm256_loadu_ps(c + r * ldc); acc[r][1] = _mm256_loadu_ps(c + r * ldc + 8); continue; }
        for (int s = 0; s < 16; ++s) t[r][s] = r < rows && s < cols ? c[r * ldc + s] : 0.0f;
        acc[r][0] = _mm256_loadu_ps(t[r]); acc[r][1] = _mm256_loadu_ps(t[r] + 8);
    }
    for (int64_t p = 0; p < kc; ++p) {
        __m256 b0 = _mm256_loadu_ps(bp + p * 16), b1 = _mm256_loadu_ps(bp + p * 16 + 8);
        for (int r = 0; r < 6; ++r) {
            __m256 a = _mm256_set1_ps(ap[p * 6 + r]);
            acc[r][0] = _mm256_fmadd_ps(a, b0, acc[r][0]);
            acc[r][1] = _mm256_fmadd_ps(a, b1, acc[r][1]);
        }
    }
    for (int r = 0; r < 6; ++r) {
        if (whole) { _mm256_storeu_ps(c + r * ldc, acc[r][0]); _mm256_storeu_ps(c + r * ldc + 8, acc[r][1]); continue; }
        if (r >= rows) continue;
        _mm256_storeu_ps(t[r], acc[r][0]); _mm256_storeu_ps(t[r] + 8, acc[r][1]);
        for (int s = 0; s < cols; ++s) c[r * ldc + s] = t[r][s];
    }
}

#define T double
#define SUF d
#define FMA fma
#define ABS fabs
#define LOG log
#define SQRT sqrt
#define TMIN DBL_MIN
#define EPS DBL_EPSILON
#define MRT 6
#define NRT 8
#define MICRO micro_d
#include "oracle_body.h"
#undef T
#undef SUF
#undef FMA
#undef ABS
#undef LOG
#undef SQRT
#undef TMIN
#undef EPS
#undef MRT
#undef NRT
#undef MICRO

#define T float
#define SUF s
#define FMA fmaf
#define ABS fabsf
#define LOG logf
#define SQRT sqrtf
#define TMIN FLT_MIN
#define EPS FLT_EPSILON
#define MRT 6
#define NRT 16
#define MICRO micro_s
#include "oracle_body.h"
