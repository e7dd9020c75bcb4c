/* Type-generic body of the oracle, included twice by oracle.c: once with T = double (suffix _d)
 * and once with T = float (suffix _s). TEST INFRASTRUCTURE ONLY (see oracle.c's header).
 *
 * Every function restates one function of rustymath (reference files cited as path:line) at the
 * level of its floating-point operations: which products are fused, in which order terms add,
 * where a reciprocal multiplies and where a quotient divides. */

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)
#define F(name) CAT(name, SUF)

/* A strided matrix view in a logical orientation, as `Matrix` in core/src/linalg.rs:646.
 * `len` is the number of items the view's slice holds from `data` on (the slice length that
 * `narrow_sums` reads, core/src/linalg.rs:921). */
typedef struct { const T* data; int64_t ld; int trans; int64_t len; } F(view_);

static inline T F(at_)(F(view_) v, int64_t i, int64_t j) { return v.trans ? v.data[j * v.ld + i] : v.data[i * v.ld + j]; }
static inline F(view_) F(offset_)(F(view_) v, int64_t row, int64_t col) {
    int64_t at = v.trans ? col * v.ld + row : row * v.ld + col;
    F(view_) r = v; r.data += at; r.len -= at; return r;
}
static inline F(view_) F(transposed_)(F(view_) v) { F(view_) r = v; r.trans = !v.trans; return r; }
static inline F(view_) F(rows_)(const T* data, int64_t ld, int64_t len) { F(view_) r = {data, ld, 0, len}; return r; }

/* `Scalar::mul_acc` on x86-64-v3: one fused multiply-add (core/src/linalg/scalar.rs:52, core/src/float.rs:5,172). */
static inline T F(mla_)(T x, T a, T y) { return FMA(x, a, y); }

/* `scale`, core/src/linalg.rs:1000-1008 (Part::Whole). */
static void F(scale_)(T* c, int64_t ldc, int64_t m, int64_t n, T beta) {
    if (beta == (T)1) return;
    for (int64_t i = 0; i < m; ++i)
        for (int64_t j = 0; j < n; ++j) { T* x = &c[i * ldc + j]; *x = beta == (T)0 ? (T)0 : beta * *x; }
}

/* `sum_of_products`/`dot`, core/src/linalg/factor.rs:460-473: eight running sums over whole chunks of
 * eight, folded in order from zero, then the remaining terms in order. */
static T F(dot_)(const T* x, const T* y, int64_t k) {
    T s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int64_t whole = k / 8 * 8;
    for (int64_t c = 0; c < whole; c += 8)
        for (int l = 0; l < 8; ++l) s[l] = F(mla_)(x[c + l], y[c + l], s[l]);
    T t = 0;
    for (int l = 0; l < 8; ++l) t = t + s[l];
    for (int64_t p = whole; p < k; ++p) t = F(mla_)(x[p], y[p], t);
    return t;
}

/* `dots`, core/src/linalg.rs:903-918: rows in quads from `first`, each with eight running sums that
 * add pairwise, then the remaining terms in order; rows past the last quad take `dot`. */
static void F(dots_)(T* out, int64_t rows, const T* data, int64_t ld, int64_t first, const T* x, int64_t k) {
    int64_t quads = rows / 4, whole = k / 8 * 8;
    for (int64_t q = 0; q < quads; ++q) {
        T s[4][8];
        for (int r = 0; r < 4; ++r) for (int l = 0; l < 8; ++l) s[r][l] = 0;
        for (int64_t c = 0; c < whole; c += 8)
            for (int r = 0; r < 4; ++r) {
                const T* a = &data[(first + 4 * q + r) * ld + c];
                for (int l = 0; l < 8; ++l) s[r][l] = F(mla_)(a[l], x[c + l], s[r][l]);
            }
        for (int r = 0; r < 4; ++r) {
            T* v = s[r];
            T t = ((v[0] + v[4]) + (v[2] + v[6])) + ((v[1] + v[5]) + (v[3] + v[7]));
            const T* row = &data[(first + 4 * q + r) * ld];
            for (int64_t p = whole; p < k; ++p) t = F(mla_)(row[p], x[p], t);
            out[4 * q + r] = t;
        }
    }
    for (int64_t i = quads * 4; i < rows; ++i) out[i] = F(dot_)(&data[(first + i) * ld], x, k);
}

/* `narrow_dots`, core/src/linalg.rs:923-938: stored rows of at most 8 items, two running sums. */
static void F(narrow_dots_)(T* out, int64_t rows, const T* data, int64_t ld, int64_t first, const T* x, int64_t k) {
    for (int64_t i = 0; i < rows; ++i) {
        const T* a = &data[(first + i) * ld];
        T s[2] = {0, 0};
        for (int64_t l = 0; l < k; ++l) s[l % 2] = F(mla_)(a[l], x[l], s[l % 2]);
        out[i] = s[0] + s[1];
    }
}

/* `narrow_sums`, core/src/linalg.rs:941-956: at most 8 results, four sets of running sums that take
 * every fourth stored row, whole rows only while 8 items stay inside the slice. */
static void F(narrow_sums_)(T* out, int64_t n, const T* data, int64_t len, int64_t ld, const T* w, int64_t wn) {
    int64_t whole;
    if (len < 8) whole = 0;
    else { whole = (len - 8) / (ld > 1 ? ld : 1) + 1; if (whole > wn) whole = wn; }
    T acc[4][8];
    for (int s = 0; s < 4; ++s) for (int t = 0; t < 8; ++t) acc[s][t] = 0;
    int64_t quads = whole / 4;
    for (int64_t q = 0; q < quads; ++q)
        for (int s = 0; s < 4; ++s) {
            const T* a = &data[(4 * q + s) * ld];
            for (int t = 0; t < 8; ++t) acc[s][t] = F(mla_)(a[t], w[4 * q + s], acc[s][t]);
        }
    for (int64_t p = quads * 4; p < whole; ++p) {
        const T* a = &data[p * ld];
        for (int t = 0; t < 8; ++t) acc[0][t] = F(mla_)(a[t], w[p], acc[0][t]);
    }
    for (int64_t p = whole; p < wn; ++p)
        for (int64_t t = 0; t < n; ++t) acc[0][t] = F(mla_)(data[p * ld + t], w[p], acc[0][t]);
    for (int64_t t = 0; t < n; ++t) out[t] = (acc[0][t] + acc[2][t]) + (acc[1][t] + acc[3][t]);
}

/* `by_vector`, core/src/linalg.rs:1020-1073, at rustymath's own kernels on one thread: a matrix times a
 * vector where m or n is 1. A transposed matrix adds its stored rows into each result in order
 * (`scaled_runs`, :883-896) or, for at most 8 results, through `narrow_sums`. Gives 0 where it declines. */
static int F(by_vector_)(T* c, int64_t ldc, F(view_) a, T alpha, F(view_) b, int64_t m, int64_t k, int64_t n, T beta) {
    F(view_) matrix, vector; int64_t rows, step;
    if (n == 1) { matrix = a; vector = F(transposed_)(b); rows = m; step = ldc; }
    else if (m == 1) { matrix = F(transposed_)(b); vector = a; rows = n; step = 1; }
    else return 0;
    int narrow = (matrix.trans ? rows : k) <= 8;
    int64_t stride = vector.trans ? vector.ld : 1;
    T* x = (T*)malloc(sizeof(T) * (size_t)k);
    for (int64_t p = 0; p < k; ++p) x[p] = vector.data[p * stride];
    int one = alpha == (T)1, start = beta == (T)0;
    T* y = (T*)calloc((size_t)rows, sizeof(T));
    if (matrix.trans) {
        if (narrow) F(narrow_sums_)(y, rows, matrix.data, matrix.len, matrix.ld, x, k);
        else
            for (int64_t t = 0; t < rows; ++t) {
                T acc = 0;
                for (int64_t p = 0; p < k; ++p) acc = F(mla_)(x[p], matrix.data[p * matrix.ld + t], acc);
                y[t] = acc;
            }
    } else if (k <= 8) F(narrow_dots_)(y, rows, matrix.data, matrix.ld, 0, x, k);
    else F(dots_)(y, rows, matrix.data, matrix.ld, 0, x, k);
    for (int64_t i = 0; i < rows; ++i) {
        T v = one ? y[i] : alpha * y[i];
        T* z = &c[i * step];
        *z = start ? v : F(mla_)(beta, *z, v);
    }
    free(x); free(y);
    return 1;
}

/* Packing for `product` (core/src/linalg.rs:1429-1434 `pack`, :737-760 `strip`): strips of MR rows of
 * `a` times `alpha` (only where alpha is not 1, as `Matrix::strip` scales, :719-727), and strips of NR
 * columns of `b`, zero past the edge. */
static void F(pack_a_)(T* out, F(view_) a, int64_t i0, int64_t mc, int64_t p0, int64_t kc, T alpha) {
    int scale = alpha != (T)1;
    for (int64_t s = 0; s * MRT < mc; ++s)
        for (int64_t p = 0; p < kc; ++p)
            for (int l = 0; l < MRT; ++l) {
                int64_t r = s * MRT + l;
                T v = r < mc ? F(at_)(a, i0 + r, p0 + p) : (T)0;
                out[(s * kc + p) * MRT + l] = r < mc && scale ? alpha * v : v;
            }
}
static void F(pack_b_)(T* out, F(view_) b, int64_t j0, int64_t nc, int64_t p0, int64_t kc) {
    for (int64_t s = 0; s * NRT < nc; ++s)
        for (int64_t p = 0; p < kc; ++p)
            for (int l = 0; l < NRT; ++l) {
                int64_t j = s * NRT + l;
                out[(s * kc + p) * NRT + l] = j < nc ? F(at_)(b, p0 + p, j0 + j) : (T)0;
            }
}

/* `product`, core/src/linalg.rs:1353-1426, with KC = 256, MC = 128, NC = 2048 (:389-395). Each result
 * starts from zero (`start`) or from `c`, then takes one fused multiply-add per step of p, in order:
 * the tile kernel reloads `c` between blocks of KC steps, so blocking never changes a result's bits.
 * Bands of MC rows split across OpenMP threads (the reference's `pieces_of`, :1389). */
static void F(product_)(T* c, int64_t ldc, F(view_) a, T alpha, F(view_) b, int64_t m, int64_t k, int64_t n, int start) {
    int64_t nstrips = (OR_NC + NRT - 1) / NRT;
    T* bp = (T*)malloc(sizeof(T) * (size_t)(nstrips * NRT * OR_KC));
    for (int64_t jc = 0; jc < n; jc += OR_NC) {
        int64_t nc = n - jc < OR_NC ? n - jc : OR_NC;
        for (int64_t pc = 0; pc < k; pc += OR_KC) {
            int64_t kc = k - pc < OR_KC ? k - pc : OR_KC;
            F(pack_b_)(bp, b, jc, nc, pc, kc);
            int load = pc > 0 || !start;
            int64_t blocks = (m + OR_MC - 1) / OR_MC;
            int64_t work = m * kc * nc;
#pragma omp parallel for schedule(dynamic) if (work >= (1 << 20))
            for (int64_t blk = 0; blk < blocks; ++blk) {
                int64_t ic = blk * OR_MC, mc = m - ic < OR_MC ? m - ic : OR_MC;
                T* ap = (T*)malloc(sizeof(T) * (size_t)(((mc + MRT - 1) / MRT) * MRT * kc));
                F(pack_a_)(ap, a, ic, mc, pc, kc, alpha);
                for (int64_t jr = 0; jr < nc; jr += NRT)
                    for (int64_t ir = 0; ir < mc; ir += MRT) {
                        int rows = (int)(mc - ir < MRT ? mc - ir : MRT), cols = (int)(nc - jr < NRT ? nc - jr : NRT);
                        MICRO(kc, ap + ir * kc, bp + jr * kc, c + (ic + ir) * ldc + jc + jr, ldc, rows, cols, load);
                    }
                free(ap);
            }
        }
    }
    free(bp);
}

/* `gemm`, core/src/linalg.rs:799-807, with `native_gemm` (:1077-1081): c = alpha a b + beta c. */
static void F(gemm_)(T* c, int64_t ldc, F(view_) a, F(view_) b, int64_t m, int64_t k, int64_t n, T alpha, T beta) {
    if (m == 0 || n == 0) return;
    if (k == 0 || alpha == (T)0) { F(scale_)(c, ldc, m, n, beta); return; }
    if (F(by_vector_)(c, ldc, a, alpha, b, m, k, n, beta)) return;
    int start = beta == (T)0;
    if (!start) F(scale_)(c, ldc, m, n, beta);
    F(product_)(c, ldc, a, alpha, b, m, k, n, start);
}

/* Row-major CBLAS-style entry: op codes 111 (Normal), 112 (Transpose), 113 (Adjoint = Transpose for
 * real items), as accelerate.rs:36-37 maps `Op`. Extents checked as `Matrix::check` does (:702-707).
 * Returns 0, or minus the position of the first bad argument. */
int F(or_gemm_)(int opa, int opb, int64_t m, int64_t n, int64_t k, T alpha, const T* a, int64_t lda,
                const T* b, int64_t ldb, T beta, T* c, int64_t ldc) {
    if (opa < 111 || opa > 113) return -1;
    if (opb < 111 || opb > 113) return -2;
    if (m < 0) return -3;
    if (n < 0) return -4;
    if (k < 0) return -5;
    int ta = opa != 111, tb = opb != 111;
    if (m > 0 && k > 0 && lda < (ta ? m : k)) return -8;
    if (k > 0 && n > 0 && ldb < (tb ? k : n)) return -10;
    if (m > 0 && n > 0 && ldc < n) return -13;
    int64_t alen = (ta ? k : m) > 0 ? ((ta ? k : m) - 1) * lda + (ta ? m : k) : 0;
    int64_t blen = (tb ? n : k) > 0 ? ((tb ? n : k) - 1) * ldb + (tb ? k : n) : 0;
    F(view_) va = {a, lda, ta, alen}, vb = {b, ldb, tb, blen};
    F(gemm_)(c, ldc, va, vb, m, k, n, alpha, beta);
    return 0;
}

/* The product's defining fold, item by item, for pinning the blocked kernels: every result starts from
 * zero or beta c and takes alpha a[i][p] b[p][j] as one fused multiply-add per p, in order. */
int F(or_gemm_naive_)(int opa, int opb, int64_t m, int64_t n, int64_t k, T alpha, const T* a, int64_t lda,
                      const T* b, int64_t ldb, T beta, T* c, int64_t ldc) {
    int ta = opa != 111, tb = opb != 111;
    for (int64_t i = 0; i < m; ++i)
        for (int64_t j = 0; j < n; ++j) {
            T acc;
            if (k == 0 || alpha == (T)0) { c[i * ldc + j] = beta == (T)1 ? c[i * ldc + j] : beta == (T)0 ? (T)0 : beta * c[i * ldc + j]; continue; }
            acc = beta == (T)0 ? (T)0 : beta == (T)1 ? c[i * ldc + j] : beta * c[i * ldc + j];
            for (int64_t p = 0; p < k; ++p) {
                T x = ta ? a[p * lda + i] : a[i * lda + p];
                T y = tb ? b[j * ldb + p] : b[p * ldb + j];
                acc = FMA(alpha == (T)1 ? x : alpha * x, y, acc);
            }
            c[i * ldc + j] = acc;
        }
    return 0;
}

/* `divide`, core/src/linalg/factor.rs:506-508: a reciprocal multiply, or a quotient for a subnormal divisor. */
static void F(divide_)(T* x, int64_t n, int64_t stride, T d) {
    if (ABS(d) >= TMIN) { T r = (T)1 / d; for (int64_t i = 0; i < n; ++i) x[i * stride] = x[i * stride] * r; }
    else for (int64_t i = 0; i < n; ++i) x[i * stride] = x[i * stride] / d;
}

/* `largest` (factor.rs:476) with `greatest` (:488) and `Core::maximum` (core/src/float.rs:200): the
 * first position of the greatest magnitude, NaN ignored, 0 where all are NaN. */
static int64_t F(largest_)(const T* x, int64_t n) {
    T top = 0;
    for (int64_t i = 0; i < n; ++i) { T v = ABS(x[i]); top = v > top ? v : top; }
    for (int64_t i = 0; i < n; ++i) if (ABS(x[i]) == top) return i;
    return 0;
}

/* `forward`, factor.rs:668-689. `x` holds rows r0..r1, `k` items each, `ldx` apart. */
static void F(forward_)(F(view_) l, int64_t r0, int64_t r1, T* x, int64_t xlen, int64_t ldx, int64_t k, int64_t block, int lower, int unit) {
    if (!lower && r1 - r0 > block) {
        int64_t mid = (r0 + r1) / 2;
        T* bottom = x + (mid - r0) * ldx;
        F(forward_)(l, r0, mid, x, (mid - r0) * ldx, ldx, k, block, 0, unit);
        F(gemm_)(bottom, ldx, F(offset_)(l, mid, r0), F(rows_)(x, ldx, (mid - r0) * ldx), r1 - mid, mid - r0, k, (T)-1, (T)1);
        F(forward_)(l, mid, r1, bottom, xlen - (mid - r0) * ldx, ldx, k, block, 0, unit);
        return;
    }
    for (int64_t i0 = r0; i0 < r1; i0 += block) {
        int64_t i1 = i0 + block < r1 ? i0 + block : r1;
        int64_t width = lower ? i1 : k;
        T* rows = x + (i0 - r0) * ldx;
        for (int64_t i = i0; i < i1; ++i) {
            T* xi = rows + (i - i0) * ldx;
            for (int64_t j = i0; j < i; ++j) {
                T a = -F(at_)(l, i, j);
                const T* xj = rows + (j - i0) * ldx;
                for (int64_t t = 0; t < width; ++t) xi[t] = F(mla_)(xj[t], a, xi[t]);
            }
            if (!unit) F(divide_)(xi, width, 1, F(at_)(l, i, i));
        }
        if (i1 < r1)
            F(gemm_)(x + (i1 - r0) * ldx, ldx, F(offset_)(l, i1, i0), F(rows_)(rows, ldx, (i1 - i0) * ldx), r1 - i1, i1 - i0, width, (T)-1, (T)1);
    }
    (void)xlen;
}

/* `backward`, factor.rs:693-707, with NS = 32 (:82). */
static void F(backward_)(F(view_) u, int64_t r0, int64_t r1, T* x, int64_t ldx, int64_t k, int unit) {
    if (r1 - r0 > 32) {
        int64_t mid = (r0 + r1) / 2;
        T* bottom = x + (mid - r0) * ldx;
        F(backward_)(u, mid, r1, bottom, ldx, k, unit);
        F(gemm_)(x, ldx, F(offset_)(u, r0, mid), F(rows_)(bottom, ldx, (r1 - mid) * ldx), mid - r0, r1 - mid, k, (T)-1, (T)1);
        F(backward_)(u, r0, mid, x, ldx, k, unit);
        return;
    }
    for (int64_t i = r1 - 1; i >= r0; --i) {
        T* xi = x + (i - r0) * ldx;
        for (int64_t j = i + 1; j < r1; ++j) {
            T a = -F(at_)(u, i, j);
            const T* xj = x + (j - r0) * ldx;
            for (int64_t t = 0; t < k; ++t) xi[t] = F(mla_)(xj[t], a, xi[t]);
        }
        if (!unit) F(divide_)(xi, k, 1, F(at_)(u, i, i));
    }
}

/* `single`, factor.rs:725-748: one contiguous right-hand side. */
static void F(single_)(T* x, F(view_) a, int64_t m, int lower, int unit) {
    const T* d = a.data; int64_t ld = a.ld;
#define SOLVED(v, i) (unit ? (v) : (v) / d[(i) * ld + (i)])
    if (a.trans && lower) {
        for (int64_t j = 0; j < m; ++j) {
            x[j] = SOLVED(x[j], j);
            T a_ = -x[j];
            for (int64_t i = j + 1; i < m; ++i) x[i] = F(mla_)(d[j * ld + i], a_, x[i]);
        }
    } else if (a.trans) {
        for (int64_t j = m - 1; j >= 0; --j) {
            x[j] = SOLVED(x[j], j);
            T a_ = -x[j];
            for (int64_t i = 0; i < j; ++i) x[i] = F(mla_)(d[j * ld + i], a_, x[i]);
        }
    } else if (lower) {
        for (int64_t i = 0; i < m; ++i) x[i] = SOLVED(x[i] - F(dot_)(&d[i * ld], x, i), i);
    } else {
        for (int64_t i = m - 1; i >= 0; --i) x[i] = SOLVED(x[i] - F(dot_)(&d[i * ld + i + 1], &x[i + 1], m - i - 1), i);
    }
#undef SOLVED
}

/* The left-side solve of `triangular_native`, factor.rs:711-721, once the effective triangle (`lower`, the stored
 * triangle's side exchanged by a transposing view) and the diagonal (`unit`) are known: `single` for one contiguous
 * column, else `forward` (leaves of LEAF = 16 rows where x is at least as wide as the matrix is tall, else NS = 32) or
 * `backward`. */
static void F(solve_left_)(T* x, int64_t ldx, int64_t xlen, F(view_) a, int64_t m, int64_t n, int lower, int unit) {
    if (m == 0 || n == 0) return;
    if (n == 1 && ldx == 1) { F(single_)(x, a, m, lower, unit); return; }
    if (lower) F(forward_)(a, 0, m, x, xlen, ldx, n, n >= m ? 16 : 32, 0, unit);
    else F(backward_)(a, 0, m, x, ldx, n, unit);
}

/* `triangular_native`, factor.rs:711-721, for a left-side solve: `stored_lower` is Triangle::Lower,
 * otherwise Triangle::UnitUpper, of the stored matrix before the view's transpose. */
static void F(triangular_)(T* x, int64_t ldx, int64_t xlen, F(view_) a, int64_t m, int64_t n, int stored_lower) {
    F(solve_left_)(x, ldx, xlen, a, m, n, stored_lower != a.trans, !stored_lower);
}

/* `Tiled::triangular` (core/src/linalg.rs:781-787) as `triangular_native` computes it, under the CBLAS codes of its
 * vendor seam (core/src/linalg/accelerate.rs:67-70): side 141 / 142, uplo 121 / 122 (the stored triangle), trans
 * 111 / 112 / 113 (113 is 112 for real items), diag 131 / 132. The reference's Triangle::Lower is (122, 131) and
 * Triangle::UnitUpper (121, 132); the other two pairs take the same orders from (lower, unit). A right-side solve
 * transposes x into a compact copy, solves on the left with the transposed view and transposes back (:713-717). x is
 * row-major m×n with ld ldx, a row-major with ld lda; 0, or -i for invalid argument i. */
int F(or_trsm_)(int side, int uplo, int trans, int diag, int64_t m, int64_t n, const T* a, int64_t lda, T* x, int64_t ldx) {
    if (side != 141 && side != 142) return -1;
    if (uplo != 121 && uplo != 122) return -2;
    if (trans != 111 && trans != 112 && trans != 113) return -3;
    if (diag != 131 && diag != 132) return -4;
    if (m < 0) return -5;
    if (n < 0) return -6;
    int64_t order = side == 141 ? m : n;
    if (m > 0 && n > 0 && lda < order) return -8;
    if (m > 0 && n > 0 && ldx < n) return -10;
    if (m == 0 || n == 0) return 0;
    int stored_lower = uplo == 122, unit = diag == 132, ta = trans != 111;
    F(view_) v = {a, lda, ta, (order - 1) * lda + order};
    if (side == 141) {
        F(solve_left_)(x, ldx, (m - 1) * ldx + n, v, m, n, stored_lower != ta, unit);
        return 0;
    }
    T* xt = (T*)malloc(sizeof(T) * (size_t)(m * n));
    for (int64_t i = 0; i < m; ++i) for (int64_t j = 0; j < n; ++j) xt[j * m + i] = x[i * ldx + j];
    v.trans = !ta;
    F(solve_left_)(xt, m, n * m, v, n, m, stored_lower != !ta, unit);
    for (int64_t i = 0; i < m; ++i) for (int64_t j = 0; j < n; ++j) x[i * ldx + j] = xt[j * m + i];
    free(xt);
    return 0;
}

static int64_t F(stride_)(int64_t n) { return n > 0 && n % 128 == 0 ? n + 8 : n; }

/* `swap_rows`, factor.rs:572-577: in each of `cols` columns `ld` apart, swap rows first+i and pivots[i]. */
static void F(swap_rows_)(T* columns, int64_t cols, int64_t ld, const int64_t* pivots, int64_t count, int64_t first) {
    for (int64_t j = 0; j < cols; ++j) {
        T* col = columns + j * ld;
        for (int64_t i = 0; i < count; ++i) { T t = col[first + i]; col[first + i] = col[pivots[i]]; col[pivots[i]] = t; }
    }
}

/* `lu_panel`, factor.rs:580-615, LEAF = 16 (:79), on the column-by-column copy `f`. */
static void F(lu_panel_)(T* f, int64_t flen, int64_t n, int64_t ld, int64_t c0, int64_t c1, int64_t* pivots) {
    if (c1 - c0 > 16) {
        int64_t mid = (c0 + c1) / 2;
        F(lu_panel_)(f, flen, n, ld, c0, mid, pivots);
        int64_t k = mid - c0, w = c1 - mid, lw = F(stride_)(w);
        T* half = f + mid * ld;
        F(swap_rows_)(half, w, ld, pivots + c0, k, c0);
        /* transpose_by_into(&half[c0..], [ld, lw], [w, k]) : scratch row p holds U12 row c0 + p. */
        T* scratch = (T*)malloc(sizeof(T) * (size_t)(k * lw));
        for (int64_t p = 0; p < k; ++p) {
            for (int64_t r = 0; r < w; ++r) scratch[p * lw + r] = half[r * ld + c0 + p];
            for (int64_t r = w; r < lw; ++r) scratch[p * lw + r] = 0;
        }
        F(view_) l11 = F(transposed_)(F(rows_)(f + c0 * ld + c0, ld, mid * ld - (c0 * ld + c0)));
        F(triangular_)(scratch, lw, k * lw, l11, k, w, 0);
        for (int64_t r = 0; r < w; ++r) for (int64_t p = 0; p < k; ++p) half[r * ld + c0 + p] = scratch[p * lw + r];
        F(view_) u12t = F(transposed_)(F(rows_)(scratch, lw, k * lw));
        F(view_) l21t = F(rows_)(f + c0 * ld + mid, ld, mid * ld - (c0 * ld + mid));
        F(gemm_)(half + mid, ld, u12t, l21t, w, k, n - mid, (T)-1, (T)1);
        free(scratch);
        F(lu_panel_)(f, flen, n, ld, mid, c1, pivots);
        F(swap_rows_)(f + c0 * ld, mid - c0, ld, pivots + mid, c1 - mid, mid);
        return;
    }
    for (int64_t c = c0; c < c1; ++c) {
        int64_t p = c + F(largest_)(f + c * ld + c, n - c);
        pivots[c] = p;
        if (p != c) for (int64_t j = c0; j < c1; ++j) { T t = f[j * ld + c]; f[j * ld + c] = f[j * ld + p]; f[j * ld + p] = t; }
        T pivot = f[c * ld + c];
        if (pivot == (T)0) continue;
        T* l = f + c * ld + c + 1;
        F(divide_)(l, n - c - 1, 1, pivot);
        for (int64_t j = c + 1; j < c1; ++j) {
            T* col = f + j * ld;
            T a = -col[c];
            for (int64_t i = c + 1; i < n; ++i) col[i] = F(mla_)(l[i - c - 1], a, col[i]);
        }
    }
}

/* `lu_factor`, factor.rs:617-630: factors (column-major, ld = n) and perm from the row-major `a`. */
int F(or_lu_)(const T* a, int64_t n, T* factors, int64_t* perm, int64_t* pivots_out) {
    if (n < 0) return -2;
    int64_t ld = F(stride_)(n);
    T* f = (T*)malloc(sizeof(T) * (size_t)(n * ld > 0 ? n * ld : 1));
    int64_t* pivots = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    for (int64_t j = 0; j < n; ++j) { for (int64_t i = 0; i < n; ++i) f[j * ld + i] = a[i * n + j]; for (int64_t i = n; i < ld; ++i) f[j * ld + i] = 0; }
    F(lu_panel_)(f, n * ld, n, ld, 0, n, pivots);
    for (int64_t j = 0; j < n; ++j) for (int64_t i = 0; i < n; ++i) factors[j * n + i] = f[j * ld + i];
    for (int64_t i = 0; i < n; ++i) perm[i] = i;
    for (int64_t c = 0; c < n; ++c) { int64_t t = perm[c]; perm[c] = perm[pivots[c]]; perm[pivots[c]] = t; }
    if (pivots_out) memcpy(pivots_out, pivots, sizeof(int64_t) * (size_t)n);
    free(f); free(pivots);
    return 0;
}

/* The ends of the reference's leaves (lu_panel's halving down to LEAF = 16 columns): end[c] is the
 * first column past the leaf holding column c. */
static void F(leaf_ends_)(int64_t c0, int64_t c1, int64_t* end) {
    if (c1 - c0 > 16) { int64_t mid = (c0 + c1) / 2; F(leaf_ends_)(c0, mid, end); F(leaf_ends_)(mid, c1, end); return; }
    for (int64_t c = c0; c < c1; ++c) end[c] = c1;
}

/* The factorization that `lu_panel`'s recursion computes, stated flat: right-looking, one column at a
 * time, every item taking its updates in order of the pivot column, one fused multiply-add each. A
 * zero pivot leaves its column undivided and skips the updates of the rest of its own leaf only, as the
 * leaf loop's `continue` does (factor.rs:606); columns past its leaf take its raw column through the
 * recursion's triangular solve and product. Same arguments as or_lu_. */
int F(or_lu_flat_)(const T* a, int64_t n, T* factors, int64_t* perm, int64_t* pivots_out) {
    T* f = (T*)malloc(sizeof(T) * (size_t)(n * n > 0 ? n * n : 1));
    int64_t* pivots = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    int64_t* end = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    for (int64_t j = 0; j < n; ++j) for (int64_t i = 0; i < n; ++i) f[j * n + i] = a[i * n + j];
    F(leaf_ends_)(0, n, end);
    for (int64_t c = 0; c < n; ++c) {
        int64_t p = c + F(largest_)(f + c * n + c, n - c);
        pivots[c] = p;
        if (p != c) for (int64_t j = 0; j < n; ++j) { T t = f[j * n + c]; f[j * n + c] = f[j * n + p]; f[j * n + p] = t; }
        T pivot = f[c * n + c];
        int zero = pivot == (T)0;
        if (!zero) F(divide_)(f + c * n + c + 1, n - c - 1, 1, pivot);
        for (int64_t j = c + 1; j < n; ++j) {
            if (zero && j < end[c]) continue;
            T* col = f + j * n;
            T u = -col[c];
            for (int64_t i = c + 1; i < n; ++i) col[i] = F(mla_)(f[c * n + i], u, col[i]);
        }
    }
    memcpy(factors, f, sizeof(T) * (size_t)(n * n));
    for (int64_t i = 0; i < n; ++i) perm[i] = i;
    for (int64_t c = 0; c < n; ++c) { int64_t t = perm[c]; perm[c] = perm[pivots[c]]; perm[pivots[c]] = t; }
    if (pivots_out) memcpy(pivots_out, pivots, sizeof(int64_t) * (size_t)n);
    free(f); free(pivots); free(end);
    return 0;
}

/* `lu_solve`, factor.rs:633-641: x = U⁻¹ L⁻¹ (rows perm of b), b and x row-major n×k. */
int F(or_lu_solve_)(const T* factors, const int64_t* perm, int64_t n, const T* b, int64_t k, T* x) {
    for (int64_t i = 0; i < n; ++i) { if (perm[i] < 0 || perm[i] >= n) return -2; memcpy(x + i * k, b + perm[i] * k, sizeof(T) * (size_t)k); }
    F(view_) p = F(transposed_)(F(rows_)(factors, n, n * n));
    F(triangular_)(x, k, n * k, p, n, k, 0);
    F(triangular_)(x, k, n * k, p, n, k, 1);
    return 0;
}

/* The transposed solve Aᵀ x = b with rustymath's LU factors (no counterpart in the reference, which has no
 * gradients; it is the adjoint's solve): Aᵀ = Uᵀ Lᵀ P, so Uᵀ y = b in increasing order of the terms, each row then
 * scaled by its diagonal as `divide` does (factor.rs:506); Lᵀ z = y in decreasing order; x[perm[i]] = z[i]. The
 * terms are fused multiply-adds x_q · (−entry) + acc, as the substitutions above take them. */
int F(or_lu_solve_t_)(const T* factors, const int64_t* perm, int64_t n, const T* b, int64_t k, T* x) {
    for (int64_t i = 0; i < n; ++i) if (perm[i] < 0 || perm[i] >= n) return -2;
    T* w = (T*)malloc(sizeof(T) * (size_t)(n * k > 0 ? n * k : 1));
    memcpy(w, b, sizeof(T) * (size_t)(n * k));
    for (int64_t c = 0; c < k; ++c) {
        for (int64_t i = 0; i < n; ++i) {  /* Uᵀ: entry (i, q) = U(q, i) = factors[i n + q] */
            T acc = w[i * k + c];
            for (int64_t q = 0; q < i; ++q) acc = FMA(w[q * k + c], -factors[i * n + q], acc);
            F(divide_)(&acc, 1, 1, factors[i * n + i]);
            w[i * k + c] = acc;
        }
        for (int64_t i = n - 1; i >= 0; --i) {  /* Lᵀ: entry (i, q) = L(q, i) = factors[i n + q], unit diagonal */
            T acc = w[i * k + c];
            for (int64_t q = n - 1; q > i; --q) acc = FMA(w[q * k + c], -factors[i * n + q], acc);
            w[i * k + c] = acc;
        }
    }
    for (int64_t i = 0; i < n; ++i) memcpy(x + perm[i] * k, w + i * k, sizeof(T) * (size_t)k);
    free(w);
    return 0;
}

/* `lu_inverse`, factor.rs:645-664, rustymath's own kernels: out is row-major n×n. */
int F(or_lu_inverse_)(const T* factors, const int64_t* perm, int64_t n, T* out) {
    int64_t ld = F(stride_)(n);
    T* x = (T*)calloc((size_t)(n * ld > 0 ? n * ld : 1), sizeof(T));
    for (int64_t i = 0; i < n; ++i) x[i * ld + i] = 1;
    F(view_) p = F(transposed_)(F(rows_)(factors, n, n * n));
    if (n > 0) F(forward_)(p, 0, n, x, n * ld, ld, n, 32, 1, 1);
    F(triangular_)(x, ld, n * ld, p, n, n, 1);
    for (int64_t r = 0; r < n; ++r) for (int64_t c = 0; c < n; ++c) out[r * n + perm[c]] = x[r * ld + c];
    free(x);
    return 0;
}

/* `cholesky_block`, factor.rs:29: columns per block of `cholesky_factor`. */
static int64_t F(cholesky_block_)(int64_t n) { int64_t w = n / 8; return w < 32 ? 32 : (w > 128 ? 128 : w); }

/* `cholesky_panel`, factor.rs:751-775: columns c0..c1 of the `rows`-row block that `panel` holds column by column,
 * `ld` apart, halving the columns down to LEAF = 16 as `lu_panel` does. 0 where a diagonal item is not > 0. */
static int F(cholesky_panel_)(T* panel, int64_t rows, int64_t ld, int64_t c0, int64_t c1) {
    if (c1 - c0 > 16) {
        int64_t mid = (c0 + c1) / 2;
        if (!F(cholesky_panel_)(panel, rows, ld, c0, mid)) return 0;
        /* l = Matrix::rows(&left[c0 * ld + mid..], ld); right[mid..] -= l.adjoint() l, [c1 - mid, mid - c0, rows - mid] */
        F(view_) l = F(rows_)(panel + c0 * ld + mid, ld, mid * ld - (c0 * ld + mid));
        F(gemm_)(panel + mid * ld + mid, ld, F(transposed_)(l), l, c1 - mid, mid - c0, rows - mid, (T)-1, (T)1);
        return F(cholesky_panel_)(panel, rows, ld, mid, c1);
    }
    for (int64_t c = c0; c < c1; ++c) {
        T d = panel[c * ld + c];
        if (!(d > (T)0)) return 0;
        d = SQRT(d);
        panel[c * ld + c] = d;
        T* lc = panel + c * ld + c + 1;
        F(divide_)(lc, rows - c - 1, 1, d);
        /* each later column j of the leaf loses L[j][c] times column c, from row j down (`axpy`, :452) */
        for (int64_t j = 0; j < c1 - c - 1; ++j) {
            T* column = panel + (c + 1 + j) * ld;
            T a = -lc[j];
            for (int64_t r = c + 1 + j; r < rows; ++r) column[r] = F(mla_)(lc[r - c - 1], a, column[r]);
        }
    }
    return 1;
}

/* `cholesky_factor`, factor.rs:777-797: L (row-major n×n, zeros above the diagonal) with L Lᵀ = a, from a's lower
 * triangle, by blocks of `cholesky_block(n)` columns: each block's columns factor in a transposed copy (columns
 * `stride(rows)` apart), then one product (`gram`, the lower part of -x xᵀ, linalg.rs:777-779, which never takes
 * `by_vector`) updates the rest. Gives 1, or 0 (l untouched) where a is not positive definite. */
int F(or_cholesky_)(const T* a, int64_t n, T* out) {
    if (n < 0) return -2;
    int64_t ld = F(stride_)(n);
    T* l = (T*)calloc((size_t)(n * ld > 0 ? n * ld : 1), sizeof(T));
    for (int64_t i = 0; i < n; ++i) memcpy(l + i * ld, a + i * n, sizeof(T) * (size_t)n);
    int64_t width = F(cholesky_block_)(n);
    int ok = 1;
    for (int64_t k0 = 0; k0 < n && ok; k0 += width) {
        int64_t k1 = k0 + width < n ? k0 + width : n;
        int64_t nb = k1 - k0, rows = n - k0, w = n - k1, pl = F(stride_)(rows);
        T* panel = (T*)calloc((size_t)(nb * pl), sizeof(T));
        for (int64_t c = 0; c < nb; ++c) for (int64_t r = 0; r < rows; ++r) panel[c * pl + r] = l[(k0 + r) * ld + k0 + c];
        ok = F(cholesky_panel_)(panel, rows, pl, 0, nb);
        if (ok) {
            for (int64_t r = 0; r < rows; ++r) for (int64_t c = 0; c < nb; ++c) l[(k0 + r) * ld + k0 + c] = panel[c * pl + r];
            if (w > 0) {
                F(view_) x = F(transposed_)(F(rows_)(panel + nb, pl, nb * pl - nb));
                F(product_)(l + k1 * ld + k1, ld, x, (T)-1, F(transposed_)(x), w, nb, w, 0);
            }
        }
        free(panel);
    }
    if (ok) for (int64_t i = 0; i < n; ++i) for (int64_t j = 0; j < n; ++j) out[i * n + j] = j <= i ? l[i * ld + j] : (T)0;
    free(l);
    return ok;
}

/* The factorization that `cholesky_factor`'s blocks and halving compute, stated flat and left-looking: item (i, j) of
 * L, i >= j, is a[i][j] less L[i][p] L[j][p] for p = 0, 1, ..., j - 1 in that order, one fused multiply-add each,
 * then the square root (i = j, which must be > 0) or `divide` by L[j][j]. Same arguments as or_cholesky_. */
int F(or_cholesky_flat_)(const T* a, int64_t n, T* out) {
    if (n < 0) return -2;
    T* l = (T*)calloc((size_t)(n * n > 0 ? n * n : 1), sizeof(T));
    for (int64_t j = 0; j < n; ++j) {
        for (int64_t i = j; i < n; ++i) {
            T acc = a[i * n + j];
            for (int64_t p = 0; p < j; ++p) acc = F(mla_)(l[i * n + p], -l[j * n + p], acc);
            if (i == j) {
                if (!(acc > (T)0)) { free(l); return 0; }
                l[j * n + j] = SQRT(acc);
            } else {
                l[i * n + j] = acc;
                F(divide_)(&l[i * n + j], 1, 1, l[j * n + j]);
            }
        }
    }
    memcpy(out, l, sizeof(T) * (size_t)(n * n));
    free(l);
    return 1;
}

/* `odd`, factor.rs:186-194. */
static int F(odd_)(const int64_t* perm, int64_t n) {
    int64_t cycles = 0;
    for (int64_t s = 0; s < n; ++s) { int64_t i = perm[s]; while (i > s) i = perm[i]; if (i == s) ++cycles; }
    return (int)((n - cycles) % 2);
}

/* `Lu::determinant`, factor.rs:174. */
T F(or_det_)(const T* factors, const int64_t* perm, int64_t n) {
    T p = 1;
    for (int64_t i = 0; i < n; ++i) p = p * factors[i * n + i];
    return F(odd_)(perm, n) ? -p : p;
}

/* `Lu::log_determinant`, factor.rs:176-182. */
void F(or_slogdet_)(const T* factors, const int64_t* perm, int64_t n, T* sign, T* logabs) {
    T s = F(odd_)(perm, n) ? (T)-1 : (T)1, l = 0;
    for (int64_t i = 0; i < n; ++i) {
        T d = factors[i * n + i], m = ABS(d);
        if (m == (T)0) { s = 0; l = -(T)INFINITY; }
        else { s = s * (d * ((T)1 / m)); l = l + LOG(m); }
    }
    *sign = s; *logabs = l;
}

/* `LuWork::solve` with `regular`, factor.rs:207-226, for a square `a`: returns 1 and x where no pivot
 * is negligible; 0 (x untouched) where the reference falls back to least squares. */
int F(or_solve_)(const T* a, const T* b, int64_t n, int64_t k, T* x) {
    if (n == 0) return 0;
    T largest = 0;
    for (int64_t i = 0; i < n * n; ++i) { T v = ABS(a[i]); largest = v > largest ? v : largest; }
    T negligible = EPS * (T)n * largest;
    T* f = (T*)malloc(sizeof(T) * (size_t)(n * n));
    int64_t* perm = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    F(or_lu_)(a, n, f, perm, NULL);
    int regular = 1;
    for (int64_t i = 0; i < n; ++i) if (ABS(f[i * n + i]) <= negligible) { regular = 0; break; }
    if (regular) F(or_lu_solve_)(f, perm, n, b, k, x);
    free(f); free(perm);
    return regular;
}

#undef F
#undef CAT
#undef CAT_
