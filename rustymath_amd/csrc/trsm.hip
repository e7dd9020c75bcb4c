// Triangular solves with one triangle of a square matrix: rustymath's `Tiled::triangular` (core/src/linalg.rs:781-787),
// whose native form is `triangular_native` (core/src/linalg/factor.rs:711-721) over `forward` (:668-689), `backward`
// (:693-707) and `single` (:725-748).
//
// The reference's orders of operations, which fix the last bits of x, are kept by keeping its structure: a lower solve
// halves its rows above `block` (LEAF = 16 where x has at least as many columns as the matrix has rows, else NS = 32),
// solves the top half, removes it from the bottom half with one product, then solves the bottom half; an upper solve
// halves above NS, bottom half first. The products go through this library's `gemm`, which routes as the reference's
// does (`by_vector` where x is one strided column, the in-order MFMA product otherwise), and a leaf substitutes row by
// row with `axpy`'s fused multiply-adds in column order, then `divide` (a reciprocal multiply, or a quotient for a
// subnormal diagonal). One contiguous column takes `single`: a triangle read by columns (a transposed view) adds each
// solved item to the rest with fused multiply-adds and divides by quotients (trsv1_columns in solve.hip); a triangle
// read by rows takes each item's `dot` (eight running sums over whole chunks of eight, folded in order, then the rest)
// and a quotient (single_rows_kernel here). A right-side solve transposes x and solves on the left with the transposed
// view, as the reference does.
#include "hs_kernels.h"
#include <algorithm>

namespace hs {

namespace {

constexpr int LEAF = 16, NS = 32;

template <typename T>
__device__ __forceinline__ T smallest_normal_t() {
    return sizeof(T) == 8 ? T(2.2250738585072014e-308) : T(1.17549435e-38f);
}

// Entry (i, j) of the matrix that the view shows: a[j lda + i] when it transposes the stored matrix, else a[i lda + j].
template <typename T>
__device__ __forceinline__ T entry(const T* a, int64_t lda, bool ta, int64_t i, int64_t j) {
    return ta ? a[j * lda + i] : a[i * lda + j];
}

// A leaf of at most 32 rows [r0, r0 + h): one thread per column of x, the leaf's items of that column in registers.
// LOWER: rows in increasing order, each taking the rows above it in the leaf in increasing order (`forward`'s leaf);
// otherwise rows from the bottom up, each taking the rows below it in increasing order (`backward`'s leaf). Each term
// is one fused multiply-add, x_i += x_j (-a_ij); then, unless unit, `divide` by the diagonal item.
template <typename T, bool LOWER>
__global__ __launch_bounds__(256) void trsm_leaf_kernel(int64_t r0, int h, int64_t k, const T* __restrict__ a, int64_t lda,
                                                        int ta, int unit, T* __restrict__ x, int64_t ldx) {
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t cc = c < k ? c : k - 1;
    T v[NS];
#pragma unroll
    for (int i = 0; i < NS; ++i) v[i] = x[(r0 + (i < h ? i : h - 1)) * ldx + cc];
    const T tiny = smallest_normal_t<T>();
    auto solve_row = [&](int i) {
        if (!unit) {
            const T d = entry(a, lda, ta != 0, r0 + i, r0 + i);
            v[i] = fabs(d) >= tiny ? v[i] * (T(1) / d) : v[i] / d;
        }
    };
    if (LOWER) {
#pragma unroll
        for (int i = 0; i < NS; ++i) {
            if (i < h) {
#pragma unroll
                for (int j = 0; j < i; ++j)
                    v[i] = __builtin_elementwise_fma(v[j], -entry(a, lda, ta != 0, r0 + i, r0 + j), v[i]);
                solve_row(i);
            }
        }
    } else {
#pragma unroll
        for (int i = NS - 1; i >= 0; --i) {
            if (i < h) {
#pragma unroll
                for (int j = i + 1; j < NS; ++j)
                    if (j < h) v[i] = __builtin_elementwise_fma(v[j], -entry(a, lda, ta != 0, r0 + i, r0 + j), v[i]);
                solve_row(i);
            }
        }
    }
    if (c < k) {
#pragma unroll
        for (int i = 0; i < NS; ++i)
            if (i < h) x[(r0 + i) * ldx + c] = v[i];
    }
}

// `single` for a triangle read by rows (a view that does not transpose): x_i = (x_i - dot(row, solved items)) / a_ii in
// solve order (increasing for LOWER, decreasing otherwise), the quotient only where not unit. `dot` is
// `sum_of_products` (factor.rs:460-473): eight running sums over the whole chunks of eight from the slice's start,
// lanes 0-7 one sum each, folded in order from zero, then the remaining terms in order. x is kept in LDS (m items).
template <typename T, bool LOWER>
__global__ __launch_bounds__(64) void single_rows_kernel(int64_t m, const T* __restrict__ a, int64_t lda, int unit,
                                                         T* __restrict__ x) {
    extern __shared__ __attribute__((aligned(16))) unsigned char dyn[];
    T* xs = reinterpret_cast<T*>(dyn);
    const int lane = threadIdx.x;
    for (int64_t i = lane; i < m; i += 64) xs[i] = x[i];
    __syncthreads();
    for (int64_t step = 0; step < m; ++step) {
        const int64_t i = LOWER ? step : m - 1 - step;
        const int64_t start = LOWER ? 0 : i + 1, len = LOWER ? i : m - 1 - i;
        const T* row = a + i * lda + start;
        const T* xv = xs + start;
        const int64_t whole = len / 8 * 8;
        T part = T(0);
        if (lane < 8) {
            for (int64_t c = lane; c < whole; c += 8) part = __builtin_elementwise_fma(row[c], xv[c], part);
        }
        T t = T(0);
#pragma unroll
        for (int l = 0; l < 8; ++l) t = t + __shfl(part, l);
        for (int64_t p = whole; p < len; ++p) t = __builtin_elementwise_fma(row[p], xv[p], t);
        T v = xs[i] - t;
        if (!unit) v = v / a[i * lda + i];
        __syncthreads();  // every lane has read what it needs of xs
        if (lane == 0) xs[i] = v;
        __syncthreads();
    }
    for (int64_t i = lane; i < m; i += 64) x[i] = xs[i];
}

template <typename T>
int leaf(bool lower, int64_t r0, int64_t h, int64_t k, const T* a, int64_t lda, bool ta, bool unit, T* x, int64_t ldx,
         hipStream_t s) {
    const unsigned grid = (unsigned)((k + 255) / 256);
    if (lower)
        trsm_leaf_kernel<T, true><<<grid, 256, 0, s>>>(r0, (int)h, k, a, lda, ta, unit, x, ldx);
    else
        trsm_leaf_kernel<T, false><<<grid, 256, 0, s>>>(r0, (int)h, k, a, lda, ta, unit, x, ldx);
    HS_HIP(hipGetLastError());
    return OK;
}

// `forward` (factor.rs:668-689) over rows [r0, r1) of x (absolute row indices).
template <typename T>
int forward_rec(const T* a, int64_t lda, bool ta, int64_t r0, int64_t r1, T* x, int64_t ldx, int64_t k, int64_t block,
                bool unit, hipStream_t s) {
    if (r1 - r0 <= block) return leaf<T>(true, r0, r1 - r0, k, a, lda, ta, unit, x, ldx, s);
    const int64_t mid = (r0 + r1) / 2;
    int rc = forward_rec<T>(a, lda, ta, r0, mid, x, ldx, k, block, unit, s);
    // The bottom rows take the top rows' terms: the view offset to (mid, r0), as `l.offset(mid, r0)`.
    const T* blk = ta ? a + r0 * lda + mid : a + mid * lda + r0;
    if (rc == OK)
        rc = gemm<T>(ta, false, r1 - mid, k, mid - r0, T(-1), blk, lda, x + r0 * ldx, ldx, T(1), x + mid * ldx, ldx, s);
    if (rc == OK) rc = forward_rec<T>(a, lda, ta, mid, r1, x, ldx, k, block, unit, s);
    return rc;
}

// `backward` (factor.rs:693-707) over rows [r0, r1).
template <typename T>
int backward_rec(const T* a, int64_t lda, bool ta, int64_t r0, int64_t r1, T* x, int64_t ldx, int64_t k, bool unit,
                 hipStream_t s) {
    if (r1 - r0 <= NS) return leaf<T>(false, r0, r1 - r0, k, a, lda, ta, unit, x, ldx, s);
    const int64_t mid = (r0 + r1) / 2;
    int rc = backward_rec<T>(a, lda, ta, mid, r1, x, ldx, k, unit, s);
    const T* blk = ta ? a + mid * lda + r0 : a + r0 * lda + mid;  // `u.offset(r0, mid)`
    if (rc == OK)
        rc = gemm<T>(ta, false, mid - r0, k, r1 - mid, T(-1), blk, lda, x + mid * ldx, ldx, T(1), x + r0 * ldx, ldx, s);
    if (rc == OK) rc = backward_rec<T>(a, lda, ta, r0, mid, x, ldx, k, unit, s);
    return rc;
}

// The left-side solve of `triangular_native` with the view's orientation (ta), the effective triangle (lower) and
// diagonal (unit) already decided; `single_ok`: x is one contiguous column.
template <typename T>
int trsm_left(bool lower, bool unit, bool ta, bool single_ok, int64_t m, int64_t k, const T* a, int64_t lda, T* x,
              int64_t ldx, hipStream_t s) {
    if (m == 0 || k == 0) return OK;
    if (single_ok && k == 1) {
        if (ta) return trsv1_columns<T>(lower, unit, m, a, lda, x, s);
        const size_t lds = (size_t)m * sizeof(T);
        if (lds > 160 * 1024) {
            set_error("a one-column triangular solve with a triangle read by rows holds x in LDS: order at most 20480 "
                      "(FP64) or 40960 (FP32)");
            return ERR_UNSUPPORTED;
        }
        if (lower) {
            HS_HIP(hipFuncSetAttribute((const void*)single_rows_kernel<T, true>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            single_rows_kernel<T, true><<<1, 64, lds, s>>>(m, a, lda, unit, x);
        } else {
            HS_HIP(hipFuncSetAttribute((const void*)single_rows_kernel<T, false>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            single_rows_kernel<T, false><<<1, 64, lds, s>>>(m, a, lda, unit, x);
        }
        HS_HIP(hipGetLastError());
        return OK;
    }
    if (lower) return forward_rec<T>(a, lda, ta, 0, m, x, ldx, k, k >= m ? LEAF : NS, unit, s);
    return backward_rec<T>(a, lda, ta, 0, m, x, ldx, k, unit, s);
}

}  // namespace

template <typename T>
int trsm(bool right, bool stored_lower, bool ta, bool unit, int64_t m, int64_t n, const T* a, int64_t lda, T* x,
         int64_t ldx, bool contiguous, hipStream_t s) {
    if (m == 0 || n == 0) return OK;
    int rc = OK;
    if (!right) {
        rc = trsm_left<T>(stored_lower != ta, unit, ta, contiguous && n == 1, m, n, a, lda, x, ldx, s);
    } else {
        // x a = b as aᵀ xᵀ = bᵀ: x transposed into a compact n×m copy, solved on the left with the transposed view.
        T* xt = (T*)pool_get(sizeof(T) * (size_t)(m * n));
        if (!xt) return hip_status(hipErrorOutOfMemory, "pool_get");
        rc = transpose<T>(m, n, x, ldx, xt, m, s);
        if (rc == OK) rc = trsm_left<T>(stored_lower != !ta, unit, !ta, m == 1, n, m, a, lda, xt, m, s);
        if (rc == OK) rc = transpose<T>(n, m, xt, m, x, ldx, s);
        hipStreamSynchronize(s);
        pool_put(xt);
    }
    if (rc == OK) HS_HIP(hipStreamSynchronize(s));
    return rc;
}

template int trsm<double>(bool, bool, bool, bool, int64_t, int64_t, const double*, int64_t, double*, int64_t, bool,
                          hipStream_t);
template int trsm<float>(bool, bool, bool, bool, int64_t, int64_t, const float*, int64_t, float*, int64_t, bool,
                         hipStream_t);

}  // namespace hs
