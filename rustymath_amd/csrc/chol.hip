// The Cholesky factorization on gfx950: `cholesky_factor` (core/src/linalg/factor.rs:777-797) with
// `cholesky_panel` (:751-775).
//
// rustymath's blocks and halving leave every item of L with one order of operations: item (i, j), i >= j, is
// a[i][j] less L[i][p] L[j][p] for p = 0, 1, ..., j - 1 in that order, one fused multiply-add each (the leaf's
// `axpy`s, the panel's `gemm` and the blocks' `gram` all add their terms in order of p, and `gram` never takes
// `by_vector`); then a square root on the diagonal (which must be > 0) or `divide` by L[j][j] below it. So the
// blocking here is free: 64-column blocks, each a diagonal block in one wave, the rows below it one per thread, and
// one in-order MFMA product for the trailing matrix.
#include "hs_kernels.h"
#include "hs_pivot.h"
#include <algorithm>

namespace hs {

namespace {

constexpr int CB = 64;  // columns per block

template <typename T>
__device__ __forceinline__ T divided(T x, T d, T r) {  // `divide` (factor.rs:506-508)
    const T tiny = sizeof(T) == 8 ? T(2.2250738585072014e-308) : T(1.17549435e-38f);
    return fabs(d) >= tiny ? x * r : x / d;
}

// The diagonal block [k0, k0 + nb)², nb <= 64, right-looking in one wave: lane r holds row k0 + r of the block in
// registers; at column c the diagonal item's square root goes to every lane, the items below it are divided, and
// each later column j takes its term c from row j down, L[j][c] read from lane j. A diagonal item that is not > 0
// sets *fail.
template <typename T>
__global__ __launch_bounds__(64) void chol_diag_kernel(T* a, int64_t lda, int64_t k0, int nb, int* fail) {
    const int lane = threadIdx.x;
    const int64_t row = k0 + (lane < nb ? lane : nb - 1);
    T v[CB];
#pragma unroll
    for (int j = 0; j < CB; ++j) v[j] = a[row * lda + k0 + (j < nb ? j : nb - 1)];
    asm volatile("" ::: "memory");
#pragma unroll
    for (int c = 0; c < CB; ++c) {
        if (c < nb) {
            const T dc = lane_value(v[c], c);
            if (!(dc > T(0)) && lane == 0) atomicExch(fail, 1);
            const T d = sqrt(dc), r = T(1) / d;
            if (lane == c) v[c] = d;
            else if (lane > c) v[c] = divided(v[c], d, r);
#pragma unroll
            for (int j = c + 1; j < CB; ++j) {
                if (j < nb) {
                    const T ljc = lane_value(v[c], j);
                    if (lane >= j) v[j] = __builtin_elementwise_fma(v[c], -ljc, v[j]);
                }
            }
        }
    }
    if (lane < nb) {
#pragma unroll
        for (int j = 0; j < CB; ++j)
            if (j <= lane && j < nb) a[row * lda + k0 + j] = v[j];
    }
}

// The rows [k1, n) of the block's columns: thread per row, item c = (a[r][c] - Σ_{p<c} L[r][p] L[c][p]) / L[c][c]
// with the block's triangle and its reciprocals in LDS (one address for every lane).
template <typename T>
__global__ __launch_bounds__(256) void chol_panel_kernel(T* a, int64_t lda, int64_t k0, int nb, int64_t k1, int64_t n) {
    __shared__ T L[CB][CB + 1];
    __shared__ T R[CB];
    const int t = threadIdx.x;
    for (int idx = t; idx < CB * CB; idx += 256) {
        const int i = idx / CB, j = idx % CB;
        L[i][j] = (i < nb && j <= i) ? a[(k0 + i) * lda + k0 + j] : T(0);
    }
    __syncthreads();
    if (t < CB) R[t] = t < nb ? T(1) / L[t][t] : T(0);
    __syncthreads();
    const int64_t r = k1 + (int64_t)blockIdx.x * 256 + t;
    if (r >= n) return;
    T v[CB];
#pragma unroll
    for (int c = 0; c < CB; ++c) v[c] = a[r * lda + k0 + (c < nb ? c : nb - 1)];
#pragma unroll
    for (int c = 0; c < CB; ++c) {
        if (c < nb) {
            T acc = v[c];
#pragma unroll
            for (int p = 0; p < c; ++p) acc = __builtin_elementwise_fma(v[p], -L[c][p], acc);
            v[c] = divided(acc, L[c][c], R[c]);
        }
    }
#pragma unroll
    for (int c = 0; c < CB; ++c)
        if (c < nb) a[r * lda + k0 + c] = v[c];
}

template <typename T>
__global__ void zero_upper(int64_t n, T* a, int64_t lda) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * n) return;
    const int64_t i = idx / n, j = idx % n;
    if (j > i) a[i * lda + j] = T(0);
}

__global__ void clear_int(int* p) { *p = 0; }

}  // namespace

template <typename T>
int potrf(int64_t n, T* a, int64_t lda, int* ok, hipStream_t s) {
    *ok = 0;
    if (n == 0) { *ok = 1; return OK; }
    int* fail = (int*)pool_get(sizeof(int));
    if (!fail) return hip_status(hipErrorOutOfMemory, "pool_get");
    clear_int<<<1, 1, 0, s>>>(fail);
    HS_HIP(hipGetLastError());
    int rc = OK;
    for (int64_t k0 = 0; k0 < n && rc == OK; k0 += CB) {
        const int64_t k1 = std::min<int64_t>(n, k0 + CB);
        const int nb = (int)(k1 - k0);
        chol_diag_kernel<T><<<1, 64, 0, s>>>(a, lda, k0, nb, fail);
        HS_HIP(hipGetLastError());
        if (k1 < n) {
            chol_panel_kernel<T><<<(unsigned)((n - k1 + 255) / 256), 256, 0, s>>>(a, lda, k0, nb, k1, n);
            HS_HIP(hipGetLastError());
            // The trailing matrix less L21 L21ᵀ, each item's terms in order (the upper items are computed too and
            // cleared at the end).
            rc = gemm<T>(false, true, n - k1, n - k1, nb, T(-1), a + k1 * lda + k0, lda, a + k1 * lda + k0, lda, T(1),
                         a + k1 * lda + k1, lda, s, true);
        }
    }
    if (rc == OK) {
        zero_upper<T><<<(unsigned)((n * n + 255) / 256), 256, 0, s>>>(n, a, lda);
        HS_HIP(hipGetLastError());
        int failed = 1;
        HS_HIP(hipMemcpyAsync(&failed, fail, sizeof(int), hipMemcpyDeviceToHost, s));
        HS_HIP(hipStreamSynchronize(s));
        *ok = failed ? 0 : 1;
    }
    pool_put(fail);
    return rc;
}

template int potrf<double>(int64_t, double*, int64_t, int*, hipStream_t);
template int potrf<float>(int64_t, float*, int64_t, int*, hipStream_t);

}  // namespace hs
