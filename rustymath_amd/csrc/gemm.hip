// Matrix products on gfx950: an MFMA GEMM with LDS-staged, double-buffered tiles (FP64 and FP32), the
// `by_vector` products of a matrix and a vector, and the scaling of `gemm`'s degenerate cases.
//
// Reference: core/src/linalg.rs gemm :799-807, scale :1000-1008, by_vector :1020-1073, product
// :1353-1426. rustymath's product gives each result as one fused multiply-add per step of p, in order,
// from 0 (beta = 0) or from beta c, with alpha folded into a's packed copy. This kernel keeps one
// accumulator per result for the whole of k and feeds the MFMA unit in order of p, so it computes the
// same chain of roundings; a k tail pads with +0 · (-0) = -0, which adds nothing to any value.
#include "hs_kernels.h"
#include "gemm_tile.h"
#include <hip/hip_ext.h>
#include <mutex>
#include <vector>
#include <algorithm>

namespace hs {

namespace {
// The live profile of the MFMA product: while on, each launch is bracketed by HIP events on its own stream.
struct GemmProfile {
    std::mutex mu;
    bool on = false;
    struct Rec { hipEvent_t a, b; double flops, bytes; };
    std::vector<Rec> recs;
} g_prof;
}  // namespace

int gemm_profile(int on, double* ms, int64_t* launches, double* flops, double* bytes) {
    std::lock_guard<std::mutex> lk(g_prof.mu);
    double total = 0, fl = 0, by = 0;
    int64_t count = 0;
    for (auto& r : g_prof.recs) {
        float e = 0;
        HS_HIP(hipEventSynchronize(r.b));
        HS_HIP(hipEventElapsedTime(&e, r.a, r.b));
        total += e; fl += r.flops; by += r.bytes; ++count;
        hipEventDestroy(r.a); hipEventDestroy(r.b);
    }
    g_prof.recs.clear();
    if (ms) *ms = total;
    if (launches) *launches = count;
    if (flops) *flops = fl;
    if (bytes) *bytes = by;
    g_prof.on = on != 0;
    return OK;
}

namespace {

constexpr int THREADS = 256;
using tile::BM;
using tile::BN;
using tile::BK;
using tile::Pad;
using tile::gemm_tile;

// Maps a launch index to an output tile so that each XCD (dispatch is round-robin over the 8 XCDs)
// works on one contiguous run of tiles, and runs of 8 tile-rows share B's column panels in its L2.
__device__ __forceinline__ bool tile_of(int w, int tiles_m, int tiles_n, int& tm, int& tn) {
    int total = tiles_m * tiles_n;
    int per = (total + 7) / 8;
    int id = (w % 8) * per + w / 8;
    if (id >= total) return false;
    tile::grouped(id, tiles_m, tiles_n, tm, tn);
    return true;
}

// blockIdx.y is the cell of a batch: A, B and C advance by sa, sb and sc items per cell (0 repeats an operand).
template <typename T, bool TA, bool TB>
__global__ __launch_bounds__(THREADS, 2) void gemm_mfma(int64_t m, int64_t n, int64_t k, T alpha,
                                                         const T* __restrict__ A, int64_t lda,
                                                         const T* __restrict__ B, int64_t ldb, T beta,
                                                         T* __restrict__ C, int64_t ldc, int tiles_m, int tiles_n,
                                                         int64_t sa, int64_t sb, int64_t sc) {
    constexpr int P = Pad<T>::v;
    __shared__ T As[2][BK][BM + P];
    __shared__ T Bs[2][BK][BN + P];
    int tm, tn;
    if (!tile_of(blockIdx.x, tiles_m, tiles_n, tm, tn)) return;
    A += blockIdx.y * sa;
    B += blockIdx.y * sb;
    C += blockIdx.y * sc;
    const int64_t m0 = (int64_t)tm * BM, n0 = (int64_t)tn * BN;
    if (m0 + BM <= m && n0 + BN <= n)
        gemm_tile<T, TA, TB, false>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, m0, n0, As, Bs, threadIdx.x, true);
    else
        gemm_tile<T, TA, TB, true>(m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, m0, n0, As, Bs, threadIdx.x, true);
}

// `scale` (core/src/linalg.rs:1000-1008): c = beta c, zeros for beta = 0.
template <typename T>
__global__ void scale_kernel(int64_t m, int64_t n, T beta, T* C, int64_t ldc, int64_t sc) {
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= m * n) return;
    C += blockIdx.y * sc;
    int64_t i = idx / n, j = idx % n;
    T* x = &C[i * ldc + j];
    *x = beta == T(0) ? T(0) : beta * *x;
}

// `by_vector` (core/src/linalg.rs:1020-1073): one thread per result, each summing in the order the
// reference's kernel for its case uses — `dots` (:903, quads of rows with eight running sums added
// pairwise; rows past the last quad take `dot`, factor.rs:460), `narrow_dots` (:923, two sums) for
// stored rows of at most 8 items, and for a transposed matrix `narrow_sums` (:941, four interleaved
// sets) for at most 8 results or an in-order fold (`scaled_runs`, :883) otherwise. `mlen` is the
// matrix slice's length from its start, which `narrow_sums` reads.
template <typename T>
__global__ void by_vector_kernel(int64_t rows, int64_t k, const T* __restrict__ mat, int64_t ld, int mtrans, int64_t mlen,
                                 const T* __restrict__ vec, int64_t vstride, T alpha, T beta, T* C, int64_t step,
                                 int64_t smat, int64_t svec, int64_t sc) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows) return;
    mat += blockIdx.y * smat;
    vec += blockIdx.y * svec;
    C += blockIdx.y * sc;
    T y;
    if (!mtrans) {
        const T* row = mat + i * ld;
        if (k <= 8) {
            T s0 = 0, s1 = 0;
            for (int64_t l = 0; l < k; ++l) {
                if (l % 2 == 0) s0 = __builtin_elementwise_fma(row[l], vec[l * vstride], s0);
                else s1 = __builtin_elementwise_fma(row[l], vec[l * vstride], s1);
            }
            y = s0 + s1;
        } else {
            T s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            int64_t whole = k / 8 * 8;
            for (int64_t c = 0; c < whole; c += 8)
#pragma unroll
                for (int l = 0; l < 8; ++l) s[l] = __builtin_elementwise_fma(row[c + l], vec[(c + l) * vstride], s[l]);
            T tsum;
            if (i < rows / 4 * 4) tsum = ((s[0] + s[4]) + (s[2] + s[6])) + ((s[1] + s[5]) + (s[3] + s[7]));
            else {
                tsum = T(0);
#pragma unroll
                for (int l = 0; l < 8; ++l) tsum = tsum + s[l];
            }
            for (int64_t p = whole; p < k; ++p) tsum = __builtin_elementwise_fma(row[p], vec[p * vstride], tsum);
            y = tsum;
        }
    } else if (rows <= 8) {
        int64_t whole;
        if (mlen < 8) whole = 0;
        else { whole = (mlen - 8) / (ld > 1 ? ld : 1) + 1; if (whole > k) whole = k; }
        T acc[4] = {0, 0, 0, 0};
        int64_t quads = whole / 4;
        for (int64_t q = 0; q < quads; ++q)
#pragma unroll
            for (int s = 0; s < 4; ++s) acc[s] = __builtin_elementwise_fma(mat[(4 * q + s) * ld + i], vec[(4 * q + s) * vstride], acc[s]);
        for (int64_t p = quads * 4; p < k; ++p) acc[0] = __builtin_elementwise_fma(mat[p * ld + i], vec[p * vstride], acc[0]);
        y = (acc[0] + acc[2]) + (acc[1] + acc[3]);
    } else {
        T a = 0;
        for (int64_t p = 0; p < k; ++p) a = __builtin_elementwise_fma(vec[p * vstride], mat[p * ld + i], a);
        y = a;
    }
    if (alpha != T(1)) y = alpha * y;
    T* z = &C[i * step];
    *z = beta == T(0) ? y : __builtin_elementwise_fma(beta, *z, y);
}

}  // namespace

namespace {
// One launch of `batch` (<= 65535) cells.
template <typename T>
int gemm_cells(bool ta, bool tb, int64_t m, int64_t n, int64_t k, T alpha, const T* a, int64_t lda, int64_t sa,
               const T* b, int64_t ldb, int64_t sb, T beta, T* c, int64_t ldc, int64_t sc, int batch, hipStream_t s,
               bool in_order) {
    if (m == 0 || n == 0 || batch == 0) return OK;
    if (k == 0 || alpha == T(0)) {
        if (beta == T(1)) return OK;
        int64_t total = m * n;
        scale_kernel<T><<<dim3((unsigned)((total + 255) / 256), (unsigned)batch), 256, 0, s>>>(m, n, beta, c, ldc, sc);
        HS_HIP(hipGetLastError());
        return OK;
    }
    if ((n == 1 || m == 1) && !in_order) {
        // As by_vector: for n = 1 the matrix is a and the vector b's column; for m = 1 the matrix is bᵀ.
        const T* mat; int64_t ld, rows, step, vstride, mlen, smat, svec; int mtrans; const T* vec;
        if (n == 1) {
            mat = a; ld = lda; mtrans = ta; rows = m; step = ldc;
            vec = b; vstride = tb ? 1 : ldb;
            mlen = ta ? (k - 1) * lda + m : (m - 1) * lda + k;
            smat = sa; svec = sb;
        } else {
            mat = b; ld = ldb; mtrans = !tb; rows = n; step = 1;
            vec = a; vstride = ta ? lda : 1;
            mlen = tb ? (n - 1) * ldb + k : (k - 1) * ldb + n;
            smat = sb; svec = sa;
        }
        by_vector_kernel<T><<<dim3((unsigned)((rows + 255) / 256), (unsigned)batch), 256, 0, s>>>(
            rows, k, mat, ld, mtrans, mlen, vec, vstride, alpha, beta, c, step, smat, svec, sc);
        HS_HIP(hipGetLastError());
        return OK;
    }
    int tiles_m = (int)((m + BM - 1) / BM), tiles_n = (int)((n + BN - 1) / BN);
    int total = tiles_m * tiles_n;
    unsigned grid = (unsigned)(((total + 7) / 8) * 8);
    // While profiling, the launch carries its own start and stop events (hipExtLaunchKernelGGL records them at the
    // dispatch's start and end), so the measured time is the kernel's alone.
#define HS_LAUNCH(TA_, TB_)                                                                                        \
    do {                                                                                                           \
        if (profiled)                                                                                              \
            hipExtLaunchKernelGGL(gemm_mfma<T, TA_, TB_>, dim3(grid, batch), dim3(THREADS), 0, s, rec.a, rec.b, 0,  \
                                  m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, tiles_m, tiles_n, sa, sb, sc);     \
        else                                                                                                       \
            gemm_mfma<T, TA_, TB_><<<dim3(grid, batch), THREADS, 0, s>>>(m, n, k, alpha, a, lda, b, ldb, beta, c,     \
                                                                       ldc, tiles_m, tiles_n, sa, sb, sc);         \
    } while (0)
    // Algorithmic bytes: a and b read once, c written once (and read once where beta is not 0).
    GemmProfile::Rec rec{nullptr, nullptr, 2.0 * batch * (double)m * (double)n * (double)k,
                         (double)sizeof(T) * batch * ((double)m * k + (double)k * n + (beta != T(0) ? 2.0 : 1.0) * m * n)};
    bool profiled = false;
    {
        std::lock_guard<std::mutex> lk(g_prof.mu);
        profiled = g_prof.on;
    }
    if (profiled) {
        HS_HIP(hipEventCreate(&rec.a));
        HS_HIP(hipEventCreate(&rec.b));
    }
    if (!ta && !tb) HS_LAUNCH(false, false);
    else if (!ta && tb) HS_LAUNCH(false, true);
    else if (ta && !tb) HS_LAUNCH(true, false);
    else HS_LAUNCH(true, true);
#undef HS_LAUNCH
    HS_HIP(hipGetLastError());
    if (profiled) {
        std::lock_guard<std::mutex> lk(g_prof.mu);
        g_prof.recs.push_back(rec);
    }
    return OK;
}
}  // namespace

template <typename T>
int gemm(bool ta, bool tb, int64_t m, int64_t n, int64_t k, T alpha, const T* a, int64_t lda, const T* b, int64_t ldb,
         T beta, T* c, int64_t ldc, hipStream_t s, bool in_order) {
    return gemm_cells<T>(ta, tb, m, n, k, alpha, a, lda, 0, b, ldb, 0, beta, c, ldc, 0, 1, s, in_order);
}

template <typename T>
int gemm_batched(bool ta, bool tb, int64_t m, int64_t n, int64_t k, T alpha, const T* a, int64_t lda, int64_t sa,
                 const T* b, int64_t ldb, int64_t sb, T beta, T* c, int64_t ldc, int64_t sc, int64_t batch,
                 hipStream_t s) {
    for (int64_t c0 = 0; c0 < batch; c0 += 65535) {
        const int cells = (int)std::min<int64_t>(65535, batch - c0);
        const int rc = gemm_cells<T>(ta, tb, m, n, k, alpha, a + c0 * sa, lda, sa, b + c0 * sb, ldb, sb, beta,
                                     c + c0 * sc, ldc, sc, cells, s, false);
        if (rc) return rc;
    }
    return OK;
}
template int gemm_batched<double>(bool, bool, int64_t, int64_t, int64_t, double, const double*, int64_t, int64_t,
                                  const double*, int64_t, int64_t, double, double*, int64_t, int64_t, int64_t,
                                  hipStream_t);
template int gemm_batched<float>(bool, bool, int64_t, int64_t, int64_t, float, const float*, int64_t, int64_t,
                                 const float*, int64_t, int64_t, float, float*, int64_t, int64_t, int64_t, hipStream_t);

template int gemm<double>(bool, bool, int64_t, int64_t, int64_t, double, const double*, int64_t, const double*, int64_t,
                          double, double*, int64_t, hipStream_t, bool);
template int gemm<float>(bool, bool, int64_t, int64_t, int64_t, float, const float*, int64_t, const float*, int64_t,
                         float, float*, int64_t, hipStream_t, bool);

}  // namespace hs
