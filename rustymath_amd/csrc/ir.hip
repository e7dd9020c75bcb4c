// Mixed-precision solves on gfx950 (no counterpart in rustymath, which has no iterative refinement): the
// matrix factors in FP32 on MFMA (or with BF16 trailing products accumulating in FP32), and FP64 iterative
// refinement on the GPU recovers FP64 accuracy: r = b - A x (FP64, one streaming pass over A), the
// correction solves with the low-precision factors, x += d in FP64, and the loop stops once the normwise
// backward error ||b - A x|| / (||A|| ||x||) (Frobenius / 2-norms, FP64) reaches `tol`.
#include "hs_kernels.h"
#include <cmath>
#include <vector>

namespace hs {

namespace {

// r = b - A x for one right-hand side, one wave per row: lanes stride along the row (coalesced 512-byte
// reads), four running sums each, added across the wave at the end. HBM-bound: one read of A.
__global__ __launch_bounds__(256) void residual_kernel(int64_t n, const double* __restrict__ a, int64_t lda,
                                                       const double* __restrict__ x, const double* __restrict__ b,
                                                       double* __restrict__ r) {
    const int lane = threadIdx.x & 63;
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= n) return;
    const double* ar = a + row * lda;
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    int64_t j = lane;
    for (; j + 192 < n; j += 256) {
        s0 = __builtin_elementwise_fma(ar[j], x[j], s0);
        s1 = __builtin_elementwise_fma(ar[j + 64], x[j + 64], s1);
        s2 = __builtin_elementwise_fma(ar[j + 128], x[j + 128], s2);
        s3 = __builtin_elementwise_fma(ar[j + 192], x[j + 192], s3);
    }
    for (; j < n; j += 64) s0 = __builtin_elementwise_fma(ar[j], x[j], s0);
    double s = (s0 + s1) + (s2 + s3);
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (lane == 0) r[row] = b[row] - s;
}

// Per-block sums of squares of count items (grid-stride), into part[blockIdx.x]; a second launch of one block
// adds the parts. Deterministic: fixed grid, fixed order.
__global__ __launch_bounds__(256) void sumsq_kernel(int64_t count, const double* __restrict__ v, double* part) {
    __shared__ double s[4];
    double acc = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        acc = __builtin_elementwise_fma(v[i], v[i], acc);
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (s[0] + s[1]) + (s[2] + s[3]);
}
__global__ void sum_parts(int count, const double* part, double* out) {
    __shared__ double s[256];
    double acc = 0;
    for (int i = threadIdx.x; i < count; i += 256) acc += part[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s[0];
}

// r32 = r * scale (scale a power of two chosen on the host from ||r||, so the cast loses nothing but rounding).
__global__ void scale_to_f32(int64_t n, const double* __restrict__ r, double scale, float* __restrict__ r32) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) r32[i] = (float)(r[i] * scale);
}
// x += d32 / scale, in FP64.
__global__ void add_correction(int64_t n, const float* __restrict__ d32, double inv_scale, double* __restrict__ x) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = __builtin_elementwise_fma((double)d32[i], inv_scale, x[i]);
}

// y = A v, one wave per row (as residual_kernel).
__global__ __launch_bounds__(256) void matvec_kernel(int64_t n, const double* __restrict__ a, int64_t lda,
                                                     const double* __restrict__ v, double* __restrict__ y) {
    const int lane = threadIdx.x & 63;
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= n) return;
    const double* ar = a + row * lda;
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    int64_t j = lane;
    for (; j + 192 < n; j += 256) {
        s0 = __builtin_elementwise_fma(ar[j], v[j], s0);
        s1 = __builtin_elementwise_fma(ar[j + 64], v[j + 64], s1);
        s2 = __builtin_elementwise_fma(ar[j + 128], v[j + 128], s2);
        s3 = __builtin_elementwise_fma(ar[j + 192], v[j + 192], s3);
    }
    for (; j < n; j += 64) s0 = __builtin_elementwise_fma(ar[j], v[j], s0);
    double s = (s0 + s1) + (s2 + s3);
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (lane == 0) y[row] = s;
}

// out[i] = V_i · w for the first gridDim.x rows of V (n items each): one block per product, a fixed reduction order.
__global__ __launch_bounds__(256) void dots_kernel(int64_t n, const double* __restrict__ V, const double* __restrict__ w,
                                                   double* __restrict__ out) {
    __shared__ double part[4];
    const double* v = V + (int64_t)blockIdx.x * n;
    double acc = 0;
    for (int64_t e = threadIdx.x; e < n; e += 256) acc = __builtin_elementwise_fma(v[e], w[e], acc);
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = (part[0] + part[1]) + (part[2] + part[3]);
}

// out += sign · Σ_{i < count} coef[i] V_i.
__global__ __launch_bounds__(256) void combine_kernel(int64_t n, const double* __restrict__ V, int count,
                                                      const double* __restrict__ coef, double sign, double* out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    double acc = 0;
    for (int i = 0; i < count; ++i) acc = __builtin_elementwise_fma(coef[i], V[(int64_t)i * n + e], acc);
    out[e] = __builtin_elementwise_fma(sign, acc, out[e]);
}

// dst = src · f.
__global__ void scale_into(int64_t n, const double* __restrict__ src, double f, double* __restrict__ dst) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) dst[e] = src[e] * f;
}
// dst = f64(src) · f.
__global__ void widen_into(int64_t n, const float* __restrict__ src, double f, double* __restrict__ dst) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) dst[e] = (double)src[e] * f;
}

int sumsq(int64_t count, const double* v, double* part, double* out, hipStream_t s) {
    sumsq_kernel<<<1024, 256, 0, s>>>(count, v, part);
    sum_parts<<<1, 256, 0, s>>>(1024, part, out);
    HS_HIP(hipGetLastError());
    return OK;
}

double pow2_near(double v) {
    if (!(v > 0) || !std::isfinite(v)) return 1.0;
    return std::ldexp(1.0, -(int)std::floor(std::log2(v)));
}

}  // namespace

// Factors A (n×n f64, row-major, ld lda) in FP32 and solves A x = b (b, x: n f64, device) with FP64
// refinement. Writes the number of refinement steps and the final backward error. A is read, never written.
int gesv_ir(int64_t n, const double* a, int64_t lda, const double* b, double* x, int max_iter, double tol,
            int* iters, double* berr, hipStream_t s) {
    if (n == 0) { if (iters) *iters = 0; if (berr) *berr = 0; return OK; }
    float* a32 = (float*)pool_get(sizeof(float) * (size_t)(n * n));
    int64_t* d_piv = (int64_t*)pool_get(sizeof(int64_t) * (size_t)n);
    int64_t* d_perm = (int64_t*)pool_get(sizeof(int64_t) * (size_t)n);
    double* r = (double*)pool_get(sizeof(double) * (size_t)n);
    float* r32 = (float*)pool_get(sizeof(float) * (size_t)n);
    double* part = (double*)pool_get(sizeof(double) * 1024 + 64);
    double* sc = (double*)pool_get(sizeof(double) * 4);
    int rc = OK;
    auto fail = [&](int e) { rc = e; };
    if (!a32 || !d_piv || !d_perm || !r || !r32 || !part || !sc) rc = hip_status(hipErrorOutOfMemory, "pool_get");
    const unsigned blocks = (unsigned)((n + 255) / 256);
    double h[3] = {0, 0, 0};
    int it = 0;
    double be = INFINITY;
    if (!rc) rc = cast_d2s(n, n, a, lda, a32, n, s);
    if (!rc) rc = getrf<float>(n, a32, n, d_piv, s);
    if (!rc) {
        std::vector<int64_t> piv(n), perm(n);
        if (hipMemcpy(piv.data(), d_piv, sizeof(int64_t) * n, hipMemcpyDeviceToHost) != hipSuccess) fail(hip_status(hipErrorUnknown, "D2H"));
        for (int64_t i = 0; i < n; ++i) perm[i] = i;
        for (int64_t c = 0; c < n && !rc; ++c) std::swap(perm[c], perm[piv[c]]);
        if (!rc && hipMemcpyAsync(d_perm, perm.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice, s) != hipSuccess)
            fail(hip_status(hipErrorUnknown, "H2D"));
    }
    if (!rc) rc = sumsq(n * n, a, part, sc, s);  // ||A||_F² (A is compact when lda == n)
    if (!rc && lda != n) { set_error("gesv_ir needs lda == n"); rc = -3; }
    // x = 0, r = b on the first pass.
    if (!rc && hipMemsetAsync(x, 0, sizeof(double) * n, s) != hipSuccess) fail(hip_status(hipErrorUnknown, "memset"));
    if (!rc && hipMemcpyAsync(r, b, sizeof(double) * n, hipMemcpyDeviceToDevice, s) != hipSuccess) fail(hip_status(hipErrorUnknown, "copy"));
    if (!rc) rc = sumsq(n, r, part, sc + 1, s);
    if (!rc && hipMemcpyAsync(h, sc, sizeof(double) * 2, hipMemcpyDeviceToHost, s) != hipSuccess) fail(hip_status(hipErrorUnknown, "D2H"));
    if (!rc && hipStreamSynchronize(s) != hipSuccess) fail(hip_status(hipErrorUnknown, "sync"));
    const double anorm = std::sqrt(h[0]);
    double rnorm = std::sqrt(h[1]);
    while (!rc) {
        // Correction: d = A⁻¹ r with the low-precision factors, r scaled by a power of two into FP32's range.
        const double scale = pow2_near(rnorm / std::sqrt((double)n));
        scale_to_f32<<<blocks, 256, 0, s>>>(n, r, scale, r32);
        rc = getrs<float>(false, n, 1, a32, n, d_perm, r32, 1, s);
        if (rc) break;
        add_correction<<<blocks, 256, 0, s>>>(n, r32, 1.0 / scale, x);
        residual_kernel<<<(unsigned)((n + 3) / 4), 256, 0, s>>>(n, a, lda, x, b, r);
        if ((rc = sumsq(n, r, part, sc + 1, s))) break;
        if ((rc = sumsq(n, x, part, sc + 2, s))) break;
        if (hipMemcpyAsync(h, sc, sizeof(double) * 3, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) { fail(hip_status(hipErrorUnknown, "D2H")); break; }
        rnorm = std::sqrt(h[1]);
        be = rnorm / (anorm * std::sqrt(h[2]));
        ++it;
        if (be <= tol || it >= max_iter || !std::isfinite(be)) break;
    }
    pool_put(a32); pool_put(d_piv); pool_put(d_perm); pool_put(r); pool_put(r32); pool_put(part); pool_put(sc);
    if (iters) *iters = it;
    if (berr) *berr = be;
    return rc;
}

// GMRES-based refinement (Carson and Higham, "A new analysis of iterative refinement and its application to accurate
// solution of ill-conditioned sparse linear systems", SIAM J. Sci. Comput. 39 (2017), and "Accelerating the solution of
// linear systems by iterative refinement in three precisions", 40 (2018)): factors that are too inexact for classical
// refinement (here BF16 trailing products, u ≈ 2⁻⁸, against cond(A) up to 1e6) still make a good preconditioner.
// Each refinement step solves A d = r by restarted GMRES on M⁻¹A d = M⁻¹r, M = the low-precision LU, in FP64 on the
// GPU: matrix-vector products and vectors in FP64, M⁻¹ applied with the FP32 factors (the residual scaled by a power of
// two into FP32's range), classical Gram-Schmidt with one reorthogonalization, Givens rotations on the host. A step's
// GMRES stops when its preconditioned residual has fallen by GMRES_TOL or after `restart` steps.
int gesv_gmres_ir(int64_t n, const double* a, int64_t lda, const double* b, double* x, int max_iter, int restart,
                  double tol, int* iters, int* inner, double* berr, hipStream_t s, int terms) {
    constexpr double GMRES_TOL = 1e-4;
    if (iters) *iters = 0;
    if (inner) *inner = 0;
    if (n == 0) { if (berr) *berr = 0; return OK; }
    if (lda != n) { set_error("gesv_gmres_ir needs lda == n"); return -3; }
    if (restart < 1 || restart > 200) { set_error("gmres restart must be in 1..200"); return -7; }
    const size_t m1 = (size_t)restart + 1;
    float* a32 = (float*)pool_get(sizeof(float) * (size_t)(n * n));
    int64_t* d_piv = (int64_t*)pool_get(sizeof(int64_t) * (size_t)n);
    int64_t* d_perm = (int64_t*)pool_get(sizeof(int64_t) * (size_t)n);
    double* r = (double*)pool_get(sizeof(double) * (size_t)n);
    double* w = (double*)pool_get(sizeof(double) * (size_t)n);
    float* r32 = (float*)pool_get(sizeof(float) * (size_t)n);
    double* V = (double*)pool_get(sizeof(double) * (size_t)n * m1);
    double* coef = (double*)pool_get(sizeof(double) * m1 + 64);
    double* part = (double*)pool_get(sizeof(double) * 1024 + 64);
    double* sc = (double*)pool_get(sizeof(double) * 4);
    int rc = OK;
    if (!a32 || !d_piv || !d_perm || !r || !w || !r32 || !V || !coef || !part || !sc)
        rc = hip_status(hipErrorOutOfMemory, "pool_get");
    const unsigned blocks = (unsigned)((n + 255) / 256);
    auto sync_read = [&](const double* dev, double* host, size_t count) -> int {
        HS_HIP(hipMemcpyAsync(host, dev, sizeof(double) * count, hipMemcpyDeviceToHost, s));
        HS_HIP(hipStreamSynchronize(s));
        return OK;
    };
    auto norm = [&](const double* v, double& out) -> int {
        int e = sumsq(n, v, part, sc, s);
        if (e) return e;
        double h = 0;
        if ((e = sync_read(sc, &h, 1))) return e;
        out = std::sqrt(h);
        return OK;
    };
    // dst = M⁻¹ src (dst may be src): scaled into FP32, both triangular solves with the low-precision factors, back.
    auto precond = [&](const double* src, double* dst) -> int {
        double nv = 0;
        int e = norm(src, nv);
        if (e) return e;
        const double scale = pow2_near(nv / std::sqrt((double)n));
        scale_to_f32<<<blocks, 256, 0, s>>>(n, src, scale, r32);
        if ((e = getrs<float>(false, n, 1, a32, n, d_perm, r32, 1, s))) return e;
        widen_into<<<blocks, 256, 0, s>>>(n, r32, 1.0 / scale, dst);
        HS_HIP(hipGetLastError());
        return OK;
    };
    if (!rc) rc = cast_d2s(n, n, a, lda, a32, n, s);
    if (!rc) rc = getrf<float>(n, a32, n, d_piv, s, terms);
    if (!rc) {
        std::vector<int64_t> piv(n), perm(n);
        if (hipMemcpy(piv.data(), d_piv, sizeof(int64_t) * n, hipMemcpyDeviceToHost) != hipSuccess)
            rc = hip_status(hipErrorUnknown, "D2H");
        for (int64_t i = 0; i < n; ++i) perm[i] = i;
        for (int64_t c = 0; c < n && !rc; ++c) std::swap(perm[c], perm[piv[c]]);
        if (!rc && hipMemcpyAsync(d_perm, perm.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice, s) != hipSuccess)
            rc = hip_status(hipErrorUnknown, "H2D");
    }
    double anorm = 0;
    if (!rc) { rc = sumsq(n * n, a, part, sc, s); }
    if (!rc) { double h = 0; rc = sync_read(sc, &h, 1); anorm = std::sqrt(h); }
    if (!rc) rc = precond(b, x);  // x0 = M⁻¹ b
    int it = 0, steps = 0;
    double be = INFINITY;
    std::vector<double> H(m1 * m1), g(m1), cs(m1), sn(m1), h(m1), h2(m1), y(m1);
    while (!rc) {
        residual_kernel<<<(unsigned)((n + 3) / 4), 256, 0, s>>>(n, a, lda, x, b, r);
        double rn = 0, xn = 0;
        if ((rc = norm(r, rn)) || (rc = norm(x, xn))) break;
        be = rn / (anorm * xn);
        if (be <= tol || it >= max_iter || !std::isfinite(be)) break;
        // GMRES on M⁻¹ A d = M⁻¹ r.
        if ((rc = precond(r, w))) break;
        double beta = 0;
        if ((rc = norm(w, beta))) break;
        if (!(beta > 0)) break;
        scale_into<<<blocks, 256, 0, s>>>(n, w, 1.0 / beta, V);
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int kdone = 0;
        for (int j = 0; j < restart && !rc; ++j) {
            double* vj = V + (size_t)j * n;
            matvec_kernel<<<(unsigned)((n + 3) / 4), 256, 0, s>>>(n, a, lda, vj, w);
            if ((rc = precond(w, w))) break;
            // Classical Gram-Schmidt, twice.
            for (int pass = 0; pass < 2 && !rc; ++pass) {
                std::vector<double>& hh = pass ? h2 : h;
                dots_kernel<<<(unsigned)(j + 1), 256, 0, s>>>(n, V, w, coef);
                if ((rc = sync_read(coef, hh.data(), (size_t)j + 1))) break;
                combine_kernel<<<blocks, 256, 0, s>>>(n, V, j + 1, coef, -1.0, w);
                HS_HIP(hipGetLastError());
            }
            if (rc) break;
            double hn = 0;
            if ((rc = norm(w, hn))) break;
            for (int i = 0; i <= j; ++i) H[i * m1 + j] = h[i] + h2[i];
            H[(j + 1) * m1 + j] = hn;
            for (int i = 0; i < j; ++i) {  // earlier rotations
                const double t1 = cs[i] * H[i * m1 + j] + sn[i] * H[(i + 1) * m1 + j];
                const double t2 = -sn[i] * H[i * m1 + j] + cs[i] * H[(i + 1) * m1 + j];
                H[i * m1 + j] = t1;
                H[(i + 1) * m1 + j] = t2;
            }
            const double d0 = H[j * m1 + j], d1 = H[(j + 1) * m1 + j], den = std::hypot(d0, d1);
            cs[j] = den > 0 ? d0 / den : 1.0;
            sn[j] = den > 0 ? d1 / den : 0.0;
            H[j * m1 + j] = den;
            H[(j + 1) * m1 + j] = 0.0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            ++steps;
            kdone = j + 1;
            if (!(hn > 0) || std::fabs(g[j + 1]) <= GMRES_TOL * beta) break;
            scale_into<<<blocks, 256, 0, s>>>(n, w, 1.0 / hn, V + (size_t)(j + 1) * n);
        }
        if (rc) break;
        for (int i = kdone - 1; i >= 0; --i) {  // H y = g, upper triangular
            double acc = g[i];
            for (int q = i + 1; q < kdone; ++q) acc -= H[i * m1 + q] * y[q];
            y[i] = acc / H[i * m1 + i];
        }
        if (hipMemcpyAsync(coef, y.data(), sizeof(double) * kdone, hipMemcpyHostToDevice, s) != hipSuccess) {
            rc = hip_status(hipErrorUnknown, "H2D");
            break;
        }
        combine_kernel<<<blocks, 256, 0, s>>>(n, V, kdone, coef, 1.0, x);
        if (hipStreamSynchronize(s) != hipSuccess) { rc = hip_status(hipErrorUnknown, "sync"); break; }
        ++it;
    }
    pool_put(a32); pool_put(d_piv); pool_put(d_perm); pool_put(r); pool_put(w); pool_put(r32); pool_put(V);
    pool_put(coef); pool_put(part); pool_put(sc);
    if (iters) *iters = it;
    if (inner) *inner = steps;
    if (berr) *berr = be;
    return rc;
}

}  // namespace hs
