// Mixed-precision solves on gfx950 (no counterpart in rustymath, which has no iterative refinement): the
// matrix factors in FP32 on MFMA (or with BF16 trailing products accumulating in FP32), and FP64 iterative
// refinement on the GPU recovers FP64 accuracy: r = b - A x (FP64, one streaming pass over A), the
// correction solves with the low-precision factors, x += d in FP64, and the loop stops once the normwise
// backward error ||b - A x|| / (||A|| ||x||) (Frobenius / 2-norms, FP64) reaches `tol`.
#include "hs_kernels.h"
#include <atomic>
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
__global__ __launch_bounds__(256) void sumsq_kernel(int64_t count, const double* __restrict__ v, double* part,
                                                    const int* done) {
    __shared__ double s[4];
    if (done && *done) return;
    double acc = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        acc = __builtin_elementwise_fma(v[i], v[i], acc);
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (s[0] + s[1]) + (s[2] + s[3]);
}
__global__ void sum_parts(int count, const double* part, double* out, const int* done) {
    __shared__ double s[256];
    if (done && *done) return;
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

__global__ void fill_zero(unsigned long long* p, int64_t count) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) p[i] = 0ull;
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
// The refinement's state in device memory: every step's kernels read it and the convergence test writes it, so that a
// run of steps goes out without the host between them; once `done` is set, every later kernel of the run returns at
// once.
struct IrState {
    double scale, inv_scale;  // the residual's power-of-two scale into FP32, and its inverse
    double anorm2, rn2, xn2;  // ||A||_F², ||r||², ||x||²
    double be, tol;           // the backward error after the last step, and the target
    int it, max_iter, done, pad;
};

int sumsq(int64_t count, const double* v, double* part, double* out, hipStream_t s, const int* done = nullptr);

// The blocked single-column substitutions of the refinement's corrections (FP32 factors left row-major by getrf, FP32
// vector): one workgroup
// of 256 threads per block of IRB rows, workgroups in solve order (a block waits only on blocks launched before it, so
// no residency is assumed). A block adds each earlier block's IRB solved values as they are published (8-byte words
// that carry the pass's epoch beside the value, stored and loaded at agent scope), its tile of the factors for the
// next block in flight meanwhile; then one wave substitutes within the diagonal block (held in LDS) and publishes.
// Thread t holds rows 4 (t % 32) .. + 3 of the block and the 16 columns 16 (t / 32) .. + 15 of each earlier block.
// There is no reference counterpart (rustymath has no refinement): the order of each row's sum is not `single`'s,
// only its accuracy matters here. LOWER: L y = P r·scale (unit diagonal, increasing order; rows gathered through
// perm and scaled into FP32 on load), y into yout. Otherwise U d = y (decreasing order, a reciprocal multiply at each
// row's turn), then x += d / scale (accumulate) or x = d / scale in FP64.
constexpr int IRB = 128;
constexpr uint64_t IR_SPIN_LIMIT = 1ull << 22;

__device__ __forceinline__ bool ir_take(unsigned long long w, unsigned epoch, float& v) {
    v = __builtin_bit_cast(float, (unsigned)w);
    return (unsigned)(w >> 32) == epoch;
}

template <bool LOWER>
__global__ __launch_bounds__(256) void ir_trsv_kernel(int64_t n, const float* __restrict__ f, int64_t ld,
                                                      const double* __restrict__ r, const int64_t* __restrict__ perm,
                                                      const float* __restrict__ yin, float* __restrict__ yout,
                                                      double* __restrict__ x, int accumulate,
                                                      unsigned long long* pub, unsigned epoch, const IrState* st,
                                                      int* status) {
    // The diagonal block, row by row (row stride IRB + 4 floats, so that 16-byte reads of the rows of a wave's lanes
    // spread over the banks): s_t[i][j] = L(r0 + i, r0 + j) below the diagonal for LOWER, U(r0 + i, r0 + j) / U(r0 + i,
    // r0 + i) above it otherwise, and 0 everywhere else, so that a substitution step needs neither a select nor a
    // divide.
    constexpr int TS = IRB + 4;
    __shared__ __attribute__((aligned(16))) float s_t[IRB][TS];
    __shared__ float s_rc[IRB];                                   // 1 / U(i, i) (1 on the padding)
    __shared__ __attribute__((aligned(16))) float s_y[IRB];       // an earlier block's values
    __shared__ float s_acc[8][IRB];                               // the rows' partial sums, by column group
    if (st->done) return;
    if (__hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    const int nb = (int)((n + IRB - 1) / IRB), u = blockIdx.x;
    const int q = LOWER ? u : nb - 1 - u;
    const int64_t r0 = (int64_t)q * IRB;
    const int h = (int)(n - r0 < IRB ? n - r0 : IRB);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int rg = tid & 31, cg = tid >> 5;  // rows 4 rg .. 4 rg + 3, columns 16 cg .. 16 cg + 15 of a block
    {
        if (!LOWER && tid < IRB) s_rc[tid] = tid < h ? 1.f / f[(r0 + tid) * ld + r0 + tid] : 1.f;
        if (!LOWER) __syncthreads();
        float tmp[IRB * IRB / 256];
#pragma unroll
        for (int k = 0; k < IRB * IRB / 256; ++k) {  // row i of the block, columns j along the threads: coalesced
            const int idx = tid + 256 * k, i = idx / IRB, j = idx % IRB;
            tmp[k] = f[(r0 + (i < h ? i : h - 1)) * ld + r0 + (j < h ? j : h - 1)];
        }
#pragma unroll
        for (int k = 0; k < IRB * IRB / 256; ++k) {
            const int idx = tid + 256 * k, i = idx / IRB, j = idx % IRB;
            const bool in = j < h && i < h && (LOWER ? i > j : i < j);
            s_t[i][j] = in ? (LOWER ? tmp[k] : tmp[k] * s_rc[i]) : 0.f;
        }
    }
    // The right-hand side of this block's rows (wave 0: rows lane and lane + 64).
    float c0v = 0.f, c1v = 0.f;
    if (wave == 0) {
        const double sc = st->scale;
        if (LOWER) {
            if (lane < h) c0v = (float)(r[perm[r0 + lane]] * sc);
            if (lane + 64 < h) c1v = (float)(r[perm[r0 + lane + 64]] * sc);
        } else {
            if (lane < h) c0v = yin[r0 + lane];
            if (lane + 64 < h) c1v = yin[r0 + lane + 64];
        }
    }
    // Earlier blocks, in solve order t = 0 .. u - 1: block p(t). The tile F(r0 + 4 rg + e, p IRB + 16 cg + j) of
    // producer t: four rows' 16 items each, four to a 16-byte load (the factors are row-major: f[row ld + col]).
    const int np = u;
    auto blk = [&](int t) { return LOWER ? t : nb - 1 - t; };
    const int64_t rowb = r0 + 4 * rg;
    const bool vec = (ld & 3) == 0;  // rows start 16-byte aligned
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    typedef float f4 __attribute__((ext_vector_type(4)));
    auto load_tile = [&](int t, f4* e) {  // e[4 r + q]: row rowb + r, columns c0 + 4 q .. + 3
        const int64_t c0 = (int64_t)blk(t) * IRB + 16 * cg;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int64_t row = rowb + r < n ? rowb + r : n - 1;
            const float* fr = f + row * ld;
            if (vec && c0 + 16 <= n) {
#pragma unroll
                for (int q = 0; q < 4; ++q) e[4 * r + q] = *reinterpret_cast<const f4*>(fr + c0 + 4 * q);
            } else {  // columns past n read column n - 1 (their values meet zeros)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    f4 v;
#pragma unroll
                    for (int m = 0; m < 4; ++m) { const int64_t c = c0 + 4 * q + m; v[m] = fr[c < n ? c : n - 1]; }
                    e[4 * r + q] = v;
                }
            }
        }
    };
    // Waits for producer t's values into s_y (wave 0 polls, two words a lane); false where the wait ran out.
    auto fetch = [&](int t) -> bool {
        if (wave == 0) {
            const int64_t p0 = (int64_t)blk(t) * IRB;
            const int ph = (int)(n - p0 < IRB ? n - p0 : IRB);
            float v0 = 0.f, v1 = 0.f;
            bool ok0 = lane >= ph, ok1 = lane + 64 >= ph;
            const int64_t i0 = p0 + (lane < ph ? lane : 0), i1 = p0 + (lane + 64 < ph ? lane + 64 : 0);
            uint64_t spins = 0;
            for (;;) {
                if (!ok0) ok0 = ir_take(__hip_atomic_load(&pub[i0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), epoch, v0);
                if (!ok1) ok1 = ir_take(__hip_atomic_load(&pub[i1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), epoch, v1);
                if (__all(ok0 && ok1)) break;
                // A wait that runs out (or one that ran out elsewhere) voids the pass: this block goes on with zeros and
                // still publishes, so that no later block waits on it.
                ++spins;
                if (spins > IR_SPIN_LIMIT ||
                    ((spins & 1023) == 0 && __hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
                    if (lane == 0) atomicExch(status, ERR_LIMIT);
                    v0 = v1 = 0.f;
                    break;
                }
            }
            s_y[lane] = lane < ph ? v0 : 0.f;
            s_y[lane + 64] = lane + 64 < ph ? v1 : 0.f;
        }
        __syncthreads();
        return true;
    };
    auto terms = [&](const f4* e) {
        const f4* yv = reinterpret_cast<const f4*>(&s_y[16 * cg]);
        f4 y4[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) y4[k] = yv[k];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const float yj = y4[j / 4][j % 4];
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[r] = __builtin_elementwise_fma(-e[4 * r + j / 4][j % 4], yj, acc[r]);
        }
    };
    {
        f4 ea[16], eb[16];
        if (np > 0) load_tile(0, ea);
        int t = 0;
        for (; t + 1 < np; t += 2) {
            load_tile(t + 1, eb);
            fetch(t);
            terms(ea);
            __syncthreads();  // s_y is read before the next fetch writes it
            if (t + 2 < np) load_tile(t + 2, ea);
            fetch(t + 1);
            terms(eb);
            __syncthreads();
        }
        if (t < np) {
            fetch(t);
            terms(ea);
        }
    }
#pragma unroll
    for (int e2 = 0; e2 < 4; ++e2) s_acc[cg][4 * rg + e2] = acc[e2];
    __syncthreads();
    if (wave != 0) return;
    // The diagonal block: wave 0, rows lane and lane + 64, one step per column in solve order; a row's entries arrive
    // four columns per 16-byte read, several steps ahead of their use.
    float t0 = c0v, t1 = c1v;
#pragma unroll
    for (int g = 0; g < 8; ++g) { t0 += s_acc[g][lane]; t1 += s_acc[g][lane + 64]; }
    if (!LOWER) { t0 *= s_rc[lane]; t1 *= s_rc[lane + 64]; }
    {
        typedef float f4 __attribute__((ext_vector_type(4)));
        const f4* row0 = reinterpret_cast<const f4*>(&s_t[lane][0]);
        const f4* row1 = reinterpret_cast<const f4*>(&s_t[lane + 64][0]);
        constexpr int NC = IRB / 4, D = 4;  // chunks of four steps, D chunks in flight
        f4 w0[D], w1[D];
        // Chunk c in solve order covers columns 4 c' .. 4 c' + 3 with c' = c (LOWER) or NC - 1 - c.
        auto chunk_col = [](int c) { return LOWER ? c : NC - 1 - c; };
#pragma unroll
        for (int c = 0; c < D; ++c) { w0[c] = row0[chunk_col(c)]; w1[c] = row1[chunk_col(c)]; }
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const f4 e0 = w0[c % D], e1 = w1[c % D];
            if (c + D < NC) { w0[c % D] = row0[chunk_col(c + D)]; w1[c % D] = row1[chunk_col(c + D)]; }
            const int cc = chunk_col(c);
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const int e = LOWER ? m : 3 - m, j = 4 * cc + e;  // this step's column
                // Row i's entry in column j is 0 unless i > j (LOWER) or i < j; the steps that reach no row of a half
                // are left out.
                const float vj = j < 64 ? __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, t0), j))
                                        : __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, t1), j - 64));
                if (LOWER ? j < 63 : j > 0) t0 = __builtin_elementwise_fma(-e0[e], vj, t0);
                if (LOWER ? j < IRB - 1 : j > 64) t1 = __builtin_elementwise_fma(-e1[e], vj, t1);
            }
        }
    }
    // Publish (the words for later blocks), then the outputs.
    const unsigned long long tag = (unsigned long long)epoch << 32;
    if (lane < h) __hip_atomic_store(&pub[r0 + lane], tag | __builtin_bit_cast(unsigned, t0), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (lane + 64 < h)
        __hip_atomic_store(&pub[r0 + lane + 64], tag | __builtin_bit_cast(unsigned, t1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (LOWER) {
        if (lane < h) yout[r0 + lane] = t0;
        if (lane + 64 < h) yout[r0 + lane + 64] = t1;
    } else {
        const double is = st->inv_scale;
        if (lane < h) { const int64_t i = r0 + lane; x[i] = accumulate ? __builtin_elementwise_fma((double)t0, is, x[i]) : (double)t0 * is; }
        if (lane + 64 < h) {
            const int64_t i = r0 + lane + 64;
            x[i] = accumulate ? __builtin_elementwise_fma((double)t1, is, x[i]) : (double)t1 * is;
        }
    }
}

// scale = 2^-floor(log2(||r|| / sqrt(n))) (1 where that is not a positive finite number), so that the scaled residual's
// items are near 1 in FP32; inv_scale = 1 / scale (exact).
__global__ void ir_scale_kernel(int64_t n, IrState* st) {
    if (st->done) return;
    const double v = sqrt(st->rn2) / sqrt((double)n);
    double sc = 1.0;
    if (v > 0 && v <= 1.79e308) { int e; frexp(v, &e); sc = ldexp(1.0, 1 - e); }
    st->scale = sc;
    st->inv_scale = 1.0 / sc;
}
// The convergence test, on the GPU: be = ||r|| / (||A||_F ||x||); done once be <= tol, after max_iter steps, or where
// be is not finite.
__global__ void ir_test_kernel(IrState* st) {
    if (st->done) return;
    const double be = sqrt(st->rn2) / (sqrt(st->anorm2) * sqrt(st->xn2));
    st->be = be;
    st->it += 1;
    if (be <= st->tol || st->it >= st->max_iter || !(be <= 1.79e308)) st->done = 1;
}
__global__ void ir_init_kernel(IrState* st, double tol, int max_iter) {
    st->scale = st->inv_scale = 1.0;
    st->rn2 = st->xn2 = 0.0;
    st->be = INFINITY;
    st->tol = tol;
    st->it = 0;
    st->max_iter = max_iter;
    st->done = 0;
}
// r = b - A x, as residual_kernel, skipped once the refinement is done.
__global__ __launch_bounds__(256) void ir_residual_kernel(int64_t n, const double* __restrict__ a, int64_t lda,
                                                          const double* __restrict__ x, const double* __restrict__ b,
                                                          double* __restrict__ r, const IrState* st) {
    if (st->done) return;
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

// The power of two 2^-floor(log2 v) (1 where v is not a positive finite number).
double pow2_near(double v) {
    if (!(v > 0) || !std::isfinite(v)) return 1.0;
    return std::ldexp(1.0, -(int)std::floor(std::log2(v)));
}
__global__ void ir_set_scale_kernel(IrState* st, double scale) {
    st->scale = scale;
    st->inv_scale = 1.0 / scale;
}

// The refinement's words: one per row, their epochs unique per pass (two passes a step) and never 0, so the cleared
// buffer and words left by earlier passes never match.
std::atomic<unsigned> g_ir_epoch{1};
unsigned next_ir_epochs() {
    unsigned e = g_ir_epoch.fetch_add(2);
    if (e == 0 || e + 1 == 0) e = g_ir_epoch.fetch_add(2);
    return e;
}

// d = A⁻¹ r·scale through the FP32 factors (the two blocked passes), then x += d / scale (accumulate) or x = d / scale.
int ir_lu_solve(int64_t n, const float* f, int64_t ld, const int64_t* perm, const double* r, float* y32, double* x,
                int accumulate, unsigned long long* pub, const IrState* st, int* status, hipStream_t s) {
    const unsigned nb = (unsigned)((n + IRB - 1) / IRB), e = next_ir_epochs();
    ir_trsv_kernel<true><<<nb, 256, 0, s>>>(n, f, ld, r, perm, nullptr, y32, nullptr, 0, pub, e, st, status);
    ir_trsv_kernel<false><<<nb, 256, 0, s>>>(n, f, ld, nullptr, nullptr, y32, nullptr, x, accumulate, pub, e + 1, st, status);
    HS_HIP(hipGetLastError());
    return OK;
}

int sumsq(int64_t count, const double* v, double* part, double* out, hipStream_t s, const int* done) {
    sumsq_kernel<<<1024, 256, 0, s>>>(count, v, part, done);
    sum_parts<<<1, 256, 0, s>>>(1024, part, out, done);
    HS_HIP(hipGetLastError());
    return OK;
}

}  // namespace

// Factors A (n×n f64, row-major, ld lda) in FP32 and solves A x = b (b, x: n f64, device) with FP64
// refinement. Writes the number of refinement steps and the final backward error. A is read, never written.
// The steps run from the GPU's own state: the host enqueues them a few at a time and reads the state once per batch.
int gesv_ir(int64_t n, const double* a, int64_t lda, const double* b, double* x, int max_iter, double tol,
            int* iters, double* berr, hipStream_t s) {
    if (n == 0) { if (iters) *iters = 0; if (berr) *berr = 0; return OK; }
    if (lda != n) { set_error("gesv_ir needs lda == n"); return -3; }
    if (max_iter < 1) max_iter = 1;
    float* a32 = (float*)pool_get(sizeof(float) * (size_t)(n * n));
    int64_t* d_piv = (int64_t*)pool_get(sizeof(int64_t) * (size_t)n);
    int64_t* d_perm = (int64_t*)pool_get(sizeof(int64_t) * (size_t)n);
    double* r = (double*)pool_get(sizeof(double) * (size_t)n);
    float* y32 = (float*)pool_get(sizeof(float) * (size_t)n);
    unsigned long long* pub = (unsigned long long*)pool_get(sizeof(unsigned long long) * (size_t)n + 64);
    double* part = (double*)pool_get(sizeof(double) * 1024 + 64);
    IrState* st = (IrState*)pool_get(sizeof(IrState) + 64);
    int* status = (int*)(st + 1);
    int rc = OK;
    if (!a32 || !d_piv || !d_perm || !r || !y32 || !pub || !part || !st) rc = hip_status(hipErrorOutOfMemory, "pool_get");
    const unsigned blocks = (unsigned)((n + 255) / 256);
    int it = 0;
    double be = INFINITY;
    // float(A), with ||A||_F² from the same pass (A is compact: lda == n); then the FP32 factors, left row-major.
    if (!rc) rc = cast_d2s_sumsq(n * n, a, a32, part, 1024, s);
    if (!rc) {
        ir_init_kernel<<<1, 1, 0, s>>>(st, tol, max_iter);
        sum_parts<<<1, 256, 0, s>>>(1024, part, &st->anorm2, nullptr);
        HS_HIP(hipGetLastError());
    }
    if (!rc) rc = getrf<float>(n, a32, n, d_piv, s, 0, d_perm, true);
    if (!rc) {
        fill_zero<<<1, 64, 0, s>>>((unsigned long long*)status, 1);
        fill_zero<<<(unsigned)std::min<int64_t>((n + 255) / 256, 1024), 256, 0, s>>>(pub, n);
        HS_HIP(hipGetLastError());
    }
    // x = 0, r = b before the first step.
    if (!rc) fill_zero<<<(unsigned)std::min<int64_t>((n + 255) / 256, 1024), 256, 0, s>>>((unsigned long long*)x, n);
    if (!rc && hipMemcpyAsync(r, b, sizeof(double) * n, hipMemcpyDeviceToDevice, s) != hipSuccess) rc = hip_status(hipErrorUnknown, "copy");
    if (!rc) rc = sumsq(n, r, part, &st->rn2, s);
    // Steps: scale, the correction through the FP32 factors (x += d / scale), r = b - A x, the norms, the test. Two
    // steps per batch at first (what the usual matrix needs), then four.
    IrState hst{};
    for (int batch = 2; !rc; batch = 4) {
        for (int k = 0; k < batch && !rc; ++k) {
            ir_scale_kernel<<<1, 1, 0, s>>>(n, st);
            rc = ir_lu_solve(n, a32, n, d_perm, r, y32, x, 1, pub, st, status, s);
            if (rc) break;
            ir_residual_kernel<<<(unsigned)((n + 3) / 4), 256, 0, s>>>(n, a, lda, x, b, r, st);
            if ((rc = sumsq(n, r, part, &st->rn2, s, &st->done))) break;
            if ((rc = sumsq(n, x, part, &st->xn2, s, &st->done))) break;
            ir_test_kernel<<<1, 1, 0, s>>>(st);
            HS_HIP(hipGetLastError());
        }
        if (rc) break;
        if (hipMemcpyAsync(&hst, st, sizeof(IrState), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) { rc = hip_status(hipErrorUnknown, "D2H"); break; }
        if (hst.done) break;
    }
    if (!rc) {
        int stv = 0;
        if (hipMemcpy(&stv, status, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) rc = hip_status(hipErrorUnknown, "D2H");
        else if (stv) { set_error("refinement solve wait ran out"); rc = ERR_LIMIT; }
    }
    if (!rc) { it = hst.it; be = hst.be; }
    (void)blocks;
    pool_put(a32); pool_put(d_piv); pool_put(d_perm); pool_put(r); pool_put(y32); pool_put(pub); pool_put(part); pool_put(st);
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
    unsigned long long* pub = (unsigned long long*)pool_get(sizeof(unsigned long long) * (size_t)n + 64);
    IrState* st = (IrState*)pool_get(sizeof(IrState) + 64);
    int* status = st ? (int*)(st + 1) : nullptr;
    int rc = OK;
    if (!a32 || !d_piv || !d_perm || !r || !w || !r32 || !V || !coef || !part || !sc || !pub || !st)
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
        ir_set_scale_kernel<<<1, 1, 0, s>>>(st, scale);
        if ((e = ir_lu_solve(n, a32, n, d_perm, src, r32, dst, 0, pub, st, status, s))) return e;
        int stv = 0;
        HS_HIP(hipMemcpyAsync(&stv, status, sizeof(int), hipMemcpyDeviceToHost, s));
        HS_HIP(hipStreamSynchronize(s));
        if (stv) { set_error("refinement solve wait ran out"); return ERR_LIMIT; }
        return OK;
    };
    if (!rc) rc = cast_d2s(n, n, a, lda, a32, n, s);
    if (!rc) rc = getrf<float>(n, a32, n, d_piv, s, terms, d_perm, true);
    if (!rc) {
        ir_init_kernel<<<1, 1, 0, s>>>(st, 0.0, 1);
        fill_zero<<<1, 64, 0, s>>>((unsigned long long*)status, 1);
        fill_zero<<<(unsigned)std::min<int64_t>((n + 255) / 256, 1024), 256, 0, s>>>(pub, n);
        if (hipGetLastError() != hipSuccess) rc = hip_status(hipErrorUnknown, "launch");
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
    pool_put(coef); pool_put(part); pool_put(sc); pool_put(pub); pool_put(st);
    if (iters) *iters = it;
    if (inner) *inner = steps;
    if (berr) *berr = be;
    return rc;
}

}  // namespace hs
