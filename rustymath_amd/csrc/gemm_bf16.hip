// The low-precision trailing product of the BF16 factorization: c (FP32) = c - a b with the items of a and b taken as
// bfloat16 and the products summed in FP32 on MFMA (v_mfma_f32_16x16x32_bf16). This form has no counterpart in
// rustymath (which factors in FP64 or FP32 only); its results are not the reference's bits, and the refinement that
// follows it recovers FP64 accuracy.
//
// TERMS = 1 rounds each item once to bfloat16 (nearest even): products carry a relative error up to 2^-8. TERMS = 3
// splits each item x into hi = bf16(x) and lo = bf16(x - hi) (x - hi is exact in FP32), and sums a_lo b_hi,
// a_hi b_lo and a_hi b_hi into the FP32 accumulator (the dropped a_lo b_lo and lo's rounding leave about 2^-16):
// three BF16 products at the BF16 MFMA rate, close to FP32 operands in accuracy.
//
// Tiles of 128x128, 4 waves of 64x64 (4x4 MFMA tiles each), k in steps of 32: a's rows and b's columns are staged in
// LDS as bfloat16, each row (column) holding its 32 items of the step contiguously, so that a lane reads its 8 items
// for an MFMA operand with one 16-byte load.
#include "hs_kernels.h"

namespace hs {

namespace {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int TM = 128, TN = 128, TK = 32, NT = 256, KP = TK + 8;  // KP: LDS row pitch (items), 16-byte aligned

__device__ __forceinline__ unsigned short to_bf16(float x) {
    const unsigned u = __builtin_bit_cast(unsigned, x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40u);  // NaN stays NaN
    return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ float from_bf16(unsigned short h) { return __builtin_bit_cast(float, (unsigned)h << 16); }

template <int TERMS>
__global__ __launch_bounds__(NT, 2) void gemm_bf16_kernel(int64_t m, int64_t n, int64_t k, const float* __restrict__ A,
                                                         int64_t lda, const float* __restrict__ B, int64_t ldb,
                                                         float* __restrict__ C, int64_t ldc, int tiles_m, int tiles_n) {
    constexpr int P = TERMS == 3 ? 2 : 1;  // parts per item: hi (and lo)
    __shared__ __attribute__((aligned(16))) unsigned short As[P][TM][KP];  // rows of a, negated
    __shared__ __attribute__((aligned(16))) unsigned short Bs[P][TN][KP];  // columns of b
    const int total = tiles_m * tiles_n, per = (total + 7) / 8;
    const int id = (blockIdx.x % 8) * per + blockIdx.x / 8;  // each XCD a contiguous run of tiles
    if (id >= total) return;
    const int64_t m0 = (int64_t)(id % tiles_m) * TM, n0 = (int64_t)(id / tiles_m) * TN;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6, wm = wave >> 1, wn = wave & 1;

    f4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int64_t r = m0 + wm * 64 + i * 16 + 4 * (lane >> 4) + e, c = n0 + wn * 64 + j * 16 + (lane & 15);
                const int64_t rr = r < m ? r : m - 1, cc = c < n ? c : n - 1;
                acc[i][j][e] = C[rr * ldc + cc];
            }

    for (int64_t p0 = 0; p0 < k; p0 += TK) {
        // a: thread -> row t / 2, 16 steps from (t % 2) * 16; b: thread -> step t / 8, 16 columns from (t % 8) * 16.
        float av[16], bv[16];
        {
            const int64_t r = m0 + (t >> 1), p = p0 + (t & 1) * 16;
            const int64_t rr = r < m ? r : m - 1;
#pragma unroll
            for (int q = 0; q < 16; ++q) av[q] = A[rr * lda + (p + q < k ? p + q : k - 1)];
#pragma unroll
            for (int q = 0; q < 16; ++q) av[q] = (r < m && p + q < k) ? -av[q] : 0.0f;
        }
        {
            const int64_t p = p0 + (t >> 3), c = n0 + (t & 7) * 16;
            const int64_t pp = p < k ? p : k - 1;
#pragma unroll
            for (int q = 0; q < 16; ++q) bv[q] = B[pp * ldb + (c + q < n ? c + q : n - 1)];
#pragma unroll
            for (int q = 0; q < 16; ++q) bv[q] = (p < k && c + q < n) ? bv[q] : 0.0f;
        }
        __syncthreads();  // the previous step's reads of As and Bs are done
        {
            const int r = t >> 1, p = (t & 1) * 16;
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const unsigned short hi = to_bf16(av[q]);
                As[0][r][p + q] = hi;
                if (TERMS == 3) As[P - 1][r][p + q] = to_bf16(av[q] - from_bf16(hi));
            }
            const int pk = t >> 3, c = (t & 7) * 16;
#pragma unroll
            for (int q = 0; q < 16; ++q) {
                const unsigned short hi = to_bf16(bv[q]);
                Bs[0][c + q][pk] = hi;
                if (TERMS == 3) Bs[P - 1][c + q][pk] = to_bf16(bv[q] - from_bf16(hi));
            }
        }
        __syncthreads();
        bf8 fa[P][4], fb[P][4];
        const int kq = 8 * (lane >> 4), li = lane & 15;
#pragma unroll
        for (int h = 0; h < P; ++h) {
#pragma unroll
            for (int i = 0; i < 4; ++i) fa[h][i] = *reinterpret_cast<const bf8*>(&As[h][wm * 64 + i * 16 + li][kq]);
#pragma unroll
            for (int j = 0; j < 4; ++j) fb[h][j] = *reinterpret_cast<const bf8*>(&Bs[h][wn * 64 + j * 16 + li][kq]);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (TERMS == 3) {  // the small products first
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[P - 1][i], fb[0][j], acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[0][i], fb[P - 1][j], acc[i][j], 0, 0, 0);
                }
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[0][i], fb[0][j], acc[i][j], 0, 0, 0);
            }
    }

#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int64_t r = m0 + wm * 64 + i * 16 + 4 * (lane >> 4) + e, c = n0 + wn * 64 + j * 16 + (lane & 15);
                if (r < m && c < n) C[r * ldc + c] = acc[i][j][e];
            }
}

}  // namespace

int gemm_bf16_update(int64_t m, int64_t n, int64_t k, const float* a, int64_t lda, const float* b, int64_t ldb, float* c,
                     int64_t ldc, hipStream_t s, int terms) {
    if (m == 0 || n == 0 || k == 0) return OK;
    const int tiles_m = (int)((m + TM - 1) / TM), tiles_n = (int)((n + TN - 1) / TN);
    const unsigned grid = (unsigned)(((tiles_m * tiles_n + 7) / 8) * 8);
    if (terms == 3) gemm_bf16_kernel<3><<<grid, NT, 0, s>>>(m, n, k, a, lda, b, ldb, c, ldc, tiles_m, tiles_n);
    else gemm_bf16_kernel<1><<<grid, NT, 0, s>>>(m, n, k, a, lda, b, ldb, c, ldc, tiles_m, tiles_n);
    HS_HIP(hipGetLastError());
    return OK;
}

}  // namespace hs
