// The MFMA product's tile, shared by the product kernel (gemm.hip) and the LU's trailing-product workers (lu.hip).
//
// Reference: core/src/linalg.rs product :1353-1426. rustymath's product gives each result as one fused multiply-add per
// step of p, in order, from 0 (beta = 0) or from beta c, with alpha folded into a's packed copy. A tile keeps one
// accumulator per result for the whole of k and feeds the MFMA unit in order of p, so it computes the same chain of
// roundings; a k tail pads with +0 · (-0) = -0, which adds nothing to any value.
#pragma once
#include "hs_common.h"

namespace hs {
namespace tile {

constexpr int BM = 128, BN = 128, BK = 16;

template <typename T> struct Pad { static constexpr int v = 2; };
template <> struct Pad<float> { static constexpr int v = 4; };

// Tile number id -> (tm, tn), in runs of 8 tile-rows that share B's column panels in L2.
__device__ __forceinline__ void grouped(int id, int tiles_m, int tiles_n, int& tm, int& tn) {
    constexpr int G = 8;
    const int group = id / (G * tiles_n);
    const int first = group * G;
    const int gm = tiles_m - first < G ? tiles_m - first : G;
    const int in = id - group * G * tiles_n;
    tm = first + in % gm;
    tn = in / gm;
}

// One 128x128 tile of c = alpha op(a) op(b) + beta c, on WAVES waves (t = 0 .. 64 WAVES - 1): four waves of 64x64
// each, or eight of 32x64. The threads may be half of a larger workgroup: every __syncthreads here is the
// workgroup's, 1 + ceil(k / BK) of them whatever the path, so two halves working on two tiles with the same k meet at
// each. EDGE selects the path for tiles that cross
// the last row or column: there every address is clamped into the matrix and the value replaced afterwards, so no
// load sits behind a per-lane branch (a guarded load makes the compiler wait for each load before the next). Interior
// tiles load without checks; a k tail (the last step when BK does not divide k) takes the clamped loads in either
// path. store = false computes the tile and writes nothing.
template <typename T, bool TA, bool TB, bool EDGE, int WAVES = 4>
__device__ __forceinline__ void gemm_tile(int64_t m, int64_t n, int64_t k, T alpha, const T* __restrict__ A, int64_t lda,
                                          const T* __restrict__ B, int64_t ldb, T beta, T* __restrict__ C, int64_t ldc,
                                          int64_t m0, int64_t n0, T (*As)[BK][BM + Pad<T>::v], T (*Bs)[BK][BN + Pad<T>::v],
                                          int t, bool store) {
    static_assert(WAVES == 4 || WAVES == 8, "four or eight waves");
    typedef typename Mfma<T>::acc Acc;
    constexpr int RW = 2 * BM / WAVES, NI = RW / 16;  // rows per wave, 16-row MFMA tiles per wave
    constexpr int NQ = BK * BM / (64 * WAVES);         // items of A (and of B) each thread brings per step
    constexpr int SPLIT = BK / NQ, GROUPS = BM / NQ;   // threads per row (column) of k; threads per step
    const int lane = t & 63, wave = t >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const bool scale_a = alpha != T(1);

    // Global loads: each thread brings NQ items of A's tile and NQ of B's per step of BK. Items outside the matrix are
    // +0 on A's side and, past k, -0 on B's side, so that padding terms add (+0)(-0) = -0, which changes no sum. alpha
    // scales only the items that exist (bits of amask), when they are stored to LDS: the loads of the next step stay in
    // flight while the MFMAs of this one run.
    T ra[NQ], rb[NQ];
    unsigned amask = 0xffu;
    auto load = [&](int64_t p0) {
        const bool checked = EDGE || p0 + BK > k;
        if (!TA) {  // A row-major m×k: thread -> row t / SPLIT, steps (t % SPLIT) NQ ..+NQ
            const int64_t r = m0 + t / SPLIT, p = p0 + (t % SPLIT) * NQ;
            if (!checked) {
#pragma unroll
                for (int q = 0; q < NQ; ++q) ra[q] = A[r * lda + p + q];
            } else {
                const int64_t rr = r < m ? r : m - 1;
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
                    const int64_t pp = p + q < k ? p + q : k - 1;
                    const T v = A[rr * lda + pp];
                    ra[q] = (r < m && p + q < k) ? v : T(0);
                }
            }
            amask = 0xffu;
            if (checked) {
#pragma unroll
                for (int q = 0; q < NQ; ++q) amask &= (r < m && p + q < k) ? 0xffu : ~(1u << q);
            }
        } else {  // A stored k×m: thread -> step t / GROUPS, rows (t % GROUPS) NQ ..+NQ
            const int64_t p = p0 + t / GROUPS, r = m0 + (t % GROUPS) * NQ;
            if (!checked) {
#pragma unroll
                for (int q = 0; q < NQ; ++q) ra[q] = A[p * lda + r + q];
            } else {
                const int64_t pp = p < k ? p : k - 1;
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
                    const int64_t rr = r + q < m ? r + q : m - 1;
                    const T v = A[pp * lda + rr];
                    ra[q] = (p < k && r + q < m) ? v : T(0);
                }
            }
            amask = 0xffu;
            if (checked) {
#pragma unroll
                for (int q = 0; q < NQ; ++q) amask &= (p < k && r + q < m) ? 0xffu : ~(1u << q);
            }
        }
        if (!TB) {  // B row-major k×n: thread -> step t / GROUPS, columns (t % GROUPS) NQ ..+NQ
            const int64_t p = p0 + t / GROUPS, c = n0 + (t % GROUPS) * NQ;
            if (!checked) {
#pragma unroll
                for (int q = 0; q < NQ; ++q) rb[q] = B[p * ldb + c + q];
            } else {
                const int64_t pp = p < k ? p : k - 1;
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
                    const int64_t cc = c + q < n ? c + q : n - 1;
                    const T v = B[pp * ldb + cc];
                    rb[q] = p < k ? ((c + q < n) ? v : T(0)) : T(-0.0);
                }
            }
        } else {  // B stored n×k: thread -> column t / SPLIT, steps (t % SPLIT) NQ ..+NQ
            const int64_t c = n0 + t / SPLIT, p = p0 + (t % SPLIT) * NQ;
            if (!checked) {
#pragma unroll
                for (int q = 0; q < NQ; ++q) rb[q] = B[c * ldb + p + q];
            } else {
                const int64_t cc = c < n ? c : n - 1;
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
                    const int64_t pp = p + q < k ? p + q : k - 1;
                    const T v = B[cc * ldb + pp];
                    rb[q] = p + q < k ? ((c < n) ? v : T(0)) : T(-0.0);
                }
            }
        }
    };
    auto stash = [&](int buf) {
        if (scale_a) {
#pragma unroll
            for (int q = 0; q < NQ; ++q) ra[q] = ((amask >> q) & 1u) ? alpha * ra[q] : ra[q];
        }
        if (!TA) {
            const int r = t / SPLIT, p = (t % SPLIT) * NQ;
#pragma unroll
            for (int q = 0; q < NQ; ++q) As[buf][p + q][r] = ra[q];
        } else {
            const int p = t / GROUPS, r = (t % GROUPS) * NQ;
#pragma unroll
            for (int q = 0; q < NQ; ++q) As[buf][p][r + q] = ra[q];
        }
        if (!TB) {
            const int p = t / GROUPS, c = (t % GROUPS) * NQ;
#pragma unroll
            for (int q = 0; q < NQ; ++q) Bs[buf][p][c + q] = rb[q];
        } else {
            const int c = t / SPLIT, p = (t % SPLIT) * NQ;
#pragma unroll
            for (int q = 0; q < NQ; ++q) Bs[buf][p + q][c] = rb[q];
        }
    };

    const int64_t steps = (k + BK - 1) / BK;
    // The first step's A and B go out before C, so that the stash below waits for them only (the wait counter counts
    // loads in order) and C's loads stay in flight until the first MFMA.
    load(0);
    // Accumulators start at 0, or at beta c (c itself where beta is 1), as `native_gemm` scales c first.
    Acc acc[NI][4];
    if (beta != T(0)) {
#pragma unroll
        for (int i = 0; i < NI; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int64_t r = m0 + wm * RW + i * 16 + Mfma<T>::row(lane, e), col = n0 + wn * 64 + j * 16 + (lane & 15);
                    T v;
                    if (EDGE) {
                        const int64_t rr = r < m ? r : m - 1, cc = col < n ? col : n - 1;
                        v = C[rr * ldc + cc];
                    } else {
                        v = C[r * ldc + col];
                    }
                    acc[i][j][e] = v;  // scaled by beta below, after the first step's stash, so nothing waits for C here
                }
    } else {
#pragma unroll
        for (int i = 0; i < NI; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int e = 0; e < 4; ++e) acc[i][j][e] = T(0);
    }

    stash(0);
    __syncthreads();
    if (beta != T(0) && beta != T(1)) {
#pragma unroll
        for (int i = 0; i < NI; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int e = 0; e < 4; ++e) acc[i][j][e] = beta * acc[i][j][e];
    }
    for (int64_t s = 0; s < steps; ++s) {
        const int buf = s & 1;
        if (s + 1 < steps) load((s + 1) * BK);
#pragma unroll
        for (int kk = 0; kk < BK / 4; ++kk) {
            T fa[NI], fb[4];
            const int kr = kk * 4 + (lane >> 4), li = lane & 15;
#pragma unroll
            for (int i = 0; i < NI; ++i) fa[i] = As[buf][kr][wm * RW + i * 16 + li];
#pragma unroll
            for (int j = 0; j < 4; ++j) fb[j] = Bs[buf][kr][wn * 64 + j * 16 + li];
#pragma unroll
            for (int i = 0; i < NI; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = Mfma<T>::step(fa[i], fb[j], acc[i][j]);
        }
        if (s + 1 < steps) stash(buf ^ 1);
        __syncthreads();
    }

#pragma unroll
    for (int i = 0; i < NI; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int64_t r = m0 + wm * RW + i * 16 + Mfma<T>::row(lane, e), col = n0 + wn * 64 + j * 16 + (lane & 15);
                if (store && (!EDGE || (r < m && col < n))) C[r * ldc + col] = acc[i][j][e];
            }
}

}  // namespace tile
}  // namespace hs
