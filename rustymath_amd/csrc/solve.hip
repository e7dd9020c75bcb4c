// Triangular solves with LU factors on gfx950: lu_solve (core/src/linalg/factor.rs:633-641) and the
// transposed solve that the adjoint needs.
//
// rustymath substitutes in three orders, and each fixes the last bits of x:
//   L y = P b           every row takes its terms in increasing column order (`forward` :668, `single` :725),
//   U x = y, one column  terms in decreasing column order, then a quotient by the diagonal (`single`),
//   U x = y, k columns   `backward` (:693): halving above NS = 32 rows, so a row takes the bottom half of
//                        each ancestor in increasing order, outermost ancestor first, then the rows of its
//                        own leaf after it in increasing order, then a reciprocal multiply (`divide` :506).
// The transposed solve takes Uᵀ in increasing order with `divide`'s scaling, then Lᵀ in decreasing order.
//
// One column (k = 1): one launch per pass, a one-wave block per 32 rows that adds earlier blocks' rows as
// they are published (trsv1_kernel). Several columns: 64-row blocks that solve among themselves in one
// workgroup (sweep_kernel for the increasing and the Lᵀ orders, bwd_node_kernel for a node of the halving),
// joined by in-order MFMA products that add a finished block's terms to the rows still to solve
// (gemm_skinny_kernel for k <= 16, the tiled product otherwise).
#include "hs_kernels.h"
#include <algorithm>
#include <atomic>
#include <mutex>

namespace hs {

namespace {

constexpr int BR = 32;  // NS of factor.rs:82: the halving's leaves have at most this many rows
constexpr uint64_t SPIN_LIMIT = 1ull << 22;

enum Mode : int {
    FWD_UNIT = 0,     // L (unit lower of F): increasing order, no division
    BWD_SINGLE = 1,   // U, one column: decreasing order, then a quotient
    BWD_LEAVES = 2,   // U, several columns: the halving order, then a reciprocal multiply
    FWD_DIAG = 3,     // Uᵀ (lower, diagonal of U): increasing order, then a reciprocal multiply
    BWD_UNIT = 4,     // Lᵀ (upper, unit): decreasing order, no division
    FWD_QUOT = 5,     // a lower triangle read by columns, one column: increasing order, then a quotient (`single`)
    BWD_COL_UNIT = 6, // a unit upper triangle read by columns, one column: decreasing order, no division (`single`)
};


template <typename T>
__global__ void gather_rows(int64_t n, int k, const T* __restrict__ b, int64_t ldb, const int64_t* __restrict__ perm, T* x, int64_t ldx) {
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * k) return;
    int64_t r = idx / k, c = idx % k;
    x[r * ldx + c] = b[perm[r] * ldb + c];
}

template <typename T>
__global__ void scatter_rows(int64_t n, int k, const T* __restrict__ w, int64_t ldw, const int64_t* __restrict__ perm, T* x, int64_t ldx) {
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n * k) return;
    int64_t r = idx / k, c = idx % k;
    x[perm[r] * ldx + c] = w[r * ldw + c];
}

// One right-hand side (the refinement's corrections, single-column lu_solve): one wave per 32-row block, in
// solve order. Lane t holds row r0 + t. A solved row is published as 8-byte words that carry the call's
// epoch beside the value (two words for FP64), stored with device-coherent (sc1) atomics for other XCDs and with
// plain stores for the next blocks on the same XCD, and loaded at device scope, so one load both tells that a row is
// ready and brings its value; the tile of matrix items a block needs from the next producer loads during the wait. A row adds the producers' rows in the reference's order, in
// increasing (L) or decreasing (U, `single`) column order, and the block then substitutes within itself
// with register broadcasts: a quotient by the diagonal at each row's turn for U.
// SCOPE is the store's: agent scope (sc1) for words that any XCD reads, wavefront scope (a plain store, which stays
// in the XCD's L2) for words that only workgroups of the writer's XCD read; readers load both kinds at agent scope,
// which reads the L2 of their own XCD.
template <typename T, int SCOPE = __HIP_MEMORY_SCOPE_AGENT>
__device__ __forceinline__ void publish1(unsigned long long* pub, int64_t row, T v, unsigned long long epoch) {
    if constexpr (sizeof(T) == 4) {
        const unsigned long long w = (epoch << 32) | __builtin_bit_cast(unsigned, v);
        __hip_atomic_store(&pub[row], w, __ATOMIC_RELAXED, SCOPE);
    } else {
        const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
        __hip_atomic_store(&pub[2 * row], (epoch << 32) | (u & 0xffffffffull), __ATOMIC_RELAXED, SCOPE);
        __hip_atomic_store(&pub[2 * row + 1], (epoch << 32) | (u >> 32), __ATOMIC_RELAXED, SCOPE);
    }
}
// A row's words are loaded (load1) and decoded (decode1) apart, so that other loads can be issued in between.
template <typename T>
__device__ __forceinline__ void load1(unsigned long long* pub, int64_t row, unsigned long long* w) {
    if constexpr (sizeof(T) == 4) {
        w[0] = __hip_atomic_load(&pub[row], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        w[0] = __hip_atomic_load(&pub[2 * row], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        w[1] = __hip_atomic_load(&pub[2 * row + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}
template <typename T>
__device__ __forceinline__ bool decode1(const unsigned long long* w, unsigned long long epoch, T& v) {
    if constexpr (sizeof(T) == 4) {
        v = __builtin_bit_cast(T, (unsigned)w[0]);
        return (w[0] >> 32) == epoch;
    } else {
        v = __builtin_bit_cast(T, (w[1] << 32) | (w[0] & 0xffffffffull));
        return (w[0] >> 32) == epoch && (w[1] >> 32) == epoch;
    }
}
template <typename T>
__device__ __forceinline__ T lane_value(T v, int l) {
    if constexpr (sizeof(T) == 8) {
        const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
        const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, l);
        const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
        return __builtin_bit_cast(T, ((unsigned long long)hi << 32) | lo);
    } else {
        return __builtin_bit_cast(T, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
    }
}

constexpr int B1 = 32;
constexpr int RUN1 = 16;  // consecutive blocks (in solve order) on one XCD

// The diagonal block of a single-column pass, rows in lanes 0..h-1 (H = 32: a full block, no row test; H = 0: h rows).
// Forward: increasing order, each row at its turn scaled (FWD_DIAG: `divide`, a reciprocal multiply unless the pivot
// is subnormal; FWD_QUOT: a quotient) and then subtracted from the rows after it; backward: decreasing order, BWD_SINGLE
// taking a quotient at each row's turn (`single`, factor.rs:725-748). Only the turn's lane takes the scaling, as a
// select, so the steps of the unit modes carry no division at all.
template <typename T, int H>
__device__ __forceinline__ T diag_steps(int mode, bool fwd, T acc, const T* d, T dg, int lane, int h = B1) {
    if (H == B1) h = B1;
    if (fwd) {
        if (mode == FWD_UNIT) {
#pragma unroll
            for (int j = 0; j < B1; ++j) {
                if (H == B1 || j < h) {
                    const T xj = lane_value(acc, j);
                    const T t = __builtin_elementwise_fma(xj, -d[j], acc);
                    acc = lane > j ? t : acc;
                }
            }
        } else {
            const T tiny = sizeof(T) == 8 ? T(2.2250738585072014e-308) : T(1.17549435e-38f);
            const T rdg = T(1) / dg;
            // Lanes whose turn takes a true quotient: every lane for FWD_QUOT, subnormal pivots for FWD_DIAG.
            const unsigned long long quot = mode == FWD_QUOT ? ~0ull : __ballot(!(fabs(dg) >= tiny));
#pragma unroll
            for (int j = 0; j < B1; ++j) {
                if (H == B1 || j < h) {
                    T v = acc * rdg;
                    if ((quot >> j) & 1ull) v = acc / dg;
                    acc = lane == j ? v : acc;
                    const T xj = lane_value(acc, j);
                    const T t = __builtin_elementwise_fma(xj, -d[j], acc);
                    acc = lane > j ? t : acc;
                }
            }
        }
    } else {
        if (mode == BWD_SINGLE) {
#pragma unroll
            for (int j = B1 - 1; j >= 0; --j) {
                if (H == B1 || j < h) {
                    const T v = acc / dg;
                    acc = lane == j ? v : acc;
                    const T xj = lane_value(acc, j);
                    const T t = __builtin_elementwise_fma(xj, -d[j], acc);
                    acc = lane < j ? t : acc;
                }
            }
        } else {
#pragma unroll
            for (int j = B1 - 1; j >= 0; --j) {
                if (H == B1 || j < h) {
                    const T xj = lane_value(acc, j);
                    const T t = __builtin_elementwise_fma(xj, -d[j], acc);
                    acc = lane < j ? t : acc;
                }
            }
        }
    }
    return acc;
}

// One column, one pass: a one-wave block per 32 rows (lane i holds row r0 + i), blocks numbered in solve order. A block
// adds each earlier block's rows as soon as they are published (tagged words, so a stale word never matches), its
// tiles of the matrix loaded one producer ahead, then substitutes within itself and publishes. Orders: FWD_UNIT (L,
// increasing), BWD_SINGLE (U, decreasing, then a quotient: `single`, factor.rs:725), and for the transposed solve
// FWD_DIAG (Uᵀ, increasing, then `divide`'s reciprocal multiply) and BWD_UNIT (Lᵀ, decreasing). The transposed modes
// read rows of F where the others read columns. FWD_QUOT and BWD_COL_UNIT are `single`'s two other cases of a
// triangle read by columns (factor.rs:730-741), for the triangular-solve entry points.
//
// Launch index w runs on XCD w % 8, and takes solve-order block u, so that runs of `run` consecutive blocks share an
// XCD: a block publishes its rows twice, to `loc` with plain stores that the next blocks of its run read from the
// XCD's L2, and to `pub` at agent scope for the other XCDs. With run > 1 a block may start before blocks it waits on,
// so the host takes run > 1 only for grids that fit on the GPU at once; run = 1 is the launch order.
template <typename T>
__global__ __launch_bounds__(64) void trsv1_kernel(int mode, int64_t n, const T* __restrict__ f, int64_t ld, T* x,
                                                   unsigned long long* pub, unsigned long long* loc, int run,
                                                   unsigned long long epoch, int* status) {
    const int nb = (int)((n + B1 - 1) / B1), lane = threadIdx.x;
    const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
    const int u = ((slot / run) * 8 + xcd) * run + slot % run;  // solve order; a bijection of the padded grid
    if (u >= nb) return;
    // After a wait that ran out (here or in the first pass) the result is void: return at once rather than wait again.
    if (__hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    const bool fwd = mode == FWD_UNIT || mode == FWD_DIAG || mode == FWD_QUOT, rowwise = mode == FWD_DIAG || mode == BWD_UNIT;
    const int q = fwd ? u : nb - 1 - u;  // this block's rows [q B1, ...)
    const int64_t r0 = (int64_t)q * B1;
    const int h = (int)(n - r0 < B1 ? n - r0 : B1);
    // FP32: producers two to a unit, tiles loaded by all 64 lanes (a block that trails the wavefront takes two
    // producers per wait on its tiles' loads); FP64: one producer per step, which measured faster for its larger
    // registers (order 16384: two passes 2.71 against 2.94 ms).
    T acc;
    const bool act = lane < h;
    const int64_t i = r0 + (act ? lane : 0);
    T ddv[B1];
    if constexpr (sizeof(T) == 4) {
        const int li = lane & 31, half = lane >> 5;
        const int64_t irow = r0 + (li < h ? li : 0);  // the row whose items this lane loads (both halves of the wave)
        acc = act ? x[i] : T(0);
        // Tiles: lane l holds entry (irow, pq B1 + 16 (l / 32) + j) in e[j], j < 16, which is F[pq B1 + ...][irow] or,
        // row-wise, F[irow][pq B1 + ...]: all 64 lanes load, half the loads per lane of a 32-lane tile. Columns past n read
        // column n - 1 here and are zeroed where the tile is used, so that no instruction touches the loaded values before.
        constexpr int HB = B1 / 2;
        auto load_tile = [&](int pq, T* e) {
            const int64_t p0 = (int64_t)pq * B1 + HB * half;
    #pragma unroll
            for (int j = 0; j < HB; ++j) {
                const int64_t pc = p0 + j < n ? p0 + j : n - 1;
                e[j] = rowwise ? f[irow * ld + pc] : f[pc * ld + irow];
            }
        };
        // A lane's row items HB .. B1 - 1 come from the lane 32 above it: v_permlane32_swap's second result holds, in lanes
        // 0-31, the source's lanes 32-63 (checked on gfx950).
        auto up32 = [](unsigned v) -> unsigned { return __builtin_amdgcn_permlane32_swap(v, v, false, false)[1]; };
        auto full_row = [&](const T* e, T* ev) {
    #pragma unroll
            for (int j = 0; j < HB; ++j) {
                ev[j] = e[j];
                if constexpr (sizeof(T) == 4) {
                    ev[HB + j] = __builtin_bit_cast(T, up32(__builtin_bit_cast(unsigned, e[j])));
                } else {
                    const unsigned long long u = __builtin_bit_cast(unsigned long long, e[j]);
                    const unsigned long long lo = up32((unsigned)u), hh = up32((unsigned)(u >> 32));
                    ev[HB + j] = __builtin_bit_cast(T, (hh << 32) | lo);
                }
            }
        };
        // Producers in solve order t = 0 .. np - 1, two to a unit: one look at both producers' rows (lanes 0-31 the first,
        // 32-63 the second), the next unit's two tiles issued behind it (none after the last unit: the load counter is in
        // order, so a poll behind them would wait for them), then the first producer's terms as soon as its rows are all
        // published, and the second's once its rows are. A block that trails the wavefront takes two producers per wait
        // on its tiles' loads.
        const int np = fwd ? q : nb - 1 - q;
        auto tile_of = [&](int t) {
            const int c = t < np ? t : np - 1;
            return fwd ? c : nb - 1 - c;
        };
        T d[HB];
        load_tile(q, d);  // the diagonal tile (its columns past n are never used), off the critical path
        auto terms = [&](T xv, int base, int ph, const T* e) {
            T ev[B1];
            full_row(e, ev);
            // Rows past the producer's end hold +0 and their items are zeroed: their terms add -0, which changes no sum.
            if (ph < B1) {
    #pragma unroll
                for (int j = 0; j < B1; ++j) if (j >= ph) ev[j] = T(0);
            }
            if (fwd) {
    #pragma unroll
                for (int j = 0; j < B1; ++j) acc = __builtin_elementwise_fma(lane_value(xv, base + j), -ev[j], acc);
            } else {
    #pragma unroll
                for (int j = B1 - 1; j >= 0; --j) acc = __builtin_elementwise_fma(lane_value(xv, base + j), -ev[j], acc);
            }
        };
        auto rows_of = [&](int t) { const int64_t p0 = (int64_t)tile_of(t) * B1; return (int)(n - p0 < B1 ? n - p0 : B1); };
        auto unit = [&](int t0, const T* e0, const T* e1, T* n0, T* n1) -> bool {
            const int t = t0 + half, tc = t < np ? t : np - 1;  // this half's producer (the first again past the end)
            const int pq = tile_of(tc);
            const int64_t p0 = (int64_t)pq * B1;
            const int ph = (int)(n - p0 < B1 ? n - p0 : B1);
            const int64_t prow = p0 + (li < ph ? li : 0);
            unsigned long long* src = ((tc / run) & 7) == xcd ? loc : pub;  // producer tc (solve order) on this XCD?
            T xv = T(0);
            bool ok = false;
            unsigned long long w[2];
            auto take = [&]() {
                T v;
                ok = li >= ph || decode1(w, epoch, v);
                xv = li < ph ? v : T(0);
            };
            load1<T>(src, prow, w);
            __builtin_amdgcn_sched_barrier(0);
            if (t0 + 2 < np) {
                load_tile(tile_of(t0 + 2), n0);
                load_tile(tile_of(t0 + 3), n1);  // past the last producer: that producer's tile again, never used
            }
            __builtin_amdgcn_sched_barrier(0);
            take();
            // Poll until the lanes in `mask` hold published words (a published word keeps its value for the pass).
            auto wait = [&](unsigned long long mask) -> bool {
                if ((__ballot(ok) & mask) == mask) return true;  // outside the loop: the first look waits for its own load
                uint64_t spins = 0;
                do {
                    __builtin_amdgcn_s_sleep(2);  // many blocks wait on the same producer: poll its words gently
                    if (++spins > SPIN_LIMIT) { if (lane == 0) atomicExch(status, ERR_LIMIT); return false; }
                    load1<T>(src, prow, w);
                    take();
                } while ((__ballot(ok) & mask) != mask);
                return true;
            };
            if (!wait(0xffffffffull)) return false;
            terms(xv, 0, rows_of(t0), e0);
            if (t0 + 1 < np) {
                if (!wait(~0ull)) return false;
                terms(xv, 32, rows_of(t0 + 1), e1);
            }
            return true;
        };
        if (np > 0) {
            T ta[2][HB], tb[2][HB];
            load_tile(tile_of(0), ta[0]);
            load_tile(tile_of(1), ta[1]);
            // Units unrolled by two, so that each pair of tiles keeps its registers.
            for (int t0 = 0;; t0 += 4) {
                if (!unit(t0, ta[0], ta[1], tb[0], tb[1])) return;
                if (t0 + 2 >= np) break;
                if (!unit(t0 + 2, tb[0], tb[1], ta[0], ta[1])) return;
                if (t0 + 4 >= np) break;
            }
        }
        full_row(d, ddv);

    } else {
        acc = act ? x[i] : T(0);
        // Tile of producer block pq: e[j] = entry (i, pq B1 + j), which is F[pq B1 + j][i] or, row-wise, F[i][pq B1 + j].
        // Columns past n read column n - 1 here and are zeroed where the tile is used, so that no instruction touches
        // the loaded values before then.
        auto load_tile = [&](int pq, T* e) {
            const int64_t p0 = (int64_t)pq * B1;
    #pragma unroll
            for (int j = 0; j < B1; ++j) {
                const int64_t pc = p0 + j < n ? p0 + j : n - 1;
                e[j] = rowwise ? f[i * ld + pc] : f[pc * ld + i];
            }
        };
        // Producers in solve order t = 0 .. np - 1, two tiles in registers: the one in use and the next one loading.
        const int np = fwd ? q : nb - 1 - q;
        auto tile_of = [&](int t) {
            const int c = t < np ? t : np - 1;
            return fwd ? c : nb - 1 - c;
        };
        T tl[2][B1], d[B1];
        load_tile(q, d);  // the diagonal tile (its columns past n are never used), off the critical path
        // One producer: look at its rows once, issue the loads of tile t + 1, then take the look (scheduling
        // barriers keep that order, so that the wait for the look leaves the tile's loads in flight), poll until every
        // row is published, and add the terms. Every lane reads a word (lanes past the producer's rows read its first
        // row and keep +0); a published word keeps its value for the rest of the pass.
        auto step = [&](int t, T* e, T* next) -> bool {
            const int pq = tile_of(t);
            const int64_t p0 = (int64_t)pq * B1;
            const int ph = (int)(n - p0 < B1 ? n - p0 : B1);
            const int64_t prow = p0 + (lane < ph ? lane : 0);
            unsigned long long* src = ((t / run) & 7) == xcd ? loc : pub;  // producer t (solve order) on this XCD?
            T xv = T(0);
            bool ok = false;
            unsigned long long w[2];
            auto take = [&]() {
                T v;
                ok = lane >= ph || decode1(w, epoch, v);
                xv = lane < ph ? v : T(0);
            };
            load1<T>(src, prow, w);
            __builtin_amdgcn_sched_barrier(0);
            // The next producer's tile; none at the last producer, whose wait is the critical one: the load counter is
            // in order, so a poll issued behind 32 tile loads would see the producer's rows only once they arrive.
            if (t + 1 < np) load_tile(tile_of(t + 1), next);
            __builtin_amdgcn_sched_barrier(0);
            take();
            if (!__all(ok)) {  // outside the polling loop, so that the first look waits for its own load alone
                uint64_t spins = 0;
                do {
                    __builtin_amdgcn_s_sleep(2);  // many blocks wait on the same producer: poll its words gently
                    if (++spins > SPIN_LIMIT) { if (lane == 0) atomicExch(status, ERR_LIMIT); return false; }
                    load1<T>(src, prow, w);
                    take();
                } while (!__all(ok));
            }
            // Rows past the producer's end hold +0 and their items are zeroed: their terms add -0, which changes no sum.
            if (ph < B1) {
    #pragma unroll
                for (int j = 0; j < B1; ++j) if (j >= ph) e[j] = T(0);
            }
            if (fwd) {
    #pragma unroll
                for (int j = 0; j < B1; ++j) acc = __builtin_elementwise_fma(lane_value(xv, j), -e[j], acc);
            } else {
    #pragma unroll
                for (int j = B1 - 1; j >= 0; --j) acc = __builtin_elementwise_fma(lane_value(xv, j), -e[j], acc);
            }
            return true;
        };
        if (np > 0) {
            load_tile(tile_of(0), tl[0]);
            if constexpr (sizeof(T) == 8) {
                // Unrolled by two, so that each tile keeps its registers (a copy would wait for the loads), and left after
                // the last producer (a step past it would only add latency): 3.1 -> 2.8 ms per solve at order 16384.
                for (int t0 = 0;; t0 += 2) {
                    if (!step(t0, tl[0], tl[1])) return;
                    if (t0 + 1 >= np) break;
                    if (!step(t0 + 1, tl[1], tl[0])) return;
                    if (t0 + 2 >= np) break;
                }
            } else {
                // FP32 is faster with the copy (2.48 ms against 2.55 ms unrolled at order 16384).
                for (int t = 0; t < np; ++t) {
                    if (!step(t, tl[0], tl[1])) return;
    #pragma unroll
                    for (int j = 0; j < B1; ++j) tl[0][j] = tl[1][j];
                }
            }
        }
    #pragma unroll
        for (int j = 0; j < B1; ++j) ddv[j] = d[j];

    }
    const T* dd = ddv;
    // Within the block: a straight run of 32 dependent steps per mode (the mode and a full block are uniform tests
    // outside the steps; only the block holding row n - 1 has h < 32), each step a register broadcast of the row
    // just finished and one fused multiply-add in the rows after it.
    T dg = T(1);
#pragma unroll
    for (int j = 0; j < B1; ++j) if (j == lane) dg = dd[j];
    if (h == B1) acc = diag_steps<T, B1>(mode, fwd, acc, dd, dg, lane);
    else acc = diag_steps<T, 0>(mode, fwd, acc, dd, dg, lane, h);
    if (act) {
        publish1<T, __HIP_MEMORY_SCOPE_WAVEFRONT>(loc, i, acc, epoch);
        publish1(pub, i, acc, epoch);
        x[i] = acc;
    }
}

// The epoch of each single-column pass: increasing across calls, so that words left by an earlier pass never
// match (the published words are also cleared before each pass).
std::atomic<unsigned long long> g_epoch1{1};

__global__ void zero_words(unsigned long long* p, int64_t count) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

// Bytes of the single-column passes' workspace for order n: the published words twice (agent scope, and the XCD's
// L2) and the status word.
template <typename T>
size_t trsv1_bytes(int64_t n) { return 2 * (size_t)n * (sizeof(T) == 8 ? 2 : 1) * 8 + 64; }

// L then U (lu_solve's order for one column) in place on x (n items, contiguous), or Uᵀ then Lᵀ (trans), with one
// status check. `buf` (trsv1_bytes(n) bytes of device memory) is the workspace, or null for one from the pool. The
// words are cleared by a fill kernel, so the sequence holds no memset node when a caller captures it.
template <typename T>
int trsv1_lu(bool trans, int64_t n, const T* f, int64_t ld, T* x, hipStream_t s, char* buf = nullptr) {
    const int nb = (int)((n + B1 - 1) / B1);
    // Runs of RUN1 blocks per XCD when the padded grid fits on the GPU at once (one wave per block, at least two per
    // CU at this kernel's register use), else the launch order, in which no block waits on one launched after it.
    const int padded = (nb + 8 * RUN1 - 1) / (8 * RUN1) * (8 * RUN1);
    // Runs of RUN1 blocks per XCD need the probed placement (launch index w on XCD w mod 8); a caller that brings
    // its own workspace (no allocation) never triggers the probe and takes runs of one until one has run.
    const int run = padded <= 2 * device_cus() && exchange_local(buf == nullptr) ? RUN1 : 1;
    const int grid = run > 1 ? padded : (nb + 7) / 8 * 8;
    const size_t words = (size_t)n * (sizeof(T) == 8 ? 2 : 1);
    const bool pooled = buf == nullptr;
    if (pooled && !(buf = (char*)pool_get(trsv1_bytes<T>(n)))) return hip_status(hipErrorOutOfMemory, "pool_get");
    unsigned long long* pub = (unsigned long long*)buf;
    unsigned long long* loc = pub + words;
    int* status = (int*)(buf + 2 * words * 8);
    // With runs, a block waits on blocks that must be resident beside it: one such solve at a time in this process
    // (a concurrent one could hold the CUs its blocks need).
    std::unique_lock<std::mutex> lk(resident_mutex(), std::defer_lock);
    if (run > 1) lk.lock();
    zero_words<<<(unsigned)std::min<int64_t>((2 * words + 8 + 255) / 256, 1024), 256, 0, s>>>(
        (unsigned long long*)buf, (int64_t)(2 * words + 8));
    HS_HIP(hipGetLastError());
    const unsigned long long e1 = (g_epoch1.fetch_add(2) % 0xfffffff0ull) | 1ull;  // odd, and e1 + 1 < 2^32
    trsv1_kernel<T><<<grid, 64, 0, s>>>(trans ? FWD_DIAG : FWD_UNIT, n, f, ld, x, pub, loc, run, e1, status);
    HS_HIP(hipGetLastError());
    trsv1_kernel<T><<<grid, 64, 0, s>>>(trans ? BWD_UNIT : BWD_SINGLE, n, f, ld, x, pub, loc, run, e1 + 1, status);
    HS_HIP(hipGetLastError());
    int st = 0;
    HS_HIP(hipMemcpyAsync(&st, status, sizeof(int), hipMemcpyDeviceToHost, s));
    HS_HIP(hipStreamSynchronize(s));
    if (pooled) pool_put(buf);
    if (st) { set_error("triangular solve wait ran out"); return ERR_LIMIT; }
    return OK;
}

constexpr int BS = 64;  // rows of a block solve

// The block's triangle [r0, r0 + m)², m <= 64, into LDS with coalesced loads: E[i][q] = entry (r0 + i, r0 + q), which
// is f[i ld + q] for the transposed modes (rows of F) and f[q ld + i] otherwise; zero outside the block.
template <typename T, int NT>
__device__ __forceinline__ void stage_triangle(bool tmode, int64_t r0, int m, const T* __restrict__ f, int64_t ld,
                                               T (*E)[BS + 1]) {
    const int t = threadIdx.x;
    T ev[BS * BS / NT];
#pragma unroll
    for (int e = 0; e < BS * BS / NT; ++e) {
        const int idx = t + NT * e, i = tmode ? idx / BS : idx % BS, q = tmode ? idx % BS : idx / BS;
        const int64_t ii = r0 + (i < m ? i : m - 1), qq = r0 + (q < m ? q : m - 1);
        ev[e] = tmode ? f[ii * ld + qq] : f[qq * ld + ii];
    }
    asm volatile("" ::: "memory");
#pragma unroll
    for (int e = 0; e < BS * BS / NT; ++e) {
        const int idx = t + NT * e, i = tmode ? idx / BS : idx % BS, q = tmode ? idx % BS : idx / BS;
        E[i][q] = (i < m && q < m) ? ev[e] : T(0);
    }
}

template <typename T>
__device__ __forceinline__ T smallest_normal() {
    return sizeof(T) == 8 ? T(2.2250738585072014e-308) : T(1.17549435e-38f);
}

// A block of at most 64 rows [r0, r0 + m) solving among themselves in a right-looking sweep, k <= 64 columns of x:
// lane i holds row r0 + i and its row of the triangle in registers; at step t row t is final, its value goes to every
// lane by readlane, and the rows it feeds add its term. FWD_UNIT and FWD_DIAG step t = 0, 1, ... (each row's terms in
// increasing order; FWD_DIAG scales a row by its diagonal as `divide` does, a reciprocal multiply or, for a subnormal
// diagonal, a quotient, once its terms are in); BWD_UNIT steps from the bottom (terms in decreasing order). Eight
// waves, two columns each at a time, so that two chains overlap in every wave.
constexpr int SW_WAVES = 8, SW_NC = 2;
template <typename T, int MODE>
__global__ __launch_bounds__(64 * SW_WAVES) void sweep_kernel(int64_t r0, int m, int64_t ktot, const T* __restrict__ f,
                                                              int64_t ld, T* x, int64_t ldx) {
    __shared__ T E[BS][BS + 1];
    x += (int64_t)blockIdx.x * BS;  // this workgroup's 64 columns
    const int k = (int)std::min<int64_t>(BS, ktot - (int64_t)blockIdx.x * BS);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    stage_triangle<T, 64 * SW_WAVES>(MODE != FWD_UNIT, r0, m, f, ld, E);
    __syncthreads();
    T e[BS];
#pragma unroll
    for (int q = 0; q < BS; ++q) e[q] = E[lane][q];
    const T d = E[lane][lane < m ? lane : 0];
    const int nrm = fabs(d) >= smallest_normal<T>();
    const T ri = T(1) / d;
    auto scale = [&](T v) { return v * ri; };
    for (int c0 = wave * SW_NC; c0 < k; c0 += SW_WAVES * SW_NC) {
        int me = lane;
        asm volatile("" : "+v"(me));  // the step masks come from this copy, so they are not all kept across passes
        T xv[SW_NC];
#pragma unroll
        for (int j = 0; j < SW_NC; ++j) {
            const int c = c0 + j < k ? c0 + j : k - 1;
            const T v = x[(r0 + (lane < m ? lane : m - 1)) * ldx + c];
            xv[j] = lane < m ? v : T(0);
        }
#pragma unroll
        for (int step = 0; step < BS; ++step) {
            const int tt = MODE == BWD_UNIT ? BS - 1 - step : step;  // rows past m hold zeros and add nothing
            T bc[SW_NC];
#pragma unroll
            for (int j = 0; j < SW_NC; ++j) {
                T v = xv[j];
                if (MODE == FWD_DIAG) v = __builtin_amdgcn_readlane(nrm, tt) ? scale(v) : v / d;
                bc[j] = lane_value(v, tt);
            }
            if (MODE == BWD_UNIT ? me < tt : me > tt) {
#pragma unroll
                for (int j = 0; j < SW_NC; ++j) xv[j] = __builtin_elementwise_fma(bc[j], -e[tt], xv[j]);
            }
        }
#pragma unroll
        for (int j = 0; j < SW_NC; ++j) {
            T v = xv[j];
            if (MODE == FWD_DIAG) v = nrm ? scale(v) : v / d;
            if (lane < m && c0 + j < k) x[(r0 + lane) * ldx + c0 + j] = v;
        }
    }
}

// A node of `backward`'s halving with at most 64 rows (BWD_LEAVES). In a leaf (at most 32 rows) a row's terms come
// in increasing order but the rows after it finish from the bottom up, so each row waits for all of them: the chains
// are serial, one per column. Wave 0 runs them with lane c for column c (k <= 64) and the leaf's rows of x in
// registers, the last row in the last slot, so that every row's terms are exactly the slots after it; the triangle's
// items come from LDS (one address for all lanes) and the reciprocals of the diagonal are ready beforehand. The slots
// before the leaf's first row compute rows that do not exist (the LDS before the triangle is a guard region for
// them), are never stored and feed no real row, so the leaf is one straight run of code. A leaf with a subnormal
// diagonal, whose rows take a quotient, runs a plain loop instead. Above 32 rows (factor.rs:693-707) the bottom leaf
// goes first, then all four waves add the bottom rows' terms to the top rows in increasing order, then the top leaf.
template <typename T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) void bwd_node_kernel(
    int64_t r0, int m, int64_t ktot, const T* __restrict__ f, int64_t ld, T* x, int64_t ldx) {
    x += (int64_t)blockIdx.x * BS;  // this workgroup's 64 columns
    const int k = (int)std::min<int64_t>(BS, ktot - (int64_t)blockIdx.x * BS);
    constexpr int GUARD = BR * (BS + 1);  // covers rows and columns down to -31
    __shared__ T Eg[GUARD + BS * (BS + 1)];
    __shared__ T xb[BS][BS + 1];  // xb[row][column]
    __shared__ T rdg[BR + BS];    // rdg[BR + i] = 1 / U(r0 + i, r0 + i)
    T(*E)[BS + 1] = reinterpret_cast<T(*)[BS + 1]>(Eg + GUARD);  // E[i][q] = U(r0 + i, r0 + q) = f[(r0 + q) ld + r0 + i]
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    stage_triangle<T, 256>(false, r0, m, f, ld, E);
    for (int idx = t; idx < BS * BS; idx += 256) {
        const int r = idx / BS, cc = idx % BS;
        xb[r][cc] = (r < m && cc < k) ? x[(r0 + r) * ldx + cc] : T(0);
    }
    __syncthreads();
    if (t < BS) rdg[BR + t] = T(1) / E[t][t];
    __syncthreads();
    const T tiny = smallest_normal<T>();
    const int split = m > BR ? (int)((2 * r0 + m) / 2 - r0) : 0;  // the node's middle, (r0 + r1) / 2, less r0
    for (int phase = 0; phase < 2; ++phase) {  // the bottom leaf (or the only one), then the top leaf
        const int lo = phase == 0 ? split : 0, hi = phase == 0 ? m : split;
        if (phase == 1) {
            if (split == 0) break;
            T acc[BR / 4];  // rows wave, wave + 4, ... of the top half (split <= 32)
#pragma unroll
            for (int j = 0; j < BR / 4; ++j) acc[j] = xb[wave + 4 * j][lane];
            for (int q = split; q < m; ++q) {
                const T xq = xb[q][lane];
#pragma unroll
                for (int j = 0; j < BR / 4; ++j) acc[j] = __builtin_elementwise_fma(xq, -E[wave + 4 * j][q], acc[j]);
            }
#pragma unroll
            for (int j = 0; j < BR / 4; ++j)
                if (wave + 4 * j < split) xb[wave + 4 * j][lane] = acc[j];
            __syncthreads();
        }
        if (wave == 0) {
            const int dl = lo + (lane < hi - lo ? lane : 0);
            if (__all(fabs(E[dl][dl]) >= tiny)) {
                const int off = BR - (hi - lo), base = lo - off;  // slot u holds row base + u
                T xr[BR];
#pragma unroll
                for (int u = 0; u < BR; ++u) xr[u] = xb[u >= off ? base + u : lo][lane];
                // One address register for the triangle and one for the reciprocals, the slots as fixed offsets.
                int eb = GUARD + base * (BS + 1) + base, rb = BR + base;
                asm volatile("" : "+v"(eb), "+v"(rb));
                // Each row's items are loaded while the row below it runs its chain.
                T ec[BR], en[BR], rc = rdg[rb + BR - 1], rn = rc;
#pragma unroll
                for (int u = BR - 1; u >= 0; --u) {
                    if (u > 0) {
#pragma unroll
                        for (int q = u; q < BR; ++q) en[q] = Eg[eb + (u - 1) * (BS + 1) + q];
                        rn = rdg[rb + u - 1];
                    }
                    __builtin_amdgcn_sched_barrier(0);  // keep those loads ahead of this row's chain
                    T acc = xr[u];
#pragma unroll
                    for (int q = u + 1; q < BR; ++q) acc = __builtin_elementwise_fma(xr[q], -ec[q], acc);
                    xr[u] = acc * rc;
#pragma unroll
                    for (int q = 0; q < BR; ++q) ec[q] = en[q];
                    rc = rn;
                }
#pragma unroll
                for (int u = 0; u < BR; ++u)
                    if (u >= off) xb[base + u][lane] = xr[u];
            } else {
                for (int row = hi - 1; row >= lo; --row) {
                    T acc = xb[row][lane];
                    for (int q = row + 1; q < hi; ++q) acc = __builtin_elementwise_fma(xb[q][lane], -E[row][q], acc);
                    const T dg = E[row][row];
                    xb[row][lane] = fabs(dg) >= tiny ? acc * rdg[BR + row] : acc / dg;
                }
            }
        }
        __syncthreads();
    }
    for (int idx = t; idx < m * BS; idx += 256) {
        const int r = idx / BS, cc = idx % BS;
        if (cc < k) x[(r0 + r) * ldx + cc] = xb[r][cc];
    }
}

// c -= A b for the substitutions' products: b and c have at most 16 columns, so one wave owns a strip of 16 rows of c
// (one MFMA accumulator) and takes the whole of k in order, 4 terms per 16x16x4 step: increasing p, or decreasing p
// for REV. The terms come in stages of 4 D, the next stage's operands in flight while the current one runs; whole
// stages step plain pointers, the last partial stage clamps and pads. A(i, p) is a[p lda + i] (ta) or a[i lda + p].
// Rows and columns past the ends read the last row or column (their results are dropped); terms past k read
// (+0)(-0) = -0, which changes no sum.
template <typename T, bool TA, bool REV, int D>
__global__ __launch_bounds__(256) void gemm_skinny_kernel(int64_t m, int nc, int64_t k, const T* __restrict__ a,
                                                          int64_t lda, const T* __restrict__ b, int64_t ldb, T* c,
                                                          int64_t ldc) {
    typedef typename Mfma<T>::acc Acc;
    const int lane = threadIdx.x & 63;
    const int64_t i0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
    if (i0 >= m) return;
    const int li = lane & 15, lk = lane >> 4;
    Acc acc;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int64_t r = i0 + Mfma<T>::row(lane, e);
        acc[e] = c[(r < m ? r : m - 1) * ldc + (li < nc ? li : nc - 1)];
    }
    const int64_t ai = i0 + li < m ? i0 + li : m - 1;
    const int bc = li < nc ? li : nc - 1;
    // Operand indices of this lane's term t (p = t, or k - 1 - t for REV), and their steps per MFMA.
    const int64_t p_first = REV ? k - 1 - lk : lk;
    const int64_t ia0 = TA ? p_first * lda + ai : ai * lda + p_first, ib0 = p_first * ldb + bc;
    const int64_t sa = (REV ? -4 : 4) * (TA ? lda : 1), sb = (REV ? -4 : 4) * ldb;
    const int64_t full = k / (4 * D) * (4 * D);  // terms in whole stages
    auto load = [&](int64_t t0, T(&av)[D], T(&bv)[D]) {  // a whole stage from term t0
        int64_t ia = ia0 + (t0 / 4) * sa, ib = ib0 + (t0 / 4) * sb;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            av[d] = TA ? a[ia] : a[ia + d * sa];  // rows of A: one add a step; columns: fixed offsets
            bv[d] = b[ib];
            if (TA) ia += sa;
            ib += sb;
            // One running index per operand, rather than a register pair for every step's address.
            asm volatile("" : "+v"(ia), "+v"(ib));
        }
    };
    auto run = [&](const T(&av)[D], const T(&bv)[D]) {
        // The stage's operands pass through an empty asm here, so that their negations are not moved up to the
        // loads (where they would wait for the loads just issued).
        T x[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            x[d] = av[d];
            asm volatile("" : "+v"(x[d]));
        }
#pragma unroll
        for (int d = 0; d < D; ++d) acc = Mfma<T>::step(-x[d], bv[d], acc);
    };
    if (full > 0) {
        // Pairs of stages in one straight block; every prefetch is issued (past the last whole stage it reloads that
        // stage), so that each wait is for the older stage only (2 D loads a stage stay under the 63 that a wait
        // can count).
        const int64_t last = full - 4 * D;
        T a0[D], b0[D], a1[D], b1[D];
        load(0, a0, b0);
        int64_t t0 = 0;
        for (; t0 + 8 * D <= full; t0 += 8 * D) {
            load(t0 + 4 * D, a1, b1);
            run(a0, b0);
            load(t0 + 8 * D < last ? t0 + 8 * D : last, a0, b0);
            run(a1, b1);
        }
        if (t0 < full) run(a0, b0);
    }
    if (full < k) {  // the partial stage
        T av[D], bv[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int64_t t = full + 4 * d + lk, tc = t < k ? t : k - 1, p = REV ? k - 1 - tc : tc;
            av[d] = TA ? a[p * lda + ai] : a[ai * lda + p];
            bv[d] = b[p * ldb + bc];
        }
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const bool in = full + 4 * d + lk < k;
            acc = Mfma<T>::step(in ? -av[d] : T(0), in ? bv[d] : T(-0.0), acc);
        }
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int64_t r = i0 + Mfma<T>::row(lane, e);
        if (r < m && li < nc) c[r * ldc + li] = acc[e];
    }
}

// c (m x nc, nc <= 16) -= A b in order of the terms (decreasing p when rev), a 4-wave workgroup per 64 rows of c.
template <typename T>
int gemm_skinny(bool ta, bool rev, int64_t m, int64_t nc, int64_t k, const T* a, int64_t lda, const T* b, int64_t ldb,
                T* c, int64_t ldc, hipStream_t s) {
    if (m == 0 || nc == 0 || k == 0) return OK;
    const unsigned grid = (unsigned)((m + 63) / 64);
#define HS_SKINNY(TA, REV, D) gemm_skinny_kernel<T, TA, REV, D><<<grid, 256, 0, s>>>(m, (int)nc, k, a, lda, b, ldb, c, ldc)
    if (k <= 64) {
        if (ta) rev ? HS_SKINNY(true, true, 16) : HS_SKINNY(true, false, 16);
        else rev ? HS_SKINNY(false, true, 16) : HS_SKINNY(false, false, 16);
    } else {
        if (ta) rev ? HS_SKINNY(true, true, 24) : HS_SKINNY(true, false, 24);
        else rev ? HS_SKINNY(false, true, 24) : HS_SKINNY(false, false, 24);
    }
#undef HS_SKINNY
    HS_HIP(hipGetLastError());
    return OK;
}

// Substitutions over many rows as blocks of rows that solve among themselves (block_solve_kernel) and MFMA products
// that add a finished block's terms to the rows still to solve, in order of the terms (in-order products: every item
// takes one fused multiply-add per term, in the order the substitution would).
template <typename T>
int block_solve(int mode, int64_t r0, int64_t m, int64_t k, const T* lu, int64_t lda, T* w, int64_t ldw, hipStream_t s) {
    const unsigned grid = (unsigned)((k + BS - 1) / BS);  // a workgroup per 64 columns
    if (mode == BWD_LEAVES) bwd_node_kernel<T><<<grid, 256, 0, s>>>(r0, (int)m, k, lu, lda, w, ldw);
    else if (mode == FWD_UNIT) sweep_kernel<T, FWD_UNIT><<<grid, 64 * SW_WAVES, 0, s>>>(r0, (int)m, k, lu, lda, w, ldw);
    else if (mode == FWD_DIAG) sweep_kernel<T, FWD_DIAG><<<grid, 64 * SW_WAVES, 0, s>>>(r0, (int)m, k, lu, lda, w, ldw);
    else sweep_kernel<T, BWD_UNIT><<<grid, 64 * SW_WAVES, 0, s>>>(r0, (int)m, k, lu, lda, w, ldw);
    HS_HIP(hipGetLastError());
    return OK;
}

template <typename T>
int fwd_blocked(int mode, int64_t n, int64_t k, const T* lu, int64_t lda, T* w, hipStream_t s) {
    for (int64_t lo = 0; lo < n; lo += BS) {
        const int64_t hi = std::min<int64_t>(n, lo + BS);
        const int rc = block_solve<T>(mode, lo, hi - lo, k, lu, lda, w, k, s);
        if (rc) return rc;
        if (hi < n) {
            // Rows [hi, n) take rows [lo, hi): entry (i, q) is L(i, q) = F[q][i] (FWD_UNIT) or Uᵀ(i, q) = F[i][q].
            const bool ta = mode == FWD_UNIT;
            const T* a = ta ? lu + lo * lda + hi : lu + hi * lda + lo;
            const int rc = k <= 16 ? gemm_skinny<T>(ta, false, n - hi, k, hi - lo, a, lda, w + lo * k, k, w + hi * k, k, s)
                                   : gemm<T>(ta, false, n - hi, k, hi - lo, T(-1), a, lda, w + lo * k, k, T(1), w + hi * k,
                                             k, s, true);
            if (rc) return rc;
        }
    }
    return OK;
}

// Lᵀ x = z from the bottom (terms in decreasing order): 64-row blocks from the end up, each solving among itself, then
// the rows above it take its rows' terms, last row first.
template <typename T>
int bwd_unit_blocked(int64_t n, int64_t k, const T* lu, int64_t lda, T* w, hipStream_t s) {
    for (int64_t hi = n; hi > 0; hi -= BS) {
        const int64_t lo = std::max<int64_t>(0, hi - BS);
        int rc = block_solve<T>(BWD_UNIT, lo, hi - lo, k, lu, lda, w, k, s);
        // Rows [0, lo) take rows [lo, hi): entry (i, q) is Lᵀ(i, lo + q) = F[i][lo + q].
        for (int64_t c0 = 0; c0 < k && lo > 0 && rc == OK; c0 += 16)
            rc = gemm_skinny<T>(false, true, lo, std::min<int64_t>(16, k - c0), hi - lo, lu + lo, lda, w + lo * k + c0, k,
                                w + c0, k, s);
        if (rc) return rc;
    }
    return OK;
}

// `backward`'s halving (factor.rs:693-707): the bottom half first, then the top half takes the bottom half's terms in
// increasing order, then the top half; nodes of at most 64 rows go to one block solve.
template <typename T>
int bwd_halving(int64_t r0, int64_t r1, int64_t n, int64_t k, const T* lu, int64_t lda, T* w, hipStream_t s) {
    if (r1 - r0 <= BS) return block_solve<T>(BWD_LEAVES, r0, r1 - r0, k, lu, lda, w, k, s);
    const int64_t mid = (r0 + r1) / 2;
    int rc = bwd_halving<T>(mid, r1, n, k, lu, lda, w, s);
    // Rows [r0, mid) take rows [mid, r1): entry (i, q) is U(i, q) = F[q][i].
    if (rc == OK)
        rc = k <= 16 ? gemm_skinny<T>(true, false, mid - r0, k, r1 - mid, lu + mid * lda + r0, lda, w + mid * k, k, w + r0 * k, k, s)
                     : gemm<T>(true, false, mid - r0, k, r1 - mid, T(-1), lu + mid * lda + r0, lda, w + mid * k, k, T(1),
                               w + r0 * k, k, s, true);
    if (rc == OK) rc = bwd_halving<T>(r0, mid, n, k, lu, lda, w, s);
    return rc;
}

}  // namespace

// `single` (factor.rs:725-748) for a triangle that a transposed view reads by columns: entry (i, j) is f[j ld + i], one
// contiguous column x of n items, in place. lower: increasing order, then (unless unit) a quotient by the diagonal;
// otherwise decreasing order, then (unless unit) a quotient.
template <typename T>
int trsv1_columns(bool lower, bool unit, int64_t n, const T* f, int64_t ld, T* x, hipStream_t s) {
    if (n == 0) return OK;
    const int nb = (int)((n + B1 - 1) / B1);
    const int padded = (nb + 8 * RUN1 - 1) / (8 * RUN1) * (8 * RUN1);
    const int run = padded <= 2 * device_cus() && exchange_local(true) ? RUN1 : 1;
    const int grid = run > 1 ? padded : (nb + 7) / 8 * 8;
    const size_t words = (size_t)n * (sizeof(T) == 8 ? 2 : 1);
    char* buf = (char*)pool_get(trsv1_bytes<T>(n));
    if (!buf) return hip_status(hipErrorOutOfMemory, "pool_get");
    unsigned long long* pub = (unsigned long long*)buf;
    int* status = (int*)(buf + 2 * words * 8);
    std::unique_lock<std::mutex> lk(resident_mutex(), std::defer_lock);
    if (run > 1) lk.lock();
    zero_words<<<(unsigned)std::min<int64_t>((2 * words + 8 + 255) / 256, 1024), 256, 0, s>>>(pub, (int64_t)(2 * words + 8));
    const int mode = lower ? (unit ? FWD_UNIT : FWD_QUOT) : (unit ? BWD_COL_UNIT : BWD_SINGLE);
    const unsigned long long e1 = (g_epoch1.fetch_add(2) % 0xfffffff0ull) | 1ull;
    trsv1_kernel<T><<<grid, 64, 0, s>>>(mode, n, f, ld, x, pub, pub + words, run, e1, status);
    HS_HIP(hipGetLastError());
    int st = 0;
    HS_HIP(hipMemcpyAsync(&st, status, sizeof(int), hipMemcpyDeviceToHost, s));
    HS_HIP(hipStreamSynchronize(s));
    pool_put(buf);
    if (st) { set_error("triangular solve wait ran out"); return ERR_LIMIT; }
    return OK;
}
template int trsv1_columns<double>(bool, bool, int64_t, const double*, int64_t, double*, hipStream_t);
template int trsv1_columns<float>(bool, bool, int64_t, const float*, int64_t, float*, hipStream_t);


namespace {
// getrs with its workspace: w (n k items) holds the right-hand sides in solve order, buf (trsv1_bytes(n) bytes) the
// single-column passes' words; null pointers take buffers from the pool.
template <typename T>
int getrs_ws(bool trans, int64_t n, int64_t k, const T* lu, int64_t lda, const int64_t* perm, T* b, int64_t ldb,
             hipStream_t s, T* w, char* buf) {
    if (n == 0 || k == 0) return OK;
    const bool pooled = w == nullptr;
    if (pooled && !(w = (T*)pool_get(sizeof(T) * (size_t)(n * k)))) return hip_status(hipErrorOutOfMemory, "pool_get");
    int rc = OK;
    const int64_t total = n * k;
    if (!trans) {
        gather_rows<T><<<(unsigned)((total + 255) / 256), 256, 0, s>>>(n, (int)k, b, ldb, perm, w, k);
        HS_HIP(hipGetLastError());
        if (k == 1) {
            rc = trsv1_lu<T>(false, n, lu, lda, w, s, buf);
        } else {
            rc = fwd_blocked<T>(FWD_UNIT, n, k, lu, lda, w, s);
            if (rc == OK) rc = bwd_halving<T>(0, n, n, k, lu, lda, w, s);
        }
        if (rc == OK) {
            HS_HIP(hipMemcpy2DAsync(b, ldb * sizeof(T), w, k * sizeof(T), k * sizeof(T), n, hipMemcpyDeviceToDevice, s));
        }
    } else {
        HS_HIP(hipMemcpy2DAsync(w, k * sizeof(T), b, ldb * sizeof(T), k * sizeof(T), n, hipMemcpyDeviceToDevice, s));
        if (k == 1) {
            rc = trsv1_lu<T>(true, n, lu, lda, w, s, buf);
        } else {
            rc = fwd_blocked<T>(FWD_DIAG, n, k, lu, lda, w, s);
            if (rc == OK) rc = bwd_unit_blocked<T>(n, k, lu, lda, w, s);
        }
        if (rc == OK) {
            scatter_rows<T><<<(unsigned)((total + 255) / 256), 256, 0, s>>>(n, (int)k, w, k, perm, b, ldb);
            HS_HIP(hipGetLastError());
        }
    }
    HS_HIP(hipStreamSynchronize(s));
    if (pooled) pool_put(w);
    return rc;
}
}  // namespace

template <typename T>
int getrs(bool trans, int64_t n, int64_t k, const T* lu, int64_t lda, const int64_t* perm, T* b, int64_t ldb, hipStream_t s) {
    return getrs_ws<T>(trans, n, k, lu, lda, perm, b, ldb, s, nullptr, nullptr);
}

namespace {
// The workspace of getrs_noalloc, in the library's own device memory: the right-hand sides in solve order (at most
// NOALLOC_ITEMS FP64 items) and the single-column passes' words (order at most NOALLOC_N), 4 MiB and 256 bytes in all.
// Every device that runs getrs_noalloc holds its own copy from its first call until the process ends.
constexpr int64_t NOALLOC_N = 65536, NOALLOC_K = 16, NOALLOC_ITEMS = int64_t(1) << 18;
constexpr size_t NOALLOC_W = (size_t)NOALLOC_ITEMS * sizeof(double);
constexpr size_t NOALLOC_BYTES = NOALLOC_W + 2 * (size_t)NOALLOC_N * 2 * 8 + 256;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

// The address of each device's copy of `scratch`: host code names the array only through these (the array's own name
// in host code is a host-side stand-in, not the device's memory).
std::mutex g_scratch_mu;
void* g_scratch_at[64];

int scratch_of_current_device(void*& at) {
    int dev = 0;
    HS_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) { set_error("device index above 63"); return hip_status(hipErrorInvalidDevice, "getrs_noalloc"); }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_at[dev]) {
        void* p = nullptr;
        HS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) { set_error("hipGetSymbolAddress gave no address"); return hip_status(hipErrorInvalidSymbol, "getrs_noalloc"); }
        g_scratch_at[dev] = p;
    }
    at = g_scratch_at[dev];
    return OK;
}
}  // namespace

// getrs for FP64 with the library's static workspace instead of allocated memory: n <= 65536, k <= 16 and n k <= 2^18
// (beyond them -2, -3 or -3 with a message). Every call on one device shares that device's workspace, so calls that
// overlap on one device race; each call synchronises `s` before it returns.
int getrs_noalloc(bool trans, int64_t n, int64_t k, const double* lu, int64_t lda, const int64_t* perm, double* b,
                  int64_t ldb, hipStream_t s) {
    if (n > NOALLOC_N) { set_error("getrs_noalloc: n above 65536"); return -2; }
    if (k > NOALLOC_K || n * k > NOALLOC_ITEMS) { set_error("getrs_noalloc: k above 16 or n k above 2^18"); return -3; }
    if (n == 0 || k == 0) return OK;
    void* at = nullptr;
    const int rc = scratch_of_current_device(at);
    if (rc) return rc;
    return getrs_ws<double>(trans, n, k, lu, lda, perm, b, ldb, s, (double*)at, (char*)at + NOALLOC_W);
}

namespace {
template <typename T>
__global__ void fill_identity(int64_t n, T* w) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < n * n) w[idx] = idx / n == idx % n ? T(1) : T(0);
}
template <typename T>
__global__ void scatter_columns(int64_t n, const T* __restrict__ w, const int64_t* __restrict__ perm, T* out,
                                int64_t ldo) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < n * n) out[(idx / n) * ldo + perm[idx % n]] = w[idx];
}
}  // namespace

// `lu_inverse` with rustymath's own kernels (factor.rs:645-664): from the identity, L⁻¹ by `forward`'s block mode
// (32-row blocks; a block's rows take only the columns up to the block's end, the rest stay zero, and a product
// carries each block to every row below it, routed as `gemm` routes it), then U⁻¹ by `backward`'s halving over all n
// columns, then out[r][perm[c]] = x[r][c]. out is row-major n×n with leading dimension ldo.
template <typename T>
int getri(int64_t n, const T* lu, int64_t lda, const int64_t* perm, T* out, int64_t ldo, hipStream_t s) {
    if (n == 0) return OK;
    T* w = (T*)pool_get(sizeof(T) * (size_t)(n * n));
    if (!w) return hip_status(hipErrorOutOfMemory, "pool_get");
    const unsigned grid = (unsigned)((n * n + 255) / 256);
    fill_identity<T><<<grid, 256, 0, s>>>(n, w);
    HS_HIP(hipGetLastError());
    int rc = OK;
    for (int64_t i0 = 0; i0 < n && rc == OK; i0 += BR) {
        const int64_t i1 = std::min<int64_t>(n, i0 + BR);
        rc = block_solve<T>(FWD_UNIT, i0, i1 - i0, i1, lu, lda, w, n, s);
        // Rows [i1, n) take rows [i0, i1) over columns [0, i1): entry (i, q) is L(i1 + i, i0 + q) = F[i0 + q][i1 + i].
        if (rc == OK && i1 < n)
            rc = gemm<T>(true, false, n - i1, i1, i1 - i0, T(-1), lu + i0 * lda + i1, lda, w + i0 * n, n, T(1), w + i1 * n,
                         n, s);
    }
    if (rc == OK) rc = bwd_halving<T>(0, n, n, n, lu, lda, w, s);
    if (rc == OK) {
        scatter_columns<T><<<grid, 256, 0, s>>>(n, w, perm, out, ldo);
        HS_HIP(hipGetLastError());
    }
    HS_HIP(hipStreamSynchronize(s));
    pool_put(w);
    return rc;
}

template int getri<double>(int64_t, const double*, int64_t, const int64_t*, double*, int64_t, hipStream_t);
template int getri<float>(int64_t, const float*, int64_t, const int64_t*, float*, int64_t, hipStream_t);
template int getrs<double>(bool, int64_t, int64_t, const double*, int64_t, const int64_t*, double*, int64_t, hipStream_t);
template int getrs<float>(bool, int64_t, int64_t, const float*, int64_t, const int64_t*, float*, int64_t, hipStream_t);

}  // namespace hs
