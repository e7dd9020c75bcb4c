// Batched solves of many independent square systems of different orders (n <= 512) in one persistent launch.
//
// rustymath solves a stack cell by cell, one `LuWork` per thread (core/src/parallel.rs:118-192,
// src/linalg/stacked.rs:88-101): each cell is `LuWork::solve` (core/src/linalg/factor.rs:207-226), the
// `regular` check around `lu_factor` and `lu_solve`. Here each workgroup takes the next system from a queue
// sorted by decreasing cost (longest processing time first) and runs the same arithmetic, item for item:
//
//   - the factorization is right-looking over panels made of whole reference leaves (lu_panel halves down
//     to 16 columns), at most BW columns each. A panel's rows sit one per thread in registers, shifted so
//     v[0] is the current column; rows never move: a swap relabels positions (`perm[position] = row`). Each
//     column costs one barrier: every wave publishes its best candidate row, and all threads pick the
//     winner from LDS. Every item takes its updates in pivot-column order, one fused multiply-add each;
//     multipliers are a · (1/pivot), a quotient for a subnormal pivot; a zero pivot skips the rest of its
//     own leaf;
//   - right of a panel, each thread solves one column's U rows with the panel's unit-lower triangle, in
//     order, and the trailing rows take the panel's product on MFMA: each wave owns 32×32 blocks and adds
//     the terms four at a time. Panels go in pairs, so that the rest of the trailing matrix takes both
//     panels' terms in one product. gfx950's FP64 and FP32 16x16x4 MFMA rounds every step of its four-term
//     sum as a fused multiply-add in order of k (measured), so the product is the in-order fold; padding
//     terms are (+0)·(−0) = −0, which leaves every sum unchanged;
//   - factored and solved in one run, the forward substitution happens in the column loop (each row carries
//     its x); the backward substitution, and both where earlier factors are reused, go in blocks of 32 rows:
//     each half of a wave solves a diagonal block for one right-hand side with register broadcasts, then
//     every row outside it adds the block's terms, in the reference's three orders (csrc/solve.hip has the
//     same three) — and the transposed solve for adjoints.
//
// A is overwritten by its factors without moving rows: row perm[i] of the matrix holds position i of the
// factors, where perm is rustymath's `perm` (factor.rs:132-134). B is overwritten by X where the system is
// regular. The result of a system does not depend on which workgroup ran it, when, or the tuning choices.
#include "hs_kernels.h"
#include "hs_pivot.h"
#include <vector>
#include <algorithm>
#include <numeric>
#include <mutex>
#include <map>

namespace hs {

namespace {

constexpr int NMAX = 512;           // the largest order
constexpr int MAXRHS = 8;           // right-hand sides per system
constexpr int NS = 32;              // `backward`'s leaf height (factor.rs:693-707)
constexpr int SB = 32;              // substitution block height

enum Mode { FACTOR_SOLVE = 0, SOLVE = 1, SOLVE_T = 2 };

struct Desc {
    long long n, k, a_off, b_off, p_off;
};

template <typename T>
struct Bits { typedef typename std::conditional<sizeof(T) == 8, unsigned long long, unsigned int>::type U; };

// One system's matrix as a buffer resource: loads at an index past its n² items return +0 and stores there are
// dropped (the buffer unit's range check), so guarded loads need no branch and stay in flight together.
typedef unsigned int u2v __attribute__((ext_vector_type(2)));
template <typename T>
struct Mat {
    __amdgpu_buffer_rsrc_t r;
    unsigned items;
    __device__ __forceinline__ Mat(T* p, int n) : items((unsigned)n * (unsigned)n) {
        r = __builtin_amdgcn_make_buffer_rsrc(p, 0, (int)(items * sizeof(T)), 0x00020000);
    }
    __device__ __forceinline__ T ld(unsigned i) const {
        if constexpr (sizeof(T) == 8) return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, i * 8u, 0, 0));
        else return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, i * 4u, 0, 0));
    }
    __device__ __forceinline__ void st(unsigned i, T v) const {
        if constexpr (sizeof(T) == 8) __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2v, v), r, i * 8u, 0, 0);
        else __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, i * 4u, 0, 0);
    }
    // i where ok, else an index past the end.
    __device__ __forceinline__ unsigned at(bool ok, int i) const { return ok ? (unsigned)i : items; }
};

// The trailing update on MFMA: the items at positions [r0, r1) × [q0, q1) take the terms of pivot columns
// [c0, c0 + w) in order, A22 += (−L21) U12, with L21 the rows' multipliers in those columns and U12 those pivot rows'
// items of U, in 32×32 blocks (2×2 MFMA tiles per wave). With at least as many 32-row blocks as waves, each wave takes
// whole rows of blocks and keeps its rows' L21 in registers across them (the update reads memory for c and U12
// only); otherwise the blocks are dealt out one by one. BW: the most terms one update takes (w <= BW).
// Where `mx` is not null, it also takes the largest |c| of the items read (NaN ignored): items the first panel's
// terms reach are the matrix's own items.
template <typename T, int BW, int WAVES>
__device__ __forceinline__ void trailing_update(const Mat<T>& A, int n, int c0, int w, int r0, int r1, int q0, int q1,
                                                const short* perm, int wave, int lane, T* mx) {
    typedef typename Mfma<T>::acc V;
    const int nbr = (r1 - r0 + 31) / 32, nbc = (q1 - q0 + 31) / 32;
    const int lr = lane & 15, lk = lane >> 4;
    // The MFMA takes l · (−u): terms outside the block or past w load as +0 on the L side and −0 on the U side, so
    // padding adds (+0)·(−0) = −0, and (−u)·l is the reference's term (its α = −1 folds into U12). All BW / 4 steps
    // run, also past a narrower w (padding then).
    T av[BW / 4][2];
    auto load_l = [&](int rb) {  // the rows' L21, every load issued before any is used
        int arow[2];
        bool aok[2];
#pragma unroll
        for (int ti = 0; ti < 2; ++ti) {
            const int r = rb + 16 * ti + lr;
            aok[ti] = r < r1;
            arow[ti] = perm[r < r1 ? r : r0] * n + c0;
        }
#pragma unroll
        for (int s = 0; s < BW / 4; ++s) {
            const int kk = 4 * s + lk;
            const bool kok = kk < w;
#pragma unroll
            for (int ti = 0; ti < 2; ++ti) av[s][ti] = A.ld(A.at(kok && aok[ti], arow[ti] + kk));
        }
    };
    auto block = [&](int rb, int cb) {  // c's block and U12's columns in flight together, then the MFMAs in order
        V acc[2][2];
        unsigned cidx[2][2][4];
#pragma unroll
        for (int ti = 0; ti < 2; ++ti)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int r = rb + 16 * ti + Mfma<T>::row(lane, e);
                const int prow = perm[r < r1 ? r : r0] * n;
#pragma unroll
                for (int tj = 0; tj < 2; ++tj) {
                    const int col = cb + 16 * tj + lr;
                    cidx[ti][tj][e] = A.at(r < r1 && col < q1, prow + col);
                }
            }
#pragma unroll
        for (int ti = 0; ti < 2; ++ti)
#pragma unroll
            for (int tj = 0; tj < 2; ++tj)
#pragma unroll
                for (int e = 0; e < 4; ++e) acc[ti][tj][e] = A.ld(cidx[ti][tj][e]);
        if (mx) {
#pragma unroll
            for (int ti = 0; ti < 2; ++ti)
#pragma unroll
                for (int tj = 0; tj < 2; ++tj)
#pragma unroll
                    for (int e = 0; e < 4; ++e) { const T v = fabs(acc[ti][tj][e]); *mx = v > *mx ? v : *mx; }
        }
        T bv[BW / 4][2];
#pragma unroll
        for (int s = 0; s < BW / 4; ++s) {
            const int kk = 4 * s + lk;
            const bool kok = kk < w;
            const int brow = perm[c0 + (kok ? kk : 0)] * n;
#pragma unroll
            for (int tj = 0; tj < 2; ++tj) {
                const int col = cb + 16 * tj + lr;
                bv[s][tj] = -A.ld(A.at(kok && col < q1, brow + col));
            }
        }
#pragma unroll
        for (int s = 0; s < BW / 4; ++s)
#pragma unroll
            for (int ti = 0; ti < 2; ++ti)
#pragma unroll
                for (int tj = 0; tj < 2; ++tj) acc[ti][tj] = Mfma<T>::step(av[s][ti], bv[s][tj], acc[ti][tj]);
#pragma unroll
        for (int ti = 0; ti < 2; ++ti)
#pragma unroll
            for (int tj = 0; tj < 2; ++tj)
#pragma unroll
                for (int e = 0; e < 4; ++e) A.st(cidx[ti][tj][e], acc[ti][tj][e]);
    };
    if (nbr >= WAVES) {
        for (int rbk = wave; rbk < nbr; rbk += WAVES) {
            load_l(r0 + 32 * rbk);
            for (int cbk = 0; cbk < nbc; ++cbk) block(r0 + 32 * rbk, q0 + 32 * cbk);
        }
    } else {
        for (int blk = wave; blk < nbr * nbc; blk += WAVES) {
            load_l(r0 + 32 * (blk / nbc));
            block(r0 + 32 * (blk / nbc), q0 + 32 * (blk % nbc));
        }
    }
}

// BW: the widest panel (16 or 32 columns); BT: the workgroup's threads (512: one row of a panel per thread, one
// workgroup per CU; 256: up to two rows per thread, two workgroups and so two systems per CU, one system's
// latency-bound pivot steps beside the other's products). Both are tuning choices; the results depend on neither.
// `mode`: factor and solve (perm and regular written), or solve A X = B / Aᵀ X = B with the factors and perm of an
// earlier run. Where `active` is not null, a system whose flag is 0 is skipped (the refinement's converged systems).
// SOLVE_ONLY: the form for the solve modes, compiled without the factorization, so that its registers leave room for
// more waves per SIMD in the latency-bound substitutions: SOLVE_WGS 256-thread workgroups (systems) per CU.
template <typename T>
constexpr int solve_wgs() { return sizeof(T) == 8 ? 3 : 4; }
template <typename T, int BW, int BT, bool SOLVE_ONLY = false>
__global__ __launch_bounds__(BT, SOLVE_ONLY ? solve_wgs<T>() : 2) void gesv_batched_kernel(int count, const Desc* __restrict__ desc, const int* __restrict__ order,
                                                          T* __restrict__ a_all, T* __restrict__ b_all,
                                                          long long* __restrict__ perm_all, int* __restrict__ regular_all,
                                                          const int* __restrict__ active, int* next, int mode,
                                                          long long* prof) {
    constexpr int WAVES = BT / 64, R = NMAX / BT;  // R: a panel's rows per thread
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // Where `prof` is not null, thread 0 of workgroup 0 adds each phase's 100 MHz ticks into prof[0..5] and counts
    // its systems in prof[7]: setup, panel columns, trailing update, regularity + forward (the forward substitution
    // only where it did not ride along with the factorization), backward, U rows.
    const bool timed = prof != nullptr && blockIdx.x == 0 && tid == 0;
    long long tl = timed ? (long long)__builtin_amdgcn_s_memrealtime() : 0;
#define HS_BPHASE(q) do { if (timed) { long long t_ = __builtin_amdgcn_s_memrealtime(); prof[q] += t_ - tl; tl = t_; } } while (0)
    __shared__ int s_sys;
    __shared__ unsigned long long s_max;
    __shared__ short perm[NMAX];       // position -> row
    __shared__ short leaf_end[NMAX];   // end of the reference leaf of each column
    __shared__ short panel_end[NMAX];  // for a panel's first column, the panel's end
    // Magnitude keys: an FP32 item's fits in 32 bits, so its wave maximum takes one DPP pass.
    typedef typename std::conditional<sizeof(T) == 8, unsigned long long, unsigned>::type Key;
    __shared__ Key s_wkv[2][WAVES];   // each wave's best candidate, double-buffered by column
    __shared__ int s_wkp[2][WAVES];
    // The panel's candidate rows [2][WAVES][CW] (a row's panel items, then its right-hand sides), its rows of U
    // [BW][BW] and unit-lower triangle [BW][BW + 1].
    constexpr int CW = BW + MAXRHS;
    constexpr int PANEL_ITEMS = 2 * WAVES * CW + BW * BW + BW * (BW + 1);
    __shared__ alignas(16) T s_work[SOLVE_ONLY ? 16 : PANEL_ITEMS];
    T (*s_cand)[WAVES][CW] = reinterpret_cast<T (*)[WAVES][CW]>(s_work);
    T (*s_u11)[BW] = reinterpret_cast<T (*)[BW]>(s_work + 2 * WAVES * CW);
    // s_l11n[q][j] = −L11[q + 1 + j][q]: column q of the unit-lower triangle below its diagonal, negated and shifted
    // to the row start, so that step q of the U12 solve reads one row from one scalar offset.
    T (*s_l11n)[BW] = reinterpret_cast<T (*)[BW]>(s_work + 2 * WAVES * CW + BW * BW);
    // The substitutions' x by position; while a system is factored and solved in one run, the forward
    // substitution's partial sums, kept with their rows (see the column loop).
    __shared__ alignas(16) T s_x[NMAX][MAXRHS];
    static_assert(CW <= 64, "a candidate row is read one item per lane");
    static_assert(BW == 16 || BW == 32, "panels hold whole 16-column leaves");
    static_assert(BT * R == NMAX && (R == 1 || R == 2), "a workgroup of 512 or 256 threads");
    const T tiny = sizeof(T) == 8 ? T(2.2250738585072014e-308) : T(1.17549435e-38f);
    const T eps = sizeof(T) == 8 ? T(2.220446049250313e-16) : T(1.1920929e-07f);

    for (;;) {
        if (tid == 0) { s_sys = atomicAdd(next, 1); s_max = 0ull; }
        __syncthreads();
        const int q = s_sys;
        if (q >= count) break;
        const int sys = order[q];
        if (active && !active[sys]) {
            __syncthreads();  // every thread has read s_sys before thread 0 takes the next system
            continue;
        }
        if (timed) prof[7] += 1;
        const Desc d = desc[sys];
        const int n = (int)d.n, k = (int)d.k;
        T* a = a_all + d.a_off;
        T* b = b_all + d.b_off;
        const Mat<T> A(a, n);
        long long* gperm = perm_all + d.p_off;
        bool solve;
        // Factored and solved in one run, the forward substitution over L (lu_solve, factor.rs:633-641) rides
        // along with the factorization: every row carries its x, and at column c each row below the pivot adds
        // the pivot row's finished x times −l, the term `forward` gives it for that column (factor.rs:668-690),
        // in the same column order and with the same fused multiply-add.
        const bool fuse = !SOLVE_ONLY && mode == FACTOR_SOLVE && k > 0;

        if constexpr (SOLVE_ONLY) {
            for (int i = tid; i < n; i += BT) perm[i] = (short)gperm[i];
            __syncthreads();
            solve = n > 0 && k > 0;
        } else if (mode == FACTOR_SOLVE) {
            for (int i = tid; i < n; i += BT) {
                perm[i] = (short)i;
                // The reference leaf holding column i: lu_panel's halving while wider than 16 (factor.rs:592-616).
                int l0 = 0, l1 = n;
                while (l1 - l0 > 16) { const int mid = (l0 + l1) / 2; if (i < mid) l1 = mid; else l0 = mid; }
                leaf_end[i] = (short)l1;
            }
            // x by position, which is row i's before any pivot: B's rows.
            if (fuse)
                for (int i = tid; i < n * MAXRHS; i += BT) {
                    const int r = i / MAXRHS, c = i % MAXRHS;
                    s_x[r][c] = c < k ? b[r * k + c] : T(0);
                }
            __syncthreads();
            if (tid == 0) {
                // Panels: greedy unions of whole leaves, at most BW columns.
                int pstart = 0;
                for (int c = 0; c < n; c = leaf_end[c]) {
                    if (leaf_end[c] - pstart > BW) { panel_end[pstart] = (short)c; pstart = c; }
                }
                if (n > 0) panel_end[pstart] = (short)n;
            }
            __syncthreads();
            HS_BPHASE(0);

            // ---- lu_factor ----
            T big = T(0);  // this thread's share of the matrix's greatest magnitude
            // Panels go in pairs: the first's terms reach only the second's columns before the second is factored;
            // the second's pivot rows then take them right of it, and the rest of the trailing matrix takes both
            // panels' terms in one product, read and written once per pair. Every item still takes its terms in
            // pivot-column order. [dk0, dk1): the pair's first panel while its terms are pending, else dk0 < 0.
            int dk0 = -1, dk1 = -1;
            for (int c0 = 0; c0 < n;) {
                const int c1 = panel_end[c0], w = c1 - c0, rows = n - c0;
                // Slot r of a thread holds the row at panel position tid + r·BT.
                bool has[R];
                int row[R], pos[R];
                T v[R][BW];
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int t = tid + r * BT;
                    has[r] = t < rows;
                    row[r] = perm[c0 + (has[r] ? t : 0)];
                    pos[r] = t;
#pragma unroll
                    for (int j = 0; j < BW; ++j) v[r][j] = A.ld(A.at(has[r] && j < w, row[r] * n + c0 + j));
                }
                // The slots' rows' x, where the forward substitution rides along.
                T x[R][MAXRHS];
#pragma unroll
                for (int r = 0; r < R; ++r)
#pragma unroll
                    for (int cc = 0; cc < MAXRHS; ++cc) x[r][cc] = fuse ? s_x[c0 + (has[r] ? pos[r] : 0)][cc] : T(0);
                // `greatest` magnitude of the matrix, NaN ignored (factor.rs:488), for `regular`: the first panel's
                // loads, its rows of U right of it and its trailing update read every item of A once, as it was.
                if (c0 == 0) {
#pragma unroll
                    for (int r = 0; r < R; ++r)
#pragma unroll
                        for (int j = 0; j < BW; ++j) { const T x = fabs(v[r][j]); big = x > big ? x : big; }
                }
#pragma clang loop unroll(disable)
                for (int c = 0; c < w; ++c) {
                    const int buf = c & 1;
                    // First greatest magnitude at positions >= c, NaN ignored; all NaN -> position c (factor.rs:476-490):
                    // the thread's best slot (larger key, then smaller position), then the wave's.
                    Key kv = 0;
                    int kp = 0x7fffffff, kr = 0;
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        const bool cnd = has[r] && pos[r] >= c;
                        const Key k2 = cnd ? (Key)mag_key(v[r][0]) : Key(0);
                        const int p2 = cnd ? pos[r] : 0x7fffffff;
                        if (k2 > kv || (k2 == kv && p2 < kp)) { kv = k2; kp = p2; kr = r; }
                    }
                    int hl = 0;
                    wave_best(kv, kp, hl);
                    // The candidate's live items (w − c of them; the rest are never read) and x, from the slot that
                    // holds it, named by a wave-uniform index so that one store sequence runs without selects.
                    auto put_cand = [&](const T (&vr)[BW], const T (&xr)[MAXRHS]) {
                        constexpr int G = 32 / sizeof(T);  // items per pair of 128-bit stores
                        T* dst = s_cand[buf][wave];
#pragma unroll
                        for (int j0 = 0; j0 < BW; j0 += G)
                            if (j0 < w - c)
#pragma unroll
                                for (int j = j0; j < j0 + G; ++j) dst[j] = vr[j];
                        if (fuse) {
#pragma unroll
                            for (int j0 = 0; j0 < MAXRHS; j0 += G)
                                if (j0 < k)
#pragma unroll
                                    for (int j = j0; j < j0 + G; ++j) dst[BW + j] = xr[j];
                        }
                    };
                    if (hl >= 0) {
                        const int ks = R == 1 ? 0 : __builtin_amdgcn_readlane(kr, hl);
                        if (lane == hl) {
                            if (ks == 0) put_cand(v[0], x[0]);
                            else put_cand(v[R - 1], x[R - 1]);
                        }
                    }
                    if (lane == 0) { s_wkv[buf][wave] = kv; s_wkp[buf][wave] = kp; }
                    __syncthreads();
                    // The waves' bests: every lane reads them all from LDS (one broadcast each) and keeps the first
                    // greatest key, then the smallest position; no cross-lane steps.
                    Key bkv = s_wkv[buf][0];
                    int p = s_wkp[buf][0], win = 0;
#pragma unroll
                    for (int u = 1; u < WAVES; ++u) {
                        const Key ku = s_wkv[buf][u];
                        const int pu = s_wkp[buf][u];
                        const bool better = ku > bkv || (ku == bkv && pu < p);
                        bkv = better ? ku : bkv;
                        p = better ? pu : p;
                        win = better ? u : win;
                    }
                    const T* piv = s_cand[buf][win];
                    // The pivot row as scalar operands: lane j reads item j (every lane active), then reads go lane
                    // to scalar.
                    T mine = piv[lane < CW ? lane : 0];
                    asm volatile("" : "+v"(mine));
                    if (tid < BW) s_u11[c][tid] = piv[tid];
                    const T pivot = lane_value(mine, 0);
                    // Rows at positions > c take the column's multiplier and update. Every slot computes (those of
                    // pivoted rows or past the panel's rows are never read again), so the update has no branch.
                    // divide (factor.rs:506-508); a zero pivot leaves the column undivided and skips the rest of its
                    // own leaf (:606): the shared pass then takes +0 for both the multiplier and the pivot row's
                    // items, and v + (+0)·(−0) = v + (−0) is v exactly, so it only shifts the items; the items of the
                    // leaf's later columns take their updates after it, in place.
                    const bool nz = pivot != T(0);
                    T l[R], lm[R];
                    bool upd[R];
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        if (has[r] && pos[r] == p) pos[r] = c;
                        else if (has[r] && pos[r] == c) pos[r] = p;
                        upd[r] = has[r] && pos[r] > c;
                        l[r] = v[r][0];
                    }
                    if (nz) {
                        if (fabs(pivot) >= tiny) {
                            const T rp = T(1) / pivot;
#pragma unroll
                            for (int r = 0; r < R; ++r) l[r] = l[r] * rp;
                        } else {
#pragma unroll
                            for (int r = 0; r < R; ++r) l[r] = l[r] / pivot;
                        }
                    }
#pragma unroll
                    for (int r = 0; r < R; ++r) lm[r] = nz ? l[r] : T(0);
                    // The forward substitution's terms of this column. A zero pivot makes the system irregular, and
                    // then no x is written.
                    if (fuse && nz) {
                        T y[MAXRHS];
#pragma unroll
                        for (int cc = 0; cc < MAXRHS; ++cc) y[cc] = cc < k ? lane_value(mine, BW + cc) : T(0);
#pragma unroll
                        for (int r = 0; r < R; ++r) {
                            if (upd[r]) {
#pragma unroll
                                for (int cc = 0; cc < MAXRHS; ++cc)
                                    if (cc < k) x[r][cc] = __builtin_elementwise_fma(y[cc], -l[r], x[r][cc]);
                            }
                        }
                    }
                    // The pivot row as the update reads it: +0 throughout for a zero pivot (one select per column, so
                    // that reading an item takes no branch).
                    const T mz = nz ? mine : T(0);
#pragma unroll
                    for (int j0 = 1; j0 < BW; j0 += 8) {
                        if (j0 < w - c) {  // items past the panel are never used
#pragma unroll
                            for (int j = j0; j < j0 + 8 && j < BW; j += 2) {
                                // Two items read out into scalars before their FMAs, so that no FMA waits for the
                                // read just before it.
                                const T u0 = lane_value(mz, j);
                                const T u1 = j + 1 < BW ? lane_value(mz, j + 1) : T(0);
                                asm volatile("" ::"s"(u0), "s"(u1));
#pragma unroll
                                for (int r = 0; r < R; ++r) v[r][j - 1] = __builtin_elementwise_fma(lm[r], -u0, v[r][j]);
                                if (j + 1 < BW) {
#pragma unroll
                                    for (int r = 0; r < R; ++r) v[r][j] = __builtin_elementwise_fma(lm[r], -u1, v[r][j + 1]);
                                }
                            }
                        }
                    }
                    if (!nz) {
                        const int skip = leaf_end[c0 + c] - c0 - c;
#pragma unroll
                        for (int j = 1; j < BW; ++j) {
                            if (j >= skip) {
                                const T uj = piv[j];
#pragma unroll
                                for (int r = 0; r < R; ++r) v[r][j - 1] = __builtin_elementwise_fma(l[r], -uj, v[r][j - 1]);
                            }
                        }
                    }
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        v[r][BW - 1] = T(0);
                        if (upd[r]) a[row[r] * n + c0 + c] = l[r];
                    }
                }
#pragma unroll
                for (int r = 0; r < R; ++r)
                    if (has[r]) {
                        perm[c0 + pos[r]] = (short)row[r];
                        if (fuse) {
#pragma unroll
                            for (int cc = 0; cc < MAXRHS; ++cc) s_x[c0 + pos[r]][cc] = x[r][cc];
                        }
                    }
                __syncthreads();
                HS_BPHASE(1);
                if (dk0 >= 0 && c1 < n) {
                    // The pair's first panel's terms for this panel's pivot rows, right of it.
                    trailing_update<T, BW, WAVES>(A, n, dk0, dk1 - dk0, c0, c1, c1, n, perm, wave, lane,
                                                  dk0 == 0 ? &big : nullptr);
                    __syncthreads();
                }
                // The panel's rows of U, from LDS to their rows.
                for (int i = tid; i < BW * BW; i += BT) {
                    const int cc = i / BW, j = i % BW;
                    if (cc < w && j < w - cc) a[perm[c0 + cc] * n + c0 + cc + j] = s_u11[cc][j];
                }
                if (c1 < n) {
                    // The panel's unit-lower triangle into LDS, every load issued before any is stored.
                    constexpr int LF = (BW * BW + BT - 1) / BT;
                    T lv[LF];
#pragma unroll
                    for (int t = 0; t < LF; ++t) {
                        const int i = tid + t * BT, r = i / BW, cq = i % BW;
                        const bool ok = i < BW * BW && r < w && cq < r;
                        lv[t] = A.ld(A.at(ok, perm[c0 + (ok ? r : 0)] * n + c0 + cq));
                    }
                    // U12: the panel's unit-lower solve for each column right of it, in order, one column per thread
                    // (two in turn at 256 threads); the first column's loads are in flight across the barrier.
                    T u[BW];
                    int col = c1 + tid;
#pragma unroll
                    for (int j = 0; j < BW; ++j) u[j] = A.ld(A.at(col < n && j < w, perm[c0 + (j < w ? j : 0)] * n + col));
#pragma unroll
                    for (int t = 0; t < LF; ++t) {
                        const int i = tid + t * BT, r = i / BW, cq = i % BW;
                        if (i < BW * BW && r > cq) s_l11n[cq][r - cq - 1] = -lv[t];
                    }
                    __syncthreads();
                    while (col < n) {
                        if (c0 == 0) {
#pragma unroll
                            for (int j = 0; j < BW; ++j) { const T x = fabs(u[j]); big = x > big ? x : big; }
                        }
                        // u[j] holds row qq + j of the column: step qq finishes row qq and shifts the rest down. In
                        // the P-th quarter of the steps only the rows below it (j < BW − P·BW/4) are live.
#pragma unroll
                        for (int P = 0; P < 4; ++P) {
                            constexpr int QW = BW / 4;
                            const int q1 = (P + 1) * QW < w ? (P + 1) * QW : w;
#pragma clang loop unroll(disable)
                            for (int qq = P * QW; qq < q1; ++qq) {
                                const T x = u[0];
                                a[perm[c0 + qq] * n + col] = x;
                                const T* lq = s_l11n[qq];
#pragma unroll
                                for (int j = 1; j < BW - P * QW; ++j) u[j - 1] = __builtin_elementwise_fma(lq[j - 1], x, u[j]);
                            }
                        }
                        col += BT;
                        if (col < n) {
#pragma unroll
                            for (int j = 0; j < BW; ++j) u[j] = A.ld(A.at(j < w, perm[c0 + (j < w ? j : 0)] * n + col));
                        }
                    }
                    __syncthreads();
                    HS_BPHASE(5);
                    const int c2 = panel_end[c1];
                    if (dk0 < 0 && c2 < n) {
                        trailing_update<T, BW, WAVES>(A, n, c0, w, c1, n, c1, c2, perm, wave, lane,
                                                      c0 == 0 ? &big : nullptr);
                        dk0 = c0;
                        dk1 = c1;
                    } else if (dk0 >= 0) {
                        trailing_update<T, 2 * BW, WAVES>(A, n, dk0, c1 - dk0, c1, n, c1, n, perm, wave, lane,
                                                          dk0 == 0 ? &big : nullptr);
                        dk0 = -1;
                    } else {
                        trailing_update<T, BW, WAVES>(A, n, c0, w, c1, n, c1, n, perm, wave, lane,
                                                      c0 == 0 ? &big : nullptr);
                    }
                }
                __syncthreads();
                HS_BPHASE(2);
                c0 = c1;
            }

            // ---- regular (factor.rs:221-226): no |u_ii| at or below eps · n · greatest ----
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) { const T o = __shfl_xor(big, off); big = o > big ? o : big; }
            if (lane == 0) atomicMax(&s_max, (unsigned long long)__builtin_bit_cast(typename Bits<T>::U, big));
            __syncthreads();
            const T largest = __builtin_bit_cast(T, (typename Bits<T>::U)s_max);
            const T negligible = eps * T(n) * largest;
            int mine_ok = 1;
            for (int i = tid; i < n; i += BT) mine_ok &= !(fabs(a[perm[i] * n + i]) <= negligible);
            const int ok = __syncthreads_and(mine_ok);
            for (int i = tid; i < n; i += BT) gperm[i] = perm[i];
            if (tid == 0) regular_all[sys] = ok && n > 0;
            solve = ok && n > 0 && k > 0;
        } else {
            for (int i = tid; i < n; i += BT) perm[i] = (short)gperm[i];
            __syncthreads();
            solve = n > 0 && k > 0;
        }

        if (solve) {
            // ---- lu_solve (factor.rs:633-641): x = rows perm of b, then L, then U; or, transposed, Uᵀ then Lᵀ,
            // then the inverse permutation ----
            const bool tr = mode == SOLVE_T;
            // Item (i, j) of the triangle being solved: F(i, j), or F(j, i) for the transposed solve.
            auto ent_at = [&](bool ok, int i, int j) -> unsigned { return A.at(ok, tr ? perm[j] * n + i : perm[i] * n + j); };
            auto ent = [&](int i, int j) -> T { return A.ld(ent_at(true, i, j)); };
            if (!fuse) {
                for (int i = tid; i < n * MAXRHS; i += BT) {
                    const int r = i / MAXRHS, c = i % MAXRHS;
                    s_x[r][c] = c < k ? b[(tr ? r : perm[r]) * k + c] : T(0);
                }
                __syncthreads();
            }
            // Forward: L (unit) or Uᵀ (a reciprocal multiply at each row's turn), increasing order. A diagonal block
            // per right-hand side, one wave each (in turn where there are more columns than waves). Done already
            // where the factorization carried it.
            for (int p0 = 0; p0 < (fuse ? 0 : n); p0 += SB) {
                const int h = n - p0 < SB ? n - p0 : SB;
                // The block's terms for this thread's first row below it, in flight across the diagonal block.
                const int i0 = p0 + h + tid, i0c = i0 < n ? i0 : n - 1;
                T e0[SB];
#pragma unroll
                for (int j = 0; j < SB; ++j) e0[j] = A.ld(ent_at(j < h && i0 < n, i0c, p0 + (j < h ? j : 0)));
                if (wave < k) {
                    // Each half of a wave solves one column: the wave's columns cc and cc + WAVES at once, with the
                    // block's items loaded once for both.
                    const int half = lane >> 5, ln = lane & 31;
                    const bool act = ln < h;
                    const int r = p0 + (act ? ln : 0);
                    T mj[SB];
#pragma unroll
                    for (int j = 0; j < SB; ++j) mj[j] = A.ld(ent_at(act && j < ln, r, p0 + (j < h ? j : 0)));
                    const T dgl = ent(r, r);
                    const T dg = (tr && act) ? dgl : T(1);
                    for (int cc0 = wave; cc0 < k; cc0 += 2 * WAVES) {
                        const int cc = cc0 + half * WAVES;
                        const bool live = act && cc < k;
                        T x = live ? s_x[r][cc] : T(0);
#pragma unroll
                        for (int j = 0; j < SB; ++j) {
                            if (j < h) {
                                if (tr && ln == j) x = fabs(dg) >= tiny ? x * (T(1) / dg) : x / dg;
                                const T lo = lane_value(x, j), hi = lane_value(x, 32 + j);
                                const T xj = half ? hi : lo;
                                if (act && ln > j) x = __builtin_elementwise_fma(xj, -mj[j], x);
                            }
                        }
                        if (live) s_x[r][cc] = x;
                    }
                }
                __syncthreads();
                for (int i = i0; i < n; i += BT) {
                    T e[SB];
#pragma unroll
                    for (int j = 0; j < SB; ++j) e[j] = -(i == i0 ? e0[j] : A.ld(ent_at(j < h, i, p0 + (j < h ? j : 0))));
#pragma unroll
                    for (int c = 0; c < MAXRHS; ++c) {
                        if (c < k) {
                            T acc = s_x[i][c];
#pragma unroll
                            for (int j = 0; j < SB; ++j) if (j < h) acc = __builtin_elementwise_fma(s_x[p0 + j][c], e[j], acc);
                            s_x[i][c] = acc;
                        }
                    }
                }
                __syncthreads();
            }
            HS_BPHASE(3);
            if (tr || k == 1) {
                // Backward, decreasing order: U with a quotient at each row's turn (`single`, factor.rs:725-748), or
                // Lᵀ (unit). Diagonal blocks sit right-aligned in the wave: lane SB - 1 holds the block's last row.
                for (int q1 = n; q1 > 0; q1 -= SB) {
                    const int q0 = q1 - SB > 0 ? q1 - SB : 0, h = q1 - q0, off = SB - h;
                    // The block's terms for this thread's first row above it, in flight across the diagonal block.
                    const int i0c = tid < q0 ? tid : 0;
                    T e0[SB];
#pragma unroll
                    for (int j = 0; j < SB; ++j) e0[j] = A.ld(ent_at(j < h && tid < q0, i0c, q0 + (j < h ? j : 0)));
                    if (wave < k) {  // two columns per wave, one per half, as in the forward substitution
                        const int half = lane >> 5, ln = lane & 31;
                        const bool act = ln >= off;
                        const int r = act ? q0 + ln - off : q0;
                        T mj[SB];
#pragma unroll
                        for (int j = 0; j < SB; ++j) mj[j] = A.ld(ent_at(act && j > ln, r, j >= off ? q0 + j - off : q0));
                        const T dgl = ent(r, r);
                        const T dg = (!tr && act) ? dgl : T(1);
                        for (int cc0 = wave; cc0 < k; cc0 += 2 * WAVES) {
                            const int cc = cc0 + half * WAVES;
                            const bool live = act && cc < k;
                            T x = live ? s_x[r][cc] : T(0);
#pragma unroll
                            for (int s = 0; s < SB; ++s) {
                                const int L = SB - 1 - s;
                                if (s < h) {
                                    if (!tr && ln == L) x = x / dg;
                                    const T lo = lane_value(x, L), hi = lane_value(x, 32 + L);
                                    const T xl = half ? hi : lo;
                                    if (act && ln < L) x = __builtin_elementwise_fma(xl, -mj[L], x);
                                }
                            }
                            if (live) s_x[r][cc] = x;
                        }
                    }
                    __syncthreads();
                    for (int i = tid; i < q0; i += BT) {
                        T e[SB];
#pragma unroll
                        for (int j = 0; j < SB; ++j) e[j] = -(i == tid ? e0[j] : A.ld(ent_at(j < h, i, q0 + (j < h ? j : 0))));
#pragma unroll
                        for (int c = 0; c < MAXRHS; ++c) {
                            if (c < k) {
                                T acc = s_x[i][c];
#pragma unroll
                                for (int j = SB - 1; j >= 0; --j) if (j < h) acc = __builtin_elementwise_fma(s_x[q0 + j][c], e[j], acc);
                                s_x[i][c] = acc;
                            }
                        }
                    }
                    __syncthreads();
                }
            } else {
                // Several columns: `backward`'s halving above NS rows (factor.rs:693-707). Leaves run bottom to top,
                // each row adding the rows after it in its leaf in increasing order, then a reciprocal multiply; once
                // the leaf starting at a split's middle is done, the split's top half adds its bottom half.
                for (int end = n; end > 0;) {
                    int r0 = 0, r1 = n;
                    while (r1 - r0 > NS) { const int mid = (r0 + r1) / 2; if (end <= mid) r1 = mid; else r0 = mid; }
                    const int h = r1 - r0, off = SB - h;
                    // Each half of a wave solves one column (the leaf's rows right-aligned in its 32 lanes): the
                    // wave's columns cc and cc + WAVES at once, with the leaf's items loaded once for both.
                    static_assert(SB == 32, "a leaf's rows fill half a wave");
                    const int half = lane >> 5, ln = lane & 31;
                    const bool act = ln >= off;
                    const int r = act ? r0 + ln - off : r0;
                    T mj[SB];
#pragma unroll
                    for (int j = 0; j < SB; ++j) mj[j] = A.ld(A.at(wave < k && act && j > ln, perm[r] * n + (j >= off ? r0 + j - off : r0)));
                    const T dgl = a[perm[r] * n + r];
                    const T dg = act ? dgl : T(1);
                    for (int cc0 = wave; cc0 < k; cc0 += 2 * WAVES) {
                        const int cc = cc0 + half * WAVES;
                        const bool live = act && cc < k;
                        T x = live ? s_x[r][cc] : T(0);
                        T xs[SB];
#pragma unroll
                        for (int s = 0; s < SB; ++s) {
                            const int L = SB - 1 - s;
                            if (s < h) {
                                if (ln == L) {
#pragma unroll
                                    for (int j = L + 1; j < SB; ++j) x = __builtin_elementwise_fma(xs[j], -mj[j], x);
                                    x = fabs(dg) >= tiny ? x * (T(1) / dg) : x / dg;
                                }
                                const T lo = lane_value(x, L), hi = lane_value(x, 32 + L);
                                xs[L] = half ? hi : lo;
                            }
                        }
                        if (live) s_x[r][cc] = x;
                    }
                    __syncthreads();
                    int a0 = 0, a1 = n;
                    while (a1 - a0 > NS) {
                        const int mid = (a0 + a1) / 2;
                        if (mid == r0) {
                            for (int i = a0 + tid; i < mid; i += BT) {
                                for (int j0 = mid; j0 < a1; j0 += SB) {
                                    T e[SB];
#pragma unroll
                                    for (int j = 0; j < SB; ++j) e[j] = -A.ld(A.at(j0 + j < a1, perm[i] * n + j0 + j));
#pragma unroll
                                    for (int c = 0; c < MAXRHS; ++c) {
                                        if (c < k) {
                                            T acc = s_x[i][c];
#pragma unroll
                                            for (int j = 0; j < SB; ++j)
                                                if (j0 + j < a1) acc = __builtin_elementwise_fma(s_x[j0 + j][c], e[j], acc);
                                            s_x[i][c] = acc;
                                        }
                                    }
                                }
                            }
                            __syncthreads();
                            break;
                        }
                        if (r0 < mid) a1 = mid; else a0 = mid;
                    }
                    end = r0;
                }
            }
            for (int i = tid; i < n * MAXRHS; i += BT) {
                const int r = i / MAXRHS, c = i % MAXRHS;
                if (c < k) b[(tr ? perm[r] : r) * k + c] = s_x[r][c];
            }
        }
        __syncthreads();
        HS_BPHASE(4);
    }
#undef HS_BPHASE
}

__global__ void zero_int(int* p) { *p = 0; }

// ---- The FP32-LU + FP64-refinement form (no counterpart in the reference, which has no refinement) ----
// Per system, one workgroup of 256 threads: the FP32 kernel above factors float32(A_i) and solves, and each refinement
// step here adds the correction in FP64, forms r = B - A X in FP64 from the original A, and decides on the GPU whether
// the normwise backward error ||R||_F / (||A||_F ||X||_F) has reached the system's tolerance (tol_per_n · n · 2^-52);
// a system that has not takes another FP32 solve of its residual, scaled by a power of two into FP32's range.
constexpr int IRT = 256;

__device__ __forceinline__ double block_sum(double v, double* red) {
    // A fixed order: within each wave by halving, then the four waves' sums in order.
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return (red[0] + red[1]) + (red[2] + red[3]);
}

// v + (v of the lane CTRL names) over the lanes ROWS enables, 0 from lanes with no source (a DPP step of a wave sum).
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_add(double v) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)u, CTRL, ROWS, 0xf, true);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), CTRL, ROWS, 0xf, true);
    return v + __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// The sum of v over a wave's 64 lanes, in lane 63: shifts by 1, 2, 4 and 8 within each 16-lane row, then lane 15 of
// each row into the next row and lane 31 into the upper half (DPP, no LDS).
__device__ __forceinline__ double wave_sum63(double v) {
    v = dpp_add<0x111, 0xf>(v);
    v = dpp_add<0x112, 0xf>(v);
    v = dpp_add<0x114, 0xf>(v);
    v = dpp_add<0x118, 0xf>(v);
    v = dpp_add<0x142, 0xa>(v);
    return dpp_add<0x143, 0xc>(v);
}

// The power of two that brings a positive finite v to [1, 2), or 1.
__device__ __forceinline__ double pow2_scale(double v) {
    if (!(v > 0) || !(v <= 1.7976931348623157e308)) return 1.0;
    return exp2(-floor(log2(v)));
}

__global__ __launch_bounds__(IRT) void ir_prep_kernel(int count, const Desc* __restrict__ desc, const double* __restrict__ a,
                                                      const double* __restrict__ b, float* __restrict__ a32,
                                                      float* __restrict__ r32, double* __restrict__ anorm,
                                                      double* __restrict__ rscale) {
    __shared__ double red[4];
    const int sys = blockIdx.x;
    if (sys >= count) return;
    const Desc d = desc[sys];
    const long long nn = d.n * d.n, nk = d.n * d.k;
    double aa = 0, bb = 0;
    for (long long i = threadIdx.x; i < nn; i += IRT) {
        const double v = a[d.a_off + i];
        a32[d.a_off + i] = (float)v;
        aa = __builtin_elementwise_fma(v, v, aa);
    }
    for (long long i = threadIdx.x; i < nk; i += IRT) {
        const double v = b[d.b_off + i];
        bb = __builtin_elementwise_fma(v, v, bb);
    }
    aa = block_sum(aa, red);
    bb = block_sum(bb, red);
    const double sc = pow2_scale(sqrt(bb / (double)(nk > 0 ? nk : 1)));
    for (long long i = threadIdx.x; i < nk; i += IRT) r32[d.b_off + i] = (float)(b[d.b_off + i] * sc);
    if (threadIdx.x == 0) { anorm[sys] = sqrt(aa); rscale[sys] = sc; }
}

// One refinement step for every system still active (all regular systems at the first step, which takes X from the
// first FP32 solve): X += correction, R = B - A X, the backward error, and the next scaled residual where needed.
__global__ __launch_bounds__(IRT) void ir_step_kernel(int count, const Desc* __restrict__ desc, const double* __restrict__ a,
                                                      const double* __restrict__ b, double* __restrict__ x,
                                                      float* __restrict__ r32, double* __restrict__ rscale,
                                                      const double* __restrict__ anorm, int* __restrict__ active,
                                                      int* __restrict__ iters, double* __restrict__ berr,
                                                      const int* __restrict__ regular, int first, int last,
                                                      double tol_per_n) {
    __shared__ double xs[NMAX][MAXRHS];
    __shared__ double red[4];
    const int sys = blockIdx.x;
    if (sys >= count) return;
    const Desc d = desc[sys];
    const int n = (int)d.n, k = (int)d.k, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (first) {
        if (tid == 0) { iters[sys] = 0; berr[sys] = regular[sys] && n > 0 ? 0.0 : __builtin_nan(""); active[sys] = 0; }
        if (!regular[sys] || n == 0 || k == 0) return;
    } else if (!active[sys]) {
        return;
    }
    const double inv = 1.0 / rscale[sys];
    for (int i = tid; i < n * k; i += IRT) {
        const double c = (double)r32[d.b_off + i];
        const double v = first ? c * inv : __builtin_elementwise_fma(c, inv, x[d.b_off + i]);
        x[d.b_off + i] = v;
        xs[i / k][i % k] = v;
    }
    __syncthreads();
    // R = B - A X: one wave per pair of rows, lanes along the rows, every item of both rows loaded before the sums
    // (n <= 512: eight per lane and row), then wave sums by DPP. The residual goes to r32 as FP32 at once (its scaling
    // by a power of two afterwards is exact), its squares into rr.
    const double* A = a + d.a_off;
    double rr = 0, xx = 0;
    constexpr int PER = NMAX / 64;
    for (int row0 = 2 * wave; row0 < n; row0 += 2 * (IRT / 64)) {
        const int row1 = row0 + 1 < n ? row0 + 1 : row0;
        const double* A0 = A + (long long)row0 * n;
        const double* A1 = A + (long long)row1 * n;
        double a0[PER], a1[PER];
#pragma unroll
        for (int t = 0; t < PER; ++t) {
            const int j = lane + 64 * t;
            a0[t] = j < n ? A0[j] : 0.0;
            a1[t] = j < n ? A1[j] : 0.0;
        }
        double acc[2][MAXRHS];
#pragma unroll
        for (int c = 0; c < MAXRHS; ++c) acc[0][c] = acc[1][c] = 0;
#pragma unroll
        for (int t = 0; t < PER; ++t) {
            const int j = lane + 64 * t;
            if (j < n) {
#pragma unroll
                for (int c = 0; c < MAXRHS; ++c) {
                    if (c < k) {
                        const double xj = xs[j][c];
                        acc[0][c] = __builtin_elementwise_fma(a0[t], xj, acc[0][c]);
                        acc[1][c] = __builtin_elementwise_fma(a1[t], xj, acc[1][c]);
                    }
                }
            }
        }
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int row = q ? row1 : row0;
            if (q == 0 || row1 != row0) {
#pragma unroll
                for (int c = 0; c < MAXRHS; ++c) {
                    if (c < k) {
                        const double v = wave_sum63(acc[q][c]);
                        if (lane == 63) {
                            const double r = b[d.b_off + (long long)row * k + c] - v;
                            r32[d.b_off + (long long)row * k + c] = (float)r;
                            rr = __builtin_elementwise_fma(r, r, rr);
                        }
                    }
                }
            }
        }
    }
    for (int i = tid; i < n * k; i += IRT) {
        const double v = xs[i / k][i % k];
        xx = __builtin_elementwise_fma(v, v, xx);
    }
    rr = block_sum(rr, red);
    xx = block_sum(xx, red);
    const double be = sqrt(rr) / (anorm[sys] * sqrt(xx));
    const double tol = tol_per_n * (double)n * 2.220446049250313e-16;
    const bool more = !last && be > tol && be <= 1.7976931348623157e308;
    if (tid == 0) {
        berr[sys] = be;
        active[sys] = more ? 1 : 0;
        if (!first) iters[sys] += 1;
    }
    if (more) {
        const double sc = pow2_scale(sqrt(rr / (double)(n * k)));
        if (tid == 0) rscale[sys] = sc;
        __syncthreads();  // every row's residual is in r32
        for (int i = tid; i < n * k; i += IRT) r32[d.b_off + i] = (float)((double)r32[d.b_off + i] * sc);
    }
}

// A plan: the systems' descriptors and their order on the device, and the launch captured in a hipGraph.
struct Plan {
    int count = 0;
    bool f64 = true;
    int mode = FACTOR_SOLVE;
    Desc* d_desc = nullptr;
    int* d_order = nullptr;
    int* d_next = nullptr;
    int grid = 0;
    int bw = 32;
    int wgs = 1;
    void *a = nullptr, *b = nullptr;
    long long* perm = nullptr;
    int* regular = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // The FP32-LU + FP64-refinement form (a, b above are then the FP32 factors and scaled residuals):
    bool ir = false;
    const double* a64 = nullptr;
    const double* b64 = nullptr;
    double* x64 = nullptr;
    int* iters = nullptr;
    double* berr = nullptr;
    double* d_anorm = nullptr;
    double* d_rscale = nullptr;
    int* d_active = nullptr;
    int max_iter = 0;
    double tol_per_n = 10.0;
};
long long* g_bprof = nullptr;  // the batched kernel's phase ticks, while profiling is on
bool g_bprof_on = false;
std::mutex g_plans_mu;
std::map<long long, Plan*> g_plans;
long long g_next_plan = 1;

template <typename T>
int enqueue_mode(Plan* p, int mode, const int* active, hipStream_t s, long long* prof) {
    zero_int<<<1, 1, 0, s>>>(p->d_next);
    if (mode != FACTOR_SOLVE) {
        // The solve modes: three (FP64) or four (FP32) 256-thread workgroups per CU (the result depends on neither).
        gesv_batched_kernel<T, 32, 256, true><<<dim3(device_cus() * solve_wgs<T>()), 256, 0, s>>>(
            p->count, p->d_desc, p->d_order, (T*)p->a, (T*)p->b, p->perm, p->regular, active, p->d_next, mode, prof);
        HS_HIP(hipGetLastError());
        return OK;
    }
    // One workgroup of 512 threads per CU, or two of 256.
    const dim3 grid(p->grid);
#define HS_BLAUNCH(bw, bt) gesv_batched_kernel<T, bw, bt><<<grid, bt, 0, s>>>(p->count, p->d_desc, p->d_order, (T*)p->a, \
        (T*)p->b, p->perm, p->regular, active, p->d_next, mode, prof)
    if (p->bw == 16) { if (p->wgs == 2) HS_BLAUNCH(16, 256); else HS_BLAUNCH(16, 512); }
    else { if (p->wgs == 2) HS_BLAUNCH(32, 256); else HS_BLAUNCH(32, 512); }
#undef HS_BLAUNCH
    HS_HIP(hipGetLastError());
    return OK;
}

// The refinement form's sequence: cast and scale, the FP32 factorization and first solve, then max_iter rounds of a
// refinement step and an FP32 solve of the systems still active, and a last step that only measures.
int enqueue_ir(Plan* p, hipStream_t s, long long* prof) {
    const unsigned g = (unsigned)(p->count > 0 ? p->count : 1);
    ir_prep_kernel<<<g, IRT, 0, s>>>(p->count, p->d_desc, p->a64, p->b64, (float*)p->a, (float*)p->b, p->d_anorm,
                                     p->d_rscale);
    HS_HIP(hipGetLastError());
    int rc = enqueue_mode<float>(p, FACTOR_SOLVE, nullptr, s, prof);
    for (int it = 0; rc == OK && it <= p->max_iter; ++it) {
        ir_step_kernel<<<g, IRT, 0, s>>>(p->count, p->d_desc, p->a64, p->b64, p->x64, (float*)p->b, p->d_rscale,
                                         p->d_anorm, p->d_active, p->iters, p->berr, p->regular, it == 0,
                                         it == p->max_iter, p->tol_per_n);
        HS_HIP(hipGetLastError());
        if (it < p->max_iter) rc = enqueue_mode<float>(p, SOLVE, p->d_active, s, prof);
    }
    return rc;
}

template <typename T>
int enqueue(Plan* p, hipStream_t s, long long* prof) {
    if (p->ir) return enqueue_ir(p, s, prof);
    return enqueue_mode<T>(p, p->mode, nullptr, s, prof);
}

Plan* find_plan(long long handle) {
    std::lock_guard<std::mutex> lk(g_plans_mu);
    auto it = g_plans.find(handle);
    return it == g_plans.end() ? nullptr : it->second;
}

}  // namespace

int batched_plan(int64_t count, const int64_t* n, const int64_t* k, const int64_t* a_off, const int64_t* b_off,
                 const int64_t* p_off, bool f64, void* a, void* b, int64_t* perm, int* regular, int mode, int panel_width,
                 int wgs_per_cu, long long* handle) {
    if (mode < FACTOR_SOLVE || mode > SOLVE_T) { set_error("batched mode must be 0, 1 or 2"); return -11; }
    if (panel_width != 16 && panel_width != 32) { set_error("batched panel width must be 16 or 32"); return -12; }
    if (wgs_per_cu < 1 || wgs_per_cu > 2) { set_error("batched workgroups per CU must be 1 or 2"); return -13; }
    std::vector<Desc> desc(count);
    std::vector<int> order(count);
    for (int64_t i = 0; i < count; ++i) {
        if (n[i] < 0 || n[i] > NMAX) { set_error("batched systems need 0 <= n <= 512"); return -2; }
        if (k[i] < 0 || k[i] > MAXRHS) { set_error("batched systems need 0 <= k <= 8"); return -3; }
        if (a_off[i] < 0 || b_off[i] < 0 || p_off[i] < 0) { set_error("batched offsets must be >= 0"); return -4; }
        desc[i] = Desc{n[i], k[i], a_off[i], b_off[i], p_off[i]};
        if (n[i] > 0 && (!a || !perm)) { set_error("batched systems need A and perm buffers"); return !a ? -7 : -9; }
        if (n[i] * k[i] > 0 && !b) { set_error("batched right-hand sides need a B buffer"); return -8; }
    }
    if (count > 0 && mode == FACTOR_SOLVE && !regular) { set_error("batched plans need a regular-flag buffer"); return -10; }
    // Longest processing time first: the factorization's n³ and the solve's n² k.
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        const double fx = mode == FACTOR_SOLVE ? (double)n[x] : 0.0, fy = mode == FACTOR_SOLVE ? (double)n[y] : 0.0;
        const double cx = (double)n[x] * n[x] * (fx + 3.0 * k[x]), cy = (double)n[y] * n[y] * (fy + 3.0 * k[y]);
        return cx > cy;
    });
    Plan* p = new Plan;
    p->count = (int)count; p->f64 = f64; p->mode = mode; p->a = a; p->b = b; p->perm = (long long*)perm;
    p->regular = regular;
    p->grid = device_cus() * wgs_per_cu;
    p->bw = panel_width;
    p->wgs = wgs_per_cu;
    const size_t cnt = count > 0 ? (size_t)count : 1;
    if (hipMalloc(&p->d_desc, sizeof(Desc) * cnt) != hipSuccess || hipMalloc(&p->d_order, sizeof(int) * cnt) != hipSuccess ||
        hipMalloc(&p->d_next, sizeof(int)) != hipSuccess) {
        hipFree(p->d_desc); hipFree(p->d_order);
        delete p;
        return hip_status(hipErrorOutOfMemory, "hipMalloc");
    }
    if (count > 0) {
        HS_HIP(hipMemcpy(p->d_desc, desc.data(), sizeof(Desc) * count, hipMemcpyHostToDevice));
        HS_HIP(hipMemcpy(p->d_order, order.data(), sizeof(int) * count, hipMemcpyHostToDevice));
    }
    std::lock_guard<std::mutex> lk(g_plans_mu);
    *handle = g_next_plan++;
    g_plans[*handle] = p;
    return OK;
}

int batched_ir_plan(int64_t count, const int64_t* n, const int64_t* k, const int64_t* a_off, const int64_t* b_off,
                    const int64_t* p_off, const double* a, const double* b, double* x, float* a32, float* r32,
                    int64_t* perm, int* regular, int* iters, double* berr, int max_iter, double tol_per_n,
                    int panel_width, int wgs_per_cu, long long* handle) {
    if (max_iter < 0 || max_iter > 100) { set_error("batched refinement: max_iter must be in 0..100"); return -16; }
    if (!(tol_per_n > 0)) { set_error("batched refinement: the tolerance factor must be > 0"); return -17; }
    for (int64_t i = 0; i < count; ++i)
        if (n[i] > 0 && (!a || !x || !iters || !berr)) { set_error("batched refinement needs a, x, iters and berr"); return -7; }
    long long h = 0;
    int rc = batched_plan(count, n, k, a_off, b_off, p_off, false, a32, r32, perm, regular, FACTOR_SOLVE, panel_width,
                          wgs_per_cu, &h);
    if (rc) return rc;
    Plan* p = find_plan(h);
    p->ir = true;
    p->a64 = a; p->b64 = b; p->x64 = x; p->iters = iters; p->berr = berr;
    p->max_iter = max_iter; p->tol_per_n = tol_per_n;
    const size_t cnt = count > 0 ? (size_t)count : 1;
    if (hipMalloc(&p->d_anorm, sizeof(double) * cnt) != hipSuccess || hipMalloc(&p->d_rscale, sizeof(double) * cnt) != hipSuccess ||
        hipMalloc(&p->d_active, sizeof(int) * cnt) != hipSuccess) {
        batched_free(h);
        return hip_status(hipErrorOutOfMemory, "hipMalloc");
    }
    *handle = h;
    return OK;
}

int batched_run(long long handle, bool use_graph, hipStream_t s) {
    Plan* p = find_plan(handle);
    if (!p) { set_error("unknown batched plan"); return -1; }
    if (!use_graph || g_bprof_on) {
        long long* prof = g_bprof_on ? g_bprof : nullptr;
        return p->f64 ? enqueue<double>(p, s, prof) : enqueue<float>(p, s, prof);
    }
    if (!p->exec) {
        // The queue reset and the persistent launch, captured once; every run replays the graph.
        hipStream_t cap;
        HS_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        HS_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
        const int rc = p->f64 ? enqueue<double>(p, cap, nullptr) : enqueue<float>(p, cap, nullptr);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(cap, &g);
        hipStreamDestroy(cap);
        if (rc) { if (g) hipGraphDestroy(g); return rc; }
        HS_HIP(e);
        p->graph = g;
        HS_HIP(hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0));
    }
    HS_HIP(hipGraphLaunch(p->exec, s));
    return OK;
}

int batched_profile(int on, long long* out) {
    if (!g_bprof) HS_HIP(hipMalloc(&g_bprof, 8 * sizeof(long long)));
    HS_HIP(hipDeviceSynchronize());
    if (out) HS_HIP(hipMemcpy(out, g_bprof, 8 * sizeof(long long), hipMemcpyDeviceToHost));
    if (on) HS_HIP(hipMemset(g_bprof, 0, 8 * sizeof(long long)));
    g_bprof_on = on != 0;
    return OK;
}

int batched_free(long long handle) {
    Plan* p;
    {
        std::lock_guard<std::mutex> lk(g_plans_mu);
        auto it = g_plans.find(handle);
        if (it == g_plans.end()) { set_error("unknown batched plan"); return -1; }
        p = it->second;
        g_plans.erase(it);
    }
    if (p->exec) hipGraphExecDestroy(p->exec);
    if (p->graph) hipGraphDestroy(p->graph);
    hipFree(p->d_desc); hipFree(p->d_order); hipFree(p->d_next);
    if (p->d_anorm) hipFree(p->d_anorm);
    if (p->d_rscale) hipFree(p->d_rscale);
    if (p->d_active) hipFree(p->d_active);
    delete p;
    return OK;
}

}  // namespace hs
