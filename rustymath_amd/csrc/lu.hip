// LU with partial pivoting on gfx950, and the solves with its factors.
//
// The factorization is right-looking over panels of at most 64 columns. rustymath's lu_panel
// (core/src/linalg/factor.rs:580-615) recurses to 16-column leaves, but every item of its result takes
// its updates in order of the pivot column, one fused multiply-add each, with l = a · (1/pivot) (a
// quotient for a subnormal pivot, :506), and a zero pivot skips only the rest of its own leaf (:606).
// This file computes exactly that arithmetic: panels are unions of whole reference leaves, the panel
// kernel updates in registers in column order, the block triangular solve and the MFMA trailing product
// add their terms in order of p. So the factors and pivots are the reference's, bit for bit, wherever
// the MFMA unit rounds each step as a fused multiply-add.
//
// Device layout: the matrix is factored row-major in place (rows contiguous, so row swaps and panel
// loads coalesce) and transposed in place at the end into the reference's column-major `factors`
// ("row j holds column j", src/linalg/stacked.rs:25).
#include "hs_kernels.h"
#include "hs_pivot.h"
#include "gemm_tile.h"
#include <hip/hip_ext.h>
#include <vector>
#include <mutex>
#include <algorithm>

namespace hs {

namespace {

constexpr int PW = 64;            // widest panel
constexpr int PT = 512;           // rows of a panel workgroup, one a thread
constexpr int SG = 16;            // columns a panel workgroup stages through LDS at a time, loading and storing its rows
constexpr int STOP_AHEAD = 8;     // columns before a panel's end at which its workers stop taking tiles
constexpr int MAX_WG = 32;        // panel workgroups: one per CU of one XCD, one flag per lane of a wave
constexpr uint64_t SPIN_LIMIT = 1ull << 22;
constexpr int ARRIVE_STRIPS = 8;  // column strips of a panel's in-block update (at most 192 columns, 32 a strip)

// The FP32 panel workgroups have a ninth wave that holds no row and only polls the exchange, so that no wave waits for
// the other workgroups' keys and pivot row behind its own share of the column's update. The FP64 panel's registers
// leave room for two waves a SIMD only.
template <typename T> constexpr bool own_poller() { return sizeof(T) == 4; }
template <typename T> constexpr int panel_threads() { return PT + (own_poller<T>() ? 64 : 0); }

template <typename T> struct Bits;
template <> struct Bits<double> { typedef unsigned long long u; };
template <> struct Bits<float> { typedef unsigned int u; };

// The exchange between the workgroups of one panel. Every word a workgroup publishes carries a tag (the panel's
// launch number within the factorization times 64, plus the column, plus one) beside its payload, so that one
// device-coherent load both tells that a word is current and brings its value: no flag, no fence, no wait for
// store completion. Per column parity and workgroup: the key (FP32: one word, tag << 46 | kv << 15 | 32767 - pos;
// FP64: three words, tag << 32 with the high and low halves of kv and with 32767 - pos) and the candidate row
// (one word per FP32 item, tag << 32 | bits; two per FP64 item, tag << 32 | half). The buffer is cleared when a
// factorization starts; tags within it never repeat.
constexpr int KEY_WORDS = 4;
constexpr int ROW_WORDS = 2 * PW;
struct Exchange {
    unsigned long long key[2][MAX_WG][KEY_WORDS];
    unsigned long long row[2][MAX_WG][ROW_WORDS];
};

// The workgroups of a panel (launch indices 0, 8, 16, ...) run on one XCD under round-robin dispatch (which XCD
// varies from launch to launch) and share its L2. A word is published with a store that carries no cache
// bits (the wavefront-scope atomic store): it passes the writer's write-through L1 into that L2 and stays there, where
// the readers' agent-scope loads (which miss every L1) find it. An agent-scope store would also write the line
// through to memory first: measured on one XCD, a round trip between two workgroups takes 0.46 us this way against
// 1.1 us with agent-scope stores. Where the host has not found launch indices 0, 8, 16, ... on one XCD (or the caller
// asked for it), `agent` is set and every word goes out at agent scope, which any XCD's loads see.
__device__ __forceinline__ void xput(unsigned long long* p, unsigned long long v, bool agent) {
    if (agent) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}
__device__ __forceinline__ unsigned long long xget(const unsigned long long* p) {
    return __hip_atomic_load(const_cast<unsigned long long*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename T>
__device__ __forceinline__ void put_key(unsigned long long* k, unsigned tag, unsigned long long kv, int kp, bool agent) {
    const unsigned long long pp = (unsigned long long)(32767 - (kp < 32767 ? kp : 32767));
    if constexpr (sizeof(T) == 4) {
        xput(&k[0], ((unsigned long long)tag << 46) | (kv << 15) | pp, agent);
    } else {
        const unsigned long long t = (unsigned long long)tag << 32;
        xput(&k[0], t | (kv >> 32), agent);
        xput(&k[1], t | (kv & 0xffffffffull), agent);
        xput(&k[2], t | pp, agent);
    }
}
// A key's words as loaded (FP32: w[0]; FP64: w[0..2]), and their decoding: false while any word is older than `tag`.
template <typename T>
__device__ __forceinline__ void key_words(const unsigned long long* k, unsigned long long* w) {
    w[0] = xget(&k[0]);
    if constexpr (sizeof(T) == 8) { w[1] = xget(&k[1]); w[2] = xget(&k[2]); }
}
template <typename T>
__device__ __forceinline__ bool key_decode(const unsigned long long* w, unsigned tag, unsigned long long& kv, int& kp) {
    if constexpr (sizeof(T) == 4) {
        kv = (w[0] >> 15) & 0x7fffffffffull;
        const int pp = (int)(w[0] & 0x7fff);
        kp = pp == 0 ? 0x7fffffff : 32767 - pp;
        return (unsigned)(w[0] >> 46) == tag;
    } else {
        kv = ((w[0] & 0xffffffffull) << 32) | (w[1] & 0xffffffffull);
        const int pp = (int)(w[2] & 0x7fff);
        kp = pp == 0 ? 0x7fffffff : 32767 - pp;
        return (unsigned)(w[0] >> 32) == tag && (unsigned)(w[1] >> 32) == tag && (unsigned)(w[2] >> 32) == tag;
    }
}
// Reads a key; false while any of its words is older than `tag`.
template <typename T>
__device__ __forceinline__ bool get_key(const unsigned long long* k, unsigned tag, unsigned long long& kv, int& kp) {
    unsigned long long w[3];
    key_words<T>(k, w);
    return key_decode<T>(w, tag, kv, kp);
}
template <typename T>
__device__ __forceinline__ void put_item(unsigned long long* r, int j, unsigned tag, T v, bool agent) {
    const unsigned long long t = (unsigned long long)tag << 32;
    if constexpr (sizeof(T) == 4) {
        xput(&r[j], t | __builtin_bit_cast(unsigned, v), agent);
    } else {
        const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
        xput(&r[2 * j], t | (u & 0xffffffffull), agent);
        xput(&r[2 * j + 1], t | (u >> 32), agent);
    }
}
template <typename T>
__device__ __forceinline__ bool get_item(const unsigned long long* r, int j, unsigned tag, T& v) {
    if constexpr (sizeof(T) == 4) {
        const unsigned long long w = xget(&r[j]);
        v = __builtin_bit_cast(T, (unsigned)w);
        return (unsigned)(w >> 32) == tag;
    } else {
        const unsigned long long w0 = xget(&r[2 * j]), w1 = xget(&r[2 * j + 1]);
        v = __builtin_bit_cast(T, (w1 << 32) | (w0 & 0xffffffffull));
        return (unsigned)(w0 >> 32) == tag && (unsigned)(w1 >> 32) == tag;
    }
}

// A trailing product left to run beside the panels (look-ahead): C (m×n) less A (m×k) B (k×n), all row-major, in
// 128x128 tiles numbered as tile::grouped numbers them. Workers take tiles in order of number from *next; the panel's
// workgroup 0 writes its launch tag to *stop when its columns are done, after which no worker takes another tile, and
// trail_rest computes the tiles from *next on. C = nullptr: no product.
template <typename T>
struct TrailJob {
    int64_t m, n, k;
    const T* A;
    int64_t lda;
    const T* B;
    int64_t ldb;
    T* C;
    int64_t ldc;
    int tiles_m, tiles_n;
    int* next;
    unsigned* stop;
    int* log;  // trail_rest writes the first tile it computes here, where not null (profiling)
};

// The panel kernel's LDS, and the workers', in one allocation.
template <typename T>
struct PanelLds {
    T s_stage[PT][SG + 1];
    unsigned long long s_kv[2][PT / 64];
    int s_kp[2][PT / 64];
    unsigned long long s_best[2];  // FP32: the workgroup's best key and position, packed, by column parity
    T s_piv[2][PW];  // the pivot rows of the last two columns
    T s_rcp[2];      // their pivots' reciprocals
    alignas(16) T s_hand[PT / 64][PW + (sizeof(T) == 4 ? 4 : 2)];  // rows of 16-byte multiples: b128 writes
    T s_u11[PW][PW];
    int s_p[2], s_abort, s_skip[PW], s_dest[PT];
};
template <typename T>
struct WorkLds {
    T As[2][tile::BK][tile::BM + tile::Pad<T>::v];
    T Bs[2][tile::BK][tile::BN + tile::Pad<T>::v];
    int id;
};
template <typename T>
constexpr size_t lds_bytes() { return sizeof(WorkLds<T>) > sizeof(PanelLds<T>) ? sizeof(WorkLds<T>) : sizeof(PanelLds<T>); }

// A worker of the panel kernel (a workgroup of 512 threads off the panel's XCD): takes the next tile of the trailing
// product, all eight waves on it, until the tiles run out or the panel's columns are done, then returns.
template <typename T>
__device__ void trail_work(const TrailJob<T>& J, unsigned stop_tag, WorkLds<T>& W) {
    const int total = J.tiles_m * J.tiles_n;
    for (;;) {
        if (threadIdx.x == 0) {
            int id = -1;
            if (__hip_atomic_load(J.stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != stop_tag) {
                id = atomicAdd(J.next, 1);
                if (id >= total) id = -1;
            }
            W.id = id;
        }
        __syncthreads();
        const int id = W.id;
        __syncthreads();  // W.id is written again at the next round
        if (id < 0) return;
        int tm, tn;
        tile::grouped(id, J.tiles_m, J.tiles_n, tm, tn);
        const int64_t m0 = (int64_t)tm * tile::BM, n0 = (int64_t)tn * tile::BN;
        if (m0 + tile::BM <= J.m && n0 + tile::BN <= J.n)
            tile::gemm_tile<T, false, false, false, 8>(J.m, J.n, J.k, T(-1), J.A, J.lda, J.B, J.ldb, T(1), J.C, J.ldc, m0,
                                                       n0, W.As, W.Bs, threadIdx.x, true);
        else
            tile::gemm_tile<T, false, false, true, 8>(J.m, J.n, J.k, T(-1), J.A, J.lda, J.B, J.ldb, T(1), J.C, J.ldc, m0, n0,
                                                      W.As, W.Bs, threadIdx.x, true);
    }
}

// Launch index w -> item of [0, count), each XCD a contiguous run of ceil(count / 8); count (past the end) for the
// launch indices that a larger grid has beyond those runs.
__device__ __forceinline__ int xcd_run_of(int w, int count) {
    const int per = (count + 7) / 8;
    return w / 8 < per ? (w % 8) * per + w / 8 : count;
}

// The tiles of a trailing product that the panels' workers left: numbers [*next, total), split over the XCDs in
// contiguous runs. Launched on the same stream after the panels, so *next no longer moves.
template <typename T>
__global__ __launch_bounds__(256, 2) void trail_rest(TrailJob<T> J) {
    __shared__ T As[2][tile::BK][tile::BM + tile::Pad<T>::v];
    __shared__ T Bs[2][tile::BK][tile::BN + tile::Pad<T>::v];
    const int total = J.tiles_m * J.tiles_n;
    const int taken = __hip_atomic_load(J.next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int first = taken < total ? taken : total;
    if (J.log && blockIdx.x == 0 && threadIdx.x == 0) *J.log = first;
    const int id = first + xcd_run_of(blockIdx.x, total - first);
    if (id >= total) return;
    int tm, tn;
    tile::grouped(id, J.tiles_m, J.tiles_n, tm, tn);
    const int64_t m0 = (int64_t)tm * tile::BM, n0 = (int64_t)tn * tile::BN;
    if (m0 + tile::BM <= J.m && n0 + tile::BN <= J.n)
        tile::gemm_tile<T, false, false, false>(J.m, J.n, J.k, T(-1), J.A, J.lda, J.B, J.ldb, T(1), J.C, J.ldc, m0, n0, As,
                                                Bs, threadIdx.x, true);
    else
        tile::gemm_tile<T, false, false, true>(J.m, J.n, J.k, T(-1), J.A, J.lda, J.B, J.ldb, T(1), J.C, J.ldc, m0, n0, As,
                                               Bs, threadIdx.x, true);
}

// Wall-clock ticks (100 MHz) for the optional per-phase profile of the panel kernel.
__device__ __forceinline__ long long ticks() { return __builtin_amdgcn_s_memrealtime(); }

// Factors the panel of columns [c0, c0 + w) over rows [c0, c0 + rows) of the row-major `a`.
//
// One thread holds one row in registers, shifted so that v[0] is always the current column: at column c,
// v[j] is the row's item in column c + j. The rows never move between threads. A swap only relabels
// positions (each thread knows its row's current position `pos`); each multiplier goes to
// `scratch[column][row]` and each pivot row's U items to LDS (every workgroup sees every pivot row), from
// which each row is written once, at its final position, when the panel is done.
//
// A column's critical path is short: once the pivot row of column c is known, each row computes its multiplier
// and only its next item x0 (column c + 1); its pivot candidacy for column c + 1 goes out at once, and the
// rest of the row's update for column c runs while the candidates are compared. Per column: a wave reduction,
// a barrier, the workgroups' best keys compared, the winner's row fetched, a second barrier.
//
// With one workgroup, the winner's row goes through LDS. With several (launch indices 0, 8, 16, ..., one XCD
// under round-robin dispatch, one workgroup per CU, so all are resident), each workgroup publishes its best key
// and, from the thread holding it, its row as tagged words (Exchange); the last wave of every workgroup reads all keys,
// picks the winner and fetches its row. The launch indices on the other XCDs return at once, or, where `job`
// holds a trailing product, work on its tiles (trail_work) until the columns are done.
//
// Writes pivots[c0 + c] (absolute rows) and `moves` (count, then pairs dst, src: the rows whose position
// changed), for the row-swap kernel. `leaf_end[c]` is the end of the reference leaf of column c; a zero
// pivot leaves its column undivided and skips the rest of that leaf (factor.rs:606). Where `prof` is not
// null, workgroup 0 adds the ticks of each phase of the column loop into prof[0..7].
template <typename T>
__global__ __launch_bounds__(panel_threads<T>()) void panel_kernel(T* __restrict__ a, int64_t lda, int64_t c0, int64_t rows, int w,
                                                   int64_t* __restrict__ pivots, long long* __restrict__ moves,
                                                   const int* __restrict__ leaf_end, T* __restrict__ scratch,
                                                   Exchange* __restrict__ ex, unsigned tag0, int nwg, int* status,
                                                   long long* prof, TrailJob<T> job, int local) {
    __shared__ __attribute__((aligned(16))) unsigned char lds[lds_bytes<T>()];
    // After a wait that ran out in an earlier launch the factorization is void: the launch returns rather than wait
    // again. The panel's workgroups test the word once their rows are in (its load then adds no latency of its own).
    const int void_run = __hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (blockIdx.x & 7) {  // off the panel's XCD: a worker of the trailing product, if there is one
        // Eight waves compute a tile; a ninth returns here, and a barrier waits only for the waves that have not ended.
        if (job.C && !void_run && threadIdx.x < PT) trail_work<T>(job, tag0, *reinterpret_cast<WorkLds<T>*>(lds));
        return;
    }
    const int g = (int)(blockIdx.x >> 3);
    if (g >= nwg) return;
    // local = 0: words go out at agent scope, and where the workgroups run does not matter.
    const bool agent = !local;
    PanelLds<T>& L = *reinterpret_cast<PanelLds<T>*>(lds);
    if (threadIdx.x == 0) { L.s_best[0] = 0ull; L.s_best[1] = 0ull; }  // before the barriers of the rows' loads
    auto& s_stage = L.s_stage;
    auto& s_kv = L.s_kv;
    auto& s_kp = L.s_kp;
    auto& s_piv = L.s_piv;
    auto& s_rcp = L.s_rcp;
    auto& s_hand = L.s_hand;
    auto& s_u11 = L.s_u11;
    auto& s_p = L.s_p;
    auto& s_abort = L.s_abort;
    auto& s_skip = L.s_skip;
    auto& s_dest = L.s_dest;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    constexpr int WAVES = PT / 64;  // the waves that hold rows
    constexpr int POLL_WAVE = own_poller<T>() ? WAVES : WAVES - 1;
    const bool rower = tid < PT;  // false for the polling wave of its own
    typedef typename std::conditional<sizeof(T) == 4, unsigned, unsigned long long>::type Key;  // a candidate's key
    const bool timed = prof != nullptr && g == 0 && threadIdx.x == 0;
    long long tp[8] = {0, 0, 0, 0, 0, 0, 0, 0}, tl = timed ? ticks() : 0;
#define HS_PHASE(q) do { if (timed) { long long t_ = ticks(); tp[q] += t_ - tl; tl = t_; } } while (0)

    const int64_t phys = (int64_t)g * PT + (rower ? tid : 0);  // this thread's row, relative to c0
    const bool has = rower && phys < rows;
    int pos = (int)phys;
    // The row's items, every load issued before any is used, from addresses clamped into the panel (a load behind a
    // per-lane condition would be issued and waited for one at a time). Items past the panel's width only ever meet
    // other such items (an item in column j takes the pivot row's item in column j), are never written, and so may
    // hold any value.
    // The rows come in through LDS, 16 columns at a time, so that the global loads run along rows (a wave reads four
    // rows' 16 items per load) rather than one item of 64 rows per load.
    T v[PW];
    const int wg_rows = (int)(rows - (int64_t)g * PT < PT ? rows - (int64_t)g * PT : PT);
#pragma unroll
    for (int ch = 0; ch < PW / SG; ++ch) {
        if (rower) {
            T tmp[SG];
#pragma unroll
            for (int q = 0; q < SG; ++q) {
                const int idx = tid + PT * q, row = idx / SG, col = ch * SG + idx % SG;
                const int64_t r = (int64_t)g * PT + (row < wg_rows ? row : wg_rows - 1);
                tmp[q] = a[(c0 + r) * lda + c0 + (col < w ? col : w - 1)];
            }
            asm volatile("" ::: "memory");
#pragma unroll
            for (int q = 0; q < SG; ++q) {
                const int idx = tid + PT * q;
                s_stage[idx / SG][idx % SG] = tmp[q];
            }
        }
        __syncthreads();
#pragma unroll
        for (int jj = 0; jj < SG; ++jj) v[ch * SG + jj] = rower ? s_stage[tid][jj] : T(0);
        __syncthreads();
    }

    if (void_run) return;  // the same word for every workgroup of the launch, so all of them return
    if (tid == 0) s_abort = 0;
    if (tid < PW) s_skip[tid] = tid < w ? leaf_end[c0 + tid] - (int)c0 - tid : 0;  // a zero pivot's skip, by column
    const T tiny = sizeof(T) == 8 ? T(2.2250738585072014e-308) : T(1.17549435e-38f);

    // x0: this row's item in the current column, fully updated; the rest of v may still owe the previous column's
    // update (pending), with multiplier l_prev, the pivot row prev and the zero-pivot skip.
    T x0 = v[0], l_prev = T(0);
    const T* prev = s_piv[1];
    T* sc = scratch + (has ? phys : 0);  // this row's multiplier of the current column: scratch[c][row]
    bool pending = false;
    int skip_prev = 0;
    int col_skip = 0;  // the previous column's zero-pivot skip, the same for every row

    HS_PHASE(5);  // the rows' loads and the set-up
#pragma clang loop unroll(disable)
    for (int c = 0; c < w; ++c) {
        const int buf = c & 1;
        const int rem = w - c;  // items of the current row that lie in the panel
        const int leaf_skip = s_skip[c];  // read now, used once the pivot is known
        const unsigned tag = tag0 + (unsigned)c;
        // This row's candidacy for column c and the wave's best.
        const bool cand = has && pos >= c;
        Key kv = cand ? (Key)mag_key(x0) : Key(0);
        int kp = cand ? pos : 0x7fffffff, hl = 0;
        wave_best(kv, kp, hl);
        // The best of the waves' keys (larger key, then smaller position). FP32: a key and 2^31 - 1 less the position
        // pack into 64 bits whose maximum is the best, which each wave adds to an LDS word with one atomic maximum.
        // FP64 (64-bit keys): lane q takes wave q's key after the barrier, and one more wave reduction picks it.
        if constexpr (sizeof(T) == 4) {
            const unsigned long long pk = ((unsigned long long)kv << 32) | (unsigned)(0x7fffffff - kp);
            if (lane == 0 && rower && pk != 0ull) atomicMax(&L.s_best[buf], pk);
        } else {
            if (lane == 0 && rower) { s_kv[buf][wave] = kv; s_kp[buf][wave] = kp; }
        }
        __syncthreads();
        HS_PHASE(0);
        Key bk;
        int bkp;
        if constexpr (sizeof(T) == 4) {
            const unsigned long long pk = L.s_best[buf];
            bk = (Key)(pk >> 32);
            bkp = 0x7fffffff - (int)(unsigned)pk;
        } else {
            int bl = 0;
            bk = (Key)s_kv[buf][lane & (WAVES - 1)];
            bkp = s_kp[buf][lane & (WAVES - 1)];
            bk = lane < WAVES ? bk : Key(0);
            bkp = lane < WAVES ? bkp : 0x7fffffff;
            wave_best(bk, bkp, bl);
        }
        const unsigned long long bkv = bk;
        if (nwg > 1 && tid == 0) put_key<T>(ex->key[buf][g], tag, bkv, bkp, agent);

        // The workgroup's best row goes out before the rest of the update runs: the holding lane writes its items,
        // still owing the previous column's update, to its wave's LDS row, and lane j applies that update to item j
        // (one fused multiply-add, the same one the holder would apply) and publishes it. That update skips the items
        // that the previous column's zero pivot skips (the same for every row), whichever rows the lanes hold.
        auto hand_item = [&](const T* h) {
            if (c == 0) return h[lane];
            const int j = lane + 1 < PW ? lane + 1 : PW - 1;
            const T u = __builtin_elementwise_fma(h[PW + 1], -prev[j], h[j]);
            return lane == 0 ? h[PW] : (lane == PW - 1 ? T(0) : (j >= col_skip ? u : h[j]));
        };
        {
            const bool holder = has && pos == bkp;
            if (__ballot(holder)) {
                if (holder) {
                    // Only the items that lanes below rem read: v[j] for j <= rem, four at a time.
#pragma unroll
                    for (int j0 = 0; j0 < PW; j0 += 4) {
                        if (j0 <= rem) {
#pragma unroll
                            for (int j = j0; j < j0 + 4; ++j) s_hand[wave][j] = v[j];
                        }
                    }
                    s_hand[wave][PW] = x0;
                    s_hand[wave][PW + 1] = l_prev;
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                const T item = hand_item(s_hand[wave]);
                if (nwg > 1) {
                    if (lane < rem) put_item<T>(ex->row[buf][g], lane, tag, item, agent);
                } else {
                    s_piv[buf][lane] = item;
                    if (lane == 0) s_rcp[buf] = T(1) / item;
                }
            }
        }

        HS_PHASE(7);  // the workgroup's best key and row published
        // The previous column's update of the rest of the row, sixteen items at a time, only as far as the panel
        // reaches (items past it are never used).
        // The pivot row reaches every lane as scalars: one LDS read per wave (lane j takes item j, with every lane
        // active, so that each lane's register holds its item), then lane j's item is read out as a uniform operand,
        // rather than every lane reading every item from LDS.
        // The last wave's first look at the workgroups' keys goes out before the update, from an address clamped into
        // the buffer, and is decoded after it: a key published by then costs no round trip of its own.
        const bool poller = nwg > 1 && wave == POLL_WAVE;
        unsigned long long kw[3] = {0ull, 0ull, 0ull};
        // (A polling wave of its own has no update: its keys go out in its polling branch below, so that the load and
        // its use share one branch.)
        if (!own_poller<T>() && poller) key_words<T>(ex->key[buf][lane < nwg ? lane : 0], kw);
        T mine = prev[lane];
        asm volatile("" : "+v"(mine));  // loaded here, by every lane, not sunk into the branch below
        if (pending) {
            const T* pp = prev;
            if (skip_prev <= 2) {
#pragma unroll
                for (int j0 = 2; j0 < PW; j0 += 16) {
                    if (j0 <= rem) {
                        if constexpr (sizeof(T) == 4) {
                            // Four items are read out into four scalars before their FMAs, so that no FMA waits
                            // for the read just before it.
#pragma unroll
                            for (int j = j0; j < j0 + 16 && j < PW; j += 4) {
                                T u[4];
#pragma unroll
                                for (int q = 0; q < 4; ++q) u[q] = j + q < PW ? lane_value(mine, j + q) : T(0);
                                asm volatile("" ::"s"(u[0]), "s"(u[1]), "s"(u[2]), "s"(u[3]));
#pragma unroll
                                for (int q = 0; q < 4; ++q)
                                    if (j + q < PW) v[j + q - 1] = __builtin_elementwise_fma(l_prev, -u[q], v[j + q]);
                            }
                        } else {
#pragma unroll
                            for (int j = j0; j < j0 + 16 && j < PW; ++j)
                                v[j - 1] = __builtin_elementwise_fma(l_prev, -lane_value(mine, j), v[j]);
                        }
                    }
                }
            } else {  // a zero pivot: the rest of its leaf (at most 16 columns) keeps its items
#pragma unroll
                for (int j = 2; j < 16; ++j) v[j - 1] = j >= skip_prev ? __builtin_elementwise_fma(l_prev, -pp[j], v[j]) : v[j];
#pragma unroll
                for (int j = 16; j < PW; ++j) v[j - 1] = __builtin_elementwise_fma(l_prev, -pp[j], v[j]);
            }
            v[PW - 1] = T(0);
            pending = false;
        }
        v[0] = x0;
        HS_PHASE(1);

        if (nwg > 1) {
            // The last wave compares the workgroups' keys, then fetches the winner's row. Not wave 0: where pivots lie
            // on the diagonal (no row moves), the pivot row of every column of a panel sits in wave 0 of workgroup 0,
            // whose hand-off of that row would otherwise delay the poll behind it.
            if (poller) {
                unsigned long long gkv = 0ull;
                int gkp = 0x7fffffff, gl = 0;
                bool ok = true;
                if (own_poller<T>()) key_words<T>(ex->key[buf][lane < nwg ? lane : 0], kw);
                if (lane < nwg && !key_decode<T>(kw, tag, gkv, gkp)) {
                    uint64_t spins = 0;
                    while (!get_key<T>(ex->key[buf][lane], tag, gkv, gkp)) {
                        if (++spins > SPIN_LIMIT) { ok = false; break; }
                    }
                }
                bool all = __all(ok);
                Key gk = (Key)gkv;
                wave_best(gk, gkp, gl);
                T item = T(0);
                if (all && lane < rem) {
                    uint64_t spins = 0;
                    while (!get_item<T>(ex->row[buf][gl], lane, tag, item)) {
                        if (++spins > SPIN_LIMIT) { ok = false; break; }
                    }
                }
                all = __all(ok);
                s_piv[buf][lane] = item;
                if (lane == 0) {
                    s_rcp[buf] = T(1) / item;
                    s_p[buf] = gkp;
                    if (!all) { s_abort = 1; atomicExch(status, ERR_LIMIT); }
                }
            }
        } else if (tid == 0) {
            s_p[buf] = bkp;
        }
        HS_PHASE(2);
        __syncthreads();
        HS_PHASE(3);
        const int p = s_p[buf];
        const T* piv = s_piv[buf];
        const T pivot = piv[0], piv1 = piv[1], rcp = s_rcp[buf];
        const int abort_now = s_abort;
        asm volatile("" ::"v"(p), "v"(pivot), "v"(piv1), "v"(rcp));  // all five reads in one LDS round trip
        if (abort_now) return;

        // The winner takes position c; its row of U is kept in LDS. The row at position c moves to p.
        if (tid < PW) s_u11[c][tid] = piv[tid];
        if (sizeof(T) == 4 && tid == 0) L.s_best[buf] = 0ull;  // every wave has read it; next used at column c + 2
        if (has && pos == p) pos = c;
        else if (has && pos == c) pos = p;
        if (g == 0 && tid == 0) pivots[c0 + c] = c0 + p;

        // Rows below the pivot: the multiplier (a quotient for a subnormal pivot, the item itself for a zero one,
        // whose update then skips the rest of its leaf) and the next item; the rest of the row waits for the next
        // column's candidates to be out.
        if (has && pos > c) {
            const bool zero = pivot == T(0);
            const int skip = zero ? leaf_skip : 0;
            const T l = zero ? x0 : (fabs(pivot) >= tiny ? x0 * rcp : x0 / pivot);
            x0 = 1 >= skip ? __builtin_elementwise_fma(l, -piv1, v[1]) : v[1];
            *sc = l;
            l_prev = l;
            skip_prev = skip;
            pending = true;
        }
        col_skip = pivot == T(0) ? leaf_skip : 0;
        prev = piv;
        sc += rows;
        // The workers take no new tile once STOP_AHEAD columns remain, so that the tiles in progress end about when
        // the columns do.
        if (job.C && g == 0 && tid == 0 && c == (w > STOP_AHEAD ? w - 1 - STOP_AHEAD : 0))
            __hip_atomic_store(job.stop, tag0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        HS_PHASE(4);
    }
    __syncthreads();  // every row of U is in s_u11

    // The row at its final position: multipliers from scratch left of it, its row of U (pivot rows) from LDS; the
    // scratch items are all loaded before any is used (as the first loads). The rows go out through LDS, 16 columns
    // at a time, each wave storing four rows' 16 items per store.
    {
#pragma unroll
        for (int j = 0; j < PW; ++j) v[j] = scratch[(int64_t)(j < w ? j : w - 1) * rows + (has ? phys : 0)];
        asm volatile("" ::: "memory");
        const int upos = pos < PW ? pos : PW - 1;
#pragma unroll
        for (int j = 0; j < PW; ++j) v[j] = j < pos ? v[j] : s_u11[upos][(j - pos) & (PW - 1)];
        if (rower) s_dest[tid] = pos;
#pragma unroll
        for (int ch = 0; ch < PW / SG; ++ch) {
            if (rower) {
#pragma unroll
                for (int jj = 0; jj < SG; ++jj) s_stage[tid][jj] = v[ch * SG + jj];
            }
            __syncthreads();
            if (rower) {
#pragma unroll
                for (int q = 0; q < SG; ++q) {
                    const int idx = tid + PT * q, row = idx / SG, col = ch * SG + idx % SG;
                    if (row < wg_rows && col < w) a[(c0 + s_dest[row]) * lda + c0 + col] = s_stage[row][idx % SG];
                }
            }
            __syncthreads();
        }
        if (has && pos != phys) {
            const long long k = atomicAdd((unsigned long long*)moves, 1ull);
            moves[1 + 2 * k] = c0 + pos;
            moves[2 + 2 * k] = c0 + phys;
        }
    }
    if (timed) {
        __builtin_amdgcn_s_waitcnt(0);  // the row's stores issued above have completed
        HS_PHASE(6);                    // the rows written back at their final positions
        for (int q = 0; q < 8; ++q) atomicAdd((unsigned long long*)&prof[q], (unsigned long long)tp[q]);
    }
#undef HS_PHASE
}

// Applies the net row permutations of `count` consecutive panels (move lists `stride` words apart), in order, to the
// columns [lo, hi) outside [x0, x1) (x0 <= x1 within [lo, hi)): every moved row takes the old item of its source row.
// Each block of 64 columns reads all of a list's sources into LDS before writing, so chains of swaps cost one read and
// one write per moved item; the block's barrier orders one list's writes before the next list's reads.
template <typename T>
__global__ __launch_bounds__(256) void swap_kernel(T* __restrict__ a, int64_t lda, int64_t lo, int64_t hi, int64_t x0,
                                                   int64_t x1, const long long* __restrict__ moves, int64_t stride,
                                                   int count) {
    __shared__ T tmp[2 * PW][64];
    __shared__ long long mv[1 + 4 * PW];
    const int tid = threadIdx.x, col_l = tid & 63, grp = tid >> 6;
    int64_t col = lo + (int64_t)blockIdx.x * 64 + col_l;  // over the columns of [lo, hi) outside [x0, x1)
    if (col >= x0) col += x1 - x0;
    const bool live = col < hi;
    const int64_t lc = live ? col : 0;
    for (int list = 0; list < count; ++list) {
        __syncthreads();  // the previous list's reads of mv and tmp, and its writes, are done
        for (int q = threadIdx.x; q < 1 + 4 * PW; q += blockDim.x) mv[q] = moves[list * stride + q];
        __syncthreads();
        const int nm = (int)mv[0];
        if (nm == 0) continue;
        // All sources in flight at once: past the last move, a load repeats the first source (addresses stay inside
        // the matrix and no load waits behind a per-lane condition).
        T vals[2 * PW / 4];
#pragma unroll
        for (int q = 0; q < 2 * PW / 4; ++q) {
            const int k = grp + 4 * q;
            vals[q] = a[mv[2 + 2 * (k < nm ? k : 0)] * lda + lc];
        }
        asm volatile("" ::: "memory");
#pragma unroll
        for (int q = 0; q < 2 * PW / 4; ++q) {
            const int k = grp + 4 * q;
            if (k < nm) tmp[k][col_l] = vals[q];
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 2 * PW / 4; ++q) {
            const int k = grp + 4 * q;
            if (k < nm && live) a[mv[1 + 2 * k] * lda + col] = tmp[k][col_l];
        }
    }
}

// Columns per workgroup of the kernels below that hold a strip of columns of up to 256 rows in LDS.
template <typename T> struct RowsCols;
template <> struct RowsCols<double> { static constexpr int v = 32; };
template <> struct RowsCols<float> { static constexpr int v = 64; };
// Columns per workgroup of trsm_rows_kernel: 32, so that at n = 16384 about two workgroups share each CU and one's
// substitution steps overlap the other's.
constexpr int ROWS_TCOL = 32;

// The update of a block's columns [c1, ce) after its panel [c0, c1 = c0 + w), in one launch: U12 = L11⁻¹ A12 for the
// panel's rows, then A22 -= L21 U12 for the rows [c1, n). Workgroup (bi, bj) takes rows c1 + 64 bi .. + 63 and columns
// c1 + TCOL bj .. + TCOL - 1. It solves its columns' U12 in LDS in four steps of 16 rows: the step's rows substitute
// among themselves (one lane per column, in order of the pivot column), then the rows below take the step's terms on
// MFMA (16x16x4 steps, each rounding its four terms as fused multiply-adds in order), its L21 and A22 tiles in flight
// meanwhile. Then each wave takes 16 rows of A22 on MFMA with k over the panel's columns in order (terms past w are
// (+0)·(−0) = −0, which changes no sum). Every workgroup of a column strip reads the panel's rows of that strip and
// solves the same U12; the last of them to have read those rows (a count per strip in `arrived`, cleared when the
// factorization starts) writes U12 back, so that no workgroup reads a row another has already overwritten. Every item
// takes its terms in order of the pivot column, one fused multiply-add each: the triangular solve and product this
// replaces, bit for bit.
template <typename T>
__global__ __launch_bounds__(256, sizeof(T) == 4 ? 2 : 1) void panel_update_kernel(T* __restrict__ a, int64_t lda, int64_t c0, int w, int64_t ce,
                                                           int64_t n, int* __restrict__ arrived) {
    constexpr int TCOL = RowsCols<T>::v;
    __shared__ T Ls[PW][PW + 1];    // strictly lower part of L11, zeros elsewhere
    __shared__ T Xs[PW][TCOL + 1];  // U12 of this workgroup's columns
    typedef typename Mfma<T>::acc Acc;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int64_t c1 = c0 + w, col0 = c1 + (int64_t)blockIdx.y * TCOL, rt = c1 + (int64_t)blockIdx.x * 64;
    const int64_t rw = rt + 16 * wave;  // this wave's 16 rows
    // L21 (A operand: lane holds row rw + lane % 16, columns c0 + 4 ks + lane / 16) and A22 (the accumulators), from
    // addresses clamped into the matrix.
    T av[PW / 4];
    Acc acc[TCOL / 16];
    {
        const int64_t ar = rw + (lane & 15) < n ? rw + (lane & 15) : n - 1;
#pragma unroll
        for (int ks = 0; ks < PW / 4; ++ks) {
            const int k = 4 * ks + (lane >> 4);
            av[ks] = a[ar * lda + c0 + (k < w ? k : w - 1)];
        }
#pragma unroll
        for (int cb = 0; cb < TCOL / 16; ++cb)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int64_t r = rw + Mfma<T>::row(lane, e), c = col0 + 16 * cb + (lane & 15);
                acc[cb][e] = a[(r < n ? r : n - 1) * lda + (c < ce ? c : ce - 1)];
            }
    }
    {
        T lv[PW * PW / 256], xv[PW * TCOL / 256];
#pragma unroll
        for (int q = 0; q < PW * PW / 256; ++q) {
            const int idx = t + 256 * q, r = idx / PW, c = idx % PW;
            lv[q] = a[(c0 + (r < w ? r : w - 1)) * lda + c0 + (c < w ? c : w - 1)];
        }
#pragma unroll
        for (int q = 0; q < PW * TCOL / 256; ++q) {
            const int idx = t + 256 * q, r = idx / TCOL, c = idx % TCOL;
            const int64_t col = col0 + c < ce ? col0 + c : ce - 1;
            xv[q] = a[(c0 + (r < w ? r : w - 1)) * lda + col];
        }
        asm volatile("" ::: "memory");
#pragma unroll
        for (int q = 0; q < PW * PW / 256; ++q) {
            const int idx = t + 256 * q, r = idx / PW, c = idx % PW;
            Ls[r][c] = (r < w && c < r) ? lv[q] : T(0);
        }
#pragma unroll
        for (int q = 0; q < PW * TCOL / 256; ++q) {
            const int idx = t + 256 * q, r = idx / TCOL, c = idx % TCOL;
            Xs[r][c] = r < w ? xv[q] : T(-0.0);  // the padding rows: −0, against +0 on the L side
        }
    }
    __shared__ int s_last;
    __syncthreads();  // every thread's loads of the panel's rows have returned (their values are in LDS)
    if (t == 0) s_last = atomicAdd(&arrived[blockIdx.y], 1) == (int)gridDim.x - 1;
    for (int kb = 0; kb < PW / 16 && 16 * kb < w; ++kb) {
        const int k0 = 16 * kb;
        if (wave == 0 && lane < TCOL) {  // the 16 rows among themselves; lane = column
            T x[16];
#pragma unroll
            for (int i = 0; i < 16; ++i) x[i] = Xs[k0 + i][lane];
#pragma unroll
            for (int q = 0; q < 16; ++q)
#pragma unroll
                for (int r = q + 1; r < 16; ++r) x[r] = __builtin_elementwise_fma(-Ls[k0 + r][k0 + q], x[q], x[r]);
#pragma unroll
            for (int i = 0; i < 16; ++i) if (k0 + i < w) Xs[k0 + i][lane] = x[i];
        }
        __syncthreads();
        const int r0 = 16 * wave;  // the rows below, one 16-row block per wave, on MFMA
        if (wave > kb && r0 < w) {
#pragma unroll
            for (int cb = 0; cb < TCOL / 16; ++cb) {
                Acc u;
#pragma unroll
                for (int e = 0; e < 4; ++e) u[e] = Xs[r0 + Mfma<T>::row(lane, e)][16 * cb + (lane & 15)];
#pragma unroll
                for (int st = 0; st < 4; ++st)
                    u = Mfma<T>::step(-Ls[r0 + (lane & 15)][k0 + 4 * st + (lane >> 4)],
                                      Xs[k0 + 4 * st + (lane >> 4)][16 * cb + (lane & 15)], u);
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int r = r0 + Mfma<T>::row(lane, e);
                    if (r < w) Xs[r][16 * cb + (lane & 15)] = u[e];
                }
            }
        }
        __syncthreads();
    }
    if (s_last) {  // U12 back to the panel's rows, by the strip's last workgroup to have read them
#pragma unroll
        for (int q = 0; q < PW * TCOL / 256; ++q) {
            const int idx = t + 256 * q, r = idx / TCOL, c = idx % TCOL;
            if (r < w && col0 + c < ce) a[(c0 + r) * lda + col0 + c] = Xs[r][c];
        }
    }
    // A22 -= L21 U12: k over the panel's columns in order; past w the L side is +0 and the U side −0.
#pragma unroll
    for (int ks = 0; ks < PW / 4; ++ks) {
        if (4 * ks < w) {
            const int k = 4 * ks + (lane >> 4);
            const T l = k < w ? -av[ks] : T(0);
#pragma unroll
            for (int cb = 0; cb < TCOL / 16; ++cb) acc[cb] = Mfma<T>::step(l, Xs[k][16 * cb + (lane & 15)], acc[cb]);
        }
    }
#pragma unroll
    for (int cb = 0; cb < TCOL / 16; ++cb)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int64_t r = rw + Mfma<T>::row(lane, e), c = col0 + 16 * cb + (lane & 15);
            if (r < n && c < ce) a[r * lda + c] = acc[cb][e];
        }
}

// U12 = L11⁻¹ A12 for a whole block of rows [r0, r0 + m), m <= 256, columns [cs, n): the rows of U right of a block,
// in one launch rather than a product and a triangular solve per panel. A workgroup keeps TCOL columns of all m rows in
// LDS and walks the rows in 16-row steps: the step's rows substitute among themselves (one lane per column, in order
// of the pivot column), then every row below takes the step's 16 terms on MFMA (16x16x4 steps, each rounding its four
// terms as fused multiply-adds in order). L11 streams through LDS one 16-column strip per step, the next strip's loads
// in flight during the step. Every item takes its terms in order of the pivot column, one fused multiply-add each,
// so the result is the per-panel product and solve's, bit for bit, however the rows group into steps.
template <typename T>
__global__ __launch_bounds__(256, 2) void trsm_rows_kernel(T* __restrict__ a, int64_t lda, int64_t r0, int m, int64_t cs,
                                                        int64_t n) {
    constexpr int TCOL = ROWS_TCOL, MR = 4 * PW;
    __shared__ T Xs[MR][TCOL + 1];
    __shared__ T Lst[2][MR][17];  // Lst[b][r][q] = L(r, 16 kb + q) below the diagonal, 0 elsewhere
    typedef typename Mfma<T>::acc Acc;
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int64_t col0 = cs + (int64_t)blockIdx.x * TCOL;
    const int nkb = (m + 15) / 16;
    T sva[16], svb[16];  // strips in flight: each step issues the loads of the strip two steps ahead
    // Strip kb (rows 0 .. 255, columns 16 kb .. +15): item i of thread t is row t / 16 + 16 i, column t % 16, so that a
    // wave's load reads four rows' 16 items (addresses clamped into the block).
    auto load_strip = [&](int kb, T* sv) {
        const int p = 16 * kb + (t & 15), pc = p < m ? p : m - 1;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int r = (t >> 4) + 16 * i;
            sv[i] = a[(r0 + (r < m ? r : m - 1)) * lda + r0 + pc];
        }
    };
    auto store_strip = [&](int kb, const T* sv, int buf) {
        const int q = t & 15, p = 16 * kb + q;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int r = (t >> 4) + 16 * i;
            Lst[buf][r][q] = (r < m && p < r) ? sv[i] : T(0);
        }
    };
    load_strip(0, sva);
    if (nkb > 1) load_strip(1, svb);
    {
        T xv[MR * TCOL / 256];
#pragma unroll
        for (int q = 0; q < MR * TCOL / 256; ++q) {
            const int idx = t + 256 * q, r = idx / TCOL, c = idx % TCOL;
            const int64_t col = col0 + c < n ? col0 + c : n - 1;
            xv[q] = a[(r0 + (r < m ? r : m - 1)) * lda + col];
        }
        asm volatile("" ::: "memory");
#pragma unroll
        for (int q = 0; q < MR * TCOL / 256; ++q) {
            const int idx = t + 256 * q, r = idx / TCOL, c = idx % TCOL;
            Xs[r][c] = xv[q];
        }
    }
    store_strip(0, sva, 0);
    __syncthreads();
    // Step kb: strip kb is in Lst[kb % 2]; `mine` held strip kb (free now) and takes strip kb + 2, `next` holds strip
    // kb + 1, which goes to LDS at the step's end.
    auto step = [&](int kb, T* mine, const T* next) {
        const int buf = kb & 1, k0 = 16 * kb;
        if (kb + 2 < nkb) load_strip(kb + 2, mine);
        if (wave == 0 && lane < TCOL) {  // the step's rows among themselves; lane = column
            T x[16];
#pragma unroll
            for (int i = 0; i < 16; ++i) x[i] = Xs[k0 + i][lane];
#pragma unroll
            for (int q = 0; q < 16; ++q)
#pragma unroll
                for (int r = q + 1; r < 16; ++r) x[r] = __builtin_elementwise_fma(-Lst[buf][k0 + r][q], x[q], x[r]);
#pragma unroll
            for (int i = 0; i < 16; ++i) Xs[k0 + i][lane] = x[i];
        }
        __syncthreads();
        for (int rb = kb + 1 + wave; rb < nkb; rb += 4) {  // the rows below, 16 at a time, on MFMA
            T av[4];
#pragma unroll
            for (int st = 0; st < 4; ++st) av[st] = -Lst[buf][16 * rb + (lane & 15)][4 * st + (lane >> 4)];
#pragma unroll
            for (int cb = 0; cb < TCOL / 16; ++cb) {
                Acc acc;
#pragma unroll
                for (int e = 0; e < 4; ++e) acc[e] = Xs[16 * rb + Mfma<T>::row(lane, e)][16 * cb + (lane & 15)];
#pragma unroll
                for (int st = 0; st < 4; ++st)
                    acc = Mfma<T>::step(av[st], Xs[k0 + 4 * st + (lane >> 4)][16 * cb + (lane & 15)], acc);
#pragma unroll
                for (int e = 0; e < 4; ++e) Xs[16 * rb + Mfma<T>::row(lane, e)][16 * cb + (lane & 15)] = acc[e];
            }
        }
        if (kb + 1 < nkb) store_strip(kb + 1, next, buf ^ 1);
        __syncthreads();
    };
    for (int kb = 0; kb < nkb; kb += 2) {
        step(kb, sva, svb);
        if (kb + 1 < nkb) step(kb + 1, svb, sva);
    }
#pragma unroll
    for (int q = 0; q < MR * TCOL / 256; ++q) {
        const int idx = t + 256 * q, r = idx / TCOL, c = idx % TCOL;
        if (r < m && col0 + c < n) a[(r0 + r) * lda + col0 + c] = Xs[r][c];
    }
}

template <typename T>
__global__ void transpose_square_inplace(T* a, int64_t lda, int64_t n) {
    __shared__ T t1[32][33], t2[32][33];
    const int bi = blockIdx.y, bj = blockIdx.x;
    if (bj < bi) return;
    const int tx = threadIdx.x, ty = threadIdx.y;  // 32 x 8
    const int64_t i0 = (int64_t)bi * 32, j0 = (int64_t)bj * 32;
    for (int r = ty; r < 32; r += 8) {
        int64_t i = i0 + r, j = j0 + tx;
        if (i < n && j < n) t1[r][tx] = a[i * lda + j];
        int64_t i2 = j0 + r, j2 = i0 + tx;
        if (i2 < n && j2 < n) t2[r][tx] = a[i2 * lda + j2];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        int64_t i = i0 + r, j = j0 + tx;           // block (bi, bj) gets transpose of block (bj, bi)
        if (i < n && j < n) a[i * lda + j] = t2[tx][r];
        int64_t i2 = j0 + r, j2 = i0 + tx;
        if (i2 < n && j2 < n) a[i2 * lda + j2] = t1[tx][r];
    }
}

template <typename T>
__global__ void transpose_kernel(int64_t rows, int64_t cols, const T* __restrict__ src, int64_t lds, T* __restrict__ dst, int64_t ldd) {
    __shared__ T t[32][33];
    const int64_t i0 = (int64_t)blockIdx.y * 32, j0 = (int64_t)blockIdx.x * 32;
    for (int r = threadIdx.y; r < 32; r += 8) {
        int64_t i = i0 + r, j = j0 + threadIdx.x;
        if (i < rows && j < cols) t[r][threadIdx.x] = src[i * lds + j];
    }
    __syncthreads();
    for (int r = threadIdx.y; r < 32; r += 8) {
        int64_t j = j0 + r, i = i0 + threadIdx.x;
        if (i < rows && j < cols) dst[j * ldd + i] = t[threadIdx.x][r];
    }
}

template <typename S, typename D>
__global__ void cast_kernel(int64_t rows, int64_t cols, const S* __restrict__ src, int64_t lds, D* __restrict__ dst, int64_t ldd) {
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= rows * cols) return;
    int64_t i = idx / cols, j = idx % cols;
    dst[i * ldd + j] = (D)src[i * lds + j];
}

// ---- Panels taller than the panel kernel holds (more than 32 x 512 rows) ----
// The panel stays in memory and is factored one column at a time in three launches: a pivot search over the column in
// blocks (each block's best key and first position), the pick (the same rule over the blocks' bests: larger magnitude,
// then smaller position, NaN never chosen over a number) with the swap of the two rows within the panel's columns, and
// the multipliers with their update of the panel's columns to the right, one row per thread. Every item takes the same
// operations in the same order as in the panel kernel: l = x · (1/pivot) (a quotient for a subnormal pivot, x itself for
// a zero one, whose update then skips the rest of its leaf), one fused multiply-add per later column.
constexpr int TALL_T = 256, TALL_G = 256;
static_assert(TALL_T == TALL_G, "the pick reduces one block best per thread");

template <typename T>
__global__ __launch_bounds__(TALL_T) void tall_colmax(const T* __restrict__ a, int64_t lda, int64_t c0, int64_t rows, int c,
                                                      unsigned long long* __restrict__ part_kv, int* __restrict__ part_kp) {
    __shared__ unsigned long long skv[TALL_T];
    __shared__ int skp[TALL_T];
    unsigned long long bkv = 0ull;
    int bkp = 0x7fffffff;
    for (int64_t r = c + (int64_t)blockIdx.x * TALL_T + threadIdx.x; r < rows; r += (int64_t)TALL_G * TALL_T) {
        const unsigned long long kv = mag_key(a[(c0 + r) * lda + c0 + c]);
        if (kv > bkv || (kv == bkv && (int)r < bkp)) { bkv = kv; bkp = (int)r; }
    }
    skv[threadIdx.x] = bkv;
    skp[threadIdx.x] = bkp;
    __syncthreads();
    for (int h = TALL_T / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            const unsigned long long ov = skv[threadIdx.x + h];
            const int op = skp[threadIdx.x + h];
            if (ov > skv[threadIdx.x] || (ov == skv[threadIdx.x] && op < skp[threadIdx.x])) {
                skv[threadIdx.x] = ov;
                skp[threadIdx.x] = op;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { part_kv[blockIdx.x] = skv[0]; part_kp[blockIdx.x] = skp[0]; }
}

template <typename T>
__global__ __launch_bounds__(TALL_T) void tall_pick_swap(T* __restrict__ a, int64_t lda, int64_t c0, int c, int w,
                                                         const unsigned long long* __restrict__ part_kv,
                                                         const int* __restrict__ part_kp, int64_t* __restrict__ pivots,
                                                         int* __restrict__ pick) {
    __shared__ unsigned long long skv[TALL_T];
    __shared__ int skp[TALL_T];
    skv[threadIdx.x] = part_kv[threadIdx.x];
    skp[threadIdx.x] = part_kp[threadIdx.x];
    __syncthreads();
    for (int h = TALL_T / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            const unsigned long long ov = skv[threadIdx.x + h];
            const int op = skp[threadIdx.x + h];
            if (ov > skv[threadIdx.x] || (ov == skv[threadIdx.x] && op < skp[threadIdx.x])) {
                skv[threadIdx.x] = ov;
                skp[threadIdx.x] = op;
            }
        }
        __syncthreads();
    }
    const int p = skp[0];
    if (threadIdx.x == 0) { pivots[c0 + c] = c0 + p; *pick = p; }
    if (p != c && (int)threadIdx.x < w) {
        T* rc = a + (c0 + c) * lda + c0;
        T* rp = a + (c0 + (int64_t)p) * lda + c0;
        const T t = rc[threadIdx.x];
        rc[threadIdx.x] = rp[threadIdx.x];
        rp[threadIdx.x] = t;
    }
}

template <typename T>
__global__ __launch_bounds__(TALL_T) void tall_update(T* __restrict__ a, int64_t lda, int64_t c0, int64_t rows, int c, int w,
                                                      const int* __restrict__ leaf_end) {
    const int64_t r = c + 1 + (int64_t)blockIdx.x * TALL_T + threadIdx.x;
    if (r >= rows) return;
    const T* prow = a + (c0 + c) * lda + c0;
    T* row = a + (c0 + r) * lda + c0;
    const T pivot = prow[c];
    const T tiny = sizeof(T) == 8 ? T(2.2250738585072014e-308) : T(1.17549435e-38f);
    const bool zero = pivot == T(0);
    const int skip = zero ? leaf_end[c0 + c] - (int)c0 - c : 0;
    const T x = row[c];
    const T l = zero ? x : (fabs(pivot) >= tiny ? x * (T(1) / pivot) : x / pivot);
    row[c] = l;
    for (int j = c + 1; j < w; ++j)
        if (j - c >= skip) row[j] = __builtin_elementwise_fma(l, -prow[j], row[j]);
}

// The panel's net row permutation as the move list that swap_kernel applies (count, then pairs dst, src): the swaps of
// columns [c0, c0 + w) in order, followed on the at most 2 w positions they touch.
__global__ void tall_moves(const int64_t* __restrict__ pivots, int64_t c0, int w, long long* __restrict__ moves) {
    if (threadIdx.x != 0) return;
    long long pos[2 * PW], orig[2 * PW];
    int cnt = 0;
    auto find = [&](long long x) {
        for (int i = 0; i < cnt; ++i) if (pos[i] == x) return i;
        pos[cnt] = x; orig[cnt] = x;
        return cnt++;
    };
    for (int c = 0; c < w; ++c) {
        const int i = find(c0 + c), j = find(pivots[c0 + c]);
        const long long t = orig[i]; orig[i] = orig[j]; orig[j] = t;
    }
    long long k = 0;
    for (int i = 0; i < cnt; ++i)
        if (orig[i] != pos[i]) { moves[1 + 2 * k] = pos[i]; moves[2 + 2 * k] = orig[i]; ++k; }
    moves[0] = k;
}

// perm (position -> original row, rustymath's `perm`, factor.rs:132-134) from the panels' move lists, in order: the
// row at position dst after a panel is the row that was at src before it. One workgroup; perm lives in LDS up to
// PERM_LDS rows, above that in memory (agent-scope accesses, so that no wave reads a line its L1 holds from before
// another wave's write).
constexpr int PERM_LDS = 16384;
template <bool IN_LDS>
__global__ __launch_bounds__(256) void moves_to_perm(int64_t n, const long long* __restrict__ moves, int64_t stride,
                                                     int lists, int64_t* __restrict__ perm) {
    __shared__ int sp[IN_LDS ? PERM_LDS : 1];
    const int tid = threadIdx.x;
    auto get = [&](int64_t i) -> int64_t {
        if constexpr (IN_LDS) return sp[i];
        else return __hip_atomic_load(&perm[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    };
    auto put = [&](int64_t i, int64_t v) {
        if constexpr (IN_LDS) sp[i] = (int)v;
        else __hip_atomic_store(&perm[i], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    };
    for (int64_t i = tid; i < n; i += 256) put(i, i);
    __syncthreads();
    for (int L = 0; L < lists; ++L) {
        const long long* mv = moves + L * stride;
        const int cnt = (int)mv[0];
        long long dst = 0;
        int64_t v = 0;
        for (int k0 = 0; k0 < cnt; k0 += 256) {  // at most 2 PW moves: one pass
            const int k = k0 + tid;
            if (k < cnt) { dst = mv[1 + 2 * k]; v = get(mv[2 + 2 * k]); }
            __syncthreads();
            if (k < cnt) put(dst, v);
            __syncthreads();
        }
    }
    if constexpr (IN_LDS)
        for (int64_t i = tid; i < n; i += 256) perm[i] = sp[i];
}

// The reference's leaves: lu_panel halves [c0, c1) at (c0 + c1) / 2 until 16 columns or fewer remain.
void leaves(int64_t c0, int64_t c1, std::vector<int64_t>& ends) {
    if (c1 - c0 > 16) { int64_t mid = (c0 + c1) / 2; leaves(c0, mid, ends); leaves(mid, c1, ends); return; }
    ends.push_back(c1);
}

// Per-device profile counters of the panel kernel.
struct PanelProfile {
    long long* prof = nullptr;
    bool on = false;
};
std::mutex g_panel_mu;
PanelProfile g_panel[16];

int panel_prof(PanelProfile*& st) {
    int dev = 0;
    HS_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_panel_mu);
    st = &g_panel[dev & 15];
    if (!st->prof) {
        HS_HIP(hipMalloc(&st->prof, 8 * sizeof(long long)));
        HS_HIP(hipMemset(st->prof, 0, 8 * sizeof(long long)));
        HS_HIP(hipDeviceSynchronize());
    }
    return OK;
}

// The live profile of trail_rest: while on, each launch is bracketed by HIP events on its stream and writes its first
// tile to a slot of `first`; the launch's flops and bytes are counted from that tile on when the profile is read.
struct TrailProfile {
    std::mutex mu;
    bool on = false;
    int* first = nullptr;  // device, SLOTS ints
    static constexpr int SLOTS = 4096;
    struct Rec {
        hipEvent_t a, b;
        int64_t m, n, k;
        int tiles_m, tiles_n;
        size_t item;
    };
    std::vector<Rec> recs;
    // The panel launches, timed the same way: their own flops (the panel's factorization) are counted per launch; the
    // tiles their workers computed are each trail_rest launch's tiles before its first.
    struct PanelRec {
        hipEvent_t a, b;
        int64_t rows;
        int w;
    };
    std::vector<PanelRec> panels;
    // Events are kept for reuse once read, so that a profiled factorization creates none after the first.
    std::vector<hipEvent_t> spare;
    hipError_t event(hipEvent_t* e) {
        if (!spare.empty()) { *e = spare.back(); spare.pop_back(); return hipSuccess; }
        return hipEventCreate(e);
    }
} g_trail;

// The flops of a panel's own factorization: per column c, the multipliers of the rows below it and their update of the
// panel's columns right of c.
double panel_flops(int64_t rows, int w) {
    double f = 0;
    for (int c = 0; c < w; ++c) {
        const double below = (double)(rows - c - 1);
        f += below > 0 ? below * (1.0 + 2.0 * (double)(w - c - 1)) : 0.0;
    }
    return f;
}

// Tile number -> (tm, tn) as tile::grouped numbers tiles, on the host.
void grouped_host(int id, int tiles_m, int tiles_n, int& tm, int& tn) {
    constexpr int G = 8;
    const int group = id / (G * tiles_n), first = group * G;
    const int gm = tiles_m - first < G ? tiles_m - first : G;
    const int in = id - group * G * tiles_n;
    tm = first + in % gm;
    tn = in / gm;
}

__global__ void fill_words(unsigned long long* p, int64_t n, unsigned long long v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace

// The panel kernel's per-phase profile (ticks of 10 ns, summed over launches): on = 1 clears and enables,
// on = 0 disables; out (8 values) receives the sums where not null.
int panel_profile(int on, long long* out) {
    PanelProfile* st;
    int e = panel_prof(st);
    if (e) return e;
    if (out) HS_HIP(hipMemcpy(out, st->prof, 8 * sizeof(long long), hipMemcpyDeviceToHost));
    if (on) HS_HIP(hipMemset(st->prof, 0, 8 * sizeof(long long)));
    st->on = on != 0;
    return OK;
}

// The LU's live profile (see TrailProfile): on = 1 clears and starts, 0 stops. out[0..3]: trail_rest's summed event
// time (ms), launches, the flops of the tiles it computed (2 rows cols k per tile) and their algorithmic bytes (each
// computed tile's items of c read and written, the rows of a and columns of b the launch's tiles touch, read once);
// out[4..7]: the panel launches' summed event time (ms), launches, their own factorization flops, and the flops of the
// trailing tiles their workers computed.
int lu_profile(int on, double* out) {
    std::lock_guard<std::mutex> lk(g_trail.mu);
    double total = 0, fl = 0, by = 0, worker = 0;
    std::vector<int> first(g_trail.recs.size());
    if (!g_trail.recs.empty())
        HS_HIP(hipMemcpy(first.data(), g_trail.first, first.size() * sizeof(int), hipMemcpyDeviceToHost));
    for (size_t r = 0; r < g_trail.recs.size(); ++r) {
        auto& R = g_trail.recs[r];
        float e = 0;
        HS_HIP(hipEventSynchronize(R.b));
        HS_HIP(hipEventElapsedTime(&e, R.a, R.b));
        g_trail.spare.push_back(R.a);
        g_trail.spare.push_back(R.b);
        total += e;
        std::vector<char> row_seen(R.tiles_m, 0), col_seen(R.tiles_n, 0);
        double c_items = 0, a_rows = 0, b_cols = 0;
        for (int id = 0; id < R.tiles_m * R.tiles_n; ++id) {
            int tm, tn;
            grouped_host(id, R.tiles_m, R.tiles_n, tm, tn);
            const double rows = (double)std::min<int64_t>(tile::BM, R.m - (int64_t)tm * tile::BM);
            const double cols = (double)std::min<int64_t>(tile::BN, R.n - (int64_t)tn * tile::BN);
            const double f = 2.0 * rows * cols * (double)R.k;
            if (id < first[r]) { worker += f; continue; }
            fl += f;
            c_items += rows * cols;
            if (!row_seen[tm]) { row_seen[tm] = 1; a_rows += rows; }
            if (!col_seen[tn]) { col_seen[tn] = 1; b_cols += cols; }
        }
        by += (double)R.item * (2.0 * c_items + (a_rows + b_cols) * (double)R.k);
    }
    double pms = 0, pfl = 0;
    for (auto& P : g_trail.panels) {
        float e = 0;
        HS_HIP(hipEventSynchronize(P.b));
        HS_HIP(hipEventElapsedTime(&e, P.a, P.b));
        g_trail.spare.push_back(P.a);
        g_trail.spare.push_back(P.b);
        pms += e;
        pfl += panel_flops(P.rows, P.w);
    }
    if (out) {
        out[0] = total; out[1] = (double)first.size(); out[2] = fl; out[3] = by;
        out[4] = pms; out[5] = (double)g_trail.panels.size(); out[6] = pfl; out[7] = worker;
    }
    g_trail.recs.clear();
    g_trail.panels.clear();
    g_trail.on = on != 0;
    if (g_trail.on && !g_trail.first) HS_HIP(hipMalloc(&g_trail.first, TrailProfile::SLOTS * sizeof(int)));
    return OK;
}

// Right-looking LU over blocks of up to OB = 256 columns, each made of panels of at most 64 (whole reference leaves).
// Within a block, each panel is factored, its row moves applied to every other column, and only the block's own
// columns right of it updated (a narrow product, k = the panel's width). After the block, the rows of U right of it
// are solved panel by panel (a product over the block's earlier panels, then the panel's unit-lower triangle), and
// the trailing matrix takes one product with k up to 256, which reads and writes it a quarter as often as per-panel
// updates would. Every item still takes its terms in order of the pivot column, one fused multiply-add each.
template <typename T>
int getrf(int64_t n, T* a, int64_t lda, int64_t* pivots, hipStream_t s, int bf16_terms, int64_t* perm_out, bool keep_rows) {
    if (n == 0) return OK;
    if (n >= (int64_t)1 << 31) { set_error("LU of order 2^31 or more is not supported"); return ERR_UNSUPPORTED; }
    // Above 32 x 512 rows the panels are factored in memory, column by column (tall_*), without look-ahead.
    const bool tall = n > (int64_t)MAX_WG * PT;
    // The panel launches' workgroups wait on one another: one factorization with such panels at a time (see
    // resident_mutex). Every path below synchronises the stream before it returns.
    std::unique_lock<std::mutex> resident(resident_mutex(), std::defer_lock);
    if (n > PT) resident.lock();
    // The panels' exchange: through one L2 where the placement probe found launch indices 0, 8, 16, ... on one XCD
    // (local = 1), else at agent scope.
    const int local = n > PT && exchange_local(true) ? 1 : 0;
    PanelProfile* st;
    int e = panel_prof(st);
    if (e) return e;
    constexpr int64_t OB = 4 * PW;
    // Panels: consecutive whole reference leaves, at most PW columns each; blocks: consecutive panels, at most OB.
    std::vector<int64_t> ends;
    leaves(0, n, ends);
    std::vector<int> leaf_end(n);
    {
        int64_t c = 0;
        for (int64_t end : ends) { for (; c < end; ++c) leaf_end[c] = (int)end; }
    }
    std::vector<std::pair<int64_t, int64_t>> panels;
    {
        int64_t start = 0, cur = 0;
        for (int64_t end : ends) {
            if (end - start > PW) { panels.push_back({start, cur}); start = cur; }
            cur = end;
        }
        panels.push_back({start, cur});
    }
    std::vector<std::pair<size_t, size_t>> blocks;  // [first panel, end panel)
    for (size_t i = 0; i < panels.size();) {
        size_t j = i + 1;
        while (j < panels.size() && panels[j].second - panels[i].first <= OB) ++j;
        blocks.push_back({i, j});
        i = j;
    }
    const size_t np = panels.size();
    const size_t move_words = 1 + 4 * PW;
    // + status and stop mark, then one tile counter per block (for the product run beside that block's panels)
    const size_t clear_words = sizeof(Exchange) / 8 + np * move_words + 2 + (blocks.size() + 1) / 2 + (TALL_G * 12 + 8) / 8 +
                               np * ARRIVE_STRIPS / 2;
    int* d_leaf = (int*)pool_get(n * sizeof(int));
    T* scratch = (T*)pool_get(sizeof(T) * (size_t)n * PW);
    char* xbuf = (char*)pool_get(clear_words * 8);
    if (!d_leaf || !scratch || !xbuf) {
        pool_put(d_leaf); pool_put(scratch); pool_put(xbuf);
        return hip_status(hipErrorOutOfMemory, "pool_get");
    }
    Exchange* ex = (Exchange*)xbuf;
    long long* moves = (long long*)(xbuf + sizeof(Exchange));
    int* status = (int*)(moves + np * move_words);
    unsigned* trail_stop = (unsigned*)(status + 2);
    int* trail_next = status + 4;  // blocks.size() counters
    // The tall panels' block bests and pick, after the counters.
    unsigned long long* tall_kv = (unsigned long long*)xbuf + (clear_words - (TALL_G * 12 + 8) / 8);
    int* tall_kp = (int*)(tall_kv + TALL_G);
    int* tall_pick = tall_kp + TALL_G;
    int* arrived = (int*)(tall_kv - np * ARRIVE_STRIPS / 2);  // panel_update_kernel's counts, ARRIVE_STRIPS a panel
    HS_HIP(hipMemcpyAsync(d_leaf, leaf_end.data(), n * sizeof(int), hipMemcpyHostToDevice, s));
    fill_words<<<(unsigned)((clear_words + 255) / 256), 256, 0, s>>>((unsigned long long*)xbuf, (int64_t)clear_words, 0ull);
    HS_HIP(hipGetLastError());

    int rc = OK;
    // The row moves of `count` consecutive panels, in order, on the columns [lo, hi) outside [x0, x1).
    auto swap = [&](const long long* mv, int64_t lo, int64_t hi, int64_t x0, int64_t x1, int count = 1) {
        const int64_t cols = (hi - lo) - (x1 - x0);
        if (cols <= 0) return OK;
        swap_kernel<T><<<(unsigned)((cols + 63) / 64), 256, 0, s>>>(a, lda, lo, hi, x0, x1, mv, (int64_t)move_words,
                                                                     count);
        HS_HIP(hipGetLastError());
        return OK;
    };
    // Look-ahead. After block b, the trailing product runs in two parts: the next block's columns at once (on the whole
    // chip), and the columns right of the next block (`job`) by the workers of the next block's panel launches, which
    // fill the other seven XCDs while the panel's workgroups run on one; trail_rest takes the tiles left when the
    // block's panels are done. Meanwhile a panel's row moves reach only its own block's columns; the other columns take them after the
    // product (row moves and the product touch every item in the same order either way). Every tile keeps its terms in
    // order of p, so every item takes the same operations as with one product.
    TrailJob<T> job{};
    const TrailJob<T> none{};
    for (size_t b = 0; b < blocks.size() && rc == OK; ++b) {
        const auto [p0, p1] = blocks[b];
        const int64_t B0 = panels[p0].first, B1 = panels[p1 - 1].second;
        const bool ahead = job.C != nullptr;
        for (size_t i = p0; i < p1 && rc == OK; ++i) {
            const auto [c0, c1] = panels[i];
            const int w = (int)(c1 - c0);
            const int64_t rows = n - c0;
            const int nwg = (int)((rows + PT - 1) / PT);
            const unsigned grid = ahead ? 8u * MAX_WG : (nwg > 1 ? (unsigned)(8 * nwg) : 1u);
            long long* mv = moves + i * move_words;
            if (tall) {
                for (int c = 0; c < w; ++c) {
                    tall_colmax<T><<<TALL_G, TALL_T, 0, s>>>(a, lda, c0, rows, c, tall_kv, tall_kp);
                    tall_pick_swap<T><<<1, TALL_T, 0, s>>>(a, lda, c0, c, w, tall_kv, tall_kp, pivots, tall_pick);
                    tall_update<T><<<(unsigned)((rows - c - 1 + TALL_T - 1) / TALL_T + 1), TALL_T, 0, s>>>(a, lda, c0, rows,
                                                                                                      c, w, d_leaf);
                }
                tall_moves<<<1, 64, 0, s>>>(pivots, c0, w, mv);
            } else {
                std::unique_lock<std::mutex> lk(g_trail.mu);
                if (g_trail.on && g_trail.panels.size() < (size_t)TrailProfile::SLOTS) {
                    TrailProfile::PanelRec P{nullptr, nullptr, rows, w};
                    HS_HIP(g_trail.event(&P.a));
                    HS_HIP(g_trail.event(&P.b));
                    hipExtLaunchKernelGGL(panel_kernel<T>, dim3(grid), dim3(panel_threads<T>()), 0, s, P.a, P.b, 0, a, lda,
                                          c0, rows, w,
                                          pivots, mv, d_leaf, scratch, ex, (unsigned)(i * PW + 1), nwg, status,
                                          st->on ? st->prof : nullptr, ahead ? job : none, local);
                    g_trail.panels.push_back(P);
                } else {
                    panel_kernel<T><<<grid, panel_threads<T>(), 0, s>>>(a, lda, c0, rows, w, pivots, mv, d_leaf, scratch, ex,
                                                         (unsigned)(i * PW + 1), nwg, status, st->on ? st->prof : nullptr,
                                                         ahead ? job : none, local);
                }
            }
            HS_HIP(hipGetLastError());
            if ((rc = ahead ? swap(mv, B0, B1, c0, c1) : swap(mv, 0, n, c0, c1))) break;
            if (c1 < B1) {  // the block's columns right of the panel: its rows of U, then the rows below
                constexpr int TCOL = RowsCols<T>::v;
                const dim3 ug((unsigned)((n - c1 + 63) / 64), (unsigned)((B1 - c1 + TCOL - 1) / TCOL));
                panel_update_kernel<T><<<ug, 256, 0, s>>>(a, lda, c0, w, B1, n, arrived + i * ARRIVE_STRIPS);
                HS_HIP(hipGetLastError());
            }
        }
        if (rc) break;
        if (ahead) {
            const unsigned grid = (unsigned)(((job.tiles_m * job.tiles_n + 7) / 8) * 8);
            std::unique_lock<std::mutex> lk(g_trail.mu);
            if (g_trail.on && g_trail.recs.size() < (size_t)TrailProfile::SLOTS) {
                TrailProfile::Rec R{nullptr, nullptr, job.m, job.n, job.k, job.tiles_m, job.tiles_n, sizeof(T)};
                HS_HIP(g_trail.event(&R.a));
                HS_HIP(g_trail.event(&R.b));
                job.log = g_trail.first + g_trail.recs.size();
                hipExtLaunchKernelGGL(trail_rest<T>, dim3(grid), dim3(256), 0, s, R.a, R.b, 0, job);
                g_trail.recs.push_back(R);
            } else {
                trail_rest<T><<<grid, 256, 0, s>>>(job);
            }
            lk.unlock();
            HS_HIP(hipGetLastError());
            job = none;
            if ((rc = swap(moves + p0 * move_words, 0, n, B0, B1, (int)(p1 - p0)))) break;
        }
        if (B1 >= n) continue;
        // Rows of U right of the block (one launch for all its panels), then the trailing product.
        {
            trsm_rows_kernel<T><<<(unsigned)((n - B1 + ROWS_TCOL - 1) / ROWS_TCOL), 256, 0, s>>>(a, lda, B0, (int)(B1 - B0), B1,
                                                                                          n);
            HS_HIP(hipGetLastError());
        }
        if constexpr (std::is_same<T, float>::value) {
            if (bf16_terms) {
                rc = gemm_bf16_update(n - B1, n - B1, B1 - B0, a + B1 * lda + B0, lda, a + B0 * lda + B1, lda,
                                      a + B1 * lda + B1, lda, s, bf16_terms);
                continue;
            }
        }
        const int64_t B2 = panels[blocks[b + 1].second - 1].second;
        if (tall || B2 >= n || n - B1 < 2) {
            rc = gemm<T>(false, false, n - B1, n - B1, B1 - B0, T(-1), a + B1 * lda + B0, lda, a + B0 * lda + B1, lda, T(1),
                         a + B1 * lda + B1, lda, s);
            continue;
        }
        // The next block's columns now; the rest beside the next block's panels.
        rc = gemm<T>(false, false, n - B1, B2 - B1, B1 - B0, T(-1), a + B1 * lda + B0, lda, a + B0 * lda + B1, lda, T(1),
                     a + B1 * lda + B1, lda, s, true);
        job.m = n - B1;
        job.n = n - B2;
        job.k = B1 - B0;
        job.A = a + B1 * lda + B0;
        job.lda = lda;
        job.B = a + B0 * lda + B2;
        job.ldb = lda;
        job.C = a + B1 * lda + B2;
        job.ldc = lda;
        job.tiles_m = (int)((job.m + tile::BM - 1) / tile::BM);
        job.tiles_n = (int)((job.n + tile::BN - 1) / tile::BN);
        job.next = trail_next + b + 1;  // the next block's panels take its tiles
        job.stop = trail_stop;
    }
    if (rc == OK) {
        if (!keep_rows) {
            dim3 grid((unsigned)((n + 31) / 32), (unsigned)((n + 31) / 32)), block(32, 8);
            transpose_square_inplace<T><<<grid, block, 0, s>>>(a, lda, n);
            HS_HIP(hipGetLastError());
        }
        if (perm_out) {
            if (n <= PERM_LDS) moves_to_perm<true><<<1, 256, 0, s>>>(n, moves, (int64_t)move_words, (int)np, perm_out);
            else moves_to_perm<false><<<1, 256, 0, s>>>(n, moves, (int64_t)move_words, (int)np, perm_out);
            HS_HIP(hipGetLastError());
        }
        int stv = 0;
        HS_HIP(hipMemcpyAsync(&stv, status, sizeof(int), hipMemcpyDeviceToHost, s));
        HS_HIP(hipStreamSynchronize(s));
        if (stv) {
            set_error("panel exchange wait ran out (workgroups not co-resident?)");
            rc = ERR_LIMIT;
        }
    } else {
        HS_HIP(hipStreamSynchronize(s));
    }
    pool_put(d_leaf);
    pool_put(scratch);
    pool_put(xbuf);
    return rc;
}

template int getrf<double>(int64_t, double*, int64_t, int64_t*, hipStream_t, int, int64_t*, bool);
template int getrf<float>(int64_t, float*, int64_t, int64_t*, hipStream_t, int, int64_t*, bool);

template <typename T>
int transpose(int64_t rows, int64_t cols, const T* src, int64_t lds, T* dst, int64_t ldd, hipStream_t s) {
    if (rows == 0 || cols == 0) return OK;
    dim3 grid((unsigned)((cols + 31) / 32), (unsigned)((rows + 31) / 32)), block(32, 8);
    transpose_kernel<T><<<grid, block, 0, s>>>(rows, cols, src, lds, dst, ldd);
    HS_HIP(hipGetLastError());
    return OK;
}
template int transpose<double>(int64_t, int64_t, const double*, int64_t, double*, int64_t, hipStream_t);
template int transpose<float>(int64_t, int64_t, const float*, int64_t, float*, int64_t, hipStream_t);

// dst = float(src) for a compact n-item array, two items per thread and step, and part[b] = block b's sum of the
// squares of src (FP64, a fixed grid and order, so the sum is the same on every run).
__global__ __launch_bounds__(256) void cast_sumsq_kernel(int64_t count, const double* __restrict__ src,
                                                         float* __restrict__ dst, double* __restrict__ part) {
    __shared__ double sp[4];
    typedef double d2 __attribute__((ext_vector_type(2)));
    typedef float f2 __attribute__((ext_vector_type(2)));
    double acc = 0;
    const int64_t pairs = count / 2, stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < pairs; i += stride) {
        const d2 v = reinterpret_cast<const d2*>(src)[i];
        f2 o;
        o.x = (float)v.x;
        o.y = (float)v.y;
        reinterpret_cast<f2*>(dst)[i] = o;
        acc = __builtin_elementwise_fma(v.x, v.x, acc);
        acc = __builtin_elementwise_fma(v.y, v.y, acc);
    }
    if ((count & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
        const double v = src[count - 1];
        dst[count - 1] = (float)v;
        acc = __builtin_elementwise_fma(v, v, acc);
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if ((threadIdx.x & 63) == 0) sp[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (sp[0] + sp[1]) + (sp[2] + sp[3]);
}

int cast_d2s_sumsq(int64_t count, const double* src, float* dst, double* part, int blocks, hipStream_t s) {
    if (!count) return OK;
    if ((reinterpret_cast<uintptr_t>(src) & 15) || (reinterpret_cast<uintptr_t>(dst) & 7)) {
        set_error("cast_d2s_sumsq needs 16-byte aligned source and 8-byte aligned destination");
        return ERR_UNSUPPORTED;
    }
    cast_sumsq_kernel<<<blocks, 256, 0, s>>>(count, src, dst, part);
    HS_HIP(hipGetLastError());
    return OK;
}

int cast_d2s(int64_t rows, int64_t cols, const double* src, int64_t lds, float* dst, int64_t ldd, hipStream_t s) {
    int64_t total = rows * cols;
    if (!total) return OK;
    cast_kernel<double, float><<<(unsigned)((total + 255) / 256), 256, 0, s>>>(rows, cols, src, lds, dst, ldd);
    HS_HIP(hipGetLastError());
    return OK;
}
int cast_s2d(int64_t rows, int64_t cols, const float* src, int64_t lds, double* dst, int64_t ldd, hipStream_t s) {
    int64_t total = rows * cols;
    if (!total) return OK;
    cast_kernel<float, double><<<(unsigned)((total + 255) / 256), 256, 0, s>>>(rows, cols, src, lds, dst, ldd);
    HS_HIP(hipGetLastError());
    return OK;
}

}  // namespace hs
