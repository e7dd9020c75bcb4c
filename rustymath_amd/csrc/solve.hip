// Triangular solves with LU factors on gfx950: lu_solve (core/src/linalg/factor.rs:633-641) and the
// transposed solve that the adjoint needs.
//
// rustymath substitutes in three orders, and each fixes the last bits of x:
//   L y = P b           every row takes its terms in increasing column order (`forward` :668, `single` :725),
//   U x = y, one column  terms in decreasing column order, then a quotient by the diagonal (`single`),
//   U x = y, k columns   `backward` (:693): halving above NS = 32 rows, so a row takes the bottom half of
//                        each ancestor in increasing order, outermost ancestor first, then the rows of its
//                        own leaf after it in increasing order, then a reciprocal multiply (`divide` :506).
// One kernel serves all of them. A workgroup owns a block of rows (a reference leaf where the order needs
// it) and all k columns; blocks are numbered in solve order, so a block only ever waits for blocks with
// smaller numbers, which were dispatched before it. Each row keeps one running sum per column and adds
// the solved rows of other blocks in the order above, one fused multiply-add each, as soon as their
// block publishes them; then the block substitutes within itself and publishes its rows.
#include "hs_kernels.h"
#include <atomic>
#include <vector>

namespace hs {

namespace {

constexpr int BR = 32;         // rows per block (NS of factor.rs:82; leaves have at most this many)
constexpr int MAXK = 16;       // right-hand sides per launch
constexpr uint64_t SPIN_LIMIT = 1ull << 22;


enum Mode : int {
    FWD_UNIT = 0,    // L (unit lower of F): increasing order, no division
    BWD_SINGLE = 1,  // U, one column: decreasing order, then a quotient
    BWD_LEAVES = 2,  // U, several columns: the halving order, then a reciprocal multiply
    FWD_DIAG = 3,    // Uᵀ (lower, diagonal of U): increasing order, then a reciprocal multiply
    BWD_UNIT = 4,    // Lᵀ (upper, unit): decreasing order, no division
};

template <typename T> struct Bits;
template <> struct Bits<double> { typedef unsigned long long u; };
template <> struct Bits<float> { typedef unsigned int u; };
// Device-coherent (sc1) loads and stores of the rows that blocks publish to each other.
template <typename T>
__device__ __forceinline__ void put(T* p, T v) {
    typedef typename Bits<T>::u U;
    __hip_atomic_store(reinterpret_cast<U*>(p), __builtin_bit_cast(U, v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ T get(const T* p) {
    typedef typename Bits<T>::u U;
    U u = __hip_atomic_load(reinterpret_cast<U*>(const_cast<T*>(p)), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return __builtin_bit_cast(T, u);
}

// Segments: seg[seg_off[b] .. seg_off[b+1]) are (first row, end row, direction) of the other blocks'
// rows that block b's rows add, in order; direction 1 = increasing rows, 0 = decreasing.
// The blocks are launch indices 0, 8, 16, ... (one XCD under round-robin dispatch; the others return at
// once): a block publishes its solved rows to x and raises its flag with device-coherent (sc1) stores, and
// the blocks after it read both with sc1 loads. Each producer block's rows of x and the 32x32 tile of
// matrix items that the consumer needs come in through LDS with coalesced loads.
template <typename T>
__global__ __launch_bounds__(BR * MAXK) void trsv_blocks(int mode, int64_t n, int k, const T* __restrict__ f, int64_t ld,
                                                         T* x, int64_t ldx, const int* __restrict__ blk_r0,
                                                         const int* __restrict__ blk_r1, const int* __restrict__ blk_of_row,
                                                         const int* __restrict__ seg_off, const int* __restrict__ seg,
                                                         unsigned long long* flags, int* status) {
    if (blockIdx.x & 7) return;
    const int b = blockIdx.x >> 3;
    const int r0 = blk_r0[b], r1 = blk_r1[b], rows = r1 - r0;
    const int tid = threadIdx.x, li = tid % BR, col = tid / BR;
    const bool live = li < rows && col < k;
    const int64_t i = r0 + li;
    const bool tmode = mode == FWD_DIAG || mode == BWD_UNIT;  // entry(i, p) = f[i*ld + p]
    __shared__ T xs[BR][MAXK];
    __shared__ T es[BR][BR + 1];  // es[row of this block][producer row]
    __shared__ int s_abort;
    if (tid == 0) s_abort = 0;

    T acc = live ? x[i * ldx + col] : T(0);
    __syncthreads();

    for (int s = seg_off[b]; s < seg_off[b + 1]; ++s) {
        const int a0 = seg[3 * s], a1 = seg[3 * s + 1], inc = seg[3 * s + 2];
        int p = inc ? a0 : a1 - 1;
        while (inc ? p < a1 : p >= a0) {
            const int pb = blk_of_row[p];
            const int q0 = blk_r0[pb], q1 = blk_r1[pb];
            const int lo = q0 > a0 ? q0 : a0, hi = q1 < a1 ? q1 : a1, cnt = hi - lo;
            if (tid == 0) {
                uint64_t spins = 0;
                while (__hip_atomic_load(&flags[pb], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0ull) {
                    if (++spins > SPIN_LIMIT) { s_abort = 1; atomicExch(status, ERR_LIMIT); break; }
                }
            }
            __syncthreads();
            if (s_abort) return;
            // Stage the producer's solved rows and this block's matrix items against them.
            for (int idx = tid; idx < cnt * k; idx += blockDim.x) {
                const int rr = idx / k, cc = idx % k;
                xs[rr][cc] = get(&x[(int64_t)(lo + rr) * ldx + cc]);
            }
            for (int idx = tid; idx < rows * cnt; idx += blockDim.x) {
                int rr, qq;
                if (tmode) { rr = idx / cnt; qq = idx % cnt; }   // row-contiguous items
                else { qq = idx / rows; rr = idx % rows; }       // column-contiguous items
                es[rr][qq] = tmode ? f[(int64_t)(r0 + rr) * ld + lo + qq] : f[(int64_t)(lo + qq) * ld + r0 + rr];
            }
            __syncthreads();
            if (live) {
                if (inc)
                    for (int q = 0; q < cnt; ++q) acc = __builtin_fma(xs[q][col], -es[li][q], acc);
                else
                    for (int q = cnt - 1; q >= 0; --q) acc = __builtin_fma(xs[q][col], -es[li][q], acc);
            }
            __syncthreads();
            p = inc ? hi : lo - 1;
        }
    }

    // Substitution within the block, with its own triangle staged in LDS.
    for (int idx = tid; idx < rows * rows; idx += blockDim.x) {
        int rr, qq;
        if (tmode) { rr = idx / rows; qq = idx % rows; }
        else { qq = idx / rows; rr = idx % rows; }
        es[rr][qq] = tmode ? f[(int64_t)(r0 + rr) * ld + r0 + qq] : f[(int64_t)(r0 + qq) * ld + r0 + rr];
    }
    __syncthreads();
    const bool down = mode == FWD_UNIT || mode == FWD_DIAG;
    const T tiny = sizeof(T) == 8 ? T(2.2250738585072014e-308) : T(1.17549435e-38f);
    if (mode == BWD_LEAVES) {
        // Row r1-1 first; each row adds the rows after it in increasing order, then divides.
        for (int t = rows - 1; t >= 0; --t) {
            if (live && li == t) {
                for (int q = t + 1; q < rows; ++q) acc = __builtin_fma(xs[q][col], -es[li][q], acc);
                const T d = es[li][li];
                acc = fabs(d) >= tiny ? acc * (T(1) / d) : acc / d;
                xs[t][col] = acc;
            }
            __syncthreads();
        }
    } else {
        for (int st = 0; st < rows; ++st) {
            const int t = down ? st : rows - 1 - st;
            if (live && li == t) {
                const T d = es[li][li];
                if (mode == BWD_SINGLE) acc = acc / d;
                else if (mode == FWD_DIAG) acc = fabs(d) >= tiny ? acc * (T(1) / d) : acc / d;
                xs[t][col] = acc;
            }
            __syncthreads();
            const bool after = down ? li > t : li < t;
            if (live && after) acc = __builtin_fma(xs[t][col], -es[li][t], acc);
        }
    }
    if (live) put(&x[i * ldx + col], acc);
    __syncthreads();  // every store of this block has completed (s_waitcnt before the barrier)
    if (tid == 0) __hip_atomic_store(&flags[b], 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void fill_u64(unsigned long long* p, int64_t n, unsigned long long v) {
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx < n) p[idx] = v;
}

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

// The bottom-half ranges of `backward`'s halving (factor.rs:693-707) that the rows of the leaf
// [l0, l1) add before their own leaf, outermost first.
void backward_ranges(int r0, int r1, int l0, std::vector<std::pair<int, int>>& out) {
    if (r1 - r0 <= BR) return;
    int mid = (r0 + r1) / 2;
    if (l0 < mid) { out.push_back({mid, r1}); backward_ranges(r0, mid, l0, out); }
    else backward_ranges(mid, r1, l0, out);
}
void backward_leaves(int r0, int r1, std::vector<std::pair<int, int>>& out) {
    if (r1 - r0 > BR) { int mid = (r0 + r1) / 2; backward_leaves(r0, mid, out); backward_leaves(mid, r1, out); return; }
    out.push_back({r0, r1});
}

// One right-hand side (the refinement's corrections, single-column lu_solve): one wave per 32-row block, in
// solve order. Lane t holds row r0 + t. A solved row is published as 8-byte words that carry the call's
// epoch beside the value (two words for FP64), stored and loaded with device-coherent (sc1) atomics, so one
// load both tells that a row is ready and brings its value; the tile of matrix items a block needs from the
// next producer is loaded before the wait. A row adds the producers' rows in the reference's order, in
// increasing (L) or decreasing (U, `single`) column order, and the block then substitutes within itself
// with register broadcasts: a quotient by the diagonal at each row's turn for U.
template <typename T>
__device__ __forceinline__ void publish1(unsigned long long* pub, int64_t row, T v, unsigned long long epoch) {
    if constexpr (sizeof(T) == 4) {
        const unsigned long long w = (epoch << 32) | __builtin_bit_cast(unsigned, v);
        __hip_atomic_store(&pub[row], w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
        __hip_atomic_store(&pub[2 * row], (epoch << 32) | (u & 0xffffffffull), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&pub[2 * row + 1], (epoch << 32) | (u >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}
template <typename T>
__device__ __forceinline__ bool fetch1(unsigned long long* pub, int64_t row, unsigned long long epoch, T& v) {
    if constexpr (sizeof(T) == 4) {
        const unsigned long long w = __hip_atomic_load(&pub[row], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        v = __builtin_bit_cast(T, (unsigned)w);
        return (w >> 32) == epoch;
    } else {
        const unsigned long long w0 = __hip_atomic_load(&pub[2 * row], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long w1 = __hip_atomic_load(&pub[2 * row + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        v = __builtin_bit_cast(T, (w1 << 32) | (w0 & 0xffffffffull));
        return (w0 >> 32) == epoch && (w1 >> 32) == epoch;
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

template <typename T>
__global__ __launch_bounds__(64) void trsv1_kernel(int mode, int64_t n, const T* __restrict__ f, int64_t ld, T* x,
                                                   unsigned long long* pub, unsigned long long epoch, int* status) {
    const int nb = (int)((n + B1 - 1) / B1), b = blockIdx.x, lane = threadIdx.x;
    const bool fwd = mode == FWD_UNIT;
    const int q = fwd ? b : nb - 1 - b;  // this block's rows [q B1, ...)
    const int64_t r0 = (int64_t)q * B1;
    const int h = (int)(n - r0 < B1 ? n - r0 : B1);
    const bool act = lane < h;
    const int64_t i = r0 + (act ? lane : 0);
    T acc = act ? x[i] : T(0);
    // Tile of producer block pq: e[j] = F(i, pq B1 + j) (column-major factors), padded with zeros.
    auto load_tile = [&](int pq, T* e) {
        const int64_t p0 = (int64_t)pq * B1;
#pragma unroll
        for (int j = 0; j < B1; ++j) {
            const int64_t pc = p0 + j < n ? p0 + j : n - 1;
            const T v = f[pc * ld + i];
            e[j] = p0 + j < n ? v : T(0);
        }
    };
    const int np = fwd ? q : nb - 1 - q;
    T e[B1], en[B1], d[B1];
    load_tile(q, d);  // the diagonal tile d[j] = F(i, r0 + j), off the critical path
    if (np > 0) load_tile(fwd ? 0 : nb - 1, e);
    for (int t = 0; t < np; ++t) {
        const int pq = fwd ? t : nb - 1 - t;
        if (t + 1 < np) load_tile(fwd ? t + 1 : nb - 2 - t, en);
        const int64_t p0 = (int64_t)pq * B1;
        const int ph = (int)(n - p0 < B1 ? n - p0 : B1);
        T xv = T(0);
        bool ok = lane >= ph;
        uint64_t spins = 0;
        while (true) {
            if (!ok) ok = fetch1(pub, p0 + lane, epoch, xv);
            if (__all(ok)) break;
            __builtin_amdgcn_s_sleep(2);  // many blocks wait on the same producer: poll its words gently
            if (++spins > SPIN_LIMIT) { if (lane == 0) atomicExch(status, ERR_LIMIT); return; }
        }
        if (fwd) {
#pragma unroll
            for (int j = 0; j < B1; ++j)
                if (j < ph) acc = __builtin_fma(lane_value(xv, j), -e[j], acc);
        } else {
#pragma unroll
            for (int j = B1 - 1; j >= 0; --j)
                if (j < ph) acc = __builtin_fma(lane_value(xv, j), -e[j], acc);
        }
#pragma unroll
        for (int j = 0; j < B1; ++j) e[j] = en[j];
    }
    // Within the block.
    if (fwd) {
#pragma unroll
        for (int j = 0; j < B1; ++j) {
            if (j < h) {
                const T xj = lane_value(acc, j);
                if (lane > j) acc = __builtin_fma(xj, -d[j], acc);
            }
        }
    } else {
        T dg = T(1);
#pragma unroll
        for (int j = 0; j < B1; ++j) if (j == lane) dg = d[j];
#pragma unroll
        for (int j = B1 - 1; j >= 0; --j) {
            if (j < h) {
                if (lane == j) acc = acc / dg;
                const T xj = lane_value(acc, j);
                if (lane < j) acc = __builtin_fma(xj, -d[j], acc);
            }
        }
    }
    if (act) {
        publish1(pub, i, acc, epoch);
        x[i] = acc;
    }
}

// The epoch of each single-column pass: increasing across calls, so that words left by an earlier pass never
// match (the published words are also cleared before each pass).
std::atomic<unsigned long long> g_epoch1{1};

// L then U (lu_solve's order for one column) in place on x (n items, contiguous), with one status check.
template <typename T>
int trsv1_lu(int64_t n, const T* f, int64_t ld, T* x, hipStream_t s) {
    const int nb = (int)((n + B1 - 1) / B1);
    const size_t words = (size_t)n * (sizeof(T) == 8 ? 2 : 1);
    char* buf = (char*)pool_get(words * 8 + 64);
    if (!buf) return hip_status(hipErrorOutOfMemory, "pool_get");
    unsigned long long* pub = (unsigned long long*)buf;
    int* status = (int*)(buf + words * 8);
    HS_HIP(hipMemsetAsync(buf, 0, words * 8 + sizeof(int), s));
    const unsigned long long e1 = (g_epoch1.fetch_add(2) % 0xfffffff0ull) | 1ull;  // odd, and e1 + 1 < 2^32
    trsv1_kernel<T><<<nb, 64, 0, s>>>(FWD_UNIT, n, f, ld, x, pub, e1, status);
    HS_HIP(hipGetLastError());
    trsv1_kernel<T><<<nb, 64, 0, s>>>(BWD_SINGLE, n, f, ld, x, pub, e1 + 1, status);
    HS_HIP(hipGetLastError());
    int st = 0;
    HS_HIP(hipMemcpyAsync(&st, status, sizeof(int), hipMemcpyDeviceToHost, s));
    HS_HIP(hipStreamSynchronize(s));
    pool_put(buf);
    if (st) { set_error("triangular solve wait ran out"); return ERR_LIMIT; }
    return OK;
}

// One triangular pass of `mode` over rows [lo, hi) of x (n×k, k <= MAXK) in place, taking only the terms of rows in
// [lo, hi): the terms of rows outside it have been added already (by products, in the same order). For BWD_LEAVES,
// [lo, hi) is a node of `backward`'s halving of [0, n), so the halving inside it is the reference's.
template <typename T>
int trsv_pass(int mode, int64_t lo, int64_t hi, int64_t n, int k, const T* f, int64_t ld, T* x, int64_t ldx,
              hipStream_t s) {
    std::vector<std::pair<int, int>> blocks;  // in solve order
    const bool down = mode == FWD_UNIT || mode == FWD_DIAG;
    if (mode == BWD_LEAVES) {
        std::vector<std::pair<int, int>> lv;
        backward_leaves((int)lo, (int)hi, lv);
        blocks.assign(lv.rbegin(), lv.rend());
    } else if (down) {
        for (int64_t r = lo; r < hi; r += BR) blocks.push_back({(int)r, (int)std::min<int64_t>(hi, r + BR)});
    } else {
        const int nb = (int)((hi - lo + BR - 1) / BR);
        for (int q = nb - 1; q >= 0; --q)
            blocks.push_back({(int)(lo + (int64_t)q * BR), (int)std::min<int64_t>(hi, lo + (int64_t)(q + 1) * BR)});
    }
    const int nb = (int)blocks.size();
    std::vector<int> r0(nb), r1(nb), of_row(n), seg_off(nb + 1), seg;
    for (int b = 0; b < nb; ++b) {
        r0[b] = blocks[b].first; r1[b] = blocks[b].second;
        for (int r = r0[b]; r < r1[b]; ++r) of_row[r] = b;
        seg_off[b] = (int)seg.size() / 3;
        if (mode == BWD_LEAVES) {
            std::vector<std::pair<int, int>> rs;
            backward_ranges((int)lo, (int)hi, r0[b], rs);
            for (auto [a0, a1] : rs) { seg.push_back(a0); seg.push_back(a1); seg.push_back(1); }
        } else if (down) {
            if (r0[b] > lo) { seg.push_back((int)lo); seg.push_back(r0[b]); seg.push_back(1); }
        } else {
            if (r1[b] < hi) { seg.push_back(r1[b]); seg.push_back((int)hi); seg.push_back(0); }
        }
    }
    seg_off[nb] = (int)seg.size() / 3;
    if (seg.empty()) { seg.push_back(0); seg.push_back(0); seg.push_back(1); }

    size_t ibytes = sizeof(int) * (2 * (size_t)nb + (size_t)n + nb + 1 + seg.size());
    char* buf = (char*)pool_get(ibytes + sizeof(unsigned long long) * nb + sizeof(int) + 64);
    if (!buf) return hip_status(hipErrorOutOfMemory, "pool_get");
    std::vector<int> host;
    host.reserve(ibytes / sizeof(int));
    host.insert(host.end(), r0.begin(), r0.end());
    host.insert(host.end(), r1.begin(), r1.end());
    host.insert(host.end(), of_row.begin(), of_row.end());
    host.insert(host.end(), seg_off.begin(), seg_off.end());
    host.insert(host.end(), seg.begin(), seg.end());
    int* d = (int*)buf;
    HS_HIP(hipMemcpyAsync(d, host.data(), ibytes, hipMemcpyHostToDevice, s));
    size_t foff = (ibytes + 63) / 64 * 64;
    unsigned long long* flags = (unsigned long long*)(buf + foff);
    int* status = (int*)(flags + nb);
    fill_u64<<<(nb + 255) / 256, 256, 0, s>>>(flags, nb, 0ull);
    HS_HIP(hipGetLastError());
    HS_HIP(hipMemsetAsync(status, 0, sizeof(int), s));
    const int* d_r0 = d;
    const int* d_r1 = d + nb;
    const int* d_of = d + 2 * nb;
    const int* d_soff = d_of + n;
    const int* d_seg = d_soff + nb + 1;
    trsv_blocks<T><<<8 * nb, BR * MAXK, 0, s>>>(mode, n, k, f, ld, x, ldx, d_r0, d_r1, d_of, d_soff, d_seg, flags, status);
    HS_HIP(hipGetLastError());
    int st = 0;
    HS_HIP(hipMemcpyAsync(&st, status, sizeof(int), hipMemcpyDeviceToHost, s));
    HS_HIP(hipStreamSynchronize(s));
    pool_put(buf);
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
__global__ __launch_bounds__(64 * SW_WAVES) void sweep_kernel(int64_t r0, int m, int k, const T* __restrict__ f,
                                                              int64_t ld, T* x, int64_t ldx) {
    __shared__ T E[BS][BS + 1];
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
                for (int j = 0; j < SW_NC; ++j) xv[j] = __builtin_fma(bc[j], -e[tt], xv[j]);
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
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 1))) void bwd_node_kernel(int64_t r0, int m, int k, const T* __restrict__ f, int64_t ld,
                                                       T* x, int64_t ldx) {
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
                for (int j = 0; j < BR / 4; ++j) acc[j] = __builtin_fma(xq, -E[wave + 4 * j][q], acc[j]);
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
                    for (int q = u + 1; q < BR; ++q) acc = __builtin_fma(xr[q], -ec[q], acc);
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
                    for (int q = row + 1; q < hi; ++q) acc = __builtin_fma(xb[q][lane], -E[row][q], acc);
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
// (one MFMA accumulator) and takes the whole of k in order, 4 terms per 16x16x4 step, eight steps' operands in flight.
// A(i, p) is a[p lda + i] (ta) or a[i lda + p]. Padding: rows and columns past the ends read +0 (results dropped),
// terms past k read (+0)(-0) = -0, which changes no sum.
template <typename T, bool TA>
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
        const T v = c[(r < m ? r : m - 1) * ldc + (li < nc ? li : nc - 1)];
        acc[e] = (r < m && li < nc) ? v : T(0);
    }
    const int64_t ai = i0 + li < m ? i0 + li : m - 1;
    const int bc = li < nc ? li : nc - 1;
    const bool aok = i0 + li < m, bok = li < nc;
    constexpr int D = 8;
    for (int64_t p0 = 0; p0 < k; p0 += 4 * D) {
        T av[D], bv[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int64_t p = p0 + 4 * d + lk, pc = p < k ? p : k - 1;
            av[d] = TA ? a[pc * lda + ai] : a[ai * lda + pc];
            bv[d] = b[pc * ldb + bc];
        }
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int64_t p = p0 + 4 * d + lk;
            const T x = (aok && p < k) ? -av[d] : T(0);
            const T y = p < k ? (bok ? bv[d] : T(0)) : T(-0.0);
            acc = Mfma<T>::step(x, y, acc);
        }
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int64_t r = i0 + Mfma<T>::row(lane, e);
        if (r < m && li < nc) c[r * ldc + li] = acc[e];
    }
}

template <typename T>
int gemm_skinny(bool ta, int64_t m, int64_t nc, int64_t k, const T* a, int64_t lda, const T* b, int64_t ldb, T* c,
                int64_t ldc, hipStream_t s) {
    if (m == 0 || nc == 0 || k == 0) return OK;
    const unsigned grid = (unsigned)((m + 63) / 64);
    if (ta) gemm_skinny_kernel<T, true><<<grid, 256, 0, s>>>(m, (int)nc, k, a, lda, b, ldb, c, ldc);
    else gemm_skinny_kernel<T, false><<<grid, 256, 0, s>>>(m, (int)nc, k, a, lda, b, ldb, c, ldc);
    HS_HIP(hipGetLastError());
    return OK;
}

// Substitutions over many rows as blocks of rows that solve among themselves (block_solve_kernel) and MFMA products
// that add a finished block's terms to the rows still to solve, in order of the terms (in-order products: every item
// takes one fused multiply-add per term, in the order the substitution would).
template <typename T>
int block_solve(int mode, int64_t r0, int64_t m, int64_t k, const T* lu, int64_t lda, T* w, hipStream_t s) {
    for (int64_t c0 = 0; c0 < k; c0 += BS) {  // at most 64 columns a launch
        const int kc = (int)std::min<int64_t>(BS, k - c0);
        if (mode == BWD_LEAVES) bwd_node_kernel<T><<<1, 256, 0, s>>>(r0, (int)m, kc, lu, lda, w + c0, k);
        else if (mode == FWD_UNIT) sweep_kernel<T, FWD_UNIT><<<1, 64 * SW_WAVES, 0, s>>>(r0, (int)m, kc, lu, lda, w + c0, k);
        else if (mode == FWD_DIAG) sweep_kernel<T, FWD_DIAG><<<1, 64 * SW_WAVES, 0, s>>>(r0, (int)m, kc, lu, lda, w + c0, k);
        else sweep_kernel<T, BWD_UNIT><<<1, 64 * SW_WAVES, 0, s>>>(r0, (int)m, kc, lu, lda, w + c0, k);
        HS_HIP(hipGetLastError());
    }
    return OK;
}

template <typename T>
int fwd_blocked(int mode, int64_t n, int64_t k, const T* lu, int64_t lda, T* w, hipStream_t s) {
    for (int64_t lo = 0; lo < n; lo += BS) {
        const int64_t hi = std::min<int64_t>(n, lo + BS);
        const int rc = block_solve<T>(mode, lo, hi - lo, k, lu, lda, w, s);
        if (rc) return rc;
        if (hi < n) {
            // Rows [hi, n) take rows [lo, hi): entry (i, q) is L(i, q) = F[q][i] (FWD_UNIT) or Uᵀ(i, q) = F[i][q].
            const bool ta = mode == FWD_UNIT;
            const T* a = ta ? lu + lo * lda + hi : lu + hi * lda + lo;
            const int rc = k <= 16 ? gemm_skinny<T>(ta, n - hi, k, hi - lo, a, lda, w + lo * k, k, w + hi * k, k, s)
                                   : gemm<T>(ta, false, n - hi, k, hi - lo, T(-1), a, lda, w + lo * k, k, T(1), w + hi * k,
                                             k, s, true);
            if (rc) return rc;
        }
    }
    return OK;
}

// `backward`'s halving (factor.rs:693-707): the bottom half first, then the top half takes the bottom half's terms in
// increasing order, then the top half; nodes of at most 64 rows go to one block solve.
template <typename T>
int bwd_halving(int64_t r0, int64_t r1, int64_t n, int64_t k, const T* lu, int64_t lda, T* w, hipStream_t s) {
    if (r1 - r0 <= BS) return block_solve<T>(BWD_LEAVES, r0, r1 - r0, k, lu, lda, w, s);
    const int64_t mid = (r0 + r1) / 2;
    int rc = bwd_halving<T>(mid, r1, n, k, lu, lda, w, s);
    // Rows [r0, mid) take rows [mid, r1): entry (i, q) is U(i, q) = F[q][i].
    if (rc == OK)
        rc = k <= 16 ? gemm_skinny<T>(true, mid - r0, k, r1 - mid, lu + mid * lda + r0, lda, w + mid * k, k, w + r0 * k, k, s)
                     : gemm<T>(true, false, mid - r0, k, r1 - mid, T(-1), lu + mid * lda + r0, lda, w + mid * k, k, T(1),
                               w + r0 * k, k, s, true);
    if (rc == OK) rc = bwd_halving<T>(r0, mid, n, k, lu, lda, w, s);
    return rc;
}

}  // namespace

template <typename T>
int getrs(bool trans, int64_t n, int64_t k, const T* lu, int64_t lda, const int64_t* perm, T* b, int64_t ldb, hipStream_t s) {
    if (n == 0 || k == 0) return OK;
    T* w = (T*)pool_get(sizeof(T) * (size_t)(n * k));
    if (!w) return hip_status(hipErrorOutOfMemory, "pool_get");
    int rc = OK;
    const int64_t total = n * k;
    if (!trans) {
        gather_rows<T><<<(unsigned)((total + 255) / 256), 256, 0, s>>>(n, (int)k, b, ldb, perm, w, k);
        HS_HIP(hipGetLastError());
        if (k == 1) {
            rc = trsv1_lu<T>(n, lu, lda, w, s);
        } else {
            rc = fwd_blocked<T>(FWD_UNIT, n, k, lu, lda, w, s);
            if (rc == OK) rc = bwd_halving<T>(0, n, n, k, lu, lda, w, s);
        }
        if (rc == OK) {
            HS_HIP(hipMemcpy2DAsync(b, ldb * sizeof(T), w, k * sizeof(T), k * sizeof(T), n, hipMemcpyDeviceToDevice, s));
        }
    } else {
        HS_HIP(hipMemcpy2DAsync(w, k * sizeof(T), b, ldb * sizeof(T), k * sizeof(T), n, hipMemcpyDeviceToDevice, s));
        rc = fwd_blocked<T>(FWD_DIAG, n, k, lu, lda, w, s);
        for (int64_t c0 = 0; c0 < k && rc == OK; c0 += MAXK) {
            const int kc = (int)std::min<int64_t>(MAXK, k - c0);
            rc = trsv_pass<T>(BWD_UNIT, 0, n, n, kc, lu, lda, w + c0, k, s);
        }
        if (rc == OK) {
            scatter_rows<T><<<(unsigned)((total + 255) / 256), 256, 0, s>>>(n, (int)k, w, k, perm, b, ldb);
            HS_HIP(hipGetLastError());
        }
    }
    HS_HIP(hipStreamSynchronize(s));
    pool_put(w);
    return rc;
}

template int getrs<double>(bool, int64_t, int64_t, const double*, int64_t, const int64_t*, double*, int64_t, hipStream_t);
template int getrs<float>(bool, int64_t, int64_t, const float*, int64_t, const int64_t*, float*, int64_t, hipStream_t);

}  // namespace hs
