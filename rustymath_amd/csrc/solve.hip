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

// One triangular pass of `mode` over all n rows of x (n×k, k <= MAXK) in place.
template <typename T>
int trsv_pass(int mode, int64_t n, int k, const T* f, int64_t ld, T* x, int64_t ldx, hipStream_t s) {
    std::vector<std::pair<int, int>> blocks;  // in solve order
    const bool down = mode == FWD_UNIT || mode == FWD_DIAG;
    if (mode == BWD_LEAVES) {
        std::vector<std::pair<int, int>> lv;
        backward_leaves(0, (int)n, lv);
        blocks.assign(lv.rbegin(), lv.rend());
    } else if (down) {
        for (int r = 0; r < n; r += BR) blocks.push_back({r, (int)std::min<int64_t>(n, r + BR)});
    } else {
        int nb = (int)((n + BR - 1) / BR);
        for (int q = nb - 1; q >= 0; --q) blocks.push_back({q * BR, (int)std::min<int64_t>(n, (int64_t)(q + 1) * BR)});
    }
    const int nb = (int)blocks.size();
    std::vector<int> r0(nb), r1(nb), of_row(n), seg_off(nb + 1), seg;
    for (int b = 0; b < nb; ++b) {
        r0[b] = blocks[b].first; r1[b] = blocks[b].second;
        for (int r = r0[b]; r < r1[b]; ++r) of_row[r] = b;
        seg_off[b] = (int)seg.size() / 3;
        if (mode == BWD_LEAVES) {
            std::vector<std::pair<int, int>> rs;
            backward_ranges(0, (int)n, r0[b], rs);
            for (auto [a0, a1] : rs) { seg.push_back(a0); seg.push_back(a1); seg.push_back(1); }
        } else if (down) {
            if (r0[b] > 0) { seg.push_back(0); seg.push_back(r0[b]); seg.push_back(1); }
        } else {
            if (r1[b] < n) { seg.push_back(r1[b]); seg.push_back((int)n); seg.push_back(0); }
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
        if (k == 1) rc = trsv1_lu<T>(n, lu, lda, w, s);
        for (int64_t c0 = 0; c0 < k && rc == OK && k > 1; c0 += MAXK) {
            int kc = (int)std::min<int64_t>(MAXK, k - c0);
            rc = trsv_pass<T>(FWD_UNIT, n, kc, lu, lda, w + c0, k, s);
            if (rc == OK) rc = trsv_pass<T>(k == 1 ? BWD_SINGLE : BWD_LEAVES, n, kc, lu, lda, w + c0, k, s);
        }
        if (rc == OK) {
            HS_HIP(hipMemcpy2DAsync(b, ldb * sizeof(T), w, k * sizeof(T), k * sizeof(T), n, hipMemcpyDeviceToDevice, s));
        }
    } else {
        HS_HIP(hipMemcpy2DAsync(w, k * sizeof(T), b, ldb * sizeof(T), k * sizeof(T), n, hipMemcpyDeviceToDevice, s));
        for (int64_t c0 = 0; c0 < k && rc == OK; c0 += MAXK) {
            int kc = (int)std::min<int64_t>(MAXK, k - c0);
            rc = trsv_pass<T>(FWD_DIAG, n, kc, lu, lda, w + c0, k, s);
            if (rc == OK) rc = trsv_pass<T>(BWD_UNIT, n, kc, lu, lda, w + c0, k, s);
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
