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
#include <vector>
#include <mutex>
#include <algorithm>

namespace hs {

namespace {

constexpr int PW = 64;            // widest panel
constexpr int PT = 256;           // threads of a panel workgroup
constexpr int MAX_WG = 64;        // panel workgroups: one wave polls one flag per lane
constexpr int PANEL_ROWS = 256;   // rows of a panel workgroup (64 per wave)
constexpr uint64_t SPIN_LIMIT = 1ull << 22;

template <typename T> struct Bits;
template <> struct Bits<double> { typedef unsigned long long u; };
template <> struct Bits<float> { typedef unsigned int u; };

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
__device__ __forceinline__ void put64(long long* p, long long v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ long long get64(const long long* p) {
    return __hip_atomic_load(const_cast<long long*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One workgroup's published candidate for one column: the magnitude (-1 for none or NaN), its row, the
// candidate row's items, and (from the workgroup holding it) the current row at the pivot position.
template <typename T> struct Slot {
    T val;
    long long row;
    T cand[PW];
    T rowc[PW];
};

// Prefers the larger magnitude, then the smaller row: `largest`'s first maximum (factor.rs:476).
template <typename T>
__device__ __forceinline__ void better(T& v, long long& r, T v2, long long r2) {
    if (v2 > v || (v2 == v && r2 < r)) { v = v2; r = r2; }
}

template <typename T>
__device__ __forceinline__ void wave_best(T& v, long long& r) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        T v2 = __shfl_xor(v, off);
        long long r2 = __shfl_xor(r, off);
        better(v, r, v2, r2);
    }
}

// Factors the panel of columns [c0, c0 + w) over rows [c0, c0 + rows) of the row-major `a`. Workgroup g
// holds rows [4R g, 4R (g + 1)) in registers, wave q of it rows [4R g + R q, + R), lane l column l: so a
// column's items sit in one lane of each wave, a row's in one wave, and every step of a column loads,
// stores and updates whole rows at once. Every column takes one exchange between the workgroups: each
// publishes its best candidate, waits for every other's, and all pick the same pivot. Writes
// pivots[c0 + c] (absolute rows) and `moves`, the net row permutation of the panel's swaps (pairs dst,
// src), for the row-swap kernel. `leaf_end[c]` is the end of the reference leaf of column c.
template <typename T, int R>
__global__ __launch_bounds__(PT) void panel_kernel(T* __restrict__ a, int64_t lda, int64_t c0, int64_t rows, int w,
                                                   int64_t* __restrict__ pivots, long long* __restrict__ moves,
                                                   const int* __restrict__ leaf_end, Slot<T>* slots,
                                                   unsigned long long* flags, unsigned long long epoch, int nwg,
                                                   int* status) {
    const int tid = threadIdx.x, col = tid & 63, grp = tid >> 6, g = blockIdx.x;
    constexpr int64_t PER = 4 * (int64_t)R;
    const int64_t base = (int64_t)g * PER + (int64_t)grp * R;  // this wave's first row, relative to c0

    T v[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
        const int64_t r = base + j;
        v[j] = (r < rows && col < w) ? a[(c0 + r) * lda + c0 + col] : T(0);
    }

    __shared__ T s_val[4];
    __shared__ long long s_row[4];
    __shared__ T s_piv[PW];
    __shared__ T s_rowc[PW];
    __shared__ T s_l[4 * R];
    __shared__ long long s_p;
    __shared__ int s_abort;
    // The net permutation of the swaps so far (workgroup 0 only): content of block position q, and the
    // positions outside the block that a swap has touched with their content.
    __shared__ long long s_blk[PW];
    __shared__ long long s_opos[PW];
    __shared__ long long s_oval[PW];
    __shared__ int s_nout;
    if (tid < PW) s_blk[tid] = c0 + tid;
    if (tid == 0) { s_nout = 0; s_abort = 0; }
    __syncthreads();

    for (int c = 0; c < w; ++c) {
        // Local candidate: the first row at or below c with the greatest magnitude, NaN never chosen.
        if (col == c) {
            T best = T(-1);
            long long brow = 0x7fffffffffffffffll;
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const int64_t r = base + j;
                if (r < rows && r >= c) {
                    T x = fabs(v[j]);
                    if (x != x) x = T(-1);
                    better(best, brow, x, (long long)r);
                }
            }
            s_val[grp] = best;
            s_row[grp] = brow;
        }
        __syncthreads();
        T wb = s_val[0];
        long long wr = s_row[0];
#pragma unroll
        for (int q = 1; q < 4; ++q) better(wb, wr, s_val[q], s_row[q]);

        // Publish: the candidate row and, from the wave holding it, the row now at position c.
        Slot<T>* mine = slots + (c & 1) * MAX_WG + g;
        if (col < w) {
            const int64_t jw = wr - base, jc = (int64_t)c - base;
            T xw = T(0), xc = T(0);
#pragma unroll
            for (int j = 0; j < R; ++j) { if (j == jw) xw = v[j]; if (j == jc) xc = v[j]; }
            if (jw >= 0 && jw < R) put(&mine->cand[col], xw);
            if (jc >= 0 && jc < R) put(&mine->rowc[col], xc);
        }
        if (tid == 0) { put(&mine->val, wb); put64(&mine->row, wr); }
        __syncthreads();
        const unsigned long long target = epoch * 1024ull + (unsigned long long)c + 1ull;
        if (tid == 0) __hip_atomic_store(&flags[g], target, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);

        // Wait for every workgroup's candidate, then agree on the pivot.
        if (grp == 0) {
            const int lane = col;
            T gv = T(-1);
            long long gr = 0x7fffffffffffffffll;
            bool ok = true;
            if (lane < nwg) {
                uint64_t spins = 0;
                while (__hip_atomic_load(&flags[lane], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
                    if (++spins > SPIN_LIMIT) { ok = false; break; }
                    __builtin_amdgcn_s_sleep(1);
                }
                if (ok) {
                    const Slot<T>* sl = slots + (c & 1) * MAX_WG + lane;
                    gv = get(&sl->val);
                    gr = get64(&sl->row);
                }
            }
            const bool all = __all(ok);
            wave_best(gv, gr);
            if (lane == 0) {
                s_p = gr;
                if (!all) { s_abort = 1; atomicExch(status, ERR_LIMIT); }
            }
        }
        __syncthreads();
        if (s_abort) return;
        const long long p = s_p;
        const int gp = (int)(p / PER), gc = (int)(c / PER);
        if (tid < w) {
            s_piv[tid] = get(&slots[(c & 1) * MAX_WG + gp].cand[tid]);
            s_rowc[tid] = get(&slots[(c & 1) * MAX_WG + gc].rowc[tid]);
        }
        if (g == 0 && tid == 0) {
            pivots[c0 + c] = c0 + p;
            // Track the net permutation: swap the contents of positions c0 + c and c0 + p.
            const long long pos = c0 + p;
            const long long x = s_blk[c];
            if (p < w) { s_blk[c] = s_blk[p]; s_blk[p] = x; }
            else {
                int k = 0;
                while (k < s_nout && s_opos[k] != pos) ++k;
                const long long y = k < s_nout ? s_oval[k] : pos;
                if (k == s_nout) { s_opos[k] = pos; ++s_nout; }
                s_oval[k] = x;
                s_blk[c] = y;
            }
        }
        __syncthreads();

        // Swap rows c and p, then divide column c below the pivot and update the columns after it.
        const T pivot = s_piv[c];
        const bool zero = pivot == T(0);
        const bool normal = fabs(pivot) >= (sizeof(T) == 8 ? T(2.2250738585072014e-308) : T(1.17549435e-38f));
        const T rcp = zero ? T(0) : T(1) / pivot;
        if (col < w) {
            const T pc = s_piv[col], rc = s_rowc[col];
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const int64_t r = base + j;
                if (r == c) v[j] = pc;
                else if (r == p) v[j] = rc;
            }
        }
        if (col == c) {
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const int64_t r = base + j;
                T l = v[j];
                if (r < rows && r > c && !zero) {
                    l = normal ? l * rcp : l / pivot;
                    v[j] = l;
                }
                s_l[grp * R + j] = l;
            }
        }
        __syncthreads();
        const int le = leaf_end[c0 + c] - (int)c0;
        if (col > c && col < w && (!zero || col >= le)) {
            const T u = -s_piv[col];
#pragma unroll
            for (int j = 0; j < R; ++j) {
                const int64_t r = base + j;
                if (r < rows && r > c) v[j] = __builtin_fma(s_l[grp * R + j], u, v[j]);
            }
        }
    }

#pragma unroll
    for (int j = 0; j < R; ++j) {
        const int64_t r = base + j;
        if (r < rows && col < w) a[(c0 + r) * lda + c0 + col] = v[j];
    }
    if (g == 0) {
        __syncthreads();
        // moves: [0] = count, then pairs (dst, src).
        if (tid == 0) moves[0] = w + s_nout;
        if (tid < w) { moves[1 + 2 * tid] = c0 + tid; moves[2 + 2 * tid] = s_blk[tid]; }
        if (tid < s_nout) { moves[1 + 2 * (w + tid)] = s_opos[tid]; moves[2 + 2 * (w + tid)] = s_oval[tid]; }
    }
}

// Applies a panel's net row permutation to the columns outside it: for columns [0, c0) and [c1, n),
// every moved row takes the old item of its source row. Each block of 64 columns reads all sources into
// LDS before writing, so chains of swaps cost one read and one write per moved item.
template <typename T>
__global__ __launch_bounds__(256) void swap_kernel(T* __restrict__ a, int64_t lda, int64_t n, int64_t c0, int64_t c1,
                                                   const long long* __restrict__ moves) {
    __shared__ T tmp[2 * PW][64];
    const int nm = (int)moves[0];
    const int tid = threadIdx.x, col_l = tid & 63, grp = tid >> 6;
    int64_t col = (int64_t)blockIdx.x * 64 + col_l;  // over the columns outside [c0, c1)
    if (col >= c0) col += c1 - c0;
    const bool live = col < n;
    for (int k = grp; k < nm; k += 4) {
        long long src = moves[2 + 2 * k];
        if (live) tmp[k][col_l] = a[src * lda + col];
    }
    __syncthreads();
    for (int k = grp; k < nm; k += 4) {
        long long dst = moves[1 + 2 * k];
        if (live) a[dst * lda + col] = tmp[k][col_l];
    }
}

// U12 = L11⁻¹ A12 for rows [c0, c0 + w), columns [c0 + w, n): a workgroup takes 128 columns, stages
// them and L11 through LDS with coalesced row reads, and each thread solves one column in registers,
// each row taking its terms in order of the pivot column (factor.rs:668-689's leaf order).
constexpr int TC = 128;
template <typename T>
__global__ __launch_bounds__(TC) void trsm_block_kernel(T* __restrict__ a, int64_t lda, int64_t n, int64_t c0, int w) {
    __shared__ T L[PW][PW + 1];
    __shared__ T X[PW][TC + 1];
    const int t = threadIdx.x;
    const int64_t col0 = c0 + w + (int64_t)blockIdx.x * TC;
    for (int idx = t; idx < w * w; idx += TC) {
        int r = idx / w, q = idx % w;
        L[r][q] = a[(c0 + r) * lda + c0 + q];
    }
    for (int r = 0; r < w; ++r) X[r][t] = col0 + t < n ? a[(c0 + r) * lda + col0 + t] : T(0);
    __syncthreads();
    T x[PW];
#pragma unroll
    for (int r = 0; r < PW; ++r) x[r] = X[r][t];
#pragma unroll
    for (int q = 0; q < PW; ++q) {
#pragma unroll
        for (int r = q + 1; r < PW; ++r)
            if (r < w) x[r] = __builtin_fma(-L[r][q], x[q], x[r]);
        __builtin_amdgcn_sched_barrier(0);  // keeps each step's LDS reads with its step
    }
#pragma unroll
    for (int r = 0; r < PW; ++r) X[r][t] = x[r];
    __syncthreads();
    for (int r = 0; r < w; ++r) if (col0 + t < n) a[(c0 + r) * lda + col0 + t] = X[r][t];
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

// The reference's leaves: lu_panel halves [c0, c1) at (c0 + c1) / 2 until 16 columns or fewer remain.
void leaves(int64_t c0, int64_t c1, std::vector<int64_t>& ends) {
    if (c1 - c0 > 16) { int64_t mid = (c0 + c1) / 2; leaves(c0, mid, ends); leaves(mid, c1, ends); return; }
    ends.push_back(c1);
}

// Per-device state of the panel exchange: slots and flags, allocated once and never freed while the
// library is loaded; the epoch makes each launch's flag targets larger than any value left behind.
struct PanelState {
    void* slots = nullptr;
    unsigned long long* flags = nullptr;
    int* status = nullptr;
    long long* moves = nullptr;
    unsigned long long epoch = 0;
};
std::mutex g_panel_mu;
PanelState g_panel[16];

int panel_state(PanelState*& st) {
    int dev = 0;
    HS_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_panel_mu);
    st = &g_panel[dev & 15];
    if (!st->slots) {
        HS_HIP(hipMalloc(&st->slots, 2 * MAX_WG * sizeof(Slot<double>)));
        HS_HIP(hipMalloc(&st->flags, MAX_WG * sizeof(unsigned long long)));
        HS_HIP(hipMalloc(&st->status, sizeof(int)));
        HS_HIP(hipMalloc(&st->moves, (1 + 4 * PW) * sizeof(long long)));
        HS_HIP(hipMemset(st->flags, 0, MAX_WG * sizeof(unsigned long long)));
        HS_HIP(hipMemset(st->status, 0, sizeof(int)));
        HS_HIP(hipDeviceSynchronize());
    }
    return OK;
}

}  // namespace

template <typename T>
int getrf(int64_t n, T* a, int64_t lda, int64_t* pivots, hipStream_t s) {
    if (n == 0) return OK;
    PanelState* st;
    int e = panel_state(st);
    if (e) return e;
    // Panels: consecutive whole reference leaves, at most PW columns each.
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
    int* d_leaf = (int*)pool_get(n * sizeof(int));
    if (!d_leaf) return hip_status(hipErrorOutOfMemory, "pool_get");
    HS_HIP(hipMemcpyAsync(d_leaf, leaf_end.data(), n * sizeof(int), hipMemcpyHostToDevice, s));

    int rc = OK;
    for (auto [c0, c1] : panels) {
        const int w = (int)(c1 - c0);
        const int64_t rows = n - c0;
        if (rows > (int64_t)MAX_WG * PANEL_ROWS) { set_error("n above 16384: the panel kernel holds at most 64 x 256 rows"); rc = -1; break; }
        const int nwg = (int)((rows + PANEL_ROWS - 1) / PANEL_ROWS);
        unsigned long long epoch;
        {
            std::lock_guard<std::mutex> lk(g_panel_mu);
            epoch = ++st->epoch;
        }
        auto slots = reinterpret_cast<Slot<T>*>(st->slots);
        panel_kernel<T, PANEL_ROWS / 4><<<nwg, PT, 0, s>>>(a, lda, c0, rows, w, pivots, st->moves, d_leaf, slots, st->flags, epoch, nwg, st->status);
        HS_HIP(hipGetLastError());
        const int64_t outside = n - w;
        if (outside > 0) {
            swap_kernel<T><<<(unsigned)((outside + 63) / 64), 256, 0, s>>>(a, lda, n, c0, c1, st->moves);
            HS_HIP(hipGetLastError());
        }
        if (c1 < n) {
            const int64_t cols = n - c1;
            trsm_block_kernel<T><<<(unsigned)((cols + TC - 1) / TC), TC, 0, s>>>(a, lda, n, c0, w);
            HS_HIP(hipGetLastError());
            rc = gemm<T>(false, false, n - c1, n - c1, w, T(-1), a + c1 * lda + c0, lda, a + c0 * lda + c1, lda, T(1),
                         a + c1 * lda + c1, lda, s);
            if (rc) break;
        }
    }
    if (rc == OK) {
        dim3 grid((unsigned)((n + 31) / 32), (unsigned)((n + 31) / 32)), block(32, 8);
        transpose_square_inplace<T><<<grid, block, 0, s>>>(a, lda, n);
        HS_HIP(hipGetLastError());
        int status = 0;
        HS_HIP(hipMemcpyAsync(&status, st->status, sizeof(int), hipMemcpyDeviceToHost, s));
        HS_HIP(hipStreamSynchronize(s));
        if (status) {
            HS_HIP(hipMemsetAsync(st->status, 0, sizeof(int), s));
            set_error("panel exchange wait ran out (workgroups not co-resident?)");
            rc = ERR_LIMIT;
        }
    } else {
        HS_HIP(hipStreamSynchronize(s));
    }
    pool_put(d_leaf);
    return rc;
}

template int getrf<double>(int64_t, double*, int64_t, int64_t*, hipStream_t);
template int getrf<float>(int64_t, float*, int64_t, int64_t*, hipStream_t);

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
