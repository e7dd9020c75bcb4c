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
constexpr int PT = 512;           // threads (rows) of a panel workgroup
constexpr int MAX_WG = 32;        // panel workgroups: one per CU of one XCD, one flag per lane of a wave
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
};

// Prefers the larger magnitude, then the smaller position: `largest`'s first maximum (factor.rs:476). Selects,
// not branches, so that a reduction step costs a few instructions.
template <typename T>
__device__ __forceinline__ bool better(T v, int r, T v2, int r2) {
    return (v2 > v) | ((v2 == v) & (r2 < r));
}

template <typename T>
__device__ __forceinline__ void wave_best(T& v, int& r) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const T v2 = __shfl_xor(v, off);
        const int r2 = __shfl_xor(r, off);
        const bool take = better(v, r, v2, r2);
        v = take ? v2 : v;
        r = take ? r2 : r;
    }
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
// With one workgroup, a column's pivot takes workgroup barriers only. With several (launch indices 0, 8,
// 16, ..., one XCD under round-robin dispatch, one workgroup per CU, so all are resident; the other launch
// indices return at once), every column takes one exchange: each workgroup publishes its best candidate
// (magnitude, position, the candidate row) with device-coherent (sc1) stores, raises its flag, waits for
// every other flag, and all pick the same pivot. Candidates compare by magnitude, NaN never chosen, then by
// position: `largest`'s first maximum (core/src/linalg/factor.rs:476).
//
// Writes pivots[c0 + c] (absolute rows) and `moves` (count, then pairs dst, src: the rows whose position
// changed), for the row-swap kernel. `leaf_end[c]` is the end of the reference leaf of column c; a zero
// pivot leaves its column undivided and skips the rest of that leaf (factor.rs:606). Where `prof` is not
// null, workgroup 0 adds the ticks of each phase of the column loop into prof[0..7].
template <typename T>
__global__ __launch_bounds__(PT) void panel_kernel(T* __restrict__ a, int64_t lda, int64_t c0, int64_t rows, int w,
                                                   int64_t* __restrict__ pivots, long long* __restrict__ moves,
                                                   const int* __restrict__ leaf_end, T* __restrict__ scratch,
                                                   char* state, unsigned long long epoch, int nwg, int* status,
                                                   long long* prof) {
    if (nwg > 1 && (blockIdx.x & 7)) return;
    const int g = nwg > 1 ? (int)(blockIdx.x >> 3) : 0;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    constexpr int WAVES = PT / 64;
    unsigned long long* flags = reinterpret_cast<unsigned long long*>(state);
    Slot<T>* slots = reinterpret_cast<Slot<T>*>(state + MAX_WG * sizeof(unsigned long long));
    const bool timed = prof != nullptr && g == 0 && threadIdx.x == 0;
    long long tp[8] = {0, 0, 0, 0, 0, 0, 0, 0}, tl = timed ? ticks() : 0;
#define HS_PHASE(q) do { if (timed) { long long t_ = ticks(); tp[q] += t_ - tl; tl = t_; } } while (0)

    const int64_t phys = (int64_t)g * PT + tid;  // this thread's row, relative to c0
    const bool has = phys < rows;
    int pos = (int)phys;
    T v[PW];
    {
        // All loads first, from addresses clamped into the panel, then the selects: a load behind a per-lane
        // condition would be issued and waited for one at a time.
        const T* src = a + (c0 + (has ? phys : 0)) * lda + c0;
#pragma unroll
        for (int j = 0; j < PW; ++j) v[j] = src[j < w ? j : 0];
        asm volatile("" ::: "memory");
#pragma unroll
        for (int j = 0; j < PW; ++j) v[j] = (has && j < w) ? v[j] : T(0);
    }

    // Double-buffered by column, so that a workgroup alone needs one barrier per column: each wave's best
    // candidate (magnitude, position, row), the pivot rows of U, and the pivot row of the current column.
    __shared__ T s_val[2][WAVES];
    __shared__ int s_key[2][WAVES];
    __shared__ T s_cand[2][WAVES][PW];
    __shared__ T s_u11[PW][PW];
    __shared__ T s_piv[PW];
    __shared__ int s_p, s_gp, s_abort;
    if (tid == 0) { s_abort = 0; if (g == 0) moves[0] = 0; }
    const T tiny = sizeof(T) == 8 ? T(2.2250738585072014e-308) : T(1.17549435e-38f);

#pragma clang loop unroll(disable)
    for (int c = 0; c < w; ++c) {
        const int buf = c & 1;
        // This wave's candidate, published with its row.
        T val = T(-1);
        int key = 0x7fffffff;
        if (has && pos >= c) {
            const T x = fabs(v[0]);
            val = x != x ? T(-1) : x;
            key = pos;
        }
        T bv = val;
        int bk = key;
        wave_best(bv, bk);
        if (bk != 0x7fffffff && key == bk) {
#pragma unroll
            for (int j = 0; j < PW; ++j) s_cand[buf][wave][j] = v[j];
        }
        if (lane == 0) { s_val[buf][wave] = bv; s_key[buf][wave] = bk; }
        __syncthreads();
        HS_PHASE(0);
        T wv[WAVES];
        int wk[WAVES];
#pragma unroll
        for (int q = 0; q < WAVES; ++q) { wv[q] = s_val[buf][q]; wk[q] = s_key[buf][q]; }
        T wb = wv[0];
        int wbk = wk[0], win = 0;
#pragma unroll
        for (int q = 1; q < WAVES; ++q) {
            const bool take = better(wb, wbk, wv[q], wk[q]);
            wb = take ? wv[q] : wb;
            wbk = take ? wk[q] : wbk;
            win = take ? q : win;
        }
        const T* cand = s_cand[buf][win];
        const T* piv = cand;
        int p = wbk;
        HS_PHASE(1);

        if (nwg > 1) {
            Slot<T>* mine = slots + (c & 1) * MAX_WG + g;
            if (tid < w - c) put(&mine->cand[tid], cand[tid]);
            if (tid == 0) { put(&mine->val, wb); put64(&mine->row, wbk); }
            // Every slot store of this workgroup reaches L2 before the flag can: the barrier alone does not wait
            // for stores.
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            const unsigned long long target = epoch * 1024ull + (unsigned long long)c + 1ull;
            if (tid == 0) __hip_atomic_store(&flags[g], target, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (wave == 0) {
                T gv = T(-1);
                int gk = 0x7fffffff, who = lane;
                bool ok = true;
                if (lane < nwg) {
                    uint64_t spins = 0;
                    while (__hip_atomic_load(&flags[lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
                        if (++spins > SPIN_LIMIT) { ok = false; break; }
                    }
                    if (ok) {
                        const Slot<T>* sl = slots + (c & 1) * MAX_WG + lane;
                        gv = get(&sl->val);
                        gk = (int)get64(&sl->row);
                    }
                }
                const bool all = __all(ok);
                // Reduce (magnitude, position) and remember which workgroup holds the winner.
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) {
                    const T v2 = __shfl_xor(gv, off);
                    const int k2 = __shfl_xor(gk, off), w2 = __shfl_xor(who, off);
                    const bool take = better(gv, gk, v2, k2);
                    gv = take ? v2 : gv;
                    gk = take ? k2 : gk;
                    who = take ? w2 : who;
                }
                if (lane == 0) {
                    s_p = gk;
                    s_gp = who;
                    if (!all) { s_abort = 1; atomicExch(status, ERR_LIMIT); }
                }
            }
            __syncthreads();
            HS_PHASE(2);
            if (s_abort) return;
            p = s_p;
            const int gp = s_gp;
            if (tid < PW) s_piv[tid] = gp == g ? cand[tid] : (tid < w - c ? get(&slots[(c & 1) * MAX_WG + gp].cand[tid]) : T(0));
            __syncthreads();
            piv = s_piv;
            HS_PHASE(3);
        }

        // The winner takes position c; its row of U is kept in LDS. The row at position c moves to p.
        if (tid < PW) s_u11[c][tid] = piv[tid];
        const T pivot = piv[0];
        if (has && pos == p) pos = c;
        else if (has && pos == c) pos = p;
        if (g == 0 && tid == 0) pivots[c0 + c] = c0 + p;

        // Rows below the pivot: divide (a quotient for a subnormal pivot, nothing for a zero one), record the
        // multiplier, update the rest of the row and shift it one column left.
        if (has && pos > c) {
            // Every item takes fma(l, -u, a). A zero pivot (rare; the same for all threads) leaves l undivided and
            // skips columns c + j, j < skip <= 16 (a leaf's width): only the first sixteen items have a second,
            // guarded path, so that the common path has no per-item condition.
            T l = v[0];
            const bool zero = pivot == T(0);
            if (!zero) {
                l = fabs(pivot) >= tiny ? l * (T(1) / pivot) : l / pivot;
#pragma unroll
                for (int j = 1; j < 16; ++j) v[j - 1] = __builtin_fma(l, -piv[j], v[j]);
            } else {
                const int skip = leaf_end[c0 + c] - (int)c0 - c;
#pragma unroll
                for (int j = 1; j < 16; ++j) v[j - 1] = j >= skip ? __builtin_fma(l, -piv[j], v[j]) : v[j];
            }
            asm volatile("" ::: "memory");
#pragma unroll
            for (int j0 = 16; j0 < PW; j0 += 16) {
#pragma unroll
                for (int j = j0; j < j0 + 16 && j < PW; ++j) v[j - 1] = __builtin_fma(l, -piv[j], v[j]);
                asm volatile("" ::: "memory");  // sixteen items of the pivot row in registers at a time
            }
            v[PW - 1] = T(0);
            scratch[(int64_t)c * rows + phys] = l;
        }
        HS_PHASE(4);
    }
    if (timed)
        for (int q = 0; q < 8; ++q) atomicAdd((unsigned long long*)&prof[q], (unsigned long long)tp[q]);
#undef HS_PHASE
    __syncthreads();  // every row of U is in s_u11

    if (has) {
        // The row at its final position: multipliers from scratch left of it, its row of U (pivot rows) from LDS.
        // The scratch items are all loaded before any is used (see the first loads).
        T* dst = a + (c0 + pos) * lda + c0;
#pragma unroll
        for (int j = 0; j < PW; ++j) v[j] = scratch[(int64_t)(j < w ? j : 0) * rows + phys];
        asm volatile("" ::: "memory");
        const int upos = pos < PW ? pos : PW - 1;
#pragma unroll
        for (int j = 0; j < PW; ++j)
            if (j < w) dst[j] = j < pos ? v[j] : s_u11[upos][(j - pos) & (PW - 1)];
        if (pos != phys) {
            const long long k = atomicAdd((unsigned long long*)moves, 1ull);
            moves[1 + 2 * k] = c0 + pos;
            moves[2 + 2 * k] = c0 + phys;
        }
    }
}

// Applies a panel's net row permutation to the columns outside it: for columns [0, c0) and [c1, n),
// every moved row takes the old item of its source row. Each block of 64 columns reads all sources into
// LDS before writing, so chains of swaps cost one read and one write per moved item.
template <typename T>
__global__ __launch_bounds__(256) void swap_kernel(T* __restrict__ a, int64_t lda, int64_t n, int64_t c0, int64_t c1,
                                                   const long long* __restrict__ moves) {
    __shared__ T tmp[2 * PW][64];
    __shared__ long long mv[1 + 4 * PW];
    for (int q = threadIdx.x; q < 1 + 4 * PW; q += blockDim.x) mv[q] = moves[q];
    __syncthreads();
    const int nm = (int)mv[0];
    if (nm == 0) return;
    const int tid = threadIdx.x, col_l = tid & 63, grp = tid >> 6;
    int64_t col = (int64_t)blockIdx.x * 64 + col_l;  // over the columns outside [c0, c1)
    if (col >= c0) col += c1 - c0;
    const bool live = col < n;
    const int64_t lc = live ? col : 0;
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
    // All loads in flight at once, from addresses clamped into the matrix; zeros then pad a narrow panel.
    T lv[PW * PW / TC], xv[PW];
    const int64_t col = col0 + t < n ? col0 + t : n - 1;
#pragma unroll
    for (int q = 0; q < PW * PW / TC; ++q) {
        const int idx = t + q * TC, r = idx / PW, c = idx % PW;
        lv[q] = a[(c0 + (r < w ? r : w - 1)) * lda + c0 + (c < w ? c : w - 1)];
    }
#pragma unroll
    for (int r = 0; r < PW; ++r) xv[r] = a[(c0 + (r < w ? r : w - 1)) * lda + col];
    asm volatile("" ::: "memory");
#pragma unroll
    for (int q = 0; q < PW * PW / TC; ++q) {
        const int idx = t + q * TC, r = idx / PW, c = idx % PW;
        L[r][c] = (r < w && c < w) ? lv[q] : T(0);
    }
#pragma unroll
    for (int r = 0; r < PW; ++r) X[r][t] = (r < w && col0 + t < n) ? xv[r] : T(0);
    __syncthreads();
    T x[PW];
#pragma unroll
    for (int r = 0; r < PW; ++r) x[r] = X[r][t];
    // Padding rows and columns are zero, so the full 64-row sweep leaves the real rows' bits unchanged.
#pragma unroll
    for (int q = 0; q < PW; ++q) {
#pragma unroll
        for (int r = q + 1; r < PW; ++r) x[r] = __builtin_fma(-L[r][q], x[q], x[r]);
    }
#pragma unroll
    for (int r = 0; r < PW; ++r) X[r][t] = x[r];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < PW; ++r)
        if (r < w && col0 + t < n) a[(c0 + r) * lda + col0 + t] = X[r][t];
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
    char* state = nullptr;  // MAX_WG flags, then 2 x MAX_WG slots
    long long* prof = nullptr;
    bool prof_on = false;
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
    if (!st->state) {
        const size_t bytes = MAX_WG * sizeof(unsigned long long) + 2 * MAX_WG * sizeof(Slot<double>);
        HS_HIP(hipMalloc(&st->state, bytes));
        HS_HIP(hipMemset(st->state, 0, bytes));
        HS_HIP(hipMalloc(&st->status, sizeof(int)));
        HS_HIP(hipMalloc(&st->prof, 8 * sizeof(long long)));
        HS_HIP(hipMemset(st->prof, 0, 8 * sizeof(long long)));
        HS_HIP(hipMalloc(&st->moves, (1 + 4 * PW) * sizeof(long long)));
        HS_HIP(hipMemset(st->status, 0, sizeof(int)));
        HS_HIP(hipDeviceSynchronize());
    }
    return OK;
}

}  // namespace

// The panel kernel's per-phase profile (ticks of 10 ns, summed over launches): on = 1 clears and enables,
// on = 0 disables; out (8 values) receives the sums where not null.
int panel_profile(int on, long long* out) {
    PanelState* st;
    int e = panel_state(st);
    if (e) return e;
    if (out) HS_HIP(hipMemcpy(out, st->prof, 8 * sizeof(long long), hipMemcpyDeviceToHost));
    if (on) HS_HIP(hipMemset(st->prof, 0, 8 * sizeof(long long)));
    st->prof_on = on != 0;
    return OK;
}

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
    T* scratch = (T*)pool_get(sizeof(T) * (size_t)n * PW);
    if (!d_leaf || !scratch) { pool_put(d_leaf); pool_put(scratch); return hip_status(hipErrorOutOfMemory, "pool_get"); }
    HS_HIP(hipMemcpyAsync(d_leaf, leaf_end.data(), n * sizeof(int), hipMemcpyHostToDevice, s));

    int rc = OK;
    for (auto [c0, c1] : panels) {
        const int w = (int)(c1 - c0);
        const int64_t rows = n - c0;
        if (rows > (int64_t)MAX_WG * PT) { set_error("n above 16384: the panel kernel holds at most 32 x 512 rows"); rc = -1; break; }
        const int nwg = (int)((rows + PT - 1) / PT);
        unsigned long long epoch;
        {
            std::lock_guard<std::mutex> lk(g_panel_mu);
            epoch = ++st->epoch;
        }
        const unsigned grid = nwg > 1 ? (unsigned)(8 * nwg) : 1u;
        panel_kernel<T><<<grid, PT, 0, s>>>(a, lda, c0, rows, w, pivots, st->moves, d_leaf, scratch, st->state, epoch, nwg,
                                             st->status, st->prof_on ? st->prof : nullptr);
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
    pool_put(scratch);
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
