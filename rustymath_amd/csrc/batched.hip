// Batched solves of many independent square systems of different orders (n <= 512) in one persistent launch.
//
// rustymath solves a stack cell by cell, one `LuWork` per thread (core/src/parallel.rs:118-192,
// src/linalg/stacked.rs:88-101): each cell is `LuWork::solve` (core/src/linalg/factor.rs:207-226), the
// `regular` check around `lu_factor` and `lu_solve`. Here each workgroup takes the next system from a queue
// sorted by decreasing cost (longest processing time first) and runs the same arithmetic, item for item:
//
//   - the factorization is right-looking over panels made of whole reference leaves (lu_panel halves down
//     to 16 columns), at most 32 columns each. A panel's rows sit one per thread in registers, shifted so
//     v[0] is the current column; rows never move: a swap relabels positions (`perm[position] = row`), as in
//     the large-matrix panel kernel. Every item takes its updates in pivot-column order, one fused
//     multiply-add each; multipliers are a · (1/pivot), a quotient for a subnormal pivot; a zero pivot
//     skips the rest of its own leaf;
//   - right of a panel, each thread owns a column: it solves the panel's unit-lower triangle for that
//     column's U rows in registers, then updates the trailing rows with the panel's multipliers, which the
//     workgroup stages in LDS a chunk of rows at a time;
//   - the solve substitutes in the reference's three orders (csrc/solve.hip has the same three).
//
// A is overwritten by its factors without moving rows: row perm[i] of the matrix holds position i of the
// factors, where perm is rustymath's `perm` (factor.rs:132-134). B is overwritten by X where the system is
// regular. The result of a system does not depend on which workgroup ran it or when.
#include "hs_kernels.h"
#include <vector>
#include <algorithm>
#include <numeric>
#include <mutex>
#include <map>

namespace hs {

namespace {

constexpr int BT = 512;    // threads of a workgroup = the largest order
constexpr int LR = 64;     // trailing rows whose multipliers are staged in LDS at once
constexpr int MAXRHS = 8;  // right-hand sides per system
constexpr int NS = 32;     // `backward`'s leaf height (factor.rs:693-707)

struct Desc {
    long long n, k, a_off, b_off, p_off;
};

template <typename T>
__device__ __forceinline__ void wave_best(T& v, int& r) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        T v2 = __shfl_xor(v, off);
        int r2 = __shfl_xor(r, off);
        if (v2 > v || (v2 == v && r2 < r)) { v = v2; r = r2; }
    }
}

template <typename T>
struct Bits { typedef typename std::conditional<sizeof(T) == 8, unsigned long long, unsigned int>::type U; };

// BW: the widest panel (16 or 32 columns, a tuning choice; the results do not depend on it).
template <typename T, int BW>
__global__ __launch_bounds__(BT) void gesv_batched_kernel(int count, const Desc* __restrict__ desc, const int* __restrict__ order,
                                                          T* __restrict__ a_all, T* __restrict__ b_all,
                                                          long long* __restrict__ perm_all, int* __restrict__ regular_all,
                                                          int* next) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    constexpr int WAVES = BT / 64;
    __shared__ int s_sys;
    __shared__ unsigned long long s_max;
    __shared__ short perm[BT];       // position -> row
    __shared__ short leaf_end[BT];   // end of the reference leaf of each column
    __shared__ short panel_end[BT];  // for a panel's first column, the panel's end
    __shared__ T s_piv[BW];
    __shared__ T s_val[WAVES];
    __shared__ int s_key[WAVES];
    __shared__ T s_L[BW][BW + 1];    // the panel's unit-lower triangle
    __shared__ T s_buf[BT * MAXRHS]; // trailing multipliers [LR][BW], or the solve's x [BT][MAXRHS]
    T (*s_l2)[BW] = reinterpret_cast<T (*)[BW]>(s_buf);
    T (*s_x)[MAXRHS] = reinterpret_cast<T (*)[MAXRHS]>(s_buf);
    static_assert(LR * BW <= BT * MAXRHS && BW >= 16, "LDS carve-up; panels hold whole 16-column leaves");
    const T tiny = sizeof(T) == 8 ? T(2.2250738585072014e-308) : T(1.17549435e-38f);
    const T eps = sizeof(T) == 8 ? T(2.220446049250313e-16) : T(1.1920929e-07f);

    for (;;) {
        if (tid == 0) { s_sys = atomicAdd(next, 1); s_max = 0ull; }
        __syncthreads();
        const int q = s_sys;
        if (q >= count) break;
        const int sys = order[q];
        const Desc d = desc[sys];
        const int n = (int)d.n, k = (int)d.k;
        T* a = a_all + d.a_off;
        T* b = b_all + d.b_off;

        if (tid < n) {
            perm[tid] = (short)tid;
            // The reference leaf holding column tid: lu_panel's halving while wider than 16 (factor.rs:592-616).
            int l0 = 0, l1 = n;
            while (l1 - l0 > 16) { const int mid = (l0 + l1) / 2; if (tid < mid) l1 = mid; else l0 = mid; }
            leaf_end[tid] = (short)l1;
        }
        {
            // `greatest` magnitude of the matrix, NaN ignored (factor.rs:488), for `regular`.
            T m = T(0);
            for (int i = tid; i < n * n; i += BT) { const T v = fabs(a[i]); m = v > m ? v : m; }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) { const T o = __shfl_xor(m, off); m = o > m ? o : m; }
            if (lane == 0) atomicMax(&s_max, (unsigned long long)__builtin_bit_cast(typename Bits<T>::U, m));
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

        // ---- lu_factor ----
        for (int c0 = 0; c0 < n;) {
            const int c1 = panel_end[c0], w = c1 - c0, rows = n - c0;
            const bool has = tid < rows;
            const int row = has ? perm[c0 + tid] : 0;
            int pos = tid;
            T v[BW];
#pragma unroll
            for (int j = 0; j < BW; ++j) v[j] = (has && j < w) ? a[row * n + c0 + j] : T(0);
#pragma clang loop unroll(disable)
            for (int c = 0; c < w; ++c) {
                // First greatest magnitude at positions >= c, NaN ignored; all NaN -> position c (factor.rs:476-490).
                T val = T(-1);
                int key = 0x7fffffff;
                if (has && pos >= c) { const T x = fabs(v[0]); val = x != x ? T(-1) : x; key = pos; }
                wave_best(val, key);
                if (lane == 0) { s_val[wave] = val; s_key[wave] = key; }
                __syncthreads();
                T wb = s_val[0];
                int p = s_key[0];
#pragma unroll
                for (int u = 1; u < WAVES; ++u)
                    if (s_val[u] > wb || (s_val[u] == wb && s_key[u] < p)) { wb = s_val[u]; p = s_key[u]; }
                if (has && pos == p) {
#pragma unroll
                    for (int j = 0; j < BW; ++j) s_piv[j] = v[j];
                }
                __syncthreads();
                const T pivot = s_piv[0];
                if (has && pos == p) {
#pragma unroll
                    for (int j = 0; j < BW; ++j) if (j < w - c) a[row * n + c0 + c + j] = v[j];
                    pos = c;
                } else if (has && pos == c) {
                    pos = p;
                }
                if (has && pos > c) {
                    // divide (factor.rs:506-508), nothing for a zero pivot, whose leaf then skips its update (:606).
                    const bool zero = pivot == T(0);
                    T l = v[0];
                    if (!zero) l = fabs(pivot) >= tiny ? l * (T(1) / pivot) : l / pivot;
                    a[row * n + c0 + c] = l;
                    const int skip = zero ? leaf_end[c0 + c] - c0 - c : 1;
#pragma unroll
                    for (int j = 1; j < BW; ++j) {
                        const T nv = __builtin_fma(l, -s_piv[j], v[j]);
                        v[j - 1] = j >= skip ? nv : v[j];
                    }
                    v[BW - 1] = T(0);
                }
            }
            if (has) perm[c0 + pos] = (short)row;
            __syncthreads();
            if (c1 < n) {
                for (int i = tid; i < BW * BW; i += BT) {
                    const int r = i / BW, cq = i % BW;
                    s_L[r][cq] = (r < w && cq < w) ? a[perm[c0 + r] * n + c0 + cq] : T(0);
                }
                // Threads split into CG column groups of 64 and G row groups.
                const int cols = n - c1, CG = (cols + 63) / 64, G = WAVES / CG;
                const int grp = wave / CG, col = c1 + (wave % CG) * 64 + lane;
                const bool live = grp < G && col < n;
                // U12 for this column: the panel's unit-lower solve, in order (every group of the column alike,
                // all reading before the barrier, group 0 writing after it).
                T u[BW];
#pragma unroll
                for (int r = 0; r < BW; ++r) u[r] = (live && r < w) ? a[perm[c0 + r] * n + col] : T(0);
                __syncthreads();
#pragma unroll
                for (int qq = 0; qq < BW; ++qq)
#pragma unroll
                    for (int r = qq + 1; r < BW; ++r) u[r] = __builtin_fma(-s_L[r][qq], u[qq], u[r]);
                if (live && grp == 0) {
#pragma unroll
                    for (int r = 0; r < BW; ++r) if (r < w) a[perm[c0 + r] * n + col] = u[r];
                }
                // Trailing rows [c1, n): each item takes the panel's w terms in order, one fused multiply-add each.
                for (int r0 = c1; r0 < n; r0 += LR) {
                    const int nr = n - r0 < LR ? n - r0 : LR;
                    __syncthreads();  // the previous chunk's multipliers are no longer read
                    for (int i = tid; i < LR * BW; i += BT) {
                        const int rr = i / BW, p = i % BW;
                        s_l2[rr][p] = (rr < nr && p < w) ? -a[perm[r0 + rr] * n + c0 + p] : T(0);
                    }
                    __syncthreads();
                    if (live) {
                        for (int rr = grp; rr < nr; rr += G) {
                            T* item = a + perm[r0 + rr] * n + col;
                            T acc = *item;
#pragma unroll
                            for (int p = 0; p < BW; ++p)
                                if (p < w) acc = __builtin_fma(s_l2[rr][p], u[p], acc);
                            *item = acc;
                        }
                    }
                }
                __syncthreads();
            }
            c0 = c1;
        }

        // ---- regular (factor.rs:221-226): no |u_ii| at or below eps · n · greatest ----
        const T largest = __builtin_bit_cast(T, (typename Bits<T>::U)s_max);
        const T negligible = eps * T(n) * largest;
        const int ok = __syncthreads_and(!(tid < n) || !(fabs(a[perm[tid] * n + tid]) <= negligible));
        if (tid < n) perm_all[d.p_off + tid] = perm[tid];
        if (tid == 0) regular_all[sys] = ok && n > 0;
        if (!ok || n == 0 || k == 0) { __syncthreads(); continue; }

        // ---- lu_solve (factor.rs:633-641): x = rows perm of b, then L, then U ----
        const bool mine = tid < n;
        const T* frow = a + (mine ? perm[tid] : 0) * n;  // this thread's factor row (position tid)
        T acc[MAXRHS];
#pragma unroll
        for (int c = 0; c < MAXRHS; ++c) acc[c] = (mine && c < k) ? b[perm[tid] * k + c] : T(0);
        // L: increasing order, unit diagonal.
        for (int p = 0; p < n; ++p) {
            if (tid == p) {
#pragma unroll
                for (int c = 0; c < MAXRHS; ++c) s_x[p][c] = acc[c];
            }
            __syncthreads();
            if (mine && tid > p) {
                const T lp = -frow[p];
#pragma unroll
                for (int c = 0; c < MAXRHS; ++c) if (c < k) acc[c] = __builtin_fma(s_x[p][c], lp, acc[c]);
            }
        }
        __syncthreads();
        if (k == 1) {
            // One column: decreasing order, then a quotient (`single`, factor.rs:725-748).
            for (int p = n - 1; p >= 0; --p) {
                if (tid == p) { acc[0] = acc[0] / frow[p]; s_x[p][0] = acc[0]; }
                __syncthreads();
                if (mine && tid < p) acc[0] = __builtin_fma(s_x[p][0], -frow[p], acc[0]);
            }
        } else {
            // Several columns: `backward`'s halving above NS rows (factor.rs:693-707). Leaves run bottom to top;
            // once the leaf starting at a split's middle is done, the split's top half adds its bottom half.
            for (int end = n; end > 0;) {
                int r0 = 0, r1 = n;
                while (r1 - r0 > NS) { const int mid = (r0 + r1) / 2; if (end <= mid) r1 = mid; else r0 = mid; }
                for (int t = r1 - 1; t >= r0; --t) {
                    if (tid == t) {
                        // The rows after it in the leaf, in increasing order, then a reciprocal multiply.
                        for (int j = t + 1; j < r1; ++j) {
                            const T uu = -frow[j];
#pragma unroll
                            for (int c = 0; c < MAXRHS; ++c) if (c < k) acc[c] = __builtin_fma(s_x[j][c], uu, acc[c]);
                        }
                        const T dd = frow[t];
                        const bool nrm = fabs(dd) >= tiny;
                        const T rcp = T(1) / dd;
#pragma unroll
                        for (int c = 0; c < MAXRHS; ++c) {
                            if (c < k) acc[c] = nrm ? acc[c] * rcp : acc[c] / dd;
                            s_x[t][c] = acc[c];
                        }
                    }
                    __syncthreads();
                }
                int a0 = 0, a1 = n;
                while (a1 - a0 > NS) {
                    const int mid = (a0 + a1) / 2;
                    if (mid == r0) {
                        if (tid >= a0 && tid < mid) {
                            for (int j = mid; j < a1; ++j) {
                                const T uu = -frow[j];
#pragma unroll
                                for (int c = 0; c < MAXRHS; ++c) if (c < k) acc[c] = __builtin_fma(s_x[j][c], uu, acc[c]);
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
        if (mine) {
#pragma unroll
            for (int c = 0; c < MAXRHS; ++c) if (c < k) b[tid * k + c] = acc[c];
        }
        __syncthreads();
    }
}

__global__ void zero_int(int* p) { *p = 0; }

// A plan: the systems' descriptors and their order on the device, and the launch captured in a hipGraph.
struct Plan {
    int count = 0;
    bool f64 = true;
    Desc* d_desc = nullptr;
    int* d_order = nullptr;
    int* d_next = nullptr;
    int grid = 0;
    int bw = 32;
    void *a = nullptr, *b = nullptr;
    long long* perm = nullptr;
    int* regular = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};
std::mutex g_plans_mu;
std::map<long long, Plan*> g_plans;
long long g_next_plan = 1;

template <typename T>
int enqueue(Plan* p, hipStream_t s) {
    zero_int<<<1, 1, 0, s>>>(p->d_next);
    if (p->bw == 16)
        gesv_batched_kernel<T, 16><<<p->grid, BT, 0, s>>>(p->count, p->d_desc, p->d_order, (T*)p->a, (T*)p->b, p->perm,
                                                         p->regular, p->d_next);
    else
        gesv_batched_kernel<T, 32><<<p->grid, BT, 0, s>>>(p->count, p->d_desc, p->d_order, (T*)p->a, (T*)p->b, p->perm,
                                                         p->regular, p->d_next);
    HS_HIP(hipGetLastError());
    return OK;
}

Plan* find_plan(long long handle) {
    std::lock_guard<std::mutex> lk(g_plans_mu);
    auto it = g_plans.find(handle);
    return it == g_plans.end() ? nullptr : it->second;
}

}  // namespace

int batched_plan(int64_t count, const int64_t* n, const int64_t* k, const int64_t* a_off, const int64_t* b_off,
                 const int64_t* p_off, bool f64, void* a, void* b, int64_t* perm, int* regular, int panel_width,
                 int wgs_per_cu, long long* handle) {
    if (panel_width != 16 && panel_width != 32) { set_error("batched panel width must be 16 or 32"); return -11; }
    if (wgs_per_cu < 1 || wgs_per_cu > 2) { set_error("batched workgroups per CU must be 1 or 2"); return -12; }
    std::vector<Desc> desc(count);
    std::vector<int> order(count);
    for (int64_t i = 0; i < count; ++i) {
        if (n[i] < 0 || n[i] > BT) { set_error("batched systems need 0 <= n <= 512"); return -2; }
        if (k[i] < 0 || k[i] > MAXRHS) { set_error("batched systems need 0 <= k <= 8"); return -3; }
        if (a_off[i] < 0 || b_off[i] < 0 || p_off[i] < 0) { set_error("batched offsets must be >= 0"); return -4; }
        desc[i] = Desc{n[i], k[i], a_off[i], b_off[i], p_off[i]};
    }
    // Longest processing time first: the factorization's n³ and the solve's n² k.
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        const double cx = (double)n[x] * n[x] * (n[x] + 3.0 * k[x]), cy = (double)n[y] * n[y] * (n[y] + 3.0 * k[y]);
        return cx > cy;
    });
    Plan* p = new Plan;
    p->count = (int)count; p->f64 = f64; p->a = a; p->b = b; p->perm = (long long*)perm; p->regular = regular;
    p->grid = device_cus() * wgs_per_cu;
    p->bw = panel_width;
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

int batched_run(long long handle, bool use_graph, hipStream_t s) {
    Plan* p = find_plan(handle);
    if (!p) { set_error("unknown batched plan"); return -1; }
    if (!use_graph) return p->f64 ? enqueue<double>(p, s) : enqueue<float>(p, s);
    if (!p->exec) {
        // The queue reset and the persistent launch, captured once; every run replays the graph.
        hipStream_t cap;
        HS_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        HS_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
        const int rc = p->f64 ? enqueue<double>(p, cap) : enqueue<float>(p, cap);
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
    delete p;
    return OK;
}

}  // namespace hs
