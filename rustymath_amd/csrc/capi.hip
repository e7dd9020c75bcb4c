// The C ABI of hip-sys (declared in include/hip_sys.h): plain pointers and sizes, no torch types.
// Host-pointer entry points are the drop-in for rustymath's `Tiled`/`Factor` seam; `hs_dev_*` entry
// points take device pointers and a stream, for callers that keep data resident in HBM.
#include "hs_kernels.h"
#include "../../include/hip_sys.h"
#include <map>
#include <mutex>
#include <vector>
#include <cstring>
#include <cmath>
#include <algorithm>

namespace hs {

namespace {
thread_local std::string t_error;
thread_local hipStream_t t_stream = nullptr;
std::mutex g_pool_mu;
std::multimap<size_t, void*> g_free;
std::map<void*, size_t> g_size;

size_t size_class(size_t bytes) {
    if (bytes <= (1u << 20)) return (bytes + 255) / 256 * 256;
    size_t c = 1u << 20;
    while (c < bytes) c += c / 4 > (64u << 20) ? (64u << 20) : c / 4;  // ~25% steps, at most 64 MiB apart
    return c;
}
}  // namespace

void set_error(const std::string& what) { t_error = what; }
const char* last_error() { return t_error.c_str(); }

int hip_status(hipError_t e, const char* where) {
    t_error = std::string(where) + ": " + hipGetErrorString(e);
    return (int)e > 0 ? (int)e : 1;
}

hipStream_t thread_stream() {
    if (!t_stream) {
        if (hipStreamCreateWithFlags(&t_stream, hipStreamNonBlocking) != hipSuccess) t_stream = nullptr;
    }
    return t_stream;
}

void* pool_get(size_t bytes) {
    size_t c = size_class(bytes ? bytes : 1);
    {
        std::lock_guard<std::mutex> lk(g_pool_mu);
        auto it = g_free.find(c);
        if (it != g_free.end()) { void* p = it->second; g_free.erase(it); return p; }
    }
    void* p = nullptr;
    if (hipMalloc(&p, c) != hipSuccess) {
        // Give back cached buffers and try once more.
        pool_release_all();
        if (hipMalloc(&p, c) != hipSuccess) return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_pool_mu);
    g_size[p] = c;
    return p;
}

void pool_put(void* p) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(g_pool_mu);
    auto it = g_size.find(p);
    if (it == g_size.end()) return;
    g_free.insert({it->second, p});
}

void pool_release_all() {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    for (auto& kv : g_free) { hipFree(kv.second); g_size.erase(kv.second); }
    g_free.clear();
}

std::mutex& resident_mutex() {
    static std::mutex mu;
    return mu;
}

int device_cus() {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 256;
    return cus;
}

namespace {

// `greatest(a, modulus)` (core/src/linalg/factor.rs:488) over the rows×cols items of a (rows ld apart): NaN ignored,
// 0 for none.
template <typename T>
__global__ void maxabs_kernel(int64_t rows, int64_t cols, int64_t ld, const T* a, unsigned long long* out) {
    T m = T(0);
    const int64_t count = rows * cols;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
        T v = fabs(a[(i / cols) * ld + i % cols]);
        m = v > m ? v : m;
    }
    for (int off = 32; off > 0; off >>= 1) { T o = __shfl_xor(m, off); m = o > m ? o : m; }
    // Magnitudes are non-negative, so their bit patterns order as their values.
    if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__builtin_bit_cast(typename std::conditional<sizeof(T) == 8, unsigned long long, unsigned int>::type, m));
}

// `regular` (factor.rs:221-226): whether every |u_ii| exceeds eps·n·largest; NaN is not negligible.
template <typename T>
__global__ void regular_kernel(int64_t n, const T* f, int64_t ld, const unsigned long long* largest_bits, int* out) {
    typedef typename std::conditional<sizeof(T) == 8, unsigned long long, unsigned int>::type U;
    T largest = __builtin_bit_cast(T, (U)*largest_bits);
    T negligible = (sizeof(T) == 8 ? T(2.220446049250313e-16) : T(1.1920929e-07f)) * T(n) * largest;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        if (fabs(f[i * ld + i]) <= negligible) atomicExch(out, 0);
}

__global__ void fill_int(int* p, int v) { *p = v; }

// The first column whose pivot is exactly zero (the diagonal of the column-major factors), as 1 + its index in *out,
// which starts at n + 1 (LAPACK's `info`: factor.rs:606 leaves such a column undivided and goes on).
template <typename T>
__global__ void first_zero_pivot(int64_t n, const T* __restrict__ f, int64_t ld, unsigned long long* out) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        if (f[i * ld + i] == T(0)) atomicMin(out, (unsigned long long)(i + 1));
}
__global__ void fill_u64(unsigned long long* p, unsigned long long v) { *p = v; }

template <typename T>
int dev_getrf_perm(int64_t n, T* a, int64_t lda, int64_t* d_perm, int64_t* h_perm_out, int64_t* h_piv_out, hipStream_t s,
                   int64_t* info = nullptr) {
    if (info) *info = 0;
    int64_t* d_piv = (int64_t*)pool_get(sizeof(int64_t) * (size_t)((n ? n : 1) + 1));
    if (!d_piv) return hip_status(hipErrorOutOfMemory, "pool_get");
    // The device perm comes from the panels' row moves, composed on the GPU; the host's from the pivots.
    int rc = getrf<T>(n, a, lda, d_piv, s, 0, d_perm);
    if (rc) { pool_put(d_piv); return rc; }
    std::vector<int64_t> piv(n + 1), perm(n);
    unsigned long long* d_zero = (unsigned long long*)(d_piv + n);
    fill_u64<<<1, 1, 0, s>>>(d_zero, (unsigned long long)n + 1);
    first_zero_pivot<T><<<(unsigned)std::min<int64_t>((n + 255) / 256, 1024), 256, 0, s>>>(n, a, lda, d_zero);
    HS_HIP(hipGetLastError());
    HS_HIP(hipMemcpyAsync(piv.data(), d_piv, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToHost, s));
    HS_HIP(hipStreamSynchronize(s));
    pool_put(d_piv);
    if (info) *info = piv[n] > n ? 0 : piv[n];
    for (int64_t i = 0; i < n; ++i) perm[i] = i;
    for (int64_t c = 0; c < n; ++c) std::swap(perm[c], perm[piv[c]]);
    if (h_perm_out) std::memcpy(h_perm_out, perm.data(), sizeof(int64_t) * n);
    if (h_piv_out) std::memcpy(h_piv_out, piv.data(), sizeof(int64_t) * n);
    HS_HIP(hipStreamSynchronize(s));
    return OK;
}

int op_of(int code, bool& t) {
    if (code == HS_NO_TRANS) { t = false; return 0; }
    if (code == HS_TRANS || code == HS_CONJ_TRANS) { t = true; return 0; }
    return 1;
}

template <typename T>
int check_gemm(int opa, int opb, int64_t m, int64_t n, int64_t k, int64_t lda, int64_t ldb, int64_t ldc, bool& ta, bool& tb) {
    if (op_of(opa, ta)) return -1;
    if (op_of(opb, tb)) return -2;
    if (m < 0) return -3;
    if (n < 0) return -4;
    if (k < 0) return -5;
    if (m > 0 && k > 0 && lda < (ta ? m : k)) return -8;
    if (k > 0 && n > 0 && ldb < (tb ? k : n)) return -10;
    if (m > 0 && n > 0 && ldc < n) return -13;
    return 0;
}

// The CBLAS codes of a triangular solve (core/src/linalg/accelerate.rs:67-70): side 141 / 142 (left / right), uplo
// 121 / 122 (upper / lower stored triangle), trans 111 / 112 / 113, diag 131 / 132 (non-unit / unit).
int check_trsm(int side, int uplo, int trans, int diag, int64_t m, int64_t n, int64_t lda, int64_t ldx, bool& right,
               bool& lower, bool& ta, bool& unit) {
    if (side != 141 && side != 142) return -1;
    if (uplo != 121 && uplo != 122) return -2;
    if (op_of(trans, ta)) return -3;
    if (diag != 131 && diag != 132) return -4;
    if (m < 0) return -5;
    if (n < 0) return -6;
    right = side == 142;
    lower = uplo == 122;
    unit = diag == 132;
    const int64_t order = right ? n : m;
    if (m > 0 && n > 0 && lda < order) return -8;
    if (m > 0 && n > 0 && ldx < n) return -10;
    return 0;
}

template <typename T>
int host_trsm(int side, int uplo, int trans, int diag, int64_t m, int64_t n, const T* a, int64_t lda, T* x, int64_t ldx) {
    bool right, lower, ta, unit;
    int e = check_trsm(side, uplo, trans, diag, m, n, lda, ldx, right, lower, ta, unit);
    if (e) return e;
    if (m == 0 || n == 0) return OK;
    const int64_t order = right ? n : m;
    hipStream_t s = thread_stream();
    T* da = (T*)pool_get(sizeof(T) * (size_t)(order * order));
    T* dx = (T*)pool_get(sizeof(T) * (size_t)(m * n));
    int rc = OK;
    if (!da || !dx) rc = hip_status(hipErrorOutOfMemory, "pool_get");
    if (!rc && (hipMemcpy2DAsync(da, order * sizeof(T), a, lda * sizeof(T), order * sizeof(T), order, hipMemcpyHostToDevice, s) != hipSuccess ||
                hipMemcpy2DAsync(dx, n * sizeof(T), x, ldx * sizeof(T), n * sizeof(T), m, hipMemcpyHostToDevice, s) != hipSuccess))
        rc = hip_status(hipErrorUnknown, "H2D");
    // x is compact on the device; whether the caller's x is one contiguous column still picks `single`'s orders.
    if (!rc) rc = trsm<T>(right, lower, ta, unit, m, n, da, order, dx, n, ldx == 1, s);
    if (!rc && hipMemcpy2DAsync(x, ldx * sizeof(T), dx, n * sizeof(T), n * sizeof(T), m, hipMemcpyDeviceToHost, s) != hipSuccess)
        rc = hip_status(hipErrorUnknown, "D2H");
    hipStreamSynchronize(s);
    pool_put(da); pool_put(dx);
    return rc;
}

template <typename T>
int dev_trsm(int side, int uplo, int trans, int diag, int64_t m, int64_t n, const T* a, int64_t lda, T* x, int64_t ldx,
             void* stream) {
    bool right, lower, ta, unit;
    int e = check_trsm(side, uplo, trans, diag, m, n, lda, ldx, right, lower, ta, unit);
    if (e) return e;
    return trsm<T>(right, lower, ta, unit, m, n, a, lda, x, ldx, ldx == 1, (hipStream_t)stream);
}

// The items a row-major r×c matrix with leading dimension ld spans from its first item: what one 1-D copy moves.
int64_t extent(int64_t r, int64_t c, int64_t ld) { return r > 0 && c > 0 ? (r - 1) * ld + c : 0; }

// `batch` products c_i = alpha op(a_i) op(b_i) + beta c_i from host memory, cell i at a + i sa, b + i sb, c + i sc
// (a stride of 0 repeats an operand): each operand's whole span goes to the device in one 1-D copy, leading dimensions
// kept, and c's span comes back the same way.
template <typename T>
int host_gemm_batched(int opa, int opb, int64_t m, int64_t n, int64_t k, T alpha, const T* A, int64_t lda, int64_t sa,
                      const T* B, int64_t ldb, int64_t sb, T beta, T* C, int64_t ldc, int64_t sc, int64_t batch) {
    bool ta, tb;
    int e = check_gemm<T>(opa, opb, m, n, k, lda, ldb, ldc, ta, tb);
    if (e) return e == -10 ? -11 : (e == -13 ? -15 : e);  // ldb and ldc sit after the strides here
    if (batch < 0) return -17;
    if (sa < 0 || sb < 0 || sc < 0) return sa < 0 ? -9 : (sb < 0 ? -12 : -16);
    if (m == 0 || n == 0 || batch == 0) return OK;
    const int64_t ar = ta ? k : m, ac = ta ? m : k, br = tb ? n : k, bc = tb ? k : n;
    const int64_t ea = extent(ar, ac, lda), eb = extent(br, bc, ldb), ec = extent(m, n, ldc);
    if (batch > 1 && sc < ec) { set_error("batched products: cells of c overlap"); return -16; }
    const int64_t na = ea ? (batch - 1) * sa + ea : 0, nb = eb ? (batch - 1) * sb + eb : 0, nc = (batch - 1) * sc + ec;
    hipStream_t s = thread_stream();
    const bool product = k > 0 && alpha != T(0);
    T* dA = product ? (T*)pool_get(sizeof(T) * (size_t)(na ? na : 1)) : nullptr;
    T* dB = product ? (T*)pool_get(sizeof(T) * (size_t)(nb ? nb : 1)) : nullptr;
    T* dC = (T*)pool_get(sizeof(T) * (size_t)nc);
    int rc = OK;
    if (!dC || (product && (!dA || !dB))) rc = hip_status(hipErrorOutOfMemory, "pool_get");
    if (!rc && product && ((na && hipMemcpyAsync(dA, A, sizeof(T) * na, hipMemcpyHostToDevice, s) != hipSuccess) ||
                           (nb && hipMemcpyAsync(dB, B, sizeof(T) * nb, hipMemcpyHostToDevice, s) != hipSuccess)))
        rc = hip_status(hipErrorUnknown, "H2D");
    // c's span goes over where beta is not 0, and where it has items between rows or cells, which must come back as
    // they were.
    const bool gaps = (m > 1 && ldc > n) || (batch > 1 && sc > ec);
    if (!rc && (beta != T(0) || gaps) && hipMemcpyAsync(dC, C, sizeof(T) * nc, hipMemcpyHostToDevice, s) != hipSuccess)
        rc = hip_status(hipErrorUnknown, "H2D");
    if (!rc) rc = gemm_batched<T>(ta, tb, m, n, k, alpha, dA, lda, sa, dB, ldb, sb, beta, dC, ldc, sc, batch, s);
    if (!rc && (hipMemcpyAsync(C, dC, sizeof(T) * nc, hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess))
        rc = hip_status(hipErrorUnknown, "copy back");
    hipStreamSynchronize(s);
    pool_put(dA); pool_put(dB); pool_put(dC);
    return rc;
}

template <typename T>
int host_gemm(int opa, int opb, int64_t m, int64_t n, int64_t k, T alpha, const T* A, int64_t lda, const T* B, int64_t ldb,
              T beta, T* C, int64_t ldc) {
    bool ta, tb;
    const int e = check_gemm<T>(opa, opb, m, n, k, lda, ldb, ldc, ta, tb);
    if (e) return e;
    return host_gemm_batched<T>(opa, opb, m, n, k, alpha, A, lda, 0, B, ldb, 0, beta, C, ldc, 0, 1);
}

template <typename T>
int host_lu_factor(const T* a, int64_t n, T* factors, int64_t* perm, int64_t* pivots, int64_t* info = nullptr) {
    if (n < 0) return -2;
    if (info) *info = 0;
    if (n == 0) return OK;
    hipStream_t s = thread_stream();
    T* d = (T*)pool_get(sizeof(T) * (size_t)(n * n));
    if (!d) return hip_status(hipErrorOutOfMemory, "pool_get");
    int rc = OK;
    if (hipMemcpyAsync(d, a, sizeof(T) * n * n, hipMemcpyHostToDevice, s) != hipSuccess) rc = hip_status(hipErrorUnknown, "H2D");
    if (!rc) rc = dev_getrf_perm<T>(n, d, n, nullptr, perm, pivots, s, info);
    if (!rc && hipMemcpyAsync(factors, d, sizeof(T) * n * n, hipMemcpyDeviceToHost, s) != hipSuccess) rc = hip_status(hipErrorUnknown, "D2H");
    hipStreamSynchronize(s);
    pool_put(d);
    return rc;
}

template <typename T>
int host_lu_solve(const T* factors, const int64_t* perm, int64_t n, const T* b, int64_t k, T* x) {
    if (n < 0) return -3;
    if (k < 0) return -5;
    for (int64_t i = 0; i < n; ++i) if (perm[i] < 0 || perm[i] >= n) return -2;
    if (n == 0 || k == 0) return OK;
    hipStream_t s = thread_stream();
    T* df = (T*)pool_get(sizeof(T) * (size_t)(n * n));
    T* db = (T*)pool_get(sizeof(T) * (size_t)(n * k));
    int64_t* dp = (int64_t*)pool_get(sizeof(int64_t) * (size_t)n);
    int rc = OK;
    if (!df || !db || !dp) rc = hip_status(hipErrorOutOfMemory, "pool_get");
    if (!rc && (hipMemcpyAsync(df, factors, sizeof(T) * n * n, hipMemcpyHostToDevice, s) != hipSuccess ||
                hipMemcpyAsync(db, b, sizeof(T) * n * k, hipMemcpyHostToDevice, s) != hipSuccess ||
                hipMemcpyAsync(dp, perm, sizeof(int64_t) * n, hipMemcpyHostToDevice, s) != hipSuccess))
        rc = hip_status(hipErrorUnknown, "H2D");
    if (!rc) rc = getrs<T>(false, n, k, df, n, dp, db, k, s);
    if (!rc && hipMemcpyAsync(x, db, sizeof(T) * n * k, hipMemcpyDeviceToHost, s) != hipSuccess) rc = hip_status(hipErrorUnknown, "D2H");
    hipStreamSynchronize(s);
    pool_put(df); pool_put(db); pool_put(dp);
    return rc;
}

// Copies the row-major n×n a to the device, factors it in place and decides `regular` (`LuWork::regular`,
// factor.rs:221-226: every pivot exceeds eps·n·max|a|, NaN counting as not negligible).
template <typename T>
int host_factor_regular(const T* a, int64_t n, T* d, int64_t* dp, unsigned long long* dm, int& reg, hipStream_t s) {
    reg = 0;
    if (hipMemcpyAsync(d, a, sizeof(T) * n * n, hipMemcpyHostToDevice, s) != hipSuccess) return hip_status(hipErrorUnknown, "H2D");
    hipMemsetAsync(dm, 0, 64, s);
    maxabs_kernel<T><<<256, 256, 0, s>>>(n, n, n, d, dm);
    int rc = dev_getrf_perm<T>(n, d, n, dp, nullptr, nullptr, s);
    if (rc) return rc;
    int* dreg = (int*)(dm + 1);
    fill_int<<<1, 1, 0, s>>>(dreg, 1);
    regular_kernel<T><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, d, n, dm, dreg);
    if (hipMemcpyAsync(&reg, dreg, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
        return hip_status(hipErrorUnknown, "regular");
    return OK;
}

template <typename T>
int host_solve(const T* a, int64_t n, const T* b, int64_t k, T* x, int* regular) {
    if (n < 0) return -2;
    if (k < 0) return -4;
    if (regular) *regular = 0;
    if (n == 0) return OK;
    hipStream_t s = thread_stream();
    T* d = (T*)pool_get(sizeof(T) * (size_t)(n * n));
    T* db = (T*)pool_get(sizeof(T) * (size_t)(n * (k ? k : 1)));
    int64_t* dp = (int64_t*)pool_get(sizeof(int64_t) * (size_t)n);
    unsigned long long* dm = (unsigned long long*)pool_get(64);
    int rc = OK, reg = 0;
    if (!d || !db || !dp || !dm) rc = hip_status(hipErrorOutOfMemory, "pool_get");
    if (!rc && k && hipMemcpyAsync(db, b, sizeof(T) * n * k, hipMemcpyHostToDevice, s) != hipSuccess)
        rc = hip_status(hipErrorUnknown, "H2D");
    if (!rc) rc = host_factor_regular<T>(a, n, d, dp, dm, reg, s);
    if (!rc && reg && k) rc = getrs<T>(false, n, k, d, n, dp, db, k, s);
    if (!rc && reg && k && hipMemcpyAsync(x, db, sizeof(T) * n * k, hipMemcpyDeviceToHost, s) != hipSuccess) rc = hip_status(hipErrorUnknown, "D2H");
    hipStreamSynchronize(s);
    pool_put(d); pool_put(db); pool_put(dp); pool_put(dm);
    if (regular) *regular = rc ? 0 : reg;
    return rc;
}

// `LuWork::solve` (factor.rs:207-226) on device memory: a (ld lda) factored in place, perm written, `regular` decided
// (every pivot exceeds eps·n·max|a|, NaN not negligible) and, where regular, b (n×k, ld ldb) solved in place.
template <typename T>
int dev_gesv(int64_t n, T* a, int64_t lda, int64_t* perm, T* b, int64_t k, int64_t ldb, int* regular, hipStream_t s) {
    *regular = 0;
    if (n == 0) return OK;
    unsigned long long* dm = (unsigned long long*)pool_get(64);
    if (!dm) return hip_status(hipErrorOutOfMemory, "pool_get");
    int rc = OK, reg = 0;
    fill_u64<<<1, 1, 0, s>>>(dm, 0ull);
    maxabs_kernel<T><<<256, 256, 0, s>>>(n, n, lda, a, dm);
    if (hipGetLastError() != hipSuccess) rc = hip_status(hipErrorUnknown, "maxabs");
    if (!rc) rc = dev_getrf_perm<T>(n, a, lda, perm, nullptr, nullptr, s);
    if (!rc) {
        int* dreg = (int*)(dm + 1);
        fill_int<<<1, 1, 0, s>>>(dreg, 1);
        regular_kernel<T><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(n, a, lda, dm, dreg);
        if (hipMemcpyAsync(&reg, dreg, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            rc = hip_status(hipErrorUnknown, "regular");
    }
    if (!rc && reg && k) rc = getrs<T>(false, n, k, a, lda, perm, b, ldb, s);
    hipStreamSynchronize(s);
    pool_put(dm);
    *regular = rc ? 0 : reg;
    return rc;
}

// The determinant from factors and perm (`Lu::determinant`, factor.rs:174): U's diagonal multiplied in order from 1,
// negated for an odd permutation (`odd`, :186-194, counting cycles at their smallest positions).
template <typename T>
T determinant_of(const T* diag, const int64_t* perm, int64_t n) {
    T d = T(1);
    for (int64_t j = 0; j < n; ++j) d = d * diag[j];
    int64_t cycles = 0;
    for (int64_t st = 0; st < n; ++st) {
        int64_t i = perm[st];
        while (i > st) i = perm[i];
        cycles += i == st;
    }
    return (n - cycles) % 2 ? -d : d;
}

// `LuWork::inverse_into` (factor.rs:215-219): where a is regular, out = a⁻¹ (row-major n×n) and *det its determinant.
template <typename T>
int host_inverse(const T* a, int64_t n, T* out, T* det, int* regular) {
    if (n < 0) return -2;
    if (regular) *regular = 0;
    if (n == 0) return OK;
    hipStream_t s = thread_stream();
    T* d = (T*)pool_get(sizeof(T) * (size_t)(n * n));
    T* dx = (T*)pool_get(sizeof(T) * (size_t)(n * n));
    int64_t* dp = (int64_t*)pool_get(sizeof(int64_t) * (size_t)n);
    unsigned long long* dm = (unsigned long long*)pool_get(64);
    int rc = OK, reg = 0;
    if (!d || !dx || !dp || !dm) rc = hip_status(hipErrorOutOfMemory, "pool_get");
    if (!rc) rc = host_factor_regular<T>(a, n, d, dp, dm, reg, s);
    if (!rc && reg) rc = getri<T>(n, d, n, dp, dx, n, s);
    if (!rc && reg) {
        std::vector<T> diag((size_t)n);
        std::vector<int64_t> perm((size_t)n);
        if (hipMemcpyAsync(out, dx, sizeof(T) * n * n, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipMemcpy2DAsync(diag.data(), sizeof(T), d, sizeof(T) * (n + 1), sizeof(T), n, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipMemcpyAsync(perm.data(), dp, sizeof(int64_t) * n, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            rc = hip_status(hipErrorUnknown, "D2H");
        else if (det)
            *det = determinant_of<T>(diag.data(), perm.data(), n);
    }
    hipStreamSynchronize(s);
    pool_put(d); pool_put(dx); pool_put(dp); pool_put(dm);
    if (regular) *regular = rc ? 0 : reg;
    return rc;
}

// `Lu::inverse` (factor.rs:150-154) from host factors (column-major n×n) and perm.
template <typename T>
int host_lu_inverse(const T* factors, const int64_t* perm, int64_t n, T* out) {
    if (n < 0) return -3;
    for (int64_t i = 0; i < n; ++i) if (perm[i] < 0 || perm[i] >= n) return -2;
    if (n == 0) return OK;
    hipStream_t s = thread_stream();
    T* df = (T*)pool_get(sizeof(T) * (size_t)(n * n));
    T* dx = (T*)pool_get(sizeof(T) * (size_t)(n * n));
    int64_t* dp = (int64_t*)pool_get(sizeof(int64_t) * (size_t)n);
    int rc = OK;
    if (!df || !dx || !dp) rc = hip_status(hipErrorOutOfMemory, "pool_get");
    if (!rc && (hipMemcpyAsync(df, factors, sizeof(T) * n * n, hipMemcpyHostToDevice, s) != hipSuccess ||
                hipMemcpyAsync(dp, perm, sizeof(int64_t) * n, hipMemcpyHostToDevice, s) != hipSuccess))
        rc = hip_status(hipErrorUnknown, "H2D");
    if (!rc) rc = getri<T>(n, df, n, dp, dx, n, s);
    if (!rc && hipMemcpyAsync(out, dx, sizeof(T) * n * n, hipMemcpyDeviceToHost, s) != hipSuccess) rc = hip_status(hipErrorUnknown, "D2H");
    hipStreamSynchronize(s);
    pool_put(df); pool_put(dx); pool_put(dp);
    return rc;
}

// `cholesky` (factor.rs:267, `cholesky_factor` :777-797) from a host matrix: l = L (row-major, zeros above) and
// *ok = 1, or *ok = 0 and l untouched where a is not positive definite.
template <typename T>
int host_cholesky(const T* a, int64_t n, T* l, int* ok) {
    if (n < 0) return -2;
    if (ok) *ok = 0;
    int okv = 1;
    if (n == 0) { if (ok) *ok = 1; return OK; }
    hipStream_t s = thread_stream();
    T* d = (T*)pool_get(sizeof(T) * (size_t)(n * n));
    int rc = OK;
    if (!d) rc = hip_status(hipErrorOutOfMemory, "pool_get");
    if (!rc && hipMemcpyAsync(d, a, sizeof(T) * n * n, hipMemcpyHostToDevice, s) != hipSuccess) rc = hip_status(hipErrorUnknown, "H2D");
    if (!rc) rc = potrf<T>(n, d, n, &okv, s);
    if (!rc && okv && hipMemcpyAsync(l, d, sizeof(T) * n * n, hipMemcpyDeviceToHost, s) != hipSuccess) rc = hip_status(hipErrorUnknown, "D2H");
    hipStreamSynchronize(s);
    if (d) pool_put(d);
    if (ok) *ok = rc ? 0 : okv;
    return rc;
}

template <typename T>
__global__ void batched_to_reference(int64_t count, int64_t n, const T* __restrict__ a, const int64_t* __restrict__ perm,
                                     T* f) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= count * n * n) return;
    const int64_t cell = idx / (n * n), rem = idx % (n * n), c = rem / n, r = rem % n;
    f[idx] = a[cell * n * n + perm[cell * n + r] * n + c];  // column c, row r of the cell's LU
}

// A stack of `count` cells of order n (and k right-hand sides): `Stacked::solve` (src/linalg/stacked.rs:88-101) when
// x is given, `Stacked::lu` (:174-187) when factors are. Orders up to 512 with at most 8 right-hand sides go to the
// batched kernel in one launch (each cell as `LuWork::solve` or `lu_factor`, bit for bit); larger cells go one by
// one through the single-matrix path.
template <typename T>
int host_stack(int64_t count, int64_t n, int64_t k, const T* a, const T* b, T* x, int* regular, T* factors,
               int64_t* perm) {
    if (count < 0) return -1;
    if (n < 0) return -2;
    if (k < 0) return -3;
    const bool solve = x != nullptr;
    if (count == 0) return OK;
    if (n == 0) {  // the reference's `regular` declines an empty matrix (factor.rs:222)
        if (regular) for (int64_t i = 0; i < count; ++i) regular[i] = 0;
        return OK;
    }
    if (n > 512 || k > 8) {
        for (int64_t i = 0; i < count; ++i) {
            int reg = 0;
            const int rc = solve ? host_solve<T>(a + i * n * n, n, b + i * n * k, k, x + i * n * k, &reg)
                                 : host_lu_factor<T>(a + i * n * n, n, factors + i * n * n, perm + i * n, nullptr);
            if (rc) return rc;
            if (solve && !reg && k) std::memcpy(x + i * n * k, b + i * n * k, sizeof(T) * (size_t)(n * k));
            if (solve && regular) regular[i] = reg;
        }
        return OK;
    }
    const int64_t kk = solve ? k : 0;
    hipStream_t s = thread_stream();
    T* da = (T*)pool_get(sizeof(T) * (size_t)(count * n * n));
    T* db = (T*)pool_get(sizeof(T) * (size_t)std::max<int64_t>(count * n * kk, 1));
    T* df = solve ? nullptr : (T*)pool_get(sizeof(T) * (size_t)(count * n * n));
    int64_t* dp = (int64_t*)pool_get(sizeof(int64_t) * (size_t)(count * n));
    int* dr = (int*)pool_get(sizeof(int) * (size_t)count);
    int rc = OK;
    if (!da || !db || (!solve && !df) || !dp || !dr) rc = hip_status(hipErrorOutOfMemory, "pool_get");
    if (!rc && (hipMemcpyAsync(da, a, sizeof(T) * count * n * n, hipMemcpyHostToDevice, s) != hipSuccess ||
                (kk && hipMemcpyAsync(db, b, sizeof(T) * count * n * kk, hipMemcpyHostToDevice, s) != hipSuccess)))
        rc = hip_status(hipErrorUnknown, "H2D");
    if (!rc) {
        std::vector<int64_t> ns((size_t)count, n), ks((size_t)count, kk), ao((size_t)count), bo((size_t)count), po((size_t)count);
        for (int64_t i = 0; i < count; ++i) { ao[i] = i * n * n; bo[i] = i * n * kk; po[i] = i * n; }
        long long h = 0;
        rc = batched_plan(count, ns.data(), ks.data(), ao.data(), bo.data(), po.data(), sizeof(T) == 8, da, db, dp, dr, 0, 32, 1,
                          &h);
        if (!rc) rc = batched_run(h, false, s);
        if (h) { const int rf = batched_free(h); if (!rc) rc = rf; }
    }
    if (!rc && solve) {
        if ((kk && hipMemcpyAsync(x, db, sizeof(T) * count * n * kk, hipMemcpyDeviceToHost, s) != hipSuccess) ||
            (regular && hipMemcpyAsync(regular, dr, sizeof(int) * count, hipMemcpyDeviceToHost, s) != hipSuccess))
            rc = hip_status(hipErrorUnknown, "D2H");
    }
    if (!rc && !solve) {
        batched_to_reference<T><<<(unsigned)((count * n * n + 255) / 256), 256, 0, s>>>(count, n, da, dp, df);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(factors, df, sizeof(T) * count * n * n, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipMemcpyAsync(perm, dp, sizeof(int64_t) * count * n, hipMemcpyDeviceToHost, s) != hipSuccess)
            rc = hip_status(hipErrorUnknown, "D2H");
    }
    hipStreamSynchronize(s);
    pool_put(da); pool_put(db); pool_put(dp); pool_put(dr);
    if (df) pool_put(df);
    return rc;
}

// Uneven batched solves from host memory: system i is a[i] (row-major n[i]×n[i]) with b[i] (row-major n[i]×k[i]), each
// `LuWork::solve` (core/src/linalg/factor.rs:207-226) as rustymath runs the cells of a stack (core/src/parallel.rs:118-192).
// Systems up to order 512 with at most 8 right-hand sides go to one persistent batched launch (packed back to back into
// two device buffers), others one by one. x[i] receives X_i where regular[i] = 1, b[i]'s copy where the reference would
// take least squares (regular[i] = 0; an empty system counts as such, factor.rs:222).
template <typename T>
int host_batched(int64_t count, const int64_t* n, const int64_t* k, const T* const* a, const T* const* b, T* const* x,
                 int* regular) {
    if (count < 0) return -1;
    if (count == 0) return OK;
    if (!n) return -2;
    for (int64_t i = 0; i < count; ++i) if (n[i] < 0) return -2;
    if (!k) return -3;
    for (int64_t i = 0; i < count; ++i) if (k[i] < 0) return -3;
    if (!a) return -4;
    for (int64_t i = 0; i < count; ++i) if (n[i] > 0 && !a[i]) return -4;
    if (!b) return -5;
    for (int64_t i = 0; i < count; ++i) if (n[i] > 0 && k[i] > 0 && !b[i]) return -5;
    if (!x) return -6;
    for (int64_t i = 0; i < count; ++i) if (n[i] > 0 && k[i] > 0 && !x[i]) return -6;
    std::vector<int64_t> idx, ns, ks, ao, bo, po;
    int64_t ta = 0, tb = 0, tp = 0;
    for (int64_t i = 0; i < count; ++i) {
        if (n[i] == 0) { if (regular) regular[i] = 0; continue; }
        if (n[i] > 512 || k[i] > 8) {  // beyond the batched kernel: this system on its own
            int reg = 0;
            const int rc = host_solve<T>(a[i], n[i], b[i], k[i], x[i], &reg);
            if (rc) return rc;
            if (!reg && k[i]) std::memcpy(x[i], b[i], sizeof(T) * (size_t)(n[i] * k[i]));
            if (regular) regular[i] = reg;
            continue;
        }
        idx.push_back(i); ns.push_back(n[i]); ks.push_back(k[i]);
        ao.push_back(ta); bo.push_back(tb); po.push_back(tp);
        ta += n[i] * n[i]; tb += n[i] * k[i]; tp += n[i];
    }
    const int64_t m = (int64_t)idx.size();
    if (m == 0) return OK;
    std::vector<T> ha((size_t)ta), hb((size_t)std::max<int64_t>(tb, 1));
    for (int64_t j = 0; j < m; ++j) {
        const int64_t i = idx[j];
        std::memcpy(ha.data() + ao[j], a[i], sizeof(T) * (size_t)(n[i] * n[i]));
        if (k[i]) std::memcpy(hb.data() + bo[j], b[i], sizeof(T) * (size_t)(n[i] * k[i]));
    }
    hipStream_t s = thread_stream();
    T* da = (T*)pool_get(sizeof(T) * (size_t)ta);
    T* db = (T*)pool_get(sizeof(T) * (size_t)std::max<int64_t>(tb, 1));
    int64_t* dp = (int64_t*)pool_get(sizeof(int64_t) * (size_t)tp);
    int* dr = (int*)pool_get(sizeof(int) * (size_t)m);
    int rc = OK;
    if (!da || !db || !dp || !dr) rc = hip_status(hipErrorOutOfMemory, "pool_get");
    if (!rc && (hipMemcpyAsync(da, ha.data(), sizeof(T) * ta, hipMemcpyHostToDevice, s) != hipSuccess ||
                (tb && hipMemcpyAsync(db, hb.data(), sizeof(T) * tb, hipMemcpyHostToDevice, s) != hipSuccess)))
        rc = hip_status(hipErrorUnknown, "H2D");
    if (!rc) {
        long long h = 0;
        rc = batched_plan(m, ns.data(), ks.data(), ao.data(), bo.data(), po.data(), sizeof(T) == 8, da, db, dp, dr, 0, 32, 2,
                          &h);
        if (!rc) rc = batched_run(h, false, s);
        if (h) { const int rf = batched_free(h); if (!rc) rc = rf; }
    }
    std::vector<int> hr((size_t)m);
    if (!rc && ((tb && hipMemcpyAsync(hb.data(), db, sizeof(T) * tb, hipMemcpyDeviceToHost, s) != hipSuccess) ||
                hipMemcpyAsync(hr.data(), dr, sizeof(int) * m, hipMemcpyDeviceToHost, s) != hipSuccess))
        rc = hip_status(hipErrorUnknown, "D2H");
    if (hipStreamSynchronize(s) != hipSuccess && !rc) rc = hip_status(hipErrorUnknown, "sync");
    pool_put(da); pool_put(db); pool_put(dp); pool_put(dr);
    if (rc) return rc;
    for (int64_t j = 0; j < m; ++j) {  // b is left on the device as it was where the system is not regular
        const int64_t i = idx[j];
        if (k[i]) std::memcpy(x[i], hb.data() + bo[j], sizeof(T) * (size_t)(n[i] * k[i]));
        if (regular) regular[i] = hr[j];
    }
    return OK;
}

// Row r of a cell's reference-layout factors (column-major, "row j holds column j") into the batched kernel's layout
// (row-major, row perm[r] holding position r): the inverse of batched_to_reference.
template <typename T>
__global__ void reference_to_batched(int64_t count, int64_t n, const T* __restrict__ f, const int64_t* __restrict__ perm,
                                     T* a) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= count * n * n) return;
    const int64_t cell = idx / (n * n), rem = idx % (n * n), c = rem / n, r = rem % n;
    a[cell * n * n + perm[cell * n + r] * n + c] = f[idx];
}

// `Stacked::lu_solve` (src/linalg/stacked.rs:189-204): x_i = lu_solve(factors_i, perm_i, b_i) for `count` cells of order
// n with k right-hand sides. Orders up to 512 with at most 8 right-hand sides, where every perm_i is a permutation, run
// as one batched launch (the batched kernel's lu_solve mode, bit for bit each cell's lu_solve); others cell by cell.
template <typename T>
int host_lu_solve_stack(int64_t count, int64_t n, int64_t k, const T* factors, const int64_t* perm, const T* b, T* x) {
    if (count < 0) return -1;
    if (n < 0) return -2;
    if (k < 0) return -3;
    for (int64_t i = 0; i < count * n; ++i) if (perm[i] < 0 || perm[i] >= n) return -5;
    if (count == 0 || n == 0 || k == 0) return OK;
    bool perms = n <= 512 && k <= 8;
    for (int64_t c = 0; c < count && perms; ++c) {
        std::vector<char> seen((size_t)n, 0);
        for (int64_t i = 0; i < n && perms; ++i) { perms = !seen[perm[c * n + i]]; seen[perm[c * n + i]] = 1; }
    }
    if (!perms) {
        for (int64_t c = 0; c < count; ++c) {
            const int rc = host_lu_solve<T>(factors + c * n * n, perm + c * n, n, b + c * n * k, k, x + c * n * k);
            if (rc) return rc;
        }
        return OK;
    }
    hipStream_t s = thread_stream();
    T* df = (T*)pool_get(sizeof(T) * (size_t)(count * n * n));
    T* da = (T*)pool_get(sizeof(T) * (size_t)(count * n * n));
    T* db = (T*)pool_get(sizeof(T) * (size_t)(count * n * k));
    int64_t* dp = (int64_t*)pool_get(sizeof(int64_t) * (size_t)(count * n));
    int rc = OK;
    if (!df || !da || !db || !dp) rc = hip_status(hipErrorOutOfMemory, "pool_get");
    if (!rc && (hipMemcpyAsync(df, factors, sizeof(T) * count * n * n, hipMemcpyHostToDevice, s) != hipSuccess ||
                hipMemcpyAsync(db, b, sizeof(T) * count * n * k, hipMemcpyHostToDevice, s) != hipSuccess ||
                hipMemcpyAsync(dp, perm, sizeof(int64_t) * count * n, hipMemcpyHostToDevice, s) != hipSuccess))
        rc = hip_status(hipErrorUnknown, "H2D");
    if (!rc) {
        reference_to_batched<T><<<(unsigned)((count * n * n + 255) / 256), 256, 0, s>>>(count, n, df, dp, da);
        if (hipGetLastError() != hipSuccess) rc = hip_status(hipErrorUnknown, "reference_to_batched");
    }
    if (!rc) {
        std::vector<int64_t> ns((size_t)count, n), ks((size_t)count, k), ao((size_t)count), bo((size_t)count), po((size_t)count);
        for (int64_t i = 0; i < count; ++i) { ao[i] = i * n * n; bo[i] = i * n * k; po[i] = i * n; }
        long long h = 0;
        rc = batched_plan(count, ns.data(), ks.data(), ao.data(), bo.data(), po.data(), sizeof(T) == 8, da, db, dp, nullptr,
                          1, 32, 1, &h);
        if (!rc) rc = batched_run(h, false, s);
        if (h) { const int rf = batched_free(h); if (!rc) rc = rf; }
    }
    if (!rc && hipMemcpyAsync(x, db, sizeof(T) * count * n * k, hipMemcpyDeviceToHost, s) != hipSuccess)
        rc = hip_status(hipErrorUnknown, "D2H");
    hipStreamSynchronize(s);
    pool_put(df); pool_put(da); pool_put(db); pool_put(dp);
    return rc;
}

template <typename T>
int dev_gemm_batched(int opa, int opb, int64_t m, int64_t n, int64_t k, T alpha, const T* A, int64_t lda, int64_t sa,
                     const T* B, int64_t ldb, int64_t sb, T beta, T* C, int64_t ldc, int64_t sc, int64_t batch,
                     void* stream) {
    bool ta, tb;
    int e = check_gemm<T>(opa, opb, m, n, k, lda, ldb, ldc, ta, tb);
    if (e) return e == -10 ? -11 : (e == -13 ? -15 : e);  // ldb and ldc sit after the strides here
    if (batch < 0) return -17;
    if (sa < 0 || sb < 0 || sc < 0) return sa < 0 ? -9 : (sb < 0 ? -12 : -16);
    if (batch > 1 && m > 0 && n > 0 && sc < extent(m, n, ldc)) { set_error("batched products: cells of c overlap"); return -16; }
    return gemm_batched<T>(ta, tb, m, n, k, alpha, A, lda, sa, B, ldb, sb, beta, C, ldc, sc, batch, (hipStream_t)stream);
}

template <typename T>
int dev_gemm(int opa, int opb, int64_t m, int64_t n, int64_t k, T alpha, const T* A, int64_t lda, const T* B, int64_t ldb,
             T beta, T* C, int64_t ldc, void* stream) {
    bool ta, tb;
    int e = check_gemm<T>(opa, opb, m, n, k, lda, ldb, ldc, ta, tb);
    if (e) return e;
    hipStream_t s = (hipStream_t)stream;
    return gemm<T>(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s);
}

}  // namespace
namespace {
__device__ int g_xcc[256];  // the XCD each of 256 launch indices ran on, written by the probe
__global__ void xcc_probe_kernel() {
    if (threadIdx.x == 0) {
        unsigned x;
        asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID, 0, 4)" : "=s"(x));
        g_xcc[blockIdx.x] = (int)x;
    }
}
std::mutex g_xmu;
int g_xforce = 0;              // 1: device scope always
signed char g_xlocal[64];      // per device: 0 unknown, 1 one L2, 2 device scope
}  // namespace

bool exchange_local(bool may_probe) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return false;
    std::lock_guard<std::mutex> lk(g_xmu);
    if (g_xforce) return false;
    if (g_xlocal[dev] == 0 && may_probe) {
        // Launch index i on the XCD of index i mod 8 (round-robin dispatch; which XCD index 0 takes varies from
        // launch to launch): the panel's workgroups (indices 0, 8, 16, ...) share one L2, and so do the
        // substitutions' runs of blocks.
        int xcc[256];
        bool ok = false;
        xcc_probe_kernel<<<256, 64, 0, 0>>>();
        if (hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess &&
            hipMemcpyFromSymbol(xcc, HIP_SYMBOL(g_xcc), sizeof(xcc)) == hipSuccess) {
            ok = true;
            for (int i = 8; i < 256; ++i) ok = ok && xcc[i] == xcc[i % 8];
        }
        g_xlocal[dev] = ok ? 1 : 2;
    }
    return g_xlocal[dev] == 1;
}

}  // namespace hs

using namespace hs;

extern "C" {

int hs_set_exchange(int mode, int* local) {
    if (mode < 0 || mode > 1) { set_error("exchange mode must be 0 (probe) or 1 (device scope)"); return -1; }
    {
        std::lock_guard<std::mutex> lk(hs::g_xmu);
        hs::g_xforce = mode;
    }
    const bool l = hs::exchange_local(true);
    if (local) *local = l ? 1 : 0;
    return OK;
}

int hs_init(void) {
    int n = 0;
    HS_HIP(hipGetDeviceCount(&n));
    if (n <= 0) { set_error("no HIP device"); return 1; }
    if (!thread_stream()) return hip_status(hipErrorUnknown, "hipStreamCreate");
    hs::exchange_local(true);
    return OK;
}

int hs_finalize(void) {
    hipDeviceSynchronize();
    pool_release_all();
    return OK;
}

const char* hs_last_error(void) { return last_error(); }

int hs_dgemm(int opa, int opb, int64_t m, int64_t n, int64_t k, double alpha, const double* a, int64_t lda, const double* b,
             int64_t ldb, double beta, double* c, int64_t ldc) {
    return host_gemm<double>(opa, opb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}
int hs_sgemm(int opa, int opb, int64_t m, int64_t n, int64_t k, float alpha, const float* a, int64_t lda, const float* b,
             int64_t ldb, float beta, float* c, int64_t ldc) {
    return host_gemm<float>(opa, opb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}
int hs_dgemm_batched(int opa, int opb, int64_t m, int64_t n, int64_t k, double alpha, const double* a, int64_t lda,
                     int64_t sa, const double* b, int64_t ldb, int64_t sb, double beta, double* c, int64_t ldc, int64_t sc,
                     int64_t batch) {
    return host_gemm_batched<double>(opa, opb, m, n, k, alpha, a, lda, sa, b, ldb, sb, beta, c, ldc, sc, batch);
}
int hs_sgemm_batched(int opa, int opb, int64_t m, int64_t n, int64_t k, float alpha, const float* a, int64_t lda,
                     int64_t sa, const float* b, int64_t ldb, int64_t sb, float beta, float* c, int64_t ldc, int64_t sc,
                     int64_t batch) {
    return host_gemm_batched<float>(opa, opb, m, n, k, alpha, a, lda, sa, b, ldb, sb, beta, c, ldc, sc, batch);
}
int hs_dev_dgemm_batched(int opa, int opb, int64_t m, int64_t n, int64_t k, double alpha, const double* a, int64_t lda,
                         int64_t sa, const double* b, int64_t ldb, int64_t sb, double beta, double* c, int64_t ldc,
                         int64_t sc, int64_t batch, void* stream) {
    return dev_gemm_batched<double>(opa, opb, m, n, k, alpha, a, lda, sa, b, ldb, sb, beta, c, ldc, sc, batch, stream);
}
int hs_dev_sgemm_batched(int opa, int opb, int64_t m, int64_t n, int64_t k, float alpha, const float* a, int64_t lda,
                         int64_t sa, const float* b, int64_t ldb, int64_t sb, float beta, float* c, int64_t ldc,
                         int64_t sc, int64_t batch, void* stream) {
    return dev_gemm_batched<float>(opa, opb, m, n, k, alpha, a, lda, sa, b, ldb, sb, beta, c, ldc, sc, batch, stream);
}
int hs_dtrsm(int side, int uplo, int trans, int diag, int64_t m, int64_t n, const double* a, int64_t lda, double* x,
             int64_t ldx) {
    return host_trsm<double>(side, uplo, trans, diag, m, n, a, lda, x, ldx);
}
int hs_strsm(int side, int uplo, int trans, int diag, int64_t m, int64_t n, const float* a, int64_t lda, float* x,
             int64_t ldx) {
    return host_trsm<float>(side, uplo, trans, diag, m, n, a, lda, x, ldx);
}
int hs_dev_dtrsm(int side, int uplo, int trans, int diag, int64_t m, int64_t n, const double* a, int64_t lda, double* x,
                 int64_t ldx, void* stream) {
    return dev_trsm<double>(side, uplo, trans, diag, m, n, a, lda, x, ldx, stream);
}
int hs_dev_strsm(int side, int uplo, int trans, int diag, int64_t m, int64_t n, const float* a, int64_t lda, float* x,
                 int64_t ldx, void* stream) {
    return dev_trsm<float>(side, uplo, trans, diag, m, n, a, lda, x, ldx, stream);
}
int hs_dlu_factor(const double* a, int64_t n, double* factors, int64_t* perm, int64_t* pivots, int64_t* info) {
    return host_lu_factor<double>(a, n, factors, perm, pivots, info);
}
int hs_slu_factor(const float* a, int64_t n, float* factors, int64_t* perm, int64_t* pivots, int64_t* info) {
    return host_lu_factor<float>(a, n, factors, perm, pivots, info);
}
int hs_dlu_solve(const double* factors, const int64_t* perm, int64_t n, const double* b, int64_t k, double* x) {
    return host_lu_solve<double>(factors, perm, n, b, k, x);
}
int hs_slu_solve(const float* factors, const int64_t* perm, int64_t n, const float* b, int64_t k, float* x) {
    return host_lu_solve<float>(factors, perm, n, b, k, x);
}
int hs_dlu_inverse(const double* factors, const int64_t* perm, int64_t n, double* out) {
    return host_lu_inverse<double>(factors, perm, n, out);
}
int hs_slu_inverse(const float* factors, const int64_t* perm, int64_t n, float* out) {
    return host_lu_inverse<float>(factors, perm, n, out);
}
int hs_dinverse(const double* a, int64_t n, double* out, double* det, int* regular) {
    return host_inverse<double>(a, n, out, det, regular);
}
int hs_sinverse(const float* a, int64_t n, float* out, float* det, int* regular) {
    return host_inverse<float>(a, n, out, det, regular);
}
int hs_dcholesky(const double* a, int64_t n, double* l, int* ok) { return host_cholesky<double>(a, n, l, ok); }
int hs_scholesky(const float* a, int64_t n, float* l, int* ok) { return host_cholesky<float>(a, n, l, ok); }
int hs_dev_dpotrf(int64_t n, double* a, int64_t lda, int* ok, void* stream) {
    if (n < 0) return -1;
    if (lda < n) return -3;
    if (!ok) return -4;
    return potrf<double>(n, a, lda, ok, (hipStream_t)stream);
}
int hs_dev_spotrf(int64_t n, float* a, int64_t lda, int* ok, void* stream) {
    if (n < 0) return -1;
    if (lda < n) return -3;
    if (!ok) return -4;
    return potrf<float>(n, a, lda, ok, (hipStream_t)stream);
}
int hs_dsolve_stack(int64_t count, int64_t n, int64_t k, const double* a, const double* b, double* x, int* regular) {
    return host_stack<double>(count, n, k, a, b, x, regular, nullptr, nullptr);
}
int hs_ssolve_stack(int64_t count, int64_t n, int64_t k, const float* a, const float* b, float* x, int* regular) {
    return host_stack<float>(count, n, k, a, b, x, regular, nullptr, nullptr);
}
int hs_dgesv_batched(int64_t count, const int64_t* n, const int64_t* k, const double* const* a, const double* const* b,
                     double* const* x, int* regular) {
    return host_batched<double>(count, n, k, a, b, x, regular);
}
int hs_sgesv_batched(int64_t count, const int64_t* n, const int64_t* k, const float* const* a, const float* const* b,
                     float* const* x, int* regular) {
    return host_batched<float>(count, n, k, a, b, x, regular);
}
int hs_dlu_stack(int64_t count, int64_t n, const double* a, double* factors, int64_t* perm) {
    return host_stack<double>(count, n, 0, a, nullptr, nullptr, nullptr, factors, perm);
}
int hs_slu_stack(int64_t count, int64_t n, const float* a, float* factors, int64_t* perm) {
    return host_stack<float>(count, n, 0, a, nullptr, nullptr, nullptr, factors, perm);
}
int hs_dlu_solve_stack(int64_t count, int64_t n, int64_t k, const double* factors, const int64_t* perm, const double* b,
                       double* x) {
    return host_lu_solve_stack<double>(count, n, k, factors, perm, b, x);
}
int hs_slu_solve_stack(int64_t count, int64_t n, int64_t k, const float* factors, const int64_t* perm, const float* b,
                       float* x) {
    return host_lu_solve_stack<float>(count, n, k, factors, perm, b, x);
}
int hs_dsolve(const double* a, int64_t n, const double* b, int64_t k, double* x, int* regular) {
    return host_solve<double>(a, n, b, k, x, regular);
}
int hs_ssolve(const float* a, int64_t n, const float* b, int64_t k, float* x, int* regular) {
    return host_solve<float>(a, n, b, k, x, regular);
}

int hs_dev_dgemm(int opa, int opb, int64_t m, int64_t n, int64_t k, double alpha, const double* a, int64_t lda, const double* b,
                 int64_t ldb, double beta, double* c, int64_t ldc, void* stream) {
    return dev_gemm<double>(opa, opb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, stream);
}
int hs_dev_sgemm(int opa, int opb, int64_t m, int64_t n, int64_t k, float alpha, const float* a, int64_t lda, const float* b,
                 int64_t ldb, float beta, float* c, int64_t ldc, void* stream) {
    return dev_gemm<float>(opa, opb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, stream);
}
int hs_dev_dgetrf(int64_t n, double* a, int64_t lda, int64_t* perm, int64_t* info, void* stream) {
    if (n < 0) return -1;
    if (lda < n) return -3;
    hipStream_t s = (hipStream_t)stream;
    return dev_getrf_perm<double>(n, a, lda, perm, nullptr, nullptr, s, info);
}
int hs_dev_sgetrf(int64_t n, float* a, int64_t lda, int64_t* perm, int64_t* info, void* stream) {
    if (n < 0) return -1;
    if (lda < n) return -3;
    hipStream_t s = (hipStream_t)stream;
    return dev_getrf_perm<float>(n, a, lda, perm, nullptr, nullptr, s, info);
}
int hs_dev_dgesv(int64_t n, double* a, int64_t lda, int64_t* perm, double* b, int64_t k, int64_t ldb, int* regular,
                 void* stream) {
    if (n < 0) return -1;
    if (n > 0 && lda < n) return -3;
    if (k < 0) return -6;
    if (n > 0 && k > 0 && ldb < k) return -7;
    if (!regular) return -8;
    return dev_gesv<double>(n, a, lda, perm, b, k, ldb, regular, (hipStream_t)stream);
}
int hs_dev_sgesv(int64_t n, float* a, int64_t lda, int64_t* perm, float* b, int64_t k, int64_t ldb, int* regular,
                 void* stream) {
    if (n < 0) return -1;
    if (n > 0 && lda < n) return -3;
    if (k < 0) return -6;
    if (n > 0 && k > 0 && ldb < k) return -7;
    if (!regular) return -8;
    return dev_gesv<float>(n, a, lda, perm, b, k, ldb, regular, (hipStream_t)stream);
}
int hs_dev_dgetrs(int trans, int64_t n, int64_t k, const double* factors, int64_t ldf, const int64_t* perm, double* b, int64_t ldb,
                  void* stream) {
    if (n < 0) return -2;
    if (k < 0) return -3;
    if (n > 0 && ldf < n) return -5;
    if (n > 0 && k > 0 && ldb < k) return -8;
    hipStream_t s = (hipStream_t)stream;
    return getrs<double>(trans != 0, n, k, factors, ldf, perm, b, ldb, s);
}
int hs_dev_sgetrs(int trans, int64_t n, int64_t k, const float* factors, int64_t ldf, const int64_t* perm, float* b, int64_t ldb,
                  void* stream) {
    if (n < 0) return -2;
    if (k < 0) return -3;
    if (n > 0 && ldf < n) return -5;
    if (n > 0 && k > 0 && ldb < k) return -8;
    hipStream_t s = (hipStream_t)stream;
    return getrs<float>(trans != 0, n, k, factors, ldf, perm, b, ldb, s);
}
int hs_dev_dgetrs_noalloc(int trans, int64_t n, int64_t k, const double* factors, int64_t ldf, const int64_t* perm,
                          double* b, int64_t ldb, void* stream) {
    if (n < 0) return -2;
    if (k < 0) return -3;
    if (n > 0 && ldf < n) return -5;
    if (n > 0 && k > 0 && ldb < k) return -8;
    return getrs_noalloc(trans != 0, n, k, factors, ldf, perm, b, ldb, (hipStream_t)stream);
}
int hs_dev_dgetri(int64_t n, const double* factors, int64_t ldf, const int64_t* perm, double* out, int64_t ldo,
                  void* stream) {
    if (n < 0) return -1;
    if (ldf < n) return -3;
    if (ldo < n) return -6;
    return getri<double>(n, factors, ldf, perm, out, ldo, (hipStream_t)stream);
}
int hs_dev_sgetri(int64_t n, const float* factors, int64_t ldf, const int64_t* perm, float* out, int64_t ldo,
                  void* stream) {
    if (n < 0) return -1;
    if (ldf < n) return -3;
    if (ldo < n) return -6;
    return getri<float>(n, factors, ldf, perm, out, ldo, (hipStream_t)stream);
}

int hs_dev_dsgesv(int64_t n, const double* a, int64_t lda, const double* b, double* x, int max_iter, double tol, int* iters,
                  double* berr, void* stream) {
    if (n < 0) return -1;
    if (lda != n) return -3;
    if (max_iter < 1) return -6;
    return gesv_ir(n, a, lda, b, x, max_iter, tol, iters, berr, (hipStream_t)stream);
}

int hs_dev_sgemm_bf16_update(int64_t m, int64_t n, int64_t k, const float* a, int64_t lda, const float* b, int64_t ldb,
                             float* c, int64_t ldc, int terms, void* stream) {
    if (m < 0) return -1;
    if (n < 0) return -2;
    if (k < 0) return -3;
    if (lda < (k > 1 ? k : 1)) return -5;
    if (ldb < (n > 1 ? n : 1)) return -7;
    if (ldc < (n > 1 ? n : 1)) return -9;
    if (terms != 1 && terms != 3) return -10;
    return gemm_bf16_update(m, n, k, a, lda, b, ldb, c, ldc, (hipStream_t)stream, terms);
}

int hs_dev_dbgesv(int64_t n, const double* a, int64_t lda, const double* b, double* x, int max_iter, int restart,
                  int terms, double tol, int* iters, int* inner, double* berr, void* stream) {
    if (n < 0) return -1;
    if (lda != n) return -3;
    if (max_iter < 1) return -6;
    if (restart < 1 || restart > 200) return -7;
    if (terms != 1 && terms != 3) return -8;
    return gesv_gmres_ir(n, a, lda, b, x, max_iter, restart, tol, iters, inner, berr, (hipStream_t)stream, terms);
}

int hs_dev_dgesv_batched_plan(int64_t count, const int64_t* n, const int64_t* k, const int64_t* a_off, const int64_t* b_off,
                              const int64_t* p_off, double* a, double* b, int64_t* perm, int* regular, int mode, int panel_width,
                              int wgs_per_cu, int64_t* plan) {
    if (count < 0) return -1;
    if (count > 0 && (!n || !k || !a_off || !b_off || !p_off)) return -2;
    if (!plan) return -14;
    long long h = 0;
    const int rc = batched_plan(count, n, k, a_off, b_off, p_off, true, a, b, perm, regular, mode, panel_width, wgs_per_cu,
                                &h);
    *plan = h;
    return rc;
}
int hs_dev_sgesv_batched_plan(int64_t count, const int64_t* n, const int64_t* k, const int64_t* a_off, const int64_t* b_off,
                              const int64_t* p_off, float* a, float* b, int64_t* perm, int* regular, int mode, int panel_width,
                              int wgs_per_cu, int64_t* plan) {
    if (count < 0) return -1;
    if (count > 0 && (!n || !k || !a_off || !b_off || !p_off)) return -2;
    if (!plan) return -14;
    long long h = 0;
    const int rc = batched_plan(count, n, k, a_off, b_off, p_off, false, a, b, perm, regular, mode, panel_width, wgs_per_cu,
                                &h);
    *plan = h;
    return rc;
}
int hs_dev_dsgesv_batched_plan(int64_t count, const int64_t* n, const int64_t* k, const int64_t* a_off,
                               const int64_t* b_off, const int64_t* p_off, const double* a, const double* b, double* x,
                               float* a32, float* r32, int64_t* perm, int* regular, int* iters, double* berr,
                               int max_iter, double tol_per_n, int panel_width, int wgs_per_cu, int64_t* plan) {
    if (count < 0) return -1;
    if (count > 0 && (!n || !k || !a_off || !b_off || !p_off)) return -2;
    if (!plan) return -20;
    long long h = 0;
    const int rc = batched_ir_plan(count, n, k, a_off, b_off, p_off, a, b, x, a32, r32, perm, regular, iters, berr, max_iter,
                                   tol_per_n, panel_width, wgs_per_cu, &h);
    *plan = h;
    return rc;
}
int hs_dev_gesv_batched_run(int64_t plan, int graph, void* stream) {
    return batched_run(plan, graph != 0, (hipStream_t)stream);
}
int hs_dev_gesv_batched_free(int64_t plan) { return batched_free(plan); }

int hs_profile_panel(int on, int64_t* out) { return panel_profile(on, (long long*)out); }
int hs_profile_batched(int on, int64_t* out) { return batched_profile(on, (long long*)out); }
int hs_profile_gemm(int on, double* ms, int64_t* launches, double* flops, double* bytes) {
    return gemm_profile(on, ms, launches, flops, bytes);
}
int hs_profile_lu(int on, double* out) { return lu_profile(on, out); }

}  // extern "C"
