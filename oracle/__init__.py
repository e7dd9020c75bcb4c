"""The oracle: a CPU restatement of rustymath's dense linear-algebra hot path (oracle.c), loaded through ctypes.

TEST INFRASTRUCTURE ONLY. `tests/`, `__graft_entry__.smoke()` and `bench.py`'s CPU-baseline leg use it as the
checker. The package `rustymath_amd` never imports it, and no product path falls back to it.

Every function takes and gives NumPy arrays in rustymath's layouts: row-major matrices, and LU factors
column-major ("row j holds column j", src/linalg/stacked.rs:25) with `perm[i]` the source row of row i
(core/src/linalg/factor.rs:132-134).
"""
import ctypes, os, subprocess
import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, 'liboracle.so')
_lib = None

OPS = {'N': 111, 'T': 112, 'C': 113}


def build():
    "Compile liboracle.so with the committed Makefile."
    subprocess.run(['make', '-s', '-C', _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB): build()
        _lib = ctypes.CDLL(_LIB)
        i64, i32, vp = ctypes.c_int64, ctypes.c_int, ctypes.c_void_p
        for s, t in (('d', ctypes.c_double), ('s', ctypes.c_float)):
            for name in ('or_gemm_', 'or_gemm_naive_'):
                f = getattr(_lib, name + s)
                f.argtypes = [i32, i32, i64, i64, i64, t, vp, i64, vp, i64, t, vp, i64]
                f.restype = i32
            for name in ('or_lu_', 'or_lu_flat_'):
                f = getattr(_lib, name + s)
                f.argtypes = [vp, i64, vp, vp, vp]
                f.restype = i32
            f = getattr(_lib, 'or_lu_solve_' + s); f.argtypes = [vp, vp, i64, vp, i64, vp]; f.restype = i32
            f = getattr(_lib, 'or_trsm_' + s); f.argtypes = [i32, i32, i32, i32, i64, i64, vp, i64, vp, i64]; f.restype = i32
            f = getattr(_lib, 'or_lu_solve_t_' + s); f.argtypes = [vp, vp, i64, vp, i64, vp]; f.restype = i32
            f = getattr(_lib, 'or_lu_inverse_' + s); f.argtypes = [vp, vp, i64, vp]; f.restype = i32
            f = getattr(_lib, 'or_det_' + s); f.argtypes = [vp, vp, i64]; f.restype = t
            f = getattr(_lib, 'or_slogdet_' + s); f.argtypes = [vp, vp, i64, vp, vp]; f.restype = None
            f = getattr(_lib, 'or_solve_' + s); f.argtypes = [vp, vp, i64, i64, vp]; f.restype = i32
            for name in ('or_cholesky_', 'or_cholesky_flat_'):
                f = getattr(_lib, name + s); f.argtypes = [vp, i64, vp]; f.restype = i32
        _lib.or_set_threads.argtypes = [i32]
        _lib.or_get_threads.restype = i32
    return _lib


def _suffix(dtype):
    dtype = np.dtype(dtype)
    if dtype == np.float64: return 'd'
    if dtype == np.float32: return 's'
    raise TypeError(f'the oracle computes at float64 and float32, not {dtype}')


def _p(x): return x.ctypes.data_as(ctypes.c_void_p)


def set_threads(n):
    "Threads for the products' row bands (the reference's `set_threads`); 0 means every core."
    lib().or_set_threads(int(n))


def get_threads():
    return lib().or_get_threads()


def gemm(a, b, alpha=1.0, beta=0.0, c=None, opa='N', opb='N', naive=False):
    """`Tiled::gemm` (core/src/linalg.rs:799): c = alpha op(a) op(b) + beta c, row-major, in place on `c`
    (a new zero matrix where `c` is None). `naive` runs the defining in-order fold instead of the blocked kernels."""
    s = _suffix(a.dtype)
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b, dtype=a.dtype)
    m, k = (a.shape[1], a.shape[0]) if opa != 'N' else a.shape
    kb, n = (b.shape[1], b.shape[0]) if opb != 'N' else b.shape
    assert k == kb, 'inner extents differ'
    if c is None: c = np.zeros((m, n), a.dtype)
    assert c.flags.c_contiguous and c.dtype == a.dtype and c.shape == (m, n)
    f = getattr(lib(), ('or_gemm_naive_' if naive else 'or_gemm_') + s)
    st = f(OPS[opa], OPS[opb], m, n, k, alpha, _p(a), max(a.shape[1], 1), _p(b), max(b.shape[1], 1), beta, _p(c), max(n, 1))
    if st: raise ValueError(f'gemm rejected argument {-st}')
    return c


def matmul(a, b):
    "`PlusTimes.inner` of two row-major matrices (or a matrix and a vector), as rustymath's `matmul`."
    a = np.asarray(a); b = np.asarray(b, dtype=a.dtype)
    av, bv = a.ndim == 1, b.ndim == 1
    a2 = a[None, :] if av else a
    b2 = b[:, None] if bv else b
    c = gemm(a2, b2)
    if av and bv: return c[0, 0]
    if av: return c[0]
    if bv: return c[:, 0]
    return c


def lu(a, flat=False):
    """`lu` (core/src/linalg/factor.rs:261): (factors column-major n×n, perm, pivots). `flat` runs the
    one-column-at-a-time statement of the same factorization."""
    a = np.ascontiguousarray(a)
    s, n = _suffix(a.dtype), a.shape[0]
    assert a.shape == (n, n)
    f = np.zeros((n, n), a.dtype); perm = np.zeros(n, np.int64); piv = np.zeros(n, np.int64)
    getattr(lib(), ('or_lu_flat_' if flat else 'or_lu_') + s)(_p(a), n, _p(f), _p(perm), _p(piv))
    return f, perm, piv


def lu_solve(factors, perm, b):
    "`lu_solve` (factor.rs:633): b is n or n×k; the result has b's shape."
    factors = np.ascontiguousarray(factors); s, n = _suffix(factors.dtype), factors.shape[0]
    b = np.asarray(b, dtype=factors.dtype); vec = b.ndim == 1
    b2 = np.ascontiguousarray(b.reshape(n, 1 if vec else b.shape[1])); k = b2.shape[1]
    perm = np.ascontiguousarray(perm, dtype=np.int64)
    x = np.zeros((n, k), factors.dtype)
    st = getattr(lib(), 'or_lu_solve_' + s)(_p(factors), _p(perm), n, _p(b2), k, _p(x))
    if st: raise IndexError('row positions out of range')
    return x[:, 0] if vec else x


SIDES = {'L': 141, 'R': 142}
UPLOS = {'U': 121, 'L': 122}
DIAGS = {'N': 131, 'U': 132}


def trsm(a, x, side='L', uplo='L', trans='N', diag='N', ldx=None):
    """`Tiled::triangular` (core/src/linalg.rs:781) as `triangular_native` (factor.rs:711-721) computes it: solves
    op(a) X = B (side 'L') or X op(a) = B ('R') for the row-major B held in `x` (m×n), reading only a's triangle `uplo`
    ('U' or 'L', with a unit diagonal for diag 'U'); trans 'N', 'T' or 'C'. `ldx` (>= n) lays x out with that leading
    dimension, as a strided view would hold it (one column with ldx > 1 takes `forward`/`backward` rather than
    `single`). Returns the solution, a new m×n array."""
    a = np.ascontiguousarray(a)
    s = _suffix(a.dtype)
    x = np.asarray(x, dtype=a.dtype)
    m, n = x.shape
    ld = n if ldx is None else ldx
    buf = np.zeros(max((m - 1) * ld + n, 1) if m and n else 1, a.dtype)
    for i in range(m): buf[i * ld:i * ld + n] = x[i]
    st = getattr(lib(), 'or_trsm_' + s)(SIDES[side], UPLOS[uplo], OPS[trans], DIAGS[diag], m, n, _p(a),
                                        max(a.shape[1], 1), _p(buf), max(ld, 1))
    if st: raise ValueError(f'trsm rejected argument {-st}')
    return np.array([buf[i * ld:i * ld + n] for i in range(m)], a.dtype).reshape(m, n)


def lu_solve_t(factors, perm, b):
    """Aᵀ x = b with the factors of A (the adjoint's solve, no counterpart in the reference): Uᵀ in increasing order with
    `divide`'s diagonal scaling, Lᵀ in decreasing order, then the inverse row permutation."""
    factors = np.ascontiguousarray(factors); s, n = _suffix(factors.dtype), factors.shape[0]
    b = np.asarray(b, dtype=factors.dtype); vec = b.ndim == 1
    b2 = np.ascontiguousarray(b.reshape(n, 1 if vec else b.shape[1])); k = b2.shape[1]
    perm = np.ascontiguousarray(perm, dtype=np.int64)
    x = np.zeros((n, k), factors.dtype)
    st = getattr(lib(), 'or_lu_solve_t_' + s)(_p(factors), _p(perm), n, _p(b2), k, _p(x))
    if st: raise IndexError('row positions out of range')
    return x[:, 0] if vec else x


def lu_inverse(factors, perm):
    "`lu_inverse` (factor.rs:645), rustymath's own kernels."
    factors = np.ascontiguousarray(factors); s, n = _suffix(factors.dtype), factors.shape[0]
    out = np.zeros((n, n), factors.dtype)
    getattr(lib(), 'or_lu_inverse_' + s)(_p(factors), _p(np.ascontiguousarray(perm, dtype=np.int64)), n, _p(out))
    return out


def cholesky(a, flat=False):
    """`cholesky_factor` (factor.rs:777-797): L (row-major, zeros above the diagonal) with L Lᵀ = a from a's lower
    triangle, or None where a is not positive definite. `flat` gives the left-looking statement of the same order."""
    a = np.ascontiguousarray(a); s, n = _suffix(a.dtype), a.shape[0]
    out = np.zeros((n, n), a.dtype)
    ok = getattr(lib(), ('or_cholesky_flat_' if flat else 'or_cholesky_') + s)(_p(a), n, _p(out))
    return out if ok == 1 else None


def det(factors, perm):
    factors = np.ascontiguousarray(factors)
    return getattr(lib(), 'or_det_' + _suffix(factors.dtype))(_p(factors), _p(np.ascontiguousarray(perm, dtype=np.int64)), factors.shape[0])


def slogdet(factors, perm):
    factors = np.ascontiguousarray(factors)
    sign, log = np.zeros(1, factors.dtype), np.zeros(1, factors.dtype)
    getattr(lib(), 'or_slogdet_' + _suffix(factors.dtype))(_p(factors), _p(np.ascontiguousarray(perm, dtype=np.int64)), factors.shape[0], _p(sign), _p(log))
    return sign[0], log[0]


def solve(a, b):
    """`LuWork::solve` for a square `a` (factor.rs:207-226): the LU solution, or None where a pivot is
    negligible and the reference would take least squares."""
    a = np.ascontiguousarray(a); s, n = _suffix(a.dtype), a.shape[0]
    b = np.asarray(b, dtype=a.dtype); vec = b.ndim == 1
    b2 = np.ascontiguousarray(b.reshape(n, 1 if vec else b.shape[1])); k = b2.shape[1]
    x = np.zeros((n, k), a.dtype)
    ok = getattr(lib(), 'or_solve_' + s)(_p(a), _p(b2), n, k, _p(x))
    if not ok: return None
    return x[:, 0] if vec else x


def lower(factors):
    "`Lu::lower` (factor.rs:158): L with a unit diagonal, row-major."
    f = np.asarray(factors).T
    return np.tril(f, -1) + np.eye(f.shape[0], dtype=f.dtype)


def upper(factors):
    "`Lu::upper` (factor.rs:166): U, row-major."
    return np.triu(np.asarray(factors).T)
