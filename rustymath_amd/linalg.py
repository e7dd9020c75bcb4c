"""rustymath's dense linear-algebra hot path on MI355X, under the names of python/rustymath/linalg.py:
`matmul`, `solve`, `lu`, the `Lu` factors with `solve`, `determinant` and `log_determinant`, and `lu_solve`.

Arguments are NumPy-convertible. Functions take stacks of matrices in the last two axes; only a one-dimensional
right-hand side is a vector. Frames pair from the front, as rustymath broadcasts (src/linalg/stacked.rs:1,
tests/test_linalg.py:111-112), where NumPy pairs them from the back. float32 and float64 compute at their own
width; float16 computes in float32 with results rounded to float16; booleans, uint8 and int16 compute as float32
and wider integers as float64 (src/linalg/stacked.rs:78-85). Every product, factorization and solve runs on the
GPU through libhip_sys.so; a square matrix with a negligible pivot, where rustymath falls back to least squares
(core/src/linalg/factor.rs:208), raises `NotImplementedError` (least squares is outside this library).
"""
from typing import NamedTuple
import ctypes
import numpy as np

from . import _lib

__all__ = ['matmul', 'solve', 'lu', 'Lu', 'lu_solve', 'lu_det', 'lu_slogdet', 'lu_lower', 'lu_upper']


def _ptr(x): return x.ctypes.data_as(ctypes.c_void_p)


def _compute_dtype(*arrays):
    "The dtypes (compute, store) of rustymath's ladder for these arguments (src/linalg/stacked.rs:78-85)."
    common = np.result_type(*arrays)
    if np.issubdtype(common, np.complexfloating):
        raise TypeError('complex matrices are outside this library (rustymath computes them on the CPU)')
    if common == np.float64: return np.float64, np.float64
    if common == np.float32: return np.float32, np.float32
    if common == np.float16: return np.float32, np.float16
    if common in (np.bool_, np.uint8, np.int8, np.int16): return np.float32, np.float32
    if np.issubdtype(common, np.integer): return np.float64, np.float64
    raise TypeError(f'linalg does not take {common}')


def _pair_frames(fa, fb):
    "Pairs two frames from the front: the shorter gains trailing unit axes, then they broadcast."
    la, lb = len(fa), len(fb)
    fa2 = tuple(fa) + (1,) * (max(la, lb) - la)
    fb2 = tuple(fb) + (1,) * (max(la, lb) - lb)
    return fa2, fb2, np.broadcast_shapes(fa2, fb2)


def _cells(x, frame, core):
    "x with its frame broadcast to `frame`, as a C-contiguous array of cells [count, *core]."
    x = np.broadcast_to(x, tuple(frame) + tuple(core))
    return np.ascontiguousarray(x.reshape((int(np.prod(frame, dtype=np.int64)),) + tuple(core)))


def _sfx(dt): return 'd' if dt == np.float64 else 's'


def _operand(x, fx, frame, count, rows, cols):
    """(array, op code, leading dimension, cell stride) through which the product reads the [*fx, rows, cols] array x
    over `frame`, as `Mapped::matrix` does (core/src/linalg.rs:658-670): in place where every cell is positive-strided
    with one unit stride (row by row, or column by column as a transpose) and the cells lie a fixed stride apart, one
    matrix broadcast over every cell with stride 0; any other layout as a compact copy of its cells."""
    item = x.itemsize
    cells = None
    if all(f == 1 for f in fx):
        cells, stride = x.reshape(rows, cols), 0
    elif tuple(fx) == tuple(frame):
        try:
            cells = x.view(); cells.shape = (count, rows, cols)  # a view, or AttributeError where a copy would be needed
            stride = cells.strides[0] // item if count > 1 else 0
            if cells.strides[0] % item or stride < 0: cells = None
        except AttributeError:
            cells = None
    if cells is not None:
        s0, s1 = cells.strides[-2], cells.strides[-1]
        if s0 % item == 0 and s1 % item == 0:
            s0, s1 = s0 // item, s1 // item
            if (cols == 1 or s1 == 1) and (rows == 1 or s0 >= cols) and s0 >= 0:
                return cells, _lib.NO_TRANS, (s0 if rows > 1 else max(cols, 1)), stride
            if (rows == 1 or s0 == 1) and (cols == 1 or s1 >= rows) and s1 >= 0:
                return cells, _lib.TRANS, (s1 if cols > 1 else max(rows, 1)), stride
    dense = _cells(x, frame, (rows, cols))
    return dense, _lib.NO_TRANS, max(cols, 1), rows * cols


def matmul(x1, x2):
    """The matrix product of x1 and x2 (py/src/linalg.rs:241-247, the `+.×` inner product). A vector on the left
    acts as one row and on the right as one column, and the result leaves that axis out. A stack's products run as one
    batched launch; operands whose cells are positive-strided with one unit stride (transposed views included) are read
    where they lie, as rustymath's `Mapped::matrix` reads them (core/src/linalg.rs:658-670)."""
    a, b = np.asarray(x1), np.asarray(x2)
    if np.issubdtype(np.result_type(a, b), np.integer) or np.result_type(a, b) == np.bool_:
        raise TypeError('integer and Boolean products (int64 `+.×`) are outside this library')
    compute, store = _compute_dtype(a, b)
    av, bv = a.ndim == 1, b.ndim == 1
    a2 = a[None, :] if av else a
    b2 = b[:, None] if bv else b
    if a2.ndim < 2 or b2.ndim < 2: raise ValueError('matmul needs arguments with at least one axis')
    m, k = a2.shape[-2:]
    kb, n = b2.shape[-2:]
    if k != kb: raise ValueError(f'matmul contracts {k} items with {kb}')
    fa, fb, frame = _pair_frames(a2.shape[:-2], b2.shape[:-2])
    count = int(np.prod(frame, dtype=np.int64))
    am, opa, lda, sa = _operand(a2.reshape(fa + (m, k)).astype(compute, copy=False), fa, frame, count, m, k)
    bm, opb, ldb, sb = _operand(b2.reshape(fb + (k, n)).astype(compute, copy=False), fb, frame, count, k, n)
    out = np.zeros((count, m, n), compute)
    if count:
        f = getattr(_lib.load(), 'hs_' + _sfx(compute) + 'gemm_batched')
        _lib.check(f(opa, opb, m, n, k, 1.0, _ptr(am), lda, sa, _ptr(bm), ldb, sb, 0.0, _ptr(out), max(n, 1), m * n,
                     count), 'matmul')
    out = out.reshape(tuple(frame) + (m, n))
    if av: out = out[..., 0, :]
    if bv: out = out[..., 0]
    out = out.astype(store, copy=False)
    return out[()] if out.ndim == 0 else out


class Lu(NamedTuple):
    "Packed LU factors in rustymath's order, where row `j` holds column `j`, and row positions (python/rustymath/linalg.py:26-51)"
    factors: np.ndarray
    p: np.ndarray

    @property
    def l(self): return lu_lower(self.factors)

    @property
    def u(self): return lu_upper(self.factors)

    def solve(self, b):
        "Solve for `b` using these factors, without factoring again or applying a cutoff"
        return lu_solve(self.factors, self.p, b)

    def inverse(self):
        "The inverse from these factors, without factoring again"
        return lu_inverse(self.factors, self.p)

    def determinant(self):
        "The determinant from these factors"
        return lu_det(self.factors, self.p)

    def log_determinant(self):
        "The determinant's sign and the natural logarithm of its magnitude, from these factors"
        return lu_slogdet(self.factors, self.p)


def _factor(a, name):
    "(factors at the compute width, perm, store dtype) of each square matrix of a stack."
    a = np.asarray(a)
    compute, store = _compute_dtype(a)
    if a.ndim < 2 or a.shape[-1] != a.shape[-2]: raise ValueError(f'{name} needs square matrices, not {a.shape[-2:]}')
    n = a.shape[-1]
    frame = a.shape[:-2]
    cells = np.ascontiguousarray(a.reshape((int(np.prod(frame, dtype=np.int64)), n, n)).astype(compute, copy=False))
    factors = np.empty_like(cells)
    perm = np.empty((cells.shape[0], n), np.int64)
    if cells.shape[0] > 1:  # a stack: one batched launch (src/linalg/stacked.rs:174-187)
        f = getattr(_lib.load(), 'hs_' + _sfx(compute) + 'lu_stack')
        _lib.check(f(cells.shape[0], n, _ptr(cells), _ptr(factors), _ptr(perm)), name)
    else:
        f = getattr(_lib.load(), 'hs_' + _sfx(compute) + 'lu_factor')
        for i in range(cells.shape[0]):
            _lib.check(f(_ptr(cells[i]), n, _ptr(factors[i]), _ptr(perm[i]), None, None), name)
    return factors.reshape(frame + (n, n)), perm.reshape(frame + (n,)), store


def lu(a):
    """The packed LU factors of each square matrix, with partial pivoting, and the positions of its rows
    (src/linalg/stacked.rs:174-187). Row `i` of `l @ u` is row `p[i]` of the matrix."""
    factors, perm, store = _factor(a, 'lu')
    return Lu(factors.astype(store, copy=False), perm)


def _positions(p, n):
    p = np.asarray(p)
    if p.shape[-1:] != (n,): raise ValueError(f'needs {n} row positions for each matrix')
    if np.issubdtype(p.dtype, np.floating):
        if not (p == np.round(p)).all(): raise IndexError('row positions must be whole numbers')
    q = p.astype(np.int64)
    if ((q < 0) | (q >= n)).any(): raise IndexError(f'row positions must lie from 0 to {n - 1}')
    return q


def _rhs(b, n):
    "(core dims, k, vector) of a right-hand side beside n-row matrices (src/linalg/stacked.rs:497-507)."
    if b.ndim == 0: raise ValueError('a right-hand side needs at least one axis')
    if b.ndim == 1:
        if b.shape[0] != n: raise ValueError(f"the right-hand side has {b.shape[0]} rows, not {n}")
        return 1, 1, True
    if b.shape[-2] != n: raise ValueError(f"the right-hand side has {b.shape[-2]} rows, not {n}")
    return 2, b.shape[-1], False


def lu_solve(factors, p, b):
    """`x` with `a @ x == b`, from the packed LU factors and row positions of each matrix `a`, as `lu` gives them,
    without factoring again (src/linalg/stacked.rs:189-204)."""
    factors, b = np.asarray(factors), np.asarray(b)
    compute, store = _compute_dtype(factors, b)
    n = factors.shape[-1]
    perm = _positions(p, n)
    core, k, vector = _rhs(b, n)
    b2 = b[:, None] if vector else b
    ff, fb, frame = _pair_frames(factors.shape[:-2], b2.shape[:-2])
    fc = _cells(factors.reshape(ff + (n, n)).astype(compute, copy=False), frame, (n, n))
    pc = _cells(perm.reshape(ff + (n,)), frame, (n,))
    bc = _cells(b2.reshape(fb + (n, k)).astype(compute, copy=False), frame, (n, k))
    out = np.zeros_like(bc)
    if fc.shape[0] > 1:  # a stack: one batched launch where the cells are small enough (stacked.rs:189-204)
        f = getattr(_lib.load(), 'hs_' + _sfx(compute) + 'lu_solve_stack')
        _lib.check(f(fc.shape[0], n, k, _ptr(fc), _ptr(np.ascontiguousarray(pc)), _ptr(bc), _ptr(out)), 'lu_solve')
    else:
        f = getattr(_lib.load(), 'hs_' + _sfx(compute) + 'lu_solve')
        for i in range(fc.shape[0]):
            _lib.check(f(_ptr(fc[i]), _ptr(pc[i]), n, _ptr(bc[i]), k, _ptr(out[i])), 'lu_solve')
    out = out.reshape(tuple(frame) + (n, k))
    if vector: out = out[..., 0]
    return out.astype(store, copy=False)


def solve(a, b):
    """The solution `x` of `a @ x == b` for each matrix `a` and right-hand side `b`, a vector or matrices
    (src/linalg/stacked.rs:88-101, core/src/linalg/factor.rs:207-226)."""
    a, b = np.asarray(a), np.asarray(b)
    compute, store = _compute_dtype(a, b)
    if a.ndim < 2: raise ValueError('solve needs matrices')
    m, n = a.shape[-2:]
    if m != n: raise NotImplementedError('non-square least squares is outside this library')
    core, k, vector = _rhs(b, n)
    b2 = b[:, None] if vector else b
    fa, fb, frame = _pair_frames(a.shape[:-2], b2.shape[:-2])
    ac = _cells(a.reshape(fa + (n, n)).astype(compute, copy=False), frame, (n, n))
    bc = _cells(b2.reshape(fb + (n, k)).astype(compute, copy=False), frame, (n, k))
    out = np.zeros_like(bc)
    if ac.shape[0] > 1:  # a stack: one batched launch where the cells are small enough (stacked.rs:88-101)
        ac, bc = np.ascontiguousarray(ac), np.ascontiguousarray(bc)
        regular = np.zeros(ac.shape[0], np.int32)
        f = getattr(_lib.load(), 'hs_' + _sfx(compute) + 'solve_stack')
        _lib.check(f(ac.shape[0], n, k, _ptr(ac), _ptr(bc), _ptr(out), _ptr(regular)), 'solve')
        ok = bool(regular.all()) or n == 0
    else:
        f = getattr(_lib.load(), 'hs_' + _sfx(compute) + 'solve')
        reg = ctypes.c_int(0)
        ok = True
        for i in range(ac.shape[0]):
            _lib.check(f(_ptr(ac[i]), n, _ptr(bc[i]), k, _ptr(out[i]), ctypes.byref(reg)), 'solve')
            ok = ok and (bool(reg.value) or n == 0)
    if not ok:
        raise NotImplementedError('a negligible LU pivot: rustymath solves this matrix by least squares, '
                                  'which is outside this library')
    out = out.reshape(tuple(frame) + (n, k))
    if vector: out = out[..., 0]
    return out.astype(store, copy=False)


def lu_inverse(factors, p):
    """The inverse of each matrix from its packed LU factors and row positions, without factoring again
    (src/linalg/stacked.rs:207-214 over core/src/linalg/factor.rs:645-664)."""
    factors = np.asarray(factors)
    compute, store = _compute_dtype(factors)
    n = factors.shape[-1]
    perm = _positions(p, n)
    frame = factors.shape[:-2]
    count = int(np.prod(frame, dtype=np.int64))
    fc = np.ascontiguousarray(factors.reshape((count, n, n)).astype(compute, copy=False))
    pc = np.ascontiguousarray(np.broadcast_to(perm, frame + (n,)).reshape((count, n)))
    out = np.empty_like(fc)
    f = getattr(_lib.load(), 'hs_' + _sfx(compute) + 'lu_inverse')
    for i in range(count):
        _lib.check(f(_ptr(fc[i]), _ptr(pc[i]), n, _ptr(out[i])), 'lu_inverse')
    return out.reshape(frame + (n, n)).astype(store, copy=False)


def _square_cells(a, name):
    a = np.asarray(a)
    compute, store = _compute_dtype(a)
    if a.ndim < 2: raise ValueError(f'{name} needs matrices')
    m, n = a.shape[-2:]
    frame = a.shape[:-2]
    cells = np.ascontiguousarray(a.reshape((int(np.prod(frame, dtype=np.int64)), m, n)).astype(compute, copy=False))
    return cells, frame, compute, store


def pinv(a):
    """The inverse of each regular square matrix, through LU with no cutoff (src/linalg/stacked.rs:125-133 over
    `LuWork::pinv_into`, core/src/linalg/factor.rs:211-219). A matrix that is not square, or has a negligible pivot,
    takes the SVD's pseudoinverse in rustymath, which is outside this library."""
    cells, frame, compute, store = _square_cells(a, 'pinv')
    m, n = cells.shape[-2:]
    if m != n: raise NotImplementedError('the pseudoinverse of a non-square matrix is outside this library')
    out = np.empty_like(cells)
    f = getattr(_lib.load(), 'hs_' + _sfx(compute) + 'inverse')
    regular = ctypes.c_int(0)
    for i in range(cells.shape[0]):
        _lib.check(f(_ptr(cells[i]), n, _ptr(out[i]), None, ctypes.byref(regular)), 'pinv')
        if n and not regular.value:
            raise NotImplementedError('a negligible LU pivot: rustymath takes the SVD pseudoinverse of this matrix, '
                                      'which is outside this library')
    return out.reshape(frame + (n, n)).astype(store, copy=False)


def det(a):
    """The determinant of each square matrix, from its LU factors at the compute width, then stored
    (src/linalg/stacked.rs:136-144, factor.rs:174)."""
    factors, perm, store = _factor(a, 'det')
    d = np.asarray(lu_det(factors, perm)).astype(store, copy=False)
    return d[()] if d.ndim == 0 else d


def slogdet(a):
    """The sign of each square matrix's determinant and the natural logarithm of its magnitude, from its LU factors
    at the compute width (src/linalg/stacked.rs:160-171, factor.rs:176-182)."""
    factors, perm, store = _factor(a, 'slogdet')
    sign, log = lu_slogdet(factors, perm)
    sign = np.asarray(sign).astype(store, copy=False)
    log = np.asarray(log).astype(store, copy=False)
    return (sign[()] if sign.ndim == 0 else sign), (log[()] if log.ndim == 0 else log)


def cholesky(a):
    """The lower triangular L with `L @ L.T == a` for each symmetric positive definite matrix, reading only its lower
    triangle (src/linalg/stacked.rs:260-271 over core/src/linalg/factor.rs:777-797). A matrix that isn't positive
    definite gives NaN, with no warning."""
    cells, frame, compute, store = _square_cells(a, 'cholesky')
    m, n = cells.shape[-2:]
    if m != n: raise ValueError(f'cholesky needs square matrices, not {(m, n)}')
    out = np.empty_like(cells)
    f = getattr(_lib.load(), 'hs_' + _sfx(compute) + 'cholesky')
    ok = ctypes.c_int(0)
    for i in range(cells.shape[0]):
        _lib.check(f(_ptr(cells[i]), n, _ptr(out[i]), ctypes.byref(ok)), 'cholesky')
        if not ok.value: out[i] = np.nan
    return out.reshape(frame + (n, n)).astype(store, copy=False)


def lu_lower(factors):
    "The unit lower triangle of each matrix's packed LU factors (core/src/linalg/factor.rs:158-165)."
    f = np.swapaxes(np.asarray(factors), -1, -2)
    n = f.shape[-1]
    return np.tril(f, -1) + np.eye(n, dtype=f.dtype)


def lu_upper(factors):
    "The upper triangle of each matrix's packed LU factors (core/src/linalg/factor.rs:166-172)."
    return np.triu(np.swapaxes(np.asarray(factors), -1, -2))


def _odd(perm):
    "Whether a permutation is odd, by counting its cycles at their smallest positions (factor.rs:186-194)."
    n, cycles = len(perm), 0
    for s in range(n):
        i = perm[s]
        while i > s: i = perm[i]
        cycles += i == s
    return (n - cycles) % 2 == 1


def lu_det(factors, p):
    """The determinant from packed factors and row positions: U's diagonal multiplied in order from 1, negated for an
    odd permutation (factor.rs:174). An O(n) fold over the factors the GPU computed."""
    factors = np.asarray(factors); n = factors.shape[-1]
    perm = _positions(p, n)
    fc = factors.reshape(int(np.prod(factors.shape[:-2], dtype=np.int64)), n, n); pc = perm.reshape(fc.shape[0], n)
    out = np.empty(fc.shape[0], factors.dtype)
    with np.errstate(over='ignore', invalid='ignore'):  # the fold overflows to ±inf as the reference's does
        for i in range(fc.shape[0]):
            d = factors.dtype.type(1)
            for j in range(n): d = d * fc[i, j, j]
            out[i] = -d if _odd(pc[i]) else d
    out = out.reshape(factors.shape[:-2])
    return out[()] if out.ndim == 0 else out


def lu_slogdet(factors, p):
    "The determinant's sign and the log of its magnitude from packed factors and row positions (factor.rs:176-182)."
    factors = np.asarray(factors); n = factors.shape[-1]
    t = factors.dtype.type
    perm = _positions(p, n)
    fc = factors.reshape(int(np.prod(factors.shape[:-2], dtype=np.int64)), n, n); pc = perm.reshape(fc.shape[0], n)
    signs = np.empty(fc.shape[0], factors.dtype); logs = np.empty(fc.shape[0], factors.dtype)
    for i in range(fc.shape[0]):
        s, l = (t(-1) if _odd(pc[i]) else t(1)), t(0)
        for j in range(n):
            d = fc[i, j, j]; m = abs(d)
            if m == 0: s, l = t(0), t(-np.inf)
            else: s, l = s * (d * (t(1) / m)), l + np.log(m)
        signs[i], logs[i] = s, l
    shape = factors.shape[:-2]
    signs, logs = signs.reshape(shape), logs.reshape(shape)
    return (signs[()], logs[()]) if not shape else (signs, logs)
