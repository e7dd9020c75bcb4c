"""Pins the oracle (oracle/oracle.c) before anything is checked against it.

Each check is one that a plausible mistake in the restatement would fail: the reference's own known answers
(exact products, the subnormal-pivot factors, the notebook's signed zeros, determinants), its invariant tests
(residual bounds, exact power-of-two scaling, zero columns), a textbook library routine that shares the pivot rule
(LAPACK getrf through SciPy), a higher-precision product, and equality between the recursive restatement and the
flat one-column-at-a-time statement of the same arithmetic.
"""
import math
import numpy as np
import pytest
import scipy.linalg

import oracle
from support import (FOLD_SIZES, KC, MC, TRSM_CODES, TRSM_SHAPES, integer_operands, random, norm, distance,
                     backward_error, triangle_system)

EPS = {np.float64: np.finfo(np.float64).eps, np.float32: np.finfo(np.float32).eps}


def bits_equal(x, y):
    """Equal bit for bit, so signed zeros count, with every NaN equal to every other: which operand's NaN an
    operation passes on, and its sign, is the hardware's choice (x86's default NaN is negative)."""
    x, y = np.array(x, copy=True), np.array(y, copy=True)
    if x.shape != y.shape or x.dtype != y.dtype: return False
    nx, ny = np.isnan(x), np.isnan(y)
    if not (nx == ny).all(): return False
    x[nx] = 0; y[ny] = 0
    return x.tobytes() == y.tobytes()


@pytest.mark.parametrize('size', FOLD_SIZES)
@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_products_match_folds_in_order(size, dtype):
    "tests/linalg.rs:160-177: small whole numbers make every sum exact, so the product equals the exact one."
    m, k, n = size
    a, b = integer_operands(m, k, n)
    want = (a.astype(np.int64) @ b.astype(np.int64)).astype(dtype)
    c = np.ones((m, n), dtype)
    oracle.gemm(a.astype(dtype), b.astype(dtype), c=c)
    assert bits_equal(c, want)
    assert bits_equal(oracle.gemm(a.astype(dtype), b.astype(dtype), naive=True), want)


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
@pytest.mark.parametrize('ops', ['NN', 'NT', 'TN', 'TT'])
@pytest.mark.parametrize('ab', [(1.0, 0.0), (-1.0, 1.0), (2.0, 0.5), (0.75, 1.0)])
def test_blocked_kernels_match_the_in_order_fold(dtype, ops, ab):
    """The blocked GotoBLAS driver with AVX2 tiles gives the same bits as the defining fold (one fused multiply-add
    per step of p, in order): KC blocks reload c exactly, so blocking never changes a result."""
    rng = np.random.default_rng(5)
    alpha, beta = ab
    for m, k, n in [(MC + 7, KC * 2 + 3, 19), (13, 300, 2053 + 16), (64, 64, 64), (6, 9, 17)]:
        a = rng.standard_normal((m, k) if ops[0] == 'N' else (k, m)).astype(dtype)
        b = rng.standard_normal((k, n) if ops[1] == 'N' else (n, k)).astype(dtype)
        c0 = rng.standard_normal((m, n)).astype(dtype)
        got = oracle.gemm(a, b, alpha, beta, c0.copy(), ops[0], ops[1])
        want = oracle.gemm(a, b, alpha, beta, c0.copy(), ops[0], ops[1], naive=True)
        assert bits_equal(got, want)


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_products_within_the_rounding_bound(dtype):
    "Each result lies within k·eps·(|A||B|) of the product in extended precision (the error bound of any fold)."
    rng = np.random.default_rng(0)
    for m, k, n in [(200, 333, 150), (1, 500, 300), (300, 500, 1), (40, 5, 40), (7, 3, 1), (1, 6, 9)]:
        a, b = rng.standard_normal((m, k)), rng.standard_normal((k, n))
        got = oracle.matmul(a.astype(dtype), b.astype(dtype)).astype(np.longdouble)
        exact = a.astype(dtype).astype(np.longdouble) @ b.astype(dtype).astype(np.longdouble)
        bound = k * EPS[dtype] * (np.abs(a.astype(dtype)).astype(np.longdouble) @ np.abs(b.astype(dtype)).astype(np.longdouble))
        assert (np.abs(got - exact) <= bound * 1.0001 + np.finfo(dtype).tiny).all()


def test_vector_products_read_vectors_as_rows_and_columns():
    "rustymath's matmul with a vector acts as one row on the left and one column on the right (py/src/linalg.rs:241)."
    a, b = integer_operands(9, 20, 7)
    v = np.arange(20.0) - 9
    assert bits_equal(oracle.matmul(a, v), a @ v)
    w = np.arange(9.0) - 4
    assert bits_equal(oracle.matmul(w, a), w @ a)
    assert oracle.matmul(v, v) == v @ v


def test_vector_products_take_the_reference_summation_orders():
    """Known answers for each of `by_vector`'s summation orders (core/src/linalg.rs:1020-1073), each chosen with
    B = 2^53 so that an in-order fold or another pairing gives a different result (B + 1 rounds to B, B + 2 does not).
    - `dots` (:903-917, k > 8): a row of 1, 0, -B, 0, B and zeros puts 1 and B in running sum 0, -B in sum 2; the
      pairwise reduction gives ((1 + B) + -B) = 0. Rows past the last quad take `dot` (factor.rs:466-473), eight sums
      folded in order from zero: 1, 1 - B, 1. Items past the last chunk of eight add after the reduction: one more 1
      gives 0 + 1 = 1, where folding it into sum 0 first, or an in-order fold, gives 2.
    - `narrow_dots` (:923-937, k <= 8): two running sums by item parity. The same 8-item row keeps 1, -B, B in sum 0:
      1 - B, then 1 (where `dots` would give 0). B, 1, 1, -B gives (B + 1) + (1 - B) = B + (1 - B) = 1; in order, 0.
    - `narrow_sums` (:941-956, a transposed matrix with at most 8 results): stored rows whose 8-item read stays in
      the data go in quads, row p of a quad into set p mod 4, the sets summed (s0 + s2) + (s1 + s3). Stored rows B, 1,
      -B, 1 of 8 items give (B - B) + (1 + 1) = 2; in order or paired by neighbours, 1. With 5 items to a row only
      the first 3 rows are whole, fewer than a quad, so every row folds into set 0 in order: 1.
    - `scaled_runs` (:882-900, more than 8 results): an in-order fold over the stored rows, running B, B, 0, 1: 1."""
    B = 2.0 ** 53
    r8 = [1.0, 0.0, -B, 0.0, B, 0.0, 0.0, 0.0]
    assert bits_equal(oracle.matmul(np.tile(r8 + [0.0] * 8, (5, 1)), np.ones(16)), np.array([0.0, 0.0, 0.0, 0.0, 1.0]))
    assert bits_equal(oracle.matmul(np.tile(r8 + [0.0] * 8 + [1.0], (4, 1)), np.ones(17)), np.ones(4))
    assert bits_equal(oracle.matmul(np.tile(r8, (3, 1)), np.ones(8)), np.ones(3))
    assert bits_equal(oracle.matmul(np.tile([B, 1.0, 1.0, -B], (3, 1)), np.ones(4)), np.ones(3))
    stored = lambda n: np.tile(np.array([B, 1.0, -B, 1.0])[:, None], (1, n))
    assert bits_equal(oracle.matmul(np.ones(4), stored(8)), np.full(8, 2.0))
    assert bits_equal(oracle.gemm(stored(8), np.ones((4, 1)), opa='T')[:, 0], np.full(8, 2.0))
    assert bits_equal(oracle.gemm(stored(5), np.ones((4, 1)), opa='T')[:, 0], np.ones(5))
    assert bits_equal(oracle.matmul(np.ones(4), stored(20)), np.ones(20))
    # FP32 takes the same orders with B = 2^24.
    b32 = np.float32(2.0 ** 24)
    row32 = np.array([1, 0, -b32, 0, b32] + [0] * 11, np.float32)
    assert bits_equal(oracle.matmul(np.tile(row32, (5, 1)), np.ones(16, np.float32)),
                      np.array([0, 0, 0, 0, 1], np.float32))


def test_notebook_known_answers():
    """nbs/linalg.ipynb cells 4, 6, 25, 27 and README.md:79-82, including the signed zeros that the order of the
    substitutions fixes: U's solve divides by a reciprocal for several right-hand sides and by a quotient for one."""
    a = np.array([[1., 2], [2, 5]])
    f, perm, _ = oracle.lu(a)
    assert perm.tolist() == [1, 0]
    assert bits_equal(oracle.lower(f), np.array([[1., 0], [0.5, 1]]))
    assert bits_equal(oracle.upper(f), np.array([[2., 5], [0, -0.5]]))
    x = oracle.lu_solve(f, perm, np.array([[1., 2], [2, 1]]))
    assert bits_equal(x, np.array([[1., 8], [-0., -3]]))
    assert bits_equal(oracle.lu_inverse(f, perm), np.array([[5., -2], [-2, 1]]))
    assert oracle.det(f, perm) == 1.0 and oracle.slogdet(f, perm) == (1.0, 0.0)
    rhs = np.array([[1., 2], [2, 1]])
    want = [[[1.0, 8.0], [-0.0, -3.0]], [[0.5, 4.0], [-0.0, -1.5]], [[0.33333333, 2.6666667], [-0.0, -1.0]]]
    vectors = [[1.0, -0.0], [4.0, -1.5], [1.0, -0.33333333]]
    for s in range(3):
        m = (s + 1) * a
        x = oracle.solve(m, rhs)
        assert np.allclose(x, want[s], rtol=1e-7, atol=0) and (np.signbit(x) == np.signbit(want[s])).all()
        v = oracle.solve(m, np.array([[1., 2], [2, 1], [1, 1]])[s])
        assert np.allclose(v, vectors[s], rtol=1e-7, atol=0) and (np.signbit(v) == np.signbit(vectors[s])).all()
    x = oracle.solve(np.array([[4., 1], [1, 3]]), np.array([1., 2]))
    assert np.allclose(x, [0.090909091, 0.63636364], rtol=1e-8)


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_subnormal_pivot(dtype):
    "tests/linalg.rs:510-511: lu([p, 1; q, 1]).factors == [p, q/p, 1, 1 - q/p], p, q below the normal range."
    tiny = np.finfo(dtype).tiny
    p, q = dtype(tiny / 1024), dtype(tiny / 1048576)
    f, perm, _ = oracle.lu(np.array([[p, 1], [q, 1]], dtype))
    want = np.array([p, q / p, 1, dtype(1) - q / p], dtype)
    assert bits_equal(f.T.reshape(-1), want) or bits_equal(f.reshape(-1), want)
    assert bits_equal(f.reshape(-1), want)


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_determinants(dtype):
    "tests/linalg.rs:553-558: a swapped diag(1, 2, 3, 4) has determinant -24; a zero column gives 0 and -inf."
    swapped = np.array([[0., 2, 0, 0], [1, 0, 0, 0], [0, 0, 3, 0], [0, 0, 0, 4]], dtype)
    assert oracle.det(*oracle.lu(swapped)[:2]) == -24
    f, perm, _ = oracle.lu(np.array([[1., 0], [3, 0]], dtype))
    assert oracle.det(f, perm) == 0 and oracle.slogdet(f, perm) == (0, -np.inf)


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
@pytest.mark.parametrize('n', [0, 1, 2, 5, 31, 32, 33, 71])
def test_decompositions(dtype, n):
    "tests/linalg.rs:373-390: L U is A with its rows permuted, and solves and the inverse have small residuals."
    bound = 4 * EPS[dtype] * max(n, 1)
    a = random(n, n, n + 1, dtype)
    aw = a.astype(np.float64)
    f, perm, _ = oracle.lu(a)
    l, u = oracle.lower(f).astype(np.float64), oracle.upper(f).astype(np.float64)
    assert distance(l @ u, aw[perm], max(norm(aw), 1)) < bound
    for k in [1, 3]:
        b = random(n, k, 99, dtype)
        x = oracle.lu_solve(f, perm, b).astype(np.float64)
        assert distance(aw @ x, b, max(norm(aw) * norm(x), 1)) < bound
    inv = oracle.lu_inverse(f, perm).astype(np.float64)
    assert distance(aw @ inv, np.eye(n), max(norm(aw) * norm(inv), 1)) < bound


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_hard_matrices_scale_exactly(dtype):
    """tests/linalg.rs:466-483: scaling by a power of two near underflow or overflow gives the same permutation,
    U scaled exactly, L unchanged, the same solves of a scaled right-hand side and an exactly scaled inverse."""
    n = 40
    fi = np.finfo(dtype)
    tiny = dtype(fi.tiny / fi.eps)
    quartic = dtype(2.0 ** (int(math.log2(float(fi.max)) / 4) + 6))
    a, b = random(n, n, 17, dtype), random(n, 3, 19, dtype)
    f, perm, _ = oracle.lu(a)
    for s in [tiny, dtype(1) / tiny, quartic]:
        g, gperm, _ = oracle.lu(a * s)
        assert (gperm == perm).all()
        scaled = f.copy()
        iu = np.triu_indices(n)            # column-major storage: row j of `f` is column j, so U is the lower part
        scaled[iu[1], iu[0]] = f[iu[1], iu[0]] * s
        assert bits_equal(g, scaled)
        assert bits_equal(oracle.lu_solve(g, gperm, b * s), oracle.lu_solve(f, perm, b))
        assert bits_equal(oracle.lu_inverse(g, gperm) * s, oracle.lu_inverse(f, perm))


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_zero_columns(dtype):
    "tests/linalg.rs:500-509: a zero column leaves finite factors with a zero pivot and a small residual."
    n = 71
    for zero in [0, n // 2, n - 1]:
        a = random(n, n, 29, dtype)
        a[:, zero] = 0
        f, perm, _ = oracle.lu(a)
        assert np.isfinite(f).all() and (np.diag(f) == 0).any()
        l, u = oracle.lower(f).astype(np.float64), oracle.upper(f).astype(np.float64)
        assert distance(l @ u, a.astype(np.float64)[perm], norm(a)) < 4 * EPS[dtype] * n


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
@pytest.mark.parametrize('n', [530, 1030])
def test_blocked_sizes(dtype, n):
    "tests/linalg.rs:569-590: LU, solve and inverse at sizes past the wider blocks, checked with random vectors."
    bound = 4 * EPS[dtype] * n
    a = random(n, n, n, dtype); aw = a.astype(np.float64)
    x = random(n, 1, 3, dtype).astype(np.float64)[:, 0]
    f, perm, _ = oracle.lu(a)
    l, u = oracle.lower(f).astype(np.float64), oracle.upper(f).astype(np.float64)
    assert distance(l @ (u @ x), (aw @ x)[perm], norm(aw) * norm(x)) < bound
    b = random(n, 3, 5, dtype)
    sol = oracle.lu_solve(f, perm, b).astype(np.float64)
    assert distance(aw @ sol, b, norm(aw) * norm(sol)) < bound
    inv = oracle.lu_inverse(f, perm).astype(np.float64)
    assert distance(aw @ (inv @ x), x, norm(aw) * norm(inv) * norm(x)) < bound


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_recursion_equals_the_flat_in_order_factorization(dtype):
    """lu_panel's recursion (halving to 16-column leaves, triangular solves and products between them) gives the
    same bits as the flat right-looking statement: every item takes its updates in pivot order, one fused
    multiply-add each, and a zero pivot skips only the rest of its own leaf. This is the arithmetic that the GPU
    factorization reproduces, so it is checked on zero columns, NaN and infinities too."""
    rng = np.random.default_rng(11)
    cases = []
    for n in [1, 2, 3, 16, 17, 33, 64, 100, 129, 257]:
        cases.append(rng.standard_normal((n, n)))
    z = rng.standard_normal((100, 100)); z[:, [3, 17, 50, 99]] = 0; cases.append(z)
    z = rng.standard_normal((70, 70)); z[:, 5] = 0; z[7, 5] = np.nan; cases.append(z)
    z = rng.standard_normal((70, 70)); z[3, 3] = np.inf; z[9, 40] = -np.inf; cases.append(z)
    z = np.zeros((40, 40)); z[np.arange(40), (np.arange(40) * 7) % 40] = 1; cases.append(z)
    for a in cases:
        a = a.astype(dtype)
        f1, p1, v1 = oracle.lu(a)
        f2, p2, v2 = oracle.lu(a, flat=True)
        assert bits_equal(f1, f2) and (p1 == p2).all() and (v1 == v2).all()


def test_pivots_match_lapack_getrf():
    """Partial pivoting with the first largest magnitude is LAPACK getrf's rule (idamax): SciPy's lu_factor picks
    the same rows wherever no near-tie exists, and random normal matrices of these sizes have none."""
    rng = np.random.default_rng(2)
    for n in [10, 64, 200, 333]:
        a = rng.standard_normal((n, n))
        _, _, piv = oracle.lu(a)
        _, lapack = scipy.linalg.lu_factor(a)
        assert (piv == lapack).all()


def test_lu_stack_fixture():
    "tests/test_linalg.py:14-34: 64 matrices base + k/8 I, solved and inverted against NumPy's LAPACK."
    base = np.array([[0., 2, 1], [3, 1, -1], [1, 0, 4]])
    b = np.arange(12.).reshape(3, 4)[:, ::-1]
    for dt in [np.float32, np.float64]:
        for k in range(64):
            a = (base + k / 8 * np.eye(3)).astype(dt)
            f, perm, _ = oracle.lu(a)
            assert np.allclose(oracle.lower(f) @ oracle.upper(f), a[perm], rtol=0.002, atol=0.002)
            wide = a.astype(np.float64)
            assert np.allclose(oracle.lu_solve(f, perm, b.astype(dt)), np.linalg.solve(wide, b), rtol=0.003, atol=0.003)
            assert np.allclose(oracle.lu_inverse(f, perm), np.linalg.inv(wide), rtol=0.003, atol=0.003)
            assert np.allclose(oracle.det(f, perm), np.linalg.det(wide), rtol=0.003, atol=0.003)


def test_regular_rule():
    "factor.rs:221-226: a negligible pivot (here an exactly singular matrix) gives no LU solution."
    assert oracle.solve(np.array([[1., 2], [2, 4]]), np.array([1., 2])) is None
    assert oracle.solve(np.zeros((3, 3)), np.ones(3)) is None
    x = oracle.solve(np.array([[2., 0], [0, 4]]), np.array([1., 1]))
    assert x.tolist() == [0.5, 0.25]


def test_reference_generator_matches_known_bits():
    """SplitMix64's published first outputs for seed 0 (Vigna's reference implementation): 0xe220a8397b1dcdaf,
    0x6e789e6aa1b965f4, 0x06c45d188009454f, which pin `support.splitmix` and so the reference tests' inputs."""
    from support import splitmix
    got = [int(v) for v in splitmix(np.arange(3, dtype=np.uint64))]
    assert got == [0xe220a8397b1dcdaf, 0x6e789e6aa1b965f4, 0x06c45d188009454f]


def test_backward_error_of_a_plain_solve():
    "The acceptance bound of the FP64 solve, 10·n·eps64 on the normwise backward error, on a 300×300 N(0,1) system."
    a = np.random.default_rng(1).standard_normal((300, 300)); b = np.random.default_rng(1).standard_normal(300)
    x = oracle.solve(a, b)
    assert backward_error(a, x, b) <= 10 * 300 * EPS[np.float64]


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_transposed_solve(dtype):
    """The adjoint's transposed solve (no reference counterpart): Aᵀ x = b within the backward-stable bound, LAPACK's
    solution of Aᵀ x = b within the forward bound cond(A)·10·n·eps, and exact on a scaled permutation (powers of two),
    where every term is exact."""
    rng = np.random.default_rng(11)
    for n in (1, 2, 7, 33, 130):
        a = rng.standard_normal((n, n)).astype(dtype)
        b = rng.standard_normal((n, 3)).astype(dtype)
        f, perm, _ = oracle.lu(a)
        x = oracle.lu_solve_t(f, perm, b).astype(np.float64)
        r = np.linalg.norm(a.T.astype(np.float64) @ x - b) / (np.linalg.norm(a) * np.linalg.norm(x))
        assert r <= 10 * n * np.finfo(dtype).eps
        want = np.linalg.solve(a.T.astype(np.float64), b.astype(np.float64))
        fwd = np.linalg.norm(x - want) / np.linalg.norm(want)
        assert fwd <= 10 * n * np.finfo(dtype).eps * np.linalg.cond(a.astype(np.float64))
    p = np.eye(6, dtype=dtype)[[3, 0, 5, 1, 2, 4]] * np.array([2, 4, 8, 0.5, 1, 16], dtype=dtype)[:, None]
    b = rng.standard_normal((6, 2)).astype(dtype)
    f, perm, _ = oracle.lu(p)
    x = oracle.lu_solve_t(f, perm, b)
    assert np.array_equal(x, np.linalg.solve(p.T.astype(np.float64), b.astype(np.float64)).astype(dtype))


def spd_of(n, dtype):
    "tests/linalg.rs:391-392: A Aᵀ + n I for A = random(n, n, n + 1), formed in float64 and narrowed."
    aw = random(n, n, n + 1, dtype).astype(np.float64)
    return (aw @ aw.T + n * np.eye(n)).astype(dtype)


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
@pytest.mark.parametrize('n', [0, 1, 2, 5, 31, 32, 33, 71, 300])
def test_cholesky(dtype, n):
    """tests/linalg.rs:391-399: L is lower triangular with L Lᵀ near A A^T + n I, and a negative definite matrix has
    no factor; the blocks and the halving (block width 32 here, 37 at n = 300) equal the flat left-looking statement
    of the same order bit for bit; NumPy's (LAPACK's) factor agrees to rounding (tests/test_linalg.py:109-110)."""
    spd = spd_of(n, dtype)
    l = oracle.cholesky(spd)
    assert l is not None and (np.triu(l, 1) == 0).all()
    lw, sw = l.astype(np.float64), spd.astype(np.float64)
    bound = 4 * np.finfo(dtype).eps * max(n, 1)
    assert distance(lw @ lw.T, sw, max(norm(sw), 1.0)) < bound
    assert bits_equal(l, oracle.cholesky(spd, flat=True))
    if n: 
        assert oracle.cholesky(-spd) is None
        assert np.allclose(l, np.linalg.cholesky(sw), rtol=100 * np.finfo(dtype).eps * n, atol=1e-30)


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_cholesky_scales_exactly(dtype):
    "tests/linalg.rs:484-486: cholesky(A c) == cholesky(A) sqrt(c) bit for bit for an even power of two c."
    n = 70
    fi = np.finfo(dtype)
    tiny = dtype(fi.tiny / fi.eps)
    quartic = dtype(2.0 ** (int(math.log2(float(fi.max)) / 4) + 6))
    spd = spd_of(n, dtype)
    l = oracle.cholesky(spd)
    for s in [tiny, dtype(1) / tiny, quartic]:
        c = s if np.sqrt(s) * np.sqrt(s) == s else s + s
        assert bits_equal(oracle.cholesky(spd * c), l * np.sqrt(c))


def test_cholesky_known_answers():
    """A hand-worked factor: [[4, 2], [2, 5]] = [[2, 0], [1, 2]] [[2, 1], [0, 2]]; only the lower triangle is read;
    a zero or NaN diagonal has no factor (tests/test_linalg.py:180-181 gives NaN for -I)."""
    assert bits_equal(oracle.cholesky(np.array([[4., 99], [2, 5]])), np.array([[2., 0], [1, 2]]))
    assert oracle.cholesky(-np.eye(2)) is None and oracle.cholesky(np.zeros((3, 3))) is None
    assert oracle.cholesky(np.array([[np.nan]])) is None


def test_float_products_round_once():
    """A float FMA whose product is a float32 midpoint and whose addend is tiny: rounding once gives the upper
    neighbour, rounding through double first (an FP64 FMA of promoted operands) lands on the midpoint and then on the
    even neighbour below. x1 = b1 - x0 L10 for L = [[1, 0], [-(1/2 + 2^-13), 1]], x0 = 2 + 2^-11, b1 = 2^-60."""
    f = np.float32
    a = np.array([[1, 0], [-(0.5 + 2.0 ** -13), 1]], f)
    b = np.array([[2 + 2.0 ** -11], [2.0 ** -60]], f)
    fac, perm, _ = oracle.lu(a)
    x = oracle.lu_solve(fac, perm, b)
    assert (perm == [0, 1]).all() and x[1, 0] == f(1 + 2.0 ** -11 + 2.0 ** -23)
    assert x[1, 0] != f(np.float64(b[0, 0]) * np.float64(0.5 + 2.0 ** -13) + np.float64(b[1, 0]))


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
@pytest.mark.parametrize('shape', TRSM_SHAPES)
def test_triangular_solves_are_exact_on_whole_numbers(dtype, shape):
    """`Tiled::triangular` (core/src/linalg.rs:781) for every code: side, stored triangle, transpose and unit diagonal.
    On whole-number systems every order of the terms is exact, so the solution is the exact one; a dropped term, a
    transposed view, the wrong triangle or a diagonal that should count as one all change it. NaN outside the named
    triangle (and 7s on a unit diagonal) must never be read."""
    side, m, n, ldx = shape
    for k, (uplo, trans, diag) in enumerate(TRSM_CODES):
        a, x, b = triangle_system(side, m, n, uplo, trans, diag, dtype, seed=k + 10 * m + n)
        got = oracle.trsm(a, b, side, uplo, trans, diag, ldx=ldx)
        assert np.array_equal(got, x), (side, m, n, ldx, uplo, trans, diag)


def test_single_row_solve_takes_dot_with_eight_running_sums():
    """`single` for a triangle read by rows (factor.rs:745-746) takes each item's `dot` (`sum_of_products`,
    factor.rs:466-473): eight running sums over whole chunks of eight, folded in order from zero, then the remaining
    terms one fused multiply-add each. A known answer that an in-order sum would miss: row 16 of a unit lower triangle
    holds 1 at column 0, -2^53 at column 1 and 2^53 at column 8, zeros elsewhere, with x = 1 above it. Sum 0 takes
    1 then 2^53 (2^53 + 1 rounds to even, 2^53) and sum 1 takes -2^53, so the dot is 0 and x_16 = b_16 - 0 = 0; in
    order, 1 - 2^53 is exact and adding 2^53 gives 1."""
    m = 17
    a = np.eye(m)
    a[16, 0], a[16, 1], a[16, 8] = 1.0, -2.0 ** 53, 2.0 ** 53
    b = np.ones((m, 1)); b[16, 0] = 0.0
    x = oracle.trsm(a, b, 'L', 'L', 'N', 'U')
    assert np.array_equal(x[:16, 0], np.ones(16))
    assert x[16, 0] == 0.0
    # The other reading order of the same triangle (columns, through `axpy`) takes the terms in order: 0 - 1 + 2^53 ...
    xt = oracle.trsm(np.ascontiguousarray(a.T), b, 'L', 'U', 'T', 'U')
    assert xt[16, 0] == -1.0


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_triangular_solves_reproduce_lu_solve(dtype):
    """lu_solve (factor.rs:633-641) is the two triangular solves of its factors' transposed view: the stored unit upper
    triangle (L) then the stored lower one (U), both with code 112. The general entry reproduces it bit for bit."""
    rng = np.random.default_rng(21)
    for n, k in ((50, 1), (70, 5), (33, 40)):
        a = rng.standard_normal((n, n)).astype(dtype)
        b = rng.standard_normal((n, k)).astype(dtype)
        f, perm, _ = oracle.lu(a)
        y = oracle.trsm(f, b[perm], 'L', 'U', 'T', 'U')
        x = oracle.trsm(f, y, 'L', 'L', 'T', 'N')
        assert bits_equal(x, oracle.lu_solve(f, perm, b))


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_triangular_solves_within_the_rounding_bound(dtype):
    """Random well-conditioned triangles (diagonal items 2 to 3 in magnitude): the residual |op(A) X - B| stays within
    the substitution's rounding bound against SciPy's solve_triangular on every code and shape."""
    rng = np.random.default_rng(4)
    eps = np.finfo(dtype).eps
    for side, m, n, ldx in TRSM_SHAPES:
        order = m if side == 'L' else n
        for uplo, trans, diag in TRSM_CODES:
            a = (rng.standard_normal((order, order)) / np.sqrt(order)).astype(dtype)
            a[np.diag_indices(order)] = rng.uniform(2, 3, order) * rng.choice([-1, 1], order)
            b = rng.standard_normal((m, n)).astype(dtype)
            x = oracle.trsm(a, b, side, uplo, trans, diag, ldx=ldx).astype(np.float64)
            t = np.tril(a, -1) if uplo == 'L' else np.triu(a, 1)
            t = (t + (np.eye(order) if diag == 'U' else np.diag(np.diag(a)))).astype(np.float64)
            op = t.T if trans != 'N' else t
            lower = (uplo == 'L') == (trans == 'N')  # the triangle op(A) has nonzeros in
            if side == 'L': want = scipy.linalg.solve_triangular(op, b.astype(np.float64), lower=lower)
            else: want = scipy.linalg.solve_triangular(op.T, b.T.astype(np.float64), lower=not lower).T
            r = (op @ x - b) if side == 'L' else (x @ op - b)
            scale = (np.abs(op) @ np.abs(x)) if side == 'L' else (np.abs(x) @ np.abs(op))
            assert (np.abs(r) <= 4 * order * eps * scale + 1e-30).all()
            assert np.abs(x - want).max() <= 100 * order * eps * max(np.abs(want).max(), 1)
