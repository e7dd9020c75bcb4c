"""GPU parity of the products, the LU factorization and the solves against the oracle, through the C ABI.

The bar: the same bits as the oracle's restatement of rustymath wherever the arithmetic is defined item by item
(exact products, pivots, factors, solves, including signed zeros), and the reference's own invariants (residual
bounds, power-of-two scaling, zero columns) beside them. Sizes stay where the oracle finishes in well under a
second, save the benchmark's own orders (FP64 8192, FP32 16384: 10 to 30 s of the oracle).
"""
import math
import numpy as np
import pytest

import oracle
from support import FOLD_SIZES, integer_operands, random, norm, distance, backward_error
from test_oracle import bits_equal

pytestmark = pytest.mark.gpu
EPS = {np.float64: np.finfo(np.float64).eps, np.float32: np.finfo(np.float32).eps}


@pytest.fixture(scope='module')
def la():
    from rustymath_amd import linalg
    return linalg


def hs_gemm(a, b, alpha=1.0, beta=0.0, c=None, opa='N', opb='N'):
    import ctypes
    from rustymath_amd import _lib
    a = np.ascontiguousarray(a); b = np.ascontiguousarray(b, dtype=a.dtype)
    m, k = (a.shape[1], a.shape[0]) if opa != 'N' else a.shape
    n = b.shape[0] if opb != 'N' else b.shape[1]
    c = np.zeros((m, n), a.dtype) if c is None else np.ascontiguousarray(c)
    f = getattr(_lib.load(), 'hs_dgemm' if a.dtype == np.float64 else 'hs_sgemm')
    p = lambda x: x.ctypes.data_as(ctypes.c_void_p)
    _lib.check(f(oracle.OPS[opa], oracle.OPS[opb], m, n, k, alpha, p(a), max(a.shape[1], 1), p(b), max(b.shape[1], 1),
                 beta, p(c), max(n, 1)), 'gemm')
    return c


@pytest.mark.parametrize('size', FOLD_SIZES)
@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_products_match_folds_in_order(la, size, dtype):
    "tests/linalg.rs:160-177 on the GPU: exact small-integer products, whatever the summation order."
    m, k, n = size
    a, b = integer_operands(m, k, n)
    want = (a.astype(np.int64) @ b.astype(np.int64)).astype(dtype)
    assert bits_equal(la.matmul(a.astype(dtype), b.astype(dtype)), want)


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
@pytest.mark.parametrize('ops', ['NN', 'NT', 'TN', 'TT'])
@pytest.mark.parametrize('ab', [(1.0, 0.0), (-1.0, 1.0), (2.0, 0.5)])
def test_products_match_the_oracle_bit_for_bit(dtype, ops, ab):
    """The MFMA product takes each result's terms in order of p, one rounding each, from 0 or beta c: the oracle's
    (and rustymath's) fold, bit for bit, on random normal operands with ragged edges."""
    rng = np.random.default_rng(7)
    alpha, beta = ab
    for m, k, n in [(131, 261, 37), (200, 1000, 129), (16, 16, 16), (3, 5, 2)]:
        a = rng.standard_normal((m, k) if ops[0] == 'N' else (k, m)).astype(dtype)
        b = rng.standard_normal((k, n) if ops[1] == 'N' else (n, k)).astype(dtype)
        c0 = rng.standard_normal((m, n)).astype(dtype)
        got = hs_gemm(a, b, alpha, beta, c0.copy(), ops[0], ops[1])
        want = oracle.gemm(a, b, alpha, beta, c0.copy(), ops[0], ops[1])
        assert bits_equal(got, want), f'{(m, k, n)}: max diff {np.abs(got - want).max()}'


def test_benchmark_product_matches_the_oracle_bit_for_bit(la):
    "The product the benchmark times, FP64 8192 x 8192 x 8192, equals the oracle's in-order fold item for item."
    n = 8192
    a = random(n, n, 0, np.float64); b = random(n, n, 1, np.float64)
    assert bits_equal(la.matmul(a, b), oracle.matmul(a, b))


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_vector_products_follow_by_vector(la, dtype):
    "`by_vector` (core/src/linalg.rs:1020): matrix times vector in the reference's summation orders, bit for bit."
    rng = np.random.default_rng(3)
    for m, k in [(9, 5), (33, 100), (1000, 1003), (7, 8)]:
        a = rng.standard_normal((m, k)).astype(dtype)
        v = rng.standard_normal(k).astype(dtype)
        w = rng.standard_normal(m).astype(dtype)
        assert bits_equal(la.matmul(a, v), oracle.matmul(a, v))
        assert bits_equal(la.matmul(w, a), oracle.matmul(w, a))


def test_vector_products_take_the_known_answer_orders(la):
    """The oracle's known answers for `by_vector`'s orders (`dots`, its `dot` tail rows and tail items,
    `narrow_dots`, `narrow_sums` with and without whole quads, `scaled_runs`): each differs from an in-order fold."""
    B = 2.0 ** 53
    r8 = [1.0, 0.0, -B, 0.0, B, 0.0, 0.0, 0.0]
    stored = lambda n: np.tile(np.array([B, 1.0, -B, 1.0])[:, None], (1, n))
    cases = [(la.matmul(np.tile(r8 + [0.0] * 8, (5, 1)), np.ones(16)), [0.0, 0.0, 0.0, 0.0, 1.0]),
             (la.matmul(np.tile(r8 + [0.0] * 8 + [1.0], (4, 1)), np.ones(17)), [1.0] * 4),
             (la.matmul(np.tile(r8, (3, 1)), np.ones(8)), [1.0] * 3),
             (la.matmul(np.tile([B, 1.0, 1.0, -B], (3, 1)), np.ones(4)), [1.0] * 3),
             (la.matmul(np.ones(4), stored(8)), [2.0] * 8),
             (hs_gemm(stored(8), np.ones((4, 1)), opa='T')[:, 0], [2.0] * 8),
             (hs_gemm(stored(5), np.ones((4, 1)), opa='T')[:, 0], [1.0] * 5),
             (la.matmul(np.ones(4), stored(20)), [1.0] * 20)]
    for i, (got, want) in enumerate(cases):
        assert bits_equal(got, np.array(want)), (i, got)


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_transposed_matrix_times_vector_follows_narrow_sums(dtype):
    """A stored k x m matrix read transposed times a vector: `narrow_sums` for m <= 8, whose quads stop at the last
    stored row with 8 items left in the data (core/src/linalg.rs:943), and `scaled_runs` past 8 results."""
    rng = np.random.default_rng(11)
    for m in [1, 3, 5, 8, 9, 20]:
        for k in [1, 3, 4, 7, 9, 37]:
            a = rng.standard_normal((k, m)).astype(dtype)
            x = rng.standard_normal((k, 1)).astype(dtype)
            c0 = rng.standard_normal((m, 1)).astype(dtype)
            for alpha, beta in [(1.0, 0.0), (-2.0, 0.5)]:
                got = hs_gemm(a, x, alpha, beta, c0.copy(), 'T', 'N')
                want = oracle.gemm(a, x, alpha, beta, c0.copy(), 'T', 'N')
                assert bits_equal(got, want), (m, k, alpha, beta)


def test_notebook_known_answers(la):
    "nbs/linalg.ipynb cells 4, 6, 25, 27 and README.md:79-82 on the GPU, signed zeros included."
    a = np.array([[1., 2], [2, 5]])
    f = la.lu(a)
    assert f.p.tolist() == [1, 0]
    assert bits_equal(f.l, np.array([[1., 0], [0.5, 1]])) and bits_equal(f.u, np.array([[2., 5], [0, -0.5]]))
    assert bits_equal(f.solve(np.array([[1., 2], [2, 1]])), np.array([[1., 8], [-0., -3]]))
    assert f.determinant() == 1.0 and f.log_determinant() == (1.0, 0.0)
    rhs = np.array([[1., 2], [2, 1]])
    stack = np.stack([a, 2 * a, 3 * a])
    x = la.solve(stack, rhs)
    want = np.array([[[1.0, 8.0], [-0.0, -3.0]], [[0.5, 4.0], [-0.0, -1.5]], [[0.33333333, 2.6666667], [-0.0, -1.0]]])
    assert np.allclose(x, want, rtol=1e-7, atol=0) and (np.signbit(x) == np.signbit(want)).all()
    v = la.solve(stack, np.array([[1., 2], [2, 1], [1, 1]])[..., None])[..., 0]
    wv = np.array([[1.0, -0.0], [4.0, -1.5], [1.0, -0.33333333]])
    assert np.allclose(v, wv, rtol=1e-7, atol=0) and (np.signbit(v) == np.signbit(wv)).all()
    assert np.allclose(la.solve(np.array([[4., 1], [1, 3]]), np.array([1., 2])), [0.090909091, 0.63636364], rtol=1e-8)


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_subnormal_pivot(la, dtype):
    "tests/linalg.rs:510-511: lu([p, 1; q, 1]).factors == [p, q/p, 1, 1 - q/p]."
    tiny = np.finfo(dtype).tiny
    p, q = dtype(tiny / 1024), dtype(tiny / 1048576)
    f = la.lu(np.array([[p, 1], [q, 1]], dtype))
    assert bits_equal(f.factors.reshape(-1), np.array([p, q / p, 1, dtype(1) - q / p], dtype))


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
@pytest.mark.parametrize('n', [1, 2, 5, 16, 17, 31, 32, 33, 63, 64, 65, 71, 100, 129, 257, 530, 1030])
def test_lu_matches_the_oracle_bit_for_bit(la, dtype, n):
    "Pivots, factors and perm equal the oracle's (rustymath's lu_factor arithmetic) item for item."
    a = random(n, n, n + 1, dtype)
    f = la.lu(a)
    want, perm, _ = oracle.lu(a)
    assert (f.p == perm).all()
    assert bits_equal(f.factors, want), f'max diff {np.abs(f.factors - want).max()}'


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
@pytest.mark.parametrize('n', [2100, 4100])
def test_lu_with_look_ahead_matches_the_oracle_bit_for_bit(la, dtype, n):
    """Sizes with many 256-column blocks: the factorization runs each block's trailing product beside the next block's
    panels (tiles taken by the panel launches' workers, the rest afterwards) and moves rows of the other columns late;
    pivots and factors still equal the oracle's item for item."""
    a = random(n, n, n + 3, dtype)
    f = la.lu(a)
    want, perm, _ = oracle.lu(a)
    assert (f.p == perm).all()
    assert bits_equal(f.factors, want), f'max diff {np.abs(f.factors - want).max()}'


@pytest.mark.parametrize('dtype,n', [(np.float64, 8192), (np.float32, 16384)])
def test_lu_at_the_benchmark_orders_matches_the_oracle_bit_for_bit(la, dtype, n):
    """The factorizations the benchmark times: FP64 of order 8192 (the plain solve) and FP32 of order 16384 (the
    mixed solve's), with every block's look-ahead and the panel kernel at its full 32 workgroups; pivots and factors
    equal the oracle's item for item (the oracle takes about 10 to 30 s here), and so do the single-column solves with
    them (the refinement's)."""
    a = random(n, n, 11, dtype)
    f = la.lu(a)
    want, perm, _ = oracle.lu(a)
    assert (f.p == perm).all()
    assert bits_equal(f.factors, want), f'max diff {np.abs(f.factors - want).max()}'
    b = random(n, 1, 13, dtype)[:, 0]
    assert bits_equal(f.solve(b), oracle.lu_solve(want, perm, b))
    # Sixteen right-hand sides (the plain FP64 benchmark's): `backward`'s halving at full size (factor.rs:693-707).
    B = np.random.default_rng(1).standard_normal((n, 16)).astype(dtype)
    assert bits_equal(f.solve(B), oracle.lu_solve(want, perm, B))


def test_the_cpu_runnable_configuration_equals_the_oracle(la):
    """The reference's CPU-runnable configuration through the Python surface (py/src/linalg.rs:29-33, :241-247):
    matmul of two 1024×1024 float64 N(0, 1) matrices (numpy seed 0, a then b) and solve of a 1024×1024 N(0, 1) system
    with b 1024×1 (seed 1, A then b); bits equal to the oracle's, and the solve's backward error within 10·n·eps64."""
    n = 1024
    rng = np.random.default_rng(0); a = rng.standard_normal((n, n)); b = rng.standard_normal((n, n))
    assert bits_equal(la.matmul(a, b), oracle.matmul(a, b))
    rng = np.random.default_rng(1); A = rng.standard_normal((n, n)); v = rng.standard_normal((n, 1))
    x = la.solve(A, v)
    assert bits_equal(x, oracle.solve(A, v))
    assert bits_equal(la.solve(A, v[:, 0]), x[:, 0])
    assert backward_error(A, x, v) <= 10 * n * EPS[np.float64]


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_lu_edge_matrices_match_the_oracle(la, dtype):
    "Zero columns (a zero pivot skips only the rest of its reference leaf), NaN, infinities and a permutation matrix."
    rng = np.random.default_rng(11)
    cases = []
    z = rng.standard_normal((100, 100)); z[:, [3, 17, 50, 99]] = 0; cases.append(z)
    z = rng.standard_normal((70, 70)); z[:, 5] = 0; z[7, 5] = np.nan; cases.append(z)
    z = rng.standard_normal((70, 70)); z[3, 3] = np.inf; z[9, 40] = -np.inf; cases.append(z)
    z = np.zeros((40, 40)); z[np.arange(40), (np.arange(40) * 7) % 40] = 1; cases.append(z)
    cases.append(np.zeros((20, 20)))
    for a in cases:
        a = a.astype(dtype)
        f = la.lu(a)
        want, perm, _ = oracle.lu(a)
        assert (f.p == perm).all() and bits_equal(f.factors, want)


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
@pytest.mark.parametrize('n', [40, 600])
def test_pivot_row_after_a_zero_pivot_keeps_its_leaf(la, dtype, n):
    """A zero pivot in column 3 skips the rest of its leaf (factor.rs:606), so the next pivot row (row 4) keeps its
    items there, whatever row 3 holds: U[4, 5] stays 2 beside an infinity above it (an update would give 0 * inf = NaN)
    and U[4, 6] stays -0 (an update would give +0). Rows 1 and 2, pivots before the zero one, share row 4's wave on the
    GPU; the items that lanes holding such rows hand over must take the skip of the column, not of their own rows.
    n = 600 runs the panel over two workgroups, whose rows go through the exchange."""
    a = np.zeros((n, n))
    a[np.arange(n), np.arange(n)] = 4.0
    a[3, 3] = 0.0
    a[3, 4], a[3, 5], a[3, 6] = 1.0, np.inf, -2.0
    a[4, 5], a[4, 6] = 2.0, -0.0
    a[5:, 4] = 0.5
    a = a.astype(dtype)
    f = la.lu(a)
    want, perm, _ = oracle.lu(a)
    assert want[5, 4] == 2.0 and np.signbit(want[6, 4])
    assert (f.p == perm).all() and bits_equal(f.factors, want)


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_hard_matrices_scale_exactly(la, dtype):
    "tests/linalg.rs:466-483 on the GPU: power-of-two scaling scales U exactly and leaves perm, L and solves alone."
    n = 40
    fi = np.finfo(dtype)
    tiny = dtype(fi.tiny / fi.eps)
    quartic = dtype(2.0 ** (int(math.log2(float(fi.max)) / 4) + 6))
    a, b = random(n, n, 17, dtype), random(n, 3, 19, dtype)
    f = la.lu(a)
    for s in [tiny, dtype(1) / tiny, quartic]:
        g = la.lu(a * s)
        assert (g.p == f.p).all()
        scaled = f.factors.copy()
        iu = np.triu_indices(n)
        scaled[iu[1], iu[0]] = f.factors[iu[1], iu[0]] * s
        assert bits_equal(g.factors, scaled)
        assert bits_equal(g.solve(b * s), f.solve(b))


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
@pytest.mark.parametrize('n', [1, 5, 33, 71, 200, 530])
@pytest.mark.parametrize('k', [1, 3, 16])
def test_lu_solve_matches_the_oracle_bit_for_bit(la, dtype, n, k):
    """lu_solve's substitution orders (increasing for L, decreasing with a quotient for one column of U, the
    halving order for several) give the oracle's bits."""
    a = random(n, n, n + 3, dtype)
    b = random(n, k, 99, dtype)
    f, perm, _ = oracle.lu(a)
    got = la.lu_solve(f, perm, b)
    want = oracle.lu_solve(f, perm, b)
    assert bits_equal(got, want), f'max diff {np.abs(got - want).max()}'
    got1 = la.lu_solve(f, perm, b[:, 0])
    assert bits_equal(got1, oracle.lu_solve(f, perm, b[:, 0]))


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_single_column_solve_above_order_16384_matches_the_oracle(la, dtype):
    """Order 16500: more 32-row blocks than the single-column solve can keep resident in runs per XCD, so its blocks
    run in launch order. Synthetic factors (diagonal in [1, 2), other items N(0, 1)/n) and a random permutation,
    against the oracle bit for bit."""
    n = 16500
    rng = np.random.default_rng(21)
    f = rng.standard_normal((n, n)).astype(dtype)
    f *= dtype(1.0 / n)
    f[np.diag_indices(n)] = (1 + rng.random(n)).astype(dtype)
    perm = rng.permutation(n).astype(np.int64)
    b = rng.standard_normal(n).astype(dtype)
    got = la.lu_solve(f, perm, b)
    want = oracle.lu_solve(f, perm, b)
    assert np.isfinite(want).all()
    assert bits_equal(got, want), f'max diff {np.abs(got - want).max()}'


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
@pytest.mark.parametrize('n', [530, 1030])
def test_blocked_sizes_have_small_residuals(la, dtype, n):
    "tests/linalg.rs:569-590 on the GPU, plus the acceptance bound 10·n·eps on the solve's backward error."
    a = random(n, n, n, dtype); aw = a.astype(np.float64)
    b = random(n, 3, 5, dtype)
    x = la.solve(a, b).astype(np.float64)
    assert distance(aw @ x, b, norm(aw) * norm(x)) < 4 * EPS[dtype] * n
    assert backward_error(aw, x, b) <= 10 * n * EPS[dtype]


def test_singular_matrix_is_not_solved_by_lu(la):
    "factor.rs:221-226: a negligible pivot sends rustymath to least squares, which this library leaves out loudly."
    with pytest.raises(NotImplementedError):
        la.solve(np.array([[1., 2], [2, 4]]), np.array([1., 2]))


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
@pytest.mark.parametrize('k', [1, 3, 16])
def test_solves_with_subnormal_diagonals_match_the_oracle(dtype, k):
    """A matrix scaled into the subnormal range: U's diagonal items are subnormal, so every row of U takes a quotient
    rather than a reciprocal multiply (`divide`, factor.rs:506), in the block solves of both orders."""
    import torch
    from rustymath_amd import device as dv
    n = 150
    rng = np.random.default_rng(5 + k)
    scale = 2.0 ** (-1060 if dtype == np.float64 else -140)
    a = (rng.standard_normal((n, n)) * scale).astype(dtype)
    b = (rng.standard_normal((n, k)) * scale).astype(dtype)
    f, perm = dv.getrf(torch.from_numpy(a).cuda())
    wf, wperm, _ = oracle.lu(a)
    assert np.abs(np.diag(wf)).max() < np.finfo(dtype).tiny
    assert np.isfinite(oracle.lu_solve(wf, wperm, b)).all() and np.isfinite(oracle.lu_solve_t(wf, wperm, b)).all()
    x = dv.getrs(f, perm, torch.from_numpy(b).cuda()).cpu().numpy()
    assert bits_equal(x, oracle.lu_solve(wf, wperm, b))
    y = dv.getrs(f, perm, torch.from_numpy(b).cuda(), trans=True).cpu().numpy()
    assert bits_equal(y, oracle.lu_solve_t(wf, wperm, b))


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
@pytest.mark.parametrize('n', [1, 5, 64, 100, 129, 513, 1030])
@pytest.mark.parametrize('k', [1, 3, 16, 21, 70])
def test_solves_and_transposed_solves_match_the_oracle_at_block_sizes(dtype, n, k):
    """Blocked substitutions: 64-row block solves joined by in-order MFMA products, for A x = b (`lu_solve`'s orders,
    the halving of `backward` across nodes) and Aᵀ x = b (the adjoint's orders), against the oracle bit for bit; k past
    16 takes the products in column groups (or the tiled product), past 64 the block solves in column chunks."""
    import torch
    from rustymath_amd import device as dv
    rng = np.random.default_rng(n * 31 + k)
    a = rng.standard_normal((n, n)).astype(dtype)
    b = rng.standard_normal((n, k)).astype(dtype)
    f, perm = dv.getrf(torch.from_numpy(a).cuda())
    wf, wperm, _ = oracle.lu(a)
    x = dv.getrs(f, perm, torch.from_numpy(b).cuda()).cpu().numpy()
    assert bits_equal(x, oracle.lu_solve(wf, wperm, b))
    y = dv.getrs(f, perm, torch.from_numpy(b).cuda(), trans=True).cpu().numpy()
    assert bits_equal(y, oracle.lu_solve_t(wf, wperm, b))


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
@pytest.mark.parametrize('n', [1, 2, 5, 31, 32, 33, 64, 65, 97, 130, 257, 530])
def test_inverse_matches_the_oracle_bit_for_bit(la, dtype, n):
    """`Lu::inverse` (factor.rs:645-664): L⁻¹ by `forward`'s 32-row block mode (a last block of one row takes its
    product through `by_vector`, as at n = 33 and 65), U⁻¹ by `backward`'s halving over all columns, then the columns
    permuted back; the same bits as the oracle, and a small residual (tests/linalg.rs:389-390)."""
    a = random(n, n, 3 * n + 1, dtype)
    f = la.lu(a)
    want = oracle.lu_inverse(f.factors, f.p)
    got = f.inverse()
    assert got.dtype == dtype and bits_equal(got, want), f'max diff {np.abs(got - want).max()}'
    aw, inv = a.astype(np.float64), got.astype(np.float64)
    bound = 4 * EPS[dtype] * n
    assert distance(aw @ inv, np.eye(n), max(norm(aw) * norm(inv), 1.0)) < bound


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_device_inverse_and_pinv_agree_with_lu_inverse(la, dtype):
    """The device entry point gives the host one's bits; `pinv` of a regular matrix is the LU inverse
    (tests/linalg.rs:546-547) and `det` is the LU determinant; a singular matrix has no LU inverse."""
    import torch
    from rustymath_amd import device as dv
    n = 300
    a = random(n, n, 77, dtype) + np.eye(n, dtype=dtype)
    f = la.lu(a)
    inv = f.inverse()
    fd, perm = dv.getrf(torch.from_numpy(a).cuda())
    assert bits_equal(dv.getri(fd, perm).cpu().numpy(), inv)
    assert bits_equal(la.pinv(a), inv)
    assert la.det(a) == f.determinant()
    stack = np.stack([a, 2 * a])
    assert bits_equal(la.pinv(stack)[1], la.lu(2 * a).inverse())
    with pytest.raises(NotImplementedError): la.pinv(np.array([[1., 2], [2, 4]], dtype))


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_inverse_scales_exactly(la, dtype):
    "tests/linalg.rs:477-483: lu(a s).inverse() s == lu(a).inverse() bit for bit, for powers of two s."
    n = 70
    fi = np.finfo(dtype)
    a = random(n, n, 23, dtype)
    f = la.lu(a).inverse()
    for s in [dtype(2.0 ** -20), dtype(2.0 ** 20), dtype(fi.tiny / fi.eps)]:
        assert bits_equal(la.lu(a * s).inverse() * s, f)


def test_inverse_and_determinant_known_answers(la):
    """nbs/linalg.ipynb: inverse [[5, -2], [-2, 1]] of [[1, 2], [2, 5]] with determinant 1; tests/linalg.rs:553-558:
    diag(1, 2, 3, 4) with two rows swapped has determinant -24; a zero column gives 0 and a log of -inf;
    tests/test_linalg.py:6-32's stack of 64 fixtures against NumPy, float16 computing in float32."""
    a = np.array([[1., 2], [2, 5]])
    assert bits_equal(la.lu(a).inverse(), np.array([[5., -2], [-2, 1]])) and la.det(a) == 1.0
    d = np.diag([1., 2, 3, 4])[[1, 0, 2, 3]]
    sign, log = la.slogdet(d)
    assert la.det(d) == -24.0 and sign == -1.0 and abs(log - math.log(24.0)) < 1e-14
    assert la.det(np.array([[1., 0], [2, 0]])) == 0.0
    assert la.slogdet(np.array([[1., 2], [2, 4]])) == (0.0, -np.inf)
    base = np.array([[0., 2, 1], [3, 1, -1], [1, 0, 4]])
    wide = np.stack([base + k / 8 * np.eye(3) for k in range(64)])
    for dt in (np.float16, np.float32, np.float64):
        f = la.lu(wide.astype(dt))
        assert np.allclose(f.inverse(), np.linalg.inv(wide), rtol=0.003, atol=0.003)
        assert np.allclose(f.determinant(), np.linalg.det(wide), rtol=0.003, atol=0.003)
        assert np.allclose(la.pinv(wide.astype(dt)), np.linalg.inv(wide), rtol=0.003, atol=0.003)
        sign, log = la.slogdet(wide.astype(dt))
        want = np.linalg.slogdet(wide)
        assert np.allclose(sign, want.sign, rtol=0.003, atol=0.003) and np.allclose(log, want.logabsdet, rtol=0.003, atol=0.003)
    h = la.det(np.eye(3, dtype=np.float16) * 2)
    assert np.asarray(h).dtype == np.float16 and h == 8


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
@pytest.mark.parametrize('n,k', [(1, 1), (17, 3), (40, 8), (100, 1), (512, 2), (530, 2), (33, 9)])
def test_stacks_match_each_cell(la, dtype, n, k):
    """`Stacked::solve` and `Stacked::lu` (src/linalg/stacked.rs:88-101, :174-187): a stack goes to the batched kernel
    in one launch where its cells allow (n <= 512, k <= 8) and cell by cell otherwise; either way every cell is the
    oracle's `solve` and `lu_factor` of that cell, bit for bit."""
    rng = np.random.default_rng(n * 10 + k)
    count = 5
    a = rng.standard_normal((count, n, n)).astype(dtype) + 2 * np.eye(n, dtype=dtype)
    b = rng.standard_normal((count, n, k)).astype(dtype)
    x = la.solve(a, b)
    f = la.lu(a)
    for i in range(count):
        wf, wperm, _ = oracle.lu(a[i])
        assert (f.p[i] == wperm).all() and bits_equal(f.factors[i], wf), f'cell {i} factors'
        assert bits_equal(x[i], oracle.solve(a[i], b[i])), f'cell {i} solution'
    assert bits_equal(f.solve(b), x)


def test_a_stack_with_a_negligible_pivot_is_left_to_least_squares(la):
    a = np.stack([np.eye(4), np.array([[1., 2, 0, 0], [2, 4, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]])])
    with pytest.raises(NotImplementedError): la.solve(a, np.ones((2, 4, 1)))
    assert la.solve(np.zeros((3, 0, 0)), np.zeros((3, 0, 2))).shape == (3, 0, 2)


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
@pytest.mark.parametrize('n', [1, 2, 5, 31, 32, 33, 63, 64, 65, 71, 100, 129, 300, 530])
def test_cholesky_matches_the_oracle_bit_for_bit(la, dtype, n):
    """`cholesky_factor` (factor.rs:777-797): every item of L in the reference's order (its terms in order of p, then
    the square root or `divide`), whatever the blocking: the oracle's bits at sizes either side of both blockings, zeros
    above the diagonal, and only the lower triangle read (tests/linalg.rs:391-396)."""
    from test_oracle import spd_of
    spd = spd_of(n, dtype)
    want = oracle.cholesky(spd)
    junk = spd.copy(); junk[np.triu_indices(n, 1)] = 7.0
    got = la.cholesky(junk)
    assert got.dtype == dtype and bits_equal(got, want), f'max diff {np.abs(got - want).max()}'


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_cholesky_failures_scaling_and_device_form(la, dtype):
    """A matrix that is not positive definite gives NaN (tests/test_linalg.py:180-181, :228-233, cell by cell in a
    stack); scaling by an even power of two scales L exactly (tests/linalg.rs:484-486); the device entry point gives
    the same bits; NumPy agrees to rounding (tests/test_linalg.py:109-110)."""
    import torch
    from rustymath_amd import device as dv
    from test_oracle import spd_of
    assert np.isnan(la.cholesky(-np.eye(2, dtype=dtype))).all()
    a = np.broadcast_to(np.eye(3, dtype=dtype), (2, 8, 3, 3)).copy()
    a[0, -1, 0, 0] = -1
    l = la.cholesky(a)
    assert np.isnan(l[0, -1]).all() and np.isfinite(l[1]).all() and np.isfinite(l[0, :-1]).all()
    n = 150
    spd = spd_of(n, dtype)
    l = la.cholesky(spd)
    fi = np.finfo(dtype)
    for s in [dtype(fi.tiny / fi.eps), dtype(fi.eps / fi.tiny)]:
        c = s if np.sqrt(s) * np.sqrt(s) == s else s + s
        assert bits_equal(la.cholesky(spd * c), l * np.sqrt(c))
    t, ok = dv.potrf(torch.from_numpy(spd.copy()).cuda())
    assert ok and bits_equal(t.cpu().numpy(), l)
    _, ok = dv.potrf(torch.from_numpy(-spd).cuda())
    assert not ok
    sw = spd.astype(np.float64)
    assert np.allclose(l, np.linalg.cholesky(sw), rtol=100 * fi.eps * n, atol=1e-30)
    h = la.cholesky(np.array([[4., 2], [2, 5]], np.float16))
    assert h.dtype == np.float16 and (h == np.array([[2, 0], [1, 2]], np.float16)).all()


@pytest.mark.parametrize('k', [1, 3])
def test_float_fused_multiply_adds_round_once(la, k):
    """Every FP32 fused multiply-add rounds once, as rustymath's: the case of test_oracle's
    test_float_products_round_once, where a product through double would round twice, through the single-column
    and the block substitutions."""
    f = np.float32
    a = np.array([[1, 0], [-(0.5 + 2.0 ** -13), 1]], f)
    b = np.repeat(np.array([[2 + 2.0 ** -11], [2.0 ** -60]], f), k, axis=1)
    x = la.solve(a, b)
    assert (x[1] == f(1 + 2.0 ** -11 + 2.0 ** -23)).all()
    fac = la.lu(a)
    assert bits_equal(fac.solve(b), oracle.lu_solve(*oracle.lu(a)[:2], b))


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_profiles_account_for_a_factorization(la, dtype):
    """The live profiles `bench.py` reads (trailing tiles, product launches, panel phases) count a factorization of
    order 2100 within its flops, and profiling leaves the factors the oracle's."""
    import ctypes
    import torch
    from rustymath_amd import _lib, device as dv
    lib = _lib.load()
    n = 2100
    a = random(n, n, 7, dtype)
    v = (ctypes.c_double * 8)()
    g = [ctypes.c_double(0), ctypes.c_int64(0), ctypes.c_double(0), ctypes.c_double(0)]
    ticks = (ctypes.c_int64 * 8)()
    assert lib.hs_profile_lu(1, None) == 0
    assert lib.hs_profile_gemm(1, None, None, None, None) == 0
    assert lib.hs_profile_panel(1, None) == 0
    f, perm = dv.getrf(torch.from_numpy(a.copy()).cuda())
    torch.cuda.synchronize()
    assert lib.hs_profile_lu(0, v) == 0
    assert lib.hs_profile_gemm(0, *[ctypes.byref(x) for x in g]) == 0
    assert lib.hs_profile_panel(0, ticks) == 0
    trail_ms, trail_n, trail_fl, trail_by, panel_ms, panel_n, panel_fl, worker_fl = list(v)
    # one trail_rest per block (at most 256 columns of whole panels) with a product beside the next block's panels
    assert 0 < trail_n < n // 128 and trail_ms > 0 and (trail_by > 0 or trail_fl == 0)
    # every panel launch timed: the panels are whole reference leaves of at most 64 columns
    assert n // 64 <= panel_n <= n // 16 and panel_ms > 0 and panel_fl > 0
    assert g[0].value > 0 and g[1].value > 0 and sum(ticks) > 0
    # the trailing tiles (workers' and trail_rest's), the product launches and the panels' own work stay within the
    # factorization's flops (the row-of-U solves are not counted)
    assert 0 <= trail_fl and 0 <= worker_fl
    assert trail_fl + worker_fl + g[2].value + panel_fl <= 2 / 3 * n ** 3
    want, wperm, _ = oracle.lu(a)
    assert bits_equal(f.cpu().numpy(), want) and (perm.cpu().numpy() == wperm).all()


def test_matrix_layouts_are_read_in_place(la):
    """tests/test_linalg.py:207-227 on the GPU: reversed, transposed and strided views, and stacks beside a single
    matrix. Views that are positive-strided with one unit stride are read where they lie (as `Mapped::matrix` reads
    them, core/src/linalg.rs:658-670), others through a compact copy; the products are NumPy's within rounding, and
    where both extents exceed 1 (the MFMA product, whose every item is the in-order fold whatever the layout) the
    oracle's product of compact copies bit for bit, cell by cell."""
    rng = np.random.default_rng(4)
    a = rng.normal(size=(64, 8, 8))[:, ::-1, ::-1]
    b = rng.normal(size=(8, 5))[::-1]
    got = la.matmul(a, b)
    assert np.allclose(got, a @ b)
    for i in range(64):
        assert bits_equal(got[i], oracle.matmul(np.ascontiguousarray(a[i]), np.ascontiguousarray(b)))
    s = rng.normal(size=(3, 8, 5))
    for dtype in (np.float32, np.float64):
        m, st = b.astype(dtype), s.astype(dtype)
        for view in (m.T, m[::-1], m[:, ::2]):
            dense = np.ascontiguousarray(view)
            p = la.matmul(view.T, view)
            assert np.allclose(p, dense.T @ dense, rtol=1e-5, atol=1e-6)
            assert bits_equal(p, oracle.matmul(np.ascontiguousarray(dense.T), dense))
        for p, q in [(la.matmul(m[:, 1], m), b[:, 1] @ b), (la.matmul(m.T, m[:, 1]), b.T @ b[:, 1]),
                     (la.matmul(st.swapaxes(-1, -2), st), s.swapaxes(-1, -2) @ s),
                     (la.matmul(st[:, ::-1], m.T), s[:, ::-1] @ b.T)]:
            assert np.allclose(p, q, rtol=1e-5, atol=1e-5)
        stack = la.matmul(st.swapaxes(-1, -2), st)
        for i in range(3):
            assert bits_equal(stack[i], oracle.matmul(np.ascontiguousarray(st[i].T), np.ascontiguousarray(st[i])))
    for x in [a[:1], a[:3], np.eye(128)[None]]:  # tests/test_linalg.py:205-206
        assert np.allclose(la.matmul(x, x), x @ x)


def test_batched_products_equal_the_single_products():
    """hs_dgemm_batched / hs_sgemm_batched: every cell of a stack, a repeated operand (stride 0) and transposed cells
    give the bits of hs_dgemm on that cell."""
    import ctypes
    from rustymath_amd import _lib
    rng = np.random.default_rng(8)
    p = lambda x: x.ctypes.data_as(ctypes.c_void_p)
    for dtype, f in ((np.float64, 'hs_dgemm_batched'), (np.float32, 'hs_sgemm_batched')):
        for m, k, n, ops in ((37, 130, 65, 'NN'), (128, 64, 1, 'TN'), (1, 50, 70, 'NT'), (200, 33, 129, 'TT')):
            cnt = 5
            a = rng.standard_normal((cnt,) + ((m, k) if ops[0] == 'N' else (k, m))).astype(dtype)
            b = rng.standard_normal((k, n) if ops[1] == 'N' else (n, k)).astype(dtype)
            c0 = rng.standard_normal((cnt, m, n)).astype(dtype)
            c = c0.copy()
            _lib.check(getattr(_lib.load(), f)(oracle.OPS[ops[0]], oracle.OPS[ops[1]], m, n, k, 0.5, p(a), a.shape[2],
                                               a.shape[1] * a.shape[2], p(b), b.shape[1], 0, -1.0, p(c), n, m * n, cnt),
                       'gemm_batched')
            for i in range(cnt):
                want = hs_gemm(a[i], b, 0.5, -1.0, c0[i].copy(), ops[0], ops[1])
                assert bits_equal(c[i], want), (dtype, m, k, n, ops, i)


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_lu_solve_on_stacks_matches_each_cell(la, dtype):
    """`Stacked::lu_solve` (src/linalg/stacked.rs:189-204): a stack of factors and positions with a stack of right-hand
    sides (one batched launch up to order 512 and 8 columns, cell by cell beyond, and for positions that are not a
    permutation), each cell the oracle's lu_solve bit for bit; a vector right-hand side broadcast over the stack."""
    rng = np.random.default_rng(12)
    for n, k, count in ((40, 3, 7), (200, 8, 4), (513, 2, 2), (64, 9, 3)):
        a = rng.standard_normal((count, n, n)).astype(dtype)
        b = rng.standard_normal((count, n, k)).astype(dtype)
        f = la.lu(a)
        x = la.lu_solve(f.factors, f.p, b)
        for i in range(count):
            assert bits_equal(x[i], oracle.lu_solve(f.factors[i], f.p[i], b[i])), (n, k, i)
    v = rng.standard_normal(40).astype(dtype)
    a = rng.standard_normal((5, 40, 40)).astype(dtype)
    f = la.lu(a)
    xv = la.lu_solve(f.factors, f.p, v)
    for i in range(5):
        assert bits_equal(xv[i], oracle.lu_solve(f.factors[i], f.p[i], v))
    p = f.p.copy(); p[:, 1] = p[:, 0]  # positions in range but not a permutation: the cell-by-cell path
    xp = la.lu_solve(f.factors, p, v)
    for i in range(5):
        assert bits_equal(xp[i], oracle.lu_solve(f.factors[i], p[i], v))


def test_lu_above_the_panel_kernels_height_matches_the_oracle(la):
    """Order 16448, above the 32 × 512 rows the panel kernel holds: the panels are factored in memory column by column
    (pivot search in blocks, pick and swap, multipliers and their update), without look-ahead. Pivots and factors equal
    the oracle's item for item, two zero columns included (one inside a leaf, one in the last panel: each leaves its
    column undivided and skips the rest of its leaf, factor.rs:606). The reference's lu_factor has no order limit
    (core/src/linalg/factor.rs:617-630)."""
    n = 16448
    a = random(n, n, 23, np.float64)
    a[:, 5] = 0.0
    a[:, 16400] = 0.0
    print(f'tall LU n={n}: GPU', flush=True)
    f = la.lu(a)
    print('tall LU: oracle', flush=True)
    want, perm, _ = oracle.lu(a)
    assert (f.p == perm).all()
    assert bits_equal(f.factors, want), f'max diff {np.nanmax(np.abs(f.factors - want))}'
