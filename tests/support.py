"""Shared test inputs, restating the reference's own test fixtures.

`random` restates tests/support/mod.rs:5-10 (SplitMix64's output at index i), `seeded` (:13) and the matrices of
`mod factor` in tests/linalg.rs:332-335: uniform numbers from -1 to 1, `2 u - 1` with `u` from rand's
`StandardUniform` for f64 (the top 53 bits of one 64-bit draw times 2^-53), narrowed to the item type.
"""
import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix(i):
    "SplitMix64's output at each index in the uint64 array `i` (tests/support/mod.rs:5-10)."
    with np.errstate(over='ignore'):
        z = (np.asarray(i, np.uint64) + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def random(m, n, seed, dtype=np.float64):
    "An m×n matrix of numbers from -1 to 1, as `random::<T>(m, n, seed)` in tests/linalg.rs:332-335."
    draws = splitmix(np.arange(seed, seed + m * n, dtype=np.uint64))
    u = (draws >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
    return (2.0 * u - 1.0).astype(dtype).reshape(m, n)


def integer_operands(m, k, n):
    "The exact small-integer operands of `products_match_folds_in_order` (tests/linalg.rs:165-166)."
    a = (np.arange(m * k) * 7 % 17 - 8).astype(np.float64).reshape(m, k)
    b = (np.arange(k * n) * 5 % 13 - 6).astype(np.float64).reshape(k, n)
    return a, b


# KC, MC and NC of core/src/linalg.rs:389-395, and the sizes of tests/linalg.rs:163, which leave partial tiles,
# give no steps or a single one, and cross each block size.
KC, MC, NC = 256, 128, 2048
FOLD_SIZES = [(0, 3, 4), (3, 0, 4), (3, 4, 0), (1, 1, 1), (5, 7, 9), (MC + 3, KC + 5, 37), (MC + 3, KC + 5, 1), (17, 3, NC + 5)]


def norm(x):
    "`norm` of tests/linalg.rs:344: the square root of the sum of squares, in float64."
    x = np.asarray(x, np.float64)
    return float(np.sqrt((x * x).sum()))


def distance(x, y, scale):
    "`distance` of tests/linalg.rs:347: the norm of x - y relative to `scale`."
    return norm(np.asarray(x, np.float64) - np.asarray(y, np.float64)) / scale


def backward_error(a, x, b):
    "The normwise backward error ||b - A x|| / (||A|| ||x||), Frobenius and 2-norms, computed in float64."
    a, x, b = (np.asarray(v, np.float64) for v in (a, x, b))
    r = b - a @ x
    return norm(r) / max(norm(a) * norm(x), np.finfo(np.float64).tiny)


# The shapes of the triangular-solve checks, (side, m, n, ldx): left solves that halve with NS = 32 (n < m) and with
# LEAF = 16 (n >= m) leaves, one contiguous column (`single`), one strided column (the halving with `by_vector`
# products), and right-side solves, whose transposed copy is one contiguous column when m = 1.
TRSM_SHAPES = [('L', 70, 3, None), ('L', 40, 45, None), ('L', 70, 1, None), ('L', 70, 1, 4), ('L', 5, 2, None),
               ('R', 3, 70, None), ('R', 1, 40, None), ('R', 45, 40, 50), ('L', 1, 1, None), ('L', 33, 1, None)]
TRSM_CODES = [(u, t, d) for u in 'UL' for t in 'NT' for d in 'NU']


def triangle_system(side, m, n, uplo, trans, diag, dtype, seed):
    """(a, x_true, b) with op(a) x_true = b (side 'L') or x_true op(a) = b ('R') holding exactly: whole numbers from -2
    to 2, diagonal items ±1, ±2 or ±4 (ignored for a unit diagonal, which holds 7s instead), and NaN outside the
    triangle that `uplo` names, which a solve must never read. Every partial sum is a whole number below 2^10 and every
    division is by a power of two, so the solve is exact in any order of its terms."""
    rng = np.random.default_rng(seed)
    order = m if side == 'L' else n
    t = rng.integers(-2, 3, size=(order, order)).astype(np.float64)
    d = rng.choice([-4.0, -2.0, -1.0, 1.0, 2.0, 4.0], size=order)
    tri = np.tril(t, -1) if uplo == 'L' else np.triu(t, 1)
    logical_diag = np.ones(order) if diag == 'U' else d
    full = tri + np.diag(logical_diag)
    strict = np.tril(np.ones((order, order), bool), -1) if uplo == 'L' else np.triu(np.ones((order, order), bool), 1)
    a = np.where(strict, tri, np.nan)
    a[np.diag_indices(order)] = 7.0 if diag == 'U' else d
    op = full.T if trans != 'N' else full
    x = rng.integers(-2, 3, size=(m, n)).astype(np.float64)
    b = op @ x if side == 'L' else x @ op
    return a.astype(dtype), x.astype(dtype), b.astype(dtype)
