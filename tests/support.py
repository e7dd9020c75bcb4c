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
