"""GPU parity of the triangular-solve entry points (hs_dtrsm / hs_strsm and their device forms) against the oracle.

`Tiled::triangular` (core/src/linalg.rs:781-787) is one of the three sites of rustymath's `dispatch!` seam; its native
form (`triangular_native`, core/src/linalg/factor.rs:711-721) fixes the orders of operations. The bar is the oracle's
bits for every code (side, stored triangle, transpose, unit diagonal) on random inputs, with NaN outside the named
triangle that no solve may read, and exactness on the whole-number systems that pin the oracle.
"""
import ctypes
import numpy as np
import pytest

import oracle
from support import TRSM_CODES, TRSM_SHAPES, triangle_system
from test_oracle import bits_equal

pytestmark = pytest.mark.gpu

BIG_SHAPES = [('L', 300, 7, None), ('L', 257, 1, None), ('L', 200, 1, 3), ('R', 5, 300, None), ('R', 1, 500, None),
              ('L', 130, 130, None), ('L', 1000, 1, None), ('L', 2100, 1, None), ('R', 64, 100, 128)]


def random_triangle(order, uplo, diag, dtype, rng):
    "A well-conditioned triangle (diagonal items 2 to 3 in magnitude) with NaN outside it and 7s on a unit diagonal."
    a = (rng.standard_normal((order, order)) / np.sqrt(max(order, 1))).astype(dtype)
    a[np.diag_indices(order)] = (rng.uniform(2, 3, order) * rng.choice([-1, 1], order)).astype(dtype)
    keep = np.tril(np.ones((order, order), bool)) if uplo == 'L' else np.triu(np.ones((order, order), bool))
    a = np.where(keep, a, np.nan).astype(dtype)
    if diag == 'U': a[np.diag_indices(order)] = 7
    return a


def hs_trsm(a, b, side, uplo, trans, diag, ldx=None):
    "hs_dtrsm / hs_strsm on host buffers, x laid out with leading dimension ldx (>= n)."
    from rustymath_amd import _lib
    m, n = b.shape
    ld = n if ldx is None else ldx
    buf = np.zeros(max((m - 1) * ld + n, 1), b.dtype)
    for i in range(m): buf[i * ld:i * ld + n] = b[i]
    a = np.ascontiguousarray(a)
    f = getattr(_lib.load(), 'hs_dtrsm' if b.dtype == np.float64 else 'hs_strsm')
    p = lambda x: x.ctypes.data_as(ctypes.c_void_p)
    _lib.check(f(oracle.SIDES[side], oracle.UPLOS[uplo], oracle.OPS[trans], oracle.DIAGS[diag], m, n, p(a),
                 max(a.shape[1], 1), p(buf), max(ld, 1)), 'trsm')
    return np.array([buf[i * ld:i * ld + n] for i in range(m)], b.dtype).reshape(m, n)


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
@pytest.mark.parametrize('shape', TRSM_SHAPES + BIG_SHAPES)
def test_triangular_solves_match_the_oracle_bit_for_bit(dtype, shape):
    side, m, n, ldx = shape
    print(f'trsm {np.dtype(dtype).name} {shape}', flush=True)
    rng = np.random.default_rng(m * 1000 + n)
    for uplo, trans, diag in TRSM_CODES:
        a = random_triangle(m if side == 'L' else n, uplo, diag, dtype, rng)
        b = rng.standard_normal((m, n)).astype(dtype)
        got = hs_trsm(a, b, side, uplo, trans, diag, ldx)
        want = oracle.trsm(a, b, side, uplo, trans, diag, ldx=ldx)
        assert bits_equal(got, want), (uplo, trans, diag, np.nanmax(np.abs(got - want)))


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
@pytest.mark.parametrize('shape', TRSM_SHAPES)
def test_triangular_solves_are_exact_on_whole_numbers(dtype, shape):
    side, m, n, ldx = shape
    for k, (uplo, trans, diag) in enumerate(TRSM_CODES):
        a, x, b = triangle_system(side, m, n, uplo, trans, diag, dtype, seed=k + 10 * m + n)
        assert np.array_equal(hs_trsm(a, b, side, uplo, trans, diag, ldx), x), (uplo, trans, diag)


def test_device_triangular_solve_matches_the_host_one():
    "hs_dev_dtrsm on device memory and a torch stream gives the host entry point's bits, strided x included."
    import torch
    from rustymath_amd import _lib
    rng = np.random.default_rng(3)
    for side, m, n, ldx in (('L', 150, 4, 9), ('R', 20, 90, None), ('L', 300, 1, None)):
        for uplo, trans, diag in (('U', 'T', 'U'), ('L', 'T', 'N'), ('L', 'N', 'N'), ('U', 'N', 'U')):
            a = random_triangle(m if side == 'L' else n, uplo, diag, np.float64, rng)
            b = rng.standard_normal((m, n))
            ld = n if ldx is None else ldx
            xd = torch.zeros((m, ld), dtype=torch.float64, device='cuda')
            xd[:, :n] = torch.from_numpy(b).cuda()
            ad = torch.from_numpy(a).cuda()
            st = _lib.load().hs_dev_dtrsm(oracle.SIDES[side], oracle.UPLOS[uplo], oracle.OPS[trans], oracle.DIAGS[diag],
                                          m, n, ad.data_ptr(), a.shape[1], xd.data_ptr(), ld,
                                          torch.cuda.current_stream().cuda_stream)
            _lib.check(st, 'dev_trsm')
            assert bits_equal(xd[:, :n].cpu().numpy(), hs_trsm(a, b, side, uplo, trans, diag, ldx))
