"""The solve that allocates nothing (hs_dev_dgetrs_noalloc) against the allocating one (hs_dev_dgetrs).

Both take device factors from the library's own LU and solve in place on device right-hand sides; the no-allocation
form keeps its workspace in a static array of the library on each device. Two calls in a row, the second smaller than
the first, in both orientations and with one and several columns, must give hs_dev_dgetrs's bits.
"""
import numpy as np
import pytest

from test_oracle import bits_equal

pytestmark = pytest.mark.gpu


def solve_with(name, factors, perm, b, trans):
    import torch
    from rustymath_amd import _lib
    n, k = b.shape
    out = b.clone()
    f = getattr(_lib.load(), name)
    _lib.check(f(1 if trans else 0, n, k, factors.data_ptr(), factors.stride(0), perm.data_ptr(), out.data_ptr(),
                 out.stride(0), torch.cuda.current_stream().cuda_stream), name)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_no_alloc_solve_equals_the_allocating_one():
    import torch
    from rustymath_amd import device as dv
    torch.cuda.set_device(0)
    rng = np.random.default_rng(5)
    for n, k in ((3000, 16), (700, 5), (2500, 1), (333, 1)):  # each call smaller than the one before it in its kind
        print(f'no-alloc solve n={n} k={k}', flush=True)
        a = torch.from_numpy(rng.standard_normal((n, n))).cuda()
        factors, perm = dv.getrf(a)
        b = torch.from_numpy(rng.standard_normal((n, k))).cuda()
        for trans in (False, True):
            want = solve_with('hs_dev_dgetrs', factors, perm, b, trans)
            got = solve_with('hs_dev_dgetrs_noalloc', factors, perm, b, trans)
            assert bits_equal(got, want), (n, k, trans)


def test_no_alloc_solve_rejects_sizes_beyond_its_workspace():
    from rustymath_amd import _lib
    lib = _lib.load()
    assert lib.hs_dev_dgetrs_noalloc(0, 65537, 1, None, 65537, None, None, 1, None) == -2
    assert lib.hs_dev_dgetrs_noalloc(0, 100, 17, None, 100, None, None, 17, None) == -3
    assert lib.hs_dev_dgetrs_noalloc(0, 20000, 16, None, 20000, None, None, 16, None) == -3
    assert lib.hs_dev_dgetrs_noalloc(0, 0, 3, None, 1, None, None, 3, None) == 0
