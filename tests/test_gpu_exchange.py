"""The LU panels' and single-column substitutions' exchange at device scope (hs_set_exchange, mode 1).

By default their workgroups exchange pivots and rows through one XCD's L2, which needs launch indices 0, 8, 16, ...
on one XCD; a probe checks that placement on each device. Where it fails, or where the caller asks, every word goes
out at device scope instead. The arithmetic does not change with the mode: factors, pivots and solutions are bit
for bit the oracle's and the default mode's.
"""
import ctypes

import numpy as np
import pytest

import oracle
from support import random
from test_oracle import bits_equal

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def lib():
    import torch
    torch.cuda.set_device(0)
    from rustymath_amd import _lib
    lib = _lib.load()
    yield lib
    lib.hs_set_exchange(0, None)


def set_mode(lib, mode):
    local = ctypes.c_int(-1)
    assert lib.hs_set_exchange(mode, ctypes.byref(local)) == 0
    return local.value


def test_probe_finds_round_robin_placement(lib):
    "On one MI355X in its default partition, launch index i runs on XCD i mod 8: the exchanges use one L2."
    assert set_mode(lib, 0) == 1
    assert set_mode(lib, 1) == 0
    assert set_mode(lib, 0) == 1
    assert lib.hs_set_exchange(2, None) != 0


@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_device_scope_lu_matches_the_oracle(lib, dtype):
    from rustymath_amd import linalg as la
    n = 2100
    print(f'device-scope exchange: LU n={n} {np.dtype(dtype).name}', flush=True)
    a = random(n, n, n + 3, dtype)
    set_mode(lib, 1)
    try:
        f = la.lu(a)
    finally:
        set_mode(lib, 0)
    want, perm, _ = oracle.lu(a)
    assert (f.p == perm).all()
    assert bits_equal(f.factors, want)


def test_device_scope_factor_and_single_column_solve_equal_the_default(lib):
    "FP32 of order 8192: the panels' exchange and the substitutions' runs differ between the modes, the bits do not."
    import torch
    from rustymath_amd import device as dv
    n = 8192
    print(f'device-scope exchange: getrf + getrs n={n} float32', flush=True)
    g = torch.Generator(device='cuda'); g.manual_seed(12)
    a = torch.randn(n, n, dtype=torch.float32, device='cuda', generator=g)
    b = torch.randn(n, 1, dtype=torch.float32, device='cuda', generator=g)
    out = []
    for mode in (0, 1, 0):
        set_mode(lib, mode)
        f, perm = dv.getrf(a.clone())
        x = b.clone()
        dv.getrs(f, perm, x)
        torch.cuda.synchronize()
        out.append((f.view(torch.int32), perm, x.view(torch.int32)))
    for got in out[1:]:
        assert all(torch.equal(u, v) for u, v in zip(out[0], got))
