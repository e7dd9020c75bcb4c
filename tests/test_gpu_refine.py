"""GPU checks of the mixed solve's own machinery: the permutation that getrf composes on the GPU from its panels' row
moves, and the refinement's blocked FP32 substitutions with the convergence test on the GPU.

The permutation is rustymath's `perm` (core/src/linalg/factor.rs:132-134), which `lu`'s host path builds from the
pivots: the device form must equal it item for item, at orders where the panels run with and without look-ahead, and
above the 16384 rows the composition keeps in LDS. The refinement has no reference counterpart (rustymath has no
iterative refinement): its contract is the reported normwise backward error, which must be the true one and at most
10·n·eps64, at orders on both sides of the substitutions' 128-row blocks.
"""
import numpy as np
import pytest

from support import random

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps


@pytest.fixture(scope='module')
def dv():
    import torch
    torch.cuda.set_device(0)
    from rustymath_amd import device
    return device


@pytest.mark.parametrize('n', [1, 17, 600, 2100, 16448])
def test_device_perm_equals_the_perm_of_the_pivots(dv, n):
    "hs_dev_dgetrf's perm (composed on the GPU) against hs_dlu_factor's (built on the host from the pivots)."
    import torch
    from rustymath_amd import linalg as la
    a = random(n, n, 40 + n, np.float64)
    print(f'device perm n={n}', flush=True)
    want = la.lu(a).p
    _, perm = dv.getrf(torch.from_numpy(a).cuda())
    assert (perm.cpu().numpy() == want).all()


def true_backward_error(a, x, b):
    """Recomputed independently (torch's product): near the rounding level it differs from the solve's own residual by
    the order of summation, so the comparisons allow 1 % of the tolerance besides 0.1 % of the value."""
    return float((b - a @ x).norm() / (a.norm() * x.norm()))


@pytest.mark.parametrize('n', [1, 2, 63, 127, 128, 129, 300, 1000, 4099])
def test_refinement_at_block_edges(dv, n):
    "FP32 LU + FP64 refinement: the reported backward error is the true one and within 10·n·eps64."
    import torch
    g = torch.Generator(device='cuda'); g.manual_seed(n)
    a = torch.randn(n, n, dtype=torch.float64, device='cuda', generator=g) + (n ** 0.5) * torch.eye(n, dtype=torch.float64, device='cuda')
    b = torch.randn(n, dtype=torch.float64, device='cuda', generator=g)
    print(f'dsgesv n={n}', flush=True)
    tol = 10 * n * EPS
    x, iters, berr = dv.dsgesv(a, b, tol=tol)
    assert 1 <= iters <= 30 and berr <= tol, (iters, berr)
    assert abs(true_backward_error(a, x, b) - berr) <= 1e-3 * berr + 1e-2 * tol


@pytest.mark.parametrize('n', [129, 1000])
def test_gmres_refinement_at_block_edges(dv, n):
    "BF16 trailing products + GMRES refinement, preconditioned by the blocked FP32 substitutions."
    import torch
    g = torch.Generator(device='cuda'); g.manual_seed(n + 7)
    a = torch.randn(n, n, dtype=torch.float64, device='cuda', generator=g)
    b = torch.randn(n, dtype=torch.float64, device='cuda', generator=g)
    print(f'dbgesv n={n}', flush=True)
    tol = 10 * n * EPS
    x, iters, inner, berr = dv.dbgesv(a, b, tol=tol)
    assert berr <= tol, (iters, inner, berr)
    assert abs(true_backward_error(a, x, b) - berr) <= 1e-3 * berr + 1e-2 * tol


def test_refinement_runs_back_to_back(dv):
    "Two solves in a row on one stream (fresh epochs, the same pooled words) both converge."
    import torch
    from bench import householder_matrix
    a, b = householder_matrix(1536, seed=9)
    tol = 10 * 1536 * EPS
    for _ in range(2):
        x, iters, berr = dv.dsgesv(a, b, tol=tol)
        assert berr <= tol and abs(true_backward_error(a, x, b) - berr) <= 1e-3 * berr
