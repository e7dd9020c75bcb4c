"""GPU checks of the mixed-precision solve (hs_dev_dsgesv): FP32 LU on MFMA, FP64 iterative refinement on the GPU.

There is no reference function to match (rustymath has no refinement), so the checks are the contract the solve
states: the normwise backward error ||b − A x|| / (||A||_F ||x||) that it reports is the true one (recomputed
independently in FP64) and is at most 10·n·eps64; the solution agrees with the plain FP64 LU solve to within what
two backward-stable solves allow (forward error ≤ 2 · cond · 10·n·eps64); the FP32 factors it refines with are the oracle's FP32 LU of
float32(A), bit for bit. Matrices: A = U diag(s) Vᵀ (8 + 8 Householder reflectors, s log-spaced 1 … 1e-6, cond
1e6) and iid N(0, 1).
"""
import numpy as np
import pytest

import oracle
from test_oracle import bits_equal

pytestmark = pytest.mark.gpu
EPS = np.finfo(np.float64).eps


@pytest.fixture(scope='module')
def dv():
    import torch
    torch.cuda.set_device(0)
    from rustymath_amd import device
    return device


def true_backward_error(a, x, b):
    return float((b - a @ x).norm() / (a.norm() * x.norm()))


def fp64_solution(dv, a, b):
    import torch
    f, perm = dv.getrf(a.clone())
    x = b.reshape(-1, 1).clone()
    dv.getrs(f, perm, x)
    return x.reshape(-1)


@pytest.mark.parametrize('n', [2048, 16384])
def test_ill_conditioned_householder_system(dv, n):
    import torch
    from bench import householder_matrix
    a, b = householder_matrix(n, seed=3)
    tol = 10 * n * EPS
    x, iters, berr = dv.dsgesv(a, b, tol=tol)
    assert 1 <= iters <= 30 and berr <= tol, (iters, berr)
    assert abs(true_backward_error(a, x, b) - berr) <= 1e-3 * berr
    x64 = fp64_solution(dv, a, b)
    fwd = float((x - x64).norm() / x64.norm())
    assert fwd <= 2 * 1e6 * tol, fwd
    torch.cuda.synchronize()


def test_well_conditioned_gaussian_system(dv):
    import torch
    n = 4096
    g = torch.Generator(device='cuda'); g.manual_seed(1)
    a = torch.randn(n, n, dtype=torch.float64, device='cuda', generator=g)
    b = torch.randn(n, dtype=torch.float64, device='cuda', generator=g)
    x, iters, berr = dv.dsgesv(a, b)
    assert berr <= 10 * n * EPS and iters <= 30
    assert abs(true_backward_error(a, x, b) - berr) <= 1e-3 * berr


def test_iteration_cap_and_reported_error(dv):
    "max_iter = 1 stops after one correction and reports that correction's backward error."
    import torch
    from bench import householder_matrix
    a, b = householder_matrix(1024, seed=5)
    x, iters, berr = dv.dsgesv(a, b, max_iter=1, tol=0.0)
    assert iters == 1
    assert abs(true_backward_error(a, x, b) - berr) <= 1e-3 * berr


def test_low_precision_factors_are_the_oracle_fp32_lu(dv):
    "The factors the refinement uses: the FP32 LU of float32(A), bit for bit the oracle's (rustymath's f32 lu)."
    import torch
    from bench import householder_matrix
    a, _ = householder_matrix(600, seed=7)
    a32 = a.float().contiguous()
    f, perm = dv.getrf(a32.clone())
    wf, wperm, _ = oracle.lu(a32.cpu().numpy())
    assert (perm.cpu().numpy() == wperm).all()
    assert bits_equal(f.cpu().numpy(), wf.reshape(600, 600))
