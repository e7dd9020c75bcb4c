"""GPU checks of the mixed-precision solves: hs_dev_dsgesv (FP32 LU on MFMA, FP64 iterative refinement on the GPU) and
hs_dev_dbgesv (BF16 trailing products, FP64 GMRES-based refinement), and of the BF16 trailing product itself.

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


@pytest.mark.parametrize('m,n,k', [(1, 1, 1), (130, 200, 70), (257, 129, 300), (512, 512, 256), (64, 1000, 33)])
def test_bf16_update_is_the_rounded_product_summed_in_fp32(dv, m, n, k):
    """c -= bf16(a) bf16(b): against the same product of the rounded operands in FP64 on the CPU, within FP32
    accumulation's bound k·u32·(|a'||b'|) (a', b' the bfloat16 operands; u32 = 2^-24) plus the final rounding of c."""
    import torch
    g = torch.Generator(device='cuda'); g.manual_seed(m * 7 + n * 3 + k)
    a = torch.randn(m, k, device='cuda', generator=g)
    b = torch.randn(k, n, device='cuda', generator=g)
    c = torch.randn(m, n, device='cuda', generator=g)
    ar, br = a.bfloat16().double().cpu(), b.bfloat16().double().cpu()
    want = c.double().cpu() - ar @ br
    got = dv.gemm_bf16_update(c.clone(), a, b).double().cpu()
    bound = (k + 2) * 2.0 ** -24 * ((ar.abs() @ br.abs()) + c.double().cpu().abs())
    assert ((got - want).abs() <= bound).all(), float(((got - want).abs() / bound).max())


@pytest.mark.parametrize('m,n,k', [(1, 1, 1), (130, 200, 70), (257, 129, 300), (512, 512, 256)])
def test_split_bf16_update_is_the_three_products_summed_in_fp32(dv, m, n, k):
    """terms = 3: each operand x split into hi = bf16(x) and lo = bf16(x - hi), and c -= a_lo b_hi + a_hi b_lo + a_hi b_hi.
    Against those three products of the split operands in FP64 on the CPU, within FP32 accumulation's bound; and
    against the FP32 operands' own product within the split's 2^-15 relative error, which a single rounding to
    bfloat16 (2^-8) would exceed."""
    import torch
    g = torch.Generator(device='cuda'); g.manual_seed(m * 5 + n * 11 + k)
    a = torch.randn(m, k, device='cuda', generator=g)
    b = torch.randn(k, n, device='cuda', generator=g)
    c = torch.randn(m, n, device='cuda', generator=g)
    split = lambda x: (x.bfloat16().float(), (x - x.bfloat16().float()).bfloat16().float())
    (ah, al), (bh, bl) = split(a), split(b)
    ah, al, bh, bl = (t.double().cpu() for t in (ah, al, bh, bl))
    c64 = c.double().cpu()
    want = c64 - (al @ bh + ah @ bl + ah @ bh)
    got = dv.gemm_bf16_update(c.clone(), a, b, terms=3).double().cpu()
    mag = (al.abs() @ bh.abs()) + (ah.abs() @ bl.abs()) + (ah.abs() @ bh.abs()) + c64.abs()
    assert ((got - want).abs() <= (3 * k + 2) * 2.0 ** -24 * mag).all()
    exact = c64 - a.double().cpu() @ b.double().cpu()
    scale = a.double().cpu().abs() @ b.double().cpu().abs() + c64.abs()
    assert ((got - exact).abs() <= 2.0 ** -15 * scale).all()
    once = dv.gemm_bf16_update(c.clone(), a, b, terms=1).double().cpu()
    if k > 1: assert (once - exact).abs().max() > 2.0 ** -15 * float(scale.max()) / 16


@pytest.mark.parametrize('n', [2048])
def test_bf16_gmres_refinement_reaches_fp64_accuracy(dv, n):
    """The BF16 path on the cond-1e6 Householder matrix: classical refinement cannot converge from BF16 factors here
    (cond · 2^-8 >> 1); GMRES-based refinement must reach 10·n·eps64, with the reported backward error the true one
    and the solution as close to the plain FP64 LU solution as two backward-stable solves allow."""
    import torch
    from bench import householder_matrix
    a, b = householder_matrix(n, seed=3)
    tol = 10 * n * EPS
    x, iters, inner, berr = dv.dbgesv(a, b, tol=tol)
    assert berr <= tol and 1 <= iters <= 30 and inner >= iters, (iters, inner, berr)
    assert abs(true_backward_error(a, x, b) - berr) <= 1e-3 * berr
    x64 = fp64_solution(dv, a, b)
    assert float((x - x64).norm() / x64.norm()) <= 2 * 1e6 * tol
    torch.cuda.synchronize()


@pytest.mark.parametrize('terms', [1, 3])
def test_bf16_gmres_refinement_on_a_gaussian_system(dv, terms):
    import torch
    n = 4096
    g = torch.Generator(device='cuda'); g.manual_seed(1)
    a = torch.randn(n, n, dtype=torch.float64, device='cuda', generator=g)
    b = torch.randn(n, dtype=torch.float64, device='cuda', generator=g)
    x, iters, inner, berr = dv.dbgesv(a, b, terms=terms)
    assert berr <= 10 * n * EPS, (iters, inner, berr)
    assert abs(true_backward_error(a, x, b) - berr) <= 1e-3 * berr


def gaussian_system(n, seed):
    "The N(0, 1) variant of the mixed-precision workload: A and b iid N(0, 1) from torch's CUDA generator."
    import torch
    g = torch.Generator(device='cuda'); g.manual_seed(seed)
    a = torch.randn(n, n, dtype=torch.float64, device='cuda', generator=g)
    b = torch.randn(n, dtype=torch.float64, device='cuda', generator=g)
    return a, b


@pytest.mark.parametrize('matrix,form', [('gaussian', 'fp32'), ('gaussian', 'bf16x3'), ('cond1e6', 'bf16x3'),
                                         ('cond1e6', 'bf16')])
def test_mixed_precision_at_the_benchmark_order(dv, matrix, form):
    """The mixed-precision workload at its full order, n = 16384, on both of its matrices (the cond-1e6 Householder
    matrix and an iid N(0, 1) one, seed 4 as the benchmark draws it): FP32 LU with classical refinement and BF16
    trailing products (split pairs, or rounded once) with GMRES-based refinement each reach 10·n·eps64, and report
    their true backward error; the solution stays within 2·cond·10·n·eps64 of the plain FP64 LU solution on the
    Householder matrix, whose condition number is 1e6 by construction."""
    import torch
    from bench import householder_matrix
    n = 16384
    print(f'mixed n={n} {matrix} {form}', flush=True)
    a, b = householder_matrix(n, seed=3) if matrix == 'cond1e6' else gaussian_system(n, 4)
    tol = 10 * n * EPS
    if form == 'fp32':
        x, iters, berr = dv.dsgesv(a, b, tol=tol)
    else:
        x, iters, inner, berr = dv.dbgesv(a, b, tol=tol, max_iter=10, terms=3 if form == 'bf16x3' else 1)
    assert berr <= tol, (form, iters, berr)
    assert abs(true_backward_error(a, x, b) - berr) <= 1e-3 * berr
    if matrix == 'cond1e6':
        x64 = fp64_solution(dv, a, b)
        assert float((x - x64).norm() / x64.norm()) <= 2 * 1e6 * tol
    del a
    torch.cuda.synchronize()
