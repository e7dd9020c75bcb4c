"""GPU checks of the hand-written adjoint (rustymath_amd.adjoint): loss = sum(w ⊙ solve(A, matmul(C, D))).

The gradients come from the library's kernels (the forward pass's LU factors reused for the transposed solve, MFMA
products for Ā = −B̄ Xᵀ, C̄ = B̄ Dᵀ, D̄ = Cᵀ B̄). They are checked three ways: torch.autograd.gradcheck (central
finite differences, step 1e-6) at n = 64; a central difference of step 1e-6 along random directions; and the
FP64 CPU gradient of the same loss through torch's own CPU solve and matmul (an independent library), to a
relative error of 1e-8, on a stack of 256 systems of order 256 and on one system of order 4096. C is n×n and D is
n×16 (the right-hand sides B = C D have 16 columns); A = G + sqrt(n) I and G, C, D, w are N(0, 1).
"""
import numpy as np
import pytest

import oracle
from test_oracle import bits_equal

pytestmark = pytest.mark.gpu
K = 16


@pytest.fixture(scope='module')
def adj():
    import torch
    torch.cuda.set_device(0)
    from rustymath_amd import adjoint
    return adjoint


def problem(n, batch, seed, device='cuda'):
    import torch
    g = torch.Generator().manual_seed(seed)
    shape = (batch,) if batch else ()
    a = torch.randn(*shape, n, n, generator=g, dtype=torch.float64) + n ** 0.5 * torch.eye(n, dtype=torch.float64)
    c = torch.randn(*shape, n, n, generator=g, dtype=torch.float64)
    d = torch.randn(*shape, n, K, generator=g, dtype=torch.float64)
    w = torch.randn(*shape, n, K, generator=g, dtype=torch.float64)
    return [x.to(device) for x in (a, c, d, w)]


def loss_fn(adj, w):
    return lambda a, c, d: (w * adj.solve(a, adj.matmul(c, d))).sum()


def cpu_gradients(a, c, d, w):
    import torch
    a, c, d, w = [x.detach().cpu().clone().requires_grad_(True) for x in (a, c, d, w)]
    (w * torch.linalg.solve(a, c @ d)).sum().backward()
    return a.grad, c.grad, d.grad


def rel(x, y):
    return float((x.cpu() - y.cpu()).norm() / y.cpu().norm())


def test_forward_is_the_library_solve_bit_for_bit(adj):
    "The forward pass is rustymath's solve of A X = C D (oracle bits), for one system and for a stack."
    import torch
    a, c, d, _ = problem(96, 0, 1)
    x = adj.solve(a, adj.matmul(c, d))
    b = oracle.matmul(c.cpu().numpy(), d.cpu().numpy())
    assert bits_equal(x.cpu().numpy(), oracle.solve(a.cpu().numpy(), b))
    a3, c3, d3, _ = problem(40, 5, 2)
    x3 = adj.solve(a3, adj.matmul(c3, d3))
    for i in range(5):
        b = oracle.matmul(c3[i].cpu().numpy(), d3[i].cpu().numpy())
        assert bits_equal(x3[i].cpu().numpy(), oracle.solve(a3[i].cpu().numpy(), b))
    torch.cuda.synchronize()


def test_gradcheck_n64(adj):
    "torch.autograd.gradcheck: central differences of step 1e-6 against the adjoint, every input item."
    import torch
    a, c, d, w = problem(64, 0, 3)
    inputs = tuple(x.clone().requires_grad_(True) for x in (a, c, d))
    assert torch.autograd.gradcheck(loss_fn(adj, w), inputs, eps=1e-6, atol=1e-6, rtol=1e-5)


def test_central_differences_along_random_directions(adj):
    import torch
    a, c, d, w = problem(64, 0, 4)
    f = loss_fn(adj, w)
    xs = [x.clone().requires_grad_(True) for x in (a, c, d)]
    f(*xs).backward()
    g = torch.Generator().manual_seed(9)
    h = 1e-6
    for i in range(3):
        v = torch.randn(xs[i].shape, generator=g, dtype=torch.float64).cuda()
        plus = [x.detach().clone() for x in xs]; minus = [x.detach().clone() for x in xs]
        plus[i] += h * v; minus[i] -= h * v
        fd = (f(*plus) - f(*minus)).item() / (2 * h)
        an = float((xs[i].grad * v).sum())
        assert abs(fd - an) <= 1e-6 * abs(an), (i, fd, an)


def test_closed_forms(adj):
    "B̄ = A⁻ᵀ w and Ā = −B̄ Xᵀ, C̄ = B̄ Dᵀ, D̄ = Cᵀ B̄, evaluated in FP64 on the host."
    import torch
    a, c, d, w = problem(80, 0, 5)
    xs = [x.clone().requires_grad_(True) for x in (a, c, d)]
    loss_fn(adj, w)(*xs).backward()
    A, C, D, W = [x.cpu().numpy() for x in (a, c, d, w)]
    X = np.linalg.solve(A, C @ D)
    Bbar = np.linalg.solve(A.T, W)
    for got, want in ((xs[0].grad, -Bbar @ X.T), (xs[1].grad, Bbar @ D.T), (xs[2].grad, C.T @ Bbar)):
        assert np.linalg.norm(got.cpu().numpy() - want) <= 1e-10 * np.linalg.norm(want)


@pytest.mark.parametrize('n,batch', [(256, 256), (4096, 0)])
def test_matches_the_fp64_cpu_gradient(adj, n, batch):
    "Relative error at most 1e-8 against torch's FP64 CPU gradient: 256 systems of order 256, one of order 4096."
    import torch
    a, c, d, w = problem(n, batch, 6)
    xs = [x.clone().requires_grad_(True) for x in (a, c, d)]
    loss_fn(adj, w)(*xs).backward()
    torch.cuda.synchronize()
    want = cpu_gradients(a, c, d, w)
    for got, ref in zip((x.grad for x in xs), want):
        assert rel(got, ref) <= 1e-8, rel(got, ref)


def test_singular_systems_raise_for_one_matrix_and_for_a_stack(adj):
    """A matrix with a negligible pivot is where rustymath's `solve` takes least squares (core/src/linalg/factor.rs:208),
    which this library leaves out: the adjoint's forward pass raises NotImplementedError for it, whether it comes alone
    (the blocked factorization) or in a stack (the batched kernel)."""
    import torch
    a = torch.ones(64, 64, dtype=torch.float64, device='cuda')
    b = torch.ones(64, 2, dtype=torch.float64, device='cuda')
    with pytest.raises(NotImplementedError):
        adj.solve(a, b)
    a3 = torch.eye(40, dtype=torch.float64, device='cuda').repeat(3, 1, 1)
    a3[1] = 1.0
    with pytest.raises(NotImplementedError):
        adj.solve(a3, torch.ones(3, 40, 2, dtype=torch.float64, device='cuda'))
    big = torch.ones(600, 600, dtype=torch.float64, device='cuda')
    with pytest.raises(NotImplementedError):
        adj.solve(big.unsqueeze(0).repeat(2, 1, 1), torch.ones(2, 600, 1, dtype=torch.float64, device='cuda'))
