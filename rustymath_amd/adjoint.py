"""Hand-written adjoints of solve and matmul on the device, as torch.autograd.Functions over the HIP library.

rustymath has no gradients for its linear algebra (core/src/derivative.rs covers scalar formulas only); these are
the standard adjoints, built so that the backward pass reuses what the forward pass computed on the GPU:

  X = solve(A, B):  B̄ = A⁻ᵀ X̄, by transposed substitutions with the forward pass's own LU factors (Uᵀ, then Lᵀ, then
                    the inverse row permutation; no second factorization), and Ā = −B̄ Xᵀ, one MFMA product;
  C = matmul(A, B): Ā = C̄ Bᵀ and B̄ = Aᵀ C̄, MFMA products with transposed operands read in place.

A stack's products (forward and backward) are one batched launch each.

Forward results are the library's (bit for bit rustymath's `lu_factor`/`lu_solve` and product folds). Inputs are
float64 or float32 CUDA tensors: A (n×n) with B (n×k) or (n,), or stacks of them (batch×n×n with batch×n×k or
batch×n). A stack of systems with n <= 512 is factored and solved by the batched kernel, one workgroup per system;
larger ones go through the blocked factorization one system at a time.
"""
import torch

from . import device as dv


def _chunks(k):
    "Column ranges of at most 8 right-hand sides, the batched kernel's limit."
    return [(c, min(c + 8, k)) for c in range(0, k, 8)] or [(0, 0)]


def _factor_solve(a, b):
    "Factors the square matrices of `a` (batch×n×n) and solves for b (batch×n×k): (factors, perms, x, layout)."
    batch, n = a.shape[0], a.shape[1]
    k = b.shape[2]
    if 0 < n <= 512 and batch > 1:
        fa = a.detach().contiguous().clone().reshape(-1)
        x = torch.empty_like(b, memory_format=torch.contiguous_format)
        perms = None
        for c0, c1 in _chunks(k):
            part = b[:, :, c0:c1].detach().contiguous().reshape(-1)
            if perms is None:
                plan = dv.BatchedSolve(fa, part, [n] * batch, [c1 - c0] * batch)
                plan.run(graph=False)
                perms, ok = plan.perm[:batch * n], plan.regular[:batch]
                if not bool(ok.all()):
                    raise NotImplementedError('a matrix of the stack is singular to working precision: rustymath would '
                                              'take least squares there')
            else:
                plan = dv.BatchedSolve(fa, part, [n] * batch, [c1 - c0] * batch, mode='lu_solve', perm=perms)
                plan.run(graph=False)
            plan.close()
            x[:, :, c0:c1] = part.view(batch, n, c1 - c0)
        return fa.view(batch, n, n), perms.view(batch, n), x, 'rows'
    fs, ps, xs = [], [], []
    for i in range(batch):
        x = b[i].detach().contiguous().clone()
        f, perm, regular = dv.gesv(a[i].detach().contiguous().clone(), x)
        if not regular:
            raise NotImplementedError('the matrix is singular to working precision: rustymath would take least squares '
                                      'there')
        fs.append(f); ps.append(perm); xs.append(x)
    return torch.stack(fs), torch.stack(ps), torch.stack(xs), 'columns'


def _solve_transposed(fs, perms, g, layout):
    "B̄ = A⁻ᵀ g for each system, with the forward pass's factors."
    batch, n, k = g.shape
    if layout == 'rows':
        out = torch.empty_like(g, memory_format=torch.contiguous_format)
        for c0, c1 in _chunks(k):
            part = g[:, :, c0:c1].detach().contiguous().reshape(-1)
            plan = dv.BatchedSolve(fs.reshape(-1), part, [n] * batch, [c1 - c0] * batch, mode='solve_t',
                                   perm=perms.reshape(-1))
            plan.run(graph=False)
            plan.close()
            out[:, :, c0:c1] = part.view(batch, n, c1 - c0)
        return out
    out = g.detach().contiguous().clone()
    for i in range(batch):
        dv.getrs(fs[i], perms[i], out[i], trans=True)
    return out


class Solve(torch.autograd.Function):
    "X = A⁻¹ B with B̄ = A⁻ᵀ X̄ (transposed substitutions on the saved factors) and Ā = −B̄ Xᵀ."

    @staticmethod
    def forward(ctx, a, b):
        if a.dim() not in (2, 3) or a.shape[-1] != a.shape[-2]: raise ValueError('solve needs square matrices')
        stacked = a.dim() == 3
        vec = b.dim() == a.dim() - 1
        a3 = a if stacked else a.unsqueeze(0)
        b3 = (b.unsqueeze(-1) if vec else b)
        b3 = b3 if stacked else b3.unsqueeze(0)
        if b3.shape[:2] != a3.shape[:2]: raise ValueError(f'solve: {tuple(a.shape)} with {tuple(b.shape)}')
        fs, perms, x, layout = _factor_solve(a3, b3)
        ctx.save_for_backward(fs, perms, x)
        ctx.layout, ctx.vec, ctx.stacked = layout, vec, stacked
        out = x.squeeze(-1) if vec else x
        return out if stacked else out[0]

    @staticmethod
    def backward(ctx, gx):
        fs, perms, x = ctx.saved_tensors
        g = gx if ctx.stacked else gx.unsqueeze(0)
        g = g.unsqueeze(-1) if ctx.vec else g
        gb = _solve_transposed(fs, perms, g, ctx.layout)
        ga = None
        if ctx.needs_input_grad[0]:  # Ā = −B̄ Xᵀ for every system, one launch
            ga = dv.gemm_batched(gb.contiguous(), x.contiguous(), alpha=-1.0, trans_b=True)
            ga = ga if ctx.stacked else ga[0]
        gb = gb.squeeze(-1) if ctx.vec else gb
        gb = gb if ctx.stacked else gb[0]
        return ga, (gb if ctx.needs_input_grad[1] else None)


class Matmul(torch.autograd.Function):
    "C = A B with Ā = C̄ Bᵀ and B̄ = Aᵀ C̄ (rustymath's `matmul` forward, core/src/linalg.rs:799-807)."

    @staticmethod
    def forward(ctx, a, b):
        if a.dim() != b.dim() or a.dim() not in (2, 3): raise ValueError('matmul takes two matrices or two stacks')
        ctx.save_for_backward(a, b)
        if a.dim() == 2: return dv.gemm(a.contiguous(), b.contiguous())
        return dv.gemm_batched(a.contiguous(), b.contiguous())

    @staticmethod
    def backward(ctx, gc):
        a, b = ctx.saved_tensors
        gc = gc.contiguous()
        ga = gb = None
        if a.dim() == 2:
            if ctx.needs_input_grad[0]: ga = dv.gemm(gc, b.contiguous(), trans_b=True)
            if ctx.needs_input_grad[1]: gb = dv.gemm(a.contiguous(), gc, trans_a=True)
            return ga, gb
        if ctx.needs_input_grad[0]: ga = dv.gemm_batched(gc, b.contiguous(), trans_b=True)
        if ctx.needs_input_grad[1]: gb = dv.gemm_batched(a.contiguous(), gc, trans_a=True)
        return ga, gb


def solve(a, b):
    "Differentiable X = A⁻¹ B on the device (see Solve)."
    return Solve.apply(a, b)


def matmul(a, b):
    "Differentiable A B on the device (see Matmul)."
    return Matmul.apply(a, b)
