"""Device-resident entry points on torch tensors (HBM): the `hs_dev_*` half of the C ABI.

torch provides device memory and the current stream; every product, factorization and solve runs in
libhip_sys.so's kernels. Matrices are row-major 2-D tensors (any leading dimension along rows, unit column
stride); factors are rustymath's column-major layout in the same buffer.
"""
import torch

from . import _lib


def _stream():
    return torch.cuda.current_stream().cuda_stream


def _mat(x, name):
    if not x.is_cuda: raise ValueError(f'{name} must be a GPU tensor')
    if x.dim() != 2 or x.stride(1) != 1: raise ValueError(f'{name} must be a 2-D tensor with unit column stride')
    if x.dtype not in (torch.float64, torch.float32): raise TypeError(f'{name} must be float64 or float32')
    return x


def _sfx(dtype): return 'd' if dtype == torch.float64 else 's'


def gemm(a, b, c=None, alpha=1.0, beta=0.0, trans_a=False, trans_b=False):
    "c = alpha op(a) op(b) + beta c in place (a new tensor where c is None), as `Tiled::gemm` (core/src/linalg.rs:799)."
    _mat(a, 'a'); _mat(b, 'b')
    m, k = (a.shape[1], a.shape[0]) if trans_a else a.shape
    kb, n = (b.shape[1], b.shape[0]) if trans_b else b.shape
    if k != kb: raise ValueError(f'gemm contracts {k} items with {kb}')
    if c is None:
        c = torch.empty((m, n), dtype=a.dtype, device=a.device)
        beta = 0.0
    _mat(c, 'c')
    f = getattr(_lib.load(), f'hs_dev_{_sfx(a.dtype)}gemm')
    st = f(_lib.TRANS if trans_a else _lib.NO_TRANS, _lib.TRANS if trans_b else _lib.NO_TRANS, m, n, k, alpha,
           a.data_ptr(), max(a.stride(0), 1), b.data_ptr(), max(b.stride(0), 1), beta, c.data_ptr(), max(c.stride(0), 1),
           _stream())
    _lib.check(st, 'gemm')
    return c


def gemm_batched(a, b, c=None, alpha=1.0, beta=0.0, trans_a=False, trans_b=False):
    """c_i = alpha op(a_i) op(b_i) + beta c_i for every cell of a stack in one launch, each cell exactly as `gemm`
    computes it. a, b and c are 3-D (cells, rows, columns) with unit column stride; a 2-D a or b is used for every cell.
    Returns c (a new stack where c is None)."""
    def cells(x, name):
        if not x.is_cuda or x.stride(-1) != 1 or x.dim() not in (2, 3):
            raise ValueError(f'{name} must be a 2-D or 3-D GPU tensor with unit column stride')
        return (x, 0) if x.dim() == 2 else (x, x.stride(0))
    (a, sa), (b, sb) = cells(a, 'a'), cells(b, 'b')
    batch = max(a.shape[0] if a.dim() == 3 else 1, b.shape[0] if b.dim() == 3 else 1)
    for x in (a, b):
        if x.dim() == 3 and x.shape[0] != batch: raise ValueError('a and b must hold as many cells')
    ar, ac = a.shape[-2:]; br, bc = b.shape[-2:]
    m, k = (ac, ar) if trans_a else (ar, ac)
    kb, n = (bc, br) if trans_b else (br, bc)
    if k != kb: raise ValueError(f'gemm contracts {k} items with {kb}')
    if c is None:
        c = torch.empty((batch, m, n), dtype=a.dtype, device=a.device)
        beta = 0.0
    if c.dim() != 3 or c.stride(-1) != 1 or tuple(c.shape) != (batch, m, n): raise ValueError('c must be cells x m x n')
    f = getattr(_lib.load(), f'hs_dev_{_sfx(a.dtype)}gemm_batched')
    ld = lambda x: max(x.stride(-2), 1)
    _lib.check(f(_lib.TRANS if trans_a else _lib.NO_TRANS, _lib.TRANS if trans_b else _lib.NO_TRANS, m, n, k, alpha,
                 a.data_ptr(), ld(a), sa, b.data_ptr(), ld(b), sb, beta, c.data_ptr(), ld(c), c.stride(0), batch,
                 _stream()), 'gemm_batched')
    return c


def getrf(a, info=False):
    """Factors the square row-major tensor `a` in place into rustymath's column-major packed factors and returns
    (a, perm) with perm an int64 GPU tensor (core/src/linalg/factor.rs:617-630); with `info`, (a, perm, info) where
    info is 0 or 1 + the first column whose pivot is exactly zero."""
    import ctypes
    _mat(a, 'a')
    n = a.shape[0]
    if a.shape[1] != n: raise ValueError('getrf needs a square matrix')
    perm = torch.empty(n, dtype=torch.int64, device=a.device)
    first_zero = ctypes.c_int64(0)
    f = getattr(_lib.load(), f'hs_dev_{_sfx(a.dtype)}getrf')
    _lib.check(f(n, a.data_ptr(), a.stride(0), perm.data_ptr(), ctypes.byref(first_zero), _stream()), 'getrf')
    return (a, perm, first_zero.value) if info else (a, perm)


def gesv(a, b):
    """`LuWork::solve` (core/src/linalg/factor.rs:207-226) in place on device tensors: factors the square row-major `a`
    and, where every pivot exceeds eps·n·max|a|, solves the row-major b (n×k) in place. Returns (a, perm, regular);
    b is untouched where regular is False (the reference would take least squares there)."""
    import ctypes
    _mat(a, 'a'); _mat(b, 'b')
    n = a.shape[0]
    if a.shape[1] != n or b.shape[0] != n: raise ValueError('gesv needs a square a and b with as many rows')
    if a.dtype != b.dtype: raise TypeError('a and b must have one dtype')
    perm = torch.empty(n, dtype=torch.int64, device=a.device)
    reg = ctypes.c_int(0)
    f = getattr(_lib.load(), f'hs_dev_{_sfx(a.dtype)}gesv')
    _lib.check(f(n, a.data_ptr(), a.stride(0), perm.data_ptr(), b.data_ptr(), b.shape[1], max(b.stride(0), 1),
                 ctypes.byref(reg), _stream()), 'gesv')
    return a, perm, bool(reg.value) or n == 0


def getrs(factors, perm, b, trans=False):
    "Solves A x = b (or Aᵀ x = b) in place on the row-major b (n×k) with factors from getrf (factor.rs:633-641)."
    _mat(factors, 'factors'); _mat(b, 'b')
    n, k = b.shape
    f = getattr(_lib.load(), f'hs_dev_{_sfx(b.dtype)}getrs')
    _lib.check(f(1 if trans else 0, n, k, factors.data_ptr(), factors.stride(0), perm.data_ptr(), b.data_ptr(),
                 b.stride(0), _stream()), 'getrs')
    return b


def getri(factors, perm):
    "The inverse (n×n, row-major) from factors and perm as getrf gives them (`lu_inverse`, factor.rs:645-664)."
    _mat(factors, 'factors')
    n = factors.shape[0]
    out = torch.empty((n, n), dtype=factors.dtype, device=factors.device)
    f = getattr(_lib.load(), f'hs_dev_{_sfx(factors.dtype)}getri')
    _lib.check(f(n, factors.data_ptr(), factors.stride(0), perm.data_ptr(), out.data_ptr(), out.stride(0), _stream()),
               'getri')
    return out


def potrf(a):
    """The Cholesky factor in place on the row-major `a` (lower triangle read, L left with zeros above) with rustymath's
    order of operations (factor.rs:777-797); returns (a, ok), ok False where `a` is not positive definite."""
    import ctypes
    _mat(a, 'a')
    ok = ctypes.c_int(0)
    f = getattr(_lib.load(), f'hs_dev_{_sfx(a.dtype)}potrf')
    _lib.check(f(a.shape[0], a.data_ptr(), a.stride(0), ctypes.byref(ok), _stream()), 'potrf')
    return a, bool(ok.value)


def dsgesv(a, b, max_iter=30, tol=None):
    """Solves A x = b for the f64 matrix `a` (n×n, compact) and vector `b` by an FP32 LU on MFMA and FP64 iterative
    refinement on the GPU. Returns (x, corrections, backward error); `tol` defaults to 10·n·eps64."""
    import ctypes
    _mat(a, 'a')
    n = a.shape[0]
    if a.dtype != torch.float64 or b.dtype != torch.float64: raise TypeError('dsgesv takes float64')
    b = b.contiguous().reshape(n)
    x = torch.empty(n, dtype=torch.float64, device=a.device)
    if tol is None: tol = 10 * n * torch.finfo(torch.float64).eps
    it, be = ctypes.c_int(0), ctypes.c_double(0)
    _lib.check(_lib.load().hs_dev_dsgesv(n, a.data_ptr(), a.stride(0), b.data_ptr(), x.data_ptr(), max_iter, tol,
                                         ctypes.byref(it), ctypes.byref(be), _stream()), 'dsgesv')
    return x, it.value, be.value


def dbgesv(a, b, max_iter=30, restart=40, tol=None, terms=3):
    """Solves A x = b for the f64 matrix `a` (n×n, compact) and vector `b` with BF16 trailing products in the
    factorization (FP32 panels, FP32 accumulation on MFMA; `terms` 3 splits each operand into a bfloat16 pair and takes
    three products, 1 rounds it once) and FP64 GMRES-based refinement preconditioned by those factors. Returns
    (x, refinement steps, GMRES steps, backward error); `tol` defaults to 10·n·eps64."""
    import ctypes
    _mat(a, 'a')
    n = a.shape[0]
    if a.dtype != torch.float64 or b.dtype != torch.float64: raise TypeError('dbgesv takes float64')
    b = b.contiguous().reshape(n)
    x = torch.empty(n, dtype=torch.float64, device=a.device)
    if tol is None: tol = 10 * n * torch.finfo(torch.float64).eps
    it, inner, be = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_double(0)
    _lib.check(_lib.load().hs_dev_dbgesv(n, a.data_ptr(), a.stride(0), b.data_ptr(), x.data_ptr(), max_iter, restart,
                                         terms, tol, ctypes.byref(it), ctypes.byref(inner), ctypes.byref(be),
                                         _stream()),
               'dbgesv')
    return x, it.value, inner.value, be.value


def gemm_bf16_update(c, a, b, terms=1):
    """c -= a b in place for FP32 tensors with the operands taken as bfloat16 (the BF16 factorization's trailing
    product): rounded once to nearest even (terms 1) or split into hi + lo pairs with three products (terms 3), the
    products summed in FP32 on MFMA."""
    for t, name in ((a, 'a'), (b, 'b'), (c, 'c')):
        _mat(t, name)
        if t.dtype != torch.float32: raise TypeError('gemm_bf16_update takes float32')
    m, k = a.shape
    if b.shape[0] != k or c.shape != (m, b.shape[1]): raise ValueError('gemm_bf16_update: shapes do not match')
    _lib.check(_lib.load().hs_dev_sgemm_bf16_update(m, b.shape[1], k, a.data_ptr(), a.stride(0), b.data_ptr(),
                                                    b.stride(0), c.data_ptr(), c.stride(0), terms, _stream()),
               'gemm_bf16_update')
    return c


class BatchedSolve:
    """Many independent systems A_i X_i = B_i (orders n_i <= 512, k_i <= 8 right-hand sides) solved by one persistent
    launch, replayed from a hipGraph: the cells of a rustymath stack, each solved as `LuWork::solve` would
    (core/src/linalg/factor.rs:207-226) and bit for bit its result.

    `a` and `b` are flat device buffers holding the systems back to back: A_i row-major (n_i² items) and B_i row-major
    (n_i k_i items), in the order of `ns`/`ks` (see `pack`). A run overwrites them: A_i with its factors, rows left in
    place (row perm_i[j] holds position j), and B_i with X_i where `regular[i]`. Modes 'lu_solve' and 'solve_t' take
    such factors and their `perm` and solve A_i X = B_i or A_iᵀ X = B_i in place on b. `panel_width` (16 or 32) and
    `wgs_per_cu` (1 or 2) change the speed, never the results; `tune` picks them."""

    MODES = {'solve': 0, 'lu_solve': 1, 'solve_t': 2}

    def __init__(self, a, b, ns, ks, panel_width=32, wgs_per_cu=1, mode='solve', perm=None):
        import ctypes
        import numpy as np
        if not (a.is_cuda and b.is_cuda): raise ValueError('batched buffers must be GPU tensors')
        if a.dtype != b.dtype or a.dtype not in (torch.float64, torch.float32): raise TypeError('a and b: float64 or float32')
        if not (a.is_contiguous() and b.is_contiguous()): raise ValueError('batched buffers must be contiguous')
        ns = np.ascontiguousarray(ns, dtype=np.int64); ks = np.ascontiguousarray(ks, dtype=np.int64)
        if ns.shape != ks.shape or ns.ndim != 1: raise ValueError('ns and ks: one order and one column count per system')
        self.ns, self.ks = ns, ks
        self.a_off = np.concatenate([[0], np.cumsum(ns * ns)[:-1]]).astype(np.int64) if len(ns) else ns.copy()
        self.b_off = np.concatenate([[0], np.cumsum(ns * ks)[:-1]]).astype(np.int64) if len(ns) else ns.copy()
        self.p_off = np.concatenate([[0], np.cumsum(ns)[:-1]]).astype(np.int64) if len(ns) else ns.copy()
        if a.numel() < int((ns * ns).sum()) or b.numel() < int((ns * ks).sum()):
            raise ValueError('the buffers are smaller than the systems they should hold')
        if mode not in self.MODES: raise ValueError(f'mode must be one of {sorted(self.MODES)}')
        if mode != 'solve' and (perm is None or perm.dtype != torch.int64 or perm.numel() < int(ns.sum())):
            raise ValueError("modes 'lu_solve' and 'solve_t' take the int64 perm of an earlier 'solve' run")
        self.a, self.b, self.mode = a, b, mode
        self.perm = perm if perm is not None else torch.empty(max(int(ns.sum()), 1), dtype=torch.int64, device=a.device)
        self.regular = torch.zeros(max(len(ns), 1), dtype=torch.int32, device=a.device)
        self.panel_width, self.wgs_per_cu = panel_width, wgs_per_cu
        h = ctypes.c_int64(0)
        p = lambda x: x.ctypes.data_as(ctypes.c_void_p)
        f = getattr(_lib.load(), f'hs_dev_{_sfx(a.dtype)}gesv_batched_plan')
        _lib.check(f(len(ns), p(ns), p(ks), p(self.a_off), p(self.b_off), p(self.p_off), a.data_ptr(), b.data_ptr(),
                     self.perm.data_ptr(), self.regular.data_ptr(), self.MODES[mode], panel_width, wgs_per_cu,
                     ctypes.byref(h)),
                   'gesv_batched_plan')
        self.handle = h.value

    def run(self, graph=True):
        "Solves every system on the current stream (asynchronous)."
        if self.handle is None: raise ValueError('the plan is closed')
        _lib.check(_lib.load().hs_dev_gesv_batched_run(self.handle, 1 if graph else 0, _stream()), 'gesv_batched_run')
        return self

    def close(self):
        if self.handle is not None:
            _lib.check(_lib.load().hs_dev_gesv_batched_free(self.handle), 'gesv_batched_free')
            self.handle = None

    def __del__(self):
        try: self.close()
        except Exception: pass

    def x(self, i):
        "X_i (n_i×k_i view of b)."
        n, k, o = int(self.ns[i]), int(self.ks[i]), int(self.b_off[i])
        return self.b[o:o + n * k].view(n, k)

    def factors(self, i):
        "(A_i's overwritten buffer n_i×n_i, perm_i): position j of the factors is row perm_i[j]."
        n, o, po = int(self.ns[i]), int(self.a_off[i]), int(self.p_off[i])
        return self.a[o:o + n * n].view(n, n), self.perm[po:po + n]


class BatchedMixedSolve:
    """The FP32-LU + FP64-refinement form of BatchedSolve: many float64 systems A_i X_i = B_i (n_i <= 512, k_i <= 8),
    each factored in FP32 on MFMA (float32(A_i), bit for bit rustymath's f32 `lu_factor`) and refined in FP64 on the
    GPU until ||B_i - A_i X_i||_F / (||A_i||_F ||X_i||_F) <= tol_per_n · n_i · 2^-52 or max_iter corrections have run;
    the convergence test runs on the GPU. One run is one sequence of persistent launches, replayed from a hipGraph.

    `a` and `b` are the flat float64 device buffers of BatchedSolve's layout; they are read, never written. After a run:
    `x(i)` is X_i, `iters[i]` the corrections applied, `berr[i]` the final backward error, `regular[i]` the
    reference's `regular` decision on the FP32 factors (NaN error and X left at zero where it is 0), and
    `factors32(i)` float32(A_i)'s factors with perm_i (row perm_i[j] holds position j)."""

    def __init__(self, a, b, ns, ks, max_iter=10, tol_per_n=10.0, panel_width=32, wgs_per_cu=1):
        import ctypes
        import numpy as np
        if not (a.is_cuda and b.is_cuda): raise ValueError('batched buffers must be GPU tensors')
        if a.dtype != torch.float64 or b.dtype != torch.float64: raise TypeError('a and b must be float64')
        if not (a.is_contiguous() and b.is_contiguous()): raise ValueError('batched buffers must be contiguous')
        ns = np.ascontiguousarray(ns, dtype=np.int64); ks = np.ascontiguousarray(ks, dtype=np.int64)
        if ns.shape != ks.shape or ns.ndim != 1: raise ValueError('ns and ks: one order and one column count per system')
        self.ns, self.ks = ns, ks
        offsets = lambda sizes: np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64) if len(ns) else ns.copy()
        self.a_off, self.b_off, self.p_off = offsets(ns * ns), offsets(ns * ks), offsets(ns)
        if a.numel() < int((ns * ns).sum()) or b.numel() < int((ns * ks).sum()):
            raise ValueError('the buffers are smaller than the systems they should hold')
        dev = a.device
        self.a, self.b = a, b
        self.xbuf = torch.zeros(max(b.numel(), 1), dtype=torch.float64, device=dev)
        self.a32 = torch.empty(max(a.numel(), 1), dtype=torch.float32, device=dev)
        self.r32 = torch.empty(max(b.numel(), 1), dtype=torch.float32, device=dev)
        self.perm = torch.empty(max(int(ns.sum()), 1), dtype=torch.int64, device=dev)
        self.regular = torch.zeros(max(len(ns), 1), dtype=torch.int32, device=dev)
        self.iters = torch.zeros(max(len(ns), 1), dtype=torch.int32, device=dev)
        self.berr = torch.zeros(max(len(ns), 1), dtype=torch.float64, device=dev)
        h = ctypes.c_int64(0)
        p = lambda x: x.ctypes.data_as(ctypes.c_void_p)
        _lib.check(_lib.load().hs_dev_dsgesv_batched_plan(
            len(ns), p(ns), p(ks), p(self.a_off), p(self.b_off), p(self.p_off), a.data_ptr(), b.data_ptr(),
            self.xbuf.data_ptr(), self.a32.data_ptr(), self.r32.data_ptr(), self.perm.data_ptr(),
            self.regular.data_ptr(), self.iters.data_ptr(), self.berr.data_ptr(), max_iter, tol_per_n, panel_width,
            wgs_per_cu, ctypes.byref(h)), 'dsgesv_batched_plan')
        self.handle = h.value

    def run(self, graph=True):
        "Solves and refines every system on the current stream (asynchronous)."
        if self.handle is None: raise ValueError('the plan is closed')
        _lib.check(_lib.load().hs_dev_gesv_batched_run(self.handle, 1 if graph else 0, _stream()), 'gesv_batched_run')
        return self

    def close(self):
        if self.handle is not None:
            _lib.check(_lib.load().hs_dev_gesv_batched_free(self.handle), 'gesv_batched_free')
            self.handle = None

    def __del__(self):
        try: self.close()
        except Exception: pass

    def x(self, i):
        n, k, o = int(self.ns[i]), int(self.ks[i]), int(self.b_off[i])
        return self.xbuf[o:o + n * k].view(n, k)

    def factors32(self, i):
        n, o, po = int(self.ns[i]), int(self.a_off[i]), int(self.p_off[i])
        return self.a32[o:o + n * n].view(n, n), self.perm[po:po + n]


def pack(mats, rhs, dtype=torch.float64, device='cuda'):
    "Flat device buffers for BatchedSolve from lists of host (numpy) square matrices and right-hand sides (n×k)."
    import numpy as np
    a = np.concatenate([np.ascontiguousarray(m).reshape(-1) for m in mats]) if mats else np.zeros(0)
    b = np.concatenate([np.ascontiguousarray(x).reshape(-1) for x in rhs]) if rhs else np.zeros(0)
    ns = [m.shape[0] for m in mats]
    ks = [x.shape[1] if x.ndim == 2 else 1 for x in rhs]
    ta = torch.from_numpy(np.ascontiguousarray(a)).to(device=device, dtype=dtype)
    tb = torch.from_numpy(np.ascontiguousarray(b)).to(device=device, dtype=dtype)
    return ta, tb, ns, ks


_TUNED = {}


def tune(a, b, ns, ks, candidates=((32, 1), (16, 1), (32, 2), (16, 2)), reps=2):
    """The batched kernel's autotuner: times each (panel width, workgroups per CU) candidate on copies of the systems
    with HIP events on the current stream, best of `reps` graph replays, and returns the fastest, with the times (ms).
    Results are remembered per dtype and workload shape (count, Σn³, Σn²k)."""
    import numpy as np
    ns64 = np.asarray(ns, dtype=np.int64); ks64 = np.asarray(ks, dtype=np.int64)
    key = (a.dtype, len(ns64), int((ns64 ** 3).sum()), int((ns64 ** 2 * ks64).sum()))
    if key in _TUNED: return _TUNED[key]
    wa, wb = a.clone(), b.clone()
    times = {}
    for pw, wg in candidates:
        plan = BatchedSolve(wa, wb, ns64, ks64, panel_width=pw, wgs_per_cu=wg)
        best = float('inf')
        for _ in range(reps + 1):          # the first run captures the graph and is not counted
            wa.copy_(a); wb.copy_(b)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(); plan.run(); e1.record(); e1.synchronize()
            best = min(best, e0.elapsed_time(e1))
        times[(pw, wg)] = best
        plan.close()
    choice = min(times, key=times.get)
    _TUNED[key] = (choice, times)
    return choice, times
