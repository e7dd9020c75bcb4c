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


def getrf(a):
    """Factors the square row-major tensor `a` in place into rustymath's column-major packed factors and returns
    (a, perm) with perm an int64 GPU tensor (core/src/linalg/factor.rs:617-630)."""
    _mat(a, 'a')
    n = a.shape[0]
    if a.shape[1] != n: raise ValueError('getrf needs a square matrix')
    perm = torch.empty(n, dtype=torch.int64, device=a.device)
    f = getattr(_lib.load(), f'hs_dev_{_sfx(a.dtype)}getrf')
    _lib.check(f(n, a.data_ptr(), a.stride(0), perm.data_ptr(), _stream()), 'getrf')
    return a, perm


def getrs(factors, perm, b, trans=False):
    "Solves A x = b (or Aᵀ x = b) in place on the row-major b (n×k) with factors from getrf (factor.rs:633-641)."
    _mat(factors, 'factors'); _mat(b, 'b')
    n, k = b.shape
    f = getattr(_lib.load(), f'hs_dev_{_sfx(b.dtype)}getrs')
    _lib.check(f(1 if trans else 0, n, k, factors.data_ptr(), factors.stride(0), perm.data_ptr(), b.data_ptr(),
                 b.stride(0), _stream()), 'getrs')
    return b


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
