"""rustymath_amd: rustymath's dense linear-algebra hot path (matrix products, LU with partial pivoting, solves)
as hand-written HIP kernels for MI355X (gfx950) behind a C ABI (include/hip_sys.h, libhip_sys.so).

`rustymath_amd.linalg` mirrors rustymath's Python names (`matmul`, `linalg.solve`, `linalg.lu`, `Lu`);
`rustymath_amd.device` keeps data resident in HBM as torch tensors for the mixed-precision, batched and adjoint
forms.
"""
from . import linalg
from .linalg import matmul

__all__ = ['linalg', 'matmul']
