"""The C ABI without a GPU: the library loads, and exports every entry point that include/hip_sys.h declares."""
import ctypes, os, re
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, 'include', 'hip_sys.h')).read()
    return sorted(set(re.findall(r'^\s*(?:int|const char\*)\s+(hs_\w+)\(', text, re.M)))


def test_header_declares_the_entry_points():
    names = declared()
    assert 'hs_dgemm' in names and 'hs_dlu_factor' in names and 'hs_dlu_solve' in names and len(names) >= 17


def test_library_exports_every_declared_symbol():
    from rustymath_amd import _lib
    lib = _lib.load()
    for name in declared():
        assert hasattr(lib, name), name
    assert set(declared()) <= set(_lib.SIGNATURES)


def test_bad_arguments_are_rejected_before_any_device_work():
    "Argument checks run on the host and name the argument, LAPACK style, without touching a GPU."
    from rustymath_amd import _lib
    lib = _lib.load()
    assert lib.hs_dgemm(7, 111, 2, 2, 2, 1.0, None, 2, None, 2, 0.0, None, 2) == -1
    assert lib.hs_dgemm(111, 111, -1, 2, 2, 1.0, None, 2, None, 2, 0.0, None, 2) == -3
    assert lib.hs_dgemm(111, 111, 4, 4, 4, 1.0, None, 3, None, 4, 0.0, None, 4) == -8
    assert lib.hs_dgemm(111, 111, 0, 4, 4, 1.0, None, 4, None, 4, 0.0, None, 4) == 0
    assert lib.hs_dlu_factor(None, -1, None, None, None, None) == -2
    bad = (ctypes.c_int64 * 3)(0, 3, 1)
    assert lib.hs_dlu_inverse(None, None, -1, None) == -3
    assert lib.hs_dlu_inverse(None, bad, 3, None) == -2
    assert lib.hs_dinverse(None, -1, None, None, None) == -2
    assert lib.hs_dsolve_stack(-1, 2, 1, None, None, None, None) == -1
    assert lib.hs_dsolve_stack(2, -2, 1, None, None, None, None) == -2
    assert lib.hs_ssolve_stack(2, 2, -1, None, None, None, None) == -3
    assert lib.hs_dlu_stack(2, -1, None, None, None) == -2
    assert lib.hs_dlu_stack(0, 5, None, None, None) == 0
    assert lib.hs_dcholesky(None, -1, None, None) == -2
    assert lib.hs_dev_dpotrf(-1, None, 1, None, None) == -1
    assert lib.hs_dev_spotrf(4, None, 3, None, None) == -3
    assert lib.hs_dev_dgetri(-1, None, 1, None, None, 1, None) == -1
    assert lib.hs_set_exchange(2, None) == -1 and lib.hs_set_exchange(-1, None) == -1
    assert lib.hs_dev_dgetri(4, None, 3, None, None, 4, None) == -3
    assert lib.hs_dev_sgetri(4, None, 4, None, None, 2, None) == -6


def test_linalg_validates_shapes_on_the_host():
    import numpy as np
    from rustymath_amd import linalg as la
    with pytest.raises(ValueError): la.matmul(np.ones((2, 3)), np.ones((4, 2)))
    with pytest.raises(ValueError): la.lu(np.ones((2, 3)))
    with pytest.raises(IndexError): la.lu_solve(np.eye(3), np.array([-1, 0, 1]), np.ones(3))
    with pytest.raises(IndexError): la.Lu(np.eye(3), np.array([-1, 0, 1])).inverse()  # tests/test_linalg.py:32
    with pytest.raises(NotImplementedError): la.pinv(np.ones((2, 3)))
    with pytest.raises(ValueError): la.det(np.ones((2, 3)))
    with pytest.raises(ValueError): la.cholesky(np.ones((2, 3)))
    with pytest.raises(TypeError): la.matmul(np.ones((2, 2), complex), np.ones((2, 2)))


def test_batched_plan_rejects_bad_systems_on_the_host():
    "The batched plan checks every system's order, column count and offsets, and its tuning knobs, before any HIP call."
    import numpy as np
    from rustymath_amd import _lib
    lib = _lib.load()
    p = lambda x: x.ctypes.data_as(ctypes.c_void_p)
    buf = np.zeros(64)
    ptr, reg = p(buf), p(np.zeros(4, np.int32))
    plan = ctypes.c_int64(0)
    def call(n, k, off=0, pw=32, wg=1, count=1):
        n = np.array([n], np.int64); k = np.array([k], np.int64); o = np.array([off], np.int64)
        return lib.hs_dev_dgesv_batched_plan(count, p(n), p(k), p(o), p(o), p(o), ptr, ptr, ptr, reg, mode, pw, wg,
                                             ctypes.byref(plan))
    mode = 0
    assert call(4, 1, count=-1) == -1
    assert call(513, 1) == -2
    assert call(4, 9) == -3
    assert call(4, 1, off=-1) == -4
    assert call(4, 1, pw=24) == -12
    assert call(4, 1, wg=3) == -13
    mode = 3
    assert call(4, 1) == -11
    assert lib.hs_dev_gesv_batched_free(12345) == -1


def test_mixed_precision_solve_rejects_bad_arguments_on_the_host():
    from rustymath_amd import _lib
    lib = _lib.load()
    assert lib.hs_dev_dsgesv(-1, None, 1, None, None, 10, 0.0, None, None, None) == -1
    assert lib.hs_dev_dsgesv(4, None, 5, None, None, 10, 0.0, None, None, None) == -3
    assert lib.hs_dev_dsgesv(4, None, 4, None, None, 0, 0.0, None, None, None) == -6
    assert lib.hs_dev_dbgesv(-1, None, 1, None, None, 10, 40, 3, 0.0, None, None, None, None) == -1
    assert lib.hs_dev_dbgesv(4, None, 5, None, None, 10, 40, 3, 0.0, None, None, None, None) == -3
    assert lib.hs_dev_dbgesv(4, None, 4, None, None, 0, 40, 3, 0.0, None, None, None, None) == -6
    assert lib.hs_dev_dbgesv(4, None, 4, None, None, 10, 0, 3, 0.0, None, None, None, None) == -7
    assert lib.hs_dev_dbgesv(4, None, 4, None, None, 10, 40, 2, 0.0, None, None, None, None) == -8
    assert lib.hs_dev_sgemm_bf16_update(-1, 1, 1, None, 1, None, 1, None, 1, 1, None) == -1
    assert lib.hs_dev_sgemm_bf16_update(2, 3, 4, None, 3, None, 3, None, 3, 1, None) == -5
    assert lib.hs_dev_sgemm_bf16_update(2, 3, 4, None, 4, None, 3, None, 3, 2, None) == -10


def test_triangular_solve_rejects_bad_codes_on_the_host():
    from rustymath_amd import _lib
    lib = _lib.load()
    assert lib.hs_dtrsm(140, 121, 111, 131, 2, 2, None, 2, None, 2) == -1
    assert lib.hs_dtrsm(141, 120, 111, 131, 2, 2, None, 2, None, 2) == -2
    assert lib.hs_dtrsm(141, 121, 110, 131, 2, 2, None, 2, None, 2) == -3
    assert lib.hs_strsm(141, 121, 111, 133, 2, 2, None, 2, None, 2) == -4
    assert lib.hs_dtrsm(141, 121, 111, 131, -1, 2, None, 2, None, 2) == -5
    assert lib.hs_dtrsm(142, 121, 111, 131, 2, -1, None, 2, None, 2) == -6
    assert lib.hs_dtrsm(142, 121, 111, 131, 2, 5, None, 4, None, 5) == -8   # right side: a is n×n
    assert lib.hs_dtrsm(141, 121, 111, 131, 5, 3, None, 5, None, 2) == -10
    assert lib.hs_dtrsm(141, 121, 111, 131, 0, 3, None, 0, None, 3) == 0
    assert lib.hs_dev_dgetrs(0, 4, 1, None, 3, None, None, 1, None) == -5
    assert lib.hs_dev_sgetrs(0, 4, 2, None, 4, None, None, 1, None) == -8


def test_batched_products_reject_bad_arguments_on_the_host():
    from rustymath_amd import _lib
    lib = _lib.load()
    assert lib.hs_dgemm_batched(111, 111, 2, 2, 2, 1.0, None, 2, 4, None, 2, 4, 0.0, None, 2, 4, -1) == -17
    assert lib.hs_dgemm_batched(111, 111, 2, 2, 2, 1.0, None, 1, 4, None, 2, 4, 0.0, None, 2, 4, 3) == -8
    assert lib.hs_dgemm_batched(111, 111, 2, 2, 2, 1.0, None, 2, 4, None, 1, 4, 0.0, None, 2, 4, 3) == -11
    assert lib.hs_sgemm_batched(111, 111, 2, 2, 2, 1.0, None, 2, 4, None, 2, 4, 0.0, None, 1, 4, 3) == -15
    assert lib.hs_dgemm_batched(111, 111, 2, 2, 2, 1.0, None, 2, -4, None, 2, 4, 0.0, None, 2, 4, 3) == -9
    assert lib.hs_dgemm_batched(111, 111, 2, 2, 2, 1.0, None, 2, 4, None, 2, 4, 0.0, None, 2, 3, 3) == -16
    assert lib.hs_dev_dgemm_batched(111, 111, 2, 2, 2, 1.0, None, 2, 4, None, 2, 4, 0.0, None, 2, 1, 3, None) == -16
    assert lib.hs_dgemm_batched(111, 111, 2, 2, 2, 1.0, None, 2, 4, None, 2, 4, 0.0, None, 2, 4, 0) == 0


def header_params():
    "name -> number of parameters, for every entry point that include/hip_sys.h declares."
    text = re.sub(r'/\*.*?\*/', '', open(os.path.join(ROOT, 'include', 'hip_sys.h')).read(), flags=re.S)
    out = {}
    for name, args in re.findall(r'^\s*(?:int|const char\*)\s+(hs_\w+)\(([^;]*?)\);', text, re.M | re.S):
        args = ' '.join(args.split())
        out[name] = 0 if args in ('', 'void') else len(args.split(','))
    return out


def test_bindings_match_the_header():
    """The ctypes table (rustymath_amd/_lib.py) and the hip-sys crate (hip-sys/src/lib.rs, the Rust binding a
    rustymath build would link) declare every entry point of the header with its number of parameters."""
    from rustymath_amd import _lib
    want = header_params()
    assert len(want) >= 50
    for name, count in want.items():
        assert len(_lib.SIGNATURES[name][1]) == count, name
    rs = open(os.path.join(ROOT, 'hip-sys', 'src', 'lib.rs')).read()
    got = {m.group(1): (0 if not m.group(2).strip() else m.group(2).count(':'))
           for m in re.finditer(r'pub fn (hs_\w+)\(([^)]*)\)', rs)}
    assert got == want


def test_uneven_batched_solve_rejects_bad_arguments_on_the_host():
    "hs_*gesv_batched checks its arguments in order before any device work (LAPACK-style negative positions)."
    import numpy as np
    from rustymath_amd import _lib
    lib = _lib.load()
    n, k = np.array([3, 2], np.int64), np.array([1, 1], np.int64)
    for f in (lib.hs_dgesv_batched, lib.hs_sgesv_batched):
        assert f(-1, None, None, None, None, None, None) == -1
        assert f(0, None, None, None, None, None, None) == 0
        assert f(2, None, k.ctypes.data, None, None, None, None) == -2
        assert f(2, np.array([3, -1], np.int64).ctypes.data, k.ctypes.data, None, None, None, None) == -2
        assert f(2, n.ctypes.data, np.array([1, -2], np.int64).ctypes.data, None, None, None, None) == -3
        assert f(2, n.ctypes.data, k.ctypes.data, None, None, None, None) == -4
