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
    assert lib.hs_dlu_factor(None, -1, None, None, None) == -2


def test_linalg_validates_shapes_on_the_host():
    import numpy as np
    from rustymath_amd import linalg as la
    with pytest.raises(ValueError): la.matmul(np.ones((2, 3)), np.ones((4, 2)))
    with pytest.raises(ValueError): la.lu(np.ones((2, 3)))
    with pytest.raises(IndexError): la.lu_solve(np.eye(3), np.array([-1, 0, 1]), np.ones(3))
    with pytest.raises(TypeError): la.matmul(np.ones((2, 2), complex), np.ones((2, 2)))
