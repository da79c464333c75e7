"""CPU checks of the C ABI: the library loads, exports every declared symbol,
and fails cleanly (an error code, not a crash) where no GPU is present."""
import ctypes as C
import os
import re

import pytest

from fourphonon_amd import runtime as rt

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "fp4.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|const char \*)\s*(fp4_\w+)\(", text, flags=re.M)))


def test_header_and_python_mirror_agree():
    assert declared_symbols() == sorted(rt.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(rt.LIB_PATH):
        pytest.skip("libfp4.so not built")
    lib = C.CDLL(rt.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_context_creation_reports_missing_device_or_succeeds(small_case):
    if not os.path.exists(rt.LIB_PATH):
        pytest.skip("libfp4.so not built")
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present; covered by the gpu tests")
    except Exception:
        pass
    with pytest.raises(rt.FP4Error) as e:
        rt.Context(small_case)
    assert e.value.code == 2


def test_bad_config_is_rejected_before_touching_the_device(small_case):
    if not os.path.exists(rt.LIB_PATH):
        pytest.skip("libfp4.so not built")
    lib = rt.load()
    cfg = rt.Config()
    cfg.nbands, cfg.natoms = 7, 2                      # not 3*natoms
    h = C.c_void_p()
    assert lib.fp4_ctx_create(C.byref(cfg), C.byref(h)) == 1
    assert b"3*natoms" in lib.fp4_last_error()


def test_phonopy_eigen_dm_checks_arguments_before_touching_the_device():
    if not os.path.exists(rt.LIB_PATH):
        pytest.skip("libfp4.so not built")
    import numpy as np
    lib = rt.load()
    pos = np.zeros((3, 2), order="F")
    masses = np.array([10.8, 74.9])
    types = np.array([1, 2], np.int32)
    h = rt.Harmonic()
    h.ngrid[:] = [2, 2, 2]
    h.natoms, h.nelements = 2, 2
    h.scell[:] = [1, 1, 1]
    h.lattvec[:] = [0.0, 0.25, 0.25, 0.25, 0.0, 0.25, 0.25, 0.25, 0.0]
    h.positions = pos.ctypes.data_as(C.c_void_p)
    h.masses = masses.ctypes.data_as(C.c_void_p)
    h.types = types.ctypes.data_as(C.c_void_p)
    fc = np.zeros(2 * 3 * 2 * 3)
    born = np.zeros(18)
    out = [np.zeros(8 * 6), np.zeros(2 * 8 * 36), np.zeros(8 * 18)]
    p = [x.ctypes.data_as(C.c_void_p) for x in out]
    # born without epsilon
    rc = lib.fp4_eigen_dm_phonopy(C.byref(h), fc.ctypes.data_as(C.c_void_p), born.ctypes.data_as(C.c_void_p),
                                  None, None, None, *p)
    assert rc == 1 and b"born and epsilon" in lib.fp4_last_error()
    # a type beyond nelements
    types[1] = 3
    rc = lib.fp4_eigen_dm_phonopy(C.byref(h), fc.ctypes.data_as(C.c_void_p), None, None, None, None, *p)
    assert rc == 1 and b"types" in lib.fp4_last_error()
    with pytest.raises(ValueError):
        rt.eigen_dm_phonopy(None, fc, (1, 1, 1), (2, 2, 2), born=np.zeros((3, 3, 2)))
