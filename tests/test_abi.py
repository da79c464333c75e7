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
