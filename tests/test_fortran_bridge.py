"""The Fortran drop-in module compiles and links against the C ABI, and the
driver fails loudly (error stop, non-zero exit, message on stderr) when no
device is available."""
import os
import shutil
import subprocess

import pytest

from fourphonon_amd import crystal
from fourphonon_amd.fortran import build_fortran
from fourphonon_amd.fortran.io import write_inputs

FLANG = build_fortran.FLANG


@pytest.mark.skipif(not os.path.exists(FLANG), reason="flang not available")
def test_bridge_builds_and_links():
    if not os.path.exists(os.path.join(build_fortran.PKG, "libfp4.so")):
        pytest.skip("libfp4.so not built")
    path = build_fortran.build()
    assert os.path.exists(path)
    syms = subprocess.run(["nm", "-C", os.path.join(build_fortran.PKG, "_build", "fortran", "fp4_processes.o")],
                          capture_output=True, text=True).stdout.lower()
    for name in ("np_driver_4ph", "ind_only_driver_4ph", "rta_driver_4ph", "rta_driver_4ph_gpu_using_ind",
                 "ind_driver_4ph", "vp_pp", "vp_pm", "vp_mm"):
        assert name in syms, name


def test_driver_reports_missing_device(tmp_path):
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present; covered by the gpu tests")
    except Exception:
        pass
    if not os.path.exists(build_fortran.DRIVER):
        pytest.skip("driver not built")
    case = crystal.config_case(1, ngrid=(2, 2, 2))
    write_inputs(case, str(tmp_path / "in"))
    os.makedirs(tmp_path / "out")
    p = subprocess.run([build_fortran.DRIVER, str(tmp_path / "in"), str(tmp_path / "out")], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode != 0
    assert "fp4_ctx_create failed" in p.stderr
