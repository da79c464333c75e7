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
                 "ind_driver_4ph", "vp_pp", "vp_pm", "vp_mm", "iteration0", "iteration_4ph", "tconduct",
                 "fp4_init_iterations", "fp4_solve_4ph"):
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


STUB_C = r"""
#include <stdint.h>
static int dummy, calls, nb_, np_, nl_, ev_null;
const char *fp4_last_error(void) { return "stub"; }
int fp4_ctx_create(const int32_t *cfg, void **out) { *out = &dummy; nb_ = cfg[3]; np_ = cfg[0] * cfg[1] * cfg[2]; return 0; }
int fp4_ctx_destroy(void *c) { return 0; }
int fp4_ctx_set_modes(void *c, const double *e, const double *v, const double *ev, const int32_t *ijk, int32_t nlist,
                      const int32_t *L) { nl_ = nlist; ev_null = (ev == 0); return 0; }
int fp4_np_4ph(void *c, int64_t *a, double *b, int64_t *cc, double *d, int64_t *e, double *f) {
  if (!ev_null) return 5;   /* NP_driver_4ph passes no eigenvectors (a null pointer) */
  for (int i = 0; i < nb_ * nl_; i++) { a[i] = cc[i] = e[i] = 0; b[i] = d[i] = f[i] = 0; }
  return 0;
}
int fp4_iteration0(void *c, int32_t nlist, const int32_t *a, const int32_t *b, const double *tau, double *F) {
  for (int t = 0; t < 3 * nb_ * np_; t++) F[t] = 1.0;
  return 0;
}
int fp4_tconduct(void *c, double V, const double *F, double *kappa, double *mode) {
  for (int t = 0; t < 9 * nb_; t++) kappa[t] = t + 1 + 100 * calls;
  calls++;
  return 0;
}
int fp4_bte_solve(void *c, double V, const double *tau, int32_t maxiter, double eps, double *F, double *k, double *h,
                  double *r, int32_t *n) {
  *n = 2;
  for (int t = 0; t < 27; t++) h[t] = 7.0 + t;
  for (int t = 0; t < 3; t++) r[t] = 0.5 * t;
  for (int t = 0; t < 9 * nb_; t++) k[t] = -1.0 - t;
  return 0;
}
"""


@pytest.mark.skipif(not os.path.exists(FLANG), reason="flang not available")
def test_driver_iterate_outputs_carry_the_abi_results(tmp_path):
    """The driver's iterate mode writes exactly what the C ABI returned
    (checked against a C stub of the library, no device needed)."""
    import numpy as np
    from fourphonon_amd.fortran.io import read_iteration_outputs
    from fourphonon_amd.runtime import EXPORTS
    objdir = os.path.join(build_fortran.PKG, "_build", "fortran")
    if not (os.path.exists(os.path.join(objdir, "fp4_processes.o")) and
            os.path.exists(os.path.join(objdir, "fp4_driver.o"))):
        pytest.skip("Fortran objects not built")
    have = {"fp4_last_error", "fp4_ctx_create", "fp4_ctx_destroy", "fp4_ctx_set_modes", "fp4_np_4ph",
            "fp4_iteration0", "fp4_tconduct", "fp4_bte_solve"}
    src = STUB_C + "".join(f"int {n}() {{ return 0; }}\n" for n in EXPORTS if n not in have)
    (tmp_path / "stub.c").write_text(src)
    subprocess.run(["gcc", "-w", "-shared", "-fPIC", "-o", str(tmp_path / "libfp4.so"), str(tmp_path / "stub.c")],
                   check=True)
    drv = str(tmp_path / "drv")
    subprocess.run([FLANG, os.path.join(objdir, "fp4_processes.o"), os.path.join(objdir, "fp4_driver.o"), "-o", drv,
                    "-L" + str(tmp_path), "-lfp4", "-Wl,-rpath," + str(tmp_path)], check=True)
    case = crystal.config_case(1, ngrid=(2, 2, 2))
    write_inputs(case, str(tmp_path / "in"))
    os.makedirs(tmp_path / "out")
    p = subprocess.run([drv, str(tmp_path / "in"), str(tmp_path / "out"), "fp64", "iterate"], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    nb, nptk = case.modes.nbands, case.modes.nptk
    got = read_iteration_outputs(str(tmp_path / "out"), nb, nptk)
    base = np.arange(1, 9 * nb + 1, dtype=float).reshape((nb, 3, 3), order="F")
    for s in range(4):
        assert np.array_equal(got["kappa_sweeps"][..., s], base + 100 * s)
    assert got["niter"] == 2
    assert np.array_equal(got["kappa_iter"].ravel(order="F"), 7.0 + np.arange(27))
    assert np.array_equal(got["relchange"], [0.0, 0.5, 1.0])
    assert np.array_equal(got["kappa"].ravel(order="F"), -1.0 - np.arange(9 * nb))
