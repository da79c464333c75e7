"""The Fortran driver, through the drop-in module fp4_processes, reproduces
the oracle: counts and lists bit-exact, FP64 rates to 1e-10, and writes
BTE.Numprocess_4ph / BTE.w_4ph in the reference's formats."""
import os
import subprocess

import numpy as np
import pytest

from fourphonon_amd import crystal
from fourphonon_amd.fortran import build_fortran
from fourphonon_amd.fortran.io import read_iteration_outputs, read_outputs, write_inputs

pytestmark = pytest.mark.gpu


def run_driver(case, tmp, mode, lists=False, extra=None):
    ind, outd = os.path.join(tmp, "in"), os.path.join(tmp, "out_" + mode)
    write_inputs(case, ind)
    os.makedirs(outd, exist_ok=True)
    args = [build_fortran.DRIVER, ind, outd, mode] + (["list"] if lists else []) + ([extra] if extra else [])
    print("\n[gpu] fortran driver", " ".join(args[2:]), flush=True)
    p = subprocess.run(args, capture_output=True, text=True, timeout=240)
    assert p.returncode == 0, p.stderr
    return outd


def test_fortran_driver_matches_oracle(tmp_path):
    from oracle.oracle import Oracle
    case = crystal.config_case(1, ngrid=(4, 4, 4))
    o = Oracle(case)
    nm = o.nmodes
    outd = run_driver(case, str(tmp_path), "fp64", lists=True)
    out = read_outputs(outd, nm, lists=True)
    No, _ = o.np_driver()
    for ch in range(3):
        assert np.array_equal(out["N"][ch], No[ch])
    want = o.ind_only_driver(No)
    for ch, name in enumerate(("pp", "pm", "mm")):
        assert np.array_equal(out["lists_" + name], np.stack(want[ch]))
    r = o.rta_driver(nthreads=8)
    got = out["rates"]
    for t, key in enumerate(("rate_4ph", "rate_pp", "rate_pm", "rate_mm", "wp4_pp", "wp4_pm", "wp4_mm")):
        w = r[key]
        scale = np.abs(w).max()
        assert np.all(np.abs(got[t] - w) <= 1e-10 * np.maximum(np.abs(w), 1e-3 * scale)), key
    # reference file formats
    lines = open(os.path.join(outd, "BTE.Numprocess_4ph")).read().split("\n")
    first = [int(x) for x in lines[0].split()]
    assert first == [int(No[0][0]), int(No[1][0]), int(No[2][0])]
    w4 = np.loadtxt(os.path.join(outd, "BTE.w_4ph"))
    assert w4.shape == (nm, 2)
    # BTE.w_4ph lines run over (branch, point); rate arrays are (Nb, Nlist)
    assert np.allclose(w4[:, 1], got[0].reshape(case.nbands, -1, order="F").ravel(), rtol=1e-9, atol=0)


def test_fortran_driver_fp32_dense(tmp_path):
    from oracle.oracle import Oracle
    case = crystal.config_case(1, ngrid=(4, 4, 4), shell=2)
    o = Oracle(case)
    outd = run_driver(case, str(tmp_path), "fp32")
    got = read_outputs(outd, o.nmodes)["rates"][0]
    want = o.rta_driver(nthreads=8)["rate_4ph"]
    scale = np.abs(want).max()
    assert np.max(np.abs(got - want) / np.maximum(np.abs(want), 1e-3 * scale)) <= 1e-3


def test_fortran_driver_iterations(tmp_path):
    """iteration0 / iteration_4ph / TConduct drop-ins and the device loop,
    called from Fortran, against the oracle's restatement of the same loop."""
    from oracle.oracle import (Oracle, bte_solve, iteration0, iteration_4ph, symmetrize_F, tconduct)
    case = crystal.config_case(1, ngrid=(4, 4, 4))
    o = Oracle(case)
    m = case.modes
    outd = run_driver(case, str(tmp_path), "fp64", extra="iterate")
    got = read_iteration_outputs(outd, m.nbands, m.nptk)
    No, _ = o.np_driver()
    lists, gammas, rates = o.ind_driver(No)
    r = rates["rate_4ph"].reshape(m.nbands, m.Nlist, order="F")
    tau = np.where(r != 0, 1.0 / np.where(r != 0, r, 1.0), 0.0)
    assert np.allclose(got["tau"], tau, rtol=1e-10, atol=0)
    tau = got["tau"]                                   # same tau on both sides from here on
    F = iteration0(case, tau)
    kap = [tconduct(case, F)]
    for _ in range(3):
        F = symmetrize_F(case, iteration_4ph(case, No, lists, gammas, tau, F))
        kap.append(tconduct(case, F))
    scale = np.abs(F).max()
    assert np.abs(got["F_n"] - F).max() <= 1e-9 * scale
    want = np.stack(kap, axis=-1)
    assert np.abs(got["kappa_sweeps"] - want).max() <= 1e-9 * np.abs(want).max()
    sol = bte_solve(case, No, lists, gammas, tau, maxiter=100)
    print(f"niter {got['niter']} (oracle {sol['niter']})")
    assert got["niter"] == sol["niter"]
    assert np.abs(got["kappa_iter"] - sol["kappa_iter"]).max() <= 1e-9 * np.abs(sol["kappa_iter"]).max()
    assert np.abs(got["kappa"] - sol["kappa"]).max() <= 1e-9 * np.abs(sol["kappa"]).max()
