"""Workload 1 at its defined size: diamond 8x8x8 with 1st-shell third- and
fourth-order IFCs, RTA 3ph + 4ph rates of every mode and the RTA kappa,
against the oracle's all-mode fixture (tests/golden/cfg1k.npz, written by
tests/golden/make_golden.py cfg1k).

Bars: FP64 per-process rates (3ph and 4ph) and kappa within 1e-10 of the
oracle; the dense MFMA forms' 4ph rates within 1e-3 per mode and kappa
within 1e-4 of the FP64 form.
"""
import os

import numpy as np
import pytest

from fourphonon_amd import crystal
from fourphonon_amd import runtime as rt

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cfg1k.npz")


def banner(msg):
    print(f"\n[gpu] {msg}", flush=True)


def rel(got, want):
    got, want = np.asarray(got, float).ravel(), np.asarray(want, float).ravel()
    nz = want != 0
    assert np.all(got[~nz] == 0)
    return float(np.max(np.abs(got[nz] - want[nz]) / np.abs(want[nz]))) if nz.any() else 0.0


def tau_of(rate):
    r = np.asarray(rate, float)
    return np.asfortranarray(np.where(r > 0, 1.0 / np.where(r > 0, r, 1.0), 0.0))


@pytest.fixture(scope="module")
def cfg1():
    g = np.load(GOLDEN)
    case = crystal.config_case(1)
    banner("workload 1: diamond 8x8x8, 1st-shell FC3/FC4, all 174 modes")
    c = rt.Context(case)
    yield g, case, c
    c.close()


def test_rates_of_every_mode(cfg1):
    g, case, c = cfg1
    r4 = c.RTA_driver_4ph(rt.PREC_FP64)
    for dk, ok in (("rate_scatt_plusplus", "rate_pp"), ("rate_scatt_plusminus", "rate_pm"),
                   ("rate_scatt_minusminus", "rate_mm"), ("WP4_plusplus", "wp4_pp"),
                   ("WP4_plusminus", "wp4_pm"), ("WP4_minusminus", "wp4_mm")):
        err = rel(r4[dk].ravel(order="F"), g["r4_" + ok])
        print(f"4ph {dk}: max rel err {err:.2e}")
        assert err <= 1e-10, dk
    r3 = c.RTA_driver()
    for dk, ok in (("rate_scatt_plus", "rate_plus"), ("rate_scatt_minus", "rate_minus"), ("WP3_plus", "wp3_plus"),
                   ("WP3_minus", "wp3_minus")):
        err = rel(r3[dk].ravel(order="F"), g["r3_" + ok])
        print(f"3ph {dk}: max rel err {err:.2e}")
        assert err <= 1e-10, dk


def test_rta_kappa_3ph_plus_4ph(cfg1):
    """kappa_RTA from the total rate (Src/ShengBTE.f90:1220, 1232-1239) on the
    device (iteration0 + TConduct) against the oracle's from its own rates;
    the MFMA forms against the FP64 form."""
    from oracle.oracle import iteration0, tconduct
    g, case, c = cfg1
    shape = (case.modes.nbands, case.modes.Nlist)
    r3 = c.RTA_driver()["rate_scatt"]
    r4 = c.RTA_driver_4ph(rt.PREC_FP64)["rate_scatt_4ph"]
    k64 = c.TConduct(c.iteration0(tau_of(r3 + r4))).sum(axis=0)
    wrate = (g["r3_rate"] + g["r4_rate_4ph"]).reshape(shape, order="F")
    kw = tconduct(case, iteration0(case, tau_of(wrate))).sum(axis=0)
    err = float(np.abs(k64 - kw).max() / np.abs(kw).max())
    print(f"kappa_RTA xx {k64[0, 0]:.6f} W/m/K, rel err vs oracle {err:.2e}")
    assert err <= 1e-10
    for prec, name in ((rt.PREC_FP32, "fp32"), (rt.PREC_BF16X3, "bf16x3"), (rt.PREC_FP16X3, "fp16x3")):
        rp = c.RTA_driver_4ph(prec)["rate_scatt_4ph"]
        rerr = rel(rp.ravel(order="F"), r4.ravel(order="F"))
        kp = c.TConduct(c.iteration0(tau_of(r3 + rp))).sum(axis=0)
        kerr = abs(np.trace(kp) - np.trace(k64)) / abs(np.trace(k64))
        print(f"{name}: 4ph rate max rel err {rerr:.2e}, kappa rel err {kerr:.2e}")
        assert rerr <= 1e-3 and kerr <= 1e-4, name
