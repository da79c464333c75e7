"""Dense (fold + contraction) form against the oracle and the FP64 form.

Bars: FP64 dense rates within 1e-9 relative of the oracle (the regrouped sums
round differently from the reference order); FP32 dense per-mode rates within
1e-3 relative and kappa within 1e-4 relative of FP64; the dense kernel visits
exactly the prefilter's processes.
"""
import os

import numpy as np
import pytest

from fourphonon_amd import crystal
from fourphonon_amd import runtime as rt

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def banner(msg):
    print(f"\n[gpu] {msg}", flush=True)


def rel_err(got, want):
    got, want = np.asarray(got, float).ravel(), np.asarray(want, float).ravel()
    scale = np.abs(want).max()
    den = np.maximum(np.abs(want), 1e-3 * scale)
    return float(np.max(np.abs(got - want) / den))


def kappa_rta(case, rate):
    """RTA kappa (W/m/K) summed over branches, from per-mode rates (Nb,Nlist)."""
    from oracle.oracle import iteration0, tconduct
    r = np.asarray(rate).reshape(case.modes.nbands, case.modes.Nlist, order="F")
    tau = np.where(r != 0, 1.0 / np.where(r != 0, r, 1.0), 0.0)
    F = iteration0(case, np.asfortranarray(tau))
    return tconduct(case, F).sum(axis=0)


def test_mfma_lane_layout():
    """The dense kernel's MFMA operand/result layout, checked on the device."""
    assert rt.selftest_mfma() == 0.0


@pytest.fixture(scope="module", params=[1, 2], ids=["shell1", "shell2"])
def pair(request):
    from oracle.oracle import Oracle
    case = crystal.config_case(1, ngrid=(4, 4, 4), shell=request.param)
    banner(f"context diamond 4x4x4 shell {request.param}")
    c = rt.Context(case)
    yield case, c, Oracle(case)
    c.close()


def test_dense_fp64_matches_oracle(pair):
    case, c, o = pair
    banner("dense FP64 vs oracle")
    got = c.RTA_driver_4ph(rt.PREC_FP64_DENSE)
    want = o.rta_driver(nthreads=8)
    for g, w in (("rate_scatt_plusplus", "rate_pp"), ("rate_scatt_plusminus", "rate_pm"),
                 ("rate_scatt_minusminus", "rate_mm"), ("WP4_plusminus", "wp4_pm"), ("WP4_minusminus", "wp4_mm")):
        assert rel_err(got[g].ravel(order="F"), want[w]) <= 1e-9, g


def test_dense_visits_exactly_the_prefilter_processes(pair):
    case, c, o = pair
    modes = np.arange(1, c.nmodes + 1)
    _, n64 = c.RTA_driver_4ph_modes(modes, rt.PREC_FP64)
    _, n32 = c.RTA_driver_4ph_modes(modes, rt.PREC_FP32)
    _, nb16 = c.RTA_driver_4ph_modes(modes, rt.PREC_BF16X3)
    _, nf16 = c.RTA_driver_4ph_modes(modes, rt.PREC_FP16X3)
    _, nh16 = c.RTA_driver_4ph_modes(modes, rt.PREC_FP16X2)
    assert c.dense_kernel() > 0          # the wave-owned-slice kernel with its sorted-x1 screen
    _, nd = c.RTA_driver_4ph_modes(modes, rt.PREC_FP64_DENSE)
    assert c.dense_kernel() == 0
    assert n64 == n32 == nb16 == nf16 == nh16 == nd > 0


def test_dense_fp32_tolerances(pair):
    case, c, o = pair
    banner("dense FP32 vs FP64")
    ref = c.RTA_driver_4ph(rt.PREC_FP64)
    got = c.RTA_driver_4ph(rt.PREC_FP32)
    err = rel_err(got["rate_scatt_4ph"], ref["rate_scatt_4ph"])
    print(f"per-mode rate max rel err {err:.3e}, shape {c.dense_shape()}")
    assert c.dense_shape()["mfma"]
    assert err <= 1e-3
    k64 = kappa_rta(case, ref["rate_scatt_4ph"])
    k32 = kappa_rta(case, got["rate_scatt_4ph"])
    kerr = abs(np.trace(k32) - np.trace(k64)) / abs(np.trace(k64))
    print(f"kappa rel err {kerr:.3e}")
    assert kerr <= 1e-4


@pytest.mark.parametrize("prec,mf", [(rt.PREC_BF16X3, 2), (rt.PREC_FP16X3, 3), (rt.PREC_FP16X2, 4)],
                         ids=["bf16x3", "fp16x3", "fp16x2"])
def test_dense_split_form_tolerances(pair, prec, mf):
    """BF16 and FP16 hi/lo MFMA forms (three products) and the FP16 form with
    hi/lo phase factors only (two products): per-mode rates within 1e-3 and
    kappa within 1e-4 of FP64."""
    case, c, o = pair
    banner(f"dense hi/lo form {prec} vs FP64")
    ref = c.RTA_driver_4ph(rt.PREC_FP64)
    got = c.RTA_driver_4ph(prec)
    assert c.dense_shape()["mfma"] == mf
    err = rel_err(got["rate_scatt_4ph"], ref["rate_scatt_4ph"])
    print(f"per-mode rate max rel err {err:.3e}")
    assert err <= 1e-3
    k64 = kappa_rta(case, ref["rate_scatt_4ph"])
    k16 = kappa_rta(case, got["rate_scatt_4ph"])
    kerr = abs(np.trace(k16) - np.trace(k64)) / abs(np.trace(k64))
    print(f"kappa rel err {kerr:.3e}")
    assert kerr <= 1e-4


def test_dense_golden_diamond_8cube():
    path = os.path.join(GOLDEN, "cfg1.npz")
    if not os.path.exists(path):
        pytest.skip("cfg1.npz not generated")
    g = np.load(path)
    banner("dense vs golden diamond 8x8x8")
    c = rt.Context(crystal.config_case(1))
    rm = g["rate_modes"]
    for prec, tol in ((rt.PREC_FP64_DENSE, 1e-9), (rt.PREC_FP32, 1e-3), (rt.PREC_BF16X3, 1e-3), (rt.PREC_FP16X3, 1e-3),
                      (rt.PREC_FP16X2, 1e-3)):
        r, n = c.RTA_driver_4ph_modes(rm, prec)
        assert n == int(g["N"][:, rm - 1].sum())
        for key, wk in (("rate_scatt_plusplus", "rate_pp"), ("rate_scatt_plusminus", "rate_pm"),
                        ("rate_scatt_minusminus", "rate_mm")):
            assert rel_err(r[key].ravel(order="F")[rm - 1], g[wk]) <= tol, (prec, key)
    c.close()


def test_dense_golden_diamond_16cube_window():
    path = os.path.join(GOLDEN, "cfg2.npz")
    if not os.path.exists(path):
        pytest.skip("cfg2.npz not generated")
    g = np.load(path)
    banner("dense vs golden diamond 16x16x16 shell 2 (q' window)")
    c = rt.Context(crystal.config_case(2))
    rm = g["rate_modes"]
    for prec, tol in ((rt.PREC_FP64_DENSE, 1e-9), (rt.PREC_FP32, 1e-3), (rt.PREC_BF16X3, 1e-3), (rt.PREC_FP16X3, 1e-3),
                      (rt.PREC_FP16X2, 1e-3)):
        r, _ = c.RTA_driver_4ph_modes(rm, prec, ii_lo=int(g["ii_lo"]), ii_hi=int(g["ii_hi"]))
        for key, wk in (("rate_scatt_plusplus", "rate_pp"), ("rate_scatt_plusminus", "rate_pm"),
                        ("rate_scatt_minusminus", "rate_mm")):
            assert rel_err(r[key].ravel(order="F")[rm - 1], g[wk]) <= tol, (prec, key)
    c.close()


def test_dense_golden_wurtzite_12x12x8_window():
    """Nb = 12 dense shapes (FP64 VALU, FP32 MFMA, BF16 hi/lo MFMA) against the
    oracle's rates of the 12x12x8 wurtzite, 2nd shell, on a q' window."""
    path = os.path.join(GOLDEN, "cfg3.npz")
    if not os.path.exists(path):
        pytest.skip("cfg3.npz not generated")
    g = np.load(path)
    banner("dense vs golden wurtzite 12x12x8 shell 2 (q' window)")
    c = rt.Context(crystal.config_case(3))
    rm = g["rate_modes"]
    for prec, tol in ((rt.PREC_FP64_DENSE, 1e-9), (rt.PREC_FP32, 1e-3), (rt.PREC_BF16X3, 1e-3), (rt.PREC_FP16X3, 1e-3),
                      (rt.PREC_FP16X2, 1e-3)):
        r, _ = c.RTA_driver_4ph_modes(rm, prec, ii_lo=int(g["ii_lo"]), ii_hi=int(g["ii_hi"]))
        for key, wk in (("rate_scatt_plusplus", "rate_pp"), ("rate_scatt_plusminus", "rate_pm"),
                        ("rate_scatt_minusminus", "rate_mm")):
            assert rel_err(r[key].ravel(order="F")[rm - 1], g[wk]) <= tol, (prec, key)
        if prec != rt.PREC_FP64_DENSE:
            assert c.dense_shape()["mfma"] == {rt.PREC_FP32: 1, rt.PREC_BF16X3: 2, rt.PREC_FP16X3: 3,
                                               rt.PREC_FP16X2: 4}[prec]
    c.close()


@pytest.mark.parametrize("prec", [rt.PREC_FP16X2, rt.PREC_BF16X3, rt.PREC_FP16X3, rt.PREC_FP32, rt.PREC_FP64_DENSE],
                         ids=["fp16x2", "bf16x3", "fp16x3", "fp32", "fp64"])
def test_autotuner_picks_a_built_shape_and_keeps_the_rates(prec):
    """The autotuner times the built shapes on a sample and keeps the fastest;
    later calls run that shape and give the same rates as the default shape
    (to rounding) and the FP64 form's within the form's bar.  Shapes of the
    MFMA forms may differ in their k-step grouping, so their FP32 sums
    round differently: the shape-to-shape bar there is 1e-5, far below the
    forms' 1e-3."""
    case = crystal.config_case(1, ngrid=(4, 4, 4), shell=2)
    c = rt.Context(case)
    try:
        modes = np.arange(1, c.nmodes + 1, dtype=np.int32)
        before, _ = c.RTA_driver_4ph_modes(modes, prec)
        banner(f"autotune precision {prec}")
        t = c.autotune_dense(prec, modes[::5], ii_lo=1, ii_hi=16)
        print(f"chosen IB {t['ib']}, TQ {t['tq']}, slots {t['slots']}, {t['ms']:.3f} ms")
        assert t["ib"] in (1, 2, 3, 6) and t["tq"] in (2, 4, 8) and t["slots"] >= 1 and t["ms"] > 0
        after, _ = c.RTA_driver_4ph_modes(modes, prec)
        assert c.dense_shape()["ib"] == t["ib"] and c.dense_shape()["tq"] == t["tq"]
        same_bar = 1e-9 if prec == rt.PREC_FP64_DENSE else 1e-5
        assert rel_err(after["rate_scatt_4ph"], before["rate_scatt_4ph"]) <= same_bar
        ref, _ = c.RTA_driver_4ph_modes(modes, rt.PREC_FP64)
        bar = 1e-9 if prec == rt.PREC_FP64_DENSE else 1e-3
        assert rel_err(after["rate_scatt_4ph"], ref["rate_scatt_4ph"]) <= bar
    finally:
        c.close()
