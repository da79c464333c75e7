"""The device drivers against the reference's own BAs example run
(tests/golden/bas_example.npz: zincblende BAs, F-43m, 2 atoms, 6 branches,
16x16x16, phonopy 2nd-order constants with the non-analytic term of its Born
charges), with omega, v and e from the phonopy front end.

The example's folder holds two runs (see tests/test_reference_bas.py): its
per-mode 4ph counts (BTE.Numprocess_4ph) are the scalebroad = 0.1 run's,
compared here for all 870 modes; its log.out 3ph totals, 3ph rates and WP3
(T300K/) are the scalebroad = 1.0 run's, compared with a second context at
1.0.  That run's 4ph totals and WP4 come from a 4ph window other than the
present source's (they match neither scalebroad) and are not compared.
Bars are the printing precision of the files (ten digits; WP five)."""
import dataclasses

import numpy as np
import pytest

from fourphonon_amd import runtime as rt
from tests.bas_example import bas_case, load
from tests.si_example import degenerate_sets

pytestmark = pytest.mark.gpu


def _rel(got, want, floor=0.0):
    got, want = np.atleast_1d(np.asarray(got, float)), np.atleast_1d(np.asarray(want, float))
    nz = want != 0
    if not nz.any():
        return float(np.abs(got).max()) if got.size else 0.0
    return float(np.max(np.abs(got[nz] - want[nz]) / np.maximum(np.abs(want[nz]), floor)))


@pytest.fixture(scope="module")
def bas():
    g = load()
    case = bas_case(g)
    print("\n[gpu] BAs example 16x16x16 context, scalebroad 0.1", flush=True)
    c = rt.Context(case)
    yield g, case, c
    c.close()


@pytest.fixture(scope="module")
def bas1(bas):
    g, case, _ = bas
    case1 = dataclasses.replace(case, scalebroad=1.0)
    print("\n[gpu] BAs example 16x16x16 context, scalebroad 1.0", flush=True)
    c = rt.Context(case1)
    yield g, case1, c
    c.close()


def test_four_phonon_counts_all_modes(bas):
    """Counts per mode equal the run's in ++ for every mode, and in +- and
    -- but for candidates within rounding of the window edge: omega and v
    agree with the run's printed values to about 1e-10, so a few of the
    5.9e9 +- candidates flip (measured: 64 modes, at most 30 on one of the
    degenerate optical modes, the total by 34).  The run's P4 files are
    Infinity wherever a mode has processes (its build's 1/nptk**3 wrapped
    to zero), so phase space is checked for being finite and positive
    exactly where the count is."""
    g, case, c = bas
    N = c.NP_driver_4ph()
    stats = []
    for ch in range(3):
        got, want = N[2 * ch].astype(np.int64), g["N4"][ch]
        d = np.abs(got - want)
        print(f"4ph channel {ch}: {int(got.sum())} vs {int(want.sum())}, {int((d > 0).sum())} modes differ, "
              f"max {int(d.max())}", flush=True)
        stats.append((got, want, d, N[2 * ch + 1].ravel(order="F")))
    for ch, (got, want, d, P) in enumerate(stats):
        if ch == 0:
            assert np.array_equal(got, want)
        assert int((d > 0).sum()) <= 100 and int(d.max()) <= 40, ch
        assert abs(int(got.sum()) - int(want.sum())) <= 1e-8 * int(want.sum()), ch
        assert np.all(np.isfinite(P)) and np.array_equal(P > 0, got > 0), ch


def test_three_phonon_counts_rates_and_wp3_at_scalebroad_1(bas1):
    """The bars of the oracle's test of the same files: most modes at the
    printing precision, all within 2e-5 (the run's eigenvectors at 41
    zone-boundary ties are not fixed by its printed omega and v)."""
    g, case, c = bas1
    N = c.NP_driver()
    assert [int(N[2 * ch].sum()) for ch in range(2)] == [int(x) for x in g["Ntotal3"]]
    r = c.RTA_driver()
    sets = degenerate_sets(case)
    single = np.array([s[0] for s in sets if len(s) == 1])
    for k, key in enumerate(("rate_scatt_plus", "rate_scatt_minus", "rate_scatt")):
        got, ref = r[key].ravel(order="F"), g["W3"][k]
        floor = 1e-10 * np.abs(ref).max()
        e = np.abs(got[single - 1] - ref[single - 1]) / np.maximum(np.abs(ref[single - 1]), floor)
        print(f"{key}: non-degenerate modes median rel {np.median(e):.2e}, max {e.max():.2e}", flush=True)
        assert np.median(e) <= 1e-8 and e.max() <= 2e-5, key
        assert abs(got.sum() / ref.sum() - 1.0) <= 1e-7, key
        for s in sets:
            if len(s) > 1:
                assert _rel(got[s - 1].sum(), ref[s - 1].sum(), floor) <= 5e-6, (key, s)
    for ch, key in enumerate(("WP3_plus", "WP3_minus")):
        assert _rel(r[key].ravel(order="F"), g["WP3"][ch], 1e-10 * np.abs(g["WP3"][ch]).max()) <= 2e-4, key

