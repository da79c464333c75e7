"""The device count, phase-space and weighted-phase-space drivers (three and
four phonons) against the reference's own Si example run at 16x16x16
(tests/golden/si_example.npz).  Agreement is statistical for the reasons
given in tests/test_reference_example.py: channel totals within 0.5 %,
per-mode medians within 1 % (counts, phase space) and a few percent
(weighted phase space, printed with five digits)."""
import numpy as np
import pytest

from fourphonon_amd import runtime as rt
from tests.si_example import load, si_case

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return abs(float(a) - float(b)) / abs(float(b))


def _median_rel(got, want):
    nz = want != 0
    return float(np.median(np.abs(got[nz] - want[nz]) / np.abs(want[nz])))


@pytest.fixture(scope="module")
def si():
    g = load()
    print("\n[gpu] Si example 16x16x16 context", flush=True)
    c = rt.Context(si_case(g))
    yield g, c
    c.close()


def test_four_phonon_counts_and_phase_space(si):
    g, c = si
    N = c.NP_driver_4ph()
    for ch in range(3):
        got, want = N[2 * ch], g["N4"][ch]
        P, Pw = N[2 * ch + 1].ravel(order="F"), g["P4"][ch].ravel(order="F")
        print(f"4ph channel {ch}: counts {int(got.sum())} vs {int(want.sum())}, "
              f"median rel {_median_rel(got, want):.2e}; P4 total rel {_rel(P.sum(), Pw.sum()):.2e}")
        assert _rel(got.sum(), want.sum()) <= 5e-3
        assert _rel(P.sum(), Pw.sum()) <= 5e-3
        assert _median_rel(got, want) <= 1e-2
        assert _median_rel(P, Pw) <= 1e-2


def test_three_phonon_counts_and_phase_space(si):
    g, c = si
    N = c.NP_driver()
    for ch in range(2):
        assert _rel(N[2 * ch].sum(), g["Ntotal3"][ch]) <= 2e-3
        P, Pw = N[2 * ch + 1].ravel(order="F"), g["P3"][ch].ravel(order="F")
        assert _rel(P.sum(), Pw.sum()) <= 2e-3
        assert _median_rel(P, Pw) <= 3e-2


def test_weighted_phase_space(si):
    g, c = si
    modes = np.arange(1, c.nmodes + 1)[::10]
    r, n = c.RTA_driver_4ph_modes(modes, rt.PREC_FP64)
    assert n > 0
    for ch, k in enumerate(("WP4_plusplus", "WP4_plusminus", "WP4_minusminus")):
        got, want = r[k].ravel(order="F")[modes - 1], g["WP4"][ch][modes - 1]
        print(f"WP4 channel {ch}: total rel {_rel(got.sum(), want.sum()):.2e}, median rel {_median_rel(got, want):.2e}")
        assert _rel(got.sum(), want.sum()) <= 3e-2 and _median_rel(got, want) <= 3e-2
    r3 = c.RTA_driver()
    for ch, k in enumerate(("WP3_plus", "WP3_minus")):
        got, want = r3[k].ravel(order="F"), g["WP3"][ch]
        assert _rel(got.sum(), want.sum()) <= 5e-2 and _median_rel(got, want) <= 1e-1
