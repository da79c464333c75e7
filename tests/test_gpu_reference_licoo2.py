"""The device drivers against the reference's own LiCoO2 example run
(tests/golden/licoo2_example.npz: 4 atoms, 12 branches, R-3m, 10x10x10,
scalebroad 0.01, phonopy 2nd-order constants): 4ph counts per mode for all
1392 modes (BTE.Numprocess_4ph, the only 12-branch four-phonon counts the
reference ships) and phase space (BTE.P4_*), 3ph counts, phase space and
RTA rates (BTE.P3_*, T300K/BTE.w_3ph*), and the weighted phase space, with
omega, v and e from the phonopy front end.  Bars are the printing precision
of the files (ten digits; WP five), as in tests/test_reference_licoo2.py."""
import numpy as np
import pytest

from fourphonon_amd import runtime as rt
from tests.licoo2_example import licoo2_case, load
from tests.si_example import degenerate_sets

pytestmark = pytest.mark.gpu


def _rel(got, want, floor=0.0):
    got, want = np.atleast_1d(np.asarray(got, float)), np.atleast_1d(np.asarray(want, float))
    nz = want != 0
    if not nz.any():
        return float(np.abs(got).max()) if got.size else 0.0
    return float(np.max(np.abs(got[nz] - want[nz]) / np.maximum(np.abs(want[nz]), floor)))


@pytest.fixture(scope="module")
def licoo2():
    g = load()
    case = licoo2_case(g)
    print("\n[gpu] LiCoO2 example 10x10x10 context (Nb = 12)", flush=True)
    c = rt.Context(case)
    yield g, case, c
    c.close()


def test_four_phonon_counts_and_phase_space_all_modes(licoo2):
    """Counts per mode equal the reference's for every mode in ++, and in +-
    and -- except for candidates within rounding of the window edge: omega
    agrees with the reference's LAPACK values to about 1e-10 only, so a few
    of the ~1e11 +- candidates flip (measured: ++ all 1392 modes equal; +-
    83 modes differ by at most 9, the total by 2 of 9.08e8)."""
    g, case, c = licoo2
    N = c.NP_driver_4ph()
    for ch in range(3):
        got, want = N[2 * ch].astype(np.int64), g["N4"][ch]
        d = np.abs(got - want)
        perr = _rel(N[2 * ch + 1].ravel(order="F"), g["P4"][ch].ravel(order="F"))
        print(f"4ph channel {ch}: {int(got.sum())} vs {int(want.sum())}, {int((d > 0).sum())} modes differ, "
              f"max {int(d.max())}; P4 max rel {perr:.2e}", flush=True)
        if ch == 0:
            assert np.array_equal(got, want)
        assert int((d > 0).sum()) <= 150 and int(d.max()) <= 16, ch
        assert abs(int(got.sum()) - int(want.sum())) <= 1e-8 * int(want.sum()), ch
        assert perr <= 2e-6, ch


def test_three_phonon_counts_phase_space_and_rates(licoo2):
    g, case, c = licoo2
    N = c.NP_driver()
    for ch in range(2):
        assert int(N[2 * ch].sum()) == int(g["Ntotal3"][ch])
        assert _rel(N[2 * ch + 1].ravel(order="F"), g["P3"][ch].ravel(order="F")) <= 1e-8
    r = c.RTA_driver()
    sets = degenerate_sets(case)
    single = np.array([s[0] for s in sets if len(s) == 1])
    for k, key in enumerate(("rate_scatt_plus", "rate_scatt_minus", "rate_scatt")):
        got, ref = r[key].ravel(order="F"), g["W3"][k]
        floor = 1e-10 * np.abs(ref).max()     # far-tail rates of 1e-23..1e-28 THz
        err = _rel(got[single - 1], ref[single - 1], floor)
        print(f"{key}: non-degenerate modes max rel {err:.2e}", flush=True)
        assert err <= 1e-8, key
        for s in sets:
            if len(s) > 1:
                assert _rel(got[s - 1].sum(), ref[s - 1].sum(), floor) <= 1e-6, (key, s)
    for ch, key in enumerate(("WP3_plus", "WP3_minus")):
        assert _rel(r[key].ravel(order="F"), g["WP3"][ch], 1e-10 * np.abs(g["WP3"][ch]).max()) <= 2e-4, key
