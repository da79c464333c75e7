"""The harmonic front end, the FORCE_CONSTANTS_3RD reader and the oracle's
three-phonon drivers against the reference's own Si example run
(tests/golden/si_example.npz, from the reference's input and output files).

The example prints ten significant digits (WP: five), so these bars are set
at the printing precision: omega 1e-9, phase space 1e-8, 3ph rates 1e-8
(per mode outside degenerate sets; summed over each degenerate set, whose
split depends on the eigenvector basis, 1e-6), counts exactly.
"""
import copy
import os

import numpy as np

from fourphonon_amd import harmonic
from oracle.oracle import Oracle
from tests.si_example import degenerate_sets, load, si_case, structure


def _rel(got, want):
    got, want = np.atleast_1d(np.asarray(got, float)), np.atleast_1d(np.asarray(want, float))
    nz = want != 0
    if not nz.any():
        return float(np.abs(got).max()) if got.size else 0.0
    return float(np.max(np.abs(got[nz] - want[nz]) / np.abs(want[nz])))


def test_harmonic_front_end():
    g = load()
    case = si_case(g)
    m = case.modes
    assert np.array_equal(m.List, g["List"]) and np.array_equal(m.Nequi, g["Nequi"])
    w, wref = m.energy[m.List - 1], g["omega_irr"]
    ok = wref > 1e-3
    assert np.max(np.abs(w[ok] - wref[ok]) / wref[ok]) <= 1e-9
    assert np.all(w[~ok] == 0.0)
    vref = g["velocity"]
    assert np.max(np.abs(m.velocity - vref)) <= 1e-8 * np.abs(vref).max()


def test_fc3_reader_round_trip(tmp_path):
    g = load()
    case = si_case(g)
    f = case.ifc3
    path = os.path.join(tmp_path, "FORCE_CONSTANTS_3RD")
    harmonic.write_force_constants_3rd(path, f)
    back = harmonic.read_force_constants_3rd(path, structure(g).lattvec)
    assert back.Ntri == f.Ntri
    assert np.array_equal(back.Index_i, f.Index_i) and np.array_equal(back.Index_k, f.Index_k)
    assert np.allclose(back.Phi, f.Phi, rtol=1e-11, atol=0)
    assert np.allclose(back.R_j, f.R_j, atol=1e-12) and np.array_equal(back.cells, f.cells)


def test_three_phonon_counts_and_phase_space():
    g = load()
    o = Oracle(si_case(g))
    N, P = o.np3_driver(nthreads=8)
    for ch in range(2):
        assert int(N[ch].sum()) == int(g["Ntotal3"][ch]), ch
        assert _rel(P[ch], g["P3"][ch].ravel(order="F")) <= 1e-8, ch


def test_three_phonon_rates_and_weighted_phase_space():
    g = load()
    case = si_case(g)
    o = Oracle(case)
    sets = degenerate_sets(case)
    single = np.array([s[0] for s in sets if len(s) == 1])
    multi = [s for s in sets if len(s) > 1]
    modes = np.unique(np.concatenate([single[::20], *multi[::4]])).astype(np.int32)
    r = o.rta3_driver(modes=modes, nthreads=8)
    for k, key in enumerate(("rate_plus", "rate_minus", "rate")):
        ref = g["W3"][k]
        sel = np.intersect1d(single, modes)
        assert _rel(r[key][sel - 1], ref[sel - 1]) <= 1e-8, key
        for s in multi[::4]:
            assert _rel(r[key][s - 1].sum(), ref[s - 1].sum()) <= 1e-6, (key, s)
    for ch, key in enumerate(("wp3_plus", "wp3_minus")):
        assert _rel(r[key][modes - 1], g["WP3"][ch][modes - 1]) <= 2e-4, key


def test_negative_control_broadening():
    g = load()
    c2 = copy.copy(si_case(g))
    c2.scalebroad = 0.5
    N, _ = Oracle(c2).np3_driver(nthreads=8)
    assert all(abs(int(N[ch].sum()) - int(g["Ntotal3"][ch])) > 0.2 * int(g["Ntotal3"][ch]) for ch in range(2))


def test_polar_correction_is_refused():
    import pytest
    g = load()
    born = np.zeros((3, 3, 2))
    born[0, 0, 0] = 1.0
    with pytest.raises(ValueError):
        harmonic.phonons_espresso(np.zeros((1, 3)), g["lattvec"], g["positions"], [float(g["mass"])], [1, 1],
                                  (6, 6, 6), g["fc2"], born=born)
