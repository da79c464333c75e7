"""The phonopy front end with a non-zero non-analytic term, and the oracle's
drivers, against the reference's own BAs example run
(tests/golden/bas_example.npz, from the reference's input and output files:
zincblende BAs, F-43m, 2 atoms, 6 branches, 16x16x16, Born charges
-0.4818 / +0.4838 and epsilon 9.823).

The example's output folder holds two runs.  BTE.omega, v_full, P3, P4 and
Numprocess_4ph come from the scalebroad = 0.1 of its CONTROL; log.out and
T300K/ (the 3ph rates, WP3) from a run at scalebroad = 1.0, whose 3ph
totals the oracle reproduces exactly at 1.0 and not at 0.1.  That run's 4ph
totals (2.2e10 / 1.4e11 / 2.7e10) and WP4 match neither scalebroad (at 1.0
the oracle's totals are 7.8e9 / 5.9e10 / 3.6e9, WP4 within 1.5%): a build
with another 4ph window, so they are not compared.  The P4 files
hold Infinity for every mode with a process and 0 otherwise: a build whose
1/nptk**3 wrapped to zero at nptk = 4096 (the reference's present
dble(nptk)**3, Src/processes.f90:2319, does not), so only that pattern is
checked.  Bars are the printing precision, as for the LiCoO2 example.
"""
import dataclasses

import numpy as np
import pytest

from fourphonon_amd import harmonic
from oracle.oracle import Oracle
from tests.bas_example import bas_case, load, modes
from tests.si_example import degenerate_sets


def _rel(got, want, floor=0.0):
    got, want = np.atleast_1d(np.asarray(got, float)), np.atleast_1d(np.asarray(want, float))
    nz = want != 0
    if not nz.any():
        return float(np.abs(got).max()) if got.size else 0.0
    return float(np.max(np.abs(got[nz] - want[nz]) / np.maximum(np.abs(want[nz]), floor)))


@pytest.fixture(scope="module")
def g():
    return load()


@pytest.fixture(scope="module")
def case(g):
    return bas_case(g)


def test_spectrum_with_the_nonanalytic_term(g, case):
    m = case.modes
    # log.out: F-43m, 24 operations; time reversal doubles them (no inversion)
    assert m.nsymm_rot == 48 and m.Nlist == 145
    assert np.array_equal(m.List, g["List"]) and np.array_equal(m.Nequi, g["Nequi"])
    w, wref = m.energy[m.List - 1], g["omega_irr"]
    assert np.max(np.abs(w - wref)) <= 1e-9 * np.abs(wref).max()
    assert np.all(w[0, :3] == 0.0)
    vref = g["velocity"]
    assert np.max(np.abs(m.velocity - vref)) <= 1e-9 * np.abs(vref).max()
    # the term is far above the bar: without it the LO branch moves by up
    # to 0.54 rad/ps (0.4%)
    st = case.structure
    kp = harmonic.grid_kpoints(st.rlattvec, g["ngrid"])[m.List[1:] - 1]
    w_no = harmonic.phonons_phonopy(kp, st, tuple(int(x) for x in g["scell"]), g["fc2"])[0]
    assert np.max(np.abs(w_no - wref[1:]) / wref[1:]) > 3e-3


def test_shortest_image_search_differs_only_at_zone_boundary_ties(g, case):
    """The front end's own search takes the example's direction everywhere
    except where several images of q are equally short (the strict
    comparison there follows the norm's last bit)."""
    m0 = modes(g, images=False)
    m = case.modes
    diff = np.flatnonzero(np.abs(m0.energy - m.energy).max(axis=1) > 1e-7)
    assert 0 < len(diff) < 200
    st = case.structure
    kp = harmonic.grid_kpoints(st.rlattvec, g["ngrid"])
    b = st.rlattvec
    shifts = np.array([(x, y, z) for x in range(-2, 3) for y in range(-2, 3) for z in range(-2, 3)], float)
    for iq in diff:
        d = np.linalg.norm(kp[iq][None, :] + shifts @ b.T, axis=1)
        assert np.sum(d <= d.min() * (1 + 1e-12)) >= 2, iq
        assert abs(np.linalg.norm(g["nac_images"][iq]) - d.min()) <= 1e-12 * d.min(), iq


def test_the_recorded_images_are_shortest_images_of_their_q(g, case):
    st = case.structure
    kp = harmonic.grid_kpoints(st.rlattvec, g["ngrid"])
    img = g["nac_images"]
    n = np.linalg.solve(st.rlattvec, (img - kp).T).T            # reciprocal-lattice coefficients
    assert np.abs(n - np.round(n)).max() <= 1e-9
    b = st.rlattvec
    shifts = np.array([(x, y, z) for x in range(-2, 3) for y in range(-2, 3) for z in range(-2, 3)], float)
    dmin = np.min(np.linalg.norm(kp[:, None, :] + (shifts @ b.T)[None, :, :], axis=2), axis=1)
    assert np.all(np.abs(np.linalg.norm(img, axis=1) - dmin) <= 1e-12 * np.maximum(dmin, 1.0))


def test_three_phonon_phase_space_at_scalebroad_0_1(g, case):
    N, P = Oracle(case).np3_driver(nthreads=8)
    for ch in range(2):
        assert _rel(P[ch], g["P3"][ch].ravel(order="F")) <= 1e-8, ch


def test_three_phonon_totals_rates_and_wp3_at_scalebroad_1(g, case):
    """log.out's Ntotal_plus / Ntotal_minus and T300K's rates and WP3 are
    the scalebroad = 1.0 run's.  The rates agree at the printing precision
    for most modes (median 1.5e-10 per non-degenerate mode, totals 3e-8) and
    within 1.2e-5 for all: at 41 zone-boundary q points two tied images give
    the same omega and the same symmetrised v but other eigenvectors, so the
    run's printed omega and v do not say which one it took; fitting those
    choices to the rates closes half of the gap only, so the rest is left
    as the bar."""
    c1 = dataclasses.replace(case, scalebroad=1.0)
    o = Oracle(c1)
    N, _ = o.np3_driver(nthreads=8)
    assert [int(N[ch].sum()) for ch in range(2)] == [int(x) for x in g["Ntotal3"]]
    sets = degenerate_sets(case)
    single = np.array([s[0] for s in sets if len(s) == 1])
    multi = [s for s in sets if len(s) > 1]
    sel_modes = np.unique(np.concatenate([single[::5], *multi[::4]])).astype(np.int32)
    r = o.rta3_driver(modes=sel_modes, nthreads=8)
    for k, key in enumerate(("rate_plus", "rate_minus", "rate")):
        ref = g["W3"][k]
        floor = 1e-10 * np.abs(ref).max()
        sel = np.intersect1d(single, sel_modes)
        e = np.abs(r[key][sel - 1] - ref[sel - 1]) / np.maximum(np.abs(ref[sel - 1]), floor)
        assert np.median(e) <= 1e-8 and e.max() <= 2e-5, key
        for s in multi[::4]:
            assert _rel(r[key][s - 1].sum(), ref[s - 1].sum(), floor) <= 5e-6, (key, s)
    for ch, key in enumerate(("wp3_plus", "wp3_minus")):
        ref = g["WP3"][ch]
        assert _rel(r[key][sel_modes - 1], ref[sel_modes - 1], 1e-10 * np.abs(ref).max()) <= 2e-4, key


def test_four_phonon_counts_of_a_sample_mode(g, case):
    """BTE.Numprocess_4ph of mode 400 (the GPU tests cover all 870)."""
    o = Oracle(case)
    mm = 400
    for ch in range(3):
        n, p = o.np_mode(ch, mm)
        assert n == int(g["N4"][ch, mm - 1]), ch
        assert np.isfinite(p) and (p > 0) == (n > 0), ch


def test_p4_files_are_the_overflowed_build(g):
    for ch in range(3):
        P, N = g["P4"][ch].ravel(order="F"), g["N4"][ch]
        assert np.all(np.isinf(P) == (N > 0)) and np.all(P[N == 0] == 0.0), ch
