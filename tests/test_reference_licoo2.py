"""The phonopy front end (FORCE_CONSTANTS_2ND reader, phonon_phonopy with
the non-analytic correction) and the oracle's drivers against the
reference's own LiCoO2 example run (tests/golden/licoo2_example.npz, from
the reference's input and output files: 4 atoms, 12 branches, R-3m,
10x10x10, scalebroad 0.01).

The example prints ten significant digits (WP: five), so the bars are set at
the printing precision: omega and v 1e-9 of the largest value, phase space
1e-8, 3ph rates 1e-8 per mode outside degenerate sets and 1e-6 summed over
each degenerate set, counts exactly.  The example's CONTROL gives no Born
charges, so its non-analytic term is zero; that term is pinned separately
by the LO-TO splitting of a rocksalt-like crystal, whose size the
Lyddane-Sachs-Teller / Born-Huang closed form fixes, with the unit constant
derived from CODATA, and by a finite-difference check of the velocities.
"""
import numpy as np
import pytest

from fourphonon_amd import crystal, harmonic
from oracle.oracle import Oracle
from tests.licoo2_example import licoo2_case, load
from tests.si_example import degenerate_sets


def _rel(got, want, floor=0.0):
    """Largest relative error; values below floor (absolute) are compared
    against floor instead of themselves."""
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
    return licoo2_case(g)


def test_phonopy_front_end_reproduces_the_spectrum(g, case):
    m = case.modes
    assert m.nsymm_rot == 12 and m.Nlist == 116        # log.out: 12 operations, Nlist = 116
    assert np.array_equal(m.List, g["List"]) and np.array_equal(m.Nequi, g["Nequi"])
    w, wref = m.energy[m.List - 1], g["omega_irr"]
    assert np.max(np.abs(w - wref)) <= 1e-9 * np.abs(wref).max()
    assert np.all(w[0, :3] == 0.0)
    vref = g["velocity"]
    assert np.max(np.abs(m.velocity - vref)) <= 1e-9 * np.abs(vref).max()


def test_three_phonon_counts_and_phase_space(g, case):
    N, P = Oracle(case).np3_driver(nthreads=8)
    for ch in range(2):
        assert int(N[ch].sum()) == int(g["Ntotal3"][ch]), ch
        assert _rel(P[ch], g["P3"][ch].ravel(order="F")) <= 1e-8, ch


def test_three_phonon_rates(g, case):
    o = Oracle(case)
    sets = degenerate_sets(case)
    single = np.array([s[0] for s in sets if len(s) == 1])
    multi = [s for s in sets if len(s) > 1]
    modes = np.unique(np.concatenate([single[::7], *multi[::3]])).astype(np.int32)
    r = o.rta3_driver(modes=modes, nthreads=8)
    for k, key in enumerate(("rate_plus", "rate_minus", "rate")):
        ref = g["W3"][k]
        # at scalebroad 0.01 a few modes have rates of 1e-23..1e-28 THz from
        # the far tail of the Gaussian, which amplifies the last digits of
        # omega; they are compared against 1e-10 of the largest rate
        floor = 1e-10 * np.abs(ref).max()
        sel = np.intersect1d(single, modes)
        assert _rel(r[key][sel - 1], ref[sel - 1], floor) <= 1e-8, key
        for s in multi[::3]:
            assert _rel(r[key][s - 1].sum(), ref[s - 1].sum(), floor) <= 1e-6, (key, s)
    for ch, key in enumerate(("wp3_plus", "wp3_minus")):
        assert _rel(r[key][modes - 1], g["WP3"][ch][modes - 1]) <= 2e-4, key


def test_the_example_files_are_one_run(g):
    """The per-mode 4ph counts sum to the log's totals (the BAs example's do
    not: its folder mixes two runs)."""
    assert [int(g["N4"][ch].sum()) for ch in range(3)] == [int(x) for x in g["Ntotal4"]]


def test_four_phonon_counts_of_sample_modes(g, case):
    """BTE.Numprocess_4ph for a few modes (the GPU tests cover all 1392)."""
    o = Oracle(case)
    for mm in (13, 200, 777):
        for ch in range(3):
            n, p = o.np_mode(ch, mm)
            assert n == int(g["N4"][ch, mm - 1]), (mm, ch)
            assert _rel(p, g["P4"][ch].ravel(order="F")[mm - 1]) <= 1e-8, (mm, ch)


def test_reader_split_index_and_fc_layout(g):
    # split_index is x fastest, then y, z, atom (Src/input.f90:83-99)
    assert harmonic.split_index(1, (3, 3, 3)) == (1, 1, 1, 1)
    assert harmonic.split_index(2, (3, 3, 3)) == (2, 1, 1, 1)
    assert harmonic.split_index(4, (3, 3, 3)) == (1, 2, 1, 1)
    assert harmonic.split_index(28, (3, 3, 3)) == (1, 1, 1, 2)
    fc = g["fc2"]
    assert fc.shape == (4, 3, 3, 3, 3, 4, 3)
    # acoustic sum rule of the stored constants: sum over all partners and cells
    # of m-weighted blocks vanishes (phonopy's constants obey it)
    m = np.asarray(g["masses"])[np.asarray(g["types"]) - 1]
    raw = fc * np.sqrt(m[:, None, None, None, None, None, None] * m[None, None, None, None, None, :, None])
    s = raw.sum(axis=(2, 3, 4, 5))
    assert np.abs(s).max() <= 1e-6 * np.abs(raw).max()


# ---------------------------------------------------------------------------
# non-analytic correction
# ---------------------------------------------------------------------------

def _rocksalt(k_spring=10.0, scell=(3, 3, 3)):
    """Two-atom rocksalt-like crystal (a = 0.56 nm, masses 22.99 / 35.45 amu)
    with nearest-neighbour central springs (k eV/A^2), as reduced phonopy
    constants fc[a, i, cell, b, j] in (rad/ps)^2."""
    a = 0.56
    lat = 0.5 * a * np.array([[0, 1, 1], [1, 0, 1], [1, 1, 0]], float).T      # columns a_i, nm
    st = crystal.Structure("rocksalt", lat, np.array([[0, 0, 0], [0.5, 0.5, 0.5]]), np.array([1, 2], np.int32),
                           np.array([22.99, 35.45]))
    cart = st.positions @ lat.T
    m = st.masses[st.types - 1]
    fc = np.zeros((2, 3, *scell, 2, 3))
    dnn = a / 2
    for ia in range(2):
        for ib in range(2):
            for n in np.ndindex(5, 5, 5):
                R = lat @ (np.array(n) - 2)
                d = cart[ib] + R - cart[ia]
                if abs(np.linalg.norm(d) - dnn) > 1e-9:
                    continue
                u = d / np.linalg.norm(d)
                blk = k_spring * np.outer(u, u)
                c = tuple(np.mod(np.array(n) - 2, scell))
                fc[ia, :, c[0], c[1], c[2], ib, :] -= blk
                fc[ia, :, 0, 0, 0, ia, :] += blk
    mm = np.sqrt(m[:, None] * m[None, :])
    return st, harmonic.PHONOPY_UNITFACTOR * fc / mm[:, None, None, None, None, :, None]


def test_nac_prefactor_is_e2_over_eps0_in_the_reference_units():
    # e^2/eps0 in amu nm^3 (rad/ps)^2 from CODATA 2018 (exact e, eps0, amu)
    e, eps0, amu = 1.602176634e-19, 8.8541878128e-12, 1.66053906660e-27
    want = e * e / eps0 / amu * 1e27 * 1e-24
    assert abs(harmonic.NAC_PREFACTOR - want) <= 1e-6 * want


def test_nac_with_zero_born_charges_changes_nothing():
    st, fc = _rocksalt()
    k = harmonic.grid_kpoints(st.rlattvec, (4, 4, 4))
    w0, v0, _ = harmonic.phonons_phonopy(k, st, (3, 3, 3), fc, nonanalytic=False)
    w1, v1, _ = harmonic.phonons_phonopy(k, st, (3, 3, 3), fc, nonanalytic=True, epsilon=np.eye(3) * 3.0,
                                         born=np.zeros((3, 3, 2)))
    assert np.array_equal(w0, w1) and np.array_equal(v0, v1)


def test_nac_lo_to_splitting_closed_form():
    """Near Gamma along q, the Born charges +Z / -Z (isotropic) and eps_inf
    split the optical triplet: the longitudinal mode obeys
    w_LO^2 - w_TO^2 = e^2 Z^2 / (eps0 eps_inf V) (1/m1 + 1/m2)."""
    st, fc = _rocksalt()
    Z, eps = 1.3, 4.2
    born = np.zeros((3, 3, 2))
    born[:, :, 0] = Z * np.eye(3)
    born[:, :, 1] = -Z * np.eye(3)
    for direction in (np.array([1.0, 0.0, 0.0]), np.array([1.0, 1.0, 0.0]) / np.sqrt(2.0)):
        q = 1e-6 * direction[None, :]
        w, _, e = harmonic.phonons_phonopy(q, st, (3, 3, 3), fc, nonanalytic=True, epsilon=eps * np.eye(3),
                                           born=born)
        w = np.sort(w[0])
        w_to, w_lo = w[3], w[5]
        assert abs(w[4] - w_to) <= 1e-6 * w_to                # the transverse pair stays degenerate
        m1, m2 = st.masses
        split = harmonic.NAC_PREFACTOR * Z * Z / (eps * st.volume) * (1.0 / m1 + 1.0 / m2)
        assert abs((w_lo ** 2 - w_to ** 2) - split) <= 1e-5 * split
        # without the correction the optical triplet is degenerate at this q
        w0 = np.sort(harmonic.phonons_phonopy(q, st, (3, 3, 3), fc)[0][0])
        assert abs(w0[5] - w_to) <= 1e-6 * w_to


def test_nac_velocities_are_frequency_gradients():
    """v = d omega / d q with the correction on, at a general q, against
    central differences of omega (the q-dependence of the dielectric term
    enters both)."""
    st, fc = _rocksalt()
    born = np.zeros((3, 3, 2))
    born[:, :, 0] = np.array([[1.1, 0.1, 0.0], [0.1, 1.3, 0.0], [0.0, 0.0, 1.2]])
    born[:, :, 1] = -born[:, :, 0]
    eps = np.array([[4.0, 0.2, 0.0], [0.2, 4.5, 0.0], [0.0, 0.0, 5.0]])
    q0 = np.array([0.7, -0.4, 1.1])
    w, v, _ = harmonic.phonons_phonopy(q0[None], st, (3, 3, 3), fc, nonanalytic=True, epsilon=eps, born=born)
    h = 1e-5
    for a in range(3):
        dq = np.zeros(3)
        dq[a] = h
        wp = harmonic.phonons_phonopy((q0 + dq)[None], st, (3, 3, 3), fc, nonanalytic=True, epsilon=eps, born=born)[0]
        wm = harmonic.phonons_phonopy((q0 - dq)[None], st, (3, 3, 3), fc, nonanalytic=True, epsilon=eps, born=born)[0]
        fd = (wp[0] - wm[0]) / (2 * h)
        assert np.max(np.abs(fd - v[0, :, a])) <= 1e-5 * np.abs(v).max(), a
