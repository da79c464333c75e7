"""The oracle against the reference's own Si sampling run
(tests/golden/si_sampling_example.npz: Si, 32x32x32, scalebroad 1, full 3ph
drivers, 4ph sampling estimators with 100000 samples per mode and channel).

The run's 3ph outputs are deterministic and compared at the printing
precision, as for the 16x16x16 Si example.  Its 4ph outputs come from a
clock-seeded random stream, so they are compared statistically: the
estimate N = INT(REAL(hits)/num*box) with box = Nb*nptk*Nb*nptk*Nb
returns each mode's hit count exactly (the printed counts are integers
times box/num to 4e-7), and two independent samplings of the same
candidates give per-mode hit counts whose difference, given their sum S,
is that of a Binomial(S, 1/2): (h1 - h2)^2 / S has mean 1 and variance
about 2, so the channel's chi-square per mode lies within 5*sqrt(2/dof) of
1, and the total hit difference within 5 sigma of 0.
"""
import numpy as np
import pytest

from oracle.oracle import Oracle
from tests.si_example import degenerate_sets
from tests.si_sampling_example import hit_statistics, hits, load, nsample, si_sampling_case


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
    return si_sampling_case(g)


def test_spectrum_at_32(g, case):
    m = case.modes
    assert m.nsymm_rot == 48 and m.Nlist == 897                    # log.out: Fd-3m, Nlist = 897
    assert np.array_equal(m.List, g["List"]) and np.array_equal(m.Nequi, g["Nequi"])
    w, wref = m.energy[m.List - 1], g["omega_irr"]
    assert np.max(np.abs(w - wref)) <= 1e-9 * np.abs(wref).max()
    v, vref = m.velocity[m.List - 1], g["v_irr"]
    assert np.max(np.abs(v - vref)) <= 1e-8 * np.abs(vref).max()


def test_three_phonon_totals_and_phase_space_at_32(g, case):
    N, P = Oracle(case).np3_driver(nthreads=8)
    assert [int(N[ch].sum()) for ch in range(2)] == [int(x) for x in g["Ntotal3"]]
    for ch in range(2):
        assert _rel(P[ch], g["P3"][ch].ravel(order="F")) <= 1e-8, ch


def test_three_phonon_rates_of_sample_modes(g, case):
    sets = degenerate_sets(case)
    single = np.array([s[0] for s in sets if len(s) == 1])
    multi = [s for s in sets if len(s) > 1]
    pick = [multi[len(multi) // 3], multi[2 * len(multi) // 3]]
    modes = np.unique(np.concatenate([single[::400], *pick])).astype(np.int32)
    r = Oracle(case).rta3_driver(modes=modes, nthreads=8)
    sel = np.intersect1d(single, modes)
    for k, key in enumerate(("rate_plus", "rate_minus", "rate")):
        ref = g["W3"][k]
        floor = 1e-10 * np.abs(ref).max()
        assert _rel(r[key][sel - 1], ref[sel - 1], floor) <= 1e-8, key
        for s in pick:
            assert _rel(r[key][s - 1].sum(), ref[s - 1].sum(), floor) <= 1e-6, (key, s)
    for ch, key in enumerate(("wp3_plus", "wp3_minus")):
        assert _rel(r[key][modes - 1], g["WP3"][ch][modes - 1], 1e-10 * np.abs(g["WP3"][ch]).max()) <= 2e-4, key


def test_four_phonon_sampling_matches_the_runs_statistics(g, case):
    """Measured with seed 12345: z -1.2 / -1.6 / 0.3 and chi-square per mode
    0.97 / 0.98 / 1.01 over 2649 / 5379 / 3440 modes; phase-space totals
    within 0.35%.  The phase-space bar is 5 sigma of the hit statistics with
    a factor 3 for the spread of the Gaussian weights."""
    num = nsample(g, "num_sample_process_4ph_phase_space")
    assert num == 100000
    N, P = Oracle(case).np_sample_driver(num, 12345, nthreads=8)
    for ch in range(3):
        ho, hr = hits(N[ch], case, num), hits(g["N4"][ch], case, num)
        z, chi, dof = hit_statistics(ho, hr)
        assert abs(z) <= 5.0, (ch, z)
        assert abs(chi - 1.0) <= 5.0 * np.sqrt(2.0 / dof), (ch, chi, dof)
        Pr = g["P4"][ch].ravel(order="F")
        bar = 5.0 * np.sqrt(2.0) * 3.0 / np.sqrt(hr.sum())
        assert abs(P[ch].sum() / Pr.sum() - 1.0) <= bar, (ch, P[ch].sum() / Pr.sum(), bar)


def test_the_statistics_reject_a_biased_window(g, case):
    """Negative control: the same estimator with the window 5% narrower
    (scalebroad 0.95) fails the chi-square or the total in every channel."""
    import dataclasses
    num = nsample(g, "num_sample_process_4ph_phase_space")
    N, _ = Oracle(dataclasses.replace(case, scalebroad=0.95)).np_sample_driver(num, 12345, nthreads=8)
    for ch in range(3):
        z, chi, dof = hit_statistics(hits(N[ch], case, num), hits(g["N4"][ch], case, num))
        assert abs(z) > 5.0 or abs(chi - 1.0) > 5.0 * np.sqrt(2.0 / dof), (ch, z, chi)
