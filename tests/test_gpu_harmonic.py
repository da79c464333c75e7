"""Device harmonic front end (fp4_eigen_dm_espresso) on the reference's Si
example: omega and v against the reference's own printed output, the
eigenvectors against the host path through gauge-invariant projectors, and
the three-phonon rates computed from the device eigenvectors against the
reference's BTE.w_3ph* (tests/golden/si_example.npz)."""
import copy

import numpy as np
import pytest

from fourphonon_amd import crystal
from fourphonon_amd import runtime as rt
from tests.si_example import degenerate_sets, load, si_case, structure

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"\n[gpu] {msg}", flush=True)


def _sets_at(E, rtol=1e-6):
    """Degenerate branch groups of one q point."""
    out, used = [], np.zeros(len(E), bool)
    for i in range(len(E)):
        if not used[i]:
            g = np.nonzero(np.abs(E - E[i]) <= rtol * max(abs(E[i]), 1.0))[0]
            used[g] = True
            out.append(g)
    return out


@pytest.fixture(scope="module")
def device_modes():
    g = load()
    host = si_case(g)
    st = structure(g)
    banner("eigenDM on the device, Si example 16x16x16")
    om, eig, vel = rt.eigen_dm_espresso(st, g["fc2"], g["scell"], g["ngrid"], host.modes.symmetrizers)
    return g, host, om, eig, vel


def test_frequencies_and_velocities(device_modes):
    g, host, om, eig, vel = device_modes
    m = host.modes
    w, wref = om[m.List - 1], g["omega_irr"]
    ok = wref > 1e-3
    assert np.max(np.abs(w[ok] - wref[ok]) / wref[ok]) <= 1e-9
    assert np.all(om[0, :3] == 0) and np.all(vel[0, :3] == 0)
    okh = m.energy > 1e-3
    assert np.max(np.abs(om[okh] - m.energy[okh]) / m.energy[okh]) <= 1e-12
    # velocities: summed over each degenerate set (their split is basis dependent)
    vref = g["velocity"]
    scale = np.abs(vref).max()
    worst = 0.0
    for q in range(m.nptk):
        for s in _sets_at(om[q]):
            worst = max(worst, float(np.abs(vel[q, s].sum(axis=0) - vref[q, s].sum(axis=0)).max()))
    print(f"velocity max abs err over degenerate-set sums {worst:.2e} (scale {scale:.2f})")
    assert worst <= 1e-8 * scale


def test_eigenvectors_orthonormal_and_same_subspaces(device_modes):
    g, host, om, eig, vel = device_modes
    m = host.modes
    worst_o, worst_p = 0.0, 0.0
    for q in range(0, m.nptk, 7):
        E = eig[q]                                   # [mode, component]
        worst_o = max(worst_o, float(np.abs(E @ E.conj().T - np.eye(m.nbands)).max()))
        for s in _sets_at(om[q]):
            Pd = E[s].T @ E[s].conj()
            Ph = m.eigenvect[q][s].T @ m.eigenvect[q][s].conj()
            worst_p = max(worst_p, float(np.abs(Pd - Ph).max()))
    print(f"orthonormality {worst_o:.2e}, projector difference {worst_p:.2e}")
    assert worst_o <= 1e-12 and worst_p <= 1e-9


def test_three_phonon_rates_from_device_eigenvectors(device_modes):
    g, host, om, eig, vel = device_modes
    case = copy.copy(host)
    m = copy.copy(host.modes)
    m.energy, m.velocity, m.eigenvect = om, vel, eig
    case.modes = m
    banner("3ph RTA from the device eigenvectors")
    c = rt.Context(case)
    try:
        r = c.RTA_driver()
    finally:
        c.close()
    # The Jacobi basis of a degenerate set differs from LAPACK's; the
    # velocities of its members, and with them sigma and the window of every
    # process that has such a member as a partner, follow the basis, so the
    # rates agree to ~1e-5 here rather than to the printing precision (the
    # host path, on LAPACK, agrees to 2e-9).
    sets = degenerate_sets(case)
    single = np.array([s[0] for s in sets if len(s) == 1])
    for k, key in enumerate(("rate_scatt_plus", "rate_scatt_minus", "rate_scatt")):
        got, ref = r[key].ravel(order="F"), g["W3"][k]
        nz = ref[single - 1] != 0
        err = float(np.max(np.abs(got[single - 1][nz] - ref[single - 1][nz]) / np.abs(ref[single - 1][nz])))
        errs = max((abs(got[s - 1].sum() - ref[s - 1].sum()) / abs(ref[s - 1].sum())
                    for s in sets if len(s) > 1 and ref[s - 1].sum() != 0), default=0.0)
        print(f"{key}: non-degenerate max rel {err:.2e}, degenerate-set sums {errs:.2e}")
        assert err <= 3e-5 and errs <= 3e-5, key
