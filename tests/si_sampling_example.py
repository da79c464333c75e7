"""The reference's Si sampling example (Example-Si-sampling: the Si example's
crystal and constants on a 32x32x32 mesh, 4ph sampling estimators with
100000 samples per mode and channel) as a hot-path case: the constants come
from tests/golden/si_example.npz (the two examples' constant files are
byte-identical), the mesh and run parameters and the outputs from
tests/golden/si_sampling_example.npz (tests/golden/make_si_sampling_example.py).
"""
import dataclasses
import os

import numpy as np

from fourphonon_amd import harmonic
from tests import si_example

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "si_sampling_example.npz")


def load():
    return np.load(FIXTURE)


def nsample(g, key):
    keys = [str(k) for k in g["nsample_keys"]]
    return int(g["nsample"][keys.index(key)])


def box(case):
    """The 4ph sampling box of one mode: Nb*nptk*Nb*nptk*Nb candidates."""
    m = case.modes
    nptk = int(np.prod(m.ngrid))
    return m.nbands ** 3 * nptk ** 2


def si_sampling_case(g=None):
    g = load() if g is None else g
    gs = si_example.load()
    base = si_example.si_case(gs)
    st = base.structure
    modes = harmonic.modes_from_espresso(st, gs["fc2"], tuple(int(x) for x in gs["scell"]),
                                         tuple(int(x) for x in g["ngrid"]))
    return dataclasses.replace(base, modes=modes, T=float(g["T"]), scalebroad=float(g["scalebroad"]),
                               meta={"workload": "Si sampling example 32x32x32"})


def hit_statistics(h_ours, h_ref):
    """(z of the total difference, chi-square per mode, dof)."""
    S = h_ours + h_ref
    nz = S > 0
    z = float((h_ours.sum() - h_ref.sum()) / np.sqrt(S.sum()))
    chi = float(((h_ours - h_ref)[nz] ** 2 / S[nz]).sum() / nz.sum())
    return z, chi, int(nz.sum())


def hits(N, case, num):
    h = np.asarray(N, float) * num / box(case)
    r = np.round(h)
    assert np.abs(h - r).max() <= 1e-5
    return r
