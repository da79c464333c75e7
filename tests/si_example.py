"""The reference's Si example (16x16x16) as a hot-path case, rebuilt from the
fixture that tests/golden/make_si_example.py writes from the reference's own
input and output files.

omega, v and the eigenvectors come from fourphonon_amd.harmonic applied to
the example's 2nd-order constants; the third-order constants are the
example's; there are no fourth-order constants in the example, so the 4ph
clusters are one zero cluster (4ph counts, phase space and WP4 need none).
"""
import os

import numpy as np

from fourphonon_amd import crystal, harmonic

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "si_example.npz")


def load():
    return np.load(FIXTURE)


def structure(g):
    return crystal.Structure("Si-example", np.asarray(g["lattvec"]), np.asarray(g["positions"]),
                             np.array([1, 1], np.int32), np.array([float(g["mass"])]))


def si_case(g=None):
    g = load() if g is None else g
    st = structure(g)
    modes = harmonic.modes_from_espresso(st, g["fc2"], tuple(int(x) for x in g["scell"]),
                                         tuple(int(x) for x in g["ngrid"]))
    idx = g["fc3_idx"]
    ifc3 = crystal.IFC3(g["fc3_phi"].shape[-1], np.asfortranarray(g["fc3_phi"]), np.asfortranarray(g["fc3_Rj"]),
                        np.asfortranarray(g["fc3_Rk"]), *[np.ascontiguousarray(idx[a], np.int32) for a in range(3)],
                        g["fc3_cells"])
    zero4 = crystal.IFC4(1, np.zeros((3, 3, 3, 3, 1), order="F"), *[np.zeros((3, 1), order="F")] * 3,
                         *[np.ones(1, np.int32)] * 4, np.zeros((1, 3, 3), np.int64))
    return crystal.Case(st, modes, zero4, T=float(g["T"]), scalebroad=float(g["scalebroad"]), ifc3=ifc3,
                        meta={"workload": "Si example 16x16x16"})


def degenerate_sets(case, rtol=1e-6):
    """Per irreducible q, groups of branches (mm indices, 1-based) whose
    frequencies agree within rtol: per-mode rates of a degenerate set depend
    on the basis a Hermitian solver picks, their sum does not."""
    m = case.modes
    E = m.energy[m.List - 1]
    sets = []
    for ll in range(m.Nlist):
        used = np.zeros(m.nbands, bool)
        for i in range(m.nbands):
            if used[i]:
                continue
            grp = np.nonzero(np.abs(E[ll] - E[ll, i]) <= rtol * max(E[ll, i], 1.0))[0]
            used[grp] = True
            sets.append(ll * m.nbands + grp + 1)
    return sets
