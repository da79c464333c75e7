"""The reference's LiCoO2 example (4 atoms, 12 branches, 10x10x10, phonopy
2nd-order constants) as a hot-path case, rebuilt from the fixture that
tests/golden/make_licoo2_example.py writes from the reference's own input
and output files.

omega, v and the eigenvectors come from fourphonon_amd.harmonic's phonopy
front end applied to the example's 2nd-order constants (with its
nonanalytic flag, Born charges and dielectric tensor); the third-order
constants are the example's; the example ships no fourth-order constants, so
the 4ph clusters are one zero cluster (4ph counts and phase space need none).
"""
import os

import numpy as np

from fourphonon_amd import crystal, harmonic

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "licoo2_example.npz")


def load():
    return np.load(FIXTURE)


def structure(g):
    return crystal.Structure("LiCoO2-example", np.asarray(g["lattvec"]), np.asarray(g["positions"]),
                             np.asarray(g["types"], np.int32), np.asarray(g["masses"]))


def modes(g=None):
    g = load() if g is None else g
    return harmonic.modes_from_phonopy(structure(g), g["fc2"], tuple(int(x) for x in g["scell"]),
                                       tuple(int(x) for x in g["ngrid"]), nonanalytic=bool(g["nonanalytic"]),
                                       epsilon=g["epsilon"], born=g["born"])


def ifcs(g):
    """(third-order constants of the fixture, one zero 4ph cluster)."""
    idx = g["fc3_idx"]
    ifc3 = crystal.IFC3(g["fc3_phi"].shape[-1], np.asfortranarray(g["fc3_phi"]), np.asfortranarray(g["fc3_Rj"]),
                        np.asfortranarray(g["fc3_Rk"]), *[np.ascontiguousarray(idx[a], np.int32) for a in range(3)],
                        g["fc3_cells"])
    zero4 = crystal.IFC4(1, np.zeros((3, 3, 3, 3, 1), order="F"), *[np.zeros((3, 1), order="F")] * 3,
                         *[np.ones(1, np.int32)] * 4, np.zeros((1, 3, 3), np.int64))
    return ifc3, zero4


def licoo2_case(g=None):
    g = load() if g is None else g
    ifc3, zero4 = ifcs(g)
    return crystal.Case(structure(g), modes(g), zero4, T=float(g["T"]), scalebroad=float(g["scalebroad"]),
                        ifc3=ifc3, meta={"workload": "LiCoO2 example 10x10x10"})
