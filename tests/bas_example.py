"""The reference's BAs example (zincblende, F-43m, 2 atoms, 6 branches,
16x16x16, phonopy 2nd-order constants with the non-analytic correction of
its Born charges and dielectric tensor) as a hot-path case, rebuilt from the
fixture that tests/golden/make_bas_example.py writes from the reference's
own input and output files.  Built like the LiCoO2 example case: omega, v
and e from fourphonon_amd.harmonic's phonopy front end, with the
non-analytic directions the example run took at the zone-boundary ties
(nac_images), the example's third-order constants, and one zero 4ph cluster
(the example ships no fourth-order constants; counts and phase space need
none).
"""
import os

import numpy as np

from fourphonon_amd import crystal, harmonic
from tests import licoo2_example

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "bas_example.npz")


def load():
    return np.load(FIXTURE)


def structure(g):
    return crystal.Structure("BAs-example", np.asarray(g["lattvec"]), np.asarray(g["positions"]),
                             np.asarray(g["types"], np.int32), np.asarray(g["masses"]))


def modes(g=None, images=True):
    """images=False: the front end's own shortest-image search."""
    g = load() if g is None else g
    return harmonic.modes_from_phonopy(structure(g), g["fc2"], tuple(int(x) for x in g["scell"]),
                                       tuple(int(x) for x in g["ngrid"]), nonanalytic=bool(g["nonanalytic"]),
                                       epsilon=g["epsilon"], born=g["born"],
                                       images=g["nac_images"] if images else None)


def bas_case(g=None):
    g = load() if g is None else g
    ifc3, zero4 = licoo2_example.ifcs(g)
    return crystal.Case(structure(g), modes(g), zero4, T=float(g["T"]), scalebroad=float(g["scalebroad"]),
                        ifc3=ifc3, meta={"workload": "BAs example 16x16x16"})
