"""The reference's Si example (16x16x16) as a hot-path case, rebuilt from the
fixture that tests/golden/make_si_example.py writes from the reference's own
output files.

Only omega, v, the grid, List and the broadening are real; eigenvectors are
zero and the clusters are a single zero cluster, so the case serves the
count, list and phase-space drivers (and weighted phase space with
onlyharmonic), which read nothing else.
"""
import os

import numpy as np

from fourphonon_amd import crystal

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "si_example.npz")


def load():
    return np.load(FIXTURE)


def si_case(g=None, onlyharmonic=True):
    g = load() if g is None else g
    ngrid = np.asarray(g["ngrid"], np.int32)
    nptk = int(np.prod(ngrid))
    nlist, nb = g["omega_irr"].shape
    energy = np.asfortranarray(g["omega_irr"][g["rep"] - 1, :])             # (nptk, Nb)
    velocity = np.asfortranarray(g["velocity"])                             # (nptk, Nb, 3)
    st = crystal.Structure("Si-example", np.asarray(g["lattvec"]), np.array([[0, 0, 0], [0.25, 0.25, 0.25]]),
                           np.array([1, 1], np.int32), np.array([28.085]))
    one = np.ones((1, nptk), np.int32, order="F")
    modes = crystal.Modes(
        ngrid=ngrid, energy=energy, velocity=velocity,
        eigenvect=np.zeros((nptk, nb, nb), np.complex128, order="F"), IJK=crystal.grid_ijk(ngrid),
        Nlist=nlist, List=np.asarray(g["List"], np.int32), Nequi=np.asarray(g["Nequi"], np.int32),
        ALLEquiList=one, TypeofSymmetry=one, nsymm_rot=1, crotations=np.asfortranarray(np.eye(3)[:, :, None]),
        qrotations=np.asfortranarray(np.eye(3)[:, :, None]),
        symmetrizers=np.asfortranarray(np.repeat(np.eye(3)[:, :, None], nptk, axis=2)),
        ID_equi=np.asfortranarray(np.arange(1, nptk + 1, dtype=np.int32)[None, :]))
    zero4 = crystal.IFC4(1, np.zeros((3, 3, 3, 3, 1), order="F"), *[np.zeros((3, 1), order="F")] * 3,
                         *[np.ones(1, np.int32)] * 4, np.zeros((1, 3, 3), np.int64))
    zero3 = crystal.IFC3(1, np.zeros((3, 3, 3, 1), order="F"), *[np.zeros((3, 1), order="F")] * 2,
                         *[np.ones(1, np.int32)] * 3, np.zeros((1, 2, 3), np.int64))
    return crystal.Case(st, modes, zero4, T=float(g["T"]), scalebroad=float(g["scalebroad"]),
                        onlyharmonic=onlyharmonic, ifc3=zero3, meta={"workload": "Si example 16x16x16"})
