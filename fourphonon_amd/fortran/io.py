"""Input/output files of the Fortran driver (raw arrays in reference memory order)."""
import os

import numpy as np


def write_inputs(case, path):
    """meta.txt plus one raw little-endian file per array, Fortran order."""
    os.makedirs(path, exist_ok=True)
    st, m, f = case.structure, case.modes, case.ifc4
    with open(os.path.join(path, "meta.txt"), "w") as fh:
        fh.write(" ".join(str(int(x)) for x in (*m.ngrid, m.nbands, st.natoms, len(st.masses), m.Nlist, f.Ntri)) + "\n")
        fh.write(f"{case.T!r} {case.scalebroad!r} {min(case.omega_max, 1e300)!r} {int(case.onlyharmonic)}\n")
        fh.write(" ".join(repr(float(x)) for x in np.asarray(st.rlattvec).ravel(order="F")) + "\n")
        fh.write(" ".join(repr(float(x)) for x in st.masses) + "\n")
        fh.write(" ".join(str(int(x)) for x in st.types) + "\n")
        fh.write(f"{m.nsymm_rot} {float(st.volume)!r}\n")
    arrays = {
        "energy": (m.energy, np.float64), "velocity": (m.velocity, np.float64),
        "eigenvect": (m.eigenvect, np.complex128), "IJK": (m.IJK, np.int32), "List": (m.List, np.int32),
        "Phi": (f.Phi, np.float64), "R_j": (f.R_j, np.float64), "R_k": (f.R_k, np.float64),
        "R_l": (f.R_l, np.float64), "Index_i": (f.Index_i, np.int32), "Index_j": (f.Index_j, np.int32),
        "Index_k": (f.Index_k, np.int32), "Index_l": (f.Index_l, np.int32),
        "Nequi": (m.Nequi, np.int32), "ALLEquiList": (m.ALLEquiList, np.int32),
        "TypeofSymmetry": (m.TypeofSymmetry, np.int32), "ID_equi": (m.ID_equi, np.int32),
        "symmetrizers": (m.symmetrizers, np.float64), "crotations": (m.crotations, np.float64),
    }
    for name, (a, dt) in arrays.items():
        np.asarray(a, dtype=dt).ravel(order="F").tofile(os.path.join(path, name + ".bin"))


def read_outputs(path, nmodes, lists=False):
    N = np.fromfile(os.path.join(path, "N.bin"), dtype=np.int64).reshape(3, nmodes)
    rates = np.fromfile(os.path.join(path, "rates.bin"), dtype=np.float64).reshape(7, nmodes)
    out = {"N": N, "rates": rates}
    if lists:
        for ch in ("pp", "pm", "mm"):
            x = np.fromfile(os.path.join(path, f"lists_{ch}.bin"), dtype=np.int64)
            out["lists_" + ch] = x.reshape(3, -1)
    return out


def read_iteration_outputs(path, nbands, nptk, sweeps=3):
    """Outputs of the driver's "iterate" mode."""
    F = np.fromfile(os.path.join(path, "iter_F.bin"), dtype=np.float64).reshape((nbands, nptk, 3), order="F")
    kap = np.fromfile(os.path.join(path, "iter_kappa.bin"), dtype=np.float64)
    kap = kap.reshape((nbands, 3, 3, sweeps + 1), order="F")
    raw = np.fromfile(os.path.join(path, "solve.bin"), dtype=np.float64)
    niter = int(raw[0])
    hist = raw[1:1 + 9 * (niter + 1)].reshape((3, 3, niter + 1), order="F")
    rel = raw[1 + 9 * (niter + 1):1 + 10 * (niter + 1)]
    kappa = raw[1 + 10 * (niter + 1):].reshape((nbands, 3, 3), order="F")
    tau = np.fromfile(os.path.join(path, "tau.bin"), dtype=np.float64).reshape((nbands, -1), order="F")
    return {"F_n": F, "kappa_sweeps": kap, "niter": niter, "kappa_iter": hist, "relchange": rel, "kappa": kappa,
            "tau": tau}
