"""Turns the reference's own LiCoO2 example outputs into a test fixture.

The reference ships a complete run of its LiCoO2 example (4 atoms, 12
branches, R-3m, 10x10x10 mesh, scalebroad = 0.01, T = 300 K, phonopy
2nd-order constants with nonanalytic = .TRUE.): inputs ``CONTROL``,
``FORCE_CONSTANTS_2ND`` (phonopy format) and ``FORCE_CONSTANTS_3RD``, and
these output files (data, read as text):

* ``BTE.qpoints`` (List, Nequi), ``BTE.omega`` (omega of the irreducible
  points, rad/ps), ``BTE.v_full`` (group velocities on the full grid, km/s);
* ``BTE.P3_plus`` / ``BTE.P3_minus``, ``BTE.P4_plusplus`` /
  ``BTE.P4_plusmius`` / ``BTE.P4_minusminus`` and ``BTE.Numprocess_4ph``
  (phase space and four-phonon counts, mode mm in row order);
* ``T300K/BTE.w_3ph*`` (3ph RTA rates, ten digits; rows branch-major) and
  ``T300K/BTE.WP3_*``, ``T300K/BTE.WP4_*`` (weighted phase space);
* the totals of ``log.out`` (3ph and 4ph process counts).

The CONTROL gives no Born charges or dielectric tensor, so the reference's
defaults hold (born = 0, epsilon = identity) and its non-analytic term is
exactly zero.  The inputs are parsed with fourphonon_amd's readers and
stored as arrays; the tests rebuild omega, v and e with the phonopy front
end and compare every stage against the outputs.  Run in the development
container (the reference tree is not on the GPU box):

    python tests/golden/make_licoo2_example.py

and commit the resulting tests/golden/licoo2_example.npz.
"""
import os
import re
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from fourphonon_amd import formats, harmonic  # noqa: E402

SRC = "/root/reference/Example-LiCoO2"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "licoo2_example.npz")


def floats(path):
    with open(path) as f:
        return np.array([float(x) for x in f.read().replace("D", "E").split()])


def convert(src, dest):
    """Parses one reference example run (inputs and outputs under src) into
    the fixture file dest."""
    out = os.path.join(src, "output")
    c = formats.read_control(os.path.join(src, "input", "CONTROL"))
    lattvec = np.asarray(c.lattvec_nm, np.float64)
    positions = np.asarray(c.positions, np.float64).T                       # (natoms, 3) fractional
    types = np.asarray(c.types, np.int32)
    masses = np.array([harmonic.average_mass(e) for e in c.elements])
    ngrid = np.asarray(c.ngrid, np.int32)
    scell = np.asarray(c.scell, np.int32)
    natoms = len(types)
    fc2 = harmonic.read_phonopy_fc2(os.path.join(src, "input", "FORCE_CONSTANTS_2ND"), natoms, scell, masses, types)
    fc3 = harmonic.read_force_constants_3rd(os.path.join(src, "input", "FORCE_CONSTANTS_3RD"), lattvec)
    nptk = int(np.prod(ngrid))
    q = floats(os.path.join(out, "BTE.qpoints")).reshape(-1, 6)
    nlist = len(q)
    List = q[:, 1].astype(np.int32)
    Nequi = q[:, 2].astype(np.int32)
    omega_irr = floats(os.path.join(out, "BTE.omega")).reshape(nlist, -1)
    nb = omega_irr.shape[1]
    v = floats(os.path.join(out, "BTE.v_full")).reshape(nb, nptk, 3)          # (branch, q, alpha)
    velocity = np.ascontiguousarray(v.transpose(1, 0, 2))                     # (q, branch, alpha)
    P3 = np.stack([floats(os.path.join(out, f)).reshape(nlist, nb).T for f in ("BTE.P3_plus", "BTE.P3_minus")])
    P4 = np.stack([floats(os.path.join(out, f)).reshape(nlist, nb).T
                   for f in ("BTE.P4_plusplus", "BTE.P4_plusmius", "BTE.P4_minusminus")])
    N4 = np.loadtxt(os.path.join(out, "BTE.Numprocess_4ph"), dtype=np.int64).T      # (3, nmodes), mm order
    t = os.path.join(out, "T%dK" % int(c.T))

    def wp(name):
        # rows run branch-major (i outer, ll inner); returned in mm order
        x = floats(os.path.join(t, name)).reshape(nb, nlist, 2)
        assert np.allclose(x[..., 0], omega_irr.T, rtol=1e-4, atol=1e-4)
        return np.ascontiguousarray(x[..., 1].T).ravel()

    WP3 = np.stack([wp(f) for f in ("BTE.WP3_plus", "BTE.WP3_minus")])
    W3 = np.stack([wp(f) for f in ("BTE.w_3ph_plus", "BTE.w_3ph_minus", "BTE.w_3ph")])
    WP4 = np.stack([wp(f) for f in ("BTE.WP4_plusplus", "BTE.WP4_plusminus", "BTE.WP4_minusminus")])
    log = open(os.path.join(out, "log.out")).read()
    n3 = [int(re.search(r"Ntotal_%s\s*=\s*(\d+)" % s, log).group(1)) for s in ("plus", "minus")]
    n4 = [int(x) for x in re.search(r"Ntotal_minusminus=\s*\n\s*(\d+)\s+(\d+)\s+(\d+)", log).groups()]
    # LiCoO2's log totals are the sums of BTE.Numprocess_4ph; the BAs example
    # ships a Numprocess_4ph from another run than its log.out (sums 7.9e8 /
    # 5.9e9 / 3.6e8 against 2.2e10 / 1.4e11 / 2.7e10), so both are kept as
    # they are and the tests say which one they compare against
    assert N4.shape == (3, nlist * nb)
    born = np.zeros((3, 3, natoms)) if c.born is None else np.asarray(c.born, np.float64)
    epsilon = np.eye(3) if c.epsilon is None else np.asarray(c.epsilon, np.float64)
    np.savez_compressed(dest, lattvec=lattvec, positions=positions, types=types, masses=masses,
                        elements=np.array(c.elements), scell=scell, fc2=fc2, nonanalytic=bool(c.nonanalytic),
                        born=born, epsilon=epsilon,
                        fc3_phi=fc3.Phi, fc3_Rj=fc3.R_j, fc3_Rk=fc3.R_k,
                        fc3_idx=np.stack([fc3.Index_i, fc3.Index_j, fc3.Index_k]), fc3_cells=fc3.cells,
                        ngrid=ngrid, T=float(c.T), scalebroad=float(c.scalebroad), List=List, Nequi=Nequi,
                        omega_irr=omega_irr, velocity=velocity, P3=P3, P4=P4, N4=N4, WP3=WP3, WP4=WP4, W3=W3,
                        Ntotal3=np.array(n3, np.int64), Ntotal4=np.array(n4, np.int64))
    print("wrote", dest, os.path.getsize(dest), "bytes")


def main():
    convert(SRC, OUT)


if __name__ == "__main__":
    sys.exit(main())
