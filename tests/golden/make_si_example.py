"""Turns the reference's own Si example outputs into a test fixture.

The reference ships a complete run of its Si example (16x16x16 mesh,
scalebroad = 1, T = 300 K): inputs ``CONTROL``, ``espresso.ifc2`` (2nd-order
constants) and ``FORCE_CONSTANTS_3RD``, and these output files (data, read
as text):

* ``BTE.qpoints`` (List, Nequi), ``BTE.omega`` (omega of the irreducible
  points, rad/ps), ``BTE.v_full`` (group velocities on the full grid, km/s),
  written by Src/ShengBTE.f90:228-311;
* ``BTE.P3_plus`` / ``BTE.P3_minus`` (NP_driver's phase space),
  ``BTE.P4_plusplus`` / ``BTE.P4_plusmius`` / ``BTE.P4_minusminus`` and
  ``BTE.Numprocess_4ph`` (NP_driver_4ph's phase space and counts, mode mm in
  row order), Src/ShengBTE.f90:500-560;
* ``T300K/BTE.WP3_*`` and ``T300K/BTE.WP4_*`` (weighted phase space, five
  digits), ``T300K/BTE.w_3ph*`` (3ph RTA rates, ten digits; rows
  branch-major), and the totals of ``log.out``.

The inputs are parsed with fourphonon_amd.harmonic's readers and stored as
arrays; the tests rebuild omega, v and e with the harmonic front end and
compare every stage against the outputs.  Run in the development container
(the reference tree is not on the GPU box):

    python tests/golden/make_si_example.py

and commit the resulting tests/golden/si_example.npz.
"""
import os
import re
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from fourphonon_amd import harmonic  # noqa: E402

SRC = "/root/reference/Example-Si"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "si_example.npz")


def floats(path):
    with open(path) as f:
        return np.array([float(x) for x in f.read().replace("D", "E").split()])


def control(path):
    text = open(path).read()
    lf = float(re.search(r"lfactor\s*=\s*([-0-9.eEdD+]+)", text).group(1))
    cols = []
    for i in (1, 2, 3):
        m = re.search(r"lattvec\(:,%d\)\s*=\s*([-0-9.eE+]+)\s+([-0-9.eE+]+)\s+([-0-9.eE+]+)" % i, text)
        cols.append([float(m.group(k)) for k in (1, 2, 3)])
    ngrid = [int(x) for x in re.search(r"ngrid\(:\)\s*=\s*(\d+)\s+(\d+)\s+(\d+)", text).groups()]
    T = float(re.search(r"\bT\s*=\s*([0-9.]+)", text).group(1))
    sb = float(re.search(r"scalebroad\s*=\s*([0-9.]+)", text).group(1))
    scell = [int(x) for x in re.search(r"scell\(:\)\s*=\s*(\d+)\s+(\d+)\s+(\d+)", text).groups()]
    elem = re.search(r"elements\s*=\s*\"(\w+)\"", text).group(1)
    pos = [[float(x) for x in m] for m in
           re.findall(r"positions\(:,\d+\)\s*=\s*([-0-9.eE+]+)\s+([-0-9.eE+]+)\s+([-0-9.eE+]+)", text)]
    return lf * np.array(cols).T, np.array(ngrid), T, sb, np.array(scell), elem, np.array(pos)


def main():
    out = os.path.join(SRC, "output")
    lattvec, ngrid, T, scalebroad, scell, elem, positions = control(os.path.join(SRC, "input", "CONTROL"))
    ifc2 = harmonic.read_espresso_ifc2(os.path.join(SRC, "input", "espresso.ifc2"))
    fc3 = harmonic.read_force_constants_3rd(os.path.join(SRC, "input", "FORCE_CONSTANTS_3RD"), lattvec)
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
    t = os.path.join(out, "T%dK" % int(T))

    def wp(name):
        # rows run branch-major (i outer, ll inner, Src/ShengBTE.f90:623-628); returned in mm order
        x = floats(os.path.join(t, name)).reshape(nb, nlist, 2)
        assert np.allclose(x[..., 0], omega_irr.T, rtol=1e-4, atol=1e-4)
        return np.ascontiguousarray(x[..., 1].T).ravel()

    WP3 = np.stack([wp(f) for f in ("BTE.WP3_plus", "BTE.WP3_minus")])
    W3 = np.stack([wp(f) for f in ("BTE.w_3ph_plus", "BTE.w_3ph_minus", "BTE.w_3ph")])
    WP4 = np.stack([wp(f) for f in ("BTE.WP4_plusplus", "BTE.WP4_plusminus", "BTE.WP4_minusminus")])
    log = open(os.path.join(out, "log.out")).read()
    n3 = [int(re.search(r"Ntotal_%s\s*=\s*(\d+)" % s, log).group(1)) for s in ("plus", "minus")]
    n4 = [int(x) for x in re.search(r"Ntotal_minusminus=\s*\n\s*(\d+)\s+(\d+)\s+(\d+)", log).groups()]
    assert N4.shape == (3, nlist * nb) and list(N4.sum(axis=1)) == n4
    np.savez_compressed(OUT, lattvec=lattvec, positions=positions,
                        mass=harmonic.average_mass(elem), scell=scell, fc2=ifc2["fc"],
                        fc3_phi=fc3.Phi, fc3_Rj=fc3.R_j, fc3_Rk=fc3.R_k,
                        fc3_idx=np.stack([fc3.Index_i, fc3.Index_j, fc3.Index_k]), fc3_cells=fc3.cells,
                        ngrid=ngrid, T=T, scalebroad=scalebroad, List=List, Nequi=Nequi,
                        omega_irr=omega_irr, velocity=velocity, P3=P3, P4=P4, N4=N4, WP3=WP3, WP4=WP4, W3=W3,
                        Ntotal3=np.array(n3, np.int64), Ntotal4=np.array(n4, np.int64))
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    sys.exit(main())
