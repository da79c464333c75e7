"""Turns the reference's own Si sampling example outputs into a test fixture.

Example-Si-sampling is the Si example's crystal and constants (its CONTROL
differs in the mesh, 32x32x32, and in the sampling parameters:
num_sample_process_3ph(_phase_space) = -1, i.e. the full 3ph drivers, and
num_sample_process_4ph(_phase_space) = 100000, the 4ph sampling estimators
of Src/processes.f90:2480-2746, 4811-5171); its espresso.ifc2 and
FORCE_CONSTANTS_3RD are byte-identical to Example-Si's, so the tests take
the constants from tests/golden/si_example.npz and this fixture holds the
run's outputs only:

* BTE.qpoints (List, Nequi), BTE.omega and BTE.v (irreducible points; rows
  branch-major);
* BTE.P3_plus / minus (full 3ph phase space), T300K/BTE.WP3_*, BTE.w_3ph*
  (full 3ph rates), and the log's exact 3ph totals;
* the 4ph estimates: BTE.Numprocess_4ph (INT(REAL(hits)/num*total) per
  mode), BTE.P4_*, T300K/BTE.WP4_*, and the log's estimated totals.

The run's random stream is clock-seeded, so the 4ph outputs are compared
statistically.  Its BTE.w_4ph* need the fourth-order constants, which the
example does not ship.  Run in the development container (the reference
tree is not on the GPU box):

    python tests/golden/make_si_sampling_example.py

and commit the resulting tests/golden/si_sampling_example.npz.
"""
import filecmp
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/Example-Si-sampling"
SI = "/root/reference/Example-Si"
OUT = os.path.join(HERE, "si_sampling_example.npz")


def floats(path):
    with open(path) as f:
        return np.array([float(x) for x in f.read().replace("D", "E").split()])


def main():
    for f in ("espresso.ifc2", "FORCE_CONSTANTS_3RD"):
        assert filecmp.cmp(os.path.join(SRC, "input", f), os.path.join(SI, "input", f), shallow=False), f
    text = open(os.path.join(SRC, "input", "CONTROL")).read()
    ngrid = np.array([int(x) for x in re.search(r"ngrid\(:\)\s*=\s*(\d+)\s+(\d+)\s+(\d+)", text).groups()])
    T = float(re.search(r"\bT\s*=\s*([0-9.]+)", text).group(1))
    scalebroad = float(re.search(r"scalebroad\s*=\s*([0-9.]+)", text).group(1))
    nsample = {k: int(re.search(r"\b%s\s*=\s*(-?\d+)" % k, text).group(1))
               for k in ("num_sample_process_3ph_phase_space", "num_sample_process_3ph",
                         "num_sample_process_4ph_phase_space", "num_sample_process_4ph")}
    out = os.path.join(SRC, "output")
    q = floats(os.path.join(out, "BTE.qpoints")).reshape(-1, 6)
    nlist = len(q)
    List = q[:, 1].astype(np.int32)
    Nequi = q[:, 2].astype(np.int32)
    omega_irr = floats(os.path.join(out, "BTE.omega")).reshape(nlist, -1)
    nb = omega_irr.shape[1]
    v_irr = floats(os.path.join(out, "BTE.v")).reshape(nb, nlist, 3).transpose(1, 0, 2)   # (ll, branch, alpha)
    P3 = np.stack([floats(os.path.join(out, f)).reshape(nlist, nb).T for f in ("BTE.P3_plus", "BTE.P3_minus")])
    P4 = np.stack([floats(os.path.join(out, f)).reshape(nlist, nb).T
                   for f in ("BTE.P4_plusplus", "BTE.P4_plusmius", "BTE.P4_minusminus")])
    N4 = np.loadtxt(os.path.join(out, "BTE.Numprocess_4ph"), dtype=np.int64).T      # (3, nmodes), mm order
    t = os.path.join(out, "T%dK" % int(T))

    def wp(name):
        x = floats(os.path.join(t, name)).reshape(nb, nlist, 2)
        assert np.allclose(x[..., 0], omega_irr.T, rtol=1e-4, atol=1e-4)
        return np.ascontiguousarray(x[..., 1].T).ravel()

    WP3 = np.stack([wp(f) for f in ("BTE.WP3_plus", "BTE.WP3_minus")])
    W3 = np.stack([wp(f) for f in ("BTE.w_3ph_plus", "BTE.w_3ph_minus", "BTE.w_3ph")])
    WP4 = np.stack([wp(f) for f in ("BTE.WP4_plusplus", "BTE.WP4_plusminus", "BTE.WP4_minusminus")])
    log = open(os.path.join(out, "log.out"), errors="replace").read()
    n3 = [int(re.search(r"Ntotal_%s\s*=\s*(\d+)" % s, log).group(1)) for s in ("plus", "minus")]
    n4 = [int(x) for x in re.search(r"Ntotal_minusminus=\s*\n\s*(\d+)\s+(\d+)\s+(\d+)", log).groups()]
    assert N4.shape == (3, nlist * nb) and list(N4.sum(axis=1)) == n4
    np.savez_compressed(OUT, ngrid=ngrid, T=T, scalebroad=scalebroad,
                        nsample=np.array([nsample[k] for k in sorted(nsample)]),
                        nsample_keys=np.array(sorted(nsample)), List=List, Nequi=Nequi, omega_irr=omega_irr,
                        v_irr=v_irr, P3=P3, P4=P4, N4=N4, WP3=WP3, WP4=WP4, W3=W3,
                        Ntotal3=np.array(n3, np.int64), Ntotal4=np.array(n4, np.int64))
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    sys.exit(main())
