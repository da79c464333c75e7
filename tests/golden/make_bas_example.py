"""Turns the reference's own BAs example outputs into a test fixture.

The reference ships a complete run of its BAs example (zincblende boron
arsenide, F-43m, 2 atoms, 6 branches, 16x16x16 mesh, scalebroad = 0.1,
T = 300 K, phonopy 2nd-order constants with nonanalytic = .TRUE. and the Born
charges and dielectric tensor of its CONTROL, FORCE_CONSTANTS_3RD).  The same
files as for the LiCoO2 example are read (see make_licoo2_example.py);
unlike LiCoO2, the non-analytic term is not zero here, so the example pins
it.

One more array is derived: nac_images (nptk, 3), the image of each q point
whose direction the run's non-analytic term took.  phonon_phonopy
(Src/phonon_routines.f90:143-159) keeps q unless an image is strictly
shorter by BLAS dnrm2; on the zone boundary several images are equally
short, and which one that comparison keeps turns on the norm's last bit in
the BLAS the run was linked with.  For every such tie the fixture records
the image whose velocities match the run's BTE.v_full (the tied images
differ there by 1e-4 .. 6e-3 km/s, the match is at the printing precision),
and the search's own image everywhere else.  The example's log.out and
T300K/ come from another run than its BTE.* files (3ph totals 671638 /
733206, those of scalebroad = 1.0, and 4ph totals 2.2e10 / 1.4e11 /
2.7e10, against 7.9e8 / 5.9e9 / 3.6e8 summed over BTE.Numprocess_4ph), so
the log's totals are stored as printed and the tests say which run they
compare against.  Run in the development container (the reference tree is
not on the GPU box):

    python tests/golden/make_bas_example.py

and commit the resulting tests/golden/bas_example.npz.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from fourphonon_amd import crystal, harmonic  # noqa: E402
from tests.golden.make_licoo2_example import convert  # noqa: E402

SRC = "/root/reference/Example-BAs"
OUT = os.path.join(HERE, "bas_example.npz")


def nac_images(g):
    st = crystal.Structure("BAs", g["lattvec"], g["positions"], g["types"], g["masses"])
    scell = tuple(int(x) for x in g["scell"])
    kp = harmonic.grid_kpoints(st.rlattvec, g["ngrid"])
    b = st.rlattvec
    shifts = [(x, y, z) for x in range(-2, 3) for y in range(-2, 3) for z in range(-2, 3)]
    symm = crystal.symmetry_tables(st, g["ngrid"])[7]           # q-point symmetrisers, as eigenDM applies them
    out = np.zeros_like(kp)
    nties = 0
    for iq, k in enumerate(kp):
        cands = [k] + [k + x * b[:, 0] + y * b[:, 1] + z * b[:, 2] for x, y, z in shifts if (x, y, z) != (0, 0, 0)]
        d = np.array([np.linalg.norm(r) for r in cands])
        tied = [cands[i] for i in np.flatnonzero(d <= d.min() * (1 + 1e-12))]
        if d.min() == 0.0 or len(tied) == 1:
            out[iq] = tied[0]
            continue
        nties += 1
        err = []
        for r in tied:
            _, v, _ = harmonic.phonons_phonopy(k[None], st, scell, g["fc2"], True, g["epsilon"], g["born"],
                                               images=r[None])
            err.append(np.abs(v[0] @ symm[:, :, iq].T - g["velocity"][iq]).max())
        err = np.array(err)
        best = int(np.argmin(err))
        assert err[best] <= 1e-6, (iq, err)
        # equal results (the tied directions give the same term) or one clear match
        assert np.all((err <= 1e-6) | (err >= 1e-5)), (iq, err)
        out[iq] = tied[best]
    print("zone-boundary ties:", nties)
    return out


def main():
    convert(SRC, OUT)
    g = dict(np.load(OUT))
    g["nac_images"] = nac_images(g)
    np.savez_compressed(OUT, **g)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    sys.exit(main())
