"""Harmonic front end and on-disk formats of the reference's inputs.

* Quantum Espresso 2nd-order constants (``espresso.ifc2``) to omega, group
  velocities and eigenvectors on a q grid, following phonon_espresso
  (Src/phonon_routines.f90:311-677) and eigenDM (:12-85): Wigner-Seitz
  weights of the supercell images, the simple momentum-conservation fix of
  the on-site block, D(q) and dD/dq by the same image sums, Hermitian
  eigenproblem, v = <e|dD/dq|e>/(2 omega), the q-point symmetrisers, and
  exactly zero acoustic modes at Gamma.  The unsuffixed constants of the
  reference (bohr2nm, toTHz, massfactor) are default reals there, so they are
  float32 values here too.
* ``FORCE_CONSTANTS_3RD`` / ``FORCE_CONSTANTS_4TH`` readers and writers
  (read3fc / read4fc, Src/input.f90:102-182): Angstrom vectors rounded to the
  nearest lattice vector and converted to nm.

The polar (non-analytic) correction is not built: a non-zero Born charge
raises ValueError.  Eigenvectors of degenerate modes are one orthonormal
basis of their subspace, as from any Hermitian solver; gauge-invariant
outputs (omega, v, |V|^2 summed over a degenerate set, rates, kappa) are
what they determine.

This module prepares inputs; nothing here is on a timed path.
"""
from __future__ import annotations

import numpy as np

from . import crystal

# default-real (float32) constants of phonon_espresso, Src/phonon_routines.f90:319-320
BOHR2NM = float(np.float32(0.052917721092))
TO_THZ = float(np.float32(20670.687))
MASSFACTOR = float(np.float32(np.float32(1.8218779) * np.float32(6.022e-4)))
WS_TOL = float(np.float32(1e-6))

# Natural-abundance isotopes (mass in amu, abundance in %) for the elements of
# the reference's examples; the average mass is the abundance-weighted mean
# (data_calc_mandg, Src/data.f90:907-929).  Standard values (NIST).
ISOTOPES = {
    "Si": [(27.976928, 92.23), (28.976496, 4.67), (29.973772, 3.1)],
    "B": [(10.012937, 19.9), (11.009305, 80.1)],
    "As": [(74.921596, 100.0)],
}


def average_mass(element: str) -> float:
    m = 0.0
    for mass, ab in ISOTOPES[element]:
        m = m + mass * ab
    return m / 100.0


# ---------------------------------------------------------------------------
# espresso.ifc2
# ---------------------------------------------------------------------------

def read_espresso_ifc2(path):
    """-> dict(ntype, nat, qscell (3,), fc (3,3,nat,nat,s1,s2,s3) in Ry/bohr^2).

    The header (cell, species, positions, dielectric data) is read and, as in
    the reference, not used: the crystal comes from the caller.
    """
    with open(path) as f:
        lines = [ln for ln in f.read().splitlines()]
    pos = 0

    def nxt():
        nonlocal pos
        while not lines[pos].strip():
            pos += 1
        ln = lines[pos]
        pos += 1
        return ln

    head = nxt().split()
    ntype, nat, ibrav = int(head[0]), int(head[1]), int(head[2])
    if ibrav == 0:
        for _ in range(3):
            nxt()
    for _ in range(ntype):
        nxt()
    for _ in range(nat):
        nxt()
    polar = nxt().split()[0].upper().startswith("T")
    if polar:
        for _ in range(3):
            nxt()
        for _ in range(nat):
            for _ in range(4):
                nxt()
    qs = [int(x) for x in nxt().split()[:3]]
    fc = np.zeros((3, 3, nat, nat, *qs), order="F")
    for _ in range(9 * nat * nat):
        ip, jp, ia, ja = (int(x) for x in nxt().split()[:4])
        for _ in range(qs[0] * qs[1] * qs[2]):
            t = nxt().split()
            fc[ip - 1, jp - 1, ia - 1, ja - 1, int(t[0]) - 1, int(t[1]) - 1, int(t[2]) - 1] = float(
                t[3].replace("D", "E"))
    return {"ntype": ntype, "nat": nat, "qscell": np.array(qs), "fc": fc, "polar": polar}


def grid_kpoints(rlattvec, ngrid):
    """kspace of eigenDM (Src/phonon_routines.f90:26-33), in grid-index order;
    (i-1.0)/ngrid(i) is a default-real (float32) quotient there."""
    ijk = crystal.grid_ijk(ngrid)                              # (3, nptk), 0-based
    frac = np.stack([(ijk[a].astype(np.float32) / np.float32(ngrid[a])).astype(np.float64) for a in range(3)])
    b = np.asarray(rlattvec)
    return (b[:, 0][:, None] * frac[0] + b[:, 1][:, None] * frac[1] + b[:, 2][:, None] * frac[2]).T


def _ws_images(cell_r, scell, rr):
    """Images t of the supercell with their Wigner-Seitz weights for one atom
    pair (Src/phonon_routines.f90:440-458, 482-504)."""
    ws = cell_r * np.asarray(scell)[:, None]
    ms = np.array([m for m in np.ndindex(5, 5, 5)]) - 2
    ms = ms[np.any(ms != 0, axis=1)]
    rws = ms @ ws                                              # (124, 3)
    rws0 = 0.5 * np.einsum("ij,ij->i", rws, rws)
    rng = [np.arange(-2 * s, 2 * s + 1) for s in scell]
    M = np.stack(np.meshgrid(*rng, indexing="ij"), -1).reshape(-1, 3)
    t = M @ cell_r                                             # (nimg, 3) bohr
    r_ws = t + rr
    ck = r_ws @ rws.T - rws0                                   # (nimg, 124)
    outside = np.any(ck > WS_TOL, axis=1)
    nreq = 1 + np.sum(np.abs(ck) < WS_TOL, axis=1)
    w = np.where(outside, 0.0, 1.0 / nreq)
    keep = w > 0
    return M[keep], t[keep], w[keep]


def phonons_espresso(kpoints, lattvec, positions, masses, types, scell, fc, born=None):
    """omega (nk,Nb) rad/ps, velocity (nk,Nb,3) km/s and eigenvect (nk,Nb,Nb)
    [mode, 3*(atom-1)+alpha] for the kpoints (nk,3) in nm^-1, before the
    q-point symmetrisers and the Gamma fix-up of eigenDM.  born (3,3,natoms),
    the CONTROL Born charges, must be zero or None: with zero charges the
    reference's non-analytic term vanishes (Src/phonon_routines.f90:545-603)."""
    if born is not None and np.any(np.asarray(born) != 0):
        raise ValueError("the non-analytic (polar) correction is not built; Born charges must be zero")
    lattvec = np.asarray(lattvec, np.float64)
    nat = len(types)
    nb = 3 * nat
    fc = np.array(fc, order="F", copy=True)
    # momentum conservation, Src/phonon_routines.f90:416-423
    for i in range(3):
        for j in range(3):
            for iat in range(nat):
                fc[i, j, iat, iat, 0, 0, 0] = fc[i, j, iat, iat, 0, 0, 0] - fc[i, j, iat].sum()
    mass = np.asarray(masses, np.float64) / MASSFACTOR
    k = np.asarray(kpoints, np.float64) * BOHR2NM
    cell_r = lattvec.T / BOHR2NM                               # rows a_i in bohr
    r = (lattvec @ np.asarray(positions, np.float64).T).T / BOHR2NM
    nk = len(k)
    dyn = np.zeros((nk, nb, nb), np.complex128)
    ddyn = np.zeros((nk, nb, nb, 3), np.complex128)
    for iat in range(nat):
        for jat in range(nat):
            rr = r[iat] - r[jat]
            M, t, w = _ws_images(cell_r, scell, rr)
            idx = np.mod(M, np.asarray(scell))                 # supercell cell of each image
            kt = k @ t.T                                       # (nk, nimg)
            ph = (np.cos(kt) - 1j * np.sin(kt)) * w            # phexp(-kt) * weight
            blk = fc[:, :, iat, jat][:, :, idx[:, 0], idx[:, 1], idx[:, 2]]   # (3,3,nimg)
            dyn[:, 3 * iat:3 * iat + 3, 3 * jat:3 * jat + 3] += np.einsum("abn,kn->kab", blk, ph)
            ddyn[:, 3 * iat:3 * iat + 3, 3 * jat:3 * jat + 3, :] -= 1j * np.einsum("abn,kn,nc->kabc", blk, ph, t)
    mm = np.sqrt(mass[np.asarray(types) - 1][:, None] * mass[np.asarray(types) - 1][None, :])
    mmb = np.repeat(np.repeat(mm, 3, axis=0), 3, axis=1)
    dyn /= mmb
    ddyn /= mmb[:, :, None]
    w2, V = np.linalg.eigh(dyn, UPLO="U")
    omega = np.sign(w2) * np.sqrt(np.abs(w2))
    vel = np.einsum("kci,kcdj,kdi->kij", V.conj(), ddyn, V).real / (2.0 * omega[:, :, None])
    eig = np.ascontiguousarray(np.transpose(V, (0, 2, 1)))      # [mode, component]
    return omega * TO_THZ, vel * TO_THZ * BOHR2NM, eig


def modes_from_espresso(st: crystal.Structure, fc, scell, ngrid, born=None):
    """Modes of eigenDM on the full grid: phonons_espresso at every grid
    point, velocities through the q-point symmetrisers, Gamma acoustic
    omega and v set to zero (Src/phonon_routines.f90:65-80), with the
    irreducible wedge and symmetry tables of crystal.symmetry_tables."""
    ngrid = np.asarray(ngrid, np.int32)
    kp = grid_kpoints(st.rlattvec, ngrid)
    omega, vel, eig = phonons_espresso(kp, st.lattvec, st.positions, st.masses, st.types, scell, fc, born)
    Lst, Nequi, ALLEq, TypeS, ns, Cr, Qr, symm, ID = crystal.symmetry_tables(st, ngrid)
    vel = np.einsum("abq,qnb->qna", symm, vel)
    omega[0, :3] = 0.0
    vel[0, :3, :] = 0.0
    return crystal.Modes(ngrid=ngrid, energy=np.asfortranarray(omega), velocity=np.asfortranarray(vel),
                         eigenvect=np.asfortranarray(eig), IJK=crystal.grid_ijk(ngrid), Nlist=len(Lst), List=Lst,
                         Nequi=Nequi, ALLEquiList=ALLEq,
                         TypeofSymmetry=TypeS, nsymm_rot=ns, crotations=Cr, qrotations=Qr, symmetrizers=symm,
                         ID_equi=ID)


# ---------------------------------------------------------------------------
# FORCE_CONSTANTS_3RD / FORCE_CONSTANTS_4TH
# ---------------------------------------------------------------------------

def _round_to_lattice(R_ang, lattvec):
    """R (3,n) in Angstrom -> nm, rounded to the nearest lattice vector
    (Src/input.f90:133-139: dgesv, then lattvec . anint(x / 10))."""
    x = np.linalg.solve(np.asarray(lattvec), np.asarray(R_ang))
    return np.asfortranarray(np.asarray(lattvec) @ np.rint(x / 10.0))


def _tokens(path):
    with open(path) as f:
        return f.read().replace("D", "E").split()


def read_force_constants_3rd(path, lattvec) -> crystal.IFC3:
    """read3fc, Src/input.f90:102-140."""
    tok = _tokens(path)
    p = 0
    n = int(tok[p]); p += 1
    Phi = np.zeros((3, 3, 3, n), order="F")
    Rj = np.zeros((3, n), order="F")
    Rk = np.zeros((3, n), order="F")
    idx = np.zeros((3, n), np.int32)
    for t in range(n):
        p += 1                                                  # block index
        Rj[:, t] = [float(x) for x in tok[p:p + 3]]; p += 3
        Rk[:, t] = [float(x) for x in tok[p:p + 3]]; p += 3
        idx[:, t] = [int(x) for x in tok[p:p + 3]]; p += 3
        # element order is the loop order (l slowest), whatever the line labels
        for a, b, c in np.ndindex(3, 3, 3):
            Phi[a, b, c, t] = float(tok[p + 3]); p += 4
    Rj_nm, Rk_nm = _round_to_lattice(Rj, lattvec), _round_to_lattice(Rk, lattvec)
    cells = np.stack([np.rint(np.linalg.solve(lattvec, Rj_nm)).T, np.rint(np.linalg.solve(lattvec, Rk_nm)).T], 1)
    return crystal.IFC3(n, Phi, Rj_nm, Rk_nm, *[np.ascontiguousarray(idx[a]) for a in range(3)],
                        cells.astype(np.int64))


def read_force_constants_4th(path, lattvec) -> crystal.IFC4:
    """read4fc, Src/input.f90:141-182."""
    tok = _tokens(path)
    p = 0
    n = int(tok[p]); p += 1
    Phi = np.zeros((3, 3, 3, 3, n), order="F")
    R = [np.zeros((3, n), order="F") for _ in range(3)]
    idx = np.zeros((4, n), np.int32)
    for t in range(n):
        p += 1
        for m in range(3):
            R[m][:, t] = [float(x) for x in tok[p:p + 3]]; p += 3
        idx[:, t] = [int(x) for x in tok[p:p + 4]]; p += 4
        for a, b, c, d in np.ndindex(3, 3, 3, 3):
            Phi[a, b, c, d, t] = float(tok[p + 4]); p += 5
    Rn = [_round_to_lattice(x, lattvec) for x in R]
    cells = np.stack([np.rint(np.linalg.solve(lattvec, x)).T for x in Rn], 1).astype(np.int64)
    return crystal.IFC4(n, Phi, *Rn, *[np.ascontiguousarray(idx[a]) for a in range(4)], cells)


def write_force_constants_3rd(path, ifc3: crystal.IFC3):
    """The sparse text format read3fc reads: Ntri, then per cluster its
    number, R_j and R_k in Angstrom, the three atoms and 27 lines
    'l m n Phi(l,m,n)' with l slowest."""
    with open(path, "w") as f:
        f.write(f"{ifc3.Ntri:12d}\n")
        for t in range(ifc3.Ntri):
            f.write(f"\n{t + 1:12d}\n")
            for R in (ifc3.R_j, ifc3.R_k):
                f.write("".join(f"{10.0 * x:20.10E}" for x in R[:, t]) + "\n")
            f.write(f"{ifc3.Index_i[t]:12d}{ifc3.Index_j[t]:12d}{ifc3.Index_k[t]:12d}\n")
            for a in range(3):
                for b in range(3):
                    for c in range(3):
                        f.write(f"{a + 1:5d}{b + 1:5d}{c + 1:5d}{ifc3.Phi[a, b, c, t]:22.12E}\n")


def write_force_constants_4th(path, ifc4: crystal.IFC4):
    """The sparse text format read4fc reads (R_j, R_k, R_l in Angstrom,
    four atoms, 81 lines 'l m n o Phi(l,m,n,o)')."""
    with open(path, "w") as f:
        f.write(f"{ifc4.Ntri:12d}\n")
        for t in range(ifc4.Ntri):
            f.write(f"\n{t + 1:12d}\n")
            for R in (ifc4.R_j, ifc4.R_k, ifc4.R_l):
                f.write("".join(f"{10.0 * x:20.10E}" for x in R[:, t]) + "\n")
            f.write(f"{ifc4.Index_i[t]:12d}{ifc4.Index_j[t]:12d}{ifc4.Index_k[t]:12d}{ifc4.Index_l[t]:12d}\n")
            for a in range(3):
                for b in range(3):
                    for c in range(3):
                        for d in range(3):
                            f.write(f"{a + 1:5d}{b + 1:5d}{c + 1:5d}{d + 1:5d}{ifc4.Phi[a, b, c, d, t]:22.12E}\n")
