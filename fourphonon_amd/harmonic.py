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

* phonopy ``FORCE_CONSTANTS_2ND`` (read2fc, Src/input.f90:18-79) to the same
  quantities, following phonon_phonopy (Src/phonon_routines.f90:84-307):
  nearest supercell images with the 1e-5 nm tie rule, and the non-analytic
  (dielectric) correction from the CONTROL Born charges and dielectric
  tensor when ``nonanalytic`` is set.

The espresso path has no polar (non-analytic) term: a non-zero Born charge
raises ValueError there.  Eigenvectors of degenerate modes are one orthonormal
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
# (data_calc_mandg, Src/data.f90:907-929).  The values are the reference's
# own table's (Src/data.f90:35-895): its boron abundances are 19.8 / 80.2 %,
# and a 19.9 / 80.1 % table moves the mass by 9e-5 and the BAs example's
# optical frequencies by 5e-5.
ISOTOPES = {
    "Si": [(27.976928, 92.23), (28.976496, 4.67), (29.973772, 3.1)],
    "B": [(10.012938, 19.8), (11.009305, 80.2)],
    "As": [(74.921596, 100.0)],
    "Li": [(6.015123, 7.42), (7.016005, 92.58)],
    "Co": [(58.933198, 100.0)],
    "O": [(15.994915, 99.76), (16.999131, 0.038), (17.999159, 0.2)],
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
    """kspace of eigenDM (Src/phonon_routines.f90:26-33), in grid-index order:
    rlattvec(:,1)*(ii-1.0)/ngrid(1) + ... evaluates left to right, so each
    term is (b * (i-1)) / N in double precision."""
    ijk = crystal.grid_ijk(ngrid).astype(np.float64)          # (3, nptk), 0-based
    b = np.asarray(rlattvec, np.float64)
    t = [(b[:, a][:, None] * ijk[a]) / float(ngrid[a]) for a in range(3)]
    return ((t[0] + t[1]) + t[2]).T


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
# phonopy FORCE_CONSTANTS_2ND, with the non-analytic correction
# ---------------------------------------------------------------------------

# default-real (float32) literals of read2fc and phonon_phonopy
# (Src/input.f90:13, Src/phonon_routines.f90:91)
PHONOPY_UNITFACTOR = float(np.float32(9648.53336213))   # eV/(A^2 amu) -> (rad/ps)^2
NAC_PREFACTOR = float(np.float32(1745.91429109))        # THz^2 amu nm^3


def split_index(index, scell):
    """Supercell atom number (1-based) -> (ix, iy, iz, iatom), 1-based, x
    fastest (split_index, Src/input.f90:83-99)."""
    t, ix = divmod(index - 1, scell[0])
    t, iy = divmod(t, scell[1])
    iatom, iz = divmod(t, scell[2])
    return ix + 1, iy + 1, iz + 1, iatom + 1


def read_phonopy_fc2(path, natoms, scell, masses, types):
    """phonopy's FORCE_CONSTANTS_2ND (read2fc, Src/input.f90:18-79) ->
    fc(natoms,3,s1,s2,s3,natoms,3) in (rad/ps)^2: the blocks whose first atom
    lies in the home cell, divided by sqrt(m_i m_j) and scaled to (rad/ps)^2.
    fc[a, i, cell, b, j] couples atom a (home cell, direction i) with atom b
    in supercell cell `cell` (direction j)."""
    s1, s2, s3 = (int(x) for x in scell)
    with open(path) as f:
        lines = f.read().split("\n")
    pos = 0
    while not lines[pos].strip():
        pos += 1
    ntot = int(lines[pos].split()[0])
    pos += 1
    if ntot != s1 * s2 * s3 * natoms:
        raise ValueError(f"{path}: {ntot} atoms, scell and natoms give {s1 * s2 * s3 * natoms}")
    fc = np.zeros((natoms, 3, s1, s2, s3, natoms, 3), order="F")
    for _ in range(ntot * ntot):
        while not lines[pos].strip():
            pos += 1
        a1, a2 = (int(x) for x in lines[pos].split()[:2])
        ix1, iy1, iz1, i1 = split_index(a1, (s1, s2, s3))
        ix2, iy2, iz2, i2 = split_index(a2, (s1, s2, s3))
        if ix1 == 1 and iy1 == 1 and iz1 == 1:
            for ip in range(3):
                fc[i1 - 1, ip, ix2 - 1, iy2 - 1, iz2 - 1, i2 - 1, :] = [float(x) for x in lines[pos + 1 + ip].split()[:3]]
        pos += 4
    m = np.asarray(masses, np.float64)[np.asarray(types) - 1]
    mm = np.sqrt(m[:, None] * m[None, :])
    mm[np.diag_indices(natoms)] = m
    fc = fc / mm[:, None, None, None, None, :, None]
    return PHONOPY_UNITFACTOR * fc


def _phonopy_images(lattvec, cart, scell):
    """Per (atom 1, atom 2, supercell cell): the supercell images of
    r = r1 - r2 + R_cell nearest to the origin, with the reference's 1e-5 nm
    tie rule applied in its loop order (Src/phonon_routines.f90:216-240)
    -> dict (a1, a2, ix, iy, iz) -> (neq, 3) array of rl + R_cell."""
    nat = cart.shape[1]
    out = {}
    shifts = [(x, y, z) for x in range(-2, 3) for y in range(-2, 3) for z in range(-2, 3)]
    for a1 in range(nat):
        for a2 in range(nat):
            for ix in range(scell[0]):
                for iy in range(scell[1]):
                    for iz in range(scell[2]):
                        rcell = lattvec @ np.array([ix, iy, iz], np.float64)
                        r = cart[:, a1] - cart[:, a2] + rcell
                        dmin = np.finfo(np.float64).max
                        rr = []
                        for x, y, z in shifts:
                            rl = (x * scell[0] * lattvec[:, 0] + y * scell[1] * lattvec[:, 1]
                                  + z * scell[2] * lattvec[:, 2])
                            rn = float(np.sqrt(np.sum((rl + r) ** 2)))
                            if abs(rn - dmin) > 1e-5:
                                if rn < dmin:
                                    dmin = rn
                                    rr = [rl + rcell]
                            else:
                                rr.append(rl + rcell)
                        out[(a1, a2, ix, iy, iz)] = np.array(rr)
    return out


def phonons_phonopy(kpoints, st: crystal.Structure, scell, fc, nonanalytic=False, epsilon=None, born=None,
                    images=None):
    """omega (nk,Nb) rad/ps, velocity (nk,Nb,3) km/s and eigenvect (nk,Nb,Nb)
    [mode, 3*(atom-1)+alpha] for the kpoints (nk,3) in nm^-1 from the
    reduced constants of read_phonopy_fc2 (phonon_phonopy,
    Src/phonon_routines.f90:84-307): D(q) = sum over cells and nearest
    supercell images of e^{-i q.R}/neq fc, dD/dq alike, and, when
    nonanalytic, the dielectric correction (Born charges born(3,3,natoms),
    dielectric tensor epsilon(3,3)) added to every cell's constants and to
    dD/dq, with q taken at its shortest image (no correction at Gamma), or
    at images (nk,3) when given."""
    lattvec = np.asarray(st.lattvec, np.float64)
    rlatt = st.rlattvec
    nat = st.natoms
    nb = 3 * nat
    s1, s2, s3 = (int(x) for x in scell)
    ncell = s1 * s2 * s3
    k = np.asarray(kpoints, np.float64)
    nk = len(k)
    m = np.asarray(st.masses, np.float64)[np.asarray(st.types) - 1]
    mm = np.sqrt(m[:, None] * m[None, :])
    mm[np.diag_indices(nat)] = m
    cart = lattvec @ np.asarray(st.positions, np.float64).T         # (3, nat)
    eps = np.eye(3) if epsilon is None else np.asarray(epsilon, np.float64)
    bz = np.zeros((3, 3, nat)) if born is None else np.asarray(born, np.float64)
    V = st.volume
    # the 1st-BZ image of each q point (Src/phonon_routines.f90:143-159): q
    # itself, replaced by an image in x, y, z order only when strictly
    # shorter.  On the zone boundary several images are equally short and
    # give non-analytic terms of different directions; the strict comparison
    # then turns on the last bit of the norm (BLAS dnrm2 in the reference),
    # so callers that must follow one particular run pass its images.
    if images is not None:
        shortest = np.array(images, np.float64).reshape(nk, 3)
    else:
        shortest = k.copy()
        best = np.sqrt(np.sum(k * k, axis=1))
        for x in range(-2, 3):
            for y in range(-2, 3):
                for z in range(-2, 3):
                    r = k + x * rlatt[:, 0] + y * rlatt[:, 1] + z * rlatt[:, 2]
                    d = np.sqrt(np.sum(r * r, axis=1))
                    better = d < best
                    best = np.where(better, d, best)
                    shortest[better] = r[better]
    # dielectric correction per q: dyn_nac and ddyn_nac (:168-197)
    dyn_nac = np.zeros((nk, nb, nb))
    ddyn_nac = np.zeros((nk, nb, nb, 3))
    nac_on = np.zeros(nk, bool)
    if nonanalytic:
        np_err = np.seterr(divide="ignore", invalid="ignore")   # Gamma rows are discarded below
        nac_on = ~np.all(shortest == 0.0, axis=1)
        sq = shortest
        tmp3 = np.einsum("ka,ab,kb->k", sq, eps, sq)
        t = np.einsum("ka,aib->kib", sq, bz)                      # t[k, i, atom] = q . born(:, i, atom)
        es = sq @ eps.T                                             # (eps q)(ip)
        for a1 in range(nat):
            for a2 in range(nat):
                for i in range(3):
                    for j in range(3):
                        t1, t2 = t[:, i, a1], t[:, j, a2]
                        r_, c_ = 3 * a1 + i, 3 * a2 + j
                        dyn_nac[:, r_, c_] = t1 * t2 / mm[a1, a2]
                        for ip in range(3):
                            ddyn_nac[:, r_, c_, ip] = (t1 * bz[ip, j, a2] + t2 * bz[ip, i, a1]
                                                       - 2. * t1 * t2 * es[:, ip] / tmp3)
                        ddyn_nac[:, r_, c_, :] /= mm[a1, a2]
        safe = np.where(nac_on, tmp3, 1.0)
        dyn_nac = np.where(nac_on[:, None, None], NAC_PREFACTOR * dyn_nac / safe[:, None, None] / V, 0.0)
        ddyn_nac = np.where(nac_on[:, None, None, None], NAC_PREFACTOR * ddyn_nac / safe[:, None, None, None] / V,
                            0.0)
        np.seterr(**np_err)
    fc_diel = dyn_nac / ncell                                       # the same for every cell (:198-208)
    images = _phonopy_images(lattvec, cart, (s1, s2, s3))
    dyn = np.zeros((nk, nb, nb), np.complex128)
    ddyn = np.zeros((nk, nb, nb, 3), np.complex128)
    kq_on = nonanalytic & ~np.all(k == 0.0, axis=1)
    for a1 in range(nat):
        for a2 in range(nat):
            star_all = np.zeros(nk, np.complex128)
            zr_all = np.zeros((nk, 3), np.complex128)
            blk = np.zeros((nk, 3, 3), np.complex128)
            dblk = np.zeros((nk, 3, 3, 3), np.complex128)
            for ix in range(s1):
                for iy in range(s2):
                    for iz in range(s3):
                        rr = images[(a1, a2, ix, iy, iz)]
                        neq = len(rr)
                        qr = k @ rr.T                                   # (nk, neq)
                        z = (np.cos(-qr) + 1j * np.sin(-qr)) / neq       # phexp(-qr)/neq
                        star = z.sum(axis=1)
                        zr = z @ rr                                     # sum_p z_p rr_p
                        # fc_total(iatom2, j, cell, iatom1, i) for (i, j)
                        f = fc[a2, :, ix, iy, iz, a1, :].T              # [i, j]
                        blk += star[:, None, None] * f[None]
                        dblk -= 1j * zr[:, None, None, :] * f[None, :, :, None]
                        star_all += star
                        zr_all += zr
            # the dielectric constants enter every cell's fc_total alike
            fd = fc_diel[:, 3 * a2:3 * a2 + 3, 3 * a1:3 * a1 + 3].transpose(0, 2, 1)   # [k, i, j]
            blk += star_all[:, None, None] * fd
            dblk -= 1j * zr_all[:, None, None, :] * fd[..., None]
            # nonanalytic part of dD/dq, star times ddyn_nac / ncell per cell (:249-257)
            dn = ddyn_nac[:, 3 * a1:3 * a1 + 3, 3 * a2:3 * a2 + 3, :]
            dblk += np.where(kq_on[:, None, None, None], star_all[:, None, None, None] * dn / ncell, 0.0)
            dyn[:, 3 * a1:3 * a1 + 3, 3 * a2:3 * a2 + 3] = blk
            ddyn[:, 3 * a1:3 * a1 + 3, 3 * a2:3 * a2 + 3, :] = dblk
    w2, Vv = np.linalg.eigh(dyn, UPLO="U")
    omega = np.sign(w2) * np.sqrt(np.abs(w2))
    vel = np.einsum("kci,kcdj,kdi->kij", Vv.conj(), ddyn, Vv).real / (2.0 * omega[:, :, None])
    eig = np.ascontiguousarray(np.transpose(Vv, (0, 2, 1)))
    return omega, vel, eig


def modes_from_phonopy(st: crystal.Structure, fc, scell, ngrid, nonanalytic=False, epsilon=None, born=None,
                       images=None):
    """eigenDM with the phonopy front end on the full grid (as
    modes_from_espresso: q-point symmetrisers on v, Gamma acoustic omega and
    v set to zero), with the irreducible wedge and symmetry tables.  images
    (nptk,3), grid-index order: the non-analytic directions, as in
    phonons_phonopy."""
    ngrid = np.asarray(ngrid, np.int32)
    kp = grid_kpoints(st.rlattvec, ngrid)
    omega, vel, eig = phonons_phonopy(kp, st, scell, fc, nonanalytic, epsilon, born, images)
    Lst, Nequi, ALLEq, TypeS, ns, Cr, Qr, symm, ID = crystal.symmetry_tables(st, ngrid)
    vel = np.einsum("abq,qnb->qna", symm, vel)
    raw_gamma = omega[0, :3].copy()
    omega[0, :3] = 0.0
    vel[0, :3, :] = 0.0
    modes = crystal.Modes(ngrid=ngrid, energy=np.asfortranarray(omega), velocity=np.asfortranarray(vel),
                          eigenvect=np.asfortranarray(eig), IJK=crystal.grid_ijk(ngrid), Nlist=len(Lst), List=Lst,
                          Nequi=Nequi, ALLEquiList=ALLEq, TypeofSymmetry=TypeS, nsymm_rot=ns, crotations=Cr,
                          qrotations=Qr, symmetrizers=symm, ID_equi=ID)
    modes.raw_gamma = raw_gamma
    return modes


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
