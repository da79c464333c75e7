"""Seeded synthetic crystals and their harmonic data in the reference's conventions.

This module produces every array the four-phonon hot path consumes, laid out
exactly as the reference holds them in memory (Fortran column-major, 1-based
indices stored in the data):

* ``energy(nptk,Nb)`` in rad/ps, ``velocity(nptk,Nb,3)`` in km/s (= nm/ps),
  ``eigenvect(nptk,Nb,Nb)`` complex with the Cartesian component
  ``3*(atom-1)+alpha`` last (``Src/ShengBTE.f90:142-149``,
  ``Src/phonon_routines.f90:289``);
* ``IJK(3,nptk)`` 0-based grid coordinates, grid index
  ``1+i1+N1*(i2+N2*i3)`` (``Src/config.f90:458-478``);
* the irreducible wedge ``List``, ``Nequi``, ``ALLEquiList``,
  ``TypeofSymmetry`` built by the reference's first-seen orbit walk
  (``Src/wedgetc.f90:20-75``) over rotations doubled by time reversal with
  duplicates dropped (``Src/config.f90:272-376``);
* fourth-order clusters ``Phi(3,3,3,3,Ntri)`` in eV/A^4, ``R_j/R_k/R_l(3,Ntri)``
  in nm and exactly on lattice vectors, ``Index_i..l(Ntri)`` 1-based atoms
  (``Src/input.f90:141-182``);
* third-order clusters ``Phi(3,3,3,Ntri)`` in eV/A^3, ``R_j/R_k(3,Ntri)`` in
  nm, ``Index_i..k(Ntri)`` (``Src/input.f90:102-140``), generated the same
  way as the fourth-order ones.

The harmonic model is a nearest-neighbour central + transverse spring model
(every bond contributes ``-(k_c n n^T + k_t (1 - n n^T))``; the on-site block
restores the acoustic sum rule), so the dynamical matrix is positive definite
away from Gamma for any connected bond graph.  Jitter is drawn once per spring
family so that the crystal keeps its full point-group symmetry.  Fourth-order
constants are seeded N(0,1) values on every ordered cluster whose atoms are
pairwise within the cutoff, projected alternately onto the index-permutation
symmetric subspace and the acoustic-sum-rule subspace until both hold to
round-off.

The dynamical-matrix phase convention is D_IJ(q) = sum_R Phi(I0;JR) e^{+iq.R}
/ sqrt(m_I m_J), the one under which the matrix element's e^{+iq'.R_j} phase
for a phonon entering with +q' is consistent (``Src/processes.f90:180-199``).

This module is input preparation for tests, the benchmark and the Fortran
driver; nothing here is on a timed path.
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass, field

import numpy as np

# eV / (A^2 amu) -> (rad/ps)^2
UNITFACTOR = 1.602176634e-19 / (1e-20 * 1.66053906660e-27) * 1e-24
SEED_BASE = 20261203


@dataclass
class Structure:
    name: str
    lattvec: np.ndarray          # (3,3) nm, columns are a_1..a_3 (Fortran lattvec(:,i))
    positions: np.ndarray        # (natoms,3) fractional
    types: np.ndarray            # (natoms,) 1-based element index
    masses: np.ndarray           # (nelements,) amu

    @property
    def natoms(self) -> int:
        return len(self.types)

    @property
    def rlattvec(self) -> np.ndarray:
        # b_i . a_j = 2 pi delta_ij, columns are b_1..b_3 (nm^-1)
        return 2.0 * np.pi * np.linalg.inv(self.lattvec).T

    @property
    def volume(self) -> float:
        return abs(np.linalg.det(self.lattvec))

    def cart(self, frac) -> np.ndarray:
        return np.asarray(frac, dtype=np.float64) @ self.lattvec.T


@dataclass
class IFC4:
    Ntri: int
    Phi: np.ndarray              # (3,3,3,3,Ntri) Fortran order, eV/A^4
    R_j: np.ndarray              # (3,Ntri) Fortran order, nm
    R_k: np.ndarray
    R_l: np.ndarray
    Index_i: np.ndarray          # (Ntri,) int32, 1-based
    Index_j: np.ndarray
    Index_k: np.ndarray
    Index_l: np.ndarray
    cells: np.ndarray            # (Ntri,3,3) int lattice coordinates of j,k,l


@dataclass
class IFC3:
    Ntri: int
    Phi: np.ndarray              # (3,3,3,Ntri) Fortran order, eV/A^3
    R_j: np.ndarray              # (3,Ntri) Fortran order, nm
    R_k: np.ndarray
    Index_i: np.ndarray          # (Ntri,) int32, 1-based
    Index_j: np.ndarray
    Index_k: np.ndarray
    cells: np.ndarray            # (Ntri,2,3) int lattice coordinates of j,k


@dataclass
class Modes:
    ngrid: np.ndarray            # (3,) int32
    energy: np.ndarray           # (nptk,Nb) F-order
    velocity: np.ndarray         # (nptk,Nb,3) F-order
    eigenvect: np.ndarray        # (nptk,Nb,Nb) F-order complex128
    IJK: np.ndarray              # (3,nptk) F-order int32
    Nlist: int
    List: np.ndarray             # (Nlist,) int32 1-based
    Nequi: np.ndarray            # (Nlist,) int32
    ALLEquiList: np.ndarray      # (Nsymm_rot,nptk) F-order int32
    TypeofSymmetry: np.ndarray   # (Nsymm_rot,nptk) F-order int32
    nsymm_rot: int
    crotations: np.ndarray       # (3,3,Nsymm_rot) F-order
    qrotations: np.ndarray       # (3,3,Nsymm_rot) F-order, reciprocal lattice coordinates
    symmetrizers: np.ndarray     # (3,3,nptk) F-order
    ID_equi: np.ndarray          # (Nsymm_rot,nptk) F-order int32 (full-grid symmetry map)

    @property
    def nptk(self) -> int:
        return int(np.prod(self.ngrid))

    @property
    def nbands(self) -> int:
        return self.energy.shape[1]


@dataclass
class Case:
    """Everything one hot-path invocation needs, plus the run parameters."""
    structure: Structure
    modes: Modes
    ifc4: IFC4
    T: float = 300.0
    scalebroad: float = 1.0
    omega_max: float = 1.0e100
    onlyharmonic: bool = False
    meta: dict = field(default_factory=dict)
    ifc3: IFC3 | None = None

    @property
    def nptk(self) -> int:
        return self.modes.nptk

    @property
    def nbands(self) -> int:
        return self.modes.nbands


# ---------------------------------------------------------------------------
# structures
# ---------------------------------------------------------------------------

def diamond(a_ang: float = 5.43, mass: float = 28.085) -> Structure:
    a = a_ang / 10.0
    lat = a * np.array([[0.0, 0.5, 0.5], [0.5, 0.0, 0.5], [0.5, 0.5, 0.0]]).T
    pos = np.array([[0.0, 0.0, 0.0], [0.25, 0.25, 0.25]])
    return Structure("diamond", lat, pos, np.array([1, 1], np.int32), np.array([mass]))


def wurtzite(a_ang: float = 3.2, c_ang: float = 5.2, m1: float = 26.98, m2: float = 14.01,
             u: float = 0.375) -> Structure:
    a, c = a_ang / 10.0, c_ang / 10.0
    lat = np.array([[a, 0.0, 0.0], [-0.5 * a, np.sqrt(3.0) / 2.0 * a, 0.0], [0.0, 0.0, c]]).T
    pos = np.array([[1 / 3, 2 / 3, 0.0], [2 / 3, 1 / 3, 0.5],
                    [1 / 3, 2 / 3, u], [2 / 3, 1 / 3, 0.5 + u]])
    return Structure("wurtzite", lat, pos, np.array([1, 1, 2, 2], np.int32), np.array([m1, m2]))


def cubic_family(kind: str, natoms: int, a_ang: float, masses) -> Structure:
    """fcc / zincblende / rocksalt-like cells with 2, 4, 6 or 8 atoms.

    The primitive fcc cell is stacked along a_3 to reach ``natoms``; the basis
    alternates between the two sublattice sites of the parent structure.
    """
    a = a_ang / 10.0
    nrep = max(1, natoms // 2)
    base = a * np.array([[0.0, 0.5, 0.5], [0.5, 0.0, 0.5], [0.5, 0.5, 0.0]]).T
    lat = base.copy()
    lat[:, 2] *= nrep
    second = {"fcc": 0.5, "zincblende": 0.25, "rocksalt": 0.5}[kind]
    frac2 = np.array([second, second, second])
    pos, types = [], []
    for r in range(nrep):
        shift = np.array([0.0, 0.0, r]) @ base.T          # Cartesian shift by r*a_3
        for s, t in ((np.zeros(3), 1), (frac2 @ base.T, 2)):
            cart = shift + s
            pos.append(np.linalg.solve(lat, cart))
            types.append(t)
    pos = np.array(pos) % 1.0
    masses = np.asarray(masses, dtype=np.float64)
    types = np.array(types, np.int32)
    if kind == "fcc":
        types[:] = 1
        masses = masses[:1]
    return Structure(f"{kind}{natoms}", lat, pos[:natoms], types[:natoms], masses)


# ---------------------------------------------------------------------------
# neighbours and clusters
# ---------------------------------------------------------------------------

def _image_table(st: Structure, rmax: float):
    """All (atom, cell) sites within rmax of some home-cell atom."""
    lat = st.lattvec
    # enough cells to cover rmax in every direction
    inv = np.linalg.inv(lat)
    span = np.ceil(rmax * np.linalg.norm(inv, axis=1)).astype(int) + 1
    rng = [np.arange(-s, s + 1) for s in span]
    cells = np.array(list(itertools.product(*rng)), dtype=np.int64)
    sites = []
    for c in cells:
        for J in range(st.natoms):
            sites.append((J, c[0], c[1], c[2]))
    sites = np.array(sites, dtype=np.int64)
    xyz = (st.positions[sites[:, 0]] + sites[:, 1:]) @ lat.T
    return sites, xyz


def shell_distances(st: Structure, nshell: int = 4) -> np.ndarray:
    sites, xyz = _image_table(st, 3.0 * np.cbrt(st.volume / st.natoms) + 0.5)
    d = []
    for I in range(st.natoms):
        r = np.linalg.norm(xyz - st.cart(st.positions[I]), axis=1)
        d.append(r[r > 1e-9])
    d = np.unique(np.round(np.concatenate(d), 6))
    return d[:nshell]


def cluster_cutoff(st: Structure, shell: int) -> float:
    d = shell_distances(st, shell + 1)
    return 0.5 * (d[shell - 1] + d[shell])


def _neighbours(st: Structure, rc: float):
    sites, xyz = _image_table(st, rc)
    out = []
    for I in range(st.natoms):
        r = np.linalg.norm(xyz - st.cart(st.positions[I]), axis=1)
        sel = np.nonzero(r <= rc + 1e-9)[0]          # includes the atom itself
        out.append((sites[sel], xyz[sel]))
    return out


def enumerate_quartets(st: Structure, rc: float):
    """Ordered quartets (I@0, j, k, l), atoms pairwise within rc, repeats allowed.

    Returns sites (n,4,4) int: [atom, c1, c2, c3] per position, first at cell 0.
    """
    out = []
    for I, (nsites, nxyz) in enumerate(_neighbours(st, rc)):
        n = len(nsites)
        dmat = np.linalg.norm(nxyz[:, None, :] - nxyz[None, :, :], axis=2)
        ok = dmat <= rc + 1e-9
        home = np.nonzero((nsites[:, 0] == I) & np.all(nsites[:, 1:] == 0, axis=1))[0][0]
        # every other site in the table is already within rc of the home atom
        for a in range(n):
            if not ok[home, a]:
                continue
            for b in range(n):
                if not (ok[home, b] and ok[a, b]):
                    continue
                cs = np.nonzero(ok[home] & ok[a] & ok[b])[0]
                for c in cs:
                    out.append((nsites[home], nsites[a], nsites[b], nsites[c]))
    return np.array(out, dtype=np.int64)


def enumerate_triplets(st: Structure, rc: float):
    """Ordered triplets (I@0, j, k), atoms pairwise within rc, repeats allowed.

    Returns sites (n,3,4) int: [atom, c1, c2, c3] per position, first at cell 0.
    """
    out = []
    for I, (nsites, nxyz) in enumerate(_neighbours(st, rc)):
        n = len(nsites)
        dmat = np.linalg.norm(nxyz[:, None, :] - nxyz[None, :, :], axis=2)
        ok = dmat <= rc + 1e-9
        home = np.nonzero((nsites[:, 0] == I) & np.all(nsites[:, 1:] == 0, axis=1))[0][0]
        for a in range(n):
            if not ok[home, a]:
                continue
            for b in np.nonzero(ok[home] & ok[a])[0]:
                out.append((nsites[home], nsites[a], nsites[b]))
    return np.array(out, dtype=np.int64)


def _canon(sites):
    """Translate each quartet so that its first site sits in cell 0."""
    s = sites.copy()
    s[:, :, 1:] -= s[:, :1, 1:]
    return s


def _keys(sites):
    return [tuple(x) for x in sites.reshape(len(sites), -1)]


def space_group_ops(st: Structure, tol: float = 1e-6):
    """Space-group operations x -> M x + t (fractional) of the crystal, each
    with the image of every basis atom: M pos[I] + t = pos[amap[I]] + shift[I].
    One translation per rotation (primitive cells)."""
    lat = st.lattvec
    ops = []
    for M in point_group(st, tol):
        mp = st.positions @ M.T
        for J in range(st.natoms):
            if st.types[J] != st.types[0]:
                continue
            t = st.positions[J] - mp[0]
            amap, shift = [], []
            for I in range(st.natoms):
                x = mp[I] + t
                d = st.positions - x
                dr = d - np.round(d)
                hit = np.nonzero((np.linalg.norm(dr @ lat.T, axis=1) < 1e-6) & (st.types == st.types[I]))[0]
                if hit.size == 0:
                    break
                amap.append(int(hit[0]))
                shift.append(np.round(x - st.positions[hit[0]]).astype(np.int64))
            if len(amap) == st.natoms:
                ops.append((M, t, np.array(amap), np.array(shift)))
                break
    return ops


def _space_group_projector(st: Structure, sites):
    """(maps, rots): for every space-group operation g, the index of the image
    quartet/triplet g.k of every cluster k (translated so its first site is in
    cell 0) and the Cartesian rotation of g.  The cluster set must be closed
    under the group (it is: clusters are chosen by distances)."""
    index = {k: i for i, k in enumerate(_keys(sites))}
    lat = st.lattvec
    maps, rots = [], []
    for M, t, amap, shift in space_group_ops(st):
        img = sites.copy()
        atoms = sites[:, :, 0]
        img[:, :, 0] = amap[atoms]
        img[:, :, 1:] = shift[atoms] + np.einsum("ab,ksb->ksa", M, sites[:, :, 1:])
        maps.append(np.array([index[k] for k in _keys(_canon(img))], dtype=np.int64))
        rots.append(lat @ M @ np.linalg.inv(lat))
    return maps, rots


def make_ifc4(st: Structure, shell: int, seed: int, sym_iters: int = 400, space_group: bool = False) -> IFC4:
    """Seeded N(0,1) fourth-order IFCs on every quartet within the shell's
    cutoff, made symmetric under permutations of the four sites and
    acoustic-sum-rule corrected (alternating projections); with space_group
    also invariant under the crystal's space group, Phi(g.k) = R_g x4 Phi(k),
    so that rates are equal over every star."""
    rc = cluster_cutoff(st, shell)
    sites = _canon(enumerate_quartets(st, rc))
    # deterministic order: by first atom, then lexicographic on the rest
    order = np.lexsort(sites.reshape(len(sites), -1).T[::-1])
    sites = sites[order]
    n = len(sites)
    index = {k: i for i, k in enumerate(_keys(sites))}
    rng = np.random.Generator(np.random.PCG64(seed))
    phi = rng.standard_normal((n, 3, 3, 3, 3))

    perms = list(itertools.permutations(range(4)))
    pmaps = []
    for p in perms:
        ps = _canon(sites[:, list(p), :])
        pmaps.append(np.array([index[k] for k in _keys(ps)], dtype=np.int64))
    inv_axes = [tuple(int(np.argsort(p)[a]) + 1 for a in range(4)) for p in perms]

    # ASR groups: all quartets sharing sites 0..2
    g_keys = _keys(sites[:, :3, :])
    gid = {}
    groups = np.array([gid.setdefault(k, len(gid)) for k in g_keys], dtype=np.int64)
    counts = np.bincount(groups)

    def symmetrise(x):
        acc = np.zeros_like(x)
        for pm, ax in zip(pmaps, inv_axes):
            acc += np.transpose(x[pm], (0,) + ax)
        return acc / len(perms)

    def asr(x):
        s = np.zeros((len(counts), 3, 3, 3, 3))
        np.add.at(s, groups, x)
        return x - (s / counts[:, None, None, None, None])[groups]

    if space_group:
        gmaps, grots = _space_group_projector(st, sites)

        def spacegroup(x):
            # Phi(k) <- average over g of R_g^T x4 Phi(g.k)
            acc = np.zeros_like(x)
            for gm, R in zip(gmaps, grots):
                acc += np.einsum("ia,jb,kc,ld,nijkl->nabcd", R, R, R, R, x[gm], optimize=True)
            return acc / len(gmaps)
    else:
        def spacegroup(x):
            return x

    for _ in range(sym_iters):
        phi = spacegroup(symmetrise(asr(phi)))
        s = np.zeros((len(counts), 3, 3, 3, 3))
        np.add.at(s, groups, phi)
        if np.abs(s).max() < 1e-13 * max(1.0, np.abs(phi).max()):
            break

    lat = st.lattvec
    cells = sites[:, 1:, 1:]                                       # (n,3,3) cells of j,k,l
    R = [np.asfortranarray((cells[:, m, :] @ lat.T).T) for m in range(3)]
    Phi = np.asfortranarray(np.transpose(phi, (1, 2, 3, 4, 0)))   # (3,3,3,3,Ntri)
    idx = [np.ascontiguousarray(sites[:, m, 0] + 1, dtype=np.int32) for m in range(4)]
    return IFC4(n, Phi, R[0], R[1], R[2], idx[0], idx[1], idx[2], idx[3], cells)


def make_ifc3(st: Structure, shell: int, seed: int, sym_iters: int = 400) -> IFC3:
    """Third-order constants in the layout of FORCE_CONSTANTS_3RD as the
    reference holds it (Src/input.f90:102-140): Phi(3,3,3,Ntri) in eV/A^3,
    R_j, R_k in nm on lattice vectors, 1-based atoms.  Seeded N(0,1) values on
    every ordered triplet whose atoms are pairwise within the cutoff, projected
    alternately onto the index-permutation symmetric subspace and the
    acoustic-sum-rule subspace (sum over the third site's atom and cell)."""
    rc = cluster_cutoff(st, shell)
    sites = _canon(enumerate_triplets(st, rc))
    order = np.lexsort(sites.reshape(len(sites), -1).T[::-1])
    sites = sites[order]
    n = len(sites)
    index = {k: i for i, k in enumerate(_keys(sites))}
    rng = np.random.Generator(np.random.PCG64(seed))
    phi = rng.standard_normal((n, 3, 3, 3))

    perms = list(itertools.permutations(range(3)))
    pmaps = []
    for p in perms:
        ps = _canon(sites[:, list(p), :])
        pmaps.append(np.array([index[k] for k in _keys(ps)], dtype=np.int64))
    inv_axes = [tuple(int(np.argsort(p)[a]) + 1 for a in range(3)) for p in perms]

    g_keys = _keys(sites[:, :2, :])
    gid = {}
    groups = np.array([gid.setdefault(k, len(gid)) for k in g_keys], dtype=np.int64)
    counts = np.bincount(groups)

    def symmetrise(x):
        acc = np.zeros_like(x)
        for pm, ax in zip(pmaps, inv_axes):
            acc += np.transpose(x[pm], (0,) + ax)
        return acc / len(perms)

    def asr(x):
        s = np.zeros((len(counts), 3, 3, 3))
        np.add.at(s, groups, x)
        return x - (s / counts[:, None, None, None])[groups]

    for _ in range(sym_iters):
        phi = symmetrise(asr(phi))
        s = np.zeros((len(counts), 3, 3, 3))
        np.add.at(s, groups, phi)
        if np.abs(s).max() < 1e-13 * max(1.0, np.abs(phi).max()):
            break

    lat = st.lattvec
    cells = sites[:, 1:, 1:]                                       # (n,2,3) cells of j,k
    R = [np.asfortranarray((cells[:, m, :] @ lat.T).T) for m in range(2)]
    Phi = np.asfortranarray(np.transpose(phi, (1, 2, 3, 0)))      # (3,3,3,Ntri)
    idx = [np.ascontiguousarray(sites[:, m, 0] + 1, dtype=np.int32) for m in range(3)]
    return IFC3(n, Phi, R[0], R[1], idx[0], idx[1], idx[2], cells)


# ---------------------------------------------------------------------------
# harmonic part
# ---------------------------------------------------------------------------

def make_fc2(st: Structure, seed: int, k_central: float = 10.0, k_transverse: float = 2.5):
    """Nearest-neighbour central + transverse springs, returns [(I, J, cell, 3x3)] in eV/A^2."""
    rng = np.random.Generator(np.random.PCG64(seed))
    npair = st.masses.size * (st.masses.size + 1) // 2
    jit_c = rng.uniform(0.95, 1.05, size=npair)
    jit_t = rng.uniform(0.95, 1.05, size=npair)

    def pair_id(a, b):
        a, b = min(a, b) - 1, max(a, b) - 1
        return a * st.masses.size - a * (a - 1) // 2 + (b - a)

    d1 = shell_distances(st, 1)[0]
    rc = 1.05 * d1
    blocks = []
    for I, (nsites, nxyz) in enumerate(_neighbours(st, rc)):
        onsite = np.zeros((3, 3))
        x0 = st.cart(st.positions[I])
        for s, x in zip(nsites, nxyz):
            dv = x - x0
            r = np.linalg.norm(dv)
            if r < 1e-9:
                continue
            nvec = dv / r
            pid = pair_id(st.types[I], st.types[s[0]])
            kc, kt = k_central * jit_c[pid], k_transverse * jit_t[pid]
            blk = -(kc * np.outer(nvec, nvec) + kt * (np.eye(3) - np.outer(nvec, nvec)))
            blocks.append((I, int(s[0]), s[1:].copy(), blk))
            onsite -= blk
        blocks.append((I, I, np.zeros(3, np.int64), onsite))
    return blocks


def grid_ijk(ngrid) -> np.ndarray:
    n1, n2, n3 = ngrid
    idx = np.arange(n1 * n2 * n3)
    ijk = np.empty((3, idx.size), dtype=np.int32, order="F")
    ijk[0] = idx % n1
    ijk[1] = (idx // n1) % n2
    ijk[2] = idx // (n1 * n2)
    return ijk


def grid_index0(ijk, ngrid):
    """0-based grid index of (possibly unreduced) integer coordinates."""
    n = np.asarray(ngrid).reshape(3, *([1] * (np.ndim(ijk) - 1)))
    m = np.mod(ijk, n)
    return m[0] + n[0] * (m[1] + n[1] * m[2])


def harmonic(st: Structure, fc2, ngrid):
    ngrid = np.asarray(ngrid, dtype=np.int32)
    nptk = int(np.prod(ngrid))
    nb = 3 * st.natoms
    ijk = grid_ijk(ngrid)
    qfrac = ijk.T / ngrid                                        # (nptk,3)
    D = np.zeros((nptk, nb, nb), dtype=np.complex128)
    dD = np.zeros((3, nptk, nb, nb), dtype=np.complex128)
    lat = st.lattvec
    for I, J, cell, blk in fc2:
        ph = np.exp(2j * np.pi * (qfrac @ cell))                # e^{+i q.R}
        w = UNITFACTOR / np.sqrt(st.masses[st.types[I] - 1] * st.masses[st.types[J] - 1])
        R = lat @ cell                                          # nm
        D[:, 3 * I:3 * I + 3, 3 * J:3 * J + 3] += w * ph[:, None, None] * blk
        for a in range(3):
            dD[a, :, 3 * I:3 * I + 3, 3 * J:3 * J + 3] += 1j * R[a] * w * ph[:, None, None] * blk
    D = 0.5 * (D + np.conj(np.transpose(D, (0, 2, 1))))
    w2, vecs = np.linalg.eigh(D)
    omega = np.sign(w2) * np.sqrt(np.abs(w2))
    vel = np.zeros((nptk, nb, 3))
    for a in range(3):
        num = np.einsum("qcb,qcd,qdb->qb", np.conj(vecs), dD[a], vecs).real
        with np.errstate(divide="ignore", invalid="ignore"):
            vel[:, :, a] = np.where(omega > 0, num / (2.0 * np.where(omega > 0, omega, 1.0)), 0.0)
    eig = np.transpose(vecs, (0, 2, 1))                          # [q, branch, comp]
    return omega, vel, eig


# ---------------------------------------------------------------------------
# symmetry
# ---------------------------------------------------------------------------

def point_group(st: Structure, tol: float = 1e-6):
    """Space-group rotations (fractional, real space) of the crystal."""
    lat = st.lattvec
    G = lat.T @ lat
    out = []
    for ent in itertools.product((-1, 0, 1), repeat=9):
        M = np.array(ent, dtype=np.int64).reshape(3, 3)
        if abs(round(np.linalg.det(M))) != 1:
            continue
        if not np.allclose(M.T @ G @ M, G, atol=tol * np.abs(G).max()):
            continue
        # look for a translation mapping the basis onto itself
        mp = st.positions @ M.T
        found = False
        for J in range(st.natoms):
            if st.types[J] != st.types[0]:
                continue
            t = st.positions[J] - mp[0]
            ok = True
            for I in range(st.natoms):
                x = mp[I] + t
                d = st.positions - x
                d -= np.round(d)
                hit = np.nonzero((np.linalg.norm(d @ lat.T, axis=1) < 1e-6) & (st.types == st.types[I]))[0]
                if hit.size == 0:
                    ok = False
                    break
            if ok:
                found = True
                break
        if found:
            out.append(M)
    # identity first, then a fixed lexicographic order
    out.sort(key=lambda m: (not np.array_equal(m, np.eye(3, dtype=np.int64)), tuple(m.ravel())))
    return out


def symmetry_tables(st: Structure, ngrid):
    ngrid = np.asarray(ngrid, dtype=np.int64)
    nptk = int(np.prod(ngrid))
    rots = point_group(st)
    lat = st.lattvec
    crot = [lat @ M @ np.linalg.inv(lat) for M in rots]
    qrot = [np.round(np.linalg.inv(M).T).astype(np.int64) for M in rots]
    # time reversal doubles the list; duplicates (and grid-incompatible ones) are dropped
    crot = crot + [-c for c in crot]
    qrot = qrot + [-q for q in qrot]
    ijk = grid_ijk(ngrid).astype(np.int64)
    keep_c, keep_q = [], []
    for c, q in zip(crot, qrot):
        if any(np.array_equal(q, k) for k in keep_q):
            continue
        img = q @ (ijk / ngrid[:, None]) * ngrid[:, None]
        if np.abs(img - np.round(img)).max() > 1e-6:
            continue
        keep_c.append(c)
        keep_q.append(q)
    ns = len(keep_q)
    ID = np.empty((ns, nptk), dtype=np.int32, order="F")
    fixes = np.zeros((ns, nptk), dtype=bool)
    for s, q in enumerate(keep_q):
        img = np.round(q @ (ijk / ngrid[:, None]) * ngrid[:, None]).astype(np.int64)
        ID[s] = grid_index0(img, ngrid) + 1
        fixes[s] = np.all(img == ijk, axis=0)
    # wedge, in the reference's first-seen order
    seen = np.zeros(nptk + 1, dtype=bool)
    List, Nequi = [], []
    ALLEq = np.zeros((ns, nptk), dtype=np.int32, order="F")
    TypeS = np.zeros((ns, nptk), dtype=np.int32, order="F")
    for ii in range(1, nptk + 1):
        if seen[ii]:
            continue
        ll = len(List)
        List.append(ii)
        eq = []
        for s in range(ns):
            t = int(ID[s, ii - 1])
            if t not in eq:
                eq.append(t)
                seen[t] = True
                ALLEq[len(eq) - 1, ll] = t
                TypeS[len(eq) - 1, ll] = s + 1
        Nequi.append(len(eq))
    symm = np.zeros((3, 3, nptk), order="F")
    for i in range(nptk):
        sel = np.nonzero(fixes[:, i])[0]
        symm[:, :, i] = sum(keep_c[s] for s in sel) / max(1, len(sel))
    C = np.asfortranarray(np.stack(keep_c, axis=2))
    Q = np.asfortranarray(np.stack(keep_q, axis=2).astype(np.float64))
    return (np.array(List, np.int32), np.array(Nequi, np.int32), ALLEq, TypeS, ns, C, Q, symm, ID)


def make_modes(st: Structure, fc2, ngrid, symmetric: bool = True) -> Modes:
    ngrid = np.asarray(ngrid, dtype=np.int32)
    omega, vel, eig = harmonic(st, fc2, ngrid)
    nptk = int(np.prod(ngrid))
    if symmetric:
        List, Nequi, ALLEq, TypeS, ns, C, Q, symm, ID = symmetry_tables(st, ngrid)
        # velocities take the symmetry of their q point (Src/phonon_routines.f90:65-68)
        vel = np.einsum("abq,qnb->qna", symm, vel)
    else:
        List = np.arange(1, nptk + 1, dtype=np.int32)
        Nequi = np.ones(nptk, np.int32)
        ns = 1
        ALLEq = np.asfortranarray(np.arange(1, nptk + 1, dtype=np.int32)[None, :])
        TypeS = np.ones((1, nptk), np.int32, order="F")
        C = np.asfortranarray(np.eye(3)[:, :, None])
        Q = C.copy()
        symm = np.asfortranarray(np.repeat(np.eye(3)[:, :, None], nptk, axis=2))
        ID = np.asfortranarray(np.arange(1, nptk + 1, dtype=np.int32)[None, :])
    # acoustic branches at Gamma are exactly zero (Src/phonon_routines.f90:79-80)
    omega[0, :3] = 0.0
    vel[0, :3, :] = 0.0
    return Modes(ngrid=ngrid, energy=np.asfortranarray(omega), velocity=np.asfortranarray(vel),
                 eigenvect=np.asfortranarray(eig), IJK=grid_ijk(ngrid), Nlist=len(List),
                 List=List, Nequi=Nequi, ALLEquiList=ALLEq, TypeofSymmetry=TypeS,
                 nsymm_rot=ns, crotations=C, qrotations=Q, symmetrizers=symm, ID_equi=ID)


# ---------------------------------------------------------------------------
# named workloads
# ---------------------------------------------------------------------------

def config_case(k: int, ngrid=None, shell=None, T: float = 300.0, scalebroad: float = 1.0,
                symmetric: bool = True) -> Case:
    """Workloads 1..3 of the benchmark description, with optional smaller meshes.

    1: diamond a=5.43 A, m=28.085, 8x8x8, clusters to the 1st shell.
    2: same crystal at 16x16x16, clusters to the 2nd shell.
    3: wurtzite-like a=3.2 A c=5.2 A, m=26.98/14.01, 12x12x8, 2nd shell.
    5: diamond at 20x20x20, 1st shell (iterative BTE workload).
    """
    if k in (1, 2, 5):
        st = diamond()
        default_grid = {1: (8, 8, 8), 2: (16, 16, 16), 5: (20, 20, 20)}[k]
        default_shell = {1: 1, 2: 2, 5: 1}[k]
    elif k == 3:
        st = wurtzite()
        default_grid, default_shell = (12, 12, 8), 2
    else:
        raise ValueError(f"unknown workload {k}")
    ngrid = tuple(default_grid if ngrid is None else ngrid)
    shell = default_shell if shell is None else shell
    s = SEED_BASE + 10 * k
    fc2 = make_fc2(st, seed=s + 2)
    modes = make_modes(st, fc2, ngrid, symmetric=symmetric)
    ifc4 = make_ifc4(st, shell, seed=s + 4)
    ifc3 = make_ifc3(st, shell, seed=s + 3)
    return Case(st, modes, ifc4, T=T, scalebroad=scalebroad, ifc3=ifc3,
                meta={"workload": k, "ngrid": list(ngrid), "shell": shell, "structure": st.name})


# the batch of 24 small crystals: kinds and cell sizes cycle so that every
# (kind, natoms) pair appears twice; the mesh range shrinks with the cell so
# that no crystal dominates the batch (candidates grow as nptk^2 Nb^4)
BATCH_SIZE = 24
BATCH_KINDS = ("fcc", "zincblende", "rocksalt")
BATCH_NATOMS = (2, 4, 6, 8)
BATCH_MESH = {2: (10, 14), 4: (8, 12), 6: (6, 10), 8: (6, 8)}


def batch_crystal(c: int) -> dict:
    """Draw crystal ``c`` (0..23) of the batch: kind, natoms, masses, a and mesh.

    Structure draw from PCG64(SEED_BASE + 1000 + c): masses U(10, 120) amu for
    the two sublattices, a U(3.5, 6.0) A, mesh edge uniform in the natoms range
    of ``BATCH_MESH`` (6 to 14 overall).
    """
    if not 0 <= c < BATCH_SIZE:
        raise ValueError(f"batch crystal index {c} outside 0..{BATCH_SIZE - 1}")
    rng = np.random.Generator(np.random.PCG64(SEED_BASE + 1000 + c))
    natoms = BATCH_NATOMS[c % 4]
    kind = BATCH_KINDS[(c // 4) % 3]
    masses = rng.uniform(10.0, 120.0, size=2)
    a_ang = float(rng.uniform(3.5, 6.0))
    lo, hi = BATCH_MESH[natoms]
    n = int(rng.integers(lo, hi + 1))
    return {"kind": kind, "natoms": natoms, "masses": masses, "a_ang": a_ang, "mesh": n}


def batch_case(c: int, ngrid=None, shell: int = 1, T: float = 300.0, scalebroad: float = 1.0,
               symmetric: bool = True) -> Case:
    """Crystal ``c`` of the 24-crystal batch as a hot-path case.

    FC2 jitter from seed SEED_BASE + 1000 + c + 100 and FC4 from + 300 (+ 200
    is the third-order seed); clusters to the 1st shell unless ``shell`` says
    otherwise; ``ngrid`` overrides the drawn mesh (tests run smaller meshes).
    """
    d = batch_crystal(c)
    st = cubic_family(d["kind"], d["natoms"], d["a_ang"], d["masses"])
    ngrid = tuple((d["mesh"],) * 3 if ngrid is None else ngrid)
    s = SEED_BASE + 1000 + c
    fc2 = make_fc2(st, seed=s + 100)
    modes = make_modes(st, fc2, ngrid, symmetric=symmetric)
    ifc4 = make_ifc4(st, shell, seed=s + 300)
    ifc3 = make_ifc3(st, shell, seed=s + 200)
    return Case(st, modes, ifc4, T=T, scalebroad=scalebroad, ifc3=ifc3,
                meta={"workload": 4, "crystal": c, "ngrid": list(ngrid), "shell": shell,
                      "structure": st.name, "a_ang": d["a_ang"], "masses": list(map(float, d["masses"]))})
