"""On-disk formats of the reference around the four-phonon path: the CONTROL
namelist reader (Src/config.f90:60-230) and the BTE.* writers of
Src/ShengBTE.f90 (counts, phase space, weighted phase space, rates, kappa),
producing the same text as the reference's Fortran edit descriptors, so that
outputs diff against a reference run line by line.

Host-side Python: these files are read or written once per run and are not
on the device path.
"""
from __future__ import annotations

import dataclasses
import os
import re

import numpy as np


# ------------------------------------------------------------ edit descriptors
def fortran_e(x: float, w: int, d: int) -> str:
    """Fortran Ew.d: 0.ddddE+xx right-justified in w columns (E dropped for
    three-digit exponents, as the standard prescribes)."""
    x = float(x)
    if x == 0.0:
        body = "0." + "0" * d + "E+00"
    else:
        m, e = f"{abs(x):.{d - 1}e}".split("e")
        digits = m.replace(".", "")
        exp = int(e) + 1
        sign_e = "+" if exp >= 0 else "-"
        body = "0." + digits + (f"E{sign_e}{abs(exp):02d}" if abs(exp) <= 99 else f"{sign_e}{abs(exp):03d}")
        if x < 0:
            body = "-" + body
    if len(body) > w:
        return "*" * w
    return body.rjust(w)


def fortran_f(x: float, w: int, d: int) -> str:
    s = f"{float(x):.{d}f}"
    return "*" * w if len(s) > w else s.rjust(w)


def fortran_i(n: int, w: int) -> str:
    s = str(int(n))
    return "*" * w if len(s) > w else s.rjust(w)


def list_directed_real8(x: float) -> str:
    """List-directed output of one REAL(8) as the reference's example files
    hold it (`write(1,*) value`): a blank, the sign position, 16 significant
    digits and a three-digit exponent, e.g. '  2.837012589647658E-003'."""
    x = float(x)
    m, e = f"{abs(x):.15E}".split("E")
    exp = int(e)
    return " " + ("-" if x < 0 else " ") + f"{m}E{'+' if exp >= 0 else '-'}{abs(exp):03d}"


def _lines(path, rows):
    with open(path, "w") as f:
        for r in rows:
            f.write(r + "\n")


# ---------------------------------------------------------------- BTE writers
def write_numprocess_4ph(path, N_pp, N_pm, N_mm, nlist, nbands):
    """BTE.Numprocess_4ph (Src/ShengBTE.f90:501-507): branch-outer, point-inner,
    '(1X,I0,1X,I0,1X,I0)'; N_* indexed ll+Nlist*(i-1) as the reference does."""
    rows = []
    for i in range(nbands):
        for ll in range(nlist):
            k = ll + nlist * i
            rows.append(f" {int(N_pp[k])} {int(N_pm[k])} {int(N_mm[k])}")
    _lines(path, rows)


def write_per_point(path, values):
    """One line of Nb values E20.10 per irreducible point: BTE.omega,
    BTE.P3*, BTE.P4* (Src/ShengBTE.f90:305-308, 511-526, 543-563).  values is
    (Nb, Nlist)."""
    v = np.asarray(values, float)
    _lines(path, ["".join(fortran_e(x, 20, 10) for x in v[:, ll]) for ll in range(v.shape[1])])


def write_list_directed(path, value):
    """BTE.P3_total and the other one-number totals (`write(1,*)`)."""
    _lines(path, [list_directed_real8(value)])


def write_energy_value(path, energy_irr, values, w, d):
    """Branch-outer, point-inner pairs (omega, value): BTE.WP3*, BTE.WP4*
    (Ew.d = E14.5, Src/ShengBTE.f90:623-676) and BTE.w_3ph*, BTE.w_4ph*
    (E20.10, :1091-1120).  energy_irr and values are (Nb, Nlist)."""
    e, v = np.asarray(energy_irr, float), np.asarray(values, float)
    rows = []
    for i in range(e.shape[0]):
        for ll in range(e.shape[1]):
            rows.append(fortran_e(e[i, ll], w, d) + fortran_e(v[i, ll], w, d))
    _lines(path, rows)


def write_velocities(path, velocity, points):
    """BTE.v / BTE.v_full (Src/ShengBTE.f90:289-301): branch-outer, the given
    grid points (0-based) inner, 3E20.10.  velocity is (nptk, Nb, 3)."""
    v = np.asarray(velocity, float)
    rows = []
    for i in range(v.shape[1]):
        for q in points:
            rows.append("".join(fortran_e(x, 20, 10) for x in v[q, i, :]))
    _lines(path, rows)


def write_qpoints(path, List, Nequi, q_reduced):
    """BTE.qpoints (Src/ShengBTE.f90:228-231): '(I9,x,I9,x,I9,x,3(E20.10,x))'
    with q0_reduced of each irreducible point ((nptk, 3) array)."""
    rows = []
    for ll in range(len(List)):
        q = q_reduced[List[ll] - 1]
        # the x after the last value is not reached: output ends with the list
        rows.append(f"{fortran_i(ll + 1, 9)} {fortran_i(List[ll], 9)} {fortran_i(Nequi[ll], 9)} "
                    + " ".join(fortran_e(x, 20, 10) for x in q))
    _lines(path, rows)


class KappaWriter:
    """BTE.kappa, BTE.kappa_tensor and BTE.kappa_scalar of one temperature
    (Src/ShengBTE.f90:1254-1303): iteration 0 is written with the formats of
    :1262-1265, later iterations with those of :1297-1303.  ThConductivity is
    (Nb, 3, 3), already symmetrised."""

    def __init__(self, directory):
        self.f = [open(os.path.join(directory, n), "w") for n in ("BTE.kappa", "BTE.kappa_tensor",
                                                                      "BTE.kappa_scalar")]

    def write(self, it, kappa):
        k = np.asarray(kappa, float)
        tot = k.sum(axis=0)
        flat = k.ravel(order="F")
        self.f[0].write(fortran_i(it, 9) + "".join(fortran_e(x, 20, 10) for x in flat) + "\n")
        self.f[1].write(fortran_i(it, 9) + "".join(fortran_e(x, 20, 10) for x in tot.ravel(order="F")) + "\n")
        self.f[2].write(fortran_i(it, 9) + fortran_e(np.trace(tot) / 3.0, 20, 10) + "\n")
        for f in self.f:
            f.flush()

    def close(self):
        for f in self.f:
            f.close()


def kappa_vs_T_line(T, kappa_total, niter=None):
    """One line of BTE.KappaTensorVsT_RTA ('(F7.1,9E14.5)') or, with niter,
    BTE.KappaTensorVsT_CONV ('(F7.1,9E14.5,I6)'), Src/ShengBTE.f90:1266, 1308."""
    s = fortran_f(T, 7, 1) + "".join(fortran_e(x, 14, 5) for x in np.asarray(kappa_total, float).ravel(order="F"))
    return s + (fortran_i(niter, 6) if niter is not None else "")


def read_table(path):
    """Numbers of a BTE.* text file, one list per line (Fortran E output,
    including the E-less three-digit exponent form)."""
    out = []
    with open(path) as f:
        for line in f:
            toks = line.split()
            if toks:
                out.append([_fnum(t) for t in toks])
    return out


def _fnum(t):
    t = t.replace("D", "E").replace("d", "e")
    m = re.fullmatch(r"([-+]?\d*\.\d*)([-+]\d{3})", t)
    if m:
        t = m.group(1) + "E" + m.group(2)
    return float(t) if any(c in t for c in ".eE") else int(t)


# --------------------------------------------------------------- CONTROL file
@dataclasses.dataclass
class Control:
    """The CONTROL namelists (&allocations, &crystal, &parameters, &flags) with
    the reference's defaults (Src/config.f90:68-167).  A default the
    reference assigns from a default-REAL literal (eps = 1e-5) keeps its
    float32 value; a value read from the file is a REAL(8) read."""
    nelements: int = 0
    natoms: int = 0
    ngrid: tuple = (0, 0, 0)
    norientations: int = 0
    lfactor: float = 1.0
    lattvec: np.ndarray = None          # (3, 3), columns a_i, before lfactor
    epsilon: np.ndarray = None
    elements: list = None
    types: np.ndarray = None
    positions: np.ndarray = None        # (3, natoms), lattice coordinates
    masses: np.ndarray = None
    gfactors: np.ndarray = None
    born: np.ndarray = None
    scell: tuple = (-1, -1, -1)
    orientations: np.ndarray = None
    T: float = 0.0
    T_min: float = 0.0
    T_max: float = 0.0
    T_step: float = 0.0
    scalebroad: float = 1.0
    rmin: float = 5.0
    rmax: float = 505.0
    dr: float = 100.0
    maxiter: int = 1000
    nticks: int = 100
    eps: float = float(np.float32(1e-5))
    omega_max: float = 1e100
    Length: float = 1.0
    Ewald: float = -1.0
    num_sample_process_3ph: int = -1
    num_sample_process_3ph_phase_space: int = -1
    num_sample_process_4ph: int = -1
    num_sample_process_4ph_phase_space: int = -1
    nonanalytic: bool = True
    convergence: bool = True
    isotopes: bool = True
    autoisotopes: bool = True
    nanowires: bool = False
    onlyharmonic: bool = False
    espresso: bool = False
    tdep: bool = False
    four_phonon: bool = False
    four_phonon_iteration: bool = False
    nanolength: bool = False

    @property
    def nbands(self):
        return 3 * self.natoms

    @property
    def nptk(self):
        return int(np.prod(self.ngrid))

    @property
    def lattvec_nm(self):
        """lattvec*lfactor (Src/config.f90:236), nm."""
        return self.lfactor * self.lattvec


_INT_KEYS = {"nelements", "natoms", "norientations", "maxiter", "nticks", "num_sample_process_3ph",
             "num_sample_process_3ph_phase_space", "num_sample_process_4ph", "num_sample_process_4ph_phase_space"}
_BOOL_KEYS = {"nonanalytic", "convergence", "isotopes", "autoisotopes", "nanowires", "onlyharmonic", "espresso",
              "tdep", "four_phonon", "four_phonon_iteration", "nanolength"}
_GROUPS = {
    "allocations": {"nelements", "natoms", "ngrid", "norientations"},
    "crystal": {"lfactor", "lattvec", "elements", "types", "positions", "masses", "gfactors", "epsilon", "born",
                "scell", "orientations"},
    "parameters": {"T", "scalebroad", "rmin", "rmax", "dr", "maxiter", "nticks", "eps", "T_min", "T_max", "T_step",
                   "omega_max", "Length", "Ewald", "num_sample_process_3ph", "num_sample_process_3ph_phase_space",
                   "num_sample_process_4ph", "num_sample_process_4ph_phase_space"},
    "flags": _BOOL_KEYS,
}


def _namelist_groups(text):
    """{group: body} of '&group ... &end' or '&group ... /' blocks; '!' comments removed."""
    text = "\n".join(_strip_comment(line) for line in text.splitlines())
    out = {}
    for m in re.finditer(r"&(\w+)(.*?)(?:&end|/)", text, re.S | re.I):
        out[m.group(1).lower()] = m.group(2)
    return out


def _strip_comment(line):
    q = None
    for i, ch in enumerate(line):
        if ch in "\"'":
            q = None if q == ch else (ch if q is None else q)
        elif ch == "!" and q is None:
            return line[:i]
    return line


def _values(s):
    """Value tokens of one assignment: quoted strings, r*value repeats,
    separated by blanks and commas."""
    toks = re.findall(r"\"[^\"]*\"|'[^']*'|[^\s,]+", s)
    out = []
    for t in toks:
        m = re.fullmatch(r"(\d+)\*(.+)", t)
        out.extend([m.group(2)] * int(m.group(1)) if m else [t])
    return out


def _scalar(tok, key):
    if tok[0] in "\"'":
        return tok[1:-1]
    if key in _BOOL_KEYS:
        t = tok.strip(".").lower()
        if t in ("t", "true"):
            return True
        if t in ("f", "false"):
            return False
        raise ValueError(f"{key}: not a logical: {tok}")
    v = float(tok.replace("d", "e").replace("D", "E"))
    return int(v) if key in _INT_KEYS or key in ("ngrid", "scell", "types", "orientations") else v


def read_control(path) -> Control:
    """Reads CONTROL as Src/config.f90:68-230 does (namelists in order, the
    reference's defaults and its consistency checks, which raise ValueError
    here where the reference prints an error)."""
    with open(path) as f:
        groups = _namelist_groups(f.read())
    c = Control()
    arrays = {}

    def shape_of(key):
        na, ne = max(c.natoms, 1), max(c.nelements, 1)
        return {"ngrid": (3,), "lattvec": (3, 3), "epsilon": (3, 3), "elements": (ne,), "types": (na,),
                "positions": (3, na), "masses": (ne,), "gfactors": (ne,), "born": (3, 3, na), "scell": (3,),
                "orientations": (3, max(c.norientations, 1))}.get(key)

    for gname in ("allocations", "crystal", "parameters", "flags"):
        if gname not in groups:
            raise ValueError(f"CONTROL: namelist &{gname} missing")
        body = groups[gname]
        pat = re.compile(r"(\w+)\s*(\(([^)]*)\))?\s*=")
        hits = list(pat.finditer(body))
        for n, m in enumerate(hits):
            key = m.group(1)
            match = [k for k in _GROUPS[gname] if k.lower() == key.lower()]
            if not match:
                raise ValueError(f"CONTROL: unknown variable {key} in &{gname}")
            key = match[0]
            end = hits[n + 1].start() if n + 1 < len(hits) else len(body)
            vals = [_scalar(t, key) for t in _values(body[m.end():end])]
            shp = shape_of(key)
            if shp is None:
                setattr(c, key, vals[0])
                continue
            if key not in arrays:
                dtype = object if key == "elements" else (int if key in ("ngrid", "scell", "types", "orientations")
                                                         else float)
                arrays[key] = np.zeros(shp, dtype=dtype, order="F")
            a = arrays[key]
            # start position of a (possibly partial) subscript, Fortran order fill
            start = [0] * len(shp)
            if m.group(3):
                subs = [s.strip() for s in m.group(3).split(",")]
                start = [0 if s == ":" else int(s) - 1 for s in subs]
            flat = np.ravel_multi_index(tuple(start), shp, order="F")
            av = a.reshape(-1, order="F")
            av[flat:flat + len(vals)] = vals
            arrays[key] = av.reshape(shp, order="F")
        if gname == "allocations":
            c.ngrid = tuple(int(x) for x in arrays.pop("ngrid", np.zeros(3, int)))
            if c.norientations < 0:
                raise ValueError("norientations must be >=0")
            if c.nelements < 1 or c.natoms < 1 or c.natoms < c.nelements:
                raise ValueError("nelements,natoms must be >0, natoms must be >=nelements")
            if any(g < 1 for g in c.ngrid):
                raise ValueError("all components of ngrid must be must be >0")
        elif gname == "crystal":
            c.lattvec = arrays.get("lattvec", np.zeros((3, 3)))
            c.epsilon = arrays.get("epsilon", np.eye(3))
            c.elements = [str(x) for x in arrays.get("elements", np.array([""] * c.nelements, object))]
            c.types = arrays.get("types", np.zeros(c.natoms, int))
            c.positions = arrays.get("positions", np.zeros((3, c.natoms)))
            c.masses = arrays.get("masses")
            c.gfactors = arrays.get("gfactors", np.zeros(c.nelements))
            c.born = arrays.get("born", np.zeros((3, 3, c.natoms)))
            c.scell = tuple(int(x) for x in arrays.get("scell", np.full(3, -1)))
            c.orientations = arrays.get("orientations")
            if not all(s > 0 for s in c.scell):
                raise ValueError("all supercell sizes must be >0")
            if not np.all(c.types != 0):
                raise ValueError("atom types must be initialized correctly")
        elif gname == "parameters":
            if c.T <= 0 and c.T_min <= 0:
                raise ValueError("T must be >0 K")
            if c.T > 0:
                c.T_min = c.T_max = c.T_step = c.T
            if c.rmin <= 0 or c.rmax <= c.rmin or c.dr <= 0:
                raise ValueError("rmin and dr must be >0, and rmax must be > rmin")
            if c.maxiter <= 0:
                raise ValueError("maxiter must be >0")
        else:
            if c.four_phonon_iteration and not c.convergence:
                raise ValueError("four_phonon_iteration=.TRUE. but convergence=.FALSE.")
            if c.four_phonon_iteration and not c.four_phonon:
                raise ValueError("four_phonon_iteration=.TRUE. but four_phonon=.FALSE.")
            if c.nanowires and c.norientations == 0:
                raise ValueError("nanowires=.TRUE. but norientations=0")
            if c.tdep and c.nonanalytic and c.Ewald == -1.0:
                raise ValueError("tdep=.TRUE. & nonanalytic=.TRUE. but Ewald parameter is not supplied")
            if not c.tdep and c.Ewald != -1.0:
                raise ValueError("tdep=.FALSE. but Ewald parameter is given")
            if not c.four_phonon and (c.num_sample_process_4ph > 0 or c.num_sample_process_4ph_phase_space > 0):
                raise ValueError("four_phonon=.false. but num_sample_process_4ph.gt.0 or "
                                 "num_sample_process_4ph_phase_space.gt.0")
            if c.convergence and c.num_sample_process_3ph > 0:
                raise ValueError("convergence=.true. but num_sample_process_3ph.gt.0")
            if c.four_phonon_iteration and c.num_sample_process_4ph > 0:
                raise ValueError("four_phonon_iteration=.true. but num_sample_process_4ph.gt.0")
    return c


# ------------------------------------------------ driver-level output blocks
def write_phase_space_3ph(directory, P_plus, P_minus, Nequi):
    """BTE.P3_plus, BTE.P3_minus, BTE.P3 = 2(P+ + P-/2)/3 per point and the
    Nequi-weighted totals (Src/ShengBTE.f90:509-541).  P_* are (Nb, Nlist)."""
    Pp, Pm = np.asarray(P_plus, float), np.asarray(P_minus, float)
    w = np.asarray(Nequi, float)[None, :]
    d = directory
    write_per_point(os.path.join(d, "BTE.P3_plus"), Pp)
    write_per_point(os.path.join(d, "BTE.P3_minus"), Pm)
    write_per_point(os.path.join(d, "BTE.P3"), 2.0 * (Pp + Pm / 2.0) / 3.0)
    tp, tm = float((Pp * w).sum()), float((Pm * w).sum())
    write_list_directed(os.path.join(d, "BTE.P3_plus_total"), tp)
    write_list_directed(os.path.join(d, "BTE.P3_minus_total"), tm)
    write_list_directed(os.path.join(d, "BTE.P3_total"), 2.0 * (tp + tm / 2.0) / 3.0)


def write_phase_space_4ph(directory, P_pp, P_pm, P_mm, Nequi):
    """BTE.P4_plusplus, BTE.P4_plusmius (the reference's file name),
    BTE.P4_minusminus, BTE.P4 = 6(P++ + P+- + P--)/7 per point and the
    Nequi-weighted totals (Src/ShengBTE.f90:542-585).  P_* are (Nb, Nlist)."""
    P = [np.asarray(x, float) for x in (P_pp, P_pm, P_mm)]
    w = np.asarray(Nequi, float)[None, :]
    d = directory
    for name, x in zip(("BTE.P4_plusplus", "BTE.P4_plusmius", "BTE.P4_minusminus"), P):
        write_per_point(os.path.join(d, name), x)
    write_per_point(os.path.join(d, "BTE.P4"), 6.0 * (P[0] + P[1] + P[2]) / 7.0)
    t = [float((x * w).sum()) for x in P]
    for name, x in zip(("BTE.P4_plusplus_total", "BTE.P4_plusminus_total", "BTE.P4_minusminus_total"), t):
        write_list_directed(os.path.join(d, name), x)
    write_list_directed(os.path.join(d, "BTE.P4_total"), 6.0 * (t[0] + t[1] + t[2]) / 7.0)


def write_weighted_phase_space(directory, energy_irr, WP3_plus, WP3_minus, WP4=None):
    """T-directory files BTE.WP3_plus, BTE.WP3_minus, BTE.WP3 (sum) and, with
    WP4 = (WP4_plusplus, WP4_plusminus, WP4_minusminus), BTE.WP4_* and BTE.WP4
    (sum), Src/ShengBTE.f90:623-676.  All arrays (Nb, Nlist)."""
    d = directory
    write_energy_value(os.path.join(d, "BTE.WP3_plus"), energy_irr, WP3_plus, 14, 5)
    write_energy_value(os.path.join(d, "BTE.WP3_minus"), energy_irr, WP3_minus, 14, 5)
    write_energy_value(os.path.join(d, "BTE.WP3"), energy_irr, np.asarray(WP3_plus) + np.asarray(WP3_minus), 14, 5)
    if WP4 is not None:
        for name, x in zip(("BTE.WP4_plusplus", "BTE.WP4_plusminus", "BTE.WP4_minusminus"), WP4):
            write_energy_value(os.path.join(d, name), energy_irr, x, 14, 5)
        write_energy_value(os.path.join(d, "BTE.WP4"), energy_irr,
                           np.asarray(WP4[0]) + np.asarray(WP4[1]) + np.asarray(WP4[2]), 14, 5)


def write_rates(directory, energy_irr, rate_3ph=None, rate_4ph=None):
    """BTE.w_3ph, BTE.w_3ph_plus, BTE.w_3ph_minus from rate_3ph = (total,
    plus, minus) and BTE.w_4ph, BTE.w_4ph_plusplus, _plusminus, _minusminus
    from rate_4ph = (total, ++, +-, --), Src/ShengBTE.f90:1089-1120."""
    d = directory
    if rate_3ph is not None:
        for name, x in zip(("BTE.w_3ph", "BTE.w_3ph_plus", "BTE.w_3ph_minus"), rate_3ph):
            write_energy_value(os.path.join(d, name), energy_irr, x, 20, 10)
    if rate_4ph is not None:
        for name, x in zip(("BTE.w_4ph", "BTE.w_4ph_plusplus", "BTE.w_4ph_plusminus", "BTE.w_4ph_minusminus"),
                           rate_4ph):
            write_energy_value(os.path.join(d, name), energy_irr, x, 20, 10)
