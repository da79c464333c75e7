"""Python mirror of the reference's four-phonon call surface over the fp4 C ABI.

The names, argument meaning and output arrays follow ``module processes``
(``Src/processes.f90``): ``NP_driver_4ph``, ``Ind_only_driver_4ph``,
``RTA_driver_4ph``, ``RTA_driver_4ph_GPU_using_Ind``, ``Ind_driver_4ph`` and
``Vp_pp``/``Vp_pm``/``Vp_mm``.  The module globals the reference reads
implicitly (nptk, Nbands, ngrid, rlattvec, masses, types, T, scalebroad,
omega_max, onlyharmonic) are given once to :class:`Context`.

Everything here runs on the GPU through ``libfp4.so``; there is no CPU
fallback.  Importing works without a GPU, but creating a context raises if
the library or a HIP device is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# FP4_PROFILE_BUILD=1 loads the profiling build (same kernels, plus the
# per-phase clock counters of the dense kernel; see build.py)
LIB_PATH = os.path.join(HERE, "libfp4prof.so" if os.environ.get("FP4_PROFILE_BUILD") == "1" else "libfp4.so")

PREC_FP64 = 0
PREC_FP64_DENSE = 1
PREC_FP32 = 2
PREC_BF16X3 = 3
PREC_FP16X3 = 4
PREC_FP16X2 = 5

ERRORS = {1: "FP4_ERR_ARG", 2: "FP4_ERR_DEVICE", 3: "FP4_ERR_NOMEM", 4: "FP4_ERR_OVERFLOW", 5: "FP4_ERR_STATE",
          6: "FP4_ERR_UNSUPPORTED"}

# every symbol include/fp4.h declares
EXPORTS = [
    "fp4_last_error", "fp4_ctx_create", "fp4_ctx_destroy", "fp4_ctx_set_temperature", "fp4_ctx_set_modes",
    "fp4_ctx_set_ifc4", "fp4_np_4ph", "fp4_ind_only_4ph", "fp4_rta_4ph", "fp4_rta_4ph_modes",
    "fp4_rta_4ph_using_ind", "fp4_ind_4ph", "fp4_vp_4ph", "fp4_np_4ph_modes", "fp4_np_4ph_modes_noalloc", "fp4_list_hash_4ph_modes", "fp4_ind_only_4ph_modes",
    "fp4_ctx_kernel_stats",
    "fp4_ctx_reset_stats", "fp4_ctx_set_batch_capacity", "fp4_ctx_dense_info", "fp4_ctx_dense_shape", "fp4_ctx_dense_kernel", "fp4_ctx_dense_candidates", "fp4_ctx_set_dense_shape", "fp4_ctx_dense_phase_profile",
    "fp4_selftest_mfma", "fp4_ctx_set_symmetry", "fp4_iteration0", "fp4_bte_set_operator", "fp4_iteration_4ph",
    "fp4_bte_iterate", "fp4_bte_set_sweep", "fp4_tconduct", "fp4_bte_solve", "fp4_ctx_set_ifc3", "fp4_np_3ph", "fp4_ind_only_3ph",
    "fp4_rta_3ph", "fp4_rta_3ph_modes", "fp4_rta_3ph_using_ind", "fp4_ind_3ph", "fp4_vp_3ph",
    "fp4_eigen_dm_espresso", "fp4_eigen_dm_phonopy", "fp4_np_4ph_sample", "fp4_rta_4ph_sample", "fp4_rta_4ph_sample_modes",
    "fp4_np_3ph_sample", "fp4_rta_3ph_sample", "fp4_rta_3ph_sample_modes", "fp4_ctx_autotune_dense", "fp4_bte_assemble_operator",
    "fp4_bte_operator_begin", "fp4_bte_operator_add", "fp4_bte_operator_end",
    "fp4_batch_create", "fp4_batch_destroy", "fp4_batch_rta_4ph",
]


class FP4Error(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Config(C.Structure):
    _fields_ = [("ngrid", C.c_int32 * 3), ("nbands", C.c_int32), ("natoms", C.c_int32), ("nelements", C.c_int32),
                ("rlattvec", C.c_double * 9), ("masses", C.c_void_p), ("types", C.c_void_p), ("T", C.c_double),
                ("scalebroad", C.c_double), ("omega_max", C.c_double), ("onlyharmonic", C.c_int32),
                ("device", C.c_int32)]


_lib = None


def load():
    """Loads libfp4.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FP4Error(2, f"{LIB_PATH} is missing: build it with fourphonon_amd/build.py")
        _lib = C.CDLL(LIB_PATH)
        _lib.fp4_last_error.restype = C.c_char_p
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _check(rc):
    if rc != 0:
        raise FP4Error(rc, load().fp4_last_error().decode())


def selftest_mfma(device: int = 0) -> float:
    """Max abs error of a 16x16x4 FP32 MFMA against the exact product."""
    err = C.c_double()
    _check(load().fp4_selftest_mfma(C.c_int32(device), C.byref(err)))
    return err.value


class Harmonic(C.Structure):
    _fields_ = [("ngrid", C.c_int32 * 3), ("natoms", C.c_int32), ("nelements", C.c_int32), ("scell", C.c_int32 * 3),
                ("lattvec", C.c_double * 9), ("positions", C.c_void_p), ("masses", C.c_void_p),
                ("types", C.c_void_p), ("device", C.c_int32)]


def eigen_dm_espresso(structure, fc, scell, ngrid, symmetrizers=None, device: int = 0):
    """Device eigenDM for espresso 2nd-order constants fc(3,3,nat,nat,s1,s2,s3)
    -> (omega (nptk,Nb), eigenvect (nptk,Nb,Nb), velocity (nptk,Nb,3)), F-order."""
    st = structure
    nat = st.natoms
    nb = 3 * nat
    nptk = int(np.prod(ngrid))
    pos = np.asfortranarray(np.asarray(st.positions, np.float64).T)          # positions(3,natoms)
    masses = np.ascontiguousarray(st.masses, np.float64)
    types = np.ascontiguousarray(st.types, np.int32)
    fcf = np.asfortranarray(fc, np.float64)
    h = Harmonic()
    h.ngrid[:] = [int(x) for x in ngrid]
    h.natoms, h.nelements = nat, len(masses)
    h.scell[:] = [int(x) for x in scell]
    h.lattvec[:] = list(np.asarray(st.lattvec, np.float64).ravel(order="F"))
    h.positions, h.masses, h.types = _p(pos), _p(masses), _p(types)
    h.device = device
    om = np.zeros((nptk, nb), order="F")
    eig = np.zeros((nptk, nb, nb), np.complex128, order="F")
    vel = np.zeros((nptk, nb, 3), order="F")
    sym = None if symmetrizers is None else np.asfortranarray(symmetrizers, np.float64)
    _check(load().fp4_eigen_dm_espresso(C.byref(h), _p(fcf), _p(sym) if sym is not None else None,
                                        _p(om), _p(eig), _p(vel)))
    return om, eig, vel


def eigen_dm_phonopy(structure, fc, scell, ngrid, born=None, epsilon=None, images=None, symmetrizers=None,
                     device: int = 0):
    """Device eigenDM for read2fc's reduced phonopy constants
    fc(nat,3,s1,s2,s3,nat,3) in (rad/ps)^2, with the dielectric correction
    when born(3,3,nat) and epsilon(3,3) are given (its direction from
    images (nptk,3) when given) -> (omega (nptk,Nb), eigenvect (nptk,Nb,Nb),
    velocity (nptk,Nb,3)), F-order."""
    if (born is None) != (epsilon is None):
        raise ValueError("born and epsilon go together")
    st = structure
    nat = st.natoms
    nb = 3 * nat
    nptk = int(np.prod(ngrid))
    pos = np.asfortranarray(np.asarray(st.positions, np.float64).T)
    masses = np.ascontiguousarray(st.masses, np.float64)
    types = np.ascontiguousarray(st.types, np.int32)
    fcf = np.asfortranarray(fc, np.float64)
    h = Harmonic()
    h.ngrid[:] = [int(x) for x in ngrid]
    h.natoms, h.nelements = nat, len(masses)
    h.scell[:] = [int(x) for x in scell]
    h.lattvec[:] = list(np.asarray(st.lattvec, np.float64).ravel(order="F"))
    h.positions, h.masses, h.types = _p(pos), _p(masses), _p(types)
    h.device = device
    om = np.zeros((nptk, nb), order="F")
    eig = np.zeros((nptk, nb, nb), np.complex128, order="F")
    vel = np.zeros((nptk, nb, 3), order="F")
    bz = None if born is None else np.asfortranarray(born, np.float64)
    ep = None if epsilon is None else np.asfortranarray(epsilon, np.float64)
    qi = None if images is None else np.ascontiguousarray(images, np.float64).reshape(nptk, 3)
    sym = None if symmetrizers is None else np.asfortranarray(symmetrizers, np.float64)
    opt = [None if x is None else _p(x) for x in (bz, ep, qi, sym)]
    _check(load().fp4_eigen_dm_phonopy(C.byref(h), _p(fcf), *opt, _p(om), _p(eig), _p(vel)))
    return om, eig, vel


class Context:
    """One device context holding a crystal's modes and fourth-order IFCs."""

    def __init__(self, case, device: int = 0):
        L = load()
        st, m, f = case.structure, case.modes, case.ifc4
        self.case = case
        self.nptk, self.nb, self.nlist = m.nptk, m.nbands, m.Nlist
        self._masses = np.ascontiguousarray(st.masses, np.float64)
        self._types = np.ascontiguousarray(st.types, np.int32)
        cfg = Config()
        cfg.ngrid[:] = [int(x) for x in m.ngrid]
        cfg.nbands, cfg.natoms, cfg.nelements = m.nbands, st.natoms, len(st.masses)
        cfg.rlattvec[:] = list(np.asarray(st.rlattvec).ravel(order="F"))
        cfg.masses, cfg.types = _p(self._masses), _p(self._types)
        cfg.T, cfg.scalebroad, cfg.omega_max = case.T, case.scalebroad, case.omega_max
        cfg.onlyharmonic = int(case.onlyharmonic)
        cfg.device = device
        h = C.c_void_p()
        _check(L.fp4_ctx_create(C.byref(cfg), C.byref(h)))
        self.h = h
        _check(L.fp4_ctx_set_modes(h, _p(np.asfortranarray(m.energy, np.float64)),
                                   _p(np.asfortranarray(m.velocity, np.float64)),
                                   _p(np.asfortranarray(m.eigenvect, np.complex128)),
                                   _p(np.asfortranarray(m.IJK, np.int32)), C.c_int32(m.Nlist),
                                   _p(np.ascontiguousarray(m.List, np.int32))))
        _check(L.fp4_ctx_set_ifc4(h, C.c_int32(f.Ntri), _p(np.asfortranarray(f.Phi, np.float64)),
                                  _p(np.asfortranarray(f.R_j, np.float64)), _p(np.asfortranarray(f.R_k, np.float64)),
                                  _p(np.asfortranarray(f.R_l, np.float64)),
                                  *[_p(np.ascontiguousarray(x, np.int32))
                                    for x in (f.Index_i, f.Index_j, f.Index_k, f.Index_l)]))
        f3 = getattr(case, "ifc3", None)
        if f3 is not None:
            _check(L.fp4_ctx_set_ifc3(h, C.c_int32(f3.Ntri), _p(np.asfortranarray(f3.Phi, np.float64)),
                                      _p(np.asfortranarray(f3.R_j, np.float64)),
                                      _p(np.asfortranarray(f3.R_k, np.float64)),
                                      *[_p(np.ascontiguousarray(x, np.int32))
                                        for x in (f3.Index_i, f3.Index_j, f3.Index_k)]))
        self.set_symmetry()

    def close(self):
        if getattr(self, "h", None):
            load().fp4_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def nmodes(self):
        return self.nb * self.nlist

    def set_temperature(self, T):
        _check(load().fp4_ctx_set_temperature(self.h, C.c_double(T)))

    # ------------------------------------------------------------ drivers
    def NP_driver_4ph(self):
        """-> (N_pp, P_pp, N_pm, P_pm, N_mm, P_mm); N int64 (Nlist*Nb), P (Nb,Nlist)."""
        nm = self.nmodes
        N = [np.zeros(nm, np.int64) for _ in range(3)]
        P = [np.zeros((self.nb, self.nlist), order="F") for _ in range(3)]
        _check(load().fp4_np_4ph(self.h, _p(N[0]), _p(P[0]), _p(N[1]), _p(P[1]), _p(N[2]), _p(P[2])))
        return N[0], P[0], N[1], P[1], N[2], P[2]

    def NP_driver_4ph_modes(self, modes):
        """Counts and phase space of a mode subset -> (N_pp, P_pp, N_pm, P_pm, N_mm, P_mm), each len(modes)."""
        modes = np.ascontiguousarray(np.sort(np.asarray(modes)), np.int32)
        N = [np.zeros(len(modes), np.int64) for _ in range(3)]
        P = [np.zeros(len(modes)) for _ in range(3)]
        _check(load().fp4_np_4ph_modes(self.h, C.c_int32(len(modes)), _p(modes), _p(N[0]), _p(P[0]), _p(N[1]),
                                       _p(P[1]), _p(N[2]), _p(P[2])))
        return N[0], P[0], N[1], P[1], N[2], P[2]

    def NP_driver_4ph_modes_noalloc(self, modes):
        """NP_driver_4ph_modes through the allocation-free entry point (fixed per-GPU scratch array)."""
        modes = np.ascontiguousarray(np.sort(np.asarray(modes)), np.int32)
        N = [np.zeros(len(modes), np.int64) for _ in range(3)]
        P = [np.zeros(len(modes)) for _ in range(3)]
        _check(load().fp4_np_4ph_modes_noalloc(self.h, C.c_int32(len(modes)), _p(modes), _p(N[0]), _p(P[0]),
                                               _p(N[1]), _p(P[1]), _p(N[2]), _p(P[2])))
        return N[0], P[0], N[1], P[1], N[2], P[2]

    def list_hash_4ph_modes(self, modes):
        """Counts and order-sensitive list hashes of a mode subset, computed on the device without
        materialising the lists -> (N[3, nmodes] int64, H[3, nmodes] uint64)."""
        modes = np.ascontiguousarray(np.sort(np.asarray(modes)), np.int32)
        N = np.zeros((3, len(modes)), np.int64)
        H = np.zeros((3, len(modes)), np.uint64)
        _check(load().fp4_list_hash_4ph_modes(self.h, C.c_int32(len(modes)), _p(modes), _p(N[0]), _p(N[1]),
                                              _p(N[2]), _p(H[0]), _p(H[1]), _p(H[2])))
        return N, H

    def Ind_only_driver_4ph(self, N_pp, N_pm, N_mm):
        """-> [[Indof2nd, Indof3rd, Indof4th] for ++, +-, --] (int64, reference encoding)."""
        Ns = [np.ascontiguousarray(n, np.int64) for n in (N_pp, N_pm, N_mm)]
        outs = [[np.zeros(max(int(n.sum()), 1), np.int64) for _ in range(3)] for n in Ns]
        _check(load().fp4_ind_only_4ph(self.h, *[_p(n) for n in Ns], *[_p(o) for ch in outs for o in ch]))
        return [[o[:int(n.sum())] for o in ch] for ch, n in zip(outs, Ns)]

    def Ind_only_driver_4ph_modes(self, modes, cap=None):
        """Lists of a mode subset -> (N_pp, N_pm, N_mm, lists)."""
        modes = np.ascontiguousarray(np.sort(np.asarray(modes)), np.int32)
        if cap is None:
            N = self.NP_driver_4ph()
            cap = max(int(N[2 * ch][modes - 1].sum()) for ch in range(3))
        cap = max(int(cap), 1)
        Ns = [np.zeros(len(modes), np.int64) for _ in range(3)]
        outs = [[np.zeros(cap, np.int64) for _ in range(3)] for _ in range(3)]
        _check(load().fp4_ind_only_4ph_modes(self.h, C.c_int32(len(modes)), _p(modes), C.c_int64(cap),
                                             *[_p(n) for n in Ns], *[_p(o) for ch in outs for o in ch]))
        return Ns[0], Ns[1], Ns[2], [[o[:int(n.sum())] for o in ch] for ch, n in zip(outs, Ns)]

    def _rate_outputs(self):
        return [np.zeros((self.nb, self.nlist), order="F") for _ in range(7)]

    @staticmethod
    def _named(out):
        return dict(zip(["rate_scatt_4ph", "rate_scatt_plusplus", "rate_scatt_plusminus", "rate_scatt_minusminus",
                         "WP4_plusplus", "WP4_plusminus", "WP4_minusminus"], out))

    def RTA_driver_4ph(self, precision=PREC_FP64):
        out = self._rate_outputs()
        _check(load().fp4_rta_4ph(self.h, C.c_int32(precision), *[_p(o) for o in out]))
        return self._named(out)

    def RTA_driver_4ph_modes(self, modes, precision=PREC_FP64, ii_lo=1, ii_hi=None):
        """RTA_driver_4ph over a mode subset (mm, 1-based) and a q' window; returns (dict, nproc)."""
        out = self._rate_outputs()
        modes = np.ascontiguousarray(np.sort(np.asarray(modes)), np.int32)
        nproc = C.c_int64()
        _check(load().fp4_rta_4ph_modes(self.h, C.c_int32(precision), C.c_int32(len(modes)), _p(modes),
                                        C.c_int32(ii_lo), C.c_int32(self.nptk if ii_hi is None else ii_hi),
                                        *[_p(o) for o in out], C.byref(nproc)))
        return self._named(out), nproc.value

    # ------------------------------------------------ sampling estimators
    def NP_driver_4ph_sample(self, nsample, seed):
        """Sampled NP_driver_4ph -> (N_pp, P_pp, N_pm, P_pm, N_mm, P_mm) estimates."""
        nm = self.nmodes
        N = [np.zeros(nm, np.int64) for _ in range(3)]
        P = [np.zeros((self.nb, self.nlist), order="F") for _ in range(3)]
        _check(load().fp4_np_4ph_sample(self.h, C.c_int32(nsample), C.c_uint64(seed), _p(N[0]), _p(P[0]), _p(N[1]),
                                        _p(P[1]), _p(N[2]), _p(P[2])))
        return N[0], P[0], N[1], P[1], N[2], P[2]

    def RTA_driver_4ph_sample(self, nsample, seed, modes=None):
        """Sampled RTA_driver_4ph (FP64 per-process form) -> (rates dict, sampled processes evaluated)."""
        out = self._rate_outputs()
        if modes is None:
            _check(load().fp4_rta_4ph_sample(self.h, C.c_int32(nsample), C.c_uint64(seed), *[_p(o) for o in out]))
            return self._named(out), None
        modes = np.ascontiguousarray(np.sort(np.asarray(modes)), np.int32)
        n = C.c_int64()
        _check(load().fp4_rta_4ph_sample_modes(self.h, C.c_int32(nsample), C.c_uint64(seed), C.c_int32(len(modes)),
                                               _p(modes), *[_p(o) for o in out], C.byref(n)))
        return self._named(out), n.value

    def RTA_driver_4ph_GPU_using_Ind(self, N_pp, N_pm, N_mm, lists):
        Ns = [np.ascontiguousarray(n, np.int64) for n in (N_pp, N_pm, N_mm)]
        L3 = [[np.ascontiguousarray(x, np.int64) if len(x) else np.zeros(1, np.int64) for x in ch] for ch in lists]
        out = self._rate_outputs()
        _check(load().fp4_rta_4ph_using_ind(self.h, *[_p(n) for n in Ns], *[_p(x) for ch in L3 for x in ch],
                                            *[_p(o) for o in out]))
        return self._named(out)

    def Ind_driver_4ph(self, N_pp, N_pm, N_mm):
        """-> (lists, [Gamma_pp, Gamma_pm, Gamma_mm], rates dict)."""
        Ns = [np.ascontiguousarray(n, np.int64) for n in (N_pp, N_pm, N_mm)]
        tot = [int(n.sum()) for n in Ns]
        lists = [[np.zeros(max(t, 1), np.int64) for _ in range(3)] for t in tot]
        gam = [np.zeros(max(t, 1)) for t in tot]
        out = self._rate_outputs()
        _check(load().fp4_ind_4ph(self.h, *[_p(n) for n in Ns], *[_p(x) for ch in lists for x in ch],
                                  *[_p(g) for g in gam], *[_p(o) for o in out]))
        return ([[x[:t] for x in ch] for ch, t in zip(lists, tot)], [g[:t] for g, t in zip(gam, tot)],
                self._named(out))

    def Vp(self, channel, rows):
        """Vp_pp (0) / Vp_pm (1) / Vp_mm (2) for rows of (i,j,k,l,q,q',q'',q''') (1-based)."""
        rows = np.ascontiguousarray(np.atleast_2d(rows), np.int32)
        out = np.zeros(len(rows), np.complex128)
        _check(load().fp4_vp_4ph(self.h, C.c_int32(channel), C.c_int64(len(rows)), _p(rows), _p(out)))
        return out

    def Vp_pp(self, rows):
        return self.Vp(0, rows)

    def Vp_pm(self, rows):
        return self.Vp(1, rows)

    def Vp_mm(self, rows):
        return self.Vp(2, rows)

    # ------------------------------------------------------------ iterations
    def set_symmetry(self):
        """Uploads Nsymm_rot, ID_equi, symmetrizers and crotations of the case's modes."""
        m = self.case.modes
        _check(load().fp4_ctx_set_symmetry(self.h, C.c_int32(m.nsymm_rot),
                                           _p(np.asfortranarray(m.ID_equi, np.int32)),
                                           _p(np.asfortranarray(m.symmetrizers, np.float64)),
                                           _p(np.asfortranarray(m.crotations, np.float64))))

    def _orbits(self):
        m = self.case.modes
        return (np.ascontiguousarray(m.Nequi, np.int32), np.asfortranarray(m.ALLEquiList, np.int32),
                np.asfortranarray(m.TypeofSymmetry, np.int32))

    def iteration0(self, tau_zero):
        """iteration0 -> F_n (Nb,nptk,3) F-order."""
        Nequi, alleq, _ = self._orbits()
        tau = np.asfortranarray(tau_zero, np.float64).reshape((self.nb, self.nlist), order="F")
        F = np.zeros((self.nb, self.nptk, 3), order="F")
        _check(load().fp4_iteration0(self.h, C.c_int32(self.nlist), _p(Nequi), _p(alleq), _p(tau), _p(F)))
        return F

    def _operator_args(self, N, lists, gammas, N3=None, lists3=None, gammas3=None):
        nm = self.nmodes
        Nequi, alleq, types = self._orbits()
        N4 = [np.ascontiguousarray(n, np.int64) for n in N]
        L4 = [[np.ascontiguousarray(x, np.int64) if len(x) else np.zeros(1, np.int64) for x in ch] for ch in lists]
        G4 = [np.ascontiguousarray(g, np.float64) if len(g) else np.zeros(1) for g in gammas]
        if N3 is None:
            N3 = [np.zeros(nm, np.int32), np.zeros(nm, np.int32)]
            lists3 = [[np.zeros(1, np.int32)] * 2] * 2
            gammas3 = [np.zeros(1), np.zeros(1)]
        N3 = [np.ascontiguousarray(n, np.int32) for n in N3]
        L3 = [[np.ascontiguousarray(x, np.int32) if len(x) else np.zeros(1, np.int32) for x in ch] for ch in lists3]
        G3 = [np.ascontiguousarray(g, np.float64) if len(g) else np.zeros(1) for g in gammas3]
        keep = (Nequi, alleq, types, N4, L4, G4, N3, L3, G3)
        args = [C.c_int32(self.nlist), _p(Nequi), _p(alleq), _p(types), _p(N3[0]), _p(N3[1]),
                C.c_int32(int(N3[0].sum())), C.c_int32(int(N3[1].sum())), *[_p(n) for n in N4],
                *[C.c_int64(int(n.sum())) for n in N4], *[_p(x) for ch in L3 for x in ch],
                *[_p(x) for ch in L4 for x in ch], _p(G3[0]), _p(G3[1]), *[_p(g) for g in G4]]
        return keep, args

    def bte_set_operator(self, N, lists, gammas, N3=None, lists3=None, gammas3=None):
        """Uploads the scattering operator (4ph counts/lists/Gamma, optional 3ph ones)."""
        keep, args = self._operator_args(N, lists, gammas, N3, lists3, gammas3)
        _check(load().fp4_bte_set_operator(self.h, *args))
        del keep

    def bte_assemble_operator(self, precision=None, with_3ph=True, with_4ph=True, rates=False):
        """Assembles the aggregated scattering operator on the device at the context's T (3ph from the
        FP64 per-process Gamma, 4ph in the given form; default FP64 per-process) -> (n3, n4) processes,
        plus (rate_3ph, rate_4ph) as (Nb, Nlist) arrays when rates=True."""
        precision = PREC_FP64 if precision is None else precision
        Nequi, alleq, types = self._orbits()
        n3, n4 = C.c_int64(), C.c_int64()
        r3 = np.zeros((self.nb, self.nlist), order="F")
        r4 = np.zeros((self.nb, self.nlist), order="F")
        _check(load().fp4_bte_assemble_operator(self.h, C.c_int32(self.nlist), _p(Nequi), _p(alleq), _p(types),
                                                C.c_int32(precision), C.c_int32(int(bool(with_3ph))),
                                                C.c_int32(int(bool(with_4ph))), C.byref(n3), C.byref(n4),
                                                _p(r3) if rates else None, _p(r4) if rates else None))
        if rates:
            return n3.value, n4.value, r3, r4
        return n3.value, n4.value

    def bte_operator_begin(self):
        Nequi, alleq, types = self._orbits()
        _check(load().fp4_bte_operator_begin(self.h, C.c_int32(self.nlist), _p(Nequi), _p(alleq), _p(types)))

    def bte_operator_add(self, modes, precision, with_3ph=True, with_4ph=True, r3=None, r4=None):
        """Adds the processes of a mode subset to the operator being assembled; rates of those modes are
        written into r3 / r4 ((Nb, Nlist) F-order arrays) when given -> (n3, n4)."""
        modes = np.ascontiguousarray(np.sort(np.asarray(modes)), np.int32)
        n3, n4 = C.c_int64(), C.c_int64()
        _check(load().fp4_bte_operator_add(self.h, C.c_int32(precision), C.c_int32(int(bool(with_3ph))),
                                           C.c_int32(int(bool(with_4ph))), C.c_int32(len(modes)), _p(modes),
                                           C.byref(n3), C.byref(n4), _p(r3) if r3 is not None else None,
                                           _p(r4) if r4 is not None else None))
        return n3.value, n4.value

    def bte_operator_end(self):
        _check(load().fp4_bte_operator_end(self.h))

    def iteration_4ph(self, N, lists, gammas, tau_zero, F_n, N3=None, lists3=None, gammas3=None):
        """iteration_4ph on a copy of F_n; returns the updated F_n."""
        keep, args = self._operator_args(N, lists, gammas, N3, lists3, gammas3)
        tau = np.asfortranarray(tau_zero, np.float64).reshape((self.nb, self.nlist), order="F")
        F = np.array(F_n, dtype=np.float64, order="F", copy=True)
        _check(load().fp4_iteration_4ph(self.h, *args, _p(tau), _p(F)))
        del keep
        return F

    def set_modes(self, modes=None, eigenvectors=True):
        """fp4_ctx_set_modes with the case's (or the given) mode data; eigenvectors=False passes none
        (counts and lists only)."""
        m = self.case.modes if modes is None else modes
        ev = _p(np.asfortranarray(m.eigenvect, np.complex128)) if eigenvectors else None
        _check(load().fp4_ctx_set_modes(self.h, _p(np.asfortranarray(m.energy, np.float64)),
                                        _p(np.asfortranarray(m.velocity, np.float64)), ev,
                                        _p(np.asfortranarray(m.IJK, np.int32)), C.c_int32(m.Nlist),
                                        _p(np.ascontiguousarray(m.List, np.int32))))

    def bte_set_sweep(self, persistent):
        """Aggregated-operator sweeps as one cooperative launch (True) or one launch per step (False)."""
        _check(load().fp4_bte_set_sweep(self.h, C.c_int32(int(bool(persistent)))))

    def bte_iterate(self, tau_zero, F_n):
        tau = np.asfortranarray(tau_zero, np.float64).reshape((self.nb, self.nlist), order="F")
        F = np.array(F_n, dtype=np.float64, order="F", copy=True)
        _check(load().fp4_bte_iterate(self.h, _p(tau), _p(F)))
        return F

    def TConduct(self, F_n, V=None, mode=False):
        """-> kappa (Nb,3,3) [, kappa_mode (nptk,Nb,3,3)] in W/(m K)."""
        V = self.case.structure.volume if V is None else V
        F = np.asfortranarray(F_n, np.float64)
        kappa = np.zeros((self.nb, 3, 3), order="F")
        km = np.zeros((self.nptk, self.nb, 3, 3), order="F") if mode else None
        _check(load().fp4_tconduct(self.h, C.c_double(V), _p(F), _p(kappa), _p(km) if mode else None))
        return (kappa, km) if mode else kappa

    def bte_solve(self, tau_zero, maxiter=1000, eps=float(np.float32(1e-5)), V=None):
        """Iterative solution with the current operator -> dict(F_n, kappa, kappa_iter, relchange, niter)."""
        V = self.case.structure.volume if V is None else V
        tau = np.asfortranarray(tau_zero, np.float64).reshape((self.nb, self.nlist), order="F")
        F = np.zeros((self.nb, self.nptk, 3), order="F")
        kappa = np.zeros((self.nb, 3, 3), order="F")
        hist = np.zeros((3, 3, maxiter + 1), order="F")
        rel = np.zeros(maxiter + 1)
        n = C.c_int32()
        _check(load().fp4_bte_solve(self.h, C.c_double(V), _p(tau), C.c_int32(maxiter), C.c_double(eps), _p(F),
                                    _p(kappa), _p(hist), _p(rel), C.byref(n)))
        k = n.value
        return {"F_n": F, "kappa": kappa, "kappa_iter": hist[:, :, :k + 1], "relchange": rel[:k + 1], "niter": k}

    # ------------------------------------------------------------ diagnostics
    KS_PREFILTER, KS_EMIT, KS_RATE64, KS_DENSE, KS_ITER, KS_SCAN3, KS_RATE3, KS_SAMPLE = 0, 1, 2, 3, 4, 5, 6, 7

    def kernel_stats(self, kind):
        ms = C.c_double()
        n = C.c_int64()
        _check(load().fp4_ctx_kernel_stats(self.h, C.c_int32(kind), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def reset_stats(self):
        _check(load().fp4_ctx_reset_stats(self.h))

    def dense_info(self):
        """-> dict(triples, items, ngroups, nfine) of the last dense-form call."""
        t, it = C.c_int64(), C.c_int64()
        g, f = C.c_int32(), C.c_int32()
        _check(load().fp4_ctx_dense_info(self.h, C.byref(t), C.byref(it), C.byref(g), C.byref(f)))
        return {"triples": t.value, "items": it.value, "ngroups": g.value, "nfine": f.value}

    def autotune_dense(self, precision, modes, ii_lo=1, ii_hi=None):
        """Runs the dense-form autotuner on a mode subset and q' window -> dict(ib, tq, slots, ms)."""
        modes = np.ascontiguousarray(np.sort(np.asarray(modes)), np.int32)
        ib, tq, sl = C.c_int32(), C.c_int32(), C.c_int32()
        ms = C.c_double()
        _check(load().fp4_ctx_autotune_dense(self.h, C.c_int32(precision), C.c_int32(len(modes)), _p(modes),
                                             C.c_int32(ii_lo), C.c_int32(self.nptk if ii_hi is None else ii_hi),
                                             C.byref(ib), C.byref(tq), C.byref(sl), C.byref(ms)))
        return {"ib": ib.value, "tq": tq.value, "slots": sl.value, "ms": ms.value}

    def dense_shape(self):
        """-> dict(mfma, tq, ib) of the last dense-form call."""
        a, b, c = C.c_int32(), C.c_int32(), C.c_int32()
        _check(load().fp4_ctx_dense_shape(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return {"mfma": a.value, "tq": b.value, "ib": c.value}   # mfma: 0 VALU, 1 FP32, 2 BF16 hi/lo, 3 FP16 hi/lo

    def dense_candidates(self, cap=64):
        """-> list of dict(kg, ib, tq, slots, ms): the candidates the last autotune timed."""
        kg, ib, tq, sl = (np.zeros(cap, np.int32) for _ in range(4))
        ms = np.zeros(cap)
        n = load().fp4_ctx_dense_candidates(self.h, C.c_int32(cap), _p(kg), _p(ib), _p(tq), _p(sl), _p(ms))
        if n < 0:
            _check(-n)
        return [{"kg": int(kg[i]), "ib": int(ib[i]), "tq": int(tq[i]), "slots": int(sl[i]), "ms": float(ms[i])}
                for i in range(min(n, cap))]

    DENSE_PHASES = ["item", "setup", "scan", "tile", "bar_tile", "filter", "bar_filter", "queue", "gemm",
                    "contract", "rate", "bar_end"]

    def dense_phase_profile(self):
        """Profiling build only: {phase: shader cycles summed over waves} of the last dense call."""
        cyc = np.zeros(len(self.DENSE_PHASES))
        n = load().fp4_ctx_dense_phase_profile(self.h, C.c_int32(len(cyc)), _p(cyc))
        if n < 0:
            _check(-n)
        return dict(zip(self.DENSE_PHASES, cyc.tolist()))

    def set_dense_shape(self, precision, kg, ib, slots=0):
        """Makes later dense calls of this precision run the given kernel (kg = 0: block-cooperative) and IB."""
        _check(load().fp4_ctx_set_dense_shape(self.h, C.c_int32(precision), C.c_int32(kg), C.c_int32(ib),
                                              C.c_int32(slots)))

    def dense_kernel(self):
        """-> coarse groups per MFMA k-step of the last dense call's kernel (0: block-cooperative kernel)."""
        k = C.c_int32()
        _check(load().fp4_ctx_dense_kernel(self.h, C.byref(k)))
        return k.value

    # ------------------------------------------------------- three phonons
    _RATE3 = ["rate_scatt", "rate_scatt_plus", "rate_scatt_minus", "WP3_plus", "WP3_minus"]

    def NP_driver(self):
        """-> (N_plus, P_plus, N_minus, P_minus); N int32 (Nlist*Nb), P (Nb,Nlist)."""
        nm = self.nmodes
        N = [np.zeros(nm, np.int32) for _ in range(2)]
        P = [np.zeros((self.nb, self.nlist), order="F") for _ in range(2)]
        _check(load().fp4_np_3ph(self.h, _p(N[0]), _p(P[0]), _p(N[1]), _p(P[1])))
        return N[0], P[0], N[1], P[1]

    def Ind_only_driver(self, N_plus, N_minus):
        """-> [[Indof2nd, Indof3rd] for +, -] (int32, reference encoding)."""
        Ns = [np.ascontiguousarray(n, np.int32) for n in (N_plus, N_minus)]
        outs = [[np.zeros(max(int(n.sum()), 1), np.int32) for _ in range(2)] for n in Ns]
        _check(load().fp4_ind_only_3ph(self.h, *[_p(n) for n in Ns], *[_p(o) for ch in outs for o in ch]))
        return [[o[:int(n.sum())] for o in ch] for ch, n in zip(outs, Ns)]

    def _rate3_outputs(self):
        return [np.zeros((self.nb, self.nlist), order="F") for _ in range(5)]

    def RTA_driver(self):
        out = self._rate3_outputs()
        _check(load().fp4_rta_3ph(self.h, *[_p(o) for o in out]))
        return dict(zip(self._RATE3, out))

    def NP_driver_sample(self, nsample, seed):
        """Sampled NP_driver -> (N_plus, P_plus, N_minus, P_minus) estimates."""
        nm = self.nmodes
        N = [np.zeros(nm, np.int32) for _ in range(2)]
        P = [np.zeros((self.nb, self.nlist), order="F") for _ in range(2)]
        _check(load().fp4_np_3ph_sample(self.h, C.c_int32(nsample), C.c_uint64(seed), _p(N[0]), _p(P[0]), _p(N[1]),
                                        _p(P[1])))
        return N[0], P[0], N[1], P[1]

    def RTA_driver_sample(self, nsample, seed, modes=None):
        """Sampled RTA_driver -> (rates dict, sampled processes evaluated or None)."""
        out = self._rate3_outputs()
        if modes is None:
            _check(load().fp4_rta_3ph_sample(self.h, C.c_int32(nsample), C.c_uint64(seed), *[_p(o) for o in out]))
            return dict(zip(self._RATE3, out)), None
        modes = np.ascontiguousarray(np.sort(np.asarray(modes)), np.int32)
        n = C.c_int64()
        _check(load().fp4_rta_3ph_sample_modes(self.h, C.c_int32(nsample), C.c_uint64(seed), C.c_int32(len(modes)),
                                               _p(modes), *[_p(o) for o in out], C.byref(n)))
        return dict(zip(self._RATE3, out)), n.value

    def RTA_driver_modes(self, modes, ii_lo=1, ii_hi=None):
        """RTA_driver over a mode subset and a q' window -> (dict, nproc)."""
        out = self._rate3_outputs()
        modes = np.ascontiguousarray(np.sort(np.asarray(modes)), np.int32)
        nproc = C.c_int64()
        _check(load().fp4_rta_3ph_modes(self.h, C.c_int32(len(modes)), _p(modes), C.c_int32(ii_lo),
                                        C.c_int32(self.nptk if ii_hi is None else ii_hi),
                                        *[_p(o) for o in out], C.byref(nproc)))
        return dict(zip(self._RATE3, out)), nproc.value

    def RTA_driver_GPU_using_Ind(self, N_plus, N_minus, lists):
        Ns = [np.ascontiguousarray(n, np.int32) for n in (N_plus, N_minus)]
        L2 = [[np.ascontiguousarray(x, np.int32) if len(x) else np.zeros(1, np.int32) for x in ch] for ch in lists]
        out = self._rate3_outputs()
        # reference argument order: 2nd+, 2nd-, 3rd+, 3rd-
        _check(load().fp4_rta_3ph_using_ind(self.h, _p(Ns[0]), _p(Ns[1]), _p(L2[0][0]), _p(L2[1][0]),
                                            _p(L2[0][1]), _p(L2[1][1]), *[_p(o) for o in out]))
        return dict(zip(self._RATE3, out))

    def Ind_driver(self, N_plus, N_minus):
        """-> (lists, [Gamma_plus, Gamma_minus], rates dict)."""
        Ns = [np.ascontiguousarray(n, np.int32) for n in (N_plus, N_minus)]
        tot = [int(n.sum()) for n in Ns]
        lists = [[np.zeros(max(t, 1), np.int32) for _ in range(2)] for t in tot]
        gam = [np.zeros(max(t, 1)) for t in tot]
        out = self._rate3_outputs()
        _check(load().fp4_ind_3ph(self.h, _p(Ns[0]), _p(Ns[1]), _p(lists[0][0]), _p(lists[0][1]), _p(gam[0]),
                                  _p(lists[1][0]), _p(lists[1][1]), _p(gam[1]), *[_p(o) for o in out]))
        return ([[x[:t] for x in ch] for ch, t in zip(lists, tot)], [g[:t] for g, t in zip(gam, tot)],
                dict(zip(self._RATE3, out)))

    def Vp3(self, channel, rows):
        """Vp_plus (0) / Vp_minus (1) for rows of (i, j, k, q, q', q'') -> complex array."""
        rows = np.ascontiguousarray(rows, np.int32).reshape(-1, 6)
        out = np.zeros((len(rows), 2))
        _check(load().fp4_vp_3ph(self.h, C.c_int32(channel), C.c_int64(len(rows)), _p(rows), _p(out)))
        return out[:, 0] + 1j * out[:, 1]

    def set_batch_capacity(self, processes):
        _check(load().fp4_ctx_set_batch_capacity(self.h, C.c_int64(processes)))


class Batch:
    """Several crystals' RTA_driver_4ph in one persistent launch over all of them (fp4_batch_*)."""

    def __init__(self, contexts):
        self.contexts = list(contexts)
        arr = (C.c_void_p * len(self.contexts))(*[c.h.value for c in self.contexts])
        h = C.c_void_p()
        _check(load().fp4_batch_create(C.c_int32(len(self.contexts)), arr, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            load().fp4_batch_destroy(self.h)
            self.h = None

    def RTA_driver_4ph(self, precision, use_graph=True, part=0, nparts=1):
        """-> (list of per-crystal dicts of (Nb, Nlist) arrays, nproc per crystal, kernel ms, launches) for
        q' window part of nparts (the parts' rates add up to the full rates)."""
        outs = [np.zeros(7 * c.nb * c.nlist) for c in self.contexts]
        ptrs = (C.c_void_p * len(outs))(*[o.ctypes.data for o in outs])
        nproc = np.zeros(len(outs), np.int64)
        ms = C.c_double()
        nl = C.c_int32()
        _check(load().fp4_batch_rta_4ph(self.h, C.c_int32(precision), C.c_int32(int(bool(use_graph))),
                                        C.c_int32(part), C.c_int32(nparts), ptrs, _p(nproc), C.byref(ms),
                                        C.byref(nl)))
        keys = ("rate_scatt_4ph", "rate_scatt_plusplus", "rate_scatt_plusminus", "rate_scatt_minusminus",
                "WP4_plusplus", "WP4_plusminus", "WP4_minusminus")
        res = []
        for c, o in zip(self.contexts, outs):
            nm = c.nb * c.nlist
            res.append({k: o[t * nm:(t + 1) * nm].reshape((c.nb, c.nlist), order="F") for t, k in enumerate(keys)})
        return res, nproc, ms.value, nl.value
