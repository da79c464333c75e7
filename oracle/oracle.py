"""ctypes front end of the CPU oracle -- TEST INFRASTRUCTURE.

Only the tests, ``__graft_entry__.smoke()`` and the benchmark's CPU-baseline
leg import this module.  It wraps ``libfp4oracle.so`` (built from
``fp4_oracle.c`` by ``oracle/Makefile``), the FP64 CPU restatement of the
reference's four-phonon path and its three-phonon companion; see
``fp4_oracle.c`` for the reference lines each routine follows.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfp4oracle.so")

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


class OrcCase(C.Structure):
    _fields_ = [
        ("ngrid", C.c_int32 * 3), ("nptk", C.c_int32), ("nbands", C.c_int32), ("nlist", C.c_int32),
        ("rlattvec", C.c_double * 9), ("T", C.c_double), ("scalebroad", C.c_double),
        ("omega_max", C.c_double), ("onlyharmonic", C.c_int32),
        ("masses", C.c_void_p), ("types", C.c_void_p), ("energy", C.c_void_p), ("velocity", C.c_void_p),
        ("eigenvect", C.c_void_p), ("IJK", C.c_void_p), ("List", C.c_void_p), ("ntri", C.c_int32),
        ("Phi", C.c_void_p), ("R_j", C.c_void_p), ("R_k", C.c_void_p), ("R_l", C.c_void_p),
        ("Index_i", C.c_void_p), ("Index_j", C.c_void_p), ("Index_k", C.c_void_p), ("Index_l", C.c_void_p),
        ("ntri3", C.c_int32), ("Phi3", C.c_void_p), ("R3_j", C.c_void_p), ("R3_k", C.c_void_p),
        ("Index3_i", C.c_void_p), ("Index3_j", C.c_void_p), ("Index3_k", C.c_void_p),
    ]


def build(quiet: bool = True) -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True,
                   stdout=subprocess.DEVNULL if quiet else None)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class Oracle:
    """Holds one Case's arrays in the layout the C oracle reads."""

    def __init__(self, case):
        m, f, st = case.modes, case.ifc4, case.structure
        self.case = case
        self.nptk, self.nb, self.nlist = m.nptk, m.nbands, m.Nlist
        keep = dict(
            masses=np.ascontiguousarray(st.masses, np.float64),
            types=np.ascontiguousarray(st.types, np.int32),
            energy=np.asfortranarray(m.energy, np.float64),
            velocity=np.asfortranarray(m.velocity, np.float64),
            eigenvect=np.asfortranarray(m.eigenvect, np.complex128),
            IJK=np.asfortranarray(m.IJK, np.int32),
            List=np.ascontiguousarray(m.List, np.int32),
            Phi=np.asfortranarray(f.Phi, np.float64),
            R_j=np.asfortranarray(f.R_j, np.float64), R_k=np.asfortranarray(f.R_k, np.float64),
            R_l=np.asfortranarray(f.R_l, np.float64),
            Index_i=np.ascontiguousarray(f.Index_i, np.int32), Index_j=np.ascontiguousarray(f.Index_j, np.int32),
            Index_k=np.ascontiguousarray(f.Index_k, np.int32), Index_l=np.ascontiguousarray(f.Index_l, np.int32),
        )
        f3 = getattr(case, "ifc3", None)
        if f3 is not None:
            keep.update(
                Phi3=np.asfortranarray(f3.Phi, np.float64),
                R3_j=np.asfortranarray(f3.R_j, np.float64), R3_k=np.asfortranarray(f3.R_k, np.float64),
                Index3_i=np.ascontiguousarray(f3.Index_i, np.int32),
                Index3_j=np.ascontiguousarray(f3.Index_j, np.int32),
                Index3_k=np.ascontiguousarray(f3.Index_k, np.int32))
        self._keep = keep
        c = OrcCase()
        c.ngrid[:] = [int(x) for x in m.ngrid]
        c.nptk, c.nbands, c.nlist = self.nptk, self.nb, self.nlist
        c.rlattvec[:] = list(np.asfortranarray(st.rlattvec).ravel(order="F"))
        c.T, c.scalebroad, c.omega_max = case.T, case.scalebroad, case.omega_max
        c.onlyharmonic = int(case.onlyharmonic)
        for k, v in keep.items():
            setattr(c, k, _ptr(v))
        c.ntri = f.Ntri
        c.ntri3 = f3.Ntri if f3 is not None else 0
        self.c = c
        self.L = lib()

    @property
    def nmodes(self):
        return self.nb * self.nlist

    # ------------------------------------------------------------------ kernels
    def base_sigma(self, v):
        v = np.ascontiguousarray(v, np.float64)
        self.L.orc_base_sigma.restype = C.c_double
        return self.L.orc_base_sigma(C.byref(self.c), _ptr(v))

    def np_mode(self, ch, mm):
        n = C.c_int64()
        p = C.c_double()
        self.L.orc_np_mode(C.byref(self.c), C.c_int(ch), C.c_int(mm), C.byref(n), C.byref(p))
        return n.value, p.value

    def ind_only_mode(self, ch, mm, cap=None):
        if cap is None:
            cap = self.np_mode(ch, mm)[0]
        out = [np.zeros(max(cap, 1), np.int64) for _ in range(3)]
        self.L.orc_ind_only_mode.restype = C.c_int64
        n = self.L.orc_ind_only_mode(C.byref(self.c), C.c_int(ch), C.c_int(mm), C.c_int64(cap),
                                     *[_ptr(o) for o in out])
        return [o[:n] for o in out]

    def list_hash_mode(self, ch, mm):
        """(count, phase space, list hash) of one mode's list, streamed without storing the list."""
        n = C.c_int64()
        p = C.c_double()
        self.L.orc_list_hash_mode.restype = C.c_uint64
        h = self.L.orc_list_hash_mode(C.byref(self.c), C.c_int(ch), C.c_int(mm), C.byref(n), C.byref(p))
        return n.value, p.value, int(h)

    def vp(self, ch, i, j, k, l, q, qp, qdp, qtp):
        out = np.zeros(2)
        self.L.orc_vp_idx(C.byref(self.c), C.c_int(ch), *[C.c_int(int(x)) for x in (i, j, k, l, q, qp, qdp, qtp)],
                          _ptr(out))
        return complex(out[0], out[1])

    def rta_mode(self, ch, mm, ii_lo=1, ii_hi=None):
        g = C.c_double()
        w = C.c_double()
        self.L.orc_rta_mode(C.byref(self.c), C.c_int(ch), C.c_int(mm), C.c_int32(ii_lo),
                            C.c_int32(self.nptk if ii_hi is None else ii_hi), C.byref(g), C.byref(w))
        return g.value, w.value

    # ------------------------------------------------------------------ drivers
    def np_driver(self, nthreads=0):
        nm = self.nmodes
        N = [np.zeros(nm, np.int64) for _ in range(3)]
        P = [np.zeros(nm, np.float64) for _ in range(3)]
        self.L.orc_np_driver(C.byref(self.c), C.c_int(nthreads), _ptr(N[0]), _ptr(P[0]), _ptr(N[1]), _ptr(P[1]),
                             _ptr(N[2]), _ptr(P[2]))
        return N, P

    def ind_only_driver(self, N, nthreads=0):
        outs = [[np.zeros(max(int(n.sum()), 1), np.int64) for _ in range(3)] for n in N]
        rc = self.L.orc_ind_only_driver(C.byref(self.c), C.c_int(nthreads), *[_ptr(n) for n in N],
                                        *[_ptr(o) for ch in outs for o in ch])
        if rc != 0:
            raise RuntimeError("oracle list longer than its count")
        return [[o[:int(n.sum())] for o in ch] for ch, n in zip(outs, N)]

    def rta_driver(self, modes=None, ii_lo=1, ii_hi=None, nthreads=0):
        nm = self.nmodes
        modes = np.arange(1, nm + 1, dtype=np.int32) if modes is None else np.ascontiguousarray(modes, np.int32)
        out = [np.zeros(nm, np.float64) for _ in range(7)]
        self.L.orc_rta_driver(C.byref(self.c), C.c_int(nthreads), C.c_int(len(modes)), _ptr(modes),
                              C.c_int32(ii_lo), C.c_int32(self.nptk if ii_hi is None else ii_hi),
                              *[_ptr(o) for o in out])
        names = ["rate_4ph", "rate_pp", "rate_pm", "rate_mm", "wp4_pp", "wp4_pm", "wp4_mm"]
        return dict(zip(names, out))

    # ---------------------------------------------------- sampling estimators
    def sample_mode(self, ch, mm, nsample, seed, want_rates=False):
        """Raw sums of one (channel, mode) -> (hits, P, Gamma, WP4) before unit and sampling factors."""
        h, p, g, w = C.c_int64(), C.c_double(), C.c_double(), C.c_double()
        self.L.orc_sample_mode(C.byref(self.c), C.c_int(ch), C.c_int(mm), C.c_int32(nsample), C.c_uint64(seed),
                               C.c_int(int(want_rates)), C.byref(h), C.byref(p), C.byref(g), C.byref(w))
        return h.value, p.value, g.value, w.value

    def np_sample_driver(self, nsample, seed, nthreads=0):
        nm = self.nmodes
        N = [np.zeros(nm, np.int64) for _ in range(3)]
        P = [np.zeros(nm, np.float64) for _ in range(3)]
        self.L.orc_np_sample_driver(C.byref(self.c), C.c_int(nthreads), C.c_int32(nsample), C.c_uint64(seed),
                                    _ptr(N[0]), _ptr(P[0]), _ptr(N[1]), _ptr(P[1]), _ptr(N[2]), _ptr(P[2]))
        return N, P

    def rta_sample_driver(self, nsample, seed, modes=None, nthreads=0):
        nm = self.nmodes
        modes = np.arange(1, nm + 1, dtype=np.int32) if modes is None else np.ascontiguousarray(modes, np.int32)
        out = [np.zeros(nm, np.float64) for _ in range(7)]
        hits = C.c_int64()
        self.L.orc_rta_sample_driver(C.byref(self.c), C.c_int(nthreads), C.c_int32(nsample), C.c_uint64(seed),
                                     C.c_int(len(modes)), _ptr(modes), *[_ptr(o) for o in out], C.byref(hits))
        names = ["rate_4ph", "rate_pp", "rate_pm", "rate_mm", "wp4_pp", "wp4_pm", "wp4_mm"]
        return dict(zip(names, out)), hits.value

    def ind_driver(self, N, nthreads=0):
        nm = self.nmodes
        tot = [int(n.sum()) for n in N]
        lists = [[np.zeros(max(t, 1), np.int64) for _ in range(3)] for t in tot]
        gam = [np.zeros(max(t, 1), np.float64) for t in tot]
        rates = [np.zeros(nm, np.float64) for _ in range(7)]
        rc = self.L.orc_ind_driver(C.byref(self.c), C.c_int(nthreads), *[_ptr(n) for n in N],
                                   *[_ptr(o) for ch in lists for o in ch], *[_ptr(g) for g in gam],
                                   *[_ptr(r) for r in rates])
        if rc != 0:
            raise RuntimeError("oracle list longer than its count")
        names = ["rate_4ph", "rate_pp", "rate_pm", "rate_mm", "wp4_pp", "wp4_pm", "wp4_mm"]
        return ([[o[:t] for o in ch] for ch, t in zip(lists, tot)], [g[:t] for g, t in zip(gam, tot)],
                dict(zip(names, rates)))


    # ------------------------------------------------------- three phonons
    def _need3(self):
        if self.c.ntri3 <= 0:
            raise ValueError("case has no third-order constants")

    def np3_mode(self, ch, mm):
        n = C.c_int64()
        p = C.c_double()
        self.L.orc3_np_mode(C.byref(self.c), C.c_int(ch), C.c_int(mm), C.byref(n), C.byref(p))
        return n.value, p.value

    def ind_only3_mode(self, ch, mm, cap=None):
        if cap is None:
            cap = self.np3_mode(ch, mm)[0]
        out = [np.zeros(max(cap, 1), np.int32) for _ in range(2)]
        self.L.orc3_ind_only_mode.restype = C.c_int64
        n = self.L.orc3_ind_only_mode(C.byref(self.c), C.c_int(ch), C.c_int(mm), C.c_int64(cap),
                                      *[_ptr(o) for o in out])
        return [o[:n] for o in out]

    def vp3(self, ch, i, j, k, q, qp, qdp):
        self._need3()
        out = np.zeros(2)
        self.L.orc3_vp_idx(C.byref(self.c), C.c_int(ch), *[C.c_int(int(x)) for x in (i, j, k, q, qp, qdp)],
                           _ptr(out))
        return complex(out[0], out[1])

    def np3_driver(self, nthreads=0):
        nm = self.nmodes
        N = [np.zeros(nm, np.int32) for _ in range(2)]
        P = [np.zeros(nm, np.float64) for _ in range(2)]
        self.L.orc3_np_driver(C.byref(self.c), C.c_int(nthreads), _ptr(N[0]), _ptr(P[0]), _ptr(N[1]), _ptr(P[1]))
        return N, P

    def ind_only3_driver(self, N, nthreads=0):
        N = [np.ascontiguousarray(n, np.int32) for n in N]
        outs = [[np.zeros(max(int(n.sum()), 1), np.int32) for _ in range(2)] for n in N]
        rc = self.L.orc3_ind_only_driver(C.byref(self.c), C.c_int(nthreads), *[_ptr(n) for n in N],
                                         *[_ptr(o) for ch in outs for o in ch])
        if rc != 0:
            raise RuntimeError("oracle list longer than its count")
        return [[o[:int(n.sum())] for o in ch] for ch, n in zip(outs, N)]

    def rta3_driver(self, modes=None, ii_lo=1, ii_hi=None, nthreads=0):
        self._need3()
        nm = self.nmodes
        modes = np.arange(1, nm + 1, dtype=np.int32) if modes is None else np.ascontiguousarray(modes, np.int32)
        out = [np.zeros(nm, np.float64) for _ in range(5)]
        self.L.orc3_rta_driver(C.byref(self.c), C.c_int(nthreads), C.c_int(len(modes)), _ptr(modes),
                               C.c_int32(ii_lo), C.c_int32(self.nptk if ii_hi is None else ii_hi),
                               *[_ptr(o) for o in out])
        return dict(zip(["rate", "rate_plus", "rate_minus", "wp3_plus", "wp3_minus"], out))

    def sample3_mode(self, ch, mm, nsample, seed, want_rates=False):
        """Raw 3ph sums of one (channel, mode) -> (hits, P, Gamma, WP3) before unit and sampling factors."""
        if want_rates:
            self._need3()
        h, p, g, w = C.c_int64(), C.c_double(), C.c_double(), C.c_double()
        self.L.orc3_sample_mode(C.byref(self.c), C.c_int(ch), C.c_int(mm), C.c_int32(nsample), C.c_uint64(seed),
                                C.c_int(int(want_rates)), C.byref(h), C.byref(p), C.byref(g), C.byref(w))
        return h.value, p.value, g.value, w.value

    def np3_sample_driver(self, nsample, seed, nthreads=0):
        nm = self.nmodes
        N = [np.zeros(nm, np.int32) for _ in range(2)]
        P = [np.zeros(nm, np.float64) for _ in range(2)]
        rc = self.L.orc3_np_sample_driver(C.byref(self.c), C.c_int(nthreads), C.c_int32(nsample), C.c_uint64(seed),
                                          _ptr(N[0]), _ptr(P[0]), _ptr(N[1]), _ptr(P[1]))
        if rc != 0:
            raise OverflowError("a 3ph count estimate exceeds int32")
        return N, P

    def rta3_sample_driver(self, nsample, seed, modes=None, nthreads=0):
        self._need3()
        nm = self.nmodes
        modes = np.arange(1, nm + 1, dtype=np.int32) if modes is None else np.ascontiguousarray(modes, np.int32)
        out = [np.zeros(nm, np.float64) for _ in range(5)]
        hits = C.c_int64()
        self.L.orc3_rta_sample_driver(C.byref(self.c), C.c_int(nthreads), C.c_int32(nsample), C.c_uint64(seed),
                                      C.c_int(len(modes)), _ptr(modes), *[_ptr(o) for o in out], C.byref(hits))
        return dict(zip(["rate", "rate_plus", "rate_minus", "wp3_plus", "wp3_minus"], out)), hits.value

    def ind3_driver(self, N, nthreads=0):
        self._need3()
        nm = self.nmodes
        N = [np.ascontiguousarray(n, np.int32) for n in N]
        tot = [int(n.sum()) for n in N]
        lists = [[np.zeros(max(t, 1), np.int32) for _ in range(2)] for t in tot]
        gam = [np.zeros(max(t, 1), np.float64) for t in tot]
        rates = [np.zeros(nm, np.float64) for _ in range(5)]
        rc = self.L.orc3_ind_driver(C.byref(self.c), C.c_int(nthreads), _ptr(N[0]), _ptr(N[1]),
                                    _ptr(lists[0][0]), _ptr(lists[0][1]), _ptr(gam[0]),
                                    _ptr(lists[1][0]), _ptr(lists[1][1]), _ptr(gam[1]), *[_ptr(r) for r in rates])
        if rc != 0:
            raise RuntimeError("oracle list longer than its count")
        return ([[o[:t] for o in ch] for ch, t in zip(lists, tot)], [g[:t] for g, t in zip(gam, tot)],
                dict(zip(["rate", "rate_plus", "rate_minus", "wp3_plus", "wp3_minus"], rates)))


def iteration0(case, tau_zero):
    m = case.modes
    F = np.zeros((m.nbands, m.nptk, 3), order="F")
    tz = np.asfortranarray(tau_zero, np.float64)
    lib().orc_iteration0(C.c_int(m.nptk), C.c_int(m.nbands), C.c_int(m.Nlist), C.c_int(m.nsymm_rot),
                         _ptr(np.ascontiguousarray(m.Nequi, np.int32)), _ptr(m.ALLEquiList), _ptr(m.energy),
                         _ptr(m.velocity), _ptr(tz), _ptr(F))
    return F


def iteration_4ph(case, N, lists, gammas, tau_zero, F_n, N3=None, lists3=None, gammas3=None, symmetrizers=None):
    """iteration_4ph on a copy of F_n.  N3/lists3/gammas3: optional 3ph counts (int32), [[i2, i3] for +, -]
    and Gamma; ``symmetrizers`` overrides the case's (3,3,nptk) array."""
    m = case.modes
    F = np.asfortranarray(F_n.copy())
    tz = np.asfortranarray(tau_zero, np.float64)
    if N3 is None:
        p3 = [None] * 8
    else:
        n3 = [np.ascontiguousarray(n, np.int32) for n in N3]
        l3 = [np.ascontiguousarray(x, np.int32) if len(x) else np.zeros(1, np.int32) for ch in lists3 for x in ch]
        g3 = [np.ascontiguousarray(g, np.float64) if len(g) else np.zeros(1) for g in gammas3]
        keep3 = (n3, l3, g3)
        p3 = [_ptr(x) for x in n3 + l3 + g3]
    S = m.symmetrizers if symmetrizers is None else np.asfortranarray(symmetrizers, np.float64)
    args = [C.c_int(m.nptk), C.c_int(m.nbands), C.c_int(m.Nlist), C.c_int(m.nsymm_rot),
            _ptr(np.ascontiguousarray(m.Nequi, np.int32)), _ptr(m.ALLEquiList), _ptr(m.TypeofSymmetry),
            _ptr(m.ID_equi)] + p3 + [_ptr(n) for n in N]
    args += [_ptr(np.ascontiguousarray(o, np.int64)) for ch in lists for o in ch]
    args += [_ptr(m.energy), _ptr(m.velocity)] + [_ptr(np.ascontiguousarray(g, np.float64)) for g in gammas]
    args += [_ptr(tz), _ptr(S), _ptr(F)]
    lib().orc_iteration_4ph(*args)
    return F


def symmetrize_F(case, F_n):
    """Symmetrisers applied to F_n(:,q,:) for every q (ShengBTE.f90:1288-1290); returns a copy."""
    m = case.modes
    F = np.array(F_n, dtype=np.float64, order="F", copy=True)
    lib().orc_symmetrize_F(C.c_int(m.nptk), C.c_int(m.nbands), _ptr(m.symmetrizers), _ptr(F))
    return F


def symmetrize_tensor(case, tensor):
    """symmetrize_tensor (config.f90:503-518) of a (3,3) array; returns a copy."""
    m = case.modes
    t = np.array(tensor, dtype=np.float64, order="F", copy=True)
    lib().orc_symmetrize_tensor(C.c_int(m.nsymm_rot), _ptr(np.asfortranarray(m.crotations, np.float64)), _ptr(t))
    return t


def twonorm3x3(a):
    """2-norm of a 3x3 matrix: the largest singular value from LAPACK gesdd, as misc.f90:52-65."""
    return float(np.linalg.svd(np.asarray(a, np.float64), compute_uv=False)[0])


def bte_solve(case, N, lists, gammas, tau_zero, maxiter=1000, eps=float(np.float32(1e-5)), N3=None, lists3=None,
              gammas3=None):
    """The iterative loop of ShengBTE.f90:1250-1307 with the four-phonon iteration:
    -> dict(F_n, kappa (Nb,3,3), kappa_iter (3,3,niter+1), relchange (niter+1), niter)."""
    m = case.modes
    F = iteration0(case, tau_zero)

    def kappa_of(F):
        k = tconduct(case, F)
        for b in range(m.nbands):
            k[b] = symmetrize_tensor(case, k[b])
        return k

    k = kappa_of(F)
    hist, rel = [k.sum(axis=0)], [0.0]
    it = 0
    for it in range(1, maxiter + 1):
        old = hist[-1]
        F = iteration_4ph(case, N, lists, gammas, tau_zero, F, N3, lists3, gammas3)
        F = symmetrize_F(case, F)
        k = kappa_of(F)
        ks = k.sum(axis=0)
        hist.append(ks)
        rel.append(twonorm3x3(ks - old) / twonorm3x3(old))
        if rel[-1] < eps:
            break
    return {"F_n": F, "kappa": k, "kappa_iter": np.stack(hist, axis=-1), "relchange": np.array(rel), "niter": it}


def tconduct(case, F_n):
    m = case.modes
    kappa = np.zeros((m.nbands, 3, 3), order="F")
    lib().orc_tconduct(C.c_int(m.nptk), C.c_int(m.nbands), C.c_double(case.T),
                       C.c_double(case.structure.volume), _ptr(m.energy), _ptr(m.velocity),
                       _ptr(np.asfortranarray(F_n)), _ptr(kappa))
    return kappa
