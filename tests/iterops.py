"""Synthetic scattering operators and a dense restatement of one iteration
sweep, shared by the oracle and device tests of the iterative solver.

The iterative update of Src/iterations.f90:95-215 is, for fixed symmetrisers
equal to the identity, one Gauss-Seidel sweep over the rows (i, q_kk) in the
order (ll, i, kk) of the linear system

    F(i, q) = tau(i, ll) [ v(q, i) omega(q, i) + sum_c A[(i, q), c] F(c) ],

whose coefficients A come from the per-process Gamma and the list entries
mapped through ID_equi(TypeofSymmetry(kk, ll), .).  ``gauss_seidel_sweep``
solves the triangular system (I - L) x_new = b + U x_old with a dense LU
solve; ``linear_solution`` solves the full system.  Neither walks the lists
in the sweep's order, so they check the in-place sequential update, the
channel signs and the index decoding by a different route.
"""
from __future__ import annotations

import dataclasses

import numpy as np


def synthetic_operator(case, seed, per_mode=4, with_3ph=True):
    """Random valid lists and positive Gamma for every mode and list type.

    Returns (N4, lists4, gammas4, N3, lists3, gammas3) in the reference's
    kinds: 4ph counts/lists int64, 3ph int32."""
    m = case.modes
    rng = np.random.default_rng(seed)
    nm, nq = m.nbands * m.Nlist, m.nptk * m.nbands
    N4 = [rng.integers(0, per_mode + 1, nm).astype(np.int64) for _ in range(3)]
    lists4 = [[rng.integers(1, nq + 1, int(n.sum())).astype(np.int64) for _ in range(3)] for n in N4]
    gammas4 = [rng.uniform(0.05, 1.0, int(n.sum())) for n in N4]
    if not with_3ph:
        return N4, lists4, gammas4, None, None, None
    N3 = [rng.integers(0, per_mode + 1, nm).astype(np.int32) for _ in range(2)]
    lists3 = [[rng.integers(1, nq + 1, int(n.sum())).astype(np.int32) for _ in range(2)] for n in N3]
    gammas3 = [rng.uniform(0.05, 1.0, int(n.sum())) for n in N3]
    return N4, lists4, gammas4, N3, lists3, gammas3


def contractive_tau(case, N4, gammas4, N3=None, gammas3=None, margin=0.25):
    """tau_zero(Nb,Nlist) with tau * (row sum of |A|) <= margin for every mode."""
    m = case.modes
    nm = m.nbands * m.Nlist
    load = np.zeros(nm)
    weights4 = (3.0, 3.0, 3.0)
    for n, g, wgt in zip(N4, gammas4, weights4):
        seg = np.repeat(np.arange(nm), n)
        load += wgt * np.bincount(seg, weights=np.abs(g), minlength=nm)
    if N3 is not None:
        for n, g, wgt in zip(N3, gammas3, (2.0, 1.0)):
            seg = np.repeat(np.arange(nm), n)
            load += wgt * np.bincount(seg, weights=np.abs(g), minlength=nm)
    return np.asfortranarray((margin / (load + 1.0)).reshape(m.nbands, m.Nlist, order="F"))


def identity_symmetrizers(case):
    """A copy of the case whose symmetrisers are all the identity."""
    m = case.modes
    eye = np.asfortranarray(np.repeat(np.eye(3)[:, :, None], m.nptk, axis=2))
    return dataclasses.replace(case, modes=dataclasses.replace(m, symmetrizers=eye))


def _system(case, N4, lists4, gammas4, N3, lists3, gammas3, tau_zero):
    """Dense (B, b, order): B = diag(tau) A, b = tau v omega per component, and
    the sweep's row order."""
    m = case.modes
    nb, nptk, ns = m.nbands, m.nptk, m.nsymm_rot
    n = nb * nptk
    ideq = np.asarray(m.ID_equi).reshape((ns, nptk), order="F")
    alleq = np.asarray(m.ALLEquiList).reshape((ns, nptk), order="F")
    types = np.asarray(m.TypeofSymmetry).reshape((ns, nptk), order="F")
    tau = np.asarray(tau_zero).reshape((nb, m.Nlist), order="F")
    B = np.zeros((n, n))
    b = np.zeros((n, 3))
    order = []
    # (sign of phonon 2, 3, 4) per list type: 3ph +, 3ph -, ++, +-, --
    signs = [(-1.0, 1.0, None), (0.5, 0.5, None), (-1.0, -1.0, 1.0), (-1.0, 1.0, 1.0), (1.0, 1.0, 1.0)]
    counts, entries, gams = [], [], []
    for t in range(5):
        if t < 2:
            if N3 is None:
                counts.append(np.zeros(nb * m.Nlist, np.int64))
                entries.append([np.zeros(0, np.int64)] * 2)
                gams.append(np.zeros(0))
                continue
            counts.append(np.asarray(N3[t], np.int64))
            entries.append([np.asarray(x, np.int64) for x in lists3[t]])
            gams.append(np.asarray(gammas3[t]))
        else:
            counts.append(np.asarray(N4[t - 2], np.int64))
            entries.append([np.asarray(x, np.int64) for x in lists4[t - 2]])
            gams.append(np.asarray(gammas4[t - 2]))
    offs = [np.concatenate([[0], np.cumsum(c)]) for c in counts]
    vel = np.asarray(m.velocity)
    en = np.asarray(m.energy)
    for ll in range(m.Nlist):
        for i in range(nb):
            mm = ll * nb + i
            for kk in range(int(m.Nequi[ll])):
                q = alleq[kk, ll] - 1
                s = types[kk, ll] - 1
                r = i + nb * q
                order.append(r)
                b[r] = tau[i, ll] * vel[q, i, :] * en[q, i]
                for t in range(5):
                    lo, hi = offs[t][mm], offs[t][mm + 1]
                    if hi == lo:
                        continue
                    g = gams[t][lo:hi] * tau[i, ll]
                    for slot, enc in enumerate(entries[t]):
                        e0 = enc[lo:hi] - 1
                        cols = e0 % nb + nb * (ideq[s, e0 // nb] - 1)
                        np.add.at(B[r], cols, signs[t][slot] * g)
    return B, b, np.array(order)


def gauss_seidel_sweep(case, N4, lists4, gammas4, N3, lists3, gammas3, tau_zero, F_n):
    """One sweep with identity symmetrisers, by a dense triangular solve."""
    m = case.modes
    B, b, order = _system(case, N4, lists4, gammas4, N3, lists3, gammas3, tau_zero)
    n = B.shape[0]
    assert np.array_equal(np.sort(order), np.arange(n)), "stars must partition the grid"
    pos = np.empty(n, np.int64)
    pos[order] = np.arange(n)
    earlier = pos[None, :] < pos[:, None]          # column visited before the row
    L = np.where(earlier, B, 0.0)
    U = B - L
    x_old = np.asarray(F_n).reshape((n, 3), order="F")
    x_new = np.linalg.solve(np.eye(n) - L, b + U @ x_old)
    return np.asfortranarray(x_new.reshape((m.nbands, m.nptk, 3), order="F"))


def linear_solution(case, N4, lists4, gammas4, N3, lists3, gammas3, tau_zero):
    """The fixed point of the sweep (identity symmetrisers)."""
    m = case.modes
    B, b, _ = _system(case, N4, lists4, gammas4, N3, lists3, gammas3, tau_zero)
    x = np.linalg.solve(np.eye(B.shape[0]) - B, b)
    return np.asfortranarray(x.reshape((m.nbands, m.nptk, 3), order="F"))
