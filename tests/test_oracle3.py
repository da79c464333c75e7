"""Pins the three-phonon part of the CPU oracle.

No reference vector exists for 3ph rates on these synthetic crystals, so each
function is pinned by a check that a plausible slip would fail:

* counts and lists: an independent vectorised restatement over the whole
  (q', j, k) candidate cube, in both loop orders the reference uses (ii outer
  for NP_*/RTA_*, j outer for Ind_only_*/Ind_*);
* matrix elements: a one-term closed form, lattice-translation covariance,
  the time-reversal map between Vp_plus and Vp_minus, and the acoustic sum
  rule at a Gamma acoustic phonon;
* rates: per-process Gamma summed per mode equals the RTA rate; the RTA rate
  equals a sum over the listed processes of the closed form
  hbarp*pi/4 * WP3 * |Vp|^2 evaluated here with numpy Bose factors and the
  oracle's own matrix elements; q' windows partition the rate.
"""
import copy

import numpy as np
import pytest

from fourphonon_amd import crystal
from oracle.oracle import Oracle

HBAR, KB = 1.05457172647e-22, 1.380648813e-23
HBARP = HBAR * float(np.float32(1e22))
UNITS3 = float(np.float32(5.60626442))


def vectorised_list3(case, ch, mm, j_outer):
    m = case.modes
    nb, nptk = m.nbands, m.nptk
    ng = np.asarray(m.ngrid)
    i = (mm - 1) % nb
    q0 = int(m.List[(mm - 1) // nb]) - 1
    omega = m.energy[q0, i]
    if omega == 0.0:
        return np.zeros((2, 0), np.int64)
    ijk = m.IJK.T.astype(np.int64)
    s = 1 if ch == 0 else -1
    qdp = (ijk[q0][None, :] + s * ijk) % ng
    ss = qdp[:, 0] + ng[0] * (qdp[:, 1] + ng[1] * qdp[:, 2])          # (nptk,)
    E, V = m.energy, m.velocity
    ii = np.arange(nptk)[:, None, None]
    j = np.arange(nb)[None, :, None]
    k = np.arange(nb)[None, None, :]
    shape = (nptk, nb, nb)
    ii3 = np.broadcast_to(ii, shape)
    j3 = np.broadcast_to(j, shape)
    k3 = np.broadcast_to(k, shape)
    ss3 = np.broadcast_to(ss[:, None, None], shape)
    op, odp = E[ii3, j3], E[ss3, k3]
    ok = (op != 0) & (odp != 0)
    dv = V[ii3, j3, :] - V[ss3, k3, :]
    b = case.structure.rlattvec
    comps = []
    for nu in range(3):
        d = b[0, nu] * dv[..., 0] + b[1, nu] * dv[..., 1] + b[2, nu] * dv[..., 2]
        t = d / float(ng[nu])
        comps.append(t * t)
    sig = case.scalebroad * np.sqrt(np.maximum(np.maximum(comps[0], comps[1]), comps[2]) / 2.0)
    delta = (omega + op - odp) if ch == 0 else (omega - op - odp)
    ok &= np.abs(delta) <= 2.0 * sig
    if j_outer:
        ok = ok.transpose(1, 0, 2)              # (j, ii, k)
        jv, iv, kv = np.nonzero(ok)
    else:
        iv, jv, kv = np.nonzero(ok)
    return np.stack([iv * nb + jv + 1, ss[iv] * nb + kv + 1]).astype(np.int64)


@pytest.fixture(scope="module")
def case3():
    return crystal.config_case(1, ngrid=(4, 4, 4))


@pytest.fixture(scope="module")
def orc3(case3):
    return Oracle(case3)


@pytest.mark.parametrize("which", ["diamond333", "diamond234", "wurtzite332"])
def test_lists_match_vectorised_enumeration(which):
    if which == "diamond333":
        case = crystal.config_case(1, ngrid=(3, 3, 3))
    elif which == "diamond234":
        case = crystal.config_case(1, ngrid=(2, 3, 4), symmetric=False)
    else:
        case = crystal.config_case(3, ngrid=(3, 3, 2), shell=1)
    o = Oracle(case)
    for ch in range(2):
        for mm in range(1, o.nmodes + 1):
            ref = vectorised_list3(case, ch, mm, j_outer=True)
            got = np.stack(o.ind_only3_mode(ch, mm)).astype(np.int64)
            assert np.array_equal(got, ref), (ch, mm)
            assert o.np3_mode(ch, mm)[0] == ref.shape[1]


def test_counts_and_lists_drivers(case3, orc3):
    o = orc3
    N, P = o.np3_driver()
    lists = o.ind_only3_driver(N)
    for ch in range(2):
        acc = np.concatenate([[0], np.cumsum(N[ch].astype(np.int64))])
        assert acc[-1] > 0
        for mm in (4, 9, 17, 40, o.nmodes):
            seg = np.stack([x[acc[mm - 1]:acc[mm]] for x in lists[ch]]).astype(np.int64)
            assert np.array_equal(seg, vectorised_list3(case3, ch, mm, j_outer=True)), (ch, mm)
    # P3: the Gaussian of each listed process over (sigma sqrt(pi) nptk^2 Nb^3)
    m = case3.modes
    for ch in range(2):
        for mm in (5, 23):
            v = vectorised_list3(case3, ch, mm, j_outer=False)
            q0 = int(m.List[(mm - 1) // m.nbands]) - 1
            w = m.energy[q0, (mm - 1) % m.nbands]
            ii, j = (v[0] - 1) // m.nbands, (v[0] - 1) % m.nbands
            ss, k = (v[1] - 1) // m.nbands, (v[1] - 1) % m.nbands
            op, odp = m.energy[ii, j], m.energy[ss, k]
            d = (w + op - odp) if ch == 0 else (w - op - odp)
            sig = np.array([o.base_sigma(m.velocity[a, b] - m.velocity[c, e])
                            for a, b, c, e in zip(ii, j, ss, k)]) * case3.scalebroad
            want = np.sum(np.exp(-d ** 2 / sig ** 2) / (sig * np.sqrt(np.pi) * m.nptk ** 2 * m.nbands ** 3))
            assert np.isclose(P[ch][mm - 1], want, rtol=1e-12), (ch, mm)


def _qdp(m, q, qp, s):
    g = (m.IJK[:, q - 1] + s * m.IJK[:, qp - 1]) % m.ngrid
    return int(g[0] + m.ngrid[0] * (g[1] + m.ngrid[1] * g[2])) + 1


def test_vp3_single_term_closed_form(case3):
    case = copy.deepcopy(case3)
    f = case.ifc3
    lat = case.structure.lattvec
    cells = np.array([[1, 0, -1], [0, 1, 1]])
    f.Ntri = 1
    f.Phi = np.zeros((3, 3, 3, 1), order="F")
    f.Phi[2, 0, 1, 0] = -1.75                           # (t,s,r) = (3,1,2)
    f.R_j = np.asfortranarray((lat @ cells[0])[:, None])
    f.R_k = np.asfortranarray((lat @ cells[1])[:, None])
    f.Index_i = np.array([2], np.int32)
    f.Index_j = np.array([1], np.int32)
    f.Index_k = np.array([2], np.int32)
    o = Oracle(case)
    m = case.modes
    ev = m.eigenvect
    mass = case.structure.masses[0]
    q, qp = 7, 30
    i, j, k = 2, 5, 4
    qf = lambda x: m.IJK[:, x - 1] / m.ngrid
    for ch, s in ((0, 1), (1, -1)):
        qdp = _qdp(m, q, qp, s)
        e2 = ev[qp - 1, j - 1, 0]
        e2 = e2 if ch == 0 else np.conj(e2)
        phase = 2 * np.pi * (s * qf(qp) @ cells[0] - qf(qdp) @ cells[1])
        want = -1.75 / mass ** 1.5 * np.exp(1j * phase) * ev[q - 1, i - 1, 3 + 2] * e2 * \
            np.conj(ev[qdp - 1, k - 1, 3 + 1])
        got = o.vp3(ch, i, j, k, q, qp, qdp)
        assert abs(got - want) <= 1e-13 * abs(want)


def _random3(m, rng, n):
    out = []
    for _ in range(n):
        q = int(m.List[rng.integers(m.Nlist)])
        qp = int(rng.integers(1, m.nptk + 1))
        out.append((q, qp, *(int(x) for x in rng.integers(1, m.nbands + 1, size=3))))
    return out


@pytest.fixture(scope="module")
def shell2_case3():
    return crystal.config_case(1, ngrid=(4, 4, 4), shell=2)


def test_vp3_lattice_translation_covariance(shell2_case3):
    """Shifting R_j and R_k by T multiplies V by exp(i (s q' - q'') . T)."""
    base = Oracle(shell2_case3)
    case = copy.deepcopy(shell2_case3)
    T = np.array([-1, 2, 1])
    shift = (case.structure.lattvec @ T)[:, None]
    case.ifc3.R_j = np.asfortranarray(case.ifc3.R_j + shift)
    case.ifc3.R_k = np.asfortranarray(case.ifc3.R_k + shift)
    moved = Oracle(case)
    m = case.modes
    qf = lambda x: m.IJK[:, x - 1] / m.ngrid
    rng = np.random.default_rng(13)
    for ch, s in ((0, 1), (1, -1)):
        got, want = [], []
        for q, qp, i, j, k in _random3(m, rng, 8):
            qdp = _qdp(m, q, qp, s)
            v0 = base.vp3(ch, i, j, k, q, qp, qdp)
            got.append(moved.vp3(ch, i, j, k, q, qp, qdp))
            want.append(np.exp(2j * np.pi * (s * qf(qp) - qf(qdp)) @ T) * v0)
        got, want = np.array(got), np.array(want)
        assert np.abs(want).max() > 0
        assert np.abs(got - want).max() <= 1e-11 * np.abs(want).max()


def test_vp3_time_reversal_map(shell2_case3):
    """With e(-q) = conj(e(q)): Vp_minus(q, q', q'') = Vp_plus(q, -q', q'')."""
    case = copy.deepcopy(shell2_case3)
    m = case.modes
    ev = m.eigenvect.copy()
    for q in range(1, m.nptk + 1):
        g = (-m.IJK[:, q - 1]) % m.ngrid
        mq = int(g[0] + m.ngrid[0] * (g[1] + m.ngrid[1] * g[2])) + 1
        if mq > q:
            ev[mq - 1] = np.conj(ev[q - 1])
    m.eigenvect = np.asfortranarray(ev)
    o = Oracle(case)
    rng = np.random.default_rng(17)
    a, b = [], []
    for q, qp, i, j, k in _random3(m, rng, 14):
        qdp = _qdp(m, q, qp, -1)
        g = (-m.IJK[:, qp - 1]) % m.ngrid
        mqp = int(g[0] + m.ngrid[0] * (g[1] + m.ngrid[1] * g[2])) + 1
        if mqp == qp:
            continue        # -q' = q': e(q') is not made real here, so the map does not apply
        a.append(o.vp3(1, i, j, k, q, qp, qdp))
        b.append(o.vp3(0, i, j, k, q, mqp, qdp))
    a, b = np.array(a), np.array(b)
    assert np.abs(a).max() > 0
    assert np.abs(a - b).max() <= 1e-11 * np.abs(a).max()


def test_vp3_vanishes_for_gamma_acoustic_phonon(shell2_case3):
    o = Oracle(shell2_case3)
    m = shell2_case3.modes
    rng = np.random.default_rng(19)
    for q, qp, i, j, k in _random3(m, rng, 4):
        typical = abs(o.vp3(0, i, j, k, q, qp, _qdp(m, q, qp, 1)))
        for jg in (1, 2, 3):
            v = o.vp3(0, i, jg, k, q, 1, _qdp(m, q, 1, 1))
            assert abs(v) <= 1e-9 * max(typical, 1e-6)


def test_rates_closed_form_and_ind_consistency(case3, orc3):
    o = orc3
    m = case3.modes
    N, _ = o.np3_driver()
    lists, gam, rind = o.ind3_driver(N)
    modes = np.arange(1, o.nmodes + 1, dtype=np.int32)[2::7]
    rta = o.rta3_driver(modes=modes)
    for key in ("rate_plus", "rate_minus", "wp3_plus", "wp3_minus", "rate"):
        assert np.allclose(rta[key][modes - 1], rind[key][modes - 1], rtol=1e-12, atol=0), key
    T = case3.T
    bose = lambda w: 1.0 / np.expm1(HBAR * w / KB / T)
    for mm in modes[:5]:
        q0 = int(m.List[(mm - 1) // m.nbands])
        i = (mm - 1) % m.nbands + 1
        w = m.energy[q0 - 1, i - 1]
        if w == 0:
            continue
        for ch in range(2):
            v = vectorised_list3(case3, ch, int(mm), j_outer=False)
            gsum, wsum = 0.0, 0.0
            for a2, a3 in v.T:
                ii, j = (a2 - 1) // m.nbands + 1, (a2 - 1) % m.nbands + 1
                ss, k = (a3 - 1) // m.nbands + 1, (a3 - 1) % m.nbands + 1
                op, odp = m.energy[ii - 1, j - 1], m.energy[ss - 1, k - 1]
                d = (w + op - odp) if ch == 0 else (w - op - odp)
                sig = case3.scalebroad * o.base_sigma(m.velocity[ii - 1, j - 1] - m.velocity[ss - 1, k - 1])
                occ = bose(op) - bose(odp) if ch == 0 else bose(op) + bose(odp) + 1
                wp3 = occ * np.exp(-d ** 2 / sig ** 2) / (sig * np.sqrt(np.pi) * w * op * odp)
                wsum += wp3
                gsum += HBARP * np.pi / 4 * wp3 * abs(o.vp3(ch, i, j, k, q0, ii, ss)) ** 2
            half = 1.0 if ch == 0 else 0.5
            key = "rate_plus" if ch == 0 else "rate_minus"
            wkey = "wp3_plus" if ch == 0 else "wp3_minus"
            assert np.isclose(rta[key][mm - 1], gsum * UNITS3 * 1e8 / m.nptk * half, rtol=1e-10), (ch, mm)
            assert np.isclose(rta[wkey][mm - 1], wsum * half / m.nptk, rtol=1e-10), (ch, mm)


def test_qprime_window_partitions_the_rate(orc3):
    o = orc3
    modes = np.array([11, 30], np.int32)
    full = o.rta3_driver(modes=modes)
    parts = [o.rta3_driver(modes=modes, ii_lo=lo, ii_hi=hi) for lo, hi in ((1, 17), (18, 40), (41, o.nptk))]
    for key in ("rate_plus", "rate_minus", "wp3_plus"):
        s = sum(p[key][modes - 1] for p in parts)
        assert np.allclose(s, full[key][modes - 1], rtol=1e-12), key


def test_onlyharmonic_and_omega_max(case3):
    c = copy.copy(case3)
    c.onlyharmonic = True
    o = Oracle(c)
    r = o.rta3_driver(modes=[8, 9])
    assert np.all(r["rate"][7:9] == 0) and np.all(r["wp3_plus"][7:9] + r["wp3_minus"][7:9] > 0)
    c2 = copy.copy(case3)
    c2.omega_max = 60.0
    o2 = Oracle(c2)
    N, _ = o2.np3_driver()
    w = case3.modes.energy[case3.modes.List - 1, :].ravel()
    hi = np.nonzero(w > 60.0)[0]
    assert hi.size > 0 and all(N[ch][hi].sum() == 0 for ch in range(2))
