"""Pins the CPU oracle before anything is compared against it.

The reference ships no test vectors for rates, matrix elements or kappa
(FORCE_CONSTANTS_4TH is stripped from every example), so each oracle
function is pinned by checks that a plausible mistake would fail:

* enumeration and lists: an independent vectorised restatement (masks over
  the full (q', q'', j, k, l) candidate cube instead of nested loops) must give
  the same per-mode lists, element for element;
* matrix elements: a one-term closed form, lattice-translation covariance,
  the ++ exchange symmetry, the time-reversal maps between ++, +- and --, and
  the acoustic sum rule;
* drivers: per-process Gamma summed per mode equals the RTA rate, counts
  equal list lengths, omega_max and onlyharmonic gate what they should;
* kappa: the constant-relaxation-time closed form.
"""
import copy

import numpy as np
import pytest

from fourphonon_amd import crystal
from oracle.oracle import Oracle, iteration0, iteration_4ph, tconduct
from tests.golden.listhash import list_hash


# ---------------------------------------------------------------------------
# vectorised enumeration (independent of the oracle's loop nest)
# ---------------------------------------------------------------------------

def vectorised_list(case, ch, mm):
    m = case.modes
    nb, nptk = m.nbands, m.nptk
    ng = np.asarray(m.ngrid)
    i = (mm - 1) % nb
    ll = (mm - 1) // nb
    q0 = int(m.List[ll]) - 1
    omega = m.energy[q0, i]
    if omega == 0.0:
        return np.zeros((3, 0), np.int64)
    ijk = m.IJK.T.astype(np.int64)
    s2 = {0: 1, 1: 1, 2: -1}[ch]
    s3 = {0: 1, 1: -1, 2: -1}[ch]
    qt = (ijk[q0][None, None, :] + s2 * ijk[:, None, :] + s3 * ijk[None, :, :]) % ng
    ss = qt[..., 0] + ng[0] * (qt[..., 1] + ng[1] * qt[..., 2])        # 0-based (nptk,nptk)
    II, JJ = np.meshgrid(np.arange(nptk), np.arange(nptk), indexing="ij")
    pair_ok = np.ones((nptk, nptk), bool)
    if ch in (0, 2):
        pair_ok &= JJ >= II
    if ch in (1, 2):
        pair_ok &= ss >= JJ
    E = m.energy
    V = m.velocity
    j = np.arange(nb)[:, None, None]
    k = np.arange(nb)[None, :, None]
    l = np.arange(nb)[None, None, :]
    shape = (nptk, nptk, nb, nb, nb)
    ii5 = np.broadcast_to(II[:, :, None, None, None], shape)
    jj5 = np.broadcast_to(JJ[:, :, None, None, None], shape)
    ss5 = np.broadcast_to(ss[:, :, None, None, None], shape)
    j5 = np.broadcast_to(j[None, None], shape)
    k5 = np.broadcast_to(k[None, None], shape)
    l5 = np.broadcast_to(l[None, None], shape)
    ok = np.broadcast_to(pair_ok[:, :, None, None, None], shape).copy()
    if ch in (0, 2):
        ok &= ~((ii5 == jj5) & (k5 < j5))
    if ch in (1, 2):
        ok &= ~((ss5 == jj5) & (l5 < k5))
    op, odp, otp = E[ii5, j5], E[jj5, k5], E[ss5, l5]
    ok &= (op != 0) & (odp != 0) & (otp != 0)
    v3 = V[jj5, k5, :]
    v4 = V[ss5, l5, :]
    dv = (-v3 + v4) if ch == 0 else (v3 - v4)
    b = case.structure.rlattvec
    comps = []
    for nu in range(3):
        d = b[0, nu] * dv[..., 0] + b[1, nu] * dv[..., 1] + b[2, nu] * dv[..., 2]
        t = d / float(ng[nu])
        comps.append(t * t)
    sig = case.scalebroad * np.sqrt(np.maximum(np.maximum(comps[0], comps[1]), comps[2]) / 2.0)
    if ch == 1:
        qi = q0
        excl = ((qi == jj5) & (i == k5)) | ((qi == ss5) & (i == l5)) | ((ii5 == jj5) & (j5 == k5)) | \
               ((ii5 == ss5) & (j5 == l5))
        ok &= ~excl
    if ch == 0:
        delta = omega + op + odp - otp
    elif ch == 1:
        delta = omega + op - odp - otp
    else:
        delta = omega - op - odp - otp
    ok &= np.abs(delta) <= sig
    idx = np.nonzero(ok)            # C order == (ii, jj, j, k, l) lexicographic
    a, bb, jv, kv, lv = idx
    return np.stack([a * nb + jv + 1, bb * nb + kv + 1, ss[a, bb] * nb + lv + 1]).astype(np.int64)


@pytest.mark.parametrize("ngrid", [(3, 3, 3), (2, 3, 4)])
def test_lists_match_vectorised_enumeration_all_modes(ngrid):
    case = crystal.config_case(1, ngrid=ngrid, symmetric=(ngrid[0] == ngrid[1] == ngrid[2]))
    o = Oracle(case)
    for ch in range(3):
        for mm in range(1, o.nmodes + 1):
            ref = vectorised_list(case, ch, mm)
            got = np.stack(o.ind_only_mode(ch, mm))
            assert got.shape == ref.shape, (ch, mm)
            assert np.array_equal(got, ref), (ch, mm)
            assert o.np_mode(ch, mm)[0] == ref.shape[1]
            # the streamed hash (golden fixtures of whole meshes use it)
            n, p, h = o.list_hash_mode(ch, mm)
            assert (n, p) == o.np_mode(ch, mm)
            assert h == list_hash(*ref), (ch, mm)


def test_lists_match_vectorised_enumeration_4cube(small_case, small_oracle):
    for ch in range(3):
        for mm in (1, 4, 7, 12, 20, 31, small_oracle.nmodes):
            ref = vectorised_list(small_case, ch, mm)
            got = np.stack(small_oracle.ind_only_mode(ch, mm))
            assert np.array_equal(got, ref), (ch, mm)


def test_wurtzite_lists_match_vectorised():
    case = crystal.config_case(3, ngrid=(3, 3, 2), shell=1)
    o = Oracle(case)
    for ch in range(3):
        for mm in (13, 20, o.nmodes):
            assert np.array_equal(np.stack(o.ind_only_mode(ch, mm)), vectorised_list(case, ch, mm))


# ---------------------------------------------------------------------------
# matrix elements
# ---------------------------------------------------------------------------

def _minus_q(m, q):
    ijk = (-m.IJK[:, q - 1]) % m.ngrid
    return int(ijk[0] + m.ngrid[0] * (ijk[1] + m.ngrid[1] * ijk[2])) + 1


def _qt(m, q, qp, qdp, s2, s3):
    g = (m.IJK[:, q - 1] + s2 * m.IJK[:, qp - 1] + s3 * m.IJK[:, qdp - 1]) % m.ngrid
    return int(g[0] + m.ngrid[0] * (g[1] + m.ngrid[1] * g[2])) + 1


def test_vp_single_term_closed_form(small_case):
    """One cluster, one non-zero IFC element: V is a single product."""
    case = copy.deepcopy(small_case)
    f = case.ifc4
    lat = case.structure.lattvec
    cells = np.array([[1, 0, 0], [0, -1, 1], [1, 1, 0]])
    f.Ntri = 1
    f.Phi = np.zeros((3, 3, 3, 3, 1), order="F")
    f.Phi[0, 1, 2, 1, 0] = 2.5                         # (u,t,s,r) = (1,2,3,2)
    f.R_j = np.asfortranarray((lat @ cells[0])[:, None])
    f.R_k = np.asfortranarray((lat @ cells[1])[:, None])
    f.R_l = np.asfortranarray((lat @ cells[2])[:, None])
    f.Index_i = np.array([1], np.int32)
    f.Index_j = np.array([2], np.int32)
    f.Index_k = np.array([2], np.int32)
    f.Index_l = np.array([1], np.int32)
    o = Oracle(case)
    m = case.modes
    ev = m.eigenvect
    mass = case.structure.masses[0]
    q, qp, qdp = 5, 22, 41
    for ch, (s2, s3, c2, c3) in enumerate([(1, 1, False, False), (1, -1, False, True), (-1, -1, True, True)]):
        qt = _qt(m, q, qp, qdp, s2, s3)
        i, j, k, l = 2, 4, 5, 6
        e2 = ev[qp - 1, j - 1, 3 + 1]
        e3 = ev[qdp - 1, k - 1, 3 + 2]
        e2 = np.conj(e2) if c2 else e2
        e3 = np.conj(e3) if c3 else e3
        qf = lambda x: m.IJK[:, x - 1] / m.ngrid
        phase = 2 * np.pi * (s2 * qf(qp) @ cells[0] + s3 * qf(qdp) @ cells[1] - qf(qt) @ cells[2])
        want = 2.5 / mass ** 2 * np.exp(1j * phase) * ev[q - 1, i - 1, 0] * e2 * e3 * np.conj(ev[qt - 1, l - 1, 1])
        got = o.vp(ch, i, j, k, l, q, qp, qdp, qt)
        assert abs(got - want) <= 1e-13 * abs(want)


@pytest.fixture(scope="module")
def shell2_case():
    return crystal.config_case(1, ngrid=(4, 4, 4), shell=2)


def _random_processes(m, rng, n):
    out = []
    for _ in range(n):
        q = int(m.List[rng.integers(m.Nlist)])
        qp, qdp = (int(x) for x in rng.integers(1, m.nptk + 1, size=2))
        out.append((q, qp, qdp, *(int(x) for x in rng.integers(1, m.nbands + 1, size=4))))
    return out


def test_vp_lattice_translation_covariance(shell2_case):
    """Shifting R_j, R_k, R_l by one lattice vector T multiplies V by
    exp(i(s2 q' + s3 q'' - q''').T), a pure phase fixed by the channel's signs."""
    base = Oracle(shell2_case)
    case = copy.deepcopy(shell2_case)
    lat = case.structure.lattvec
    T = np.array([1, -2, 1])
    shift = (lat @ T)[:, None]
    case.ifc4.R_j = np.asfortranarray(case.ifc4.R_j + shift)
    case.ifc4.R_k = np.asfortranarray(case.ifc4.R_k + shift)
    case.ifc4.R_l = np.asfortranarray(case.ifc4.R_l + shift)
    moved = Oracle(case)
    m = case.modes
    rng = np.random.default_rng(7)
    for ch, (s2, s3) in enumerate([(1, 1), (1, -1), (-1, -1)]):
        got, want = [], []
        for q, qp, qdp, i, j, k, l in _random_processes(m, rng, 6):
            qt = _qt(m, q, qp, qdp, s2, s3)
            v0 = base.vp(ch, i, j, k, l, q, qp, qdp, qt)
            got.append(moved.vp(ch, i, j, k, l, q, qp, qdp, qt))
            qf = lambda x: m.IJK[:, x - 1] / m.ngrid
            want.append(np.exp(2j * np.pi * (s2 * qf(qp) + s3 * qf(qdp) - qf(qt)) @ T) * v0)
        got, want = np.array(got), np.array(want)
        assert np.abs(want).max() > 0
        assert np.abs(got - want).max() <= 1e-11 * np.abs(want).max()


def test_vp_plusplus_exchange_symmetry(shell2_case):
    o = Oracle(shell2_case)
    m = shell2_case.modes
    rng = np.random.default_rng(3)
    a, b = [], []
    for q, qp, qdp, i, j, k, l in _random_processes(m, rng, 8):
        qt = _qt(m, q, qp, qdp, 1, 1)
        a.append(o.vp(0, i, j, k, l, q, qp, qdp, qt))
        b.append(o.vp(0, i, k, j, l, q, qdp, qp, qt))
    a, b = np.array(a), np.array(b)
    assert np.abs(a - b).max() <= 1e-11 * np.abs(a).max()


def _time_reversal_case(case):
    """Eigenvectors with e(-q) = conj(e(q)) exactly (a valid choice, since
    D(-q) = conj(D(q)))."""
    c = copy.deepcopy(case)
    m = c.modes
    ev = m.eigenvect.copy()
    for q in range(1, m.nptk + 1):
        mq = _minus_q(m, q)
        if mq > q:
            ev[mq - 1] = np.conj(ev[q - 1])
    m.eigenvect = np.asfortranarray(ev)
    return c


def test_vp_time_reversal_maps_between_channels(shell2_case):
    """V_pm(q,q',q'',q''') = V_pp(q,q',-q'',q''') and
    V_mm(q,q',q'',q''') = V_pp(q,-q',-q'',q''') when e(-q) = conj(e(q))."""
    case = _time_reversal_case(shell2_case)
    o = Oracle(case)
    m = case.modes
    rng = np.random.default_rng(11)
    pm, pp1, mmv, pp2 = [], [], [], []
    for q, qp, qdp, i, j, k, l in _random_processes(m, rng, 8):
        qt = _qt(m, q, qp, qdp, 1, -1)
        pm.append(o.vp(1, i, j, k, l, q, qp, qdp, qt))
        pp1.append(o.vp(0, i, j, k, l, q, qp, _minus_q(m, qdp), qt))
        qt = _qt(m, q, qp, qdp, -1, -1)
        mmv.append(o.vp(2, i, j, k, l, q, qp, qdp, qt))
        pp2.append(o.vp(0, i, j, k, l, q, _minus_q(m, qp), _minus_q(m, qdp), qt))
    for x, y in ((pm, pp1), (mmv, pp2)):
        x, y = np.array(x), np.array(y)
        assert np.abs(x).max() > 0
        assert np.abs(x - y).max() <= 1e-11 * np.abs(x).max()


def test_vp_vanishes_for_gamma_acoustic_phonon(shell2_case):
    o = Oracle(shell2_case)
    m = shell2_case.modes
    rng = np.random.default_rng(5)
    for q, qp, qdp, i, j, k, l in _random_processes(m, rng, 4):
        typical = abs(o.vp(0, i, j, k, l, q, qp, qdp, _qt(m, q, qp, qdp, 1, 1)))
        for jg in (1, 2, 3):
            # Gamma acoustic mode as the second phonon
            v = o.vp(0, i, jg, k, l, q, 1, qdp, _qt(m, q, 1, qdp, 1, 1))
            assert abs(v) <= 1e-9 * max(typical, 1e-6)


# ---------------------------------------------------------------------------
# drivers
# ---------------------------------------------------------------------------

def test_driver_consistency(small_case, small_oracle):
    o = small_oracle
    N, P = o.np_driver()
    lists = o.ind_only_driver(N)
    for ch in range(3):
        acc = np.concatenate([[0], np.cumsum(N[ch])])
        for mm in (4, 7, 15, 33):
            seg = [x[acc[mm - 1]:acc[mm]] for x in lists[ch]]
            ref = o.ind_only_mode(ch, mm)
            assert all(np.array_equal(a, b) for a, b in zip(seg, ref))
    ind_lists, gam, rates_ind = o.ind_driver(N)
    for ch in range(3):
        assert all(np.array_equal(a, b) for a, b in zip(ind_lists[ch], lists[ch]))
    modes = np.arange(1, o.nmodes + 1, dtype=np.int32)[::5]
    rta = o.rta_driver(modes=modes)
    for key in ("rate_pp", "rate_pm", "rate_mm", "wp4_pp", "wp4_pm", "wp4_mm"):
        a, b = rta[key][modes - 1], rates_ind[key][modes - 1]
        assert np.allclose(a, b, rtol=1e-12, atol=0), key
    assert np.all(rta["rate_4ph"][modes - 1] > 0) or np.any(rta["rate_4ph"][modes - 1] == 0)


def test_qprime_window_partitions_the_rate(small_oracle):
    o = small_oracle
    mm = 11
    full = [o.rta_mode(ch, mm) for ch in range(3)]
    parts = [[o.rta_mode(ch, mm, lo, hi) for (lo, hi) in ((1, 20), (21, 45), (46, o.nptk))] for ch in range(3)]
    for ch in range(3):
        assert np.isclose(sum(p[0] for p in parts[ch]), full[ch][0], rtol=1e-12)


def test_onlyharmonic_and_omega_max(small_case):
    case = copy.copy(small_case)
    case.onlyharmonic = True
    o = Oracle(case)
    r = o.rta_driver(modes=[8, 9])
    assert np.all(r["rate_4ph"][7:9] == 0) and np.all(r["wp4_pm"][7:9] > 0)
    case2 = copy.copy(small_case)
    case2.omega_max = 60.0
    o2 = Oracle(case2)
    N, _ = o2.np_driver()
    w = small_case.modes.energy[small_case.modes.List - 1, :].ravel()   # mm order: (ll, i)
    hi = np.nonzero(w > 60.0)[0]
    assert hi.size > 0 and all(N[ch][hi].sum() == 0 for ch in range(3))


# ---------------------------------------------------------------------------
# kappa and the iterative update
# ---------------------------------------------------------------------------

def test_kappa_constant_tau_closed_form(small_case):
    m = small_case.modes
    tau = 3.0
    tz = np.full((m.nbands, m.Nlist), tau, order="F")
    F = iteration0(small_case, tz)
    kap = tconduct(small_case, F)
    hbar, kb = 1.05457172647e-22, 1.380648813e-23
    T, V = small_case.T, small_case.structure.volume
    w = m.energy[1:]                                             # q = 1 is skipped
    v = m.velocity[1:]
    x = hbar * w / kb / T
    f = 1.0 / np.expm1(x)
    want = np.einsum("qb,qba,qbc->bac", f * (f + 1) * w * w * tau, v, v)
    want *= float(np.float32(1e21)) * hbar ** 2 / (kb * T * T * V * m.nptk)
    assert np.allclose(kap, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
    tot = kap.sum(axis=0)
    assert np.allclose(tot, tot.T, rtol=1e-10, atol=1e-10 * abs(tot).max())


def test_iteration_without_scattering_is_a_fixed_point(small_case, small_oracle):
    m = small_case.modes
    N = [np.zeros(small_oracle.nmodes, np.int64) for _ in range(3)]
    lists = [[np.zeros(0, np.int64)] * 3 for _ in range(3)]
    gam = [np.zeros(0)] * 3
    tz = np.full((m.nbands, m.Nlist), 2.0, order="F")
    F0 = iteration0(small_case, tz)
    F1 = iteration_4ph(small_case, N, lists, gam, tz, F0)
    sym = np.einsum("abq,nqb->nqa", m.symmetrizers, F0)
    assert np.allclose(F1, sym, rtol=1e-13, atol=1e-13 * abs(F0).max())
