"""Three-phonon path on the device against the CPU oracle, through the C ABI.

Bars: counts and lists bit-exact; phase space within 1e-12 relative; FP64
rates, WP3, per-process Gamma and matrix elements within 1e-10 relative
(summation order and exact root-of-unity phases are the only differences).
"""
import numpy as np
import pytest

from fourphonon_amd import crystal
from fourphonon_amd import runtime as rt

pytestmark = pytest.mark.gpu
RTOL = 1e-10
KEYS = (("rate_scatt_plus", "rate_plus"), ("rate_scatt_minus", "rate_minus"), ("rate_scatt", "rate"),
        ("WP3_plus", "wp3_plus"), ("WP3_minus", "wp3_minus"))


def banner(msg):
    print(f"\n[gpu] {msg}", flush=True)


def assert_close(got, want, rtol=RTOL, what=""):
    got, want = np.asarray(got, float).ravel(), np.asarray(want, float).ravel()
    scale = max(np.abs(want).max(), 1e-300)
    err = np.abs(got - want)
    bad = err > rtol * np.maximum(np.abs(want), 1e-3 * scale)
    assert not bad.any(), f"{what}: max rel err {np.max(err / np.maximum(np.abs(want), 1e-300)):.3e}"


CASES = {
    "diamond444": lambda: crystal.config_case(1, ngrid=(4, 4, 4)),
    "diamond8cube": lambda: crystal.config_case(1),
    "wurtzite332": lambda: crystal.config_case(3, ngrid=(3, 3, 2), shell=1),
    "batch7_nb24": lambda: crystal.batch_case(7, ngrid=(3, 3, 3)),
}


@pytest.fixture(scope="module", params=list(CASES))
def trio(request):
    from oracle.oracle import Oracle
    case = CASES[request.param]()
    banner(f"3ph context {request.param} Nb {case.nbands}")
    c = rt.Context(case)
    o = Oracle(case)
    No, Po = o.np3_driver(nthreads=8)
    yield request.param, case, c, o, No, Po
    c.close()


def test_counts_and_lists_bitexact(trio):
    name, case, c, o, No, Po = trio
    banner(f"NP_driver / Ind_only_driver {name}")
    N = c.NP_driver()
    for ch in range(2):
        assert np.array_equal(N[2 * ch], No[ch]), f"counts channel {ch}"
        assert_close(N[2 * ch + 1].ravel(order="F"), Po[ch], 1e-12, f"P channel {ch}")
    assert sum(int(n.sum()) for n in No) > 0
    want = o.ind_only3_driver(No, nthreads=8)
    got = c.Ind_only_driver(*No)
    for ch in range(2):
        for t in range(2):
            assert np.array_equal(got[ch][t], want[ch][t]), (ch, t)
    c.set_batch_capacity(1024)
    try:
        got2 = c.Ind_only_driver(*No)
    finally:
        c.set_batch_capacity(1 << 26)
    for ch in range(2):
        for t in range(2):
            assert np.array_equal(got2[ch][t], want[ch][t]), (ch, t)


def test_rta_rates(trio):
    name, case, c, o, No, Po = trio
    banner(f"RTA_driver {name}")
    nm = c.nmodes
    modes = np.arange(1, nm + 1) if nm <= 400 else np.arange(1, nm + 1)[::9]
    want = o.rta3_driver(modes=modes, nthreads=8)
    if len(modes) == nm:
        got = c.RTA_driver()
    else:
        got, _ = c.RTA_driver_modes(modes)
    for g, w in KEYS:
        assert_close(got[g].ravel(order="F")[modes - 1], want[w][modes - 1], RTOL, g)
    assert np.abs(want["rate"][modes - 1]).max() > 0
    # a q' window and a small batch capacity change nothing but the summation split
    lo, hi = 2, max(2, c.nptk // 3)
    wwin = o.rta3_driver(modes=modes[:6], ii_lo=lo, ii_hi=hi, nthreads=8)
    c.set_batch_capacity(1024)
    try:
        gwin, n = c.RTA_driver_modes(modes[:6], ii_lo=lo, ii_hi=hi)
    finally:
        c.set_batch_capacity(1 << 26)
    for g, w in KEYS:
        assert_close(gwin[g].ravel(order="F")[modes[:6] - 1], wwin[w][modes[:6] - 1], RTOL, g + " window")


def test_ind_driver_and_using_ind(trio):
    name, case, c, o, No, Po = trio
    if c.nmodes > 400:
        pytest.skip("covered on the smaller meshes")
    banner(f"Ind_driver / RTA_driver_GPU_using_Ind {name}")
    lists, gam, rates = c.Ind_driver(*No)
    wl, wg, wr = o.ind3_driver(No, nthreads=8)
    for ch in range(2):
        for t in range(2):
            assert np.array_equal(lists[ch][t], wl[ch][t])
        assert_close(gam[ch], wg[ch], RTOL, f"Gamma channel {ch}")
    for g, w in KEYS:
        assert_close(rates[g].ravel(order="F"), wr[w], RTOL, g)
    base = c.RTA_driver()
    via = c.RTA_driver_GPU_using_Ind(*No, lists)
    for k in base:
        assert np.array_equal(base[k], via[k]), k


def test_vp3_matches_oracle(trio):
    name, case, c, o, No, Po = trio
    banner(f"Vp_plus / Vp_minus {name}")
    m = case.modes
    rng = np.random.default_rng(23)
    for ch, s in ((0, 1), (1, -1)):
        rows = []
        for _ in range(40):
            q = int(m.List[rng.integers(m.Nlist)])
            qp = int(rng.integers(1, m.nptk + 1))
            g = (m.IJK[:, q - 1] + s * m.IJK[:, qp - 1]) % m.ngrid
            qdp = int(g[0] + m.ngrid[0] * (g[1] + m.ngrid[1] * g[2])) + 1
            rows.append([*(int(x) for x in rng.integers(1, m.nbands + 1, size=3)), q, qp, qdp])
        rows = np.array(rows, np.int32)
        got = c.Vp3(ch, rows)
        want = np.array([o.vp3(ch, *r) for r in rows])
        scale = np.abs(want).max()
        assert scale > 0
        assert np.abs(got - want).max() <= 1e-12 * scale, ch


def test_onlyharmonic_and_omega_max():
    from oracle.oracle import Oracle
    banner("3ph onlyharmonic / omega_max")
    case = crystal.config_case(1, ngrid=(4, 4, 4))
    case.onlyharmonic = True
    case.omega_max = 70.0
    c = rt.Context(case)
    o = Oracle(case)
    try:
        N = c.NP_driver()
        No, _ = o.np3_driver()
        for ch in range(2):
            assert np.array_equal(N[2 * ch], No[ch])
        got = c.RTA_driver()
        want = o.rta3_driver()
        assert np.all(got["rate_scatt"] == 0)
        assert_close(got["WP3_minus"].ravel(order="F"), want["wp3_minus"], RTOL, "WP3 -")
    finally:
        c.close()


def test_diamond_16cube_shell2_subset():
    """Full-size diamond 16^3 with 2nd-shell third-order clusters: all counts
    bit-exact, lists of a mode subset bit-exact, rates of a subset."""
    from oracle.oracle import Oracle
    banner("3ph diamond 16x16x16 shell 2")
    case = crystal.config_case(2)
    c = rt.Context(case)
    o = Oracle(case)
    try:
        N = c.NP_driver()
        No, Po = o.np3_driver(nthreads=8)
        for ch in range(2):
            assert np.array_equal(N[2 * ch], No[ch]), ch
            assert_close(N[2 * ch + 1].ravel(order="F"), Po[ch], 1e-12, f"P {ch}")
        lists = c.Ind_only_driver(*No)
        for ch in range(2):
            acc = np.concatenate([[0], np.cumsum(No[ch].astype(np.int64))])
            for mm in (4, 100, 333, 611, c.nmodes):
                want = o.ind_only3_mode(ch, mm)
                for t in range(2):
                    assert np.array_equal(lists[ch][t][acc[mm - 1]:acc[mm]], want[t]), (ch, mm)
        modes = np.array([5, 200, 611, c.nmodes], np.int32)
        got, _ = c.RTA_driver_modes(modes)
        want = o.rta3_driver(modes=modes, nthreads=8)
        for g, w in KEYS:
            assert_close(got[g].ravel(order="F")[modes - 1], want[w][modes - 1], RTOL, g)
    finally:
        c.close()
