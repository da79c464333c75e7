"""Pins of the oracle's sampling estimators (NP_*_sample, RTA_*_sample,
Src/processes.f90:2480-2746, 4811-5171) against the full drivers, which the
other oracle tests pin.

The estimator is unbiased, so with fixed seeds these are deterministic
tests whose bars are set from the known sampling spread: for counts the
exact binomial sigma from the full per-mode counts; for phase space and
rates the standard error over independent seeds.  A dropped or inverted
half-space condition doubles or halves the estimates; a wrong box size or
scaling factor moves them by a factor of Nb or more; a mistake on the
diagonal (k >= j against k > j) is a ~10-sigma effect in the wide-window
case below.  The reference's own sampling cannot be rerun (it reseeds from
the clock), so parity with it is in distribution only.
"""
import copy

import numpy as np

from fourphonon_amd import crystal
from oracle.oracle import Oracle

SEEDS = [11, 12, 13, 14, 15, 16, 17, 18]


def _case(scalebroad=None):
    case = crystal.config_case(1, ngrid=(4, 4, 4))
    if scalebroad is not None:
        case = copy.copy(case)
        case.scalebroad = scalebroad
    return case


def _total(case):
    return case.modes.nbands ** 3 * case.modes.nptk ** 2


def _count_bar(Nfull, total, nsample, nsig):
    """nsig binomial sigmas of sum over modes of the count estimate, plus the
    truncation of INT() (at most 1 per mode)."""
    p = Nfull.astype(float) / total
    var = (total ** 2 / nsample * p * (1 - p)).sum()
    return nsig * np.sqrt(var) + len(Nfull)


def test_counts_unbiased_against_full_driver():
    case = _case()
    o = Oracle(case)
    Nf, Pf = o.np_driver(nthreads=8)
    total, n = _total(case), 100000
    Ns, Ps = [], []
    for seed in SEEDS:
        N, P = o.np_sample_driver(n, seed, nthreads=8)
        Ns.append([x.copy() for x in N])
        Ps.append([x.copy() for x in P])
    for ch in range(3):
        # every seed: the channel total within 5 binomial sigmas
        bar = _count_bar(Nf[ch], total, n, 5.0)
        for k in range(len(SEEDS)):
            d = abs(int(Ns[k][ch].sum()) - int(Nf[ch].sum()))
            assert d <= bar, (ch, k, d, bar)
        # phase space: mean over seeds within 5 standard errors
        tot = np.array([Ps[k][ch].sum() for k in range(len(SEEDS))])
        se = tot.std(ddof=1) / np.sqrt(len(tot))
        assert abs(tot.mean() - Pf[ch].sum()) <= 5 * se, (ch, tot.mean(), Pf[ch].sum(), se)
        assert se < 0.05 * Pf[ch].sum()


def test_wide_window_pins_the_half_space():
    """With a window 1e9 times wider, every candidate in the visited part of
    the box passes unless its velocity difference (and so sigma) is exactly
    zero; hit fractions are about 1/2 (++, +-) and 1/6 (--), and the channel
    totals have a relative sigma below 6e-4, against a 1.6e-2 relative
    change for a slip on the diagonal (k > j instead of k >= j)."""
    case = _case(scalebroad=1e9)
    o = Oracle(case)
    Nf, _ = o.np_driver(nthreads=8)
    total, n = _total(case), 400000
    for ch in range(3):
        frac = Nf[ch].sum() / (total * len(Nf[ch]))
        print(f"channel {ch}: hit fraction {frac:.3f}")
        assert 0.1 < frac < 0.6
    N, _ = o.np_sample_driver(n, 5, nthreads=8)
    for ch in range(3):
        bar = _count_bar(Nf[ch], total, n, 5.0)
        d = abs(int(N[ch].sum()) - int(Nf[ch].sum()))
        print(f"channel {ch}: full {int(Nf[ch].sum())}, sampled {int(N[ch].sum())}, bar {bar:.0f}")
        assert d <= bar, (ch, d, bar)
        assert bar < 3e-3 * Nf[ch].sum()


def test_rates_unbiased_against_full_driver():
    case = _case()
    o = Oracle(case)
    modes = np.array([5, 17, 30, 44], np.int32)
    full = o.rta_driver(modes=modes, nthreads=8)
    est = [o.rta_sample_driver(100000, seed, modes=modes, nthreads=8)[0] for seed in SEEDS]
    for key in ("rate_pm", "rate_mm", "wp4_pm", "wp4_mm", "rate_4ph"):
        want = full[key][modes - 1].sum()
        got = np.array([e[key][modes - 1].sum() for e in est])
        se = got.std(ddof=1) / np.sqrt(len(got))
        print(f"{key}: full {want:.6e}, mean of estimates {got.mean():.6e}, se {se:.2e}")
        assert want > 0 and abs(got.mean() - want) <= 5 * se, key
        assert se < 0.1 * want, key


def test_deterministic_and_keyed_by_mode():
    case = _case()
    o = Oracle(case)
    a = o.sample_mode(1, 17, 5000, 42, want_rates=True)
    assert o.sample_mode(1, 17, 5000, 42, want_rates=True) == a
    assert o.sample_mode(1, 17, 5000, 43, want_rates=True) != a
    # a mode's estimate does not depend on which other modes are asked
    r1, _ = o.rta_sample_driver(3000, 9, modes=np.array([17], np.int32), nthreads=2)
    r2, _ = o.rta_sample_driver(3000, 9, modes=np.array([3, 17, 40], np.int32), nthreads=2)
    for k in r1:
        assert r1[k][16] == r2[k][16]
    # more samples: hits grow in proportion
    h1 = o.sample_mode(1, 17, 20000, 42)[0]
    h2 = o.sample_mode(1, 17, 80000, 42)[0]
    assert 3.0 < h2 / h1 < 5.0


# ------------------------------------------------------------- three phonons
def test_3ph_counts_and_rates_unbiased_against_full_drivers():
    """Every one of the Nb*nptk*Nb candidates per mode is one the full 3ph
    loops visit, so hits/nsample estimates N/(Nb*nptk*Nb) directly."""
    case = _case()
    o = Oracle(case)
    Nf, Pf = o.np3_driver(nthreads=8)
    total, n = case.modes.nbands ** 2 * case.modes.nptk, 20000
    for seed in SEEDS[:2]:
        N, P = o.np3_sample_driver(n, seed, nthreads=8)
        for ch in range(2):
            bar = _count_bar(Nf[ch].astype(np.int64), total, n, 5.0)
            d = abs(int(N[ch].sum()) - int(Nf[ch].sum()))
            print(f"3ph channel {ch}: full {int(Nf[ch].sum())}, sampled {int(N[ch].sum())}, bar {bar:.0f}")
            assert d <= bar and bar < 0.05 * Nf[ch].sum(), (ch, d, bar)
    modes = np.array([5, 17, 30, 44], np.int32)
    full = o.rta3_driver(modes=modes, nthreads=8)
    est = [o.rta3_sample_driver(4000, seed, modes=modes, nthreads=8)[0] for seed in SEEDS]
    for key in ("rate_plus", "rate_minus", "wp3_plus", "wp3_minus", "rate"):
        want = full[key][modes - 1].sum()
        got = np.array([e[key][modes - 1].sum() for e in est])
        se = got.std(ddof=1) / np.sqrt(len(got))
        print(f"3ph {key}: full {want:.6e}, mean {got.mean():.6e}, se {se:.2e}")
        assert want > 0 and abs(got.mean() - want) <= 5 * se and se < 0.1 * want, key


def test_3ph_sampling_with_every_candidate_drawn_often():
    """With nsample >> Nb*nptk*Nb each candidate is drawn ~nsample/total
    times, so the estimate converges to the full value: at nsample = 200*total
    the relative error of the channel totals is below 1%."""
    case = _case()
    o = Oracle(case)
    Nf, _ = o.np3_driver(nthreads=8)
    total = case.modes.nbands ** 2 * case.modes.nptk
    N, _ = o.np3_sample_driver(200 * total, 3, nthreads=8)
    for ch in range(2):
        assert abs(int(N[ch].sum()) - int(Nf[ch].sum())) <= 0.01 * Nf[ch].sum(), ch
