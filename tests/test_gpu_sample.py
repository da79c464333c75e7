"""Device sampling estimators (fp4_np_4ph_sample, fp4_rta_4ph_sample[_modes])
against the oracle's restatement at the same seed and sample count: the
same draws and the same exact window give identical hit sets, so the count
estimates are equal and phase space and rates agree to 1e-12 (the device
sums hits in a different order).  Then the estimator against the device's
own full FP64 rates over several seeds (unbiased within 5 standard
errors)."""
import numpy as np
import pytest

from fourphonon_amd import crystal
from fourphonon_amd import runtime as rt

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"\n[gpu] {msg}", flush=True)


def rel(got, want):
    got, want = np.asarray(got, float), np.asarray(want, float)
    scale = np.abs(want).max()
    return float(np.abs(got - want).max() / scale) if scale > 0 else float(np.abs(got).max())


@pytest.fixture(scope="module")
def ctx():
    case = crystal.config_case(1, ngrid=(4, 4, 4))
    banner("sampling context diamond 4x4x4")
    c = rt.Context(case)
    yield case, c
    c.close()


@pytest.mark.parametrize("nsample,seed", [(1000, 1), (37717, 2024)])
def test_np_sample_matches_oracle(ctx, nsample, seed):
    from oracle.oracle import Oracle
    case, c = ctx
    banner(f"NP_driver_4ph sampled, nsample {nsample}, seed {seed}")
    got = c.NP_driver_4ph_sample(nsample, seed)
    Nw, Pw = Oracle(case).np_sample_driver(nsample, seed, nthreads=8)
    for ch in range(3):
        assert np.array_equal(got[2 * ch], Nw[ch]), ch
        assert rel(got[2 * ch + 1].ravel(order="F"), Pw[ch]) <= 1e-12, ch
    assert sum(int(got[2 * ch].sum()) for ch in range(3)) > 0


def test_rta_sample_matches_oracle(ctx):
    from oracle.oracle import Oracle
    case, c = ctx
    nsample, seed = 20000, 77
    banner(f"RTA_driver_4ph sampled, nsample {nsample}")
    r, _ = c.RTA_driver_4ph_sample(nsample, seed)
    modes = np.arange(1, c.nmodes + 1, dtype=np.int32)
    w, hits = Oracle(case).rta_sample_driver(nsample, seed, modes=modes, nthreads=8)
    pairs = [("rate_scatt_4ph", "rate_4ph"), ("rate_scatt_plusplus", "rate_pp"),
             ("rate_scatt_plusminus", "rate_pm"), ("rate_scatt_minusminus", "rate_mm"),
             ("WP4_plusplus", "wp4_pp"), ("WP4_plusminus", "wp4_pm"), ("WP4_minusminus", "wp4_mm")]
    for dk, ok in pairs:
        err = rel(r[dk].ravel(order="F"), w[ok])
        print(f"{dk}: max rel err {err:.2e}")
        assert err <= 1e-12, dk
    # a subset gives the same per-mode values, bit for bit
    sub = np.array([3, 17, 30, 44], np.int32)
    rs, n = c.RTA_driver_4ph_sample(nsample, seed, modes=sub)
    assert n > 0
    for dk, _ in pairs:
        assert np.array_equal(rs[dk].ravel(order="F")[sub - 1], r[dk].ravel(order="F")[sub - 1]), dk
    ms, launches = c.kernel_stats(rt.Context.KS_SAMPLE)
    assert launches >= 2 and ms > 0


def test_rta_sample_unbiased_against_full_fp64(ctx):
    case, c = ctx
    modes = np.array([5, 17, 30, 44], np.int32)
    banner("sampled vs full FP64 rates, 8 seeds")
    full, _ = c.RTA_driver_4ph_modes(modes, rt.PREC_FP64)
    est = []
    for seed in range(100, 108):
        r, _ = c.RTA_driver_4ph_sample(100000, seed, modes=modes)
        est.append(r)
    for key in ("rate_scatt_plusminus", "rate_scatt_minusminus", "WP4_plusminus", "WP4_minusminus",
                "rate_scatt_4ph"):
        want = full[key].ravel(order="F")[modes - 1].sum()
        got = np.array([e[key].ravel(order="F")[modes - 1].sum() for e in est])
        se = got.std(ddof=1) / np.sqrt(len(got))
        print(f"{key}: full {want:.6e}, mean {got.mean():.6e}, se {se:.2e}")
        assert want > 0 and abs(got.mean() - want) <= 5 * se, key


def test_bad_sample_count_is_rejected(ctx):
    case, c = ctx
    with pytest.raises(rt.FP4Error):
        c.NP_driver_4ph_sample(0, 1)
    with pytest.raises(rt.FP4Error):
        c.RTA_driver_4ph_sample(-5, 1)


# ------------------------------------------------------------- three phonons
@pytest.mark.parametrize("nsample,seed", [(500, 3), (24001, 99)])
def test_np3_sample_matches_oracle(ctx, nsample, seed):
    from oracle.oracle import Oracle
    case, c = ctx
    banner(f"NP_driver sampled, nsample {nsample}, seed {seed}")
    got = c.NP_driver_sample(nsample, seed)
    Nw, Pw = Oracle(case).np3_sample_driver(nsample, seed, nthreads=8)
    for ch in range(2):
        assert np.array_equal(got[2 * ch], Nw[ch]), ch
        assert rel(got[2 * ch + 1].ravel(order="F"), Pw[ch]) <= 1e-12, ch


def test_rta3_sample_matches_oracle_and_is_unbiased(ctx):
    from oracle.oracle import Oracle
    case, c = ctx
    banner("RTA_driver sampled vs oracle and vs the full FP64 rates")
    r, _ = c.RTA_driver_sample(5000, 8)
    w, _ = Oracle(case).rta3_sample_driver(5000, 8, nthreads=8)
    for dk, ok in (("rate_scatt", "rate"), ("rate_scatt_plus", "rate_plus"), ("rate_scatt_minus", "rate_minus"),
                   ("WP3_plus", "wp3_plus"), ("WP3_minus", "wp3_minus")):
        err = rel(r[dk].ravel(order="F"), w[ok])
        print(f"{dk}: max rel err {err:.2e}")
        assert err <= 1e-12, dk
    modes = np.array([5, 17, 30, 44], np.int32)
    full, _ = c.RTA_driver_modes(modes)
    est = [c.RTA_driver_sample(4000, seed, modes=modes)[0] for seed in range(200, 208)]
    for key in ("rate_scatt_plus", "rate_scatt_minus", "WP3_plus", "WP3_minus"):
        want = full[key].ravel(order="F")[modes - 1].sum()
        got = np.array([e[key].ravel(order="F")[modes - 1].sum() for e in est])
        se = got.std(ddof=1) / np.sqrt(len(got))
        print(f"{key}: full {want:.6e}, mean {got.mean():.6e}, se {se:.2e}")
        assert want > 0 and abs(got.mean() - want) <= 5 * se, key
