"""Pins of the oracle's iterative-solver routines (iterations.f90,
conductivity.f90, ShengBTE.f90:1250-1307) by independent routes:

* one iteration_4ph sweep (identity symmetrisers) equals the dense
  Gauss-Seidel triangular solve of the same linear system, for 3ph and 4ph
  list types, with the real star/symmetry tables of a symmetric mesh;
* the converged loop equals the dense linear solution;
* symmetrize_tensor under the 48 operations of the diamond point group gives
  trace/3 times the identity (the only invariant rank-2 tensor of O_h);
* the symmetrisers are projectors.
"""
import numpy as np
import pytest

from oracle.oracle import (bte_solve, iteration0, iteration_4ph, symmetrize_F, symmetrize_tensor, tconduct,
                           twonorm3x3)
from tests import iterops


@pytest.fixture(scope="module")
def sym_case(small_case):
    assert small_case.modes.nsymm_rot == 48 and small_case.modes.Nequi.max() > 1
    return small_case


@pytest.mark.parametrize("with_3ph", [False, True], ids=["4ph", "3ph+4ph"])
def test_sweep_is_gauss_seidel_on_the_dense_system(sym_case, with_3ph):
    case = iterops.identity_symmetrizers(sym_case)
    N4, L4, G4, N3, L3, G3 = iterops.synthetic_operator(case, seed=11, with_3ph=with_3ph)
    tau = iterops.contractive_tau(case, N4, G4, N3, G3, margin=0.9)
    F = iteration0(case, tau)
    for sweep in range(2):
        got = iteration_4ph(case, N4, L4, G4, tau, F, N3, L3, G3)
        want = iterops.gauss_seidel_sweep(case, N4, L4, G4, N3, L3, G3, tau, F)
        scale = np.abs(want).max()
        assert np.abs(got - want).max() <= 1e-12 * scale, sweep
        F = got


def test_sweep_differs_from_jacobi(sym_case):
    """The in-place order matters: a Jacobi-style update (all reads from the
    old F) gives a different result, so the check above is not vacuous."""
    case = iterops.identity_symmetrizers(sym_case)
    N4, L4, G4, N3, L3, G3 = iterops.synthetic_operator(case, seed=12)
    tau = iterops.contractive_tau(case, N4, G4, N3, G3, margin=0.9)
    F = iteration0(case, tau)
    B, b, _ = iterops._system(case, N4, L4, G4, N3, L3, G3, tau)
    jac = (b + B @ F.reshape((-1, 3), order="F")).reshape(F.shape, order="F")
    got = iteration_4ph(case, N4, L4, G4, tau, F, N3, L3, G3)
    assert np.abs(got - jac).max() > 1e-6 * np.abs(jac).max()


def test_converged_loop_is_the_linear_solution(sym_case):
    case = iterops.identity_symmetrizers(sym_case)
    N4, L4, G4, N3, L3, G3 = iterops.synthetic_operator(case, seed=13)
    tau = iterops.contractive_tau(case, N4, G4, N3, G3, margin=0.25)
    res = bte_solve(case, N4, L4, G4, tau, maxiter=200, eps=1e-14, N3=N3, lists3=L3, gammas3=G3)
    want = iterops.linear_solution(case, N4, L4, G4, N3, L3, G3, tau)
    assert res["niter"] < 200
    assert np.abs(res["F_n"] - want).max() <= 1e-11 * np.abs(want).max()
    k = tconduct(case, want)
    ks = np.sum([symmetrize_tensor(case, k[b]) for b in range(k.shape[0])], axis=0)
    assert np.allclose(res["kappa_iter"][:, :, -1], ks, rtol=1e-10, atol=1e-10 * abs(ks).max())
    assert res["relchange"][0] == 0.0 and res["relchange"][-1] < 1e-14


def test_symmetrize_tensor_cubic_closed_form(sym_case):
    rng = np.random.default_rng(5)
    for _ in range(3):
        t = rng.normal(size=(3, 3))
        got = symmetrize_tensor(sym_case, t)
        assert np.allclose(got, np.trace(t) / 3 * np.eye(3), rtol=0, atol=1e-14)


def test_symmetrizers_are_projectors(sym_case):
    rng = np.random.default_rng(6)
    m = sym_case.modes
    F = np.asfortranarray(rng.normal(size=(m.nbands, m.nptk, 3)))
    once = symmetrize_F(sym_case, F)
    twice = symmetrize_F(sym_case, once)
    assert np.allclose(once, twice, rtol=0, atol=1e-13)


def test_twonorm3x3_closed_forms():
    assert twonorm3x3(np.diag([1.0, -4.0, 2.0])) == pytest.approx(4.0, rel=1e-15)
    a = np.outer([1.0, 2.0, 2.0], [0.0, 3.0, 4.0])          # rank one: |u| |v|
    assert twonorm3x3(a) == pytest.approx(15.0, rel=1e-15)
