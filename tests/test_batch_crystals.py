"""The 24-crystal batch generator: seeded draws inside their stated ranges,
every (kind, natoms) pair present, and the harmonic data obeying the
invariants the hot path relies on (acoustic zeros at Gamma, orthonormal
eigenvectors, the wedge covering the grid, IFC4 index symmetry and acoustic
sum rule)."""
import numpy as np
import pytest

from fourphonon_amd import crystal


def test_draws_in_range_and_coverage():
    seen = set()
    for c in range(crystal.BATCH_SIZE):
        d = crystal.batch_crystal(c)
        again = crystal.batch_crystal(c)
        assert all(np.array_equal(d[k], again[k]) for k in d)
        assert 3.5 <= d["a_ang"] <= 6.0
        assert np.all((10.0 <= d["masses"]) & (d["masses"] <= 120.0))
        lo, hi = crystal.BATCH_MESH[d["natoms"]]
        assert 6 <= lo <= d["mesh"] <= hi <= 14
        seen.add((d["kind"], d["natoms"]))
    assert len(seen) == 12
    with pytest.raises(ValueError):
        crystal.batch_crystal(crystal.BATCH_SIZE)


@pytest.mark.parametrize("c", [0, 5, 10, 15])
def test_batch_case_invariants(c):
    case = crystal.batch_case(c, ngrid=(3, 3, 3))
    d = crystal.batch_crystal(c)
    m = case.modes
    assert case.nbands == 3 * d["natoms"] == 3 * case.structure.natoms
    assert np.all(m.energy[0, :3] == 0.0) and np.all(m.energy[1:, :] > 0.0)
    # eigenvectors orthonormal at every q
    e = np.asarray(m.eigenvect)
    for q in range(m.nptk):
        assert np.allclose(e[q].conj().T @ e[q], np.eye(case.nbands), atol=1e-10)
    # the wedge's orbits cover the grid exactly once
    assert int(m.Nequi.sum()) == m.nptk
    # IFC4: acoustic sum rule over the last index, symmetry under index swaps
    f = case.ifc4
    assert f.Ntri > 0
    phi = np.asarray(f.Phi)
    key = {}
    for t in range(f.Ntri):
        key.setdefault((f.Index_i[t], f.Index_j[t], f.Index_k[t], tuple(f.cells[t, 0]), tuple(f.cells[t, 1])),
                       []).append(t)
    worst = max(np.abs(phi[..., ts].sum(axis=-1)).max() for ts in key.values())
    assert worst <= 1e-8 * np.abs(phi).max()
