"""Host logic of the CIPSI driver (no GPU): S^2 completion, sorting and
generator/selector counts."""
from __future__ import annotations

import numpy as np

from quantum_package_amd.cipsi import make_s2_eigenfunction, n_leading, sort_by_weight, det_keys
from quantum_package_amd.hamiltonian import fci_space, bitstring_to_list


def _pattern(d):
    a, b = int(d[0, 0]), int(d[1, 0])
    return (a & b, a ^ b)


def test_make_s2_eigenfunction_brute_force():
    M, na, nb = 8, 3, 3
    full = fci_space(M, na, nb)
    rng = np.random.default_rng(1)
    psi = full[rng.choice(full.shape[0], 25, replace=False)]
    got = make_s2_eigenfunction(psi, na)
    pats = {_pattern(d) for d in psi}
    inpsi = {tuple(d.ravel()) for d in psi}
    want = {tuple(d.ravel()) for d in full if _pattern(d) in pats and tuple(d.ravel()) not in inpsi}
    assert {tuple(d.ravel()) for d in got} == want
    assert len(got) == len(want)


def test_make_s2_eigenfunction_two_words():
    M, na = 70, 3
    from quantum_package_amd.hamiltonian import list_to_bitstring
    d = np.stack([list_to_bitstring([1, 65, 66], 2), list_to_bitstring([1, 2, 67], 2)])[None]
    got = make_s2_eigenfunction(d, na)
    # singly occupied: 2, 65, 66, 67 with 2 alpha among them -> C(4,2)=6 patterns, one is psi
    assert got.shape[0] == 5
    for g in got:
        a = bitstring_to_list(g[0]); b = bitstring_to_list(g[1])
        assert 1 in a and 1 in b and len(a) == 3 and len(b) == 3


def test_sort_and_leading():
    dets = fci_space(6, 2, 2)[:10]
    coef = np.linspace(0.1, 1.0, 10)[:, None]
    ds, cs, contrib = sort_by_weight(dets, coef)
    assert np.all(np.diff(contrib) <= 0)
    assert abs(contrib.sum() - 1) < 1e-12
    assert n_leading(contrib, 0.0) == 1
    assert n_leading(contrib, 1.0) <= 10
    assert det_keys(ds).shape == (10,)


def test_make_s2_eigenfunction_nothing_to_add():
    from quantum_package_amd.hamiltonian import hf_determinant
    d = hf_determinant(8, 3, 3)[None]
    assert make_s2_eigenfunction(d, 3).shape == (0, 2, 1)
    assert det_keys(np.zeros((0, 2, 1), dtype=np.uint64)).shape == (0,)


def test_cipsi_two_states_on_oracle():
    """Two states from HF: the second state has no vector until psi holds
    two determinants; the final energies respect the variational bound of
    the full-CI eigenvalues (min-max)."""
    from test_gpu_cipsi import OracleBackend
    from quantum_package_amd.cipsi import cipsi
    from quantum_package_amd.hamiltonian import fci_space, synthetic_hamiltonian
    from oracle import oracle as O
    H = synthetic_hamiltonian(8, 3, 3, seed=11)
    r = cipsi(OracleBackend(), H, n_det_max=400, n_states=2, s2_eig=True, log=lambda m: None)
    assert r.coef.shape[1] == 2 and np.all(np.isfinite(r.energies))
    dets = fci_space(8, 3, 3)
    n = dets.shape[0]
    ham, bil = O.Ham(H), O.Bilinear(dets)
    Hm, _ = O.sigma(ham, bil, np.eye(n))
    w = np.linalg.eigvalsh(0.5 * (Hm + Hm.T))
    assert r.energies[0] >= w[0] - 1e-10 and r.energies[1] >= w[1] - 1e-10
    assert r.energies[0] <= r.energies[1]
