"""Selection/EN-PT2 parity: libqpx on the GPU against the oracle's
restatement of select_singles_and_doubles (same psi, same E0).

The couplings are accumulated per cell in selector order with the
reference's operation order, so the selected determinants and their keys
are compared bit for bit; PT2 sums (different summation trees) to 1e-12
relative.
"""
from __future__ import annotations

import numpy as np
import pytest

from quantum_package_amd.hamiltonian import hf_determinant, synthetic_hamiltonian
from oracle import oracle as O
from test_gpu_sigma import _random_space, _say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from quantum_package_amd.api import Context
    c = Context(0)
    yield c
    c.close()


def _sorted_psi(M, na, nb, n, seed, n_states):
    dets = _random_space(M, na, nb, n, seed)
    rng = np.random.default_rng(seed + 100)
    hf = hf_determinant(M, na, nb)
    coef = rng.standard_normal((dets.shape[0], n_states)) * 0.05
    ih = np.where(np.all(dets.reshape(dets.shape[0], -1) == hf.ravel()[None, :], axis=1))[0][0]
    coef[ih, 0] = 1.0
    if n_states > 1:
        coef[(ih + 1) % dets.shape[0], 1] = 0.8
    coef /= np.linalg.norm(coef, axis=0)
    w = (coef ** 2).mean(axis=1)
    order = np.argsort(-w, kind="stable")
    return dets[order], np.asfortranarray(coef[order])


CASES = [
    ("M10-3a3b", 10, 3, 3, 60, 21, 1, None, None),
    ("M10-3a3b-2states", 10, 3, 3, 60, 22, 2, None, None),
    ("M12-4a3b-gen-subset", 12, 4, 3, 150, 23, 1, 40, 100),
    ("M70-nint2", 70, 3, 3, 120, 24, 1, None, None),
    ("M16-5a5b", 16, 5, 5, 400, 25, 1, 200, 400),
    # three and four words per spin string; 130 orbitals keep the coupling
    # matrix in LDS for one state and in HBM for two; 200 orbitals in HBM
    ("M130-nint3", 130, 3, 3, 80, 26, 1, None, None),
    ("M130-nint3-2states-subset", 130, 4, 3, 150, 27, 2, 60, 120),
    ("M200-nint4", 200, 3, 2, 60, 28, 1, None, None),
]


@pytest.mark.parametrize("name,M,na,nb,n,seed,ns,n_gen,n_sel", CASES)
def test_selection_vs_oracle(ctx, name, M, na, nb, n, seed, ns, n_gen, n_sel):
    H = synthetic_hamiltonian(M, na, nb, seed=seed)
    psi, coef = _sorted_psi(M, na, nb, n, seed, ns)
    nd = psi.shape[0]
    n_gen = n_gen or nd
    n_sel = n_sel or nd
    ham = O.Ham(H)
    bil = O.Bilinear(psi)
    u = coef
    v, _ = O.sigma(ham, bil, u)
    E0 = np.array([u[:, k] @ v[:, k] for k in range(ns)])
    n_keep = 50
    _say(f"[selection] case={name} n_det={nd} n_gen={n_gen} n_sel={n_sel} ns={ns}")
    ctx.set_hamiltonian(H)
    ctx.set_wavefunction(psi)
    pt2, dets, vals = ctx.ZMQ_selection(coef, n_gen, n_sel, E0, n_keep)
    st = ctx.select_stats()
    ref = O.selection(ham, psi, coef, n_gen, n_sel, E0, n_keep)
    _say(f"  gpu pt2 {pt2} oracle {ref['pt2']} cand {st['candidates']} / {ref['n_candidates']}")
    assert np.allclose(pt2, ref["pt2"], rtol=1e-12, atol=1e-15)
    assert st["candidates"] == ref["n_candidates"]
    assert st["blocks"] == ref["n_blocks"]
    assert dets.shape == ref["dets"].shape
    assert np.array_equal(vals, ref["vals"])
    assert np.array_equal(dets, ref["dets"])


def test_selection_subset(ctx):
    M, na, nb, seed = 12, 4, 4, 31
    H = synthetic_hamiltonian(M, na, nb, seed=seed)
    psi, coef = _sorted_psi(M, na, nb, 120, seed, 1)
    ham = O.Ham(H)
    v, _ = O.sigma(ham, O.Bilinear(psi), coef)
    E0 = np.array([coef[:, 0] @ v[:, 0]])
    ctx.set_hamiltonian(H)
    ctx.set_wavefunction(psi)
    nd = psi.shape[0]
    _say("[selection subset] fragments sum to the full PT2")
    full, _, _ = ctx.ZMQ_selection(coef, nd, nd, E0, 10)
    tot = np.zeros(1)
    for sub in range(1, na * na + 1):
        p, _, _ = ctx.ZMQ_selection(coef, nd, nd, E0, 10, subset=sub)
        tot += p
    assert np.allclose(tot, full, rtol=1e-12)
