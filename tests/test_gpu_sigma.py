"""Sigma parity: libqpx on the GPU against the CPU oracle (same inputs).

FP64 tolerance: 1e-11 absolute on sigma entries of O(1) vectors (the two
sides sum the same terms in different orders).  FP32: 2e-5 relative to the
vector norm.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

from quantum_package_amd.hamiltonian import fci_space, hf_determinant, list_to_bitstring, bitstring_to_list, \
    synthetic_hamiltonian
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    from quantum_package_amd.api import Context
    c = Context(0)
    yield c
    c.close()


def _random_space(M, na, nb, n, seed):
    """n distinct determinants: HF and random excitations of degree <= 4"""
    rng = np.random.default_rng(seed)
    n_int = (M + 63) // 64
    hf = hf_determinant(M, na, nb)
    out = {tuple(hf.ravel().tolist()): hf}
    while len(out) < n:
        d = hf.copy()
        for _ in range(int(rng.integers(1, 5))):
            s = int(rng.integers(2))
            occ = bitstring_to_list(d[s])
            vir = [p for p in range(1, M + 1) if p not in occ]
            h = int(rng.choice(occ))
            p = int(rng.choice(vir))
            occ = sorted([o for o in occ if o != h] + [p])
            d[s] = list_to_bitstring(occ, n_int)
        out.setdefault(tuple(d.ravel().tolist()), d)
    dets = np.array(list(out.values()), dtype=np.uint64)
    return dets[rng.permutation(len(dets))]


CASES = [
    ("fci-m8-4a4b", 8, 4, 4, None, 1001),
    ("fci-M10-3a3b", 10, 3, 3, None, 5),
    ("fci-M9-4a3b", 9, 4, 3, None, 6),
    ("sparse-M12", 12, 4, 4, 3000, 7),
    ("nint2-M70", 70, 3, 3, 2500, 8),
    ("nint2-M96", 96, 5, 5, 2000, 9),
    ("fci-M12-3a1b", 12, 3, 1, None, 10),
    # three and four words per spin string (the dense map does not apply;
    # above 160 orbitals the sparse walk reads the packed integrals, no M^4 table)
    ("nint3-M130", 130, 4, 3, 1500, 11),
    ("nint4-M200", 200, 3, 3, 1000, 12),
]


@pytest.mark.parametrize("name,M,na,nb,n,seed", CASES)
@pytest.mark.parametrize("path", [0, 1, 2, 3])
def test_sigma_fp64_vs_oracle(ctx, name, M, na, nb, n, seed, path):
    H = synthetic_hamiltonian(M, na, nb, seed=seed)
    dets = fci_space(M, na, nb) if n is None else _random_space(M, na, nb, n, seed)
    if n is None:
        dets = dets[np.random.default_rng(seed).permutation(dets.shape[0])]
    _say(f"[sigma fp64] case={name} path={path} n_det={dets.shape[0]}")
    ctx.set_hamiltonian(H)
    ctx.set_wavefunction(dets)
    rng = np.random.default_rng(seed + 1)
    nst = 4 if path != 2 else 3
    u = rng.standard_normal((dets.shape[0], nst))
    if path == 1:
        try:
            ctx._force_sigma_path(1)
            v, s = ctx.H_S2_u_0_nstates(u)
        except Exception as e:  # dense path not applicable to this size
            ctx._force_sigma_path(0)
            pytest.skip(f"dense-map path not applicable: {e}")
    else:
        ctx._force_sigma_path(path)
        v, s = ctx.H_S2_u_0_nstates(u)
    used = ctx._force_sigma_path(0)
    _say(f"  path used = {used}")
    ham = O.Ham(H)
    bil = O.Bilinear(dets)
    vo, so = O.sigma(ham, bil, u)
    assert np.max(np.abs(v - vo)) < 1e-11 * max(1.0, np.max(np.abs(vo)))
    assert np.max(np.abs(s - so)) < 1e-11 * max(1.0, np.max(np.abs(so)))


@pytest.mark.parametrize("path", [0, 3])
@pytest.mark.parametrize("name,M,na,nb,n,seed", CASES[:1] + CASES[3:5])
def test_sigma_fp32(ctx, name, M, na, nb, n, seed, path):
    H = synthetic_hamiltonian(M, na, nb, seed=seed)
    dets = fci_space(M, na, nb) if n is None else _random_space(M, na, nb, n, seed)
    _say(f"[sigma fp32] case={name} path={path} n_det={dets.shape[0]}")
    ctx.set_hamiltonian(H)
    ctx.set_wavefunction(dets)
    u = np.random.default_rng(3).standard_normal((dets.shape[0], 2))
    ctx._force_sigma_path(path)
    v, s = ctx.H_S2_u_0_nstates(u, prec=1)
    ctx._force_sigma_path(0)
    ham = O.Ham(H)
    vo, so = O.sigma(ham, O.Bilinear(dets), u)
    for k in range(2):
        assert np.linalg.norm(v[:, k] - vo[:, k]) < 2e-5 * np.linalg.norm(vo[:, k])
        assert np.linalg.norm(s[:, k] - so[:, k]) < 2e-5 * np.linalg.norm(so[:, k])


def test_sigma_sparse_without_integral_table(ctx):
    """Above 160 orbitals the sparse path reads the pair matrix instead of
    the M^4 table; three words per spin string."""
    M, na, nb, seed = 170, 3, 3, 12
    H = synthetic_hamiltonian(M, na, nb, seed=seed)
    dets = _random_space(M, na, nb, 1500, seed)
    _say(f"[sigma fp64] case=nint3-M170 path=3 n_det={dets.shape[0]}")
    ctx.set_hamiltonian(H)
    ctx.set_wavefunction(dets)
    u = np.random.default_rng(seed).standard_normal((dets.shape[0], 4))
    ctx._force_sigma_path(3)
    v, s = ctx.H_S2_u_0_nstates(u)
    assert ctx._force_sigma_path(0) == 3
    vo, so = O.sigma(O.Ham(H), O.Bilinear(dets), u)
    assert np.max(np.abs(v - vo)) < 1e-11 * max(1.0, np.max(np.abs(vo)))
    assert np.max(np.abs(s - so)) < 1e-11 * max(1.0, np.max(np.abs(so)))


def test_sigma_sparse_vs_scan_on_a_large_space(ctx):
    """60 000 determinants within four excitations of HF, 96 orbitals 10a10b
    (N_int = 2): the bucket walk and the scanning path, two independent
    algorithms each checked against the oracle above, agree to 1e-11 in
    FP64; the FP32 bucket walk agrees to 2e-5 of the vector norm."""
    M, na, nb, seed = 96, 10, 10, 13
    H = synthetic_hamiltonian(M, na, nb, seed=seed)
    dets = _random_space(M, na, nb, 60000, seed)
    _say(f"[sigma fp64] case=nint2-M96-10a10b sparse vs scan n_det={dets.shape[0]}")
    ctx.set_hamiltonian(H)
    ctx.set_wavefunction(dets)
    u = np.random.default_rng(seed).standard_normal((dets.shape[0], 4))
    out = {}
    for path in (3, 2):
        ctx._force_sigma_path(path)
        out[path] = ctx.H_S2_u_0_nstates(u)
        assert ctx._force_sigma_path(0) == path
    (v3, s3), (v2, s2) = out[3], out[2]
    assert np.max(np.abs(v3 - v2)) < 1e-11 * max(1.0, np.max(np.abs(v2)))
    assert np.max(np.abs(s3 - s2)) < 1e-11 * max(1.0, np.max(np.abs(s2)))
    ctx._force_sigma_path(3)
    vf, sf = ctx.H_S2_u_0_nstates(u, prec=1)
    ctx._force_sigma_path(0)
    for k in range(4):
        assert np.linalg.norm(vf[:, k] - v2[:, k]) < 2e-5 * np.linalg.norm(v2[:, k])
        assert np.linalg.norm(sf[:, k] - s2[:, k]) < 2e-5 * np.linalg.norm(s2[:, k])


@pytest.mark.parametrize("path", [0, 1, 2, 3])
def test_sigma_single_and_pair(ctx, path):
    """The smallest wavefunctions: one determinant (v = H_II u, s = S^2_II u)
    and an open-shell pair of spin partners (2x2 blocks) on every path."""
    from quantum_package_amd.hamiltonian import list_to_bitstring
    H = synthetic_hamiltonian(10, 3, 3, seed=40)
    ham = O.Ham(H)
    ctx.set_hamiltonian(H)
    one = hf_determinant(10, 3, 3)[None]
    a = list_to_bitstring([1, 2, 4], 1)
    b = list_to_bitstring([1, 2, 5], 1)
    pair = np.array([[a, b], [b, a]], dtype=np.uint64)
    _say(f"[sigma edge] one determinant and a spin pair, path={path}")
    for dets in (one, pair):
        ctx.set_wavefunction(dets)
        u = np.random.default_rng(3).standard_normal((dets.shape[0], 2))
        ctx._force_sigma_path(path)
        try:
            v, s = ctx.H_S2_u_0_nstates(u)
        except Exception as e:
            ctx._force_sigma_path(0)
            if path in (1, 3):
                pytest.skip(f"path {path} not applicable: {e}")
            raise
        ctx._force_sigma_path(0)
        vo, so = O.sigma(ham, O.Bilinear(dets), u)
        assert np.max(np.abs(v - vo)) < 1e-12 and np.max(np.abs(s - so)) < 1e-12
    # the pair: <a b|S^2|b a> = -1 (spin exchange), diagonal S^2 = 1
    assert np.allclose(so, np.array([[1, -1], [-1, 1]]) @ u)


def test_empty_wavefunction_is_rejected(ctx):
    from quantum_package_amd.api import QpxError
    ctx.set_hamiltonian(synthetic_hamiltonian(8, 4, 4, seed=1))
    with pytest.raises(QpxError):
        ctx.set_wavefunction(np.zeros((0, 2, 1), dtype=np.uint64))


def test_diag_h(ctx):
    H = synthetic_hamiltonian(70, 3, 3, seed=8)
    dets = _random_space(70, 3, 3, 500, 8)
    _say("[diag_h] M=70 n_det=500")
    ctx.set_hamiltonian(H)
    ctx.set_wavefunction(dets)
    d = ctx.diag_H()
    ham = O.Ham(H)
    want = np.array([ham.diag(x) for x in dets])
    assert np.max(np.abs(d - want)) < 1e-11


def test_bad_inputs_raise(ctx):
    from quantum_package_amd.api import QpxError
    H = synthetic_hamiltonian(8, 4, 4, seed=1)
    ctx.set_hamiltonian(H)
    bad = fci_space(8, 3, 4)[:10]          # wrong alpha count
    _say("[errors] wrong electron count / duplicates")
    with pytest.raises(QpxError):
        ctx.set_wavefunction(bad)
    dup = np.concatenate([fci_space(8, 4, 4)[:5]] * 2)
    with pytest.raises(QpxError):
        ctx.set_wavefunction(dup)
