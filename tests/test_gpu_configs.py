"""Parity at the sizes of the benchmark workloads (libqpx on the GPU against
the CPU oracle on the same inputs).

* 14 orbitals, 6 alpha + 6 beta electrons, full CI (9,018,009 determinants,
  4 vectors): the dense-map sigma kernels the benchmark times -- 48 alpha
  singles per string, FP64 and FP32 -- on 2,000+ strided rows probed by the
  oracle's work routine, and the 3-root Davidson's eigenpairs checked on
  those rows against the oracle's H;
* one full-vector oracle case per alpha-singles bucket of the dense-map
  column kernel (<= 48, <= 64, <= 96 singles);
* 96 orbitals, 10 + 10 electrons, N_int = 2, a CIPSI-grown wavefunction of
  more than 10^5 determinants: sparse sigma rows (FP64, FP32) against the
  oracle, mixed-precision against FP64 Davidson per root, and the
  eigenpairs checked on the oracle's rows;
* 128 orbitals, 12 + 12 electrons, N_int = 2, a selected wavefunction of
  10^5 determinants (all of them selectors): the selection task farm over 8
  leading and 8 strided generators against the oracle's -- every
  candidate count, the EN-PT2 of each task, and the 10^5 selected
  determinants and their keys bit for bit; fragments (`subset`) and the
  per-generator PT2 entry point against the oracle as well.

Tolerances: FP64 sigma 1e-11 relative to the largest entry; FP32 sigma 2e-5
of the largest entry of the sampled rows; Davidson |dE| (mixed against
FP64) 1e-9 Eh per root; PT2 1e-12 relative; selected determinants and keys
exact.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from quantum_package_amd.hamiltonian import fci_space, hf_determinant, synthetic_hamiltonian
from oracle import oracle as O
from test_gpu_sigma import _say

pytestmark = pytest.mark.gpu

NTHR = min(16, os.cpu_count() or 1)


@pytest.fixture(scope="module")
def ctx():
    from quantum_package_amd.api import Context
    c = Context(0)
    yield c
    c.close()


def _check_rows(v, vo, tol, what):
    scale = max(1.0, float(np.max(np.abs(vo))))
    err = float(np.max(np.abs(v - vo)))
    _say(f"  {what}: max |d| = {err:.3e} (scale {scale:.3e})")
    assert err <= tol * scale, (what, err, scale)


def _hf_index(dets, M, na, nb):
    hf = hf_determinant(M, na, nb)
    return int(np.nonzero(np.all(dets.reshape(dets.shape[0], -1) == hf.ravel()[None, :], axis=1))[0][0])


# ---------------------------------------------------------------- FCI M=14
def test_fci14_sigma_rows_and_davidson(ctx):
    M, na, nb, seed, nst = 14, 6, 6, 1002, 4
    H = synthetic_hamiltonian(M, na, nb, seed=seed)
    dets = fci_space(M, na, nb)
    n = dets.shape[0]
    _say(f"[config fci14] n_det={n} (sigma rows FP64/FP32 + 3-root Davidson)")
    ctx.set_hamiltonian(H)
    ctx.set_wavefunction(dets)
    ham, bil = O.Ham(H), O.Bilinear(dets)
    # the context's bilinear order is the reference's (psi_bilinear_matrix_order)
    assert np.array_equal(ctx.bilinear_order(), bil.order())
    u = np.random.default_rng(seed).standard_normal((n, nst))
    istep = n // 2003
    rows, vo, so = O.sigma_rows(ham, bil, u, istep, NTHR)
    assert rows.size >= 2000
    for prec, tol, name in ((0, 1e-11, "fp64"), (1, 2e-5, "fp32")):
        v, s = ctx.H_S2_u_0_nstates(u, prec=prec)
        assert ctx._force_sigma_path(0) == 1          # the dense-map kernels
        _check_rows(v[rows], vo, tol, f"sigma {name} v")
        _check_rows(s[rows], so, tol, f"sigma {name} s")
        del v, s
    # Davidson, lowest 3 roots + 1 (HF + seeded random guess, no S^2 filter)
    g = np.zeros((n, 3))
    g[_hf_index(dets, M, na, nb), 0] = 1.0
    g[:, 1:] = np.random.default_rng(1002).standard_normal((n, 2))
    e64, _, x64, n64 = ctx.davidson_diag_hs2(g, N_st=3, N_st_diag=4, threshold=1e-10, s2_eig=False)
    emx, _, _, nmx = ctx.davidson_diag_hs2(g, N_st=3, N_st_diag=4, threshold=1e-10, s2_eig=False, mixed=True,
                                           switch_threshold=1e-6)
    _say(f"  E fp64 {e64[:3]} ({n64}); mixed {emx[:3]} ({nmx})")
    assert nmx[1] > 0
    assert np.max(np.abs(e64[:3] - emx[:3])) < 1e-9
    # the eigenpairs on the oracle's H: (H x - lambda x) on the sampled rows
    x = x64[:, :3] / np.linalg.norm(x64[:, :3], axis=0)
    rows, hx, _ = O.sigma_rows(ham, bil, x, istep, NTHR)
    r = hx - x[rows] * e64[None, :3]
    rr = float(np.max(np.abs(r)))
    # squared residual norm extrapolated from the sample
    r2 = float(np.max((r ** 2).sum(axis=0) * n / rows.size))
    _say(f"  oracle-row residual max |r_i| {rr:.2e}, extrapolated |r|^2 {r2:.2e}")
    assert rr < 1e-6 and r2 < 1e-8


@pytest.mark.parametrize("name,M,na,nb,nst,singles", [
    ("fci-M14-3a3b", 14, 3, 3, 4, 33),     # 33 alpha singles per string: the <= 48 bucket
    ("fci-M24-3a1b", 24, 3, 1, 4, 63),     # the <= 64 bucket
    ("fci-M30-3a1b", 30, 3, 1, 1, 81),     # the <= 96 bucket
])
def test_dense_map_singles_buckets(ctx, name, M, na, nb, nst, singles):
    """Every instantiation of the dense-map column kernel (by the number of
    alpha singles a string has), FP64 and FP32, against the oracle's full
    sigma."""
    assert singles == na * (M - na)
    H = synthetic_hamiltonian(M, na, nb, seed=77)
    dets = fci_space(M, na, nb)
    dets = dets[np.random.default_rng(M).permutation(dets.shape[0])]
    _say(f"[dense buckets] case={name} n_det={dets.shape[0]} alpha singles={singles} nst={nst}")
    ctx.set_hamiltonian(H)
    ctx.set_wavefunction(dets)
    u = np.random.default_rng(5).standard_normal((dets.shape[0], nst))
    vo, so = O.sigma(O.Ham(H), O.Bilinear(dets), u, NTHR)
    for prec, tol, pname in ((0, 1e-11, "fp64"), (1, 2e-5, "fp32")):
        v, s = ctx.H_S2_u_0_nstates(u, prec=prec)
        assert ctx._force_sigma_path(0) == 1
        _check_rows(v, vo, tol, f"{pname} v")
        _check_rows(s, so, tol, f"{pname} s")


# ------------------------------------------------------- selected psi M=96
@pytest.fixture(scope="module")
def selected96(ctx):
    """96 orbitals, 10 + 10 electrons, 2 states: the CIPSI loop from HF to
    8,000 determinants, one selection step of 60,000 and the S^2
    completion."""
    from quantum_package_amd.cipsi import cipsi, make_s2_eigenfunction, n_leading, sort_by_weight
    M, na, nb, seed, ns = 96, 10, 10, 1003, 2
    H = synthetic_hamiltonian(M, na, nb, seed=seed)
    r = cipsi(ctx, H, n_det_max=8000, n_states=ns, s2_eig=True, log=lambda m: None)
    ds, cs, contrib = sort_by_weight(r.dets, r.coef)
    n_gen = n_leading(contrib, 0.99)
    n_sel = max(n_leading(contrib, 0.999, True), n_gen)
    ctx.set_wavefunction(ds)
    E0 = np.array([ctx.u_0_H_u_0(cs[:, k])[0] for k in range(ns)])
    _, new, _ = ctx.ZMQ_selection(np.asfortranarray(cs), n_gen, n_sel, E0, 60000)
    dets = np.concatenate([ds, new])
    dets = np.concatenate([dets, make_s2_eigenfunction(dets, na)])
    g = np.zeros((dets.shape[0], ns))
    g[: ds.shape[0]] = cs
    return H, dets, g


def test_selected96_sparse_sigma_rows_and_mixed_davidson(ctx, selected96):
    H, dets, g = selected96
    n = dets.shape[0]
    assert n >= 100000
    _say(f"[config selected96] n_det={n}")
    ctx.set_hamiltonian(H)
    ctx.set_wavefunction(dets)
    ham, bil = O.Ham(H), O.Bilinear(dets)
    assert np.array_equal(ctx.bilinear_order(), bil.order())
    u = np.random.default_rng(3).standard_normal((n, 4))
    istep = n // 2003
    rows, vo, so = O.sigma_rows(ham, bil, u, istep, NTHR)
    ctx._force_sigma_path(3)
    for prec, tol, name in ((0, 1e-11, "fp64"), (1, 2e-5, "fp32")):
        v, s = ctx.H_S2_u_0_nstates(u, prec=prec)
        assert ctx._force_sigma_path(3) == 3
        _check_rows(v[rows], vo, tol, f"sparse sigma {name} v")
        _check_rows(s[rows], so, tol, f"sparse sigma {name} s")
    ctx._force_sigma_path(0)
    e64, s64, x64, n64 = ctx.davidson_diag_hs2(g, N_st=2, N_st_diag=4, threshold=1e-10, s2_eig=True)
    emx, smx, _, nmx = ctx.davidson_diag_hs2(g, N_st=2, N_st_diag=4, threshold=1e-10, s2_eig=True, mixed=True,
                                             switch_threshold=1e-6)
    d = np.abs(e64[:2] - emx[:2])
    _say(f"  E fp64 {e64[:2]} ({n64}); mixed {emx[:2]} ({nmx}); |dE| per root {d}")
    assert nmx[1] > 0 and np.all(d < 1e-9)
    assert np.all(np.abs(s64[:2]) < 1e-6)      # singlets (s2_eig)
    x = x64[:, :2] / np.linalg.norm(x64[:, :2], axis=0)
    rows, hx, _ = O.sigma_rows(ham, bil, x, istep, NTHR)
    r = hx - x[rows] * e64[None, :2]
    _say(f"  oracle-row residual max |r_i| {np.max(np.abs(r)):.2e}")
    assert np.max(np.abs(r)) < 1e-5


# ---------------------------------------------------- selection at M = 128
@pytest.fixture(scope="module")
def selected128(ctx):
    """128 orbitals, 12 + 12 electrons: HF and one selection step to 10^5
    determinants, diagonalised, in psi_det_sorted order."""
    from quantum_package_amd.cipsi import sort_by_weight
    M, na, nb, seed = 128, 12, 12, 1004
    H = synthetic_hamiltonian(M, na, nb, seed=seed)
    ctx.set_hamiltonian(H)
    hf = hf_determinant(M, na, nb)[None]
    ctx.set_wavefunction(hf)
    E_hf = ctx.u_0_H_u_0(np.ones(1))
    _, new, _ = ctx.ZMQ_selection(np.ones((1, 1), order="F"), 1, 1, E_hf, 99999)
    dets = np.concatenate([hf, new])
    ctx.set_wavefunction(dets)
    g = np.zeros((dets.shape[0], 1))
    g[0, 0] = 1.0
    e, _, u, _ = ctx.davidson_diag_hs2(g, N_st=1, N_st_diag=4, threshold=1e-8, s2_eig=False)
    ds, cs, _ = sort_by_weight(dets, u[:, :1])
    ctx.set_wavefunction(ds)
    E0 = ctx.u_0_H_u_0(cs[:, 0])
    return H, ds, np.asfortranarray(cs), E0


def test_selection128_task_farm_vs_oracle(ctx, selected128):
    H, ds, cs, E0 = selected128
    n = ds.shape[0]
    assert n == 100000
    gens = list(range(1, 9)) + [int(x) for x in np.linspace(n // 8, n, 8).astype(int)]
    tasks = [(g, 0) for g in gens]
    N = n
    _say(f"[config selection128] n_det={n}, generators {gens}, to_select {N}")
    ctx.set_hamiltonian(H)
    ctx.set_wavefunction(ds)
    pt2, sel, vals, tpt2 = ctx.select_tasks(cs, n, E0, N, tasks)
    st = ctx.select_stats()
    ref = O.selection(O.Ham(H), ds, cs, max(gens), n, E0, N, nthreads=NTHR, tasks=tasks)
    _say(f"  gpu pt2 {pt2} oracle {ref['pt2']}; candidates {st['candidates']} / {ref['n_candidates']}; "
         f"selected {sel.shape[0]} / {ref['dets'].shape[0]}")
    assert st["candidates"] == ref["n_candidates"]
    assert np.allclose(pt2, ref["pt2"], rtol=1e-12, atol=0)
    assert np.allclose(tpt2, ref["task_pt2"], rtol=1e-12, atol=1e-15)
    assert sel.shape == ref["dets"].shape and sel.shape[0] == N
    assert np.array_equal(vals, ref["vals"])
    assert np.array_equal(sel, ref["dets"])
    # the per-generator PT2 entry point (the stochastic PT2's task)
    gp = ctx.select_generators(cs, n, E0, gens)
    assert np.allclose(gp, ref["task_pt2"], rtol=1e-12, atol=1e-15)


def test_selection128_fragments_vs_oracle(ctx, selected128):
    """subset > 0 (selection.irp.f:3-10,561-562) at 128 orbitals: fragments
    of the HF generator and of later generators, in one merged buffer."""
    H, ds, cs, E0 = selected128
    n = ds.shape[0]
    tasks = [(1, 1), (1, 77), (1, 144), (2, 5), (3000, 12), (3000, 13), (60000, 100)]
    _say(f"[config selection128] fragments {tasks}")
    ctx.set_hamiltonian(H)
    ctx.set_wavefunction(ds)
    pt2, sel, vals, tpt2 = ctx.select_tasks(cs, n, E0, 5000, tasks)
    ref = O.selection(O.Ham(H), ds, cs, 60000, n, E0, 5000, nthreads=NTHR, tasks=tasks)
    _say(f"  gpu pt2 {pt2} oracle {ref['pt2']}; task pt2 {tpt2[:, 0]}")
    assert np.allclose(pt2, ref["pt2"], rtol=1e-12, atol=0)
    assert np.allclose(tpt2, ref["task_pt2"], rtol=1e-12, atol=1e-15)
    assert np.array_equal(vals, ref["vals"])
    assert np.array_equal(sel, ref["dets"])


def test_select_tasks_rejects_bad_lists(ctx, selected128):
    from quantum_package_amd.api import QpxError
    H, ds, cs, E0 = selected128
    ctx.set_hamiltonian(H)
    ctx.set_wavefunction(ds)
    n = ds.shape[0]
    for bad in ([(0, 0)], [(n + 1, 0)], [(5, 0), (5, 3)], [(5, 2), (5, 2)], [(5, 145)]):
        with pytest.raises(QpxError):
            ctx.select_tasks(cs, n, E0, 10, bad)


def test_selection128_two_states_hbm_matrix(ctx, selected128):
    """Two states at 128 orbitals: N_states * M^2 doubles (256 KiB) exceed
    the LDS of a workgroup, so the coupling matrices live in HBM (the
    reference allocates mat(N_states, M, M) on the heap for any size,
    selection.irp.f:409).  Same bit-exact bar as one state."""
    H, ds, cs, E0 = selected128
    n = ds.shape[0]
    rng = np.random.default_rng(8)
    c2 = rng.standard_normal(n) * np.abs(cs[:, 0]) + 1e-3 * rng.standard_normal(n)
    c2 -= (c2 @ cs[:, 0]) * cs[:, 0]
    coef = np.asfortranarray(np.column_stack([cs[:, 0], c2 / np.linalg.norm(c2)]))
    E0_2 = np.array([E0[0], E0[0] + 0.5])
    tasks = [(1, 0), (2, 0), (7, 0), (40000, 0), (99000, 0)]
    _say(f"[config selection128] two states, tasks {tasks}")
    ctx.set_hamiltonian(H)
    ctx.set_wavefunction(ds)
    pt2, sel, vals, tpt2 = ctx.select_tasks(coef, n, E0_2, 20000, tasks)
    ref = O.selection(O.Ham(H), ds, coef, 99000, n, E0_2, 20000, nthreads=NTHR, tasks=tasks)
    _say(f"  gpu pt2 {pt2} oracle {ref['pt2']}")
    assert np.allclose(pt2, ref["pt2"], rtol=1e-12, atol=0)
    assert np.allclose(tpt2, ref["task_pt2"], rtol=1e-12, atol=1e-15)
    assert np.array_equal(vals, ref["vals"])
    assert np.array_equal(sel, ref["dets"])
