"""Pins the CPU oracle against independent brute force (no GPU).

* integral key: closed form and 8-fold symmetry;
* excitation degree / phase: second-quantization sign counting;
* <I|H|J>, <I|H|I>, <I|S^2|J>: textbook Slater-Condon on spin-orbital lists;
* sigma: Knowles-Handy string formulation of the full-CI H (independent of
  the determinant-driven traversal), on the full space and a sparse subset;
* Davidson: Lanczos (scipy eigsh) on the Knowles-Handy operator;
* selection/EN-PT2: explicit enumeration of the first-order space.
"""
from __future__ import annotations

import itertools

import numpy as np
import pytest

from quantum_package_amd.hamiltonian import (all_strings, bitstring_to_list, fci_space, hf_determinant,
                                             integral_key, list_to_bitstring, n_packed, synthetic_hamiltonian)
from oracle import oracle as O
import brute


def _det(alpha, beta, n_int):
    return np.stack([list_to_bitstring(alpha, n_int), list_to_bitstring(beta, n_int)])


def _random_det(rng, M, na, nb):
    n_int = (M + 63) // 64
    a = sorted(rng.choice(np.arange(1, M + 1), na, replace=False).tolist())
    b = sorted(rng.choice(np.arange(1, M + 1), nb, replace=False).tolist())
    return _det(a, b, n_int)


def _excite(rng, det, M, degree):
    """random excitation of given degree from det"""
    d = det.copy()
    n_int = d.shape[1]
    for _ in range(degree):
        s = rng.integers(2)
        occ = bitstring_to_list(d[s])
        vir = [p for p in range(1, M + 1) if p not in occ]
        h = rng.choice(occ)
        p = rng.choice(vir)
        occ = [o for o in occ if o != h] + [int(p)]
        d[s] = list_to_bitstring(sorted(occ), n_int)
    return d


def test_integral_key_closed_form():
    M = 6
    keys = set()
    for i, j, k, l in itertools.product(range(1, M + 1), repeat=4):
        key = int(integral_key(i, j, k, l))
        # 8-fold symmetry of <ij|kl> for real orbitals
        for perm in ((k, j, i, l), (i, l, k, j), (k, l, i, j), (j, i, l, k), (l, i, j, k), (j, k, l, i), (l, k, j, i)):
            assert int(integral_key(*perm)) == key
        keys.add(key)
    assert keys == set(range(1, n_packed(M) + 1))


def test_synthetic_hamiltonian_properties():
    H = synthetic_hamiltonian(8, 4, 4, seed=1001)
    assert np.allclose(H.h, H.h.T)
    g = H.eri_dense_chemist()
    assert np.allclose(g, g.transpose(1, 0, 2, 3))
    assert np.allclose(g, g.transpose(2, 3, 0, 1))
    sup = g.reshape(64, 64)
    assert np.linalg.eigvalsh(sup).min() > -1e-12
    # the generator is deterministic
    H2 = synthetic_hamiltonian(8, 4, 4, seed=1001)
    assert np.array_equal(H.eri, H2.eri) and np.array_equal(H.h, H2.h)
    # <ij|kl> = (ik|jl)
    for (i, j, k, l) in [(1, 2, 3, 4), (2, 2, 5, 1), (8, 3, 3, 8)]:
        assert H.integral(i, j, k, l) == pytest.approx(g[i - 1, k - 1, j - 1, l - 1], abs=0)


@pytest.mark.parametrize("M,na,nb", [(10, 3, 3), (12, 5, 4), (70, 4, 4), (130, 5, 5)])
def test_degree_and_phase_second_quantization(M, na, nb):
    rng = np.random.default_rng(M)
    for trial in range(300):
        d1 = _random_det(rng, M, na, nb)
        deg = int(rng.integers(0, 4))
        d2 = _excite(rng, d1, M, deg)
        pc = sum(bin(int(x) ^ int(y)).count("1") for x, y in zip(d1.ravel(), d2.ravel()))
        # oracle degree = popcount(xor)/2
        n_int = d1.shape[1]
        lib = O.lib()
        dd1 = np.ascontiguousarray(d1)
        dd2 = np.ascontiguousarray(d2)
        got = lib.oq_excitation_degree(O._u64(dd1), O._u64(dd2), n_int)
        assert got == pc // 2
        if got in (1, 2):
            exc = np.zeros(12, dtype=np.int32)
            dg = O.C.c_int(0)
            ph = O.C.c_double(0)
            lib.oq_get_excitation(O._u64(dd1), O._u64(dd2), exc.ctypes.data_as(O.c_ip), O.C.byref(dg), O.C.byref(ph), n_int)
            # brute: d2 = phase * (a+ particles)(a holes) d1, pairing sorted holes
            # with sorted particles within each spin, spin alpha first
            occ = brute.det_to_spinorbs(d1)
            sign = 1
            e = exc.reshape(2, 2, 3)  # [spin, hp, n]
            for s in range(2):
                for n in range(1, e[s, 0, 0] + 1):
                    h = e[s, 0, n] - 1 + (100000 if s else 0)
                    p = e[s, 1, n] - 1 + (100000 if s else 0)
                    occ, s1 = brute.apply_annihilate(occ, h)
                    occ, s2 = brute.apply_create(occ, p)
                    sign *= s1 * s2
            assert occ == sorted(brute.det_to_spinorbs(d2))
            assert sign == ph.value, (d1, d2, exc)


@pytest.mark.parametrize("M,na,nb,seed", [(10, 3, 3, 11), (10, 4, 3, 12), (70, 3, 3, 13)])
def test_matrix_elements_slater_condon(M, na, nb, seed):
    H = synthetic_hamiltonian(M, na, nb, seed=seed)
    ham = O.Ham(H)
    rng = np.random.default_rng(seed)
    ref = hf_determinant(M, na, nb)
    for trial in range(150):
        d1 = _excite(rng, ref, M, int(rng.integers(0, 3)))
        d2 = _excite(rng, d1, M, int(rng.integers(0, 3)))
        want = brute.slater_condon(H, d1, d2)
        assert ham.i_H_j(d1, d2) == pytest.approx(want, abs=1e-12)
        assert ham.i_H_j(d2, d1) == pytest.approx(want, abs=1e-12)
        assert ham.diag(d1) == pytest.approx(brute.slater_condon(H, d1, d1), abs=1e-11)
        # diag_H_mat_elem_fock relative to a generator reproduces <D|H|D>
        f = ham.fock_tmp(d1)
        if ham.degree(d1, d2) <= 2:
            assert ham.diag_fock(d1, d2, f) == pytest.approx(ham.diag(d2), abs=1e-11)
        # the specialised sigma elements agree with i_H_j
        n_int = d1.shape[1]
        da, db = np.ascontiguousarray(d1), np.ascontiguousarray(d2)
        xa = sum(bin(int(x) ^ int(y)).count("1") for x, y in zip(d1[0], d2[0]))
        xb = sum(bin(int(x) ^ int(y)).count("1") for x, y in zip(d1[1], d2[1]))
        if xa == 2 and xb == 2:
            assert O.lib().oq_i_H_j_double_alpha_beta(ham.ptr, O._u64(da), O._u64(db)) == pytest.approx(want, abs=1e-12)
        if xa == 2 and xb == 0:
            assert O.lib().oq_i_H_j_mono_spin(ham.ptr, O._u64(da), O._u64(db), 1) == pytest.approx(want, abs=1e-12)
        if xa == 0 and xb == 2:
            assert O.lib().oq_i_H_j_mono_spin(ham.ptr, O._u64(da), O._u64(db), 2) == pytest.approx(want, abs=1e-12)
        if xa == 4 and xb == 0:
            assert O.lib().oq_i_H_j_double_spin(ham.ptr, O._u64(da[0].copy()), O._u64(db[0].copy())) == pytest.approx(want, abs=1e-12)
        del n_int


def test_s2_matrix_spectrum():
    # S^2 on the full (Sz=0) space of 6 orbitals, 3+3 electrons: eigenvalues s(s+1)
    M, na, nb = 6, 3, 3
    dets = fci_space(M, na, nb)
    n = dets.shape[0]
    S = np.zeros((n, n))
    lib = O.lib()
    D = np.ascontiguousarray(dets)
    for i in range(n):
        for j in range(n):
            S[i, j] = lib.oq_get_s2(O._u64(D[i]), O._u64(D[j]), 1)
    sz = 0.5 * (na - nb)
    S += np.eye(n) * sz * (sz - 1.0)   # S_z2_Sz added by the Davidson driver
    w = np.linalg.eigvalsh(S)
    allowed = np.array([s * (s + 1) for s in (0, 1, 2, 3)])
    assert np.all(np.min(np.abs(w[:, None] - allowed[None, :]), axis=1) < 1e-10)
    assert np.allclose(S, S.T)


def _kh(H):
    a = all_strings(H.M, H.n_alpha)
    b = all_strings(H.M, H.n_beta)
    return brute.KnowlesHandy(H, a, b), a, b


def test_sigma_full_ci_knowles_handy():
    H = synthetic_hamiltonian(8, 4, 4, seed=1001)
    kh, a, b = _kh(H)
    dets = fci_space(8, 4, 4)                     # alpha-major product order
    rng = np.random.default_rng(3)
    u = rng.standard_normal((dets.shape[0], 3))
    ham = O.Ham(H)
    bil = O.Bilinear(dets)
    v, s = O.sigma(ham, bil, u)
    for k in range(3):
        want = kh.sigma(u[:, k].reshape(a.size, b.size)).ravel()
        assert np.max(np.abs(v[:, k] - want)) < 1e-11


def test_sigma_sparse_subset():
    H = synthetic_hamiltonian(8, 4, 4, seed=1001)
    kh, a, b = _kh(H)
    dets = fci_space(8, 4, 4)
    rng = np.random.default_rng(5)
    keep = np.sort(rng.choice(dets.shape[0], 1700, replace=False))
    perm = rng.permutation(keep)
    sub = dets[perm]
    u = rng.standard_normal((sub.shape[0], 2))
    ham = O.Ham(H)
    bil = O.Bilinear(sub)
    v, s = O.sigma(ham, bil, u)
    for k in range(2):
        full = np.zeros(dets.shape[0])
        full[perm] = u[:, k]
        want = kh.sigma(full.reshape(a.size, b.size)).ravel()[perm]
        assert np.max(np.abs(v[:, k] - want)) < 1e-11


def test_davidson_fci_m8_lanczos():
    from scipy.sparse.linalg import LinearOperator, eigsh
    H = synthetic_hamiltonian(8, 4, 4, seed=1001)
    kh, a, b = _kh(H)
    n = a.size * b.size
    op = LinearOperator((n, n), matvec=lambda x: kh.sigma(np.asarray(x).reshape(a.size, b.size)).ravel(),
                        dtype=np.float64)
    w = eigsh(op, k=6, which="SA", tol=1e-13, return_eigenvectors=False)
    w = np.sort(w)
    dets = fci_space(8, 4, 4)
    ham = O.Ham(H)
    bil = O.Bilinear(dets)
    guess = np.zeros(n)
    hf = np.where((dets[:, 0, 0] == 15) & (dets[:, 1, 0] == 15))[0][0]
    guess[hf] = 1.0
    e, s2, u, nsig = O.davidson(ham, bil, dets, guess, n_st=1, n_st_diag=4, threshold=1e-10)
    # lowest singlet: the S^2 filter keeps singlets; compare with the lowest
    # eigenvalue whose S^2 is 0
    assert abs(s2[0]) < 1e-6
    # find singlet eigenvalues of the full H via dense S^2 projection check
    assert min(abs(e[0] - x) for x in w) < 1e-9
    # residual of the returned vector
    x = u[:, 0] / np.linalg.norm(u[:, 0])
    r = kh.sigma(x.reshape(a.size, b.size)).ravel() - e[0] * x
    assert np.linalg.norm(r) < 1e-4


def _first_order_space(psi, M):
    """all singles/doubles of the dets of psi that are not in psi"""
    n_int = psi.shape[2]
    inpsi = {tuple(d.ravel().tolist()) for d in psi}
    out = set()
    for d in psi:
        occ = [bitstring_to_list(d[0]), bitstring_to_list(d[1])]
        vir = [[p for p in range(1, M + 1) if p not in occ[s]] for s in range(2)]
        singles = []
        for s in range(2):
            for h in occ[s]:
                for p in vir[s]:
                    singles.append((s, h, p))
        cand = []
        for (s, h, p) in singles:
            cand.append([(s, h, p)])
        for i1 in range(len(singles)):
            for i2 in range(i1 + 1, len(singles)):
                (s1, h1, p1), (s2, h2, p2) = singles[i1], singles[i2]
                if s1 == s2 and (h1 == h2 or p1 == p2):
                    continue
                cand.append([singles[i1], singles[i2]])
        for ops in cand:
            o = [list(occ[0]), list(occ[1])]
            for (s, h, p) in ops:
                o[s].remove(h)
                o[s].append(p)
            key = tuple(np.stack([list_to_bitstring(sorted(o[0]), n_int),
                                  list_to_bitstring(sorted(o[1]), n_int)]).ravel().tolist())
            if key not in inpsi:
                out.add(key)
    return [np.array(k, dtype=np.uint64).reshape(2, n_int) for k in sorted(out)]


@pytest.mark.parametrize("n_states", [1, 2])
def test_selection_brute_force(n_states):
    M, na, nb = 8, 3, 3
    H = synthetic_hamiltonian(M, na, nb, seed=2024)
    ham = O.Ham(H)
    rng = np.random.default_rng(11 + n_states)
    hf = hf_determinant(M, na, nb)
    psi = [hf]
    seen = {tuple(hf.ravel().tolist())}
    while len(psi) < 14:
        d = _excite(rng, hf, M, int(rng.integers(1, 3)))
        k = tuple(d.ravel().tolist())
        if k not in seen:
            seen.add(k)
            psi.append(d)
    psi = np.array(psi, dtype=np.uint64)
    coef = rng.standard_normal((psi.shape[0], n_states)) * 0.1
    coef[0, 0] = 1.0
    if n_states > 1:
        coef[1, 1] = 0.9
    coef /= np.linalg.norm(coef, axis=0)
    # sort by decreasing state-averaged weight (psi_det_sorted)
    w = (coef ** 2).mean(axis=1)
    order = np.argsort(-w, kind="stable")
    psi, coef = psi[order], coef[order]
    n = psi.shape[0]
    Hm = np.array([[ham.i_H_j(psi[i], psi[j]) for j in range(n)] for i in range(n)])
    E0 = np.array([coef[:, k] @ Hm @ coef[:, k] for k in range(n_states)])
    res = O.selection(ham, psi, coef, n_gen=n, n_sel=n, E0=E0, n_keep=25)
    ext = _first_order_space(psi, M)
    pt2 = np.zeros(n_states)
    vals = []
    for D in ext:
        V = np.array([sum(coef[i, k] * ham.i_H_j(D, psi[i]) for i in range(n)) for k in range(n_states)])
        if np.all(V == 0):
            continue
        Hii = ham.diag(D)
        dE = E0 - Hii
        e = 0.5 * (np.sign(dE) * np.sqrt(dE * dE + 4 * V * V) - dE)
        pt2 += e
        vals.append((min(e.min(), 0.0), D))
    assert np.allclose(res["pt2"], pt2, atol=1e-12, rtol=1e-10)
    vals.sort(key=lambda t: t[0])
    want = {tuple(d.ravel().tolist()) for v, d in vals[:25]}
    got = {tuple(d.ravel().tolist()) for d in res["dets"]}
    cut = vals[24][0]
    near = {tuple(d.ravel().tolist()) for v, d in vals if abs(v - cut) <= 4 * np.spacing(abs(cut))}
    assert (want ^ got) <= near
    assert res["n_candidates"] >= len(vals)


# ------------------------------------------------- generator/selector split
def _holes(det):
    return [bitstring_to_list(det[0]), bitstring_to_list(det[1])]


def _block_index(gen):
    """maskInd of every (s1, h1, s2, h2) block of one generator, in the
    reference's loop order (selection.irp.f:412-420,556-557): s1 = 1..2, i1
    from the last hole down, s2 = s1..2, i2 from the last hole down to
    i1+1 (same spin) or 1."""
    h = _holes(gen)
    out, m = {}, -1
    for s1 in (1, 2):
        for i1 in range(len(h[s1 - 1]), 0, -1):
            for s2 in range(s1, 3):
                ib = i1 + 1 if s1 == s2 else 1
                for i2 in range(len(h[s2 - 1]), ib - 1, -1):
                    m += 1
                    out[(s1, h[s1 - 1][i1 - 1], s2, h[s2 - 1][i2 - 1])] = m
    return out


def _block_of(gen, D):
    """The block of generator gen that enumerates the external determinant
    D.  Doubles sit in the block of their two holes; single excitations are
    folded into the alpha-beta blocks: an alpha single h -> p where the beta
    hole is gen's last beta orbital, a beta single where the alpha hole is
    gen's last alpha orbital (the monoAdo / monoBdo unbans, :551-559)."""
    hg, hd = _holes(gen), _holes(D)
    holes = [(s + 1, o) for s in range(2) for o in hg[s] if o not in hd[s]]
    if len(holes) == 2:
        (s1, a), (s2, b) = sorted(holes)
        return (s1, a, s2, b)
    (s, o), = holes
    return (1, o, 2, hg[1][-1]) if s == 1 else (1, hg[0][-1], 2, o)


def _brute_selection(ham, psi, coef, n_gen, n_sel, E0, M, subset=0, fragment_count=None):
    """EN-PT2 and keys from the definition: every determinant within two
    excitations of a generator and not in psi, counted once -- by the first
    generator it is connected to, in the block that enumerates it there --
    with V = sum over all selectors of c_S <D|H|S>."""
    ns = coef.shape[1]
    inpsi = {tuple(d.ravel().tolist()) for d in psi}
    ext = [D for D in _first_order_space(psi[:n_gen], M) if tuple(D.ravel().tolist()) not in inpsi]
    pt2 = np.zeros(ns)
    vals = []
    for D in ext:
        g = next(i for i in range(n_gen) if ham.degree(D, psi[i]) <= 2)
        if subset:
            m = _block_index(psi[g])[_block_of(psi[g], D)]
            if m % fragment_count != subset - 1:
                continue
        V = np.array([sum(coef[i, k] * ham.i_H_j(D, psi[i]) for i in range(n_sel)) for k in range(ns)])
        if np.all(V == 0):
            continue
        dE = E0 - ham.diag(D)
        e = 0.5 * (np.sign(dE) * np.sqrt(dE * dE + 4 * V * V) - dE)
        pt2 += e
        vals.append((min(e.min(), 0.0), D))
    vals.sort(key=lambda t: t[0])
    return pt2, vals


def _check_top(res, vals, n_keep):
    want = {tuple(d.ravel().tolist()) for v, d in vals[:n_keep]}
    got = {tuple(d.ravel().tolist()) for d in res["dets"]}
    cut = vals[min(n_keep, len(vals)) - 1][0]
    near = {tuple(d.ravel().tolist()) for v, d in vals if abs(v - cut) <= 4 * np.spacing(abs(cut))}
    assert (want ^ got) <= near


def _split_psi(seed, n_states, n=16, M=8, na=3, nb=3):
    H = synthetic_hamiltonian(M, na, nb, seed=seed)
    ham = O.Ham(H)
    rng = np.random.default_rng(seed)
    hf = hf_determinant(M, na, nb)
    psi, seen = [hf], {tuple(hf.ravel().tolist())}
    while len(psi) < n:
        d = _excite(rng, hf, M, int(rng.integers(1, 4)))
        k = tuple(d.ravel().tolist())
        if k not in seen:
            seen.add(k)
            psi.append(d)
    psi = np.array(psi, dtype=np.uint64)
    coef = rng.standard_normal((n, n_states)) * 0.1
    coef[0, 0] = 1.0
    if n_states > 1:
        coef[1, 1] = 0.9
    coef /= np.linalg.norm(coef, axis=0)
    order = np.argsort(-(coef ** 2).mean(axis=1), kind="stable")
    psi, coef = psi[order], coef[order]
    Hm = np.array([[ham.i_H_j(psi[i], psi[j]) for j in range(n)] for i in range(n)])
    E0 = np.array([coef[:, k] @ Hm @ coef[:, k] for k in range(n_states)])
    return H, ham, psi, coef, E0


@pytest.mark.parametrize("n_states,n_gen,n_sel", [(1, 5, 12), (2, 4, 16), (1, 9, 9)])
def test_selection_brute_force_generator_selector_split(n_states, n_gen, n_sel):
    """n_gen < n_sel: the generators own the external space (each external
    determinant counted by its first connected generator), the couplings run
    over every selector (selection.irp.f:50-79,320-403,1088-1171)."""
    H, ham, psi, coef, E0 = _split_psi(300 + n_gen, n_states)
    res = O.selection(ham, psi, coef, n_gen=n_gen, n_sel=n_sel, E0=E0, n_keep=20)
    pt2, vals = _brute_selection(ham, psi, coef, n_gen, n_sel, E0, H.M)
    assert np.allclose(res["pt2"], pt2, atol=1e-12, rtol=1e-10)
    _check_top(res, vals, 20)


@pytest.mark.parametrize("subset", [1, 2, 5, 9])
def test_selection_brute_force_subset(subset):
    """subset > 0 keeps the blocks with maskInd mod fragment_count ==
    subset-1 (selection.irp.f:3-10,556-557); fragment_count = n_alpha^2."""
    H, ham, psi, coef, E0 = _split_psi(400 + subset, 1)
    n_gen, n_sel = 6, psi.shape[0]
    fc = H.n_alpha ** 2
    tasks = [(g, subset) for g in range(1, n_gen + 1)]
    res = O.selection(ham, psi, coef, n_gen=n_gen, n_sel=n_sel, E0=E0, n_keep=15, tasks=tasks)
    pt2, vals = _brute_selection(ham, psi, coef, n_gen, n_sel, E0, H.M, subset=subset, fragment_count=fc)
    assert np.allclose(res["pt2"], pt2, atol=1e-12, rtol=1e-10)
    if vals:
        _check_top(res, vals, 15)


def test_selection_subsets_partition_the_generator():
    """The fragments of one generator partition its blocks: their PT2 sum to
    the whole generator's (oracle task farm against itself, all subsets)."""
    H, ham, psi, coef, E0 = _split_psi(500, 1)
    n = psi.shape[0]
    whole = O.selection(ham, psi, coef, n_gen=n, n_sel=n, E0=E0, n_keep=5, tasks=[(3, 0)])["pt2"]
    parts = sum(O.selection(ham, psi, coef, n_gen=n, n_sel=n, E0=E0, n_keep=5, tasks=[(3, s)])["pt2"]
                for s in range(1, H.n_alpha ** 2 + 1))
    assert np.allclose(parts, whole, rtol=1e-12, atol=1e-15)


# ------------------------------------------------------------ block network
def test_mlp_reference_against_autograd():
    """oracle/mlp_ref.forward_backward with the bf16 rounding switched off is
    the exact gradient of its loss: compared with torch.autograd on the same
    network written independently (float64)."""
    torch = pytest.importorskip("torch")
    from oracle import mlp_ref
    rng = np.random.default_rng(5)
    n, n_orb = 37, 11
    P = {"W1": rng.standard_normal((80, 256)) * 0.1, "W2": rng.standard_normal((256, 256)) * 0.06,
         "W3": rng.standard_normal((256, 256)) * 0.06, "b1": rng.standard_normal(256) * 0.1,
         "b2": rng.standard_normal(256) * 0.1, "b3": rng.standard_normal(256) * 0.1,
         "w4": rng.standard_normal(256) * 0.06, "b4": np.array([0.3]), "emb": rng.standard_normal((n_orb, 32)) * 0.3}
    desc = np.stack([rng.integers(0, n_orb, n), rng.integers(0, n_orb, n), rng.integers(0, 3, n),
                     rng.integers(0, 5, n)], axis=1)
    fdesc = rng.standard_normal((n, 3))
    target = rng.standard_normal(n)
    y, loss, g = mlp_ref.forward_backward(P, desc, fdesc, target, round_bf16=False)

    T = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in P.items()}
    d = torch.tensor(desc)
    X = torch.cat([T["emb"][d[:, 0]], T["emb"][d[:, 1]], torch.tensor(fdesc),
                   torch.nn.functional.one_hot(d[:, 2], 3).double(), torch.nn.functional.one_hot(d[:, 3], 5).double(),
                   torch.zeros(n, 5, dtype=torch.float64)], dim=1)
    h = torch.relu(X @ T["W1"] + T["b1"])
    h = torch.relu(h @ T["W2"] + T["b2"])
    h = torch.relu(h @ T["W3"] + T["b3"])
    yt = h @ T["w4"] + T["b4"]
    lt = ((yt - torch.tensor(target)) ** 2).mean()
    lt.backward()
    assert np.allclose(y, yt.detach().numpy(), rtol=1e-12, atol=1e-12)
    assert abs(loss - lt.item()) <= 1e-12 * max(1.0, abs(loss))
    for k in P:
        want = T[k].grad.numpy()
        assert np.allclose(g[k], want, rtol=1e-10, atol=1e-12), k
