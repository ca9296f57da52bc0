"""CIPSI driver: the loop of the reference's ``fci_zmq`` program
(plugins/Full_CI_ZMQ/fci_zmq.irp.f:57-170) around the GPU hot path.

Per iteration:
  1. diagonalize_CI: dense up to n_det_max_jacobi determinants (the
     reference's Lapack branch with its S^2 state choice), davidson_diag_hs2
     above (src/Davidson/diagonalize_CI.irp.f:57-145);
  2. psi_det_sorted: determinants by decreasing state-averaged weight
     (src/Determinants/determinants.irp.f:228-293); generators / selectors
     are the leading determinants up to threshold_generators /
     threshold_selectors of the norm (plugins/Generators_full/generators.irp.f:3-23,
     plugins/Selectors_full/selectors.irp.f:3-24);
  3. ZMQ_selection(to_select, pt2) with E0 = psi_energy
     (plugins/Full_CI_ZMQ/energy.irp.f:9-24), to_select = N_det/2+1 when
     s2_eig else N_det (fci_zmq.irp.f:148-156);
  4. fill_H_apply_buffer_no_selection + copy_H_apply_buffer_to_wf: the new
     determinants enter psi with zero coefficients and psi is normalised
     (src/Determinants/H_apply.irp.f:113-198,275-316);
  5. make_s2_eigenfunction (on the GPU, Context.s2_complete) when s2_eig or
     N_states > 1: every determinant
     sharing an occupation pattern with psi is added
     (src/Determinants/occ_pattern.irp.f:251-309).

With a proposer (quantum_package_amd.proposal.Proposer) step 3 enumerates
only the (generator, hole-pair) blocks the network proposes; every
full_sweep_every-th iteration (and until the network is trained) the full
enumeration runs instead, trains the network on its exact block PT2 and
records the PT2 fraction the proposal would have missed.

The backend is libqpx (quantum_package_amd.api.Context); the same loop runs
on any object exposing set_wavefunction / davidson_diag_hs2 / ZMQ_selection.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np


def popcount64(x: np.ndarray) -> np.ndarray:
    x = x.astype(np.uint64)
    c = np.zeros(x.shape, dtype=np.int64)
    for sh in range(0, 64, 8):
        c += _POP8[((x >> np.uint64(sh)) & np.uint64(0xFF)).astype(np.int64)]
    return c


_POP8 = np.array([bin(i).count("1") for i in range(256)], dtype=np.int64)


def det_keys(dets: np.ndarray) -> np.ndarray:
    """one hashable row per determinant (structured void view)"""
    d = np.ascontiguousarray(dets.reshape(dets.shape[0], int(np.prod(dets.shape[1:]))))
    return d.view(np.dtype((np.void, d.dtype.itemsize * d.shape[1]))).ravel()


def sort_by_weight(dets, coef, state_weights=None):
    """psi_det_sorted: decreasing state-averaged weight, stable."""
    ns = coef.shape[1]
    w = np.ones(ns) / ns if state_weights is None else np.asarray(state_weights)
    contrib = (coef ** 2) @ w
    contrib = contrib / contrib.sum()
    order = np.argsort(-contrib, kind="stable")
    return dets[order], coef[order], contrib[order]


def n_leading(contrib_sorted, threshold, strict=False):
    """number of leading determinants whose cumulated weight reaches threshold"""
    c = np.cumsum(contrib_sorted)
    idx = np.nonzero(c > threshold if strict else c >= threshold)[0]
    return int(idx[0]) + 1 if idx.size else len(contrib_sorted)


def make_s2_eigenfunction(dets: np.ndarray, n_alpha: int) -> np.ndarray:
    """Determinants sharing an occupation pattern with dets and not in dets
    (occ_pattern.irp.f:251-309), from libqpx's host routine
    qpx_make_s2_eigenfunction.  dets (N, 2, n_int) uint64."""
    import ctypes as C
    from . import _lib
    L = _lib.load()
    d = np.ascontiguousarray(dets, dtype=np.uint64)
    n_int = d.shape[2]
    n = C.c_int64(0)
    cap = max(1024, 4 * d.shape[0])
    while True:
        out = np.empty((cap, 2, n_int), dtype=np.uint64)
        rc = L.qpx_make_s2_eigenfunction(n_int, d.shape[0], d.ctypes.data_as(_lib.c_u64p), int(n_alpha), cap,
                                         out.ctypes.data_as(_lib.c_u64p), C.byref(n))
        if rc == 2:
            cap = n.value
            continue
        if rc != 0:
            raise ValueError("qpx_make_s2_eigenfunction: bad arguments")
        return out[:n.value].copy()


@dataclass
class CipsiResult:
    dets: np.ndarray
    coef: np.ndarray
    energies: np.ndarray
    pt2: np.ndarray
    history: list = field(default_factory=list)
    # True when pt2 comes from a sweep over the proposed blocks only: it then
    # lacks the contribution of the blocks left out and is a lower bound of
    # the EN-PT2 magnitude, not the EN-PT2 itself
    pt2_partial: bool = False


def cipsi(ctx, H, n_det_max: int, n_states: int = 1, n_states_diag: int = 4, s2_eig: bool = True,
          threshold_davidson: float = 1e-10, threshold_generators: float = 0.99,
          threshold_selectors: float = 0.999, pt2_max: float = 1e-4, guess_dets=None, log=print,
          mixed: bool = False, proposer=None, full_sweep_every: int = 2, audit_proposal: bool = False,
          n_det_max_jacobi: int = 2000):
    """Run the CIPSI loop on the GPU context ctx from HF (or guess_dets)."""
    from .hamiltonian import hf_determinant
    M, na, nb = H.M, H.n_alpha, H.n_beta
    if guess_dets is None:
        dets = hf_determinant(M, na, nb)[None, :, :]
    else:
        dets = np.asarray(guess_dets, dtype=np.uint64)
    ctx.set_hamiltonian(H)
    coef = np.zeros((dets.shape[0], n_states))
    coef[0, 0] = 1.0
    history = []
    pt2 = np.full(n_states, -np.inf)
    pt2_partial = False
    energies = None

    def diagonalize(dets, coef):
        ctx.set_wavefunction(dets)
        n = dets.shape[0]
        if n <= n_det_max_jacobi or n < 3 * n_states_diag:
            # dense diagonalisation, as the reference's "Lapack" branch
            # (src/Davidson/diagonalize_CI.irp.f:70-145) for N_det <= N_det_max_jacobi
            Hm = np.zeros((n, n))
            Sm = np.zeros((n, n))
            for k0 in range(0, n, 8):
                k1 = min(n, k0 + 8)
                e = np.zeros((n, k1 - k0))
                e[np.arange(k0, k1), np.arange(k1 - k0)] = 1.0
                hv, sv = ctx.H_S2_u_0_nstates(e)
                Hm[:, k0:k1] = hv
                Sm[:, k0:k1] = sv
            w, V = np.linalg.eigh(0.5 * (Hm + Hm.T))
            s2v = np.einsum("ik,ij,jk->k", V, 0.5 * (Sm + Sm.T), V)
            order = np.arange(n)
            if s2_eig:
                sz = 0.5 * (na - nb)
                good = np.abs(s2v - sz * (sz + 1.0)) <= 0.5
                sel = np.nonzero(good)[0][:n_states]
                if sel.size:
                    order = np.concatenate([sel, np.array([j for j in range(n) if j not in set(sel.tolist())],
                                                          dtype=np.int64)])
            w, V = w[order], V[:, order]
            if n < n_states:
                # fewer determinants than states: the missing states get
                # zero vectors (no PT2) until psi has grown
                w = np.concatenate([w, np.full(n_states - n, w[-1])])
                V = np.concatenate([V, np.zeros((n, n_states - n))], axis=1)
            return w[:n_states], V[:, :n_states]
        g = np.zeros((n, n_states))
        g[:, :] = coef[:, :n_states]
        for k in range(n_states):
            if not np.any(g[:, k]):
                g[k % n, k] = 1.0
        e, s2, u, _ = ctx.davidson_diag_hs2(g, N_st=n_states, N_st_diag=max(n_states_diag, n_states),
                                            threshold=threshold_davidson, s2_eig=s2_eig, mixed=mixed)
        return e[:n_states], u[:, :n_states]

    while True:
        t0 = time.time()
        energies, coef = diagonalize(dets, coef)
        n_det = dets.shape[0]
        rec = dict(n_det=n_det, energies=energies.tolist(), pt2=None, t_diag=time.time() - t0)
        # the PT2 stop test needs the full EN-PT2: a proposed sweep's partial
        # sum could only understate it
        if n_det >= n_det_max or (not pt2_partial and np.max(np.abs(pt2)) <= pt2_max):
            history.append(rec)
            log(f"cipsi: N_det={n_det} E={energies} (stop)")
            break
        dets_s, coef_s, contrib = sort_by_weight(dets, coef)
        n_gen = n_leading(contrib, threshold_generators)
        n_sel = n_det if threshold_generators >= 1.0 else max(n_leading(contrib, threshold_selectors, True), n_gen)
        ctx.set_wavefunction(dets_s)
        E0 = np.array([ctx.u_0_H_u_0(coef_s[:, k])[0] if np.any(coef_s[:, k]) else energies[k]
                       for k in range(n_states)])
        to_select = (n_det // 2 + 1) if s2_eig else n_det
        to_select = min(to_select, n_det_max - n_det)
        t1 = time.time()
        if proposer is None:
            pt2, new_dets, vals = ctx.ZMQ_selection(np.asfortranarray(coef_s), n_gen, n_sel, E0, to_select)
        else:
            pt2, new_dets = _proposed_selection(ctx, H, proposer, dets_s, coef_s, n_gen, n_sel, E0, to_select,
                                                len(history), full_sweep_every, rec, log, audit_proposal)
        pt2_partial = rec.get("sweep") == "proposed"
        rec.update(pt2=pt2.tolist(), n_gen=n_gen, n_sel=n_sel, n_new=int(new_dets.shape[0]),
                   t_sel=time.time() - t1)
        # new determinants with zero coefficients, psi normalised
        dets = np.concatenate([dets_s, new_dets], axis=0)
        coef = np.concatenate([coef_s, np.zeros((new_dets.shape[0], n_states))], axis=0)
        if s2_eig or n_states > 1:
            extra = ctx.s2_complete(dets, na)
            if extra.shape[0]:
                dets = np.concatenate([dets, extra], axis=0)
                coef = np.concatenate([coef, np.zeros((extra.shape[0], n_states))], axis=0)
            rec["n_s2_added"] = int(extra.shape[0])
        nrm = np.linalg.norm(coef, axis=0)
        coef /= np.where(nrm > 0, nrm, 1.0)
        history.append(rec)
        log(f"cipsi: N_det={n_det} E={energies} PT2={pt2} gen={n_gen} sel={n_sel} -> {dets.shape[0]} "
            f"({rec['t_diag']:.1f}s diag, {rec['t_sel']:.1f}s sel)")
    return CipsiResult(dets=dets, coef=coef, energies=energies, pt2=pt2, history=history, pt2_partial=pt2_partial)


def _proposed_selection(ctx, H, proposer, dets_s, coef_s, n_gen, n_sel, E0, to_select, it, full_every, rec, log,
                        audit=False):
    """Selection restricted to the proposed blocks, or a full sweep that
    trains the proposer and measures what it would have missed.  A proposal
    that cannot supply to_select determinants falls back to the full sweep.
    audit: also run the full sweep on proposed iterations and record the
    PT2 fraction the proposal actually missed (for tests and measurement)."""
    from .proposal import block_descriptors
    desc, fdesc = block_descriptors(H, dets_s, coef_s, n_gen)
    cf = np.asfortranarray(coef_s)
    if proposer.ready and it % full_every != 0:
        keep = proposer.propose(desc, fdesc)
        pt2, new_dets, _, _, _ = ctx.select_blocks(cf, n_gen, n_sel, E0, to_select, keep=keep)
        if new_dets.shape[0] >= to_select:
            rec.update(sweep="proposed", kept=float(keep.mean()))
            if audit:
                pt2_full, _, _ = ctx.ZMQ_selection(cf, n_gen, n_sel, E0, to_select)
                rec["missed_pt2_fraction"] = float(1.0 - pt2[0] / pt2_full[0])
            return pt2, new_dets
        rec["fallback"] = True
    pt2, new_dets, _, bpt2, _ = ctx.select_blocks(cf, n_gen, n_sel, E0, to_select)
    rec["sweep"] = "full"
    if proposer.ready:
        keep = proposer.propose(desc, fdesc)
        tot = np.abs(bpt2[:, 0]).sum()
        rec["missed_pt2_fraction_if_kept"] = float(1.0 - np.abs(bpt2[keep == 1, 0]).sum() / max(tot, 1e-300))
        rec["would_keep"] = float(keep.mean())
    t = time.time()
    if proposer.fit(desc, fdesc, bpt2):
        rec["t_train"] = time.time() - t
        log(f"cipsi: full sweep, proposer trained on {desc.shape[0]} blocks ({rec['t_train']:.1f}s)"
            + (f"; kept on, it would have missed {rec['missed_pt2_fraction_if_kept']:.2e} of PT2 keeping "
               f"{rec['would_keep']:.3f}" if "missed_pt2_fraction_if_kept" in rec else ""))
    return pt2, new_dets
