"""End-to-end CIPSI (fci_zmq loop) on the GPU against the same loop driven
by the CPU oracle: same growth, variational energies within 1e-8 Eh per
iteration.  PT2 is linear in the Davidson coefficients, which both sides
converge along different orthogonalisation paths; with the squared
(preconditioned) residual driven to 1e-13 PT2 agrees to 1e-6 Eh."""
from __future__ import annotations

import numpy as np
import pytest

from quantum_package_amd.cipsi import cipsi, make_s2_eigenfunction
from quantum_package_amd.hamiltonian import synthetic_hamiltonian
from oracle import oracle as O
from test_gpu_sigma import _say

pytestmark = pytest.mark.gpu


class OracleBackend:
    """The oracle behind the Context interface used by the CIPSI driver."""

    def set_hamiltonian(self, H):
        self.H = H
        self.ham = O.Ham(H)

    def set_wavefunction(self, dets):
        self.dets = np.ascontiguousarray(dets)
        self.bil = O.Bilinear(self.dets)

    def H_S2_u_0_nstates(self, u):
        u = np.asarray(u, dtype=np.float64)
        v, s = O.sigma(self.ham, self.bil, u if u.ndim == 2 else u[:, None])
        return (v, s) if u.ndim == 2 else (v[:, 0], s[:, 0])

    def u_0_H_u_0(self, u):
        v, _ = self.H_S2_u_0_nstates(u)
        return np.array([np.dot(v, u)])

    def davidson_diag_hs2(self, g, N_st, N_st_diag, threshold, s2_eig, mixed=False):
        e, s2, u, n = O.davidson(self.ham, self.bil, self.dets, g, n_st=N_st, n_st_diag=N_st_diag,
                                 threshold=threshold, s2_eig=s2_eig)
        return e, s2, u, (n, 0)

    def ZMQ_selection(self, coef, n_gen, n_sel, E0, N):
        r = O.selection(self.ham, self.dets, coef, n_gen, n_sel, E0, N)
        return r["pt2"], r["dets"], r["vals"]

    def s2_complete(self, dets, n_alpha):
        # the host routine, pinned by brute force in test_cpu_cipsi.py
        return make_s2_eigenfunction(dets, n_alpha)


@pytest.mark.parametrize("s2_eig", [False, True])
def test_cipsi_gpu_vs_oracle(s2_eig):
    from quantum_package_amd.api import Context
    H = synthetic_hamiltonian(12, 4, 4, seed=77)
    _say(f"[cipsi] M=12 4a4b s2_eig={s2_eig}")
    ctx = Context(0)
    rg = cipsi(ctx, H, n_det_max=600, s2_eig=s2_eig, threshold_davidson=1e-13, log=_say, n_det_max_jacobi=64)
    ro = cipsi(OracleBackend(), H, n_det_max=600, s2_eig=s2_eig, threshold_davidson=1e-13, log=_say,
               n_det_max_jacobi=64)
    ctx.close()
    assert len(rg.history) == len(ro.history)
    for a, b in zip(rg.history, ro.history):
        assert a["n_det"] == b["n_det"]
        assert abs(a["energies"][0] - b["energies"][0]) < 1e-8
        if a["pt2"] is not None:
            assert abs(a["pt2"][0] - b["pt2"][0]) < 1e-6
