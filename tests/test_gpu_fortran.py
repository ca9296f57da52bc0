"""The Fortran binding (quantum_package_amd/fortran/qpx_mod.f90): a plain
Fortran host program calls qpx_H_S2_u_0_nstates and qpx_davidson_diag with
the reference's array layouts; results are checked against the oracle."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from quantum_package_amd.hamiltonian import fci_space, synthetic_hamiltonian
from oracle import oracle as O
from test_gpu_sigma import _say

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "quantum_package_amd", "qpx_sigma_driver")


def test_fortran_host_sigma_and_davidson(tmp_path):
    if not os.path.exists(DRIVER):
        pytest.skip("Fortran driver not built (amdflang missing at build time)")
    H = synthetic_hamiltonian(8, 4, 4, seed=1001)
    dets = fci_space(8, 4, 4)
    n = dets.shape[0]
    rng = np.random.default_rng(2)
    u = rng.standard_normal((n, 2))
    u[:, 0] = 0.0
    hf = np.where((dets[:, 0, 0] == 15) & (dets[:, 1, 0] == 15))[0][0]
    u[hf, 0] = 1.0
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        np.array([8, 4, 4, 1], dtype=np.int32).tofile(f)
        np.array([n], dtype=np.int64).tofile(f)
        np.array([2], dtype=np.int32).tofile(f)
        np.array([H.eri.size], dtype=np.int64).tofile(f)
        np.asfortranarray(H.h).ravel(order="F").tofile(f)
        H.eri.tofile(f)
        np.ascontiguousarray(dets).tofile(f)          # det(N_int,2,N_det) Fortran order
        u.ravel(order="F").tofile(f)
    _say("[fortran] qpx_sigma_driver on FCI M=8 4a4b")
    r = subprocess.run([DRIVER, str(fin), str(fout)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = np.fromfile(fout, dtype=np.uint8)
    out = raw[: 8 * (4 * n + 1)].view(np.float64)
    tail = raw[8 * (4 * n + 1):].view(np.int64)
    v = out[: 2 * n].reshape((n, 2), order="F")
    s = out[2 * n: 4 * n].reshape((n, 2), order="F")
    e = out[4 * n]
    ham = O.Ham(H)
    bil = O.Bilinear(dets)
    vo, so = O.sigma(ham, bil, u)
    assert np.max(np.abs(v - vo)) < 1e-11
    assert np.max(np.abs(s - so)) < 1e-11
    eo, _, _, _ = O.davidson(ham, bil, dets, u[:, 0], n_st=1, n_st_diag=4, threshold=1e-10)
    assert abs(e - eo[0]) < 1e-9
    # S^2 completion of the first third of psi, called from Fortran with
    # det(N_int,2,N) arrays: the multiplet partners missing from it, each
    # sharing an occupation pattern with it
    from quantum_package_amd.cipsi import make_s2_eigenfunction
    part = dets[: n // 3]
    want = make_s2_eigenfunction(part, 4)
    n_new = int(tail[0])
    got = tail[1: 1 + 2 * n_new].astype(np.uint64).reshape(n_new, 2, 1)
    pt2_f = tail[1 + 2 * n_new:].view(np.float64)
    assert n_new == want.shape[0] > 0
    assert np.array_equal(got, want)
    pats = {(int(d[0, 0]) & int(d[1, 0]), int(d[0, 0]) ^ int(d[1, 0])) for d in part}
    assert all((int(d[0, 0]) & int(d[1, 0]), int(d[0, 0]) ^ int(d[1, 0])) in pats for d in got)
    # per-generator PT2 of that third (u(:,1) as coefficients) through the
    # Fortran wrapper, against the Python entry point
    from quantum_package_amd.api import Context
    ctx = Context(0)
    ctx.set_hamiltonian(H)
    ctx.set_wavefunction(part)
    c = u[: n // 3, :1]
    hc, _ = ctx.H_S2_u_0_nstates(c)
    e0 = float(c[:, 0] @ hc[:, 0] / (c[:, 0] @ c[:, 0]))
    want_pt2 = ctx.select_generators(c, n // 3, np.array([e0]), [1, 2, 3, 4])[:, 0]
    ctx.close()
    assert pt2_f.shape == (4,)
    np.testing.assert_allclose(pt2_f, want_pt2, rtol=1e-12, atol=1e-15)
    assert np.any(pt2_f != 0.0)
