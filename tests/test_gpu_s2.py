"""make_s2_eigenfunction on the GPU (qpx_s2_complete) against the host
routine qpx_make_s2_eigenfunction, which the CPU tests pin by brute force
over whole determinant spaces (occ_pattern.irp.f:251-309).  The two must
return the same determinants in the same order."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from quantum_package_amd import _lib
from quantum_package_amd.cipsi import make_s2_eigenfunction
from quantum_package_amd.hamiltonian import fci_space, hf_determinant, list_to_bitstring
from test_gpu_sigma import _say

pytestmark = pytest.mark.gpu


def _random_excited(M, na, nb, n, max_exc, seed):
    """n distinct determinants with up to max_exc excitations per spin from
    the aufbau determinant (orbitals 1..na, 1..nb)."""
    rng = np.random.default_rng(seed)
    nint = (M + 63) // 64
    seen, out = set(), []
    while len(out) < n:
        occ = []
        for ne in (na, nb):
            o = set(range(1, ne + 1))
            for _ in range(rng.integers(0, max_exc + 1)):
                o.remove(int(rng.choice(sorted(o))))
                o.add(int(rng.choice([x for x in range(ne + 1, M + 1) if x not in o])))
            occ.append(tuple(sorted(o)))
        key = tuple(occ)
        if key in seen:
            continue
        seen.add(key)
        out.append(np.stack([list_to_bitstring(list(occ[0]), nint), list_to_bitstring(list(occ[1]), nint)]))
    return np.stack(out)


@pytest.fixture(scope="module")
def ctx():
    from quantum_package_amd.api import Context
    c = Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("M,na,nb,n,max_exc,seed", [
    (8, 3, 3, 25, 3, 1),        # one word, random subset of a small space
    (12, 4, 3, 200, 2, 2),      # open-shell reference (na != nb)
    (70, 3, 3, 300, 2, 3),      # two words
    (96, 10, 10, 20000, 2, 4),  # the 96-orbital system's size class
    (130, 6, 5, 3000, 2, 5),    # three words
    (200, 5, 4, 3000, 2, 6),    # four words
])
def test_s2_complete_matches_host(ctx, M, na, nb, n, max_exc, seed):
    _say(f"[s2] M={M} {na}a{nb}b n={n} seed={seed}")
    if M == 8:
        full = fci_space(M, na, nb)
        psi = full[np.random.default_rng(seed).choice(full.shape[0], n, replace=False)]
    else:
        psi = _random_excited(M, na, nb, n, max_exc, seed)
    want = make_s2_eigenfunction(psi, na)
    got = ctx.s2_complete(psi, na)
    _say(f"  added {want.shape[0]}")
    assert got.shape == want.shape
    assert np.array_equal(got, want)


def test_s2_complete_edge_cases(ctx):
    M, na, nb = 10, 3, 3
    full = fci_space(M, na, nb)
    # a complete space adds nothing
    assert ctx.s2_complete(full, na).shape[0] == 0
    # closed shells only: nothing to add
    assert ctx.s2_complete(hf_determinant(M, na, nb)[None], na).shape[0] == 0
    # the count-only call, then a buffer one short of it
    psi = full[:7]
    L = _lib.load()
    d = np.ascontiguousarray(psi, dtype=np.uint64)
    n = C.c_int64(0)
    assert L.qpx_s2_complete(ctx.ptr, 1, d.shape[0], d.ctypes.data_as(_lib.c_u64p), na, 0, None, C.byref(n)) == 0
    want = make_s2_eigenfunction(psi, na)
    assert n.value == want.shape[0] > 0
    out = np.zeros((n.value, 2, 1), dtype=np.uint64)
    rc = L.qpx_s2_complete(ctx.ptr, 1, d.shape[0], d.ctypes.data_as(_lib.c_u64p), na, n.value - 1,
                           out.ctypes.data_as(_lib.c_u64p), C.byref(n))
    assert rc == 4 and n.value == want.shape[0] and not out.any()
    # bad arguments raise
    from quantum_package_amd.api import QpxError
    with pytest.raises(QpxError):
        ctx.s2_complete(np.zeros((2, 2, 5), dtype=np.uint64), na)
