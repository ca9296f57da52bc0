"""Stochastic PT2 on the GPU against the exact full selection.

With every determinant a generator and a selector, the EN-PT2 of a state
does not depend on the generator order (each external determinant is
counted once, by its first generator), so one full selection gives the
exact PT2 of every state.  The stochastic driver must give it back exactly
when forced to compute every generator, and within its error bar when it
stops early."""
from __future__ import annotations

import numpy as np
import pytest

from quantum_package_amd.cipsi import cipsi, sort_by_weight
from quantum_package_amd.hamiltonian import synthetic_hamiltonian
from quantum_package_amd.pt2_stoch import ZMQ_pt2
from test_gpu_sigma import _say

pytestmark = pytest.mark.gpu


def test_stochastic_pt2_against_the_full_selection():
    from quantum_package_amd.api import Context
    H = synthetic_hamiltonian(24, 5, 5, seed=606)
    ctx = Context(0)
    _say("[pt2_stoch] CIPSI M=24 5a5b, 2 states, to 1200 dets")
    r = cipsi(ctx, H, n_det_max=1200, n_states=2, s2_eig=True, log=_say)
    ds, cs, _ = sort_by_weight(r.dets, r.coef)
    n = ds.shape[0]
    ctx.set_wavefunction(ds)
    E0 = np.array([ctx.u_0_H_u_0(cs[:, k])[0] for k in range(2)])
    exact, _, _ = ctx.ZMQ_selection(np.asfortranarray(cs), n, n, E0, 1)
    _say(f"  n_det={n} exact PT2 = {exact}")

    pt2, err, info = ZMQ_pt2(ctx, ds, cs, E0, relative_error=0.0, absolute_error=0.0, seed=3)
    _say(f"  all generators: {pt2} {info}")
    assert all(i["generators_computed"] == n for i in info)
    assert np.all(err == 0.0)
    np.testing.assert_allclose(pt2, exact, rtol=1e-11)

    pt2, err, info = ZMQ_pt2(ctx, ds, cs, E0, relative_error=2e-3, absolute_error=0.0, seed=4,
                             combs_per_batch=8, log=_say)
    _say(f"  stochastic: {pt2} +/- {err} {info}")
    for k in range(2):
        assert info[k]["generators_computed"] < n
        assert 0.0 < err[k] < 5e-3 * abs(exact[k])
        assert abs(pt2[k] - exact[k]) < 5 * err[k]
    ctx.close()


def test_block_pt2_does_not_depend_on_the_other_blocks():
    """The stochastic estimate relies on a generator's exact PT2 being the
    same whatever else is enumerated in the call: block PT2 under a sparse
    keep mask equals, bit for bit, the same blocks of the full sweep, and a
    repeated sweep gives the same values."""
    from quantum_package_amd.api import Context
    H = synthetic_hamiltonian(24, 5, 5, seed=607)
    ctx = Context(0)
    r = cipsi(ctx, H, n_det_max=800, n_states=1, s2_eig=False, log=lambda m: None)
    ds, cs, _ = sort_by_weight(r.dets, r.coef)
    n = ds.shape[0]
    ctx.set_wavefunction(ds)
    E0 = ctx.u_0_H_u_0(cs[:, 0])
    cf = np.asfortranarray(cs)
    _say(f"[pt2_stoch] block PT2 under keep masks, n_det={n}")
    _, _, _, full, info = ctx.select_blocks(cf, n, n, E0, 1)
    _, _, _, again, _ = ctx.select_blocks(cf, n, n, E0, 1, info=False)
    assert np.array_equal(full, again)
    nbk = full.shape[0] // n
    rng = np.random.default_rng(5)
    for trial in range(3):
        gens = np.sort(rng.choice(np.arange(1, n + 1), size=max(3, n // 20), replace=False))
        n_gen = int(gens.max())
        keep = np.zeros((n_gen, nbk), dtype=np.uint8)
        keep[gens - 1] = 1
        _, _, _, part, _ = ctx.select_blocks(cf, n_gen, n, E0, 1, keep=keep.ravel(), info=False)
        sel = keep.ravel().astype(bool)
        assert np.array_equal(part[sel], full[: n_gen * nbk][sel])
        assert np.all(part[~sel] == 0.0)
        # per-generator entry point: each generator's blocks summed in block order
        order = rng.permutation(gens)
        got = ctx.select_generators(cf, n, E0, order)[:, 0]
        for g, v in zip(order, got):
            acc = 0.0
            for x in full[(g - 1) * nbk: g * nbk, 0]:
                acc += float(x)
            assert v == acc
    ctx.close()
