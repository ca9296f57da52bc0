"""Block-proposal network: BF16 MFMA kernels against the float64 reference
(oracle/mlp_ref.py) and the learned proposal against exact selection.

Tolerances: the GEMM sees bf16-rounded inputs on both sides and
accumulates in fp32 on the GPU, so 1e-5 relative; gradients differ through
bf16 roundings of intermediate activations, 3e-2 relative (Frobenius) per
tensor.
"""
from __future__ import annotations

import numpy as np
import pytest

from quantum_package_amd.hamiltonian import synthetic_hamiltonian
from oracle import mlp_ref
from test_gpu_sigma import _say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mlp():
    from quantum_package_amd.api import BlockMLP
    m = BlockMLP(n_orb=40, seed=3)
    yield m
    m.close()


@pytest.mark.parametrize("M,N,K", [(16, 16, 16), (70, 90, 50), (256, 256, 256), (1000, 256, 80), (256, 80, 1000)])
def test_mfma_gemm_layout(mlp, M, N, K):
    rng = np.random.default_rng(M + N + K)
    A = rng.standard_normal((M, K)).astype(np.float32)
    B = rng.standard_normal((K, N)).astype(np.float32)
    _say(f"[mlp gemm] {M}x{N}x{K}")
    C = mlp.gemm_check(A, B)
    want = mlp_ref.bf16(A) @ mlp_ref.bf16(B)
    assert np.max(np.abs(C - want)) <= 1e-5 * np.max(np.abs(want)) + 1e-6


@pytest.mark.parametrize("M,N,K,split", [(70, 90, 50, 0), (64, 64, 32, 0), (256, 256, 256, 0), (1000, 256, 80, 0),
                                         (256, 256, 2048, 1), (100, 130, 700, 1), (256, 1, 700, 1)])
@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_mfma_gemm_operand_layouts(mlp, M, N, K, split, ta, tb):
    """Every operand layout the network uses (A^T for the weight gradients,
    B^T for the input gradients), aligned and unaligned leading dimensions,
    ragged tiles, and the K split with its ordered reduction."""
    rng = np.random.default_rng(7 * M + 5 * N + K + 3 * ta + tb)
    A = rng.standard_normal((M, K)).astype(np.float32)
    B = rng.standard_normal((K, N)).astype(np.float32)
    _say(f"[mlp gemm layout] {M}x{N}x{K} ta={ta} tb={tb} split={split}")
    C = mlp._gemm_layout(A, B, ta, tb, split)
    want = mlp_ref.bf16(A).astype(np.float64) @ mlp_ref.bf16(B).astype(np.float64)
    # fp32 accumulation over K terms of bf16 products
    assert np.max(np.abs(C - want)) <= 1e-6 * np.sqrt(K) * np.max(np.abs(want)) + 1e-5


def _random_batch(n, n_orb, rng):
    desc = np.zeros((n, 4), dtype=np.int32)
    desc[:, 0] = rng.integers(0, n_orb, n)
    desc[:, 1] = rng.integers(0, n_orb, n)
    desc[:, 2] = rng.integers(0, 3, n)
    desc[:, 3] = rng.integers(0, 5, n)
    fdesc = rng.standard_normal((n, 3)).astype(np.float32)
    y = rng.standard_normal(n).astype(np.float32)
    return desc, fdesc, y


def test_gradients_vs_reference(mlp):
    rng = np.random.default_rng(5)
    P = {k: mlp.get(k).astype(np.float64) for k in ["W1", "W2", "W3", "b1", "b2", "b3", "w4", "emb", "b4"]}
    P["b1"] = rng.standard_normal(256) * 0.1
    mlp.set("b1", P["b1"])
    P["b4"] = np.array([0.3])
    mlp.set("b4", P["b4"])
    P["b4"] = float(P["b4"][0])
    desc, fdesc, y = _random_batch(700, 40, rng)
    _say("[mlp grads] n=700")
    loss = mlp.train_step(desc, fdesc, y, lr=0.0)
    yr, lr_, g = mlp_ref.forward_backward(P, desc, fdesc.astype(np.float64), y.astype(np.float64))
    assert abs(loss - lr_) <= 1e-2 * abs(lr_)
    pred = mlp.predict(desc, fdesc)
    assert np.max(np.abs(pred - yr)) <= 2e-2 * np.max(np.abs(yr)) + 1e-3
    for k in ["W1", "W2", "W3", "b1", "b2", "b3", "w4", "emb", "b4"]:
        gg = mlp.get(k, grad=True).astype(np.float64).reshape(np.shape(g[k]))
        rel = np.linalg.norm(gg - g[k]) / (np.linalg.norm(g[k]) + 1e-30)
        _say(f"  grad {k}: rel err {rel:.2e}")
        assert rel < 3e-2, k


def test_train_epochs_matches_single_steps():
    """The device-resident epoch loop (batches gathered on the device, the
    output bias updated on the device, losses read once per epoch) against
    the same full-batch Adam steps taken one call at a time.  Only the order
    of the batch rows differs, so the sums round differently: 1e-4 on the
    losses, 1e-3 on each parameter tensor."""
    from quantum_package_amd.api import BlockMLP
    rng = np.random.default_rng(21)
    desc, fdesc, y = _random_batch(512, 30, rng)
    a = BlockMLP(n_orb=30, seed=4)
    b = BlockMLP(n_orb=30, seed=4)
    _say("[mlp epochs] 3 full-batch epochs of 512 against 3 single steps")
    la = a.train_epochs(desc, fdesc, y, 1e-3, 3, 512, seed=1)
    lb = np.array([b.train_step(desc, fdesc, y, lr=1e-3) for _ in range(3)], dtype=np.float32)
    np.testing.assert_allclose(la, lb, rtol=1e-4)
    for k in ["W1", "W2", "W3", "b1", "b2", "b3", "w4", "emb", "b4"]:
        pa, pb = a.get(k), b.get(k)
        assert np.linalg.norm(pa - pb) <= 1e-3 * np.linalg.norm(pb) + 1e-6, k
    # three Adam steps of lr = 1e-3 from zero move the bias, by at most ~3 lr
    assert 0.0 < abs(float(a.get("b4").ravel()[0])) < 4e-3
    a.close()
    b.close()


def test_training_reduces_loss():
    from quantum_package_amd.api import BlockMLP
    rng = np.random.default_rng(9)
    m = BlockMLP(n_orb=30, seed=11)
    desc, fdesc, _ = _random_batch(8192, 30, rng)
    # a learnable target: smooth function of the descriptors
    y = (np.sin(desc[:, 0] / 5.0) + 0.5 * np.cos(desc[:, 1] / 3.0) + 0.3 * fdesc[:, 2] - 0.2 * desc[:, 2]).astype(
        np.float32)
    _say("[mlp train] 8192 samples")
    l0 = m.train_step(desc, fdesc, y, lr=0.0)
    for ep in range(40):
        for b in range(0, 8192, 1024):
            m.train_step(desc[b:b + 1024], fdesc[b:b + 1024], y[b:b + 1024], lr=2e-3)
    l1 = m.train_step(desc, fdesc, y, lr=0.0)
    _say(f"  loss {l0:.4f} -> {l1:.4f}")
    m.close()
    assert l1 < 0.1 * l0


def test_out_of_range_descriptors_are_rejected(mlp):
    """A descriptor row outside the embedding table or the one-hot slots is
    an error returned before anything reaches the device."""
    from quantum_package_amd.api import QpxError
    rng = np.random.default_rng(2)
    n = 64
    desc = np.stack([rng.integers(0, 40, n), rng.integers(0, 40, n), rng.integers(0, 3, n),
                     rng.integers(0, 5, n)], axis=1).astype(np.int32)
    fdesc = rng.standard_normal((n, 3)).astype(np.float32)
    y = rng.standard_normal(n).astype(np.float32)
    _say("[mlp] out-of-range descriptors")
    before = mlp.get("W1").copy()
    for col, bad in ((0, 40), (1, -1), (2, 3), (3, 5)):
        d = desc.copy()
        d[7, col] = bad
        with pytest.raises(QpxError):
            mlp.train_step(d, fdesc, y, 1e-3)
        with pytest.raises(QpxError):
            mlp.predict(d, fdesc)
        with pytest.raises(QpxError):
            mlp.train_epochs(d, fdesc, y, 1e-3, 1, 16, 0)
    assert np.array_equal(mlp.get("W1"), before)
    assert np.isfinite(mlp.predict(desc, fdesc)).all()
