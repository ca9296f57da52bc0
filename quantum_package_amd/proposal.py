"""Learned block proposal for CIPSI selection.

The exact selection (select_singles_and_doubles,
plugins/Full_CI_ZMQ/selection.irp.f:258-578) spends its time in
(generator, hole-pair) blocks; the reference already lets a caller run a
subset of them (`subset` / fragment_count, :561-562).  Here a small MLP
(BlockMLP, BF16 MFMA kernels) learns, from the per-block EN-PT2 sums that
the exact selection itself produces, which blocks carry the PT2; only the
blocks it proposes are enumerated, exactly.  A periodic full sweep measures
the PT2 the proposal missed, and the full enumeration stays the fallback.
"""
from __future__ import annotations

import numpy as np

LABEL_SPAN = 15.0
CALIBRATION_TARGET = 5e-4


def block_order(na: int, nb: int):
    """(s1, i1, s2, i2) of every block of one generator, in the reference's
    loop order (selection.irp.f:412-420): s1 = 1,2; i1 descending; s2 >= s1;
    i2 descending down to i1+1 (same spin) or 1.  i1/i2 index the
    generator's sorted hole list of that spin (1-based)."""
    out = []
    nh = {1: na, 2: nb}
    for s1 in (1, 2):
        for i1 in range(nh[s1], 0, -1):
            for s2 in range(s1, 3):
                ib = i1 + 1 if s1 == s2 else 1
                for i2 in range(nh[s2], ib - 1, -1):
                    out.append((s1, i1, s2, i2))
    return np.array(out, dtype=np.int64)


def occupied_lists(dets: np.ndarray, n_occ: int, spin: int) -> np.ndarray:
    """sorted 1-based occupied orbitals (n_det, n_occ) of one spin"""
    n, _, n_int = dets.shape
    out = np.zeros((n, n_occ), dtype=np.int64)
    cur = np.zeros(n, dtype=np.int64)
    one = np.uint64(1)
    from .cipsi import popcount64
    for w in range(n_int):
        x = dets[:, spin, w].copy()
        while True:
            nz = x != 0
            if not nz.any():
                break
            with np.errstate(over="ignore"):
                lb = x & (~x + one)
                p = popcount64(lb - one) + 64 * w + 1
                rows = np.nonzero(nz)[0]
                out[rows, cur[rows]] = p[rows]
                cur[rows] += 1
                x = np.where(nz, x & (x - one), x)
    return out


def block_descriptors(H, dets_sorted: np.ndarray, coef_sorted: np.ndarray, n_gen: int):
    """desc (n_blocks, 4) int32 (h1, h2 0-based, spin class, degree) and
    fdesc (n_blocks, 3) float32 (h_11, h_22, log|c|) for the blocks of the
    first n_gen generators, in qpx_select_blocks order."""
    from .hamiltonian import hf_determinant
    from .cipsi import popcount64
    na, nb = H.n_alpha, H.n_beta
    gens = dets_sorted[:n_gen]
    occ = {1: occupied_lists(gens, na, 0), 2: occupied_lists(gens, nb, 1)}
    pat = block_order(na, nb)
    nbk = pat.shape[0]
    s1, i1, s2, i2 = pat[:, 0], pat[:, 1], pat[:, 2], pat[:, 3]
    h1 = np.empty((n_gen, nbk), dtype=np.int64)
    h2 = np.empty((n_gen, nbk), dtype=np.int64)
    for s in (1, 2):
        m1 = s1 == s
        h1[:, m1] = occ[s][:, i1[m1] - 1]
        m2 = s2 == s
        h2[:, m2] = occ[s][:, i2[m2] - 1]
    spin = np.where(s1 != s2, 2, s1 - 1)
    hf = hf_determinant(H.M, na, nb)
    deg = (popcount64(gens[:, 0, :] & ~hf[0][None, :]).sum(axis=1) +
           popcount64(gens[:, 1, :] & ~hf[1][None, :]).sum(axis=1))
    deg = np.minimum(deg, 4)
    logc = np.log(np.abs(coef_sorted[:n_gen, 0]) + 1e-300)
    desc = np.empty((n_gen * nbk, 4), dtype=np.int32)
    desc[:, 0] = (h1 - 1).ravel()
    desc[:, 1] = (h2 - 1).ravel()
    desc[:, 2] = np.tile(spin, n_gen)
    desc[:, 3] = np.repeat(deg, nbk)
    hd = np.diag(H.h)
    fdesc = np.empty((n_gen * nbk, 3), dtype=np.float32)
    fdesc[:, 0] = hd[desc[:, 0]]
    fdesc[:, 1] = hd[desc[:, 1]]
    fdesc[:, 2] = np.repeat(logc, nbk)
    return desc, fdesc


def labels_from_blocks(block_pt2: np.ndarray, span: float = LABEL_SPAN) -> np.ndarray:
    """log|e| of each block (state 0), floored at max - span: blocks more
    than e^span below the largest carry no PT2 that matters, and the floor
    keeps the regression on the range where the ranking is decided."""
    y = np.log(np.abs(block_pt2[:, 0]) + 1e-300)
    return np.maximum(y, y.max() - span).astype(np.float32)


class Trainer:
    """Standardised regression of block labels with Adam on the GPU."""

    def __init__(self, mlp, lr=1e-3, batch=4096, seed=0):
        self.mlp = mlp
        self.lr = lr
        self.batch = batch
        self.rng = np.random.default_rng(seed)
        self.mu = None
        self.sd = None
        self.fmu = None
        self.fsd = None

    def _norm_f(self, fdesc):
        return ((fdesc - self.fmu) / self.fsd).astype(np.float32)

    def fit(self, desc, fdesc, y, epochs=3, log=None):
        if self.mu is None:
            self.mu, self.sd = float(y.mean()), float(y.std() + 1e-6)
            self.fmu = fdesc.mean(axis=0)
            self.fsd = fdesc.std(axis=0) + 1e-6
        t = ((y - self.mu) / self.sd).astype(np.float32)
        f = self._norm_f(fdesc)
        n = desc.shape[0]
        if n >= self.batch:
            # the whole training set on the device, batches gathered there
            seed = int(self.rng.integers(1 << 62))
            losses = self.mlp.train_epochs(desc, f, t, self.lr, epochs, self.batch, seed)
            if log:
                for ep, l in enumerate(losses):
                    log(f"mlp epoch {ep}: loss {l:.4f}")
            return list(losses)
        losses = []
        for ep in range(epochs):
            losses.append(self.mlp.train_step(desc, f, t, self.lr))
            if log:
                log(f"mlp epoch {ep}: loss {losses[-1]:.4f}")
        return losses

    def predict(self, desc, fdesc):
        return self.mlp.predict(desc, self._norm_f(fdesc)) * self.sd + self.mu


def calibrate_threshold(pred_log: np.ndarray, block_pt2: np.ndarray, target: float = CALIBRATION_TARGET) -> float:
    """Threshold tau on the predicted log|e| of a fully swept iteration such
    that the blocks predicted below tau carry at most `target` of its exact
    PT2.  Applied to the next iterations it misses several times `target`
    (psi moves on), hence a target well below the allowed missed fraction."""
    e = np.abs(block_pt2[:, 0])
    order = np.argsort(pred_log, kind="stable")
    cum = np.cumsum(e[order]) / max(e.sum(), 1e-300)
    j = int(np.searchsorted(cum, target, side="right"))
    return float(pred_log[order[j]]) if j < order.size else float("inf")


def choose_blocks(pred_log: np.ndarray, tau: float) -> np.ndarray:
    """keep mask (uint8, one per block): predicted log|e| >= tau"""
    return (pred_log >= tau).astype(np.uint8)


class Proposer:
    """Train on a full sweep (exact block PT2), propose blocks afterwards."""

    def __init__(self, mlp, epochs: int = 30, target: float = CALIBRATION_TARGET, lr: float = 1e-3,
                 batch: int = 2048, seed: int = 0, min_blocks: int = 4096, max_samples: int = 262144):
        self.trainer = Trainer(mlp, lr=lr, batch=batch, seed=seed)
        self.epochs = epochs
        self.target = target
        self.min_blocks = min_blocks
        self.max_samples = max_samples
        self.rng = np.random.default_rng(seed + 1)
        self.tau = None

    @property
    def ready(self) -> bool:
        return self.tau is not None

    def fit(self, desc, fdesc, block_pt2, log=None) -> bool:
        """Train on one full sweep; sweeps smaller than min_blocks are too
        small to learn from (and to standardise the labels on) and are
        skipped.  Returns whether it trained."""
        if desc.shape[0] < self.min_blocks:
            return False
        y = labels_from_blocks(block_pt2)
        # training on a uniform sample of at most max_samples blocks; the
        # threshold is calibrated on all of them
        idx = np.arange(desc.shape[0])
        if idx.size > self.max_samples:
            idx = np.sort(self.rng.choice(idx.size, self.max_samples, replace=False))
        self.trainer.fit(desc[idx], fdesc[idx], y[idx], epochs=self.epochs, log=log)
        self.tau = calibrate_threshold(self.trainer.predict(desc, fdesc), block_pt2, self.target)
        return True

    def propose(self, desc, fdesc) -> np.ndarray:
        return choose_blocks(self.trainer.predict(desc, fdesc), self.tau)
