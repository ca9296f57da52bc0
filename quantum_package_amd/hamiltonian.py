"""Synthetic Hamiltonians and determinant helpers.

The reference reads MO integrals from text files
(``plugins/read_integral/read_integrals_mo.irp.f:1-116``: ``kinetic_mo``,
``nuclear_mo`` as ``i j value``, ``bielec_mo`` as ``i j k l value``).  This
module produces the same inputs from a seed instead, so every machine can
rebuild a workload without files:

* a SplitMix64 stream, U(0,1) from the top 53 bits, Box-Muller normals
  (one normal per pair of uniforms, cosine branch);
* draw order: X (M x M, column-major), then Y_L (M x M, column-major) for
  L = 1..16;
* h_pq = eps_p delta_pq + (X_pq + X_qp)/2 with X ~ N(0, 0.05^2),
  eps_p = -2 + 0.5 p (p = 1..M);
* (pq|rs) = sum_L B^L_pq B^L_rs + 0.3 delta_pq delta_rs with
  B^L = (Y_L + Y_L^T)/2, Y ~ N(0, 0.1^2)  -- 8-fold symmetric and PSD.

Two-electron integrals are stored 8-fold packed exactly as the reference's
map keys them (``src/Integrals_Bielec/map_integrals.irp.f:19-34``): the
physicist integral <ij|kl> = (ik|jl) has key tri(tri(i,k), tri(j,l)) with
tri(a,b) = min + max(max-1)/2 on 1-based indices; the packed array holds
key-1.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

_MASK64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n consecutive SplitMix64 outputs starting from state ``seed``."""
    with np.errstate(over="ignore"):
        k = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def normals(seed: int, n: int) -> np.ndarray:
    """n standard normals: Box-Muller on consecutive uniform pairs."""
    u = (splitmix64(seed, 2 * n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    u1 = 1.0 - u[0::2]
    u2 = u[1::2]
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


def n_pairs(M: int) -> int:
    return M * (M + 1) // 2


def n_packed(M: int) -> int:
    p = n_pairs(M)
    return p * (p + 1) // 2


def tri1(a, b):
    """tri(a,b) on 1-based indices (map_integrals.irp.f:23-28)."""
    a = np.asarray(a, dtype=np.int64)
    b = np.asarray(b, dtype=np.int64)
    lo = np.minimum(a, b)
    hi = np.maximum(a, b)
    return lo + (hi * hi - hi) // 2


def integral_key(i, j, k, l):
    """bielec_integrals_index: key of <ij|kl> (1-based)."""
    return tri1(tri1(i, k), tri1(j, l))


@dataclass
class Hamiltonian:
    M: int
    n_alpha: int
    n_beta: int
    h: np.ndarray      # (M, M) float64, symmetric
    eri: np.ndarray    # (n_packed(M),) float64, packed by integral_key - 1
    seed: int = 0

    @property
    def n_int(self) -> int:
        return (self.M + 63) // 64

    def integral(self, i, j, k, l):
        """<ij|kl> with 1-based orbital indices (mo_bielec_integral)."""
        return self.eri[integral_key(i, j, k, l) - 1]

    def eri_dense_chemist(self) -> np.ndarray:
        """(pq|rs) as a dense M^4 array, 0-based (small M only)."""
        M = self.M
        idx = np.arange(1, M + 1)
        P = tri1(idx[:, None], idx[None, :])            # pair index of (p,q)
        key = tri1(P[:, :, None, None], P[None, None, :, :])
        return self.eri[key - 1]


def synthetic_hamiltonian(M: int, n_alpha: int, n_beta: int, seed: int) -> Hamiltonian:
    """The seeded synthetic Hamiltonian described in the module docstring."""
    nL = 16
    z = normals(seed, M * M * (1 + nL))
    X = 0.05 * z[: M * M].reshape(M, M, order="F")
    eps = -2.0 + 0.5 * np.arange(1, M + 1, dtype=np.float64)
    h = np.diag(eps) + 0.5 * (X + X.T)
    Y = 0.1 * z[M * M:].reshape(nL, M, M).transpose(0, 2, 1)  # each Y_L column-major
    B = 0.5 * (Y + Y.transpose(0, 2, 1))
    # pair-vectors over packed pair index P = tri(p,q) (1-based, p<=q)
    npair = n_pairs(M)
    # enumerate pairs in packed order: P = p + q(q-1)/2 for p<=q, 1-based
    pp = np.empty(npair, dtype=np.int64)
    qq = np.empty(npair, dtype=np.int64)
    for qi in range(1, M + 1):
        base = qi * (qi - 1) // 2
        for pi in range(1, qi + 1):
            pp[base + pi - 1] = pi
            qq[base + pi - 1] = qi
    Bp = B[:, pp - 1, qq - 1]                      # (nL, npair)
    delta = (pp == qq).astype(np.float64)
    V = Bp.T @ Bp + 0.3 * np.outer(delta, delta)    # (npair, npair) = (PQ|RS)
    # packed: key = tri(P, R) for P <= R -> R(R-1)/2 + P, so for each R the
    # keys of P = 1..R are contiguous (V is symmetric: row R-1 = column R-1)
    eri = np.empty(n_packed(M), dtype=np.float64)
    for c in range(npair):
        o = c * (c + 1) // 2
        eri[o:o + c + 1] = V[c, :c + 1]
    return Hamiltonian(M=M, n_alpha=n_alpha, n_beta=n_beta, h=np.ascontiguousarray(h), eri=eri, seed=seed)


# ----------------------------------------------------------------------------
# determinants: det(N_int, 2) int64 words, orbital p is bit (p-1)%64 of word
# (p-1)//64 (src/Bitmask/bitmasks_routines.irp.f:1-63)
# ----------------------------------------------------------------------------

def list_to_bitstring(occ, n_int: int) -> np.ndarray:
    s = np.zeros(n_int, dtype=np.uint64)
    for o in occ:
        s[(o - 1) // 64] |= np.uint64(1) << np.uint64((o - 1) % 64)
    return s


def bitstring_to_list(s) -> list:
    out = []
    for w, x in enumerate(np.asarray(s, dtype=np.uint64)):
        x = int(x)
        while x:
            b = (x & -x).bit_length() - 1
            out.append(b + 1 + 64 * w)
            x &= x - 1
    return out


def hf_determinant(M: int, n_alpha: int, n_beta: int) -> np.ndarray:
    """HF_bitmask: orbitals 1..n_alpha and 1..n_beta (bitmasks.irp.f:71-87)."""
    n_int = (M + 63) // 64
    d = np.zeros((2, n_int), dtype=np.uint64)
    d[0] = list_to_bitstring(range(1, n_alpha + 1), n_int)
    d[1] = list_to_bitstring(range(1, n_beta + 1), n_int)
    return d


def all_strings(M: int, n: int) -> np.ndarray:
    """All spin strings of n electrons in M orbitals (M <= 64), ascending."""
    from itertools import combinations
    assert M <= 64
    vals = []
    for c in combinations(range(M), n):
        v = 0
        for o in c:
            v |= 1 << o
        vals.append(v)
    vals.sort()
    return np.array(vals, dtype=np.uint64)


def fci_space(M: int, n_alpha: int, n_beta: int) -> np.ndarray:
    """Full-CI determinants as (N_det, 2, 1) uint64, alpha-major product order."""
    a = all_strings(M, n_alpha)
    b = all_strings(M, n_beta)
    dets = np.empty((a.size * b.size, 2, 1), dtype=np.uint64)
    dets[:, 0, 0] = np.repeat(a, b.size)
    dets[:, 1, 0] = np.tile(b, a.size)
    return dets
