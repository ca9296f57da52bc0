"""MO integrals as the text triplet ``kinetic_mo``, ``nuclear_mo``, ``bielec_mo``,
and the two-electron integral work file ``work/mo_ints`` (at the end).

The reference's ``read_integrals`` program
(``plugins/read_integral/read_integrals_mo.irp.f:1-116``) is its route for
integrals that come from outside: list-directed text, ``i j value`` for the
one-electron files and ``i j k l value`` (physicists' <ij|kl>) for the
two-electron file.  This module reads the same files into a
:class:`~quantum_package_amd.hamiltonian.Hamiltonian` (whose packed array the
C-ABI takes as is) and writes them back, with the reference's semantics:

* the orbital count is the largest first index in ``nuclear_mo`` (:18-25;
  the EZFIO ``mo_tot_num`` wins when the caller passes ``mo_num``);
* one array A is zeroed once, filled from ``kinetic_mo``, then overwritten
  entry by entry from ``nuclear_mo`` (:65-84), so ``mo_ne_integral`` keeps the
  kinetic value wherever ``nuclear_mo`` lists no entry; the one-electron
  Hamiltonian is the sum of the two (no pseudopotential);
* every ``bielec_mo`` line goes to key tri(tri(i,k), tri(j,l))
  (``bielec_integrals_index``), the map is sorted and ``map_unique`` keeps one
  value per key (:98-108, ``src/Utils/map_module.f90:301-321``); which of
  several lines for one key survives depends on the reference's sort, here
  it is the first line in the file; keys never listed are zero.
"""
from __future__ import annotations

import os

import numpy as np

from .hamiltonian import Hamiltonian, integral_key, n_packed, n_pairs, tri1

FILES = ("kinetic_mo", "nuclear_mo", "bielec_mo")


def _records(path: str, n_int: int) -> tuple[np.ndarray, np.ndarray]:
    """List-directed records of n_int integers and one real per line (extra
    values on a line are ignored, Fortran D exponents accepted).  Returns
    (int64 (n, n_int), float64 (n,))."""
    with open(path, "r") as f:
        text = f.read().replace("D", "E").replace("d", "e")
    lines = [ln for ln in text.splitlines() if ln.strip()]
    ncol = n_int + 1
    if not lines:
        return np.zeros((0, n_int), dtype=np.int64), np.zeros(0)
    tok = np.array([len(ln.split()) for ln in lines])
    if np.any(tok < ncol):
        bad = int(np.argmax(tok < ncol))
        raise ValueError(f"{path}: line {bad + 1} has fewer than {ncol} values")
    if not np.all(tok == ncol):
        lines = [" ".join(ln.split()[:ncol]) for ln in lines]
    # indices parse exactly as doubles (< 2^53)
    rows = np.fromstring(" ".join(lines), sep=" ").reshape(-1, ncol)
    idx = rows[:, :n_int].astype(np.int64)
    if not np.array_equal(idx, rows[:, :n_int]):
        raise ValueError(f"{path}: non-integer orbital index")
    return idx, rows[:, n_int].copy()


def _assign_last(A: np.ndarray, i: np.ndarray, j: np.ndarray, v: np.ndarray) -> None:
    """A[i, j] = v line by line: where a pair repeats, the last line wins."""
    lin = i * A.shape[1] + j
    rev = lin[::-1]
    _, pos = np.unique(rev, return_index=True)
    last = len(lin) - 1 - pos
    A.ravel()[lin[last]] = v[last]


def read_integrals_mo(directory: str, n_alpha: int, n_beta: int, mo_num: int | None = None,
                      seed: int = 0) -> Hamiltonian:
    """Hamiltonian from ``directory``/{kinetic_mo, nuclear_mo, bielec_mo}
    (read_integrals_mo.irp.f:1-116)."""
    kin_i, kin_v = _records(os.path.join(directory, "kinetic_mo"), 2)
    nuc_i, nuc_v = _records(os.path.join(directory, "nuclear_mo"), 2)
    if mo_num is None:
        if nuc_i.shape[0] == 0:
            raise ValueError("nuclear_mo is empty: the orbital count cannot be derived")
        mo_num = int(nuc_i[:, 0].max())
    M = int(mo_num)
    for name, ii in (("kinetic_mo", kin_i), ("nuclear_mo", nuc_i)):
        if ii.size and (ii.min() < 1 or ii.max() > M):
            raise ValueError(f"{name}: orbital index outside 1..{M}")
    A = np.zeros((M, M))
    _assign_last(A, kin_i[:, 0] - 1, kin_i[:, 1] - 1, kin_v)
    kinetic = A.copy()
    _assign_last(A, nuc_i[:, 0] - 1, nuc_i[:, 1] - 1, nuc_v)
    nuclear = A
    h = kinetic + nuclear

    bi_i, bi_v = _records(os.path.join(directory, "bielec_mo"), 4)
    eri = np.zeros(n_packed(M))
    if bi_i.shape[0]:
        if bi_i.min() < 1 or bi_i.max() > M:
            raise ValueError(f"bielec_mo: orbital index outside 1..{M}")
        key = integral_key(bi_i[:, 0], bi_i[:, 1], bi_i[:, 2], bi_i[:, 3])
        # one value per key: the first line that names it
        uk, first = np.unique(key, return_index=True)
        eri[uk - 1] = bi_v[first]
    if not (0 <= n_beta <= n_alpha <= M):
        raise ValueError("need 0 <= n_beta <= n_alpha <= mo_num")
    return Hamiltonian(M=M, n_alpha=n_alpha, n_beta=n_beta, h=np.ascontiguousarray(h), eri=eri, seed=seed)


def canonical_indices(M: int) -> np.ndarray:
    """(n_packed(M), 4) int64: for each packed key (in key order) the 1-based
    (i, j, k, l) with <ij|kl> on that key, i <= k, j <= l, tri(i,k) <= tri(j,l)."""
    npair = n_pairs(M)
    lo = np.empty(npair, dtype=np.int64)
    hi = np.empty(npair, dtype=np.int64)
    for q in range(1, M + 1):
        base = q * (q - 1) // 2
        lo[base:base + q] = np.arange(1, q + 1)
        hi[base:base + q] = q
    # key = P + R(R-1)/2 with P <= R (pair indices, 1-based)
    R = np.repeat(np.arange(1, npair + 1, dtype=np.int64), np.arange(1, npair + 1))
    P = np.arange(1, n_packed(M) + 1, dtype=np.int64) - R * (R - 1) // 2
    i, k = lo[P - 1], hi[P - 1]
    j, l = lo[R - 1], hi[R - 1]
    out = np.stack([i, j, k, l], axis=1)
    assert np.array_equal(tri1(tri1(i, k), tri1(j, l)), np.arange(1, n_packed(M) + 1))
    return out


def write_integrals_mo(H: Hamiltonian, directory: str, zero_threshold: float = 0.0) -> None:
    """Write H as the text triplet: the whole one-electron matrix in
    ``kinetic_mo``, zeros in ``nuclear_mo`` (every entry listed, so the
    reader's orbital count and overwrite rule give h back exactly), one
    ``bielec_mo`` line per packed key with |value| > zero_threshold.  Values
    are written with 17 significant digits, enough to read back the same
    doubles."""
    os.makedirs(directory, exist_ok=True)
    M = H.M
    ii, jj = np.meshgrid(np.arange(1, M + 1), np.arange(1, M + 1), indexing="ij")
    with open(os.path.join(directory, "kinetic_mo"), "w") as f:
        for a, b, v in zip(ii.ravel(), jj.ravel(), H.h.ravel()):
            f.write(f"{a} {b} {v:.17g}\n")
    with open(os.path.join(directory, "nuclear_mo"), "w") as f:
        for a, b in zip(ii.ravel(), jj.ravel()):
            f.write(f"{a} {b} 0.0\n")
    idx = canonical_indices(M)
    keep = np.abs(H.eri) > zero_threshold
    with open(os.path.join(directory, "bielec_mo"), "w") as f:
        for (i, j, k, l), v in zip(idx[keep], H.eri[keep]):
            f.write(f"{i} {j} {k} {l} {v:.17g}\n")


# ---------------------------------------------------------------------------
# The integral work file: the consolidated map that the reference writes to
# <ezfio>/work/mo_ints (src/Integrals_Bielec/mo_bi_integrals.irp.f:35,149;
# src/Utils/map_functions.irp.f:1-131).  A key (bielec_integrals_index) goes
# to bucket key >> 15 (map_shift) and is stored there as its low 15 bits, an
# integer*2 (map_module.f90:26-27, map_append :489-517); map_size = keymax >>
# 15 with keymax the key of <MM|MM> (map_integrals.irp.f:286-294).  Three
# files of raw native-endian arrays:
#   <base>_consolidated_idx    int64 (map_size + 2): 1-based start of each
#                              bucket's entries, then n_elements + 1;
#   <base>_consolidated_key    int16 (n_elements): the low 15 bits, ascending
#                              within a bucket (map_sort before the save);
#   <base>_consolidated_value  float64 (n_elements).
# Keys absent from the map are zero integrals.

MAP_SHIFT = 15


def _map_paths(base: str) -> tuple[str, str, str]:
    return base + "_consolidated_idx", base + "_consolidated_key", base + "_consolidated_value"


def write_mo_ints(eri: np.ndarray, base: str, zero_threshold: float = 0.0) -> int:
    """Write packed MO integrals (element key-1 holds <ij|kl>) as the
    consolidated map files at base (e.g. <ezfio>/work/mo_ints).  Entries with
    |value| <= zero_threshold are left out, as the reference keeps only the
    integrals above its threshold.  Returns the number of stored entries."""
    eri = np.ascontiguousarray(eri, dtype=np.float64)
    n8 = eri.size
    map_size = n8 >> MAP_SHIFT
    keys = np.nonzero(np.abs(eri) > zero_threshold)[0].astype(np.int64) + 1
    counts = np.bincount(keys >> MAP_SHIFT, minlength=map_size + 1)
    idx = np.empty(map_size + 2, dtype="<i8")
    idx[0] = 1
    idx[1:] = 1 + np.cumsum(counts)
    d = os.path.dirname(base)
    if d:
        os.makedirs(d, exist_ok=True)
    p_idx, p_key, p_val = _map_paths(base)
    idx.tofile(p_idx)
    (keys & ((1 << MAP_SHIFT) - 1)).astype("<i2").tofile(p_key)
    eri[keys - 1].astype("<f8").tofile(p_val)
    return int(keys.size)


def read_mo_ints(base: str, mo_num: int) -> np.ndarray:
    """Packed MO integrals (n_packed(mo_num) doubles, element key-1) from the
    consolidated map files at base, with map_load_from_disk's checks: the
    index file must hold map_size + 2 entries, a cache key must not be
    negative and a value must not exceed 1e30 ('Error in integrals file')."""
    n8 = n_packed(mo_num)
    map_size = n8 >> MAP_SHIFT
    p_idx, p_key, p_val = _map_paths(base)
    idx = np.fromfile(p_idx, dtype="<i8")
    if idx.size != map_size + 2:
        raise ValueError(f"{p_idx}: {idx.size} entries, {map_size + 2} expected for mo_num = {mo_num}")
    n_el = int(idx[-1] - 1)
    if idx[0] != 1 or n_el < 0 or np.any(np.diff(idx) < 0):
        raise ValueError(f"{p_idx}: bucket starts are not 1-based and non-decreasing")
    ckey = np.fromfile(p_key, dtype="<i2", count=n_el)
    vals = np.fromfile(p_val, dtype="<f8", count=n_el)
    if ckey.size != n_el or vals.size != n_el:
        raise ValueError(f"{base}: fewer than the {n_el} entries its index announces")
    if np.any(ckey < 0) or np.any(vals > 1.0e30):
        raise ValueError(f"{base}: error in integrals file")
    bucket = np.repeat(np.arange(map_size + 1, dtype=np.int64), np.diff(idx))
    keys = (bucket << MAP_SHIFT) | ckey.astype(np.int64)
    if np.any(keys < 1) or np.any(keys > n8):
        raise ValueError(f"{base}: a key lies outside 1..{n8}")
    eri = np.zeros(n8)
    eri[keys - 1] = vals
    return eri
