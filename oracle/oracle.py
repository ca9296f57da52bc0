"""ctypes front end of the CPU oracle (test infrastructure only).

Only the test suite, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this module.  The product (libqpx) never calls it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

c_dp = C.POINTER(C.c_double)
c_u64p = C.POINTER(C.c_uint64)
c_ip = C.POINTER(C.c_int)


def build() -> None:
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.oq_ham_create.restype = C.c_void_p
        L.oq_ham_create.argtypes = [C.c_int, C.c_int, C.c_int, c_dp, c_dp, C.c_int64]
        L.oq_ham_destroy.argtypes = [C.c_void_p]
        L.oq_get_integral.restype = C.c_double
        L.oq_get_integral.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
        L.oq_ref_energy.restype = C.c_double
        L.oq_ref_energy.argtypes = [C.c_void_p]
        L.oq_get_fock_cs.argtypes = [C.c_void_p, c_dp]
        L.oq_excitation_degree.restype = C.c_int
        L.oq_excitation_degree.argtypes = [c_u64p, c_u64p, C.c_int]
        L.oq_get_excitation.argtypes = [c_u64p, c_u64p, c_ip, c_ip, c_dp, C.c_int]
        for f in ("oq_i_H_j",):
            getattr(L, f).restype = C.c_double
            getattr(L, f).argtypes = [C.c_void_p, c_u64p, c_u64p]
        L.oq_i_H_j_double_alpha_beta.restype = C.c_double
        L.oq_i_H_j_double_alpha_beta.argtypes = [C.c_void_p, c_u64p, c_u64p]
        L.oq_i_H_j_double_spin.restype = C.c_double
        L.oq_i_H_j_double_spin.argtypes = [C.c_void_p, c_u64p, c_u64p]
        L.oq_i_H_j_mono_spin.restype = C.c_double
        L.oq_i_H_j_mono_spin.argtypes = [C.c_void_p, c_u64p, c_u64p, C.c_int]
        L.oq_diag_H_mat_elem.restype = C.c_double
        L.oq_diag_H_mat_elem.argtypes = [C.c_void_p, c_u64p]
        L.oq_diag_S_mat_elem.restype = C.c_double
        L.oq_diag_S_mat_elem.argtypes = [c_u64p, C.c_int]
        L.oq_get_s2.restype = C.c_double
        L.oq_get_s2.argtypes = [c_u64p, c_u64p, C.c_int]
        L.oq_build_fock_tmp.argtypes = [C.c_void_p, c_u64p, c_dp]
        L.oq_diag_H_mat_elem_fock.restype = C.c_double
        L.oq_diag_H_mat_elem_fock.argtypes = [C.c_void_p, c_u64p, c_u64p, c_dp]
        L.oq_bilinear_create.restype = C.c_void_p
        L.oq_bilinear_create.argtypes = [c_u64p, C.c_int64, C.c_int]
        L.oq_bilinear_destroy.argtypes = [C.c_void_p]
        L.oq_H_S2_u_0_nstates.argtypes = [C.c_void_p, C.c_void_p, c_dp, c_dp, c_dp, C.c_int, C.c_int64, C.c_int]
        L.oq_H_S2_u_0_work.argtypes = [C.c_void_p, C.c_void_p, c_dp, c_dp, c_dp, C.c_int,
                                       C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_int]
        L.oq_sigma_count_pairs.restype = C.c_int64
        L.oq_sigma_count_pairs.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
        L.oq_jacobi_eigh.argtypes = [C.c_int, c_dp, c_dp, c_dp]
        L.oq_davidson_diag_hs2.restype = C.c_int
        L.oq_davidson_diag_hs2.argtypes = [C.c_void_p, C.c_void_p, c_u64p, c_dp, C.c_int64, C.c_int64,
                                           C.c_int, C.c_int, C.c_void_p, c_dp, c_dp]
        L.oq_sel_buffer_create.restype = C.c_void_p
        L.oq_sel_buffer_create.argtypes = [C.c_int, C.c_int, C.c_int]
        L.oq_sel_buffer_destroy.argtypes = [C.c_void_p]
        L.oq_select_all.argtypes = [C.c_void_p, C.c_void_p, c_dp, c_dp, C.c_void_p, C.c_int]
        L.oq_select_tasks.argtypes = [C.c_void_p, C.c_void_p, c_dp, c_dp, C.c_void_p, C.c_int, C.c_int,
                                      C.POINTER(C.c_int), C.POINTER(C.c_int), c_dp]
        L.oq_select_connected.argtypes = [C.c_void_p, C.c_void_p, C.c_int, c_dp, c_dp, C.c_void_p, C.c_int]
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(c_dp)


def _u64(a):
    return a.ctypes.data_as(c_u64p)


def _dets(d):
    d = np.ascontiguousarray(np.asarray(d, dtype=np.uint64))
    return d


class Ham:
    """oq_ham wrapper; dets are (..., 2, n_int) uint64 arrays."""

    def __init__(self, H):
        self.H = H
        self._h = np.asfortranarray(H.h, dtype=np.float64)
        self._eri = np.ascontiguousarray(H.eri, dtype=np.float64)
        self.ptr = lib().oq_ham_create(H.M, H.n_alpha, H.n_beta, _dp(self._h), _dp(self._eri), self._eri.size)
        if not self.ptr:
            raise ValueError("oq_ham_create failed")
        self.n_int = H.n_int

    def __del__(self):
        if getattr(self, "ptr", None) and _LIB is not None:
            _LIB.oq_ham_destroy(self.ptr)
            self.ptr = None

    def integral(self, i, j, k, l):
        return lib().oq_get_integral(self.ptr, i, j, k, l)

    def ref_energy(self):
        return lib().oq_ref_energy(self.ptr)

    def fock_cs(self):
        out = np.zeros((self.H.M, self.H.M), order="F")
        lib().oq_get_fock_cs(self.ptr, _dp(out))
        return out

    def degree(self, a, b):
        a, b = _dets(a), _dets(b)
        return lib().oq_excitation_degree(_u64(a), _u64(b), self.n_int)

    def excitation(self, a, b):
        a, b = _dets(a), _dets(b)
        exc = np.zeros(12, dtype=np.int32)
        deg = C.c_int(0)
        ph = C.c_double(1.0)
        lib().oq_get_excitation(_u64(a), _u64(b), exc.ctypes.data_as(c_ip), C.byref(deg), C.byref(ph), self.n_int)
        return deg.value, exc.reshape(2, 2, 3), ph.value   # exc[spin, hp, n]

    def i_H_j(self, a, b):
        a, b = _dets(a), _dets(b)
        return lib().oq_i_H_j(self.ptr, _u64(a), _u64(b))

    def diag(self, a):
        a = _dets(a)
        return lib().oq_diag_H_mat_elem(self.ptr, _u64(a))

    def s2(self, a, b):
        a, b = _dets(a), _dets(b)
        return lib().oq_get_s2(_u64(a), _u64(b), self.n_int)

    def diag_s(self, a):
        a = _dets(a)
        return lib().oq_diag_S_mat_elem(_u64(a), self.n_int)

    def fock_tmp(self, det):
        det = _dets(det)
        f = np.zeros(2 * (self.H.M + 1))
        lib().oq_build_fock_tmp(self.ptr, _u64(det), _dp(f))
        return f

    def diag_fock(self, ref, pert, f):
        ref, pert = _dets(ref), _dets(pert)
        return lib().oq_diag_H_mat_elem_fock(self.ptr, _u64(ref), _u64(pert), _dp(f))


class _BilinearStruct(C.Structure):
    # oq_bilinear (qp_oracle.h): 1-based int arrays as in the reference
    _fields_ = [("n_det", C.c_int64), ("nint", C.c_int), ("n_alpha", C.c_int), ("n_beta", C.c_int),
                ("alpha_unique", c_u64p), ("beta_unique", c_u64p),
                ("rows", c_ip), ("columns", c_ip), ("order", c_ip), ("order_reverse", c_ip),
                ("columns_loc", c_ip), ("transp_rows", c_ip), ("transp_columns", c_ip), ("transp_order", c_ip),
                ("transp_rows_loc", c_ip), ("order_transp_reverse", c_ip)]


class Bilinear:
    def __init__(self, dets):
        self.dets = _dets(dets)
        self.n_det = self.dets.shape[0]
        self.n_int = self.dets.shape[-1]
        self.ptr = lib().oq_bilinear_create(_u64(self.dets), self.n_det, self.n_int)

    def order(self) -> np.ndarray:
        """psi_bilinear_matrix_order - 1: the psi_det index (0-based) of
        each bilinear position"""
        b = _BilinearStruct.from_address(self.ptr)
        return np.ctypeslib.as_array(b.order, shape=(self.n_det,)).astype(np.int64) - 1

    def __del__(self):
        if getattr(self, "ptr", None) and _LIB is not None:
            _LIB.oq_bilinear_destroy(self.ptr)
            self.ptr = None


def sigma(ham: Ham, bil: Bilinear, u: np.ndarray, nthreads: int = 0):
    """H_S2_u_0_nstates_openmp: u is (N_det, N_st); returns v, s of the same shape."""
    u = np.asfortranarray(u, dtype=np.float64)
    if u.ndim == 1:
        u = u[:, None].copy(order="F")
    sze, nst = u.shape
    v = np.zeros((sze, nst), order="F")
    s = np.zeros((sze, nst), order="F")
    lib().oq_H_S2_u_0_nstates(ham.ptr, bil.ptr, _dp(v), _dp(s), _dp(u), nst, sze, nthreads)
    return v, s


def sigma_work_sample(ham: Ham, bil: Bilinear, nst: int, istep: int, nthreads: int = 0):
    """Run the work routine on every istep-th row (bilinear order); returns
    (rows_done, v_t) with v_t (N_det, nst) state-fastest.  Used to time a
    bounded sample of a sigma call."""
    n = bil.n_det
    rng = np.random.default_rng(7)
    u_t = np.ascontiguousarray(rng.standard_normal((n, nst)))
    v_t = np.zeros((n, nst))
    s_t = np.zeros((n, nst))
    lib().oq_H_S2_u_0_work(ham.ptr, bil.ptr, _dp(v_t), _dp(s_t), _dp(u_t), nst, 1, n, 0, istep, nthreads)
    return len(range(1, n + 1, istep)), v_t


def sigma_rows(ham: Ham, bil: Bilinear, u: np.ndarray, istep: int, nthreads: int = 0):
    """Rows of v = H u and s = S^2 u for every istep-th bilinear position,
    by the work routine on those rows only; u (N_det, N_st) in psi_det order.
    Returns (psi indices of the rows, v rows, s rows)."""
    u = np.asarray(u, dtype=np.float64)
    if u.ndim == 1:
        u = u[:, None]
    n, nst = u.shape
    order = bil.order()
    u_t = np.ascontiguousarray(u[order])          # (N_det, N_st) state-fastest, bilinear order
    v_t = np.zeros((n, nst))
    s_t = np.zeros((n, nst))
    lib().oq_H_S2_u_0_work(ham.ptr, bil.ptr, _dp(v_t), _dp(s_t), _dp(u_t), nst, 1, n, 0, istep, nthreads)
    k = np.arange(0, n, istep)
    return order[k], v_t[k], s_t[k]


def count_pairs(ham: Ham, bil: Bilinear, nthreads: int = 0) -> int:
    return lib().oq_sigma_count_pairs(ham.ptr, bil.ptr, nthreads)


def jacobi_eigh(A):
    A = np.asfortranarray(A, dtype=np.float64)
    n = A.shape[0]
    w = np.zeros(n)
    V = np.zeros((n, n), order="F")
    lib().oq_jacobi_eigh(n, _dp(A), _dp(w), _dp(V))
    return w, V


class _DavParams(C.Structure):
    _fields_ = [("threshold", C.c_double), ("sze_max", C.c_int), ("s2_eig", C.c_int),
                ("state_following", C.c_int), ("expected_s2", C.c_double), ("s_z2_sz", C.c_double),
                ("nthreads", C.c_int), ("max_iter", C.c_int), ("verbose", C.c_int)]


def davidson(ham: Ham, bil: Bilinear, dets, u_guess, n_st: int, n_st_diag: int, threshold=1e-10,
             sze_max=8, s2_eig=True, state_following=False, nthreads=0, max_iter=0, verbose=False):
    """davidson_diag_hs2; u_guess (N_det, >= n_st) guess columns.  Returns
    (energies[n_st_diag], s2[n_st_diag], u[N_det, n_st_diag], n_sigma)."""
    dets = _dets(dets)
    sze = dets.shape[0]
    u = np.zeros((sze, n_st_diag), order="F")
    g = np.asarray(u_guess, dtype=np.float64)
    if g.ndim == 1:
        g = g[:, None]
    u[:, : g.shape[1]] = g
    na, nb = ham.H.n_alpha, ham.H.n_beta
    sz = 0.5 * (na - nb)
    P = _DavParams(threshold, sze_max, int(s2_eig), int(state_following), sz * (sz + 1.0), sz * (sz - 1.0),
                   nthreads, max_iter, int(verbose))
    e = np.zeros(n_st_diag)
    s2 = np.zeros(n_st_diag)
    rc = lib().oq_davidson_diag_hs2(ham.ptr, bil.ptr, _u64(dets), _dp(u), sze, sze, n_st, n_st_diag,
                                    C.byref(P), _dp(e), _dp(s2))
    if rc < 0:
        raise RuntimeError(f"oracle Davidson failed ({rc})")
    return e, s2, u, rc


class _SelBuf(C.Structure):
    _fields_ = [("N", C.c_int), ("cur", C.c_int), ("cap", C.c_int), ("nint", C.c_int),
                ("det", c_u64p), ("val", c_dp), ("mini", C.c_double)]


class _SelCtx(C.Structure):
    _fields_ = [("n_det", C.c_int64), ("n_states", C.c_int), ("nint", C.c_int),
                ("psi_det_sorted", c_u64p), ("coef_transp", c_dp), ("n_gen", C.c_int), ("n_sel", C.c_int),
                ("B", C.c_void_p), ("fragment_count", C.c_int),
                ("gen_hole_mask", C.c_uint64 * 8), ("gen_part_mask", C.c_uint64 * 8),
                ("n_candidates", C.c_int64), ("n_gathers", C.c_int64), ("n_blocks", C.c_int64),
                ("n_blocks_total", C.c_int64),
                ("block_log", c_dp), ("block_log_cap", C.c_int64), ("block_log_n", C.c_int64),
                ("block_keep", C.POINTER(C.c_uint8)), ("block_keep_offsets", C.POINTER(C.c_int64))]


def selection(ham: Ham, dets_sorted, coefs_sorted, n_gen: int, n_sel: int, E0, n_keep: int,
              nthreads: int = 0, block_log_cap: int = 0, block_keep=None, block_keep_offsets=None, tasks=None):
    """ZMQ_selection on a psi already sorted by decreasing weight.

    dets_sorted (N_det, 2, n_int); coefs_sorted (N_det, N_states).
    tasks: optional (n, 2) int array of (i_generator, subset) pairs, the
    reference's selection tasks; default every generator whole.
    Returns dict(pt2, dets, vals, n_candidates, n_gathers, n_blocks, block_log)."""
    dets = _dets(dets_sorted)
    n_det, _, nint = dets.shape
    coefs = np.asarray(coefs_sorted, dtype=np.float64)
    if coefs.ndim == 1:
        coefs = coefs[:, None]
    ns = coefs.shape[1]
    ct = np.ascontiguousarray(coefs)  # (N_det, N_states) C-order == (N_states, N_det) Fortran
    bil = Bilinear(dets)
    ctx = _SelCtx()
    ctx.n_det = n_det
    ctx.n_states = ns
    ctx.nint = nint
    ctx.psi_det_sorted = _u64(dets)
    ctx.coef_transp = _dp(ct)
    ctx.n_gen = n_gen
    ctx.n_sel = n_sel
    ctx.B = bil.ptr
    ctx.fragment_count = ham.H.n_alpha ** 2
    M = ham.H.M
    full = [0] * 8
    for w in range(nint):
        nb = min(64, M - 64 * w)
        full[w] = (1 << nb) - 1 if nb < 64 else 0xFFFFFFFFFFFFFFFF
        full[nint + w] = full[w]
    for i in range(8):
        ctx.gen_hole_mask[i] = full[i]
        ctx.gen_part_mask[i] = full[i]
    log = None
    if block_log_cap:
        log = np.zeros((block_log_cap, 8))
        ctx.block_log = _dp(log)
        ctx.block_log_cap = block_log_cap
    if block_keep is not None:
        bk = np.ascontiguousarray(block_keep, dtype=np.uint8)
        bo = np.ascontiguousarray(block_keep_offsets, dtype=np.int64)
        ctx.block_keep = bk.ctypes.data_as(C.POINTER(C.c_uint8))
        ctx.block_keep_offsets = bo.ctypes.data_as(C.POINTER(C.c_int64))
    E0 = np.ascontiguousarray(np.atleast_1d(np.asarray(E0, dtype=np.float64)))
    pt2 = np.zeros(ns)
    task_pt2 = None
    buf = lib().oq_sel_buffer_create(n_keep, 2 * n_keep, nint)
    if tasks is None:
        lib().oq_select_all(ham.ptr, C.byref(ctx), _dp(E0), _dp(pt2), buf, nthreads)
    else:
        tk = np.asarray(tasks, dtype=np.int32).reshape(-1, 2)
        tg = np.ascontiguousarray(tk[:, 0])
        tsb = np.ascontiguousarray(tk[:, 1])
        task_pt2 = np.zeros((tk.shape[0], ns))
        ip = C.POINTER(C.c_int)
        lib().oq_select_tasks(ham.ptr, C.byref(ctx), _dp(E0), _dp(pt2), buf, nthreads, tk.shape[0],
                              tg.ctypes.data_as(ip), tsb.ctypes.data_as(ip), _dp(task_pt2))
    b = _SelBuf.from_address(buf)
    cur = b.cur
    out_d = np.ctypeslib.as_array(b.det, shape=(b.cap * 2 * nint,))[: cur * 2 * nint].reshape(cur, 2, nint).copy()
    out_v = np.ctypeslib.as_array(b.val, shape=(b.cap,))[:cur].copy()
    lib().oq_sel_buffer_destroy(buf)
    res = dict(pt2=pt2, dets=out_d, vals=out_v, n_candidates=ctx.n_candidates, n_gathers=ctx.n_gathers,
               n_blocks=ctx.n_blocks, n_blocks_total=ctx.n_blocks_total, task_pt2=task_pt2)
    if log is not None:
        res["block_log"] = log[: ctx.block_log_n].copy()
    return res
