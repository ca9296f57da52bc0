"""Host-side mirror of the reference's provider/subroutine interface for the
determinant-driven CI hot path, backed by libqpx on the GPU.

Names follow the reference:
  H_S2_u_0_nstates   -> H_S2_u_0_nstates_openmp   (src/Davidson/u0Hu0.irp.f:15)
  u_0_H_u_0          -> u_0_H_u_0                 (src/Davidson/diagonalization_hs2.irp.f:446)
  davidson_diag_hs2  -> davidson_diag_hs2         (src/Davidson/diagonalization_hs2.irp.f:1)
  ZMQ_selection      -> ZMQ_selection             (plugins/Full_CI_ZMQ/zmq_selection.irp.f:1)
Arrays follow the reference layouts: psi_det (N_det, 2, N_int) uint64 (i.e.
Fortran det(N_int,2,N_det)), vectors (sze, N_st) column-major.  Errors raise
QpxError, the counterpart of the reference's ``stop`` on fatal errors.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

FP64, FP32 = 0, 1


class QpxError(RuntimeError):
    pass


def _dp(a):
    return a.ctypes.data_as(_lib.c_dp)


class Context:
    def __init__(self, device: int = 0):
        self.L = _lib.load()
        self.ptr = C.c_void_p()
        rc = self.L.qpx_create(device, C.byref(self.ptr))
        if rc != 0:
            raise QpxError(f"qpx_create failed (code {rc}): no usable HIP device {device}")
        self.H = None
        self.n_det = 0
        self.n_int = 0

    def _check(self, rc):
        if rc != 0:
            buf = C.create_string_buffer(1024)
            self.L.qpx_last_error(self.ptr, buf, 1024)
            raise QpxError(buf.value.decode(errors="replace") or f"libqpx error {rc}")

    def close(self):
        if self.ptr:
            self.L.qpx_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------------- inputs
    def set_verbosity(self, level: int):
        """1: print the Davidson iteration table to stderr."""
        self._check(self.L.qpx_set_verbosity(self.ptr, int(level)))

    def set_hamiltonian(self, H):
        h = np.asfortranarray(H.h, dtype=np.float64)
        eri = np.ascontiguousarray(H.eri, dtype=np.float64)
        self._check(self.L.qpx_set_hamiltonian(self.ptr, H.M, H.n_alpha, H.n_beta, _dp(h), _dp(eri), eri.size))
        self.H = H

    def set_wavefunction(self, psi_det):
        d = np.ascontiguousarray(psi_det, dtype=np.uint64)
        if d.ndim != 3 or d.shape[1] != 2:
            raise QpxError("psi_det must have shape (N_det, 2, N_int)")
        self._check(self.L.qpx_set_wavefunction(self.ptr, d.shape[2], d.shape[0], d.ctypes.data_as(_lib.c_u64p)))
        self.n_det = d.shape[0]
        self.n_int = d.shape[2]

    def bilinear_order(self) -> np.ndarray:
        perm = np.zeros(self.n_det, dtype=np.int64)
        self._check(self.L.qpx_get_bilinear_order(self.ptr, perm.ctypes.data_as(_lib.c_i64p)))
        return perm

    def diag_H(self) -> np.ndarray:
        out = np.zeros(self.n_det)
        self._check(self.L.qpx_diag_h(self.ptr, _dp(out)))
        return out

    # ----------------------------------------------------------------- sigma
    def H_S2_u_0_nstates(self, u_0, prec: int = FP64):
        """v_0 = H u_0, s_0 = S^2 u_0; u_0 (sze, N_st) or (sze,)."""
        u = np.asfortranarray(u_0, dtype=np.float64)
        squeeze = u.ndim == 1
        if squeeze:
            u = u[:, None].copy(order="F")
        sze, nst = u.shape
        v = np.zeros((sze, nst), order="F")
        s = np.zeros((sze, nst), order="F")
        self._check(self.L.qpx_sigma(self.ptr, prec, nst, sze, _dp(u), _dp(v), _dp(s)))
        if squeeze:
            return v[:, 0], s[:, 0]
        return v, s

    def u_0_H_u_0(self, u_0, prec: int = FP64):
        """e_0 = <u_0|H|u_0> per state (no division by the norm when it is
        nonzero, as the reference)."""
        u = np.asfortranarray(u_0, dtype=np.float64)
        if u.ndim == 1:
            u = u[:, None]
        v, _ = self.H_S2_u_0_nstates(u, prec)
        e = np.zeros(u.shape[1])
        for i in range(u.shape[1]):
            if np.dot(u[:, i], u[:, i]) != 0.0:
                e[i] = np.dot(v[:, i], u[:, i])
        return e

    # -------------------------------------------------------------- Davidson
    def davidson_diag_hs2(self, u_in, N_st: int, N_st_diag: int, threshold: float = 1e-10, sze_max: int = 8,
                          s2_eig: bool = True, state_following: bool = False, mixed: bool = False,
                          switch_threshold: float = 1e-7, seed: int = 0x5EED):
        """Returns (energies[N_st_diag], s2[N_st_diag], u[sze, N_st_diag], n_sigma(fp64, fp32))."""
        g = np.asarray(u_in, dtype=np.float64)
        if g.ndim == 1:
            g = g[:, None]
        sze = g.shape[0]
        u = np.zeros((sze, N_st_diag), order="F")
        u[:, : min(g.shape[1], N_st_diag)] = g[:, :N_st_diag]
        e = np.zeros(N_st_diag)
        s2 = np.zeros(N_st_diag)
        ns = (C.c_int * 2)()
        rc = self.L.qpx_davidson(self.ptr, N_st, N_st_diag, threshold, sze_max, int(s2_eig), int(state_following),
                                 int(mixed), switch_threshold, seed, _dp(u), sze, _dp(e), _dp(s2),
                                 C.cast(ns, _lib.c_ip))
        self._check(rc)
        return e, s2, u, (ns[0], ns[1])

    # ------------------------------------------------------------- selection
    def ZMQ_selection(self, coef_sorted, n_gen: int, n_sel: int, E0, N: int, subset: int = 0):
        """Selection + EN-PT2 over the current wavefunction (set in
        psi_det_sorted order).  Returns (pt2[N_states], dets[n, 2, N_int],
        vals[n])."""
        coef = np.asfortranarray(coef_sorted, dtype=np.float64)
        if coef.ndim == 1:
            coef = coef[:, None].copy(order="F")
        ns = coef.shape[1]
        E0 = np.ascontiguousarray(np.atleast_1d(E0), dtype=np.float64)
        pt2 = np.zeros(ns)
        out_d = np.zeros((max(N, 1), 2, self.n_int), dtype=np.uint64)
        out_v = np.zeros(max(N, 1))
        n_out = C.c_int64(0)
        self._check(self.L.qpx_select(self.ptr, ns, _dp(coef), n_gen, n_sel, _dp(E0), N, subset, _dp(pt2),
                                      out_d.ctypes.data_as(_lib.c_u64p), _dp(out_v), C.byref(n_out)))
        n = n_out.value
        return pt2, out_d[:n].copy(), out_v[:n].copy()

    def select_generators(self, coef_sorted, n_sel: int, E0, gens) -> np.ndarray:
        """EN-PT2 contribution of each listed generator (1-based indices into
        the current psi_det_sorted): (len(gens), N_states)."""
        coef = np.asfortranarray(coef_sorted, dtype=np.float64)
        if coef.ndim == 1:
            coef = coef[:, None].copy(order="F")
        ns = coef.shape[1]
        g = np.ascontiguousarray(gens, dtype=np.int32)
        E0 = np.ascontiguousarray(np.atleast_1d(E0), dtype=np.float64)
        out = np.zeros((g.size, ns))
        self._check(self.L.qpx_select_generators(self.ptr, ns, _dp(coef), n_sel, _dp(E0), g.size,
                                                 g.ctypes.data_as(_lib.c_ip), _dp(out)))
        return out

    def s2_complete(self, dets, n_alpha: int):
        """make_s2_eigenfunction on the GPU (occ_pattern.irp.f:251-309): the
        determinants sharing an occupation pattern with dets (N, 2, n_int)
        and not in it, in the order of the host routine
        (cipsi.make_s2_eigenfunction)."""
        d = np.ascontiguousarray(dets, dtype=np.uint64)
        n_int = d.shape[2]
        n = C.c_int64(0)
        self._check(self.L.qpx_s2_complete(self.ptr, n_int, d.shape[0], d.ctypes.data_as(_lib.c_u64p), int(n_alpha),
                                           0, None, C.byref(n)))
        out = np.empty((n.value, 2, n_int), dtype=np.uint64)
        if n.value:
            self._check(self.L.qpx_s2_complete(self.ptr, n_int, d.shape[0], d.ctypes.data_as(_lib.c_u64p),
                                               int(n_alpha), n.value, out.ctypes.data_as(_lib.c_u64p),
                                               C.byref(n)))
        return out

    def select_tasks(self, coef_sorted, n_sel: int, E0, N: int, tasks):
        """The selection task farm over explicit (i_generator, subset) tasks
        (1-based generator indices into the current psi_det_sorted; subset 0
        = every block).  Returns (pt2[N_states], dets[n, 2, N_int], vals[n],
        task_pt2[n_tasks, N_states])."""
        coef = np.asfortranarray(coef_sorted, dtype=np.float64)
        if coef.ndim == 1:
            coef = coef[:, None].copy(order="F")
        ns = coef.shape[1]
        tk = np.asarray(tasks, dtype=np.int32).reshape(-1, 2)
        tg = np.ascontiguousarray(tk[:, 0])
        tsb = np.ascontiguousarray(tk[:, 1])
        E0 = np.ascontiguousarray(np.atleast_1d(E0), dtype=np.float64)
        pt2 = np.zeros(ns)
        out_d = np.zeros((max(N, 1), 2, self.n_int), dtype=np.uint64)
        out_v = np.zeros(max(N, 1))
        n_out = C.c_int64(0)
        tpt2 = np.zeros((tk.shape[0], ns))
        self._check(self.L.qpx_select_tasks(self.ptr, ns, _dp(coef), n_sel, _dp(E0), N, tk.shape[0],
                                            tg.ctypes.data_as(_lib.c_ip), tsb.ctypes.data_as(_lib.c_ip), _dp(pt2),
                                            out_d.ctypes.data_as(_lib.c_u64p), _dp(out_v), C.byref(n_out),
                                            _dp(tpt2)))
        n = n_out.value
        return pt2, out_d[:n].copy(), out_v[:n].copy(), tpt2

    def select_blocks(self, coef_sorted, n_gen: int, n_sel: int, E0, N: int, keep=None, info: bool = True):
        """Selection restricted to the blocks with keep[b] != 0 (one byte per
        block of the first n_gen generators, in the reference's block order;
        None = all).  Returns (pt2, dets, vals, block_pt2[n_blocks, N_states],
        block_info[n_blocks, 6] = gen, s1, h1, s2, h2 (1-based), status);
        block_info is None with info=False (it is 48 bytes per block)."""
        coef = np.asfortranarray(coef_sorted, dtype=np.float64)
        if coef.ndim == 1:
            coef = coef[:, None].copy(order="F")
        ns = coef.shape[1]
        nb = C.c_int64(0)
        self._check(self.L.qpx_select_count_blocks(self.ptr, n_gen, C.byref(nb)))
        nb = nb.value
        E0 = np.ascontiguousarray(np.atleast_1d(E0), dtype=np.float64)
        pt2 = np.zeros(ns)
        out_d = np.zeros((max(N, 1), 2, self.n_int), dtype=np.uint64)
        out_v = np.zeros(max(N, 1))
        n_out = C.c_int64(0)
        bpt2 = np.zeros((nb, ns))
        binfo = np.zeros((nb, 6), dtype=np.int64) if info else None
        kp = None
        if keep is not None:
            k = np.ascontiguousarray(keep, dtype=np.uint8)
            if k.size != nb:
                raise QpxError(f"keep mask has {k.size} entries, expected {nb}")
            kp = k.ctypes.data_as(C.POINTER(C.c_uint8))
        self._check(self.L.qpx_select_blocks(self.ptr, ns, _dp(coef), n_gen, n_sel, _dp(E0), N, kp, _dp(pt2),
                                             out_d.ctypes.data_as(_lib.c_u64p), _dp(out_v), C.byref(n_out),
                                             _dp(bpt2),
                                             binfo.ctypes.data_as(_lib.c_i64p) if info else None))
        n = n_out.value
        return pt2, out_d[:n].copy(), out_v[:n].copy(), bpt2, binfo

    def select_stats(self):
        st = np.zeros(4, dtype=np.int64)
        self._check(self.L.qpx_select_stats(self.ptr, st.ctypes.data_as(_lib.c_i64p)))
        return dict(candidates=int(st[0]), blocks=int(st[1]), blocks_visited=int(st[2]), gathers=int(st[3]))

    # ----------------------------------------------------------------- tests
    def _force_sigma_path(self, path: int) -> int:
        last = C.c_int(0)
        self._check(self.L.qpx_debug_force_sigma_path(self.ptr, path, C.byref(last)))
        return last.value

    def _select_block_ticks(self, n_blocks: int):
        """Per-block workgroup durations (100 MHz ticks) of the last selection
        call, in block enumeration order."""
        t = np.zeros(n_blocks, dtype=np.int64)
        self._check(self.L.qpx_debug_select_block_ticks(self.ptr, t.ctypes.data_as(_lib.c_i64p), n_blocks))
        return t

    def _sparse_stats(self):
        """Entries the sparse sigma walk visits per class (sum of squared
        bucket sizes): ab, a1, a2, b1, b2."""
        v = np.zeros(5, dtype=np.int64)
        self._check(self.L.qpx_debug_sparse_stats(self.ptr, v.ctypes.data_as(_lib.c_i64p)))
        return dict(zip(("ab", "a1", "a2", "b1", "b2"), (int(x) for x in v)))


# ---------------------------------------------------------------------------
# block-proposal network (BF16 MFMA kernels in libqpx)
# ---------------------------------------------------------------------------
_c_fp = C.POINTER(C.c_float)
PARAM_NAMES = ["W1", "W2", "W3", "b1", "b2", "b3", "w4", "emb", "b4"]


class BlockMLP:
    """MLP over (generator, hole-pair) block descriptors predicting the
    block's log|EN-PT2 contribution|.  desc (n, 4) int32: h1, h2 (0-based),
    spin class (0 aa, 1 bb, 2 ab), generator excitation degree from HF;
    fdesc (n, 3) float32: h_11, h_22, log|c_gen|."""

    def __init__(self, n_orb: int, seed: int = 1, device: int = 0):
        self.L = _lib.load()
        self.ptr = C.c_void_p()
        rc = self.L.qpx_mlp_create(device, n_orb, seed, C.byref(self.ptr))
        if rc != 0:
            raise QpxError(f"qpx_mlp_create failed ({rc})")
        self.n_orb = n_orb
        self.shapes = {"W1": (80, 256), "W2": (256, 256), "W3": (256, 256), "b1": (256,), "b2": (256,),
                       "b3": (256,), "w4": (256,), "emb": (n_orb, 32), "b4": (1,)}

    def _check(self, rc):
        if rc != 0:
            buf = C.create_string_buffer(1024)
            self.L.qpx_mlp_last_error(self.ptr, buf, 1024)
            raise QpxError(buf.value.decode(errors="replace"))

    def close(self):
        if self.ptr:
            self.L.qpx_mlp_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _prep(desc, fdesc):
        d = np.ascontiguousarray(desc, dtype=np.int32)
        f = np.ascontiguousarray(fdesc, dtype=np.float32)
        return d, f

    def train_step(self, desc, fdesc, target, lr):
        d, f = self._prep(desc, fdesc)
        t = np.ascontiguousarray(target, dtype=np.float32)
        loss = C.c_float(0)
        self._check(self.L.qpx_mlp_train_step(self.ptr, d.shape[0], d.ctypes.data_as(_lib.c_ip),
                                              f.ctypes.data_as(_c_fp), t.ctypes.data_as(_c_fp), lr,
                                              C.byref(loss)))
        return loss.value

    def train_epochs(self, desc, fdesc, target, lr, epochs: int, batch: int, seed: int = 0):
        """epochs of Adam steps on batches of a device-resident training set;
        returns the mean batch loss of each epoch."""
        d, f = self._prep(desc, fdesc)
        t = np.ascontiguousarray(target, dtype=np.float32)
        losses = np.zeros(epochs, dtype=np.float32)
        self._check(self.L.qpx_mlp_train_epochs(self.ptr, d.shape[0], d.ctypes.data_as(_lib.c_ip),
                                                f.ctypes.data_as(_c_fp), t.ctypes.data_as(_c_fp), lr, epochs,
                                                batch, seed, losses.ctypes.data_as(_c_fp)))
        return losses

    def predict(self, desc, fdesc):
        d, f = self._prep(desc, fdesc)
        out = np.zeros(d.shape[0], dtype=np.float32)
        self._check(self.L.qpx_mlp_predict(self.ptr, d.shape[0], d.ctypes.data_as(_lib.c_ip),
                                           f.ctypes.data_as(_c_fp), out.ctypes.data_as(_c_fp)))
        return out

    def get(self, name, grad=False):
        shp = self.shapes[name]
        a = np.zeros(int(np.prod(shp)), dtype=np.float32)
        self._check(self.L.qpx_mlp_param(self.ptr, PARAM_NAMES.index(name), int(grad), 0,
                                         a.ctypes.data_as(_c_fp), a.size))
        return a.reshape(shp)

    def set(self, name, value):
        a = np.ascontiguousarray(value, dtype=np.float32).ravel()
        self._check(self.L.qpx_mlp_param(self.ptr, PARAM_NAMES.index(name), 0, 1, a.ctypes.data_as(_c_fp), a.size))

    def _gemm_layout(self, A, B, ta, tb, split):
        """C = A B through the network's GEMM with A passed transposed in
        memory (ta) and/or B (tb), and the K split enabled (split)."""
        M, K = A.shape
        N = B.shape[1]
        Ad = np.ascontiguousarray(A.T if ta else A, dtype=np.float32)
        Bd = np.ascontiguousarray(B.T if tb else B, dtype=np.float32)
        Cm = np.zeros((M, N), dtype=np.float32)
        self._check(self.L.qpx_debug_mlp_gemm(self.ptr, M, N, K, Ad.ctypes.data_as(_c_fp), Bd.ctypes.data_as(_c_fp),
                                              int(ta), int(tb), int(split), Cm.ctypes.data_as(_c_fp)))
        return Cm

    def gemm_check(self, A, B):
        A = np.ascontiguousarray(A, dtype=np.float32)
        B = np.ascontiguousarray(B, dtype=np.float32)
        Cm = np.zeros((A.shape[0], B.shape[1]), dtype=np.float32)
        self._check(self.L.qpx_mlp_gemm_check(self.ptr, A.shape[0], B.shape[1], A.shape[1], A.ctypes.data_as(_c_fp),
                                              B.ctypes.data_as(_c_fp), Cm.ctypes.data_as(_c_fp)))
        return Cm
