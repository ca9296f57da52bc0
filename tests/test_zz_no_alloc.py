"""The allocation-free prediction entry point (qpx_mlp_predict_static)
against qpx_mlp_predict, called through the C-ABI: the same kernels on the
same inputs, so the predictions agree bit for bit.  Two calls in a row, the
second on a smaller input, and the device's free memory is unchanged by
them; sizes beyond the scratch are rejected."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from test_gpu_sigma import _say

pytestmark = pytest.mark.gpu

_fp = C.POINTER(C.c_float)


def _static_predict(m, desc, fdesc):
    d = np.ascontiguousarray(desc, dtype=np.int32)
    f = np.ascontiguousarray(fdesc, dtype=np.float32)
    out = np.zeros(d.shape[0], dtype=np.float32)
    rc = m.L.qpx_mlp_predict_static(m.ptr, d.shape[0], d.ctypes.data_as(C.POINTER(C.c_int)), f.ctypes.data_as(_fp),
                                    out.ctypes.data_as(_fp))
    return rc, out


def _free_bytes():
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(0), C.c_size_t(0)
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def _inputs(rng, n, n_orb):
    desc = np.stack([rng.integers(0, n_orb, n), rng.integers(0, n_orb, n), rng.integers(0, 3, n),
                     rng.integers(0, 5, n)], axis=1).astype(np.int32)
    return desc, rng.standard_normal((n, 3)).astype(np.float32)


def test_predict_static_matches_predict():
    from quantum_package_amd.api import BlockMLP
    rng = np.random.default_rng(12)
    n_orb = 48
    m = BlockMLP(n_orb=n_orb, seed=5)
    # a few training steps so that every parameter differs from its start
    d, f = _inputs(rng, 4096, n_orb)
    m.train_epochs(d, f, rng.standard_normal(4096).astype(np.float32), 1e-3, 2, 512, 3)
    for n in (2048, 311):
        desc, fdesc = _inputs(rng, n, n_orb)
        _say(f"[no-alloc] predict_static n={n}")
        want = m.predict(desc, fdesc)
        before = _free_bytes()
        rc, got = _static_predict(m, desc, fdesc)
        after = _free_bytes()
        assert rc == 0
        assert after == before
        assert np.array_equal(got, want)
    for bad in (0, 2049):
        desc, fdesc = _inputs(rng, max(bad, 1), n_orb)
        rc, _ = _static_predict(m, desc[:bad] if bad else desc[:0], fdesc[:bad] if bad else fdesc[:0])
        assert rc != 0
    m.close()
