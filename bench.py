"""Benchmark of the determinant-driven CI hot path on one MI355X.

A step is one sigma call v = H u, s = S^2 u over the full-CI space of the
synthetic 14-orbital, 6 alpha + 6 beta electron Hamiltonian (9,018,009
determinants, N_int = 1) with 4 vectors (3 roots + 1 extra, the Davidson
block size) in FP64, inputs resident in HBM.  value = determinants / s.

The JSON line also carries:
  roofline      the dominant (column) kernel against its bound, the LDS
                read bandwidth measured on this GPU: the alpha-beta pairs'
                u-row and integral bytes it reads from LDS per launch over
                its event-timed duration; with its measured HBM traffic
                (rocprofv3 counters, calibrated) and the whole call's;
  cpu_baseline  the CPU oracle (a C/OpenMP restatement of the reference's
                H_S2_u_0_nstates_openmp_work, the reference itself cannot be
                built here) timed on a bounded, strided sample of the same
                rows on this host's cores;
  davidson      davidson_diag_hs2 on the same space (lowest 3 roots, 4
                vectors, threshold 1e-10): plain FP64 against mixed precision (FP32
                correction-vector sigma, FP64 everything else), wall time,
                sigma calls and |dE| per root;
  pt2           CIPSI selection + EN-PT2 candidates/s (FP64) on the
                128-orbital, 12 alpha + 12 beta electron system (N_int = 2):
                a selected wavefunction of --pt2-ndet determinants (one
                selection step from HF, then diagonalised), the leading
                generators by weight timed on the GPU and, on a smaller
                generator sample, on the CPU oracle.
Legs after the first are skipped with --no-davidson / --no-pt2 / --no-cipsi; a leg that
fails records its error in its own field and the line is still printed.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "H·Ψ sigma dets/s + PT2 candidates/s at |ΔE|<1 µEh vs FP64 plain path"
HBM_PEAK = 8.0e12


def sigma_alg_bytes(n_det, n_pairs, n_st, w, n_int):
    """B = P (8 N_st w + 8 w) + N_det (24 N_st w + 16 N_int)"""
    return n_pairs * (8 * n_st * w + 8 * w) + n_det * (24 * n_st * w + 16 * n_int)


def fci_pairs(M, na):
    ns = na * (M - na)
    nd = math.comb(na, 2) * math.comb(M - na, 2)
    return ns * ns + 2 * (ns + nd)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--prec", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--quiet", action="store_true")
    ap.add_argument("--no-davidson", action="store_true")
    ap.add_argument("--no-pt2", action="store_true")
    ap.add_argument("--pt2-ndet", type=int, default=100000)
    ap.add_argument("--pt2-seconds", type=float, default=20.0)
    ap.add_argument("--pt2-per-stratum", type=int, default=16)
    ap.add_argument("--no-pt2-stoch", action="store_true")
    ap.add_argument("--no-selected", action="store_true")
    ap.add_argument("--sel-small", type=int, default=8000,
                    help="CIPSI loop size before the large selection step of the selected-psi leg")
    ap.add_argument("--sel-select", type=int, default=930000,
                    help="determinants selected in one step before the S^2 completion (about 2^21 after it)")
    ap.add_argument("--sel-no-davidson", action="store_true",
                    help="selected-psi leg: sigma timings only (for counter runs)")
    ap.add_argument("--no-mlp", action="store_true")
    ap.add_argument("--no-cipsi", action="store_true")
    ap.add_argument("--cipsi-ndet", type=int, default=65536)
    ap.add_argument("--calib-copy", action="store_true",
                    help="only run a 1 GiB device-to-device copy (counter calibration for rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only (the hot path is measured on one "
              "MI355X); multi-GPU sharding is a design note, not a launched path.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    if args.calib_copy:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda:0")   # 1 GiB
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calib_copy_bytes": src.numel() * 8}), flush=True)
        return
    from quantum_package_amd.api import Context, FP32, FP64
    from quantum_package_amd.hamiltonian import fci_space, synthetic_hamiltonian

    def log(msg):
        if not args.quiet:
            print(msg, file=sys.stderr, flush=True)

    M, na, nb, seed, nst = 14, 6, 6, 1002, 4
    prec = FP64 if args.prec == "fp64" else FP32
    w = 1.0 if prec == FP64 else 0.5
    H = synthetic_hamiltonian(M, na, nb, seed=seed)
    t0 = time.time()
    dets = fci_space(M, na, nb)
    n_det = dets.shape[0]
    log(f"bench: FCI space {n_det} dets built in {time.time() - t0:.1f}s")
    ctx = Context(0)
    ctx.set_hamiltonian(H)
    t0 = time.time()
    ctx.set_wavefunction(dets)
    log(f"bench: wavefunction set up in {time.time() - t0:.1f}s")

    dev = torch.device("cuda:0")
    dt = torch.float64 if prec == FP64 else torch.float32
    g = torch.Generator(device=dev)
    g.manual_seed(7)
    u = torch.randn(n_det * nst, dtype=dt, device=dev, generator=g)
    v = torch.empty_like(u)
    s = torch.empty_like(u)
    # a dedicated stream (default priority): the kernels and the timing
    # events must share it
    stream = torch.cuda.Stream(device=dev)
    sptr = stream.cuda_stream
    torch.cuda.synchronize()

    def step():
        ctx._check(ctx.L.qpx_sigma_device(ctx.ptr, prec, nst, u.data_ptr(), v.data_ptr(), s.data_ptr(), sptr))

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t_wall0 = time.time()
    ev0.record(stream)
    for _ in range(args.steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    t_wall = time.time() - t_wall0
    ms = ev0.elapsed_time(ev1) / args.steps
    log(f"bench: sigma {ms:.3f} ms/step (wall {1e3 * t_wall / args.steps:.3f} ms/step), path {ctx._force_sigma_path(0)}")
    dets_per_s = n_det / (ms * 1e-3)

    # per-kernel split of one sigma call, measured with events on the same stream
    kms = np.zeros(2, dtype=np.float32)
    ctx._check(ctx.L.qpx_debug_sigma_kernel_times(ctx.ptr, prec, nst, u.data_ptr(), v.data_ptr(), s.data_ptr(),
                                                  sptr, kms.ctypes.data_as(__import__("ctypes").POINTER(
                                                      __import__("ctypes").c_float)), 3))
    P = fci_pairs(M, na) * n_det
    B = sigma_alg_bytes(n_det, P, nst, w, 1)
    traffic = traffic_col = None
    tpath = os.path.join(ROOT, "profiles", "sigma_dense_fci14_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        traffic = tj.get("hbm_bytes_per_sigma")
        for name, kv in tj.get("kernels", {}).items():
            if "sigma_dense_col_kernel" in name and ("double" if prec == FP64 else "float") in name:
                traffic_col = kv.get("hbm_bytes")
    # Roofline of the dominant kernel, the column kernel (about 80 % of the
    # call).  It serves the alpha-beta pairs' reads -- per pair a row of u
    # (8 N_st w bytes) and an integral (8 w) -- from LDS, where the survey's
    # gather model counts them as HBM gathers; its bound is the LDS read
    # bandwidth, measured on this GPU by a conflict-free LDS read kernel.
    lds_peak = __import__("ctypes").c_double(0.0)
    ctx._check(ctx.L.qpx_debug_lds_bandwidth(ctx.ptr, __import__("ctypes").byref(lds_peak)))
    lds_peak = lds_peak.value
    n_s = na * (M - na)
    P_ab = n_det * n_s * n_s
    lds_bytes = P_ab * (8 * nst + 8) * w
    t_col = float(kms[0]) * 1e-3
    achieved = lds_bytes / t_col

    out = {
        "metric": METRIC,
        "value": dets_per_s,
        "unit": "dets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == FP64 else "f32",
        "data": "synthetic",
        "config": {"workload": "full-CI sigma H.Psi and S^2.Psi, synthetic Hamiltonian seed 1002: 14 MOs, "
                               "6 alpha + 6 beta electrons, 9,018,009 determinants, N_int=1, 4 vectors "
                               "(3 roots + 1, Davidson block)",
                   "n_det": n_det, "n_states_per_call": nst, "connected_pairs": P,
                   "kernel_ms": {"sigma_dense_col": float(kms[0]), "sigma_dense_row": float(kms[1])}},
        "roofline": {"bound": "lds", "kernel": f"sigma_dense_col_kernel ({'FP64' if prec == FP64 else 'FP32'}, "
                                               f"{nst} vectors, {n_s} alpha singles per string)",
                     "achieved": achieved / 1e9, "peak": lds_peak / 1e9, "unit": "GB/s", "frac": achieved / lds_peak,
                     "traffic": traffic_col,
                     "hbm_frac": (traffic_col / t_col / HBM_PEAK) if traffic_col else None,
                     "alg_bytes_per_launch": lds_bytes,
                     "note": ("achieved = alpha-beta pairs x (u row + integral) bytes read from LDS per launch / "
                              "the column kernel's event-timed duration; peak = LDS read bandwidth measured on this "
                              "GPU; traffic = the launch's HBM bytes (rocprofv3 FETCH_SIZE + WRITE_SIZE, "
                              "calibrated), hbm_frac = traffic / duration / 8 TB/s"),
                     "call": {"survey_gather_model_bytes": B, "traffic": traffic,
                              "hbm_frac": (traffic / (ms * 1e-3) / HBM_PEAK) if traffic else None}},
    }

    if not args.no_cpu:
        from oracle import oracle as O
        ham = O.Ham(H)
        bil = O.Bilinear(dets)
        nthr = args.cpu_threads
        # calibrate on a small strided sample, then run ~cpu_seconds of work
        istep = max(1, n_det // 2000)
        t0 = time.time()
        rows, _ = O.sigma_work_sample(ham, bil, nst, istep, nthr)
        dt0 = time.time() - t0
        rate0 = rows / max(dt0, 1e-6)
        target_rows = int(rate0 * args.cpu_seconds)
        istep = max(1, n_det // max(target_rows, 1))
        t0 = time.time()
        rows, _ = O.sigma_work_sample(ham, bil, nst, istep, nthr)
        dt1 = time.time() - t0
        out["cpu_baseline"] = {"value": rows / dt1, "unit": "dets/s", "cores": nthr, "kind": "port",
                               "sample": f"{rows} of {n_det} rows (every {istep}-th, bilinear order) of one "
                                         f"FP64 4-vector sigma call, reference algorithm restated in C/OpenMP "
                                         f"schedule(dynamic,64); {dt1:.1f} s"}
        log(f"bench: cpu oracle {rows / dt1:.1f} dets/s on {nthr} threads")
        # the same restatement on one thread, ~1/3 of the time budget
        istep1 = max(1, n_det // max(int(rows / dt1 / nthr * args.cpu_seconds / 3), 1))
        t0 = time.time()
        rows1, _ = O.sigma_work_sample(ham, bil, nst, istep1, 1)
        dt2 = time.time() - t0
        out["cpu_baseline_1thread"] = {"value": rows1 / dt2, "unit": "dets/s", "cores": 1, "kind": "port",
                                       "sample": f"{rows1} rows (every {istep1}-th), one thread; {dt2:.1f} s"}
        out["host"] = _host_info(nthr)
        log(f"bench: cpu oracle {rows1 / dt2:.1f} dets/s on 1 thread")
    del u, v, s
    torch.cuda.empty_cache()
    if not args.no_davidson:
        out["davidson"] = _guarded(davidson_leg, ctx, H, dets, log)
    ctx.close()
    if not args.no_selected:
        out["selected_davidson"] = _guarded(selected_davidson_leg, args, log)
    if not args.no_pt2:
        out["pt2"] = _guarded(pt2_leg, args, log)
    if not args.no_mlp:
        out["mlp"] = _guarded(mlp_leg, args, log)
    if not args.no_cipsi:
        out["cipsi_step"] = _guarded(cipsi_step_leg, args, log)
    print(json.dumps(out), flush=True)


def _host_info(threads_used):
    """CPU model and logical CPU count of the host the CPU baseline ran on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None
    return {"cpu_model": model, "logical_cpus": os.cpu_count(), "cpus_available": aff,
            "threads_used": threads_used}


def _guarded(fn, *a):
    try:
        return fn(*a)
    except Exception as e:   # the leg's failure is part of the record
        return {"error": f"{type(e).__name__}: {e}"}


def davidson_leg(ctx, H, dets, log):
    """FP64 vs mixed-precision davidson_diag_hs2, 3 roots on the full-CI space."""
    import numpy as np
    from quantum_package_amd.hamiltonian import hf_determinant
    n = dets.shape[0]
    hf = hf_determinant(H.M, H.n_alpha, H.n_beta)
    g = np.zeros((n, 3))
    i_hf = np.where(np.all(dets.reshape(n, -1) == hf.ravel()[None, :], axis=1))[0][0]
    g[i_hf, 0] = 1.0
    g[:, 1:] = np.random.default_rng(1002).standard_normal((n, 2))
    res = {}
    ctx.set_verbosity(1)   # iteration table on stderr
    for name, mixed in (("fp64", False), ("mixed", True)):
        log(f"bench: davidson {name} ...")
        t0 = time.time()
        e, s2, u, ns = ctx.davidson_diag_hs2(g, N_st=3, N_st_diag=4, threshold=1e-10, s2_eig=False, mixed=mixed,
                                             switch_threshold=1e-6)
        dt = time.time() - t0
        res[name] = {"seconds": dt, "energies": e[:3].tolist(), "s2": s2[:3].tolist(),
                     "sigma_calls_fp64": ns[0], "sigma_calls_fp32": ns[1]}
        log(f"bench: davidson {name}: {dt:.2f} s, E = {e[:3]}, sigma calls {ns}")
    d = np.abs(np.array(res["mixed"]["energies"]) - np.array(res["fp64"]["energies"]))
    res["abs_dE_per_root"] = d.tolist()
    res["speedup_mixed"] = res["fp64"]["seconds"] / res["mixed"]["seconds"]
    res["workload"] = ("davidson_diag_hs2, lowest 3 roots of H (no S^2 filter: seeded random guess vectors are "
                       "not spin-adapted) + 1 extra vector, threshold 1e-10 on the squared residual, HF + seeded "
                       "random guess, same 9,018,009-determinant space")
    return res


MFMA_BF16_PEAK = 2.5e15   # MI355X dense bf16 FLOP/s


def mlp_leg(args, log):
    """Block-proposal network throughput: Adam training epochs (batch 2048)
    and prediction on 262,144 synthetic block descriptors of the 96-orbital
    system, against the dense bf16 MFMA peak."""
    import numpy as np
    from quantum_package_amd.api import BlockMLP
    rng = np.random.default_rng(3)
    n, M = 262144, 96
    desc = np.stack([rng.integers(0, 10, n), rng.integers(0, 10, n), rng.integers(0, 3, n),
                     rng.integers(0, 5, n)], axis=1).astype(np.int32)
    fdesc = rng.standard_normal((n, 3)).astype(np.float32)
    y = rng.standard_normal(n).astype(np.float32)
    m = BlockMLP(n_orb=M, seed=7)
    m.train_epochs(desc[:16384], fdesc[:16384], y[:16384], 1e-3, 1, 2048, 1)   # warm-up
    macs = 80 * 256 + 256 * 256 * 2 + 256            # per sample, forward
    fwd = 2.0 * macs
    res = {}
    for batch in (2048, 16384):
        t = time.time()
        m.train_epochs(desc, fdesc, y, 1e-3, 2, batch, 2)
        dt = time.time() - t
        steps = 2 * (n // batch)
        fl = 3.0 * fwd * batch * steps
        res[f"train_batch{batch}"] = {"seconds": dt, "ms_per_step": 1e3 * dt / steps,
                                      "samples_per_s": batch * steps / dt, "tflops": fl / dt / 1e12}
        log(f"bench: mlp training, batch {batch}: {1e3 * dt / steps:.3f} ms/step, {fl / dt / 1e12:.2f} TFLOP/s")
    t = time.time()
    m.predict(desc, fdesc)
    dt = time.time() - t
    res["predict"] = {"seconds": dt, "samples_per_s": n / dt, "tflops": fwd * n / dt / 1e12,
                      "note": "host descriptors in, predictions out (chunks of 65536), transfers included"}
    log(f"bench: mlp prediction of {n} blocks: {1e3 * dt:.1f} ms ({fwd * n / dt / 1e12:.2f} TFLOP/s)")
    best = max(res["train_batch2048"]["tflops"], res["train_batch16384"]["tflops"])
    res["roofline"] = {"bound": "mfma", "achieved": best, "peak": MFMA_BF16_PEAK / 1e12, "unit": "TFLOP/s",
                       "frac": best * 1e12 / MFMA_BF16_PEAK}
    res["workload"] = ("x(80) -> 256 -> 256 -> 256 -> 1, BF16 weights, FP32 accumulation; 3.0e5 FLOP per block "
                       "forward, x3 for a training step; 262,144 synthetic descriptors (96 orbitals)")
    m.close()
    return res


def sparse_alg_bytes(n_det, visited, na, nb, nst, w):
    """Algorithmic bytes of one sparse sigma call.  From the visited-entry
    counts of the five bucket classes: S_a = V_a1 - N na ordered alpha-single
    pairs (S_b likewise); a pair of alpha singles shares n_b ab buckets and
    n_a - 1 aa buckets, a determinant meets itself in every slot, so the
    connected pairs are ab = V_ab - N na nb - S_b na - S_a nb and
    aa = V_aa - N na(na-1)/2 - S_a (na-1).  Bytes: 8 per visited payload, 8 per
    (determinant, slot) range, an nst*w row of u per connected pair, w per
    ab integral, 2w per same-spin double, (na+nb+1) w per single, and the
    determinant's own u, v, s rows."""
    V = visited
    sa = V["a1"] - n_det * na
    sb = V["b1"] - n_det * nb
    ab = V["ab"] - n_det * na * nb - sb * na - sa * nb
    aa = V["a2"] - n_det * na * (na - 1) // 2 - sa * (na - 1)
    bb = V["b2"] - n_det * nb * (nb - 1) // 2 - sb * (nb - 1)
    slots = na * nb + na + nb + na * (na - 1) // 2 + nb * (nb - 1) // 2
    pairs = ab + aa + bb + sa + sb
    b = (8 * sum(V.values()) + 8 * n_det * slots + pairs * nst * w + ab * w + 2 * (aa + bb) * w
         + (sa + sb) * (na + nb + 1) * w + 3 * n_det * nst * w)
    return b, {"ab": ab, "aa": aa, "bb": bb, "a": sa, "b": sb}


def selected_davidson_leg(args, log):
    """Mixed-precision vs FP64 davidson_diag_hs2 on a selected wavefunction
    of about 2^21 determinants (sparse sigma path), sigma kernel timings and
    the CPU oracle on a sample of rows."""
    import ctypes
    import numpy as np
    import torch
    from quantum_package_amd.api import FP32, FP64, Context
    from quantum_package_amd.cipsi import cipsi, n_leading, sort_by_weight
    from quantum_package_amd.hamiltonian import synthetic_hamiltonian
    M, na, nb, seed, ns, nst = 96, 10, 10, 1003, 2, 4
    H = synthetic_hamiltonian(M, na, nb, seed=seed)
    ctx = Context(0)
    t0 = time.time()
    r = cipsi(ctx, H, n_det_max=args.sel_small, n_states=ns, s2_eig=True, log=lambda m: None)
    ds, cs, contrib = sort_by_weight(r.dets, r.coef)
    n = ds.shape[0]
    n_gen = n_leading(contrib, 0.99)
    n_sel = max(n_leading(contrib, 0.999, True), n_gen)
    ctx.set_wavefunction(ds)
    E0 = np.array([ctx.u_0_H_u_0(cs[:, k])[0] for k in range(ns)])
    _, new, _ = ctx.ZMQ_selection(np.asfortranarray(cs), n_gen, n_sel, E0, args.sel_select)
    dets = np.concatenate([ds, new])
    t_s2 = time.time()
    extra = ctx.s2_complete(dets, na)   # make_s2_eigenfunction on the GPU
    t_s2 = time.time() - t_s2
    dets = np.concatenate([dets, extra])
    nd = dets.shape[0]
    t_build = time.time() - t0
    log(f"bench: selected psi {nd} dets ({n} by the CIPSI loop, {new.shape[0]} selected in one step, rest S^2 "
        f"completion) in {t_build:.1f}s")
    ctx.set_wavefunction(dets)
    t0 = time.time()
    visited = ctx._sparse_stats()
    t_buckets = time.time() - t0
    path = ctx._force_sigma_path(0)
    res = {"n_det": nd, "psi_build_seconds": t_build, "s2_complete": {"seconds": t_s2, "input": int(nd - extra.shape[0]),
                                                                       "added": int(extra.shape[0])},
           "bucket_build_seconds": t_buckets, "visited": visited}
    # sigma kernel timings, 4 vectors, FP64 and FP32; measured HBM bytes per
    # call (rocprofv3 FETCH_SIZE/WRITE_SIZE of this leg, calibrated) when the
    # committed summary is for this psi
    traffic = {}
    tpath = os.path.join(ROOT, "profiles", "sigma_sparse_selected_m96_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        if f"{nd:,}" in tj.get("workload", ""):
            traffic = tj["hbm_bytes_per_sigma"]
    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(device=dev)
    kern = {}
    for prec, dt, w, name in ((FP64, torch.float64, 8, "fp64"), (FP32, torch.float32, 4, "fp32")):
        gen = torch.Generator(device=dev)
        gen.manual_seed(11)
        u = torch.randn(nd * nst, dtype=dt, device=dev, generator=gen)
        v = torch.empty_like(u)
        s = torch.empty_like(u)
        torch.cuda.synchronize()
        kms = np.zeros(2, dtype=np.float32)
        ctx._check(ctx.L.qpx_debug_sigma_kernel_times(ctx.ptr, prec, nst, u.data_ptr(), v.data_ptr(), s.data_ptr(),
                                                      stream.cuda_stream,
                                                      kms.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), 5))
        b, pairs = sparse_alg_bytes(nd, visited, na, nb, nst, w)
        ms = float(kms[0])
        kern[name] = {"kernel_ms": ms, "dets_per_s": nd / (ms * 1e-3),
                      "roofline": {"bound": "hbm", "achieved": b / (ms * 1e-3) / 1e9, "peak": HBM_PEAK / 1e9,
                                   "unit": "GB/s", "frac": b / (ms * 1e-3) / HBM_PEAK, "alg_bytes": b,
                                   "traffic": traffic.get(name)}}
        log(f"bench: sparse sigma {name} {nst} vectors: {ms:.2f} ms ({nd / (ms * 1e-3):.3e} dets/s)")
        del u, v, s
    res["sigma_sparse"] = kern
    res["connected_pairs"] = pairs
    res["sigma_path"] = path
    g = np.zeros((nd, ns))
    g[:n] = cs
    dav = {}
    for name, mixed in (() if args.sel_no_davidson else (("fp64", False), ("mixed", True))):
        t = time.time()
        e, s2, _, nsig = ctx.davidson_diag_hs2(g, N_st=ns, N_st_diag=4, threshold=1e-10, s2_eig=True, mixed=mixed,
                                               switch_threshold=1e-6)
        dav[name] = {"seconds": time.time() - t, "energies": e[:ns].tolist(), "s2": s2[:ns].tolist(),
                     "sigma_calls_fp64": nsig[0], "sigma_calls_fp32": nsig[1]}
        log(f"bench: selected davidson {name}: {dav[name]['seconds']:.2f}s, E = {e[:ns]}, calls {nsig}")
    if dav:
        res["davidson"] = dav
        res["abs_dE_per_root"] = np.abs(np.array(dav["mixed"]["energies"]) -
                                        np.array(dav["fp64"]["energies"])).tolist()
        res["speedup_mixed"] = dav["fp64"]["seconds"] / dav["mixed"]["seconds"]
    if not args.no_cpu:
        from oracle import oracle as O
        ham = O.Ham(H)
        bil = O.Bilinear(dets)
        nthr = args.cpu_threads
        # calibrate on growing strided samples until one takes over a second
        # (the first call pays the thread start-up), then run ~cpu_seconds
        istep = max(1, nd // 256)
        while True:
            t0 = time.time()
            rows, _ = O.sigma_work_sample(ham, bil, nst, istep, nthr)
            dt0 = time.time() - t0
            if dt0 > 1.0 or istep == 1:
                break
            istep = max(1, istep // 4)
        rate0 = rows / max(dt0, 1e-6)
        istep = max(1, nd // max(int(rate0 * args.cpu_seconds), 1))
        t0 = time.time()
        rows, _ = O.sigma_work_sample(ham, bil, nst, istep, nthr)
        dt1 = time.time() - t0
        res["cpu_baseline"] = {"value": rows / dt1, "unit": "dets/s", "cores": nthr, "kind": "port",
                               "sample": f"{rows} of {nd} rows (every {istep}-th, bilinear order) of one FP64 "
                                         f"{nst}-vector sigma call, reference algorithm restated in C/OpenMP; "
                                         f"{dt1:.1f} s"}
        log(f"bench: selected psi cpu oracle {rows / dt1:.1f} dets/s on {nthr} threads")
    res["workload"] = (f"davidson_diag_hs2 on a selected wavefunction, synthetic Hamiltonian seed {seed}: {M} MOs, "
                       f"{na} alpha + {nb} beta electrons, N_int=2, {nd} determinants (CIPSI loop from HF to "
                       f"{n}, one selection step of {new.shape[0]}, S^2 completion); {ns} roots + 2 extra vectors, "
                       f"s2_eig, threshold 1e-10; mixed = FP32 sigma for correction vectors until the residual "
                       f"falls under 1e-6, then FP64")
    ctx.close()
    return res


def cipsi_step_leg(args, log):
    """Plain vs learned CIPSI step on a CIPSI-grown psi."""
    import numpy as np
    from quantum_package_amd.api import BlockMLP, Context
    from quantum_package_amd.cipsi import cipsi, n_leading, sort_by_weight
    from quantum_package_amd.hamiltonian import synthetic_hamiltonian
    from quantum_package_amd.proposal import Proposer, block_descriptors
    M, na, nb, seed, ns = 96, 10, 10, 1003, 2
    H = synthetic_hamiltonian(M, na, nb, seed=seed)
    ctx = Context(0)
    mlp = BlockMLP(n_orb=M, seed=7)
    prop = Proposer(mlp, epochs=20)
    t0 = time.time()
    r = cipsi(ctx, H, n_det_max=args.cipsi_ndet, n_states=ns, s2_eig=True, proposer=prop, full_sweep_every=2,
              log=lambda m: log("bench: " + m))
    log(f"bench: cipsi grown to {r.dets.shape[0]} dets in {time.time() - t0:.1f}s")
    ds, cs, contrib = sort_by_weight(r.dets, r.coef)
    n = ds.shape[0]
    n_gen = n_leading(contrib, 0.99)
    n_sel = max(n_leading(contrib, 0.999, True), n_gen)
    ctx.set_wavefunction(ds)
    E0 = np.array([ctx.u_0_H_u_0(cs[:, k])[0] for k in range(ns)])
    cf = np.asfortranarray(cs)
    to_select = n // 2 + 1
    # selection: full sweep, then the proposed blocks (network trained on the
    # loop's last full sweep)
    t = time.time()
    pt2_full, new_full, _, bpt2, _ = ctx.select_blocks(cf, n_gen, n_sel, E0, to_select)
    t_full = time.time() - t
    ticks = ctx._select_block_ticks(bpt2.shape[0]).astype(np.float64)
    t = time.time()
    desc, fdesc = block_descriptors(H, ds, cs, n_gen)
    keep = prop.propose(desc, fdesc)
    t_pred = time.time() - t
    t = time.time()
    pt2_prop, new_prop, _, _, _ = ctx.select_blocks(cf, n_gen, n_sel, E0, to_select, keep=keep)
    t_prop = time.time() - t
    missed = float(1.0 - pt2_prop[0] / pt2_full[0])
    log(f"bench: selection full {t_full:.2f}s, proposed {t_prop:.2f}s + {t_pred:.2f}s descriptors/prediction, "
        f"kept {keep.mean():.3f}, missed PT2 {missed:.2e}")
    # what any block subset could do: with the exact block PT2 and the
    # measured block durations of the full sweep, the least kernel time that
    # keeps 1 - m of the PT2 (blocks ranked by |PT2| per tick)
    e = np.abs(bpt2[:, 0])
    order = np.argsort(-(e / np.maximum(ticks, 1.0)), kind="stable")
    ce = np.cumsum(e[order]) / e.sum()
    ct = np.cumsum(ticks[order]) / ticks.sum()
    bound = {}
    for m in (max(missed, 1e-6), 5e-3, 5e-4):
        k = min(int(np.searchsorted(ce, 1.0 - m)), ce.size - 1)
        bound[f"{m:.2e}"] = {"kernel_time_fraction": float(ct[k]), "selection_speedup": float(1.0 / ct[k])}
    keep_b = keep.astype(bool)
    log(f"bench: proposed blocks hold {ticks[keep_b].sum() / ticks.sum():.3f} of the kernel time; exact-ranking "
        f"bound at 5e-3 missed: {bound['5.00e-03']['selection_speedup']:.2f}x")
    # Davidson on the grown psi (the full sweep's determinants), FP64 then mixed
    dets = np.concatenate([ds, new_full])
    extra = ctx.s2_complete(dets, na)
    dets = np.concatenate([dets, extra])
    g = np.zeros((dets.shape[0], ns))
    g[:n] = cs
    ctx.set_wavefunction(dets)
    # the wavefunction's one-time sigma set-up (sparse buckets) before either
    # timing, so that the two Davidson runs are compared on equal terms
    t = time.time()
    ctx.H_S2_u_0_nstates(np.zeros((dets.shape[0], 1)))
    t_setup = time.time() - t
    res = {}
    for name, mixed in (("fp64", False), ("mixed", True)):
        t = time.time()
        e, s2, _, nsig = ctx.davidson_diag_hs2(g, N_st=ns, N_st_diag=4, threshold=1e-10, s2_eig=True, mixed=mixed,
                                               switch_threshold=1e-6)
        res[name] = (time.time() - t, e[:ns], nsig)
        log(f"bench: davidson {name} on {dets.shape[0]} dets: {res[name][0]:.2f}s, E = {e[:ns]}, calls {nsig}")
    t_plain = res["fp64"][0] + t_full
    t_learned = res["mixed"][0] + t_prop + t_pred
    # the network is trained on each full sweep (every second iteration):
    # charged here to the proposed iteration that follows it
    trains = [h.get("t_train") for h in r.history if isinstance(h, dict) and h.get("t_train")]
    t_train = float(trains[-1]) if trains else 0.0
    mlp.close()
    ctx.close()
    return {"workload": f"one CIPSI step, synthetic Hamiltonian seed {seed}: {M} MOs, {na} alpha + {nb} beta "
                        f"electrons, N_int=2, {ns} states, s2_eig; psi of {n} determinants grown by the CIPSI "
                        f"loop (full sweeps every second iteration train the block proposer), selection of "
                        f"N_det/2+1 = {to_select} with {n_gen} generators and {n_sel} selectors, then Davidson "
                        f"(threshold 1e-10) on the {dets.shape[0]} determinants of the grown psi",
            "plain_seconds": t_plain, "learned_seconds": t_learned, "speedup": t_plain / t_learned,
            "training_seconds_last_sweep": t_train,
            "speedup_with_training": t_plain / (t_learned + t_train),
            "selection_full_seconds": t_full, "selection_proposed_seconds": t_prop,
            "proposal_overhead_seconds": t_pred, "blocks_kept": float(keep.mean()), "missed_pt2_fraction": missed,
            "kept_kernel_time_fraction": float(ticks[keep_b].sum() / ticks.sum()),
            "exact_ranking_bound": bound,
            "note": ("exact_ranking_bound: the selection speed-up of the best block subset at each missed-PT2 "
                     "fraction, from the full sweep's exact block PT2 and measured block durations; the full "
                     "sweep is timed through the block-level entry point (it also returns the block arrays)"),
            "davidson_fp64_seconds": res["fp64"][0], "davidson_mixed_seconds": res["mixed"][0],
            "sigma_setup_seconds": t_setup,
            "davidson_sigma_calls": {"fp64": list(res["fp64"][2]), "mixed": list(res["mixed"][2])},
            "abs_dE_per_root": np.abs(res["mixed"][1] - res["fp64"][1]).tolist()}


def _selected_psi(ctx, H, n_target, log):
    """HF, one selection step of n_target-1 determinants, diagonalised:
    returns (dets, coef) in psi_det_sorted order and E."""
    import numpy as np
    from quantum_package_amd.cipsi import sort_by_weight
    from quantum_package_amd.hamiltonian import hf_determinant
    hf = hf_determinant(H.M, H.n_alpha, H.n_beta)[None]
    ctx.set_wavefunction(hf)
    E0 = ctx.u_0_H_u_0(np.ones(1))
    _, new, _ = ctx.ZMQ_selection(np.ones((1, 1), order="F"), 1, 1, E0, n_target - 1)
    dets = np.concatenate([hf, new])
    ctx.set_wavefunction(dets)
    g = np.zeros((dets.shape[0], 1))
    g[0, 0] = 1.0
    e, _, u, ns = ctx.davidson_diag_hs2(g, N_st=1, N_st_diag=4, threshold=1e-8, s2_eig=False)
    ds, cs, _ = sort_by_weight(dets, u[:, :1])
    log(f"bench: selected psi {dets.shape[0]} dets, E = {e[0]:.10f} ({ns} sigma calls)")
    return ds, np.asfortranarray(cs), e[0]


def pt2_leg(args, log):
    """Selection + EN-PT2 throughput, candidates/s."""
    import numpy as np
    from quantum_package_amd.api import Context
    from quantum_package_amd.hamiltonian import synthetic_hamiltonian
    M, na, nb, seed = 128, 12, 12, 1004
    t0 = time.time()
    H = synthetic_hamiltonian(M, na, nb, seed=seed)
    ctx = Context(0)
    ctx.set_hamiltonian(H)
    log(f"bench: pt2 system M={M} {na}a{nb}b set up in {time.time() - t0:.1f}s")
    ds, cs, _ = _selected_psi(ctx, H, args.pt2_ndet, log)
    n = ds.shape[0]
    ctx.set_wavefunction(ds)
    E0 = ctx.u_0_H_u_0(cs[:, 0])

    def run_tasks(gens):
        t = time.time()
        pt2, sel, _, _ = ctx.select_tasks(cs, n, E0, n, [(int(g), 0) for g in gens])
        return time.time() - t, pt2, ctx.select_stats(), sel.shape[0]

    log("bench: pt2 warm-up on generators 1..8 ...")
    run_tasks(range(1, 9))
    # Stratified sample of the whole generator range: the 64 leading
    # generators (the heaviest, HF among them) all, then 10 strata of equal
    # width over the rest with args.pt2_per_stratum generators drawn in each;
    # each stratum is one selection call (to_select = N_det) timed on its own
    # and weighted by (stratum width / drawn) to project the full sweep.
    rng = np.random.default_rng(seed)
    head = np.arange(1, min(n, 64) + 1)
    edges = np.linspace(head[-1] + 1, n + 1, 11).astype(int)
    strata = [(head, 1.0)]
    for a, b in zip(edges[:-1], edges[1:]):
        if b > a:
            k = min(args.pt2_per_stratum, b - a)
            strata.append((np.sort(rng.choice(np.arange(a, b), size=k, replace=False)), (b - a) / k))
    rows = []
    t_tot = c_tot = 0.0
    t_proj = c_proj = 0.0
    g_tot = b_tot = 0
    for gens, wgt in strata:
        dt, _, st, _ = run_tasks(gens)
        rows.append({"first": int(gens[0]), "last": int(gens[-1]), "drawn": int(gens.size), "weight": float(wgt),
                     "seconds": dt, "candidates": st["candidates"], "blocks": st["blocks"]})
        t_tot += dt
        c_tot += st["candidates"]
        t_proj += wgt * dt
        c_proj += wgt * st["candidates"]
        g_tot += st["gathers"]
        b_tot += st["blocks"]
        log(f"bench: pt2 stratum {int(gens[0])}..{int(gens[-1])} ({gens.size} drawn): {dt:.2f} s, "
            f"{st['candidates']} candidates")
    n_drawn = int(sum(r["drawn"] for r in rows))
    res = {"value": c_tot / t_tot, "unit": "candidates/s", "dtype": "f64", "seconds": t_tot,
           "generators": n_drawn, "selectors": n, "candidates": int(c_tot), "blocks": b_tot,
           "projected_full_sweep": {"seconds": t_proj, "candidates": c_proj, "candidates_per_s": c_proj / t_proj,
                                    "generators": n},
           "strata": rows,
           "workload": f"selection + EN-PT2, synthetic Hamiltonian seed {seed}: {M} MOs, {na} alpha + {nb} beta "
                       f"electrons, N_int=2; psi = {n} determinants (HF + one selection step, diagonalised), all of "
                       f"them selectors, to_select = N_det; a stratified sample of {n_drawn} of the {n} generators "
                       f"(the {head.size} leading ones, then {args.pt2_per_stratum} drawn in each of 10 equal "
                       f"strata), one timed selection call per stratum"}
    # algorithmic bytes of the selection: 8 per integral/coefficient gather,
    # (8 N_states + 16) per evaluated candidate, 8 N_states M^2 per block tile
    b_alg = 8 * g_tot + c_tot * (8 * 1 + 16) + b_tot * 8 * 1 * M * M
    res["roofline"] = {"bound": "hbm", "achieved": b_alg / t_tot / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                       "frac": b_alg / t_tot / HBM_PEAK, "alg_bytes": b_alg,
                       "note": "whole qpx_select_tasks calls of the sample (host block enumeration and final "
                               "sort included)"}
    log(f"bench: pt2 sample of {n_drawn} generators: {t_tot:.2f} s, {res['value']:.3e} candidates/s; projected "
        f"full sweep {t_proj:.1f} s")
    if not args.no_pt2_stoch:
        # the full EN-PT2 of this psi (every determinant a generator) is out of
        # reach exactly; the hybrid stochastic estimate, capped in time
        from quantum_package_amd.pt2_stoch import ZMQ_pt2
        t = time.time()
        est, err, info = ZMQ_pt2(ctx, ds, cs, np.atleast_1d(E0), relative_error=1e-3, absolute_error=1e-9,
                                 seed=11, combs_per_batch=5, max_seconds=args.pt2_seconds,
                                 log=lambda m: log("bench: " + m))
        res["stochastic"] = {"pt2": est.tolist(), "error": err.tolist(), "seconds": time.time() - t,
                             "combs": info[0]["combs"], "generators_computed": info[0]["generators_computed"],
                             "generators": info[0]["generators"], "deterministic_set": info[0]["deterministic"],
                             "workload": "hybrid stochastic EN-PT2 of the same psi, every determinant a generator "
                                         "and a selector, 100 teeth, 5 combs per batch, stopped at relative "
                                         f"error 1e-3 or after {args.pt2_seconds:.0f} s"}
        log(f"bench: stochastic PT2 {est[0]:.8f} +/- {err[0]:.2e} from {info[0]['generators_computed']} of "
            f"{info[0]['generators']} generators in {time.time() - t:.1f} s")
    if not args.no_cpu:
        # the reference's task farm: (generator, fragment) tasks, fragment
        # count n_alpha^2 (selection.irp.f:561-562), leading generators first
        from oracle import oracle as O
        ham = O.Ham(H)
        nfrag = na * na

        def cpu(k):
            tasks = np.array([[1 + i // nfrag, 1 + i % nfrag] for i in range(k)], dtype=np.int32)
            t = time.time()
            r = O.selection(ham, ds, cs, int(tasks[-1, 0]), n, E0, n, nthreads=args.cpu_threads, tasks=tasks)
            return time.time() - t, r

        k = 2 * args.cpu_threads
        dtc, r = cpu(k)
        k2 = int(min(64 * nfrag, k * args.cpu_seconds / max(dtc, 1e-3)))
        if k2 > k:
            k = k2
            dtc, r = cpu(k)
        res["cpu_baseline"] = {"value": r["n_candidates"] / dtc, "unit": "candidates/s", "cores": args.cpu_threads,
                               "kind": "port",
                               "sample": f"{k} selection tasks (generator, fragment of {nfrag}) over the leading "
                                         f"generators of the same psi, reference selection restated in "
                                         f"C/OpenMP; {r['n_candidates']} candidates in {dtc:.1f} s"}
        log(f"bench: pt2 cpu oracle {res['cpu_baseline']['value']:.3e} candidates/s on {args.cpu_threads} threads")
    ctx.close()
    return res


if __name__ == "__main__":
    main()
