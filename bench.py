"""Benchmark: whole-volume SHORE-8 FISTA-L1 fit on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp16|fp32|bf16]

A step is one whole-volume fit (``shorefit_fit``) of the whole-brain-shaped
workload: a 96x96x60 ellipsoid mask (250,176 voxels), 258 measurements (1 b0 +
257 Cartesian q-points, bmax 5000), radial order 8 (145 coefficients),
FISTA-L1 with fixed alpha = 1e-3, max_iter 500, tol 1e-6. Inputs (DWI volume,
mask) and outputs are resident in HBM before the timed region. Prints one JSON
line with the throughput (voxel-fits/s), the roofline of the dominant kernel,
the accuracy of a sample of voxels against the FP64 oracle, and the oracle's
CPU throughput on that sample (the CPU-baseline leg).
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "voxel fits/s (SHORE-8 FISTA, coef err≤1e-4 vs FP64) + % MFMA/HBM roofline"
PEAK = {"fp32": 157.3, "fp16": 2500.0, "bf16": 2500.0}  # dense MFMA TFLOP/s, MI355X
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
KERNEL = {"fp32": "fista_f32_persist_kernel<10,8,true>",
          "fp16": "fista_mixed_kernel<10,8,false,true,false,false,false,145>",
          "bf16": "fista_mixed_kernel<10,8,true,true,false,false,false,145>"}
# HBM bytes per launch of the solve kernel on this workload, from rocprofv3
# --pmc FETCH_SIZE and WRITE_SIZE passes (KiB per dispatch), each scaled by
# what the counter reads for a plain 1 GiB device copy (bench.py
# --copy-calibration 1024); the summaries under profiles/ hold the commands.
TRAFFIC_FILE = ROOT / "profiles" / "traffic.json"


def traffic_bytes(precision: str):
    try:
        t = json.loads(TRAFFIC_FILE.read_text())[precision]
    except (OSError, KeyError, ValueError):
        return None
    return round(t["fetch_kib"] * t["fetch_scale"] * 1024 + t["write_kib"] * t["write_scale"] * 1024)


def cpu_threads():
    """(threads for the CPU baseline, CPUs this process may run on, host CPU
    count). The threads are the CPUs in this process's affinity mask, capped
    by OMP_NUM_THREADS when the environment sets it (the CPU share a shared
    GPU host allots to one GPU's job)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return (min(avail, share) if share > 0 else avail), avail, os.cpu_count() or avail


def cpu_fit_rate(orc, S, A, N, K, *, threads, **kw):
    """FP64 oracle fits of the rows of S (one voxel per task on `threads`
    threads): (fits/s, wall s, result dict)."""
    t0 = time.perf_counter()
    ref = orc.fit_volume(S.reshape(len(S), 1, 1, N), None, A, n_threads=threads, **kw)
    wall = time.perf_counter() - t0
    return len(S) / wall, wall, ref


def err_stats(c, ref, its=None, its_ref=None) -> dict:
    """Per-voxel relative coefficient error of c against ref (rows)."""
    c, ref = np.asarray(c, np.float64), np.asarray(ref, np.float64)
    e = np.linalg.norm(c - ref, axis=1) / np.maximum(np.linalg.norm(ref, axis=1), 1e-30)
    out = {"sample": len(e), "coef_rel_err_median": float(np.median(e)), "coef_rel_err_p99": float(np.quantile(e, 0.99)),
           "coef_rel_err_max": float(e.max()), "frac_le_1e-4": float(np.mean(e <= 1e-4))}
    if its is not None:
        out["iterations_equal_frac"] = float(np.mean(np.asarray(its) == np.asarray(its_ref)))
    return out


def refresh_used(args, Kp: int) -> int:
    """The refresh period the library applies: --refresh, else its default
    for the arithmetic (FP16 with split increments, padded K <= 160: 40;
    otherwise 10)."""
    if args.refresh:
        return args.refresh
    return 40 if (args.precision == "fp16" and Kp <= 160) else 10


def d_Kp(K: int, precision: str) -> int:
    """Padded coefficient count the kernels use (multiples of 16 / 32)."""
    sizes = [16, 32, 48, 64, 80, 96, 128, 160, 192, 256] if precision == "fp32" else [32, 64, 96, 128, 160, 256]
    return next(s for s in sizes if s >= K)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--precision", default="fp16", choices=["fp32", "fp16", "bf16"])
    p.add_argument("--refresh", type=int, default=None,
                   help="exact-refresh period of the mixed kernels (default: the library's for the arithmetic)")
    p.add_argument("--cpu-sample", type=int, default=32768)
    p.add_argument("--cpu-single-sample", type=int, default=1024,
                   help="voxels of the single-thread CPU-baseline run")
    p.add_argument("--equal-iter-sample", type=int, default=4096,
                   help="voxels of the equal-iteration (tol = 0) accuracy check")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-postfit", action="store_true", help="skip the post-fit map timings")
    p.add_argument("--precision-report", action="store_true",
                   help="only print the precision report of the 16x16x8 order-6 volume (device fits at FP32, "
                        "BF16 and FP16 with refresh 5/10/20 against the FP64 oracle fit) and exit")
    p.add_argument("--max-iter", type=int, default=500)
    p.add_argument("--alpha-mode", default="fixed", choices=["fixed", "alpha-ratio", "path-bic", "l2-anchored"],
                   help="alpha strategy (the headline line is the fixed-alpha fit)")
    p.add_argument("--workload", default="shore8", choices=["shore8", "tikhonov", "graddev", "adjoint", "io"],
                   help="shore8: the shared-design whole-brain fit (the headline line); tikhonov: the same volume "
                        "fitted with Tikhonov L2 (lambda_N = lambda_L = 1e-2); io: the volume's NIfTI read and "
                        "coefficient write path; graddev: the same volume "
                        "with per-voxel gradient-deviation tensors and mixed SNR (one line per --orders entry)")
    p.add_argument("--orders", default="6,8,10", help="radial orders of the graddev workload")
    p.add_argument("--tune", action="store_true", help="graddev: autotune the tile shapes first")
    p.add_argument("--graddev-sample", type=int, default=256,
                   help="graddev: voxels checked against the FP64 oracle (also the CPU-baseline sample)")
    p.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                   help="only run a plain device copy of MIB MiB (three launches) and exit: the known-size "
                        "reference for scaling HBM byte counters")
    return p.parse_args()


def copy_calibration(mib: int):
    import torch
    n = mib << 20
    src = torch.ones(n, dtype=torch.uint8, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {"bytes_read": n, "bytes_written": n, "launches": 3}}), flush=True)


def precision_report():
    from cs_dmri_amd import workloads as W
    from cs_dmri_amd.precision import report
    from oracle import load
    wl = W.config1()
    g = wl.gtab
    orc = load()
    A = orc.design_matrix(6, 700.0, g.bvals, g.bvecs, g.big_delta, g.small_delta)
    t0 = time.perf_counter()
    ref = orc.fit_volume(wl.data, wl.mask, A, alpha_mode="fixed", alpha=1e-3, max_iter=500, tol=1e-6,
                         n_threads=cpu_threads()[0])["coefficients"]
    rep = report(wl.data, wl.mask, g, ref)
    rep["workload"] = "shore6-cartesian101-16x16x8 (2048 voxels), fista-l1 alpha=1e-3, max_iter 500, tol 1e-6"
    rep["fp64_reference_s"] = round(time.perf_counter() - t0, 2)
    print(json.dumps({"precision_report": rep}), flush=True)


def postfit_timings(L, d_coef, d_mask, wl, K, N, dev):
    """Post-fit maps on the fitted volume, device-resident: the Tournier SH
    projection (lmax 8), synthesis on the fit's own table and the six
    microstructure maps (with fit-failure rejection). Mean of 3 calls after
    one warm-up, each a whole entry-point call (set-up, kernels, stream sync).
    These maps are HBM-bound (a K-vector in, a few floats out per voxel)."""
    import torch
    nx, ny, nz = wl.mask.shape
    nv = nx * ny * nz
    g = wl.gtab
    sh = torch.empty((nx, ny, nz, 45), dtype=torch.float32, device=dev)
    pred = torch.empty((nx, ny, nz, N), dtype=torch.float32, device=dev)
    maps = [torch.empty((nx, ny, nz), dtype=torch.float32, device=dev) for _ in range(6)]
    grads, _ = L.gradients(g.bvals, g.bvecs, g.big_delta, g.small_delta, g.b0_threshold)
    calls = {
        "odf_sh": (lambda: L.lib().shorefit_odf_sh(d_coef.data_ptr(), d_mask.data_ptr(), nx, ny, nz, 8, wl.zeta, 8,
                                                   sh.data_ptr(), 1, None),
                   wl.n_voxels * 4 * (K + 45)),
        "predict": (lambda: L.lib().shorefit_predict(d_coef.data_ptr(), nx, ny, nz, 8, wl.zeta, C.byref(grads),
                                                     pred.data_ptr(), 1, None),
                    nv * 4 * (K + N)),
        "microstructure": (lambda: L.lib().shorefit_microstructure(d_coef.data_ptr(), d_mask.data_ptr(), nx, ny, nz,
                                                                   8, wl.zeta, None, L.UNITS_UM, 10.0,
                                                                   *[m.data_ptr() for m in maps], 1, None),
                           wl.n_voxels * 4 * (K + 6)),
    }
    out = {}
    for name, (fn, nbytes) in calls.items():
        L.check(fn())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(3):
            L.check(fn())
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / 3 * 1e3
        out[name] = {"ms": round(ms, 3), "algorithmic_bytes": nbytes, "GB_s": round(nbytes / ms / 1e6, 1)}
    return out


def io_bench(args):
    """The on-disk path around the fit at whole-brain size: the 96x96x60x258
    series stored as int16 with scl_slope / scl_inter (the usual scanner
    export) and as float32, read into the voxel-major f32 volume the fit
    uses, and the (X, Y, Z, 145) coefficient volume written back. The device
    kernel (``to_voxel_major_kernel``, a type conversion and a 4-axis
    reversal) is HBM-bound: raw bytes in + 4 bytes out per element; its time
    is measured with HIP events on its stream. Wall times of the whole calls
    (disk, PCIe, kernel) are reported beside it; the CPU leg is numpy doing
    the same conversion (astype + scale + transpose) on the host."""
    import tempfile

    import torch

    from cs_dmri_amd import io as sio
    from cs_dmri_amd import nifti as nf
    from cs_dmri_amd import workloads as W
    from cs_dmri_amd import _lib as L

    wl = W.config2()
    data = wl.data  # (96, 96, 60, 258) float32
    nx, ny, nz, nt = data.shape
    slope, inter = 0.05, 0.0
    q = np.clip(np.rint(data / slope), -32768, 32767).astype(np.int16).transpose(3, 2, 1, 0)
    dev = torch.device("cuda:0")
    res = {"metric": "NIfTI DWI read to voxel-major f32 (GB/s of the device conversion)", "unit": "GB/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "int16->f32", "data": "synthetic (the config-2 phantom, quantised)",
           "config": {"workload": "dwi-96x96x60x258-int16-scl", "elements": int(data.size)}}
    with tempfile.TemporaryDirectory() as td:
        p16 = os.path.join(td, "dwi16.nii")
        nf.save(p16, nf.new_header(data.shape, np.diag([2.0, 2.0, 2.0, 1.0]), datatype=4, bitpix=16), q)
        h = nf.read_header(p16)
        h["scl_slope"], h["scl_inter"] = slope, inter
        nf.save(p16, h, q)
        img = nf.load(p16)
        raw = np.require(img.data.reshape(-1), requirements=["C", "W"])
        d_raw = torch.from_numpy(raw.view(np.uint8)).to(dev)
        out = torch.empty((nx, ny, nz, nt), dtype=torch.float32, device=dev)
        stream = torch.cuda.current_stream(dev)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        for _ in range(args.warmup):
            L.check(L.lib().shorefit_volume_to_voxel_major(d_raw.data_ptr(), 4, nx, ny, nz, nt, slope, inter,
                                                           out.data_ptr(), 1, stream.cuda_stream))
        torch.cuda.synchronize()
        ms = []
        for _ in range(args.steps):
            e0.record(stream)
            L.check(L.lib().shorefit_volume_to_voxel_major(d_raw.data_ptr(), 4, nx, ny, nz, nt, slope, inter,
                                                           out.data_ptr(), 1, stream.cuda_stream))
            e1.record(stream)
            torch.cuda.synchronize()
            ms.append(e0.elapsed_time(e1))
        k_ms = float(np.mean(ms))
        nbytes = data.size * (2 + 4)
        res["value"] = round(nbytes / k_ms / 1e6, 1)
        res["ms_per_step"] = round(k_ms, 4)
        res["roofline"] = {"bound": "hbm", "achieved": res["value"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": round(nbytes / k_ms / 1e6 / HBM_PEAK_GBS, 4), "traffic": None,
                           "kernel": "to_voxel_major_kernel<int16_t>", "kernel_ms": round(k_ms, 4),
                           "algorithmic_bytes": nbytes}
        want = (q.astype(np.float32) * np.float32(slope) + np.float32(inter)).transpose(3, 2, 1, 0)
        res["parity"] = {"bit_identical_to_numpy": bool(np.array_equal(out.cpu().numpy(), want))}
        # whole calls: file -> device volume, host volume, coefficient file write
        t0 = time.perf_counter()
        v_dev = sio.read_dwi_volume(p16, device=dev)
        torch.cuda.synchronize()
        t_read_dev = time.perf_counter() - t0
        t0 = time.perf_counter()
        sio.read_dwi_volume(p16)
        t_read_host = time.perf_counter() - t0
        coef = torch.zeros((nx, ny, nz, 145), dtype=torch.float32, device=dev)
        coef[..., :] = v_dev[..., :145]
        side = {"n_coefficients": 145}
        t0 = time.perf_counter()
        sio.write_coefficients(os.path.join(td, "coef.nii"), coef, p16, side)
        t_write = time.perf_counter() - t0
        res["wall_s"] = {"read_to_device": round(t_read_dev, 3), "read_to_host": round(t_read_host, 3),
                         "write_coefficients_from_device": round(t_write, 3)}
        if not args.no_cpu:
            t0 = time.perf_counter()
            ref = np.ascontiguousarray((img.data.astype(np.float32) * np.float32(slope) + np.float32(inter))
                                       .transpose(3, 2, 1, 0))
            cpu_s = time.perf_counter() - t0
            res["cpu_baseline"] = {"value": round(nbytes / cpu_s / 1e9, 2), "unit": "GB/s", "cores": 1,
                                   "kind": "port", "sample": f"the whole series, numpy astype + scale + transpose, "
                                                             f"{cpu_s:.2f} s"}
            res["parity"]["cpu_equal"] = bool(np.array_equal(ref, want))
    print(json.dumps(res), flush=True)


def tikhonov_bench(args):
    """The whole-brain volume (250,176 voxels, order 8) fitted with the
    Tikhonov L2 solver, lambda_N = lambda_L = 1e-2 (src/solver/tikhonov.rs:26-76).
    A step is one whole ``shorefit_fit``: design, G, Cholesky of G + Lambda,
    P = (G + Lambda)^-1 A^T (FP64, on the device), then c = P s and the
    residual per voxel (``tikhonov_h16_kernel``: FP16 MFMA on hi/lo splits of
    both operands, FP32 accumulation, signals read once). The roofline is
    that kernel's compulsory HBM bytes, 4 (N + K + 1) per voxel (signals in,
    coefficients and residual out), against 8 TB/s; the FP16 MFMA flops it
    executes (3 products per split pair, padded shapes) are reported beside
    it against the FP16 peak."""
    import torch

    from cs_dmri_amd import _lib as L
    from cs_dmri_amd import workloads as W
    from cs_dmri_amd.models import _spec

    dev = torch.device("cuda:0")
    wl = W.config2()
    g = wl.gtab
    K = L.lib().shorefit_n_coeffs(wl.radial_order)
    N = len(g.bvals)
    V = wl.n_voxels
    shape = wl.mask.shape
    d_data = torch.from_numpy(wl.data).to(dev)
    d_mask = torch.from_numpy(wl.mask.astype(np.uint8)).to(dev)
    d_coef = torch.empty(shape + (K,), dtype=torch.float32, device=dev)
    d_res = torch.empty(shape, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    spec = _spec(wl.radial_order, wl.zeta, "l2", "fixed", 1e-3, 1e-3, 20, None, 0.05, args.max_iter, 1e-6,
                 False, 1e-2, 1e-2, True, False, "fp32", args.refresh)
    grads, keep = L.gradients(g.bvals, g.bvecs, g.big_delta, g.small_delta, g.b0_threshold)
    vol = L.ShorefitVolume(d_data.data_ptr(), d_mask.data_ptr(), shape[0], shape[1], shape[2], 1, 0)

    def step():
        out = L.ShorefitOutput()
        out.coefficients = d_coef.data_ptr()
        out.residual_l2 = d_res.data_ptr()
        L.check(L.lib().shorefit_fit(C.byref(grads), C.byref(vol), C.byref(spec), C.byref(out), None))
        return out

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    kernel_ms = []
    for _ in range(args.steps):
        kernel_ms.append(step().fit_ms)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_s = float(np.mean(kernel_ms)) * 1e-3
    flops = 4.0 * N * K * V
    nbytes = 4.0 * (N + K + 1) * V
    Kp, Np2 = 16 * ((K + 15) // 16), 288  # padded shapes of the apply kernel
    exec_flops = 3 * 2.0 * (2 * Kp * Np2) * V
    result = {
        "metric": "voxel fits/s (SHORE-8 Tikhonov L2)", "value": round(V * args.steps / wall, 1),
        "unit": "voxel-fits/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp16 hi/lo split, fp32 accumulation",
        "data": "synthetic (seeded multi-tensor + Rician phantom)",
        "config": {"workload": "shore8-cartesian258-96x96x60-ellipsoid", "voxels": V, "n_grads": N, "n_coeffs": K,
                   "radial_order": 8, "solver": "tikhonov-l2 lambda_n=lambda_l=1e-2"},
        "roofline": {"bound": "hbm", "achieved": round(nbytes / kern_s / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(nbytes / kern_s / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": traffic_bytes("tikhonov"),
                     "kernel": "tikhonov_h16_kernel<10,18,4>", "kernel_ms": round(kern_s * 1e3, 3),
                     "algorithmic_bytes": nbytes, "algorithmic_flops": flops,
                     "executed": {"mfma_flops_per_launch": exec_flops,
                                  "tflops": round(exec_flops / kern_s / 1e12, 3),
                                  "frac": round(exec_flops / kern_s / 1e12 / PEAK["fp16"], 4)}},
    }
    if not args.no_cpu:
        from oracle import load
        orc = load()
        A = orc.design_matrix(wl.radial_order, wl.zeta, g.bvals, g.bvecs, g.big_delta, g.small_delta)
        dn, dl = orc.regularization(wl.radial_order)
        S = wl.data[wl.mask]
        pick = np.arange(V)  # a Tikhonov fit is one solve per voxel: the whole mask takes about a second
        threads, _, host = cpu_threads()
        t0 = time.perf_counter()
        ref = orc.fit_volume(S[pick].reshape(len(pick), 1, 1, N), None, A, regularization="l2", dn=dn, dl=dl,
                             lambda_n=1e-2, lambda_l=1e-2, n_threads=threads)
        cpu_wall = time.perf_counter() - t0
        cr = ref["coefficients"].reshape(len(pick), K).astype(np.float64)
        cg = d_coef.cpu().numpy()[wl.mask][pick].astype(np.float64)
        err = np.linalg.norm(cg - cr, axis=1) / np.maximum(np.linalg.norm(cr, axis=1), 1e-30)
        rg = d_res.cpu().numpy()[wl.mask][pick].astype(np.float64)
        rr = ref["residual_l2"].reshape(len(pick)).astype(np.float64)
        result["cpu_baseline"] = {"value": round(len(pick) / cpu_wall, 2), "unit": "voxel-fits/s", "cores": threads,
                                  "host_cores": host, "kind": "port",
                                  "sample": f"all {len(pick)} masked voxels, same fit settings, "
                                            f"FP64 oracle (C++), {cpu_wall:.2f} s wall"}
        result["accuracy_vs_fp64"] = {"sample": len(pick), "coef_rel_err_median": float(np.median(err)),
                                      "coef_rel_err_max": float(err.max()),
                                      "residual_rel_err_max": float(np.max(np.abs(rg - rr) / np.maximum(rr, 1e-30)))}
    print(json.dumps(result), flush=True)


def graddev_bench(args):
    """Per-voxel-design workload: the whole-brain volume with T = I + E per
    masked voxel (E symmetric, N(0, 0.02^2) entries) and per-voxel SNR
    U(5, 40); every voxel is fitted with its own effective table and design.
    One JSON line per radial order."""
    import torch

    from cs_dmri_amd import _lib as L
    from cs_dmri_amd import workloads as W
    from cs_dmri_amd.models import _graddev_struct, _spec

    dev = torch.device("cuda:0")
    wl, field = W.config4()
    g = wl.gtab
    N = len(g.bvals)
    V = wl.n_voxels
    shape = wl.mask.shape
    d_data = torch.from_numpy(wl.data).to(dev)
    d_mask = torch.from_numpy(wl.mask.astype(np.uint8)).to(dev)
    d_field = torch.from_numpy(np.ascontiguousarray(field.reshape(shape + (9,)))).to(dev)
    d_iter = torch.empty(shape, dtype=torch.int32, device=dev)
    d_res = torch.empty(shape, dtype=torch.float32, device=dev)
    grads, keep = L.gradients(g.bvals, g.bvecs, g.big_delta, g.small_delta, g.b0_threshold)
    vol = L.ShorefitVolume(d_data.data_ptr(), d_mask.data_ptr(), shape[0], shape[1], shape[2], 1, 0)
    pick = np.linspace(0, V - 1, args.graddev_sample).astype(np.int64)
    flat_idx = np.flatnonzero(wl.mask.ravel())[pick]
    for order in [int(o) for o in args.orders.split(",")]:
        K = L.lib().shorefit_n_coeffs(order)
        d_coef = torch.empty(shape + (K,), dtype=torch.float32, device=dev)
        spec = _spec(order, wl.zeta, "l1", "fixed", 1e-3, 1e-3, 20, None, 0.05, args.max_iter, 1e-6, False, 1e-8,
                     1e-8, True, False, "fp32", 10)
        gd, _keep = _graddev_struct(field, shape, N)
        gd.field = d_field.data_ptr()
        tuned = None
        if args.tune:
            n = L.lib().shorefit_graddev_variants(order, -1, None, None, None, None)
            ms = np.zeros(n)
            best = L.lib().shorefit_graddev_tune(C.byref(grads), C.byref(vol), C.byref(gd), C.byref(spec), 16384,
                                                 L.ptr(ms, L._pd), None)
            if best < 0:
                L.check(-best)
            tuned = {"variant": best, "ms_16384_voxels": [round(x, 3) for x in ms]}

        def step():
            out = L.ShorefitOutput()
            out.coefficients = d_coef.data_ptr()
            out.iterations = d_iter.data_ptr()
            out.residual_l2 = d_res.data_ptr()
            L.check(L.lib().shorefit_fit_graddev(C.byref(grads), C.byref(vol), C.byref(gd), C.byref(spec),
                                                 C.byref(out), None))
            return out

        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        phases = np.zeros(3)
        for _ in range(args.steps):
            o = step()
            phases += np.array(o.graddev_ms[:])
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        phases /= args.steps
        its = d_iter.cpu().numpy()[wl.mask].astype(np.int64)
        per_it = 4.0 * N * K + N + 12.0 * K
        flops = float(its.sum()) * per_it
        solve_s = phases[1] * 1e-3
        variant = tuned["variant"] if tuned else 0
        vv = [C.c_int32() for _ in range(4)]
        L.lib().shorefit_graddev_variants(order, variant, *[C.byref(x) for x in vv])
        Kp, W_f = vv[0].value, vv[3].value
        result = {
            "metric": f"voxel fits/s (SHORE-{order} FISTA, per-voxel graddev designs)", "value": round(V * args.steps / wall, 1),
            "unit": "voxel-fits/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(wall / args.steps * 1e3, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "fp32 (FP64 design / Gram / residual)",
            "data": "synthetic (seeded multi-tensor + Rician phantom, SNR U(5,40), T = I + E)",
            "config": {"workload": f"shore{order}-graddev-cartesian258-96x96x60-ellipsoid", "voxels": V, "n_grads": N,
                       "n_coeffs": K, "radial_order": order, "solver": "fista-l1 fixed alpha=1e-3",
                       "max_iter": args.max_iter, "tol": 1e-6, "variant": {"index": variant, "kp": Kp,
                                                                          "rows": vv[1].value,
                                                                          "gram_waves": vv[2].value,
                                                                          "fista_waves": W_f}},
            "phase_ms": {"gram": round(phases[0], 3), "fista": round(phases[1], 3), "residual": round(phases[2], 3)},
            "iterations": {"mean": float(its.mean()), "min": int(its.min()), "p50": float(np.median(its)),
                           "p99": float(np.quantile(its, 0.99)), "max": int(its.max()),
                           "at_cap": float(np.mean(its >= args.max_iter))},
            "roofline": {"bound": "fp32 vector (register-resident Gram, VALU)", "kernel": f"gd_fista_kernel<{Kp},{W_f}>",
                         "achieved": round(flops / solve_s / 1e12, 3), "peak": PEAK["fp32"], "unit": "TFLOP/s",
                         "frac": round(flops / solve_s / 1e12 / PEAK["fp32"], 4),
                         "executed": {"flops_per_launch": float(its.sum()) * (2.0 * Kp * Kp + 10.0 * Kp),
                                      "tflops": round(float(its.sum()) * (2.0 * Kp * Kp + 10.0 * Kp) / solve_s / 1e12,
                                                      3)},
                         "gram": {"fp64_mfma_flops_per_voxel": 2.0 * N * Kp * (Kp + 16) / 2.0,
                                  "tflops": round(V * 2.0 * N * Kp * (Kp + 16) / 2.0 / (phases[0] * 1e-3) / 1e12, 3),
                                  "peak": 78.6}},
        }
        if tuned:
            result["autotune"] = tuned
        if not args.no_cpu:
            from oracle import load
            orc = load()
            sub = wl.data.reshape(-1, N)[flat_idx].reshape(len(pick), 1, 1, N)
            fs = field.reshape(-1, 3, 3)[flat_idx].reshape(len(pick), 1, 1, 3, 3)
            threads, _, host = cpu_threads()
            t0 = time.perf_counter()
            ref = orc.fit_graddev(sub, None, fs, g.bvals, g.bvecs, big_delta=g.big_delta, small_delta=g.small_delta,
                                  radial_order=order, max_iter=args.max_iter, tol=1e-6, n_threads=threads)
            cpu_wall = time.perf_counter() - t0
            cr = ref["coefficients"].reshape(len(pick), K).astype(np.float64)
            cg = d_coef.cpu().numpy()[wl.mask][pick].astype(np.float64)
            err = np.linalg.norm(cg - cr, axis=1) / np.maximum(np.linalg.norm(cr, axis=1), 1e-30)
            result["cpu_baseline"] = {"value": round(len(pick) / cpu_wall, 2), "unit": "voxel-fits/s",
                                      "cores": threads, "host_cores": host, "kind": "port",
                                      "sample": f"{len(pick)} voxels evenly spaced over the mask, same fit settings "
                                                f"(per-voxel design, Jacobi Lipschitz), FP64 oracle (C++), "
                                                f"{cpu_wall:.1f} s wall"}
            result["accuracy_vs_fp64"] = {"sample": len(pick), "coef_rel_err_median": float(np.median(err)),
                                          "coef_rel_err_p99": float(np.quantile(err, 0.99)),
                                          "coef_rel_err_max": float(err.max()),
                                          "frac_le_1e-4": float(np.mean(err <= 1e-4)),
                                          "iterations_equal_frac": float(np.mean(ref["iterations"].ravel() ==
                                                                                 its[pick]))}
        print(json.dumps(result), flush=True)
        del d_coef


def adjoint_bench(args):
    """Differentiation workload: 65,536 voxels, order 8, 10% of the
    diffusion-weighted q-points held out (seed 6, 232 training rows). A step
    is one FP32 FISTA fit of every voxel on the training rows (alpha = 1e-3)
    plus one adjoint pass for the held-out-MSE gradient (d/ds and d/dalpha).
    Also reported: the adjoint against the FP64 oracle and against finite
    differences of KKT-verified FP64 minimisers, and the per-voxel alpha
    found by hypergradient descent against the 20-point grid search."""
    import torch

    from cs_dmri_amd import design_matrix
    from cs_dmri_amd import workloads as W
    from cs_dmri_amd.autograd import LassoDesign
    from cs_dmri_amd.hypergrad import grid_search, heldout_mse, hypergradient_search

    dev = torch.device("cuda:0")
    V = 65536
    gt, sig, tr, te = W.config5(V)
    A = design_matrix(gt.bvals, gt.bvecs, radial_order=8, big_delta=gt.big_delta, small_delta=gt.small_delta)
    A_tr, A_te = np.ascontiguousarray(A[tr]), np.ascontiguousarray(A[te])
    K = A.shape[1]
    d = LassoDesign(A_tr)
    S = torch.from_numpy(np.ascontiguousarray(sig[:, tr], np.float32)).to(dev)
    St = torch.from_numpy(np.ascontiguousarray(sig[:, te], np.float32)).to(dev)
    At = torch.from_numpy(A_te.astype(np.float32)).to(dev)
    nte = len(te)

    def step():
        coef, its = d.forward(S, 1e-3, max_iter=args.max_iter, tol=1e-6)
        e1.record()
        g = (2.0 / nte) * (coef @ At.T - St) @ At  # d(held-out MSE)/d(coef)
        e2.record()
        gs, ga, st = d.adjoint(coef, g)
        return coef, its, g, gs, ga, st

    e0, e1, e2, e3 = (torch.cuda.Event(enable_timing=True) for _ in range(4))
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t_fwd, t_adj = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        e0.record()
        coef, its, g, gs, ga, st = step()
        e3.record()
        torch.cuda.synchronize()
        t_fwd.append(e0.elapsed_time(e1))
        t_adj.append(e2.elapsed_time(e3))
    wall = (time.perf_counter() - t0) / args.steps
    its_np = its.cpu().numpy()
    fwd_ms, adj_ms = float(np.mean(t_fwd)), float(np.mean(t_adj))
    nnz = (coef != 0).sum(dim=1).cpu().numpy()
    # adjoint work: Cholesky m^3/3 + two solves 2 m^2 + A_S w 2 N m (FP64)
    adj_flops = float(np.sum(nnz.astype(np.float64) ** 3 / 3 + 2.0 * nnz ** 2 + 2.0 * len(tr) * nnz))
    result = {
        "metric": "voxel fit+adjoint/s (SHORE-8 FISTA-L1, implicit differentiation on the active set)",
        "value": round(V / wall, 1), "unit": "voxel-gradients/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32 fit, fp64 adjoint", "data": "synthetic (seeded multi-tensor + Rician, SNR 20)",
        "config": {"workload": "shore8-cartesian258-65536-heldout10", "voxels": V, "n_train": len(tr), "n_heldout": nte,
                   "n_coeffs": K, "alpha": 1e-3, "max_iter": args.max_iter, "tol": 1e-6},
        "phase_ms": {"forward": round(fwd_ms, 3), "adjoint": round(adj_ms, 3)},
        "active_set": {"mean": float(nnz.mean()), "max": int(nnz.max())},
        "iterations": {"mean": float(its_np.mean()), "max": int(its_np.max())},
        "adjoint_roofline": {"bound": "fp64 (Cholesky in registers for |S| <= 64, in LDS above; latency-bound)", "flops": adj_flops,
                             "tflops": round(adj_flops / (adj_ms * 1e-3) / 1e12, 3), "peak": 78.6},
        "status_not_pd": int((st != 0).sum()),
    }
    # --- accuracy: FP64 oracle adjoint and finite differences ------------------
    if not args.no_cpu:
        from oracle import load
        orc = load()
        pick = np.linspace(0, V - 1, 256).astype(np.int64)
        c_np = coef[pick].cpu().numpy().astype(np.float64)
        g_np = g[pick].cpu().numpy().astype(np.float64)
        t1 = time.perf_counter()
        rs, ra = orc.lasso_adjoint(A_tr, c_np, g_np)
        cpu_s = time.perf_counter() - t1
        es = np.linalg.norm(gs[pick].cpu().numpy() - rs, axis=1) / np.maximum(np.linalg.norm(rs, axis=1), 1e-30)
        ea = np.abs(ga[pick].cpu().numpy() - ra) / np.maximum(np.abs(ra), 1e-30)
        rng = np.random.default_rng(7)
        fd_err, fd_skip = [], 0
        for v in pick[:64]:
            s_v = sig[v, tr].astype(np.float32).astype(np.float64)
            c, stp = orc.lasso_polish(A_tr, s_v, 1e-3, coef[v].cpu().numpy().astype(np.float64))
            if stp != 0:
                fd_skip += 1
                continue
            gg = rng.standard_normal(K).astype(np.float32)
            gsv, _, _ = d.adjoint(torch.from_numpy(c[None].astype(np.float32)).to(dev), torch.from_numpy(gg[None]).to(dev))
            dirn = rng.standard_normal(len(tr))
            h = 1e-4 * np.linalg.norm(s_v) / np.linalg.norm(dirn)
            cp, sp = orc.lasso_polish(A_tr, s_v + h * dirn, 1e-3, c)
            cm, sm = orc.lasso_polish(A_tr, s_v - h * dirn, 1e-3, c)
            if sp or sm:
                fd_skip += 1
                continue
            fd = (gg.astype(np.float64) @ cp - gg.astype(np.float64) @ cm) / (2 * h)
            an = float(gsv[0].cpu().numpy().astype(np.float64) @ dirn)
            fd_err.append(abs(an - fd) / max(abs(fd), 1e-30))
        result["accuracy"] = {
            "vs_fp64_adjoint": {"sample": len(pick), "grad_signal_rel_max": float(es.max()),
                                "grad_alpha_rel_max": float(ea.max())},
            "vs_finite_differences": {"h": "1e-4 relative", "checked": len(fd_err), "skipped": fd_skip,
                                      "rel_err_median": float(np.median(fd_err)) if fd_err else None,
                                      "rel_err_max": float(np.max(fd_err)) if fd_err else None}}
        # CPU baseline: the FP64 oracle's fit (16 threads) and adjoint of a
        # sample of the same voxels, per voxel-gradient
        ns = 4096
        cs = np.linspace(0, V - 1, ns).astype(np.int64)
        sub = np.ascontiguousarray(sig[cs][:, tr], np.float32).reshape(ns, 1, 1, len(tr))
        threads, _, host = cpu_threads()
        t1 = time.perf_counter()
        ref = orc.fit_volume(sub, None, A_tr, alpha_mode="fixed", alpha=1e-3, max_iter=args.max_iter, tol=1e-6,
                             n_threads=threads)
        rc = ref["coefficients"].reshape(ns, K).astype(np.float64)
        rg = (2.0 / nte) * (rc @ A_te.T - sig[cs][:, te]) @ A_te
        orc.lasso_adjoint(A_tr, rc, rg)
        cpu_s = time.perf_counter() - t1
        result["cpu_baseline"] = {"value": round(ns / cpu_s, 1), "unit": "voxel-gradients/s", "cores": threads,
                                  "host_cores": host, "kind": "port",
                                  "sample": f"{ns} voxels evenly spaced: FP64 oracle fit ({threads} threads) and "
                                            f"FP64 adjoint (1 thread) of the held-out MSE, {cpu_s:.1f} s"}
    # --- held-out alpha: hypergradient descent vs the 20-point grid -----------
    base, _ = d.forward(S, 1e-3, max_iter=args.max_iter, tol=1e-6)
    mse0 = float(heldout_mse(base, At, St).mean())
    grid = grid_search(d, S, At, St, np.logspace(-5, -1, 20), max_iter=args.max_iter)
    hg = hypergradient_search(d, S, At, St, alpha0=1e-3, steps=30, max_iter=args.max_iter)
    result["alpha_selection"] = {
        "heldout_mse_alpha0": mse0,
        "grid20": {"mean_mse": float(grid["mse"].mean()), "seconds": round(grid["seconds"], 3), "fits": 20},
        "hypergradient": {"mean_mse_best": float(hg["best_mse"].mean()), "mean_mse_final": float(hg["mse"].mean()),
                          "seconds": round(hg["seconds"], 3), "steps": 30,
                          "history_first_last": [hg["history"][0], hg["history"][-1]]},
        "alpha_median": {"grid": float(grid["alpha"].median()), "hypergradient": float(hg["best_alpha"].median())}}
    print(json.dumps(result), flush=True)


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU; multi-GPU runs are not supported "
              "(the voxel loop would shard by voxel ranges with no communication).", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.precision_report:
        precision_report()
        return
    if args.workload == "tikhonov":
        tikhonov_bench(args)
        return
    if args.workload == "io":
        io_bench(args)
        return
    if args.workload == "graddev":
        graddev_bench(args)
        return
    if args.workload == "adjoint":
        adjoint_bench(args)
        return

    import torch

    from cs_dmri_amd import _lib as L
    from cs_dmri_amd import workloads as W

    dev = torch.device("cuda:0")
    wl = W.config2()
    g = wl.gtab
    K = L.lib().shorefit_n_coeffs(wl.radial_order)
    N = len(g.bvals)
    V = wl.n_voxels
    shape = wl.mask.shape

    d_data = torch.from_numpy(wl.data).to(dev)
    d_mask = torch.from_numpy(wl.mask.astype(np.uint8)).to(dev)
    d_coef = torch.empty(shape + (K,), dtype=torch.float32, device=dev)
    d_iter = torch.empty(shape, dtype=torch.int32, device=dev)
    d_res = torch.empty(shape, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()

    from cs_dmri_amd.models import _spec
    spec = _spec(wl.radial_order, wl.zeta, "l1", args.alpha_mode, 1e-3, 1e-3, 20, None, 0.05, args.max_iter, 1e-6,
                 False, 1e-8, 1e-8, True, False, args.precision, args.refresh)
    grads, keep = L.gradients(g.bvals, g.bvecs, g.big_delta, g.small_delta, g.b0_threshold)
    vol = L.ShorefitVolume(d_data.data_ptr(), d_mask.data_ptr(), shape[0], shape[1], shape[2], 1, 0)

    def step():
        out = L.ShorefitOutput()
        out.coefficients = d_coef.data_ptr()
        out.iterations = d_iter.data_ptr()
        out.residual_l2 = d_res.data_ptr()
        L.check(L.lib().shorefit_fit(C.byref(grads), C.byref(vol), C.byref(spec), C.byref(out), None))
        return out

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    kernel_ms = []
    for _ in range(args.steps):
        o = step()
        kernel_ms.append(o.fit_ms)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall / args.steps * 1e3
    value = V * args.steps / wall

    # --- roofline of the dominant kernel (the FISTA solve kernel) -------------
    # Algorithmic work per voxel-iteration is the reference algorithm's
    # (src/solver/fista.rs:89-107, residual form): A y and A^T r, 4 N K
    # MAC-flops, plus N + 12 K element-wise flops. The kernel executes the
    # equivalent Gram form (2 Kp^2 per increment, 3 x 2 Kp^2 per split-FP16
    # refresh, FP32 MFMA for fp32), reported beside it as "executed".
    its = d_iter.cpu().numpy()[wl.mask].astype(np.int64)
    its_sum = float(its.sum())
    per_it = 4.0 * N * K + N + 12.0 * K
    flops = its_sum * per_it
    kern_s = float(np.mean(kernel_ms)) * 1e-3
    achieved = flops / kern_s / 1e12
    peak = PEAK[args.precision]
    Kp = d_Kp(K, args.precision)
    if args.precision == "fp32":
        executed = its_sum * 2.0 * Kp * Kp
    else:
        # increments on FP16/BF16 MFMA, two passes (both halves of the split
        # operator) at padded K <= 160; refresh every k iterations: one FP32
        # MFMA pass, or three FP16 passes (fp16 at padded K > 160)
        k = refresh_used(args, Kp)
        split = Kp <= 160
        refresh_passes = 3.0 if (args.precision == "fp16" and not split) else 1.0
        executed = its_sum * 2.0 * Kp * Kp * ((2.0 if split else 1.0) * (k - 1) / k + refresh_passes / k)
    roofline = {"bound": "mfma", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
                "frac": round(achieved / peak, 4), "traffic": traffic_bytes(args.precision),
                "kernel": KERNEL[args.precision], "kernel_ms": round(kern_s * 1e3, 3),
                "flops_per_launch": flops, "flops_per_voxel_iteration": per_it,
                "executed": {"mfma_flops_per_launch": executed,
                             "tflops": round(executed / kern_s / 1e12, 3),
                             "frac": round(executed / kern_s / 1e12 / peak, 4)}}

    postfit = postfit_timings(L, d_coef, d_mask, wl, K, N, dev) if not args.no_postfit else None
    coef = d_coef.cpu().numpy()[wl.mask]
    result = {
        "metric": METRIC, "value": round(value, 1), "unit": "voxel-fits/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": args.precision,
        "data": "synthetic (seeded multi-tensor + Rician phantom)",
        "config": {"workload": "shore8-cartesian258-96x96x60-ellipsoid", "voxels": V, "n_grads": N,
                   "n_coeffs": K, "radial_order": 8, "solver": ("fista-l1 fixed alpha=1e-3" if args.alpha_mode == "fixed"
                                                             else f"fista-l1 {args.alpha_mode}"),
                   "max_iter": args.max_iter, "tol": 1e-6, "precision": args.precision,
                   "refresh_every": refresh_used(args, Kp) if args.precision != "fp32" else None},
        "iterations": {"mean": float(its.mean()), "p50": float(np.median(its)), "p99": float(np.quantile(its, 0.99)),
                       "max": int(its.max()), "at_cap": float(np.mean(its >= args.max_iter))},
        "roofline": roofline,
    }

    # --- CPU baseline leg and accuracy: the FP64 oracle on bounded samples -----
    if not args.no_cpu and args.alpha_mode == "fixed":
        from oracle import load

        from cs_dmri_amd import fista_batch
        orc = load()
        A = orc.design_matrix(wl.radial_order, wl.zeta, g.bvals, g.bvecs, g.big_delta, g.small_delta)
        S = wl.data[wl.mask]
        fit_kw = dict(alpha_mode="fixed", alpha=1e-3, max_iter=args.max_iter, tol=1e-6)
        threads, avail, host = cpu_threads()
        pick = np.linspace(0, V - 1, args.cpu_sample).astype(np.int64)
        rate, cpu_wall, ref = cpu_fit_rate(orc, S[pick], A, N, K, threads=threads, **fit_kw)
        one = pick[:: max(1, len(pick) // args.cpu_single_sample)][: args.cpu_single_sample]
        rate1, wall1, _ = cpu_fit_rate(orc, S[one], A, N, K, threads=1, **fit_kw)
        result["cpu_baseline"] = {
            "value": round(rate, 2), "unit": "voxel-fits/s", "cores": threads, "threads": threads,
            "host_cores": host, "cpus_available": avail, "single_thread_value": round(rate1, 2), "kind": "port",
            "sample": f"{len(pick)} voxels evenly spaced over the mask, same fit settings, FP64 oracle (C++), one "
                      f"voxel per task on {threads} threads, {cpu_wall:.1f} s wall; single thread: {len(one)} of "
                      f"those voxels, {wall1:.1f} s"}
        cr = ref["coefficients"].reshape(len(pick), K)
        itr = ref["iterations"].ravel()
        acc = {"stopping_rule": err_stats(coef[pick], cr, its[pick], itr)}
        # the FP32 form on the same voxels: the yardstick for the mixed forms
        if args.precision != "fp32":
            c32, d32 = fista_batch(A, S[pick], 1e-3, max_iter=args.max_iter, tol=1e-6, precision="fp32")
            acc["stopping_rule_fp32_form"] = err_stats(c32, cr, d32["iterations"], itr)
        # trajectory parity: equal iteration counts (tol = 0)
        tp = pick[:: max(1, len(pick) // args.equal_iter_sample)][: args.equal_iter_sample]
        ct, _ = fista_batch(A, S[tp], 1e-3, max_iter=args.max_iter, tol=0.0, precision=args.precision,
                            refresh_every=args.refresh)
        _, _, rt = cpu_fit_rate(orc, S[tp], A, N, K, threads=threads, alpha_mode="fixed", alpha=1e-3,
                                max_iter=args.max_iter, tol=0.0)
        eq = err_stats(ct, rt["coefficients"].reshape(len(tp), K))
        eq["iterations"] = args.max_iter
        acc["equal_iterations"] = eq
        result["accuracy_vs_fp64"] = acc
    if postfit is not None:
        result["postfit"] = postfit
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
