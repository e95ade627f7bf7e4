#!/usr/bin/env python3
"""Benchmark of the MI355X Foldy-Lax path: log-det evaluations per second at N=1000.

Workload (one "step"): the trace-log tr log M(k) / N of the multiple-scattering
matrix for d=3, N=1000 scatterers in the reference's unit-density ball (seed 1,
the medium msmain builds), hard spheres alpha=0.05, at the 64 x 64 = 4096 wave
numbers of a complex map over Re k in [5, 30] x Im k in [-2, 0] (map order of
get_arg_cmap).  Positions and the k grid are uploaded once per call (24 KB +
64 KB); every matrix is assembled, factored and reduced on the device.

Paths timed on the same map and GPU:
  fp64 batched   complex128 assembly + blocked LU on MFMA, all cells batched
  mixed          complex64 LU on MFMA for the whole map, complex128 re-evaluation
                 of cells with |det| < 1e-6 max|det| (the near-zero cells)
  fp64 plain     one matrix per call in complex128 (the plain FP64 GPU path)
value is the throughput of the fastest batched path whose error against the
complex128 map meets |d tracelog| <= 1e-8 (value_path names it).
Also: p50/p99 latency of a stream of resonance requests, the roofline of the
dominant kernel from HIP events, and the reference msmain (gcc -O2, MKL) on a
bounded sample of the same workload as CPU baseline.
"""
import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_MFMA = 157.3e12
PEAK_FP64_MFMA = 78.6e12
PEAK_HBM = 8.0e12


def kgrid(nx, ny, xmin, xmax, ymin, ymax):
    """Row-major from the top-left corner, as get_arg_cmap (complex_map_util.c:27-31)."""
    dx, dy = (xmax - xmin) / (nx - 1), (ymax - ymin) / (ny - 1)
    l = np.arange(nx * ny)
    return (xmin + (l % nx) * dx) + 1j * (ymax - (l // nx) * dy)


def cpu_baseline(n, nx, ny):
    """Times the reference msmain (oracle/_ref, built from the MSModel sources with MKL) on a k-map sample."""
    exe = os.path.join(ROOT, "oracle", "_ref", "msmain")
    if not os.path.exists(exe):
        return None
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    tmp = tempfile.mkdtemp(prefix="msbench_")
    try:
        conf = os.path.join(tmp, "map.conf")
        with open(conf, "w") as f:
            f.write("[medium]\ndimension=3\nnatom=%d\nmodel=hardsphere 0.05\nshape=ball\nratio=1\n\n"
                    "[kplane]\nnseed=1\nxrange=5:30\nyrange=-2:0\nxsample=%d\nysample=%d\n"
                    "color=temperature div 5\ntitle=bench\n" % (n, nx, ny))
        env = dict(os.environ, MKL_THREADING_LAYER="GNU", MKL_NUM_THREADS="1", OMP_NUM_THREADS=str(cores))
        subprocess.run([exe, "-p", str(cores), conf], cwd=tmp, env=env, check=True, stdout=subprocess.DEVNULL,
                       stderr=subprocess.DEVNULL, timeout=300)
        dat = None
        for dp, _, files in os.walk(os.path.join(tmp, "out")):
            for fn in files:
                if fn.endswith(".dat"):
                    dat = os.path.join(dp, fn)
        rt = None
        for line in open(dat):
            if line.startswith("realtime="):
                rt = float(line.split("=")[1])
        evals = nx * ny
        return {"value": evals / rt, "unit": "log-det evals/s", "cores": cores, "kind": "reference",
                "sample": "msmain -p %d (gcc -O2, MKL sequential per thread): d=3 N=%d unit-density ball seed 1, "
                          "hardsphere 0.05, %dx%d k-map over Re[5,30] x Im[-2,0] = %d log-dets in %.2f s"
                          % (cores, n, nx, ny, evals, rt)}
    except Exception as ex:  # the baseline is reported, never fatal to the GPU measurement
        return {"value": None, "unit": "log-det evals/s", "cores": cores, "kind": "reference",
                "sample": "failed: %s" % ex}
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def resonance_latencies(G, nreq):
    """A stream of resonance requests: each is an Aberth search with 8 starts on its own medium
    (d=2, N in [100, 400] log-uniform, unit-density disk, alpha=0.1), mixed precision."""
    rng = np.random.default_rng(11)
    lat = []
    for r in range(nreq):
        n = int(round(np.exp(rng.uniform(np.log(100), np.log(400)))))
        med = G.Medium(2, n, G.HARDSPHERE, (0.1,)).fill(100 + r)
        dom = (1.0, 12.0, -1.0, 0.0)
        starts = dom[0] + (dom[1] - dom[0]) * rng.uniform(0.05, 0.95, 8) + \
            1j * (dom[2] + (dom[3] - dom[2]) * rng.uniform(0.05, 0.95, 8))
        t0 = time.perf_counter()
        G.resonances(med, starts, dom, 4, 40, 1e-10, G.MIXED)
        lat.append(time.perf_counter() - t0)
    lat = np.array(lat) * 1e3
    return float(np.percentile(lat, 50)), float(np.percentile(lat, 99))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1000)
    ap.add_argument("--grid", type=int, default=64)
    ap.add_argument("--plain", type=int, default=24, help="cells timed on the one-matrix-per-call FP64 path")
    ap.add_argument("--requests", type=int, default=12, help="resonance requests for p50/p99")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run this many device copies of 1 GiB (calibrates the HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        sys.exit("bench.py: this build runs on one GPU; --gpus %d is not supported (the k points of a map "
                 "shard across GPUs without exchange, see DESIGN.md)" % args.gpus)

    if args.copy_probe:
        import torch
        src = torch.empty(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        src.fill_(1.0)
        for _ in range(args.copy_probe):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": src.numel() * 8, "copies": args.copy_probe}))
        return

    import msmodel_amd as G
    G.init()
    n, nk = args.n, args.grid * args.grid
    med = G.Medium(3, n, G.HARDSPHERE, (0.05,)).fill(1)
    ks = kgrid(args.grid, args.grid, 5.0, 30.0, -2.0, 0.0)

    def timed_path(mode):
        for _ in range(args.warmup):
            G.tracelog_batch(med, ks, mode, 1e-6)
        G.profile_begin()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            out, flags = G.tracelog_batch(med, ks, mode, 1e-6)
        dt = time.perf_counter() - t0
        return out, flags, dt, G.profile_end()

    out64, _, t64, prof64 = timed_path(G.FP64)
    fp64_batched = args.steps * nk / t64
    out_mixed, flags, t_mixed, prof32 = timed_path(G.MIXED)
    mixed_rate = args.steps * nk / t_mixed
    # plain FP64 GPU path: one matrix per call
    sub = ks[:: max(1, nk // args.plain)][: args.plain]
    G.tracelog_batch(med, sub[:1], G.FP64)
    t0 = time.perf_counter()
    for k in sub:
        G.tracelog_batch(med, np.array([k]), G.FP64)
    fp64_plain = len(sub) / (time.perf_counter() - t0)
    # arg det = N Im(tracelog) is defined modulo 2 pi: compare Im wrapped
    dre = np.abs(out_mixed.real - out64.real)
    dim = np.abs((n * (out_mixed.imag - out64.imag) + np.pi) % (2 * np.pi) - np.pi) / n
    dtl = np.hypot(dre, dim)
    dtl_re = float(dre.max())

    p50, p99 = resonance_latencies(G, args.requests) if args.requests > 0 else (None, None)

    mixed_ok = float(dtl.max()) <= 1e-8
    if mixed_ok and mixed_rate > fp64_batched:
        value, value_path, prof, peak, kname = mixed_rate, "mixed", prof32, PEAK_FP32_MFMA, "gemm_sub_kernel<float>"
        ms_step = 1e3 * t_mixed / args.steps
    else:
        value, value_path, prof, peak, kname = fp64_batched, "fp64_batched", prof64, PEAK_FP64_MFMA, \
            "gemm_sub_kernel<double>"
        ms_step = 1e3 * t64 / args.steps
    g = prof["gemm"]
    achieved = g["flops"] / (g["ms"] * 1e-3) if g["ms"] > 0 else 0.0
    roofline = {"bound": "mfma", "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": None,
                "kernel": kname + " (LU trailing update, MFMA 16x16x4)",
                "launches": g["launches"], "avg_ms": g["ms"] / max(1, g["launches"])}
    g32 = prof32["gemm"]
    roofline_c64 = {"achieved": g32["flops"] / max(1e-12, g32["ms"] * 1e-3) / 1e12, "peak": PEAK_FP32_MFMA / 1e12,
                    "frac": g32["flops"] / max(1e-12, g32["ms"] * 1e-3) / PEAK_FP32_MFMA}
    breakdown = {c: round(v["ms"] / max(1e-9, sum(x["ms"] for x in prof.values())), 4) for c, v in prof.items()}

    cpu = None if args.no_cpu_baseline else cpu_baseline(n, 16, 8)
    line = {
        "metric": "log-det evals/s at N=1000 (mixed vs FP64, same GPU); p50/p99 resonance latency",
        "value": value, "unit": "log-det evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c128" if value_path == "fp64_batched" else "c64+c128", "data": "synthetic",
        "config": {"workload": "tracelog map d=3 N=%d unit-density ball seed 1 hardsphere 0.05, %dx%d k over "
                               "Re[5,30] x Im[-2,0]" % (n, args.grid, args.grid),
                   "cells_per_step": nk},
        "value_path": value_path,
        "fp64_batched_evals_per_s": fp64_batched, "mixed_evals_per_s": mixed_rate,
        "fp64_plain_evals_per_s": fp64_plain, "value_over_fp64_plain": value / fp64_plain,
        "mixed_reevaluated_cells": int(flags.sum()), "mixed_reeval_rel": 1e-6,
        "mixed_max_abs_dtracelog": float(dtl.max()), "mixed_median_abs_dtracelog": float(np.median(dtl)),
        "mixed_max_abs_dtracelog_re": dtl_re, "roofline_c64_gemm": roofline_c64,
        "resonance_latency_ms": {"p50": p50, "p99": p99, "requests": args.requests},
        "kernel_time_fraction": breakdown,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
