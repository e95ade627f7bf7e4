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
  fp64 ldlt      complex128, pivot-free LDL^T of the symmetric M (half the update
                 flops, no pivot search) on the cells whose growth is predicted
                 small, the blocked LU on the others and on any cell whose measured
                 growth max |L| exceeds 1e3
  mixed          complex64 LU on MFMA for the whole map, complex128 re-evaluation
                 of cells with |det| < 1e-6 max|det| (the near-zero cells)
  fp64 plain     one matrix per call in complex128 (the plain FP64 GPU path)
value is the throughput of the fastest batched path whose error against the
complex128 LU map meets |d log det| <= 1e-8 in every cell (log det = N tracelog;
value_path names it).
Also reported: a Poisson stream of mixed requests (log-det tiles, solves with a psi
grid, resonance searches) through the serving loop with p50/p99 latency, a
resonance search over 64 media of uneven N in one call, the N=1000 solve and
1024^2 psi grid, the roofline of the dominant kernel from HIP events (HBM traffic
from the recorded counter runs), and the reference msmain (gcc -O2, MKL) on a
bounded sample of the same map as CPU baseline.
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


def counter_record(kernel, path):
    """The counter-run record of a kernel in the serialized pass of a map path at this
    workload (profiles/traffic.json, from the runs of profiles/pmc_map_n1000.md): HBM bytes
    per launch read and written (FETCH_SIZE and WRITE_SIZE scaled by a 1 GiB copy) and the
    average duration of the kernel alone on the GPU; None if not recorded."""
    try:
        rec = json.load(open(os.path.join(ROOT, "profiles", "traffic.json")))
        return rec["paths"][path]["kernels"]["msg::" + kernel]
    except (OSError, KeyError, ValueError):
        return None


def measured_traffic(kernel, path):
    rec = counter_record(kernel, path)
    return None if rec is None else rec["read_bytes_per_launch"] + rec["write_bytes_per_launch"]


def run_msmain(exe, n, nx, ny, threads, extra_env=None):
    """Runs an msmain build on a d=3 unit-density k-map config; returns its realtime (s)."""
    tmp = tempfile.mkdtemp(prefix="msbench_")
    try:
        conf = os.path.join(tmp, "map.conf")
        with open(conf, "w") as f:
            f.write("[medium]\ndimension=3\nnatom=%d\nmodel=hardsphere 0.05\nshape=ball\nratio=1\n\n"
                    "[kplane]\nnseed=1\nxrange=5:30\nyrange=-2:0\nxsample=%d\nysample=%d\n"
                    "color=temperature div 5\ntitle=bench\n" % (n, nx, ny))
        env = dict(os.environ, MKL_THREADING_LAYER="GNU", MKL_NUM_THREADS="1", OMP_NUM_THREADS=str(threads))
        env.update(extra_env or {})
        subprocess.run([exe, "-p", str(threads), conf], cwd=tmp, env=env, check=True, stdout=subprocess.DEVNULL,
                       stderr=subprocess.DEVNULL, timeout=300)
        for dp, _, files in os.walk(os.path.join(tmp, "out")):
            for fn in files:
                if fn.endswith(".dat"):
                    for line in open(os.path.join(dp, fn)):
                        if line.startswith("realtime="):
                            return float(line.split("=")[1])
        raise RuntimeError("no realtime in the output of %s" % exe)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def host_cores():
    """Threads for the host legs: the CPU share the process may use (OMP_NUM_THREADS when
    set, as on a GPU box whose share of the host is fixed per GPU), else every core in
    the affinity mask."""
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    share = os.environ.get("OMP_NUM_THREADS", "")
    if share.isdigit() and int(share) > 0:
        cores = min(cores, int(share))
    return max(1, cores)


def host_description():
    """The host CPU as /proc/cpuinfo names it (first processor): model name, and the family,
    model and stepping numbers where the name is generic."""
    info = {}
    try:
        for line in open("/proc/cpuinfo"):
            if not line.strip():
                break
            key, _, val = line.partition(":")
            info[key.strip()] = val.strip()
    except OSError:
        pass
    model = info.get("model name", "unknown")
    ids = ", ".join("%s %s" % (k, info[k]) for k in ("cpu family", "model", "stepping") if k in info)
    return {"cpu_model": model + (" (%s)" % ids if ids else ""), "cpus_online": os.cpu_count(),
            "lapack": "Intel MKL via /opt/conda/lib/libmkl_rt.so.1 (sequential per thread)"}


def cpu_baseline(n, nx, ny):
    """Times the reference msmain (oracle/_ref, built from the MSModel sources with MKL) on a k-map sample."""
    exe = os.path.join(ROOT, "oracle", "_ref", "msmain")
    if not os.path.exists(exe):
        return None
    cores = host_cores()
    try:
        rt = run_msmain(exe, n, nx, ny, cores)
        evals = nx * ny
        line = {"value": evals / rt, "unit": "log-det evals/s", "cores": cores, "kind": "reference",
                "sample": "msmain -p %d (gcc -O2, MKL sequential per thread): d=3 N=%d unit-density ball seed 1, "
                          "hardsphere 0.05, %dx%d k-map over Re[5,30] x Im[-2,0] = %d log-dets in %.2f s"
                          % (cores, n, nx, ny, evals, rt)}
        line.update(host_description())
        return line
    except Exception as ex:  # the baseline is reported, never fatal to the GPU measurement
        return {"value": None, "unit": "log-det evals/s", "cores": cores, "kind": "reference",
                "sample": "failed: %s" % ex}


def dropin_msmain(n, nx, ny):
    """The drop-in path: the reference msmain relinked against libmsgpu.so (oracle/_ref/msmain_gpu;
    its matrix functions run on the GPU, concurrent tracelog_lu calls are coalesced), same map."""
    exe = os.path.join(ROOT, "oracle", "_ref", "msmain_gpu")
    if not os.path.exists(exe):
        return None
    threads = host_cores()
    try:
        rt = run_msmain(exe, n, nx, ny, threads)
        rk = run_msmain(exe, n, nx, ny, threads, {"MSGPU_TRACELOG_KEEP_INPUT": "1"})
        return {"evals_per_s": nx * ny / rt, "threads": threads,
                "evals_per_s_keep_input": nx * ny / rk,
                "sample": "msmain_gpu -p %d: d=3 N=%d unit-density ball seed 1, hardsphere 0.05, %dx%d k-map "
                          "= %d log-dets in %.2f s (%.2f s with MSGPU_TRACELOG_KEEP_INPUT=1: tracelog_lu leaves "
                          "the matrix as given)" % (threads, n, nx, ny, nx * ny, rt, rk)}
    except Exception as ex:
        return {"evals_per_s": None, "threads": threads, "sample": "failed: %s" % ex}


def run_msmain_conf(exe, conf_text, threads, section, extra_env=None):
    """Runs an msmain build on a configuration; returns the realtime (s) msmain writes for
    the given section of its output."""
    tmp = tempfile.mkdtemp(prefix="msbench_")
    try:
        conf = os.path.join(tmp, "t.conf")
        with open(conf, "w") as f:
            f.write(conf_text)
        env = dict(os.environ, MKL_THREADING_LAYER="GNU", MKL_NUM_THREADS="1", OMP_NUM_THREADS=str(threads))
        env.update(extra_env or {})
        subprocess.run([exe, "-p", str(threads), conf], cwd=tmp, env=env, check=True, stdout=subprocess.DEVNULL,
                       stderr=subprocess.DEVNULL, timeout=600)
        for dp, _, files in os.walk(os.path.join(tmp, "out")):
            for fn in files:
                if fn.endswith(".dat"):
                    cur = None
                    for line in open(os.path.join(dp, fn)):
                        line = line.strip()
                        if line.startswith("["):
                            cur = line
                        elif cur == "[%s]" % section and line.startswith("realtime="):
                            return float(line.split("=")[1])
        raise RuntimeError("no realtime for [%s] in the output of %s" % (section, exe))
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def dropin_wavefunction(n, grids):
    """msmain's [wavefunction] task (the pixel loop of run_wave_function_omp calling
    eval_total_wave once per pixel from every thread): the reference msmain (MKL, host
    threads) against msmain relinked to the library, whose calls the resident psi kernel
    serves (and, for comparison, with MSGPU_WAVE_RESIDENT=0: concurrent calls coalesced
    into one launch per batch); d=3, N unit-density ball, seed 1, hardsphere 0.1, plane
    wave k=10."""
    ref = os.path.join(ROOT, "oracle", "_ref", "msmain")
    gpu = os.path.join(ROOT, "oracle", "_ref", "msmain_gpu")
    if not (os.path.exists(ref) and os.path.exists(gpu)):
        return None
    threads = host_cores()
    res = {"threads": threads, "workload": "d=3 N=%d unit-density ball seed 1, hardsphere 0.1, plane wave k=10" % n}
    for g in grids:
        conf = ("[medium]\ndimension=3\nnatom=%d\nmodel=hardsphere 0.1\nshape=ball\nratio=1\n\n"
                "[wavefunction]\nnseed=1\niseed=1\nk=10.+0i\nsource=plane\nnx=%d\nny=%d\nxrange=-8:8\n"
                "yrange=-8:8\ncolor1=temperature div 5\ncolor2=temperature div 5\n" % (n, g, g))
        try:
            tr = run_msmain_conf(ref, conf, threads, "wavefunction")
            tg = run_msmain_conf(gpu, conf, threads, "wavefunction")
            tb = run_msmain_conf(gpu, conf, threads, "wavefunction", {"MSGPU_WAVE_RESIDENT": "0"})
            res["grid_%d" % g] = {"reference_s": tr, "dropin_s": tg, "dropin_over_reference": tr / tg,
                                  "dropin_pixels_per_s": g * g / tg, "dropin_batched_launches_s": tb}
        except Exception as ex:
            res["grid_%d" % g] = {"failed": str(ex)}
    return res


def serving_stream(G, nreq, rate, nmin, nmax, seed=5):
    """Poisson stream of mixed requests through the serving loop (msgpu_server_*).

    Each request draws d in {2, 3}, N log-uniform in [nmin, nmax], a medium seed
    (unit-density ball, hard spheres alpha=0.1) and one of: a 64-point log-det tile
    (8x8 over Re k in [1,10] x Im k in [-1,0], mixed), a Foldy-Lax solve at real k in
    [2,10] (mixed, refined to 1e-12) with a 256x256 psi grid over the medium, or a
    32-start resonance search in (1,10) x (-1,0) (mixed, toler 1e-10).  Latency is
    submit -> done as the server measures it."""
    import threading
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(nreq):
        d = int(rng.choice([2, 3]))
        n = int(round(np.exp(rng.uniform(np.log(nmin), np.log(nmax)))))
        med = G.Medium(d, n, G.HARDSPHERE, (0.1,)).fill(int(rng.integers(1, 1 << 30)))
        kind = int(rng.integers(0, 3))
        if kind == 0:
            arg = kgrid(8, 8, 1.0, 10.0, -1.0, 0.0)
        elif kind == 1:
            r = float(np.abs(med.pos).max()) + 0.5
            arg = (float(rng.uniform(2.0, 10.0)), (256, 256, -r, r, -r, r))
        else:
            arg = rng.uniform(1.2, 9.8, 32) + 1j * rng.uniform(-0.95, -0.05, 32)
        specs.append((kind, med, arg, float(rng.exponential(1.0 / rate))))
    srv = G.Server(hbm_gb=64.0, max_active=64)
    ids = [None] * nreq

    def submitter():
        for j, (kind, med, arg, gap) in enumerate(specs):
            time.sleep(gap)
            if kind == 0:
                ids[j] = srv.submit_logdet(med, arg, G.MIXED, 1e-6)
            elif kind == 1:
                ids[j] = srv.submit_solve(med, arg[0], G.PLANE, G.MIXED, 1e-12, arg[1])
            else:
                ids[j] = srv.submit_resonance(med, arg, (1.0, 10.0, -1.0, 0.0), 4, 30, 1e-10, G.MIXED)

    t0 = time.perf_counter()
    th = threading.Thread(target=submitter)
    th.start()
    th.join()
    lat = {0: [], 1: [], 2: []}
    nbig = {0: 0, 1: 0, 2: 0}
    for j, (kind, med, _, _) in enumerate(specs):
        lat[kind].append(srv.wait(ids[j])[1]["total_ms"])
        nbig[kind] += med.n >= 4000
        if nmax > 2000:
            progress("request %d/%d (kind %d, d=%d, N=%d) done after %.0f ms" % (j + 1, nreq, kind, med.c.d, med.n,
                                                                              lat[kind][-1]))
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    allv = np.array(sum(lat.values(), []))

    def pct(v):
        return {"p50": float(np.percentile(v, 50)), "p99": float(np.percentile(v, 99)), "n": len(v)} if v else None

    return {"requests": nreq, "arrival_rate_per_s": rate, "N_range": [nmin, nmax], "wall_s": wall,
            "throughput_req_per_s": nreq / wall, "latency_ms": pct(list(allv)),
            "by_kind_ms": {"logdet_tile64": pct(lat[0]), "solve_psi256": pct(lat[1]), "resonance32": pct(lat[2])},
            "requests_with_N_ge_4000_by_kind": {"logdet_tile64": nbig[0], "solve_psi256": nbig[1], "resonance32": nbig[2]},
            "steps": st["steps"], "batched_matrices": st["batched_matrices"],
            "max_requests_per_step": st["max_requests_per_step"], "resident_media_at_end": st["resident_media"],
            "kept_factor_hits": st["factor_hits"], "kept_factor_misses": st["factor_misses"]}


def kept_factor_solves(G, n=2000):
    """Complex128 solves + 256^2 psi through the server over one medium (d=3, N unit-density
    ball, seed 3, alpha=0.1) at one k = 8 - 0.05i: the first factors M(k) and keeps the
    factor on the device; the next ones (other sources and grids) solve with it."""
    med = G.Medium(3, n, G.HARDSPHERE, (0.1,)).fill(3)
    r = float(np.abs(med.pos).max()) + 0.5
    grids = [(256, 256, -r, r, -r, r), (256, 256, -r / 2, r / 2, -r / 2, r / 2)]
    srv = G.Server(hbm_gb=64.0)
    lat = []
    for j, (src, grid) in enumerate([(G.PLANE, grids[0]), (G.SPHERICAL, grids[0]), (G.PLANE, grids[1]),
                                     (G.SPHERICAL, grids[1])]):
        lat.append(srv.wait(srv.submit_solve(med, 8.0 - 0.05j, src, G.FP64, 0.0, grid))[1]["total_ms"])
    st = srv.stats()
    srv.close()
    return {"workload": "d=3 N=%d unit-density ball seed 3, alpha 0.1, k = 8 - 0.05i, complex128 solve + 256^2 "
                        "psi, sources plane/spherical, two grids" % n,
            "first_ms": lat[0], "with_kept_factor_ms": lat[1:], "factor_hits": st["factor_hits"],
            "factor_misses": st["factor_misses"], "kept_factor_bytes": st["kept_factor_bytes"]}


def uneven_media(G, nmed, total, nmin, nmax):
    """The uneven-batch resonance search's media and starts: d=2, nmed media (seeds 100..)
    with N log-uniform in [nmin, nmax] in unit disks, alpha=0.1, total starts uniform in
    Re k in [1,40] x Im k in [-1,0], as evenly spread over the media as they divide."""
    rng = np.random.default_rng(100)
    meds, starts, doms = [], [], []
    for i, s in enumerate(range(100, 100 + nmed)):
        n = int(round(np.exp(rng.uniform(np.log(nmin), np.log(nmax)))))
        per = total // nmed + (1 if i < total % nmed else 0)
        meds.append(G.Medium(2, n, G.HARDSPHERE, (0.1,)).fill(s, radius=1.0))
        starts.append(rng.uniform(1.0, 40.0, per) + 1j * rng.uniform(-1.0, 0.0, per))
        doms.append((1.0, 40.0, -1.0, 0.0))
    return meds, starts, doms


def uneven_search(G, nmed, total, nmin, nmax):
    """Resonance search over media of uneven size in one call (msgpu_resonances_multi):
    complex64 sweeps then complex128 polish to |dk| < 1e-12 |k|."""
    meds, starts, doms = uneven_media(G, nmed, total, nmin, nmax)
    t0 = time.perf_counter()
    roots, sweeps = G.resonances_multi(meds, starts, doms, 4, 40, 1e-12, G.MIXED)
    dt = time.perf_counter() - t0
    ns = np.array([m.n for m in meds])
    return {"media": nmed, "starts": int(sum(len(z) for z in starts)), "N_range": [nmin, nmax],
            "N_of_media_min_median_max": [int(ns.min()), int(np.median(ns)), int(ns.max())],
            "sweeps": sweeps, "wall_s": dt, "roots_found": int(sum(len(r) for r in roots)),
            "starts_per_s": sum(len(z) for z in starts) / dt}


def tune_entry(G, np_, prec):
    """The tuning-table entry for padded size np_ and precision prec (0 complex128, 1
    complex64) as a dict, or None when the table has none."""
    import tempfile
    keys = ("swap_variant", "chunk", "streams", "row_major", "panel_variant", "graph", "block128")
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "t.txt")
        G.tune_save(path)
        for line in open(path):
            f = line.split()
            if line.startswith("#") or len(f) < 9:
                continue
            if int(f[0]) == np_ and int(f[1]) == prec:
                return dict(zip(keys, map(int, f[2:9])))
    return None


def set_entry(G, np_, prec, e):
    import tempfile
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "t.txt")
        with open(path, "w") as f:
            f.write("%d %d %d %d %d %d %d %d %d\n" % (np_, prec, e["swap_variant"], e["chunk"], e["streams"],
                                                      e["row_major"], e["panel_variant"], e["graph"], e["block128"]))
        G.tune_load(path)


def small_map(G, steps, tune):
    """The reference's CPU-runnable map: d=3, N=200 in a ball of radius 1 (seed 1),
    alpha=0.1, 128x64 k over Re[1,20] x Im[-1,0], complex128; with the chunk pipeline
    launched eagerly and replayed from captured HIP graphs (tuned launch shapes)."""
    med = G.Medium(3, 200, G.HARDSPHERE, (0.1,)).fill(1, radius=1.0)
    ks = kgrid(128, 64, 1.0, 20.0, -1.0, 0.0)
    tuned = tune_entry(G, 256, 0)
    if tune or tuned is None:
        G.autotune(200, G.FP64, 1024)
        tuned = tune_entry(G, 256, 0)
    out = {}
    for graph in (0, 1):
        set_entry(G, 256, 0, dict(tuned, graph=graph))
        G.tracelog_batch(med, ks, G.FP64)
        t0 = time.perf_counter()
        for _ in range(steps):
            res, _ = G.tracelog_batch(med, ks, G.FP64)
        out["graph" if graph else "eager"] = steps * ks.size / (time.perf_counter() - t0)
        out["tracelog_%d" % graph] = res
    lu = out.pop("tracelog_0")
    same = bool(np.array_equal(lu, out.pop("tracelog_1")))
    set_entry(G, 256, 0, tuned)
    # the pivot-free LDL^T path on the same map (LU where its growth is large)
    G.tracelog_batch(med, ks, G.FP64_LDLT)
    t0 = time.perf_counter()
    for _ in range(steps):
        res, flags = G.tracelog_batch(med, ks, G.FP64_LDLT)
    ldlt_rate = steps * ks.size / (time.perf_counter() - t0)
    dlog = np.hypot(200 * (res.real - lu.real), (200 * (res.imag - lu.imag) + np.pi) % (2 * np.pi) - np.pi)
    return {"workload": "tracelog map d=3 N=200 ball R=1 seed 1 hardsphere 0.1, 128x64 k over Re[1,20] x Im[-1,0]",
            "evals_per_s_eager": out["eager"], "evals_per_s_graph": out["graph"], "graph_equals_eager": same,
            "evals_per_s_ldlt": ldlt_rate, "ldlt_lu_cells": int(flags.sum()),
            "ldlt_max_abs_dlogdet": float(dlog.max()), "launch_shapes": tuned}


def serialized_pass(G, med, ks, np_, mode):
    """One pass of the map in the given mode on a single stream (the tuning entry of np_
    with streams=1), every launch bracketed by HIP events on that stream: the kernels run
    one after the other, so each launch's event interval is its duration alone on the GPU."""
    tuned = tune_entry(G, np_, 0)
    if tuned is None:
        return None
    set_entry(G, np_, 0, dict(tuned, streams=1))
    try:
        G.tracelog_batch(med, ks, mode)
        G.profile_begin()
        t0 = time.perf_counter()
        G.tracelog_batch(med, ks, mode)
        wall = time.perf_counter() - t0
        prof = G.profile_end()
    finally:
        set_entry(G, np_, 0, tuned)
    return prof, wall


def progress(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def wavefunction_case(G):
    """Solve + psi grid: d=3, N=1000 in a ball of radius 5 (seed 2), alpha=0.1, plane
    wave at real k=10, psi on 1024x1024 in the (x0, x1) plane over [-6, 6]^2."""
    med = G.Medium(3, 1000, G.HARDSPHERE, (0.1,)).fill(2, radius=5.0)
    G.solve(med, 10.0, G.PLANE, G.FP64)
    t0 = time.perf_counter()
    a64, _ = G.solve(med, 10.0, G.PLANE, G.FP64)
    t_fp64 = time.perf_counter() - t0
    t0 = time.perf_counter()
    a32, steps = G.solve(med, 10.0, G.PLANE, G.MIXED, 1e-12)
    t_mixed = time.perf_counter() - t0
    G.wavefunction_grid(med, 10.0, a64, 64, 64, -6, 6, -6, 6)
    G.profile_begin()
    t0 = time.perf_counter()
    psi = G.wavefunction_grid(med, 10.0, a64, 1024, 1024, -6, 6, -6, 6)
    t_psi = time.perf_counter() - t0
    kern = G.profile_end()["psi"]
    evals = 1024 * 1024 * 1000
    return {"solve_fp64_ms": 1e3 * t_fp64, "solve_mixed_ms": 1e3 * t_mixed, "mixed_refinement_steps": steps,
            "max_rel_da_mixed_vs_fp64": float(np.abs(a32 - a64).max() / np.abs(a64).max()),
            "psi_grid_ms": 1e3 * t_psi, "green_evals_per_s": evals / t_psi,
            "psi_kernel_ms": kern["ms"], "green_evals_per_s_kernel": evals / max(1e-12, kern["ms"] * 1e-3),
            "psi_timing": "psi_grid_ms: the call (positions and amplitudes up, kernel, 16.8 MB of psi down into "
                          "host memory); psi_kernel_ms: HIP events around the kernel on its stream",
            "psi_finite": bool(np.isfinite(psi).all())}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1000)
    ap.add_argument("--grid", type=int, default=64)
    ap.add_argument("--plain", type=int, default=24, help="cells timed on the one-matrix-per-call FP64 path")
    ap.add_argument("--requests", type=int, default=48, help="requests of the serving stream (0: skip)")
    ap.add_argument("--rate", type=float, default=8.0, help="Poisson arrival rate of the serving stream (1/s)")
    ap.add_argument("--serve-nmax", type=int, default=2000, help="largest N of the serving stream")
    ap.add_argument("--uneven", type=int, default=64, help="media of the uneven resonance search (0: skip)")
    ap.add_argument("--uneven-starts", type=int, default=10000, help="starts of the uneven search (all media)")
    ap.add_argument("--no-wave", action="store_true", help="skip the solve + 1024^2 psi case")
    ap.add_argument("--tune", action="store_true", help="run the autotuner for N before timing")
    ap.add_argument("--tune-file", default="", help="with --tune: where to save the tuning table")
    ap.add_argument("--tune-sizes", default="200,500,2000", help="with --tune: further N to tune (both precisions)")
    ap.add_argument("--no-small", action="store_true", help="skip the N=200 map (eager vs HIP graph)")
    ap.add_argument("--no-mixed", action="store_true", help="time the FP64 path only (for profiler runs)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--roofline-path", default="fp64_ldlt", choices=("fp64_batched", "fp64_ldlt"),
                    help="the path whose serialized pass --roofline-only runs")
    ap.add_argument("--roofline-only", action="store_true",
                    help="only the serialized pass of the map (one stream), for the rocprof kernel-trace run")
    ap.add_argument("--no-dropin", action="store_true", help="skip timing msmain relinked against the library")
    ap.add_argument("--dropin-wave-grids", default="256",
                    help="grid sizes of the drop-in [wavefunction] timing (comma-separated)")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run this many device copies of 1 GiB (calibrates the HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        sys.exit("bench.py: measures one GPU; --gpus %d is not supported (the k points of a map "
                 "shard across GPUs without any exchange: each GPU would run this benchmark on its "
                 "share)" % args.gpus)

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
    # launch shapes (set-up, untimed): the tuning table loaded at initialisation (the one
    # shipped beside the library, or MSGPU_TUNE_FILE), or with --tune the autotuner run
    # here for this N in both precisions, saved to --tune-file when given
    if args.tune:
        # (batches of 2048 matrices up to N = 1024: enough chunks for the stream count to show)
        items = lambda m: 2048 if m <= 1024 else 512
        tuned = {"fp64": G.autotune(n, G.FP64, items(n)), "fp32": G.autotune(n, G.FP32, items(n))}
        for m in sorted(set(int(x) for x in args.tune_sizes.split(",") if x) - {n}):
            for mode in (G.FP64, G.FP32):
                G.autotune(m, mode, items(m))
            progress("tuned N=%d" % m)
        if args.tune_file:
            G.tune_save(args.tune_file)
    else:
        tp = G.tuning_table_path()
        root = os.path.dirname(os.path.abspath(__file__))
        tuned = {"table": os.path.relpath(tp, root) if tp and os.path.abspath(tp).startswith(root + os.sep) else tp}
    progress("launch shapes", tuned)

    if args.roofline_only:
        np_ = (n + 63) // 64 * 64
        mode = {"fp64_batched": G.FP64, "fp64_ldlt": G.FP64_LDLT}[args.roofline_path]
        prof, wall = serialized_pass(G, med, ks, np_, mode)     # its warm-up pass is serialized too
        g = prof["gemm"]
        print(json.dumps({"serialized_pass_ms": 1e3 * wall, "gemm_launches": g["launches"],
                          "gemm_avg_ms": g["ms"] / max(1, g["launches"]),
                          "gemm_alg_bytes_per_launch": g["bytes"] / max(1, g["launches"]),
                          "gemm_flops_per_launch": g["flops"] / max(1, g["launches"]),
                          "gemm_tflops": g["flops"] / max(1e-12, g["ms"] * 1e-3) / 1e12,
                          "kernel_ms": {c: v["ms"] for c, v in prof.items() if v["launches"]},
                          "launches": {c: v["launches"] for c, v in prof.items() if v["launches"]}}))
        return

    def timed_path(mode, rel=1e-6):
        for _ in range(args.warmup):
            G.tracelog_batch(med, ks, mode, rel)
        G.profile_begin()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            out, flags = G.tracelog_batch(med, ks, mode, rel)
        dt = time.perf_counter() - t0
        return out, flags, dt, G.profile_end()

    out64, _, t64, prof64 = timed_path(G.FP64)
    progress("fp64 map: %.1f evals/s" % (args.steps * nk / t64))
    fp64_batched = args.steps * nk / t64
    outL, flL, tL, profL = timed_path(G.FP64_LDLT)
    ldlt_rate = args.steps * nk / tL
    progress("fp64 ldlt map: %.1f evals/s, %d cells through the LU" % (ldlt_rate, int(flL.sum())))
    if args.no_mixed:   # (profiler runs of the FP64 path alone)
        out_mixed, flags, t_mixed, prof32 = out64, np.zeros(nk, np.int32), t64, None
        mixed_rate = None
    else:
        out_mixed, flags, t_mixed, prof32 = timed_path(G.MIXED)
        mixed_rate = args.steps * nk / t_mixed
        progress("mixed map: %.1f evals/s" % mixed_rate)
    # plain FP64 GPU path: one matrix per call
    fp64_plain = None
    if args.plain > 0:
        sub = ks[:: max(1, nk // args.plain)][: args.plain]
        G.tracelog_batch(med, sub[:1], G.FP64)
        t0 = time.perf_counter()
        for k in sub:
            G.tracelog_batch(med, np.array([k]), G.FP64)
        fp64_plain = len(sub) / (time.perf_counter() - t0)
    # arg det = N Im(tracelog) is defined modulo 2 pi: compare Im wrapped
    def dtracelog(out):
        dre = np.abs(out.real - out64.real)
        dim = np.abs((n * (out.imag - out64.imag) + np.pi) % (2 * np.pi) - np.pi) / n
        return np.hypot(dre, dim), dre

    dtl, dre = dtracelog(out_mixed)
    dtl_re = float(dre.max())
    # the pivot-growth gate (complex64 factor; complex128 where 16 eps32 exp(max log|U_ii| -
    # min log|U_ii|) > tol, tol on |d tracelog|) at the log det target 1e-8 (tol = 1e-8 / N)
    # and at a looser tolerance
    gated = None
    if not args.no_mixed:
        gated = {}
        for tol in (1e-8 / n, 1e-5):
            og, fg, tg, _ = timed_path(G.MIXED_GATED, tol)
            e, _ = dtracelog(og)
            gated["tol_%g" % tol] = {
                "evals_per_s": args.steps * nk / tg, "recomputed_cells": int(fg.sum()), "cells": nk,
                "max_abs_dtracelog": float(e.max()), "max_abs_dlogdet": float(n * e.max()),
                "median_abs_dtracelog": float(np.median(e)),
                "p99_abs_dtracelog": float(np.percentile(e, 99)), "cells_within_tol": int((e <= tol).sum()),
                "max_abs_dtracelog_kept_cells": float(e[fg == 0].max()) if (fg == 0).any() else 0.0}
            progress("gated mixed map, tol %g: %.1f evals/s, %d cells recomputed" % (tol, args.steps * nk / tg,
                                                                                   int(fg.sum())))

    serving = uneven = wave = small = None
    if args.requests > 0:
        serving = serving_stream(G, args.requests, args.rate, 100, args.serve_nmax)
        progress("serving stream done")
    if args.uneven > 0:
        uneven = uneven_search(G, args.uneven, args.uneven_starts, 200, 2000)
        progress("uneven search: %d starts in %.1f s" % (uneven["starts"], uneven["wall_s"]))
        progress("uneven resonance search done")
    kept = None
    if not args.no_wave:
        wave = wavefunction_case(G)
        kept = kept_factor_solves(G)
        progress("kept-factor solves done")
        progress("solve + psi grid done")
    if not args.no_small:
        small = small_map(G, 3, args.tune)
        progress("N=200 map done")

    # the bound is on log det = N tracelog
    mixed_ok = n * float(dtl.max()) <= 1e-8
    dtl_ldlt, dre_ldlt = dtracelog(outL)
    ldlt_ok = n * float(dtl_ldlt.max()) <= 1e-8
    # the fastest batched path within 1e-8 in log det of the complex128 LU map
    cands = [(fp64_batched, "fp64_batched", prof64, PEAK_FP64_MFMA, "gemm_sub_kernel<double>", t64, G.FP64)]
    if ldlt_ok:
        cands.append((ldlt_rate, "fp64_ldlt", profL, PEAK_FP64_MFMA, "gemm_sub_kernel<double>", tL, G.FP64_LDLT))
    if mixed_rate is not None and mixed_ok:
        cands.append((mixed_rate, "mixed", prof32, PEAK_FP32_MFMA, "gemm_sub_kernel<float>", t_mixed, G.MIXED))
    value, value_path, prof, peak, kname, t_value, value_mode = max(cands, key=lambda c: c[0])
    ms_step = 1e3 * t_value / args.steps
    # The dominant kernel's roofline from a serialized pass (one stream: each launch alone on
    # the GPU, durations from HIP events on its stream), measured here; the same launches in
    # the timed region (four streams sharing the GPU, which stretches each launch's event
    # interval) are reported beside it as *_in_mix.
    g = prof["gemm"]
    ser = serialized_pass(G, med, ks, (n + 63) // 64 * 64, value_mode) if value_path != "mixed" else None
    gs = ser[0]["gemm"] if ser else g
    achieved = gs["flops"] / (gs["ms"] * 1e-3) if gs["ms"] > 0 else 0.0
    achieved_mix = g["flops"] / (g["ms"] * 1e-3) if g["ms"] > 0 else 0.0
    traffic = measured_traffic(kname, value_path)
    roofline = {"bound": "mfma", "achieved": achieved / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": traffic,
                "traffic_source": "profiles/traffic.json: counter runs of the serialized pass of this path "
                                  "(profiles/pmc_map_n1000.md; rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, per "
                                  "launch, scaled by a 1 GiB copy)" if traffic is not None else
                                  "no counter run of this path's launches is recorded",
                "kernel": kname + (" (trailing updates of the LDL^T (lower tiles) and of the LU, MFMA 16x16x4)"
                                   if value_path == "fp64_ldlt" else " (LU trailing update, MFMA 16x16x4)"),
                "path": value_path,
                "timing": "serialized pass: one stream, HIP events around every launch" if ser else
                          "timed region (four streams)",
                "launches": gs["launches"], "avg_ms": gs["ms"] / max(1, gs["launches"]),
                "flops_per_launch": gs["flops"] / max(1, gs["launches"]),
                "algorithmic_bytes": gs["bytes"] / max(1, gs["launches"]),
                "serialized_pass_evals_per_s": nk / ser[1] if ser else None,
                "achieved_in_mix": achieved_mix / 1e12, "frac_in_mix": achieved_mix / peak,
                "avg_ms_in_mix": g["ms"] / max(1, g["launches"]), "launches_in_mix": g["launches"]}
    prof = {c: v for c, v in prof.items() if v["launches"]}
    if ser:
        ser = ({c: v for c, v in ser[0].items() if v["launches"]}, ser[1])
        tot = sum(v["ms"] for v in ser[0].values())
        roofline["serialized_kernel_ms"] = {c: round(v["ms"], 3) for c, v in ser[0].items()}
        roofline["serialized_kernel_fraction"] = {c: round(v["ms"] / max(1e-9, tot), 4) for c, v in ser[0].items()}
    roofline_c64 = None
    if prof32 is not None:
        g32 = prof32["gemm"]
        roofline_c64 = {"achieved": g32["flops"] / max(1e-12, g32["ms"] * 1e-3) / 1e12,
                        "peak": PEAK_FP32_MFMA / 1e12, "frac": g32["flops"] / max(1e-12, g32["ms"] * 1e-3) / PEAK_FP32_MFMA,
                        "timing": "in the mix (four streams; each launch shares the GPU): compare with "
                                  "roofline.frac_in_mix, not roofline.frac"}
    breakdown = {c: round(v["ms"] / max(1e-9, sum(x["ms"] for x in prof.values())), 4) for c, v in prof.items()}

    cpu = None if args.no_cpu_baseline else cpu_baseline(n, 48, 48)
    if cpu:
        progress("cpu baseline done")
    dropin = None if args.no_dropin else dropin_msmain(n, 48, 48)
    if dropin:
        progress("drop-in msmain done")
    dropin_wave = None if args.no_dropin else dropin_wavefunction(n, [int(x) for x in args.dropin_wave_grids.split(",") if x])
    if dropin_wave:
        progress("drop-in wavefunction done")
    line = {
        "metric": "log-det evals/s at N=1000 (mixed vs FP64, same GPU); p50/p99 resonance latency",
        "value": value, "unit": "log-det evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c64+c128" if value_path == "mixed" else "c128", "data": "synthetic",
        "config": {"workload": "tracelog map d=3 N=%d unit-density ball seed 1 hardsphere 0.05, %dx%d k over "
                               "Re[5,30] x Im[-2,0]" % (n, args.grid, args.grid),
                   "cells_per_step": nk},
        "value_path": value_path,
        "fp64_batched_evals_per_s": fp64_batched, "mixed_evals_per_s": mixed_rate,
        "fp64_ldlt": {"evals_per_s": ldlt_rate, "lu_cells": int(flL.sum()), "cells": nk,
                      "max_abs_dtracelog": float(dtl_ldlt.max()), "max_abs_dlogdet": float(n * dtl_ldlt.max()),
                      "max_abs_dtracelog_re": float(dre_ldlt.max())},
        "fp64_plain_evals_per_s": fp64_plain, "value_over_fp64_plain": value / fp64_plain if fp64_plain else None,
        "mixed_reevaluated_cells": int(flags.sum()), "mixed_reeval_rel": 1e-6,
        "mixed_max_abs_dtracelog": float(dtl.max()), "mixed_max_abs_dlogdet": float(n * dtl.max()),
        "mixed_median_abs_dtracelog": float(np.median(dtl)),
        "mixed_p90_p99_abs_dtracelog": [float(np.percentile(dtl, 90)), float(np.percentile(dtl, 99))],
        "mixed_cells_within_1e-8_logdet": int((n * dtl <= 1e-8).sum()),
        "mixed_max_abs_dtracelog_re": dtl_re, "roofline_c64_gemm": roofline_c64,
        "mixed_gated": gated,
        "resonance_latency_ms": serving["by_kind_ms"]["resonance32"] if serving else None,
        "serving": serving,
        "uneven_resonance_search": uneven,
        "solve_wavefunction_n1000": wave,
        "server_kept_factor_solves": kept,
        "map_n200": small,
        "kernel_time_fraction": breakdown,
        "autotune": tuned,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "dropin_msmain": dropin,
        "dropin_wavefunction": dropin_wave,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
