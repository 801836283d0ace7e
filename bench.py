"""Benchmark of the MI355X sweep engine: DG cell-angle-group solves/s.

Workload (default): the 1,296,000-tet precision problem -- unit cube,
60^3 hexes x 6 Kuhn tets, interior nodes jittered +-0.15h (seeded), S8
product quadrature (128 angles), 16 groups, sigma_t_g = 1 + 0.1 g,
sigma_s = 0.9 sigma_t, q = 1, vacuum boundaries; synthetic, generated here.

One step = one source-iteration step on the device: total source
q + sigma_s/(4 pi) phi, one full transport sweep over every (cell, angle,
group), and phi = D psi.  The timed region starts with all inputs resident in
HBM; K steps are bracketed by HIP events on the engine's stream.

The JSON line also carries
  roofline      the sweep kernels (dominant): algorithmic bytes per solve
                B = w (4 + 3 f_in) + (4 w_src + 32) / N_a with
                f_in = 2 - 1/n (the mean count of upwind interior faces),
                divided by their average launch duration measured with HIP
                events over a run of sweeps, against 8 TB/s HBM;
  cpu_baseline  the CPU oracle (a port of the reference's FP64 sweep) on the
                host cores, over a bounded sample of the same sweep (the
                first S angles of the same mesh, all 16 groups).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


class HipEvents:
    """hipEvent timing on a given stream (the engine's own stream)."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so")

    def _ok(self, rc):
        if rc != 0:
            raise RuntimeError(f"HIP event call failed ({rc})")

    def create(self):
        e = ctypes.c_void_p()
        self._ok(self.hip.hipEventCreate(ctypes.byref(e)))
        return e

    def record(self, e, stream):
        self._ok(self.hip.hipEventRecord(e, ctypes.c_void_p(stream)))

    def elapsed_ms(self, a, b):
        self._ok(self.hip.hipEventSynchronize(b))
        ms = ctypes.c_float()
        self._ok(self.hip.hipEventElapsedTime(ctypes.byref(ms), a, b))
        return float(ms.value)


def alg_bytes_per_solve(width: int, n: int, na: int) -> float:
    f_in = 2.0 - 1.0 / n
    return width * (4.0 + 3.0 * f_in) + (4.0 * width + 4.0 * 8.0) / na


def time_region(ev, s, fn, count):
    start, stop = ev.create(), ev.create()
    stream = s.stream()
    s.synchronize()
    ev.record(start, stream)
    for _ in range(count):
        fn()
    ev.record(stop, stream)
    return ev.elapsed_ms(start, stop)


def gpu_leg(p, precision, steps, warmup, ev):
    from tycho2_amd.sweeper import SweeperHIP
    t0 = time.time()
    s = SweeperHIP.from_problem(p, precision)
    setup_s = time.time() - t0
    try:
        info = s.info()
        for _ in range(warmup):
            s.sweep_step()
        s.synchronize()
        ms_steps = time_region(ev, s, s.sweep_step, steps)
        ms_sweeps = time_region(ev, s, s.sweep_only, steps)
        s.synchronize()
    finally:
        s.close()
    return dict(ms_per_step=ms_steps / steps, ms_per_sweep=ms_sweeps / steps,
                n_steps=info["n_steps"], min_levels=info["min_levels"],
                psi_gb=info["psi_bytes"] / 1e9, setup_s=setup_s)


def cpu_leg(p, n_angles, threads):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as orc
    m = orc.build_mesh(p.smesh)
    src = np.ones((p.nc, n_angles, 4, p.ng))
    psi = np.zeros_like(src)
    t0 = time.time()
    orc.sweep(m, p.omega[:n_angles], p.ng, p.sigma_t, src, psi=psi,
              nthreads=threads)
    dt = time.time() - t0
    return p.nc * n_angles * p.ng / dt, dt


def counter_run(p, precision, sweeps):
    """A short program for rocprofv3 --pmc: `sweeps` sweeps, then a 4 GiB
    device-to-device copy whose counter readings calibrate the sweep's."""
    from tycho2_amd.sweeper import SweeperHIP
    hip = ctypes.CDLL("libamdhip64.so")
    with SweeperHIP.from_problem(p, precision) as s:
        for _ in range(sweeps):
            s.sweep_only()
        s.synchronize()
    nbytes = 4 << 30
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) or \
            hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
    hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3)   # device to device
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(f"# counter run: {sweeps} sweeps, calibration copy {nbytes} B",
          file=sys.stderr, flush=True)


def traffic_from_profiles(workload):
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="precision",
                    choices=["plumbing", "plain", "precision", "scattering"])
    ap.add_argument("--cpu-angles", type=int, default=0,
                    help="angles in the CPU-baseline sample (0: 2 x cores)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--counter-run", choices=["fp64", "fp32"],
                    help="only a short run for rocprofv3 --pmc")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for "
              f"--gpus {args.gpus}); the sharded design is described in "
              f"DESIGN.md", file=sys.stderr)
        return 2

    from tycho2_amd import problems
    from tycho2_amd.sweeper import FP32, FP64
    t0 = time.time()
    p = getattr(problems, args.config)()
    n = round((p.nc / 6) ** (1.0 / 3.0))
    gen_s = time.time() - t0
    print(f"# {args.config}: {p.nc} cells x {p.na} angles x {p.ng} groups "
          f"= {p.solves_per_sweep} solves/sweep (mesh {gen_s:.1f}s)",
          file=sys.stderr, flush=True)
    if args.counter_run:
        counter_run(p, FP64 if args.counter_run == "fp64" else FP32,
                    max(1, args.steps))
        return 0
    ev = HipEvents()
    r64 = gpu_leg(p, FP64, args.steps, args.warmup, ev)
    print(f"# fp64: {r64}", file=sys.stderr, flush=True)
    r32 = None
    if not args.no_fp32:
        r32 = gpu_leg(p, FP32, args.steps, args.warmup, ev)
        print(f"# fp32: {r32}", file=sys.stderr, flush=True)

    solves = p.solves_per_sweep
    value = solves / (r64["ms_per_step"] * 1e-3)
    b64 = alg_bytes_per_solve(8, n, p.na)
    launch_ms = r64["ms_per_sweep"] / r64["n_steps"]
    achieved = solves * b64 / (r64["ms_per_sweep"] * 1e-3) / 1e9
    roof = {"bound": "hbm", "kernel": "sweep_step_kernel<double,16,1>",
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic_from_profiles(args.config),
            "alg_bytes_per_solve": round(b64, 3),
            "alg_bytes_per_launch": round(solves * b64 / r64["n_steps"]),
            "avg_launch_ms": round(launch_ms, 5),
            "launches_per_sweep": r64["n_steps"]}
    out = {"metric": "DG cell-angle-group solves/s (FP64 sweep, source-"
                     "iteration step)",
           "value": round(value, 1), "unit": "solves/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(r64["ms_per_step"], 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"{args.config}: {p.nc} tets (Kuhn cube "
                                  f"{n}^3 x 6, jitter 0.15h), {p.na} angles "
                                  f"(S{p.sn}), {p.ng} groups",
                      "cells": p.nc, "angles": p.na, "groups": p.ng,
                      "solves_per_step": solves},
           "roofline": roof,
           "sweep_only_solves_per_s": round(
               solves / (r64["ms_per_sweep"] * 1e-3), 1)}
    if r32:
        b32 = alg_bytes_per_solve(4, n, p.na)
        a32 = solves * b32 / (r32["ms_per_sweep"] * 1e-3) / 1e9
        out["fp32"] = {"solves_per_s_step": round(
                           solves / (r32["ms_per_step"] * 1e-3), 1),
                       "solves_per_s_sweep": round(
                           solves / (r32["ms_per_sweep"] * 1e-3), 1),
                       "achieved_gbs": round(a32, 1),
                       "frac": round(a32 / HBM_PEAK_GBS, 4)}
    if not args.no_cpu:
        cores = min(16, len(os.sched_getaffinity(0)))
        na_cpu = args.cpu_angles or min(p.na, 2 * cores)
        rate, dt = cpu_leg(p, na_cpu, cores)
        out["cpu_baseline"] = {
            "value": round(rate, 1), "unit": "solves/s", "cores": cores,
            "kind": "port",
            "sample": f"oracle FP64 sweep of angles 0..{na_cpu - 1} of the "
                      f"same mesh, {p.ng} groups, {cores} OpenMP threads "
                      f"(one angle each), {dt:.1f}s"}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
