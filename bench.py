"""Benchmark of the MI355X sweep engine: DG cell-angle-group solves/s.

Workload (default): the 1,296,000-tet precision problem -- unit cube,
60^3 hexes x 6 Kuhn tets, interior nodes jittered +-0.15h (seeded), S8
product quadrature (128 angles), 16 groups, sigma_t_g = 1 + 0.1 g,
sigma_s = 0.9 sigma_t, q = 1, vacuum boundaries; synthetic, generated here.

Two paths run on the same device, each autotuned first:
  plain    FP64 sweeps, FP64 psi; GMRES(20) to 1e-10
  refined  FP32 sweeps, FP32 psi, inside FP64 defect correction
           (t2_solve_refined); its scalar flux is compared with the plain
           path's after the FP32 stage and after the correction.

A step is one source-iteration step on the device (total source, one full
sweep of every (cell, angle, group), phi = D psi); K steps are bracketed by
HIP events on the engine's stream with every input resident in HBM.
`value` is the refined path's solves/s over its whole solve when it meets
the 1e-10 bar (all of its sweeps' solves / its wall time); the plain path's
numbers sit beside it.

Also on the line:
  roofline      the sweep kernels (dominant): algorithmic bytes per solve
                B = w (4 + 3 f_in) + (4 w + 32) / N_a with f_in = 2 - 1/n
                (mean upwind interior faces), over their average launch
                duration measured with HIP events, against 8 TB/s HBM;
  cpu_baseline  the CPU oracle (a port of the reference's FP64 sweep) on the
                host cores, over a bounded sample of the same sweep (the
                first S angles of the same mesh, all groups).
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


def alg_bytes_per_solve(width: int, n: int, na: int) -> float:
    f_in = 2.0 - 1.0 / n
    return width * (4.0 + 3.0 * f_in) + (4.0 * width + 4.0 * 8.0) / na


def log(msg):
    print(f"# {msg}", file=sys.stderr, flush=True)


def fixed_launch(s, launch, batches=1, concurrent=False):
    """Apply a "block,groups_per_lane" launch shape (threads per block) and
    an angle-batch schedule (one after the other or concurrent)."""
    if launch:
        vals = [int(x) for x in launch.split(",")]
        block, gpl = vals[0], vals[1]
        s.set_items_per_lane(1)
        s.set_launch(block // (s.ng // gpl), gpl)
        if len(vals) > 2:
            s.set_items_per_lane(vals[2])
    if batches > 1:
        s.set_schedule(batches)
    s.set_concurrent_batches(concurrent)


def path_leg(p, precision, steps, warmup, ev, tune, solve, launch="",
             batches=1, concurrent=False):
    """Autotune (or apply a fixed launch shape), time steps and sweeps, and
    (optionally) solve to 1e-10."""
    import numpy as np
    from tycho2_amd import autotune
    from tycho2_amd.sweeper import FP32, SweeperHIP
    from tycho2_amd.timing import time_region
    t0 = time.time()
    s = SweeperHIP.from_problem(p, precision)
    out = dict(setup_s=round(time.time() - t0, 2))
    log(f"precision {precision}: handle ready in {out['setup_s']}s")
    try:
        if tune:
            # angle batches > 2 and the pure queue never win at this size
            # (measured); the tuner still checks the launch shapes and the
            # hybrid's queued tail
            best, _ = autotune.tune(s, reps=2, batches=(1, 2),
                                    engines=(0, 2), log=log)
            out["tuned"] = best
            log(f"autotune: {best}")
        else:
            fixed_launch(s, launch, batches, concurrent)
        info = s.info()
        out["engine"] = info["engine"]
        out["queue_first_step"] = info["queue_first_step"]
        out["streams"] = (info["angle_batches"]
                          if info["concurrent_batches"] else 1)
        for _ in range(warmup):
            s.sweep_step()
        s.synchronize()
        out["ms_per_step"] = time_region(ev, s, s.sweep_step, steps) / steps
        out["ms_per_sweep"] = time_region(ev, s, s.sweep_only, steps) / steps
        log(f"precision {precision}: {out['ms_per_step']:.2f} ms/step, "
            f"{out['ms_per_sweep']:.2f} ms/sweep")
        # The sweep kernel's own launch duration: the same schedule and
        # launches replayed on one stream (concurrent angle batches off), so
        # that the sweep's time over its launches is the kernel's average
        # duration -- what a kernel trace of the run reports (the tracer
        # serialises the streams of the concurrent form).
        out["ms_per_sweep_serial"] = out["ms_per_sweep"]
        if out["streams"] > 1:
            s.set_concurrent_batches(False)
            s.sweep_only()
            s.synchronize()
            out["ms_per_sweep_serial"] = time_region(ev, s, s.sweep_only,
                                                     steps) / steps
            s.set_concurrent_batches(True)
            log(f"precision {precision}: {out['ms_per_sweep_serial']:.2f} "
                f"ms/sweep with the batches' launches on one stream")
        out["n_steps"] = info["n_steps"]
        out["psi_gb"] = info["psi_bytes"] / 1e9
        if solve:
            cfg = p.solver
            holder = {}
            if precision == FP32:
                # untimed first solve: device allocations, and phi after the
                # FP32 stage for the error report
                holder["info"], holder["phi1"] = s.solve_refined(
                    cfg["iter_max"], 1e-6, 1e-5, cfg["restart"],
                    want_phi1=True)
                out["phi1"] = holder["phi1"]
                sweeps0 = s.info()["sweeps"]

                def run():
                    holder["info"], _ = s.solve_refined(
                        cfg["iter_max"], 1e-6, 1e-5, cfg["restart"])
                ms = time_region(ev, s, run, 1)
                out["refine"] = holder["info"]
            else:
                s.solve_krylov(cfg["iter_max"], cfg["tol"], cfg["restart"])
                sweeps0 = s.info()["sweeps"]

                def run():
                    holder["hist"] = s.solve_krylov(cfg["iter_max"],
                                                    cfg["tol"],
                                                    cfg["restart"])
                ms = time_region(ev, s, run, 1)
                out["iters"] = len(holder["hist"])
                out["residual"] = float(holder["hist"][-1]) \
                    if len(holder["hist"]) else 0.0
            out["solve_ms"] = ms
            out["sweeps"] = s.info()["sweeps"] - sweeps0
            out["phi"] = s.phi()
            log(f"precision {precision}: solve {out['solve_ms']:.1f} ms, "
                f"{out['sweeps']} sweeps")
            if precision != FP32:
                # the reference for every error: FP64 GMRES far below 1e-10
                s.solve_krylov(cfg["iter_max"], 1e-13, cfg["restart"])
                out["phi_ref"] = s.phi()
                log("reference phi: FP64 GMRES to 1e-13")
            _ = np
    finally:
        s.close()
    return out


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cores() -> tuple[int, int]:
    """(threads to use, logical CPUs in this process's affinity): the
    affinity, bounded by OMP_NUM_THREADS where the environment sets it (the
    host's CPU share for this process)."""
    logical = len(os.sched_getaffinity(0))
    n = logical
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return n, logical


def cpu_leg(p, n_angles, threads, reps=5):
    """Median solves/s of `reps` timed oracle sweeps of the first n_angles
    angles (all groups) on `threads` host threads; returns (rate, total s)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as orc
    m = orc.build_mesh(p.smesh)
    src = np.ones((p.nc, n_angles, 4, p.ng))
    psi = np.zeros_like(src)
    times = []
    for _ in range(reps):
        t0 = time.time()
        orc.sweep(m, p.omega[:n_angles], p.ng, p.sigma_t, src, psi=psi,
                  nthreads=threads)
        times.append(time.time() - t0)
    return p.nc * n_angles * p.ng / float(np.median(times)), sum(times)


def solve_run(p, precision, launch="", batches=1, concurrent=False):
    """A short program for a kernel trace of one path's solve: one solve
    (FP64 GMRES(20) to the problem's tolerance, or the refined solve) after
    one untimed solve that makes the device allocations."""
    from tycho2_amd.sweeper import FP32, SweeperHIP
    cfg = p.solver
    with SweeperHIP.from_problem(p, precision) as s:
        fixed_launch(s, launch, batches, concurrent)
        for rep in range(2):
            t0 = time.time()
            if precision == FP32:
                info, _ = s.solve_refined(cfg["iter_max"], 1e-6, 1e-5,
                                          cfg["restart"])
            else:
                info = len(s.solve_krylov(cfg["iter_max"], cfg["tol"],
                                          cfg["restart"]))
            s.synchronize()
            log(f"solve run {rep}: {time.time() - t0:.3f} s, {info}")


def counter_run(p, precision, sweeps, launch="", batches=1,
                concurrent=False):
    """A short program for rocprofv3 --pmc: `sweeps` sweeps, then a 4 GiB
    device-to-device copy whose counter readings calibrate the sweep's."""
    from tycho2_amd.sweeper import SweeperHIP
    hip = ctypes.CDLL("libamdhip64.so")
    with SweeperHIP.from_problem(p, precision) as s:
        fixed_launch(s, launch, batches, concurrent)
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
    log(f"counter run: {sweeps} sweeps, calibration copy {nbytes} B")


def batch_bytes(p) -> int:
    """Device bytes of one FP64 batch member: psi, item records, a merged
    copy of the records (level form), ready-queue tables, Krylov space."""
    items = p.nc * p.na
    return items * (32 * p.ng + 80 + 80 + 20) + p.nc * 4 * p.ng * 8 * 30


def batch_leg(args, ev):
    """The 48-problem workload: waves of problems that fit the device, each
    swept as one batch by both engines and solved by lock-step GMRES."""
    import numpy as np
    from tycho2_amd import problems
    from tycho2_amd.sweeper import (ENGINE_LEVELS, ENGINE_QUEUE, FP64,
                                    SweepBatch, SweeperHIP)
    from tycho2_amd.timing import time_region
    t0 = time.time()
    ps = problems.batch(count=args.batch_count)
    log(f"batch: {len(ps)} problems generated in {time.time() - t0:.0f}s")
    budget = args.batch_gb * 1e9
    waves, cur, used = [], [], 0
    for p in sorted(ps, key=batch_bytes, reverse=True):
        if cur and used + batch_bytes(p) > budget:
            waves.append(cur)
            cur, used = [], 0
        cur.append(p)
        used += batch_bytes(p)
    waves.append(cur)
    out = dict(problems=len(ps), waves=[], per_problem=[])
    tot = dict(levels=0.0, queue=0.0, best=0.0, solves=0, solve_s=0.0)
    rule = (tuple(int(x) for x in args.batch_priority.split(","))
            if args.batch_priority else None)
    only = ({int(x) for x in args.batch_waves.split(",")}
            if args.batch_waves else None)
    for w, wave in enumerate(waves):
        if only is not None and w not in only:
            continue
        t1 = time.time()
        hs = []
        for p in wave:
            hs.append(SweeperHIP.from_problem(p, FP64))
            log(f"wave {w}: handle {len(hs)}/{len(wave)} ({p.nc} cells, "
                f"{p.na} angles, {p.ng} groups)")
        log(f"wave {w}: {len(wave)} problems, handles in "
            f"{time.time() - t1:.0f}s")
        solves = sum(p.solves_per_sweep for p in wave)
        rec = dict(problems=len(wave), solves_per_sweep=solves)
        # engines and queue chunk sizes timed by the project's autotuner
        from tycho2_amd import autotune

        def tuned(b, prefix, engines):
            _, trials = autotune.tune_batch(b, reps=args.steps,
                                            engines=engines, log=log)
            out = {}
            for t in trials:
                if t["engine"] == ENGINE_LEVELS:
                    name = (f"levels_c{t['concurrent']}" if t["concurrent"]
                            else "levels")
                else:
                    name = f"{prefix}_r{t['chunk_rounds']}"
                out[name] = t["ms"]
                log(f"wave {w}: {name} {t['ms']:.2f} ms/sweep, "
                    f"{solves / t['ms'] / 1e6:.1f} G solves/s")
            return out

        timings = {"levels": float("inf")}
        with SweepBatch(hs) as b:
            timings.update(tuned(b, "queue", (ENGINE_LEVELS, ENGINE_QUEUE)))
        if rule:
            # the ready queue in the order of the reference's traversal
            # priorities (members re-scheduled, so a new batch)
            for h in hs:
                h.set_priority(*rule)
            with SweepBatch(hs) as b:
                timings.update(tuned(b, "queue_prio", (ENGINE_QUEUE,)))
        best = min(timings, key=timings.get)
        if rule and not best.startswith("queue_prio"):
            for h in hs:
                h.set_priority(-1)
        rec["ms_per_sweep"] = {k: (round(v, 3) if v != float("inf")
                                   else None)
                               for k, v in timings.items()}
        rec["engine"] = best
        tot["levels"] += timings["levels"]
        tot["queue"] += min(timings[k] for k in timings
                            if k.startswith("queue_r"))
        if rule:
            tot["queue_prio"] = tot.get("queue_prio", 0.0) + min(
                timings[k] for k in timings if k.startswith("queue_prio"))
        tot["levels_conc"] = tot.get("levels_conc", 0.0) + min(
            [v for k, v in timings.items() if k.startswith("levels_c")] +
            [float("inf")])
        tot["best"] += timings[best]
        tot["solves"] += solves
        with SweepBatch(hs) as b:
            if best.startswith("levels"):
                b.set_engine(ENGINE_LEVELS, 0, 0)
                b.set_concurrent(0 if best == "levels"
                                 else int(best[len("levels_c"):]))
            else:
                b.set_engine(ENGINE_QUEUE, 0,
                             int(best.rsplit("_r", 1)[1]))
            t2 = time.time()
            its, sec = b.solve_krylov(300, 1e-10, 20)
            rec["solve_s"] = round(time.time() - t2, 3)
            tot["solve_s"] += rec["solve_s"]
            log(f"wave {w}: batched GMRES(20) to 1e-10 on {best} in "
                f"{rec['solve_s']}s, iterations {int(its.min())}-"
                f"{int(its.max())}")
        for p, h, k, t in zip(wave, hs, its, sec):
            out["per_problem"].append(dict(
                name=p.name, cells=p.nc, angles=p.na, groups=p.ng,
                gmres_iters=int(k), time_to_solution_s=round(float(t), 3)))
            h.close()
        out["waves"].append(rec)
    tts = np.array([q["time_to_solution_s"] for q in out["per_problem"]])
    res = {"metric": "DG cell-angle-group solves/s, 48 independent problems "
                     "batched on one GPU (FP64; each wave on its faster "
                     "engine)",
           "value": round(tot["solves"] / (tot["best"] * 1e-3)),
           "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
           "higher_is_better": True, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"{len(ps)} seeded Kuhn cubes, n in "
                                  "[10, 45], S4/6/8/12, 1/4/16/32 groups",
                      "waves": len(waves),
                      "waves_run": sorted(only) if only is not None
                      else list(range(len(waves))),
                      "schedule": ("level sets" if rule is None else
                                   "level sets, and for the queue also "
                                   "priorities intraAngleP/interAngleP/"
                                   f"maxCellsPerStep = {list(rule)}")},
           "levels_solves_per_s": round(tot["solves"] /
                                        (tot["levels"] * 1e-3)),
           "levels_concurrent_solves_per_s": round(
               tot["solves"] / (tot["levels_conc"] * 1e-3)),
           "queue_solves_per_s": round(tot["solves"] / (tot["queue"] * 1e-3)),
           "queue_priority_solves_per_s": (
               round(tot["solves"] / (tot["queue_prio"] * 1e-3))
               if rule else None),
           "solve_wall_s": round(tot["solve_s"], 3),
           "time_to_solution_s": {"min": float(tts.min()),
                                  "median": float(np.median(tts)),
                                  "max": float(tts.max())},
           "detail": out}
    print(json.dumps(res), flush=True)


def schedule_leg(p, args, ev):
    """Sweep time of the level sets against the reference's traversal
    priorities (the ready queue's ticket order), per engine, on one FP64
    handle at the default launch shape."""
    from tycho2_amd.sweeper import (ENGINE_HYBRID, ENGINE_LEVELS,
                                    ENGINE_QUEUE, FP64, SweeperHIP)
    from tycho2_amd.timing import time_region
    widths = [int(w) for w in args.study_widths.split(",")]
    rules = [(-1, 0, 0)] + [(i, j, w) for w in widths
                            for (i, j) in ((1, 0), (3, 1), (4, 2), (0, 0))]
    rows = []
    with SweeperHIP.from_problem(p, FP64) as s:
        for rule in rules:
            t0 = time.time()
            s.set_priority(*rule)
            build_s = round(time.time() - t0, 2)
            n_steps = s.info()["n_steps"]
            engines = ((ENGINE_LEVELS, ENGINE_QUEUE, ENGINE_HYBRID)
                       if rule[0] < 0 else (ENGINE_QUEUE, ENGINE_LEVELS))
            for engine in engines:
                s.set_engine(engine)
                if engine == ENGINE_LEVELS and n_steps > 20000:
                    continue    # one launch per step: not a useful pairing
                s.sweep_only()
                s.synchronize()
                ms = time_region(ev, s, s.sweep_only, args.steps) / args.steps
                row = dict(intra=rule[0], inter=rule[1], width=rule[2],
                           engine=engine, steps=n_steps, build_s=build_s,
                           ms_per_sweep=round(ms, 3),
                           solves_per_s=round(p.solves_per_sweep /
                                              (ms * 1e-3)))
                rows.append(row)
                log(f"schedule study {row}")
    base = {r["engine"]: r["ms_per_sweep"] for r in rows if r["intra"] < 0}
    best = min(rows, key=lambda r: r["ms_per_sweep"])
    res = {"metric": "DG cell-angle-group solves/s by sweep schedule "
                     "(level sets vs traversal priorities), FP64 sweep",
           "value": best["solves_per_s"], "unit": "solves/s", "n_gpus": 1,
           "steps": args.steps, "higher_is_better": True, "dtype": "f64",
           "data": "synthetic",
           "config": {"workload": f"{args.config}: {p.nc} tets, {p.na} "
                                  f"angles, {p.ng} groups",
                      "solves_per_sweep": p.solves_per_sweep},
           "level_sets_ms_per_sweep": base, "best": best, "rows": rows}
    print(json.dumps(res), flush=True)


def host_leg(p, args):
    """The drop-in path: t2_sweep_host on host arrays in the reference
    layout (source uploaded, swept, psi downloaded; ordinary pageable
    memory, as the reference's PsiData is), wall-clock timed."""
    import numpy as np
    from tycho2_amd.sweeper import FP64, SweeperHIP
    shape = (p.nc, p.na, 4, p.ng)
    src = np.full(shape, 1.0)
    psi = np.zeros(shape)
    with SweeperHIP.from_problem(p, FP64) as s:
        s.sweep(psi, src)                        # warm-up: staging buffers
        t0 = time.perf_counter()
        for _ in range(args.steps):
            s.sweep(psi, src)
        sec = (time.perf_counter() - t0) / args.steps
        s.sweep_only()
        s.synchronize()
        from tycho2_amd.timing import HipEvents, time_region
        ev = HipEvents()
        dev_ms = time_region(ev, s, s.sweep_only, args.steps) / args.steps
    moved = 2 * psi.nbytes
    res = {"metric": "DG cell-angle-group solves/s of the drop-in host call "
                     "(t2_sweep_host, PCIe-inclusive), FP64",
           "value": round(p.solves_per_sweep / sec), "unit": "solves/s",
           "n_gpus": 1, "steps": args.steps, "higher_is_better": True,
           "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"{args.config}: {p.nc} tets, {p.na} "
                                  f"angles, {p.ng} groups"},
           "seconds_per_call": round(sec, 4),
           "host_bytes_moved_per_call": moved,
           "host_link_gb_per_s": round(moved / sec / 1e9, 2),
           "device_sweep_ms": round(dev_ms, 3),
           "device_only_solves_per_s": round(p.solves_per_sweep /
                                             (dev_ms * 1e-3))}
    print(json.dumps(res), flush=True)


def batch_refined_leg(args, ev):
    """The 48-problem workload solved twice per wave, each member to <=1e-10:
    FP64 members by lock-step batched GMRES(20), and FP32 members by the
    lock-step refined solve (FP32 sweeps, FP64 defect correction); errors
    are max-relative phi against FP64 GMRES to 1e-13 on the same members."""
    import numpy as np
    from tycho2_amd import autotune, problems
    from tycho2_amd.sweeper import FP32, FP64, SweepBatch, SweeperHIP
    t0 = time.time()
    ps = problems.batch(count=args.batch_count)
    log(f"batch: {len(ps)} problems generated in {time.time() - t0:.0f}s")
    budget = args.batch_gb * 1e9
    waves, cur, used = [], [], 0
    for p in sorted(ps, key=batch_bytes, reverse=True):
        if cur and used + batch_bytes(p) > budget:
            waves.append(cur)
            cur, used = [], 0
        cur.append(p)
        used += batch_bytes(p)
    waves.append(cur)
    only = ({int(x) for x in args.batch_waves.split(",")}
            if args.batch_waves else None)
    per, tot = [], dict(plain_s=0.0, refined_s=0.0, plain_solves=0,
                        refined_solves=0)
    for w, wave in enumerate(waves):
        if only is not None and w not in only:
            continue
        hs = [SweeperHIP.from_problem(p, FP64) for p in wave]
        with SweepBatch(hs) as b:
            best64, _ = autotune.tune_batch(b, reps=2, log=log)
            b.solve_krylov(300, 1e-10, 20)                  # warm-up
            sw0 = [h.info()["sweeps"] for h in hs]
            t1 = time.perf_counter()
            its, sec64 = b.solve_krylov(300, 1e-10, 20)
            wall64 = time.perf_counter() - t1
            sweeps64 = [h.info()["sweeps"] - s0 for h, s0 in zip(hs, sw0)]
            phi64 = [h.phi() for h in hs]
            b.solve_krylov(600, 1e-13, 20)
            ref = [h.phi() for h in hs]
        for h in hs:
            h.close()
        log(f"wave {w}: FP64 batched GMRES to 1e-10 in {wall64:.3f}s "
            f"(engine {best64['engine']})")
        hs = [SweeperHIP.from_problem(p, FP32) for p in wave]
        with SweepBatch(hs) as b:
            best, _ = autotune.tune_batch(b, reps=2, log=log)
            b.solve_refined(300, 1e-6, 1e-5, 20)           # warm-up
            sw0 = [h.info()["sweeps"] for h in hs]
            t1 = time.perf_counter()
            infos, sec32 = b.solve_refined(300, 1e-6, 1e-5, 20)
            wall32 = time.perf_counter() - t1
            sweeps32 = [h.info()["sweeps"] - s0 for h, s0 in zip(hs, sw0)]
            phi32 = [h.phi() for h in hs]
        for h in hs:
            h.close()
        log(f"wave {w}: FP32 batched refined solve in {wall32:.3f}s "
            f"(engine {best['engine']})")
        tot["plain_s"] += wall64
        tot["refined_s"] += wall32
        for i, p in enumerate(wave):
            tot["plain_solves"] += sweeps64[i] * p.solves_per_sweep
            tot["refined_solves"] += sweeps32[i] * p.solves_per_sweep
            per.append(dict(
                name=p.name, cells=p.nc, angles=p.na, groups=p.ng,
                plain_tts_s=round(float(sec64[i]), 3),
                refined_tts_s=round(float(sec32[i]), 3),
                plain_err=maxrel(phi64[i], ref[i]),
                refined_err=maxrel(phi32[i], ref[i]),
                refined_err_bar_met=bool(maxrel(phi32[i], ref[i]) <= 1e-10),
                sweeps=[sweeps64[i], sweeps32[i]]))
    ref_sps = tot["refined_solves"] / tot["refined_s"]
    plain_sps = tot["plain_solves"] / tot["plain_s"]
    errs = np.array([q["refined_err"] for q in per])
    res = {"metric": "DG cell-angle-group solves/s, 48 independent problems "
                     "batched on one GPU: FP32 sweeps + FP64 defect "
                     "correction (lock-step refined solve) vs FP64 "
                     "batched GMRES, each to <=1e-10 phi err",
           "value": round(ref_sps), "unit": "solves/s", "n_gpus": 1,
           "higher_is_better": True, "dtype": "f32+f64",
           "data": "synthetic",
           "config": {"workload": f"{len(ps)} seeded Kuhn cubes, n in "
                                  "[10, 45], S4/6/8/12, 1/4/16/32 groups",
                      "waves": len(waves)},
           "plain_fp64_solves_per_s": round(plain_sps),
           "speedup_solves_per_s": round(ref_sps / plain_sps, 3),
           "wall_s": {"plain": round(tot["plain_s"], 3),
                      "refined": round(tot["refined_s"], 3)},
           "speedup_time_to_solution": round(tot["plain_s"] /
                                             tot["refined_s"], 3),
           "refined_err_max": float(errs.max()),
           "all_members_meet_1e10": bool((errs <= 1e-10).all()),
           "per_problem": per}
    print(json.dumps(res), flush=True)


def traffic_from_profiles(workload, precision, launches_per_sweep=None):
    """Measured HBM bytes of one sweep-kernel launch: the committed
    counter figures per sweep, over this run's launches per sweep (a
    different angle-batch schedule splits the same traffic over more
    launches)."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f).get(precision)
    if not rec:
        return None
    if launches_per_sweep is None:
        return rec.get("hbm_bytes_per_launch")
    return round(rec["hbm_bytes_per_sweep"] / launches_per_sweep)


def maxrel(a, b):
    import numpy as np
    return float(np.abs(a - b).max() / np.abs(b).max())


ELEM_FLOOR = 1e-6   # of max |phi_ref|: the floor of the elementwise error


def phi_errors(a, b):
    """The other two error measures of phi against the reference: L1
    relative, sum |a - b| / sum |b|, and elementwise relative,
    max |a - b| / max(|b|, ELEM_FLOOR max |b|)."""
    import numpy as np
    d = np.abs(a - b)
    ab = np.abs(b)
    floor = ELEM_FLOOR * ab.max()
    return {"l1_rel": float(d.sum() / ab.sum()),
            "elem_maxrel": float((d / np.maximum(ab, floor)).max()),
            "elem_floor": ELEM_FLOOR}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="precision",
                    choices=["plumbing", "plain", "precision", "scattering",
                             "batch"])
    ap.add_argument("--batch-count", type=int, default=48)
    ap.add_argument("--batch-gb", type=float, default=200.0,
                    help="device memory per wave of the batch workload")
    ap.add_argument("--batch-refined", action="store_true",
                    help="with --config batch: FP32 refined vs FP64 batched "
                         "solves")
    ap.add_argument("--batch-priority", default="",
                    help="intraAngleP,interAngleP,maxCellsPerStep for every "
                         "batch member (default: level sets)")
    ap.add_argument("--batch-waves", default="",
                    help="comma-separated waves of the batch workload to run "
                         "(default: all)")
    ap.add_argument("--cpu-angles", type=int, default=0,
                    help="angles in the CPU-baseline sample (at least one "
                         "per core)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solve", action="store_true",
                    help="skip the two solves to 1e-10 (steps only)")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--batches", type=int, default=1,
                    help="with --no-tune: angle batches of the schedule")
    ap.add_argument("--concurrent", action="store_true",
                    help="with --no-tune: sweep the angle batches on "
                         "parallel streams")
    ap.add_argument("--launch", default="",
                    help="with --no-tune: threads per block, groups per "
                         "lane and optionally items per lane of the sweep "
                         "kernels, e.g. 64,1 or 64,1,2")
    ap.add_argument("--host-path", action="store_true",
                    help="only time the drop-in host call (PCIe-inclusive)")
    ap.add_argument("--schedule-study", action="store_true",
                    help="only time level sets against traversal-priority "
                         "schedules on each engine")
    ap.add_argument("--study-widths", default="32768,262144",
                    help="maxCellsPerStep values of the schedule study")
    ap.add_argument("--counter-run", choices=["fp64", "fp32"],
                    help="only a short run for rocprofv3 --pmc")
    ap.add_argument("--solve-run", choices=["fp64", "fp32"],
                    help="only two solves of one path (for a kernel trace; "
                         "the fixed launch flags apply)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: measures one GPU only (asked for --gpus "
              f"{args.gpus}). Several GPUs would each sweep a share of the "
              f"energy groups (groups are decoupled, so shards exchange only "
              f"convergence norms); that path is not built.", file=sys.stderr)
        return 2

    from tycho2_amd import problems
    from tycho2_amd.sweeper import FP32, FP64
    from tycho2_amd.timing import HipEvents
    if args.config == "batch":
        if args.batch_refined:
            batch_refined_leg(args, HipEvents())
        else:
            batch_leg(args, HipEvents())
        return 0
    t0 = time.time()
    p = getattr(problems, args.config)()
    n = round((p.nc / 6) ** (1.0 / 3.0))
    log(f"{args.config}: {p.nc} cells x {p.na} angles x {p.ng} groups = "
        f"{p.solves_per_sweep} solves/sweep (mesh {time.time() - t0:.1f}s)")
    if args.counter_run:
        counter_run(p, FP64 if args.counter_run == "fp64" else FP32,
                    max(1, args.steps), args.launch, args.batches,
                    args.concurrent)
        return 0
    if args.solve_run:
        solve_run(p, FP64 if args.solve_run == "fp64" else FP32, args.launch,
                  args.batches, args.concurrent)
        return 0
    ev = HipEvents()
    if args.schedule_study:
        schedule_leg(p, args, ev)
        return 0
    if args.host_path:
        host_leg(p, args)
        return 0
    solve = not args.no_solve
    tune = not args.no_tune
    r64 = path_leg(p, FP64, args.steps, args.warmup, ev, tune, solve,
                   args.launch, args.batches, args.concurrent)
    log("fp64: " + str({k: v for k, v in r64.items()
                        if k not in ("phi", "phi_ref")}))
    r32 = path_leg(p, FP32, args.steps, args.warmup, ev, tune, solve,
                   args.launch, args.batches, args.concurrent)
    log("fp32: " + str({k: v for k, v in r32.items()
                        if k not in ("phi", "phi1")}))

    solves = p.solves_per_sweep
    b64, b32 = alg_bytes_per_solve(8, n, p.na), alg_bytes_per_solve(4, n, p.na)

    def roof(r, width, bytes_per_solve):
        t = r.get("tuned") or {}
        gpl = t.get("groups_per_lane", 1)
        ty = "double" if width == 8 else "float"
        if r.get("engine", 0) == 1:
            kernel = f"queue_sweep_kernel<{ty},{p.ng},{gpl}>"
            launches = 1
        elif r.get("engine", 0) == 2:
            kernel = (f"sweep_step_kernel<{ty},{p.ng},{gpl}> (steps before "
                      f"{r['queue_first_step']}) + queue_sweep_kernel")
            launches = r["queue_first_step"] + 1
        else:
            kernel = f"sweep_step_kernel<{ty},{p.ng},{gpl}>"
            launches = r["n_steps"]
        # concurrent angle batches: `streams` chains of launches
        streams = r.get("streams", 1) if r.get("engine", 0) == 0 else 1
        per_launch = solves * bytes_per_solve / launches
        # the kernel's average launch duration: the schedule's launches on
        # one stream (= the sweep's time / launches, no gaps in a replay)
        avg_ms = r["ms_per_sweep_serial"] / launches
        ach = per_launch / (avg_ms * 1e-3) / 1e9
        # the tuned concurrent form: launches overlap, so its time per launch
        # is an interval, not a kernel duration
        ach_c = solves * bytes_per_solve / (r["ms_per_sweep"] * 1e-3) / 1e9
        return {"bound": "hbm", "kernel": kernel, "achieved": round(ach, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4),
                "traffic": traffic_from_profiles(args.config,
                                                 f"fp{8 * width}",
                                                 launches),
                "alg_bytes_per_solve": round(bytes_per_solve, 3),
                "alg_bytes_per_launch": round(per_launch),
                "avg_launch_ms": round(avg_ms, 5),
                "launches_per_sweep": launches,
                "streams": streams,
                "concurrent_interval_ms": round(r["ms_per_sweep"] / launches,
                                                5),
                "achieved_concurrent": round(ach_c, 1),
                "frac_concurrent": round(ach_c / HBM_PEAK_GBS, 4)}

    plain = {"solves_per_s_step": round(solves / (r64["ms_per_step"] * 1e-3)),
             "solves_per_s_sweep": round(solves /
                                         (r64["ms_per_sweep"] * 1e-3)),
             "ms_per_sweep": round(r64["ms_per_sweep"], 3),
             "tuned": r64.get("tuned"),
             "roofline": roof(r64, 8, b64)}
    refined = {"solves_per_s_step": round(solves / (r32["ms_per_step"] * 1e-3)),
               "solves_per_s_sweep": round(solves /
                                           (r32["ms_per_sweep"] * 1e-3)),
               "ms_per_sweep": round(r32["ms_per_sweep"], 3),
               "tuned": r32.get("tuned"),
               "roofline": roof(r32, 4, b32)}
    value = plain["solves_per_s_step"]
    metric_path = "FP64 plain, source-iteration step"
    ms_per_step = r64["ms_per_step"]
    dtype = "f64"
    roofline = plain["roofline"]
    if solve:
        plain.update(solve_s=round(r64["solve_ms"] / 1e3, 3),
                     gmres_iters=r64["iters"], sweeps=r64["sweeps"],
                     solves_per_s_solve=round(r64["sweeps"] * solves /
                                              (r64["solve_ms"] * 1e-3)))
        ref = r64["phi_ref"]
        e1 = maxrel(r32["phi1"], ref)
        e2 = maxrel(r32["phi"], ref)
        plain["phi_maxrel_err"] = maxrel(r64["phi"], ref)
        plain["phi_err_other"] = phi_errors(r64["phi"], ref)
        plain["error_reference"] = ("FP64 GMRES(20) to 1e-13 on the same "
                                    "handle (untimed)")
        refined.update(solve_s=round(r32["solve_ms"] / 1e3, 3),
                       sweeps=r32["sweeps"],
                       solves_per_s_solve=round(r32["sweeps"] * solves /
                                                (r32["solve_ms"] * 1e-3)),
                       phi_maxrel_err_after_fp32_stage=e1,
                       phi_maxrel_err_after_correction=e2,
                       phi_err_other_after_fp32_stage=phi_errors(
                           r32["phi1"], ref),
                       phi_err_other_after_correction=phi_errors(
                           r32["phi"], ref),
                       gmres_iters=[r32["refine"]["iters_sweep"],
                                    r32["refine"]["iters_correction"]])
        refined["speedup_solves_per_s"] = round(
            refined["solves_per_s_solve"] / plain["solves_per_s_solve"], 3)
        refined["speedup_time_to_solution"] = round(
            plain["solve_s"] / refined["solve_s"], 3)
        if e2 <= 1e-10:
            value = refined["solves_per_s_solve"]
            metric_path = ("FP32 sweeps + FP64 defect correction, whole "
                           "solve to <=1e-10 max-rel phi err vs FP64 GMRES "
                           "to 1e-13")
            ms_per_step = r32["solve_ms"] / r32["sweeps"]
            dtype = "f32+f64"
            roofline = refined["roofline"]
    out = {"metric": "DG cell-angle-group solves/s, FP64 plain vs "
                     "FP32+FP64-refined at <=1e-10 phi err "
                     f"[value: {metric_path}]",
           "value": float(value), "unit": "solves/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(ms_per_step, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": dtype, "data": "synthetic",
           "config": {"workload": f"{args.config}: {p.nc} tets (Kuhn cube "
                                  f"{n}^3 x 6, jitter 0.15h), {p.na} angles "
                                  f"(S{p.sn}), {p.ng} groups",
                      "cells": p.nc, "angles": p.na, "groups": p.ng,
                      "solves_per_sweep": solves},
           "roofline": roofline, "plain_fp64": plain,
           "refined_fp32_fp64": refined}
    if not args.no_cpu:
        cores, logical = host_cores()
        # at least one angle per thread (the oracle's threads take angles)
        na_cpu = min(p.na, max(args.cpu_angles, cores))
        rate, dt = cpu_leg(p, na_cpu, cores)
        out["cpu_baseline"] = {
            "value": round(rate, 1), "unit": "solves/s", "cores": cores,
            "kind": "port", "cpu_model": cpu_model(),
            "logical_cpus_in_affinity": logical,
            "sample": f"median of 5 oracle FP64 sweeps of angles "
                      f"0..{na_cpu - 1} of the same mesh, {p.ng} groups, "
                      f"{cores} OpenMP threads (every core of this "
                      f"process's CPU share; at least one angle each), "
                      f"{dt:.1f}s in total"}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
