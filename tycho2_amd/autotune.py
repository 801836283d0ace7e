"""Autotuner of the sweep schedule and launch shape.

For one sweeper handle it times full sweeps (HIP events on the engine's
stream) over
  * angle batches  -- how many octant-ordered groups of angles are swept one
    after the other (smaller per-step working set, more steps), or side by
    side on parallel streams,
  * cells per block and groups per lane -- the launch shape of the sweep
    kernels,
  * the engine -- per-step launches replayed from a hipGraph, one
    persistent ready-queue launch, or per-step launches with the schedule's
    thin tail on the ready queue (and the queue's chunk size),
  * optionally, the reference's traversal priorities (intraAngleP,
    interAngleP, maxCellsPerStep) as the ready queue's order,
keeps the fastest setting on the handle and returns the measurements.
Results do not depend on any of these knobs (bitwise, tested).
"""
from __future__ import annotations

from .sweeper import ENGINE_HYBRID, ENGINE_LEVELS, ENGINE_QUEUE
from .timing import HipEvents, time_region

BATCHES = (1, 2, 4, 8, 16)
CHUNK_ROUNDS = (2, 4, 8)          # ready-queue chunk size, in rounds
BLOCKS = (256, 128, 64)          # threads per block
GROUPS_PER_LANE = (1, 2)
GAUSS_NOPIVOT = 1
CHAINS = (0, 1, 2, 4)             # batch level form: parallel chains


def _time_sweep(ev, s, reps):
    s.sweep_only()
    s.synchronize()
    return time_region(ev, s, s.sweep_only, reps) / reps


def tune(s, reps: int = 3, batches=BATCHES, blocks=BLOCKS,
         groups_per_lane=GROUPS_PER_LANE,
         engines=(ENGINE_LEVELS, ENGINE_QUEUE, ENGINE_HYBRID), log=None,
         priorities=()):
    """Tune handle `s` in place; returns (best, trials) where best is a dict
    {engine, batches, block, groups_per_lane, ms[, chunk_rounds,
    priority]} and trials lists every timing; log (a callable) receives one
    line per trial.  priorities: (intraAngleP, interAngleP, maxCellsPerStep)
    rules tried on the ready queue after the rest (each rebuilds the
    schedule on the host)."""
    ev = HipEvents()
    ng = s.ng
    trials = []
    best = None
    s.set_engine(ENGINE_LEVELS)
    s.set_items_per_lane(1)
    # schedule first (it rebuilds the records), with the default launch;
    # several angle batches one after the other or side by side
    for nb in batches:
        if nb > s.na:
            continue
        s.set_schedule(nb)
        s.set_launch(256 // ng if ng <= 256 else 1, 1)
        for conc in ((False, True) if nb > 1 else (False,)):
            s.set_concurrent_batches(conc)
            ms = _time_sweep(ev, s, reps)
            t = dict(engine=ENGINE_LEVELS, batches=nb, concurrent=conc,
                     block=256, groups_per_lane=1, items_per_lane=1, ms=ms)
            trials.append(t)
            if log:
                log(f"tune {t}")
            if best is None or ms < best["ms"]:
                best = t
    s.set_schedule(best["batches"])
    s.set_concurrent_batches(best["concurrent"])
    for gpl in groups_per_lane:
        if gpl > ng:
            continue
        lanes = ng // gpl
        for block in blocks:
            if block % lanes or (block, gpl) == (256, 1):
                continue
            s.set_launch(block // lanes, gpl)
            ms = _time_sweep(ev, s, reps)
            t = dict(engine=ENGINE_LEVELS, batches=best["batches"],
                     concurrent=best["concurrent"], block=block,
                     groups_per_lane=gpl, items_per_lane=1, ms=ms)
            trials.append(t)
            if log:
                log(f"tune {t}")
            if ms < best["ms"]:
                best = t
    # two items per lane (one group per lane, NoPivot)
    if s.info()["gauss"] == GAUSS_NOPIVOT:
        for block in blocks:
            if block % ng:
                continue
            s.set_launch(block // ng, 1)
            s.set_items_per_lane(2)
            ms = _time_sweep(ev, s, reps)
            s.set_items_per_lane(1)
            t = dict(engine=ENGINE_LEVELS, batches=best["batches"],
                     concurrent=best["concurrent"], block=block,
                     groups_per_lane=1, items_per_lane=2, ms=ms)
            trials.append(t)
            if log:
                log(f"tune {t}")
            if ms < best["ms"]:
                best = t
    lanes = ng // best["groups_per_lane"]
    s.set_launch(best["block"] // lanes, best["groups_per_lane"])
    s.set_items_per_lane(best["items_per_lane"])
    # the persistent engine at the best shape, over the schedules it may
    # prefer (fewer steps matter less to it)
    level_best = best
    for engine in (e for e in engines if e != ENGINE_LEVELS):
        for rounds in CHUNK_ROUNDS:
            s.set_engine(engine, chunk_rounds=rounds)
            for nb in sorted({1, level_best["batches"]}):
                s.set_schedule(nb)
                ms = _time_sweep(ev, s, reps)
                t = dict(engine=engine, batches=nb, concurrent=False,
                         block=level_best["block"],
                         groups_per_lane=level_best["groups_per_lane"],
                         items_per_lane=level_best["items_per_lane"],
                         chunk_rounds=rounds, ms=ms)
                trials.append(t)
                if log:
                    log(f"tune {t}")
                if ms < best["ms"]:
                    best = t
    # the reference's traversal priorities as the queue's order, at the best
    # chunk size found for the queue (level sets otherwise)
    if priorities:
        q_trials = [t for t in trials if t["engine"] == ENGINE_QUEUE]
        rounds = (min(q_trials, key=lambda t: t["ms"])["chunk_rounds"]
                  if q_trials else 0)
        s.set_engine(ENGINE_QUEUE, chunk_rounds=rounds)
        s.set_schedule(1)
        for rule in priorities:
            s.set_priority(*rule)
            ms = _time_sweep(ev, s, reps)
            t = dict(engine=ENGINE_QUEUE, batches=1, concurrent=False,
                     block=level_best["block"],
                     groups_per_lane=level_best["groups_per_lane"],
                     items_per_lane=level_best["items_per_lane"],
                     chunk_rounds=rounds, priority=tuple(rule), ms=ms)
            trials.append(t)
            if log:
                log(f"tune {t}")
            if ms < best["ms"]:
                best = t
        s.set_priority(*best.get("priority", (-1, 0, 0)))
    s.set_engine(best["engine"], chunk_rounds=best.get("chunk_rounds", 0))
    s.set_schedule(best["batches"])
    s.set_concurrent_batches(best["concurrent"])
    return best, trials


def tune_batch(b, reps: int = 3, rounds=CHUNK_ROUNDS,
               engines=(ENGINE_LEVELS, ENGINE_QUEUE), log=None,
               blocks=BLOCKS):
    """Tune batch `b` in place over its engine (merged per-step launches,
    their group counts interleaved or on parallel streams, or one
    persistent ready-queue launch), the level form's block size (at its
    best chain count) and the queue's chunk size; returns (best, trials)
    like tune().  A level form that refuses the members' schedules (too
    many steps) is skipped."""
    from .sweeper import T2Error
    ev = HipEvents()
    trials = []

    def trial(engine, r, conc, block):
        b.set_engine(engine, 0, r)
        b.set_concurrent(conc)
        b.set_launch(block)
        try:
            b.sweep_only()
        except T2Error as e:
            if log:
                log(f"tune batch engine={engine}: {e}")
            return
        b.synchronize()
        ms = time_region(ev, b, b.sweep_only, reps) / reps
        t = dict(engine=engine, chunk_rounds=r, concurrent=conc,
                 block=block, ms=ms)
        trials.append(t)
        if log:
            log(f"tune batch {t}")

    for engine in engines:
        if engine == ENGINE_QUEUE:
            for r in rounds:
                trial(engine, r, 0, blocks[0])
            continue
        for conc in CHAINS:
            trial(engine, 0, conc, blocks[0])
        level = [t for t in trials if t["engine"] == engine]
        if level:
            conc = min(level, key=lambda t: t["ms"])["concurrent"]
            for block in blocks[1:]:
                trial(engine, 0, conc, block)
    if not trials:
        raise RuntimeError("no batch engine could sweep these members")
    best = min(trials, key=lambda t: t["ms"])
    b.set_engine(best["engine"], 0, best["chunk_rounds"])
    b.set_concurrent(best["concurrent"])
    b.set_launch(best["block"])
    return best, trials
