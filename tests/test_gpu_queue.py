"""Ready-queue sweep engine (one persistent launch with atomic upwind
counters) against the per-step level engine and the oracle.

The two engines run the same cell solve on the same records, so every result
must agree bit for bit; the wait bound of the queue is exercised from the
host with a tiny poll limit, which makes the kernel give up and exit.
"""
import numpy as np
import pytest

import oracle as orc
from tycho2_amd import mesh as meshmod
from tycho2_amd.quadrature import product_quadrature
from tycho2_amd.sweeper import (ENGINE_HYBRID, ENGINE_LEVELS, ENGINE_QUEUE,
                                ERR_TIMEOUT, FP32, FP64, SweeperHIP, T2Error)

pytestmark = pytest.mark.gpu

SPIN = 1 << 20      # polls per wait: far above any wait of these sizes


def problem(n, sn, ng, seed, jitter=0.15):
    sm = meshmod.kuhn_cube(n, jitter, seed=seed)
    conn = meshmod.tycho_connectivity(sm)
    omega, wt = (orc.quadrature(sn) if sn == 2 else product_quadrature(sn))
    rng = np.random.default_rng(seed)
    st = rng.uniform(0.5, 3.0, (sm.n_cells, ng))
    ss = st * rng.uniform(0.1, 0.9, (sm.n_cells, ng))
    q = rng.uniform(0.0, 2.0, (sm.n_cells, 4, ng))
    return sm, conn, omega, wt, st, ss, q


def run_steps(args, precision, engine, steps=3, batches=1, launch=None,
              rounds=0, split=None):
    sm, conn, omega, wt, st, ss, q = args
    ng = q.shape[2]
    with SweeperHIP(sm, conn, omega, wt, ng, st, ss, precision) as s:
        s.set_engine(engine, SPIN, rounds)
        if split is not None:
            n = s.info()["n_steps"]
            s.set_queue_split({"half": n // 2, "all": n}.get(split, split))
        if batches != 1:
            s.set_schedule(batches)
        if launch:
            s.set_launch(*launch)
        s.set_source_isotropic(q)
        for _ in range(steps):
            s.sweep_step()
        return s.phi(), s.psi_angle(omega.shape[0] - 1)


@pytest.mark.parametrize("engine", [ENGINE_LEVELS, ENGINE_QUEUE,
                                    ENGINE_HYBRID])
@pytest.mark.parametrize("precision", [FP64, FP32])
@pytest.mark.parametrize("lagged", [False, True], ids=["acyclic", "lagged"])
def test_groups_per_lane_bitwise(engine, precision, lagged):
    """Two and four groups per lane (all groups' loads issued before any
    group is solved and stored) give one group per lane's results bit for
    bit on every engine, with and without lagged faces."""
    print(f"groups per lane engine={engine} precision={precision} "
          f"lagged={lagged}", flush=True)
    args = (problem(3, 2, 8, seed=3, jitter=0.2) if lagged
            else problem(4, 4, 8, seed=44))
    ref = run_steps(args, precision, engine, launch=(8, 1))
    for launch in [(16, 2), (32, 4)]:
        got = run_steps(args, precision, engine, launch=launch)
        assert np.array_equal(ref[0], got[0])
        assert np.array_equal(ref[1], got[1])


@pytest.mark.parametrize("engine", [ENGINE_LEVELS, ENGINE_HYBRID])
@pytest.mark.parametrize("precision,ng", [(FP64, 1), (FP64, 8), (FP32, 4),
                                          (FP32, 8), (FP32, 16), (FP32, 64)])
@pytest.mark.parametrize("lagged", [False, True], ids=["acyclic", "lagged"])
def test_two_items_per_lane_bitwise(engine, precision, ng, lagged):
    """Two items per lane (both items' loads issued before either solve)
    give one item per lane's results bit for bit, also on odd item counts
    where the last wavefront holds one set."""
    print(f"two items per lane engine={engine} precision={precision} "
          f"ng={ng} lagged={lagged}", flush=True)
    args = (problem(3, 2, ng, seed=3, jitter=0.2) if lagged
            else problem(4, 4, ng, seed=44))
    ref = run_steps(args, precision, engine)
    sm, conn, omega, wt, st, ss, q = args
    with SweeperHIP(sm, conn, omega, wt, ng, st, ss, precision) as s:
        s.set_engine(engine, SPIN)
        s.set_launch(max(1, 64 // ng), 1)
        s.set_items_per_lane(2)
        with pytest.raises(T2Error):
            s.set_launch(0, 2)             # several groups per lane
        s.set_source_isotropic(q)
        for _ in range(3):
            s.sweep_step()
        got = (s.phi(), s.psi_angle(omega.shape[0] - 1))
    assert np.array_equal(ref[0], got[0])
    assert np.array_equal(ref[1], got[1])


@pytest.mark.parametrize("precision,ng", [(FP64, 1), (FP64, 4), (FP64, 16),
                                          (FP32, 1), (FP32, 16)])
def test_queue_matches_levels_bitwise(precision, ng):
    print(f"queue vs levels precision={precision} ng={ng}", flush=True)
    args = problem(5, 4, ng, seed=40 + ng)
    ref = run_steps(args, precision, ENGINE_LEVELS)
    got = run_steps(args, precision, ENGINE_QUEUE)
    assert np.array_equal(ref[0], got[0])
    assert np.array_equal(ref[1], got[1])


@pytest.mark.parametrize("batches,launch,rounds", [(2, None, 1),
                                                   (1, (8, 2), 4),
                                                   (4, (4, 1), 16),
                                                   (1, None, 64)])
def test_queue_schedules_and_shapes_bitwise(batches, launch, rounds):
    print(f"queue batches={batches} launch={launch} rounds={rounds}",
          flush=True)
    args = problem(4, 4, 16, seed=52)
    ref = run_steps(args, FP64, ENGINE_LEVELS)
    got = run_steps(args, FP64, ENGINE_QUEUE, batches=batches, launch=launch,
                    rounds=rounds)
    assert np.array_equal(ref[0], got[0])
    assert np.array_equal(ref[1], got[1])


@pytest.mark.parametrize("precision,ng,split", [
    (FP64, 16, "half"), (FP32, 16, "half"), (FP64, 1, 0), (FP64, 4, "all"),
    (FP64, 4, -1), (FP32, 1, 3)])
def test_hybrid_matches_levels_bitwise(precision, ng, split):
    """Per-step launches up to the split step, the ready queue after it."""
    print(f"hybrid precision={precision} ng={ng} split={split}", flush=True)
    args = problem(5, 4, ng, seed=80 + ng)
    ref = run_steps(args, precision, ENGINE_LEVELS)
    got = run_steps(args, precision, ENGINE_HYBRID, split=split)
    assert np.array_equal(ref[0], got[0])
    assert np.array_equal(ref[1], got[1])


@pytest.mark.parametrize("engine", [ENGINE_QUEUE, ENGINE_HYBRID])
def test_queue_lagged_sweep_matches_oracle(engine):
    """Cyclic mesh: lagged faces read the pre-sweep psi in the queue too."""
    print(f"queue cyclic sweep engine={engine}", flush=True)
    sm = meshmod.kuhn_cube(3, 0.2, seed=3)
    conn = meshmod.tycho_connectivity(sm)
    omega, wt = orc.quadrature(2)
    rng = np.random.default_rng(8)
    st = rng.uniform(1.0, 2.0, (sm.n_cells, 1))
    src = rng.uniform(0.1, 1.0, (sm.n_cells, omega.shape[0], 4, 1))
    prev = rng.uniform(0.0, 1.0, src.shape)
    psi = prev.copy()
    with SweeperHIP(sm, conn, omega, wt, 1, st, 0.5 * st, FP64) as s:
        s.set_engine(engine, SPIN)
        s.set_queue_split(s.info()["n_steps"] // 2)
        assert s.info()["n_lagged"] > 0
        s.sweep(psi, src)
        if engine == ENGINE_HYBRID:
            assert s.info()["queue_first_step"] == s.info()["n_steps"] // 2
    ref, _ = orc.sweep(orc.build_mesh(sm), omega, 1, st, src, psi=prev.copy())
    assert np.abs(psi - ref).max() <= 1e-12 * np.abs(ref).max()


def test_queue_krylov_bitwise():
    print("queue krylov", flush=True)
    sm, conn, omega, wt, st, ss, q = problem(4, 4, 4, seed=61)
    out = []
    for engine in (ENGINE_LEVELS, ENGINE_QUEUE):
        with SweeperHIP(sm, conn, omega, wt, 4, st, ss, FP64) as s:
            s.set_engine(engine, SPIN)
            s.set_source_isotropic(q)
            hist = s.solve_krylov(200, 1e-10, 20)
            out.append((s.phi(), len(hist)))
    assert out[0][1] == out[1][1]
    assert np.array_equal(out[0][0], out[1][0])


def test_queue_wait_limit_abandons_and_recovers():
    """With a poll limit of 1 some wait runs out: the kernel exits, the next
    synchronising call reports it, and later sweeps are exact again."""
    print("queue wait limit", flush=True)
    args = problem(6, 4, 4, seed=70)
    ref = run_steps(args, FP64, ENGINE_LEVELS, steps=1)
    sm, conn, omega, wt, st, ss, q = args
    with SweeperHIP(sm, conn, omega, wt, 4, st, ss, FP64) as s:
        s.set_source_isotropic(q)
        s.set_engine(ENGINE_QUEUE, 1)
        timeouts = 0
        for _ in range(5):
            try:
                s.sweep_only()
                s.synchronize()
            except T2Error as e:
                assert e.code == ERR_TIMEOUT
                timeouts += 1
        assert timeouts >= 1
        s.set_engine(ENGINE_QUEUE, SPIN)
        s.set_phi(np.zeros_like(ref[0]))
        s.sweep_step()
        assert np.array_equal(s.phi(), ref[0])


@pytest.mark.parametrize("engine", [ENGINE_LEVELS, ENGINE_QUEUE])
def test_face_order_bitwise(engine):
    """Items sorted by incoming-face pattern within a step: same results."""
    from tycho2_amd.sweeper import ORDER_CELL, ORDER_FACES
    print(f"face order engine={engine}", flush=True)
    sm, conn, omega, wt, st, ss, q = problem(5, 4, 16, seed=90)
    out = []
    for order in (ORDER_CELL, ORDER_FACES):
        with SweeperHIP(sm, conn, omega, wt, 16, st, ss, FP64) as s:
            s.set_engine(engine, SPIN)
            s.set_order(order)
            s.set_source_isotropic(q)
            for _ in range(2):
                s.sweep_step()
            out.append((s.phi(), s.psi_angle(5)))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("engine,intra,inter,width", [
    (ENGINE_QUEUE, 3, 1, 100), (ENGINE_QUEUE, 1, 0, 0),
    (ENGINE_QUEUE, 4, 2, 4096), (ENGINE_QUEUE, 0, 0, 512),
    (ENGINE_LEVELS, 3, 1, 1000), (ENGINE_HYBRID, 1, 1, 2000),
    (ENGINE_LEVELS, 2, 2, 0)])
def test_priority_schedule_bitwise(engine, intra, inter, width):
    """The reference's heap priorities as the schedule (for the ready queue,
    its ticket order): results equal the level sets' bit for bit."""
    print(f"priority engine={engine} intra={intra} inter={inter} "
          f"width={width}", flush=True)
    args = problem(5, 4, 16, seed=95)
    ref = run_steps(args, FP64, ENGINE_LEVELS)
    sm, conn, omega, wt, st, ss, q = args
    with SweeperHIP(sm, conn, omega, wt, 16, st, ss, FP64) as s:
        s.set_engine(engine, SPIN)
        s.set_priority(intra, inter, width)
        inf = s.info()
        assert (inf["intra_angle_p"], inf["inter_angle_p"],
                inf["max_cells_per_step"]) == (intra, inter, width)
        s.set_source_isotropic(q)
        for _ in range(3):
            s.sweep_step()
        assert np.array_equal(s.phi(), ref[0])
        assert np.array_equal(s.psi_angle(omega.shape[0] - 1), ref[1])
        s.set_priority(-1)
        assert s.info()["intra_angle_p"] == -1


@pytest.mark.parametrize("engine", [ENGINE_LEVELS, ENGINE_QUEUE])
def test_priority_schedule_cyclic_and_fp32(engine):
    """Lagged faces (cyclic mesh) and FP32 storage under a priority
    schedule: the oracle's sweep, and the level sets bit for bit."""
    print(f"priority cyclic engine={engine}", flush=True)
    sm = meshmod.kuhn_cube(3, 0.2, seed=3)
    conn = meshmod.tycho_connectivity(sm)
    omega, wt = orc.quadrature(2)
    rng = np.random.default_rng(9)
    st = rng.uniform(1.0, 2.0, (sm.n_cells, 1))
    src = rng.uniform(0.1, 1.0, (sm.n_cells, omega.shape[0], 4, 1))
    prev = rng.uniform(0.0, 1.0, src.shape)
    psi = prev.copy()
    with SweeperHIP(sm, conn, omega, wt, 1, st, 0.5 * st, FP64) as s:
        s.set_engine(engine, SPIN)
        s.set_priority(3, 1, 64)
        assert s.info()["n_lagged"] > 0
        s.sweep(psi, src)
    ref, _ = orc.sweep(orc.build_mesh(sm), omega, 1, st, src, psi=prev.copy())
    assert np.abs(psi - ref).max() <= 1e-12 * np.abs(ref).max()
    args = problem(4, 4, 4, seed=96)
    ref32 = run_steps(args, FP32, ENGINE_LEVELS)
    sm, conn, omega, wt, st, ss, q = args
    with SweeperHIP(sm, conn, omega, wt, 4, st, ss, FP32) as s:
        s.set_engine(engine, SPIN)
        s.set_priority(1, 2, 300)
        s.set_source_isotropic(q)
        for _ in range(3):
            s.sweep_step()
        assert np.array_equal(s.phi(), ref32[0])


def test_level_engine_refuses_overlong_schedules():
    """A priority schedule of one item per step has more steps than the
    level engine launches per sweep: it refuses (before capturing a graph),
    for one handle and for a batch, and the ready queue runs the same
    schedule with the level sets' results."""
    from tycho2_amd.sweeper import SweepBatch
    print("overlong schedule", flush=True)
    args = problem(6, 4, 1, seed=97)
    ref = run_steps(args, FP64, ENGINE_LEVELS, steps=1)
    sm, conn, omega, wt, st, ss, q = args
    with SweeperHIP(sm, conn, omega, wt, 1, st, ss, FP64) as s:
        s.set_priority(1, 0, 1)
        assert s.info()["n_steps"] == sm.n_cells * omega.shape[0] > 24576
        s.set_source_isotropic(q)
        with pytest.raises(T2Error) as e:
            s.sweep_step()
        assert e.value.code == -1
        with SweepBatch([s]) as b:
            with pytest.raises(T2Error):
                b.sweep_only()
            b.set_engine(ENGINE_QUEUE, SPIN)
            s.set_phi(np.zeros_like(ref[0]))
            b.sweep_step()
            b.synchronize()
        assert np.array_equal(s.phi(), ref[0])
        s.set_engine(ENGINE_QUEUE, SPIN)
        s.set_phi(np.zeros_like(ref[0]))
        s.sweep_step()
        assert np.array_equal(s.phi(), ref[0])


def test_autotune_keeps_results_and_applies_best():
    """The tuner times every knob (engines, angle batches, launch shapes,
    chunk sizes, priority rules), leaves the fastest on the handle, and the
    handle's results afterwards equal the untuned ones bit for bit."""
    from tycho2_amd import autotune
    print("autotune", flush=True)
    args = problem(4, 4, 4, seed=98)
    ref = run_steps(args, FP64, ENGINE_LEVELS)
    sm, conn, omega, wt, st, ss, q = args
    with SweeperHIP(sm, conn, omega, wt, 4, st, ss, FP64) as s:
        best, trials = autotune.tune(s, reps=1, batches=(1, 2),
                                     blocks=(256, 64),
                                     priorities=[(1, 0, 256), (3, 1, 0)])
        assert best == min(trials, key=lambda t: t["ms"])
        engines = {t["engine"] for t in trials}
        assert engines == {ENGINE_LEVELS, ENGINE_QUEUE, ENGINE_HYBRID}
        assert any("priority" in t for t in trials)
        inf = s.info()
        assert inf["engine"] == best["engine"]
        assert inf["angle_batches"] == best["batches"]
        assert inf["groups_per_lane"] == best["groups_per_lane"]
        assert any(t["items_per_lane"] == 2 for t in trials)
        assert inf["concurrent_batches"] == int(best["concurrent"])
        want = best.get("priority", (-1, 0, 0))
        assert (inf["intra_angle_p"], inf["inter_angle_p"],
                inf["max_cells_per_step"]) == want
        s.set_source_isotropic(q)
        s.set_phi(np.zeros_like(ref[0]))
        for _ in range(3):
            s.sweep_step()
        assert np.array_equal(s.phi(), ref[0])
        assert np.array_equal(s.psi_angle(omega.shape[0] - 1), ref[1])


@pytest.mark.parametrize("precision,batches", [(FP64, 2), (FP32, 2),
                                               (FP64, 4)])
def test_concurrent_angle_batches_bitwise(precision, batches):
    """Angle batches swept side by side on parallel streams (forked and
    joined inside the captured graph) give the one-batch sweep's results,
    also with lagged faces and through a GMRES solve."""
    print(f"concurrent batches precision={precision} batches={batches}",
          flush=True)
    args = problem(5, 4, 16, seed=99)
    ref = run_steps(args, precision, ENGINE_LEVELS)
    sm, conn, omega, wt, st, ss, q = args
    with SweeperHIP(sm, conn, omega, wt, 16, st, ss, precision) as s:
        s.set_schedule(batches)
        s.set_concurrent_batches(True)
        assert s.info()["concurrent_batches"] == 1
        s.set_source_isotropic(q)
        for _ in range(3):
            s.sweep_step()
        assert np.array_equal(s.phi(), ref[0])
        assert np.array_equal(s.psi_angle(omega.shape[0] - 1), ref[1])
    # cyclic mesh: lagged faces read the pre-sweep psi in every branch
    sm = meshmod.kuhn_cube(3, 0.2, seed=3)
    conn = meshmod.tycho_connectivity(sm)
    omega, wt = orc.quadrature(2)
    rng = np.random.default_rng(10)
    st = rng.uniform(1.0, 2.0, (sm.n_cells, 1))
    src = rng.uniform(0.1, 1.0, (sm.n_cells, omega.shape[0], 4, 1))
    prev = rng.uniform(0.0, 1.0, src.shape)
    psi = prev.copy()
    with SweeperHIP(sm, conn, omega, wt, 1, st, 0.5 * st, FP64) as s:
        s.set_schedule(2)
        s.set_concurrent_batches(True)
        assert s.info()["n_lagged"] > 0
        s.sweep(psi, src)
    ref_psi, _ = orc.sweep(orc.build_mesh(sm), omega, 1, st, src,
                           psi=prev.copy())
    assert np.abs(psi - ref_psi).max() <= 1e-12 * np.abs(ref_psi).max()


def test_concurrent_batches_krylov_and_no_graph_bitwise(monkeypatch):
    """The same fork/join also runs uncaptured (direct launches), and a
    GMRES solve on concurrent batches equals the one-batch solve."""
    print("concurrent batches krylov", flush=True)
    sm, conn, omega, wt, st, ss, q = problem(4, 4, 4, seed=61)
    out = []
    for conc, no_graph in ((False, "0"), (True, "0"), (True, "1")):
        monkeypatch.setenv("T2_NO_GRAPH", no_graph)
        with SweeperHIP(sm, conn, omega, wt, 4, st, ss, FP64) as s:
            if conc:
                s.set_schedule(2)
                s.set_concurrent_batches(True)
            s.set_source_isotropic(q)
            hist = s.solve_krylov(200, 1e-10, 20)
            out.append((s.phi(), len(hist)))
    for phi, n in out[1:]:
        assert n == out[0][1]
        assert np.array_equal(phi, out[0][0])
