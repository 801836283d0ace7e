"""Parity of the HIP sweep engine against the CPU oracle (FP64 bar 1e-12).

All calls go through the C ABI (tycho2_amd/libtycho2_sweep.so); the oracle
only checks.
"""
import numpy as np
import pytest

import oracle as orc
from tycho2_amd import mesh as meshmod
from tycho2_amd import problems
from tycho2_amd.psi_io import rel_diff
from tycho2_amd.quadrature import product_quadrature
from tycho2_amd.sweeper import (FP32, FP64, GAUSS_CRAMER_GLU,
                                GAUSS_CRAMER_INTEL, GAUSS_ORIGINAL,
                                SweeperHIP, T2Error)

pytestmark = pytest.mark.gpu


def maxrel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def random_problem(n=3, sn=4, ng=4, seed=0, jitter=0.15):
    sm = meshmod.kuhn_cube(n, jitter, seed=seed)
    conn = meshmod.tycho_connectivity(sm)
    omega, wt = product_quadrature(sn)
    rng = np.random.default_rng(seed)
    st = rng.uniform(0.5, 3.0, (sm.n_cells, ng))
    ss = st * rng.uniform(0.1, 0.9, (sm.n_cells, ng))
    return sm, conn, omega, wt, st, ss, rng


@pytest.mark.parametrize("ng", [1, 2, 4, 8, 16, 32])
def test_single_sweep_fp64(ng):
    print(f"single sweep fp64 ng={ng}", flush=True)
    sm, conn, omega, wt, st, ss, rng = random_problem(ng=ng, seed=ng)
    src = rng.uniform(-0.5, 2.0, (sm.n_cells, omega.shape[0], 4, ng))
    psi = np.zeros_like(src)
    with SweeperHIP(sm, conn, omega, wt, ng, st, ss, FP64) as s:
        s.sweep(psi, src)
        assert s.info()["n_lagged"] == 0
    ref, lag = orc.sweep(orc.build_mesh(sm), omega, ng, st, src)
    assert lag == 0
    assert maxrel(psi, ref) <= 1e-12


@pytest.mark.parametrize("ng", [1, 16])
def test_single_sweep_fp32(ng):
    print(f"single sweep fp32 ng={ng}", flush=True)
    sm, conn, omega, wt, st, ss, rng = random_problem(ng=ng, seed=10 + ng)
    src = rng.uniform(0.1, 2.0, (sm.n_cells, omega.shape[0], 4, ng))
    psi = np.zeros_like(src)
    with SweeperHIP(sm, conn, omega, wt, ng, st, ss, FP32) as s:
        s.sweep(psi, src)
    ref, _ = orc.sweep(orc.build_mesh(sm), omega, ng, st, src)
    assert maxrel(psi, ref) <= 2e-5


@pytest.mark.parametrize("gauss,ng,precision", [
    (GAUSS_ORIGINAL, 4, FP64), (GAUSS_CRAMER_GLU, 16, FP64),
    (GAUSS_CRAMER_INTEL, 1, FP64), (GAUSS_ORIGINAL, 8, FP32),
    (GAUSS_CRAMER_GLU, 4, FP32)])
def test_cell_solvers_match_oracle(gauss, ng, precision):
    """The pivoting and Cramer cell solves against the oracle's same
    method (the reference's GaussElim options)."""
    print(f"cell solver {gauss} ng={ng} precision={precision}", flush=True)
    sm, conn, omega, wt, st, ss, rng = random_problem(ng=ng, seed=100 + ng)
    src = rng.uniform(-0.5, 2.0, (sm.n_cells, omega.shape[0], 4, ng))
    psi = np.zeros_like(src)
    with SweeperHIP(sm, conn, omega, wt, ng, st, ss, precision,
                    gauss=gauss) as s:
        with pytest.raises(T2Error):
            s.set_engine(1)
        if ng > 1:
            with pytest.raises(T2Error):
                s.set_launch(0, 2)
        s.sweep(psi, src)
    ref, _ = orc.sweep(orc.build_mesh(sm), omega, ng, st, src, method=gauss)
    assert maxrel(psi, ref) <= (1e-12 if precision == FP64 else 2e-5)


def test_launch_shapes_bitwise():
    print("launch shapes", flush=True)
    ng = 16
    sm, conn, omega, wt, st, ss, rng = random_problem(n=4, ng=ng, seed=3)
    src = rng.uniform(0.0, 1.0, (sm.n_cells, omega.shape[0], 4, ng))
    outs = []
    with SweeperHIP(sm, conn, omega, wt, ng, st, ss, FP64) as s:
        for ipb, gpl in [(0, 1), (4, 1), (8, 1), (16, 2), (32, 4), (64, 4)]:
            s.set_launch(ipb, gpl)
            psi = np.zeros_like(src)
            s.sweep(psi, src)
            outs.append(psi)
        with pytest.raises(Exception):
            s.set_launch(64, 1)          # 1024 threads: above the bound
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])


def test_gold_source_iteration(cube208, gold):
    """The reference's regression deck end to end on the device."""
    print("gold source iteration", flush=True)
    p = problems.gold(cube208)
    om = orc.build_mesh(cube208)
    src = orc.hat_source(om, p.omega, p.ng, 10.0, 5.0)
    with SweeperHIP.from_problem(p) as s:
        s.set_source_angular(src)
        hist = s.solve_source_iteration(100, 1e-10)
        psi = s.psi()
    rpsi, _, rhist = orc.source_iteration(om, p.omega, p.wt, p.ng, 10.0, 5.0,
                                          src, 100, 1e-10)
    assert len(hist) == len(rhist)
    assert maxrel(psi, rpsi) <= 1e-12
    # against the reference's own answer, with its test/diff.py metric
    got = psi[..., 0][:, gold["psi_angles"], :]
    l1, linf = rel_diff(got, gold["psi_g0"])
    assert l1 < 1e-10 and linf < 1e-10
    assert np.array_equal(psi[..., 0], psi[..., 1])
    assert abs(np.abs(psi).sum() / gold["psi_l1"] - 1) < 1e-10


def test_plumbing_source_iteration():
    """750 tets, 32 angles, 1 group, SI to 1e-12."""
    print("plumbing SI", flush=True)
    p = problems.plumbing()
    om = orc.build_mesh(p.smesh)
    src = orc.isotropic_source(p.nc, p.na, p.ng, np.ones((p.nc, p.ng)))
    with SweeperHIP.from_problem(p) as s:
        hist = s.solve_source_iteration(1000, 1e-12)
        phi = s.phi()
    _, rphi, rhist = orc.source_iteration(om, p.omega, p.wt, p.ng, 1.0, 0.5,
                                          src, 1000, 1e-12)
    assert len(hist) == len(rhist)
    assert maxrel(phi, rphi) <= 1e-12


def test_krylov_matches_source_iteration():
    print("krylov vs SI", flush=True)
    ng = 4
    sm, conn, omega, wt, st, ss, rng = random_problem(n=4, ng=ng, seed=5)
    q = np.ones((sm.n_cells, 4, ng))
    src = orc.isotropic_source(sm.n_cells, omega.shape[0], ng,
                               np.ones((sm.n_cells, ng)))
    _, rphi, _ = orc.source_iteration(orc.build_mesh(sm), omega, wt, ng, st,
                                      ss, src, 2000, 1e-14)
    for prec, tol in ((FP64, 1e-10), (FP32, 1e-5)):
        with SweeperHIP(sm, conn, omega, wt, ng, st, ss, prec) as s:
            s.set_source_isotropic(q)
            hist = s.solve_krylov(200, 1e-12, 20)
            phi = s.phi()
        assert len(hist) < 200
        assert maxrel(phi, rphi) <= tol


def test_sweep_step_matches_oracle():
    print("sweep step", flush=True)
    ng = 8
    sm, conn, omega, wt, st, ss, rng = random_problem(n=4, ng=ng, seed=9)
    q = rng.uniform(0.0, 1.0, (sm.n_cells, 4, ng))
    phi0 = rng.uniform(0.0, 1.0, (sm.n_cells, 4, ng))
    with SweeperHIP(sm, conn, omega, wt, ng, st, ss, FP64) as s:
        s.set_source_isotropic(q)
        s.set_phi(phi0)
        s.sweep_step()
        phi = s.phi()
    src = np.ascontiguousarray(np.broadcast_to(
        q[:, None], (sm.n_cells, omega.shape[0], 4, ng)))
    m = orc.build_mesh(sm)
    tot = src + (ss / (4 * np.pi))[:, None, None, :] * phi0[:, None]
    psi, _ = orc.sweep(m, omega, ng, st, tot)
    assert maxrel(phi, orc.psi_to_phi(psi, wt)) <= 1e-12


def test_full_size_plain_angle_planes():
    """384,000-tet, 128-angle, 16-group sweep; the oracle checks four angle
    planes (sweeps of different angles are independent)."""
    print("full-size plain sweep", flush=True)
    p = problems.plain()
    rng = np.random.default_rng(1)
    phi0 = rng.uniform(0.0, 1.0, (p.nc, 4, p.ng))
    angles = (0, 37, 90, 127)
    with SweeperHIP.from_problem(p) as s:
        s.set_phi(phi0)
        s.sweep_step()
        got = [s.psi_angle(a) for a in angles]
    m = orc.build_mesh(p.smesh)
    tot = (p.q_iso + (p.sigma_s / (4 * np.pi))[:, None, :] * phi0)
    for a, psi_a in zip(angles, got):
        ref, _ = orc.sweep(m, p.omega[a:a + 1], p.ng, p.sigma_t,
                           tot[:, None])
        assert maxrel(psi_a, ref[:, 0]) <= 1e-12


def test_schedule_batches_bitwise():
    """Sweeping the angles in batches changes the order, not the result."""
    print("schedule batches", flush=True)
    ng = 4
    sm, conn, omega, wt, st, ss, rng = random_problem(n=4, sn=4, ng=ng, seed=6)
    src = rng.uniform(0.0, 1.0, (sm.n_cells, omega.shape[0], 4, ng))
    outs = []
    with SweeperHIP(sm, conn, omega, wt, ng, st, ss, FP64) as s:
        for batches in (1, 2, 8, omega.shape[0]):
            s.set_schedule(batches)
            psi = np.zeros_like(src)
            s.sweep(psi, src)
            outs.append(psi)
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])


def cyclic_problem():
    # S2 directions run along the Kuhn diagonal faces: cycles appear
    sm = meshmod.kuhn_cube(3, 0.2, seed=3)
    conn = meshmod.tycho_connectivity(sm)
    omega, wt = orc.quadrature(2)
    rng = np.random.default_rng(8)
    st = rng.uniform(1.0, 2.0, (sm.n_cells, 1))
    return sm, conn, omega, wt, st, 0.5 * st, rng


def test_cyclic_mesh_lagged_sweep_matches_oracle():
    """Lagged faces read the previous psi in both implementations."""
    print("cyclic sweep", flush=True)
    sm, conn, omega, wt, st, ss, rng = cyclic_problem()
    src = rng.uniform(0.1, 1.0, (sm.n_cells, omega.shape[0], 4, 1))
    prev = rng.uniform(0.0, 1.0, src.shape)
    psi = prev.copy()
    with SweeperHIP(sm, conn, omega, wt, 1, st, ss, FP64) as s:
        n_lag = s.info()["n_lagged"]
        s.sweep(psi, src)
    ref, lag = orc.sweep(orc.build_mesh(sm), omega, 1, st, src,
                         psi=prev.copy())
    assert n_lag == lag > 0
    assert maxrel(psi, ref) <= 1e-12


def test_cyclic_mesh_source_iteration():
    print("cyclic SI", flush=True)
    sm, conn, omega, wt, st, ss, rng = cyclic_problem()
    q = np.ones((sm.n_cells, 4, 1))
    src = orc.isotropic_source(sm.n_cells, omega.shape[0], 1,
                               np.ones((sm.n_cells, 1)))
    with SweeperHIP(sm, conn, omega, wt, 1, st, ss, FP64) as s:
        s.set_source_isotropic(q)
        hist = s.solve_source_iteration(500, 1e-13)
        phi = s.phi()
    _, rphi, rhist = orc.source_iteration(orc.build_mesh(sm), omega, wt, 1,
                                          st, ss, src, 500, 1e-13)
    assert len(hist) == len(rhist)
    assert maxrel(phi, rphi) <= 1e-12


@pytest.mark.parametrize("ratio,bar", [(0.9, 1e-10), (0.0, 1e-12)])
def test_refined_solve_reaches_1e10(ratio, bar):
    """FP32 sweeps with FP32 psi inside FP64 defect correction: the scalar
    flux within 1e-10 (max-relative) of the FP64 oracle, where the FP32
    stage alone is not.  Without scattering one correction solves the
    exact defect equation L delta = r, so only FP32 rounding of the
    correction itself is left (~1e-14): a defect with a wrong term or scale
    would stay near the FP32 stage's 1e-7."""
    print(f"refined solve ratio={ratio}", flush=True)
    ng = 4
    sm = meshmod.kuhn_cube(6, 0.15, seed=21)
    conn = meshmod.tycho_connectivity(sm)
    omega, wt = product_quadrature(4)
    st = np.broadcast_to(1.0 + 0.1 * np.arange(ng), (sm.n_cells, ng))
    ss = ratio * st
    q = np.ones((sm.n_cells, 4, ng))
    src = orc.isotropic_source(sm.n_cells, omega.shape[0], ng,
                               np.ones((sm.n_cells, ng)))
    rpsi, rphi, _ = orc.krylov(orc.build_mesh(sm), omega, wt, ng, st, ss,
                               src, 400, 1e-14, 30)
    with SweeperHIP(sm, conn, omega, wt, ng, st, ss, FP32) as s:
        s.set_source_isotropic(q)
        info, phi1 = s.solve_refined(300, 1e-6, 1e-5, 20, want_phi1=True)
        phi = s.phi()
        psi = s.psi()
        psi_a = s.psi_angle(5)
        # a plain sweep afterwards replaces psi: FP32 accuracy again
        s.sweep_step()
        psi32 = s.psi()
    e1, e2 = maxrel(phi1, rphi), maxrel(phi, rphi)
    ep, ep32 = maxrel(psi, rpsi), maxrel(psi32, rpsi)
    print(f"refined: FP32 stage {e1:.2e}, after correction {e2:.2e}, "
          f"psi {ep:.2e} (a plain FP32 sweep: {ep32:.2e}), {info}")
    assert e2 <= bar
    assert e1 > 100 * e2
    # psi = psi32 + its FP32 correction meets the phi bar too; the two
    # accessors agree
    assert ep <= bar
    assert ep32 > 100 * ep
    assert np.array_equal(psi_a, psi[:, 5])
    # right-hand sides 2, psi32 twice, the correction 1, and one residual
    # apply per GMRES(20) restart
    restarts = (info["iters_sweep"] - 1) // 20 + \
        (info["iters_correction"] - 1) // 20
    assert info["sweeps"] == \
        info["iters_sweep"] + info["iters_correction"] + 5 + restarts


def test_iteration_limits_report_noconv():
    """Every solve that stops at its iteration limit above its tolerance
    returns T2_ERR_NOCONV with the last iterate in place (and the call's
    results on the exception); one that converges returns T2_OK."""
    from tycho2_amd.sweeper import NoConvergence, SweepBatch
    print("iteration limits", flush=True)
    ng = 4
    sm, conn, omega, wt, st, ss, rng = random_problem(n=3, ng=ng, seed=41)
    ss = 0.9 * st
    q = rng.uniform(0.5, 1.5, (sm.n_cells, 4, ng))
    with SweeperHIP(sm, conn, omega, wt, ng, st, ss, FP64) as s:
        s.set_source_isotropic(q)
        with pytest.raises(NoConvergence) as e:
            s.solve_source_iteration(3, 1e-14)
        assert e.value.code == -6 and len(e.value.result) == 3
        assert e.value.result[-1] > 1e-14
        with pytest.raises(NoConvergence) as e:
            s.solve_krylov(2, 1e-14, 20)
        assert len(e.value.result) == 2
        assert len(s.solve_krylov(200, 1e-12, 20)) < 200
    with SweeperHIP(sm, conn, omega, wt, ng, st, ss, FP32) as s:
        s.set_source_isotropic(q)
        # FP32 sweeps cannot take the first stage to 1e-12
        with pytest.raises(NoConvergence) as e:
            s.solve_refined(6, 1e-12, 1e-5, 20)
        info, _ = e.value.result
        assert info["iters_sweep"] == 6 and info["residual_sweep"] > 1e-12
    mem = [SweeperHIP(sm, conn, omega, wt, ng, st, ss, FP64) for _ in range(2)]
    for m in mem:
        m.set_source_isotropic(q)
    with SweepBatch(mem) as b:
        with pytest.raises(NoConvergence) as e:
            b.solve_krylov(2, 1e-14, 10)
        its, _ = e.value.result
        assert list(its) == [2, 2]
        with pytest.raises(NoConvergence):
            b.solve_source_iteration(2, 1e-14)
    for m in mem:
        m.close()


def batch_members(precision, groups=(4, 4, 4)):
    """Three problems of different mesh, quadrature and cross sections; the
    third has lagged faces."""
    specs = [(3, 4, 0.15, 31), (5, 2, 0.1, 32), (3, 2, 0.2, 3)]
    out = []
    for (n, sn, jit, seed), ng in zip(specs, groups):
        sm = meshmod.kuhn_cube(n, jit, seed=seed)
        conn = meshmod.tycho_connectivity(sm)
        omega, wt = (orc.quadrature(sn) if sn == 2 else
                     product_quadrature(sn))
        rng = np.random.default_rng(seed)
        st = rng.uniform(0.5, 3.0, (sm.n_cells, ng))
        ss = st * rng.uniform(0.1, 0.9, (sm.n_cells, ng))
        q = rng.uniform(0.0, 2.0, (sm.n_cells, 4, ng))
        s = SweeperHIP(sm, conn, omega, wt, ng, st, ss, precision)
        s.set_source_isotropic(q)
        out.append(s)
    return out


@pytest.mark.parametrize("precision,engine,groups", [
    (FP64, 0, (4, 4, 4)), (FP32, 0, (4, 4, 4)), (FP64, 0, (1, 16, 4)),
    (FP64, 1, (1, 16, 4)), (FP32, 1, (32, 2, 4)), (FP64, 1, (64, 8, 1)),
    (FP64, 2, (1, 16, 4)), (FP32, 2, (32, 2, 32)), (FP64, 3, (4, 4, 4)),
    (FP32, 3, (16, 4, 16))])
def test_batch_sweep_bitwise(precision, engine, groups):
    """A batch (merged per-step launches, or one persistent ready-queue
    launch for all members) gives every member what its own source
    iteration steps give, bit for bit."""
    from tycho2_amd.sweeper import SweepBatch
    print(f"batch sweep precision={precision} engine={engine} "
          f"groups={groups}", flush=True)
    own = batch_members(precision, groups)
    mem = batch_members(precision, groups)
    assert mem[2].info()["n_lagged"] > 0
    for s in own:
        for _ in range(3):
            s.sweep_step()
    with SweepBatch(mem) as b:
        # engine 2 / 3 here: the level form on parallel streams, one chain
        # per group count / up to two chains per group count
        b.set_engine(engine if engine < 2 else 0, 1 << 20)
        b.set_concurrent({2: 1, 3: 2}.get(engine, 0))
        inf = b.info()
        assert inf["n_steps"] == max(s.info()["n_steps"] for s in mem)
        for _ in range(3):
            b.sweep_step()
        b.synchronize()
    for s, m in zip(own, mem):
        assert np.array_equal(s.phi(), m.phi())
        assert np.array_equal(s.psi_angle(1), m.psi_angle(1))
    for s in own + mem:
        s.close()


@pytest.mark.parametrize("precision,engine,method", [
    (FP64, 0, "krylov"), (FP64, 1, "krylov"), (FP32, 1, "krylov"),
    (FP64, 1, "si"), (FP64, 0, "si"), (FP64, 2, "krylov"),
    (FP64, 2, "si")])
def test_batch_solves_bitwise(precision, engine, method):
    """Lock-step batched solves give every member its own solve's phi and
    psi bit for bit, and its own iteration count."""
    from tycho2_amd.sweeper import SweepBatch
    print(f"batch solve precision={precision} engine={engine} {method}",
          flush=True)
    groups = (4, 16, 1)
    own = batch_members(precision, groups)
    mem = batch_members(precision, groups)
    counts = []
    for s in own:
        if method == "krylov":
            counts.append(len(s.solve_krylov(100, 1e-10, 10)))
        else:
            counts.append(len(s.solve_source_iteration(300, 1e-10)))
    with SweepBatch(mem) as b:
        # engine 2 here: the level form, group counts on parallel streams
        b.set_engine(engine if engine < 2 else 0, 1 << 20)
        b.set_concurrent(2 if engine == 2 else 0)
        if method == "krylov":
            its, sec = b.solve_krylov(100, 1e-10, 10)
        else:
            its, sec = b.solve_source_iteration(300, 1e-10)
    print(f"  iterations {list(its)} seconds {[round(x, 3) for x in sec]}")
    assert list(its) == counts
    assert all(x > 0 for x in sec)
    for s, m in zip(own, mem):
        assert np.array_equal(s.phi(), m.phi())
        assert np.array_equal(s.psi_angle(2), m.psi_angle(2))
    for s in own + mem:
        s.close()


@pytest.mark.parametrize("engine", [0, 1])
def test_batch_refined_bitwise(engine):
    """The lock-step refined solve (each member's correction sweep carries
    its own FP32 residual source) gives every member its own refined solve's
    phi, psi and counts bit for bit; a later batch sweep carries no residual
    source."""
    from tycho2_amd.sweeper import SweepBatch
    print(f"batch refined engine={engine}", flush=True)
    groups = (4, 16, 1)
    own = batch_members(FP32, groups)
    mem = batch_members(FP32, groups)
    infos = [s.solve_refined(100, 1e-6, 1e-5, 10)[0] for s in own]
    with SweepBatch(mem) as b:
        b.set_engine(engine, 1 << 20)
        got, sec = b.solve_refined(100, 1e-6, 1e-5, 10)
        print(f"  {got}", flush=True)
        assert got == infos
        assert all(x > 0 for x in sec)
        for s, m in zip(own, mem):
            assert np.array_equal(s.phi(), m.phi())
            assert np.array_equal(s.psi_angle(2), m.psi_angle(2))
        for s in own:
            s.sweep_step()
        b.sweep_step()
        b.synchronize()
    for s, m in zip(own, mem):
        assert np.array_equal(s.phi(), m.phi())
    for s in own + mem:
        s.close()


def test_batch_refined_needs_fp32():
    from tycho2_amd.sweeper import SweepBatch, T2Error
    mem = batch_members(FP64)
    with SweepBatch(mem) as b:
        with pytest.raises(T2Error):
            b.solve_refined(10)
    for s in mem:
        s.close()


def test_batch_rejects_changed_member():
    from tycho2_amd.sweeper import SweepBatch, T2Error
    print("batch stale member", flush=True)
    mem = batch_members(FP64)
    b = SweepBatch(mem)
    b.sweep_only()
    mem[0].set_schedule(2)
    with pytest.raises(T2Error):
        b.sweep_only()
    b.close()
    for s in mem:
        s.close()


@pytest.mark.parametrize("engine", [0, 1])
def test_batch_priority_members_bitwise(engine):
    """Members on priority schedules (the reference's b-level and DFDS
    rules, two widths) beside a member on level sets: every member's batched
    steps equal its own, bit for bit."""
    from tycho2_amd.sweeper import SweepBatch
    print(f"batch priority members engine={engine}", flush=True)
    groups = (4, 16, 1)
    rules = [(1, 0, 256), (3, 1, 1000), (-1, 0, 0)]
    own = batch_members(FP64, groups)
    mem = batch_members(FP64, groups)
    for s, m, r in zip(own, mem, rules):
        s.set_priority(*r)
        m.set_priority(*r)
        for _ in range(2):
            s.sweep_step()
    with SweepBatch(mem) as b:
        b.set_engine(engine, 1 << 20)
        for _ in range(2):
            b.sweep_step()
        b.synchronize()
    for s, m in zip(own, mem):
        assert np.array_equal(s.phi(), m.phi())
        assert np.array_equal(s.psi_angle(1), m.psi_angle(1))
    for s in own + mem:
        s.close()


@pytest.mark.parametrize("precision", [FP64, FP32])
def test_batch_block_sizes_bitwise(precision):
    """The level form's merged launches give the members' own sweeps bit
    for bit at 64, 128 and 256 threads per block, with the group counts
    interleaved and on parallel streams."""
    from tycho2_amd.sweeper import SweepBatch
    print(f"batch block sizes precision={precision}", flush=True)
    groups = (4, 16, 1)
    own = batch_members(precision, groups)
    for s in own:
        s.sweep_step()
    want = [s.phi() for s in own]
    for block, chains in [(64, 0), (128, 1), (256, 0)]:
        mem = batch_members(precision, groups)
        with SweepBatch(mem) as b:
            b.set_engine(0, 1 << 20)
            b.set_concurrent(chains)
            b.set_launch(block)
            b.sweep_step()
            b.synchronize()
        for w, m in zip(want, mem):
            assert np.array_equal(w, m.phi())
        for m in mem:
            m.close()
    for s in own:
        s.close()


def test_batch_autotune_keeps_results():
    """The batch tuner times both engines, the level form's chain counts
    and block sizes and the queue's chunk sizes, leaves the fastest on the
    batch, and the batch's steps afterwards equal the members' own bit for
    bit (whatever block size it kept)."""
    from tycho2_amd import autotune
    from tycho2_amd.sweeper import SweepBatch
    print("batch autotune", flush=True)
    groups = (4, 16, 1)
    own = batch_members(FP64, groups)
    mem = batch_members(FP64, groups)
    for s in own:
        s.sweep_step()
    with SweepBatch(mem) as b:
        best, trials = autotune.tune_batch(b, reps=1)
        assert best == min(trials, key=lambda t: t["ms"])
        assert {t["engine"] for t in trials} == {0, 1}
        assert len(trials) == len(autotune.CHAINS) + len(
            autotune.CHUNK_ROUNDS) + len(autotune.BLOCKS) - 1
        assert {t["concurrent"] for t in trials if t["engine"] == 0} == \
            set(autotune.CHAINS)
        assert {t["block"] for t in trials if t["engine"] == 0} == \
            set(autotune.BLOCKS)
        for m in mem:
            m.set_phi(np.zeros_like(m.phi()))
        b.sweep_step()
        b.synchronize()
    for s, m in zip(own, mem):
        assert np.array_equal(s.phi(), m.phi())
    for s in own + mem:
        s.close()


def test_refined_on_concurrent_paths_bitwise():
    """The refined solve on a handle with two concurrent angle batches, and
    the lock-step refined solve on a batch with two chains per group count,
    equal the plain one-stream solves bit for bit."""
    from tycho2_amd.sweeper import SweepBatch
    print("refined on concurrent paths", flush=True)
    groups = (4, 16, 1)
    own = batch_members(FP32, groups)
    ref = [s.solve_refined(100, 1e-6, 1e-5, 10)[0] for s in own]
    phis = [s.phi() for s in own]
    conc = batch_members(FP32, groups)
    for s, info, phi in zip(conc, ref, phis):
        s.set_schedule(2)
        s.set_concurrent_batches(True)
        got, _ = s.solve_refined(100, 1e-6, 1e-5, 10)
        assert got == info
        assert np.array_equal(s.phi(), phi)
    mem = batch_members(FP32, groups)
    with SweepBatch(mem) as b:
        b.set_concurrent(2)
        infos, _ = b.solve_refined(100, 1e-6, 1e-5, 10)
    assert infos == ref
    for m, phi in zip(mem, phis):
        assert np.array_equal(m.phi(), phi)
    for s in own + conc + mem:
        s.close()


@pytest.mark.parametrize("engine", [0, 1])
def test_dependency_tolerance_matches_oracle(engine):
    """The precomputed schedule's rule (faces with |Omega.n| <= 1e-10 are
    not dependencies and read the pre-sweep psi; SweepType
    OriginalTycho1/2) on an unjittered Kuhn cube with S2 directions along
    its diagonal faces: the lagged faces, one sweep and a source iteration
    equal the oracle's under the same rule, on the level engine and the
    ready queue."""
    from tycho2_amd.sweeper import ENGINE_QUEUE
    print(f"dependency tolerance engine={engine}", flush=True)
    sm = meshmod.kuhn_cube(3, 0.0)
    conn = meshmod.tycho_connectivity(sm)
    omega, wt = orc.quadrature(2)
    ng = 4
    rng = np.random.default_rng(13)
    st = rng.uniform(1.0, 2.0, (sm.n_cells, ng))
    ss = 0.5 * st
    src = rng.uniform(0.1, 1.0, (sm.n_cells, omega.shape[0], 4, ng))
    prev = rng.uniform(0.0, 1.0, src.shape)
    m = orc.build_mesh(sm)
    with orc.dependency_tolerance(1e-10):
        ref, lag = orc.sweep(m, omega, ng, st, src, psi=prev.copy())
        q1 = orc.isotropic_source(sm.n_cells, omega.shape[0], ng,
                                  np.ones((sm.n_cells, ng)))
        _, rphi, rhist = orc.source_iteration(m, omega, wt, ng, st, ss, q1,
                                              500, 1e-13)
    with SweeperHIP(sm, conn, omega, wt, ng, st, ss, FP64) as s:
        n0 = s.info()["n_lagged"]
        s.set_dependency_tolerance(1e-10)
        if engine:
            s.set_engine(ENGINE_QUEUE)
        assert s.info()["n_lagged"] == lag > n0
        psi = prev.copy()
        s.sweep(psi, src)
        assert maxrel(psi, ref) <= 1e-12
        s.set_source_isotropic(np.ones((sm.n_cells, 4, ng)))
        hist = s.solve_source_iteration(500, 1e-13)
        assert len(hist) == len(rhist)
        assert maxrel(s.phi(), rphi) <= 1e-12
        # back to the traverser's rule
        s.set_dependency_tolerance(-1)
        assert s.info()["n_lagged"] == n0
