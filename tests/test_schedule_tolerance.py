"""The precomputed schedule's dependency rule (SweepType OriginalTycho1/2):
faces with |Omega.n| <= 1e-10 are not sweep dependencies
(src/SweepSchedule.cc:57,108-125) and read the psi their upwind cell held
before the sweep.

Pinned on an unjittered Kuhn cube with S2 directions, which run along the
diagonal faces (|Omega.n| ~ 1e-16 there, exactly 0 on some): the number of
faces the rule lags is counted directly from the geometry, engine and oracle
lag the same faces, and source iteration with the rule converges to the
solution of the fully coupled DG system of every direction (a dense global
solve, tests/_dg.py), so the lagging changes the iteration, not the answer.
"""
import numpy as np
import pytest

import oracle as orc
from tycho2_amd import mesh as meshmod
from tycho2_amd import sweeper

from _dg import global_sweep_solution

TOL = 1e-10


def s2_cube(n=3):
    sm = meshmod.kuhn_cube(n, 0.0)
    m = orc.build_mesh(sm)
    om, w = orc.quadrature(2)
    return sm, m, om, w


def test_rule_lags_exactly_the_near_parallel_faces():
    sm, m, om, w = s2_cube()
    dots = np.einsum("ad,ckd->ack", om, m.normal)
    interior = (m.adj >= 0)[None]
    want = int((interior & (dots <= 0) & (dots >= -TOL)).sum())
    assert want > 0
    # the near-parallel faces are round-off: nothing between 1e-15 and 0.1
    near = np.abs(dots[np.broadcast_to(interior, dots.shape)])
    assert near[near < 0.1].max() < 1e-15
    src = orc.isotropic_source(m.nc, om.shape[0], 1, np.ones((m.nc, 1)))
    st = np.ones((m.nc, 1))
    with orc.dependency_tolerance(TOL):
        _, lag = orc.sweep(m, om, 1, st, src)
    _, n_lag = sweeper.sweep_levels(m.adj, m.area, m.normal, om,
                                    dep_tol=TOL)
    assert lag == n_lag == want
    # the graph traverser's rule lags fewer faces (only to break cycles)
    _, lag0 = orc.sweep(m, om, 1, st, src)
    _, n0 = sweeper.sweep_levels(m.adj, m.area, m.normal, om)
    assert lag0 == n0 < want


def test_levels_respect_remaining_dependencies():
    """With the rule, every face with Omega.n < -tol still orders its two
    cells: the upwind cell sits on a lower level."""
    sm, m, om, w = s2_cube(4)
    levels, _ = sweeper.sweep_levels(m.adj, m.area, m.normal, om,
                                     dep_tol=TOL)
    dots = np.einsum("ad,ckd->ack", om, m.normal)
    for a in range(om.shape[0]):
        for c in range(m.nc):
            for k in range(4):
                u = m.adj[c, k]
                if u >= 0 and dots[a, c, k] < -TOL:
                    assert levels[a, u] < levels[a, c]


@pytest.mark.parametrize("tol", [TOL, 1e-3])
def test_source_iteration_with_rule_converges_to_coupled_solution(tol):
    """SI with lagged near-parallel faces reaches the fully coupled
    solution of every direction.  tol = 1e-3 lags no further face here (the
    next |Omega.n| is 0.57), so both tolerances give one answer."""
    sm, m, om, w = s2_cube()
    ng = 1
    rng = np.random.default_rng(9)
    st = rng.uniform(1.0, 2.0, (m.nc, ng))
    ss = 0.5 * st
    src = orc.isotropic_source(m.nc, om.shape[0], ng, np.ones((m.nc, ng)))
    with orc.dependency_tolerance(tol):
        psi, phi, hist = orc.source_iteration(m, om, w, ng, st, ss, src,
                                              2000, 1e-15)
    assert hist[-1] <= 1e-15
    tot = src + (ss / (4 * np.pi))[:, None, None, :] * phi[:, None]
    for a in range(om.shape[0]):
        want = global_sweep_solution(sm.coords, sm.cell_nodes, m.adj, om[a],
                                     st[:, 0], tot[:, a, :, 0])
        assert np.allclose(psi[:, a, :, 0], want, rtol=1e-12,
                           atol=1e-13 * np.abs(want).max())


def test_rule_leaves_a_single_sweep_within_the_coupling_size():
    """One sweep from psi = 0 with the rule differs from the traverser's
    sweep only through the lagged faces' coupling, A |Omega.n| <= ~1e-16
    here: far below 1e-12."""
    sm, m, om, w = s2_cube()
    rng = np.random.default_rng(10)
    st = rng.uniform(1.0, 2.0, (m.nc, 1))
    src = rng.uniform(0.0, 1.0, (m.nc, om.shape[0], 4, 1))
    ref, _ = orc.sweep(m, om, 1, st, src)
    with orc.dependency_tolerance(TOL):
        got, _ = orc.sweep(m, om, 1, st, src)
    assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max()
