"""Traversal priorities and the priority list schedule on the host: the
oracle's restatement of the reference's Priorities::calcPriorities is pinned
by facts it does not compute the same way (b-levels are the level sets of
the reversed directions; DFDS on one rank is a breadth-first distance; the
random rule is MT19937 seeded with 0), and the library's host code must
equal the oracle exactly; every schedule it builds is a valid list
schedule."""
import numpy as np
import pytest

import priority_oracle as opr
from tycho2_amd import mesh as meshmod
from tycho2_amd import sweeper
from tycho2_amd.quadrature import product_quadrature

MASK = (1 << 64) - 1


def geometry(n, jitter, seed):
    sm = meshmod.kuhn_cube(n, jitter, seed=seed)
    con = meshmod.tycho_connectivity(sm)
    _, area, nrm = sweeper.mesh_geometry(sm, con)
    return con.adj_cell, area, nrm


@pytest.fixture(scope="module")
def acyclic():
    """A jittered mesh and directions with no dependency cycle (the
    reference's traversal only finishes on those)."""
    adj, area, nrm = geometry(3, 0.1, 4)
    omega, _ = product_quadrature(4)
    for om in (omega, -omega):
        _, lag = sweeper.sweep_levels(adj, area, nrm, om)
        assert lag == 0
    return adj, area, nrm, omega


def test_blevels_are_levels_of_the_reversed_directions(acyclic):
    adj, area, nrm, omega = acyclic
    blev = np.array(opr.blevels(adj, nrm, omega))
    rev, _ = sweeper.sweep_levels(adj, area, nrm, -omega)
    assert np.array_equal(blev, rev)
    assert blev.max() > 3


def test_dfds_is_minus_breadth_first_distance_to_a_leaf(acyclic):
    """One rank, parent shift -1 in uint64: a cell with no dependant gets 0,
    any other 2^64 - (fewest steps to one)."""
    adj, _, nrm, omega = acyclic
    p3 = opr.calc_priorities(adj, nrm, omega, 3, 0)
    p4 = opr.calc_priorities(adj, nrm, omega, 4, 0)
    assert p3 == p4
    for a in range(len(omega)):
        down = opr.downwind(adj, nrm, omega[a])
        dist = {c: 0 for c in range(len(adj)) if not down[c]}
        frontier = list(dist)
        parents = [[] for _ in range(len(adj))]
        for c, d in enumerate(down):
            for u in d:
                parents[u].append(c)
        while frontier:        # breadth-first search from the leaves
            nxt = []
            for u in frontier:
                for c in parents[u]:
                    if c not in dist:
                        dist[c] = dist[u] + 1
                        nxt.append(c)
            frontier = nxt
        for c in range(len(adj)):
            want = 0 if dist[c] == 0 else (1 << 64) - dist[c]
            assert p3[a][c] == want


def test_bfds_is_zero_and_random_is_mt19937(acyclic):
    adj, _, nrm, omega = acyclic
    assert all(v == 0 for row in opr.calc_priorities(adj, nrm, omega, 2, 0)
               for v in row)
    r = opr.random_priorities(2, 3)
    # first outputs of std::mt19937 seeded with 0
    assert [r[0][0], r[0][1], r[0][2]] == [2357136044 >> 1,
                                           2546248239 >> 1,
                                           3071714933 >> 1]


def test_angle_rules(acyclic):
    adj, _, nrm, omega = acyclic
    nc, na = len(adj), len(omega)
    base = opr.calc_priorities(adj, nrm, omega, 1, 0)
    nlev = max(max(r) for r in base)
    glob = opr.calc_priorities(adj, nrm, omega, 1, 1)
    for a in range(na):
        assert all(glob[a][c] - base[a][c] == a * nlev * nc
                   for c in range(nc))
    loc = opr.calc_priorities(adj, nrm, omega, 1, 2)
    shifts = [loc[a][0] - base[a][0] for a in range(na)]
    assert sorted(shifts) == [k * nlev * nc for k in range(na)]
    # the angle with the longest critical path gets the top shift
    top = max(range(na), key=lambda a: (max(base[a]), a))
    assert shifts[top] == (na - 1) * nlev * nc


@pytest.mark.parametrize("intra", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("inter", [0, 1, 2])
def test_library_priorities_equal_oracle(acyclic, intra, inter):
    adj, area, nrm, omega = acyclic
    got = sweeper.sweep_priorities(adj, area, nrm, omega, intra, inter)
    want = opr.calc_priorities(adj, nrm, omega, intra, inter)
    assert got.dtype == np.uint64
    assert [[int(x) for x in row] for row in got] == want


def check_list_schedule(adj, nrm, omega, items, off, width):
    """Every pair once; every step within the width; every upwind pair of a
    pair in an earlier step."""
    nc, na = len(adj), len(omega)
    assert sorted(items.tolist()) == list(range(nc * na))
    step = np.empty(nc * na, np.int64)
    for s in range(len(off) - 1):
        assert 0 < off[s + 1] - off[s] <= (width or nc * na)
        step[items[off[s]:off[s + 1]]] = s
    for a in range(na):
        down = opr.downwind(adj, nrm, omega[a])
        for c in range(nc):
            for u in down[c]:
                assert step[a * nc + c] < step[a * nc + u]


@pytest.mark.parametrize("intra,inter,width", [
    (1, 0, 0), (1, 1, 7), (3, 1, 100), (4, 2, 16), (2, 0, 33), (0, 2, 50),
    (3, 0, 1)])
def test_library_schedule_equals_oracle(acyclic, intra, inter, width):
    adj, area, nrm, omega = acyclic
    items, off = sweeper.priority_schedule(adj, area, nrm, omega, intra,
                                           inter, width)
    prio = opr.calc_priorities(adj, nrm, omega, intra, inter)
    oi, oo = opr.list_schedule(adj, nrm, omega, prio, width)
    assert np.array_equal(items, oi)
    assert np.array_equal(off, oo)
    check_list_schedule(adj, nrm, omega, items, off, width)


def test_unlimited_width_gives_level_sets(acyclic):
    """With no step limit every step holds exactly the pairs of one level."""
    adj, area, nrm, omega = acyclic
    items, off = sweeper.priority_schedule(adj, area, nrm, omega, 3, 1, 0)
    levels, _ = sweeper.sweep_levels(adj, area, nrm, omega)
    nc = len(adj)
    for s in range(len(off) - 1):
        it = items[off[s]:off[s + 1]]
        assert (levels[it // nc, it % nc] == s).all()
    assert len(off) - 1 == levels.max() + 1


def test_face_order_keeps_each_step(acyclic):
    adj, area, nrm, omega = acyclic
    a, ao = sweeper.priority_schedule(adj, area, nrm, omega, 1, 1, 40)
    b, bo = sweeper.priority_schedule(adj, area, nrm, omega, 1, 1, 40,
                                      order=sweeper.ORDER_FACES)
    assert np.array_equal(ao, bo)
    for s in range(len(ao) - 1):
        assert sorted(a[ao[s]:ao[s + 1]]) == sorted(b[bo[s]:bo[s + 1]])


def test_cyclic_mesh_schedule_places_every_pair():
    """Dependency cycles: the lagged faces drop out of the graph; the
    schedule still holds every pair once and follows the level sets'
    dependencies."""
    adj, area, nrm = geometry(3, 0.2, 3)
    import oracle as orc
    omega, _ = orc.quadrature(2)
    levels, lag = sweeper.sweep_levels(adj, area, nrm, omega)
    assert lag > 0
    nc, na = len(adj), len(omega)
    for intra in (1, 3):
        items, off = sweeper.priority_schedule(adj, area, nrm, omega, intra,
                                               1, 10)
        assert sorted(items.tolist()) == list(range(nc * na))
        assert (np.diff(off) <= 10).all()


def test_bad_rules_rejected(acyclic):
    adj, area, nrm, omega = acyclic
    with pytest.raises(sweeper.T2Error):
        sweeper.sweep_priorities(adj, area, nrm, omega, 5, 0)
    with pytest.raises(sweeper.T2Error):
        sweeper.priority_schedule(adj, area, nrm, omega, 1, 3, 0)
    with pytest.raises(sweeper.T2Error):
        sweeper.priority_schedule(adj, area, nrm, omega, 1, 0, -1)
