"""Traversal priorities and the priority list schedule -- TEST
INFRASTRUCTURE ONLY (importable from ``tests/`` alone).

A pure-Python restatement, for small meshes, of the reference's
``Priorities::calcPriorities`` on one rank (src/Priorities.cc:396-435) and
of the order in which ``GraphTraverser::traverse`` pops ready (cell, angle)
pairs from its priority heap (src/GraphTraverser.cc:467-644), cut into
steps of at most ``max_cells_per_step`` pairs that are all ready when the
step starts.

Dependencies follow the reference: face k of cell c is outgoing for
direction Omega when Omega . n_k > 0 (src/TychoMesh.hh:107-108); then the
neighbour across it depends on c.  Like the reference, this restatement
needs an acyclic dependency graph (it raises on a cycle).

Unsigned 64-bit arithmetic is reproduced with Python integers masked to 64
bits.  The heap order among equal priorities is fixed here as lower angle,
then lower cell (std::priority_queue leaves it to the implementation).
"""
from __future__ import annotations

import heapq

import numpy as np

MASK = (1 << 64) - 1


def _omega_dot_n(om, n):
    # products and sums rounded separately, in the reference's order
    return (om[0] * n[..., 0] + om[1] * n[..., 1]) + om[2] * n[..., 2]


def downwind(adj, normal, om):
    """down[c] = the cells that depend on c for direction om."""
    adj = np.asarray(adj)
    out = np.asarray(_omega_dot_n(om, np.asarray(normal))) > 0
    return [[int(adj[c, k]) for k in range(4) if adj[c, k] >= 0 and out[c, k]]
            for c in range(adj.shape[0])]


def _reverse_topological(down):
    """Cells ordered so that every cell follows all of its downwind cells
    (the backward traversal, src/Priorities.cc:380-384)."""
    nc = len(down)
    n_children = [len(d) for d in down]
    parents = [[] for _ in range(nc)]
    for c, d in enumerate(down):
        for u in d:
            parents[u].append(c)
    ready = [c for c in range(nc) if n_children[c] == 0]
    order = []
    while ready:
        c = ready.pop()
        order.append(c)
        for p in parents[c]:
            n_children[p] -= 1
            if n_children[p] == 0:
                ready.append(p)
    if len(order) != nc:
        raise ValueError("dependency cycle: the reference never finishes")
    return order


def blevels(adj, normal, omega):
    """b-level of every (angle, cell), blev[a][c]: 0 for a cell with no
    downwind dependant, else 1 + the largest b-level downwind
    (BLevelData::update, src/Priorities.cc:120-143)."""
    na, nc = len(omega), len(adj)
    blev = [[0] * nc for _ in range(na)]
    for a in range(na):
        down = downwind(adj, normal, omega[a])
        for c in _reverse_topological(down):
            blev[a][c] = max([blev[a][u] + 1 for u in down[c]], default=0)
    return blev


def neighbor_priorities(adj, normal, omega, parent_shift):
    """NeighborPriorityData::update (src/Priorities.cc:220-242) on one rank:
    no face is a partition side, so the bound terms never apply and
    prio(c) = max(0, max over downwind u of prio(u) + parent_shift), with
    the shift added in unsigned 64-bit arithmetic."""
    na, nc = len(omega), len(adj)
    shift = parent_shift & MASK
    prio = [[0] * nc for _ in range(na)]
    for a in range(na):
        down = downwind(adj, normal, omega[a])
        for c in _reverse_topological(down):
            v = 0
            for u in down[c]:
                v = max(v, (prio[a][u] + shift) & MASK)
            prio[a][c] = v
    return prio


def random_priorities(na, nc):
    """A 31-bit draw per (angle, cell), angle-outer, from MT19937 seeded
    with 0 (numpy's legacy RandomState seeds it as std::mt19937(0) does);
    the reference draws rand() after srand(0) (src/Priorities.cc:283-291)."""
    raw = np.random.RandomState(0).randint(0, 2**32, size=na * nc,
                                           dtype=np.uint32)
    raw = (raw >> 1).astype(np.uint64).reshape(na, nc)
    return [[int(x) for x in row] for row in raw]


def angle_priorities(prio, inter, nlevels):
    """anglePriorities (src/Priorities.cc:328-375), in place."""
    na, nc = len(prio), len(prio[0])
    span = (nlevels * nc) & MASK
    if inter == 1:
        for a in range(na):
            for c in range(nc):
                prio[a][c] = (prio[a][c] + a * span) & MASK
    elif inter == 2:
        ranked = sorted((max(prio[a]), a) for a in range(na))
        for order, (_, a) in enumerate(reversed(ranked), start=1):
            for c in range(nc):
                prio[a][c] = (prio[a][c] + (na - order) * span) & MASK
    return prio


def calc_priorities(adj, normal, omega, intra, inter):
    """Priorities::calcPriorities (src/Priorities.cc:396-435): prio[a][c]."""
    blev = blevels(adj, normal, omega)
    nlevels = max(max(row) for row in blev)
    if intra == 0:
        prio = random_priorities(len(omega), len(adj))
    elif intra == 1:
        prio = [list(row) for row in blev]
    elif intra == 2:
        prio = neighbor_priorities(adj, normal, omega, 0)
    else:  # 3 DFDS, 4 DFHDS: the same on one rank
        prio = neighbor_priorities(adj, normal, omega, -1)
    return angle_priorities(prio, inter, nlevels)


def list_schedule(adj, normal, omega, prio, max_cells_per_step):
    """(items, step_off): pairs popped from a max-heap of ready pairs, each
    step taking up to max_cells_per_step (0: all) of those ready at its
    start; items as angle * n_cells + cell."""
    na, nc = len(omega), len(adj)
    downs = [downwind(adj, normal, omega[a]) for a in range(na)]
    indeg = [[0] * nc for _ in range(na)]
    for a in range(na):
        for c in range(nc):
            for u in downs[a][c]:
                indeg[a][u] += 1
    heap = [(-prio[a][c], a, c) for a in range(na) for c in range(nc)
            if indeg[a][c] == 0]
    heapq.heapify(heap)
    items, off = [], [0]
    while heap:
        step = []
        while heap and (max_cells_per_step == 0 or
                        len(step) < max_cells_per_step):
            step.append(heapq.heappop(heap))
        for _, a, c in step:
            items.append(a * nc + c)
        off.append(len(items))
        for _, a, c in step:
            for u in downs[a][c]:
                indeg[a][u] -= 1
                if indeg[a][u] == 0:
                    heapq.heappush(heap, (-prio[a][u], a, u))
    if len(items) != na * nc:
        raise ValueError("dependency cycle: the reference never finishes")
    return np.array(items, dtype=np.int64), np.array(off, dtype=np.int64)
