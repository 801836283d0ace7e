"""The C-ABI library on a host without a GPU: it loads, exports every
symbol of include/tycho2_sweep.h, its host helpers agree with the oracle,
and its entry points reject bad arguments with status codes."""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle as orc
from tycho2_amd import mesh as meshmod
from tycho2_amd import sweeper
from tycho2_amd.quadrature import product_quadrature

from conftest import ROOT


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"\b(t2_\w+)\s*\(", text))
    return sorted(names)


def test_every_declared_symbol_is_exported_and_bound():
    lib = sweeper.lib()
    names = declared_symbols()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
        assert name in sweeper.SIGNATURES, name
    assert lib.t2_abi_version() == 3


def test_host_geometry_bitwise_equal_to_oracle(cube208):
    for sm in (cube208, meshmod.kuhn_cube(5, 0.15, seed=9)):
        con = meshmod.tycho_connectivity(sm)
        vol, area, nrm = sweeper.mesh_geometry(sm, con)
        m = orc.build_mesh(sm)
        assert np.array_equal(vol, m.volume)
        assert np.array_equal(area, m.area)
        assert np.array_equal(nrm, m.normal)


def test_levels_are_a_topological_order(cube208):
    sm = cube208
    con = meshmod.tycho_connectivity(sm)
    _, area, nrm = sweeper.mesh_geometry(sm, con)
    omega, _ = product_quadrature(8)
    levels, lag = sweeper.sweep_levels(con.adj_cell, area, nrm, omega)
    assert lag == 0 and levels.min() == 0
    adj = con.adj_cell
    for a in range(omega.shape[0]):
        odn = nrm @ omega[a]                     # (nc, 4)
        c, k = np.nonzero((adj >= 0) & ~(odn > 0))
        assert (levels[a, adj[c, k]] < levels[a, c]).all()
        # every cell above level 0 waits on an upwind cell one level below
        c0 = np.nonzero(levels[a] > 0)[0]
        for cc in c0:
            ups = [adj[cc, kk] for kk in range(4)
                   if adj[cc, kk] >= 0 and not odn[cc, kk] > 0]
            assert max(levels[a, u] for u in ups) == levels[a, cc] - 1


@pytest.mark.parametrize("n,sn,jit,seed", [(3, 2, 0.2, 3), (2, 6, 0.1, 0),
                                           (4, 4, 0.15, 1)])
def test_lagged_faces_agree_with_oracle(n, sn, jit, seed):
    sm = meshmod.kuhn_cube(n, jit, seed=seed)
    con = meshmod.tycho_connectivity(sm)
    _, area, nrm = sweeper.mesh_geometry(sm, con)
    om, _ = orc.quadrature(sn)
    _, lag = sweeper.sweep_levels(con.adj_cell, area, nrm, om)
    m = orc.build_mesh(sm)
    _, olag = orc.sweep(m, om, 1, 1.0, np.ones((m.nc, om.shape[0], 4, 1)))
    assert lag == olag


def test_full_size_schedule_statistics():
    """Level sets of the 1,296,000-tet, 128-angle workload: acyclic for its
    seed, every (angle, cell) placed once."""
    from tycho2_amd import problems
    p = problems.precision()
    _, area, nrm = sweeper.mesh_geometry(p.smesh, p.conn)
    levels, lag = sweeper.sweep_levels(p.conn.adj_cell, area, nrm, p.omega)
    assert lag == 0
    assert levels.shape == (128, p.nc)
    per_angle = levels.max(axis=1) + 1
    assert 200 < per_angle.min() <= per_angle.max() < 1200


def _create(**kw):
    sm = meshmod.kuhn_cube(1)
    con = meshmod.tycho_connectivity(sm)
    om, wt = product_quadrature(2)
    args = dict(ng=1, precision=0, gauss=1)
    args.update(kw)
    ng = args["ng"]
    h = ctypes.c_void_p()
    f = lambda a, t: np.ascontiguousarray(a, dtype=t)
    xyz, cn = f(sm.coords, np.float64), f(sm.cell_nodes, np.int64)
    adj, f2c = f(con.adj_cell, np.int64), f(con.f2c, np.int64)
    st = np.ones((sm.n_cells, max(ng, 1)))
    rc = sweeper.lib().t2_sweeper_create(
        ctypes.byref(h), 0, sm.n_cells, sm.n_nodes, sweeper._ptr(xyz),
        sweeper._ptr(cn, sweeper._ip), sweeper._ptr(adj, sweeper._ip),
        sweeper._ptr(f2c, sweeper._ip), om.shape[0], sweeper._ptr(om),
        sweeper._ptr(wt), ng, sweeper._ptr(st), sweeper._ptr(st),
        args["precision"], args["gauss"])
    return rc, sweeper.lib().t2_last_error().decode()


def test_create_rejects_bad_arguments():
    rc, msg = _create(ng=3)
    assert rc == -1 and "power of two" in msg
    rc, msg = _create(precision=7)
    assert rc == -1 and "precision" in msg
    rc, msg = _create(gauss=4)
    assert rc == -1 and "T2_GAUSS" in msg


def test_null_handle_errors():
    lib = sweeper.lib()
    assert lib.t2_sweep_step(None) == -1
    assert b"null" in lib.t2_last_error()
    info = sweeper.T2Info()
    assert lib.t2_sweeper_info(None, ctypes.byref(info)) == -1
    assert lib.t2_sweeper_destroy(None) == 0
    # knobs and batch calls added with the schedules and concurrency
    assert lib.t2_set_priority(None, 1, 0, 0) == -1
    assert lib.t2_set_concurrent_batches(None, 1) == -1
    assert lib.t2_batch_set_concurrent(None, 1) == -1
    assert lib.t2_batch_set_launch(None, 64) == -1
    assert lib.t2_set_items_per_lane(None, 2) == -1
    assert lib.t2_batch_solve_refined(None, 10, 1e-6, 1e-5, 10, None,
                                      None) == -1
    assert lib.t2_defect(None, None, None) == -1
    assert lib.t2_krylov_apply(None, None, None) == -1
    assert lib.t2_krylov_rhs(None, None) == -1
    assert lib.t2_krylov_apply_device(None, None, None) == -1
    assert lib.t2_batch_destroy(None) == 0


def test_geometry_rejects_bad_face_map():
    sm = meshmod.kuhn_cube(1)
    con = meshmod.tycho_connectivity(sm)
    bad = con.f2c.copy()
    bad[0, 0, 0] = 0          # face 0 cannot contain vertex 0
    with pytest.raises(sweeper.T2Error):
        sweeper.mesh_geometry(sm, meshmod.Connectivity(con.adj_cell, bad,
                                                       con.nbr_vrtx))
    dup = con.f2c.copy()
    dup[2, 1, 2] = dup[2, 1, 0]   # a face vertex twice: maps not inverse
    with pytest.raises(sweeper.T2Error, match="vertex map"):
        sweeper.mesh_geometry(sm, meshmod.Connectivity(con.adj_cell, dup,
                                                       con.nbr_vrtx))
