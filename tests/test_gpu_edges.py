"""Edge cases of the device sweep against the oracle: void (sigma_t = 0) and
optically thick cells, a one-tet mesh (every face on the vacuum boundary),
and the largest group count."""
import numpy as np
import pytest

import oracle as orc
from tycho2_amd import mesh as meshmod
from tycho2_amd.quadrature import product_quadrature
from tycho2_amd.sweeper import (FP32, FP64, GAUSS_CRAMER_GLU, GAUSS_NOPIVOT,
                                GAUSS_ORIGINAL, SweeperHIP)

pytestmark = pytest.mark.gpu


def maxrel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def sweep_both(sm, omega, wt, ng, st, src, precision=FP64,
               gauss=GAUSS_NOPIVOT):
    conn = meshmod.tycho_connectivity(sm)
    psi = np.zeros_like(src)
    with SweeperHIP(sm, conn, omega, wt, ng, st, 0.5 * st, precision,
                    gauss=gauss) as s:
        s.sweep(psi, src)
    ref, _ = orc.sweep(orc.build_mesh(sm), omega, ng, st, src,
                       method=gauss if gauss != GAUSS_NOPIVOT else orc.NOPIVOT)
    return psi, ref


@pytest.mark.parametrize("gauss", [GAUSS_NOPIVOT, GAUSS_ORIGINAL,
                                   GAUSS_CRAMER_GLU])
def test_void_and_thick_cells(gauss):
    """sigma_t = 0 in a third of the cells, 1e3 in some others."""
    print(f"void and thick cells gauss={gauss}", flush=True)
    sm = meshmod.kuhn_cube(4, 0.15, seed=5)
    omega, wt = product_quadrature(4)
    rng = np.random.default_rng(0)
    st = rng.uniform(0.5, 2.0, (sm.n_cells, 4))
    st[::3] = 0.0
    st[1::7] = 1e3
    src = rng.uniform(0.0, 1.0, (sm.n_cells, omega.shape[0], 4, 4))
    psi, ref = sweep_both(sm, omega, wt, 4, st, src, gauss=gauss)
    assert np.isfinite(psi).all()
    assert maxrel(psi, ref) <= 1e-12


def test_void_and_thick_cells_fp32():
    print("void and thick cells fp32", flush=True)
    sm = meshmod.kuhn_cube(4, 0.15, seed=6)
    omega, wt = product_quadrature(4)
    rng = np.random.default_rng(1)
    st = rng.uniform(0.5, 2.0, (sm.n_cells, 1))
    st[::3] = 0.0
    st[2::9] = 1e3
    src = rng.uniform(0.0, 1.0, (sm.n_cells, omega.shape[0], 4, 1))
    psi, ref = sweep_both(sm, omega, wt, 1, st, src, precision=FP32)
    assert maxrel(psi, ref) <= 2e-5


def one_tet() -> meshmod.SerialMesh:
    coords = np.array([[0.0, 0.0, 0.0], [1.0, 0.1, 0.0], [0.2, 0.9, 0.1],
                       [0.1, 0.2, 1.1]])
    # face k opposite node k
    faces = np.array([[1, 2, 3], [0, 2, 3], [0, 1, 3], [0, 1, 2]])
    return meshmod.SerialMesh(
        cell_faces=np.array([[0, 1, 2, 3]]),
        cell_nodes=np.array([[0, 1, 2, 3]]), material=np.zeros(1, np.int64),
        face_cells=np.array([[0, -1]] * 4), face_nodes=faces, coords=coords)


@pytest.mark.parametrize("ng", [1, 64])
def test_one_tet_mesh(ng):
    """One cell: every face on the vacuum boundary, one step per sweep."""
    print(f"one tet ng={ng}", flush=True)
    sm = one_tet()
    omega, wt = product_quadrature(4)
    rng = np.random.default_rng(2)
    st = rng.uniform(0.5, 2.0, (1, ng))
    src = rng.uniform(0.0, 1.0, (1, omega.shape[0], 4, ng))
    psi, ref = sweep_both(sm, omega, wt, ng, st, src)
    assert maxrel(psi, ref) <= 1e-12


def test_sixty_four_groups():
    """The largest group count: a lane per group, one item per wavefront."""
    print("64 groups", flush=True)
    sm = meshmod.kuhn_cube(3, 0.15, seed=7)
    omega, wt = product_quadrature(4)
    rng = np.random.default_rng(3)
    st = rng.uniform(0.5, 3.0, (sm.n_cells, 64))
    src = rng.uniform(-0.5, 2.0, (sm.n_cells, omega.shape[0], 4, 64))
    psi, ref = sweep_both(sm, omega, wt, 64, st, src)
    assert maxrel(psi, ref) <= 1e-12


def zero_face_directions(sm, m):
    """Directions with Omega . n exactly zero on some faces of an unjittered
    Kuhn cube: two in the xy-plane (zero on every z-normal face) and two
    along Omega_x = Omega_y (zero on the faces whose normal has n_x =
    -n_y, n_z = 0)."""
    c, s = 0.6, 0.8
    d = 1.0 / np.sqrt(3.0)
    omega = np.array([[c, s, 0.0], [-s, c, 0.0], [d, d, d], [-d, -d, d]])
    wt = np.full(4, np.pi)
    dots = np.einsum("ad,ckd->ack", omega, m.normal)
    return omega, wt, dots


def test_exactly_zero_omega_dot_n_faces():
    """Caller-supplied directions perpendicular to face normals: the
    traversal counts an Omega . n = 0 face as incoming on both sides
    (src/TychoMesh.hh:107-110), which makes a cycle that is broken by
    lagging it; the cell solve gives it no coupling (A (Omega . n) = 0,
    src/Transport.cc:113,184).  Device and oracle agree, whatever psi the
    lagged faces read."""
    print("exactly zero Omega.n faces", flush=True)
    sm = meshmod.kuhn_cube(3, 0.0)
    m = orc.build_mesh(sm)
    omega, wt, dots = zero_face_directions(sm, m)
    interior = m.adj >= 0
    zeros = [(dots[a] == 0.0)[interior].sum() for a in range(4)]
    print(f"  exact-zero interior faces per direction: {zeros}")
    assert all(z > 0 for z in zeros)
    ng = 4
    rng = np.random.default_rng(11)
    st = rng.uniform(0.5, 2.0, (sm.n_cells, ng))
    src = rng.uniform(0.0, 1.0, (sm.n_cells, 4, 4, ng))
    prev = rng.uniform(0.0, 1.0, src.shape)
    conn = meshmod.tycho_connectivity(sm)
    psi = prev.copy()
    with SweeperHIP(sm, conn, omega, wt, ng, st, 0.5 * st, FP64) as s:
        n_lag = s.info()["n_lagged"]
        s.sweep(psi, src)
        # the lagged faces' previous psi has no effect: a sweep from zeros
        psi0 = np.zeros_like(src)
        s.sweep(psi0, src)
    ref, lag = orc.sweep(m, omega, ng, st, src, psi=prev.copy())
    print(f"  lagged faces: engine {n_lag}, oracle {lag}")
    assert n_lag == lag > 0
    assert maxrel(psi, ref) <= 1e-12
    assert np.array_equal(psi, psi0)


def hip_free_bytes():
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


@pytest.mark.parametrize("precision,lagged", [(FP64, True), (FP32, False)])
def test_device_allocation_failures_are_clean(precision, lagged):
    """Every device allocation of a handle's creation failed in turn by the
    testing hook (t2_set_alloc_fail): creation returns T2_ERR_NOMEM and
    releases what it had taken, and the next creation succeeds and sweeps
    as the oracle does."""
    from tycho2_amd.sweeper import T2Error, set_alloc_fail
    print(f"allocation failures precision={precision} lagged={lagged}",
          flush=True)
    if lagged:   # S2 along the Kuhn diagonal faces: lagged faces
        sm = meshmod.kuhn_cube(3, 0.2, seed=3)
        omega, wt = orc.quadrature(2)
    else:
        sm = meshmod.kuhn_cube(3, 0.15, seed=3)
        omega, wt = product_quadrature(4)
    conn = meshmod.tycho_connectivity(sm)
    ng = 2
    rng = np.random.default_rng(12)
    st = rng.uniform(0.5, 2.0, (sm.n_cells, ng))
    with SweeperHIP(sm, conn, omega, wt, ng, st, 0.5 * st, precision):
        pass                                   # warm the runtime's pools
    free0 = hip_free_bytes()
    failures = 0
    try:
        for nth in range(1, 200):
            set_alloc_fail(nth)
            try:
                s = SweeperHIP(sm, conn, omega, wt, ng, st, 0.5 * st,
                               precision)
            except T2Error as e:
                assert e.code == -3 and "t2_set_alloc_fail" in str(e), e
                failures += 1
                continue
            s.close()
            break
        else:
            pytest.fail("creation never succeeded")
    finally:
        set_alloc_fail(0)
    print(f"  {failures} allocations of a creation failed in turn")
    assert failures >= 10
    assert abs(hip_free_bytes() - free0) <= 64 << 20
    src = rng.uniform(0.0, 1.0, (sm.n_cells, omega.shape[0], 4, ng))
    psi, ref = sweep_both(sm, omega, wt, ng, st, src, precision=precision)
    assert maxrel(psi, ref) <= (1e-12 if precision == FP64 else 2e-5)
