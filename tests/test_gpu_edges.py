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
