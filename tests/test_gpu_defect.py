"""The FP64 defect r = q + M S phi - L psi (t2_defect) against the oracle's
defect of the same psi, total source and cross sections.

The oracle's defect is pinned on the CPU against an independently assembled
global DG system; here the device's operator apply (the one the refined
solve uses) must give the same nodal density for an arbitrary psi, on a mesh
without and with lagged faces, for isotropic and angular sources, FP64 and
FP32 handles.  All calls go through the C ABI.
"""
import numpy as np
import pytest

import oracle as orc
from tycho2_amd import mesh as meshmod
from tycho2_amd.quadrature import product_quadrature
from tycho2_amd.sweeper import FP32, FP64, SweeperHIP

pytestmark = pytest.mark.gpu

FOUR_PI = 4.0 * np.pi


def maxrel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def setup(n, sn, jitter, seed, ng, lagged):
    sm = meshmod.kuhn_cube(n, jitter, seed=seed)
    conn = meshmod.tycho_connectivity(sm)
    omega, wt = orc.quadrature(sn) if lagged else product_quadrature(sn)
    rng = np.random.default_rng(seed)
    st = rng.uniform(0.5, 3.0, (sm.n_cells, ng))
    ss = st * rng.uniform(0.1, 0.9, (sm.n_cells, ng))
    return sm, conn, omega, wt, st, ss, rng


def total_source(q_ang, ss, phi):
    """q[c][a][v][g] + sigma_s / (4 pi) phi[c][v][g], angle-expanded."""
    return q_ang + (ss[:, None, :] * phi)[:, None, :, :] / FOUR_PI


# (mesh n, quadrature order, jitter, seed, lagged faces expected)
MESHES = [(3, 4, 0.15, 5, False), (3, 2, 0.2, 3, True)]


@pytest.mark.parametrize("mesh", MESHES, ids=["acyclic", "lagged"])
@pytest.mark.parametrize("angular", [False, True], ids=["iso", "ang"])
@pytest.mark.parametrize("ng", [1, 4, 16])
def test_defect_fp64_matches_oracle(mesh, angular, ng):
    n, sn, jit, seed, lagged = mesh
    print(f"defect fp64 mesh={mesh} angular={angular} ng={ng}", flush=True)
    sm, conn, omega, wt, st, ss, rng = setup(n, sn, jit, seed, ng, lagged)
    nc, na = sm.n_cells, omega.shape[0]
    phi = rng.uniform(0.0, 2.0, (nc, 4, ng))
    psi = rng.uniform(0.0, 1.0, (nc, na, 4, ng))
    with SweeperHIP(sm, conn, omega, wt, ng, st, ss, FP64) as s:
        assert (s.info()["n_lagged"] > 0) == lagged
        if angular:
            q_ang = rng.uniform(-0.5, 2.0, (nc, na, 4, ng))
            s.set_source_angular(q_ang)
        else:
            q = rng.uniform(0.0, 2.0, (nc, 4, ng))
            s.set_source_isotropic(q)
            q_ang = np.broadcast_to(q[:, None], (nc, na, 4, ng))
        s.set_phi(phi)
        got = s.defect(psi)
        again = s.defect(None)          # psi is now the handle's own
    ref = orc.defect(orc.build_mesh(sm), omega, ng, st,
                     total_source(q_ang, ss, phi), psi)
    assert np.abs(ref).max() > 1e-2     # psi is far from the solution
    assert maxrel(got, ref) <= 1e-12
    assert np.array_equal(again, got)


@pytest.mark.parametrize("ng", [1, 16])
def test_defect_fp32_handle_rounds_an_fp64_defect(ng):
    """FP32 handles keep psi in FP32 but form the defect from FP64
    coefficients: the result is the FP64 defect of the FP32-rounded psi,
    rounded once to FP32."""
    print(f"defect fp32 ng={ng}", flush=True)
    sm, conn, omega, wt, st, ss, rng = setup(3, 4, 0.15, 7, ng, False)
    nc, na = sm.n_cells, omega.shape[0]
    phi = rng.uniform(0.0, 2.0, (nc, 4, ng))
    psi = rng.uniform(0.0, 1.0, (nc, na, 4, ng))
    q = rng.uniform(0.0, 2.0, (nc, 4, ng))
    with SweeperHIP(sm, conn, omega, wt, ng, st, ss, FP32) as s:
        s.set_source_isotropic(q)
        s.set_phi(phi)
        got = s.defect(psi)
    psi32 = psi.astype(np.float32).astype(np.float64)
    ref = orc.defect(orc.build_mesh(sm), omega, ng, st,
                     total_source(np.broadcast_to(q[:, None], psi.shape), ss,
                                  phi), psi32)
    # one FP32 rounding (2^-24 relative) on top of FP64 round-off
    tol = 2.0 ** -24 * np.abs(ref) + 1e-13 * np.abs(ref).max()
    assert np.all(np.abs(got - ref) <= tol)
    assert np.array_equal(got, got.astype(np.float32).astype(np.float64))


def test_defect_of_a_converged_solve_vanishes():
    """After GMRES to 1e-13 and its final sweep psi = L^-1 (q + M S phi),
    the defect of the handle's own psi is round-off, and a source added
    afterwards shows up in the defect as itself."""
    print("defect of a converged solve", flush=True)
    ng = 4
    sm, conn, omega, wt, st, ss, rng = setup(4, 4, 0.15, 9, ng, False)
    nc, na = sm.n_cells, omega.shape[0]
    q = rng.uniform(0.5, 2.0, (nc, 4, ng))
    with SweeperHIP(sm, conn, omega, wt, ng, st, ss, FP64) as s:
        s.set_source_isotropic(q)
        s.solve_krylov(200, 1e-13, 20)
        d0 = s.defect(None)
        extra = rng.uniform(0.0, 1.0, (nc, 4, ng))
        s.set_source_isotropic(q + extra)
        d1 = s.defect(None)
    assert np.abs(d0).max() <= 1e-12 * np.abs(q).max()
    want = np.broadcast_to(extra[:, None], (nc, na, 4, ng))
    assert np.allclose(d1, want, rtol=1e-11, atol=1e-12)
