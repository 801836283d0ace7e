"""The Krylov operator applies (t2_krylov_apply, t2_krylov_rhs) against the
oracle's sweep and moment, and an outside Krylov solver driving the engine
through them (as the reference's KrylovSolver drives lhsOperator,
src/SourceIteration.cc:76-114, 216-285), converging to the oracle's GMRES
answer.  All calls go through the C ABI.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as sla

import oracle as orc
from tycho2_amd import mesh as meshmod
from tycho2_amd.quadrature import product_quadrature
from tycho2_amd.sweeper import FP32, FP64, SweeperHIP

pytestmark = pytest.mark.gpu

FOUR_PI = 4.0 * np.pi


def maxrel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def setup(ng, seed, n=3, sn=4):
    sm = meshmod.kuhn_cube(n, 0.15, seed=seed)
    conn = meshmod.tycho_connectivity(sm)
    omega, wt = product_quadrature(sn)
    rng = np.random.default_rng(seed)
    st = rng.uniform(0.5, 3.0, (sm.n_cells, ng))
    ss = st * rng.uniform(0.1, 0.9, (sm.n_cells, ng))
    return sm, conn, omega, wt, st, ss, rng


def oracle_d_linv(sm, omega, wt, ng, st, src):
    psi, lag = orc.sweep(orc.build_mesh(sm), omega, ng, st, src)
    assert lag == 0
    return orc.psi_to_phi(psi, wt)


@pytest.mark.parametrize("precision,ng,bar", [(FP64, 1, 1e-12),
                                              (FP64, 16, 1e-12),
                                              (FP32, 4, 2e-5)])
def test_apply_and_rhs_match_oracle(precision, ng, bar):
    print(f"krylov apply precision={precision} ng={ng}", flush=True)
    sm, conn, omega, wt, st, ss, rng = setup(ng, 40 + ng)
    nc, na = sm.n_cells, omega.shape[0]
    x = rng.uniform(-0.5, 2.0, (nc, 4, ng))
    q = rng.uniform(0.0, 2.0, (nc, 4, ng))
    phi0 = rng.uniform(0.0, 1.0, (nc, 4, ng))
    with SweeperHIP(sm, conn, omega, wt, ng, st, ss, precision) as s:
        assert s.info()["n_lagged"] == 0
        s.set_source_isotropic(q)
        s.set_phi(phi0)
        b = s.krylov_apply(x)
        rhs = s.krylov_rhs()
        assert np.array_equal(s.phi(), phi0)      # phi is left alone
    ms_x = (ss[:, None, :] * x)[:, None] / FOUR_PI
    want_b = x - oracle_d_linv(sm, omega, wt, ng, st,
                               np.broadcast_to(ms_x, (nc, na, 4, ng)))
    want_rhs = oracle_d_linv(sm, omega, wt, ng, st,
                             np.broadcast_to(q[:, None], (nc, na, 4, ng)))
    assert maxrel(b, want_b) <= bar
    assert maxrel(rhs, want_rhs) <= bar


def test_outside_gmres_drives_the_engine():
    """scipy's GMRES with the engine as its operator (only phi-sized vectors
    cross the host link) reaches the oracle's converged phi, and the final
    sweep gives the oracle's psi, as the reference's krylov() does."""
    print("outside gmres", flush=True)
    ng = 4
    sm, conn, omega, wt, st, ss, rng = setup(ng, 61, n=4)
    nc, na = sm.n_cells, omega.shape[0]
    q = rng.uniform(0.5, 2.0, (nc, 4, ng))
    n = nc * 4 * ng
    src = np.ascontiguousarray(np.broadcast_to(q[:, None], (nc, na, 4, ng)))
    rpsi, rphi, _ = orc.krylov(orc.build_mesh(sm), omega, wt, ng, st, ss,
                               src, 400, 1e-14, 30)
    with SweeperHIP(sm, conn, omega, wt, ng, st, ss, FP64) as s:
        assert s.info()["n_lagged"] == 0
        s.set_source_isotropic(q)
        op = sla.LinearOperator(
            (n, n), dtype=np.float64,
            matvec=lambda v: s.krylov_apply(v.reshape(nc, 4, ng)).ravel())
        b = s.krylov_rhs().ravel()
        x, info = sla.gmres(op, b, rtol=1e-13, atol=0.0, restart=30,
                            maxiter=20)
        assert info == 0
        phi = x.reshape(nc, 4, ng)
        s.set_phi(phi)
        s.sweep_step()                      # psi = L^-1 (M S phi + q)
        psi = s.psi()
    assert maxrel(phi, rphi) <= 1e-11
    assert maxrel(psi, rpsi) <= 1e-11


@pytest.mark.parametrize("lagged", [False, True], ids=["acyclic", "lagged"])
@pytest.mark.parametrize("precision", [FP64, FP32])
def test_solve_ignores_the_previous_psi(precision, lagged):
    """GMRES (FP64 handles) and the refined solve (FP32 handles) give the
    same phi and psi bit for bit whatever psi the handle held before: with
    lagged faces the solve starts from psi = 0, without them its first
    sweep writes all of psi before anything reads it."""
    print(f"solve ignores previous psi precision={precision} "
          f"lagged={lagged}", flush=True)
    ng = 4
    if lagged:
        sm = meshmod.kuhn_cube(3, 0.2, seed=3)
        omega, wt = orc.quadrature(2)
        rng = np.random.default_rng(8)
        st = rng.uniform(0.5, 3.0, (sm.n_cells, ng))
        ss = st * rng.uniform(0.1, 0.9, (sm.n_cells, ng))
        conn = meshmod.tycho_connectivity(sm)
    else:
        sm, conn, omega, wt, st, ss, rng = setup(ng, 52)
    nc, na = sm.n_cells, omega.shape[0]
    q = rng.uniform(0.5, 2.0, (nc, 4, ng))
    with SweeperHIP(sm, conn, omega, wt, ng, st, ss, precision) as s:
        assert (s.info()["n_lagged"] > 0) == lagged
        s.set_source_isotropic(q)

        def solve():
            if precision == FP64:
                s.solve_krylov(100, 1e-10, 20)
            else:
                s.solve_refined(100, 1e-6, 1e-5, 20)
            return s.phi(), s.psi()

        first = solve()
        s.defect(rng.uniform(-5.0, 5.0, (nc, na, 4, ng)))  # plant a psi
        second = solve()
    for a, b in zip(first, second):
        assert np.array_equal(a, b)


def test_device_apply_equals_host_apply():
    """The apply on device vectors (allocated with the HIP runtime the
    library uses) gives the host apply's b bit for bit; aliasing or
    overlapping vectors are refused."""
    import ctypes
    from tycho2_amd.sweeper import T2Error
    print("krylov apply on device vectors", flush=True)
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p,
                              ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    h2d, d2h = 1, 2
    ng = 4
    sm, conn, omega, wt, st, ss, rng = setup(ng, 77)
    x = rng.uniform(-0.5, 2.0, (sm.n_cells, 4, ng))
    got = np.zeros_like(x)
    dx, db = ctypes.c_void_p(), ctypes.c_void_p()
    with SweeperHIP(sm, conn, omega, wt, ng, st, ss, FP64) as s:
        assert s.info()["n_lagged"] == 0
        want = s.krylov_apply(x)
        assert hip.hipMalloc(ctypes.byref(dx), x.nbytes) == 0
        assert hip.hipMalloc(ctypes.byref(db), x.nbytes) == 0
        try:
            assert hip.hipMemcpy(dx, x.ctypes.data, x.nbytes, h2d) == 0
            s.synchronize()
            s.krylov_apply_device(dx.value, db.value)
            s.synchronize()
            assert hip.hipMemcpy(got.ctypes.data, db, x.nbytes, d2h) == 0
            with pytest.raises(T2Error):
                s.krylov_apply_device(dx.value, dx.value)
            with pytest.raises(T2Error):
                s.krylov_apply_device(dx.value, dx.value + 8)
        finally:
            hip.hipFree(dx)
            hip.hipFree(db)
    assert np.array_equal(got, want)
