"""The allocation-free source iteration (t2_solve_source_iteration_noalloc)
against the ordinary entry point (t2_solve_source_iteration): phi, psi, the
iteration count and the error history bit for bit, on two calls in a row
(the second on a smaller problem), with the library's device allocations
armed to fail while the allocation-free calls run."""
import ctypes

import numpy as np
import pytest

from tycho2_amd import mesh as meshmod
from tycho2_amd.quadrature import product_quadrature
from tycho2_amd.sweeper import (ENGINE_QUEUE, FP32, FP64, SweeperHIP,
                                T2Error, set_alloc_fail)

pytestmark = pytest.mark.gpu


def make(n, sn, ng, seed, precision, lagged=False):
    """Handle on a jittered n^3 x 6 Kuhn cube; lagged: S2 along the Kuhn
    diagonal faces, so that the sweep has lagged faces."""
    sm = meshmod.kuhn_cube(n, 0.2 if lagged else 0.15, seed=seed)
    omega, wt = product_quadrature(sn)
    if lagged:
        import oracle as orc
        omega, wt = orc.quadrature(2)
    rng = np.random.default_rng(seed)
    st = rng.uniform(0.5, 2.0, (sm.n_cells, ng))
    q = rng.uniform(0.0, 1.0, (sm.n_cells, 4, ng))
    s = SweeperHIP(sm, meshmod.tycho_connectivity(sm), omega, wt, ng, st,
                   0.8 * st, precision)
    s.set_source_isotropic(q)
    return s


def hip_free_bytes():
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def iterate(s, no_alloc):
    # FP32 sweeps stall near 1e-7 relative change
    err_max = 1e-12 if s.info()["precision"] == FP64 else 1e-5
    return s.solve_source_iteration(200, err_max, no_alloc=no_alloc)


def solve(s, no_alloc):
    hist = iterate(s, no_alloc)
    return hist, s.phi(), s.psi()


def assert_same(a, b):
    (h1, phi1, psi1), (h2, phi2, psi2) = a, b
    assert len(h1) == len(h2)
    assert np.array_equal(h1, h2)
    assert np.array_equal(phi1, phi2)
    assert np.array_equal(psi1, psi2)


@pytest.mark.parametrize("precision", [FP64, FP32])
def test_two_calls_in_a_row_match_the_ordinary_entry_point(precision):
    """Allocation-free solves of a larger, then a smaller problem (fresh
    handles: their sweeps run without a captured graph), with every library
    device allocation armed to fail; then the ordinary solves of the same
    handles (graph replay) give the same bits."""
    print(f"no-alloc source iteration precision={precision}", flush=True)
    big = make(6, 4, 8, 21, precision)
    small = make(3, 2, 2, 22, precision)
    try:
        free0 = hip_free_bytes()
        set_alloc_fail(1)          # the next library allocation would fail
        try:
            hist_big = iterate(big, True)
            hist_small = iterate(small, True)
        finally:
            set_alloc_fail(0)
        assert abs(hip_free_bytes() - free0) <= 4 << 20
        # (reading psi back allocates a staging array: after the calls)
        got_big = (hist_big, big.phi(), big.psi())
        got_small = (hist_small, small.phi(), small.psi())
        print(f"  big: {len(got_big[0])} iterations, small: "
              f"{len(got_small[0])} iterations", flush=True)
        assert_same(got_big, solve(big, False))
        assert_same(got_small, solve(small, False))
        # and once more on the allocation-free path, now with graphs
        # captured by the ordinary calls: the same bits again
        assert_same(solve(big, True), got_big)
        assert_same(solve(small, True), got_small)
    finally:
        big.close()
        small.close()


def test_lagged_faces_and_concurrent_batches():
    """Lagged faces (their pre-sweep snapshot) and concurrent angle batches
    (streams made by the handle's first ordinary solve)."""
    print("no-alloc source iteration, lagged faces + concurrent batches",
          flush=True)
    with make(3, 2, 4, 3, FP64, lagged=True) as s:
        assert s.info()["n_lagged"] > 0
        s.set_schedule(2)
        s.set_concurrent_batches(True)
        with pytest.raises(T2Error) as e:
            s.solve_source_iteration(50, 1e-12, no_alloc=True)
        assert e.value.code == -4 and "streams" in str(e.value)
        ref = solve(s, False)
        assert_same(solve(s, True), ref)


def test_queue_engine_needs_its_tables_first():
    """The ready-queue tables are built by an ordinary call; before that the
    allocation-free call refuses, after it the results are the same bits."""
    print("no-alloc source iteration, queue engine", flush=True)
    with make(4, 4, 4, 5, FP64) as s:
        s.set_engine(ENGINE_QUEUE)
        with pytest.raises(T2Error) as e:
            s.solve_source_iteration(50, 1e-12, no_alloc=True)
        assert e.value.code == -4 and "ready-queue" in str(e.value)
        ref = solve(s, False)
        assert_same(solve(s, True), ref)
