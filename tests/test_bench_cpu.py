"""bench.py's host-side pieces: the algorithmic-bytes model of the roofline
(its mean-upwind-face term checked against a direct count on a mesh), the
traffic figures it reads from the committed profiles, and its one-GPU
contract."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT
from tycho2_amd import mesh as meshmod
from tycho2_amd import sweeper
from tycho2_amd.quadrature import product_quadrature

sys.path.insert(0, ROOT)
import bench  # noqa: E402


@pytest.mark.parametrize("n,jitter", [(4, 0.0), (5, 0.15)])
def test_mean_incoming_interior_faces_is_two_minus_one_over_n(n, jitter):
    """Every interior face is incoming for exactly one of its two cells, so
    the mean over (cell, angle) is 2 - 1/n on the Kuhn cube, whatever the
    jitter and direction."""
    sm = meshmod.kuhn_cube(n, jitter, seed=2)
    con = meshmod.tycho_connectivity(sm)
    _, _, nrm = sweeper.mesh_geometry(sm, con)
    omega, _ = product_quadrature(4)
    adj = con.adj_cell
    counts = []
    for om in omega:
        odn = nrm @ om
        counts.append(((adj >= 0) & (odn < 0)).sum())
    mean = np.mean(counts) / sm.n_cells
    assert mean == pytest.approx(2.0 - 1.0 / n, abs=1e-12)


def test_algorithmic_bytes_per_solve():
    """w (4 + 3 f_in) + (4 w + 32) / N_a: psi written (4 values), upwind
    psi read (3 values per incoming interior face), the isotropic source
    read and the FP64 phi written once per cell, group and sweep."""
    n, na = 60, 128
    f_in = 2.0 - 1.0 / n
    for w in (8, 4):
        want = 4 * w + 3 * f_in * w + (4 * w) / na + (4 * 8) / na
        assert bench.alg_bytes_per_solve(w, n, na) == pytest.approx(want)
    assert bench.alg_bytes_per_solve(8, n, na) == pytest.approx(80.1, abs=0.01)
    assert bench.alg_bytes_per_solve(4, n, na) == pytest.approx(40.175,
                                                                 abs=1e-3)


def test_traffic_figures_from_profiles():
    path = os.path.join(ROOT, "profiles", "traffic_precision.json")
    prof = json.load(open(path))
    for prec in ("fp64", "fp32"):
        t = bench.traffic_from_profiles("precision", prec)
        assert t == prof[prec]["hbm_bytes_per_launch"]
        # a schedule of 867 launches splits the same per-sweep traffic
        t867 = bench.traffic_from_profiles("precision", prec, 867)
        assert t867 == round(prof[prec]["hbm_bytes_per_sweep"] / 867)
        # actual traffic at least the algorithmic bytes
        assert (prof[prec]["hbm_bytes_per_sweep"] >=
                prof[prec]["algorithmic_bytes_per_sweep"])
    assert bench.traffic_from_profiles("no-such-workload", "fp64") is None


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"),
                        "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_phi_error_measures():
    """A uniform relative perturbation e gives e in every measure; an
    entry of the reference below the floor is measured against the floor,
    and the L1 measure weighs entries by size."""
    rng = np.random.default_rng(2)
    ref = rng.uniform(0.5, 2.0, (30, 4, 3))
    e = 1e-9
    got = bench.phi_errors(ref * (1 + e), ref)
    assert got["l1_rel"] == pytest.approx(e, rel=1e-6)
    assert got["elem_maxrel"] == pytest.approx(e, rel=1e-6)
    ref2 = ref.copy()
    ref2[3, 1, 2] = 0.0
    a = ref2.copy()
    a[3, 1, 2] = 1e-9
    floor = bench.ELEM_FLOOR * np.abs(ref2).max()
    got = bench.phi_errors(a, ref2)
    assert got["elem_maxrel"] == pytest.approx(1e-9 / floor, rel=1e-12)
    assert got["l1_rel"] == pytest.approx(1e-9 / np.abs(ref2).sum(),
                                          rel=1e-12)
    assert bench.maxrel(a, ref2) == pytest.approx(
        1e-9 / np.abs(ref2).max(), rel=1e-12)
