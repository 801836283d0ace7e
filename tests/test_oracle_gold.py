"""The oracle against the reference's own golden answer (test/regression/
gold.psi on util/cube-208.smesh, input-gold.deck) and against the
reference's own quadrature built from its sources (oracle/_ref)."""
import ctypes
import os

import numpy as np
import pytest

import oracle as orc
from tycho2_amd.psi_io import rel_diff

REF_QUAD = os.path.join(os.path.dirname(orc.__file__), "_ref",
                        "libref_quadrature.so")


def test_gold_fixture_header(gold):
    assert int(gold["n_cells"]) == 208 and int(gold["n_angles"]) == 128
    assert int(gold["n_groups"]) == 2 and bool(gold["groups_identical"])


@pytest.fixture(scope="module")
def gold_run(cube208, gold):
    m = orc.build_mesh(cube208)
    om, wt = gold["ref_omega"], gold["ref_weights"]
    src = orc.hat_source(m, om, 2, 10.0, 5.0)
    psi, phi, hist = orc.source_iteration(m, om, wt, 2, 10.0, 5.0, src,
                                          int(gold["iter_max"]),
                                          float(gold["err_max"]))
    return psi, phi, hist


def test_oracle_reproduces_gold_psi(gold_run, gold):
    """With the reference's quadrature the oracle's source iteration gives
    the reference's psi to a few ulps (the reference's bar is 1e-10,
    test/regression.py:4)."""
    psi, _, hist = gold_run
    got = psi[..., 0][:, gold["psi_angles"], :]
    l1, linf = rel_diff(got, gold["psi_g0"])
    assert l1 < 2e-15 and linf < 5e-15
    assert np.array_equal(psi[..., 0], psi[..., 1])
    assert abs(np.abs(psi).sum() / gold["psi_l1"] - 1) < 1e-14
    assert abs(np.abs(psi).max() / gold["psi_linf"] - 1) < 1e-14
    assert abs((psi * psi).sum() / gold["psi_sumsq"] - 1) < 1e-14
    assert 10 < len(hist) < 100 and hist[-1] <= 1e-10


def test_oracle_gold_scalar_flux(gold_run, gold):
    _, phi, _ = gold_run
    l1, linf = rel_diff(phi[:, :, 0], gold["phi_g0"])
    assert l1 < 2e-15 and linf < 5e-15


def test_gold_with_own_quadrature(cube208, gold):
    """The oracle's own quadrature tables (correctly rounded) stay within
    the reference's 1e-10 bar."""
    m = orc.build_mesh(cube208)
    om, wt = orc.quadrature(8)
    src = orc.hat_source(m, om, 2, 10.0, 5.0)
    psi, _, _ = orc.source_iteration(m, om, wt, 2, 10.0, 5.0, src, 100,
                                     1e-10)
    got = psi[..., 0][:, gold["psi_angles"], :]
    l1, linf = rel_diff(got, gold["psi_g0"])
    assert l1 < 1e-13 and linf < 1e-13


@pytest.mark.skipif(not os.path.exists(REF_QUAD),
                    reason="oracle/_ref not built (no /root/reference here)")
@pytest.mark.parametrize("sn", list(range(2, 31, 2)))
def test_quadrature_against_reference_build(sn):
    """The reference's Quadrature, compiled from its own sources, against
    the oracle: identical construction, tables equal to ~1e-15 (the
    reference's literals came from an older numpy)."""
    lib = ctypes.CDLL(REF_QUAD)
    na = 2 * sn * sn
    r = [np.empty(na) for _ in range(4)]
    dp = ctypes.POINTER(ctypes.c_double)
    assert lib.ref_quadrature(sn, *[a.ctypes.data_as(dp) for a in r]) == na
    om, wt = orc.quadrature(sn)
    assert np.abs(np.stack(r[:3], axis=1) - om).max() < 5e-15
    assert np.abs(r[3] - wt).max() < 5e-15


def test_reference_quadrature_in_fixture(gold):
    """The fixture keeps the reference's S8 set; the oracle's agrees."""
    om, wt = orc.quadrature(8)
    assert np.abs(gold["ref_omega"] - om).max() < 5e-15
    assert np.abs(gold["ref_weights"] - wt).max() < 5e-15
