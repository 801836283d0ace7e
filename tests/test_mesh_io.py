"""Mesh generator, .smesh and psi file formats, connectivity."""
import os

import numpy as np
import pytest

import oracle as orc
from tycho2_amd import mesh as meshmod
from tycho2_amd import psi_io

from conftest import golden


def test_smesh_roundtrip_is_byte_identical(tmp_path):
    for name in ("cube-12.smesh", "cube-208.smesh"):
        src = golden(name)
        sm = meshmod.read_smesh(src)
        out = tmp_path / name
        meshmod.write_smesh(sm, str(out))
        with open(src, "rb") as a, open(out, "rb") as b:
            assert a.read() == b.read()


def test_cube208_counts(cube208):
    assert (cube208.n_cells, cube208.n_faces, cube208.n_nodes) == (208, 464, 67)
    assert (cube208.face_cells[:, 1] < 0).sum() == 96


def test_read_smesh_rejects_garbage(tmp_path):
    p = tmp_path / "bad.smesh"
    p.write_bytes(b"not a mesh" + bytes(100))
    with pytest.raises(ValueError):
        meshmod.read_smesh(str(p))


@pytest.mark.parametrize("n", [1, 2, 5])
def test_kuhn_topology(n):
    sm = meshmod.kuhn_cube(n)
    assert sm.n_cells == 6 * n ** 3
    assert sm.n_nodes == (n + 1) ** 3
    boundary = (sm.face_cells[:, 1] < 0).sum()
    assert boundary == 12 * n * n
    assert sm.n_faces == (4 * sm.n_cells + boundary) // 2
    vol = orc.build_mesh(sm).volume
    assert abs(vol.sum() - 1.0) < 1e-14


def test_kuhn_jitter_seeded_and_valid():
    a = meshmod.kuhn_cube(6, 0.15, seed=11)
    b = meshmod.kuhn_cube(6, 0.15, seed=11)
    c = meshmod.kuhn_cube(6, 0.15, seed=12)
    assert np.array_equal(a.coords, b.coords)
    assert not np.array_equal(a.coords, c.coords)
    base = meshmod.kuhn_cube(6)
    d = np.abs(a.coords - base.coords)
    assert d.max() <= 0.15 / 6 + 1e-15
    on_boundary = ((base.coords == 0) | (base.coords == 1)).any(axis=1)
    assert (d[on_boundary] == 0).all()
    vol = orc.build_mesh(a).volume
    assert (vol > 0).all() and abs(vol.sum() - 1.0) < 1e-13


def test_connectivity_matches_oracle(cube208):
    for sm in (cube208, meshmod.kuhn_cube(4, 0.15, seed=2)):
        con = meshmod.tycho_connectivity(sm)
        m = orc.build_mesh(sm)
        assert np.array_equal(con.adj_cell, m.adj)
        assert np.array_equal(con.f2c, m.f2c)
        assert np.array_equal(con.nbr_vrtx, m.nbrv)


def test_psi_file_roundtrip(tmp_path):
    rng = np.random.default_rng(0)
    psi = rng.normal(size=(7, 8, 4, 3))
    p = str(tmp_path / "x.psi")
    psi_io.write_psi(p, psi)
    raw = np.fromfile(p)
    assert raw.size == 8 + psi.size
    assert raw[:4].tobytes().rstrip(b"\0") == b"Tycho 2 Psi Output"
    assert list(raw[4:8].view(np.uint64)) == [1, 7, 8, 3]
    assert np.array_equal(psi_io.read_psi(p), psi)


def test_rel_diff_metric():
    b = np.array([1.0, -2.0, 4.0])
    a = b + np.array([0.1, 0.0, -0.2])
    l1, linf = psi_io.rel_diff(a, b)
    assert l1 == pytest.approx(0.3 / 7.0)
    assert linf == pytest.approx(0.2 / 4.0)
