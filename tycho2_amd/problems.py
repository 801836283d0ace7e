"""Problem definitions of the benchmark workloads and the reference's gold
deck.

Every problem is a unit cube of Kuhn tetrahedra (``mesh.kuhn_cube``), a
product quadrature, per-(cell, group) cross sections, an isotropic source
q[cell][vertex][group] (or the reference's angle-dependent hat source) and
vacuum boundaries.  All random draws are seeded and recorded.

  plumbing        5^3 x 6 = 750 tets, unjittered, S4 (32 angles), G=1,
                  sigma_t=1, sigma_s=0.5, q=1; SI to 1e-12
  plain           40^3 x 6 = 384,000 tets, jitter 0.15h, S8 (128), G=16,
                  sigma_t_g = 1 + 0.1 g, sigma_s = 0.5 sigma_t; SI + GMRES(20)
  precision       60^3 x 6 = 1,296,000 tets, jitter, S8, G=16,
                  sigma_t_g = 1 + 0.1 g (our stated choice), sigma_s/sigma_t=0.9
  scattering      50^3 x 6 = 750,000 tets, jitter, S12 (288), G=8,
                  sigma_t ~ logU[0.1, 10] per cell, sigma_s = 0.99 sigma_t,
                  q=1 inside r < 0.2 about the centre; GMRES(30)
  batch           48 problems, n ~ U{10..45}, angles from {32,72,128,288},
                  groups from {1,4,16,32}, sigma_t ~ logU[0.5,5],
                  sigma_s/sigma_t ~ U[0.3, 0.95]
  gold            the reference's regression deck on cube-208.smesh: S8,
                  G=2, sigma_t=10, sigma_s=5, hat source (src/Problem.cc:47-82)
"""
from __future__ import annotations

import dataclasses
from typing import Optional

import numpy as np

from . import mesh as meshmod
from .quadrature import product_quadrature

JITTER_SEED = 20240611
XS_SEED = 7
BATCH_SEED = 11


@dataclasses.dataclass
class Problem:
    name: str
    smesh: meshmod.SerialMesh
    conn: meshmod.Connectivity
    sn: int
    omega: np.ndarray
    wt: np.ndarray
    ng: int
    sigma_t: np.ndarray            # (nc, ng)
    sigma_s: np.ndarray            # (nc, ng)
    q_iso: Optional[np.ndarray]    # (nc, 4, ng) or None
    solver: dict

    @property
    def nc(self) -> int:
        return self.smesh.n_cells

    @property
    def na(self) -> int:
        return int(self.omega.shape[0])

    @property
    def solves_per_sweep(self) -> int:
        return self.nc * self.na * self.ng


def _make(name, smesh, sn, ng, sigma_t, sigma_s, q_iso, solver) -> Problem:
    omega, wt = product_quadrature(sn)
    conn = meshmod.tycho_connectivity(smesh)
    nc = smesh.n_cells
    st = np.ascontiguousarray(np.broadcast_to(sigma_t, (nc, ng)), np.float64)
    ss = np.ascontiguousarray(np.broadcast_to(sigma_s, (nc, ng)), np.float64)
    return Problem(name, smesh, conn, sn, omega, wt, ng, st, ss, q_iso, solver)


def _uniform_q(nc, ng, value=1.0):
    return np.full((nc, 4, ng), value, dtype=np.float64)


def plumbing(n: int = 5) -> Problem:
    sm = meshmod.kuhn_cube(n, 0.0)
    return _make("plumbing", sm, 4, 1, 1.0, 0.5, _uniform_q(sm.n_cells, 1),
                 dict(method="si", iter_max=1000, tol=1e-12))


def plain(n: int = 40, sn: int = 8, ng: int = 16) -> Problem:
    sm = meshmod.kuhn_cube(n, 0.15, seed=JITTER_SEED)
    st = 1.0 + 0.1 * np.arange(ng)
    return _make("plain", sm, sn, ng, st[None, :], 0.5 * st[None, :],
                 _uniform_q(sm.n_cells, ng),
                 dict(method="krylov", iter_max=200, tol=1e-10, restart=20))


def precision(n: int = 60, sn: int = 8, ng: int = 16) -> Problem:
    sm = meshmod.kuhn_cube(n, 0.15, seed=JITTER_SEED)
    st = 1.0 + 0.1 * np.arange(ng)
    return _make("precision", sm, sn, ng, st[None, :], 0.9 * st[None, :],
                 _uniform_q(sm.n_cells, ng),
                 dict(method="krylov", iter_max=300, tol=1e-10, restart=20))


def _centroids(sm: meshmod.SerialMesh) -> np.ndarray:
    return sm.coords[sm.cell_nodes].mean(axis=1)


def scattering(n: int = 50, sn: int = 12, ng: int = 8) -> Problem:
    sm = meshmod.kuhn_cube(n, 0.15, seed=JITTER_SEED)
    rng = np.random.default_rng(XS_SEED)
    st = np.exp(rng.uniform(np.log(0.1), np.log(10.0), sm.n_cells))
    cen = _centroids(sm)
    inside = np.linalg.norm(cen - 0.5, axis=1) < 0.2
    q = np.zeros((sm.n_cells, 4, ng))
    q[inside] = 1.0
    return _make("scattering", sm, sn, ng, st[:, None], 0.99 * st[:, None], q,
                 dict(method="krylov", iter_max=600, tol=1e-10, restart=30))


def batch(count: int = 48, seed: int = BATCH_SEED, n_range=(10, 45),
          angle_orders=(4, 6, 8, 12), groups=(1, 4, 16, 32), only=None):
    """Seeded list of independent problems of uneven size.  ``only`` (a set
    of indices) builds just those members, with the same draws."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.integers(n_range[0], n_range[1] + 1))
        sn = int(rng.choice(angle_orders))
        ng = int(rng.choice(groups))
        nc = 6 * n ** 3
        st = np.exp(rng.uniform(np.log(0.5), np.log(5.0), nc))
        ratio = rng.uniform(0.3, 0.95, nc)
        if only is not None and i not in only:
            continue
        sm = meshmod.kuhn_cube(n, 0.15, seed=JITTER_SEED + i)
        assert sm.n_cells == nc
        out.append(_make(f"batch{i}", sm, sn, ng, st[:, None],
                         (ratio * st)[:, None], _uniform_q(sm.n_cells, ng),
                         dict(method="krylov", iter_max=300, tol=1e-10,
                              restart=20)))
    return out


def batch_sizes(count: int = 48, seed: int = BATCH_SEED, n_range=(10, 45),
                angle_orders=(4, 6, 8, 12), groups=(1, 4, 16, 32)):
    """(hexes per side, angles, groups) of every member of ``batch`` without
    building a mesh."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(rng.integers(n_range[0], n_range[1] + 1))
        sn = int(rng.choice(angle_orders))
        ng = int(rng.choice(groups))
        rng.uniform(size=6 * n ** 3)     # sigma_t draws
        rng.uniform(size=6 * n ** 3)     # sigma_s / sigma_t draws
        out.append((n, 2 * sn * sn, ng))
    return out


def gold(smesh: meshmod.SerialMesh) -> Problem:
    """The reference's input-gold.deck on cube-208.smesh (hat source)."""
    p = _make("gold", smesh, 8, 2, 10.0, 5.0, None,
              dict(method="si", iter_max=100, tol=1e-10))
    return p


def hat_source(p: Problem) -> np.ndarray:
    """Angle-dependent manufactured source of the reference on [0,100]^3
    (src/Problem.cc:47-82), layout [cell][angle][vertex][group]."""
    xyz = p.smesh.coords[p.smesh.cell_nodes] - 50.0          # (nc, 4, 3)
    r = np.linalg.norm(xyz, axis=2)                           # (nc, 4)
    with np.errstate(divide="ignore", invalid="ignore"):
        d = xyz / (30.0 * r[..., None])
    stream = -(d[:, None, :, 0] * p.omega[None, :, None, 0]) \
        - d[:, None, :, 1] * p.omega[None, :, None, 1] \
        - d[:, None, :, 2] * p.omega[None, :, None, 2]
    absorb = (p.sigma_t[:, 0] - p.sigma_s[:, 0])[:, None] * (1.0 - r / 30.0)
    val = np.where((r <= 30.0)[:, None, :], stream + absorb[:, None, :], 0.0)
    return np.ascontiguousarray(
        np.broadcast_to(val[..., None], (p.nc, p.na, 4, p.ng)))


def hat_l2_error(smesh: meshmod.SerialMesh, psi: np.ndarray) -> float:
    """L2 relative error of the cell-average psi against the hat solution
    1 - r/30 (r <= 30, else 0) at the cell centroid of [0,100]^3, over all
    (cell, angle, group): the reference's end-of-run check
    (src/Problem.cc:91-134, printed by src/Main.cc:340-343)."""
    psi = np.asarray(psi, dtype=np.float64)
    xyz = (smesh.coords[smesh.cell_nodes] - 50.0).sum(axis=1) / 4.0
    c = np.sqrt((xyz * xyz).sum(axis=1))                       # (nc,)
    act = np.where(c <= 30.0, 1.0 - c / 30.0, 0.0)
    val = psi.sum(axis=2) / 4.0                                # (nc, na, ng)
    diff = val - act[:, None, None]
    return float(np.sqrt((diff * diff).sum() / (val * val).sum()))

