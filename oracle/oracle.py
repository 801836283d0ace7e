"""ctypes front end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this module.  It loads ``oracle/liboracle.so`` (built by
``oracle/Makefile``) and wraps the plain-C restatement of the reference's
FP64 sweep path with numpy arrays in the reference's layouts.
"""
from __future__ import annotations

import ctypes
import dataclasses
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

NOPIVOT = 1
ORIGINAL = 0
CRAMER_GLU = 2
CRAMER_INTEL = 3

_d = ctypes.POINTER(ctypes.c_double)
_i = ctypes.POINTER(ctypes.c_int64)


class _OrMesh(ctypes.Structure):
    _fields_ = [("nc", ctypes.c_int64), ("adj", _i), ("c2f", _i),
                ("nbrv", _i), ("volume", _d), ("area", _d), ("normal", _d)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing: run `make -C oracle`")
        _LIB = ctypes.CDLL(path)
    return _LIB


def _p(a, kind):
    return a.ctypes.data_as(_d if kind == "d" else _i)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


@dataclasses.dataclass
class Mesh:
    """TychoMesh-equivalent arrays built by the oracle from a serial mesh."""
    adj: np.ndarray
    f2c: np.ndarray
    c2f: np.ndarray
    nbrv: np.ndarray
    volume: np.ndarray
    area: np.ndarray
    normal: np.ndarray
    cell_nodes: np.ndarray
    coords: np.ndarray

    @property
    def nc(self) -> int:
        return int(self.volume.shape[0])

    def cstruct(self) -> _OrMesh:
        return _OrMesh(self.nc, _p(self.adj, "i"), _p(self.c2f, "i"),
                       _p(self.nbrv, "i"), _p(self.volume, "d"),
                       _p(self.area, "d"), _p(self.normal, "d"))


def build_mesh(smesh) -> Mesh:
    """Restatement of TychoMesh construction for a one-partition mesh."""
    nc, nf, nn = smesh.n_cells, smesh.n_faces, smesh.n_nodes
    cf, cn = _i64(smesh.cell_faces), _i64(smesh.cell_nodes)
    fc, fn = _i64(smesh.face_cells), _i64(smesh.face_nodes)
    xyz = _f64(smesh.coords)
    adj = np.empty((nc, 4), np.int64)
    f2c = np.empty((nc, 4, 3), np.int64)
    c2f = np.empty((nc, 4, 4), np.int64)
    nbrv = np.empty((nc, 4, 3), np.int64)
    vol = np.empty(nc)
    area = np.empty((nc, 4))
    nrm = np.empty((nc, 4, 3))
    rc = lib().or_mesh_build(
        ctypes.c_int64(nc), ctypes.c_int64(nf), ctypes.c_int64(nn),
        _p(cf, "i"), _p(cn, "i"), _p(fc, "i"), _p(fn, "i"), _p(xyz, "d"),
        _p(adj, "i"), _p(f2c, "i"), _p(c2f, "i"), _p(nbrv, "i"),
        _p(vol, "d"), _p(area, "d"), _p(nrm, "d"))
    if rc != 0:
        raise ValueError(f"or_mesh_build failed ({rc})")
    return Mesh(adj, f2c, c2f, nbrv, vol, area, nrm, cn, xyz)


def quadrature(sn: int):
    """(omega (na,3), weights (na,)) of the reference's product set."""
    na = 2 * sn * sn
    xi, eta, mu, wt = (np.empty(na) for _ in range(4))
    if lib().or_quadrature(ctypes.c_int(sn), _p(xi, "d"), _p(eta, "d"),
                           _p(mu, "d"), _p(wt, "d")) != 0:
        raise ValueError(f"quadrature order {sn} not supported")
    return np.stack([xi, eta, mu], axis=1), wt


def cell_solve(volume, abar, sigma_t, q, upwind, method=NOPIVOT):
    out = np.empty(4)
    ab, qq = _f64(abar), _f64(q)
    up = _f64(upwind).reshape(4, 4)
    lib().or_cell_solve(ctypes.c_int(method), ctypes.c_double(volume),
                        _p(ab, "d"), ctypes.c_double(sigma_t), _p(qq, "d"),
                        _p(up, "d"), _p(out, "d"))
    return out


def _sig(sig, nc, ng):
    s = np.asarray(sig, dtype=np.float64)
    if s.ndim == 0:
        s = np.full((nc, ng), float(s))
    elif s.ndim == 1 and s.shape[0] == ng and nc != ng:
        s = np.broadcast_to(s[None, :], (nc, ng))
    elif s.ndim == 1:
        s = np.broadcast_to(s[:, None], (nc, ng))
    return _f64(s)


class dependency_tolerance:
    """Context manager: the oracle's sweeps inside use the precomputed
    schedule's dependency rule with this tolerance (faces with
    -tol <= Omega.n <= 0 lagged from the start; SweepSchedule.cc:57)."""

    def __init__(self, tol: float):
        self.tol = float(tol)

    def __enter__(self):
        lib().or_set_dependency_tolerance(ctypes.c_double(self.tol))
        return self

    def __exit__(self, *exc):
        lib().or_set_dependency_tolerance(ctypes.c_double(-1.0))


def sweep(mesh: Mesh, omega, ng, sigma_t, source, psi=None, method=NOPIVOT,
          nthreads=0, angles=None):
    """psi[c][a][v][g] = L^-1 source.  Returns (psi, n_lagged)."""
    om = _f64(omega)
    na = om.shape[0]
    st = _sig(sigma_t, mesh.nc, ng)
    src = _f64(source).reshape(mesh.nc, na, 4, ng)
    if psi is None:
        psi = np.zeros((mesh.nc, na, 4, ng))
    lag = ctypes.c_int64(0)
    m = mesh.cstruct()
    a0, a1 = (0, na) if angles is None else angles
    rc = lib().or_sweep_angles(
        ctypes.byref(m), ctypes.c_int64(na), _p(om, "d"), ctypes.c_int64(ng),
        _p(st, "d"), _p(src, "d"), _p(psi, "d"), ctypes.c_int(method),
        ctypes.c_int(nthreads), ctypes.c_int64(a0), ctypes.c_int64(a1),
        ctypes.byref(lag))
    if rc != 0:
        raise RuntimeError("or_sweep failed")
    return psi, lag.value


def defect(mesh: Mesh, omega, ng, sigma_t, source, psi):
    """Nodal density d[c][a][v][g] of the defect r = source - L psi (weak
    form V/20 (I + J) d = r), every interior face reading the current psi."""
    om = _f64(omega)
    na = om.shape[0]
    st = _sig(sigma_t, mesh.nc, ng)
    src = _f64(source).reshape(mesh.nc, na, 4, ng)
    ps = _f64(psi).reshape(mesh.nc, na, 4, ng)
    out = np.empty((mesh.nc, na, 4, ng))
    m = mesh.cstruct()
    lib().or_defect(ctypes.byref(m), ctypes.c_int64(na), _p(om, "d"),
                    ctypes.c_int64(ng), _p(st, "d"), _p(src, "d"),
                    _p(ps, "d"), _p(out, "d"))
    return out


def psi_to_phi(psi, wt):
    nc, na, _, ng = psi.shape
    phi = np.empty((nc, 4, ng))
    w = _f64(wt)
    lib().or_psi_to_phi(ctypes.c_int64(nc), ctypes.c_int64(na),
                        ctypes.c_int64(ng), _p(w, "d"),
                        _p(_f64(psi), "d"), _p(phi, "d"))
    return phi


def source_iteration(mesh: Mesh, omega, wt, ng, sigma_t, sigma_s, source,
                     iter_max=100, err_max=1e-10, method=NOPIVOT,
                     nthreads=0):
    """Returns (psi, phi, err_history)."""
    om, w = _f64(omega), _f64(wt)
    na = om.shape[0]
    st, ss = _sig(sigma_t, mesh.nc, ng), _sig(sigma_s, mesh.nc, ng)
    src = _f64(source).reshape(mesh.nc, na, 4, ng)
    psi = np.zeros((mesh.nc, na, 4, ng))
    phi = np.empty((mesh.nc, 4, ng))
    hist = np.zeros(max(int(iter_max), 1))
    nit = ctypes.c_int64(0)
    m = mesh.cstruct()
    rc = lib().or_source_iteration(
        ctypes.byref(m), ctypes.c_int64(na), _p(om, "d"), _p(w, "d"),
        ctypes.c_int64(ng), _p(st, "d"), _p(ss, "d"), _p(src, "d"),
        ctypes.c_int64(iter_max), ctypes.c_double(err_max),
        ctypes.c_int(method), ctypes.c_int(nthreads), _p(psi, "d"),
        _p(phi, "d"), _p(hist, "d"), ctypes.byref(nit))
    if rc != 0:
        raise RuntimeError("or_source_iteration failed")
    return psi, phi, hist[:nit.value].copy()


def krylov(mesh: Mesh, omega, wt, ng, sigma_t, sigma_s, source,
           iter_max=100, rtol=1e-10, restart=30, method=NOPIVOT, nthreads=0):
    """Returns (psi, phi, relative-residual history)."""
    om, w = _f64(omega), _f64(wt)
    na = om.shape[0]
    st, ss = _sig(sigma_t, mesh.nc, ng), _sig(sigma_s, mesh.nc, ng)
    src = _f64(source).reshape(mesh.nc, na, 4, ng)
    psi = np.zeros((mesh.nc, na, 4, ng))
    phi = np.empty((mesh.nc, 4, ng))
    hist = np.zeros(max(int(iter_max), 1))
    nit = ctypes.c_int64(0)
    m = mesh.cstruct()
    rc = lib().or_krylov(
        ctypes.byref(m), ctypes.c_int64(na), _p(om, "d"), _p(w, "d"),
        ctypes.c_int64(ng), _p(st, "d"), _p(ss, "d"), _p(src, "d"),
        ctypes.c_int64(iter_max), ctypes.c_double(rtol),
        ctypes.c_int64(restart), ctypes.c_int(method),
        ctypes.c_int(nthreads), _p(psi, "d"), _p(phi, "d"), _p(hist, "d"),
        ctypes.byref(nit))
    if rc != 0:
        raise RuntimeError("or_krylov failed")
    return psi, phi, hist[:nit.value].copy()


def gmres_dense(a, b, restart=30, iter_max=1000, rtol=1e-12):
    a, b = _f64(a), _f64(b)
    n = b.shape[0]
    x = np.empty(n)
    nit = ctypes.c_int64(0)
    lib().or_gmres_dense(ctypes.c_int64(n), _p(a, "d"), _p(b, "d"),
                         _p(x, "d"), ctypes.c_int64(restart),
                         ctypes.c_int64(iter_max), ctypes.c_double(rtol),
                         ctypes.byref(nit))
    return x, nit.value


def hat_source(mesh: Mesh, omega, ng, sigma_t_cell, sigma_s_cell):
    om = _f64(omega)
    na = om.shape[0]
    xi, eta, mu = (_f64(om[:, k]) for k in range(3))
    st = _f64(np.broadcast_to(np.asarray(sigma_t_cell, float), (mesh.nc,)))
    ss = _f64(np.broadcast_to(np.asarray(sigma_s_cell, float), (mesh.nc,)))
    out = np.empty((mesh.nc, na, 4, ng))
    lib().or_hat_source(ctypes.c_int64(mesh.nc), _p(mesh.cell_nodes, "i"),
                        _p(mesh.coords, "d"), ctypes.c_int64(na),
                        _p(xi, "d"), _p(eta, "d"), _p(mu, "d"),
                        ctypes.c_int64(ng), _p(st, "d"), _p(ss, "d"),
                        _p(out, "d"))
    return out


def isotropic_source(nc, na, ng, q_cell_group):
    """Angle-expanded [c][a][v][g] source from an isotropic q[c][g]."""
    q = np.asarray(q_cell_group, dtype=np.float64).reshape(nc, 1, 1, ng)
    return np.ascontiguousarray(np.broadcast_to(q, (nc, na, 4, ng)))
