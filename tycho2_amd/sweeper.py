"""Python front end of the MI355X sweep engine (``libtycho2_sweep.so``).

``SweeperHIP`` mirrors the reference's sweeper plug-in interface
(src/SweeperAbstract.hh:47-70): ``sweep(psi, source, zero_psi_bound)``
overwrites ``psi`` with L^-1 ``source`` on host arrays in the reference
layout [cell][angle][vertex][group], and ``solve()`` runs the outer
iteration (source iteration or GMRES, src/SourceIteration.cc) on the device.
Every call goes through the C ABI declared in include/tycho2_sweep.h; there
is no fallback: a missing library or a failing call raises.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

import numpy as np

from . import build as _build

FP64 = 0
FP32 = 1
GAUSS_ORIGINAL = 0
GAUSS_NOPIVOT = 1
GAUSS_CRAMER_GLU = 2
GAUSS_CRAMER_INTEL = 3
ENGINE_LEVELS = 0
ENGINE_QUEUE = 1
ENGINE_HYBRID = 2
ORDER_CELL = 0
ORDER_FACES = 1
ERR_TIMEOUT = -7

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_LIB = None

# Every exported symbol of include/tycho2_sweep.h with its argument types.
SIGNATURES = {
    "t2_last_error": (ctypes.c_char_p, []),
    "t2_abi_version": (ctypes.c_int, []),
    "t2_mesh_geometry": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int64, _dp,
                                        _ip, _ip, _dp, _dp, _dp]),
    "t2_sweep_levels": (ctypes.c_int, [ctypes.c_int64, _ip, _dp, _dp,
                                       ctypes.c_int64, _dp, _i32p, _ip]),
    "t2_sweep_priorities": (ctypes.c_int, [ctypes.c_int64, _ip, _dp, _dp,
                                           ctypes.c_int64, _dp, ctypes.c_int,
                                           ctypes.c_int, _u64p]),
    "t2_priority_schedule": (ctypes.c_int, [
        ctypes.c_int64, _ip, _dp, _dp, ctypes.c_int64, _dp, ctypes.c_int,
        ctypes.c_int, ctypes.c_int64, ctypes.c_int, _ip, _ip, _ip]),
    "t2_sweeper_create": (ctypes.c_int, [
        ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_int64,
        ctypes.c_int64, _dp, _ip, _ip, _ip, ctypes.c_int64, _dp, _dp,
        ctypes.c_int64, _dp, _dp, ctypes.c_int, ctypes.c_int]),
    "t2_sweeper_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "t2_sweeper_info": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "t2_set_launch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int,
                                     ctypes.c_int]),
    "t2_set_schedule": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "t2_set_order": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "t2_set_dependency_tolerance": (ctypes.c_int, [ctypes.c_void_p,
                                                   ctypes.c_double]),
    "t2_sweep_levels_tol": (ctypes.c_int, [ctypes.c_int64, _ip, _dp, _dp,
                                           ctypes.c_int64, _dp,
                                           ctypes.c_double, _i32p, _ip]),
    "t2_set_items_per_lane": (ctypes.c_int, [ctypes.c_void_p,
                                             ctypes.c_int]),
    "t2_set_concurrent_batches": (ctypes.c_int, [ctypes.c_void_p,
                                                 ctypes.c_int]),
    "t2_set_priority": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int64]),
    "t2_set_engine": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int64]),
    "t2_set_queue_split": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64]),
    "t2_solve_refined": (ctypes.c_int, [
        ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.c_double,
        ctypes.c_int64, ctypes.c_void_p, _dp]),
    "t2_defect": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp]),
    "t2_krylov_apply": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp]),
    "t2_krylov_rhs": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "t2_krylov_apply_device": (ctypes.c_int, [ctypes.c_void_p,
                                              ctypes.c_void_p,
                                              ctypes.c_void_p]),
    "t2_sweep_host": (ctypes.c_int, [ctypes.c_void_p, _dp, _dp,
                                     ctypes.c_int]),
    "t2_set_source_isotropic": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "t2_set_source_angular": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "t2_solve_source_iteration": (ctypes.c_int, [
        ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, _ip, _dp]),
    "t2_solve_source_iteration_noalloc": (ctypes.c_int, [
        ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, _ip, _dp]),
    "t2_solve_krylov": (ctypes.c_int, [
        ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.c_int64,
        _ip, _dp]),
    "t2_sweep_step": (ctypes.c_int, [ctypes.c_void_p]),
    "t2_sweep_only": (ctypes.c_int, [ctypes.c_void_p]),
    "t2_get_phi": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "t2_get_psi": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "t2_set_phi": (ctypes.c_int, [ctypes.c_void_p, _dp]),
    "t2_get_psi_angle": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64,
                                        _dp]),
    "t2_set_alloc_fail": (ctypes.c_int, [ctypes.c_int64]),
    "t2_synchronize": (ctypes.c_int, [ctypes.c_void_p]),
    "t2_get_stream": (ctypes.c_int, [ctypes.c_void_p,
                                     ctypes.POINTER(ctypes.c_void_p)]),
    "t2_batch_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p),
                                       ctypes.POINTER(ctypes.c_void_p),
                                       ctypes.c_int]),
    "t2_batch_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "t2_batch_set_engine": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int,
                                           ctypes.c_int, ctypes.c_int64]),
    "t2_batch_set_concurrent": (ctypes.c_int, [ctypes.c_void_p,
                                               ctypes.c_int]),
    "t2_batch_set_launch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "t2_batch_sweep_step": (ctypes.c_int, [ctypes.c_void_p]),
    "t2_batch_sweep_only": (ctypes.c_int, [ctypes.c_void_p]),
    "t2_batch_info": (ctypes.c_int, [ctypes.c_void_p, _ip, _ip]),
    "t2_batch_solve_krylov": (ctypes.c_int, [
        ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.c_int64,
        _ip, _dp]),
    "t2_batch_solve_source_iteration": (ctypes.c_int, [
        ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, _ip, _dp]),
    "t2_batch_solve_refined": (ctypes.c_int, [
        ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.c_double,
        ctypes.c_int64, ctypes.c_void_p, _dp]),
    "t2_batch_synchronize": (ctypes.c_int, [ctypes.c_void_p]),
    "t2_batch_get_stream": (ctypes.c_int, [ctypes.c_void_p,
                                           ctypes.POINTER(ctypes.c_void_p)]),
}


class T2Info(ctypes.Structure):
    _fields_ = [("n_cells", ctypes.c_int64), ("n_angles", ctypes.c_int64),
                ("n_groups", ctypes.c_int64), ("n_steps", ctypes.c_int64),
                ("min_levels", ctypes.c_int64), ("n_lagged", ctypes.c_int64),
                ("psi_bytes", ctypes.c_int64),
                ("device_bytes", ctypes.c_int64),
                ("precision", ctypes.c_int32), ("gauss", ctypes.c_int32),
                ("sweeps", ctypes.c_int64), ("engine", ctypes.c_int32),
                ("angle_batches", ctypes.c_int32),
                ("items_per_block", ctypes.c_int32),
                ("groups_per_lane", ctypes.c_int32),
                ("chunk_rounds", ctypes.c_int32),
                ("queue_first_step", ctypes.c_int32),
                ("intra_angle_p", ctypes.c_int32),
                ("inter_angle_p", ctypes.c_int32),
                ("max_cells_per_step", ctypes.c_int64),
                ("concurrent_batches", ctypes.c_int32)]


class T2RefineInfo(ctypes.Structure):
    _fields_ = [("iters_sweep", ctypes.c_int64),
                ("iters_correction", ctypes.c_int64),
                ("sweeps", ctypes.c_int64),
                ("residual_sweep", ctypes.c_double),
                ("residual_correction", ctypes.c_double),
                ("correction_rel", ctypes.c_double)]


class T2Error(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


ERR_NOCONV = -6


class NoConvergence(T2Error):
    """T2_ERR_NOCONV: a solve stopped at its iteration limit above its
    tolerance.  The handle holds the last iterate; ``result`` is what the
    call returns on success (iteration history, counts, infos)."""
    result = None


def lib():
    """Load the engine library (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        path = _build.LIBPATH
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing; build it with "
                               "`python -c 'import __graft_entry__ as g; "
                               "g.build()'`")
        L = ctypes.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def library_path() -> str:
    return _build.LIBPATH


def _check(rc: int) -> None:
    if rc != 0:
        msg = lib().t2_last_error()
        cls = NoConvergence if rc == ERR_NOCONV else T2Error
        raise cls(rc, msg.decode() if msg else "")


def _solved(rc: int, result):
    """`result` of a solve call, or NoConvergence carrying it."""
    try:
        _check(rc)
    except NoConvergence as e:
        e.result = result
        raise
    return result


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def _ptr(a, kind=_dp):
    return a.ctypes.data_as(kind)


def set_alloc_fail(nth: int) -> None:
    """Testing hook: the nth device allocation from now fails with
    T2_ERR_NOMEM (0 disables)."""
    _check(lib().t2_set_alloc_fail(int(nth)))


def mesh_geometry(smesh, conn):
    """(volume, area, normal) of every cell, computed by the engine's host
    code exactly as the reference's TychoMesh constructor does."""
    nc = smesh.n_cells
    vol, area, nrm = np.empty(nc), np.empty((nc, 4)), np.empty((nc, 4, 3))
    xyz, cn, f2c = _f64(smesh.coords), _i64(smesh.cell_nodes), _i64(conn.f2c)
    _check(lib().t2_mesh_geometry(nc, smesh.n_nodes, _ptr(xyz),
                                  _ptr(cn, _ip), _ptr(f2c, _ip), _ptr(vol),
                                  _ptr(area), _ptr(nrm)))
    return vol, area, nrm


def sweep_levels(adj, area, normal, omega, dep_tol=None):
    """(levels[a][c], n_lagged) of the engine's per-angle schedule, with the
    graph traverser's dependency rule or (dep_tol >= 0) the precomputed
    schedule's."""
    adj, ar, nrm, om = _i64(adj), _f64(area), _f64(normal), _f64(omega)
    nc, na = adj.shape[0], om.shape[0]
    levels = np.empty((na, nc), dtype=np.int32)
    lag = ctypes.c_int64(0)
    if dep_tol is None:
        _check(lib().t2_sweep_levels(nc, _ptr(adj, _ip), _ptr(ar), _ptr(nrm),
                                     na, _ptr(om), _ptr(levels, _i32p),
                                     ctypes.byref(lag)))
    else:
        _check(lib().t2_sweep_levels_tol(
            nc, _ptr(adj, _ip), _ptr(ar), _ptr(nrm), na, _ptr(om),
            float(dep_tol), _ptr(levels, _i32p), ctypes.byref(lag)))
    return levels, lag.value


def sweep_priorities(adj, area, normal, omega, intra: int, inter: int):
    """The reference's traversal priorities prio[a][c] (uint64) on one rank
    for deck values intraAngleP / interAngleP."""
    adj, ar, nrm, om = _i64(adj), _f64(area), _f64(normal), _f64(omega)
    nc, na = adj.shape[0], om.shape[0]
    prio = np.empty((na, nc), dtype=np.uint64)
    _check(lib().t2_sweep_priorities(nc, _ptr(adj, _ip), _ptr(ar), _ptr(nrm),
                                     na, _ptr(om), int(intra), int(inter),
                                     _ptr(prio, _u64p)))
    return prio


def priority_schedule(adj, area, normal, omega, intra: int, inter: int,
                      max_cells_per_step: int, order: int = ORDER_CELL):
    """(items, step_off): the priority list schedule, items as
    angle * n_cells + cell in schedule order, steps items[off[i]:off[i+1]]."""
    adj, ar, nrm, om = _i64(adj), _f64(area), _f64(normal), _f64(omega)
    nc, na = adj.shape[0], om.shape[0]
    items = np.empty(nc * na, dtype=np.int64)
    off = np.empty(nc * na + 1, dtype=np.int64)
    n = ctypes.c_int64(0)
    _check(lib().t2_priority_schedule(
        nc, _ptr(adj, _ip), _ptr(ar), _ptr(nrm), na, _ptr(om), int(intra),
        int(inter), int(max_cells_per_step), int(order), _ptr(items, _ip),
        _ptr(off, _ip), ctypes.byref(n)))
    return items, off[:n.value + 1].copy()


class SweeperHIP:
    """Device sweeper for one mesh, quadrature and cross-section set."""

    def __init__(self, smesh, conn, omega, wt, ng, sigma_t, sigma_s,
                 precision: int = FP64, device: int = 0,
                 gauss: int = GAUSS_NOPIVOT):
        nc = smesh.n_cells
        self.nc, self.na, self.ng = nc, int(np.shape(omega)[0]), int(ng)
        st = _f64(np.broadcast_to(sigma_t, (nc, ng)))
        ss = _f64(np.broadcast_to(sigma_s, (nc, ng)))
        xyz, cn = _f64(smesh.coords), _i64(smesh.cell_nodes)
        adj, f2c = _i64(conn.adj_cell), _i64(conn.f2c)
        om, w = _f64(omega), _f64(wt)
        h = ctypes.c_void_p()
        _check(lib().t2_sweeper_create(
            ctypes.byref(h), device, nc, smesh.n_nodes, _ptr(xyz),
            _ptr(cn, _ip), _ptr(adj, _ip), _ptr(f2c, _ip), self.na, _ptr(om),
            _ptr(w), self.ng, _ptr(st), _ptr(ss), precision, int(gauss)))
        self._h = h

    @classmethod
    def from_problem(cls, p, precision: int = FP64, device: int = 0):
        s = cls(p.smesh, p.conn, p.omega, p.wt, p.ng, p.sigma_t, p.sigma_s,
                precision, device)
        if p.q_iso is not None:
            s.set_source_isotropic(p.q_iso)
        return s

    # -- lifetime
    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().t2_sweeper_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        if not self._h:
            raise T2Error(-4, "sweeper is closed")
        return self._h

    def info(self) -> dict:
        i = T2Info()
        _check(lib().t2_sweeper_info(self.handle, ctypes.byref(i)))
        return {k: getattr(i, k) for k, _ in T2Info._fields_}

    def set_launch(self, items_per_block: int = 0, groups_per_lane: int = 0):
        _check(lib().t2_set_launch(self.handle, items_per_block,
                                   groups_per_lane))

    def set_items_per_lane(self, items: int) -> None:
        _check(lib().t2_set_items_per_lane(self.handle, int(items)))

    def set_schedule(self, angle_batches: int) -> None:
        _check(lib().t2_set_schedule(self.handle, int(angle_batches)))

    def set_order(self, order: int) -> None:
        """ORDER_CELL or ORDER_FACES: item order within a step."""
        _check(lib().t2_set_order(self.handle, int(order)))

    def set_dependency_tolerance(self, tol: float) -> None:
        """Faces with |Omega.n| <= tol are not sweep dependencies and read
        the pre-sweep psi (SweepType OriginalTycho1/2 use 1e-10); tol < 0
        returns to the graph traverser's rule."""
        _check(lib().t2_set_dependency_tolerance(self.handle, float(tol)))

    def set_concurrent_batches(self, on: bool) -> None:
        """Level engine: angle batches on parallel streams (True) or one
        after the other (False, default)."""
        _check(lib().t2_set_concurrent_batches(self.handle, int(bool(on))))

    def set_priority(self, intra_angle_p: int, inter_angle_p: int = 0,
                     max_cells_per_step: int = 0) -> None:
        """Priority list schedule (deck keys intraAngleP, interAngleP,
        maxCellsPerStep) in place of the level sets; intra_angle_p = -1
        returns to level sets."""
        _check(lib().t2_set_priority(self.handle, int(intra_angle_p),
                                     int(inter_angle_p),
                                     int(max_cells_per_step)))

    def set_engine(self, engine: int, spin_limit: int = 0,
                   chunk_rounds: int = 0) -> None:
        """ENGINE_LEVELS (per-step launches in a hipGraph) or ENGINE_QUEUE
        (one persistent ready-queue launch); spin_limit bounds each queue
        wait in polls, chunk_rounds sets the queue's chunk size (0: library
        defaults)."""
        _check(lib().t2_set_engine(self.handle, int(engine), int(chunk_rounds),
                                   int(spin_limit)))

    def set_queue_split(self, first_step: int = -1) -> None:
        """ENGINE_HYBRID: first step run by the ready queue (-1: the thin
        tail, chosen by the library)."""
        _check(lib().t2_set_queue_split(self.handle, int(first_step)))

    # -- SweeperAbstract::sweep
    def sweep(self, psi: np.ndarray, source: np.ndarray,
              zero_psi_bound: bool = False) -> np.ndarray:
        """psi[:] = L^-1 source (host arrays, [cell][angle][vertex][group])."""
        shape = (self.nc, self.na, 4, self.ng)
        if psi.shape != shape or psi.dtype != np.float64 or \
                not psi.flags.c_contiguous:
            raise ValueError(f"psi must be a C-contiguous float64 {shape}")
        src = _f64(source).reshape(shape)
        _check(lib().t2_sweep_host(self.handle, _ptr(psi), _ptr(src),
                                   int(bool(zero_psi_bound))))
        return psi

    # -- fixed source and outer iterations
    def set_source_isotropic(self, q: Optional[np.ndarray]) -> None:
        if q is None:
            _check(lib().t2_set_source_isotropic(self.handle, None))
            return
        qq = _f64(q).reshape(self.nc, 4, self.ng)
        _check(lib().t2_set_source_isotropic(self.handle, _ptr(qq)))

    def set_source_angular(self, source: np.ndarray) -> None:
        src = _f64(source).reshape(self.nc, self.na, 4, self.ng)
        _check(lib().t2_set_source_angular(self.handle, _ptr(src)))

    def solve_source_iteration(self, iter_max=100, err_max=1e-10,
                               no_alloc=False):
        """Source iteration on the device; no_alloc=True calls the
        allocation-free entry point (same results)."""
        hist = np.zeros(max(int(iter_max), 1))
        nit = ctypes.c_int64(0)
        fn = (lib().t2_solve_source_iteration_noalloc if no_alloc
              else lib().t2_solve_source_iteration)
        rc = fn(self.handle, int(iter_max), float(err_max), ctypes.byref(nit),
                _ptr(hist))
        return _solved(rc, hist[:nit.value].copy())

    def solve_krylov(self, iter_max=100, rtol=1e-10, restart=30):
        hist = np.zeros(max(int(iter_max), 1))
        nit = ctypes.c_int64(0)
        rc = lib().t2_solve_krylov(self.handle, int(iter_max), float(rtol),
                                   int(restart), ctypes.byref(nit),
                                   _ptr(hist))
        return _solved(rc, hist[:nit.value].copy())

    def solve_refined(self, iter_max=300, rtol_sweep=1e-6,
                      rtol_correction=1e-5, restart=20, want_phi1=False):
        """FP32 sweeps + FP64 defect correction (FP32 handles).  Returns
        (info dict, phi after the FP32 stage or None)."""
        info = T2RefineInfo()
        phi1 = np.empty((self.nc, 4, self.ng)) if want_phi1 else None
        rc = lib().t2_solve_refined(
            self.handle, int(iter_max), float(rtol_sweep),
            float(rtol_correction), int(restart), ctypes.byref(info),
            _ptr(phi1) if want_phi1 else None)
        return _solved(rc, ({k: getattr(info, k)
                             for k, _ in T2RefineInfo._fields_}, phi1))

    def krylov_apply(self, x: np.ndarray) -> np.ndarray:
        """b = (I - D L^-1 M S) x for phi-shaped x [c][v][g] (the
        reference's lhsOperator); one sweep on the device."""
        p = _f64(x).reshape(self.nc, 4, self.ng)
        out = np.empty((self.nc, 4, self.ng))
        _check(lib().t2_krylov_apply(self.handle, _ptr(p), _ptr(out)))
        return out

    def krylov_apply_device(self, x_ptr: int, b_ptr: int) -> None:
        """b = (I - D L^-1 M S) x on device vectors (addresses of n_cells x 4
        x n_groups doubles on the handle's device), enqueued on the
        handle's stream; synchronize before reading b."""
        _check(lib().t2_krylov_apply_device(self.handle,
                                            ctypes.c_void_p(x_ptr),
                                            ctypes.c_void_p(b_ptr)))

    def krylov_rhs(self) -> np.ndarray:
        """b = D L^-1 q with the handle's source (one sweep)."""
        out = np.empty((self.nc, 4, self.ng))
        _check(lib().t2_krylov_rhs(self.handle, _ptr(out)))
        return out

    def defect(self, psi: Optional[np.ndarray] = None) -> np.ndarray:
        """Nodal density d[c][a][v][g] of r = q + M S phi - L psi in FP64
        (weak form V/20 (I + J) d = r), for the given psi (which becomes the
        handle's angular flux) or, with None, the handle's own."""
        out = np.empty((self.nc, self.na, 4, self.ng))
        p = None
        if psi is not None:
            p = _f64(psi).reshape(self.nc, self.na, 4, self.ng)
        _check(lib().t2_defect(self.handle, _ptr(p) if p is not None
                               else None, _ptr(out)))
        return out

    def sweep_step(self) -> None:
        _check(lib().t2_sweep_step(self.handle))

    def sweep_only(self) -> None:
        _check(lib().t2_sweep_only(self.handle))

    def synchronize(self) -> None:
        _check(lib().t2_synchronize(self.handle))

    def stream(self) -> int:
        s = ctypes.c_void_p()
        _check(lib().t2_get_stream(self.handle, ctypes.byref(s)))
        return s.value or 0

    # -- results
    def phi(self) -> np.ndarray:
        out = np.empty((self.nc, 4, self.ng))
        _check(lib().t2_get_phi(self.handle, _ptr(out)))
        return out

    def set_phi(self, phi: np.ndarray) -> None:
        p = _f64(phi).reshape(self.nc, 4, self.ng)
        _check(lib().t2_set_phi(self.handle, _ptr(p)))

    def psi_angle(self, angle: int) -> np.ndarray:
        out = np.empty((self.nc, 4, self.ng))
        _check(lib().t2_get_psi_angle(self.handle, int(angle), _ptr(out)))
        return out

    def psi(self) -> np.ndarray:
        out = np.empty((self.nc, self.na, 4, self.ng))
        _check(lib().t2_get_psi(self.handle, _ptr(out)))
        return out


class SweepBatch:
    """Several sweepers (any meshes, quadratures and group counts; one
    precision) swept together: merged per-step launches (ENGINE_LEVELS) or
    one persistent ready-queue launch for all of them (ENGINE_QUEUE).

    Members keep their own psi, phi and sources; each batch sweep gives every
    member exactly what its own sweep would.  Members must stay alive and
    keep their schedule and source kind while the batch is used (the
    library refuses a batch whose members changed)."""

    def __init__(self, members):
        self.members = list(members)
        arr = (ctypes.c_void_p * len(self.members))(
            *[m.handle.value for m in self.members])
        h = ctypes.c_void_p()
        _check(lib().t2_batch_create(ctypes.byref(h), arr, len(self.members)))
        self._h = h

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().t2_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        if not self._h:
            raise T2Error(-4, "batch is closed")
        return self._h

    def set_engine(self, engine: int, spin_limit: int = 0,
                   chunk_rounds: int = 0) -> None:
        _check(lib().t2_batch_set_engine(self.handle, int(engine),
                                         int(chunk_rounds), int(spin_limit)))

    def set_launch(self, block: int) -> None:
        """Threads per block of the level form's merged launches."""
        _check(lib().t2_batch_set_launch(self.handle, int(block)))

    def set_concurrent(self, chains: int) -> None:
        """Level form: independent merged schedules on parallel streams,
        up to `chains` per group count (1: one per group count), or 0 to
        interleave every launch on one stream (default)."""
        _check(lib().t2_batch_set_concurrent(self.handle, int(chains)))

    def sweep_step(self) -> None:
        _check(lib().t2_batch_sweep_step(self.handle))

    def sweep_only(self) -> None:
        _check(lib().t2_batch_sweep_only(self.handle))

    def synchronize(self) -> None:
        _check(lib().t2_batch_synchronize(self.handle))

    def stream(self) -> int:
        s = ctypes.c_void_p()
        _check(lib().t2_batch_get_stream(self.handle, ctypes.byref(s)))
        return s.value or 0

    def solve_krylov(self, iter_max: int = 200, rtol: float = 1e-10,
                     restart: int = 20):
        """Every member's GMRES solve in lock step; returns (iterations,
        seconds to converged phi) per member."""
        n = len(self.members)
        its = np.zeros(n, np.int64)
        sec = np.zeros(n)
        rc = lib().t2_batch_solve_krylov(self.handle, int(iter_max),
                                         float(rtol), int(restart),
                                         _ptr(its, _ip), _ptr(sec))
        return _solved(rc, (its, sec))

    def solve_source_iteration(self, iter_max: int = 1000,
                               err_max: float = 1e-10):
        n = len(self.members)
        its = np.zeros(n, np.int64)
        sec = np.zeros(n)
        rc = lib().t2_batch_solve_source_iteration(
            self.handle, int(iter_max), float(err_max), _ptr(its, _ip),
            _ptr(sec))
        return _solved(rc, (its, sec))

    def solve_refined(self, iter_max: int = 300, rtol_sweep: float = 1e-6,
                      rtol_correction: float = 1e-5, restart: int = 20):
        """Every FP32 member's refined solve (FP32 sweeps + FP64 defect
        correction), stage by stage in lock step; returns (info dict per
        member, seconds to the end of each member's correction)."""
        n = len(self.members)
        infos = (T2RefineInfo * n)()
        sec = np.zeros(n)
        rc = lib().t2_batch_solve_refined(
            self.handle, int(iter_max), float(rtol_sweep),
            float(rtol_correction), int(restart), infos, _ptr(sec))
        return _solved(rc, ([{k: getattr(x, k)
                              for k, _ in T2RefineInfo._fields_}
                             for x in infos], sec))

    def info(self) -> dict:
        n_steps, n_items = ctypes.c_int64(), ctypes.c_int64()
        _check(lib().t2_batch_info(self.handle, ctypes.byref(n_steps),
                                   ctypes.byref(n_items)))
        return {"n_steps": n_steps.value, "n_items": n_items.value}
