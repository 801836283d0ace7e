"""Angular-flux output file and relative-difference metric.

``write_psi`` / ``read_psi`` use the reference's "Tycho 2 Psi Output" v1
layout (src/PsiData.cc:46-106): a 32-byte name padded with zeros, uint64
{version=1, cells, angles, groups}, then psi[cell][angle][vertex][group] as
doubles in global cell order.  ``rel_diff`` is the metric of the reference's
regression script (test/diff.py:21-53): L1 and Linf differences relative to
the second argument, both required below the tolerance.
"""
from __future__ import annotations

import struct

import numpy as np

PSI_NAME = b"Tycho 2 Psi Output"


def write_psi(path: str, psi: np.ndarray) -> None:
    nc, na, nv, ng = psi.shape
    if nv != 4:
        raise ValueError("psi must be [cell][angle][4][group]")
    with open(path, "wb") as f:
        f.write(PSI_NAME.ljust(32, b"\0"))
        f.write(struct.pack("<4Q", 1, nc, na, ng))
        f.write(np.ascontiguousarray(psi, dtype=np.float64).tobytes())


def read_psi(path: str) -> np.ndarray:
    raw = np.fromfile(path, dtype=np.float64)
    if raw.size < 8 or raw[:4].tobytes().rstrip(b"\0") != PSI_NAME:
        raise ValueError(f"{path}: not a Tycho 2 psi file")
    version, nc, na, ng = (int(v) for v in raw[4:8].view(np.uint64))
    if version != 1 or raw.size != 8 + nc * na * 4 * ng:
        raise ValueError(f"{path}: bad psi header")
    return raw[8:].reshape(nc, na, 4, ng)


def rel_diff(a: np.ndarray, b: np.ndarray):
    """(L1, Linf) of a - b relative to b."""
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    if a.size != b.size:
        raise ValueError("vector lengths differ")
    d = a - b
    return (float(np.abs(d).sum() / np.abs(b).sum()),
            float(np.abs(d).max() / np.abs(b).max()))


def diff_between_groups(psi: np.ndarray) -> float:
    """max |psi_g - psi_0| over groups g > 0, relative to max |psi_0|, for
    psi[cell][angle][vertex][group]: the reference's end-of-run check
    (src/Util.cc:55-79, printed by src/Main.cc:341-344)."""
    psi = np.asarray(psi, dtype=np.float64)
    p0 = psi[..., 0]
    d = np.abs(psi[..., 1:] - p0[..., None]).max() if psi.shape[-1] > 1 \
        else 0.0
    return float(d / np.abs(p0).max())

