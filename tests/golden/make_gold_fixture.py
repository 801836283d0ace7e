"""Derive ``gold_cube208.npz`` from the reference's regression answer.

The reference's only golden vector is test/regression/gold.psi: the angular
flux of input-gold.deck (snOrder 8, 2 groups, sigma_t 10, sigma_s 5, hat
source, source iteration to 1e-10, NoPivot) on util/cube-208.smesh, stored as
"Tycho 2 Psi Output" v1 (src/PsiData.cc:46-106): 32-byte name, 4 x uint64
{version, cells, angles, groups}, then psi[cell][angle][vertex][group].

This script keeps a small, sufficient part of it:
  * psi of group 0 at every 8th angle (the file's two groups are identical),
  * whole-file statistics (L1 and Linf norms, sum, sum of squares),
  * the scalar flux phi[cell][vertex] of group 0, accumulated angle by angle
    with the reference's own quadrature weights (oracle/_ref build).

Usage (in a container holding /root/reference):
    make -C oracle ref && python tests/golden/make_gold_fixture.py
"""
import ctypes
import os

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def main() -> None:
    raw = np.fromfile(os.path.join(REF, "test/regression/gold.psi"))
    header = raw[:8]
    name = header[:4].tobytes().rstrip(b"\0")
    version, nc, na, ng = header[4:].view(np.uint64)
    assert name == b"Tycho 2 Psi Output" and version == 1
    psi = raw[8:].reshape(int(nc), int(na), 4, int(ng))
    lib = ctypes.CDLL(os.path.join(ROOT, "oracle/_ref/libref_quadrature.so"))
    q = [np.empty(int(na)) for _ in range(4)]
    dp = ctypes.POINTER(ctypes.c_double)
    got = lib.ref_quadrature(8, *[a.ctypes.data_as(dp) for a in q])
    assert got == na
    wt = q[3]
    phi = np.zeros((int(nc), 4))
    for a in range(int(na)):
        phi += psi[:, a, :, 0] * wt[a]
    angles = np.arange(0, int(na), 8)
    np.savez_compressed(
        os.path.join(HERE, "gold_cube208.npz"),
        n_cells=int(nc), n_angles=int(na), n_groups=int(ng),
        sn_order=8, sigma_t=10.0, sigma_s=5.0, iter_max=100, err_max=1e-10,
        groups_identical=bool(np.array_equal(psi[..., 0], psi[..., 1])),
        psi_angles=angles,
        psi_g0=np.ascontiguousarray(psi[:, :, :, 0][:, angles, :]),
        psi_l1=np.abs(psi).sum(), psi_linf=np.abs(psi).max(),
        psi_sum=psi.sum(), psi_sumsq=(psi * psi).sum(),
        phi_g0=phi, ref_weights=wt,
        ref_omega=np.stack(q[:3], axis=1))


if __name__ == "__main__":
    main()
