"""Chebyshev-Legendre product quadrature of the reference
(src/Quadrature.cc:150-244).

Order N gives 2 N^2 directions indexed ``i * 2N + k``: polar cosine
mu = x_i (the i-th Gauss-Legendre node on [-1, 1]), azimuth
phi_k = (2 pi)/(2N) (k + 1/2), xi = sqrt(1 - mu^2) cos phi_k,
eta = sqrt(1 - mu^2) sin phi_k, weight w = g_i (2 pi)/(2N), so that the
weights sum to 4 pi.  The Gauss-Legendre nodes and weights are rounded to 16
decimals as in the reference's tables; here they are computed by Newton's
method in extended precision (numpy longdouble) before rounding.
"""
from __future__ import annotations

import functools
import math

import numpy as np

SUPPORTED = tuple(range(2, 31, 2))


def _legendre(n: int, x):
    p0 = np.longdouble(1)
    p1 = x
    for k in range(2, n + 1):
        p0, p1 = p1, ((2 * k - 1) * x * p1 - (k - 1) * p0) / k
    dp = n * (x * p1 - p0) / (x * x - 1)
    return p1, dp


@functools.lru_cache(maxsize=None)
def gauss_legendre(n: int):
    """Ascending nodes and weights, each rounded to 16 decimals."""
    nodes, weights = [], []
    for i in range(n):
        x = np.longdouble(math.cos(math.pi * (i + 0.75) / (n + 0.5)))
        for _ in range(100):
            p, dp = _legendre(n, x)
            dx = p / dp
            x = x - dx
            if abs(dx) < np.longdouble(1e-19):
                break
        _, dp = _legendre(n, x)
        w = np.longdouble(2) / ((1 - x * x) * dp * dp)
        nodes.append(float(np.format_float_positional(x, precision=16,
                                                      unique=False)))
        weights.append(float(np.format_float_positional(w, precision=16,
                                                        unique=False)))
    order = np.argsort(nodes)
    return np.array(nodes)[order], np.array(weights)[order]


def product_quadrature(sn: int):
    """(omega[na][3], weights[na]) with na = 2 sn^2."""
    if sn not in SUPPORTED:
        raise ValueError(f"quadrature order {sn} not supported "
                         f"(even orders 2..30)")
    gn, gw = gauss_legendre(sn)
    na = 2 * sn * sn
    omega = np.empty((na, 3))
    wt = np.empty(na)
    for i in range(sn):
        s = math.sqrt(1.0 - gn[i] * gn[i])
        for k in range(2 * sn):
            phi = (2.0 * math.pi) / (2 * sn) * (k + 0.5)
            idx = i * 2 * sn + k
            omega[idx] = (s * math.cos(phi), s * math.sin(phi), gn[i])
            wt[idx] = gw[i] * (2.0 * math.pi) / (2 * sn)
    return omega, wt
