"""Shared inputs and comparison helpers for the parity tests."""
import functools

import numpy as np

import workloads as W

# Support radius used for parity: 3.75 h (the configs' "40-50 neighbours").
# At 2.5 h the reference's own RKPM solve reports a pivot collapse on the
# straight-crease sheet (tested as an error-path case).
RHO_FACTOR = 3.75


@functools.lru_cache(maxsize=None)
def sheet(name: str):
    if name == "straight21":
        return W.config1(rho_factor=RHO_FACTOR)
    if name == "straight21_rho2.5":
        return W.config1(rho_factor=2.5)
    if name == "straight21_rho6":
        return W.config1(rho_factor=6.0)
    if name == "straight21_rho8.5":
        return W.config1(rho_factor=8.5)
    if name == "straight11":
        return W.config1(rho_factor=RHO_FACTOR, n_side=11)
    if name == "straight21_bspline":
        return W.config1(rho_factor=RHO_FACTOR, kernel=1)
    if name == "arc41":
        return W.config2(rho_factor=RHO_FACTOR, n_side=41, nseg=64)
    if name == "arc201":
        return W.config2(rho_factor=RHO_FACTOR)
    raise KeyError(name)


def pair_gp(off):
    return np.repeat(np.arange(off.size - 1), np.diff(off))


def normwise(a, r, g, ng):
    """Per-GP max_I |a - r| / max_I |r| over the listed arrays."""
    num = np.zeros(ng); den = np.zeros(ng)
    for x, y in zip(a, r):
        np.maximum.at(num, g, np.abs(x - y))
        np.maximum.at(den, g, np.abs(y))
    return num / np.where(den > 0, den, 1.0)


EPS = np.finfo(np.float64).eps


def moment_cond(gx, gy, X, Y, rho, off, ids, psi, order):
    """cond_2 of each GP's moment matrix M = sum_I w(s_I) p_I p_I^T over its
    visible neighbours (psi != 0), with the quartic window and the basis of
    the given order in xi = (x - X_I) / rho: the conditioning that bounds the
    round-off two correct evaluations of the shape functions can differ by."""
    g = pair_gp(off)
    xi = (gx[g] - X[ids]) / rho
    eta = (gy[g] - Y[ids]) / rho
    s = np.sqrt(xi * xi + eta * eta)
    w = np.where(s < 1, 1 - 6 * s ** 2 + 8 * s ** 3 - 3 * s ** 4, 0.0) * (psi != 0)
    P = [xi ** (d - k) * eta ** k for d in range(order + 1) for k in range(d + 1)]
    ng = off.size - 1
    M = np.zeros((ng, len(P), len(P)))
    for a in range(len(P)):
        for b in range(a, len(P)):
            M[:, a, b] = M[:, b, a] = np.bincount(g, w * P[a] * P[b], ng)
    return np.linalg.cond(M)


def sf_tolerance(cond, floor=1e-12):
    """Per-GP normwise bars for (psi, gradient, Hessian): max(floor, C eps
    cond(M)) with C = 4 / 16 / 64."""
    return tuple(np.maximum(floor, C * EPS * cond) for C in (4, 16, 64))
