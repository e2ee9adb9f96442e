"""Independent brute-force ERIs by McMurchie-Davidson (textbook, tiny inputs).

Used only to pin the oracle's VRR/HRR, normalisation and ``cons`` handling: a
different algorithm (Hermite expansion, McMurchie & Davidson 1978) evaluated
with scipy's incomplete gamma for the Boys function, so a dropped term, wrong
sign or transposed index in the oracle's Obara-Saika path shows up here.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import gammainc, gamma


def boys(m: int, T: float) -> float:
    if T < 1e-12:
        return 1.0 / (2 * m + 1) - T / (2 * m + 3)
    return gamma(m + 0.5) * gammainc(m + 0.5, T) / (2.0 * T ** (m + 0.5))


def hermite_E(i, j, t, Qx, a, b):
    """E_t^{ij} for the 1-D overlap distribution, Qx = A_x - B_x."""
    p = a + b
    q = a * b / p
    if t < 0 or t > i + j:
        return 0.0
    if i == j == t == 0:
        return math.exp(-q * Qx * Qx)
    if j == 0:
        return (1 / (2 * p)) * hermite_E(i - 1, j, t - 1, Qx, a, b) - (q * Qx / a) * hermite_E(i - 1, j, t, Qx, a, b) \
            + (t + 1) * hermite_E(i - 1, j, t + 1, Qx, a, b)
    return (1 / (2 * p)) * hermite_E(i, j - 1, t - 1, Qx, a, b) + (q * Qx / b) * hermite_E(i, j - 1, t, Qx, a, b) \
        + (t + 1) * hermite_E(i, j - 1, t + 1, Qx, a, b)


def hermite_R(t, u, v, n, p, PCx, PCy, PCz, RPC):
    T = p * RPC * RPC
    if t == u == v == 0:
        return (-2 * p) ** n * boys(n, T)
    if t == u == 0:
        val = 0.0
        if v > 1:
            val += (v - 1) * hermite_R(t, u, v - 2, n + 1, p, PCx, PCy, PCz, RPC)
        return val + PCz * hermite_R(t, u, v - 1, n + 1, p, PCx, PCy, PCz, RPC)
    if t == 0:
        val = 0.0
        if u > 1:
            val += (u - 1) * hermite_R(t, u - 2, v, n + 1, p, PCx, PCy, PCz, RPC)
        return val + PCy * hermite_R(t, u - 1, v, n + 1, p, PCx, PCy, PCz, RPC)
    val = 0.0
    if t > 1:
        val += (t - 1) * hermite_R(t - 2, u, v, n + 1, p, PCx, PCy, PCz, RPC)
    return val + PCx * hermite_R(t - 1, u, v, n + 1, p, PCx, PCy, PCz, RPC)


def prim_eri(a, lmn1, A, b, lmn2, B, c, lmn3, C, d, lmn4, D):
    """Unnormalised primitive (ab|cd)."""
    l1, m1, n1 = lmn1
    l2, m2, n2 = lmn2
    l3, m3, n3 = lmn3
    l4, m4, n4 = lmn4
    p, q = a + b, c + d
    alpha = p * q / (p + q)
    P = (a * A + b * B) / p
    Q = (c * C + d * D) / q
    RPQ = np.linalg.norm(P - Q)
    val = 0.0
    for t in range(l1 + l2 + 1):
        e1 = hermite_E(l1, l2, t, A[0] - B[0], a, b)
        for u in range(m1 + m2 + 1):
            e2 = hermite_E(m1, m2, u, A[1] - B[1], a, b)
            for v in range(n1 + n2 + 1):
                e3 = hermite_E(n1, n2, v, A[2] - B[2], a, b)
                for tau in range(l3 + l4 + 1):
                    f1 = hermite_E(l3, l4, tau, C[0] - D[0], c, d)
                    for nu in range(m3 + m4 + 1):
                        f2 = hermite_E(m3, m4, nu, C[1] - D[1], c, d)
                        for phi in range(n3 + n4 + 1):
                            f3 = hermite_E(n3, n4, phi, C[2] - D[2], c, d)
                            val += e1 * e2 * e3 * f1 * f2 * f3 * (-1) ** (tau + nu + phi) * \
                                hermite_R(t + tau, u + nu, v + phi, 0, alpha, *(P - Q), RPQ)
    return 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q)) * val


def contracted(fx, i):
    """(center, lmn, exps, coefs) of basis function i (0-based) from shell tables.

    The ERI path uses, for every component of a shell's Qshell, the contraction
    coefficients of the function at ksumtype+Qstart-type offset (QUICK's
    Xcoeff selection) times cons.
    """
    ns = fx.nshell
    for s in range(ns):
        if fx.ksumtype[s] - 1 <= i < fx.ksumtype[s + 1] - 1:
            break
    L = None
    for t in range(fx.Qstart[s], fx.Qfinal[s] + 1):
        if fx.ksumtype[s] - 1 + fx.Qsbasis[s, t] <= i <= fx.ksumtype[s] - 1 + fx.Qfbasis[s, t]:
            L = t
    npr = fx.kprim[s]
    cf = fx.gccoeff[:npr, fx.ksumtype[s] - 1 + L] * fx.cons[i]
    ex = fx.gcexpo[:npr, fx.ksumtype[s] - 1 + L]
    return fx.xyz[:, fx.katom[s] - 1], tuple(int(v) for v in fx.KLMN[:, i]), ex, cf


def eri(fx, i, j, k, l):
    fs = [contracted(fx, x) for x in (i, j, k, l)]
    tot = 0.0
    (A, la, ea, ca), (B, lb, eb, cb), (Cc, lc, ec, cc), (Dd, ld, ed, cd) = fs
    for a, x1 in zip(ea, ca):
        for b, x2 in zip(eb, cb):
            for c, x3 in zip(ec, cc):
                for d, x4 in zip(ed, cd):
                    tot += x1 * x2 * x3 * x4 * prim_eri(a, la, A, b, lb, B, c, lc, Cc, d, ld, Dd)
    return tot
