"""Independent McMurchie-Davidson reference for tiny inputs (test helper).

Uses Hermite-Gaussian expansion coefficients E_t^{ij} and Hermite Coulomb
integrals R_{tuv}^n (Helgaker, Jorgensen & Olsen, Molecular Electronic-Structure
Theory, sections 9.5 and 9.9) with the Boys function from mpmath's incomplete
gamma function.  It shares no recurrence with the Obara-Saika/HGP code it
checks, so agreement pins both the arithmetic and the conventions
(component order, Libint normalisation).
"""
from __future__ import annotations

import math
from functools import lru_cache

import mpmath
import numpy as np

mpmath.mp.dps = 30


def boys_mp(m: int, T: float) -> float:
    if T < 1e-300:
        return 1.0 / (2 * m + 1)
    T = mpmath.mpf(T)
    h = mpmath.mpf(m) + mpmath.mpf(1) / 2
    return float(mpmath.gammainc(h, 0, T) / (2 * T ** h))


def cart_components(l):
    out = []
    for i in range(l + 1):
        lx = l - i
        for j in range(i + 1):
            out.append((lx, i - j, j))
    return out


def _df(n):
    r = 1.0
    while n > 1:
        r *= n
        n -= 2
    return r


def normalize(l, exps, coefs):
    """x^l primitive normalisation, then unit self-overlap of the contraction."""
    exps = np.asarray(exps, float)
    c = np.asarray(coefs, float) * np.array(
        [math.sqrt(2 ** l * (2 * a) ** (l + 1.5) / (math.pi ** 1.5 * _df(2 * l - 1))) for a in exps])
    norm = 0.0
    for i, a in enumerate(exps):
        for j, b in enumerate(exps):
            norm += _df(2 * l - 1) * math.pi ** 1.5 * c[i] * c[j] / (2 ** l * (a + b) ** (l + 1.5))
    return c / math.sqrt(norm)


def hermite_E(i, j, a, b, XAB):
    """E_t^{ij} for t = 0..i+j (1D Hermite expansion of a Gaussian product)."""
    p = a + b
    mu = a * b / p
    XPA = -b / p * XAB
    XPB = a / p * XAB

    @lru_cache(None)
    def E(i, j, t):
        if t < 0 or t > i + j:
            return 0.0
        if i == 0 and j == 0:
            return math.exp(-mu * XAB * XAB) if t == 0 else 0.0
        if i > 0:
            return (1 / (2 * p)) * E(i - 1, j, t - 1) + XPA * E(i - 1, j, t) + (t + 1) * E(i - 1, j, t + 1)
        return (1 / (2 * p)) * E(i, j - 1, t - 1) + XPB * E(i, j - 1, t) + (t + 1) * E(i, j - 1, t + 1)

    return [E(i, j, t) for t in range(i + j + 1)]


def hermite_R(tmax, umax, vmax, alpha, PQ):
    """R_{tuv}^0 by the Hermite recursion from Boys values."""
    X, Y, Z = PQ
    T = alpha * (X * X + Y * Y + Z * Z)
    nmax = tmax + umax + vmax
    Fv = [boys_mp(n, T) for n in range(nmax + 1)]

    @lru_cache(None)
    def R(t, u, v, n):
        if t < 0 or u < 0 or v < 0:
            return 0.0
        if t == 0 and u == 0 and v == 0:
            return (-2 * alpha) ** n * Fv[n]
        if t > 0:
            return (t - 1) * R(t - 2, u, v, n + 1) + X * R(t - 1, u, v, n + 1)
        if u > 0:
            return (u - 1) * R(t, u - 2, v, n + 1) + Y * R(t, u - 1, v, n + 1)
        return (v - 1) * R(t, u, v - 2, n + 1) + Z * R(t, u, v - 1, n + 1)

    return R


def eri_primitive(la, a, A, lb, b, B, lc, c, C, ld, d, D):
    """Cartesian primitive block (na, nb, nc, nd), unnormalised monomials."""
    A, B, C, D = map(np.asarray, (A, B, C, D))
    p, q = a + b, c + d
    P = (a * A + b * B) / p
    Q = (c * C + d * D) / q
    alpha = p * q / (p + q)
    R = hermite_R(la + lb + lc + ld, la + lb + lc + ld, la + lb + lc + ld, alpha, P - Q)
    Eab = [[hermite_E(i, j, a, b, A[k] - B[k]) for j in range(lb + 1) for i in range(la + 1)] for k in range(3)]
    Ecd = [[hermite_E(i, j, c, d, C[k] - D[k]) for j in range(ld + 1) for i in range(lc + 1)] for k in range(3)]

    def Eget(E, l1, i, j, k):
        return E[k][j * (l1 + 1) + i]

    ca, cb, cc, cd = (cart_components(l) for l in (la, lb, lc, ld))
    out = np.zeros((len(ca), len(cb), len(cc), len(cd)))
    pref = 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q))
    for ia, (ax, ay, az) in enumerate(ca):
        for ib, (bx, by, bz) in enumerate(cb):
            ex = Eget(Eab, la, ax, bx, 0)
            ey = Eget(Eab, la, ay, by, 1)
            ez = Eget(Eab, la, az, bz, 2)
            for ic, (cx, cy, cz) in enumerate(cc):
                for idd, (dx, dy, dz) in enumerate(cd):
                    fx = Eget(Ecd, lc, cx, dx, 0)
                    fy = Eget(Ecd, lc, cy, dy, 1)
                    fz = Eget(Ecd, lc, cz, dz, 2)
                    s = 0.0
                    for t in range(ax + bx + 1):
                        for u in range(ay + by + 1):
                            for v in range(az + bz + 1):
                                e1 = ex[t] * ey[u] * ez[v]
                                if e1 == 0.0:
                                    continue
                                for tau in range(cx + dx + 1):
                                    for nu in range(cy + dy + 1):
                                        for phi in range(cz + dz + 1):
                                            sgn = -1.0 if (tau + nu + phi) % 2 else 1.0
                                            s += e1 * sgn * fx[tau] * fy[nu] * fz[phi] * R(t + tau, u + nu, v + phi, 0)
                    out[ia, ib, ic, idd] = pref * s
    return out


def eri_contracted(shells):
    """shells: 4 tuples (l, exps, raw_coefs, centre).  Libint-normalised block."""
    blocks = None
    norms = [normalize(l, e, c) for (l, e, c, _) in shells]
    (la, ea, _, A), (lb, eb, _, B), (lc, ec, _, C), (ld, ed, _, D) = shells
    for i, a in enumerate(ea):
        for j, b in enumerate(eb):
            for k, c in enumerate(ec):
                for m, d in enumerate(ed):
                    w = norms[0][i] * norms[1][j] * norms[2][k] * norms[3][m]
                    blk = w * eri_primitive(la, a, A, lb, b, B, lc, c, C, ld, d, D)
                    blocks = blk if blocks is None else blocks + blk
    return blocks
