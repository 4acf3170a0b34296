"""Independent McMurchie-Davidson reference for tiny inputs (test helper).

Uses Hermite-Gaussian expansion coefficients E_t^{ij} and Hermite Coulomb
integrals R_{tuv}^n (Helgaker, Jorgensen & Olsen, Molecular Electronic-Structure
Theory, sections 9.5 and 9.9) with the Boys function from mpmath's incomplete
gamma function.  It shares no recurrence with the Obara-Saika/HGP code it
checks, so agreement pins both the arithmetic and the conventions
(component order, Libint normalisation).
"""
from __future__ import annotations

import itertools
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


def _hermite_R_array(L, alpha, PQ):
    """R_{tuv}^0 for t + u + v <= L as an (L+1)^3 array (zero outside)."""
    R = hermite_R(L, L, L, alpha, PQ)
    out = np.zeros((L + 1, L + 1, L + 1))
    for t in range(L + 1):
        for u in range(L + 1 - t):
            for v in range(L + 1 - t - u):
                out[t, u, v] = R(t, u, v, 0)
    return out


def eri_block_md(la, a, A, lb, b, B, lc, c, C, ld, d, D):
    """Same primitive block as eri_primitive, vectorised over the Hermite
    sums: (ab|cd) = pref * sum E^ab_tuv (-1)^(tau+nu+phi) E^cd_{tau nu phi}
    R_{t+tau, u+nu, v+phi}."""
    A, B, C, D = (np.asarray(x, float) for x in (A, B, C, D))
    p, q = a + b, c + d
    P = (a * A + b * B) / p
    Q = (c * C + d * D) / q
    alpha = p * q / (p + q)
    L = la + lb + lc + ld
    R = _hermite_R_array(L, alpha, P - Q)
    n1, n2 = la + lb + 1, lc + ld + 1
    t1, t2 = np.arange(n1), np.arange(n2)
    idx = t1[:, None] + t2[None, :]
    # H[t, u, v, tau, nu, phi] = R[t+tau, u+nu, v+phi]
    H = R[idx[:, None, None, :, None, None], idx[None, :, None, None, :, None], idx[None, None, :, None, None, :]]

    def evec(l1, l2, e1, e2, X1, X2, comps1, comps2, n, sign):
        Ek = [{(i, j): np.pad(np.array(hermite_E(i, j, e1, e2, X1[k] - X2[k])), (0, n - i - j - 1))
               for i in range(l1 + 1) for j in range(l2 + 1)} for k in range(3)]
        out = np.zeros((len(comps1), len(comps2), n, n, n))
        s = (-1.0) ** np.arange(n) if sign else np.ones(n)
        for i1, c1 in enumerate(comps1):
            for i2, c2 in enumerate(comps2):
                ex, ey, ez = (Ek[k][(c1[k], c2[k])] * s for k in range(3))
                out[i1, i2] = ex[:, None, None] * ey[None, :, None] * ez[None, None, :]
        return out

    Eab = evec(la, lb, a, b, A, B, cart_components(la), cart_components(lb), n1, False)
    Ecd = evec(lc, ld, c, d, C, D, cart_components(lc), cart_components(ld), n2, True)
    pref = 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q))
    return pref * np.einsum("abtuv,tuvxyz,cdxyz->abcd", Eab, H, Ecd, optimize=True)


def _shift_index(l):
    """Cartesian component tuple -> index within shell l."""
    return {c: i for i, c in enumerate(cart_components(l))}


def eri_deriv1_contracted(shells):
    """First derivatives d/dX_k of the Libint-normalised block w.r.t. the four
    centres: array (4, 3, na, nb, nc, nd), from the Gaussian identity
    d/dA_k x^a exp(-alpha |r-A|^2) = 2 alpha x^(a+1_k) ... - a_k x^(a-1_k) ...
    applied to primitive MD blocks of the raised and lowered shells."""
    norms = [normalize(l, e, c) for (l, e, c, _) in shells]
    ls = [s[0] for s in shells]
    nfun = [len(cart_components(l)) for l in ls]
    out = np.zeros((4, 3) + tuple(nfun))
    exps = [s[1] for s in shells]
    cen = [np.asarray(s[3], float) for s in shells]
    for i0, e0 in enumerate(exps[0]):
        for i1, e1 in enumerate(exps[1]):
            for i2, e2 in enumerate(exps[2]):
                for i3, e3 in enumerate(exps[3]):
                    w = norms[0][i0] * norms[1][i1] * norms[2][i2] * norms[3][i3]
                    ee = (e0, e1, e2, e3)
                    for k in range(4):
                        def blk(lk):
                            lv = list(ls)
                            lv[k] = lk
                            return eri_block_md(lv[0], ee[0], cen[0], lv[1], ee[1], cen[1], lv[2], ee[2], cen[2],
                                                lv[3], ee[3], cen[3])
                        up = blk(ls[k] + 1)
                        dn = blk(ls[k] - 1) if ls[k] > 0 else None
                        iu = _shift_index(ls[k] + 1)
                        idn = _shift_index(ls[k] - 1) if ls[k] > 0 else None
                        for f, comp in enumerate(cart_components(ls[k])):
                            for x in range(3):
                                cu = list(comp)
                                cu[x] += 1
                                val = 2 * ee[k] * np.take(up, iu[tuple(cu)], axis=k)
                                if comp[x] > 0:
                                    cd = list(comp)
                                    cd[x] -= 1
                                    val = val - comp[x] * np.take(dn, idn[tuple(cd)], axis=k)
                                sl = [slice(None)] * 4
                                sl[k] = f
                                out[(k, x) + tuple(sl)] += w * val
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


def eri_deriv2_contracted(shells):
    """Second derivatives with respect to pairs of the 12 centre coordinates
    (centre-major, then x, y, z; pairs i <= j in lexicographic order): array
    (78, na, nb, nc, nd).  The Gaussian shift identity applied twice -- the
    second operator acts on the shifted functions the first produced -- to
    McMurchie-Davidson blocks of the shifted classes."""
    norms = [normalize(l, e, c) for (l, e, c, _) in shells]
    ls = [s[0] for s in shells]
    comps = [cart_components(l) for l in ls]
    out = np.zeros((78,) + tuple(len(c) for c in comps))
    exps = [s[1] for s in shells]
    cen = [np.asarray(s[3], float) for s in shells]
    pairs = [(i, j) for i in range(12) for j in range(i, 12)]

    def op(terms, k, x, ee):
        res = []
        for coef, cv in terms:
            up = [list(c) for c in cv]
            up[k][x] += 1
            res.append((coef * 2 * ee[k], tuple(tuple(c) for c in up)))
            if cv[k][x] > 0:
                dn = [list(c) for c in cv]
                dn[k][x] -= 1
                res.append((-coef * cv[k][x], tuple(tuple(c) for c in dn)))
        return res

    for i0, e0 in enumerate(exps[0]):
        for i1, e1 in enumerate(exps[1]):
            for i2, e2 in enumerate(exps[2]):
                for i3, e3 in enumerate(exps[3]):
                    w = norms[0][i0] * norms[1][i1] * norms[2][i2] * norms[3][i3]
                    ee = (e0, e1, e2, e3)
                    cache = {}

                    def val(cv):
                        lv = tuple(sum(c) for c in cv)
                        if lv not in cache:
                            cache[lv] = eri_block_md(lv[0], e0, cen[0], lv[1], e1, cen[1], lv[2], e2, cen[2],
                                                     lv[3], e3, cen[3])
                        idx = tuple(_shift_index(lv[t])[cv[t]] for t in range(4))
                        return cache[lv][idx]

                    for f in itertools.product(*[range(len(c)) for c in comps]):
                        cv = tuple(comps[t][f[t]] for t in range(4))
                        for q, (I, J) in enumerate(pairs):
                            terms = op(op([(1.0, cv)], I // 3, I % 3, ee), J // 3, J % 3, ee)
                            out[(q,) + f] += w * sum(c * val(v) for c, v in terms)
    return out
