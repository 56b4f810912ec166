"""Independent ERI algorithm for tests: McMurchie-Davidson (Hermite
expansion) over unnormalised primitive Cartesian Gaussians, with the Boys
function from the regularised incomplete gamma function.

Shares no code or recursion with the Obara-Saika engines under test, so an
error in either (a dropped term, a wrong sign, a transposed centre) shows up
as a mismatch.  Pure Python; meant for a few hundred primitive quartets.
"""
import math

import numpy as np
from scipy.special import gammainc, gamma


def boys(n, t):
    if t < 1e-10:
        return 1.0 / (2 * n + 1) - t / (2 * n + 3)
    a = n + 0.5
    return gamma(a) * gammainc(a, t) / (2.0 * t ** a)


def hermite_e(i, j, a, b, xa, xb):
    """E^{ij}_t for t = 0..i+j along one axis (two centres at xa, xb)."""
    p = a + b
    mu = a * b / p
    xp = (a * xa + b * xb) / p
    xpa, xpb = xp - xa, xp - xb
    E = {}
    E[(0, 0, 0)] = math.exp(-mu * (xa - xb) ** 2)

    def get(ii, jj, t):
        if t < 0 or t > ii + jj or ii < 0 or jj < 0:
            return 0.0
        return E[(ii, jj, t)]

    for ii in range(i + 1):
        for jj in range(j + 1):
            if ii == 0 and jj == 0:
                continue
            for t in range(ii + jj + 1):
                if ii > 0:
                    v = get(ii - 1, jj, t - 1) / (2 * p) + xpa * get(ii - 1, jj, t) + (t + 1) * get(ii - 1, jj, t + 1)
                else:
                    v = get(ii, jj - 1, t - 1) / (2 * p) + xpb * get(ii, jj - 1, t) + (t + 1) * get(ii, jj - 1, t + 1)
                E[(ii, jj, t)] = v
    return [E[(i, j, t)] for t in range(i + j + 1)]


def hermite_r(L, alpha, pq):
    """R_{tuv} for t+u+v <= L."""
    x, y, z = pq
    T = alpha * (x * x + y * y + z * z)
    R = {}
    for n in range(L + 1):
        R[(0, 0, 0, n)] = (-2 * alpha) ** n * boys(n, T)

    def get(t, u, v, n):
        if t < 0 or u < 0 or v < 0:
            return 0.0
        return R[(t, u, v, n)]

    for tot in range(1, L + 1):
        for n in range(L - tot + 1):
            for t in range(tot + 1):
                for u in range(tot - t + 1):
                    v = tot - t - u
                    if t > 0:
                        val = (t - 1) * get(t - 2, u, v, n + 1) + x * get(t - 1, u, v, n + 1)
                    elif u > 0:
                        val = (u - 1) * get(t, u - 2, v, n + 1) + y * get(t, u - 1, v, n + 1)
                    else:
                        val = (v - 1) * get(t, u, v - 2, n + 1) + z * get(t, u, v - 1, n + 1)
                    R[(t, u, v, n)] = val
    return lambda t, u, v: R[(t, u, v, 0)]


def eri_prim(la, a, A, lb, b, B, lc, c, C, ld, d, D):
    """(ab|cd) for unnormalised primitives; la.. are (lx,ly,lz) tuples."""
    p, q = a + b, c + d
    P = [(a * A[k] + b * B[k]) / p for k in range(3)]
    Q = [(c * C[k] + d * D[k]) / q for k in range(3)]
    alpha = p * q / (p + q)
    L = sum(la) + sum(lb) + sum(lc) + sum(ld)
    R = hermite_r(L, alpha, [P[k] - Q[k] for k in range(3)])
    Eab = [hermite_e(la[k], lb[k], a, b, A[k], B[k]) for k in range(3)]
    Ecd = [hermite_e(lc[k], ld[k], c, d, C[k], D[k]) for k in range(3)]
    s = 0.0
    for t, e1 in enumerate(Eab[0]):
        for u, e2 in enumerate(Eab[1]):
            for v, e3 in enumerate(Eab[2]):
                eab = e1 * e2 * e3
                if eab == 0.0:
                    continue
                for tau, f1 in enumerate(Ecd[0]):
                    for nu, f2 in enumerate(Ecd[1]):
                        for phi, f3 in enumerate(Ecd[2]):
                            sgn = -1.0 if (tau + nu + phi) % 2 else 1.0
                            s += eab * sgn * f1 * f2 * f3 * R(t + tau, u + nu, v + phi)
    return 2.0 * math.pi ** 2.5 / (p * q * math.sqrt(p + q)) * s


def contracted_eri(basis, i, j, k, l):
    """(ij|kl) over the reference's contracted functions (CC rows)."""
    cc = basis.cc

    def prims(c):
        idx = np.nonzero(cc[c])[0]
        return [(int(pp), cc[c, pp]) for pp in idx]

    def geo(pp):
        return ((int(basis.pax[pp]), int(basis.pay[pp]), int(basis.paz[pp])), basis.zt[pp],
                (basis.pgsx[pp], basis.pgsy[pp], basis.pgsz[pp]))

    tot = 0.0
    for p1, c1 in prims(i):
        g1 = geo(p1)
        for p2, c2 in prims(j):
            g2 = geo(p2)
            for p3, c3 in prims(k):
                g3 = geo(p3)
                for p4, c4 in prims(l):
                    g4 = geo(p4)
                    tot += c1 * c2 * c3 * c4 * eri_prim(*g1, *g2, *g3, *g4)
    return tot
