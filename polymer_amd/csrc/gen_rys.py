"""Generate the Rys-quadrature pieces of libpolyeri.

rys_tables.h -- roots and weights of the Rys quadrature, n = 1..RYS_NMAX:
  the n-point Gauss rule for the weight exp(-T t^2) on t in [0, 1] written
  in u = t^2, i.e. nodes u_r and weights w_r with
      sum_r w_r u_r^m = F_m(T),   m = 0 .. 2n - 1
  (F_m the Boys function).  For T < RYS_TMAX(n) = 32 + 8 n the nodes and
  weights are Chebyshev series of degree RYS_DEG on the unit intervals
  [i, i + 1); above it they are the half-range Gauss-Hermite rule scaled by
  T: u_r = x_r^2 / T, w_r = W_r / sqrt(T), whose error there is below
  1e-17 relative.  The reference values are computed here in 90-digit
  decimal arithmetic: the moments F_m(T) by their convergent series, the
  monic recurrence coefficients by the Chebyshev algorithm, the nodes by
  Sturm-sequence bisection of the orthogonal polynomial, the weights as
  Christoffel numbers.  Double-precision evaluation of the fitted series
  reproduces the nodes and weights to about 3e-15 relative.

rys_generated.h -- per shell-quartet class and lane variant, one primitive
  quartet's contribution to a lane's share of the contracted integrals (see
  emit_class below).

Run by __graft_entry__.build(), which regenerates rys_generated.h (files
are rewritten only when their text changes); the tables, whose 90-digit
reference values take minutes, only with --tables or when rys_tables.h is
missing.
"""
import math
import os
import sys
from decimal import Decimal as D, getcontext

getcontext().prec = 90

RYS_NMAX = 5
RYS_DEG = 11          # Chebyshev degree per unit interval (12 coefficients)

PI = D("3.14159265358979323846264338327950288419716939937510582097494459230781640628620899862803482534211706798")


def rys_tmax(n):
    return 32 + 8 * n


def boys_hp(mmax, T):
    """F_0..F_mmax(T) = e^-T sum_k (2T)^k / ((2m+1)(2m+3)...(2m+2k+1))."""
    T = D(T)
    e = (-T).exp()
    out = []
    for m in range(mmax + 1):
        term = D(1) / (2 * m + 1)
        s = term
        k = 1
        while True:
            term = term * 2 * T / (2 * m + 2 * k + 1)
            s += term
            if term < s * D(10) ** -88:
                break
            k += 1
        out.append(e * s)
    return out


def recurrence(mom, n):
    """Monic three-term recurrence p_{k+1} = (x - a_k) p_k - b_k p_{k-1} from
    the ordinary moments mom[0 .. 2n-1] (Chebyshev algorithm)."""
    alpha = [D(0)] * n
    beta = [D(0)] * n
    sig_prev = [D(0)] * (2 * n)
    sig = list(mom[:2 * n])
    alpha[0] = mom[1] / mom[0]
    beta[0] = mom[0]
    for k in range(1, n):
        new = [D(0)] * (2 * n)
        for l in range(k, 2 * n - k):
            new[l] = sig[l + 1] - alpha[k - 1] * sig[l] - beta[k - 1] * sig_prev[l]
        alpha[k] = new[k + 1] / new[k] - sig[k] / sig[k - 1]
        beta[k] = new[k] / sig[k - 1]
        sig_prev, sig = sig, new
    return alpha, beta


def _poly(x, alpha, beta, n):
    p = [D(1), x - alpha[0]]
    for k in range(1, n):
        p.append((x - alpha[k]) * p[k] - beta[k] * p[k - 1])
    return p


def gauss_from_moments(mom, n, hi):
    """Nodes (ascending, in (0, hi)) and weights of the n-point Gauss rule."""
    alpha, beta = recurrence(mom, n)

    def below(x):   # number of nodes < x: n minus the sign changes of p_0..p_n
        p = _poly(x, alpha, beta, n)
        return n - sum(1 for i in range(n) if (p[i] > 0) != (p[i + 1] > 0))

    nodes = []
    for r in range(n):
        lo, up = D(0), D(hi)
        for _ in range(300):
            mid = (lo + up) / 2
            if below(mid) > r:
                up = mid
            else:
                lo = mid
        nodes.append((lo + up) / 2)
    norms = [beta[0]]
    for k in range(1, n):
        norms.append(norms[-1] * beta[k])
    weights = []
    for x in nodes:
        p = _poly(x, alpha, beta, n)
        weights.append(1 / sum(p[k] * p[k] / norms[k] for k in range(n)))
    return nodes, weights


def rys_hp(T, n):
    return gauss_from_moments(boys_hp(2 * n - 1, T), n, 1)


def hermite_half_hp(n):
    """Nodes y_r = x_r^2 and weights W_r with sum W_r y_r^m = Gamma(m + 1/2) / 2."""
    mom = []
    g = PI.sqrt()
    for m in range(2 * n):
        mom.append(g / 2)
        g = g * (D(m) + D("0.5"))
    return gauss_from_moments(mom, n, 4 * n + 10)


def cheb_interval(a, n):
    """Chebyshev coefficients (RYS_DEG + 1 each) of the n nodes and n weights
    on T in [a, a + 1]."""
    N = RYS_DEG + 1
    xs = [math.cos(math.pi * (j + 0.5) / N) for j in range(N)]
    samples = [rys_hp(D(a) + (1 + D(repr(x))) / 2, n) for x in xs]
    out = []
    for kind in (0, 1):
        for r in range(n):
            for k in range(N):
                s = sum(D(repr(math.cos(k * math.acos(xs[j])))) * samples[j][kind][r] for j in range(N)) * 2 / N
                if k == 0:
                    s /= 2
                out.append(float(s))
    # layout per root r: 12 node coefficients then 12 weight coefficients
    res = []
    for r in range(n):
        res += out[r * N:(r + 1) * N] + out[(n + r) * N:(n + r + 1) * N]
    return res


def write_tables(path):
    lines = ["// Generated by gen_rys.py -- do not edit.",
             "// Rys quadrature nodes u_r and weights w_r (sum_r w_r u_r^m = F_m(T)).",
             "// RYS_TABLE_n: for T in [i, i + 1), i < RYS_TMAX(n), and root r, the",
             "// RYS_DEG + 1 Chebyshev coefficients of u_r then those of w_r in the",
             "// variable x = 2 (T - i) - 1.  RYS_HERM_n: y_r then W_r of the half-range",
             "// Gauss-Hermite rule (u_r = y_r / T, w_r = W_r / sqrt(T) for T >= RYS_TMAX(n)).",
             "#pragma once",
             f"#define RYS_NMAX {RYS_NMAX}",
             f"#define RYS_DEG {RYS_DEG}",
             "#define RYS_TMAX(n) (32 + 8 * (n))"]
    for n in range(1, RYS_NMAX + 1):
        vals = []
        for i in range(rys_tmax(n)):
            vals += cheb_interval(i, n)
        lines.append(f"#define RYS_TABLE_{n} {{ \\")
        for k in range(0, len(vals), 4):
            lines.append("    " + ", ".join(repr(v) for v in vals[k:k + 4]) + ", \\")
        lines.append("}")
        y, w = hermite_half_hp(n)
        lines.append(f"#define RYS_HERM_{n} {{ " + ", ".join(repr(float(v)) for v in y + w) + " }")
    text = "\n".join(lines) + "\n"
    if not os.path.exists(path) or open(path).read() != text:
        with open(path, "w") as f:
            f.write(text)




# ---------------------------------------------------------------------------
# Per-class code.  A lane of the Rys kernel owns a share of one contracted
# quartet's integrals: the components of some shells ("fixed" shells) are
# held at one value, the other shells' components all run.  The tuples of
# fixed components fall into orbits under permutations of the x, y, z axes;
# per orbit ("variant") the code is written once for a representative tuple
# in a local frame, and the lane of orbit member m works in the frame whose
# local axis l is the global axis sigma_m(l) -- same instructions, permuted
# inputs.  Every integral of the quartet belongs to exactly one (variant,
# member).
# ---------------------------------------------------------------------------
import itertools

PERMS = [p for p in itertools.permutations(range(3))]
RYS_CLASSES = {
    # class: shells whose components a lane fixes (0 a, 1 b, 2 c, 3 d)
    (2, 1, 1, 1): (2,),
    (2, 1, 2, 1): (1, 3),
    (2, 2, 1, 1): (2, 3),
    (2, 2, 2, 0): (0,),
    (2, 2, 2, 1): (0, 3),
    (2, 2, 2, 2): (0, 1),
}


def comps(l):
    out = []
    for i in range(l, -1, -1):
        for k in range(0, l - i + 1):
            out.append((i, l - i - k, k))
    return out


def permute(c, p):
    """Global exponents of the local component c in the frame whose local
    axis l is global axis p[l]."""
    g = [0, 0, 0]
    for l in range(3):
        g[p[l]] = c[l]
    return tuple(g)


def variants(cls):
    ls = cls
    fixed = RYS_CLASSES[cls]
    tuples = list(itertools.product(*[comps(ls[s]) for s in fixed]))
    seen, out = set(), []
    for t in tuples:
        if t in seen:
            continue
        members = []
        for p in PERMS:
            img = tuple(permute(c, p) for c in t)
            if img not in [m[0] for m in members]:
                members.append((img, p))
        for img, _ in members:
            seen.add(img)
        out.append((t, [m[1] for m in members]))
    assert sum(len(m) for _, m in out) == len(tuples)
    return fixed, out


def emit_variant(cls, vi, rep, perms, fixed):
    la, lb, lc, ld = cls
    L = la + lb + lc + ld
    nroot = L // 2 + 1
    shells = []
    for s, l in enumerate(cls):
        shells.append([rep[fixed.index(s)]] if s in fixed else comps(l))
    outs = list(itertools.product(*shells))
    ci = {}
    for s, l in enumerate(cls):
        ci[s] = [comps(l).index(c) for c in shells[s]]
    # demand-driven 2D nodes
    code = []
    memo = {}

    def g2(d, i, k):
        key = ("g", d, i, k)
        if key in memo:
            return memo[key]
        nm = f"g{d}_{i}_{k}"
        if i == 0 and k == 0:
            return "W" if d == 2 else "ONE"

        def term(c, x):    # c * x with the unit factor folded away
            return c if x == "ONE" else f"{c} * {x}"

        def coef(n, v):
            return v if n == 1 else f"real({n}) * {v}"
        if k == 0:
            e = term(f"C00_{d}", g2(d, i - 1, 0))
            if i >= 2:
                e = f"fmar({coef(i - 1, 'B10')}, {g2(d, i - 2, 0)}, {e})"
        else:
            e = term(f"D00_{d}", g2(d, i, k - 1))
            if k >= 2:
                e = f"fmar({coef(k - 1, 'B01')}, {g2(d, i, k - 2)}, {e})"
            if i >= 1:
                e = f"fmar({coef(i, 'B00')}, {g2(d, i - 1, k - 1)}, {e})"
        code.append(f"        const real {nm} = {e};")
        memo[key] = nm
        return nm

    def hrr(d, a, b, c, e):
        key = ("h", d, a, b, c, e)
        if key in memo:
            return memo[key]
        if b == 0 and e == 0:
            return g2(d, a, c)
        nm = f"h{d}_{a}{b}{c}{e}"
        if b > 0:
            x = f"fmar(AB_{d}, {hrr(d, a, b - 1, c, e)}, {hrr(d, a + 1, b - 1, c, e)})"
        else:
            x = f"fmar(CD_{d}, {hrr(d, a, 0, c, e - 1)}, {hrr(d, a, 0, c + 1, e - 1)})"
        code.append(f"        const real {nm} = {x};")
        memo[key] = nm
        return nm

    acc_lines = []
    for o, (ca, cb, cc, cd) in enumerate(outs):
        f = [hrr(d, ca[d], cb[d], cc[d], cd[d]) for d in range(3)]
        terms = [t for t in f if t != "ONE"]
        if len(terms) == 0:
            acc_lines.append(f"        acc[{o}] += ONE;")
        elif len(terms) == 1:
            acc_lines.append(f"        acc[{o}] += {terms[0]};")
        elif len(terms) == 2:
            acc_lines.append(f"        acc[{o}] = fmar({terms[0]}, {terms[1]}, acc[{o}]);")
        else:
            acc_lines.append(f"        acc[{o}] = fmar({terms[0]} * {terms[1]}, {terms[2]}, acc[{o}]);")
    nm = f"RysVar<{la}, {lb}, {lc}, {ld}, {vi}>"
    pl = ", ".join("{" + ", ".join(str(x) for x in p) + "}" for p in perms)
    lists = []
    for s, key in enumerate("ABCD"):
        lists.append(f"    static constexpr int N{key} = {len(ci[s])};")
        lists.append(f"    static constexpr int C{key}[{len(ci[s])}] = {{{', '.join(str(x) for x in ci[s])}}};")
    text = [f"template <>", f"struct {nm} {{",
            f"    static constexpr int NOUT = {len(outs)}, NROOT = {nroot}, NMEMB = {len(perms)};",
            f"    // local axis l of member m is global axis PERM[m][l]",
            f"    static constexpr int PERM[{len(perms)}][3] = {{{pl}}};",
            "    // local component indices each shell runs over (outputs in nested a, b, c, d order)"]
    text += lists
    sw = ["    static POLY_HD void perm(int m, int& p0, int& p1, int& p2) {", "        switch (m) {"]
    for m, pm in enumerate(perms):
        lbl = "default" if m == len(perms) - 1 else f"case {m}"
        sw.append(f"        {lbl}: p0 = {pm[0]}; p1 = {pm[1]}; p2 = {pm[2]}; break;")
    sw += ["        }", "    }"]
    text += sw
    text += ["    // adds one Rys node's contribution: W = weight x prefactor; the",
             "    // recursion coefficients and geometry are in the local frame; real:",
             "    // the recursion's and the accumulators' type (double, or float for",
             "    // the quartets of the mixed-precision lists)",
             "    template <typename real>",
             "    static POLY_HD void node(real W, real B00, real B10, real B01, real C00_0, real C00_1, real C00_2,",
             "                             real D00_0, real D00_1, real D00_2, real AB_0, real AB_1, real AB_2,",
             "                             real CD_0, real CD_1, real CD_2, real* __restrict__ acc) {",
             "        constexpr real ONE = real(1);"]
    text += code + acc_lines + ["    }", "};"]
    return "\n".join(text), len(outs), nroot, len(perms)


def write_classes(path):
    parts = ["// Generated by gen_rys.py -- do not edit.",
             "// Rys-quadrature code per shell-quartet class and lane variant (see gen_rys.py).",
             "#pragma once", "",
             "POLY_HD double fmar(double a, double b, double c) { return __builtin_fma(a, b, c); }",
             "POLY_HD float fmar(float a, float b, float c) { return __builtin_fmaf(a, b, c); }", "",
             "template <int LA, int LB, int LC, int LD, int V>", "struct RysVar;", ""]
    table = []
    for cls in RYS_CLASSES:
        fixed, vs = variants(cls)
        for vi, (rep, perms) in enumerate(vs):
            t, nout, nroot, nm = emit_variant(cls, vi, rep, perms, fixed)
            parts.append(t)
            parts.append("")
        table.append(f"template <> struct RysClass<{cls[0]}, {cls[1]}, {cls[2]}, {cls[3]}> {{ static constexpr int NVAR = {len(vs)}; }};")
    parts.insert(6, "template <int LA, int LB, int LC, int LD>\nstruct RysClass { static constexpr int NVAR = 0; };\n")
    parts += table
    text = "\n".join(parts) + "\n"
    if not os.path.exists(path) or open(path).read() != text:
        with open(path, "w") as f:
            f.write(text)


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    tab = os.path.join(here, "rys_tables.h")
    if "--tables" in sys.argv or not os.path.exists(tab):
        write_tables(tab)
    write_classes(os.path.join(here, "rys_generated.h"))


if __name__ == "__main__":
    main()
