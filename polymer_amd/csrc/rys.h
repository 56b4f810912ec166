// rys.h -- Rys quadrature nodes and weights (host and device).
//
// For a primitive quartet with Boys argument T, the n-point rule with nodes
// u_r in (0, 1) and weights w_r satisfies sum_r w_r u_r^m = F_m(T) for
// m = 0 .. 2n - 1, so an integral that the Obara-Saika recursion writes as a
// combination of F_0 .. F_L (L < 2n) is the weighted sum over the nodes of
// products of three one-dimensional factors (rys_generated.h).  The nodes and
// weights are Chebyshev series on unit intervals of T, and the scaled
// half-range Gauss-Hermite rule above RYS_TMAX(n) (gen_rys.py).
#pragma once

#include "eri_engine.h"
#include "rys_tables.h"

// Chebyshev series of degree RYS_DEG at x in [-1, 1] (Clenshaw).
POLY_HD double rys_cheb(const double* __restrict__ c, double x) {
    const double x2 = 2.0 * x;
    double b1 = c[RYS_DEG], b2 = 0.0;
#pragma unroll
    for (int k = RYS_DEG - 1; k >= 1; --k) {
        const double t = __builtin_fma(x2, b1, c[k] - b2);
        b2 = b1;
        b1 = t;
    }
    return __builtin_fma(x, b1, c[0] - b2);
}

// Node r and its weight of the N-point rule at T.  tab: the RYS_TABLE_N
// coefficients; herm: RYS_HERM_N.
template <int N>
POLY_HD void rys_node(double T, int r, const double* __restrict__ tab, const double* __restrict__ herm, double& u,
                      double& w) {
    if (T < RYS_TMAX(N)) {
        const int i = (int)T;
        const double x = 2.0 * (T - (double)i) - 1.0;
        const double* c = tab + ((size_t)i * N + r) * (2 * (RYS_DEG + 1));
        u = rys_cheb(c, x);
        w = rys_cheb(c + RYS_DEG + 1, x);
    } else {
        const double rs = rsqrt_fp64(T);
        u = herm[r] * rs * rs;
        w = herm[N + r] * rs;
    }
}
