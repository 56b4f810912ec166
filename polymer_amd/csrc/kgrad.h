// kgrad.h -- two-electron part of the nuclear gradient (the quantity
// FOURINDEX_ERI_GRADIENT accumulates, src/pfourindexgradient.f90:1334-1420):
//
//   dE2/dR = sum over shell quartets (ab|cd) of sum_{mu nu lam sig}
//            Gamma(mu nu lam sig) d(mu nu|lam sig)/dR,
//   Gamma  = f [4 D_mu,nu D_lam,sig - x (D_mu,lam D_nu,sig + D_mu,sig D_nu,lam)]
//
// over the symmetry-unique quartets (a >= b, c >= d, pair ab >= pair cd;
// f = 1/2 for each of a == b, c == d, ab == cd), with D the density in this
// library's contracted Cartesian basis and x the exchange scale.  That is
// E2 = 1/2 sum D D (mu nu|lam sig) - x/4 sum D D (mu lam|nu sig) over all
// function quadruples, differentiated at fixed D.  The derivative with
// respect to centre A of a primitive Cartesian Gaussian is
// 2 alpha [a + 1] - a_x [a - 1] (componentwise), taken on A, B and C; D's
// follows from translational invariance, -(dA + dB + dC).
//
// Integrals by Rys quadrature with the library's node tables (rys.h): for
// each primitive quartet and each of the n = (L + 1) / 2 + 1 roots, the
// one-dimensional factors G(i, k) of the two-electron recursion, then the
// horizontal transfers to I(i, j, k, l) for i + j <= La + Lb + 1 and
// k + l <= Lc + Ld + 1 -- one level above the quartet, for the derivatives.
//
// One thread per unique contracted quartet (grid-stride) over a pair list
// sorted by class and primitive count; quartets with Q_ab Q_cd, or Q_ab Q_cd
// times the largest |Gamma| of the quartet, below the threshold skipped, and
// primitive pairs below the J/K path's primitive threshold; per-atom sums in
// LDS, one global atomic per atom coordinate per block.  The same code runs on the
// host (polyeri_grad_host) for the CPU tests.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "eri_kernels.h"
#include "rys.h"

namespace {

#define GRAD_LMAX 2                      // highest shell L of the library's basis (d)
#define GMAX (2 * GRAD_LMAX + 2)         // G(i, k): i, k <= 2 LMAX + 1
#define IMAX (GRAD_LMAX + 2)             // I(i, j, k, l): i, j, k <= LMAX + 1
#define LMAXD (GRAD_LMAX + 1)            //                l <= LMAX

struct RysTabs {
    const double* tab[RYS_NMAX + 1];
    const double* herm[RYS_NMAX + 1];
};

__device__ const double g_gtab1[] = RYS_TABLE_1;
__device__ const double g_gtab2[] = RYS_TABLE_2;
__device__ const double g_gtab3[] = RYS_TABLE_3;
__device__ const double g_gtab4[] = RYS_TABLE_4;
__device__ const double g_gtab5[] = RYS_TABLE_5;
__device__ const double g_gherm1[] = RYS_HERM_1;
__device__ const double g_gherm2[] = RYS_HERM_2;
__device__ const double g_gherm3[] = RYS_HERM_3;
__device__ const double g_gherm4[] = RYS_HERM_4;
__device__ const double g_gherm5[] = RYS_HERM_5;
static const double h_gtab1[] = RYS_TABLE_1, h_gtab2[] = RYS_TABLE_2, h_gtab3[] = RYS_TABLE_3,
                    h_gtab4[] = RYS_TABLE_4, h_gtab5[] = RYS_TABLE_5;
static const double h_gherm1[] = RYS_HERM_1, h_gherm2[] = RYS_HERM_2, h_gherm3[] = RYS_HERM_3,
                    h_gherm4[] = RYS_HERM_4, h_gherm5[] = RYS_HERM_5;

POLY_HD void rys_nodes_n(int n, double T, const RysTabs& rt, double* u, double* w) {
    switch (n) {
    case 1: rys_node<1>(T, 0, rt.tab[1], rt.herm[1], u[0], w[0]); break;
    case 2: for (int r = 0; r < 2; ++r) rys_node<2>(T, r, rt.tab[2], rt.herm[2], u[r], w[r]); break;
    case 3: for (int r = 0; r < 3; ++r) rys_node<3>(T, r, rt.tab[3], rt.herm[3], u[r], w[r]); break;
    case 4: for (int r = 0; r < 4; ++r) rys_node<4>(T, r, rt.tab[4], rt.herm[4], u[r], w[r]); break;
    default: for (int r = 0; r < 5; ++r) rys_node<5>(T, r, rt.tab[5], rt.herm[5], u[r], w[r]); break;
    }
}

POLY_HD void comp_of(int l, int q, int& lx, int& ly, int& lz) {
    // xx, xy, xz, yy, yz, zz order (x exponent descending, then y)
    int n = 0;
    for (int i = l; i >= 0; --i)
        for (int k = 0; k <= l - i; ++k) {
            if (n == q) { lx = i; ly = l - i - k; lz = k; return; }
            ++n;
        }
    lx = ly = lz = 0;
}

// binomial coefficients C(n, k), n <= 3 (the transfers go one level above d)
POLY_HD double binom(int n, int k) {
    return n == 3 ? (k == 0 || k == 3 ? 1.0 : 3.0) : n == 2 ? (k == 1 ? 2.0 : 1.0) : 1.0;
}

// One-dimensional factors of one axis for one root: I[i][j][k][l] for
// i <= ia, j <= jb, k <= kc, l <= ld with i + j <= na and k + l <= nc (the
// derivatives need one level above the quartet on one index at a time).
POLY_HD void axis_table(double PA, double QC, double PQ, double AB, double CD, double u, double ze, double et,
                        double rho, int na, int nc, int ia, int jb, int kc, int ld,
                        double (&I)[IMAX][IMAX][IMAX][LMAXD]) {
    double G[GMAX][GMAX];
    const double c00 = PA - rho / ze * PQ * u, d00 = QC + rho / et * PQ * u;
    const double b10 = 0.5 / ze * (1.0 - rho * u / ze), b01 = 0.5 / et * (1.0 - rho * u / et);
    const double b00 = 0.5 * u / (ze + et);
    G[0][0] = 1.0;
    for (int i = 0; i < na; ++i) G[i + 1][0] = c00 * G[i][0] + (i ? i * b10 * G[i - 1][0] : 0.0);
    for (int k = 0; k < nc; ++k)
        for (int i = 0; i <= na; ++i)
            G[i][k + 1] = d00 * G[i][k] + (k ? k * b01 * G[i][k - 1] : 0.0) + (i ? i * b00 * G[i - 1][k] : 0.0);
    // ket transfer: H[i][k][l] = sum_s C(l, s) CD^(l - s) G[i][k + s]
    double H[GMAX][IMAX][LMAXD];
    for (int i = 0; i <= na; ++i)
        for (int k = 0; k <= kc; ++k)
            for (int l = 0; l <= ld && k + l <= nc; ++l) {
                double s = 0.0, p = 1.0;
                for (int t = l; t >= 0; --t) {      // CD^(l - t) G[i][k + t]
                    s += binom(l, t) * p * G[i][k + t];
                    p *= CD;
                }
                H[i][k][l] = s;
            }
    for (int i = 0; i <= ia; ++i)
        for (int j = 0; j <= jb && i + j <= na; ++j)
            for (int k = 0; k <= kc; ++k)
                for (int l = 0; l <= ld && k + l <= nc; ++l) {
                    double s = 0.0, p = 1.0;
                    for (int t = j; t >= 0; --t) {
                        s += binom(j, t) * p * H[i + t][k][l];
                        p *= AB;
                    }
                    I[i][j][k][l] = s;
                }
}

// Shell data of one quartet position.
struct Sh {
    int L, atom, f0, p0, np;
    double x, y, z;
};

POLY_HD Sh shell_of(const GradArgs& g, int s) {
    Sh r;
    r.L = g.L[s]; r.atom = g.atom[s]; r.f0 = g.f0[s]; r.p0 = g.p0[s]; r.np = g.np[s];
    r.x = g.xyz[3 * s]; r.y = g.xyz[3 * s + 1]; r.z = g.xyz[3 * s + 2];
    return r;
}

// max over components of |(mu nu|mu nu)| for the Schwarz bound of pair (a, b)
POLY_HD double pair_bound(const GradArgs& g, const RysTabs& rt, int a, int b) {
    const Sh A = shell_of(g, a), B = shell_of(g, b);
    const int na = A.L + B.L, n = (2 * na) / 2 + 1;
    const int ca = ncart(A.L), cb = ncart(B.L);
    double acc[36];
    for (int q = 0; q < ca * cb; ++q) acc[q] = 0.0;
    const double AB[3] = {A.x - B.x, A.y - B.y, A.z - B.z};
    const double ab2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
    double I3[3][IMAX][IMAX][IMAX][LMAXD];
    for (int i = 0; i < A.np; ++i)
        for (int j = 0; j < B.np; ++j) {
            const double al = g.alpha[A.p0 + i], be = g.alpha[B.p0 + j];
            const double ze = al + be, kab = std::exp(-al * be / ze * ab2);
            const double cab = g.coef[A.p0 + i] * g.coef[B.p0 + j] * kab;
            if (std::fabs(cab) < g.ptol) continue;
            const double P[3] = {(al * A.x + be * B.x) / ze, (al * A.y + be * B.y) / ze, (al * A.z + be * B.z) / ze};
            for (int k = 0; k < A.np; ++k)
                for (int l = 0; l < B.np; ++l) {
                    const double ga = g.alpha[A.p0 + k], de = g.alpha[B.p0 + l];
                    const double et = ga + de, kcd = std::exp(-ga * de / et * ab2);
                    const double ccd = g.coef[A.p0 + k] * g.coef[B.p0 + l] * kcd;
                    if (std::fabs(ccd) < g.ptol) continue;
                    const double Q[3] = {(ga * A.x + de * B.x) / et, (ga * A.y + de * B.y) / et,
                                         (ga * A.z + de * B.z) / et};
                    const double PQ[3] = {P[0] - Q[0], P[1] - Q[1], P[2] - Q[2]};
                    const double rho = ze * et / (ze + et);
                    const double T = rho * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
                    const double pref = 2.0 * POLY_PI * POLY_PI * POLY_SQRT_PI / (ze * et * std::sqrt(ze + et)) * cab * ccd;
                    double u[RYS_NMAX], w[RYS_NMAX];
                    rys_nodes_n(n, T, rt, u, w);
                    for (int r = 0; r < n; ++r) {
                        const double Pc[3] = {P[0], P[1], P[2]};
                        const double Ac[3] = {A.x, A.y, A.z};
                        for (int ax = 0; ax < 3; ++ax)
                            axis_table(Pc[ax] - Ac[ax], Q[ax] - Ac[ax], PQ[ax], AB[ax], AB[ax], u[r], ze, et, rho, na,
                                       na, A.L, B.L, A.L, B.L, I3[ax]);
                        for (int p = 0; p < ca; ++p)
                            for (int q = 0; q < cb; ++q) {
                                int x1, y1, z1, x2, y2, z2;
                                comp_of(A.L, p, x1, y1, z1);
                                comp_of(B.L, q, x2, y2, z2);
                                acc[p * cb + q] += w[r] * pref * I3[0][x1][x2][x1][x2] * I3[1][y1][y2][y1][y2] *
                                                   I3[2][z1][z2][z1][z2];
                            }
                    }
                }
        }
    double m = 0.0;
    for (int q = 0; q < ca * cb; ++q) m = std::fmax(m, std::fabs(acc[q]));
    return std::sqrt(m);
}

// One-dimensional factors of one axis for one root with compile-time
// ranges: I[i][j][k][l], i + j <= NA, k + l <= NC, i <= LA + 1, j <= LB + 1,
// k <= LC + 1, l <= LD.
template <int LA, int LB, int LC, int LD>
POLY_HD void axis_table_t(double PA, double QC, double PQ, double AB, double CD, double u, double ze, double et,
                          double rho, double (&I)[LA + 2][LB + 2][LC + 2][LD + 1]) {
    constexpr int NA = LA + LB + 1, NC = LC + LD + 1;
    double G[NA + 1][NC + 1];
    const double c00 = PA - rho / ze * PQ * u, d00 = QC + rho / et * PQ * u;
    const double b10 = 0.5 / ze * (1.0 - rho * u / ze), b01 = 0.5 / et * (1.0 - rho * u / et);
    const double b00 = 0.5 * u / (ze + et);
    G[0][0] = 1.0;
#pragma unroll
    for (int i = 0; i < NA; ++i) G[i + 1][0] = c00 * G[i][0] + (i ? i * b10 * G[i - 1][0] : 0.0);
#pragma unroll
    for (int k = 0; k < NC; ++k)
#pragma unroll
        for (int i = 0; i <= NA; ++i)
            G[i][k + 1] = d00 * G[i][k] + (k ? k * b01 * G[i][k - 1] : 0.0) + (i ? i * b00 * G[i - 1][k] : 0.0);
    double H[NA + 1][LC + 2][LD + 1];
#pragma unroll
    for (int i = 0; i <= NA; ++i)
#pragma unroll
        for (int k = 0; k <= LC + 1; ++k)
#pragma unroll
            for (int l = 0; l <= LD; ++l) {
                if (k + l > NC) continue;
                double sum = 0.0, pw = 1.0;
#pragma unroll
                for (int t = l; t >= 0; --t) {
                    sum += binom(l, t) * pw * G[i][k + t];
                    pw *= CD;
                }
                H[i][k][l] = sum;
            }
#pragma unroll
    for (int i = 0; i <= LA + 1; ++i)
#pragma unroll
        for (int j = 0; j <= LB + 1; ++j) {
            if (i + j > NA) continue;
#pragma unroll
            for (int k = 0; k <= LC + 1; ++k)
#pragma unroll
                for (int l = 0; l <= LD; ++l) {
                    if (k + l > NC) continue;
                    double sum = 0.0, pw = 1.0;
#pragma unroll
                    for (int t = j; t >= 0; --t) {
                        sum += binom(j, t) * pw * H[i + t][k][l];
                        pw *= AB;
                    }
                    I[i][j][k][l] = sum;
                }
        }
}

// Gradient contributions of one unique contracted quartet, shells a, b, c, d
// of L = LA, LB, LC, LD, to centres A, B, C (D's by invariance).  The
// quartet's Gamma block is formed once, then the primitive quartets and
// roots run over it.
template <int LA, int LB, int LC, int LD>
POLY_HD void grad_quartet_t(const GradArgs& g, const RysTabs& rt, int a, int b, int c, int d, double fdeg,
                            double (&gA)[3], double (&gB)[3], double (&gC)[3]) {
    constexpr int CA = ncart(LA), CB = ncart(LB), CC = ncart(LC), CD_ = ncart(LD);
    constexpr int NR = (LA + LB + LC + LD + 1) / 2 + 1;
    const Sh A = shell_of(g, a), B = shell_of(g, b), C = shell_of(g, c), D = shell_of(g, d);
    const double AB[3] = {A.x - B.x, A.y - B.y, A.z - B.z}, CDv[3] = {C.x - D.x, C.y - D.y, C.z - D.z};
    const double ab2 = AB[0] * AB[0] + AB[1] * AB[1] + AB[2] * AB[2];
    const double cd2 = CDv[0] * CDv[0] + CDv[1] * CDv[1] + CDv[2] * CDv[2];
    const double Ac[3] = {A.x, A.y, A.z}, Cc[3] = {C.x, C.y, C.z};
    const double* Dm = g.D;
    const size_t nbf = (size_t)g.nbf;
    double gam[CA][CB][CC][CD_];
#pragma unroll
    for (int p = 0; p < CA; ++p)
#pragma unroll
        for (int q = 0; q < CB; ++q) {
            const double dab = Dm[(A.f0 + p) * nbf + B.f0 + q];
#pragma unroll
            for (int s2 = 0; s2 < CC; ++s2) {
                const double dac = Dm[(A.f0 + p) * nbf + C.f0 + s2], dbc = Dm[(B.f0 + q) * nbf + C.f0 + s2];
#pragma unroll
                for (int t = 0; t < CD_; ++t) {
                    const double dcd = Dm[(C.f0 + s2) * nbf + D.f0 + t], dbd = Dm[(B.f0 + q) * nbf + D.f0 + t];
                    const double dad = Dm[(A.f0 + p) * nbf + D.f0 + t];
                    gam[p][q][s2][t] = fdeg * (4.0 * dab * dcd - g.xfac * (dac * dbd + dad * dbc));
                }
            }
        }
    for (int e = 0; e < 3; ++e) gA[e] = gB[e] = gC[e] = 0.0;
    double I3[3][LA + 2][LB + 2][LC + 2][LD + 1];
    for (int i = 0; i < A.np; ++i)
        for (int j = 0; j < B.np; ++j) {
            const double al = g.alpha[A.p0 + i], be = g.alpha[B.p0 + j];
            const double ze = al + be;
            const double cab = g.coef[A.p0 + i] * g.coef[B.p0 + j] * std::exp(-al * be / ze * ab2);
            if (std::fabs(cab) < g.ptol) continue;          // the J/K pairs' primitive screen
            const double P[3] = {(al * A.x + be * B.x) / ze, (al * A.y + be * B.y) / ze, (al * A.z + be * B.z) / ze};
            for (int k = 0; k < C.np; ++k)
                for (int l = 0; l < D.np; ++l) {
                    const double ga = g.alpha[C.p0 + k], de = g.alpha[D.p0 + l];
                    const double et = ga + de;
                    const double ccd = g.coef[C.p0 + k] * g.coef[D.p0 + l] * std::exp(-ga * de / et * cd2);
                    if (std::fabs(ccd) < g.ptol) continue;
                    const double Q[3] = {(ga * C.x + de * D.x) / et, (ga * C.y + de * D.y) / et,
                                         (ga * C.z + de * D.z) / et};
                    const double PQ[3] = {P[0] - Q[0], P[1] - Q[1], P[2] - Q[2]};
                    const double rho = ze * et / (ze + et);
                    const double T = rho * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
                    const double pref = 2.0 * POLY_PI * POLY_PI * POLY_SQRT_PI / (ze * et * std::sqrt(ze + et)) * cab * ccd;
                    double u[NR], w[NR];
#pragma unroll
                    for (int r = 0; r < NR; ++r) rys_node<NR>(T, r, rt.tab[NR], rt.herm[NR], u[r], w[r]);
#pragma unroll
                    for (int r = 0; r < NR; ++r) {
#pragma unroll
                        for (int ax = 0; ax < 3; ++ax)
                            axis_table_t<LA, LB, LC, LD>(P[ax] - Ac[ax], Q[ax] - Cc[ax], PQ[ax], AB[ax], CDv[ax], u[r],
                                                         ze, et, rho, I3[ax]);
                        const double wr = w[r] * pref;
#pragma unroll
                        for (int p = 0; p < CA; ++p) {
                            int a3[3];
                            comp_of(LA, p, a3[0], a3[1], a3[2]);
#pragma unroll
                            for (int q = 0; q < CB; ++q) {
                                int b3[3];
                                comp_of(LB, q, b3[0], b3[1], b3[2]);
#pragma unroll
                                for (int s2 = 0; s2 < CC; ++s2) {
                                    int c3[3];
                                    comp_of(LC, s2, c3[0], c3[1], c3[2]);
#pragma unroll
                                    for (int t = 0; t < CD_; ++t) {
                                        int d3[3];
                                        comp_of(LD, t, d3[0], d3[1], d3[2]);
                                        double v[3], dA[3], dB[3], dC[3];
#pragma unroll
                                        for (int e = 0; e < 3; ++e) {
                                            const auto& Ie = I3[e];
                                            const int x1 = a3[e], x2 = b3[e], x3 = c3[e], x4 = d3[e];
                                            v[e] = Ie[x1][x2][x3][x4];
                                            dA[e] = 2.0 * al * Ie[x1 + 1][x2][x3][x4] - (x1 ? x1 * Ie[x1 - 1][x2][x3][x4] : 0.0);
                                            dB[e] = 2.0 * be * Ie[x1][x2 + 1][x3][x4] - (x2 ? x2 * Ie[x1][x2 - 1][x3][x4] : 0.0);
                                            dC[e] = 2.0 * ga * Ie[x1][x2][x3 + 1][x4] - (x3 ? x3 * Ie[x1][x2][x3 - 1][x4] : 0.0);
                                        }
                                        const double f = wr * gam[p][q][s2][t];
                                        gA[0] += f * dA[0] * v[1] * v[2];
                                        gA[1] += f * v[0] * dA[1] * v[2];
                                        gA[2] += f * v[0] * v[1] * dA[2];
                                        gB[0] += f * dB[0] * v[1] * v[2];
                                        gB[1] += f * v[0] * dB[1] * v[2];
                                        gB[2] += f * v[0] * v[1] * dB[2];
                                        gC[0] += f * dC[0] * v[1] * v[2];
                                        gC[1] += f * v[0] * dC[1] * v[2];
                                        gC[2] += f * v[0] * v[1] * dC[2];
                                    }
                                }
                            }
                        }
                    }
                }
        }
}

// index t of the lower triangle (row r, column c <= r): t = r (r + 1) / 2 + c
POLY_HD void tri_index(unsigned long long t, int& r, int& c) {
    long long rr = (long long)((std::sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
    while ((unsigned long long)(rr * (rr + 1) / 2) > t) --rr;
    while ((unsigned long long)((rr + 1) * (rr + 2) / 2) <= t) ++rr;
    r = (int)rr;
    c = (int)(t - (unsigned long long)(rr * (rr + 1) / 2));
}

// Contribution of the quartet of sorted pairs pij (class LA LB) and pkl
// (class LC LD) to the per-atom gradient through add(atom, axis, value).
template <int LA, int LB, int LC, int LD, typename Add>
POLY_HD void quartet_cls(const GradArgs& g, const RysTabs& rt, int pij, int pkl, Add add) {
    const double qq = g.Qp[pij] * g.Qp[pkl];
    if (qq < g.tol) return;
    const int a = g.pa[pij], b = g.pb[pij], c = g.pa[pkl], d = g.pb[pkl];
    // the density-weighted bound: Q_ab Q_cd max |Gamma| over the quartet
    const int ns = g.nsh;
    const double gmax = 4.0 * g.dsh[a * ns + b] * g.dsh[c * ns + d] +
                        std::fabs(g.xfac) * (g.dsh[a * ns + c] * g.dsh[b * ns + d] + g.dsh[a * ns + d] * g.dsh[b * ns + c]);
    if (qq * gmax < g.tol) return;
    const double fdeg = (a == b ? 0.5 : 1.0) * (c == d ? 0.5 : 1.0) * (pij == pkl ? 0.5 : 1.0);
    double gA[3], gB[3], gC[3];
    grad_quartet_t<LA, LB, LC, LD>(g, rt, a, b, c, d, fdeg, gA, gB, gC);
    const int at[4] = {g.atom[a], g.atom[b], g.atom[c], g.atom[d]};
    for (int e = 0; e < 3; ++e) {
        add(at[0], e, gA[e]);
        add(at[1], e, gB[e]);
        add(at[2], e, gC[e]);
        add(at[3], e, -(gA[e] + gB[e] + gC[e]));
    }
}

// quartet t of the class pair: (x0 + i, y0 + j).  With chunks: the chunk
// whose range holds t (binary search; chunks are never empty, so their
// offsets increase strictly), i its row and j its first column plus the
// distance from its start.  Without: j <= i when the classes are equal (lower
// triangle), else i = t / ny, j = t % ny.
POLY_HD void class_pair_index(const GradArgs& g, unsigned long long t, int same, int ny, int& i, int& j) {
    if (g.coff) {
        int lo = 0, hi = g.nchunk - 1;         // largest c with coff[c] <= t
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (g.coff[mid] <= t) lo = mid;
            else hi = mid - 1;
        }
        i = g.cij[2 * lo];
        j = g.cij[2 * lo + 1] + (int)(t - g.coff[lo]);
        return;
    }
    if (same) {
        tri_index(t, i, j);
    } else {
        i = (int)(t / (unsigned long long)ny);
        j = (int)(t % (unsigned long long)ny);
    }
}

__device__ RysTabs device_tabs() {
    RysTabs rt;
    rt.tab[0] = rt.herm[0] = nullptr;
    rt.tab[1] = g_gtab1; rt.tab[2] = g_gtab2; rt.tab[3] = g_gtab3; rt.tab[4] = g_gtab4; rt.tab[5] = g_gtab5;
    rt.herm[1] = g_gherm1; rt.herm[2] = g_gherm2; rt.herm[3] = g_gherm3; rt.herm[4] = g_gherm4; rt.herm[5] = g_gherm5;
    return rt;
}

RysTabs host_tabs() {
    RysTabs rt;
    rt.tab[0] = rt.herm[0] = nullptr;
    rt.tab[1] = h_gtab1; rt.tab[2] = h_gtab2; rt.tab[3] = h_gtab3; rt.tab[4] = h_gtab4; rt.tab[5] = h_gtab5;
    rt.herm[1] = h_gherm1; rt.herm[2] = h_gherm2; rt.herm[3] = h_gherm3; rt.herm[4] = h_gherm4; rt.herm[5] = h_gherm5;
    return rt;
}

// per-atom sums in LDS (3 natom doubles of dynamic shared memory), one
// global atomic per element per block
template <int LA, int LB, int LC, int LD>
__global__ void __launch_bounds__(256) k_eri_grad(GradArgs g, int x0, int y0, int ny, int same, unsigned long long nq,
                                                  double* __restrict__ grad) {
    extern __shared__ double sg[];
    for (int i = threadIdx.x; i < 3 * g.natom; i += blockDim.x) sg[i] = 0.0;
    __syncthreads();
    const RysTabs rt = device_tabs();
    for (unsigned long long t = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; t < nq;
         t += (unsigned long long)gridDim.x * blockDim.x) {
        int i, j;
        class_pair_index(g, t, same, ny, i, j);
        quartet_cls<LA, LB, LC, LD>(g, rt, x0 + i, y0 + j, [&](int at, int e, double v) {
            if (v != 0.0) atomicAdd(&sg[3 * at + e], v);
        });
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 3 * g.natom; i += blockDim.x)
        if (sg[i] != 0.0) atomicAdd(&grad[i], sg[i]);
}

// one class pair's launch (device) or loop (host)
struct GradLaunch {
    const GradArgs& gq;
    int x0, y0, ny, same;
    unsigned long long nq;
    unsigned grid;
    size_t lds;
    double* d;
    hipStream_t s;
    template <int A, int B, int C, int D>
    void operator()() const {
        hipLaunchKernelGGL((k_eri_grad<A, B, C, D>), dim3(grid), dim3(256), lds, s, gq, x0, y0, ny, same, nq, d);
    }
};

struct GradHostRun {
    const GradArgs& gq;
    const RysTabs& rt;
    int x0, y0, ny, same;
    unsigned long long nq;
    double* grad;
    template <int A, int B, int C, int D>
    void operator()() const {
        for (unsigned long long t = 0; t < nq; ++t) {
            int i, j;
            class_pair_index(gq, t, same, ny, i, j);
            quartet_cls<A, B, C, D>(gq, rt, x0 + i, y0 + j, [&](int at, int e, double v) { grad[3 * at + e] += v; });
        }
    }
};

// Runs class pair (X, Y), Y <= X, of the pair classes
// ((ss) (ps) (pp) (ds) (dp) (dd) = 0 .. 5) through f.
constexpr int kGradPL[6][2] = {{0, 0}, {1, 0}, {1, 1}, {2, 0}, {2, 1}, {2, 2}};
template <int X, typename F>
void grad_class_row(int Y, F&& f) {
    switch (Y) {
    case 0: f.template operator()<kGradPL[X][0], kGradPL[X][1], 0, 0>(); break;
    case 1: if constexpr (X >= 1) f.template operator()<kGradPL[X][0], kGradPL[X][1], 1, 0>(); break;
    case 2: if constexpr (X >= 2) f.template operator()<kGradPL[X][0], kGradPL[X][1], 1, 1>(); break;
    case 3: if constexpr (X >= 3) f.template operator()<kGradPL[X][0], kGradPL[X][1], 2, 0>(); break;
    case 4: if constexpr (X >= 4) f.template operator()<kGradPL[X][0], kGradPL[X][1], 2, 1>(); break;
    case 5: if constexpr (X >= 5) f.template operator()<kGradPL[X][0], kGradPL[X][1], 2, 2>(); break;
    default: break;
    }
}

}  // namespace

// the kernels of bra pair class X, one translation unit each (kgrad/kgrad_xX.hip)
#define GRAD_ROW_DECL(X)                                                                                        \
    void polyeri_grad_row_launch_##X(const GradArgs& gq, int Y, int x0, int y0, int ny, int same,               \
                                     unsigned long long nq, unsigned grid, size_t lds, double* d, hipStream_t s); \
    void polyeri_grad_row_host_##X(const GradArgs& gq, int Y, int x0, int y0, int ny, int same,                 \
                                   unsigned long long nq, double* grad);
GRAD_ROW_DECL(0) GRAD_ROW_DECL(1) GRAD_ROW_DECL(2) GRAD_ROW_DECL(3) GRAD_ROW_DECL(4) GRAD_ROW_DECL(5)
#undef GRAD_ROW_DECL
