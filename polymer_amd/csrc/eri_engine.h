// eri_engine.h -- per-shell-quartet electron-repulsion integrals for gfx950.
//
// One contracted shell quartet (ab|cd) is evaluated by looping over its
// primitive pairs: Boys function (downward recursion from a Taylor table),
// Obara-Saika VRR to [e0|f0] accumulated over primitives, then HRR to the
// final Cartesian components (code from gen_eri.py).  The integrals are the
// ones src/pfourindexdirect.f90:528-690 evaluates, over contracted shells
// instead of primitive composite shells.
#pragma once

#include <math.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define POLY_HD __host__ __device__ __forceinline__
#else
#define POLY_HD inline
#endif

// src/pmodules.f90:7 -- polymer's value of pi (used by its normalisation and
// ERI prefactors; kept so results agree with the reference to the last bits).
#define POLY_PI 3.14159265358979

// Boys table: F_m(T) on T = i/20, i = 0..BOYS_NT-1, m = 0..BOYS_NM-1,
// row-major [i][m] so the seven values of one Taylor step are contiguous.
#define BOYS_NT 801
#define BOYS_NM 16
#define BOYS_TMAX 40.0
// The thread kernels read a second layout of the same values: for each
// L < BOYS_LA a table of 64-byte rows F_L(T_i) .. F_{L+6}(T_i) (and one pad),
// so the seven values of one lookup arrive in four 16-byte loads.  It follows
// the BOYS_NT x BOYS_NM table in the same allocation (BOYS_AOFF doubles in).
#define BOYS_LA 9
#define BOYS_AOFF (BOYS_NT * BOYS_NM)
// A third layout, small enough for LDS, for the kernels that copy it there:
// for each L < BOYS_LC, rows F_L(T_j) .. F_{L+9}(T_j) on the coarse grid
// T_j = j/5 (every fourth point of the main grid), j = 0..BOYS_CNT-1, read
// with a ten-term Taylor step (|h| <= 0.1, truncation below 1e-18).  It
// follows the aligned tables (BOYS_COFF doubles in); BOYS_CROWS doubles
// (16 080 bytes) per L.
#define BOYS_LC 7
#define BOYS_CNT 201
#define BOYS_CROWS (BOYS_CNT * 10)
#define BOYS_COFF (BOYS_AOFF + BOYS_LA * BOYS_NT * 8)

// Primitive pair of a shell pair (a,b): exponent sum, prefactor with both
// contraction coefficients, centre P and P - A.  64 bytes.
struct PrimPair {
    double zeta;
    double K;   // c_a c_b sqrt(2) pi^{5/4} exp(-ab/zeta |AB|^2) / zeta
    double Px, Py, Pz;
    double PAx, PAy, PAz;
};

// Generally contracted pair (classes (ls|), s shells that share exponents):
// the pair's primitive pairs carry K without contraction coefficients, and
// each of its m (<= 4) member shell pairs weighs primitive pair p with
// w[j] = c_a(p) c_b(p) of its own two shells.  Member j pairs shell
// ((code >> 2j) & 1) of the first group (first functions fa0 / fa1) with
// shell ((code >> 2j+1) & 1) of the second group (fb0 / fb1).  A member
// quartet is digested only when q_bra q_ket reaches the Schwarz threshold,
// the same test the plain lists apply.
struct GcMeta {
    int prim_off, nprim, m, code;
    int fa0, fa1, fb0, fb1;
    int sa0, sa1, sb0, sb1;   // shell indices of the same four shells (-1: none)
    double q[4];   // Schwarz bound of each member pair (0 where the pair has no primitive pairs)
};
struct GcPrim {
    PrimPair pp;   // K = sqrt(2) pi^{5/4} exp(-ab/zeta |AB|^2) / zeta
    double w[4];
};
#define GC_MAXM 4

POLY_HD constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }

// sqrt(POLY_PI) rounded to double
#define POLY_SQRT_PI 1.7724538509055152

// 1/sqrt(x) for x > 0.  On the device: the hardware estimate refined by two
// Newton-Raphson steps r <- r + r (1 - x r^2) / 2 (each step squares the
// relative error, so the estimate's ~2^-22 ends below the double ulp),
// one instruction sequence instead of a division and a square root.
POLY_HD double rsqrt_fp64(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rsq(x);
    double e = __builtin_fma(-x * r, r, 1.0);
    r = __builtin_fma(0.5 * r, e, r);
    e = __builtin_fma(-x * r, r, 1.0);
    return __builtin_fma(0.5 * r, e, r);
#else
    return 1.0 / sqrt(x);
#endif
}

// Primitive pairs of the ket: an array of PrimPair, or the bucket-interleaved
// field arrays (primitive q of the ket at p[q * nk], field f at f * fs), in
// which neighbouring threads' kets are neighbouring doubles.
struct AosKet {
    const PrimPair* p;
    POLY_HD PrimPair load(int q) const { return p[q]; }
};
// A ket translated by dx along x (periodic lattice sums): P moves, P - A not.
struct ShiftKet {
    const PrimPair* p;
    double dx;
    POLY_HD PrimPair load(int q) const {
        PrimPair r = p[q];
        r.Px += dx;
        return r;
    }
};
struct SoaKet {
    const double* p;
    int nk;
    long long fs;
    POLY_HD PrimPair load(int q) const {
        const double* x = p + (long long)q * nk;
        PrimPair r;
        r.zeta = x[0]; r.K = x[fs]; r.Px = x[2 * fs]; r.Py = x[3 * fs]; r.Pz = x[4 * fs];
        r.PAx = x[5 * fs]; r.PAy = x[6 * fs]; r.PAz = x[7 * fs];
        return r;
    }
};

// F_0..F_L (T) scaled by pref.  FP64: 7-term Taylor step for F_L around the
// nearest grid point, then the downward recursion
// F_{m-1} = (2T F_m + e^-T) / (2m-1); large T: asymptotic F_0 and the
// upward recursion without the (negligible) e^-T term.
// Two table values loaded as one 16-byte word.
struct alignas(16) BoysPair {
    double x, y;
};

// tab: the aligned per-L tables (see BOYS_LA).
template <int L>
POLY_HD void boys_fp64(double T, double pref, const double* __restrict__ tab,
                       double* __restrict__ F) {
    static_assert(L < BOYS_LA, "Boys table too short for this class");
    if (T < BOYS_TMAX) {
        const int ts = (int)(T * 20.0 + 0.5);
        const double h = 0.05 * ts - T;
        const BoysPair* g = reinterpret_cast<const BoysPair*>(tab + ((size_t)L * BOYS_NT + ts) * 8);
        const BoysPair g01 = g[0], g23 = g[1], g45 = g[2];
        double f = tab[((size_t)L * BOYS_NT + ts) * 8 + 6];
        f = g45.y + h * (1.0 / 6.0) * f;
        f = g45.x + h * (1.0 / 5.0) * f;
        f = g23.y + h * (1.0 / 4.0) * f;
        f = g23.x + h * (1.0 / 3.0) * f;
        f = g01.y + h * (1.0 / 2.0) * f;
        f = g01.x + h * f;
        F[L] = f;
        if (L > 0) {
            const double e = exp(-T);
            const double t2 = 2.0 * T;
#pragma unroll
            for (int m = L - 1; m >= 0; --m) F[m] = (t2 * F[m + 1] + e) * (1.0 / (2 * m + 1));
        }
    } else {
        const double r = rsqrt_fp64(T), it = r * r;
        F[0] = 0.5 * POLY_SQRT_PI * r;
#pragma unroll
        for (int m = 0; m < L; ++m) F[m + 1] = F[m] * (m + 0.5) * it;
    }
#pragma unroll
    for (int m = 0; m <= L; ++m) F[m] *= pref;
}

// The same from the coarse rows of one L (row0: BOYS_CNT rows of ten
// values, in LDS), same downward recursion and asymptotic form.
template <int L>
POLY_HD void boys_coarse(double T, double pref, const double* __restrict__ row0, double* __restrict__ F) {
    static_assert(L < BOYS_LC, "coarse Boys table too short for this class");
    if (T < BOYS_TMAX) {
        const int ts = (int)(T * 5.0 + 0.5);
        const double h = 0.2 * ts - T;
        const BoysPair* g = reinterpret_cast<const BoysPair*>(row0 + ts * 10);
        const BoysPair g01 = g[0], g23 = g[1], g45 = g[2], g67 = g[3], g89 = g[4];
        double f = g89.y;
        f = g89.x + h * (1.0 / 9.0) * f;
        f = g67.y + h * (1.0 / 8.0) * f;
        f = g67.x + h * (1.0 / 7.0) * f;
        f = g45.y + h * (1.0 / 6.0) * f;
        f = g45.x + h * (1.0 / 5.0) * f;
        f = g23.y + h * (1.0 / 4.0) * f;
        f = g23.x + h * (1.0 / 3.0) * f;
        f = g01.y + h * (1.0 / 2.0) * f;
        f = g01.x + h * f;
        F[L] = f;
        if (L > 0) {
            const double e = exp(-T);
            const double t2 = 2.0 * T;
#pragma unroll
            for (int m = L - 1; m >= 0; --m) F[m] = (t2 * F[m + 1] + e) * (1.0 / (2 * m + 1));
        }
    } else {
        const double r = rsqrt_fp64(T), it = r * r;
        F[0] = 0.5 * POLY_SQRT_PI * r;
#pragma unroll
        for (int m = 0; m < L; ++m) F[m + 1] = F[m] * (m + 0.5) * it;
    }
#pragma unroll
    for (int m = 0; m <= L; ++m) F[m] *= pref;
}

#include "eri_generated.h"

// Dispatch to the generated per-class recursion.
template <int LA, int LB, int LC, int LD, typename real>
struct ClassCode;

#define POLY_CLASS(a, b, c, d)                                                          \
    template <typename real>                                                            \
    struct ClassCode<a, b, c, d, real> {                                                \
        static POLY_HD void vrr(real PAx, real PAy, real PAz, real WPx, real WPy,       \
                                real WPz, real QCx, real QCy, real QCz, real WQx,       \
                                real WQy, real WQz, real oo2z, real oo2e, real oo2ze,   \
                                real roz, real roe, const real* F, real* acc) {         \
            vrr_##a##b##c##d<real>(PAx, PAy, PAz, WPx, WPy, WPz, QCx, QCy, QCz, WQx, WQy, \
                                   WQz, oo2z, oo2e, oo2ze, roz, roe, F, acc);           \
        }                                                                               \
        template <typename Cb>                                                          \
        static POLY_HD void hrr(real ABx, real ABy, real ABz, real CDx, real CDy,       \
                                real CDz, const real* acc, Cb& cb) {                    \
            hrr_##a##b##c##d<real, Cb>(ABx, ABy, ABz, CDx, CDy, CDz, acc, cb);           \
        }                                                                               \
    };
POLY_CLASS(0, 0, 0, 0)
POLY_CLASS(1, 0, 0, 0)
POLY_CLASS(1, 0, 1, 0)
POLY_CLASS(1, 1, 0, 0)
POLY_CLASS(1, 1, 1, 0)
POLY_CLASS(1, 1, 1, 1)
POLY_CLASS(2, 0, 0, 0)
POLY_CLASS(2, 0, 1, 0)
POLY_CLASS(2, 0, 1, 1)
POLY_CLASS(2, 0, 2, 0)
POLY_CLASS(2, 1, 0, 0)
POLY_CLASS(2, 1, 1, 0)
POLY_CLASS(2, 1, 1, 1)
POLY_CLASS(2, 1, 2, 0)
POLY_CLASS(2, 1, 2, 1)
POLY_CLASS(2, 2, 0, 0)
POLY_CLASS(2, 2, 1, 0)
POLY_CLASS(2, 2, 1, 1)
POLY_CLASS(2, 2, 2, 0)
POLY_CLASS(2, 2, 2, 1)
POLY_CLASS(2, 2, 2, 2)
#undef POLY_CLASS

// One primitive quartet: the VRR's [e0|f0] values, prefactors K_ab K_cd
// included, added to acc.  COARSE: boys_tab is the coarse rows of L.
template <typename real, int LA, int LB, int LC, int LD, bool COARSE = false>
POLY_HD void prim_vrr(const PrimPair& bp, const PrimPair& kq, const double* __restrict__ boys_tab, real* acc) {
    constexpr int L = LA + LB + LC + LD;
    const double ze = bp.zeta + kq.zeta;
    const double rs = rsqrt_fp64(ze);
    const double ize = rs * rs;
    const double rho = bp.zeta * kq.zeta * ize;
    const double PQx = bp.Px - kq.Px, PQy = bp.Py - kq.Py, PQz = bp.Pz - kq.Pz;
    const double T = rho * (PQx * PQx + PQy * PQy + PQz * PQz);
    const double pref = bp.K * kq.K * rs;
    double F[L + 1];
    if constexpr (COARSE) boys_coarse<L>(T, pref, boys_tab, F);
    else boys_fp64<L>(T, pref, boys_tab, F);
    real Fr[L + 1];
#pragma unroll
    for (int m = 0; m <= L; ++m) Fr[m] = (real)F[m];
    const double roz = kq.zeta * ize;   // rho / zeta
    const double roe = bp.zeta * ize;   // rho / eta
    ClassCode<LA, LB, LC, LD, real>::vrr(
        (real)bp.PAx, (real)bp.PAy, (real)bp.PAz, (real)(-roz * PQx), (real)(-roz * PQy), (real)(-roz * PQz),
        (real)kq.PAx, (real)kq.PAy, (real)kq.PAz, (real)(roe * PQx), (real)(roe * PQy), (real)(roe * PQz),
        (real)(0.5 / bp.zeta), (real)(0.5 / kq.zeta), (real)(0.5 * ize), (real)roz, (real)roe, Fr, acc);
}

// Contracted (ab|cd) for one shell quartet.  bra/ket point at the
// primitive pairs of the two shell pairs; AB = A - B, CD = C - D.  Every
// final integral is passed to cb(ia, ib, ic, id, value) with the component
// indices of gen_eri.py (compile-time constants after inlining).  The pair
// quantities, T and the Boys function are FP64 in both precisions; real is
// the type of the VRR / HRR recursion and of the integrals (double, or float
// for the quartets the mixed-precision mode evaluates in FP32).
template <typename real, int LA, int LB, int LC, int LD, typename Ket, typename Cb>
POLY_HD void quartet(const PrimPair* __restrict__ bra, int nbra, const Ket& ket, int nket, double ABx,
                     double ABy, double ABz, double CDx, double CDy, double CDz,
                     const double* __restrict__ boys_tab, Cb& cb) {
    // components of every l in [LA, LA+LB] (and [LC, LC+LD])
    constexpr int NE = ((LA + LB + 1) * (LA + LB + 2) * (LA + LB + 3) - LA * (LA + 1) * (LA + 2)) / 6;
    constexpr int NF = ((LC + LD + 1) * (LC + LD + 2) * (LC + LD + 3) - LC * (LC + 1) * (LC + 2)) / 6;
    real acc[NE * NF];
#pragma unroll
    for (int i = 0; i < NE * NF; ++i) acc[i] = real(0);
    for (int p = 0; p < nbra; ++p) {
        const PrimPair bp = bra[p];
        for (int q = 0; q < nket; ++q) prim_vrr<real, LA, LB, LC, LD>(bp, ket.load(q), boys_tab, acc);
    }
    ClassCode<LA, LB, LC, LD, real>::hrr((real)ABx, (real)ABy, (real)ABz, (real)CDx, (real)CDy, (real)CDz,
                                         acc, cb);
}

template <int LA, int LB, int LC, int LD, typename Cb>
POLY_HD void quartet_fp64(const PrimPair* __restrict__ bra, int nbra, const PrimPair* __restrict__ ket,
                          int nket, double ABx, double ABy, double ABz, double CDx, double CDy,
                          double CDz, const double* __restrict__ boys_tab, Cb& cb) {
    quartet<double, LA, LB, LC, LD>(bra, nbra, AosKet{ket}, nket, ABx, ABy, ABz, CDx, CDy, CDz, boys_tab, cb);
}

// Store variant: out[((ia*NB + ib)*NC + ic)*ND + id].
template <int NB, int NC, int ND, typename real = double>
struct StoreCb {
    real* out;
    POLY_HD void operator()(int a, int b, int c, int d, real v) {
        out[((a * NB + b) * NC + c) * ND + d] = v;
    }
};
