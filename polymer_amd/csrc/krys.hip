// krys.hip -- Rys-quadrature J/K kernels for the d-heavy shell-quartet classes.
//
// The quartet (ab|cd) of a primitive quartet is, per Rys node (u_r, w_r),
// a product of three one-dimensional factors I_x I_y I_z, each from a short
// two-term recursion in (i, k) followed by the horizontal transfer to the
// four shells' exponents; summed over the n = L/2 + 1 nodes it equals the
// Obara-Saika integral that src/pfourindexdirect.f90:528-690 builds from
// F_0 .. F_L (sum_r w_r u_r^m = F_m(T)).  The integrals of one contracted
// quartet are shared among the lanes of its "members": the components of
// some shells are fixed per lane (rys_generated.h), and the tuples of fixed
// components are grouped by axis-permutation orbits, so that every lane of
// a wavefront runs the same instructions in its own permuted frame.
//
// Work: the bra-row segments of the shell-pair lists (as k_eri_seg).  A
// block owns one segment at a time, so the bra pair is fixed: the exchange
// contributions to the rows of the two bra shells and J_ab are summed in
// LDS and added to the XCD's J/K copy once per segment; J_cd is added per
// quartet.  The digestion (src/pfourindexdirect.f90:691-721) uses the same
// conventions as RegDigest: J = 2 (Jacc + Jacc^T), K = Kacc + Kacc^T after
// the build, each unique quartet weighted by its permutational degeneracy.
#include "eri_kernels.h"
#include "rys.h"
#include "rys_generated.h"

__device__ const double g_rys_tab3[] = RYS_TABLE_3;
__device__ const double g_rys_tab4[] = RYS_TABLE_4;
__device__ const double g_rys_tab5[] = RYS_TABLE_5;
__device__ const double g_rys_herm3[] = RYS_HERM_3;
__device__ const double g_rys_herm4[] = RYS_HERM_4;
__device__ const double g_rys_herm5[] = RYS_HERM_5;

template <int N>
__device__ __forceinline__ const double* rys_tab() {
    return N == 3 ? g_rys_tab3 : N == 4 ? g_rys_tab4 : g_rys_tab5;
}
template <int N>
__device__ __forceinline__ const double* rys_herm() {
    return N == 3 ? g_rys_herm3 : N == 4 ? g_rys_herm4 : g_rys_herm5;
}

// Global component index of local component k of a shell of angular momentum
// L, in the frame whose local axis l is global axis p[l] (component order
// xx, xy, xz, yy, yz, zz / x, y, z).
template <int L>
POLY_HD int glob_comp(int k, const int* p) {
    if (L == 0) return 0;
    if (L == 1) return p[k];
    const int i = k < 3 ? 0 : k < 5 ? 1 : 2;               // local axes of the d component
    const int j = k < 3 ? k : k < 5 ? k - 2 : 2;
    int a = p[i], b = p[j];
    if (a > b) { const int t = a; a = b; b = t; }
    return a * (5 - a) / 2 + b;
}

POLY_HD double sel3(const double* v, int i) { return i == 0 ? v[0] : i == 1 ? v[1] : v[2]; }

// One primitive quartet's contribution to a lane's integrals: the bra
// primitive pair bq (PA in the lane's frame: PA0..PA2), the ket primitive
// pair kq, the lane's axis permutation p, A - B and C - D in that frame;
// tab / herm: the N-point node tables of the class (R::NROOT).
template <typename R, typename real>
POLY_HD void rys_prim(const PrimPair& bq, double PA0, double PA1, double PA2, const PrimPair& kq, const int* p,
                      double ABl0, double ABl1, double ABl2, double CDl0, double CDl1, double CDl2,
                      const double* __restrict__ tab, const double* __restrict__ herm, real* __restrict__ accr) {
    const double ze = bq.zeta + kq.zeta;
    const double rs = rsqrt_fp64(ze);
    const double ize = rs * rs;
    const double PQg[3] = {bq.Px - kq.Px, bq.Py - kq.Py, bq.Pz - kq.Pz};
    const double T = bq.zeta * kq.zeta * ize * (PQg[0] * PQg[0] + PQg[1] * PQg[1] + PQg[2] * PQg[2]);
    const double pref = bq.K * kq.K * rs;
    const double QCg[3] = {kq.PAx, kq.PAy, kq.PAz};
    const double PQ0 = sel3(PQg, p[0]), PQ1 = sel3(PQg, p[1]), PQ2 = sel3(PQg, p[2]);
    const double QC0 = sel3(QCg, p[0]), QC1 = sel3(QCg, p[1]), QC2 = sel3(QCg, p[2]);
    const double roz = kq.zeta * ize, roe = bq.zeta * ize;   // rho/zeta, rho/eta
    const double h2z = 0.5 / bq.zeta, h2e = 0.5 / kq.zeta, h2ze = 0.5 * ize;
#pragma unroll 1
    for (int r = 0; r < R::NROOT; ++r) {
        double u, w;
        rys_node<R::NROOT>(T, r, tab, herm, u, w);
        const double uz = roz * u, ue = roe * u;
        R::template node<real>((real)(w * pref), (real)(u * h2ze), (real)(h2z * (1.0 - uz)),
                               (real)(h2e * (1.0 - ue)), (real)(PA0 - uz * PQ0), (real)(PA1 - uz * PQ1),
                               (real)(PA2 - uz * PQ2), (real)(QC0 + ue * PQ0), (real)(QC1 + ue * PQ1),
                               (real)(QC2 + ue * PQ2), (real)ABl0, (real)ABl1, (real)ABl2, (real)CDl0, (real)CDl1,
                               (real)CDl2, accr);
    }
}

// One variant of the class over the items of one segment that this wave
// owns: chunk c of the segment's items goes to wave (c + base) % nw.  MODE:
// MODE_JK (FP64) or MODE_JK32 (recursion and per-lane integrals in FP32;
// nodes, weights, pair data and the digestion in FP64).
template <int LA, int LB, int LC, int LD, int V, int MODE>
__device__ __forceinline__ void rys_variant(const EriArgs& g, const SegDesc& seg, int bp, int fa, int fb,
                                            const PrimPair* __restrict__ bra, int nbra, double* __restrict__ Kr,
                                            double* __restrict__ Jl, const double* __restrict__ Dl, double* J,
                                            int wave, int nw, int& base) {
    using R = RysVar<LA, LB, LC, LD, V>;
    using real = typename std::conditional<MODE == MODE_JK32, float, double>::type;
    constexpr int NA = ncart(LA), NB = ncart(LB);
    const int lane = threadIdx.x & 63;
    constexpr int IPW = 64 / R::NMEMB;
    const int chunks = (seg.n + IPW - 1) / IPW;
    const int n = g.nbf;
    const int memb = lane % R::NMEMB, slot = lane / R::NMEMB;
    int p[3];
    R::perm(memb, p[0], p[1], p[2]);
    int gA[R::NA], gB[R::NB], gC[R::NC], gD[R::ND];
#pragma unroll
    for (int i = 0; i < R::NA; ++i) gA[i] = glob_comp<LA>(R::CA[i], p);
#pragma unroll
    for (int i = 0; i < R::NB; ++i) gB[i] = glob_comp<LB>(R::CB[i], p);
#pragma unroll
    for (int i = 0; i < R::NC; ++i) gC[i] = glob_comp<LC>(R::CC[i], p);
#pragma unroll
    for (int i = 0; i < R::ND; ++i) gD[i] = glob_comp<LD>(R::CD[i], p);
    const double ABg[3] = {g.ab[4 * bp], g.ab[4 * bp + 1], g.ab[4 * bp + 2]};
    const double ABl0 = sel3(ABg, p[0]), ABl1 = sel3(ABg, p[1]), ABl2 = sel3(ABg, p[2]);
    double jab[R::NA * R::NB];
#pragma unroll
    for (int i = 0; i < R::NA * R::NB; ++i) jab[i] = 0.0;
    const int first = ((wave - base) % nw + nw) % nw;
    const bool screen = g.dtol > 0.0;
    for (int c = first; c < chunks; c += nw) {
        const int item = c * IPW + slot;
        bool live = slot < IPW && item < seg.n;
        int pos = 0, b = 0, kp = 0;
        if (live) {
            decode_in_row(g, seg.row, seg.t0 + item, pos, b);
            kp = g.ket_ids[pos];
            if (screen) {   // density screening: Q_ab Q_cd max|D| over the six blocks
                const int2 ab = g.psh[bp], cd = g.psh[kp];
                const size_t ns = g.nsh, sa = ab.x, sb = ab.y, sc = cd.x, sd = cd.y;
                double m = fmax(g.dsh[sc * ns + sd], g.dsh[sa * ns + sb]);
                m = fmax(m, fmax(g.dsh[sb * ns + sd], g.dsh[sb * ns + sc]));
                m = fmax(m, fmax(g.dsh[sa * ns + sd], g.dsh[sa * ns + sc]));
                live = g.Qs[bp] * g.Qs[kp] * m >= g.dtol;
            }
        }
        if (screen) {   // one count per item (member 0)
            const unsigned long long mask = __ballot(live && memb == 0);
            if (lane == 0 && mask) atomicAdd(g.kept, (unsigned)__popcll(mask));
        }
        if (!live) continue;
        const PairMeta mk = g.meta[kp];
        const int fc = mk.fa, fd = mk.fb;
        const SoaKet kv{g.ksoa + g.bucket_soa[b] + (pos - g.bucket_start[b]),
                        g.bucket_start[b + 1] - g.bucket_start[b], g.kfs};
        const double CDg[3] = {g.ab[4 * kp], g.ab[4 * kp + 1], g.ab[4 * kp + 2]};
        const double CDl0 = sel3(CDg, p[0]), CDl1 = sel3(CDg, p[1]), CDl2 = sel3(CDg, p[2]);
        real accr[R::NOUT];
#pragma unroll
        for (int i = 0; i < R::NOUT; ++i) accr[i] = real(0);
        for (int pi = 0; pi < nbra; ++pi) {
            const PrimPair bq = bra[pi];
            const double PAg[3] = {bq.PAx, bq.PAy, bq.PAz};
            const double PA0 = sel3(PAg, p[0]), PA1 = sel3(PAg, p[1]), PA2 = sel3(PAg, p[2]);
            for (int qi = 0; qi < mk.nprim; ++qi)
                rys_prim<R, real>(bq, PA0, PA1, PA2, kv.load(qi), p, ABl0, ABl1, ABl2, CDl0, CDl1, CDl2,
                                  rys_tab<R::NROOT>(), rys_herm<R::NROOT>(), accr);
        }
        // digestion of the lane's integrals v(a, b, c, d) (local components)
        const double s = (fa == fb ? 0.5 : 1.0) * (fc == fd ? 0.5 : 1.0) * (bp == kp ? 0.5 : 1.0);
        const double* D = g.D;
        double acc[R::NOUT];
#pragma unroll
        for (int i = 0; i < R::NOUT; ++i) acc[i] = s * (double)accr[i];
        // J_ab += v D_cd (registers, per lane); J_cd += v D_ab (one atomic per element)
#pragma unroll
        for (int ic = 0; ic < R::NC; ++ic)
#pragma unroll
            for (int id = 0; id < R::ND; ++id) {
                const double dcd = D[(size_t)(fc + gC[ic]) * n + fd + gD[id]];
                double jcd = 0.0;
#pragma unroll
                for (int ia = 0; ia < R::NA; ++ia)
#pragma unroll
                    for (int ib = 0; ib < R::NB; ++ib) {
                        const double x = acc[((ia * R::NB + ib) * R::NC + ic) * R::ND + id];
                        jab[ia * R::NB + ib] = __builtin_fma(x, dcd, jab[ia * R::NB + ib]);
                        jcd = __builtin_fma(x, Dl[gA[ia] * NB + gB[ib]], jcd);
                    }
                jk_add(&J[(size_t)(fc + gC[ic]) * n + fd + gD[id]], jcd, g.jk_copies);
            }
        // K_ac += v D_bd, K_ad += v D_bc (rows of shell a)
#pragma unroll
        for (int ia = 0; ia < R::NA; ++ia) {
#pragma unroll
            for (int ic = 0; ic < R::NC; ++ic) {
                double k = 0.0;
#pragma unroll
                for (int ib = 0; ib < R::NB; ++ib)
#pragma unroll
                    for (int id = 0; id < R::ND; ++id)
                        k = __builtin_fma(acc[((ia * R::NB + ib) * R::NC + ic) * R::ND + id],
                                          D[(size_t)(fb + gB[ib]) * n + fd + gD[id]], k);
                atomicAdd(&Kr[(size_t)gA[ia] * n + fc + gC[ic]], k);
            }
#pragma unroll
            for (int id = 0; id < R::ND; ++id) {
                double k = 0.0;
#pragma unroll
                for (int ib = 0; ib < R::NB; ++ib)
#pragma unroll
                    for (int ic = 0; ic < R::NC; ++ic)
                        k = __builtin_fma(acc[((ia * R::NB + ib) * R::NC + ic) * R::ND + id],
                                          D[(size_t)(fb + gB[ib]) * n + fc + gC[ic]], k);
                atomicAdd(&Kr[(size_t)gA[ia] * n + fd + gD[id]], k);
            }
        }
        // K_bc += v D_ad, K_bd += v D_ac (rows of shell b)
#pragma unroll
        for (int ib = 0; ib < R::NB; ++ib) {
#pragma unroll
            for (int ic = 0; ic < R::NC; ++ic) {
                double k = 0.0;
#pragma unroll
                for (int ia = 0; ia < R::NA; ++ia)
#pragma unroll
                    for (int id = 0; id < R::ND; ++id)
                        k = __builtin_fma(acc[((ia * R::NB + ib) * R::NC + ic) * R::ND + id],
                                          D[(size_t)(fa + gA[ia]) * n + fd + gD[id]], k);
                atomicAdd(&Kr[(size_t)(NA + gB[ib]) * n + fc + gC[ic]], k);
            }
#pragma unroll
            for (int id = 0; id < R::ND; ++id) {
                double k = 0.0;
#pragma unroll
                for (int ia = 0; ia < R::NA; ++ia)
#pragma unroll
                    for (int ic = 0; ic < R::NC; ++ic)
                        k = __builtin_fma(acc[((ia * R::NB + ib) * R::NC + ic) * R::ND + id],
                                          D[(size_t)(fa + gA[ia]) * n + fc + gC[ic]], k);
                atomicAdd(&Kr[(size_t)(NA + gB[ib]) * n + fd + gD[id]], k);
            }
        }
    }
    // this lane's J_ab share over the variant's items
    if (slot < IPW) {
#pragma unroll
        for (int ia = 0; ia < R::NA; ++ia)
#pragma unroll
            for (int ib = 0; ib < R::NB; ++ib)
                if (jab[ia * R::NB + ib] != 0.0) atomicAdd(&Jl[gA[ia] * NB + gB[ib]], jab[ia * R::NB + ib]);
    }
    base += chunks;
}

template <int LA, int LB, int LC, int LD, int V, int MODE>
__device__ __forceinline__ void rys_variants(const EriArgs& g, const SegDesc& seg, int bp, int fa, int fb,
                                             const PrimPair* bra, int nbra, double* Kr, double* Jl, const double* Dl,
                                             double* J, int wave, int nw, int& base) {
    if constexpr (V < RysClass<LA, LB, LC, LD>::NVAR) {
        rys_variant<LA, LB, LC, LD, V, MODE>(g, seg, bp, fa, fb, bra, nbra, Kr, Jl, Dl, J, wave, nw, base);
        rys_variants<LA, LB, LC, LD, V + 1, MODE>(g, seg, bp, fa, fb, bra, nbra, Kr, Jl, Dl, J, wave, nw, base);
    }
}

// LDS: exchange rows of the two bra shells ((NA + NB) nbf doubles), J_ab and
// the bra's density block D_ab (NA NB doubles each).
template <int LA, int LB, int LC, int LD, int MODE>
__global__ void __launch_bounds__(256) k_eri_rys(EriArgs g) {
    constexpr int NA = ncart(LA), NB = ncart(LB);
    extern __shared__ double rys_lds[];
    const int n = g.nbf;
    double* Kr = rys_lds;
    double* Jl = rys_lds + (size_t)(NA + NB) * n;
    double* Dl = Jl + NA * NB;
    const int total = (NA + NB) * n + NA * NB;
    const size_t off = (size_t)xcd_slot(g.jk_copies) * n * n;
    double* J = g.J + off;
    double* K = g.K + off;
    const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (int sg = blockIdx.x; sg < g.nseg; sg += gridDim.x) {
        const SegDesc seg = g.segs[sg];
        const int bp = g.bra_ids[seg.row];
        if (g.dtol > 0.0 && g.segq && g.Qs[bp] * g.segq[sg] * g.dmax[0] < g.dtol) continue;   // block-uniform
        const PairMeta mb = g.meta[bp];
        const int fa = mb.fa, fb = mb.fb;
        for (int i = threadIdx.x; i < total; i += blockDim.x) rys_lds[i] = 0.0;
        for (int i = threadIdx.x; i < NA * NB; i += blockDim.x) Dl[i] = g.D[(size_t)(fa + i / NB) * n + fb + i % NB];
        __syncthreads();
        int base = 0;
        rys_variants<LA, LB, LC, LD, 0, MODE>(g, seg, bp, fa, fb, g.prims + mb.prim_off, mb.nprim, Kr, Jl, Dl, J,
                                              wave, nw, base);
        __syncthreads();
        for (int i = threadIdx.x; i < (NA + NB) * n; i += blockDim.x) {
            const double x = Kr[i];
            if (x != 0.0) {
                const int r = i / n, col = i - r * n;
                const int f = r < NA ? fa + r : fb + (r - NA);
                jk_add(&K[(size_t)f * n + col], x, g.jk_copies);
            }
        }
        for (int i = threadIdx.x; i < NA * NB; i += blockDim.x)
            jk_add(&J[(size_t)(fa + i / NB) * n + fb + i % NB], Jl[i], g.jk_copies);
        __syncthreads();
    }
}

size_t polyeri_rys_lds_bytes(int la, int lb, int nbf) {
    const size_t na = ncart(la), nb = ncart(lb);
    return 8 * ((na + nb) * nbf + 2 * na * nb);
}

bool polyeri_rys_class(int la, int lb, int lc, int ld) {
    switch (la * 1000 + lb * 100 + lc * 10 + ld) {
    case 2111: case 2121: case 2211: case 2220: case 2221: case 2222: return true;
    }
    return false;
}

template <int MODE>
static void (*rys_kernel(int la, int lb, int lc, int ld))(EriArgs) {
    switch (la * 1000 + lb * 100 + lc * 10 + ld) {
    case 2111: return k_eri_rys<2, 1, 1, 1, MODE>;
    case 2121: return k_eri_rys<2, 1, 2, 1, MODE>;
    case 2211: return k_eri_rys<2, 2, 1, 1, MODE>;
    case 2220: return k_eri_rys<2, 2, 2, 0, MODE>;
    case 2221: return k_eri_rys<2, 2, 2, 1, MODE>;
    case 2222: return k_eri_rys<2, 2, 2, 2, MODE>;
    }
    return nullptr;
}

int polyeri_rys_launch(const EriArgs& a, int mode, int la, int lb, int lc, int ld, int block, int max_grid,
                       hipStream_t stream) {
    if (a.nseg <= 0) return 0;
    void (*k)(EriArgs) = mode == MODE_JK ? rys_kernel<MODE_JK>(la, lb, lc, ld)
                       : mode == MODE_JK32 ? rys_kernel<MODE_JK32>(la, lb, lc, ld) : nullptr;
    if (!k) return -1;
    if (block != 64 && block != 128 && block != 256) block = 256;
    unsigned grid = (unsigned)a.nseg;
    if (max_grid > 0 && grid > (unsigned)max_grid) grid = (unsigned)max_grid;
    hipLaunchKernelGGL(k, dim3(grid), dim3(block), polyeri_rys_lds_bytes(la, lb, a.nbf), stream, a);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

// Host evaluation of the same nodes and weights (tests).
static const double h_rys_tab1[] = RYS_TABLE_1, h_rys_tab2[] = RYS_TABLE_2, h_rys_tab3[] = RYS_TABLE_3,
                    h_rys_tab4[] = RYS_TABLE_4, h_rys_tab5[] = RYS_TABLE_5;
static const double h_rys_herm1[] = RYS_HERM_1, h_rys_herm2[] = RYS_HERM_2, h_rys_herm3[] = RYS_HERM_3,
                    h_rys_herm4[] = RYS_HERM_4, h_rys_herm5[] = RYS_HERM_5;

int polyeri_rys_nodes_host(int n, double T, double* u, double* w) {
    if (!(T >= 0.0) || !u || !w) return -1;
    for (int r = 0; r < n; ++r) {
        switch (n) {
        case 1: rys_node<1>(T, r, h_rys_tab1, h_rys_herm1, u[r], w[r]); break;
        case 2: rys_node<2>(T, r, h_rys_tab2, h_rys_herm2, u[r], w[r]); break;
        case 3: rys_node<3>(T, r, h_rys_tab3, h_rys_herm3, u[r], w[r]); break;
        case 4: rys_node<4>(T, r, h_rys_tab4, h_rys_herm4, u[r], w[r]); break;
        case 5: rys_node<5>(T, r, h_rys_tab5, h_rys_herm5, u[r], w[r]); break;
        default: return -1;
        }
    }
    return 0;
}

// Host evaluation of one primitive quartet's Cartesian block through the
// same per-lane code the kernels run (every variant, every member), for the
// CPU tests of the generated recursions.
static const double* host_tab(int n) {
    return n == 1 ? h_rys_tab1 : n == 2 ? h_rys_tab2 : n == 3 ? h_rys_tab3 : n == 4 ? h_rys_tab4 : h_rys_tab5;
}
static const double* host_herm(int n) {
    return n == 1 ? h_rys_herm1 : n == 2 ? h_rys_herm2 : n == 3 ? h_rys_herm3 : n == 4 ? h_rys_herm4 : h_rys_herm5;
}

template <int LA, int LB, int LC, int LD, int V>
static void rys_block_host_var(const PrimPair& bq, const PrimPair& kq, const double* AB, const double* CD, double* out,
                               int* hits) {
    if constexpr (V < RysClass<LA, LB, LC, LD>::NVAR) {
        using R = RysVar<LA, LB, LC, LD, V>;
        constexpr int NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
        for (int m = 0; m < R::NMEMB; ++m) {
            int p[3];
            R::perm(m, p[0], p[1], p[2]);
            const double PAg[3] = {bq.PAx, bq.PAy, bq.PAz};
            double acc[R::NOUT];
            for (int i = 0; i < R::NOUT; ++i) acc[i] = 0.0;
            rys_prim<R, double>(bq, sel3(PAg, p[0]), sel3(PAg, p[1]), sel3(PAg, p[2]), kq, p, sel3(AB, p[0]),
                                sel3(AB, p[1]), sel3(AB, p[2]), sel3(CD, p[0]), sel3(CD, p[1]), sel3(CD, p[2]),
                                host_tab(R::NROOT), host_herm(R::NROOT), acc);
            for (int ia = 0; ia < R::NA; ++ia)
                for (int ib = 0; ib < R::NB; ++ib)
                    for (int ic = 0; ic < R::NC; ++ic)
                        for (int id = 0; id < R::ND; ++id) {
                            const int o = ((glob_comp<LA>(R::CA[ia], p) * NB + glob_comp<LB>(R::CB[ib], p)) * NC +
                                           glob_comp<LC>(R::CC[ic], p)) * ND + glob_comp<LD>(R::CD[id], p);
                            out[o] = acc[((ia * R::NB + ib) * R::NC + ic) * R::ND + id];
                            ++hits[o];
                        }
        }
        rys_block_host_var<LA, LB, LC, LD, V + 1>(bq, kq, AB, CD, out, hits);
    }
}

// Primitive pair of exponents a (centre A) and b (centre B), unit
// coefficients: K = sqrt(2) pi^{5/4} exp(-ab/zeta |AB|^2) / zeta.
static PrimPair host_prim_pair(double a, const double* A, double b, const double* B) {
    PrimPair q;
    q.zeta = a + b;
    const double d2 = (A[0] - B[0]) * (A[0] - B[0]) + (A[1] - B[1]) * (A[1] - B[1]) + (A[2] - B[2]) * (A[2] - B[2]);
    q.K = std::sqrt(2.0) * std::pow(3.141592653589793, 1.25) * std::exp(-a * b / q.zeta * d2) / q.zeta;
    q.Px = (a * A[0] + b * B[0]) / q.zeta;
    q.Py = (a * A[1] + b * B[1]) / q.zeta;
    q.Pz = (a * A[2] + b * B[2]) / q.zeta;
    q.PAx = q.Px - A[0];
    q.PAy = q.Py - A[1];
    q.PAz = q.Pz - A[2];
    return q;
}

int polyeri_rys_block_host(int la, int lb, int lc, int ld, const double* xyz, const double* ex, double* out) {
    if (!polyeri_rys_class(la, lb, lc, ld) || !xyz || !ex || !out) return -1;
    const PrimPair bq = host_prim_pair(ex[0], xyz, ex[1], xyz + 3);
    const PrimPair kq = host_prim_pair(ex[2], xyz + 6, ex[3], xyz + 9);
    const double AB[3] = {xyz[0] - xyz[3], xyz[1] - xyz[4], xyz[2] - xyz[5]};
    const double CD[3] = {xyz[6] - xyz[9], xyz[7] - xyz[10], xyz[8] - xyz[11]};
    const int nout = ncart(la) * ncart(lb) * ncart(lc) * ncart(ld);
    std::vector<int> hits(nout, 0);
    switch (la * 1000 + lb * 100 + lc * 10 + ld) {
    case 2111: rys_block_host_var<2, 1, 1, 1, 0>(bq, kq, AB, CD, out, hits.data()); break;
    case 2121: rys_block_host_var<2, 1, 2, 1, 0>(bq, kq, AB, CD, out, hits.data()); break;
    case 2211: rys_block_host_var<2, 2, 1, 1, 0>(bq, kq, AB, CD, out, hits.data()); break;
    case 2220: rys_block_host_var<2, 2, 2, 0, 0>(bq, kq, AB, CD, out, hits.data()); break;
    case 2221: rys_block_host_var<2, 2, 2, 1, 0>(bq, kq, AB, CD, out, hits.data()); break;
    case 2222: rys_block_host_var<2, 2, 2, 2, 0>(bq, kq, AB, CD, out, hits.data()); break;
    }
    // every integral of the block produced by exactly one (variant, member, output)
    for (int i = 0; i < nout; ++i)
        if (hits[i] != 1) return -2;
    return 0;
}
