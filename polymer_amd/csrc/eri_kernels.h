// eri_kernels.h -- the shell-quartet kernel template of libpolyeri.
//
// One thread per contracted shell quartet (ab|cd) of one class: primitive
// loop, Boys + VRR + HRR (eri_engine.h), then the quartet's contribution
// to the Coulomb and exchange matrices with FP64 atomics (8-fold
// permutational symmetry: each unique quartet is computed once).  Each
// class is instantiated in its own translation unit (kgen_*.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>
#include <vector>

#include "eri_engine.h"

// MODE_JK32: J/K of the quartets below the mixed-precision cut, recursion in FP32.
// MODE_PBC: periodic lattice sums (pairs with cell offsets, no permutational
// symmetry, work item = (bra, ket, Q3) with Q3 innermost).
enum { MODE_JK = 0, MODE_SCHWARZ = 1, MODE_TENSOR = 2, MODE_JK32 = 3, MODE_PBC = 4 };

struct PairMeta {
    int prim_off, nprim, fa, fb;
};

// A segment of one bra row: items t0 .. t0 + n - 1 of the list, all with
// the bra pair of row `row`.
struct SegDesc {
    long long t0;
    int row, n;
};

struct EriArgs {
    const PrimPair* prims;
    const PairMeta* meta;
    const double* ab;          // 4 per pair: ABx ABy ABz 0
    const int* bra_ids;
    const int* ket_ids;
    const long long* row_off;  // nrows + 1
    int nrows;
    long long nwork;
    const int* rb_off;         // nrows x (nbucket+1): ket-bucket offsets inside each row
    const int* bucket_start;   // nbucket: first ket of each bucket in ket_ids
    const int* rb_first;       // nrows x nbucket: first ket of the row's range inside each
                               // bucket (null: every range starts at the bucket start)
    int nbucket;
    const double* boys;        // BOYS_NT x BOYS_NM (wavefront kernel)
    const double* boysA;       // aligned per-L tables (thread kernels)
    const double* D;
    double* J;
    double* K;
    int nbf;
    double* out;               // tensor (MODE_TENSOR, MODE_PBC) or Schwarz (MODE_SCHWARZ)
    int jk_copies;             // 1: one J/K (device-scope atomics); 8: one per XCD
    int tpack;                 // MODE_TENSOR: the pair-packed layout of TensorCb
    const int2* items;         // explicit (bra pair, ket position) items of a density-screened list, or null
    const unsigned* count;     // number of items (device), with items
    const double* ksoa;        // ket primitive pairs, bucket-interleaved fields (8 x kfs doubles)
    long long kfs;             // field stride of ksoa
    const int* bucket_soa;     // nbucket: offset of each ket bucket in ksoa
    // periodic (MODE_PBC): nwork = items x n3, Q3 = t % n3 - cel2
    int n3, cel1, cel2, both;  // both: digest every integral also with bra and ket roles exchanged
    double period;
    const int2* pcell;         // per pair {cell offset Q of the second shell, swap}
    const double* Dc;          // (2 cel1 + 1) x nbf^2 density cells, P(Q)[k][l] = P(k_0, l_Q)
    double *Jc, *Kc;           // jk_copies x (2 cel1 + 1) x nbf^2
    const SegDesc* segs;       // segment kernel: bra-row segments of the list
    int nseg;
    // segment kernel, density screening in the kernel (dtol > 0): skip items
    // with Q_ab Q_cd max|D| (six shell blocks) < dtol, count the rest in kept
    double dtol;
    const double* Qs;
    const int2* psh;           // shells of each pair
    const double* dsh;         // nsh x nsh shell-block maxima of |D|
    int nsh;
    unsigned* kept;
    // with dtol > 0: per segment the largest ket bound (segq, indexed like
    // segs) and the largest |D| block (*dmax); a segment with bra bound x
    // segq x dmax < dtol holds no surviving item and is skipped
    const double* segq;
    const double* dmax;
    // generally contracted lists (k_eri_gc): pair data, bra primitive pairs,
    // ket fields 8 .. 11 of ksoa = the member weights
    const GcMeta* gmeta;
    const GcPrim* gprims;
    double qtol;               // Schwarz threshold of the member quartets
    int qloop;                 // MODE_PBC J/K, thread kernel: Q3 values per thread (0: one, t = item x n3 + Q3)
};

// Number of work items of a launch: the device count of a screened list, or
// the static list size.
__device__ __forceinline__ long long work_count(const EriArgs& g) {
    return g.items ? (long long)*g.count : g.nwork;
}

// J/K accumulation.  With one copy per XCD (jk_copies == 8) a block adds into
// the copy of the XCD it runs on (hardware XCC_ID register, masked to 0..7)
// with workgroup-scope atomics, which complete in that XCD's L2; the copies
// are summed afterwards.  With one copy the atomics are device scope.
__device__ __forceinline__ int xcd_slot(int copies) {
    if (copies <= 1) return 0;
    const int xcc = __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20);   // HW_REG_XCC_ID[3:0]
    return xcc & (copies - 1);
}

__device__ __forceinline__ void jk_add(double* p, double v, int copies) {
    if (copies > 1)
        __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    else
        atomicAdd(p, v);
}

// Fixed-point accumulation (periodic J/K): the element holds a two's
// complement integer in units of 2^-JK_FX_BITS, each contribution is rounded
// to that grid and added with an integer atomic.  Integer addition is
// associative, so the sum (and the sum of the per-XCD copies) does not depend
// on the order in which the atomics arrive: two builds of the same density
// give bit-identical matrices.  Resolution 2^-50 = 8.9e-16 per contribution;
// the final elements must lie within +-2^13 (intermediate sums may wrap).
#define JK_FX_BITS 50
__device__ __forceinline__ void jk_add_fx(double* p, double v, int copies) {
    const unsigned long long x = (unsigned long long)__double2ll_rn(v * 0x1p50);
    if (copies > 1)
        __hip_atomic_fetch_add(reinterpret_cast<unsigned long long*>(p), x, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
    else
        __hip_atomic_fetch_add(reinterpret_cast<unsigned long long*>(p), x, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
}

// Digestion with the quartet's six density / Fock blocks held in registers
// (light classes).  With s the permutational degeneracy weight:
// Jacc[ab] += s v D[cd], Jacc[cd] += s v D[ab], Kacc[ac] += s v D[bd],
// Kacc[ad] += s v D[bc], Kacc[bc] += s v D[ad], Kacc[bd] += s v D[ac];
// afterwards J = 2(Jacc + Jacc^T) and K = Kacc + Kacc^T.
template <int NA, int NB, int NC, int ND, typename real = double>
struct RegDigest {
    real Dab[NA * NB], Dcd[NC * ND], Dac[NA * NC], Dad[NA * ND], Dbc[NB * NC], Dbd[NB * ND];
    real Jab[NA * NB], Jcd[NC * ND], Kac[NA * NC], Kad[NA * ND], Kbc[NB * NC], Kbd[NB * ND];
    __device__ __forceinline__ void load(const double* __restrict__ D, int n, int fa, int fb, int fc, int fd) {
#pragma unroll
        for (int a = 0; a < NA; ++a) {
#pragma unroll
            for (int b = 0; b < NB; ++b) { Dab[a * NB + b] = (real)D[(size_t)(fa + a) * n + fb + b]; Jab[a * NB + b] = 0; }
#pragma unroll
            for (int c = 0; c < NC; ++c) { Dac[a * NC + c] = (real)D[(size_t)(fa + a) * n + fc + c]; Kac[a * NC + c] = 0; }
#pragma unroll
            for (int d = 0; d < ND; ++d) { Dad[a * ND + d] = (real)D[(size_t)(fa + a) * n + fd + d]; Kad[a * ND + d] = 0; }
        }
#pragma unroll
        for (int b = 0; b < NB; ++b) {
#pragma unroll
            for (int c = 0; c < NC; ++c) { Dbc[b * NC + c] = (real)D[(size_t)(fb + b) * n + fc + c]; Kbc[b * NC + c] = 0; }
#pragma unroll
            for (int d = 0; d < ND; ++d) { Dbd[b * ND + d] = (real)D[(size_t)(fb + b) * n + fd + d]; Kbd[b * ND + d] = 0; }
        }
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int d = 0; d < ND; ++d) { Dcd[c * ND + d] = (real)D[(size_t)(fc + c) * n + fd + d]; Jcd[c * ND + d] = 0; }
    }
    __device__ __forceinline__ void operator()(int a, int b, int c, int d, real x) {
        Jab[a * NB + b] += x * Dcd[c * ND + d];
        Jcd[c * ND + d] += x * Dab[a * NB + b];
        Kac[a * NC + c] += x * Dbd[b * ND + d];
        Kad[a * ND + d] += x * Dbc[b * NC + c];
        Kbc[b * NC + c] += x * Dad[a * ND + d];
        Kbd[b * ND + d] += x * Dac[a * NC + c];
    }
    __device__ __forceinline__ void flush(double* J, double* K, int n, int fa, int fb, int fc, int fd, double s,
                                          int copies) {
#pragma unroll
        for (int a = 0; a < NA; ++a) {
#pragma unroll
            for (int b = 0; b < NB; ++b) jk_add(&J[(size_t)(fa + a) * n + fb + b], s * (double)Jab[a * NB + b], copies);
#pragma unroll
            for (int c = 0; c < NC; ++c) jk_add(&K[(size_t)(fa + a) * n + fc + c], s * (double)Kac[a * NC + c], copies);
#pragma unroll
            for (int d = 0; d < ND; ++d) jk_add(&K[(size_t)(fa + a) * n + fd + d], s * (double)Kad[a * ND + d], copies);
        }
#pragma unroll
        for (int b = 0; b < NB; ++b) {
#pragma unroll
            for (int c = 0; c < NC; ++c) jk_add(&K[(size_t)(fb + b) * n + fc + c], s * (double)Kbc[b * NC + c], copies);
#pragma unroll
            for (int d = 0; d < ND; ++d) jk_add(&K[(size_t)(fb + b) * n + fd + d], s * (double)Kbd[b * ND + d], copies);
        }
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int d = 0; d < ND; ++d) jk_add(&J[(size_t)(fc + c) * n + fd + d], s * (double)Jcd[c * ND + d], copies);
    }
};

// Digestion for heavy classes: each integral goes straight to the six
// matrix elements it touches (density read through the cache), so no block
// of integrals or accumulators is held per thread.
struct StreamDigest {
    const double* __restrict__ D;
    double* J;
    double* K;
    int n, fa, fb, fc, fd;
    double s;
    int copies;
    __device__ __forceinline__ void operator()(int a, int b, int c, int d, double x) {
        const size_t A = fa + a, B = fb + b, C = fc + c, E = fd + d;
        const double sx = s * x;
        jk_add(&J[A * n + B], sx * D[C * n + E], copies);
        jk_add(&J[C * n + E], sx * D[A * n + B], copies);
        jk_add(&K[A * n + C], sx * D[B * n + E], copies);
        jk_add(&K[A * n + E], sx * D[B * n + C], copies);
        jk_add(&K[B * n + C], sx * D[A * n + E], copies);
        jk_add(&K[B * n + E], sx * D[A * n + C], copies);
    }
};

struct SchwarzCb {
    double m = 0.0;
    __device__ __forceinline__ void operator()(int a, int b, int c, int d, double x) {
        if (a == c && b == d) m = fmax(m, fabs(x));
    }
};

// Writes one integral (ij|kl) to every position of the ERI tensor it fills:
// the full n^4 tensor at ((i n + j) n + k) n + l (8 positions), or, pk, the
// tensor with the first pair packed, A(P, k, l) at (P n + k) n + l with
// P = max(i, j) (max(i, j) + 1) / 2 + min(i, j) (4 positions, half the size).
struct TensorCb {
    double* out;
    size_t n;
    int fa, fb, fc, fd;
    int pk;
    __device__ __forceinline__ void operator()(int a, int b, int c, int d, double x) {
        const size_t n1 = n, n2 = n1 * n, n3 = n2 * n;
        const size_t i = fa + a, j = fb + b, k = fc + c, l = fd + d;
        if (pk) {
            const size_t pij = i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i;
            const size_t pkl = k >= l ? k * (k + 1) / 2 + l : l * (l + 1) / 2 + k;
            out[pij * n2 + k * n1 + l] = x;
            out[pij * n2 + l * n1 + k] = x;
            out[pkl * n2 + i * n1 + j] = x;
            out[pkl * n2 + j * n1 + i] = x;
            return;
        }
        out[i * n3 + j * n2 + k * n1 + l] = x;
        out[j * n3 + i * n2 + k * n1 + l] = x;
        out[i * n3 + j * n2 + l * n1 + k] = x;
        out[j * n3 + i * n2 + l * n1 + k] = x;
        out[k * n3 + l * n2 + i * n1 + j] = x;
        out[l * n3 + k * n2 + i * n1 + j] = x;
        out[k * n3 + l * n2 + j * n1 + i] = x;
        out[l * n3 + k * n2 + j * n1 + i] = x;
    }
};

// Work item t of a class pair -> (bra pair, ket position, ket bucket).
// Rows are bra pairs in Schwarz order; inside a row the surviving kets are
// grouped in buckets of equal primitive-pair count (each bucket a Schwarz
// prefix, or a range from rb_first), so neighbouring threads run primitive
// loops of the same length.  The ket pair is ket_ids[pos].
// Item t of bra row `row` -> (ket position, ket bucket).
__device__ __forceinline__ void decode_in_row(const EriArgs& g, int row, long long t, int& pos, int& b) {
    const int local = (int)(t - g.row_off[row]);
    const int* rb = g.rb_off + (size_t)row * (g.nbucket + 1);
    b = 0;
    while (b + 1 < g.nbucket && rb[b + 1] <= local) ++b;
    const int first = g.rb_first ? g.rb_first[(size_t)row * g.nbucket + b] : 0;
    pos = g.bucket_start[b] + first + (local - rb[b]);
}

__device__ __forceinline__ void decode_work(const EriArgs& g, long long t, int& bp, int& pos, int& b) {
    int lo = 0, hi = g.nrows - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (g.row_off[mid] <= t) lo = mid; else hi = mid - 1;
    }
    bp = g.bra_ids[lo];
    decode_in_row(g, lo, t, pos, b);
}

// Bucket of ket position pos.
__device__ __forceinline__ int bucket_of(const EriArgs& g, int pos) {
    int b = 0;
    while (b + 1 < g.nbucket && g.bucket_start[b + 1] <= pos) ++b;
    return b;
}

// Periodic digestion of one integral (i_0 j_Q1 | k_Q3 l_Q3+Q2) (src/pfourindexdirect.f90:691-704):
// J(Q1)[i][j] += v P(Q2)[k][l]; K(Q3+Q2)[i][l] += v P(Q1-Q3)[k][j] when both
// cells are within [-cel1, cel1].
__device__ __forceinline__ void pbc_add(const EriArgs& g, double* J, double* K, int q1, int q2, int q3, int fi,
                                        int fj, int fk, int fl, double v) {
    const size_t n = g.nbf, n2 = n * n;
    const int c1 = g.cel1;
    jk_add_fx(&J[(size_t)(q1 + c1) * n2 + fi * n + fj], v * g.Dc[(size_t)(q2 + c1) * n2 + fk * n + fl], g.jk_copies);
    const int qk = q3 + q2, qd = q1 - q3;
    if (qk >= -c1 && qk <= c1 && qd >= -c1 && qd <= c1)
        jk_add_fx(&K[(size_t)(qk + c1) * n2 + fi * n + fl], v * g.Dc[(size_t)(qd + c1) * n2 + fk * n + fj],
               g.jk_copies);
}

// Lattice tensor store (MODE_PBC with out set, src/pfourindexstore.f90:258-747):
// T[i][Q1][j][Q3][k][Q2][l] = (i_0 j_Q1 | k_Q3 l_Q3+Q2), Q1 and Q2 in
// [-cel1, cel1] (the launch passes the reduced pair range), Q3 in
// [-cel2, cel2].  Each integral has exactly one producer, so plain stores.
__device__ __forceinline__ void pbc_store(const EriArgs& g, int q1, int q2, int q3, int fi, int fj, int fk, int fl,
                                          double v) {
    const size_t n = g.nbf, n1 = 2 * g.cel1 + 1;
    size_t x = (size_t)fi * n1 + (q1 + g.cel1);
    x = (x * n + fj) * g.n3 + (q3 + g.cel2);
    x = (x * n + fk) * n1 + (q2 + g.cel1);
    g.out[x * n + fl] = v;
}

// Integral (a b | c d) of a stored bra pair (first functions fa, fb; cells cb)
// and a stored ket pair (fc, fd; ck) translated by s cells: as the reference
// quartet with this bra and ket, and (both) as the one with the roles
// exchanged, which is the same integral translated by -s cells.
__device__ __forceinline__ void pbc_integral(const EriArgs& g, double* J, double* K, int2 cb, int2 ck, int s, int fa,
                                             int fb, int fc, int fd, int a, int b, int c, int d, double v) {
    const int fi = cb.y ? fb + b : fa + a, fj = cb.y ? fa + a : fb + b;
    const int fk = ck.y ? fd + d : fc + c, fl = ck.y ? fc + c : fd + d;
    if (g.out) {
        pbc_store(g, cb.x, ck.x, s, fi, fj, fk, fl, v);
        if (g.both) pbc_store(g, ck.x, cb.x, -s, fk, fl, fi, fj, v);
        return;
    }
    pbc_add(g, J, K, cb.x, ck.x, s, fi, fj, fk, fl, v);
    if (g.both) pbc_add(g, J, K, ck.x, cb.x, -s, fk, fl, fi, fj, v);
}

// Periodic digestion of one (quartet, Q3) in registers, one atomic per
// matrix element (the same sums as pbc_integral over the quartet's
// integrals).  SB / SK: the stored bra / ket pair is swapped, so the
// reference's i (cell 0) and j (cell Q1) functions are b and a (and l, k
// are c, d).  Blocks: J(Q1)[i][j] += v D(Q2)[k][l]; K(Q3+Q2)[i][l] += v
// D(Q1-Q3)[k][j]; with both roles also J(Q2)[k][l] += v D(Q1)[i][j] and
// K(Q1-Q3)[k][j] += v D(Q2+Q3)[i][l].
template <int NA, int NB, int NC, int ND, bool SB, bool SK>
__device__ __forceinline__ void pbc_digest(const EriArgs& g, const double* v, int fa, int fb, int fc, int fd, int q1,
                                           int q2, int s, double* J, double* K) {
    constexpr int NI = SB ? NB : NA, NJ = SB ? NA : NB, NK = SK ? ND : NC, NL = SK ? NC : ND;
    const int fi = SB ? fb : fa, fj = SB ? fa : fb, fk = SK ? fd : fc, fl = SK ? fc : fd;
    const size_t n = g.nbf, n2 = n * n;
    const int c1 = g.cel1, cp = g.jk_copies;
    auto V = [&](int i, int j, int k, int l) -> double {
        const int a = SB ? j : i, b = SB ? i : j, c = SK ? l : k, d = SK ? k : l;
        return v[((a * NB + b) * NC + c) * ND + d];
    };
    {
        const double* D = g.Dc + (size_t)(q2 + c1) * n2;
        double dkl[NK * NL], acc[NI * NJ];
#pragma unroll
        for (int k = 0; k < NK; ++k)
#pragma unroll
            for (int l = 0; l < NL; ++l) dkl[k * NL + l] = D[(size_t)(fk + k) * n + fl + l];
#pragma unroll
        for (int i = 0; i < NI; ++i)
#pragma unroll
            for (int j = 0; j < NJ; ++j) {
                double x = 0.0;
#pragma unroll
                for (int k = 0; k < NK; ++k)
#pragma unroll
                    for (int l = 0; l < NL; ++l) x += V(i, j, k, l) * dkl[k * NL + l];
                acc[i * NJ + j] = x;
            }
        double* Jq = J + (size_t)(q1 + c1) * n2;
#pragma unroll
        for (int i = 0; i < NI; ++i)
#pragma unroll
            for (int j = 0; j < NJ; ++j) jk_add_fx(&Jq[(size_t)(fi + i) * n + fj + j], acc[i * NJ + j], cp);
    }
    const int qk = s + q2, qd = q1 - s;
    if (qk >= -c1 && qk <= c1 && qd >= -c1 && qd <= c1) {
        const double* D = g.Dc + (size_t)(qd + c1) * n2;
        double dkj[NK * NJ], acc[NI * NL];
#pragma unroll
        for (int k = 0; k < NK; ++k)
#pragma unroll
            for (int j = 0; j < NJ; ++j) dkj[k * NJ + j] = D[(size_t)(fk + k) * n + fj + j];
#pragma unroll
        for (int i = 0; i < NI; ++i)
#pragma unroll
            for (int l = 0; l < NL; ++l) {
                double x = 0.0;
#pragma unroll
                for (int j = 0; j < NJ; ++j)
#pragma unroll
                    for (int k = 0; k < NK; ++k) x += V(i, j, k, l) * dkj[k * NJ + j];
                acc[i * NL + l] = x;
            }
        double* Kq = K + (size_t)(qk + c1) * n2;
#pragma unroll
        for (int i = 0; i < NI; ++i)
#pragma unroll
            for (int l = 0; l < NL; ++l) jk_add_fx(&Kq[(size_t)(fi + i) * n + fl + l], acc[i * NL + l], cp);
    }
    if (!g.both) return;
    {
        const double* D = g.Dc + (size_t)(q1 + c1) * n2;
        double dij[NI * NJ], acc[NK * NL];
#pragma unroll
        for (int i = 0; i < NI; ++i)
#pragma unroll
            for (int j = 0; j < NJ; ++j) dij[i * NJ + j] = D[(size_t)(fi + i) * n + fj + j];
#pragma unroll
        for (int k = 0; k < NK; ++k)
#pragma unroll
            for (int l = 0; l < NL; ++l) {
                double x = 0.0;
#pragma unroll
                for (int i = 0; i < NI; ++i)
#pragma unroll
                    for (int j = 0; j < NJ; ++j) x += V(i, j, k, l) * dij[i * NJ + j];
                acc[k * NL + l] = x;
            }
        double* Jq = J + (size_t)(q2 + c1) * n2;
#pragma unroll
        for (int k = 0; k < NK; ++k)
#pragma unroll
            for (int l = 0; l < NL; ++l) jk_add_fx(&Jq[(size_t)(fk + k) * n + fl + l], acc[k * NL + l], cp);
    }
    const int qk2 = q1 - s, qd2 = q2 + s;
    if (qk2 >= -c1 && qk2 <= c1 && qd2 >= -c1 && qd2 <= c1) {
        const double* D = g.Dc + (size_t)(qd2 + c1) * n2;
        double dil[NI * NL], acc[NK * NJ];
#pragma unroll
        for (int i = 0; i < NI; ++i)
#pragma unroll
            for (int l = 0; l < NL; ++l) dil[i * NL + l] = D[(size_t)(fi + i) * n + fl + l];
#pragma unroll
        for (int k = 0; k < NK; ++k)
#pragma unroll
            for (int j = 0; j < NJ; ++j) {
                double x = 0.0;
#pragma unroll
                for (int i = 0; i < NI; ++i)
#pragma unroll
                    for (int l = 0; l < NL; ++l) x += V(i, j, k, l) * dil[i * NL + l];
                acc[k * NJ + j] = x;
            }
        double* Kq = K + (size_t)(qk2 + c1) * n2;
#pragma unroll
        for (int k = 0; k < NK; ++k)
#pragma unroll
            for (int j = 0; j < NJ; ++j) jk_add_fx(&Kq[(size_t)(fk + k) * n + fj + j], acc[k * NJ + j], cp);
    }
}

// Periodic J/K of one (bra, ket) item over Q3 in [s0, s1] in one thread:
// J(Q1) (and, with both roles, J(Q2)) is the same matrix block for every
// Q3, so it is summed in registers and added once; the exchange blocks go
// to a different cell per Q3 and are added per Q3.  Same sums as
// pbc_digest.
template <int LA, int LB, int LC, int LD, bool SB, bool SK>
__device__ __forceinline__ void pbc_loop(const EriArgs& g, int bp, int kp, const PairMeta& mb, const PairMeta& mk,
                                         int q1, int q2, int s0, int s1) {
    constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    constexpr int NI = SB ? NB : NA, NJ = SB ? NA : NB, NK = SK ? ND : NC, NL = SK ? NC : ND;
    const int fi = SB ? mb.fb : mb.fa, fj = SB ? mb.fa : mb.fb, fk = SK ? mk.fb : mk.fa, fl = SK ? mk.fa : mk.fb;
    const size_t n = g.nbf, n2 = n * n;
    const int c1 = g.cel1, cp = g.jk_copies;
    const size_t off = (size_t)xcd_slot(cp) * (2 * c1 + 1) * n2;
    double* J = g.Jc + off;
    double* K = g.Kc + off;
    double dkl[NK * NL], jij[NI * NJ], dij[NI * NJ], jkl[NK * NL];
    {
        const double* D = g.Dc + (size_t)(q2 + c1) * n2;
#pragma unroll
        for (int k = 0; k < NK; ++k)
#pragma unroll
            for (int l = 0; l < NL; ++l) dkl[k * NL + l] = D[(size_t)(fk + k) * n + fl + l];
        const double* E = g.Dc + (size_t)(q1 + c1) * n2;
#pragma unroll
        for (int i = 0; i < NI; ++i)
#pragma unroll
            for (int j = 0; j < NJ; ++j) {
                dij[i * NJ + j] = g.both ? E[(size_t)(fi + i) * n + fj + j] : 0.0;
                jij[i * NJ + j] = 0.0;
            }
#pragma unroll
        for (int i = 0; i < NK * NL; ++i) jkl[i] = 0.0;
    }
    const double ABx = g.ab[4 * bp], ABy = g.ab[4 * bp + 1], ABz = g.ab[4 * bp + 2];
    const double CDx = g.ab[4 * kp], CDy = g.ab[4 * kp + 1], CDz = g.ab[4 * kp + 2];
    for (int s = s0; s <= s1; ++s) {
        double v[NA * NB * NC * ND];
        StoreCb<NB, NC, ND, double> st{v};
        const ShiftKet kv{g.prims + mk.prim_off, s * g.period};
        quartet<double, LA, LB, LC, LD>(g.prims + mb.prim_off, mb.nprim, kv, mk.nprim, ABx, ABy, ABz, CDx, CDy, CDz,
                                        g.boysA, st);
        auto V = [&](int i, int j, int k, int l) -> double {
            const int a = SB ? j : i, b = SB ? i : j, c = SK ? l : k, d = SK ? k : l;
            return v[((a * NB + b) * NC + c) * ND + d];
        };
#pragma unroll
        for (int i = 0; i < NI; ++i)
#pragma unroll
            for (int j = 0; j < NJ; ++j) {
                double x = 0.0;
#pragma unroll
                for (int k = 0; k < NK; ++k)
#pragma unroll
                    for (int l = 0; l < NL; ++l) x += V(i, j, k, l) * dkl[k * NL + l];
                jij[i * NJ + j] += x;
            }
        const int qk = s + q2, qd = q1 - s;
        if (qk >= -c1 && qk <= c1 && qd >= -c1 && qd <= c1) {
            const double* D = g.Dc + (size_t)(qd + c1) * n2;
            double dkj[NK * NJ];
#pragma unroll
            for (int k = 0; k < NK; ++k)
#pragma unroll
                for (int j = 0; j < NJ; ++j) dkj[k * NJ + j] = D[(size_t)(fk + k) * n + fj + j];
            double* Kq = K + (size_t)(qk + c1) * n2;
#pragma unroll
            for (int i = 0; i < NI; ++i)
#pragma unroll
                for (int l = 0; l < NL; ++l) {
                    double x = 0.0;
#pragma unroll
                    for (int j = 0; j < NJ; ++j)
#pragma unroll
                        for (int k = 0; k < NK; ++k) x += V(i, j, k, l) * dkj[k * NJ + j];
                    jk_add_fx(&Kq[(size_t)(fi + i) * n + fl + l], x, cp);
                }
        }
        if (!g.both) continue;
#pragma unroll
        for (int k = 0; k < NK; ++k)
#pragma unroll
            for (int l = 0; l < NL; ++l) {
                double x = 0.0;
#pragma unroll
                for (int i = 0; i < NI; ++i)
#pragma unroll
                    for (int j = 0; j < NJ; ++j) x += V(i, j, k, l) * dij[i * NJ + j];
                jkl[k * NL + l] += x;
            }
        const int qk2 = q1 - s, qd2 = q2 + s;
        if (qk2 >= -c1 && qk2 <= c1 && qd2 >= -c1 && qd2 <= c1) {
            const double* D = g.Dc + (size_t)(qd2 + c1) * n2;
            double dil[NI * NL];
#pragma unroll
            for (int i = 0; i < NI; ++i)
#pragma unroll
                for (int l = 0; l < NL; ++l) dil[i * NL + l] = D[(size_t)(fi + i) * n + fl + l];
            double* Kq = K + (size_t)(qk2 + c1) * n2;
#pragma unroll
            for (int k = 0; k < NK; ++k)
#pragma unroll
                for (int j = 0; j < NJ; ++j) {
                    double x = 0.0;
#pragma unroll
                    for (int i = 0; i < NI; ++i)
#pragma unroll
                        for (int l = 0; l < NL; ++l) x += V(i, j, k, l) * dil[i * NL + l];
                    jk_add_fx(&Kq[(size_t)(fk + k) * n + fj + j], x, cp);
                }
        }
    }
    double* J1 = J + (size_t)(q1 + c1) * n2;
#pragma unroll
    for (int i = 0; i < NI; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j) jk_add_fx(&J1[(size_t)(fi + i) * n + fj + j], jij[i * NJ + j], cp);
    if (!g.both) return;
    double* J2 = J + (size_t)(q2 + c1) * n2;
#pragma unroll
    for (int k = 0; k < NK; ++k)
#pragma unroll
        for (int l = 0; l < NL; ++l) jk_add_fx(&J2[(size_t)(fk + k) * n + fl + l], jkl[k * NL + l], cp);
}

// Work index t = item x nchunk + chunk, chunk of g.qloop Q3 values.
template <int LA, int LB, int LC, int LD>
__device__ __forceinline__ void pbc_item_loop(const EriArgs& g, long long t) {
    const int nchunk = (g.n3 + g.qloop - 1) / g.qloop;
    const long long u = t / nchunk;
    const int chunk = (int)(t - u * nchunk);
    const int s0 = -g.cel2 + chunk * g.qloop, s1 = min(s0 + g.qloop - 1, g.cel2);
    int bp, pos, b;
    decode_work(g, u, bp, pos, b);
    const int kp = g.ket_ids[pos];
    const PairMeta mb = g.meta[bp], mk = g.meta[kp];
    const int2 cb = g.pcell[bp], ck = g.pcell[kp];
    switch (cb.y * 2 + ck.y) {
    case 0: pbc_loop<LA, LB, LC, LD, false, false>(g, bp, kp, mb, mk, cb.x, ck.x, s0, s1); break;
    case 1: pbc_loop<LA, LB, LC, LD, false, true>(g, bp, kp, mb, mk, cb.x, ck.x, s0, s1); break;
    case 2: pbc_loop<LA, LB, LC, LD, true, false>(g, bp, kp, mb, mk, cb.x, ck.x, s0, s1); break;
    default: pbc_loop<LA, LB, LC, LD, true, true>(g, bp, kp, mb, mk, cb.x, ck.x, s0, s1); break;
    }
}

template <int LA, int LB, int LC, int LD>
__device__ __forceinline__ void pbc_item(const EriArgs& g, long long t) {
    const long long u = t / g.n3;
    const int s = (int)(t - u * g.n3) - g.cel2;
    int bp, pos, b;
    decode_work(g, u, bp, pos, b);
    const int kp = g.ket_ids[pos];
    const PairMeta mb = g.meta[bp], mk = g.meta[kp];
    constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    double v[NA * NB * NC * ND];
    StoreCb<NB, NC, ND, double> st{v};
    const ShiftKet kv{g.prims + mk.prim_off, s * g.period};
    quartet<double, LA, LB, LC, LD>(g.prims + mb.prim_off, mb.nprim, kv, mk.nprim, g.ab[4 * bp], g.ab[4 * bp + 1],
                                    g.ab[4 * bp + 2], g.ab[4 * kp], g.ab[4 * kp + 1], g.ab[4 * kp + 2], g.boysA, st);
    const size_t off = (size_t)xcd_slot(g.jk_copies) * (2 * g.cel1 + 1) * g.nbf * g.nbf;
    double* J = g.Jc + off;
    double* K = g.Kc + off;
    const int2 cb = g.pcell[bp], ck = g.pcell[kp];
    if (g.out) {   // lattice tensor: one store per integral
        for (int a = 0; a < NA; ++a)
            for (int bb = 0; bb < NB; ++bb)
                for (int c = 0; c < NC; ++c)
                    for (int d = 0; d < ND; ++d)
                        pbc_integral(g, J, K, cb, ck, s, mb.fa, mb.fb, mk.fa, mk.fb, a, bb, c, d,
                                     v[((a * NB + bb) * NC + c) * ND + d]);
        return;
    }
    switch (cb.y * 2 + ck.y) {
    case 0: pbc_digest<NA, NB, NC, ND, false, false>(g, v, mb.fa, mb.fb, mk.fa, mk.fb, cb.x, ck.x, s, J, K); break;
    case 1: pbc_digest<NA, NB, NC, ND, false, true>(g, v, mb.fa, mb.fb, mk.fa, mk.fb, cb.x, ck.x, s, J, K); break;
    case 2: pbc_digest<NA, NB, NC, ND, true, false>(g, v, mb.fa, mb.fb, mk.fa, mk.fb, cb.x, ck.x, s, J, K); break;
    default: pbc_digest<NA, NB, NC, ND, true, true>(g, v, mb.fa, mb.fb, mk.fa, mk.fb, cb.x, ck.x, s, J, K); break;
    }
}

template <int LA, int LB, int LC, int LD, int MODE>
__device__ __forceinline__ void eri_item(const EriArgs& g, long long t) {
    if (MODE == MODE_PBC) {
        if (g.qloop) pbc_item_loop<LA, LB, LC, LD>(g, t);
        else pbc_item<LA, LB, LC, LD>(g, t);
        return;
    }
    int bp, kp, pos = 0, b = 0;
    if (MODE == MODE_SCHWARZ) {
        bp = kp = g.bra_ids[t];
    } else {
        if (g.items) {
            const int2 it = g.items[t];
            bp = it.x; pos = it.y;
            b = bucket_of(g, pos);
        } else {
            decode_work(g, t, bp, pos, b);
        }
        kp = g.ket_ids[pos];
    }
    const PairMeta mb = g.meta[bp], mk = g.meta[kp];
    constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    const double ABx = g.ab[4 * bp], ABy = g.ab[4 * bp + 1], ABz = g.ab[4 * bp + 2];
    const double CDx = g.ab[4 * kp], CDy = g.ab[4 * kp + 1], CDz = g.ab[4 * kp + 2];
    const PrimPair* bra = g.prims + mb.prim_off;
    const PrimPair* ket = g.prims + mk.prim_off;
    const int fa = mb.fa, fb = mb.fb, fc = mk.fa, fd = mk.fb;
    const int n = g.nbf;
    if (MODE == MODE_SCHWARZ) {
        SchwarzCb cb;
        quartet_fp64<LA, LB, LC, LD>(bra, mb.nprim, ket, mk.nprim, ABx, ABy, ABz, CDx, CDy, CDz, g.boysA, cb);
        g.out[bp] = sqrt(cb.m);
        return;
    }
    const SoaKet kv{g.ksoa + g.bucket_soa[b] + (pos - g.bucket_start[b]), g.bucket_start[b + 1] - g.bucket_start[b],
                    g.kfs};
    if (MODE == MODE_TENSOR) {
        TensorCb cb{g.out, (size_t)n, fa, fb, fc, fd, g.tpack};
        quartet<double, LA, LB, LC, LD>(bra, mb.nprim, kv, mk.nprim, ABx, ABy, ABz, CDx, CDy, CDz, g.boysA, cb);
    } else {
        const double s = (fa == fb ? 0.5 : 1.0) * (fc == fd ? 0.5 : 1.0) * (bp == kp ? 0.5 : 1.0);
        // light classes only (heavier ones run in kcoop.hip): the (ab|cd)
        // block in registers, then the six J/K blocks
        using real = typename std::conditional<MODE == MODE_JK32, float, double>::type;
        real v[NA * NB * NC * ND];
        StoreCb<NB, NC, ND, real> st{v};
        quartet<real, LA, LB, LC, LD>(bra, mb.nprim, kv, mk.nprim, ABx, ABy, ABz, CDx, CDy, CDz, g.boysA, st);
        RegDigest<NA, NB, NC, ND, real> cb;
        cb.load(g.D, n, fa, fb, fc, fd);
        const size_t off = (size_t)xcd_slot(g.jk_copies) * n * n;
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int c = 0; c < NC; ++c)
#pragma unroll
                    for (int d = 0; d < ND; ++d) cb(a, b, c, d, v[((a * NB + b) * NC + c) * ND + d]);
        cb.flush(g.J + off, g.K + off, n, fa, fb, fc, fd, s, g.jk_copies);
    }
}

// One thread per work item; grid-stride, so a density-screened list (item
// count known only on the device) runs with a fixed grid.
template <int LA, int LB, int LC, int LD, int MODE>
__global__ void __launch_bounds__(256) k_eri(EriArgs g) {
    const long long nw = work_count(g);
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < nw;
         t += (long long)gridDim.x * blockDim.x)
        eri_item<LA, LB, LC, LD, MODE>(g, t);
}

// Segment kernel for the light classes (J/K modes): one block per segment of
// a bra row, so every item of the block shares the bra pair (a b|.  The
// exchange contributions of a quartet all land in the rows of the two bra
// shells (K[a][c], K[a][d], K[b][c], K[b][d]) and J_ab in one fixed block:
// those are summed in LDS (rows NA + NB of width nbf, LDS atomics) and in
// registers (J_ab, reduced over the wavefront), and added to the XCD's J/K
// copy once per segment, one atomic per nonzero element.  Only J_cd is added
// per quartet.  LDS: ((NA + NB) nbf + NA NB) doubles.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <int LA, int LB, int LC, int LD, int MODE>
__global__ void __launch_bounds__(256) k_eri_seg(EriArgs g) {
    constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC), ND = ncart(LD);
    using real = typename std::conditional<MODE == MODE_JK32, float, double>::type;
    extern __shared__ double seg_lds[];
    const int n = g.nbf;
    double* Kr = seg_lds;                        // rows fa.. (NA) then fb.. (NB), width n
    double* Jl = seg_lds + (size_t)(NA + NB) * n;
    const int total = (NA + NB) * n + NA * NB;
    const size_t off = (size_t)xcd_slot(g.jk_copies) * n * n;
    double* J = g.J + off;
    double* K = g.K + off;
    const int lane = threadIdx.x & 63;
    for (int sg = blockIdx.x; sg < g.nseg; sg += gridDim.x) {
        const SegDesc seg = g.segs[sg];
        const int bp = g.bra_ids[seg.row];
        if (g.dtol > 0.0 && g.segq && g.Qs[bp] * g.segq[sg] * g.dmax[0] < g.dtol) continue;   // block-uniform
        for (int i = threadIdx.x; i < total; i += blockDim.x) seg_lds[i] = 0.0;
        __syncthreads();
        const PairMeta mb = g.meta[bp];
        const int fa = mb.fa, fb = mb.fb;
        const double ABx = g.ab[4 * bp], ABy = g.ab[4 * bp + 1], ABz = g.ab[4 * bp + 2];
        const PrimPair* bra = g.prims + mb.prim_off;
        double jab[NA * NB];
#pragma unroll
        for (int i = 0; i < NA * NB; ++i) jab[i] = 0.0;
        const bool screen = g.dtol > 0.0;
        for (long long t0 = seg.t0; t0 < seg.t0 + seg.n; t0 += blockDim.x) {
            const long long t = t0 + threadIdx.x;
            int pos = 0, b = 0, kp = 0;
            bool live = t < seg.t0 + seg.n;
            if (live) {
                decode_in_row(g, seg.row, t, pos, b);
                kp = g.ket_ids[pos];
                if (screen) {
                    const int2 ab = g.psh[bp], cd = g.psh[kp];
                    const size_t ns = g.nsh, a = ab.x, bb = ab.y, c = cd.x, d = cd.y;
                    double m = fmax(g.dsh[c * ns + d], g.dsh[a * ns + bb]);
                    m = fmax(m, fmax(g.dsh[bb * ns + d], g.dsh[bb * ns + c]));
                    m = fmax(m, fmax(g.dsh[a * ns + d], g.dsh[a * ns + c]));
                    live = g.Qs[bp] * g.Qs[kp] * m >= g.dtol;
                }
            }
            if (screen) {
                const unsigned long long mask = __ballot(live);
                if (lane == 0 && mask) atomicAdd(g.kept, (unsigned)__popcll(mask));
            }
            if (!live) continue;
            const PairMeta mk = g.meta[kp];
            const int fc = mk.fa, fd = mk.fb;
            const SoaKet kv{g.ksoa + g.bucket_soa[b] + (pos - g.bucket_start[b]),
                            g.bucket_start[b + 1] - g.bucket_start[b], g.kfs};
            real v[NA * NB * NC * ND];
            StoreCb<NB, NC, ND, real> st{v};
            quartet<real, LA, LB, LC, LD>(bra, mb.nprim, kv, mk.nprim, ABx, ABy, ABz, g.ab[4 * kp],
                                          g.ab[4 * kp + 1], g.ab[4 * kp + 2], g.boysA, st);
            RegDigest<NA, NB, NC, ND, real> cb;
            cb.load(g.D, n, fa, fb, fc, fd);
#pragma unroll
            for (int a = 0; a < NA; ++a)
#pragma unroll
                for (int bb = 0; bb < NB; ++bb)
#pragma unroll
                    for (int c = 0; c < NC; ++c)
#pragma unroll
                        for (int d = 0; d < ND; ++d) cb(a, bb, c, d, v[((a * NB + bb) * NC + c) * ND + d]);
            const double s = (fa == fb ? 0.5 : 1.0) * (fc == fd ? 0.5 : 1.0) * (bp == kp ? 0.5 : 1.0);
#pragma unroll
            for (int i = 0; i < NA * NB; ++i) jab[i] += s * (double)cb.Jab[i];
#pragma unroll
            for (int a = 0; a < NA; ++a) {
#pragma unroll
                for (int c = 0; c < NC; ++c) atomicAdd(&Kr[(size_t)a * n + fc + c], s * (double)cb.Kac[a * NC + c]);
#pragma unroll
                for (int d = 0; d < ND; ++d) atomicAdd(&Kr[(size_t)a * n + fd + d], s * (double)cb.Kad[a * ND + d]);
            }
#pragma unroll
            for (int bb = 0; bb < NB; ++bb) {
#pragma unroll
                for (int c = 0; c < NC; ++c)
                    atomicAdd(&Kr[(size_t)(NA + bb) * n + fc + c], s * (double)cb.Kbc[bb * NC + c]);
#pragma unroll
                for (int d = 0; d < ND; ++d)
                    atomicAdd(&Kr[(size_t)(NA + bb) * n + fd + d], s * (double)cb.Kbd[bb * ND + d]);
            }
#pragma unroll
            for (int c = 0; c < NC; ++c)
#pragma unroll
                for (int d = 0; d < ND; ++d)
                    jk_add(&J[(size_t)(fc + c) * n + fd + d], s * (double)cb.Jcd[c * ND + d], g.jk_copies);
        }
#pragma unroll
        for (int i = 0; i < NA * NB; ++i) {
            const double w = wave_sum(jab[i]);
            if (lane == 0) atomicAdd(&Jl[i], w);
        }
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

// Segment kernel of the generally contracted lists, classes (ls|l's): as
// k_eri_seg, but an item is a pair of generally contracted pairs (GcMeta)
// and computes all member quartets from one primitive loop: for each ket
// primitive, the primitive quartets' [a0|c0] blocks are summed over the bra
// primitives with each bra member's weights, and that sum is added to every
// member quartet with the ket members' weights.  Each member quartet
// that passes the Schwarz test is digested with its own degeneracy weight; a list row whose bra equals
// the ket visits both orders of two different members, each at half
// weight.  LDS: the exchange rows of the up to two shells of each bra group
// ((2 NA + 2) nbf doubles) and J_ab of every bra member (MB NA doubles).
// Largest shell-block density maximum between shells {x0, x1} and {y0, y1}
// (x1 / y1 < 0: absent).
__device__ __forceinline__ double gc_dmax(const EriArgs& g, int x0, int x1, int y0, int y1) {
    const size_t ns = g.nsh;
    double m = g.dsh[x0 * ns + y0];
    if (y1 >= 0) m = fmax(m, g.dsh[x0 * ns + y1]);
    if (x1 >= 0) {
        m = fmax(m, g.dsh[x1 * ns + y0]);
        if (y1 >= 0) m = fmax(m, g.dsh[x1 * ns + y1]);
    }
    return m;
}

// With density screening (dtol > 0) an item is skipped when the group
// bound (largest member Q of each side times the largest density block over
// the groups' shells) is below dtol, and a member quartet is digested only
// when its own bound Q_ab Q_cd max|D| (six shell blocks) reaches dtol, the
// test the shell-pair lists apply; digested member quartets are counted in
// kept.  BT: the block copies the coarse Boys rows of its L into LDS (in
// front of the rows) and reads them there instead of the aligned global
// table, whose row gathers miss the L1 on most primitive iterations.
// TS: each member quartet's integrals are written to the ERI tensor g.out
// (TensorCb, layout g.tpack) instead of digested; no LDS rows, no J/K.
template <int LA, int LC, bool BT, bool TS = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(LA == 1 && LC == 0 ? 3 : 1)))
k_eri_gc(EriArgs g) {
    constexpr int NA = ncart(LA), NC = ncart(LC), NI = NA * NC;
    constexpr int MB = LA == 0 ? 4 : 2, MK = LC == 0 ? 4 : 2;
    // factored member weights for (ss|ss) (measured faster there, no change
    // for the other classes, whose bra sums cost registers)
    constexpr bool GC_FACTOR = LA == 0 && LC == 0;
    extern __shared__ double lds_all[];
    const double* btab = g.boysA;
    double* seg_lds = lds_all;
    if (BT) {
        const double* src = g.boys + BOYS_COFF + (size_t)(LA + LC) * BOYS_CROWS;
        for (int i = threadIdx.x; i < BOYS_CROWS; i += blockDim.x) lds_all[i] = src[i];
        btab = lds_all;
        seg_lds = lds_all + BOYS_CROWS;      // the first zeroing pass below is followed by a barrier
    }
    const int n = g.nbf;
    double* Kr = seg_lds;                        // rows fa0.. fa1.. (NA each), fb0, fb1
    double* Jl = seg_lds + (size_t)(2 * NA + 2) * n;
    const int total = (2 * NA + 2) * n + MB * NA;
    const size_t off = (size_t)xcd_slot(g.jk_copies) * n * n;
    double* J = g.J + off;
    double* K = g.K + off;
    const int lane = threadIdx.x & 63;
    const long long fs = g.kfs;
    for (int sg = blockIdx.x; sg < g.nseg; sg += gridDim.x) {
        const SegDesc seg = g.segs[sg];
        const int bp = g.bra_ids[seg.row];
        const GcMeta mb = g.gmeta[bp];
        if (g.dtol > 0.0 && g.segq &&
            fmax(fmax(mb.q[0], mb.q[1]), fmax(mb.q[2], mb.q[3])) * g.segq[sg] * g.dmax[0] < g.dtol)
            continue;   // block-uniform: no item of the segment survives
        if constexpr (!TS)
            for (int i = threadIdx.x; i < total; i += blockDim.x) seg_lds[i] = 0.0;
        __syncthreads();
        const GcPrim* bra = g.gprims + mb.prim_off;
        double jab[MB * NA];
#pragma unroll
        for (int i = 0; i < MB * NA; ++i) jab[i] = 0.0;
        const bool screen = g.dtol > 0.0;
        unsigned nkept = 0;
        for (long long t0 = seg.t0; t0 < seg.t0 + seg.n; t0 += blockDim.x) {
            const long long t = t0 + threadIdx.x;
            if (t >= seg.t0 + seg.n) continue;
            int pos, b;
            decode_in_row(g, seg.row, t, pos, b);
            const int kp = g.ket_ids[pos];
            const GcMeta mk = g.gmeta[kp];
            if (screen) {
                double qb = 0.0, qk = 0.0;
#pragma unroll
                for (int i = 0; i < 4; ++i) { qb = fmax(qb, mb.q[i]); qk = fmax(qk, mk.q[i]); }
                double m = gc_dmax(g, mk.sa0, mk.sa1, mk.sb0, mk.sb1);
                m = fmax(m, gc_dmax(g, mb.sa0, mb.sa1, mb.sb0, mb.sb1));
                m = fmax(m, gc_dmax(g, mb.sb0, mb.sb1, mk.sb0, mk.sb1));
                m = fmax(m, gc_dmax(g, mb.sb0, mb.sb1, mk.sa0, mk.sa1));
                m = fmax(m, gc_dmax(g, mb.sa0, mb.sa1, mk.sb0, mk.sb1));
                m = fmax(m, gc_dmax(g, mb.sa0, mb.sa1, mk.sa0, mk.sa1));
                if (qb * qk * m < g.dtol) continue;
            }
            const int nk = g.bucket_start[b + 1] - g.bucket_start[b];
            const double* kb = g.ksoa + g.bucket_soa[b] + (pos - g.bucket_start[b]);
            double acc[MB * MK * NI];
#pragma unroll
            for (int i = 0; i < MB * MK * NI; ++i) acc[i] = 0.0;
            // ket primitive outer (its fields stay in registers), bra inner
            // (the same for the whole block: scalar loads)
            for (int q = 0; q < mk.nprim; ++q) {
                const double* x = kb + (long long)q * nk;
                PrimPair kq;
                kq.zeta = x[0]; kq.K = x[fs]; kq.Px = x[2 * fs]; kq.Py = x[3 * fs]; kq.Pz = x[4 * fs];
                kq.PAx = x[5 * fs]; kq.PAy = x[6 * fs]; kq.PAz = x[7 * fs];
                if (GC_FACTOR) {
                    // sum over the bra primitives with the bra members'
                    // weights, then apply the ket members' weights once per
                    // ket primitive
                    double sb[MB * NI];
#pragma unroll
                    for (int i = 0; i < MB * NI; ++i) sb[i] = 0.0;
                    for (int p = 0; p < mb.nprim; ++p) {
                        const GcPrim bq = bra[p];
                        double tv[NI];
#pragma unroll
                        for (int i = 0; i < NI; ++i) tv[i] = 0.0;
                        prim_vrr<double, LA, 0, LC, 0, BT>(bq.pp, kq, btab, tv);
#pragma unroll
                        for (int i = 0; i < MB; ++i)
#pragma unroll
                            for (int e = 0; e < NI; ++e) sb[i * NI + e] += bq.w[i] * tv[e];
                    }
#pragma unroll
                    for (int j = 0; j < MK; ++j) {
                        const double wk = x[(8 + j) * fs];
#pragma unroll
                        for (int i = 0; i < MB; ++i)
#pragma unroll
                            for (int e = 0; e < NI; ++e) acc[(i * MK + j) * NI + e] += wk * sb[i * NI + e];
                    }
                } else {
                    double wk[MK];
#pragma unroll
                    for (int j = 0; j < MK; ++j) wk[j] = x[(8 + j) * fs];
                    for (int p = 0; p < mb.nprim; ++p) {
                        const GcPrim bq = bra[p];
                        double tv[NI];
#pragma unroll
                        for (int i = 0; i < NI; ++i) tv[i] = 0.0;
                        prim_vrr<double, LA, 0, LC, 0, BT>(bq.pp, kq, btab, tv);
                        double u[MB * NI];
#pragma unroll
                        for (int i = 0; i < MB; ++i)
#pragma unroll
                            for (int e = 0; e < NI; ++e) u[i * NI + e] = bq.w[i] * tv[e];
#pragma unroll
                        for (int j = 0; j < MK; ++j)
#pragma unroll
                            for (int i = 0; i < MB; ++i)
#pragma unroll
                                for (int e = 0; e < NI; ++e) acc[(i * MK + j) * NI + e] += wk[j] * u[i * NI + e];
                    }
                }
            }
            const double sq = bp == kp ? 0.5 : 1.0;
            // ket member outer: its J_cd block is summed over the bra members
            // in registers and added once
#pragma unroll
            for (int j = 0; j < MK; ++j) {
                if (j >= mk.m) break;
                const int kc = (mk.code >> (2 * j)) & 1, kd = (mk.code >> (2 * j + 1)) & 1;
                const int fc = kc ? mk.fa1 : mk.fa0, fd = kd ? mk.fb1 : mk.fb0;
                double jcd[NC];
#pragma unroll
                for (int c = 0; c < NC; ++c) jcd[c] = 0.0;
                bool any = false;
#pragma unroll
                for (int i = 0; i < MB; ++i) {
                    if (i >= mb.m) break;
                    if (mb.q[i] * mk.q[j] < g.qtol) continue;
                    const int ia = (mb.code >> (2 * i)) & 1, ib = (mb.code >> (2 * i + 1)) & 1;
                    const int fa = ia ? mb.fa1 : mb.fa0, fb = ib ? mb.fb1 : mb.fb0;
                    if (screen) {
                        const size_t ns = g.nsh;
                        const int a = ia ? mb.sa1 : mb.sa0, bb = ib ? mb.sb1 : mb.sb0;
                        const int c = kc ? mk.sa1 : mk.sa0, d = kd ? mk.sb1 : mk.sb0;
                        double m = fmax(g.dsh[c * ns + d], g.dsh[a * ns + bb]);
                        m = fmax(m, fmax(g.dsh[bb * ns + d], g.dsh[bb * ns + c]));
                        m = fmax(m, fmax(g.dsh[a * ns + d], g.dsh[a * ns + c]));
                        if (mb.q[i] * mk.q[j] * m < g.dtol) continue;
                        nkept += bp != kp || j <= i;   // both orders of a bra == ket row are one quartet
                    }
                    double v[NI];
                    StoreCb<1, NC, 1, double> st{v};
                    ClassCode<LA, 0, LC, 0, double>::hrr(0.0, 0.0, 0.0, 0.0, 0.0, 0.0, acc + (i * MK + j) * NI, st);
                    if constexpr (TS) {
                        TensorCb tc{g.out, (size_t)n, fa, fb, fc, fd, g.tpack};
#pragma unroll
                        for (int a = 0; a < NA; ++a)
#pragma unroll
                            for (int c = 0; c < NC; ++c) tc(a, 0, c, 0, v[a * NC + c]);
                        continue;
                    }
                    RegDigest<NA, 1, NC, 1, double> cb;
                    cb.load(g.D, n, fa, fb, fc, fd);
#pragma unroll
                    for (int a = 0; a < NA; ++a)
#pragma unroll
                        for (int c = 0; c < NC; ++c) cb(a, 0, c, 0, v[a * NC + c]);
                    const double s = (fa == fb ? 0.5 : 1.0) * (fc == fd ? 0.5 : 1.0) * sq;
#pragma unroll
                    for (int a = 0; a < NA; ++a) {
                        jab[i * NA + a] += s * cb.Jab[a];
                        double* row = Kr + (size_t)(ia * NA + a) * n;
#pragma unroll
                        for (int c = 0; c < NC; ++c) atomicAdd(&row[fc + c], s * cb.Kac[a * NC + c]);
                        atomicAdd(&row[fd], s * cb.Kad[a]);
                    }
                    double* rowb = Kr + (size_t)(2 * NA + ib) * n;
#pragma unroll
                    for (int c = 0; c < NC; ++c) atomicAdd(&rowb[fc + c], s * cb.Kbc[c]);
                    atomicAdd(&rowb[fd], s * cb.Kbd[0]);
#pragma unroll
                    for (int c = 0; c < NC; ++c) jcd[c] += s * cb.Jcd[c];
                    any = true;
                }
                if (any) {
#pragma unroll
                    for (int c = 0; c < NC; ++c) jk_add(&J[(size_t)(fc + c) * n + fd], jcd[c], g.jk_copies);
                }
            }
        }
        if constexpr (!TS) {
#pragma unroll
            for (int i = 0; i < MB * NA; ++i) {
                const double w = wave_sum(jab[i]);
                if (lane == 0 && i < mb.m * NA) atomicAdd(&Jl[i], w);
            }
        }
        if (screen) {
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) nkept += __shfl_xor(nkept, o, 64);
            if (lane == 0 && nkept) atomicAdd(g.kept, nkept);
        }
        __syncthreads();
        if constexpr (!TS) {
            for (int i = threadIdx.x; i < (2 * NA + 2) * n; i += blockDim.x) {
                const double x = Kr[i];
                if (x != 0.0) {
                    const int r = i / n, col = i - r * n;
                    const int f =
                        r < 2 * NA ? (r < NA ? mb.fa0 + r : mb.fa1 + (r - NA)) : (r == 2 * NA ? mb.fb0 : mb.fb1);
                    jk_add(&K[(size_t)f * n + col], x, g.jk_copies);
                }
            }
            for (int i = threadIdx.x; i < mb.m * NA; i += blockDim.x) {
                const int j = i / NA, a = i - j * NA;
                const int fa = ((mb.code >> (2 * j)) & 1) ? mb.fa1 : mb.fa0;
                const int fb = ((mb.code >> (2 * j + 1)) & 1) ? mb.fb1 : mb.fb0;
                jk_add(&J[(size_t)(fa + a) * n + fb], Jl[i], g.jk_copies);
            }
        }
        __syncthreads();
    }
}

// Segment kernel for classes (ab|c s) whose bra is a shell pair and whose
// kets are generally contracted pairs ((pp|ss) (pp|ps) (ds|ps) (dp|ss)):
// bra and ket classes differ, so no bra-ket symmetry is involved.  For each
// ket primitive the VRR sums over the bra primitives (the block's shared
// bra), and that block is added to every ket member with its weight; after
// the loops the bra HRR runs per ket member, whose quartet is digested as in
// k_eri_seg when it passes Q_ab q_cd >= qtol (and, screened, the density
// test of the shell-pair lists).  LDS as k_eri_seg.
// One member quartet of k_eri_gck: bra HRR of its [e0|c0] block, then the
// segment kernel's digestion (exchange rows of the bra shells in LDS, J_ab
// in jab, J_cd to the global copy).
template <int LA, int LB, int LC>
__device__ __forceinline__ void gck_digest(const EriArgs& g, const double* acc, double ABx, double ABy, double ABz,
                                           int fa, int fb, int fc, int fd, double* Kr, double* J, double* jab) {
    constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC);
    const int n = g.nbf;
    double v[NA * NB * NC];
    StoreCb<NB, NC, 1, double> st{v};
    ClassCode<LA, LB, LC, 0, double>::hrr(ABx, ABy, ABz, 0.0, 0.0, 0.0, acc, st);
    RegDigest<NA, NB, NC, 1, double> cb;
    cb.load(g.D, n, fa, fb, fc, fd);
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int bb = 0; bb < NB; ++bb)
#pragma unroll
            for (int c = 0; c < NC; ++c) cb(a, bb, c, 0, v[(a * NB + bb) * NC + c]);
    const double s = (fa == fb ? 0.5 : 1.0) * (fc == fd ? 0.5 : 1.0);
#pragma unroll
    for (int i = 0; i < NA * NB; ++i) jab[i] += s * cb.Jab[i];
#pragma unroll
    for (int a = 0; a < NA; ++a) {
#pragma unroll
        for (int c = 0; c < NC; ++c) atomicAdd(&Kr[(size_t)a * n + fc + c], s * cb.Kac[a * NC + c]);
        atomicAdd(&Kr[(size_t)a * n + fd], s * cb.Kad[a]);
    }
#pragma unroll
    for (int bb = 0; bb < NB; ++bb) {
#pragma unroll
        for (int c = 0; c < NC; ++c) atomicAdd(&Kr[(size_t)(NA + bb) * n + fc + c], s * cb.Kbc[bb * NC + c]);
        atomicAdd(&Kr[(size_t)(NA + bb) * n + fd], s * cb.Kbd[bb]);
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) jk_add(&J[(size_t)(fc + c) * n + fd], s * cb.Jcd[c], g.jk_copies);
}

// BT: the coarse Boys rows in LDS, as in k_eri_gc.
template <int LA, int LB, int LC, bool BT>
__global__ void __launch_bounds__(256) k_eri_gck(EriArgs g) {
    constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC);
    constexpr int NE = ((LA + LB + 1) * (LA + LB + 2) * (LA + LB + 3) - LA * (LA + 1) * (LA + 2)) / 6;
    constexpr int NF = ncart(LC);
    constexpr int MK = LC == 0 ? 4 : 2;
    extern __shared__ double lds_all[];
    const double* btab = g.boysA;
    double* seg_lds = lds_all;
    if (BT) {
        const double* src = g.boys + BOYS_COFF + (size_t)(LA + LB + LC) * BOYS_CROWS;
        for (int i = threadIdx.x; i < BOYS_CROWS; i += blockDim.x) lds_all[i] = src[i];
        btab = lds_all;
        seg_lds = lds_all + BOYS_CROWS;      // the first zeroing pass below is followed by a barrier
    }
    const int n = g.nbf;
    double* Kr = seg_lds;                        // rows fa.. (NA) then fb.. (NB), width n
    double* Jl = seg_lds + (size_t)(NA + NB) * n;
    const int total = (NA + NB) * n + NA * NB;
    const size_t off = (size_t)xcd_slot(g.jk_copies) * n * n;
    double* J = g.J + off;
    double* K = g.K + off;
    const int lane = threadIdx.x & 63;
    const long long fs = g.kfs;
    for (int sg = blockIdx.x; sg < g.nseg; sg += gridDim.x) {
        const SegDesc seg = g.segs[sg];
        const int bp = g.bra_ids[seg.row];
        if (g.dtol > 0.0 && g.segq && g.Qs[bp] * g.segq[sg] * g.dmax[0] < g.dtol) continue;   // block-uniform
        for (int i = threadIdx.x; i < total; i += blockDim.x) seg_lds[i] = 0.0;
        __syncthreads();
        const PairMeta mb = g.meta[bp];
        const int fa = mb.fa, fb = mb.fb;
        const double ABx = g.ab[4 * bp], ABy = g.ab[4 * bp + 1], ABz = g.ab[4 * bp + 2];
        const PrimPair* bra = g.prims + mb.prim_off;
        const double qb = g.Qs[bp];
        const bool screen = g.dtol > 0.0;
        double jab[NA * NB];
#pragma unroll
        for (int i = 0; i < NA * NB; ++i) jab[i] = 0.0;
        unsigned nkept = 0;
        for (long long t0 = seg.t0; t0 < seg.t0 + seg.n; t0 += blockDim.x) {
            const long long t = t0 + threadIdx.x;
            if (t >= seg.t0 + seg.n) continue;
            int pos, b;
            decode_in_row(g, seg.row, t, pos, b);
            const int kp = g.ket_ids[pos];
            const GcMeta mk = g.gmeta[kp];
            const int2 ab = screen ? g.psh[bp] : make_int2(0, 0);
            if (screen) {
                double qk = 0.0;
#pragma unroll
                for (int i = 0; i < 4; ++i) qk = fmax(qk, mk.q[i]);
                double m = gc_dmax(g, mk.sa0, mk.sa1, mk.sb0, mk.sb1);
                m = fmax(m, g.dsh[(size_t)ab.x * g.nsh + ab.y]);
                m = fmax(m, gc_dmax(g, ab.y, -1, mk.sb0, mk.sb1));
                m = fmax(m, gc_dmax(g, ab.y, -1, mk.sa0, mk.sa1));
                m = fmax(m, gc_dmax(g, ab.x, -1, mk.sb0, mk.sb1));
                m = fmax(m, gc_dmax(g, ab.x, -1, mk.sa0, mk.sa1));
                if (qb * qk * m < g.dtol) continue;
            }
            const int nk = g.bucket_start[b + 1] - g.bucket_start[b];
            const double* kb = g.ksoa + g.bucket_soa[b] + (pos - g.bucket_start[b]);
            double acc[MK * NE * NF];
#pragma unroll
            for (int i = 0; i < MK * NE * NF; ++i) acc[i] = 0.0;
            for (int q = 0; q < mk.nprim; ++q) {
                const double* x = kb + (long long)q * nk;
                PrimPair kq;
                kq.zeta = x[0]; kq.K = x[fs]; kq.Px = x[2 * fs]; kq.Py = x[3 * fs]; kq.Pz = x[4 * fs];
                kq.PAx = x[5 * fs]; kq.PAy = x[6 * fs]; kq.PAz = x[7 * fs];
                double sb[NE * NF];
#pragma unroll
                for (int i = 0; i < NE * NF; ++i) sb[i] = 0.0;
                for (int p = 0; p < mb.nprim; ++p) prim_vrr<double, LA, LB, LC, 0, BT>(bra[p], kq, btab, sb);
#pragma unroll
                for (int j = 0; j < MK; ++j) {
                    const double wk = x[(8 + j) * fs];
#pragma unroll
                    for (int e = 0; e < NE * NF; ++e) acc[j * NE * NF + e] += wk * sb[e];
                }
            }
            // one digestion instance: member j is always in acc[0 ..), the
            // accumulators move down one member after each step
#pragma unroll 1
            for (int j = 0; j < mk.m; ++j) {
                const int kc = (mk.code >> (2 * j)) & 1, kd = (mk.code >> (2 * j + 1)) & 1;
                const int fc = kc ? mk.fa1 : mk.fa0, fd = kd ? mk.fb1 : mk.fb0;
                bool live = qb * mk.q[j] >= g.qtol;
                if (screen && live) {
                    const size_t ns = g.nsh;
                    const int c = kc ? mk.sa1 : mk.sa0, d = kd ? mk.sb1 : mk.sb0;
                    double m = fmax(g.dsh[c * ns + d], g.dsh[ab.x * ns + ab.y]);
                    m = fmax(m, fmax(g.dsh[ab.y * ns + d], g.dsh[ab.y * ns + c]));
                    m = fmax(m, fmax(g.dsh[ab.x * ns + d], g.dsh[ab.x * ns + c]));
                    live = qb * mk.q[j] * m >= g.dtol;
                    nkept += live;
                }
                if (live) gck_digest<LA, LB, LC>(g, acc, ABx, ABy, ABz, fa, fb, fc, fd, Kr, J, jab);
#pragma unroll
                for (int jj = 0; jj + 1 < MK; ++jj)
#pragma unroll
                    for (int e = 0; e < NE * NF; ++e) acc[jj * NE * NF + e] = acc[(jj + 1) * NE * NF + e];
            }
        }
#pragma unroll
        for (int i = 0; i < NA * NB; ++i) {
            const double w = wave_sum(jab[i]);
            if (lane == 0) atomicAdd(&Jl[i], w);
        }
        if (screen) {
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) nkept += __shfl_xor(nkept, o, 64);
            if (lane == 0 && nkept) atomicAdd(g.kept, nkept);
        }
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

typedef void (*EriKernel)(EriArgs);
// Getter modes of polyeri_class_kernel beyond the k_eri modes: the segment
// kernel for MODE_JK and MODE_JK32.
enum { GET_SEG_JK = 5, GET_SEG_JK32 = 6 };
// Strided batched FP64 products on the MFMA units (kgemm.hip):
// C_b(m,n) = sum_k A_b(m,k) B_b(k,n), X_b(r,c) = X[b*sX + r*sXr + c*sXc].
enum { GEMM_F64 = 1, GEMM_BF16X3 = 2 };
int polyeri_gemm_launch(int mode, int M, int N, int K, const double* A, long sam, long sak, long sA, const double* B,
                        long sbk, long sbn, long sB, double* C, long scm, long scn, long sC, int batch,
                        hipStream_t stream);
// Planar complex products C_b = A_b B_b (A real when Ai is null), same
// strides for the real and imaginary planes (kgemm.hip).  nci > 0 splits
// the columns of B and C as (n / nci, n % nci) with outer strides sbo / sco
// (and inner strides sbn / scn).
int polyeri_zgemm_launch(int mode, int M, int N, int K, const double* Ar, const double* Ai, long sam, long sak,
                         long sA, const double* Br, const double* Bi, long sbk, long sbn, long sB, double* Cr,
                         double* Ci, long scm, long scn, long sC, int batch, hipStream_t stream, int nci = 0,
                         long sbo = 0, long sco = 0);
// Density screening (kscreen.hip): shell-block maxima of |D|, and the
// compaction of one work list to the items whose bound
// Q_ab Q_cd max(|D| over the six blocks the quartet reads) reaches tol.
int polyeri_shell_dmax(const double* D, int nbf, const int* shell_f0, const int* shell_nf, int nsh,
                       double* dsh, hipStream_t stream);
// *out = the largest of the n values x (one block).
int polyeri_max_reduce(const double* x, size_t n, double* out, hipStream_t stream);
int polyeri_screen_list(const EriArgs& a, const double* Q, const int2* pair_shells, const double* dsh,
                        int nsh, double tol, int2* items, unsigned* count, hipStream_t stream);
// Cooperative (one wavefront per quartet, LDS-resident recursion) kernels;
// primitive quartets are prepared in batches of 2^pbs (pbs 0..3).
int polyeri_coop_nslot(int la, int lb, int lc, int ld, int pbs);
int polyeri_coop_launch(const EriArgs& a, int mode, int la, int lb, int lc, int ld, int pbs, int max_grid,
                        hipStream_t stream);
// default batch: eight for the classes whose LDS is small enough that the
// records cost no occupancy, one otherwise
inline int polyeri_coop_default_pbs(int nint) { return nint < 200 ? 3 : 0; }
// Generally contracted segment kernel (kgc.hip): LDS bytes, whether the
// class (la s|lc s) has one, and the launch over a.segs (grid blocks).
size_t polyeri_gc_lds_bytes(int la, int nbf);
bool polyeri_gc_class(int la, int lc);
int polyeri_gc_launch(const EriArgs& a, int la, int lc, int block, int grid, hipStream_t stream, bool tensor = false);
// The same for k_eri_gck (shell-pair bra (la lb|, generally contracted kets |lc s)).
size_t polyeri_gck_lds_bytes(int la, int lb, int nbf);
bool polyeri_gck_class(int la, int lb, int lc);
int polyeri_gck_launch(const EriArgs& a, int la, int lb, int lc, int block, int grid, hipStream_t stream);
// Rys-quadrature segment kernels (krys.hip) for the classes (dp|pp) (dp|dp)
// (dd|pp) (dd|ds) (dd|dp) (dd|dd): J/K over bra-row segments, mode MODE_JK
// or MODE_JK32 (FP32 recursion).
bool polyeri_rys_class(int la, int lb, int lc, int ld);
size_t polyeri_rys_lds_bytes(int la, int lb, int nbf);
int polyeri_rys_launch(const EriArgs& a, int mode, int la, int lb, int lc, int ld, int block, int max_grid,
                       hipStream_t stream);
// Nodes u[0..n-1] and weights w[0..n-1] of the n-point Rys rule at T, n <= 5
// (host evaluation of the kernels' tables).
int polyeri_rys_nodes_host(int n, double T, double* u, double* w);
// One primitive quartet's Cartesian block through the kernels' per-lane Rys
// code on the host (0; -1 bad class or pointer; -2 an integral not produced
// exactly once).
int polyeri_rys_block_host(int la, int lb, int lc, int ld, const double* xyz, const double* ex, double* out);

// Two-electron part of the nuclear gradient (kgrad.hip): shells of the
// library's basis (L, atom, first function, first primitive, primitive
// count, centre), primitive exponents and coefficients, the density in the
// library's basis (nbf x nbf) and its largest |element| per shell block
// (nsh x nsh), the per-pair Schwarz bounds (filled by the launcher), the
// threshold on Q_ab Q_cd and on Q_ab Q_cd max |Gamma|, the exchange scale.
// Shell pairs (pa[i], pb[i]) with L(pa) >= L(pb), grouped by pair class
// ((ss) (ps) (pp) (ds) (dp) (dd) = 0 .. 5: pairs seg[X] .. seg[X + 1] - 1),
// by primitive count inside a class; the quartets of class pair X >= Y run
// as one launch of the kernel for that class.
struct GradArgs {
    int nsh, natom, nbf, npair;
    int seg[7];
    const int *L, *atom, *f0, *p0, *np;
    const int *pa, *pb;
    const double *xyz, *alpha, *coef;
    const double* D;
    const double* dsh;
    const double* Qp;
    double tol, xfac, ptol;
    // the Schwarz survivors of one class pair as chunks: chunk c holds the
    // quartets coff[c] .. coff[c + 1] - 1, bra pair x0 + cij[2c] and ket
    // pairs y0 + cij[2c + 1] + (t - coff[c]); null: the full product
    const unsigned long long* coff = nullptr;
    const int* cij = nullptr;
    int nchunk = 0;
};

// Schwarz-surviving quartets of every class pair (X, Y <= X) of pairs sorted
// by polyeri_grad_sort_pairs: inside a class, pairs of one primitive count
// (a group) are in descending bound order, so the kets of a group that
// survive with a given bra are a prefix of the group.  Class pair k = 6X + Y
// has n[k] chunks, offsets off[off0[k]] .. off[off0[k] + n[k]] and
// (row, first column) pairs ij[ij0[k] ..].
struct GradChunks {
    std::vector<unsigned long long> off;
    std::vector<int> ij;
    size_t off0[36], ij0[36];
    int n[36];
    unsigned long long nq[36];
};
// grad (3 natom, zeroed by the caller) += dE2/dR on the device; d_Qp: npair
// doubles, the pair bounds when have_qp, else computed here.  ch (optional,
// with d_off / d_ij its off / ij arrays on the device): launch only the
// Schwarz survivors it lists.
int polyeri_grad_launch(const GradArgs& g, double* d_Qp, double* d_grad, hipStream_t stream, bool have_qp = false,
                        const GradChunks* ch = nullptr, const unsigned long long* d_off = nullptr,
                        const int* d_ij = nullptr);
// Orders the pairs of every class segment by primitive count (descending,
// as they come) and, inside one count, by bound (descending); pa, pb and qp
// permuted together.  False (nothing changed) when a bound is not finite.
bool polyeri_grad_sort_pairs(int* pa, int* pb, double* qp, const int* np, const int* seg);
// The chunks of sorted pairs whose bound product is >= tol.
void polyeri_grad_chunks(const int* pa, const int* pb, const int* np, const double* qp, const int* seg, double tol,
                         GradChunks& ch);
// the same sum on the host (host pointers), grad overwritten
int polyeri_grad_host(const GradArgs& g, double* grad);
// One record of the reference's .ao files (kstored.hip): packed cells,
// packed 1-based indices I J K L, the integral.
struct StoredEri {
    int32_t cells, ijkl;
    double v;
};
// J/K blocks (reference layout, 2 cel1 + 1 cells or 1) from n records.
int polyeri_jk_stored_launch(const StoredEri* rec, long long n, int nc, int cel1, int cel2, const double* P, double* J,
                             double* K, hipStream_t s);
// Internal-function AO tensor of a molecule from n records.
int polyeri_tensor_stored_launch(const StoredEri* rec, long long n, const int* inv, const double* scale, int nbf,
                                 double* T, hipStream_t s);
// Classes with more than this many integrals per quartet run cooperatively.
#define POLYERI_COOP_MIN_INTEGRALS 108
// Kernel of class (la lb|lc ld) and mode, or nullptr (generated getters).
EriKernel polyeri_class_kernel(int la, int lb, int lc, int ld, int mode);
