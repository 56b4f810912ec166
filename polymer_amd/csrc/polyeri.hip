// polyeri.hip -- libpolyeri: MI355X (gfx950) direct two-electron path.
//
// Device side: one kernel instantiation per shell-quartet class
// (la lb|lc ld), one thread per contracted shell quartet, primitive loop,
// Boys + VRR + HRR from eri_engine.h, then the quartet's contribution to the
// Coulomb and exchange matrices with FP64 atomics (8-fold permutational
// symmetry, so each unique quartet is computed once).
// Host side: context set-up, class-sorted Schwarz-screened work lists, the
// J/K build that replaces FOURINDEX_ERI, the SCF loop and MP2.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <array>
#include <mutex>
#include <string>
#include <vector>
#include <unordered_map>

#include "../../include/polyeri.h"
#include "polyeri_internal.h"

using namespace polyeri;

static thread_local std::string g_err;

const char* polyeri_last_error(void) { return g_err.c_str(); }

static int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

int polyeri_record_error(int code, const char* msg) { return fail(code, msg ? msg : ""); }

#define HIPCHK(x)                                                                        \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess)                                                            \
            return fail(e_ == hipErrorOutOfMemory ? POLYERI_E_NOMEM : POLYERI_E_HIP,     \
                        std::string(#x) + ": " + hipGetErrorString(e_));                 \
    } while (0)

#define RBCHK(x)                                                                         \
    do {                                                                                 \
        int s_ = (int)(x);                                                               \
        if (s_ != 0) return fail(POLYERI_E_LIB, std::string(#x) + " failed, status " + std::to_string(s_)); \
    } while (0)

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------
#include "eri_kernels.h"
#include "kcc.h"

__global__ void k_fill(double* p, size_t n, double v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// out = x + a y
__global__ void k_axpby(const double* __restrict__ x, const double* __restrict__ y, double a, size_t n,
                        double* __restrict__ out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = x[i] + a * y[i];
}

// D_int[mu][nu] = s_mu s_nu P_C(ref mu, ref nu)
__global__ void k_pack_density(const double* __restrict__ pc, int nc, const int* __restrict__ ref,
                               const double* __restrict__ sc, int nbf, double* __restrict__ D) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)nbf * nbf) return;
    int mu = (int)(t / nbf), nu = (int)(t % nbf);
    D[t] = sc[mu] * sc[nu] * pc[ref[mu] + (size_t)nc * ref[nu]];
}

// C_C(ref mu, ref nu) = s s 2(J + J^T), X_C = s s (K + K^T), summed over the
// per-XCD copies
__global__ void k_unpack_jk(const double* __restrict__ J, const double* __restrict__ K, int nbf, int copies,
                            const int* __restrict__ ref, const double* __restrict__ sc, int nc,
                            double* __restrict__ jc, double* __restrict__ kc) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)nbf * nbf) return;
    int mu = (int)(t / nbf), nu = (int)(t % nbf);
    double f = sc[mu] * sc[nu];
    size_t r = ref[mu] + (size_t)nc * ref[nu];
    const size_t n2 = (size_t)nbf * nbf;
    double js = 0.0, ks = 0.0;
    for (int x = 0; x < copies; ++x) {
        js += J[x * n2 + (size_t)mu * nbf + nu] + J[x * n2 + (size_t)nu * nbf + mu];
        ks += K[x * n2 + (size_t)mu * nbf + nu] + K[x * n2 + (size_t)nu * nbf + mu];
    }
    jc[r] = f * 2.0 * js;
    kc[r] = f * ks;
}

// F = T + V + J - hfx/2 K, then the Hermitian average (src/pfock.f90:44, :224-233)
__global__ void k_fock(const double* __restrict__ T, const double* __restrict__ V,
                       const double* __restrict__ J, const double* __restrict__ K, double hfx, int n,
                       double* __restrict__ F) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n * n) return;
    int i = (int)(t % n), j = (int)(t / n);
    size_t ij = t, ji = (size_t)i * n + j;
    double a = T[ij] + V[ij] + J[ij] - 0.5 * hfx * K[ij];
    double b = T[ji] + V[ji] + J[ji] - 0.5 * hfx * K[ji];
    F[ij] = 0.5 * (a + b);
}

// X(:, k) = U(:, keep[k]) / sqrt(e[keep[k]])
__global__ void k_orth(const double* __restrict__ U, const double* __restrict__ e,
                       const int* __restrict__ keep, int n, int nk, double* __restrict__ X) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n * nk) return;
    int r = (int)(t % n), k = (int)(t / n);
    X[t] = U[r + (size_t)n * keep[k]] / sqrt(e[keep[k]]);
}

// W(:, l) = occ[l] * C(:, l)
__global__ void k_scale_cols(const double* __restrict__ C, const double* __restrict__ occ, int n,
                             int m, double* __restrict__ W) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n * m) return;
    W[t] = C[t] * occ[t / n];
}

// Cint[mu, i] = s_mu CO[ref mu, i]
__global__ void k_mo_int(const double* __restrict__ co, int nc, const int* __restrict__ ref,
                         const double* __restrict__ sc, int nbf, int nmo, double* __restrict__ C) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)nbf * nmo) return;
    int mu = (int)(t % nbf), i = (int)(t / nbf);
    C[t] = sc[mu] * co[ref[mu] + (size_t)nc * i];
}

// Plain streaming copy (16 B per lane per step): the reference point that
// scales the FETCH_SIZE / WRITE_SIZE counters to bytes.
__global__ void k_copy(const double2* __restrict__ src, double2* __restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// E(2) singlet and triplet (src/pmp2energy.f90:395-433 at one k-point) from
// T[b][a'][i'][j'] = (ia|jb), j' fastest: j' and i' over the "occupied"
// range icore.. (nop wide), a' over the "virtual" range starting aoff below
// iocc (nvp wide), b over the nv virtuals; the energy uses i, j < no and the
// true virtuals a' = aoff + a.
// One block per virtual a and a stride of virtuals b: the no x no block of
// (ia|jb) of each pair (a, b) is read once, coalesced, into LDS (TILE; else
// straight from global memory) so that (ja|ib) is its transpose.
template <bool TILE>
__global__ void __launch_bounds__(256) k_mp2_energy(const double* __restrict__ T, const double* __restrict__ eps,
                                                    int no, int nv, int nop, int nvp, int aoff, int icore, int iocc,
                                                    double* __restrict__ acc) {
    extern __shared__ double tile[];                     // no x (no + 1)
    const int a = blockIdx.x, ld = no + 1, nn = no * no;
    double es = 0.0, et = 0.0;
    for (int b = blockIdx.y; b < nv; b += gridDim.y) {
        const double* __restrict__ Tab = T + (size_t)nop * nop * ((size_t)aoff + a + (size_t)nvp * b);
        const double eab = eps[iocc + a] + eps[iocc + b];
        if (TILE) {
            __syncthreads();                                 // the previous pair's reads are done
            for (int e = threadIdx.x; e < nn; e += blockDim.x) {
                const int i = e / no, j = e - i * no;
                tile[i * ld + j] = Tab[j + (size_t)nop * i];
            }
            __syncthreads();
        }
        for (int e = threadIdx.x; e < nn; e += blockDim.x) {
            const int i = e / no, j = e - i * no;
            const double x = TILE ? tile[i * ld + j] : Tab[j + (size_t)nop * i];     // (ia|jb)
            const double xp = TILE ? tile[j * ld + i] : Tab[i + (size_t)nop * j];    // (ja|ib)
            const double den = eps[icore + i] + eps[icore + j] - eab;
            es += (0.5 * x * x + 0.5 * x * xp) / den;
            et += (1.5 * x * x - 1.5 * x * xp) / den;
        }
    }
    __shared__ double ss[256], st[256];
    ss[threadIdx.x] = es;
    st[threadIdx.x] = et;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            ss[threadIdx.x] += ss[threadIdx.x + w];
            st[threadIdx.x] += st[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicAdd(&acc[0], ss[0]);
        atomicAdd(&acc[1], st[0]);
    }
}

// MP2 quasi-particle sums of band g = g0 + blockIdx.y (src/pmp2energy.f90:445-522
// at one k-point), same T layout; g sits in the i' slot (first sum) and in the
// a' slot (second sum):
//   q(g) = sum_{b,j,a} [2 (ga|jb)^2 - (ga|jb)(ja|gb)] / (e_g + e_j - e_a - e_b)
//        + sum_{b,j,i} [2 (ig|jb)^2 - (ig|jb)(jg|ib)] / (e_g + e_b - e_i - e_j)
__global__ void k_mp2_qp(const double* __restrict__ T, const double* __restrict__ eps, int no, int nv, int nop,
                         int nvp, int aoff, int icore, int iocc, int vlo, int g0, double* __restrict__ q) {
    const int g = g0 + blockIdx.y;
    const int gi = g - icore, ga = g - vlo;
    const size_t n1 = (size_t)nv * no * nv, n2 = (size_t)nv * no * no;
    double s = 0.0;
    const size_t nop2 = (size_t)nop * nop;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n1 + n2; t += (size_t)gridDim.x * blockDim.x) {
        if (t < n1) {
            const int j = (int)(t % no);
            const size_t r = t / no;
            const int a = (int)(r % nv), b = (int)(r / nv);
            const size_t base = nop2 * ((size_t)aoff + a + (size_t)nvp * b);
            const double x = T[base + j + (size_t)nop * gi];
            const double xs = T[base + gi + (size_t)nop * j];
            s += (2.0 * x * x - x * xs) / (eps[g] + eps[icore + j] - eps[iocc + a] - eps[iocc + b]);
        } else {
            const size_t u = t - n1;
            const int j = (int)(u % no);
            const size_t r = u / no;
            const int i = (int)(r % no), b = (int)(r / no);
            const size_t base = nop2 * ((size_t)ga + (size_t)nvp * b);
            const double x = T[base + j + (size_t)nop * i];
            const double xs = T[base + i + (size_t)nop * j];
            s += (2.0 * x * x - x * xs) / (eps[g] + eps[iocc + b] - eps[icore + i] - eps[icore + j]);
        }
    }
    __shared__ double sh[256];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(&q[blockIdx.y], sh[0]);
}

// ---------------------------------------------------------------------------
// Launch table
// ---------------------------------------------------------------------------
template <int MODE>
static EriKernel kernel_for(int la, int lb, int lc, int ld) {
    return polyeri_class_kernel(la, lb, lc, ld, MODE);
}

static const int kPairL[NPAIRCLASS][2] = {{0, 0}, {1, 0}, {1, 1}, {2, 0}, {2, 1}, {2, 2}};

static int ncart_h(int l) { return (l + 1) * (l + 2) / 2; }

// Algorithmic FLOPs of one primitive quartet of class (la lb|lc ld): Boys +
// pair assembly, the OS VRR targets (8 flops each), and J/K digestion under
// 8-fold symmetry.  The same model prices the CPU baseline.
static double flop_model(int la, int lb, int lc, int ld) {
    int L = la + lb + lc + ld;
    double f = 25 + 3 * L, v = 0;
    for (int e = 0; e <= la + lb; ++e)
        for (int g = 0; g <= lc + ld; ++g)
            if (e + g >= 1) v += (double)ncart_h(e) * ncart_h(g) * (L - e - g + 1);
    return f + 8 * v + 12.0 * ncart_h(la) * ncart_h(lb) * ncart_h(lc) * ncart_h(ld);
}

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
// How one work list is launched.  kind: LAUNCH_AUTO picks, for classes of
// at most POLYERI_COOP_MIN_INTEGRALS integrals, the segment kernel (bra-row
// segments with LDS-accumulated exchange rows) when its LDS fits and the
// list is not density-screened, else the register kernel, and the
// cooperative kernel above; block: threads per block of the register or
// segment kernel; grid: block cap (0 = one block per 256 items for the
// register kernel, one per segment for the segment kernel, 4096 for the
// cooperative one).  The autotuner fills these per list.
enum { LAUNCH_AUTO = 0, LAUNCH_REG = 1, LAUNCH_COOP = 2, LAUNCH_SEG = 3, LAUNCH_RYS = 5 };
static const int RYS_NMAX_EXPORT = 5;
// Items per bra-row segment of the segment kernel (at most), and the number
// of segments a list's primitive quartets are spread over (at least).
static const int kSegItems = 2048, kSegMin = 256;
static const double kSegTarget = 4096.0;

// Bra-row segments of a work list for the segment kernels: each segment is
// part of one row, at most kSegItems items and about `target` primitive
// quartets (the row's bra primitive count times each ket's), with target =
// the list's primitive quartets / kSegTarget, but at least kSegMin items (one
// pass of a block: shorter segments pay more for their row flush than they
// gain in balance) where the row has them; ordered by
// primitive quartets, largest first, so that the blocks launched last are
// the short ones.  pairs / bra / kets / bstart: the list's pair data, bra
// ids by row, ket ids and bucket starts of the ket class.  balanced = false:
// rows cut every kSegItems items, in row order (the autotuner keeps
// whichever form is faster per list).
static std::vector<SegDesc> make_segments(const WorkList& l, const std::vector<PairInfo>& pairs,
                                          const std::vector<int>& bra, const std::vector<int>& kets,
                                          const std::vector<int>& bstart, bool balanced = true) {
    const int nb = l.nbucket;
    if (!balanced) {
        std::vector<SegDesc> out;
        for (int r = 0; r < l.nrows; ++r)
            for (long long t = l.row[r]; t < l.row[r + 1]; t += kSegItems)
                out.push_back(SegDesc{t, r, (int)std::min<long long>(kSegItems, l.row[r + 1] - t)});
        return out;
    }
    std::vector<double> knp(nb, 0.0);
    for (int b = 0; b + 1 < (int)bstart.size() && b < nb; ++b) knp[b] = pairs[kets[bstart[b]]].nprim;
    const double target = l.prim_quartets / kSegTarget;
    std::vector<std::pair<double, SegDesc>> segs;
    for (int r = 0; r < l.nrows; ++r) {
        const double bp = pairs[bra[r]].nprim;
        const int* rb = l.rb.data() + (size_t)r * (nb + 1);
        long long t = l.row[r];
        int n = 0;
        double cost = 0.0;
        for (int b = 0; b < nb; ++b) {
            int cnt = rb[b + 1] - rb[b];
            const double ci = bp * knp[b];
            while (cnt > 0) {
                int take = std::min(cnt, kSegItems - n);
                if (ci > 0.0)
                    take = (int)std::min<double>(take, std::max<double>(kSegMin - n, std::ceil((target - cost) / ci)));
                take = std::max(take, 1);
                n += take; cost += take * ci; cnt -= take;
                if (n >= kSegItems || (n >= kSegMin && cost >= target)) {
                    segs.push_back({cost, SegDesc{t, r, n}});
                    t += n; n = 0; cost = 0.0;
                }
            }
        }
        if (n > 0) segs.push_back({cost, SegDesc{t, r, n}});
    }
    std::stable_sort(segs.begin(), segs.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
    std::vector<SegDesc> out(segs.size());
    for (size_t i = 0; i < segs.size(); ++i) out[i] = segs[i].second;
    return out;
}
// Largest ket Schwarz bound kq over the items of each segment.  A row's
// items cover one contiguous range of ket positions per bucket (the
// decode_in_row mapping), so a sparse table of range maxima over the ket
// order answers each range.  With a density-screened build a segment whose
// bra bound x this x the largest |D| block is below the threshold holds no
// surviving item and is skipped whole.
static std::vector<double> segment_ket_bounds(const WorkList& l, const std::vector<SegDesc>& segs,
                                              const std::vector<int>& kets, const std::vector<int>& bstart,
                                              const std::vector<double>& kq) {
    const int P = (int)kets.size();
    std::vector<double> out(segs.size(), 0.0);
    if (P == 0) return out;
    std::vector<std::vector<double>> sp(1, std::vector<double>(P));
    for (int i = 0; i < P; ++i) sp[0][i] = kq[kets[i]];
    for (int k = 1; (1 << k) <= P; ++k) {
        sp.emplace_back(P - (1 << k) + 1);
        for (int i = 0; i + (1 << k) <= P; ++i) sp[k][i] = std::max(sp[k - 1][i], sp[k - 1][i + (1 << (k - 1))]);
    }
    const int nb = l.nbucket;
    for (size_t s = 0; s < segs.size(); ++s) {
        const SegDesc& sg = segs[s];
        const int* rb = l.rb.data() + (size_t)sg.row * (nb + 1);
        const long long l0 = sg.t0 - l.row[sg.row], l1 = l0 + sg.n;
        double m = 0.0;
        for (int b = 0; b < nb && b + 1 < (int)bstart.size(); ++b) {
            const long long lo = std::max<long long>(l0, rb[b]), hi = std::min<long long>(l1, rb[b + 1]);
            if (lo >= hi) continue;
            const int first = l.rf.empty() ? 0 : l.rf[(size_t)sg.row * nb + b];
            const int p0 = bstart[b] + first + (int)(lo - rb[b]), p1 = p0 + (int)(hi - lo);
            const int k = 31 - __builtin_clz((unsigned)(p1 - p0));
            m = std::max(m, std::max(sp[k][p0], sp[k][p1 - (1 << k)]));
        }
        out[s] = m;
    }
    return out;
}

struct LaunchCfg {
    int kind = LAUNCH_AUTO, block = 256, grid = 0;
    int segv = 0;   // segment kernels: 0 cost-balanced segments, 1 segments of kSegItems items
    int pbs = -1;   // cooperative kernel: log2 of the primitive batch (-1: the class default)
};

struct ClassPairWork {
    int X, Y;
    int la, lb, lc, ld;
    int* d_ket;            // ket ids of class Y grouped by primitive count
    int* d_bra;            // bra ids (class X, sorted by Q desc)
    long long* d_row;
    int* d_rb;             // nrows x (nbucket+1) bucket offsets inside each row
    int* d_bstart;         // bucket starts of class Y
    int* d_bsoa;           // bucket offsets of class Y in the interleaved ket data
    int* d_rf;             // FP32 list: nrows x nbucket first kets (else null)
    int2* d_items;         // density-screened items (capacity nwork), allocated on first use
    SegDesc* d_seg;        // bra-row segments (molecular lists): nseg cost-balanced ones, then
    int nseg, nsegc;       // nsegc of kSegItems items each (make_segments)
    double* d_segq;        // per segment of d_seg: the largest ket bound (segment_ket_bounds)
    LaunchCfg cfg;         // launch shape (autotuned or default)
    int prec;              // 0: FP64, 1: FP32 (mixed precision)
    int both;              // periodic lists of two different classes: both bra/ket roles
    int nbucket;
    int nrows;
    long long nwork;
    double prim_quartets;
    double flops;
    // The same list over generally contracted pairs (classes (ls|l's) of a
    // basis with two-shell s groups, FP64 molecular contexts): used by the
    // unscreened builds while gc_on (gc_block threads per block).
    struct {
        long long nwork = 0;
        int nrows = 0, nbucket = 1, nseg = 0, nsegc = 0;
        const int *d_bra = nullptr, *d_ket = nullptr, *d_bstart = nullptr, *d_bsoa = nullptr;
        long long* d_row = nullptr;
        int* d_rb = nullptr;
        SegDesc* d_seg = nullptr;
        double* d_segq = nullptr;
        double prim_quartets = 0;
        int kind = 0;   // 0: group pairs on both sides (k_eri_gc); 1: shell-pair bra, group-pair kets (k_eri_gck)
    } gc;
    int gc_on = 0, gc_block = 256, gc_segv = 0;
    double est_ms = 0;     // launch time of the chosen form (autotuner), else 0: model FLOPs order the lists
};

struct ScfArena {
    std::vector<std::pair<void*, size_t>> bufs;
    size_t next = 0;
    int get_bytes(void** p, size_t bytes);
    template <typename T>
    int get(T** p, size_t n) {
        void* q = nullptr;
        const int rc = get_bytes(&q, n * sizeof(T));
        *p = static_cast<T*>(q);
        return rc;
    }
    void release();
};

struct polyeri_ctx {
    int tpack = 0;                 // MODE_TENSOR launches write the pair-packed tensor
    RefBasis rb;
    IntBasis ib;
    std::vector<PairInfo> pairs;
    std::vector<PrimPair> prims;
    std::vector<double> Q;
    polyeri_opts opts;
    int device = 0;
    hipStream_t stream = nullptr;
    PrimPair* d_prims = nullptr;
    PairMeta* d_meta = nullptr;
    double* d_ab = nullptr;
    double* d_boys = nullptr;
    double* d_Q = nullptr;
    int* d_class_ids[NPAIRCLASS] = {};
    std::vector<int> class_ids[NPAIRCLASS];
    // ket order of each class: buckets of equal primitive-pair count, each
    // bucket in Schwarz order
    int* d_ket_ids[NPAIRCLASS] = {};
    int* d_bstart[NPAIRCLASS] = {};
    std::vector<int> ket_ids[NPAIRCLASS], bstart[NPAIRCLASS];
    // ket primitive pairs bucket-interleaved: field f of primitive q of the
    // i-th ket of bucket b of class Y at d_ksoa[f * nprims + bsoa[Y][b] + q * nk_b + i]
    double* d_ksoa = nullptr;
    int* d_bsoa[NPAIRCLASS] = {};
    std::vector<ClassPairWork> work;
    // generally contracted pairs: metadata, bra primitive pairs, ket data
    // bucket-interleaved as d_ksoa with the four member weights as fields 8-11
    GcMeta* d_gmeta = nullptr;
    GcPrim* d_gprims = nullptr;
    double* d_gksoa = nullptr;
    long long gkfs = 0;
    int* d_gbra[NPAIRCLASS] = {};
    int* d_gket[NPAIRCLASS] = {};
    int* d_gbstart[NPAIRCLASS] = {};
    int* d_gbsoa[NPAIRCLASS] = {};
    int* d_ref = nullptr;
    double* d_scale = nullptr;
    double *d_D = nullptr, *d_J = nullptr, *d_K = nullptr;
    double *d_pc = nullptr, *d_jc = nullptr, *d_kc = nullptr;
    // one-electron (reference layout)
    std::vector<double> S_C, T_C, N_C;
    double e_nuc = 0.0;
    // stats
    long long quartets = 0;
    double prim_quartets = 0, model_flops = 0;
    bool profiling = false;
    std::vector<hipEvent_t> ev;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_jk_ms = 0;
    std::vector<double> last_class_ms;
    rocblas_handle blas = nullptr;
    double *d_T = nullptr, *d_V = nullptr;   // T_C, N_C on the device (Fock assembly)
    int jk_copies = 8;                       // J/K accumulation copies (one per XCD)
    // density screening: shells of every pair, shell extents, shell-block
    // maxima of |D|, compacted item lists (allocated on first use) and counts
    int2* d_psh = nullptr;
    int *d_sh_f0 = nullptr, *d_sh_nf = nullptr;
    double* d_dsh = nullptr;
    double* d_dmax = nullptr;   // largest entry of d_dsh (screened builds)
    unsigned* d_counts = nullptr;
    std::vector<unsigned> h_counts;
    bool screened_last = false, stored_last = false;
    long long last_quartets = 0;
    // J/K builds replayed from captured HIP graphs, one per (density, J, K,
    // screening) combination; dropped when the launch shapes change
    struct GraphEntry {
        const double* pc;
        double *jc, *kc;
        double tol;
        hipGraphExec_t exec;
    };
    bool use_graph = true;
    std::vector<GraphEntry> graphs;
    // captured builds fork the class launches into kBranches streams (the
    // context's own and kBranches - 1 side streams), joined before unpacking
    hipStream_t side[3] = {};
    hipEvent_t ev_fork = nullptr, ev_join[3] = {};
    // stored integrals (DIRECT .F.): .ao records in HBM, used by the J/K
    // builds and the MP2 tensor while use_stored; reference -> internal map
    StoredEri* d_ao = nullptr;
    long long n_ao = 0;
    int* d_inv = nullptr;
    bool use_stored = false;
    double* d_noalloc = nullptr;   // this device's address of `scratch`
    // the last molecular MP2: device time of the AO tensor and of the four
    // quarter transforms (incl. the MO coefficient set-up), their FLOPs
    double last_mp2_tensor_ms = 0, last_mp2_transform_ms = 0, last_mp2_flops = 0;
    // polyeri_scf: work buffers, and the one-electron matrices, C2S and the
    // canonical orthogonaliser of this geometry (kept per (spherical, lindep))
    ScfArena scf_arena;
    bool scf_orth = false;
    int scf_spherical = 0, scf_nind = 0;
    double scf_lindep = 0.0;
    double *d_scfT = nullptr, *d_scfV = nullptr, *d_scfS = nullptr, *d_scfC2S = nullptr, *d_scfX = nullptr;
    bool tuned = false;
    // periodic lattice sums along x (HELIX = 0): cells of the density and
    // of J/K in [-cel1, cel1], Coulomb sum over Q3 in [-cel2, cel2], pair
    // cells in [-rc1, rc1] (rc1: last cell with contracted overlap > 1e-15)
    bool pbc = false;
    double period = 0.0;
    int cel1 = 0, cel2 = 0, rc1 = 0;
    std::vector<int2_host> pcell;
    int2* d_pcell = nullptr;
    double *d_Dc = nullptr, *d_Jc = nullptr, *d_Kc = nullptr;   // J/K cells: fixed-point integers (jk_add_fx)
    unsigned* d_fx_overflow = nullptr;                            // set when a J/K cell sum left the fixed-point range
    double *d_pcc = nullptr, *d_jcc = nullptr, *d_kcc = nullptr;
};

static void drop_graphs(polyeri_ctx* c) {
    for (auto& g : c->graphs) (void)hipGraphExecDestroy(g.exec);
    c->graphs.clear();
}

static int ensure_onee_device(polyeri_ctx* c) {
    if (c->d_T) return 0;
    size_t n2 = c->T_C.size();
    HIPCHK(hipMalloc((void**)&c->d_T, n2 * sizeof(double)));
    HIPCHK(hipMalloc((void**)&c->d_V, n2 * sizeof(double)));
    HIPCHK(hipMemcpy(c->d_T, c->T_C.data(), n2 * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_V, c->N_C.data(), n2 * sizeof(double), hipMemcpyHostToDevice));
    return 0;
}

template <typename T>
static int dalloc(T** p, size_t n) {
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    if (e != hipSuccess) return fail(POLYERI_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    return 0;
}

static unsigned nblk(size_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// LDS bytes of the segment kernel for bra shells la, lb.
static size_t seg_lds_bytes(int la, int lb, int nbf) {
    return 8 * ((size_t)(ncart_h(la) + ncart_h(lb)) * nbf + (size_t)ncart_h(la) * ncart_h(lb));
}

// The launch kind LAUNCH_AUTO resolves to.
static int auto_kind(const polyeri_ctx* c, int la, int lb, int lc, int ld, bool seg_ok) {
    if (ncart_h(la) * ncart_h(lb) * ncart_h(lc) * ncart_h(ld) > POLYERI_COOP_MIN_INTEGRALS)
        return seg_ok && polyeri_rys_class(la, lb, lc, ld) && polyeri_rys_lds_bytes(la, lb, c->ib.nbf) <= 65536
                   ? LAUNCH_RYS
                   : LAUNCH_COOP;
    return seg_ok && seg_lds_bytes(la, lb, c->ib.nbf) <= 65536 ? LAUNCH_SEG : LAUNCH_REG;
}

// Density screening done inside the segment kernel.
struct SegScreen {
    const double* Q;
    const int2* psh;
    const double* dsh;
    int nsh;
    double tol;
    unsigned* kept;
    const double* segq;   // per segment of the launch: largest ket bound (null: no segment test)
    const double* dmax;   // device scalar: largest |D| block
};

// Whether a list's J/K launch (unscreened) runs the segment kernel.
// (The Rys kernel is a segment kernel too: both screen inside the kernel and
// keep no item lists.)
static bool list_uses_seg(const polyeri_ctx* c, const ClassPairWork& cw) {
    if (cw.nseg <= 0 || seg_lds_bytes(cw.la, cw.lb, c->ib.nbf) > 65536) return false;
    const int kind = cw.cfg.kind == LAUNCH_AUTO ? auto_kind(c, cw.la, cw.lb, cw.lc, cw.ld, true) : cw.cfg.kind;
    if (kind == LAUNCH_RYS)
        return polyeri_rys_class(cw.la, cw.lb, cw.lc, cw.ld) && polyeri_rys_lds_bytes(cw.la, cw.lb, c->ib.nbf) <= 65536;
    const int mode = cw.prec ? MODE_JK32 : MODE_JK;
    if (!polyeri_class_kernel(cw.la, cw.lb, cw.lc, cw.ld, mode)) return false;
    return kind == LAUNCH_SEG;
}

static int launch_class(polyeri_ctx* c, int mode, int la, int lb, int lc, int ld, const int* bra,
                        const int* ket, const long long* row, int nrows, long long nwork, double* out,
                        const int* rb = nullptr, const int* bstart = nullptr, int nbucket = 0,
                        const int* rf = nullptr, const int2* items = nullptr, const unsigned* count = nullptr,
                        LaunchCfg cfg = LaunchCfg(), const int* bsoa = nullptr, const SegDesc* segs = nullptr,
                        int nseg = 0, const SegScreen* scr = nullptr) {
    if (nwork <= 0) return 0;
    EriArgs a{};
    a.ksoa = c->d_ksoa; a.kfs = (long long)c->prims.size(); a.bucket_soa = bsoa;
    if (c->pbc) {
        a.n3 = 2 * c->cel2 + 1; a.cel1 = c->cel1; a.cel2 = c->cel2; a.period = c->period;
        a.pcell = c->d_pcell; a.Dc = c->d_Dc; a.Jc = c->d_Jc; a.Kc = c->d_Kc;
    }
    a.prims = c->d_prims; a.meta = c->d_meta; a.ab = c->d_ab;
    a.bra_ids = bra; a.ket_ids = ket; a.row_off = row; a.nrows = nrows; a.nwork = nwork;
    a.rb_off = rb; a.bucket_start = bstart; a.nbucket = nbucket; a.rb_first = rf;
    a.items = items; a.count = count;
    a.boys = c->d_boys; a.boysA = c->d_boys + BOYS_AOFF;
    a.D = c->d_D; a.J = c->d_J; a.K = c->d_K; a.nbf = c->ib.nbf; a.out = out;
    a.jk_copies = c->jk_copies;
    a.tpack = mode == MODE_TENSOR ? c->tpack : 0;
    EriKernel k = mode == MODE_JK ? kernel_for<MODE_JK>(la, lb, lc, ld)
                : mode == MODE_JK32 ? kernel_for<MODE_JK32>(la, lb, lc, ld)
                : mode == MODE_SCHWARZ ? kernel_for<MODE_SCHWARZ>(la, lb, lc, ld)
                : kernel_for<MODE_TENSOR>(la, lb, lc, ld);
    // the segment kernel needs the static list (no screened items), a J/K mode and LDS for the rows
    const bool seg_ok = segs && nseg > 0 && !items && (mode == MODE_JK || mode == MODE_JK32) &&
                        seg_lds_bytes(la, lb, c->ib.nbf) <= 65536;
    int kind = cfg.kind;
    if (kind == LAUNCH_AUTO) {
        kind = auto_kind(c, la, lb, lc, ld, seg_ok);   // Rys, wavefront, segment or register kernel
        if ((kind == LAUNCH_REG || kind == LAUNCH_SEG) && !k) kind = LAUNCH_COOP;
    }
    if (kind == LAUNCH_SEG && !seg_ok) kind = k ? LAUNCH_REG : LAUNCH_COOP;
    if (kind == LAUNCH_REG && !k) return fail(POLYERI_E_ARG, "no register kernel for class");
    if (kind == LAUNCH_SEG && mode == MODE_JK32)
        return fail(POLYERI_E_ARG, "the FP32 segment kernel is switched off (unexplained fault)");
    if (kind == LAUNCH_RYS) {
        const bool rys_ok = seg_ok && polyeri_rys_class(la, lb, lc, ld) &&
                            polyeri_rys_lds_bytes(la, lb, c->ib.nbf) <= 65536;
        if (!rys_ok) {
            kind = LAUNCH_COOP;
        } else {
            a.segs = segs;
            a.nseg = nseg;
            if (scr) {
                a.dtol = scr->tol; a.Qs = scr->Q; a.psh = scr->psh; a.dsh = scr->dsh; a.nsh = scr->nsh;
                a.kept = scr->kept; a.segq = scr->segq; a.dmax = scr->dmax;
            }
            if (polyeri_rys_launch(a, mode, la, lb, lc, ld, cfg.block, cfg.grid, c->stream))
                return fail(POLYERI_E_HIP, "Rys kernel launch failed");
            return 0;
        }
    }
    if (kind == LAUNCH_SEG) {
        EriKernel ks = polyeri_class_kernel(la, lb, lc, ld, mode == MODE_JK ? GET_SEG_JK : GET_SEG_JK32);
        if (!ks) return fail(POLYERI_E_ARG, "no segment kernel for class");
        a.segs = segs;
        a.nseg = nseg;
        if (scr) {
            a.dtol = scr->tol; a.Qs = scr->Q; a.psh = scr->psh; a.dsh = scr->dsh; a.nsh = scr->nsh;
            a.kept = scr->kept; a.segq = scr->segq; a.dmax = scr->dmax;
        }
        const int block = cfg.block == 64 || cfg.block == 128 ? cfg.block : 256;
        unsigned grid = (unsigned)nseg;
        if (cfg.grid > 0 && grid > (unsigned)cfg.grid) grid = (unsigned)cfg.grid;
        hipLaunchKernelGGL(ks, dim3(grid), dim3(block), seg_lds_bytes(la, lb, c->ib.nbf), c->stream, a);
        HIPCHK(hipGetLastError());
        return 0;
    }
    if (kind == LAUNCH_COOP) {
        const int pbs = cfg.pbs >= 0 ? cfg.pbs
                                     : polyeri_coop_default_pbs(ncart_h(la) * ncart_h(lb) * ncart_h(lc) * ncart_h(ld));
        const int grid = cfg.grid > 0 ? cfg.grid : 4096;
        if (polyeri_coop_launch(a, mode, la, lb, lc, ld, pbs, grid, c->stream))
            return fail(POLYERI_E_HIP, "cooperative kernel launch failed");
        return 0;
    }
    const int block = cfg.block == 64 || cfg.block == 128 ? cfg.block : 256;
    // one pass over the list, or a fixed (persistent) grid that strides; a
    // screened list holds at most nwork items and always strides
    unsigned grid = nblk((size_t)nwork, block);
    unsigned cap = cfg.grid > 0 ? (unsigned)cfg.grid : (items ? 8192u * (256 / block) : 0u);
    if (cap && grid > cap) grid = cap;
    hipLaunchKernelGGL(k, dim3(grid), dim3(block), 0, c->stream, a);
    HIPCHK(hipGetLastError());
    return 0;
}

// Whether a list's J/K launches run its generally contracted form.
static bool list_uses_gc(const ClassPairWork& cw) { return cw.gc_on && cw.gc.nwork > 0; }

// In a density-screened build: whether the list screens inside its segment
// (or generally contracted, or Rys) kernel; the others have their surviving
// items compacted first (k_screen) for the thread or wavefront kernel.  (A
// compacting variant for every list with a thread kernel measured 3x
// slower on the (H2O)16 incremental builds: the compaction pass and the
// per-quartet J/K atomics cost more than the idle lanes it removes.)
static bool screens_in_kernel(const polyeri_ctx* c, const ClassPairWork& cw) {
    return list_uses_seg(c, cw) || list_uses_gc(cw);
}

// J/K launch of a list's generally contracted form (one block per segment),
// density-screened inside the kernel when scr is given; with tensor (lists of
// shell-pair-group bras only), the integrals written to it instead.
static int launch_gc(polyeri_ctx* c, const ClassPairWork& cw, int block, const SegScreen* scr = nullptr,
                     double* tensor = nullptr) {
    const auto& g = cw.gc;
    if (g.nwork <= 0) return 0;
    EriArgs a{};
    a.out = tensor;
    a.tpack = tensor ? c->tpack : 0;
    a.bra_ids = g.d_bra; a.ket_ids = g.d_ket; a.row_off = g.d_row; a.nrows = g.nrows; a.nwork = g.nwork;
    a.rb_off = g.d_rb; a.bucket_start = g.d_bstart; a.nbucket = g.nbucket; a.rb_first = nullptr;
    a.ksoa = c->d_gksoa; a.kfs = c->gkfs; a.bucket_soa = g.d_bsoa;
    a.boys = c->d_boys; a.boysA = c->d_boys + BOYS_AOFF;
    a.D = c->d_D; a.J = c->d_J; a.K = c->d_K; a.nbf = c->ib.nbf; a.jk_copies = c->jk_copies;
    a.segs = cw.gc_segv ? g.d_seg + g.nseg : g.d_seg;
    a.nseg = cw.gc_segv ? g.nsegc : g.nseg;
    a.gmeta = c->d_gmeta; a.gprims = c->d_gprims; a.qtol = c->opts.schwarz_tol;
    if (scr) {
        a.dtol = scr->tol; a.dsh = scr->dsh; a.nsh = scr->nsh; a.kept = scr->kept; a.dmax = scr->dmax;
        a.segq = g.d_segq ? g.d_segq + (cw.gc_segv ? g.nseg : 0) : nullptr;
    }
    if (block != 64 && block != 128) block = 256;
    if (g.kind == 1) {   // shell-pair bra
        if (tensor) return fail(POLYERI_E_ARG, "generally contracted ket lists have no tensor form");
        a.prims = c->d_prims; a.meta = c->d_meta; a.ab = c->d_ab; a.Qs = c->d_Q; a.psh = c->d_psh;
        if (polyeri_gck_launch(a, cw.la, cw.lb, cw.lc, block, a.nseg, c->stream))
            return fail(POLYERI_E_HIP, "generally contracted ket kernel launch failed");
        return 0;
    }
    if (polyeri_gc_launch(a, cw.la, cw.lc, block, a.nseg, c->stream, tensor != nullptr))
        return fail(POLYERI_E_HIP, "generally contracted kernel launch failed");
    return 0;
}

// tensor: store the integrals T[i][Q1][j][Q3][k][Q2][l] instead of digesting
// them into J/K (pair cells in the reduced range [-rc1, rc1]).
static int launch_class_pbc(polyeri_ctx* c, const ClassPairWork& cw, long long nwork, double* tensor = nullptr) {
    EriArgs a{};
    a.prims = c->d_prims; a.meta = c->d_meta; a.ab = c->d_ab;
    a.bra_ids = cw.d_bra; a.ket_ids = cw.d_ket; a.row_off = cw.d_row; a.nrows = cw.nrows; a.nwork = nwork;
    a.rb_off = cw.d_rb; a.bucket_start = cw.d_bstart; a.nbucket = cw.nbucket;
    a.boys = c->d_boys; a.boysA = c->d_boys + BOYS_AOFF; a.nbf = c->ib.nbf; a.jk_copies = c->jk_copies;
    a.n3 = 2 * c->cel2 + 1; a.cel1 = c->cel1; a.cel2 = c->cel2; a.period = c->period; a.both = cw.both;
    a.pcell = c->d_pcell; a.Dc = c->d_Dc; a.Jc = c->d_Jc; a.Kc = c->d_Kc;
    if (tensor) { a.out = tensor; a.cel1 = c->rc1; }
    if (nwork <= 0) return 0;
    EriKernel k = kernel_for<MODE_PBC>(cw.la, cw.lb, cw.lc, cw.ld);
    const int nint4 = ncart_h(cw.la) * ncart_h(cw.lb) * ncart_h(cw.lc) * ncart_h(cw.ld);
    if (!k || nint4 > POLYERI_COOP_MIN_INTEGRALS) {
        if (polyeri_coop_launch(a, MODE_PBC, cw.la, cw.lb, cw.lc, cw.ld, polyeri_coop_default_pbs(nint4), 4096,
                                c->stream))
            return fail(POLYERI_E_HIP, "cooperative kernel launch failed");
        return 0;
    }
    if (!tensor) {
        // J/K: each thread loops over a chunk of Q3 values and adds its J
        // block once; (ss|ss), whose one-integral digestion gains little
        // and whose primitive loops are long, keeps more and shorter chunks
        const int nint = ncart_h(cw.la) * ncart_h(cw.lb) * ncart_h(cw.lc) * ncart_h(cw.ld);
        a.qloop = nint == 1 ? 3 : a.n3;
        a.nwork = nwork / a.n3 * ((a.n3 + a.qloop - 1) / a.qloop);
    }
    hipLaunchKernelGGL(k, dim3(nblk((size_t)a.nwork)), dim3(256), 0, c->stream, a);
    HIPCHK(hipGetLastError());
    return 0;
}

// One stream (one hardware queue) per device shared by every context: the
// device's scratch reservation for the spilling high-l kernels is then made
// once, not once per context.
static hipStream_t shared_stream(int device, int* rc) {
    static std::mutex mu;
    static std::vector<hipStream_t> streams;
    std::lock_guard<std::mutex> lk(mu);
    if ((int)streams.size() <= device) streams.resize(device + 1, nullptr);
    if (!streams[device]) {
        hipError_t e = hipStreamCreateWithFlags(&streams[device], hipStreamNonBlocking);
        if (e != hipSuccess) { *rc = fail(POLYERI_E_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e)); return nullptr; }
    }
    *rc = 0;
    return streams[device];
}

// Fixed device workspace of polyeri_jk_build_device_noalloc: the packed
// density and the per-XCD J/K accumulation copies of one build (instead of
// the context's own), so that entry point allocates nothing.  Each device
// holds its own copy of the array; host code uses only the address that
// hipGetSymbolAddress returns on that device (kept per device below).
static constexpr size_t kNoAllocDoubles = size_t(1) << 22;   // 32 MiB
static __device__ __align__(256) double scratch[kNoAllocDoubles];

static int noalloc_scratch(int device, double** out) {
    static std::mutex mu;
    static std::vector<double*> addr;
    std::lock_guard<std::mutex> lk(mu);
    if (device < 0) return fail(POLYERI_E_ARG, "bad device");
    if ((int)addr.size() <= device) addr.resize(device + 1, nullptr);
    if (!addr[device]) {   // the caller has made `device` current
        double* p = nullptr;
        hipError_t e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(POLYERI_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        addr[device] = p;
    }
    *out = addr[device];
    return 0;
}

// Side streams of the captured J/K builds (default priority, shared by the
// contexts of a device like the main stream).
static const int kBranches = 4;
static int shared_side_streams(int device, hipStream_t* out) {
    static std::mutex mu;
    static std::vector<std::array<hipStream_t, kBranches - 1>> streams;
    std::lock_guard<std::mutex> lk(mu);
    if ((int)streams.size() <= device) streams.resize(device + 1, std::array<hipStream_t, kBranches - 1>{});
    for (int i = 0; i < kBranches - 1; ++i) {
        if (!streams[device][i]) {
            hipError_t e = hipStreamCreateWithFlags(&streams[device][i], hipStreamNonBlocking);
            if (e != hipSuccess) return fail(POLYERI_E_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
        }
        out[i] = streams[device][i];
    }
    return 0;
}

static void free_ctx(polyeri_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    (void)hipFree(c->d_prims); (void)hipFree(c->d_meta); (void)hipFree(c->d_ab); (void)hipFree(c->d_boys); (void)hipFree(c->d_Q);
    for (int i = 0; i < NPAIRCLASS; ++i) {
        (void)hipFree(c->d_class_ids[i]); (void)hipFree(c->d_ket_ids[i]); (void)hipFree(c->d_bstart[i]); (void)hipFree(c->d_bsoa[i]);
    }
    (void)hipFree(c->d_ksoa);
    for (auto& w : c->work) {
        (void)hipFree(w.d_row); (void)hipFree(w.d_rb);
        if (w.d_rf) (void)hipFree(w.d_rf);
        if (w.d_items) (void)hipFree(w.d_items);
        if (w.d_seg) (void)hipFree(w.d_seg);
        (void)hipFree(w.d_segq); (void)hipFree(w.gc.d_segq);
        if (w.gc.d_row) (void)hipFree(w.gc.d_row);
        if (w.gc.d_rb) (void)hipFree(w.gc.d_rb);
        if (w.gc.d_seg) (void)hipFree(w.gc.d_seg);
    }
    (void)hipFree(c->d_gmeta); (void)hipFree(c->d_gprims); (void)hipFree(c->d_gksoa);
    for (int i = 0; i < NPAIRCLASS; ++i) {
        (void)hipFree(c->d_gbra[i]); (void)hipFree(c->d_gket[i]); (void)hipFree(c->d_gbstart[i]); (void)hipFree(c->d_gbsoa[i]);
    }
    (void)hipFree(c->d_pcell); (void)hipFree(c->d_Dc); (void)hipFree(c->d_Jc); (void)hipFree(c->d_Kc);
    (void)hipFree(c->d_fx_overflow);
    (void)hipFree(c->d_pcc); (void)hipFree(c->d_jcc); (void)hipFree(c->d_kcc);
    (void)hipFree(c->d_psh); (void)hipFree(c->d_sh_f0); (void)hipFree(c->d_sh_nf); (void)hipFree(c->d_dsh); (void)hipFree(c->d_counts); (void)hipFree(c->d_dmax);
    (void)hipFree(c->d_ref); (void)hipFree(c->d_scale); (void)hipFree(c->d_D); (void)hipFree(c->d_J); (void)hipFree(c->d_K);
    (void)hipFree(c->d_pc); (void)hipFree(c->d_jc); (void)hipFree(c->d_kc); (void)hipFree(c->d_T); (void)hipFree(c->d_V);
    for (auto e : c->ev) (void)hipEventDestroy(e);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    (void)hipFree(c->d_ao); (void)hipFree(c->d_inv);
    c->scf_arena.release();
    (void)hipFree(c->d_scfT); (void)hipFree(c->d_scfV); (void)hipFree(c->d_scfS); (void)hipFree(c->d_scfC2S);
    (void)hipFree(c->d_scfX);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    for (auto e : c->ev_join)
        if (e) (void)hipEventDestroy(e);
    if (c->blas) rocblas_destroy_handle(c->blas);
    drop_graphs(c);
    delete c;
}

void polyeri_ctx_destroy(polyeri_ctx* ctx) { free_ctx(ctx); }

// Last cell offset q in [0, cel1] whose contracted overlap with cell 0 has
// an element above 1e-15 (REDUCED_CEL1X, src/ponee.f90:359-371).
static int reduced_cel1(const IntBasis& ib, double period, int cel1, const std::vector<double>& boys) {
    int rc1 = 0;
    for (int q = 1; q <= cel1; ++q) {
        IntBasis two = ib;
        for (const Shell& sh : ib.shells) {
            Shell t = sh;
            t.x += q * period;
            t.f0 += ib.nbf;
            two.shells.push_back(t);
        }
        two.nbf = 2 * ib.nbf;
        std::vector<double> S, T, V;
        one_electron(two, boys, S, T, V);
        double mx = 0.0;
        for (int a = 0; a < ib.nbf; ++a)
            for (int b = 0; b < ib.nbf; ++b)
                mx = std::max(mx, std::fabs(ib.scale[a] * ib.scale[b] * S[(size_t)a * two.nbf + ib.nbf + b]));
        if (mx > 1.0e-15) rc1 = q;
    }
    return rc1;
}

// Generally contracted lists (build_gc_pairs): Schwarz bound of a pair =
// the largest of its members', class lists in that order, the work lists
// of host.cpp, and each list with a generally contracted kernel attached to
// the FP64 list of the same class pair (segments as for the plain lists).
// Quartet counts stay those of the plain lists.
static int build_gc_lists(polyeri_ctx* c) {
    std::vector<PairInfo> gp;
    std::vector<GcMeta> gm;
    std::vector<GcPrim> gprim;
    std::vector<int> mp;
    if (build_gc_pairs(c->ib, c->opts.prim_tol, c->pairs, gp, gm, gprim, mp) == 0) return 0;
    const int ng = (int)gp.size();
    std::vector<double> gQ(ng, 0.0);
    for (int i = 0; i < ng; ++i)
        for (int k = 0; k < GC_MAXM; ++k) {
            const int pk = mp[(size_t)GC_MAXM * i + k];
            gm[i].q[k] = pk >= 0 ? c->Q[pk] : 0.0;
            gQ[i] = std::max(gQ[i], gm[i].q[k]);
        }
    std::vector<int> cls[NPAIRCLASS], kets[NPAIRCLASS], bst[NPAIRCLASS];
    for (int i = 0; i < ng; ++i) cls[gp[i].cls].push_back(i);
    for (int X = 0; X < NPAIRCLASS; ++X)
        std::stable_sort(cls[X].begin(), cls[X].end(), [&](int a, int b) { return gQ[a] > gQ[b]; });
    std::vector<WorkList> wl;
    build_work_lists(gp, gQ, c->opts.schwarz_tol, 0.0, cls, kets, bst, wl);
    int rc;
    const size_t P = gprim.size();
    if ((rc = dalloc(&c->d_gmeta, gm.size())) || (rc = dalloc(&c->d_gprims, P))) return rc;
    HIPCHK(hipMemcpy(c->d_gmeta, gm.data(), gm.size() * sizeof(GcMeta), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_gprims, gprim.data(), P * sizeof(GcPrim), hipMemcpyHostToDevice));
    std::vector<double> soa(12 * P, 0.0);
    size_t off = 0;
    for (int Y = 0; Y < NPAIRCLASS; ++Y) {
        const auto& kl = kets[Y];
        const auto& bs = bst[Y];
        std::vector<int> bo(std::max<size_t>(bs.size(), 1), 0);
        for (size_t b = 0; b + 1 < bs.size(); ++b) {
            const int nk = bs[b + 1] - bs[b];
            const int np_b = gp[kl[bs[b]]].nprim;
            bo[b] = (int)off;
            for (int i = 0; i < nk; ++i) {
                const PairInfo& pi = gp[kl[bs[b] + i]];
                for (int q = 0; q < np_b; ++q) {
                    const GcPrim& x = gprim[pi.prim_off + q];
                    const size_t at = off + (size_t)q * nk + i;
                    const double f[12] = {x.pp.zeta, x.pp.K, x.pp.Px, x.pp.Py, x.pp.Pz, x.pp.PAx, x.pp.PAy, x.pp.PAz,
                                          x.w[0], x.w[1], x.w[2], x.w[3]};
                    for (int k = 0; k < 12; ++k) soa[k * P + at] = f[k];
                }
            }
            off += (size_t)np_b * nk;
        }
        if ((rc = dalloc(&c->d_gbra[Y], cls[Y].size())) || (rc = dalloc(&c->d_gket[Y], kl.size())) ||
            (rc = dalloc(&c->d_gbstart[Y], bs.size())) || (rc = dalloc(&c->d_gbsoa[Y], bo.size())))
            return rc;
        if (!cls[Y].empty()) {
            HIPCHK(hipMemcpy(c->d_gbra[Y], cls[Y].data(), cls[Y].size() * sizeof(int), hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(c->d_gket[Y], kl.data(), kl.size() * sizeof(int), hipMemcpyHostToDevice));
        }
        HIPCHK(hipMemcpy(c->d_gbstart[Y], bs.data(), bs.size() * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->d_gbsoa[Y], bo.data(), bo.size() * sizeof(int), hipMemcpyHostToDevice));
    }
    if (off != P) return fail(POLYERI_E_BASIS, "generally contracted ket layout does not cover every primitive pair");
    if ((rc = dalloc(&c->d_gksoa, soa.size()))) return rc;
    HIPCHK(hipMemcpy(c->d_gksoa, soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice));
    c->gkfs = (long long)P;
    for (const WorkList& l : wl) {
        if (l.nwork <= 0) continue;
        const int la = kPairL[l.X][0], lc = kPairL[l.Y][0];
        if (kPairL[l.X][1] != 0 || kPairL[l.Y][1] != 0 || !polyeri_gc_class(la, lc) ||
            polyeri_gc_lds_bytes(la, c->ib.nbf) > 65536)
            continue;
        ClassPairWork* w = nullptr;
        for (ClassPairWork& x : c->work)
            if (x.X == l.X && x.Y == l.Y && x.prec == 0) w = &x;
        if (!w) continue;
        std::vector<SegDesc> segs = make_segments(l, gp, cls[l.X], kets[l.Y], bst[l.Y]);
        const std::vector<SegDesc> segc = make_segments(l, gp, cls[l.X], kets[l.Y], bst[l.Y], false);
        if (segs.size() + segc.size() > 0x7fffffffu) return fail(POLYERI_E_ARG, "too many segments");
        auto& g = w->gc;
        g.nwork = l.nwork; g.nrows = l.nrows; g.nbucket = l.nbucket;
        g.nseg = (int)segs.size(); g.nsegc = (int)segc.size();
        segs.insert(segs.end(), segc.begin(), segc.end());
        {
            const std::vector<double> sq = segment_ket_bounds(l, segs, kets[l.Y], bst[l.Y], gQ);
            if ((rc = dalloc(&g.d_segq, sq.size()))) return rc;
            HIPCHK(hipMemcpy(g.d_segq, sq.data(), sq.size() * sizeof(double), hipMemcpyHostToDevice));
        }
        g.d_bra = c->d_gbra[l.X]; g.d_ket = c->d_gket[l.Y]; g.d_bstart = c->d_gbstart[l.Y];
        g.d_bsoa = c->d_gbsoa[l.Y];
        g.prim_quartets = l.prim_quartets;
        if ((rc = dalloc(&g.d_row, l.row.size())) || (rc = dalloc(&g.d_rb, l.rb.size())) ||
            (rc = dalloc(&g.d_seg, segs.size())))
            return rc;
        HIPCHK(hipMemcpy(g.d_row, l.row.data(), l.row.size() * sizeof(long long), hipMemcpyHostToDevice));
        if (!l.rb.empty()) HIPCHK(hipMemcpy(g.d_rb, l.rb.data(), l.rb.size() * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(g.d_seg, segs.data(), segs.size() * sizeof(SegDesc), hipMemcpyHostToDevice));
        w->gc_on = 1;
    }
    // Shell-pair bras of the classes (pp) (ds) (dp) against group-pair kets
    // of (ss) (ps): one list set over the shell pairs followed by the group
    // pairs, so the ket order, buckets and interleaved ket data are the
    // group-pair lists' own.  Off until the autotuner finds it faster.
    const int np = (int)c->pairs.size();
    std::vector<PairInfo> cp(c->pairs);
    cp.insert(cp.end(), gp.begin(), gp.end());
    std::vector<double> cq(c->Q);
    cq.insert(cq.end(), gQ.begin(), gQ.end());
    std::vector<int> cid[NPAIRCLASS], ck[NPAIRCLASS], cb[NPAIRCLASS];
    for (int X : {2, 3, 4}) cid[X] = c->class_ids[X];
    for (int Y : {0, 1})
        for (int i : cls[Y]) cid[Y].push_back(np + i);
    std::vector<WorkList> wk;
    build_work_lists(cp, cq, c->opts.schwarz_tol, 0.0, cid, ck, cb, wk);
    for (int Y : {0, 1}) {
        if (ck[Y].size() != kets[Y].size() || cb[Y] != bst[Y])
            return fail(POLYERI_E_BASIS, "generally contracted ket order differs between list sets");
        for (size_t i = 0; i < ck[Y].size(); ++i)
            if (ck[Y][i] != np + kets[Y][i])
                return fail(POLYERI_E_BASIS, "generally contracted ket order differs between list sets");
    }
    for (const WorkList& l : wk) {
        if (l.nwork <= 0 || l.Y > 1 || l.X < 2) continue;
        const int la = kPairL[l.X][0], lb = kPairL[l.X][1], lc = kPairL[l.Y][0];
        if (!polyeri_gck_class(la, lb, lc) || polyeri_gck_lds_bytes(la, lb, c->ib.nbf) > 65536) continue;
        ClassPairWork* w = nullptr;
        for (ClassPairWork& x : c->work)
            if (x.X == l.X && x.Y == l.Y && x.prec == 0) w = &x;
        if (!w || w->gc.nwork > 0) continue;
        std::vector<SegDesc> segs = make_segments(l, cp, cid[l.X], ck[l.Y], cb[l.Y]);
        const std::vector<SegDesc> segc = make_segments(l, cp, cid[l.X], ck[l.Y], cb[l.Y], false);
        if (segs.size() + segc.size() > 0x7fffffffu) return fail(POLYERI_E_ARG, "too many segments");
        auto& g = w->gc;
        g.kind = 1;
        {
            std::vector<SegDesc> all(segs);
            all.insert(all.end(), segc.begin(), segc.end());
            const std::vector<double> sq = segment_ket_bounds(l, all, ck[l.Y], cb[l.Y], cq);
            if ((rc = dalloc(&g.d_segq, sq.size()))) return rc;
            HIPCHK(hipMemcpy(g.d_segq, sq.data(), sq.size() * sizeof(double), hipMemcpyHostToDevice));
        }
        g.nwork = l.nwork; g.nrows = l.nrows; g.nbucket = l.nbucket;
        g.nseg = (int)segs.size(); g.nsegc = (int)segc.size();
        segs.insert(segs.end(), segc.begin(), segc.end());
        g.d_bra = c->d_class_ids[l.X]; g.d_ket = c->d_gket[l.Y]; g.d_bstart = c->d_gbstart[l.Y];
        g.d_bsoa = c->d_gbsoa[l.Y];
        g.prim_quartets = l.prim_quartets;
        if ((rc = dalloc(&g.d_row, l.row.size())) || (rc = dalloc(&g.d_rb, l.rb.size())) ||
            (rc = dalloc(&g.d_seg, segs.size())))
            return rc;
        HIPCHK(hipMemcpy(g.d_row, l.row.data(), l.row.size() * sizeof(long long), hipMemcpyHostToDevice));
        if (!l.rb.empty()) HIPCHK(hipMemcpy(g.d_rb, l.rb.data(), l.rb.size() * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(g.d_seg, segs.data(), segs.size() * sizeof(SegDesc), hipMemcpyHostToDevice));
        w->gc_on = 0;
    }
    return 0;
}

static int build_ctx(polyeri_ctx* c) {
    std::string err;
    if (extract_shells(c->rb, c->ib, err)) return fail(POLYERI_E_BASIS, err);
    if (c->pbc) {
        std::vector<double> boys0;
        boys_table(boys0);
        c->rc1 = reduced_cel1(c->ib, c->period, c->cel1, boys0);
        build_pairs_cells(c->ib, c->opts.prim_tol, c->period, c->rc1, c->pairs, c->prims, c->pcell);
    } else {
        build_pairs(c->ib, c->opts.prim_tol, c->pairs, c->prims);
    }
    HIPCHK(hipSetDevice(c->device));
    int src = 0;
    c->stream = shared_stream(c->device, &src);
    if (src) return src;
    HIPCHK(hipEventCreate(&c->ev0));
    HIPCHK(hipEventCreate(&c->ev1));
    if ((src = shared_side_streams(c->device, c->side))) return src;
    HIPCHK(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
    for (auto& e : c->ev_join) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    int np = (int)c->pairs.size(), nbf = c->ib.nbf, nc = c->rb.ncgs;
    std::vector<PairMeta> meta(np);
    std::vector<double> ab(4 * (size_t)np);
    for (int i = 0; i < np; ++i) {
        const PairInfo& p = c->pairs[i];
        meta[i] = {p.prim_off, p.nprim, c->ib.shells[p.a].f0, c->ib.shells[p.b].f0};
        ab[4 * i] = p.ABx; ab[4 * i + 1] = p.ABy; ab[4 * i + 2] = p.ABz; ab[4 * i + 3] = 0;
    }
    std::vector<double> boys;
    boys_table(boys);
    int rc;
    if ((rc = dalloc(&c->d_prims, c->prims.size()))) return rc;
    if ((rc = dalloc(&c->d_meta, np))) return rc;
    if ((rc = dalloc(&c->d_ab, 4 * (size_t)np))) return rc;
    // device copy: the BOYS_NT x BOYS_NM table, then the aligned per-L tables
    // of the thread kernels (BOYS_LA x BOYS_NT rows of F_L .. F_{L+6}, pad),
    // then the coarse per-L rows (BOYS_LC x BOYS_CNT rows of F_L .. F_{L+9}
    // at every fourth grid point)
    static_assert(BOYS_LC - 1 + 9 < BOYS_NM && 4 * (BOYS_CNT - 1) < BOYS_NT, "coarse Boys rows out of the table");
    std::vector<double> boys_dev(boys);
    boys_dev.resize((size_t)BOYS_COFF + (size_t)BOYS_LC * BOYS_CROWS, 0.0);
    for (int L = 0; L < BOYS_LA; ++L)
        for (int i = 0; i < BOYS_NT; ++i)
            for (int j = 0; j < 7; ++j)
                boys_dev[(size_t)BOYS_AOFF + ((size_t)L * BOYS_NT + i) * 8 + j] = boys[(size_t)i * BOYS_NM + L + j];
    for (int L = 0; L < BOYS_LC; ++L)
        for (int i = 0; i < BOYS_CNT; ++i)
            for (int j = 0; j < 10; ++j)
                boys_dev[(size_t)BOYS_COFF + (size_t)L * BOYS_CROWS + i * 10 + j] = boys[(size_t)4 * i * BOYS_NM + L + j];
    if ((rc = dalloc(&c->d_boys, boys_dev.size()))) return rc;
    if ((rc = dalloc(&c->d_Q, np))) return rc;
    if ((rc = dalloc(&c->d_ref, nbf))) return rc;
    if ((rc = dalloc(&c->d_scale, nbf))) return rc;
    if ((rc = dalloc(&c->d_D, (size_t)nbf * nbf))) return rc;
    {
        const char* e = getenv("POLYERI_JK_COPIES");
        c->jk_copies = e ? atoi(e) : 8;
        const char* gr = getenv("POLYERI_GRAPH");
        c->use_graph = !(gr && gr[0] == '0');
        if (c->jk_copies != 1 && c->jk_copies != 8) c->jk_copies = 8;
    }
    if ((rc = dalloc(&c->d_J, (size_t)nbf * nbf * c->jk_copies))) return rc;
    if ((rc = dalloc(&c->d_K, (size_t)nbf * nbf * c->jk_copies))) return rc;
    if ((rc = dalloc(&c->d_pc, (size_t)nc * nc))) return rc;
    if ((rc = dalloc(&c->d_jc, (size_t)nc * nc))) return rc;
    if ((rc = dalloc(&c->d_kc, (size_t)nc * nc))) return rc;
    HIPCHK(hipMemcpy(c->d_prims, c->prims.data(), c->prims.size() * sizeof(PrimPair), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_meta, meta.data(), np * sizeof(PairMeta), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_ab, ab.data(), ab.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_boys, boys_dev.data(), boys_dev.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_ref, c->ib.ref_of_int.data(), nbf * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->d_scale, c->ib.scale.data(), nbf * sizeof(double), hipMemcpyHostToDevice));
    // Schwarz bounds Q_ab = sqrt(max |(ab|ab)|), one class at a time
    for (int X = 0; X < NPAIRCLASS; ++X) {
        c->class_ids[X].clear();
        for (int i = 0; i < np; ++i) if (c->pairs[i].cls == X) c->class_ids[X].push_back(i);
        if ((rc = dalloc(&c->d_class_ids[X], c->class_ids[X].size()))) return rc;
        if (c->class_ids[X].empty()) continue;
        HIPCHK(hipMemcpy(c->d_class_ids[X], c->class_ids[X].data(), c->class_ids[X].size() * sizeof(int),
                         hipMemcpyHostToDevice));
        int la = kPairL[X][0], lb = kPairL[X][1];
        if ((rc = launch_class(c, MODE_SCHWARZ, la, lb, la, lb, c->d_class_ids[X], nullptr, nullptr, 0,
                               (long long)c->class_ids[X].size(), c->d_Q)))
            return rc;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    c->Q.resize(np);
    HIPCHK(hipMemcpy(c->Q.data(), c->d_Q, np * sizeof(double), hipMemcpyDeviceToHost));
    for (int i = 0; i < np; ++i)
        if (!std::isfinite(c->Q[i])) return fail(POLYERI_E_NONFINITE, "non-finite Schwarz bound");
    // class lists sorted by Q descending (Schwarz screening becomes a prefix)
    for (int X = 0; X < NPAIRCLASS; ++X) {
        auto& ids = c->class_ids[X];
        std::stable_sort(ids.begin(), ids.end(), [&](int a, int b) { return c->Q[a] > c->Q[b]; });
        if (!ids.empty())
            HIPCHK(hipMemcpy(c->d_class_ids[X], ids.data(), ids.size() * sizeof(int), hipMemcpyHostToDevice));
    }
    // class-pair work lists (host.cpp build_work_lists).  Mixed precision
    // splits the class pairs of the wavefront kernel only: the light classes
    // keep one FP64 list each and run the segment and generally contracted
    // kernels, as in FP64 contexts (an FP32 recursion issues at the FP64 rate
    // on gfx950 without packed FP32 math, so it saves them nothing, and the
    // FP32 segment kernel stays switched off; see launch_class).
    std::vector<WorkList> wl;
    const double cut = c->opts.precision == POLYERI_MIXED ? c->opts.mixed_cut : 0.0;
    uint64_t split = 0;
    for (int X = 0; X < NPAIRCLASS; ++X)
        for (int Y = 0; Y <= X; ++Y)
            if (ncart_h(kPairL[X][0]) * ncart_h(kPairL[X][1]) * ncart_h(kPairL[Y][0]) * ncart_h(kPairL[Y][1]) >
                POLYERI_COOP_MIN_INTEGRALS)
                split |= 1ull << (X * NPAIRCLASS + Y);
    build_work_lists(c->pairs, c->Q, c->opts.schwarz_tol, c->pbc ? 0.0 : cut, c->class_ids, c->ket_ids, c->bstart,
                     wl, c->pbc, split);
    for (int Y = 0; Y < NPAIRCLASS; ++Y) {
        if ((rc = dalloc(&c->d_ket_ids[Y], c->ket_ids[Y].size()))) return rc;
        if ((rc = dalloc(&c->d_bstart[Y], c->bstart[Y].size()))) return rc;
        if (!c->ket_ids[Y].empty())
            HIPCHK(hipMemcpy(c->d_ket_ids[Y], c->ket_ids[Y].data(), c->ket_ids[Y].size() * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->d_bstart[Y], c->bstart[Y].data(), c->bstart[Y].size() * sizeof(int), hipMemcpyHostToDevice));
    }
    {
        const size_t P = c->prims.size();
        std::vector<double> soa(8 * P, 0.0);
        size_t off = 0;
        for (int Y = 0; Y < NPAIRCLASS; ++Y) {
            const auto& kl = c->ket_ids[Y];
            const auto& bs = c->bstart[Y];
            std::vector<int> bo(std::max<size_t>(bs.size(), 1), 0);
            for (size_t b = 0; b + 1 < bs.size(); ++b) {
                const int nk = bs[b + 1] - bs[b];
                const int np_b = c->pairs[kl[bs[b]]].nprim;
                bo[b] = (int)off;
                for (int i = 0; i < nk; ++i) {
                    const PairInfo& pi = c->pairs[kl[bs[b] + i]];
                    for (int q = 0; q < np_b; ++q) {
                        const PrimPair& pp = c->prims[pi.prim_off + q];
                        const size_t at = off + (size_t)q * nk + i;
                        const double f[8] = {pp.zeta, pp.K, pp.Px, pp.Py, pp.Pz, pp.PAx, pp.PAy, pp.PAz};
                        for (int k = 0; k < 8; ++k) soa[k * P + at] = f[k];
                    }
                }
                off += (size_t)np_b * nk;
            }
            if ((rc = dalloc(&c->d_bsoa[Y], bo.size()))) return rc;
            HIPCHK(hipMemcpy(c->d_bsoa[Y], bo.data(), bo.size() * sizeof(int), hipMemcpyHostToDevice));
        }
        if (off != P) return fail(POLYERI_E_BASIS, "interleaved ket layout does not cover every primitive pair");
        if ((rc = dalloc(&c->d_ksoa, soa.size()))) return rc;
        HIPCHK(hipMemcpy(c->d_ksoa, soa.data(), soa.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    c->quartets = 0; c->prim_quartets = 0; c->model_flops = 0;
    for (const WorkList& l : wl) {
        ClassPairWork w;
        w.X = l.X; w.Y = l.Y;
        w.la = kPairL[l.X][0]; w.lb = kPairL[l.X][1]; w.lc = kPairL[l.Y][0]; w.ld = kPairL[l.Y][1];
        w.d_bra = c->d_class_ids[l.X]; w.d_ket = c->d_ket_ids[l.Y]; w.d_bstart = c->d_bstart[l.Y];
        w.d_bsoa = c->d_bsoa[l.Y];
        w.nbucket = l.nbucket;
        w.nrows = l.nrows;
        w.nwork = l.nwork;
        w.prec = l.prec;
        w.both = c->pbc && l.X != l.Y;
        w.d_rf = nullptr;
        w.d_items = nullptr;
        w.d_seg = nullptr;
        w.d_segq = nullptr;
        w.nseg = w.nsegc = 0;
        w.prim_quartets = l.prim_quartets;
        w.flops = l.prim_quartets * flop_model(w.la, w.lb, w.lc, w.ld);
        w.d_row = nullptr;
        w.d_rb = nullptr;
        if ((rc = dalloc(&w.d_row, l.row.size()))) return rc;
        if ((rc = dalloc(&w.d_rb, l.rb.size()))) return rc;
        HIPCHK(hipMemcpy(w.d_row, l.row.data(), l.row.size() * sizeof(long long), hipMemcpyHostToDevice));
        if (!l.rb.empty()) HIPCHK(hipMemcpy(w.d_rb, l.rb.data(), l.rb.size() * sizeof(int), hipMemcpyHostToDevice));
        // Segment lists for the FP64 lists of molecular contexts, and for the
        // FP32 lists of the Rys classes (their own kernel).  The light FP32
        // lists get none: the FP32 light-class segment kernel once stopped on
        // an illegal memory access ((H2O)4 cc-pVDZ, cut 1e-5) whose cause is
        // not established (a host replay of its index path over those lists
        // stays in bounds), so it is not launched, and mixed contexts do not
        // split the light classes.
        if (!c->pbc && l.nwork > 0 && (!l.prec || polyeri_rys_class(w.la, w.lb, w.lc, w.ld))) {
            std::vector<SegDesc> segs =
                make_segments(l, c->pairs, c->class_ids[l.X], c->ket_ids[l.Y], c->bstart[l.Y]);
            const std::vector<SegDesc> segc =
                make_segments(l, c->pairs, c->class_ids[l.X], c->ket_ids[l.Y], c->bstart[l.Y], false);
            if (segs.size() + segc.size() > 0x7fffffffu) return fail(POLYERI_E_ARG, "too many segments");
            w.nseg = (int)segs.size();
            w.nsegc = (int)segc.size();
            segs.insert(segs.end(), segc.begin(), segc.end());
            if ((rc = dalloc(&w.d_seg, segs.size()))) return rc;
            HIPCHK(hipMemcpy(w.d_seg, segs.data(), segs.size() * sizeof(SegDesc), hipMemcpyHostToDevice));
            const std::vector<double> sq = segment_ket_bounds(l, segs, c->ket_ids[l.Y], c->bstart[l.Y], c->Q);
            if ((rc = dalloc(&w.d_segq, sq.size()))) return rc;
            HIPCHK(hipMemcpy(w.d_segq, sq.data(), sq.size() * sizeof(double), hipMemcpyHostToDevice));
        }
        if (l.prec) {
            if ((rc = dalloc(&w.d_rf, l.rf.size()))) return rc;
            if (!l.rf.empty())
                HIPCHK(hipMemcpy(w.d_rf, l.rf.data(), l.rf.size() * sizeof(int), hipMemcpyHostToDevice));
        }
        c->quartets += w.nwork;
        c->prim_quartets += w.prim_quartets;
        c->model_flops += w.flops;
        c->work.push_back(w);
    }
    if (!c->pbc && (rc = build_gc_lists(c))) return rc;
    // density-screening tables
    {
        const int nsh = (int)c->ib.shells.size();
        std::vector<int2> psh(np);
        for (int i = 0; i < np; ++i) psh[i] = make_int2(c->pairs[i].a, c->pairs[i].b);
        std::vector<int> f0(nsh), nf(nsh);
        for (int i = 0; i < nsh; ++i) { f0[i] = c->ib.shells[i].f0; nf[i] = ncart(c->ib.shells[i].L); }
        if ((rc = dalloc(&c->d_psh, np)) || (rc = dalloc(&c->d_sh_f0, nsh)) || (rc = dalloc(&c->d_sh_nf, nsh)) ||
            (rc = dalloc(&c->d_dsh, (size_t)nsh * nsh)) || (rc = dalloc(&c->d_counts, c->work.size())) ||
            (rc = dalloc(&c->d_dmax, 1)))
            return rc;
        HIPCHK(hipMemcpy(c->d_psh, psh.data(), np * sizeof(int2), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->d_sh_f0, f0.data(), nsh * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(c->d_sh_nf, nf.data(), nsh * sizeof(int), hipMemcpyHostToDevice));
        c->h_counts.assign(c->work.size(), 0);
    }
    if (c->pbc) {
        const size_t ncell = 2 * (size_t)c->cel1 + 1, n2 = (size_t)nbf * nbf, c2 = (size_t)nc * nc;
        if ((rc = dalloc(&c->d_pcell, c->pcell.size())) || (rc = dalloc(&c->d_Dc, ncell * n2)) ||
            (rc = dalloc(&c->d_Jc, ncell * n2 * c->jk_copies)) || (rc = dalloc(&c->d_Kc, ncell * n2 * c->jk_copies)) ||
            (rc = dalloc(&c->d_pcc, ncell * c2)) || (rc = dalloc(&c->d_jcc, ncell * c2)) ||
            (rc = dalloc(&c->d_kcc, ncell * c2)) || (rc = dalloc(&c->d_fx_overflow, 1)))
            return rc;
        HIPCHK(hipMemcpy(c->d_pcell, c->pcell.data(), c->pcell.size() * sizeof(int2_host), hipMemcpyHostToDevice));
        HIPCHK(hipMemset(c->d_fx_overflow, 0, sizeof(unsigned)));
        // work counts: every (bra, ket) item runs for 2 cel2 + 1 cells Q3
        c->quartets = 0;
        for (ClassPairWork& w : c->work) c->quartets += w.nwork * (2LL * c->cel2 + 1);
    }
    c->ev.resize(c->work.size() + 1);
    for (auto& e : c->ev) HIPCHK(hipEventCreate(&e));
    c->last_class_ms.assign(c->work.size(), 0.0);
    // one-electron matrices in the reference layout
    std::vector<double> S, T, V;
    one_electron(c->ib, boys, S, T, V);
    c->S_C.assign((size_t)nc * nc, 0.0); c->T_C.assign((size_t)nc * nc, 0.0); c->N_C.assign((size_t)nc * nc, 0.0);
    for (int a = 0; a < nbf; ++a)
        for (int b = 0; b < nbf; ++b) {
            size_t r = c->ib.ref_of_int[a] + (size_t)nc * c->ib.ref_of_int[b];
            double f = c->ib.scale[a] * c->ib.scale[b];
            c->S_C[r] = f * S[(size_t)a * nbf + b];
            c->T_C[r] = f * T[(size_t)a * nbf + b];
            c->N_C[r] = f * V[(size_t)a * nbf + b];
        }
    c->e_nuc = nuclear_repulsion(c->ib);
    return 0;
}

static void default_opts(polyeri_opts* o) {
    o->schwarz_tol = 1e-12;
    o->prim_tol = 1e-20;
    o->precision = POLYERI_FP64;
    o->mixed_cut = 1e-5;
    o->device = 0;
}

int polyeri_ctx_create(int npgs, int ncgs, const int* pax, const int* pay, const int* paz,
                       const double* pgsx, const double* pgsy, const double* pgsz, const double* zt,
                       const double* cc, int natom, const int* iatom, const double* xyz,
                       const polyeri_opts* opts, polyeri_ctx** out) {
    if (!out || npgs <= 0 || ncgs <= 0 || !pax || !pay || !paz || !pgsx || !pgsy || !pgsz || !zt || !cc ||
        natom <= 0 || !iatom || !xyz)
        return fail(POLYERI_E_ARG, "polyeri_ctx_create: null or empty argument");
    *out = nullptr;
    polyeri_ctx* c = new polyeri_ctx();
    if (opts) c->opts = *opts; else default_opts(&c->opts);
    c->device = c->opts.device;
    RefBasis& rb = c->rb;
    rb.natom = natom;
    rb.iatom.assign(iatom, iatom + natom);
    for (int i = 0; i < natom; ++i) {
        rb.ax.push_back(xyz[3 * i]); rb.ay.push_back(xyz[3 * i + 1]); rb.az.push_back(xyz[3 * i + 2]);
    }
    rb.npgs = npgs; rb.ncgs = ncgs;
    rb.pax.assign(pax, pax + npgs); rb.pay.assign(pay, pay + npgs); rb.paz.assign(paz, paz + npgs);
    rb.pgsx.assign(pgsx, pgsx + npgs); rb.pgsy.assign(pgsy, pgsy + npgs); rb.pgsz.assign(pgsz, pgsz + npgs);
    rb.zt.assign(zt, zt + npgs);
    rb.cc.assign(cc, cc + (size_t)npgs * ncgs);
    rb.pgs_atom.assign(npgs, 0);
    int rc = build_ctx(c);
    if (rc) { free_ctx(c); return rc; }
    *out = c;
    return 0;
}

int polyeri_ctx_create_periodic(int npgs, int ncgs, const int* pax, const int* pay, const int* paz,
                                const double* pgsx, const double* pgsy, const double* pgsz, const double* zt,
                                const double* cc, int natom, const int* iatom, const double* xyz, double period,
                                int cel1, int cel2, const polyeri_opts* opts, polyeri_ctx** out) {
    if (!out || npgs <= 0 || ncgs <= 0 || !pax || !pay || !paz || !pgsx || !pgsy || !pgsz || !zt || !cc ||
        natom <= 0 || !iatom || !xyz)
        return fail(POLYERI_E_ARG, "polyeri_ctx_create_periodic: null or empty argument");
    if (!(period > 0.0) || cel1 < 0 || cel2 < 0 || cel1 > 64 || cel2 > 256)
        return fail(POLYERI_E_ARG, "polyeri_ctx_create_periodic: bad period or cutoffs");
    *out = nullptr;
    polyeri_ctx* c = new polyeri_ctx();
    if (opts) c->opts = *opts; else default_opts(&c->opts);
    c->device = c->opts.device;
    c->pbc = true;
    c->period = period;
    c->cel1 = cel1;
    c->cel2 = cel2;
    RefBasis& rb = c->rb;
    rb.natom = natom;
    rb.iatom.assign(iatom, iatom + natom);
    for (int i = 0; i < natom; ++i) {
        rb.ax.push_back(xyz[3 * i]); rb.ay.push_back(xyz[3 * i + 1]); rb.az.push_back(xyz[3 * i + 2]);
    }
    rb.npgs = npgs; rb.ncgs = ncgs;
    rb.pax.assign(pax, pax + npgs); rb.pay.assign(pay, pay + npgs); rb.paz.assign(paz, paz + npgs);
    rb.pgsx.assign(pgsx, pgsx + npgs); rb.pgsy.assign(pgsy, pgsy + npgs); rb.pgsz.assign(pgsz, pgsz + npgs);
    rb.zt.assign(zt, zt + npgs);
    rb.cc.assign(cc, cc + (size_t)npgs * ncgs);
    rb.pgs_atom.assign(npgs, 0);
    int rc = build_ctx(c);
    if (rc) { free_ctx(c); return rc; }
    *out = c;
    return 0;
}

int polyeri_rys_nodes(int n, double T, double* u, double* w) {
    if (n < 1 || n > RYS_NMAX_EXPORT || !(T >= 0.0) || !u || !w)
        return fail(POLYERI_E_ARG, "polyeri_rys_nodes: n must be 1..5, T >= 0 and u, w non-null");
    return polyeri_rys_nodes_host(n, T, u, w) ? fail(POLYERI_E_ARG, "polyeri_rys_nodes: bad argument") : 0;
}

int polyeri_rys_block(int la, int lb, int lc, int ld, const double* xyz, const double* ex, double* out) {
    const int rc = polyeri_rys_block_host(la, lb, lc, ld, xyz, ex, out);
    if (rc == -1) return fail(POLYERI_E_ARG, "polyeri_rys_block: not a Rys class, or a null pointer");
    if (rc) return fail(POLYERI_E_ARG, "polyeri_rys_block: the lane variants do not cover the block exactly once");
    return 0;
}

int polyeri_basis_read(const char* path, int natom, const int* iatom, const double* xyz, int* npgs,
                       int* ncgs, int* npshell, int* pax, int* pay, int* paz, int* pgs_atom,
                       double* pgsx, double* pgsy, double* pgsz, double* zt, double* cc, double* cgsx,
                       double* cgsy, double* cgsz, int* p_shell, int* p_shell_ang) {
    if (!path || natom <= 0 || !iatom || !xyz || !npgs || !ncgs || !npshell)
        return fail(POLYERI_E_ARG, "polyeri_basis_read: bad argument");
    RefBasis rb;
    std::string err;
    if (read_basis(path, natom, iatom, xyz, rb, err)) return fail(POLYERI_E_BASIS, err);
    *npgs = rb.npgs; *ncgs = rb.ncgs; *npshell = rb.npshell;
    auto cpi = [](int* d, const std::vector<int>& s) { if (d) std::copy(s.begin(), s.end(), d); };
    auto cpd = [](double* d, const std::vector<double>& s) { if (d) std::copy(s.begin(), s.end(), d); };
    cpi(pax, rb.pax); cpi(pay, rb.pay); cpi(paz, rb.paz); cpi(pgs_atom, rb.pgs_atom);
    cpd(pgsx, rb.pgsx); cpd(pgsy, rb.pgsy); cpd(pgsz, rb.pgsz); cpd(zt, rb.zt); cpd(cc, rb.cc);
    cpd(cgsx, rb.cgsx); cpd(cgsy, rb.cgsy); cpd(cgsz, rb.cgsz);
    cpi(p_shell, rb.p_shell); cpi(p_shell_ang, rb.p_shell_ang);
    return 0;
}

// Per-list launch shapes chosen by timing: the register kernel with 256,
// 128 or 64 threads per block (one pass, or a persistent grid of 2048 blocks
// striding over the list) and, for classes of 36 or more integrals, the
// cooperative kernel with grids of 2048, 4096 or 8192 wavefronts.  Each
// candidate runs once untimed and reps times between events on the
// context's density buffer (zeroed: the cost does not depend on values).
// Event timings are diagnostics: where the runtime cannot time a pair of
// events (under a counter-collecting profiler, for instance) the time reads
// NaN and the pending error is cleared instead of failing the build.
static double event_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) {
        (void)hipGetLastError();
        return std::nan("");
    }
    return ms;
}

int polyeri_autotune(polyeri_ctx* ctx, int reps) {
    if (!ctx) return fail(POLYERI_E_ARG, "polyeri_autotune: null context");
    if (ctx->pbc) return fail(POLYERI_E_ARG, "polyeri_autotune: periodic context");
    if (reps < 1) reps = 1;
    HIPCHK(hipSetDevice(ctx->device));
    polyeri_ctx* c = ctx;
    size_t n2 = (size_t)c->ib.nbf * c->ib.nbf;
    hipLaunchKernelGGL(k_fill, dim3(nblk(n2)), dim3(256), 0, c->stream, c->d_D, n2, 0.0);
    HIPCHK(hipGetLastError());
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    int rc = 0;
    for (ClassPairWork& cw : c->work) {
        if (cw.nwork <= 0) continue;
        const int mode = cw.prec ? MODE_JK32 : MODE_JK;
        const int nint = ncart_h(cw.la) * ncart_h(cw.lb) * ncart_h(cw.lc) * ncart_h(cw.ld);
        std::vector<LaunchCfg> cand;
        const bool reg = polyeri_class_kernel(cw.la, cw.lb, cw.lc, cw.ld, mode) != nullptr;
        if (reg) {
            for (int b : {256, 128, 64}) { LaunchCfg x; x.kind = LAUNCH_REG; x.block = b; cand.push_back(x); }
            LaunchCfg x; x.kind = LAUNCH_REG; x.block = 256; x.grid = 2048; cand.push_back(x);
            if (cw.nseg > 0 && seg_lds_bytes(cw.la, cw.lb, c->ib.nbf) <= 65536)
                for (int v : {0, 1})
                    for (int b : {256, 128, 64}) {
                        LaunchCfg y; y.kind = LAUNCH_SEG; y.block = b; y.segv = v; cand.push_back(y);
                    }
        }
        if (cw.nseg > 0 && polyeri_rys_class(cw.la, cw.lb, cw.lc, cw.ld) &&
            polyeri_rys_lds_bytes(cw.la, cw.lb, c->ib.nbf) <= 65536)
            for (int v : {0, 1})
                for (int b : {256, 128, 64}) {
                    LaunchCfg y; y.kind = LAUNCH_RYS; y.block = b; y.segv = v; cand.push_back(y);
                }
        if (!reg || nint >= 36)
            for (int g : {2048, 4096, 8192})
                for (int pbs : {0, 1, 2, 3}) {
                    LaunchCfg x; x.kind = LAUNCH_COOP; x.grid = g; x.pbs = pbs; cand.push_back(x);
                }
        double best = 1e300;
        LaunchCfg pick = cw.cfg;
        for (const LaunchCfg& x : cand) {
            for (int r = 0; r <= reps && !rc; ++r) {
                if (r == 1) HIPCHK(hipEventRecord(e0, c->stream));
                rc = launch_class(c, mode, cw.la, cw.lb, cw.lc, cw.ld, cw.d_bra, cw.d_ket, cw.d_row, cw.nrows,
                                  cw.nwork, nullptr, cw.d_rb, cw.d_bstart, cw.nbucket, cw.d_rf, nullptr, nullptr, x,
                                  cw.d_bsoa, x.segv ? cw.d_seg + cw.nseg : cw.d_seg, x.segv ? cw.nsegc : cw.nseg);
            }
            if (rc) break;
            HIPCHK(hipEventRecord(e1, c->stream));
            HIPCHK(hipEventSynchronize(e1));
            const double ms = event_ms(e0, e1);     // NaN where untimable: keeps the current pick
            if (ms < best) { best = ms; pick = x; }
        }
        if (rc) break;
        cw.cfg = pick;
        // the generally contracted form, kept where it beats the best plain launch
        if (cw.gc.nwork > 0) {
            double gbest = 1e300;
            int gblock = cw.gc_block, gsegv = cw.gc_segv;
            for (int v : {0, 1})
                for (int blk : {256, 128, 64}) {
                    cw.gc_segv = v;
                    for (int r = 0; r <= reps && !rc; ++r) {
                        if (r == 1) HIPCHK(hipEventRecord(e0, c->stream));
                        rc = launch_gc(c, cw, blk);
                    }
                    if (rc) break;
                    HIPCHK(hipEventRecord(e1, c->stream));
                    HIPCHK(hipEventSynchronize(e1));
                    const double ms = event_ms(e0, e1);
                    if (ms < gbest) { gbest = ms; gblock = blk; gsegv = v; }
                }
            cw.gc_segv = gsegv;
            if (rc) break;
            if (gbest < best) { cw.gc_on = 1; cw.gc_block = gblock; }
            else if (best < 1e300) cw.gc_on = 0;
            if (gbest < best) best = gbest;
        }
        cw.est_ms = best < 1e300 ? best / std::max(reps, 1) : 0.0;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    drop_graphs(c);
    c->tuned = true;
    return 0;
}

int polyeri_set_graph(polyeri_ctx* ctx, int on) {
    if (!ctx) return fail(POLYERI_E_ARG, "null context");
    ctx->use_graph = on != 0;
    if (!ctx->use_graph) drop_graphs(ctx);
    return 0;
}

int polyeri_set_general_contraction(polyeri_ctx* ctx, int on) {
    if (!ctx) return fail(POLYERI_E_ARG, "null context");
    for (ClassPairWork& w : ctx->work) w.gc_on = on && w.gc.nwork > 0;
    drop_graphs(ctx);
    return 0;
}

// Launch configuration of work list `list` as 9 ints: kind (0 auto, 1
// register, 2 wavefront, 3 segment, 5 Rys), block, grid, segment variant,
// primitive batch (-1 default), generally contracted form on, its block,
// its segment variant, and the autotuned time in microseconds (0 if none).
int polyeri_get_launch_config(const polyeri_ctx* ctx, int list, int* cfg) {
    if (!ctx || !cfg || list < 0 || list >= (int)ctx->work.size())
        return fail(POLYERI_E_ARG, "polyeri_get_launch_config: bad context, list or pointer");
    const ClassPairWork& w = ctx->work[list];
    const int v[9] = {w.cfg.kind, w.cfg.block, w.cfg.grid, w.cfg.segv, w.cfg.pbs, w.gc_on, w.gc_block, w.gc_segv,
                      (int)std::lround(w.est_ms * 1e3)};
    std::copy(v, v + 9, cfg);
    return 0;
}

// Sets what polyeri_get_launch_config returns (e.g. the autotuner's picks
// of an earlier process); rejects kinds that the list cannot run.
int polyeri_set_launch_config(polyeri_ctx* ctx, int list, const int* cfg) {
    if (!ctx || !cfg || list < 0 || list >= (int)ctx->work.size())
        return fail(POLYERI_E_ARG, "polyeri_set_launch_config: bad context, list or pointer");
    ClassPairWork& w = ctx->work[list];
    const int kind = cfg[0];
    if (kind != LAUNCH_AUTO && kind != LAUNCH_REG && kind != LAUNCH_COOP && kind != LAUNCH_SEG && kind != LAUNCH_RYS)
        return fail(POLYERI_E_ARG, "polyeri_set_launch_config: unknown launch kind");
    const int mode = w.prec ? MODE_JK32 : MODE_JK;
    const bool reg = polyeri_class_kernel(w.la, w.lb, w.lc, w.ld, mode) != nullptr;
    if ((kind == LAUNCH_REG && !reg) || (kind == LAUNCH_SEG && (!reg || w.nseg <= 0 || w.prec)) ||
        (kind == LAUNCH_RYS && (w.nseg <= 0 || !polyeri_rys_class(w.la, w.lb, w.lc, w.ld))) ||
        (cfg[5] && w.gc.nwork <= 0))
        return fail(POLYERI_E_ARG, "polyeri_set_launch_config: the list cannot run this launch kind");
    w.cfg.kind = kind;
    w.cfg.block = cfg[1];
    w.cfg.grid = std::max(0, cfg[2]);
    w.cfg.segv = cfg[3] ? 1 : 0;
    w.cfg.pbs = std::max(-1, std::min(3, cfg[4]));
    w.gc_on = cfg[5] ? 1 : 0;
    w.gc_block = cfg[6];
    w.gc_segv = cfg[7] ? 1 : 0;
    w.est_ms = std::max(0, cfg[8]) * 1e-3;
    drop_graphs(ctx);
    return 0;
}

int polyeri_set_profiling(polyeri_ctx* ctx, int on) {
    if (!ctx) return fail(POLYERI_E_ARG, "null context");
    ctx->profiling = on != 0;
    return 0;
}

// J/K from the context's device density (reference layout, ncgs^2).
// Periodic: C_C(ref i, ref j, Q) = s_i s_j J(Q)[i][j] summed over the per-XCD
// copies (no symmetrisation: the lattice sums run over every quartet).  The
// copies hold fixed-point integers (jk_add_fx): they are summed as integers
// (wrapping, so intermediate overflow cancels) and converted once, so the
// result does not depend on the order of the atomics.  A sum beyond the
// fixed-point range is reported through *overflow.
__global__ void k_unpack_cells(const double* __restrict__ J, const double* __restrict__ K, int nbf, int ncell,
                               int copies, const int* __restrict__ ref, const double* __restrict__ sc, int nc,
                               double* __restrict__ jc, double* __restrict__ kc, unsigned* __restrict__ overflow) {
    const size_t n2 = (size_t)nbf * nbf;
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n2 * ncell) return;
    const int cell = (int)(t / n2);
    const size_t e = t % n2;
    const int mu = (int)(e / nbf), nu = (int)(e % nbf);
    const unsigned long long* Ji = reinterpret_cast<const unsigned long long*>(J);
    const unsigned long long* Ki = reinterpret_cast<const unsigned long long*>(K);
    unsigned long long js = 0, ks = 0;
    for (int x = 0; x < copies; ++x) {
        js += Ji[((size_t)x * ncell + cell) * n2 + e];
        ks += Ki[((size_t)x * ncell + cell) * n2 + e];
    }
    const double jv = (double)(long long)js * 0x1p-50, kv = (double)(long long)ks * 0x1p-50;
    // a sum near the edge of the +-2^13 range has most likely wrapped
    if (fabs(jv) > 0x1p12 || fabs(kv) > 0x1p12) *overflow = 1u;
    const size_t r = (size_t)cell * nc * nc + ref[mu] + (size_t)nc * ref[nu];
    jc[r] = sc[mu] * sc[nu] * jv;
    kc[r] = sc[mu] * sc[nu] * kv;
}

__global__ void k_fill_u32(unsigned* p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// The kernel sequence of one J/K build of the contracted density d_pc into
// d_jc / d_kc (reference layout): density packing, zeroing (fill kernels,
// so the sequence can be captured in a graph), the optional density
// screening, the class kernels, and the unpacking.  dens_tol > 0 drops the
// quartets whose density-weighted Schwarz bound is below dens_tol
// (kscreen.hip); the surviving items of every list are compacted on the
// device and the class kernels stride over them.
static int jk_enqueue(polyeri_ctx* c, const double* d_pc, double* d_jc, double* d_kc, double dens_tol,
                      bool capturing = false) {
    int nbf = c->ib.nbf, nc = c->rb.ncgs;
    size_t n2 = (size_t)nbf * nbf, c2 = (size_t)nc * nc;
    const bool screen = dens_tol > 0.0;
    hipLaunchKernelGGL(k_pack_density, dim3(nblk(n2)), dim3(256), 0, c->stream, d_pc, nc, c->d_ref,
                       c->d_scale, nbf, c->d_D);
    hipLaunchKernelGGL(k_fill, dim3(nblk(n2 * c->jk_copies)), dim3(256), 0, c->stream, c->d_J, n2 * c->jk_copies, 0.0);
    hipLaunchKernelGGL(k_fill, dim3(nblk(n2 * c->jk_copies)), dim3(256), 0, c->stream, c->d_K, n2 * c->jk_copies, 0.0);
    HIPCHK(hipGetLastError());
    if (screen) {
        const int nsh = (int)c->ib.shells.size();
        hipLaunchKernelGGL(k_fill_u32, dim3(nblk(c->work.size())), dim3(256), 0, c->stream, c->d_counts, c->work.size());
        if (polyeri_shell_dmax(c->d_D, nbf, c->d_sh_f0, c->d_sh_nf, nsh, c->d_dsh, c->stream) ||
            polyeri_max_reduce(c->d_dsh, (size_t)nsh * nsh, c->d_dmax, c->stream))
            return fail(POLYERI_E_HIP, "shell density maxima launch failed");
        for (size_t w = 0; w < c->work.size(); ++w) {
            const ClassPairWork& cw = c->work[w];
            if (cw.nwork <= 0 || screens_in_kernel(c, cw)) continue;   // these screen in the kernel
            EriArgs a{};
            a.bra_ids = cw.d_bra; a.ket_ids = cw.d_ket; a.row_off = cw.d_row; a.nrows = cw.nrows; a.nwork = cw.nwork;
            a.rb_off = cw.d_rb; a.bucket_start = cw.d_bstart; a.nbucket = cw.nbucket; a.rb_first = cw.d_rf;
            if (polyeri_screen_list(a, c->d_Q, c->d_psh, c->d_dsh, nsh, dens_tol, cw.d_items, c->d_counts + w,
                                    c->stream))
                return fail(POLYERI_E_HIP, "screening launch failed");
        }
    }
    // Branches: the class launches only add into J/K (the additions commute),
    // so a captured build runs them on kBranches streams, the lists assigned
    // largest first to the least loaded branch (autotuned times, else model
    // FLOPs), and joins them before unpacking.  Profiled builds stay on one
    // stream (per-list events between the launches).
    const bool fork = capturing && !c->profiling && c->work.size() > 1;
    std::vector<int> branch(c->work.size(), 0);
    hipStream_t main = c->stream;
    if (fork) {
        std::vector<int> order(c->work.size());
        for (size_t w = 0; w < order.size(); ++w) order[w] = (int)w;
        const bool timed = std::all_of(c->work.begin(), c->work.end(),
                                       [](const ClassPairWork& x) { return x.nwork <= 0 || x.est_ms > 0; });
        auto cost = [&](int w) { return timed ? c->work[w].est_ms : c->work[w].flops; };
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost(a) > cost(b); });
        double load[kBranches] = {};
        for (int w : order) {
            int bmin = 0;
            for (int i = 1; i < kBranches; ++i)
                if (load[i] < load[bmin]) bmin = i;
            branch[w] = bmin;
            load[bmin] += cost(w);
        }
        HIPCHK(hipEventRecord(c->ev_fork, main));
        for (int i = 0; i < kBranches - 1; ++i) HIPCHK(hipStreamWaitEvent(c->side[i], c->ev_fork, 0));
    }
    int rc = 0;
    for (size_t w = 0; w < c->work.size() && !rc; ++w) {
        const ClassPairWork& cw = c->work[w];
        if (c->profiling) HIPCHK(hipEventRecord(c->ev[w], c->stream));
        c->stream = branch[w] ? c->side[branch[w] - 1] : main;   // every launch below goes to c->stream
        const SegScreen scr{c->d_Q, c->d_psh, c->d_dsh, (int)c->ib.shells.size(), dens_tol, c->d_counts + w,
                            cw.d_segq ? cw.d_segq + (cw.cfg.segv ? cw.nseg : 0) : nullptr, c->d_dmax};
        const bool compact = screen && !screens_in_kernel(c, cw);
        if (list_uses_gc(cw) && !compact) {
            rc = launch_gc(c, cw, cw.gc_block, screen ? &scr : nullptr);
        } else {
            const bool seg = list_uses_seg(c, cw) && !compact;
            rc = launch_class(c, cw.prec ? MODE_JK32 : MODE_JK, cw.la, cw.lb, cw.lc, cw.ld, cw.d_bra, cw.d_ket,
                              cw.d_row, cw.nrows, cw.nwork, nullptr, cw.d_rb, cw.d_bstart, cw.nbucket, cw.d_rf,
                              screen && !seg ? cw.d_items : nullptr, screen && !seg ? c->d_counts + w : nullptr,
                              cw.cfg, cw.d_bsoa, cw.cfg.segv ? cw.d_seg + cw.nseg : cw.d_seg,
                              cw.cfg.segv ? cw.nsegc : cw.nseg, screen && seg ? &scr : nullptr);
        }
        c->stream = main;
    }
    if (fork)   // join every branch, also after a failed launch (capture must end joined)
        for (int i = 0; i < kBranches - 1; ++i) {
            HIPCHK(hipEventRecord(c->ev_join[i], c->side[i]));
            HIPCHK(hipStreamWaitEvent(main, c->ev_join[i], 0));
        }
    if (rc) return rc;
    if (c->profiling) HIPCHK(hipEventRecord(c->ev[c->work.size()], c->stream));
    hipLaunchKernelGGL(k_fill, dim3(nblk(c2)), dim3(256), 0, c->stream, d_jc, c2, 0.0);
    hipLaunchKernelGGL(k_fill, dim3(nblk(c2)), dim3(256), 0, c->stream, d_kc, c2, 0.0);
    hipLaunchKernelGGL(k_unpack_jk, dim3(nblk(n2)), dim3(256), 0, c->stream, c->d_J, c->d_K, nbf, c->jk_copies,
                       c->d_ref, c->d_scale, nc, d_jc, d_kc);
    HIPCHK(hipGetLastError());
    return 0;
}

// One J/K build.  Without per-class profiling the kernel sequence is
// captured once per (density, J, K, screening) combination into a HIP graph
// and replayed; the timing events stay outside the graph.
// J/K from the stored records (DIRECT .F.) into the reference-layout
// blocks d_jc / d_kc (ncell blocks: 1 for a molecule).
static int jk_stored(polyeri_ctx* c, const double* d_pc, double* d_jc, double* d_kc) {
    const int nc = c->rb.ncgs, ncell = c->pbc ? 2 * c->cel1 + 1 : 1;
    const size_t c2 = (size_t)nc * nc * ncell;
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    hipLaunchKernelGGL(k_fill, dim3(nblk(c2)), dim3(256), 0, c->stream, d_jc, c2, 0.0);
    hipLaunchKernelGGL(k_fill, dim3(nblk(c2)), dim3(256), 0, c->stream, d_kc, c2, 0.0);
    HIPCHK(hipGetLastError());
    if (polyeri_jk_stored_launch(c->d_ao, c->n_ao, nc, c->pbc ? c->cel1 : 0, c->pbc ? c->cel2 : 0, d_pc, d_jc, d_kc,
                                 c->stream))
        return fail(POLYERI_E_HIP, "stored-integral J/K launch failed");
    c->screened_last = false;
    c->stored_last = true;
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    return 0;
}

static int jk_device(polyeri_ctx* c, const double* d_pc, double* d_jc, double* d_kc, double dens_tol = 0.0) {
    if (c->pbc) return fail(POLYERI_E_ARG, "periodic context: use polyeri_jk_build_cells");
    if (c->use_stored) return jk_stored(c, d_pc, d_jc, d_kc);   // every stored record, no density screening
    c->stored_last = false;
    const bool screen = dens_tol > 0.0;
    if (screen)
        for (ClassPairWork& w : c->work)
            if (!w.d_items && w.nwork > 0 && !screens_in_kernel(c, w)) {   // the others keep no items
                int rc = dalloc(&w.d_items, (size_t)w.nwork);
                if (rc) return rc;
            }
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    if (c->use_graph && !c->profiling) {
        hipGraphExec_t exec = nullptr;
        for (auto& g : c->graphs)
            if (g.pc == d_pc && g.jc == d_jc && g.kc == d_kc && g.tol == dens_tol) exec = g.exec;
        if (!exec) {
            if (c->graphs.size() >= 8) {
                (void)hipGraphExecDestroy(c->graphs.front().exec);
                c->graphs.erase(c->graphs.begin());
            }
            HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            int rc = jk_enqueue(c, d_pc, d_jc, d_kc, dens_tol, true);
            hipGraph_t graph = nullptr;
            hipError_t e = hipStreamEndCapture(c->stream, &graph);
            if (rc) { if (graph) (void)hipGraphDestroy(graph); return rc; }
            if (e != hipSuccess) return fail(POLYERI_E_HIP, std::string("graph capture: ") + hipGetErrorString(e));
            e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess) return fail(POLYERI_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
            c->graphs.push_back({d_pc, d_jc, d_kc, dens_tol, exec});
        }
        HIPCHK(hipGraphLaunch(exec, c->stream));
    } else {
        int rc = jk_enqueue(c, d_pc, d_jc, d_kc, dens_tol);
        if (rc) return rc;
    }
    if (screen && !c->work.empty())
        HIPCHK(hipMemcpyAsync(c->h_counts.data(), c->d_counts, c->work.size() * sizeof(unsigned),
                              hipMemcpyDeviceToHost, c->stream));
    c->screened_last = screen;
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    return 0;
}

static int jk_collect_timing(polyeri_ctx* c) {
    HIPCHK(hipEventSynchronize(c->ev1));
    c->last_jk_ms = event_ms(c->ev0, c->ev1);
    c->last_quartets = 0;
    if (c->stored_last) {   // records digested (no class launches)
        c->last_quartets = c->n_ao;
        return 0;
    }
    for (size_t w = 0; w < c->work.size(); ++w)
        c->last_quartets += c->screened_last ? (long long)c->h_counts[w] : c->work[w].nwork;
    if (c->profiling)
        for (size_t w = 0; w < c->work.size(); ++w) c->last_class_ms[w] = event_ms(c->ev[w], c->ev[w + 1]);
    return 0;
}

int polyeri_jk_build_device(polyeri_ctx* ctx, const double* d_p_c, double* d_c_c, double* d_x_c) {
    if (!ctx || !d_p_c || !d_c_c || !d_x_c) return fail(POLYERI_E_ARG, "polyeri_jk_build_device: null argument");
    HIPCHK(hipSetDevice(ctx->device));
    return jk_device(ctx, d_p_c, d_c_c, d_x_c);
}

// The same build with the density staging and the J/K accumulation copies
// in the fixed workspace `scratch` (8 copies while 17 nbf^2 doubles fit,
// else one copy with device-scope atomics), eager launches (no graph), no
// allocation.  Every call on a device shares that device's workspace:
// calls that overlap on one device race.
int polyeri_jk_build_device_noalloc(polyeri_ctx* ctx, const double* d_p_c, double* d_c_c, double* d_x_c) {
    if (!ctx || !d_p_c || !d_c_c || !d_x_c) return fail(POLYERI_E_ARG, "polyeri_jk_build_device_noalloc: null argument");
    if (ctx->pbc) return fail(POLYERI_E_ARG, "polyeri_jk_build_device_noalloc: periodic context");
    const size_t n2 = (size_t)ctx->ib.nbf * ctx->ib.nbf;
    const int copies = 17 * n2 <= kNoAllocDoubles ? 8 : 3 * n2 <= kNoAllocDoubles ? 1 : 0;
    if (!copies)
        return fail(POLYERI_E_ARG, "polyeri_jk_build_device_noalloc: 3 nbf^2 doubles exceed the fixed workspace of "
                                   "2^22 doubles (nbf <= 1182)");
    HIPCHK(hipSetDevice(ctx->device));
    if (!ctx->d_noalloc) {   // first call on this context: the device's workspace address (once per device)
        int rc = noalloc_scratch(ctx->device, &ctx->d_noalloc);
        if (rc) return rc;
    }
    if (ctx->use_stored) return jk_stored(ctx, d_p_c, d_c_c, d_x_c);   // uses no workspace
    ctx->stored_last = false;
    double *sd = ctx->d_D, *sj = ctx->d_J, *sk = ctx->d_K;
    const int sc = ctx->jk_copies;
    ctx->d_D = ctx->d_noalloc;
    ctx->d_J = ctx->d_noalloc + n2;
    ctx->d_K = ctx->d_noalloc + n2 + copies * n2;
    ctx->jk_copies = copies;
    hipError_t e = hipEventRecord(ctx->ev0, ctx->stream);
    int rc = e == hipSuccess ? jk_enqueue(ctx, d_p_c, d_c_c, d_x_c, 0.0) : fail(POLYERI_E_HIP, hipGetErrorString(e));
    ctx->d_D = sd; ctx->d_J = sj; ctx->d_K = sk; ctx->jk_copies = sc;   // launches took their arguments by value
    if (rc) return rc;
    ctx->screened_last = false;
    HIPCHK(hipEventRecord(ctx->ev1, ctx->stream));
    return 0;
}

int polyeri_sync(polyeri_ctx* ctx) {
    if (!ctx) return fail(POLYERI_E_ARG, "null context");
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return jk_collect_timing(ctx);
}

static int jk_build_host(polyeri_ctx* ctx, const double* p_c, double dens_tol, double* c_c, double* x_c) {
    HIPCHK(hipSetDevice(ctx->device));
    size_t c2 = (size_t)ctx->rb.ncgs * ctx->rb.ncgs;
    HIPCHK(hipMemcpyAsync(ctx->d_pc, p_c, c2 * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    int rc = jk_device(ctx, ctx->d_pc, ctx->d_jc, ctx->d_kc, dens_tol);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(c_c, ctx->d_jc, c2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(x_c, ctx->d_kc, c2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if ((rc = jk_collect_timing(ctx))) return rc;
    for (size_t i = 0; i < c2; ++i)
        if (!std::isfinite(c_c[i]) || !std::isfinite(x_c[i])) return fail(POLYERI_E_NONFINITE, "non-finite J/K");
    return 0;
}

// Periodic J/K of the contracted density cells (reference layout, cell Q at
// block Q + cel1) into the J/K cells.
static int jk_device_cells(polyeri_ctx* c, const double* d_pc, double* d_jc, double* d_kc) {
    if (c->use_stored) return jk_stored(c, d_pc, d_jc, d_kc);
    c->stored_last = false;
    const int nbf = c->ib.nbf, nc = c->rb.ncgs, ncell = 2 * c->cel1 + 1;
    const size_t n2 = (size_t)nbf * nbf, c2 = (size_t)nc * nc;
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    for (int q = 0; q < ncell; ++q)
        hipLaunchKernelGGL(k_pack_density, dim3(nblk(n2)), dim3(256), 0, c->stream, d_pc + q * c2, nc, c->d_ref,
                           c->d_scale, nbf, c->d_Dc + q * n2);
    const size_t nj = n2 * ncell * c->jk_copies;
    hipLaunchKernelGGL(k_fill, dim3(nblk(nj)), dim3(256), 0, c->stream, c->d_Jc, nj, 0.0);
    hipLaunchKernelGGL(k_fill, dim3(nblk(nj)), dim3(256), 0, c->stream, c->d_Kc, nj, 0.0);
    HIPCHK(hipGetLastError());
    const long long n3 = 2LL * c->cel2 + 1;
    for (size_t w = 0; w < c->work.size(); ++w) {
        const ClassPairWork& cw = c->work[w];
        if (c->profiling) HIPCHK(hipEventRecord(c->ev[w], c->stream));
        int rc = launch_class_pbc(c, cw, cw.nwork * n3);
        if (rc) return rc;
    }
    if (c->profiling) HIPCHK(hipEventRecord(c->ev[c->work.size()], c->stream));
    hipLaunchKernelGGL(k_fill, dim3(nblk(c2 * ncell)), dim3(256), 0, c->stream, d_jc, c2 * ncell, 0.0);
    hipLaunchKernelGGL(k_fill, dim3(nblk(c2 * ncell)), dim3(256), 0, c->stream, d_kc, c2 * ncell, 0.0);
    hipLaunchKernelGGL(k_unpack_cells, dim3(nblk(n2 * ncell)), dim3(256), 0, c->stream, c->d_Jc, c->d_Kc, nbf, ncell,
                       c->jk_copies, c->d_ref, c->d_scale, nc, d_jc, d_kc, c->d_fx_overflow);
    HIPCHK(hipGetLastError());
    c->screened_last = false;
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    return 0;
}

// Fails when a periodic J/K cell sum has left the fixed-point range
// (|element| >= 2^12 Eh); the flag is cleared for the next build.
static int check_fx_overflow(polyeri_ctx* c) {
    unsigned f = 0;
    HIPCHK(hipMemcpy(&f, c->d_fx_overflow, sizeof(unsigned), hipMemcpyDeviceToHost));
    if (!f) return 0;
    HIPCHK(hipMemset(c->d_fx_overflow, 0, sizeof(unsigned)));
    return fail(POLYERI_E_NONFINITE, "periodic J/K element beyond the fixed-point accumulation range");
}

int polyeri_jk_build_cells(polyeri_ctx* ctx, const double* p_c, double* c_c, double* x_c) {
    if (!ctx || !p_c || !c_c || !x_c) return fail(POLYERI_E_ARG, "polyeri_jk_build_cells: null argument");
    if (!ctx->pbc) return fail(POLYERI_E_ARG, "polyeri_jk_build_cells: not a periodic context");
    HIPCHK(hipSetDevice(ctx->device));
    const size_t n = (size_t)ctx->rb.ncgs * ctx->rb.ncgs * (2 * ctx->cel1 + 1);
    HIPCHK(hipMemcpyAsync(ctx->d_pcc, p_c, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    int rc = jk_device_cells(ctx, ctx->d_pcc, ctx->d_jcc, ctx->d_kcc);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(c_c, ctx->d_jcc, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(x_c, ctx->d_kcc, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if ((rc = jk_collect_timing(ctx))) return rc;
    if ((rc = check_fx_overflow(ctx))) return rc;
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(c_c[i]) || !std::isfinite(x_c[i])) return fail(POLYERI_E_NONFINITE, "non-finite J/K");
    return 0;
}

int polyeri_jk_build(polyeri_ctx* ctx, const double* p_c, double* c_c, double* x_c) {
    if (!ctx || !p_c || !c_c || !x_c) return fail(POLYERI_E_ARG, "polyeri_jk_build: null argument");
    return jk_build_host(ctx, p_c, 0.0, c_c, x_c);
}

int polyeri_jk_build_screened(polyeri_ctx* ctx, const double* p_c, double dens_tol, double* c_c, double* x_c) {
    if (!ctx || !p_c || !c_c || !x_c) return fail(POLYERI_E_ARG, "polyeri_jk_build_screened: null argument");
    if (!(dens_tol >= 0.0)) return fail(POLYERI_E_ARG, "polyeri_jk_build_screened: dens_tol < 0");
    return jk_build_host(ctx, p_c, dens_tol, c_c, x_c);
}

int polyeri_device_alloc(polyeri_ctx* ctx, size_t bytes, void** ptr) {
    if (!ctx || !ptr) return fail(POLYERI_E_ARG, "polyeri_device_alloc: null argument");
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipMalloc(ptr, bytes ? bytes : 1));
    return 0;
}

int polyeri_device_free(polyeri_ctx* ctx, void* ptr) {
    if (!ctx) return fail(POLYERI_E_ARG, "polyeri_device_free: null context");
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    HIPCHK(hipFree(ptr));
    return 0;
}

int polyeri_memcpy(polyeri_ctx* ctx, void* dst, const void* src, size_t bytes, int to_device) {
    if (!ctx || !dst || !src) return fail(POLYERI_E_ARG, "polyeri_memcpy: null argument");
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipMemcpyAsync(dst, src, bytes, to_device ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int polyeri_device_copy(polyeri_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!ctx || !dst || !src || bytes % 16) return fail(POLYERI_E_ARG, "polyeri_device_copy: bad argument");
    HIPCHK(hipSetDevice(ctx->device));
    size_t n = bytes / 16;
    hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, ctx->stream, (const double2*)src, (double2*)dst, n);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int polyeri_fock_build_device(polyeri_ctx* ctx, const double* d_p_c, double hfexchange, double* d_f_c,
                              double* d_c_c, double* d_x_c) {
    if (!ctx || !d_p_c || !d_f_c || !d_c_c || !d_x_c) return fail(POLYERI_E_ARG, "polyeri_fock_build_device: null argument");
    HIPCHK(hipSetDevice(ctx->device));
    int rc = ensure_onee_device(ctx);
    if (rc) return rc;
    if ((rc = jk_device(ctx, d_p_c, d_c_c, d_x_c))) return rc;
    int nc = ctx->rb.ncgs;
    hipLaunchKernelGGL(k_fock, dim3(nblk((size_t)nc * nc)), dim3(256), 0, ctx->stream, ctx->d_T, ctx->d_V, d_c_c, d_x_c,
                       hfexchange, nc, d_f_c);
    HIPCHK(hipGetLastError());
    return 0;
}

int polyeri_one_electron(polyeri_ctx* ctx, double* s_c, double* t_c, double* n_c, double* e_nuc) {
    if (!ctx) return fail(POLYERI_E_ARG, "null context");
    if (s_c) std::copy(ctx->S_C.begin(), ctx->S_C.end(), s_c);
    if (t_c) std::copy(ctx->T_C.begin(), ctx->T_C.end(), t_c);
    if (n_c) std::copy(ctx->N_C.begin(), ctx->N_C.end(), n_c);
    if (e_nuc) *e_nuc = ctx->e_nuc;
    return 0;
}

// Internal ERI tensor on the device, screened like the J/K build: the full
// nbf^4 tensor, or (packed, computed integrals only) the tensor with its
// first pair packed, nbf (nbf + 1) / 2 x nbf^2 (TensorCb).
static int tensor_device(polyeri_ctx* c, double* d_t, bool packed = false) {
    if (c->pbc) return fail(POLYERI_E_ARG, "periodic context: no molecular ERI tensor");
    if (packed && c->use_stored) return fail(POLYERI_E_ARG, "packed tensor: computed integrals only");
    size_t n = c->ib.nbf;
    size_t n4 = packed ? n * (n + 1) / 2 * n * n : n * n * n * n;
    hipLaunchKernelGGL(k_fill, dim3(nblk(n4)), dim3(256), 0, c->stream, d_t, n4, 0.0);
    HIPCHK(hipGetLastError());
    if (c->use_stored) {   // DIRECT .F.: the tensor from the stored records
        if (polyeri_tensor_stored_launch(c->d_ao, c->n_ao, c->d_inv, c->d_scale, (int)n, d_t, c->stream))
            return fail(POLYERI_E_HIP, "stored-integral tensor launch failed");
        return 0;
    }
    c->tpack = packed;
    for (const ClassPairWork& cw : c->work) {
        // lists with a generally contracted form of shell-pair groups: one
        // primitive loop per group quartet for all its member quartets
        const bool gc = cw.gc.nwork > 0 && cw.gc.kind == 0;
        int rc = gc ? launch_gc(c, cw, cw.gc_block, nullptr, d_t)
                    : launch_class(c, MODE_TENSOR, cw.la, cw.lb, cw.lc, cw.ld, cw.d_bra, cw.d_ket, cw.d_row,
                                   cw.nrows, cw.nwork, d_t, cw.d_rb, cw.d_bstart, cw.nbucket, cw.d_rf, nullptr,
                                   nullptr, LaunchCfg(), cw.d_bsoa);
        if (rc) { c->tpack = 0; return rc; }
    }
    c->tpack = 0;
    return 0;
}

int polyeri_eri_tensor(polyeri_ctx* ctx, double* feri) {
    if (!ctx || !feri) return fail(POLYERI_E_ARG, "polyeri_eri_tensor: null argument");
    HIPCHK(hipSetDevice(ctx->device));
    size_t n = ctx->ib.nbf, nc = ctx->rb.ncgs;
    double* d_t = nullptr;
    int rc;
    if ((rc = dalloc(&d_t, n * n * n * n))) return rc;
    rc = tensor_device(ctx, d_t);
    std::vector<double> t;
    if (!rc) {
        t.resize(n * n * n * n);
        hipError_t e = hipMemcpyAsync(t.data(), d_t, t.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = fail(POLYERI_E_HIP, hipGetErrorString(e));
    }
    (void)hipFree(d_t);
    if (rc) return rc;
    std::fill(feri, feri + nc * nc * nc * nc, 0.0);
    const auto& ref = ctx->ib.ref_of_int;
    const auto& sc = ctx->ib.scale;
    for (size_t a = 0; a < n; ++a)
        for (size_t b = 0; b < n; ++b)
            for (size_t cc = 0; cc < n; ++cc)
                for (size_t d = 0; d < n; ++d)
                    feri[ref[a] + nc * (ref[b] + nc * (ref[cc] + nc * ref[d]))] =
                        sc[a] * sc[b] * sc[cc] * sc[d] * t[((a * n + b) * n + cc) * n + d];
    return 0;
}

int polyeri_get_stats(polyeri_ctx* ctx, polyeri_stats* st) {
    if (!ctx || !st) return fail(POLYERI_E_ARG, "null argument");
    memset(st, 0, sizeof(*st));
    st->nshell = (int)ctx->ib.shells.size();
    st->nbf = ctx->ib.nbf;
    st->npair = (int)ctx->pairs.size();
    st->nprim_pair = (int)ctx->prims.size();
    st->quartets = ctx->quartets;
    st->prim_quartets = ctx->prim_quartets;
    st->model_flops = ctx->model_flops;
    st->last_jk_ms = ctx->last_jk_ms;
    st->last_quartets = ctx->last_quartets;
    st->nclass = (int)std::min<size_t>(POLYERI_MAX_LISTS, ctx->work.size());
    st->last_mp2_tensor_ms = ctx->last_mp2_tensor_ms;
    st->last_mp2_transform_ms = ctx->last_mp2_transform_ms;
    st->last_mp2_flops = ctx->last_mp2_flops;
    for (int i = 0; i < st->nclass; ++i) {
        const ClassPairWork& w = ctx->work[i];
        st->last_class_ms[i] = ctx->last_class_ms[i];
        st->class_quartets[i] = w.nwork;
        st->class_prec[i] = w.prec;
        const int kind = w.gc_on && w.gc.nwork > 0 ? 4
                       : w.cfg.kind != LAUNCH_AUTO ? w.cfg.kind
                       : auto_kind(ctx, w.la, w.lb, w.lc, w.ld, w.nseg > 0 && !ctx->pbc);
        st->class_kind[i] = kind;
        st->class_block[i] = kind == LAUNCH_COOP ? 64 : kind == 4 ? w.gc_block : w.cfg.block;
        st->class_grid[i] = kind == 4 ? (w.gc_segv ? w.gc.nsegc : w.gc.nseg)
                          : kind == LAUNCH_SEG || kind == LAUNCH_RYS ? (w.cfg.segv ? w.nsegc : w.nseg) : w.cfg.grid;
        st->class_prim_quartets[i] = w.prim_quartets;
        st->class_flops[i] = w.flops;
        if (kind == 4) {   // executed arithmetic: every member combination per primitive quartet
            const int ni = ncart_h(w.la) * ncart_h(w.lc);
            const int mm = (w.la == 0 ? 4 : 2) * (w.lc == 0 ? 4 : 2);
            st->class_prim_quartets[i] = w.gc.prim_quartets;
            st->class_flops[i] = w.gc.kind == 1 ? w.gc.prim_quartets * flop_model(w.la, w.lb, w.lc, 0)
                                                : w.gc.prim_quartets * (flop_model(w.la, 0, w.lc, 0) + 2.0 * mm * ni);
            st->prim_quartets += w.gc.prim_quartets - w.prim_quartets;
            st->model_flops += st->class_flops[i] - w.flops;
        }
        st->class_l[i] = w.la * 1000 + w.lb * 100 + w.lc * 10 + w.ld;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// SCF on the device
// ---------------------------------------------------------------------------
// src/pmodules.f90:12
static const double kBoltzmann = 1.0 / 3.1577464e5;

// Solve the DIIS system with partial pivoting and one refinement step (the
// role of LUDCMP / LUBKSB / MPROVE, src/pdensitymatrix.f90:385-512).
static void lu_solve(int n, const std::vector<double>& A0, const std::vector<double>& b0, std::vector<double>& x) {
    std::vector<double> A = A0, sc(n, 1.0), r(n);
    std::vector<int> piv(n);
    for (int i = 0; i < n; ++i) {
        double m = 0;
        for (int j = 0; j < n; ++j) m = std::max(m, std::fabs(A[i + n * j]));
        sc[i] = m > 0 ? 1.0 / m : 1.0;
    }
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = -1;
        for (int i = k; i < n; ++i) {
            double v = std::fabs(A[i + n * k]) * sc[i];
            if (v > best) { best = v; p = i; }
        }
        piv[k] = p;
        if (p != k) {
            for (int j = 0; j < n; ++j) std::swap(A[k + n * j], A[p + n * j]);
            std::swap(sc[k], sc[p]);
        }
        if (A[k + n * k] == 0.0) A[k + n * k] = 1e-20;
        for (int i = k + 1; i < n; ++i) {
            double f = A[i + n * k] / A[k + n * k];
            A[i + n * k] = f;
            for (int j = k + 1; j < n; ++j) A[i + n * j] -= f * A[k + n * j];
        }
    }
    x.assign(n, 0.0);
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 0) r = b0;
        else
            for (int i = 0; i < n; ++i) {
                double s = -b0[i];
                for (int j = 0; j < n; ++j) s += A0[i + n * j] * x[j];
                r[i] = s;
            }
        for (int k = 0; k < n; ++k) if (piv[k] != k) std::swap(r[k], r[piv[k]]);
        for (int i = 0; i < n; ++i) for (int j = 0; j < i; ++j) r[i] -= A[i + n * j] * r[j];
        for (int i = n - 1; i >= 0; --i) {
            for (int j = i + 1; j < n; ++j) r[i] -= A[i + n * j] * r[j];
            r[i] /= A[i + n * i];
        }
        if (pass == 0) x = r;
        else for (int i = 0; i < n; ++i) x[i] -= r[i];
    }
}

// FORM_C2S (src/pbasis.f90:31-130) for d shells.
static int form_c2s(const RefBasis& rb, int spherical, std::vector<double>& c2s) {
    int nc = rb.ncgs, np = rb.npgs;
    c2s.assign((size_t)nc * nc, 0.0);
    for (int i = 0; i < nc; ++i) c2s[i + (size_t)nc * i] = 1.0;
    if (!spherical) return 0;
    std::vector<int> p2c(np, 0);
    for (int p = 0; p < np; ++p) {
        int k = 0;
        for (int c = 0; c < nc; ++c)
            if (rb.cc[c + (size_t)nc * p] != 0.0) { ++k; p2c[p] = c; }
        if (k > 1) return fail(POLYERI_E_BASIS, "GENERAL CONTRACTION NOT YET IMPLEMENTED");
    }
    auto C = [&](int i, int j) -> double& { return c2s[i + (size_t)nc * j]; };
    for (int p = 0; p < np; ++p) {
        int x = rb.pax[p], y = rb.pay[p], z = rb.paz[p];
        if (x + y + z == 3) return fail(POLYERI_E_BASIS, "spherical f functions not supported");
        if (x == 2 && y == 0 && z == 0) C(p2c[p], p2c[p]) = 0.0;
        else if (x == 0 && y == 2 && z == 0) {
            C(p2c[p], p2c[p]) = -std::sqrt(3.0) / 2.0;
            C(p2c[p], p2c[p - 1]) = std::sqrt(3.0) / 2.0;
        } else if (x == 0 && y == 0 && z == 2) {
            C(p2c[p], p2c[p - 2]) = -0.5;
            C(p2c[p], p2c[p - 1]) = -0.5;
        }
    }
    return 0;
}

struct DevBuf {
    std::vector<void*> ptrs;
    ~DevBuf() { for (void* p : ptrs) (void)hipFree(p); }
    template <typename T>
    int get(T** p, size_t n) {
        int rc = dalloc(p, n);
        if (!rc) ptrs.push_back(*p);
        return rc;
    }
};

// Device buffers of polyeri_scf kept by the context between calls: the k-th
// request of a call gets the k-th buffer (re-allocated only when it must
// grow), so repeated SCFs of one geometry allocate nothing after the first.
// Every call initialises what it reads.
int ScfArena::get_bytes(void** p, size_t bytes) {
    if (bytes == 0) bytes = 8;
    if (next < bufs.size() && bufs[next].second >= bytes) {
        *p = bufs[next++].first;
        return 0;
    }
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) return fail(POLYERI_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    if (next < bufs.size()) {
        (void)hipFree(bufs[next].first);
        bufs[next] = {q, bytes};
    } else {
        bufs.push_back({q, bytes});
    }
    ++next;
    *p = q;
    return 0;
}

void ScfArena::release() {
    for (auto& b : bufs) (void)hipFree(b.first);
    bufs.clear();
    next = 0;
}

static int ensure_blas(polyeri_ctx* c) {
    if (!c->blas) {
        RBCHK(rocblas_create_handle(&c->blas));
        RBCHK(rocblas_set_stream(c->blas, c->stream));
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Zero-temperature density without diagonalisation: trace-correcting
// purification (TC2; A. M. N. Niklasson, Phys. Rev. B 66, 155115 (2002)).
// X_0 = (e_max I - F') / (e_max - e_min) from Gershgorin bounds of the
// orthonormal-basis Fock matrix F' maps the spectrum into [0, 1] with the
// occupied levels near 1; each step replaces X by X^2 or by 2X - X^2,
// whichever brings the trace closer to the number of occupied orbitals, and
// the sequence converges quadratically to the projector on the nocc lowest
// eigenvectors, so D' = 2 X equals the aufbau density of the diagonalisation
// whenever the HOMO-LUMO gap is open.  Every step is one FP64 product and
// two small kernels; the step decision stays on the device.
// ---------------------------------------------------------------------------
// Gershgorin interval of column i of a symmetric n x n matrix (one
// wavefront per column): lohi[2i] = a_ii - r_i, lohi[2i+1] = a_ii + r_i.
__global__ void __launch_bounds__(64) k_gershgorin(const double* __restrict__ A, int n, double* __restrict__ lohi) {
    const int i = blockIdx.x;
    const double* col = A + (size_t)n * i;   // column i = row i
    double r = 0.0;
    for (int j = threadIdx.x; j < n; j += 64) r += j == i ? 0.0 : fabs(col[j]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) r += __shfl_xor(r, o, 64);
    if (threadIdx.x == 0) {
        lohi[2 * i] = col[i] - r;
        lohi[2 * i + 1] = col[i] + r;
    }
}

// Block-wide min / max of the n Gershgorin intervals (every block reduces
// them itself, 2n loads from L2).
__device__ void gersh_bounds(const double* __restrict__ lohi, int n, double* lo, double* hi) {
    __shared__ double sl[256], sh[256];
    double mn = 1e300, mx = -1e300;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        mn = fmin(mn, lohi[2 * i]);
        mx = fmax(mx, lohi[2 * i + 1]);
    }
    sl[threadIdx.x] = mn;
    sh[threadIdx.x] = mx;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            sl[threadIdx.x] = fmin(sl[threadIdx.x], sl[threadIdx.x + h]);
            sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + h]);
        }
        __syncthreads();
    }
    *lo = sl[0];
    *hi = sh[0];
}

// X = (e_max I - F) / (e_max - e_min); blocks of 256 threads.
__global__ void __launch_bounds__(256) k_tc2_init(const double* __restrict__ F, int n, const double* __restrict__ lohi,
                                                  double* __restrict__ X) {
    double lo, hi;
    gersh_bounds(lohi, n, &lo, &hi);
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n * n) return;
    const double w = hi - lo;
    X[t] = ((t % n == t / n ? hi : 0.0) - F[t]) / (w > 0.0 ? w : 1.0);
}

// One TC2 step after Y = X X, into Z (X and Y are only read, so no block
// overwrites a diagonal element another block still sums): every block
// sums the two diagonals itself (2n loads), takes Z = X^2 when tr X^2 is
// nearer to nocc than 2 tr X - tr X^2, else Z = 2X - X^2; block 0 records
// the idempotency error tr X - tr X^2 of X in st[1].  Blocks of 256.
__global__ void __launch_bounds__(256) k_tc2_step(const double* __restrict__ X, const double* __restrict__ Y, int n,
                                                  double nocc, double* __restrict__ st, double* __restrict__ Z) {
    __shared__ double sx[256], sy[256];
    double a = 0.0, c = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        a += X[(size_t)i * (n + 1)];
        c += Y[(size_t)i * (n + 1)];
    }
    sx[threadIdx.x] = a;
    sy[threadIdx.x] = c;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            sx[threadIdx.x] += sx[threadIdx.x + h];
            sy[threadIdx.x] += sy[threadIdx.x + h];
        }
        __syncthreads();
    }
    const double tx = sx[0], ty = sy[0];
    const bool square = fabs(ty - nocc) < fabs(2.0 * tx - ty - nocc);
    if (blockIdx.x == 0 && threadIdx.x == 0) st[1] = tx - ty;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n * n) return;
    Z[t] = square ? Y[t] : 2.0 * X[t] - Y[t];
}

// E = B - B^T (n x n): the commutator F P S - S P F from B = F P S.
__global__ void k_antisym(const double* __restrict__ B, int n, double* __restrict__ E) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n * n) return;
    const size_t i = t % n, j = t / n;
    E[t] = B[t] - B[i * n + j];
}

// *result <- the projector on the nocc lowest eigenvectors of Fp (one of
// the n x n buffers X, Z); Y, st (4 doubles) and lohi (2n) are work space.  *ok = 0 when tr X - tr X^2 did not fall
// below tol within 96 steps (a closed gap): the caller then diagonalises.
// The host reads the idempotency error first after *steps - 4 steps (the
// count the previous call needed; at least 8), then every 4 steps, and
// returns the count used in *steps.
static int purify_tc2(rocblas_handle h, hipStream_t s, const double* Fp, int n, int nocc, double* X, double* Y,
                      double* Z, double* st, double* lohi, int* ok, int* steps, double** result) {
    const size_t n2 = (size_t)n * n;
    const double one = 1.0, zero = 0.0, tol = 1e-11;
    *ok = 0;
    hipLaunchKernelGGL(k_gershgorin, dim3(n), dim3(64), 0, s, Fp, n, lohi);
    hipLaunchKernelGGL(k_tc2_init, dim3(nblk(n2)), dim3(256), 0, s, Fp, n, lohi, X);
    HIPCHK(hipGetLastError());
    double err = 0.0;
    int done = 0, next = std::max(8, *steps - 4);
    while (!*ok && done < 96) {
        for (; done < next; ++done) {
            RBCHK(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, n, n, n, &one, X, n, X, n, &zero, Y,
                                n));
            hipLaunchKernelGGL(k_tc2_step, dim3(nblk(n2)), dim3(256), 0, s, X, Y, n, (double)nocc, st, Z);
            std::swap(X, Z);   // the step's result is the next X
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(&err, st + 1, sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (std::isfinite(err) && std::fabs(err) < tol) *ok = 1;
        next = done + 4;
    }
    *steps = done;
    *result = X;
    return 0;
}

int polyeri_scf(polyeri_ctx* ctx, const polyeri_scf_opts* o, double* p_c, double* eps_out, double* co_out,
                double* trace, polyeri_scf_result* res) {
    if (!ctx || !o || !p_c || !res) return fail(POLYERI_E_ARG, "polyeri_scf: null argument");
    if (o->scfcycles < 1) return fail(POLYERI_E_ARG, "polyeri_scf: scfcycles < 1");
    auto t_start = std::chrono::steady_clock::now();
    HIPCHK(hipSetDevice(ctx->device));
    int rc;
    if ((rc = ensure_blas(ctx))) return rc;
    rocblas_handle h = ctx->blas;
    const RefBasis& rb = ctx->rb;
    int nc = rb.ncgs;
    size_t n2 = (size_t)nc * nc;
    hipStream_t s = ctx->stream;
    // POLYERI_SCF_PHASES=1: wall time per phase (the stream drained at each
    // boundary), printed to stderr at the end -- a diagnostic
    static const bool phases = std::getenv("POLYERI_SCF_PHASES") != nullptr;
    double ph[6] = {0, 0, 0, 0, 0, 0};   // J/K + Fock, F', DIIS, density, bookkeeping + energy, final orbitals
    auto tick = [&]() {
        if (phases) (void)hipStreamSynchronize(s);
        return std::chrono::steady_clock::now();
    };
    auto tnow = tick();
    auto lap = [&](int k) {
        if (!phases) return;
        const auto t = tick();
        ph[k] += std::chrono::duration<double>(t - tnow).count();
        tnow = t;
    };
    ScfArena& buf = ctx->scf_arena;
    buf.next = 0;
    double *dF, *dW1, *dW2, *dE, *dWk, *dCO, *dP, *dPout, *dOcc, *dFp;
    int* dInfo;
    if ((rc = buf.get(&dF, n2)) || (rc = buf.get(&dW1, n2)) || (rc = buf.get(&dW2, n2)) ||
        (rc = buf.get(&dE, nc)) || (rc = buf.get(&dWk, nc)) || (rc = buf.get(&dCO, n2)) || (rc = buf.get(&dP, n2)) ||
        (rc = buf.get(&dPout, n2)) || (rc = buf.get(&dOcc, nc)) || (rc = buf.get(&dFp, n2)) || (rc = buf.get(&dInfo, 1)))
        return rc;
    // incremental builds: density / J / K of the last build, the difference;
    // density-screened builds keep the input density of the last build too
    // (for the unscreened build at convergence)
    double *dPprev = nullptr, *dJprev = nullptr, *dKprev = nullptr, *dDel = nullptr, *dJd = nullptr, *dKd = nullptr;
    const double dens_tol = o->dens_tol > 0 ? o->dens_tol : 0.0;
    if (o->incremental > 0 || dens_tol > 0.0)
        if ((rc = buf.get(&dPprev, n2))) return rc;
    if (o->incremental > 0)
        if ((rc = buf.get(&dJprev, n2)) || (rc = buf.get(&dKprev, n2)) || (rc = buf.get(&dDel, n2)) ||
            (rc = buf.get(&dJd, n2)) || (rc = buf.get(&dKd, n2)))
            return rc;
    int since_full = 0;
    const int HMAX = 9;   // DIIS uses at most 8 previous vectors
    double *dHP, *dHD;
    if ((rc = buf.get(&dHP, n2 * HMAX)) || (rc = buf.get(&dHD, n2 * HMAX))) return rc;
    const double one = 1.0, zero = 0.0;
    // T, N, S, C2S and the canonical orthogonaliser of C2S S C2S^T
    // (src/pfock.f90:270-281): computed once per geometry and (spherical,
    // LINDEP), kept by the context for later SCFs
    if (!ctx->scf_orth || ctx->scf_spherical != o->spherical || ctx->scf_lindep != o->lindep) {
        ctx->scf_orth = false;
        for (double** d : {&ctx->d_scfT, &ctx->d_scfV, &ctx->d_scfS, &ctx->d_scfC2S, &ctx->d_scfX})
            if (!*d && (rc = dalloc(d, n2))) return rc;
        std::vector<double> c2s;
        if ((rc = form_c2s(rb, o->spherical, c2s))) return rc;
        HIPCHK(hipMemcpy(ctx->d_scfT, ctx->T_C.data(), n2 * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ctx->d_scfV, ctx->N_C.data(), n2 * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ctx->d_scfS, ctx->S_C.data(), n2 * 8, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ctx->d_scfC2S, c2s.data(), n2 * 8, hipMemcpyHostToDevice));
        RBCHK(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, nc, nc, nc, &one, ctx->d_scfC2S, nc,
                            ctx->d_scfS, nc, &zero, dW1, nc));
        RBCHK(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_transpose, nc, nc, nc, &one, dW1, nc,
                            ctx->d_scfC2S, nc, &zero, dW2, nc));
        RBCHK(rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, nc, dW2, nc, dE, dWk, dInfo));
        std::vector<double> e(nc);
        HIPCHK(hipMemcpyAsync(e.data(), dE, nc * 8, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        std::vector<int> keep;
        for (int j = 0; j < nc; ++j) if (e[j] > o->lindep) keep.push_back(j);
        if (keep.empty()) return fail(POLYERI_E_ARG, "no linearly independent functions");
        DevBuf once;   // not from the arena, so its slots keep their sizes from call to call
        int* dKeep;
        if ((rc = once.get(&dKeep, keep.size()))) return rc;
        HIPCHK(hipMemcpy(dKeep, keep.data(), keep.size() * sizeof(int), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_orth, dim3(nblk((size_t)nc * keep.size())), dim3(256), 0, s, dW2, dE, dKeep, nc,
                           (int)keep.size(), ctx->d_scfX);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(s));   // before `once` frees dKeep
        ctx->scf_nind = (int)keep.size();
        ctx->scf_spherical = o->spherical;
        ctx->scf_lindep = o->lindep;
        ctx->scf_orth = true;
    }
    double *dT = ctx->d_scfT, *dV = ctx->d_scfV, *dS = ctx->d_scfS, *dC2S = ctx->d_scfC2S, *dX = ctx->d_scfX;
    const int nind = ctx->scf_nind;
    // electrons (src/porbitals.f90:15-19)
    int nel = 0;
    for (int z : rb.iatom) nel += z;
    nel -= o->charge;
    if (nel % 2) return fail(POLYERI_E_ARG, "ODD NUMBER OF ELECTRONS");
    int iocc = nel / 2;
    // solver options (no counterpart in the reference): purification of the
    // zero-temperature density instead of the diagonalisation after the
    // first iteration, commutator DIIS on the Fock matrices every iteration
    const bool want_purify = o->solver == 1 && (o->temphf > 0 ? o->temphf : 1.0e-4) <= 1.0 && iocc > 0 && iocc < nind;
    const bool cdiis = o->diis_mode == 1;
    double *dXp = nullptr, *dYp = nullptr, *dSt = nullptr, *dFpLast = nullptr, *dHF = nullptr, *dHE = nullptr,
           *dDots = nullptr;
    double* dEn = nullptr;
    if ((rc = buf.get(&dEn, 5))) return rc;
    double *dZp = nullptr, *dLohi = nullptr;
    if (want_purify &&
        ((rc = buf.get(&dXp, n2)) || (rc = buf.get(&dYp, n2)) || (rc = buf.get(&dSt, 4)) || (rc = buf.get(&dFpLast, n2)) ||
         (rc = buf.get(&dZp, n2)) || (rc = buf.get(&dLohi, 2 * (size_t)nc))))
        return rc;
    if (cdiis && ((rc = buf.get(&dHF, n2 * 8)) || (rc = buf.get(&dHE, n2 * 8)) || (rc = buf.get(&dDots, 8))))
        return rc;
    bool purify_ok = want_purify, last_purified = false;
    int tc2_steps = 24;
    std::vector<double> cB(64 * 64, 0.0);   // commutator DIIS: <e_i, e_j> of the last 8 iterations (ring)
    // guess
    std::vector<double> P(n2, 0.0);
    if (o->guess == 1) {
        std::copy(p_c, p_c + n2, P.begin());
    } else {
        // GUESS_DENSITYMATRIX (src/pdensitymatrix.f90:83-127): Z / (number of
        // non-ghost functions on the atom) on the diagonal
        std::vector<int> fatom(nc, -1);
        for (int c = 0; c < nc; ++c) {
            int mu = ctx->ib.int_of_ref[c];
            if (mu < 0) continue;
            for (const Shell& sh : ctx->ib.shells)
                if (mu >= sh.f0 && mu < sh.f0 + ncart(sh.L))
                    for (int a = 0; a < rb.natom; ++a)
                        if (sh.x == rb.ax[a] && sh.y == rb.ay[a] && sh.z == rb.az[a]) { fatom[c] = a; break; }
        }
        for (int a = 0; a < rb.natom; ++a) {
            int L = 0;
            for (int c = 0; c < nc; ++c) L += fatom[c] == a;
            if (!L) return fail(POLYERI_E_ARG, "atom without basis functions");
            for (int c = 0; c < nc; ++c) if (fatom[c] == a) P[c + (size_t)nc * c] = (double)rb.iatom[a] / L;
        }
    }
    HIPCHK(hipMemcpyAsync(dP, P.data(), n2 * 8, hipMemcpyHostToDevice, s));
    double temphf = o->temphf > 0 ? o->temphf : 1.0e-4;
    std::vector<double> pulay((size_t)(o->scfcycles + 2) * (o->scfcycles + 2), 0.0);
    int ld = o->scfcycles + 2;
    std::vector<double> eps(nc, 1e99), occ(nc, 0.0);
    double etot = 0, ek = 0, ev = 0, ej = 0, ex = 0, fermi = 0;
    int converged = 0, iscf;
    double jk_total = 0, fock_total = 0;
    long long build_quartets = 0;
    hipEvent_t ef0, ef1;
    HIPCHK(hipEventCreate(&ef0));
    HIPCHK(hipEventCreate(&ef1));
    // eigenvalues, coefficients, Fermi occupations and the output density
    // of F' (dFp, overwritten): the reference's FOCK_BUILD path
    auto diagonalise = [&]() -> int {
        RBCHK(rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, nind, dFp, nind, dE, dWk, dInfo));
        // C = X C', CO = C2S^T C
        RBCHK(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, nc, nind, nind, &one, dX, nc, dFp, nind, &zero, dW1, nc));
        RBCHK(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, nc, nind, nc, &one, dC2S, nc, dW1, nc, &zero, dCO, nc));
        HIPCHK(hipMemcpyAsync(eps.data(), dE, nind * 8, hipMemcpyDeviceToHost, s));
        int info = 0;
        HIPCHK(hipMemcpyAsync(&info, dInfo, sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (info != 0) return fail(POLYERI_E_LIB, "EIGENVECTORS NOT CONVERGED");
        for (int j = nind; j < nc; ++j) eps[j] = 1.0e99;
        // Fermi level by bisection (src/pfock.f90:339-423)
        double mine = 1e99, maxe = -1e99;
        for (int j = 0; j < nc; ++j) { maxe = std::max(maxe, eps[j]); mine = std::min(mine, eps[j]); }
        double f1 = mine, f2 = maxe, f12 = 0.5 * (mine + maxe);
        for (int it = 0; it < 100000; ++it) {
            double n1 = 0, nn2 = 0, n12 = 0;
            for (int j = 0; j < nc; ++j) {
                n1 += 2.0 / (std::exp((eps[j] - f1) / kBoltzmann / temphf) + 1.0);
                nn2 += 2.0 / (std::exp((eps[j] - f2) / kBoltzmann / temphf) + 1.0);
                n12 += 2.0 / (std::exp((eps[j] - f12) / kBoltzmann / temphf) + 1.0);
            }
            n1 -= nel; nn2 -= nel; n12 -= nel;
            if (std::fabs(n12) < 1e-10) { fermi = f12; break; }
            else if (n1 > 0 && nn2 > 0) { f2 = f1; f1 -= 0.1; }
            else if (n1 < 0 && nn2 < 0) { f1 = f2; f2 += 0.1; }
            else if (n1 < 0 && n12 >= 0) f2 = f12;
            else if (n12 <= 0 && nn2 > 0) f1 = f12;
            else break;
            f12 = 0.5 * (f1 + f2);
        }
        for (int j = 0; j < nc; ++j) occ[j] = j < nind ? 2.0 / (std::exp((eps[j] - fermi) / kBoltzmann / temphf) + 1.0) : 0.0;
        HIPCHK(hipMemcpyAsync(dOcc, occ.data(), nind * 8, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_scale_cols, dim3(nblk((size_t)nc * nind)), dim3(256), 0, s, dCO, dOcc, nc, nind, dW1);
        RBCHK(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_transpose, nc, nc, nind, &one, dCO, nc, dW1, nc, &zero, dPout, nc));
        return 0;
    };
    for (iscf = 1; iscf <= o->scfcycles; ++iscf) {
        lap(4);
        HIPCHK(hipEventRecord(ef0, s));
        const bool inc_now = o->incremental > 0 && iscf > 1 && since_full < o->incremental;
        if (inc_now) {
            // J/K(P) = J/K(P_prev) + J/K(P - P_prev)
            hipLaunchKernelGGL(k_axpby, dim3(nblk(n2)), dim3(256), 0, s, dP, dPprev, -1.0, n2, dDel);
            if ((rc = jk_device(ctx, dDel, dJd, dKd, dens_tol))) break;
            hipLaunchKernelGGL(k_axpby, dim3(nblk(n2)), dim3(256), 0, s, dJprev, dJd, 1.0, n2, ctx->d_jc);
            hipLaunchKernelGGL(k_axpby, dim3(nblk(n2)), dim3(256), 0, s, dKprev, dKd, 1.0, n2, ctx->d_kc);
            ++since_full;
        } else {
            if ((rc = jk_device(ctx, dP, ctx->d_jc, ctx->d_kc, dens_tol))) break;
            since_full = 1;
        }
        if (dPprev) HIPCHK(hipMemcpyAsync(dPprev, dP, n2 * 8, hipMemcpyDeviceToDevice, s));
        if (o->incremental > 0) {
            HIPCHK(hipMemcpyAsync(dJprev, ctx->d_jc, n2 * 8, hipMemcpyDeviceToDevice, s));
            HIPCHK(hipMemcpyAsync(dKprev, ctx->d_kc, n2 * 8, hipMemcpyDeviceToDevice, s));
        }

        hipLaunchKernelGGL(k_fock, dim3(nblk(n2)), dim3(256), 0, s, dT, dV, ctx->d_jc, ctx->d_kc, o->hfexchange, nc, dF);
        HIPCHK(hipEventRecord(ef1, s));
        lap(0);
        // C2S F C2S^T, X^T (.) X
        RBCHK(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, nc, nc, nc, &one, dC2S, nc, dF, nc, &zero, dW1, nc));
        RBCHK(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_transpose, nc, nc, nc, &one, dW1, nc, dC2S, nc, &zero, dW2, nc));
        RBCHK(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, nc, nind, nc, &one, dW2, nc, dX, nc, &zero, dW1, nc));
        RBCHK(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, nind, nind, nc, &one, dX, nc, dW1, nc, &zero, dFp, nind));
        lap(1);
        // commutator DIIS (diis_mode 1): error F P S - S P F of this
        // iteration's input density, in the orthonormal basis; F' is replaced
        // by the extrapolation over the last (up to 8) iterations
        if (cdiis) {
            const int slot8 = (iscf - 1) % 8;
            double* he = dHE + n2 * slot8;
            RBCHK(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, nc, nc, nc, &one, dF, nc, dP, nc, &zero, dW1, nc));
            RBCHK(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, nc, nc, nc, &one, dW1, nc, dS, nc, &zero, dW2, nc));
            hipLaunchKernelGGL(k_antisym, dim3(nblk(n2)), dim3(256), 0, s, dW2, nc, dW1);
            RBCHK(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, nc, nc, nc, &one, dC2S, nc, dW1, nc, &zero, dW2, nc));
            RBCHK(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_transpose, nc, nc, nc, &one, dW2, nc, dC2S, nc, &zero, dW1, nc));
            RBCHK(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, nc, nind, nc, &one, dW1, nc, dX, nc, &zero, dW2, nc));
            RBCHK(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, nind, nind, nc, &one, dX, nc, dW2, nc, &zero, he, nind));
            HIPCHK(hipMemcpyAsync(dHF + n2 * slot8, dFp, (size_t)nind * nind * 8, hipMemcpyDeviceToDevice, s));
            const int nt = std::min(8, iscf);
            RBCHK(rocblas_set_pointer_mode(h, rocblas_pointer_mode_device));
            for (int i = 0; i < nt; ++i)
                RBCHK(rocblas_ddot(h, nind * nind, dHE + n2 * ((iscf - 1 - i) % 8), 1, he, 1, dDots + i));
            RBCHK(rocblas_set_pointer_mode(h, rocblas_pointer_mode_host));
            double dots[8];
            HIPCHK(hipMemcpyAsync(dots, dDots, nt * 8, hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            // B[a][b] over iterations a, b (ring of 64 x 64 by iteration mod 64)
            for (int i = 0; i < nt; ++i) {
                const int a = (iscf - 1) % 64, b = (iscf - 1 - i) % 64;
                cB[a * 64 + b] = cB[b * 64 + a] = dots[i];
            }
            if (nt > 1) {
                const int m = nt + 1;
                std::vector<double> B((size_t)m * m, 0.0), rhs(m, 0.0), cv;
                for (int i = 0; i < nt; ++i) {
                    for (int j = 0; j < nt; ++j) B[j + (size_t)m * i] = cB[((iscf - 1 - i) % 64) * 64 + (iscf - 1 - j) % 64];
                    B[i + (size_t)m * nt] = -1.0;
                    B[nt + (size_t)m * i] = -1.0;
                }
                rhs[nt] = -1.0;
                lu_solve(m, B, rhs, cv);
                bool finite = true;
                for (int i = 0; i < nt; ++i) finite = finite && std::isfinite(cv[i]);
                if (finite) {
                    hipLaunchKernelGGL(k_fill, dim3(nblk((size_t)nind * nind)), dim3(256), 0, s, dFp, (size_t)nind * nind, 0.0);
                    for (int i = 0; i < nt; ++i) {
                        const double w = cv[i];
                        RBCHK(rocblas_daxpy(h, nind * nind, &w, dHF + n2 * ((iscf - 1 - i) % 8), 1, dFp, 1));
                    }
                }
            }
        }
        lap(2);
        // the density: purification while it converges (an open gap), else
        // the diagonalisation; after a diagonalisation purification resumes
        // only if the HOMO-LUMO gap exceeds 1 mEh
        int purified = 0;
        double* dProj = nullptr;
        if (purify_ok) {
            HIPCHK(hipMemcpyAsync(dFpLast, dFp, (size_t)nind * nind * 8, hipMemcpyDeviceToDevice, s));
            if ((rc = purify_tc2(h, s, dFp, nind, iocc, dXp, dYp, dZp, dSt, dLohi, &purified, &tc2_steps, &dProj)))
                break;
        }
        if (purified) {
            // P = C2S^T X (2 X_p) X^T C2S
            const double two = 2.0;
            RBCHK(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, nc, nind, nind, &two, dX, nc, dProj, nind, &zero, dW1, nc));
            RBCHK(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_transpose, nc, nc, nind, &one, dW1, nc, dX, nc, &zero, dW2, nc));
            RBCHK(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, nc, nc, nc, &one, dC2S, nc, dW2, nc, &zero, dW1, nc));
            RBCHK(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, nc, nc, nc, &one, dW1, nc, dC2S, nc, &zero, dPout, nc));
        } else {
            if ((rc = diagonalise())) break;
            if (want_purify)   // eps ascending
                purify_ok = std::isfinite(eps[iocc]) && eps[iocc] < 1e90 && eps[iocc] - eps[iocc - 1] > 1e-3;
        }
        last_purified = purified != 0;
        lap(3);
        {
            fock_total += event_ms(ef0, ef1) * 1e-3;
            jk_total += event_ms(ctx->ev0, ctx->ev1) * 1e-3;
            if ((rc = jk_collect_timing(ctx))) break;
            build_quartets += ctx->last_quartets;
        }
        // STORE_DENSITYMATRIX: Delta P, RMS change, Pulay row (src/pdensitymatrix.f90:131-208)
        double dchg = 0.0;
        bool have_energy = false;
        if (cdiis) {
            // no density history: the change and this iteration's energy
            // (input density, its own J/K) in one synchronisation
            hipLaunchKernelGGL(k_axpby, dim3(nblk(n2)), dim3(256), 0, s, dPout, dP, -1.0, n2, dW1);
            RBCHK(rocblas_set_pointer_mode(h, rocblas_pointer_mode_device));
            RBCHK(rocblas_ddot(h, (int)n2, dW1, 1, dW1, 1, dEn + 4));
            RBCHK(rocblas_ddot(h, (int)n2, dP, 1, dT, 1, dEn));
            RBCHK(rocblas_ddot(h, (int)n2, dP, 1, dV, 1, dEn + 1));
            RBCHK(rocblas_ddot(h, (int)n2, dP, 1, ctx->d_jc, 1, dEn + 2));
            RBCHK(rocblas_ddot(h, (int)n2, dP, 1, ctx->d_kc, 1, dEn + 3));
            RBCHK(rocblas_set_pointer_mode(h, rocblas_pointer_mode_host));
            double en[5];
            HIPCHK(hipMemcpyAsync(en, dEn, sizeof en, hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            ek = en[0]; ev = en[1]; ej = en[2]; ex = en[3];
            dchg = std::sqrt(en[4] / (double)n2);
            have_energy = true;
        } else {
            int slot = (iscf - 1) % HMAX;
            double* hp = dHP + n2 * slot;
            double* hd = dHD + n2 * slot;
            HIPCHK(hipMemcpyAsync(hp, dP, n2 * 8, hipMemcpyDeviceToDevice, s));
            HIPCHK(hipMemcpyAsync(hd, dPout, n2 * 8, hipMemcpyDeviceToDevice, s));
            const double mone = -1.0;
            RBCHK(rocblas_daxpy(h, (int)n2, &mone, dP, 1, hd, 1));
            RBCHK(rocblas_set_pointer_mode(h, rocblas_pointer_mode_host));
            RBCHK(rocblas_ddot(h, (int)n2, hd, 1, hd, 1, &dchg));
            pulay[iscf + (size_t)ld * iscf] = dchg;
            for (int i = std::max(1, iscf - HMAX + 1); i < iscf; ++i) {
                double v = 0.0;
                RBCHK(rocblas_ddot(h, (int)n2, dHD + n2 * ((i - 1) % HMAX), 1, hd, 1, &v));
                pulay[i + (size_t)ld * iscf] = v;
                pulay[iscf + (size_t)ld * i] = v;
            }
            dchg = std::sqrt(dchg / (double)n2);
        }
        // DIIS every other iteration from IOPTN(16) on, else relaxation
        // (commutator DIIS: the next input density is P_out of the
        // extrapolated Fock matrix, taken after the energy below)
        if (cdiis) {
        } else if (o->diis != 0 && iscf >= o->diis && ((iscf - o->diis) % 2) == 0) {
            int nt = std::min(8, iscf), n = nt + 1;
            std::vector<double> B((size_t)n * n, 0.0), rhs(n, 0.0), cv;
            for (int i = 0; i < nt; ++i) {
                for (int j = 0; j < nt; ++j) B[j + (size_t)n * i] = pulay[(iscf - j) + (size_t)ld * (iscf - i)];
                B[i + (size_t)n * nt] = -1.0;
                B[nt + (size_t)n * i] = -1.0;
            }
            rhs[nt] = -1.0;
            lu_solve(n, B, rhs, cv);
            hipLaunchKernelGGL(k_fill, dim3(nblk(n2)), dim3(256), 0, s, dP, n2, 0.0);
            for (int i = iscf - nt + 1; i <= iscf; ++i) {
                double w = cv[iscf - i];
                RBCHK(rocblas_daxpy(h, (int)n2, &w, dHP + n2 * ((i - 1) % HMAX), 1, dP, 1));
            }
        } else {
            double a = o->scfrelax, b = 1.0 - o->scfrelax;
            RBCHK(rocblas_dscal(h, (int)n2, &b, dP, 1));
            RBCHK(rocblas_daxpy(h, (int)n2, &a, dPout, 1, dP, 1));
        }
        // Converged on an incremental or density-screened build: the energy
        // below takes J/K of this iteration's input density from one full,
        // unscreened build, so it carries no screening error.
        if (dchg < o->scfconv && (inc_now || dens_tol > 0.0)) {
            if ((rc = jk_device(ctx, dPprev, ctx->d_jc, ctx->d_kc, 0.0))) break;
            HIPCHK(hipStreamSynchronize(s));
            if ((rc = jk_collect_timing(ctx))) break;
            jk_total += ctx->last_jk_ms * 1e-3;
            build_quartets += ctx->last_quartets;
            have_energy = false;
        }
        // ENERGY_CALC with the updated density (src/pselfconsistentfield.f90:487-574);
        // with commutator DIIS the input density is idempotent (the previous
        // P_out) and the energy is that of the input density and its own J/K,
        // stationary to second order
        if (!have_energy) {   // four products, one synchronisation
            RBCHK(rocblas_set_pointer_mode(h, rocblas_pointer_mode_device));
            RBCHK(rocblas_ddot(h, (int)n2, dP, 1, dT, 1, dEn));
            RBCHK(rocblas_ddot(h, (int)n2, dP, 1, dV, 1, dEn + 1));
            RBCHK(rocblas_ddot(h, (int)n2, dP, 1, ctx->d_jc, 1, dEn + 2));
            RBCHK(rocblas_ddot(h, (int)n2, dP, 1, ctx->d_kc, 1, dEn + 3));
            RBCHK(rocblas_set_pointer_mode(h, rocblas_pointer_mode_host));
            double en[4];
            HIPCHK(hipMemcpyAsync(en, dEn, sizeof en, hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            ek = en[0]; ev = en[1]; ej = en[2]; ex = en[3];
        }
        ej *= 0.5;
        ex *= -0.25 * o->hfexchange;
        etot = ctx->e_nuc + ek + ev + ej + ex;
        if (!std::isfinite(etot)) { rc = fail(POLYERI_E_NONFINITE, "non-finite SCF energy"); break; }
        if (trace) { trace[2 * (iscf - 1)] = dchg; trace[2 * (iscf - 1) + 1] = etot; }
        if (dchg < o->scfconv) { converged = 1; break; }
        if (cdiis) HIPCHK(hipMemcpyAsync(dP, dPout, n2 * 8, hipMemcpyDeviceToDevice, s));
    }
    lap(4);
    if (!rc && last_purified) {   // orbital energies and coefficients of the last Fock matrix
        hipError_t e = hipMemcpyAsync(dFp, dFpLast, (size_t)nind * nind * 8, hipMemcpyDeviceToDevice, s);
        rc = e == hipSuccess ? diagonalise() : fail(POLYERI_E_HIP, hipGetErrorString(e));
    }
    lap(5);
    if (phases)
        fprintf(stderr, "polyeri_scf phases (s): J/K+Fock %.4f  F' %.4f  DIIS %.4f  density %.4f  rest %.4f  final %.4f\n",
                ph[0], ph[1], ph[2], ph[3], ph[4], ph[5]);
    (void)hipEventDestroy(ef0);
    (void)hipEventDestroy(ef1);
    if (rc) return rc;
    if (iscf > o->scfcycles) iscf = o->scfcycles;
    HIPCHK(hipMemcpyAsync(p_c, dP, n2 * 8, hipMemcpyDeviceToHost, s));
    if (co_out) HIPCHK(hipMemcpyAsync(co_out, dCO, (size_t)nc * nind * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (co_out) std::fill(co_out + (size_t)nc * nind, co_out + n2, 0.0);
    if (eps_out) std::copy(eps.begin(), eps.end(), eps_out);
    res->e_total = etot; res->e_nuc = ctx->e_nuc; res->e_kin = ek; res->e_nuc_att = ev;
    res->e_coul = ej; res->e_hfx = ex; res->fermi = fermi;
    res->iterations = iscf; res->converged = converged; res->n_indep = nind; res->iocc = iocc;
    res->jk_seconds = jk_total;
    res->quartets = build_quartets;
    res->fock_seconds = fock_total;
    res->scf_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    return converged ? 0 : fail(POLYERI_E_NOTCONV, "SCF CONVERGENCE NOT MET");
}

// ---------------------------------------------------------------------------
// MP2 (one k-point): AO integrals on the device, four quarter transforms
// ---------------------------------------------------------------------------
// Y(s, l, a, i) = X2(P(s, l), i, a): the half-transformed integrals (sl|ia)
// of the packed bra pair P(s, l) = max (max + 1) / 2 + min, X2 at
// (P no + i) nv + a, spread to Y at ((s n + l) nv + a) no + i.  One block per
// P: the contiguous no x nv block is read once, in 32 x 32 tiles through LDS,
// and written transposed (nv x no, contiguous) to (s, l) and (l, s).
__global__ void __launch_bounds__(256) k_mp2_unpack(const double* __restrict__ X2, int n, int no, int nv,
                                                    double* __restrict__ Y) {
    __shared__ double tile[32][33];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const size_t nb = (size_t)no * nv, npk = (size_t)n * (n + 1) / 2;
    const int ti = (no + 31) / 32, ta = (nv + 31) / 32;
    for (size_t P = blockIdx.x; P < npk; P += gridDim.x) {
        int s = (int)((std::sqrt(8.0 * (double)P + 1.0) - 1.0) * 0.5);
        while ((size_t)s * (s + 1) / 2 > P) --s;
        while ((size_t)(s + 1) * (s + 2) / 2 <= P) ++s;
        const int l = (int)(P - (size_t)s * (s + 1) / 2);
        const double* src = X2 + P * nb;
        double* d0 = Y + ((size_t)s * n + l) * nb;
        double* d1 = Y + ((size_t)l * n + s) * nb;
        for (int t = 0; t < ti * ta; ++t) {
            const int i0 = (t / ta) * 32, a0 = (t % ta) * 32;
            for (int r = ty; r < 32; r += 8) {
                const int i = i0 + r, a = a0 + tx;
                if (i < no && a < nv) tile[r][tx] = src[(size_t)i * nv + a];
            }
            __syncthreads();
            for (int r = ty; r < 32; r += 8) {
                const int a = a0 + r, i = i0 + tx;
                if (i < no && a < nv) {
                    const double v = tile[tx][r];
                    d0[(size_t)a * no + i] = v;
                    if (s != l) d1[(size_t)a * no + i] = v;
                }
            }
            __syncthreads();
        }
    }
}

// nqp > 0: also the quasi-particle bands (IOPTN(44)): the occupied-side
// transforms run over icore .. min(iocc + nqp, nall - ivirtcore) - 1 and the
// virtual side of the third quarter from max(iocc - nqp, icore), as the
// reference's TOK / TVOKK / TOVOKK ranges (src/pmp2energy.f90:140-154).
static int mp2_energy(polyeri_ctx* ctx, const double* co, const double* eps, int icore, int iocc, int nall,
                      int ivirtcore, int gemm, double* e2s, double* e2t, int nqp = 0, double* qeps = nullptr) {
    if (!ctx || !co || !eps || !e2s || !e2t) return fail(POLYERI_E_ARG, "polyeri_mp2_energy: null argument");
    const int hi = nall - ivirtcore;
    int no = iocc - icore, nv = hi - iocc;
    if (no <= 0 || nv <= 0 || icore < 0) return fail(POLYERI_E_ARG, "polyeri_mp2_energy: empty orbital space");
    if (nqp < 0 || (nqp > 0 && !qeps)) return fail(POLYERI_E_ARG, "polyeri_mp2_energy: bad quasi-particle request");
    if (gemm != POLYERI_GEMM_ROCBLAS && gemm != POLYERI_GEMM_MFMA_F64 && gemm != POLYERI_GEMM_MFMA_BF16X3)
        return fail(POLYERI_E_ARG, "polyeri_mp2_energy: unknown transform precision");
    const int ohi = std::min(iocc + nqp, hi), vlo = std::max(iocc - nqp, icore);
    const int nop = ohi - icore, nvp = hi - vlo, aoff = iocc - vlo;
    HIPCHK(hipSetDevice(ctx->device));
    int rc;
    if ((rc = ensure_blas(ctx))) return rc;
    rocblas_handle h = ctx->blas;
    hipStream_t s = ctx->stream;
    int n = ctx->ib.nbf, nc = ctx->rb.ncgs;
    // computed integrals: the tensor with its bra pair packed (half the
    // memory and half the scattered writes); stored records: the full tensor
    const bool packed = !ctx->use_stored;
    const size_t npk = (size_t)n * (n + 1) / 2;
    const size_t n4 = packed ? npk * n * n : (size_t)n * n * n * n;
    if (packed && (npk * n > (size_t)INT_MAX || (size_t)n * nvp * nop > (size_t)INT_MAX))
        return fail(POLYERI_E_ARG, "polyeri_mp2_energy: basis too large");
    DevBuf buf;
    double *dTen, *dCO, *dC, *dX1, *dX2, *dX3, *dT4, *dEps, *dAcc, *dQ, *dY = nullptr;
    // occupied-side indices first: the largest intermediate is nop n^3 (full
    // tensor) or nop n npk (packed)
    const size_t x1 = packed ? npk * nop * n : (size_t)nop * n * n * n;
    const size_t x2 = packed ? npk * nop * nvp : (size_t)nop * nop * n * n;
    const size_t x3 = (size_t)nop * nop * nvp * n, x4 = (size_t)nop * nop * nvp * nv;
    const size_t ny = packed ? (size_t)n * n * nvp * nop : 0;
    if ((rc = buf.get(&dTen, n4)) || (rc = buf.get(&dCO, (size_t)nc * nc)) || (rc = buf.get(&dC, (size_t)n * nc)) ||
        (rc = buf.get(&dX1, x1)) || (rc = buf.get(&dX2, x2)) || (rc = buf.get(&dX3, x3)) || (rc = buf.get(&dT4, x4)) ||
        (rc = buf.get(&dEps, nc)) || (rc = buf.get(&dAcc, 2)) || (rc = buf.get(&dQ, (size_t)nc)))
        return rc;
    if (packed) {   // Y reuses the tensor's memory (read by the first transform only) when it fits
        if (ny <= n4) dY = dTen;
        else if ((rc = buf.get(&dY, ny))) return rc;
    }
    HIPCHK(hipEventRecord(ctx->ev0, s));
    if ((rc = tensor_device(ctx, dTen, packed))) return rc;
    hipEvent_t et = nullptr;
    HIPCHK(hipEventCreate(&et));
    struct EvGuard { hipEvent_t e; ~EvGuard() { if (e) (void)hipEventDestroy(e); } } eg{et};
    HIPCHK(hipEventRecord(et, s));
    HIPCHK(hipMemcpyAsync(dCO, co, (size_t)nc * nc * 8, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dEps, eps, (size_t)nc * 8, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_mo_int, dim3(nblk((size_t)n * nc)), dim3(256), 0, s, dCO, nc, ctx->d_ref, ctx->d_scale, n, nc, dC);
    const double one = 1.0, zero = 0.0;
    const double* Cv = dC + (size_t)n * iocc;    // virtual columns (b)
    const double* Co = dC + (size_t)n * icore;   // occupied-side columns (j', i')
    const double* Ca = dC + (size_t)n * vlo;     // virtual-side columns (a')
    // Tensor A(s, l, u, m) = (sl|um) at s + n (l + n (u + n m)).  Its 8-fold
    // symmetry makes the reversed reading A[m + n (u + n (l + n s))] the same
    // tensor, so the first transform reads s as the slowest index: rows of n^3
    // consecutive elements, streamed in long runs.  Quarter transforms:
    //   X1(i, m, u, l) = sum_s Co(s, i) A(s, l, u, m)       X1[i + nop (m + n (u + n l))]
    //   X2(j, i, m, l) = sum_u X1(i, m, u, l) Co(u, j)      [batched over l]
    //   X3(j, i, a, m) = sum_l X2(j, i, m, l) Ca(l, a)      [batched over m]
    //   T(j, i, a, b)  = sum_m X3(j, i, a, m) Cv(m, b)  = (ia|jb)
    const long N = n, NO = nop, NA = nvp, NO2 = (long)nop * nop, N3 = (long)n * n * n;
    // Packed tensor A(P(s, l), u, m) = (sl|um) = A(P, m, u): the ket indices
    // first, the packed pair spread out after the second quarter:
    //   X1(P, m, i) = sum_u A(P, m, u) Co(u, i)      X1[(P n + m) nop + i]   [rows (P, m): one product]
    //   X2(P, i, a) = sum_m X1(P, m, i) Ca(m, a)     X2[(P nop + i) nvp + a] [batched over P]
    //   Y(s, l, a, i) = X2(P(s, l), i, a)            Y[((s n + l) nvp + a) nop + i]
    //   X3(j, i, a, l) = sum_s Co(s, j) Y(s, l, a, i)  X3[j + nop (i + nop (a + nvp l))]
    //   T(j, i, a, b) = sum_l X3(j, i, a, l) Cv(l, b)  = (jb|ia) = (ia|jb)
    const long NPK = (long)npk, NYC = (long)n * nvp * nop;
    auto unpack = [&]() -> int {
        const unsigned blocks = (unsigned)std::min<size_t>(npk, 65536);
        hipLaunchKernelGGL(k_mp2_unpack, dim3(blocks), dim3(256), 0, s, dX2, n, nop, nvp, dY);
        return hipGetLastError() == hipSuccess ? 0 : -2;
    };
    if (packed && gemm == POLYERI_GEMM_ROCBLAS) {
        RBCHK(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, nop, (int)(npk * n), n, &one, Co, n,
                            dTen, n, &zero, dX1, nop));
        RBCHK(rocblas_dgemm_strided_batched(h, rocblas_operation_transpose, rocblas_operation_transpose, nvp, nop, n,
                                            &one, Ca, n, 0, dX1, nop, (rocblas_stride)N * NO, &zero, dX2, nvp,
                                            (rocblas_stride)NO * NA, (int)npk));
        if (unpack()) return fail(POLYERI_E_HIP, "MP2 unpack launch failed");
        RBCHK(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_transpose, nop, (int)NYC, n, &one, Co, n,
                            dY, (int)NYC, &zero, dX3, nop));
        RBCHK(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, nop * nop * nvp, nv, n, &one, dX3,
                            nop * nop * nvp, Cv, n, &zero, dT4, nop * nop * nvp));
    } else if (packed) {
        const int mode = gemm == POLYERI_GEMM_MFMA_F64 ? GEMM_F64 : GEMM_BF16X3;
        int r = polyeri_gemm_launch(mode, (int)(npk * n), nop, n, dTen, N, 1, 0, Co, 1, N, 0, dX1, NO, 1, 0, 1, s);
        if (!r) r = polyeri_gemm_launch(mode, nop, nvp, n, dX1, 1, NO, N * NO, Ca, 1, N, 0, dX2, NA, 1, NO * NA,
                                        (int)npk, s);
        if (!r) r = unpack();
        if (!r) r = polyeri_gemm_launch(mode, nop, (int)NYC, n, Co, N, 1, 0, dY, NYC, 1, 0, dX3, 1, NO, 0, 1, s);
        if (!r) r = polyeri_gemm_launch(mode, nop * nop * nvp, nv, n, dX3, 1, NO2 * NA, 0, Cv, 1, N, 0, dT4, 1,
                                        NO2 * NA, 0, 1, s);
        if (r) return fail(POLYERI_E_HIP, "MP2 transform launch failed");
    } else if (gemm == POLYERI_GEMM_ROCBLAS) {
        // column-major views: X1 is nop x n^3, X2 (as its transpose) nop x (nop n) per l, X3 nop^2 x nvp per m
        RBCHK(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_transpose, nop, n * n * n, n, &one, Co, n,
                            dTen, n * n * n, &zero, dX1, nop));
        RBCHK(rocblas_dgemm_strided_batched(h, rocblas_operation_transpose, rocblas_operation_transpose, nop, nop * n, n,
                                            &one, Co, n, 0, dX1, nop * n, (rocblas_stride)nop * n * n, &zero, dX2, nop,
                                            (rocblas_stride)NO2 * n, n));
        RBCHK(rocblas_dgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, nop * nop, nvp, n, &one,
                                            dX2, nop * nop * n, (rocblas_stride)NO2, Ca, n, 0, &zero, dX3, nop * nop,
                                            (rocblas_stride)NO2 * nvp, n));
        RBCHK(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, nop * nop * nvp, nv, n, &one, dX3,
                            nop * nop * nvp, Cv, n, &zero, dT4, nop * nop * nvp));
    } else {
        // the same four quarter transforms on this library's MFMA kernels
        const int mode = gemm == POLYERI_GEMM_MFMA_F64 ? GEMM_F64 : GEMM_BF16X3;
        int r = polyeri_gemm_launch(mode, nop, n * n * n, n, Co, N, 1, 0, dTen, N3, 1, 0, dX1, 1, NO, 0, 1, s);
        if (!r) r = polyeri_gemm_launch(mode, nop * n, nop, n, dX1, 1, NO * N, NO * N * N, Co, 1, N, 0, dX2, NO, 1,
                                        NO2 * N, n, s);
        if (!r) r = polyeri_gemm_launch(mode, nop * nop, nvp, n, dX2, 1, NO2 * N, NO2, Ca, 1, N, 0, dX3, 1, NO2,
                                        NO2 * NA, n, s);
        if (!r) r = polyeri_gemm_launch(mode, nop * nop * nvp, nv, n, dX3, 1, NO2 * NA, 0, Cv, 1, N, 0, dT4, 1,
                                        NO2 * NA, 0, 1, s);
        if (r) return fail(POLYERI_E_HIP, "MP2 transform launch failed");
    }
    HIPCHK(hipEventRecord(ctx->ev1, s));
    HIPCHK(hipEventSynchronize(ctx->ev1));
    ctx->last_mp2_tensor_ms = event_ms(ctx->ev0, et);
    ctx->last_mp2_transform_ms = event_ms(et, ctx->ev1);
    ctx->last_mp2_flops =
        packed ? 2.0 * (double)npk * n * n * nop + 2.0 * (double)npk * nop * n * nvp +
                     2.0 * nop * (double)n * n * nvp * nop + 2.0 * nop * (double)nop * nvp * nv * n
               : 2.0 * nop * (double)n * n * n * n + 2.0 * nop * (double)nop * n * n * n +
                     2.0 * nop * (double)nop * nvp * n * n + 2.0 * nop * (double)nop * nvp * nv * n;
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(256), 0, s, dAcc, (size_t)2, 0.0);
    {
        const size_t tile = (size_t)no * (no + 1) * 8;
        const dim3 eg((unsigned)nv, (unsigned)std::min(nv, 16));
        if (tile <= 56 * 1024)
            hipLaunchKernelGGL(k_mp2_energy<true>, eg, dim3(256), tile, s, dT4, dEps, no, nv, nop, nvp, aoff, icore,
                               iocc, dAcc);
        else
            hipLaunchKernelGGL(k_mp2_energy<false>, eg, dim3(256), 0, s, dT4, dEps, no, nv, nop, nvp, aoff, icore,
                               iocc, dAcc);
    }
    HIPCHK(hipGetLastError());
    const int ng = nqp > 0 ? ohi - vlo : 0;
    if (ng > 0) {
        hipLaunchKernelGGL(k_fill, dim3(nblk((size_t)ng)), dim3(256), 0, s, dQ, (size_t)ng, 0.0);
        const size_t terms = (size_t)nv * no * (nv + no);
        const unsigned gx = (unsigned)std::min<size_t>(nblk(terms), 1024);
        hipLaunchKernelGGL(k_mp2_qp, dim3(gx, (unsigned)ng), dim3(256), 0, s, dT4, dEps, no, nv, nop, nvp, aoff, icore,
                           iocc, vlo, vlo, dQ);
        HIPCHK(hipGetLastError());
    }
    double acc[2];
    std::vector<double> q((size_t)std::max(ng, 1), 0.0);
    HIPCHK(hipMemcpyAsync(acc, dAcc, 16, hipMemcpyDeviceToHost, s));
    if (ng > 0) HIPCHK(hipMemcpyAsync(q.data(), dQ, (size_t)ng * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    *e2s = acc[0];
    *e2t = acc[1];
    if (!std::isfinite(acc[0]) || !std::isfinite(acc[1])) return fail(POLYERI_E_NONFINITE, "non-finite E(2)");
    if (qeps) {
        // QEPSILON = EPSILON + correction (src/pmp2energy.f90:549-558; one k-point: no k normalisation)
        for (int k = 0; k < nc; ++k) qeps[k] = eps[k];
        for (int k = 0; k < ng; ++k) qeps[vlo + k] = eps[vlo + k] + q[k];
    }
    return 0;
}

int polyeri_mp2_energy(polyeri_ctx* ctx, const double* co, const double* eps, int icore, int iocc, int nall,
                       int ivirtcore, double* e2s, double* e2t) {
    return mp2_energy(ctx, co, eps, icore, iocc, nall, ivirtcore, POLYERI_GEMM_MFMA_F64, e2s, e2t);
}

int polyeri_mp2_energy_ex(polyeri_ctx* ctx, const double* co, const double* eps, int icore, int iocc, int nall,
                          int ivirtcore, int transform, double* e2s, double* e2t) {
    return mp2_energy(ctx, co, eps, icore, iocc, nall, ivirtcore, transform, e2s, e2t);
}

int polyeri_mp2_energy_qp(polyeri_ctx* ctx, const double* co, const double* eps, int icore, int iocc, int nall,
                          int ivirtcore, int nqp, int transform, double* e2s, double* e2t, double* qeps) {
    if (nqp < 1) return fail(POLYERI_E_ARG, "polyeri_mp2_energy_qp: nqp must be positive");
    return mp2_energy(ctx, co, eps, icore, iocc, nall, ivirtcore, transform, e2s, e2t, nqp, qeps);
}

// ---------------------------------------------------------------------------
// Full MO integral set and coupled-cluster energies (src/pcc.f90)
// ---------------------------------------------------------------------------
// g[p][q][r][s] = (pq|rs) from the quarter-transform result
// G[b][a][i][j] = (ia|jb) of the four products below.
__global__ void k_mo_order(const double* __restrict__ G, double* __restrict__ g, int m) {
    const size_t m4 = (size_t)m * m * m * m;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < m4; t += (size_t)gridDim.x * blockDim.x) {
        const int sidx = (int)(t % m), r = (int)(t / m % m), q = (int)(t / ((size_t)m * m) % m),
                  p = (int)(t / ((size_t)m * m * m));
        g[t] = G[(((size_t)sidx * m + q) * m + p) * m + r];
    }
}

// INTEGRAL_TRANSFORM_DIRECTACCESS (src/pcc.f90:247-436) at one k-point: the
// AO tensor on the device and four quarter transforms over the MOs lo .. hi-1
// on the FP64 MFMA kernel (the MP2 transform with every index over the full
// active range), then reordered to dG[p][q][r][s] = (pq|rs).  Buffers are
// owned by buf; *ms_tensor / *ms_transform get the device times.
static int mo_tensor_device(polyeri_ctx* ctx, const double* co, int lo, int hi, DevBuf& buf, double** dG,
                            double* ms_tensor, double* ms_transform) {
    if (ctx->pbc) return fail(POLYERI_E_ARG, "MO integrals: periodic context (molecules only)");
    const int n = ctx->ib.nbf, nc = ctx->rb.ncgs, m = hi - lo;
    if (lo < 0 || m < 1 || hi > nc) return fail(POLYERI_E_ARG, "MO integrals: bad orbital range");
    const long N = n, M = m, N3 = (long)n * n * n, M2 = (long)m * m;
    if ((double)M2 * M2 >= 2147483648.0)
        return fail(POLYERI_E_ARG, "MO integrals: more than 2^31 MO integrals");
    hipStream_t s = ctx->stream;
    double *dTen, *dCO, *dC, *dX1, *dX2, *dX3, *dT;
    int rc;
    if ((rc = buf.get(&dTen, (size_t)N3 * n)) || (rc = buf.get(&dCO, (size_t)nc * nc)) ||
        (rc = buf.get(&dC, (size_t)n * nc)) || (rc = buf.get(&dX1, (size_t)m * N3)) ||
        (rc = buf.get(&dX2, (size_t)M2 * n * n)) || (rc = buf.get(&dX3, (size_t)M2 * m * n)) ||
        (rc = buf.get(&dT, (size_t)M2 * M2)) || (rc = buf.get(dG, (size_t)M2 * M2)))
        return rc;
    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventCreate(&e2));
    struct Ev { hipEvent_t a, b, c; ~Ev() { (void)hipEventDestroy(a); (void)hipEventDestroy(b); (void)hipEventDestroy(c); } } ev{e0, e1, e2};
    HIPCHK(hipEventRecord(e0, s));
    if ((rc = tensor_device(ctx, dTen))) return rc;
    HIPCHK(hipEventRecord(e1, s));
    HIPCHK(hipMemcpyAsync(dCO, co, (size_t)nc * nc * 8, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_mo_int, dim3(nblk((size_t)n * nc)), dim3(256), 0, s, dCO, nc, ctx->d_ref, ctx->d_scale, n, nc, dC);
    const double* Cm = dC + (size_t)n * lo;
    // X1(i, m, u, l) = sum_s C(s, i) A(s, l, u, m); X2(j, i, m, l); X3(j, i, a, m); T(j, i, a, b) = (ia|jb)
    int r = polyeri_gemm_launch(GEMM_F64, m, n * n * n, n, Cm, N, 1, 0, dTen, N3, 1, 0, dX1, 1, M, 0, 1, s);
    if (!r) r = polyeri_gemm_launch(GEMM_F64, m * n, m, n, dX1, 1, M * N, M * N * N, Cm, 1, N, 0, dX2, M, 1, M2 * N, n, s);
    if (!r) r = polyeri_gemm_launch(GEMM_F64, m * m, m, n, dX2, 1, M2 * N, M2, Cm, 1, N, 0, dX3, 1, M2, M2 * M, n, s);
    if (!r) r = polyeri_gemm_launch(GEMM_F64, m * m * m, m, n, dX3, 1, M2 * M, 0, Cm, 1, N, 0, dT, 1, M2 * M, 0, 1, s);
    if (r) return fail(POLYERI_E_HIP, "MO integral transform launch failed");
    hipLaunchKernelGGL(k_mo_order, dim3(nblk((size_t)M2 * M2)), dim3(256), 0, s, dT, *dG, m);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(e2, s));
    HIPCHK(hipEventSynchronize(e2));
    if (ms_tensor) *ms_tensor = event_ms(e0, e1);
    if (ms_transform) *ms_transform = event_ms(e1, e2);
    return 0;
}

int polyeri_mo_integrals(polyeri_ctx* ctx, const double* co, int lo, int hi, double* g) {
    if (!ctx || !co || !g) return fail(POLYERI_E_ARG, "polyeri_mo_integrals: null argument");
    HIPCHK(hipSetDevice(ctx->device));
    DevBuf buf;
    double* dG = nullptr;
    int rc = mo_tensor_device(ctx, co, lo, hi, buf, &dG, nullptr, nullptr);
    if (rc) return rc;
    const size_t m = (size_t)(hi - lo);
    HIPCHK(hipMemcpyAsync(g, dG, m * m * m * m * 8, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return 0;
}

int polyeri_cc(polyeri_ctx* ctx, const char* method, const double* co, const double* eps, int icore, int iocc,
               int nall, int ivirtcore, double conv, int maxiter, int ndiis, polyeri_cc_result* res) {
    if (!ctx || !method || !co || !eps || !res) return fail(POLYERI_E_ARG, "polyeri_cc: null argument");
    static const char* const names[] = {"CCSD", "CCD", "LCCD", "LCCSD", "QCISD"};
    int m = -1;
    for (int q = 0; q < 5; ++q)
        if (std::string(method) == names[q]) m = q;
    if (m < 0) return fail(POLYERI_E_ARG, "UNKNOWN OR UNIMPLEMENTED COUPLED-CLUSTER METHOD");
    const int hi = nall - ivirtcore, no = iocc - icore;
    if (icore < 0 || no < 1 || ivirtcore < 0 || hi <= iocc || nall > ctx->rb.ncgs)
        return fail(POLYERI_E_ARG, "polyeri_cc: empty occupied or virtual space");
    if (!(conv > 0.0) || maxiter < 1) return fail(POLYERI_E_ARG, "polyeri_cc: bad convergence request");
    auto t0 = std::chrono::steady_clock::now();
    HIPCHK(hipSetDevice(ctx->device));
    DevBuf buf;
    double* dG = nullptr;
    double mst = 0.0, msx = 0.0;
    int rc = mo_tensor_device(ctx, co, icore, hi, buf, &dG, &mst, &msx);
    if (rc) return rc;
    CcSolveOut o{};
    std::string err;
    rc = polyeri_cc_solve(dG, hi - icore, no, eps + icore, m, conv, maxiter, ndiis, &o, ctx->stream, &err);
    if (rc) return fail(rc == -2 ? POLYERI_E_NOMEM : rc == -1 ? POLYERI_E_ARG : POLYERI_E_HIP, "polyeri_cc: " + err);
    res->e2 = o.e2;
    res->e3 = o.e3;
    res->e_singles = o.e_singles;
    res->e_doubles = o.e_doubles;
    res->e_corr = o.e_corr;
    res->residual = o.residual;
    res->iterations = o.iterations;
    res->converged = o.converged;
    res->tensor_ms = mst;
    res->transform_ms = msx;
    res->iter_ms = o.iter_ms;
    res->iter_gflop = o.iter_gflop;
    res->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (!std::isfinite(res->e_corr)) return fail(POLYERI_E_NONFINITE, "polyeri_cc: non-finite energy");
    // the reference aborts here (src/pcc.f90:218); res holds the last iterate
    if (!res->converged) return fail(POLYERI_E_NOTCONV, "CC ITERATION DID NOT CONVERGE");
    return 0;
}

int polyeri_ccsd(polyeri_ctx* ctx, const double* co, const double* eps, int icore, int iocc, int nall, int ivirtcore,
                 double conv, int maxiter, int ndiis, polyeri_cc_result* res) {
    return polyeri_cc(ctx, "CCSD", co, eps, icore, iocc, nall, ivirtcore, conv, maxiter, ndiis, res);
}

// ---------------------------------------------------------------------------
// Periodic SCF (1-D, HELIX = 0): SELF_CONSISTENT_FIELD with the k-point
// FOCK_BUILD of src/pfock.f90:180-470
// ---------------------------------------------------------------------------
// S_C, T_C, N_C of every cell Q in [-cel1, cel1] (reference layout, cell Q at
// block Q + cel1): the pair (i in cell 0, j in cell Q) from a two-cell basis,
// the nuclear attraction summed over the nuclei of cells -cel2..cel2
// (src/ponee.f90:486-560); and NUCLEAR_REPULSION_CALC per unit cell
// (src/pselfconsistentfield.f90:652-702).
static void one_electron_cells(const polyeri_ctx* c, std::vector<double>& S, std::vector<double>& T,
                               std::vector<double>& N, double& enuc) {
    const IntBasis& ib = c->ib;
    const int nbf = ib.nbf, nc = c->rb.ncgs, ncell = 2 * c->cel1 + 1;
    const size_t c2 = (size_t)nc * nc;
    S.assign(c2 * ncell, 0.0); T.assign(c2 * ncell, 0.0); N.assign(c2 * ncell, 0.0);
    std::vector<double> boys;
    boys_table(boys);
    std::vector<int> z;
    std::vector<double> xyz;
    for (int k = -c->cel2; k <= c->cel2; ++k)
        for (int a = 0; a < c->rb.natom; ++a) {
            z.push_back(c->rb.iatom[a]);
            xyz.push_back(c->rb.ax[a] + k * c->period);
            xyz.push_back(c->rb.ay[a]);
            xyz.push_back(c->rb.az[a]);
        }
    for (int q = -c->cel1; q <= c->cel1; ++q) {
        IntBasis two = ib;
        for (const Shell& sh : ib.shells) {
            Shell t = sh;
            t.x += q * c->period;
            t.f0 += nbf;
            two.shells.push_back(t);
        }
        two.nbf = 2 * nbf;
        two.natom_z = z;
        two.atom_xyz = xyz;
        std::vector<double> s2, t2, v2;
        one_electron(two, boys, s2, t2, v2);
        const size_t off = (size_t)(q + c->cel1) * c2;
        for (int a = 0; a < nbf; ++a)
            for (int b = 0; b < nbf; ++b) {
                const size_t r = off + ib.ref_of_int[a] + (size_t)nc * ib.ref_of_int[b];
                const size_t e = (size_t)a * two.nbf + nbf + b;
                const double f = ib.scale[a] * ib.scale[b];
                S[r] = f * s2[e]; T[r] = f * t2[e]; N[r] = f * v2[e];
            }
    }
    enuc = 0.0;
    for (int i = 0; i < c->rb.natom; ++i)
        for (int j = 0; j < c->rb.natom; ++j)
            for (int k = -c->cel2; k <= c->cel2; ++k) {
                if (k == 0 && i == j) continue;
                const double dx = c->rb.ax[j] + k * c->period - c->rb.ax[i], dy = c->rb.ay[j] - c->rb.ay[i],
                             dz = c->rb.az[j] - c->rb.az[i];
                enuc += 0.5 * c->rb.iatom[i] * c->rb.iatom[j] / std::sqrt(dx * dx + dy * dy + dz * dz);
            }
}

// Mk[k](K,J) = sum_Q M(K,J,Q) exp(i pi IX Q / kvc), IX = k - kvc; Hermitian part.
__global__ void k_bloch(const double* __restrict__ M, int nc, int ncell, int cel1, int nk, int kvc,
                        rocblas_double_complex* __restrict__ Mk) {
    const size_t c2 = (size_t)nc * nc;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= c2 * nk) return;
    const int k = (int)(t / c2);
    const size_t e = t % c2;
    const int K = (int)(e % nc), J = (int)(e / nc);
    const double ph = kvc > 0 ? POLY_PI * (double)(k - kvc) / kvc : 0.0;
    double re = 0.0, im = 0.0;
    for (int q = -cel1; q <= cel1; ++q) {
        const double a = M[(size_t)(q + cel1) * c2 + K + (size_t)nc * J];
        const double b = M[(size_t)(q + cel1) * c2 + J + (size_t)nc * K];
        double sn, cs;
        sincos(ph * q, &sn, &cs);
        // (M(K,J) e^{i ph q} + conj(M(J,K) e^{i ph q})) / 2
        re += 0.5 * (a + b) * cs;
        im += 0.5 * (a - b) * sn;
    }
    Mk[t] = rocblas_double_complex(re, im);
}

// F = T + V + J - hfx K / 2 elementwise over every cell (the Bloch sums
// take the Hermitian part, src/pfock.f90:216-226)
__global__ void k_fock_flat(const double* __restrict__ T, const double* __restrict__ V, const double* __restrict__ J,
                            const double* __restrict__ K, double hfx, size_t n, double* __restrict__ F) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) F[t] = T[t] + V[t] + J[t] - 0.5 * hfx * K[t];
}

// P(L,K,Q) = w Re sum_k V_k(L,K) exp(-i pi IX Q / kvc), V_k = conj(W_k)
__global__ void k_unbloch(const rocblas_double_complex* __restrict__ Wk, int nc, int ncell, int cel1, int nk, int kvc,
                          double w, double* __restrict__ P) {
    const size_t c2 = (size_t)nc * nc;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= c2 * ncell) return;
    const int q = (int)(t / c2) - cel1;
    const size_t e = t % c2;
    double s = 0.0;
    for (int k = 0; k < nk; ++k) {
        const double ph = kvc > 0 ? POLY_PI * (double)(k - kvc) / kvc : 0.0;
        double sn, cs;
        sincos(ph * q, &sn, &cs);
        const rocblas_double_complex v = Wk[(size_t)k * c2 + e];
        s += v.real() * cs + v.imag() * sn;
    }
    P[t] = w * s;
}

// Orthogonaliser columns: X(:,j) = U(:,j)/sqrt(e_j) for e_j > lindep, else 0.
__global__ void k_orth_c(const rocblas_double_complex* __restrict__ U, const double* __restrict__ e, int nc, int nk,
                         double lindep, rocblas_double_complex* __restrict__ X) {
    const size_t c2 = (size_t)nc * nc;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= c2 * nk) return;
    const int k = (int)(t / c2), j = (int)((t % c2) / nc);
    const double ev = e[(size_t)k * nc + j];
    X[t] = ev > lindep ? U[t] * (1.0 / sqrt(ev)) : rocblas_double_complex(0.0, 0.0);
}

// Columns dropped by the orthogonaliser get a large diagonal so their
// eigenvalues sort above every physical one.
__global__ void k_shift_dropped(rocblas_double_complex* __restrict__ F, const double* __restrict__ e, int nc, int nk,
                                double lindep) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nc * nk) return;
    const int k = t / nc, j = t % nc;
    if (!(e[(size_t)k * nc + j] > lindep))
        F[(size_t)k * nc * nc + j + (size_t)nc * j] = rocblas_double_complex(1.0e6, 0.0);
}

// CO(:,l) scaled by occupation: Y(K,l) = occ(k,l) CO(K,l)
__global__ void k_scale_occ(const rocblas_double_complex* __restrict__ C, const double* __restrict__ occ, int nc,
                            int nk, rocblas_double_complex* __restrict__ Y) {
    const size_t c2 = (size_t)nc * nc;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= c2 * nk) return;
    const int k = (int)(t / c2), l = (int)((t % c2) / nc);
    Y[t] = C[t] * occ[(size_t)k * nc + l];
}

__global__ void k_real_to_complex(const double* __restrict__ a, size_t n, rocblas_double_complex* __restrict__ b) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) b[t] = rocblas_double_complex(a[t], 0.0);
}

int polyeri_scf_periodic(polyeri_ctx* ctx, const polyeri_scf_opts* o, int kvc, double* p_c, double* eps_out,
                         double* trace, polyeri_scf_result* res, double* co_out, int* nall_out) {
    if (!ctx || !o || !p_c || !res) return fail(POLYERI_E_ARG, "polyeri_scf_periodic: null argument");
    if (!ctx->pbc) return fail(POLYERI_E_ARG, "polyeri_scf_periodic: not a periodic context");
    if (o->scfcycles < 1 || kvc < 0 || kvc > 512) return fail(POLYERI_E_ARG, "polyeri_scf_periodic: bad options");
    auto t_start = std::chrono::steady_clock::now();
    HIPCHK(hipSetDevice(ctx->device));
    int rc;
    if ((rc = ensure_blas(ctx))) return rc;
    rocblas_handle h = ctx->blas;
    hipStream_t s = ctx->stream;
    const int nc = ctx->rb.ncgs, ncell = 2 * ctx->cel1 + 1, nk = kvc > 0 ? 2 * kvc : 1;
    const size_t c2 = (size_t)nc * nc, nv = c2 * ncell, ck = c2 * nk;
    // one-electron cells and nuclear repulsion (host), C2S
    std::vector<double> Sh, Th, Nh, c2s;
    double enuc = 0.0;
    one_electron_cells(ctx, Sh, Th, Nh, enuc);
    if ((rc = form_c2s(ctx->rb, o->spherical, c2s))) return rc;
    DevBuf buf;
    double *dS, *dT, *dN, *dF, *dP, *dPout, *dE, *dOcc, *dSE;
    rocblas_double_complex *dSk, *dFk, *dW1, *dW2, *dX, *dC2S, *dCk;
    int* dInfo;
    double* dWk;
    if ((rc = buf.get(&dS, nv)) || (rc = buf.get(&dT, nv)) || (rc = buf.get(&dN, nv)) || (rc = buf.get(&dF, nv)) ||
        (rc = buf.get(&dP, nv)) || (rc = buf.get(&dPout, nv)) || (rc = buf.get(&dE, (size_t)nc * nk)) ||
        (rc = buf.get(&dSE, (size_t)nc * nk)) || (rc = buf.get(&dOcc, (size_t)nc * nk)) || (rc = buf.get(&dSk, ck)) ||
        (rc = buf.get(&dFk, ck)) || (rc = buf.get(&dW1, ck)) || (rc = buf.get(&dW2, ck)) || (rc = buf.get(&dX, ck)) ||
        (rc = buf.get(&dC2S, c2)) || (rc = buf.get(&dCk, ck)) || (rc = buf.get(&dInfo, nk)) ||
        (rc = buf.get(&dWk, (size_t)nc * nk)))
        return rc;
    const int HMAX = 9;
    double *dHP, *dHD;
    if ((rc = buf.get(&dHP, nv * HMAX)) || (rc = buf.get(&dHD, nv * HMAX))) return rc;
    HIPCHK(hipMemcpy(dS, Sh.data(), nv * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dT, Th.data(), nv * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dN, Nh.data(), nv * 8, hipMemcpyHostToDevice));
    {
        double* dtmp;
        if ((rc = buf.get(&dtmp, c2))) return rc;
        HIPCHK(hipMemcpy(dtmp, c2s.data(), c2 * 8, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_real_to_complex, dim3(nblk(c2)), dim3(256), 0, s, dtmp, c2, dC2S);
    }
    const rocblas_double_complex one(1.0, 0.0), zero(0.0, 0.0);
    const rocblas_stride sc2 = (rocblas_stride)c2;
    // S(k), C2S S(k) C2S^T, eigenvectors -> orthogonaliser X(k) (src/pfock.f90:236-281 per k)
    hipLaunchKernelGGL(k_bloch, dim3(nblk(ck)), dim3(256), 0, s, dS, nc, ncell, ctx->cel1, nk, kvc, dSk);
    RBCHK(rocblas_zgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, nc, nc, nc, &one, dC2S, nc,
                                        0, dSk, nc, sc2, &zero, dW1, nc, sc2, nk));
    RBCHK(rocblas_zgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_transpose, nc, nc, nc, &one, dW1,
                                        nc, sc2, dC2S, nc, 0, &zero, dW2, nc, sc2, nk));
    RBCHK(rocsolver_zheevd_strided_batched(h, rocblas_evect_original, rocblas_fill_upper, nc, dW2, nc, sc2, dSE, nc,
                                           dWk, nc, dInfo, nk));
    hipLaunchKernelGGL(k_orth_c, dim3(nblk(ck)), dim3(256), 0, s, dW2, dSE, nc, nk, o->lindep, dX);
    std::vector<double> se((size_t)nc * nk);
    HIPCHK(hipMemcpyAsync(se.data(), dSE, se.size() * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    std::vector<int> nind(nk, 0);
    for (int k = 0; k < nk; ++k)
        for (int j = 0; j < nc; ++j) nind[k] += se[(size_t)k * nc + j] > o->lindep;
    int nel = 0;
    for (int z : ctx->rb.iatom) nel += z;
    nel -= o->charge;
    if (nel % 2) return fail(POLYERI_E_ARG, "ODD NUMBER OF ELECTRONS");
    // guess
    std::vector<double> P(nv, 0.0);
    if (o->guess == 1) {
        std::copy(p_c, p_c + nv, P.begin());
    } else {
        // GUESS_DENSITYMATRIX in cell 0 (src/pdensitymatrix.f90:83-127)
        std::vector<int> fatom(nc, -1);
        for (int c = 0; c < nc; ++c) {
            const int mu = ctx->ib.int_of_ref[c];
            if (mu < 0) continue;
            for (const Shell& sh : ctx->ib.shells)
                if (mu >= sh.f0 && mu < sh.f0 + ncart(sh.L))
                    for (int a = 0; a < ctx->rb.natom; ++a)
                        if (sh.x == ctx->rb.ax[a] && sh.y == ctx->rb.ay[a] && sh.z == ctx->rb.az[a]) fatom[c] = a;
        }
        for (int a = 0; a < ctx->rb.natom; ++a) {
            int L = 0;
            for (int c = 0; c < nc; ++c) L += fatom[c] == a;
            if (!L) return fail(POLYERI_E_ARG, "atom without basis functions");
            for (int c = 0; c < nc; ++c)
                if (fatom[c] == a) P[(size_t)ctx->cel1 * c2 + c + (size_t)nc * c] = (double)ctx->rb.iatom[a] / L;
        }
    }
    HIPCHK(hipMemcpyAsync(dP, P.data(), nv * 8, hipMemcpyHostToDevice, s));
    const double temphf = o->temphf > 0 ? o->temphf : 1.0e-4;
    std::vector<double> pulay((size_t)(o->scfcycles + 2) * (o->scfcycles + 2), 0.0);
    const int ld = o->scfcycles + 2;
    std::vector<double> eps((size_t)nc * nk), occ((size_t)nc * nk);
    double etot = 0, ek = 0, ev = 0, ej = 0, ex = 0, fermi = 0, jk_total = 0;
    long long quartets = 0;
    int converged = 0, iscf;
    for (iscf = 1; iscf <= o->scfcycles; ++iscf) {
        if ((rc = jk_device_cells(ctx, dP, ctx->d_jcc, ctx->d_kcc))) break;
        hipLaunchKernelGGL(k_fock_flat, dim3(nblk(nv)), dim3(256), 0, s, dT, dN, ctx->d_jcc, ctx->d_kcc,
                           o->hfexchange, nv, dF);
        HIPCHK(hipGetLastError());
        hipLaunchKernelGGL(k_bloch, dim3(nblk(ck)), dim3(256), 0, s, dF, nc, ncell, ctx->cel1, nk, kvc, dFk);
        // C2S F(k) C2S^T, X^H (.) X
        RBCHK(rocblas_zgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, nc, nc, nc, &one, dC2S,
                                            nc, 0, dFk, nc, sc2, &zero, dW1, nc, sc2, nk));
        RBCHK(rocblas_zgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_transpose, nc, nc, nc, &one,
                                            dW1, nc, sc2, dC2S, nc, 0, &zero, dW2, nc, sc2, nk));
        RBCHK(rocblas_zgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, nc, nc, nc, &one, dW2,
                                            nc, sc2, dX, nc, sc2, &zero, dW1, nc, sc2, nk));
        RBCHK(rocblas_zgemm_strided_batched(h, rocblas_operation_conjugate_transpose, rocblas_operation_none, nc, nc,
                                            nc, &one, dX, nc, sc2, dW1, nc, sc2, &zero, dFk, nc, sc2, nk));
        hipLaunchKernelGGL(k_shift_dropped, dim3(nblk((size_t)nc * nk)), dim3(256), 0, s, dFk, dSE, nc, nk, o->lindep);
        RBCHK(rocsolver_zheevd_strided_batched(h, rocblas_evect_original, rocblas_fill_upper, nc, dFk, nc, sc2, dE, nc,
                                               dWk, nc, dInfo, nk));
        // C(k) = X C'(k), CO(k) = C2S^T C(k)
        RBCHK(rocblas_zgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_none, nc, nc, nc, &one, dX, nc,
                                            sc2, dFk, nc, sc2, &zero, dW1, nc, sc2, nk));
        RBCHK(rocblas_zgemm_strided_batched(h, rocblas_operation_transpose, rocblas_operation_none, nc, nc, nc, &one,
                                            dC2S, nc, 0, dW1, nc, sc2, &zero, dCk, nc, sc2, nk));
        std::vector<int> info(nk, 0);
        HIPCHK(hipMemcpyAsync(eps.data(), dE, eps.size() * 8, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(info.data(), dInfo, nk * sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if ((rc = jk_collect_timing(ctx))) break;
        jk_total += ctx->last_jk_ms * 1e-3;
        quartets += ctx->quartets;   // every pair item runs over the 2 cel2 + 1 cells Q3
        bool bad = false;
        for (int k = 0; k < nk; ++k) bad |= info[k] != 0;
        if (bad) { rc = fail(POLYERI_E_LIB, "EIGENVECTORS NOT CONVERGED"); break; }
        for (int k = 0; k < nk; ++k)
            for (int j = nind[k]; j < nc; ++j) eps[(size_t)k * nc + j] = 1.0e99;
        // Fermi level over every k (src/pfock.f90:339-423), weights 1/nk
        double mine = 1e99, maxe = -1e99;
        for (double e : eps) { maxe = std::max(maxe, e); mine = std::min(mine, e); }
        double f1 = mine, f2 = maxe, f12 = 0.5 * (mine + maxe);
        auto count = [&](double mu) {
            double nn = 0;
            for (double e : eps) nn += 2.0 / (std::exp((e - mu) / kBoltzmann / temphf) + 1.0);
            return nn / nk - nel;
        };
        for (int it = 0; it < 100000; ++it) {
            const double n1 = count(f1), nn2 = count(f2), n12 = count(f12);
            if (std::fabs(n12) < 1e-10) { fermi = f12; break; }
            else if (n1 > 0 && nn2 > 0) { f2 = f1; f1 -= 0.1; }
            else if (n1 < 0 && nn2 < 0) { f1 = f2; f2 += 0.1; }
            else if (n1 < 0 && n12 >= 0) f2 = f12;
            else if (n12 <= 0 && nn2 > 0) f1 = f12;
            else break;
            f12 = 0.5 * (f1 + f2);
        }
        for (size_t i = 0; i < eps.size(); ++i) occ[i] = 2.0 / (std::exp((eps[i] - fermi) / kBoltzmann / temphf) + 1.0);
        HIPCHK(hipMemcpyAsync(dOcc, occ.data(), occ.size() * 8, hipMemcpyHostToDevice, s));
        // W_k(K,J) = sum_l occ conj(CO(K,l)) CO(J,l): conj(CO) occ CO^T; P_out(L,K,Q) = Re sum w W_k(L,K) e^{ikQ}
        hipLaunchKernelGGL(k_scale_occ, dim3(nblk(ck)), dim3(256), 0, s, dCk, dOcc, nc, nk, dW1);
        // W2 = Y CO^H = conj(W_k), so Re(W_k e^{ikQ}) = Re(W2 e^{-ikQ})
        RBCHK(rocblas_zgemm_strided_batched(h, rocblas_operation_none, rocblas_operation_conjugate_transpose, nc, nc,
                                            nc, &one, dW1, nc, sc2, dCk, nc, sc2, &zero, dW2, nc, sc2, nk));
        hipLaunchKernelGGL(k_unbloch, dim3(nblk(nv)), dim3(256), 0, s, dW2, nc, ncell, ctx->cel1, nk, kvc, 1.0 / nk,
                           dPout);
        HIPCHK(hipGetLastError());
        // STORE_DENSITYMATRIX / DIIS / RELAX over all cells (src/pdensitymatrix.f90:131-339)
        const int slot = (iscf - 1) % HMAX;
        double* hp = dHP + nv * slot;
        double* hd = dHD + nv * slot;
        HIPCHK(hipMemcpyAsync(hp, dP, nv * 8, hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(hd, dPout, nv * 8, hipMemcpyDeviceToDevice, s));
        const double mone = -1.0;
        RBCHK(rocblas_daxpy(h, (int)nv, &mone, dP, 1, hd, 1));
        RBCHK(rocblas_set_pointer_mode(h, rocblas_pointer_mode_host));
        double dchg = 0.0;
        RBCHK(rocblas_ddot(h, (int)nv, hd, 1, hd, 1, &dchg));
        pulay[iscf + (size_t)ld * iscf] = dchg;
        for (int i = std::max(1, iscf - HMAX + 1); i < iscf; ++i) {
            double v = 0.0;
            RBCHK(rocblas_ddot(h, (int)nv, dHD + nv * ((i - 1) % HMAX), 1, hd, 1, &v));
            pulay[i + (size_t)ld * iscf] = v;
            pulay[iscf + (size_t)ld * i] = v;
        }
        dchg = std::sqrt(dchg / (double)nv);
        if (o->diis != 0 && iscf >= o->diis && ((iscf - o->diis) % 2) == 0) {
            int nt = std::min(8, iscf), n = nt + 1;
            std::vector<double> B((size_t)n * n, 0.0), rhs(n, 0.0), cv;
            for (int i = 0; i < nt; ++i) {
                for (int j = 0; j < nt; ++j) B[j + (size_t)n * i] = pulay[(iscf - j) + (size_t)ld * (iscf - i)];
                B[i + (size_t)n * nt] = -1.0;
                B[nt + (size_t)n * i] = -1.0;
            }
            rhs[nt] = -1.0;
            lu_solve(n, B, rhs, cv);
            hipLaunchKernelGGL(k_fill, dim3(nblk(nv)), dim3(256), 0, s, dP, nv, 0.0);
            for (int i = iscf - nt + 1; i <= iscf; ++i) {
                double w = cv[iscf - i];
                RBCHK(rocblas_daxpy(h, (int)nv, &w, dHP + nv * ((i - 1) % HMAX), 1, dP, 1));
            }
        } else {
            double a = o->scfrelax, b = 1.0 - o->scfrelax;
            RBCHK(rocblas_dscal(h, (int)nv, &b, dP, 1));
            RBCHK(rocblas_daxpy(h, (int)nv, &a, dPout, 1, dP, 1));
        }
        // ENERGY_CALC over the cells (src/pselfconsistentfield.f90:487-574)
        RBCHK(rocblas_ddot(h, (int)nv, dP, 1, dT, 1, &ek));
        RBCHK(rocblas_ddot(h, (int)nv, dP, 1, dN, 1, &ev));
        RBCHK(rocblas_ddot(h, (int)nv, dP, 1, ctx->d_jcc, 1, &ej));
        RBCHK(rocblas_ddot(h, (int)nv, dP, 1, ctx->d_kcc, 1, &ex));
        ej *= 0.5;
        ex *= -0.25 * o->hfexchange;
        etot = enuc + ek + ev + ej + ex;
        if (!std::isfinite(etot)) { rc = fail(POLYERI_E_NONFINITE, "non-finite SCF energy"); break; }
        if (trace) { trace[2 * (iscf - 1)] = dchg; trace[2 * (iscf - 1) + 1] = etot; }
        if (dchg < o->scfconv) { converged = 1; break; }
    }
    if (rc) return rc;
    if ((rc = check_fx_overflow(ctx))) return rc;
    if (iscf > o->scfcycles) iscf = o->scfcycles;
    HIPCHK(hipMemcpyAsync(p_c, dP, nv * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (eps_out) std::copy(eps.begin(), eps.end(), eps_out);
    if (co_out) {
        HIPCHK(hipMemcpyAsync(co_out, dCk, ck * sizeof(rocblas_double_complex), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    if (nall_out) std::copy(nind.begin(), nind.end(), nall_out);
    memset(res, 0, sizeof(*res));
    res->e_total = etot; res->e_nuc = enuc; res->e_kin = ek; res->e_nuc_att = ev;
    res->e_coul = ej; res->e_hfx = ex; res->fermi = fermi;
    res->iterations = iscf; res->converged = converged; res->n_indep = nind[0];
    res->iocc = nel / 2;
    res->jk_seconds = jk_total;
    res->quartets = quartets;
    res->scf_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    if (!converged) return fail(POLYERI_E_NOTCONV, "SCF CONVERGENCE NOT MET");
    return 0;
}

// ---------------------------------------------------------------------------
// k-point MP2 (MP2_ENERGY_ALG1 with HELIX = 0, MODULO 1; src/pmp2energy.f90:1-548)
// ---------------------------------------------------------------------------
// The stored lattice integrals of STORE_FOURINDEX_ERI (src/pfourindexstore.f90:
// 258-747) computed on the device: T[i][Q1][j][Q3][k][Q2][l] = (i_0 j_Q1 |
// k_Q3 l_Q3+Q2), Q1, Q2 in [-rc1, rc1], Q3 in [-cel2, cel2], internal
// function order; quartets below the context's Schwarz threshold are zero.
static int tensor_device_cells(polyeri_ctx* c, double* dT, size_t nt) {
    hipLaunchKernelGGL(k_fill, dim3(nblk(nt)), dim3(256), 0, c->stream, dT, nt, 0.0);
    HIPCHK(hipGetLastError());
    const long long n3 = 2LL * c->cel2 + 1;
    for (const ClassPairWork& cw : c->work) {
        int rc = launch_class_pbc(c, cw, cw.nwork * n3, dT);
        if (rc) return rc;
    }
    return 0;
}

// CO(k) (reference order, complex, column-major per k) -> planar internal
// C[k][f][p] = scale(f) CO(ref(f), p; k).
__global__ void k_co_planar(const double2* __restrict__ co, int nc, const int* __restrict__ ref,
                            const double* __restrict__ sc, int nbf, int nk, double* __restrict__ Cr,
                            double* __restrict__ Ci) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)nk * nbf * nc) return;
    const int p = (int)(t % nc);
    const size_t r = t / nc;
    const int f = (int)(r % nbf), k = (int)(r / nbf);
    const double2 v = co[((size_t)k * nc + p) * nc + ref[f]];
    Cr[t] = sc[f] * v.x;
    Ci[t] = sc[f] * v.y;
}

// Band operand of one quarter transform:
//   X[t][qi][f][p] = conj?(C[k][f][p0 + p]) exp(i pi q (sk (k - kvc) + sf (kf - kvc)) / kvc)
// with k = (k0 + t) mod nk, q = qi - qh; zero for p0 + p >= plim[k] (the
// bands IALL(k) - IVIRTCORE kept at that k).  kvc = 0: the Gamma point, no phase.
__global__ void k_mp2k_operand(const double* __restrict__ Cr, const double* __restrict__ Ci, int nk, int nbf,
                               int nc, int kvc, int k0, int nt, int nq, int qh, int np, int p0, int sk, int kf,
                               int sf, int conj, const int* __restrict__ plim, double* __restrict__ Xr,
                               double* __restrict__ Xi) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)nt * nq * nbf * np) return;
    const int p = (int)(t % np);
    size_t r = t / np;
    const int f = (int)(r % nbf); r /= nbf;
    const int qi = (int)(r % nq);
    const int tt = (int)(r / nq);
    const int k = (k0 + tt) % nk;
    if (p0 + p >= plim[k]) { Xr[t] = 0.0; Xi[t] = 0.0; return; }
    const size_t e = ((size_t)k * nbf + f) * nc + p0 + p;
    const double cr = Cr[e], ci = conj ? -Ci[e] : Ci[e];
    double sn = 0.0, cs = 1.0;
    if (kvc > 0) {
        const int q = qi - qh;
        // integer phase numerator, reduced mod 2 kvc before the angle is formed
        long long m = (long long)q * (sk * (k - kvc) + sf * (kf - kvc));
        m %= 2LL * kvc;
        sincospi((double)m / kvc, &sn, &cs);
    }
    Xr[t] = cr * cs - ci * sn;
    Xi[t] = cr * sn + ci * cs;
}

// E(2) of one k_b (src/pmp2energy.f90:395-433) from
// TOVOKK[kj][ka][b][jo][a'][io] (planar; jo, io < nop from ICORE, a' < nvp
// from band iocc - aoff), T2 = TOVOKK(j, a, i) at (ka, ki), ki = ka - kj + kb
// (mod nk); eps[k][band]; the sum runs over jo, io < no and the virtual
// a < IALL(ka) - IVIRTCORE - IOCC.
__global__ void k_mp2k_energy(const double* __restrict__ Tr, const double* __restrict__ Ti,
                              const double* __restrict__ eps, const int* __restrict__ plim, int nk, int kb, int nb,
                              int no, int nv, int nop, int nvp, int aoff, int nc, int icore, int iocc,
                              double* __restrict__ acc) {
    const size_t tot = (size_t)nk * nk * nb * no * nv * no;
    double es = 0.0, et = 0.0;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < tot; t += (size_t)gridDim.x * blockDim.x) {
        const int io = (int)(t % no);
        size_t r = t / no;
        const int a = (int)(r % nv); r /= nv;
        const int jo = (int)(r % no); r /= no;
        const int b = (int)(r % nb); r /= nb;
        const int ka = (int)(r % nk);
        const int kj = (int)(r / nk);
        if (iocc + a >= plim[ka]) continue;
        const int ki = (ka - kj + kb + nk) % nk;
        const size_t t1 = (((((size_t)kj * nk + ka) * nb + b) * nop + jo) * nvp + aoff + a) * nop + io;
        const double x = Tr[t1], y = Ti[t1];
        const size_t t2 = (((((size_t)ki * nk + ka) * nb + b) * nop + io) * nvp + aoff + a) * nop + jo;
        const double x2 = Tr[t2], y2 = Ti[t2];
        const double aa = x * x + y * y, ab = x * x2 + y * y2;
        const double den = eps[(size_t)ki * nc + icore + io] + eps[(size_t)kj * nc + icore + jo] -
                           eps[(size_t)ka * nc + iocc + a] - eps[(size_t)kb * nc + iocc + b];
        es += (0.5 * aa + 0.5 * ab) / den;
        et += (1.5 * aa - 1.5 * ab) / den;
    }
    __shared__ double ss[256], st[256];
    ss[threadIdx.x] = es;
    st[threadIdx.x] = et;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            ss[threadIdx.x] += ss[threadIdx.x + w];
            st[threadIdx.x] += st[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicAdd(&acc[0], ss[0]);
        atomicAdd(&acc[1], st[0]);
    }
}

// Quasi-particle sums of one k_b (IOPTN(44), src/pmp2energy.f90:445-522) on
// the TOVOKK layout of k_mp2k_energy: block (g', kg) adds to dq[kg][g'] the
// band g = vlo + g' at kg (g < min(iocc + nqp, IALL(kg) - IVIRTCORE))
//   sum over (kj, b, j, a), ka = kg + kj - kb:
//     [2 |T(g,a,j;ka,kj)|^2 - Re T(g,a,j;ka,kj) T*(j,a,g;ka,kg)] / (e_g + e_j - e_a - e_b)
//   sum over (kj, b, j, i), ki = kg - kj + kb:
//     [2 |T(i,g,j;kg,kj)|^2 - Re T(i,g,j;kg,kj) T*(j,g,i;kg,ki)] / (e_g + e_b - e_i - e_j)
// One block per output and the k_b launches in stream order: no atomics,
// the same sum on every run.
__global__ void __launch_bounds__(256) k_mp2k_qp(const double* __restrict__ Tr, const double* __restrict__ Ti,
                                                 const double* __restrict__ eps, const int* __restrict__ plim,
                                                 int nk, int kb, int nb, int no, int nv, int nop, int nvp, int aoff,
                                                 int nc, int icore, int iocc, int vlo, int ghi, int ng,
                                                 double* __restrict__ dq) {
    const int gp = blockIdx.x, kg = blockIdx.y;
    const int g = vlo + gp;
    if (g >= min(ghi, plim[kg])) return;                      // whole block
    const double eg = eps[(size_t)kg * nc + g];
    const size_t sK = (size_t)nb * nop * nvp * nop;          // one (kj, ka) block of TOVOKK
    double q = 0.0;
    const size_t n1 = (size_t)nk * nb * no * nv;
    for (size_t t = threadIdx.x; t < n1; t += blockDim.x) {
        const int a = (int)(t % nv);
        size_t r = t / nv;
        const int jo = (int)(r % no); r /= no;
        const int b = (int)(r % nb);
        const int kj = (int)(r / nb);
        const int ka = (kg + kj - kb + nk) % nk;
        if (iocc + a >= plim[ka]) continue;
        const size_t o1 = ((size_t)kj * nk + ka) * sK + (((size_t)b * nop + jo) * nvp + aoff + a) * nop + (g - icore);
        const size_t o2 = ((size_t)kg * nk + ka) * sK + (((size_t)b * nop + (g - icore)) * nvp + aoff + a) * nop + jo;
        const double x = Tr[o1], y = Ti[o1], x2 = Tr[o2], y2 = Ti[o2];
        const double den = eg + eps[(size_t)kj * nc + icore + jo] - eps[(size_t)ka * nc + iocc + a] -
                           eps[(size_t)kb * nc + iocc + b];
        q += (2.0 * (x * x + y * y) - (x * x2 + y * y2)) / den;
    }
    const size_t n2 = (size_t)nk * nb * no * no;
    for (size_t t = threadIdx.x; t < n2; t += blockDim.x) {
        const int io = (int)(t % no);
        size_t r = t / no;
        const int jo = (int)(r % no); r /= no;
        const int b = (int)(r % nb);
        const int kj = (int)(r / nb);
        const int ki = (kg - kj + kb + nk) % nk;
        const size_t o1 = ((size_t)kj * nk + kg) * sK + (((size_t)b * nop + jo) * nvp + (g - vlo)) * nop + io;
        const size_t o2 = ((size_t)ki * nk + kg) * sK + (((size_t)b * nop + io) * nvp + (g - vlo)) * nop + jo;
        const double x = Tr[o1], y = Ti[o1], x2 = Tr[o2], y2 = Ti[o2];
        const double den = eg + eps[(size_t)kb * nc + iocc + b] - eps[(size_t)ki * nc + icore + io] -
                           eps[(size_t)kj * nc + icore + jo];
        q += (2.0 * (x * x + y * y) - (x * x2 + y * y2)) / den;
    }
    __shared__ double sq[256];
    sq[threadIdx.x] = q;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sq[threadIdx.x] += sq[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) dq[(size_t)kg * ng + gp] += sq[0];
}

// nqp > 0: also the quasi-particle bands (IOPTN(44)) into qeps[k][band]
// (ncgs per k): the j and i sides of the transforms run over icore ..
// min(iocc + nqp, max IALL - IVIRTCORE) - 1 and the a side from
// max(iocc - nqp, icore) (src/pmp2energy.f90:137-153), then k_mp2k_qp.
static int mp2_periodic(polyeri_ctx* ctx, int kvc, const double* co, const double* eps, const int* nall, int icore,
                        int iocc, int ivirtcore, int transform, double* e2s, double* e2t, double* seconds, int nqp,
                        double* qeps) {
    if (!ctx || !co || !eps || !nall || !e2s || !e2t)
        return fail(POLYERI_E_ARG, "polyeri_mp2_energy_periodic: null argument");
    if (nqp < 0 || (nqp > 0 && !qeps))
        return fail(POLYERI_E_ARG, "polyeri_mp2_energy_periodic: bad quasi-particle request");
    if (!ctx->pbc) return fail(POLYERI_E_ARG, "polyeri_mp2_energy_periodic: not a periodic context");
    if (kvc < 0 || kvc > 512) return fail(POLYERI_E_ARG, "polyeri_mp2_energy_periodic: bad KPOINTS");
    if (transform != POLYERI_GEMM_MFMA_F64 && transform != POLYERI_GEMM_MFMA_BF16X3)
        return fail(POLYERI_E_ARG, "polyeri_mp2_energy_periodic: transform must be MFMA_F64 or MFMA_BF16X3");
    const int nk = kvc > 0 ? 2 * kvc : 1, nc = ctx->rb.ncgs, n = ctx->ib.nbf;
    const int no = iocc - icore;
    if (icore < 0 || no <= 0 || ivirtcore < 0) return fail(POLYERI_E_ARG, "polyeri_mp2_energy_periodic: empty occupied space");
    int nvmax = 0;
    std::vector<int> plim(nk);
    for (int k = 0; k < nk; ++k) {
        if (nall[k] < 0 || nall[k] > nc) return fail(POLYERI_E_ARG, "polyeri_mp2_energy_periodic: bad IALL");
        plim[k] = nall[k] - ivirtcore;
        if (plim[k] < iocc)
            return fail(POLYERI_E_ARG, "polyeri_mp2_energy_periodic: fewer bands than occupied orbitals at a k-point");
        nvmax = std::max(nvmax, plim[k] - iocc);
    }
    if (nvmax <= 0) return fail(POLYERI_E_ARG, "polyeri_mp2_energy_periodic: empty virtual space");
    const int pmax = iocc + nvmax;                      // max IALL - IVIRTCORE
    const int ghi = std::min(iocc + nqp, pmax), vlo = std::max(iocc - nqp, icore);
    const int nop = ghi - icore, nvp = pmax - vlo, aoff = iocc - vlo, ng = nqp > 0 ? ghi - vlo : 0;
    auto t0 = std::chrono::steady_clock::now();
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    const int mode = transform == POLYERI_GEMM_MFMA_F64 ? GEMM_F64 : GEMM_BF16X3;
    const long N = n, n1 = 2 * ctx->rc1 + 1, n3 = 2 * ctx->cel2 + 1, NK = nk, NO = nop, NV = nvp;
    const long NB = nvmax;                              // virtual bands of one k_b, at most
    const long M1 = N * n1 * N * n3 * N, M2 = N * n1 * N;
    if (M1 > INT_MAX || (long)NK * NB * NO > INT_MAX || (long)NB * NO * NV > INT_MAX)
        return fail(POLYERI_E_ARG, "polyeri_mp2_energy_periodic: lattice tensor too large for one pass");
    const size_t nT = (size_t)M1 * n1 * N;
    const size_t nTN = (size_t)M1 * NB, nTOK = (size_t)NK * NB * NO * M2, nTV = (size_t)NK * NK * NB * NO * NV * N,
                 nTOV = (size_t)NK * NK * NB * NO * NV * NO;
    const size_t nB1 = (size_t)n1 * N * NB, nB2 = (size_t)NK * n3 * N * NO, nB3 = (size_t)NK * n1 * N * NV,
                 nB4 = (size_t)2 * NK * N * NO, nC = (size_t)NK * N * nc;
    const size_t need = 8 * (nT + 2 * (nTN + nTOK + nTV + nTOV + nB1 + nB2 + nB3 + nB4 + nC) + (size_t)nc * nk) +
                        sizeof(double2) * (size_t)nc * nc * nk;
    size_t freeb = 0, totalb = 0;
    HIPCHK(hipMemGetInfo(&freeb, &totalb));
    if (need > freeb - freeb / 16)
        return fail(POLYERI_E_NOMEM, "polyeri_mp2_energy_periodic: " + std::to_string(need >> 20) +
                                         " MiB of transform buffers exceed free device memory");
    DevBuf buf;
    double *dT, *dTNr, *dTNi, *dTOKr, *dTOKi, *dTVr, *dTVi, *dTOVr, *dTOVi, *dB1r, *dB1i, *dB2r, *dB2i, *dB3r, *dB3i,
        *dB4r, *dB4i, *dCr, *dCi, *dEps, *dAcc, *dQ = nullptr;
    double2* dCO;
    int* dPlim;
    int rc;
    if ((rc = buf.get(&dT, nT)) || (rc = buf.get(&dTNr, nTN)) || (rc = buf.get(&dTNi, nTN)) ||
        (rc = buf.get(&dTOKr, nTOK)) || (rc = buf.get(&dTOKi, nTOK)) || (rc = buf.get(&dTVr, nTV)) ||
        (rc = buf.get(&dTVi, nTV)) || (rc = buf.get(&dTOVr, nTOV)) || (rc = buf.get(&dTOVi, nTOV)) ||
        (rc = buf.get(&dB1r, nB1)) || (rc = buf.get(&dB1i, nB1)) || (rc = buf.get(&dB2r, nB2)) ||
        (rc = buf.get(&dB2i, nB2)) || (rc = buf.get(&dB3r, nB3)) || (rc = buf.get(&dB3i, nB3)) ||
        (rc = buf.get(&dB4r, nB4)) || (rc = buf.get(&dB4i, nB4)) || (rc = buf.get(&dCr, nC)) ||
        (rc = buf.get(&dCi, nC)) || (rc = buf.get(&dEps, (size_t)nc * nk)) || (rc = buf.get(&dAcc, 2)) ||
        (rc = buf.get(&dCO, (size_t)nc * nc * nk)) || (rc = buf.get(&dPlim, (size_t)nk)) ||
        (ng > 0 && (rc = buf.get(&dQ, (size_t)nk * ng))))
        return rc;
    if (ng > 0) hipLaunchKernelGGL(k_fill, dim3(nblk((size_t)nk * ng)), dim3(256), 0, s, dQ, (size_t)nk * ng, 0.0);
    HIPCHK(hipMemcpyAsync(dCO, co, sizeof(double2) * (size_t)nc * nc * nk, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dEps, eps, 8 * (size_t)nc * nk, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dPlim, plim.data(), sizeof(int) * nk, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_co_planar, dim3(nblk(nC)), dim3(256), 0, s, dCO, nc, ctx->d_ref, ctx->d_scale, n, nk, dCr,
                       dCi);
    HIPCHK(hipGetLastError());
    if ((rc = tensor_device_cells(ctx, dT, nT))) return rc;
    // k-independent operands: B3[ka][(Q1,j)][a] = CO(j,a,ka) e^{i ka Q1} (:323-329),
    // B4[t][i][io] = conj(CO(i,io,t mod nk)), t < 2 nk (:350-386)
    hipLaunchKernelGGL(k_mp2k_operand, dim3(nblk(nB3)), dim3(256), 0, s, dCr, dCi, nk, n, nc, kvc, 0, nk, (int)n1,
                       ctx->rc1, nvp, vlo, 1, 0, 0, 0, dPlim, dB3r, dB3i);
    hipLaunchKernelGGL(k_mp2k_operand, dim3(nblk(nB4)), dim3(256), 0, s, dCr, dCi, nk, n, nc, kvc, 0, 2 * nk, 1, 0,
                       nop, icore, 0, 0, 0, 1, dPlim, dB4r, dB4i);
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(256), 0, s, dAcc, (size_t)2, 0.0);
    HIPCHK(hipGetLastError());
    auto gemm = [&](int M, int Nn, int K, const double* Ar, const double* Ai, long sam, long sak, long sA,
                    const double* Br, const double* Bi, long sbk, long sbn, long sB, double* Cr, double* Ci, long scm,
                    long scn, long sC, int batch, int nci = 0, long sbo = 0, long sco = 0) {
        return polyeri_zgemm_launch(mode, M, Nn, K, Ar, Ai, sam, sak, sA, Br, Bi, sbk, sbn, sB, Cr, Ci, scm, scn, sC,
                                    batch, s, nci, sbo, sco);
    };
    for (int kb = 0; kb < nk; ++kb) {
        const int nb = plim[kb] - iocc;
        if (nb <= 0) continue;
        const long NBk = nb;
        // Q1 quarter: TN[b][(i,Q1,j,Q3,k)] = sum_{Q2,l} T CO(l,b,kb) e^{i kb Q2} (:253-258;
        // the e^{i kb Q3} factor moves into the next operand); b outermost so the
        // next quarter reads its contraction index (Q3,k) contiguously
        hipLaunchKernelGGL(k_mp2k_operand, dim3(nblk(nB1)), dim3(256), 0, s, dCr, dCi, nk, n, nc, kvc, kb, 1,
                           (int)n1, ctx->rc1, nb, iocc, 1, 0, 0, 0, dPlim, dB1r, dB1i);
        // B2[kj][(Q3,k)][jo] = conj(CO(k,jo,kj)) e^{i (kb - kj) Q3} (:294-299)
        hipLaunchKernelGGL(k_mp2k_operand, dim3(nblk(nB2)), dim3(256), 0, s, dCr, dCi, nk, n, nc, kvc, 0, nk,
                           (int)n3, ctx->cel2, nop, icore, -1, kb, 1, 1, dPlim, dB2r, dB2i);
        HIPCHK(hipGetLastError());
        if (gemm((int)M1, nb, (int)(n1 * N), dT, nullptr, n1 * N, 1, 0, dB1r, dB1i, NBk, 1, 0, dTNr, dTNi, 1, M1, 0,
                 1))
            return fail(POLYERI_E_HIP, "k-point MP2: first quarter launch failed");
        // Q2 quarter, every kj in one product (columns (kj, jo)), so TN is
        // read once: TOK[kj][b][jo][(i,Q1,j)] = sum_{Q3,k} TN B2[kj]
        if (gemm((int)M2, nk * nop, (int)(n3 * N), dTNr, dTNi, n3 * N, 1, M1, dB2r, dB2i, NO, 1, 0, dTOKr, dTOKi, 1,
                 M2, NO * M2, nb, nop, n3 * N * NO, NBk * NO * M2))
            return fail(POLYERI_E_HIP, "k-point MP2: second quarter launch failed");
        // Q3 quarter, every ka in one product (columns (ka, a)):
        // TV[ka][(kj,b,jo)][a][i] = sum_{Q1,j} TOK B3[ka]
        if (gemm((int)N, nk * nvp, (int)(n1 * N), dTOKr, dTOKi, n1 * N, 1, M2, dB3r, dB3i, NV, 1, 0, dTVr, dTVi, 1,
                 N, NV * N, nk * nb * nop, nvp, n1 * N * NV, NK * NBk * NO * NV * N))
            return fail(POLYERI_E_HIP, "k-point MP2: third quarter launch failed");
        // Q4 quarter, per kj, batched over ka with ki = ka + (kb - kj mod nk):
        // TOV[kj][ka][(b,jo,a)][io] = sum_i TV[ka][kj] conj(CO(i,io,ki))
        for (int kj = 0; kj < nk; ++kj) {
            const long d = ((kb - kj) % nk + nk) % nk;
            if (gemm(nb * nop * nvp, nop, (int)N, dTVr + kj * NBk * NO * NV * N, dTVi + kj * NBk * NO * NV * N, N, 1,
                     NK * NBk * NO * NV * N, dB4r + d * N * NO, dB4i + d * N * NO, NO, 1, N * NO,
                     dTOVr + kj * NK * NBk * NO * NV * NO, dTOVi + kj * NK * NBk * NO * NV * NO, NO, 1,
                     NBk * NO * NV * NO, nk))
                return fail(POLYERI_E_HIP, "k-point MP2: fourth quarter launch failed");
        }
        const size_t ne = (size_t)nk * nk * nb * no * nvmax * no;
        const size_t grid = std::min<size_t>(nblk(ne), 4096);
        hipLaunchKernelGGL(k_mp2k_energy, dim3((unsigned)grid), dim3(256), 0, s, dTOVr, dTOVi, dEps, dPlim, nk, kb, nb,
                           no, nvmax, nop, nvp, aoff, nc, icore, iocc, dAcc);
        if (ng > 0)
            hipLaunchKernelGGL(k_mp2k_qp, dim3(ng, nk), dim3(256), 0, s, dTOVr, dTOVi, dEps, dPlim, nk, kb, nb, no,
                               nvmax, nop, nvp, aoff, nc, icore, iocc, vlo, ghi, ng, dQ);
        HIPCHK(hipGetLastError());
    }
    double acc[2];
    std::vector<double> hq((size_t)nk * ng);
    HIPCHK(hipMemcpyAsync(acc, dAcc, 16, hipMemcpyDeviceToHost, s));
    if (ng > 0) HIPCHK(hipMemcpyAsync(hq.data(), dQ, 8 * hq.size(), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    const double norm = (double)nk * nk * nk;
    *e2s = acc[0] / norm;
    *e2t = acc[1] / norm;
    if (nqp > 0) {
        // QEPSILON = EPSILON + sum / (2 KVC)^2 (src/pmp2energy.f90:550-558)
        std::copy(eps, eps + (size_t)nc * nk, qeps);
        for (int k = 0; k < nk; ++k)
            for (int g = vlo; g < std::min(ghi, plim[k]); ++g)
                qeps[(size_t)k * nc + g] = eps[(size_t)k * nc + g] + hq[(size_t)k * ng + g - vlo] / ((double)nk * nk);
    }
    if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (!std::isfinite(acc[0]) || !std::isfinite(acc[1])) return fail(POLYERI_E_NONFINITE, "non-finite E(2)");
    return 0;
}

int polyeri_mp2_energy_periodic(polyeri_ctx* ctx, int kvc, const double* co, const double* eps, const int* nall,
                                int icore, int iocc, int ivirtcore, int transform, double* e2s, double* e2t,
                                double* seconds) {
    return mp2_periodic(ctx, kvc, co, eps, nall, icore, iocc, ivirtcore, transform, e2s, e2t, seconds, 0, nullptr);
}

int polyeri_reduced_cel1(const polyeri_ctx* ctx) {
    if (!ctx) return fail(POLYERI_E_ARG, "polyeri_reduced_cel1: null context");
    if (!ctx->pbc) return fail(POLYERI_E_ARG, "polyeri_reduced_cel1: not a periodic context");
    return ctx->rc1;
}

int polyeri_eri_tensor_cells(polyeri_ctx* ctx, double* t_c) {
    if (!ctx || !t_c) return fail(POLYERI_E_ARG, "polyeri_eri_tensor_cells: null argument");
    if (!ctx->pbc) return fail(POLYERI_E_ARG, "polyeri_eri_tensor_cells: not a periodic context");
    HIPCHK(hipSetDevice(ctx->device));
    const size_t n = ctx->ib.nbf, nc = ctx->rb.ncgs, n1 = 2 * ctx->rc1 + 1, n3 = 2 * ctx->cel2 + 1;
    const size_t nt = n * n1 * n * n3 * n * n1 * n, nc4 = nc * nc * nc * nc;
    double* d_t = nullptr;
    int rc;
    if ((rc = dalloc(&d_t, nt))) return rc;
    rc = tensor_device_cells(ctx, d_t, nt);
    std::vector<double> t;
    if (!rc) {
        t.resize(nt);
        hipError_t e = hipMemcpyAsync(t.data(), d_t, nt * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) rc = fail(POLYERI_E_HIP, hipGetErrorString(e));
    }
    (void)hipFree(d_t);
    if (rc) return rc;
    std::fill(t_c, t_c + nc4 * n1 * n1 * n3, 0.0);
    const auto& ref = ctx->ib.ref_of_int;
    const auto& sc = ctx->ib.scale;
    // device T[i][Q1][j][Q3][k][Q2][l] -> block (Q1, Q2, Q3) of column-major FERI(I,J,K,L)
    size_t x = 0;
    for (size_t a = 0; a < n; ++a)
        for (size_t q1 = 0; q1 < n1; ++q1)
            for (size_t b = 0; b < n; ++b)
                for (size_t q3 = 0; q3 < n3; ++q3)
                    for (size_t c = 0; c < n; ++c)
                        for (size_t q2 = 0; q2 < n1; ++q2)
                            for (size_t d = 0; d < n; ++d, ++x) {
                                double* blk = t_c + ((q1 * n1 + q2) * n3 + q3) * nc4;
                                blk[ref[a] + nc * (ref[b] + nc * (ref[c] + nc * ref[d]))] =
                                    sc[a] * sc[b] * sc[c] * sc[d] * t[x];
                            }
    return 0;
}

// ---------------------------------------------------------------------------
// The reference's .ao ERI files (STORE_FOURINDEX_ERI / RESTORE_FOURINDEX_ERI)
// ---------------------------------------------------------------------------
// Fortran unformatted sequential records as gfortran / amdflang write them:
// a 4-byte little-endian length before and after each payload.  A payload is
// 1000 entries: ICASHE1X, ICASHE1Y, ICASHE1Z (int32: cells packed
// ((Q1+CEL1)*256 + Q2+CEL1)*256 + Q3+CEL2), ICASHE2 (int32: ((I*256+J)*256+K)
// *256+L, 1-based) and DCASHE (float64); unused entries of the last record
// are -1 (src/pfourindexstore.f90:731-747, 781).
namespace {
const int kAoCache = 1000;

struct AoWriter {
    FILE* f = nullptr;
    std::vector<int32_t> x, y, z, ijkl;
    std::vector<double> v;
    int n = 0;
    int reset() {
        std::fill(x.begin(), x.end(), -1); std::fill(y.begin(), y.end(), -1); std::fill(z.begin(), z.end(), -1);
        std::fill(ijkl.begin(), ijkl.end(), -1); std::fill(v.begin(), v.end(), 0.0);
        n = 0;
        return 0;
    }
    bool flush() {
        const int32_t len = kAoCache * (4 * 4 + 8);
        bool ok = fwrite(&len, 4, 1, f) == 1;
        ok = ok && fwrite(x.data(), 4, kAoCache, f) == (size_t)kAoCache;
        ok = ok && fwrite(y.data(), 4, kAoCache, f) == (size_t)kAoCache;
        ok = ok && fwrite(z.data(), 4, kAoCache, f) == (size_t)kAoCache;
        ok = ok && fwrite(ijkl.data(), 4, kAoCache, f) == (size_t)kAoCache;
        ok = ok && fwrite(v.data(), 8, kAoCache, f) == (size_t)kAoCache;
        ok = ok && fwrite(&len, 4, 1, f) == 1;
        reset();
        return ok;
    }
};
}  // namespace

int polyeri_store_fourindex_eri(polyeri_ctx* ctx, const char* path, double inttol, long long* stored) {
    if (!ctx || !path) return fail(POLYERI_E_ARG, "polyeri_store_fourindex_eri: null argument");
    const size_t nc = ctx->rb.ncgs;
    if (nc >= 256) return fail(POLYERI_E_ARG, "polyeri_store_fourindex_eri: NCGS must be below 256");
    if (ctx->pbc && (2 * ctx->cel1 + 1 > 255 || 2 * ctx->cel2 + 1 > 255))
        return fail(POLYERI_E_ARG, "polyeri_store_fourindex_eri: cell offsets must pack into 8 bits");
    // molecule: one cell triple (0, 0, 0); periodic: Q1, Q2 in [-rc1, rc1], Q3 in [-cel2, cel2]
    const int rc1 = ctx->pbc ? ctx->rc1 : 0, cel1 = ctx->pbc ? ctx->cel1 : 0, cel2 = ctx->pbc ? ctx->cel2 : 0;
    const size_t n1 = 2 * rc1 + 1, n3 = 2 * cel2 + 1, nc4 = nc * nc * nc * nc;
    std::vector<double> feri(nc4 * n1 * n1 * n3);
    int rc = ctx->pbc ? polyeri_eri_tensor_cells(ctx, feri.data()) : polyeri_eri_tensor(ctx, feri.data());
    if (rc) return rc;
    AoWriter w;
    w.f = fopen(path, "wb");
    if (!w.f) return fail(POLYERI_E_ARG, std::string("cannot open ") + path);
    w.x.resize(kAoCache); w.y.resize(kAoCache); w.z.resize(kAoCache); w.ijkl.resize(kAoCache); w.v.resize(kAoCache);
    w.reset();
    long long count = 0;
    bool ok = true;
    // Q1, Q2, Q3 outer, then I outermost and L innermost (src/pfourindexstore.f90:258-262, 378, 727-750);
    // ICASHE1X = ((Q1+CEL1)*256 + Q2+CEL1)*256 + Q3+CEL2; the y and z cells are 0
    for (int q1 = -rc1; q1 <= rc1 && ok; ++q1)
        for (int q2 = -rc1; q2 <= rc1 && ok; ++q2)
            for (int q3 = -cel2; q3 <= cel2 && ok; ++q3) {
                const double* blk = feri.data() + (((size_t)(q1 + rc1) * n1 + (q2 + rc1)) * n3 + (q3 + cel2)) * nc4;
                const int32_t cx = ((q1 + cel1) * 256 + (q2 + cel1)) * 256 + (q3 + cel2);
                for (size_t i = 0; i < nc && ok; ++i)
                    for (size_t j = 0; j < nc && ok; ++j)
                        for (size_t k = 0; k < nc && ok; ++k)
                            for (size_t l = 0; l < nc && ok; ++l) {
                                const double v = blk[i + nc * (j + nc * (k + nc * l))];
                                if (!(std::fabs(v) > inttol)) continue;
                                w.x[w.n] = cx; w.y[w.n] = 0; w.z[w.n] = 0;
                                w.ijkl[w.n] = (int32_t)((((i + 1) * 256 + (j + 1)) * 256 + (k + 1)) * 256 + (l + 1));
                                w.v[w.n] = v;
                                ++count;
                                if (++w.n == kAoCache) ok = w.flush();
                            }
            }
    if (ok && w.n) ok = w.flush();
    ok = (fclose(w.f) == 0) && ok;
    if (!ok) return fail(POLYERI_E_ARG, std::string("write error on ") + path);
    if (stored) *stored = count;
    return 0;
}

static int restore_ao(const char* path, int ncgs, int cel1, int cel2, double* feri, long long* restored);

int polyeri_restore_fourindex_eri(const char* path, int ncgs, double* feri, long long* restored) {
    return restore_ao(path, ncgs, -1, 0, feri, restored);
}

int polyeri_restore_fourindex_eri_cells(const char* path, int ncgs, int cel1, int cel2, double* t_c,
                                        long long* restored) {
    if (cel1 < 0 || cel2 < 0 || 2 * cel1 + 1 > 255 || 2 * cel2 + 1 > 255)
        return fail(POLYERI_E_ARG, "polyeri_restore_fourindex_eri_cells: bad cell range");
    return restore_ao(path, ncgs, cel1, cel2, t_c, restored);
}

// Records of one .ao file, appended to out.  cel1 < 0: a molecular file
// (every record in cell 0); else the file's cells are packed with cel1 /
// cel2 and must satisfy Q1, Q2 in [-q12, q12], Q3 in [-cel2, cel2] (the
// reference's degraded-file test, src/pfourindexstore.f90:901-906), and
// every index lies in 1 .. ncgs.
static int read_ao_records(const char* path, int ncgs, int cel1, int q12, int cel2, std::vector<StoredEri>& out) {
    const bool cells = cel1 >= 0;
    FILE* f = fopen(path, "rb");
    if (!f) return fail(POLYERI_E_ARG, std::string("cannot open ") + path);
    std::vector<int32_t> x(kAoCache), y(kAoCache), z(kAoCache), ijkl(kAoCache);
    std::vector<double> v(kAoCache);
    int rc = 0;
    for (;;) {
        int32_t len = 0, len2 = 0;
        if (fread(&len, 4, 1, f) != 1) break;
        if (len != kAoCache * 24) { rc = fail(POLYERI_E_ARG, "INTEGRAL FILE HAS BEEN DEGRADED"); break; }
        bool ok = fread(x.data(), 4, kAoCache, f) == (size_t)kAoCache && fread(y.data(), 4, kAoCache, f) == (size_t)kAoCache &&
                  fread(z.data(), 4, kAoCache, f) == (size_t)kAoCache &&
                  fread(ijkl.data(), 4, kAoCache, f) == (size_t)kAoCache && fread(v.data(), 8, kAoCache, f) == (size_t)kAoCache &&
                  fread(&len2, 4, 1, f) == 1 && len2 == len;
        if (!ok) { rc = fail(POLYERI_E_ARG, "INTEGRAL FILE HAS BEEN DEGRADED"); break; }
        for (int e = 0; e < kAoCache; ++e) {
            if (x[e] == -1) break;
            if (!cells) {
                if (x[e] != 0 || y[e] != 0 || z[e] != 0) { rc = fail(POLYERI_E_ARG, "periodic .ao records in a molecular read"); break; }
            } else {
                // src/pmp2energy.f90:221-252: unpack and range-check the cells
                const int q1 = ((x[e] >> 16) & 255) - cel1, q2 = ((x[e] >> 8) & 255) - cel1, q3 = (x[e] & 255) - cel2;
                if ((x[e] >> 24) != 0 || y[e] != 0 || z[e] != 0 || q1 < -q12 || q1 > q12 || q2 < -q12 || q2 > q12 ||
                    q3 < -cel2 || q3 > cel2) { rc = fail(POLYERI_E_ARG, "INTEGRAL FILE HAS BEEN DEGRADED"); break; }
            }
            const int i = ((ijkl[e] >> 24) & 255), j = ((ijkl[e] >> 16) & 255), k = ((ijkl[e] >> 8) & 255),
                      l = (ijkl[e] & 255);
            if (i < 1 || j < 1 || k < 1 || l < 1 || i > ncgs || j > ncgs || k > ncgs || l > ncgs) {
                rc = fail(POLYERI_E_ARG, "INTEGRAL FILE HAS BEEN DEGRADED");
                break;
            }
            out.push_back(StoredEri{x[e], ijkl[e], v[e]});
        }
        if (rc) break;
    }
    fclose(f);
    return rc;
}

// cel1 < 0: molecular file (every record in cell 0); else Q1, Q2 in
// [-cel1, cel1], Q3 in [-cel2, cel2] into (2 cel1 + 1)^2 (2 cel2 + 1) blocks.
static int restore_ao(const char* path, int ncgs, int cel1, int cel2, double* feri, long long* restored) {
    if (!path || !feri || ncgs <= 0 || ncgs >= 256) return fail(POLYERI_E_ARG, "polyeri_restore_fourindex_eri: bad argument");
    const bool cells = cel1 >= 0;
    const size_t n1 = cells ? 2 * cel1 + 1 : 1, n3 = cells ? 2 * cel2 + 1 : 1;
    const size_t nc = ncgs, nc4 = nc * nc * nc * nc;
    std::vector<StoredEri> rec;
    int rc = read_ao_records(path, ncgs, cel1, cel1, cel2, rec);
    if (rc) return rc;
    std::fill(feri, feri + nc4 * n1 * n1 * n3, 0.0);
    for (const StoredEri& r : rec) {
        size_t blk = 0;
        if (cells) {
            const int q1 = ((r.cells >> 16) & 255) - cel1, q2 = ((r.cells >> 8) & 255) - cel1, q3 = (r.cells & 255) - cel2;
            blk = (((size_t)(q1 + cel1) * n1 + (q2 + cel1)) * n3 + (q3 + cel2)) * nc4;
        }
        const size_t i = ((r.ijkl >> 24) & 255) - 1, j = ((r.ijkl >> 16) & 255) - 1, k = ((r.ijkl >> 8) & 255) - 1,
                     l = (r.ijkl & 255) - 1;
        feri[blk + i + nc * (j + nc * (k + nc * l))] = r.v;
    }
    if (restored) *restored = (long long)rec.size();
    return 0;
}

// DIRECT .F.: the records of an .ao file into the context (appended to
// those already loaded), and the context's J/K builds switched to them.
int polyeri_load_fourindex_eri(polyeri_ctx* ctx, const char* path, long long* restored) {
    if (!ctx || !path) return fail(POLYERI_E_ARG, "polyeri_load_fourindex_eri: null argument");
    const int nc = ctx->rb.ncgs;
    if (nc >= 256) return fail(POLYERI_E_ARG, "polyeri_load_fourindex_eri: the .ao format holds at most 255 functions");
    if (ctx->pbc && (2 * ctx->cel1 + 1 > 255 || 2 * ctx->cel2 + 1 > 255))
        return fail(POLYERI_E_ARG, "polyeri_load_fourindex_eri: cell range beyond the .ao format");
    std::vector<StoredEri> rec;
    int rc = read_ao_records(path, nc, ctx->pbc ? ctx->cel1 : -1, ctx->pbc ? ctx->rc1 : 0, ctx->pbc ? ctx->cel2 : 0,
                             rec);
    if (rc) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    const size_t tot = (size_t)ctx->n_ao + rec.size();
    StoredEri* d = nullptr;
    if ((rc = dalloc(&d, tot))) return rc;
    if (ctx->n_ao) HIPCHK(hipMemcpy(d, ctx->d_ao, ctx->n_ao * sizeof(StoredEri), hipMemcpyDeviceToDevice));
    if (!rec.empty())
        HIPCHK(hipMemcpy(d + ctx->n_ao, rec.data(), rec.size() * sizeof(StoredEri), hipMemcpyHostToDevice));
    (void)hipFree(ctx->d_ao);
    ctx->d_ao = d;
    ctx->n_ao = (long long)tot;
    if (!ctx->d_inv) {
        std::vector<int> inv(nc);
        for (int c = 0; c < nc; ++c) inv[c] = ctx->ib.int_of_ref[c];
        if ((rc = dalloc(&ctx->d_inv, nc))) return rc;
        HIPCHK(hipMemcpy(ctx->d_inv, inv.data(), nc * sizeof(int), hipMemcpyHostToDevice));
    }
    ctx->use_stored = true;
    if (restored) *restored = (long long)rec.size();
    return 0;
}

int polyeri_set_direct(polyeri_ctx* ctx, int direct) {
    if (!ctx) return fail(POLYERI_E_ARG, "null context");
    if (!direct && ctx->n_ao <= 0) return fail(POLYERI_E_ARG, "polyeri_set_direct: no stored integrals loaded");
    ctx->use_stored = !direct;
    return 0;
}

// ---------------------------------------------------------------------------
// Two-electron part of the nuclear gradient (FOURINDEX_ERI_GRADIENT,
// src/pfourindexgradient.f90:1-1478, molecule): kgrad.hip
// ---------------------------------------------------------------------------
struct GradHost {
    std::vector<int> L, atom, f0, p0, np, pa, pb;
    std::vector<double> xyz, alpha, coef, D, dsh;
    int seg[NPAIRCLASS + 1];
};

// Shells, primitives and the density in the library's basis:
// D(mu, nu) = s_mu s_nu P_C(ref mu, ref nu).  The reference's arrays carry
// no atom index per primitive, so each shell's atom is the one at its
// centre; -1 when no atom is there.
static int grad_host_arrays(const IntBasis& ib, const RefBasis& rb, const double* p_c, GradHost& h) {
    for (const Shell& s : ib.shells) {
        int at = -1;
        for (int a = 0; a < rb.natom && at < 0; ++a)
            if (std::fabs(rb.ax[a] - s.x) < 1e-8 && std::fabs(rb.ay[a] - s.y) < 1e-8 && std::fabs(rb.az[a] - s.z) < 1e-8)
                at = a;
        if (at < 0) return fail(POLYERI_E_BASIS, "eri gradient: a shell centre is not at an atom");
        h.L.push_back(s.L);
        h.atom.push_back(at);
        h.f0.push_back(s.f0);
        h.p0.push_back((int)h.alpha.size());
        h.np.push_back((int)s.alpha.size());
        h.xyz.insert(h.xyz.end(), {s.x, s.y, s.z});
        h.alpha.insert(h.alpha.end(), s.alpha.begin(), s.alpha.end());
        h.coef.insert(h.coef.end(), s.coef.begin(), s.coef.end());
    }
    const int nbf = ib.nbf, nc = rb.ncgs;
    h.D.resize((size_t)nbf * nbf);
    for (int mu = 0; mu < nbf; ++mu)
        for (int nu = 0; nu < nbf; ++nu)
            h.D[(size_t)mu * nbf + nu] =
                ib.scale[mu] * ib.scale[nu] * p_c[ib.ref_of_int[mu] + (size_t)nc * ib.ref_of_int[nu]];
    const int nsh = (int)ib.shells.size();
    h.dsh.assign((size_t)nsh * nsh, 0.0);
    for (int a = 0; a < nsh; ++a)
        for (int b = 0; b < nsh; ++b) {
            double m = 0.0;
            for (int p = 0; p < ncart(ib.shells[a].L); ++p)
                for (int q = 0; q < ncart(ib.shells[b].L); ++q)
                    m = std::max(m, std::fabs(h.D[(size_t)(ib.shells[a].f0 + p) * nbf + ib.shells[b].f0 + q]));
            h.dsh[(size_t)a * nsh + b] = m;
        }
    // every unordered shell pair once, higher L first, grouped by pair class
    // (ascending) and by primitive count inside a class
    std::vector<std::pair<int, int>> pr;
    for (int a = 0; a < nsh; ++a)
        for (int b = 0; b <= a; ++b)
            pr.push_back(h.L[b] >= h.L[a] ? std::make_pair(b, a) : std::make_pair(a, b));
    auto cls = [&](const std::pair<int, int>& p) { return pair_class(h.L[p.first], h.L[p.second]); };
    std::stable_sort(pr.begin(), pr.end(), [&](const auto& x, const auto& y) {
        if (cls(x) != cls(y)) return cls(x) < cls(y);
        return h.np[x.first] * h.np[x.second] > h.np[y.first] * h.np[y.second];
    });
    for (int X = 0; X <= NPAIRCLASS; ++X) h.seg[X] = 0;
    for (const auto& p : pr) {
        h.pa.push_back(p.first);
        h.pb.push_back(p.second);
        ++h.seg[cls(p) + 1];
    }
    for (int X = 0; X < NPAIRCLASS; ++X) h.seg[X + 1] += h.seg[X];
    return 0;
}

static int grad_check(polyeri_ctx* ctx, const double* p_c, double* grad, const char* who) {
    if (!ctx || !p_c || !grad) return fail(POLYERI_E_ARG, std::string(who) + ": null argument");
    if (ctx->pbc) return fail(POLYERI_E_ARG, std::string(who) + ": molecular contexts only");
    if (3 * ctx->rb.natom * 8 > 48 * 1024) return fail(POLYERI_E_ARG, std::string(who) + ": more than 2048 atoms");
    return 0;
}

int polyeri_eri_gradient(polyeri_ctx* ctx, const double* p_c, double hfexchange, double* grad) {
    int rc;
    if ((rc = grad_check(ctx, p_c, grad, "polyeri_eri_gradient"))) return rc;
    HIPCHK(hipSetDevice(ctx->device));
    GradHost h;
    if ((rc = grad_host_arrays(ctx->ib, ctx->rb, p_c, h))) return rc;
    const int nsh = (int)h.L.size(), natom = ctx->rb.natom, nbf = ctx->ib.nbf;
    const size_t npair = h.pa.size();
    DevBuf buf;
    int *dL, *dAt, *dF0, *dP0, *dNp, *dPa, *dPb;
    double *dXyz, *dAl, *dCo, *dD, *dDsh, *dQp, *dG;
    if ((rc = buf.get(&dL, nsh)) || (rc = buf.get(&dAt, nsh)) || (rc = buf.get(&dF0, nsh)) ||
        (rc = buf.get(&dP0, nsh)) || (rc = buf.get(&dNp, nsh)) || (rc = buf.get(&dXyz, 3 * (size_t)nsh)) ||
        (rc = buf.get(&dAl, h.alpha.size())) || (rc = buf.get(&dCo, h.coef.size())) ||
        (rc = buf.get(&dD, h.D.size())) || (rc = buf.get(&dDsh, h.dsh.size())) || (rc = buf.get(&dQp, npair)) ||
        (rc = buf.get(&dG, 3 * (size_t)natom)) || (rc = buf.get(&dPa, npair)) || (rc = buf.get(&dPb, npair)))
        return rc;
    hipStream_t s = ctx->stream;
    HIPCHK(hipMemcpyAsync(dL, h.L.data(), nsh * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dAt, h.atom.data(), nsh * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dF0, h.f0.data(), nsh * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dP0, h.p0.data(), nsh * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dNp, h.np.data(), nsh * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dXyz, h.xyz.data(), h.xyz.size() * 8, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dAl, h.alpha.data(), h.alpha.size() * 8, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dCo, h.coef.data(), h.coef.size() * 8, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dD, h.D.data(), h.D.size() * 8, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dDsh, h.dsh.data(), h.dsh.size() * 8, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_fill, dim3(nblk(3 * (size_t)natom)), dim3(256), 0, s, dG, 3 * (size_t)natom, 0.0);
    // pair bounds: the context's Schwarz bounds Q_ab = sqrt(max |(ab|ab)|) of
    // the same shell pairs (pairs the context dropped by primitive screening: 0)
    std::vector<double> qp(npair, 0.0);
    {
        std::unordered_map<long long, double> qmap;
        qmap.reserve(ctx->pairs.size() * 2);
        for (size_t i = 0; i < ctx->pairs.size(); ++i) {
            const int a = ctx->pairs[i].a, b = ctx->pairs[i].b;
            qmap[(long long)std::min(a, b) * nsh + std::max(a, b)] = ctx->Q[i];
        }
        for (size_t t = 0; t < npair; ++t) {
            auto it = qmap.find((long long)std::min(h.pa[t], h.pb[t]) * nsh + std::max(h.pa[t], h.pb[t]));
            if (it != qmap.end()) qp[t] = it->second;
        }
    }
    // only the quartets with Q_ab Q_cd >= tol are launched (pairs reordered
    // by bound inside each class and primitive count)
    const bool sorted = polyeri_grad_sort_pairs(h.pa.data(), h.pb.data(), qp.data(), h.np.data(), h.seg);
    GradChunks ch;
    unsigned long long* dOff = nullptr;
    int* dIj = nullptr;
    if (sorted) {
        polyeri_grad_chunks(h.pa.data(), h.pb.data(), h.np.data(), qp.data(), h.seg, ctx->opts.schwarz_tol, ch);
        if ((rc = buf.get(&dOff, ch.off.size())) || (rc = buf.get(&dIj, std::max<size_t>(ch.ij.size(), 1)))) return rc;
        HIPCHK(hipMemcpyAsync(dOff, ch.off.data(), ch.off.size() * 8, hipMemcpyHostToDevice, s));
        if (!ch.ij.empty())
            HIPCHK(hipMemcpyAsync(dIj, ch.ij.data(), ch.ij.size() * sizeof(int), hipMemcpyHostToDevice, s));
    }
    HIPCHK(hipMemcpyAsync(dPa, h.pa.data(), npair * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dPb, h.pb.data(), npair * sizeof(int), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dQp, qp.data(), npair * sizeof(double), hipMemcpyHostToDevice, s));
    GradArgs g{nsh, natom, nbf, (int)npair, {0}, dL, dAt, dF0, dP0, dNp, dPa, dPb, dXyz, dAl, dCo, dD, dDsh,
               nullptr, ctx->opts.schwarz_tol, hfexchange, ctx->opts.prim_tol};
    std::copy(h.seg, h.seg + NPAIRCLASS + 1, g.seg);
    if (polyeri_grad_launch(g, dQp, dG, s, true, sorted ? &ch : nullptr, dOff, dIj))
        return fail(POLYERI_E_HIP, "polyeri_eri_gradient: launch failed");
    HIPCHK(hipMemcpyAsync(grad, dG, 3 * (size_t)natom * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (int i = 0; i < 3 * natom; ++i)
        if (!std::isfinite(grad[i])) return fail(POLYERI_E_NONFINITE, "polyeri_eri_gradient: non-finite gradient");
    return 0;
}

int polyeri_eri_gradient_host(int npgs, int ncgs, const int* pax, const int* pay, const int* paz,
                              const double* pgsx, const double* pgsy, const double* pgsz, const double* zt,
                              const double* cc, int natom, const int* iatom, const double* xyz, double schwarz_tol,
                              const double* p_c, double hfexchange, double* grad) {
    if (npgs <= 0 || ncgs <= 0 || !pax || !pay || !paz || !pgsx || !pgsy || !pgsz || !zt || !cc || natom <= 0 ||
        !iatom || !xyz || !p_c || !grad)
        return fail(POLYERI_E_ARG, "polyeri_eri_gradient_host: null or empty argument");
    RefBasis rb;
    rb.natom = natom;
    rb.iatom.assign(iatom, iatom + natom);
    for (int i = 0; i < natom; ++i) {
        rb.ax.push_back(xyz[3 * i]); rb.ay.push_back(xyz[3 * i + 1]); rb.az.push_back(xyz[3 * i + 2]);
    }
    rb.npgs = npgs; rb.ncgs = ncgs;
    rb.pax.assign(pax, pax + npgs); rb.pay.assign(pay, pay + npgs); rb.paz.assign(paz, paz + npgs);
    rb.pgsx.assign(pgsx, pgsx + npgs); rb.pgsy.assign(pgsy, pgsy + npgs); rb.pgsz.assign(pgsz, pgsz + npgs);
    rb.zt.assign(zt, zt + npgs);
    rb.cc.assign(cc, cc + (size_t)npgs * ncgs);
    rb.pgs_atom.assign(npgs, 0);
    IntBasis ib;
    std::string err;
    if (extract_shells(rb, ib, err)) return fail(POLYERI_E_BASIS, err);
    GradHost h;
    if (int rc = grad_host_arrays(ib, rb, p_c, h)) return rc;
    GradArgs g{(int)h.L.size(), natom, ib.nbf, (int)h.pa.size(), {0}, h.L.data(), h.atom.data(), h.f0.data(),
               h.p0.data(), h.np.data(), h.pa.data(), h.pb.data(), h.xyz.data(), h.alpha.data(), h.coef.data(),
               h.D.data(), h.dsh.data(), nullptr, schwarz_tol, hfexchange, 1e-20};
    std::copy(h.seg, h.seg + NPAIRCLASS + 1, g.seg);
    return polyeri_grad_host(g, grad) ? fail(POLYERI_E_ARG, "polyeri_eri_gradient_host failed") : 0;
}

// One-electron and nuclear parts of the nuclear gradient (OVERLAP_KINETIC(1),
// NUCLEAR_ATTRACTION(1), HELLMANN_FEYNMAN_ATTRACTION / _REPULSION): host code
// over the library's basis (host.cpp one_electron_gradient).
static int one_electron_gradient_ib(const IntBasis& ib, const RefBasis& rb, const double* p_c, const double* w_c,
                                    double* grad) {
    std::vector<int> at(ib.shells.size(), -1);
    for (size_t s = 0; s < ib.shells.size(); ++s) {
        const Shell& sh = ib.shells[s];
        for (int a = 0; a < rb.natom && at[s] < 0; ++a)
            if (std::fabs(rb.ax[a] - sh.x) < 1e-8 && std::fabs(rb.ay[a] - sh.y) < 1e-8 && std::fabs(rb.az[a] - sh.z) < 1e-8)
                at[s] = a;
        if (at[s] < 0) return fail(POLYERI_E_BASIS, "one-electron gradient: a shell centre is not at an atom");
    }
    const int nbf = ib.nbf, nc = rb.ncgs;
    std::vector<double> D((size_t)nbf * nbf), W((size_t)nbf * nbf);
    for (int mu = 0; mu < nbf; ++mu)
        for (int nu = 0; nu < nbf; ++nu) {
            const double sc = ib.scale[mu] * ib.scale[nu];
            const size_t r = ib.ref_of_int[mu] + (size_t)nc * ib.ref_of_int[nu];
            D[(size_t)mu * nbf + nu] = sc * p_c[r];
            W[(size_t)mu * nbf + nu] = sc * w_c[r];
        }
    std::vector<double> boys;
    boys_table(boys);
    one_electron_gradient(ib, at, boys, D.data(), W.data(), grad);
    for (int i = 0; i < 3 * rb.natom; ++i)
        if (!std::isfinite(grad[i])) return fail(POLYERI_E_NONFINITE, "one-electron gradient: non-finite value");
    return 0;
}

int polyeri_one_electron_gradient(polyeri_ctx* ctx, const double* p_c, const double* w_c, double* grad) {
    if (!ctx || !p_c || !w_c || !grad) return fail(POLYERI_E_ARG, "polyeri_one_electron_gradient: null argument");
    if (ctx->pbc) return fail(POLYERI_E_ARG, "polyeri_one_electron_gradient: molecular contexts only");
    return one_electron_gradient_ib(ctx->ib, ctx->rb, p_c, w_c, grad);
}

int polyeri_one_electron_gradient_host(int npgs, int ncgs, const int* pax, const int* pay, const int* paz,
                                       const double* pgsx, const double* pgsy, const double* pgsz, const double* zt,
                                       const double* cc, int natom, const int* iatom, const double* xyz,
                                       const double* p_c, const double* w_c, double* grad) {
    if (npgs <= 0 || ncgs <= 0 || !pax || !pay || !paz || !pgsx || !pgsy || !pgsz || !zt || !cc || natom <= 0 ||
        !iatom || !xyz || !p_c || !w_c || !grad)
        return fail(POLYERI_E_ARG, "polyeri_one_electron_gradient_host: null or empty argument");
    RefBasis rb;
    rb.natom = natom;
    rb.iatom.assign(iatom, iatom + natom);
    for (int i = 0; i < natom; ++i) {
        rb.ax.push_back(xyz[3 * i]); rb.ay.push_back(xyz[3 * i + 1]); rb.az.push_back(xyz[3 * i + 2]);
    }
    rb.npgs = npgs; rb.ncgs = ncgs;
    rb.pax.assign(pax, pax + npgs); rb.pay.assign(pay, pay + npgs); rb.paz.assign(paz, paz + npgs);
    rb.pgsx.assign(pgsx, pgsx + npgs); rb.pgsy.assign(pgsy, pgsy + npgs); rb.pgsz.assign(pgsz, pgsz + npgs);
    rb.zt.assign(zt, zt + npgs);
    rb.cc.assign(cc, cc + (size_t)npgs * ncgs);
    rb.pgs_atom.assign(npgs, 0);
    IntBasis ib;
    std::string err;
    if (extract_shells(rb, ib, err)) return fail(POLYERI_E_BASIS, err);
    return one_electron_gradient_ib(ib, rb, p_c, w_c, grad);
}

// MP2_ENERGY_ALG1 with MODULO = nmod > 1 (IOPTN(99), src/pmp2energy.f90:64-80):
// the k-point MP2 on the sub-grid of every nmod-th k-point of the SCF grid,
// k = kb nmod for kb in [-kvc / nmod, kvc / nmod - 1], with that sub-grid's
// phases and its (2 kvc / nmod)^3 normalisation -- the k-point MP2 of the
// selected bands.  Quasi-particle bands (nqp > 0) come back at those
// k-points of qeps (EPSILON's shape); the other k-points copy eps.
static int mp2_periodic_mod(polyeri_ctx* ctx, int kvc, int nmod, const double* co, const double* eps,
                            const int* nall, int icore, int iocc, int ivirtcore, int transform, double* e2s,
                            double* e2t, double* seconds, int nqp, double* qeps) {
    if (!ctx || !co || !eps || !nall || !e2s || !e2t)
        return fail(POLYERI_E_ARG, "polyeri_mp2_energy_periodic_mod: null argument");
    if (nmod < 1) return fail(POLYERI_E_ARG, "ILLEGAL MODULO PARAMETER");
    if (kvc < 0 || kvc / nmod * nmod != kvc) return fail(POLYERI_E_ARG, "MODULO MUST EVENLY DIVIDE KPOINTS ALONG X");
    if (nmod == 1)
        return mp2_periodic(ctx, kvc, co, eps, nall, icore, iocc, ivirtcore, transform, e2s, e2t, seconds, nqp,
                            qeps);
    if (nqp < 0 || (nqp > 0 && !qeps))
        return fail(POLYERI_E_ARG, "polyeri_mp2_energy_periodic: bad quasi-particle request");
    const int kvc2 = kvc / nmod, nk2 = kvc2 > 0 ? 2 * kvc2 : 1, nc = ctx->rb.ncgs;
    const size_t cstride = 2 * (size_t)nc * nc;           // complex, column-major per k
    std::vector<double> co2(cstride * nk2), eps2((size_t)nc * nk2), q2(nqp > 0 ? (size_t)nc * nk2 : 0);
    std::vector<int> nall2(nk2), kfull(nk2);
    for (int j = 0; j < nk2; ++j) {
        const int kb = kvc2 > 0 ? j - kvc2 : 0;
        const int k = kb * nmod + kvc;                    // index of k-point kb nmod on the SCF grid
        kfull[j] = k;
        std::copy(co + cstride * k, co + cstride * (k + 1), co2.begin() + cstride * j);
        std::copy(eps + (size_t)nc * k, eps + (size_t)nc * (k + 1), eps2.begin() + (size_t)nc * j);
        nall2[j] = nall[k];
    }
    const int rc = mp2_periodic(ctx, kvc2, co2.data(), eps2.data(), nall2.data(), icore, iocc, ivirtcore, transform,
                                e2s, e2t, seconds, nqp, nqp > 0 ? q2.data() : nullptr);
    if (rc || nqp <= 0) return rc;
    const int nk = kvc > 0 ? 2 * kvc : 1;
    std::copy(eps, eps + (size_t)nc * nk, qeps);
    for (int j = 0; j < nk2; ++j)
        std::copy(q2.begin() + (size_t)nc * j, q2.begin() + (size_t)nc * (j + 1), qeps + (size_t)nc * kfull[j]);
    return 0;
}

int polyeri_mp2_energy_periodic_mod(polyeri_ctx* ctx, int kvc, int nmod, const double* co, const double* eps,
                                    const int* nall, int icore, int iocc, int ivirtcore, int transform, double* e2s,
                                    double* e2t, double* seconds) {
    return mp2_periodic_mod(ctx, kvc, nmod, co, eps, nall, icore, iocc, ivirtcore, transform, e2s, e2t, seconds, 0,
                            nullptr);
}

int polyeri_mp2_energy_periodic_qp(polyeri_ctx* ctx, int kvc, int nmod, const double* co, const double* eps,
                                   const int* nall, int icore, int iocc, int ivirtcore, int nqp, int transform,
                                   double* e2s, double* e2t, double* qeps, double* seconds) {
    if (nqp < 1) return fail(POLYERI_E_ARG, "polyeri_mp2_energy_periodic_qp: nqp must be positive");
    return mp2_periodic_mod(ctx, kvc, nmod, co, eps, nall, icore, iocc, ivirtcore, transform, e2s, e2t, seconds,
                            nqp, qeps);
}
