// kcc.hip -- spin-restricted CCSD, MP2 and MP3 on the device
// (COUPLED_CLUSTER_ENERGY with COPTN(70) = 'CCSD', MP2ENERGY_DISK and
// MP3ENERGY_DISK, src/pcc.f90:1-243, 503-807, at one k-point).
//
// The reference streams amplitude and integral records through direct-access
// files and contracts them in scalar loops.  Here the active MO integrals,
// the amplitudes and every intermediate stay in HBM, and each term of the
// amplitude equations is one tensor contraction C[lc] = alpha A[la] B[lb] +
// beta C[lc] mapped onto the library's FP64 MFMA GEMM: index groups that are
// already contiguous in an operand are passed to the GEMM as strides, other
// operands are first transposed by a gather kernel.  The equations are the
// spin-adapted closed-shell CCSD for canonical HF orbitals written with the
// intermediates F_oo, F_vv, F_ov, L_oo, L_vv, W_oooo, W_vvvv, W_voov, W_vovo;
// DIIS over the amplitude vectors.  CCD, LCCD, LCCSD and QCISD run the same
// code with term switches (no singles; no t2^2 terms; no t1-nonlinear or
// t1 t2 doubles terms).  Labels i, j, k, l run over the occupied
// orbitals, a, b, c, d over the virtual ones; integrals are in chemists'
// notation, (ia|jb) stored [i][a][j][b].
#include "kcc.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <map>
#include <set>
#include <string>
#include <vector>

#include "eri_kernels.h"

namespace {

constexpr int kMaxRank = 6;
struct Perm {
    int rank;
    long long dim[kMaxRank];
    long long sst[kMaxRank];
};

// dst[t] = alpha src[off(t)] + beta dst[t], t over dst in row-major order,
// off the source offset of the same multi-index (gather transpose).
// Tensors here stay below 2^31 elements (checked by the solver), so the
// index arithmetic runs in 32-bit unsigned integers.
__global__ void __launch_bounds__(256) k_cc_permute(const double* __restrict__ src, double* __restrict__ dst, Perm p,
                                                    long long n, double alpha, double beta) {
    for (unsigned t = blockIdx.x * 256u + threadIdx.x; t < (unsigned)n; t += gridDim.x * 256u) {
        unsigned r = t, off = 0;
        for (int d = p.rank - 1; d >= 0; --d) {
            const unsigned dd = (unsigned)p.dim[d];
            const unsigned q = r / dd;
            off += (r - q * dd) * (unsigned)p.sst[d];
            r = q;
        }
        const double v = alpha * src[off];
        dst[t] = beta == 0.0 ? v : beta * dst[t] + v;
    }
}

// out[t] = sum_s part[s mn + t], s = 0 .. ns-1 in order (split-K partials)
__global__ void __launch_bounds__(256) k_cc_sumk(const double* __restrict__ part, int ns, long long mn,
                                                 double* __restrict__ out) {
    for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < mn; t += (long long)gridDim.x * 256) {
        double v = 0.0;
        for (int q = 0; q < ns; ++q) v += part[q * mn + t];
        out[t] = v;
    }
}

// y = a x + b y (b = 0: y is not read)
__global__ void __launch_bounds__(256) k_cc_axpby(const double* __restrict__ x, double* __restrict__ y, long long n,
                                                  double a, double b) {
    for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long long)gridDim.x * 256)
        y[t] = b == 0.0 ? a * x[t] : a * x[t] + b * y[t];
}

__global__ void __launch_bounds__(256) k_cc_fill(double* __restrict__ y, long long n, double v) {
    for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long long)gridDim.x * 256) y[t] = v;
}

// New amplitudes t = r / D over [t1 | t2]: D_ia = e_i - e_a,
// D_ijab = e_i + e_j - e_a - e_b (eps on the device, occupied first).
__global__ void __launch_bounds__(256) k_cc_denom(const double* __restrict__ r, double* __restrict__ t,
                                                  const double* __restrict__ eps, int no, int nv, long long n1,
                                                  long long n) {
    for (long long x = (long long)blockIdx.x * 256 + threadIdx.x; x < n; x += (long long)gridDim.x * 256) {
        double d;
        if (x < n1) {
            const int a = (int)(x % nv), i = (int)(x / nv);
            d = eps[i] - eps[no + a];
        } else {
            long long y = x - n1;
            const int b = (int)(y % nv); y /= nv;
            const int a = (int)(y % nv); y /= nv;
            const int j = (int)(y % no), i = (int)(y / no);
            d = eps[i] + eps[j] - eps[no + a] - eps[no + b];
        }
        t[x] = r[x] / d;
    }
}

constexpr int kRedBlocks = 256;
constexpr long long kSplitMax = 512;            // split-K chunks, at most
constexpr long long kSplitBuf = 1LL << 21;      // doubles of split-K partials

// Per-block partial sums of a . b (mode 0) or of max |a - b| (mode 1);
// the host adds the kRedBlocks partials in a fixed order.
__global__ void __launch_bounds__(256) k_cc_reduce(const double* __restrict__ a, const double* __restrict__ b,
                                                   long long n, int mode, double* __restrict__ part) {
    double v = 0.0;
    for (long long t = (long long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long long)gridDim.x * 256)
        v = mode == 0 ? v + a[t] * b[t] : fmax(v, fabs(a[t] - b[t]));
    __shared__ double s[256];
    s[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] = mode == 0 ? s[threadIdx.x] + s[threadIdx.x + w]
                                                             : fmax(s[threadIdx.x], s[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

unsigned grid_of(long long n) {
    const long long b = (n + 255) / 256;
    return (unsigned)std::max<long long>(1, std::min<long long>(b, 4096));
}

struct Dev {
    std::vector<void*> ptrs;
    ~Dev() {
        for (void* p : ptrs) (void)hipFree(p);
    }
    double* get(size_t n) {
        void* p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(double)) != hipSuccess) return nullptr;
        ptrs.push_back(p);
        return (double*)p;
    }
};

struct Engine {
    hipStream_t s;
    long long no, nv;
    double *w0, *w1, *w2, *part, *wk;
    std::string* err;
    double flops = 0.0;   // GEMM flops issued (2 m n k per contraction)
    Dev* dev = nullptr;   // owner of cached layouts
    std::set<const double*> consts;               // tensors fixed for the whole solve
    std::map<std::string, double*> cache;         // their transposed copies, by (pointer, layout)

    long long dim(char c) const { return (c >= 'i' && c <= 'l') ? no : nv; }
    long long size(const std::string& l) const {
        long long n = 1;
        for (char c : l) n *= dim(c);
        return n;
    }
    int bad(const std::string& m) {
        if (err) *err = m;
        return -1;
    }

    // dst[ld] = alpha src[ls] + beta dst[ld] (same label set)
    int permute(double alpha, const double* src, const std::string& ls, double beta, double* dst,
                const std::string& ld) {
        if (ls.size() != ld.size() || (int)ld.size() > kMaxRank) return bad("cc permute: rank mismatch " + ls + "->" + ld);
        Perm p{};
        p.rank = (int)ld.size();
        for (int d = 0; d < p.rank; ++d) {
            const size_t pos = ls.find(ld[d]);
            if (pos == std::string::npos) return bad("cc permute: label " + std::string(1, ld[d]) + " missing in " + ls);
            long long st = 1;
            for (size_t e = pos + 1; e < ls.size(); ++e) st *= dim(ls[e]);
            p.dim[d] = dim(ld[d]);
            p.sst[d] = st;
        }
        const long long n = size(ld);
        hipLaunchKernelGGL(k_cc_permute, dim3(grid_of(n)), dim3(256), 0, s, src, dst, p, n, alpha, beta);
        return hipGetLastError() == hipSuccess ? 0 : bad("cc permute launch failed");
    }

    // C[lc] = alpha sum_K A[la] B[lb] + beta C[lc]; every label of C is in
    // exactly one operand, every other label in both.
    int contract(double alpha, const double* A, std::string la, const double* B, std::string lb, double beta,
                 double* C, const std::string& lc) {
        if (lc.empty()) return bad("cc contract: empty output");
        if (la.find(lc[0]) == std::string::npos) {
            std::swap(A, B);
            std::swap(la, lb);
        }
        std::string M, N, K;
        for (char c : la) {
            const bool inb = lb.find(c) != std::string::npos, inc = lc.find(c) != std::string::npos;
            if (inc && !inb) M += c;
            else if (inb && !inc) K += c;
            else return bad("cc contract: label " + std::string(1, c) + " of " + la + "," + lb + "->" + lc);
        }
        for (char c : lb) {
            const bool ina = la.find(c) != std::string::npos, inc = lc.find(c) != std::string::npos;
            if (inc && !ina) N += c;
            else if (!(ina && !inc)) return bad("cc contract: label " + std::string(1, c) + " of " + lb);
        }
        if (M.size() + N.size() != lc.size()) return bad("cc contract: output labels " + lc);
        // the contracted labels in A's order or in B's, whichever leaves
        // fewer elements to transpose this call (cached layouts are free)
        {
            std::string KB;
            for (char c : lb)
                if (K.find(c) != std::string::npos) KB += c;
            auto cost = [&](const std::string& Ko) {
                long long c = 0;
                if (!(la == M + Ko || la == Ko + M) && !cached(A, la, M + Ko)) c += size(la);
                if (!(lb == Ko + N || lb == N + Ko) && !cached(B, lb, Ko + N)) c += size(lb);
                return c;
            };
            if (KB != K && cost(KB) < cost(K)) K = KB;
        }
        const long long m = size(M), n = size(N), k = size(K);
        if (m > INT32_MAX || n > INT32_MAX || k > INT32_MAX) return bad("cc contract: dimension too large");
        const double* Ap = A;
        long sam, sak;
        if (la == M + K) { sam = (long)k; sak = 1; }
        else if (la == K + M) { sam = 1; sak = (long)m; }
        else {
            if (!(Ap = layout(A, la, M + K, w0))) return -1;
            sam = (long)k; sak = 1;
        }
        const double* Bp = B;
        long sbk, sbn;
        if (lb == K + N) { sbk = (long)n; sbn = 1; }
        else if (lb == N + K) { sbk = 1; sbn = (long)k; }
        else {
            if (!(Bp = layout(B, lb, K + N, w1))) return -1;
            sbk = (long)n; sbn = 1;
        }
        flops += 2.0 * (double)m * (double)n * (double)k;
        // Split-K for a small output with a long reduction (the GEMM spreads
        // only output tiles over the chip): ns chunks of kc as one batched
        // product into partials, the remainder as one more, summed in order.
        const long long tiles = ((m + 15) / 16) * ((n + 15) / 16);
        long long ns = 1;
        if (k >= 2048 && tiles < 1024) {
            ns = std::min<long long>(k / 512, std::max<long long>(1, 4096 / tiles));
            ns = std::min<long long>(ns, kSplitMax);
            while (ns > 1 && (ns + 1) * m * n > kSplitBuf) --ns;
        }
        if (ns > 1) {
            const long long kc = k / ns, rem = k - ns * kc;
            if (polyeri_gemm_launch(GEMM_F64, (int)m, (int)n, (int)kc, Ap, sam, sak, (long)(kc * sak), Bp, sbk, sbn,
                                    (long)(kc * sbk), wk, (long)n, 1, (long)(m * n), (int)ns, s))
                return bad("cc contract: split-K GEMM launch failed for " + la + "," + lb + "->" + lc);
            if (rem > 0 && polyeri_gemm_launch(GEMM_F64, (int)m, (int)n, (int)rem, Ap + ns * kc * sak, sam, sak, 0,
                                               Bp + ns * kc * sbk, sbk, sbn, 0, wk + ns * m * n, (long)n, 1, 0, 1, s))
                return bad("cc contract: split-K remainder launch failed");
            hipLaunchKernelGGL(k_cc_sumk, dim3(grid_of(m * n)), dim3(256), 0, s, wk, (int)(ns + (rem > 0)), m * n, w2);
            if (hipGetLastError() != hipSuccess) return bad("cc split-K sum launch failed");
            return permute(alpha, w2, M + N, beta, C, lc);
        }
        const bool direct = alpha == 1.0 && beta == 0.0 && (lc == M + N || lc == N + M);
        double* Cp = direct ? C : w2;
        long scm = (long)n, scn = 1;
        if (direct && lc != M + N) { scm = 1; scn = (long)m; }
        if (polyeri_gemm_launch(GEMM_F64, (int)m, (int)n, (int)k, Ap, sam, sak, 0, Bp, sbk, sbn, 0, Cp, scm, scn, 0, 1,
                                s))
            return bad("cc contract: GEMM launch failed for " + la + "," + lb + "->" + lc);
        if (!direct) return permute(alpha, w2, M + N, beta, C, lc);
        return 0;
    }

    bool cached(const double* p, const std::string& from, const std::string& to) const {
        return consts.count(p) && cache.count(key(p, from, to));
    }
    static std::string key(const double* p, const std::string& from, const std::string& to) {
        return std::to_string((unsigned long long)(uintptr_t)p) + ":" + from + ">" + to;
    }
    // p[from] transposed to the order `to`: a cached copy for a constant
    // tensor (made on first use), else a transpose into the workspace ws.
    const double* layout(const double* p, const std::string& from, const std::string& to, double* ws) {
        if (!consts.count(p)) return permute(1.0, p, from, 0.0, ws, to) ? nullptr : ws;
        const std::string kk = key(p, from, to);
        auto it = cache.find(kk);
        if (it != cache.end()) return it->second;
        double* q = dev->get((size_t)size(to));
        if (!q) {
            bad("cc: device allocation of a cached layout failed");
            return nullptr;
        }
        if (permute(1.0, p, from, 0.0, q, to)) return nullptr;
        cache[kk] = q;
        return q;
    }

    int axpby(const double* x, double* y, long long n, double a, double b) {
        hipLaunchKernelGGL(k_cc_axpby, dim3(grid_of(n)), dim3(256), 0, s, x, y, n, a, b);
        return hipGetLastError() == hipSuccess ? 0 : bad("cc axpby launch failed");
    }
    // y += x + x^T(jiba)  (the pair permutation P(ia, jb))
    int add_pair_sym(const double* x, double* y) {
        if (axpby(x, y, no * no * nv * nv, 1.0, 1.0)) return -1;
        return permute(1.0, x, "jiba", 1.0, y, "ijab");
    }
    int reduce(const double* a, const double* b, long long n, int mode, double* out) {
        hipLaunchKernelGGL(k_cc_reduce, dim3(kRedBlocks), dim3(256), 0, s, a, b, n, mode, part);
        if (hipGetLastError() != hipSuccess) return bad("cc reduction launch failed");
        double h[kRedBlocks];
        if (hipMemcpyAsync(h, part, sizeof(h), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return bad("cc reduction copy failed");
        double v = 0.0;
        for (int q = 0; q < kRedBlocks; ++q) v = mode == 0 ? v + h[q] : std::max(v, h[q]);
        *out = v;
        return 0;
    }
};

// Small dense solve (DIIS), partial pivoting; false when singular.
bool solve_dense(std::vector<double> A, std::vector<double>& x, int n) {
    for (int c = 0; c < n; ++c) {
        int p = c;
        for (int r = c + 1; r < n; ++r)
            if (std::fabs(A[r * n + c]) > std::fabs(A[p * n + c])) p = r;
        if (std::fabs(A[p * n + c]) < 1e-300) return false;
        if (p != c) {
            for (int q = 0; q < n; ++q) std::swap(A[c * n + q], A[p * n + q]);
            std::swap(x[c], x[p]);
        }
        for (int r = c + 1; r < n; ++r) {
            const double f = A[r * n + c] / A[c * n + c];
            for (int q = c; q < n; ++q) A[r * n + q] -= f * A[c * n + q];
            x[r] -= f * x[c];
        }
    }
    for (int c = n - 1; c >= 0; --c) {
        double v = x[c];
        for (int q = c + 1; q < n; ++q) v -= A[c * n + q] * x[q];
        x[c] = v / A[c * n + c];
    }
    return true;
}

}  // namespace

int polyeri_cc_solve(const double* dG, int nmo, int no_, const double* eps, int method, double conv, int maxiter,
                     int ndiis, CcSolveOut* out, hipStream_t s, std::string* err) {
    const long long no = no_, nv = nmo - no_, n = nmo;
    if (method < CC_CCSD || method > CC_QCISD) {
        if (err) *err = "UNKNOWN OR UNIMPLEMENTED COUPLED-CLUSTER METHOD";
        return -1;
    }
    // singles; t2^2 terms; t1-nonlinear and t1 t2 doubles terms (src/pcc.f90:39-66)
    const bool S = method == CC_CCSD || method == CC_LCCSD || method == CC_QCISD;
    const bool Q2 = method == CC_CCSD || method == CC_CCD || method == CC_QCISD;
    const bool NL = method == CC_CCSD;
    if (no < 1 || nv < 1 || !out || !eps) {
        if (err) *err = "cc: empty occupied or virtual space";
        return -1;
    }
    // the transposes index with 32-bit unsigned integers: every tensor,
    // the MO integrals included, below 2^31 elements
    const long long big = std::max({nv * nv * nv * nv, no * nv * nv * nv, no * no * nv * nv, no * no * no * no});
    if (std::max(big, n * n * n * n) >= (1LL << 31)) {
        if (err) *err = "cc: active space too large for one pass (nmo^4 or nv^4 elements beyond 2^31)";
        return -1;
    }
    ndiis = std::max(1, std::min(ndiis, 16));
    const long long o2v2 = no * no * nv * nv, n1 = no * nv, n12 = n1 + o2v2;
    Dev dv;
    Engine E{s, no, nv, nullptr, nullptr, nullptr, nullptr, nullptr, err};
    double *ovov, *oovv, *ovvo, *ovoo, *ovvv, *O2, *V2, *L, *Lp, *dEps;
    double *T, *Tn, *R, *tau, *tauF, *tk, *tl, *x, *u;
    double *Foo, *Fvv, *Fov, *Loo, *Lvv, *Xoo, *Xov, *Woooo, *Wvoov, *Wvovo, *Vd, *Ed;
    const long long wn = big;
    if (!(E.w0 = dv.get(wn)) || !(E.w1 = dv.get(wn)) || !(E.w2 = dv.get(wn)) || !(E.part = dv.get(kRedBlocks)) ||
        !(E.wk = dv.get(kSplitBuf)) ||
        !(ovov = dv.get(o2v2)) || !(oovv = dv.get(o2v2)) || !(ovvo = dv.get(o2v2)) ||
        !(ovoo = dv.get(no * no * no * nv)) || !(ovvv = dv.get(no * nv * nv * nv)) ||
        !(O2 = dv.get(no * no * no * no)) || !(V2 = dv.get(nv * nv * nv * nv)) || !(L = dv.get(o2v2)) ||
        !(Lp = dv.get(o2v2)) || !(dEps = dv.get(n)) || !(T = dv.get(n12)) || !(Tn = dv.get(n12)) ||
        !(R = dv.get(n12)) || !(tau = dv.get(o2v2)) || !(tauF = dv.get(o2v2)) || !(tk = dv.get(o2v2)) || !(tl = dv.get(o2v2)) ||
        !(x = dv.get(o2v2)) || !(u = dv.get(wn)) || !(Foo = dv.get(no * no)) || !(Fvv = dv.get(nv * nv)) ||
        !(Fov = dv.get(n1)) || !(Loo = dv.get(no * no)) || !(Lvv = dv.get(nv * nv)) || !(Xoo = dv.get(no * no)) ||
        !(Xov = dv.get(n1)) || !(Woooo = dv.get(no * no * no * no)) ||
        !(Wvoov = dv.get(o2v2)) || !(Wvovo = dv.get(o2v2)) || !(Vd = dv.get(ndiis * n12)) ||
        !(Ed = dv.get(ndiis * n12))) {
        if (err) *err = "cc: device allocation failed";
        return -2;
    }
    if (hipMemcpyAsync(dEps, eps, 8 * n, hipMemcpyHostToDevice, s) != hipSuccess) {
        if (err) *err = "cc: copy of eps failed";
        return -2;
    }
    // blocks of g[p][q][r][s]: pattern of o/v per index, source strides n^3 .. 1
    auto extract = [&](double* dst, const char* pat) -> int {
        Perm p{};
        p.rank = 4;
        long long off = 0, st = n * n * n;
        for (int d = 0; d < 4; ++d, st /= n) {
            p.dim[d] = pat[d] == 'o' ? no : nv;
            p.sst[d] = st;
            if (pat[d] == 'v') off += no * st;
        }
        const long long cnt = p.dim[0] * p.dim[1] * p.dim[2] * p.dim[3];
        hipLaunchKernelGGL(k_cc_permute, dim3(grid_of(cnt)), dim3(256), 0, s, dG + off, dst, p, cnt, 1.0, 0.0);
        return hipGetLastError() == hipSuccess ? 0 : E.bad("cc: block extraction failed");
    };
    double *oooo = u, *vvvv = E.w2;   // raw blocks, only to build O2 / V2 below
    int rc = extract(ovov, "ovov") || extract(oovv, "oovv") || extract(ovvo, "ovvo") || extract(ovoo, "ovoo") ||
             extract(ovvv, "ovvv") || extract(oooo, "oooo") || E.permute(1.0, oooo, "kilj", 0.0, O2, "klij") ||
             extract(vvvv, "vvvv") ||
             E.permute(1.0, vvvv, "acbd", 0.0, V2, "abcd") ||
             // L[i,a,j,b] = 2 (ia|jb) - (ib|ja); Lp its [i][j][a][b] order
             E.axpby(ovov, L, o2v2, 2.0, 0.0) || E.permute(-1.0, ovov, "ibja", 1.0, L, "iajb") ||
             E.permute(1.0, L, "iajb", 0.0, Lp, "ijab");
    if (rc) return -3;
    E.dev = &dv;
    for (const double* c : {(const double*)ovov, (const double*)oovv, (const double*)ovvo, (const double*)ovoo,
                            (const double*)ovvv, (const double*)O2, (const double*)V2, (const double*)L,
                            (const double*)Lp})
        E.consts.insert(c);
    double* t1 = T;
    double* t2 = T + n1;
    double* r1 = R;
    double* r2 = R + n1;
    auto zero = [&](double* y, long long cnt) -> int {
        hipLaunchKernelGGL(k_cc_fill, dim3(grid_of(cnt)), dim3(256), 0, s, y, cnt, 0.0);
        return hipGetLastError() == hipSuccess ? 0 : E.bad("cc: fill launch failed");
    };
    auto denom = [&](const double* r, double* t, long long cnt) -> int {
        hipLaunchKernelGGL(k_cc_denom, dim3(grid_of(cnt)), dim3(256), 0, s, r, t, dEps, (int)no, (int)nv, n1, cnt);
        return hipGetLastError() == hipSuccess ? 0 : E.bad("cc: denominator launch failed");
    };
    // MP2: t2(1) = (ia|jb) / D, E(2) = sum L t2
    if (zero(R, n1) || E.permute(1.0, ovov, "iajb", 0.0, r2, "ijab") || denom(R, Tn, n12)) return -3;
    if (E.reduce(Lp, Tn + n1, o2v2, 0, &out->e2)) return -3;
    // MP3: E(3) = sum L t2(2), t2(2) = (linear doubles residual of t2(1)) / D
    {
        const double* a = Tn + n1;
        double* r = r2;
        if (E.contract(1.0, O2, "klij", a, "klab", 0.0, r, "ijab") ||
            E.contract(1.0, a, "ijcd", V2, "abcd", 1.0, r, "ijab") ||
            E.contract(2.0, ovvo, "kcai", a, "kjcb", 0.0, x, "ijab") ||
            E.contract(-1.0, oovv, "kiac", a, "kjcb", 1.0, x, "ijab") ||
            E.contract(-1.0, ovvo, "kcai", a, "kjbc", 1.0, x, "ijab") ||
            E.contract(-1.0, oovv, "kibc", a, "kjac", 1.0, x, "ijab") || E.add_pair_sym(x, r))
            return -3;
        if (zero(R, n1) || denom(R, R, n12)) return -3;
        if (E.reduce(Lp, R + n1, o2v2, 0, &out->e3)) return -3;
    }
    // CCSD from zero amplitudes (the first iteration gives the MP2 amplitudes)
    if (zero(T, n12)) return -3;
    std::vector<double> B((size_t)ndiis * ndiis, 0.0);
    int nvec = 0, slot = 0, it = 0;
    double resid = 0.0;
    out->converged = 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, s);
    double f_it = 0.0;
    for (it = 1; it <= maxiter; ++it) {
        const double f0 = E.flops;
        // Term switches of the method: S singles; q2 the terms quadratic in
        // t2; nl the terms nonlinear in t1 and the t1 t2 terms of the doubles
        // (CCSD keeps all).  K skips a term whose coefficient is zero.
        auto K = [&](double a, const double* A, const char* la, const double* B, const char* lb, double b, double* C,
                     const char* lc) -> int {
            if (a != 0.0) return E.contract(a, A, la, B, lb, b, C, lc);
            return b == 0.0 ? zero(C, E.size(lc)) : 0;
        };
        const double q2 = Q2 ? 1.0 : 0.0, nl = NL ? 1.0 : 0.0, fF = (Q2 || NL) ? 1.0 : 0.0;
        // tau = t2 + nl t1 t1 (against W_oooo and (ac|bd)); tauF = q2 t2 + nl t1 t1
        // (the F intermediates); tk[ikcd] = t2 + nl t1_ic t1_kd; tl[klac] = t2 + nl t1_ka t1_lc
        rc = E.axpby(t2, tau, o2v2, 1.0, 0.0) || K(nl, t1, "ia", t1, "jb", 1.0, tau, "ijab") ||
             E.axpby(t2, tauF, o2v2, q2, 0.0) || K(nl, t1, "ia", t1, "jb", 1.0, tauF, "ijab");
        if (S)
            rc = rc || E.axpby(t2, tk, o2v2, 1.0, 0.0) || K(nl, t1, "ic", t1, "kd", 1.0, tk, "ikcd") ||
                 E.axpby(t2, tl, o2v2, 1.0, 0.0) || K(nl, t1, "ka", t1, "lc", 1.0, tl, "klac");
        // one-body intermediates
        rc = rc || K(fF, L, "kcld", tauF, "ilcd", 0.0, Foo, "ki") ||
             K(-fF, L, "kcld", tauF, "klad", 0.0, Fvv, "ac") || E.axpby(Foo, Loo, no * no, 1.0, 0.0) ||
             K(2.0 * nl, ovoo, "lcki", t1, "lc", 1.0, Loo, "ki") || K(-nl, ovoo, "kcli", t1, "lc", 1.0, Loo, "ki") ||
             E.axpby(Fvv, Lvv, nv * nv, 1.0, 0.0) || K(2.0 * nl, ovvv, "kdac", t1, "kd", 1.0, Lvv, "ac") ||
             K(-nl, ovvv, "kcad", t1, "kd", 1.0, Lvv, "ac");
        // singles residual
        if (S)
            rc = rc || K(fF, L, "kcld", t1, "ld", 0.0, Fov, "kc") || K(fF, Fvv, "ac", t1, "ic", 0.0, r1, "ia") ||
                 K(-fF, Foo, "ki", t1, "ka", 1.0, r1, "ia") || K(2.0 * q2, Fov, "kc", t2, "kica", 1.0, r1, "ia") ||
                 K(-q2, Fov, "kc", t2, "ikca", 1.0, r1, "ia") || K(nl, Fov, "kc", t1, "ic", 0.0, Xoo, "ki") ||
                 K(nl, Xoo, "ki", t1, "ka", 1.0, r1, "ia") ||
                 E.contract(2.0, ovvo, "kcai", t1, "kc", 1.0, r1, "ia") ||
                 E.contract(-1.0, oovv, "kiac", t1, "kc", 1.0, r1, "ia") ||
                 E.contract(2.0, ovvv, "kdac", tk, "ikcd", 1.0, r1, "ia") ||
                 E.contract(-1.0, ovvv, "kcad", tk, "ikcd", 1.0, r1, "ia") ||
                 E.contract(-2.0, ovoo, "lcki", tl, "klac", 1.0, r1, "ia") ||
                 E.contract(1.0, ovoo, "kcli", tl, "klac", 1.0, r1, "ia");
        else
            rc = rc || zero(r1, n1);
        // two-body intermediates
        rc = rc || E.axpby(O2, Woooo, no * no * no * no, 1.0, 0.0) ||
             K(nl, ovoo, "lcki", t1, "jc", 1.0, Woooo, "klij") || K(nl, ovoo, "kclj", t1, "ic", 1.0, Woooo, "klij") ||
             K(fF, ovov, "kcld", tauF, "ijcd", 1.0, Woooo, "klij") ||
             E.permute(1.0, ovvo, "kcai", 0.0, Wvoov, "akic") ||
             K(nl, ovvv, "kcad", t1, "id", 1.0, Wvoov, "akic") || K(-nl, ovoo, "kcli", t1, "la", 1.0, Wvoov, "akic") ||
             // y[ilda] = -q2 t2/2 - nl t1_id t1_la, z = y + q2 t2[ilad]: the amplitude
             // combinations the ovov terms of W_voov and W_vovo share
             K(-nl, t1, "id", t1, "la", 0.0, tl, "ilda") || E.axpby(t2, tl, o2v2, -0.5 * q2, 1.0) ||
             E.axpby(tl, tk, o2v2, 1.0, 0.0) || (q2 != 0.0 && E.permute(q2, t2, "ilad", 1.0, tk, "ilda")) ||
             K(fF, ovov, "ldkc", tk, "ilda", 1.0, Wvoov, "akic") ||
             K(-0.5 * q2, ovov, "lckd", t2, "ilad", 1.0, Wvoov, "akic") ||
             E.permute(1.0, oovv, "kiac", 0.0, Wvovo, "akci") ||
             K(nl, ovvv, "kdac", t1, "id", 1.0, Wvovo, "akci") || K(-nl, ovoo, "lcki", t1, "la", 1.0, Wvovo, "akci") ||
             K(fF, ovov, "lckd", tl, "ilda", 1.0, Wvovo, "akci");
        // doubles residual
        rc = rc || E.permute(1.0, ovov, "iajb", 0.0, r2, "ijab");
        if (S)
            rc = rc || E.contract(1.0, ovvv, "iacb", t1, "jc", 0.0, x, "ijab") ||
                 K(nl, oovv, "kibc", t1, "jc", 0.0, u, "kibj") || K(-nl, u, "kibj", t1, "ka", 1.0, x, "ijab") ||
                 E.add_pair_sym(x, r2) || E.contract(-1.0, ovoo, "iajk", t1, "kb", 0.0, x, "ijab") ||
                 K(nl, ovvo, "kcai", t1, "jc", 0.0, u, "kaij") || K(-nl, u, "kaij", t1, "kb", 1.0, x, "ijab") ||
                 E.add_pair_sym(x, r2);
        rc = rc || E.contract(1.0, Woooo, "klij", tau, "klab", 1.0, r2, "ijab") ||
             // W_vvvv = (ac|bd) - sum_k (kd|ac) t_kb - sum_k (kc|bd) t_ka against tau:
             // the t1 parts as Y[kaij] = sum_cd (kd|ac) tau_ijcd, then -P(t_kb Y)
             E.contract(1.0, tau, "ijcd", V2, "abcd", 1.0, r2, "ijab") ||
             (NL && (E.contract(1.0, ovvv, "kdac", tau, "ijcd", 0.0, u, "kaij") ||
                     E.contract(-1.0, u, "kaij", t1, "kb", 0.0, x, "ijab") || E.add_pair_sym(x, r2))) ||
             K(fF, Lvv, "ac", t2, "ijcb", 0.0, x, "ijab") || K(-fF, Loo, "ki", t2, "kjab", 1.0, x, "ijab") ||
             // 2 W_voov[akic] - W_vovo[akci] against the same amplitudes
             E.axpby(Wvoov, tk, o2v2, 2.0, 0.0) || E.permute(-1.0, Wvovo, "akci", 1.0, tk, "akic") ||
             E.contract(1.0, tk, "akic", t2, "kjcb", 1.0, x, "ijab") ||
             E.contract(-1.0, Wvoov, "akic", t2, "kjbc", 1.0, x, "ijab") ||
             E.contract(-1.0, Wvovo, "bkci", t2, "kjac", 1.0, x, "ijab") || E.add_pair_sym(x, r2);
        if (rc || denom(R, Tn, n12)) return -3;
        f_it = E.flops - f0;
        // DIIS over [t1 | t2]: vector Tn, error Tn - T
        double* vs = Vd + slot * n12;
        double* es = Ed + slot * n12;
        if (E.axpby(Tn, vs, n12, 1.0, 0.0) || E.axpby(Tn, es, n12, 1.0, 0.0) || E.axpby(T, es, n12, -1.0, 1.0))
            return -3;
        if (E.reduce(Tn, T, n12, 1, &resid)) return -3;
        nvec = std::min(nvec + 1, ndiis);
        for (int q = 0; q < nvec; ++q) {
            double d;
            if (E.reduce(es, Ed + q * n12, n12, 0, &d)) return -3;
            B[(size_t)slot * ndiis + q] = B[(size_t)q * ndiis + slot] = d;
        }
        if (nvec > 1) {
            const int m = nvec + 1;
            std::vector<double> A((size_t)m * m, -1.0), w(m, 0.0);
            A[(size_t)m * m - 1] = 0.0;
            for (int a = 0; a < nvec; ++a)
                for (int b = 0; b < nvec; ++b) A[(size_t)a * m + b] = B[(size_t)a * ndiis + b];
            w[nvec] = -1.0;
            if (solve_dense(A, w, m)) {
                for (int a = 0; a < nvec; ++a)
                    if (E.axpby(Vd + a * n12, T, n12, w[a], a == 0 ? 0.0 : 1.0)) return -3;
            } else if (E.axpby(Tn, T, n12, 1.0, 0.0)) {
                return -3;
            }
        } else if (E.axpby(Tn, T, n12, 1.0, 0.0)) {
            return -3;
        }
        slot = (slot + 1) % ndiis;
        if (resid < conv) {
            out->converged = 1;
            break;
        }
    }
    (void)hipEventRecord(e1, s);
    (void)hipEventSynchronize(e1);
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    out->iterations = std::min(it, maxiter);
    out->iter_ms = ms / std::max(1, out->iterations);
    out->residual = resid;
    out->iter_gflop = f_it * 1e-9;
    // E(SINGLES) = sum L_iajb t1_ia t1_jb, E(DOUBLES) = sum L t2
    if (E.contract(1.0, L, "iajb", t1, "jb", 0.0, Xov, "ia") || E.reduce(Xov, t1, n1, 0, &out->e_singles) ||
        E.reduce(Lp, t2, o2v2, 0, &out->e_doubles))
        return -3;
    // E(CORRELATION): singles + doubles for CCSD, the doubles part otherwise (src/pcc.f90:202-207)
    out->e_corr = method == CC_CCSD ? out->e_singles + out->e_doubles : out->e_doubles;
    return 0;
}
