// kgemm.hip -- strided, batched FP64 matrix products on the MFMA units for
// the MP2 four-index transform.
//
//   C_b(m,n) = sum_k A_b(m,k) B_b(k,n),  X_b(r,c) = X[b*sX + r*sXr + c*sXc]
//
// GEMM_F64: v_mfma_f64_16x16x4f64, FP64 products and accumulation.
// GEMM_BF16X3: each FP64 operand split into three BF16 pieces
// (a = a0 + a1 + a2 + O(2^-24 |a|)); the six products a_i b_j with
// i + j <= 2 on v_mfma_f32_16x16x16bf16_1k with FP32 accumulation, the
// result converted to FP64 (about FP32 accuracy at BF16 MFMA rate).
//
// BF16x3: one wavefront per 16 x 16 tile of C (four per block); lane l
// holds row l % 16 of the A slice and column l % 16 of the B slice (k =
// 4 (l / 16) .. +3) and receives rows 4q .. 4q+3 of column l % 16 of the
// result, q = l / 16.  The v_mfma_f64_16x16x4f64 lane layout of the FP64
// kernel: A row l % 16, B column l % 16, k = l / 16; result rows q, q+4,
// q+8, q+12 of column l % 16.  Rows, columns and k out of range are zero,
// so any size is accepted.
#include <hip/hip_runtime.h>

#include "eri_kernels.h"

namespace {

typedef __attribute__((__vector_size__(4 * sizeof(double)))) double doublex4;
typedef __attribute__((__vector_size__(4 * sizeof(short)))) short bf16x4;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;

struct GemmArgs {
    int M, N, K;
    const double* A;
    long sam, sak, sA;
    const double* B;
    long sbk, sbn, sB;
    double* C;
    long scm, scn, sC;
};

// LDS-tiled FP64 product.  A block of 256 threads (four wavefronts in a
// WM x WN grid) computes a BM x BN tile of C, each wavefront TMW x TNW tiles
// of 16 x 16 on v_mfma_f64_16x16x4f64, so an A fragment read from LDS feeds
// TNW products and a B fragment TMW.  The tile shape follows the product: a
// short M (up to 64 rows, the first quarter transform's occupied orbitals)
// takes one row of wavefronts over 128 columns; a narrow N (up to 64
// columns) one column of wavefronts over 128 rows; otherwise 64 x 128, or
// 32 x 160 when N fits 160 (one column block, so A is streamed once).
//
// K advances 16 at a time through two LDS buffers: the slices one and two
// steps ahead are loaded from global memory into registers while the current
// one feeds the MFMA units, the next one then stored into the other buffer,
// one barrier per step.  Staging runs along the unit-stride index of each
// operand (see the staging map in the kernel).  The loads are buffer loads
// relative to a wavefront-uniform base with 32-bit byte offsets fixed per
// thread; a row, column or k out of range gets an offset past the buffer's
// record count, which the hardware returns as zero without a memory access,
// so the loop has no branches.  LDS rows are padded to 16 mod 32 doubles so
// that the four k rows of a fragment read fall on different bank halves.  At
// most 80 KB of LDS and 256 registers per wavefront, so two blocks share a
// CU: one block's barrier and load waits are covered by the other's products.
constexpr unsigned kOob = 0x80000000u;          // byte offset past kRecords: reads as zero
constexpr int kRecords = 0x7fffffff;
constexpr int kRsrcWord3 = 0x00020000;          // raw buffer, 32-bit data format (gfx9 family)

__device__ __forceinline__ double buf_load(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}

// LDS column swizzle of k row kk for operands staged k fastest (mode 1):
// 16 lanes store one column's 16 k rows; with rows 16 mod 32 doubles apart
// they would fall on two bank pairs.  XOR-ing the column's low three bits
// with kk / 2 spreads them over 16 (the even rows over the first 16 pairs,
// the odd ones, shifted by the row pitch, over the other 16), and a
// fragment read (16 consecutive columns of one row) still covers the same
// 16 aligned columns.
template <int MODE>
__device__ __forceinline__ int swz(int kk) {
    return MODE == 1 ? (kk >> 1) & 7 : 0;
}

template <int WM, int WN, int TMW, int TNW, bool AKC, bool BKC>
__global__ void __launch_bounds__(64 * WM * WN, 2) k_gemm_f64_tiled(GemmArgs g, int mblocks) {
    constexpr int NT = 64 * WM * WN, BM = WM * TMW * 16, BN = WN * TNW * 16, TK = 16;
    constexpr int LA = BM + (48 - BM % 32) % 32, LB = BN + (48 - BN % 32) % 32;
    constexpr int SA = BM * TK, SB = BN * TK;                     // slice elements
    constexpr int NA = (SA + NT - 1) / NT, NB = (SB + NT - 1) / NT;  // per thread (the last one partial)
    static_assert(2 * TK * (LA + LB) * 8 <= 80 * 1024, "two blocks per CU");
    __shared__ double As[2][TK][LA];
    __shared__ double Bs[2][TK][LB];
    const int mb = (int)(blockIdx.x % (unsigned)mblocks), nb = (int)(blockIdx.x / (unsigned)mblocks);
    const long bat = blockIdx.y;
    const int m0 = mb * BM, n0 = nb * BN;
    const double* A = g.A + bat * g.sA + (long)m0 * g.sam;
    const double* B = g.B + bat * g.sB + (long)n0 * g.sbn;
    const int t = threadIdx.x, lane = t & 63;
    const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
    static_assert(NT == 256 && TK == 16, "four wavefronts, 16-deep slices");
    // Staging map of the A slice (TK x BM): element p of this thread has k
    // index a_kk(p) and row a_mm(p).  Three forms:
    //  AMODE 1 (sak == 1): k = t & 15, rows t / 16 + 16 p: 16 lanes read one
    //          row's 16 consecutive k;
    //  AMODE 2 (rows contiguous, BM a multiple of 64): wavefront w stages k
    //          rows 4w .. 4w + 3, each in BM / 64 runs of 64 consecutive rows;
    //          the buffer base sits at row 4w, so the per-lane offsets span
    //          three k strides at most;
    //  AMODE 3 (otherwise): element t + 256 p in k-major order.
    // Likewise for B (column b_nn).
    constexpr int AMODE = AKC ? 1 : (BM % 64 == 0 ? 2 : 3);
    constexpr int BMODE = BKC ? 1 : (BN % 64 == 0 ? 2 : 3);
    auto a_kk = [&](int p) -> int {
        if constexpr (AMODE == 1) return t & 15;
        else if constexpr (AMODE == 2) return 4 * wave + p / (BM / 64);
        else return (t + NT * p) / BM;
    };
    auto a_mm = [&](int p) -> int {
        if constexpr (AMODE == 1) return (t >> 4) + p * (NT / 16);
        else if constexpr (AMODE == 2) return (p % (BM / 64)) * 64 + lane;
        else return (t + NT * p) % BM;
    };
    auto b_kk = [&](int p) -> int {
        if constexpr (BMODE == 1) return t & 15;
        else if constexpr (BMODE == 2) return 4 * wave + p / (BN / 64);
        else return (t + NT * p) / BN;
    };
    auto b_nn = [&](int p) -> int {
        if constexpr (BMODE == 1) return (t >> 4) + p * (NT / 16);
        else if constexpr (BMODE == 2) return (p % (BN / 64)) * 64 + lane;
        else return (t + NT * p) % BN;
    };
    const int ka0 = AMODE == 2 ? 4 * wave : 0, kb0 = BMODE == 2 ? 4 * wave : 0;   // k row of the buffer base
    unsigned aoff[NA], boff[NB];
#pragma unroll
    for (int p = 0; p < NA; ++p)
        aoff[p] = t + NT * p < SA && m0 + a_mm(p) < g.M
                      ? (unsigned)(8 * ((long)a_mm(p) * g.sam + (long)(a_kk(p) - ka0) * g.sak)) : kOob;
#pragma unroll
    for (int p = 0; p < NB; ++p)
        boff[p] = t + NT * p < SB && n0 + b_nn(p) < g.N
                      ? (unsigned)(8 * ((long)b_nn(p) * g.sbn + (long)(b_kk(p) - kb0) * g.sbk)) : kOob;
    // two register sets: the slice two steps ahead is in flight while the
    // next one is stored
    double ra0[NA], rb0[NB], ra1[NA], rb1[NB];
    auto load = [&](int k0, double (&ra)[NA], double (&rb)[NB]) {
        const __amdgpu_buffer_rsrc_t rsa = __builtin_amdgcn_make_buffer_rsrc(
            (void*)(A + (long)(k0 + ka0) * g.sak), (short)0, kRecords, kRsrcWord3);
        const __amdgpu_buffer_rsrc_t rsb = __builtin_amdgcn_make_buffer_rsrc(
            (void*)(B + (long)(k0 + kb0) * g.sbk), (short)0, kRecords, kRsrcWord3);
#pragma unroll
        for (int p = 0; p < NA; ++p) ra[p] = buf_load(rsa, k0 + a_kk(p) < g.K ? aoff[p] : kOob);
#pragma unroll
        for (int p = 0; p < NB; ++p) rb[p] = buf_load(rsb, k0 + b_kk(p) < g.K ? boff[p] : kOob);
    };
    auto store = [&](int buf, const double (&ra)[NA], const double (&rb)[NB]) {
#pragma unroll
        for (int p = 0; p < NA; ++p)
            if (SA % NT == 0 || t + NT * p < SA) As[buf][a_kk(p)][a_mm(p) ^ swz<AMODE>(a_kk(p))] = ra[p];
#pragma unroll
        for (int p = 0; p < NB; ++p)
            if (SB % NT == 0 || t + NT * p < SB) Bs[buf][b_kk(p)][b_nn(p) ^ swz<BMODE>(b_kk(p))] = rb[p];
    };
    const int r = lane & 15, q = lane >> 4;
    const int wr = (wave % WM) * TMW * 16, wc = (wave / WM) * TNW * 16;
    doublex4 acc[TMW][TNW];
#pragma unroll
    for (int i = 0; i < TMW; ++i)
#pragma unroll
        for (int j = 0; j < TNW; ++j) acc[i][j] = doublex4{0.0, 0.0, 0.0, 0.0};
    auto compute = [&](int buf) {
#pragma unroll
        for (int ks = 0; ks < TK; ks += 4) {
            double a[TMW], b[TNW];
#pragma unroll
            for (int i = 0; i < TMW; ++i) a[i] = As[buf][ks + q][(wr + i * 16 + r) ^ swz<AMODE>(ks + q)];
#pragma unroll
            for (int j = 0; j < TNW; ++j) b[j] = Bs[buf][ks + q][(wc + j * 16 + r) ^ swz<BMODE>(ks + q)];
#pragma unroll
            for (int i = 0; i < TMW; ++i)
#pragma unroll
                for (int j = 0; j < TNW; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
        }
    };
    // Slices at k >= K load as zero without a memory access, so the steps
    // past the end need no test.  Step s computes LDS buffer s & 1, stores
    // the slice of step s + 1 (register set (s + 1) & 1) into the other
    // buffer, which every wavefront finished reading before the previous
    // barrier, and loads the slice of step s + 2 into set s & 1.
    const int nk = (g.K + TK - 1) / TK;
    load(0, ra0, rb0);
    load(TK, ra1, rb1);
    store(0, ra0, rb0);
    __syncthreads();
    for (int it = 0; it < nk; it += 2) {
        load((it + 2) * TK, ra0, rb0);
        compute(0);
        store(1, ra1, rb1);
        __syncthreads();
        if (it + 1 >= nk) break;                       // block-uniform
        load((it + 3) * TK, ra1, rb1);
        compute(1);
        store(0, ra0, rb0);
        __syncthreads();
    }
    double* C = g.C + bat * g.sC;
#pragma unroll
    for (int i = 0; i < TMW; ++i)
#pragma unroll
        for (int j = 0; j < TNW; ++j)
#pragma unroll
            for (int x = 0; x < 4; ++x) {   // rows q, q + 4, q + 8, q + 12 of the tile, column r
                const int row = m0 + wr + i * 16 + q + 4 * x, col = n0 + wc + j * 16 + r;
                if (row < g.M && col < g.N) C[(long)row * g.scm + (long)col * g.scn] = acc[i][j][x];
            }
}

template <int WM, int WN, int TMW, int TNW>
static int launch_f64_tiled(const GemmArgs& g, int nbat, hipStream_t stream) {
    const bool akc = g.sak == 1, bkc = g.sbk == 1;
    constexpr int BM = WM * TMW * 16, BN = WN * TNW * 16;
    const long mbk = (g.M + BM - 1) / BM, nbk = (g.N + BN - 1) / BN;
    if (mbk * nbk > 0x7fffffffL) return -1;
    // per-thread byte offsets within one K slice must stay below the record count
    const long ka = akc ? 15 : (BM % 64 == 0 ? 3 : 15), kb = bkc ? 15 : (BN % 64 == 0 ? 3 : 15);   // k rows a lane spans
    const long amax = 8 * ((long)(BM - 1) * g.sam + ka * g.sak), bmax = 8 * ((long)(BN - 1) * g.sbn + kb * g.sbk);
    if (g.sam < 0 || g.sak < 0 || g.sbk < 0 || g.sbn < 0 || amax >= (long)kRecords - 8 || bmax >= (long)kRecords - 8)
        return -1;
    const dim3 grid((unsigned)(mbk * nbk), (unsigned)nbat), block(64 * WM * WN);
    if (akc && bkc) hipLaunchKernelGGL((k_gemm_f64_tiled<WM, WN, TMW, TNW, true, true>), grid, block, 0, stream, g, (int)mbk);
    else if (akc) hipLaunchKernelGGL((k_gemm_f64_tiled<WM, WN, TMW, TNW, true, false>), grid, block, 0, stream, g, (int)mbk);
    else if (bkc) hipLaunchKernelGGL((k_gemm_f64_tiled<WM, WN, TMW, TNW, false, true>), grid, block, 0, stream, g, (int)mbk);
    else hipLaunchKernelGGL((k_gemm_f64_tiled<WM, WN, TMW, TNW, false, false>), grid, block, 0, stream, g, (int)mbk);
    return 0;
}

// Tile shape for an M x N product (see k_gemm_f64_tiled).
static int launch_f64(const GemmArgs& g, int nbat, hipStream_t stream) {
    const int tm = (g.M + 15) / 16, tn = (g.N + 15) / 16;
    if (tm <= 4) {
        switch (tm) {
            case 1: return launch_f64_tiled<1, 4, 1, 2>(g, nbat, stream);
            case 2: return launch_f64_tiled<1, 4, 2, 2>(g, nbat, stream);
            case 3: return launch_f64_tiled<1, 4, 3, 2>(g, nbat, stream);
            default: return launch_f64_tiled<1, 4, 4, 2>(g, nbat, stream);
        }
    }
    if (tn <= 4) {
        switch (tn) {
            case 1: return launch_f64_tiled<4, 1, 2, 1>(g, nbat, stream);
            case 2: return launch_f64_tiled<4, 1, 2, 2>(g, nbat, stream);
            case 3: return launch_f64_tiled<4, 1, 2, 3>(g, nbat, stream);
            default: return launch_f64_tiled<4, 1, 2, 4>(g, nbat, stream);
        }
    }
    if (tn > 8 && tn <= 10) return launch_f64_tiled<2, 2, 1, 5>(g, nbat, stream);
    return launch_f64_tiled<2, 2, 2, 4>(g, nbat, stream);
}

__device__ __forceinline__ short bf16_bits(float x) {
    unsigned u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}
__device__ __forceinline__ float bf16_value(short h) { return __uint_as_float(((unsigned)(unsigned short)h) << 16); }

// a = p0 + p1 + p2 + O(2^-24 |a|), each piece a BF16 number.
__device__ __forceinline__ void split3(double a, short& p0, short& p1, short& p2) {
    p0 = bf16_bits((float)a);
    const double r1 = a - (double)bf16_value(p0);
    p1 = bf16_bits((float)r1);
    const double r2 = r1 - (double)bf16_value(p1);
    p2 = bf16_bits((float)r2);
}

__global__ void __launch_bounds__(256) k_gemm_bf16x3(GemmArgs g) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int tiles_n = (g.N + 15) >> 4;
    const long tile = (long)blockIdx.x * 4 + wave;
    const int tm = (int)(tile / tiles_n), tn = (int)(tile % tiles_n);
    if ((long)tm * 16 >= g.M) return;
    const long b = blockIdx.y;
    const double* A = g.A + b * g.sA;
    const double* B = g.B + b * g.sB;
    const int r = lane & 15, q = lane >> 4;
    const int m = tm * 16 + r, n = tn * 16 + r;
    floatx4 acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < g.K; k0 += 16) {
        bf16x4 a0, a1, a2, b0, b1, b2;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k = k0 + 4 * q + j;
            const double av = (m < g.M && k < g.K) ? A[m * g.sam + k * g.sak] : 0.0;
            const double bv = (n < g.N && k < g.K) ? B[k * g.sbk + n * g.sbn] : 0.0;
            short x0, x1, x2;
            split3(av, x0, x1, x2);
            a0[j] = x0; a1[j] = x1; a2[j] = x2;
            split3(bv, x0, x1, x2);
            b0[j] = x0; b1[j] = x1; b2[j] = x2;
        }
        // small terms first
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a0, b2, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1, b1, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a2, b0, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a0, b1, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a1, b0, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a0, b0, acc, 0, 0, 0);
    }
    double* C = g.C + b * g.sC;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int row = tm * 16 + 4 * q + j, col = tn * 16 + r;
        if (row < g.M && col < g.N) C[row * g.scm + col * g.scn] = (double)acc[j];
    }
}

// Planar complex products for the k-point MP2 transform:
//   C_b = A_b B_b with A = Ar + i Ai (Ai null: A real), B = Br + i Bi,
//   C = Cr + i Ci, all planes sharing one set of strides.
// Real and imaginary parts accumulate in two MFMA accumulators:
//   Cr += Ar Br - Ai Bi,  Ci += Ar Bi + Ai Br.
// Column n of B and C is (n / nci, n % nci), at offset
// (n / nci) sbo + (n % nci) sbn in B and (n / nci) sco + (n % nci) scn in C,
// so that several products sharing A run as one (nci = N: plain columns).
struct ZGemmArgs {
    int M, N, K;
    const double *Ar, *Ai;
    long sam, sak, sA;
    const double *Br, *Bi;
    long sbk, sbn, sB;
    double *Cr, *Ci;
    long scm, scn, sC;
    int nci;
    long sbo, sco;
};

// Each wavefront computes NT column tiles of one 16-row strip, so the A
// value it loads per k step feeds NT (complex: 2 NT or 4 NT) products.
template <int NT>
__global__ void __launch_bounds__(256) k_zgemm_f64(ZGemmArgs g) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int groups = (g.N + 16 * NT - 1) / (16 * NT);
    const long tile = (long)blockIdx.x * 4 + wave;
    const int tm = (int)(tile / groups), tg = (int)(tile % groups);
    if ((long)tm * 16 >= g.M) return;
    const long b = blockIdx.y;
    const bool cplx = g.Ai != nullptr;
    const long oa = b * g.sA, ob = b * g.sB;
    const int r = lane & 15, q = lane >> 4;
    const int m = tm * 16 + r;
    doublex4 cr[NT], ci[NT];
    long bcol[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        cr[t] = doublex4{0.0, 0.0, 0.0, 0.0};
        ci[t] = doublex4{0.0, 0.0, 0.0, 0.0};
        const int n = (tg * NT + t) * 16 + r;
        bcol[t] = (long)(n / g.nci) * g.sbo + (long)(n % g.nci) * g.sbn;
    }
    for (int k0 = 0; k0 < g.K; k0 += 4) {
        const int k = k0 + q;
        const bool va = m < g.M && k < g.K;
        const long ia = oa + m * g.sam + k * g.sak;
        const double ar = va ? g.Ar[ia] : 0.0;
        const double ai = (cplx && va) ? g.Ai[ia] : 0.0;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int n = (tg * NT + t) * 16 + r;
            const bool vb = n < g.N && k < g.K;
            const long ib = ob + k * g.sbk + bcol[t];
            const double br = vb ? g.Br[ib] : 0.0, bi = vb ? g.Bi[ib] : 0.0;
            cr[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, br, cr[t], 0, 0, 0);
            ci[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, bi, ci[t], 0, 0, 0);
            if (cplx) {
                cr[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(-ai, bi, cr[t], 0, 0, 0);
                ci[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, br, ci[t], 0, 0, 0);
            }
        }
    }
    const long oc = b * g.sC;
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int row = tm * 16 + q + 4 * j, col = (tg * NT + t) * 16 + r;
            if (row < g.M && col < g.N) {
                const long ic = oc + row * g.scm + (long)(col / g.nci) * g.sco + (long)(col % g.nci) * g.scn;
                g.Cr[ic] = cr[t][j];
                g.Ci[ic] = ci[t][j];
            }
        }
}

// acc += x y with x, y given as three BF16 pieces each (six products, small first)
__device__ __forceinline__ floatx4 mfma_split3(const bf16x4* x, const bf16x4* y, floatx4 acc) {
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(x[0], y[2], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(x[1], y[1], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(x[2], y[0], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(x[0], y[1], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(x[1], y[0], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(x[0], y[0], acc, 0, 0, 0);
    return acc;
}

template <int NT>
__global__ void __launch_bounds__(256) k_zgemm_bf16x3(ZGemmArgs g) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int groups = (g.N + 16 * NT - 1) / (16 * NT);
    const long tile = (long)blockIdx.x * 4 + wave;
    const int tm = (int)(tile / groups), tg = (int)(tile % groups);
    if ((long)tm * 16 >= g.M) return;
    const long b = blockIdx.y;
    const bool cplx = g.Ai != nullptr;
    const long oa = b * g.sA, ob = b * g.sB;
    const int r = lane & 15, q = lane >> 4;
    const int m = tm * 16 + r;
    floatx4 cr[NT], ci[NT];
    long bcol[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        cr[t] = floatx4{0.f, 0.f, 0.f, 0.f};
        ci[t] = floatx4{0.f, 0.f, 0.f, 0.f};
        const int n = (tg * NT + t) * 16 + r;
        bcol[t] = (long)(n / g.nci) * g.sbo + (long)(n % g.nci) * g.sbn;
    }
    for (int k0 = 0; k0 < g.K; k0 += 16) {
        bf16x4 ar[3], an[3], ai[3];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int k = k0 + 4 * q + j;
            const bool va = m < g.M && k < g.K;
            const long ia = oa + m * g.sam + k * g.sak;
            short x0, x1, x2;
            split3(va ? g.Ar[ia] : 0.0, x0, x1, x2);
            ar[0][j] = x0; ar[1][j] = x1; ar[2][j] = x2;
            const double aiv = (cplx && va) ? g.Ai[ia] : 0.0;
            split3(aiv, x0, x1, x2);
            ai[0][j] = x0; ai[1][j] = x1; ai[2][j] = x2;
            split3(-aiv, x0, x1, x2);
            an[0][j] = x0; an[1][j] = x1; an[2][j] = x2;
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            bf16x4 br[3], bi[3];
            const int n = (tg * NT + t) * 16 + r;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int k = k0 + 4 * q + j;
                const bool vb = n < g.N && k < g.K;
                const long ib = ob + k * g.sbk + bcol[t];
                short x0, x1, x2;
                split3(vb ? g.Br[ib] : 0.0, x0, x1, x2);
                br[0][j] = x0; br[1][j] = x1; br[2][j] = x2;
                split3(vb ? g.Bi[ib] : 0.0, x0, x1, x2);
                bi[0][j] = x0; bi[1][j] = x1; bi[2][j] = x2;
            }
            cr[t] = mfma_split3(ar, br, cr[t]);
            ci[t] = mfma_split3(ar, bi, ci[t]);
            if (cplx) {
                cr[t] = mfma_split3(an, bi, cr[t]);
                ci[t] = mfma_split3(ai, br, ci[t]);
            }
        }
    }
    const long oc = b * g.sC;
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int row = tm * 16 + 4 * q + j, col = (tg * NT + t) * 16 + r;
            if (row < g.M && col < g.N) {
                const long ic = oc + row * g.scm + (long)(col / g.nci) * g.sco + (long)(col % g.nci) * g.scn;
                g.Cr[ic] = (double)cr[t][j];
                g.Ci[ic] = (double)ci[t][j];
            }
        }
}

}  // namespace

int polyeri_zgemm_launch(int mode, int M, int N, int K, const double* Ar, const double* Ai, long sam, long sak,
                         long sA, const double* Br, const double* Bi, long sbk, long sbn, long sB, double* Cr,
                         double* Ci, long scm, long scn, long sC, int batch, hipStream_t stream, int nci, long sbo,
                         long sco) {
    if (M < 1 || N < 1 || K < 1 || batch < 1 || !Ar || !Br || !Bi || !Cr || !Ci) return -1;
    if (nci <= 0) nci = N;
    if (mode != GEMM_F64 && mode != GEMM_BF16X3) return -1;
    // column tiles per wavefront: enough for N (up to 4)
    const int ntile = (N + 15) / 16;
    const int nt = ntile >= 4 ? 4 : ntile;
    const long tiles = (long)((M + 15) / 16) * ((N + 16 * nt - 1) / (16 * nt));
    if ((tiles + 3) / 4 > 0x7fffffffL) return -1;
    for (int b0 = 0; b0 < batch; b0 += 65535) {       // grid.y limit
        const int nb = batch - b0 < 65535 ? batch - b0 : 65535;
        ZGemmArgs g{M, N, K, Ar + b0 * sA, Ai ? Ai + b0 * sA : nullptr, sam, sak, sA, Br + b0 * sB, Bi + b0 * sB,
                    sbk, sbn, sB, Cr + b0 * sC, Ci + b0 * sC, scm, scn, sC, nci, sbo, sco};
        const dim3 grid((unsigned)((tiles + 3) / 4), (unsigned)nb);
        if (mode == GEMM_F64) {
            if (nt == 1) hipLaunchKernelGGL(k_zgemm_f64<1>, grid, dim3(256), 0, stream, g);
            else if (nt == 2) hipLaunchKernelGGL(k_zgemm_f64<2>, grid, dim3(256), 0, stream, g);
            else if (nt == 3) hipLaunchKernelGGL(k_zgemm_f64<3>, grid, dim3(256), 0, stream, g);
            else hipLaunchKernelGGL(k_zgemm_f64<4>, grid, dim3(256), 0, stream, g);
        } else {
            if (nt == 1) hipLaunchKernelGGL(k_zgemm_bf16x3<1>, grid, dim3(256), 0, stream, g);
            else if (nt == 2) hipLaunchKernelGGL(k_zgemm_bf16x3<2>, grid, dim3(256), 0, stream, g);
            else if (nt == 3) hipLaunchKernelGGL(k_zgemm_bf16x3<3>, grid, dim3(256), 0, stream, g);
            else hipLaunchKernelGGL(k_zgemm_bf16x3<4>, grid, dim3(256), 0, stream, g);
        }
        if (hipGetLastError() != hipSuccess) return -2;
    }
    return 0;
}

int polyeri_gemm_launch(int mode, int M, int N, int K, const double* A, long sam, long sak, long sA, const double* B,
                        long sbk, long sbn, long sB, double* C, long scm, long scn, long sC, int batch,
                        hipStream_t stream) {
    if (M < 1 || N < 1 || K < 1 || batch < 1) return -1;
    if (mode != GEMM_F64 && mode != GEMM_BF16X3) return -1;
    const long tiles = (long)((M + 15) / 16) * ((N + 15) / 16);
    for (int b0 = 0; b0 < batch; b0 += 65535) {       // grid.y limit
        const int nb = batch - b0 < 65535 ? batch - b0 : 65535;
        GemmArgs g{M, N, K, A + b0 * sA, sam, sak, sA, B + b0 * sB, sbk, sbn, sB, C + b0 * sC, scm, scn, sC};
        const dim3 grid((unsigned)((tiles + 3) / 4), (unsigned)nb);
        if (mode == GEMM_F64) {
            if (launch_f64(g, nb, stream)) return -1;
        } else {
            hipLaunchKernelGGL(k_gemm_bf16x3, grid, dim3(256), 0, stream, g);
        }
        if (hipGetLastError() != hipSuccess) return -2;
    }
    return 0;
}

extern "C" int polyeri_gemm_f64(int mode, int M, int N, int K, const double* d_a, long sam, long sak, const double* d_b,
                                long sbk, long sbn, double* d_c, long scm, long scn) {
    if (M < 1 || N < 1 || K < 1 || !d_a || !d_b || !d_c) return -1;
    const int r = polyeri_gemm_launch(mode, M, N, K, d_a, sam, sak, 0, d_b, sbk, sbn, 0, d_c, scm, scn, 0, 1, nullptr);
    if (r) return -1;
    return hipDeviceSynchronize() == hipSuccess ? 0 : -2;
}
