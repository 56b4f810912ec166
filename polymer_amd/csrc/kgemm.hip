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

// LDS-tiled FP64 product: a block of 256 threads computes a 64 x 128 tile
// of C; the four wavefronts take 32 x 64 quarters, each 2 x 4 tiles of
// 16 x 16 on v_mfma_f64_16x16x4f64, so every A fragment a lane reads from LDS
// feeds four products and every B fragment two.  K advances 16 at a time,
// the A (64 x 16) and B (16 x 128) slices staged in LDS by all threads with
// the unit-stride index fastest (AK1 / BK1: sak / sbk == 1), out-of-range
// rows, columns and k read as zero.  Block order: the blocks of one column
// block (all row blocks) are 8 launch slots apart, i.e. on one XCD, so the
// B slice they share is fetched into one L2.
constexpr int TM = 64, TN = 128, TK = 16;

template <bool AK1, bool BK1>
__global__ void __launch_bounds__(256) k_gemm_f64_lds(GemmArgs g, int mblocks, int nblocks) {
    __shared__ double As[TK][TM + 1];
    __shared__ double Bs[TK][TN + 1];
    const int bid = blockIdx.x, xcd = bid & 7, slot = bid >> 3;
    const int mb = slot % mblocks, nb = (slot / mblocks) * 8 + xcd;
    if (nb >= nblocks) return;                       // block-uniform, before any barrier
    const long bat = blockIdx.y;
    const double* A = g.A + bat * g.sA;
    const double* B = g.B + bat * g.sB;
    const int m0 = mb * TM, n0 = nb * TN;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wr = wave & 1, wc = wave >> 1;
    const int r = lane & 15, q = lane >> 4;
    doublex4 acc[2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = doublex4{0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < g.K; k0 += TK) {
#pragma unroll
        for (int e = threadIdx.x; e < TM * TK; e += 256) {
            const int kk = AK1 ? e % TK : e / TM, mm = AK1 ? e / TK : e % TM;
            const int m = m0 + mm, k = k0 + kk;
            As[kk][mm] = (m < g.M && k < g.K) ? A[m * g.sam + k * g.sak] : 0.0;
        }
#pragma unroll
        for (int e = threadIdx.x; e < TN * TK; e += 256) {
            const int kk = BK1 ? e % TK : e / TN, nn = BK1 ? e / TK : e % TN;
            const int n = n0 + nn, k = k0 + kk;
            Bs[kk][nn] = (n < g.N && k < g.K) ? B[k * g.sbk + n * g.sbn] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < TK; ks += 4) {
            double a[2], b[4];
#pragma unroll
            for (int i = 0; i < 2; ++i) a[i] = As[ks + q][wr * 32 + i * 16 + r];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = Bs[ks + q][wc * 64 + j * 16 + r];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    double* C = g.C + bat * g.sC;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int x = 0; x < 4; ++x) {   // rows q, q + 4, q + 8, q + 12 of the tile, column r
                const int row = m0 + wr * 32 + i * 16 + q + 4 * x, col = n0 + wc * 64 + j * 16 + r;
                if (row < g.M && col < g.N) C[row * g.scm + col * g.scn] = acc[i][j][x];
            }
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
            const int mbk = (M + TM - 1) / TM, nbk = (N + TN - 1) / TN, nb8 = (nbk + 7) / 8 * 8;
            if ((long)mbk * nb8 > 0x7fffffffL) return -1;
            const dim3 tg((unsigned)(mbk * nb8), (unsigned)nb);
            if (sak == 1 && sbk == 1) hipLaunchKernelGGL((k_gemm_f64_lds<true, true>), tg, dim3(256), 0, stream, g, mbk, nbk);
            else if (sak == 1) hipLaunchKernelGGL((k_gemm_f64_lds<true, false>), tg, dim3(256), 0, stream, g, mbk, nbk);
            else if (sbk == 1) hipLaunchKernelGGL((k_gemm_f64_lds<false, true>), tg, dim3(256), 0, stream, g, mbk, nbk);
            else hipLaunchKernelGGL((k_gemm_f64_lds<false, false>), tg, dim3(256), 0, stream, g, mbk, nbk);
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
