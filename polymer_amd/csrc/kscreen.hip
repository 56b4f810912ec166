// kscreen.hip -- density-weighted screening of the J/K work lists.
//
// A quartet (ab|cd) adds v D[cd] to J[ab], v D[ab] to J[cd] and v times
// D[bd], D[bc], D[ad], D[ac] to the four K blocks, with |v| <= Q_ab Q_cd.
// A quartet whose Q_ab Q_cd times the largest |D| over those six shell
// blocks is below tol changes no J/K element by more than tol and is
// dropped.  With a density difference (incremental Fock builds) the bound
// shrinks as the SCF converges, so later iterations touch few quartets.
#include "eri_kernels.h"

// dsh[A * nsh + B] = max |D[a, b]| over a in shell A, b in shell B.
__global__ void k_shell_dmax(const double* __restrict__ D, int nbf, const int* __restrict__ f0,
                             const int* __restrict__ nf, int nsh, double* __restrict__ dsh) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nsh * nsh) return;
    const int A = t / nsh, B = t % nsh;
    double m = 0.0;
    for (int a = f0[A]; a < f0[A] + nf[A]; ++a)
        for (int b = f0[B]; b < f0[B] + nf[B]; ++b) m = fmax(m, fabs(D[(size_t)a * nbf + b]));
    dsh[t] = m;
}

// Compaction of one list: every wavefront takes 64 consecutive items, keeps
// the survivors in order and reserves their slots with one atomic, so runs
// of equal primitive-loop length stay together.
__global__ void __launch_bounds__(256) k_screen(EriArgs g, const double* __restrict__ Q,
                                                const int2* __restrict__ psh, const double* __restrict__ dsh,
                                                int nsh, double tol, int2* __restrict__ items,
                                                unsigned* __restrict__ count) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    const int lane = threadIdx.x & 63;
    for (long long base = (long long)blockIdx.x * blockDim.x + (threadIdx.x & ~63); base < g.nwork;
         base += stride) {
        const long long t = base + lane;
        int bp = 0, pos = 0;
        bool keep = false;
        if (t < g.nwork) {
            int bk;
            decode_work(g, t, bp, pos, bk);
            const int kp = g.ket_ids[pos];
            const int2 ab = psh[bp], cd = psh[kp];
            const size_t a = ab.x, b = ab.y, c = cd.x, d = cd.y;
            double m = fmax(dsh[c * nsh + d], dsh[a * nsh + b]);
            m = fmax(m, fmax(dsh[b * nsh + d], dsh[b * nsh + c]));
            m = fmax(m, fmax(dsh[a * nsh + d], dsh[a * nsh + c]));
            keep = Q[bp] * Q[kp] * m >= tol;
        }
        const unsigned long long mask = __ballot(keep);
        const unsigned below = __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                                         __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
        unsigned slot = 0;
        if (lane == 0 && mask) slot = atomicAdd(count, (unsigned)__popcll(mask));
        slot = __shfl(slot, 0, 64);
        if (keep) items[slot + below] = make_int2(bp, pos);
    }
}

int polyeri_shell_dmax(const double* D, int nbf, const int* shell_f0, const int* shell_nf, int nsh,
                       double* dsh, hipStream_t stream) {
    const int n = nsh * nsh;
    if (n <= 0) return 0;
    hipLaunchKernelGGL(k_shell_dmax, dim3((n + 255) / 256), dim3(256), 0, stream, D, nbf, shell_f0, shell_nf, nsh,
                       dsh);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

__global__ void __launch_bounds__(256) k_max_reduce(const double* __restrict__ x, size_t n, double* __restrict__ out) {
    __shared__ double m[256];
    double v = 0.0;
    for (size_t i = threadIdx.x; i < n; i += blockDim.x) v = fmax(v, x[i]);
    m[threadIdx.x] = v;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) m[threadIdx.x] = fmax(m[threadIdx.x], m[threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = m[0];
}

int polyeri_max_reduce(const double* x, size_t n, double* out, hipStream_t stream) {
    hipLaunchKernelGGL(k_max_reduce, dim3(1), dim3(256), 0, stream, x, n, out);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

int polyeri_screen_list(const EriArgs& a, const double* Q, const int2* pair_shells, const double* dsh,
                        int nsh, double tol, int2* items, unsigned* count, hipStream_t stream) {
    if (a.nwork <= 0) return 0;
    long long blocks = (a.nwork + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_screen, dim3((unsigned)blocks), dim3(256), 0, stream, a, Q, pair_shells, dsh, nsh, tol,
                       items, count);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}
