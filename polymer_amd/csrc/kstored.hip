// kstored.hip -- Coulomb/exchange matrices from stored integrals (the
// reference's DIRECT .F. path, RESTORE_FOURINDEX_ERI,
// src/pfourindexstore.f90:805-1004).
//
// The records of the .ao files are held in HBM for the life of the context
// (the reference re-reads the files every SCF iteration).  One record is one
// integral (I J_Q1 | K_Q3 L_Q3+Q2) of the stored loop, which runs over every
// index order, and adds
//     C(I, J, Q1)      += P(K, L, Q2) v
//     X(I, L, Q3 + Q2) += P(K, J, Q1 - Q3) v     (both cells in [-CEL1, CEL1])
// in the reference's layout (column-major ncgs x ncgs blocks, cell Q at
// block Q + CEL1; one block for a molecule).  HBM-bound: 16 bytes per record
// read, two FP64 atomics into J/K blocks that stay in L2.
#include <hip/hip_runtime.h>

#include "eri_kernels.h"

// cells: (Q1 + cel1) << 16 | (Q2 + cel1) << 8 | (Q3 + cel2) as in the file;
// ijkl: 1-based I, J, K, L in bytes 3 .. 0.
__global__ void __launch_bounds__(256) k_jk_stored(const StoredEri* __restrict__ rec, long long n, int nc,
                                                  int cel1, int cel2, const double* __restrict__ P,
                                                  double* __restrict__ J, double* __restrict__ K) {
    const size_t c2 = (size_t)nc * nc;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        const StoredEri r = rec[t];
        const int q1 = ((r.cells >> 16) & 255) - cel1, q2 = ((r.cells >> 8) & 255) - cel1,
                  q3 = (r.cells & 255) - cel2;
        const size_t i = ((r.ijkl >> 24) & 255) - 1, j = ((r.ijkl >> 16) & 255) - 1,
                     k = ((r.ijkl >> 8) & 255) - 1, l = (r.ijkl & 255) - 1;
        atomicAdd(&J[(size_t)(q1 + cel1) * c2 + i + nc * j], P[(size_t)(q2 + cel1) * c2 + k + nc * l] * r.v);
        const int qx = q3 + q2, qp = q1 - q3;
        if (qx >= -cel1 && qx <= cel1 && qp >= -cel1 && qp <= cel1)
            atomicAdd(&K[(size_t)(qx + cel1) * c2 + i + nc * l], P[(size_t)(qp + cel1) * c2 + k + nc * j] * r.v);
    }
}

// Dense AO tensor of a molecule from its records, in the internal
// functions (T[((mu n + nu) n + la) n + si] = (mu nu | la si)): inv maps a
// reference function to its internal function, scale is the internal
// function's factor (reference function = scale x internal function).
__global__ void __launch_bounds__(256) k_tensor_stored(const StoredEri* __restrict__ rec, long long n,
                                                      const int* __restrict__ inv, const double* __restrict__ scale,
                                                      int nbf, double* __restrict__ T) {
    const size_t n1 = nbf, n2 = n1 * n1, n3 = n2 * n1;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
         t += (long long)gridDim.x * blockDim.x) {
        const StoredEri r = rec[t];
        const int a = inv[((r.ijkl >> 24) & 255) - 1], b = inv[((r.ijkl >> 16) & 255) - 1],
                  c = inv[((r.ijkl >> 8) & 255) - 1], d = inv[(r.ijkl & 255) - 1];
        if (a < 0 || b < 0 || c < 0 || d < 0) continue;   // ghost functions carry no integrals
        T[a * n3 + b * n2 + c * n1 + d] = r.v / (scale[a] * scale[b] * scale[c] * scale[d]);
    }
}

int polyeri_jk_stored_launch(const StoredEri* rec, long long n, int nc, int cel1, int cel2, const double* P, double* J,
                             double* K, hipStream_t s) {
    if (n <= 0) return 0;
    const unsigned grid = (unsigned)std::min<long long>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_jk_stored, dim3(grid), dim3(256), 0, s, rec, n, nc, cel1, cel2, P, J, K);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int polyeri_tensor_stored_launch(const StoredEri* rec, long long n, const int* inv, const double* scale, int nbf,
                                 double* T, hipStream_t s) {
    if (n <= 0) return 0;
    const unsigned grid = (unsigned)std::min<long long>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_tensor_stored, dim3(grid), dim3(256), 0, s, rec, n, inv, scale, nbf, T);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
