// kgc.hip -- launches of the generally contracted segment kernel
// (k_eri_gc, eri_kernels.h) for the classes (ss|ss) (ps|ss) (ps|ps) (ds|ss).
#include "eri_kernels.h"

// LDS bytes of k_eri_gc for bra class (la s|.
size_t polyeri_gc_lds_bytes(int la, int nbf) {
    const int na = ncart(la);
    return 8 * ((size_t)(2 * na + 2) * nbf + (size_t)(la == 0 ? 4 : 2) * na);
}

// Whether (la s|lc s) has a generally contracted kernel.
bool polyeri_gc_class(int la, int lc) {
    return (la == 0 && lc == 0) || (la == 1 && lc <= 1) || (la == 2 && lc == 0);
}

int polyeri_gc_launch(const EriArgs& a, int la, int lc, int block, int grid, hipStream_t stream) {
    if (!polyeri_gc_class(la, lc) || a.nseg <= 0 || grid <= 0) return -1;
    const size_t lds = polyeri_gc_lds_bytes(la, a.nbf);
    if (lds > 65536) return -1;
    void (*k)(EriArgs) = nullptr;
    if (la == 0) k = k_eri_gc<0, 0>;
    else if (la == 1 && lc == 0) k = k_eri_gc<1, 0>;
    else if (la == 1) k = k_eri_gc<1, 1>;
    else k = k_eri_gc<2, 0>;
    hipLaunchKernelGGL(k, dim3((unsigned)grid), dim3(block), lds, stream, a);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}
