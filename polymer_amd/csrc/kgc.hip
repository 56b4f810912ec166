// kgc.hip -- launches of the generally contracted segment kernels
// (eri_kernels.h): k_eri_gc for the classes (ss|ss) (ps|ss) (ps|ps) (ds|ss),
// k_eri_gck for (pp|ss) (pp|ps) (ds|ps) (dp|ss).
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

// tensor: the form that writes the integrals to a.out (no LDS rows)
int polyeri_gc_launch(const EriArgs& a, int la, int lc, int block, int grid, hipStream_t stream, bool tensor) {
    if (!polyeri_gc_class(la, lc) || a.nseg <= 0 || grid <= 0) return -1;
    size_t lds = tensor ? 0 : polyeri_gc_lds_bytes(la, a.nbf);
    if (lds > 65536) return -1;
    // the coarse Boys rows go into LDS where they fit beside the rows
    const bool bt = lds + BOYS_CROWS * sizeof(double) <= 65536;
    if (bt) lds += BOYS_CROWS * sizeof(double);
    void (*k)(EriArgs) = nullptr;
    if (tensor) {
        if (!bt) return -1;
        if (la == 0) k = k_eri_gc<0, 0, true, true>;
        else if (la == 1 && lc == 0) k = k_eri_gc<1, 0, true, true>;
        else if (la == 1) k = k_eri_gc<1, 1, true, true>;
        else k = k_eri_gc<2, 0, true, true>;
    } else if (la == 0) k = bt ? k_eri_gc<0, 0, true> : k_eri_gc<0, 0, false>;
    else if (la == 1 && lc == 0) k = bt ? k_eri_gc<1, 0, true> : k_eri_gc<1, 0, false>;
    else if (la == 1) k = bt ? k_eri_gc<1, 1, true> : k_eri_gc<1, 1, false>;
    else k = bt ? k_eri_gc<2, 0, true> : k_eri_gc<2, 0, false>;
    hipLaunchKernelGGL(k, dim3((unsigned)grid), dim3(block), lds, stream, a);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}

size_t polyeri_gck_lds_bytes(int la, int lb, int nbf) {
    return 8 * ((size_t)(ncart(la) + ncart(lb)) * nbf + (size_t)ncart(la) * ncart(lb));
}

bool polyeri_gck_class(int la, int lb, int lc) {
    return (la == 1 && lb == 1 && lc <= 1) || (la == 2 && lb == 0 && lc == 1) || (la == 2 && lb == 1 && lc == 0);
}

int polyeri_gck_launch(const EriArgs& a, int la, int lb, int lc, int block, int grid, hipStream_t stream) {
    if (!polyeri_gck_class(la, lb, lc) || a.nseg <= 0 || grid <= 0) return -1;
    size_t lds = polyeri_gck_lds_bytes(la, lb, a.nbf);
    if (lds > 65536) return -1;
    const bool bt = lds + BOYS_CROWS * sizeof(double) <= 65536;
    if (bt) lds += BOYS_CROWS * sizeof(double);
    void (*k)(EriArgs) = nullptr;
    if (la == 1 && lc == 0) k = bt ? k_eri_gck<1, 1, 0, true> : k_eri_gck<1, 1, 0, false>;
    else if (la == 1) k = bt ? k_eri_gck<1, 1, 1, true> : k_eri_gck<1, 1, 1, false>;
    else if (lb == 0) k = bt ? k_eri_gck<2, 0, 1, true> : k_eri_gck<2, 0, 1, false>;
    else k = bt ? k_eri_gck<2, 1, 0, true> : k_eri_gck<2, 1, 0, false>;
    hipLaunchKernelGGL(k, dim3((unsigned)grid), dim3(block), lds, stream, a);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}
