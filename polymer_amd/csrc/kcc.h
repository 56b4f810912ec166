// kcc.h -- closed-shell coupled-cluster solver on the device (kcc.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <string>

struct CcSolveOut {
    double e2, e3;              // E(2), E(3)
    double e_singles, e_doubles; // sum L t1 t1, sum L t2
    double e_corr;              // E(CORRELATION) of the method
    double residual;            // max |t_new - t_old| of the last iteration
    int iterations, converged;
    double iter_ms;             // device time per iteration (average)
    double iter_gflop;          // GEMM work of one iteration, GFLOP
};

// CCSD, CCD, LCCD, LCCSD or QCISD (method) and MP2, MP3 from the active MO integrals g[p][q][r][s] = (pq|rs)
// (device, nmo^4, occupied orbitals first, no of them) and the orbital
// energies eps (host, nmo) of canonical closed-shell HF orbitals.  Every
// contraction runs on the library's FP64 MFMA GEMM.  Returns 0, or a
// negative code with *err set.
enum { CC_CCSD = 0, CC_CCD = 1, CC_LCCD = 2, CC_LCCSD = 3, CC_QCISD = 4 };
int polyeri_cc_solve(const double* dG, int nmo, int no, const double* eps, int method, double conv, int maxiter,
                     int ndiis, CcSolveOut* out, hipStream_t s, std::string* err);
