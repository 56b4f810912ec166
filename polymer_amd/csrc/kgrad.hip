// kgrad.hip -- launchers of the two-electron gradient kernels (kgrad.h):
// the Schwarz bounds of the sorted pairs, then one launch per class pair,
// whose kernels live in kgrad/kgrad_xX.hip (one translation unit per bra
// pair class, compiled in parallel).
#include "kgrad.h"

namespace {

__global__ void __launch_bounds__(256) k_grad_schwarz(GradArgs g, double* Qp) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.npair) return;
    Qp[t] = pair_bound(g, device_tabs(), g.pa[t], g.pb[t]);
}


using RowLaunch = void (*)(const GradArgs&, int, int, int, int, int, unsigned long long, unsigned, size_t, double*,
                           hipStream_t);
using RowHost = void (*)(const GradArgs&, int, int, int, int, int, unsigned long long, double*);
const RowLaunch kRowLaunch[6] = {polyeri_grad_row_launch_0, polyeri_grad_row_launch_1, polyeri_grad_row_launch_2,
                                 polyeri_grad_row_launch_3, polyeri_grad_row_launch_4, polyeri_grad_row_launch_5};
const RowHost kRowHost[6] = {polyeri_grad_row_host_0, polyeri_grad_row_host_1, polyeri_grad_row_host_2,
                             polyeri_grad_row_host_3, polyeri_grad_row_host_4, polyeri_grad_row_host_5};

}  // namespace

int polyeri_grad_launch(const GradArgs& g, double* d_Qp, double* d_grad, hipStream_t stream, bool have_qp) {
    const int npair = g.npair;
    if (g.nsh <= 0 || npair <= 0 || g.natom <= 0 || 3 * g.natom * 8 > 48 * 1024) return -1;
    if (!have_qp) {
        hipLaunchKernelGGL(k_grad_schwarz, dim3((npair + 255) / 256), dim3(256), 0, stream, g, d_Qp);
        if (hipGetLastError() != hipSuccess) return -2;
    }
    GradArgs gq = g;
    gq.Qp = d_Qp;
    for (int X = 0; X < 6; ++X)
        for (int Y = 0; Y <= X; ++Y) {
            const int nx = g.seg[X + 1] - g.seg[X], ny = g.seg[Y + 1] - g.seg[Y];
            if (nx <= 0 || ny <= 0) continue;
            const int same = X == Y;
            const unsigned long long nq =
                same ? (unsigned long long)nx * (nx + 1) / 2 : (unsigned long long)nx * (unsigned long long)ny;
            const unsigned long long want = (nq + 255) / 256;
            const unsigned grid = (unsigned)(want < 16384ULL ? want : 16384ULL);
            const size_t lds = 3 * (size_t)g.natom * sizeof(double);
            kRowLaunch[X](gq, Y, g.seg[X], g.seg[Y], ny, same, nq, grid, lds, d_grad, stream);
            if (hipGetLastError() != hipSuccess) return -2;
        }
    return 0;
}

int polyeri_grad_host(const GradArgs& g, double* grad) {
    const RysTabs rt = host_tabs();
    const int npair = g.npair;
    std::vector<double> Qp(npair);
    for (int t = 0; t < npair; ++t) Qp[t] = pair_bound(g, rt, g.pa[t], g.pb[t]);
    GradArgs gq = g;
    gq.Qp = Qp.data();
    for (int i = 0; i < 3 * g.natom; ++i) grad[i] = 0.0;
    for (int X = 0; X < 6; ++X)
        for (int Y = 0; Y <= X; ++Y) {
            const int nx = g.seg[X + 1] - g.seg[X], ny = g.seg[Y + 1] - g.seg[Y];
            if (nx <= 0 || ny <= 0) continue;
            const int same = X == Y;
            const unsigned long long nq =
                same ? (unsigned long long)nx * (nx + 1) / 2 : (unsigned long long)nx * (unsigned long long)ny;
            kRowHost[X](gq, Y, g.seg[X], g.seg[Y], ny, same, nq, grad);
        }
    return 0;
}
