// Two-electron gradient kernels of bra pair class 2 (kgrad.h).
#include "../kgrad.h"

void polyeri_grad_row_launch_2(const GradArgs& gq, int Y, int x0, int y0, int ny, int same, unsigned long long nq,
                               unsigned grid, size_t lds, double* d, hipStream_t s) {
    grad_class_row<2>(Y, GradLaunch{gq, x0, y0, ny, same, nq, grid, lds, d, s});
}

void polyeri_grad_row_host_2(const GradArgs& gq, int Y, int x0, int y0, int ny, int same, unsigned long long nq,
                             double* grad) {
    const RysTabs rt = host_tabs();
    grad_class_row<2>(Y, GradHostRun{gq, rt, x0, y0, ny, same, nq, grad});
}
