// kgrad.hip -- launchers of the two-electron gradient kernels (kgrad.h):
// the Schwarz bounds of the sorted pairs, the list of surviving quartets
// (pairs ordered by bound inside each primitive count, so the survivors of a
// bra are prefixes, enumerated as chunks), then one launch per class pair,
// whose kernels live in kgrad/kgrad_xX.hip (one translation unit per bra
// pair class, compiled in parallel).
#include "kgrad.h"

#include <algorithm>

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

bool polyeri_grad_sort_pairs(int* pa, int* pb, double* qp, const int* np, const int* seg) {
    for (int t = 0; t < seg[6]; ++t)
        if (!std::isfinite(qp[t])) return false;
    std::vector<int> idx, ta, tb;
    std::vector<double> tq;
    for (int X = 0; X < 6; ++X) {
        const int s0 = seg[X], n = seg[X + 1] - seg[X];
        if (n <= 1) continue;
        idx.resize(n);
        for (int k = 0; k < n; ++k) idx[k] = s0 + k;
        auto prim = [&](int t) { return np[pa[t]] * np[pb[t]]; };
        std::stable_sort(idx.begin(), idx.end(), [&](int u, int v) {
            if (prim(u) != prim(v)) return prim(u) > prim(v);
            return qp[u] > qp[v];
        });
        ta.resize(n); tb.resize(n); tq.resize(n);
        for (int k = 0; k < n; ++k) { ta[k] = pa[idx[k]]; tb[k] = pb[idx[k]]; tq[k] = qp[idx[k]]; }
        std::copy(ta.begin(), ta.end(), pa + s0);
        std::copy(tb.begin(), tb.end(), pb + s0);
        std::copy(tq.begin(), tq.end(), qp + s0);
    }
    return true;
}

void polyeri_grad_chunks(const int* pa, const int* pb, const int* np, const double* qp, const int* seg, double tol,
                         GradChunks& ch) {
    ch.off.clear();
    ch.ij.clear();
    for (int X = 0; X < 6; ++X)
        for (int Y = 0; Y < 6; ++Y) {
            const int k = 6 * X + Y;
            ch.off0[k] = ch.ij0[k] = 0;
            ch.n[k] = 0;
            ch.nq[k] = 0;
            if (Y > X) continue;
            const int x0 = seg[X], nx = seg[X + 1] - seg[X], y0 = seg[Y], ny = seg[Y + 1] - seg[Y];
            ch.off0[k] = ch.off.size();
            ch.ij0[k] = ch.ij.size();
            if (nx <= 0 || ny <= 0) continue;
            // groups of Y: runs of one primitive count
            std::vector<int> gs{0};
            for (int j = 1; j < ny; ++j)
                if (np[pa[y0 + j]] * np[pb[y0 + j]] != np[pa[y0 + j - 1]] * np[pb[y0 + j - 1]]) gs.push_back(j);
            gs.push_back(ny);
            unsigned long long tot = 0;
            for (int i = 0; i < nx; ++i) {
                const double qi = qp[x0 + i];
                for (size_t gr = 0; gr + 1 < gs.size(); ++gr) {
                    const int lo = gs[gr], hi = X == Y ? std::min(gs[gr + 1], i + 1) : gs[gr + 1];
                    if (hi <= lo) continue;
                    // first j of [lo, hi) with qi Q_j < tol (the products fall with j)
                    int a = lo, b = hi;
                    while (a < b) {
                        const int m = (a + b) >> 1;
                        if (qi * qp[y0 + m] >= tol) a = m + 1;
                        else b = m;
                    }
                    if (a == lo) continue;
                    ch.off.push_back(tot);
                    ch.ij.push_back(i);
                    ch.ij.push_back(lo);
                    tot += (unsigned long long)(a - lo);
                    ++ch.n[k];
                }
            }
            ch.off.push_back(tot);
            ch.nq[k] = tot;
        }
}

int polyeri_grad_launch(const GradArgs& g, double* d_Qp, double* d_grad, hipStream_t stream, bool have_qp,
                        const GradChunks* ch, const unsigned long long* d_off, const int* d_ij) {
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
            unsigned long long nq =
                same ? (unsigned long long)nx * (nx + 1) / 2 : (unsigned long long)nx * (unsigned long long)ny;
            GradArgs gx = gq;
            if (ch) {
                const int k = 6 * X + Y;
                nq = ch->nq[k];
                if (nq == 0) continue;
                gx.coff = d_off + ch->off0[k];
                gx.cij = d_ij + ch->ij0[k];
                gx.nchunk = ch->n[k];
            }
            const unsigned long long want = (nq + 255) / 256;
            const unsigned grid = (unsigned)(want < 16384ULL ? want : 16384ULL);
            const size_t lds = 3 * (size_t)g.natom * sizeof(double);
            kRowLaunch[X](gx, Y, g.seg[X], g.seg[Y], ny, same, nq, grid, lds, d_grad, stream);
            if (hipGetLastError() != hipSuccess) return -2;
        }
    return 0;
}

int polyeri_grad_host(const GradArgs& g, double* grad) {
    const RysTabs rt = host_tabs();
    const int npair = g.npair;
    std::vector<double> Qp(npair);
    for (int t = 0; t < npair; ++t) Qp[t] = pair_bound(g, rt, g.pa[t], g.pb[t]);
    // the survivors only, as on the device
    std::vector<int> pa(g.pa, g.pa + npair), pb(g.pb, g.pb + npair);
    const bool sorted = polyeri_grad_sort_pairs(pa.data(), pb.data(), Qp.data(), g.np, g.seg);
    GradChunks ch;
    if (sorted) polyeri_grad_chunks(pa.data(), pb.data(), g.np, Qp.data(), g.seg, g.tol, ch);
    GradArgs gq = g;
    gq.Qp = Qp.data();
    gq.pa = pa.data();
    gq.pb = pb.data();
    for (int i = 0; i < 3 * g.natom; ++i) grad[i] = 0.0;
    for (int X = 0; X < 6; ++X)
        for (int Y = 0; Y <= X; ++Y) {
            const int nx = g.seg[X + 1] - g.seg[X], ny = g.seg[Y + 1] - g.seg[Y];
            if (nx <= 0 || ny <= 0) continue;
            const int same = X == Y;
            unsigned long long nq =
                same ? (unsigned long long)nx * (nx + 1) / 2 : (unsigned long long)nx * (unsigned long long)ny;
            GradArgs gx = gq;
            if (sorted) {
                const int k = 6 * X + Y;
                nq = ch.nq[k];
                gx.coff = ch.off.data() + ch.off0[k];
                gx.cij = ch.ij.data() + ch.ij0[k];
                gx.nchunk = ch.n[k];
            }
            if (nq) kRowHost[X](gx, Y, g.seg[X], g.seg[Y], ny, same, nq, grad);
        }
    return 0;
}
