// kcoop.hip -- cooperative shell-quartet kernel for the d-heavy classes.
//
// One wavefront (64 lanes) evaluates one contracted shell quartet: the
// class's VRR / HRR recursion is executed level by level from generated op
// tables (eri_programs.h); every node value lives in LDS, so no per-lane
// array has to fit in registers or spill to scratch.  Same integrals and
// the same J/K contributions as k_eri, gathered per output element.
#include "eri_kernels.h"
#include "eri_programs.h"

// LDS record of one prepared primitive quartet: 40 VRR coefficients, then
// F_0 .. F_L
constexpr int COOP_PREC = 40 + 4 * POLYERI_LMAX + 1;

__device__ __forceinline__ double boys_taylor(int m, double T, const double* __restrict__ tab) {
    if (T < BOYS_TMAX) {
        const int ts = (int)(T * 20.0 + 0.5);
        const double h = 0.05 * ts - T;
        const double* gg = tab + ts * BOYS_NM + m;
        double f = gg[6];
        f = gg[5] + h * (1.0 / 6.0) * f;
        f = gg[4] + h * (1.0 / 5.0) * f;
        f = gg[3] + h * (1.0 / 4.0) * f;
        f = gg[2] + h * (1.0 / 3.0) * f;
        f = gg[1] + h * (1.0 / 2.0) * f;
        return gg[0] + h * f;
    }
    const double it = 1.0 / T;
    double f = 0.5 * sqrt(POLY_PI * it);
    for (int k = 0; k < m; ++k) f *= (k + 0.5) * it;
    return f;
}

// floor(t / d) for 0 <= t < 4096 and d <= 64 from r = 1/d (FP32): t/d is
// at least 1/(2d) away from the next integer after the +1/2 shift, far
// beyond the rounding of the product.
__device__ __forceinline__ int qdiv(int t, float r) { return (int)(((float)t + 0.5f) * r); }

// NIT > 0: each lane loads its VRR op words once per wavefront into
// registers (at most NIT wavefront passes over the levels: 12 for the
// classes with L <= 6, 20 for L = 7), so no op load sits on a level's
// critical path.  NH > 0 (J/K, L = 5: at most 7 VRR passes): the HRR op
// words (at most NH passes) and the final-integral map (at most 3 passes)
// are held in registers too.
template <int MODE, int NIT, int NH = 0>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(NIT == 0 ? 1 : NIT <= 12 ? 4 : 3)))
k_eri_coop(EriArgs g, int prog, int na, int nb, int nc, int nd, int pbs) {
    // node values in the recursion's precision (FP32 for MODE_JK32)
    using real = typename std::conditional<MODE == MODE_JK32, float, double>::type;
    extern __shared__ double sl_raw[];
    real* sl = reinterpret_cast<real*>(sl_raw);
    const ProgDesc& d = g_prog_desc[prog];
    const int lane = threadIdx.x;
    const float rb = 1.0f / nb, rc = 1.0f / nc, rd = 1.0f / nd, rcd = 1.0f / (nc * nd);
    // pass k of the VRR: op word of this lane (vmask bit k: the lane has an
    // op), a barrier after it (brk bit k: the level ends)
    constexpr bool REGOPS = NIT > 0;
    uint4 opw[REGOPS ? NIT : 1];
    unsigned vmask = 0, brk = 0;
    int nit = 0;
    if (REGOPS) {
        int lv = 0, base = d.vrr_lv[0];
#pragma unroll
        for (int k = 0; k < NIT; ++k) {
            const bool live = lv < d.nvrr_lv;
            const int end = live ? d.vrr_lv[lv + 1] : 0;
            const bool has = live && base + lane < end;
            opw[k] = has ? g_prog_vrr[base + lane] : make_uint4(0u, 0u, 0u, 0u);
            vmask |= (unsigned)has << k;
            nit += live;
            if (live && base + 64 >= end) {
                brk |= 1u << k;
                ++lv;
                base = lv < d.nvrr_lv ? d.vrr_lv[lv] : 0;
            } else {
                base += 64;
            }
        }
    }
    // HRR passes (x, y, w of the op word) and final-integral map entries
    constexpr int NHW = NH > 0 ? NH : 1, NFM = NH > 0 ? 3 : 1;
    uint32_t hx[NHW], hy[NHW], hw[NHW], fmap[NFM];
    unsigned hmask = 0, hbrk = 0;
    int nhit = 0;
    if (NH > 0) {
        int lv = 0, base = d.hrr_lv[0];
#pragma unroll
        for (int k = 0; k < NH; ++k) {
            const bool live = lv < d.nhrr_lv;
            const int end = live ? d.hrr_lv[lv + 1] : 0;
            const bool has = live && base + lane < end;
            const uint4 op = has ? g_prog_hrr[base + lane] : make_uint4(0u, 0u, 0u, 0u);
            hx[k] = op.x; hy[k] = op.y; hw[k] = op.w;
            hmask |= (unsigned)has << k;
            nhit += live;
            if (live && base + 64 >= end) {
                hbrk |= 1u << k;
                ++lv;
                base = lv < d.nhrr_lv ? d.hrr_lv[lv] : 0;
            } else {
                base += 64;
            }
        }
#pragma unroll
        for (int j = 0; j < NFM; ++j) fmap[j] = lane + 64 * j < d.nfin ? g_prog_aux[d.fin_off + lane + 64 * j] : 0u;
    }
    const long long nw = work_count(g);
    for (long long t = blockIdx.x; t < nw; t += gridDim.x) {
        int bp, kp, s3 = 0;
        if (MODE == MODE_PBC) {
            const long long u = t / g.n3;
            s3 = (int)(t - u * g.n3) - g.cel2;
            int pos, b;
            decode_work(g, u, bp, pos, b);
            kp = g.ket_ids[pos];
        } else if (MODE == MODE_SCHWARZ) {
            bp = kp = g.bra_ids[t];
        } else if (g.items) {
            const int2 it = g.items[t];
            bp = it.x; kp = g.ket_ids[it.y];
        } else {
            int pos, b;
            decode_work(g, t, bp, pos, b);
            kp = g.ket_ids[pos];
        }
        const PairMeta mb = g.meta[bp], mk = g.meta[kp];
        // Primitive quartets are prepared 2^pbs at a time, 64 >> pbs lanes
        // each (the pair data and Boys lookups of a batch in flight together):
        // record r holds the VRR coefficients, indexed by the ops' (side,
        // direction) -- [3 typ + i]: P - A (bra) / Q - C (ket); [6 + 3 typ + i]:
        // W - P / W - Q; [14 + typ]: rho/zeta, rho/eta; [16 + 8 typ + n]:
        // n/2zeta, n/2eta; [32 + n]: n/2(zeta+eta) -- then F_0 .. F_L times the
        // prefactor.  Read from LDS so that the op loop has no select of the
        // two reciprocals (which the compiler turns into a division per op)
        // and no integer-to-double conversions.
        real* pbuf = sl + d.nslot + na * nb * nc * nd + na * nb + nc * nd + na * nc + na * nd + nb * nc + nb * nd;
        for (int k = lane; k <= d.nacc; k += 64) sl[k] = real(0);
        const int npq = mb.nprim * mk.nprim;
        const int pb = 1 << pbs, lpr = 64 >> pbs;
        for (int p0 = 0; p0 < npq; p0 += pb) {
            const int nbt = npq - p0 < pb ? npq - p0 : pb;
            {
                const int j = lane >> (6 - pbs), part = lane & (lpr - 1);
                if (j < nbt) {
                    const int pp = p0 + j;
                    const PrimPair bq = g.prims[mb.prim_off + pp / mk.nprim];
                    PrimPair kq = g.prims[mk.prim_off + pp % mk.nprim];
                    if (MODE == MODE_PBC) kq.Px += s3 * g.period;
                    const double ze = bq.zeta + kq.zeta, ize = 1.0 / ze;
                    const double rho = bq.zeta * kq.zeta * ize;
                    const double PQx = bq.Px - kq.Px, PQy = bq.Py - kq.Py, PQz = bq.Pz - kq.Pz;
                    const double T = rho * (PQx * PQx + PQy * PQy + PQz * PQz);
                    const double pref = bq.K * kq.K * sqrt(ize);
                    const double roz = kq.zeta * ize, roe = bq.zeta * ize;
                    real* rec = pbuf + j * COOP_PREC;
                    for (int e = part; e < 40; e += lpr) {
                        const int typ = e >= 16 ? (e >> 3) & 1 : e >= 12 ? e & 1 : (e % 6) / 3;
                        const int i = e % 3;
                        const double pq = i == 0 ? PQx : (i == 1 ? PQy : PQz);
                        double v;
                        if (e < 6) v = typ ? (i == 0 ? kq.PAx : (i == 1 ? kq.PAy : kq.PAz))
                                           : (i == 0 ? bq.PAx : (i == 1 ? bq.PAy : bq.PAz));
                        else if (e < 12) v = typ ? roe * pq : -roz * pq;
                        else if (e < 14) v = 0.0;
                        else if (e < 16) v = typ ? roe : roz;
                        else if (e < 32) v = (e & 7) * 0.5 / (typ ? kq.zeta : bq.zeta);
                        else v = (e - 32) * 0.5 * ize;
                        rec[e] = (real)v;
                    }
                    for (int m = part; m <= d.L; m += lpr) rec[40 + m] = (real)(pref * boys_taylor(m, T, g.boys));
                }
            }
            __syncthreads();
            for (int q = 0; q < nbt; ++q) {
                const real* cf = pbuf + q * COOP_PREC;
                if (lane <= d.L) sl[d.base + lane] = cf[40 + lane];
                __syncthreads();
                // op words hold LDS byte offsets of 8-byte slots (halved for
                // FP32 slots); a target node is also added to its accumulator
                char* lb = reinterpret_cast<char*>(sl);
                constexpr int SH = sizeof(real) == 8 ? 0 : 1;
#define SLOT(off) (*reinterpret_cast<real*>(lb + ((off) >> SH)))
#define VRR_OP(op)                                                                                         \
    {                                                                                                      \
        const int typ = (op).w & 1, ti = 3 * typ + (((op).w >> 1) & 3);                                    \
        const real k3 = cf[16 + 8 * typ + (((op).w >> 3) & 7)], k5 = cf[32 + (((op).w >> 6) & 7)];         \
        const real v = cf[ti] * SLOT((op).x >> 16) + cf[6 + ti] * SLOT((op).y & 0xffff) +                  \
                       k3 * (SLOT((op).y >> 16) - cf[14 + typ] * SLOT((op).z & 0xffff)) +                  \
                       k5 * SLOT((op).z >> 16);                                                            \
        SLOT((op).x & 0xffff) = v;                                                                         \
        if ((op).w & 512) SLOT((op).w >> 16) += v;                                                         \
    }
                if (REGOPS) {
#pragma unroll
                    for (int k = 0; k < NIT; ++k) {
                        if (k < nit) {
                            // the empty asm keeps the compiler from hoisting
                            // the decoded fields of all passes out of the
                            // primitive loop (which would cost ~150 VGPRs)
                            uint4 op = opw[k];
                            asm volatile("" : "+v"(op.x), "+v"(op.y), "+v"(op.z), "+v"(op.w));
                            if ((vmask >> k) & 1) VRR_OP(op);
                            if ((brk >> k) & 1) __syncthreads();
                        }
                    }
                } else {
                    for (int lv = 0; lv < d.nvrr_lv; ++lv) {
                        for (int o = d.vrr_lv[lv] + lane; o < d.vrr_lv[lv + 1]; o += 64) {
                            const uint4 op = g_prog_vrr[o];
                            VRR_OP(op);
                        }
                        __syncthreads();
                    }
                }
#undef VRR_OP
#undef SLOT
                if (d.nacc_map > 0) {   // targets no VRR op produces
                    for (int k = lane; k < d.nacc_map; k += 64) {
                        const uint32_t a = g_prog_aux[d.acc_off + k];
                        sl[a & 0xffff] += sl[a >> 16];
                    }
                    __syncthreads();
                }
            }
        }
        const real ab[3] = {(real)g.ab[4 * bp], (real)g.ab[4 * bp + 1], (real)g.ab[4 * bp + 2]};
        const real cd[3] = {(real)g.ab[4 * kp], (real)g.ab[4 * kp + 1], (real)g.ab[4 * kp + 2]};
#define HRR_OP(ox, oy, ow)                                                          \
    {                                                                               \
        const int typ = (ow) & 1, i = ((ow) >> 1) & 3;                              \
        const real gv = typ ? (i == 0 ? cd[0] : (i == 1 ? cd[1] : cd[2]))           \
                            : (i == 0 ? ab[0] : (i == 1 ? ab[1] : ab[2]));          \
        sl[(ox) & 0xffff] = sl[(ox) >> 16] + gv * sl[(oy)];                         \
    }
        if (NH > 0) {
#pragma unroll
            for (int k = 0; k < NHW; ++k) {
                if (k < nhit) {
                    uint32_t ox = hx[k], oy = hy[k], ow = hw[k];
                    asm volatile("" : "+v"(ox), "+v"(oy), "+v"(ow));
                    if ((hmask >> k) & 1) HRR_OP(ox, oy, ow);
                    if ((hbrk >> k) & 1) __syncthreads();
                }
            }
        } else {
            for (int lv = 0; lv < d.nhrr_lv; ++lv) {
                for (int o = d.hrr_lv[lv] + lane; o < d.hrr_lv[lv + 1]; o += 64) {
                    const uint4 op = g_prog_hrr[o];
                    HRR_OP(op.x, op.y, op.w);
                }
                __syncthreads();
            }
        }
#undef HRR_OP
        const int fa = mb.fa, fb = mb.fb, fc = mk.fa, fd = mk.fb;
        const int n = g.nbf;
        if (MODE == MODE_SCHWARZ) {
            double m = 0.0;
            for (int k = lane; k < d.nfin; k += 64) {
                const uint32_t f = g_prog_aux[d.fin_off + k];
                const int a = (f >> 16) & 15, b = (f >> 20) & 15, c = (f >> 24) & 15, e = f >> 28;
                if (a == c && b == e) m = fmax(m, fabs((double)sl[f & 0xffff]));
            }
            for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
            if (lane == 0) g.out[bp] = sqrt(m);
        } else if (MODE == MODE_PBC && g.out) {   // lattice tensor: one store per integral
            const size_t off = (size_t)xcd_slot(g.jk_copies) * (2 * g.cel1 + 1) * n * n;
            const int2 cb = g.pcell[bp], ck = g.pcell[kp];
            for (int k = lane; k < d.nfin; k += 64) {
                const uint32_t f = g_prog_aux[d.fin_off + k];
                pbc_integral(g, g.Jc + off, g.Kc + off, cb, ck, s3, fa, fb, fc, fd, (f >> 16) & 15, (f >> 20) & 15,
                             (f >> 24) & 15, f >> 28, (double)sl[f & 0xffff]);
            }
        } else if (MODE == MODE_PBC) {
            // Gather digestion of the lattice sums (the blocks of pbc_digest):
            // integrals in LDS in the reference's (i j|k l) roles (i in cell 0,
            // j in cell Q1, k in cell Q3, l in cell Q3+Q2; a stored pair with
            // its swap bit set has its shells the other way round), the
            // density blocks in LDS, then one atomic per output element.
            const int2 cb = g.pcell[bp], ck = g.pcell[kp];
            const int q1 = cb.x, q2 = ck.x, c1 = g.cel1;
            const int ni = cb.y ? nb : na, nj = cb.y ? na : nb, nk = ck.y ? nd : nc, nl = ck.y ? nc : nd;
            const int fi = cb.y ? fb : fa, fj = cb.y ? fa : fb, fk = ck.y ? fd : fc, fl = ck.y ? fc : fd;
            const int qk = s3 + q2, qd = q1 - s3, qk2 = q1 - s3, qd2 = q2 + s3;
            const bool kin = qk >= -c1 && qk <= c1 && qd >= -c1 && qd <= c1;
            const bool kin2 = g.both && qk2 >= -c1 && qk2 <= c1 && qd2 >= -c1 && qd2 <= c1;
            const size_t n2 = (size_t)n * n;
            real* dv = sl + d.nslot;
            real* D1 = dv + d.nfin;          // D(Q2)[k][l]       nk x nl
            real* D2 = D1 + nk * nl;         // D(Q1-Q3)[k][j]    nk x nj
            real* D3 = D2 + nk * nj;         // D(Q1)[i][j]       ni x nj
            real* D4 = D3 + ni * nj;         // D(Q2+Q3)[i][l]    ni x nl
            for (int k = lane; k < d.nfin; k += 64) {
                const uint32_t f = g_prog_aux[d.fin_off + k];
                const int a = (f >> 16) & 15, b = (f >> 20) & 15, c = (f >> 24) & 15, e = f >> 28;
                const int i = cb.y ? b : a, j = cb.y ? a : b, kk = ck.y ? e : c, l = ck.y ? c : e;
                dv[((i * nj + j) * nk + kk) * nl + l] = sl[f & 0xffff];
            }
            const double* Dq2 = g.Dc + (size_t)(q2 + c1) * n2;
            for (int k = lane; k < nk * nl; k += 64) D1[k] = Dq2[(size_t)(fk + k / nl) * n + fl + k % nl];
            if (kin) {
                const double* Dqd = g.Dc + (size_t)(qd + c1) * n2;
                for (int k = lane; k < nk * nj; k += 64) D2[k] = Dqd[(size_t)(fk + k / nj) * n + fj + k % nj];
            }
            if (g.both) {
                const double* Dq1 = g.Dc + (size_t)(q1 + c1) * n2;
                for (int k = lane; k < ni * nj; k += 64) D3[k] = Dq1[(size_t)(fi + k / nj) * n + fj + k % nj];
            }
            if (kin2) {
                const double* Dqd2 = g.Dc + (size_t)(qd2 + c1) * n2;
                for (int k = lane; k < ni * nl; k += 64) D4[k] = Dqd2[(size_t)(fi + k / nl) * n + fl + k % nl];
            }
            __syncthreads();
            const size_t off = (size_t)xcd_slot(g.jk_copies) * (2 * c1 + 1) * n2;
            double* J = g.Jc + off;
            double* K = g.Kc + off;
            // the blocks' sums as tasks over the lanes (one output element and
            // one outer index each) meeting in an LDS buffer, as in the
            // molecular digestion below
            const int o1 = ni * nj, o2 = o1 + (kin ? ni * nl : 0), o3 = o2 + (g.both ? nk * nl : 0),
                      o4 = o3 + (kin2 ? nk * nj : 0);
            real* out = D1 + (na * nb + nc * nd + na * nc + na * nd + nb * nc + nb * nd) + (COOP_PREC << pbs);
            const int nkl = nk * nl;
            const float rnj = 1.0f / nj, rnk = 1.0f / nk, rnl = 1.0f / nl, rnkl = 1.0f / nkl;
            for (int k = lane; k < o4; k += 64) out[k] = real(0);
            __syncthreads();
            // J(Q1)[i][j] += sum_kl v D(Q2)[k][l]: task (ij, k), inner l
            for (int t = lane; t < ni * nj * nk; t += 64) {
                const int ij = qdiv(t, rnk), k = t - ij * nk;
                const real* v = dv + t * nl;
                real acc = 0;
                for (int l = 0; l < nl; ++l) acc += v[l] * D1[k * nl + l];
                atomicAdd(&out[ij], acc);
            }
            // K(Q3+Q2)[i][l] += sum_jk v D(Q1-Q3)[k][j]: task (i, j, l), inner k
            if (kin)
                for (int t = lane; t < ni * nj * nl; t += 64) {
                    const int ij = qdiv(t, rnl), l = t - ij * nl, i = qdiv(ij, rnj), j = ij - i * nj;
                    const real* v = dv + ij * nkl + l;
                    real acc = 0;
                    for (int k = 0; k < nk; ++k) acc += v[k * nl] * D2[k * nj + j];
                    atomicAdd(&out[o1 + i * nl + l], acc);
                }
            // J(Q2)[k][l] += sum_ij v D(Q1)[i][j]: task (i, kl), inner j
            if (g.both)
                for (int t = lane; t < ni * nkl; t += 64) {
                    const int i = qdiv(t, rnkl), kl = t - i * nkl;
                    const real* v = dv + i * nj * nkl + kl;
                    real acc = 0;
                    for (int j = 0; j < nj; ++j) acc += v[j * nkl] * D3[i * nj + j];
                    atomicAdd(&out[o2 + kl], acc);
                }
            // K(Q1-Q3)[k][j] += sum_il v D(Q2+Q3)[i][l]: task (i, j, k), inner l
            if (kin2)
                for (int t = lane; t < ni * nj * nk; t += 64) {
                    const int ij = qdiv(t, rnk), k = t - ij * nk, i = qdiv(ij, rnj), j = ij - i * nj;
                    const real* v = dv + t * nl;
                    real acc = 0;
                    for (int l = 0; l < nl; ++l) acc += v[l] * D4[i * nl + l];
                    atomicAdd(&out[o3 + k * nj + j], acc);
                }
            __syncthreads();
            for (int o = lane; o < o4; o += 64) {
                double* dst;
                if (o < o1) dst = &J[(size_t)(q1 + c1) * n2 + (size_t)(fi + o / nj) * n + fj + o % nj];
                else if (o < o2) dst = &K[(size_t)(qk + c1) * n2 + (size_t)(fi + (o - o1) / nl) * n + fl + (o - o1) % nl];
                else if (o < o3) dst = &J[(size_t)(q2 + c1) * n2 + (size_t)(fk + (o - o2) / nl) * n + fl + (o - o2) % nl];
                else dst = &K[(size_t)(qk2 + c1) * n2 + (size_t)(fk + (o - o3) / nj) * n + fj + (o - o3) % nj];
                jk_add_fx(dst, (double)out[o], g.jk_copies);
            }
        } else if (MODE == MODE_TENSOR) {
            TensorCb cb{g.out, (size_t)n, fa, fb, fc, fd, g.tpack};
            for (int k = lane; k < d.nfin; k += 64) {
                const uint32_t f = g_prog_aux[d.fin_off + k];
                cb((f >> 16) & 15, (f >> 20) & 15, (f >> 24) & 15, f >> 28, (double)sl[f & 0xffff]);
            }
        } else {
            // Gather digestion: the quartet's integrals are copied into a dense
            // (a,b,c,d)-ordered array and the six density blocks into LDS.  Each
            // output element of the six J/K blocks is a two-index sum; a task is
            // one output element and one value of the outer index (the inner
            // index summed in registers), the tasks of a block are spread over
            // the lanes, and the partial sums meet in an LDS buffer (LDS
            // atomics), so every lane runs the same short loop; then one global
            // atomic per output element.
            const double s = (fa == fb ? 0.5 : 1.0) * (fc == fd ? 0.5 : 1.0) * (bp == kp ? 0.5 : 1.0);
            real* dv = sl + d.nslot;
            real* Dab = dv + d.nfin;
            real* Dcd = Dab + na * nb;
            real* Dac = Dcd + nc * nd;
            real* Dad = Dac + na * nc;
            real* Dbc = Dad + na * nd;
            real* Dbd = Dbc + nb * nc;
            const int o1 = na * nb, o2 = o1 + nc * nd, o3 = o2 + na * nc, o4 = o3 + na * nd, o5 = o4 + nb * nc,
                      o6 = o5 + nb * nd;
            real* out = Dbd + nb * nd + (COOP_PREC << pbs);
            if (NH > 0) {
#pragma unroll
                for (int j = 0; j < NFM; ++j) {
                    if (lane + 64 * j < d.nfin) {
                        uint32_t f = fmap[j];
                        asm volatile("" : "+v"(f));
                        const int a = (f >> 16) & 15, b = (f >> 20) & 15, c = (f >> 24) & 15, e = f >> 28;
                        dv[((a * nb + b) * nc + c) * nd + e] = sl[f & 0xffff];
                    }
                }
            } else {
                for (int k = lane; k < d.nfin; k += 64) {
                    const uint32_t f = g_prog_aux[d.fin_off + k];
                    const int a = (f >> 16) & 15, b = (f >> 20) & 15, c = (f >> 24) & 15, e = f >> 28;
                    dv[((a * nb + b) * nc + c) * nd + e] = sl[f & 0xffff];
                }
            }
            for (int k = lane; k < na * nb; k += 64) Dab[k] = (real)g.D[(size_t)(fa + k / nb) * n + fb + k % nb];
            for (int k = lane; k < nc * nd; k += 64) Dcd[k] = (real)g.D[(size_t)(fc + k / nd) * n + fd + k % nd];
            for (int k = lane; k < na * nc; k += 64) Dac[k] = (real)g.D[(size_t)(fa + k / nc) * n + fc + k % nc];
            for (int k = lane; k < na * nd; k += 64) Dad[k] = (real)g.D[(size_t)(fa + k / nd) * n + fd + k % nd];
            for (int k = lane; k < nb * nc; k += 64) Dbc[k] = (real)g.D[(size_t)(fb + k / nc) * n + fc + k % nc];
            for (int k = lane; k < nb * nd; k += 64) Dbd[k] = (real)g.D[(size_t)(fb + k / nd) * n + fd + k % nd];
            for (int k = lane; k < o6; k += 64) out[k] = real(0);
            __syncthreads();
            const int nbcd = nb * nc * nd, ncd = nc * nd;
            // J[a,b] += sum_cd v(ab|cd) D[c,d]: task (ab, c), inner d
            for (int t = lane; t < na * nb * nc; t += 64) {
                const int ab = qdiv(t, rc), c = t - ab * nc;
                const real* v = dv + t * nd;
                const real* Dp = Dcd + c * nd;
                real acc = 0;
                for (int e = 0; e < nd; ++e) acc += v[e] * Dp[e];
                atomicAdd(&out[ab], acc);
            }
            // J[c,d] += sum_ab v(ab|cd) D[a,b]: task (a, cd), inner b
            for (int t = lane; t < na * ncd; t += 64) {
                const int a = qdiv(t, rcd), cd = t - a * ncd;
                const real* v = dv + a * nbcd + cd;
                const real* Dp = Dab + a * nb;
                real acc = 0;
                for (int b = 0; b < nb; ++b) acc += v[b * ncd] * Dp[b];
                atomicAdd(&out[o1 + cd], acc);
            }
            // K[a,c] += sum_bd v(ab|cd) D[b,d]: task (a, b, c), inner d
            for (int t = lane; t < na * nbcd / nd; t += 64) {
                const int ab = qdiv(t, rc), c = t - ab * nc, a = qdiv(ab, rb), b = ab - a * nb;
                const real* v = dv + t * nd;
                const real* Dp = Dbd + b * nd;
                real acc = 0;
                for (int e = 0; e < nd; ++e) acc += v[e] * Dp[e];
                atomicAdd(&out[o2 + a * nc + c], acc);
            }
            // K[a,d] += sum_bc v(ab|cd) D[b,c]: task (a, b, d), inner c
            for (int t = lane; t < na * nb * nd; t += 64) {
                const int ab = qdiv(t, rd), e = t - ab * nd, a = qdiv(ab, rb), b = ab - a * nb;
                const real* v = dv + ab * ncd + e;
                const real* Dp = Dbc + b * nc;
                real acc = 0;
                for (int c = 0; c < nc; ++c) acc += v[c * nd] * Dp[c];
                atomicAdd(&out[o3 + a * nd + e], acc);
            }
            // K[b,c] += sum_ad v(ab|cd) D[a,d]: task (a, b, c), inner d
            for (int t = lane; t < na * nbcd / nd; t += 64) {
                const int ab = qdiv(t, rc), c = t - ab * nc, a = qdiv(ab, rb), b = ab - a * nb;
                const real* v = dv + t * nd;
                const real* Dp = Dad + a * nd;
                real acc = 0;
                for (int e = 0; e < nd; ++e) acc += v[e] * Dp[e];
                atomicAdd(&out[o4 + b * nc + c], acc);
            }
            // K[b,d] += sum_ac v(ab|cd) D[a,c]: task (a, b, d), inner c
            for (int t = lane; t < na * nb * nd; t += 64) {
                const int ab = qdiv(t, rd), e = t - ab * nd, a = qdiv(ab, rb), b = ab - a * nb;
                const real* v = dv + ab * ncd + e;
                const real* Dp = Dac + a * nc;
                real acc = 0;
                for (int c = 0; c < nc; ++c) acc += v[c * nd] * Dp[c];
                atomicAdd(&out[o5 + b * nd + e], acc);
            }
            __syncthreads();
            const size_t off = (size_t)xcd_slot(g.jk_copies) * n * n;
            double* J = g.J + off;
            double* K = g.K + off;
            for (int o = lane; o < o6; o += 64) {
                double* dst;
                if (o < o1) dst = &J[(size_t)(fa + o / nb) * n + fb + o % nb];
                else if (o < o2) dst = &J[(size_t)(fc + (o - o1) / nd) * n + fd + (o - o1) % nd];
                else if (o < o3) dst = &K[(size_t)(fa + (o - o2) / nc) * n + fc + (o - o2) % nc];
                else if (o < o4) dst = &K[(size_t)(fa + (o - o3) / nd) * n + fd + (o - o3) % nd];
                else if (o < o5) dst = &K[(size_t)(fb + (o - o4) / nc) * n + fc + (o - o4) % nc];
                else dst = &K[(size_t)(fb + (o - o5) / nd) * n + fd + (o - o5) % nd];
                jk_add(dst, s * (double)out[o], g.jk_copies);
            }
        }
        __syncthreads();
    }
}

// LDS doubles one wavefront of the class needs: the program's node slots,
// the dense integral block and the six density blocks (J/K), the records of
// one batch of 2^pbs primitive quartets, and the J/K output elements.
int polyeri_coop_nslot(int la, int lb, int lc, int ld, int pbs) {
    const int p = prog_index(la, lb, lc, ld);
    if (p < 0) return -1;
    const int na = ncart(la), nb = ncart(lb), nc = ncart(lc), nd = ncart(ld);
    return g_prog_nslot[p] + na * nb * nc * nd + na * nb + nc * nd + na * nc + na * nd + nb * nc + nb * nd + (COOP_PREC << pbs) +
           na * nb + nc * nd + na * nc + na * nd + nb * nc + nb * nd;
}

int polyeri_coop_launch(const EriArgs& a, int mode, int la, int lb, int lc, int ld, int pbs, int max_grid,
                        hipStream_t stream) {
    const int p = prog_index(la, lb, lc, ld);
    if (p < 0 || pbs < 0 || pbs > 3 || a.nwork <= 0 || max_grid <= 0)
        return p < 0 || pbs < 0 || pbs > 3 || max_grid <= 0 ? -1 : 0;
    const int nslot = polyeri_coop_nslot(la, lb, lc, ld, pbs);
    const long long grid = a.nwork < max_grid ? a.nwork : max_grid;
    const size_t lds = (size_t)nslot * sizeof(double);
    const int na = ncart(la), nb = ncart(lb), nc = ncart(lc), nd = ncart(ld);
    // every class with L <= 6 needs at most 12 wavefront passes over its
    // VRR levels, L = 7 20 (checked against the generated programs)
    const int L = la + lb + lc + ld;
    void (*k)(EriArgs, int, int, int, int, int, int) = nullptr;
    // L = 5: at most 7 VRR passes, 6 HRR passes and 192 final integrals
    if (mode == MODE_JK)
        k = L == 5 ? k_eri_coop<MODE_JK, 7, 6> : L <= 6 ? k_eri_coop<MODE_JK, 12> : L == 7 ? k_eri_coop<MODE_JK, 20>
                                                                                           : k_eri_coop<MODE_JK, 0>;
    else if (mode == MODE_JK32) k = L <= 6 ? k_eri_coop<MODE_JK32, 12> : k_eri_coop<MODE_JK32, 0>;
    else if (mode == MODE_PBC) k = L <= 6 ? k_eri_coop<MODE_PBC, 12> : k_eri_coop<MODE_PBC, 0>;
    else if (mode == MODE_SCHWARZ) k = k_eri_coop<MODE_SCHWARZ, 0>;
    else k = k_eri_coop<MODE_TENSOR, 0>;
    hipLaunchKernelGGL(k, dim3((unsigned)grid), dim3(64), lds, stream, a, p, na, nb, nc, nd, pbs);
    return hipGetLastError() == hipSuccess ? 0 : -2;
}
