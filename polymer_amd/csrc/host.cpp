// host.cpp -- host-side set-up of libpolyeri: basis-set reader with the
// reference's layout, contracted-shell extraction, shell pairs, Boys table
// and one-electron integrals.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <tuple>

#include "polyeri_internal.h"

namespace polyeri {

static const char* kSym[36] = {
    "H ", "He", "Li", "Be", "B ", "C ", "N ", "O ", "F ", "Ne", "Na", "Mg",
    "Al", "Si", "P ", "S ", "Cl", "Ar", "K ", "Ca", "Sc", "Ti", "V ", "Cr",
    "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr"};

void cart_comp(int l, int k, int& lx, int& ly, int& lz) {
    // xx, xy, xz, yy, yz, zz ordering (x exponent descending, then y)
    int n = 0;
    for (int i = l; i >= 0; --i)
        for (int kk = 0; kk <= l - i; ++kk) {
            if (n == k) { lx = i; ly = l - i - kk; lz = kk; return; }
            ++n;
        }
    lx = ly = lz = 0;
}

static double dfact(int n) {  // (2n-1)!!
    double r = 1.0;
    for (int k = 2 * n - 1; k > 1; k -= 2) r *= k;
    return r;
}

static double fnum(const char* s) {
    char b[64];
    size_t n = std::min<size_t>(strlen(s), 63);
    for (size_t i = 0; i < n; ++i) b[i] = (s[i] == 'D' || s[i] == 'd') ? 'E' : s[i];
    b[n] = 0;
    return strtod(b, nullptr);
}

// Gaussian-format basis file -> the reference's primitive / contracted /
// composite-shell arrays.  Layout rules follow READ3 (src/pread.f90:
// 584-1482): primitive normalisation (2a/pi)^{3/4} (4a)^{l/2} / sqrt(df)
// folded into CC, contracted functions not renormalised, composite shells
// anchored on s primitives, missing lower-l partners filled with ghost
// primitives and ghost contracted functions (CC = 0).
int read_basis(const std::string& path, int natom, const int* iatom, const double* xyz,
               RefBasis& rb, std::string& err) {
    FILE* f = fopen(path.c_str(), "r");
    if (!f) { err = "cannot open basis file " + path; return -1; }
    rb = RefBasis();
    rb.natom = natom;
    for (int i = 0; i < natom; ++i) {
        rb.iatom.push_back(iatom[i]);
        rb.ax.push_back(xyz[3 * i]); rb.ay.push_back(xyz[3 * i + 1]); rb.az.push_back(xyz[3 * i + 2]);
    }
    struct CCE { int c, p; double v; };
    std::vector<CCE> ccs;
    auto newc = [&](double x, double y, double z) {
        rb.cgsx.push_back(x); rb.cgsy.push_back(y); rb.cgsz.push_back(z);
        return rb.ncgs++;
    };
    auto newp = [&](int lx, int ly, int lz, double x, double y, double z, int at, double zt) {
        rb.pax.push_back(lx); rb.pay.push_back(ly); rb.paz.push_back(lz);
        rb.pgsx.push_back(x); rb.pgsy.push_back(y); rb.pgsz.push_back(z);
        rb.pgs_atom.push_back(at); rb.zt.push_back(zt);
        return rb.npgs++;
    };
    static const int D6[6][3] = {{2,0,0},{0,2,0},{0,0,2},{1,1,0},{0,1,1},{1,0,1}};
    static const int F10[10][3] = {{3,0,0},{0,3,0},{0,0,3},{2,1,0},{0,2,1},
                                   {1,0,2},{1,2,0},{0,1,2},{2,0,1},{1,1,1}};
    char line[512];
    for (int ia = 0; ia < natom; ++ia) {
        if (iatom[ia] < 1 || iatom[ia] > 36) { fclose(f); err = "bad atomic number"; return -2; }
        const char* sym = kSym[iatom[ia] - 1];
        double ax = xyz[3 * ia], ay = xyz[3 * ia + 1], az = xyz[3 * ia + 2];
        rewind(f);
        bool after_sep = false, found = false;
        while (fgets(line, sizeof line, f)) {
            char c0 = line[0], c1 = (line[0] && line[1] && line[1] != '\n' && line[1] != '\r') ? line[1] : ' ';
            if (after_sep && c0 == sym[0] && c1 == sym[1]) {
                found = true;
                while (fgets(line, sizeof line, f)) {
                    if (line[0] == '*' && line[1] == '*') break;
                    char typ[8], ssc[64];
                    int nprim;
                    if (sscanf(line, "%7s %d %63s", typ, &nprim, ssc) != 3) break;
                    double s = fnum(ssc);
                    int ncomp;
                    if (!strcmp(typ, "S")) ncomp = 1;
                    else if (!strcmp(typ, "P")) ncomp = 3;
                    else if (!strcmp(typ, "SP")) ncomp = 4;
                    else if (!strcmp(typ, "D")) ncomp = 6;
                    else if (!strcmp(typ, "F")) ncomp = 10;
                    else { fclose(f); err = std::string("angular momentum not supported: ") + typ; return -3; }
                    int base = rb.ncgs;
                    for (int k = 0; k < ncomp; ++k) newc(ax, ay, az);
                    for (int k = 0; k < nprim; ++k) {
                        if (!fgets(line, sizeof line, f)) break;
                        char sa[64], sb[64], sd[64];
                        int nr = sscanf(line, "%63s %63s %63s", sa, sb, sd);
                        double a = fnum(sa), b = fnum(sb), d = nr >= 3 ? fnum(sd) : 0.0;
                        double nrm = std::exp(0.75 * std::log(2.0 * a / POLY_PI));
                        double zt = a * s, q = std::sqrt(4.0 * a);
                        if (ncomp == 1) {
                            ccs.push_back({base, newp(0, 0, 0, ax, ay, az, ia, zt), b * nrm});
                        } else if (ncomp == 3 || ncomp == 4) {
                            int off = 0;
                            double pc = b;
                            if (ncomp == 4) {
                                ccs.push_back({base, newp(0, 0, 0, ax, ay, az, ia, zt), b * nrm});
                                off = 1; pc = d;
                            }
                            for (int dd = 0; dd < 3; ++dd)
                                ccs.push_back({base + off + dd, newp(dd == 0, dd == 1, dd == 2, ax, ay, az, ia, zt),
                                               pc * nrm * q});
                        } else if (ncomp == 6) {
                            for (int dd = 0; dd < 6; ++dd) {
                                double df = std::sqrt(dfact(D6[dd][0]) * dfact(D6[dd][1]) * dfact(D6[dd][2]));
                                ccs.push_back({base + dd, newp(D6[dd][0], D6[dd][1], D6[dd][2], ax, ay, az, ia, zt),
                                               b * nrm * q * q / df});
                            }
                        } else {
                            for (int dd = 0; dd < 10; ++dd) {
                                double df = std::sqrt(dfact(F10[dd][0]) * dfact(F10[dd][1]) * dfact(F10[dd][2]));
                                ccs.push_back({base + dd, newp(F10[dd][0], F10[dd][1], F10[dd][2], ax, ay, az, ia, zt),
                                               b * nrm * q * q * q / df});
                            }
                        }
                    }
                }
                break;
            }
            after_sep = (c0 == '*' && c1 == '*');
        }
        if (!found) { fclose(f); err = std::string("orbital basis set not found for ") + sym; return -4; }
    }
    fclose(f);
    // composite shells and ghosts
    std::vector<std::vector<int>> ps;   // 125 slots per shell
    std::vector<int> pang;
    auto slot = [](int x, int y, int z) { return x * 25 + y * 5 + z; };
    for (int i = 0; i < rb.npgs; ++i)
        if (rb.pax[i] + rb.pay[i] + rb.paz[i] == 0) {
            ps.emplace_back(125, -1);
            ps.back()[0] = i;
            pang.push_back(0);
        }
    int np0 = rb.npgs;
    for (int i = 0; i < np0; ++i) {
        int l = rb.pax[i] + rb.pay[i] + rb.paz[i];
        if (l == 0) continue;
        bool attached = false;
        int sl = slot(rb.pax[i], rb.pay[i], rb.paz[i]);
        for (size_t k = 0; k < ps.size(); ++k) {
            int a0 = ps[k][0];
            if (rb.pgsx[a0] == rb.pgsx[i] && rb.pgsy[a0] == rb.pgsy[i] && rb.pgsz[a0] == rb.pgsz[i] &&
                rb.zt[a0] == rb.zt[i] && ps[k][sl] == -1) {
                ps[k][sl] = i;
                pang[k] = l;
                attached = true;   // attaches to every matching anchor (no early exit)
            }
        }
        if (!attached) {
            double x = rb.pgsx[i], y = rb.pgsy[i], z = rb.pgsz[i], zt = rb.zt[i];
            int at = rb.pgs_atom[i];
            ps.emplace_back(125, -1);
            auto& s = ps.back();
            newc(x, y, z);
            s[0] = newp(0, 0, 0, x, y, z, at, zt);
            if (l >= 2)
                for (int dd = 0; dd < 3; ++dd) {
                    newc(x, y, z);
                    s[slot(dd == 0, dd == 1, dd == 2)] = newp(dd == 0, dd == 1, dd == 2, x, y, z, at, zt);
                }
            if (l >= 3)
                for (int dd = 0; dd < 6; ++dd) {
                    newc(x, y, z);
                    s[slot(D6[dd][0], D6[dd][1], D6[dd][2])] = newp(D6[dd][0], D6[dd][1], D6[dd][2], x, y, z, at, zt);
                }
            s[sl] = i;
            pang.push_back(l);
        }
    }
    rb.npshell = (int)ps.size();
    rb.cc.assign((size_t)rb.ncgs * rb.npgs, 0.0);
    for (auto& e : ccs) rb.cc[e.c + (size_t)rb.ncgs * e.p] = e.v;
    rb.p_shell.resize((size_t)rb.npshell * 125);
    for (int k = 0; k < rb.npshell; ++k) std::copy(ps[k].begin(), ps[k].end(), rb.p_shell.begin() + (size_t)k * 125);
    rb.p_shell_ang = pang;
    return 0;
}

// Contracted shells from the reference arrays: each non-ghost contracted
// function is a fixed linear combination of primitives with one Cartesian
// component and one centre; functions with the same centre, exponents and
// normalised coefficients form a shell.
int extract_shells(const RefBasis& rb, IntBasis& ib, std::string& err) {
    ib = IntBasis();
    int nc = rb.ncgs, np = rb.npgs;
    struct Fn { int c, L, lx, ly, lz, atom; double x, y, z; std::vector<double> a, v; };
    std::vector<Fn> fns;
    ib.int_of_ref.assign(nc, -1);
    for (int c = 0; c < nc; ++c) {
        Fn fn{c, -1, 0, 0, 0, 0, 0, 0, 0, {}, {}};
        for (int p = 0; p < np; ++p) {
            double v = rb.cc[c + (size_t)nc * p];
            if (v == 0.0) continue;
            int L = rb.pax[p] + rb.pay[p] + rb.paz[p];
            if (fn.L < 0) {
                fn.L = L; fn.lx = rb.pax[p]; fn.ly = rb.pay[p]; fn.lz = rb.paz[p];
                fn.x = rb.pgsx[p]; fn.y = rb.pgsy[p]; fn.z = rb.pgsz[p]; fn.atom = rb.pgs_atom[p];
            } else if (rb.pax[p] != fn.lx || rb.pay[p] != fn.ly || rb.paz[p] != fn.lz ||
                       rb.pgsx[p] != fn.x || rb.pgsy[p] != fn.y || rb.pgsz[p] != fn.z) {
                err = "contracted function mixes components or centres";
                return -1;
            }
            double df = std::sqrt(dfact(rb.pax[p]) * dfact(rb.pay[p]) * dfact(rb.paz[p]));
            fn.a.push_back(rb.zt[p]);
            fn.v.push_back(v * df);
        }
        if (fn.L < 0) continue;  // ghost
        if (fn.L > MAXL) { err = "angular momentum above d is not supported by this build"; return -2; }
        fns.push_back(fn);
    }
    // group
    std::vector<int> used(fns.size(), 0);
    std::vector<Shell> shells;
    std::vector<std::vector<int>> members;   // per shell: fns index per component
    for (size_t i = 0; i < fns.size(); ++i) {
        if (used[i]) continue;
        const Fn& f = fns[i];
        int ncomp = ncart(f.L);
        std::vector<int> mem(ncomp, -1);
        for (size_t j = i; j < fns.size(); ++j) {
            if (used[j]) continue;
            const Fn& g = fns[j];
            if (g.L != f.L || g.x != f.x || g.y != f.y || g.z != f.z || g.a.size() != f.a.size()) continue;
            bool same = true;
            for (size_t k = 0; k < f.a.size() && same; ++k)
                if (g.a[k] != f.a[k] || std::fabs(g.v[k] - f.v[k]) > 1e-12 * std::fabs(f.v[k])) same = false;
            if (!same) continue;
            int k = -1;
            for (int q = 0; q < ncomp; ++q) {
                int lx, ly, lz;
                cart_comp(f.L, q, lx, ly, lz);
                if (lx == g.lx && ly == g.ly && lz == g.lz) k = q;
            }
            if (k < 0 || mem[k] >= 0) continue;
            mem[k] = (int)j;
            used[j] = 1;
            bool full = true;
            for (int q = 0; q < ncomp; ++q) if (mem[q] < 0) full = false;
            if (full) break;
        }
        for (int q = 0; q < ncomp; ++q)
            if (mem[q] < 0) { err = "incomplete Cartesian shell in the basis"; return -3; }
        Shell s;
        s.L = f.L; s.atom = f.atom; s.x = f.x; s.y = f.y; s.z = f.z;
        s.alpha = f.a;
        s.coef = f.v;
        shells.push_back(s);
        members.push_back(mem);
    }
    // order: L descending, then atom, then input order (stable)
    std::vector<int> ord(shells.size());
    for (size_t i = 0; i < ord.size(); ++i) ord[i] = (int)i;
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
        if (shells[a].L != shells[b].L) return shells[a].L > shells[b].L;
        return shells[a].atom < shells[b].atom;
    });
    int f0 = 0;
    for (int k : ord) {
        Shell s = shells[k];
        s.f0 = f0;
        for (int q = 0; q < ncart(s.L); ++q) {
            const Fn& g = fns[members[k][q]];
            // scale = CC / (coef of the shell) = 1 / sqrt(df) of the component
            int lx, ly, lz;
            cart_comp(s.L, q, lx, ly, lz);
            double sc = 1.0 / std::sqrt(dfact(lx) * dfact(ly) * dfact(lz));
            ib.ref_of_int.push_back(g.c);
            ib.scale.push_back(sc);
            ib.int_of_ref[g.c] = f0 + q;
        }
        f0 += ncart(s.L);
        ib.shells.push_back(s);
    }
    ib.nbf = f0;
    for (int a = 0; a < rb.natom; ++a) {
        ib.natom_z.push_back(rb.iatom[a]);
        ib.atom_xyz.push_back(rb.ax[a]); ib.atom_xyz.push_back(rb.ay[a]); ib.atom_xyz.push_back(rb.az[a]);
    }
    return 0;
}

void build_pairs(const IntBasis& ib, double primtol, std::vector<PairInfo>& pairs,
                 std::vector<PrimPair>& prims) {
    pairs.clear();
    prims.clear();
    const double c0 = std::sqrt(2.0) * std::pow(POLY_PI, 1.25);
    int ns = (int)ib.shells.size();
    for (int a = 0; a < ns; ++a)
        for (int b = a; b < ns; ++b) {
            const Shell& A = ib.shells[a];
            const Shell& B = ib.shells[b];
            PairInfo pi;
            pi.a = a; pi.b = b;
            pi.cls = pair_class(A.L, B.L);
            pi.ABx = A.x - B.x; pi.ABy = A.y - B.y; pi.ABz = A.z - B.z;
            double r2 = pi.ABx * pi.ABx + pi.ABy * pi.ABy + pi.ABz * pi.ABz;
            pi.prim_off = (int)prims.size();
            for (size_t i = 0; i < A.alpha.size(); ++i)
                for (size_t j = 0; j < B.alpha.size(); ++j) {
                    double al = A.alpha[i], be = B.alpha[j], z = al + be;
                    double e = std::exp(-al * be / z * r2);
                    double cc = A.coef[i] * B.coef[j];
                    if (std::fabs(cc) * e < primtol) continue;
                    PrimPair pp;
                    pp.zeta = z;
                    pp.K = cc * e * c0 / z;
                    pp.Px = (al * A.x + be * B.x) / z;
                    pp.Py = (al * A.y + be * B.y) / z;
                    pp.Pz = (al * A.z + be * B.z) / z;
                    pp.PAx = pp.Px - A.x; pp.PAy = pp.Py - A.y; pp.PAz = pp.Pz - A.z;
                    prims.push_back(pp);
                }
            pi.nprim = (int)prims.size() - pi.prim_off;
            if (pi.nprim > 0) pairs.push_back(pi);
        }
}

int build_gc_pairs(const IntBasis& ib, double primtol, const std::vector<PairInfo>& pairs,
                   std::vector<PairInfo>& gpairs, std::vector<GcMeta>& gmeta, std::vector<GcPrim>& gprims,
                   std::vector<int>& member_pair) {
    gpairs.clear(); gmeta.clear(); gprims.clear(); member_pair.clear();
    const int ns = (int)ib.shells.size();
    // groups, with each member's coefficients over the first shell's exponents
    std::vector<std::vector<int>> groups;
    std::vector<std::vector<std::vector<double>>> gcoef;
    std::vector<int> gof(ns, -1);
    int ngc = 0;
    for (int a = 0; a < ns; ++a) {
        if (gof[a] >= 0) continue;
        const Shell& A = ib.shells[a];
        gof[a] = (int)groups.size();
        groups.push_back({a});
        gcoef.push_back({A.coef});
        if (A.L != 0) continue;
        bool distinct = true;
        for (size_t i = 0; i < A.alpha.size(); ++i)
            for (size_t j = 0; j < i; ++j) distinct = distinct && A.alpha[i] != A.alpha[j];
        if (!distinct) continue;
        for (int b = a + 1; b < ns; ++b) {
            const Shell& B = ib.shells[b];
            if (gof[b] >= 0 || B.L != 0 || B.x != A.x || B.y != A.y || B.z != A.z) continue;
            std::vector<double> cb(A.alpha.size(), 0.0);
            bool sub = true;
            for (size_t j = 0; j < B.alpha.size() && sub; ++j) {
                const auto it = std::find(A.alpha.begin(), A.alpha.end(), B.alpha[j]);
                if (it == A.alpha.end()) sub = false;
                else cb[it - A.alpha.begin()] += B.coef[j];
            }
            if (!sub) continue;
            gof[b] = gof[a];
            groups.back().push_back(b);
            gcoef.back().push_back(cb);
            ++ngc;
            break;
        }
    }
    if (ngc == 0) return 0;
    std::vector<int> pidx((size_t)ns * ns, -1);
    for (size_t i = 0; i < pairs.size(); ++i) {
        pidx[(size_t)pairs[i].a * ns + pairs[i].b] = (int)i;
        pidx[(size_t)pairs[i].b * ns + pairs[i].a] = (int)i;
    }
    const double c0 = std::sqrt(2.0) * std::pow(POLY_PI, 1.25);
    const int ng = (int)groups.size();
    for (int g1 = 0; g1 < ng; ++g1)
        for (int g2 = g1; g2 < ng; ++g2) {
            const std::vector<int>& G1 = groups[g1];
            const std::vector<int>& G2 = groups[g2];
            const Shell& A = ib.shells[G1[0]];
            const Shell& B = ib.shells[G2[0]];
            if (B.L != 0) continue;
            int mem[GC_MAXM][2], m = 0;
            for (size_t i = 0; i < G1.size(); ++i)
                for (size_t j = 0; j < G2.size(); ++j)
                    if (g1 != g2 || j >= i) { mem[m][0] = (int)i; mem[m][1] = (int)j; ++m; }
            PairInfo pi;
            pi.a = G1[0]; pi.b = G2[0];
            pi.cls = pair_class(A.L, 0);
            pi.ABx = A.x - B.x; pi.ABy = A.y - B.y; pi.ABz = A.z - B.z;
            const double r2 = pi.ABx * pi.ABx + pi.ABy * pi.ABy + pi.ABz * pi.ABz;
            pi.prim_off = (int)gprims.size();
            for (size_t i = 0; i < A.alpha.size(); ++i)
                for (size_t j = 0; j < B.alpha.size(); ++j) {
                    const double al = A.alpha[i], be = B.alpha[j], z = al + be;
                    const double e = std::exp(-al * be / z * r2);
                    GcPrim gp{};
                    double wmax = 0.0;
                    for (int k = 0; k < m; ++k) {
                        gp.w[k] = gcoef[g1][mem[k][0]][i] * gcoef[g2][mem[k][1]][j];
                        wmax = std::max(wmax, std::fabs(gp.w[k]));
                    }
                    if (wmax * e < primtol) continue;
                    gp.pp.zeta = z;
                    gp.pp.K = e * c0 / z;
                    gp.pp.Px = (al * A.x + be * B.x) / z;
                    gp.pp.Py = (al * A.y + be * B.y) / z;
                    gp.pp.Pz = (al * A.z + be * B.z) / z;
                    gp.pp.PAx = gp.pp.Px - A.x; gp.pp.PAy = gp.pp.Py - A.y; gp.pp.PAz = gp.pp.Pz - A.z;
                    gprims.push_back(gp);
                }
            pi.nprim = (int)gprims.size() - pi.prim_off;
            if (pi.nprim == 0) continue;
            GcMeta gm{};
            gm.prim_off = pi.prim_off; gm.nprim = pi.nprim; gm.m = m;
            for (int k = 0; k < m; ++k) gm.code |= (mem[k][0] << (2 * k)) | (mem[k][1] << (2 * k + 1));
            gm.fa0 = ib.shells[G1[0]].f0; gm.fa1 = G1.size() > 1 ? ib.shells[G1[1]].f0 : -1;
            gm.fb0 = ib.shells[G2[0]].f0; gm.fb1 = G2.size() > 1 ? ib.shells[G2[1]].f0 : -1;
            gm.sa0 = G1[0]; gm.sa1 = G1.size() > 1 ? G1[1] : -1;
            gm.sb0 = G2[0]; gm.sb1 = G2.size() > 1 ? G2[1] : -1;
            for (int k = 0; k < GC_MAXM; ++k)
                member_pair.push_back(k < m ? pidx[(size_t)G1[mem[k][0]] * ns + G2[mem[k][1]]] : -1);
            gpairs.push_back(pi);
            gmeta.push_back(gm);
        }
    return ngc;
}

// Periodic pairs: for every cell offset Q in [-rc1, rc1] and every ordered
// shell pair (s1, s2), s1 in cell 0 and s2 in cell Q (translated by Q*period
// along x).  Stored with the higher angular momentum first, so the pair is
// (s1_0, s2_Q) or, with swap = 1, (s2_Q, s1_0).
void build_pairs_cells(const IntBasis& ib, double primtol, double period, int rc1, std::vector<PairInfo>& pairs,
                       std::vector<PrimPair>& prims, std::vector<int2_host>& cells) {
    pairs.clear();
    prims.clear();
    cells.clear();
    const double c0 = std::sqrt(2.0) * std::pow(POLY_PI, 1.25);
    const int ns = (int)ib.shells.size();
    for (int q = -rc1; q <= rc1; ++q)
        for (int s1 = 0; s1 < ns; ++s1)
            for (int s2 = 0; s2 < ns; ++s2) {
                const bool swap = ib.shells[s1].L < ib.shells[s2].L;
                const int a = swap ? s2 : s1, b = swap ? s1 : s2;
                const Shell& A = ib.shells[a];
                const Shell& B = ib.shells[b];
                const double axs = A.x + (swap ? q * period : 0.0), bxs = B.x + (swap ? 0.0 : q * period);
                PairInfo pi;
                pi.a = a; pi.b = b;
                pi.cls = pair_class(A.L, B.L);
                pi.ABx = axs - bxs; pi.ABy = A.y - B.y; pi.ABz = A.z - B.z;
                const double r2 = pi.ABx * pi.ABx + pi.ABy * pi.ABy + pi.ABz * pi.ABz;
                pi.prim_off = (int)prims.size();
                for (size_t i = 0; i < A.alpha.size(); ++i)
                    for (size_t j = 0; j < B.alpha.size(); ++j) {
                        const double al = A.alpha[i], be = B.alpha[j], z = al + be;
                        const double e = std::exp(-al * be / z * r2);
                        const double cc = A.coef[i] * B.coef[j];
                        if (std::fabs(cc) * e < primtol) continue;
                        PrimPair pp;
                        pp.zeta = z;
                        pp.K = cc * e * c0 / z;
                        pp.Px = (al * axs + be * bxs) / z;
                        pp.Py = (al * A.y + be * B.y) / z;
                        pp.Pz = (al * A.z + be * B.z) / z;
                        pp.PAx = pp.Px - axs; pp.PAy = pp.Py - A.y; pp.PAz = pp.Pz - A.z;
                        prims.push_back(pp);
                    }
                pi.nprim = (int)prims.size() - pi.prim_off;
                if (pi.nprim > 0) {
                    pairs.push_back(pi);
                    cells.push_back({q, swap ? 1 : 0});
                }
            }
}

void build_work_lists(const std::vector<PairInfo>& pairs, const std::vector<double>& Q, double tol,
                      double cut, const std::vector<int> class_ids[], std::vector<int> ket_ids[],
                      std::vector<int> bstart[], std::vector<WorkList>& out, bool nosym,
                      uint64_t split_mask) {
    out.clear();
    std::vector<int> rank(pairs.size(), 0);
    for (int X = 0; X < NPAIRCLASS; ++X)
        for (size_t r = 0; r < class_ids[X].size(); ++r) rank[class_ids[X][r]] = (int)r;
    for (int Y = 0; Y < NPAIRCLASS; ++Y) {
        auto& kl = ket_ids[Y];
        kl = class_ids[Y];
        std::stable_sort(kl.begin(), kl.end(), [&](int a, int b) {
            if (pairs[a].nprim != pairs[b].nprim) return pairs[a].nprim > pairs[b].nprim;
            return rank[a] < rank[b];
        });
        auto& bs = bstart[Y];
        bs.clear();
        for (size_t j = 0; j < kl.size(); ++j)
            if (j == 0 || pairs[kl[j]].nprim != pairs[kl[j - 1]].nprim) bs.push_back((int)j);
        bs.push_back((int)kl.size());
    }
    for (int X = 0; X < NPAIRCLASS; ++X)
        for (int Y = 0; Y <= X; ++Y) {
            const bool split = cut > tol && ((split_mask >> (X * NPAIRCLASS + Y)) & 1);
            for (int prec = 0; prec <= (split ? 1 : 0); ++prec) {
                const auto& bx = class_ids[X];
                const auto& ky = ket_ids[Y];
                const auto& bs = bstart[Y];
                const int nb = (int)bs.size() - 1;
                WorkList w;
                w.X = X; w.Y = Y; w.prec = prec;
                w.nbucket = std::max(nb, 1);
                w.nrows = (int)bx.size();
                w.row.assign(bx.size() + 1, 0);
                w.rb.assign(bx.size() * (size_t)(w.nbucket + 1), 0);
                if (prec) w.rf.assign(bx.size() * (size_t)w.nbucket, 0);
                // prefix of bucket [lo, hi) with qi * Q >= thr
                auto prefix = [&](double qi, int lo, int hi, double thr) {
                    int a0 = lo, a1 = hi;
                    while (a0 < a1) {
                        const int mid = (a0 + a1) / 2;
                        if (qi * Q[ky[mid]] >= thr) a0 = mid + 1; else a1 = mid;
                    }
                    return a0 - lo;
                };
                for (size_t i = 0; i < bx.size(); ++i) {
                    const double qi = Q[bx[i]];
                    const int ri = rank[bx[i]];
                    int* r = w.rb.data() + i * (size_t)(w.nbucket + 1);
                    long long tot = 0;
                    double kp = 0;
                    for (int b = 0; b < nb; ++b) {
                        const int lo = bs[b], hi = bs[b + 1];
                        int lim = hi - lo;
                        if (X == Y && !nosym) {  // rank_j <= rank_i (a prefix: buckets are in rank order)
                            int b0 = lo, b1 = hi;
                            while (b0 < b1) {
                                const int mid = (b0 + b1) / 2;
                                if (rank[ky[mid]] <= ri) b0 = mid + 1; else b1 = mid;
                            }
                            lim = b0 - lo;
                        }
                        const int n_all = std::min(lim, prefix(qi, lo, hi, tol));
                        const int n_hi = split ? std::min(n_all, prefix(qi, lo, hi, cut)) : n_all;
                        const int first = prec ? n_hi : 0;
                        const int cnt = prec ? n_all - n_hi : n_hi;
                        if (prec) w.rf[i * (size_t)w.nbucket + b] = first;
                        r[b] = (int)tot;
                        tot += cnt;
                        if (cnt) kp += (double)cnt * pairs[ky[lo]].nprim;
                    }
                    for (int b = std::max(nb, 0); b <= w.nbucket; ++b) r[b] = (int)tot;
                    w.row[i + 1] = w.row[i] + tot;
                    w.prim_quartets += pairs[bx[i]].nprim * kp;
                }
                w.nwork = w.row[bx.size()];
                out.push_back(std::move(w));
            }
        }
}

// F_m(T) on the grid by the convergent series
// F_m(T) = e^-T sum_i (2T)^i / ((2m+1)(2m+3)...(2m+2i+1)), all terms > 0.
void boys_table(std::vector<double>& tab) {
    tab.assign((size_t)BOYS_NT * BOYS_NM, 0.0);
    for (int i = 0; i < BOYS_NT; ++i) {
        double T = i / 20.0;
        for (int m = 0; m < BOYS_NM; ++m) {
            double term = 1.0 / (2 * m + 1), sum = term;
            for (int k = 1; k < 2000; ++k) {
                term *= 2.0 * T / (2 * m + 2 * k + 1);
                sum += term;
                if (term < 1e-18 * sum) break;
            }
            tab[(size_t)i * BOYS_NM + m] = std::exp(-T) * sum;
        }
    }
}

static void boys_host(int L, double T, const std::vector<double>& tab, double* F) {
    if (T < BOYS_TMAX) {
        int ts = (int)(T * 20.0 + 0.5);
        double h = 0.05 * ts - T;
        const double* g = tab.data() + (size_t)ts * BOYS_NM + L;
        double f = g[6];
        for (int k = 5; k >= 0; --k) f = g[k] + h / (k + 1) * f;
        F[L] = f;
        double e = std::exp(-T);
        for (int m = L - 1; m >= 0; --m) F[m] = (2 * T * F[m + 1] + e) / (2 * m + 1);
    } else {
        F[0] = 0.5 * std::sqrt(POLY_PI / T);
        for (int m = 0; m < L; ++m) F[m + 1] = F[m] * (m + 0.5) / T;
    }
}

// Cartesian triple index for |t| <= 7
struct Tri {
    int idx[8][8][8];
    int n = 0;
    Tri() {
        for (int l = 0; l <= 7; ++l)
            for (int i = l; i >= 0; --i)
                for (int k = 0; k <= l - i; ++k) idx[i][l - i - k][k] = n++;
    }
};
static const Tri kTri;

// Two-centre HRR: (a|b) from e-table E over triples (|e| <= la+lb).
static double hrr2(const double* E, int ax, int ay, int az, int bx, int by, int bz,
                   double ABx, double ABy, double ABz) {
    if (bx + by + bz == 0) return E[kTri.idx[ax][ay][az]];
    if (bx > 0) return hrr2(E, ax + 1, ay, az, bx - 1, by, bz, ABx, ABy, ABz) + ABx * hrr2(E, ax, ay, az, bx - 1, by, bz, ABx, ABy, ABz);
    if (by > 0) return hrr2(E, ax, ay + 1, az, bx, by - 1, bz, ABx, ABy, ABz) + ABy * hrr2(E, ax, ay, az, bx, by - 1, bz, ABx, ABy, ABz);
    return hrr2(E, ax, ay, az + 1, bx, by, bz - 1, ABx, ABy, ABz) + ABz * hrr2(E, ax, ay, az, bx, by, bz - 1, ABx, ABy, ABz);
}

void one_electron(const IntBasis& ib, const std::vector<double>& boys, std::vector<double>& S,
                  std::vector<double>& Tm, std::vector<double>& V) {
    int n = ib.nbf, ns = (int)ib.shells.size();
    S.assign((size_t)n * n, 0.0); Tm.assign((size_t)n * n, 0.0); V.assign((size_t)n * n, 0.0);
    int natom = (int)ib.natom_z.size();
#pragma omp parallel for schedule(dynamic)
    for (int a = 0; a < ns; ++a) {
        std::vector<double> es(kTri.n), ev(kTri.n * 8);
        for (int b = a; b < ns; ++b) {
            const Shell& A = ib.shells[a];
            const Shell& B = ib.shells[b];
            int la = A.L, lb = B.L, na = ncart(la), nb = ncart(lb);
            double ABx = A.x - B.x, ABy = A.y - B.y, ABz = A.z - B.z;
            double r2 = ABx * ABx + ABy * ABy + ABz * ABz;
            std::vector<double> s(na * nb, 0.0), t(na * nb, 0.0), v(na * nb, 0.0);
            for (size_t i = 0; i < A.alpha.size(); ++i)
                for (size_t j = 0; j < B.alpha.size(); ++j) {
                    double al = A.alpha[i], be = B.alpha[j], z = al + be;
                    double cc = A.coef[i] * B.coef[j] * std::exp(-al * be / z * r2);
                    double P[3] = {(al * A.x + be * B.x) / z, (al * A.y + be * B.y) / z, (al * A.z + be * B.z) / z};
                    double PA[3] = {P[0] - A.x, P[1] - A.y, P[2] - A.z};
                    int Le = la + lb + 2;
                    // overlap e-ladder: S(e+1_i) = PA_i S(e) + e_i/(2z) S(e-1_i)
                    es[0] = std::pow(POLY_PI / z, 1.5);
                    for (int l = 1; l <= Le; ++l)
                        for (int ix = l; ix >= 0; --ix)
                            for (int iz = 0; iz <= l - ix; ++iz) {
                                int iy = l - ix - iz;
                                int e[3] = {ix, iy, iz};
                                int d = ix > 0 ? 0 : (iy > 0 ? 1 : 2);
                                int e1[3] = {ix, iy, iz}; e1[d]--;
                                double val = PA[d] * es[kTri.idx[e1[0]][e1[1]][e1[2]]];
                                if (e1[d] > 0) {
                                    int e2[3] = {e1[0], e1[1], e1[2]}; e2[d]--;
                                    val += e1[d] / (2 * z) * es[kTri.idx[e2[0]][e2[1]][e2[2]]];
                                }
                                es[kTri.idx[e[0]][e[1]][e[2]]] = val;
                            }
                    // nuclear attraction ladder with m
                    int Lv = la + lb;
                    std::vector<double> vv(kTri.n, 0.0);
                    for (int k = 0; k < natom; ++k) {
                        double C[3] = {ib.atom_xyz[3 * k], ib.atom_xyz[3 * k + 1], ib.atom_xyz[3 * k + 2]};
                        double PC[3] = {P[0] - C[0], P[1] - C[1], P[2] - C[2]};
                        double TT = z * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
                        double F[16];
                        boys_host(Lv, TT, boys, F);
                        double pre = -ib.natom_z[k] * 2.0 * POLY_PI / z;
                        for (int m = 0; m <= Lv; ++m) ev[m * kTri.n] = pre * F[m];
                        for (int l = 1; l <= Lv; ++l)
                            for (int ix = l; ix >= 0; --ix)
                                for (int iz = 0; iz <= l - ix; ++iz) {
                                    int iy = l - ix - iz;
                                    int d = ix > 0 ? 0 : (iy > 0 ? 1 : 2);
                                    int e1[3] = {ix, iy, iz}; e1[d]--;
                                    int i1 = kTri.idx[e1[0]][e1[1]][e1[2]];
                                    int i2 = -1;
                                    if (e1[d] > 0) { int e2[3] = {e1[0], e1[1], e1[2]}; e2[d]--; i2 = kTri.idx[e2[0]][e2[1]][e2[2]]; }
                                    int i0 = kTri.idx[ix][iy][iz];
                                    for (int m = 0; m <= Lv - l; ++m) {
                                        double val = PA[d] * ev[m * kTri.n + i1] - PC[d] * ev[(m + 1) * kTri.n + i1];
                                        if (i2 >= 0) val += e1[d] / (2 * z) * (ev[m * kTri.n + i2] - ev[(m + 1) * kTri.n + i2]);
                                        ev[m * kTri.n + i0] = val;
                                    }
                                }
                        for (int q = 0; q < kTri.n; ++q) vv[q] += ev[q];
                    }
                    for (int p = 0; p < na; ++p) {
                        int ax, ay, az;
                        cart_comp(la, p, ax, ay, az);
                        for (int q = 0; q < nb; ++q) {
                            int bx, by, bz;
                            cart_comp(lb, q, bx, by, bz);
                            double sv = hrr2(es.data(), ax, ay, az, bx, by, bz, ABx, ABy, ABz);
                            // kinetic: -1/2 sum_d [b_d(b_d-1)<a|b-2d> - 2be(2b_d+1)<a|b> + 4be^2<a|b+2d>]
                            int bb[3] = {bx, by, bz};
                            double kin = 0.0;
                            for (int d = 0; d < 3; ++d) {
                                int up[3] = {bx, by, bz}; up[d] += 2;
                                double k3 = 4 * be * be * hrr2(es.data(), ax, ay, az, up[0], up[1], up[2], ABx, ABy, ABz);
                                double k2 = -2 * be * (2 * bb[d] + 1) * sv;
                                double k1 = 0.0;
                                if (bb[d] >= 2) {
                                    int dn[3] = {bx, by, bz}; dn[d] -= 2;
                                    k1 = bb[d] * (bb[d] - 1) * hrr2(es.data(), ax, ay, az, dn[0], dn[1], dn[2], ABx, ABy, ABz);
                                }
                                kin += -0.5 * (k1 + k2 + k3);
                            }
                            double vx = hrr2(vv.data(), ax, ay, az, bx, by, bz, ABx, ABy, ABz);
                            s[p * nb + q] += cc * sv;
                            t[p * nb + q] += cc * kin;
                            v[p * nb + q] += cc * vx;
                        }
                    }
                }
            for (int p = 0; p < na; ++p)
                for (int q = 0; q < nb; ++q) {
                    size_t i1 = (size_t)(A.f0 + p) * n + (B.f0 + q), i2 = (size_t)(B.f0 + q) * n + (A.f0 + p);
                    S[i1] = S[i2] = s[p * nb + q];
                    Tm[i1] = Tm[i2] = t[p * nb + q];
                    V[i1] = V[i2] = v[p * nb + q];
                }
        }
    }
}

// One-electron and nuclear parts of the nuclear gradient at fixed D and W
// (library basis, nbf x nbf row-major): sum D (dT + dV) - W dS over the
// basis-function derivatives on both centres (OVERLAP_KINETIC(1),
// NUCLEAR_ATTRACTION(1), src/ponee.f90:157-214, 418-), the derivative of
// the attraction operator at each nucleus (HELLMANN_FEYNMAN_ATTRACTION,
// src/phellmannfeynman.f90:86-), taken through translational invariance of
// each nucleus' term, -(d/dA + d/dB), and the nuclear repulsion
// (HELLMANN_FEYNMAN_REPULSION, src/phellmannfeynman.f90:1-85).  The
// derivative of a primitive Cartesian Gaussian on its centre is
// 2 alpha [a + 1] - a_x [a - 1].  shell_atom: the atom of each shell.
void one_electron_gradient(const IntBasis& ib, const std::vector<int>& shell_atom, const std::vector<double>& boys,
                           const double* D, const double* W, double* grad) {
    const int n = ib.nbf, ns = (int)ib.shells.size(), natom = (int)ib.natom_z.size();
    for (int i = 0; i < 3 * natom; ++i) grad[i] = 0.0;
    std::vector<std::vector<double>> part(ns, std::vector<double>());
#pragma omp parallel for schedule(dynamic)
    for (int a = 0; a < ns; ++a) {
        std::vector<double> g(3 * natom, 0.0);
        std::vector<double> es(kTri.n), ev((size_t)kTri.n * 12), vk((size_t)natom * kTri.n);
        const Shell& A = ib.shells[a];
        for (int b = a; b < ns; ++b) {
            const Shell& B = ib.shells[b];
            const int la = A.L, lb = B.L, na = ncart(la), nb = ncart(lb);
            const double ABx = A.x - B.x, ABy = A.y - B.y, ABz = A.z - B.z;
            const double r2 = ABx * ABx + ABy * ABy + ABz * ABz;
            const double fac = a == b ? 1.0 : 2.0;
            const int atA = shell_atom[a], atB = shell_atom[b];
            for (size_t i = 0; i < A.alpha.size(); ++i)
                for (size_t j = 0; j < B.alpha.size(); ++j) {
                    const double al = A.alpha[i], be = B.alpha[j], z = al + be;
                    const double cc = A.coef[i] * B.coef[j] * std::exp(-al * be / z * r2);
                    if (std::fabs(cc) < 1e-17) continue;   // overlap factor of a distant pair
                    const double P[3] = {(al * A.x + be * B.x) / z, (al * A.y + be * B.y) / z,
                                         (al * A.z + be * B.z) / z};
                    const double PA[3] = {P[0] - A.x, P[1] - A.y, P[2] - A.z};
                    const int Le = la + lb + 3;          // +1 derivative, +2 kinetic
                    es[0] = std::pow(POLY_PI / z, 1.5);
                    for (int l = 1; l <= Le; ++l)
                        for (int ix = l; ix >= 0; --ix)
                            for (int iz = 0; iz <= l - ix; ++iz) {
                                const int iy = l - ix - iz;
                                const int d = ix > 0 ? 0 : (iy > 0 ? 1 : 2);
                                int e1[3] = {ix, iy, iz};
                                e1[d]--;
                                double val = PA[d] * es[kTri.idx[e1[0]][e1[1]][e1[2]]];
                                if (e1[d] > 0) {
                                    int e2[3] = {e1[0], e1[1], e1[2]};
                                    e2[d]--;
                                    val += e1[d] / (2 * z) * es[kTri.idx[e2[0]][e2[1]][e2[2]]];
                                }
                                es[kTri.idx[ix][iy][iz]] = val;
                            }
                    // attraction ladders of every nucleus, one level above the pair
                    const int Lv = la + lb + 1;
                    for (int k = 0; k < natom; ++k) {
                        const double PC[3] = {P[0] - ib.atom_xyz[3 * k], P[1] - ib.atom_xyz[3 * k + 1],
                                              P[2] - ib.atom_xyz[3 * k + 2]};
                        const double TT = z * (PC[0] * PC[0] + PC[1] * PC[1] + PC[2] * PC[2]);
                        double F[16];
                        boys_host(Lv, TT, boys, F);
                        const double pre = -ib.natom_z[k] * 2.0 * POLY_PI / z;
                        for (int m = 0; m <= Lv; ++m) ev[(size_t)m * kTri.n] = pre * F[m];
                        for (int l = 1; l <= Lv; ++l)
                            for (int ix = l; ix >= 0; --ix)
                                for (int iz = 0; iz <= l - ix; ++iz) {
                                    const int iy = l - ix - iz;
                                    const int d = ix > 0 ? 0 : (iy > 0 ? 1 : 2);
                                    int e1[3] = {ix, iy, iz};
                                    e1[d]--;
                                    const int i1 = kTri.idx[e1[0]][e1[1]][e1[2]];
                                    int i2 = -1;
                                    if (e1[d] > 0) {
                                        int e2[3] = {e1[0], e1[1], e1[2]};
                                        e2[d]--;
                                        i2 = kTri.idx[e2[0]][e2[1]][e2[2]];
                                    }
                                    const int i0 = kTri.idx[ix][iy][iz];
                                    for (int m = 0; m <= Lv - l; ++m) {
                                        double val = PA[d] * ev[(size_t)m * kTri.n + i1] - PC[d] * ev[(size_t)(m + 1) * kTri.n + i1];
                                        if (i2 >= 0)
                                            val += e1[d] / (2 * z) * (ev[(size_t)m * kTri.n + i2] - ev[(size_t)(m + 1) * kTri.n + i2]);
                                        ev[(size_t)m * kTri.n + i0] = val;
                                    }
                                }
                        for (int q = 0; q < kTri.n; ++q) vk[(size_t)k * kTri.n + q] = ev[q];
                    }
                    auto S = [&](const int* x, const int* y) {
                        return hrr2(es.data(), x[0], x[1], x[2], y[0], y[1], y[2], ABx, ABy, ABz);
                    };
                    auto T = [&](const int* x, const int* y) {   // kinetic operator on the second function
                        double t = 0.0;
                        const double sv = S(x, y);
                        for (int d = 0; d < 3; ++d) {
                            int up[3] = {y[0], y[1], y[2]};
                            up[d] += 2;
                            double v = 4 * be * be * S(x, up) - 2 * be * (2 * y[d] + 1) * sv;
                            if (y[d] >= 2) {
                                int dn[3] = {y[0], y[1], y[2]};
                                dn[d] -= 2;
                                v += y[d] * (y[d] - 1) * S(x, dn);
                            }
                            t += -0.5 * v;
                        }
                        return t;
                    };
                    auto V = [&](int k, const int* x, const int* y) {
                        return hrr2(vk.data() + (size_t)k * kTri.n, x[0], x[1], x[2], y[0], y[1], y[2], ABx, ABy, ABz);
                    };
                    for (int p = 0; p < na; ++p) {
                        int x[3];
                        cart_comp(la, p, x[0], x[1], x[2]);
                        for (int q = 0; q < nb; ++q) {
                            int y[3];
                            cart_comp(lb, q, y[0], y[1], y[2]);
                            const size_t ipq = (size_t)(A.f0 + p) * n + B.f0 + q;
                            const double dpq = fac * cc * D[ipq], wpq = fac * cc * W[ipq];
                            if (dpq == 0.0 && wpq == 0.0) continue;
                            for (int e = 0; e < 3; ++e) {
                                // derivatives of the bra function (centre A) and the ket function (centre B)
                                int xu[3] = {x[0], x[1], x[2]}, xd[3] = {x[0], x[1], x[2]};
                                int yu[3] = {y[0], y[1], y[2]}, yd[3] = {y[0], y[1], y[2]};
                                xu[e]++; xd[e]--; yu[e]++; yd[e]--;
                                const double sA = 2 * al * S(xu, y) - (x[e] ? x[e] * S(xd, y) : 0.0);
                                const double sB = 2 * be * S(x, yu) - (y[e] ? y[e] * S(x, yd) : 0.0);
                                const double tA = 2 * al * T(xu, y) - (x[e] ? x[e] * T(xd, y) : 0.0);
                                const double tB = 2 * be * T(x, yu) - (y[e] ? y[e] * T(x, yd) : 0.0);
                                double vA = 0.0, vB = 0.0;
                                for (int k = 0; k < natom; ++k) {
                                    const double ak = 2 * al * V(k, xu, y) - (x[e] ? x[e] * V(k, xd, y) : 0.0);
                                    const double bk = 2 * be * V(k, x, yu) - (y[e] ? y[e] * V(k, x, yd) : 0.0);
                                    vA += ak;
                                    vB += bk;
                                    g[3 * k + e] -= dpq * (ak + bk);          // the operator at nucleus k
                                }
                                g[3 * atA + e] += dpq * (tA + vA) - wpq * sA;
                                g[3 * atB + e] += dpq * (tB + vB) - wpq * sB;
                            }
                        }
                    }
                }
        }
        part[a] = std::move(g);
    }
    for (int a = 0; a < ns; ++a)
        for (int i = 0; i < 3 * natom; ++i) grad[i] += part[a][i];
    // nuclear repulsion
    for (int i = 0; i < natom; ++i)
        for (int j = 0; j < natom; ++j) {
            if (i == j) continue;
            const double d[3] = {ib.atom_xyz[3 * i] - ib.atom_xyz[3 * j], ib.atom_xyz[3 * i + 1] - ib.atom_xyz[3 * j + 1],
                                 ib.atom_xyz[3 * i + 2] - ib.atom_xyz[3 * j + 2]};
            const double r = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
            for (int e = 0; e < 3; ++e) grad[3 * i + e] -= ib.natom_z[i] * ib.natom_z[j] * d[e] / (r * r * r);
        }
}

double nuclear_repulsion(const IntBasis& ib) {
    double e = 0.0;
    int n = (int)ib.natom_z.size();
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j) {
            double dx = ib.atom_xyz[3 * i] - ib.atom_xyz[3 * j], dy = ib.atom_xyz[3 * i + 1] - ib.atom_xyz[3 * j + 1],
                   dz = ib.atom_xyz[3 * i + 2] - ib.atom_xyz[3 * j + 2];
            e += ib.natom_z[i] * ib.natom_z[j] / std::sqrt(dx * dx + dy * dy + dz * dz);
        }
    return e;
}

}  // namespace polyeri
