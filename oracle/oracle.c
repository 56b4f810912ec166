/*
 * oracle.c -- CPU restatement of polymer's direct two-electron path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the MI355X product.  It is
 * loaded by tests/, by smoke() and by bench.py's cpu_baseline leg, never by
 * the product.  Each routine names the reference routine (file:line) whose
 * algorithm it restates; the code is this project's own.
 *
 * Conventions: 0-based indices, column-major matrices (a[i + n*j]).
 */
#include "oracle.h"

#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* src/pmodules.f90:7 -- the reference's (truncated) value of pi. */
#define PI_REF 3.14159265358979
/* src/pmodules.f90:12 */
#define BOLTZMANN (1.0 / 3.1577464e5)

static const char *CATOM[36] = {
    "H ", "He", "Li", "Be", "B ", "C ", "N ", "O ", "F ", "Ne", "Na", "Mg",
    "Al", "Si", "P ", "S ", "Cl", "Ar", "K ", "Ca", "Sc", "Ti", "V ", "Cr",
    "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Ga", "Ge", "As", "Se", "Br", "Kr"};

/* ------------------------------------------------------------------ */
/* Boys function: src/pbasis.f90:1-27 and src/pmodules.f90:153-161    */
/* ------------------------------------------------------------------ */
#define NGRID 1701
#define NMROW 28
static double g_igamma[NGRID * NMROW]; /* IGAMMA(i,j) = g_igamma[i + NGRID*j] */
static double g_f1[NMROW], g_f2[NMROW];
static int g_boys_ready = 0;
/* TF(0:21), src/pmodules.f90:156-159: switch points to the asymptotic form */
static const double TF[22] = {33, 37, 41, 43, 46, 49, 51, 54, 56, 58, 61,
                              63, 66, 68, 70, 72, 74, 76, 78, 80, 82, 84};

static void boys_init(void) {
    if (g_boys_ready) return;
    for (int i = 0; i < NGRID; ++i) {
        double a = i / 20.0, b = exp(-a), c = 0.0;
        int k = (int)(50 + i / 3.0);
        for (int j = k; j >= 0; --j) {
            c = (2.0 * a * c + b) / (double)(2 * j + 1);
            if (j < NMROW) g_igamma[i + NGRID * j] = c;
        }
    }
    /* asymptotic coefficients, src/pfourindexdirect.f90:74-81 */
    g_f1[0] = 1.0 / (2.0 / sqrt(PI_REF));
    g_f2[0] = -0.5;
    for (int i = 1; i < NMROW; ++i) {
        g_f1[i] = g_f1[i - 1] * (double)(2 * i - 1) / 2.0;
        g_f2[i] = g_f2[i - 1] - 1.0;
    }
    g_boys_ready = 1;
}

void or_boys_table(double *igamma) {
    boys_init();
    memcpy(igamma, g_igamma, sizeof(g_igamma));
}

/* src/pfourindexdirect.f90:543-553 -- 7-term Taylor expansion about the
 * nearest 0.05 grid point, or the large-T asymptotic form. */
double or_boys(double t, int m) {
    boys_init();
    if (t < TF[m]) {
        int ts = (int)floor(t * 20.0 + 0.5);
        double h = 0.05 * ts - t;
        const double *g = g_igamma + ts;
#define IG(k) g[NGRID * (m + (k))]
        return ((((((IG(6) * h * 0.166666666666667 + IG(5)) * h * 0.2 + IG(4)) *
                       h * 0.25 + IG(3)) * h * 0.333333333333333 + IG(2)) *
                     h * 0.5 + IG(1)) * h + IG(0));
#undef IG
    }
    return g_f1[m] * exp(g_f2[m] * log(t));
}

/* ------------------------------------------------------------------ */
/* Basis reader: src/pread.f90:584-1482 (READ3)                       */
/* ------------------------------------------------------------------ */
#define MAXP 8192
#define MAXC 8192

typedef struct {
    int npgs, ncgs, npshell;
    int pax[MAXP], pay[MAXP], paz[MAXP], atom[MAXP];
    double x[MAXP], y[MAXP], z[MAXP], zt[MAXP];
    double cx[MAXC], cy[MAXC], cz[MAXC];
    /* sparse CC: for each primitive at most a few contracted functions    */
    int nc_of_p[MAXP];
    int c_of_p[MAXP][4];
    double v_of_p[MAXP][4];
    int *pshell; /* MAXP x 125 */
    int pang[MAXP];
} rbuild;

static void rb_set_cc(rbuild *r, int c, int p, double v) {
    int n = r->nc_of_p[p];
    for (int k = 0; k < n; ++k)
        if (r->c_of_p[p][k] == c) { r->v_of_p[p][k] = v; return; }
    if (n < 4) { r->c_of_p[p][n] = c; r->v_of_p[p][n] = v; r->nc_of_p[p] = n + 1; }
}

static int rb_new_cgs(rbuild *r, double x, double y, double z) {
    int c = r->ncgs++;
    r->cx[c] = x; r->cy[c] = y; r->cz[c] = z;
    return c;
}

static int rb_new_pgs(rbuild *r, int ax, int ay, int az, double x, double y,
                      double z, int atom, double zt) {
    int p = r->npgs++;
    r->pax[p] = ax; r->pay[p] = ay; r->paz[p] = az;
    r->x[p] = x; r->y[p] = y; r->z[p] = z;
    r->atom[p] = atom; r->zt[p] = zt; r->nc_of_p[p] = 0;
    return p;
}

static double fdouble(const char *s) {
    char buf[64];
    size_t n = strlen(s);
    if (n > 63) n = 63;
    for (size_t i = 0; i < n; ++i) buf[i] = (s[i] == 'D' || s[i] == 'd') ? 'E' : s[i];
    buf[n] = 0;
    return strtod(buf, NULL);
}

#define PSH(r, s, i, j, k) ((r)->pshell[(s) * 125 + ((i) * 25 + (j) * 5 + (k))])

/* Cartesian component orders and normalisation divisors of the reader
 * (src/pread.f90:692-1021): p x,y,z; d xx,yy,zz,xy,yz,xz;
 * f xxx,yyy,zzz,xxy,yyz,xzz,xyy,yzz,xxz,xyz. */
static const int DCOMP[6][3] = {{2,0,0},{0,2,0},{0,0,2},{1,1,0},{0,1,1},{1,0,1}};
static const double DDIV[6] = {1.7320508075688772, 1.7320508075688772,
                               1.7320508075688772, 1, 1, 1};
static const int FCOMP[10][3] = {{3,0,0},{0,3,0},{0,0,3},{2,1,0},{0,2,1},
                                 {1,0,2},{1,2,0},{0,1,2},{2,0,1},{1,1,1}};
static const double FDIV[10] = {3.872983346207417, 3.872983346207417,
                                3.872983346207417, 1.7320508075688772,
                                1.7320508075688772, 1.7320508075688772,
                                1.7320508075688772, 1.7320508075688772,
                                1.7320508075688772, 1};

int or_basis_build(const char *path, int natom, const int *iatom,
                   const double *xyz, or_basis **out) {
    FILE *f = fopen(path, "r");
    if (!f) return -1;
    rbuild *r = (rbuild *)calloc(1, sizeof(rbuild));
    r->pshell = (int *)malloc(sizeof(int) * MAXP * 125);
    char line[512];
    for (int ia = 0; ia < natom; ++ia) {
        const char *sym = CATOM[iatom[ia] - 1];
        double ax = xyz[3 * ia], ay = xyz[3 * ia + 1], az = xyz[3 * ia + 2];
        rewind(f);
        int latom = 0, found = 0;
        while (fgets(line, sizeof line, f)) {
            char c2[3] = {line[0], (line[0] && line[1] && line[1] != '\n') ? line[1] : ' ', 0};
            if (c2[1] == '\r') c2[1] = ' ';
            if (latom && c2[0] == sym[0] && c2[1] == sym[1]) {
                found = 1;
                /* shell blocks until the next '**' */
                while (fgets(line, sizeof line, f)) {
                    char typ[8]; int nprim; char ssc[64];
                    if (line[0] == '*' && line[1] == '*') break;
                    if (sscanf(line, "%7s %d %63s", typ, &nprim, ssc) != 3) break;
                    double s = fdouble(ssc);
                    int base = r->ncgs;
                    int ncomp = 0;
                    if (!strcmp(typ, "S")) ncomp = 1;
                    else if (!strcmp(typ, "P")) ncomp = 3;
                    else if (!strcmp(typ, "SP")) ncomp = 4;
                    else if (!strcmp(typ, "D")) ncomp = 6;
                    else if (!strcmp(typ, "F")) ncomp = 10;
                    else { fclose(f); free(r->pshell); free(r); return -2; }
                    for (int k = 0; k < ncomp; ++k) rb_new_cgs(r, ax, ay, az);
                    for (int k = 0; k < nprim; ++k) {
                        char sa[64], sb[64], sd[64];
                        if (!fgets(line, sizeof line, f)) break;
                        int nread = sscanf(line, "%63s %63s %63s", sa, sb, sd);
                        double a = fdouble(sa), bcoef = fdouble(sb);
                        double dcoef = (nread >= 3) ? fdouble(sd) : 0.0;
                        double nrm = exp(0.75 * log(2.0 * a / PI_REF));
                        double zt = a * s;
                        if (ncomp == 1) {
                            int p = rb_new_pgs(r, 0, 0, 0, ax, ay, az, ia, zt);
                            rb_set_cc(r, base, p, bcoef * nrm);
                        } else if (ncomp == 3 || ncomp == 4) {
                            int off = 0;
                            double pc = bcoef;
                            if (ncomp == 4) {
                                int p = rb_new_pgs(r, 0, 0, 0, ax, ay, az, ia, zt);
                                rb_set_cc(r, base, p, bcoef * nrm);
                                off = 1;
                                pc = dcoef;
                            }
                            for (int d = 0; d < 3; ++d) {
                                int p = rb_new_pgs(r, d == 0, d == 1, d == 2, ax, ay, az, ia, zt);
                                rb_set_cc(r, base + off + d, p, pc * nrm * sqrt(4.0 * a));
                            }
                        } else if (ncomp == 6) {
                            for (int d = 0; d < 6; ++d) {
                                int p = rb_new_pgs(r, DCOMP[d][0], DCOMP[d][1], DCOMP[d][2],
                                                   ax, ay, az, ia, zt);
                                double q = sqrt(4.0 * a);
                                rb_set_cc(r, base + d, p, bcoef * nrm * q * q / DDIV[d]);
                            }
                        } else {
                            for (int d = 0; d < 10; ++d) {
                                int p = rb_new_pgs(r, FCOMP[d][0], FCOMP[d][1], FCOMP[d][2],
                                                   ax, ay, az, ia, zt);
                                double q = sqrt(4.0 * a);
                                rb_set_cc(r, base + d, p, bcoef * nrm * q * q * q / FDIV[d]);
                            }
                        }
                    }
                }
                break;
            } else if (c2[0] == '*' && c2[1] == '*') {
                latom = 1;
            } else {
                latom = 0;
            }
        }
        if (!found) { fclose(f); free(r->pshell); free(r); return -3; }
    }
    fclose(f);

    /* shell structure, src/pread.f90:1076-1421 */
    for (int i = 0; i < MAXP * 125; ++i) r->pshell[i] = -1;
    int ns = 0;
    for (int i = 0; i < r->npgs; ++i) {
        if (r->pax[i] + r->pay[i] + r->paz[i] == 0) {
            PSH(r, ns, 0, 0, 0) = i;
            r->pang[ns] = 0;
            ++ns;
        }
    }
    int npgs0 = r->npgs; /* the loop bound is fixed at entry (Fortran DO) */
    for (int i = 0; i < npgs0; ++i) {
        int l = r->pax[i] + r->pay[i] + r->paz[i];
        if (l == 0) continue;
        int flag = 0;
        for (int k = 0; k < ns; ++k) {
            int a0 = PSH(r, k, 0, 0, 0);
            if (r->x[a0] == r->x[i] && r->y[a0] == r->y[i] && r->z[a0] == r->z[i] &&
                r->zt[a0] == r->zt[i] && PSH(r, k, r->pax[i], r->pay[i], r->paz[i]) == -1) {
                flag = 1;
                PSH(r, k, r->pax[i], r->pay[i], r->paz[i]) = i;
                r->pang[k] = l;
            }
        }
        if (!flag) {
            /* ghost s (and lower-L) primitives with zero contraction */
            int s = ns++;
            int c = rb_new_cgs(r, r->x[i], r->y[i], r->z[i]);
            int p = rb_new_pgs(r, 0, 0, 0, r->x[i], r->y[i], r->z[i], r->atom[i], r->zt[i]);
            (void)c;
            PSH(r, s, 0, 0, 0) = p;
            if (l >= 2) {
                for (int d = 0; d < 3; ++d) {
                    rb_new_cgs(r, r->x[i], r->y[i], r->z[i]);
                    int q = rb_new_pgs(r, d == 0, d == 1, d == 2, r->x[i], r->y[i], r->z[i],
                                       r->atom[i], r->zt[i]);
                    PSH(r, s, d == 0, d == 1, d == 2) = q;
                }
            }
            if (l >= 3) {
                for (int d = 0; d < 6; ++d) {
                    rb_new_cgs(r, r->x[i], r->y[i], r->z[i]);
                    int q = rb_new_pgs(r, DCOMP[d][0], DCOMP[d][1], DCOMP[d][2], r->x[i],
                                       r->y[i], r->z[i], r->atom[i], r->zt[i]);
                    PSH(r, s, DCOMP[d][0], DCOMP[d][1], DCOMP[d][2]) = q;
                }
            }
            PSH(r, s, r->pax[i], r->pay[i], r->paz[i]) = i;
            r->pang[s] = l;
        }
    }
    r->npshell = ns;

    or_basis *b = (or_basis *)calloc(1, sizeof(or_basis));
    b->natom = natom;
    b->iatom = (int *)malloc(sizeof(int) * natom);
    b->atomx = (double *)malloc(sizeof(double) * natom);
    b->atomy = (double *)malloc(sizeof(double) * natom);
    b->atomz = (double *)malloc(sizeof(double) * natom);
    for (int i = 0; i < natom; ++i) {
        b->iatom[i] = iatom[i];
        b->atomx[i] = xyz[3 * i]; b->atomy[i] = xyz[3 * i + 1]; b->atomz[i] = xyz[3 * i + 2];
    }
    int np = r->npgs, nc = r->ncgs;
    b->npgs = np; b->ncgs = nc; b->npshell = ns;
    b->pax = (int *)malloc(sizeof(int) * np); b->pay = (int *)malloc(sizeof(int) * np);
    b->paz = (int *)malloc(sizeof(int) * np); b->pgs_atom = (int *)malloc(sizeof(int) * np);
    b->pgsx = (double *)malloc(sizeof(double) * np); b->pgsy = (double *)malloc(sizeof(double) * np);
    b->pgsz = (double *)malloc(sizeof(double) * np); b->zt = (double *)malloc(sizeof(double) * np);
    b->cc = (double *)calloc((size_t)nc * np, sizeof(double));
    b->cgsx = (double *)malloc(sizeof(double) * nc); b->cgsy = (double *)malloc(sizeof(double) * nc);
    b->cgsz = (double *)malloc(sizeof(double) * nc);
    b->p_shell = (int *)malloc(sizeof(int) * ns * 125);
    b->p_shell_ang = (int *)malloc(sizeof(int) * ns);
    for (int p = 0; p < np; ++p) {
        b->pax[p] = r->pax[p]; b->pay[p] = r->pay[p]; b->paz[p] = r->paz[p];
        b->pgs_atom[p] = r->atom[p];
        b->pgsx[p] = r->x[p]; b->pgsy[p] = r->y[p]; b->pgsz[p] = r->z[p]; b->zt[p] = r->zt[p];
        for (int k = 0; k < r->nc_of_p[p]; ++k)
            b->cc[r->c_of_p[p][k] + (size_t)nc * p] = r->v_of_p[p][k];
    }
    for (int c = 0; c < nc; ++c) { b->cgsx[c] = r->cx[c]; b->cgsy[c] = r->cy[c]; b->cgsz[c] = r->cz[c]; }
    memcpy(b->p_shell, r->pshell, sizeof(int) * ns * 125);
    memcpy(b->p_shell_ang, r->pang, sizeof(int) * ns);
    free(r->pshell);
    free(r);
    *out = b;
    return 0;
}

void or_basis_free(or_basis *b) {
    if (!b) return;
    free(b->iatom); free(b->atomx); free(b->atomy); free(b->atomz);
    free(b->pax); free(b->pay); free(b->paz); free(b->pgs_atom);
    free(b->pgsx); free(b->pgsy); free(b->pgsz); free(b->zt); free(b->cc);
    free(b->cgsx); free(b->cgsy); free(b->cgsz); free(b->p_shell); free(b->p_shell_ang);
    free(b);
}

void or_basis_dims(const or_basis *b, int *npgs, int *ncgs, int *npshell) {
    *npgs = b->npgs; *ncgs = b->ncgs; *npshell = b->npshell;
}

void or_basis_export(const or_basis *b, int *pax, int *pay, int *paz, int *pgs_atom,
                     double *pgsx, double *pgsy, double *pgsz, double *zt, double *cc,
                     int *p_shell, int *p_shell_ang, double *cgsx, double *cgsy,
                     double *cgsz) {
    int np = b->npgs, nc = b->ncgs, ns = b->npshell;
    if (pax) memcpy(pax, b->pax, sizeof(int) * np);
    if (pay) memcpy(pay, b->pay, sizeof(int) * np);
    if (paz) memcpy(paz, b->paz, sizeof(int) * np);
    if (pgs_atom) memcpy(pgs_atom, b->pgs_atom, sizeof(int) * np);
    if (pgsx) memcpy(pgsx, b->pgsx, sizeof(double) * np);
    if (pgsy) memcpy(pgsy, b->pgsy, sizeof(double) * np);
    if (pgsz) memcpy(pgsz, b->pgsz, sizeof(double) * np);
    if (zt) memcpy(zt, b->zt, sizeof(double) * np);
    if (cc) memcpy(cc, b->cc, sizeof(double) * (size_t)nc * np);
    if (p_shell) memcpy(p_shell, b->p_shell, sizeof(int) * ns * 125);
    if (p_shell_ang) memcpy(p_shell_ang, b->p_shell_ang, sizeof(int) * ns);
    if (cgsx) memcpy(cgsx, b->cgsx, sizeof(double) * nc);
    if (cgsy) memcpy(cgsy, b->cgsy, sizeof(double) * nc);
    if (cgsz) memcpy(cgsz, b->cgsz, sizeof(double) * nc);
}

/* ------------------------------------------------------------------ */
/* Obara-Saika index tables for composite pairs                        */
/* (the role of src/pfourindexdirect.f90:101-221)                      */
/* ------------------------------------------------------------------ */
typedef struct {
    int n;                 /* components of the composite pair          */
    int lsum[400];         /* |a|+|b|                                   */
    int a[400][3], b[400][3];
    int dir[400];          /* reduction direction, -1 for (s,s)         */
    int on_b[400];         /* reduction acts on the second centre       */
    int d3[400];           /* one step down in dir                      */
    int d4[400];           /* two steps down on the same centre, -1     */
    double c4[400];
    int d5[400];           /* down on both centres, -1                  */
    double c5[400];
    int x1[400][3];        /* first centre down in direction k, -1      */
    double cx1[400][3];
    int x2[400][3];        /* second centre down in direction k, -1     */
    double cx2[400][3];
} pairtab;

static pairtab g_pt[(OR_LMAX + 1) * (OR_LMAX + 1)];
static int g_pt_ready = 0;

static int pt_index(const pairtab *t, const int *a, const int *b) {
    for (int i = 0; i < t->n; ++i)
        if (t->a[i][0] == a[0] && t->a[i][1] == a[1] && t->a[i][2] == a[2] &&
            t->b[i][0] == b[0] && t->b[i][1] == b[1] && t->b[i][2] == b[2])
            return i;
    return -1;
}

static void pairtab_init(void) {
    if (g_pt_ready) return;
    for (int la = 0; la <= OR_LMAX; ++la)
        for (int lb = 0; lb <= OR_LMAX; ++lb) {
            pairtab *t = &g_pt[la * (OR_LMAX + 1) + lb];
            int n = 0;
            for (int i1 = 0; i1 <= la; ++i1)
                for (int i2 = 0; i2 <= la - i1; ++i2)
                    for (int i3 = 0; i3 <= la - i1 - i2; ++i3)
                        for (int j1 = 0; j1 <= lb; ++j1)
                            for (int j2 = 0; j2 <= lb - j1; ++j2)
                                for (int j3 = 0; j3 <= lb - j1 - j2; ++j3) {
                                    t->a[n][0] = i1; t->a[n][1] = i2; t->a[n][2] = i3;
                                    t->b[n][0] = j1; t->b[n][1] = j2; t->b[n][2] = j3;
                                    t->lsum[n] = i1 + i2 + i3 + j1 + j2 + j3;
                                    ++n;
                                }
            t->n = n;
            for (int c = 0; c < n; ++c) {
                int a[3] = {t->a[c][0], t->a[c][1], t->a[c][2]};
                int b[3] = {t->b[c][0], t->b[c][1], t->b[c][2]};
                t->d4[c] = t->d5[c] = -1;
                t->c4[c] = t->c5[c] = 0;
                for (int k = 0; k < 3; ++k) {
                    t->x1[c][k] = t->x2[c][k] = -1;
                    t->cx1[c][k] = t->cx2[c][k] = 0;
                    if (a[k] >= 1) {
                        int a2[3] = {a[0], a[1], a[2]}; a2[k]--;
                        t->x1[c][k] = pt_index(t, a2, b);
                        t->cx1[c][k] = a[k] / 2.0;
                    }
                    if (b[k] >= 1) {
                        int b2[3] = {b[0], b[1], b[2]}; b2[k]--;
                        t->x2[c][k] = pt_index(t, a, b2);
                        t->cx2[c][k] = b[k] / 2.0;
                    }
                }
                int dir = -1, onb = 0;
                for (int k = 0; k < 3 && dir < 0; ++k) if (a[k] >= 1) dir = k;
                if (dir < 0) for (int k = 0; k < 3 && dir < 0; ++k) if (b[k] >= 1) { dir = k; onb = 1; }
                t->dir[c] = dir; t->on_b[c] = onb;
                if (dir < 0) { t->d3[c] = -1; continue; }
                if (!onb) {
                    int a2[3] = {a[0], a[1], a[2]}; a2[dir]--;
                    t->d3[c] = pt_index(t, a2, b);
                    if (a[dir] >= 2) {
                        int a3[3] = {a[0], a[1], a[2]}; a3[dir] -= 2;
                        t->d4[c] = pt_index(t, a3, b); t->c4[c] = (a[dir] - 1) / 2.0;
                    }
                    if (b[dir] >= 1) {
                        int b3[3] = {b[0], b[1], b[2]}; b3[dir]--;
                        t->d5[c] = pt_index(t, a2, b3); t->c5[c] = b[dir] / 2.0;
                    }
                } else {
                    int b2[3] = {b[0], b[1], b[2]}; b2[dir]--;
                    t->d3[c] = pt_index(t, a, b2);
                    if (b[dir] >= 2) {
                        int b3[3] = {b[0], b[1], b[2]}; b3[dir] -= 2;
                        t->d4[c] = pt_index(t, a, b3); t->c4[c] = (b[dir] - 1) / 2.0;
                    }
                }
            }
        }
    g_pt_ready = 1;
}

/* Geometry of one primitive composite pair (src/pfourindexdirect.f90:
 * 260-315, 367-416). */
typedef struct {
    double zeta, px, py, pz;
    double ax, ay, az, bx, by, bz;
    double ov;               /* S_P of the two s anchors              */
    int type, lang;
    const pairtab *t;
    int pi[400], pj[400];    /* primitive of each component            */
    double dca[400];         /* P-A (or P-B) along the reduction dir   */
    double bound;            /* IJIJ / KLKL                            */
    int real;                /* both anchors non-ghost shells          */
} cpair;

static double s_anchor(const or_basis *b, int p, int q) {
    double ax = b->pgsx[p], ay = b->pgsy[p], az = b->pgsz[p];
    double bx = b->pgsx[q], by = b->pgsy[q], bz = b->pgsz[q];
    double r = (ax - bx) * (ax - bx) + (ay - by) * (ay - by) + (az - bz) * (az - bz);
    double z = b->zt[p] + b->zt[q];
    return sqrt(PI_REF * PI_REF * PI_REF) / sqrt(z * z * z) * exp(-b->zt[p] * b->zt[q] / z * r);
}

static void cpair_make(const or_basis *b, int I, int J, cpair *cp) {
    pairtab_init();
    int la = b->p_shell_ang[I], lb = b->p_shell_ang[J];
    int ii = b->p_shell[I * 125], jj = b->p_shell[J * 125];
    cp->ax = b->pgsx[ii]; cp->ay = b->pgsy[ii]; cp->az = b->pgsz[ii];
    cp->bx = b->pgsx[jj]; cp->by = b->pgsy[jj]; cp->bz = b->pgsz[jj];
    double z = b->zt[ii] + b->zt[jj];
    cp->zeta = z;
    cp->px = (cp->ax * b->zt[ii] + cp->bx * b->zt[jj]) / z;
    cp->py = (cp->ay * b->zt[ii] + cp->by * b->zt[jj]) / z;
    cp->pz = (cp->az * b->zt[ii] + cp->bz * b->zt[jj]) / z;
    cp->ov = s_anchor(b, ii, jj);
    cp->type = la * (OR_LMAX + 1) + lb;
    cp->lang = la + lb;
    cp->t = &g_pt[cp->type];
    const pairtab *t = cp->t;
    double pa[3] = {cp->px - cp->ax, cp->py - cp->ay, cp->pz - cp->az};
    double pb[3] = {cp->px - cp->bx, cp->py - cp->by, cp->pz - cp->bz};
    for (int c = 0; c < t->n; ++c) {
        cp->pi[c] = b->p_shell[I * 125 + t->a[c][0] * 25 + t->a[c][1] * 5 + t->a[c][2]];
        cp->pj[c] = b->p_shell[J * 125 + t->b[c][0] * 25 + t->b[c][1] * 5 + t->b[c][2]];
        cp->dca[c] = (t->dir[c] < 0) ? 0.0 : (t->on_b[c] ? pb[t->dir[c]] : pa[t->dir[c]]);
    }
    cp->real = 1;
}

static double wall_seconds(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* Composite quartet OS recursion: ket first ([00|kl]), then bra, as in
 * src/pfourindexdirect.f90:554-690.  e holds ERI(m, kl, ij) at
 * e[(ij*nkl + kl)*ms + m] with ms = L+1. */
static void os_quartet(const cpair *ij, const cpair *kl, double c2, double t,
                       int diag, double *e) {
    const pairtab *ti = ij->t, *tk = kl->t;
    int L = ij->lang + kl->lang;
    int ms = L + 1;
    int nkl = tk->n, nij = ti->n;
    double zeta = ij->zeta, eta = kl->zeta, ze = zeta + eta;
    for (int m = 0; m <= L; ++m)
        e[m] = diag ? c2 / (double)(2 * m + 1) : c2 * or_boys(t, m);
    if (L == 0) return;
    double wx = (ij->px * zeta + kl->px * eta) / ze;
    double wy = (ij->py * zeta + kl->py * eta) / ze;
    double wz = (ij->pz * zeta + kl->pz * eta) / ze;
    double wp[3] = {wx - ij->px, wy - ij->py, wz - ij->pz};
    double wq[3] = {wx - kl->px, wy - kl->py, wz - kl->pz};
    if (diag) { wp[0] = wp[1] = wp[2] = wq[0] = wq[1] = wq[2] = 0.0; }
    double rho_eta = zeta / ze, rho_zeta = eta / ze;
#define E(I, K, M) e[((size_t)(I) * nkl + (K)) * ms + (M)]
    for (int k = 1; k < nkl; ++k) {
        int d = tk->dir[k], k3 = tk->d3[k], k4 = tk->d4[k], k5 = tk->d5[k];
        int mmax = L - tk->lsum[k];
        for (int m = 0; m <= mmax; ++m) {
            double v = kl->dca[k] * E(0, k3, m) + wq[d] * E(0, k3, m + 1);
            if (k4 >= 0) v += tk->c4[k] / eta * (E(0, k4, m) - rho_eta * E(0, k4, m + 1));
            if (k5 >= 0) v += tk->c5[k] / eta * (E(0, k5, m) - rho_eta * E(0, k5, m + 1));
            E(0, k, m) = v;
        }
    }
    for (int i = 1; i < nij; ++i) {
        int d = ti->dir[i], i3 = ti->d3[i], i4 = ti->d4[i], i5 = ti->d5[i];
        for (int k = 0; k < nkl; ++k) {
            int mmax = L - ti->lsum[i] - tk->lsum[k];
            int x1 = tk->x1[k][d], x2 = tk->x2[k][d];
            for (int m = 0; m <= mmax; ++m) {
                double v = ij->dca[i] * E(i3, k, m) + wp[d] * E(i3, k, m + 1);
                if (i4 >= 0) v += ti->c4[i] / zeta * (E(i4, k, m) - rho_zeta * E(i4, k, m + 1));
                if (i5 >= 0) v += ti->c5[i] / zeta * (E(i5, k, m) - rho_zeta * E(i5, k, m + 1));
                if (x1 >= 0) v += tk->cx1[k][d] / ze * E(i3, x1, m + 1);
                if (x2 >= 0) v += tk->cx2[k][d] / ze * E(i3, x2, m + 1);
                E(i, k, m) = v;
            }
        }
    }
#undef E
}

/* MAXC(prim) = max |CC(:,prim)|, src/pfourindexdirect.f90:94-99 */
static double *maxc_make(const or_basis *b) {
    double *mc = (double *)calloc(b->npgs, sizeof(double));
    for (int p = 0; p < b->npgs; ++p)
        for (int c = 0; c < b->ncgs; ++c) {
            double v = fabs(b->cc[c + (size_t)b->ncgs * p]);
            if (v > mc[p]) mc[p] = v;
        }
    return mc;
}

/* Schwarz-type pair bound, src/pfourindexdirect.f90:317-364 / :418-462:
 * the s-s component is excluded when the pair carries angular momentum. */
static void cpair_bound(cpair *cp, const double *maxc, double *scratch) {
    double rho = 0.5 * cp->zeta;
    double c2 = 2.0 * sqrt(rho / PI_REF) * cp->ov * cp->ov;
    if (cp->lang == 0) {
        cp->bound = sqrt(c2 * maxc[cp->pi[0]] * maxc[cp->pj[0]]);
        return;
    }
    os_quartet(cp, cp, c2, 0.0, 1, scratch);
    int n = cp->t->n, ms = 2 * cp->lang + 1;
    double best = 0.0;
    for (int c = 1; c < n; ++c) {
        double v = scratch[((size_t)c * n + c) * ms];
        double a = sqrt(v * maxc[cp->pi[c]] * maxc[cp->pj[c]]);
        if (a > best) best = a;
    }
    cp->bound = best;
}

static int ncart(int l) { return (l + 1) * (l + 2) / 2; }

double or_flop_model(int la, int lb, int lc, int ld) {
    int L = la + lb + lc + ld;
    double f = 25 + 3 * L;
    double v = 0;
    for (int e = 0; e <= la + lb; ++e)
        for (int g = 0; g <= lc + ld; ++g)
            if (e + g >= 1) v += (double)ncart(e) * ncart(g) * (L - e - g + 1);
    return f + 8 * v + 12.0 * ncart(la) * ncart(lb) * ncart(lc) * ncart(ld);
}

/* Angular-momentum levels of a composite shell that carry contracted
 * (non-ghost) primitives, as a bit mask: an SP shell is {0,1}, a p shell
 * anchored on a ghost s primitive is {1}. */
static int shell_real_mask(const or_basis *b, int s, const double *maxc) {
    int mask = 0;
    for (int i = 0; i < 125; ++i) {
        int p = b->p_shell[s * 125 + i];
        if (p >= 0 && maxc[p] != 0.0) mask |= 1 << (b->pax[p] + b->pay[p] + b->paz[p]);
    }
    return mask;
}

/* ------------------------------------------------------------------ */
/* Direct J/K: src/pfourindexdirect.f90:1-814 (molecule: Q1=Q2=Q3=0)  */
/* ------------------------------------------------------------------ */
/* One bra row (I,J) of the loop nest: every KL of the table against the
 * row's composite pair, digested into the primitive-basis C and X. */
static void eri_row(const or_basis *b, long long sij, const double *p_p, double inttol, int recompute_kl,
                    const double *maxc, const int *treal, cpair *kl, double *e, double *c_p, double *x_p,
                    long long *nq, long long *nreal, double *flops, int want_flops) {
    int ns = b->npshell, np = b->npgs;
    size_t nsq = (size_t)ns * ns;
    if (sij < 0 || sij >= (long long)nsq) return;
    const double c1f = 2.0 / sqrt(PI_REF);
    int I = (int)(sij / ns), J = (int)(sij % ns);
    cpair ij;
    cpair_make(b, I, J, &ij);
    cpair_bound(&ij, maxc, e);
    if (recompute_kl) {
        /* the reference rebuilds the KL pair table inside every (I,J,Q3) */
        for (int k = 0; k < ns; ++k)
            for (int l = 0; l < ns; ++l) {
                double bd = kl[(size_t)k * ns + l].bound;
                cpair_make(b, k, l, &kl[(size_t)k * ns + l]);
                kl[(size_t)k * ns + l].bound = bd;
            }
    }
    double c1 = ij.ov * c1f;
    for (size_t skl = 0; skl < nsq; ++skl) {
        const cpair *q = &kl[skl];
        if (ij.bound * q->bound < inttol) continue;
        ++*nq;
        int K = (int)(skl / ns), Lq = (int)(skl % ns);
        for (int la = 0; la <= OR_LMAX; ++la) if (treal[I] >> la & 1)
        for (int lb = 0; lb <= OR_LMAX; ++lb) if (treal[J] >> lb & 1)
        for (int lc = 0; lc <= OR_LMAX; ++lc) if (treal[K] >> lc & 1)
        for (int ld = 0; ld <= OR_LMAX; ++ld) if (treal[Lq] >> ld & 1) {
            ++*nreal;
            if (want_flops) *flops += or_flop_model(la, lb, lc, ld);
        }
        double ze = ij.zeta + q->zeta;
        double rho = ij.zeta * q->zeta / ze;
        double c2 = c1 * sqrt(rho) * q->ov;
        double dx = q->px - ij.px, dy = q->py - ij.py, dz = q->pz - ij.pz;
        double t = rho * (dx * dx + dy * dy + dz * dz);
        os_quartet(&ij, q, c2, t, 0, e);
        int nij = ij.t->n, nk = q->t->n, ms = ij.lang + q->lang + 1;
        for (int a = 0; a < nij; ++a) {
            int pi = ij.pi[a], pj = ij.pj[a];
            for (int c = 0; c < nk; ++c) {
                double v = e[((size_t)a * nk + c) * ms];
                int pk = q->pi[c], pl = q->pj[c];
                c_p[pi + (size_t)np * pj] += p_p[pk + (size_t)np * pl] * v;
                x_p[pi + (size_t)np * pl] += p_p[pk + (size_t)np * pj] * v;
            }
        }
    }
}

long long or_fourindex_eri_rows(const or_basis *b, const double *s_p, const double *p_p,
                                double inttol, int recompute_kl, const long long *rows,
                                long long nrows, double *c_p, double *x_p,
                                long long *real_quartets, double *flop_model,
                                double *setup_seconds, double *row_seconds) {
    (void)s_p;
    boys_init();
    pairtab_init();
    int ns = b->npshell;
    double *maxc = maxc_make(b);
    int *treal = (int *)malloc(sizeof(int) * ns);
    for (int s = 0; s < ns; ++s) treal[s] = shell_real_mask(b, s, maxc);
    size_t nsq = (size_t)ns * ns;
    cpair *kl = (cpair *)malloc(sizeof(cpair) * nsq);
    double *e = (double *)malloc(sizeof(double) * 400 * 400 * (4 * OR_LMAX + 2));
    double t0 = wall_seconds();
    for (int k = 0; k < ns; ++k)
        for (int l = 0; l < ns; ++l) {
            cpair_make(b, k, l, &kl[(size_t)k * ns + l]);
            cpair_bound(&kl[(size_t)k * ns + l], maxc, e);
        }
    double t1 = wall_seconds();
    if (setup_seconds) *setup_seconds = t1 - t0;
    long long nq = 0, nreal = 0;
    double flops = 0;
    for (long long r = 0; r < nrows; ++r)
        eri_row(b, rows[r], p_p, inttol, recompute_kl, maxc, treal, kl, e, c_p, x_p, &nq, &nreal, &flops,
                flop_model != NULL);
    if (row_seconds) *row_seconds = wall_seconds() - t1;
    free(e); free(kl); free(maxc); free(treal);
    if (real_quartets) *real_quartets = nreal;
    if (flop_model) *flop_model = flops;
    return nq;
}

/* The same loop nest with the rows spread over nthreads OpenMP threads
 * (dynamic schedule, one row at a time), each thread with its own KL
 * table, work buffer and C / X accumulators (summed at the end), so every
 * thread does exactly the single-threaded per-row work.  setup_seconds: the
 * tables' construction (concurrent); row_seconds: the wall time of the
 * rows. */
long long or_fourindex_eri_rows_mt(const or_basis *b, const double *p_p, double inttol, int recompute_kl,
                                   const long long *rows, long long nrows, int nthreads, double *c_p,
                                   double *x_p, long long *real_quartets, double *setup_seconds,
                                   double *row_seconds) {
    boys_init();
    pairtab_init();
    int ns = b->npshell, np = b->npgs;
    double *maxc = maxc_make(b);
    int *treal = (int *)malloc(sizeof(int) * ns);
    for (int s = 0; s < ns; ++s) treal[s] = shell_real_mask(b, s, maxc);
    size_t nsq = (size_t)ns * ns, np2 = (size_t)np * np;
    long long nq = 0, nreal = 0;
    double t0 = 0, t1 = 0, t2 = 0;
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads) reduction(+ : nq, nreal)
    {
        cpair *kl = (cpair *)malloc(sizeof(cpair) * nsq);
        double *e = (double *)malloc(sizeof(double) * 400 * 400 * (4 * OR_LMAX + 2));
        double *cl = (double *)calloc(np2, sizeof(double)), *xl = (double *)calloc(np2, sizeof(double));
        double fl = 0;
#pragma omp barrier
#pragma omp master
        t0 = wall_seconds();
        for (int k = 0; k < ns; ++k)
            for (int l = 0; l < ns; ++l) {
                cpair_make(b, k, l, &kl[(size_t)k * ns + l]);
                cpair_bound(&kl[(size_t)k * ns + l], maxc, e);
            }
#pragma omp barrier
#pragma omp master
        t1 = wall_seconds();
#pragma omp barrier
#pragma omp for schedule(dynamic, 1)
        for (long long r = 0; r < nrows; ++r)
            eri_row(b, rows[r], p_p, inttol, recompute_kl, maxc, treal, kl, e, cl, xl, &nq, &nreal, &fl, 0);
#pragma omp master
        t2 = wall_seconds();
#pragma omp critical
        for (size_t i = 0; i < np2; ++i) {
            c_p[i] += cl[i];
            x_p[i] += xl[i];
        }
        free(cl); free(xl); free(e); free(kl);
    }
    free(maxc); free(treal);
    if (setup_seconds) *setup_seconds = t1 - t0;
    if (row_seconds) *row_seconds = t2 - t1;
    if (real_quartets) *real_quartets = nreal;
    return nq;
}

static void contract_cc(const or_basis *b, const double *m_p, double *m_c);

/* ------------------------------------------------------------------ */
/* Periodic direct J/K: src/pfourindexdirect.f90:222-750 with HELIX = 0 */
/* and the periodic axis x (the reference's only 1-D axis).            */
/* ------------------------------------------------------------------ */
/* cpair_make with the two anchors translated along x by sI and sJ
 * (src/pfourindexdirect.f90:262-270, 369-373, 470-479: shells of cell Q
 * sit at PGSX + Q*PERIODX). */
static void cpair_make_shift(const or_basis *b, int I, double sI, int J, double sJ, cpair *cp) {
    cpair_make(b, I, J, cp);
    if (sI == 0.0 && sJ == 0.0) return;
    int ii = b->p_shell[I * 125], jj = b->p_shell[J * 125];
    cp->ax += sI;
    cp->bx += sJ;
    double z = cp->zeta;
    cp->px = (cp->ax * b->zt[ii] + cp->bx * b->zt[jj]) / z;
    double r = (cp->ax - cp->bx) * (cp->ax - cp->bx) + (cp->ay - cp->by) * (cp->ay - cp->by) +
               (cp->az - cp->bz) * (cp->az - cp->bz);
    cp->ov = sqrt(PI_REF * PI_REF * PI_REF) / sqrt(z * z * z) * exp(-b->zt[ii] * b->zt[jj] / z * r);
    const pairtab *t = cp->t;
    double pa[3] = {cp->px - cp->ax, cp->py - cp->ay, cp->pz - cp->az};
    double pb[3] = {cp->px - cp->bx, cp->py - cp->by, cp->pz - cp->bz};
    for (int c = 0; c < t->n; ++c)
        cp->dca[c] = (t->dir[c] < 0) ? 0.0 : (t->on_b[c] ? pb[t->dir[c]] : pa[t->dir[c]]);
}

long long or_fourindex_eri_cells(const or_basis *b, double period, int rc1, int cel1, int cel2,
                                 const double *p_p, double inttol, double *c_p, double *x_p) {
    boys_init();
    pairtab_init();
    int ns = b->npshell, np = b->npgs, ncell = 2 * cel1 + 1, n1 = 2 * rc1 + 1;
    size_t np2 = (size_t)np * np, nsq = (size_t)ns * ns;
    double *maxc = maxc_make(b);
    const double c1f = 2.0 / sqrt(PI_REF);
    long long nq = 0;
    /* (Q1, Q2) tickets spread over the host threads as the reference spreads
     * them over MPI ranks (:224-240); each thread sums into its own J/K
     * cells, added in thread order afterwards (its MPI_ALLREDUCE, :757-764) */
    int nt = omp_get_max_threads();
    double *cpt = (double *)calloc((size_t)nt * ncell * np2, sizeof(double));
    double *xpt = (double *)calloc((size_t)nt * ncell * np2, sizeof(double));
#pragma omp parallel num_threads(nt) reduction(+ : nq)
    {
        double *cpl = cpt + (size_t)omp_get_thread_num() * ncell * np2;
        double *xpl = xpt + (size_t)omp_get_thread_num() * ncell * np2;
        cpair *kl0 = (cpair *)malloc(sizeof(cpair) * nsq);
        cpair ij, q;
        double *e = (double *)malloc(sizeof(double) * 400 * 400 * (4 * OR_LMAX + 2));
#pragma omp for schedule(static, 1)
        for (int tk = 0; tk < n1 * n1; ++tk) {
            const int q1 = tk / n1 - rc1, q2 = tk % n1 - rc1;
            /* KLKL with K in cell 0 and L in cell Q2 (:258-364) */
            for (int k = 0; k < ns; ++k)
                for (int l = 0; l < ns; ++l) {
                    cpair_make_shift(b, k, 0.0, l, q2 * period, &kl0[(size_t)k * ns + l]);
                    cpair_bound(&kl0[(size_t)k * ns + l], maxc, e);
                }
            const double *pq2 = p_p + (size_t)(q2 + cel1) * np2;
            for (int I = 0; I < ns; ++I)
                for (int J = 0; J < ns; ++J) {
                    cpair_make_shift(b, I, 0.0, J, q1 * period, &ij);
                    cpair_bound(&ij, maxc, e);
                    double c1 = ij.ov * c1f;
                    double *cq1 = cpl + (size_t)(q1 + cel1) * np2;
                    for (int q3 = -cel2; q3 <= cel2; ++q3) {
                        /* LCEL1 (:247-259): exchange only into stored cells */
                        int lx = (q3 + q2 >= -cel1 && q3 + q2 <= cel1 && q1 - q3 >= -cel1 && q1 - q3 <= cel1);
                        double *xq = lx ? xpl + (size_t)(q3 + q2 + cel1) * np2 : NULL;
                        const double *pq13 = lx ? p_p + (size_t)(q1 - q3 + cel1) * np2 : NULL;
                        for (size_t skl = 0; skl < nsq; ++skl) {
                            if (ij.bound * kl0[skl].bound < inttol) continue;
                            int K = (int)(skl / ns), L = (int)(skl % ns);
                            cpair_make_shift(b, K, q3 * period, L, (q3 + q2) * period, &q);
                            ++nq;
                            double ze = ij.zeta + q.zeta;
                            double rho = ij.zeta * q.zeta / ze;
                            double c2 = c1 * sqrt(rho) * q.ov;
                            double dx = q.px - ij.px, dy = q.py - ij.py, dz = q.pz - ij.pz;
                            double t = rho * (dx * dx + dy * dy + dz * dz);
                            os_quartet(&ij, &q, c2, t, 0, e);
                            int nij = ij.t->n, nk = q.t->n, ms = ij.lang + q.lang + 1;
                            for (int a = 0; a < nij; ++a) {
                                int pi = ij.pi[a], pj = ij.pj[a];
                                for (int c = 0; c < nk; ++c) {
                                    double v = e[((size_t)a * nk + c) * ms];
                                    int pk = q.pi[c], pl = q.pj[c];
                                    cq1[pi + (size_t)np * pj] += pq2[pk + (size_t)np * pl] * v;
                                    if (lx) xq[pi + (size_t)np * pl] += pq13[pk + (size_t)np * pj] * v;
                                }
                            }
                        }
                    }
                }
        }
        free(e); free(kl0);
    }
    for (int th = 0; th < nt; ++th)
        for (size_t i = 0; i < (size_t)ncell * np2; ++i) {
            c_p[i] += cpt[(size_t)th * ncell * np2 + i];
            x_p[i] += xpt[(size_t)th * ncell * np2 + i];
        }
    free(cpt); free(xpt); free(maxc);
    return nq;
}

/* Contracted periodic J/K (C_C, X_C per cell, src/pfourindexdirect.f90:
 * 765-781): p_c, c_c, x_c hold 2 cel1 + 1 column-major ncgs^2 blocks. */
long long or_fourindex_eri_cells_c(const or_basis *b, double period, int rc1, int cel1, int cel2,
                                   const double *p_c, double inttol, double *c_c, double *x_c) {
    int np = b->npgs, nc = b->ncgs, ncell = 2 * cel1 + 1;
    size_t np2 = (size_t)np * np, nc2 = (size_t)nc * nc;
    double *pp = (double *)calloc(np2 * ncell, sizeof(double));
    double *cp = (double *)calloc(np2 * ncell, sizeof(double));
    double *xp = (double *)calloc(np2 * ncell, sizeof(double));
    for (int c = 0; c < ncell; ++c) or_decontract(b, p_c + nc2 * c, pp + np2 * c);
    long long n = or_fourindex_eri_cells(b, period, rc1, cel1, cel2, pp, inttol, cp, xp);
    for (int c = 0; c < ncell; ++c) {
        contract_cc(b, cp + np2 * c, c_c + nc2 * c);
        contract_cc(b, xp + np2 * c, x_c + nc2 * c);
    }
    free(pp); free(cp); free(xp);
    return n;
}

long long or_fourindex_eri_range(const or_basis *b, const double *s_p, const double *p_p,
                                 double inttol, int recompute_kl, long long ij_begin,
                                 long long ij_end, double *c_p, double *x_p,
                                 long long *real_quartets, double *flop_model) {
    long long n = ij_end > ij_begin ? ij_end - ij_begin : 0;
    long long *rows = (long long *)malloc(sizeof(long long) * (n ? n : 1));
    for (long long i = 0; i < n; ++i) rows[i] = ij_begin + i;
    long long r = or_fourindex_eri_rows(b, s_p, p_p, inttol, recompute_kl, rows, n, c_p, x_p,
                                        real_quartets, flop_model, NULL, NULL);
    free(rows);
    return r;
}

static void contract_cc(const or_basis *b, const double *m_p, double *m_c) {
    /* M_C = CC M_P CC^T, src/pfourindexdirect.f90:765-781 */
    int np = b->npgs, nc = b->ncgs;
    double *w = (double *)calloc((size_t)nc * np, sizeof(double));
    for (int j = 0; j < np; ++j)
        for (int k = 0; k < np; ++k) {
            double v = m_p[k + (size_t)np * j];
            if (v == 0.0) continue;
            for (int i = 0; i < nc; ++i) w[i + (size_t)nc * j] += b->cc[i + (size_t)nc * k] * v;
        }
    for (int i = 0; i < nc * nc; ++i) m_c[i] = 0.0;
    for (int j = 0; j < nc; ++j)
        for (int k = 0; k < np; ++k) {
            double ccv = b->cc[j + (size_t)nc * k];
            if (ccv == 0.0) continue;
            for (int i = 0; i < nc; ++i) m_c[i + (size_t)nc * j] += w[i + (size_t)nc * k] * ccv;
        }
    free(w);
}

long long or_fourindex_eri(const or_basis *b, const double *s_p, const double *p_p,
                           double inttol, int recompute_kl, double *c_c, double *x_c,
                           long long *real_quartets, double *flop_model) {
    int np = b->npgs;
    double *cp = (double *)calloc((size_t)np * np, sizeof(double));
    double *xp = (double *)calloc((size_t)np * np, sizeof(double));
    long long n = or_fourindex_eri_range(b, s_p, p_p, inttol, recompute_kl, 0,
                                         (long long)b->npshell * b->npshell, cp, xp,
                                         real_quartets, flop_model);
    contract_cc(b, cp, c_c);
    contract_cc(b, xp, x_c);
    free(cp); free(xp);
    return n;
}

void or_eri_tensor(const or_basis *b, const double *s_p, double *feri) {
    (void)s_p;
    boys_init();
    pairtab_init();
    int ns = b->npshell, nc = b->ncgs;
    size_t nc2 = (size_t)nc * nc;
    memset(feri, 0, sizeof(double) * nc2 * nc2);
    double *maxc = maxc_make(b);
    size_t nsq = (size_t)ns * ns;
    cpair *pr = (cpair *)malloc(sizeof(cpair) * nsq);
    double *e = (double *)malloc(sizeof(double) * 400 * 400 * (4 * OR_LMAX + 2));
    for (size_t s = 0; s < nsq; ++s) cpair_make(b, (int)(s / ns), (int)(s % ns), &pr[s]);
    /* contracted functions of each primitive */
    int np = b->npgs;
    int *ncp = (int *)calloc(np, sizeof(int));
    int *cid = (int *)malloc(sizeof(int) * np * 4);
    double *cv = (double *)malloc(sizeof(double) * np * 4);
    for (int p = 0; p < np; ++p)
        for (int c = 0; c < nc; ++c) {
            double v = b->cc[c + (size_t)nc * p];
            if (v != 0.0 && ncp[p] < 4) { cid[p * 4 + ncp[p]] = c; cv[p * 4 + ncp[p]] = v; ncp[p]++; }
        }
    const double c1f = 2.0 / sqrt(PI_REF);
    for (size_t sij = 0; sij < nsq; ++sij) {
        const cpair *ij = &pr[sij];
        double c1 = ij->ov * c1f;
        for (size_t skl = 0; skl < nsq; ++skl) {
            const cpair *q = &pr[skl];
            double ze = ij->zeta + q->zeta, rho = ij->zeta * q->zeta / ze;
            double c2 = c1 * sqrt(rho) * q->ov;
            double dx = q->px - ij->px, dy = q->py - ij->py, dz = q->pz - ij->pz;
            os_quartet(ij, q, c2, rho * (dx * dx + dy * dy + dz * dz), 0, e);
            int nij = ij->t->n, nk = q->t->n, ms = ij->lang + q->lang + 1;
            for (int a = 0; a < nij; ++a) {
                int pi = ij->pi[a], pj = ij->pj[a];
                if (!ncp[pi] || !ncp[pj]) continue;
                for (int c = 0; c < nk; ++c) {
                    int pk = q->pi[c], pl = q->pj[c];
                    if (!ncp[pk] || !ncp[pl]) continue;
                    double v = e[((size_t)a * nk + c) * ms];
                    for (int u = 0; u < ncp[pi]; ++u)
                        for (int w = 0; w < ncp[pj]; ++w)
                            for (int x = 0; x < ncp[pk]; ++x)
                                for (int y = 0; y < ncp[pl]; ++y) {
                                    double f = cv[pi * 4 + u] * cv[pj * 4 + w] * cv[pk * 4 + x] *
                                               cv[pl * 4 + y];
                                    size_t idx = cid[pi * 4 + u] + nc * (cid[pj * 4 + w] +
                                                 (size_t)nc * (cid[pk * 4 + x] +
                                                 (size_t)nc * cid[pl * 4 + y]));
                                    feri[idx] += f * v;
                                }
                }
            }
        }
    }
    free(ncp); free(cid); free(cv); free(e); free(pr); free(maxc);
}

/* The lattice integrals STORE_FOURINDEX_ERI writes to the .ao files
 * (src/pfourindexstore.f90:258-747, unscreened):
 * T[Q1][Q2][Q3][i][j][k][l] = (i_0 j_Q1 | k_Q3 l_Q3+Q2), Q1, Q2 in
 * [-rc1, rc1], Q3 in [-cel2, cel2], contracted functions in reference
 * order, the last index fastest. */
void or_eri_tensor_cells(const or_basis *b, double period, int rc1, int cel2, double *T) {
    boys_init();
    pairtab_init();
    int ns = b->npshell, nc = b->ncgs, np = b->npgs;
    const int n1 = 2 * rc1 + 1, n3 = 2 * cel2 + 1;
    const size_t nc4 = (size_t)nc * nc * nc * nc;
    memset(T, 0, sizeof(double) * nc4 * n1 * n1 * n3);
    size_t nsq = (size_t)ns * ns;
    cpair *bra = (cpair *)malloc(sizeof(cpair) * nsq);
    cpair q;
    double *e = (double *)malloc(sizeof(double) * 400 * 400 * (4 * OR_LMAX + 2));
    int *ncp = (int *)calloc(np, sizeof(int));
    int *cid = (int *)malloc(sizeof(int) * np * 4);
    double *cv = (double *)malloc(sizeof(double) * np * 4);
    for (int p = 0; p < np; ++p)
        for (int c = 0; c < nc; ++c) {
            double v = b->cc[c + (size_t)nc * p];
            if (v != 0.0 && ncp[p] < 4) { cid[p * 4 + ncp[p]] = c; cv[p * 4 + ncp[p]] = v; ncp[p]++; }
        }
    const double c1f = 2.0 / sqrt(PI_REF);
    for (int q1 = -rc1; q1 <= rc1; ++q1) {
        for (size_t sij = 0; sij < nsq; ++sij)
            cpair_make_shift(b, (int)(sij / ns), 0.0, (int)(sij % ns), q1 * period, &bra[sij]);
        for (int q2 = -rc1; q2 <= rc1; ++q2)
            for (int q3 = -cel2; q3 <= cel2; ++q3) {
                double *tq = T + (((size_t)(q1 + rc1) * n1 + (q2 + rc1)) * n3 + (q3 + cel2)) * nc4;
                for (size_t skl = 0; skl < nsq; ++skl) {
                    cpair_make_shift(b, (int)(skl / ns), q3 * period, (int)(skl % ns), (q3 + q2) * period, &q);
                    for (size_t sij = 0; sij < nsq; ++sij) {
                        const cpair *ij = &bra[sij];
                        double ze = ij->zeta + q.zeta, rho = ij->zeta * q.zeta / ze;
                        double c2 = ij->ov * c1f * sqrt(rho) * q.ov;
                        double dx = q.px - ij->px, dy = q.py - ij->py, dz = q.pz - ij->pz;
                        os_quartet(ij, &q, c2, rho * (dx * dx + dy * dy + dz * dz), 0, e);
                        int nij = ij->t->n, nk = q.t->n, ms = ij->lang + q.lang + 1;
                        for (int a = 0; a < nij; ++a) {
                            int pi = ij->pi[a], pj = ij->pj[a];
                            if (!ncp[pi] || !ncp[pj]) continue;
                            for (int c = 0; c < nk; ++c) {
                                int pk = q.pi[c], pl = q.pj[c];
                                if (!ncp[pk] || !ncp[pl]) continue;
                                double v = e[((size_t)a * nk + c) * ms];
                                for (int u = 0; u < ncp[pi]; ++u)
                                    for (int w = 0; w < ncp[pj]; ++w)
                                        for (int x = 0; x < ncp[pk]; ++x)
                                            for (int y = 0; y < ncp[pl]; ++y) {
                                                double f = cv[pi * 4 + u] * cv[pj * 4 + w] * cv[pk * 4 + x] *
                                                           cv[pl * 4 + y];
                                                size_t idx = (((size_t)cid[pi * 4 + u] * nc + cid[pj * 4 + w]) * nc +
                                                              cid[pk * 4 + x]) * nc + cid[pl * 4 + y];
                                                tq[idx] += f * v;
                                            }
                            }
                        }
                    }
                }
            }
    }
    free(ncp); free(cid); free(cv); free(e); free(bra);
}

/* ------------------------------------------------------------------ */
/* One-electron integrals: src/ponee.f90:1-414, :418-728              */
/* ------------------------------------------------------------------ */
/* 1-D Obara-Saika overlap and kinetic tables for a primitive pair, then
 * components by products -- the reference runs the same recursions on the
 * full Cartesian index (src/ponee.f90:83-142); both give identical values. */
static void os_overlap_1d(double pa, double pb, double z, int la, int lb, double s[6][6]) {
    /* s[i][j] = <i|j> / s00 along one axis */
    for (int i = 0; i <= la + 1; ++i) for (int j = 0; j <= lb + 1; ++j) s[i][j] = 0.0;
    s[0][0] = 1.0;
    for (int i = 0; i <= la + 1; ++i)
        for (int j = 0; j <= lb + 1; ++j) {
            if (i == 0 && j == 0) continue;
            double v;
            if (i > 0) {
                v = pa * s[i - 1][j];
                if (i > 1) v += (i - 1) / (2.0 * z) * s[i - 2][j];
                if (j > 0) v += j / (2.0 * z) * s[i - 1][j - 1];
            } else {
                v = pb * s[i][j - 1];
                if (j > 1) v += (j - 1) / (2.0 * z) * s[i][j - 2];
            }
            s[i][j] = v;
        }
}

void or_one_electron(const or_basis *b, double *s_p, double *s_c, double *t_c, double *n_c) {
    boys_init();
    pairtab_init();
    int ns = b->npshell, np = b->npgs;
    double *sp = s_p, *tp, *vp;
    tp = (double *)calloc((size_t)np * np, sizeof(double));
    vp = (double *)calloc((size_t)np * np, sizeof(double));
    for (size_t i = 0; i < (size_t)np * np; ++i) sp[i] = 0.0;
    for (int I = 0; I < ns; ++I)
        for (int J = 0; J < ns; ++J) {
            cpair cp;
            cpair_make(b, I, J, &cp);
            int ii = b->p_shell[I * 125], jj = b->p_shell[J * 125];
            double ai = b->zt[ii], bj = b->zt[jj], z = cp.zeta;
            double pa[3] = {cp.px - cp.ax, cp.py - cp.ay, cp.pz - cp.az};
            double pb[3] = {cp.px - cp.bx, cp.py - cp.by, cp.pz - cp.bz};
            int la = b->p_shell_ang[I], lb = b->p_shell_ang[J];
            double s1[3][6][6];
            for (int d = 0; d < 3; ++d) os_overlap_1d(pa[d], pb[d], z, la + 1, lb + 1, s1[d]);
            const pairtab *t = cp.t;
            /* kinetic: T = sum_d <a|-1/2 d^2/dx_d^2|b> via the standard
             * identity -1/2 <i|d2|j> = -1/2 [j(j-1)<i|j-2> - 2b(2j+1)<i|j>
             *                                 + 4 b^2 <i|j+2>] */
            for (int c = 0; c < t->n; ++c) {
                int pi = cp.pi[c], pj = cp.pj[c];
                const int *A = t->a[c], *B = t->b[c];
                double sx = s1[0][A[0]][B[0]], sy = s1[1][A[1]][B[1]], sz = s1[2][A[2]][B[2]];
                double s = cp.ov * sx * sy * sz;
                sp[pi + (size_t)np * pj] = s;
                double kin = 0.0;
                for (int d = 0; d < 3; ++d) {
                    int j = B[d], i = A[d];
                    double o = 1.0;
                    for (int e = 0; e < 3; ++e) if (e != d) o *= s1[e][A[e]][B[e]];
                    double k1 = (j >= 2) ? j * (j - 1) * s1[d][i][j - 2] : 0.0;
                    double k2 = -2.0 * bj * (2 * j + 1) * s1[d][i][j];
                    double k3 = 4.0 * bj * bj * s1[d][i][j + 2];
                    kin += -0.5 * (k1 + k2 + k3) * o;
                }
                tp[pi + (size_t)np * pj] = cp.ov * kin;
            }
            (void)ai;
            /* nuclear attraction (src/ponee.f90:504-600): OS on the full
             * Cartesian index with the Taylor-table Boys function. */
            int L = la + lb;
            double e1 = 2.0 / sqrt(PI_REF) * sqrt(z) * cp.ov;
            /* N1 over components x m, reusing the composite tables */
            static double n1[400][16];
            for (int c = 0; c < t->n; ++c) vp[cp.pi[c] + (size_t)np * cp.pj[c]] += 0.0;
            for (int k = 0; k < b->natom; ++k) {
                double e2 = -(double)b->iatom[k] * e1;
                double pc[3] = {cp.px - b->atomx[k], cp.py - b->atomy[k], cp.pz - b->atomz[k]};
                double T = z * (pc[0] * pc[0] + pc[1] * pc[1] + pc[2] * pc[2]);
                for (int m = 0; m <= L; ++m) n1[0][m] = e2 * or_boys(T, m);
                for (int c = 1; c < t->n; ++c) {
                    int d = t->dir[c], c3 = t->d3[c], c4 = t->d4[c], c5 = t->d5[c];
                    int mmax = L - t->lsum[c];
                    double dd = t->on_b[c] ? pb[d] : pa[d];
                    for (int m = 0; m <= mmax; ++m) {
                        double v = dd * n1[c3][m] - pc[d] * n1[c3][m + 1];
                        if (c4 >= 0) v += t->c4[c] / z * (n1[c4][m] - n1[c4][m + 1]);
                        if (c5 >= 0) v += t->c5[c] / z * (n1[c5][m] - n1[c5][m + 1]);
                        n1[c][m] = v;
                    }
                }
                for (int c = 0; c < t->n; ++c) vp[cp.pi[c] + (size_t)np * cp.pj[c]] += n1[c][0];
            }
        }
    contract_cc(b, sp, s_c);
    contract_cc(b, tp, t_c);
    contract_cc(b, vp, n_c);
    free(tp); free(vp);
}

double or_nuclear_repulsion(const or_basis *b) {
    double e = 0.0;
    for (int i = 0; i < b->natom; ++i)
        for (int j = 0; j < b->natom; ++j) {
            if (i == j) continue;
            double dx = b->atomx[i] - b->atomx[j], dy = b->atomy[i] - b->atomy[j],
                   dz = b->atomz[i] - b->atomz[j];
            e += 0.5 * b->iatom[i] * b->iatom[j] / sqrt(dx * dx + dy * dy + dz * dz);
        }
    return e;
}

void or_decontract(const or_basis *b, const double *p_c, double *p_p) {
    int np = b->npgs, nc = b->ncgs;
    double *w = (double *)calloc((size_t)nc * np, sizeof(double));
    /* W2 = P_C CC ; P_P = CC^T W2 */
    for (int j = 0; j < np; ++j)
        for (int k = 0; k < nc; ++k) {
            double ccv = b->cc[k + (size_t)nc * j];
            if (ccv == 0.0) continue;
            for (int i = 0; i < nc; ++i) w[i + (size_t)nc * j] += p_c[i + (size_t)nc * k] * ccv;
        }
    for (int j = 0; j < np; ++j)
        for (int i = 0; i < np; ++i) {
            double s = 0.0;
            for (int k = 0; k < nc; ++k) s += b->cc[k + (size_t)nc * i] * w[k + (size_t)nc * j];
            p_p[i + (size_t)np * j] = s;
        }
    free(w);
}

void or_form_c2s(const or_basis *b, int spherical, double *c2s) {
    int nc = b->ncgs, np = b->npgs;
    memset(c2s, 0, sizeof(double) * nc * nc);
    for (int i = 0; i < nc; ++i) c2s[i + nc * i] = 1.0;
    if (!spherical) return;
    int *p2c = (int *)malloc(sizeof(int) * np);
    for (int p = 0; p < np; ++p) {
        p2c[p] = 0;
        for (int c = 0; c < nc; ++c) if (b->cc[c + (size_t)nc * p] != 0.0) p2c[p] = c;
    }
#define C2S(i, j) c2s[(i) + nc * (j)]
    /* d: 6 -> 5, src/pbasis.f90:69-80 (primitive order xx,yy,zz,...) */
    for (int p = 0; p < np; ++p) {
        int x = b->pax[p], y = b->pay[p], z = b->paz[p];
        if (x == 2 && y == 0 && z == 0) C2S(p2c[p], p2c[p]) = 0.0;
        else if (x == 0 && y == 2 && z == 0) {
            C2S(p2c[p], p2c[p]) = -sqrt(3.0) / 2.0;
            C2S(p2c[p], p2c[p - 1]) = sqrt(3.0) / 2.0;
        } else if (x == 0 && y == 0 && z == 2) {
            C2S(p2c[p], p2c[p - 2]) = -0.5;
            C2S(p2c[p], p2c[p - 1]) = -0.5;
        }
    }
#undef C2S
    free(p2c);
}

/* ------------------------------------------------------------------ */
/* Dense linear algebra for the oracle                                 */
/* ------------------------------------------------------------------ */
/* Cyclic Jacobi eigen-decomposition of a symmetric n x n matrix a
 * (destroyed).  Eigenvalues ascending in w, eigenvectors in columns of v.
 * Stands in for the EISPACK HTRIDI/TQL2 solver of src/pfock.f90:641-903;
 * any exact eigensolver gives the same invariant subspaces. */
static void jacobi_eig(int n, double *a, double *w, double *v) {
    for (int i = 0; i < n * n; ++i) v[i] = 0.0;
    for (int i = 0; i < n; ++i) v[i + n * i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, diag = 0.0;
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                if (i != j) off += a[i + n * j] * a[i + n * j];
                else diag += a[i + n * j] * a[i + n * j];
            }
        if (off <= 1e-30 * (diag + 1e-300)) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                double apq = a[p + n * q];
                if (fabs(apq) < 1e-300) continue;
                double app = a[p + n * p], aqq = a[q + n * q];
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {
                    double akp = a[k + n * p], akq = a[k + n * q];
                    a[k + n * p] = c * akp - s * akq;
                    a[k + n * q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    double apk = a[p + n * k], aqk = a[q + n * k];
                    a[p + n * k] = c * apk - s * aqk;
                    a[q + n * k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    double vkp = v[k + n * p], vkq = v[k + n * q];
                    v[k + n * p] = c * vkp - s * vkq;
                    v[k + n * q] = s * vkp + c * vkq;
                }
            }
    }
    for (int i = 0; i < n; ++i) w[i] = a[i + n * i];
    /* sort ascending */
    for (int i = 0; i < n; ++i) {
        int k = i;
        for (int j = i + 1; j < n; ++j) if (w[j] < w[k]) k = j;
        if (k != i) {
            double t = w[i]; w[i] = w[k]; w[k] = t;
            for (int r = 0; r < n; ++r) { t = v[r + n * i]; v[r + n * i] = v[r + n * k]; v[r + n * k] = t; }
        }
    }
}

/* Scaled partial-pivot LU solve with one step of iterative refinement: the
 * roles of LUDCMP/LUBKSB/MPROVE, src/pdensitymatrix.f90:385-512. */
static void lu_solve_refined(int n, const double *a0, const double *b0, double *x) {
    double *a = (double *)malloc(sizeof(double) * n * n);
    int *piv = (int *)malloc(sizeof(int) * n);
    memcpy(a, a0, sizeof(double) * n * n);
    double *sc = (double *)malloc(sizeof(double) * n);
    for (int i = 0; i < n; ++i) {
        double m = 0;
        for (int j = 0; j < n; ++j) if (fabs(a[i + n * j]) > m) m = fabs(a[i + n * j]);
        sc[i] = m > 0 ? 1.0 / m : 1.0;
    }
    for (int k = 0; k < n; ++k) {
        int p = k; double best = -1;
        for (int i = k; i < n; ++i) {
            double v = fabs(a[i + n * k]) * sc[i];
            if (v > best) { best = v; p = i; }
        }
        piv[k] = p;
        if (p != k) {
            for (int j = 0; j < n; ++j) { double t = a[k + n * j]; a[k + n * j] = a[p + n * j]; a[p + n * j] = t; }
            double t = sc[k]; sc[k] = sc[p]; sc[p] = t;
        }
        if (a[k + n * k] == 0.0) a[k + n * k] = 1e-20;
        for (int i = k + 1; i < n; ++i) {
            double f = a[i + n * k] / a[k + n * k];
            a[i + n * k] = f;
            for (int j = k + 1; j < n; ++j) a[i + n * j] -= f * a[k + n * j];
        }
    }
    double *r = (double *)malloc(sizeof(double) * n);
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 0) memcpy(r, b0, sizeof(double) * n);
        else
            for (int i = 0; i < n; ++i) {
                double s = -b0[i];
                for (int j = 0; j < n; ++j) s += a0[i + n * j] * x[j];
                r[i] = s;
            }
        for (int k = 0; k < n; ++k) if (piv[k] != k) { double t = r[k]; r[k] = r[piv[k]]; r[piv[k]] = t; }
        for (int i = 0; i < n; ++i) for (int j = 0; j < i; ++j) r[i] -= a[i + n * j] * r[j];
        for (int i = n - 1; i >= 0; --i) {
            for (int j = i + 1; j < n; ++j) r[i] -= a[i + n * j] * r[j];
            r[i] /= a[i + n * i];
        }
        if (pass == 0) memcpy(x, r, sizeof(double) * n);
        else for (int i = 0; i < n; ++i) x[i] -= r[i];
    }
    free(a); free(piv); free(sc); free(r);
}

static void matmul(int m, int n, int k, const double *a, int ta, const double *b, int tb,
                   double *c) {
    /* c (m x n) = op(a) op(b), column-major */
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < m; ++i) {
            double s = 0.0;
            for (int l = 0; l < k; ++l) {
                double x = ta ? a[l + k * i] : a[i + m * l];
                double y = tb ? b[j + n * l] : b[l + k * j];
                s += x * y;
            }
            c[i + m * j] = s;
        }
}

/* ------------------------------------------------------------------ */
/* SCF: src/pselfconsistentfield.f90:1-162, src/pfock.f90:1-506,       */
/*      src/pdensitymatrix.f90:83-339                                  */
/* ------------------------------------------------------------------ */
static void jk_from_tensor(int nc, const double *feri, const double *p, double *j, double *k) {
    size_t n2 = (size_t)nc * nc;
    for (size_t i = 0; i < n2; ++i) { j[i] = 0.0; k[i] = 0.0; }
    for (int l = 0; l < nc; ++l)
        for (int kk = 0; kk < nc; ++kk) {
            double pkl = p[kk + (size_t)nc * l];
            const double *blk = feri + n2 * (kk + (size_t)nc * l);
            for (size_t ij = 0; ij < n2; ++ij) j[ij] += blk[ij] * pkl;
        }
    /* K(i,l) = sum_jk (ij|kl) P(k,j) */
    for (int l = 0; l < nc; ++l)
        for (int kk = 0; kk < nc; ++kk) {
            const double *blk = feri + n2 * (kk + (size_t)nc * l);
            for (int jj = 0; jj < nc; ++jj) {
                double pkj = p[kk + (size_t)nc * jj];
                if (pkj == 0.0) continue;
                for (int i = 0; i < nc; ++i) k[i + (size_t)nc * l] += blk[i + (size_t)nc * jj] * pkj;
            }
        }
}

int or_scf(const or_basis *b, const or_scf_opts *o, double *p_c_io, double *eps_out,
           double *co_out, double *trace, or_scf_result *res) {
    int nc = b->ncgs, np = b->npgs;
    size_t n2 = (size_t)nc * nc;
    double *s_p = (double *)malloc(sizeof(double) * np * np);
    double *s_c = (double *)malloc(sizeof(double) * n2), *t_c = (double *)malloc(sizeof(double) * n2);
    double *n_c = (double *)malloc(sizeof(double) * n2);
    or_one_electron(b, s_p, s_c, t_c, n_c);
    double enuc = or_nuclear_repulsion(b);
    double *c2s = (double *)malloc(sizeof(double) * n2);
    or_form_c2s(b, o->spherical, c2s);
    double *feri = NULL;
    if (!o->direct) {
        feri = (double *)malloc(sizeof(double) * n2 * n2);
        or_eri_tensor(b, s_p, feri);
    }
    /* electrons and orbital classes, src/porbitals.f90:15-19 */
    int nel = 0;
    for (int i = 0; i < b->natom; ++i) nel += b->iatom[i];
    nel -= o->charge;
    int iocc = nel / 2;

    double *p = (double *)calloc(n2, sizeof(double));
    if (o->guess == 1) memcpy(p, p_c_io, sizeof(double) * n2);
    else {
        /* GUESS_DENSITYMATRIX, src/pdensitymatrix.f90:83-127 */
        for (int a = 0; a < b->natom; ++a) {
            int L = 0;
            for (int c = 0; c < nc; ++c) {
                int dummy = 1;
                for (int q = 0; q < np; ++q) if (b->cc[c + (size_t)nc * q] != 0.0) { dummy = 0; break; }
                if (!dummy && b->cgsx[c] == b->atomx[a] && b->cgsy[c] == b->atomy[a] &&
                    b->cgsz[c] == b->atomz[a]) L++;
            }
            double v = (double)b->iatom[a] / (double)L;
            for (int c = 0; c < nc; ++c) {
                int dummy = 1;
                for (int q = 0; q < np; ++q) if (b->cc[c + (size_t)nc * q] != 0.0) { dummy = 0; break; }
                if (!dummy && b->cgsx[c] == b->atomx[a] && b->cgsy[c] == b->atomy[a] &&
                    b->cgsz[c] == b->atomz[a]) p[c + (size_t)nc * c] = v;
            }
        }
    }
    /* canonical orthogonaliser from C2S S C2S^T (src/pfock.f90:236-281) */
    double *tmp = (double *)malloc(sizeof(double) * n2), *tmp2 = (double *)malloc(sizeof(double) * n2);
    double *ws = (double *)malloc(sizeof(double) * n2), *w = (double *)malloc(sizeof(double) * nc);
    double *u = (double *)malloc(sizeof(double) * n2);
    matmul(nc, nc, nc, c2s, 0, s_c, 0, tmp);
    matmul(nc, nc, nc, tmp, 0, c2s, 1, ws);
    jacobi_eig(nc, ws, w, u);
    int nind = 0;
    double *x = (double *)malloc(sizeof(double) * n2);
    for (int j = 0; j < nc; ++j)
        if (w[j] > o->lindep) {
            for (int k = 0; k < nc; ++k) x[k + (size_t)nc * nind] = u[k + (size_t)nc * j] / sqrt(w[j]);
            nind++;
        }
    int maxit = o->scfcycles;
    double *hist_p = (double *)malloc(sizeof(double) * n2 * (maxit + 1));
    double *hist_d = (double *)malloc(sizeof(double) * n2 * (maxit + 1));
    double *pulay = (double *)calloc((size_t)(maxit + 2) * (maxit + 2), sizeof(double));
    double *pout = (double *)malloc(sizeof(double) * n2);
    double *jc = (double *)malloc(sizeof(double) * n2), *kc = (double *)malloc(sizeof(double) * n2);
    double *f = (double *)malloc(sizeof(double) * n2);
    double *pp = (double *)malloc(sizeof(double) * np * np);
    double *eps = (double *)malloc(sizeof(double) * nc);
    double *co = (double *)calloc(n2, sizeof(double));
    double temphf = o->temphf > 0 ? o->temphf : 1.0e-4;
    int converged = 0, iscf;
    double etot = 0, egrand = 0, ek = 0, ev = 0, ej = 0, ex = 0, fermi = 0;
    for (iscf = 1; iscf <= maxit; ++iscf) {
        /* J, K */
        if (o->direct) {
            or_decontract(b, p, pp);
            or_fourindex_eri(b, s_p, pp, o->inttol, 0, jc, kc, NULL, NULL);
        } else {
            jk_from_tensor(nc, feri, p, jc, kc);
        }
        /* FOCK_BUILD, src/pfock.f90:44 and :186-337 at k = 0 */
        for (size_t i = 0; i < n2; ++i) f[i] = t_c[i] + n_c[i] + jc[i] - 0.5 * kc[i] * o->hfexchange;
        for (int j = 0; j < nc; ++j)
            for (int k = 0; k < nc; ++k) {
                double av = 0.5 * (f[k + (size_t)nc * j] + f[j + (size_t)nc * k]);
                f[k + (size_t)nc * j] = av; f[j + (size_t)nc * k] = av;
            }
        matmul(nc, nc, nc, c2s, 0, f, 0, tmp);
        matmul(nc, nc, nc, tmp, 0, c2s, 1, tmp2); /* WF = C2S F C2S^T */
        /* F' = X^T WF X */
        double *fx = (double *)malloc(sizeof(double) * nc * nind);
        double *fp = (double *)malloc(sizeof(double) * nind * nind);
        double *cv = (double *)malloc(sizeof(double) * nind * nind);
        matmul(nc, nind, nc, tmp2, 0, x, 0, fx);
        matmul(nind, nind, nc, x, 1, fx, 0, fp);
        jacobi_eig(nind, fp, eps, cv);
        for (int j = nind; j < nc; ++j) eps[j] = 1.0e99;
        /* CO = C2S^T (X C') */
        double *wf = (double *)malloc(sizeof(double) * nc * nind);
        matmul(nc, nind, nind, x, 0, cv, 0, wf);
        memset(co, 0, sizeof(double) * n2);
        matmul(nc, nind, nc, c2s, 1, wf, 0, co);
        free(fx); free(fp); free(cv); free(wf);
        /* Fermi level by bisection, src/pfock.f90:339-423 */
        double mine = 1e99, maxe = -1e99;
        for (int j = 0; j < nc; ++j) { if (eps[j] > maxe) maxe = eps[j]; if (eps[j] < mine) mine = eps[j]; }
        double f1 = mine, f2 = maxe, f12 = 0.5 * (mine + maxe);
        for (int it = 0; it < 100000; ++it) {
            double n1 = 0, nn2 = 0, n12 = 0;
            for (int j = 0; j < nc; ++j) {
                n1 += 2.0 / (exp((eps[j] - f1) / BOLTZMANN / temphf) + 1.0);
                nn2 += 2.0 / (exp((eps[j] - f2) / BOLTZMANN / temphf) + 1.0);
                n12 += 2.0 / (exp((eps[j] - f12) / BOLTZMANN / temphf) + 1.0);
            }
            n1 -= nel; nn2 -= nel; n12 -= nel;
            if (fabs(n12) < 1e-10) { fermi = f12; break; }
            else if (n1 > 0 && nn2 > 0) { f2 = f1; f1 = f1 - 0.1; }
            else if (n1 < 0 && nn2 < 0) { f1 = f2; f2 = f2 + 0.1; }
            else if (n1 < 0 && n12 >= 0) f2 = f12;
            else if (n12 <= 0 && nn2 > 0) f1 = f12;
            else break;
            f12 = 0.5 * (f1 + f2);
        }
        /* new density, src/pfock.f90:471-485 */
        for (size_t i = 0; i < n2; ++i) pout[i] = 0.0;
        for (int l = 0; l < nc; ++l) {
            double occ = 2.0 / (exp((eps[l] - fermi) / BOLTZMANN / temphf) + 1.0);
            if (occ == 0.0) continue;
            for (int j = 0; j < nc; ++j)
                for (int k = 0; k < nc; ++k)
                    pout[k + (size_t)nc * j] += occ * co[k + (size_t)nc * l] * co[j + (size_t)nc * l];
        }
        /* STORE_DENSITYMATRIX, src/pdensitymatrix.f90:131-208 */
        double dchg = 0.0;
        double *dcur = hist_d + n2 * (iscf - 1), *pcur = hist_p + n2 * (iscf - 1);
        for (size_t i = 0; i < n2; ++i) { dcur[i] = pout[i] - p[i]; pcur[i] = p[i]; dchg += dcur[i] * dcur[i]; }
        int ld = maxit + 2;
        pulay[iscf + ld * iscf] = dchg;
        for (int i = 1; i < iscf; ++i) {
            double s = 0.0;
            const double *di = hist_d + n2 * (i - 1);
            for (size_t q = 0; q < n2; ++q) s += di[q] * dcur[q];
            pulay[i + ld * iscf] = s; pulay[iscf + ld * i] = s;
        }
        dchg = sqrt(dchg / (double)n2);
        /* DIIS or relaxation, src/pselfconsistentfield.f90:59-65 */
        if (o->diis != 0 && iscf >= o->diis && ((iscf - o->diis) % 2) == 0) {
            int nt = iscf < 8 ? iscf : 8;
            int n = nt + 1;
            double *bm = (double *)calloc(n * n, sizeof(double)), *rhs = (double *)calloc(n, sizeof(double));
            double *c = (double *)calloc(n, sizeof(double));
            for (int i = 0; i < nt; ++i) {
                for (int j = 0; j < nt; ++j) bm[j + n * i] = pulay[(iscf - j) + ld * (iscf - i)];
                bm[i + n * nt] = -1.0; bm[nt + n * i] = -1.0;
            }
            rhs[nt] = -1.0;
            lu_solve_refined(n, bm, rhs, c);
            for (size_t q = 0; q < n2; ++q) {
                double s = 0.0;
                for (int i = 1; i <= iscf; ++i)
                    if (iscf - i + 1 <= nt) s += c[iscf - i] * hist_p[n2 * (i - 1) + q];
                p[q] = s;
            }
            free(bm); free(rhs); free(c);
        } else {
            for (size_t q = 0; q < n2; ++q) p[q] = o->scfrelax * pout[q] + (1.0 - o->scfrelax) * p[q];
        }
        /* ENERGY_CALC, src/pselfconsistentfield.f90:487-648 */
        ek = ev = ej = ex = 0.0;
        for (size_t q = 0; q < n2; ++q) {
            ek += p[q] * t_c[q]; ev += p[q] * n_c[q];
            ej += p[q] * 0.5 * jc[q]; ex -= p[q] * 0.25 * kc[q] * o->hfexchange;
        }
        etot = enuc + ek + ev + ej + ex;
        egrand = etot;
        if (o->temphf > 0) {
            egrand += -2.0 * iocc * fermi;
            double sb = 0.0;
            for (int j = 0; j < nind; ++j) {
                double fm = 1.0 / (exp((eps[j] - fermi) / BOLTZMANN / temphf) + 1.0), fp2 = 1.0 - fm;
                if (fm > 0 && fp2 > 0) sb += 2.0 * (fm * log(fm) + fp2 * log(fp2)) * BOLTZMANN * temphf;
            }
            egrand += sb;
        }
        if (trace) { trace[2 * (iscf - 1)] = dchg; trace[2 * (iscf - 1) + 1] = etot; }
        if (dchg < o->scfconv) { converged = 1; break; }
    }
    if (iscf > maxit) iscf = maxit;
    memcpy(p_c_io, p, sizeof(double) * n2);
    if (eps_out) memcpy(eps_out, eps, sizeof(double) * nc);
    if (co_out) memcpy(co_out, co, sizeof(double) * n2);
    res->e_total = etot; res->e_nuc = enuc; res->e_kin = ek; res->e_nuc_att = ev;
    res->e_coul = ej; res->e_hfx = ex; res->fermi = fermi; res->e_grand = egrand;
    res->iterations = iscf; res->converged = converged; res->n_indep = nind; res->iocc = iocc;
    free(s_p); free(s_c); free(t_c); free(n_c); free(c2s); free(feri); free(p);
    free(tmp); free(tmp2); free(ws); free(w); free(u); free(x); free(hist_p); free(hist_d);
    free(pulay); free(pout); free(jc); free(kc); free(f); free(pp); free(eps); free(co);
    return converged ? 0 : -1;
}

/* ------------------------------------------------------------------ */
/* MP2: src/pmp2energy.f90:395-444 at a single k-point                 */
/* ------------------------------------------------------------------ */
void or_mp2(int nc, const double *feri, const double *co, const double *eps, int icore,
            int iocc, int nall, int ivirtcore, double *e2s, double *e2t) {
    int no = iocc - icore, nv = nall - ivirtcore - iocc;
    size_t n2 = (size_t)nc * nc;
    /* (ia|jb) by four quarter transforms over the dense tensor */
    double *t1 = (double *)calloc(n2 * nc * (size_t)nv, sizeof(double)); /* (pq|r b) */
    for (int b = 0; b < nv; ++b)
        for (int s = 0; s < nc; ++s) {
            double c = co[s + (size_t)nc * (iocc + b)];
            if (c == 0.0) continue;
            const double *blk = feri + n2 * nc * (size_t)s;
            double *dst = t1 + n2 * nc * (size_t)b;
            for (size_t q = 0; q < n2 * nc; ++q) dst[q] += c * blk[q];
        }
    double *t2 = (double *)calloc(n2 * (size_t)no * nv, sizeof(double)); /* (pq|j b) */
    for (int b = 0; b < nv; ++b)
        for (int j = 0; j < no; ++j)
            for (int r = 0; r < nc; ++r) {
                double c = co[r + (size_t)nc * (icore + j)];
                if (c == 0.0) continue;
                const double *src = t1 + n2 * nc * (size_t)b + n2 * r;
                double *dst = t2 + n2 * ((size_t)b * no + j);
                for (size_t q = 0; q < n2; ++q) dst[q] += c * src[q];
            }
    free(t1);
    double *t3 = (double *)calloc((size_t)nc * nv * no * nv, sizeof(double)); /* (p a|j b) */
    for (int jb = 0; jb < no * nv; ++jb)
        for (int a = 0; a < nv; ++a)
            for (int q = 0; q < nc; ++q) {
                double c = co[q + (size_t)nc * (iocc + a)];
                if (c == 0.0) continue;
                const double *src = t2 + n2 * jb + (size_t)nc * q;
                double *dst = t3 + ((size_t)jb * nv + a) * nc;
                for (int pp = 0; pp < nc; ++pp) dst[pp] += c * src[pp];
            }
    free(t2);
    /* T(i,a,j,b) */
    double *tt = (double *)calloc((size_t)no * nv * no * nv, sizeof(double));
    for (int jb = 0; jb < no * nv; ++jb)
        for (int a = 0; a < nv; ++a)
            for (int i = 0; i < no; ++i) {
                const double *src = t3 + ((size_t)jb * nv + a) * nc;
                double s = 0.0;
                for (int pp = 0; pp < nc; ++pp) s += co[pp + (size_t)nc * (icore + i)] * src[pp];
                tt[i + (size_t)no * (a + (size_t)nv * jb)] = s;
            }
    free(t3);
    double es = 0, et = 0;
#define TIAJB(i, a, j, b) tt[(i) + (size_t)no * ((a) + (size_t)nv * ((j) + (size_t)no * (b)))]
    for (int b = 0; b < nv; ++b)
        for (int j = 0; j < no; ++j)
            for (int a = 0; a < nv; ++a)
                for (int i = 0; i < no; ++i) {
                    double t = TIAJB(i, a, j, b), tp = TIAJB(j, a, i, b);
                    double den = eps[icore + i] + eps[icore + j] - eps[iocc + a] - eps[iocc + b];
                    es += (0.5 * t * t + 0.5 * t * tp) / den;
                    et += (1.5 * t * t - 1.5 * t * tp) / den;
                }
#undef TIAJB
    free(tt);
    *e2s = es; *e2t = et;
}
