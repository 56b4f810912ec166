/*
 * oracle.h -- CPU restatement of polymer's direct two-electron path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the MI355X
 * product (libpolyeri) is compared against.  Only tests/, the smoke() entry
 * and the cpu_baseline leg of bench.py may load it.  Nothing in the product
 * links it or falls back to it.
 *
 * Every function restates one routine of the reference (polymer, Fortran 90)
 * and cites the file:line it follows.  Matrices use the reference's
 * column-major layout: element (i,j) of an n x m matrix is a[i + n*j].
 * Primitive/contracted/shell indices are 0-based here (1-based in Fortran).
 */
#ifndef POLYMER_ORACLE_H
#define POLYMER_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define OR_LMAX 3            /* BASIS_LMAX, src/pmodules.f90:69            */
#define OR_SLOT 5            /* 0:BASIS_LMAX+1 per Cartesian direction      */

typedef struct or_basis {
    int natom;
    int *iatom;                          /* atomic numbers                  */
    double *atomx, *atomy, *atomz;       /* bohr                            */
    int npgs, ncgs, npshell;
    int *pax, *pay, *paz, *pgs_atom;     /* per primitive                   */
    double *pgsx, *pgsy, *pgsz, *zt;     /* per primitive                   */
    double *cc;                          /* ncgs x npgs, column-major       */
    double *cgsx, *cgsy, *cgsz;          /* per contracted function         */
    int *p_shell;                        /* npshell x 5x5x5, -1 = empty     */
    int *p_shell_ang;                    /* per shell                       */
} or_basis;

/* src/pread.f90:584-1482 (READ3).  Returns 0 on success, negative on error. */
int or_basis_build(const char *path, int natom, const int *iatom,
                   const double *xyz_bohr, or_basis **out);
void or_basis_free(or_basis *b);
void or_basis_dims(const or_basis *b, int *npgs, int *ncgs, int *npshell);
void or_basis_export(const or_basis *b, int *pax, int *pay, int *paz,
                     int *pgs_atom, double *pgsx, double *pgsy, double *pgsz,
                     double *zt, double *cc, int *p_shell, int *p_shell_ang,
                     double *cgsx, double *cgsy, double *cgsz);

/* src/pbasis.f90:1-27 (FMT_PRELIMINARY): IGAMMA(0:1700,0:27), column-major. */
void or_boys_table(double *igamma);
/* src/pfourindexdirect.f90:543-553: Taylor-table / asymptotic F_m(T).      */
double or_boys(double t, int m);

/* src/ponee.f90:1-414 and :418-728 for a molecule (no cells).
 * S_P (npgs^2, unnormalised primitive overlap), S_C, T_C, N_C (ncgs^2).   */
void or_one_electron(const or_basis *b, double *s_p, double *s_c,
                     double *t_c, double *n_c);

/* src/pselfconsistentfield.f90:652-704 for a molecule.                     */
double or_nuclear_repulsion(const or_basis *b);

/* src/pfourindexdirect.f90:1-814 for a molecule: J_C (C_C) and K_C (X_C)
 * from the primitive density P_P.  recompute_kl != 0 keeps the reference's
 * per-(I,J) recomputation of the KL pair table (:469-525).  Returns the
 * number of composite primitive shell quartets that passed the screen;
 * *real_quartets receives the subset whose four shells are not ghost-
 * anchored, and flop_model the algorithmic FLOP count of those. */
long long or_fourindex_eri(const or_basis *b, const double *s_p,
                           const double *p_p, double inttol, int recompute_kl,
                           double *c_c, double *x_c, long long *real_quartets,
                           double *flop_model);

/* Same engine limited to bra shell pairs (I,J) with I*npshell+J in
 * [ij_begin, ij_end): used by the bounded CPU baseline sample. */
long long or_fourindex_eri_range(const or_basis *b, const double *s_p,
                                 const double *p_p, double inttol,
                                 int recompute_kl, long long ij_begin,
                                 long long ij_end, double *c_p, double *x_p,
                                 long long *real_quartets, double *flop_model);

/* Same engine over an explicit list of bra rows (I*npshell+J); the KL pair
 * table and its bounds are built once (setup_seconds) and every row's work,
 * the per-row KL rebuild included, is timed together (row_seconds). */
long long or_fourindex_eri_rows(const or_basis *b, const double *s_p, const double *p_p,
                                double inttol, int recompute_kl, const long long *rows,
                                long long nrows, double *c_p, double *x_p,
                                long long *real_quartets, double *flop_model,
                                double *setup_seconds, double *row_seconds);
/* The same over nthreads OpenMP threads, each with its own tables and
 * accumulators; row_seconds is the wall time of all rows. */
long long or_fourindex_eri_rows_mt(const or_basis *b, const double *p_p, double inttol, int recompute_kl,
                                   const long long *rows, long long nrows, int nthreads, double *c_p,
                                   double *x_p, long long *real_quartets, double *setup_seconds,
                                   double *row_seconds);

/* Contracted ERI tensor FERI(i,j,k,l) = sum CC.CC.CC.CC (pq|rs), the
 * quantity src/pfourindexstore.f90:704-708 accumulates (no screening). */
/* Periodic direct J/K (src/pfourindexdirect.f90:222-750, HELIX = 0, axis x):
 * J(Q1)[i,j] = sum_{Q2,Q3} (i_0 j_Q1 | k_Q3 l_Q3+Q2) P(Q2)[k,l] and
 * K(Q3+Q2)[i,l] += (i_0 j_Q1 | k_Q3 l_Q3+Q2) P(Q1-Q3)[k,j] when both cells
 * lie in [-cel1, cel1]; Q1, Q2 in [-rc1, rc1], Q3 in [-cel2, cel2], no
 * permutational symmetry, IJIJ*KLKL < inttol skipped.  Primitive-basis
 * arrays hold 2 cel1 + 1 npgs^2 blocks (cell Q at Q + cel1); the _c form
 * takes and returns contracted ncgs^2 blocks.  Returns quartets computed. */
long long or_fourindex_eri_cells(const or_basis *b, double period, int rc1, int cel1, int cel2,
                                 const double *p_p, double inttol, double *c_p, double *x_p);
long long or_fourindex_eri_cells_c(const or_basis *b, double period, int rc1, int cel1, int cel2,
                                   const double *p_c, double inttol, double *c_c, double *x_c);
void or_eri_tensor(const or_basis *b, const double *s_p, double *feri);
void or_eri_tensor_cells(const or_basis *b, double period, int rc1, int cel2, double *T);

/* src/pdensitymatrix.f90:343-381 (DECONTRACT): P_P = CC^T P_C CC.         */
void or_decontract(const or_basis *b, const double *p_c, double *p_p);

/* src/pbasis.f90:31-130 (FORM_C2S).                                         */
void or_form_c2s(const or_basis *b, int spherical, double *c2s);

/* SCF options and results (src/pselfconsistentfield.f90:1-162).            */
typedef struct or_scf_opts {
    double inttol;      /* DOPTN(12)                                         */
    double scfconv;     /* DOPTN(14), RMS density change                     */
    double scfrelax;    /* DOPTN(15)                                         */
    double hfexchange;  /* DOPTN(19)                                         */
    double lindep;      /* DOPTN(85)                                         */
    double temphf;      /* DOPTN(108), kelvin; 0 -> 1e-4 K                   */
    int scfcycles;      /* IOPTN(13)                                         */
    int diis;           /* IOPTN(16)                                         */
    int charge;         /* IOPTN(5)                                          */
    int spherical;      /* LOPTN(69)                                         */
    int direct;         /* LOPTN(25): 1 = FOURINDEX_ERI every iteration,
                           0 = contracted ERIs once (RESTORE path)           */
    int guess;          /* 0 = GUESS_DENSITYMATRIX, 1 = caller's P_C        */
} or_scf_opts;

typedef struct or_scf_result {
    double e_total, e_nuc, e_kin, e_nuc_att, e_coul, e_hfx;
    double fermi;
    /* TEMPHF > 0: e_total + chemical-potential and entropy terms as in the
     * source's ENERGY_CALC (src/pselfconsistentfield.f90:596-614).  The
     * shipped finite-T output predates those terms and prints e_total. */
    double e_grand;
    int iterations;
    int converged;
    int n_indep;        /* IALL after LINDEP                                */
    int iocc;
} or_scf_result;

/* Runs the SCF.  p_c (ncgs^2) is the guess on input when opts->guess == 1
 * and the converged density on output; eps (ncgs) and co (ncgs^2, the
 * Cartesian MO coefficients, columns = orbitals) are outputs (may be NULL).
 * trace (length >= 2*scfcycles, may be NULL) receives per-iteration
 * (density change, energy).  Returns 0 on success, -1 if not converged. */
int or_scf(const or_basis *b, const or_scf_opts *opts, double *p_c,
           double *eps, double *co, double *trace, or_scf_result *res);

/* Molecular closed-shell MP2 from the contracted ERI tensor and Cartesian
 * MO coefficients (src/pmp2energy.f90:395-444 at one k-point).  Orbitals
 * icore..iocc-1 are active occupied, iocc..nall-1-ivirtcore virtual. */
void or_mp2(int ncgs, const double *feri, const double *co, const double *eps,
            int icore, int iocc, int nall, int ivirtcore, double *e2s,
            double *e2t);

/* Algorithmic FLOP model per primitive quartet of real shells. */
double or_flop_model(int la, int lb, int lc, int ld);

#ifdef __cplusplus
}
#endif
#endif
