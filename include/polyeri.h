/*
 * polyeri.h -- C ABI of libpolyeri, the MI355X (gfx950) direct two-electron
 * path for polymer: shell-quartet ERIs contracted on the fly into Coulomb /
 * exchange matrices, the SCF loop around them, and the MP2 AO->MO transform.
 *
 * Plain pointers and sizes only.  Matrices use the reference's Fortran
 * column-major layout over the contracted functions INCLUDING ghost
 * functions (NCGS of src/pread.f90 READ3): element (i,j) is a[i + ncgs*j].
 * Every entry point returns 0 on success or a negative POLYERI_E_* code;
 * polyeri_last_error() then holds a message (thread-local).
 */
#ifndef POLYERI_H
#define POLYERI_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define POLYERI_OK 0
#define POLYERI_E_ARG (-1)       /* invalid argument / layout              */
#define POLYERI_E_BASIS (-2)     /* basis file or shell structure          */
#define POLYERI_E_HIP (-3)       /* HIP runtime / device error             */
#define POLYERI_E_NOMEM (-4)     /* device allocation failed               */
#define POLYERI_E_NOTCONV (-5)   /* SCF CONVERGENCE NOT MET                */
#define POLYERI_E_NONFINITE (-6) /* non-finite result                      */
#define POLYERI_E_LIB (-7)       /* rocBLAS / rocSOLVER error              */

typedef struct polyeri_ctx polyeri_ctx;

/* Precision modes of the J/K build. */
#define POLYERI_FP64 0
#define POLYERI_MIXED 1   /* FP32 for quartets whose Schwarz bound < mixed_cut */

typedef struct polyeri_opts {
    double schwarz_tol;  /* drop quartets with Q_ab Q_cd < tol (INTTOL)       */
    double prim_tol;     /* drop primitive pairs with |c_a c_b| e^{-mu R^2} < */
    int precision;       /* POLYERI_FP64 or POLYERI_MIXED                      */
    double mixed_cut;    /* Schwarz bound below which quartets run in FP32     */
    int device;          /* HIP device ordinal                                 */
} polyeri_opts;

/* Reads a Gaussian-format basis-set file for the given atoms and returns the
 * reference's arrays (READ3, src/pread.f90:584-1482).  Call once with all
 * array pointers NULL to get the sizes, then with arrays of those sizes.
 * xyz: 3*natom bohr.  cc: ncgs*npgs column-major.  p_shell: npshell*125
 * (slot lx*25+ly*5+lz, -1 = empty). */
int polyeri_basis_read(const char *path, int natom, const int *iatom, const double *xyz,
                       int *npgs, int *ncgs, int *npshell, int *pax, int *pay, int *paz,
                       int *pgs_atom, double *pgsx, double *pgsy, double *pgsz, double *zt,
                       double *cc, double *cgsx, double *cgsy, double *cgsz, int *p_shell,
                       int *p_shell_ang);

/* Builds a context from the reference's module arrays (the inputs
 * FOURINDEX_ERI reads from BASISSET / STRUCTURE, src/pfourindexdirect.f90:
 * 1-99, src/pmodules.f90:64-81): primitive angular exponents, centres,
 * exponents, the contraction matrix CC and the atoms.  Derives contracted
 * shells, primitive pairs and Schwarz bounds on the device.  opts may be
 * NULL (FP64, tol 1e-12). */
int polyeri_ctx_create(int npgs, int ncgs, const int *pax, const int *pay, const int *paz,
                       const double *pgsx, const double *pgsy, const double *pgsz,
                       const double *zt, const double *cc, int natom, const int *iatom,
                       const double *xyz, const polyeri_opts *opts, polyeri_ctx **out);

/* Periodic context (the reference's 1-D polymer path with HELIX = 0): the
 * basis arrays of one unit cell, translation period along x (bohr), CUTOFF1
 * (density / Fock cells -cel1..cel1) and CUTOFF2 (Coulomb lattice sum
 * -cel2..cel2), as CEL1X / CEL2X of src/pfourindexdirect.f90:222-750.  The
 * pair cells are reduced to the overlap range as REDUCED_CEL1X
 * (src/ponee.f90:359-371). */
int polyeri_ctx_create_periodic(int npgs, int ncgs, const int *pax, const int *pay, const int *paz,
                                const double *pgsx, const double *pgsy, const double *pgsz, const double *zt,
                                const double *cc, int natom, const int *iatom, const double *xyz, double period,
                                int cel1, int cel2, const polyeri_opts *opts, polyeri_ctx **ctx);

/* Periodic CALL FOURINDEX_ERI (src/pselfconsistentfield.f90:42 with
 * cells): p_c, c_c, x_c hold 2 cel1 + 1 column-major ncgs x ncgs blocks,
 * P_C(:,:,-CEL1X..CEL1X) in order; C_C(i,j,Q1) = sum (i_0 j_Q1|k_Q3 l_Q3+Q2)
 * P_C(k,l,Q2), X_C(i,l,Q3+Q2) += (i_0 j_Q1|k_Q3 l_Q3+Q2) P_C(k,j,Q1-Q3)
 * (src/pfourindexdirect.f90:691-704, 765-781). */
int polyeri_jk_build_cells(polyeri_ctx *ctx, const double *p_c, double *c_c, double *x_c);

void polyeri_ctx_destroy(polyeri_ctx *ctx);

/* Drop-in for CALL FOURINDEX_ERI (src/pselfconsistentfield.f90:42): from
 * the contracted density P_C (ncgs x ncgs, host) build the Coulomb matrix
 * C_C = J and the exchange matrix X_C = K (src/pfourindexdirect.f90:773,
 * 777), host arrays, ghost rows/columns zero. */
int polyeri_jk_build(polyeri_ctx *ctx, const double *p_c, double *c_c, double *x_c);

/* Same on device pointers (ncgs x ncgs each), enqueued on the context's
 * stream; returns after the kernels are enqueued.  polyeri_sync waits. */
int polyeri_jk_build_device(polyeri_ctx *ctx, const double *d_p_c, double *d_c_c, double *d_x_c);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  The same
 * molecular J/K build as polyeri_jk_build_device, but with the density
 * staging and the J/K accumulation copies in a fixed, library-wide device
 * workspace of 2^22 doubles (32 MiB per device, held until the process
 * ends) and eager launches: it calls no allocator.  Limit: nbf <= 1182
 * (larger contexts get POLYERI_E_ARG).  Every call on a device shares that
 * device's workspace, so calls that overlap on one device race. */
int polyeri_jk_build_device_noalloc(polyeri_ctx *ctx, const double *d_p_c, double *d_c_c, double *d_x_c);

/* polyeri_jk_build with density screening (no counterpart in the reference,
 * whose screen is IJIJ*KLKL < INTTOL only, src/pfourindexdirect.f90:529):
 * quartets whose Schwarz bound times the largest |P| over the six density
 * blocks they read is below dens_tol are skipped.  Used with density
 * differences in incremental Fock builds; dens_tol = 0 is polyeri_jk_build. */
int polyeri_jk_build_screened(polyeri_ctx *ctx, const double *p_c, double dens_tol, double *c_c,
                              double *x_c);
int polyeri_sync(polyeri_ctx *ctx);

/* Device memory for callers without their own allocator (e.g. a Fortran
 * driver): allocation on the context's device, and blocking copies
 * (to_device = 1: host -> device, 0: device -> host). */
int polyeri_device_alloc(polyeri_ctx *ctx, size_t bytes, void **ptr);
int polyeri_device_free(polyeri_ctx *ctx, void *ptr);
int polyeri_memcpy(polyeri_ctx *ctx, void *dst, const void *src, size_t bytes, int to_device);
/* Device-to-device copy with the library's own streaming kernel (bytes a
 * multiple of 16); blocking.  A known-size reference for hardware counters. */
int polyeri_device_copy(polyeri_ctx *ctx, void *dst, const void *src, size_t bytes);

/* Drop-in for FOURINDEX_ERI followed by the Fock assembly of FOCK_BUILD
 * (src/pfock.f90:44, with the Hermitian average of :224-233):
 * F_C = T_C + N_C + C_C - hfexchange/2 X_C on device pointers (ncgs^2
 * each), enqueued on the context's stream. */
int polyeri_fock_build_device(polyeri_ctx *ctx, const double *d_p_c, double hfexchange,
                              double *d_f_c, double *d_c_c, double *d_x_c);

/* One-electron matrices S_C, T_C, N_C (src/ponee.f90:1-414, :418-728),
 * host, ncgs x ncgs; the nuclear repulsion energy
 * (src/pselfconsistentfield.f90:652-704). */
int polyeri_one_electron(polyeri_ctx *ctx, double *s_c, double *t_c, double *n_c,
                         double *e_nuc);

/* Contracted ERI tensor FERI(i,j,k,l) = feri[i + n(j + n(k + n l))], n =
 * ncgs, the values src/pfourindexstore.f90:704-708 accumulates; for tests
 * and small systems (ncgs^4 doubles on host and device). */
int polyeri_eri_tensor(polyeri_ctx *ctx, double *feri);

/* The lattice integrals of a periodic context (src/pfourindexstore.f90:
 * 258-747): (I_0 J_Q1 | K_Q3 L_Q3+Q2) for Q1, Q2 in [-rc1, rc1] (rc1 =
 * polyeri_reduced_cel1) and Q3 in [-cel2, cel2], as (2 rc1 + 1)^2 (2 cel2 + 1)
 * column-major ncgs^4 FERI blocks, block ((Q1 + rc1)(2 rc1 + 1) + Q2 + rc1)
 * (2 cel2 + 1) + Q3 + cel2; quartets below the Schwarz threshold are zero. */
int polyeri_eri_tensor_cells(polyeri_ctx *ctx, double *t_c);
/* REDUCED_CEL1X of a periodic context (src/ponee.f90:359-371): the last
 * cell whose contracted overlap with cell 0 exceeds 1e-15; < 0 on error. */
int polyeri_reduced_cel1(const polyeri_ctx *ctx);

/* STORE_FOURINDEX_ERI's disk format (src/pfourindexstore.f90:1-801): the
 * contracted ERIs with |(ij|kl)| > inttol written to path as the reference's
 * .ao.NNNNN records (Fortran unformatted, 1000 entries of packed cells,
 * packed 1-based ijkl and the value per record), NCGS < 256.  Periodic
 * contexts write every cell triple (Q1, Q2 in [-rc1, rc1], Q3 in
 * [-cel2, cel2]) in the reference's loop order.  stored = the number of
 * integrals written (the count the reference prints, e.g.
 * data/H2O_CCSD_ccpVDZ_out:212). */
int polyeri_store_fourindex_eri(polyeri_ctx *ctx, const char *path, double inttol, long long *stored);
/* RESTORE_FOURINDEX_ERI's reader of the same format into FERI (ncgs^4,
 * column-major), molecular (cell 0) records. */
int polyeri_restore_fourindex_eri(const char *path, int ncgs, double *feri, long long *restored);
/* The same reader for periodic records (src/pmp2energy.f90:207-252): cells
 * unpacked with CUTOFF1 = cel1 and CUTOFF2 = cel2 into (2 cel1 + 1)^2
 * (2 cel2 + 1) FERI blocks laid out as polyeri_eri_tensor_cells with rc1 =
 * cel1; a cell outside the ranges is "INTEGRAL FILE HAS BEEN DEGRADED". */
int polyeri_restore_fourindex_eri_cells(const char *path, int ncgs, int cel1, int cel2, double *t_c,
                                        long long *restored);
/* DIRECT .F. (RESTORE_FOURINDEX_ERI as the SCF's J/K, src/pfourindexstore.f90:
 * 805-1004, called at src/pselfconsistentfield.f90:45): appends the records
 * of the .ao file at path to the context (kept in device memory; call once
 * per file .ao.00000, .ao.00001, ...) and switches every J/K build of the
 * context -- polyeri_jk_build*, polyeri_fock_build_device, polyeri_scf,
 * polyeri_jk_build_cells, polyeri_scf_periodic -- and the molecular MP2
 * tensor to them: C(I,J,Q1) += P(K,L,Q2) v, X(I,L,Q3+Q2) += P(K,J,Q1-Q3) v
 * per record.  A periodic context reads the cells with its own CUTOFF1 /
 * CUTOFF2 and rejects a Q1 or Q2 beyond REDUCED_CEL1X ("INTEGRAL FILE HAS
 * BEEN DEGRADED"). restored = the records of this file. */
int polyeri_load_fourindex_eri(polyeri_ctx *ctx, const char *path, long long *restored);
/* direct = 1: J/K builds evaluate the integrals again (the default);
 * direct = 0: they use the loaded records (error if none are loaded). */
int polyeri_set_direct(polyeri_ctx *ctx, int direct);

/* Two-electron part of the nuclear gradient (FOURINDEX_ERI_GRADIENT,
 * src/pfourindexgradient.f90:1-1478, called at src/psinglepoint.f90:380;
 * molecule, HELIX 0): grad[3*atom + axis] = dE2/dR_atom, the sum that
 * routine adds into FORCEX / FORCEY / FORCEZ
 * (src/pfourindexgradient.f90:1334-1420), of
 * E2 = 1/2 sum P P (mu nu|lam sig) - hfexchange/4 sum P P (mu lam|nu sig)
 * at fixed P_C (ncgs x ncgs column-major, the SCF density; hfexchange =
 * DOPTN(19)), in Eh/bohr.  Quartets below the context's Schwarz threshold
 * are skipped; molecular contexts, at most 2048 atoms. */
int polyeri_eri_gradient(polyeri_ctx *ctx, const double *p_c, double hfexchange, double *grad);
/* One-electron and nuclear parts of the nuclear gradient, in Eh/bohr into
 * grad[3*atom + axis]: sum P_C (dT + dV) - W_C dS over the derivatives of the
 * basis functions (OVERLAP_KINETIC(1), NUCLEAR_ATTRACTION(1),
 * src/ponee.f90:157-214 and 418-), the derivative of the attraction
 * operator at each nucleus (HELLMANN_FEYNMAN_ATTRACTION,
 * src/phellmannfeynman.f90:86-) and the nuclear repulsion
 * (HELLMANN_FEYNMAN_REPULSION, src/phellmannfeynman.f90:1-85); w_c is the
 * energy-weighted density W_C = sum_l f_l e_l C_l C_l^T
 * (src/pfock.f90:478-497).  Host code (a few ms); molecular contexts.
 * polyeri_one_electron_gradient + polyeri_eri_gradient = dE_SCF/dR. */
int polyeri_one_electron_gradient(polyeri_ctx *ctx, const double *p_c, const double *w_c, double *grad);
/* The same from the basis arrays of polyeri_ctx_create, without a context or
 * a GPU.  No counterpart in the reference. */
int polyeri_one_electron_gradient_host(int npgs, int ncgs, const int *pax, const int *pay, const int *paz,
                                       const double *pgsx, const double *pgsy, const double *pgsz, const double *zt,
                                       const double *cc, int natom, const int *iatom, const double *xyz,
                                       const double *p_c, const double *w_c, double *grad);
/* The same sum evaluated on the host by the same code, from the basis
 * arrays of polyeri_ctx_create and a Schwarz threshold, without a context
 * or a GPU (slow; for checks).  No counterpart in the reference. */
int polyeri_eri_gradient_host(int npgs, int ncgs, const int *pax, const int *pay, const int *paz,
                              const double *pgsx, const double *pgsy, const double *pgsz, const double *zt,
                              const double *cc, int natom, const int *iatom, const double *xyz, double schwarz_tol,
                              const double *p_c, double hfexchange, double *grad);

typedef struct polyeri_scf_opts {
    double scfconv;     /* DOPTN(14) RMS density change                     */
    double scfrelax;    /* DOPTN(15)                                        */
    double hfexchange;  /* DOPTN(19)                                        */
    double lindep;      /* DOPTN(85)                                        */
    double temphf;      /* DOPTN(108), K; 0 -> 1e-4 K                       */
    int scfcycles;      /* IOPTN(13)                                        */
    int diis;           /* IOPTN(16)                                        */
    int charge;         /* IOPTN(5)                                         */
    int spherical;      /* LOPTN(69)                                        */
    int guess;          /* 0: GUESS_DENSITYMATRIX; 1: caller's P_C          */
    /* Incremental Fock builds (no counterpart in the reference): with
     * incremental = n > 0, J/K of iteration i are J/K of the previous
     * iteration plus J/K of the density difference, with a full build every
     * n iterations; 0 = full build every iteration (the reference).  dens_tol
     * > 0 drops quartets whose Schwarz bound times the largest |density
     * (difference)| they touch is below dens_tol; 0 = no density screening. */
    int incremental;
    double dens_tol;
    /* Solver choices (no counterpart in the reference; 0 = the reference's):
     * solver = 1: after the first iteration's diagonalisation, and while
     * the HOMO-LUMO gap exceeds 1 mEh at TEMPHF <= 1 K, the density is the
     * trace-correcting purification of the orthonormal-basis Fock matrix
     * (FP64 products only); the orbital energies and coefficients returned
     * come from one diagonalisation of the last Fock matrix.
     * diis_mode = 1: commutator DIIS (F P S - S P F) over the Fock matrices
     * of the last 8 iterations, every iteration, in place of the density
     * DIIS / relaxation of IOPTN(16) / DOPTN(15). */
    int solver;
    int diis_mode;
} polyeri_scf_opts;

typedef struct polyeri_scf_result {
    double e_total, e_nuc, e_kin, e_nuc_att, e_coul, e_hfx, fermi;
    int iterations, converged, n_indep, iocc;
    double jk_seconds;   /* total J/K build time (device events)            */
    double fock_seconds; /* total J/K + Fock assembly time                  */
    double scf_seconds;  /* wall time of the whole loop                     */
    long long quartets;  /* shell quartets evaluated over all J/K builds    */
} polyeri_scf_result;

/* SELF_CONSISTENT_FIELD (src/pselfconsistentfield.f90:1-162) with FOCK_BUILD
 * (src/pfock.f90, one k-point), DIIS / relaxation and the energy of
 * ENERGY_CALC, all on the device.  p_c (ncgs^2) is the guess when
 * opts->guess == 1 and the converged density on return.  eps (ncgs), co
 * (ncgs^2, Cartesian MO coefficients) and trace (2*scfcycles: density
 * change, energy per iteration) may be NULL. */
int polyeri_scf(polyeri_ctx *ctx, const polyeri_scf_opts *opts, double *p_c, double *eps,
                double *co, double *trace, polyeri_scf_result *res);

/* SELF_CONSISTENT_FIELD for a periodic context (1-D, HELIX = 0) with the
 * k-point FOCK_BUILD of src/pfock.f90:180-470: 2 kvc k-points
 * k a = pi IX / kvc, IX = -kvc .. kvc-1 (KPOINTS kvc; kvc = 0: the Gamma
 * point only), Bloch sums of F and S over the cells, canonical
 * orthogonalisation and Hermitian diagonalisation per k (rocSOLVER
 * zheevd), one Fermi level over all k, P(Q) from the occupied bands, DIIS
 * over all cells, energy per unit cell.  p_c: 2 cel1 + 1 blocks as in
 * polyeri_jk_build_cells (guess on input when opts->guess == 1, converged
 * density on return); eps: ncgs x 2 kvc band energies (may be NULL); co:
 * the band coefficients CO(ncgs, ncgs, k) as complex (re, im) pairs,
 * column-major per k (2 ncgs^2 2 kvc doubles, may be NULL); nall: IALL(k),
 * the bands kept by LINDEP at each k (may be NULL). */
int polyeri_scf_periodic(polyeri_ctx *ctx, const polyeri_scf_opts *opts, int kvc, double *p_c, double *eps,
                         double *trace, polyeri_scf_result *res, double *co, int *nall);

/* Drop-in for CALL MP2_ENERGY_ALG1 (src/psinglepoint.f90:422) of a
 * periodic run (src/pmp2energy.f90:1-548, HELIX = 0, MODULO 1): the stored
 * lattice integrals (i_0 j_Q1 | k_Q3 l_Q3+Q2) of STORE_FOURINDEX_ERI
 * (src/pfourindexstore.f90:258-747) computed on the device instead of read
 * from .ao files, then the four quarter transforms for every virtual band
 * of one k_b at once as complex products on this library's MFMA kernels
 * (transform POLYERI_GEMM_MFMA_F64 or POLYERI_GEMM_MFMA_BF16X3), E(2)
 * singlet and triplet per unit cell (normalised by (2 kvc)^3).  co, eps,
 * nall as returned by polyeri_scf_periodic; bands icore..iocc-1 occupied,
 * iocc..nall(k)-1-ivirtcore virtual; every k must hold iocc bands.
 * seconds (may be NULL): wall time of the call. */
int polyeri_mp2_energy_periodic(polyeri_ctx *ctx, int kvc, const double *co, const double *eps, const int *nall,
                                int icore, int iocc, int ivirtcore, int transform, double *e2s, double *e2t,
                                double *seconds);
/* The same with MODULO = nmod (IOPTN(99), src/pmp2energy.f90:64-80): the
 * k-point MP2 over every nmod-th k-point of the 2 kvc-point SCF grid, with
 * the sub-grid's phases and (2 kvc / nmod)^3 normalisation; nmod must divide
 * kvc ("MODULO MUST EVENLY DIVIDE KPOINTS ALONG X" otherwise). */
int polyeri_mp2_energy_periodic_mod(polyeri_ctx *ctx, int kvc, int nmod, const double *co, const double *eps,
                                    const int *nall, int icore, int iocc, int ivirtcore, int transform, double *e2s,
                                    double *e2t, double *seconds);
/* MP2_ENERGY_ALG1 of a periodic run with the quasi-particle bands
 * (IOPTN(44) = nqp > 0, src/pmp2energy.f90:137-153, 445-522, 549-558; with
 * MODULO = nmod as polyeri_mp2_energy_periodic_mod): E(2) as that call and
 * QEPSILON(band, k) in qeps (EPSILON's shape, ncgs per k-point of the SCF
 * grid) for bands max(iocc - nqp, icore) .. min(iocc + nqp, IALL(k) -
 * ivirtcore) - 1 (0-based) at the k-points the MP2 samples; every other
 * entry copies eps. */
int polyeri_mp2_energy_periodic_qp(polyeri_ctx *ctx, int kvc, int nmod, const double *co, const double *eps,
                                   const int *nall, int icore, int iocc, int ivirtcore, int nqp, int transform,
                                   double *e2s, double *e2t, double *qeps, double *seconds);

/* Drop-in for CALL MP2_ENERGY_ALG1 (src/psinglepoint.f90:422,
 * src/pmp2energy.f90:1-571) at one k-point: AO integrals computed on the
 * device (no .ao files), four quarter transforms (occupied indices first) on
 * this library's FP64 MFMA kernel, singlet and triplet E(2).
 * co: ncgs x ncgs Cartesian MO coefficients (columns = orbitals); orbitals
 * icore..iocc-1 active occupied, iocc..nall-1-ivirtcore virtual. */
int polyeri_mp2_energy(polyeri_ctx *ctx, const double *co, const double *eps, int icore,
                       int iocc, int nall, int ivirtcore, double *e2s, double *e2t);

/* polyeri_mp2_energy with the quarter transforms' products chosen: rocBLAS
 * dgemm, this library's FP64 MFMA kernel (v_mfma_f64_16x16x4f64), or its
 * BF16x3-split MFMA kernel (three BF16 pieces per operand, six products,
 * FP32 accumulation).  No counterpart in the reference (scalar loops,
 * src/pmp2energy.f90:187-430). */
#define POLYERI_GEMM_ROCBLAS 0
#define POLYERI_GEMM_MFMA_F64 1
#define POLYERI_GEMM_MFMA_BF16X3 2
int polyeri_mp2_energy_ex(polyeri_ctx *ctx, const double *co, const double *eps, int icore, int iocc,
                          int nall, int ivirtcore, int transform, double *e2s, double *e2t);
/* MP2_ENERGY_ALG1 with the quasi-particle bands (IOPTN(44) = nqp > 0,
 * src/pmp2energy.f90:445-522, 549-558) at one k-point: E(2) as
 * polyeri_mp2_energy_ex and QEPSILON for bands max(iocc - nqp, icore) ..
 * min(iocc + nqp, nall - ivirtcore) - 1 (0-based) in qeps (ncgs entries;
 * the other bands copy eps), the second-order diagonal, frequency-
 * independent band energies the reference prints as "MP2 QUASI-PARTICLE
 * ENERGY BANDS". */
int polyeri_mp2_energy_qp(polyeri_ctx *ctx, const double *co, const double *eps, int icore, int iocc, int nall,
                          int ivirtcore, int nqp, int transform, double *e2s, double *e2t, double *qeps);
/* INTEGRAL_TRANSFORM_DIRECTACCESS (src/pcc.f90:247-436) at one k-point: the
 * full MO integral set g[p + m (q + m (r + m s))] = (pq|rs) over the MOs
 * lo .. hi-1 (0-based columns of co, m = hi - lo), from AO integrals
 * computed on the device and four quarter transforms on this library's FP64
 * MFMA kernel (no .mo4 file). */
int polyeri_mo_integrals(polyeri_ctx *ctx, const double *co, int lo, int hi, double *g);
/* COUPLED_CLUSTER_ENERGY with COPTN(70) = 'CCSD' at one k-point
 * (src/pcc.f90:1-243; MP2ENERGY_DISK, MP3ENERGY_DISK :503-807): spin-
 * restricted CCSD on the active MOs icore .. nall - ivirtcore - 1 of canonical
 * closed-shell HF orbitals (co, eps as polyeri_mp2_energy), every amplitude
 * contraction on the device's FP64 MFMA GEMM, DIIS over ndiis vectors
 * (IOPTN(54)), converged when max |t_new - t_old| < conv.  E(SINGLES) and
 * E(DOUBLES) as the reference prints them (src/pcc.f90:1041-1166).
 * Returns POLYERI_E_NOTCONV ("CC ITERATION DID NOT CONVERGE") after maxiter
 * iterations without convergence, with res holding the last iterate. */
typedef struct {
    double e2, e3;                /* E(2), E(3) */
    double e_singles, e_doubles;  /* sum L t1 t1, sum L t2 (L_iajb = 2 (ia|jb) - (ib|ja)) */
    double e_corr;                /* E(CORRELATION) = e_singles + e_doubles */
    double residual;              /* max |t_new - t_old| of the last iteration */
    int iterations, converged;
    double tensor_ms, transform_ms, iter_ms, seconds;  /* device times; wall time of the call */
    double iter_gflop;            /* GEMM work of one CCSD iteration (GFLOP) */
} polyeri_cc_result;
int polyeri_ccsd(polyeri_ctx *ctx, const double *co, const double *eps, int icore, int iocc, int nall, int ivirtcore,
                 double conv, int maxiter, int ndiis, polyeri_cc_result *res);
/* COUPLED_CLUSTER_ENERGY with COPTN(70) = method ("CCSD", "CCD", "LCCD",
 * "LCCSD" or "QCISD"; src/pcc.f90:39-66): the same solver with the terms of
 * that method; E(CORRELATION) is singles + doubles for CCSD and the doubles
 * part for the others (src/pcc.f90:202-207).  Other names fail with
 * "UNKNOWN OR UNIMPLEMENTED COUPLED-CLUSTER METHOD". */
int polyeri_cc(polyeri_ctx *ctx, const char *method, const double *co, const double *eps, int icore, int iocc,
               int nall, int ivirtcore, double conv, int maxiter, int ndiis, polyeri_cc_result *res);
/* One product C = A B on those kernels (transform = POLYERI_GEMM_MFMA_F64 or
 * _BF16X3); A(m,k) = a[m*sam + k*sak], B(k,n) = b[k*sbk + n*sbn],
 * C(m,n) = c[m*scm + n*scn]; device pointers. */
int polyeri_gemm_f64(int transform, int m, int n, int k, const double *d_a, long sam, long sak,
                     const double *d_b, long sbk, long sbn, double *d_c, long scm, long scn);

/* One entry per work list: class pair (X >= Y) and, in mixed precision,
 * FP64 and FP32 lists of each. */
#define POLYERI_MAX_LISTS 64
typedef struct polyeri_stats {
    int nshell, nbf, npair, nprim_pair;
    long long quartets;        /* surviving contracted shell quartets (8-fold unique) */
    double prim_quartets;      /* their primitive quartets                   */
    double model_flops;        /* algorithmic FLOPs of one J/K build         */
    double last_jk_ms;         /* device time of the last J/K build          */
    long long last_quartets;   /* quartets the last build evaluated (after density screening) */
    double last_class_ms[POLYERI_MAX_LISTS];  /* per list kernel time of the last build */
    double class_prim_quartets[POLYERI_MAX_LISTS];
    double class_flops[POLYERI_MAX_LISTS];
    long long class_quartets[POLYERI_MAX_LISTS];
    int class_l[POLYERI_MAX_LISTS];     /* la*1000+lb*100+lc*10+ld           */
    int class_prec[POLYERI_MAX_LISTS];  /* 0: FP64, 1: FP32                  */
    int class_kind[POLYERI_MAX_LISTS];  /* 1: thread per quartet, 2: wavefront per quartet,
                                           3: thread per quartet, bra-row segments (LDS exchange rows),
                                           4: as 3 over generally contracted pairs (s shells sharing exponents),
                                           5: Rys quadrature, lanes sharing a quartet, bra-row segments */
    int class_block[POLYERI_MAX_LISTS]; /* threads per block                 */
    int class_grid[POLYERI_MAX_LISTS];  /* block cap, 0 = one pass; kinds 3 and 4: segments (blocks) launched */
    int nclass;
    /* the last molecular MP2 (polyeri_mp2_energy*): device time of the AO
     * tensor, of the four quarter transforms, and their algorithmic FLOPs
     * 2VN^4 + 2OVN^3 + 2OV^2N^2 + 2O^2V^2N */
    double last_mp2_tensor_ms, last_mp2_transform_ms, last_mp2_flops;
} polyeri_stats;
int polyeri_get_stats(polyeri_ctx *ctx, polyeri_stats *st);

/* Per-class kernel timing on/off (adds events between launches; J/K builds
 * then run eagerly instead of from a captured HIP graph). */
int polyeri_set_profiling(polyeri_ctx *ctx, int on);

/* Graph replay of J/K builds on (default) or off (every build launched
 * eagerly).  No counterpart in the reference. */
int polyeri_set_graph(polyeri_ctx *ctx, int on);

/* Generally contracted evaluation of the (ls|l's) classes on or off.  With a
 * basis whose s shells on one centre share exponents (cc-pVDZ 1s/2s), those
 * lists can run over pairs of shell groups, one primitive loop for all member
 * quartets (on by default where such groups exist; polyeri_autotune keeps it
 * per list where it is faster).  Same quartets, same results to rounding.
 * No counterpart in the reference. */
int polyeri_set_general_contraction(polyeri_ctx *ctx, int on);

/* Autotuner (no counterpart in the reference): times every work list with
 * each candidate launch shape (thread-per-quartet and bra-row segment
 * kernels with 256/128/64 threads per block, one pass or a persistent grid,
 * two segmentations; wavefront-per-quartet kernel with three grid sizes and
 * primitive batches of 1/2/4/8; the generally contracted form with
 * 256/128/64 threads and two segmentations) reps times and keeps the
 * fastest.  Results are unchanged to rounding; later builds use the chosen
 * shapes (polyeri_stats). */
int polyeri_autotune(polyeri_ctx *ctx, int reps);
/* The launch configuration of work list `list` (0 .. nclass - 1 of
 * polyeri_get_stats) as 9 ints: kind (0 default, 1 thread per quartet,
 * 2 wavefront per quartet, 3 bra-row segments, 5 Rys segments), threads per
 * block, grid cap, segment variant, primitive batch (-1 default), generally
 * contracted form on, its block, its segment variant, autotuned time in
 * microseconds.  set restores such a record (e.g. the autotuner's picks saved
 * by an earlier process, so that a counter run times the same shapes).  No
 * counterpart in the reference. */
int polyeri_get_launch_config(const polyeri_ctx *ctx, int list, int *cfg);
int polyeri_set_launch_config(polyeri_ctx *ctx, int list, const int *cfg);

/* Nodes u[0..n-1] (ascending, in (0, 1)) and weights w[0..n-1] of the
 * n-point Rys quadrature at Boys argument T >= 0, n = 1..5: the Gauss rule
 * of the weight exp(-T t^2) on t in [0, 1] in the variable u = t^2, so that
 * sum_r w_r u_r^m = F_m(T) for m < 2n.  These are the nodes the library's
 * Rys-quadrature kernels use for the d-heavy quartet classes; host
 * evaluation, no device needed.  No counterpart in the reference, whose
 * integrals come from F_m(T) through the Obara-Saika recursion
 * (src/pfourindexdirect.f90:543-690). */
int polyeri_rys_nodes(int n, double T, double *u, double *w);

/* The Cartesian block (ab|cd) of one primitive quartet of unnormalised
 * Gaussians (centres xyz[0..11] = A, B, C, D; exponents ex[0..3]), for a
 * class (la lb|lc ld) that the Rys kernels evaluate ((dp|pp) (dp|dp) (dd|pp)
 * (dd|ds) (dd|dp) (dd|dd)), computed on the host by the same per-lane code
 * as the kernels: every lane variant and axis-permutation member, mapped to
 * the global components.  out: ncart(la) ncart(lb) ncart(lc) ncart(ld)
 * doubles, nested a, b, c, d, component order xx xy xz yy yz zz / x y z.
 * Fails when the variants do not produce every integral exactly once.  Host
 * only (tests of the generated recursions); no counterpart in the reference,
 * whose integrals come from the Obara-Saika recursion
 * (src/pfourindexdirect.f90:543-690). */
int polyeri_rys_block(int la, int lb, int lc, int ld, const double *xyz, const double *ex, double *out);

const char *polyeri_last_error(void);

/* ---- Learned SCF guess: multilayer perceptron on the device ------------
 * No counterpart in the reference (its guess is GUESS_DENSITYMATRIX,
 * src/pdensitymatrix.f90:83-127).  Layer l: Z = A W^T + b, A = silu(Z),
 * the last layer linear; GEMMs in BF16 on the MFMA units with FP32
 * accumulation, FP32 master weights, Adam.  Device pointers are row-major
 * FP32 (batch x width). */
typedef struct polyeri_mlp polyeri_mlp;
int polyeri_mlp_create(int nlayers, const int *dims, unsigned long long seed, polyeri_mlp **mlp);
void polyeri_mlp_destroy(polyeri_mlp *mlp);
long polyeri_mlp_nparam(const polyeri_mlp *mlp);
/* set = 1: host -> device (and reset Adam state); 0: device -> host */
int polyeri_mlp_params(polyeri_mlp *mlp, float *host, int set);
/* gradients of the last train step (update = 0 leaves the weights alone) */
int polyeri_mlp_grads(polyeri_mlp *mlp, float *host);
int polyeri_mlp_forward(polyeri_mlp *mlp, const float *d_x, int batch, float *d_y);
/* loss = sum w (y - t)^2 / sum w over the batch; update = 1 takes an Adam step */
int polyeri_mlp_train_step(polyeri_mlp *mlp, const float *d_x, const float *d_t, const float *d_w, int batch,
                           float lr, int update, float *loss);
/* C = alpha A B + beta C on the MFMA units (operands rounded to BF16, FP32
 * accumulation); A(m,k) = a[m*sam + k*sak], B(k,n) = b[k*sbk + n*sbn], C
 * row-major with leading dimension ldc. */
int polyeri_gemm_bf16(int m, int n, int k, const float *d_a, long sam, long sak, const float *d_b, long sbk,
                      long sbn, float alpha, float beta, float *d_c, long ldc);

#ifdef __cplusplus
}
#endif
#endif
