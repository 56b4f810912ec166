! polyeri_mod.f90 -- ISO_C_BINDING interface to libpolyeri (include/polyeri.h)
! for the Fortran side of polymer.  A maintainer adds this module to
! src/Makefile and replaces
!     CALL FOURINDEX_ERI              (src/pselfconsistentfield.f90:42)
! by a call of POLYERI_JK_BUILD on P_C that fills C_C and X_C, and
!     CALL MP2_ENERGY_ALG1            (src/psinglepoint.f90:422)
! by POLYERI_MP2_ENERGY.  Arrays are passed as the module arrays already are
! (column-major, contracted functions including ghosts); the context is
! built once per geometry from the BASISSET / STRUCTURE module arrays.
MODULE POLYERI
   USE ISO_C_BINDING
   IMPLICIT NONE

   INTEGER(C_INT), PARAMETER :: POLYERI_FP64 = 0, POLYERI_MIXED = 1
   INTEGER(C_INT), PARAMETER :: POLYERI_GEMM_ROCBLAS = 0, POLYERI_GEMM_MFMA_F64 = 1, POLYERI_GEMM_MFMA_BF16X3 = 2

   TYPE, BIND(C) :: POLYERI_OPTS
      REAL(C_DOUBLE) :: SCHWARZ_TOL
      REAL(C_DOUBLE) :: PRIM_TOL
      INTEGER(C_INT) :: PRECISION
      REAL(C_DOUBLE) :: MIXED_CUT
      INTEGER(C_INT) :: DEVICE
   END TYPE

   TYPE, BIND(C) :: POLYERI_SCF_OPTS
      REAL(C_DOUBLE) :: SCFCONV, SCFRELAX, HFEXCHANGE, LINDEP, TEMPHF
      INTEGER(C_INT) :: SCFCYCLES, DIIS, CHARGE, SPHERICAL, GUESS
      INTEGER(C_INT) :: INCREMENTAL
      REAL(C_DOUBLE) :: DENS_TOL
      INTEGER(C_INT) :: SOLVER, DIIS_MODE
   END TYPE

   TYPE, BIND(C) :: POLYERI_SCF_RESULT
      REAL(C_DOUBLE) :: E_TOTAL, E_NUC, E_KIN, E_NUC_ATT, E_COUL, E_HFX, FERMI
      INTEGER(C_INT) :: ITERATIONS, CONVERGED, N_INDEP, IOCC
      REAL(C_DOUBLE) :: JK_SECONDS, FOCK_SECONDS, SCF_SECONDS
      INTEGER(C_LONG_LONG) :: QUARTETS
   END TYPE

   INTERFACE
      INTEGER(C_INT) FUNCTION POLYERI_BASIS_READ(PATH, NATOM, IATOM, XYZ, NPGS, NCGS, NPSHELL, &
            PAX, PAY, PAZ, PGS_ATOM, PGSX, PGSY, PGSZ, ZT, CC, CGSX, CGSY, CGSZ, P_SHELL, &
            P_SHELL_ANG) BIND(C, NAME="polyeri_basis_read")
         IMPORT :: C_INT, C_DOUBLE, C_CHAR, C_PTR
         CHARACTER(KIND=C_CHAR), DIMENSION(*) :: PATH
         INTEGER(C_INT), VALUE :: NATOM
         INTEGER(C_INT) :: IATOM(*), NPGS, NCGS, NPSHELL
         REAL(C_DOUBLE) :: XYZ(*)
         TYPE(C_PTR), VALUE :: PAX, PAY, PAZ, PGS_ATOM, PGSX, PGSY, PGSZ, ZT, CC, CGSX, CGSY, CGSZ, &
                               P_SHELL, P_SHELL_ANG
      END FUNCTION

      INTEGER(C_INT) FUNCTION POLYERI_CTX_CREATE(NPGS, NCGS, PAX, PAY, PAZ, PGSX, PGSY, PGSZ, ZT, CC, &
            NATOM, IATOM, XYZ, OPTS, CTX) BIND(C, NAME="polyeri_ctx_create")
         IMPORT :: C_INT, C_DOUBLE, C_PTR, POLYERI_OPTS
         INTEGER(C_INT), VALUE :: NPGS, NCGS, NATOM
         INTEGER(C_INT) :: PAX(*), PAY(*), PAZ(*), IATOM(*)
         REAL(C_DOUBLE) :: PGSX(*), PGSY(*), PGSZ(*), ZT(*), CC(*), XYZ(*)
         TYPE(POLYERI_OPTS) :: OPTS
         TYPE(C_PTR) :: CTX
      END FUNCTION

      SUBROUTINE POLYERI_CTX_DESTROY(CTX) BIND(C, NAME="polyeri_ctx_destroy")
         IMPORT :: C_PTR
         TYPE(C_PTR), VALUE :: CTX
      END SUBROUTINE

      INTEGER(C_INT) FUNCTION POLYERI_JK_BUILD(CTX, P_C, C_C, X_C) BIND(C, NAME="polyeri_jk_build")
         IMPORT :: C_INT, C_DOUBLE, C_PTR
         TYPE(C_PTR), VALUE :: CTX
         REAL(C_DOUBLE) :: P_C(*), C_C(*), X_C(*)
      END FUNCTION

      INTEGER(C_INT) FUNCTION POLYERI_JK_BUILD_SCREENED(CTX, P_C, DENS_TOL, C_C, X_C) &
            BIND(C, NAME="polyeri_jk_build_screened")
         IMPORT :: C_INT, C_DOUBLE, C_PTR
         TYPE(C_PTR), VALUE :: CTX
         REAL(C_DOUBLE) :: P_C(*), C_C(*), X_C(*)
         REAL(C_DOUBLE), VALUE :: DENS_TOL
      END FUNCTION

      INTEGER(C_INT) FUNCTION POLYERI_ONE_ELECTRON(CTX, S_C, T_C, N_C, E_NUC) &
            BIND(C, NAME="polyeri_one_electron")
         IMPORT :: C_INT, C_DOUBLE, C_PTR
         TYPE(C_PTR), VALUE :: CTX
         REAL(C_DOUBLE) :: S_C(*), T_C(*), N_C(*), E_NUC
      END FUNCTION

      INTEGER(C_INT) FUNCTION POLYERI_SCF(CTX, OPTS, P_C, EPS, CO, TRACE, RES) BIND(C, NAME="polyeri_scf")
         IMPORT :: C_INT, C_DOUBLE, C_PTR, POLYERI_SCF_OPTS, POLYERI_SCF_RESULT
         TYPE(C_PTR), VALUE :: CTX
         TYPE(POLYERI_SCF_OPTS) :: OPTS
         REAL(C_DOUBLE) :: P_C(*), EPS(*), CO(*), TRACE(*)
         TYPE(POLYERI_SCF_RESULT) :: RES
      END FUNCTION

      INTEGER(C_INT) FUNCTION POLYERI_ERI_GRADIENT(CTX, P_C, HFEXCHANGE, GRAD) &
            BIND(C, NAME="polyeri_eri_gradient")
         IMPORT :: C_INT, C_DOUBLE, C_PTR
         TYPE(C_PTR), VALUE :: CTX
         REAL(C_DOUBLE) :: P_C(*), GRAD(*)
         REAL(C_DOUBLE), VALUE :: HFEXCHANGE
      END FUNCTION

      INTEGER(C_INT) FUNCTION POLYERI_MP2_ENERGY(CTX, CO, EPS, ICORE, IOCC, NALL, IVIRTCORE, E2S, E2T) &
            BIND(C, NAME="polyeri_mp2_energy")
         IMPORT :: C_INT, C_DOUBLE, C_PTR
         TYPE(C_PTR), VALUE :: CTX
         REAL(C_DOUBLE) :: CO(*), EPS(*), E2S, E2T
         INTEGER(C_INT), VALUE :: ICORE, IOCC, NALL, IVIRTCORE
      END FUNCTION

      INTEGER(C_INT) FUNCTION POLYERI_MP2_ENERGY_EX(CTX, CO, EPS, ICORE, IOCC, NALL, IVIRTCORE, TRANSFORM, &
            E2S, E2T) BIND(C, NAME="polyeri_mp2_energy_ex")
         IMPORT :: C_INT, C_DOUBLE, C_PTR
         TYPE(C_PTR), VALUE :: CTX
         REAL(C_DOUBLE) :: CO(*), EPS(*), E2S, E2T
         INTEGER(C_INT), VALUE :: ICORE, IOCC, NALL, IVIRTCORE, TRANSFORM
      END FUNCTION

      INTEGER(C_INT) FUNCTION POLYERI_MP2_ENERGY_QP(CTX, CO, EPS, ICORE, IOCC, NALL, IVIRTCORE, NQP, TRANSFORM, &
            E2S, E2T, QEPS) BIND(C, NAME="polyeri_mp2_energy_qp")
         IMPORT :: C_INT, C_DOUBLE, C_PTR
         TYPE(C_PTR), VALUE :: CTX
         REAL(C_DOUBLE) :: CO(*), EPS(*), E2S, E2T, QEPS(*)
         INTEGER(C_INT), VALUE :: ICORE, IOCC, NALL, IVIRTCORE, NQP, TRANSFORM
      END FUNCTION

      ! periodic (1-D, HELIX = 0): cells -CEL1..CEL1 of P_C, C_C, X_C are
      ! contiguous NCGS x NCGS blocks, k-point arrays hold IX = -KVC..KVC-1
      INTEGER(C_INT) FUNCTION POLYERI_CTX_CREATE_PERIODIC(NPGS, NCGS, PAX, PAY, PAZ, PGSX, PGSY, PGSZ, ZT, &
            CC, NATOM, IATOM, XYZ, PERIOD, CEL1, CEL2, OPTS, CTX) BIND(C, NAME="polyeri_ctx_create_periodic")
         IMPORT :: C_INT, C_DOUBLE, C_PTR, POLYERI_OPTS
         INTEGER(C_INT), VALUE :: NPGS, NCGS, NATOM, CEL1, CEL2
         INTEGER(C_INT) :: PAX(*), PAY(*), PAZ(*), IATOM(*)
         REAL(C_DOUBLE) :: PGSX(*), PGSY(*), PGSZ(*), ZT(*), CC(*), XYZ(*)
         REAL(C_DOUBLE), VALUE :: PERIOD
         TYPE(POLYERI_OPTS) :: OPTS
         TYPE(C_PTR) :: CTX
      END FUNCTION

      INTEGER(C_INT) FUNCTION POLYERI_JK_BUILD_CELLS(CTX, P_C, C_C, X_C) BIND(C, NAME="polyeri_jk_build_cells")
         IMPORT :: C_INT, C_DOUBLE, C_PTR
         TYPE(C_PTR), VALUE :: CTX
         REAL(C_DOUBLE) :: P_C(*), C_C(*), X_C(*)
      END FUNCTION

      INTEGER(C_INT) FUNCTION POLYERI_SCF_PERIODIC(CTX, OPTS, KVC, P_C, EPS, TRACE, RES, CO, NALL) &
            BIND(C, NAME="polyeri_scf_periodic")
         IMPORT :: C_INT, C_DOUBLE, C_DOUBLE_COMPLEX, C_PTR, POLYERI_SCF_OPTS, POLYERI_SCF_RESULT
         TYPE(C_PTR), VALUE :: CTX
         TYPE(POLYERI_SCF_OPTS) :: OPTS
         INTEGER(C_INT), VALUE :: KVC
         REAL(C_DOUBLE) :: P_C(*), EPS(*), TRACE(*)
         TYPE(POLYERI_SCF_RESULT) :: RES
         COMPLEX(C_DOUBLE_COMPLEX) :: CO(*)
         INTEGER(C_INT) :: NALL(*)
      END FUNCTION

      INTEGER(C_INT) FUNCTION POLYERI_MP2_ENERGY_PERIODIC(CTX, KVC, CO, EPS, NALL, ICORE, IOCC, IVIRTCORE, &
            TRANSFORM, E2S, E2T, SECONDS) BIND(C, NAME="polyeri_mp2_energy_periodic")
         IMPORT :: C_INT, C_DOUBLE, C_DOUBLE_COMPLEX, C_PTR
         TYPE(C_PTR), VALUE :: CTX
         INTEGER(C_INT), VALUE :: KVC, ICORE, IOCC, IVIRTCORE, TRANSFORM
         COMPLEX(C_DOUBLE_COMPLEX) :: CO(*)
         REAL(C_DOUBLE) :: EPS(*), E2S, E2T, SECONDS
         INTEGER(C_INT) :: NALL(*)
      END FUNCTION

      INTEGER(C_INT) FUNCTION POLYERI_MP2_ENERGY_PERIODIC_MOD(CTX, KVC, NMOD, CO, EPS, NALL, ICORE, IOCC, &
            IVIRTCORE, TRANSFORM, E2S, E2T, SECONDS) BIND(C, NAME="polyeri_mp2_energy_periodic_mod")
         IMPORT :: C_INT, C_DOUBLE, C_DOUBLE_COMPLEX, C_PTR
         TYPE(C_PTR), VALUE :: CTX
         INTEGER(C_INT), VALUE :: KVC, NMOD, ICORE, IOCC, IVIRTCORE, TRANSFORM
         COMPLEX(C_DOUBLE_COMPLEX) :: CO(*)
         REAL(C_DOUBLE) :: EPS(*), E2S, E2T, SECONDS
         INTEGER(C_INT) :: NALL(*)
      END FUNCTION

      INTEGER(C_INT) FUNCTION POLYERI_REDUCED_CEL1(CTX) BIND(C, NAME="polyeri_reduced_cel1")
         IMPORT :: C_INT, C_PTR
         TYPE(C_PTR), VALUE :: CTX
      END FUNCTION

      ! .ao integral files; PATH is NUL-terminated (TRIM(NAME)//C_NULL_CHAR)
      INTEGER(C_INT) FUNCTION POLYERI_STORE_FOURINDEX_ERI(CTX, PATH, INTTOL, STORED) &
            BIND(C, NAME="polyeri_store_fourindex_eri")
         IMPORT :: C_INT, C_DOUBLE, C_PTR, C_CHAR, C_LONG_LONG
         TYPE(C_PTR), VALUE :: CTX
         CHARACTER(KIND=C_CHAR) :: PATH(*)
         REAL(C_DOUBLE), VALUE :: INTTOL
         INTEGER(C_LONG_LONG) :: STORED
      END FUNCTION

      INTEGER(C_INT) FUNCTION POLYERI_RESTORE_FOURINDEX_ERI(PATH, NCGS, FERI, RESTORED) &
            BIND(C, NAME="polyeri_restore_fourindex_eri")
         IMPORT :: C_INT, C_DOUBLE, C_CHAR, C_LONG_LONG
         CHARACTER(KIND=C_CHAR) :: PATH(*)
         INTEGER(C_INT), VALUE :: NCGS
         REAL(C_DOUBLE) :: FERI(*)
         INTEGER(C_LONG_LONG) :: RESTORED
      END FUNCTION

      FUNCTION POLYERI_LAST_ERROR() BIND(C, NAME="polyeri_last_error")
         IMPORT :: C_PTR
         TYPE(C_PTR) :: POLYERI_LAST_ERROR
      END FUNCTION
   END INTERFACE

CONTAINS

   ! The message of the last failed call, for PABORT (src/pmain.f90:74-103).
   FUNCTION POLYERI_ERROR_MESSAGE() RESULT(MSG)
      CHARACTER(LEN=256) :: MSG
      CHARACTER(KIND=C_CHAR), POINTER :: S(:)
      TYPE(C_PTR) :: P
      INTEGER :: I
      MSG = ' '
      P = POLYERI_LAST_ERROR()
      IF (.NOT. C_ASSOCIATED(P)) RETURN
      CALL C_F_POINTER(P, S, [256])
      DO I = 1, 256
         IF (S(I) == C_NULL_CHAR) EXIT
         MSG(I:I) = S(I)
      ENDDO
   END FUNCTION

END MODULE POLYERI
