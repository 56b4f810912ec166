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

      INTEGER(C_INT) FUNCTION POLYERI_MP2_ENERGY(CTX, CO, EPS, ICORE, IOCC, NALL, IVIRTCORE, E2S, E2T) &
            BIND(C, NAME="polyeri_mp2_energy")
         IMPORT :: C_INT, C_DOUBLE, C_PTR
         TYPE(C_PTR), VALUE :: CTX
         REAL(C_DOUBLE) :: CO(*), EPS(*), E2S, E2T
         INTEGER(C_INT), VALUE :: ICORE, IOCC, NALL, IVIRTCORE
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
