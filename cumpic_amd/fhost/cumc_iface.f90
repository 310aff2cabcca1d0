!-----------------------------------------------------------------------------------------------
! cumc_iface.f90 -- ISO_C_BINDING interface of libcumc.so (include/cumc.h) for a Fortran host.
!
! The reference calls its step as argument-free subroutines on MODULE DEFINITION globals
! (solver/main.f90:130 finddt, :138 Rungekutta; solver/initial.f90:83 initial_update;
! solver/mhd.f90:15 find_divb).  This module declares the C entry points that replace them;
! arrays keep the reference's Fortran layout prim(1:8,-2:nx+3,-2:ny+3,-2:nz+3).
!-----------------------------------------------------------------------------------------------
MODULE CUMC_IFACE
USE ISO_C_BINDING
IMPLICIT NONE

! cumc_config of include/cumc.h
TYPE, BIND(C) :: cumc_config
  INTEGER(C_INT) :: nx, ny, nz
  INTEGER(C_INT) :: coordinate
  INTEGER(C_INT) :: boundary_flag(6)
  INTEGER(C_INT) :: recon
  INTEGER(C_INT) :: model
  INTEGER(C_INT) :: geom_source_repeat
  REAL(C_DOUBLE) :: ggas, cfl
  REAL(C_DOUBLE) :: rho_floor, eps_floor
  REAL(C_DOUBLE) :: r_sh
  INTEGER(C_INT) :: solver
  INTEGER(C_INT) :: precision
END TYPE

INTERFACE
  INTEGER(C_INT) FUNCTION cumc_create(cfg, ctx) BIND(C, NAME="cumc_create")
    IMPORT :: C_INT, C_PTR, cumc_config
    TYPE(cumc_config), INTENT(IN) :: cfg
    TYPE(C_PTR), INTENT(OUT) :: ctx
  END FUNCTION
  INTEGER(C_INT) FUNCTION cumc_destroy(ctx) BIND(C, NAME="cumc_destroy")
    IMPORT :: C_INT, C_PTR
    TYPE(C_PTR), VALUE :: ctx
  END FUNCTION
  INTEGER(C_INT) FUNCTION cumc_set_grid(ctx, xf, yf, zf) BIND(C, NAME="cumc_set_grid")
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR), VALUE :: ctx
    REAL(C_DOUBLE), INTENT(IN) :: xf(*), yf(*), zf(*)
  END FUNCTION
  INTEGER(C_INT) FUNCTION cumc_upload(ctx, prim, eps) BIND(C, NAME="cumc_upload")
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR), VALUE :: ctx
    REAL(C_DOUBLE), INTENT(IN) :: prim(*), eps(*)
  END FUNCTION
  INTEGER(C_INT) FUNCTION cumc_download(ctx, prim, eps, bcell) BIND(C, NAME="cumc_download")
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR), VALUE :: ctx
    REAL(C_DOUBLE), INTENT(OUT) :: prim(*), eps(*), bcell(*)
  END FUNCTION
  INTEGER(C_INT) FUNCTION cumc_initial_update(ctx) BIND(C, NAME="cumc_initial_update")
    IMPORT :: C_INT, C_PTR
    TYPE(C_PTR), VALUE :: ctx
  END FUNCTION
  INTEGER(C_INT) FUNCTION cumc_finddt(ctx, dt) BIND(C, NAME="cumc_finddt")
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR), VALUE :: ctx
    REAL(C_DOUBLE), INTENT(OUT) :: dt
  END FUNCTION
  INTEGER(C_INT) FUNCTION cumc_rungekutta(ctx, dt) BIND(C, NAME="cumc_rungekutta")
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR), VALUE :: ctx
    REAL(C_DOUBLE), VALUE :: dt
  END FUNCTION
  INTEGER(C_INT) FUNCTION cumc_find_divb(ctx, max_signed, max_abs) BIND(C, NAME="cumc_find_divb")
    IMPORT :: C_INT, C_PTR, C_DOUBLE
    TYPE(C_PTR), VALUE :: ctx
    REAL(C_DOUBLE), INTENT(OUT) :: max_signed, max_abs
  END FUNCTION
  TYPE(C_PTR) FUNCTION cumc_error_string(code) BIND(C, NAME="cumc_error_string")
    IMPORT :: C_INT, C_PTR
    INTEGER(C_INT), VALUE :: code
  END FUNCTION
  TYPE(C_PTR) FUNCTION cumc_last_error(ctx) BIND(C, NAME="cumc_last_error")
    IMPORT :: C_PTR
    TYPE(C_PTR), VALUE :: ctx
  END FUNCTION
  ! profile output in the reference's HDF5 layout (h5profile.c; replaces dataio.f90:7-77)
  INTEGER(C_INT) FUNCTION cumc_h5_profile(dir, n_iter, nx, ny, nz, xf, yf, zf, prim, eps) &
      BIND(C, NAME="cumc_h5_profile")
    IMPORT :: C_INT, C_CHAR, C_DOUBLE
    CHARACTER(KIND=C_CHAR), INTENT(IN) :: dir(*)
    INTEGER(C_INT), VALUE :: n_iter, nx, ny, nz
    REAL(C_DOUBLE), INTENT(IN) :: xf(*), yf(*), zf(*), prim(*), eps(*)
  END FUNCTION
END INTERFACE

CONTAINS

! STOP on a non-zero return code, as the reference STOPs on configuration errors
SUBROUTINE CUMC_CHECK(ierr, what, ctx)
  INTEGER(C_INT), INTENT(IN) :: ierr
  CHARACTER(*), INTENT(IN) :: what
  TYPE(C_PTR), INTENT(IN) :: ctx
  IF (ierr /= 0) THEN
    WRITE(*,*) 'libcumc error in ', what, ': ', TRIM(c_string(cumc_error_string(ierr))), ' ', &
               TRIM(c_string(cumc_last_error(ctx)))
    STOP 1
  END IF
END SUBROUTINE

FUNCTION c_string(p) RESULT(s)
  TYPE(C_PTR), INTENT(IN) :: p
  CHARACTER(LEN=256) :: s
  CHARACTER(KIND=C_CHAR), POINTER :: a(:)
  INTEGER :: i
  s = ' '
  IF (.NOT. C_ASSOCIATED(p)) RETURN
  CALL C_F_POINTER(p, a, [256])
  DO i = 1, 256
    IF (a(i) == C_NULL_CHAR) EXIT
    s(i:i) = a(i)
  END DO
END FUNCTION

END MODULE CUMC_IFACE
