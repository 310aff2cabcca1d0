!-----------------------------------------------------------------------------------------------
! model/riemann/riemann.f90 -- the riemann model plugin of the Fortran host: its three named
! problems computed in place, as the reference's plugin does (model/riemann/src/defs.f90:20-26,
! model.f90:4-57, riemann1d.f90:15-89, riemann2d.f90:15-63):
!   riemann_test = 1  Toro's shock tube 1: rho, p = 1, 1 for x <= 0.5, else 0.125, 0.1
!   riemann_test = 2  Brio-Wu: that gas with By = +1 / -1 on the two sides and Bx = 0.75
!   riemann_test = 3  MHD rotor: a dense disc r <= 0.1 about (0.5, 0.5) spinning at unit rim
!                     speed, a linear taper to r = 0.115, p = 0.5, Bx = 2.5/sqrt(4 pi)
! on the unit interval (tests 1, 2: n x 1 x 1 cells) or square (test 3: n x n x 1), outflow on
! every side, gamma = 5/3 (defs.f90:14), GET_MODEL filling [0,n]^3 (model.f90:15-21) at
! GET_COORD's cell centres (coords.f90:63-65).
!
! The reference fixes these as compile-time PARAMETERs; here MODEL_SETUP reads them from a
! namelist file (all optional):
!   &riemann riemann_test = 3, n = 400, cfl = 0.7, recon = 4, solver = 2, precision = 0,
!            max_steps = 1000000, total_time = (0.2 | 0.1 | 0.15 by test), output_profiletime = 100.0 /
! CUSTOM_GRID keeps GETGRID's uniform faces (the plugin's CUSTOM_GRID is empty, custom.f90:34-39),
! and OPENFILE_CUSTOM writes nothing (custom.f90:114-119).  The in-loop hooks (EOS and the
! eps >= 0 floor of CUSTOM_CHECKRHO) run on the GPU as cfg%model = CUMC_MODEL_RIEMANN.
!-----------------------------------------------------------------------------------------------
MODULE MODEL_PLUGIN
USE ISO_C_BINDING
USE CUMC_IFACE
IMPLICIT NONE

CHARACTER(LEN=*), PARAMETER :: model_name = 'riemann'
! defs.f90:14 and :20-22
REAL(C_DOUBLE), PARAMETER :: ggas = 5.0D0/3.0D0
INTEGER, PARAMETER :: toro_1 = 1, brio_wu = 2, mhd_rotor = 3
! definition.f90:48
REAL(C_DOUBLE), PARAMETER :: pi = 4.D0*DATAN(1.D0)
INTEGER, PRIVATE :: riemann_test = mhd_rotor

CONTAINS

SUBROUTINE MODEL_SETUP(path, cfg, max_steps, total_time, output_time, lo, hi)
  CHARACTER(*), INTENT(IN) :: path
  TYPE(cumc_config), INTENT(OUT) :: cfg
  INTEGER, INTENT(OUT) :: max_steps
  REAL(C_DOUBLE), INTENT(OUT) :: total_time, output_time, lo(3), hi(3)
  INTEGER :: n, recon, solver, precision, u, ios
  REAL(C_DOUBLE) :: cfl, output_profiletime
  NAMELIST /riemann/ riemann_test, n, cfl, recon, solver, precision, max_steps, total_time, output_profiletime

  n = 400
  cfl = 0.7D0
  recon = 4
  solver = 2
  precision = 0
  max_steps = 1000000
  total_time = -1.0D0
  output_profiletime = 100.0D0
  OPEN(NEWUNIT=u, FILE=path, STATUS='OLD', ACTION='READ', IOSTAT=ios)
  IF (ios /= 0) THEN
    WRITE(*,*) 'riemann plugin: cannot open the parameter file ', TRIM(path)
    STOP 2
  END IF
  READ(u, NML=riemann, IOSTAT=ios)
  CLOSE(u)
  IF (ios /= 0) THEN
    WRITE(*,*) 'riemann plugin: bad &riemann namelist in ', TRIM(path)
    STOP 2
  END IF
  IF (riemann_test < toro_1 .OR. riemann_test > mhd_rotor .OR. n < 1) THEN
    WRITE(*,*) 'riemann plugin: riemann_test must be 1, 2 or 3 and n >= 1'
    STOP 2
  END IF
  IF (total_time < 0.0D0) THEN
    SELECT CASE (riemann_test)
    CASE (toro_1)
      total_time = 0.2D0
    CASE (brio_wu)
      total_time = 0.1D0
    CASE DEFAULT
      total_time = 0.15D0
    END SELECT
  END IF
  output_time = output_profiletime
  cfg%nx = n
  cfg%ny = 1
  IF (riemann_test == mhd_rotor) cfg%ny = n
  cfg%nz = 1
  cfg%coordinate = 1
  cfg%boundary_flag = 1
  cfg%recon = recon
  cfg%model = 0
  cfg%geom_source_repeat = 8
  cfg%ggas = ggas
  cfg%cfl = cfl
  cfg%rho_floor = 0.0D0
  cfg%eps_floor = 0.0D0
  cfg%r_sh = 1.0D0
  cfg%solver = solver
  cfg%precision = precision
  lo = 0.0D0
  hi = 1.0D0
END SUBROUTINE

SUBROUTINE CUSTOM_GRID(cfg, xf, yf, zf)
  TYPE(cumc_config), INTENT(IN) :: cfg
  REAL(C_DOUBLE), INTENT(INOUT) :: xf(-3:), yf(-3:), zf(-3:)
END SUBROUTINE

! model.f90:4-25
SUBROUTINE GET_MODEL(cfg, xf, yf, zf, prim, eps)
  TYPE(cumc_config), INTENT(IN) :: cfg
  REAL(C_DOUBLE), INTENT(IN) :: xf(-3:), yf(-3:), zf(-3:)
  REAL(C_DOUBLE), INTENT(INOUT) :: prim(1:, -2:, -2:, -2:), eps(-2:, -2:, -2:)
  INTEGER :: j, k, l
  DO l = 0, cfg%nz
    DO k = 0, cfg%ny
      DO j = 0, cfg%nx
        CALL RIEMANN_PROBLEM(0.5D0*(xf(j) + xf(j-1)), 0.5D0*(yf(k) + yf(k-1)), prim(:,j,k,l), eps(j,k,l))
      END DO
    END DO
  END DO
END SUBROUTINE

! model.f90:32-57 with the cell centre (x, y) of GET_COORD passed in
SUBROUTINE RIEMANN_PROBLEM(x, y, p_out, eps_out)
  REAL(C_DOUBLE), INTENT(IN) :: x, y
  REAL(C_DOUBLE), INTENT(INOUT) :: p_out(8)
  REAL(C_DOUBLE), INTENT(OUT) :: eps_out
  REAL(C_DOUBLE) :: r, fr, r0, r1
  SELECT CASE (riemann_test)
  CASE (toro_1, brio_wu)
    ! riemann1d.f90:15-47, :54-89
    IF (x <= 0.5D0) THEN
      p_out(1) = 1.0D0
      p_out(5) = 1.0D0
      IF (riemann_test == brio_wu) p_out(7) = 1.0D0
    ELSE
      p_out(1) = 0.125D0
      p_out(5) = 0.1D0
      IF (riemann_test == brio_wu) p_out(7) = -1.0D0
    END IF
    IF (riemann_test == brio_wu) p_out(6) = 0.75D0
  CASE (mhd_rotor)
    ! riemann2d.f90:15-63
    r0 = 0.1D0
    r1 = 0.115D0
    r = DSQRT((x - 0.5D0)**2 + (y - 0.5D0)**2)
    fr = (r1 - r)/(r1 - r0)
    IF (r <= r0) THEN
      p_out(1) = 10.0D0
      p_out(2) = -1.0D0*(y - 0.5D0)/r0
      p_out(3) = 1.0D0*(x - 0.5D0)/r0
    ELSE IF (r >= r1) THEN
      p_out(1) = 1.0D0
      p_out(2) = 0.0D0
      p_out(3) = 0.0D0
    ELSE
      p_out(1) = 1.0D0 + 9.0D0*fr
      p_out(2) = -fr*1.0D0*(y - 0.5D0)/r
      p_out(3) = fr*1.0D0*(x - 0.5D0)/r
    END IF
    p_out(5) = 0.5D0
    p_out(6) = 2.5D0/DSQRT(4.0D0*pi)
  END SELECT
  eps_out = p_out(5)/p_out(1)/(ggas - 1.0D0)
END SUBROUTINE

SUBROUTINE OPENFILE_CUSTOM()
END SUBROUTINE

END MODULE MODEL_PLUGIN
