!-----------------------------------------------------------------------------------------------
! model/file/model_file.f90 -- the run-file model plugin of the Fortran host: every parameter, the
! face coordinates and the initial state come from a run file written by the initial-condition
! tool (cumpic_amd/fhost/runfile.py), for workloads whose GET_MODEL lives in Python
! (cumpic_amd/ics.py).  Same plugin interface as model/riemann: MODEL_SETUP (the runtime form of
! the compile-time PARAMETERs of solver/definition.f90 and model/<m>/src/defs.f90), CUSTOM_GRID
! (coords.f90:37), GET_MODEL (model/<m>/src/model.f90) and OPENFILE_CUSTOM (main.f90:112).  The
! in-loop hooks (EOS, CUSTOM_CHECKRHO, CUSTOM_SOURCE, CUSTOM_BOUNDARY) run on the GPU, selected
! by cfg%model.
!
! Run file (little-endian stream): 16 int32  nx ny nz coordinate bflag(6) recon model
!   geom_source_repeat precision max_steps solver (0: HLL) ; 7 float64 ggas cfl rho_floor eps_floor r_sh
!   total_time output_profiletime ; xf(nx+7) yf(ny+7) zf(nz+7) ; prim(8,nx+6,ny+6,nz+6) ;
!   eps(nx+6,ny+6,nz+6)
!-----------------------------------------------------------------------------------------------
MODULE MODEL_PLUGIN
USE ISO_C_BINDING
USE CUMC_IFACE
IMPLICIT NONE

CHARACTER(LEN=*), PARAMETER :: model_name = 'file'
REAL(C_DOUBLE), ALLOCATABLE, PRIVATE :: fx(:), fy(:), fz(:), p0(:,:,:,:), e0(:,:,:)

CONTAINS

SUBROUTINE MODEL_SETUP(path, cfg, max_steps, total_time, output_time, lo, hi)
  CHARACTER(*), INTENT(IN) :: path
  TYPE(cumc_config), INTENT(OUT) :: cfg
  INTEGER, INTENT(OUT) :: max_steps
  REAL(C_DOUBLE), INTENT(OUT) :: total_time, output_time, lo(3), hi(3)
  INTEGER(C_INT) :: hdr(16)
  REAL(C_DOUBLE) :: rh(7)
  INTEGER :: u, nx, ny, nz

  OPEN(NEWUNIT=u, FILE=path, ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD', ACTION='READ')
  READ(u) hdr
  READ(u) rh
  nx = hdr(1); ny = hdr(2); nz = hdr(3)
  cfg%nx = nx; cfg%ny = ny; cfg%nz = nz
  cfg%coordinate = hdr(4)
  cfg%boundary_flag = hdr(5:10)
  cfg%recon = hdr(11)
  cfg%model = hdr(12)
  cfg%geom_source_repeat = hdr(13)
  cfg%precision = hdr(14)
  max_steps = hdr(15)
  cfg%solver = hdr(16)
  cfg%ggas = rh(1); cfg%cfl = rh(2)
  cfg%rho_floor = rh(3); cfg%eps_floor = rh(4); cfg%r_sh = rh(5)
  total_time = rh(6)
  output_time = rh(7)
  ALLOCATE(fx(-3:nx+3), fy(-3:ny+3), fz(-3:nz+3))
  ALLOCATE(p0(1:8, -2:nx+3, -2:ny+3, -2:nz+3))
  ALLOCATE(e0(-2:nx+3, -2:ny+3, -2:nz+3))
  READ(u) fx
  READ(u) fy
  READ(u) fz
  READ(u) p0
  READ(u) e0
  CLOSE(u)
  ! the default grid is overridden by CUSTOM_GRID below
  lo = (/ fx(0), fy(0), fz(0) /)
  hi = (/ fx(nx), fy(ny), fz(nz) /)
END SUBROUTINE

! the run file's faces replace GETGRID's uniform ones
SUBROUTINE CUSTOM_GRID(cfg, xf, yf, zf)
  TYPE(cumc_config), INTENT(IN) :: cfg
  REAL(C_DOUBLE), INTENT(INOUT) :: xf(-3:), yf(-3:), zf(-3:)
  xf(-3:cfg%nx+3) = fx
  yf(-3:cfg%ny+3) = fy
  zf(-3:cfg%nz+3) = fz
END SUBROUTINE

SUBROUTINE GET_MODEL(cfg, xf, yf, zf, prim, eps)
  TYPE(cumc_config), INTENT(IN) :: cfg
  REAL(C_DOUBLE), INTENT(IN) :: xf(-3:), yf(-3:), zf(-3:)
  REAL(C_DOUBLE), INTENT(INOUT) :: prim(1:, -2:, -2:, -2:), eps(-2:, -2:, -2:)
  prim = p0
  eps = e0
  DEALLOCATE(p0, e0)
END SUBROUTINE

SUBROUTINE OPENFILE_CUSTOM()
END SUBROUTINE

END MODULE MODEL_PLUGIN
