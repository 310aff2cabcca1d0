!-----------------------------------------------------------------------------------------------
! model_file.f90 -- model plugin of the Fortran host: GET_MODEL reads a run file written by the
! initial-condition tool (cumpic_amd/fhost/runfile.py) instead of computing the state in place
! (the reference's GET_MODEL, model/<m>/src/model.f90, fills prim/eps from compile-time
! PARAMETERs).  The in-loop hooks (EOS, CUSTOM_CHECKRHO, CUSTOM_SOURCE, CUSTOM_BOUNDARY) run on
! the GPU and are selected by cfg%model.
!
! Run file (little-endian stream): 16 int32  nx ny nz coordinate bflag(6) recon model
!   geom_source_repeat precision max_steps solver (0: HLL) ; 7 float64 ggas cfl rho_floor eps_floor r_sh
!   total_time output_profiletime ; xf(nx+7) yf(ny+7) zf(nz+7) ; prim(8,nx+6,ny+6,nz+6) ;
!   eps(nx+6,ny+6,nz+6)
!-----------------------------------------------------------------------------------------------
MODULE MODEL_FILE
USE ISO_C_BINDING
USE CUMC_IFACE
IMPLICIT NONE

CONTAINS

SUBROUTINE GET_MODEL(path, cfg, xf, yf, zf, prim, eps, max_steps, total_time, output_time)
  CHARACTER(*), INTENT(IN) :: path
  TYPE(cumc_config), INTENT(OUT) :: cfg
  REAL(C_DOUBLE), ALLOCATABLE, INTENT(OUT) :: xf(:), yf(:), zf(:)
  REAL(C_DOUBLE), ALLOCATABLE, INTENT(OUT) :: prim(:,:,:,:), eps(:,:,:)
  INTEGER, INTENT(OUT) :: max_steps
  REAL(C_DOUBLE), INTENT(OUT) :: total_time, output_time
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
  ALLOCATE(xf(-3:nx+3), yf(-3:ny+3), zf(-3:nz+3))
  ALLOCATE(prim(1:8, -2:nx+3, -2:ny+3, -2:nz+3))
  ALLOCATE(eps(-2:nx+3, -2:ny+3, -2:nz+3))
  READ(u) xf
  READ(u) yf
  READ(u) zf
  READ(u) prim
  READ(u) eps
  CLOSE(u)
END SUBROUTINE

END MODULE MODEL_FILE
