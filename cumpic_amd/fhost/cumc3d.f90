!-----------------------------------------------------------------------------------------------
! cumc3d.f90 -- Fortran host driver of the MI355X CUMC3D step.
!
! Same shape as the reference driver (solver/main.f90): the model plugin's parameters, GETGRID's
! uniform faces and the plugin's CUSTOM_GRID (solver/coords.f90:11-41), the plugin's GET_MODEL
! (solver/initial.f90:7-75), initial_update, the first profile and OPENFILE_CUSTOM
! (main.f90:100-112), then DO WHILE (global_time < total_time): finddt, clamp dt to the end time,
! Rungekutta, advance, print "n_step, dt, global_time"; finally report the divergence of B.  The
! hot path runs in libcumc.so through the ISO_C_BINDING interface of cumc_iface.f90.
!
! The model plugin (MODULE MODEL_PLUGIN: MODEL_SETUP, CUSTOM_GRID, GET_MODEL, OPENFILE_CUSTOM) is
! chosen at build time from model/<name>/source.list, as the reference's makefile takes
! model/<name>/src/source.list: model/riemann computes the riemann plugin's named problems,
! model/file reads every input from a run file.
!
! Usage: cumc3d_<model> [--model-only] <parameter or run file> <output file> [<profile directory>]
! (--model-only: write the plugin's grid and initial state, int32 nx, ny, nz ; xf ; yf ; zf ; prim ;
!  eps, and stop before any device call)
! Output (stream): int32 n_step ; float64 global_time, divb_signed, divb_abs ;
!                  prim(8,nx+6,ny+6,nz+6) ; eps(nx+6,ny+6,nz+6) ; dt(n_step) ;
!                  then the state GET_MODEL produced: int32 nx, ny, nz ; prim ; eps
! With a profile directory, HDF5 profiles star_weno_data_<n_iter>.hdf5 in the reference's layout
! (dataio.f90) are written there at the reference's points (main.f90:107-110, 150-160, 165-168):
! after initial_update, whenever |global_time - time of the last profile| >= output_profiletime,
! and at the end.
!-----------------------------------------------------------------------------------------------
PROGRAM CUMC3D
USE ISO_C_BINDING
USE CUMC_IFACE
USE MODEL_PLUGIN
IMPLICIT NONE

TYPE(cumc_config) :: cfg
TYPE(C_PTR) :: ctx
REAL(C_DOUBLE), ALLOCATABLE :: xf(:), yf(:), zf(:), prim(:,:,:,:), eps(:,:,:), bcell(:,:,:,:)
REAL(C_DOUBLE), ALLOCATABLE :: dts(:), prim0(:,:,:,:), eps0(:,:,:)
REAL(C_DOUBLE) :: dt, global_time, total_time, maxs, maxa, output_time, output_last, lo(3), hi(3)
INTEGER :: max_steps, n_step, u, n_iter, j
INTEGER(C_INT) :: ierr
CHARACTER(LEN=1024) :: infile, outfile, profdir
LOGICAL :: model_only

CALL GET_COMMAND_ARGUMENT(1, infile)
model_only = TRIM(infile) == '--model-only'
j = 0
IF (model_only) j = 1
CALL GET_COMMAND_ARGUMENT(1 + j, infile)
CALL GET_COMMAND_ARGUMENT(2 + j, outfile)
CALL GET_COMMAND_ARGUMENT(3 + j, profdir)
IF (LEN_TRIM(infile) == 0 .OR. LEN_TRIM(outfile) == 0) THEN
  WRITE(*,*) 'usage: cumc3d_', model_name, ' <parameter or run file> <output file> [<profile directory>]'
  STOP 2
END IF

! setup (main.f90:57-99, initial.f90:7-75): the plugin's parameters, GETGRID's uniform faces
! x_start + j dx (coords.f90:24-32) and the plugin's CUSTOM_GRID, then GET_MODEL on zeroed arrays
CALL MODEL_SETUP(TRIM(infile), cfg, max_steps, total_time, output_time, lo, hi)
ALLOCATE(xf(-3:cfg%nx+3), yf(-3:cfg%ny+3), zf(-3:cfg%nz+3))
DO j = -3, cfg%nx + 3
  xf(j) = lo(1) + DBLE(j)*((hi(1) - lo(1))/DBLE(cfg%nx))
END DO
DO j = -3, cfg%ny + 3
  yf(j) = lo(2) + DBLE(j)*((hi(2) - lo(2))/DBLE(cfg%ny))
END DO
DO j = -3, cfg%nz + 3
  zf(j) = lo(3) + DBLE(j)*((hi(3) - lo(3))/DBLE(cfg%nz))
END DO
CALL CUSTOM_GRID(cfg, xf, yf, zf)
ALLOCATE(prim(1:8, -2:cfg%nx+3, -2:cfg%ny+3, -2:cfg%nz+3), eps(-2:cfg%nx+3, -2:cfg%ny+3, -2:cfg%nz+3))
prim = 0.0D0
eps = 0.0D0
CALL GET_MODEL(cfg, xf, yf, zf, prim, eps)
ALLOCATE(prim0, SOURCE=prim)
ALLOCATE(eps0, SOURCE=eps)
IF (model_only) THEN
  ! the plugin's grid and initial state only (no device): int32 nx, ny, nz ; xf ; yf ; zf ; prim ; eps
  OPEN(NEWUNIT=u, FILE=TRIM(outfile), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE', ACTION='WRITE')
  WRITE(u) cfg%nx, cfg%ny, cfg%nz
  WRITE(u) xf, yf, zf
  WRITE(u) prim0
  WRITE(u) eps0
  CLOSE(u)
  STOP
END IF
ctx = C_NULL_PTR
ierr = cumc_create(cfg, ctx)
CALL CUMC_CHECK(ierr, 'cumc_create', ctx)
CALL CUMC_CHECK(cumc_set_grid(ctx, xf, yf, zf), 'cumc_set_grid', ctx)
CALL CUMC_CHECK(cumc_upload(ctx, prim, eps), 'cumc_upload', ctx)
CALL CUMC_CHECK(cumc_initial_update(ctx), 'cumc_initial_update', ctx)
ALLOCATE(dts(MAX(max_steps, 1)))
ALLOCATE(bcell(1:3, -2:cfg%nx+3, -2:cfg%ny+3, -2:cfg%nz+3))

! print primitive profiles (main.f90:107-110), custom analysis files (main.f90:112)
n_iter = 0
output_last = 0.0D0
CALL PRINT_PROFILE()
CALL OPENFILE_CUSTOM()

WRITE(*,*) 'We now start the time evolution'
WRITE(*,*) '------------------------------------'
WRITE(*,*) 'iteration, timestep, simulation time'
WRITE(*,*) '------------------------------------'

! time loop (main.f90:127-163) with a step cap
global_time = 0.0D0
n_step = 0
DO WHILE (global_time < total_time .AND. n_step < max_steps)
  CALL CUMC_CHECK(cumc_finddt(ctx, dt), 'cumc_finddt', ctx)
  IF (global_time + dt > total_time) THEN
    dt = total_time - global_time
  END IF
  CALL CUMC_CHECK(cumc_rungekutta(ctx, dt), 'cumc_rungekutta', ctx)
  n_step = n_step + 1
  global_time = global_time + dt
  dts(n_step) = dt
  WRITE(*,*) n_step, dt, global_time
  ! data I/O (main.f90:150-160)
  IF (ABS(global_time - output_last) >= output_time) THEN
    output_last = global_time
    CALL PRINT_PROFILE()
  END IF
END DO

! output again (main.f90:165-168), without advancing the counter
IF (LEN_TRIM(profdir) > 0) THEN
  CALL CUMC_CHECK(cumc_download(ctx, prim, eps, bcell), 'cumc_download', ctx)
  IF (cumc_h5_profile(TRIM(profdir) // C_NULL_CHAR, INT(n_iter, C_INT), cfg%nx, cfg%ny, cfg%nz, &
                      xf, yf, zf, prim, eps) /= 0) THEN
    WRITE(*,*) 'cumc3d: HDF5 profile output failed (libhdf5 not found or directory not writable)'
    STOP 1
  END IF
END IF
CALL CUMC_CHECK(cumc_download(ctx, prim, eps, bcell), 'cumc_download', ctx)
CALL CUMC_CHECK(cumc_find_divb(ctx, maxs, maxa), 'cumc_find_divb', ctx)
WRITE(*,*)
WRITE(*,*) 'Maximum divergence B', maxs, ' max |div B|', maxa
WRITE(*,*)

OPEN(NEWUNIT=u, FILE=TRIM(outfile), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE', ACTION='WRITE')
WRITE(u) INT(n_step, C_INT)
WRITE(u) global_time, maxs, maxa
WRITE(u) prim
WRITE(u) eps
IF (n_step > 0) WRITE(u) dts(1:n_step)
WRITE(u) cfg%nx, cfg%ny, cfg%nz
WRITE(u) prim0
WRITE(u) eps0
CLOSE(u)

ierr = cumc_destroy(ctx)
WRITE(*,*) '----------------'
WRITE(*,*) '-Simulation End-'
WRITE(*,*) '----------------'

CONTAINS

! print_hydroprofile (dataio.f90:7-77): the state to the host, one HDF5 file, n_iter + 1
SUBROUTINE PRINT_PROFILE()
  IF (LEN_TRIM(profdir) == 0) RETURN
  CALL CUMC_CHECK(cumc_download(ctx, prim, eps, bcell), 'cumc_download', ctx)
  IF (cumc_h5_profile(TRIM(profdir) // C_NULL_CHAR, INT(n_iter, C_INT), cfg%nx, cfg%ny, cfg%nz, &
                      xf, yf, zf, prim, eps) /= 0) THEN
    WRITE(*,*) 'cumc3d: HDF5 profile output failed (libhdf5 not found or directory not writable)'
    STOP 1
  END IF
  n_iter = n_iter + 1
END SUBROUTINE
END PROGRAM CUMC3D
