! fftp_hip.f90 -- GHOST's fftp interface on the gfx950 library, one rank.
!
! Drop-in for the four files of a GHOST FFT backend directory
! (fftp-3/fprecision_mod.fpp, fftp_mod.fpp, fftp3D.fpp, fftp.f90):
!   MODULE fprecision (GP), MODULE fftplans (TYPE FFTPLAN, FFTW_* constants,
!   timers ffttime/tratime/comtime/tottime), MODULE threads (nth),
!   SUBROUTINE fftp3d_init_threads, fftp3d_create_plan, fftp3d_destroy_plan,
!   fftp3d_real_to_complex, fftp3d_complex_to_real, range.
! The transforms keep GHOST's layouts and argument order: REAL(GP)
! in(nx,ny,nz) -> COMPLEX(GP) out(nz,ny,nx/2+1), unnormalised, on host
! arrays (copied to and from the device per call).  fftp3d_*_dev take device
! pointers for callers that keep their fields resident.  Unlike GHOST's
! complex_to_real, the input is preserved.
MODULE fprecision
  USE iso_c_binding, ONLY: C_DOUBLE
  INTEGER, PARAMETER :: GP = C_DOUBLE
END MODULE fprecision

MODULE threads
  INTEGER :: nth = 1
END MODULE threads

MODULE fftplans
  USE iso_c_binding
  USE fprecision
  IMPLICIT NONE
  ! fftp-3/fftp_mod.fpp:27-28 and FFTW's planner flags (ignored on the GPU)
  INTEGER, PARAMETER :: FFTW_REAL_TO_COMPLEX = -1
  INTEGER, PARAMETER :: FFTW_COMPLEX_TO_REAL = 1
  INTEGER, PARAMETER :: FFTW_MEASURE = 0
  INTEGER, PARAMETER :: FFTW_EXHAUSTIVE = 8
  INTEGER, PARAMETER :: FFTW_PATIENT = 32
  INTEGER, PARAMETER :: FFTW_ESTIMATE = 64
  DOUBLE PRECISION :: comtime = 0.0
  DOUBLE PRECISION :: ffttime = 0.0
  DOUBLE PRECISION :: tratime = 0.0
  DOUBLE PRECISION :: tottime = 0.0
  TYPE FFTPLAN
    TYPE(C_PTR) :: h = C_NULL_PTR
    INTEGER :: nx = 0, ny = 0, nz = 0
    INTEGER :: fftdir = 0
  END TYPE FFTPLAN
  SAVE
END MODULE fftplans

SUBROUTINE fftp3d_init_threads(err)
  ! fftp-3/fftp3D.fpp:35-54: nothing to set up; the GPU needs no host threads.
  IMPLICIT NONE
  INTEGER, INTENT(OUT) :: err
  err = 1
END SUBROUTINE fftp3d_init_threads

SUBROUTINE fftp3d_create_plan(plan, n, fftdir, flags)
  ! fftp-3/fftp3D.fpp:56-125
  USE ghx_c
  USE fftplans
  IMPLICIT NONE
  TYPE(FFTPLAN), INTENT(OUT) :: plan
  INTEGER, INTENT(IN) :: n(3)
  INTEGER, INTENT(IN) :: fftdir, flags
  CALL ghx_check(ghx_plan_create(plan%h, INT(n(1), C_INT), INT(n(2), C_INT), INT(n(3), C_INT), &
                                 INT(fftdir, C_INT), INT(flags, C_INT)), 'fftp3d_create_plan')
  plan%nx = n(1)
  plan%ny = n(2)
  plan%nz = n(3)
  plan%fftdir = fftdir
  CALL ghx_check(ghx_plan_reset_timers(plan%h), 'fftp3d_create_plan')
END SUBROUTINE fftp3d_create_plan

SUBROUTINE fftp3d_destroy_plan(plan)
  ! fftp-3/fftp3D.fpp:126-156
  USE ghx_c
  USE fftplans
  IMPLICIT NONE
  TYPE(FFTPLAN), INTENT(INOUT) :: plan
  CALL ghx_check(ghx_plan_destroy(plan%h), 'fftp3d_destroy_plan')
  plan%h = C_NULL_PTR
END SUBROUTINE fftp3d_destroy_plan

SUBROUTINE fftp_hip_add_times(plan)
  ! Accumulates the plan's transform time into ffttime/tottime as the
  ! GTStart/GTAcc pairs of fftp-3/fftp3D.fpp:244-312 do (no transposes or
  ! communication on one GPU: tratime and comtime stay 0).
  USE ghx_c
  USE fftplans
  IMPLICIT NONE
  TYPE(FFTPLAN), INTENT(IN) :: plan
  REAL(C_DOUBLE) :: f, t, c, tot
  CALL ghx_check(ghx_plan_timers(plan%h, f, t, c, tot), 'fftp timers')
  ffttime = ffttime + f
  tottime = tottime + tot
  CALL ghx_check(ghx_plan_reset_timers(plan%h), 'fftp timers')
END SUBROUTINE fftp_hip_add_times

SUBROUTINE fftp3d_real_to_complex(plan, in, out, comm)
  ! fftp-3/fftp3D.fpp:204-315 on host arrays
  USE ghx_c
  USE fftplans
  IMPLICIT NONE
  TYPE(FFTPLAN), INTENT(IN) :: plan
  REAL(KIND=GP), INTENT(IN), TARGET :: in(plan%nx, plan%ny, plan%nz)
  COMPLEX(KIND=GP), INTENT(OUT), TARGET :: out(plan%nz, plan%ny, plan%nx/2 + 1)
  INTEGER, INTENT(IN) :: comm
  CALL ghx_check(ghx_r2c_host(plan%h, C_LOC(in), C_LOC(out)), 'fftp3d_real_to_complex')
  CALL fftp_hip_add_times(plan)
END SUBROUTINE fftp3d_real_to_complex

SUBROUTINE fftp3d_complex_to_real(plan, in, out, comm)
  ! fftp-3/fftp3D.fpp:318-431 on host arrays (the input is preserved)
  USE ghx_c
  USE fftplans
  IMPLICIT NONE
  TYPE(FFTPLAN), INTENT(IN) :: plan
  COMPLEX(KIND=GP), INTENT(IN), TARGET :: in(plan%nz, plan%ny, plan%nx/2 + 1)
  REAL(KIND=GP), INTENT(OUT), TARGET :: out(plan%nx, plan%ny, plan%nz)
  INTEGER, INTENT(IN) :: comm
  CALL ghx_check(ghx_c2r_host(plan%h, C_LOC(in), C_LOC(out)), 'fftp3d_complex_to_real')
  CALL fftp_hip_add_times(plan)
END SUBROUTINE fftp3d_complex_to_real

SUBROUTINE fftp3d_real_to_complex_dev(plan, in, out)
  ! as fftp3d_real_to_complex on device pointers, asynchronous
  USE ghx_c
  USE fftplans
  IMPLICIT NONE
  TYPE(FFTPLAN), INTENT(IN) :: plan
  TYPE(C_PTR), INTENT(IN) :: in, out
  CALL ghx_check(ghx_r2c(plan%h, in, out), 'fftp3d_real_to_complex_dev')
END SUBROUTINE fftp3d_real_to_complex_dev

SUBROUTINE fftp3d_complex_to_real_dev(plan, in, out)
  USE ghx_c
  USE fftplans
  IMPLICIT NONE
  TYPE(FFTPLAN), INTENT(IN) :: plan
  TYPE(C_PTR), INTENT(IN) :: in, out
  CALL ghx_check(ghx_c2r(plan%h, in, out), 'fftp3d_complex_to_real_dev')
END SUBROUTINE fftp3d_complex_to_real_dev

SUBROUTINE range(n1, n2, nprocs, irank, ista, iend)
  ! fftp-3/fftp.f90:17-48: block distribution of n1..n2 over nprocs ranks
  IMPLICIT NONE
  INTEGER, INTENT(IN) :: n1, n2, nprocs, irank
  INTEGER, INTENT(OUT) :: ista, iend
  INTEGER :: iwork1, iwork2
  iwork1 = (n2 - n1 + 1) / nprocs
  iwork2 = MOD(n2 - n1 + 1, nprocs)
  ista = irank * iwork1 + n1 + MIN(irank, iwork2)
  iend = ista + iwork1 - 1
  IF (iwork2 .GT. irank) iend = iend + 1
END SUBROUTINE range
