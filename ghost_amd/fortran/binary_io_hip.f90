! binary_io_hip.f90 -- drop-in for GHOST's mpiio/binary_io.f90 and
! mpiio/binary_mod.f90 on one rank, over the C ABI of include/ghx.h.
!
! MODULE iovar and MODULE iompi keep GHOST's names and variables (iswap,
! oswap; disp, bmangle); TYPE(IOPLAN) keeps its components and adds the
! library's handle.  The external io_init/io_read/io_write/io_initc/
! io_readc/io_writec keep GHOST's argument lists and write the same files:
! the global array in Fortran order, raw native FP64, no header.  Host
! arrays are written directly; MODULE binary_io_hip adds io_read_dev and
! io_write_dev for fields resident on the GPU (TYPE(ghx_field)), staged
! through pinned buffers.
MODULE iovar
  USE iso_c_binding, ONLY: C_PTR, C_NULL_PTR
  IMPLICIT NONE
  TYPE IOPLAN
    INTEGER :: ksta, kend
    INTEGER :: nx, ny, nz
    INTEGER :: iotype = 0
    TYPE(C_PTR) :: h = C_NULL_PTR
  END TYPE IOPLAN
  INTEGER :: iswap = 0
  INTEGER :: oswap = 0
  SAVE
END MODULE iovar

MODULE iompi
  USE iso_c_binding, ONLY: C_LONG_LONG
  IMPLICIT NONE
  INTEGER, SAVE :: ioerr = 0
  INTEGER(KIND=C_LONG_LONG), SAVE :: disp = 0
  INTEGER, SAVE :: bmangle = 1
END MODULE iompi

MODULE binary_io_hip
  USE iso_c_binding
  USE fprecision
  USE iovar
  USE iompi
  USE ghx_c
  USE pseudospec_hip, ONLY: ghx_field
  IMPLICIT NONE

CONTAINS

  ! the library holds iswap, bmangle and disp per plan: pass the module values
  SUBROUTINE io_sync_options(plan)
    TYPE(IOPLAN), INTENT(IN) :: plan
    CALL ghx_check(ghx_io_set_options(plan%h, INT(iswap, C_INT), INT(bmangle, C_INT), disp), 'io options')
  END SUBROUTINE io_sync_options

  SUBROUTINE io_plan_create(myrank, n, ksta, kend, plan, cplx)
    INTEGER, INTENT(IN) :: myrank, n(3), ksta, kend
    TYPE(IOPLAN), INTENT(OUT) :: plan
    LOGICAL, INTENT(IN) :: cplx
    INTEGER(C_INT) :: nn(3)
    plan%nx = n(1)
    plan%ny = n(2)
    plan%nz = n(3)
    plan%ksta = ksta
    plan%kend = kend
    nn = INT(n, C_INT)
    IF (cplx) THEN
      CALL ghx_check(ghx_io_initc(plan%h, INT(myrank, C_INT), nn, INT(ksta, C_INT), INT(kend, C_INT)), 'io_initc')
    ELSE
      CALL ghx_check(ghx_io_init(plan%h, INT(myrank, C_INT), nn, INT(ksta, C_INT), INT(kend, C_INT)), 'io_init')
    END IF
  END SUBROUTINE io_plan_create

  ! io_read / io_write on a field resident on the GPU (real, R(nx,ny,nz))
  SUBROUTINE io_read_dev(unit, dir, fname, nmb, plan, var)
    INTEGER, INTENT(IN) :: unit
    CHARACTER(len=*), INTENT(IN) :: dir, fname, nmb
    TYPE(IOPLAN), INTENT(IN) :: plan
    TYPE(ghx_field), INTENT(IN) :: var
    CALL io_sync_options(plan)
    CALL ghx_check(ghx_io_read(plan%h, TRIM(dir)//C_NULL_CHAR, fname//C_NULL_CHAR, nmb//C_NULL_CHAR, var%p), &
                   'io_read: '//TRIM(dir)//'/'//fname)
  END SUBROUTINE io_read_dev

  SUBROUTINE io_write_dev(unit, dir, fname, nmb, plan, var)
    INTEGER, INTENT(IN) :: unit
    CHARACTER(len=*), INTENT(IN) :: dir, fname, nmb
    TYPE(IOPLAN), INTENT(IN) :: plan
    TYPE(ghx_field), INTENT(IN) :: var
    CALL io_sync_options(plan)
    CALL ghx_check(ghx_io_write(plan%h, TRIM(dir)//C_NULL_CHAR, fname//C_NULL_CHAR, nmb//C_NULL_CHAR, var%p), &
                   'io_write: '//TRIM(dir)//'/'//fname)
  END SUBROUTINE io_write_dev

END MODULE binary_io_hip

! ---- GHOST's external routines (mpiio/binary_io.f90) on host arrays -------

SUBROUTINE io_init(myrank, n, ksta, kend, plan)
  USE iovar
  USE binary_io_hip, ONLY: io_plan_create
  IMPLICIT NONE
  INTEGER, INTENT(IN) :: myrank, n(3), ksta, kend
  TYPE(IOPLAN), INTENT(OUT) :: plan
  CALL io_plan_create(myrank, n, ksta, kend, plan, .FALSE.)
END SUBROUTINE io_init

SUBROUTINE io_initc(myrank, n, ksta, kend, plan)
  USE iovar
  USE binary_io_hip, ONLY: io_plan_create
  IMPLICIT NONE
  INTEGER, INTENT(IN) :: myrank, n(3), ksta, kend
  TYPE(IOPLAN), INTENT(OUT) :: plan
  CALL io_plan_create(myrank, n, ksta, kend, plan, .TRUE.)
END SUBROUTINE io_initc

SUBROUTINE io_read(unit, dir, fname, nmb, plan, var)
  USE iso_c_binding
  USE fprecision
  USE iovar
  USE ghx_c
  USE binary_io_hip, ONLY: io_sync_options
  IMPLICIT NONE
  TYPE(IOPLAN), INTENT(IN) :: plan
  REAL(KIND=GP), INTENT(OUT), TARGET :: var(plan%nx, plan%ny, plan%ksta:plan%kend)
  INTEGER, INTENT(IN) :: unit
  CHARACTER(len=100), INTENT(IN) :: dir
  CHARACTER(len=*), INTENT(IN) :: nmb
  CHARACTER(len=*), INTENT(IN) :: fname
  CALL io_sync_options(plan)
  CALL ghx_check(ghx_io_read(plan%h, TRIM(dir)//C_NULL_CHAR, fname//C_NULL_CHAR, nmb//C_NULL_CHAR, C_LOC(var)), &
                 'io_read: cannot read '//TRIM(dir)//'/'//fname)
END SUBROUTINE io_read

SUBROUTINE io_write(unit, dir, fname, nmb, plan, var)
  USE iso_c_binding
  USE fprecision
  USE iovar
  USE ghx_c
  USE binary_io_hip, ONLY: io_sync_options
  IMPLICIT NONE
  TYPE(IOPLAN), INTENT(IN) :: plan
  REAL(KIND=GP), INTENT(INOUT), TARGET :: var(plan%nx, plan%ny, plan%ksta:plan%kend)
  INTEGER, INTENT(IN) :: unit
  CHARACTER(len=100), INTENT(IN) :: dir
  CHARACTER(len=*), INTENT(IN) :: nmb
  CHARACTER(len=*), INTENT(IN) :: fname
  CALL io_sync_options(plan)
  CALL ghx_check(ghx_io_write(plan%h, TRIM(dir)//C_NULL_CHAR, fname//C_NULL_CHAR, nmb//C_NULL_CHAR, C_LOC(var)), &
                 'io_write: '//TRIM(dir)//'/'//fname)
END SUBROUTINE io_write

SUBROUTINE io_readc(unit, dir, fname, nmb, plan, var)
  USE iso_c_binding
  USE fprecision
  USE iovar
  USE ghx_c
  USE binary_io_hip, ONLY: io_sync_options
  IMPLICIT NONE
  TYPE(IOPLAN), INTENT(IN) :: plan
  COMPLEX(KIND=GP), INTENT(OUT), TARGET :: var(plan%nx, plan%ny, plan%ksta:plan%kend)
  INTEGER, INTENT(IN) :: unit
  CHARACTER(len=100), INTENT(IN) :: dir
  CHARACTER(len=*), INTENT(IN) :: nmb
  CHARACTER(len=*), INTENT(IN) :: fname
  CALL io_sync_options(plan)
  CALL ghx_check(ghx_io_readc(plan%h, TRIM(dir)//C_NULL_CHAR, fname//C_NULL_CHAR, nmb//C_NULL_CHAR, C_LOC(var)), &
                 'io_readc: cannot read '//TRIM(dir)//'/'//fname)
END SUBROUTINE io_readc

SUBROUTINE io_writec(unit, dir, fname, nmb, plan, var)
  USE iso_c_binding
  USE fprecision
  USE iovar
  USE ghx_c
  USE binary_io_hip, ONLY: io_sync_options
  IMPLICIT NONE
  TYPE(IOPLAN), INTENT(IN) :: plan
  COMPLEX(KIND=GP), INTENT(IN), TARGET :: var(plan%nx, plan%ny, plan%ksta:plan%kend)
  INTEGER, INTENT(IN) :: unit
  CHARACTER(len=100), INTENT(IN) :: dir
  CHARACTER(len=*), INTENT(IN) :: nmb
  CHARACTER(len=*), INTENT(IN) :: fname
  CALL io_sync_options(plan)
  CALL ghx_check(ghx_io_writec(plan%h, TRIM(dir)//C_NULL_CHAR, fname//C_NULL_CHAR, nmb//C_NULL_CHAR, C_LOC(var)), &
                 'io_writec: '//TRIM(dir)//'/'//fname)
END SUBROUTINE io_writec
