! pseudospec_hip_host.f90 -- host-array drop-ins for GHOST's pseudospectral
! routines: same names, argument order and array shapes as
! pseudo/pseudospec3D_hd.f90, _filt.f90 and module_dns.f90 (one rank:
! COMPLEX(GP) a(nz,ny,nx/2+1)).  Each call copies its fields to the device,
! runs the gfx950 kernel and copies the results back, so callers link these
! in place of GHOST's versions without a source change; for speed, USE
! pseudospec_hip and keep the fields on the device instead.
! ghx_pseudospec_init must have been called.
MODULE pseudospec_hip_staging
  USE iso_c_binding
  USE ghx_c
  USE pseudospec_hip, ONLY: gnc, ghx_grid_h
  IMPLICIT NONE
  INTEGER, PARAMETER :: NSTAGE = 7
  TYPE(C_PTR), SAVE :: dbuf(NSTAGE) = C_NULL_PTR
  INTEGER(C_SIZE_T), SAVE :: dbytes = 0
CONTAINS
  SUBROUTINE staging_ready()
    INTEGER :: k
    IF (.NOT. C_ASSOCIATED(ghx_grid_h)) THEN
      WRITE(*, '(A)') 'ghx: call ghx_pseudospec_init before the pseudospectral routines'
      STOP 1
    END IF
    IF (dbytes .EQ. gnc * 16_C_SIZE_T) RETURN
    DO k = 1, NSTAGE
      IF (C_ASSOCIATED(dbuf(k))) CALL ghx_check(ghx_free(dbuf(k)), 'staging free')
      CALL ghx_check(ghx_malloc(dbuf(k), gnc * 16_C_SIZE_T), 'staging alloc')
    END DO
    dbytes = gnc * 16_C_SIZE_T
  END SUBROUTINE staging_ready
  SUBROUTINE put(k, host)
    INTEGER, INTENT(IN) :: k
    TYPE(C_PTR), INTENT(IN) :: host
    CALL ghx_check(ghx_memcpy_h2d(dbuf(k), host, dbytes), 'copy to device')
  END SUBROUTINE put
  SUBROUTINE get(k, host)
    INTEGER, INTENT(IN) :: k
    TYPE(C_PTR), INTENT(IN) :: host
    CALL ghx_check(ghx_memcpy_d2h(host, dbuf(k), dbytes), 'copy to host')
  END SUBROUTINE get
END MODULE pseudospec_hip_staging

SUBROUTINE derivk3(a, b, dir)
  USE pseudospec_hip_staging
  USE fprecision
  IMPLICIT NONE
  COMPLEX(KIND=GP), INTENT(IN), TARGET :: a(*)
  COMPLEX(KIND=GP), INTENT(OUT), TARGET :: b(*)
  INTEGER, INTENT(IN) :: dir
  CALL staging_ready()
  CALL put(1, C_LOC(a))
  CALL ghx_check(ghx_derivk3(ghx_grid_h, dbuf(1), dbuf(2), INT(dir, C_INT)), 'derivk3')
  CALL get(2, C_LOC(b))
END SUBROUTINE derivk3

SUBROUTINE laplak3(a, b)
  USE pseudospec_hip_staging
  USE fprecision
  IMPLICIT NONE
  COMPLEX(KIND=GP), INTENT(IN), TARGET :: a(*)
  COMPLEX(KIND=GP), INTENT(OUT), TARGET :: b(*)
  CALL staging_ready()
  CALL put(1, C_LOC(a))
  CALL ghx_check(ghx_laplak3(ghx_grid_h, dbuf(1), dbuf(2)), 'laplak3')
  CALL get(2, C_LOC(b))
END SUBROUTINE laplak3

SUBROUTINE rotor3(a, b, c, dir)
  USE pseudospec_hip_staging
  USE fprecision
  IMPLICIT NONE
  COMPLEX(KIND=GP), INTENT(IN), TARGET :: a(*), b(*)
  COMPLEX(KIND=GP), INTENT(OUT), TARGET :: c(*)
  INTEGER, INTENT(IN) :: dir
  CALL staging_ready()
  CALL put(1, C_LOC(a))
  CALL put(2, C_LOC(b))
  CALL ghx_check(ghx_rotor3(ghx_grid_h, dbuf(1), dbuf(2), dbuf(3), INT(dir, C_INT)), 'rotor3')
  CALL get(3, C_LOC(c))
END SUBROUTINE rotor3

SUBROUTINE prodre3(a, b, c, d, e, f)
  USE pseudospec_hip_staging
  USE fprecision
  IMPLICIT NONE
  COMPLEX(KIND=GP), INTENT(IN), TARGET :: a(*), b(*), c(*)
  COMPLEX(KIND=GP), INTENT(OUT), TARGET :: d(*), e(*), f(*)
  CALL staging_ready()
  CALL put(1, C_LOC(a))
  CALL put(2, C_LOC(b))
  CALL put(3, C_LOC(c))
  CALL ghx_check(ghx_prodre3(ghx_grid_h, dbuf(1), dbuf(2), dbuf(3), dbuf(4), dbuf(5), dbuf(6)), 'prodre3')
  CALL get(4, C_LOC(d))
  CALL get(5, C_LOC(e))
  CALL get(6, C_LOC(f))
END SUBROUTINE prodre3

SUBROUTINE nonlhd3(a, b, c, g, dir)
  USE pseudospec_hip_staging
  USE fprecision
  IMPLICIT NONE
  COMPLEX(KIND=GP), INTENT(IN), TARGET :: a(*), b(*), c(*)
  COMPLEX(KIND=GP), INTENT(OUT), TARGET :: g(*)
  INTEGER, INTENT(IN) :: dir
  CALL staging_ready()
  CALL put(1, C_LOC(a))
  CALL put(2, C_LOC(b))
  CALL put(3, C_LOC(c))
  CALL ghx_check(ghx_nonlhd3(ghx_grid_h, dbuf(1), dbuf(2), dbuf(3), dbuf(4), INT(dir, C_INT)), 'nonlhd3')
  CALL get(4, C_LOC(g))
END SUBROUTINE nonlhd3

SUBROUTINE energy(a, b, c, d, kin)
  USE pseudospec_hip_staging
  USE fprecision
  IMPLICIT NONE
  COMPLEX(KIND=GP), INTENT(IN), TARGET :: a(*), b(*), c(*)
  DOUBLE PRECISION, INTENT(OUT) :: d
  INTEGER, INTENT(IN) :: kin
  REAL(C_DOUBLE) :: r
  CALL staging_ready()
  CALL put(1, C_LOC(a))
  CALL put(2, C_LOC(b))
  CALL put(3, C_LOC(c))
  CALL ghx_check(ghx_energy(ghx_grid_h, dbuf(1), dbuf(2), dbuf(3), INT(kin, C_INT), r), 'energy')
  d = r
END SUBROUTINE energy

SUBROUTINE helicity(a, b, c, d)
  USE pseudospec_hip_staging
  USE fprecision
  IMPLICIT NONE
  COMPLEX(KIND=GP), INTENT(IN), TARGET :: a(*), b(*), c(*)
  DOUBLE PRECISION, INTENT(OUT) :: d
  REAL(C_DOUBLE) :: r
  CALL staging_ready()
  CALL put(1, C_LOC(a))
  CALL put(2, C_LOC(b))
  CALL put(3, C_LOC(c))
  CALL ghx_check(ghx_helicity(ghx_grid_h, dbuf(1), dbuf(2), dbuf(3), r), 'helicity')
  d = r
END SUBROUTINE helicity

SUBROUTINE cross(a, b, c, d, e, f, g, kin)
  USE pseudospec_hip_staging
  USE fprecision
  IMPLICIT NONE
  COMPLEX(KIND=GP), INTENT(IN), TARGET :: a(*), b(*), c(*), d(*), e(*), f(*)
  DOUBLE PRECISION, INTENT(OUT) :: g
  INTEGER, INTENT(IN) :: kin
  REAL(C_DOUBLE) :: r
  CALL staging_ready()
  CALL put(1, C_LOC(a))
  CALL put(2, C_LOC(b))
  CALL put(3, C_LOC(c))
  CALL put(4, C_LOC(d))
  CALL put(5, C_LOC(e))
  CALL put(6, C_LOC(f))
  CALL ghx_check(ghx_cross(ghx_grid_h, dbuf(1), dbuf(2), dbuf(3), dbuf(4), dbuf(5), dbuf(6), INT(kin, C_INT), r), &
                 'cross')
  g = r
END SUBROUTINE cross

SUBROUTINE phaseshift(a, cdump)
  USE pseudospec_hip_staging
  USE fprecision
  IMPLICIT NONE
  COMPLEX(KIND=GP), INTENT(INOUT), TARGET :: a(*)
  COMPLEX(KIND=GP), INTENT(IN) :: cdump
  CALL staging_ready()
  CALL put(1, C_LOC(a))
  CALL ghx_check(ghx_phaseshift(ghx_grid_h, dbuf(1), REAL(REAL(cdump), C_DOUBLE), REAL(AIMAG(cdump), C_DOUBLE)), &
                 'phaseshift')
  CALL get(1, C_LOC(a))
END SUBROUTINE phaseshift

SUBROUTINE normalize(fx, fy, fz, f0, kin, comm)
  USE pseudospec_hip_staging
  USE fprecision
  IMPLICIT NONE
  COMPLEX(KIND=GP), INTENT(INOUT), TARGET :: fx(*), fy(*), fz(*)
  REAL(KIND=GP), INTENT(IN) :: f0
  INTEGER, INTENT(IN) :: kin, comm
  CALL staging_ready()
  CALL put(1, C_LOC(fx))
  CALL put(2, C_LOC(fy))
  CALL put(3, C_LOC(fz))
  CALL ghx_check(ghx_normalize(ghx_grid_h, dbuf(1), dbuf(2), dbuf(3), REAL(f0, C_DOUBLE), INT(kin, C_INT)), &
                 'normalize')
  CALL get(1, C_LOC(fx))
  CALL get(2, C_LOC(fy))
  CALL get(3, C_LOC(fz))
END SUBROUTINE normalize

SUBROUTINE sharp(a, b, delta)
  USE pseudospec_hip_staging
  USE fprecision
  IMPLICIT NONE
  COMPLEX(KIND=GP), INTENT(IN), TARGET :: a(*)
  COMPLEX(KIND=GP), INTENT(OUT), TARGET :: b(*)
  REAL(KIND=GP), INTENT(IN) :: delta
  CALL staging_ready()
  CALL put(1, C_LOC(a))
  CALL ghx_check(ghx_sharp(ghx_grid_h, dbuf(1), dbuf(2), REAL(delta, C_DOUBLE)), 'sharp')
  CALL get(2, C_LOC(b))
END SUBROUTINE sharp
