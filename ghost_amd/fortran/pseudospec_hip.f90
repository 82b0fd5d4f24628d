! pseudospec_hip.f90 -- GHOST's pseudospectral routines on device-resident
! fields of the gfx950 library, one rank.
!
! A source file that USEs pseudospec_hip gets module procedures with GHOST's
! names and argument order (pseudo/pseudospec3D_hd.f90, _mhd.f90, _filt.f90,
! module_dns.f90) whose field arguments are TYPE(ghx_field), i.e. device
! arrays, so a solver keeps its fields in HBM across substeps.  Files that do
! not USE it keep calling the external host-array versions of
! pseudospec_hip_host.f90.  ghx_pseudospec_init sets up the grid (the
! wavenumbers and dealias bounds of main3D.fpp:737-749, 1727-1800) once.
MODULE pseudospec_hip
  USE iso_c_binding
  USE fprecision
  USE ghx_c
  IMPLICIT NONE

  TYPE ghx_field
    TYPE(C_PTR) :: p = C_NULL_PTR
    LOGICAL :: spectral = .TRUE.
  END TYPE ghx_field

  TYPE(C_PTR), SAVE :: ghx_grid_h = C_NULL_PTR
  INTEGER, SAVE :: gnx = 0, gny = 0, gnz = 0
  INTEGER(C_SIZE_T), SAVE :: gnc = 0, gnr = 0
  ! shell width and count of the 3D spectra, wavenumber steps of the 1D ones
  INTEGER, SAVE :: gnmax = 0
  DOUBLE PRECISION, SAVE :: gDkk = 1.0D0, gDk(3) = 1.0D0

CONTAINS

  SUBROUTINE ghx_pseudospec_init(nx, ny, nz, Lx, Ly, Lz, arbsize, Dkk)
    INTEGER, INTENT(IN) :: nx, ny, nz, arbsize
    REAL(KIND=GP), INTENT(IN) :: Lx, Ly, Lz, Dkk
    REAL(C_DOUBLE) :: kmax, tiny, dk
    INTEGER(C_INT) :: nmax, anis
    IF (C_ASSOCIATED(ghx_grid_h)) CALL ghx_check(ghx_grid_destroy(ghx_grid_h), 'grid destroy')
    CALL ghx_check(ghx_grid_create(ghx_grid_h, INT(nx, C_INT), INT(ny, C_INT), INT(nz, C_INT), &
                                   REAL(Lx, C_DOUBLE), REAL(Ly, C_DOUBLE), REAL(Lz, C_DOUBLE), &
                                   INT(arbsize, C_INT), REAL(Dkk, C_DOUBLE)), 'ghx_pseudospec_init')
    gnx = nx
    gny = ny
    gnz = nz
    gnc = INT(nx/2 + 1, C_SIZE_T) * INT(ny, C_SIZE_T) * INT(nz, C_SIZE_T)
    gnr = INT(nx, C_SIZE_T) * INT(ny, C_SIZE_T) * INT(nz, C_SIZE_T)
    CALL ghx_check(ghx_grid_info(ghx_grid_h, kmax, tiny, nmax, dk, anis), 'ghx_pseudospec_init')
    gnmax = nmax
    gDkk = dk
    gDk = 1.0D0
    IF (arbsize .EQ. 1) gDk = (/1.0D0/Lx, 1.0D0/Ly, 1.0D0/Lz/)
  END SUBROUTINE ghx_pseudospec_init

  SUBROUTINE ghx_pseudospec_finalize()
    IF (C_ASSOCIATED(ghx_grid_h)) CALL ghx_check(ghx_grid_destroy(ghx_grid_h), 'grid destroy')
    ghx_grid_h = C_NULL_PTR
  END SUBROUTINE ghx_pseudospec_finalize

  ! -- fields ----------------------------------------------------------------
  SUBROUTINE field_alloc(f, spectral)
    TYPE(ghx_field), INTENT(OUT) :: f
    LOGICAL, INTENT(IN) :: spectral
    f%spectral = spectral
    CALL ghx_check(ghx_malloc(f%p, field_bytes(f)), 'field_alloc')
  END SUBROUTINE field_alloc

  SUBROUTINE field_free(f)
    TYPE(ghx_field), INTENT(INOUT) :: f
    CALL ghx_check(ghx_free(f%p), 'field_free')
    f%p = C_NULL_PTR
  END SUBROUTINE field_free

  INTEGER(C_SIZE_T) FUNCTION field_bytes(f)
    TYPE(ghx_field), INTENT(IN) :: f
    IF (f%spectral) THEN
      field_bytes = gnc * 16_C_SIZE_T
    ELSE
      field_bytes = gnr * 8_C_SIZE_T
    END IF
  END FUNCTION field_bytes

  ! host COMPLEX(GP) C(nz,ny,nx/2+1) or REAL(GP) R(nx,ny,nz), passed by address
  SUBROUTINE field_upload(f, host)
    TYPE(ghx_field), INTENT(INOUT) :: f
    TYPE(C_PTR), INTENT(IN) :: host
    CALL ghx_check(ghx_memcpy_h2d(f%p, host, field_bytes(f)), 'field_upload')
  END SUBROUTINE field_upload

  SUBROUTINE field_download(f, host)
    TYPE(ghx_field), INTENT(IN) :: f
    TYPE(C_PTR), INTENT(IN) :: host
    CALL ghx_check(ghx_memcpy_d2h(host, f%p, field_bytes(f)), 'field_download')
  END SUBROUTINE field_download

  ! -- transforms on the module grid ------------------------------------------
  SUBROUTINE r2c(rin, cout)
    TYPE(ghx_field), INTENT(IN) :: rin
    TYPE(ghx_field), INTENT(INOUT) :: cout
    CALL ghx_check(ghx_grid_r2c(ghx_grid_h, rin%p, cout%p), 'r2c')
  END SUBROUTINE r2c

  SUBROUTINE c2r(cin, rout)
    TYPE(ghx_field), INTENT(IN) :: cin
    TYPE(ghx_field), INTENT(INOUT) :: rout
    CALL ghx_check(ghx_grid_c2r(ghx_grid_h, cin%p, rout%p), 'c2r')
  END SUBROUTINE c2r

  ! -- pseudo/pseudospec3D_hd.f90 ----------------------------------------------
  SUBROUTINE derivk3(a, b, dir)
    TYPE(ghx_field), INTENT(IN) :: a
    TYPE(ghx_field), INTENT(INOUT) :: b
    INTEGER, INTENT(IN) :: dir
    CALL ghx_check(ghx_derivk3(ghx_grid_h, a%p, b%p, INT(dir, C_INT)), 'derivk3')
  END SUBROUTINE derivk3

  SUBROUTINE laplak3(a, b)
    TYPE(ghx_field), INTENT(IN) :: a
    TYPE(ghx_field), INTENT(INOUT) :: b
    CALL ghx_check(ghx_laplak3(ghx_grid_h, a%p, b%p), 'laplak3')
  END SUBROUTINE laplak3

  SUBROUTINE rotor3(a, b, c, dir)
    TYPE(ghx_field), INTENT(IN) :: a, b
    TYPE(ghx_field), INTENT(INOUT) :: c
    INTEGER, INTENT(IN) :: dir
    CALL ghx_check(ghx_rotor3(ghx_grid_h, a%p, b%p, c%p, INT(dir, C_INT)), 'rotor3')
  END SUBROUTINE rotor3

  SUBROUTINE prodre3(a, b, c, d, e, f)
    TYPE(ghx_field), INTENT(IN) :: a, b, c
    TYPE(ghx_field), INTENT(INOUT) :: d, e, f
    CALL ghx_check(ghx_prodre3(ghx_grid_h, a%p, b%p, c%p, d%p, e%p, f%p), 'prodre3')
  END SUBROUTINE prodre3

  SUBROUTINE nonlhd3(a, b, c, g, dir)
    TYPE(ghx_field), INTENT(IN) :: a, b, c
    TYPE(ghx_field), INTENT(INOUT) :: g
    INTEGER, INTENT(IN) :: dir
    CALL ghx_check(ghx_nonlhd3(ghx_grid_h, a%p, b%p, c%p, g%p, INT(dir, C_INT)), 'nonlhd3')
  END SUBROUTINE nonlhd3

  SUBROUTINE energy(a, b, c, d, kin)
    TYPE(ghx_field), INTENT(IN) :: a, b, c
    DOUBLE PRECISION, INTENT(OUT) :: d
    INTEGER, INTENT(IN) :: kin
    REAL(C_DOUBLE) :: r
    CALL ghx_check(ghx_energy(ghx_grid_h, a%p, b%p, c%p, INT(kin, C_INT), r), 'energy')
    d = r
  END SUBROUTINE energy

  SUBROUTINE helicity(a, b, c, d)
    TYPE(ghx_field), INTENT(IN) :: a, b, c
    DOUBLE PRECISION, INTENT(OUT) :: d
    REAL(C_DOUBLE) :: r
    CALL ghx_check(ghx_helicity(ghx_grid_h, a%p, b%p, c%p, r), 'helicity')
    d = r
  END SUBROUTINE helicity

  SUBROUTINE cross(a, b, c, d, e, f, g, kin)
    TYPE(ghx_field), INTENT(IN) :: a, b, c, d, e, f
    DOUBLE PRECISION, INTENT(OUT) :: g
    INTEGER, INTENT(IN) :: kin
    REAL(C_DOUBLE) :: r
    CALL ghx_check(ghx_cross(ghx_grid_h, a%p, b%p, c%p, d%p, e%p, f%p, INT(kin, C_INT), r), 'cross')
    g = r
  END SUBROUTINE cross

  SUBROUTINE maxabs(a, b, c, d, kin)
    TYPE(ghx_field), INTENT(IN) :: a, b, c
    REAL(KIND=GP), INTENT(OUT) :: d
    INTEGER, INTENT(IN) :: kin
    REAL(C_DOUBLE) :: r
    CALL ghx_check(ghx_maxabs(ghx_grid_h, a%p, b%p, c%p, INT(kin, C_INT), r), 'maxabs')
    d = r
  END SUBROUTINE maxabs

  ! Ektot, Hktot: shells 1..nmax/2+1 as GHOST's spectrumc (:1190-1415)
  SUBROUTINE spectrumc(a, b, c, kin, hel, Ektot, Hktot)
    TYPE(ghx_field), INTENT(IN) :: a, b, c
    INTEGER, INTENT(IN) :: kin, hel
    DOUBLE PRECISION, INTENT(OUT) :: Ektot(*), Hktot(*)
    CALL ghx_check(ghx_spectrumc(ghx_grid_h, a%p, b%p, c%p, INT(kin, C_INT), INT(hel, C_INT), Ektot, Hktot), &
                   'spectrumc')
  END SUBROUTINE spectrumc

  ! The routines below write GHOST's text files on this rank, in the working
  ! directory, with GHOST's names and formats.

  ! spectrum (:1108-1187): kspectrum/mspectrum and k/m/ghelicity.nmb.txt
  SUBROUTINE spectrum(a, b, c, nmb, kin, hel)
    TYPE(ghx_field), INTENT(IN) :: a, b, c
    INTEGER, INTENT(IN) :: kin, hel
    CHARACTER(len=*), INTENT(IN) :: nmb
    DOUBLE PRECISION :: Ek(gnmax/2 + 1), Hk(gnmax/2 + 1)
    CALL spectrumc(a, b, c, kin, hel, Ek, Hk)
    IF (kin .EQ. 1) CALL shell_file('kspectrum.'//nmb//'.txt', Ek, 0.5D0)
    IF (kin .EQ. 0) CALL shell_file('mspectrum.'//nmb//'.txt', Ek, 0.5D0)
    IF (hel .EQ. 1) THEN
      IF (kin .EQ. 1) THEN
        CALL shell_file('khelicity.'//nmb//'.txt', Hk, 1.0D0)
      ELSE IF (kin .EQ. 0) THEN
        CALL shell_file('mhelicity.'//nmb//'.txt', Hk, 1.0D0)
      ELSE
        CALL shell_file('ghelicity.'//nmb//'.txt', Hk, 1.0D0)
      END IF
    END IF
  END SUBROUTINE spectrum

  ! entrans (:1629-1820): kin = 0 m-, 1 k-, 2 j-transfer, 3 mcrostran, 4 kcrostran
  SUBROUTINE entrans(a, b, c, d, e, f, nmb, kin)
    TYPE(ghx_field), INTENT(IN) :: a, b, c, d, e, f
    CHARACTER(len=*), INTENT(IN) :: nmb
    INTEGER, INTENT(IN) :: kin
    DOUBLE PRECISION :: Ek(gnmax/2 + 1)
    CALL ghx_check(ghx_entrans(ghx_grid_h, a%p, b%p, c%p, d%p, e%p, f%p, INT(kin, C_INT), Ek), 'entrans')
    SELECT CASE (kin)
    CASE (0)
      CALL shell_file('mtransfer.'//nmb//'.txt', Ek, 1.0D0)
    CASE (1)
      CALL shell_file('ktransfer.'//nmb//'.txt', Ek, 1.0D0)
    CASE (2)
      CALL shell_file('jtransfer.'//nmb//'.txt', Ek, 1.0D0)
    CASE (3)
      CALL shell_file('mcrostran.'//nmb//'.txt', Ek, 1.0D0)
    CASE (4)
      CALL shell_file('kcrostran.'//nmb//'.txt', Ek, 1.0D0)
    END SELECT
  END SUBROUTINE entrans

  ! heltrans (:1823-1949): hmtransfer (kin = 0) or hktransfer.nmb.txt
  SUBROUTINE heltrans(a, b, c, d, e, f, nmb, kin)
    TYPE(ghx_field), INTENT(IN) :: a, b, c, d, e, f
    CHARACTER(len=*), INTENT(IN) :: nmb
    INTEGER, INTENT(IN) :: kin
    DOUBLE PRECISION :: Hk(gnmax/2 + 1)
    CALL ghx_check(ghx_heltrans(ghx_grid_h, a%p, b%p, c%p, d%p, e%p, f%p, Hk), 'heltrans')
    IF (kin .EQ. 0) THEN
      CALL shell_file('hmtransfer.'//nmb//'.txt', Hk, 1.0D0)
    ELSE
      CALL shell_file('hktransfer.'//nmb//'.txt', Hk, 1.0D0)
    END IF
  END SUBROUTINE heltrans

  ! spectr1d (:1418-1626): spref//'spec1d'//IJ.nmb.txt, k_dir and E_comp(k_dir)
  SUBROUTINE spectr1d(a, nmb, spref, cmp, dir)
    TYPE(ghx_field), INTENT(IN) :: a
    CHARACTER(len=*), INTENT(IN) :: nmb, spref
    INTEGER, INTENT(IN) :: cmp, dir
    CHARACTER(len=3), PARAMETER :: coord = 'xyz'
    DOUBLE PRECISION :: Ek(gnmax/2 + 1)
    INTEGER :: n(3), i, u
    CALL ghx_check(ghx_spectr1d(ghx_grid_h, a%p, INT(dir, C_INT), Ek), 'spectr1d')
    n = (/gnx, gny, gnz/)
    OPEN(NEWUNIT=u, FILE=TRIM(spref)//'spec1d'//coord(cmp:cmp)//coord(dir:dir)//'.'//nmb//'.txt')
    DO i = 1, n(dir)/2 + 1
      WRITE(u, FMT='(E13.6,E23.15)') REAL(gDk(dir)*(i - 1), GP), REAL(0.5D0*Ek(i)/gDk(dir), GP)
    END DO
    CLOSE(u)
  END SUBROUTINE spectr1d

  ! crosspecc (pseudospec3D_mhd.f90:603-712): reduced shell sums, no file
  SUBROUTINE crosspecc(a, b, c, d, e, f, Cktot, shift)
    TYPE(ghx_field), INTENT(IN) :: a, b, c, d, e, f
    DOUBLE PRECISION, INTENT(OUT) :: Cktot(*)
    REAL(KIND=GP), INTENT(IN) :: shift
    CALL ghx_check(ghx_crosspecc(ghx_grid_h, a%p, b%p, c%p, d%p, e%p, f%p, REAL(shift, C_DOUBLE), Cktot), &
                   'crosspecc')
  END SUBROUTINE crosspecc

  ! hdcheck (:984-1105): appends balance.txt, helicity.txt, divergence.txt
  SUBROUTINE hdcheck(a, b, c, d, e, f, t, dt, hel, chk)
    TYPE(ghx_field), INTENT(IN) :: a, b, c, d, e, f
    INTEGER, INTENT(IN) :: t, hel, chk
    REAL(KIND=GP), INTENT(IN) :: dt
    REAL(C_DOUBLE) :: r(5)
    INTEGER :: u
    CALL ghx_check(ghx_hdcheck(ghx_grid_h, a%p, b%p, c%p, d%p, e%p, f%p, INT(hel, C_INT), INT(chk, C_INT), r), &
                   'hdcheck')
    OPEN(NEWUNIT=u, FILE='balance.txt', POSITION='append')
    WRITE(u, '(E13.6,E26.18,E26.18,E26.18)') (t - 1)*dt, r(1), r(2), r(3)
    CLOSE(u)
    IF (hel .EQ. 1) THEN
      OPEN(NEWUNIT=u, FILE='helicity.txt', POSITION='append')
      WRITE(u, FMT='(E13.6,E26.18)') (t - 1)*dt, r(4)
      CLOSE(u)
    END IF
    IF (chk .EQ. 1) THEN
      OPEN(NEWUNIT=u, FILE='divergence.txt', POSITION='append')
      WRITE(u, FMT='(E13.6,E26.18)') (t - 1)*dt, r(5)
      CLOSE(u)
    END IF
  END SUBROUTINE hdcheck

  ! mhdcheck (pseudospec3D_mhd.f90:339-547): appends balance, energy, helicity,
  ! cross and divergence.txt; crs = 2 (Hall) stops with an error
  SUBROUTINE mhdcheck(a, b, c, ma, mb, mc, t, dt, hel, crs, chk)
    TYPE(ghx_field), INTENT(IN) :: a, b, c, ma, mb, mc
    INTEGER, INTENT(IN) :: t, hel, crs, chk
    REAL(KIND=GP), INTENT(IN) :: dt
    CHARACTER(len=*), PARAMETER :: f10 = '(E13.6,E22.14,E22.14,E22.14)'
    REAL(C_DOUBLE) :: r(11)
    INTEGER :: u
    CALL ghx_check(ghx_mhdcheck(ghx_grid_h, a%p, b%p, c%p, ma%p, mb%p, mc%p, INT(hel, C_INT), INT(crs, C_INT), &
                                INT(chk, C_INT), r), 'mhdcheck')
    OPEN(NEWUNIT=u, FILE='balance.txt', POSITION='append')
    WRITE(u, f10) (t - 1)*dt, r(1), r(2), r(3)
    CLOSE(u)
    OPEN(NEWUNIT=u, FILE='energy.txt', POSITION='append')
    WRITE(u, f10) (t - 1)*dt, r(4), r(5)
    CLOSE(u)
    IF (hel .EQ. 1) THEN
      OPEN(NEWUNIT=u, FILE='helicity.txt', POSITION='append')
      WRITE(u, f10) (t - 1)*dt, r(6), r(7)
      CLOSE(u)
    END IF
    IF (crs .EQ. 1) THEN
      OPEN(NEWUNIT=u, FILE='cross.txt', POSITION='append')
      WRITE(u, f10) (t - 1)*dt, r(8), r(9)
      CLOSE(u)
    END IF
    IF (chk .EQ. 1) THEN
      OPEN(NEWUNIT=u, FILE='divergence.txt', POSITION='append')
      WRITE(u, f10) (t - 1)*dt, r(10), r(11)
      CLOSE(u)
    END IF
  END SUBROUTINE mhdcheck

  ! shells 1..nmax/2+1 as GHOST writes them: Dkk*i, scale*S(i)/Dkk
  SUBROUTINE shell_file(fname, S, scale)
    CHARACTER(len=*), INTENT(IN) :: fname
    DOUBLE PRECISION, INTENT(IN) :: S(*), scale
    INTEGER :: i, u
    OPEN(NEWUNIT=u, FILE=fname)
    DO i = 1, gnmax/2 + 1
      WRITE(u, FMT='(E13.6,E23.15)') REAL(gDkk*i, GP), scale*S(i)/gDkk
    END DO
    CLOSE(u)
  END SUBROUTINE shell_file

  SUBROUTINE phaseshift(a, cdump)
    TYPE(ghx_field), INTENT(INOUT) :: a
    COMPLEX(KIND=GP), INTENT(IN) :: cdump
    CALL ghx_check(ghx_phaseshift(ghx_grid_h, a%p, REAL(REAL(cdump), C_DOUBLE), REAL(AIMAG(cdump), C_DOUBLE)), &
                   'phaseshift')
  END SUBROUTINE phaseshift

  ! module_dns.f90:13-41 (comm is accepted and ignored: one rank)
  SUBROUTINE normalize(fx, fy, fz, f0, kin, comm)
    TYPE(ghx_field), INTENT(INOUT) :: fx, fy, fz
    REAL(KIND=GP), INTENT(IN) :: f0
    INTEGER, INTENT(IN) :: kin, comm
    CALL ghx_check(ghx_normalize(ghx_grid_h, fx%p, fy%p, fz%p, REAL(f0, C_DOUBLE), INT(kin, C_INT)), 'normalize')
  END SUBROUTINE normalize

  ! -- pseudo/pseudospec3D_mhd.f90 ---------------------------------------------
  SUBROUTINE vector3(a, b, c, d, e, f, x, y, z)
    TYPE(ghx_field), INTENT(IN) :: a, b, c, d, e, f
    TYPE(ghx_field), INTENT(INOUT) :: x, y, z
    CALL ghx_check(ghx_vector3(ghx_grid_h, a%p, b%p, c%p, d%p, e%p, f%p, x%p, y%p, z%p), 'vector3')
  END SUBROUTINE vector3

  SUBROUTINE nonlin3(a, b, c, d, e, f, g, dir)
    TYPE(ghx_field), INTENT(IN) :: a, b, c, d, e, f
    TYPE(ghx_field), INTENT(INOUT) :: g
    INTEGER, INTENT(IN) :: dir
    CALL ghx_check(ghx_nonlin3(ghx_grid_h, a%p, b%p, c%p, d%p, e%p, f%p, g%p, INT(dir, C_INT)), 'nonlin3')
  END SUBROUTINE nonlin3

  SUBROUTINE gauge3(a, b, c, g, dir)
    TYPE(ghx_field), INTENT(IN) :: a, b, c
    TYPE(ghx_field), INTENT(INOUT) :: g
    INTEGER, INTENT(IN) :: dir
    CALL ghx_check(ghx_gauge3(ghx_grid_h, a%p, b%p, c%p, g%p, INT(dir, C_INT)), 'gauge3')
  END SUBROUTINE gauge3

  ! -- pseudo/pseudospec3D_filt.f90 --------------------------------------------
  SUBROUTINE sharp(a, b, delta)
    TYPE(ghx_field), INTENT(IN) :: a
    TYPE(ghx_field), INTENT(INOUT) :: b
    REAL(KIND=GP), INTENT(IN) :: delta
    CALL ghx_check(ghx_sharp(ghx_grid_h, a%p, b%p, REAL(delta, C_DOUBLE)), 'sharp')
  END SUBROUTINE sharp

  ! -- include/hd_rkstep2.f90: the body of the RK loop for stage o -------------
  ! v(1:3) <- C + dt (nu lap v + N(v) + f) / o on the dealiased modes, with the
  ! products, projection and update fused; f may be absent (no forcing).
  SUBROUTINE hd_rkstep2(o, dt, nu, v, c, f)
    INTEGER, INTENT(IN) :: o
    REAL(KIND=GP), INTENT(IN) :: dt, nu
    TYPE(ghx_field), INTENT(INOUT) :: v(3)
    TYPE(ghx_field), INTENT(IN) :: c(3)
    TYPE(ghx_field), INTENT(IN), OPTIONAL :: f(3)
    TYPE(C_PTR), TARGET :: pv(3), pc(3), pf(3)
    INTEGER :: k
    DO k = 1, 3
      pv(k) = v(k)%p
      pc(k) = c(k)%p
    END DO
    IF (PRESENT(f)) THEN
      DO k = 1, 3
        pf(k) = f(k)%p
      END DO
      CALL ghx_check(ghx_hd_rkstep2(ghx_grid_h, INT(o, C_INT), REAL(dt, C_DOUBLE), REAL(nu, C_DOUBLE), pv, pc, &
                                    C_LOC(pf), pv), 'hd_rkstep2')
    ELSE
      CALL ghx_check(ghx_hd_rkstep2(ghx_grid_h, INT(o, C_INT), REAL(dt, C_DOUBLE), REAL(nu, C_DOUBLE), pv, pc, &
                                    C_NULL_PTR, pv), 'hd_rkstep2')
    END IF
  END SUBROUTINE hd_rkstep2

  ! -- include/roth_rkstep2.f90: hd_rkstep2 in a frame rotating about z ---------
  SUBROUTINE roth_rkstep2(o, dt, nu, omegaz, v, c, f)
    INTEGER, INTENT(IN) :: o
    REAL(KIND=GP), INTENT(IN) :: dt, nu, omegaz
    TYPE(ghx_field), INTENT(INOUT) :: v(3)
    TYPE(ghx_field), INTENT(IN) :: c(3)
    TYPE(ghx_field), INTENT(IN), OPTIONAL :: f(3)
    TYPE(C_PTR), TARGET :: pv(3), pc(3), pf(3)
    INTEGER :: k
    DO k = 1, 3
      pv(k) = v(k)%p
      pc(k) = c(k)%p
    END DO
    IF (PRESENT(f)) THEN
      DO k = 1, 3
        pf(k) = f(k)%p
      END DO
      CALL ghx_check(ghx_roth_rkstep2(ghx_grid_h, INT(o, C_INT), REAL(dt, C_DOUBLE), REAL(nu, C_DOUBLE), &
                                      REAL(omegaz, C_DOUBLE), pv, pc, C_LOC(pf), pv), 'roth_rkstep2')
    ELSE
      CALL ghx_check(ghx_roth_rkstep2(ghx_grid_h, INT(o, C_INT), REAL(dt, C_DOUBLE), REAL(nu, C_DOUBLE), &
                                      REAL(omegaz, C_DOUBLE), pv, pc, C_NULL_PTR, pv), 'roth_rkstep2')
    END IF
  END SUBROUTINE roth_rkstep2

END MODULE pseudospec_hip
