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
  DOUBLE PRECISION, SAVE :: gDkk = 1.0D0, gDk(3) = 1.0D0, gL(3) = 1.0D0
  DOUBLE PRECISION, PARAMETER :: gpi = 3.14159265358979323846D0

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
    gL = 1.0D0
    IF (arbsize .EQ. 1) THEN
      gDk = (/1.0D0/Lx, 1.0D0/Ly, 1.0D0/Lz/)
      gL = (/DBLE(Lx), DBLE(Ly), DBLE(Lz)/)
    END IF
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

  ! -- pseudo/pseudospec3D_phd.f90: passive-scalar diagnostics --------------
  SUBROUTINE variance(a, b, kin)
    TYPE(ghx_field), INTENT(IN) :: a
    DOUBLE PRECISION, INTENT(OUT) :: b
    INTEGER, INTENT(IN) :: kin
    REAL(C_DOUBLE) :: r
    CALL ghx_check(ghx_variance(ghx_grid_h, a%p, INT(kin, C_INT), r), 'variance')
    b = r
  END SUBROUTINE variance

  SUBROUTINE product(a, b, c)
    TYPE(ghx_field), INTENT(IN) :: a, b
    DOUBLE PRECISION, INTENT(OUT) :: c
    REAL(C_DOUBLE) :: r
    CALL ghx_check(ghx_product(ghx_grid_h, a%p, b%p, r), 'product')
    c = r
  END SUBROUTINE product

  ! pscheck (:286-333): appends scalar.txt
  SUBROUTINE pscheck(a, b, t, dt)
    TYPE(ghx_field), INTENT(IN) :: a, b
    INTEGER, INTENT(IN) :: t
    REAL(KIND=GP), INTENT(IN) :: dt
    REAL(C_DOUBLE) :: r(3)
    INTEGER :: u
    CALL ghx_check(ghx_pscheck(ghx_grid_h, a%p, b%p, r), 'pscheck')
    OPEN(NEWUNIT=u, FILE='scalar.txt', POSITION='append')
    WRITE(u, '(E13.6,E22.14,E22.14,E22.14)') (t - 1)*dt, r(1), r(2), r(3)
    CLOSE(u)
  END SUBROUTINE pscheck

  ! spectrsc (:450-514): sspectrum, sNspectrum (isc > 0) or rhospectrum (isc = -1)
  SUBROUTINE spectrsc(a, nmb, isc)
    TYPE(ghx_field), INTENT(IN) :: a
    CHARACTER(len=*), INTENT(IN) :: nmb
    INTEGER, INTENT(IN) :: isc
    DOUBLE PRECISION :: Ek(gnmax/2 + 1)
    CHARACTER(len=1) :: si
    CALL ghx_check(ghx_spectrscc(ghx_grid_h, a%p, 0.0_C_DOUBLE, Ek), 'spectrsc')
    IF (isc .GT. 0) THEN
      WRITE(si, '(i1.1)') isc
      CALL shell_file('s'//si//'spectrum.'//nmb//'.txt', Ek, 1.0D0)
    ELSE IF (isc .EQ. 0) THEN
      CALL shell_file('sspectrum.'//nmb//'.txt', Ek, 1.0D0)
    ELSE IF (isc .EQ. -1) THEN
      CALL shell_file('rhospectrum.'//nmb//'.txt', Ek, 1.0D0)
    ELSE
      WRITE(*, '(A)') 'SPECTRSC: Invalid option'
      STOP 1
    END IF
  END SUBROUTINE spectrsc

  ! spectrscc (:516-610): shell sums, no file
  SUBROUTINE spectrscc(a, Ektot, shift)
    TYPE(ghx_field), INTENT(IN) :: a
    DOUBLE PRECISION, INTENT(OUT) :: Ektot(*)
    REAL(KIND=GP), INTENT(IN) :: shift
    CALL ghx_check(ghx_spectrscc(ghx_grid_h, a%p, REAL(shift, C_DOUBLE), Ektot), 'spectrscc')
  END SUBROUTINE spectrscc

  ! sctrans (:612-727): stransfer or sNtransfer (isc > 0)
  SUBROUTINE sctrans(a, b, nmb, isc)
    TYPE(ghx_field), INTENT(IN) :: a, b
    CHARACTER(len=*), INTENT(IN) :: nmb
    INTEGER, INTENT(IN) :: isc
    DOUBLE PRECISION :: Ek(gnmax/2 + 1)
    CHARACTER(len=1) :: si
    CALL ghx_check(ghx_sctrans(ghx_grid_h, a%p, b%p, Ek), 'sctrans')
    IF (isc .GT. 0) THEN
      WRITE(si, '(i1.1)') isc
      CALL shell_file('s'//TRIM(si)//'transfer.'//nmb//'.txt', Ek, 1.0D0)
    ELSE
      CALL shell_file('stransfer.'//nmb//'.txt', Ek, 1.0D0)
    END IF
  END SUBROUTINE sctrans

  ! difucx (:729-828) / difucz (:830-940): profilex / profilez.nmb.txt
  SUBROUTINE difucx(a, b, nmb)
    TYPE(ghx_field), INTENT(IN) :: a, b
    CHARACTER(len=*), INTENT(IN) :: nmb
    DOUBLE PRECISION :: p(3*gnx)
    INTEGER :: i, u
    CALL ghx_check(ghx_difucx(ghx_grid_h, a%p, b%p, p), 'difucx')
    OPEN(NEWUNIT=u, FILE='profilex.'//nmb//'.txt')
    DO i = 1, gnx
      WRITE(u, '(E23.15,E23.15,E23.15,E23.15)') REAL(2*gpi*gL(1)*(i - 1)/gnx, GP), p(i), p(gnx + i), p(2*gnx + i)
    END DO
    CLOSE(u)
  END SUBROUTINE difucx

  SUBROUTINE difucz(a, b, nmb)
    TYPE(ghx_field), INTENT(IN) :: a, b
    CHARACTER(len=*), INTENT(IN) :: nmb
    DOUBLE PRECISION :: p(3*gnz)
    INTEGER :: k, u
    CALL ghx_check(ghx_difucz(ghx_grid_h, a%p, b%p, p), 'difucz')
    OPEN(NEWUNIT=u, FILE='profilez.'//nmb//'.txt')
    DO k = 1, gnz
      WRITE(u, '(E23.15,E23.15,E23.15,E23.15)') REAL(2*gpi*gL(3)*(k - 1)/gnz, GP), p(k), p(gnz + k), p(2*gnz + k)
    END DO
    CLOSE(u)
  END SUBROUTINE difucz

  ! -- pseudo/pseudospec3D_rot.f90: spectra about z --------------------------
  ! specpara (:25-335): k/mspecpara (kin 1/0) and k/m/ghelipara (hel 1)
  SUBROUTINE specpara(a, b, c, nmb, kin, hel)
    TYPE(ghx_field), INTENT(IN) :: a, b, c
    CHARACTER(len=*), INTENT(IN) :: nmb
    INTEGER, INTENT(IN) :: kin, hel
    INTEGER(C_INT) :: np, nq
    DOUBLE PRECISION, ALLOCATABLE :: E(:), H(:)
    CALL ghx_check(ghx_axis_bins(ghx_grid_h, np, nq), 'specpara')
    ALLOCATE(E(3*np), H(3*np))
    CALL ghx_check(ghx_specpara(ghx_grid_h, a%p, b%p, c%p, INT(kin, C_INT), INT(hel, C_INT), E, H), 'specpara')
    IF (kin .EQ. 1) CALL axis_file('kspecpara.'//nmb//'.txt', E, np, 0, 0.5D0)
    IF (kin .EQ. 0) CALL axis_file('mspecpara.'//nmb//'.txt', E, np, 0, 0.5D0)
    IF (hel .EQ. 1) CALL axis_file(hel_name(kin)//'helipara.'//nmb//'.txt', H, np, 0, 1.0D0)
    DEALLOCATE(E, H)
  END SUBROUTINE specpara

  ! specperp (:338-686): k/mspecperp and k/m/gheliperp
  SUBROUTINE specperp(a, b, c, nmb, kin, hel)
    TYPE(ghx_field), INTENT(IN) :: a, b, c
    CHARACTER(len=*), INTENT(IN) :: nmb
    INTEGER, INTENT(IN) :: kin, hel
    INTEGER(C_INT) :: np, nq
    DOUBLE PRECISION, ALLOCATABLE :: E(:), H(:)
    CALL ghx_check(ghx_axis_bins(ghx_grid_h, np, nq), 'specperp')
    ALLOCATE(E(3*nq), H(3*nq))
    CALL ghx_check(ghx_specperp(ghx_grid_h, a%p, b%p, c%p, INT(kin, C_INT), INT(hel, C_INT), E, H), 'specperp')
    IF (kin .EQ. 1) CALL axis_file('kspecperp.'//nmb//'.txt', E, nq, 1, 0.5D0)
    IF (kin .EQ. 0) CALL axis_file('mspecperp.'//nmb//'.txt', E, nq, 1, 0.5D0)
    IF (hel .EQ. 1) CALL axis_file(hel_name(kin)//'heliperp.'//nmb//'.txt', H, nq, 1, 1.0D0)
    DEALLOCATE(E, H)
  END SUBROUTINE specperp

  ! entpara (:689-875) / entperp (:878-1064): m/k/jtranpara, m/k/jtranperp
  SUBROUTINE entpara(a, b, c, d, e, f, nmb, kin)
    TYPE(ghx_field), INTENT(IN) :: a, b, c, d, e, f
    CHARACTER(len=*), INTENT(IN) :: nmb
    INTEGER, INTENT(IN) :: kin
    INTEGER(C_INT) :: np, nq
    DOUBLE PRECISION, ALLOCATABLE :: T(:)
    CALL ghx_check(ghx_axis_bins(ghx_grid_h, np, nq), 'entpara')
    ALLOCATE(T(np))
    CALL ghx_check(ghx_entpara(ghx_grid_h, a%p, b%p, c%p, d%p, e%p, f%p, INT(kin, C_INT), T), 'entpara')
    CALL axis_file(tran_name(kin)//'tranpara.'//nmb//'.txt', T, np, 0, 1.0D0, 1)
    DEALLOCATE(T)
  END SUBROUTINE entpara

  SUBROUTINE entperp(a, b, c, d, e, f, nmb, kin)
    TYPE(ghx_field), INTENT(IN) :: a, b, c, d, e, f
    CHARACTER(len=*), INTENT(IN) :: nmb
    INTEGER, INTENT(IN) :: kin
    INTEGER(C_INT) :: np, nq
    DOUBLE PRECISION, ALLOCATABLE :: T(:)
    CALL ghx_check(ghx_axis_bins(ghx_grid_h, np, nq), 'entperp')
    ALLOCATE(T(nq))
    CALL ghx_check(ghx_entperp(ghx_grid_h, a%p, b%p, c%p, d%p, e%p, f%p, INT(kin, C_INT), T), 'entperp')
    CALL axis_file(tran_name(kin)//'tranperp.'//nmb//'.txt', T, nq, 1, 1.0D0, 1)
    DEALLOCATE(T)
  END SUBROUTINE entperp

  ! heltpara (:1067-1196) / heltperp (:1199-1328): hm/hktranpara, hm/hktranperp
  SUBROUTINE heltpara(a, b, c, d, e, f, nmb, kin)
    TYPE(ghx_field), INTENT(IN) :: a, b, c, d, e, f
    CHARACTER(len=*), INTENT(IN) :: nmb
    INTEGER, INTENT(IN) :: kin
    INTEGER(C_INT) :: np, nq
    DOUBLE PRECISION, ALLOCATABLE :: T(:)
    CALL ghx_check(ghx_axis_bins(ghx_grid_h, np, nq), 'heltpara')
    ALLOCATE(T(np))
    CALL ghx_check(ghx_heltpara(ghx_grid_h, a%p, b%p, c%p, d%p, e%p, f%p, T), 'heltpara')
    IF (kin .EQ. 0) THEN
      CALL axis_file('hmtranpara.'//nmb//'.txt', T, np, 0, 1.0D0, 1)
    ELSE
      CALL axis_file('hktranpara.'//nmb//'.txt', T, np, 0, 1.0D0, 1)
    END IF
    DEALLOCATE(T)
  END SUBROUTINE heltpara

  SUBROUTINE heltperp(a, b, c, d, e, f, nmb, kin)
    TYPE(ghx_field), INTENT(IN) :: a, b, c, d, e, f
    CHARACTER(len=*), INTENT(IN) :: nmb
    INTEGER, INTENT(IN) :: kin
    INTEGER(C_INT) :: np, nq
    DOUBLE PRECISION, ALLOCATABLE :: T(:)
    CALL ghx_check(ghx_axis_bins(ghx_grid_h, np, nq), 'heltperp')
    ALLOCATE(T(nq))
    CALL ghx_check(ghx_heltperp(ghx_grid_h, a%p, b%p, c%p, d%p, e%p, f%p, T), 'heltperp')
    IF (kin .EQ. 0) THEN
      CALL axis_file('hmtranperp.'//nmb//'.txt', T, nq, 1, 1.0D0, 1)
    ELSE
      CALL axis_file('hktranperp.'//nmb//'.txt', T, nq, 1, 1.0D0, 1)
    END IF
    DEALLOCATE(T)
  END SUBROUTINE heltperp

  ! spec2D (:1331-1629): unformatted dir/k,mspec2D.nmb.out and k,m,gheli2D,
  ! one record of REAL(GP) E(kperp, kpara)
  SUBROUTINE spec2D(a, b, c, nmb, dir, kin, hel)
    TYPE(ghx_field), INTENT(IN) :: a, b, c
    CHARACTER(len=*), INTENT(IN) :: nmb
    CHARACTER(len=100), INTENT(IN) :: dir
    INTEGER, INTENT(IN) :: kin, hel
    INTEGER(C_INT) :: np, nq
    DOUBLE PRECISION, ALLOCATABLE :: E(:), H(:)
    INTEGER :: u
    CALL ghx_check(ghx_axis_bins(ghx_grid_h, np, nq), 'spec2D')
    ALLOCATE(E(np*nq), H(np*nq))
    CALL ghx_check(ghx_spec2d(ghx_grid_h, a%p, b%p, c%p, INT(kin, C_INT), INT(hel, C_INT), E, H), 'spec2D')
    IF (kin .LE. 1) THEN
      IF (kin .EQ. 1) THEN
        OPEN(NEWUNIT=u, FILE=TRIM(dir)//'/'//'kspec2D.'//nmb//'.out', FORM='unformatted')
      ELSE
        OPEN(NEWUNIT=u, FILE=TRIM(dir)//'/'//'mspec2D.'//nmb//'.out', FORM='unformatted')
      END IF
      WRITE(u) REAL(E, GP)
      CLOSE(u)
    END IF
    IF (hel .EQ. 1) THEN
      OPEN(NEWUNIT=u, FILE=TRIM(dir)//'/'//hel_name(kin)//'heli2D.'//nmb//'.out', FORM='unformatted')
      WRITE(u) REAL(H, GP)
      CLOSE(u)
    END IF
    DEALLOCATE(E, H)
  END SUBROUTINE spec2D

  ! -- pseudo/pseudospec3D_bouss.f90: stratified-flow diagnostics -------------
  SUBROUTINE havgcomp(gsh, u, v, w, s, fo, bv, itype)
    DOUBLE PRECISION, INTENT(INOUT) :: gsh(*)
    TYPE(ghx_field), INTENT(IN) :: u, v, w, s
    REAL(KIND=GP), INTENT(IN) :: fo, bv
    INTEGER, INTENT(IN) :: itype
    CALL ghx_check(ghx_havgcomp(ghx_grid_h, u%p, v%p, w%p, s%p, REAL(fo, C_DOUBLE), REAL(bv, C_DOUBLE), &
                                INT(itype, C_INT), gsh), 'havgcomp')
  END SUBROUTINE havgcomp

  ! havgwrite (:557-610): spref.nmb.txt, z and the profile
  SUBROUTINE havgwrite(itype, spref, nmb, u, v, w, s, fo, bv)
    INTEGER, INTENT(IN) :: itype
    CHARACTER(len=*), INTENT(IN) :: spref, nmb
    TYPE(ghx_field), INTENT(IN) :: u, v, w, s
    REAL(KIND=GP), INTENT(IN) :: fo, bv
    DOUBLE PRECISION :: gsh(gnz)
    INTEGER :: k, un
    CALL havgcomp(gsh, u, v, w, s, fo, bv, itype)
    OPEN(NEWUNIT=un, FILE=TRIM(spref)//'.'//TRIM(nmb)//'.txt')
    DO k = 1, gnz
      WRITE(un, '(E23.15,E23.15)') REAL(2*gpi*gL(3)*(k - 1)/gnz, GP), gsh(k)
    END DO
    CLOSE(un)
  END SUBROUTINE havgwrite

  ! tbouss (:612-709): appends tboussavg, tboussmax, tboussmin.txt
  SUBROUTINE tbouss(u, v, w, s, t, dt, fo, bv)
    TYPE(ghx_field), INTENT(IN) :: u, v, w, s
    INTEGER, INTENT(IN) :: t
    REAL(KIND=GP), INTENT(IN) :: dt, fo, bv
    REAL(C_DOUBLE) :: avg(11), mx(11), mn(11)
    INTEGER :: un
    CALL ghx_check(ghx_tbouss(ghx_grid_h, u%p, v%p, w%p, s%p, REAL(fo, C_DOUBLE), REAL(bv, C_DOUBLE), avg, mx, mn), &
                   'tbouss')
    OPEN(NEWUNIT=un, FILE='tboussavg.txt', POSITION='append')
    WRITE(un, '(E13.6,1x,11(E26.18,1x))') (t - 1)*dt, avg
    CLOSE(un)
    OPEN(NEWUNIT=un, FILE='tboussmax.txt', POSITION='append')
    WRITE(un, '(E13.6,1x,11(E26.18,1x))') (t - 1)*dt, mx
    CLOSE(un)
    OPEN(NEWUNIT=un, FILE='tboussmin.txt', POSITION='append')
    WRITE(un, '(E13.6,1x,11(E26.18,1x))') (t - 1)*dt, mn
    CLOSE(un)
  END SUBROUTINE tbouss

  ! spectpv (:712-812): pvspectrum.nmb.txt, kz and the PV spectrum
  SUBROUTINE spectpv(u, v, w, s, nmb)
    TYPE(ghx_field), INTENT(IN) :: u, v, w, s
    CHARACTER(len=*), INTENT(IN) :: nmb
    DOUBLE PRECISION :: Ek(gnz/2 + 1)
    INTEGER :: k, un
    CALL ghx_check(ghx_spectpv(ghx_grid_h, u%p, v%p, w%p, s%p, Ek), 'spectpv')
    OPEN(NEWUNIT=un, FILE='pvspectrum.'//nmb//'.txt')
    DO k = 1, gnz/2 + 1
      WRITE(un, FMT='(E13.6,E23.15)') REAL(gDk(3)*(k - 1), GP), Ek(k)
    END DO
    CLOSE(un)
  END SUBROUTINE spectpv

  ! file-name prefixes of the helicity (k, m, g by kin) and transfer (m, k, j) files
  CHARACTER(len=1) FUNCTION hel_name(kin)
    INTEGER, INTENT(IN) :: kin
    hel_name = 'g'
    IF (kin .EQ. 1) hel_name = 'k'
    IF (kin .EQ. 0) hel_name = 'm'
  END FUNCTION hel_name
  CHARACTER(len=1) FUNCTION tran_name(kin)
    INTEGER, INTENT(IN) :: kin
    tran_name = 'j'
    IF (kin .EQ. 1) tran_name = 'k'
    IF (kin .EQ. 0) tran_name = 'm'
  END FUNCTION tran_name

  ! plane (bins 0: kz = Dkz (q-1), values times Lz) or cylinder (bins 1:
  ! kperp = Dkk (q-1), values over Dkk) files; S holds 3 n values {total, xy,
  ! z} written as four columns, or ncol = 1: n values as two columns
  SUBROUTINE axis_file(fname, S, n, bins, scale, ncol)
    CHARACTER(len=*), INTENT(IN) :: fname
    DOUBLE PRECISION, INTENT(IN) :: S(*), scale
    INTEGER(C_INT), INTENT(IN) :: n
    INTEGER, INTENT(IN) :: bins
    INTEGER, INTENT(IN), OPTIONAL :: ncol
    DOUBLE PRECISION :: k, f
    INTEGER :: q, u
    OPEN(NEWUNIT=u, FILE=fname)
    DO q = 1, n
      IF (bins .EQ. 0) THEN
        k = gDk(3)*(q - 1)
        f = scale*gL(3)
      ELSE
        k = gDkk*(q - 1)
        f = scale/gDkk
      END IF
      IF (PRESENT(ncol)) THEN
        WRITE(u, FMT='(E13.6,E23.15)') REAL(k, GP), f*S(q)
      ELSE
        WRITE(u, FMT='(E13.6,E23.15,E23.15,E23.15)') REAL(k, GP), f*S(q), f*S(n + q), f*S(2*n + q)
      END IF
    END DO
    CLOSE(u)
  END SUBROUTINE axis_file

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

  ! -- include/mhd_rkstep2.f90: the fused MHD substep for stage o -------------
  ! v(1:3), a(1:3) <- C + dt (diffusion + nonlinear terms + forcing) / o;
  ! c(1:6) holds (v, a) at the step start; f and m (mechanical and
  ! electromotive forcing) are both present or both absent.
  SUBROUTINE mhd_rkstep2(o, dt, nu, mu, v, a, c, f, m)
    INTEGER, INTENT(IN) :: o
    REAL(KIND=GP), INTENT(IN) :: dt, nu, mu
    TYPE(ghx_field), INTENT(INOUT) :: v(3), a(3)
    TYPE(ghx_field), INTENT(IN) :: c(6)
    TYPE(ghx_field), INTENT(IN), OPTIONAL :: f(3), m(3)
    TYPE(C_PTR), TARGET :: pv(6), pc(6), pf(3), pm(3)
    INTEGER :: k
    DO k = 1, 3
      pv(k) = v(k)%p
      pv(3 + k) = a(k)%p
    END DO
    DO k = 1, 6
      pc(k) = c(k)%p
    END DO
    IF (PRESENT(f) .AND. PRESENT(m)) THEN
      DO k = 1, 3
        pf(k) = f(k)%p
        pm(k) = m(k)%p
      END DO
      CALL ghx_check(ghx_mhd_rkstep2(ghx_grid_h, INT(o, C_INT), REAL(dt, C_DOUBLE), REAL(nu, C_DOUBLE), &
                                     REAL(mu, C_DOUBLE), pv, pc, C_LOC(pf), C_LOC(pm), pv), 'mhd_rkstep2')
    ELSE
      CALL ghx_check(ghx_mhd_rkstep2(ghx_grid_h, INT(o, C_INT), REAL(dt, C_DOUBLE), REAL(nu, C_DOUBLE), &
                                     REAL(mu, C_DOUBLE), pv, pc, C_NULL_PTR, C_NULL_PTR, pv), 'mhd_rkstep2')
    END IF
  END SUBROUTINE mhd_rkstep2

  ! -- include/bouss_rkstep2.f90 and rotbouss_rkstep2.f90 ----------------------
  ! v(1:3), th <- C + dt (...) / o; c(1:4) holds (v, th) at the step start;
  ! f(1:3), fs the forcings (both present or both absent).  ord is GHOST's
  ! order (main3D's variable): with ord = 3 and o = 1 the correction of
  ! :73-145 follows (c is left unchanged, where GHOST overwrites C1-C3).
  SUBROUTINE bouss_rkstep2(o, dt, nu, kappa, xmom, xtemp, v, th, c, f, fs, ord)
    INTEGER, INTENT(IN) :: o
    REAL(KIND=GP), INTENT(IN) :: dt, nu, kappa, xmom, xtemp
    TYPE(ghx_field), INTENT(INOUT) :: v(3), th
    TYPE(ghx_field), INTENT(IN) :: c(4)
    TYPE(ghx_field), INTENT(IN), OPTIONAL :: f(3), fs
    INTEGER, INTENT(IN), OPTIONAL :: ord
    TYPE(C_PTR), TARGET :: pv(3), pc(4)
    INTEGER :: k
    CALL bouss_body(o, dt, nu, kappa, xmom, xtemp, .FALSE., 0.0_GP, v, th, c, f, fs)
    IF (PRESENT(ord)) THEN
      IF (ord .EQ. 3 .AND. o .LE. 1) THEN
        DO k = 1, 3
          pv(k) = v(k)%p
        END DO
        DO k = 1, 4
          pc(k) = c(k)%p
        END DO
        CALL ghx_check(ghx_bouss_ord3(ghx_grid_h, REAL(dt, C_DOUBLE), REAL(xmom, C_DOUBLE), pc, pv), &
                       'bouss_ord3')
      END IF
    END IF
  END SUBROUTINE bouss_rkstep2

  SUBROUTINE rotbouss_rkstep2(o, dt, nu, kappa, xmom, xtemp, omegaz, v, th, c, f, fs)
    INTEGER, INTENT(IN) :: o
    REAL(KIND=GP), INTENT(IN) :: dt, nu, kappa, xmom, xtemp, omegaz
    TYPE(ghx_field), INTENT(INOUT) :: v(3), th
    TYPE(ghx_field), INTENT(IN) :: c(4)
    TYPE(ghx_field), INTENT(IN), OPTIONAL :: f(3), fs
    CALL bouss_body(o, dt, nu, kappa, xmom, xtemp, .TRUE., omegaz, v, th, c, f, fs)
  END SUBROUTINE rotbouss_rkstep2

  SUBROUTINE bouss_body(o, dt, nu, kappa, xmom, xtemp, rot, omegaz, v, th, c, f, fs)
    INTEGER, INTENT(IN) :: o
    REAL(KIND=GP), INTENT(IN) :: dt, nu, kappa, xmom, xtemp, omegaz
    LOGICAL, INTENT(IN) :: rot
    TYPE(ghx_field), INTENT(INOUT) :: v(3), th
    TYPE(ghx_field), INTENT(IN) :: c(4)
    TYPE(ghx_field), INTENT(IN), OPTIONAL :: f(3), fs
    TYPE(C_PTR), TARGET :: pv(4), pc(4), pf(4)
    TYPE(C_PTR) :: fp
    INTEGER :: k
    DO k = 1, 3
      pv(k) = v(k)%p
    END DO
    pv(4) = th%p
    DO k = 1, 4
      pc(k) = c(k)%p
    END DO
    fp = C_NULL_PTR
    IF (PRESENT(f) .AND. PRESENT(fs)) THEN
      DO k = 1, 3
        pf(k) = f(k)%p
      END DO
      pf(4) = fs%p
      fp = C_LOC(pf)
    END IF
    IF (rot) THEN
      CALL ghx_check(ghx_rotbouss_rkstep2(ghx_grid_h, INT(o, C_INT), REAL(dt, C_DOUBLE), REAL(nu, C_DOUBLE), &
                                          REAL(kappa, C_DOUBLE), REAL(xmom, C_DOUBLE), REAL(xtemp, C_DOUBLE), &
                                          REAL(omegaz, C_DOUBLE), pv, pc, fp, pv), 'rotbouss_rkstep2')
    ELSE
      CALL ghx_check(ghx_bouss_rkstep2(ghx_grid_h, INT(o, C_INT), REAL(dt, C_DOUBLE), REAL(nu, C_DOUBLE), &
                                       REAL(kappa, C_DOUBLE), REAL(xmom, C_DOUBLE), REAL(xtemp, C_DOUBLE), &
                                       pv, pc, fp, pv), 'bouss_rkstep2')
    END IF
  END SUBROUTINE bouss_body

END MODULE pseudospec_hip
