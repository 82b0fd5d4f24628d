! test_fftp_hip.f90 -- the Fortran host side, end to end on the GPU.
!
! Usage: test_fftp_hip OUTDIR
! 1. fftp3d_* on host arrays (GHOST's calls, unchanged): round trip and
!    Taylor-Green closed forms through the host-array pseudospec drop-ins.
! 2. Device-resident fields through pseudospec_hip: TG enstrophy, and an RK2
!    loop written as GHOST's main3D (C = v; DO o = ord,1,-1: hd_rkstep2).
! 2c. GHOST's MHD loop through mhd_rkstep2 on resident fields.
! 3. The resident solver with tracers attached.
! 4. fftp3d_* on host arrays with 512-point x and z lines: OUTDIR/
!    fortran_fftp_<nx>x<ny>x<nz>.bin holds the input and the r2c output.
! 2b. GHOST's diagnostic files (spectrum, entrans, heltrans, spectr1d, hdcheck,
!    the spectra about z, the scalar and the Boussinesq diagnostics) from the
!    resident fields, written to the working directory (spec2D to OUTDIR).
! Writes OUTDIR/fortran_hd32.bin (v after the RK loop, 3 x C(nz,ny,nx/2+1))
! and OUTDIR/fortran_tracers32.bin (initial and final positions) for the
! comparison with the CPU oracle; prints one line per check and stops with a
! non-zero status on the first failure.
PROGRAM test_fftp_hip
  USE iso_c_binding
  USE fprecision
  USE fftplans
  USE ghx_c
  USE pseudospec_hip, ONLY: ghx_pseudospec_init, ghx_pseudospec_finalize, ghx_field, field_alloc, field_free, &
                            field_upload, field_download, ghx_grid_h, d_rotor3 => rotor3, d_energy => energy, &
                            hd_rkstep2, spectrum, entrans, heltrans, spectr1d, hdcheck, specpara, specperp, &
                            entpara, entperp, heltpara, spec2D, pscheck, spectrsc, sctrans, difucz, tbouss, &
                            havgwrite, spectpv, mhd_rkstep2
  USE iovar, ONLY: IOPLAN
  IMPLICIT NONE
  INTEGER, PARAMETER :: n = 32, nh = n/2 + 1, nsteps = 20, npart = 1000, psteps = 5
  REAL(KIND=GP), PARAMETER :: pi = 3.14159265358979323846_GP
  REAL(KIND=GP), PARAMETER :: dt = 1.0E-3_GP, nu = 1.0E-2_GP
  TYPE(FFTPLAN) :: planrc, plancr
  REAL(KIND=GP), ALLOCATABLE, TARGET :: ux(:,:,:), uy(:,:,:), uz(:,:,:), r(:,:,:)
  COMPLEX(KIND=GP), ALLOCATABLE, TARGET :: c1(:,:,:), c2(:,:,:), c3(:,:,:), c4(:,:,:)
  REAL(KIND=GP), ALLOCATABLE, TARGET :: pos0(:,:), pos1(:,:)
  TYPE(ghx_field) :: v(3), cst(3), w(3), c6(6)
  TYPE(C_PTR) :: hd, hv(3), hf(3), tr, pp(3)
  CHARACTER(LEN=512) :: outdir
  CHARACTER(LEN=100) :: iodir
  TYPE(IOPLAN) :: planio
  DOUBLE PRECISION :: e, err, x, y, z, rn
  INTEGER :: i, j, k, o, q, u
  INTEGER :: ierr
  INTEGER(C_INT32_T) :: seed

  CALL GET_COMMAND_ARGUMENT(1, outdir)
  IF (LEN_TRIM(outdir) .EQ. 0) outdir = '.'
  IF (ghx_device_ok() .EQ. 0) THEN
    WRITE(*, '(A)') 'no GPU visible'
    STOP 2
  END IF

  CALL fftp3d_init_threads(ierr)
  CALL fftp3d_create_plan(planrc, (/n, n, n/), FFTW_REAL_TO_COMPLEX, FFTW_ESTIMATE)
  CALL fftp3d_create_plan(plancr, (/n, n, n/), FFTW_COMPLEX_TO_REAL, FFTW_ESTIMATE)
  CALL ghx_pseudospec_init(n, n, n, 1.0_GP, 1.0_GP, 1.0_GP, 0, 0.0_GP)

  ALLOCATE(ux(n, n, n), uy(n, n, n), uz(n, n, n), r(n, n, n))
  ALLOCATE(c1(n, n, nh), c2(n, n, nh), c3(n, n, nh), c4(n, n, nh))
  DO k = 1, n
    z = 2*pi*(k - 1)/n
    DO j = 1, n
      y = 2*pi*(j - 1)/n
      DO i = 1, n
        x = 2*pi*(i - 1)/n
        ux(i, j, k) = SIN(x)*COS(y)*COS(z)
        uy(i, j, k) = -COS(x)*SIN(y)*COS(z)
        uz(i, j, k) = 0.0_GP
      END DO
    END DO
  END DO

  ! -- 1. GHOST's host-array calls -------------------------------------------
  CALL fftp3d_real_to_complex(planrc, ux, c1, 0)
  CALL fftp3d_real_to_complex(planrc, uy, c2, 0)
  CALL fftp3d_real_to_complex(planrc, uz, c3, 0)
  CALL fftp3d_complex_to_real(plancr, c1, r, 0)
  err = MAXVAL(ABS(r/REAL(n, GP)**3 - ux))
  CALL report('fftp round trip', err, 1.0E-13_GP)

  CALL derivk3(c1, c4, 1)                    ! d/dx (sin x cos y cos z)
  CALL fftp3d_complex_to_real(plancr, c4, r, 0)
  err = 0.0
  DO k = 1, n
    DO j = 1, n
      DO i = 1, n
        err = MAX(err, ABS(r(i, j, k)/REAL(n, GP)**3 - COS(2*pi*(i - 1)/n)*COS(2*pi*(j - 1)/n)*COS(2*pi*(k - 1)/n)))
      END DO
    END DO
  END DO
  CALL report('derivk3 closed form', err, 1.0E-12_GP)

  CALL energy(c1, c2, c3, e, 1)              ! <u^2> of unit TG = 1/4
  CALL report('energy of TG', ABS(e - 0.25D0), 1.0E-13_GP)

  ! -- 2. device-resident fields ---------------------------------------------
  DO q = 1, 3
    CALL field_alloc(v(q), .TRUE.)
    CALL field_alloc(cst(q), .TRUE.)
    CALL field_alloc(w(q), .TRUE.)
  END DO
  CALL field_upload(v(1), C_LOC(c1))
  CALL field_upload(v(2), C_LOC(c2))
  CALL field_upload(v(3), C_LOC(c3))
  CALL d_rotor3(v(2), v(3), w(1), 1)
  CALL d_rotor3(v(1), v(3), w(2), 2)
  CALL d_rotor3(v(1), v(2), w(3), 3)
  CALL d_energy(w(1), w(2), w(3), e, 1)      ! <omega^2> of unit TG = 3/4
  CALL report('enstrophy of TG (device fields)', ABS(e - 0.75D0), 1.0E-12_GP)

  ! GHOST's TG initial condition (examples/initialv.f90_tg, u0 = 0.5) and the
  ! RK loop of main3D.fpp:3202-3231 on resident fields
  CALL ghx_check(ghx_initialv_tg(ghx_grid_h, 1.0_C_DOUBLE, 1.0_C_DOUBLE, 0.5_C_DOUBLE, v(1)%p, v(2)%p, v(3)%p), &
                 'initialv_tg')
  DO q = 1, nsteps
    DO k = 1, 3
      CALL ghx_check(ghx_memcpy_d2d(cst(k)%p, v(k)%p, INT(n, C_SIZE_T)**2*nh*16), 'copy C = v')
    END DO
    DO o = 2, 1, -1
      CALL hd_rkstep2(o, dt, nu, v, cst)
    END DO
  END DO
  OPEN(NEWUNIT=u, FILE=TRIM(outdir)//'/fortran_hd32.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
  DO k = 1, 3
    CALL field_download(v(k), C_LOC(c4))
    WRITE(u) c4
  END DO
  CLOSE(u)
  WRITE(*, '(A)') 'ok   RK loop written (fortran_hd32.bin)'

  ! GHOST's diagnostic files from resident fields, in the working directory:
  ! a k^4 exp(-k^2/k0^2) field (k0 = 3, seed 11) against a random forcing
  ! field (shells 1..6, seed 29)
  seed = 11
  CALL ghx_check(ghx_initialv_k4exp(ghx_grid_h, 3.0_C_DOUBLE, 1.0_C_DOUBLE, seed, w(1)%p, w(2)%p, w(3)%p), &
                 'initialv_k4exp')
  seed = 29
  CALL ghx_check(ghx_initialfv_random(ghx_grid_h, 1.0_C_DOUBLE, 6.0_C_DOUBLE, 1.0_C_DOUBLE, seed, cst(1)%p, &
                                      cst(2)%p, cst(3)%p), 'initialfv_random')
  CALL spectrum(w(1), w(2), w(3), '0001', 1, 1)
  CALL entrans(w(1), w(2), w(3), cst(1), cst(2), cst(3), '0001', 1)
  CALL entrans(w(1), w(2), w(3), cst(1), cst(2), cst(3), '0001', 0)
  CALL heltrans(w(1), w(2), w(3), cst(1), cst(2), cst(3), '0001', 1)
  CALL spectr1d(w(1), '0001', 'k', 1, 3)
  CALL hdcheck(w(1), w(2), w(3), cst(1), cst(2), cst(3), nsteps + 1, dt, 1, 1)
  CALL specpara(w(1), w(2), w(3), '0001', 1, 1)
  CALL specperp(w(1), w(2), w(3), '0001', 1, 1)
  CALL entpara(w(1), w(2), w(3), cst(1), cst(2), cst(3), '0001', 1)
  CALL entperp(w(1), w(2), w(3), cst(1), cst(2), cst(3), '0001', 0)
  CALL heltpara(w(1), w(2), w(3), cst(1), cst(2), cst(3), '0001', 1)
  iodir = outdir
  CALL spec2D(w(1), w(2), w(3), '0001', iodir, 1, 0)
  CALL pscheck(w(1), cst(1), nsteps + 1, dt)
  CALL spectrsc(w(1), '0001', 0)
  CALL sctrans(w(1), cst(1), '0001', 2)
  CALL difucz(w(1), w(2), '0001')
  CALL tbouss(w(1), w(2), w(3), cst(1), nsteps + 1, dt, 1.3_GP, 0.8_GP)
  CALL havgwrite(7, 'pvprof', '0001', w(1), w(2), w(3), cst(1), 1.3_GP, 0.8_GP)
  CALL spectpv(w(1), w(2), w(3), cst(1), '0001')
  WRITE(*, '(A)') 'ok   diagnostic files written'

  ! GHOST's MHD loop (include/mhd_rkstep1/2.f90) on resident fields: v = the
  ! k^4 exp field, a = the random field above, 3 unforced RK2 steps
  DO q = 1, 6
    CALL field_alloc(c6(q), .TRUE.)
  END DO
  DO q = 1, 3
    DO o = 2, 1, -1
      IF (o .EQ. 2) THEN
        DO k = 1, 3
          CALL ghx_check(ghx_memcpy_d2d(c6(k)%p, w(k)%p, INT(n, C_SIZE_T)**2*nh*16), 'copy C = v')
          CALL ghx_check(ghx_memcpy_d2d(c6(3 + k)%p, cst(k)%p, INT(n, C_SIZE_T)**2*nh*16), 'copy C = a')
        END DO
      END IF
      CALL mhd_rkstep2(o, dt, nu, nu, w, cst, c6)
    END DO
  END DO
  OPEN(NEWUNIT=u, FILE=TRIM(outdir)//'/fortran_mhd32.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
  DO k = 1, 3
    CALL field_download(w(k), C_LOC(c4))
    WRITE(u) c4
  END DO
  DO k = 1, 3
    CALL field_download(cst(k), C_LOC(c4))
    WRITE(u) c4
  END DO
  CLOSE(u)
  DO q = 1, 6
    CALL field_free(c6(q))
  END DO
  WRITE(*, '(A)') 'ok   MHD loop written (fortran_mhd32.bin)'

  ! GHOST's binary output of v (main3D.fpp:2675-2710) from the resident
  ! fields, and the same file written from host arrays through GHOST's own
  ! calls (C = v*rmp, fftp3d_complex_to_real, io_write), then read back
  iodir = outdir
  CALL io_init(0, (/n, n, n/), 1, n, planio)
  CALL ghx_check(ghx_hd_output(ghx_grid_h, planio%h, TRIM(iodir)//C_NULL_CHAR, '0001'//C_NULL_CHAR, 0_C_INT, &
                               v(1)%p, v(2)%p, v(3)%p), 'hd_output')
  CALL field_download(v(1), C_LOC(c4))
  c4 = c4*(1.0_GP/(REAL(n, GP)*REAL(n, GP)*REAL(n, GP)))
  CALL fftp3d_complex_to_real(plancr, c4, r, 0)
  CALL io_write(1, iodir, 'hx', '0001', planio, r)
  CALL io_read(1, iodir, 'vx', '0001', planio, ux)
  CALL report('io_write/io_read: device output == host output', MAXVAL(ABS(ux - r)), 0.0_GP)
  WRITE(*, '(A)') 'ok   binary output written (vx.0001.out, hx.0001.out)'

  ! -- 3. resident solver with tracers ---------------------------------------
  ALLOCATE(pos0(npart, 3), pos1(npart, 3))
  DO q = 1, npart                           ! a fixed low-discrepancy set in [0, n)
    rn = REAL(q, GP)
    pos0(q, 1) = n*(rn*0.7548776662466927_GP - AINT(rn*0.7548776662466927_GP))
    pos0(q, 2) = n*(rn*0.5698402909980532_GP - AINT(rn*0.5698402909980532_GP))
    pos0(q, 3) = n*(rn*0.4301597090019468_GP - AINT(rn*0.4301597090019468_GP))
  END DO
  CALL ghx_check(ghx_hd_create(hd, ghx_grid_h, 2_C_INT, dt, nu), 'hd_create')
  CALL ghx_check(ghx_hd_fields(hd, hv, hf), 'hd_fields')
  CALL ghx_check(ghx_initialv_tg(ghx_grid_h, 1.0_C_DOUBLE, 1.0_C_DOUBLE, 0.5_C_DOUBLE, hv(1), hv(2), hv(3)), &
                 'initialv_tg')
  CALL ghx_check(ghx_hd_touch(hd), 'hd_touch')
  CALL ghx_check(ghx_hd_set_forcing(hd, 0_C_INT), 'hd_set_forcing')
  CALL ghx_check(ghx_tracers_create(tr, ghx_grid_h, INT(npart, C_LONG_LONG), 1_C_INT), 'tracers_create')
  CALL ghx_check(ghx_tracers_positions(tr, pp(1), pp(2), pp(3)), 'tracers_positions')
  DO q = 1, 3
    CALL ghx_check(ghx_memcpy_h2d(pp(q), C_LOC(pos0(1, q)), INT(npart, C_SIZE_T)*8), 'positions in')
  END DO
  CALL ghx_check(ghx_hd_set_tracers(hd, tr), 'hd_set_tracers')
  CALL ghx_check(ghx_hd_step(hd, INT(psteps, C_INT)), 'hd_step')
  DO q = 1, 3
    CALL ghx_check(ghx_memcpy_d2h(C_LOC(pos1(1, q)), pp(q), INT(npart, C_SIZE_T)*8), 'positions out')
  END DO
  err = MAXVAL(ABS(pos1 - pos0))
  IF (err .LE. 0.0 .OR. MINVAL(pos1) .LT. 0.0 .OR. MAXVAL(pos1) .GE. n) THEN
    WRITE(*, '(A)') 'FAIL tracers did not move or left the box'
    STOP 1
  END IF
  OPEN(NEWUNIT=u, FILE=TRIM(outdir)//'/fortran_tracers32.bin', ACCESS='STREAM', FORM='UNFORMATTED', &
       STATUS='REPLACE')
  WRITE(u) pos0
  WRITE(u) pos1
  CLOSE(u)
  WRITE(*, '(A)') 'ok   solver with tracers (fortran_tracers32.bin)'

  CALL ghx_check(ghx_hd_set_tracers(hd, C_NULL_PTR), 'hd_set_tracers')
  CALL ghx_check(ghx_tracers_destroy(tr), 'tracers_destroy')
  CALL ghx_check(ghx_hd_destroy(hd), 'hd_destroy')
  DO q = 1, 3
    CALL field_free(v(q))
    CALL field_free(cst(q))
    CALL field_free(w(q))
  END DO
  CALL fftp3d_destroy_plan(planrc)
  CALL fftp3d_destroy_plan(plancr)
  CALL ghx_pseudospec_finalize()
  WRITE(*, '(A,ES10.3,A)') 'ok   ffttime accumulated: ', ffttime, ' s'

  ! -- 4. 512-point lines through GHOST's fftp calls ---------------------------
  CALL fftp_shape(512, 8, 16)
  CALL fftp_shape(16, 8, 512)
  WRITE(*, '(A)') 'ALL OK'

CONTAINS

  SUBROUTINE fftp_shape(mx, my, mz)
    ! a fixed aperiodic real field, r2c, c2r round trip, both arrays to a file
    INTEGER, INTENT(IN) :: mx, my, mz
    TYPE(FFTPLAN) :: prc, pcr
    REAL(KIND=GP), ALLOCATABLE, TARGET :: a(:,:,:), b(:,:,:)
    COMPLEX(KIND=GP), ALLOCATABLE, TARGET :: s(:,:,:)
    CHARACTER(LEN=64) :: fname
    INTEGER :: ii, jj, kk, uu
    CALL fftp3d_create_plan(prc, (/mx, my, mz/), FFTW_REAL_TO_COMPLEX, FFTW_ESTIMATE)
    CALL fftp3d_create_plan(pcr, (/mx, my, mz/), FFTW_COMPLEX_TO_REAL, FFTW_ESTIMATE)
    ALLOCATE(a(mx, my, mz), b(mx, my, mz), s(mz, my, mx/2 + 1))
    DO kk = 1, mz
      DO jj = 1, my
        DO ii = 1, mx
          a(ii, jj, kk) = SIN(0.37_GP*ii + 1.3_GP*jj*jj) + COS(0.011_GP*kk*ii) - 0.5_GP*SIN(0.7_GP*kk*kk + jj)
        END DO
      END DO
    END DO
    CALL fftp3d_real_to_complex(prc, a, s, 0)
    CALL fftp3d_complex_to_real(pcr, s, b, 0)
    WRITE(fname, '(A,I0,A,I0,A,I0,A)') 'fortran_fftp_', mx, 'x', my, 'x', mz, '.bin'
    CALL report('fftp round trip '//TRIM(fname), MAXVAL(ABS(b/(REAL(mx, GP)*my*mz) - a)), 1.0E-13_GP)
    OPEN(NEWUNIT=uu, FILE=TRIM(outdir)//'/'//TRIM(fname), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
    WRITE(uu) a
    WRITE(uu) s
    CLOSE(uu)
    DEALLOCATE(a, b, s)
    CALL fftp3d_destroy_plan(prc)
    CALL fftp3d_destroy_plan(pcr)
  END SUBROUTINE fftp_shape

  SUBROUTINE report(what, err, tol)
    CHARACTER(LEN=*), INTENT(IN) :: what
    DOUBLE PRECISION, INTENT(IN) :: err
    REAL(KIND=GP), INTENT(IN) :: tol
    IF (err .LE. tol) THEN
      WRITE(*, '(A,A,A,ES10.3)') 'ok   ', what, ': ', err
    ELSE
      WRITE(*, '(A,A,A,ES10.3,A,ES10.3)') 'FAIL ', what, ': ', err, ' > ', tol
      STOP 1
    END IF
  END SUBROUTINE report

END PROGRAM test_fftp_hip
