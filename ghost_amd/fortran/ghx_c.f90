! ghx_c.f90 -- iso_c_binding interfaces to the C ABI of include/ghx.h.
!
! Every ghx_* function returns 0 on success or a negative status; use
! ghx_check to turn a failure into a message and a non-zero STOP.  Field
! arguments are device pointers (TYPE(C_PTR)) unless the name ends in _host.
MODULE ghx_c
  USE iso_c_binding
  IMPLICIT NONE

  INTEGER(C_INT), PARAMETER :: GHX_OK = 0

  INTERFACE
    FUNCTION ghx_last_error() BIND(C, name="ghx_last_error")
      IMPORT :: C_PTR
      TYPE(C_PTR) :: ghx_last_error
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_device_ok() BIND(C, name="ghx_device_ok")
      IMPORT :: C_INT
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_malloc(ptr, bytes) BIND(C, name="ghx_malloc")
      IMPORT :: C_INT, C_PTR, C_SIZE_T
      TYPE(C_PTR) :: ptr
      INTEGER(C_SIZE_T), VALUE :: bytes
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_free(ptr) BIND(C, name="ghx_free")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: ptr
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_memcpy_h2d(dst, src, bytes) BIND(C, name="ghx_memcpy_h2d")
      IMPORT :: C_INT, C_PTR, C_SIZE_T
      TYPE(C_PTR), VALUE :: dst, src
      INTEGER(C_SIZE_T), VALUE :: bytes
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_memcpy_d2h(dst, src, bytes) BIND(C, name="ghx_memcpy_d2h")
      IMPORT :: C_INT, C_PTR, C_SIZE_T
      TYPE(C_PTR), VALUE :: dst, src
      INTEGER(C_SIZE_T), VALUE :: bytes
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_memcpy_d2d(dst, src, bytes) BIND(C, name="ghx_memcpy_d2d")
      IMPORT :: C_INT, C_PTR, C_SIZE_T
      TYPE(C_PTR), VALUE :: dst, src
      INTEGER(C_SIZE_T), VALUE :: bytes
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_sync() BIND(C, name="ghx_sync")
      IMPORT :: C_INT
    END FUNCTION

    ! plans (fftp-3/fftp3D.fpp:56-155)
    INTEGER(C_INT) FUNCTION ghx_plan_create(plan, nx, ny, nz, fftdir, flags) BIND(C, name="ghx_plan_create")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR) :: plan
      INTEGER(C_INT), VALUE :: nx, ny, nz, fftdir, flags
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_plan_destroy(plan) BIND(C, name="ghx_plan_destroy")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: plan
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_r2c(plan, rin, cout) BIND(C, name="ghx_r2c")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: plan, rin, cout
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_c2r(plan, cin, rout) BIND(C, name="ghx_c2r")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: plan, cin, rout
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_r2c_host(plan, rin, cout) BIND(C, name="ghx_r2c_host")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: plan, rin, cout
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_c2r_host(plan, cin, rout) BIND(C, name="ghx_c2r_host")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: plan, cin, rout
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_plan_timers(plan, ffttime, tratime, comtime, tottime) BIND(C, name="ghx_plan_timers")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: plan
      REAL(C_DOUBLE) :: ffttime, tratime, comtime, tottime
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_plan_reset_timers(plan) BIND(C, name="ghx_plan_reset_timers")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: plan
    END FUNCTION

    ! grid (main3D.fpp:737-749, 1727-1800)
    INTEGER(C_INT) FUNCTION ghx_grid_create(grid, nx, ny, nz, Lx, Ly, Lz, arbsize, Dkk) BIND(C, name="ghx_grid_create")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR) :: grid
      INTEGER(C_INT), VALUE :: nx, ny, nz, arbsize
      REAL(C_DOUBLE), VALUE :: Lx, Ly, Lz, Dkk
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_grid_destroy(grid) BIND(C, name="ghx_grid_destroy")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: grid
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_grid_r2c(grid, rin, cout) BIND(C, name="ghx_grid_r2c")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: grid, rin, cout
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_grid_c2r(grid, cin, rout) BIND(C, name="ghx_grid_c2r")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: grid, cin, rout
    END FUNCTION

    ! operators (pseudo/pseudospec3D_hd.f90, _mhd.f90, _filt.f90)
    INTEGER(C_INT) FUNCTION ghx_derivk3(g, a, b, dir) BIND(C, name="ghx_derivk3")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: g, a, b
      INTEGER(C_INT), VALUE :: dir
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_laplak3(g, a, b) BIND(C, name="ghx_laplak3")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: g, a, b
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_rotor3(g, a, b, c, dir) BIND(C, name="ghx_rotor3")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: g, a, b, c
      INTEGER(C_INT), VALUE :: dir
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_prodre3(g, a, b, c, d, e, f) BIND(C, name="ghx_prodre3")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: g, a, b, c, d, e, f
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_nonlhd3(g, a, b, c, gout, dir) BIND(C, name="ghx_nonlhd3")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: g, a, b, c, gout
      INTEGER(C_INT), VALUE :: dir
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_vector3(g, a, b, c, d, e, f, x, y, z) BIND(C, name="ghx_vector3")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: g, a, b, c, d, e, f, x, y, z
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_nonlin3(g, a, b, c, d, e, f, gout, dir) BIND(C, name="ghx_nonlin3")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: g, a, b, c, d, e, f, gout
      INTEGER(C_INT), VALUE :: dir
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_gauge3(g, a, b, c, gout, dir) BIND(C, name="ghx_gauge3")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: g, a, b, c, gout
      INTEGER(C_INT), VALUE :: dir
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_energy(g, a, b, c, kin, out) BIND(C, name="ghx_energy")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: g, a, b, c
      INTEGER(C_INT), VALUE :: kin
      REAL(C_DOUBLE) :: out
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_helicity(g, a, b, c, out) BIND(C, name="ghx_helicity")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: g, a, b, c
      REAL(C_DOUBLE) :: out
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_cross(g, a, b, c, d, e, f, kin, out) BIND(C, name="ghx_cross")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: g, a, b, c, d, e, f
      INTEGER(C_INT), VALUE :: kin
      REAL(C_DOUBLE) :: out
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_maxabs(g, a, b, c, kin, out) BIND(C, name="ghx_maxabs")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: g, a, b, c
      INTEGER(C_INT), VALUE :: kin
      REAL(C_DOUBLE) :: out
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_spectrumc(g, a, b, c, kin, hel, Ek, Hk) BIND(C, name="ghx_spectrumc")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: g, a, b, c
      INTEGER(C_INT), VALUE :: kin, hel
      REAL(C_DOUBLE) :: Ek(*), Hk(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_grid_info(g, kmax, tiny, nmax, Dkk, anis) BIND(C, name="ghx_grid_info")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: g
      REAL(C_DOUBLE) :: kmax, tiny, Dkk
      INTEGER(C_INT) :: nmax, anis
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_entrans(g, a, b, c, d, e, f, kin, Ek) BIND(C, name="ghx_entrans")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: g, a, b, c, d, e, f
      INTEGER(C_INT), VALUE :: kin
      REAL(C_DOUBLE) :: Ek(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_heltrans(g, a, b, c, d, e, f, Hk) BIND(C, name="ghx_heltrans")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: g, a, b, c, d, e, f
      REAL(C_DOUBLE) :: Hk(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_crosspecc(g, a, b, c, d, e, f, shift, Ck) BIND(C, name="ghx_crosspecc")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: g, a, b, c, d, e, f
      REAL(C_DOUBLE), VALUE :: shift
      REAL(C_DOUBLE) :: Ck(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_spectr1d(g, a, dir, Ek) BIND(C, name="ghx_spectr1d")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: g, a
      INTEGER(C_INT), VALUE :: dir
      REAL(C_DOUBLE) :: Ek(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_hdcheck(g, a, b, c, d, e, f, hel, chk, out) BIND(C, name="ghx_hdcheck")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: g, a, b, c, d, e, f
      INTEGER(C_INT), VALUE :: hel, chk
      REAL(C_DOUBLE) :: out(5)
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_mhdcheck(g, a, b, c, ma, mb, mc, hel, crs, chk, out) BIND(C, name="ghx_mhdcheck")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: g, a, b, c, ma, mb, mc
      INTEGER(C_INT), VALUE :: hel, crs, chk
      REAL(C_DOUBLE) :: out(11)
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_variance(g, a, kin, out) &
        BIND(C, name="ghx_variance")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: g, a
      INTEGER(C_INT), VALUE :: kin
      REAL(C_DOUBLE) :: out
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_product(g, a, b, out) &
        BIND(C, name="ghx_product")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: g, a, b
      REAL(C_DOUBLE) :: out
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_pscheck(g, a, b, out) &
        BIND(C, name="ghx_pscheck")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: g, a, b
      REAL(C_DOUBLE) :: out(3)
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_spectrscc(g, a, shift, Ek) &
        BIND(C, name="ghx_spectrscc")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: g, a
      REAL(C_DOUBLE), VALUE :: shift
      REAL(C_DOUBLE) :: Ek(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_sctrans(g, a, b, Ek) &
        BIND(C, name="ghx_sctrans")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: g, a, b
      REAL(C_DOUBLE) :: Ek(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_difucx(g, a, b, out) &
        BIND(C, name="ghx_difucx")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: g, a, b
      REAL(C_DOUBLE) :: out(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_difucz(g, a, b, out) &
        BIND(C, name="ghx_difucz")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: g, a, b
      REAL(C_DOUBLE) :: out(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_axis_bins(g, nbpara, nbperp) &
        BIND(C, name="ghx_axis_bins")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: g
      INTEGER(C_INT) :: nbpara, nbperp
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_specpara(g, a, b, c, kin, hel, Ek, Hk) &
        BIND(C, name="ghx_specpara")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: g, a, b, c
      INTEGER(C_INT), VALUE :: kin, hel
      REAL(C_DOUBLE) :: Ek(*), Hk(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_specperp(g, a, b, c, kin, hel, Ek, Hk) &
        BIND(C, name="ghx_specperp")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: g, a, b, c
      INTEGER(C_INT), VALUE :: kin, hel
      REAL(C_DOUBLE) :: Ek(*), Hk(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_spec2d(g, a, b, c, kin, hel, Ek, Hk) &
        BIND(C, name="ghx_spec2d")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: g, a, b, c
      INTEGER(C_INT), VALUE :: kin, hel
      REAL(C_DOUBLE) :: Ek(*), Hk(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_entpara(g, a, b, c, d, e, f, kin, Ek) &
        BIND(C, name="ghx_entpara")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: g, a, b, c, d, e, f
      INTEGER(C_INT), VALUE :: kin
      REAL(C_DOUBLE) :: Ek(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_entperp(g, a, b, c, d, e, f, kin, Ek) &
        BIND(C, name="ghx_entperp")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: g, a, b, c, d, e, f
      INTEGER(C_INT), VALUE :: kin
      REAL(C_DOUBLE) :: Ek(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_heltpara(g, a, b, c, d, e, f, Hk) &
        BIND(C, name="ghx_heltpara")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: g, a, b, c, d, e, f
      REAL(C_DOUBLE) :: Hk(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_heltperp(g, a, b, c, d, e, f, Hk) &
        BIND(C, name="ghx_heltperp")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: g, a, b, c, d, e, f
      REAL(C_DOUBLE) :: Hk(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_havgcomp(g, u, v, w, s, fo, bv, itype, gsh) &
        BIND(C, name="ghx_havgcomp")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: g, u, v, w, s
      REAL(C_DOUBLE), VALUE :: fo, bv
      INTEGER(C_INT), VALUE :: itype
      REAL(C_DOUBLE) :: gsh(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_tbouss(g, u, v, w, s, fo, bv, avg, mx, mn) &
        BIND(C, name="ghx_tbouss")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: g, u, v, w, s
      REAL(C_DOUBLE), VALUE :: fo, bv
      REAL(C_DOUBLE) :: avg(11), mx(11), mn(11)
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_spectpv(g, u, v, w, s, Ek) &
        BIND(C, name="ghx_spectpv")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: g, u, v, w, s
      REAL(C_DOUBLE) :: Ek(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_phaseshift(g, a, cre, cim) BIND(C, name="ghx_phaseshift")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: g, a
      REAL(C_DOUBLE), VALUE :: cre, cim
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_normalize(g, fx, fy, fz, f0, kin) BIND(C, name="ghx_normalize")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: g, fx, fy, fz
      REAL(C_DOUBLE), VALUE :: f0
      INTEGER(C_INT), VALUE :: kin
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_sharp(g, a, b, delta) BIND(C, name="ghx_sharp")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: g, a, b
      REAL(C_DOUBLE), VALUE :: delta
    END FUNCTION

    ! forcing and initial conditions
    REAL(C_DOUBLE) FUNCTION ghx_randu(seed) BIND(C, name="ghx_randu")
      IMPORT :: C_DOUBLE, C_INT32_T
      INTEGER(C_INT32_T) :: seed
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_initialv_tg(g, kdn, kup, u0, vx, vy, vz) BIND(C, name="ghx_initialv_tg")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: g, vx, vy, vz
      REAL(C_DOUBLE), VALUE :: kdn, kup, u0
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_initialfv_random(g, kdn, kup, f0, seed, fx, fy, fz) &
        BIND(C, name="ghx_initialfv_random")
      IMPORT :: C_INT, C_PTR, C_DOUBLE, C_INT32_T
      TYPE(C_PTR), VALUE :: g, fx, fy, fz
      REAL(C_DOUBLE), VALUE :: kdn, kup, f0
      INTEGER(C_INT32_T) :: seed
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_initialv_k4exp(g, k0, u0, seed, vx, vy, vz) BIND(C, name="ghx_initialv_k4exp")
      IMPORT :: C_INT, C_PTR, C_DOUBLE, C_INT32_T
      TYPE(C_PTR), VALUE :: g, vx, vy, vz
      REAL(C_DOUBLE), VALUE :: k0, u0
      INTEGER(C_INT32_T) :: seed
    END FUNCTION

    ! RK substep and resident solvers (include/hd_rkstep2.f90, main3D.fpp:3202-3267)
    INTEGER(C_INT) FUNCTION ghx_mhd_rkstep2(g, o, dt, nu, mu, v_in, c, f, m, v_out) &
        BIND(C, name="ghx_mhd_rkstep2")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: g
      INTEGER(C_INT), VALUE :: o
      REAL(C_DOUBLE), VALUE :: dt, nu, mu
      TYPE(C_PTR) :: v_in(6), c(6), v_out(6)
      TYPE(C_PTR), VALUE :: f, m
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_bouss_rkstep2(g, o, dt, nu, kappa, xmom, xtemp, v_in, c, f, v_out) &
        BIND(C, name="ghx_bouss_rkstep2")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: g
      INTEGER(C_INT), VALUE :: o
      REAL(C_DOUBLE), VALUE :: dt, nu, kappa, xmom, xtemp
      TYPE(C_PTR) :: v_in(4), c(4), v_out(4)
      TYPE(C_PTR), VALUE :: f
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_rotbouss_rkstep2(g, o, dt, nu, kappa, xmom, xtemp, omegaz, v_in, c, f, v_out) &
        BIND(C, name="ghx_rotbouss_rkstep2")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: g
      INTEGER(C_INT), VALUE :: o
      REAL(C_DOUBLE), VALUE :: dt, nu, kappa, xmom, xtemp, omegaz
      TYPE(C_PTR) :: v_in(4), c(4), v_out(4)
      TYPE(C_PTR), VALUE :: f
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_bouss_ord3(g, dt, xmom, c, v) BIND(C, name="ghx_bouss_ord3")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: g
      REAL(C_DOUBLE), VALUE :: dt, xmom
      TYPE(C_PTR) :: c(4), v(3)
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_roth_rkstep2(g, o, dt, nu, omegaz, v_in, c, f, v_out) &
        BIND(C, name="ghx_roth_rkstep2")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: g
      INTEGER(C_INT), VALUE :: o
      REAL(C_DOUBLE), VALUE :: dt, nu, omegaz
      TYPE(C_PTR) :: v_in(3), c(3), v_out(3)
      TYPE(C_PTR), VALUE :: f
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_hd_rkstep2(g, o, dt, nu, v_in, c, f, v_out) BIND(C, name="ghx_hd_rkstep2")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: g
      INTEGER(C_INT), VALUE :: o
      REAL(C_DOUBLE), VALUE :: dt, nu
      TYPE(C_PTR) :: v_in(3), c(3), v_out(3)
      TYPE(C_PTR), VALUE :: f
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_hd_create(s, g, ord, dt, nu) BIND(C, name="ghx_hd_create")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR) :: s
      TYPE(C_PTR), VALUE :: g
      INTEGER(C_INT), VALUE :: ord
      REAL(C_DOUBLE), VALUE :: dt, nu
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_hd_destroy(s) BIND(C, name="ghx_hd_destroy")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: s
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_hd_fields(s, v, f) BIND(C, name="ghx_hd_fields")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: s
      TYPE(C_PTR) :: v(3), f(3)
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_hd_touch(s) BIND(C, name="ghx_hd_touch")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: s
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_hd_set_rotation(s, omegaz) BIND(C, name="ghx_hd_set_rotation")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: s
      REAL(C_DOUBLE), VALUE :: omegaz
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_hd_set_forcing(s, enabled) BIND(C, name="ghx_hd_set_forcing")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: s
      INTEGER(C_INT), VALUE :: enabled
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_hd_step(s, nsteps) BIND(C, name="ghx_hd_step")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: s
      INTEGER(C_INT), VALUE :: nsteps
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_hd_set_tracers(s, t) BIND(C, name="ghx_hd_set_tracers")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: s, t
    END FUNCTION

    ! Lagrangian tracers (lagpart/gparts_mod.f90)
    INTEGER(C_INT) FUNCTION ghx_tracers_create(t, g, n, zcompat) BIND(C, name="ghx_tracers_create")
      IMPORT :: C_INT, C_PTR, C_LONG_LONG
      TYPE(C_PTR) :: t
      TYPE(C_PTR), VALUE :: g
      INTEGER(C_LONG_LONG), VALUE :: n
      INTEGER(C_INT), VALUE :: zcompat
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_tracers_destroy(t) BIND(C, name="ghx_tracers_destroy")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: t
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_tracers_positions(t, px, py, pz) BIND(C, name="ghx_tracers_positions")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: t
      TYPE(C_PTR) :: px, py, pz
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_tracers_set_step(t) BIND(C, name="ghx_tracers_set_step")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: t
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_tracers_step(t, v, dt, xk) BIND(C, name="ghx_tracers_step")
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: t
      TYPE(C_PTR) :: v(3)
      REAL(C_DOUBLE), VALUE :: dt, xk
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_tracers_euler_to_lag(t, field, out) BIND(C, name="ghx_tracers_euler_to_lag")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: t, field, out
    END FUNCTION

    ! binary field files (mpiio/binary_io.f90); strings are NUL-terminated
    INTEGER(C_INT) FUNCTION ghx_io_init(plan, myrank, n, ksta, kend) BIND(C, name="ghx_io_init")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR) :: plan
      INTEGER(C_INT), VALUE :: myrank, ksta, kend
      INTEGER(C_INT) :: n(3)
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_io_initc(plan, myrank, n, ksta, kend) BIND(C, name="ghx_io_initc")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR) :: plan
      INTEGER(C_INT), VALUE :: myrank, ksta, kend
      INTEGER(C_INT) :: n(3)
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_io_destroy(plan) BIND(C, name="ghx_io_destroy")
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: plan
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_io_set_options(plan, iswap, bmangle, disp) BIND(C, name="ghx_io_set_options")
      IMPORT :: C_INT, C_PTR, C_LONG_LONG
      TYPE(C_PTR), VALUE :: plan
      INTEGER(C_INT), VALUE :: iswap, bmangle
      INTEGER(C_LONG_LONG), VALUE :: disp
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_io_read(plan, dir, fname, nmb, var) BIND(C, name="ghx_io_read")
      IMPORT :: C_INT, C_PTR, C_CHAR
      TYPE(C_PTR), VALUE :: plan, var
      CHARACTER(KIND=C_CHAR) :: dir(*), fname(*), nmb(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_io_write(plan, dir, fname, nmb, var) BIND(C, name="ghx_io_write")
      IMPORT :: C_INT, C_PTR, C_CHAR
      TYPE(C_PTR), VALUE :: plan, var
      CHARACTER(KIND=C_CHAR) :: dir(*), fname(*), nmb(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_io_readc(plan, dir, fname, nmb, var) BIND(C, name="ghx_io_readc")
      IMPORT :: C_INT, C_PTR, C_CHAR
      TYPE(C_PTR), VALUE :: plan, var
      CHARACTER(KIND=C_CHAR) :: dir(*), fname(*), nmb(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_io_writec(plan, dir, fname, nmb, var) BIND(C, name="ghx_io_writec")
      IMPORT :: C_INT, C_PTR, C_CHAR
      TYPE(C_PTR), VALUE :: plan, var
      CHARACTER(KIND=C_CHAR) :: dir(*), fname(*), nmb(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_hd_output(grid, io, odir, ext, outs, vx, vy, vz) BIND(C, name="ghx_hd_output")
      IMPORT :: C_INT, C_PTR, C_CHAR
      TYPE(C_PTR), VALUE :: grid, io, vx, vy, vz
      CHARACTER(KIND=C_CHAR) :: odir(*), ext(*)
      INTEGER(C_INT), VALUE :: outs
    END FUNCTION
    INTEGER(C_INT) FUNCTION ghx_hd_restart(grid, io, idir, ext, vx, vy, vz) BIND(C, name="ghx_hd_restart")
      IMPORT :: C_INT, C_PTR, C_CHAR
      TYPE(C_PTR), VALUE :: grid, io, vx, vy, vz
      CHARACTER(KIND=C_CHAR) :: idir(*), ext(*)
    END FUNCTION
  END INTERFACE

CONTAINS

  ! GHOST has no status codes; a failing call here prints the library's
  ! message and stops with a non-zero exit status.
  SUBROUTINE ghx_check(rc, what)
    INTEGER(C_INT), INTENT(IN) :: rc
    CHARACTER(LEN=*), INTENT(IN) :: what
    CHARACTER(KIND=C_CHAR), POINTER :: msg(:)
    CHARACTER(LEN=512) :: text
    INTEGER :: i
    IF (rc .EQ. GHX_OK) RETURN
    text = ''
    CALL C_F_POINTER(ghx_last_error(), msg, [512])
    DO i = 1, 512
      IF (msg(i) .EQ. C_NULL_CHAR) EXIT
      text(i:i) = msg(i)
    END DO
    WRITE(*, '(A,A,A,I0,A,A)') 'ghx: ', what, ' failed (status ', rc, '): ', TRIM(text)
    STOP 1
  END SUBROUTINE ghx_check

END MODULE ghx_c
