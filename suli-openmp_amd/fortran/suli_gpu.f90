! ISO_C_BINDING shim between the SuLi-IC Fortran driver and libsuli.so.
!
! The reference driver (1_principal.f90:64-108) calls argument-less
! subroutines that work on module globals.  This module keeps the same
! subroutine names and the global time step dt (0_modules.f90:9, mutated per
! RK stage at 1_principal.f90:78), and forwards each call to the C ABI of
! include/suli.h.  Errors print the library's message and STOP, the
! reference's own convention (e.g. 6_convdiff.f90:81-82).
module suli_gpu
  use iso_c_binding
  implicit none

  type, bind(C) :: suli_params
    integer(c_int32_t) :: nx, ny, nz
    real(c_double) :: dx, dy, dz
    integer(c_int32_t) :: ccx0, ccxf, ccy0, ccyf, ccz0, cczf
    integer(c_int32_t) :: m_turb
    real(c_double) :: rho_f1, rho_f2, mi_f1, mi_f2, sigma, alpha1
    real(c_double) :: csmag, gx, gz, chezy
    integer(c_int32_t) :: precision, poisson, precond, max_it
    real(c_double) :: tol_abs, tol_rel
    integer(c_int32_t) :: device
  end type suli_params

  type, bind(C) :: suli_poisson_stats
    integer(c_int32_t) :: iters, outer, fallbacks, converged
    real(c_double) :: res0, res, ms
  end type suli_poisson_stats

  type, bind(C) :: suli_ls_stats
    real(c_double) :: vol_ini, vol_ins
    integer(c_int32_t) :: reinit_iters, mass_corr_iters
  end type suli_ls_stats

  ! field identifiers (include/suli.h)
  integer(c_int), parameter :: SULI_U = 0, SULI_V = 1, SULI_W = 2, SULI_PRD1 = 3, SULI_RHO = 4, &
                               SULI_LS = 6, SULI_HS = 9, SULI_NUT = 16

  interface
    integer(c_int) function c_suli_init(p) bind(C, name='suli_init')
      import :: suli_params, c_int
      type(suli_params), intent(in) :: p
    end function
    integer(c_int) function c_suli_finalize() bind(C, name='suli_finalize')
      import :: c_int
    end function
    integer(c_int) function c_suli_last_error(buf, n) bind(C, name='suli_last_error')
      import :: c_int, c_char
      character(kind=c_char), dimension(*) :: buf
      integer(c_int), value :: n
    end function
    integer(c_int) function c_suli_upload(f, a, n) bind(C, name='suli_upload')
      import :: c_int, c_double, c_int64_t
      integer(c_int), value :: f
      real(c_double), dimension(*), intent(in) :: a
      integer(c_int64_t), value :: n
    end function
    integer(c_int) function c_suli_download(f, a, n) bind(C, name='suli_download')
      import :: c_int, c_double, c_int64_t
      integer(c_int), value :: f
      real(c_double), dimension(*) :: a
      integer(c_int64_t), value :: n
    end function
    integer(c_int) function c_suli_init_dambreak(w, h, v) bind(C, name='suli_init_dambreak')
      import :: c_int, c_double
      real(c_double), value :: w, h
      real(c_double) :: v
    end function
    integer(c_int) function c_suli_level_set(dt, st) bind(C, name='suli_level_set')
      import :: c_int, c_double, suli_ls_stats
      real(c_double), value :: dt
      type(suli_ls_stats) :: st
    end function
    integer(c_int) function c_suli_contorno(nlock) bind(C, name='suli_contorno')
      import :: c_int
      integer(c_int), value :: nlock
    end function
    integer(c_int) function c_suli_visco() bind(C, name='suli_visco')
      import :: c_int
    end function
    integer(c_int) function c_suli_convdiff() bind(C, name='suli_convdiff')
      import :: c_int
    end function
    integer(c_int) function c_suli_tempo(tt, dt) bind(C, name='suli_tempo')
      import :: c_int, c_double
      integer(c_int), value :: tt
      real(c_double), value :: dt
    end function
    integer(c_int) function c_suli_graddin(dt, st) bind(C, name='suli_graddin')
      import :: c_int, c_double, suli_poisson_stats
      real(c_double), value :: dt
      type(suli_poisson_stats) :: st
    end function
    integer(c_int) function c_suli_posdin(dt) bind(C, name='suli_posdin')
      import :: c_int, c_double
      real(c_double), value :: dt
    end function
    integer(c_int) function c_suli_est_div(d) bind(C, name='suli_est_div')
      import :: c_int, c_double
      real(c_double) :: d
    end function
    integer(c_int) function c_suli_sync() bind(C, name='suli_sync')
      import :: c_int
    end function
    integer(c_int) function c_suli_set_time_scheme(t_tempo) bind(C, name='suli_set_time_scheme')
      import :: c_int
      integer(c_int), value :: t_tempo
    end function
    integer(c_int) function c_suli_set_obstacle(ku, kv, kw) bind(C, name='suli_set_obstacle')
      import :: c_int, c_int32_t
      integer(c_int32_t), intent(in) :: ku(*), kv(*), kw(*)
    end function
    integer(c_int) function c_suli_obstacle_columns(obst_t, nx, ny, nz, dx, dy, dz, ku, kv, kw) &
        bind(C, name='suli_obstacle_columns')
      import :: c_int, c_int32_t, c_double
      integer(c_int), value :: obst_t, nx, ny, nz
      real(c_double), value :: dx, dy, dz
      integer(c_int32_t) :: ku(*), kv(*), kw(*)
    end function
    integer(c_int) function c_suli_set_advection(adv_type, der) bind(C, name='suli_set_advection')
      import :: c_int
      integer(c_int), value :: adv_type, der
    end function
    integer(c_int) function c_suli_set_time(t) bind(C, name='suli_set_time')
      import :: c_int, c_double
      real(c_double), value :: t
    end function
    integer(c_int) function c_suli_set_forcing(amp, period, phase) bind(C, name='suli_set_forcing')
      import :: c_int, c_double
      real(c_double), value :: amp, period, phase
    end function
    integer(c_int) function c_suli_write_frame(path, t_plot) bind(C, name='suli_write_frame')
      import :: c_int, c_char
      character(kind=c_char), dimension(*) :: path
      integer(c_int), value :: t_plot
    end function
    integer(c_int) function c_suli_restart_write(path) bind(C, name='suli_restart_write')
      import :: c_int, c_char
      character(kind=c_char), dimension(*) :: path
    end function
    integer(c_int) function c_suli_restart_read(path) bind(C, name='suli_restart_read')
      import :: c_int, c_char
      character(kind=c_char), dimension(*) :: path
    end function
    integer(c_int) function c_suli_mms_set(mms_t, a, h0, coef) bind(C, name='suli_mms_set')
      import :: c_int, c_double
      integer(c_int), value :: mms_t
      real(c_double), value :: a, h0, coef
    end function
    integer(c_int) function c_suli_mms_i() bind(C, name='suli_mms_i')
      import :: c_int
    end function
    integer(c_int) function c_suli_mms_bc() bind(C, name='suli_mms_bc')
      import :: c_int
    end function
    integer(c_int) function c_suli_termo_fonte1() bind(C, name='suli_termo_fonte1')
      import :: c_int
    end function
    integer(c_int) function c_suli_termo_fonte2() bind(C, name='suli_termo_fonte2')
      import :: c_int
    end function
    integer(c_int) function c_suli_mms(err) bind(C, name='suli_mms')
      import :: c_int, c_double
      real(c_double), dimension(4) :: err
    end function
    integer(c_int) function c_suli_waves_coef(wave_t, a_w, p_w, h0_f, coef) bind(C, name='suli_waves_coef')
      import :: c_int, c_double
      integer(c_int), value :: wave_t
      real(c_double), value :: a_w, p_w, h0_f
      real(c_double), dimension(14) :: coef
    end function
    integer(c_int) function c_suli_boundary_waves() bind(C, name='suli_boundary_waves')
      import :: c_int
    end function
    integer(c_int) function c_suli_level_set_ini(v) bind(C, name='suli_level_set_ini')
      import :: c_int, c_double
      real(c_double) :: v
    end function
    integer(c_int) function c_suli_write_xdmf(path, nx, ny, nz, first, last, t_plot, dt_frame) &
        bind(C, name='suli_write_xdmf')
      import :: c_int, c_char, c_double
      character(kind=c_char), dimension(*) :: path
      integer(c_int), value :: nx, ny, nz, first, last, t_plot
      real(c_double), value :: dt_frame
    end function
  end interface

  ! driver state mirroring module disc (0_modules.f90:9,39-41)
  real(8) :: dt, t
  integer :: it, tt, ntt
  real(8), dimension(3) :: a_dt
  type(suli_params) :: prm
  type(suli_ls_stats) :: ls_st
  type(suli_poisson_stats) :: pois_st
  integer(8) :: pois_iters = 0, pois_solves = 0
  ! field output (module tempo, 0_modules.f90:155; t_plot, 0_modules.f90:22)
  integer(8) :: cont = 10
  integer :: t_plot = 1
  ! time scheme (0_modules.f90:24): 0 Euler, 1 RK2, 2 RK3, 3 AB2
  integer :: t_tempo = 1
  ! advection (0_modules.f90:26-27): der 1 upwind, 2 central, 3 upwind 2nd
  ! order; adv_type 1 classic, 2 rotational, 3 antisymmetric
  integer :: der = 1, adv_type = 1
  ! manufactured solution (0_modules.f90:37, module mms_m :204-219): mode,
  ! constants, accumulated mean error erro_t and the last errors of mms()
  integer :: mms_t = 0
  real(8) :: mms_a = 0.0d0, mms_h0 = 2.0d0, mms_coef = 0.0d0
  real(8) :: erro_t = 0.0d0
  real(8), dimension(4) :: mms_err = 0.0d0
  ! wave inflow (0_modules.f90:35, module wave_c :164-174): type, the
  ! constants of waves_coef (3_contorno.f90:1086-1088) and its coefficients
  integer :: wave_t = 0
  real(8) :: a_w = 0.03d0, p_w = 2.5d0, h0_f = 1.0d0
  real(8), dimension(14) :: wave_coef = 0.0d0

contains

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    character(kind=c_char), dimension(512) :: buf
    integer :: n, k
    integer(c_int) :: r2
    if (rc == 0) return
    buf = c_null_char
    r2 = c_suli_last_error(buf, 512_c_int)
    n = 0
    do k = 1, 512
      if (buf(k) == c_null_char) exit
      n = k
    enddo
    write(*,*) 'suli: ', what, ' failed: ', (buf(k), k = 1, n)
    stop 1
  end subroutine check

  ! ---- the reference subroutine names (1_principal.f90:74-92) ----------------
  subroutine level_set()
    call check(c_suli_level_set(dt, ls_st), 'level_set')
  end subroutine level_set

  subroutine contorno(nlock)
    integer, intent(in) :: nlock
    call check(c_suli_contorno(int(nlock, c_int)), 'contorno')
  end subroutine contorno

  subroutine visco()
    call check(c_suli_visco(), 'visco')
  end subroutine visco

  subroutine convdiff()
    ! the time of the step for a time-dependent body force (suli_set_forcing)
    call check(c_suli_set_time(t), 'set_time')
    call check(c_suli_convdiff(), 'convdiff')
  end subroutine convdiff

  ! The time-scheme part of iniciais (2_iniciais.f90:85-102): stage count
  ! ntt and stage steps a_dt from dt, and the scheme of the library's tempo;
  ! with it the advection scheme of the library's convdiff.
  subroutine tempo_ini()
    if ((t_tempo == 0) .or. (t_tempo == 3)) then
      ntt = 1
      a_dt = 1.0d0 * dt
    else
      ntt = 2
      if (t_tempo == 2) ntt = 3
      a_dt(1) = 0.5d0 * dt
      a_dt(2) = 1.0d0 * dt
      a_dt(3) = 1.0d0 * dt
    endif
    call check(c_suli_set_time_scheme(int(t_tempo, c_int)), 'set_time_scheme')
    call check(c_suli_set_advection(int(adv_type, c_int), int(der, c_int)), 'set_advection')
  end subroutine tempo_ini

  subroutine tempo()
    call check(c_suli_tempo(int(tt, c_int), dt), 'tempo')
  end subroutine tempo

  subroutine graddin()
    call check(c_suli_graddin(dt, pois_st), 'graddin')
    pois_iters = pois_iters + pois_st%iters
    pois_solves = pois_solves + 1
  end subroutine graddin

  subroutine posdin()
    call check(c_suli_posdin(dt), 'posdin')
  end subroutine posdin

  ! est (4_plot.f90:500-525): max abs divergence and volume, one line per step
  subroutine est(div)
    real(8), intent(out) :: div
    call check(c_suli_est_div(div), 'est')
  end subroutine est

  ! field output of plot_i / plot_f (4_plot.f90:131,184-215,260-363): the
  ! record arquivos/campos_NNNNN of the current fields, NNNNN = cont; plot_f
  ! advances the counter first, as the reference does
  subroutine plot_i()
    call write_campos()
  end subroutine plot_i

  subroutine plot_f()
    cont = cont + 1
    call write_campos()
  end subroutine plot_f

  subroutine write_campos()
    character(len=32) :: name
    call execute_command_line('mkdir -p arquivos')
    write(name, '(A,I5.5,A)') 'arquivos/campos_', cont, c_null_char
    call check(c_suli_write_frame(name, int(t_plot, c_int)), 'plot')
  end subroutine write_campos

  ! visu (12_paraview_3D.f90:1-179): the ParaView index of campos_00010..cont
  subroutine visu(dt_frame)
    real(8), intent(in) :: dt_frame
    call check(c_suli_write_xdmf('arquivos/visu.xdmf'//c_null_char, prm%nx, prm%ny, prm%nz, 10_c_int, &
                                 int(cont, c_int), int(t_plot, c_int), dt_frame), 'visu')
  end subroutine visu

  ! restart_salva / restart_ini (11_diversos.f90:232-329): the reference's
  ! restart file dados/arquivorestart, written from and read into the GPU state
  subroutine restart_salva()
    call execute_command_line('mkdir -p dados')
    call check(c_suli_restart_write('dados/arquivorestart'//c_null_char), 'restart_salva')
    write(*,*) 'Salvou restart'
  end subroutine restart_salva

  subroutine restart_ini()
    call check(c_suli_restart_read('dados/arquivorestart'//c_null_char), 'restart_ini')
    it = 0
    t = 0.
    cont = 10
  end subroutine restart_ini

  ! ---- the manufactured solution (13_mms.f90) --------------------------------
  ! mms_i (:12-212) with level_set_ini's tail (14_level_set.f90:205-217)
  subroutine mms_i()
    real(8) :: vol
    call check(c_suli_mms_set(int(mms_t, c_int), mms_a, mms_h0, mms_coef), 'mms_set')
    call check(c_suli_set_time(t), 'set_time')
    call check(c_suli_mms_i(), 'mms_i')
    call check(c_suli_level_set_ini(vol), 'level_set_ini')
    erro_t = 0.0d0
  end subroutine mms_i

  ! mms_bc (:468-816); contorno(1|2) also sets the planes first
  subroutine mms_bc()
    call check(c_suli_set_time(t), 'set_time')
    call check(c_suli_mms_bc(), 'mms_bc')
  end subroutine mms_bc

  subroutine termo_fonte1()
    call check(c_suli_set_time(t), 'set_time')
    call check(c_suli_termo_fonte1(), 'termo_fonte1')
  end subroutine termo_fonte1

  subroutine termo_fonte2()
    call check(c_suli_set_time(t), 'set_time')
    call check(c_suli_termo_fonte2(), 'termo_fonte2')
  end subroutine termo_fonte2

  ! mms (:218-465): the RMS errors of u, v, w, p and the running mean of the
  ! velocity errors (erro_t, :459)
  subroutine mms()
    call check(c_suli_set_time(t), 'set_time')
    call check(c_suli_mms(mms_err), 'mms')
    erro_t = (mms_err(1) + mms_err(2) + mms_err(3)) / 3.0d0 + erro_t
  end subroutine mms

  ! ---- wave inflow (3_contorno.f90:952-1205) ----------------------------------
  subroutine waves_coef()
    call check(c_suli_waves_coef(int(wave_t, c_int), a_w, p_w, h0_f, wave_coef), 'waves_coef')
  end subroutine waves_coef

  ! boundary_waves at the step's time (the driver calls it after tempo when
  ! wave_t > 0, 1_principal.f90:84; suli_stage does the same by itself)
  subroutine boundary_waves()
    call check(c_suli_set_time(t), 'set_time')
    call check(c_suli_boundary_waves(), 'boundary_waves')
  end subroutine boundary_waves

  ! host copy of a field in the reference extents, for output
  subroutine download(field, a)
    integer(c_int), intent(in) :: field
    real(8), dimension(:,:,:), intent(inout) :: a
    call check(c_suli_download(field, a, int(size(a), c_int64_t)), 'download')
  end subroutine download

  subroutine upload(field, a)
    integer(c_int), intent(in) :: field
    real(8), dimension(:,:,:), intent(in) :: a
    call check(c_suli_upload(field, a, int(size(a), c_int64_t)), 'upload')
  end subroutine upload

end module suli_gpu
