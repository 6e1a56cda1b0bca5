! SuLi-IC driver on the MI355X library: the time loop of the reference
! PROGRAM PNH (1_principal.f90:64-108) with its CALL lines unchanged; the
! subroutines come from module suli_gpu and run on the GPU.
!
!   principal_gpu NX NY NZ STEPS [PRECOND [WIDTH HEIGHT [FRAME [T_TEMPO [OBST_T [MMS_T]]]]]]
!
! Tank 1.0 x 0.5 x 0.5 m scaled to the grid (h = 1/NX), dam break of a water
! column WIDTH x HEIGHT in the x = 0 corner, water/air 1000/1.2 kg/m^3, no-slip
! walls, Smagorinsky Cs = 0.1, mixed precision, FCG projection to 1e-8 with
! PRECOND 0 Jacobi, 1 multigrid (default), 2 learned.  With FRAME > 0 the
! field output of the reference is written: arquivos/campos_00010 at the
! start (plot_i), one record every FRAME steps (plot_f, 4_plot.f90:260) and
! arquivos/visu.xdmf at the end (visu); every frame the velocities, pressure
! and level set are also copied back into the host arrays u, v, w, prd1, ls
! with the reference extents (0_modules.f90:92-122).  T_TEMPO selects the
! time scheme (0_modules.f90:24): 0 Euler, 1 RK2 (default), 2 RK3, 3 AB2.
! OBST_T > 0 puts the reference's bottom obstacle of that type (obstaculo,
! 3_contorno.f90:494-760) into the tank.  MMS_T > 0 runs the manufactured
! solution of 13_mms.f90 instead (1 steady, 2 unsteady with coef = 1): tank
! 2 pi x 2 pi x 2.5 on NX x NY x NZ cells, depth h0 = 2, a = 0 (the
! reference's setting), every wall prescribed (code 3), no turbulence model,
! no gravity, dt = 0.05 dx, no projection; prints the errors of mms() per step.
program principal_gpu
  use suli_gpu
  implicit none
  integer :: nx, ny, nz, ts, precond, frame, nargs, obst_t, mms_mode
  integer(c_int32_t), allocatable :: ku(:,:), kv(:,:), kw(:,:)
  real(8) :: width, height, h, div, vol_ini
  character(len=64) :: arg
  integer(8) :: c0, c1, crate
  real(8), allocatable :: u(:,:,:), v(:,:,:), w(:,:,:), prd1(:,:,:), ls(:,:,:)

  nargs = command_argument_count()
  if (nargs < 4) then
    write(*,*) 'usage: principal_gpu NX NY NZ STEPS [PRECOND [WIDTH HEIGHT [FRAME [T_TEMPO [OBST_T [MMS_T]]]]]]'
    stop 2
  endif
  call get_command_argument(1, arg); read(arg, *) nx
  call get_command_argument(2, arg); read(arg, *) ny
  call get_command_argument(3, arg); read(arg, *) nz
  call get_command_argument(4, arg); read(arg, *) ts
  precond = 1
  width = 0.25d0
  height = 0.25d0
  frame = 0
  if (nargs >= 5) then
    call get_command_argument(5, arg); read(arg, *) precond
  endif
  if (nargs >= 7) then
    call get_command_argument(6, arg); read(arg, *) width
    call get_command_argument(7, arg); read(arg, *) height
  endif
  if (nargs >= 8) then
    call get_command_argument(8, arg); read(arg, *) frame
  endif
  if (nargs >= 9) then
    call get_command_argument(9, arg); read(arg, *) t_tempo
  endif
  obst_t = 0
  if (nargs >= 10) then
    call get_command_argument(10, arg); read(arg, *) obst_t
  endif
  mms_mode = 0
  if (nargs >= 11) then
    call get_command_argument(11, arg); read(arg, *) mms_mode
  endif
  h = 1.0d0 / nx

  ! iniciais (2_iniciais.f90:10-106) and level_set_ini (14_level_set.f90:10-219)
  prm%nx = nx; prm%ny = ny; prm%nz = nz
  prm%dx = h; prm%dy = h; prm%dz = h
  prm%ccx0 = 2; prm%ccxf = 2; prm%ccy0 = 2; prm%ccyf = 2; prm%ccz0 = 2; prm%cczf = 2
  prm%m_turb = 1
  prm%rho_f1 = 1.2d0; prm%rho_f2 = 1000.0d0; prm%mi_f1 = 1.8d-5; prm%mi_f2 = 1.0d-3
  prm%sigma = 0.0728d0; prm%alpha1 = 1.5d0
  prm%csmag = 0.1d0; prm%gx = 0.0d0; prm%gz = 9.81d0; prm%chezy = 44.5d0
  prm%precision = 1; prm%poisson = 1; prm%precond = precond; prm%max_it = 0
  prm%tol_abs = 0.0d0; prm%tol_rel = 1.0d-8; prm%device = 0
  if (mms_mode > 0) then
    call run_mms()
    stop
  endif
  call check(c_suli_init(prm), 'init')
  call check(c_suli_init_dambreak(width, height, vol_ini), 'init_dambreak')
  if (obst_t > 0) then  ! obstaculo (2_iniciais.f90:49) and the library's columns
    allocate(ku(0:nx+2,0:ny+1), kv(0:nx+1,0:ny+2), kw(0:nx+1,0:ny+1))
    call check(c_suli_obstacle_columns(int(obst_t, c_int), int(nx, c_int), int(ny, c_int), int(nz, c_int), &
                                       h, h, h, ku, kv, kw), 'obstacle_columns')
    call check(c_suli_set_obstacle(ku, kv, kw), 'set_obstacle')
  endif
  dt = 0.2d0 * h / (2.0d0 * sqrt(prm%gz * height))
  call tempo_ini()
  allocate(u(0:nx+2,0:ny+1,0:nz+1), v(0:nx+1,0:ny+2,0:nz+1), w(0:nx+1,0:ny+1,0:nz+2))
  allocate(prd1(0:nx+1,0:ny+1,0:nz+1), ls(nx,ny,nz))

  call contorno(1)
  call contorno(3)
  if (frame > 0) call plot_i()

  call system_clock(c0, crate)
  do it = 1, ts
    t = it * dt

    call level_set()
    call contorno(3)

    do tt = 1, ntt
      dt = a_dt(tt)

      call visco()
      call convdiff()
      call tempo()

      call contorno(2)

      call graddin()
      call posdin()
      call contorno(1)
    enddo

    call est(div)
    write(*,'(A,I6,A,ES12.5,A,ES12.5,A,ES12.5,A,I5)') 'step', it, ' t=', t, ' vol_err=', &
      (ls_st%vol_ins - ls_st%vol_ini) / ls_st%vol_ini, ' maxdiv=', div, ' poisson_iters=', pois_st%iters
    if (frame > 0) then
      if (mod(it, frame) == 0) then
        call plot_f()
        call download(SULI_U, u)
        call download(SULI_V, v)
        call download(SULI_W, w)
        call download(SULI_PRD1, prd1)
        call download(SULI_LS, ls)
        write(*,'(A,I6,A,ES12.5,A,ES12.5)') 'frame', it, ' max|u|=', maxval(abs(u)), ' min ls=', minval(ls)
      endif
    endif
  enddo
  call check(c_suli_sync(), 'sync')
  call system_clock(c1)
  write(*,'(A,F10.3,A,F10.2,A,F8.2)') 'seconds=', real(c1 - c0, 8) / real(crate, 8), &
    ' Mcell-steps/s=', real(nx, 8) * ny * nz * ts / (real(c1 - c0, 8) / real(crate, 8)) / 1.0d6, &
    ' poisson_iters/solve=', real(pois_iters, 8) / max(pois_solves, 1_8)
  if (frame > 0) call visu(frame * dt)
  call check(c_suli_finalize(), 'finalize')

contains

  ! the manufactured-solution run of 1_principal.f90:39-108 with mms_t > 0
  subroutine run_mms()
    real(8), parameter :: pi = acos(-1.0d0)
    prm%dx = 2.0d0 * pi / nx; prm%dy = 2.0d0 * pi / ny; prm%dz = 2.5d0 / nz
    prm%ccx0 = 3; prm%ccxf = 3; prm%ccy0 = 3; prm%ccyf = 3; prm%ccz0 = 3; prm%cczf = 3
    prm%m_turb = 0; prm%gz = 0.0d0
    prm%precision = 0  ! FP64
    call check(c_suli_init(prm), 'init')
    mms_t = mms_mode
    mms_a = 0.0d0; mms_h0 = 2.0d0
    mms_coef = merge(1.0d0, 0.0d0, mms_t == 2)
    dt = 0.05d0 * prm%dx
    call tempo_ini()
    t = 0.0d0
    call mms_i()
    call contorno(1)
    call contorno(3)
    call mms()
    do it = 1, ts
      t = it * dt
      if ((mms_t == 1) .and. (it == 1)) call termo_fonte1()
      if (mms_t == 2) call termo_fonte2()
      call level_set()
      call contorno(3)
      do tt = 1, ntt
        dt = a_dt(tt)
        call visco()
        call convdiff()
        call tempo()
        call contorno(2)
      enddo
      call mms()
      write(*,'(A,I6,A,ES24.16,A,4ES24.16)') 'mms', it, ' t=', t, ' err=', mms_err
    enddo
    write(*,'(A,ES24.16)') 'erro total acumulado=', erro_t / ts
    call check(c_suli_finalize(), 'finalize')
  end subroutine run_mms
end program principal_gpu
