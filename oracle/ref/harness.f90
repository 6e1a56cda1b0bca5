! Test harness around the reference SuLi-IC routines (oracle only).
!
! Replaces the reference PROGRAM (1_principal.f90) for golden-vector and
! CPU-baseline runs: it loads a full solver state from a raw little-endian
! stream file, calls one reference subroutine (or the time step of
! 1_principal.f90:64-108), and writes the state back in the same format.
!
!   harness IN OUT CMD [ARG]
!
! CMD: ls_ini | level_set | heaviside | mod_ls1 | reinic | der_weno | weno1 ARG(ihs)
!      visco | convdiff | tempo ARG(tt) | contorno ARG(nlock) | graddin | posdin
!      stage ARG(tt) | step ARG(nsteps) | bench | plot | visu | restart_salva | restart_ini
!      obstaculo | est | mms_src | mms_run ARG(nsteps) | waves | wave_run ARG(nsteps)
! (plot and visu write under arquivos/, the restart commands use dados/, of
! the working directory)
!
! The mms and wave commands (builds with mms_t > 0, 13_mms.f90, or
! wave_t > 0, 3_contorno.f90:952-1205) also write OUT.mms:
! tf_u, tf_v, tf_w, tf_px, tf_py, tf_pz, the 24 boundary planes in the order
! of their declaration (0_modules.f90:96-120), erro_u, erro_v, erro_w,
! erro_p, each in Fortran order with its module extents.
!
! State file: int32 nx,ny,nz; 10 float64 scalars (dt, t, vol_ini, vol_ins,
! rho_f1, rho_f2, mi_f1, mi_f2, sigma, alpha1); then the arrays listed in
! io_state, each in Fortran order with the reference extents (0_modules.f90).
program harness
  use disc
  use velpre
  use vartempo
  use smag
  use ls_param
  use obst
  use tempo_mod
  use mms_m
  use wave_c
  implicit none
  character(len=512) :: fin, fout, cmd, arg
  integer :: iarg, nsteps, k, i, j
  integer(8) :: c0, c1, crate, g0, g1, gsum, ls0, ls1, lssum
  real(8) :: dt_base
  real(8), dimension(nx,ny,nz) :: ta1, tb1, tc1, td1, te1, tf1
  real(8), dimension(nx,ny,nz) :: s1, s2, s3
  integer :: ihs

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  call get_command_argument(3, cmd)
  iarg = 0
  if (command_argument_count() >= 4) then
    call get_command_argument(4, arg)
    read(arg, *) iarg
  endif

  call io_state(fin, 0)

  ! RK3-TVD coefficients and the stage steps of the time scheme t_tempo as
  ! set up by 14_level_set.f90:34-44 and 2_iniciais.f90:85-102.
  adtl(1) = 1.;     bdtl(1) = 0.;     gdtl(1) = 1.
  adtl(2) = 3./4.;  bdtl(2) = 1./4.;  gdtl(2) = 1./4.
  adtl(3) = 1./3.;  bdtl(3) = 2./3.;  gdtl(3) = 2./3.
  ls_m = 0.
  dt_base = dt
  if ((t_tempo == 0) .or. (t_tempo == 3)) then
    ntt = 1
    a_dt = 1.0*dt_base
  else
    ntt = 2
    if (t_tempo == 2) ntt = 3
    a_dt(1) = 0.5*dt_base
    a_dt(2) = 1.0*dt_base
    a_dt(3) = 1.0*dt_base
  endif
  tt = 1
  it = 1

  select case (trim(cmd))
  case ('ls_ini')
    call heaviside()
    vol_ini = 0.
    do k = 1, nz
    do j = 1, ny
    do i = 1, nx
      vol_ini = vol_ini + (1.-(1.-hs(i,j,k)))*dx*dy*dz
    enddo
    enddo
    enddo
    call mod_ls1()
  case ('level_set')
    call level_set()
  case ('heaviside')
    call heaviside()
  case ('mod_ls1')
    call mod_ls1()
  case ('reinic')
    call reinic_weno(s1, s2, s3)
  case ('der_weno')
    call der_weno(ls, ta1, tb1, tc1, td1, te1, tf1, ihs)
    hsx = ta1; hsy = tb1; hsz = tc1
    ddlsdx = td1; ddlsdy = te1; ddlsdz = tf1
    vol_ins = ihs
  case ('weno1')
    ihs = iarg
    call weno1(ddlsdx(:,1,1), ddlsdy(:,1,1), nx, dx, ls(:,1,1), ihs)
  case ('visco')
    call visco()
  case ('convdiff')
    call convdiff()
  case ('tempo')
    tt = iarg
    call tempo()
  case ('contorno')
    call contorno(iarg)
  case ('graddin')
    call graddin()
  case ('posdin')
    call posdin()
  case ('stage')
    tt = iarg
    dt = a_dt(tt)
    call visco()
    call convdiff()
    call tempo()
    call contorno(2)
    call graddin()
    call posdin()
    call contorno(1)
    dt = dt_base
  case ('step')
    nsteps = max(iarg, 1)
    gsum = 0
    lssum = 0
    call system_clock(c0, crate)
    do it = 1, nsteps
      t = it * dt
      call system_clock(ls0)
      call level_set()
      call system_clock(ls1)
      lssum = lssum + (ls1 - ls0)
      call contorno(3)
      do tt = 1, ntt
        dt = a_dt(tt)
        call visco()
        call convdiff()
        call tempo()
        call contorno(2)
        call system_clock(g0)
        call graddin()
        call system_clock(g1)
        gsum = gsum + (g1 - g0)
        call posdin()
        call contorno(1)
      enddo
    enddo
    call system_clock(c1)
    write(*,'(A,I0,A,ES16.8,A,ES16.8,A,ES16.8)') 'REF_TIMING steps=', nsteps, &
      ' seconds=', real(c1-c0,8)/real(crate,8), ' graddin_seconds=', real(gsum,8)/real(crate,8), &
      ' level_set_seconds=', real(lssum,8)/real(crate,8)
  case ('bench')
    ! CPU-baseline sample: rebuild the level-set fields from ls
    ! (14_level_set.f90:205,217), set the ghost and friction planes
    ! (1_principal.f90:46), then time one level_set and one RK stage.
    call heaviside()
    call mod_ls1()
    call contorno(1)
    call system_clock(c0, crate)
    call level_set()
    call contorno(3)
    call system_clock(ls1)
    write(*,'(A,ES16.8)') 'REF_BENCH level_set_seconds=', real(ls1-c0,8)/real(crate,8)
    flush(6)
    tt = 1
    dt = a_dt(tt)
    call visco()
    call convdiff()
    call tempo()
    call contorno(2)
    call system_clock(g0)
    call graddin()
    call system_clock(g1)
    call posdin()
    call contorno(1)
    call system_clock(c1)
    dt = dt_base
    write(*,'(A,ES16.8,A,ES16.8)') 'REF_BENCH stage1_seconds=', real(c1-ls1,8)/real(crate,8), &
      ' graddin_seconds=', real(g1-g0,8)/real(crate,8)
    flush(6)
  case ('plot')
    ! one frame of plot_f (4_plot.f90:260-363): the file counter starts at 10
    ! as plot_i leaves it (4_plot.f90:131) and `it` is a frame step, so the
    ! record goes to arquivos/campos_00011 of the working directory
    cont = 10
    it = ceiling(dt_frame/dt)
    agora = 0
    agora1 = 0
    prev = 0.
    call plot_f()
  case ('visu')
    ! the ParaView index arquivos/visu.xdmf (12_paraview_3D.f90:1-179)
    call visu()
  case ('restart_salva')
    ! dados/arquivorestart of the working directory (11_diversos.f90:303-329)
    call restart_salva()
  case ('est')
    ! est (4_plot.f90:500-525): the divergence field goes to kurv and its
    ! largest magnitude to the vol_ins slot of the state file (est itself
    ! logs the value to unit 100002 of the working directory)
    call est(s1)
    kurv = s1
    vol_ins = maxval(abs(s1))
  case ('obstaculo')
    ! the obstacle columns ku, kv, kw of obst_t (3_contorno.f90:494-760)
    call obstaculo()
  case ('mms_src')
    ! the boundary planes (mms_bc, 13_mms.f90:468-816; for mms_t = 1 only at
    ! t = 0) and the momentum sources of mms_t (termo_fonte1 / termo_fonte2,
    ! :819-1322) at the time t of the state file
    call mms_bc()
    if (mms_t == 1) then
      call termo_fonte1()
    else
      call termo_fonte2()
    endif
    call io_mms(trim(fout)//'.mms')
  case ('mms_run')
    ! the manufactured-solution run of 1_principal.f90:39-108 with mms_t > 0
    ! from t = 0: the initial fields of mms_i (level_set_ini with tipo = 6,
    ! 14_level_set.f90:178-218), contorno(1), contorno(3), mms(); then per
    ! step the sources, level_set, contorno(3), the stages without the
    ! projection (visco, convdiff, tempo, contorno(2)) and mms().  The errors
    ! of mms() are printed per step (MMS_ERR it t erro_u1 erro_v1 erro_w1
    ! erro_p1, recomputed here from the module arrays as mms() does).
    nsteps = max(iarg, 0)
    t = 0.
    it = 0
    call mms_i()
    call heaviside()
    vol_ini = 0.
    do k = 1, nz
    do j = 1, ny
    do i = 1, nx
      vol_ini = vol_ini + (1.-(1.-hs(i,j,k)))*dx*dy*dz
    enddo
    enddo
    enddo
    call mod_ls1()
    call contorno(1)
    call contorno(3)
    call mms()
    call mms_print()
    do it = 1, nsteps
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
      call mms_print()
    enddo
    call io_mms(trim(fout)//'.mms')
  case ('waves')
    ! waves_coef (3_contorno.f90:1070-1205; 2_iniciais.f90:52-54) and one
    ! boundary_waves (:952-1068) at the time t of the state file: the level
    ! set (rounded through real(4), its first column from the wave) and the
    ! inflow planes bxx0, bxx1, bxz0 (bxy0 takes v1, which the reference
    ! never sets); the wave coefficients are printed
    call waves_coef()
    call boundary_waves()
    write(*,'(A,14ES24.16)') 'WAVE_COEF ', l_w, n_w, f_w, c_w, aeta1, aeta2, aeta3, aeta4, aeta5, &
      avel1, avel2, avel3, avel4, avel5
    call io_mms(trim(fout)//'.mms')
  case ('wave_run')
    ! the time loop of 1_principal.f90:64-95 with wave_t > 0: boundary_waves
    ! after tempo in every stage (1_principal.f90:84)
    nsteps = max(iarg, 1)
    call waves_coef()
    do it = 1, nsteps
      t = it * dt
      call level_set()
      call contorno(3)
      do tt = 1, ntt
        dt = a_dt(tt)
        call visco()
        call convdiff()
        call tempo()
        call boundary_waves()
        call contorno(2)
        call graddin()
        call posdin()
        call contorno(1)
      enddo
    enddo
    dt = dt_base
    call io_mms(trim(fout)//'.mms')
  case ('restart_ini')
    ! reads dados/arquivorestart (11_diversos.f90:232-301)
    call restart_ini()
  case default
    write(*,*) 'unknown command ', trim(cmd)
    stop 2
  end select

  if (trim(fout) /= '-') call io_state(fout, 1)

contains

  subroutine mms_print()
    ! the four errors of mms() (13_mms.f90:441-444) from its module arrays
    real(8) :: e(4)
    integer :: nxc, nyc, nzc, npc
    real(8), dimension(nx1,ny,nz) :: lsx
    real(8), dimension(nx,ny1,nz) :: lsy
    real(8), dimension(nx,ny,nz1) :: lsz
    call interpx_cf(ls,nx,ny,nz,lsx)
    call interpy_cf(ls,nx,ny,nz,lsy)
    call interpz_cf(ls,nx,ny,nz,lsz)
    nxc = count(lsx > 0.); nyc = count(lsy > 0.); nzc = count(lsz > 0.); npc = count(ls > 0.)
    e(1) = sqrt(sum(erro_u)/nxc); e(2) = sqrt(sum(erro_v)/nyc)
    e(3) = sqrt(sum(erro_w)/nzc); e(4) = sqrt(sum(erro_p)/npc)
    write(*,'(A,I0,5ES24.16)') 'MMS_ERR ', it, t, e
  end subroutine mms_print

  subroutine io_mms(fname)
    character(len=*), intent(in) :: fname
    open(unit=23, file=fname, access='stream', form='unformatted', status='replace')
    write(23) tf_u, tf_v, tf_w, tf_px, tf_py, tf_pz
    write(23) bxx0, bxx1, bxy0, bxz0, bxxf, bxxf1, bxyf, bxzf, &
              byx0, byy0, byy1, byz0, byxf, byyf, byyf1, byzf, &
              bzx0, bzy0, bzz0, bzz1, bzxf, bzyf, bzzf, bzzf1
    write(23) erro_u, erro_v, erro_w, erro_p
    close(23)
  end subroutine io_mms

  subroutine io_state(fname, mode)
    character(len=*), intent(in) :: fname
    integer, intent(in) :: mode
    real(8) :: dku(0:nx1+1,0:ny+1), dkv(0:nx+1,0:ny1+1), dkw(0:nx+1,0:ny+1)  ! ku, kv, kw as float64
    integer :: hdr(3)
    real(8) :: sc(10)
    if (mode == 0) then
      open(unit=21, file=trim(fname), access='stream', form='unformatted', status='old')
      read(21) hdr
      if (hdr(1) /= nx .or. hdr(2) /= ny .or. hdr(3) /= nz) then
        write(*,*) 'grid mismatch: file ', hdr, ' build ', nx, ny, nz
        stop 3
      endif
      read(21) sc
      dt = sc(1); t = sc(2); vol_ini = sc(3); vol_ins = sc(4)
      rho_f1 = sc(5); rho_f2 = sc(6); mi_f1 = sc(7); mi_f2 = sc(8)
      sigma = sc(9); alpha1 = sc(10)
      read(21) u, v, w, prd1, rho, ls_nu, ls, mod_ls, kurv, hs, hsx, hsy, hsz, &
               ddlsdx, ddlsdy, ddlsdz, nut, xnut, ynut, znut, Fu, Fv, Fw, u0, v0, w0, ub, vb, wb, &
               fu0, fv0, fw0, fu1, fv1, fw1, dku, dkv, dkw
      ku = nint(dku); kv = nint(dkv); kw = nint(dkw)
      close(21)
    else
      open(unit=22, file=trim(fname), access='stream', form='unformatted', status='replace')
      hdr = (/ nx, ny, nz /)
      sc = (/ dt, t, vol_ini, vol_ins, rho_f1, rho_f2, mi_f1, mi_f2, sigma, alpha1 /)
      write(22) hdr
      write(22) sc
      dku = ku; dkv = kv; dkw = kw
      write(22) u, v, w, prd1, rho, ls_nu, ls, mod_ls, kurv, hs, hsx, hsy, hsz, &
                ddlsdx, ddlsdy, ddlsdz, nut, xnut, ynut, znut, Fu, Fv, Fw, u0, v0, w0, ub, vb, wb, &
                fu0, fv0, fw0, fu1, fv1, fw1, dku, dkv, dkw
      close(22)
    endif
  end subroutine io_state

end program harness
