! ORACLE TEST INFRASTRUCTURE -- never part of the product.
!
! Driver for the reference's own hot-path subroutines (RHS0_u1/2/3, nonlinear, solveU, meanflow_ctU,
! solveP, v_corr, meanflow_ctP from littleharsh.f90; the FFTPACK wrappers of FOU3D.f90; LU_buildP from
! tridLU_3D.f90), compiled unmodified from the reference sources.  The reference's own set-up routine
! (start.f90) cannot be compiled in this image (it needs a Fortran `mpi` module that the installed MPI
! does not provide for this compiler), so this program restates the set-up for one MPI rank:
!   ygrid        start.f90:665-806      (polynomial stretched grid, dthdy, 3-point D2 weights)
!   def_k        start.f90:809-862      (k1F = i*k, k2F = -k^2)
!   RK3 coeffs   start.f90:439-453      (Le & Moin 1991)
!   dtv          start.f90:476
!   columns      start.f90:930-1040     (global column list k outer, i inner; dk shift)
!   planes       start.f90:1079-1219    (np=1: planelim u/v = 1..nyu, p = 2..nyu)
!   weights      start.f90:518-537
!   allocations  start.f90:543-608      (all plane buffers zero-initialised)
!   LU_buildP    start.f90:636
!   nstat/iwrite are set so that the statistics, spectra and record_out branches never fire.
! The time loop follows littleharsh.f90:163-233 without the hard-coded 4-step limit.
!
! Input : params.txt (list-directed), state_in.bin (stream, complex(8) column fields, j fastest)
! Output: state_out.bin, hist.txt, and for mode 2 stages.bin; for mode 3 fft_out.bin; for mode 5
! stats_out.bin
program ref_harness
  use declaration
  implicit none
  include 'mpif.h'
  integer status(MPI_STATUS_SIZE), ierr, myid
  integer mode, ny_in, nsteps, istep, kRK_in, c, i, k, j, nplanes, ip
  real(8) dt_in, QxT_in, t_begin, t_end
  complex(8), allocatable :: u1(:,:), u2(:,:), u3(:,:), p(:,:), psi(:,:), div(:,:)
  complex(8), allocatable :: Nu1(:,:), Nu2(:,:), Nu3(:,:), du1(:,:), du2(:,:), du3(:,:)
  real(8), allocatable :: a_ugrid(:,:), a_vgrid(:,:), pl(:,:,:)
  integer :: ncol

  call MPI_INIT(ierr)
  call MPI_COMM_SIZE(MPI_COMM_WORLD, np, ierr)
  call MPI_COMM_RANK(MPI_COMM_WORLD, myid, ierr)
  if (np /= 1) then
    write(0,*) 'ref_harness runs on one rank only'
    call MPI_ABORT(MPI_COMM_WORLD, 1, ierr)
  end if

  open(41, file='params.txt', form='formatted', status='old')
  read(41,*) mode
  read(41,*) Ngal_x, Ngal_z, ny_in
  read(41,*) Re, alp, bet, CFL, mpgx
  read(41,*) dyq, ppp, flag_ctpress
  read(41,*) t, iter, nsteps, kRK_in, dt_in, QxT_in
  close(41)

  if (mode == 3) then
    call fft_mode()
    call MPI_FINALIZE(ierr)
    stop
  end if

  call setup(ny_in)
  ncol = columns_num(0)

  allocate(u1(jlim(1,ugrid):jlim(2,ugrid),ncol), u2(jlim(1,vgrid):jlim(2,vgrid),ncol))
  allocate(u3(jlim(1,ugrid):jlim(2,ugrid),ncol), p(jlim(1,pgrid):jlim(2,pgrid),ncol))
  allocate(psi(jlim(1,pgrid):jlim(2,pgrid),ncol), div(jlim(1,pgrid):jlim(2,pgrid),ncol))
  allocate(du1(jlim(1,ugrid):jlim(2,ugrid),ncol), du2(jlim(1,vgrid):jlim(2,vgrid),ncol))
  allocate(du3(jlim(1,ugrid):jlim(2,ugrid),ncol))
  allocate(Nu1(jlim(1,ugrid)+1:jlim(2,ugrid)-1,ncol), Nu2(jlim(1,vgrid)+1:jlim(2,vgrid)-1,ncol))
  allocate(Nu3(jlim(1,ugrid)+1:jlim(2,ugrid)-1,ncol))
  allocate(a_ugrid(3,jlim(1,ugrid):jlim(2,ugrid)), a_vgrid(3,jlim(1,vgrid):jlim(2,vgrid)))
  u1 = 0d0; u2 = 0d0; u3 = 0d0; p = 0d0; psi = 0d0; div = 0d0
  du1 = 0d0; du2 = 0d0; du3 = 0d0; Nu1 = 0d0; Nu2 = 0d0; Nu3 = 0d0
  a_ugrid = 0d0; a_vgrid = 0d0

  open(42, file='state_in.bin', form='unformatted', access='stream', status='old')
  read(42) u1, u2, u3, p
  if (mode == 2) read(42) Nu1, Nu2, Nu3
  close(42)

  ! getini bookkeeping (start.f90:1255-1316): flow rate of mode (0,0), divergence of the initial field
  call flowrateIm(Qx, u1(:,1))
  if (flag_ctpress /= 1) then
    QxT = Qx
    if (QxT_in /= 0d0) QxT = QxT_in
  end if
  iter0 = iter
  iwrite = huge(iwrite)
  nstat = huge(nstat)
  nextqt = huge(nextqt)
  call divergence(div, u1, u2, u3, 0)

  if (mode == 5) then
    ! statistics and spectra of this state: the reference's own nonlinear at kRK = 1 with flagst = 1
    ! (spectra FOU3D.f90:138-141, omega_x and stats FOU3D.f90:307-343); ppPL holds p in spectral planes
    ! as after record_out's modes_to_planes_UVP (FOU3D.f90:151-152), the state of a write step.
    allocate(Um(0:nyu+1), U2m(0:nyu+1), Wm(0:nyu+1), W2m(0:nyu+1), UVm(0:nyu+1), UWm(0:nyu+1))
    allocate(Vm(0:nyu+1), V2m(0:nyu+1), VWm(0:nyu+1), wxm(0:nyu+1), wx2m(0:nyu+1))
    allocate(Pm(1:nyu+1), P2m(1:nyu+1))
    Um = 0d0; U2m = 0d0; Wm = 0d0; W2m = 0d0; UVm = 0d0; UWm = 0d0
    Vm = 0d0; V2m = 0d0; VWm = 0d0; wxm = 0d0; wx2m = 0d0; Pm = 0d0; P2m = 0d0
    allocate(spU(jlim(1,ugrid):jlim(2,ugrid),ncol), spV(jlim(1,vgrid):jlim(2,vgrid),ncol))
    allocate(spW(jlim(1,ugrid):jlim(2,ugrid),ncol), spUV(jlim(1,ugrid):jlim(2,ugrid),ncol))
    allocate(spP(jlim(1,pgrid):jlim(2,pgrid),ncol))
    spU = 0d0; spV = 0d0; spW = 0d0; spUV = 0d0; spP = 0d0
    allocate(wx(Nspec_x+2,Nspec_z,jgal(vgrid,1)-1:jgal(vgrid,2)+1))
    allocate(u2PLN(Nspec_x+2,Nspec_z,jgal(vgrid,1)-1:jgal(vgrid,2)+1))
    allocate(du3dy_planes(Nspec_x+2,Nspec_z,jgal(vgrid,1)-1:jgal(vgrid,2)+1))
    allocate(du3dy_columns(jlim(1,vgrid):jlim(2,vgrid),ncol))
    wx = 0d0; u2PLN = 0d0; du3dy_planes = 0d0; du3dy_columns = 0d0
    istat = 0
    nstat = 1
    iter0 = iter - 1
    ppPL = 0d0
    call modes_to_planes_UVP(ppPL, p, 3, nyp, nyp_LB, 0, status, ierr)
    kRK = 1
    call RHS0_u1(du1, u1, Nu1, p, 0)
    call RHS0_u2(du2, u2, Nu2, p, 0)
    call RHS0_u3(du3, u3, Nu3, p, 0)
    call nonlinear(Nu1, Nu2, Nu3, u1, u2, u3, du1, du2, du3, p, div, 0, status, ierr)
    open(50, file='stats_out.bin', form='unformatted', access='stream', status='replace')
    write(50) istat
    write(50) Um, U2m, Wm, W2m, UVm, UWm, Vm, V2m, VWm, wxm, wx2m, Pm, P2m
    write(50) spU, spV, spW, spUV, spP
    close(50)
    call MPI_FINALIZE(ierr)
    stop
  end if

  if (mode == 4) then
    ! timing only: nsteps substeps (kRK = 1, 2, 3, 1, ...) of the main loop body, wall time of the loop
    t_begin = MPI_Wtime()
    do istep = 1, nsteps
      kRK = mod(istep-1, 3) + 1
      call RHS0_u1(du1, u1, Nu1, p, 0)
      call RHS0_u2(du2, u2, Nu2, p, 0)
      call RHS0_u3(du3, u3, Nu3, p, 0)
      call nonlinear(Nu1, Nu2, Nu3, u1, u2, u3, du1, du2, du3, p, div, 0, status, ierr)
      call solveU(u1, du1, ugrid, a_ugrid, 0)
      call solveU(u2, du2, vgrid, a_vgrid, 0)
      call solveU(u3, du3, ugrid, a_ugrid, 0)
      if (flag_ctpress == 0) call meanflow_ctU(u1, 0)
      call solveP(p, psi, u1, u2, u3, div, 0)
      call v_corr(u1, u2, u3, psi, div, 0, status, ierr)
    end do
    t_end = MPI_Wtime()
    open(49, file='timing.txt', form='formatted')
    write(49,'(I6,ES24.15)') nsteps, t_end - t_begin
    close(49)
    call MPI_FINALIZE(ierr)
    stop
  end if

  if (mode == 1) then
    ! dt_in > 0: a run at fixed dt.  The reference has only the CFL rule (FOU3D.f90:290-303), which sets
    ! dt, dti and t inside nonlinear at kRK = 1; the harness replaces them right after that call, before
    ! anything reads them (solveU is the first user of dt).
    open(43, file='hist.txt', form='formatted')
    do istep = 1, nsteps
      do kRK = 1, 3
        call RHS0_u1(du1, u1, Nu1, p, 0)
        call RHS0_u2(du2, u2, Nu2, p, 0)
        call RHS0_u3(du3, u3, Nu3, p, 0)
        t_begin = t
        call nonlinear(Nu1, Nu2, Nu3, u1, u2, u3, du1, du2, du3, p, div, 0, status, ierr)
        if (kRK == 1 .and. dt_in > 0d0) then
          dt = dt_in
          dti = 1d0/dt
          t = t_begin + dt
        end if
        call solveU(u1, du1, ugrid, a_ugrid, 0)
        call solveU(u2, du2, vgrid, a_vgrid, 0)
        call solveU(u3, du3, ugrid, a_ugrid, 0)
        if (flag_ctpress == 0) call meanflow_ctU(u1, 0)
        call solveP(p, psi, u1, u2, u3, div, 0)
        call v_corr(u1, u2, u3, psi, div, 0, status, ierr)
      end do
      if (flag_ctpress == 1) call meanflow_ctP(u1, 0)
      iter = iter + 1
      write(43,'(I10,7ES26.17E3)') iter, t, dt, dtc, Qx, mpgx, dgx, maxval(abs(div))
    end do
    close(43)
  else if (mode == 2) then
    kRK = kRK_in
    if (kRK /= 1) then
      dt = dt_in
      dti = 1d0/dt
    end if
    open(44, file='stages.bin', form='unformatted', access='stream', status='replace')
    call RHS0_u1(du1, u1, Nu1, p, 0)
    call RHS0_u2(du2, u2, Nu2, p, 0)
    call RHS0_u3(du3, u3, Nu3, p, 0)
    write(44) du1(1:nyu,:), du2(1:nyv,:), du3(1:nyu,:)
    call nonlinear(Nu1, Nu2, Nu3, u1, u2, u3, du1, du2, du3, p, div, 0, status, ierr)
    write(44) Nu1, Nu2, Nu3
    write(44) du1(1:nyu,:), du2(1:nyv,:), du3(1:nyu,:)
    write(44) uv_f, vv_c, wv_f
    call solveU(u1, du1, ugrid, a_ugrid, 0)
    call solveU(u2, du2, vgrid, a_vgrid, 0)
    call solveU(u3, du3, ugrid, a_ugrid, 0)
    write(44) u1, u2, u3
    if (flag_ctpress == 0) call meanflow_ctU(u1, 0)
    write(44) u1
    call solveP(p, psi, u1, u2, u3, div, 0)
    write(44) p, psi
    call v_corr(u1, u2, u3, psi, div, 0, status, ierr)
    write(44) u1, u2, u3, div
    write(44) dt, t, mpgx, dgx, dtc
    close(44)
  end if

  open(45, file='state_out.bin', form='unformatted', access='stream', status='replace')
  write(45) u1, u2, u3, p, Nu1, Nu2, Nu3
  write(45) t, dt, mpgx, Qx, dgx
  close(45)

  ! grid and constant set-up, for checking the restated set-up against the fixtures
  open(46, file='setup_out.bin', form='unformatted', access='stream', status='replace')
  write(46) yu, yv, dthdyu, dthdyv, dyu2i, dyv2i, DG, dtv, gridweighting, gridweighting_interp, &
            gridweighting_bc_u1, gridweighting_bc_u3
  close(46)

  call MPI_FINALIZE(ierr)

contains

  subroutine setup(ny)
    integer, intent(in) :: ny
    integer col, g
    real(8) aaa, bbb, qqq

    pnodes = 2
    nyv = ny - 2
    nyu = nyv + 1
    nyp = nyv
    nyu_LB = 0
    nyv_LB = 0
    nyp_LB = 1
    Nspec_x = (2*Ngal_x)/3
    Nspec_z = (2*Ngal_z)/3
    Lx = 2d0*pi/alp
    Lz = 2d0*pi/bet
    Ly = 2d0
    nwrite = 1

    ! --- stretched y grid (restates start.f90:665-806)
    nn = nyv + 1
    qqq = nn/2d0
    bbb = (dyq-1d0)/(qqq*(1d0-qqq**(ppp-1)+dyq*((1d0-qqq)*qqq**(ppp-1)-(1d0-qqq)**ppp)))
    aaa = (1d0-bbb*qqq**ppp)/qqq
    dtheta = 1d0
    dthetai = 1d0/dtheta
    ddthetavi = 1.0d0/dtheta
    allocate(yu(0:nn+1), dyu2i(3,0:nn+1), dthdyu(0:nn+1))
    allocate(yv(0:nn), dyv2i(3,0:nn), dthdyv(0:nn))
    yu(0) = aaa*(-1d0+.5d0-qqq)+bbb*(-1d0+.5d0-qqq)**ppp
    do j = 0, nn
      yv(j) = aaa*(j-qqq)+bbb*(j-qqq)**ppp
      yu(j+1) = aaa*(j+.5d0-qqq)+bbb*(j+.5d0-qqq)**ppp
    end do
    dthdyu(0) = 1d0/(aaa+ppp*bbb*(-1d0+.5d0-qqq)**(ppp-1))
    do j = 0, nn
      dthdyv(j) = 1d0/(aaa+ppp*bbb*(j-qqq)**(ppp-1))
      dthdyu(j+1) = 1d0/(aaa+ppp*bbb*(j+.5d0-qqq)**(ppp-1))
    end do
    j = 0
    dyu2i(1,j) = 0d0
    dyu2i(2,j) = -2d0/(yu(j+1)-yu(j))**2
    dyu2i(3,j) = 1d0/(yu(j+1)-yu(j))**2
    dyv2i(1,j) = 0d0
    dyv2i(2,j) = -2d0/(yv(j+1)-yv(j))**2
    dyv2i(3,j) = 1d0/(yv(j+1)-yv(j))**2
    do j = 1, nn-1
      dyu2i(1,j) = 2d0/((yu(j)-yu(j-1))*(yu(j+1)-yu(j-1)))
      dyu2i(2,j) = -2d0/((yu(j+1)-yu(j))*(yu(j)-yu(j-1)))
      dyu2i(3,j) = 2d0/((yu(j+1)-yu(j))*(yu(j+1)-yu(j-1)))
      dyv2i(1,j) = 2d0/((yv(j)-yv(j-1))*(yv(j+1)-yv(j-1)))
      dyv2i(2,j) = -2d0/((yv(j+1)-yv(j))*(yv(j)-yv(j-1)))
      dyv2i(3,j) = 2d0/((yv(j+1)-yv(j))*(yv(j+1)-yv(j-1)))
    end do
    j = nn
    dyu2i(1,j) = 2d0/((yu(j)-yu(j-1))*(yu(j+1)-yu(j-1)))
    dyu2i(2,j) = -2d0/((yu(j+1)-yu(j))*(yu(j)-yu(j-1)))
    dyu2i(3,j) = 2d0/((yu(j+1)-yu(j))*(yu(j+1)-yu(j-1)))
    dyv2i(1,j) = 1d0/(yv(j)-yv(j-1))**2
    dyv2i(2,j) = -2d0/(yv(j)-yv(j-1))**2
    dyv2i(3,j) = 0d0
    j = nn+1
    dyu2i(1,j) = 1d0/(yu(j)-yu(j-1))**2
    dyu2i(2,j) = -2d0/(yu(j)-yu(j-1))**2
    dyu2i(3,j) = 0d0
    nn = nyv
    allocate(u11(0:nn+2))
    u11 = 0d0

    nstat = min(20, nwrite)

    ! --- FFT work arrays (start.f90:414-432)
    allocate(buffR_x(2*Nspec_x+18), buffC_z(10*Nspec_z+19), buffRal_x(2*Ngal_x+18), buffCal_z(10*Ngal_z+19))
    buffR_x = 0d0; buffC_z = 0d0; buffRal_x = 0d0; buffCal_z = 0d0
    call rfti(Nspec_x, buffR_x)
    call cfti(Nspec_z, buffC_z)
    call rfti(Ngal_x, buffRal_x)
    call cfti(Ngal_z, buffCal_z)

    ! --- wavenumbers (restates start.f90:809-862)
    allocate(k1F_x(0:Nspec_x/2), k2F_x(0:Nspec_x/2), k1F_z(1:Nspec_z), k2F_z(1:Nspec_z))
    do i = 0, Nspec_x/2
      k1F_x(i) = im*alp*i
      k2F_x(i) = -(alp*i)**2
    end do
    do k = 1, Nspec_z/2
      k1F_z(k) = im*bet*(k-1)
      k2F_z(k) = -(bet*(k-1))**2
    end do
    do k = Nspec_z/2+1, Nspec_z
      k1F_z(k) = -im*bet*(Nspec_z-k+1)
      k2F_z(k) = -(bet*(Nspec_z-k+1))**2
    end do

    ! --- RK3 coefficients (start.f90:439-453)
    aRK = (/ 4d0/15d0, 1d0/15d0, 1d0/6d0 /)
    bRK = (/ 4d0/15d0, 1d0/15d0, 1d0/6d0 /)
    gRK = (/ 8d0/15d0, 5d0/12d0, 3d0/4d0 /)
    cRK = (/ 0d0, -17d0/60d0, -5d0/12d0 /)
    dRK = (/ 8d0/15d0, 2d0/15d0, 1d0/3d0 /)
    dtv = Re/(-k2F_x(Nspec_x/2)-k2F_z(Nspec_z/2)+4d0/(yv(nyv_LB+1)-yv(nyv_LB))**2)

    ! --- column list for one rank (start.f90:930-1040)
    allocate(columns_num(0:0))
    columns_num(0) = (Nspec_x/2+1)*Nspec_z
    allocate(columns_i(columns_num(0),0:0), columns_k(columns_num(0),0:0))
    col = 0
    do k = 1, Nspec_z/2
      do i = 0, Nspec_x/2
        col = col + 1
        columns_i(col,0) = i
        columns_k(col,0) = k
      end do
    end do
    do k = Nspec_z-Nspec_z/2+1, Nspec_z
      do i = 0, Nspec_x/2
        col = col + 1
        columns_i(col,0) = i
        columns_k(col,0) = k
      end do
    end do
    allocate(jlim(2,3))
    jlim(1,ugrid) = nyu_LB;   jlim(2,ugrid) = nyu+1
    jlim(1,vgrid) = nyv_LB;   jlim(2,vgrid) = nyv+1
    jlim(1,pgrid) = nyu_LB+1; jlim(2,pgrid) = nyu
    allocate(dk(columns_num(0),0:0), dk_phys(columns_num(0),0:0))
    dk = 0
    dk_phys = 0
    do col = 1, columns_num(0)
      if (columns_k(col,0) > Nspec_z/2) dk(col,0) = Nspec_z-Ngal_z
    end do

    ! --- plane limits for one rank (start.f90:1079-1219): the iproc==0 branch applies, not the np-1 one
    allocate(planelim(3,2,0:0), limPL_incw(3,2,0:0), limPL_excw(3,2,0:0), limPL_FFT(3,2,0:0))
    do g = 1, 3
      planelim(g,1,0) = 1
      planelim(g,2,0) = nyu
    end do
    planelim(pgrid,1,0) = 2
    limPL_excw = planelim
    limPL_incw = planelim
    limPL_incw(:,1,0) = planelim(:,1,0) - 1
    limPL_incw(:,2,0) = planelim(:,2,0) + 1
    igal = Ngal_x+2
    kgal = Ngal_z
    do g = 1, 3
      jgal(g,1) = limPL_excw(g,1,0)
      jgal(g,2) = limPL_excw(g,2,0)
    end do

    ! --- wall weights (start.f90:518-537)
    allocate(gridweighting(2), gridweighting_interp(2))
    gridweighting(1) = -(yv(nyv_LB)-yu(nyu_LB))/(yv(nyv_LB)-yu(nyu_LB+1))
    gridweighting(2) = (yu(nyu+1)-yv(nyv+1))/(yv(nyv+1)-yu(nyu))
    gridweighting_bc_u1 = (yu(0)-yv(0))/(yu(1)-yv(0))
    gridweighting_bc_u3 = (yu(0)-yv(0))/(yu(1)-yv(0))
    gridweighting_interp(1) = (yu(nyu_LB)-yu(nyu_LB+1))/(yv(nyv_LB)-yu(nyu_LB+1))
    gridweighting_interp(2) = (yu(nyu+1)-yu(nyu))/(yv(nyv+1)-yu(nyu))

    ! --- plane buffers (start.f90:543-608); zero-initialised
    allocate(u1PL(igal,kgal,jgal(ugrid,1)-1:jgal(ugrid,2)+1)); u1PL = 0d0
    allocate(u2PL(igal,kgal,jgal(vgrid,1)-1:jgal(vgrid,2)+1)); u2PL = 0d0
    allocate(u3PL(igal,kgal,jgal(ugrid,1)-1:jgal(ugrid,2)+1)); u3PL = 0d0
    allocate(u1PL_itp(igal,kgal,jgal(vgrid,1)-1:jgal(vgrid,2)+1)); u1PL_itp = 0d0
    allocate(u2PL_itp(igal,kgal,jgal(ugrid,1)-1:jgal(ugrid,2)+1)); u2PL_itp = 0d0
    allocate(u3PL_itp(igal,kgal,jgal(vgrid,1)-1:jgal(vgrid,2)+1)); u3PL_itp = 0d0
    allocate(ppPL(igal,kgal,jgal(pgrid,1)-1:jgal(pgrid,2)+1)); ppPL = 0d0
    allocate(Nu1PL(igal,kgal,jgal(ugrid,1):jgal(ugrid,2))); Nu1PL = 0d0
    allocate(Nu2PL(igal,kgal,jgal(vgrid,1):jgal(vgrid,2))); Nu2PL = 0d0
    allocate(Nu3PL(igal,kgal,jgal(ugrid,1):jgal(ugrid,2))); Nu3PL = 0d0
    allocate(uu_cPL(igal,kgal,jgal(ugrid,1)-1:jgal(ugrid,2)+1)); uu_cPL = 0d0
    allocate(uv_fPL(igal,kgal,jgal(vgrid,1)-1:jgal(vgrid,2)+1)); uv_fPL = 0d0
    allocate(uw_cPL(igal,kgal,jgal(ugrid,1)-1:jgal(ugrid,2)+1)); uw_cPL = 0d0
    allocate(vv_cPL(igal,kgal,jgal(ugrid,1)-1:jgal(ugrid,2)+1)); vv_cPL = 0d0
    allocate(wv_fPL(igal,kgal,jgal(vgrid,1)-1:jgal(vgrid,2)+1)); wv_fPL = 0d0
    allocate(ww_cPL(igal,kgal,jgal(ugrid,1)-1:jgal(ugrid,2)+1)); ww_cPL = 0d0
    allocate(u1_itp(jlim(1,vgrid):jlim(2,vgrid),columns_num(0))); u1_itp = 0d0
    allocate(u2_itp(jlim(1,ugrid):jlim(2,ugrid),columns_num(0))); u2_itp = 0d0
    allocate(u3_itp(jlim(1,vgrid):jlim(2,vgrid),columns_num(0))); u3_itp = 0d0
    allocate(Nu1_dy(jlim(1,ugrid)+1:jlim(2,ugrid)-1,columns_num(0))); Nu1_dy = 0d0
    allocate(Nu2_dy(jlim(1,vgrid)+1:jlim(2,vgrid)-1,columns_num(0))); Nu2_dy = 0d0
    allocate(Nu3_dy(jlim(1,ugrid)+1:jlim(2,ugrid)-1,columns_num(0))); Nu3_dy = 0d0
    allocate(uv_f(jlim(1,vgrid):jlim(2,vgrid),columns_num(0))); uv_f = 0d0
    allocate(vv_c(jlim(1,ugrid):jlim(2,ugrid),columns_num(0))); vv_c = 0d0
    allocate(wv_f(jlim(1,vgrid):jlim(2,vgrid),columns_num(0))); wv_f = 0d0
    allocate(DG(3,jlim(1,pgrid):jlim(2,pgrid),columns_num(0))); DG = 0d0

    call LU_buildP(jlim(1,pgrid), jlim(2,pgrid), 0, DG)
  end subroutine

  subroutine fft_mode()
    ! Applies four_to_phys_du (FOU3D.f90:681-698) then phys_to_four_du (FOU3D.f90:1654-1672) to each plane.
    integer ip2
    Nspec_x = (2*Ngal_x)/3
    Nspec_z = (2*Ngal_z)/3
    allocate(buffRal_x(2*Ngal_x+18), buffCal_z(10*Ngal_z+19))
    buffRal_x = 0d0; buffCal_z = 0d0
    call rfti(Ngal_x, buffRal_x)
    call cfti(Ngal_z, buffCal_z)
    nplanes = nsteps
    allocate(pl(Ngal_x+2, Ngal_z, 2*nplanes))
    open(47, file='fft_in.bin', form='unformatted', access='stream', status='old')
    read(47) pl(:,:,1:nplanes)
    close(47)
    do ip2 = 1, nplanes
      call four_to_phys_du(pl(1,1,ip2))
      pl(:,:,nplanes+ip2) = pl(:,:,ip2)
      call phys_to_four_du(pl(1,1,nplanes+ip2))
    end do
    open(48, file='fft_out.bin', form='unformatted', access='stream', status='replace')
    write(48) pl
    close(48)
  end subroutine

end program
