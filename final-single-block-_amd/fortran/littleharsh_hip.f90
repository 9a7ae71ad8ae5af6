! Drop-in Fortran driver: the reference's main loop (littleharsh.f90:163-233) calling subroutines with the
! reference's names and argument lists, each of which forwards to the HIP implementation through the
! ISO_C_BINDING interface (dns_hip_mod.f90; wrappers in lh_wrappers.f90).  Fields live in GPU memory between
! calls; the host arrays in the argument lists keep the reference's shapes and are synchronised for output
! (and after every wrapper call with DNS_HIP_MIRROR=1).
!
!   littleharsh_hip [nsteps] [fused]
!
! Reads input.in in the reference's positional format (start.f90:86-252) and the initial field files
! ``<dirin>/{u1,u2,u3,p}<fnmeb minus its 2-char prefix>`` (start.f90:1545-1997), runs until t >= maxt or
! nsteps steps.  As record_out (FOU3D.f90:2003-2200) it writes the fields to
! ``<dirout>/{u1,u2,u3,p}_NNNNxNNNNxNNNN_tTTTTT.dat`` and appends a record to the history file
! ``<dirout>/hist_NNNNxNNNNxNNNN.dat`` at the start of every nwrite-th step (iwrite, start.f90:1312 and
! littleharsh.f90:433), and once more at the end of the run.  Every nstat = min(20, nwrite) steps
! (start.f90:410, FOU3D.f90:77-81) the plane statistics and spectra are accumulated on the device; at each
! write step they go to ``<dirout>/stats_NNNNxNNNNxNNNN_TTTTT.dat`` (stats.f90:244-413) and
! ``<dirout>/sp{U,V,W,UV,P}_NNNNxNNNNxNNNN_TTTTT.dat`` (spectra.f90:71-188) and are reset.

! ---------------------------------------------------------------------------------------------------
program littleharsh_hip
  use lh_state
  implicit none
  complex(8), allocatable :: u1(:,:), u2(:,:), u3(:,:), p(:,:), psi(:,:), div(:,:)
  complex(8), allocatable :: Nu1(:,:), Nu2(:,:), Nu3(:,:), du1(:,:), du2(:,:), du3(:,:)
  real(8), allocatable :: a_ugrid(:,:), a_vgrid(:,:)
  type(dns_params) :: prm
  integer :: myid = 0, status(8), ierr, nmax, nargs, iter_start
  logical :: fused
  character(len=32) :: arg
  real(8) :: sc(10)

  nmax = huge(nmax)
  fused = .false.
  nargs = command_argument_count()
  if (nargs >= 1) then
    call get_command_argument(1, arg)
    read(arg, *) nmax
  end if
  if (nargs >= 2) then
    call get_command_argument(2, arg)
    fused = trim(arg) == 'fused'
  end if

  call lh_mirror_init()
  call read_input('input.in')
  prm%Mx = Ngal_x; prm%Mz = Ngal_z; prm%ny = ny; prm%ppp = ppp
  prm%flag_ctpress = flag_ctpress; prm%precision = DNS_PREC_FP64
  prm%Re = Re; prm%alp = alp; prm%bet = bet; prm%CFL = CFL; prm%mpgx = mpgx; prm%dyq = dyq
  prm%fixed_dt = 0d0; prm%plane_batch = 0; prm%members = 1; prm%reserved = 0
  call dns_hip_check(dns_create(prm, hdns), 'start')

  allocate(u1(0:nyu+1, ncol), u2(0:nyv+1, ncol), u3(0:nyu+1, ncol), p(1:nyu, ncol))
  allocate(psi(1:nyu, ncol), div(1:nyu, ncol))
  allocate(du1(0:nyu+1, ncol), du2(0:nyv+1, ncol), du3(0:nyu+1, ncol))
  allocate(Nu1(1:nyu, ncol), Nu2(1:nyv, ncol), Nu3(1:nyu, ncol))
  allocate(a_ugrid(3, 0:nyu+1), a_vgrid(3, 0:nyv+1))
  u1 = 0d0; u2 = 0d0; u3 = 0d0; p = 0d0; psi = 0d0; div = 0d0

  call read_fields(u1, u2, u3, p)
  if (flag_init == 1) then   ! start.f90:1251-1267: a borrowed field restarts the clock
    t = 0d0
    iter = 0
  end if
  call dns_hip_check(dns_upload_fields(hdns, u1, u2, u3, p, t, int(iter, c_int), 0d0), 'getini')
  call dns_hip_check(dns_refresh_stats_pressure(hdns), 'getini')   ! ppPL from the field file (start.f90:1387)
  iter_start = iter
  iwrite = iter
  iter0 = iter
  nstat = min(20, nwrite)
  write(*,'(A,I0,A,I0,A,I0,A,ES12.5)') ' littleharsh_hip: ', Nspec_x, ' x ', Nspec_z, ' x ', ny, '  t0 = ', t
  call open_history()

  if (fused) then
    do while (t < maxt .and. iter - iter_start < nmax)
      call step_output(u1, u2, u3, p)
      call dns_hip_check(dns_rk3_step(hdns, 1_c_int), 'rk3_step')
      iter = iter + 1
      call dns_hip_check(dns_get_scalars(hdns, sc), 'scalars')
      t = sc(1)
    end do
  else
    ! main loop, littleharsh.f90:163-233
    do while (t < maxt .and. iter - iter_start < nmax)
      call step_output(u1, u2, u3, p)
      do kRK = 1, 3
        call RHS0_u1(du1, u1, Nu1, p, myid)
        call RHS0_u2(du2, u2, Nu2, p, myid)
        call RHS0_u3(du3, u3, Nu3, p, myid)
        call nonlinear(Nu1, Nu2, Nu3, u1, u2, u3, du1, du2, du3, p, div, myid, status, ierr)
        call solveU(u1, du1, 2, a_ugrid, myid)
        call solveU(u2, du2, 1, a_vgrid, myid)
        call solveU(u3, du3, 2, a_ugrid, myid)
        if (flag_ctpress == 0) call meanflow_ctU(u1, myid)
        call solveP(p, psi, u1, u2, u3, div, myid)
        call v_corr(u1, u2, u3, psi, div, myid, status, ierr)
      end do
      if (flag_ctpress == 1) call meanflow_ctP(u1, myid)
      iter = iter + 1
      call dns_hip_check(dns_get_scalars(hdns, sc), 'scalars')
      t = sc(1)
    end do
  end if

  call dns_hip_check(dns_get_scalars(hdns, sc), 'scalars')
  write(*,'(A,I0,A,ES24.16,A,ES24.16,A,ES24.16)') ' iter ', iter, '  t = ', sc(1), '  dt = ', sc(2), &
    '  Qx = ', sc(5)
  iwrite = iter
  call dns_hip_check(dns_refresh_stats_pressure(hdns), 'finalize')   ! littleharsh.f90:439-443
  call record_out_hip(u1, u2, u3, p)
  close(30)
  call dns_hip_check(dns_destroy(hdns), 'finalize')

contains

  ! input.in, positional with the reference's FORMATs (start.f90:86-252, 485-488)
  subroutine read_input(fname)
    character(*), intent(in) :: fname
    integer :: i
10  format(7X, D10.1)
20  format(7X, I10)
21  format(12X, I5)
30  format(8X, A50)
    open(40, file=fname, form='formatted', status='old')
    do i = 1, 9
      read(40, 10)
    end do
    read(40, 10) Re
    read(40, 10)
    read(40, 10)
    read(40, 10) alp
    read(40, 10) bet
    read(40, 20) Ngal_x
    read(40, 20)
    read(40, 20) Ngal_z
    read(40, 20)
    read(40, 20) ny
    read(40, 10) dyq
    read(40, 20) ppp
    read(40, 10)
    read(40, 10)
    read(40, 20) flag_init
    read(40, 10)
    read(40, 10)
    read(40, 10)
    read(40, 30) fnameimb
    read(40, 10)
    read(40, 30) dirin
    read(40, 30) dirout
    read(40, 10)
    read(40, 10)
    read(40, 10) CFL
    read(40, 20) nwrite
    read(40, 21) flag_ctpress
    read(40, 10)
    if (flag_ctpress == 1) then
      read(40, 10) mpgx
    else if (flag_ctpress == 0) then
      read(40, 10)
      mpgx = 0d0
    else
      write(0, *) 'Wrong presure flag'
      error stop 1
    end if
    read(40, 10)
    read(40, 10)
    read(40, 10)
    read(40, 10)
    read(40, 10)
    read(40, 10) maxt
    close(40)
    Nspec_x = (2*Ngal_x)/3
    Nspec_z = (2*Ngal_z)/3
    nyv = ny - 2
    nyu = nyv + 1
    ncol = (Nspec_x/2 + 1)*Nspec_z
  end subroutine

  ! field files (start.f90:1545-1997 reads the same records); planes placed by their stored index j
  subroutine read_one(name, x, j0, j1, tt, it)
    character(*), intent(in) :: name
    integer, intent(in) :: j0, j1
    complex(8), intent(inout) :: x(j0:j1, ncol)
    real(8), intent(out) :: tt
    integer, intent(out) :: it
    real(8) :: Re2, alp2, bet2, dumm, yj
    integer :: nband2, dummint(88), jin, idd, nx, nz, ios, k, i, col
    integer, allocatable :: N2(:,:)
    real(8), allocatable :: buff(:,:)
    character(len=240) :: path
    path = trim(dirin)//'/'//name//trim(fnameimb(3:))
    open(10, file=path, form='unformatted', status='old')
    read(10) tt, Re2, alp2, bet2, dumm, nband2, it, dummint
    allocate(N2(4, 0:nband2+1))
    read(10) N2
    read(10)
    do
      read(10, iostat=ios) jin, idd, nx, nz, yj
      if (ios /= 0) exit
      backspace(10)
      allocate(buff(nx, nz))
      read(10) jin, idd, nx, nz, yj, buff
      if (jin >= j0 .and. jin <= j1) then
        do k = 1, min(nz, Nspec_z)
          do i = 0, min(nx/2, Nspec_x/2+1) - 1
            col = (k-1)*(Nspec_x/2+1) + i + 1
            x(jin, col) = dcmplx(buff(2*i+1, k), buff(2*i+2, k))
          end do
        end do
      end if
      deallocate(buff)
    end do
    close(10)
  end subroutine

  subroutine read_fields(u1, u2, u3, p)
    complex(8), intent(inout) :: u1(0:nyu+1, ncol), u2(0:nyv+1, ncol), u3(0:nyu+1, ncol), p(1:nyu, ncol)
    real(8) :: tt
    integer :: it
    call read_one('u1', u1, 0, nyu+1, tt, it)
    t = tt
    iter = it
    call read_one('u2', u2, 0, nyv+1, tt, it)
    call read_one('u3', u3, 0, nyu+1, tt, it)
    call read_one('p', p, 1, nyu, tt, it)
  end subroutine

  ! record_out's file layout (FOU3D.f90:2063-2161) for one field: planes j0..jw, those past the field's
  ! rows (j > j1) all zero.  One rank writes the p planes limPL_incw(pgrid) = 1..nyu+1 (start.f90:1187-1194,
  ! FOU3D.f90:2150-2153); modes_to_planes_UVP fills ppPL planes 1..nyu only, and ppPL was zeroed before it
  ! (FOU3D.f90:151-152, littleharsh.f90:439-443), so plane nyu+1 is a zero plane.
  subroutine write_one(name, x, j0, j1, jw, y, dthdy, gridid)
    character(*), intent(in) :: name
    integer, intent(in) :: j0, j1, jw, gridid
    complex(8), intent(in) :: x(j0:j1, ncol)
    real(8), intent(in) :: y(0:), dthdy(0:)
    integer :: N(4, 0:4), dummint(88), j, k, i, col
    real(8), allocatable :: buff(:,:)
    character(len=240) :: path
    character(len=4) :: e1, e2, e3
    character(len=5) :: e4
    write(e1, '(i4.4)') Nspec_x
    write(e2, '(i4.4)') Nspec_z
    write(e3, '(i4.4)') ny
    write(e4, '(i5.5)') int(100d0*t)
    path = trim(dirout)//'/'//name//'_'//e1//'x'//e2//'x'//e3//'_t'//e4//'.dat'
    N = 0
    N(1, 1:3) = Nspec_x
    N(1, 4) = -2
    N(2, 1:3) = Nspec_z
    N(3, 3) = nyv
    N(4, 3) = nyu
    dummint = 0
    allocate(buff(Nspec_x+2, Nspec_z))
    open(10, file=path, form='unformatted', status='replace')
    write(10) t, Re, alp, bet, mpgx, 3, iter, dummint
    write(10) N
    write(10) y, 1d0, dthdy
    do j = j0, jw
      buff = 0d0
      if (j <= j1) then
        do k = 1, Nspec_z
          do i = 0, Nspec_x/2
            col = (k-1)*(Nspec_x/2+1) + i + 1
            buff(2*i+1, k) = dble(x(j, col))
            buff(2*i+2, k) = dimag(x(j, col))
          end do
        end do
      end if
      write(10) j, gridid, Nspec_x+2, Nspec_z, y(j), buff
    end do
    close(10)
  end subroutine

  ! history file: opened as start.f90:273-284 (appended to when it exists, else created with the header
  ! t, Re, alp, bet, mpgx, nband and record_out's N)
  subroutine open_history()
    character(len=240) :: path
    character(len=4) :: e1, e2, e3
    integer :: N(4, 0:4)
    logical :: exists
    write(e1, '(i4.4)') Nspec_x
    write(e2, '(i4.4)') Nspec_z
    write(e3, '(i4.4)') ny
    path = trim(dirout)//'/hist_'//e1//'x'//e2//'x'//e3//'.dat'
    inquire(file=path, exist=exists)
    if (exists) then
      open(30, file=path, form='unformatted', status='old', position='append', action='write')
    else
      open(30, file=path, form='unformatted', status='new', action='write')
      N = 0
      N(1, 1:3) = Nspec_x
      N(1, 4) = -2
      N(2, 1:3) = Nspec_z
      N(3, 3) = nyv
      N(4, 3) = nyu
      write(30) t, Re, alp, bet, mpgx, 3
      write(30) N
    end if
  end subroutine

  ! the kRK = 1 output of FOU3D.f90:77-90, 138-154, 338-348: statistics every nstat steps; at write steps
  ! ppPL takes the current p before the statistics (FOU3D.f90:149-152 precede stats at :340), then
  ! record_out and the statistics and spectra files, after which the sums start again
  subroutine step_output(u1, u2, u3, p)
    complex(8), intent(inout) :: u1(0:nyu+1, ncol), u2(0:nyv+1, ncol), u3(0:nyu+1, ncol), p(1:nyu, ncol)
    logical :: flagwr
    flagwr = iter >= iwrite
    if (flagwr) call dns_hip_check(dns_refresh_stats_pressure(hdns), 'ppPL')
    if (iter - iter0 >= nstat) then
      iter0 = iter
      call dns_hip_check(dns_accumulate_stats(hdns), 'stats')
    end if
    if (flagwr) then
      call record_out_hip(u1, u2, u3, p)
      call write_stats_files()
      call dns_hip_check(dns_reset_stats(hdns), 'write_stats')
    end if
  end subroutine

  ! stats.f90:340-362 and spectra.f90:101-188 for one process holding every column
  subroutine write_stats_files()
    real(8), allocatable, target :: st(:,:), sp(:)
    real(8), allocatable :: yu(:), yv(:), du(:), dv(:)
    integer(c_int), target :: istat
    integer :: N(4, 0:4), dummint(88), q, off
    integer, parameter :: qorder(13) = [1, 2, 7, 8, 3, 4, 12, 13, 5, 6, 9, 10, 11]
    character(len=240) :: tag
    allocate(st(0:nyu+1, 13), sp(ncol * (4*nyu + nyv + 8)))
    call dns_hip_check(dns_get_stats(hdns, 0_c_int, c_loc(st), c_loc(sp), istat), 'write_stats')
    allocate(yu(0:ny), yv(0:ny-1), du(0:ny), dv(0:ny-1))
    call dns_hip_check(dns_ygrid(int(ny, c_int), dyq, int(ppp, c_int), yu, yv, du, dv), 'ygrid')
    tag = file_tag()
    call header_n(N)
    dummint = 0
    open(10, file=trim(dirout)//'/stats_'//trim(tag)//'.dat', form='unformatted', status='replace')
    write(10) t, Re, alp, bet, mpgx, 3, iter, dummint
    N(1, 1:3) = Ngal_x   ! the physical grid (Ngal) in place of the spectral sizes
    N(2, 1:3) = Ngal_z
    write(10) N
    write(10) yu, 1d0, du
    write(10) yv, 1d0, dv
    write(10) int(istat)
    do q = 1, 13
      select case (qorder(q))
      case (7, 8, 9, 10, 11)          ! v grid: planes 0..nyv+1
        write(10) st(0:nyv+1, qorder(q))
      case (12, 13)                   ! p grid: planes 1..nyu
        write(10) st(1:nyu, qorder(q))
      case default                    ! u grid: planes 0..nyu+1
        write(10) st(0:nyu+1, qorder(q))
      end select
    end do
    close(10)
    call header_n(N)
    off = 0
    call write_spectrum('spU', nyu + 2, 0, sp, off, tag, N, yu, du, yv, dv, int(istat))
    call write_spectrum('spV', nyv + 2, 0, sp, off, tag, N, yu, du, yv, dv, int(istat))
    call write_spectrum('spW', nyu + 2, 0, sp, off, tag, N, yu, du, yv, dv, int(istat))
    call write_spectrum('spUV', nyu + 2, 0, sp, off, tag, N, yu, du, yv, dv, int(istat))
    call write_spectrum('spP', nyu, 1, sp, off, tag, N, yu, du, yv, dv, int(istat))
  end subroutine

  ! recvwrspec (spectra.f90:125-188) for rows j0..j0+nr-1 of one spectrum (sp(off+1:) holds sp(j, column)):
  ! the column (i, k) goes to (i, kk) with kk = k for k <= Nz/2 and Nz+2-k above, +kz and -kz adding;
  ! entries 1..Nx/2 x 1..Nz/2 are doubled
  subroutine write_spectrum(name, nr, j0, sp, off, tag, N, yu, du, yv, dv, istat)
    character(*), intent(in) :: name, tag
    integer, intent(in) :: nr, j0, N(4, 0:4), istat
    integer, intent(inout) :: off
    real(8), intent(in) :: sp(:), yu(:), du(:), yv(:), dv(:)
    real(8), allocatable :: buff(:,:,:)
    integer :: col, i, k, kk, j, dummint(88)
    allocate(buff(0:Nspec_x/2, 1:Nspec_z/2+1, j0:j0+nr-1))
    buff = 0d0
    do col = 1, ncol
      i = mod(col - 1, Nspec_x/2 + 1)
      k = (col - 1) / (Nspec_x/2 + 1) + 1
      kk = k
      if (k > Nspec_z/2) kk = Nspec_z + 2 - k
      do j = 0, nr - 1
        buff(i, kk, j0 + j) = buff(i, kk, j0 + j) + sp(off + (col - 1) * nr + j + 1)
      end do
    end do
    buff(1:Nspec_x/2, 1:Nspec_z/2, :) = 2d0 * buff(1:Nspec_x/2, 1:Nspec_z/2, :)
    dummint = 0
    open(10, file=trim(dirout)//'/'//name//'_'//trim(tag)//'.dat', form='unformatted', status='replace')
    write(10) t, Re, alp, bet, mpgx, 3, iter, dummint
    write(10) N
    write(10) yu, 1d0, du
    write(10) yv, 1d0, dv
    write(10) istat
    write(10) buff
    close(10)
    off = off + nr * ncol
  end subroutine

  ! NNNNxNNNNxNNNN_TTTTT of the statistics files: the sizes and int(100 t) of record_out
  ! (start.f90:263-265, FOU3D.f90:2063)
  function file_tag() result(tag)
    character(len=240) :: tag
    character(len=4) :: e1, e2, e3
    character(len=5) :: e4
    write(e1, '(i4.4)') Nspec_x
    write(e2, '(i4.4)') Nspec_z
    write(e3, '(i4.4)') ny
    write(e4, '(i5.5)') int(100d0*t)
    tag = e1//'x'//e2//'x'//e3//'_'//e4
  end function

  subroutine header_n(N)
    integer, intent(out) :: N(4, 0:4)
    N = 0
    N(1, 1:3) = Nspec_x
    N(1, 4) = -2
    N(2, 1:3) = Nspec_z
    N(3, 3) = nyv
    N(4, 3) = nyu
  end subroutine

  ! record_out (FOU3D.f90:2003-2200): field files and one history record (FOU3D.f90:2196-2200)
  subroutine record_out_hip(u1, u2, u3, p)
    complex(8), intent(inout) :: u1(0:nyu+1, ncol), u2(0:nyv+1, ncol), u3(0:nyu+1, ncol), p(1:nyu, ncol)
    real(8) :: sc(10), h(13)
    call dns_hip_check(dns_get_scalars(hdns, sc), 'record_out')
    t = sc(1); dt = sc(2); mpgx = sc(4)
    call dns_hip_check(dns_download_fields(hdns, u1, u2, u3, p), 'record_out')
    call write_fields(u1, u2, u3, p)
    call dns_hip_check(dns_history_record(hdns, 0_c_int, h), 'record_out')
    if (flag_ctpress == 0) then
      write(30) flag_ctpress, iter, h(3), h(4), h(5), h(6), h(7), h(8), h(9), h(10), h(11), h(12), h(13)
    else
      write(30) flag_ctpress, iter, h(3), h(4), h(5), h(6), h(7), h(8), h(10), h(12), h(13)
    end if
    flush(30)
    iwrite = iwrite + nwrite
  end subroutine

  subroutine write_fields(u1, u2, u3, p)
    complex(8), intent(in) :: u1(0:nyu+1, ncol), u2(0:nyv+1, ncol), u3(0:nyu+1, ncol), p(1:nyu, ncol)
    real(8), allocatable :: yu(:), yv(:), du(:), dv(:)
    allocate(yu(0:ny), yv(0:ny-1), du(0:ny), dv(0:ny-1))
    call dns_hip_check(dns_ygrid(int(ny, c_int), dyq, int(ppp, c_int), yu, yv, du, dv), 'ygrid')
    call write_one('u1', u1, 0, nyu+1, nyu+1, yu, du, 1)
    call write_one('u2', u2, 0, nyv+1, nyv+1, yv, dv, 2)
    call write_one('u3', u3, 0, nyu+1, nyu+1, yu, du, 3)
    call write_one('p', p, 1, nyu, nyu+1, yu, du, 4)
  end subroutine

end program
