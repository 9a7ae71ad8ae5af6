! ISO_C_BINDING interface to the C-ABI of include/dns_hip.h (libdns_hip.so).
!
! The reference calls its hot path as external Fortran subroutines with implicit interfaces from the
! main loop (littleharsh.f90:163-233).  This module binds every C entry point by name, and
! dns_hip_check turns a non-zero status into `error stop` with the library's message, which keeps the
! reference's fatal-stop error behaviour (start.f90:203-204, rft_buff.f:308-310).
module dns_hip
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: DNS_OK = 0, DNS_PREC_FP64 = 0, DNS_PREC_MIXED = 1

  type, bind(C) :: dns_params
    integer(c_int) :: Mx, Mz, ny, ppp, flag_ctpress, precision
    real(c_double) :: Re, alp, bet, CFL, mpgx, dyq, fixed_dt
    integer(c_int) :: plane_batch
    integer(c_int) :: members
    integer(c_int) :: column_path
    integer(c_int) :: reserved(5)
  end type

  interface
    function dns_last_error() bind(C, name='dns_last_error')
      import :: c_ptr
      type(c_ptr) :: dns_last_error
    end function
    integer(c_int) function dns_ygrid(ny, dyq, ppp, yu, yv, dthdyu, dthdyv) bind(C, name='dns_ygrid')
      import :: c_int, c_double
      integer(c_int), value :: ny, ppp
      real(c_double), value :: dyq
      real(c_double) :: yu(*), yv(*), dthdyu(*), dthdyv(*)
    end function
    integer(c_int) function dns_create(prm, h) bind(C, name='dns_create')
      import :: c_int, c_ptr, dns_params
      type(dns_params), intent(in) :: prm
      type(c_ptr), intent(out) :: h
    end function
    integer(c_int) function dns_destroy(h) bind(C, name='dns_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function dns_sizes(h, out) bind(C, name='dns_sizes')
      import :: c_int, c_ptr, c_long_long
      type(c_ptr), value :: h
      integer(c_long_long) :: out(10)
    end function
    integer(c_int) function dns_setup_arrays(h, yu, yv, dthdyu, dthdyv, dyu2i, dyv2i, w) &
        bind(C, name='dns_setup_arrays')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double) :: yu(*), yv(*), dthdyu(*), dthdyv(*), dyu2i(*), dyv2i(*), w(*)
    end function
    integer(c_int) function dns_upload_fields(h, u1, u2, u3, p, t, iter, QxT) bind(C, name='dns_upload_fields')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      complex(c_double), intent(in) :: u1(*), u2(*), u3(*), p(*)
      real(c_double), value :: t, QxT
      integer(c_int), value :: iter
    end function
    integer(c_int) function dns_download_fields(h, u1, u2, u3, p) bind(C, name='dns_download_fields')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      complex(c_double) :: u1(*), u2(*), u3(*), p(*)
    end function
    integer(c_int) function dns_download_nonlinear(h, Nu1, Nu2, Nu3) bind(C, name='dns_download_nonlinear')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      complex(c_double) :: Nu1(*), Nu2(*), Nu3(*)
    end function
    integer(c_int) function dns_upload_nonlinear(h, Nu1, Nu2, Nu3) bind(C, name='dns_upload_nonlinear')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      complex(c_double), intent(in) :: Nu1(*), Nu2(*), Nu3(*)
    end function
    integer(c_int) function dns_download_rhs(h, du1, du2, du3) bind(C, name='dns_download_rhs')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      complex(c_double) :: du1(*), du2(*), du3(*)
    end function
    integer(c_int) function dns_download_flux(h, uv_f, vv_c, wv_f) bind(C, name='dns_download_flux')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      complex(c_double) :: uv_f(*), vv_c(*), wv_f(*)
    end function
    integer(c_int) function dns_download_pressure_correction(h, psi, div) &
        bind(C, name='dns_download_pressure_correction')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      complex(c_double) :: psi(*), div(*)
    end function
    integer(c_int) function dns_get_scalars(h, s) bind(C, name='dns_get_scalars')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double) :: s(10)
    end function
    integer(c_int) function dns_set_dt(h, dt) bind(C, name='dns_set_dt')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), value :: dt
    end function
    integer(c_int) function dns_rhs0(h, kRK) bind(C, name='dns_rhs0')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: kRK
    end function
    integer(c_int) function dns_nonlinear(h, kRK) bind(C, name='dns_nonlinear')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: kRK
    end function
    integer(c_int) function dns_solveU(h, kRK) bind(C, name='dns_solveU')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: kRK
    end function
    integer(c_int) function dns_meanflow_ctU(h, kRK) bind(C, name='dns_meanflow_ctU')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: kRK
    end function
    integer(c_int) function dns_solveP(h, kRK) bind(C, name='dns_solveP')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: kRK
    end function
    integer(c_int) function dns_vcorr(h, kRK) bind(C, name='dns_vcorr')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: kRK
    end function
    integer(c_int) function dns_meanflow_ctP(h) bind(C, name='dns_meanflow_ctP')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function dns_rk3_substep(h, kRK) bind(C, name='dns_rk3_substep')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: kRK
    end function
    integer(c_int) function dns_rk3_step(h, nsteps) bind(C, name='dns_rk3_step')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: nsteps
    end function
    integer(c_int) function dns_rk3_step_graph(h, nsteps) bind(C, name='dns_rk3_step_graph')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: nsteps
    end function
    integer(c_int) function dns_plane_fft_roundtrip(h, d_in, d_phys, d_back, nplanes) &
        bind(C, name='dns_plane_fft_roundtrip')
      import :: c_int, c_ptr
      type(c_ptr), value :: h, d_in, d_phys, d_back
      integer(c_int), value :: nplanes
    end function
    integer(c_int) function dns_modes_to_planes(h, d_cols, nrows, d_planes) bind(C, name='dns_modes_to_planes')
      import :: c_int, c_ptr
      type(c_ptr), value :: h, d_cols, d_planes
      integer(c_int), value :: nrows
    end function
    integer(c_int) function dns_planes_to_modes(h, d_planes, nrows, d_cols) bind(C, name='dns_planes_to_modes')
      import :: c_int, c_ptr
      type(c_ptr), value :: h, d_planes, d_cols
      integer(c_int), value :: nrows
    end function
    integer(c_int) function dns_history_record(h, m, rec) bind(C, name='dns_history_record')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: m
      real(c_double), intent(out) :: rec(13)
    end function
    integer(c_int) function dns_accumulate_stats(h) bind(C, name='dns_accumulate_stats')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function dns_reset_stats(h) bind(C, name='dns_reset_stats')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    ! ppPL = modes_to_planes_UVP(p) (FOU3D.f90:149-152): Pm, P2m sum this copy until the next call
    integer(c_int) function dns_refresh_stats_pressure(h) bind(C, name='dns_refresh_stats_pressure')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    ! stats(nyu+2, 13) in stats.f90's order; spectra: spU, spV, spW, spUV, spP one after another, each
    ! sp(rows, columns) as in spectra.f90 (rows nyu+2, nyv+2, nyu+2, nyu+2, nyu); c_null_ptr skips either
    integer(c_int) function dns_get_stats(h, m, stats, spectra, istat) bind(C, name='dns_get_stats')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: m
      type(c_ptr), value :: stats, spectra
      integer(c_int), intent(out) :: istat
    end function
    ! fin(Nx+2, Nz, nplanes) spectral planes -> fout(Mx+2, Mz, nplanes) physical planes
    integer(c_int) function dns_four_to_phys(Nx, Nz, Mx, Mz, nplanes, fin, fout, ms) &
        bind(C, name='dns_four_to_phys')
      import :: c_int, c_double
      integer(c_int), value :: Nx, Nz, Mx, Mz, nplanes
      real(c_double), intent(in) :: fin(*)
      real(c_double), intent(out) :: fout(*)
      real(c_double), intent(out) :: ms
    end function
    ! ensemble of several sizes: handles(1:n) are type(c_ptr) handles from dns_create
    integer(c_int) function dns_ensemble_create(handles, n, e) bind(C, name='dns_ensemble_create')
      import :: c_int, c_ptr
      type(c_ptr), intent(in) :: handles(*)
      integer(c_int), value :: n
      type(c_ptr), intent(out) :: e
    end function
    integer(c_int) function dns_ensemble_destroy(e) bind(C, name='dns_ensemble_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: e
    end function
    integer(c_int) function dns_ensemble_step(e, nsteps) bind(C, name='dns_ensemble_step')
      import :: c_int, c_ptr
      type(c_ptr), value :: e
      integer(c_int), value :: nsteps
    end function
    integer(c_int) function dns_ensemble_step_graph(e, nsteps) bind(C, name='dns_ensemble_step_graph')
      import :: c_int, c_ptr
      type(c_ptr), value :: e
      integer(c_int), value :: nsteps
    end function
    integer(c_int) function dns_ensemble_sync(e) bind(C, name='dns_ensemble_sync')
      import :: c_int, c_ptr
      type(c_ptr), value :: e
    end function
    integer(c_int) function dns_ensemble_set_timing(e, on) bind(C, name='dns_ensemble_set_timing')
      import :: c_int, c_ptr
      type(c_ptr), value :: e
      integer(c_int), value :: on
    end function
    integer(c_int) function dns_ensemble_get_timing(e, t, n) bind(C, name='dns_ensemble_get_timing')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: e
      real(c_double), intent(out) :: t(7)
      integer(c_int), intent(out) :: n
    end function
    ! mode 0 grouped launches, 1 one persistent launch per substep; units: c_null_ptr or c_loc of 10 ints
    integer(c_int) function dns_ensemble_set_launch(e, mode, units) bind(C, name='dns_ensemble_set_launch')
      import :: c_int, c_ptr
      type(c_ptr), value :: e
      integer(c_int), value :: mode
      type(c_ptr), value :: units
    end function
    integer(c_int) function dns_ensemble_get_launch(e, mode, nwg, ntask) bind(C, name='dns_ensemble_get_launch')
      import :: c_int, c_ptr
      type(c_ptr), value :: e
      integer(c_int), intent(out) :: mode, nwg
      integer(c_int), intent(out) :: ntask(3)
    end function
    ! tile shapes of the grouped launches: var(6), wg(5)
    integer(c_int) function dns_ensemble_set_tiles(e, var, wg) bind(C, name='dns_ensemble_set_tiles')
      import :: c_int, c_ptr
      type(c_ptr), value :: e
      integer(c_int), intent(in) :: var(6), wg(5)
    end function
    integer(c_int) function dns_ensemble_get_tiles(e, var, wg) bind(C, name='dns_ensemble_get_tiles')
      import :: c_int, c_ptr
      type(c_ptr), value :: e
      integer(c_int), intent(out) :: var(6), wg(5)
    end function
    ! var, wg and ms may be c_null_ptr (ms: DNS_ENS_TUNE_MS = 28 doubles)
    integer(c_int) function dns_ensemble_autotune(e, iters, var, wg, ms) bind(C, name='dns_ensemble_autotune')
      import :: c_int, c_ptr
      type(c_ptr), value :: e
      integer(c_int), value :: iters
      type(c_ptr), value :: var, wg, ms
    end function
    ! device pointers (type(c_ptr)) and a stream (c_null_ptr: the null stream); no device allocation
    integer(c_int) function dns_four_to_phys_noalloc(Nx, Nz, Mx, Mz, nplanes, d_in, d_out, stream) &
        bind(C, name='dns_four_to_phys_noalloc')
      import :: c_int, c_ptr
      integer(c_int), value :: Nx, Nz, Mx, Mz, nplanes
      type(c_ptr), value :: d_in, d_out, stream
    end function
    integer(c_int) function dns_set_column_path(h, path) bind(C, name='dns_set_column_path')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: path
    end function
    integer(c_int) function dns_get_column_path(h, path) bind(C, name='dns_get_column_path')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), intent(out) :: path
    end function
    integer(c_int) function dns_sync(h) bind(C, name='dns_sync')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function dns_set_timing(h, on) bind(C, name='dns_set_timing')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: on
    end function
    integer(c_int) function dns_get_timing(h, t, n) bind(C, name='dns_get_timing')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double) :: t(7)
      integer(c_int) :: n
    end function
    integer(c_int) function dns_upload_member(h, m, u1, u2, u3, p, t, iter, QxT) bind(C, name='dns_upload_member')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: m, iter
      real(c_double) :: u1(*), u2(*), u3(*), p(*)
      real(c_double), value :: t, QxT
    end function
    integer(c_int) function dns_download_member(h, m, u1, u2, u3, p) bind(C, name='dns_download_member')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: m
      real(c_double) :: u1(*), u2(*), u3(*), p(*)
    end function
    integer(c_int) function dns_set_member_params(h, m, Re, mpgx) bind(C, name='dns_set_member_params')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: m
      real(c_double), value :: Re, mpgx
    end function
    integer(c_int) function dns_get_member_scalars(h, m, s) bind(C, name='dns_get_member_scalars')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: m
      real(c_double) :: s(10)
    end function
    integer(c_int) function dns_get_plane_variants(h, var, count) bind(C, name='dns_get_plane_variants')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int) :: var(9), count(9)
    end function
    integer(c_int) function dns_set_plane_variants(h, var) bind(C, name='dns_set_plane_variants')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), intent(in) :: var(9)
    end function
    integer(c_int) function dns_autotune(h, iters, chosen, ms) bind(C, name='dns_autotune')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: iters
      integer(c_int) :: chosen(9)
      real(c_double) :: ms(27)
    end function
    function c_strlen(s) bind(C, name='strlen')
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
      integer(c_size_t) :: c_strlen
    end function
  end interface

contains

  ! status check with the reference's fatal-stop semantics
  subroutine dns_hip_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (rc == DNS_OK) return
    n = int(c_strlen(dns_last_error()))
    call c_f_pointer(dns_last_error(), msg, [n])
    write(0,'(A,A,I0,A,*(A))') trim(where), ': status ', rc, ': ', msg(1:n)
    error stop 1
  end subroutine

end module
