! Module state of the drop-in driver and the wrappers with the reference's subroutine names and argument
! lists (littleharsh.f90:682-993), each forwarding to the HIP library through ISO_C_BINDING.
!
! The fields live in GPU memory between calls.  Two safeguards keep a caller from reading stale host arrays
! without knowing:
!  * call order: within a substep the wrappers must come in the reference's order (littleharsh.f90:165-222):
!    RHS0_u1, RHS0_u2, RHS0_u3, nonlinear, solveU x3, [meanflow_ctU], solveP, v_corr, and meanflow_ctP only
!    between substeps; any other order stops with a message naming the call and the one expected.
!  * mirror mode (environment DNS_HIP_MIRROR=1): after every wrapper call the arrays that the reference's
!    subroutine would have written (du, Nu, u, p, psi, div) are copied back from the GPU into the caller's
!    arrays, so the reference's debug prints between calls (littleharsh.f90:174-176) see current values.
!    Without it only record_out's arrays are synchronised (at output).  Host-side changes to the arrays
!    between calls are not seen by the GPU in either mode: dns_upload_fields is the way in.
module lh_state
  use iso_c_binding
  use dns_hip
  implicit none
  type(c_ptr) :: hdns = c_null_ptr
  integer :: kRK = 1, solveU_calls = 0
  integer :: Nspec_x, Nspec_z, Ngal_x, Ngal_z, nyu, nyv, ny, ncol, flag_ctpress, flag_init, nwrite, ppp
  real(8) :: Re, alp, bet, CFL, mpgx, dyq, maxt, t, dt
  integer :: iter, iwrite, iter0, nstat
  character(len=120) :: fnameimb, dirin, dirout
  ! call order within a substep: 0 expects RHS0_u1 (or meanflow_ctP between substeps), 1 RHS0_u2,
  ! 2 RHS0_u3, 3 nonlinear, 4..6 the three solveU, 7 meanflow_ctU or solveP, 8 solveP (after
  ! meanflow_ctU), 9 v_corr
  integer :: stage = 0
  logical :: mirror = .false.
  complex(8), allocatable :: m1(:,:), m2(:,:), m3(:,:), mp(:,:)   ! mirror copies (fields or du)
contains
  subroutine lh_mirror_init()
    character(len=16) :: v
    integer :: st
    call get_environment_variable('DNS_HIP_MIRROR', v, status=st)
    mirror = st == 0 .and. trim(v) == '1'
  end subroutine

  subroutine lh_order(allowed1, allowed2, next, who)
    integer, intent(in) :: allowed1, allowed2, next
    character(*), intent(in) :: who
    character(len=*), parameter :: names(0:9) = [character(len=32) :: 'RHS0_u1 (or meanflow_ctP)', 'RHS0_u2', &
      'RHS0_u3', 'nonlinear', 'solveU (first)', 'solveU (second)', 'solveU (third)', &
      'meanflow_ctU or solveP', 'solveP', 'v_corr']
    if (stage /= allowed1 .and. stage /= allowed2) then
      write(0, '(A)') 'littleharsh_hip: '//who//' called out of the reference''s order; expected '// &
        trim(names(stage))//' (the fields live on the GPU and are advanced only in that order)'
      error stop 2
    end if
    stage = next
  end subroutine

  ! the current fields into m1 (u1), m2 (u2), m3 (u3), mp (p)
  subroutine lh_mirror_fields()
    if (.not. allocated(m1)) allocate(m1(0:nyu+1, ncol), m2(0:nyv+1, ncol), m3(0:nyu+1, ncol), mp(1:nyu, ncol))
    call dns_hip_check(dns_download_fields(hdns, m1, m2, m3, mp), 'mirror')
  end subroutine
end module

! ---- wrappers with the reference's names (littleharsh.f90:682-993) ----
subroutine RHS0_u1(du1, u1, Nu1, p, myid)
  use lh_state
  implicit none
  integer myid
  complex(8) du1(0:nyu+1, ncol), u1(*), Nu1(*), p(*)
  complex(8), allocatable :: d1(:,:)
  call lh_order(0, 0, 1, 'RHS0_u1')
  solveU_calls = 0
  call dns_hip_check(dns_rhs0(hdns, int(kRK, c_int)), 'RHS0_u1')   ! RHS0 of u1, u2 and u3
  if (mirror) then   ! du of the three velocities: du1 here, du2 and du3 kept for RHS0_u2/u3
    if (allocated(m1)) deallocate(m1, m2, m3, mp)
    allocate(d1(1:nyu, ncol), m2(1:nyv, ncol), m3(1:nyu, ncol))
    call dns_hip_check(dns_download_rhs(hdns, d1, m2, m3), 'mirror')
    du1(1:nyu, :) = d1
  end if
end subroutine

subroutine RHS0_u2(du2, u2, Nu2, p, myid)
  use lh_state
  implicit none
  integer myid
  complex(8) du2(0:nyv+1, ncol), u2(*), Nu2(*), p(*)
  call lh_order(1, 1, 2, 'RHS0_u2')   ! computed together with RHS0_u1 on the GPU
  if (mirror) du2(1:nyv, :) = m2
end subroutine

subroutine RHS0_u3(du3, u3, Nu3, p, myid)
  use lh_state
  implicit none
  integer myid
  complex(8) du3(0:nyu+1, ncol), u3(*), Nu3(*), p(*)
  call lh_order(2, 2, 3, 'RHS0_u3')   ! computed together with RHS0_u1 on the GPU
  if (mirror) then
    du3(1:nyu, :) = m3
    deallocate(m2, m3)
  end if
end subroutine

subroutine nonlinear(Nu1, Nu2, Nu3, u1, u2, u3, du1, du2, du3, p, div, myid, status, ierr)
  use lh_state
  implicit none
  integer myid, status(*), ierr
  complex(8) Nu1(1:nyu, ncol), Nu2(1:nyv, ncol), Nu3(1:nyu, ncol), u1(*), u2(*), u3(*), p(*), div(*)
  complex(8) du1(0:nyu+1, ncol), du2(0:nyv+1, ncol), du3(0:nyu+1, ncol)
  complex(8), allocatable :: d1(:,:), d2(:,:), d3(:,:)
  call lh_order(3, 3, 4, 'nonlinear')
  call dns_hip_check(dns_nonlinear(hdns, int(kRK, c_int)), 'nonlinear')
  if (mirror) then
    allocate(d1(1:nyu, ncol), d2(1:nyv, ncol), d3(1:nyu, ncol))
    call dns_hip_check(dns_download_nonlinear(hdns, Nu1, Nu2, Nu3), 'mirror')
    call dns_hip_check(dns_download_rhs(hdns, d1, d2, d3), 'mirror')
    du1(1:nyu, :) = d1
    du2(1:nyv, :) = d2
    du3(1:nyu, :) = d3
  end if
  ierr = 0
end subroutine

subroutine solveU(u, du, grid, a_grid, myid)
  use lh_state
  implicit none
  integer grid, myid
  complex(8) u(*), du(*)
  real(8) a_grid(*)
  call lh_order(4 + solveU_calls, 4 + solveU_calls, 5 + solveU_calls, 'solveU')
  ! the three velocity solves of a substep run in one launch at the first call
  if (solveU_calls == 0) call dns_hip_check(dns_solveU(hdns, int(kRK, c_int)), 'solveU')
  if (mirror) then   ! the calls come as u1, u2, u3 (littleharsh.f90:197-199)
    if (solveU_calls == 0) call lh_mirror_fields()
    if (solveU_calls == 0) call copy_u(u, m1, nyu + 2)
    if (solveU_calls == 1) call copy_u(u, m2, nyv + 2)
    if (solveU_calls == 2) call copy_u(u, m3, nyu + 2)
  end if
  solveU_calls = solveU_calls + 1
contains
  subroutine copy_u(dst, src, nr)
    integer, intent(in) :: nr
    complex(8), intent(inout) :: dst(nr, ncol)
    complex(8), intent(in) :: src(nr, ncol)
    dst = src
  end subroutine
end subroutine

subroutine meanflow_ctU(u1, myid)
  use lh_state
  implicit none
  integer myid
  complex(8) u1(0:nyu+1, ncol)
  call lh_order(7, 7, 8, 'meanflow_ctU')
  call dns_hip_check(dns_meanflow_ctU(hdns, int(kRK, c_int)), 'meanflow_ctU')
  if (mirror) then
    call lh_mirror_fields()
    u1 = m1
  end if
end subroutine

subroutine solveP(p, psi, u1, u2, u3, div, myid)
  use lh_state
  implicit none
  integer myid
  complex(8) p(1:nyu, ncol), psi(1:nyu, ncol), div(1:nyu, ncol), u1(*), u2(*), u3(*)
  call lh_order(7, 8, 9, 'solveP')
  call dns_hip_check(dns_solveP(hdns, int(kRK, c_int)), 'solveP')
  if (mirror) then
    call lh_mirror_fields()
    p = mp
    call dns_hip_check(dns_download_pressure_correction(hdns, psi, div), 'mirror')
  end if
end subroutine

subroutine v_corr(u1, u2, u3, psi, div, myid, status, ierr)
  use lh_state
  implicit none
  integer myid, status(*), ierr
  complex(8) u1(0:nyu+1, ncol), u2(0:nyv+1, ncol), u3(0:nyu+1, ncol), psi(1:nyu, ncol), div(1:nyu, ncol)
  call lh_order(9, 9, 0, 'v_corr')
  call dns_hip_check(dns_vcorr(hdns, int(kRK, c_int)), 'v_corr')
  if (mirror) then
    call lh_mirror_fields()
    u1 = m1
    u2 = m2
    u3 = m3
    call dns_hip_check(dns_download_pressure_correction(hdns, psi, div), 'mirror')
  end if
  ierr = 0
end subroutine

subroutine meanflow_ctP(u1, myid)
  use lh_state
  implicit none
  integer myid
  complex(8) u1(*)
  call lh_order(0, 0, 0, 'meanflow_ctP')
  call dns_hip_check(dns_meanflow_ctP(hdns), 'meanflow_ctP')
end subroutine
