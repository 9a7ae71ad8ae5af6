! Checks the drop-in wrappers' call-order guard without a GPU: a substep that starts with RHS0_u2 (not
! RHS0_u1) must stop with a message before any call reaches the library.
program lh_order_check
  use lh_state
  implicit none
  complex(8) :: a(4)
  integer :: myid = 0
  nyu = 1
  nyv = 0
  ncol = 1
  call RHS0_u2(a, a, a, a, myid)
  write(*, '(A)') 'lh_order_check: the out-of-order call was not stopped'
end program
