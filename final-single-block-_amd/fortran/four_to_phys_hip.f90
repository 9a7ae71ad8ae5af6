! The post-processors four_to_phys.f90 and four_to_phys_N.f90 of the reference, with the per-plane
! resize + cft + rft done on the GPU by dns_four_to_phys (one call per field, every plane at once).
!
!   four_to_phys_hip        -> physical planes on the 3/2 grid (Ngal = 3N/2), as four_to_phys
!   four_to_phys_hip N      -> physical planes on the spectral grid (Ngal = N), as four_to_phys_N
!
! Reads input_FtP.txt (nx, nz, ny, t with the reference's FORMATs; files are named by int(10 t)),
! converts output/{u1,u2,u3,p}_NNNNxNNNNxNNNN_tTTTTT.dat to output/{u1,u2,u3,p}_phys_..., writing the
! records in the reference's order: its header records, then for each band (of N in the file) the plane at
! the band interface (again, at bands 2 and 3), the band's planes and, after the last band, the top plane;
! p by the planes ranges of four_to_phys.f90:268-302 (four_to_phys_N.f90:267-300).
! Bands of different sizes are not supported (every band of the single-block code has the same N).

module ftp_state
  use iso_c_binding
  use dns_hip
  implicit none
  integer :: nband, iter, nfx, nfz
  integer, allocatable :: N(:,:), Ngal(:,:), dummint(:)
  real(8) :: t, Re, alp, bet, mpgx, dthetai
  real(8), allocatable :: yu(:), dthdyu(:), yv(:), dthdyv(:)
  logical :: nvar = .false.
  character(len=4) :: ext1, ext2, ext3
  character(len=5) :: ext4

contains

  ! one field: kind 1 u1, 2 u2, 3 u3, 4 p
  subroutine convert(name, kind)
    character(*), intent(in) :: name
    integer, intent(in) :: kind
    character(len=120) :: fin, fout
    integer :: nx, nz, nx2, nz2, iband, j, row, nops, cap, jd, id, jlast, idlast, ios, k
    real(8) :: dumm_y, ms
    real(8), allocatable :: spec(:,:,:), phys(:,:,:), ylist(:), grow(:,:,:), plane(:,:)
    integer, allocatable :: jlist(:), idlist(:)
    integer(c_int) :: rc

    fin = 'output/'//name//'_'//ext1//'x'//ext2//'x'//ext3//'_t'//ext4//'.dat'
    fout = 'output/'//name//'_phys_'//ext1//'x'//ext2//'x'//ext3//'_t'//ext4//'.dat'
    open(10, file=fin, form='unformatted', status='old')
    open(20, file=fout, form='unformatted', status='replace')
    if (kind == 1) then
      allocate(dummint(88))
      read(10) t, Re, alp, bet, mpgx, nband, iter, dummint
      allocate(N(4, 0:nband+1), Ngal(4, 0:nband+1))
      read(10) N
      Ngal = N
      do iband = 1, nband
        if (.not. nvar) then
          Ngal(1, iband) = N(1, iband)*3/2
          Ngal(2, iband) = N(2, iband)*3/2
        end if
        if (N(1, iband) /= N(1, 1) .or. N(2, iband) /= N(2, 1)) then
          write(0, '(A)') 'four_to_phys_hip: bands of different sizes are not supported'
          error stop 1
        end if
      end do
      allocate(yu(N(4,0):N(4,nband)+1), dthdyu(N(4,0):N(4,nband)+1))
      read(10) yu, dthetai, dthdyu
    else
      read(10)
      read(10)
    end if
    write(20) t, Re, alp, bet, mpgx, nband, iter, dummint
    write(20) Ngal
    if (kind == 2) then
      allocate(yv(N(3,0):N(3,nband)+1), dthdyv(N(3,0):N(3,nband)+1))
      read(10) yv, dthetai, dthdyv
      write(20) yv, dthetai, dthdyv
    else
      if (kind /= 1) read(10)
      write(20) yu, dthetai, dthdyu
    end if

    nx = N(1, 1) + 2
    nz = N(2, 1)
    nx2 = Ngal(1, 1) + 2
    nz2 = Ngal(2, 1)
    row = 4
    if (kind == 2) row = 3
    cap = N(row, nband) + 2*nband + 4
    allocate(spec(nx, nz, cap), jlist(cap), idlist(cap), ylist(cap), plane(nx, nz))
    nops = 0
    jlast = 0
    idlast = 0
    if (kind /= 4) then
      do iband = 1, nband
        if (iband == 1) then
          call read_plane()
        else   ! the previous band's last plane again, at the interface
          if (kind == 2) then
            call emit(jlast, idlast, yv(N(row, iband-1)))
          else
            call emit(N(row, iband-1), 1, yu(N(row, iband-1)))
          end if
        end if
        do j = N(row, iband-1) + 1, N(row, iband)
          call read_plane()
        end do
        if (iband == nband) call read_plane()
      end do
    else
      do iband = 1, nband
        if (iband == 1) then
          do j = N(4,0) + 1, N(4,1) - 1
            call read_plane()
          end do
        else if (iband == 3) then
          if (nvar) then
            do j = N(4,2) + 2, N(4,3)
              call read_plane()
            end do
          else
            do j = N(4,2) + 1, N(4,3) - 1
              call read_plane()
            end do
          end if
        else
          if (nvar) then
            do j = N(4,1), N(4,2) + 1
              call read_plane()
            end do
          else
            do j = N(4,1), N(4,2)
              call read_plane()
            end do
          end if
        end if
      end do
    end if
    close(10)

    allocate(phys(nx2, nz2, max(nops, 1)))
    rc = dns_four_to_phys(int(nx - 2, c_int), int(nz, c_int), int(nx2 - 2, c_int), int(nz2, c_int), &
                          int(nops, c_int), spec, phys, ms)
    call dns_hip_check(rc, 'four_to_phys')
    do k = 1, nops
      write(20) jlist(k), idlist(k), nx2, nz2, ylist(k), phys(:, :, k)
    end do
    close(20)
    write(*, '(A,A,A,I0,A,F9.3,A)') ' Writing ', trim(fout), ' (', nops, ' planes, ', ms, ' ms on the GPU)'

  contains

    subroutine read_plane()
      integer :: nxr, nzr
      read(10, iostat=ios) jd, id, nxr, nzr, dumm_y, plane
      if (ios /= 0) then
        write(0, '(A,A,A,I0)') 'four_to_phys_hip: ', trim(fin), ': end of file or read error after plane ', nops
        error stop 1
      end if
      jlast = jd
      idlast = id
      call emit(jd, id, dumm_y)
    end subroutine

    ! the last plane read, written with record header (jj, ii, yy)
    subroutine emit(jj, ii, yy)
      integer, intent(in) :: jj, ii
      real(8), intent(in) :: yy
      if (nops == cap) then
        call grow_to(2*cap)
      end if
      nops = nops + 1
      spec(:, :, nops) = plane
      jlist(nops) = jj
      idlist(nops) = ii
      ylist(nops) = yy
    end subroutine

    subroutine grow_to(ncap)
      integer, intent(in) :: ncap
      integer, allocatable :: itmp(:)
      real(8), allocatable :: rtmp(:)
      allocate(grow(nx, nz, ncap))
      grow(:, :, 1:cap) = spec
      call move_alloc(grow, spec)
      allocate(itmp(ncap)); itmp(1:cap) = jlist; call move_alloc(itmp, jlist)
      allocate(itmp(ncap)); itmp(1:cap) = idlist; call move_alloc(itmp, idlist)
      allocate(rtmp(ncap)); rtmp(1:cap) = ylist; call move_alloc(rtmp, ylist)
      cap = ncap
    end subroutine
  end subroutine

end module

program four_to_phys_hip
  use ftp_state
  implicit none
  character(len=16) :: arg
  real(8) :: tin
  integer :: nyin

  if (command_argument_count() >= 1) then
    call get_command_argument(1, arg)
    nvar = trim(arg) == 'N'
  end if
  open(40, file='input_FtP.txt', form='formatted', status='old')
  read(40, 20) nfx
  read(40, 20) nfz
  read(40, 20) nyin
  read(40, 10) tin
  close(40)
  write(ext1, '(i4.4)') nfx
  write(ext2, '(i4.4)') nfz
  write(ext3, '(i4.4)') nyin
  write(ext4, '(i5.5)') int(10d0*tin)

  call convert('u1', 1)
  call convert('u2', 2)
  call convert('u3', 3)
  call convert('p', 4)

10 format(7X, D10.1)
20 format(7X, I10)

end program
