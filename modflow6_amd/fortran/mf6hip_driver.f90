! mf6hip_driver.f90 -- Fortran host that solves one IMS system on the GPU
! through mf6hip_module, as sln_ls would (NumericalSolution.f90:2604-2613).
!
! usage: mf6hip_driver <system.bin> <x_out.bin>
! system.bin (stream, native endian): nlay nrow ncol neq nja ilinmeth iter1
!   (int32), dvclose rclose relax (float64), ia(neq+1) ja(nja) (int32,
!   1-based), amat(nja) rhs(neq) x0(neq) (float64).
! x_out.bin: icnvg in_iter (int32) then x(neq) (float64).
program mf6hip_driver
  use iso_c_binding, only: c_int, c_double, c_ptr, c_null_ptr, c_char
  use mf6hip_module
  implicit none
  character(len=4096) :: fin, fout
  integer(c_int) :: nlay, nrow, ncol, neq, nja, ilinmeth, iter1
  real(c_double) :: dvclose, rclose, relax
  integer(c_int), allocatable :: ia(:), ja(:)
  real(c_double), allocatable :: amat(:), rhs(:), x(:)
  type(mf6hip_ims_settings) :: s
  type(c_ptr) :: h
  integer(c_int) :: rc, icnvg, in_iter
  character(kind=c_char) :: msg(512)
  integer :: u, i

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open (newunit=u, file=trim(fin), access='stream', form='unformatted', &
        status='old')
  read (u) nlay, nrow, ncol, neq, nja, ilinmeth, iter1
  read (u) dvclose, rclose, relax
  allocate (ia(neq + 1), ja(nja), amat(nja), rhs(neq), x(neq))
  read (u) ia, ja, amat, rhs, x
  close (u)

  s%dvclose = dvclose
  s%rclose = rclose
  s%icnvgopt = 0
  s%iter1 = iter1
  s%ilinmeth = ilinmeth
  s%iscl = 0
  s%iord = 0
  s%north = 0
  s%relax = relax
  s%level = 0
  s%droptol = 0.0_c_double
  s%precision = MF6HIP_FP64
  s%ordering = MF6HIP_ORDER_RB
  s%nlay = nlay
  s%nrow = nrow
  s%ncol = ncol
  s%inner_fp32_max = 0
  s%inner_fp32_rtol = 0.0_c_double
  s%kernel_variant = 0
  s%preconditioner = 0
  s%coarse_size = 0

  rc = mf6hip_create(neq, nja, ia, ja, s, h)
  if (rc /= 0) call fail(rc)
  rc = mf6hip_solve(h, 1, 1, amat, rhs, x, icnvg, in_iter, c_null_ptr)
  if (rc /= 0) call fail(rc)
  rc = mf6hip_destroy(h)

  open (newunit=u, file=trim(fout), access='stream', form='unformatted', &
        status='replace')
  write (u) icnvg, in_iter
  write (u) x
  close (u)
  write (*, '(a,i0,a,i0)') 'mf6hip_driver: icnvg=', icnvg, ' inner=', in_iter

contains

  subroutine fail(code)
    integer(c_int), intent(in) :: code
    integer(c_int) :: r2
    r2 = mf6hip_last_error(msg, 512)
    write (*, '(a,i0,a,512a)') 'mf6hip error ', code, ': ', &
      (msg(i), i=1, index_of_nul())
    stop 1
  end subroutine fail

  integer function index_of_nul()
    integer :: k
    index_of_nul = 0
    do k = 1, 512
      if (msg(k) == char(0)) exit
      index_of_nul = k
    end do
  end function index_of_nul
end program mf6hip_driver
