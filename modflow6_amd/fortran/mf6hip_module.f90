! mf6hip_module.f90 -- ISO_C_BINDING interface to libmf6hip.so
! (include/mf6hip.h).  This is the module a Fortran host (MODFLOW 6's
! NumericalSolution / a LinearSolverBaseType extension) uses to call the
! gfx950 solver: every argument is c_int / c_double / c_ptr, CSR arrays are
! passed 1-based exactly as the memory manager holds them.
module mf6hip_module
  use iso_c_binding, only: c_int, c_double, c_ptr, c_char
  implicit none
  private

  integer(c_int), parameter, public :: MF6HIP_FP64 = 0
  integer(c_int), parameter, public :: MF6HIP_MIXED = 1
  integer(c_int), parameter, public :: MF6HIP_MIXED_FP16 = 2
  integer(c_int), parameter, public :: MF6HIP_ORDER_RB = 0

  !> mf6hip_ims_settings: ImsLinearSettingsType fields + device opt-ins
  type, bind(C), public :: mf6hip_ims_settings
    real(c_double) :: dvclose
    real(c_double) :: rclose
    integer(c_int) :: icnvgopt
    integer(c_int) :: iter1
    integer(c_int) :: ilinmeth
    integer(c_int) :: iscl
    integer(c_int) :: iord
    integer(c_int) :: north
    real(c_double) :: relax
    integer(c_int) :: level
    real(c_double) :: droptol
    integer(c_int) :: precision
    integer(c_int) :: ordering
    integer(c_int) :: nlay
    integer(c_int) :: nrow
    integer(c_int) :: ncol
    integer(c_int) :: inner_fp32_max
    real(c_double) :: inner_fp32_rtol
    integer(c_int) :: kernel_variant
  end type mf6hip_ims_settings

  public :: mf6hip_create, mf6hip_solve, mf6hip_destroy, mf6hip_last_error

  interface
    function mf6hip_create(neq, nja, ia, ja, settings, handle) result(rc) &
      bind(C, name='mf6hip_create')
      import :: c_int, c_ptr, mf6hip_ims_settings
      integer(c_int), value :: neq, nja
      integer(c_int), intent(in) :: ia(*), ja(*)
      type(mf6hip_ims_settings), intent(in) :: settings
      type(c_ptr), intent(out) :: handle
      integer(c_int) :: rc
    end function

    function mf6hip_solve(handle, kstp, kiter, amat, rhs, x, icnvg, in_iter, &
                          out) result(rc) bind(C, name='mf6hip_solve')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: handle
      integer(c_int), value :: kstp, kiter
      real(c_double), intent(in) :: amat(*), rhs(*)
      real(c_double), intent(inout) :: x(*)
      integer(c_int), intent(out) :: icnvg, in_iter
      type(c_ptr), value :: out
      integer(c_int) :: rc
    end function

    function mf6hip_destroy(handle) result(rc) bind(C, name='mf6hip_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: handle
      integer(c_int) :: rc
    end function

    function mf6hip_last_error(buf, buflen) result(rc) &
      bind(C, name='mf6hip_last_error')
      import :: c_int, c_char
      character(kind=c_char), intent(out) :: buf(*)
      integer(c_int), value :: buflen
      integer(c_int) :: rc
    end function
  end interface
end module mf6hip_module
