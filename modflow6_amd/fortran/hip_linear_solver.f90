! hip_linear_solver.f90 -- the reference-side plugin: a LinearSolverBaseType
! extension (src/Solution/LinearSolverBase.f90:17-61, the interface the
! PETSc solver implements in src/Solution/PETSc/PetscSolver.F90) whose solve
! runs the gfx950 IMS solver of libmf6hip.so through mf6hip_module.
!
! A MODFLOW 6 build adds this module next to LinearSolverFactory.F90 and
! selects it where sln_ar picks the linear solver
! (src/Solution/NumericalSolution.f90:929-932); sln_ls then calls
! solve(kiter, rhs, x, summary) and reads iteration_number / is_converged
! (NumericalSolution.f90:2601-2614) exactly as for the PETSc solver.  It
! compiles against the reference's own modules (KindModule, MatrixBase,
! SparseMatrix, VectorBase, ImsLinearSettings, ConvergenceSummary).
module HipLinearSolverModule
  use KindModule, only: DP, I4B
  use iso_c_binding, only: c_int, c_double, c_ptr, c_null_ptr, c_char, &
                           c_null_char, c_associated
  use LinearSolverBaseModule, only: LinearSolverBaseType
  use MatrixBaseModule, only: MatrixBaseType
  use VectorBaseModule, only: VectorBaseType
  use SparseMatrixModule, only: SparseMatrixType
  use ImsLinearSettingsModule, only: ImsLinearSettingsType
  use ConvergenceSummaryModule, only: ConvergenceSummaryType
  use mf6hip_module
  implicit none
  private
  public :: HipLinearSolverType

  type, extends(LinearSolverBaseType) :: HipLinearSolverType
    type(c_ptr) :: handle = c_null_ptr
    class(SparseMatrixType), pointer :: mat => null()
    type(ImsLinearSettingsType), pointer :: settings => null()
    integer(I4B) :: nlay = 0, nrow = 0, ncol = 0   !< DIS shape of the model
    integer(I4B) :: precision = MF6HIP_FP64        !< or MF6HIP_MIXED(_FP16)
    integer(I4B) :: preconditioner = 0             !< 0 ILU0/MILU0, 2 / 3 two-level
    integer(I4B) :: kstp = 1                       !< time step of the next solve
    integer(I4B) :: status = 0                     !< last libmf6hip status
  contains
    procedure :: initialize => hip_initialize
    procedure :: print_summary => hip_print_summary
    procedure :: solve => hip_solve
    procedure :: destroy => hip_destroy
    procedure :: create_matrix => hip_create_matrix
    procedure :: set_dis => hip_set_dis
    procedure :: message => hip_message
  end type HipLinearSolverType

contains

  !> DIS shape of the (single) model the solution holds; sln_ar would take
  !! it from the model's discretization before initialize
  subroutine hip_set_dis(this, nlay, nrow, ncol)
    class(HipLinearSolverType) :: this
    integer(I4B), intent(in) :: nlay, nrow, ncol
    this%nlay = nlay
    this%nrow = nrow
    this%ncol = ncol
  end subroutine hip_set_dis

  subroutine hip_initialize(this, matrix, linear_settings, convergence_summary)
    class(HipLinearSolverType) :: this
    class(MatrixBaseType), pointer :: matrix
    type(ImsLinearSettingsType), pointer :: linear_settings
    type(ConvergenceSummaryType), pointer :: convergence_summary
    type(mf6hip_ims_settings) :: s
    select type (matrix)
    class is (SparseMatrixType)
      this%mat => matrix
    end select
    this%settings => linear_settings
    this%nitermax = linear_settings%iter1
    this%iteration_number = 0
    this%is_converged = 0
    s%dvclose = linear_settings%dvclose
    s%rclose = linear_settings%rclose
    s%icnvgopt = linear_settings%icnvgopt
    s%iter1 = linear_settings%iter1
    s%ilinmeth = linear_settings%ilinmeth
    s%iscl = linear_settings%iscl
    s%iord = linear_settings%iord
    s%north = linear_settings%north
    s%relax = linear_settings%relax
    s%level = linear_settings%level
    s%droptol = linear_settings%droptol
    s%precision = this%precision
    s%ordering = MF6HIP_ORDER_RB
    s%nlay = this%nlay
    s%nrow = this%nrow
    s%ncol = this%ncol
    s%inner_fp32_max = 0
    s%inner_fp32_rtol = 0.0_c_double
    s%kernel_variant = 0
    s%preconditioner = this%preconditioner
    s%coarse_size = 0
    if (.not. associated(this%mat)) then
      this%status = 1
      return
    end if
    this%status = mf6hip_create(this%mat%nrow, this%mat%nja, this%mat%ia, &
                                this%mat%ja, s, this%handle)
  end subroutine hip_initialize

  !> sln_ls's call: solve A x = rhs in place (x holds the start heads), with
  !! the matrix values of the solution's SparseMatrixType
  subroutine hip_solve(this, kiter, rhs, x, cnvg_summary)
    class(HipLinearSolverType) :: this
    integer(I4B) :: kiter
    class(VectorBaseType), pointer :: rhs
    class(VectorBaseType), pointer :: x
    type(ConvergenceSummaryType) :: cnvg_summary
    real(DP), dimension(:), pointer, contiguous :: r, xx
    integer(c_int) :: icnvg, in_iter
    this%iteration_number = 0
    this%is_converged = 0
    if (.not. c_associated(this%handle)) then
      this%status = 1
      return
    end if
    r => rhs%get_array()
    xx => x%get_array()
    icnvg = 0
    in_iter = 0
    this%status = mf6hip_solve(this%handle, this%kstp, kiter, this%mat%amat, r, &
                               xx, icnvg, in_iter, c_null_ptr)
    this%iteration_number = in_iter
    this%is_converged = icnvg
  end subroutine hip_solve

  subroutine hip_print_summary(this)
    class(HipLinearSolverType) :: this
  end subroutine hip_print_summary

  subroutine hip_destroy(this)
    class(HipLinearSolverType) :: this
    integer(c_int) :: rc
    if (c_associated(this%handle)) rc = mf6hip_destroy(this%handle)
    this%handle = c_null_ptr
  end subroutine hip_destroy

  !> the system matrix type this solver works on (the reference's CSR
  !! SparseMatrixType; PetscSolverType%create_matrix returns its own type)
  function hip_create_matrix(this) result(matrix)
    class(HipLinearSolverType) :: this
    class(MatrixBaseType), pointer :: matrix
    class(SparseMatrixType), pointer :: spm
    allocate (spm)
    matrix => spm
  end function hip_create_matrix

  !> text of libmf6hip's last error
  function hip_message(this) result(msg)
    class(HipLinearSolverType) :: this
    character(len=512) :: msg
    character(kind=c_char) :: buf(512)
    integer(c_int) :: rc
    integer :: i
    rc = mf6hip_last_error(buf, 512_c_int)
    msg = ' '
    do i = 1, 512
      if (buf(i) == c_null_char) exit
      msg(i:i) = buf(i)
    end do
  end function hip_message

end module HipLinearSolverModule
