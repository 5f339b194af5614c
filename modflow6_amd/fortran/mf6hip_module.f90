! mf6hip_module.f90 -- ISO_C_BINDING interface to libmf6hip.so
! (include/mf6hip.h).  This is the module a Fortran host (MODFLOW 6's
! NumericalSolution / a LinearSolverBaseType extension) uses to call the
! gfx950 solver: every argument is c_int / c_double / c_ptr, CSR arrays are
! passed 1-based exactly as the memory manager holds them.
module mf6hip_module
  use iso_c_binding, only: c_int, c_double, c_ptr, c_char
  implicit none
  private

  integer(c_int), parameter, public :: MF6HIP_OK = 0
  integer(c_int), parameter, public :: MF6HIP_EINVAL = 1
  integer(c_int), parameter, public :: MF6HIP_EHIP = 2
  integer(c_int), parameter, public :: MF6HIP_EUNSUPPORTED = 3
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
    integer(c_int) :: preconditioner
    integer(c_int) :: coarse_size
  end type mf6hip_ims_settings

  integer(c_int), parameter, public :: MF6HIP_PKG_WEL = 1
  integer(c_int), parameter, public :: MF6HIP_PKG_DRN = 2
  integer(c_int), parameter, public :: MF6HIP_PKG_RCH = 3
  integer(c_int), parameter, public :: MF6HIP_PKG_CHD = 4
  integer(c_int), parameter, public :: MF6HIP_TERM_STO_SS = 0
  integer(c_int), parameter, public :: MF6HIP_TERM_STO_SY = 1
  integer(c_int), parameter, public :: MF6HIP_TERM_WEL = 2
  integer(c_int), parameter, public :: MF6HIP_TERM_DRN = 3
  integer(c_int), parameter, public :: MF6HIP_TERM_RCH = 4
  integer(c_int), parameter, public :: MF6HIP_TERM_CHD = 5

  !> mf6hip_gwf_model: DIS/NPF/STO arrays of one GWF model (c_loc of the
  !! memory-manager arrays; natural node order)
  type, bind(C), public :: mf6hip_gwf_model
    integer(c_int) :: nlay
    integer(c_int) :: nrow
    integer(c_int) :: ncol
    type(c_ptr) :: delr
    type(c_ptr) :: delc
    type(c_ptr) :: top
    type(c_ptr) :: bot
    type(c_ptr) :: k11
    type(c_ptr) :: k33
    type(c_ptr) :: icelltype
    type(c_ptr) :: ibound
    integer(c_int) :: inewton
    integer(c_int) :: inewtonur
    integer(c_int) :: ivarcv
    integer(c_int) :: idewatcv
    integer(c_int) :: icellavg
    integer(c_int) :: has_sto
    type(c_ptr) :: iconvert
    type(c_ptr) :: ss
    type(c_ptr) :: sy
    type(c_ptr) :: k22
    type(c_ptr) :: angle1
    type(c_ptr) :: angle2
    type(c_ptr) :: angle3
    integer(c_int) :: istor_coef
    integer(c_int) :: iconf_ss
    integer(c_int) :: iorig_ss
    integer(c_int) :: ithickstrt
    type(c_ptr) :: strt
  end type mf6hip_gwf_model

  !> mf6hip_ptc: pseudo-transient continuation control and state
  type, bind(C), public :: mf6hip_ptc
    integer(c_int) :: iptc
    integer(c_int) :: iallowptc
    integer(c_int) :: iptcopt
    integer(c_int) :: kstp
    integer(c_int) :: kper
    real(c_double) :: ptcf
    real(c_double) :: ptcdel0
    real(c_double) :: ptcexp
    real(c_double) :: ptcdel
    real(c_double) :: l2norm0
  end type mf6hip_ptc

  !> mf6hip_underrelax: NONLINEAR under-relaxation settings and Cooley state
  type, bind(C), public :: mf6hip_underrelax
    integer(c_int) :: nonmeth
    real(c_double) :: gamma
    real(c_double) :: theta
    real(c_double) :: akappa
    real(c_double) :: amomentum
    real(c_double) :: bigch
    real(c_double) :: bigchold
    real(c_double) :: relaxold
  end type mf6hip_underrelax

  !> mf6hip_package_options: WEL AUTO_FLOW_REDUCE, DRN AUXDEPTHNAME /
  !! DEV_CUBIC_SCALING (drndepth: c_loc of the depth column), AUXMULTNAME
  !! (rejected)
  type, bind(C), public :: mf6hip_package_options
    real(c_double) :: flowred
    integer(c_int) :: iflowred
    integer(c_int) :: idrndepth
    integer(c_int) :: icubic_scaling
    integer(c_int) :: has_auxmult
    type(c_ptr) :: drndepth
  end type mf6hip_package_options

  !> mf6hip_nonlinear: the IMS NONLINEAR block for the XMI-style outer loop
  type, bind(C), public :: mf6hip_nonlinear
    real(c_double) :: outer_dvclose
    integer(c_int) :: outer_maximum
    integer(c_int) :: nonmeth
    real(c_double) :: gamma
    real(c_double) :: theta
    real(c_double) :: akappa
    real(c_double) :: amomentum
    integer(c_int) :: backtracking_number
    real(c_double) :: backtracking_tolerance
    real(c_double) :: backtracking_reduction_factor
    real(c_double) :: backtracking_residual_limit
    integer(c_int) :: iallowptc
    integer(c_int) :: iptcopt
    real(c_double) :: ptcexp
    real(c_double) :: ptcdel0
  end type mf6hip_nonlinear

  !> mf6hip_xmi_iteration: what one mf6hip_xmi_solve call did
  type, bind(C), public :: mf6hip_xmi_iteration
    integer(c_int) :: kiter
    integer(c_int) :: inner
    integer(c_int) :: icnvg_inner
    real(c_double) :: hncg
    integer(c_int) :: lrch
    integer(c_int) :: converged
    integer(c_int) :: nur
    integer(c_int) :: nbt
    integer(c_int) :: iptc
  end type mf6hip_xmi_iteration

  public :: mf6hip_create, mf6hip_solve, mf6hip_destroy, mf6hip_last_error
  public :: mf6hip_gwf_set_package_ex, mf6hip_gwf_set_hfb, mf6hip_gwf_set_gnc, &
            mf6hip_xmi_set_nonlinear, &
            mf6hip_xmi_prepare_solve, mf6hip_xmi_solve, mf6hip_xmi_last_iteration, &
            mf6hip_xmi_finalize_solve
  public :: mf6hip_gwf_create, mf6hip_gwf_set_package, mf6hip_gwf_set_x, &
            mf6hip_gwf_get_x, mf6hip_gwf_advance, mf6hip_gwf_formulate, &
            mf6hip_gwf_residual_norms, mf6hip_gwf_backtrack, &
            mf6hip_gwf_linear_solve, mf6hip_gwf_outer_update, &
            mf6hip_gwf_budget, mf6hip_gwf_set_x_device, mf6hip_gwf_get_x_device, &
            mf6hip_gwf_flow_term

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

    function mf6hip_gwf_create(handle, model) result(rc) &
      bind(C, name='mf6hip_gwf_create')
      import :: c_int, c_ptr, mf6hip_gwf_model
      type(c_ptr), value :: handle
      type(mf6hip_gwf_model), intent(in) :: model
      integer(c_int) :: rc
    end function

    function mf6hip_gwf_set_package(handle, pkgtype, nbound, nodelist, v1, &
                                    v2) result(rc) &
      bind(C, name='mf6hip_gwf_set_package')
      import :: c_int, c_ptr
      type(c_ptr), value :: handle
      integer(c_int), value :: pkgtype, nbound
      type(c_ptr), value :: nodelist, v1, v2
      integer(c_int) :: rc
    end function

    function mf6hip_gwf_set_x(handle, x) result(rc) &
      bind(C, name='mf6hip_gwf_set_x')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: handle
      real(c_double), intent(in) :: x(*)
      integer(c_int) :: rc
    end function

    function mf6hip_gwf_get_x(handle, x) result(rc) &
      bind(C, name='mf6hip_gwf_get_x')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: handle
      real(c_double), intent(out) :: x(*)
      integer(c_int) :: rc
    end function

    function mf6hip_gwf_advance(handle, iss, delt) result(rc) &
      bind(C, name='mf6hip_gwf_advance')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: handle
      integer(c_int), value :: iss
      real(c_double), value :: delt
      integer(c_int) :: rc
    end function

    function mf6hip_gwf_formulate(handle, kiter, inewton, amat_out, rhs_out) &
      result(rc) bind(C, name='mf6hip_gwf_formulate')
      import :: c_int, c_ptr
      type(c_ptr), value :: handle
      integer(c_int), value :: kiter, inewton
      type(c_ptr), value :: amat_out, rhs_out
      integer(c_int) :: rc
    end function

    function mf6hip_gwf_budget(handle, flowja, terms, pct) result(rc) &
      bind(C, name='mf6hip_gwf_budget')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: handle
      real(c_double), intent(out) :: flowja(*)
      real(c_double), intent(out) :: terms(12)
      real(c_double), intent(out) :: pct
      integer(c_int) :: rc
    end function

    function mf6hip_gwf_flow_term(handle, term, values, count) result(rc) &
      bind(C, name='mf6hip_gwf_flow_term')
      import :: c_int, c_ptr
      type(c_ptr), value :: handle
      integer(c_int), value :: term
      type(c_ptr), value :: values
      integer(c_int), intent(out) :: count
      integer(c_int) :: rc
    end function

    function mf6hip_gwf_set_x_device(handle, x_dev) result(rc) &
      bind(C, name='mf6hip_gwf_set_x_device')
      import :: c_int, c_ptr
      type(c_ptr), value :: handle
      type(c_ptr), value :: x_dev
      integer(c_int) :: rc
    end function

    function mf6hip_gwf_get_x_device(handle, x_dev) result(rc) &
      bind(C, name='mf6hip_gwf_get_x_device')
      import :: c_int, c_ptr
      type(c_ptr), value :: handle
      type(c_ptr), value :: x_dev
      integer(c_int) :: rc
    end function

    function mf6hip_gwf_residual_norms(handle, l2norm, rvmax) result(rc) &
      bind(C, name='mf6hip_gwf_residual_norms')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: handle
      real(c_double), intent(out) :: l2norm, rvmax
      integer(c_int) :: rc
    end function

    function mf6hip_gwf_backtrack(handle, breduc, dvclose, btflag) result(rc) &
      bind(C, name='mf6hip_gwf_backtrack')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: handle
      real(c_double), value :: breduc, dvclose
      integer(c_int), intent(out) :: btflag
      integer(c_int) :: rc
    end function

    function mf6hip_gwf_linear_solve(handle, kstp, kiter, ptc, icnvg, in_iter, &
                                     hncg, lrch, out) result(rc) &
      bind(C, name='mf6hip_gwf_linear_solve')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: handle
      integer(c_int), value :: kstp, kiter
      type(c_ptr), value :: ptc
      integer(c_int), intent(out) :: icnvg, in_iter, lrch
      real(c_double), intent(out) :: hncg
      type(c_ptr), value :: out
      integer(c_int) :: rc
    end function

    function mf6hip_gwf_outer_update(handle, kiter, ur, inewtonur, dxold_max, &
                                     hncg, lrch) result(rc) &
      bind(C, name='mf6hip_gwf_outer_update')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: handle
      integer(c_int), value :: kiter
      type(c_ptr), value :: ur
      integer(c_int), intent(out) :: inewtonur, lrch
      real(c_double), intent(out) :: dxold_max, hncg
      integer(c_int) :: rc
    end function

    function mf6hip_gwf_set_package_ex(handle, pkgtype, nbound, nodelist, v1, &
                                       v2, opt) result(rc) &
      bind(C, name='mf6hip_gwf_set_package_ex')
      import :: c_int, c_double, c_ptr, mf6hip_package_options
      type(c_ptr), value :: handle
      integer(c_int), value :: pkgtype, nbound
      integer(c_int), intent(in) :: nodelist(*)
      real(c_double), intent(in) :: v1(*), v2(*)
      type(mf6hip_package_options), intent(in) :: opt
      integer(c_int) :: rc
    end function

    function mf6hip_gwf_set_hfb(handle, nhfb, noden, nodem, hydchr) result(rc) &
      bind(C, name='mf6hip_gwf_set_hfb')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: handle
      integer(c_int), value :: nhfb
      integer(c_int), intent(in) :: noden(*), nodem(*)
      real(c_double), intent(in) :: hydchr(*)
      integer(c_int) :: rc
    end function

    function mf6hip_gwf_set_gnc(handle, nexg, numjs, noden, nodem, nodesj, &
                                alphasj, implicit) result(rc) &
      bind(C, name='mf6hip_gwf_set_gnc')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: handle
      integer(c_int), value :: nexg, numjs, implicit
      integer(c_int), intent(in) :: noden(*), nodem(*), nodesj(*)
      real(c_double), intent(in) :: alphasj(*)
      integer(c_int) :: rc
    end function

    function mf6hip_xmi_set_nonlinear(handle, nl) result(rc) &
      bind(C, name='mf6hip_xmi_set_nonlinear')
      import :: c_int, c_ptr, mf6hip_nonlinear
      type(c_ptr), value :: handle
      type(mf6hip_nonlinear), intent(in) :: nl
      integer(c_int) :: rc
    end function

    function mf6hip_xmi_prepare_solve(handle, kper, kstp, iss, delt) result(rc) &
      bind(C, name='mf6hip_xmi_prepare_solve')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: handle
      integer(c_int), value :: kper, kstp, iss
      real(c_double), value :: delt
      integer(c_int) :: rc
    end function

    function mf6hip_xmi_solve(handle, has_converged) result(rc) &
      bind(C, name='mf6hip_xmi_solve')
      import :: c_int, c_ptr
      type(c_ptr), value :: handle
      integer(c_int), intent(out) :: has_converged
      integer(c_int) :: rc
    end function

    function mf6hip_xmi_last_iteration(handle, rec) result(rc) &
      bind(C, name='mf6hip_xmi_last_iteration')
      import :: c_int, c_ptr, mf6hip_xmi_iteration
      type(c_ptr), value :: handle
      type(mf6hip_xmi_iteration), intent(out) :: rec
      integer(c_int) :: rc
    end function

    function mf6hip_xmi_finalize_solve(handle, converged, outer_iterations) result(rc) &
      bind(C, name='mf6hip_xmi_finalize_solve')
      import :: c_int, c_ptr
      type(c_ptr), value :: handle
      integer(c_int), intent(out) :: converged, outer_iterations
      integer(c_int) :: rc
    end function
  end interface
end module mf6hip_module
