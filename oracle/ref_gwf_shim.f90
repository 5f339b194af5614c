! ref_gwf_shim.f90 -- C-callable entry points into the reference's GWF
! package and NumericalSolution procedures.
!
! TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together
! with the module closure of gwf.f90 / NumericalSolution.f90, read in place
! from /root/reference/src, into oracle/_ref/libmf6refgwf.so.  The tests use
! it to pin the C restatement in oracle/mf6_oracle.c bit for bit.
!
! No input system is involved: each entry point allocates the reference's
! own derived type (DisBaseType + ConnectionsType, GwfNpfType, GwfStoType,
! DrnType, NumericalSolutionType, GwfModelType, SparseMatrixType,
! SeqVectorType), fills the public members the called procedure reads, and
! calls that type-bound procedure.  Nothing here restates reference
! arithmetic; the shim only copies arrays in and out.
!
! Index convention: ia / ja / jas / isym / nodelist cross this boundary
! 1-based, as the reference stores them.
module mf6ref_gwf_shim
  use iso_c_binding, only: c_int, c_double
  use KindModule, only: DP, I4B
  use BaseDisModule, only: DisBaseType
  use ConnectionsModule, only: ConnectionsType
  use MatrixBaseModule, only: MatrixBaseType
  use SparseMatrixModule, only: SparseMatrixType
  use VectorBaseModule, only: VectorBaseType
  use SeqVectorModule, only: SeqVectorType
  use GwfNpfModule, only: GwfNpfType
  use GwfStoModule, only: GwfStoType
  use DrnModule, only: DrnType
  use BndModule, only: BndType
  use WelModule, only: wel_create
  use GwfHfbModule, only: GwfHfbType
  use GhostNodeModule, only: GhostNodeType
  use NumericalSolutionModule, only: NumericalSolutionType
  use GwfModule, only: GwfModelType
  use TdisModule, only: delt
  implicit none
  private

  type(DisBaseType), pointer, save :: dis => null()
  type(SparseMatrixType), pointer, save :: spm => null()
  class(MatrixBaseType), pointer, save :: mat => null()
  integer(I4B), allocatable, save :: idxglo(:)
  type(NumericalSolutionType), pointer, save :: sln => null()

contains

  ! ------------------------------------------------------------------
  ! Grid and solution matrix
  ! ------------------------------------------------------------------

  !> DisBaseType (DiscretizationBase.f90:30) with its ConnectionsType
  !! (Connections.f90:19) and a SparseMatrixType (SparseMatrix.f90:12) of the
  !! same CSR; idxglo is the identity (one model in the solution).
  subroutine ref_gwf_dis(nodes, nja, njas, ia, ja, jas, ihc, isym, hwva, &
                         cl1, cl2, top, bot, area) bind(C, name='ref_gwf_dis')
    integer(c_int), value :: nodes, nja, njas
    integer(c_int), intent(in) :: ia(nodes + 1), ja(nja), jas(nja), isym(nja)
    integer(c_int), intent(in) :: ihc(njas)
    real(c_double), intent(in) :: hwva(njas), cl1(njas), cl2(njas)
    real(c_double), intent(in) :: top(nodes), bot(nodes), area(nodes)
    integer(I4B) :: i
    call free_dis()
    allocate (dis)
    allocate (dis%nodes, dis%nja, dis%njas)
    dis%nodes = nodes
    dis%nja = nja
    dis%njas = njas
    allocate (dis%top(nodes), dis%bot(nodes), dis%area(nodes))
    dis%top = top
    dis%bot = bot
    dis%area = area
    allocate (dis%con)
    allocate (dis%con%nodes, dis%con%nja, dis%con%njas)
    dis%con%nodes = nodes
    dis%con%nja = nja
    dis%con%njas = njas
    allocate (dis%con%ia(nodes + 1), dis%con%ja(nja), dis%con%jas(nja), &
              dis%con%isym(nja), dis%con%mask(nja))
    dis%con%ia = ia
    dis%con%ja = ja
    dis%con%jas = jas
    dis%con%isym = isym
    dis%con%mask = 1
    allocate (dis%con%ihc(njas), dis%con%hwva(njas), dis%con%cl1(njas), &
              dis%con%cl2(njas))
    dis%con%ihc = ihc
    dis%con%hwva = hwva
    dis%con%cl1 = cl1
    dis%con%cl2 = cl2
    allocate (spm)
    allocate (spm%nrow, spm%ncol, spm%nja)
    spm%nrow = nodes
    spm%ncol = nodes
    spm%nja = nja
    allocate (spm%ia(nodes + 1), spm%ja(nja), spm%amat(nja))
    spm%ia = ia
    spm%ja = ja
    spm%amat = 0.0_DP
    mat => spm
    allocate (idxglo(nja))
    do i = 1, nja
      idxglo(i) = i
    end do
  end subroutine ref_gwf_dis

  subroutine free_dis()
    if (associated(dis)) then
      deallocate (dis%con%ia, dis%con%ja, dis%con%jas, dis%con%isym, &
                  dis%con%mask, dis%con%ihc, dis%con%hwva, dis%con%cl1, &
                  dis%con%cl2, dis%con%nodes, dis%con%nja, dis%con%njas)
      deallocate (dis%con)
      deallocate (dis%top, dis%bot, dis%area, dis%nodes, dis%nja, dis%njas)
      deallocate (dis)
      deallocate (spm%ia, spm%ja, spm%amat, spm%nrow, spm%ncol, spm%nja)
      deallocate (spm)
      deallocate (idxglo)
      mat => null()
    end if
  end subroutine free_dis

  ! ------------------------------------------------------------------
  ! NPF
  ! ------------------------------------------------------------------

  !> A GwfNpfType (gwf-npf.f90:34) with the members npf_cf / npf_fc / npf_fn
  !! / hy_eff / npf_nur read: no XT3D, no perched flow, no rewetting,
  !! isotropic horizontal K (ik22 = 0, no angles).
  function new_npf(inewton, ivarcv, idewatcv, icellavg, satomega, icelltype, &
                   ibound, k11, k33, condsat, sat) result(npf)
    integer(c_int), intent(in) :: inewton, ivarcv, idewatcv, icellavg
    real(c_double), intent(in) :: satomega
    integer(c_int), intent(in) :: icelltype(:), ibound(:)
    real(c_double), intent(in) :: k11(:), k33(:), condsat(:), sat(:)
    type(GwfNpfType), pointer :: npf
    integer(I4B) :: nodes
    nodes = dis%nodes
    allocate (npf)
    npf%dis => dis
    allocate (npf%inewton, npf%ixt3d, npf%iperched, npf%ivarcv, npf%idewatcv, &
              npf%icellavg, npf%satomega, npf%irewet, npf%iangle2, npf%ik22, &
              npf%iangle3, npf%iavgkeff)
    npf%inewton = inewton
    npf%ixt3d = 0
    npf%iperched = 0
    npf%ivarcv = ivarcv
    npf%idewatcv = idewatcv
    npf%icellavg = icellavg
    npf%satomega = satomega
    npf%irewet = 0
    npf%iangle2 = 0
    npf%iangle3 = 0
    npf%ik22 = 0
    npf%iavgkeff = 0
    allocate (npf%icelltype(nodes), npf%ibound(nodes), npf%k11(nodes), &
              npf%k22(nodes), npf%k33(nodes), npf%sat(nodes), &
              npf%condsat(dis%njas))
    npf%icelltype = icelltype
    npf%ibound = ibound
    npf%k11 = k11
    npf%k22 = k11
    npf%k33 = k33
    npf%sat = sat
    npf%condsat = condsat
  end function new_npf

  subroutine free_npf(npf)
    type(GwfNpfType), pointer :: npf
    deallocate (npf%icelltype, npf%ibound, npf%k11, npf%k22, npf%k33, &
                npf%sat, npf%condsat)
    if (associated(npf%ibotnode)) deallocate (npf%ibotnode)
    deallocate (npf%inewton, npf%ixt3d, npf%iperched, npf%ivarcv, &
                npf%idewatcv, npf%icellavg, npf%satomega, npf%irewet, &
                npf%iangle2, npf%ik22, npf%iangle3, npf%iavgkeff)
    deallocate (npf)
  end subroutine free_npf

  !> gwf-npf.f90:444 npf_cf (saturation of convertible cells)
  subroutine ref_npf_cf(nodes, inewton, satomega, icelltype, ibound, hnew, &
                        sat) bind(C, name='ref_npf_cf')
    integer(c_int), value :: nodes, inewton
    real(c_double), value :: satomega
    integer(c_int), intent(in) :: icelltype(nodes), ibound(nodes)
    real(c_double), intent(in) :: hnew(nodes)
    real(c_double), intent(inout) :: sat(nodes)
    type(GwfNpfType), pointer :: npf
    real(DP), allocatable :: h(:), kdum(:), cdum(:)
    allocate (h(nodes), kdum(nodes), cdum(dis%njas))
    h = hnew
    kdum = 1.0_DP
    cdum = 0.0_DP
    npf => new_npf(inewton, 0, 0, 0, satomega, icelltype, ibound, kdum, kdum, &
                   cdum, sat)
    call npf%npf_cf(1, nodes, h)
    sat = npf%sat
    call free_npf(npf)
  end subroutine ref_npf_cf

  !> gwf-npf.f90:474 npf_fc (iflag 0) or :578 npf_fn (iflag 1) on amat / rhs
  subroutine ref_npf_fill(iflag, nodes, nja, kiter, inewton, ivarcv, idewatcv, &
                          icellavg, satomega, icelltype, ibound, k11, k33, &
                          condsat, sat, hnew, amat, rhs) &
    bind(C, name='ref_npf_fill')
    integer(c_int), value :: iflag, nodes, nja, kiter, inewton, ivarcv, &
                             idewatcv, icellavg
    real(c_double), value :: satomega
    integer(c_int), intent(in) :: icelltype(nodes), ibound(nodes)
    real(c_double), intent(in) :: k11(nodes), k33(nodes), condsat(*), &
                                  sat(nodes), hnew(nodes)
    real(c_double), intent(inout) :: amat(nja), rhs(nodes)
    type(GwfNpfType), pointer :: npf
    real(DP), allocatable :: h(:), r(:)
    allocate (h(nodes), r(nodes))
    h = hnew
    r = rhs
    npf => new_npf(inewton, ivarcv, idewatcv, icellavg, satomega, icelltype, &
                   ibound, k11, k33, condsat(1:dis%njas), sat)
    spm%amat = amat
    if (iflag == 0) then
      call npf%npf_fc(kiter, mat, idxglo, r, h)
    else
      call npf%npf_fn(kiter, mat, idxglo, r, h)
    end if
    amat = spm%amat
    rhs = r
    call free_npf(npf)
  end subroutine ref_npf_fill

  !> gwf-npf.f90:745 npf_cq (flowja of every connection; diagonal untouched)
  subroutine ref_npf_cq(nodes, nja, inewton, ivarcv, idewatcv, icellavg, &
                        satomega, icelltype, ibound, k11, k33, condsat, sat, &
                        hnew, flowja) bind(C, name='ref_npf_cq')
    integer(c_int), value :: nodes, nja, inewton, ivarcv, idewatcv, icellavg
    real(c_double), value :: satomega
    integer(c_int), intent(in) :: icelltype(nodes), ibound(nodes)
    real(c_double), intent(in) :: k11(nodes), k33(nodes), condsat(*), &
                                  sat(nodes), hnew(nodes)
    real(c_double), intent(inout) :: flowja(nja)
    type(GwfNpfType), pointer :: npf
    real(DP), allocatable :: h(:), f(:)
    allocate (h(nodes), f(nja))
    h = hnew
    f = flowja
    npf => new_npf(inewton, ivarcv, idewatcv, icellavg, satomega, icelltype, &
                   ibound, k11, k33, condsat(1:dis%njas), sat)
    call npf%npf_cq(h, f)
    flowja = f
    call free_npf(npf)
  end subroutine ref_npf_cq

  !> gwf-npf.f90:705 npf_nur; returns inewtonur, updates dxmax / locmax
  !! (locmax 1-based)
  function ref_npf_nur(nodes, icelltype, ibound, ibotnode, x, xtemp, dx, &
                       dxmax, locmax) result(inur) bind(C, name='ref_npf_nur')
    integer(c_int), value :: nodes
    integer(c_int), intent(in) :: icelltype(nodes), ibound(nodes), &
                                  ibotnode(nodes)
    real(c_double), intent(inout) :: x(nodes), dx(nodes)
    real(c_double), intent(in) :: xtemp(nodes)
    real(c_double), intent(inout) :: dxmax
    integer(c_int), intent(inout) :: locmax
    integer(c_int) :: inur
    type(GwfNpfType), pointer :: npf
    real(DP), allocatable :: kdum(:), cdum(:), sdum(:)
    integer(I4B) :: inewtonur
    allocate (kdum(nodes), cdum(dis%njas), sdum(nodes))
    kdum = 1.0_DP
    cdum = 0.0_DP
    sdum = 1.0_DP
    npf => new_npf(1, 0, 0, 0, 1.0e-6_DP, icelltype, ibound, kdum, kdum, &
                   cdum, sdum)
    allocate (npf%ibotnode(nodes))
    npf%ibotnode = ibotnode
    inewtonur = 0
    call npf%npf_nur(nodes, x, xtemp, dx, inewtonur, dxmax, locmax)
    inur = inewtonur
    call free_npf(npf)
  end function ref_npf_nur

  ! ------------------------------------------------------------------
  ! STO
  ! ------------------------------------------------------------------

  !> A GwfStoType (gwf-sto.f90:30) with the members sto_fc / sto_fn /
  !! sto_cq read (no TVS, no integration of storage changes)
  function new_sto(deltv, iss, istor_coef, iconf_ss, iorig_ss, satomega, &
                   iconvert, ibound, ss, sy) result(sto)
    real(c_double), intent(in) :: deltv, satomega
    integer(c_int), intent(in) :: iss, istor_coef, iconf_ss, iorig_ss
    integer(c_int), intent(in) :: iconvert(:), ibound(:)
    real(c_double), intent(in) :: ss(:), sy(:)
    type(GwfStoType), pointer :: sto
    integer(I4B) :: nodes
    nodes = dis%nodes
    if (.not. associated(delt)) allocate (delt)
    delt = deltv
    allocate (sto)
    sto%dis => dis
    allocate (sto%iss, sto%istor_coef, sto%iconf_ss, sto%iorig_ss, &
              sto%satomega, sto%integratechanges, sto%iusesy)
    sto%iss = iss
    sto%istor_coef = istor_coef
    sto%iconf_ss = iconf_ss
    sto%iorig_ss = iorig_ss
    sto%satomega = satomega
    sto%integratechanges = 0
    sto%iusesy = 0
    if (any(iconvert /= 0)) sto%iusesy = 1
    allocate (sto%iconvert(nodes), sto%ibound(nodes), sto%ss(nodes), &
              sto%sy(nodes), sto%strgss(nodes), sto%strgsy(nodes))
    sto%iconvert = iconvert
    sto%ibound = ibound
    sto%ss = ss
    sto%sy = sy
    sto%strgss = 0.0_DP
    sto%strgsy = 0.0_DP
  end function new_sto

  subroutine free_sto(sto)
    type(GwfStoType), pointer :: sto
    deallocate (sto%iconvert, sto%ibound, sto%ss, sto%sy, sto%strgss, &
                sto%strgsy)
    deallocate (sto%iss, sto%istor_coef, sto%iconf_ss, sto%iorig_ss, &
                sto%satomega, sto%integratechanges, sto%iusesy)
    deallocate (sto)
  end subroutine free_sto

  !> gwf-sto.f90:226 sto_fc (iflag 0) or :353 sto_fn (iflag 1)
  subroutine ref_sto_fill(iflag, nodes, nja, kiter, deltv, iss, istor_coef, &
                          iconf_ss, iorig_ss, satomega, iconvert, ibound, ss, &
                          sy, hold, hnew, amat, rhs) bind(C, name='ref_sto_fill')
    integer(c_int), value :: iflag, nodes, nja, kiter, iss, istor_coef, &
                             iconf_ss, iorig_ss
    real(c_double), value :: deltv, satomega
    integer(c_int), intent(in) :: iconvert(nodes), ibound(nodes)
    real(c_double), intent(in) :: ss(nodes), sy(nodes), hold(nodes), &
                                  hnew(nodes)
    real(c_double), intent(inout) :: amat(nja), rhs(nodes)
    type(GwfStoType), pointer :: sto
    real(DP), allocatable :: r(:)
    allocate (r(nodes))
    r = rhs
    sto => new_sto(deltv, iss, istor_coef, iconf_ss, iorig_ss, satomega, &
                   iconvert, ibound, ss, sy)
    spm%amat = amat
    if (iflag == 0) then
      call sto%sto_fc(kiter, hold, hnew, mat, idxglo, r)
    else
      call sto%sto_fn(kiter, hold, hnew, mat, idxglo, r)
    end if
    amat = spm%amat
    rhs = r
    call free_sto(sto)
  end subroutine ref_sto_fill

  !> gwf-sto.f90:447 sto_cq: per-cell STO-SS / STO-SY rates, added to the
  !! flowja diagonal
  subroutine ref_sto_cq(nodes, nja, deltv, iss, istor_coef, iconf_ss, iorig_ss, &
                        satomega, iconvert, ibound, ss, sy, hold, hnew, flowja, &
                        strgss, strgsy) bind(C, name='ref_sto_cq')
    integer(c_int), value :: nodes, nja, iss, istor_coef, iconf_ss, iorig_ss
    real(c_double), value :: deltv, satomega
    integer(c_int), intent(in) :: iconvert(nodes), ibound(nodes)
    real(c_double), intent(in) :: ss(nodes), sy(nodes), hold(nodes), &
                                  hnew(nodes)
    real(c_double), intent(inout) :: flowja(nja), strgss(nodes), &
                                    strgsy(nodes)
    type(GwfStoType), pointer :: sto
    real(DP), allocatable :: f(:)
    allocate (f(nja))
    f = flowja
    sto => new_sto(deltv, iss, istor_coef, iconf_ss, iorig_ss, satomega, iconvert, &
                   ibound, ss, sy)
    call sto%sto_cq(f, hnew, hold)
    flowja = f
    strgss = sto%strgss
    strgsy = sto%strgsy
    call free_sto(sto)
  end subroutine ref_sto_cq

  ! ------------------------------------------------------------------
  ! DRN
  ! ------------------------------------------------------------------

  !> gwf-drn.f90:340 drn_cf, then :377 drn_fc and :418 drn_fn on amat / rhs,
  !! and BoundaryPackage.f90:545 bnd_cq (simvals, flowja diagonal).
  !! idepth /= 0 gives the drains a depth auxiliary (auxvar column 1);
  !! icubic selects cubic conductance scaling.  hcof / brhs return the
  !! package's own terms.
  subroutine ref_drn(nodes, nja, nbound, nodelist, elev, cond, idepth, depth, &
                     icubic, ibound, xnew, hcof, brhs, amat, rhs, flowja, &
                     simvals) bind(C, name='ref_drn')
    integer(c_int), value :: nodes, nja, nbound, idepth, icubic
    integer(c_int), intent(in) :: nodelist(nbound), ibound(nodes)
    real(c_double), intent(in) :: elev(nbound), cond(nbound), depth(nbound), &
                                  xnew(nodes)
    real(c_double), intent(inout) :: hcof(nbound), brhs(nbound)
    real(c_double), intent(inout) :: amat(nja), rhs(nodes), flowja(nja), &
                                    simvals(nbound)
    type(DrnType), pointer :: drn
    real(DP), allocatable :: r(:), f(:)
    allocate (r(nodes), f(nja))
    r = rhs
    f = flowja
    allocate (drn)
    drn%dis => dis
    allocate (drn%nbound, drn%iauxmultcol, drn%iauxddrncol, &
              drn%icubic_scaling, drn%imover, drn%naux)
    drn%nbound = nbound
    drn%iauxmultcol = 0
    drn%iauxddrncol = 0
    if (idepth /= 0) drn%iauxddrncol = 1
    drn%icubic_scaling = icubic
    drn%imover = 0
    drn%naux = 1
    allocate (drn%nodelist(nbound), drn%elev(nbound), drn%cond(nbound), &
              drn%hcof(nbound), drn%rhs(nbound), drn%auxvar(1, nbound), &
              drn%ibound(nodes), drn%xnew(nodes), drn%simvals(nbound))
    drn%nodelist = nodelist
    drn%elev = elev
    drn%cond = cond
    drn%auxvar(1, :) = depth
    drn%hcof = 0.0_DP
    drn%rhs = 0.0_DP
    drn%ibound = ibound
    drn%xnew = xnew
    call drn%bnd_cf()
    hcof = drn%hcof
    brhs = drn%rhs
    spm%amat = amat
    call drn%bnd_fc(r, dis%con%ia, idxglo, mat)
    call drn%bnd_fn(r, dis%con%ia, idxglo, mat)
    amat = spm%amat
    rhs = r
    drn%simvals = 0.0_DP
    call drn%bnd_cq(xnew, f)
    flowja = f
    simvals = drn%simvals
    deallocate (drn%nodelist, drn%elev, drn%cond, drn%hcof, drn%rhs, &
                drn%auxvar, drn%ibound, drn%xnew, drn%simvals)
    deallocate (drn%nbound, drn%iauxmultcol, drn%iauxddrncol, &
                drn%icubic_scaling, drn%imover, drn%naux)
    deallocate (drn)
  end subroutine ref_drn

  ! ------------------------------------------------------------------
  ! WEL
  ! ------------------------------------------------------------------

  !> gwf-wel.f90:296 wel_cf, then :340 wel_fc and :378 wel_fn on amat /
  !! rhs, and BoundaryPackage.f90:545 bnd_cq (simvals, flowja diagonal).
  !! iflowred / flowred: AUTO_FLOW_REDUCE as wel_options leaves them;
  !! icelltype the NPF cell types the package points at.  WelType is
  !! private to its module, so the package is made by wel_create (:73) and
  !! filled through the memory manager: the model, NPF (ICELLTYPE) and
  !! input-context variables its allocate_scalars / allocate_arrays
  !! (:135-178) read are created first, under a new model name per call.
  subroutine ref_wel(nodes, nja, nbound, nodelist, q, iflowred, flowred, &
                     icelltype, ibound, xnew, hcof, brhs, amat, rhs, flowja, &
                     simvals) bind(C, name='ref_wel')
    use MemoryManagerModule, only: mem_allocate, mem_setptr
    use MemoryHelperModule, only: create_mem_path
    use SimVariablesModule, only: idm_context
    use CharacterStringModule, only: CharacterStringType
    use ConstantsModule, only: LENBOUNDNAME, LENMODELNAME, LENMEMPATH
    integer(c_int), value :: nodes, nja, nbound, iflowred
    real(c_double), value :: flowred
    integer(c_int), intent(in) :: nodelist(nbound), ibound(nodes)
    integer(c_int), intent(in) :: icelltype(nodes)
    real(c_double), intent(in) :: q(nbound), xnew(nodes)
    real(c_double), intent(inout) :: hcof(nbound), brhs(nbound)
    real(c_double), intent(inout) :: amat(nja), rhs(nodes), flowja(nja), &
                                    simvals(nbound)
    class(BndType), pointer :: pk
    integer(I4B), pointer :: iv
    real(DP), pointer :: rv
    integer(I4B), dimension(:, :), pointer, contiguous :: cid
    type(CharacterStringType), dimension(:), pointer, contiguous :: bnm
    real(DP), dimension(:, :), pointer, contiguous :: aux
    real(DP), dimension(:), pointer, contiguous :: qp
    integer(I4B), dimension(:), pointer, contiguous :: ict, ib
    real(DP), dimension(:), pointer, contiguous :: xn
    character(len=LENMODELNAME) :: mname
    character(len=LENMEMPATH) :: mpath, inpath
    integer(I4B), save :: ncall = 0
    real(DP), allocatable :: r(:), f(:)
    allocate (r(nodes), f(nja))
    r = rhs
    f = flowja
    ncall = ncall + 1
    write (mname, '(a,i0)') 'SHIMWEL', ncall
    mpath = create_mem_path(mname)
    inpath = create_mem_path(mname, 'WEL', idm_context)
    call mem_allocate(iv, 'INEWTON', mpath)
    iv = 1
    call mem_allocate(iv, 'IPRPAK', mpath)
    iv = 0
    call mem_allocate(iv, 'IPRFLOW', mpath)
    iv = 0
    call mem_allocate(iv, 'IPAKCB', mpath)
    iv = 0
    call mem_allocate(iv, 'IPER', inpath)
    iv = 1
    call mem_allocate(cid, 3, nbound, 'CELLID', inpath)
    call mem_allocate(bnm, LENBOUNDNAME, nbound, 'BOUNDNAME', inpath)
    call mem_allocate(aux, 0, nbound, 'AUXVAR', inpath)
    call mem_allocate(qp, nbound, 'Q', inpath)
    call mem_allocate(ict, nodes, 'ICELLTYPE', create_mem_path(mname, 'NPF'))
    ict = icelltype
    call wel_create(pk, 1, 1, 0, 0, mname, 'WEL', inpath)
    pk%maxbound = nbound
    pk%ncolbnd = 1
    call pk%allocate_arrays()
    pk%nbound = nbound
    pk%nodelist(1:nbound) = nodelist
    qp = q
    call mem_setptr(iv, 'IFLOWRED', pk%memoryPath)
    iv = iflowred
    call mem_setptr(rv, 'FLOWRED', pk%memoryPath)
    rv = flowred
    allocate (ib(nodes), xn(nodes))
    ib = ibound
    xn = xnew
    pk%ibound => ib
    pk%xnew => xn
    pk%dis => dis
    call pk%bnd_cf()
    hcof = pk%hcof(1:nbound)
    brhs = pk%rhs(1:nbound)
    spm%amat = amat
    call pk%bnd_fc(r, dis%con%ia, idxglo, mat)
    call pk%bnd_fn(r, dis%con%ia, idxglo, mat)
    amat = spm%amat
    rhs = r
    pk%simvals = 0.0_DP
    call pk%bnd_cq(xnew, f)
    flowja = f
    simvals = pk%simvals(1:nbound)
    nullify (pk%ibound, pk%xnew)
    deallocate (ib, xn)
  end subroutine ref_wel

  ! ------------------------------------------------------------------
  ! HGeoUtil
  ! ------------------------------------------------------------------

  !> HGeoUtil.f90:29 hyeff (the effective conductivity along a direction)
  function ref_hyeff(k11, k22, k33, ang1, ang2, ang3, vg1, vg2, vg3, &
                     iavgmeth) result(k) bind(C, name='ref_hyeff')
    use HGeoUtilModule, only: hyeff
    real(c_double), value :: k11, k22, k33, ang1, ang2, ang3, vg1, vg2, vg3
    integer(c_int), value :: iavgmeth
    real(c_double) :: k
    k = hyeff(k11, k22, k33, ang1, ang2, ang3, vg1, vg2, vg3, iavgmeth)
  end function ref_hyeff

  !> Constants.f90:130 DPIO180 times a value (the NPF angle conversion)
  function ref_dpio180(a) result(r) bind(C, name='ref_dpio180')
    use ConstantsModule, only: DPIO180
    real(c_double), value :: a
    real(c_double) :: r
    r = a * DPIO180
  end function ref_dpio180

  ! ------------------------------------------------------------------
  ! GNC (one model, explicit)
  ! ------------------------------------------------------------------

  !> GhostNode.f90:172 gnc_mc (matrix positions), :280 gnc_fc with
  !! gnc_fmsav :251, under NEWTON :340 gnc_fn, then :478 gnc_cq.  nodesj /
  !! alphasj hold numjs entries per GNC (node 0: none); cond returns the
  !! conductances gnc_fmsav saved.
  subroutine ref_gnc(nodes, nja, nexg, numjs, noden, nodem, nodesj, alphasj, &
                     ibound, x, inewton, ivarcv, icelltype, condsat, amat, &
                     rhs, cond, flowja) bind(C, name='ref_gnc')
    integer(c_int), value :: nodes, nja, nexg, numjs, inewton, ivarcv
    integer(c_int), intent(in) :: noden(nexg), nodem(nexg), nodesj(numjs, nexg)
    real(c_double), intent(in) :: alphasj(numjs, nexg)
    integer(c_int), intent(in) :: ibound(nodes), icelltype(nodes)
    real(c_double), intent(in) :: x(nodes), condsat(*)
    real(c_double), intent(inout) :: amat(nja), rhs(nodes), cond(nexg), &
                                    flowja(nja)
    type(GhostNodeType), pointer :: gnc
    type(GwfModelType), pointer :: gwf
    integer(I4B), allocatable :: ict(:)
    real(DP), allocatable :: f(:)
    allocate (gwf)
    allocate (gwf%moffset)
    gwf%moffset = 0
    gwf%dis => dis
    allocate (gwf%ibound(nodes), gwf%x(nodes), gwf%rhs(nodes))
    gwf%ibound = ibound
    gwf%x = x
    gwf%rhs = rhs
    allocate (gnc)
    allocate (gnc%smgnc, gnc%implicit, gnc%i2kn, gnc%nexg, gnc%numjs)
    gnc%smgnc = .true.
    gnc%implicit = .false.
    gnc%i2kn = .false.
    gnc%nexg = nexg
    gnc%numjs = numjs
    gnc%m1 => gwf
    gnc%m2 => gwf
    allocate (gnc%nodem1(nexg), gnc%nodem2(nexg), gnc%nodesj(numjs, nexg), &
              gnc%alphasj(numjs, nexg), gnc%cond(nexg), gnc%idxglo(nexg), &
              gnc%idxsymglo(nexg), gnc%idiagn(nexg), gnc%idiagm(nexg), &
              gnc%jposinrown(numjs, nexg), gnc%jposinrowm(numjs, nexg))
    gnc%nodem1 = noden
    gnc%nodem2 = nodem
    gnc%nodesj = nodesj
    gnc%alphasj = alphasj
    gnc%cond = 0.0_DP
    call gnc%gnc_mc(mat)
    spm%amat = amat
    call gnc%gnc_fc(1, mat)
    if (inewton /= 0) then
      allocate (ict(nodes))
      ict = icelltype
      call gnc%gnc_fn(1, mat, condsat(1:dis%njas), ivarcv_opt=ivarcv, &
                      ictm1_opt=ict, ictm2_opt=ict)
      deallocate (ict)
    end if
    amat = spm%amat
    rhs = gwf%rhs
    cond = gnc%cond
    allocate (f(nja))
    f = flowja
    call gnc%gnc_cq(f)
    flowja = f
    deallocate (gnc%nodem1, gnc%nodem2, gnc%nodesj, gnc%alphasj, gnc%cond, &
                gnc%idxglo, gnc%idxsymglo, gnc%idiagn, gnc%idiagm, &
                gnc%jposinrown, gnc%jposinrowm)
    deallocate (gnc%smgnc, gnc%implicit, gnc%i2kn, gnc%nexg, gnc%numjs)
    deallocate (gnc)
    deallocate (gwf%ibound, gwf%x, gwf%rhs, gwf%moffset)
    deallocate (gwf)
  end subroutine ref_gnc

  ! ------------------------------------------------------------------
  ! HFB
  ! ------------------------------------------------------------------

  !> gwf-hfb.f90:211 hfb_fc (no XT3D, no viscosity) on amat: barriers between
  !! noden / nodem (1-based) at the ja positions idxloc (1-based, as
  !! check_data stores them); condsav returns what hfb_fc saves for hfb_cq.
  subroutine ref_hfb_fc(nodes, nja, nhfb, inewton, noden, nodem, idxloc, &
                        hydchr, icelltype, ibound, hnew, amat, condsav) &
    bind(C, name='ref_hfb_fc')
    integer(c_int), value :: nodes, nja, nhfb, inewton
    integer(c_int), intent(in) :: noden(nhfb), nodem(nhfb), idxloc(nhfb)
    real(c_double), intent(in) :: hydchr(nhfb)
    integer(c_int), intent(in) :: icelltype(nodes), ibound(nodes)
    real(c_double), intent(in) :: hnew(nodes)
    real(c_double), intent(inout) :: amat(nja)
    real(c_double), intent(out) :: condsav(nhfb)
    type(GwfHfbType), pointer :: hfb
    real(DP), allocatable :: r(:), h(:)
    allocate (r(nodes), h(nodes))
    r = 0.0_DP
    h = hnew
    allocate (hfb)
    hfb%dis => dis
    allocate (hfb%inewton, hfb%nhfb, hfb%ivsc)
    hfb%inewton = inewton
    hfb%nhfb = nhfb
    hfb%ivsc = 0
    allocate (hfb%xt3d)
    allocate (hfb%noden(nhfb), hfb%nodem(nhfb), hfb%idxloc(nhfb), &
              hfb%hydchr(nhfb), hfb%condsav(nhfb))
    hfb%noden = noden
    hfb%nodem = nodem
    hfb%idxloc = idxloc
    hfb%hydchr = hydchr
    hfb%condsav = 0.0_DP
    allocate (hfb%ibound(nodes), hfb%icelltype(nodes))
    hfb%ibound = ibound
    hfb%icelltype = icelltype
    hfb%ihc => dis%con%ihc
    hfb%ia => dis%con%ia
    hfb%ja => dis%con%ja
    hfb%jas => dis%con%jas
    hfb%isym => dis%con%isym
    hfb%hwva => dis%con%hwva
    hfb%top => dis%top
    hfb%bot => dis%bot
    spm%amat = amat
    call hfb%hfb_fc(1, mat, idxglo, r, h)
    amat = spm%amat
    condsav = hfb%condsav
    deallocate (hfb%noden, hfb%nodem, hfb%idxloc, hfb%hydchr, hfb%condsav, &
                hfb%ibound, hfb%icelltype)
    deallocate (hfb%xt3d)
    deallocate (hfb%inewton, hfb%nhfb, hfb%ivsc)
    deallocate (hfb)
  end subroutine ref_hfb_fc

  !> gwf-hfb.f90:355 hfb_cq without NEWTON: the flows across barriers next
  !! to a convertible cell from the conductances hfb_fc saved (condsav)
  subroutine ref_hfb_cq(nodes, nja, nhfb, noden, nodem, idxloc, icelltype, &
                        ibound, hnew, condsav, flowja) bind(C, name='ref_hfb_cq')
    integer(c_int), value :: nodes, nja, nhfb
    integer(c_int), intent(in) :: noden(nhfb), nodem(nhfb), idxloc(nhfb)
    integer(c_int), intent(in) :: icelltype(nodes), ibound(nodes)
    real(c_double), intent(in) :: hnew(nodes), condsav(nhfb)
    real(c_double), intent(inout) :: flowja(nja)
    type(GwfHfbType), pointer :: hfb
    real(DP), allocatable :: h(:), f(:)
    allocate (h(nodes), f(nja))
    h = hnew
    f = flowja
    allocate (hfb)
    hfb%dis => dis
    allocate (hfb%inewton, hfb%nhfb, hfb%ivsc)
    hfb%inewton = 0
    hfb%nhfb = nhfb
    hfb%ivsc = 0
    allocate (hfb%xt3d)
    allocate (hfb%noden(nhfb), hfb%nodem(nhfb), hfb%idxloc(nhfb), hfb%condsav(nhfb))
    hfb%noden = noden
    hfb%nodem = nodem
    hfb%idxloc = idxloc
    hfb%condsav = condsav
    allocate (hfb%ibound(nodes), hfb%icelltype(nodes))
    hfb%ibound = ibound
    hfb%icelltype = icelltype
    call hfb%hfb_cq(h, f)
    flowja = f
    deallocate (hfb%noden, hfb%nodem, hfb%idxloc, hfb%condsav, hfb%ibound, &
                hfb%icelltype)
    deallocate (hfb%xt3d)
    deallocate (hfb%inewton, hfb%nhfb, hfb%ivsc)
    deallocate (hfb)
  end subroutine ref_hfb_cq

  ! ------------------------------------------------------------------
  ! NumericalSolution: under-relaxation and backtracking
  ! ------------------------------------------------------------------

  !> A NumericalSolutionType (NumericalSolution.f90:57) holding the
  !! under-relaxation and backtracking members, kept between calls so the
  !! Cooley and delta-bar-delta recurrences run over several outer
  !! iterations.
  subroutine ref_sln_new(neq, nonmeth, gamma, theta, akappa, amomentum, &
                         breduc, dvclose) bind(C, name='ref_sln_new')
    integer(c_int), value :: neq, nonmeth
    real(c_double), value :: gamma, theta, akappa, amomentum, breduc, dvclose
    call ref_sln_free()
    allocate (sln)
    allocate (sln%neq, sln%nonmeth, sln%gamma, sln%theta, sln%akappa, &
              sln%amomentum, sln%bigch, sln%bigchold, sln%relaxold, &
              sln%breduc, sln%dvclose)
    sln%neq = neq
    sln%nonmeth = nonmeth
    sln%gamma = gamma
    sln%theta = theta
    sln%akappa = akappa
    sln%amomentum = amomentum
    sln%bigch = 0.0_DP
    sln%bigchold = 0.0_DP
    sln%relaxold = 0.0_DP
    sln%breduc = breduc
    sln%dvclose = dvclose
    allocate (sln%dxold(neq), sln%wsave(neq), sln%hchold(neq), &
              sln%deold(neq), sln%x(neq), sln%xtemp(neq), sln%active(neq))
    sln%dxold = 0.0_DP
    sln%wsave = 0.0_DP
    sln%hchold = 0.0_DP
    sln%deold = 0.0_DP
  end subroutine ref_sln_new

  subroutine ref_sln_free() bind(C, name='ref_sln_free')
    if (associated(sln)) then
      deallocate (sln%dxold, sln%wsave, sln%hchold, sln%deold, sln%x, &
                  sln%xtemp, sln%active)
      deallocate (sln%neq, sln%nonmeth, sln%gamma, sln%theta, sln%akappa, &
                  sln%amomentum, sln%bigch, sln%bigchold, sln%relaxold, &
                  sln%breduc, sln%dvclose)
      deallocate (sln)
    end if
  end subroutine ref_sln_free

  !> NumericalSolution.f90:2989 sln_underrelax; returns the state after the
  !! call (bigch, bigchold, relaxold and the four state vectors)
  subroutine ref_sln_underrelax(kiter, bigch, neq, active, x, xtemp, scal, &
                                dxold, wsave, hchold, deold) &
    bind(C, name='ref_sln_underrelax')
    integer(c_int), value :: kiter, neq
    real(c_double), value :: bigch
    integer(c_int), intent(in) :: active(neq)
    real(c_double), intent(inout) :: x(neq)
    real(c_double), intent(in) :: xtemp(neq)
    real(c_double), intent(inout) :: scal(3)
    real(c_double), intent(inout) :: dxold(neq), wsave(neq), hchold(neq), &
                                    deold(neq)
    integer(I4B), allocatable :: act(:)
    real(DP), allocatable :: xx(:), xt(:)
    allocate (act(neq), xx(neq), xt(neq))
    act = active
    xx = x
    xt = xtemp
    call sln%sln_underrelax(kiter, bigch, neq, act, xx, xt)
    x = xx
    scal(1) = sln%bigch
    scal(2) = sln%bigchold
    scal(3) = sln%relaxold
    dxold = sln%dxold
    wsave = sln%wsave
    hchold = sln%hchold
    deold = sln%deold
  end subroutine ref_sln_underrelax

  !> NumericalSolution.f90:2783 sln_backtracking_xupdate (get_backtracking_flag
  !! :2800 and apply_backtracking :2830); returns bt_flag
  function ref_sln_bt_xupdate(neq, active, x, xtemp) result(flag) &
    bind(C, name='ref_sln_bt_xupdate')
    integer(c_int), value :: neq
    integer(c_int), intent(in) :: active(neq)
    real(c_double), intent(inout) :: x(neq)
    real(c_double), intent(in) :: xtemp(neq)
    integer(c_int) :: flag
    integer(I4B) :: btflag
    sln%active = active
    sln%x = x
    sln%xtemp = xtemp
    btflag = 0
    call sln%sln_backtracking_xupdate(btflag)
    x = sln%x
    flag = btflag
  end function ref_sln_bt_xupdate

  ! ------------------------------------------------------------------
  ! Residual and pseudo-transient continuation
  ! ------------------------------------------------------------------

  !> NumericalSolution.f90:2966 sln_calc_residual's steps (SparseMatrix
  !! multiply, SeqVector axpy, inactive rows zeroed) and SeqVector norm2,
  !! then gwf.f90:625 gwf_ptc on that residual.  Returns the l2 norm in
  !! scal(1), ptcf in scal(2) and iptc.
  function ref_gwf_ptc(nodes, nja, amat, x, rhs, active, iss, inewton, &
                       npf_inewton, ibound, deltv, resid, scal) result(iptc) &
    bind(C, name='ref_gwf_ptc')
    integer(c_int), value :: nodes, nja, iss, inewton, npf_inewton
    real(c_double), intent(in) :: amat(nja), x(nodes), rhs(nodes)
    integer(c_int), intent(in) :: active(nodes), ibound(nodes)
    real(c_double), value :: deltv
    real(c_double), intent(inout) :: resid(nodes), scal(2)
    integer(c_int) :: iptc
    type(SeqVectorType), pointer :: vx, vb, vr
    class(VectorBaseType), pointer :: pvx, pvb, pvr
    type(GwfModelType), pointer :: gwf
    type(GwfNpfType), pointer :: npf
    real(DP), allocatable :: kdum(:), cdum(:), sdum(:)
    integer(I4B) :: n, iptcv
    real(DP) :: ptcf
    allocate (vx, vb, vr)
    call vx%create(nodes)
    call vb%create(nodes)
    call vr%create(nodes)
    vx%array = x
    vb%array = rhs
    pvx => vx
    pvb => vb
    pvr => vr
    spm%amat = amat
    call spm%multiply(pvx, pvr)
    call vr%axpy(-1.0_DP, pvb)
    do n = 1, nodes
      if (active(n) < 1) call vr%set_value_local(n, 0.0_DP)
    end do
    resid = vr%array
    scal(1) = vr%norm2()
    if (.not. associated(delt)) allocate (delt)
    delt = deltv
    allocate (kdum(nodes), cdum(dis%njas), sdum(nodes))
    kdum = 1.0_DP
    cdum = 0.0_DP
    sdum = 1.0_DP
    npf => new_npf(npf_inewton, 0, 0, 0, 1.0e-6_DP, ibound, ibound, kdum, &
                   kdum, cdum, sdum)
    allocate (gwf)
    allocate (gwf%iss, gwf%inewton)
    gwf%iss = iss
    gwf%inewton = inewton
    gwf%npf => npf
    gwf%dis => dis
    iptcv = 0
    ptcf = 0.0_DP
    call gwf%model_ptc(pvr, iptcv, ptcf)
    iptc = iptcv
    scal(2) = ptcf
    deallocate (gwf%iss, gwf%inewton)
    deallocate (gwf)
    call free_npf(npf)
    call vx%destroy()
    call vb%destroy()
    call vr%destroy()
    deallocate (vx, vb, vr)
  end function ref_gwf_ptc

end module mf6ref_gwf_shim
