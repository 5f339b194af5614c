! ref_shim.f90 -- C-callable entry points into the reference's own Fortran.
!
! TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together
! with the reference's IMS / GWF-utility sources, read in place from
! /root/reference/src, into oracle/_ref/libmf6ref.so.  The tests use it to
! pin the C restatement in oracle/mf6_oracle.c bit for bit.  Nothing here
! re-implements reference arithmetic: every routine only marshals arguments
! and calls the reference procedure named in its comment.
!
! Index convention: arrays cross this boundary 1-based, exactly as the
! reference stores them (ia/ja/iapc/japc hold Fortran positions).
module mf6ref_shim
  use iso_c_binding, only: c_int, c_double, c_char
  use KindModule, only: DP, I4B
  use IMSLinearBaseModule, only: ims_base_pccrs, ims_base_pcu, &
                                 ims_base_ilu0a, ims_base_cg, ims_base_bcgs, &
                                 ims_base_residual, ims_base_testcnvg, &
                                 ims_base_epfact, ims_base_scale
  use ConvergenceSummaryModule, only: ConvergenceSummaryType
  use MathUtilModule, only: is_close
  use GwfConductanceUtilsModule, only: hcond, vcond, condmean
  use SmoothingModule, only: sQuadraticSaturation, &
                             sQuadraticSaturationDerivative, sQSaturation, &
                             sQSaturationDerivative
  use GwfStorageUtilsModule, only: SsTerms, SyTerms, SsCapacity, SyCapacity
  use InputOutputModule, only: ulasav, ubdsv1, ubdsv06, ubdsvd
  implicit none
  private

contains

  ! ImsLinearBase.f90:1207 ims_base_pccrs
  subroutine ref_pccrs(neq, nja, ia, ja, iapc, japc) bind(C, name='ref_pccrs')
    integer(c_int), value :: neq, nja
    integer(c_int), intent(in) :: ia(neq + 1), ja(nja)
    integer(c_int), intent(inout) :: iapc(neq + 1), japc(nja)
    call ims_base_pccrs(neq, nja, ia, ja, iapc, japc)
  end subroutine ref_pccrs

  ! ImsLinearBase.f90:761 ims_base_pcu (ILU0 / MILU0)
  subroutine ref_pcu(neq, nja, ipc, relax, amat, ia, ja, apc, iapc, japc) &
    bind(C, name='ref_pcu')
    integer(c_int), value :: neq, nja, ipc
    real(c_double), value :: relax
    real(c_double), intent(in) :: amat(nja)
    integer(c_int), intent(in) :: ia(neq + 1), ja(nja)
    real(c_double), intent(inout) :: apc(nja)
    integer(c_int), intent(inout) :: iapc(neq + 1), japc(nja)
    integer(I4B), allocatable :: iw(:), jlu(:), jw(:)
    real(DP), allocatable :: w(:), wlu(:)
    allocate (iw(neq), w(neq), jlu(1), jw(1), wlu(1))
    iw = 0
    w = 0.0_DP
    call ims_base_pcu(6, nja, neq, neq, nja, ipc, relax, amat, ia, ja, apc, &
                      iapc, japc, iw, w, 0, 0.0_DP, 1, 1, 1, jlu, jw, wlu)
    deallocate (iw, w, jlu, jw, wlu)
  end subroutine ref_pcu

  ! ImsLinearBase.f90:1049 ims_base_ilu0a
  subroutine ref_ilu0a(neq, nja, apc, iapc, japc, r, d) &
    bind(C, name='ref_ilu0a')
    integer(c_int), value :: neq, nja
    real(c_double), intent(in) :: apc(nja), r(neq)
    integer(c_int), intent(in) :: iapc(neq + 1), japc(nja)
    real(c_double), intent(inout) :: d(neq)
    call ims_base_ilu0a(nja, neq, apc, iapc, japc, r, d)
  end subroutine ref_ilu0a

  ! ImsLinearBase.f90:1291 ims_base_residual
  subroutine ref_residual(neq, nja, x, b, d, a, ia, ja) &
    bind(C, name='ref_residual')
    integer(c_int), value :: neq, nja
    real(c_double), intent(in) :: x(neq), b(neq), a(nja)
    real(c_double), intent(inout) :: d(neq)
    integer(c_int), intent(in) :: ia(neq + 1), ja(nja)
    call ims_base_residual(neq, nja, x, b, d, a, ia, ja)
  end subroutine ref_residual

  ! sparsekit amux / blas ddot / dnrm2
  subroutine ref_amux(n, x, y, a, ja, ia) bind(C, name='ref_amux')
    integer(c_int), value :: n
    real(c_double), intent(in) :: x(*), a(*)
    real(c_double), intent(inout) :: y(n)
    integer(c_int), intent(in) :: ja(*), ia(n + 1)
    call amux(n, x, y, a, ja, ia)
  end subroutine ref_amux

  function ref_ddot(n, x, y) result(s) bind(C, name='ref_ddot')
    integer(c_int), value :: n
    real(c_double), intent(in) :: x(n), y(n)
    real(c_double) :: s
    real(DP), external :: ddot
    s = ddot(n, x, 1, y, 1)
  end function ref_ddot

  function ref_dnrm2(n, x) result(s) bind(C, name='ref_dnrm2')
    integer(c_int), value :: n
    real(c_double), intent(in) :: x(n)
    real(c_double) :: s
    real(DP), external :: dnrm2
    s = dnrm2(n, x, 1)
  end function ref_dnrm2

  ! ImsLinear.f90:617-750 imslinear_ap sequence without scaling/reordering,
  ! written out here around the reference routines it calls: epfact, pcu,
  ! zero D/P/Q/Z, residual, dnrm2, then ims_base_cg or ims_base_bcgs.  The
  ! per-iteration ConvergenceSummary record (NCONV > 1) is copied out.
  subroutine ref_ims_ap(neq, nja, ia, ja, amat, rhs, x, ilinmeth, ipc, &
                        icnvgopt, north, iter1, dvclose, rclose, relax, &
                        kstp, icnvg, innerit, l2norm0, hdv, hr, hldv, hlr) &
    bind(C, name='ref_ims_ap')
    integer(c_int), value :: neq, nja, ilinmeth, ipc, icnvgopt, north, iter1
    real(c_double), value :: dvclose, rclose, relax
    integer(c_int), value :: kstp
    integer(c_int), intent(in) :: ia(neq + 1), ja(nja)
    real(c_double), intent(in) :: amat(nja)
    real(c_double), intent(inout) :: rhs(neq), x(neq)
    integer(c_int), intent(out) :: icnvg, innerit
    real(c_double), intent(out) :: l2norm0
    real(c_double), intent(out) :: hdv(iter1), hr(iter1)
    integer(c_int), intent(out) :: hldv(iter1), hlr(iter1)
    real(DP), allocatable :: a(:)
    allocate (a(nja))
    a = amat
    call ref_ims_ap_scl(neq, nja, ia, ja, a, rhs, x, ilinmeth, ipc, &
                        icnvgopt, north, iter1, dvclose, rclose, relax, &
                        kstp, 0, icnvg, innerit, l2norm0, hdv, hr, hldv, hlr)
  end subroutine ref_ims_ap

  ! ImsLinearBase.f90:619-754 ims_base_scale, as called by imslinear_ap.
  subroutine ref_ims_scale(iopt, iscl, neq, nja, ia, ja, amat, x, b, &
                           dscale, dscale2) bind(C, name='ref_ims_scale')
    integer(c_int), value :: iopt, iscl, neq, nja
    integer(c_int), intent(in) :: ia(neq + 1), ja(nja)
    real(c_double), intent(inout) :: amat(nja), x(neq), b(neq)
    real(c_double), intent(inout) :: dscale(neq), dscale2(neq)
    call ims_base_scale(iopt, iscl, neq, nja, ia, ja, amat, x, b, dscale, &
                        dscale2)
  end subroutine ref_ims_scale

  ! ref_ims_ap with SCALING_METHOD iscl: ImsLinear.f90:646-652 scales A, x
  ! and b before pcu, ims_base_bcgs receives ISCL and DSCALE, and
  ! ImsLinear.f90:741-747 unscales after the solve (A and b in place).
  subroutine ref_ims_ap_scl(neq, nja, ia, ja, amat, rhs, x, ilinmeth, ipc, &
                            icnvgopt, north, iter1, dvclose, rclose, relax, &
                            kstp, iscl, icnvg, innerit, l2norm0, hdv, hr, &
                            hldv, hlr) bind(C, name='ref_ims_ap_scl')
    integer(c_int), value :: neq, nja, ilinmeth, ipc, icnvgopt, north, iter1
    real(c_double), value :: dvclose, rclose, relax
    integer(c_int), value :: kstp, iscl
    integer(c_int), intent(in) :: ia(neq + 1), ja(nja)
    real(c_double), intent(inout) :: amat(nja)
    real(c_double), intent(inout) :: rhs(neq), x(neq)
    integer(c_int), intent(out) :: icnvg, innerit
    real(c_double), intent(out) :: l2norm0
    real(c_double), intent(out) :: hdv(iter1), hr(iter1)
    integer(c_int), intent(out) :: hldv(iter1), hlr(iter1)
    ! local
    integer(I4B), allocatable :: iapc(:), japc(:), iw(:), jlu(:), jw(:)
    real(DP), allocatable :: apc(:), w(:), wlu(:), dscale(:), dscale2(:)
    real(DP), allocatable :: d(:), p(:), q(:), z(:), t(:), v(:), dhat(:), &
                             phat(:), qhat(:)
    integer(I4B) :: convmodstart(2), nconv, itmax, i
    character(len=31), allocatable :: caccel(:)
    type(ConvergenceSummaryType), pointer :: summary
    real(DP) :: epfact
    real(DP), external :: dnrm2

    allocate (iapc(neq + 1), japc(nja), iw(neq), jlu(1), jw(1))
    allocate (apc(nja), w(neq), wlu(1), dscale(neq), dscale2(neq))
    allocate (d(neq), p(neq), q(neq), z(neq), t(neq), v(neq), dhat(neq), &
              phat(neq), qhat(neq))
    nconv = max(2, iter1)
    allocate (caccel(nconv))
    allocate (summary)
    summary%iter_cnt = 0
    allocate (summary%locdv(1), summary%dvmax(1), summary%locr(1), &
              summary%rmax(1), summary%itinner(nconv))
    allocate (summary%convlocdv(1, nconv), summary%convdvmax(1, nconv), &
              summary%convlocr(1, nconv), summary%convrmax(1, nconv))
    summary%convlocdv = 0
    summary%convlocr = 0
    summary%convdvmax = 0.0_DP
    summary%convrmax = 0.0_DP
    summary%locdv = 0
    summary%locr = 0
    convmodstart(1) = 1
    convmodstart(2) = neq + 1
    iw = 0
    w = 0.0_DP
    apc = 0.0_DP
    dscale = 1.0_DP
    dscale2 = 1.0_DP
    d = 0.0_DP; p = 0.0_DP; q = 0.0_DP; z = 0.0_DP
    t = 0.0_DP; v = 0.0_DP; dhat = 0.0_DP; phat = 0.0_DP; qhat = 0.0_DP

    call ims_base_pccrs(neq, nja, ia, ja, iapc, japc)
    epfact = ims_base_epfact(icnvgopt, kstp)
    if (iscl /= 0) then
      call ims_base_scale(0, iscl, neq, nja, ia, ja, amat, x, rhs, dscale, &
                          dscale2)
    end if
    call ims_base_pcu(6, nja, neq, neq, nja, ipc, relax, amat, ia, ja, apc, &
                      iapc, japc, iw, w, 0, 0.0_DP, 1, 1, 1, jlu, jw, wlu)
    icnvg = 0
    call ims_base_residual(neq, nja, x, rhs, d, amat, ia, ja)
    l2norm0 = dnrm2(neq, d, 1)
    itmax = iter1
    if (l2norm0 == 0.0_DP) then
      itmax = 0
      icnvg = 1
    end if
    innerit = 0
    if (ilinmeth == 1) then
      call ims_base_cg(icnvg, itmax, innerit, neq, nja, neq, nja, ipc, &
                       icnvgopt, north, dvclose, rclose, l2norm0, epfact, &
                       ia, ja, amat, iapc, japc, apc, x, rhs, d, p, q, z, &
                       1, iw, jlu, nconv, 1, convmodstart, caccel, summary)
    else
      call ims_base_bcgs(icnvg, itmax, innerit, neq, nja, neq, nja, ipc, &
                         icnvgopt, north, iscl, dscale, dvclose, rclose, &
                         l2norm0, epfact, ia, ja, amat, iapc, japc, apc, x, &
                         rhs, d, p, q, t, v, dhat, phat, qhat, 1, iw, jlu, &
                         nconv, 1, convmodstart, caccel, summary)
    end if
    if (iscl /= 0) then
      call ims_base_scale(1, iscl, neq, nja, ia, ja, amat, x, rhs, dscale, &
                          dscale2)
    end if
    do i = 1, min(innerit, iter1)
      hdv(i) = summary%convdvmax(1, i)
      hr(i) = summary%convrmax(1, i)
      hldv(i) = summary%convlocdv(1, i)
      hlr(i) = summary%convlocr(1, i)
    end do
    deallocate (summary%locdv, summary%dvmax, summary%locr, summary%rmax, &
                summary%itinner, summary%convlocdv, summary%convdvmax, &
                summary%convlocr, summary%convrmax)
    deallocate (summary)
  end subroutine ref_ims_ap_scl

  ! ImsLinearBase.f90:1101 ims_base_testcnvg
  subroutine ref_testcnvg(icnvgopt, icnvg, iiter, dvmax, rmax, rmax0, &
                          epfact, dvclose, rclose) bind(C, name='ref_testcnvg')
    integer(c_int), value :: icnvgopt, iiter
    integer(c_int), intent(inout) :: icnvg
    real(c_double), value :: dvmax, rmax, rmax0, epfact, dvclose, rclose
    call ims_base_testcnvg(icnvgopt, icnvg, iiter, dvmax, rmax, rmax0, &
                           epfact, dvclose, rclose)
  end subroutine ref_testcnvg

  function ref_epfact(icnvgopt, kstp) result(e) bind(C, name='ref_epfact')
    integer(c_int), value :: icnvgopt, kstp
    real(c_double) :: e
    e = ims_base_epfact(icnvgopt, kstp)
  end function ref_epfact

  function ref_is_close(a, b) result(r) bind(C, name='ref_is_close')
    real(c_double), value :: a, b
    integer(c_int) :: r
    r = 0
    if (is_close(a, b)) r = 1
  end function ref_is_close

  ! GwfConductanceUtils.f90:43 hcond
  function ref_hcond(ibdn, ibdm, ictn, ictm, iupstream, ihc, icellavg, &
                     condsat, hn, hm, satn, satm, hkn, hkm, topn, topm, &
                     botn, botm, cln, clm, fawidth) result(c) &
    bind(C, name='ref_hcond')
    integer(c_int), value :: ibdn, ibdm, ictn, ictm, iupstream, ihc, icellavg
    real(c_double), value :: condsat, hn, hm, satn, satm, hkn, hkm, topn, &
                             topm, botn, botm, cln, clm, fawidth
    real(c_double) :: c
    c = hcond(ibdn, ibdm, ictn, ictm, iupstream, ihc, icellavg, condsat, &
              hn, hm, satn, satm, hkn, hkm, topn, topm, botn, botm, cln, clm, &
              fawidth)
  end function ref_hcond

  ! GwfConductanceUtils.f90:149 vcond
  function ref_vcond(ibdn, ibdm, ictn, ictm, inewton, ivarcv, idewatcv, &
                     condsat, hn, hm, vkn, vkm, satn, satm, topn, topm, &
                     botn, botm, flowarea) result(c) bind(C, name='ref_vcond')
    integer(c_int), value :: ibdn, ibdm, ictn, ictm, inewton, ivarcv, idewatcv
    real(c_double), value :: condsat, hn, hm, vkn, vkm, satn, satm, topn, &
                             topm, botn, botm, flowarea
    real(c_double) :: c
    c = vcond(ibdn, ibdm, ictn, ictm, inewton, ivarcv, idewatcv, condsat, &
              hn, hm, vkn, vkm, satn, satm, topn, topm, botn, botm, flowarea)
  end function ref_vcond

  ! GwfConductanceUtils.f90:226 condmean
  function ref_condmean(k1, k2, thick1, thick2, cl1, cl2, width, iavg) &
    result(c) bind(C, name='ref_condmean')
    real(c_double), value :: k1, k2, thick1, thick2, cl1, cl2, width
    integer(c_int), value :: iavg
    real(c_double) :: c
    c = condmean(k1, k2, thick1, thick2, cl1, cl2, width, iavg)
  end function ref_condmean

  ! SmoothingFunctions.f90:275 / :364
  function ref_sqsat(top, bot, x, eps) result(y) bind(C, name='ref_sqsat')
    real(c_double), value :: top, bot, x, eps
    real(c_double) :: y
    y = sQuadraticSaturation(top, bot, x, eps)
  end function ref_sqsat

  function ref_sqsat_deriv(top, bot, x, eps) result(y) &
    bind(C, name='ref_sqsat_deriv')
    real(c_double), value :: top, bot, x, eps
    real(c_double) :: y
    y = sQuadraticSaturationDerivative(top, bot, x, eps)
  end function ref_sqsat_deriv

  ! SmoothingFunctions.f90:412 sQSaturation / :466 sQSaturationDerivative
  function ref_sqsat_cubic(top, bot, x, c1, c2) result(y) &
    bind(C, name='ref_sqsat_cubic')
    real(c_double), value :: top, bot, x, c1, c2
    real(c_double) :: y
    y = sQSaturation(top, bot, x, c1, c2)
  end function ref_sqsat_cubic

  function ref_sqsat_cubic_deriv(top, bot, x, c1, c2) result(y) &
    bind(C, name='ref_sqsat_cubic_deriv')
    real(c_double), value :: top, bot, x, c1, c2
    real(c_double) :: y
    y = sQSaturationDerivative(top, bot, x, c1, c2)
  end function ref_sqsat_cubic_deriv

  ! GwfStorageUtils.f90:32 SsTerms, :98 SyTerms, :143 SsCapacity, :169
  subroutine ref_ssterms(iconvert, iorig_ss, iconf_ss, top, bot, rho1, &
                         rho1old, snnew, snold, hnew, hold, aterm, rhsterm) &
    bind(C, name='ref_ssterms')
    integer(c_int), value :: iconvert, iorig_ss, iconf_ss
    real(c_double), value :: top, bot, rho1, rho1old, snnew, snold, hnew, hold
    real(c_double), intent(out) :: aterm, rhsterm
    call SsTerms(iconvert, iorig_ss, iconf_ss, top, bot, rho1, rho1old, &
                 snnew, snold, hnew, hold, aterm, rhsterm)
  end subroutine ref_ssterms

  subroutine ref_syterms(top, bot, rho2, rho2old, snnew, snold, aterm, &
                         rhsterm) bind(C, name='ref_syterms')
    real(c_double), value :: top, bot, rho2, rho2old, snnew, snold
    real(c_double), intent(out) :: aterm, rhsterm
    call SyTerms(top, bot, rho2, rho2old, snnew, snold, aterm, rhsterm)
  end subroutine ref_syterms

  function ref_sscapacity(istor_coef, top, bot, area, ss) result(s) &
    bind(C, name='ref_sscapacity')
    integer(c_int), value :: istor_coef
    real(c_double), value :: top, bot, area, ss
    real(c_double) :: s
    s = SsCapacity(istor_coef, top, bot, area, ss)
  end function ref_sscapacity

  function ref_sycapacity(area, sy) result(s) bind(C, name='ref_sycapacity')
    real(c_double), value :: area, sy
    real(c_double) :: s
    s = SyCapacity(area, sy)
  end function ref_sycapacity

  ! --- binary output: the reference's own record writers on a unit opened
  ! as the reference opens binary output (OpenSpec.f90:17-18:
  ! FORM = 'UNFORMATTED', ACCESS = 'STREAM'); append /= 0 adds to the file

  function open_stream(path, plen, append) result(iu)
    character(kind=c_char), intent(in) :: path(*)
    integer(c_int), intent(in) :: plen, append
    integer(I4B) :: iu
    character(len=plen) :: fname
    integer(I4B) :: i
    do i = 1, plen
      fname(i:i) = path(i)
    end do
    if (append /= 0) then
      open (newunit=iu, file=fname, form='UNFORMATTED', access='STREAM', &
            status='OLD', position='APPEND')
    else
      open (newunit=iu, file=fname, form='UNFORMATTED', access='STREAM', &
            status='REPLACE')
    end if
  end function open_stream

  function text16(c) result(t)
    character(kind=c_char), intent(in) :: c(16)
    character(len=16) :: t
    integer(I4B) :: i
    do i = 1, 16
      t(i:i) = c(i)
    end do
  end function text16

  ! InputOutput.f90:924 ulasav
  subroutine ref_ulasav(path, plen, append, buf, text, kstp, kper, pertim, &
                        totim, ncol, nrow, ilay) bind(C, name='ref_ulasav')
    character(kind=c_char), intent(in) :: path(*), text(16)
    integer(c_int), value :: plen, append, kstp, kper, ncol, nrow, ilay
    real(c_double), value :: pertim, totim
    real(c_double), intent(in) :: buf(ncol, nrow)
    character(len=16) :: t
    integer(I4B) :: iu
    real(DP) :: pt, tt
    t = text16(text)
    pt = pertim
    tt = totim
    iu = open_stream(path, plen, append)
    call ulasav(buf, t, kstp, kper, pt, tt, ncol, nrow, ilay, iu)
    close (iu)
  end subroutine ref_ulasav

  ! InputOutput.f90:944 ubdsv1
  subroutine ref_ubdsv1(path, plen, append, kstp, kper, text, buff, ncol, nrow, &
                        nlay, delt, pertim, totim) bind(C, name='ref_ubdsv1')
    character(kind=c_char), intent(in) :: path(*), text(16)
    integer(c_int), value :: plen, append, kstp, kper, ncol, nrow, nlay
    real(c_double), value :: delt, pertim, totim
    real(c_double), intent(in) :: buff(ncol * nrow * nlay)
    integer(I4B) :: iu
    iu = open_stream(path, plen, append)
    call ubdsv1(kstp, kper, text16(text), iu, buff, ncol, nrow, nlay, 0, &
                delt, pertim, totim)
    close (iu)
  end subroutine ref_ubdsv1

  ! InputOutput.f90:979 ubdsv06 header, then one :1053 ubdsvd entry per
  ! list item (no auxiliary variables)
  subroutine ref_ubdsv06_list(path, plen, append, kstp, kper, text, model1, &
                              package1, model2, package2, ncol, nrow, nlay, &
                              nlist, nodes, nodes2, q, delt, pertim, totim) &
    bind(C, name='ref_ubdsv06_list')
    character(kind=c_char), intent(in) :: path(*), text(16), model1(16), &
                                          package1(16), model2(16), package2(16)
    integer(c_int), value :: plen, append, kstp, kper, ncol, nrow, nlay, nlist
    integer(c_int), intent(in) :: nodes(nlist), nodes2(nlist)
    real(c_double), intent(in) :: q(nlist)
    real(c_double), value :: delt, pertim, totim
    character(len=16), dimension(0) :: auxtxt
    real(DP), dimension(1) :: aux
    integer(I4B) :: iu, i
    aux = 0.0_DP
    iu = open_stream(path, plen, append)
    call ubdsv06(kstp, kper, text16(text), text16(model1), text16(package1), &
                 text16(model2), text16(package2), iu, 0, auxtxt, ncol, nrow, &
                 nlay, nlist, 0, delt, pertim, totim)
    do i = 1, nlist
      call ubdsvd(iu, nodes(i), nodes2(i), q(i), 0, aux)
    end do
    close (iu)
  end subroutine ref_ubdsv06_list

end module mf6ref_shim
