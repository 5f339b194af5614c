// mf6hip_xmi.hip -- the XMI-style outer loop of one device-resident GWF
// model: prepare_solve / solve(has_converged) / finalize_solve
// (srcbmi/mf6xmi.F90:220-340 xmi_prepare_solve, xmi_solve,
// xmi_finalize_solve), one NumericalSolution outer iteration per solve call
// (NumericalSolutionType%solve, src/Solution/NumericalSolution.f90
// :1482-1837), driving the library's own device entry points:
//
//   prepare_solve  gwf_ad + chd_ad (mf6hip_gwf_advance), iteration counter 0
//   solve          backtracking (sln_backtracking :2680-2776) when
//                  BACKTRACKING_NUMBER > 0; sln_buildsystem (:1941,
//                  mf6hip_gwf_formulate); pseudo-transient continuation for
//                  steady-state Newton models (sln_calc_ptc :2936-2962 ->
//                  gwf_ptc, gwf.f90:625-679); sln_ls + sln_get_dxmax
//                  (mf6hip_gwf_linear_solve); the outer test |hncg| <=
//                  OUTER_DVCLOSE (:1625-1629); when not converged,
//                  sln_underrelax / sln_calcdx and npf_nur
//                  (mf6hip_gwf_outer_update) with the NUR convergence test
//                  of sln_nur_has_converged (:3205-3218)
//   finalize_solve convergence of the time step and the outer iterations
//                  taken
//
// Heads stay in HBM; only scalars cross to the host.  Host code only.
#include <cmath>
#include <string>

#include "mf6hip_handle.h"

#define XCK(x)                        \
  do {                                \
    int rc_ = (x);                    \
    if (rc_ != MF6HIP_OK) return rc_; \
  } while (0)

namespace {

int xmi_check(Mf6Handle *h) {
  if (!h) { mf6hip_set_error("null handle"); return MF6HIP_EINVAL; }
  if (!h->has_model) { mf6hip_set_error("no GWF model on the handle (mf6hip_gwf_create)"); return MF6HIP_EINVAL; }
  if (!h->xmi.nl_set) { mf6hip_set_error("NONLINEAR settings not set (mf6hip_xmi_set_nonlinear)"); return MF6HIP_EINVAL; }
  return MF6HIP_OK;
}

// sln_backtracking: refill without Newton terms, compare the L2 norm of the
// residual with the last outer iteration's and pull x back towards xtemp
int xmi_backtracking(Mf6Handle *h, void *hv, int kiter, int *nbt) {
  Mf6Handle::Xmi &X = h->xmi;
  const mf6hip_nonlinear &nl = X.nl;
  double l2 = 0.0, rv = 0.0;
  *nbt = 0;
  XCK(mf6hip_gwf_formulate(hv, kiter, 0, nullptr, nullptr));
  XCK(mf6hip_gwf_residual_norms(hv, &l2, &rv));
  if (kiter == 1) {
    X.res_prev = l2;
    return MF6HIP_OK;
  }
  X.res_new = l2;
  if (X.res_new > X.res_prev * nl.backtracking_tolerance) {
    for (int nb = 1; nb <= nl.backtracking_number; nb++) {
      int flag = 0;
      XCK(mf6hip_gwf_backtrack(hv, nl.backtracking_reduction_factor, nl.outer_dvclose, &flag));
      if (flag == 0) break;
      *nbt = nb;
      XCK(mf6hip_gwf_formulate(hv, kiter, 0, nullptr, nullptr));
      XCK(mf6hip_gwf_residual_norms(hv, &l2, &rv));
      X.res_new = l2;
      if (nb == nl.backtracking_number) break;
      if (X.res_new < X.res_prev * nl.backtracking_tolerance) break;
      if (X.res_new < nl.backtracking_residual_limit) break;
    }
  }
  X.res_prev = X.res_new;
  return MF6HIP_OK;
}

}  // namespace

extern "C" int mf6hip_xmi_set_nonlinear(void *handle, const mf6hip_nonlinear *nl) {
  Mf6Handle *h = (Mf6Handle *)handle;
  if (!h || !nl) { mf6hip_set_error("null argument"); return MF6HIP_EINVAL; }
  if (nl->outer_maximum < 1 || !(nl->outer_dvclose > 0.0)) {
    mf6hip_set_error("OUTER_MAXIMUM must be >= 1 and OUTER_DVCLOSE > 0");
    return MF6HIP_EINVAL;
  }
  if (nl->nonmeth < 0 || nl->nonmeth > 3) { mf6hip_set_error("unknown UNDER_RELAXATION"); return MF6HIP_EINVAL; }
  if (nl->backtracking_number < 0) { mf6hip_set_error("BACKTRACKING_NUMBER must be >= 0"); return MF6HIP_EINVAL; }
  if (nl->iallowptc < -1 || nl->iallowptc > 1) { mf6hip_set_error("iallowptc must be -1, 0 or 1"); return MF6HIP_EINVAL; }
  Mf6Handle::Xmi &X = h->xmi;
  X.nl = *nl;
  X.nl_set = true;
  // sln_ar: THETA below 1e-3 is raised to 1e-3 (NumericalSolution.f90:868-870)
  X.ur = mf6hip_underrelax{nl->nonmeth, nl->gamma, nl->theta < 1e-3 ? 1e-3 : nl->theta,
                           nl->akappa, nl->amomentum, 0.0, 0.0, 0.0};
  X.ptc = mf6hip_ptc{0, nl->iallowptc, nl->iptcopt, 0, 0, 0.0, nl->ptcdel0, nl->ptcexp, 0.0, 0.0};
  X.res_prev = X.res_new = 0.0;
  return MF6HIP_OK;
}

extern "C" int mf6hip_xmi_prepare_solve(void *handle, int kper, int kstp, int iss, double delt) {
  Mf6Handle *h = (Mf6Handle *)handle;
  XCK(xmi_check(h));
  if (!(delt > 0.0)) { mf6hip_set_error("delt must be > 0"); return MF6HIP_EINVAL; }
  Mf6Handle::Xmi &X = h->xmi;
  XCK(mf6hip_gwf_advance(handle, iss, delt));
  X.kper = kper;
  X.kstp = kstp;
  X.iss = iss;
  X.delt = delt;
  X.kiter = 0;
  X.converged = 0;
  X.prepared = true;
  return MF6HIP_OK;
}

extern "C" int mf6hip_xmi_solve(void *handle, int *has_converged) {
  Mf6Handle *h = (Mf6Handle *)handle;
  XCK(xmi_check(h));
  if (!has_converged) { mf6hip_set_error("null argument"); return MF6HIP_EINVAL; }
  Mf6Handle::Xmi &X = h->xmi;
  if (!X.prepared) { mf6hip_set_error("solve before prepare_solve"); return MF6HIP_EINVAL; }
  const mf6hip_nonlinear &nl = X.nl;
  const int kiter = ++X.kiter;
  mf6hip_xmi_iteration &R = X.last;
  R = mf6hip_xmi_iteration{};
  R.kiter = kiter;
  if (nl.backtracking_number > 0) XCK(xmi_backtracking(h, handle, kiter, &R.nbt));
  XCK(mf6hip_gwf_formulate(handle, kiter, 1, nullptr, nullptr));
  // pseudo-transient continuation of steady-state Newton models
  mf6hip_ptc *ptc = nullptr;
  if (X.iss > 0 && h->m.inewton) {
    double l2 = 0.0, ptcf = 0.0;
    XCK(mf6hip_gwf_residual_norms(handle, &l2, &ptcf));
    if (ptcf == 0.0) ptcf = 1.0 / (X.delt * 10.0);
    X.ptc.iptc = 1;
    X.ptc.ptcf = ptcf;
    X.ptc.kstp = X.kstp;
    X.ptc.kper = X.kper;
    ptc = &X.ptc;
  }
  XCK(mf6hip_gwf_linear_solve(handle, X.kstp, kiter, ptc, &R.icnvg_inner, &R.inner, &R.hncg,
                              &R.lrch, nullptr));
  R.iptc = ptc ? X.ptc.iptc : 0;
  int conv = std::fabs(R.hncg) <= nl.outer_dvclose;
  if (!conv) {
    int inur = 0, l2 = 0;
    double dxo = 0.0, h2 = 0.0;
    X.ur.bigch = R.hncg;
    XCK(mf6hip_gwf_outer_update(handle, kiter, &X.ur, &inur, &dxo, &h2, &l2));
    if (inur) {
      R.nur = 1;
      if (std::fabs(dxo) <= nl.outer_dvclose && std::fabs(R.hncg) <= nl.outer_dvclose) {
        conv = 1;
        R.hncg = h2;
        R.lrch = l2;
      }
    }
  }
  R.converged = conv;
  X.converged = conv;
  *has_converged = conv;
  return MF6HIP_OK;
}

extern "C" int mf6hip_xmi_last_iteration(void *handle, mf6hip_xmi_iteration *rec) {
  Mf6Handle *h = (Mf6Handle *)handle;
  if (!h || !rec) { mf6hip_set_error("null argument"); return MF6HIP_EINVAL; }
  *rec = h->xmi.last;
  return MF6HIP_OK;
}

extern "C" int mf6hip_xmi_finalize_solve(void *handle, int *converged, int *outer_iterations) {
  Mf6Handle *h = (Mf6Handle *)handle;
  XCK(xmi_check(h));
  Mf6Handle::Xmi &X = h->xmi;
  if (!X.prepared) { mf6hip_set_error("finalize_solve before prepare_solve"); return MF6HIP_EINVAL; }
  if (converged) *converged = X.converged;
  if (outer_iterations) *outer_iterations = X.kiter;
  X.prepared = false;
  return MF6HIP_OK;
}
