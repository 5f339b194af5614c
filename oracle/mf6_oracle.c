/*
 * mf6_oracle.c -- CPU restatement of the MODFLOW 6 IMS / GWF hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle: it is
 * linked by tests/, by __graft_entry__.smoke() and by the cpu_baseline leg
 * of bench.py, and by nothing else.  The product library (libmf6hip.so)
 * never calls it.
 *
 * Every routine restates one reference routine, statement order preserved,
 * so that the floating-point results are bit-identical to the reference
 * when compiled without FMA contraction (-ffp-contract=off).  Array indices
 * are 0-based here; the reference is 1-based.  CSR rows keep the reference
 * layout: diagonal first, then ascending columns
 * (src/Utilities/Sparse.f90:217-239).
 *
 * Pinning: tests/test_oracle_vs_ref.py compares these routines bit-for-bit
 * with the reference's own Fortran (src/Solution/LinearMethods/
 * ImsLinearBase.f90, GwfConductanceUtils.f90, SmoothingFunctions.f90,
 * GwfStorageUtils.f90, sparsekit amux, blas ddot/dnrm2) compiled from
 * /root/reference by oracle/Makefile into oracle/_ref.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define DZERO 0.0
#define DONE 1.0
#define DHALF 0.5
#define DTHREE 3.0
#define DEM3 1.0e-3
#define DEM6 1.0e-6
#define DEM15 1.0e-15
#define DEM20 1.0e-20
#define DP9 0.9
/* src/Utilities/Constants.f90:120-122 */
#define DPREC 2.220446049250313080847263336181640625e-16
#define DSAME (100.0 * DPREC)
/* src/Utilities/Constants.f90:125-126 */
#define DLNLOW 0.995
#define DLNHIGH 1.005

/* Fortran SIGN(a, b): |a| carrying the sign bit of b. */
static double fsign(double a, double b) { return copysign(fabs(a), b); }

/* ------------------------------------------------------------------ */
/* BLAS-1 / SPARSKIT                                                   */
/* ------------------------------------------------------------------ */

/* src/Utilities/Libraries/blas/blas1_d.f90:295 ddot = dot_product, summed
 * left to right. */
double orc_ddot(int n, const double *dx, const double *dy)
{
  double s = 0.0;
  for (int i = 0; i < n; i++) s = s + dx[i] * dy[i];
  return s;
}

/* src/Utilities/Libraries/blas/blas1_d.f90:387-478 dnrm2, scaled sum of
 * squares. */
double orc_dnrm2(int n, const double *x)
{
  if (n < 1) return 0.0;
  if (n == 1) return fabs(x[0]);
  double scale = 0.0, ssq = 1.0;
  for (int i = 0; i < n; i++) {
    if (x[i] != 0.0) {
      double absxi = fabs(x[i]);
      if (scale < absxi) {
        double t = scale / absxi;
        ssq = 1.0 + ssq * (t * t);
        scale = absxi;
      } else {
        double t = absxi / scale;
        ssq = ssq + t * t;
      }
    }
  }
  return scale * sqrt(ssq);
}

/* src/Utilities/Libraries/sparsekit/sparsekit.f90:1-59 amux: y = A x,
 * row-dot form, t starts at 0. */
void orc_amux(int n, const double *x, double *y, const double *a,
              const int *ja, const int *ia)
{
  for (int i = 0; i < n; i++) {
    double t = 0.0;
    for (int k = ia[i]; k < ia[i + 1]; k++) t = t + a[k] * x[ja[k]];
    y[i] = t;
  }
}

/* src/Utilities/MathUtil.f90:45-86 is_close with rtol = DSAME, atol = 0,
 * symmetric test. */
int orc_is_close(double a, double b)
{
  if (a == b) return 1;
  double m = fmax(fabs(a), fabs(b));
  return fabs(a - b) <= fmax(DSAME * m, 0.0);
}

/* ------------------------------------------------------------------ */
/* IMS linear accelerator (ImsLinearBase.f90)                          */
/* ------------------------------------------------------------------ */

/* ImsLinearBase.f90:1316-1333 ims_base_epfact.  The 0.01 / 0.10 literals
 * are default-kind (single precision) reals in the reference. */
double orc_epfact(int icnvgopt, int kstp)
{
  if (icnvgopt == 2) return kstp == 1 ? (double)0.01f : (double)0.10f;
  if (icnvgopt == 4) return 1.0e-4;
  return DONE;
}

/* ImsLinearBase.f90:1101-1146 ims_base_testcnvg */
void orc_testcnvg(int icnvgopt, int *icnvg, int iiter, double dvmax,
                  double rmax, double rmax0, double epfact, double dvclose,
                  double rclose)
{
  if (icnvgopt == 0) {
    if (fabs(dvmax) <= dvclose && fabs(rmax) <= rclose) *icnvg = 1;
  } else if (icnvgopt == 1) {
    if (fabs(dvmax) <= dvclose && fabs(rmax) <= rclose) *icnvg = iiter == 1 ? 1 : -1;
  } else if (icnvgopt == 2) {
    if (fabs(dvmax) <= dvclose || rmax <= rclose) *icnvg = 1;
    else if (rmax <= rmax0 * epfact) *icnvg = -1;
  } else if (icnvgopt == 3) {
    if (fabs(dvmax) <= dvclose) *icnvg = 1;
    else if (rmax <= rmax0 * rclose) *icnvg = -1;
  } else if (icnvgopt == 4) {
    if (fabs(dvmax) <= dvclose && rmax <= rclose) *icnvg = 1;
    else if (rmax <= rmax0 * epfact) *icnvg = -1;
  }
}

/* ImsLinearBase.f90:1207-1261 ims_base_pccrs.  0-based: APC[0:neq] hold the
 * inverse pivots, APC[neq:nja] the off-diagonal factors; iapc[n] is the
 * first off-diagonal slot of row n, japc[n] (n < neq) the first upper slot,
 * japc[k] (k >= neq) the column. */
void orc_pccrs(int neq, int nja, const int *ia, const int *ja, int *iapc,
               int *japc)
{
  int ip = neq;
  int maxlen = 0;
  for (int n = 0; n < neq; n++)
    if (ia[n + 1] - ia[n] > maxlen) maxlen = ia[n + 1] - ia[n];
  int *iarr = (int *)malloc(sizeof(int) * (size_t)(maxlen + 1));
  for (int n = 0; n < neq; n++) {
    int ic = 0;
    for (int j = ia[n]; j < ia[n + 1]; j++) {
      if (ja[j] == n) continue;
      iarr[ic++] = ja[j];
    }
    /* ims_base_isort (:1268-1283) */
    for (int i = 0; i < ic - 1; i++)
      for (int j = i + 1; j < ic; j++)
        if (iarr[i] > iarr[j]) { int t = iarr[j]; iarr[j] = iarr[i]; iarr[i] = t; }
    iapc[n] = ip;
    for (int j = 0; j < ic; j++) japc[ip++] = iarr[j];
  }
  iapc[neq] = nja;
  for (int n = 0; n < neq; n++) {
    japc[n] = iapc[n + 1];
    for (int j = iapc[n]; j < iapc[n + 1]; j++)
      if (japc[j] > n) { japc[n] = j; break; }
  }
  free(iarr);
}

/* ImsLinearBase.f90:928-1042 ims_base_pcilu0.  Returns the row at which the
 * factorisation stopped early (pivot sign change / zero pivot), or -1.  As
 * in the reference (:1040-1041) *ipcflag is reset to 0 on every exit, so
 * rows from the failing one on keep their previous APC values. */
int orc_pcilu0(int nja, int neq, const double *amat, const int *ia,
               const int *ja, double *apc, const int *iapc, const int *japc,
               int *iw, double *w, double relax, int *ipcflag, double delta)
{
  int stopped = -1;
  (void)nja;
  for (int n = 0; n < neq; n++) { iw[n] = 0; w[n] = DZERO; }
  for (int n = 0; n < neq; n++) {
    for (int j = ia[n]; j < ia[n + 1]; j++) {
      int jcol = ja[j];
      iw[jcol] = 1;
      w[jcol] = w[jcol] + amat[j];
    }
    int ic0 = iapc[n], ic1 = iapc[n + 1];
    int iu = japc[n];
    double rs = DZERO;
    for (int j = ic0; j < iu; j++) {
      int jcol = japc[j];
      int iic1 = iapc[jcol + 1];
      int iiu = japc[jcol];
      double tl = w[jcol] * apc[jcol];
      w[jcol] = tl;
      for (int jj = iiu; jj < iic1; jj++) {
        int jjcol = japc[jj];
        if (iw[jjcol] != 0) w[jjcol] = w[jjcol] - tl * apc[jj];
        else rs = rs + tl * apc[jj];
      }
    }
    double d = w[n];
    double tl = (DONE + delta) * d - (relax * rs);
    double sd1 = fsign(d, tl);
    if (sd1 != d) {
      if (*ipcflag > 1) tl = fsign(DEM6, d);
      else { *ipcflag = 1; stopped = n; break; }
    }
    if (fabs(tl) == DZERO) {
      if (*ipcflag > 1) tl = fsign(DEM6, d);
      else { *ipcflag = 1; stopped = n; break; }
    }
    apc[n] = DONE / tl;
    iw[n] = 0;
    w[n] = DZERO;
    for (int j = ic0; j < ic1; j++) {
      int jcol = japc[j];
      apc[j] = w[jcol];
      iw[jcol] = 0;
      w[jcol] = DZERO;
    }
  }
  *ipcflag = 0;
  return stopped;
}

/* ImsLinearBase.f90:761-864 ims_base_pcu restricted to ILU0/MILU0 (IPC 1,2).
 * Because pcilu0 always returns IPCFLAG = 0 the delta-shift retry never runs
 * for these preconditioners; the loop is kept for fidelity. */
int orc_pcu(int nja, int neq, int ipc, double relax, const double *amat,
            const int *ia, const int *ja, double *apc, const int *iapc,
            const int *japc, int *iw, double *w)
{
  int ipcflag = 0, icount = 0, stopped = -1;
  double delta = DZERO;
  for (;;) {
    if (ipc == 1 || ipc == 2)
      stopped = orc_pcilu0(nja, neq, amat, ia, ja, apc, iapc, japc, iw, w,
                           relax, &ipcflag, delta);
    else
      ipcflag = 0;
    if (ipcflag < 1) break;
    delta = 1.5 * delta + DEM3;
    ipcflag = 0;
    if (delta > DHALF) { delta = DHALF; ipcflag = 2; }
    icount++;
    if (icount > 10) break;
  }
  return stopped;
}

/* ImsLinearBase.f90:1049-1092 ims_base_ilu0a */
void orc_ilu0a(int nja, int neq, const double *apc, const int *iapc,
               const int *japc, const double *r, double *d)
{
  (void)nja;
  for (int n = 0; n < neq; n++) {
    double tv = r[n];
    for (int j = iapc[n]; j < japc[n]; j++) tv = tv - apc[j] * d[japc[j]];
    d[n] = tv;
  }
  for (int n = neq - 1; n >= 0; n--) {
    double tv = d[n];
    for (int j = japc[n]; j < iapc[n + 1]; j++) tv = tv - apc[j] * d[japc[j]];
    d[n] = tv * apc[n];
  }
}

/* ImsLinearBase.f90:1291-1312 ims_base_residual: d = b - A x */
void orc_residual(int neq, const double *x, const double *b, double *d,
                  const double *a, const int *ia, const int *ja)
{
  orc_amux(neq, x, d, a, ja, ia);
  for (int n = 0; n < neq; n++) d[n] = b[n] - d[n];
}

/* Per-iteration record kept by the Krylov routines (the reference keeps
 * the same values in ConvergenceSummaryType, ImsLinearBase.f90:186-197). */
typedef struct {
  double dvmax;  /* signed largest |x change| */
  double rmax;   /* signed largest |residual| */
  double l2norm;
  double alpha;
  double omega;  /* BiCGSTAB only */
  int locdv;     /* 0-based cell of dvmax */
  int locr;      /* 0-based cell of rmax */
} orc_iter_rec;

/* ImsLinearBase.f90:30-240 ims_base_cg, single model (CONVNMOD = 1). */
void orc_cg(int *icnvg, int itmax, int *innerit, int neq, int nja,
            int icnvgopt, int north, double dvclose, double rclose,
            double l2norm0, double epfact, const int *ia0, const int *ja0,
            const double *a0, const int *iapc, const int *japc,
            const double *apc, double *x, const double *b, double *d,
            double *p, double *q, double *z, orc_iter_rec *hist)
{
  double rho0 = DZERO, rho = DZERO;
  *innerit = 0;
  for (int iiter = 1; iiter <= itmax; iiter++) {
    *innerit = *innerit + 1;
    orc_ilu0a(nja, neq, apc, iapc, japc, d, z);
    rho = orc_ddot(neq, d, z);
    if (iiter == 1) {
      for (int n = 0; n < neq; n++) p[n] = z[n];
    } else {
      double beta = rho / rho0;
      for (int n = 0; n < neq; n++) p[n] = z[n] + beta * p[n];
    }
    orc_amux(neq, p, q, a0, ja0, ia0);
    double denominator = orc_ddot(neq, p, q);
    denominator = denominator + fsign(DPREC, denominator);
    double alpha = rho / denominator;
    double deltax = DZERO, rmax = DZERO, l2norm = DZERO;
    int xloc = -1, rloc = -1;
    for (int n = 0; n < neq; n++) {
      double tv = alpha * p[n];
      x[n] = x[n] + tv;
      if (fabs(tv) > fabs(deltax)) { deltax = tv; xloc = n; }
      tv = d[n];
      tv = tv - alpha * q[n];
      d[n] = tv;
      if (fabs(tv) > fabs(rmax)) { rmax = tv; rloc = n; }
      l2norm = l2norm + tv * tv;
    }
    l2norm = sqrt(l2norm);
    if (hist) {
      orc_iter_rec *h = &hist[iiter - 1];
      h->dvmax = deltax; h->rmax = rmax; h->l2norm = l2norm;
      h->alpha = alpha; h->omega = 0.0; h->locdv = xloc; h->locr = rloc;
    }
    double rcnvg = (icnvgopt == 2 || icnvgopt == 3 || icnvgopt == 4) ? l2norm : rmax;
    orc_testcnvg(icnvgopt, icnvg, *innerit, deltax, rcnvg, l2norm0, epfact,
                 dvclose, rclose);
    if (rcnvg == DZERO) *icnvg = 1;
    if (*icnvg != 0) break;
    if (orc_is_close(rho, rho0)) break;
    if (north > 0 && ((iiter + 1) % north) == 0)
      orc_residual(neq, x, b, d, a0, ia0, ja0);
    if (rho == DZERO) break;
    rho0 = rho;
  }
  if (*icnvg < 0) *icnvg = 0;
}

/* ImsLinearBase.f90:249-549 ims_base_bcgs, single model, ISCL = 0. */
void orc_bcgs(int *icnvg, int itmax, int *innerit, int neq, int nja,
              int icnvgopt, int north, double dvclose, double rclose,
              double l2norm0, double epfact, const int *ia0, const int *ja0,
              const double *a0, const int *iapc, const int *japc,
              const double *apc, double *x, const double *b, double *d,
              double *p, double *q, double *t, double *v, double *dhat,
              double *phat, double *qhat, const double *dscale,
              orc_iter_rec *hist)
{
  double alpha = DZERO, alpha0 = DZERO, beta = DZERO, rho = DZERO;
  double rho0 = DZERO, omega = DZERO, omega0 = DZERO;
  *innerit = 0;
  for (int n = 0; n < neq; n++) dhat[n] = d[n];
  for (int iiter = 1; iiter <= itmax; iiter++) {
    *innerit = *innerit + 1;
    rho = orc_ddot(neq, dhat, d);
    if (iiter == 1) {
      for (int n = 0; n < neq; n++) p[n] = d[n];
    } else {
      beta = (rho / rho0) * (alpha0 / omega0);
      for (int n = 0; n < neq; n++) p[n] = d[n] + beta * (p[n] - omega0 * v[n]);
    }
    orc_ilu0a(nja, neq, apc, iapc, japc, p, phat);
    orc_amux(neq, phat, v, a0, ja0, ia0);
    double denominator = orc_ddot(neq, dhat, v);
    denominator = denominator + fsign(DPREC, denominator);
    alpha = rho / denominator;
    for (int n = 0; n < neq; n++) q[n] = d[n] - alpha * v[n];
    orc_ilu0a(nja, neq, apc, iapc, japc, q, qhat);
    orc_amux(neq, qhat, t, a0, ja0, ia0);
    double numerator = orc_ddot(neq, t, q);
    denominator = orc_ddot(neq, t, t);
    denominator = denominator + fsign(DPREC, denominator);
    omega = numerator / denominator;
    double deltax = DZERO, rmax = DZERO, l2norm = DZERO;
    int xloc = -1, rloc = -1;
    for (int n = 0; n < neq; n++) {
      double tv = alpha * phat[n] + omega * qhat[n];
      x[n] = x[n] + tv;
      if (dscale) tv = tv * dscale[n];   /* ImsLinearBase.f90:452-454 */
      if (fabs(tv) > fabs(deltax)) { deltax = tv; xloc = n; }
      tv = q[n] - omega * t[n];
      d[n] = tv;
      if (dscale) tv = tv / dscale[n];   /* ImsLinearBase.f90:467-469 */
      if (fabs(tv) > fabs(rmax)) { rmax = tv; rloc = n; }
      l2norm = l2norm + tv * tv;
    }
    l2norm = sqrt(l2norm);
    if (hist) {
      orc_iter_rec *h = &hist[iiter - 1];
      h->dvmax = deltax; h->rmax = rmax; h->l2norm = l2norm;
      h->alpha = alpha; h->omega = omega; h->locdv = xloc; h->locr = rloc;
    }
    double rcnvg = (icnvgopt == 2 || icnvgopt == 3 || icnvgopt == 4) ? l2norm : rmax;
    orc_testcnvg(icnvgopt, icnvg, *innerit, deltax, rcnvg, l2norm0, epfact,
                 dvclose, rclose);
    if (rcnvg == DZERO) *icnvg = 1;
    if (*icnvg != 0) break;
    if (orc_is_close(rho, rho0)) break;
    if (orc_is_close(alpha, alpha0)) break;
    if (orc_is_close(omega, omega0)) break;
    if (north > 0 && ((iiter + 1) % north) == 0)
      orc_residual(neq, x, b, d, a0, ia0, ja0);
    if (rho * omega == DZERO) break;
    rho0 = rho;
    alpha0 = alpha;
    omega0 = omega;
  }
  if (*icnvg < 0) *icnvg = 0;
}

/* ImsLinearBase.f90:619-754 ims_base_scale.  iopt 0 scales: iscl 1 takes
 * dscale = dscale2 = 1/sqrt|a_nn| and a_nm = (c_n a_nm) c_m; iscl 2 scales
 * each row to unit L2 norm (dscale, 1 for a zero row), then each column of
 * the row-scaled matrix (dscale2, the squares summed row by row in CSR
 * order).  Both then set x = x / dscale2 and b = b dscale.  iopt 1 undoes
 * it: a_nm = ((1/c_n) a_nm) (1/c_m), x = x dscale2, b = b / dscale. */
void orc_ims_scale(int iopt, int iscl, int neq, const int *ia, const int *ja,
                   double *amat, double *x, double *b, double *dscale,
                   double *dscale2)
{
  if (iopt == 0) {
    if (iscl == 1) {
      for (int n = 0; n < neq; n++) {
        double c1 = DONE / sqrt(fabs(amat[ia[n]]));
        dscale[n] = c1;
        dscale2[n] = c1;
      }
      for (int n = 0; n < neq; n++) {
        double c1 = dscale[n];
        for (int i = ia[n]; i < ia[n + 1]; i++)
          amat[i] = c1 * amat[i] * dscale2[ja[i]];
      }
    } else if (iscl == 2) {
      for (int n = 0; n < neq; n++) {
        double c1 = DZERO;
        for (int i = ia[n]; i < ia[n + 1]; i++) c1 = c1 + amat[i] * amat[i];
        c1 = sqrt(c1);
        c1 = (c1 == DZERO) ? DONE : DONE / c1;
        dscale[n] = c1;
        for (int i = ia[n]; i < ia[n + 1]; i++) amat[i] = c1 * amat[i];
      }
      for (int n = 0; n < neq; n++) dscale2[n] = DZERO;
      for (int n = 0; n < neq; n++)
        for (int i = ia[n]; i < ia[n + 1]; i++)
          dscale2[ja[i]] = dscale2[ja[i]] + amat[i] * amat[i];
      for (int n = 0; n < neq; n++) {
        double c2 = dscale2[n];
        dscale2[n] = (c2 == DZERO) ? DONE : DONE / sqrt(c2);
      }
      for (int n = 0; n < neq; n++)
        for (int i = ia[n]; i < ia[n + 1]; i++)
          amat[i] = dscale2[ja[i]] * amat[i];
    }
    for (int n = 0; n < neq; n++) {
      x[n] = x[n] / dscale2[n];
      b[n] = b[n] * dscale[n];
    }
  } else {
    for (int n = 0; n < neq; n++) {
      double c1 = dscale[n];
      for (int i = ia[n]; i < ia[n + 1]; i++)
        amat[i] = (DONE / c1) * amat[i] * (DONE / dscale2[ja[i]]);
      x[n] = x[n] * dscale2[n];
      b[n] = b[n] / c1;
    }
  }
}

/* ImsLinear.f90:617-750 imslinear_ap without reordering (IORD = 0): epfact,
 * scaling (iscl 1 or 2: A, x and b in place, undone at the end with the
 * reference's rounding; dscale and dscale2 neq doubles each), pcu, residual,
 * L2 norm, then CG or BCGS (BCGS measures its dx and residual maxima in
 * unscaled units, ImsLinearBase.f90:452-469).  work must hold 9*neq doubles
 * (BCGS) or 4*neq (CG); iwork neq ints; apc nja doubles; iapc neq+1, japc
 * nja ints (already filled by orc_pccrs).  Returns the number of inner
 * iterations; *l2norm0_out gets the initial residual norm, *ilu_stop the
 * early-exit row of the factor. */
int orc_ims_ap_scl(int neq, int nja, const int *ia, const int *ja,
                   double *amat, double *rhs, double *x, int ilinmeth,
                   int ipc, int icnvgopt, int north, int iter1, double dvclose,
                   double rclose, double relax, int kstp, int iscl,
                   double *dscale, double *dscale2, double *apc,
                   const int *iapc, const int *japc, int *iwork, double *work,
                   int *icnvg, double *l2norm0_out, int *ilu_stop,
                   orc_iter_rec *hist)
{
  double epfact = orc_epfact(icnvgopt, kstp);
  if (iscl != 0)
    orc_ims_scale(0, iscl, neq, ia, ja, amat, x, rhs, dscale, dscale2);
  int stop = orc_pcu(nja, neq, ipc, relax, amat, ia, ja, apc, iapc, japc,
                     iwork, work);
  if (ilu_stop) *ilu_stop = stop;
  double *d = work, *p = work + neq, *q = work + 2 * (size_t)neq,
         *z = work + 3 * (size_t)neq;
  for (int n = 0; n < neq; n++) { d[n] = 0; p[n] = 0; q[n] = 0; z[n] = 0; }
  *icnvg = 0;
  orc_residual(neq, x, rhs, d, amat, ia, ja);
  double l2norm0 = orc_dnrm2(neq, d);
  if (l2norm0_out) *l2norm0_out = l2norm0;
  int itmax = iter1;
  if (l2norm0 == DZERO) { itmax = 0; *icnvg = 1; }
  int innerit = 0;
  if (ilinmeth == 1) {
    orc_cg(icnvg, itmax, &innerit, neq, nja, icnvgopt, north, dvclose, rclose,
           l2norm0, epfact, ia, ja, amat, iapc, japc, apc, x, rhs, d, p, q, z,
           hist);
  } else {
    double *t = work + 4 * (size_t)neq, *v = work + 5 * (size_t)neq,
           *dhat = work + 6 * (size_t)neq, *phat = work + 7 * (size_t)neq,
           *qhat = work + 8 * (size_t)neq;
    for (int n = 0; n < neq; n++) { t[n] = 0; v[n] = 0; dhat[n] = 0; phat[n] = 0; qhat[n] = 0; }
    orc_bcgs(icnvg, itmax, &innerit, neq, nja, icnvgopt, north, dvclose,
             rclose, l2norm0, epfact, ia, ja, amat, iapc, japc, apc, x, rhs, d,
             p, q, t, v, dhat, phat, qhat, iscl != 0 ? dscale : NULL, hist);
  }
  if (iscl != 0)
    orc_ims_scale(1, iscl, neq, ia, ja, amat, x, rhs, dscale, dscale2);
  return innerit;
}

/* orc_ims_ap_scl without scaling (ISCL = 0): A and b are left as given. */
int orc_ims_ap(int neq, int nja, const int *ia, const int *ja,
               const double *amat, const double *rhs, double *x, int ilinmeth,
               int ipc, int icnvgopt, int north, int iter1, double dvclose,
               double rclose, double relax, int kstp, double *apc,
               const int *iapc, const int *japc, int *iwork, double *work,
               int *icnvg, double *l2norm0_out, int *ilu_stop,
               orc_iter_rec *hist)
{
  return orc_ims_ap_scl(neq, nja, ia, ja, (double *)amat, (double *)rhs, x,
                        ilinmeth, ipc, icnvgopt, north, iter1, dvclose, rclose,
                        relax, kstp, 0, NULL, NULL, apc, iapc, japc, iwork,
                        work, icnvg, l2norm0_out, ilu_stop, hist);
}

/* ------------------------------------------------------------------ */
/* Stateless GWF physics helpers                                       */
/* ------------------------------------------------------------------ */

/* src/Utilities/SmoothingFunctions.f90:275-324 sQuadraticSaturation */
double orc_sqsat(double top, double bot, double x, double eps)
{
  double teps = eps, y;
  double b = top - bot;
  if (b > DZERO) {
    double br;
    if (x < bot) br = DZERO;
    else if (x > top) br = DONE;
    else br = (x - bot) / b;
    double av = DONE / (DONE - teps);
    double bri = DONE - br;
    if (br < teps) y = av * DHALF * (br * br) / teps;
    else if (br < (DONE - teps)) y = av * br + DHALF * (DONE - av);
    else if (br < DONE) y = DONE - ((av * DHALF * (bri * bri)) / teps);
    else y = DONE;
  } else {
    y = (x < bot) ? DZERO : DONE;
  }
  return y;
}

/* src/Utilities/SmoothingFunctions.f90:364-406
 * sQuadraticSaturationDerivative */
double orc_sqsat_deriv(double top, double bot, double x, double eps)
{
  double teps = eps, y, br;
  double b = top - bot;
  if (x < bot) br = DZERO;
  else if (x > top) br = DONE;
  else br = (x - bot) / b;
  double av = DONE / (DONE - teps);
  double bri = DONE - br;
  if (br < teps) y = av * br / teps;
  else if (br < (DONE - teps)) y = av;
  else if (br < DONE) y = av * bri / teps;
  else y = DZERO;
  y = y / b;
  return y;
}

/* src/Utilities/SmoothingFunctions.f90:412-460 sQSaturation (cubic
 * smoothing from bot to top, coefficients c1 / b^3 and c2 / b^2; the
 * reference's defaults are c1 = -2, c2 = 3).  The reference's x**3.0 and
 * x**2.0 compile to (x x) x and x x, which these products reproduce. */
double orc_sqsat_cubic(double top, double bot, double x, double c1, double c2)
{
  double w = x - bot;
  double b = top - bot;
  double s = w / b;
  double cof1 = c1 / ((b * b) * b);
  double cof2 = c2 / (b * b);
  if (s < DZERO) return DZERO;
  if (s < DONE) return cof1 * ((w * w) * w) + cof2 * (w * w);
  return DONE;
}

/* SmoothingFunctions.f90:466-516 sQSaturationDerivative */
double orc_sqsat_cubic_deriv(double top, double bot, double x, double c1, double c2)
{
  double w = x - bot;
  double b = top - bot;
  double s = w / b;
  double cof1 = c1 * 3.0 / ((b * b) * b);
  double cof2 = c2 * 2.0 / (b * b);
  if (s < DZERO) return DZERO;
  if (s < DONE) return cof1 * (w * w) + cof2 * w;
  return DZERO;
}

/* GwfConductanceUtils.f90:290-305 logmean */
static double orc_logmean(double d1, double d2)
{
  double drat = d2 / d1;
  if (drat <= DLNLOW || drat >= DLNHIGH) return (d2 - d1) / log(drat);
  return DHALF * (d1 + d2);
}

/* GwfConductanceUtils.f90:226-284 condmean */
double orc_condmean(double k1, double k2, double thick1, double thick2,
                    double cl1, double cl2, double width, int iavgmeth)
{
  double t1 = k1 * thick1;
  double t2 = k2 * thick2;
  double c = DZERO;
  switch (iavgmeth) {
  case 0: /* CCOND_HMEAN */
    if (t1 * t2 > DZERO) c = width * t1 * t2 / (t1 * cl2 + t2 * cl1);
    else c = DZERO;
    break;
  case 1: { /* CCOND_LMEAN */
    double tmean = (t1 * t2 > DZERO) ? orc_logmean(t1, t2) : DZERO;
    c = tmean * width / (cl1 + cl2);
    break;
  }
  case 2: { /* CCOND_AMTLMK */
    double kmean = (k1 * k2 > DZERO) ? orc_logmean(k1, k2) : DZERO;
    c = kmean * DHALF * (thick1 + thick2) * width / (cl1 + cl2);
    break;
  }
  case 3: { /* CCOND_AMTHMK */
    double denom = (k1 * cl2 + k2 * cl1);
    double kmean = denom > DZERO ? k1 * k2 / denom : DZERO;
    c = kmean * DHALF * (thick1 + thick2) * width;
    break;
  }
  }
  return c;
}

/* GwfConductanceUtils.f90:43-143 hcond (ihc 1, no staggered connections). */
double orc_hcond(int ibdn, int ibdm, int ictn, int ictm, int iupstream,
                 int ihc, int icellavg, double condsat, double hn, double hm,
                 double satn, double satm, double hkn, double hkm,
                 double topn, double topm, double botn, double botm,
                 double cln, double clm, double fawidth)
{
  (void)ihc;
  if (ibdn == 0 || ibdm == 0) return DZERO;
  if (ictn == 0 && ictm == 0) return condsat;
  if (iupstream == 1) {
    double sat_up = (hn > hm) ? satn : satm;
    return sat_up * condsat;
  }
  double thksatn = satn * (topn - botn);
  double thksatm = satm * (topm - botm);
  return orc_condmean(hkn, hkm, thksatn, thksatm, cln, clm, fawidth, icellavg);
}

/* GwfConductanceUtils.f90:149-222 vcond */
double orc_vcond(int ibdn, int ibdm, int ictn, int ictm, int inewton,
                 int ivarcv, int idewatcv, double condsat, double hn,
                 double hm, double vkn, double vkm, double satn, double satm,
                 double topn, double topm, double botn, double botm,
                 double flowarea)
{
  (void)inewton;
  if (ibdn == 0 || ibdm == 0) return DZERO;
  if (ivarcv == 0) return condsat;
  if (ictn == 0 && ictm == 0) return condsat;
  if (hn >= topn && hm >= topm) return condsat;
  double satntmp = satn, satmtmp = satm;
  if (idewatcv == 0) {
    if (botn > botm) satmtmp = DONE;
    else satntmp = DONE;
  }
  double bovk1 = satntmp * (topn - botn) * DHALF / vkn;
  double bovk2 = satmtmp * (topm - botm) * DHALF / vkm;
  double denom = (bovk1 + bovk2);
  return denom != DZERO ? flowarea / denom : DZERO;
}

/* GwfStorageUtils.f90:143-160 SsCapacity */
double orc_sscapacity(int istor_coef, double top, double bot, double area,
                      double ss)
{
  double thick = istor_coef == 0 ? top - bot : DONE;
  return ss * thick * area;
}

/* GwfStorageUtils.f90:169-180 SyCapacity */
double orc_sycapacity(double area, double sy) { return sy * area; }

/* GwfStorageUtils.f90:32-91 SsTerms (rate not needed) */
void orc_ssterms(int iconvert, int iorig_ss, int iconf_ss, double top,
                 double bot, double rho1, double rho1old, double snnew,
                 double snold, double hnew, double hold, double *aterm,
                 double *rhsterm)
{
  (void)hnew;
  *aterm = -rho1 * snnew;
  *rhsterm = DZERO;
  if (iconvert != 0) {
    if (iorig_ss == 0) {
      if (iconf_ss == 0) {
        double tthk = top - bot;
        double zold = bot + DHALF * tthk * snold;
        double znew = bot + DHALF * tthk * snnew;
        *rhsterm = -rho1old * snold * (hold - zold) - rho1 * snnew * znew;
      } else {
        if (snold == DONE) *rhsterm = *rhsterm - rho1old * (hold - top);
        if (snnew == DONE) *rhsterm = *rhsterm - rho1 * top;
        else *aterm = DZERO;
      }
    } else {
      *rhsterm = -rho1old * snold * hold;
    }
  } else {
    *rhsterm = -rho1old * snold * hold;
  }
}

/* GwfStorageUtils.f90:98-137 SyTerms (rate not needed) */
void orc_syterms(double top, double bot, double rho2, double rho2old,
                 double snnew, double snold, double *aterm, double *rhsterm)
{
  *aterm = DZERO;
  double tthk = top - bot;
  if (snnew < DONE) {
    if (snnew > DZERO) {
      *aterm = -rho2;
      *rhsterm = -rho2old * tthk * snold - rho2 * bot;
    } else {
      *rhsterm = tthk * (DZERO - rho2old * snold);
    }
  } else {
    *rhsterm = tthk * (rho2 * snnew - rho2old * snold);
  }
}

/* ------------------------------------------------------------------ */
/* GWF formulate (scatter form, reference statement order)             */
/* ------------------------------------------------------------------ */

/* Model arrays of one structured GWF model, reference naming. */
typedef struct {
  int nodes, nja;
  const int *ia, *ja, *jas, *isym, *ihc;     /* con%... (0-based) */
  const double *cl1, *cl2, *hwva;            /* by jas */
  const double *top, *bot, *area;
  const double *k11, *k33;
  const int *icelltype;
  int *ibound;
  double *condsat;                           /* by jas */
  double *sat;
  int inewton, ivarcv, idewatcv, icellavg;
  double satomega;                           /* NPF: 1e-6 under Newton */
} orc_npf;

/* gwf-hfb.f90:789-832 condsat_modify: barrier b joins noden[b] and
 * nodem[b] (0-based) through idxloc[b], the ja position of nodem in row
 * noden (check_data, :711-766).  In barrier order (a second barrier on a
 * face combines with the first's result), csatsav[b] keeps the face's
 * condsat before the barrier; under NEWTON or when both cells are confined
 * the face's condsat becomes cond condhfb / (cond + condhfb), condhfb =
 * hydchr fawidth faheight, or -cond hydchr for hydchr <= 0. */
void orc_hfb_condsat_modify(int nhfb, const int *noden, const int *nodem,
                            const int *idxloc, const double *hydchr, int inewton,
                            const int *icelltype, const int *jas, const int *ihc,
                            const double *hwva, const double *top, const double *bot,
                            double *condsat, double *csatsav)
{
  for (int b = 0; b < nhfb; b++) {
    int js = jas[idxloc[b]];
    double cond = condsat[js];
    csatsav[b] = cond;
    int n = noden[b], m = nodem[b];
    if (inewton == 1 || (icelltype[n] == 0 && icelltype[m] == 0)) {
      double topn = top[n], topm = top[m], botn = bot[n], botm = bot[m];
      double faheight;
      if (ihc[js] == 2)
        faheight = (topn < topm ? topn : topm) - (botn > botm ? botn : botm);
      else
        faheight = DHALF * ((topn - botn) + (topm - botm));
      if (hydchr[b] > DZERO) {
        double fawidth = hwva[js];
        double condhfb = hydchr[b] * fawidth * faheight;
        cond = cond * condhfb / (cond + condhfb);
      } else {
        cond = -cond * hydchr[b];
      }
      condsat[js] = cond;
    }
  }
}

/* gwf-hfb.f90:770-781 condsat_reset: each barrier's face back to the value
 * saved before it, in barrier order */
void orc_hfb_condsat_reset(int nhfb, const int *idxloc, const int *jas,
                           const double *csatsav, double *condsat)
{
  for (int b = 0; b < nhfb; b++) condsat[jas[idxloc[b]]] = csatsav[b];
}

/* src/Utilities/HGeoUtil.f90:29-108 hyeff: the conductivity along the unit
 * vector vg of the tensor with principal values k11, k22, k33 rotated by
 * ang1, ang2, ang3 (radians); iavgmeth 0 harmonic (the NPF default), 1
 * arithmetic.  The reference's ve**2 compiles to ve * ve. */
double orc_hyeff(double k11, double k22, double k33, double ang1, double ang2,
                 double ang3, double vg1, double vg2, double vg3, int iavgmeth)
{
  /* each sin / cos pair from one sincos call: the reference's compiled
   * hyeff merges them so, and sincos can differ from separate sin and cos
   * calls in the last bit */
  double s1, c1, s2, c2, s3, c3;
  sincos(ang1, &s1, &c1);
  sincos(ang2, &s2, &c2);
  sincos(ang3, &s3, &c3);
  double r11 = c1 * c2;
  double r12 = c1 * s2 * s3 - s1 * c3;
  double r13 = -c1 * s2 * c3 - s1 * s3;
  double r21 = s1 * c2;
  double r22 = s1 * s2 * s3 + c1 * c3;
  double r23 = -s1 * s2 * c3 + c1 * s3;
  double r31 = s2;
  double r32 = -c2 * s3;
  double r33 = c2 * c3;
  double ve1 = r11 * vg1 + r21 * vg2 + r31 * vg3;
  double ve2 = r12 * vg1 + r22 * vg2 + r32 * vg3;
  double ve3 = r13 * vg1 + r23 * vg2 + r33 * vg3;
  double k = DZERO;
  if (iavgmeth == 0) {
    double dnum = DONE;
    double d1 = ve1 * ve1, d2 = ve2 * ve2, d3 = ve3 * ve3;
    if (ve1 != DZERO) { dnum = dnum * k11; d2 = d2 * k11; d3 = d3 * k11; }
    if (ve2 != DZERO) { dnum = dnum * k22; d1 = d1 * k22; d3 = d3 * k22; }
    if (ve3 != DZERO) { dnum = dnum * k33; d1 = d1 * k33; d2 = d2 * k33; }
    double denom = d1 + d2 + d3;
    if (denom > DZERO) k = dnum / denom;
  } else if (iavgmeth == 1) {
    k = ve1 * ve1 * k11 + ve2 * ve2 * k22 + ve3 * ve3 * k33;
  }
  return k;
}

/* gwf-hfb.f90:211-348 hfb_fc without NEWTON or XT3D: each barrier next to a
 * convertible cell (ICELLTYPE 1), in barrier order, replaces the matrix
 * value aterm at (n, m) and (m, n) by its series combination with the
 * barrier, cond = aterm condhfb / (aterm + condhfb), condhfb = hydchr w h
 * with h the mean wetted thickness of the two cells (tops clipped at the
 * heads of convertible cells), or -hydchr aterm for hydchr <= 0, and adds
 * aterm - cond to both diagonals; condsav keeps cond for hfb_cq.  idxloc:
 * the CSR position of (n, m) in row n. */
void orc_hfb_fc(int nhfb, const int *noden, const int *nodem, const int *idxloc,
                const double *hydchr, const int *ibound, const int *icelltype,
                const int *ia, const int *isym, const int *jas, const double *hwva,
                const double *top, const double *bot, const double *hnew,
                double *amat, double *condsav)
{
  for (int b = 0; b < nhfb; b++) {
    int ipos = idxloc[b];
    double aterm = amat[ipos];
    int n = noden[b], m = nodem[b];
    if (ibound[n] == 0 || ibound[m] == 0) continue;
    if (icelltype[n] != 1 && icelltype[m] != 1) continue;
    double topn = top[n], topm = top[m], botn = bot[n], botm = bot[m];
    if (icelltype[n] == 1 && hnew[n] < topn) topn = hnew[n];
    if (icelltype[m] == 1 && hnew[m] < topm) topm = hnew[m];
    double faheight = DHALF * ((topn - botn) + (topm - botm));
    double cond;
    if (hydchr[b] > DZERO) {
      double fawidth = hwva[jas[ipos]];
      double condhfb = hydchr[b] * DONE * fawidth * faheight;
      cond = aterm * condhfb / (aterm + condhfb);
    } else {
      cond = -aterm * hydchr[b];
    }
    condsav[b] = cond;
    amat[ia[n]] = amat[ia[n]] + (aterm - cond);
    amat[ipos] = cond;
    amat[ia[m]] = amat[ia[m]] + (aterm - cond);
    amat[isym[ipos]] = cond;
  }
}

/* gwf-hfb.f90:355-450 hfb_cq without NEWTON: the flow across such a barrier
 * is condsav (h(m) - h(n)) at (n, m) and its negative at (m, n) */
void orc_hfb_cq(int nhfb, const int *noden, const int *nodem, const int *idxloc,
                const int *isym, const int *ibound, const int *icelltype,
                const double *hnew, const double *condsav, double *flowja)
{
  for (int b = 0; b < nhfb; b++) {
    int n = noden[b], m = nodem[b];
    if (ibound[n] == 0 || ibound[m] == 0) continue;
    if (icelltype[n] != 1 && icelltype[m] != 1) continue;
    double qnm = condsav[b] * (hnew[m] - hnew[n]);
    flowja[idxloc[b]] = qnm;
    flowja[isym[idxloc[b]]] = -qnm;
  }
}

/* gwf-npf.f90:1950-2037 calc_condsat (upperOnly, saturation 1, no
 * THICKSTRT) for every node. */
void orc_calc_condsat(int nodes, const int *ia, const int *ja, const int *jas,
                      const int *ihc, const double *cl1, const double *cl2,
                      const double *hwva, const double *top, const double *bot,
                      const double *k11, const double *k22, int ncol,
                      const double *k33, int icellavg, const double *sat0,
                      const double *hn0, double *condsat)
{
  for (int node = 0; node < nodes; node++) {
    for (int ii = ia[node] + 1; ii < ia[node + 1]; ii++) {
      int m = ja[ii];
      if (m < node) continue;
      int n = node;
      int jj = jas[ii];
      /* calc_initial_sat and THICKSTRT's heads (:1997-2006): sat0 / hn0
       * are 1 and the cell top except for THICKSTRT cells */
      double satn = sat0[n], satm = sat0[m];
      double topn = top[n], botn = bot[n], topm = top[m], botm = bot[m];
      int ihcj = ihc[jj];
      double hyn = ihcj == 0 ? k33[n] : (abs(m - n) == ncol ? k22[n] : k11[n]);
      double hym = ihcj == 0 ? k33[m] : (abs(m - n) == ncol ? k22[m] : k11[m]);
      double hn = hn0[n], hm = hn0[m];
      double csat;
      if (ihcj == 0) {
        csat = orc_vcond(1, 1, 1, 1, 0, 1, 1, DONE, botn, botm, hyn, hym,
                         satn, satm, topn, topm, botn, botm, hwva[jj]);
      } else {
        csat = orc_hcond(1, 1, 1, 1, 0, ihcj, icellavg, DONE, hn, hm, satn,
                         satm, hyn, hym, topn, topm, botn, botm, cl1[jj],
                         cl2[jj], hwva[jj]);
      }
      condsat[jj] = csat;
    }
  }
}

/* gwf-npf.f90:444-470 npf_cf with sgwf_npf_thksat :775-794 (no wet/dry). */
void orc_npf_cf(int nodes, int inewton, double satomega, const int *icelltype,
                const int *ibound, const double *top, const double *bot,
                const double *hnew, double *sat)
{
  for (int n = 0; n < nodes; n++) {
    if (icelltype[n] != 0) {
      double satn;
      if (ibound[n] == 0) {
        satn = DZERO;
      } else {
        if (hnew[n] >= top[n]) satn = DONE;
        else satn = (hnew[n] - bot[n]) / (top[n] - bot[n]);
        if (inewton != 0) satn = orc_sqsat(top[n], bot[n], hnew[n], satomega);
      }
      sat[n] = satn;
    }
  }
}

/* gwf-npf.f90:474-574 npf_fc, idxglo = identity (one DIS model), no XT3D,
 * no perched flow. */
void orc_npf_fc(int nodes, const int *ia, const int *ja, const int *jas,
                const int *isym, const int *ihc, const double *cl1,
                const double *cl2, const double *hwva, const double *condsat,
                const int *ibound, const int *icelltype, int inewton,
                int ivarcv, int idewatcv, int icellavg, const double *k11,
                const double *k22, int ncol, const double *k33,
                const double *top, const double *bot,
                const double *sat, const double *hnew, double *amat)
{
  for (int n = 0; n < nodes; n++) {
    for (int ii = ia[n] + 1; ii < ia[n + 1]; ii++) {
      int m = ja[ii];
      if (m < n) continue;
      int jj = jas[ii];
      int ihcj = ihc[jj];
      double hyn = ihcj == 0 ? k33[n] : (abs(m - n) == ncol ? k22[n] : k11[n]);
      double hym = ihcj == 0 ? k33[m] : (abs(m - n) == ncol ? k22[m] : k11[m]);
      double cond;
      if (ihcj == 0) {
        cond = orc_vcond(ibound[n], ibound[m], icelltype[n], icelltype[m],
                         inewton, ivarcv, idewatcv, condsat[jj], hnew[n],
                         hnew[m], hyn, hym, sat[n], sat[m], top[n], top[m],
                         bot[n], bot[m], hwva[jj]);
      } else {
        cond = orc_hcond(ibound[n], ibound[m], icelltype[n], icelltype[m],
                         inewton, ihcj, icellavg, condsat[jj], hnew[n],
                         hnew[m], sat[n], sat[m], hyn, hym, top[n], top[m],
                         bot[n], bot[m], cl1[jj], cl2[jj], hwva[jj]);
      }
      int idiag = ia[n];
      amat[ii] = amat[ii] + cond;
      amat[idiag] = amat[idiag] + (-cond);
      int isymcon = isym[ii];
      int idiagm = ia[m];
      amat[isymcon] = amat[isymcon] + cond;
      amat[idiagm] = amat[idiagm] + (-cond);
    }
  }
}

/* gwf-npf.f90:578-698 npf_fn (no XT3D); the get/set_value_pos pairs of the
 * reference write back the value they read and are omitted. */
void orc_npf_fn(int nodes, const int *ia, const int *ja, const int *jas,
                const int *isym, const int *ihc, const double *condsat,
                const int *ibound, const int *icelltype, int ivarcv,
                const double *top, const double *bot, double satomega,
                const double *hnew, double *amat, double *rhs)
{
  for (int n = 0; n < nodes; n++) {
    int idiag = ia[n];
    for (int ii = ia[n] + 1; ii < ia[n + 1]; ii++) {
      int m = ja[ii];
      int isymcon = isym[ii];
      if (m < n) continue;
      int jj = jas[ii];
      if (ihc[jj] == 0 && ivarcv == 0) continue;
      int iups = m;
      if (hnew[m] < hnew[n]) iups = n;
      int idn = n;
      if (iups == n) idn = m;
      if (icelltype[iups] == 0) continue;
      double topup = top[iups];
      double botup = bot[iups];
      if (ihc[jj] == 2) {
        topup = fmin(top[n], top[m]);
        botup = fmax(bot[n], bot[m]);
      }
      double cond = condsat[jj];
      double consterm = -cond * (hnew[iups] - hnew[idn]);
      double derv = orc_sqsat_deriv(topup, botup, hnew[iups], satomega);
      int idiagm = ia[m];
      if (iups == n) {
        double term = consterm * derv;
        rhs[n] = rhs[n] + term * hnew[n];
        rhs[m] = rhs[m] - term * hnew[n];
        amat[idiag] = amat[idiag] + term;
        if (ibound[m] > 0) amat[isymcon] = amat[isymcon] + (-term);
      } else {
        double term = -consterm * derv;
        rhs[n] = rhs[n] + term * hnew[m];
        rhs[m] = rhs[m] - term * hnew[m];
        if (ibound[n] > 0) amat[ii] = amat[ii] + term;
        amat[idiagm] = amat[idiagm] + (-term);
      }
    }
  }
}

/* gwf-sto.f90:226-345 sto_fc (transient only; istor_coef, iorig_ss,
 * iconf_ss and integratechanges all 0).  satomega is STO's own: 1e-6 when the
 * model uses NEWTON (source_options, gwf-sto.f90:826-829), else 0. */
void orc_sto_fc(int nodes, const int *ia, const int *ibound,
                const int *iconvert, const double *top, const double *bot,
                const double *area, const double *ss, const double *sy,
                int istor_coef, int iconf_ss, int iorig_ss,
                double delt, double satomega, const double *hold,
                const double *hnew, double *amat, double *rhs)
{
  double tled = DONE / delt;
  for (int n = 0; n < nodes; n++) {
    int idiag = ia[n];
    if (ibound[n] < 1) continue;
    double tp = top[n], bt = bot[n];
    double snold, snnew;
    if (iconvert[n] == 0) { snold = DONE; snnew = DONE; }
    else {
      snold = orc_sqsat(tp, bt, hold[n], satomega);
      snnew = orc_sqsat(tp, bt, hnew[n], satomega);
    }
    double sc1 = orc_sscapacity(istor_coef, tp, bt, area[n], ss[n]);
    double rho1 = sc1 * tled;
    double rho1old = rho1;
    double aterm, rhsterm;
    orc_ssterms(iconvert[n], iorig_ss, iconf_ss, tp, bt, rho1, rho1old, snnew, snold,
                hnew[n], hold[n], &aterm, &rhsterm);
    amat[idiag] = amat[idiag] + aterm;
    rhs[n] = rhs[n] + rhsterm;
    if (iconvert[n] != 0) {
      rhsterm = DZERO;
      double sc2 = orc_sycapacity(area[n], sy[n]);
      double rho2 = sc2 * tled;
      double rho2old = rho2;
      orc_syterms(tp, bt, rho2, rho2old, snnew, snold, &aterm, &rhsterm);
      amat[idiag] = amat[idiag] + aterm;
      rhs[n] = rhs[n] + rhsterm;
    }
  }
}

/* gwf-sto.f90:353-445 sto_fn.  sQuadraticSaturation and its derivative are
 * called without eps here, so the default 1e-6 applies (not STO's 0). */
void orc_sto_fn(int nodes, const int *ia, const int *ibound,
                const int *iconvert, const double *top, const double *bot,
                const double *area, const double *ss, const double *sy,
                int istor_coef, int iconf_ss, int iorig_ss,
                double delt, const double *hnew, double *amat, double *rhs)
{
  double tled = DONE / delt;
  for (int n = 0; n < nodes; n++) {
    int idiag = ia[n];
    if (ibound[n] <= 0) continue;
    double tp = top[n], bt = bot[n];
    double tthk = tp - bt;
    double h = hnew[n];
    double snnew = orc_sqsat(tp, bt, h, DEM6);
    double sc1 = orc_sscapacity(istor_coef, tp, bt, area[n], ss[n]);
    double sc2 = orc_sycapacity(area[n], sy[n]);
    double rho1 = sc1 * tled;
    double rho2 = sc2 * tled;
    if (iconvert[n] != 0) {
      double derv = orc_sqsat_deriv(tp, bt, h, DEM6);
      /* iconf_ss == 0, iorig_ss == 0 */
      double drterm;
      if (iconf_ss == 0) {
        if (iorig_ss == 0) drterm = -rho1 * derv * (h - bt) + rho1 * tthk * snnew * derv;
        else drterm = -(rho1 * derv * h);
        amat[idiag] = amat[idiag] + drterm;
        rhs[n] = rhs[n] + drterm * h;
      }
      if (snnew < DONE) {
        if (snnew > DZERO) {
          double rterm = -rho2 * tthk * snnew;
          drterm = -rho2 * tthk * derv;
          amat[idiag] = amat[idiag] + (drterm + rho2);
          rhs[n] = rhs[n] - rterm + drterm * h + rho2 * bt;
        }
      }
    }
  }
}

/* BoundaryPackage.f90:453-479 bnd_fc (also wel_fc / drn_fc / rch_fc without
 * mover).  nodelist entries < 0 are skipped (rch_fc :370). */
void orc_bnd_fc(int nbound, const int *nodelist, const double *hcof,
                const double *brhs, const int *ia, double *amat, double *rhs)
{
  for (int i = 0; i < nbound; i++) {
    int n = nodelist[i];
    if (n < 0) continue;
    rhs[n] = rhs[n] + brhs[i];
    amat[ia[n]] = amat[ia[n]] + hcof[i];
  }
}

/* Ghost-node correction of one model in explicit form (GhostNode.f90).
 * GNC g corrects the flow between noden[g] and nodem[g] (0-based, connected
 * at CSR position ipos[g] of row n) with the heads of nodesj[g*numjs + i]
 * (-1: none) weighted by alphasj[g*numjs + i]. */

/* GhostNode.f90:280-324 gnc_fc (explicit) with gnc_fmsav :251-273: cond is
 * the matrix value at (n, m) when gnc_fc runs (the npf conductance), saved
 * for the budget; rhs(n) -= alpha cond (x(n) - x(j)), rhs(m) += the same. */
void orc_gnc_fc(int nexg, int numjs, const int *noden, const int *nodem,
                const int *nodesj, const double *alphasj, const int *ipos,
                const int *ibound, const double *x, const double *amat,
                double *cond, double *rhs)
{
  for (int g = 0; g < nexg; g++) cond[g] = amat[ipos[g]];
  for (int g = 0; g < nexg; g++) {
    int n = noden[g], m = nodem[g];
    if (ibound[n] == 0 || ibound[m] == 0) continue;
    double c = cond[g];
    for (int i = 0; i < numjs; i++) {
      int j = nodesj[g * numjs + i];
      if (j < 0) continue;
      double alpha = alphasj[g * numjs + i];
      if (alpha == DZERO) continue;
      double aterm = alpha * c;
      double rterm = aterm * (x[n] - x[j]);
      rhs[n] = rhs[n] - rterm;
      rhs[m] = rhs[m] + rterm;
    }
  }
}

/* GhostNode.f90:340-443 gnc_fn of one model: the derivative of the
 * correction with respect to the upstream head of a convertible upstream
 * cell (horizontal connections, or vertical ones with VARIABLECV), with the
 * saturated conductance of the face. */
void orc_gnc_fn(int nexg, int numjs, const int *noden, const int *nodem,
                const int *nodesj, const double *alphasj, const int *ia,
                const int *ipos, const int *isym, const int *jas,
                const int *ihc, const double *condsat, int ivarcv,
                const int *icelltype, const double *top, const double *bot,
                const int *ibound, const double *x, double *amat, double *rhs)
{
  for (int g = 0; g < nexg; g++) {
    int n = noden[g], m = nodem[g];
    if (ibound[n] == 0 || ibound[m] == 0) continue;
    int js = jas[ipos[g]];
    int hc = ihc[js];
    double csat = condsat[js];
    if (hc == 0 && ivarcv == 0) continue;
    int iups = x[m] > x[n] ? 1 : 0;
    int up = iups == 0 ? n : m;
    double topup = top[up], botup = bot[up], xup = x[up];
    if (icelltype[up] == 0) continue;
    for (int i = 0; i < numjs; i++) {
      int j = nodesj[g * numjs + i];
      if (j < 0) continue;
      if (ibound[j] == 0) continue;
      double alpha = alphasj[g * numjs + i];
      if (alpha == DZERO) continue;
      double consterm = csat * alpha * (x[n] - x[j]);
      double derv = orc_sqsat_deriv(topup, botup, xup, DEM6);
      double term = consterm * derv;
      if (iups == 0) {
        amat[ia[n]] = amat[ia[n]] + term;
        if (ibound[m] > 0) amat[isym[ipos[g]]] = amat[isym[ipos[g]]] - term;
        rhs[n] = rhs[n] + term * x[n];
        rhs[m] = rhs[m] - term * x[n];
      } else {
        amat[ia[m]] = amat[ia[m]] - term;
        if (ibound[n] > 0) amat[ipos[g]] = amat[ipos[g]] + term;
        rhs[n] = rhs[n] + term * x[m];
        rhs[m] = rhs[m] - term * x[m];
      }
    }
  }
}

/* GhostNode.f90:478-503 gnc_cq with deltaQgnc :507-543: the correction flow
 * (sum_j alpha_j) x(n) - sum_j alpha_j x(j), times the saved cond, added to
 * flowja at (n, m) and taken from (m, n). */
void orc_gnc_cq(int nexg, int numjs, const int *noden, const int *nodem,
                const int *nodesj, const double *alphasj, const int *ipos,
                const int *isym, const int *ibound, const double *x,
                const double *cond, double *flowja)
{
  for (int g = 0; g < nexg; g++) {
    int n = noden[g], m = nodem[g];
    double dq = DZERO;
    if (ibound[n] != 0 && ibound[m] != 0) {
      double sigalj = DZERO, hd = DZERO;
      for (int i = 0; i < numjs; i++) {
        int j = nodesj[g * numjs + i];
        if (j < 0) continue;
        if (ibound[j] == 0) continue;
        double alpha = alphasj[g * numjs + i];
        sigalj = sigalj + alpha;
        hd = hd + alpha * x[j];
      }
      double aterm = sigalj * x[n] - hd;
      dq = aterm * cond[g];
    }
    flowja[ipos[g]] = flowja[ipos[g]] + dq;
    flowja[isym[ipos[g]]] = flowja[isym[ipos[g]]] - dq;
  }
}

/* ------------------------------------------------------------------ */
/* GWF budget: gwf_cq + gwf_bd (gwf.f90:741-800)                        */
/* ------------------------------------------------------------------ */

/* gwf-npf.f90:745-771 npf_cq with sgwf_npf_qcalc (:797-862), no XT3D and no
 * perched flow: qnm = cond(n, m) (hm - hn) for m > n, flowja(ipos) = qnm,
 * flowja(isym) = -qnm (positive into the cell).  The conductance is the one
 * npf_fc computes from the same arguments (sat from the last npf_cf). */
void orc_npf_cq(int nodes, const int *ia, const int *ja, const int *jas,
                const int *isym, const int *ihc, const double *cl1,
                const double *cl2, const double *hwva, const double *condsat,
                const int *ibound, const int *icelltype, int inewton,
                int ivarcv, int idewatcv, int icellavg, const double *k11,
                const double *k22, int ncol, const double *k33,
                const double *top, const double *bot,
                const double *sat, const double *hnew, double *flowja)
{
  for (int n = 0; n < nodes; n++) {
    for (int ii = ia[n] + 1; ii < ia[n + 1]; ii++) {
      int m = ja[ii];
      if (m < n) continue;
      int jj = jas[ii];
      int ihcj = ihc[jj];
      double hyn = ihcj == 0 ? k33[n] : (abs(m - n) == ncol ? k22[n] : k11[n]);
      double hym = ihcj == 0 ? k33[m] : (abs(m - n) == ncol ? k22[m] : k11[m]);
      double cond;
      if (ihcj == 0) {
        cond = orc_vcond(ibound[n], ibound[m], icelltype[n], icelltype[m],
                         inewton, ivarcv, idewatcv, condsat[jj], hnew[n],
                         hnew[m], hyn, hym, sat[n], sat[m], top[n], top[m],
                         bot[n], bot[m], hwva[jj]);
      } else {
        cond = orc_hcond(ibound[n], ibound[m], icelltype[n], icelltype[m],
                         inewton, ihcj, icellavg, condsat[jj], hnew[n],
                         hnew[m], sat[n], sat[m], hyn, hym, top[n], top[m],
                         bot[n], bot[m], cl1[jj], cl2[jj], hwva[jj]);
      }
      double qnm = cond * (hnew[m] - hnew[n]);
      flowja[ii] = qnm;
      flowja[isym[ii]] = -qnm;
    }
  }
}

/* gwf-sto.f90:447-564 sto_cq (transient only; istor_coef, iorig_ss,
 * iconf_ss and integratechanges 0, STO satomega 0) with the rates of SsTerms
 * (GwfStorageUtils.f90:32-91: aterm hnew - rhsterm) and SyTerms (:98-137:
 * rho2old tthk snold - rho2 tthk snnew), each added to the diagonal. */
void orc_sto_cq(int nodes, const int *ia, const int *ibound,
                const int *iconvert, const double *top, const double *bot,
                const double *area, const double *ss, const double *sy,
                int istor_coef, int iconf_ss, int iorig_ss,
                double delt, double satomega, const double *hold,
                const double *hnew, double *flowja, double *strgss,
                double *strgsy)
{
  for (int n = 0; n < nodes; n++) { strgss[n] = DZERO; strgsy[n] = DZERO; }
  double tled = DONE / delt;
  for (int n = 0; n < nodes; n++) {
    if (ibound[n] <= 0) continue;
    double tp = top[n], bt = bot[n];
    double snold, snnew;
    if (iconvert[n] == 0) { snold = DONE; snnew = DONE; }
    else {
      snold = orc_sqsat(tp, bt, hold[n], satomega);
      snnew = orc_sqsat(tp, bt, hnew[n], satomega);
    }
    double sc1 = orc_sscapacity(istor_coef, tp, bt, area[n], ss[n]);
    double rho1 = sc1 * tled;
    double rho1old = rho1;
    double aterm, rhsterm;
    orc_ssterms(iconvert[n], iorig_ss, iconf_ss, tp, bt, rho1, rho1old, snnew, snold,
                hnew[n], hold[n], &aterm, &rhsterm);
    double rate = aterm * hnew[n] - rhsterm;
    strgss[n] = rate;
    int idiag = ia[n];
    flowja[idiag] = flowja[idiag] + rate;
    rate = DZERO;
    if (iconvert[n] != 0) {
      double sc2 = orc_sycapacity(area[n], sy[n]);
      double rho2 = sc2 * tled;
      double rho2old = rho2;
      double tthk = tp - bt;
      rate = rho2old * tthk * snold - rho2 * tthk * snnew;
    }
    strgsy[n] = rate;
    flowja[idiag] = flowja[idiag] + rate;
  }
}

/* BoundaryPackage.f90:583-619 bnd_cq_simrate without mover: rate hcof h -
 * rhs into active cells, added to the diagonal; simvals per boundary. */
void orc_bnd_cq(int nbound, const int *nodelist, const double *hcof,
                const double *brhs, const int *ibound, const double *hnew,
                const int *ia, double *flowja, double *simvals)
{
  for (int i = 0; i < nbound; i++) {
    int node = nodelist[i];
    double rrate = DZERO;
    if (node >= 0) {
      int idiag = ia[node];
      if (ibound[node] > 0) rrate = hcof[i] * hnew[node] - brhs[i];
      flowja[idiag] = flowja[idiag] + rrate;
    }
    simvals[i] = rrate;
  }
}

/* Sparse.f90:262-281 csr_diagsum */
void orc_csr_diagsum(int nodes, const int *ia, double *flowja)
{
  for (int n = 0; n < nodes; n++) {
    int iposdiag = ia[n];
    for (int ipos = ia[n] + 1; ipos < ia[n + 1]; ipos++)
      flowja[iposdiag] = flowja[iposdiag] + flowja[ipos];
  }
}

/* gwf-chd.f90:264-320 calc_chd_rate: the CHD rate of a cell is minus the
 * flow into it from its neighbours; in / out split over active neighbours */
void orc_chd_rate(int nbound, const int *nodelist, const int *ia,
                  const int *ja, const int *ibound, double *flowja,
                  double *simvals, double *ratein, double *rateout)
{
  for (int i = 0; i < nbound; i++) {
    int node = nodelist[i];
    int idiag = ia[node];
    double rate = DZERO, rin = DZERO, rout = DZERO;
    for (int ipos = ia[node] + 1; ipos < ia[node + 1]; ipos++) {
      double q = flowja[ipos];
      rate = rate - q;
      int n2 = ja[ipos];
      if (ibound[n2] > 0) {
        if (q < DZERO) rin = rin - q;
        else rout = rout + q;
      }
    }
    simvals[i] = rate;
    ratein[i] = rin;
    rateout[i] = rout;
    flowja[idiag] = flowja[idiag] + rate;
  }
}

/* Budget.f90:631-648 rate_accumulator */
void orc_rate_accumulator(int n, const double *flow, double *rin, double *rout)
{
  *rin = DZERO;
  *rout = DZERO;
  for (int i = 0; i < n; i++) {
    if (flow[i] < DZERO) *rout = *rout - flow[i];
    else *rin = *rin + flow[i];
  }
}

/* gwf-wel.f90:296-332 wel_cf.  With AUTO_FLOW_REDUCE (iflowred /= 0;
 * flowred as wel_options :205-217 leaves it, in (0, 1]) a withdrawal in a
 * convertible cell is scaled by sQSaturation over the bottom flowred of the
 * cell's thickness. */
void orc_wel_cf(int nbound, const int *nodelist, const double *q,
                const int *ibound, int iflowred, double flowred,
                const int *icelltype, const double *top, const double *bot,
                const double *xnew, double *hcof, double *brhs)
{
  for (int i = 0; i < nbound; i++) {
    int node = nodelist[i];
    hcof[i] = DZERO;
    if (ibound[node] <= 0) { brhs[i] = DZERO; continue; }
    double qi = q[i];
    if (iflowred != 0 && qi < DZERO && icelltype[node] != 0) {
      double tp = top[node], bt = bot[node];
      double thick = tp - bt;
      tp = bt + flowred * thick;
      qi = qi * orc_sqsat_cubic(tp, bt, xnew[node], -2.0, 3.0);
    }
    brhs[i] = -qi;
  }
}

/* gwf-wel.f90:378-424 wel_fn: the Newton term of a reduced withdrawal,
 * d(q sQSaturation)/dh = q sQSaturationDerivative, on the diagonal and the
 * rhs (brhs: wel_cf's rhs). */
void orc_wel_fn(int nbound, const int *nodelist, const double *q,
                const double *brhs, const int *ibound, int iflowred,
                double flowred, const int *icelltype, const double *top,
                const double *bot, const double *xnew, const int *ia,
                double *amat, double *rhs)
{
  for (int i = 0; i < nbound; i++) {
    int node = nodelist[i];
    if (ibound[node] <= 0) continue;
    if (iflowred == 0 || icelltype[node] == 0) continue;
    double qi = -brhs[i];
    if (qi < DZERO) {
      double tp = top[node], bt = bot[node];
      double thick = tp - bt;
      tp = bt + flowred * thick;
      double drterm = orc_sqsat_cubic_deriv(tp, bt, xnew[node], -2.0, 3.0);
      drterm = drterm * q[i];
      amat[ia[node]] = amat[ia[node]] + drterm;
      rhs[node] = rhs[node] + drterm * xnew[node];
    }
  }
}

/* gwf-rch.f90:303-353 rch_cf; nodelist already holds the receiving cell
 * (fixed_cell, or the top-most active cell for array input). */
void orc_rch_cf(int nbound, const int *nodelist, const double *recharge,
                const double *area, const int *ibound, double *hcof,
                double *brhs)
{
  for (int i = 0; i < nbound; i++) {
    int node = nodelist[i];
    if (node < 0) { hcof[i] = DZERO; brhs[i] = DZERO; continue; }
    hcof[i] = DZERO;
    brhs[i] = -recharge[i] * area[node];
    if (ibound[node] <= 0) brhs[i] = DZERO;
  }
}

/* gwf-drn.f90:501-531 get_drain_elevations: with a depth d /= 0 the
 * conductance is scaled between elev and elev + d */
static void drn_elevations(int idepth, const double *depth, const double *elev,
                           int i, double *drndepth, double *drntop, double *drnbot)
{
  double e = elev[i];
  *drndepth = idepth != 0 ? depth[i] : DZERO;
  if (*drndepth != DZERO) {
    double el = e + *drndepth;
    *drntop = el > e ? el : e;
    *drnbot = el < e ? el : e;
  } else {
    *drntop = e;
    *drnbot = e;
  }
}

/* gwf-drn.f90:340-373 drn_cf with get_drain_factor :534-574: a step at the
 * drain elevation, or with a depth sQSaturation (c1 = -1, c2 = 2;
 * icubic: DEV_CUBIC_SCALING, by default on under NEWTON, :130-135) or
 * sQuadraticSaturation with eps = 0 (linear) between the drain bottom and
 * top. */
void orc_drn_cf(int nbound, const int *nodelist, const double *elev,
                const double *cond, int idepth, const double *depth, int icubic,
                const int *ibound, const double *xnew, double *hcof, double *brhs)
{
  for (int i = 0; i < nbound; i++) {
    int node = nodelist[i];
    if (ibound[node] <= 0) { hcof[i] = DZERO; brhs[i] = DZERO; continue; }
    double cdrn = cond[i];
    double drndepth, drntop, drnbot, fact;
    drn_elevations(idepth, depth, elev, i, &drndepth, &drntop, &drnbot);
    double x = xnew[node];
    if (drndepth != DZERO) {
      if (icubic != 0) fact = orc_sqsat_cubic(drntop, drnbot, x, -DONE, 2.0);
      else fact = orc_sqsat(drntop, drnbot, x, DZERO);
    } else {
      fact = x <= drnbot ? DZERO : DONE;
    }
    brhs[i] = -fact * cdrn * drnbot;
    hcof[i] = -fact * cdrn;
  }
}

/* gwf-drn.f90:418-470 drn_fn: with a depth auxiliary, the derivative of the
 * cubic scaling (whichever scaling drn_cf used) times cdrn (drnbot - h) on
 * the diagonal and the rhs. */
void orc_drn_fn(int nbound, const int *nodelist, const double *elev,
                const double *cond, int idepth, const double *depth,
                const int *ibound, const double *xnew, const int *ia,
                double *amat, double *rhs)
{
  if (idepth == 0) return;
  for (int i = 0; i < nbound; i++) {
    int node = nodelist[i];
    if (ibound[node] <= 0) continue;
    double cdrn = cond[i];
    double x = xnew[node];
    double drndepth, drntop, drnbot;
    drn_elevations(idepth, depth, elev, i, &drndepth, &drntop, &drnbot);
    if (drndepth != DZERO) {
      double drterm = orc_sqsat_cubic_deriv(drntop, drnbot, x, -DONE, 2.0);
      drterm = drterm * cdrn * (drnbot - x);
      amat[ia[node]] = amat[ia[node]] + drterm;
      rhs[node] = rhs[node] + drterm * x;
    }
  }
}

/* ------------------------------------------------------------------ */
/* NumericalSolution pieces                                            */
/* ------------------------------------------------------------------ */

/* NumericalSolution.f90:2434-2478 sln_ls row fixes (one model, serial).
 * isymmetric = 1 for CG. */
void orc_sln_ls_fixes(int neq, const int *ia, const int *ja,
                      const int *active, int isymmetric, double *amat,
                      double *rhs, const double *x, double *xtemp)
{
  for (int ieq = 0; ieq < neq; ieq++) {
    xtemp[ieq] = x[ieq];
    if (active[ieq] > 0) {
      double diagval = -DONE;
      double adiag = fabs(amat[ia[ieq]]);
      if (adiag < DEM15) {
        amat[ia[ieq]] = diagval;
        rhs[ieq] = rhs[ieq] + diagval * x[ieq];
      }
    } else {
      amat[ia[ieq]] = DONE;
      for (int k = ia[ieq] + 1; k < ia[ieq + 1]; k++) amat[k] = DZERO;
      rhs[ieq] = x[ieq];
    }
  }
  if (isymmetric == 1) {
    for (int ieq = 0; ieq < neq; ieq++) {
      if (active[ieq] > 0) {
        for (int ipos = ia[ieq]; ipos < ia[ieq + 1]; ipos++) {
          int jcol = ja[ipos];
          if (jcol == ieq) continue;
          if (active[jcol] < 0) {
            rhs[ieq] = rhs[ieq] - (amat[ipos] * x[jcol]);
            amat[ipos] = DZERO;
          }
        }
      }
    }
  }
}

/* NumericalSolution.f90:3122-3153 sln_get_dxmax; returns the 0-based
 * location (or -1) and the signed change in *hncg. */
int orc_get_dxmax(int neq, const int *active, const double *x,
                  const double *xtemp, double *hncg)
{
  int nb = -1;
  double bigch = DZERO, abigch = DZERO;
  for (int n = 0; n < neq; n++) {
    if (active[n] < 1) continue;
    double hdif = x[n] - xtemp[n];
    double ahdif = fabs(hdif);
    if (ahdif > abigch) { bigch = hdif; abigch = ahdif; nb = n; }
  }
  *hncg = bigch;
  return nb;
}

/* NumericalSolution.f90:2912-2929 sln_calcdx */
void orc_calcdx(int neq, const int *active, const double *x,
                const double *xtemp, double *dx)
{
  for (int n = 0; n < neq; n++) dx[n] = active[n] < 1 ? DZERO : x[n] - xtemp[n];
}

/* NumericalSolution.f90:2875-2905 sln_maxval */
double orc_maxval(int nsize, const double *v)
{
  double vmax = v[0];
  for (int n = 1; n < nsize; n++) {
    double d = v[n];
    double denom = fabs(vmax);
    if (denom == DZERO) denom = DPREC;
    double dnorm = fabs(d) / denom;
    if (dnorm > DONE) vmax = d;
  }
  return vmax;
}

/* Under-relaxation state (NumericalSolution.f90:2989-3114). */
typedef struct {
  int nonmeth;
  double gamma, theta, akappa, amomentum;
  double bigch, bigchold, relaxold;
} orc_ur_state;

/* NumericalSolution.f90:2989-3114 sln_underrelax; wsave/hchold/deold are
 * neq-long state arrays for delta-bar-delta. */
void orc_underrelax(orc_ur_state *st, int kiter, double bigch, int neq,
                    const int *active, double *x, const double *xtemp,
                    double *dxold, double *wsave, double *hchold,
                    double *deold)
{
  if (st->nonmeth == 1) {
    for (int n = 0; n < neq; n++) {
      if (active[n] < 1) continue;
      double delx = x[n] - xtemp[n];
      dxold[n] = delx;
      x[n] = xtemp[n] + st->gamma * delx;
    }
  } else if (st->nonmeth == 2) {
    double relax;
    st->bigch = bigch;
    if (kiter == 1) {
      relax = DONE;
      st->relaxold = DONE;
      st->bigchold = bigch;
    } else {
      double es = st->bigch / (st->bigchold * st->relaxold);
      double aes = fabs(es);
      if (es < -DONE) relax = DHALF / aes;
      else relax = (DTHREE + es) / (DTHREE + aes);
    }
    st->relaxold = relax;
    st->bigchold = (DONE - st->gamma) * st->bigch + st->gamma * st->bigchold;
    if (relax < DONE) {
      for (int n = 0; n < neq; n++) {
        if (active[n] < 1) continue;
        double delx = x[n] - xtemp[n];
        dxold[n] = delx;
        x[n] = xtemp[n] + relax * delx;
      }
    }
  } else if (st->nonmeth == 3) {
    for (int n = 0; n < neq; n++) {
      if (active[n] < 1) continue;
      double delx = x[n] - xtemp[n];
      if (kiter == 1) { wsave[n] = DONE; hchold[n] = DEM20; deold[n] = DZERO; }
      double ww = wsave[n];
      if (deold[n] * delx < DZERO) ww = st->theta * wsave[n];
      else ww = wsave[n] + st->akappa;
      if (ww > DONE) ww = DONE;
      wsave[n] = ww;
      if (kiter == 1) hchold[n] = delx;
      else hchold[n] = (DONE - st->gamma) * delx + st->gamma * hchold[n];
      deold[n] = delx;
      dxold[n] = delx;
      double amom = DZERO;
      if (kiter > 4) amom = st->amomentum;
      delx = delx * ww + amom * hchold[n];
      x[n] = xtemp[n] + delx;
    }
  }
}

/* gwf-npf.f90:705-741 npf_nur.  Returns inewtonur (0/1); updates *dxmax /
 * *locmax (0-based) as the reference does. */
int orc_npf_nur(int nodes, const int *ibound, const int *icelltype,
                const int *ibotnode, const double *bot, double *x,
                const double *xtemp, double *dx, double *dxmax, int *locmax)
{
  int inewtonur = 0;
  for (int n = 0; n < nodes; n++) {
    if (ibound[n] < 1) continue;
    if (icelltype[n] > 0) {
      double botm = bot[ibotnode[n]];
      if (x[n] < botm) {
        inewtonur = 1;
        double xx = xtemp[n] * (DONE - DP9) + botm * DP9;
        double dxx = x[n] - xx;
        if (fabs(dxx) > fabs(*dxmax)) { *locmax = n; *dxmax = dxx; }
        x[n] = xx;
        dx[n] = DZERO;
      }
    }
  }
  return inewtonur;
}

/* NumericalSolution.f90:2966-2982 sln_calc_residual r = A x - b with
 * inactive rows zeroed, and SeqVector.f90:152-164 norm2 (plain sqrt of the
 * sum of squares); used by backtracking (sln_l2norm :2855). */
double orc_sln_l2norm(int neq, const int *ia, const int *ja,
                      const double *amat, const double *x, const double *rhs,
                      const int *active, double *r)
{
  orc_amux(neq, x, r, amat, ja, ia);
  for (int n = 0; n < neq; n++) r[n] = r[n] + (-1.0) * rhs[n];
  for (int n = 0; n < neq; n++) if (active[n] < 1) r[n] = 0.0;
  double s = 0.0;
  for (int n = 0; n < neq; n++) s = s + r[n] * r[n];
  return sqrt(s);
}

/* NumericalSolution.f90:2800-2826 get_backtracking_flag: 1 when
 * breduc * max|x - xtemp| over active cells >= dvclose. */
int orc_bt_flag(int neq, const int *active, const double *x,
                const double *xtemp, double breduc, double dvclose)
{
  double dmax = 0.0;
  for (int n = 0; n < neq; n++) {
    if (active[n] < 1) continue;
    double a = fabs(x[n] - xtemp[n]);
    if (a > dmax) dmax = a;
  }
  return breduc * dmax >= dvclose ? 1 : 0;
}

/* NumericalSolution.f90:2830-2843 apply_backtracking */
void orc_bt_apply(int neq, const int *active, double *x, const double *xtemp,
                  double breduc)
{
  for (int n = 0; n < neq; n++) {
    if (active[n] < 1) continue;
    double delx = breduc * (x[n] - xtemp[n]);
    x[n] = xtemp[n] + delx;
  }
}

/* gwf.f90:625-679 gwf_ptc on the residual r = A x - b of the formulated
 * system (sln_calc_ptc, NumericalSolution.f90:2936-2962).  The maximum cell
 * volume is area * (top - bot) (head at the cell top).  Returns iptc. */
int orc_gwf_ptc(int nodes, int iss, int inewton, const int *ibound,
                const double *area, const double *top, const double *bot,
                const double *resid, double delt, double *ptcf)
{
  int iptct = (iss > 0 && inewton > 0) ? inewton : 0;
  *ptcf = DZERO;
  if (iptct > 0) {
    for (int n = 0; n < nodes; n++) {
      if (ibound[n] < 1) continue;
      double v = area[n] * (top[n] - bot[n]);
      double r = fabs(resid[n]) / v;
      if (r > *ptcf) *ptcf = r;
    }
    if (*ptcf == DZERO) *ptcf = DONE / (delt * 10.0);
  }
  return iptct > 0 ? 1 : 0;
}

/* Pseudo-transient continuation state of NumericalSolutionType. */
typedef struct {
  int iallowptc;     /* 1 allowed, 0 NO_PTC ALL, -1 NO_PTC FIRST */
  int iptcopt;       /* 0: ptcdel = 1/ptcf ; 1: bnorm / l2norm */
  double ptcdel0, ptcexp;
  double ptcdel, l2norm0;
} orc_ptc_state;

/* NumericalSolution.f90:2480-2573: the PTC block of sln_ls, on the system
 * after the row fixes.  Returns 1 when PTC terms were added (iptc after the
 * l2norm test times iallowptc); r is neq work. */
int orc_sln_ptc(orc_ptc_state *st, int neq, const int *ia, const int *ja,
                double *amat, double *rhs, const double *x,
                const int *active, int kiter, int kstp, int kper, int iptc,
                double ptcf, double *r)
{
  int iallowptc;
  if (st->iallowptc < 0) iallowptc = kper > 1 ? 1 : 0;
  else iallowptc = st->iallowptc;
  int iptct = iptc * iallowptc;
  double l2norm = 0.0;
  if (iptct != 0) {
    l2norm = orc_sln_l2norm(neq, ia, ja, amat, x, rhs, active, r);
    if (kiter == 1) {
      if (kper > 1 || kstp > 1) {
        if (l2norm <= st->l2norm0) iptc = 0;
      }
    } else {
      if (orc_is_close(l2norm, st->l2norm0)) iptc = 0;
    }
  }
  iptct = iptc * iallowptc;
  if (iptct == 0) return 0;
  {
    if (kiter == 1) {
      if (st->ptcdel0 > DZERO) {
        st->ptcdel = st->ptcdel0;
      } else if (st->iptcopt == 0) {
        st->ptcdel = DONE / ptcf;
      } else {
        double bnorm = DZERO;
        for (int n = 0; n < neq; n++)
          if (active[n] > 0) bnorm = bnorm + rhs[n] * rhs[n];
        bnorm = sqrt(bnorm);
        st->ptcdel = bnorm / l2norm;
      }
    } else {
      if (l2norm > DZERO) st->ptcdel = st->ptcdel * pow(st->l2norm0 / l2norm, st->ptcexp);
      else st->ptcdel = DZERO;
    }
    double ptcval = st->ptcdel > DZERO ? DONE / st->ptcdel : DONE;
    for (int n = 0; n < neq; n++) {
      if (active[n] > 0) {
        amat[ia[n]] = amat[ia[n]] + (-ptcval);
        rhs[n] = rhs[n] - ptcval * x[n];
      }
    }
    st->l2norm0 = l2norm;
  }
  return 1;
}
