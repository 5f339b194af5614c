/*
 * mf6hip.h -- C ABI of libmf6hip.so, the MI355X (gfx950) implementation of
 * the MODFLOW 6 IMS linear solve and GWF matrix fill.
 *
 * Conventions (copied from the reference's own C-ABI, srcbmi/mf6bmi.f90 and
 * srcbmi/mf6xmi.F90): every entry point returns int 0 on success and a
 * non-zero status on failure; the message of the last failure is fetched
 * with mf6hip_last_error (cf. get_last_bmi_error, srcbmi/mf6bmiError.f90:57).
 * A linear solve that does not converge is NOT a failure: it reports
 * *icnvg = 0 and the outer loop decides (NumericalSolution.f90:1626-1629).
 *
 * Arrays cross the boundary in host memory, in the reference's own layout:
 * CSR with the diagonal first in each row followed by ascending columns
 * (src/Utilities/Matrix/SparseMatrix.f90:16, src/Utilities/Sparse.f90:217),
 * 1-based ia/ja exactly as the Fortran memory manager holds them, FP64
 * values, int32 indices.  Cell numbering is the DIS node number
 * icol + ncol*(irow-1) + (ilay-1)*nrow*ncol (src/Utilities/GeomUtil.f90:90).
 *
 * One handle = one IMS solution with one structured (DIS) GWF model; calls
 * on a handle are strictly sequential, as in MF6 (single-threaded host).
 * The library holds the system resident in HBM between calls.
 */
#ifndef MF6HIP_H
#define MF6HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes of every entry point (0 = success, as the reference's BMI
 * functions return). */
#define MF6HIP_OK 0
#define MF6HIP_EINVAL 1        /* invalid argument or input the call cannot take */
#define MF6HIP_EHIP 2          /* a HIP runtime call failed */
#define MF6HIP_EUNSUPPORTED 3  /* an option or input the device path does not model */

/* --- IMS settings ------------------------------------------------------ */

/* precision of the device solve (opt-in; stock .ims files get MF6HIP_FP64) */
#define MF6HIP_FP64 0        /* everything FP64 */
#define MF6HIP_MIXED 1       /* FP32 ILU0 + FP32 SpMV: CG with the FP32
                                residual replaced by the FP64 b - A x every
                                inner_fp32_max iterations (symmetric systems),
                                or FP64 iterative refinement around FP32
                                inner solves; heads, true residual and the
                                convergence test stay FP64 */
#define MF6HIP_MIXED_FP16 2  /* as MIXED, ILU0 pivots stored in FP16 */

/* ordering of the device ILU0 (the reference factors in natural order) */
#define MF6HIP_ORDER_RB 0    /* column red-black: vertical cell columns
                                coloured (irow + icol) mod 2 */

/* Mirrors ImsLinearSettingsType (src/Solution/LinearMethods/
 * ImsLinearSettings.f90:13-32) field for field, plus the device opt-ins.
 * Declared interoperable (bind(C)) on the Fortran side, see
 * modflow6_amd/fortran/mf6hip_module.f90. */
typedef struct mf6hip_ims_settings {
  double dvclose;   /* INNER_DVCLOSE */
  double rclose;    /* INNER_RCLOSE */
  int icnvgopt;     /* 0 default, 1 STRICT, 2 L2NORM_RCLOSE, 3 RELATIVE, 4 */
  int iter1;        /* INNER_MAXIMUM */
  int ilinmeth;     /* 1 CG, 2 BICGSTAB */
  int iscl;         /* SCALING_METHOD: 0 none, 1 symmetric, 2 L2 norm (FP64, ILU0/MILU0 only) */
  int iord;         /* REORDERING_METHOD: must be 0 on the device path */
  int north;        /* NUMBER_ORTHOGONALIZATIONS */
  double relax;     /* RELAXATION_FACTOR: 0 -> ILU0, > 0 -> MILU0 */
  int level;        /* PRECONDITIONER_LEVELS: must be 0 (no ILUT) */
  double droptol;   /* PRECONDITIONER_DROP_TOLERANCE: must be 0 */
  int precision;    /* MF6HIP_FP64 / MF6HIP_MIXED / MF6HIP_MIXED_FP16 */
  int ordering;     /* MF6HIP_ORDER_RB */
  int nlay, nrow, ncol; /* DIS shape of the (single) model */
  int inner_fp32_max;   /* MIXED: FP32 CG iterations between replacements of the
                           FP32 residual by the FP64 b - A x (0 = 64) */
  double inner_fp32_rtol; /* MIXED, unsymmetric or unfused path: relative FP32
                             residual reduction per refinement step (0 = 1e-4) */
  int kernel_variant;   /* 0: fused thread-per-column CG when the system is
                           symmetric (default); 1: one kernel per reference
                           vector pass */
  int preconditioner;   /* 0: ILU0 / MILU0 (IMS); 1: ILU0 followed by the
                           learned correction of an attached network, in
                           flexible CG (CG only); 2: two-level, ILU0 / MILU0
                           plus an additive coarse correction on column
                           aggregates (symmetric systems, CG, fused kernels,
                           any precision); 3: the same coarse space in the
                           deflated form, z = M^-1 (r - A P yc) + P yc with
                           yc = Ac^-1 P^T r (A-DEF1; same conditions as 2).
                           This operator is not symmetric: r.z can change
                           sign between iterations, and the CG then
                           restarts its direction (p = z) instead of
                           forming beta */
  int coarse_size;      /* preconditioner 2, 3: most coarse unknowns (aggregates
                           of ha x ha columns, all layers); 0 = sqrt(2 N)
                           within [16, 4096]; at most 6144 */
} mf6hip_ims_settings;

/* Per-inner-iteration record, the values ConvergenceSummaryType keeps for
 * CSV_INNER_OUTPUT (src/Solution/ConvergenceSummary.f90:13-34,
 * ImsLinearBase.f90:186-197).  Locations are 1-based node numbers. */
typedef struct mf6hip_cnvg_out {
  int capacity;     /* in: length of the arrays below (0 = none wanted) */
  int count;        /* out: records written */
  double *dvmax;    /* signed largest x change of the iteration */
  double *rmax;     /* signed largest residual of the iteration */
  double *alpha;    /* CG / BiCGSTAB alpha */
  double *omega;    /* BiCGSTAB omega (0 for CG) */
  int *locdv;
  int *locr;
} mf6hip_cnvg_out;

/* --- drop-in linear solve (replaces imslinear_ap) ---------------------- */

/* Create a solver for a neq x neq system with CSR structure ia/ja (1-based,
 * diagonal first).  Replaces imslinear_ar (src/Solution/LinearMethods/
 * ImsLinear.f90:111-339) / LinearSolverBaseType%initialize
 * (src/Solution/LinearSolverBase.f90:32).  The structure must be the
 * 7-point DIS stencil of nlay x nrow x ncol with every cell present. */
int mf6hip_create(int neq, int nja, const int *ia, const int *ja,
                  const mf6hip_ims_settings *settings, void **handle);

/* One linear solve: factor, initial residual, CG or BiCGSTAB until the
 * IMS convergence test (ImsLinearBase.f90:1101) passes or INNER_MAXIMUM is
 * reached.  amat/rhs are the system after the sln_ls row fixes
 * (NumericalSolution.f90:2434-2478); x holds the start heads on entry and the
 * solution on return.  Replaces imslinear_ap (ImsLinear.f90:617-750) as
 * called from sln_ls (NumericalSolution.f90:2604-2608), and
 * LinearSolverBaseType%solve (LinearSolverBase.f90:44). */
int mf6hip_solve(void *handle, int kstp, int kiter, const double *amat,
                 const double *rhs, double *x, int *icnvg, int *in_iter,
                 mf6hip_cnvg_out *out);

/* Release the handle (LinearSolverBaseType%destroy, LinearSolverBase.f90:53). */
int mf6hip_destroy(void *handle);

/* Copy the last error message (NUL-terminated, truncated to buflen). */
int mf6hip_last_error(char *buf, int buflen);

/* --- device-resident GWF formulate + solve ----------------------------- */

/* Model arrays of one DIS GWF model, natural node order, host memory; the
 * library copies them to HBM.  Names follow the reference's NPF / STO /
 * DIS members (gwf-npf.f90, gwf-sto.f90, Dis.f90). */
typedef struct mf6hip_gwf_model {
  int nlay, nrow, ncol;
  const double *delr;       /* ncol */
  const double *delc;       /* nrow */
  const double *top;        /* nodes: cell top (layer 1 TOP, then BOTM above) */
  const double *bot;        /* nodes: cell bottom */
  const double *k11;        /* nodes: NPF K */
  const double *k33;        /* nodes: NPF K33 (K when not given) */
  const int *icelltype;     /* nodes: NPF ICELLTYPE as read (see ithickstrt) */
  const int *ibound;        /* nodes: IDOMAIN/ibound before CHD (1 or 0) */
  int inewton;              /* GWF NEWTON */
  int inewtonur;            /* GWF NEWTON UNDER_RELAXATION */
  int ivarcv;               /* NPF VARIABLECV */
  int idewatcv;             /* NPF VARIABLECV DEWATERED */
  int icellavg;             /* NPF ALTERNATIVE_CELL_AVERAGING: 0 harmonic,
                               1 LOGARITHMIC, 2 AMT-LMK, 3 AMT-HMK
                               (gwf-npf.f90 condmean cases) */
  int has_sto;              /* STO package present */
  const int *iconvert;      /* nodes (STO ICONVERT), may be NULL without STO */
  const double *ss;         /* nodes (STO SS) */
  const double *sy;         /* nodes (STO SY) */
  const double *k22;        /* nodes: NPF K22; NULL: K11 */
  const double *angle1;     /* nodes: NPF ANGLE1, degrees as read; NULL: none */
  const double *angle2;     /* nodes: NPF ANGLE2 (needs ANGLE1 and ANGLE3) */
  const double *angle3;     /* nodes: NPF ANGLE3 (needs ANGLE1 and ANGLE2) */
  int istor_coef;           /* STO STORAGECOEFFICIENT (gwf-sto.f90:32) */
  int iconf_ss;             /* STO SS_CONFINED_ONLY (:33) */
  int iorig_ss;             /* STO original specific storage formulation (:34) */
  int ithickstrt;           /* NPF THICKSTRT: a negative ICELLTYPE marks a confined
                               cell with the saturated thickness of its starting
                               head (gwf-npf.f90:1836-1886); without it a
                               negative ICELLTYPE means convertible */
  const double *strt;       /* nodes: IC STRT (needed with THICKSTRT) */
  /* The faces' conductivities follow hy_eff (gwf-npf.f90:2280-2355) with the
   * DIS connection normals: column- and row-direction faces K11 unless K22
   * is given, then the rotated tensor's value along the face normal
   * (HGeoUtil.f90:29-108 hyeff, harmonic); vertical faces K33 unless ANGLE2
   * is given.  Computed once on the host at mf6hip_gwf_create. */
} mf6hip_gwf_model;

/* package types, in the bndlist order MF6 fixes by type (gwf.f90:119-124,
 * src/Utilities/Idm/ModelPackageInputs.f90:246-310): WEL, DRN, RCH, CHD */
#define MF6HIP_PKG_WEL 1
#define MF6HIP_PKG_DRN 2
#define MF6HIP_PKG_RCH 3
#define MF6HIP_PKG_CHD 4

/* Upload the model (replaces npf_ar/sto_ar + calc_condsat,
 * gwf-npf.f90:1748-2037: the saturated conductances are computed on the
 * device). */
int mf6hip_gwf_create(void *handle, const mf6hip_gwf_model *model);

/* Set one boundary package's period data (bnd_rp).  nodelist is 1-based.
 * WEL: v1 = q.  DRN: v1 = elevation, v2 = conductance.  RCH: v1 = recharge
 * rate (nodelist = receiving cells).  CHD: v1 = head.  Packages are applied
 * in MF6HIP_PKG order regardless of call order; a second call with the same
 * type replaces that package's data and options (this call: none). */
int mf6hip_gwf_set_package(void *handle, int pkgtype, int nbound,
                           const int *nodelist, const double *v1,
                           const double *v2);

/* Package options of the reference that change a package's formulation
 * (mf6hip_gwf_set_package_ex).
 *  - WEL AUTO_FLOW_REDUCE (iflowred = 1, flowred the value as read: <= 0
 *    becomes 0.1, above 1 becomes 1, as wel_options stores it,
 *    src/Model/GroundWaterFlow/gwf-wel.f90:205-217): withdrawals in
 *    convertible cells are scaled by sQSaturation over the bottom flowred
 *    of the cell, with the Newton term under NEWTON (wel_cf / wel_fn,
 *    :296-332, :378-424).
 *  - DRN AUXDEPTHNAME (idrndepth = 1, drndepth the named auxiliary column,
 *    nbound values): conductance scaled between elev and elev + depth
 *    (get_drain_elevations / get_drain_factor, gwf-drn.f90:501-574),
 *    linearly or, under NEWTON or with DEV_CUBIC_SCALING (icubic_scaling =
 *    1), by sQSaturation; drn_fn adds the cubic Newton term (:418-470).
 *  - AUXMULTNAME is rejected (MF6HIP_EUNSUPPORTED): the host applies the
 *    multiplier to v1 / v2 itself. */
typedef struct mf6hip_package_options {
  double flowred;          /* WEL AUTO_FLOW_REDUCE value as read */
  int iflowred;            /* WEL: AUTO_FLOW_REDUCE given */
  int idrndepth;           /* DRN: AUXDEPTHNAME given (drndepth holds it) */
  int icubic_scaling;      /* DRN: DEV_CUBIC_SCALING given */
  int has_auxmult;         /* AUXMULTNAME given (rejected) */
  const double *drndepth;  /* DRN: nbound drain depths, host memory */
} mf6hip_package_options;

/* mf6hip_gwf_set_package with the package's options (NULL: none).  The
 * options are checked before the package data are replaced: AUXMULTNAME
 * gives MF6HIP_EUNSUPPORTED; an option of another package type, a missing
 * drndepth (nbound > 0) or a NaN flowred MF6HIP_EINVAL. */
int mf6hip_gwf_set_package_ex(void *handle, int pkgtype, int nbound,
                              const int *nodelist, const double *v1,
                              const double *v2, const mf6hip_package_options *opt);

/* Horizontal flow barriers (the HFB package): a new set replaces the last
 * one, as hfb_rp does when a PERIOD block is read (gwf-hfb.f90:149-201:
 * condsat_reset, read_data, condsat_modify :770-832).  noden / nodem are
 * 1-based node numbers of horizontally adjacent cells, hydchr the barriers'
 * hydraulic characteristics (> 0: conductance per unit area, combined in
 * series with the face's conductance; <= 0: -hydchr multiplies it).  Under
 * NEWTON, or when both cells are confined, the barrier is folded into the
 * face's saturated conductance, as the reference does; every later
 * formulate and budget sees it.  Without NEWTON a barrier next to a
 * convertible cell is applied at every formulate as hfb_fc does (:211-348,
 * the face's conductance combined with the barrier over the cells' wetted
 * thickness) and in the budget as hfb_cq (:355-450).  nhfb = 0 removes the
 * barriers.  MF6HIP_EINVAL for cells that are not horizontally connected
 * (check_data, :711-766). */
int mf6hip_gwf_set_hfb(void *handle, int nhfb, const int *noden, const int *nodem,
                       const double *hydchr);

/* Ghost-node correction of the model (the GNC package of one model,
 * src/Exchange/GhostNode.f90): a new set replaces the last one, as gnc_df /
 * read_data / gnc_mc (:84-246) set it up.  GNC g corrects the flow between
 * the connected cells noden[g] and nodem[g] (1-based) with the heads of
 * numjs contributing cells nodesj[g * numjs + i] (0: none) weighted by
 * alphasj[g * numjs + i].  Every formulate adds gnc_fc's explicit rhs terms
 * (:280-324, conductances read from the matrix after npf_fc as gnc_fmsav
 * :251-273 does) and, under NEWTON, gnc_fn (:340-443); every budget adds
 * gnc_cq's correction flows (:478-543).  nexg = 0 removes them.
 * MF6HIP_EUNSUPPORTED for implicit != 0: the implicit form puts the
 * contributing cells into the matrix rows of n and m (gnc_ac :134-161),
 * outside a DIS grid's 7-point rows.  MF6HIP_EINVAL for cells that are not
 * connected or out of range. */
int mf6hip_gwf_set_gnc(void *handle, int nexg, int numjs, const int *noden,
                       const int *nodem, const int *nodesj, const double *alphasj,
                       int implicit);

/* Heads in / out (natural order, host memory). */
int mf6hip_gwf_set_x(void *handle, const double *x);
int mf6hip_gwf_get_x(void *handle, double *x);
/* The same exchange with a device buffer of N doubles in DIS node order (a
 * coupler's GPU array, e.g. a torch tensor): no host copy.  Replaces reading
 * / writing SLN_1/X through get_value_ptr (srcbmi/mf6bmi.f90:707) between
 * prepare_solve and finalize_solve (srcbmi/mf6xmi.F90:221-340) when the
 * coupler's data already lives on the GPU.  Both return after the copy has
 * completed on the library's stream. */
int mf6hip_gwf_set_x_device(void *handle, const double *x_dev);
int mf6hip_gwf_get_x_device(void *handle, double *x_dev);

/* Start a time step (gwf_ad, gwf.f90:396-421 with chd_ad): xold = x, CHD
 * cells take their head; iss = 1 for a steady-state period; delt the step
 * length. */
int mf6hip_gwf_advance(void *handle, int iss, double delt);

/* sln_buildsystem(kiter, inewton) (NumericalSolution.f90:1941-1991) for the
 * GWF model: amat = 0, rhs = 0, then npf_cf, bnd_cf, npf_fc, sto_fc, bnd_fc
 * and, when inewton = 1, npf_fn and sto_fn, in gwf_fc order (gwf.f90:446-555,
 * inwtflag), bit-identical to the reference.  Backtracking refills with
 * inewton = 0 (:2698).  If amat_out / rhs_out are not NULL the assembled CSR
 * values (reference layout, nja) and rhs are copied back for inspection. */
int mf6hip_gwf_formulate(void *handle, int kiter, int inewton, double *amat_out,
                         double *rhs_out);

/* Budget of the current heads: gwf_cq + gwf_bd (gwf.f90:741-800), i.e.
 * npf_cq / qcalc (gwf-npf.f90:745-862), sto_cq (gwf-sto.f90:447-564),
 * bnd_cq (BoundaryPackage.f90:545-619), csr_diagsum (Sparse.f90:262-281)
 * and calc_chd_rate (gwf-chd.f90:264-320).  flowja (may be NULL): nja values
 * in the CSR order of mf6hip_create, positive into the cell, the diagonal the
 * cell's flow residual.  terms[12]: (rate in, rate out) of STO-SS, STO-SY,
 * WEL, DRN, RCH, CHD (zeros for an absent term); pct_discrepancy as
 * Budget.f90:259-267.  The conductances use the sat of the last
 * mf6hip_gwf_formulate, as npf_cq uses the sat of the last npf_cf. */
int mf6hip_gwf_budget(void *handle, double *flowja, double *terms, double *pct_discrepancy);

/* Flow terms of mf6hip_gwf_budget, in the order of its terms[] pairs. */
#define MF6HIP_TERM_STO_SS 0
#define MF6HIP_TERM_STO_SY 1
#define MF6HIP_TERM_WEL 2
#define MF6HIP_TERM_DRN 3
#define MF6HIP_TERM_RCH 4
#define MF6HIP_TERM_CHD 5

/* One flow term of the last mf6hip_gwf_budget, as the reference saves it to
 * the cell-by-cell budget file (gwf_ot_flow, gwf.f90:911-958): STO-SS /
 * STO-SY are the storage rates of every cell in natural node order
 * (sto%strgss / strgsy, written by sto_save_model_flows, gwf-sto.f90:602-640);
 * WEL / DRN / RCH / CHD are the package's simvals in bound order (bnd_cq,
 * BoundaryPackage.f90:545-619; CHD: calc_chd_rate, gwf-chd.f90:264-320),
 * written by save_print_model_flows (BoundaryPackage.f90:1753-1902).
 * *count receives the number of values (nodes, or the package's nbound at
 * that budget); out may be NULL to query the count only.  MF6HIP_EINVAL if
 * no budget has been computed on the handle. */
int mf6hip_gwf_flow_term(void *handle, int term, double *out, int *count);

/* r = A x - b of the current device system with inactive rows zeroed
 * (sln_calc_residual, NumericalSolution.f90:2966-2982): *l2norm = ||r||_2
 * (sln_l2norm, :2855-2871) and *rvmax = max |r(n)| / (area (top - bot))
 * over cells with ibound >= 1, the reciprocal pseudo-time step of gwf_ptc
 * (gwf.f90:625-679) before its zero-residual protection. */
int mf6hip_gwf_residual_norms(void *handle, double *l2norm, double *rvmax);

/* sln_backtracking_xupdate (NumericalSolution.f90:2784-2843): *btflag = 1
 * when breduc * max|x - xtemp| (active cells) >= dvclose, and then
 * x = xtemp + breduc (x - xtemp) on active cells. */
int mf6hip_gwf_backtrack(void *handle, double breduc, double dvclose,
                         int *btflag);

/* Pseudo-transient continuation of sln_ls (NumericalSolution.f90:2480-2573).
 * The caller fills iptc and ptcf from sln_calc_ptc / gwf_ptc for this outer
 * iteration and keeps ptcdel / l2norm0 between outer iterations. */
typedef struct mf6hip_ptc {
  int iptc;          /* in: PTC requested (gwf_ptc); out: 1 if PTC terms were
                        added (after NO_PTC and the l2norm test) */
  int iallowptc;     /* 1 allowed, 0 NO_PTC ALL, -1 NO_PTC FIRST */
  int iptcopt;       /* 0: ptcdel = 1 / ptcf ; 1: ||b|| / ||r|| (DEV_PTC_OPTION) */
  int kstp, kper;
  double ptcf;       /* in: reciprocal pseudo-time step from gwf_ptc */
  double ptcdel0;    /* DEV_PTC_DEL0 (0 = not set) */
  double ptcexp;     /* DEV_PTC_EXPONENT (1) */
  double ptcdel;     /* in/out state */
  double l2norm0;    /* in/out state */
} mf6hip_ptc;

/* sln_ls (NumericalSolution.f90:2404-2615) on the device-resident system:
 * xtemp = x, row fixes, pseudo-transient continuation when ptc is not NULL,
 * the linear solve, then sln_get_dxmax (:3122).  hncg receives the signed
 * largest head change, lrch its 1-based node. */
int mf6hip_gwf_linear_solve(void *handle, int kstp, int kiter, mf6hip_ptc *ptc,
                            int *icnvg, int *in_iter, double *hncg, int *lrch,
                            mf6hip_cnvg_out *out);

/* IMS NONLINEAR under-relaxation (sln_underrelax, NumericalSolution.f90
 * :2989-3114): nonmeth 1 SIMPLE, 2 COOLEY, 3 DBD. */
typedef struct mf6hip_underrelax {
  int nonmeth;
  double gamma, theta, akappa, amomentum;
  double bigch;      /* in: hncg of this outer iteration */
  double bigchold;   /* in/out: Cooley state */
  double relaxold;   /* in/out: Cooley state */
} mf6hip_underrelax;

/* The tail of NumericalSolution%solve (:1722-1760) for one model when the
 * outer iteration has not converged: sln_underrelax when ur is not NULL and
 * ur->nonmeth > 0, else sln_calcdx; then npf_nur (gwf-npf.f90:705-741) if
 * Newton under-relaxation is on.  On return *inewtonur tells whether NUR
 * moved a head, *dxold_max is sln_maxval of dxold, and, when NUR applied,
 * *hncg / *lrch are recomputed (sln_get_dxmax). */
int mf6hip_gwf_outer_update(void *handle, int kiter, mf6hip_underrelax *ur,
                            int *inewtonur, double *dxold_max, double *hncg,
                            int *lrch);

/* --- XMI-style outer loop (srcbmi/mf6xmi.F90:220-340) -------------------
 *
 * prepare_solve / solve / finalize_solve of the coupling interface for the
 * handle's one GWF model: each solve call is one NumericalSolution outer
 * iteration (NumericalSolution.f90:1482-1837) run on the device entry points
 * above -- backtracking, formulate, pseudo-transient continuation (steady
 * Newton models), sln_ls, the OUTER_DVCLOSE test, under-relaxation and
 * Newton under-relaxation -- so a coupler loops
 *     prepare_solve; do { solve(&c) } while (!c && k < OUTER_MAXIMUM);
 *     finalize_solve
 * and can exchange heads with mf6hip_gwf_get_x_device / set_x_device between
 * the calls, as it reads and writes SLN_1/X through get_value_ptr
 * (srcbmi/mf6bmi.f90:707). */

/* IMS NONLINEAR block (sln_ar, NumericalSolution.f90:790-900 with the
 * complexity presets of sln_setouter :2623-2671 already applied). */
typedef struct mf6hip_nonlinear {
  double outer_dvclose;     /* OUTER_DVCLOSE */
  int outer_maximum;        /* OUTER_MAXIMUM */
  int nonmeth;              /* UNDER_RELAXATION: 0 NONE, 1 SIMPLE, 2 COOLEY, 3 DBD */
  double gamma, theta, akappa, amomentum;   /* UNDER_RELAXATION_GAMMA / THETA /
                                               KAPPA / MOMENTUM */
  int backtracking_number;  /* BACKTRACKING_NUMBER (0: none) */
  double backtracking_tolerance, backtracking_reduction_factor,
      backtracking_residual_limit;
  int iallowptc;            /* 1 PTC allowed, 0 NO_PTC ALL, -1 NO_PTC FIRST */
  int iptcopt;              /* DEV_PTC_OPTION */
  double ptcexp, ptcdel0;   /* DEV_PTC_EXPONENT, DEV_PTC_DEL0 */
} mf6hip_nonlinear;

/* What one solve call did (the ConvergenceSummary entries of an outer
 * iteration plus the NONLINEAR actions). */
typedef struct mf6hip_xmi_iteration {
  int kiter;          /* outer iteration (1-based within the time step) */
  int inner;          /* inner (linear) iterations */
  int icnvg_inner;    /* linear solve converged */
  double hncg;        /* signed largest head change */
  int lrch;           /* its 1-based node */
  int converged;      /* outer test passed */
  int nur;            /* Newton under-relaxation moved a head */
  int nbt;            /* backtracking steps */
  int iptc;           /* pseudo-transient continuation terms added */
} mf6hip_xmi_iteration;

/* Set the NONLINEAR block (once, before the first prepare_solve); resets
 * the Cooley / PTC / backtracking state. */
int mf6hip_xmi_set_nonlinear(void *handle, const mf6hip_nonlinear *nl);
/* xmi_prepare_solve: start time step kstp of period kper (iss = 1 steady,
 * delt its length): gwf_ad + chd_ad, outer iteration counter reset. */
int mf6hip_xmi_prepare_solve(void *handle, int kper, int kstp, int iss, double delt);
/* xmi_solve: one outer iteration; *has_converged = 1 when it met the outer
 * convergence test. */
int mf6hip_xmi_solve(void *handle, int *has_converged);
/* The record of the last solve call. */
int mf6hip_xmi_last_iteration(void *handle, mf6hip_xmi_iteration *rec);
/* xmi_finalize_solve: whether the time step converged and how many outer
 * iterations it took (either pointer may be NULL).  Non-convergence is not
 * an error (BMI_SUCCESS in the reference). */
int mf6hip_xmi_finalize_solve(void *handle, int *converged, int *outer_iterations);

/* Timing of the last mf6hip_gwf_linear_solve on the device stream (HIP
 * events): total milliseconds and the average duration of its dominant
 * kernel (the SpMV) with its launch count. */
int mf6hip_last_timing(void *handle, double *solve_ms, double *spmv_avg_ms,
                       int *spmv_launches);

/* --- learned preconditioner (no counterpart in the reference) ----------
 *
 * A 3D U-Net of nlev levels (2..8; channels c0 = 16, c1.. in {16, 32, 48,
 * 64}) with 1x2x2 mean pooling (level l+1 has ceil(nrow_l / 2) rows and
 * ceil(ncol_l / 2) columns) and nearest upsampling, BF16 weights and
 * activations and FP32 accumulation on MFMA, trained on the device on
 * (residual, error) pairs.  Inputs are nin planes of nlay*nrow*ncol FP32
 * values in DIS node order; parameters are one FP32 vector: per 3x3x3
 * convolution (encoder e0a e0b e1a e1b ... two per level, then decoder
 * d_{nlev-2} ... d0) W[27][cin][cout] then b[cout], then the 1x1x1 output
 * layer w[c0], b. */
int mf6hip_unet_create(int nlay, int nrow, int ncol, int nin, int nlev, const int *channels,
                       void **net);
int mf6hip_unet_destroy(void *net);
int mf6hip_unet_num_params(void *net, long long *n);
int mf6hip_unet_set_params(void *net, const float *params);
int mf6hip_unet_get_params(void *net, float *params);
/* y = network(x) */
int mf6hip_unet_forward(void *net, const float *x, float *y);
/* relative squared error sum (y - t)^2 / sum t^2 of one sample, and its
 * gradient with respect to the parameters */
int mf6hip_unet_grad(void *net, const float *x, const float *t, float *loss,
                     float *grad);
/* store n training samples on the device: [n][nin + 1][nodes] FP32 planes
 * (inputs, then the target); kept in BF16 */
int mf6hip_unet_set_data(void *net, int n, const float *samples);
/* Adam: nsteps steps over batches of stored samples in the given order
 * (order has nsteps * batch entries); loss[nsteps] gets each step's mean */
int mf6hip_unet_train(void *net, int nsteps, int batch, const int *order,
                      float lr, float *loss);
/* Attach a trained network (3 inputs) to a solver created with
 * preconditioner = 1; features = 2 planes (standardised ln K and
 * ln(K33 / K)) in DIS node order; fcg_m = directions the flexible CG keeps
 * (1..64); the network corrects every nn_every-th iteration's ILU0 step.
 * net = NULL detaches (the flexible CG then runs on ILU0). */
int mf6hip_attach_unet(void *handle, void *net, const float *features, int fcg_m,
                       int nn_every);
/* Iterations of the last solve that used the learned correction (the rest
 * ran on the ILU0 fallback). */
int mf6hip_last_learned_iterations(void *handle, int *n);

/* --- ensembles of independent systems (no counterpart in the reference:
 * MF6 solves one model per solution; an ensemble is many IMS solves of
 * different DIS models at once) ------------------------------------------
 *
 * Mixed-precision CG (FP32 ILU0 / SpMV, FP64 heads and residual
 * replacement) for every system in each kernel launch; the iteration
 * sequence is captured as a HIP graph and replayed until every system has
 * stopped.  Systems must use CG, ILU0 / MILU0 and ICNVGOPT 0. */
int mf6hip_batch_create(int capacity, void **batch);
/* register a system: settings carry its DIS shape and IMS values */
int mf6hip_batch_add(void *batch, const mf6hip_ims_settings *settings);
/* allocate device storage for every registered system */
int mf6hip_batch_finalize(void *batch);
/* system data in the reference layout (1-based DIS CSR, amat after the
 * sln_ls row fixes, rhs, start heads).  ia / ja must be the DIS 7-point
 * structure of the registered shape (diagonal first) and amat symmetric:
 * checked on the host before any staging, MF6HIP_EINVAL /
 * MF6HIP_EUNSUPPORTED otherwise.  Needs mf6hip_batch_finalize first. */
int mf6hip_batch_set_system(void *batch, int isys, const int *ia, const int *ja,
                            int nja, const double *amat, const double *rhs,
                            const double *x0);
/* pick columns per thread and the persistent grid by timing short replays;
 * report receives a text summary */
int mf6hip_batch_autotune(void *batch, char *report, int report_len);
/* solve every system; icnvg / iters per system, total device time */
int mf6hip_batch_solve(void *batch, int *icnvg, int *iters, double *solve_ms);
int mf6hip_batch_get_x(void *batch, int isys, double *x);
int mf6hip_batch_destroy(void *batch);

/* --- diagnostics (no counterpart in the reference; used by the parity
 * tests and the benchmark to reach single operators) --------------------- */

/* enable (1) / disable (0) per-solve HIP-event timing; sample = number of
 * SpMV launches timed individually per solve (at most 256) */
int mf6hip_set_timing(void *handle, int enable, int sample);

/* One device operator on host vectors in natural node order:
 * op 0: y = A x (amux order, src/Utilities/Libraries/sparsekit/
 *       sparsekit.f90:1-59);
 * op 1: y = M^-1 x, the ILU0/MILU0 apply (ims_base_ilu0a, src/Solution/
 *       LinearMethods/ImsLinearBase.f90:1049-1092) of the factor on the
 *       device, in the device ordering;
 * op 2: y = the inverse pivots of that factor (x unused);
 * op 3: y = the additive two-level preconditioner applied to x
 *       (preconditioner 2 or 3): ILU0 apply plus P (P^T A P)^-1 P^T x;
 * op 4: y = A x through the fused CG's SpMV (FP64 planes, amux order);
 * op 5: y = A x through the fused CG's FP32 SpMV, A and x rounded to FP32
 *       (mixed-precision handles), y the FP32 result widened;
 * op 6: y = the deflated two-level preconditioner applied to x (FP64,
 *       preconditioner 3): yc = (P^T A P)^-1 P^T x,
 *       y = M^-1 (x - A P yc) + P yc;
 * op 7: y = M^-1 x through the fused CG's ILU0 sweeps (FP64, preconditioner
 *       0): equal to op 1 for symmetric A;
 * op 8: the same sweeps in the mixed handle's precision (FP32 planes and
 *       pivots, or FP16 planes scaled by a power of two with FP32 pivots),
 *       x rounded to FP32.
 * op 9: (preconditioner 3) the restriction P^T x over the coupled cells two
 *       ways: y[0, nc) as the deflated iteration forms it from column sums
 *       (P^T (r - alpha q) with r = q = x and alpha = 0.75, i.e. 0.25 P^T x),
 *       y[nc, 2 nc) directly; the rest of y is zero (nc = the number of
 *       aggregates; needs 2 nc <= N).
 * A and the factor are those of the last mf6hip_solve / mf6hip_load_factor. */
int mf6hip_apply_op(void *handle, int op, const double *x, double *y);

/* Upload amat (reference CSR layout) and factor it (ims_base_pcilu0,
 * ImsLinearBase.f90:928-1042) on the device without solving. */
int mf6hip_load_factor(void *handle, const double *amat);

/* Copy the device-resident system of the last solve back (reference CSR
 * layout, after the sln_ls row fixes); either pointer may be NULL. */
int mf6hip_get_system(void *handle, double *amat_out, double *rhs_out);

/* ---------------------------------------------------------------------
 * Residual norms without allocation.  No counterpart in the reference: a
 * variant of mf6hip_gwf_residual_norms for callers that cannot allocate
 * device memory (code under HIP graph capture, tight coupling loops, hosts
 * that forbid hipMalloc).  It computes r = A x - b (sln_calc_residual,
 * NumericalSolution.f90:2966-2982) of a system the caller holds on the
 * device in the reference's CSR layout (1-based ia/ja, diagonal first),
 * over rows with active[n] >= 1 (every row when active is NULL), and
 * writes, asynchronously on `stream` (NULL: the null stream), four
 * doubles to the device array `out`: ||r||_2 (sln_l2norm, :2855-2871),
 * ||b||_2 over the same rows, the signed r of largest magnitude and its
 * 1-based row (0 when no row is active).
 * It calls no allocator: its per-block partial sums live in a fixed device
 * array of 4 x 2048 doubles (64 KiB) that each GPU the library runs on
 * holds until the process ends.  neq must be in [1, 2^30] (MF6HIP_EINVAL
 * otherwise).  Every call on a GPU shares that GPU's array: calls that
 * overlap on one GPU (two streams, two threads) race and must be
 * serialised by the caller. */
int mf6hip_csr_residual_norms_noalloc(int neq, int nja, const int *ia, const int *ja,
                                      const double *amat, const double *rhs,
                                      const double *x, const int *active,
                                      double *out, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* MF6HIP_H */
