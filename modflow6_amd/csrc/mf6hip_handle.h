// mf6hip_handle.h -- the per-solution handle behind the C ABI.
#pragma once
#include <string>
#include <vector>

#include "../../include/mf6hip.h"
#include "mf6hip_common.h"

void mf6hip_set_error(const std::string &msg);

#define HIPCHK(call)                                                        \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess) {                                                 \
      mf6hip_set_error(std::string("HIP error '") + hipGetErrorString(e_) + \
                       "' at " + __FILE__ + ":" + std::to_string(__LINE__)); \
      return MF6HIP_EHIP;                                                   \
    }                                                                       \
  } while (0)

// Package data on the device (one per MF6HIP_PKG type).
struct DevPackage {
  int nbound = 0;
  int *nodes = nullptr;    // natural, 0-based
  double *v1 = nullptr;
  double *v2 = nullptr;
  double *hcof = nullptr;  // bnd_cf results
  double *brhs = nullptr;
  // options (mf6hip_package_options): WEL AUTO_FLOW_REDUCE, DRN depth
  int iflowred = 0;
  double flowred = 0.0;    // as wel_options stores it, (0, 1]
  int idepth = 0;
  int icubic = 0;          // DEV_CUBIC_SCALING given
  double *depth = nullptr; // DRN depths (idepth)
};

struct Mf6Handle {
  Layout L{};
  mf6hip_ims_settings s{};
  int neq = 0, nja = 0;
  hipStream_t stream = nullptr;
  int nblk = 0;           // blocks of the reduction kernels

  // CSR structure (0-based row starts) for host <-> plane conversion
  int *d_ia = nullptr;
  std::vector<int> h_ia;

  // FP64 system in column red-black layout
  double *A = nullptr;    // 7 planes
  double *b = nullptr;    // rhs
  double *x = nullptr;    // heads
  double *xtemp = nullptr;
  double *invpiv = nullptr;   // ILU0 / MILU0 inverse pivots
  double *invpiv_tmp = nullptr;
  int *d_fail = nullptr;      // first failing permuted row of the factor
  // Krylov work vectors (FP64)
  double *d = nullptr, *p = nullptr, *p2 = nullptr, *q = nullptr, *z = nullptr,
         *zf = nullptr;
  double *t = nullptr, *v = nullptr, *dhat = nullptr, *phat = nullptr,
         *qhat = nullptr;
  // learned preconditioner (flexible CG)
  void *unet = nullptr;          // attached network (not owned)
  float *unet_feat = nullptr;    // [2][N] features, natural order
  int learned_iters = 0;         // iterations of the last solve that used it
  int fcg_m = 8;                 // flexible CG: directions kept
  int nn_every = 1;              // flexible CG: network on every nn_every-th iteration
  double *fcg_store = nullptr;   // [2][fcg_m][NP]: directions p_j and A p_j
  // mixed precision (FP32 inner solve)
  float *A32 = nullptr, *invpiv32 = nullptr;
  unsigned short *invpiv16 = nullptr;
  // MIXED_FP16 fused CG: ILU0 planes in FP16 scaled by pc_scale = 2^-k, with
  // the FP32 inverse pivots scaled by 2^k (invpiv32s); pc_scale keeps the
  // largest off-diagonal below 2^14
  _Float16 *A16 = nullptr;
  float *invpiv32s = nullptr;
  float pc_scale = 1.0f;
  double *d_amax = nullptr;
  float *r32 = nullptr, *e32 = nullptr, *p32 = nullptr, *p232 = nullptr,
        *q32 = nullptr, *z32 = nullptr, *zf32 = nullptr;
  float *dhat32 = nullptr, *qhat32 = nullptr, *t32 = nullptr;   // mixed BiCGSTAB
  // reductions / state
  double *d_part = nullptr;   // [8][MF6_MAXBLK]
  MaxLoc *d_mpart = nullptr;  // [4][MF6_MAXBLK]
  KState *d_state = nullptr;
  KState *h_state = nullptr;  // pinned mirror
  double *d_hist = nullptr;   // [4][hist_cap]  dvmax rmax alpha omega
  double *d_dscale = nullptr; // [2][NP] SCALING_METHOD row and column factors
  int *d_hloc = nullptr;      // [2][hist_cap]
  int hist_cap = 0;
  // scratch for host transfers in natural order
  double *d_tmp_nat = nullptr;
  double *d_amat_csr = nullptr;

  // GWF model (device-resident formulate)
  bool has_model = false;
  mf6hip_gwf_model m{};
  double *top = nullptr, *bot = nullptr, *area = nullptr, *k11 = nullptr, *k22 = nullptr,
         *sat0 = nullptr, *hn0 = nullptr,
         *k33 = nullptr, *ss = nullptr, *sy = nullptr;
  double *delr = nullptr, *delc = nullptr;
  int *icelltype = nullptr, *iconvert = nullptr;
  int *ibound = nullptr;      // positions; CHD cells negative
  int *ibound0 = nullptr;     // as uploaded (before CHD)
  double *condsat = nullptr;  // 3 planes: right, front, down (per position)
  // horizontal flow barriers folded into condsat (gwf-hfb.f90 condsat_modify):
  // the condsat index of each barrier's face and its value before that
  // barrier, for the reset before the next set (condsat_reset)
  int hfb_n = 0;
  int *d_hfb_face = nullptr;
  double *d_hfb_sav = nullptr;
  // barriers next to a convertible cell without NEWTON (hfb_fc / hfb_cq at
  // every formulate and budget): positions of n and m, direction index of m
  // seen from n, hydchr, face width, hfb_fc's condsav; per position its
  // barriers (b * 2 + role) in barrier order
  int hfbd_n = 0;
  int *d_hfbd_np = nullptr, *d_hfbd_mp = nullptr, *d_hfbd_s = nullptr;
  double *d_hfbd_hy = nullptr, *d_hfbd_w = nullptr, *d_hfbd_sav = nullptr;
  int *d_hfbd_off = nullptr, *d_hfbd_ent = nullptr;
  std::vector<double> h_delr, h_delc;
  // ghost-node correction (explicit GNC of the model): per GNC the positions
  // of n and m, the direction index of m seen from n (0..5: up, back, left,
  // right, front, down), numjs contributing positions (-1: none) and
  // weights, the conductance gnc_fmsav saved; per position the GNCs that
  // touch its row (g * 2 + role, role 1: the row is m), in GNC order
  int gnc_n = 0, gnc_numjs = 0;
  int *d_gnc_np = nullptr, *d_gnc_mp = nullptr, *d_gnc_s = nullptr, *d_gnc_jp = nullptr;
  double *d_gnc_a = nullptr, *d_gnc_cond = nullptr;
  int *d_gnc_off = nullptr, *d_gnc_ent = nullptr;
  std::vector<int> h_icelltype;   // natural order, host copy for the checks
  double *sat = nullptr;      // npf sat of the last formulate (per position)
  // budget (gwf_cq + gwf_bd) work arrays, allocated on first use
  double *flowja = nullptr;   // nja, CSR order
  double *strg = nullptr;     // [2][NP]: STO-SS, STO-SY rates
  double *bsim = nullptr;     // per-bound rates of WEL, DRN, RCH, then CHD in / out / rate
  size_t bsim_n = 0;
  // the per-bound segments of bsim at the last budget (WEL, DRN, RCH, CHD
  // rate): offset and length; budget_done once a budget has run
  size_t bsim_seg_off[4] = {0, 0, 0, 0};
  int bsim_seg_len[4] = {0, 0, 0, 0};
  int budget_done = 0;
  int *chd_flag = nullptr;    // per position: a CHD cell
  double *xold = nullptr;
  double *dxold = nullptr;
  double *wsave = nullptr, *hchold = nullptr, *deold = nullptr;   // DBD state
  int *ibotnode = nullptr;    // position of the lowest connected cell
  int *bnd_off = nullptr;     // per position: start into bnd_ent (N+1)
  int *bnd_ent = nullptr;     // entries: (pkg << 28) | index
  int nbnd_ent = 0;
  bool bnd_dirty = true;
  DevPackage pk[5];
  std::vector<int> pk_host_nodes[5];
  int iss = 1;
  double delt = 1.0;

  // timing
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<hipEvent_t> spmv_ev;
  int spmv_sample = 16;
  double last_solve_ms = 0, last_spmv_ms = 0;
  int last_spmv_n = 0;
  bool timing = false;
  // fused FP64 CG: eight iterations captured once per handle as a HIP graph
  // (used when the per-launch SpMV timing is off)
  hipGraphExec_t cg_graph = nullptr;
  bool cg_graph_failed = false;
  // two-level preconditioner (preconditioner = 2): coarse shape, per-column
  // coupled-cell masks, dense coarse matrix / Cholesky workspace G, the
  // inverse (FP64 for the FP64 solve, FP32 for the mixed ones), restricted
  // residual and coarse correction
  TwoLevel tl{};
  bool tl_ready = false;         // every two-level buffer below is allocated
  unsigned *tl_cmask = nullptr;
  double *tl_G = nullptr, *tl_A64 = nullptr, *tl_rc = nullptr, *tl_yc64 = nullptr;
  float *tl_A32 = nullptr, *tl_yc32 = nullptr;
  int *tl_info = nullptr;
  double *tl_colsum = nullptr;   // per grid column: residual summed over coupled cells
  double *tl_S3 = nullptr;       // compact rows of S = -Ac (diagonal, west, north)
  double *tl_Lc = nullptr;       // banded Cholesky factor of S by columns, nc x (naj + 1)
  double *tl_Lr = nullptr;       // the same factor by rows
  // deflated form (preconditioner = 3): per-cell row sums over the coupled
  // cells, couplings across aggregate edges, per-column sums of q (colour
  // layout, c HP + h2)
  float *tl_s32 = nullptr, *tl_ew = nullptr, *tl_ns = nullptr;
  double *tl_colsq = nullptr;
  void *tl_blas = nullptr;    // rocblas_handle
  int spmv_calls = 0;
  // XMI-style outer loop (mf6hip_xmi.hip): NONLINEAR settings, the time
  // step being solved, the outer iteration counter and the state the
  // NONLINEAR options keep between outer iterations
  struct Xmi {
    mf6hip_nonlinear nl{};
    bool nl_set = false;
    bool prepared = false;
    int kper = 0, kstp = 0, iss = 1, kiter = 0, converged = 0;
    double delt = 1.0;
    mf6hip_underrelax ur{};
    mf6hip_ptc ptc{};
    double res_prev = 0.0, res_new = 0.0;
    mf6hip_xmi_iteration last{};
  } xmi;
};

// two-level preconditioner (mf6hip_coarse.hip)
TwoLevel mf6hip_tl_shape(const Layout &L, int target);
int mf6hip_tl_setup_slice(const Layout &L, const double *A, const TwoLevel &T,
                          unsigned *cmask, double *S3, double *Lc, double *Lr, double *G,
                          float *A32, int *dinfo, int *info, hipStream_t s);
int mf6hip_tl_setup(Mf6Handle *h);
void mf6hip_tl_free(Mf6Handle *h);
template <typename TV>
void mf6hip_tl_enqueue(Mf6Handle *h, const TV *d, TV *yc, double *part, int nslot,
                       const KState *st, const double *alpha = nullptr);
template <typename TV>
void mf6hip_tl_colsum(Mf6Handle *h, const TV *d);
int mf6hip_tl_apply64(Mf6Handle *h, const double *r, double *z);
int mf6hip_tl_prolong64(Mf6Handle *h, double *z);
int mf6hip_tl_gather_check(Mf6Handle *h, const double *x, double *out);
// partial slots of rc.yc the coarse GEMV writes (of nslot cleared)
int mf6hip_tl_nslot(const Mf6Handle *h, int nslot);
// parity helpers on the fused CG's operators (ops 4-8 of mf6hip_apply_op)
int mf6hip_fused_op(Mf6Handle *h, int op, const double *x, double *y);

// host checks of a system in the reference's CSR layout (1-based ia / ja)
int mf6hip_check_structure(const Layout &L, int neq, int nja, const int *ia, const int *ja);
int mf6hip_check_symmetric_csr(const Layout &L, const int *ia, const double *amat);

// solver entry points used by the GWF part
int mf6hip_device_solve(Mf6Handle *h, int kstp, int kiter, int *icnvg,
                        int *in_iter, mf6hip_cnvg_out *out);
int mf6hip_launch_row_fixes(Mf6Handle *h);
int mf6hip_copy_system(Mf6Handle *h, double *amat_out, double *rhs_out);
int mf6hip_run_cg_fused(Mf6Handle *h);
int mf6hip_run_cg_reliable(Mf6Handle *h, int *icnvg, int *iters);
int mf6hip_is_symmetric(Mf6Handle *h, int *sym);
int mf6hip_dxmax(Mf6Handle *h, double *hncg, int *lrch, const double *xa,
                 const double *xb, bool active_only);

// solver pieces shared with the flexible CG (mf6hip_fcg.hip)
// z = ILU0(r); the kernels return at once when st is given and st->stop is set
int mf6hip_ilu_apply64(Mf6Handle *h, const double *r, double *z, const KState *st);
int mf6hip_spmv64(Mf6Handle *h, const double *x, double *y);
int mf6hip_run_fcg(Mf6Handle *h);
int mf6hip_stage_system(const Layout &L, const int *ia1, int nja, const double *amat,
                        const double *rhs, const double *x0, double relax, double *A64,
                        double *b64, double *x64, double *piv64, hipStream_t s);
// network entry points used by the solver (mf6hip_unet.hip)
int mf6hip_unet_apply_device(void *net, hipStream_t s);
unsigned short *mf6hip_unet_input(void *net);
float *mf6hip_unet_output(void *net);
int mf6hip_unet_dims(void *net, int *nlay, int *nrow, int *ncol, int *nin);
