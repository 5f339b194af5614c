// mf6hip_gwf.hip -- device-resident GWF formulate and the outer-loop pieces.
//
// Reference routines (src/Model/GroundWaterFlow, src/Model/ModelUtilities,
// src/Solution/NumericalSolution.f90):
//   calc_condsat gwf-npf.f90:1950-2037, npf_cf :444-470 (thksat :775-794),
//   npf_fc :474-574, npf_fn :578-698, npf_nur :705-741,
//   hcond / vcond / condmean GwfConductanceUtils.f90:43-284,
//   sto_fc gwf-sto.f90:226-345, sto_fn :353-445, SsTerms / SyTerms /
//   SsCapacity / SyCapacity GwfStorageUtils.f90:32-180,
//   sQuadraticSaturation(Derivative) SmoothingFunctions.f90:275-406,
//   bnd_fc BoundaryPackage.f90:453-479, wel_cf gwf-wel.f90:296-332,
//   rch_cf gwf-rch.f90:303-353, drn_cf gwf-drn.f90:340-373,
//   gwf_fc order gwf.f90:466-555, sln_ls row fixes NumericalSolution.f90:
//   2434-2478, sln_get_dxmax :3122-3153, sln_calcdx :2912, sln_maxval :2875.
//
// The reference scatters each connection's terms into both rows while
// looping over nodes in ascending order.  Here one thread owns one row and
// gathers: the diagonal and rhs of a row receive their terms in ascending
// neighbour order -- the order the scatter loop delivers them -- and every
// pair term is computed with the reference's argument order (lower node
// first), so each entry is bit-identical to the reference's (no FMA
// contraction; build with -ffp-contract=off).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "mf6hip_handle.h"

#define CK(x)                         \
  do {                                \
    int rc_ = (x);                    \
    if (rc_ != MF6HIP_OK) return rc_; \
  } while (0)

static inline int ggrid(int n) {
  int g = (n + MF6_BLOCK - 1) / MF6_BLOCK;
  if (g > MF6_MAXBLK) g = MF6_MAXBLK;
  return g < 1 ? 1 : g;
}

template <typename T>
static int galloc(T **p, size_t n) {
  if (*p) return MF6HIP_OK;
  if (hipMalloc((void **)p, n * sizeof(T) + 256) != hipSuccess) {
    mf6hip_set_error("hipMalloc failed (GWF model arrays)");
    return MF6HIP_EHIP;
  }
  return MF6HIP_OK;
}

// ---------------------------------------------------------------------------
// stateless helpers (reference statement order)
// ---------------------------------------------------------------------------

__host__ __device__ inline double sqsat(double top, double bot, double x, double eps) {
  double teps = eps, y;
  double b = top - bot;
  if (b > 0.0) {
    double br;
    if (x < bot) br = 0.0;
    else if (x > top) br = 1.0;
    else br = (x - bot) / b;
    double av = 1.0 / (1.0 - teps);
    double bri = 1.0 - br;
    if (br < teps) y = av * 0.5 * (br * br) / teps;
    else if (br < (1.0 - teps)) y = av * br + 0.5 * (1.0 - av);
    else if (br < 1.0) y = 1.0 - ((av * 0.5 * (bri * bri)) / teps);
    else y = 1.0;
  } else {
    y = (x < bot) ? 0.0 : 1.0;
  }
  return y;
}

__device__ inline double sqsat_deriv(double top, double bot, double x, double eps) {
  double teps = eps, y, br;
  double b = top - bot;
  if (x < bot) br = 0.0;
  else if (x > top) br = 1.0;
  else br = (x - bot) / b;
  double av = 1.0 / (1.0 - teps);
  double bri = 1.0 - br;
  if (br < teps) y = av * br / teps;
  else if (br < (1.0 - teps)) y = av;
  else if (br < 1.0) y = av * bri / teps;
  else y = 0.0;
  return y / b;
}

// GwfStorageUtils.f90:32-96 SsTerms with STO's flags (iorig_ss: the original
// formulation; iconf_ss: SS_CONFINED_ONLY); rho1 from SsCapacity
// (:143-160: ss times the thickness, or times 1 with STORAGECOEFFICIENT)
__device__ inline void ss_terms(int iconvert, int iorig_ss, int iconf_ss, double top,
                                double bot, double rho1, double rho1old, double snnew,
                                double snold, double hold, double &aterm, double &rhsterm) {
  aterm = -rho1 * snnew;
  rhsterm = 0.0;
  if (iconvert != 0 && iorig_ss == 0) {
    if (iconf_ss == 0) {
      const double tthk = top - bot;
      const double zold = bot + 0.5 * tthk * snold;
      const double znew = bot + 0.5 * tthk * snnew;
      rhsterm = -rho1old * snold * (hold - zold) - rho1 * snnew * znew;
    } else {
      if (snold == 1.0) rhsterm = rhsterm - rho1old * (hold - top);
      if (snnew == 1.0) rhsterm = rhsterm - rho1 * top;
      else aterm = 0.0;
    }
  } else {
    rhsterm = -rho1old * snold * hold;
  }
}

// SmoothingFunctions.f90:412-460 sQSaturation (cubic smoothing from bot
// to top with coefficients c1 / b^3, c2 / b^2); the reference's x**3.0 and
// x**2.0 compile to (x x) x and x x
__device__ inline double sqsat_cubic(double top, double bot, double x, double c1, double c2) {
  const double w = x - bot, b = top - bot;
  const double s = w / b;
  const double cof1 = c1 / ((b * b) * b), cof2 = c2 / (b * b);
  if (s < 0.0) return 0.0;
  if (s < 1.0) return cof1 * ((w * w) * w) + cof2 * (w * w);
  return 1.0;
}

// SmoothingFunctions.f90:466-516 sQSaturationDerivative
__device__ inline double sqsat_cubic_deriv(double top, double bot, double x, double c1,
                                           double c2) {
  const double w = x - bot, b = top - bot;
  const double s = w / b;
  const double cof1 = c1 * 3.0 / ((b * b) * b), cof2 = c2 * 2.0 / (b * b);
  if (s < 0.0) return 0.0;
  if (s < 1.0) return cof1 * (w * w) + cof2 * w;
  return 0.0;
}

__device__ inline double logmean(double d1, double d2) {
  double drat = d2 / d1;
  if (drat <= 0.995 || drat >= 1.005) return (d2 - d1) / log(drat);
  return 0.5 * (d1 + d2);
}

__device__ inline double condmean(double k1, double k2, double thick1, double thick2,
                                  double cl1, double cl2, double width, int iavg) {
  double t1 = k1 * thick1;
  double t2 = k2 * thick2;
  if (iavg == 0) return (t1 * t2 > 0.0) ? width * t1 * t2 / (t1 * cl2 + t2 * cl1) : 0.0;
  if (iavg == 1) {
    double tm = (t1 * t2 > 0.0) ? logmean(t1, t2) : 0.0;
    return tm * width / (cl1 + cl2);
  }
  if (iavg == 2) {
    double km = (k1 * k2 > 0.0) ? logmean(k1, k2) : 0.0;
    return km * 0.5 * (thick1 + thick2) * width / (cl1 + cl2);
  }
  double denom = (k1 * cl2 + k2 * cl1);
  double km = denom > 0.0 ? k1 * k2 / denom : 0.0;
  return km * 0.5 * (thick1 + thick2) * width;
}

__device__ inline double hcond(int ibdn, int ibdm, int ictn, int ictm, int iupstream,
                               int icellavg, double condsat, double hn, double hm,
                               double satn, double satm, double hkn, double hkm,
                               double topn, double topm, double botn, double botm,
                               double cln, double clm, double fawidth) {
  if (ibdn == 0 || ibdm == 0) return 0.0;
  if (ictn == 0 && ictm == 0) return condsat;
  if (iupstream == 1) {
    double sat_up = (hn > hm) ? satn : satm;
    return sat_up * condsat;
  }
  double thksatn = satn * (topn - botn);
  double thksatm = satm * (topm - botm);
  return condmean(hkn, hkm, thksatn, thksatm, cln, clm, fawidth, icellavg);
}

__device__ inline double vcond(int ibdn, int ibdm, int ictn, int ictm, int ivarcv,
                               int idewatcv, double condsat, double hn, double hm,
                               double vkn, double vkm, double satn, double satm,
                               double topn, double topm, double botn, double botm,
                               double flowarea) {
  if (ibdn == 0 || ibdm == 0) return 0.0;
  if (ivarcv == 0) return condsat;
  if (ictn == 0 && ictm == 0) return condsat;
  if (hn >= topn && hm >= topm) return condsat;
  double satntmp = satn, satmtmp = satm;
  if (idewatcv == 0) {
    if (botn > botm) satmtmp = 1.0;
    else satntmp = 1.0;
  }
  double bovk1 = satntmp * (topn - botn) * 0.5 / vkn;
  double bovk2 = satmtmp * (topm - botm) * 0.5 / vkm;
  double denom = (bovk1 + bovk2);
  return denom != 0.0 ? flowarea / denom : 0.0;
}

// ---------------------------------------------------------------------------
// model arrays (position order) and connection geometry
// ---------------------------------------------------------------------------

struct GwfDev {
  Layout L;
  const double *top, *bot, *area, *k11, *k22, *k33, *ss, *sy, *delr, *delc, *condsat;
  double *sat;   // npf_cf's sat, written by the formulate (read by the budget)
  const double *sat0, *hn0;   // calc_condsat's initial sat and heads (THICKSTRT)
  const int *icelltype, *iconvert, *ibound;
  int inewton, ivarcv, idewatcv, icellavg, has_sto;
  int istor_coef, iconf_ss, iorig_ss;   // STO options
  int fn;   // add the Newton terms (npf_fn, sto_fn): gwf_fc inwtflag
  double satomega;
};

__device__ inline double cell_sat(const GwfDev &g, int p, double h) {
  // npf_cf: sat of convertible cells; confined cells keep sat = 1
  if (g.icelltype[p] == 0) return 1.0;
  if (g.ibound[p] == 0) return 0.0;
  double s;
  if (h >= g.top[p]) s = 1.0;
  else s = (h - g.bot[p]) / (g.top[p] - g.bot[p]);
  if (g.inewton != 0) s = sqsat(g.top[p], g.bot[p], h, g.satomega);
  return s;
}

// geometry of the connection between a (lower node) and b (higher node);
// dir 0 = right (j+1), 1 = front (i+1), 2 = down (k+1)
struct Conn {
  int ihc;
  double cl1, cl2, hwva;
};

__device__ inline Conn conn_geom(const GwfDev &g, int dir, int pa, int pb, int ia_, int ja_) {
  Conn c;
  if (dir == 0) {
    c.ihc = 1; c.cl1 = 0.5 * g.delr[ja_]; c.cl2 = 0.5 * g.delr[ja_ + 1]; c.hwva = g.delc[ia_];
  } else if (dir == 1) {
    c.ihc = 1; c.cl1 = 0.5 * g.delc[ia_]; c.cl2 = 0.5 * g.delc[ia_ + 1]; c.hwva = g.delr[ja_];
  } else {
    c.ihc = 0;
    c.cl1 = 0.5 * (g.top[pa] - g.bot[pa]);
    c.cl2 = 0.5 * (g.top[pb] - g.bot[pb]);
    c.hwva = g.delr[ja_] * g.delc[ia_];
  }
  return c;
}

// horizontal conductivity of a face direction: hy_eff (gwf-npf.f90:
// 2280-2355) without rotation angles is K22 on row-direction faces (dir 1)
// and K11 on column-direction ones
__device__ inline double hk(const GwfDev &g, int dir, int p) {
  return dir == 1 ? g.k22[p] : g.k11[p];
}

// npf_fc conductance of the pair (a < b) as the reference computes it
__device__ inline double pair_cond(const GwfDev &g, int dir, int pa, int pb, int ia_,
                                   int ja_, double ha, double hb, double csat) {
  Conn c = conn_geom(g, dir, pa, pb, ia_, ja_);
  double sata = cell_sat(g, pa, ha), satb = cell_sat(g, pb, hb);
  if (c.ihc == 0) {
    return vcond(g.ibound[pa], g.ibound[pb], g.icelltype[pa], g.icelltype[pb], g.ivarcv,
                 g.idewatcv, csat, ha, hb, g.k33[pa], g.k33[pb], sata, satb, g.top[pa],
                 g.top[pb], g.bot[pa], g.bot[pb], c.hwva);
  }
  return hcond(g.ibound[pa], g.ibound[pb], g.icelltype[pa], g.icelltype[pb], g.inewton,
               g.icellavg, csat, ha, hb, sata, satb, hk(g, dir, pa), hk(g, dir, pb), g.top[pa],
               g.top[pb], g.bot[pa], g.bot[pb], c.cl1, c.cl2, c.hwva);
}

// qcalc's conductance (gwf-npf.f90:797-862): npf_fc's formula with the sat of
// the last npf_cf and the heads of the budget
__device__ inline double pair_cond_sat(const GwfDev &g, int dir, int pa, int pb, int ia_,
                                       int ja_, double ha, double hb, double sata,
                                       double satb, double csat) {
  Conn c = conn_geom(g, dir, pa, pb, ia_, ja_);
  if (c.ihc == 0) {
    return vcond(g.ibound[pa], g.ibound[pb], g.icelltype[pa], g.icelltype[pb], g.ivarcv,
                 g.idewatcv, csat, ha, hb, g.k33[pa], g.k33[pb], sata, satb, g.top[pa],
                 g.top[pb], g.bot[pa], g.bot[pb], c.hwva);
  }
  return hcond(g.ibound[pa], g.ibound[pb], g.icelltype[pa], g.icelltype[pb], g.inewton,
               g.icellavg, csat, ha, hb, sata, satb, hk(g, dir, pa), hk(g, dir, pb), g.top[pa],
               g.top[pb], g.bot[pa], g.bot[pb], c.cl1, c.cl2, c.hwva);
}

// calc_condsat (upper connections, saturation 1): plane dir at position of
// the lower node a
__global__ void k_condsat(GwfDev g, double *__restrict__ condsat) {
  const Layout &L = g.L;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < L.NP; p += gridDim.x * blockDim.x) {
    int c, k, i, j, l;
    if (!pos_to_cell(L, p, c, k, i, j, l)) continue;
    for (int dir = 0; dir < 3; dir++) {
      int kb = k + (dir == 2), ib = i + (dir == 1), jb = j + (dir == 0);
      double cs = 0.0;
      if (kb < L.nlay && ib < L.nrow && jb < L.ncol) {
        int pb = cell_pos(L, kb, ib, jb);
        Conn cn = conn_geom(g, dir, p, pb, i, j);
        double topn = g.top[p], botn = g.bot[p], topm = g.top[pb], botm = g.bot[pb];
        // calc_initial_sat and THICKSTRT's heads (gwf-npf.f90:1997-2006):
        // 1 and the cell top except for THICKSTRT cells
        const double satn = g.sat0[p], satm = g.sat0[pb];
        if (cn.ihc == 0) {
          cs = vcond(1, 1, 1, 1, 1, 1, 1.0, botn, botm, g.k33[p], g.k33[pb], satn, satm, topn,
                     topm, botn, botm, cn.hwva);
        } else {
          cs = hcond(1, 1, 1, 1, 0, g.icellavg, 1.0, g.hn0[p], g.hn0[pb], satn, satm,
                     hk(g, dir, p), hk(g, dir, pb), topn, topm, botn, botm, cn.cl1, cn.cl2,
                     cn.hwva);
        }
      }
      condsat[dir * L.NP + p] = cs;
    }
  }
}

// Horizontal flow barriers folded into condsat (gwf-hfb.f90:789-832
// condsat_modify), in barrier order (a second barrier on a face combines
// with the first's result): the face's condsat is saved, then replaced,
// under NEWTON or when both cells are confined, by the series combination
// with the barrier, cond * condhfb / (cond + condhfb), condhfb = hydchr *
// width * (mean saturated thickness), or by -cond * hydchr for hydchr <= 0
// (a multiplier).  One thread: the list is short and the order matters.
// nn / mm are 0-based natural node numbers of horizontally adjacent cells.
__global__ void k_hfb_apply(GwfDev g, int nhfb, const int *__restrict__ nn,
                            const int *__restrict__ mm, const double *__restrict__ hydchr,
                            double *__restrict__ condsat, int *__restrict__ face,
                            double *__restrict__ sav) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const Layout &L = g.L;
  for (int b = 0; b < nhfb; b++) {
    const int n = nn[b], m = mm[b];
    const int a = n < m ? n : m, c = n < m ? m : n;
    const int r = a % L.NRC, i = r / L.ncol, j = r - i * L.ncol, k = a / L.NRC;
    const int dir = c - a == 1 ? 0 : 1;   // right or front neighbour
    const int pa = cell_pos(L, k, i, j);
    const int pc = dir == 0 ? cell_pos(L, k, i, j + 1) : cell_pos(L, k, i + 1, j);
    const int f = dir * L.NP + pa;
    face[b] = f;
    double cond = condsat[f];
    sav[b] = cond;
    const int pn = n == a ? pa : pc, pm = n == a ? pc : pa;
    if (g.inewton == 1 || (g.icelltype[pn] == 0 && g.icelltype[pm] == 0)) {
      const double topn = g.top[pn], topm = g.top[pm], botn = g.bot[pn], botm = g.bot[pm];
      // DIS horizontal connections have ihc = 1: the mean thickness
      const double faheight = 0.5 * ((topn - botn) + (topm - botm));
      if (hydchr[b] > 0.0) {
        const double fawidth = conn_geom(g, dir, pa, pc, i, j).hwva;
        const double condhfb = hydchr[b] * fawidth * faheight;
        cond = cond * condhfb / (cond + condhfb);
      } else {
        cond = -cond * hydchr[b];
      }
      condsat[f] = cond;
    }
  }
}

// condsat_reset (gwf-hfb.f90:770-781): every face back to the value saved
// before its barrier, in barrier order
__global__ void k_hfb_reset(int nhfb, const int *__restrict__ face,
                            const double *__restrict__ sav, double *__restrict__ condsat) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  for (int b = 0; b < nhfb; b++) condsat[face[b]] = sav[b];
}

// One package's data and options as the kernels see them.
struct PkgDev {
  int kind, nb;
  const int *nodes;
  const double *v1, *v2, *depth;
  int iflowred;
  double flowred;
  int idepth, icubic;   // icubic: cubic drain scaling (NEWTON or DEV_CUBIC_SCALING)
};

// get_drain_elevations (gwf-drn.f90:501-531): with a depth d /= 0 the
// conductance is scaled between elev and elev + d
__device__ inline void drn_elevations(const PkgDev &P, int e, double &drndepth, double &drntop,
                                      double &drnbot) {
  const double el = P.v1[e];
  drndepth = P.idepth ? P.depth[e] : 0.0;
  if (drndepth != 0.0) {
    const double t = el + drndepth;
    drntop = t > el ? t : el;
    drnbot = t < el ? t : el;
  } else {
    drntop = el;
    drnbot = el;
  }
}

// bnd_cf of the WEL / DRN / RCH packages: wel_cf (gwf-wel.f90:296-332, with
// AUTO_FLOW_REDUCE), drn_cf with get_drain_factor (gwf-drn.f90:340-373,
// :534-574), rch_cf
__global__ void k_bnd_cf(PkgDev P, GwfDev g, const double *__restrict__ x,
                         double *__restrict__ hcof, double *__restrict__ brhs) {
  const Layout &L = g.L;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < P.nb; e += gridDim.x * blockDim.x) {
    int p = pos_of_natural(L, P.nodes[e]);
    if (P.kind == MF6HIP_PKG_WEL) {
      hcof[e] = 0.0;
      double q = P.v1[e];
      if (P.iflowred != 0 && q < 0.0 && g.icelltype[p] != 0) {
        const double bt = g.bot[p];
        const double tp = bt + P.flowred * (g.top[p] - bt);
        q = q * sqsat_cubic(tp, bt, x[p], -2.0, 3.0);
      }
      brhs[e] = g.ibound[p] <= 0 ? 0.0 : -q;
    } else if (P.kind == MF6HIP_PKG_DRN) {
      if (g.ibound[p] <= 0) {
        hcof[e] = 0.0;
        brhs[e] = 0.0;
      } else {
        double cdrn = P.v2[e];
        double drndepth, drntop, drnbot, fact;
        drn_elevations(P, e, drndepth, drntop, drnbot);
        const double h = x[p];
        if (drndepth != 0.0) {
          fact = P.icubic ? sqsat_cubic(drntop, drnbot, h, -1.0, 2.0)
                          : sqsat(drntop, drnbot, h, 0.0);
        } else {
          fact = h <= drnbot ? 0.0 : 1.0;
        }
        brhs[e] = -fact * cdrn * drnbot;
        hcof[e] = -fact * cdrn;
      }
    } else if (P.kind == MF6HIP_PKG_RCH) {
      hcof[e] = 0.0;
      double rr = -P.v1[e] * g.area[p];
      brhs[e] = g.ibound[p] <= 0 ? 0.0 : rr;
    } else {
      hcof[e] = 0.0;
      brhs[e] = 0.0;
    }
  }
}

struct BndDev {
  const int *off, *ent;
  const double *hcof[5], *brhs[5];
  PkgDev pk[5];
};

// Barriers that hfb_fc / hfb_cq apply (the handle's d_hfbd_* arrays).
struct HfbDev {
  const int *off, *ent;     // per position: b * 2 + role (role 1: the row is m)
  const int *np, *mp, *s;   // positions of n and m; direction index of m from n
  const double *hy, *w;     // hydchr, face width
  double *sav;              // hfb_fc's condsav (written by the row of n)
};

// The model's explicit ghost-node corrections as the kernels see them
// (the handle's d_gnc_* arrays).
struct GncDev {
  int n, numjs;
  const int *off, *ent;     // per position: g * 2 + role (role 1: the row is m)
  const int *np, *mp, *s;   // positions of n and m; direction index of m from n
  const int *jp;            // numjs contributing positions per GNC (-1: none)
  const double *a;          // their weights
  double *cond;             // the conductance gnc_fmsav saves
};

// sln_buildsystem for one model: one thread per row
__global__ void __launch_bounds__(MF6_BLOCK) k_formulate(GwfDev g, BndDev bd, GncDev G, HfbDev F,
                                                         int iss,
                                                         double delt,
                                                         const double *__restrict__ x,
                                                         const double *__restrict__ xold,
                                                         double *__restrict__ A,
                                                         double *__restrict__ rhs) {
  const Layout &L = g.L;
  const int N = L.NP;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < L.NP; p += gridDim.x * blockDim.x) {
    int c, k, i, j, l;
    if (!pos_to_cell(L, p, c, k, i, j, l)) continue;
    const double hcell = x[p];
    g.sat[p] = cell_sat(g, p, hcell);   // npf_cf (gwf-npf.f90:444-470)
    // neighbour table in ascending natural order: up, back, left, right,
    // front, down.  nb = position, dir = connection plane, lower = pair
    // (nb, n) has nb first
    int npos[6];
    int pres[6];
    pres[0] = k > 0;
    pres[1] = i > 0;
    pres[2] = j > 0;
    pres[3] = j < L.ncol - 1;
    pres[4] = i < L.nrow - 1;
    pres[5] = k < L.nlay - 1;
    npos[0] = pres[0] ? p - L.HP : -1;
    npos[5] = pres[5] ? p + L.HP : -1;
    {
      const int ob = (1 - c) * L.H + k * L.HP;
      npos[1] = pres[1] ? ob + hn(l, -L.ncol) : -1;
      npos[2] = pres[2] ? ob + hn(l, -1) : -1;
      npos[3] = pres[3] ? ob + hn(l, 1) : -1;
      npos[4] = pres[4] ? ob + hn(l, L.ncol) : -1;
    }
    const int dirs[6] = {2, 1, 0, 0, 1, 2};
    double off[6];
    double diag = 0.0;
    // --- npf_fc
#pragma unroll
    for (int s = 0; s < 6; s++) {
      off[s] = 0.0;
      if (!pres[s]) continue;
      int m = npos[s];
      int dir = dirs[s];
      double cond;
      if (s < 3) {   // m < n: pair (m, n), condsat stored at m
        int im = i - (s == 1), jm = j - (s == 2);
        cond = pair_cond(g, dir, m, p, im, jm, x[m], hcell, g.condsat[dir * N + m]);
      } else {
        cond = pair_cond(g, dir, p, m, i, j, hcell, x[m], g.condsat[dir * N + p]);
      }
      off[s] = off[s] + cond;
      diag = diag + (-cond);
    }
    // --- hfb_fc without NEWTON (gwf-hfb.f90:211-348): each barrier next to
    // a convertible cell, in barrier order, replaces the face's conductance
    // aterm by its series combination with the barrier (wetted thickness:
    // tops clipped at the heads of convertible cells) and adds aterm - cond
    // to the diagonal
    for (int e = F.off ? F.off[p] : 0, e1 = F.off ? F.off[p + 1] : 0; e < e1; e++) {
      const int code = F.ent[e];
      const int bi = code >> 1, role = code & 1;
      const int pn = F.np[bi], pm = F.mp[bi];
      if (g.ibound[pn] == 0 || g.ibound[pm] == 0) continue;
      if (g.icelltype[pn] != 1 && g.icelltype[pm] != 1) continue;
      const int s = role == 0 ? F.s[bi] : 5 - F.s[bi];
      const double aterm = off[s];
      double topn = g.top[pn], topm = g.top[pm];
      const double botn = g.bot[pn], botm = g.bot[pm];
      if (g.icelltype[pn] == 1 && x[pn] < topn) topn = x[pn];
      if (g.icelltype[pm] == 1 && x[pm] < topm) topm = x[pm];
      const double faheight = 0.5 * ((topn - botn) + (topm - botm));
      double cond;
      if (F.hy[bi] > 0.0) {
        const double condhfb = F.hy[bi] * 1.0 * F.w[bi] * faheight;
        cond = aterm * condhfb / (aterm + condhfb);
      } else {
        cond = -aterm * F.hy[bi];
      }
      if (role == 0) F.sav[bi] = cond;
      diag = diag + (aterm - cond);
      off[s] = cond;
    }
    double b = 0.0;
    // --- gnc_fc, explicit (GhostNode.f90:280-324), with gnc_fmsav
    // (:251-273): the conductance is the matrix value at (n, m) after
    // npf_fc; rhs(n) -= alpha cond (x(n) - x(j)), rhs(m) += the same
    for (int e = G.off ? G.off[p] : 0, e1 = G.off ? G.off[p + 1] : 0; e < e1; e++) {
      const int code = G.ent[e];
      const int gi = code >> 1, role = code & 1;
      const double cond = off[role == 0 ? G.s[gi] : 5 - G.s[gi]];
      if (role == 0) G.cond[gi] = cond;
      const int pn = G.np[gi], pm = G.mp[gi];
      if (g.ibound[pn] == 0 || g.ibound[pm] == 0) continue;
      const double xn = x[pn];
      for (int jj = 0; jj < G.numjs; jj++) {
        const int pj = G.jp[gi * G.numjs + jj];
        if (pj < 0) continue;
        const double alpha = G.a[gi * G.numjs + jj];
        if (alpha == 0.0) continue;
        const double aterm = alpha * cond;
        const double rterm = aterm * (xn - x[pj]);
        if (role == 0) b = b - rterm;
        else b = b + rterm;
      }
    }
    // --- sto_fc
    if (g.has_sto && !iss && g.ibound[p] >= 1) {
      double tled = 1.0 / delt;
      double tp = g.top[p], bt = g.bot[p];
      double snold, snnew;
      if (g.iconvert[p] == 0) {
        snold = 1.0;
        snnew = 1.0;
      } else {
        // STO's satomega: 1e-6 under NEWTON, else 0 (gwf-sto.f90:826-829)
        snold = sqsat(tp, bt, xold[p], g.satomega);
        snnew = sqsat(tp, bt, hcell, g.satomega);
      }
      double sc1 = g.ss[p] * (g.istor_coef == 0 ? tp - bt : 1.0) * g.area[p];
      double rho1 = sc1 * tled;
      double rho1old = rho1;
      double aterm, rhsterm;
      ss_terms(g.iconvert[p], g.iorig_ss, g.iconf_ss, tp, bt, rho1, rho1old, snnew, snold,
               xold[p], aterm, rhsterm);
      diag = diag + aterm;
      b = b + rhsterm;
      if (g.iconvert[p] != 0) {
        double rho2 = g.sy[p] * g.area[p] * tled;
        double rho2old = rho2;
        double tthk = tp - bt;
        double at = 0.0, rt;
        if (snnew < 1.0) {
          if (snnew > 0.0) {
            at = -rho2;
            rt = -rho2old * tthk * snold - rho2 * bt;
          } else {
            rt = tthk * (0.0 - rho2old * snold);
          }
        } else {
          rt = tthk * (rho2 * snnew - rho2old * snold);
        }
        diag = diag + at;
        b = b + rt;
      }
    }
    // --- bnd_fc (WEL, DRN, RCH in bndlist order; CHD adds nothing)
    for (int e = bd.off[p]; e < bd.off[p + 1]; e++) {
      int code = bd.ent[e];
      int pk = code >> 28, idx = code & 0x0fffffff;
      b = b + bd.brhs[pk][idx];
      diag = diag + bd.hcof[pk][idx];
    }
    // --- npf_fn
    if (g.fn) {
#pragma unroll
      for (int s = 0; s < 6; s++) {
        if (!pres[s]) continue;
        int dir = dirs[s];
        if (dir == 2 && g.ivarcv == 0) continue;
        int m = npos[s];
        int pa, pb_;
        double csat;
        if (s < 3) { pa = m; pb_ = p; csat = g.condsat[dir * N + m]; }
        else { pa = p; pb_ = m; csat = g.condsat[dir * N + p]; }
        double ha = x[pa], hb = x[pb_];
        int iups = pb_;
        if (hb < ha) iups = pa;
        int idn = (iups == pa) ? pb_ : pa;
        if (g.icelltype[iups] == 0) continue;
        double topup = g.top[iups], botup = g.bot[iups];
        double consterm = -csat * (x[iups] - x[idn]);
        double derv = sqsat_deriv(topup, botup, x[iups], g.satomega);
        if (iups == pa) {
          double term = consterm * derv;
          if (pa == p) {          // this row is n (upstream)
            b = b + term * ha;
            diag = diag + term;
          } else {                // this row is m
            b = b - term * ha;
            if (g.ibound[p] > 0) off[s] = off[s] + (-term);
          }
        } else {
          double term = -consterm * derv;
          if (pa == p) {          // row n, m upstream
            b = b + term * hb;
            if (g.ibound[p] > 0) off[s] = off[s] + term;
          } else {                // row m (upstream)
            b = b - term * hb;
            diag = diag + (-term);
          }
        }
      }
      // --- gnc_fn (GhostNode.f90:340-443): the correction's derivative
      // with respect to a convertible upstream cell's head, with the face's
      // saturated conductance (horizontal faces, vertical ones with
      // VARIABLECV)
      for (int e = G.off ? G.off[p] : 0, e1 = G.off ? G.off[p + 1] : 0; e < e1; e++) {
        const int code = G.ent[e];
        const int gi = code >> 1, role = code & 1;
        const int pn = G.np[gi], pm = G.mp[gi];
        if (g.ibound[pn] == 0 || g.ibound[pm] == 0) continue;
        const int sn = G.s[gi];
        if ((sn == 0 || sn == 5) && g.ivarcv == 0) continue;
        const double csat = sn < 3 ? g.condsat[dirs[sn] * N + pm] : g.condsat[dirs[sn] * N + pn];
        const double xn = x[pn], xm = x[pm];
        const int iups = xm > xn ? 1 : 0;
        const int up = iups ? pm : pn;
        if (g.icelltype[up] == 0) continue;
        const double topup = g.top[up], botup = g.bot[up], xup = x[up];
        for (int jj = 0; jj < G.numjs; jj++) {
          const int pj = G.jp[gi * G.numjs + jj];
          if (pj < 0) continue;
          if (g.ibound[pj] == 0) continue;
          const double alpha = G.a[gi * G.numjs + jj];
          if (alpha == 0.0) continue;
          const double consterm = csat * alpha * (xn - x[pj]);
          const double derv = sqsat_deriv(topup, botup, xup, 1.0e-6);
          const double term = consterm * derv;
          if (iups == 0) {
            if (role == 0) {
              diag = diag + term;
              b = b + term * xn;
            } else {
              if (g.ibound[pm] > 0) off[5 - sn] = off[5 - sn] - term;
              b = b - term * xn;
            }
          } else {
            if (role == 1) {
              diag = diag - term;
              b = b - term * xm;
            } else {
              if (g.ibound[pn] > 0) off[sn] = off[sn] + term;
              b = b + term * xm;
            }
          }
        }
      }
      // --- sto_fn
      if (g.has_sto && !iss && g.ibound[p] > 0 && g.iconvert[p] != 0) {
        double tled = 1.0 / delt;
        double tp = g.top[p], bt = g.bot[p];
        double tthk = tp - bt;
        double h = hcell;
        double snnew = sqsat(tp, bt, h, 1.0e-6);
        double rho1 = g.ss[p] * (g.istor_coef == 0 ? tp - bt : 1.0) * g.area[p] * tled;
        double rho2 = g.sy[p] * g.area[p] * tled;
        double derv = sqsat_deriv(tp, bt, h, 1.0e-6);
        double drterm;
        if (g.iconf_ss == 0) {   // gwf-sto.f90:414-423
          if (g.iorig_ss == 0) drterm = -rho1 * derv * (h - bt) + rho1 * tthk * snnew * derv;
          else drterm = -(rho1 * derv * h);
          diag = diag + drterm;
          b = b + drterm * h;
        }
        if (snnew < 1.0) {
          if (snnew > 0.0) {
            double rterm = -rho2 * tthk * snnew;
            drterm = -rho2 * tthk * derv;
            diag = diag + (drterm + rho2);
            b = b - rterm + drterm * h + rho2 * bt;
          }
        }
      }
      // --- bnd_fn in bndlist order (gwf.f90:545-553): wel_fn
      // (gwf-wel.f90:378-424) for AUTO_FLOW_REDUCE withdrawals in
      // convertible cells, drn_fn (gwf-drn.f90:418-470) for drains with a
      // depth
      if (g.ibound[p] > 0) {
        for (int e = bd.off[p]; e < bd.off[p + 1]; e++) {
          const int code = bd.ent[e];
          const int pk = code >> 28, idx = code & 0x0fffffff;
          const PkgDev &P = bd.pk[pk];
          if (pk == MF6HIP_PKG_WEL) {
            if (P.iflowred == 0 || g.icelltype[p] == 0) continue;
            const double q = -bd.brhs[pk][idx];
            if (q < 0.0) {
              const double bt = g.bot[p];
              const double tp = bt + P.flowred * (g.top[p] - bt);
              double drterm = sqsat_cubic_deriv(tp, bt, hcell, -2.0, 3.0);
              drterm = drterm * P.v1[idx];
              diag = diag + drterm;
              b = b + drterm * hcell;
            }
          } else if (pk == MF6HIP_PKG_DRN && P.idepth) {
            double drndepth, drntop, drnbot;
            drn_elevations(P, idx, drndepth, drntop, drnbot);
            if (drndepth != 0.0) {
              double drterm = sqsat_cubic_deriv(drntop, drnbot, hcell, -1.0, 2.0);
              drterm = drterm * P.v2[idx] * (drnbot - hcell);
              diag = diag + drterm;
              b = b + drterm * hcell;
            }
          }
        }
      }
    }
    A[p] = diag;
#pragma unroll
    for (int s = 0; s < 6; s++) A[(s + 1) * N + p] = off[s];
    rhs[p] = b;
  }
}

// sln_ls row fixes (NumericalSolution.f90:2434-2478) for one model
__global__ void __launch_bounds__(MF6_BLOCK) k_row_fixes(Layout L, const int *__restrict__ active,
                                                         int isymmetric, double *__restrict__ A,
                                                         double *__restrict__ rhs,
                                                         const double *__restrict__ x,
                                                         double *__restrict__ xtemp) {
  const int N = L.NP;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < L.NP; p += gridDim.x * blockDim.x) {
    int c, k, i, j, l;
    if (!pos_to_cell(L, p, c, k, i, j, l)) continue;
    double xn = x[p];
    xtemp[p] = xn;
    double b = rhs[p];
    if (active[p] > 0) {
      double adiag = fabs(A[p]);
      if (adiag < 1.0e-15) {
        A[p] = -1.0;
        b = b + (-1.0) * xn;
      }
      if (isymmetric) {
        int pres[6] = {k > 0, i > 0, j > 0, j < L.ncol - 1, i < L.nrow - 1, k < L.nlay - 1};
        const int ob = (1 - c) * L.H + k * L.HP;
        int npos[6] = {p - L.HP, ob + hn(l, -L.ncol), ob + hn(l, -1), ob + hn(l, 1),
                       ob + hn(l, L.ncol), p + L.HP};
        for (int s = 0; s < 6; s++) {
          if (!pres[s]) continue;
          int m = npos[s];
          if (active[m] < 0) {
            b = b - (A[(s + 1) * N + p] * x[m]);
            A[(s + 1) * N + p] = 0.0;
          }
        }
      }
    } else {
      A[p] = 1.0;
      for (int s = 1; s < 7; s++) A[s * N + p] = 0.0;
      b = xn;
    }
    rhs[p] = b;
  }
}

// gwf_ad: xold = x, CHD cells take their head and become ibound < 0
__global__ void k_advance(Layout L, const double *__restrict__ x, double *__restrict__ xold,
                          const int *__restrict__ ibound0, int *__restrict__ ibound) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < L.NP; p += gridDim.x * blockDim.x) {
    xold[p] = x[p];
    ibound[p] = ibound0[p];
  }
}

__global__ void k_chd_set(int nb, const int *__restrict__ nodes, const double *__restrict__ head,
                          Layout L, double *__restrict__ x, double *__restrict__ xold,
                          int *__restrict__ ibound) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nb; e += gridDim.x * blockDim.x) {
    int p = pos_of_natural(L, nodes[e]);
    if (ibound[p] != 0) {
      x[p] = head[e];
      xold[p] = head[e];
      ibound[p] = -(e + 1);
    }
  }
}

// sln_get_dxmax partials: signed largest |xa - xb| over active cells
__global__ void __launch_bounds__(MF6_BLOCK) k_dxmax(Layout L, const int *__restrict__ active,
                                                     const double *__restrict__ xa,
                                                     const double *__restrict__ xb,
                                                     MaxLoc *__restrict__ mpart) {
  __shared__ MaxLoc mlds[8];
  MaxLoc m{0.0, -1};
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < L.NP; p += gridDim.x * blockDim.x) {
    if (active[p] < 1) continue;
    int c, k, i, j, l;
    if (!pos_to_cell(L, p, c, k, i, j, l)) continue;
    double hd = xa[p] - xb[p];
    MaxLoc cand{hd, natural(L, k, i, j)};
    double fa = fabs(m.v), fb = fabs(hd);
    if (fb > fa || (fb == fa && fb > 0.0 && cand.loc < m.loc)) m = cand;
  }
  // block reduce
  for (int o = 32; o > 0; o >>= 1) {
    MaxLoc b2;
    b2.v = __shfl_xor(m.v, o);
    b2.loc = __shfl_xor(m.loc, o);
    m = maxloc_pick(m, b2);
  }
  int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) mlds[w] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    MaxLoc r{0.0, -1};
    for (int t = 0; t < (int)(blockDim.x >> 6); t++) r = maxloc_pick(r, mlds[t]);
    mpart[blockIdx.x] = r;
  }
}

// one block: the block partials' maximum.  maxloc_pick orders candidates by
// |v|, then by the smaller location, so the result does not depend on the
// reduction order (the reference's first-occurrence rule).
__global__ void __launch_bounds__(MF6_BLOCK) k_maxloc_final(const MaxLoc *mpart, int nb,
                                                            MaxLoc *out) {
  __shared__ MaxLoc mlds[MF6_BLOCK / 64];
  MaxLoc m{0.0, -1};
  for (int t = threadIdx.x; t < nb; t += blockDim.x) m = maxloc_pick(m, mpart[t]);
  for (int o = 32; o > 0; o >>= 1) {
    MaxLoc b2;
    b2.v = __shfl_xor(m.v, o);
    b2.loc = __shfl_xor(m.loc, o);
    m = maxloc_pick(m, b2);
  }
  if ((threadIdx.x & 63) == 0) mlds[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x != 0) return;
  MaxLoc r{0.0, -1};
  for (int w = 0; w < (int)(blockDim.x >> 6); w++) r = maxloc_pick(r, mlds[w]);
  *out = r;
}

// r = A x - b of the current system (amux slot order), inactive rows zero
// (sln_calc_residual): per-block partials of sum r^2, of max |r| / cell
// volume over ibound >= 1 (gwf_ptc) and of sum b^2 over active rows (the
// bnorm of DEV_PTC_OPTION).
__global__ void __launch_bounds__(MF6_BLOCK) k_resid_norms(
    Layout L, const double *__restrict__ A, const double *__restrict__ x,
    const double *__restrict__ b, const int *__restrict__ ibound,
    const double *__restrict__ area, const double *__restrict__ top,
    const double *__restrict__ bot, double *part, int nblk) {
  __shared__ double lds[3][MF6_BLOCK / 64];
  const int NP = L.NP;
  double ss = 0.0, bb = 0.0, mx = 0.0;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < NP; p += gridDim.x * blockDim.x) {
    int c, k, i, j, l;
    if (!pos_to_cell(L, p, c, k, i, j, l)) continue;
    if (ibound[p] < 1) continue;
    const int ob = (1 - c) * L.H + k * L.HP;
    double t = 0.0;
    t = t + A[p] * x[p];
    if (k > 0) t = t + A[NP + p] * x[p - L.HP];
    if (i > 0) t = t + A[2 * NP + p] * x[ob + hn(l, -L.ncol)];
    if (j > 0) t = t + A[3 * NP + p] * x[ob + hn(l, -1)];
    if (j < L.ncol - 1) t = t + A[4 * NP + p] * x[ob + hn(l, 1)];
    if (i < L.nrow - 1) t = t + A[5 * NP + p] * x[ob + hn(l, L.ncol)];
    if (k < L.nlay - 1) t = t + A[6 * NP + p] * x[p + L.HP];
    double r = t + (-1.0) * b[p];
    ss = ss + r * r;
    bb = bb + b[p] * b[p];
    double rv = fabs(r) / (area[p] * (top[p] - bot[p]));
    if (rv > mx) mx = rv;
  }
  for (int o = 32; o > 0; o >>= 1) {
    ss = ss + __shfl_xor(ss, o);
    bb = bb + __shfl_xor(bb, o);
    mx = fmax(mx, __shfl_xor(mx, o));
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { lds[0][w] = ss; lds[1][w] = bb; lds[2][w] = mx; }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  for (int v = 0; v < (int)(blockDim.x >> 6); v++) {
    a0 = a0 + lds[0][v];
    a1 = a1 + lds[1][v];
    a2 = fmax(a2, lds[2][v]);
  }
  part[blockIdx.x] = a0;
  part[nblk + blockIdx.x] = a1;
  part[2 * nblk + blockIdx.x] = a2;
}

__global__ void __launch_bounds__(MF6_BLOCK) k_resid_final(const double *part, int nblk,
                                                           double *out) {
  __shared__ double lds[3][MF6_BLOCK / 64];
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  for (int t = threadIdx.x; t < nblk; t += blockDim.x) {
    a0 = a0 + part[t];
    a1 = a1 + part[nblk + t];
    a2 = fmax(a2, part[2 * nblk + t]);
  }
  for (int o = 32; o > 0; o >>= 1) {
    a0 = a0 + __shfl_xor(a0, o);
    a1 = a1 + __shfl_xor(a1, o);
    a2 = fmax(a2, __shfl_xor(a2, o));
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { lds[0][w] = a0; lds[1][w] = a1; lds[2][w] = a2; }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int v = 0; v < (int)(blockDim.x >> 6); v++) {
    s0 = s0 + lds[0][v];
    s1 = s1 + lds[1][v];
    s2 = fmax(s2, lds[2][v]);
  }
  out[0] = sqrt(s0);
  out[1] = sqrt(s1);
  out[2] = s2;
}

// PTC terms of sln_ls (NumericalSolution.f90:2548-2558): diag -= ptcval,
// rhs -= ptcval x on active rows
__global__ void k_ptc_apply(Layout L, const int *__restrict__ ibound, double ptcval,
                            double *__restrict__ A, double *__restrict__ b,
                            const double *__restrict__ x) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < L.NP; p += gridDim.x * blockDim.x) {
    if (ibound[p] < 1) continue;
    A[p] = A[p] + (-ptcval);
    b[p] = b[p] - ptcval * x[p];
  }
}

// apply_backtracking (NumericalSolution.f90:2830-2843)
__global__ void k_bt_apply(Layout L, const int *__restrict__ ibound, double breduc,
                           double *__restrict__ x, const double *__restrict__ xtemp) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < L.NP; p += gridDim.x * blockDim.x) {
    if (ibound[p] < 1) continue;
    double delx = breduc * (x[p] - xtemp[p]);
    x[p] = xtemp[p] + delx;
  }
}

// sln_underrelax SIMPLE (gamma) and COOLEY (relax) (:3008-3061)
__global__ void k_ur_relax(Layout L, const int *__restrict__ ibound, double relax,
                           double *__restrict__ x, const double *__restrict__ xtemp,
                           double *__restrict__ dxold) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < L.NP; p += gridDim.x * blockDim.x) {
    if (ibound[p] < 1) continue;
    double delx = x[p] - xtemp[p];
    dxold[p] = delx;
    x[p] = xtemp[p] + relax * delx;
  }
}

// sln_underrelax delta-bar-delta (:3064-3111)
__global__ void k_ur_dbd(Layout L, const int *__restrict__ ibound, int kiter, double theta,
                         double akappa, double gamma, double amomentum,
                         double *__restrict__ x, const double *__restrict__ xtemp,
                         double *__restrict__ dxold, double *__restrict__ wsave,
                         double *__restrict__ hchold, double *__restrict__ deold) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < L.NP; p += gridDim.x * blockDim.x) {
    if (ibound[p] < 1) continue;
    double delx = x[p] - xtemp[p];
    if (kiter == 1) {
      wsave[p] = 1.0;
      hchold[p] = 1.0e-20;
      deold[p] = 0.0;
    }
    double ww = wsave[p];
    if (deold[p] * delx < 0.0) ww = theta * wsave[p];
    else ww = wsave[p] + akappa;
    if (ww > 1.0) ww = 1.0;
    wsave[p] = ww;
    if (kiter == 1) hchold[p] = delx;
    else hchold[p] = (1.0 - gamma) * delx + gamma * hchold[p];
    deold[p] = delx;
    dxold[p] = delx;
    double amom = 0.0;
    if (kiter > 4) amom = amomentum;
    delx = delx * ww + amom * hchold[p];
    x[p] = xtemp[p] + delx;
  }
}

// sln_calcdx and npf_nur for one model
__global__ void k_calcdx(Layout L, const int *__restrict__ active, const double *__restrict__ x,
                         const double *__restrict__ xtemp, double *__restrict__ dx) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < L.NP; p += gridDim.x * blockDim.x)
    dx[p] = active[p] < 1 ? 0.0 : x[p] - xtemp[p];
}

__global__ void k_nur(Layout L, const int *__restrict__ ibound, const int *__restrict__ icelltype,
                      const int *__restrict__ ibotnode, const double *__restrict__ bot,
                      double *__restrict__ x, const double *__restrict__ xtemp,
                      double *__restrict__ dx, int *flag) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < L.NP; p += gridDim.x * blockDim.x) {
    if (ibound[p] < 1) continue;
    if (icelltype[p] > 0) {
      double botm = bot[ibotnode[p]];
      if (x[p] < botm) {
        double xx = xtemp[p] * (1.0 - 0.9) + botm * 0.9;
        x[p] = xx;
        dx[p] = 0.0;
        atomicOr(flag, 1);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// budget: gwf_cq + gwf_bd (gwf.f90:741-800)
// ---------------------------------------------------------------------------

// bnd_cq_simrate (BoundaryPackage.f90:583-619) without mover: the rate into
// the model of every bound, hcof h - rhs for active cells
__global__ void k_bnd_rate(int nb, const int *__restrict__ nodes, Layout L,
                           const int *__restrict__ ibound, const double *__restrict__ x,
                           const double *__restrict__ hcof, const double *__restrict__ brhs,
                           double *__restrict__ rate) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nb; e += gridDim.x * blockDim.x) {
    const int p = pos_of_natural(L, nodes[e]);
    rate[e] = ibound[p] > 0 ? hcof[e] * x[p] - brhs[e] : 0.0;
  }
}

struct BndRates {
  const int *off, *ent;
  const double *rate[5];
};

// One thread per cell: the face flows of its row in CSR order (npf_cq /
// qcalc: qnm = cond (hm - hn) computed for the lower node first, the upper
// node's entry its negative), then the diagonal in the reference's order of
// accumulation: storage (sto_cq: SsTerms then SyTerms rates), the package
// rates in bndlist order (bnd_cq), the face flows in CSR order
// (csr_diagsum) and, for a CHD cell, its CHD rate (calc_chd_rate: minus the
// inflow from its neighbours).  strg receives the STO-SS / STO-SY rates.
__global__ void __launch_bounds__(MF6_BLOCK) k_budget_cells(
    GwfDev g, BndRates br, GncDev G, HfbDev F, const int *__restrict__ ia,
    const int *__restrict__ chd_cell,
    int iss, double delt, const double *__restrict__ x, const double *__restrict__ xold,
    double *__restrict__ flowja, double *__restrict__ strg) {
  const Layout &L = g.L;
  const int N = L.NP;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < L.NP; p += gridDim.x * blockDim.x) {
    int c, k, i, j, l;
    if (!pos_to_cell(L, p, c, k, i, j, l)) continue;
    const int n = natural(L, k, i, j);
    const double hcell = x[p];
    int npos[6], pres[6];
    pres[0] = k > 0;
    pres[1] = i > 0;
    pres[2] = j > 0;
    pres[3] = j < L.ncol - 1;
    pres[4] = i < L.nrow - 1;
    pres[5] = k < L.nlay - 1;
    npos[0] = pres[0] ? p - L.HP : -1;
    npos[5] = pres[5] ? p + L.HP : -1;
    {
      const int ob = (1 - c) * L.H + k * L.HP;
      npos[1] = pres[1] ? ob + hn(l, -L.ncol) : -1;
      npos[2] = pres[2] ? ob + hn(l, -1) : -1;
      npos[3] = pres[3] ? ob + hn(l, 1) : -1;
      npos[4] = pres[4] ? ob + hn(l, L.ncol) : -1;
    }
    const int dirs[6] = {2, 1, 0, 0, 1, 2};
    double q[6];
    int ipos = ia[n] + 1;
#pragma unroll
    for (int s = 0; s < 6; s++) {
      q[s] = 0.0;
      if (!pres[s]) continue;
      const int m = npos[s];
      const int dir = dirs[s];
      if (s < 3) {   // m < n: qcalc(m, n) and this entry is its negative
        const int im = i - (s == 1), jm = j - (s == 2);
        const double cond = pair_cond_sat(g, dir, m, p, im, jm, x[m], hcell, g.sat[m],
                                          g.sat[p], g.condsat[dir * N + m]);
        q[s] = -(cond * (hcell - x[m]));
      } else {
        const double cond = pair_cond_sat(g, dir, p, m, i, j, hcell, x[m], g.sat[p],
                                          g.sat[m], g.condsat[dir * N + p]);
        q[s] = cond * (x[m] - hcell);
      }
    }
    // --- hfb_cq without NEWTON (gwf-hfb.f90:355-450): across a barrier next
    // to a convertible cell the flow is hfb_fc's condsav times the head
    // difference (set, in barrier order)
    for (int e = F.off ? F.off[p] : 0, e1 = F.off ? F.off[p + 1] : 0; e < e1; e++) {
      const int code = F.ent[e];
      const int bi = code >> 1, role = code & 1;
      const int pn = F.np[bi], pm = F.mp[bi];
      if (g.ibound[pn] == 0 || g.ibound[pm] == 0) continue;
      if (g.icelltype[pn] != 1 && g.icelltype[pm] != 1) continue;
      const double qnm = F.sav[bi] * (x[pm] - x[pn]);
      if (role == 0) q[F.s[bi]] = qnm;
      else q[5 - F.s[bi]] = -qnm;
    }
    // --- gnc_cq (GhostNode.f90:478-503, deltaQgnc :507-543): the
    // correction flow with the conductance saved at the last formulate,
    // added at (n, m) and taken from (m, n), in GNC order
    for (int e = G.off ? G.off[p] : 0, e1 = G.off ? G.off[p + 1] : 0; e < e1; e++) {
      const int code = G.ent[e];
      const int gi = code >> 1, role = code & 1;
      const int pn = G.np[gi], pm = G.mp[gi];
      double dq = 0.0;
      if (g.ibound[pn] != 0 && g.ibound[pm] != 0) {
        double sigalj = 0.0, hd = 0.0;
        for (int jj = 0; jj < G.numjs; jj++) {
          const int pj = G.jp[gi * G.numjs + jj];
          if (pj < 0) continue;
          if (g.ibound[pj] == 0) continue;
          const double alpha = G.a[gi * G.numjs + jj];
          sigalj = sigalj + alpha;
          hd = hd + alpha * x[pj];
        }
        const double aterm = sigalj * x[pn] - hd;
        dq = aterm * G.cond[gi];
      }
      if (role == 0) q[G.s[gi]] = q[G.s[gi]] + dq;
      else q[5 - G.s[gi]] = q[5 - G.s[gi]] - dq;
    }
#pragma unroll
    for (int s = 0; s < 6; s++)
      if (pres[s]) flowja[ipos++] = q[s];
    double diag = 0.0;
    // --- sto_cq
    double rss = 0.0, rsy = 0.0;
    if (g.has_sto && !iss && g.ibound[p] > 0) {
      const double tled = 1.0 / delt;
      const double tp = g.top[p], bt = g.bot[p];
      double snold, snnew;
      if (g.iconvert[p] == 0) {
        snold = 1.0;
        snnew = 1.0;
      } else {
        snold = sqsat(tp, bt, xold[p], g.satomega);
        snnew = sqsat(tp, bt, hcell, g.satomega);
      }
      const double sc1 = g.ss[p] * (g.istor_coef == 0 ? tp - bt : 1.0) * g.area[p];
      const double rho1 = sc1 * tled;
      const double rho1old = rho1;
      double aterm, rhsterm;
      ss_terms(g.iconvert[p], g.iorig_ss, g.iconf_ss, tp, bt, rho1, rho1old, snnew, snold,
               xold[p], aterm, rhsterm);
      rss = aterm * hcell - rhsterm;
      diag = diag + rss;
      if (g.iconvert[p] != 0) {
        const double rho2 = g.sy[p] * g.area[p] * tled;
        const double rho2old = rho2;
        const double tthk = tp - bt;
        rsy = rho2old * tthk * snold - rho2 * tthk * snnew;
      }
      diag = diag + rsy;
    }
    strg[p] = rss;
    strg[N + p] = rsy;
    // --- bnd_cq (WEL, DRN, RCH in bndlist order)
    for (int e = br.off[p]; e < br.off[p + 1]; e++) {
      const int code = br.ent[e];
      diag = diag + br.rate[code >> 28][code & 0x0fffffff];
    }
    // --- csr_diagsum
#pragma unroll
    for (int s = 0; s < 6; s++)
      if (pres[s]) diag = diag + q[s];
    // --- calc_chd_rate
    if (chd_cell[p]) {
      double rate = 0.0;
#pragma unroll
      for (int s = 0; s < 6; s++)
        if (pres[s]) rate = rate - q[s];
      diag = diag + rate;
    }
    flowja[ia[n]] = diag;
  }
}

// calc_chd_rate per CHD bound: the rate (minus the flows of the cell's row,
// every neighbour) and its in / out split (active neighbours only); the
// row's flows in CSR order: up, back, left, right, front, down
__global__ void k_chd_inout(int nb, const int *__restrict__ nodes, const int *__restrict__ ia,
                            Layout L, const int *__restrict__ ibound,
                            const double *__restrict__ flowja, double *__restrict__ rin,
                            double *__restrict__ rout, double *__restrict__ rsim) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nb; e += gridDim.x * blockDim.x) {
    const int n = nodes[e];
    const int k = n / L.NRC, r = n - k * L.NRC, i = r / L.ncol, j = r - i * L.ncol;
    const int kk[6] = {k - 1, k, k, k, k, k + 1};
    const int ii[6] = {i, i - 1, i, i, i + 1, i};
    const int jj[6] = {j, j, j - 1, j + 1, j, j};
    double a = 0.0, b = 0.0, rate = 0.0;
    int ip = ia[n] + 1;
    for (int s = 0; s < 6; s++) {
      if (kk[s] < 0 || kk[s] >= L.nlay || ii[s] < 0 || ii[s] >= L.nrow || jj[s] < 0 ||
          jj[s] >= L.ncol)
        continue;
      const double qv = flowja[ip++];
      rate = rate - qv;
      if (ibound[cell_pos(L, kk[s], ii[s], jj[s])] > 0) {
        if (qv < 0.0) a = a - qv;
        else b = b + qv;
      }
    }
    rin[e] = a;
    rout[e] = b;
    rsim[e] = rate;
  }
}

// flag[pos(node)] = 1 for the listed nodes
__global__ void k_mark_cells(int nb, const int *__restrict__ nodes, Layout L,
                             int *__restrict__ flag) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < nb; e += gridDim.x * blockDim.x)
    flag[pos_of_natural(L, nodes[e])] = 1;
}

// rate_accumulator (Budget.f90:631-648) over n values with a fixed-shape
// reduction: part[2 b] = sum of positives, part[2 b + 1] = - sum of negatives
__global__ void __launch_bounds__(MF6_BLOCK) k_rate_acc(int n, const double *__restrict__ f,
                                                        double *part) {
  __shared__ double lds[2][MF6_BLOCK / 64];
  double a = 0.0, b = 0.0;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    const double v = f[e];
    if (v < 0.0) b = b - v;
    else a = a + v;
  }
  for (int o = 32; o > 0; o >>= 1) {
    a = a + __shfl_xor(a, o);
    b = b + __shfl_xor(b, o);
  }
  if ((threadIdx.x & 63) == 0) { lds[0][threadIdx.x >> 6] = a; lds[1][threadIdx.x >> 6] = b; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double sa = 0.0, sb = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) { sa = sa + lds[0][w]; sb = sb + lds[1][w]; }
    part[2 * blockIdx.x] = sa;
    part[2 * blockIdx.x + 1] = sb;
  }
}

__global__ void k_rate_final(const double *part, int nb, double *out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double a = 0.0, b = 0.0;
  for (int i = 0; i < nb; i++) { a = a + part[2 * i]; b = b + part[2 * i + 1]; }
  out[0] = a;
  out[1] = b;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

static GwfDev make_dev(Mf6Handle *h) {
  GwfDev g;
  g.L = h->L;
  g.top = h->top; g.bot = h->bot; g.area = h->area; g.k11 = h->k11; g.k22 = h->k22;
  g.k33 = h->k33;
  g.ss = h->ss; g.sy = h->sy; g.delr = h->delr; g.delc = h->delc; g.condsat = h->condsat;
  g.sat = h->sat;
  g.sat0 = h->sat0; g.hn0 = h->hn0;
  g.icelltype = h->icelltype; g.iconvert = h->iconvert; g.ibound = h->ibound;
  g.inewton = h->m.inewton; g.ivarcv = h->m.ivarcv; g.idewatcv = h->m.idewatcv;
  g.icellavg = h->m.icellavg; g.has_sto = h->m.has_sto;
  g.istor_coef = h->m.istor_coef; g.iconf_ss = h->m.iconf_ss; g.iorig_ss = h->m.iorig_ss;
  g.satomega = h->m.inewton ? 1.0e-6 : 0.0;   // NPF gwf-npf.f90:1394-1396, STO gwf-sto.f90:826-829
  g.fn = h->m.inewton;
  return g;
}

template <typename T>
static int upload_pos(Mf6Handle *h, T **dst, const T *src) {
  const Layout &L = h->L;
  CK(galloc(dst, (size_t)L.NP));
  std::vector<T> tmp(L.NP, (T)0);
  for (int n = 0; n < L.N; n++) tmp[pos_of_natural(L, n)] = src[n];
  HIPCHK(hipMemcpy(*dst, tmp.data(), sizeof(T) * L.NP, hipMemcpyHostToDevice));
  return MF6HIP_OK;
}

// sin and cos of one angle from one libm sincos call, as the reference's
// compiled hyeff gets them (its compiler merges each sin / cos pair; sincos
// differs from separate sin and cos calls in the last bit for some
// arguments)
static void host_sincos(double a, double *s, double *c) { ::sincos(a, s, c); }

// HGeoUtil.f90:29-108 hyeff (harmonic, iavgmeth 0): the conductivity along
// the unit vector vg of the tensor k11, k22, k33 rotated by ang1..3
// (radians); ve**2 of the reference compiles to ve * ve
static double hyeff_host(double k11, double k22, double k33, double ang1, double ang2,
                         double ang3, double vg1, double vg2, double vg3) {
  double s1, c1, s2, c2, s3, c3;
  host_sincos(ang1, &s1, &c1);
  host_sincos(ang2, &s2, &c2);
  host_sincos(ang3, &s3, &c3);
  const double r11 = c1 * c2, r12 = c1 * s2 * s3 - s1 * c3, r13 = -c1 * s2 * c3 - s1 * s3;
  const double r21 = s1 * c2, r22 = s1 * s2 * s3 + c1 * c3, r23 = -s1 * s2 * c3 + c1 * s3;
  const double r31 = s2, r32 = -c2 * s3, r33 = c2 * c3;
  const double ve1 = r11 * vg1 + r21 * vg2 + r31 * vg3;
  const double ve2 = r12 * vg1 + r22 * vg2 + r32 * vg3;
  const double ve3 = r13 * vg1 + r23 * vg2 + r33 * vg3;
  double dnum = 1.0, d1 = ve1 * ve1, d2 = ve2 * ve2, d3 = ve3 * ve3;
  if (ve1 != 0.0) { dnum = dnum * k11; d2 = d2 * k11; d3 = d3 * k11; }
  if (ve2 != 0.0) { dnum = dnum * k22; d1 = d1 * k22; d3 = d3 * k22; }
  if (ve3 != 0.0) { dnum = dnum * k33; d1 = d1 * k33; d2 = d2 * k33; }
  const double denom = d1 + d2 + d3;
  return denom > 0.0 ? dnum / denom : 0.0;
}

// The conductivities of column-direction (kx), row-direction (ky) and
// vertical (kz) faces, hy_eff (gwf-npf.f90:2280-2355) with the DIS face
// normals along the axes (hyeff depends on them only through squares and
// zero tests); angles converted as prepcheck does (:1686-1731, DPIO180 =
// atan(1) / 45, Constants.f90:130).
static int directional_k(const mf6hip_gwf_model *m, int n, std::vector<double> &kx,
                         std::vector<double> &ky, std::vector<double> &kz) {
  const bool i1 = m->angle1, i2 = m->angle2, i3 = m->angle3;
  if ((i2 && !i1) || (i3 && !i1) || (i2 != i3)) {
    mf6hip_set_error("NPF angles: ANGLE2 and ANGLE3 need ANGLE1 and each other");
    return MF6HIP_EINVAL;
  }
  const double dpio180 = std::atan(1.0) / 45.0;
  kx.assign(m->k11, m->k11 + n);
  ky.assign(m->k11, m->k11 + n);
  kz.assign(m->k33, m->k33 + n);
  const double *k22 = m->k22 ? m->k22 : m->k11;
  for (int c = 0; c < n; c++) {
    const double a1 = i1 ? m->angle1[c] * dpio180 : 0.0;
    const double a2 = i2 ? m->angle2[c] * dpio180 : 0.0;
    const double a3 = i3 ? m->angle3[c] * dpio180 : 0.0;
    if (m->k22) {
      kx[c] = hyeff_host(m->k11[c], k22[c], m->k33[c], a1, a2, a3, 1.0, 0.0, 0.0);
      ky[c] = hyeff_host(m->k11[c], k22[c], m->k33[c], a1, a2, a3, 0.0, 1.0, 0.0);
    }
    if (i2) kz[c] = hyeff_host(m->k11[c], k22[c], m->k33[c], a1, a2, a3, 0.0, 0.0, 1.0);
  }
  return MF6HIP_OK;
}

extern "C" int mf6hip_gwf_create(void *handle, const mf6hip_gwf_model *m) {
  Mf6Handle *h = (Mf6Handle *)handle;
  if (!h || !m) { mf6hip_set_error("null argument"); return MF6HIP_EINVAL; }
  const Layout &L = h->L;
  if (m->nlay != L.nlay || m->nrow != L.nrow || m->ncol != L.ncol) {
    mf6hip_set_error("model shape differs from the solver's DIS shape");
    return MF6HIP_EINVAL;
  }
  if (!m->delr || !m->delc || !m->top || !m->bot || !m->k11 || !m->k33 || !m->icelltype ||
      !m->ibound) {
    mf6hip_set_error("model array missing");
    return MF6HIP_EINVAL;
  }
  if (m->has_sto && (!m->iconvert || !m->ss || !m->sy)) {
    mf6hip_set_error("STO arrays missing");
    return MF6HIP_EINVAL;
  }
  if (m->icellavg < 0 || m->icellavg > 3) {
    mf6hip_set_error("icellavg must be 0 (harmonic), 1 (LOGARITHMIC), 2 (AMT-LMK) or 3 "
                     "(AMT-HMK)");
    return MF6HIP_EINVAL;
  }
  if (m->ithickstrt && !m->strt) {
    mf6hip_set_error("THICKSTRT needs the starting heads (strt)");
    return MF6HIP_EINVAL;
  }
  h->m = *m;
  CK(upload_pos(h, &h->top, m->top));
  CK(upload_pos(h, &h->bot, m->bot));
  {
    std::vector<double> kx, ky, kz;
    CK(directional_k(m, L.N, kx, ky, kz));
    CK(upload_pos(h, &h->k11, kx.data()));
    CK(upload_pos(h, &h->k22, ky.data()));
    CK(upload_pos(h, &h->k33, kz.data()));
  }
  {
    // THICKSTRT (gwf-npf.f90:1836-1886): cells entered with a negative
    // ICELLTYPE become confined, with the saturated thickness of their
    // starting head in calc_condsat; without it a negative ICELLTYPE means
    // convertible.  Every cell is active here (the model's ibound is 1 or 0
    // only through IDOMAIN, and THICKSTRT cells with IDOMAIN 0 keep sat 1).
    std::vector<int> ict(m->icelltype, m->icelltype + L.N);
    std::vector<double> sat0(L.N, 1.0), hn0(m->top, m->top + L.N);
    for (int n = 0; n < L.N; n++) {
      if (ict[n] >= 0) continue;
      if (!m->ithickstrt) { ict[n] = 1; continue; }
      ict[n] = 0;
      hn0[n] = m->strt[n];
      if (m->ibound[n] == 0) continue;
      const double tp = m->top[n], bt = m->bot[n], hs = m->strt[n];
      if (m->inewton) sat0[n] = sqsat(tp, bt, hs, 1.0e-6);   // thksat :775-794
      else sat0[n] = hs >= tp ? 1.0 : (hs - bt) / (tp - bt);
    }
    h->h_icelltype = ict;
    CK(upload_pos(h, &h->icelltype, ict.data()));
    CK(upload_pos(h, &h->sat0, sat0.data()));
    CK(upload_pos(h, &h->hn0, hn0.data()));
  }
  CK(upload_pos(h, &h->ibound0, m->ibound));
  {
    std::vector<double> area(L.N);
    for (int n = 0; n < L.N; n++) {
      int r = n % L.NRC, i = r / L.ncol, j = r - i * L.ncol;
      area[n] = m->delr[j] * m->delc[i];
    }
    CK(upload_pos(h, &h->area, area.data()));
  }
  if (m->has_sto) {
    CK(upload_pos(h, &h->iconvert, m->iconvert));
    CK(upload_pos(h, &h->ss, m->ss));
    CK(upload_pos(h, &h->sy, m->sy));
  }
  CK(galloc(&h->delr, (size_t)L.ncol));
  CK(galloc(&h->delc, (size_t)L.nrow));
  HIPCHK(hipMemcpy(h->delr, m->delr, sizeof(double) * L.ncol, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(h->delc, m->delc, sizeof(double) * L.nrow, hipMemcpyHostToDevice));
  h->h_delr.assign(m->delr, m->delr + L.ncol);
  h->h_delc.assign(m->delc, m->delc + L.nrow);
  CK(galloc(&h->ibound, (size_t)L.NP));
  HIPCHK(hipMemcpy(h->ibound, h->ibound0, sizeof(int) * L.NP, hipMemcpyDeviceToDevice));
  CK(galloc(&h->condsat, 3 * (size_t)L.NP));
  CK(galloc(&h->sat, (size_t)L.NP));
  {
    // preprocess (gwf-npf.f90:1850): sat = 1 until the first npf_cf
    std::vector<double> one(L.NP, 1.0);
    HIPCHK(hipMemcpy(h->sat, one.data(), sizeof(double) * L.NP, hipMemcpyHostToDevice));
  }
  CK(galloc(&h->xold, (size_t)L.NP));
  CK(galloc(&h->dxold, (size_t)L.NP));
  HIPCHK(hipMemset(h->condsat, 0, sizeof(double) * 3 * L.NP));
  HIPCHK(hipMemset(h->xold, 0, sizeof(double) * L.NP));
  HIPCHK(hipMemset(h->dxold, 0, sizeof(double) * L.NP));
  // ibotnode (gwf-npf.f90:1905-1939): walk down while the bottom decreases
  {
    std::vector<int> ib(L.N);
    for (int n = 0; n < L.N; n++) {
      int cur = n;
      while (cur + L.NRC < L.N && m->bot[cur + L.NRC] < m->bot[cur]) cur += L.NRC;
      ib[n] = pos_of_natural(L, cur);
    }
    CK(upload_pos(h, &h->ibotnode, ib.data()));
  }
  h->has_model = true;
  GwfDev g = make_dev(h);
  k_condsat<<<ggrid(L.N), MF6_BLOCK, 0, h->stream>>>(g, h->condsat);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(h->stream));
  return MF6HIP_OK;
}

// A new set of horizontal flow barriers (hfb_rp's condsat_reset, read_data
// and condsat_modify, gwf-hfb.f90:149-201): noden / nodem are 1-based node
// numbers, hydchr the barriers' hydraulic characteristics.  nhfb = 0 removes
// the barriers.  The checks of check_data (gwf-hfb.f90:711-766) apply: the
// cells must be connected, horizontally.  Without NEWTON a barrier next to a
// convertible cell is applied by hfb_fc at every formulate instead
// (gwf-hfb.f90:211-348), which the device path does not model: rejected.
static void hfbd_free(Mf6Handle *h) {
  void *bufs[] = {h->d_hfbd_np, h->d_hfbd_mp, h->d_hfbd_s, h->d_hfbd_hy, h->d_hfbd_w,
                  h->d_hfbd_sav, h->d_hfbd_off, h->d_hfbd_ent};
  for (void *b : bufs)
    if (b) hipFree(b);
  h->d_hfbd_np = h->d_hfbd_mp = h->d_hfbd_s = h->d_hfbd_off = h->d_hfbd_ent = nullptr;
  h->d_hfbd_hy = h->d_hfbd_w = h->d_hfbd_sav = nullptr;
  h->hfbd_n = 0;
}

// The barriers (list positions dn of the new set) that hfb_fc / hfb_cq
// apply: per barrier the positions of n and m, the direction index of m
// seen from n, hydchr and the face width; per position the barriers of its
// row (b * 2 + role, role 1: the row is m) in barrier order.
static int hfb_dynamic(Mf6Handle *h, int nhfb, const std::vector<int> &nn,
                       const std::vector<int> &mm, const double *hydchr,
                       const std::vector<int> &dn) {
  (void)nhfb;
  const Layout &L = h->L;
  HIPCHK(hipStreamSynchronize(h->stream));
  hfbd_free(h);
  const int nd = (int)dn.size();
  if (nd == 0) return MF6HIP_OK;
  std::vector<int> np(nd), mp(nd), sd(nd), cnt(L.NP + 1, 0);
  std::vector<double> hy(nd), w(nd);
  for (int q = 0; q < nd; q++) {
    const int n = nn[dn[q]], m = mm[dn[q]];
    const int in = (n % L.NRC) / L.ncol, jn = (n % L.NRC) % L.ncol;
    const int im = (m % L.NRC) / L.ncol, jm = (m % L.NRC) % L.ncol;
    int dir;   // back, left, right, front
    if (im == in - 1) dir = 1;
    else if (jm == jn - 1) dir = 2;
    else if (jm == jn + 1) dir = 3;
    else dir = 4;
    np[q] = pos_of_natural(L, n);
    mp[q] = pos_of_natural(L, m);
    sd[q] = dir;
    hy[q] = hydchr[dn[q]];
    w[q] = (dir == 2 || dir == 3) ? h->h_delc[in] : h->h_delr[jn];   // the face width
    cnt[np[q] + 1]++;
    cnt[mp[q] + 1]++;
  }
  for (int p = 0; p < L.NP; p++) cnt[p + 1] += cnt[p];
  std::vector<int> ent(2 * (size_t)nd);
  {
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for (int q = 0; q < nd; q++) {
      ent[fill[np[q]]++] = q * 2;
      ent[fill[mp[q]]++] = q * 2 + 1;
    }
  }
  CK(galloc(&h->d_hfbd_np, (size_t)nd));
  CK(galloc(&h->d_hfbd_mp, (size_t)nd));
  CK(galloc(&h->d_hfbd_s, (size_t)nd));
  CK(galloc(&h->d_hfbd_hy, (size_t)nd));
  CK(galloc(&h->d_hfbd_w, (size_t)nd));
  CK(galloc(&h->d_hfbd_sav, (size_t)nd));
  CK(galloc(&h->d_hfbd_off, (size_t)L.NP + 1));
  CK(galloc(&h->d_hfbd_ent, 2 * (size_t)nd));
  HIPCHK(hipMemcpy(h->d_hfbd_np, np.data(), sizeof(int) * nd, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(h->d_hfbd_mp, mp.data(), sizeof(int) * nd, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(h->d_hfbd_s, sd.data(), sizeof(int) * nd, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(h->d_hfbd_hy, hy.data(), sizeof(double) * nd, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(h->d_hfbd_w, w.data(), sizeof(double) * nd, hipMemcpyHostToDevice));
  HIPCHK(hipMemset(h->d_hfbd_sav, 0, sizeof(double) * nd));
  HIPCHK(hipMemcpy(h->d_hfbd_off, cnt.data(), sizeof(int) * (L.NP + 1), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(h->d_hfbd_ent, ent.data(), sizeof(int) * 2 * nd, hipMemcpyHostToDevice));
  h->hfbd_n = nd;
  return MF6HIP_OK;
}

extern "C" int mf6hip_gwf_set_hfb(void *handle, int nhfb, const int *noden, const int *nodem,
                                  const double *hydchr) {
  Mf6Handle *h = (Mf6Handle *)handle;
  if (!h || !h->has_model) { mf6hip_set_error("no GWF model on this handle"); return MF6HIP_EINVAL; }
  if (nhfb < 0 || (nhfb > 0 && (!noden || !nodem || !hydchr))) {
    mf6hip_set_error("null argument or negative nhfb");
    return MF6HIP_EINVAL;
  }
  const Layout &L = h->L;
  std::vector<int> nn(nhfb), mm(nhfb);
  for (int b = 0; b < nhfb; b++) {
    const int n = noden[b] - 1, m = nodem[b] - 1;
    if (n < 0 || n >= L.N || m < 0 || m >= L.N) {
      mf6hip_set_error("HFB no. " + std::to_string(b + 1) + ": node number out of range");
      return MF6HIP_EINVAL;
    }
    const int kn = n / L.NRC, km = m / L.NRC;
    const int rn = n % L.NRC, rm = m % L.NRC;
    const int in = rn / L.ncol, jn = rn % L.ncol, im = rm / L.ncol, jm = rm % L.ncol;
    const bool horiz = kn == km && ((in == im && (jn - jm == 1 || jm - jn == 1)) ||
                                    (jn == jm && (in - im == 1 || im - in == 1)));
    const bool vert = rn == rm && (kn - km == 1 || km - kn == 1);
    if (!horiz) {
      mf6hip_set_error("HFB no. " + std::to_string(b + 1) +
                       (vert ? " is between two cells not horizontally connected: "
                             : " is between two unconnected cells: ") +
                       std::to_string(n + 1) + " and " + std::to_string(m + 1));
      return MF6HIP_EINVAL;
    }
    nn[b] = n;
    mm[b] = m;
  }
  // Without NEWTON a barrier next to a convertible cell (ICELLTYPE 1) is
  // applied by hfb_fc at every formulate and hfb_cq in the budget
  // (gwf-hfb.f90:211-450); the others are folded into condsat as
  // condsat_modify does (:789-832).  Faces hold barriers of one kind only.
  std::vector<int> fn, fm, dn;
  std::vector<double> fh;
  for (int b = 0; b < nhfb; b++) {
    const bool dyn = !h->m.inewton && (h->h_icelltype[nn[b]] == 1 || h->h_icelltype[mm[b]] == 1);
    if (dyn) {
      dn.push_back(b);
    } else {
      fn.push_back(nn[b]);
      fm.push_back(mm[b]);
      fh.push_back(hydchr[b]);
    }
  }
  CK(hfb_dynamic(h, nhfb, nn, mm, hydchr, dn));
  nhfb = (int)fn.size();
  nn = fn;
  mm = fm;
  hydchr = fh.data();
  hipStream_t s = h->stream;
  if (h->hfb_n > 0) k_hfb_reset<<<1, 64, 0, s>>>(h->hfb_n, h->d_hfb_face, h->d_hfb_sav, h->condsat);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s));
  h->hfb_n = 0;
  if (h->d_hfb_face) { hipFree(h->d_hfb_face); h->d_hfb_face = nullptr; }
  if (h->d_hfb_sav) { hipFree(h->d_hfb_sav); h->d_hfb_sav = nullptr; }
  if (nhfb == 0) return MF6HIP_OK;
  int *dnn = nullptr, *dmm = nullptr;
  double *dhy = nullptr;
  CK(galloc(&h->d_hfb_face, (size_t)nhfb));
  CK(galloc(&h->d_hfb_sav, (size_t)nhfb));
  if (hipMalloc((void **)&dnn, sizeof(int) * nhfb) != hipSuccess ||
      hipMalloc((void **)&dmm, sizeof(int) * nhfb) != hipSuccess ||
      hipMalloc((void **)&dhy, sizeof(double) * nhfb) != hipSuccess) {
    if (dnn) hipFree(dnn);
    if (dmm) hipFree(dmm);
    if (dhy) hipFree(dhy);
    mf6hip_set_error("hipMalloc failed (HFB lists)");
    return MF6HIP_EHIP;
  }
  hipError_t e = hipMemcpy(dnn, nn.data(), sizeof(int) * nhfb, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dmm, mm.data(), sizeof(int) * nhfb, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dhy, hydchr, sizeof(double) * nhfb, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    k_hfb_apply<<<1, 64, 0, s>>>(make_dev(h), nhfb, dnn, dmm, dhy, h->condsat, h->d_hfb_face,
                                 h->d_hfb_sav);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
  }
  hipFree(dnn);
  hipFree(dmm);
  hipFree(dhy);
  if (e != hipSuccess) {
    mf6hip_set_error(std::string("HFB: ") + hipGetErrorString(e));
    return MF6HIP_EHIP;
  }
  h->hfb_n = nhfb;
  return MF6HIP_OK;
}

// Package options that change the formulation (wel_options,
// gwf-wel.f90:185-230; drn_options, gwf-drn.f90:187-241): validated before
// the package data are replaced, stored with them after.
extern "C" int mf6hip_gwf_set_package_ex(void *handle, int pkgtype, int nbound,
                                         const int *nodelist, const double *v1,
                                         const double *v2,
                                         const mf6hip_package_options *opt) {
  if (opt) {
    if (opt->has_auxmult != 0) {
      mf6hip_set_error("AUXMULTNAME: multiply the package values on the host before "
                       "mf6hip_gwf_set_package_ex");
      return MF6HIP_EUNSUPPORTED;
    }
    if (opt->iflowred != 0 && pkgtype != MF6HIP_PKG_WEL) {
      mf6hip_set_error("AUTO_FLOW_REDUCE is a WEL option");
      return MF6HIP_EINVAL;
    }
    if ((opt->idrndepth != 0 || opt->icubic_scaling != 0) && pkgtype != MF6HIP_PKG_DRN) {
      mf6hip_set_error("AUXDEPTHNAME / DEV_CUBIC_SCALING are DRN options");
      return MF6HIP_EINVAL;
    }
    if (opt->idrndepth != 0 && nbound > 0 && !opt->drndepth) {
      mf6hip_set_error("DRN AUXDEPTHNAME needs the depths (drndepth)");
      return MF6HIP_EINVAL;
    }
    if (opt->iflowred != 0 && !(opt->flowred == opt->flowred)) {
      mf6hip_set_error("AUTO_FLOW_REDUCE value is not a number");
      return MF6HIP_EINVAL;
    }
  }
  CK(mf6hip_gwf_set_package(handle, pkgtype, nbound, nodelist, v1, v2));
  if (!opt) return MF6HIP_OK;
  Mf6Handle *h = (Mf6Handle *)handle;
  DevPackage &pk = h->pk[pkgtype];
  if (opt->iflowred != 0) {
    // wel_options: a value <= 0 becomes 0.1, one above 1 becomes 1
    double f = opt->flowred;
    pk.iflowred = 1;
    pk.flowred = f <= 0.0 ? 0.1 : (f > 1.0 ? 1.0 : f);
  }
  if (opt->idrndepth != 0) {
    pk.idepth = 1;
    if (nbound > 0) {
      CK(galloc(&pk.depth, (size_t)nbound));
      HIPCHK(hipMemcpy(pk.depth, opt->drndepth, sizeof(double) * nbound,
                       hipMemcpyHostToDevice));
    }
  }
  pk.icubic = opt->icubic_scaling != 0 ? 1 : 0;
  return MF6HIP_OK;
}

extern "C" int mf6hip_gwf_set_package(void *handle, int pkgtype, int nbound,
                                      const int *nodelist, const double *v1,
                                      const double *v2) {
  Mf6Handle *h = (Mf6Handle *)handle;
  if (!h || !h->has_model) { mf6hip_set_error("no model on this handle"); return MF6HIP_EINVAL; }
  if (pkgtype < MF6HIP_PKG_WEL || pkgtype > MF6HIP_PKG_CHD) { mf6hip_set_error("unknown package type"); return MF6HIP_EINVAL; }
  if (nbound < 0 || (nbound > 0 && (!nodelist || !v1))) { mf6hip_set_error("bad package arrays"); return MF6HIP_EINVAL; }
  if (pkgtype == MF6HIP_PKG_DRN && nbound > 0 && !v2) { mf6hip_set_error("DRN needs conductance (v2)"); return MF6HIP_EINVAL; }
  if (nbound >= (1 << 28)) { mf6hip_set_error("too many bounds"); return MF6HIP_EUNSUPPORTED; }
  DevPackage &pk = h->pk[pkgtype];
  void *old[] = {pk.nodes, pk.v1, pk.v2, pk.hcof, pk.brhs, pk.depth};
  for (void *b : old)
    if (b) hipFree(b);
  pk = DevPackage();
  pk.nbound = nbound;
  if (nbound > 0) {
    std::vector<int> nd(nbound);
    for (int e = 0; e < nbound; e++) {
      if (nodelist[e] < 1 || nodelist[e] > h->L.N) { mf6hip_set_error("package node out of range"); return MF6HIP_EINVAL; }
      nd[e] = nodelist[e] - 1;
    }
    CK(galloc(&pk.nodes, (size_t)nbound));
    CK(galloc(&pk.v1, (size_t)nbound));
    CK(galloc(&pk.hcof, (size_t)nbound));
    CK(galloc(&pk.brhs, (size_t)nbound));
    HIPCHK(hipMemcpy(pk.nodes, nd.data(), sizeof(int) * nbound, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(pk.v1, v1, sizeof(double) * nbound, hipMemcpyHostToDevice));
    if (v2) {
      CK(galloc(&pk.v2, (size_t)nbound));
      HIPCHK(hipMemcpy(pk.v2, v2, sizeof(double) * nbound, hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemset(pk.hcof, 0, sizeof(double) * nbound));
    HIPCHK(hipMemset(pk.brhs, 0, sizeof(double) * nbound));
  }
  // host copy of the node lists for the per-cell entry table
  h->bnd_dirty = true;
  static_assert(MF6HIP_PKG_CHD == 4, "package numbering");
  // keep nodelists on the host handle for rebuilding the table
  h->pk_host_nodes[pkgtype].assign(nbound, 0);
  for (int e = 0; e < nbound; e++) h->pk_host_nodes[pkgtype][e] = nodelist[e] - 1;
  return MF6HIP_OK;
}

// per-position list of bound entries in bndlist order (WEL, DRN, RCH; CHD
// contributes nothing to the matrix) and list order
static int rebuild_bnd_table(Mf6Handle *h) {
  const Layout &L = h->L;
  std::vector<int> cnt(L.NP + 1, 0);
  auto pos_of = [&](int n) { return pos_of_natural(L, n); };
  const int order[3] = {MF6HIP_PKG_WEL, MF6HIP_PKG_DRN, MF6HIP_PKG_RCH};
  for (int t : order)
    for (int n : h->pk_host_nodes[t]) cnt[pos_of(n) + 1]++;
  for (int p = 0; p < L.NP; p++) cnt[p + 1] += cnt[p];
  int tot = cnt[L.NP];
  std::vector<int> ent(tot > 0 ? tot : 1);
  std::vector<int> fill(cnt.begin(), cnt.end() - 1);
  for (int t : order) {
    const auto &nd = h->pk_host_nodes[t];
    for (int e = 0; e < (int)nd.size(); e++) ent[fill[pos_of(nd[e])]++] = (t << 28) | e;
  }
  if (h->bnd_off) hipFree(h->bnd_off);
  if (h->bnd_ent) hipFree(h->bnd_ent);
  h->bnd_off = nullptr;
  h->bnd_ent = nullptr;
  CK(galloc(&h->bnd_off, (size_t)L.NP + 1));
  CK(galloc(&h->bnd_ent, ent.size()));
  HIPCHK(hipMemcpy(h->bnd_off, cnt.data(), sizeof(int) * (L.NP + 1), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(h->bnd_ent, ent.data(), sizeof(int) * ent.size(), hipMemcpyHostToDevice));
  h->nbnd_ent = tot;
  h->bnd_dirty = false;
  return MF6HIP_OK;
}

extern "C" int mf6hip_gwf_set_x(void *handle, const double *x) {
  Mf6Handle *h = (Mf6Handle *)handle;
  if (!h || !x) { mf6hip_set_error("null argument"); return MF6HIP_EINVAL; }
  const Layout &L = h->L;
  std::vector<double> tmp(L.NP, 0.0);
  for (int n = 0; n < L.N; n++) tmp[pos_of_natural(L, n)] = x[n];
  HIPCHK(hipMemcpy(h->x, tmp.data(), sizeof(double) * L.NP, hipMemcpyHostToDevice));
  return MF6HIP_OK;
}

extern "C" int mf6hip_gwf_get_x(void *handle, double *x) {
  Mf6Handle *h = (Mf6Handle *)handle;
  if (!h || !x) { mf6hip_set_error("null argument"); return MF6HIP_EINVAL; }
  const Layout &L = h->L;
  std::vector<double> tmp(L.NP);
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipMemcpy(tmp.data(), h->x, sizeof(double) * L.NP, hipMemcpyDeviceToHost));
  for (int n = 0; n < L.N; n++) x[n] = tmp[pos_of_natural(L, n)];
  return MF6HIP_OK;
}

namespace {
// x_nat[n] = x[pos(n)]
__global__ void k_x_to_natural(Layout L, const double *__restrict__ x, double *__restrict__ xn) {
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < L.N; n += gridDim.x * blockDim.x)
    xn[n] = x[pos_of_natural(L, n)];
}
// x[pos(n)] = x_nat[n] (padding positions untouched)
__global__ void k_x_from_natural(Layout L, const double *__restrict__ xn, double *__restrict__ x) {
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < L.N; n += gridDim.x * blockDim.x)
    x[pos_of_natural(L, n)] = xn[n];
}
}  // namespace

extern "C" int mf6hip_gwf_set_x_device(void *handle, const double *x_dev) {
  Mf6Handle *h = (Mf6Handle *)handle;
  if (!h || !x_dev) { mf6hip_set_error("null argument"); return MF6HIP_EINVAL; }
  const Layout &L = h->L;
  k_x_from_natural<<<ggrid(L.N), MF6_BLOCK, 0, h->stream>>>(L, x_dev, h->x);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(h->stream));
  return MF6HIP_OK;
}

extern "C" int mf6hip_gwf_get_x_device(void *handle, double *x_dev) {
  Mf6Handle *h = (Mf6Handle *)handle;
  if (!h || !x_dev) { mf6hip_set_error("null argument"); return MF6HIP_EINVAL; }
  const Layout &L = h->L;
  k_x_to_natural<<<ggrid(L.N), MF6_BLOCK, 0, h->stream>>>(L, h->x, x_dev);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(h->stream));
  return MF6HIP_OK;
}

extern "C" int mf6hip_gwf_advance(void *handle, int iss, double delt) {
  Mf6Handle *h = (Mf6Handle *)handle;
  if (!h || !h->has_model) { mf6hip_set_error("no model on this handle"); return MF6HIP_EINVAL; }
  if (!iss && !(delt > 0.0)) {
    mf6hip_set_error("DETECTED TIME STEP LENGTH OF ZERO.  GWF STORAGE PACKAGE CANNOT BE USED UNLESS DELT IS NON-ZERO.");
    return MF6HIP_EINVAL;
  }
  const Layout &L = h->L;
  h->iss = iss;
  h->delt = delt;
  hipStream_t s = h->stream;
  k_advance<<<ggrid(L.NP), MF6_BLOCK, 0, s>>>(L, h->x, h->xold, h->ibound0, h->ibound);
  DevPackage &chd = h->pk[MF6HIP_PKG_CHD];
  if (chd.nbound > 0)
    k_chd_set<<<ggrid(chd.nbound), MF6_BLOCK, 0, s>>>(chd.nbound, chd.nodes, chd.v1, L, h->x,
                                                      h->xold, h->ibound);
  HIPCHK(hipGetLastError());
  return MF6HIP_OK;
}

// the kernels' view of the barriers hfb_fc / hfb_cq apply (none: null
// tables)
static HfbDev hfb_dev(const Mf6Handle *h) {
  HfbDev F;
  F.off = h->hfbd_n > 0 ? h->d_hfbd_off : nullptr;
  F.ent = h->d_hfbd_ent;
  F.np = h->d_hfbd_np;
  F.mp = h->d_hfbd_mp;
  F.s = h->d_hfbd_s;
  F.hy = h->d_hfbd_hy;
  F.w = h->d_hfbd_w;
  F.sav = h->d_hfbd_sav;
  return F;
}

// the kernels' view of the model's ghost-node corrections (none: null
// tables)
static GncDev gnc_dev(const Mf6Handle *h) {
  GncDev G;
  G.n = h->gnc_n;
  G.numjs = h->gnc_numjs;
  const bool on = h->gnc_n > 0;
  G.off = on ? h->d_gnc_off : nullptr;
  G.ent = h->d_gnc_ent;
  G.np = h->d_gnc_np;
  G.mp = h->d_gnc_mp;
  G.s = h->d_gnc_s;
  G.jp = h->d_gnc_jp;
  G.a = h->d_gnc_a;
  G.cond = h->d_gnc_cond;
  return G;
}

static void gnc_free(Mf6Handle *h) {
  void *bufs[] = {h->d_gnc_np, h->d_gnc_mp, h->d_gnc_s, h->d_gnc_jp,
                  h->d_gnc_a, h->d_gnc_cond, h->d_gnc_off, h->d_gnc_ent};
  for (void *b : bufs)
    if (b) hipFree(b);
  h->d_gnc_np = h->d_gnc_mp = h->d_gnc_s = h->d_gnc_jp = nullptr;
  h->d_gnc_off = h->d_gnc_ent = nullptr;
  h->d_gnc_a = h->d_gnc_cond = nullptr;
  h->gnc_n = h->gnc_numjs = 0;
}

// The model's GNC package (GhostNode.f90 gnc_df / read_data / gnc_mc): a new
// set replaces the last one.  Explicit form only: the implicit form adds
// the contributing cells to rows n and m of the matrix (gnc_ac :134-161),
// which a DIS grid's 7-point rows do not hold.
extern "C" int mf6hip_gwf_set_gnc(void *handle, int nexg, int numjs, const int *noden,
                                  const int *nodem, const int *nodesj, const double *alphasj,
                                  int implicit) {
  Mf6Handle *h = (Mf6Handle *)handle;
  if (!h || !h->has_model) { mf6hip_set_error("no model on this handle"); return MF6HIP_EINVAL; }
  if (nexg < 0 || numjs < 0 || (nexg > 0 && (numjs < 1 || !noden || !nodem || !nodesj ||
                                             !alphasj))) {
    mf6hip_set_error("bad GNC arrays");
    return MF6HIP_EINVAL;
  }
  if (implicit != 0 && nexg > 0) {
    mf6hip_set_error("implicit GNC adds the contributing cells to the matrix rows of n and m "
                     "(GhostNode.f90:134-161), outside a DIS grid's 7-point rows: use the "
                     "EXPLICIT option");
    return MF6HIP_EUNSUPPORTED;
  }
  const Layout &L = h->L;
  if ((long long)nexg * numjs >= (1LL << 30) || nexg >= (1 << 29)) {
    mf6hip_set_error("too many GNC entries");
    return MF6HIP_EUNSUPPORTED;
  }
  std::vector<int> np(nexg), mp(nexg), sd(nexg), jp((size_t)nexg * numjs);
  std::vector<int> cnt(L.NP + 1, 0);
  for (int g = 0; g < nexg; g++) {
    const int n = noden[g] - 1, m = nodem[g] - 1;
    if (n < 0 || n >= L.N || m < 0 || m >= L.N) {
      mf6hip_set_error("GNC " + std::to_string(g + 1) + ": node out of range");
      return MF6HIP_EINVAL;
    }
    const int kn = n / L.NRC, rn = n - kn * L.NRC, in = rn / L.ncol, jn = rn - in * L.ncol;
    const int km = m / L.NRC, rm = m - km * L.NRC, im = rm / L.ncol, jm = rm - im * L.ncol;
    int dir = -1;   // m seen from n: up, back, left, right, front, down
    if (rn == rm && km == kn - 1) dir = 0;
    else if (kn == km && jn == jm && im == in - 1) dir = 1;
    else if (kn == km && in == im && jm == jn - 1) dir = 2;
    else if (kn == km && in == im && jm == jn + 1) dir = 3;
    else if (kn == km && jn == jm && im == in + 1) dir = 4;
    else if (rn == rm && km == kn + 1) dir = 5;
    if (dir < 0) {
      mf6hip_set_error("GNC " + std::to_string(g + 1) + ": cell " + std::to_string(n + 1) +
                       " is not connected to cell " + std::to_string(m + 1));
      return MF6HIP_EINVAL;
    }
    np[g] = pos_of_natural(L, n);
    mp[g] = pos_of_natural(L, m);
    sd[g] = dir;
    for (int q = 0; q < numjs; q++) {
      const int jn1 = nodesj[(size_t)g * numjs + q];
      if (jn1 < 0 || jn1 > L.N) {
        mf6hip_set_error("GNC " + std::to_string(g + 1) + ": contributing node out of range");
        return MF6HIP_EINVAL;
      }
      jp[(size_t)g * numjs + q] = jn1 == 0 ? -1 : pos_of_natural(L, jn1 - 1);
    }
    cnt[np[g] + 1]++;
    cnt[mp[g] + 1]++;
  }
  for (int p = 0; p < L.NP; p++) cnt[p + 1] += cnt[p];
  std::vector<int> ent(2 * (size_t)(nexg > 0 ? nexg : 1));
  {
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for (int g = 0; g < nexg; g++) {
      ent[fill[np[g]]++] = g * 2;
      ent[fill[mp[g]]++] = g * 2 + 1;
    }
  }
  CK(hipStreamSynchronize(h->stream) == hipSuccess ? MF6HIP_OK : MF6HIP_EHIP);
  gnc_free(h);
  if (nexg == 0) return MF6HIP_OK;
  CK(galloc(&h->d_gnc_np, (size_t)nexg));
  CK(galloc(&h->d_gnc_mp, (size_t)nexg));
  CK(galloc(&h->d_gnc_s, (size_t)nexg));
  CK(galloc(&h->d_gnc_jp, (size_t)nexg * numjs));
  CK(galloc(&h->d_gnc_a, (size_t)nexg * numjs));
  CK(galloc(&h->d_gnc_cond, (size_t)nexg));
  CK(galloc(&h->d_gnc_off, (size_t)L.NP + 1));
  CK(galloc(&h->d_gnc_ent, 2 * (size_t)nexg));
  HIPCHK(hipMemcpy(h->d_gnc_np, np.data(), sizeof(int) * nexg, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(h->d_gnc_mp, mp.data(), sizeof(int) * nexg, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(h->d_gnc_s, sd.data(), sizeof(int) * nexg, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(h->d_gnc_jp, jp.data(), sizeof(int) * jp.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(h->d_gnc_a, alphasj, sizeof(double) * (size_t)nexg * numjs,
                   hipMemcpyHostToDevice));
  HIPCHK(hipMemset(h->d_gnc_cond, 0, sizeof(double) * nexg));
  HIPCHK(hipMemcpy(h->d_gnc_off, cnt.data(), sizeof(int) * (L.NP + 1), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(h->d_gnc_ent, ent.data(), sizeof(int) * 2 * nexg, hipMemcpyHostToDevice));
  h->gnc_n = nexg;
  h->gnc_numjs = numjs;
  return MF6HIP_OK;
}

// the kernels' view of package t (drains scale cubically under NEWTON or
// with DEV_CUBIC_SCALING, gwf-drn.f90:130-135)
static PkgDev pkg_dev(const Mf6Handle *h, int t) {
  const DevPackage &pk = h->pk[t];
  PkgDev P;
  P.kind = t;
  P.nb = pk.nbound;
  P.nodes = pk.nodes;
  P.v1 = pk.v1;
  P.v2 = pk.v2;
  P.depth = pk.depth;
  P.iflowred = pk.iflowred;
  P.flowred = pk.flowred;
  P.idepth = pk.idepth;
  P.icubic = (pk.icubic || h->m.inewton) ? 1 : 0;
  return P;
}

extern "C" int mf6hip_gwf_formulate(void *handle, int kiter, int inewton, double *amat_out,
                                    double *rhs_out) {
  (void)kiter;
  Mf6Handle *h = (Mf6Handle *)handle;
  if (!h || !h->has_model) { mf6hip_set_error("no model on this handle"); return MF6HIP_EINVAL; }
  if (inewton != 0 && inewton != 1) { mf6hip_set_error("inewton must be 0 or 1"); return MF6HIP_EINVAL; }
  const Layout &L = h->L;
  hipStream_t s = h->stream;
  if (h->bnd_dirty) CK(rebuild_bnd_table(h));
  for (int t = MF6HIP_PKG_WEL; t <= MF6HIP_PKG_RCH; t++) {
    DevPackage &pk = h->pk[t];
    if (pk.nbound > 0)
      k_bnd_cf<<<ggrid(pk.nbound), MF6_BLOCK, 0, s>>>(pkg_dev(h, t), make_dev(h), h->x, pk.hcof,
                                                      pk.brhs);
  }
  BndDev bd;
  bd.off = h->bnd_off;
  bd.ent = h->bnd_ent;
  for (int t = 0; t < 5; t++) {
    bd.hcof[t] = h->pk[t].hcof;
    bd.brhs[t] = h->pk[t].brhs;
    bd.pk[t] = pkg_dev(h, t);
  }
  GwfDev g = make_dev(h);
  g.fn = h->m.inewton && inewton;
  k_formulate<<<ggrid(L.N), MF6_BLOCK, 0, s>>>(g, bd, gnc_dev(h), hfb_dev(h), h->iss, h->delt, h->x,
                                               h->xold, h->A, h->b);
  HIPCHK(hipGetLastError());
  if (amat_out || rhs_out) CK(mf6hip_copy_system(h, amat_out, rhs_out));
  return MF6HIP_OK;
}

int mf6hip_dxmax(Mf6Handle *h, double *hncg, int *lrch, const double *xa, const double *xb,
                 bool) {
  const Layout &L = h->L;
  hipStream_t s = h->stream;
  int nb = ggrid(L.N);
  k_dxmax<<<nb, MF6_BLOCK, 0, s>>>(L, h->ibound, xa, xb, h->d_mpart);
  MaxLoc *out = (MaxLoc *)(h->d_part + 7 * MF6_MAXBLK);
  k_maxloc_final<<<1, MF6_BLOCK, 0, s>>>(h->d_mpart, nb, out);
  MaxLoc r;
  HIPCHK(hipMemcpyAsync(&r, out, sizeof(MaxLoc), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  *hncg = r.v;
  *lrch = r.loc + 1;
  return MF6HIP_OK;
}

// r = A x - b norms of the current device system: out[0] = ||r||_2 (active
// rows), out[1] = ||b||_2 (active rows), out[2] = max |r| / volume
static int resid_norms(Mf6Handle *h, double out[3]) {
  const Layout &L = h->L;
  hipStream_t s = h->stream;
  const int nb = ggrid(L.NP);
  double *dout = h->d_part + 7 * MF6_MAXBLK;
  k_resid_norms<<<nb, MF6_BLOCK, 0, s>>>(L, h->A, h->x, h->b, h->ibound, h->area, h->top, h->bot,
                                         h->d_part, nb);
  k_resid_final<<<1, MF6_BLOCK, 0, s>>>(h->d_part, nb, dout);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, dout, 3 * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return MF6HIP_OK;
}

extern "C" int mf6hip_gwf_residual_norms(void *handle, double *l2norm, double *rvmax) {
  Mf6Handle *h = (Mf6Handle *)handle;
  if (!h || !h->has_model || !l2norm || !rvmax) {
    mf6hip_set_error("null argument or no model");
    return MF6HIP_EINVAL;
  }
  double o[3];
  CK(resid_norms(h, o));
  *l2norm = o[0];
  *rvmax = o[2];
  return MF6HIP_OK;
}

extern "C" int mf6hip_gwf_backtrack(void *handle, double breduc, double dvclose, int *btflag) {
  Mf6Handle *h = (Mf6Handle *)handle;
  if (!h || !h->has_model || !btflag) {
    mf6hip_set_error("null argument or no model");
    return MF6HIP_EINVAL;
  }
  double dmax;
  int loc;
  CK(mf6hip_dxmax(h, &dmax, &loc, h->x, h->xtemp, true));
  *btflag = breduc * fabs(dmax) >= dvclose ? 1 : 0;
  if (*btflag) {
    const Layout &L = h->L;
    k_bt_apply<<<ggrid(L.NP), MF6_BLOCK, 0, h->stream>>>(L, h->ibound, breduc, h->x, h->xtemp);
    HIPCHK(hipGetLastError());
  }
  return MF6HIP_OK;
}

// the PTC block of sln_ls (NumericalSolution.f90:2480-2573) on the fixed system
static int apply_ptc(Mf6Handle *h, int kiter, mf6hip_ptc *pt) {
  int iallowptc = pt->iallowptc < 0 ? (pt->kper > 1 ? 1 : 0) : pt->iallowptc;
  int iptct = pt->iptc * iallowptc;
  double o[3] = {0.0, 0.0, 0.0};
  double l2norm = 0.0;
  if (iptct != 0) {
    CK(resid_norms(h, o));
    l2norm = o[0];
    if (kiter == 1) {
      if (pt->kper > 1 || pt->kstp > 1) {
        if (l2norm <= pt->l2norm0) pt->iptc = 0;
      }
    } else {
      double a = l2norm, b = pt->l2norm0;   // is_close, MathUtil.f90:45-86
      bool same = a == b ||
                  fabs(a - b) <= fmax(100.0 * 2.220446049250313080847263336181640625e-16 *
                                          fmax(fabs(a), fabs(b)), 0.0);
      if (same) pt->iptc = 0;
    }
  }
  iptct = pt->iptc * iallowptc;
  pt->iptc = iptct;   // out: 1 when the PTC terms are added
  if (iptct == 0) return MF6HIP_OK;
  if (kiter == 1) {
    if (pt->ptcdel0 > 0.0) pt->ptcdel = pt->ptcdel0;
    else if (pt->iptcopt == 0) pt->ptcdel = 1.0 / pt->ptcf;
    else pt->ptcdel = o[1] / l2norm;
  } else {
    if (l2norm > 0.0) pt->ptcdel = pt->ptcdel * pow(pt->l2norm0 / l2norm, pt->ptcexp);
    else pt->ptcdel = 0.0;
  }
  double ptcval = pt->ptcdel > 0.0 ? 1.0 / pt->ptcdel : 1.0;
  const Layout &L = h->L;
  k_ptc_apply<<<ggrid(L.NP), MF6_BLOCK, 0, h->stream>>>(L, h->ibound, ptcval, h->A, h->b, h->x);
  HIPCHK(hipGetLastError());
  pt->l2norm0 = l2norm;
  return MF6HIP_OK;
}

extern "C" int mf6hip_gwf_linear_solve(void *handle, int kstp, int kiter, mf6hip_ptc *ptc,
                                       int *icnvg, int *in_iter, double *hncg, int *lrch,
                                       mf6hip_cnvg_out *out) {
  Mf6Handle *h = (Mf6Handle *)handle;
  if (!h || !h->has_model || !icnvg || !in_iter || !hncg || !lrch) {
    mf6hip_set_error("null argument or no model");
    return MF6HIP_EINVAL;
  }
  const Layout &L = h->L;
  k_row_fixes<<<ggrid(L.N), MF6_BLOCK, 0, h->stream>>>(L, h->ibound, h->s.ilinmeth == 1, h->A,
                                                        h->b, h->x, h->xtemp);
  HIPCHK(hipGetLastError());
  if (ptc) CK(apply_ptc(h, kiter, ptc));
  CK(mf6hip_device_solve(h, kstp, kiter, icnvg, in_iter, out));
  CK(mf6hip_dxmax(h, hncg, lrch, h->x, h->xtemp, true));
  return MF6HIP_OK;
}

static int maxval_host(Mf6Handle *h, double *vmax) {
  // sln_maxval: sequential normalised scan (NumericalSolution.f90:2875-2905)
  const Layout &L = h->L;
  std::vector<double> tmp(L.NP);
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipMemcpy(tmp.data(), h->dxold, sizeof(double) * L.NP, hipMemcpyDeviceToHost));
  std::vector<double> nat(L.N);
  for (int n = 0; n < L.N; n++) nat[n] = tmp[pos_of_natural(L, n)];
  double v = nat[0];
  for (int n = 1; n < L.N; n++) {
    double d = nat[n];
    double denom = fabs(v);
    if (denom == 0.0) denom = 2.220446049250313080847263336181640625e-16;
    if (fabs(d) / denom > 1.0) v = d;
  }
  *vmax = v;
  return MF6HIP_OK;
}

extern "C" int mf6hip_gwf_outer_update(void *handle, int kiter, mf6hip_underrelax *ur,
                                       int *inewtonur, double *dxold_max, double *hncg,
                                       int *lrch) {
  Mf6Handle *h = (Mf6Handle *)handle;
  if (!h || !h->has_model || !inewtonur || !dxold_max || !hncg || !lrch) {
    mf6hip_set_error("null argument or no model");
    return MF6HIP_EINVAL;
  }
  const Layout &L = h->L;
  hipStream_t s = h->stream;
  const int nonmeth = ur ? ur->nonmeth : 0;
  if (nonmeth < 0 || nonmeth > 3) { mf6hip_set_error("nonmeth must be 0..3"); return MF6HIP_EINVAL; }
  if (nonmeth == 1) {
    k_ur_relax<<<ggrid(L.NP), MF6_BLOCK, 0, s>>>(L, h->ibound, ur->gamma, h->x, h->xtemp,
                                                 h->dxold);
  } else if (nonmeth == 2) {
    // Cooley (NumericalSolution.f90:3023-3061): scalar recurrence on the host
    double relax;
    double bigch = ur->bigch;
    if (kiter == 1) {
      relax = 1.0;
      ur->relaxold = 1.0;
      ur->bigchold = bigch;
    } else {
      double es = bigch / (ur->bigchold * ur->relaxold);
      double aes = fabs(es);
      if (es < -1.0) relax = 0.5 / aes;
      else relax = (3.0 + es) / (3.0 + aes);
    }
    ur->relaxold = relax;
    ur->bigchold = (1.0 - ur->gamma) * bigch + ur->gamma * ur->bigchold;
    if (relax < 1.0)
      k_ur_relax<<<ggrid(L.NP), MF6_BLOCK, 0, s>>>(L, h->ibound, relax, h->x, h->xtemp,
                                                   h->dxold);
  } else if (nonmeth == 3) {
    if (!h->wsave) {
      CK(galloc(&h->wsave, (size_t)L.NP));
      CK(galloc(&h->hchold, (size_t)L.NP));
      CK(galloc(&h->deold, (size_t)L.NP));
      HIPCHK(hipMemsetAsync(h->wsave, 0, sizeof(double) * L.NP, s));
      HIPCHK(hipMemsetAsync(h->hchold, 0, sizeof(double) * L.NP, s));
      HIPCHK(hipMemsetAsync(h->deold, 0, sizeof(double) * L.NP, s));
    }
    double theta = ur->theta < 1.0e-3 ? 1.0e-3 : ur->theta;   // sln_ar :868-870
    k_ur_dbd<<<ggrid(L.NP), MF6_BLOCK, 0, s>>>(L, h->ibound, kiter, theta, ur->akappa,
                                               ur->gamma, ur->amomentum, h->x, h->xtemp,
                                               h->dxold, h->wsave, h->hchold, h->deold);
  } else {
    k_calcdx<<<ggrid(L.N), MF6_BLOCK, 0, s>>>(L, h->ibound, h->x, h->xtemp, h->dxold);
  }
  *inewtonur = 0;
  *dxold_max = 0.0;
  if (h->m.inewton && h->m.inewtonur) {
    int *flag = h->d_fail;
    HIPCHK(hipMemsetAsync(flag, 0, sizeof(int), s));
    k_nur<<<ggrid(L.N), MF6_BLOCK, 0, s>>>(L, h->ibound, h->icelltype, h->ibotnode, h->bot, h->x,
                                           h->xtemp, h->dxold, flag);
    int hf = 0;
    HIPCHK(hipMemcpyAsync(&hf, flag, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    *inewtonur = hf;
    if (hf) {
      CK(maxval_host(h, dxold_max));
      CK(mf6hip_dxmax(h, hncg, lrch, h->x, h->xtemp, true));
    }
  }
  HIPCHK(hipGetLastError());
  return MF6HIP_OK;
}

// gwf_cq + gwf_bd (gwf.f90:741-800) of the current heads: flowja in the CSR
// order of mf6hip_create (optional), the budget terms (rate in, rate out) of
// STO-SS, STO-SY, WEL, DRN, RCH, CHD and the percent discrepancy
// (Budget.f90:259-267).
extern "C" int mf6hip_gwf_budget(void *handle, double *flowja_out, double *terms,
                                 double *pct_discrepancy) {
  Mf6Handle *h = (Mf6Handle *)handle;
  if (!h || !h->has_model) { mf6hip_set_error("no model on this handle"); return MF6HIP_EINVAL; }
  if (!terms || !pct_discrepancy) { mf6hip_set_error("null argument"); return MF6HIP_EINVAL; }
  const Layout &L = h->L;
  hipStream_t s = h->stream;
  if (h->bnd_dirty) CK(rebuild_bnd_table(h));
  const int pkorder[3] = {MF6HIP_PKG_WEL, MF6HIP_PKG_DRN, MF6HIP_PKG_RCH};
  size_t need = 0;
  for (int t : pkorder) need += (size_t)h->pk[t].nbound;
  need += 3 * (size_t)h->pk[MF6HIP_PKG_CHD].nbound;
  CK(galloc(&h->flowja, (size_t)h->nja));
  if (!h->strg) {
    // the storage rates are summed over every position: the padding slots
    // of an odd nrow * ncol layout, which k_budget_cells never writes, must
    // hold zero
    CK(galloc(&h->strg, 2 * (size_t)L.NP));
    HIPCHK(hipMemsetAsync(h->strg, 0, sizeof(double) * 2 * (size_t)L.NP, s));
  }
  CK(galloc(&h->chd_flag, (size_t)L.NP));
  if (need > h->bsim_n) {
    if (h->bsim) hipFree(h->bsim);
    h->bsim = nullptr;
    CK(galloc(&h->bsim, need));
    h->bsim_n = need;
  }
  // bnd_cf with the current heads, then the rates (bnd_cq)
  BndRates br;
  br.off = h->bnd_off;
  br.ent = h->bnd_ent;
  for (int t = 0; t < 5; t++) br.rate[t] = nullptr;
  double *seg = h->bsim;
  const double *segs[4] = {nullptr, nullptr, nullptr, nullptr};
  for (int q = 0; q < 3; q++) {
    const int t = pkorder[q];
    DevPackage &pk = h->pk[t];
    segs[q] = seg;
    h->bsim_seg_off[q] = (size_t)(seg - h->bsim);
    h->bsim_seg_len[q] = pk.nbound;
    if (pk.nbound > 0) {
      k_bnd_cf<<<ggrid(pk.nbound), MF6_BLOCK, 0, s>>>(pkg_dev(h, t), make_dev(h), h->x, pk.hcof,
                                                      pk.brhs);
      k_bnd_rate<<<ggrid(pk.nbound), MF6_BLOCK, 0, s>>>(pk.nbound, pk.nodes, L, h->ibound, h->x,
                                                        pk.hcof, pk.brhs, seg);
    }
    br.rate[t] = seg;
    seg += pk.nbound;
  }
  DevPackage &chd = h->pk[MF6HIP_PKG_CHD];
  double *chd_in = seg, *chd_out = seg + chd.nbound, *chd_sim = seg + 2 * (size_t)chd.nbound;
  h->bsim_seg_off[3] = (size_t)(chd_sim - h->bsim);
  h->bsim_seg_len[3] = chd.nbound;
  HIPCHK(hipMemsetAsync(h->chd_flag, 0, sizeof(int) * L.NP, s));
  if (chd.nbound > 0)
    k_mark_cells<<<ggrid(chd.nbound), MF6_BLOCK, 0, s>>>(chd.nbound, chd.nodes, L, h->chd_flag);
  GwfDev g = make_dev(h);
  k_budget_cells<<<ggrid(L.N), MF6_BLOCK, 0, s>>>(g, br, gnc_dev(h), hfb_dev(h), h->d_ia,
                                                  h->chd_flag,
                                                  h->iss, h->delt, h->x, h->xold, h->flowja,
                                                  h->strg);
  if (chd.nbound > 0)
    k_chd_inout<<<ggrid(chd.nbound), MF6_BLOCK, 0, s>>>(chd.nbound, chd.nodes, h->d_ia, L,
                                                       h->ibound, h->flowja, chd_in, chd_out,
                                                       chd_sim);
  HIPCHK(hipGetLastError());
  // rate_accumulator of each term: STO-SS, STO-SY, WEL, DRN, RCH, then the
  // CHD in and out lists (all non-negative: their "in" sums)
  const double *arr[7] = {h->strg, h->strg + L.NP, segs[0], segs[1], segs[2], chd_in, chd_out};
  const int len[7] = {L.NP, L.NP, h->pk[MF6HIP_PKG_WEL].nbound, h->pk[MF6HIP_PKG_DRN].nbound,
                      h->pk[MF6HIP_PKG_RCH].nbound, chd.nbound, chd.nbound};
  double *dout = h->d_part + 7 * MF6_MAXBLK;
  for (int t = 0; t < 7; t++) {
    if (len[t] == 0) continue;
    const int nb = ggrid(len[t]) > 1024 ? 1024 : ggrid(len[t]);
    k_rate_acc<<<nb, MF6_BLOCK, 0, s>>>(len[t], arr[t], h->d_part);
    k_rate_final<<<1, 64, 0, s>>>(h->d_part, nb, dout + 2 * t);
  }
  HIPCHK(hipGetLastError());
  double acc[14] = {0.0};
  HIPCHK(hipMemcpyAsync(acc, dout, sizeof(acc), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  for (int t = 0; t < 7; t++)
    if (len[t] == 0) { acc[2 * t] = 0.0; acc[2 * t + 1] = 0.0; }
  for (int t = 0; t < 5; t++) { terms[2 * t] = acc[2 * t]; terms[2 * t + 1] = acc[2 * t + 1]; }
  terms[10] = acc[10];   // CHD in = sum of ratechdin
  terms[11] = acc[12];   // CHD out = sum of ratechdout
  double totin = 0.0, totout = 0.0;
  for (int t = 0; t < 6; t++) {
    totin = totin + terms[2 * t];
    totout = totout + terms[2 * t + 1];
  }
  const double avg = (totin + totout) / 2.0;
  *pct_discrepancy = avg != 0.0 ? 100.0 * (totin - totout) / avg : 0.0;
  if (flowja_out)
    HIPCHK(hipMemcpy(flowja_out, h->flowja, sizeof(double) * h->nja, hipMemcpyDeviceToHost));
  h->budget_done = 1;
  return MF6HIP_OK;
}

// One flow term of the last mf6hip_gwf_budget: the STO-SS / STO-SY rate of
// every cell (natural order; sto%strgss / strgsy, gwf-sto.f90:447-564) or a
// package's simvals in bound order (bnd_cq, BoundaryPackage.f90:545-619;
// CHD: calc_chd_rate, gwf-chd.f90:264-320).  *count receives the number of
// values; out may be NULL to query it.
extern "C" int mf6hip_gwf_flow_term(void *handle, int term, double *out, int *count) {
  Mf6Handle *h = (Mf6Handle *)handle;
  if (!h || !h->has_model) { mf6hip_set_error("no model on this handle"); return MF6HIP_EINVAL; }
  if (!count) { mf6hip_set_error("null argument"); return MF6HIP_EINVAL; }
  if (term < MF6HIP_TERM_STO_SS || term > MF6HIP_TERM_CHD) {
    mf6hip_set_error("unknown flow term");
    return MF6HIP_EINVAL;
  }
  if (!h->budget_done) {
    mf6hip_set_error("no budget has been computed on this handle");
    return MF6HIP_EINVAL;
  }
  const Layout &L = h->L;
  HIPCHK(hipStreamSynchronize(h->stream));
  if (term <= MF6HIP_TERM_STO_SY) {
    *count = L.N;
    if (!out) return MF6HIP_OK;
    std::vector<double> tmp(L.NP);
    HIPCHK(hipMemcpy(tmp.data(), h->strg + (size_t)(term - MF6HIP_TERM_STO_SS) * L.NP,
                     sizeof(double) * L.NP, hipMemcpyDeviceToHost));
    for (int n = 0; n < L.N; n++) out[n] = tmp[pos_of_natural(L, n)];
    return MF6HIP_OK;
  }
  const int q = term - MF6HIP_TERM_WEL;
  *count = h->bsim_seg_len[q];
  if (out && *count > 0)
    HIPCHK(hipMemcpy(out, h->bsim + h->bsim_seg_off[q], sizeof(double) * (size_t)*count,
                     hipMemcpyDeviceToHost));
  return MF6HIP_OK;
}
