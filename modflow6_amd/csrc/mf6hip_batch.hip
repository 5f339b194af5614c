// mf6hip_batch.hip -- an ensemble of independent steady systems (uneven DIS
// shapes) solved together: mixed-precision PCG with reliable updates (FP32
// matrix, pivots and Krylov vectors; FP64 heads and FP64 residual
// replacement), every kernel launch covering all systems, the iteration
// sequence captured once as a HIP graph and replayed until every system has
// stopped.
//
// Work is cut into tiles of columns: a tile is (system, first column) of
// 256 * cpt columns (cpt = columns per thread); red, black and interleaved
// (both colours, thread t -> colour t & 1) tile lists exist for the three
// kinds of stencil kernels.  Kernels walk the tile lists with a persistent
// grid (gridDim blocks, tile = block, block + gridDim, ...), skip systems
// that have stopped, and write one reduction partial per tile; one block per
// system then reduces its contiguous tile range.  cpt and the grid size are
// chosen by an autotuner that times a short replay of each candidate.
//
// Per system the arithmetic is that of the single-system fused CG
// (ims_base_cg, src/Solution/LinearMethods/ImsLinearBase.f90:30-240, on the
// column red-black ordering), including the IMS convergence test.
//
// With preconditioner 2 every system also carries the two-level coarse
// correction of the single-system solver (mf6hip_coarse.hip): its coarse
// inverse is built when the system is staged, the black / red backward
// kernels write its per-column residual sums, kb_tl_gather / kb_tl_gemv
// form yc = Ainv P^T d over per-system tasks of four coarse rows, the
// direction update adds P yc, the round kernel adds rc.yc to rho, and each
// residual replacement restarts the directions.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "mf6hip_handle.h"

namespace {

#define BCK(x)                        \
  do {                                \
    int rc_ = (x);                    \
    if (rc_ != MF6HIP_OK) return rc_; \
  } while (0)

struct BSys {
  Layout L;
  long long aoff;   // first of the 7 planes (each NP long)
  long long voff;   // first vector entry
  int tr0, ntr;     // red tiles
  int tb0, ntb;     // black tiles
  int ti0, nti;     // interleaved tiles
  // two-level preconditioner: coarse shape; offsets of the per-column
  // coupled-cell masks (2 HP), the per-grid-column residual sums (NRC), the
  // coarse vectors (ld) and the FP32 coarse inverse (nc ld); the system's
  // coarse GEMV blocks tg0 .. tg0 + ntg - 1
  TwoLevel tl;
  long long cmo, cso, cvo, aio;
  int tg0, ntg;
};

struct BFState {
  double apend_x, apend_d, pmax;
  int ploc, pcur;
  int restart;      // the next round starts a new direction (beta = 0)
};

struct Tile {
  int sys, base;
};

enum { P_SS_R = 0, P_SS_B = 1, P_RHO_B = 2, P_RHO_R = 3, P_PQ = 4, P_RHO_C = 5 };
enum { M_D_R = 0, M_D_B = 1, M_P = 2, M_TRUE = 3 };

template <typename T>
__device__ inline T wsum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

__device__ inline MaxLoc wmax(MaxLoc a) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    MaxLoc b;
    b.v = __shfl_xor(a.v, o);
    b.loc = __shfl_xor(a.loc, o);
    a = maxloc_pick(a, b);
  }
  return a;
}

// block sum / max, result valid in thread 0
__device__ inline double bsum(double v, double *lds) {
  v = wsum(v);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) lds[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) s = s + lds[i];
  return s;
}

__device__ inline MaxLoc bmax(MaxLoc a, MaxLoc *lds) {
  a = wmax(a);
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) lds[w] = a;
  __syncthreads();
  MaxLoc r{0.0, -1};
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) r = maxloc_pick(r, lds[i]);
  return r;
}

// ---------------------------------------------------------------------------
// stencil kernels (FP32)
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(MF6_BLOCK) kb_red_fwd(const BSys *sys, const Tile *tiles,
                                                        int ntiles, int cpt, const float *A32,
                                                        const float *piv, float *d,
                                                        const float *q, float *zf, double *part,
                                                        MaxLoc *mpart, int T, const KState *st,
                                                        const BFState *fs) {
  __shared__ double lds[8];
  __shared__ MaxLoc mlds[8];
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int s = tiles[tile].sys;
    if (st[s].stop) continue;
    const Layout L = sys[s].L;
    const float *A = A32 + sys[s].aoff;
    const float *pv = piv + sys[s].voff;
    float *dd = d + sys[s].voff;
    const float *qq = q + sys[s].voff;
    float *zz = zf + sys[s].voff;
    const int NP = L.NP;
    const float alpha = (float)fs[s].apend_d;
    double ss = 0.0;
    MaxLoc md{0.0, -1};
    for (int c = 0; c < cpt; c++) {
      const int h2 = tiles[tile].base + c * blockDim.x + threadIdx.x;
      if (h2 >= L.HP0) continue;
      const int l = col_l(L, 0, h2);
      float zup = 0.f, lup = 0.f;
      for (int k = 0; k < L.nlay; k++) {
        const int pos = k * L.HP + h2;
        float dv = dd[pos];
        dv = dv - alpha * qq[pos];
        dd[pos] = dv;
        md = maxloc_pick(md, MaxLoc{(double)dv, k * L.NRC + l});
        ss = ss + (double)dv * (double)dv;
        float tv = dv;
        if (k > 0) tv = tv - lup * zup;
        zz[pos] = tv;
        zup = tv;
        lup = A[6 * NP + pos] * pv[pos];
      }
    }
    ss = bsum(ss, lds);
    if (threadIdx.x == 0) part[P_SS_R * T + tile] = ss;
    md = bmax(md, mlds);
    if (threadIdx.x == 0) mpart[M_D_R * T + tile] = md;
  }
}

__global__ void __launch_bounds__(MF6_BLOCK) kb_black(const BSys *sys, const Tile *tiles,
                                                      int ntiles, int cpt, const float *A32,
                                                      const float *piv, float *d, const float *q,
                                                      const float *zf, float *z, double *part,
                                                      MaxLoc *mpart, int T, const KState *st,
                                                      const BFState *fs,
                                                      const unsigned *cmask, double *colsum) {
  __shared__ double lds[8];
  __shared__ MaxLoc mlds[8];
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int s = tiles[tile].sys;
    if (st[s].stop) continue;
    const Layout L = sys[s].L;
    const float *A = A32 + sys[s].aoff;
    const float *pv = piv + sys[s].voff;
    const unsigned *cms = colsum ? cmask + sys[s].cmo : nullptr;
    double *css = colsum ? colsum + sys[s].cso : nullptr;
    float *dd = d + sys[s].voff;
    const float *qq = q + sys[s].voff;
    const float *zr = zf + sys[s].voff;
    float *zz = z + sys[s].voff;
    const int NP = L.NP;
    const float alpha = (float)fs[s].apend_d;
    double ss = 0.0, rho = 0.0;
    MaxLoc md{0.0, -1};
    for (int c = 0; c < cpt; c++) {
      const int h2 = tiles[tile].base + c * blockDim.x + threadIdx.x;
      if (h2 >= L.HP1) continue;
      const int l = col_l(L, 1, h2);
      const int i = l / L.ncol, j = l - i * L.ncol;
      // off-grid neighbours read position 0 (in range); a select drops
      // their term, so no lane waits inside a branch
      const bool hb = i > 0, hl = j > 0, hr = j < L.ncol - 1, hf = i < L.nrow - 1;
      const int ib = hb ? hn(l, -L.ncol) : 0;
      const int il = hl ? hn(l, -1) : 0;
      const int ir = hr ? hn(l, 1) : 0;
      const int ifr = hf ? hn(l, L.ncol) : 0;
      float zup = 0.f, lup = 0.f;
      const unsigned cm = css ? cms[L.HP + h2] : 0u;
      double cs = 0.0;
      for (int k = 0; k < L.nlay; k++) {
        const int pos = L.H + k * L.HP + h2;
        const int rb = k * L.HP;
        float dv = dd[pos];
        dv = dv - alpha * qq[pos];
        dd[pos] = dv;
        if ((cm >> k) & 1u) cs = cs + (double)dv;
        md = maxloc_pick(md, MaxLoc{(double)dv, k * L.NRC + l});
        ss = ss + (double)dv * (double)dv;
        float tv = dv;
        {
          const auto t_ = (A[5 * NP + rb + ib] * pv[rb + ib]) * zr[rb + ib];
          tv = hb ? tv - t_ : tv;
        }
        {
          const auto t_ = (A[4 * NP + rb + il] * pv[rb + il]) * zr[rb + il];
          tv = hl ? tv - t_ : tv;
        }
        {
          const auto t_ = (A[4 * NP + pos] * pv[rb + ir]) * zr[rb + ir];
          tv = hr ? tv - t_ : tv;
        }
        {
          const auto t_ = (A[5 * NP + pos] * pv[rb + ifr]) * zr[rb + ifr];
          tv = hf ? tv - t_ : tv;
        }
        if (k > 0) tv = tv - lup * zup;
        zz[pos] = tv;
        zup = tv;
        lup = A[6 * NP + pos] * pv[pos];
      }
      if (css) css[l] = cs;
      float zdn = 0.f;
      for (int k = L.nlay - 1; k >= 0; k--) {
        const int pos = L.H + k * L.HP + h2;
        float tv = zz[pos];
        if (k < L.nlay - 1) tv = tv - A[6 * NP + pos] * zdn;
        tv = tv * pv[pos];
        zz[pos] = tv;
        zdn = tv;
        rho = rho + (double)dd[pos] * (double)tv;
      }
    }
    ss = bsum(ss, lds);
    if (threadIdx.x == 0) part[P_SS_B * T + tile] = ss;
    rho = bsum(rho, lds);
    if (threadIdx.x == 0) part[P_RHO_B * T + tile] = rho;
    md = bmax(md, mlds);
    if (threadIdx.x == 0) mpart[M_D_B * T + tile] = md;
  }
}

__global__ void __launch_bounds__(MF6_BLOCK) kb_red_bwd(const BSys *sys, const Tile *tiles,
                                                        int ntiles, int cpt, const float *A32,
                                                        const float *piv, const float *d,
                                                        const float *zf, float *z, double *part,
                                                        int T, const KState *st,
                                                        const unsigned *cmask, double *colsum) {
  __shared__ double lds[8];
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int s = tiles[tile].sys;
    if (st[s].stop) continue;
    const Layout L = sys[s].L;
    const float *A = A32 + sys[s].aoff;
    const float *pv = piv + sys[s].voff;
    const unsigned *cms = colsum ? cmask + sys[s].cmo : nullptr;
    double *css = colsum ? colsum + sys[s].cso : nullptr;
    const float *dd = d + sys[s].voff;
    const float *zr = zf + sys[s].voff;
    float *zz = z + sys[s].voff;
    const int NP = L.NP;
    double rho = 0.0;
    for (int c = 0; c < cpt; c++) {
      const int h2 = tiles[tile].base + c * blockDim.x + threadIdx.x;
      if (h2 >= L.HP0) continue;
      const int l = col_l(L, 0, h2);
      const int i = l / L.ncol, j = l - i * L.ncol;
      // off-grid neighbours read position 0 (in range); a select drops
      // their term, so no lane waits inside a branch
      const bool hb = i > 0, hl = j > 0, hr = j < L.ncol - 1, hf = i < L.nrow - 1;
      const int ib = hb ? hn(l, -L.ncol) : 0;
      const int il = hl ? hn(l, -1) : 0;
      const int ir = hr ? hn(l, 1) : 0;
      const int ifr = hf ? hn(l, L.ncol) : 0;
      float zdn = 0.f;
      const unsigned cm = css ? cms[h2] : 0u;
      double cs = 0.0;
      for (int k = L.nlay - 1; k >= 0; k--) {
        const int pos = k * L.HP + h2;
        const int bb = L.H + k * L.HP;
        float tv = zr[pos];
        if (k < L.nlay - 1) tv = tv - A[6 * NP + pos] * zdn;
        {
          const auto t_ = A[5 * NP + bb + ib] * zz[bb + ib];
          tv = hb ? tv - t_ : tv;
        }
        {
          const auto t_ = A[4 * NP + bb + il] * zz[bb + il];
          tv = hl ? tv - t_ : tv;
        }
        {
          const auto t_ = A[4 * NP + pos] * zz[bb + ir];
          tv = hr ? tv - t_ : tv;
        }
        {
          const auto t_ = A[5 * NP + pos] * zz[bb + ifr];
          tv = hf ? tv - t_ : tv;
        }
        tv = tv * pv[pos];
        zz[pos] = tv;
        zdn = tv;
        const float dv = dd[pos];
        rho = rho + (double)dv * (double)tv;
        if ((cm >> k) & 1u) cs = cs + (double)dv;
      }
      if (css) css[l] = cs;
    }
    rho = bsum(rho, lds);
    if (threadIdx.x == 0) part[P_RHO_R * T + tile] = rho;
  }
}

// p_out = z (first iteration) or z + beta p_in ; e += apend_x p_in ; max |p|
__global__ void __launch_bounds__(MF6_BLOCK) kb_pupdate(const BSys *sys, const Tile *tiles,
                                                        int ntiles, int cpt, const float *z,
                                                        float *p0, float *p1, float *e,
                                                        MaxLoc *mpart, int T, const KState *st,
                                                        const BFState *fs, const unsigned *cmask,
                                                        const float *yc) {
  __shared__ MaxLoc mlds[8];
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int s = tiles[tile].sys;
    if (st[s].stop) continue;
    const Layout L = sys[s].L;
    const long long vo = sys[s].voff;
    const float *pin = (fs[s].pcur ? p1 : p0) + vo;
    float *pout = (fs[s].pcur ? p0 : p1) + vo;
    const float *zz = z + vo;
    float *ee = e + vo;
    const bool first = st[s].iter == 0;
    const float beta = (float)st[s].beta;
    const float ax = (float)fs[s].apend_x;
    MaxLoc mp{0.0, -1};
    for (int c = 0; c < cpt; c++) {
      const int t = tiles[tile].base + c * blockDim.x + threadIdx.x;
      const int cc = t & 1, h2 = t >> 1;
      if (h2 >= ncols_of(L, cc)) continue;
      const int l = col_l(L, cc, h2);
      const int own = cc * L.H + h2;
      // two-level: z completed by the prolongation of the coarse correction
      const unsigned cm = yc ? cmask[sys[s].cmo + cc * L.HP + h2] : 0u;
      const float yv = cm ? yc[sys[s].cvo + tl_agg(sys[s].tl, L, l)] : 0.f;
      for (int k = 0; k < L.nlay; k++) {
        const int pos = own + k * L.HP;
        float zv = zz[pos];
        if ((cm >> k) & 1u) zv = zv + yv;
        float pv;
        if (first) {
          pv = zv;
        } else {
          const float po = pin[pos];
          pv = zv + beta * po;
          ee[pos] = ee[pos] + ax * po;
        }
        pout[pos] = pv;
        mp = maxloc_pick(mp, MaxLoc{(double)pv, k * L.NRC + l});
      }
    }
    mp = bmax(mp, mlds);
    if (threadIdx.x == 0) mpart[M_P * T + tile] = mp;
  }
}

// q = A p (symmetric planes), partial p.q
__global__ void __launch_bounds__(MF6_BLOCK) kb_spmv(const BSys *sys, const Tile *tiles,
                                                     int ntiles, int cpt, const float *A32,
                                                     const float *p0, const float *p1, float *q,
                                                     double *part, int T, const KState *st,
                                                     const BFState *fs) {
  __shared__ double lds[8];
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int s = tiles[tile].sys;
    if (st[s].stop) continue;
    const Layout L = sys[s].L;
    const long long vo = sys[s].voff;
    const float *A = A32 + sys[s].aoff;
    const float *p = (fs[s].pcur ? p0 : p1) + vo;   // the direction just formed
    float *qq = q + vo;
    const int NP = L.NP;
    double sum = 0.0;
    for (int c = 0; c < cpt; c++) {
      const int t = tiles[tile].base + c * blockDim.x + threadIdx.x;
      const int cc = t & 1, h2 = t >> 1;
      if (h2 >= ncols_of(L, cc)) continue;
      const int l = col_l(L, cc, h2);
      const int i = l / L.ncol, j = l - i * L.ncol;
      // off-grid neighbours read position 0 (in range); a select drops
      // their term, so no lane waits inside a branch
      const bool hb = i > 0, hl = j > 0, hr = j < L.ncol - 1, hf = i < L.nrow - 1;
      const int ib = hb ? hn(l, -L.ncol) : 0;
      const int il = hl ? hn(l, -1) : 0;
      const int ir = hr ? hn(l, 1) : 0;
      const int ifr = hf ? hn(l, L.ncol) : 0;
      const int own = cc * L.H + h2;
      const int oth = (1 - cc) * L.H;
      float pprev = 0.f, pcur = p[own], aup = 0.f;
      for (int k = 0; k < L.nlay; k++) {
        const int pos = own + k * L.HP;
        const int ob = oth + k * L.HP;
        const float pnext = (k + 1 < L.nlay) ? p[pos + L.HP] : 0.f;
        const float adn = A[6 * NP + pos];
        float tq = 0.f;
        tq = tq + A[pos] * pcur;
        if (k > 0) tq = tq + aup * pprev;
        {
          const auto t_ = A[5 * NP + ob + ib] * p[ob + ib];
          tq = hb ? tq + t_ : tq;
        }
        {
          const auto t_ = A[4 * NP + ob + il] * p[ob + il];
          tq = hl ? tq + t_ : tq;
        }
        {
          const auto t_ = A[4 * NP + pos] * p[ob + ir];
          tq = hr ? tq + t_ : tq;
        }
        {
          const auto t_ = A[5 * NP + pos] * p[ob + ifr];
          tq = hf ? tq + t_ : tq;
        }
        if (k < L.nlay - 1) tq = tq + adn * pnext;
        qq[pos] = tq;
        sum = sum + (double)pcur * (double)tq;
        aup = adn;
        pprev = pcur;
        pcur = pnext;
      }
    }
    sum = bsum(sum, lds);
    if (threadIdx.x == 0) part[P_PQ * T + tile] = sum;
  }
}

// ---------------------------------------------------------------------------
// per-system scalar kernels (one block per system)
// ---------------------------------------------------------------------------

__device__ inline void testcnvg0(int *icnvg, double dvmax, double rmax, double dvclose,
                                 double rclose) {
  if (fabs(dvmax) <= dvclose && fabs(rmax) <= rclose) *icnvg = 1;
}

__device__ inline int is_close_b(double a, double b) {
  if (a == b) return 1;
  return fabs(a - b) <= fmax(100.0 * 2.220446049250313080847263336181640625e-16 *
                                 fmax(fabs(a), fabs(b)), 0.0);
}

__global__ void __launch_bounds__(MF6_BLOCK) kb_round(const BSys *sys, KState *st, BFState *fs,
                                                      const double *part, const MaxLoc *mpart,
                                                      int T) {
  __shared__ double lds[8];
  __shared__ MaxLoc mlds[8];
  const int s = blockIdx.x;
  if (st[s].stop) return;
  const BSys S = sys[s];
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  MaxLoc mm[2] = {{0.0, -1}, {0.0, -1}};
  for (int i = threadIdx.x; i < S.ntr; i += blockDim.x) {
    a[0] = a[0] + part[P_SS_R * T + S.tr0 + i];
    a[3] = a[3] + part[P_RHO_R * T + S.tr0 + i];
    mm[0] = maxloc_pick(mm[0], mpart[M_D_R * T + S.tr0 + i]);
  }
  for (int i = threadIdx.x; i < S.ntb; i += blockDim.x) {
    a[1] = a[1] + part[P_SS_B * T + S.tb0 + i];
    a[2] = a[2] + part[P_RHO_B * T + S.tb0 + i];
    mm[1] = maxloc_pick(mm[1], mpart[M_D_B * T + S.tb0 + i]);
  }
  // two-level: rho gains rc.yc (partials of the system's coarse GEMV blocks)
  for (int i = threadIdx.x; i < S.ntg; i += blockDim.x) a[3] = a[3] + part[P_RHO_C * T + S.tg0 + i];
  const double ssr = bsum(a[0], lds);
  const double ssb = bsum(a[1], lds);
  const double rb = bsum(a[2], lds);
  const double rr = bsum(a[3], lds);
  const MaxLoc mdr = bmax(mm[0], mlds);
  const MaxLoc mdb = bmax(mm[1], mlds);
  if (threadIdx.x != 0) return;
  KState &k = st[s];
  BFState &f = fs[s];
  const MaxLoc md = maxloc_pick(mdr, mdb);
  const double rho_new = rb + rr;
  const int it = k.iter;
  f.apend_d = 0.0;
  if (it > 0) {
    const double deltax = k.alpha * f.pmax;
    k.deltax = deltax;
    k.xloc = f.ploc;
    k.rmax = md.v;
    k.rloc = md.loc;
    k.l2norm = sqrt(ssr + ssb);
    int icnvg = k.icnvg;
    testcnvg0(&icnvg, deltax, md.v, k.dvclose, k.rclose);
    if (md.v == 0.0) icnvg = 1;
    k.icnvg = icnvg;
    if (icnvg != 0) { k.stop = 1; return; }
    if (is_close_b(k.rho, k.rho0)) { k.stop = 2; return; }
    if (k.rho == 0.0) { k.stop = 2; return; }
    k.rho0 = k.rho;
    if (it >= k.itmax) { k.stop = 3; return; }
  }
  k.rho = rho_new;
  if (it > 0) k.beta = rho_new / k.rho0;
  if (f.restart) {
    k.beta = 0.0;
    f.restart = 0;
  }
}

__global__ void __launch_bounds__(MF6_BLOCK) kb_alpha(const BSys *sys, KState *st, BFState *fs,
                                                      const double *part, const MaxLoc *mpart,
                                                      int T) {
  __shared__ double lds[8];
  __shared__ MaxLoc mlds[8];
  const int s = blockIdx.x;
  if (st[s].stop) return;
  const BSys S = sys[s];
  double a = 0.0;
  MaxLoc mm{0.0, -1};
  for (int i = threadIdx.x; i < S.nti; i += blockDim.x) {
    a = a + part[P_PQ * T + S.ti0 + i];
    mm = maxloc_pick(mm, mpart[M_P * T + S.ti0 + i]);
  }
  const double pq = bsum(a, lds);
  const MaxLoc mp = bmax(mm, mlds);
  if (threadIdx.x != 0) return;
  KState &k = st[s];
  BFState &f = fs[s];
  double den = pq;
  den = den + copysign(2.220446049250313080847263336181640625e-16, den);
  k.pq = pq;
  k.alpha = k.rho / den;
  k.iter = k.iter + 1;
  f.apend_x = k.alpha;
  f.apend_d = k.alpha;
  f.pmax = mp.v;
  f.ploc = mp.loc;
  f.pcur = 1 - f.pcur;   // the direction just formed is the current one
}

// ---------------------------------------------------------------------------
// residual replacement (every system, stopped or not)
// ---------------------------------------------------------------------------

// e += apend_x p_cur (the pending head update), x64 += e, e = 0
__global__ void kb_fold(const BSys *sys, const Tile *tiles, int ntiles, int cpt, const float *p0,
                        const float *p1, float *e, double *x, BFState *fs, const int *active) {
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int s = tiles[tile].sys;
    if (!active[s]) continue;
    const Layout L = sys[s].L;
    const long long vo = sys[s].voff;
    const float *p = (fs[s].pcur ? p1 : p0) + vo;
    const float ax = (float)fs[s].apend_x;
    for (int c = 0; c < cpt; c++) {
      const int t = tiles[tile].base + c * blockDim.x + threadIdx.x;
      const int cc = t & 1, h2 = t >> 1;
      if (h2 >= ncols_of(L, cc)) continue;
      const int own = cc * L.H + h2;
      for (int k = 0; k < L.nlay; k++) {
        const long long pos = vo + own + k * L.HP;
        float ev = e[pos];
        if (ax != 0.f) ev = ev + ax * p[own + k * L.HP];
        x[pos] = x[pos] + (double)ev;
        e[pos] = 0.f;
      }
    }
  }
}

// d32 = b - A64 x64 (FP64, amux order) ; partial max |r| per tile
__global__ void __launch_bounds__(MF6_BLOCK) kb_residual(const BSys *sys, const Tile *tiles,
                                                         int ntiles, int cpt, const double *A64,
                                                         const double *x, const double *b,
                                                         float *d32, MaxLoc *mpart, int T,
                                                         const int *active) {
  __shared__ MaxLoc mlds[8];
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int s = tiles[tile].sys;
    if (!active[s]) continue;
    const Layout L = sys[s].L;
    const long long vo = sys[s].voff;
    const double *A = A64 + sys[s].aoff;
    const double *xx = x + vo;
    const int NP = L.NP;
    MaxLoc m{0.0, -1};
    for (int c = 0; c < cpt; c++) {
      const int t = tiles[tile].base + c * blockDim.x + threadIdx.x;
      const int cc = t & 1, h2 = t >> 1;
      if (h2 >= ncols_of(L, cc)) continue;
      const int l = col_l(L, cc, h2);
      const int i = l / L.ncol, j = l - i * L.ncol;
      // off-grid neighbours read position 0 (in range); a select drops
      // their term, so no lane waits inside a branch
      const bool hb = i > 0, hl = j > 0, hr = j < L.ncol - 1, hf = i < L.nrow - 1;
      const int ib = hb ? hn(l, -L.ncol) : 0;
      const int il = hl ? hn(l, -1) : 0;
      const int ir = hr ? hn(l, 1) : 0;
      const int ifr = hf ? hn(l, L.ncol) : 0;
      const int own = cc * L.H + h2;
      const int oth = (1 - cc) * L.H;
      for (int k = 0; k < L.nlay; k++) {
        const int pos = own + k * L.HP;
        const int ob = oth + k * L.HP;
        double t2 = 0.0;
        t2 = t2 + A[pos] * xx[pos];
        if (k > 0) t2 = t2 + A[NP + pos] * xx[pos - L.HP];
        {
          const auto t_ = A[2 * NP + pos] * xx[ob + ib];
          t2 = hb ? t2 + t_ : t2;
        }
        {
          const auto t_ = A[3 * NP + pos] * xx[ob + il];
          t2 = hl ? t2 + t_ : t2;
        }
        {
          const auto t_ = A[4 * NP + pos] * xx[ob + ir];
          t2 = hr ? t2 + t_ : t2;
        }
        {
          const auto t_ = A[5 * NP + pos] * xx[ob + ifr];
          t2 = hf ? t2 + t_ : t2;
        }
        if (k < L.nlay - 1) t2 = t2 + A[6 * NP + pos] * xx[pos + L.HP];
        const double r = b[vo + pos] - t2;
        d32[vo + pos] = (float)r;
        m = maxloc_pick(m, MaxLoc{r, k * L.NRC + l});
      }
    }
    m = bmax(m, mlds);
    if (threadIdx.x == 0) mpart[M_TRUE * T + tile] = m;
  }
}

// per system: no pending update; the true max |r| into out[s]
__global__ void __launch_bounds__(MF6_BLOCK) kb_after(const BSys *sys, BFState *fs,
                                                      const MaxLoc *mpart, int T, double *out,
                                                      const int *active, int restart) {
  __shared__ MaxLoc mlds[8];
  const int s = blockIdx.x;
  if (!active[s]) return;
  const BSys S = sys[s];
  MaxLoc m{0.0, -1};
  for (int i = threadIdx.x; i < S.nti; i += blockDim.x)
    m = maxloc_pick(m, mpart[M_TRUE * T + S.ti0 + i]);
  m = bmax(m, mlds);
  if (threadIdx.x != 0) return;
  fs[s].apend_d = 0.0;
  fs[s].apend_x = 0.0;
  fs[s].restart = restart;
  out[s] = m.v;
}

// two-level: tasks of four coarse rows (aggregates) of one system, in
// system order; a system's tasks are tg0 .. tg0 + ntg - 1
struct GTask {
  int sys, row0;
};

// rc of every running system: one wavefront per aggregate sums the column
// sums of its ha x ha columns (fixed order)
__global__ void __launch_bounds__(MF6_BLOCK) kb_tl_gather(const BSys *sys, const GTask *tasks,
                                                          int ntasks, const double *colsum,
                                                          double *rc, const KState *st) {
  const int w = threadIdx.x >> 6, ln = threadIdx.x & 63;
  for (int task = blockIdx.x; task < ntasks; task += gridDim.x) {
    const int s = tasks[task].sys;
    if (st[s].stop) continue;
    const BSys &S = sys[s];
    const TwoLevel &Tl = S.tl;
    const int I = tasks[task].row0 + w;
    if (I >= Tl.nc) continue;
    const int ai = I / Tl.naj, aj = I - ai * Tl.naj;
    const double *cs = colsum + S.cso;
    double v = 0.0;
    for (int t = ln; t < Tl.ha * Tl.ha; t += 64) {
      const int i = ai * Tl.ha + t / Tl.ha, j = aj * Tl.ha + t % Tl.ha;
      if (i < S.L.nrow && j < S.L.ncol) v = v + cs[i * S.L.ncol + j];
    }
    v = wsum(v);
    if (ln == 0) rc[S.cvo + I] = v;
  }
}

// yc = Ainv rc for the task's four rows (one wavefront per row, 16-byte row
// loads, rc staged in LDS: ld doubles, zero beyond nc) and the task's
// partial of rc.yc
__global__ void __launch_bounds__(MF6_BLOCK) kb_tl_gemv(const BSys *sys, const GTask *tasks,
                                                        int ntasks, const double *rc,
                                                        const float *Ainv, float *yc,
                                                        double *part, int T, const KState *st) {
  extern __shared__ double xs[];
  __shared__ double lds[8];
  const int w = threadIdx.x >> 6, ln = threadIdx.x & 63;
  for (int task = blockIdx.x; task < ntasks; task += gridDim.x) {
    const int s = tasks[task].sys;
    if (st[s].stop) continue;
    const BSys &S = sys[s];
    const TwoLevel &Tl = S.tl;
    __syncthreads();   // the previous task's xs is no longer read
    for (int c = threadIdx.x; c < Tl.ld; c += blockDim.x)
      xs[c] = c < Tl.nc ? rc[S.cvo + c] : 0.0;
    __syncthreads();
    double acc = 0.0;
    const int I = tasks[task].row0 + w;
    if (I < Tl.nc) {
      const float *row = Ainv + S.aio + (size_t)I * Tl.ld;
      double s0 = 0.0;
      for (int c = 4 * ln; c < Tl.ld; c += 256) {
        const float4 a = *reinterpret_cast<const float4 *>(row + c);
        s0 = s0 + (double)a.x * xs[c] + (double)a.y * xs[c + 1] + (double)a.z * xs[c + 2] +
             (double)a.w * xs[c + 3];
      }
      s0 = wsum(s0);
      if (ln == 0) {
        yc[S.cvo + I] = (float)s0;
        acc = xs[I] * (double)(float)s0;
      }
    }
    acc = bsum(acc, lds);
    if (threadIdx.x == 0) part[P_RHO_C * T + task] = acc;
  }
}

__global__ void kb_convert(long long n, const double *a, float *b) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

struct BatchHandle {
  int nsys = 0, cap = 0;
  std::vector<BSys> sys;
  std::vector<mf6hip_ims_settings> set;
  long long ptot = 0, vtot = 0;        // plane / vector lengths
  double *A64 = nullptr, *b64 = nullptr, *x64 = nullptr, *x0 = nullptr, *piv64 = nullptr;
  float *A32 = nullptr, *piv32 = nullptr, *d32 = nullptr, *z32 = nullptr, *zf32 = nullptr,
        *q32 = nullptr, *p0 = nullptr, *p1 = nullptr, *e32 = nullptr;
  // the tile set in use (pointers into tsets[cpt - 1])
  BSys *dsys = nullptr;
  Tile *dtr = nullptr, *dtb = nullptr, *dti = nullptr;
  int ntr = 0, ntb = 0, nti = 0, T = 0;
  double *part = nullptr;
  MaxLoc *mpart = nullptr;
  // per columns-per-thread value (1, 2): its tile lists, the systems' tile
  // offsets, the partial arrays, and the captured iteration graphs per grid
  // size; built once, kept until destroy, so a graph stays valid for reuse
  struct TileSet {
    bool built = false;
    BSys *dsys = nullptr;
    Tile *dtr = nullptr, *dtb = nullptr, *dti = nullptr;
    int ntr = 0, ntb = 0, nti = 0, T = 0;
    double *part = nullptr;
    MaxLoc *mpart = nullptr;
    std::vector<int> grids;
    std::vector<hipGraphExec_t> execs;
  } tsets[2];
  int *h_act = nullptr;                // pinned: active flags staged to the device
  KState *st = nullptr;
  BFState *fs = nullptr;
  int *active = nullptr;
  double *rtrue = nullptr;
  KState *h_st = nullptr;
  double *h_rtrue = nullptr;
  int cpt = 1, grid = 0, R = 32;
  // two-level preconditioner (every system, or none)
  bool tl = false;
  long long cmtot = 0, cstot = 0, cvtot = 0, aitot = 0;
  int maxnc = 0, maxld = 0;
  unsigned *tl_cmask = nullptr;
  double *tl_colsum = nullptr, *tl_rc = nullptr;
  float *tl_yc = nullptr, *tl_A32 = nullptr;
  double *tl_S3 = nullptr, *tl_Lc = nullptr, *tl_Lr = nullptr, *tl_G = nullptr;
  int *tl_dinfo = nullptr;
  GTask *dtg = nullptr;
  int ntg = 0;
  double tl_setup_ms = 0.0;   // coarse setups at staging, summed
  hipStream_t stream = nullptr;
  hipGraphExec_t gexec = nullptr;      // the graph of the selected (cpt, grid)
  bool finalized = false;
  double last_ms = 0.0;
  std::vector<int> iters, icnvg;
  std::string tune;
};

namespace {

// every tile set and captured graph (destroy only)
void batch_free_tilesets(BatchHandle *b) {
  for (auto &t : b->tsets) {
    for (hipGraphExec_t e : t.execs)
      if (e) hipGraphExecDestroy(e);
    void *bufs[] = {t.dsys, t.dtr, t.dtb, t.dti, t.part, t.mpart};
    for (void *p : bufs)
      if (p) hipFree(p);
    t = BatchHandle::TileSet{};
  }
  b->gexec = nullptr;
}

template <typename T>
int balloc(T **p, size_t n) {
  if (hipMalloc((void **)p, sizeof(T) * (n ? n : 1)) != hipSuccess) {
    mf6hip_set_error("hipMalloc failed (ensemble)");
    return MF6HIP_EHIP;
  }
  if (hipMemset(*p, 0, sizeof(T) * (n ? n : 1)) != hipSuccess) {
    mf6hip_set_error("hipMemset failed (ensemble)");
    return MF6HIP_EHIP;
  }
  return MF6HIP_OK;
}

// tile lists for a columns-per-thread value: built on the first request,
// then only selected (the device arrays and the graphs captured with them
// stay valid)
int build_tiles(BatchHandle *b, int cpt) {
  BatchHandle::TileSet &ts = b->tsets[cpt - 1];
  if (!ts.built) {
    std::vector<Tile> tr, tb, ti;
    std::vector<BSys> sys = b->sys;
    const int w = MF6_BLOCK * cpt;
    for (int s = 0; s < b->nsys; s++) {
      BSys &S = sys[s];
      S.tr0 = (int)tr.size();
      for (int h = 0; h < S.L.HP0; h += w) tr.push_back(Tile{s, h});
      S.ntr = (int)tr.size() - S.tr0;
      S.tb0 = (int)tb.size();
      for (int h = 0; h < S.L.HP1; h += w) tb.push_back(Tile{s, h});
      S.ntb = (int)tb.size() - S.tb0;
      S.ti0 = (int)ti.size();
      for (int h = 0; h < 2 * S.L.HP0; h += w) ti.push_back(Tile{s, h});
      S.nti = (int)ti.size() - S.ti0;
    }
    int T = (int)tr.size();
    if ((int)tb.size() > T) T = (int)tb.size();
    if ((int)ti.size() > T) T = (int)ti.size();
    if (b->ntg > T) T = b->ntg;
    BCK(balloc(&ts.dsys, sys.size()));
    BCK(balloc(&ts.dtr, tr.size()));
    BCK(balloc(&ts.dtb, tb.size()));
    BCK(balloc(&ts.dti, ti.size()));
    BCK(balloc(&ts.part, 6 * (size_t)T));
    BCK(balloc(&ts.mpart, 4 * (size_t)T));
    HIPCHK(hipMemcpy(ts.dtr, tr.data(), sizeof(Tile) * tr.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ts.dtb, tb.data(), sizeof(Tile) * tb.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ts.dti, ti.data(), sizeof(Tile) * ti.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ts.dsys, sys.data(), sizeof(BSys) * sys.size(), hipMemcpyHostToDevice));
    ts.ntr = (int)tr.size();
    ts.ntb = (int)tb.size();
    ts.nti = (int)ti.size();
    ts.T = T;
    ts.built = true;
  }
  b->dsys = ts.dsys;
  b->dtr = ts.dtr;
  b->dtb = ts.dtb;
  b->dti = ts.dti;
  b->ntr = ts.ntr;
  b->ntb = ts.ntb;
  b->nti = ts.nti;
  b->T = ts.T;
  b->part = ts.part;
  b->mpart = ts.mpart;
  b->cpt = cpt;
  b->gexec = nullptr;
  for (size_t g = 0; g < ts.grids.size(); g++)
    if (ts.grids[g] == b->grid) b->gexec = ts.execs[g];
  return MF6HIP_OK;
}

// one CG iteration for every running system
void launch_iteration(BatchHandle *b) {
  hipStream_t s = b->stream;
  const int g = b->grid, T = b->T, c = b->cpt;
  kb_red_fwd<<<g, MF6_BLOCK, 0, s>>>(b->dsys, b->dtr, b->ntr, c, b->A32, b->piv32, b->d32, b->q32,
                                     b->zf32, b->part, b->mpart, T, b->st, b->fs);
  double *cs = b->tl ? b->tl_colsum : nullptr;
  kb_black<<<g, MF6_BLOCK, 0, s>>>(b->dsys, b->dtb, b->ntb, c, b->A32, b->piv32, b->d32, b->q32,
                                   b->zf32, b->z32, b->part, b->mpart, T, b->st, b->fs,
                                   b->tl_cmask, cs);
  kb_red_bwd<<<g, MF6_BLOCK, 0, s>>>(b->dsys, b->dtr, b->ntr, c, b->A32, b->piv32, b->d32,
                                     b->zf32, b->z32, b->part, T, b->st, b->tl_cmask, cs);
  if (b->tl) {
    kb_tl_gather<<<g, MF6_BLOCK, 0, s>>>(b->dsys, b->dtg, b->ntg, b->tl_colsum, b->tl_rc, b->st);
    kb_tl_gemv<<<g, MF6_BLOCK, sizeof(double) * (size_t)b->maxld, s>>>(
        b->dsys, b->dtg, b->ntg, b->tl_rc, b->tl_A32, b->tl_yc, b->part, T, b->st);
  }
  kb_round<<<b->nsys, MF6_BLOCK, 0, s>>>(b->dsys, b->st, b->fs, b->part, b->mpart, T);
  kb_pupdate<<<g, MF6_BLOCK, 0, s>>>(b->dsys, b->dti, b->nti, c, b->z32, b->p0, b->p1, b->e32,
                                     b->mpart, T, b->st, b->fs, b->tl_cmask,
                                     b->tl ? b->tl_yc : nullptr);
  kb_spmv<<<g, MF6_BLOCK, 0, s>>>(b->dsys, b->dti, b->nti, c, b->A32, b->p0, b->p1, b->q32,
                                  b->part, T, b->st, b->fs);
  kb_alpha<<<b->nsys, MF6_BLOCK, 0, s>>>(b->dsys, b->st, b->fs, b->part, b->mpart, T);
}

// fold the FP32 correction into the FP64 heads and replace the residual
void launch_replacement(BatchHandle *b) {
  hipStream_t s = b->stream;
  const int g = b->grid;
  kb_fold<<<g, MF6_BLOCK, 0, s>>>(b->dsys, b->dti, b->nti, b->cpt, b->p0, b->p1, b->e32, b->x64,
                                  b->fs, b->active);
  kb_residual<<<g, MF6_BLOCK, 0, s>>>(b->dsys, b->dti, b->nti, b->cpt, b->A64, b->x64, b->b64,
                                      b->d32, b->mpart, b->T, b->active);
  // with the coarse correction the direction restarts at each replacement
  kb_after<<<b->nsys, MF6_BLOCK, 0, s>>>(b->dsys, b->fs, b->mpart, b->T, b->rtrue, b->active,
                                         b->tl ? 1 : 0);
}

// The graph of the selected tile set and grid: captured (kernel launches
// only) and instantiated on the first request, then reused.
int capture(BatchHandle *b) {
  BatchHandle::TileSet &ts = b->tsets[b->cpt - 1];
  for (size_t g = 0; g < ts.grids.size(); g++)
    if (ts.grids[g] == b->grid) {
      b->gexec = ts.execs[g];
      return MF6HIP_OK;
    }
  hipGraph_t graph = nullptr;
  HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  for (int i = 0; i < b->R; i++) launch_iteration(b);
  launch_replacement(b);
  hipError_t e = hipStreamEndCapture(b->stream, &graph);
  if (e != hipSuccess || !graph) { mf6hip_set_error("graph capture failed"); return MF6HIP_EHIP; }
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  if (e != hipSuccess) { mf6hip_set_error("graph instantiation failed"); return MF6HIP_EHIP; }
  ts.grids.push_back(b->grid);
  ts.execs.push_back(exec);
  b->gexec = exec;
  return MF6HIP_OK;
}

// initial state of every system: true residual, KState from the settings
int reset_states(BatchHandle *b) {
  for (int s = 0; s < b->nsys; s++) {
    KState &k = b->h_st[s];
    memset(&k, 0, sizeof(KState));
    k.dvclose = b->set[s].dvclose;
    k.rclose = b->set[s].rclose;
    k.itmax = b->set[s].iter1;
    k.icnvgopt = 0;
    k.epfact = 1.0;
  }
  HIPCHK(hipMemcpyAsync(b->st, b->h_st, sizeof(KState) * b->nsys, hipMemcpyHostToDevice, b->stream));
  HIPCHK(hipMemsetAsync(b->fs, 0, sizeof(BFState) * b->nsys, b->stream));
  HIPCHK(hipMemsetAsync(b->q32, 0, sizeof(float) * b->vtot, b->stream));
  HIPCHK(hipMemsetAsync(b->e32, 0, sizeof(float) * b->vtot, b->stream));
  for (int s = 0; s < b->nsys; s++) b->h_act[s] = 1;
  HIPCHK(hipMemcpyAsync(b->active, b->h_act, sizeof(int) * b->nsys, hipMemcpyHostToDevice, b->stream));
  launch_replacement(b);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(b->stream));
  return MF6HIP_OK;
}

}  // namespace

// Create and add touch no device (the stream is created by finalize), so
// registration and the host checks of mf6hip_batch_set_system also run
// where no GPU is present.
extern "C" int mf6hip_batch_create(int capacity, void **handle) {
  if (!handle || capacity < 1) { mf6hip_set_error("bad argument"); return MF6HIP_EINVAL; }
  BatchHandle *b = new BatchHandle();
  b->cap = capacity;
  *handle = b;
  return MF6HIP_OK;
}

// Register a system (shape + settings); storage is allocated by finalize.
extern "C" int mf6hip_batch_add(void *handle, const mf6hip_ims_settings *settings) {
  BatchHandle *b = (BatchHandle *)handle;
  if (!b || !settings) { mf6hip_set_error("null argument"); return MF6HIP_EINVAL; }
  if (b->finalized) { mf6hip_set_error("ensemble already finalized"); return MF6HIP_EINVAL; }
  if (b->nsys >= b->cap) { mf6hip_set_error("ensemble capacity exceeded"); return MF6HIP_EINVAL; }
  const mf6hip_ims_settings &st = *settings;
  if (st.ilinmeth != 1 || st.iscl || st.iord || st.level > 0 || st.droptol > 0.0 ||
      st.icnvgopt != 0 || st.north != 0) {
    mf6hip_set_error("ensemble systems must use CG + ILU0/MILU0 with the default convergence test");
    return MF6HIP_EUNSUPPORTED;
  }
  if (st.nlay < 1 || st.nlay > 32 || st.nrow < 1 || st.ncol < 1 ||
      (long long)st.nlay * st.nrow * st.ncol > (1LL << 28)) {
    mf6hip_set_error("bad DIS shape");
    return MF6HIP_EINVAL;
  }
  if (st.preconditioner != 0 && st.preconditioner != 2) {
    mf6hip_set_error("ensemble systems use preconditioner 0 (ILU0) or 2 (two-level)");
    return MF6HIP_EUNSUPPORTED;
  }
  if (b->nsys > 0 && (st.preconditioner == 2) != b->tl) {
    mf6hip_set_error("ensemble systems must all use the same preconditioner");
    return MF6HIP_EINVAL;
  }
  if (st.coarse_size < 0 || st.coarse_size > 6144) {
    mf6hip_set_error("coarse_size must be in [0, 6144]");
    return MF6HIP_EINVAL;
  }
  BSys S{};
  S.L = make_layout(st.nlay, st.nrow, st.ncol);
  S.aoff = b->ptot;
  S.voff = b->vtot;
  b->tl = st.preconditioner == 2;
  if (b->tl) {
    S.tl = mf6hip_tl_shape(S.L, st.coarse_size);
    if (S.tl.naj > 63) {
      mf6hip_set_error("ensemble two-level: the aggregate grid is wider than 63 (raise coarse_size's aggregates)");
      return MF6HIP_EUNSUPPORTED;
    }
    S.cmo = b->cmtot;
    S.cso = b->cstot;
    S.cvo = b->cvtot;
    S.aio = b->aitot;
    b->cmtot += 2 * (long long)S.L.HP;
    b->cstot += S.L.NRC;
    b->cvtot += S.tl.ld;
    b->aitot += (long long)S.tl.nc * S.tl.ld;
    if (S.tl.nc > b->maxnc) b->maxnc = S.tl.nc;
    if (S.tl.ld > b->maxld) b->maxld = S.tl.ld;
  }
  b->ptot += (long long)MF6_NSLOT * S.L.NP;
  b->vtot += S.L.NP;
  b->sys.push_back(S);
  b->set.push_back(st);
  b->nsys++;
  return MF6HIP_OK;
}

extern "C" int mf6hip_batch_finalize(void *handle) {
  BatchHandle *b = (BatchHandle *)handle;
  if (!b || b->nsys < 1) { mf6hip_set_error("empty ensemble"); return MF6HIP_EINVAL; }
  if (b->finalized) return MF6HIP_OK;
  if (!b->stream && hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    b->stream = nullptr;
    mf6hip_set_error("hipStreamCreate failed");
    return MF6HIP_EHIP;
  }
  BCK(balloc(&b->A64, b->ptot));
  BCK(balloc(&b->A32, b->ptot));
  const size_t v = b->vtot;
  BCK(balloc(&b->b64, v));
  BCK(balloc(&b->x64, v));
  BCK(balloc(&b->x0, v));
  BCK(balloc(&b->piv64, v));
  BCK(balloc(&b->piv32, v));
  BCK(balloc(&b->d32, v));
  BCK(balloc(&b->z32, v));
  BCK(balloc(&b->zf32, v));
  BCK(balloc(&b->q32, v));
  BCK(balloc(&b->p0, v));
  BCK(balloc(&b->p1, v));
  BCK(balloc(&b->e32, v));
  BCK(balloc(&b->dsys, b->nsys));
  BCK(balloc(&b->st, b->nsys));
  BCK(balloc(&b->fs, b->nsys));
  BCK(balloc(&b->active, b->nsys));
  BCK(balloc(&b->rtrue, b->nsys));
  HIPCHK(hipHostMalloc((void **)&b->h_st, sizeof(KState) * b->nsys, hipHostMallocDefault));
  HIPCHK(hipHostMalloc((void **)&b->h_rtrue, sizeof(double) * b->nsys, hipHostMallocDefault));
  HIPCHK(hipHostMalloc((void **)&b->h_act, sizeof(int) * b->nsys, hipHostMallocDefault));
  if (b->tl) {
    std::vector<GTask> tg;
    for (int s = 0; s < b->nsys; s++) {
      BSys &S = b->sys[s];
      S.tg0 = (int)tg.size();
      for (int r = 0; r < S.tl.nc; r += 4) tg.push_back(GTask{s, r});
      S.ntg = (int)tg.size() - S.tg0;
    }
    b->ntg = (int)tg.size();
    BCK(balloc(&b->dtg, tg.size()));
    HIPCHK(hipMemcpy(b->dtg, tg.data(), sizeof(GTask) * tg.size(), hipMemcpyHostToDevice));
    BCK(balloc(&b->tl_cmask, b->cmtot));
    BCK(balloc(&b->tl_colsum, b->cstot));
    BCK(balloc(&b->tl_rc, b->cvtot));
    BCK(balloc(&b->tl_yc, b->cvtot));
    BCK(balloc(&b->tl_A32, b->aitot));
    BCK(balloc(&b->tl_S3, 3 * (size_t)b->maxnc));
    BCK(balloc(&b->tl_Lc, 64 * (size_t)b->maxnc));
    BCK(balloc(&b->tl_Lr, 64 * (size_t)b->maxnc));
    BCK(balloc(&b->tl_G, (size_t)b->maxnc * b->maxld));
    BCK(balloc(&b->tl_dinfo, 1));
  }
  BCK(build_tiles(b, 1));
  int ncu = 256;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, 0) == hipSuccess && prop.multiProcessorCount > 0)
    ncu = prop.multiProcessorCount;
  b->grid = 8 * ncu;
  b->finalized = true;
  return MF6HIP_OK;
}

// System data in the reference layout (host): 1-based DIS CSR, amat after
// the sln_ls row fixes, rhs, start heads.  ia / ja must be the DIS 7-point
// structure of the system's shape and amat symmetric (the batched CG reads
// every coupling from the upper planes); both are checked on the host
// before anything is staged.
extern "C" int mf6hip_batch_set_system(void *handle, int isys, const int *ia, const int *ja,
                                       int nja, const double *amat, const double *rhs,
                                       const double *x0) {
  BatchHandle *b = (BatchHandle *)handle;
  if (!b || !ia || !ja || !amat || !rhs || !x0) { mf6hip_set_error("bad argument"); return MF6HIP_EINVAL; }
  if (isys < 0 || isys >= b->nsys) { mf6hip_set_error("system index out of range"); return MF6HIP_EINVAL; }
  const BSys &S = b->sys[isys];
  BCK(mf6hip_check_structure(S.L, S.L.N, nja, ia, ja));
  BCK(mf6hip_check_symmetric_csr(S.L, ia, amat));
  if (!b->finalized) { mf6hip_set_error("ensemble not finalized"); return MF6HIP_EINVAL; }
  BCK(mf6hip_stage_system(S.L, ia, nja, amat, rhs, x0, b->set[isys].relax, b->A64 + S.aoff,
                          b->b64 + S.voff, b->x64 + S.voff, b->piv64 + S.voff, b->stream));
  HIPCHK(hipMemcpyAsync(b->x0 + S.voff, b->x64 + S.voff, sizeof(double) * S.L.NP,
                        hipMemcpyDeviceToDevice, b->stream));
  const long long np = (long long)MF6_NSLOT * S.L.NP;
  kb_convert<<<1024, 256, 0, b->stream>>>(np, b->A64 + S.aoff, b->A32 + S.aoff);
  kb_convert<<<256, 256, 0, b->stream>>>(S.L.NP, b->piv64 + S.voff, b->piv32 + S.voff);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(b->stream));
  if (b->tl) {
    // the coarse operator of this system, with the ILU0 factor (staging)
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    HIPCHK(hipEventRecord(e0, b->stream));
    int info = 0;
    BCK(mf6hip_tl_setup_slice(S.L, b->A64 + S.aoff, S.tl, b->tl_cmask + S.cmo, b->tl_S3,
                              b->tl_Lc, b->tl_Lr, b->tl_G, b->tl_A32 + S.aio, b->tl_dinfo,
                              &info, b->stream));
    HIPCHK(hipEventRecord(e1, b->stream));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    b->tl_setup_ms += ms;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (info != 0) {
      mf6hip_set_error("ensemble two-level: the coarse matrix of system " + std::to_string(isys) +
                       " is not definite");
      return MF6HIP_EINVAL;
    }
  }
  return MF6HIP_OK;
}

// Autotune (columns per thread, persistent grid) on the loaded systems:
// time one replay of R iterations per candidate from the start state.
extern "C" int mf6hip_batch_autotune(void *handle, char *report, int report_len) {
  BatchHandle *b = (BatchHandle *)handle;
  if (!b || !b->finalized) { mf6hip_set_error("ensemble not finalized"); return MF6HIP_EINVAL; }
  int ncu = 256;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, 0) == hipSuccess && prop.multiProcessorCount > 0)
    ncu = prop.multiProcessorCount;
  // the trial replays start from the staged start heads
  const double *xsave = b->x0;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  const int cpts[2] = {1, 2};
  const int gmul[3] = {4, 8, 16};
  float best = 1e30f;
  int bc = 1, bg = 8 * ncu;
  std::string rep;
  for (int ci = 0; ci < 2; ci++) {
    BCK(build_tiles(b, cpts[ci]));
    for (int gi = 0; gi < 3; gi++) {
      b->grid = gmul[gi] * ncu;
      HIPCHK(hipMemcpy(b->x64, xsave, sizeof(double) * b->vtot, hipMemcpyDeviceToDevice));
      BCK(reset_states(b));
      BCK(capture(b));
      HIPCHK(hipGraphLaunch(b->gexec, b->stream));   // warm
      HIPCHK(hipStreamSynchronize(b->stream));
      HIPCHK(hipMemcpy(b->x64, xsave, sizeof(double) * b->vtot, hipMemcpyDeviceToDevice));
      BCK(reset_states(b));
      HIPCHK(hipEventRecord(e0, b->stream));
      HIPCHK(hipGraphLaunch(b->gexec, b->stream));
      HIPCHK(hipEventRecord(e1, b->stream));
      HIPCHK(hipEventSynchronize(e1));
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, e0, e1));
      rep += "cpt " + std::to_string(cpts[ci]) + " grid " + std::to_string(b->grid) + ": " +
             std::to_string(ms) + " ms; ";
      if (ms < best) { best = ms; bc = cpts[ci]; bg = b->grid; }
    }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  b->grid = bg;
  BCK(build_tiles(b, bc));
  HIPCHK(hipMemcpy(b->x64, xsave, sizeof(double) * b->vtot, hipMemcpyDeviceToDevice));
  rep += "chosen cpt " + std::to_string(bc) + " grid " + std::to_string(bg);
  if (b->tl) rep += "; coarse setups at staging " + std::to_string(b->tl_setup_ms) + " ms";
  b->tune = rep;
  if (report && report_len > 0) {
    strncpy(report, rep.c_str(), report_len - 1);
    report[report_len - 1] = 0;
  }
  return MF6HIP_OK;
}

// Solve every system: replays of the captured graph (R iterations + the
// FP64 residual replacement) until each system has converged (FP32 test
// passed and the FP64 true residual within RCLOSE) or reached INNER_MAXIMUM.
extern "C" int mf6hip_batch_solve(void *handle, int *icnvg, int *iters, double *solve_ms) {
  BatchHandle *b = (BatchHandle *)handle;
  if (!b || !b->finalized || !icnvg || !iters) { mf6hip_set_error("bad argument"); return MF6HIP_EINVAL; }
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, b->stream));
  // every solve starts from the staged start heads
  HIPCHK(hipMemcpyAsync(b->x64, b->x0, sizeof(double) * b->vtot, hipMemcpyDeviceToDevice,
                        b->stream));
  BCK(reset_states(b));
  BCK(capture(b));
  std::vector<int> done(b->nsys, 0);
  int *act = b->h_act;   // pinned; all 1 after reset_states
  for (int s = 0; s < b->nsys; s++) { icnvg[s] = 0; iters[s] = 0; }
  int ndone = 0, replays = 0;
  const int max_replays = 1000000;
  while (ndone < b->nsys && replays < max_replays) {
    HIPCHK(hipGraphLaunch(b->gexec, b->stream));
    replays++;
    HIPCHK(hipMemcpyAsync(b->h_st, b->st, sizeof(KState) * b->nsys, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipMemcpyAsync(b->h_rtrue, b->rtrue, sizeof(double) * b->nsys, hipMemcpyDeviceToHost, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    bool changed = false;
    for (int s = 0; s < b->nsys; s++) {
      if (done[s]) continue;
      KState &k = b->h_st[s];
      const bool true_ok = fabs(b->h_rtrue[s]) <= b->set[s].rclose;
      if (k.stop == 0 && k.iter < k.itmax) continue;
      if (k.stop == 1 && true_ok) { icnvg[s] = 1; done[s] = 1; }
      else if (k.stop == 2 && true_ok && fabs(k.deltax) <= b->set[s].dvclose) { icnvg[s] = 1; done[s] = 1; }
      else if (k.stop == 2 || k.stop == 3 || k.iter >= k.itmax) { done[s] = 1; }
      else {
        // FP32 test passed, the true residual not: continue from it
        k.stop = 0;
        k.icnvg = 0;
        changed = true;
      }
      if (done[s]) {
        iters[s] = k.iter;
        ndone++;
        act[s] = 0;
        k.stop = k.stop ? k.stop : 3;
        changed = true;
      }
    }
    if (changed) {
      HIPCHK(hipMemcpyAsync(b->st, b->h_st, sizeof(KState) * b->nsys, hipMemcpyHostToDevice, b->stream));
      HIPCHK(hipMemcpyAsync(b->active, act, sizeof(int) * b->nsys, hipMemcpyHostToDevice, b->stream));
    }
  }
  for (int s = 0; s < b->nsys; s++)
    if (!done[s]) iters[s] = b->h_st[s].iter;
  HIPCHK(hipEventRecord(e1, b->stream));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  b->last_ms = ms;
  if (solve_ms) *solve_ms = ms;
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  return MF6HIP_OK;
}

// Heads of one system (natural order, host).
extern "C" int mf6hip_batch_get_x(void *handle, int isys, double *x) {
  BatchHandle *b = (BatchHandle *)handle;
  if (!b || !b->finalized || !x) { mf6hip_set_error("bad argument"); return MF6HIP_EINVAL; }
  if (isys < 0 || isys >= b->nsys) { mf6hip_set_error("system index out of range"); return MF6HIP_EINVAL; }
  const Layout &L = b->sys[isys].L;
  std::vector<double> tmp(L.NP);
  HIPCHK(hipMemcpy(tmp.data(), b->x64 + b->sys[isys].voff, sizeof(double) * L.NP, hipMemcpyDeviceToHost));
  for (int n = 0; n < L.N; n++) x[n] = tmp[pos_of_natural(L, n)];
  return MF6HIP_OK;
}

extern "C" int mf6hip_batch_destroy(void *handle) {
  BatchHandle *b = (BatchHandle *)handle;
  if (!b) return MF6HIP_OK;
  if (b->stream) hipStreamSynchronize(b->stream);
  batch_free_tilesets(b);
  void *bufs[] = {b->A64, b->b64, b->x64, b->x0, b->piv64, b->A32, b->piv32, b->d32, b->z32, b->zf32,
                  b->q32, b->p0, b->p1, b->e32, b->st, b->fs, b->active, b->rtrue, b->dtg, b->tl_cmask,
                  b->tl_colsum, b->tl_rc, b->tl_yc, b->tl_A32, b->tl_S3, b->tl_Lc, b->tl_Lr,
                  b->tl_G, b->tl_dinfo};
  for (void *p : bufs)
    if (p) hipFree(p);
  if (b->h_st) hipHostFree(b->h_st);
  if (b->h_rtrue) hipHostFree(b->h_rtrue);
  if (b->h_act) hipHostFree(b->h_act);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
  return MF6HIP_OK;
}
