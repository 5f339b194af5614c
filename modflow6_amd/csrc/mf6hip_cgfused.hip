// mf6hip_cgfused.hip -- fused preconditioned CG (ims_base_cg,
// src/Solution/LinearMethods/ImsLinearBase.f90:30-240) for symmetric
// systems, one thread per vertical cell column.
//
// One CG iteration is five streaming kernels and two one-block scalar
// kernels instead of the reference's nine vector passes:
//
//   pc_red_fwd  d -= alpha q (red cells) ; forward ILU0 sweep of red columns
//   pc_black    d -= alpha q (black)     ; forward + backward sweep of black
//               columns ; partial rho = d.z
//   pc_red_bwd  backward sweep of red columns ; partial rho
//   s_round     convergence test of the finished iteration (testcnvg,
//               is_close, rho == 0, INNER_MAXIMUM), rho, beta
//   pupdate     p = z + beta p ; x += alpha_prev p_prev ; partial max |p|
//   spmv        q = A p ; partial p.q
//   s_alpha     alpha = rho / (p.q + sign(DPREC, p.q)) ; deltax = alpha pmax
//
// The arithmetic of every vector element is the reference's (x = x + alpha
// p, d = d - alpha q, p = z + beta p, amux order, ilu0a order); only the
// point in time at which x and d receive their update moves, and dot
// products are summed in a different (fixed) order.  Symmetry lets every
// off-diagonal be read from the upper planes (right, front, down) of the
// owning cell, so the lower planes are never streamed.
//
// The same kernels run in FP32 (matrix, pivots, Krylov vectors) for the
// mixed-precision solve: FP32 CG with reliable updates -- every R iterations
// and whenever the FP32 test passes, the FP32 correction is added to the
// FP64 heads and the FP32 residual is replaced by the FP64 true residual
// b - A x, keeping the search direction (residual replacement, so the
// Krylov space is not restarted).
#include <hip/hip_ext.h>
#include <hip/hip_fp16.h>

#include <cmath>
#include <cstring>
#include <string>

#include "mf6hip_handle.h"

#define DPREC_F 2.220446049250313080847263336181640625e-16
#define DSAME_F (100.0 * DPREC_F)

#define CK(x)                         \
  do {                                \
    int rc_ = (x);                    \
    if (rc_ != MF6HIP_OK) return rc_; \
  } while (0)

namespace {

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

__device__ inline double bsum(double v, double *lds) {
  v = wsum(v);
  int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) lds[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) s = s + lds[i];
  return s;
}

__device__ inline MaxLoc bmax(MaxLoc a, MaxLoc *lds) {
  a = wmax(a);
  int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) lds[w] = a;
  __syncthreads();
  MaxLoc r{0.0, -1};
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) r = maxloc_pick(r, lds[i]);
  return r;
}

__device__ inline double rsum(const double *part, int nb, double *lds) {
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) s = s + part[i];
  return bsum(s, lds);
}

__device__ inline MaxLoc rmaxl(const MaxLoc *part, int nb, MaxLoc *lds) {
  MaxLoc a{0.0, -1};
  for (int i = threadIdx.x; i < nb; i += blockDim.x) a = maxloc_pick(a, part[i]);
  return bmax(a, lds);
}

__device__ inline int is_close_(double a, double b) {
  if (a == b) return 1;
  return fabs(a - b) <= fmax(DSAME_F * fmax(fabs(a), fabs(b)), 0.0);
}

template <typename TP>
__device__ inline double ldp(const TP *p, int e) { return (double)p[e]; }
template <>
__device__ inline double ldp<unsigned short>(const unsigned short *p, int e) {
  unsigned short u = p[e];
  __half h = *reinterpret_cast<__half *>(&u);
  return (double)__half2float(h);
}

// fused-CG state beyond KState (kept in the same struct array: index 2)
struct FState {
  double apend_x;   // alpha of the x update still to apply
  double apend_d;   // alpha of the d update still to apply
  double pmax;      // signed max |p| of the last direction
  int ploc;
  int reliable_due;
  int restart;      // the next round starts a new direction (beta = 0)
};

// W consecutive values of one array, loaded and stored as one access
// (16 bytes for 4 FP32 or 2 FP64 values)
template <typename T, int W>
struct alignas(sizeof(T) * W) VecW {
  T v[W];
};

template <int W, typename T>
__device__ inline VecW<T, W> ldw(const T *p) {
  return *reinterpret_cast<const VecW<T, W> *>(p);
}
template <int W, typename T>
__device__ inline void stw(T *p, const VecW<T, W> &v) {
  *reinterpret_cast<VecW<T, W> *>(p) = v;
}

}  // namespace

// partial slots in part[] / mpart[] (each nblk long)
enum { P_SS_R = 0, P_SS_B = 1, P_RHO_B = 2, P_RHO_R = 3, P_PQ = 4, P_RHO_C = 5 };
enum { M_D_R = 0, M_D_B = 1, M_P = 2 };

// ---------------------------------------------------------------------------
// precondition kernels
// ---------------------------------------------------------------------------

// Deflated two-level preconditioner (preconditioner = 3): the sweeps run on
// r - A v, v = P yc (yc the coarse correction of this iteration's residual).
// For a coupled cell n of aggregate c,
//   (A v)_n = s_n yc_c + sum over coupled horizontal neighbours m in another
//             aggregate c' of A(n, m) (yc_c' - yc_c),
// s_n = A(n, n) + the couplings of n to coupled cells.  Decoupled cells have
// s = 0 and no stored edge couplings, so (A v)_n = 0 there, as in A v.
// ew: couplings across the aggregate edges in j, [(k nrow + i) 2 naj + 2 aj
// + e] (e = 0: to j - 1 at j = aj ha; e = 1: to j + 1 at j = aj ha + ha - 1);
// ns: across the edges in i, [(k 2 nai + 2 ai + e) ncol + jc], jc the
// column's place among its row's cells of its colour (colour 0 first).
struct DeflPlanes {
  const float *s = nullptr, *ew = nullptr, *ns = nullptr;
};

// Branch-free: a column off an edge reads a slot of its own aggregate row
// (always in range) and weights it by a zero difference, so every lane
// issues the same loads and none waits inside a divergent branch.  With
// aggregates at least two columns wide a column is on at most one edge per
// direction: one load per direction, slot and yc difference picked per
// column.  One-column aggregates (every column on both edges) add the second
// slot behind a grid-uniform flag.
template <typename TV>
struct DeflCol {
  TV y, dj, di, dj2, di2;   // yc of the column's aggregate; neighbour yc differences
  int ej, ni;               // offsets into ew / ns at layer 0 of the picked slots
  int ew0, ns0;             // offsets of the first slots (second pass, ha = 1)
  bool two;                 // ha = 1

  __device__ DeflCol() {}
  __device__ DeflCol(const TV *__restrict__ yc, const TwoLevel &T, const Layout &L, int i,
                     int j) {
    const int ai = i / T.ha, aj = j / T.ha, c = ai * T.naj + aj;
    const int ri = i - ai * T.ha, rj = j - aj * T.ha;
    const bool w = rj == 0 && j > 0, e = rj == T.ha - 1 && j + 1 < L.ncol;
    const bool n = ri == 0 && i > 0, so = ri == T.ha - 1 && i + 1 < L.nrow;
    two = T.ha == 1;
    y = yc[c];
    ew0 = i * 2 * T.naj + 2 * aj;
    // ns rows hold a grid row's colour-0 columns, then its colour-1 columns
    const int jc = ((i + j) & 1) == 0 ? (j >> 1)
                                      : ((i & 1) ? (L.ncol >> 1) : ((L.ncol + 1) >> 1)) + (j >> 1);
    ns0 = 2 * ai * L.ncol + jc;
    // w, else e (the first pass); e again when both (ha = 1, second pass)
    dj = yc[w ? c - 1 : (e ? c + 1 : c)] - y;
    ej = ew0 + (w ? 0 : 1);
    dj2 = yc[(w && e) ? c + 1 : c] - y;
    di = yc[n ? c - T.naj : (so ? c + T.naj : c)] - y;
    ni = ns0 + (n ? 0 : L.ncol);
    di2 = yc[(n && so) ? c + T.naj : c] - y;
  }
  // (A v) at layer k, cell position pos
  __device__ TV av(const float *__restrict__ sp, const float *__restrict__ ewp,
                   const float *__restrict__ nsp, const TwoLevel &T, const Layout &L, int k,
                   int pos) const {
    return av_s(sp[pos], ewp, nsp, T, L, k);
  }
  // the same with the cell's row sum s already loaded
  __device__ TV av_s(float sv, const float *__restrict__ ewp, const float *__restrict__ nsp,
                     const TwoLevel &T, const Layout &L, int k) const {
    return av_e(sv, edges(ewp, nsp, T, L, k));
  }
  // the edge couplings of layer k (loaded ahead of their use by the
  // chunked sweeps), and (A v) from them
  struct Edge {
    float ew, ns, ew2, ns2;
  };
  __device__ Edge edges(const float *__restrict__ ewp, const float *__restrict__ nsp,
                        const TwoLevel &T, const Layout &L, int k) const {
    const size_t ke = (size_t)k * L.nrow * 2 * T.naj;
    const size_t kn = (size_t)k * 2 * T.nai * L.ncol;
    Edge g;
    g.ew = ewp[ke + ej];
    g.ns = nsp[kn + ni];
    g.ew2 = two ? ewp[ke + ew0 + 1] : 0.0f;
    g.ns2 = two ? nsp[kn + ns0 + L.ncol] : 0.0f;
    return g;
  }
  __device__ TV av_e(float sv, const Edge &g) const {
    TV r = (TV)sv * y;
    r = r + (TV)g.ew * dj;
    r = r + (TV)g.ns * di;
    if (two) {
      r = r + (TV)g.ew2 * dj2;
      r = r + (TV)g.ns2 * di2;
    }
    return r;
  }
};

template <typename TM, typename TV, typename TP, bool DEFL>
__global__ void __launch_bounds__(MF6_BLOCK) k_f_red_fwd(Layout L, const TM *__restrict__ A,
                                                         const TP *__restrict__ piv,
                                                         TV *__restrict__ d,
                                                         const TV *__restrict__ q,
                                                         TV *__restrict__ zf, double *part,
                                                         MaxLoc *mpart, const KState *st,
                                                         const FState *fs, int nblk,
                                                         const TV *__restrict__ yc,
                                                         const float *__restrict__ dsp,
                                                         const float *__restrict__ dew,
                                                         const float *__restrict__ dns,
                                                         TwoLevel T) {
  const int lb = xcd_block();
  __shared__ double lds[8];
  __shared__ MaxLoc mlds[8];
  if (st->stop) return;
  const int NP = L.NP;
  const TV alpha = (TV)fs->apend_d;
  double ss = 0.0;
  MaxLoc md{0.0, -1};
  for (int h2 = lb * blockDim.x + threadIdx.x; h2 < L.HP0;
       h2 += gridDim.x * blockDim.x) {
    int l = col_l(L, 0, h2);
    TV zup = (TV)0;
    TV lup = (TV)0;   // A(n, up) * piv(up)
    DeflCol<TV> dc;
    if constexpr (DEFL) {
      const int i = l / L.ncol;
      dc = DeflCol<TV>(yc, T, L, i, l - i * L.ncol);
    }
    for (int k = 0; k < L.nlay; k++) {
      int pos = k * L.HP + h2;
      TV dv = d[pos];
      dv = dv - alpha * q[pos];
      TV tv = dv;
      if constexpr (DEFL) tv = tv - dc.av(dsp, dew, dns, T, L, k, pos);
      d[pos] = dv;
      md = maxloc_pick(md, MaxLoc{(double)dv, k * L.NRC + l});
      ss = ss + (double)dv * (double)dv;
      if (k > 0) tv = tv - lup * zup;
      zf[pos] = tv;
      zup = tv;
      lup = (TV)((TV)A[6 * NP + pos] * (TV)ldp(piv, pos));
    }
  }
  ss = bsum(ss, lds);
  if (threadIdx.x == 0) part[P_SS_R * nblk + lb] = ss;
  md = bmax(md, mlds);
  if (threadIdx.x == 0) mpart[M_D_R * nblk + lb] = md;
}

// The preconditioner planes may be stored scaled by sc = 2^-k with the
// inverse pivots scaled by 2^k (FP16 planes): the forward products A * piv
// are then unscaled, and each backward sweep starts from its forward value
// times sc; sc = 1 otherwise (multiplying by 1 changes nothing).
//
// Black columns: forward sweep (red neighbours + cell above) written to z,
// then the backward sweep (cell below) over the same column.  Staging the
// forward values in z instead of registers keeps the register count low
// and the kernel memory-parallel.
template <typename TM, typename TV, typename TP, bool DEFL>
__global__ void __launch_bounds__(MF6_BLOCK) k_f_black(Layout L, const TM *__restrict__ A,
                                                       const TP *__restrict__ piv,
                                                       TV *__restrict__ d,
                                                       const TV *__restrict__ q,
                                                       const TV *__restrict__ zf,
                                                       TV *__restrict__ z, double *part,
                                                       MaxLoc *mpart, const KState *st,
                                                       const FState *fs, int nblk, TV sc,
                                                       const unsigned *__restrict__ cmask,
                                                       double *__restrict__ colsum,
                                                       const TV *__restrict__ yc,
                                                       const float *__restrict__ dsp,
                                                       const float *__restrict__ dew,
                                                       const float *__restrict__ dns,
                                                       TwoLevel T) {
  const int lb = xcd_block();
  __shared__ double lds[8];
  __shared__ MaxLoc mlds[8];
  if (st->stop) return;
  const int NP = L.NP;
  const TV alpha = (TV)fs->apend_d;
  double ss = 0.0, rho = 0.0;
  MaxLoc md{0.0, -1};
  for (int h2 = lb * blockDim.x + threadIdx.x; h2 < L.HP1;
       h2 += gridDim.x * blockDim.x) {
    int l = col_l(L, 1, h2);
    int i = l / L.ncol, j = l - i * L.ncol;
    // neighbours off the grid read position 0 of their plane (in range) and
    // their term is dropped by a select: every lane issues the same loads
    // and none waits inside a branch
    const bool hb = i > 0, hl = j > 0, hr = j < L.ncol - 1, hf = i < L.nrow - 1;
    const int ib = hb ? hn(l, -L.ncol) : 0;
    const int il = hl ? hn(l, -1) : 0;
    const int ir = hr ? hn(l, 1) : 0;
    const int ifr = hf ? hn(l, L.ncol) : 0;
    TV zup = (TV)0, lup = (TV)0;
    const unsigned cm = colsum ? cmask[L.HP + h2] : 0u;
    double cs = 0.0;
    DeflCol<TV> dc;
    if constexpr (DEFL) dc = DeflCol<TV>(yc, T, L, i, j);
    for (int k = 0; k < L.nlay; k++) {
      int pos = L.H + k * L.HP + h2;
      int rb = k * L.HP;
      // L(n, r) = A(n, r) piv(r); A(n, back) = front(back), A(n, left) =
      // right(left), A(n, right) = right(n), A(n, front) = front(n).  The
      // neighbour terms come first so that their loads go out with d's.
      const TV tb = (TV)((TV)A[5 * NP + rb + ib] * (TV)ldp(piv, rb + ib)) * zf[rb + ib];
      const TV tl = (TV)((TV)A[4 * NP + rb + il] * (TV)ldp(piv, rb + il)) * zf[rb + il];
      const TV tr = (TV)((TV)A[4 * NP + pos] * (TV)ldp(piv, rb + ir)) * zf[rb + ir];
      const TV tf = (TV)((TV)A[5 * NP + pos] * (TV)ldp(piv, rb + ifr)) * zf[rb + ifr];
      TV dv = d[pos];
      dv = dv - alpha * q[pos];
      TV tv = dv;
      if constexpr (DEFL) tv = tv - dc.av(dsp, dew, dns, T, L, k, pos);
      d[pos] = dv;
      cs = ((cm >> k) & 1u) ? cs + (double)dv : cs;
      md = maxloc_pick(md, MaxLoc{(double)dv, k * L.NRC + l});
      ss = ss + (double)dv * (double)dv;
      tv = hb ? tv - tb : tv;
      tv = hl ? tv - tl : tv;
      tv = hr ? tv - tr : tv;
      tv = hf ? tv - tf : tv;
      if (k > 0) tv = tv - lup * zup;
      z[pos] = tv;
      zup = tv;
      lup = (TV)((TV)A[6 * NP + pos] * (TV)ldp(piv, pos));
    }
    if (colsum) colsum[l] = cs;
    TV zdn = (TV)0;
    for (int k = L.nlay - 1; k >= 0; k--) {
      int pos = L.H + k * L.HP + h2;
      TV tv = z[pos] * sc;
      if (k < L.nlay - 1) tv = tv - (TV)A[6 * NP + pos] * zdn;
      tv = tv * (TV)ldp(piv, pos);
      z[pos] = tv;
      zdn = tv;
      rho = rho + (double)d[pos] * (double)tv;
    }
  }
  ss = bsum(ss, lds);
  if (threadIdx.x == 0) part[P_SS_B * nblk + lb] = ss;
  rho = bsum(rho, lds);
  if (threadIdx.x == 0) part[P_RHO_B * nblk + lb] = rho;
  md = bmax(md, mlds);
  if (threadIdx.x == 0) mpart[M_D_B * nblk + lb] = md;
}

template <typename TM, typename TV, typename TP>
__global__ void __launch_bounds__(MF6_BLOCK) k_f_red_bwd(Layout L, const TM *__restrict__ A,
                                                         const TP *__restrict__ piv,
                                                         const TV *__restrict__ d,
                                                         const TV *__restrict__ zf,
                                                         TV *__restrict__ z, double *part,
                                                         const KState *st, int nblk, TV sc,
                                                         const unsigned *__restrict__ cmask,
                                                         double *__restrict__ colsum) {
  const int lb = xcd_block();
  __shared__ double lds[8];
  if (st->stop) return;
  const int NP = L.NP;
  double rho = 0.0;
  for (int h2 = lb * blockDim.x + threadIdx.x; h2 < L.HP0;
       h2 += gridDim.x * blockDim.x) {
    int l = col_l(L, 0, h2);
    int i = l / L.ncol, j = l - i * L.ncol;
    // off-grid neighbours: in-range position 0, term dropped by a select
    const bool hb = i > 0, hl = j > 0, hr = j < L.ncol - 1, hf = i < L.nrow - 1;
    const int ib = hb ? hn(l, -L.ncol) : 0;
    const int il = hl ? hn(l, -1) : 0;
    const int ir = hr ? hn(l, 1) : 0;
    const int ifr = hf ? hn(l, L.ncol) : 0;
    TV zdn = (TV)0;
    const unsigned cm = colsum ? cmask[h2] : 0u;
    double cs = 0.0;
    for (int k = L.nlay - 1; k >= 0; k--) {
      int pos = k * L.HP + h2;
      int bb = L.H + k * L.HP;
      TV tv = zf[pos] * sc;
      if (k < L.nlay - 1) tv = tv - (TV)A[6 * NP + pos] * zdn;
      const TV tb = (TV)A[5 * NP + bb + ib] * z[bb + ib];
      const TV tl = (TV)A[4 * NP + bb + il] * z[bb + il];
      const TV tr = (TV)A[4 * NP + pos] * z[bb + ir];
      const TV tf = (TV)A[5 * NP + pos] * z[bb + ifr];
      tv = hb ? tv - tb : tv;
      tv = hl ? tv - tl : tv;
      tv = hr ? tv - tr : tv;
      tv = hf ? tv - tf : tv;
      tv = tv * (TV)ldp(piv, pos);
      z[pos] = tv;
      zdn = tv;
      const TV dv = d[pos];
      rho = rho + (double)dv * (double)tv;
      cs = ((cm >> k) & 1u) ? cs + (double)dv : cs;
    }
    if (colsum) colsum[l] = cs;
  }
  rho = bsum(rho, lds);
  if (threadIdx.x == 0) part[P_RHO_R * nblk + lb] = rho;
}

// ---------------------------------------------------------------------------
// The same three kernels with W columns of one colour per thread
// ---------------------------------------------------------------------------
//
// For ncol a multiple of 2W (groups never straddle a grid row; HP is then a
// multiple of W), each stream is one W-wide access.  Column m of colour c in
// grid row i sits at h = i ncol/2 + m; its neighbours in the other colour
// are at h - ncol/2 (back), h + ncol/2 (front), h - 1 + s (left) and h + s
// (right), s = (i + c) mod 2.  A group's back and front values are one
// aligned access each; its left and right values are the W-wide access at
// h0 plus one edge value (h0 - 1 for s = 0, h0 + W for s = 1).  Per cell
// the arithmetic is that of the single-column kernels, in the same order.
// The grid is sized for the groups (a W-th of the columns); the colour
// partials are then reduced over that grid by k_f_round.

// geometry of a group of W columns of colour c starting at h0
template <int W>
struct GroupGeo {
  int i, sf, ib, ifr, ee;
  bool hb, hf, lok, rok;
  __device__ GroupGeo(const Layout &L, int c, int h0) {
    const int half = L.ncol >> 1;
    i = h0 / half;
    const int m0 = h0 - i * half;
    sf = (i & 1) ^ c;
    hb = i > 0;
    hf = i < L.nrow - 1;
    lok = sf == 1 || m0 > 0;          // the first column has a left neighbour
    rok = sf == 0 || m0 + W < half;   // the last column has a right neighbour
    // edge position, clamped into the layer (its term is dropped when the
    // neighbour is off the grid)
    ee = sf == 0 ? (h0 > 0 ? h0 - 1 : 0) : (h0 + W < L.HP ? h0 + W : L.HP - 1);
    ib = hb ? h0 - half : h0;
    ifr = hf ? h0 + half : h0;
  }
  // grid column (l) of the group's column e
  __device__ int col(int h0, int e) const { return 2 * (h0 + e) + sf; }
  // element e's left / right neighbour value from the W-wide load at h0 (w)
  // and the edge value (x)
  template <typename T>
  __device__ T left(const VecW<T, W> &w, T x, int e) const {
    return sf ? w.v[e] : (e > 0 ? w.v[e > 0 ? e - 1 : 0] : x);
  }
  template <typename T>
  __device__ T right(const VecW<T, W> &w, T x, int e) const {
    return sf ? (e < W - 1 ? w.v[e < W - 1 ? e + 1 : e] : x) : w.v[e];
  }
  __device__ bool has_left(int e) const { return e > 0 || lok; }
  __device__ bool has_right(int e) const { return e < W - 1 || rok; }
};

// The layers are walked in chunks of CH (MF6_SWEEP_CH, default 2) whose
// loads (d, q, the down plane, the pivots and, deflated, the row sums and
// edge couplings) are all issued before the chunk's arithmetic; the
// arithmetic and its order are those of the one-layer-at-a-time walk.  On
// the 10 x 1000 x 1000 steady solve chunks of 4 are slower.
#ifndef MF6_SWEEP_CH
#define MF6_SWEEP_CH 2
#endif
template <typename TM, typename TV, typename TP, bool DEFL, int W, int CH = MF6_SWEEP_CH>
__global__ void __launch_bounds__(MF6_BLOCK) k_f_red_fwdw(Layout L, const TM *__restrict__ A,
                                                          const TP *__restrict__ piv,
                                                          TV *__restrict__ d,
                                                          const TV *__restrict__ q,
                                                          TV *__restrict__ zf, double *part,
                                                          MaxLoc *mpart, const KState *st,
                                                          const FState *fs, int nblk,
                                                          const TV *__restrict__ yc,
                                                          const float *__restrict__ dsp,
                                                          const float *__restrict__ dew,
                                                          const float *__restrict__ dns,
                                                          TwoLevel T) {
  const int lb = xcd_block();
  __shared__ double lds[8];
  __shared__ MaxLoc mlds[8];
  if (st->stop) return;
  const int NP = L.NP, HP = L.HP;
  const TV alpha = (TV)fs->apend_d;
  double ss = 0.0;
  MaxLoc md{0.0, -1};
  for (int t = lb * blockDim.x + threadIdx.x; t < HP / W; t += gridDim.x * blockDim.x) {
    const int h0 = t * W;
    const GroupGeo<W> G(L, 0, h0);
    DeflCol<TV> dc[W];
    if constexpr (DEFL) {
#pragma unroll
      for (int e = 0; e < W; e++) dc[e] = DeflCol<TV>(yc, T, L, G.i, G.col(h0, e) - G.i * L.ncol);
    }
    TV zup[W], lup[W];
#pragma unroll
    for (int e = 0; e < W; e++) zup[e] = lup[e] = (TV)0;
    for (int k0 = 0; k0 < L.nlay; k0 += CH) {
      VecW<TV, W> dv[CH], qv[CH];
      VecW<TM, W> a6[CH];
      VecW<TP, W> pv[CH];
      VecW<float, W> sv[CH];
      typename DeflCol<TV>::Edge eg[CH][W];
#pragma unroll
      for (int u = 0; u < CH; u++) {
        const int k = k0 + u;
        if (k < L.nlay) {
          const int pos = k * HP + h0;
          dv[u] = ldw<W>(d + pos);
          qv[u] = ldw<W>(q + pos);
          a6[u] = ldw<W>(A + 6 * NP + pos);
          pv[u] = ldw<W>(piv + pos);
          if constexpr (DEFL) {
            sv[u] = ldw<W>(dsp + pos);
#pragma unroll
            for (int e = 0; e < W; e++) eg[u][e] = dc[e].edges(dew, dns, T, L, k);
          }
        }
      }
#pragma unroll
      for (int u = 0; u < CH; u++) {
        const int k = k0 + u;
        if (k < L.nlay) {
          const int pos = k * HP + h0;
          VecW<TV, W> zo;
#pragma unroll
          for (int e = 0; e < W; e++) {
            const TV x = dv[u].v[e] - alpha * qv[u].v[e];
            TV tv = x;
            if constexpr (DEFL) tv = tv - dc[e].av_e(sv[u].v[e], eg[u][e]);
            dv[u].v[e] = x;
            md = maxloc_pick(md, MaxLoc{(double)x, k * L.NRC + G.col(h0, e)});
            ss = ss + (double)x * (double)x;
            if (k > 0) tv = tv - lup[e] * zup[e];
            zo.v[e] = tv;
            zup[e] = tv;
            lup[e] = (TV)((TV)a6[u].v[e] * (TV)pv[u].v[e]);
          }
          stw<W>(d + pos, dv[u]);
          stw<W>(zf + pos, zo);
        }
      }
    }
  }
  ss = bsum(ss, lds);
  if (threadIdx.x == 0) part[P_SS_R * nblk + lb] = ss;
  md = bmax(md, mlds);
  if (threadIdx.x == 0) mpart[M_D_R * nblk + lb] = md;
}

template <typename TM, typename TV, typename TP, bool DEFL, int W>
__global__ void __launch_bounds__(MF6_BLOCK) k_f_blackw(Layout L, const TM *__restrict__ A,
                                                        const TP *__restrict__ piv,
                                                        TV *__restrict__ d,
                                                        const TV *__restrict__ q,
                                                        const TV *__restrict__ zf,
                                                        TV *__restrict__ z, double *part,
                                                        MaxLoc *mpart, const KState *st,
                                                        const FState *fs, int nblk, TV sc,
                                                        const unsigned *__restrict__ cmask,
                                                        double *__restrict__ colsum,
                                                        const TV *__restrict__ yc,
                                                        const float *__restrict__ dsp,
                                                        const float *__restrict__ dew,
                                                        const float *__restrict__ dns,
                                                        TwoLevel T) {
  const int lb = xcd_block();
  __shared__ double lds[8];
  __shared__ MaxLoc mlds[8];
  if (st->stop) return;
  const int NP = L.NP, HP = L.HP;
  const TV alpha = (TV)fs->apend_d;
  double ss = 0.0, rho = 0.0;
  MaxLoc md{0.0, -1};
  for (int t = lb * blockDim.x + threadIdx.x; t < HP / W; t += gridDim.x * blockDim.x) {
    const int h0 = t * W;
    const GroupGeo<W> G(L, 1, h0);
    VecW<unsigned, W> cm;
    double cs[W];
#pragma unroll
    for (int e = 0; e < W; e++) {
      cm.v[e] = 0u;
      cs[e] = 0.0;
    }
    if (colsum) cm = ldw<W>(cmask + HP + h0);
    DeflCol<TV> dc[W];
    if constexpr (DEFL) {
#pragma unroll
      for (int e = 0; e < W; e++) dc[e] = DeflCol<TV>(yc, T, L, G.i, G.col(h0, e) - G.i * L.ncol);
    }
    TV zup[W], lup[W];
#pragma unroll
    for (int e = 0; e < W; e++) zup[e] = lup[e] = (TV)0;
    for (int k = 0; k < L.nlay; k++) {
      const int pos = L.H + k * HP + h0;
      const int rb = k * HP;
      // L(n, r) = A(n, r) piv(r) with A(n, back) = front(back), A(n, left) =
      // right(left), A(n, right) = right(n), A(n, front) = front(n); the
      // neighbour loads go out with d's
      const VecW<TM, W> a5b = ldw<W>(A + 5 * NP + rb + G.ib);
      const VecW<TP, W> pb = ldw<W>(piv + rb + G.ib);
      const VecW<TV, W> zb = ldw<W>(zf + rb + G.ib);
      const VecW<TM, W> a4w = ldw<W>(A + 4 * NP + rb + h0);
      const TM a4x = A[4 * NP + rb + G.ee];
      const VecW<TP, W> pw = ldw<W>(piv + rb + h0);
      const TP px = piv[rb + G.ee];
      const VecW<TV, W> zw = ldw<W>(zf + rb + h0);
      const TV zx = zf[rb + G.ee];
      const VecW<TP, W> pf = ldw<W>(piv + rb + G.ifr);
      const VecW<TV, W> zfr = ldw<W>(zf + rb + G.ifr);
      const VecW<TM, W> a4 = ldw<W>(A + 4 * NP + pos);
      const VecW<TM, W> a5 = ldw<W>(A + 5 * NP + pos);
      const VecW<TM, W> a6 = ldw<W>(A + 6 * NP + pos);
      const VecW<TP, W> po = ldw<W>(piv + pos);
      VecW<TV, W> dv = ldw<W>(d + pos);
      const VecW<TV, W> qv = ldw<W>(q + pos);
      VecW<float, W> sv;
      typename DeflCol<TV>::Edge eg[W];
      if constexpr (DEFL) {
        sv = ldw<W>(dsp + pos);
#pragma unroll
        for (int e = 0; e < W; e++) eg[e] = dc[e].edges(dew, dns, T, L, k);
      }
      VecW<TV, W> zo;
#pragma unroll
      for (int e = 0; e < W; e++) {
        const TV tb = (TV)((TV)a5b.v[e] * (TV)pb.v[e]) * zb.v[e];
        const TV tl = (TV)((TV)G.left(a4w, a4x, e) * (TV)G.left(pw, px, e)) * G.left(zw, zx, e);
        const TV tr = (TV)((TV)a4.v[e] * (TV)G.right(pw, px, e)) * G.right(zw, zx, e);
        const TV tf = (TV)((TV)a5.v[e] * (TV)pf.v[e]) * zfr.v[e];
        const TV x = dv.v[e] - alpha * qv.v[e];
        TV tv = x;
        if constexpr (DEFL) tv = tv - dc[e].av_e(sv.v[e], eg[e]);
        dv.v[e] = x;
        cs[e] = ((cm.v[e] >> k) & 1u) ? cs[e] + (double)x : cs[e];
        md = maxloc_pick(md, MaxLoc{(double)x, k * L.NRC + G.col(h0, e)});
        ss = ss + (double)x * (double)x;
        tv = G.hb ? tv - tb : tv;
        tv = G.has_left(e) ? tv - tl : tv;
        tv = G.has_right(e) ? tv - tr : tv;
        tv = G.hf ? tv - tf : tv;
        if (k > 0) tv = tv - lup[e] * zup[e];
        zo.v[e] = tv;
        zup[e] = tv;
        lup[e] = (TV)((TV)a6.v[e] * (TV)po.v[e]);
      }
      stw<W>(d + pos, dv);
      stw<W>(z + pos, zo);
    }
    if (colsum) {
#pragma unroll
      for (int e = 0; e < W; e++) colsum[G.col(h0, e)] = cs[e];
    }
    TV zdn[W];
#pragma unroll
    for (int e = 0; e < W; e++) zdn[e] = (TV)0;
    for (int k = L.nlay - 1; k >= 0; k--) {
      const int pos = L.H + k * HP + h0;
      VecW<TV, W> zv = ldw<W>(z + pos);
      const VecW<TM, W> a6 = ldw<W>(A + 6 * NP + pos);
      const VecW<TP, W> po = ldw<W>(piv + pos);
      const VecW<TV, W> dv = ldw<W>(d + pos);
#pragma unroll
      for (int e = 0; e < W; e++) {
        TV tv = zv.v[e] * sc;
        if (k < L.nlay - 1) tv = tv - (TV)a6.v[e] * zdn[e];
        tv = tv * (TV)po.v[e];
        zv.v[e] = tv;
        zdn[e] = tv;
        rho = rho + (double)dv.v[e] * (double)tv;
      }
      stw<W>(z + pos, zv);
    }
  }
  ss = bsum(ss, lds);
  if (threadIdx.x == 0) part[P_SS_B * nblk + lb] = ss;
  rho = bsum(rho, lds);
  if (threadIdx.x == 0) part[P_RHO_B * nblk + lb] = rho;
  md = bmax(md, mlds);
  if (threadIdx.x == 0) mpart[M_D_B * nblk + lb] = md;
}

template <typename TM, typename TV, typename TP, int W>
__global__ void __launch_bounds__(MF6_BLOCK) k_f_red_bwdw(Layout L, const TM *__restrict__ A,
                                                          const TP *__restrict__ piv,
                                                          const TV *__restrict__ d,
                                                          const TV *__restrict__ zf,
                                                          TV *__restrict__ z, double *part,
                                                          const KState *st, int nblk, TV sc,
                                                          const unsigned *__restrict__ cmask,
                                                          double *__restrict__ colsum) {
  const int lb = xcd_block();
  __shared__ double lds[8];
  if (st->stop) return;
  const int NP = L.NP, HP = L.HP;
  double rho = 0.0;
  for (int t = lb * blockDim.x + threadIdx.x; t < HP / W; t += gridDim.x * blockDim.x) {
    const int h0 = t * W;
    const GroupGeo<W> G(L, 0, h0);
    VecW<unsigned, W> cm;
    double cs[W];
    TV zdn[W];
#pragma unroll
    for (int e = 0; e < W; e++) {
      cm.v[e] = 0u;
      cs[e] = 0.0;
      zdn[e] = (TV)0;
    }
    if (colsum) cm = ldw<W>(cmask + h0);
    for (int k = L.nlay - 1; k >= 0; k--) {
      const int pos = k * HP + h0;
      const int bb = L.H + k * HP;
      const VecW<TV, W> zfv = ldw<W>(zf + pos);
      const VecW<TM, W> a6 = ldw<W>(A + 6 * NP + pos);
      const VecW<TM, W> a5b = ldw<W>(A + 5 * NP + bb + G.ib);
      const VecW<TV, W> zb = ldw<W>(z + bb + G.ib);
      const VecW<TM, W> a4w = ldw<W>(A + 4 * NP + bb + h0);
      const TM a4x = A[4 * NP + bb + G.ee];
      const VecW<TV, W> zw = ldw<W>(z + bb + h0);
      const TV zx = z[bb + G.ee];
      const VecW<TV, W> zfr = ldw<W>(z + bb + G.ifr);
      const VecW<TM, W> a4 = ldw<W>(A + 4 * NP + pos);
      const VecW<TM, W> a5 = ldw<W>(A + 5 * NP + pos);
      const VecW<TP, W> po = ldw<W>(piv + pos);
      const VecW<TV, W> dv = ldw<W>(d + pos);
      VecW<TV, W> zo;
#pragma unroll
      for (int e = 0; e < W; e++) {
        TV tv = zfv.v[e] * sc;
        if (k < L.nlay - 1) tv = tv - (TV)a6.v[e] * zdn[e];
        const TV tb = (TV)a5b.v[e] * zb.v[e];
        const TV tl = (TV)G.left(a4w, a4x, e) * G.left(zw, zx, e);
        const TV tr = (TV)a4.v[e] * G.right(zw, zx, e);
        const TV tf = (TV)a5.v[e] * zfr.v[e];
        tv = G.hb ? tv - tb : tv;
        tv = G.has_left(e) ? tv - tl : tv;
        tv = G.has_right(e) ? tv - tr : tv;
        tv = G.hf ? tv - tf : tv;
        tv = tv * (TV)po.v[e];
        zo.v[e] = tv;
        zdn[e] = tv;
        rho = rho + (double)dv.v[e] * (double)tv;
        cs[e] = ((cm.v[e] >> k) & 1u) ? cs[e] + (double)dv.v[e] : cs[e];
      }
      stw<W>(z + pos, zo);
    }
    if (colsum) {
#pragma unroll
      for (int e = 0; e < W; e++) colsum[G.col(h0, e)] = cs[e];
    }
  }
  rho = bsum(rho, lds);
  if (threadIdx.x == 0) part[P_RHO_R * nblk + lb] = rho;
}

// ---------------------------------------------------------------------------
// direction update with the deferred head update, then the SpMV
// ---------------------------------------------------------------------------

// p = z (first iteration) or z + beta p_prev ; x += alpha_prev p_prev ;
// partial max |p| (signed, first occurrence) for deltax = alpha pmax.
// With the two-level preconditioner (yc not null) z is first completed by
// the prolongation of the coarse correction: z + yc[aggregate] on the
// coupled cells (bit k of the column's cmask).
template <typename TV, typename TX>
__global__ void __launch_bounds__(MF6_BLOCK) k_f_pupdate(Layout L, const TV *__restrict__ z,
                                                         const TV *__restrict__ pin,
                                                         TV *__restrict__ pout, TX *__restrict__ x,
                                                         MaxLoc *mpart, const KState *st,
                                                         const FState *fs, int nblk,
                                                         const TV *__restrict__ yc,
                                                         const unsigned *__restrict__ cmask,
                                                         TwoLevel T) {
  const int lb = xcd_block();
  __shared__ MaxLoc mlds[8];
  if (st->stop) return;
  const bool first = st->iter == 0;
  const TV beta = (TV)st->beta;
  const TV ax = (TV)fs->apend_x;
  MaxLoc mp{0.0, -1};
  const int nthreads = 2 * L.HP0;
  for (int t = lb * blockDim.x + threadIdx.x; t < nthreads;
       t += gridDim.x * blockDim.x) {
    int c = t & 1;
    int h2 = t >> 1;
    if (h2 >= ncols_of(L, c)) continue;
    int l = col_l(L, c, h2);
    const int own = c * L.H + h2;
    const unsigned cm = yc ? cmask[c * L.HP + h2] : 0u;
    const TV yv = cm ? yc[tl_agg(T, L, l)] : (TV)0;
    for (int k = 0; k < L.nlay; k++) {
      int pos = own + k * L.HP;
      TV zv = z[pos];
      if ((cm >> k) & 1u) zv = zv + yv;
      TV pv;
      if (first) {
        pv = zv;
      } else {
        TV po = pin[pos];
        pv = zv + beta * po;
        x[pos] = (TX)(x[pos] + (TX)(ax * po));
      }
      pout[pos] = pv;
      mp = maxloc_pick(mp, MaxLoc{(double)pv, k * L.NRC + l});
    }
  }
  mp = bmax(mp, mlds);
  if (threadIdx.x == 0) mpart[M_P * nblk + lb] = mp;
}

// The same update with W columns of one colour per thread (W-wide
// accesses; ncol a multiple of 2W so that a group never straddles a grid
// row), on its own grid; slots gridDim.x .. nblk-1 of its max |p| partials
// are cleared.
template <typename TV, typename TX, int W>
__global__ void __launch_bounds__(MF6_BLOCK) k_f_pupdatew(Layout L, const TV *__restrict__ z,
                                                          const TV *__restrict__ pin,
                                                          TV *__restrict__ pout,
                                                          TX *__restrict__ x, MaxLoc *mpart,
                                                          const KState *st, const FState *fs,
                                                          int nblk, const TV *__restrict__ yc,
                                                          const unsigned *__restrict__ cmask,
                                                          TwoLevel T) {
  const int lb = xcd_block();
  __shared__ MaxLoc mlds[8];
  if (st->stop) return;
  const bool first = st->iter == 0;
  const TV beta = (TV)st->beta;
  const TV ax = (TV)fs->apend_x;
  const int HP = L.HP, half = L.ncol >> 1;
  MaxLoc mp{0.0, -1};
  for (int t = lb * blockDim.x + threadIdx.x; t < 2 * (HP / W); t += gridDim.x * blockDim.x) {
    const int c = t & 1;
    const int h0 = (t >> 1) * W;
    const int sf = ((h0 / half) & 1) ^ c;
    const int own = c * L.H + h0;
    VecW<unsigned, W> cm;
    TV yv[W];
#pragma unroll
    for (int e = 0; e < W; e++) {
      cm.v[e] = 0u;
      yv[e] = (TV)0;
    }
    if (yc) {
      cm = ldw<W>(cmask + c * HP + h0);
#pragma unroll
      for (int e = 0; e < W; e++)
        if (cm.v[e]) yv[e] = yc[tl_agg(T, L, 2 * (h0 + e) + sf)];
    }
    for (int k = 0; k < L.nlay; k++) {
      const int pos = own + k * HP;
      VecW<TV, W> zv = ldw<W>(z + pos);
      VecW<TV, W> pv;
      if (first) {
#pragma unroll
        for (int e = 0; e < W; e++) pv.v[e] = ((cm.v[e] >> k) & 1u) ? zv.v[e] + yv[e] : zv.v[e];
      } else {
        const VecW<TV, W> po = ldw<W>(pin + pos);
        VecW<TX, W> xv = ldw<W>(x + pos);
#pragma unroll
        for (int e = 0; e < W; e++) {
          const TV zz = ((cm.v[e] >> k) & 1u) ? zv.v[e] + yv[e] : zv.v[e];
          pv.v[e] = zz + beta * po.v[e];
          xv.v[e] = (TX)(xv.v[e] + (TX)(ax * po.v[e]));
        }
        stw<W>(x + pos, xv);
      }
      stw<W>(pout + pos, pv);
#pragma unroll
      for (int e = 0; e < W; e++)
        mp = maxloc_pick(mp, MaxLoc{(double)pv.v[e], k * L.NRC + 2 * (h0 + e) + sf});
    }
  }
  mp = bmax(mp, mlds);
  if (threadIdx.x == 0) {
    mpart[M_P * nblk + lb] = mp;
    for (int e = (int)gridDim.x + lb; e < nblk; e += (int)gridDim.x)
      mpart[M_P * nblk + e] = MaxLoc{0.0, -1};
  }
}

// Columns per thread of the grouped kernels for a layout: 16-byte accesses
// (4 FP32 or 2 FP64 values) where the grid rows allow, else 2, else 0 (the
// single-column kernels).
template <typename TV>
static inline int col_group(const Layout &L) {
  constexpr int WMAX = 16 / (int)sizeof(TV);
  if (L.odd) return 0;
  if (L.ncol % (2 * WMAX) == 0) return WMAX;
  if (WMAX > 2 && L.ncol % 4 == 0) return 2;
  return 0;
}

// grid of the colour kernels (one thread per group of one colour); k_f_round
// reduces their partials over this many slots
template <typename TV>
static inline int colour_grid(const Layout &L) {
  const int w = col_group<TV>(L);
  int n = w ? L.HP / w : L.HP0;
  int g = (n + MF6_BLOCK - 1) / MF6_BLOCK;
  if (g > MF6_MAXBLK) g = MF6_MAXBLK;
  return g < 1 ? 1 : g;
}

template <typename TV, typename TX>
static void launch_f_pupdate(const Layout &L, int nb, hipStream_t s, const TV *z, const TV *pin,
                             TV *pout, TX *x, MaxLoc *mpart, const KState *st,
                             const FState *fs, const TV *yc = nullptr,
                             const unsigned *cmask = nullptr, TwoLevel T = TwoLevel{}) {
  constexpr int WMAX = 16 / (int)sizeof(TV);
  const int w = col_group<TV>(L);
  if (w) {
    int g = (2 * (L.HP / w) + MF6_BLOCK - 1) / MF6_BLOCK;
    if (g > nb) g = nb;
    if (g < 1) g = 1;
    if (w == WMAX)
      k_f_pupdatew<TV, TX, WMAX><<<g, MF6_BLOCK, 0, s>>>(L, z, pin, pout, x, mpart, st, fs, nb,
                                                         yc, cmask, T);
    else
      k_f_pupdatew<TV, TX, 2><<<g, MF6_BLOCK, 0, s>>>(L, z, pin, pout, x, mpart, st, fs, nb, yc,
                                                      cmask, T);
  } else {
    k_f_pupdate<TV, TX><<<nb, MF6_BLOCK, 0, s>>>(L, z, pin, pout, x, mpart, st, fs, nb, yc,
                                                 cmask, T);
  }
}

// q = A p (amux order, symmetric planes), partial p.q.  The layer walk is
// unrolled over a compile-time bound NLMAX in chunks of four layers whose
// loads are all issued before their arithmetic (more loads in flight per
// thread than the runtime-bounded loop gets).
template <typename TM, typename TV, int NLMAX, bool CS>
__global__ void __launch_bounds__(MF6_BLOCK) k_f_spmv(Layout L, const TM *__restrict__ A,
                                                      const TV *__restrict__ p,
                                                      TV *__restrict__ q, double *part,
                                                      const KState *st, int nblk,
                                                      const unsigned *__restrict__ cmask,
                                                      double *__restrict__ colsq) {
  constexpr int CH = 4;
  const int lb = xcd_block();
  __shared__ double lds[8];
  if (st->stop) return;
  const int NP = L.NP;
  double s = 0.0;
  // adjacent threads take the red and the black column of the same index,
  // so the neighbour values one colour gathers from the other are the ones
  // its partner thread streams
  const int nthreads = 2 * L.HP0;
  for (int t = lb * blockDim.x + threadIdx.x; t < nthreads;
       t += gridDim.x * blockDim.x) {
    int c = t & 1;
    int h2 = t >> 1;
    if (h2 >= ncols_of(L, c)) continue;
    int l = col_l(L, c, h2);
    int i = l / L.ncol, j = l - i * L.ncol;
    const bool hb = i > 0, hl = j > 0, hr = j < L.ncol - 1, hf = i < L.nrow - 1;
    const int ib = hb ? hn(l, -L.ncol) : 0;
    const int il = hl ? hn(l, -1) : 0;
    const int ir = hr ? hn(l, 1) : 0;
    const int ifr = hf ? hn(l, L.ncol) : 0;
    const int own = c * L.H + h2;
    const int oth = (1 - c) * L.H;
    TV pprev = (TV)0;
    TV pcur = p[own];
    TV aup = (TV)0;
    unsigned cm = 0u;
    if constexpr (CS) cm = cmask[c * L.HP + h2];
    double cs = 0.0;
#pragma unroll
    for (int k0 = 0; k0 < NLMAX; k0 += CH) {
      if (k0 >= L.nlay) break;
      TV a0[CH], a4[CH], a5[CH], a6[CH], ab[CH], al[CH], pb[CH], pl[CH], pr[CH], pf[CH],
          pn[CH];
#pragma unroll
      for (int u = 0; u < CH; u++) {
        const int k = k0 + u;
        if (k < L.nlay) {
          const int pos = own + k * L.HP, ob = oth + k * L.HP;
          const bool dn = k < L.nlay - 1;
          a0[u] = (TV)A[pos];
          a6[u] = dn ? (TV)A[6 * NP + pos] : (TV)0;
          pn[u] = dn ? p[pos + L.HP] : (TV)0;
          ab[u] = hb ? (TV)A[5 * NP + ob + ib] : (TV)0;
          pb[u] = hb ? p[ob + ib] : (TV)0;
          al[u] = hl ? (TV)A[4 * NP + ob + il] : (TV)0;
          pl[u] = hl ? p[ob + il] : (TV)0;
          a4[u] = hr ? (TV)A[4 * NP + pos] : (TV)0;
          pr[u] = hr ? p[ob + ir] : (TV)0;
          a5[u] = hf ? (TV)A[5 * NP + pos] : (TV)0;
          pf[u] = hf ? p[ob + ifr] : (TV)0;
        }
      }
#pragma unroll
      for (int u = 0; u < CH; u++) {
        const int k = k0 + u;
        if (k < L.nlay) {
          TV tq = (TV)0;
          tq = tq + a0[u] * pcur;
          if (k > 0) tq = tq + aup * pprev;
          if (hb) tq = tq + ab[u] * pb[u];
          if (hl) tq = tq + al[u] * pl[u];
          if (hr) tq = tq + a4[u] * pr[u];
          if (hf) tq = tq + a5[u] * pf[u];
          if (k < L.nlay - 1) tq = tq + a6[u] * pn[u];
          q[own + k * L.HP] = tq;
          if constexpr (CS)
            if ((cm >> k) & 1u) cs = cs + (double)tq;
          s = s + (double)pcur * (double)tq;
          aup = a6[u];
          pprev = pcur;
          pcur = pn[u];
        }
      }
    }
    if constexpr (CS) colsq[c * L.HP + h2] = cs;
  }
  s = bsum(s, lds);
  if (threadIdx.x == 0) part[P_PQ * nblk + lb] = s;
}

// The same product with W columns of one colour per thread (positions h0 ..
// h0 + W - 1, h0 a multiple of W): every stream is one W-wide load, the own
// planes and p as in the single-column kernel and the other colour's
// neighbour values as well.  With ncol even, column m of colour c in grid
// row i sits at position h = i ncol/2 + m and its neighbours in the other
// colour at h - ncol/2 (back), h + ncol/2 (front), and h - 1 + s (left) /
// h + s (right), s = (i + c) mod 2.  The W columns' back and front values are
// therefore one aligned load each, and their left and right values lie in
// the W-wide load at h0 plus one edge value (h0 - 1 for s = 0, h0 + W for
// s = 1).  Needs ncol a multiple of 2W, so that a group never straddles a
// grid row (HP is then a multiple of W).  Same amux order per cell; the p.q
// partial adds the group's cells per layer.
// It runs on its own grid (a W-th of the columns of k_f_spmv; a grid sized
// for the single-column kernel would leave the XCDs that receive the upper
// part of the logical blocks idle); slots gridDim.x .. nblk-1 of its p.q
// partials (read by k_f_alpha over nblk slots) are cleared.
template <typename TV, int W, int NLMAX, int CH, bool CS>
__global__ void __launch_bounds__(MF6_BLOCK) k_f_spmvw(Layout L, const TV *__restrict__ A,
                                                       const TV *__restrict__ p,
                                                       TV *__restrict__ q, double *part,
                                                       const KState *st, int nblk,
                                                       const unsigned *__restrict__ cmask,
                                                       double *__restrict__ colsq) {
  using VT = VecW<TV, W>;
  const int lb = xcd_block();
  __shared__ double lds[8];
  if (st->stop) return;
  const int NP = L.NP, HP = L.HP, half = L.ncol >> 1;
  const int ngrp = 2 * (HP / W);   // groups of both colours, interleaved
  double s = 0.0;
  for (int t = lb * blockDim.x + threadIdx.x; t < ngrp; t += gridDim.x * blockDim.x) {
    const int c = t & 1;
    const int h0 = (t >> 1) * W;
    const int i = h0 / half, m0 = h0 - i * half;
    const int sf = (i & 1) ^ c;
    const bool hb = i > 0, hf = i < L.nrow - 1;
    const bool lok = sf == 1 || m0 > 0;          // the first column has a left neighbour
    const bool rok = sf == 0 || m0 + W < half;   // the last column has a right neighbour
    // edge value position, clamped into the layer (its term is dropped when
    // the neighbour is off the grid)
    const int ee = sf == 0 ? (h0 > 0 ? h0 - 1 : 0) : (h0 + W < HP ? h0 + W : HP - 1);
    const int own = c * L.H + h0, oth = (1 - c) * L.H;
    const int ib = hb ? h0 - half : h0, ifr = hf ? h0 + half : h0;
    VT pcur = *reinterpret_cast<const VT *>(p + own);
    VT pprev, aup;
#pragma unroll
    for (int e = 0; e < W; e++) {
      pprev.v[e] = (TV)0;
      aup.v[e] = (TV)0;
    }
    VecW<unsigned, W> cm;
    double cs[W];
#pragma unroll
    for (int e = 0; e < W; e++) {
      cm.v[e] = 0u;
      cs[e] = 0.0;
    }
    if constexpr (CS) cm = *reinterpret_cast<const VecW<unsigned, W> *>(cmask + c * HP + h0);
#pragma unroll
    for (int k0 = 0; k0 < NLMAX; k0 += CH) {
      if (k0 >= L.nlay) break;
      VT a0[CH], a4[CH], a5[CH], a6[CH], pn[CH], po[CH], a4o[CH], a5b[CH], pb[CH], pf[CH];
      TV pe[CH], a4e[CH];
#pragma unroll
      for (int u = 0; u < CH; u++) {
        const int k = k0 + u;
        if (k < L.nlay) {
          const int pos = own + k * HP, ob = oth + k * HP;
          // the bottom layer's down plane is zero and its p below is not
          // used: both loads stay in range (own position)
          const int pdn = k + 1 < L.nlay ? pos + HP : pos;
          a0[u] = *reinterpret_cast<const VT *>(A + pos);
          a4[u] = *reinterpret_cast<const VT *>(A + 4 * NP + pos);
          a5[u] = *reinterpret_cast<const VT *>(A + 5 * NP + pos);
          a6[u] = *reinterpret_cast<const VT *>(A + 6 * NP + pos);
          pn[u] = *reinterpret_cast<const VT *>(p + pdn);
          po[u] = *reinterpret_cast<const VT *>(p + ob + h0);
          pe[u] = p[ob + ee];
          a4o[u] = *reinterpret_cast<const VT *>(A + 4 * NP + ob + h0);
          a4e[u] = A[4 * NP + ob + ee];
          a5b[u] = *reinterpret_cast<const VT *>(A + 5 * NP + ob + ib);
          pb[u] = *reinterpret_cast<const VT *>(p + ob + ib);
          pf[u] = *reinterpret_cast<const VT *>(p + ob + ifr);
        }
      }
#pragma unroll
      for (int u = 0; u < CH; u++) {
        const int k = k0 + u;
        if (k < L.nlay) {
          const bool dn = k + 1 < L.nlay;
          VT qv;
#pragma unroll
          for (int e = 0; e < W; e++) {
            // left: A(left, right) p(left); right: A(n, right) p(right)
            const TV al = sf ? a4o[u].v[e] : (e > 0 ? a4o[u].v[e > 0 ? e - 1 : 0] : a4e[u]);
            const TV pl = sf ? po[u].v[e] : (e > 0 ? po[u].v[e > 0 ? e - 1 : 0] : pe[u]);
            const TV pr = sf ? (e < W - 1 ? po[u].v[e < W - 1 ? e + 1 : e] : pe[u]) : po[u].v[e];
            TV tq = (TV)0;
            tq = tq + a0[u].v[e] * pcur.v[e];
            if (k > 0) tq = tq + aup.v[e] * pprev.v[e];
            if (hb) tq = tq + a5b[u].v[e] * pb[u].v[e];
            if (e > 0 || lok) tq = tq + al * pl;
            if (e < W - 1 || rok) tq = tq + a4[u].v[e] * pr;
            if (hf) tq = tq + a5[u].v[e] * pf[u].v[e];
            if (dn) tq = tq + a6[u].v[e] * pn[u].v[e];
            qv.v[e] = tq;
            if constexpr (CS)
              if ((cm.v[e] >> k) & 1u) cs[e] = cs[e] + (double)tq;
            s = s + (double)pcur.v[e] * (double)tq;
          }
          *reinterpret_cast<VT *>(q + own + k * HP) = qv;
          aup = a6[u];
          pprev = pcur;
          pcur = pn[u];
        }
      }
    }
    if constexpr (CS) {
      VecW<double, W> cv;
#pragma unroll
      for (int e = 0; e < W; e++) cv.v[e] = cs[e];
      *reinterpret_cast<VecW<double, W> *>(colsq + c * HP + h0) = cv;
    }
  }
  s = bsum(s, lds);
  if (threadIdx.x == 0) {
    part[P_PQ * nblk + lb] = s;
    for (int e = (int)gridDim.x + lb; e < nblk; e += (int)gridDim.x) part[P_PQ * nblk + e] = 0.0;
  }
}

// layers unrolled per chunk of the grouped SpMV (loads of a chunk are all
// issued before its arithmetic)
#ifndef MF6_SPMV_CH
#define MF6_SPMV_CH 2
#endif

// A launch of MF6_BLOCK threads; with e0 / e1 the events are recorded at the
// kernel's own start and end (hipExtLaunchKernelGGL), so their interval is
// the kernel's execution, as a profiler's kernel trace measures it.
template <typename... KArgs, typename... Args>
static void launch_timed(void (*kern)(KArgs...), int grid, hipStream_t s, hipEvent_t e0,
                         hipEvent_t e1, Args... args) {
  if (e0)
    hipExtLaunchKernelGGL(kern, dim3(grid), dim3(MF6_BLOCK), 0, s, e0, e1, 0, args...);
  else
    hipLaunchKernelGGL(kern, dim3(grid), dim3(MF6_BLOCK), 0, s, args...);
}

template <typename TV, int W, bool CS>
static void launch_f_spmvw(const Layout &L, int nb, hipStream_t s, const TV *A, const TV *p,
                           TV *q, double *part, const KState *st, const unsigned *cmask,
                           double *colsq, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
  int g = (2 * (L.HP / W) + MF6_BLOCK - 1) / MF6_BLOCK;
  if (g > nb) g = nb;
  if (g < 1) g = 1;
  constexpr int CH = MF6_SPMV_CH;
  if (L.nlay <= 4)
    launch_timed(k_f_spmvw<TV, W, 4, CH, CS>, g, s, e0, e1, L, A, p, q, part, st, nb, cmask,
                 colsq);
  else if (L.nlay <= 8)
    launch_timed(k_f_spmvw<TV, W, 8, CH, CS>, g, s, e0, e1, L, A, p, q, part, st, nb, cmask,
                 colsq);
  else if (L.nlay <= 16)
    launch_timed(k_f_spmvw<TV, W, 16, CH, CS>, g, s, e0, e1, L, A, p, q, part, st, nb, cmask,
                 colsq);
  else
    launch_timed(k_f_spmvw<TV, W, 32, CH, CS>, g, s, e0, e1, L, A, p, q, part, st, nb, cmask,
                 colsq);
}

// cmask / colsq given (the deflated two-level form): the SpMV also writes
// each column's sum of q over its coupled cells, colsq[c HP + h2] (colour
// layout, so that a group's sums are one store).
template <typename TM, typename TV, bool CS>
static void launch_f_spmv_t(const Layout &L, int nb, hipStream_t s, const TM *A, const TV *p,
                            TV *q, double *part, const KState *st, const unsigned *cmask,
                            double *colsq, hipEvent_t e0, hipEvent_t e1) {
  constexpr int WMAX = 16 / (int)sizeof(TV);
  const int w = col_group<TV>(L);
  if (w == WMAX) {
    launch_f_spmvw<TV, WMAX, CS>(L, nb, s, A, p, q, part, st, cmask, colsq, e0, e1);
    return;
  }
  if (w == 2) {
    launch_f_spmvw<TV, 2, CS>(L, nb, s, A, p, q, part, st, cmask, colsq, e0, e1);
    return;
  }
  if (L.nlay <= 4)
    launch_timed(k_f_spmv<TM, TV, 4, CS>, nb, s, e0, e1, L, A, p, q, part, st, nb, cmask, colsq);
  else if (L.nlay <= 8)
    launch_timed(k_f_spmv<TM, TV, 8, CS>, nb, s, e0, e1, L, A, p, q, part, st, nb, cmask, colsq);
  else if (L.nlay <= 16)
    launch_timed(k_f_spmv<TM, TV, 16, CS>, nb, s, e0, e1, L, A, p, q, part, st, nb, cmask,
                 colsq);
  else
    launch_timed(k_f_spmv<TM, TV, 32, CS>, nb, s, e0, e1, L, A, p, q, part, st, nb, cmask,
                 colsq);
}

template <typename TM, typename TV>
static void launch_f_spmv(const Layout &L, int nb, hipStream_t s, const TM *A, const TV *p,
                          TV *q, double *part, const KState *st,
                          const unsigned *cmask = nullptr, double *colsq = nullptr,
                          hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
  if (colsq)
    launch_f_spmv_t<TM, TV, true>(L, nb, s, A, p, q, part, st, cmask, colsq, e0, e1);
  else
    launch_f_spmv_t<TM, TV, false>(L, nb, s, A, p, q, part, st, nullptr, nullptr, e0, e1);
}

// x += apend_x p (the last iteration's head update), then no update pending
template <typename TV, typename TX>
__global__ void k_f_flush(int np, TX *__restrict__ x, const TV *__restrict__ p,
                          const FState *fs) {
  const TV ax = (TV)fs->apend_x;
  if (ax == (TV)0) return;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < np; e += gridDim.x * blockDim.x)
    x[e] = (TX)(x[e] + (TX)(ax * p[e]));
}

// ---------------------------------------------------------------------------
// scalar kernels
// ---------------------------------------------------------------------------

__device__ inline void testcnvg_(int icnvgopt, int *icnvg, int iiter, double dvmax, double rmax,
                                 double rmax0, double epfact, double dvclose, double rclose) {
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

// End of a round: test the iteration whose residual the preconditioner
// kernels just produced, then rho / beta of the next iteration.
// Colour partials sit at part[slot * nblk + b] for b < ncol (the colour
// kernels' grid).
// The one-block scalar kernels run MF6_RBLOCK threads, so each thread adds
// only a few of the (thousands of) block partials: their loads come from
// other XCDs' writes and the loop over them is latency-bound.
#define MF6_RBLOCK 1024

__global__ void __launch_bounds__(MF6_RBLOCK) k_f_round(KState *st, FState *fs,
                                                        const double *part,
                                                        const MaxLoc *mpart, int nblk,
                                                        int ncol, double *hist, int *hloc,
                                                        int ncoarse = 0, int nonsym = 0) {
  __shared__ double lds[4][MF6_RBLOCK / 64];
  __shared__ MaxLoc mlds[2][MF6_RBLOCK / 64];
  // the state is read before the partials (and the stop flag is only acted
  // on after them), so its load latency overlaps theirs
  const int stop = st->stop;
  KState S;
  double pmax = 0.0;
  int ploc = -1;
  if (threadIdx.x == 0) {
    S = *st;
    pmax = fs->pmax;
    ploc = fs->ploc;
  }
  // the six reductions in one pass over the partials (independent loads in
  // flight together); per slot the summation order is that of rsum/rmaxl
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  MaxLoc mm[2] = {{0.0, -1}, {0.0, -1}};
  for (int i = threadIdx.x; i < ncol; i += blockDim.x) {
    a[0] = a[0] + part[P_SS_R * nblk + i];
    a[1] = a[1] + part[P_SS_B * nblk + i];
    a[2] = a[2] + part[P_RHO_B * nblk + i];
    a[3] = a[3] + part[P_RHO_R * nblk + i];
    mm[0] = maxloc_pick(mm[0], mpart[M_D_R * nblk + i]);
    mm[1] = maxloc_pick(mm[1], mpart[M_D_B * nblk + i]);
  }
  // two-level preconditioner: rho gains rc.yc (partials of the coarse GEMV)
  for (int i = threadIdx.x; i < ncoarse; i += blockDim.x) a[3] = a[3] + part[P_RHO_C * nblk + i];
  const int w = threadIdx.x >> 6, ln = threadIdx.x & 63;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    a[q] = wsum(a[q]);
    if (ln == 0) lds[q][w] = a[q];
  }
#pragma unroll
  for (int q = 0; q < 2; q++) {
    mm[q] = wmax(mm[q]);
    if (ln == 0) mlds[q][w] = mm[q];
  }
  __syncthreads();
  if (threadIdx.x != 0 || stop) return;
  double tot[4] = {0.0, 0.0, 0.0, 0.0};
  MaxLoc mt[2] = {{0.0, -1}, {0.0, -1}};
  for (int v = 0; v < (int)(blockDim.x >> 6); v++) {
#pragma unroll
    for (int q = 0; q < 4; q++) tot[q] = tot[q] + lds[q][v];
    mt[0] = maxloc_pick(mt[0], mlds[0][v]);
    mt[1] = maxloc_pick(mt[1], mlds[1][v]);
  }
  const double ssr = tot[0], ssb = tot[1], rb = tot[2], rr = tot[3];
  MaxLoc md = maxloc_pick(mt[0], mt[1]);
  double rho_new = rb + rr;
  int it = S.iter;   // iteration whose residual was just formed (0: initial)
  fs->apend_d = 0.0;
  if (it > 0) {
    double l2norm = sqrt(ssr + ssb);
    double deltax = S.alpha * pmax;
    st->deltax = deltax;
    st->xloc = ploc;
    st->rmax = md.v;
    st->rloc = md.loc;
    st->l2norm = l2norm;
    if (it <= S.hist_cap) {
      int cap = S.hist_cap;
      hist[it - 1] = deltax;
      hist[cap + it - 1] = md.v;
      hist[2 * cap + it - 1] = S.alpha;
      hist[3 * cap + it - 1] = 0.0;
      hloc[it - 1] = ploc;
      hloc[cap + it - 1] = md.loc;
    }
    int o = S.icnvgopt;
    double rcnvg = (o == 2 || o == 3 || o == 4) ? l2norm : md.v;
    int icnvg = S.icnvg;
    testcnvg_(o, &icnvg, it, deltax, rcnvg, S.l2norm0, S.epfact, S.dvclose, S.rclose);
    if (rcnvg == 0.0) icnvg = 1;
    st->icnvg = icnvg;
    if (icnvg != 0) { st->stop = 1; return; }
    if (is_close_(S.rho, S.rho0)) { st->stop = 2; return; }
    if (S.rho == 0.0) { st->stop = 2; return; }
    st->rho0 = S.rho;
    if (it >= S.itmax) { st->stop = 3; return; }
  }
  st->rho = rho_new;
  if (it > 0) st->beta = rho_new / S.rho;   // rho0 was just set to the old rho
  if (fs->restart) {
    st->beta = 0.0;
    fs->restart = 0;
  }
  // The deflated preconditioner z = M^-1 (r - A Q r) + Q r (A-DEF1) is not
  // symmetric, so r.z need not keep the sign it has for a definite
  // preconditioner (negative here: the IMS matrices are negative definite,
  // and so is their ILU0 factor).  When rho changes sign from one iteration
  // to the next (beta <= 0) the direction restarts, p = z: its step alpha =
  // r.z / p.Ap is then the exact minimiser of the energy-norm error along p,
  // so the error does not grow.
  if (nonsym && it > 0 && rho_new * S.rho <= 0.0) st->beta = 0.0;
}

__global__ void __launch_bounds__(MF6_RBLOCK) k_f_alpha(KState *st, FState *fs,
                                                        const double *part,
                                                        const MaxLoc *mpart, int nblk) {
  __shared__ double lds[MF6_RBLOCK / 64];
  __shared__ MaxLoc mlds[MF6_RBLOCK / 64];
  const int stop = st->stop;
  double rho = 0.0;
  int iter = 0;
  if (threadIdx.x == 0) {
    rho = st->rho;
    iter = st->iter;
  }
  double a = 0.0;
  MaxLoc mm{0.0, -1};
  for (int i = threadIdx.x; i < nblk; i += blockDim.x) {
    a = a + part[P_PQ * nblk + i];
    mm = maxloc_pick(mm, mpart[M_P * nblk + i]);
  }
  const int w = threadIdx.x >> 6, ln = threadIdx.x & 63;
  a = wsum(a);
  mm = wmax(mm);
  if (ln == 0) { lds[w] = a; mlds[w] = mm; }
  __syncthreads();
  if (threadIdx.x != 0 || stop) return;
  double pq = 0.0;
  MaxLoc mp{0.0, -1};
  for (int v = 0; v < (int)(blockDim.x >> 6); v++) {
    pq = pq + lds[v];
    mp = maxloc_pick(mp, mlds[v]);
  }
  double den = pq;
  den = den + copysign(DPREC_F, den);
  st->pq = pq;
  st->alpha = rho / den;
  st->iter = iter + 1;
  fs->apend_x = st->alpha;
  fs->apend_d = st->alpha;
  fs->pmax = mp.v;
  fs->ploc = mp.loc;
}

// ---------------------------------------------------------------------------
// mixed precision: reliable update (residual replacement)
// ---------------------------------------------------------------------------

// x64 += xhat (FP32 correction), xhat = 0
__global__ void k_r_accumulate(int np, double *__restrict__ x, float *__restrict__ xhat) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < np; e += gridDim.x * blockDim.x) {
    x[e] = x[e] + (double)xhat[e];
    xhat[e] = 0.0f;
  }
}

// d32 = b - A x (FP64 row sums, amux order), partial max |r|
__global__ void __launch_bounds__(MF6_BLOCK) k_r_residual(Layout L, const double *__restrict__ A,
                                                          const double *__restrict__ x,
                                                          const double *__restrict__ b,
                                                          float *__restrict__ d32,
                                                          MaxLoc *mpart) {
  const int lb = xcd_block();
  __shared__ MaxLoc mlds[8];
  const int NP = L.NP;
  MaxLoc m{0.0, -1};
  const int ncolumns = L.HP0 + L.HP1;
  for (int t = lb * blockDim.x + threadIdx.x; t < ncolumns;
       t += gridDim.x * blockDim.x) {
    int c = t >= L.HP0;
    int h2 = t - c * L.HP0;
    int l = col_l(L, c, h2);
    int i = l / L.ncol, j = l - i * L.ncol;
    int ib = i > 0 ? hn(l, -L.ncol) : -1;
    int il = j > 0 ? hn(l, -1) : -1;
    int ir = j < L.ncol - 1 ? hn(l, 1) : -1;
    int ifr = i < L.nrow - 1 ? hn(l, L.ncol) : -1;
    const int own = c * L.H + h2;
    const int oth = (1 - c) * L.H;
    for (int k = 0; k < L.nlay; k++) {
      int pos = own + k * L.HP;
      int ob = oth + k * L.HP;
      double t2 = 0.0;
      t2 = t2 + A[pos] * x[pos];
      if (k > 0) t2 = t2 + A[NP + pos] * x[pos - L.HP];
      if (ib >= 0) t2 = t2 + A[2 * NP + pos] * x[ob + ib];
      if (il >= 0) t2 = t2 + A[3 * NP + pos] * x[ob + il];
      if (ir >= 0) t2 = t2 + A[4 * NP + pos] * x[ob + ir];
      if (ifr >= 0) t2 = t2 + A[5 * NP + pos] * x[ob + ifr];
      if (k < L.nlay - 1) t2 = t2 + A[6 * NP + pos] * x[pos + L.HP];
      double r = b[pos] - t2;
      d32[pos] = (float)r;
      m = maxloc_pick(m, MaxLoc{r, k * L.NRC + l});
    }
  }
  m = bmax(m, mlds);
  if (threadIdx.x == 0) mpart[lb] = m;
}

// after a replacement: no pending d / x update; a pending FP32 test result
// is re-examined on the true residual.  restart: the next direction is the
// preconditioned true residual alone (beta = 0).
__global__ void k_r_after(KState *st, FState *fs, const MaxLoc *mpart, int nblk, double *out,
                          int restart) {
  __shared__ MaxLoc mlds[8];
  MaxLoc m = rmaxl(mpart, nblk, mlds);
  if (threadIdx.x != 0) return;
  fs->apend_d = 0.0;
  fs->apend_x = 0.0;
  fs->restart = restart;
  out[0] = m.v;
  out[1] = (double)m.loc;
}

// ---------------------------------------------------------------------------
// host drivers
// ---------------------------------------------------------------------------

static inline int gridc(int n) {
  int g = (n + MF6_BLOCK - 1) / MF6_BLOCK;
  if (g > MF6_MAXBLK) g = MF6_MAXBLK;
  return g < 1 ? 1 : g;
}

// the grouped colour kernels (W columns per thread) on nbc blocks
template <typename TM, typename TV, typename TP, int W>
static void launch_pc_w(Mf6Handle *h, const TM *A, const TP *piv, TV *d, const TV *q, TV *zf,
                        TV *z, KState *st, FState *fs, int nbc, int nb, TV sc,
                        const unsigned *cm, double *cs, const TV *yc, const DeflPlanes &D) {
  const Layout &L = h->L;
  hipStream_t s = h->stream;
  if (yc) {
    k_f_red_fwdw<TM, TV, TP, true, W><<<nbc, MF6_BLOCK, 0, s>>>(
        L, A, piv, d, q, zf, h->d_part, h->d_mpart, st, fs, nb, yc, D.s, D.ew, D.ns, h->tl);
    k_f_blackw<TM, TV, TP, true, W><<<nbc, MF6_BLOCK, 0, s>>>(
        L, A, piv, d, q, zf, z, h->d_part, h->d_mpart, st, fs, nb, sc, cm, cs, yc, D.s, D.ew,
        D.ns, h->tl);
  } else {
    k_f_red_fwdw<TM, TV, TP, false, W><<<nbc, MF6_BLOCK, 0, s>>>(
        L, A, piv, d, q, zf, h->d_part, h->d_mpart, st, fs, nb, nullptr, nullptr, nullptr,
        nullptr, h->tl);
    k_f_blackw<TM, TV, TP, false, W><<<nbc, MF6_BLOCK, 0, s>>>(
        L, A, piv, d, q, zf, z, h->d_part, h->d_mpart, st, fs, nb, sc, cm, cs, nullptr, nullptr,
        nullptr, nullptr, h->tl);
  }
  k_f_red_bwdw<TM, TV, TP, W><<<nbc, MF6_BLOCK, 0, s>>>(L, A, piv, d, zf, z, h->d_part, st, nb,
                                                       sc, cm, cs);
}

// The colour kernels run on nbc blocks (one thread per column of one
// colour); their partials use the slot stride nb of the interleaved kernels.
// A grid sized for both colours would leave the XCDs that receive the upper
// half of the logical blocks idle.
// With the two-level preconditioner the black and red backward kernels also
// write each column's sum of the residual over its coupled cells (colsum,
// natural column order), the first stage of the restriction.
template <typename TM, typename TV, typename TP>
static void launch_pc(Mf6Handle *h, const TM *A, const TP *piv, TV *d, const TV *q, TV *zf,
                      TV *z, KState *st, FState *fs, int nbc, int nb, TV sc = (TV)1) {
  const Layout &L = h->L;
  hipStream_t s = h->stream;
  const bool tl = h->s.preconditioner >= 2;
  const unsigned *cm = tl ? h->tl_cmask : nullptr;
  double *cs = tl ? h->tl_colsum : nullptr;
  // deflated form: the sweeps take r - A P yc
  const TV *yc = nullptr;
  DeflPlanes D;
  if (h->s.preconditioner == 3) {
    yc = sizeof(TV) == sizeof(double) ? (const TV *)h->tl_yc64 : (const TV *)h->tl_yc32;
    D.s = h->tl_s32;
    D.ew = h->tl_ew;
    D.ns = h->tl_ns;
  }
  const int w = col_group<TV>(L);
  constexpr int WMAX = 16 / (int)sizeof(TV);
  if (w == WMAX) {
    launch_pc_w<TM, TV, TP, WMAX>(h, A, piv, d, q, zf, z, st, fs, nbc, nb, sc, cm, cs, yc, D);
    return;
  }
  if (w == 2) {
    launch_pc_w<TM, TV, TP, 2>(h, A, piv, d, q, zf, z, st, fs, nbc, nb, sc, cm, cs, yc, D);
    return;
  }
  if (yc) {
    k_f_red_fwd<TM, TV, TP, true><<<nbc, MF6_BLOCK, 0, s>>>(
        L, A, piv, d, q, zf, h->d_part, h->d_mpart, st, fs, nb, yc, D.s, D.ew, D.ns, h->tl);
    k_f_black<TM, TV, TP, true><<<nbc, MF6_BLOCK, 0, s>>>(L, A, piv, d, q, zf, z, h->d_part,
                                                          h->d_mpart, st, fs, nb, sc, cm, cs,
                                                          yc, D.s, D.ew, D.ns, h->tl);
  } else {
    k_f_red_fwd<TM, TV, TP, false><<<nbc, MF6_BLOCK, 0, s>>>(
        L, A, piv, d, q, zf, h->d_part, h->d_mpart, st, fs, nb, nullptr, nullptr, nullptr,
        nullptr, h->tl);
    k_f_black<TM, TV, TP, false><<<nbc, MF6_BLOCK, 0, s>>>(L, A, piv, d, q, zf, z, h->d_part,
                                                           h->d_mpart, st, fs, nb, sc, cm, cs,
                                                           nullptr, nullptr, nullptr, nullptr,
                                                           h->tl);
  }
  k_f_red_bwd<TM, TV, TP><<<nbc, MF6_BLOCK, 0, s>>>(L, A, piv, d, zf, z, h->d_part, st, nb, sc,
                                                    cm, cs);
}

static int read_states(Mf6Handle *h) {
  HIPCHK(hipMemcpyAsync(h->h_state, h->d_state, 3 * sizeof(KState), hipMemcpyDeviceToHost,
                        h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return MF6HIP_OK;
}

// Enqueue n fused FP64 CG iterations starting at an even round (the two
// direction buffers alternate; n even keeps the parity).
static void enqueue_fused64(Mf6Handle *h, int n, int nbc, int nb) {
  const Layout &L = h->L;
  hipStream_t s = h->stream;
  KState *st = h->d_state;
  FState *fs = reinterpret_cast<FState *>(h->d_state + 2);
  double *pbuf[2] = {h->p, h->p2};
  const bool tl = h->s.preconditioner >= 2, defl = h->s.preconditioner == 3;
  for (int round = 0; round < n; round++) {
    // additive form: the coarse correction of the residual the sweeps
    // produce; deflated form: of the residual before the sweeps, from the
    // last sweeps' column sums and the SpMV's (P^T r - alpha P^T q)
    if (defl)
      mf6hip_tl_enqueue<double>(h, (const double *)nullptr, h->tl_yc64, h->d_part + P_RHO_C * nb,
                                nb, st, &fs->apend_d);
    launch_pc<double, double, double>(h, h->A, h->invpiv, h->d, h->q, h->zf, h->z, st, fs, nbc,
                                      nb);
    if (tl && !defl)
      mf6hip_tl_enqueue<double>(h, (const double *)nullptr, h->tl_yc64, h->d_part + P_RHO_C * nb,
                                nb, st);
    k_f_round<<<1, MF6_RBLOCK, 0, s>>>(st, fs, h->d_part, h->d_mpart, nb, nbc, h->d_hist,
                                      h->d_hloc, tl ? mf6hip_tl_nslot(h, nb) : 0, defl ? 1 : 0);
    double *pin = pbuf[round & 1], *pout = pbuf[(round + 1) & 1];
    launch_f_pupdate<double, double>(L, nb, s, h->z, pin, pout, h->x, h->d_mpart, st, fs,
                                     tl ? h->tl_yc64 : nullptr, h->tl_cmask, h->tl);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (h->timing && (h->spmv_calls++ % h->spmv_sample) == 0 &&
        2 * (h->last_spmv_n + 1) <= (int)h->spmv_ev.size()) {
      const int slot = h->last_spmv_n++;
      e0 = h->spmv_ev[2 * slot];
      e1 = h->spmv_ev[2 * slot + 1];
    }
    launch_f_spmv<double, double>(L, nb, s, h->A, pout, h->q, h->d_part, st,
                                  defl ? h->tl_cmask : nullptr, defl ? h->tl_colsq : nullptr,
                                  e0, e1);
    k_f_alpha<<<1, MF6_RBLOCK, 0, s>>>(st, fs, h->d_part, h->d_mpart, nb);
  }
}

// Capture FUSED_GRAPH_ITERS iterations as a graph (once per handle; the
// buffers and grid sizes are fixed for the handle's life).  The graph holds
// kernel launches only.
constexpr int FUSED_GRAPH_ITERS = 8;

static bool fused64_graph(Mf6Handle *h, int nbc, int nb) {
  if (h->cg_graph) return true;
  if (h->cg_graph_failed) return false;
  hipStream_t s = h->stream;
  hipGraph_t g = nullptr;
  bool ok = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess;
  if (ok) {
    enqueue_fused64(h, FUSED_GRAPH_ITERS, nbc, nb);
    ok = hipStreamEndCapture(s, &g) == hipSuccess && g != nullptr;
  }
  if (ok) ok = hipGraphInstantiate(&h->cg_graph, g, nullptr, nullptr, 0) == hipSuccess;
  if (g) (void)hipGraphDestroy(g);
  if (!ok) {
    h->cg_graph = nullptr;
    h->cg_graph_failed = true;
    (void)hipGetLastError();
  }
  return ok;
}

// Fused FP64 CG on the device system.  Requires: d = b - A x in h->d,
// l2norm0 in state[0], q zeroed, A symmetric.  Without the per-launch
// timing the iterations are replayed from a captured graph.
int mf6hip_run_cg_fused(Mf6Handle *h) {
  const Layout &L = h->L;
  hipStream_t s = h->stream;
  FState *fs = reinterpret_cast<FState *>(h->d_state + 2);
  HIPCHK(hipMemsetAsync(fs, 0, sizeof(FState), s));
  HIPCHK(hipMemsetAsync(h->q, 0, sizeof(double) * L.NP, s));
  if (h->s.preconditioner == 3) mf6hip_tl_colsum<double>(h, h->d);
  // one grid size for every kernel of a round, so that each partial array
  // has the same number of slots for the round kernel to reduce
  const int nbc = colour_grid<double>(L);
  const int nbs = gridc(2 * L.HP0);
  const int nb = nbs > nbc ? nbs : nbc;
  double *pbuf[2] = {h->p, h->p2};
  const bool graph = !h->timing && fused64_graph(h, nbc, nb);
  int batch = graph ? FUSED_GRAPH_ITERS : 4;
  // round 0: preconditioner of the initial residual; batches are even
  while (true) {
    if (graph) {
      for (int r = 0; r < batch / FUSED_GRAPH_ITERS; r++) HIPCHK(hipGraphLaunch(h->cg_graph, s));
    } else {
      enqueue_fused64(h, batch, nbc, nb);
    }
    HIPCHK(hipGetLastError());
    CK(read_states(h));
    if (h->h_state->stop) break;
    if (batch < 64) batch *= 2;
  }
  // apply the pending head update of the last iteration: the direction of
  // iteration m lives in pbuf[m & 1]
  int m = h->h_state->iter;
  if (m > 0) {
    k_f_flush<double, double><<<gridc(L.NP), MF6_BLOCK, 0, s>>>(L.NP, h->x, pbuf[m & 1], fs);
    HIPCHK(hipGetLastError());
  }
  return MF6HIP_OK;
}

// Mixed precision: FP32 fused CG with reliable updates.  Requires the FP32
// planes (A32) and pivots; x (FP64) holds the start heads, b the rhs.
template <typename TMP, typename TP>
static int run_cg_reliable(Mf6Handle *h, const TMP *Apc, const TP *piv32, float sc, int *icnvg,
                           int *iters) {
  const Layout &L = h->L;
  hipStream_t s = h->stream;
  KState *st = h->d_state;
  FState *fs = reinterpret_cast<FState *>(h->d_state + 2);
  const int nbc = colour_grid<float>(L);
  const int nbs = gridc(2 * L.HP0);
  const int nb = nbs > nbc ? nbs : nbc;
  float *pbuf[2] = {h->p32, h->p232};
  double *rout = h->d_part + 7 * MF6_MAXBLK;
  // iterations between replacements.  With the two-level preconditioner
  // the direction restarts at every replacement (keeping it, as the
  // one-level solve does, lets the FP32 recurrences stall against the
  // coarse correction), so replacements come more often.
  const int R = h->s.inner_fp32_max > 0 ? h->s.inner_fp32_max
                                        : (h->s.preconditioner >= 2 ? 32 : 64);
  HIPCHK(hipMemsetAsync(fs, 0, sizeof(FState), s));
  HIPCHK(hipMemsetAsync(h->q32, 0, sizeof(float) * L.NP, s));
  HIPCHK(hipMemsetAsync(h->e32, 0, sizeof(float) * L.NP, s));
  // true residual into d32 (r32)
  k_r_residual<<<nb, MF6_BLOCK, 0, s>>>(L, h->A, h->x, h->b, h->r32, h->d_mpart);
  k_r_after<<<1, MF6_BLOCK, 0, s>>>(st, fs, h->d_mpart, nb, rout, 0);
  int round = 0, last_replace = 0;
  int total_budget = h->s.iter1;
  bool done = false;
  const bool tl = h->s.preconditioner >= 2, defl = h->s.preconditioner == 3;
  if (defl) mf6hip_tl_colsum<float>(h, h->r32);
  *icnvg = 0;
  while (!done) {
    int batch = 8;
    while (true) {
      for (int it = 0; it < batch; it++, round++) {
        if (defl)
          mf6hip_tl_enqueue<float>(h, (const float *)nullptr, h->tl_yc32,
                                   h->d_part + P_RHO_C * nb, nb, st, &fs->apend_d);
        launch_pc<TMP, float, TP>(h, Apc, piv32, h->r32, h->q32, h->zf32, h->z32, st, fs, nbc,
                                  nb, sc);
        if (tl && !defl)
          mf6hip_tl_enqueue<float>(h, (const float *)nullptr, h->tl_yc32,
                                   h->d_part + P_RHO_C * nb, nb, st);
        k_f_round<<<1, MF6_RBLOCK, 0, s>>>(st, fs, h->d_part, h->d_mpart, nb, nbc, h->d_hist,
                                          h->d_hloc, tl ? mf6hip_tl_nslot(h, nb) : 0, defl ? 1 : 0);
        float *pin = pbuf[round & 1], *pout = pbuf[(round + 1) & 1];
        launch_f_pupdate<float, float>(L, nb, s, h->z32, pin, pout, h->e32, h->d_mpart, st, fs,
                                       tl ? h->tl_yc32 : nullptr, h->tl_cmask, h->tl);
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (h->timing && (h->spmv_calls++ % h->spmv_sample) == 0 &&
            2 * (h->last_spmv_n + 1) <= (int)h->spmv_ev.size()) {
          const int slot = h->last_spmv_n++;
          e0 = h->spmv_ev[2 * slot];
          e1 = h->spmv_ev[2 * slot + 1];
        }
        launch_f_spmv<float, float>(L, nb, s, h->A32, pout, h->q32, h->d_part, st,
                                    defl ? h->tl_cmask : nullptr, defl ? h->tl_colsq : nullptr,
                                    e0, e1);
        k_f_alpha<<<1, MF6_RBLOCK, 0, s>>>(st, fs, h->d_part, h->d_mpart, nb);
      }
      HIPCHK(hipGetLastError());
      CK(read_states(h));
      if (h->h_state->stop) break;
      if (h->h_state->iter - last_replace >= R) break;
      if (batch < 32) batch *= 2;
    }
    KState hs = h->h_state[0];
    int m = hs.iter;
    // replace the residual: x += xhat + pending alpha p ; d32 = b - A x
    // the head update of iteration m (direction p_m) is still pending, and
    // so is alpha q for d when no test ran yet: fold both into the FP64
    // heads and recompute the residual from them
    if (m > 0)
      k_f_flush<float, float><<<gridc(L.NP), MF6_BLOCK, 0, s>>>(L.NP, h->e32, pbuf[m & 1], fs);
    round = m;
    k_r_accumulate<<<gridc(L.NP), MF6_BLOCK, 0, s>>>(L.NP, h->x, h->e32);
    k_r_residual<<<nb, MF6_BLOCK, 0, s>>>(L, h->A, h->x, h->b, h->r32, h->d_mpart);
    k_r_after<<<1, MF6_BLOCK, 0, s>>>(st, fs, h->d_mpart, nb, rout, tl ? 1 : 0);
    if (defl) mf6hip_tl_colsum<float>(h, h->r32);
    double hr[2];
    HIPCHK(hipMemcpyAsync(hr, rout, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    last_replace = m;
    if (hs.stop != 0) {
      bool true_ok = fabs(hr[0]) <= h->s.rclose;
      if (hs.stop == 1 && true_ok) { *icnvg = 1; done = true; break; }
      if (hs.stop == 3 || m >= total_budget) { done = true; break; }
      if (hs.stop == 2 && true_ok && fabs(hs.deltax) <= h->s.dvclose) { *icnvg = 1; done = true; break; }
      if (hs.stop == 2) { done = true; break; }
      // FP32 test passed but the true residual did not: continue from the
      // replaced residual, re-testing this iteration with it
      h->h_state[0].stop = 0;
      h->h_state[0].icnvg = 0;
      h->h_state[0].iter = m;
      HIPCHK(hipMemcpyAsync(st, h->h_state, sizeof(KState), hipMemcpyHostToDevice, s));
    }
  }
  *iters = h->h_state[0].iter;
  return MF6HIP_OK;
}

int mf6hip_run_cg_reliable(Mf6Handle *h, int *icnvg, int *iters) {
  if (h->s.precision == MF6HIP_MIXED)
    return run_cg_reliable<float, float>(h, h->A32, h->invpiv32, 1.0f, icnvg, iters);
  // MIXED_FP16: FP16 preconditioner planes (scaled), FP32 SpMV planes
  return run_cg_reliable<_Float16, float>(h, h->A16, h->invpiv32s, h->pc_scale, icnvg, iters);
}

// bitwise symmetry of the planes: A(n, up) == A(up, n), A(n, back) ==
// A(back, n), A(n, left) == A(left, n)
__global__ void k_check_sym(Layout L, const double *__restrict__ A, int *flag) {
  const int NP = L.NP;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < NP; p += gridDim.x * blockDim.x) {
    int c, k, i, j, l;
    if (!pos_to_cell(L, p, c, k, i, j, l)) continue;
    const int ob = (1 - c) * L.H + k * L.HP;
    bool ok = true;
    if (k > 0) ok = ok && (A[NP + p] == A[6 * NP + p - L.HP]);
    if (i > 0) ok = ok && (A[2 * NP + p] == A[5 * NP + ob + hn(l, -L.ncol)]);
    if (j > 0) ok = ok && (A[3 * NP + p] == A[4 * NP + ob + hn(l, -1)]);
    if (!ok) atomicOr(flag, 1);
  }
}

int mf6hip_is_symmetric(Mf6Handle *h, int *sym) {
  const Layout &L = h->L;
  hipStream_t s = h->stream;
  int *flag = h->d_fail + 1;
  HIPCHK(hipMemsetAsync(flag, 0, sizeof(int), s));
  k_check_sym<<<gridc(L.NP), MF6_BLOCK, 0, s>>>(L, h->A, flag);
  int hf = 0;
  HIPCHK(hipMemcpyAsync(&hf, flag, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  *sym = hf == 0;
  return MF6HIP_OK;
}

// ---------------------------------------------------------------------------
// parity helpers: the fused CG's operators on device vectors (positions)
// ---------------------------------------------------------------------------

template <typename TI, typename TO>
__global__ void k_op_convert(int n, const TI *__restrict__ in, TO *__restrict__ out) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x)
    out[e] = (TO)in[e];
}

// op 4: y = A x with the fused CG's SpMV on the FP64 planes;
// op 5: y = A32 x32 with the fused CG's SpMV in FP32 (A and x rounded to
//       FP32; needs a mixed-precision handle);
// op 6: y = the deflated two-level preconditioner applied to x (FP64 handle,
//       preconditioner 3): yc = (P^T A P)^-1 P^T x, y = M^-1 (x - A P yc) +
//       P yc, M the ILU0 factor, through the fused CG's sweeps;
// op 7: y = M^-1 x through the fused CG's sweeps (FP64, preconditioner 0);
// op 8: the same in the mixed handle's precision (FP32 planes and pivots,
//       or the scaled FP16 planes with FP32 pivots), x rounded to FP32.
// x and y are device vectors in positions; y may not alias x.  A state slot
// of its own (index 3) keeps the solve's states untouched.
int mf6hip_fused_op(Mf6Handle *h, int op, const double *x, double *y) {
  const Layout &L = h->L;
  hipStream_t s = h->stream;
  const int nbc = colour_grid<double>(L);
  const int nbs = gridc(2 * L.HP0);
  const int nb = nbs > nbc ? nbs : nbc;
  KState *st = h->d_state + 3;
  HIPCHK(hipMemsetAsync(st, 0, sizeof(KState), s));
  if (op == 4) {
    launch_f_spmv<double, double>(L, nb, s, h->A, x, y, h->d_part, st);
  } else if (op == 5) {
    if (!h->A32 || !h->r32 || !h->q32) {
      mf6hip_set_error("op 5 needs a mixed-precision handle");
      return MF6HIP_EINVAL;
    }
    k_op_convert<double, float><<<gridc(MF6_NSLOT * L.NP), MF6_BLOCK, 0, s>>>(
        MF6_NSLOT * L.NP, h->A, h->A32);
    k_op_convert<double, float><<<gridc(L.NP), MF6_BLOCK, 0, s>>>(L.NP, x, h->r32);
    launch_f_spmv<float, float>(L, nb, s, h->A32, h->r32, h->q32, h->d_part, st);
    k_op_convert<float, double><<<gridc(L.NP), MF6_BLOCK, 0, s>>>(L.NP, h->q32, y);
  } else if (op == 6) {
    if (h->s.preconditioner != 3 || h->s.precision != MF6HIP_FP64 || !h->tl_cmask ||
        !h->tl_yc64) {
      mf6hip_set_error("op 6 needs an FP64 handle with preconditioner 3 and a loaded system");
      return MF6HIP_EINVAL;
    }
    // the sweeps take d - alpha q with alpha = 0 (pending update of a fresh
    // state), q = 0 and column sums of q = 0
    FState *fz = reinterpret_cast<FState *>(h->d_part + 7 * (size_t)MF6_MAXBLK + 64);
    HIPCHK(hipMemsetAsync(fz, 0, sizeof(FState), s));
    HIPCHK(hipMemcpyAsync(h->d, x, sizeof(double) * L.NP, hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemsetAsync(h->q, 0, sizeof(double) * L.NP, s));
    HIPCHK(hipMemsetAsync(h->tl_colsq, 0, sizeof(double) * 2 * (size_t)L.HP, s));
    mf6hip_tl_colsum<double>(h, h->d);
    mf6hip_tl_enqueue<double>(h, (const double *)nullptr, h->tl_yc64, h->d_part + P_RHO_C * nb,
                              nb, st, &fz->apend_d);
    launch_pc<double, double, double>(h, h->A, h->invpiv, h->d, h->q, h->zf, y, st, fz, nbc,
                                      nb);
    CK(mf6hip_tl_prolong64(h, y));
  } else if (op == 7 || op == 8) {
    if (h->s.preconditioner != 0) {
      mf6hip_set_error("ops 7 and 8 need a handle with preconditioner 0");
      return MF6HIP_EINVAL;
    }
    FState *fz = reinterpret_cast<FState *>(h->d_part + 7 * (size_t)MF6_MAXBLK + 64);
    HIPCHK(hipMemsetAsync(fz, 0, sizeof(FState), s));
    if (op == 7) {
      HIPCHK(hipMemcpyAsync(h->d, x, sizeof(double) * L.NP, hipMemcpyDeviceToDevice, s));
      HIPCHK(hipMemsetAsync(h->q, 0, sizeof(double) * L.NP, s));
      launch_pc<double, double, double>(h, h->A, h->invpiv, h->d, h->q, h->zf, y, st, fz, nbc,
                                        nb);
    } else {
      if (h->s.precision == MF6HIP_FP64) {
        mf6hip_set_error("op 8 needs a mixed-precision handle");
        return MF6HIP_EINVAL;
      }
      const int nbc32 = colour_grid<float>(L);
      const int nbs32 = gridc(2 * L.HP0);
      const int nb32 = nbs32 > nbc32 ? nbs32 : nbc32;
      k_op_convert<double, float><<<gridc(L.NP), MF6_BLOCK, 0, s>>>(L.NP, x, h->r32);
      HIPCHK(hipMemsetAsync(h->q32, 0, sizeof(float) * L.NP, s));
      if (h->s.precision == MF6HIP_MIXED) {
        k_op_convert<double, float><<<gridc(MF6_NSLOT * L.NP), MF6_BLOCK, 0, s>>>(
            MF6_NSLOT * L.NP, h->A, h->A32);
        launch_pc<float, float, float>(h, h->A32, h->invpiv32, h->r32, h->q32, h->zf32, h->z32,
                                       st, fz, nbc32, nb32);
      } else {
        launch_pc<_Float16, float, float>(h, h->A16, h->invpiv32s, h->r32, h->q32, h->zf32,
                                          h->z32, st, fz, nbc32, nb32, h->pc_scale);
      }
      k_op_convert<float, double><<<gridc(L.NP), MF6_BLOCK, 0, s>>>(L.NP, h->z32, y);
    }
  } else {
    mf6hip_set_error("unknown fused operator");
    return MF6HIP_EINVAL;
  }
  HIPCHK(hipGetLastError());
  return MF6HIP_OK;
}
