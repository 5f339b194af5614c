// mf6hip_fcg.hip -- flexible preconditioned CG with the learned correction.
//
// Preconditioner (variable, hence flexible CG), a two-level cycle with the
// network as the coarse correction:
//   z  = ILU0(r)                      pre-smoothing: the device ILU0
//   t  = r - A z                      what the ILU0 step leaves
//   y  = s * net([t / s, features])   s = rms(t): the network sees unit-scale
//                                     residuals; its output is rescaled
//   z += w y,  w = y.t / y.A y        energy-optimal step along y
//   z += ILU0(r - A z)                post-smoothing: damps the rough part of
//                                     the correction (BF16 noise of y)
// FCG(m) (Notay 2000): p = z - sum_j (z.q_j / p_j.q_j) p_j over the last m
// directions, alpha = p.r / p.q, x += alpha p, r -= alpha q.  The IMS
// convergence test (ImsLinearBase.f90:1101-1146) is applied to every
// iteration with the signed largest head change alpha p and residual; when
// it passes, the FP64 true residual b - A x is recomputed and must pass it
// too (else the iteration continues from the true residual).  The network
// runs on every nn_every-th iteration (ILU0 alone on the others).
// Fallback: the ILU0-only step replaces the learned one for the rest of the
// solve when z.r changes sign against the ILU0 step (not a descent
// direction) or when ten applications of the network (10 nn_every
// iterations) reduce ||r||_2 by less than 10 %.
//
// Every scalar of the iteration lives in a device FcgState: the host
// enqueues the kernels of a group of iterations without waiting and reads
// the state once per group; once the solve has stopped (converged) the
// state-changing kernels of the rest of the group return at once.
#include <cmath>
#include <vector>

#include "mf6hip_handle.h"

namespace {

constexpr int FCG_MAXM = 64;   // directions kept, at most
constexpr int FCG_NDOT = 8;    // dot products per pass
constexpr int FCG_GROUP = 8;   // iterations enqueued between state reads

struct FcgState {
  double tt, zr_ilu, yt, yay, yr, w;   // learned step
  double zr_l;                         // z.r of the smoothed learned step
  double pr;                           // p.r
  double rnorm, dvmax, rmax;
  double rn_window;
  double c[FCG_MAXM];                  // z.q_j, by ring slot
  double pqs[FCG_MAXM];                // p_j.q_j, by ring slot
  int learned;                         // network steps still allowed
  int cand;                            // the IMS test passed: check b - A x
  int icnvg, iter, learned_iters, window;
};

struct FcgSlots {
  int n;
  int slot[FCG_MAXM];
};

struct FcgTest {
  double dvclose, rclose;
  int icnvgopt, nn_every;
};

__device__ inline double wave_sum(double s) {
  for (int o = 32; o > 0; o >>= 1) s = s + __shfl_xor(s, o);
  return s;
}

__global__ void k_fcg_init(FcgState *fs, KState *ks, int learned) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  fs->learned = learned;
  fs->cand = 0;
  fs->icnvg = 0;
  fs->iter = 0;
  fs->learned_iters = 0;
  fs->window = 0;
  fs->rn_window = -1.0;
  for (int j = 0; j < FCG_MAXM; j++) { fs->c[j] = 0.0; fs->pqs[j] = 0.0; }
  ks->stop = 0;
}

// t = r - t
__global__ void k_fcg_sub(int n, const double *__restrict__ r, double *__restrict__ t,
                          const KState *ks) {
  if (ks->stop) return;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x)
    t[e] = r[e] - t[e];
}

// network input (BF16, 16 channels, natural order): ch0 = t / s with
// s = rms(t) = sqrt(tt / N), ch1..2 = features, ch3..15 = 0
__global__ void k_fcg_nn_input(Layout L, const double *__restrict__ t, const FcgState *fs,
                               const float *__restrict__ feat, unsigned short *__restrict__ in0) {
  const double tt = fs->tt;
  const double inv_s = tt > 0.0 ? sqrt((double)L.N / tt) : 0.0;
  const int n16 = L.N * 16;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += gridDim.x * blockDim.x) {
    const int c = i & 15, n = i >> 4;
    float v = 0.f;
    if (c == 0) v = (float)(t[pos_of_natural(L, n)] * inv_s);
    else if (c < 3) v = feat[(size_t)(c - 1) * L.N + n];
    unsigned u = __float_as_uint(v);
    u = u + 0x7fffu + ((u >> 16) & 1u);
    in0[i] = (unsigned short)(u >> 16);
  }
}

// yv (positions) = s * y (natural order); padding slots 0
__global__ void k_fcg_nn_out(Layout L, const float *__restrict__ y, const FcgState *fs,
                             double *__restrict__ yv) {
  const double tt = fs->tt;
  const double sc = tt > 0.0 ? sqrt(tt / (double)L.N) : 0.0;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < L.NP; e += gridDim.x * blockDim.x) {
    int c, k, i, j, l;
    yv[e] = pos_to_cell(L, e, c, k, i, j, l) ? sc * (double)y[natural(L, k, i, j)] : 0.0;
  }
}

// the weight of the learned step: w = y.t / y.A y, or 0 when the network is
// off or the step is degenerate
__global__ void k_fcg_decide(FcgState *fs, const KState *ks) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double w = 0.0;
  if (!ks->stop && fs->learned && fs->tt > 0.0 && fs->yay != 0.0 && isfinite(fs->yt) &&
      isfinite(fs->yay))
    w = fs->yt / fs->yay;
  fs->w = w;
}

// zl = z + w y (zl = z when the step was declined: the post-smoothing that
// follows then works on defined values even though its result is dropped)
__global__ void k_fcg_axpy_w(int n, const FcgState *fs, const double *__restrict__ z,
                             const double *__restrict__ y, double *__restrict__ zl) {
  const double w = fs->w;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x)
    zl[e] = w == 0.0 ? z[e] : z[e] + w * y[e];
}

// zl += zs (post-smoothing correction)
__global__ void k_fcg_add(int n, const FcgState *fs, const double *__restrict__ zs,
                          double *__restrict__ zl) {
  if (fs->w == 0.0) return;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x)
    zl[e] = zl[e] + zs[e];
}

// keep the learned z (zl.r has the sign of the ILU0 step's z.r: a descent
// direction) or fall back to the ILU0 step for the rest of the solve
__global__ void k_fcg_accept(FcgState *fs, const KState *ks) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (ks->stop || fs->w == 0.0) { fs->w = 0.0; return; }
  if (fs->zr_l * fs->zr_ilu > 0.0 && isfinite(fs->zr_l)) {
    fs->learned_iters++;
  } else {
    fs->w = 0.0;
    fs->learned = 0;
  }
}

// z = zl when the learned step was accepted
__global__ void k_fcg_select(int n, const FcgState *fs, const double *__restrict__ zl,
                             double *__restrict__ z) {
  if (fs->w == 0.0) return;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x)
    z[e] = zl[e];
}

// up to FCG_NDOT dot products a . b_k in one pass (b_k = B + slot_k * ld,
// or the explicit pointers when B is null): part[k * gridDim.x + block]
struct DotArgs {
  int n;
  const double *b[FCG_NDOT];
};

__global__ void __launch_bounds__(MF6_BLOCK) k_fcg_mdots(int n, const double *__restrict__ a,
                                                         DotArgs d, double *part,
                                                         const KState *ks) {
  if (ks->stop) return;
  __shared__ double lds[FCG_NDOT][MF6_BLOCK / 64];
  double s[FCG_NDOT];
#pragma unroll
  for (int k = 0; k < FCG_NDOT; k++) s[k] = 0.0;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    const double av = a[e];
#pragma unroll
    for (int k = 0; k < FCG_NDOT; k++)
      if (k < d.n) s[k] = s[k] + av * d.b[k][e];
  }
  const int w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < FCG_NDOT; k++) {
    const double t = wave_sum(s[k]);
    if ((threadIdx.x & 63) == 0) lds[k][w] = t;
  }
  __syncthreads();
  if (threadIdx.x < d.n) {
    double t = 0.0;
    for (int v = 0; v < (int)(blockDim.x >> 6); v++) t = t + lds[threadIdx.x][v];
    part[threadIdx.x * gridDim.x + blockIdx.x] = t;
  }
}

// pairs of vectors: part[k * gridDim.x + block] = sum a_k . b_k, k < 3
__global__ void __launch_bounds__(MF6_BLOCK) k_fcg_dots(int n, int ndot, const double *a0,
                                                        const double *b0, const double *a1,
                                                        const double *b1, const double *a2,
                                                        const double *b2, double *part,
                                                        const KState *ks) {
  if (ks->stop) return;
  __shared__ double lds[3][MF6_BLOCK / 64];
  double s[3] = {0.0, 0.0, 0.0};
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    s[0] = s[0] + a0[e] * b0[e];
    if (ndot > 1) s[1] = s[1] + a1[e] * b1[e];
    if (ndot > 2) s[2] = s[2] + a2[e] * b2[e];
  }
  const int w = threadIdx.x >> 6;
  for (int k = 0; k < 3; k++) {
    const double t = wave_sum(s[k]);
    if ((threadIdx.x & 63) == 0) lds[k][w] = t;
  }
  __syncthreads();
  if (threadIdx.x < ndot) {
    double t = 0.0;
    for (int v = 0; v < (int)(blockDim.x >> 6); v++) t = t + lds[threadIdx.x][v];
    part[threadIdx.x * gridDim.x + blockIdx.x] = t;
  }
}

// out_k = sum of the nb partials of dot k (fixed order), k < ndot
struct SumOut {
  double *o[FCG_NDOT];
};

__global__ void __launch_bounds__(MF6_BLOCK) k_fcg_sum(const double *part, int nb, int ndot,
                                                       SumOut out, const KState *ks) {
  if (ks->stop) return;
  __shared__ double lds[MF6_BLOCK / 64];
  for (int k = 0; k < ndot; k++) {
    double s = 0.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) s = s + part[k * nb + i];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int v = 0; v < (int)(blockDim.x >> 6); v++) t = t + lds[v];
      *out.o[k] = t;
    }
    __syncthreads();
  }
}

// p = z - sum_j (c_j / pq_j) P_j over the stored slots
__global__ void k_fcg_dir(int n, const double *__restrict__ z, const double *__restrict__ P,
                          size_t ld, FcgSlots sl, const FcgState *fs, double *__restrict__ p,
                          const KState *ks) {
  if (ks->stop) return;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    double v = z[e];
    for (int j = 0; j < sl.n; j++) {
      const int k = sl.slot[j];
      const double pq = fs->pqs[k];
      if (pq != 0.0) v = v - (fs->c[k] / pq) * P[(size_t)k * ld + e];
    }
    p[e] = v;
  }
}

// x += alpha p ; r -= alpha q with alpha = p.r / p.q ; partials: sum r^2,
// signed max |alpha p| and max |r| with natural locations (first occurrence
// on ties)
__global__ void __launch_bounds__(MF6_BLOCK) k_fcg_update(Layout L, const FcgState *fs, int head,
                                                          double *__restrict__ x,
                                                          const double *__restrict__ p,
                                                          double *__restrict__ r,
                                                          const double *__restrict__ q,
                                                          double *part, MaxLoc *mpart,
                                                          const KState *ks) {
  if (ks->stop) return;
  __shared__ double lds[MF6_BLOCK / 64];
  __shared__ MaxLoc ml[2][MF6_BLOCK / 64];
  double den = fs->pqs[head];
  den = den + copysign(2.220446049250313080847263336181640625e-16, den);
  const double alpha = fs->pr / den;
  double ss = 0.0;
  MaxLoc mx{0.0, -1}, mr{0.0, -1};
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < L.NP; e += gridDim.x * blockDim.x) {
    int c, k, i, j, l;
    if (!pos_to_cell(L, e, c, k, i, j, l)) continue;
    const int nat = natural(L, k, i, j);
    const double dx = alpha * p[e];
    x[e] = x[e] + dx;
    const double rv = r[e] - alpha * q[e];
    r[e] = rv;
    ss = ss + rv * rv;
    mx = maxloc_pick(mx, MaxLoc{dx, nat});
    mr = maxloc_pick(mr, MaxLoc{rv, nat});
  }
  for (int o = 32; o > 0; o >>= 1) {
    ss = ss + __shfl_xor(ss, o);
    MaxLoc b;
    b.v = __shfl_xor(mx.v, o); b.loc = __shfl_xor(mx.loc, o); mx = maxloc_pick(mx, b);
    b.v = __shfl_xor(mr.v, o); b.loc = __shfl_xor(mr.loc, o); mr = maxloc_pick(mr, b);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { lds[w] = ss; ml[0][w] = mx; ml[1][w] = mr; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    MaxLoc a{0.0, -1}, b{0.0, -1};
    for (int v = 0; v < (int)(blockDim.x >> 6); v++) {
      t = t + lds[v];
      a = maxloc_pick(a, ml[0][v]);
      b = maxloc_pick(b, ml[1][v]);
    }
    part[blockIdx.x] = t;
    mpart[blockIdx.x] = a;
    mpart[gridDim.x + blockIdx.x] = b;
  }
}

// end of an iteration: norms, the IMS test (-> cand), the stall check of the
// learned step
__global__ void __launch_bounds__(MF6_BLOCK) k_fcg_upd_final(const double *part,
                                                             const MaxLoc *mpart, int nb,
                                                             FcgState *fs, KState *ks,
                                                             FcgTest tst) {
  if (ks->stop) return;
  __shared__ double lds[MF6_BLOCK / 64];
  __shared__ MaxLoc ml[2][MF6_BLOCK / 64];
  double t = 0.0;
  MaxLoc a{0.0, -1}, b{0.0, -1};
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    t = t + part[i];
    a = maxloc_pick(a, mpart[i]);
    b = maxloc_pick(b, mpart[nb + i]);
  }
  for (int o = 32; o > 0; o >>= 1) {
    t = t + __shfl_xor(t, o);
    MaxLoc c;
    c.v = __shfl_xor(a.v, o); c.loc = __shfl_xor(a.loc, o); a = maxloc_pick(a, c);
    c.v = __shfl_xor(b.v, o); c.loc = __shfl_xor(b.loc, o); b = maxloc_pick(b, c);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { lds[w] = t; ml[0][w] = a; ml[1][w] = b; }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double s = 0.0;
  MaxLoc ra{0.0, -1}, rb{0.0, -1};
  for (int v = 0; v < (int)(blockDim.x >> 6); v++) {
    s = s + lds[v];
    ra = maxloc_pick(ra, ml[0][v]);
    rb = maxloc_pick(rb, ml[1][v]);
  }
  const double rnorm = sqrt(s), dvmax = ra.v, rmax = rb.v;
  fs->rnorm = rnorm;
  fs->dvmax = dvmax;
  fs->rmax = rmax;
  fs->iter++;
  // IMS test on the signed maxima (ICNVGOPT 0 / 1), or with the L2 norm
  bool conv = fabs(dvmax) <= tst.dvclose && fabs(rmax) <= tst.rclose;
  if (tst.icnvgopt == 2 || tst.icnvgopt == 3 || tst.icnvgopt == 4)
    conv = fabs(dvmax) <= tst.dvclose && (tst.icnvgopt == 3 || rnorm <= tst.rclose);
  fs->cand = conv ? 1 : 0;
  // stall check of the learned step over windows of ten applications
  if (fs->learned) {
    if (fs->window == 0) fs->rn_window = rnorm;
    if (++fs->window == 10 * tst.nn_every) {
      if (!(rnorm < 0.9 * fs->rn_window)) fs->learned = 0;
      fs->window = 0;
    }
  }
}

// r = b - A x (amux order), the FP64 true residual, when the IMS test passed;
// partials of max |r|
__global__ void __launch_bounds__(MF6_BLOCK) k_fcg_resid(Layout L, const double *__restrict__ A,
                                                         const double *__restrict__ x,
                                                         const double *__restrict__ b,
                                                         double *__restrict__ r,
                                                         const FcgState *fs, const KState *ks,
                                                         double *part) {
  if (ks->stop || !fs->cand) return;
  __shared__ double lds[MF6_BLOCK / 64];
  const int NP = L.NP;
  double m = 0.0;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < NP; e += gridDim.x * blockDim.x) {
    int c, k, i, j, l;
    if (!pos_to_cell(L, e, c, k, i, j, l)) { r[e] = 0.0; continue; }
    const int ob = (1 - c) * L.H + k * L.HP;
    double t = 0.0;
    t = t + A[e] * x[e];
    if (k > 0) t = t + A[NP + e] * x[e - L.HP];
    if (i > 0) t = t + A[2 * NP + e] * x[ob + hn(l, -L.ncol)];
    if (j > 0) t = t + A[3 * NP + e] * x[ob + hn(l, -1)];
    if (j < L.ncol - 1) t = t + A[4 * NP + e] * x[ob + hn(l, 1)];
    if (i < L.nrow - 1) t = t + A[5 * NP + e] * x[ob + hn(l, L.ncol)];
    if (k < L.nlay - 1) t = t + A[6 * NP + e] * x[e + L.HP];
    const double rv = b[e] - t;
    r[e] = rv;
    m = fmax(m, fabs(rv));
  }
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int v = 0; v < (int)(blockDim.x >> 6); v++) t = fmax(t, lds[v]);
    part[blockIdx.x] = t;
  }
}

// the true residual decides: stop when max |b - A x| <= rclose (or the
// test ignores the residual)
__global__ void __launch_bounds__(MF6_BLOCK) k_fcg_resid_final(const double *part, int nb,
                                                               FcgState *fs, KState *ks,
                                                               FcgTest tst) {
  if (ks->stop || !fs->cand) return;
  __shared__ double lds[MF6_BLOCK / 64];
  double m = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) m = fmax(m, part[i]);
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x != 0) return;
  double t = 0.0;
  for (int v = 0; v < (int)(blockDim.x >> 6); v++) t = fmax(t, lds[v]);
  fs->cand = 0;
  if (t <= tst.rclose || tst.icnvgopt == 3) {
    fs->icnvg = 1;
    ks->stop = 1;
  }
}

inline int gsz(int n) {
  int g = (n + MF6_BLOCK - 1) / MF6_BLOCK;
  if (g > 2048) g = 2048;
  return g < 1 ? 1 : g;
}

}  // namespace

#define FCK(x)                        \
  do {                                \
    int rc_ = (x);                    \
    if (rc_ != MF6HIP_OK) return rc_; \
  } while (0)

// device dots of up to 3 pairs into state fields
static void dots(Mf6Handle *h, double *part, int ndot, const double *a0, const double *b0,
                 const double *a1, const double *b1, const double *a2, const double *b2,
                 double *o0, double *o1, double *o2) {
  const int nb = gsz(h->L.NP);
  hipStream_t s = h->stream;
  k_fcg_dots<<<nb, MF6_BLOCK, 0, s>>>(h->L.NP, ndot, a0, b0, a1, b1, a2, b2, part, h->d_state);
  SumOut so{};
  so.o[0] = o0;
  so.o[1] = o1;
  so.o[2] = o2;
  k_fcg_sum<<<1, MF6_BLOCK, 0, s>>>(part, nb, ndot, so, h->d_state);
}

// z = ILU0(r), plus the learned step when `learned` (enqueued only; the
// weight is decided on the device)
static int precondition(Mf6Handle *h, FcgState *fs, double *part, const double *r, double *z,
                        bool learned) {
  const Layout &L = h->L;
  hipStream_t s = h->stream;
  FCK(mf6hip_ilu_apply64(h, r, z, h->d_state));
  if (!learned) return MF6HIP_OK;
  double *t = h->v, *yv = h->dhat, *ay = h->phat, *zl = h->qhat;
  FCK(mf6hip_spmv64(h, z, t));
  k_fcg_sub<<<gsz(L.NP), MF6_BLOCK, 0, s>>>(L.NP, r, t, h->d_state);
  dots(h, part, 2, t, t, z, r, nullptr, nullptr, &fs->tt, &fs->zr_ilu, nullptr);
  k_fcg_nn_input<<<gsz(L.N * 16), MF6_BLOCK, 0, s>>>(L, t, fs, h->unet_feat,
                                                     mf6hip_unet_input(h->unet));
  FCK(mf6hip_unet_apply_device(h->unet, s));
  k_fcg_nn_out<<<gsz(L.NP), MF6_BLOCK, 0, s>>>(L, mf6hip_unet_output(h->unet), fs, yv);
  FCK(mf6hip_spmv64(h, yv, ay));
  dots(h, part, 2, yv, t, yv, ay, nullptr, nullptr, &fs->yt, &fs->yay, nullptr);
  k_fcg_decide<<<1, 64, 0, s>>>(fs, h->d_state);
  k_fcg_axpy_w<<<gsz(L.NP), MF6_BLOCK, 0, s>>>(L.NP, fs, z, yv, zl);
  // post-smoothing: zl += ILU0(r - A zl)
  FCK(mf6hip_spmv64(h, zl, t));
  k_fcg_sub<<<gsz(L.NP), MF6_BLOCK, 0, s>>>(L.NP, r, t, h->d_state);
  FCK(mf6hip_ilu_apply64(h, t, yv, h->d_state));
  k_fcg_add<<<gsz(L.NP), MF6_BLOCK, 0, s>>>(L.NP, fs, yv, zl);
  dots(h, part, 1, zl, r, nullptr, nullptr, nullptr, nullptr, &fs->zr_l, nullptr, nullptr);
  k_fcg_accept<<<1, 64, 0, s>>>(fs, h->d_state);
  k_fcg_select<<<gsz(L.NP), MF6_BLOCK, 0, s>>>(L.NP, fs, zl, z);
  HIPCHK(hipGetLastError());
  return MF6HIP_OK;
}

// Flexible CG on the device system; h->d = b - A x, the ILU0 factor and
// the initial state (l2norm0, settings) are in place.  The new direction is
// A-orthogonalised against the last fcg_m directions (FCG(m)).
int mf6hip_run_fcg(Mf6Handle *h) {
  const Layout &L = h->L;
  hipStream_t s = h->stream;
  KState *hs = h->h_state;
  const int m = h->fcg_m;
  if (!h->fcg_store) {
    const size_t bytes = sizeof(double) * 2 * (size_t)m * L.NP + sizeof(FcgState) +
                         sizeof(double) * FCG_NDOT * 2048;
    if (hipMalloc((void **)&h->fcg_store, bytes) != hipSuccess) {
      mf6hip_set_error("hipMalloc failed (flexible CG directions)");
      return MF6HIP_EHIP;
    }
  }
  double *r = h->d, *z = h->z;
  double *P = h->fcg_store, *Q = h->fcg_store + (size_t)m * L.NP;
  FcgState *fs = (FcgState *)(h->fcg_store + 2 * (size_t)m * L.NP);
  double *part = (double *)(fs + 1);
  const int nb = gsz(L.NP);
  FcgTest tst{h->s.dvclose, h->s.rclose, h->s.icnvgopt, h->nn_every};
  bool learned = h->unet != nullptr;
  k_fcg_init<<<1, 64, 0, s>>>(fs, h->d_state, learned ? 1 : 0);
  HIPCHK(hipGetLastError());
  int nstored = 0, head = 0, it = 0;
  const int itmax = h->s.iter1;
  FcgState hf{};
  while (it < itmax) {
    const int ng = itmax - it < FCG_GROUP ? itmax - it : FCG_GROUP;
    for (int g = 0; g < ng; g++) {
      it++;
      const bool use_net = learned && ((it - 1) % h->nn_every == 0);
      FCK(precondition(h, fs, part, r, z, use_net));
      // c_j = z.q_j over the stored slots (newest first), FCG_NDOT a pass
      FcgSlots sl{};
      sl.n = nstored;
      for (int j = 0; j < nstored; j++) sl.slot[j] = (head - 1 - j + 2 * m) % m;
      for (int j0 = 0; j0 < nstored; j0 += FCG_NDOT) {
        const int nd = nstored - j0 < FCG_NDOT ? nstored - j0 : FCG_NDOT;
        DotArgs da{};
        da.n = nd;
        SumOut so{};
        for (int k = 0; k < nd; k++) {
          da.b[k] = Q + (size_t)sl.slot[j0 + k] * L.NP;
          so.o[k] = &fs->c[sl.slot[j0 + k]];
        }
        k_fcg_mdots<<<nb, MF6_BLOCK, 0, s>>>(L.NP, z, da, part, h->d_state);
        k_fcg_sum<<<1, MF6_BLOCK, 0, s>>>(part, nb, nd, so, h->d_state);
      }
      // the new direction goes straight into the ring slot `head`
      double *p = P + (size_t)head * L.NP, *q = Q + (size_t)head * L.NP;
      k_fcg_dir<<<gsz(L.NP), MF6_BLOCK, 0, s>>>(L.NP, z, P, (size_t)L.NP, sl, fs, p, h->d_state);
      FCK(mf6hip_spmv64(h, p, q));
      dots(h, part, 2, p, q, p, r, nullptr, nullptr, &fs->pqs[head], &fs->pr, nullptr);
      k_fcg_update<<<nb, MF6_BLOCK, 0, s>>>(L, fs, head, h->x, p, r, q, h->d_part, h->d_mpart,
                                            h->d_state);
      k_fcg_upd_final<<<1, MF6_BLOCK, 0, s>>>(h->d_part, h->d_mpart, nb, fs, h->d_state, tst);
      k_fcg_resid<<<nb, MF6_BLOCK, 0, s>>>(L, h->A, h->x, h->b, r, fs, h->d_state, part);
      k_fcg_resid_final<<<1, MF6_BLOCK, 0, s>>>(part, nb, fs, h->d_state, tst);
      HIPCHK(hipGetLastError());
      head = (head + 1) % m;
      if (nstored < m) nstored++;
    }
    HIPCHK(hipMemcpyAsync(&hf, fs, sizeof(FcgState), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (hf.icnvg) break;
    learned = learned && hf.learned;
  }
  hs->icnvg = hf.icnvg;
  hs->iter = hf.iter;
  hs->deltax = hf.dvmax;
  hs->rmax = hf.rmax;
  hs->l2norm = hf.rnorm;
  hs->stop = 1;
  h->learned_iters = hf.learned_iters;
  HIPCHK(hipMemcpyAsync(h->d_state, hs, sizeof(KState), hipMemcpyHostToDevice, s));
  return MF6HIP_OK;
}

extern "C" int mf6hip_attach_unet(void *handle, void *net, const float *features, int fcg_m,
                                  int nn_every) {
  Mf6Handle *h = (Mf6Handle *)handle;
  if (!h) { mf6hip_set_error("null handle"); return MF6HIP_EINVAL; }
  if (fcg_m < 1 || fcg_m > 64) { mf6hip_set_error("fcg_m must be 1..64"); return MF6HIP_EINVAL; }
  if (nn_every < 1) { mf6hip_set_error("nn_every must be >= 1"); return MF6HIP_EINVAL; }
  h->nn_every = nn_every;
  if (fcg_m != h->fcg_m && h->fcg_store) {
    hipFree(h->fcg_store);
    h->fcg_store = nullptr;
  }
  h->fcg_m = fcg_m;
  if (!net) { h->unet = nullptr; return MF6HIP_OK; }
  if (!features) { mf6hip_set_error("features required"); return MF6HIP_EINVAL; }
  int nl, nr, nc, nin;
  FCK(mf6hip_unet_dims(net, &nl, &nr, &nc, &nin));
  if (nl != h->L.nlay || nr != h->L.nrow || nc != h->L.ncol || nin != 3) {
    mf6hip_set_error("network grid or inputs do not match the solver");
    return MF6HIP_EINVAL;
  }
  if (!h->unet_feat) {
    if (hipMalloc((void **)&h->unet_feat, sizeof(float) * 2 * (size_t)h->L.N) != hipSuccess) {
      mf6hip_set_error("hipMalloc failed");
      return MF6HIP_EHIP;
    }
  }
  HIPCHK(hipMemcpy(h->unet_feat, features, sizeof(float) * 2 * (size_t)h->L.N, hipMemcpyHostToDevice));
  h->unet = net;
  return MF6HIP_OK;
}

extern "C" int mf6hip_last_learned_iterations(void *handle, int *n) {
  Mf6Handle *h = (Mf6Handle *)handle;
  if (!h || !n) { mf6hip_set_error("null argument"); return MF6HIP_EINVAL; }
  *n = h->learned_iters;
  return MF6HIP_OK;
}
