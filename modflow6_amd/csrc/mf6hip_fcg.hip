// mf6hip_fcg.hip -- flexible preconditioned CG with the learned correction.
//
// Preconditioner (variable, hence flexible CG with one stored direction):
//   z  = ILU0(r)                      the device ILU0 of the IMS solve
//   t  = r - A z                      what the ILU0 step leaves
//   y  = s * net([t / s, features])   s = rms(t): the network sees unit-scale
//                                     residuals; its output is rescaled
//   z += w y,  w = y.t / y.A y        energy-optimal step along y
// FCG(m) (Notay 2000): p = z - sum_j (z.q_j / p_j.q_j) p_j over the last m
// directions, alpha = p.r / p.q, x += alpha p, r -= alpha q.  The IMS convergence test
// (ImsLinearBase.f90:1101-1146) is applied to every iteration with the signed
// largest head change alpha p and residual; when it passes, the FP64 true
// residual b - A x is recomputed and must pass it too (else the iteration
// continues from the true residual).  The network runs on every nn_every-th
// iteration (ILU0 alone on the others).  Fallback: the ILU0-only step
// replaces the learned one for the rest of the solve when z.r changes sign
// against the ILU0 step (not a descent direction) or when ten applications
// of the network (10 nn_every iterations) reduce ||r||_2 by less than 10 %.
#include <cmath>
#include <vector>

#include "mf6hip_handle.h"

namespace {

__global__ void k_fcg_sub(int n, const double *__restrict__ r, double *__restrict__ t) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x)
    t[e] = r[e] - t[e];
}

// network input (BF16, 16 channels, natural order): ch0 = t / s, ch1..2 =
// features, ch3..15 = 0
__global__ void k_fcg_nn_input(Layout L, const double *__restrict__ t, double inv_s,
                               const float *__restrict__ feat, unsigned short *__restrict__ in0) {
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
__global__ void k_fcg_nn_out(Layout L, const float *__restrict__ y, double s,
                             double *__restrict__ yv) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < L.NP; e += gridDim.x * blockDim.x) {
    int c, k, i, j, l;
    yv[e] = pos_to_cell(L, e, c, k, i, j, l) ? s * (double)y[natural(L, k, i, j)] : 0.0;
  }
}

// z += w y
__global__ void k_fcg_axpy(int n, double w, const double *__restrict__ y, double *__restrict__ z) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x)
    z[e] = z[e] + w * y[e];
}

// up to 3 dot products in one pass: part[k * nb + block]
__global__ void __launch_bounds__(MF6_BLOCK) k_fcg_dots(int n, int ndot, const double *a0,
                                                        const double *b0, const double *a1,
                                                        const double *b1, const double *a2,
                                                        const double *b2, double *part) {
  __shared__ double lds[3][MF6_BLOCK / 64];
  double s[3] = {0.0, 0.0, 0.0};
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    s[0] = s[0] + a0[e] * b0[e];
    if (ndot > 1) s[1] = s[1] + a1[e] * b1[e];
    if (ndot > 2) s[2] = s[2] + a2[e] * b2[e];
  }
  for (int k = 0; k < 3; k++)
    for (int o = 32; o > 0; o >>= 1) s[k] = s[k] + __shfl_xor(s[k], o);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int k = 0; k < 3; k++) lds[k][w] = s[k];
  __syncthreads();
  if (threadIdx.x == 0)
    for (int k = 0; k < ndot; k++) {
      double t = 0.0;
      for (int v = 0; v < (int)(blockDim.x >> 6); v++) t = t + lds[k][v];
      part[k * gridDim.x + blockIdx.x] = t;
    }
}

__global__ void k_fcg_sum(const double *part, int nb, int ndot, double *out) {
  __shared__ double lds[MF6_BLOCK];
  for (int k = 0; k < ndot; k++) {
    double s = 0.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) s = s + part[k * nb + i];
    lds[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int i = 0; i < (int)blockDim.x; i++) t = t + lds[i];
      out[k] = t;
    }
    __syncthreads();
  }
}

// p = z + beta p_prev
__global__ void k_fcg_dir(int n, const double *__restrict__ z, double beta,
                          const double *__restrict__ pp, double *__restrict__ p) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x)
    p[e] = beta == 0.0 ? z[e] : z[e] + beta * pp[e];
}

// x += alpha p ; r -= alpha q ; partials: sum r^2, signed max |alpha p| and
// max |r| with natural locations (first occurrence on ties)
__global__ void __launch_bounds__(MF6_BLOCK) k_fcg_update(Layout L, double alpha,
                                                          double *__restrict__ x,
                                                          const double *__restrict__ p,
                                                          double *__restrict__ r,
                                                          const double *__restrict__ q,
                                                          double *part, MaxLoc *mpart) {
  __shared__ double lds[MF6_BLOCK / 64];
  __shared__ MaxLoc ml[2][MF6_BLOCK / 64];
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

__global__ void __launch_bounds__(MF6_BLOCK) k_fcg_upd_final(const double *part,
                                                             const MaxLoc *mpart, int nb,
                                                             double *out) {
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
  out[0] = s;
  out[1] = ra.v;
  out[2] = rb.v;
}

// r = b - A x (amux order), the FP64 true residual
__global__ void k_fcg_resid(Layout L, const double *__restrict__ A, const double *__restrict__ x,
                            const double *__restrict__ b, double *__restrict__ r) {
  const int NP = L.NP;
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
    r[e] = b[e] - t;
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

// dots of up to 3 pairs (host result)
static int dots(Mf6Handle *h, int ndot, const double *a0, const double *b0, const double *a1,
                const double *b1, const double *a2, const double *b2, double *out) {
  const int nb = gsz(h->L.NP);
  hipStream_t s = h->stream;
  double *dout = h->d_part + 7 * MF6_MAXBLK;
  k_fcg_dots<<<nb, MF6_BLOCK, 0, s>>>(h->L.NP, ndot, a0, b0, a1, b1, a2, b2, h->d_part);
  k_fcg_sum<<<1, MF6_BLOCK, 0, s>>>(h->d_part, nb, ndot, dout);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, dout, sizeof(double) * ndot, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return MF6HIP_OK;
}

// z = ILU0(r) + w y, y = s net(t / s), t = r - A ILU0(r), with the
// energy-optimal weight w = (y.t) / (y.A y) (the step along y that
// minimises the A-norm of the error left by ILU0, so a poor prediction costs
// iterations but cannot push the iterate away); *used = 1 when the
// network ran, *zr_ilu = ILU0(r).r (the sign a descent direction must keep)
static int precondition(Mf6Handle *h, const double *r, double *z, bool learned, int *used,
                        double *zr_ilu) {
  const Layout &L = h->L;
  hipStream_t s = h->stream;
  *used = 0;
  FCK(mf6hip_ilu_apply64(h, r, z));
  if (!learned) return MF6HIP_OK;
  double *t = h->v, *yv = h->dhat, *ay = h->phat;
  FCK(mf6hip_spmv64(h, z, t));
  k_fcg_sub<<<gsz(L.NP), MF6_BLOCK, 0, s>>>(L.NP, r, t);
  double d2[2];
  FCK(dots(h, 2, t, t, z, r, nullptr, nullptr, d2));
  *zr_ilu = d2[1];
  if (!(d2[0] > 0.0)) return MF6HIP_OK;
  const double sc = sqrt(d2[0] / L.N);
  k_fcg_nn_input<<<gsz(L.N * 16), MF6_BLOCK, 0, s>>>(L, t, 1.0 / sc, h->unet_feat,
                                                     mf6hip_unet_input(h->unet));
  FCK(mf6hip_unet_apply_device(h->unet, s));
  k_fcg_nn_out<<<gsz(L.NP), MF6_BLOCK, 0, s>>>(L, mf6hip_unet_output(h->unet), sc, yv);
  FCK(mf6hip_spmv64(h, yv, ay));
  double w2[2];
  FCK(dots(h, 2, yv, t, yv, ay, nullptr, nullptr, w2));
  if (w2[1] != 0.0 && std::isfinite(w2[0]) && std::isfinite(w2[1])) {
    k_fcg_axpy<<<gsz(L.NP), MF6_BLOCK, 0, s>>>(L.NP, w2[0] / w2[1], yv, z);
    *used = 1;
  }
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
    if (hipMalloc((void **)&h->fcg_store, sizeof(double) * 2 * (size_t)m * L.NP) != hipSuccess) {
      mf6hip_set_error("hipMalloc failed (flexible CG directions)");
      return MF6HIP_EHIP;
    }
  }
  double *r = h->d, *z = h->z, *p = h->p, *q = h->q;
  double *P = h->fcg_store, *Q = h->fcg_store + (size_t)m * L.NP;
  std::vector<double> pqs(m, 0.0);
  int nstored = 0, head = 0;
  const int nb = gsz(L.NP);
  double *dout = h->d_part + 7 * MF6_MAXBLK;
  bool learned = h->unet != nullptr;
  h->learned_iters = 0;
  double rn_window = -1.0;
  int window = 0;
  int it = 0, icnvg = 0;
  const int itmax = h->s.iter1;
  while (it < itmax) {
    it++;
    int used = 0;
    double zr_ilu = 0.0;
    const bool use_net = learned && ((it - 1) % h->nn_every == 0);
    FCK(precondition(h, r, z, use_net, &used, &zr_ilu));
    double zr;
    FCK(dots(h, 1, z, r, nullptr, nullptr, nullptr, nullptr, &zr));
    if (used && !(zr * zr_ilu > 0.0)) {
      // z.r changed sign against the ILU0 step (M is definite with the
      // sign of A): fall back to the ILU0 step for good
      learned = false;
      FCK(precondition(h, r, z, false, &used, &zr_ilu));
    }
    if (used) h->learned_iters++;
    // p = z - sum_j (z.q_j / p_j.q_j) p_j over the stored directions
    std::vector<double> c(nstored, 0.0);
    for (int j0 = 0; j0 < nstored; j0 += 3) {
      const int nd = nstored - j0 < 3 ? nstored - j0 : 3;
      const double *qj[3] = {nullptr, nullptr, nullptr};
      for (int k = 0; k < nd; k++) qj[k] = Q + (size_t)((head - 1 - j0 - k + 2 * m) % m) * L.NP;
      FCK(dots(h, nd, z, qj[0], z, qj[1], z, qj[2], c.data() + j0));
    }
    k_fcg_dir<<<gsz(L.NP), MF6_BLOCK, 0, s>>>(L.NP, z, 0.0, z, p);
    for (int j = 0; j < nstored; j++) {
      const int slot = (head - 1 - j + 2 * m) % m;
      if (pqs[slot] != 0.0)
        k_fcg_axpy<<<gsz(L.NP), MF6_BLOCK, 0, s>>>(L.NP, -c[j] / pqs[slot],
                                                   P + (size_t)slot * L.NP, p);
    }
    FCK(mf6hip_spmv64(h, p, q));
    double pq[2];
    FCK(dots(h, 2, p, q, p, r, nullptr, nullptr, pq));
    double den = pq[0];
    den = den + copysign(2.220446049250313080847263336181640625e-16, den);
    const double alpha = pq[1] / den;
    k_fcg_update<<<nb, MF6_BLOCK, 0, s>>>(L, alpha, h->x, p, r, q, h->d_part, h->d_mpart);
    k_fcg_upd_final<<<1, MF6_BLOCK, 0, s>>>(h->d_part, h->d_mpart, nb, dout);
    HIPCHK(hipGetLastError());
    // store the direction in the ring (the oldest one is replaced)
    HIPCHK(hipMemcpyAsync(P + (size_t)head * L.NP, p, sizeof(double) * L.NP,
                          hipMemcpyDeviceToDevice, s));
    HIPCHK(hipMemcpyAsync(Q + (size_t)head * L.NP, q, sizeof(double) * L.NP,
                          hipMemcpyDeviceToDevice, s));
    pqs[head] = pq[0];
    head = (head + 1) % m;
    if (nstored < m) nstored++;
    double u3[3];
    HIPCHK(hipMemcpyAsync(u3, dout, sizeof(u3), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    const double rnorm = sqrt(u3[0]), dvmax = u3[1], rmax = u3[2];
    hs->deltax = dvmax;
    hs->rmax = rmax;
    hs->l2norm = rnorm;
    // IMS test on the signed maxima (ICNVGOPT 0 / 1), or with the L2 norm
    bool conv = fabs(dvmax) <= h->s.dvclose && fabs(rmax) <= h->s.rclose;
    if (h->s.icnvgopt == 2 || h->s.icnvgopt == 3 || h->s.icnvgopt == 4) {
      conv = fabs(dvmax) <= h->s.dvclose && (h->s.icnvgopt == 3 || rnorm <= h->s.rclose);
    }
    if (conv) {
      // the FP64 true residual is the judge
      k_fcg_resid<<<gsz(L.NP), MF6_BLOCK, 0, s>>>(L, h->A, h->x, h->b, r);
      HIPCHK(hipGetLastError());
      std::vector<double> tmp(L.NP);
      HIPCHK(hipMemcpyAsync(tmp.data(), r, sizeof(double) * L.NP, hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      double tmax = 0.0;
      for (double v : tmp) tmax = fmax(tmax, fabs(v));
      if (tmax <= h->s.rclose || h->s.icnvgopt == 3) { icnvg = 1; break; }
    }
    // stall check of the learned step over windows of ten applications
    if (learned) {
      if (window == 0) rn_window = rnorm;
      if (++window == 10 * h->nn_every) {
        if (!(rnorm < 0.9 * rn_window)) learned = false;
        window = 0;
      }
    }
  }
  hs->icnvg = icnvg;
  hs->iter = it;
  hs->stop = 1;
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
