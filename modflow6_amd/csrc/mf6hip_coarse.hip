// mf6hip_coarse.hip -- two-level preconditioner for the fused CG: the IMS
// ILU0 / MILU0 preconditioner plus an additive coarse-space correction on
// column aggregates (no counterpart in the reference, whose preconditioners
// are the one-level ILU0 / MILU0 / ILUT of ImsLinearBase.f90:761-1092).
//
//   M^-1 r = ILU0^-1 r + P Ac^-1 P^T r,   Ac = P^T A P
//
// P is piecewise constant: aggregate I holds every layer of an ha x ha block
// of grid columns, and P maps I to each of its cells that is coupled to a
// neighbour (a row with a nonzero off-diagonal).  Decoupled rows -- CHD and
// inactive cells after the sln_ls row fixes, identity rows -- stay with the
// ILU0 part alone, so their heads are left exactly as the one-level solve
// leaves them.  With A symmetric both parts are symmetric and the sum is a
// fixed linear operator: plain PCG applies.
//
// Setup (once per linear solve, the matrix changes every outer iteration):
//   k_tl_cmask    per column: bit k set when cell (k, column) is coupled
//   k_tl_build    one block per aggregate: its row of -Ac (five nonzeros on
//                 the aggregate grid) into a dense symmetric matrix
//   rocSOLVER     dpotrf + dpotri: (-Ac)^-1, FP64
//   k_tl_finish   Ainv = -(-Ac)^-1 as a full row-major matrix, FP64 and FP32
// Per CG iteration (after the preconditioner kernels, before the round):
//   the black / red backward preconditioner kernels write each column's sum
//   of the residual over its coupled cells (colsum)
//   k_tl_gather   one block per aggregate: rc = P^T d from colsum
//   k_tl_gemv     one wavefront per coarse row: yc = Ainv rc, partial rc.yc
// and the direction update adds the prolongation: p = z + P yc + beta p.
// rho = d.(z + P yc) = d.z + rc.yc, so the round kernel adds the partials of
// rc.yc to those of d.z.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <cmath>
#include <string>
#include <vector>

#include "mf6hip_handle.h"

#define CK(x)                         \
  do {                                \
    int rc_ = (x);                    \
    if (rc_ != MF6HIP_OK) return rc_; \
  } while (0)

namespace {

__device__ inline double wsum64(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// sum over the block (every thread gets it)
__device__ inline double block_sum(double v, double *lds) {
  v = wsum64(v);
  const int w = threadIdx.x >> 6, ln = threadIdx.x & 63;
  __syncthreads();
  if (ln == 0) lds[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < (int)(blockDim.x >> 6); i++) s = s + lds[i];
  return s;
}

// bit k of cmask[c * HP + h2]: cell (k, column) has a nonzero off-diagonal
__global__ void k_tl_cmask(Layout L, const double *__restrict__ A, unsigned *__restrict__ cmask) {
  const int NP = L.NP;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < 2 * L.HP;
       t += gridDim.x * blockDim.x) {
    const int c = t >= L.HP;
    const int h2 = t - c * L.HP;
    unsigned m = 0u;
    if (h2 < ncols_of(L, c)) {
      for (int k = 0; k < L.nlay; k++) {
        const int pos = c * L.H + k * L.HP + h2;
        bool cp = false;
        for (int s = S_UP; s <= S_DOWN; s++) cp = cp || A[(size_t)s * NP + pos] != 0.0;
        if (cp) m |= 1u << k;
      }
    }
    cmask[t] = m;
  }
}

// Row I of -Ac into the dense symmetric matrix G (leading dimension ld): the
// diagonal (cell diagonals plus the couplings inside the aggregate) and the
// couplings to the four neighbouring aggregates.  An aggregate without a
// coupled cell gets a unit row (its restricted residual is always zero).
__global__ void __launch_bounds__(MF6_BLOCK) k_tl_build(Layout L, const double *__restrict__ A,
                                                       const unsigned *__restrict__ cmask,
                                                       TwoLevel T, double *__restrict__ G) {
  __shared__ double lds[MF6_BLOCK / 64];
  const int I = blockIdx.x;
  if (I >= T.nc) return;
  const int ai = I / T.naj, aj = I - ai * T.naj;
  const int i0 = ai * T.ha, j0 = aj * T.ha;
  const int NP = L.NP;
  double sd = 0.0, sw = 0.0, se = 0.0, sn = 0.0, ss = 0.0, cnt = 0.0;
  for (int t = threadIdx.x; t < T.ha * T.ha; t += blockDim.x) {
    const int i = i0 + t / T.ha, j = j0 + t % T.ha;
    if (i >= L.nrow || j >= L.ncol) continue;
    const int l = i * L.ncol + j, c = (i + j) & 1, h2 = l >> 1;
    const unsigned m = cmask[c * L.HP + h2];
    for (int k = 0; k < L.nlay; k++) {
      if (!((m >> k) & 1u)) continue;
      const int pos = c * L.H + k * L.HP + h2;
      cnt = cnt + 1.0;
      // up / down stay in the column, hence in the aggregate
      sd = sd + A[pos] + A[(size_t)S_UP * NP + pos] + A[(size_t)S_DOWN * NP + pos];
      const double ab = A[(size_t)S_BACK * NP + pos], al = A[(size_t)S_LEFT * NP + pos];
      const double ar = A[(size_t)S_RIGHT * NP + pos], af = A[(size_t)S_FRONT * NP + pos];
      if (i > i0) sd = sd + ab; else sn = sn + ab;
      if (i < i0 + T.ha - 1) sd = sd + af; else ss = ss + af;
      if (j > j0) sd = sd + al; else sw = sw + al;
      if (j < j0 + T.ha - 1) sd = sd + ar; else se = se + ar;
    }
  }
  sd = block_sum(sd, lds);
  sw = block_sum(sw, lds);
  se = block_sum(se, lds);
  sn = block_sum(sn, lds);
  ss = block_sum(ss, lds);
  cnt = block_sum(cnt, lds);
  if (threadIdx.x != 0) return;
  const size_t ld = (size_t)T.ld;
  if (cnt == 0.0) {
    G[(size_t)I * ld + I] = 1.0;
    return;
  }
  G[(size_t)I * ld + I] = -sd;
  if (aj > 0) G[(size_t)I * ld + (I - 1)] = -sw;
  if (aj < T.naj - 1) G[(size_t)I * ld + (I + 1)] = -se;
  if (ai > 0) G[(size_t)I * ld + (I - T.naj)] = -sn;
  if (ai < T.nai - 1) G[(size_t)I * ld + (I + T.naj)] = -ss;
}

// Ainv = -(G^-1) from the lower triangle potri left in G (column-major:
// element (r, c), r >= c, at G[r + c ld]); written in full, FP64 and FP32.
__global__ void k_tl_finish(TwoLevel T, const double *__restrict__ G, double *__restrict__ A64,
                            float *__restrict__ A32) {
  const size_t n = (size_t)T.nc * T.ld;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (size_t)gridDim.x * blockDim.x) {
    const int r = (int)(e / T.ld), c = (int)(e - (size_t)r * T.ld);
    double v = 0.0;
    if (c < T.nc) v = -(r >= c ? G[r + (size_t)c * T.ld] : G[c + (size_t)r * T.ld]);
    if (A64) A64[e] = v;
    if (A32) A32[e] = (float)v;
  }
}

// rc[I] = sum of d over the coupled cells of aggregate I
template <typename TV>
__global__ void __launch_bounds__(MF6_BLOCK) k_tl_restrict(Layout L, const TV *__restrict__ d,
                                                          const unsigned *__restrict__ cmask,
                                                          TwoLevel T, double *__restrict__ rc,
                                                          const KState *st) {
  __shared__ double lds[MF6_BLOCK / 64];
  if (st && st->stop) return;
  for (int I = blockIdx.x; I < T.nc; I += gridDim.x) {
    const int ai = I / T.naj, aj = I - ai * T.naj;
    double s = 0.0;
    for (int t = threadIdx.x; t < T.ha * T.ha; t += blockDim.x) {
      const int i = ai * T.ha + t / T.ha, j = aj * T.ha + t % T.ha;
      if (i >= L.nrow || j >= L.ncol) continue;
      const int l = i * L.ncol + j, c = (i + j) & 1, h2 = l >> 1;
      const unsigned m = cmask[c * L.HP + h2];
      const int base = c * L.H + h2;
      for (int k = 0; k < L.nlay; k++)
        if ((m >> k) & 1u) s = s + (double)d[base + k * L.HP];
    }
    s = block_sum(s, lds);
    if (threadIdx.x == 0) rc[I] = s;
  }
}

// rc[I] = sum of colsum over the columns of aggregate I (fixed order)
__global__ void __launch_bounds__(MF6_BLOCK) k_tl_gather(Layout L, const double *__restrict__ colsum,
                                                        TwoLevel T, double *__restrict__ rc,
                                                        const KState *st) {
  __shared__ double lds[MF6_BLOCK / 64];
  if (st && st->stop) return;
  for (int I = blockIdx.x; I < T.nc; I += gridDim.x) {
    const int ai = I / T.naj, aj = I - ai * T.naj;
    double s = 0.0;
    for (int t = threadIdx.x; t < T.ha * T.ha; t += blockDim.x) {
      const int i = ai * T.ha + t / T.ha, j = aj * T.ha + t % T.ha;
      if (i < L.nrow && j < L.ncol) s = s + colsum[i * L.ncol + j];
    }
    s = block_sum(s, lds);
    if (threadIdx.x == 0) rc[I] = s;
  }
}

template <typename TA> struct Row4;
template <> struct Row4<float> {
  float4 v;
  __device__ static Row4 load(const float *p) { return Row4{*reinterpret_cast<const float4 *>(p)}; }
  __device__ double dot(const double *x) const {
    return (double)v.x * x[0] + (double)v.y * x[1] + (double)v.z * x[2] + (double)v.w * x[3];
  }
};
template <> struct Row4<double> {
  double2 a, b;
  __device__ static Row4 load(const double *p) {
    return Row4{reinterpret_cast<const double2 *>(p)[0], reinterpret_cast<const double2 *>(p)[1]};
  }
  __device__ double dot(const double *x) const {
    return a.x * x[0] + a.y * x[1] + b.x * x[2] + b.y * x[3];
  }
};

// yc = Ainv rc, one wavefront per row with 16-byte (FP32) / 32-byte (FP64)
// row loads; rc is staged in LDS (ld doubles, zero beyond nc; the row
// entries beyond nc are zero as well).  part[b] = this block's sum of
// rc[I] yc[I]; slots gridDim.x .. nslot-1 are cleared.
template <typename TA, typename TV>
__global__ void __launch_bounds__(MF6_BLOCK) k_tl_gemv(TwoLevel T, const TA *__restrict__ Ainv,
                                                      const double *__restrict__ rc,
                                                      TV *__restrict__ yc, double *part,
                                                      int nslot, const KState *st) {
  extern __shared__ double xs[];
  __shared__ double lds[MF6_BLOCK / 64];
  if (st && st->stop) return;
  for (int c = threadIdx.x; c < T.ld; c += blockDim.x) xs[c] = c < T.nc ? rc[c] : 0.0;
  __syncthreads();
  const int w = threadIdx.x >> 6, ln = threadIdx.x & 63;
  const int wpb = blockDim.x >> 6;
  double acc = 0.0;
  for (int I = blockIdx.x * wpb + w; I < T.nc; I += gridDim.x * wpb) {
    const TA *row = Ainv + (size_t)I * T.ld;
    double s0 = 0.0, s1 = 0.0;
    int c = 4 * ln;
    for (; c + 256 < T.ld; c += 512) {
      const Row4<TA> r0 = Row4<TA>::load(row + c);
      const Row4<TA> r1 = Row4<TA>::load(row + c + 256);
      s0 = s0 + r0.dot(xs + c);
      s1 = s1 + r1.dot(xs + c + 256);
    }
    if (c < T.ld) s0 = s0 + Row4<TA>::load(row + c).dot(xs + c);
    const double sr = wsum64(s0 + s1);
    if (ln == 0) {
      yc[I] = (TV)sr;
      acc = acc + xs[I] * (double)(TV)sr;
    }
  }
  acc = block_sum(acc, lds);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = acc;
    for (int e = (int)gridDim.x + blockIdx.x; e < nslot; e += (int)gridDim.x) part[e] = 0.0;
  }
}

// z += P yc (coupled cells)
template <typename TV>
__global__ void k_tl_prolong(Layout L, const unsigned *__restrict__ cmask, TwoLevel T,
                             const TV *__restrict__ yc, TV *__restrict__ z) {
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < 2 * L.HP;
       t += gridDim.x * blockDim.x) {
    const int c = t >= L.HP;
    const int h2 = t - c * L.HP;
    if (h2 >= ncols_of(L, c)) continue;
    const unsigned m = cmask[t];
    if (!m) continue;
    const int l = col_l(L, c, h2);
    const TV y = yc[tl_agg(T, L, l)];
    for (int k = 0; k < L.nlay; k++)
      if ((m >> k) & 1u) {
        const int pos = c * L.H + k * L.HP + h2;
        z[pos] = z[pos] + y;
      }
  }
}

inline int tl_grid(size_t n) {
  size_t g = (n + MF6_BLOCK - 1) / MF6_BLOCK;
  if (g > 65535) g = 65535;
  return g < 1 ? 1 : (int)g;
}

}  // namespace

// Aggregate width and grid: the smallest ha whose aggregate grid has at most
// target coarse unknowns.
TwoLevel mf6hip_tl_shape(const Layout &L, int target) {
  if (target <= 0) target = 4096;
  TwoLevel T{};
  int ha = (int)std::ceil(std::sqrt((double)L.NRC / (double)target));
  if (ha < 1) ha = 1;
  while (((L.nrow + ha - 1) / ha) * ((L.ncol + ha - 1) / ha) > target) ha++;
  T.ha = ha;
  T.nai = (L.nrow + ha - 1) / ha;
  T.naj = (L.ncol + ha - 1) / ha;
  T.nc = T.nai * T.naj;
  T.ld = (T.nc + 63) / 64 * 64;
  return T;
}

int mf6hip_tl_alloc(Mf6Handle *h) {
  if (h->tl_cmask) return MF6HIP_OK;
  const Layout &L = h->L;
  h->tl = mf6hip_tl_shape(L, h->s.coarse_size);
  const size_t nn = (size_t)h->tl.nc * h->tl.ld;
  if (hipMalloc((void **)&h->tl_cmask, sizeof(unsigned) * 2 * (size_t)L.HP) != hipSuccess ||
      hipMalloc((void **)&h->tl_G, sizeof(double) * nn) != hipSuccess ||
      hipMalloc((void **)&h->tl_rc, sizeof(double) * h->tl.ld) != hipSuccess ||
      hipMalloc((void **)&h->tl_yc64, sizeof(double) * h->tl.ld) != hipSuccess ||
      hipMalloc((void **)&h->tl_yc32, sizeof(float) * h->tl.ld) != hipSuccess ||
      hipMalloc((void **)&h->tl_info, sizeof(int)) != hipSuccess ||
      hipMalloc((void **)&h->tl_colsum, sizeof(double) * (size_t)L.NRC) != hipSuccess) {
    mf6hip_set_error("hipMalloc failed (two-level preconditioner)");
    return MF6HIP_EHIP;
  }
  if (h->s.precision == MF6HIP_FP64) {
    if (hipMalloc((void **)&h->tl_A64, sizeof(double) * nn) != hipSuccess) {
      mf6hip_set_error("hipMalloc failed (two-level preconditioner)");
      return MF6HIP_EHIP;
    }
  } else if (hipMalloc((void **)&h->tl_A32, sizeof(float) * nn) != hipSuccess) {
    mf6hip_set_error("hipMalloc failed (two-level preconditioner)");
    return MF6HIP_EHIP;
  }
  rocblas_handle rb = nullptr;
  if (rocblas_create_handle(&rb) != rocblas_status_success) {
    mf6hip_set_error("rocblas_create_handle failed");
    return MF6HIP_EHIP;
  }
  h->tl_blas = rb;
  if (rocblas_set_stream(rb, h->stream) != rocblas_status_success) {
    mf6hip_set_error("rocblas_set_stream failed");
    return MF6HIP_EHIP;
  }
  return MF6HIP_OK;
}

void mf6hip_tl_free(Mf6Handle *h) {
  if (h->tl_blas) rocblas_destroy_handle((rocblas_handle)h->tl_blas);
  h->tl_blas = nullptr;
  void *bufs[] = {h->tl_cmask, h->tl_G, h->tl_rc, h->tl_yc64, h->tl_yc32, h->tl_info,
                  h->tl_A64, h->tl_A32, h->tl_colsum};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  h->tl_cmask = nullptr;
  h->tl_G = h->tl_A64 = h->tl_rc = h->tl_yc64 = nullptr;
  h->tl_A32 = h->tl_yc32 = nullptr;
  h->tl_info = nullptr;
  h->tl_colsum = nullptr;
}

// Coarse operator of the current FP64 planes: mask, -Ac, its inverse.
int mf6hip_tl_setup(Mf6Handle *h) {
  CK(mf6hip_tl_alloc(h));
  const Layout &L = h->L;
  const TwoLevel &T = h->tl;
  hipStream_t s = h->stream;
  k_tl_cmask<<<tl_grid(2 * (size_t)L.HP), MF6_BLOCK, 0, s>>>(L, h->A, h->tl_cmask);
  HIPCHK(hipMemsetAsync(h->tl_G, 0, sizeof(double) * (size_t)T.nc * T.ld, s));
  k_tl_build<<<T.nc, MF6_BLOCK, 0, s>>>(L, h->A, h->tl_cmask, T, h->tl_G);
  HIPCHK(hipGetLastError());
  rocblas_handle rb = (rocblas_handle)h->tl_blas;
  if (rocsolver_dpotrf(rb, rocblas_fill_lower, T.nc, h->tl_G, T.ld, h->tl_info) !=
      rocblas_status_success) {
    mf6hip_set_error("rocsolver_dpotrf failed");
    return MF6HIP_EHIP;
  }
  int info = 0;
  HIPCHK(hipMemcpyAsync(&info, h->tl_info, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (info != 0) {
    mf6hip_set_error("two-level preconditioner: the coarse matrix is not definite (leading minor " +
                     std::to_string(info) + "); the system must be symmetric definite");
    return MF6HIP_EINVAL;
  }
  if (rocsolver_dpotri(rb, rocblas_fill_lower, T.nc, h->tl_G, T.ld, h->tl_info) !=
      rocblas_status_success) {
    mf6hip_set_error("rocsolver_dpotri failed");
    return MF6HIP_EHIP;
  }
  k_tl_finish<<<tl_grid((size_t)T.nc * T.ld), MF6_BLOCK, 0, s>>>(T, h->tl_G, h->tl_A64,
                                                                h->tl_A32);
  HIPCHK(hipGetLastError());
  return MF6HIP_OK;
}

// rc = P^T d (from the preconditioner kernels' column sums when d is null,
// else from d itself) ; yc = Ainv rc ; partial rc.yc into part[0 .. nslot)
template <typename TV>
void mf6hip_tl_enqueue(Mf6Handle *h, const TV *d, TV *yc, double *part, int nslot,
                       const KState *st) {
  const TwoLevel &T = h->tl;
  hipStream_t s = h->stream;
  int gr = T.nc < 4096 ? T.nc : 4096;
  if (d)
    k_tl_restrict<TV><<<gr, MF6_BLOCK, 0, s>>>(h->L, d, h->tl_cmask, T, h->tl_rc, st);
  else
    k_tl_gather<<<gr, MF6_BLOCK, 0, s>>>(h->L, h->tl_colsum, T, h->tl_rc, st);
  int gg = (T.nc + 3) / 4;
  if (gg > nslot) gg = nslot;
  if (gg < 1) gg = 1;
  const size_t shm = sizeof(double) * (size_t)T.ld;
  if (sizeof(TV) == sizeof(double))
    k_tl_gemv<double, TV><<<gg, MF6_BLOCK, shm, s>>>(T, h->tl_A64, h->tl_rc, yc, part, nslot,
                                                     st);
  else
    k_tl_gemv<float, TV><<<gg, MF6_BLOCK, shm, s>>>(T, h->tl_A32, h->tl_rc, yc, part, nslot,
                                                    st);
}

template void mf6hip_tl_enqueue<double>(Mf6Handle *, const double *, double *, double *, int,
                                        const KState *);
template void mf6hip_tl_enqueue<float>(Mf6Handle *, const float *, float *, double *, int,
                                       const KState *);

// z += P Ainv P^T r on device vectors (FP64): the coarse part of M^-1 r for
// the parity helper mf6hip_apply_op
int mf6hip_tl_apply64(Mf6Handle *h, const double *r, double *z) {
  if (!h->tl_A64) {
    mf6hip_set_error("the FP64 coarse inverse exists for precision FP64 only");
    return MF6HIP_EINVAL;
  }
  double *part = h->d_part + 5 * (size_t)MF6_MAXBLK;
  mf6hip_tl_enqueue<double>(h, r, h->tl_yc64, part, MF6_MAXBLK, nullptr);
  k_tl_prolong<double><<<tl_grid(2 * (size_t)h->L.HP), MF6_BLOCK, 0, h->stream>>>(
      h->L, h->tl_cmask, h->tl, h->tl_yc64, z);
  HIPCHK(hipGetLastError());
  return MF6HIP_OK;
}
