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
//   k_tl_cmask     per column: bit k set when cell (k, column) is coupled
//   k_tl_build     one block per aggregate: its row of S = -Ac (diagonal,
//                  west and north couplings on the aggregate grid)
//   k_tl_band_chol S = L L^T in one workgroup: S is banded (bandwidth naj,
//                  the aggregate grid's width), so L is too; the trailing
//                  window of the right-looking factorisation lives in LDS
//   k_tl_band_inv  one wavefront per column j of S^-1: banded forward and
//                  backward substitution of e_j, the band held across lanes
//   k_tl_finish_rows  Ainv = -S^-1, FP64 or FP32
// (aggregate grids wider than 63 take rocSOLVER's dense dpotrf + dpotri)
// Per CG iteration (after the preconditioner kernels, before the round):
//   the black / red backward preconditioner kernels write each column's sum
//   of the residual over its coupled cells (colsum)
//   k_tl_gather   one block per aggregate: rc = P^T d from colsum
//   k_tl_gemv     one wavefront per coarse row: yc = Ainv rc, partial rc.yc
// and the direction update adds the prolongation: p = z + P yc + beta p.
// rho = d.(z + P yc) = d.z + rc.yc, so the round kernel adds the partials of
// rc.yc to those of d.z.
// Deflated form (preconditioner = 3): z = M^-1 (r - A P yc) + P yc with
// yc = Ac^-1 P^T r, the coarse solve before the sweeps.  P^T r of the
// iteration's new residual is P^T r_old - alpha P^T q: the sweeps' column
// sums of r_old and the SpMV's column sums of q (k_tl_gather with colsq).
// k_tl_defl_setup stores what A P yc needs per cell (the row sum over the
// coupled cells and the couplings across aggregate edges), k_tl_colsum the
// column sums of a replaced residual.
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

// Row I of S = -Ac, compact: S3[3 I] = S(I, I) (cell diagonals plus the
// couplings inside the aggregate), S3[3 I + 1] = S(I, I - 1) (west),
// S3[3 I + 2] = S(I, I - naj) (north); zero where the neighbour does not
// exist.  An aggregate without a coupled cell gets a unit row (its
// restricted residual is always zero).
__global__ void __launch_bounds__(MF6_BLOCK) k_tl_build(Layout L, const double *__restrict__ A,
                                                       const unsigned *__restrict__ cmask,
                                                       TwoLevel T, double *__restrict__ S3) {
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
  (void)se;
  (void)ss;
  S3[3 * I] = cnt == 0.0 ? 1.0 : -sd;
  S3[3 * I + 1] = (cnt != 0.0 && aj > 0) ? -sw : 0.0;
  S3[3 * I + 2] = (cnt != 0.0 && ai > 0) ? -sn : 0.0;
}

// lower-band entry S(i, l), l <= i, from the compact rows
__device__ inline double s_entry(const TwoLevel &T, const double *S3, int i, int l) {
  double v = 0.0;
  if (l == i) v = S3[3 * i];
  if (l == i - 1) v = v + S3[3 * i + 1];
  if (l == i - T.naj) v = v + S3[3 * i + 2];
  return v;
}

// dense symmetric S (row-major, leading dimension ld; zeroed beforehand),
// for rocSOLVER
__global__ void k_tl_expand(TwoLevel T, const double *__restrict__ S3, double *__restrict__ G) {
  const size_t ld = (size_t)T.ld;
  for (int I = blockIdx.x * blockDim.x + threadIdx.x; I < T.nc; I += gridDim.x * blockDim.x) {
    G[(size_t)I * ld + I] = S3[3 * I];
    if (I % T.naj > 0) {
      G[(size_t)I * ld + I - 1] = S3[3 * I + 1];
      G[(size_t)(I - 1) * ld + I] = S3[3 * I + 1];
    }
    if (I >= T.naj) {
      G[(size_t)I * ld + I - T.naj] = S3[3 * I + 2];
      G[(size_t)(I - T.naj) * ld + I] = S3[3 * I + 2];
    }
  }
}

// S = L L^T for S banded with bandwidth b = naj <= 63, in one workgroup.
// Right-looking: step j takes the pivot, scales column j (rows j+1 .. j+b)
// and subtracts its outer product from the trailing (b+1) x (b+1) window,
// which lives in LDS with rows and columns taken modulo 64; the row j+b+1
// enters the window from S as row j leaves it.  Lc[j (b+1) + t] = L(j+t, j)
// (by columns), Lr[i (b+1) + t] = L(i, i-t) (by rows), Lc[j (b+1)] holding
// the reciprocal pivot 1 / L(j, j) for the substitutions.  (A variant with
// one barrier per column, forming column j+1 during step j's update, was
// slower: the step is bound by the window's LDS traffic, not the barriers.)
// info[0] = 1 + the first column whose pivot is not positive, 0 if none.
constexpr int TL_CHOL_THREADS = 512;
__global__ void __launch_bounds__(TL_CHOL_THREADS) k_tl_band_chol(TwoLevel T,
                                                                  const double *__restrict__ S3,
                                                                  double *__restrict__ Lc,
                                                                  double *__restrict__ Lr,
                                                                  int *info) {
  __shared__ double W[64][65];
  __shared__ double col[64];
  __shared__ int bad;
  const int n = T.nc, b = T.naj, B1 = b + 1;
  for (int e = threadIdx.x; e < 64 * 64; e += blockDim.x) W[e >> 6][e & 63] = 0.0;
  if (threadIdx.x == 0) bad = 0;
  __syncthreads();
  for (int e = threadIdx.x; e < B1 * B1; e += blockDim.x) {
    const int i = e / B1, l = e - i * B1;
    if (i < n && l <= i) W[i & 63][l & 63] = s_entry(T, S3, i, l);
  }
  // this thread's (t1, t2) pairs of the update, 1 <= t2 <= t1 <= b
  constexpr int MAXP = (63 * 64 / 2 + TL_CHOL_THREADS - 1) / TL_CHOL_THREADS;
  int pt1[MAXP], pt2[MAXP];
  const int npairs = b * B1 / 2;
#pragma unroll
  for (int q = 0; q < MAXP; q++) {
    const int p = (int)threadIdx.x + q * TL_CHOL_THREADS;
    int t1 = 0, t2 = 0;
    if (p < npairs) {
      t1 = (int)((1.0 + sqrt(1.0 + 8.0 * (double)p)) * 0.5);
      while (t1 * (t1 - 1) / 2 > p) t1--;
      while ((t1 + 1) * t1 / 2 <= p) t1++;
      t2 = p - t1 * (t1 - 1) / 2 + 1;
    }
    pt1[q] = t1;
    pt2[q] = t2;
  }
  __syncthreads();
  for (int j = 0; j < n; j++) {
    if (threadIdx.x == 0) {
      double w = W[j & 63][j & 63];
      if (!(w > 0.0)) {
        if (!bad) bad = j + 1;
        w = 1.0;
      }
      const double d = sqrt(w);
      col[0] = d;
      Lc[(size_t)j * B1] = 1.0 / d;
    }
    __syncthreads();
    const int t = (int)threadIdx.x + 1;
    if (t <= b) {
      const double v = j + t < n ? W[(j + t) & 63][j & 63] / col[0] : 0.0;
      col[t] = v;
      if (j + t < n) {
        Lc[(size_t)j * B1 + t] = v;
        Lr[(size_t)(j + t) * B1 + t] = v;
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MAXP; q++) {
      const int t1 = pt1[q], t2 = pt2[q];
      if (t1 > 0 && j + t1 < n) W[(j + t1) & 63][(j + t2) & 63] -= col[t1] * col[t2];
    }
    const int inew = j + B1;
    if (inew < n && (int)threadIdx.x < B1) {
      const int l = j + 1 + (int)threadIdx.x;
      W[inew & 63][l & 63] = s_entry(T, S3, inew, l);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) info[0] = bad;
}

// lane `lane`'s value in every lane (two 32-bit readlanes; lane is uniform)
__device__ inline double readlane64(double v, int lane) {
  const unsigned long long u = __double_as_longlong(v);
  const unsigned lo = __builtin_amdgcn_readlane((unsigned)u, lane);
  const unsigned hi = __builtin_amdgcn_readlane((unsigned)(u >> 32), lane);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Row j of X = S^-1 (= column j, S symmetric) by banded substitution of e_j,
// both passes in axpy form with the band spread over the lanes: row r lives
// in lane r mod 64, so a step reads its pivot row with one readlane and no
// value moves between lanes.
//   forward  L y = e_j from row j on: step m takes y_m from lane m mod 64,
//            lanes of rows m+1 .. m+b subtract L(row, m) y_m (column m of L
//            from Lc), and lane m mod 64 then holds row m + 64 (zero);
//   backward L^T x = y from the last row: step i takes x_i, lanes of rows
//            i-b .. i-1 subtract L(i, row) x_i (row i of L from Lr), and the
//            row i-1-b enters (its lane holds a finished row, or row i when
//            b = 63) with y_{i-1-b} from the forward pass.
// One wavefront per j, four per block; y is kept in X's row j between the
// passes.  The steps run in chunks of TL_CH whose band values are loaded
// together, one chunk ahead (the loads are L2 hits whose latency is many
// steps long).
constexpr int TL_CH = 8;
__global__ void __launch_bounds__(256) k_tl_band_inv(TwoLevel T, const double *__restrict__ Lc,
                                                    const double *__restrict__ Lr,
                                                    double *__restrict__ X) {
  const int ln = threadIdx.x & 63;
  const int j = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
  if (j >= T.nc) return;
  const int n = T.nc, b = T.naj, B1 = b + 1;
  double *xr = X + (size_t)j * T.ld;
  auto lcol = [&](int m) {
    const int t = (ln - m) & 63;
    return (m < n && t >= 1 && t <= b && m + t < n) ? Lc[(size_t)m * B1 + t] : 0.0;
  };
  auto rpiv = [&](int m) { return (m >= 0 && m < n) ? Lc[(size_t)m * B1] : 1.0; };
  double w = ln == (j & 63) ? 1.0 : 0.0;
  double lv[TL_CH], rd[TL_CH], lvn[TL_CH], rdn[TL_CH];
#pragma unroll
  for (int u = 0; u < TL_CH; u++) {
    lv[u] = lcol(j + u);
    rd[u] = rpiv(j + u);
  }
  for (int m0 = j; m0 < n; m0 += TL_CH) {
#pragma unroll
    for (int u = 0; u < TL_CH; u++) {
      lvn[u] = lcol(m0 + TL_CH + u);
      rdn[u] = rpiv(m0 + TL_CH + u);
    }
#pragma unroll
    for (int u = 0; u < TL_CH; u++) {
      const int m = m0 + u;
      if (m < n) {
        const double ym = readlane64(w, m & 63) * rd[u];
        if (ln == 0) xr[m] = ym;
        w = w - lv[u] * ym;
        if (ln == (m & 63)) w = 0.0;
      }
    }
#pragma unroll
    for (int u = 0; u < TL_CH; u++) {
      lv[u] = lvn[u];
      rd[u] = rdn[u];
    }
  }
  // the backward pass reads, in every lane, the y that lane 0 stored
  __threadfence_block();
  auto lrow = [&](int i) {
    const int t = (i - ln) & 63;
    return (i >= 0 && t >= 1 && t <= b && i - t >= 0) ? Lr[(size_t)i * B1 + t] : 0.0;
  };
  auto yraw = [&](int m) { return (m >= j && m >= 0) ? xr[m] : 0.0; };
  {
    const int t = (n - 1 - ln) & 63;
    w = t <= b ? yraw(n - 1 - t) : 0.0;
  }
  double yi[TL_CH], yin[TL_CH];
#pragma unroll
  for (int u = 0; u < TL_CH; u++) {
    lv[u] = lrow(n - 1 - u);
    rd[u] = rpiv(n - 1 - u);
    yi[u] = yraw(n - 2 - b - u);
  }
  for (int i0 = n - 1; i0 >= 0; i0 -= TL_CH) {
#pragma unroll
    for (int u = 0; u < TL_CH; u++) {
      lvn[u] = lrow(i0 - TL_CH - u);
      rdn[u] = rpiv(i0 - TL_CH - u);
      yin[u] = yraw(i0 - TL_CH - 1 - b - u);
    }
#pragma unroll
    for (int u = 0; u < TL_CH; u++) {
      const int i = i0 - u;
      if (i >= 0) {
        const double xi = readlane64(w, i & 63) * rd[u];
        if (ln == 0) xr[i] = xi;
        w = w - lv[u] * xi;
        if (ln == ((i - 1 - b) & 63)) w = yi[u];
      }
    }
#pragma unroll
    for (int u = 0; u < TL_CH; u++) {
      lv[u] = lvn[u];
      rd[u] = rdn[u];
      yi[u] = yin[u];
    }
  }
  for (int c = n + ln; c < T.ld; c += 64) xr[c] = 0.0;
}

// Ainv = -X row by row (zero beyond nc), FP64 and / or FP32
__global__ void k_tl_finish_rows(TwoLevel T, const double *__restrict__ X,
                                 double *__restrict__ A64, float *__restrict__ A32) {
  const size_t n = (size_t)T.nc * T.ld;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (size_t)gridDim.x * blockDim.x) {
    const int c = (int)(e % T.ld);
    const double v = c < T.nc ? -X[e] : 0.0;
    if (A64) A64[e] = v;
    if (A32) A32[e] = (float)v;
  }
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
                                                        const KState *st,
                                                        const double *__restrict__ colsq = nullptr,
                                                        const double *alpha = nullptr) {
  __shared__ double lds[MF6_BLOCK / 64];
  if (st && st->stop) return;
  // colsq given (colour layout, c HP + l / 2): P^T (r - alpha q); alpha = 0
  // leaves colsq unread
  const double a = colsq ? *alpha : 0.0;
  for (int I = blockIdx.x; I < T.nc; I += gridDim.x) {
    const int ai = I / T.naj, aj = I - ai * T.naj;
    double s = 0.0;
    for (int t = threadIdx.x; t < T.ha * T.ha; t += blockDim.x) {
      const int i = ai * T.ha + t / T.ha, j = aj * T.ha + t % T.ha;
      if (i < L.nrow && j < L.ncol) {
        const int l = i * L.ncol + j;
        double v = colsum[l];
        if (a != 0.0) v = v - a * colsq[((i + j) & 1) * L.HP + (l >> 1)];
        s = s + v;
      }
    }
    s = block_sum(s, lds);
    if (threadIdx.x == 0) rc[I] = s;
  }
}

// The same sums by aggregate-row strips: block (ai, chunk) takes TL_GCH
// consecutive grid columns (whole aggregates) of the strip's ha grid rows,
// one column per lane and every fourth row per wavefront, so each grid
// row's sums are read as whole 64-column segments (and each colour's as two
// interleaved ones) rather than as ha-wide pieces.  Each lane adds its rows
// in order, the four wavefronts' column totals are added in order, and the
// aggregate sums add ha consecutive column totals (fixed order).
constexpr int TL_GCH = 64;
__global__ void __launch_bounds__(4 * TL_GCH) k_tl_gather_rows(Layout L,
                                                              const double *__restrict__ colsum,
                                                              TwoLevel T, double *__restrict__ rc,
                                                              const KState *st,
                                                              const double *__restrict__ colsq,
                                                              const double *alpha) {
  __shared__ double acc[4][TL_GCH];
  if (st && st->stop) return;
  const double a = colsq ? *alpha : 0.0;
  const int per = (TL_GCH / T.ha) * T.ha;   // columns of whole aggregates per block
  const int nag = per / T.ha;
  const int nchunk = (L.ncol + per - 1) / per;
  const int w = threadIdx.x >> 6, ln = threadIdx.x & 63;
  for (int blk = blockIdx.x; blk < T.nai * nchunk; blk += gridDim.x) {
    const int ai = blk / nchunk, ch = blk - ai * nchunk;
    const int j = ch * per + ln;
    const int i0 = ai * T.ha;
    const int i1 = i0 + T.ha < L.nrow ? i0 + T.ha : L.nrow;
    double sum = 0.0;
    if (ln < per && j < L.ncol) {
#pragma unroll 4
      for (int i = i0 + w; i < i1; i += 4) {
        const int l = i * L.ncol + j;
        double v = colsum[l];
        if (a != 0.0) v = v - a * colsq[((i + j) & 1) * L.HP + (l >> 1)];
        sum = sum + v;
      }
    }
    acc[w][ln] = sum;
    __syncthreads();
    if ((int)threadIdx.x < nag) {
      const int aj = ch * nag + (int)threadIdx.x;
      if (aj < T.naj) {
        double t = 0.0;
        for (int q = 0; q < T.ha; q++) {
          const int cq = threadIdx.x * T.ha + q;
          t = t + (((acc[0][cq] + acc[1][cq]) + acc[2][cq]) + acc[3][cq]);
        }
        rc[ai * T.naj + aj] = t;
      }
    }
    __syncthreads();
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
constexpr int TL_GEMV_THREADS = MF6_BLOCK;
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

// colsum[l] = sum of d over the coupled cells of grid column l
template <typename TV>
__global__ void k_tl_colsum(Layout L, const TV *__restrict__ d, const unsigned *__restrict__ cmask,
                            double *__restrict__ colsum) {
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < L.HP0 + L.HP1;
       t += gridDim.x * blockDim.x) {
    const int c = t >= L.HP0;
    const int h2 = t - c * L.HP0;
    const unsigned m = cmask[c * L.HP + h2];
    double sum = 0.0;
    for (int k = 0; k < L.nlay; k++)
      if ((m >> k) & 1u) sum = sum + (double)d[c * L.H + k * L.HP + h2];
    colsum[col_l(L, c, h2)] = sum;
  }
}

// the same column sums in the colour layout (c HP + l / 2) the SpMV writes
// the column sums of q in, for the gather check of mf6hip_apply_op
template <typename TV>
__global__ void k_tl_colsum_colour(Layout L, const TV *__restrict__ d,
                                   const unsigned *__restrict__ cmask,
                                   double *__restrict__ colsq) {
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < L.HP0 + L.HP1;
       t += gridDim.x * blockDim.x) {
    const int c = t >= L.HP0;
    const int h2 = t - c * L.HP0;
    const unsigned m = cmask[c * L.HP + h2];
    double sum = 0.0;
    for (int k = 0; k < L.nlay; k++)
      if ((m >> k) & 1u) sum = sum + (double)d[c * L.H + k * L.HP + h2];
    colsq[c * L.HP + h2] = sum;
  }
}

// Position of grid column (i, j) in a row of ns: colour by colour within the
// row (the row's colour-0 cells first), so the lanes of one colour kernel
// read consecutive words.
__host__ __device__ inline int ns_col(const Layout &L, int i, int j) {
  if (((i + j) & 1) == 0) return j >> 1;
  return ((i & 1) ? (L.ncol >> 1) : ((L.ncol + 1) >> 1)) + (j >> 1);
}

// Per cell n (FP64 planes A): s[n] = A(n, n) + sum of A(n, m) over the
// coupled neighbours m (0 on decoupled cells), and the couplings A(n, m) to
// coupled neighbours across aggregate edges into ew / ns (layouts as the
// fused CG reads them: ew [(k nrow + i) 2 naj + 2 aj + e], e = 0 to j - 1
// at the aggregate's first column, 1 to j + 1 at its last; ns [(k 2 nai +
// 2 ai + e) ncol + ns_col(i, j)], the same in i).
__global__ void k_tl_defl_setup(Layout L, const double *__restrict__ A,
                                const unsigned *__restrict__ cmask, TwoLevel T,
                                float *__restrict__ sdiag, float *__restrict__ ew,
                                float *__restrict__ ns) {
  const size_t NP = (size_t)L.NP;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < L.HP0 + L.HP1;
       t += gridDim.x * blockDim.x) {
    const int c = t >= L.HP0;
    const int h2 = t - c * L.HP0;
    const int l = col_l(L, c, h2);
    const int i = l / L.ncol, j = l - i * L.ncol;
    const int ai = i / T.ha, aj = j / T.ha;
    const int ri = i - ai * T.ha, rj = j - aj * T.ha;
    const bool hb = i > 0, hf = i < L.nrow - 1, hl = j > 0, hr = j < L.ncol - 1;
    const unsigned m = cmask[c * L.HP + h2];
    const unsigned *om = cmask + (1 - c) * L.HP;
    const unsigned mb = hb ? om[hn(l, -L.ncol)] : 0u, mf = hf ? om[hn(l, L.ncol)] : 0u;
    const unsigned ml = hl ? om[hn(l, -1)] : 0u, mr = hr ? om[hn(l, 1)] : 0u;
    for (int k = 0; k < L.nlay; k++) {
      const size_t pos = (size_t)c * L.H + (size_t)k * L.HP + h2;
      const bool own = (m >> k) & 1u;
      const double ab = ((mb >> k) & 1u) ? A[2 * NP + pos] : 0.0;
      const double al = ((ml >> k) & 1u) ? A[3 * NP + pos] : 0.0;
      const double ar = ((mr >> k) & 1u) ? A[4 * NP + pos] : 0.0;
      const double af = ((mf >> k) & 1u) ? A[5 * NP + pos] : 0.0;
      const double au = (k > 0 && ((m >> (k - 1)) & 1u)) ? A[1 * NP + pos] : 0.0;
      const double ad = (k + 1 < L.nlay && ((m >> (k + 1)) & 1u)) ? A[6 * NP + pos] : 0.0;
      double sum = A[pos];
      sum = sum + au;
      sum = sum + ab;
      sum = sum + al;
      sum = sum + ar;
      sum = sum + af;
      sum = sum + ad;
      sdiag[pos] = own ? (float)sum : 0.0f;
      const size_t e0 = ((size_t)k * L.nrow + i) * 2 * T.naj + 2 * aj;
      const size_t n0 = ((size_t)k * 2 * T.nai + 2 * ai) * L.ncol + ns_col(L, i, j);
      if (rj == 0 && hl) ew[e0] = own ? (float)al : 0.0f;
      if (rj == T.ha - 1 && hr) ew[e0 + 1] = own ? (float)ar : 0.0f;
      if (ri == 0 && hb) ns[n0] = own ? (float)ab : 0.0f;
      if (ri == T.ha - 1 && hf) ns[n0 + L.ncol] = own ? (float)af : 0.0f;
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
// target coarse unknowns and, where that costs at most a doubling of ha, at
// most 63 aggregates per row (the bandwidth the banded setup handles).  The
// default target, sqrt(2 N) within [16, 4096], keeps the coarse GEMV (nc^2
// values per iteration) well below the fine iteration's bytes (~24 N).
TwoLevel mf6hip_tl_shape(const Layout &L, int target) {
  if (target <= 0) {
    target = (int)std::sqrt(2.0 * (double)L.N);
    if (target < 16) target = 16;
    if (target > 4096) target = 4096;
  }
  TwoLevel T{};
  int ha = (int)std::ceil(std::sqrt((double)L.NRC / (double)target));
  if (ha < 1) ha = 1;
  while (((L.nrow + ha - 1) / ha) * ((L.ncol + ha - 1) / ha) > target) ha++;
  const int ha0 = ha;
  while ((L.ncol + ha - 1) / ha > 63 && ha < 2 * ha0) ha++;
  T.ha = ha;
  T.nai = (L.nrow + ha - 1) / ha;
  T.naj = (L.ncol + ha - 1) / ha;
  T.nc = T.nai * T.naj;
  T.ld = (T.nc + 63) / 64 * 64;
  return T;
}

// Allocates every buffer of the two-level preconditioner; tl_ready is set
// only once all of them exist, and a failure frees what was allocated, so
// the next call starts over instead of running on missing buffers.
static int tl_alloc_all(Mf6Handle *h);

int mf6hip_tl_alloc(Mf6Handle *h) {
  if (h->tl_ready) return MF6HIP_OK;
  mf6hip_tl_free(h);
  const int rc = tl_alloc_all(h);
  if (rc != MF6HIP_OK) {
    mf6hip_tl_free(h);
    return rc;
  }
  h->tl_ready = true;
  return MF6HIP_OK;
}

static int tl_alloc_all(Mf6Handle *h) {
  const Layout &L = h->L;
  h->tl = mf6hip_tl_shape(L, h->s.coarse_size);
  const size_t nn = (size_t)h->tl.nc * h->tl.ld;
  if (hipMalloc((void **)&h->tl_cmask, sizeof(unsigned) * 2 * (size_t)L.HP) != hipSuccess ||
      hipMalloc((void **)&h->tl_G, sizeof(double) * nn) != hipSuccess ||
      hipMalloc((void **)&h->tl_rc, sizeof(double) * h->tl.ld) != hipSuccess ||
      hipMalloc((void **)&h->tl_yc64, sizeof(double) * h->tl.ld) != hipSuccess ||
      hipMalloc((void **)&h->tl_yc32, sizeof(float) * h->tl.ld) != hipSuccess ||
      hipMalloc((void **)&h->tl_info, sizeof(int)) != hipSuccess ||
      hipMalloc((void **)&h->tl_colsum, sizeof(double) * (size_t)L.NRC) != hipSuccess ||
      hipMalloc((void **)&h->tl_S3, sizeof(double) * 3 * (size_t)h->tl.nc) != hipSuccess ||
      hipMalloc((void **)&h->tl_Lc, sizeof(double) * 64 * (size_t)h->tl.nc) != hipSuccess ||
      hipMalloc((void **)&h->tl_Lr, sizeof(double) * 64 * (size_t)h->tl.nc) != hipSuccess) {
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
  if (h->s.preconditioner == 3) {
    const TwoLevel &T = h->tl;
    const size_t new_ = (size_t)L.nlay * L.nrow * 2 * T.naj;
    const size_t nns = (size_t)L.nlay * 2 * T.nai * L.ncol;
    if (hipMalloc((void **)&h->tl_s32, sizeof(float) * (size_t)L.NP) != hipSuccess ||
        hipMalloc((void **)&h->tl_ew, sizeof(float) * new_) != hipSuccess ||
        hipMalloc((void **)&h->tl_ns, sizeof(float) * nns) != hipSuccess ||
        hipMalloc((void **)&h->tl_colsq, sizeof(double) * 2 * (size_t)L.HP) != hipSuccess) {
      mf6hip_set_error("hipMalloc failed (two-level preconditioner)");
      return MF6HIP_EHIP;
    }
    HIPCHK(hipMemsetAsync(h->tl_s32, 0, sizeof(float) * (size_t)L.NP, h->stream));
    HIPCHK(hipMemsetAsync(h->tl_ew, 0, sizeof(float) * new_, h->stream));
    HIPCHK(hipMemsetAsync(h->tl_ns, 0, sizeof(float) * nns, h->stream));
    HIPCHK(hipMemsetAsync(h->tl_colsq, 0, sizeof(double) * 2 * (size_t)L.HP, h->stream));
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
  h->tl_ready = false;
  if (h->tl_blas) rocblas_destroy_handle((rocblas_handle)h->tl_blas);
  h->tl_blas = nullptr;
  void *bufs[] = {h->tl_cmask, h->tl_G, h->tl_rc, h->tl_yc64, h->tl_yc32, h->tl_info,
                  h->tl_A64, h->tl_A32, h->tl_colsum, h->tl_S3, h->tl_Lc, h->tl_Lr,
                  h->tl_s32, h->tl_ew, h->tl_ns, h->tl_colsq};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  h->tl_cmask = nullptr;
  h->tl_G = h->tl_A64 = h->tl_rc = h->tl_yc64 = nullptr;
  h->tl_A32 = h->tl_yc32 = nullptr;
  h->tl_info = nullptr;
  h->tl_colsum = nullptr;
  h->tl_S3 = h->tl_Lc = h->tl_Lr = nullptr;
  h->tl_s32 = h->tl_ew = h->tl_ns = nullptr;
  h->tl_colsq = nullptr;
}

// Coarse operator of the current FP64 planes: mask, S = -Ac, its inverse.
int mf6hip_tl_setup(Mf6Handle *h) {
  CK(mf6hip_tl_alloc(h));
  const Layout &L = h->L;
  const TwoLevel &T = h->tl;
  hipStream_t s = h->stream;
  k_tl_cmask<<<tl_grid(2 * (size_t)L.HP), MF6_BLOCK, 0, s>>>(L, h->A, h->tl_cmask);
  k_tl_build<<<T.nc, MF6_BLOCK, 0, s>>>(L, h->A, h->tl_cmask, T, h->tl_S3);
  if (h->s.preconditioner == 3)
    k_tl_defl_setup<<<tl_grid((size_t)L.HP0 + L.HP1), MF6_BLOCK, 0, s>>>(
        L, h->A, h->tl_cmask, T, h->tl_s32, h->tl_ew, h->tl_ns);
  HIPCHK(hipGetLastError());
  int info = 0;
  if (T.naj <= 63) {
    k_tl_band_chol<<<1, TL_CHOL_THREADS, 0, s>>>(T, h->tl_S3, h->tl_Lc, h->tl_Lr, h->tl_info);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(&info, h->tl_info, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (info == 0) {
      k_tl_band_inv<<<(T.nc + 3) / 4, 256, 0, s>>>(T, h->tl_Lc, h->tl_Lr, h->tl_G);
      k_tl_finish_rows<<<tl_grid((size_t)T.nc * T.ld), MF6_BLOCK, 0, s>>>(T, h->tl_G, h->tl_A64,
                                                                         h->tl_A32);
      HIPCHK(hipGetLastError());
    }
  } else {
    HIPCHK(hipMemsetAsync(h->tl_G, 0, sizeof(double) * (size_t)T.nc * T.ld, s));
    k_tl_expand<<<tl_grid(T.nc), MF6_BLOCK, 0, s>>>(T, h->tl_S3, h->tl_G);
    HIPCHK(hipGetLastError());
    rocblas_handle rb = (rocblas_handle)h->tl_blas;
    if (rocsolver_dpotrf(rb, rocblas_fill_lower, T.nc, h->tl_G, T.ld, h->tl_info) !=
        rocblas_status_success) {
      mf6hip_set_error("rocsolver_dpotrf failed");
      return MF6HIP_EHIP;
    }
    HIPCHK(hipMemcpyAsync(&info, h->tl_info, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (info == 0) {
      if (rocsolver_dpotri(rb, rocblas_fill_lower, T.nc, h->tl_G, T.ld, h->tl_info) !=
          rocblas_status_success) {
        mf6hip_set_error("rocsolver_dpotri failed");
        return MF6HIP_EHIP;
      }
      k_tl_finish<<<tl_grid((size_t)T.nc * T.ld), MF6_BLOCK, 0, s>>>(T, h->tl_G, h->tl_A64,
                                                                    h->tl_A32);
      HIPCHK(hipGetLastError());
    }
  }
  if (info != 0) {
    mf6hip_set_error("two-level preconditioner: the coarse matrix is not definite (leading minor " +
                     std::to_string(info) + "); the system must be symmetric definite");
    return MF6HIP_EINVAL;
  }
  return MF6HIP_OK;
}

// The banded coarse setup of one system whose FP64 planes are A (layout
// L), into caller-owned arrays: cmask (2 HP), S3 (3 nc), Lc / Lr (64 nc),
// scratch G (nc ld doubles) and the FP32 inverse A32 (nc ld).  For the
// ensemble solver; aggregate grids wider than 63 are not supported there.
// Returns the factorisation's info code through *info (0: definite).
int mf6hip_tl_setup_slice(const Layout &L, const double *A, const TwoLevel &T,
                          unsigned *cmask, double *S3, double *Lc, double *Lr, double *G,
                          float *A32, int *dinfo, int *info, hipStream_t s) {
  if (T.naj > 63) {
    mf6hip_set_error("ensemble two-level: aggregate grid wider than 63");
    return MF6HIP_EUNSUPPORTED;
  }
  k_tl_cmask<<<tl_grid(2 * (size_t)L.HP), MF6_BLOCK, 0, s>>>(L, A, cmask);
  k_tl_build<<<T.nc, MF6_BLOCK, 0, s>>>(L, A, cmask, T, S3);
  k_tl_band_chol<<<1, TL_CHOL_THREADS, 0, s>>>(T, S3, Lc, Lr, dinfo);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(info, dinfo, sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  if (*info != 0) return MF6HIP_OK;
  k_tl_band_inv<<<(T.nc + 3) / 4, 256, 0, s>>>(T, Lc, Lr, G);
  k_tl_finish_rows<<<tl_grid((size_t)T.nc * T.ld), MF6_BLOCK, 0, s>>>(T, G, nullptr, A32);
  HIPCHK(hipGetLastError());
  return MF6HIP_OK;
}

// blocks of the coarse GEMV (each writes one partial of rc.yc)
static int tl_gemv_grid(const TwoLevel &T, int nslot) {
  const int rows = TL_GEMV_THREADS / 64;
  int gg = (T.nc + rows - 1) / rows;
  if (gg > nslot) gg = nslot;
  return gg < 1 ? 1 : gg;
}

int mf6hip_tl_nslot(const Mf6Handle *h, int nslot) { return tl_gemv_grid(h->tl, nslot); }

// rc = P^T (colsum - alpha colsq) into tl_rc (alpha null: P^T colsum)
static void launch_tl_gather(Mf6Handle *h, const KState *st, const double *alpha) {
  const TwoLevel &T = h->tl;
  const Layout &L = h->L;
  const double *cq = alpha ? h->tl_colsq : nullptr;
  if (T.ha <= TL_GCH) {
    const int per = (TL_GCH / T.ha) * T.ha;
    int gr = T.nai * ((L.ncol + per - 1) / per);
    if (gr > 4096) gr = 4096;
    k_tl_gather_rows<<<gr, 4 * TL_GCH, 0, h->stream>>>(L, h->tl_colsum, T, h->tl_rc, st, cq,
                                                       alpha);
  } else {
    const int gr = T.nc < 4096 ? T.nc : 4096;
    k_tl_gather<<<gr, MF6_BLOCK, 0, h->stream>>>(L, h->tl_colsum, T, h->tl_rc, st, cq, alpha);
  }
}

// rc = P^T d (from the preconditioner kernels' column sums when d is null,
// else from d itself) ; yc = Ainv rc ; partial rc.yc into part[0 .. nslot)
// With alpha (a device pointer, the deflated form): rc = P^T (d - alpha q)
// from the column sums of d and of q.
template <typename TV>
void mf6hip_tl_enqueue(Mf6Handle *h, const TV *d, TV *yc, double *part, int nslot,
                       const KState *st, const double *alpha) {
  const TwoLevel &T = h->tl;
  hipStream_t s = h->stream;
  int gr = T.nc < 4096 ? T.nc : 4096;
  if (d)
    k_tl_restrict<TV><<<gr, MF6_BLOCK, 0, s>>>(h->L, d, h->tl_cmask, T, h->tl_rc, st);
  else
    launch_tl_gather(h, st, alpha);
  const int gg = tl_gemv_grid(T, nslot);
  const size_t shm = sizeof(double) * (size_t)T.ld;
  if (sizeof(TV) == sizeof(double))
    k_tl_gemv<double, TV><<<gg, TL_GEMV_THREADS, shm, s>>>(T, h->tl_A64, h->tl_rc, yc, part,
                                                           nslot, st);
  else
    k_tl_gemv<float, TV><<<gg, TL_GEMV_THREADS, shm, s>>>(T, h->tl_A32, h->tl_rc, yc, part,
                                                          nslot, st);
}

template void mf6hip_tl_enqueue<double>(Mf6Handle *, const double *, double *, double *, int,
                                        const KState *, const double *);
template void mf6hip_tl_enqueue<float>(Mf6Handle *, const float *, float *, double *, int,
                                       const KState *, const double *);

// column sums of a (replaced) residual d over the coupled cells, for the
// deflated form's first coarse solve after it
template <typename TV>
void mf6hip_tl_colsum(Mf6Handle *h, const TV *d) {
  const Layout &L = h->L;
  k_tl_colsum<TV><<<tl_grid((size_t)L.HP0 + L.HP1), MF6_BLOCK, 0, h->stream>>>(
      L, d, h->tl_cmask, h->tl_colsum);
}

template void mf6hip_tl_colsum<double>(Mf6Handle *, const double *);
template void mf6hip_tl_colsum<float>(Mf6Handle *, const float *);

// z += P yc (the FP64 coarse correction of the last mf6hip_tl_enqueue)
int mf6hip_tl_prolong64(Mf6Handle *h, double *z) {
  k_tl_prolong<double><<<tl_grid(2 * (size_t)h->L.HP), MF6_BLOCK, 0, h->stream>>>(
      h->L, h->tl_cmask, h->tl, h->tl_yc64, z);
  HIPCHK(hipGetLastError());
  return MF6HIP_OK;
}

// Restriction P^T x over the coupled cells two ways, for the parity helper
// mf6hip_apply_op (op 9): out[0, nc) by the deflated iteration's route, the
// gather of column sums with the colour-layout column sums of x as q and
// alpha = 0.75 (so P^T (x - 0.75 x)); out[nc, 2 nc) by k_tl_restrict.  x is
// a device vector in the position layout; out a device array of 2 nc.
int mf6hip_tl_gather_check(Mf6Handle *h, const double *x, double *out) {
  if (!h->tl_cmask || !h->tl_colsq || !h->tl_colsum) {
    mf6hip_set_error("op 9 needs preconditioner = 3 and a loaded system");
    return MF6HIP_EINVAL;
  }
  const Layout &L = h->L;
  const TwoLevel &T = h->tl;
  hipStream_t s = h->stream;
  static const double alpha_host = 0.75;
  double *alpha = h->d_part + 7 * (size_t)MF6_MAXBLK + 16;
  HIPCHK(hipMemcpyAsync(alpha, &alpha_host, sizeof(double), hipMemcpyHostToDevice, s));
  mf6hip_tl_colsum<double>(h, x);
  k_tl_colsum_colour<double><<<tl_grid((size_t)L.HP0 + L.HP1), MF6_BLOCK, 0, s>>>(
      L, x, h->tl_cmask, h->tl_colsq);
  launch_tl_gather(h, nullptr, alpha);
  HIPCHK(hipMemcpyAsync(out, h->tl_rc, sizeof(double) * T.nc, hipMemcpyDeviceToDevice, s));
  const int gr = T.nc < 4096 ? T.nc : 4096;
  k_tl_restrict<double><<<gr, MF6_BLOCK, 0, s>>>(L, x, h->tl_cmask, T, out + T.nc, nullptr);
  HIPCHK(hipGetLastError());
  return MF6HIP_OK;
}

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
