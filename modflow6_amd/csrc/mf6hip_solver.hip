// mf6hip_solver.hip -- IMS linear solve on gfx950.
//
// Reference algorithms (src/Solution/LinearMethods/ImsLinearBase.f90):
//   ims_base_cg :30-240, ims_base_bcgs :249-549, ims_base_pcilu0 :928-1042,
//   ims_base_ilu0a :1049-1092, ims_base_testcnvg :1101-1146,
//   ims_base_residual :1291-1312, imslinear_ap (ImsLinear.f90:617-750),
//   amux (src/Utilities/Libraries/sparsekit/sparsekit.f90:1-59).
//
// The device ILU0 / MILU0 is the reference's factorisation applied in the
// column red-black order: all cells of vertical columns with (i + j) even,
// column by column, top to bottom, then the odd columns.  On the 7-point
// stencil (a bipartite graph) ILU0 keeps no fill, so the factor is a vector
// of inverse pivots and L / U are the matrix entries scaled on the fly; the
// factor and apply kernels reproduce ims_base_pcilu0 / ims_base_ilu0a on the
// permuted matrix bit for bit (no FMA contraction, same operation order).
// Row sums, dot products and norms use fixed-shape tree reductions, so a
// solve is deterministic run to run.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include <hip/hip_fp16.h>

#include "mf6hip_handle.h"

#define DPREC_D 2.220446049250313080847263336181640625e-16
#define DSAME_D (100.0 * DPREC_D)

static std::mutex g_err_mu;
static std::string g_err;

void mf6hip_set_error(const std::string &msg) {
  std::lock_guard<std::mutex> lk(g_err_mu);
  g_err = msg;
}

extern "C" int mf6hip_last_error(char *buf, int buflen) {
  std::lock_guard<std::mutex> lk(g_err_mu);
  if (!buf || buflen <= 0) return MF6HIP_EINVAL;
  int n = (int)g_err.size();
  if (n > buflen - 1) n = buflen - 1;
  memcpy(buf, g_err.data(), n);
  buf[n] = 0;
  return MF6HIP_OK;
}

// ---------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------

template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

__device__ inline MaxLoc wave_maxloc(MaxLoc a) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    MaxLoc b;
    b.v = __shfl_xor(a.v, o);
    b.loc = __shfl_xor(a.loc, o);
    a = maxloc_pick(a, b);
  }
  return a;
}

// Sum over the block (256 threads = 4 waves), result valid in thread 0.
__device__ inline double block_sum(double v, double *lds) {
  v = wave_sum(v);
  int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) lds[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) s = s + lds[i];
  }
  return s;
}

__device__ inline MaxLoc block_maxloc(MaxLoc a, MaxLoc *lds) {
  a = wave_maxloc(a);
  int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) lds[w] = a;
  __syncthreads();
  MaxLoc r{0.0, -1};
  if (threadIdx.x == 0) {
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) r = maxloc_pick(r, lds[i]);
  }
  return r;
}

// Final reduction of nb partial sums inside a single block.
__device__ inline double reduce_parts(const double *part, int nb, double *lds) {
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) s = s + part[i];
  return block_sum(s, lds);
}

__device__ inline MaxLoc reduce_mparts(const MaxLoc *part, int nb, MaxLoc *lds) {
  MaxLoc a{0.0, -1};
  for (int i = threadIdx.x; i < nb; i += blockDim.x) a = maxloc_pick(a, part[i]);
  return block_maxloc(a, lds);
}

// ---------------------------------------------------------------------------
// layout conversion
// ---------------------------------------------------------------------------

// CSR row of natural node n (reference layout) -> 7 planes at position p.
__global__ void k_csr_to_planes(Layout L, const int *__restrict__ ia,
                                const double *__restrict__ amat,
                                double *__restrict__ A) {
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < L.N;
       n += gridDim.x * blockDim.x) {
    int k = n / L.NRC, r = n - k * L.NRC, i = r / L.ncol, j = r - i * L.ncol;
    int p = cell_pos(L, k, i, j);
    int e = ia[n];
    const int N = L.NP;
    A[p] = amat[e++];
    A[N + p] = k > 0 ? amat[e++] : 0.0;
    A[2 * N + p] = i > 0 ? amat[e++] : 0.0;
    A[3 * N + p] = j > 0 ? amat[e++] : 0.0;
    A[4 * N + p] = j < L.ncol - 1 ? amat[e++] : 0.0;
    A[5 * N + p] = i < L.nrow - 1 ? amat[e++] : 0.0;
    A[6 * N + p] = k < L.nlay - 1 ? amat[e++] : 0.0;
  }
}

__global__ void k_planes_to_csr(Layout L, const int *__restrict__ ia,
                                const double *__restrict__ A,
                                double *__restrict__ amat) {
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < L.N;
       n += gridDim.x * blockDim.x) {
    int k = n / L.NRC, r = n - k * L.NRC, i = r / L.ncol, j = r - i * L.ncol;
    int p = cell_pos(L, k, i, j);
    int e = ia[n];
    const int N = L.NP;
    amat[e++] = A[p];
    if (k > 0) amat[e++] = A[N + p];
    if (i > 0) amat[e++] = A[2 * N + p];
    if (j > 0) amat[e++] = A[3 * N + p];
    if (j < L.ncol - 1) amat[e++] = A[4 * N + p];
    if (i < L.nrow - 1) amat[e++] = A[5 * N + p];
    if (k < L.nlay - 1) amat[e++] = A[6 * N + p];
  }
}

template <typename TI, typename TO>
__global__ void k_nat_to_pos(Layout L, const TI *__restrict__ in, TO *__restrict__ out) {
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < L.N;
       n += gridDim.x * blockDim.x) {
    int k = n / L.NRC, r = n - k * L.NRC, i = r / L.ncol, j = r - i * L.ncol;
    out[cell_pos(L, k, i, j)] = (TO)in[n];
  }
}

template <typename TI, typename TO>
__global__ void k_pos_to_nat(Layout L, const TI *__restrict__ in, TO *__restrict__ out) {
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < L.N;
       n += gridDim.x * blockDim.x) {
    int k = n / L.NRC, r = n - k * L.NRC, i = r / L.ncol, j = r - i * L.ncol;
    out[n] = (TO)in[cell_pos(L, k, i, j)];
  }
}

// max |x| over n values, as the bit pattern of a non-negative double (whose
// integer order is the value order)
// (one atomic per block: reduced over the block in LDS first)
__global__ void __launch_bounds__(MF6_BLOCK) k_maxabs_bits(int n, const double *__restrict__ x,
                                                           unsigned long long *out) {
  __shared__ double lds[MF6_BLOCK / 64];
  double m = 0.0;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x)
    m = fmax(m, fabs(x[e]));
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = fmax(m, lds[w]);
  if (m > 0.0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

// y = (FP16)(x * scale) (scale a power of two: exact before the rounding)
__global__ void k_planes_to_half(int n, const double *__restrict__ x, double scale,
                                 _Float16 *__restrict__ y) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x)
    y[e] = (_Float16)(x[e] * scale);
}

// y = (float)(x * scale)
__global__ void k_scale_to_float(int n, const double *__restrict__ x, double scale,
                                 float *__restrict__ y) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x)
    y[e] = (float)(x[e] * scale);
}

template <typename TI, typename TO>
__global__ void k_convert(int n, const TI *__restrict__ in, TO *__restrict__ out) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x)
    out[e] = (TO)in[e];
}

__global__ void k_to_half(int n, const double *__restrict__ in, unsigned short *__restrict__ out) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    __half hv = __float2half_rn((float)in[e]);
    out[e] = *reinterpret_cast<unsigned short *>(&hv);
  }
}

// ---------------------------------------------------------------------------
// SpMV  y = A x  (amux order: t = 0, then slots in CSR order)
// ---------------------------------------------------------------------------

template <typename TM, typename TV>
__device__ inline TV row_dot(const Layout &L, const TM *__restrict__ A,
                             const TV *__restrict__ x, int p, int c, int k,
                             int i, int j, int l) {
  const int N = L.NP;
  const TV *xo = x + (1 - c) * L.H + k * L.HP;
  TV t = (TV)0;
  t = t + (TV)A[p] * x[p];
  if (k > 0) t = t + (TV)A[N + p] * x[p - L.HP];
  if (i > 0) t = t + (TV)A[2 * N + p] * xo[hn(l, -L.ncol)];
  if (j > 0) t = t + (TV)A[3 * N + p] * xo[hn(l, -1)];
  if (j < L.ncol - 1) t = t + (TV)A[4 * N + p] * xo[hn(l, 1)];
  if (i < L.nrow - 1) t = t + (TV)A[5 * N + p] * xo[hn(l, L.ncol)];
  if (k < L.nlay - 1) t = t + (TV)A[6 * N + p] * x[p + L.HP];
  return t;
}

template <typename TM, typename TV>
__global__ void __launch_bounds__(MF6_BLOCK) k_spmv(Layout L, const TM *__restrict__ A,
                                                    const TV *__restrict__ x,
                                                    TV *__restrict__ y) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < L.NP;
       p += gridDim.x * blockDim.x) {
    int c, k, i, j, l;
    if (!pos_to_cell(L, p, c, k, i, j, l)) continue;
    y[p] = row_dot(L, A, x, p, c, k, i, j, l);
  }
}

// d = b - A x (ims_base_residual), optionally gated by the solve state
template <typename TM, typename TV>
__global__ void __launch_bounds__(MF6_BLOCK) k_residual(Layout L, const TM *__restrict__ A,
                                                        const TV *__restrict__ x,
                                                        const TV *__restrict__ b,
                                                        TV *__restrict__ d,
                                                        KState *st) {
  if (st) {
    if (st->stop || !st->north_due) return;
  }
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < L.NP;
       p += gridDim.x * blockDim.x) {
    int c, k, i, j, l;
    if (!pos_to_cell(L, p, c, k, i, j, l)) continue;
    TV t = row_dot(L, A, x, p, c, k, i, j, l);
    d[p] = b[p] - t;
  }
}

// ---------------------------------------------------------------------------
// ILU0 / MILU0 factor, column red-black order (ims_base_pcilu0 on the
// permuted matrix).  Permuted index of column h2, layer k of colour c:
// c*H + h2*nlay + k (colour 0 first, column by column, top to bottom).
// ---------------------------------------------------------------------------

__device__ inline int perm_index(const Layout &L, int c, int k, int h2) {
  return c * L.H + h2 * L.nlay + k;
}

// Red columns: the only lower entry of a cell is the one above it.
__global__ void k_ilu_factor_red(Layout L, const double *__restrict__ A, double relax,
                                 double *__restrict__ piv, int *fail) {
  const int N = L.NP;
  for (int h2 = blockIdx.x * blockDim.x + threadIdx.x; h2 < L.HP0;
       h2 += gridDim.x * blockDim.x) {
    int l = col_l(L, 0, h2);
    int i = l / L.ncol, j = l - i * L.ncol;
    double pivup = 0.0;
    for (int k = 0; k < L.nlay; k++) {
      int p = k * L.HP + h2;
      double w = 0.0;
      w = w + A[p];
      double rs = 0.0;
      if (k > 0) {
        int pu = p - L.HP;
        double tl = A[N + p] * pivup;            // W(up) * APC(up)
        w = w - tl * A[6 * N + pu];              // U(up, n): down slot of up
        if (relax != 0.0) {                      // dropped fill of row up
          if (i > 0) rs = rs + tl * A[2 * N + pu];
          if (j > 0) rs = rs + tl * A[3 * N + pu];
          if (j < L.ncol - 1) rs = rs + tl * A[4 * N + pu];
          if (i < L.nrow - 1) rs = rs + tl * A[5 * N + pu];
        }
      }
      double d = w;
      double t = (1.0 + 0.0) * d - (relax * rs);
      if (copysign(fabs(d), t) != d || fabs(t) == 0.0) {
        atomicMin(fail, perm_index(L, 0, k, h2));
        break;
      }
      pivup = 1.0 / t;
      piv[p] = pivup;
    }
  }
}

// Black columns: lower entries are the four red neighbours (back, left,
// right, front) and the cell above.
__global__ void k_ilu_factor_black(Layout L, const double *__restrict__ A, double relax,
                                   double *__restrict__ piv, int *fail) {
  if (*fail < L.H) return;   // a red row failed: the reference stopped there
  const int N = L.NP;
  for (int h2 = blockIdx.x * blockDim.x + threadIdx.x; h2 < L.HP1;
       h2 += gridDim.x * blockDim.x) {
    int l = col_l(L, 1, h2);
    int i = l / L.ncol, j = l - i * L.ncol;
    double pivup = 0.0;
    for (int k = 0; k < L.nlay; k++) {
      int p = L.H + k * L.HP + h2;
      const int rb = k * L.HP;   // red plane base at this layer
      double w = 0.0;
      w = w + A[p];
      double rs = 0.0;
      // red neighbours in ascending permuted order: back, left, right, front
      int nb[4];
      int slot_rn[4];   // slot of the neighbour in n's row
      int cnt = 0;
      if (i > 0) { nb[cnt] = rb + hn(l, -L.ncol); slot_rn[cnt] = S_BACK; cnt++; }
      if (j > 0) { nb[cnt] = rb + hn(l, -1); slot_rn[cnt] = S_LEFT; cnt++; }
      if (j < L.ncol - 1) { nb[cnt] = rb + hn(l, 1); slot_rn[cnt] = S_RIGHT; cnt++; }
      if (i < L.nrow - 1) { nb[cnt] = rb + hn(l, L.ncol); slot_rn[cnt] = S_FRONT; cnt++; }
      for (int t = 0; t < cnt; t++) {
        int r = nb[t];
        int srn = slot_rn[t];
        int snr = S_DIAG + S_FRONT + S_BACK - srn;   // back<->front, left<->right
        double tl = A[srn * N + p] * piv[r];
        // upper entries of red row r, ascending: down (k+1), then its black
        // neighbours back, left, right, front (one of them is n)
        if (relax != 0.0) {
          if (k < L.nlay - 1) rs = rs + tl * A[6 * N + r];
        }
        int ir = (srn == S_BACK) ? i - 1 : (srn == S_FRONT ? i + 1 : i);
        int jr = (srn == S_LEFT) ? j - 1 : (srn == S_RIGHT ? j + 1 : j);
#pragma unroll
        for (int sl = S_BACK; sl <= S_FRONT; sl++) {
          bool present = (sl == S_BACK) ? ir > 0
                       : (sl == S_LEFT) ? jr > 0
                       : (sl == S_RIGHT) ? jr < L.ncol - 1
                                         : ir < L.nrow - 1;
          if (!present) continue;
          if (sl == snr) {
            w = w - tl * A[sl * N + r];
          } else if (relax != 0.0) {
            rs = rs + tl * A[sl * N + r];
          }
        }
      }
      if (k > 0) {
        int pu = p - L.HP;
        double tl = A[N + p] * pivup;
        w = w - tl * A[6 * N + pu];
      }
      double d = w;
      double tt = (1.0 + 0.0) * d - (relax * rs);
      if (copysign(fabs(d), tt) != d || fabs(tt) == 0.0) {
        atomicMin(fail, perm_index(L, 1, k, h2));
        break;
      }
      pivup = 1.0 / tt;
      piv[p] = pivup;
    }
  }
}

// Commit the new pivots of every row before the first failing one; the
// others keep their previous values (ImsLinearBase.f90:1040-1041: the
// reference leaves APC of the remaining rows untouched).
__global__ void k_ilu_commit(Layout L, const double *__restrict__ tmp,
                             double *__restrict__ piv, const int *fail) {
  int f = *fail;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < L.NP;
       p += gridDim.x * blockDim.x) {
    int c, k, i, j, l;
    if (!pos_to_cell(L, p, c, k, i, j, l)) continue;
    if (perm_index(L, c, k, l >> 1) < f) piv[p] = tmp[p];
  }
}

// ---------------------------------------------------------------------------
// ILU0 apply (ims_base_ilu0a on the permuted matrix)
// ---------------------------------------------------------------------------

template <typename TP>
__device__ inline double ldpiv(const TP *p, int e) { return (double)p[e]; }
template <>
__device__ inline double ldpiv<unsigned short>(const unsigned short *p, int e) {
  unsigned short u = p[e];
  __half h = *reinterpret_cast<__half *>(&u);
  return (double)__half2float(h);
}

// forward sweep of the red columns: zf = L^-1 r restricted to red rows
template <typename TM, typename TV, typename TP>
__global__ void __launch_bounds__(MF6_BLOCK) k_ilu_fwd_red(Layout L, const TM *__restrict__ A,
                                                           const TP *__restrict__ piv,
                                                           const TV *__restrict__ r,
                                                           TV *__restrict__ zf,
                                                           const KState *st) {
  if (st && st->stop) return;
  const int N = L.NP;
  for (int h2 = blockIdx.x * blockDim.x + threadIdx.x; h2 < L.HP0;
       h2 += gridDim.x * blockDim.x) {
    TV zup = (TV)0;
    for (int k = 0; k < L.nlay; k++) {
      int p = k * L.HP + h2;
      TV tv = r[p];
      if (k > 0) tv = tv - (TV)((TV)A[N + p] * (TV)ldpiv(piv, p - L.HP)) * zup;
      zf[p] = tv;
      zup = tv;
    }
  }
}

// black columns: forward (red neighbours + cell above) then backward (cell
// below), whole column in registers
template <typename TM, typename TV, typename TP, int NLMAX>
__global__ void __launch_bounds__(MF6_BLOCK) k_ilu_black(Layout L, const TM *__restrict__ A,
                                                         const TP *__restrict__ piv,
                                                         const TV *__restrict__ r,
                                                         const TV *__restrict__ zf,
                                                         TV *__restrict__ z,
                                                         const KState *st) {
  if (st && st->stop) return;
  const int N = L.NP;
  for (int h2 = blockIdx.x * blockDim.x + threadIdx.x; h2 < L.HP1;
       h2 += gridDim.x * blockDim.x) {
    int l = col_l(L, 1, h2);
    int i = l / L.ncol, j = l - i * L.ncol;
    int ib = i > 0 ? hn(l, -L.ncol) : -1;
    int il = j > 0 ? hn(l, -1) : -1;
    int ir = j < L.ncol - 1 ? hn(l, 1) : -1;
    int ifr = i < L.nrow - 1 ? hn(l, L.ncol) : -1;
    TV col[NLMAX];
    TV zup = (TV)0;
#pragma unroll
    for (int k = 0; k < NLMAX; k++) {
      if (k < L.nlay) {
        int p = L.H + k * L.HP + h2;
        int rb = k * L.HP;
        TV tv = r[p];
        if (ib >= 0) tv = tv - (TV)((TV)A[2 * N + p] * (TV)ldpiv(piv, rb + ib)) * zf[rb + ib];
        if (il >= 0) tv = tv - (TV)((TV)A[3 * N + p] * (TV)ldpiv(piv, rb + il)) * zf[rb + il];
        if (ir >= 0) tv = tv - (TV)((TV)A[4 * N + p] * (TV)ldpiv(piv, rb + ir)) * zf[rb + ir];
        if (ifr >= 0) tv = tv - (TV)((TV)A[5 * N + p] * (TV)ldpiv(piv, rb + ifr)) * zf[rb + ifr];
        if (k > 0) tv = tv - (TV)((TV)A[N + p] * (TV)ldpiv(piv, p - L.HP)) * zup;
        col[k] = tv;
        zup = tv;
      }
    }
    TV zdn = (TV)0;
#pragma unroll
    for (int k = NLMAX - 1; k >= 0; k--) {
      if (k < L.nlay) {
        int p = L.H + k * L.HP + h2;
        TV tv = col[k];
        if (k < L.nlay - 1) tv = tv - (TV)A[6 * N + p] * zdn;
        tv = tv * (TV)ldpiv(piv, p);
        z[p] = tv;
        zdn = tv;
      }
    }
  }
}

// backward sweep of the red columns: cell below, then black neighbours
template <typename TM, typename TV, typename TP>
__global__ void __launch_bounds__(MF6_BLOCK) k_ilu_bwd_red(Layout L, const TM *__restrict__ A,
                                                           const TP *__restrict__ piv,
                                                           const TV *__restrict__ zf,
                                                           TV *__restrict__ z,
                                                           const KState *st) {
  if (st && st->stop) return;
  const int N = L.NP;
  for (int h2 = blockIdx.x * blockDim.x + threadIdx.x; h2 < L.HP0;
       h2 += gridDim.x * blockDim.x) {
    int l = col_l(L, 0, h2);
    int i = l / L.ncol, j = l - i * L.ncol;
    int ib = i > 0 ? hn(l, -L.ncol) : -1;
    int il = j > 0 ? hn(l, -1) : -1;
    int ir = j < L.ncol - 1 ? hn(l, 1) : -1;
    int ifr = i < L.nrow - 1 ? hn(l, L.ncol) : -1;
    TV zdn = (TV)0;
    for (int k = L.nlay - 1; k >= 0; k--) {
      int p = k * L.HP + h2;
      const TV *zb = z + L.H + k * L.HP;
      TV tv = zf[p];
      if (k < L.nlay - 1) tv = tv - (TV)A[6 * N + p] * zdn;
      if (ib >= 0) tv = tv - (TV)A[2 * N + p] * zb[ib];
      if (il >= 0) tv = tv - (TV)A[3 * N + p] * zb[il];
      if (ir >= 0) tv = tv - (TV)A[4 * N + p] * zb[ir];
      if (ifr >= 0) tv = tv - (TV)A[5 * N + p] * zb[ifr];
      tv = tv * (TV)ldpiv(piv, p);
      z[p] = tv;
      zdn = tv;
    }
  }
}

// ---------------------------------------------------------------------------
// vector kernels
// ---------------------------------------------------------------------------

// partial dot products of two vectors
template <typename TV>
__global__ void __launch_bounds__(MF6_BLOCK) k_dot(int n, const TV *__restrict__ a,
                                                   const TV *__restrict__ b,
                                                   double *__restrict__ part,
                                                   const KState *st) {
  __shared__ double lds[8];
  if (st && st->stop) return;
  double s = 0.0;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x)
    s = s + (double)a[e] * (double)b[e];
  s = block_sum(s, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// CG: p_out = z (first) or z + beta p_in; q = A p_out; partial p.q
template <typename TM, typename TV>
__global__ void __launch_bounds__(MF6_BLOCK) k_cg_spmv(Layout L, const TM *__restrict__ A,
                                                       const TV *__restrict__ z,
                                                       const TV *__restrict__ pin,
                                                       TV *__restrict__ pout,
                                                       TV *__restrict__ q,
                                                       double *__restrict__ part,
                                                       const KState *st) {
  __shared__ double lds[8];
  if (st->stop) return;
  const bool first = st->iter == 0;
  const TV beta = (TV)st->beta;
  const int N = L.NP;
  double s = 0.0;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < L.NP;
       p += gridDim.x * blockDim.x) {
    int c, k, i, j, l;
    if (!pos_to_cell(L, p, c, k, i, j, l)) continue;
#define PV(e) (first ? z[e] : (TV)(z[e] + beta * pin[e]))
    const int ob = (1 - c) * L.H + k * L.HP;
    TV pp = PV(p);
    TV t = (TV)0;
    t = t + (TV)A[p] * pp;
    if (k > 0) t = t + (TV)A[N + p] * PV(p - L.HP);
    if (i > 0) t = t + (TV)A[2 * N + p] * PV(ob + hn(l, -L.ncol));
    if (j > 0) t = t + (TV)A[3 * N + p] * PV(ob + hn(l, -1));
    if (j < L.ncol - 1) t = t + (TV)A[4 * N + p] * PV(ob + hn(l, 1));
    if (i < L.nrow - 1) t = t + (TV)A[5 * N + p] * PV(ob + hn(l, L.ncol));
    if (k < L.nlay - 1) t = t + (TV)A[6 * N + p] * PV(p + L.HP);
#undef PV
    pout[p] = pp;
    q[p] = t;
    s = s + (double)pp * (double)t;
  }
  s = block_sum(s, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// CG update: x += alpha p ; d -= alpha q ; maxima and sum of squares
template <typename TV, typename TX>
__global__ void __launch_bounds__(MF6_BLOCK) k_cg_update(Layout L, TX *__restrict__ x,
                                                         const TV *__restrict__ p,
                                                         TV *__restrict__ d,
                                                         const TV *__restrict__ q,
                                                         double *__restrict__ part,
                                                         MaxLoc *__restrict__ mpart,
                                                         const KState *st, int nblk) {
  __shared__ double lds[8];
  __shared__ MaxLoc mlds[8];
  if (st->stop) return;
  const TV alpha = (TV)st->alpha;
  double s = 0.0;
  MaxLoc mx{0.0, -1}, mr{0.0, -1};
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < L.NP;
       e += gridDim.x * blockDim.x) {
    int c, k, i, j, l;
    if (!pos_to_cell(L, e, c, k, i, j, l)) continue;
    int nat = natural(L, k, i, j);
    TV tv = alpha * p[e];
    x[e] = (TX)(x[e] + (TX)tv);
    mx = maxloc_pick(mx, MaxLoc{(double)tv, nat});
    TV dv = d[e];
    dv = dv - alpha * q[e];
    d[e] = dv;
    mr = maxloc_pick(mr, MaxLoc{(double)dv, nat});
    s = s + (double)dv * (double)dv;
  }
  s = block_sum(s, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
  mx = block_maxloc(mx, mlds);
  if (threadIdx.x == 0) mpart[blockIdx.x] = mx;
  mr = block_maxloc(mr, mlds);
  if (threadIdx.x == 0) mpart[nblk + blockIdx.x] = mr;
}

// BiCGSTAB: p = d (first) or d + beta (p - omega0 v)
template <typename TV>
__global__ void __launch_bounds__(MF6_BLOCK) k_bcgs_p(int n, const TV *__restrict__ d,
                                                      TV *__restrict__ p,
                                                      const TV *__restrict__ v,
                                                      const KState *st) {
  if (st->stop) return;
  const bool first = st->iter == 0;
  const TV beta = (TV)st->beta, omega0 = (TV)st->omega0;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x)
    p[e] = first ? d[e] : d[e] + beta * (p[e] - omega0 * v[e]);
}

// BiCGSTAB: v = A phat with partial dhat.v
template <typename TM, typename TV>
__global__ void __launch_bounds__(MF6_BLOCK) k_spmv_dot(Layout L, const TM *__restrict__ A,
                                                        const TV *__restrict__ x,
                                                        TV *__restrict__ y,
                                                        const TV *__restrict__ w,
                                                        double *__restrict__ part,
                                                        const KState *st) {
  __shared__ double lds[8];
  if (st->stop) return;
  double s = 0.0;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < L.NP;
       p += gridDim.x * blockDim.x) {
    int c, k, i, j, l;
    if (!pos_to_cell(L, p, c, k, i, j, l)) continue;
    TV t = row_dot(L, A, x, p, c, k, i, j, l);
    y[p] = t;
    s = s + (double)w[p] * (double)t;
  }
  s = block_sum(s, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// BiCGSTAB: q = d - alpha v
template <typename TV>
__global__ void __launch_bounds__(MF6_BLOCK) k_bcgs_q(int n, const TV *__restrict__ d,
                                                      const TV *__restrict__ v,
                                                      TV *__restrict__ q,
                                                      const KState *st) {
  if (st->stop) return;
  const TV alpha = (TV)st->alpha;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x)
    q[e] = d[e] - alpha * v[e];
}

// BiCGSTAB: t = A qhat with partial t.q and t.t
template <typename TM, typename TV>
__global__ void __launch_bounds__(MF6_BLOCK) k_bcgs_t(Layout L, const TM *__restrict__ A,
                                                      const TV *__restrict__ qhat,
                                                      TV *__restrict__ t,
                                                      const TV *__restrict__ q,
                                                      double *__restrict__ part,
                                                      const KState *st, int nblk) {
  __shared__ double lds[8];
  if (st->stop) return;
  double s1 = 0.0, s2 = 0.0;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < L.NP;
       p += gridDim.x * blockDim.x) {
    int c, k, i, j, l;
    if (!pos_to_cell(L, p, c, k, i, j, l)) continue;
    TV tv = row_dot(L, A, qhat, p, c, k, i, j, l);
    t[p] = tv;
    s1 = s1 + (double)tv * (double)q[p];
    s2 = s2 + (double)tv * (double)tv;
  }
  s1 = block_sum(s1, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = s1;
  s2 = block_sum(s2, lds);
  if (threadIdx.x == 0) part[nblk + blockIdx.x] = s2;
}

// BiCGSTAB update: x += alpha phat + omega qhat ; d = q - omega t
template <typename TV, typename TX>
__global__ void __launch_bounds__(MF6_BLOCK) k_bcgs_update(Layout L, TX *__restrict__ x,
                                                           const TV *__restrict__ phat,
                                                           const TV *__restrict__ qhat,
                                                           TV *__restrict__ d,
                                                           const TV *__restrict__ q,
                                                           const TV *__restrict__ t,
                                                           double *__restrict__ part,
                                                           MaxLoc *__restrict__ mpart,
                                                           const KState *st, int nblk,
                                                           const double *__restrict__ dscale) {
  __shared__ double lds[8];
  __shared__ MaxLoc mlds[8];
  if (st->stop) return;
  const TV alpha = (TV)st->alpha, omega = (TV)st->omega;
  double s = 0.0;
  MaxLoc mx{0.0, -1}, mr{0.0, -1};
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < L.NP;
       e += gridDim.x * blockDim.x) {
    int c, k, i, j, l;
    if (!pos_to_cell(L, e, c, k, i, j, l)) continue;
    int nat = natural(L, k, i, j);
    TV tv = alpha * phat[e] + omega * qhat[e];
    x[e] = (TX)(x[e] + (TX)tv);
    double dv = (double)tv;
    if (dscale) dv = dv * dscale[e];   // ImsLinearBase.f90:452-454
    mx = maxloc_pick(mx, MaxLoc{dv, nat});
    tv = q[e] - omega * t[e];
    d[e] = tv;
    double rv = (double)tv;
    if (dscale) rv = rv / dscale[e];   // ImsLinearBase.f90:467-469
    mr = maxloc_pick(mr, MaxLoc{rv, nat});
    s = s + rv * rv;
  }
  s = block_sum(s, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
  mx = block_maxloc(mx, mlds);
  if (threadIdx.x == 0) mpart[blockIdx.x] = mx;
  mr = block_maxloc(mr, mlds);
  if (threadIdx.x == 0) mpart[nblk + blockIdx.x] = mr;
}

// max |v| with location, and sum of squares (scaled), over a vector
__global__ void __launch_bounds__(MF6_BLOCK) k_maxloc_ss(Layout L, const double *__restrict__ v,
                                                         double *__restrict__ part,
                                                         MaxLoc *__restrict__ mpart,
                                                         double scale_inv) {
  __shared__ double lds[8];
  __shared__ MaxLoc mlds[8];
  double s = 0.0;
  MaxLoc m{0.0, -1};
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < L.NP;
       e += gridDim.x * blockDim.x) {
    int c, k, i, j, l;
    if (!pos_to_cell(L, e, c, k, i, j, l)) continue;
    double t = v[e];
    m = maxloc_pick(m, MaxLoc{t, natural(L, k, i, j)});
    double u = t * scale_inv;
    s = s + u * u;
  }
  s = block_sum(s, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
  m = block_maxloc(m, mlds);
  if (threadIdx.x == 0) mpart[blockIdx.x] = m;
}

// ---------------------------------------------------------------------------
// scalar kernels (one block)
// ---------------------------------------------------------------------------

__device__ inline int is_close_d(double a, double b) {
  if (a == b) return 1;
  return fabs(a - b) <= fmax(DSAME_D * fmax(fabs(a), fabs(b)), 0.0);
}

__device__ inline void testcnvg(int icnvgopt, int *icnvg, int iiter, double dvmax,
                                double rmax, double rmax0, double epfact,
                                double dvclose, double rclose) {
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

// rho = sum(part); beta = rho / rho0 (CG)
__global__ void k_s_cg_rho(KState *st, const double *part, int nb) {
  __shared__ double lds[8];
  if (st->stop) return;
  double s = reduce_parts(part, nb, lds);
  if (threadIdx.x == 0) {
    st->rho = s;
    if (st->iter > 0) st->beta = s / st->rho0;
  }
}

// alpha = rho / (p.q + sign(DPREC, p.q))
__global__ void k_s_cg_alpha(KState *st, const double *part, int nb) {
  __shared__ double lds[8];
  if (st->stop) return;
  double s = reduce_parts(part, nb, lds);
  if (threadIdx.x == 0) {
    double den = s;
    den = den + copysign(DPREC_D, den);
    st->pq = s;
    st->alpha = st->rho / den;
  }
}

// end of a CG / BiCGSTAB iteration: convergence test and exits
__global__ void k_s_iter_end(KState *st, const double *part, const MaxLoc *mpart,
                             int nb, double *hist, int *hloc, int bcgs) {
  __shared__ double lds[8];
  __shared__ MaxLoc mlds[8];
  if (st->stop) return;
  double ss = reduce_parts(part, nb, lds);
  MaxLoc mx = reduce_mparts(mpart, nb, mlds);
  MaxLoc mr = reduce_mparts(mpart + nb, nb, mlds);
  if (threadIdx.x != 0) return;
  st->iter = st->iter + 1;
  double l2norm = sqrt(ss);
  st->deltax = mx.v;
  st->xloc = mx.loc;
  st->rmax = mr.v;
  st->rloc = mr.loc;
  st->l2norm = l2norm;
  int it = st->iter;
  if (it <= st->hist_cap) {
    int cap = st->hist_cap;
    hist[it - 1] = mx.v;
    hist[cap + it - 1] = mr.v;
    hist[2 * cap + it - 1] = st->alpha;
    hist[3 * cap + it - 1] = bcgs ? st->omega : 0.0;
    hloc[it - 1] = mx.loc;
    hloc[cap + it - 1] = mr.loc;
  }
  int o = st->icnvgopt;
  double rcnvg = (o == 2 || o == 3 || o == 4) ? l2norm : mr.v;
  int icnvg = st->icnvg;
  testcnvg(o, &icnvg, it, mx.v, rcnvg, st->l2norm0, st->epfact, st->dvclose, st->rclose);
  if (rcnvg == 0.0) icnvg = 1;
  st->icnvg = icnvg;
  if (icnvg != 0) { st->stop = 1; return; }
  if (is_close_d(st->rho, st->rho0)) { st->stop = 1; return; }
  if (bcgs) {
    if (is_close_d(st->alpha, st->alpha0)) { st->stop = 1; return; }
    if (is_close_d(st->omega, st->omega0)) { st->stop = 1; return; }
  }
  if (st->north > 0 && ((it + 1) % st->north) == 0) st->north_due = 1;
  if (bcgs) {
    if (st->rho * st->omega == 0.0) { st->stop = 1; return; }
  } else {
    if (st->rho == 0.0) { st->stop = 1; return; }
  }
  st->rho0 = st->rho;
  if (bcgs) {
    st->alpha0 = st->alpha;
    st->omega0 = st->omega;
  }
  if (it >= st->itmax) st->stop = 1;
}

__global__ void k_s_clear_north(KState *st) { st->north_due = 0; }

// BiCGSTAB: rho = dhat.d, beta
__global__ void k_s_bcgs_rho(KState *st, const double *part, int nb) {
  __shared__ double lds[8];
  if (st->stop) return;
  double s = reduce_parts(part, nb, lds);
  if (threadIdx.x == 0) {
    st->rho = s;
    if (st->iter > 0) st->beta = (s / st->rho0) * (st->alpha0 / st->omega0);
  }
}

__global__ void k_s_bcgs_alpha(KState *st, const double *part, int nb) {
  __shared__ double lds[8];
  if (st->stop) return;
  double s = reduce_parts(part, nb, lds);
  if (threadIdx.x == 0) {
    double den = s + copysign(DPREC_D, s);
    st->alpha = st->rho / den;
  }
}

__global__ void k_s_bcgs_omega(KState *st, const double *part, int nb) {
  __shared__ double lds[8];
  if (st->stop) return;
  double tq = reduce_parts(part, nb, lds);
  double tt = reduce_parts(part + nb, nb, lds);
  if (threadIdx.x == 0) {
    double den = tt + copysign(DPREC_D, tt);
    st->tq = tq;
    st->tt = tt;
    st->omega = tq / den;
  }
}

// initial residual norm: l2norm0 = scale * sqrt(sum((d/scale)^2)), scale =
// max|d| (dnrm2 scaling, deterministic order)
__global__ void k_s_l2norm0(KState *st, const double *part, const MaxLoc *mpart, int nb,
                            double scale) {
  __shared__ double lds[8];
  double ss = reduce_parts(part, nb, lds);
  if (threadIdx.x == 0) {
    st->l2norm0 = scale * sqrt(ss);
    if (st->l2norm0 == 0.0) {
      st->icnvg = 1;
      st->stop = 1;
    }
  }
}

__global__ void k_s_maxloc(const MaxLoc *mpart, int nb, MaxLoc *out) {
  __shared__ MaxLoc mlds[8];
  MaxLoc m = reduce_mparts(mpart, nb, mlds);
  if (threadIdx.x == 0) *out = m;
}

// ---------------------------------------------------------------------------
// host orchestration
// ---------------------------------------------------------------------------

static inline int grid_for(int n) {
  int g = (n + MF6_BLOCK - 1) / MF6_BLOCK;
  if (g > MF6_MAXBLK) g = MF6_MAXBLK;
  if (g < 1) g = 1;
  return g;
}

template <typename T>
static int dalloc(T **p, size_t n) {
  if (*p) return MF6HIP_OK;
  hipError_t e = hipMalloc((void **)p, n * sizeof(T) + 256);
  if (e != hipSuccess) {
    mf6hip_set_error(std::string("hipMalloc failed: ") + hipGetErrorString(e));
    return MF6HIP_EHIP;
  }
  return MF6HIP_OK;
}

#define CK(x)                              \
  do {                                     \
    int rc_ = (x);                         \
    if (rc_ != MF6HIP_OK) return rc_;      \
  } while (0)

// ILU0 / MILU0 factor of the FP64 planes into h->invpiv
static int ilu_factor(Mf6Handle *h) {
  const Layout &L = h->L;
  hipStream_t s = h->stream;
  int big = 0x7fffffff;
  HIPCHK(hipMemcpyAsync(h->d_fail, &big, sizeof(int), hipMemcpyHostToDevice, s));
  int g = (L.HP + MF6_BLOCK - 1) / MF6_BLOCK;
  k_ilu_factor_red<<<g, MF6_BLOCK, 0, s>>>(L, h->A, h->s.relax, h->invpiv_tmp, h->d_fail);
  k_ilu_factor_black<<<g, MF6_BLOCK, 0, s>>>(L, h->A, h->s.relax, h->invpiv_tmp, h->d_fail);
  k_ilu_commit<<<grid_for(L.NP), MF6_BLOCK, 0, s>>>(L, h->invpiv_tmp, h->invpiv, h->d_fail);
  HIPCHK(hipGetLastError());
  if (h->s.precision == MF6HIP_MIXED) {
    k_convert<double, float><<<grid_for(L.NP), MF6_BLOCK, 0, s>>>(L.NP, h->invpiv, h->invpiv32);
  } else if (h->s.precision == MF6HIP_MIXED_FP16) {
    k_to_half<<<grid_for(L.NP), MF6_BLOCK, 0, s>>>(L.NP, h->invpiv, h->invpiv16);
    // scaled FP16 planes for the fused CG's preconditioner: 2^-k with k the
    // smallest integer that brings max |off-diagonal| to 2^14 or below
    HIPCHK(hipMemsetAsync(h->d_amax, 0, sizeof(double), s));
    k_maxabs_bits<<<grid_for(3 * L.NP) < 2048 ? grid_for(3 * L.NP) : 2048, MF6_BLOCK, 0, s>>>(
        3 * L.NP, h->A + 4 * (size_t)L.NP, (unsigned long long *)h->d_amax);
    double amax = 0.0;
    HIPCHK(hipMemcpyAsync(&amax, h->d_amax, sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    int k = 0;
    if (amax > 0.0 && std::isfinite(amax)) {
      k = (int)std::ceil(std::log2(amax / 16384.0));
      if (k < -60) k = -60;
      if (k > 60) k = 60;
    }
    h->pc_scale = std::ldexp(1.0f, -k);
    k_planes_to_half<<<grid_for(3 * L.NP), MF6_BLOCK, 0, s>>>(
        3 * L.NP, h->A + 4 * (size_t)L.NP, std::ldexp(1.0, -k), h->A16 + 4 * (size_t)L.NP);
    k_scale_to_float<<<grid_for(L.NP), MF6_BLOCK, 0, s>>>(L.NP, h->invpiv, std::ldexp(1.0, k),
                                                          h->invpiv32s);
  }
  HIPCHK(hipGetLastError());
  return MF6HIP_OK;
}

template <typename TM, typename TV, typename TP>
static void launch_ilu_apply(Mf6Handle *h, const TM *A, const TP *piv, const TV *r, TV *zf,
                             TV *z, const KState *st) {
  const Layout &L = h->L;
  hipStream_t s = h->stream;
  int g = (L.HP + MF6_BLOCK - 1) / MF6_BLOCK;
  k_ilu_fwd_red<TM, TV, TP><<<g, MF6_BLOCK, 0, s>>>(L, A, piv, r, zf, st);
  if (L.nlay <= 4)
    k_ilu_black<TM, TV, TP, 4><<<g, MF6_BLOCK, 0, s>>>(L, A, piv, r, zf, z, st);
  else if (L.nlay <= 8)
    k_ilu_black<TM, TV, TP, 8><<<g, MF6_BLOCK, 0, s>>>(L, A, piv, r, zf, z, st);
  else if (L.nlay <= 16)
    k_ilu_black<TM, TV, TP, 16><<<g, MF6_BLOCK, 0, s>>>(L, A, piv, r, zf, z, st);
  else
    k_ilu_black<TM, TV, TP, 32><<<g, MF6_BLOCK, 0, s>>>(L, A, piv, r, zf, z, st);
  k_ilu_bwd_red<TM, TV, TP><<<g, MF6_BLOCK, 0, s>>>(L, A, piv, zf, z, st);
}

template <typename TM, typename TV>
static void launch_spmv(Mf6Handle *h, const TM *A, const TV *x, TV *y) {
  k_spmv<TM, TV><<<grid_for(h->L.NP), MF6_BLOCK, 0, h->stream>>>(h->L, A, x, y);
}

// One system of an ensemble into device planes: CSR (reference layout,
// host) -> 7 planes, rhs / start heads -> positions, ILU0 / MILU0 factor.
// The plane and vector ranges must be zeroed by the caller (padding).
int mf6hip_stage_system(const Layout &L, const int *ia1, int nja, const double *amat,
                        const double *rhs, const double *x0, double relax, double *A64,
                        double *b64, double *x64, double *piv64, hipStream_t s) {
  int *dia = nullptr, *dfail = nullptr;
  double *dcsr = nullptr, *dvec = nullptr, *dtmp = nullptr;
  std::vector<int> ia0(L.N + 1);
  for (int n = 0; n <= L.N; n++) ia0[n] = ia1[n] - 1;
  int rc = MF6HIP_OK;
  if (hipMalloc((void **)&dia, sizeof(int) * (L.N + 1)) != hipSuccess ||
      hipMalloc((void **)&dfail, sizeof(int)) != hipSuccess ||
      hipMalloc((void **)&dcsr, sizeof(double) * nja) != hipSuccess ||
      hipMalloc((void **)&dvec, sizeof(double) * L.N) != hipSuccess ||
      hipMalloc((void **)&dtmp, sizeof(double) * L.NP) != hipSuccess) {
    mf6hip_set_error("hipMalloc failed (ensemble staging)");
    rc = MF6HIP_EHIP;
  }
  const int big = 0x7fffffff;
  if (!rc && (hipMemcpyAsync(dia, ia0.data(), sizeof(int) * (L.N + 1), hipMemcpyHostToDevice, s) != hipSuccess ||
              hipMemcpyAsync(dcsr, amat, sizeof(double) * nja, hipMemcpyHostToDevice, s) != hipSuccess ||
              hipMemcpyAsync(dfail, &big, sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess ||
              hipMemsetAsync(dtmp, 0, sizeof(double) * L.NP, s) != hipSuccess))
    rc = MF6HIP_EHIP;
  if (!rc) {
    k_csr_to_planes<<<grid_for(L.N), MF6_BLOCK, 0, s>>>(L, dia, dcsr, A64);
    if (hipMemcpyAsync(dvec, rhs, sizeof(double) * L.N, hipMemcpyHostToDevice, s) != hipSuccess) rc = MF6HIP_EHIP;
  }
  if (!rc) {
    k_nat_to_pos<double, double><<<grid_for(L.N), MF6_BLOCK, 0, s>>>(L, dvec, b64);
    // the vector staging buffer is reused: wait for the rhs conversion
    if (hipStreamSynchronize(s) != hipSuccess ||
        hipMemcpyAsync(dvec, x0, sizeof(double) * L.N, hipMemcpyHostToDevice, s) != hipSuccess)
      rc = MF6HIP_EHIP;
  }
  if (!rc) {
    k_nat_to_pos<double, double><<<grid_for(L.N), MF6_BLOCK, 0, s>>>(L, dvec, x64);
    int g = (L.HP + MF6_BLOCK - 1) / MF6_BLOCK;
    k_ilu_factor_red<<<g, MF6_BLOCK, 0, s>>>(L, A64, relax, dtmp, dfail);
    k_ilu_factor_black<<<g, MF6_BLOCK, 0, s>>>(L, A64, relax, dtmp, dfail);
    k_ilu_commit<<<grid_for(L.NP), MF6_BLOCK, 0, s>>>(L, dtmp, piv64, dfail);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) rc = MF6HIP_EHIP;
  }
  if (rc == MF6HIP_EHIP) mf6hip_set_error("ensemble staging failed");
  if (dia) hipFree(dia);
  if (dfail) hipFree(dfail);
  if (dcsr) hipFree(dcsr);
  if (dvec) hipFree(dvec);
  if (dtmp) hipFree(dtmp);
  return rc;
}

int mf6hip_ilu_apply64(Mf6Handle *h, const double *r, double *z, const KState *st) {
  launch_ilu_apply<double, double, double>(h, h->A, h->invpiv, r, h->zf, z, st);
  HIPCHK(hipGetLastError());
  return MF6HIP_OK;
}

int mf6hip_spmv64(Mf6Handle *h, const double *x, double *y) {
  launch_spmv<double, double>(h, h->A, x, y);
  HIPCHK(hipGetLastError());
  return MF6HIP_OK;
}

static void timed_spmv_begin(Mf6Handle *h, int &slot) {
  slot = -1;
  if (!h->timing) return;
  int call = h->spmv_calls++;
  if (call % h->spmv_sample == 0 && 2 * (h->last_spmv_n + 1) <= (int)h->spmv_ev.size()) {
    slot = h->last_spmv_n++;
    hipEventRecord(h->spmv_ev[2 * slot], h->stream);
  }
}

static void timed_spmv_end(Mf6Handle *h, int slot) {
  if (slot >= 0) hipEventRecord(h->spmv_ev[2 * slot + 1], h->stream);
}

// initialise the solve state; d = b - A x must already be in h->d
static int init_state(Mf6Handle *h, int kstp) {
  KState *hs = h->h_state;
  memset(hs, 0, sizeof(KState));
  const mf6hip_ims_settings &s = h->s;
  hs->icnvgopt = s.icnvgopt;
  hs->north = s.north;
  hs->itmax = s.iter1;
  hs->dvclose = s.dvclose;
  hs->rclose = s.rclose;
  // ims_base_epfact (ImsLinearBase.f90:1316-1333), single-precision literals
  if (s.icnvgopt == 2) hs->epfact = kstp == 1 ? (double)0.01f : (double)0.10f;
  else if (s.icnvgopt == 4) hs->epfact = 1.0e-4;
  else hs->epfact = 1.0;
  hs->hist_cap = h->hist_cap;
  HIPCHK(hipMemcpyAsync(h->d_state, hs, sizeof(KState), hipMemcpyHostToDevice, h->stream));
  return MF6HIP_OK;
}

// L2NORM0 of h->d (two passes: max|d|, then scaled sum of squares)
static int compute_l2norm0(Mf6Handle *h) {
  const Layout &L = h->L;
  hipStream_t s = h->stream;
  int nb = h->nblk;
  k_maxloc_ss<<<nb, MF6_BLOCK, 0, s>>>(L, h->d, h->d_part, h->d_mpart, 0.0);
  MaxLoc *mres = (MaxLoc *)(h->d_part + 7 * MF6_MAXBLK);
  k_s_maxloc<<<1, MF6_BLOCK, 0, s>>>(h->d_mpart, nb, mres);
  MaxLoc hm;
  HIPCHK(hipMemcpyAsync(&hm, mres, sizeof(MaxLoc), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  double scale = fabs(hm.v);
  k_maxloc_ss<<<nb, MF6_BLOCK, 0, s>>>(L, h->d, h->d_part, h->d_mpart,
                                        scale > 0.0 ? 1.0 / scale : 0.0);
  k_s_l2norm0<<<1, MF6_BLOCK, 0, s>>>(h->d_state, h->d_part, h->d_mpart, nb, scale);
  HIPCHK(hipGetLastError());
  return MF6HIP_OK;
}

static int read_state(Mf6Handle *h) {
  HIPCHK(hipMemcpyAsync(h->h_state, h->d_state, sizeof(KState), hipMemcpyDeviceToHost,
                        h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return MF6HIP_OK;
}

// FP64 preconditioned CG (ims_base_cg) on the device system
static int run_cg(Mf6Handle *h) {
  const Layout &L = h->L;
  hipStream_t s = h->stream;
  int nb = h->nblk;
  int g = grid_for(L.N);
  KState *st = h->d_state;
  double *pin = h->p, *pout = h->p2;
  int batch = 4;
  for (int launched = 0; launched < h->s.iter1;) {
    int nb_it = batch;
    if (launched + nb_it > h->s.iter1) nb_it = h->s.iter1 - launched;
    for (int it = 0; it < nb_it; it++) {
      launch_ilu_apply<double, double, double>(h, h->A, h->invpiv, h->d, h->zf, h->z, st);
      k_dot<double><<<nb, MF6_BLOCK, 0, s>>>(L.NP, h->d, h->z, h->d_part, st);
      k_s_cg_rho<<<1, MF6_BLOCK, 0, s>>>(st, h->d_part, nb);
      int slot;
      timed_spmv_begin(h, slot);
      k_cg_spmv<double, double><<<nb, MF6_BLOCK, 0, s>>>(L, h->A, h->z, pin, pout, h->q,
                                                         h->d_part, st);
      timed_spmv_end(h, slot);
      k_s_cg_alpha<<<1, MF6_BLOCK, 0, s>>>(st, h->d_part, nb);
      k_cg_update<double, double><<<nb, MF6_BLOCK, 0, s>>>(L, h->x, pout, h->d, h->q,
                                                           h->d_part, h->d_mpart, st, nb);
      k_s_iter_end<<<1, MF6_BLOCK, 0, s>>>(st, h->d_part, h->d_mpart, nb, h->d_hist,
                                           h->d_hloc, 0);
      if (h->s.north > 0) {
        k_residual<double, double><<<g, MF6_BLOCK, 0, s>>>(L, h->A, h->x, h->b, h->d, st);
        k_s_clear_north<<<1, 1, 0, s>>>(st);
      }
      double *t = pin; pin = pout; pout = t;
    }
    HIPCHK(hipGetLastError());
    launched += nb_it;
    CK(read_state(h));
    if (h->h_state->stop) break;
    if (batch < 64) batch *= 2;
  }
  return MF6HIP_OK;
}

// FP64 preconditioned BiCGSTAB (ims_base_bcgs, ISCL = 0)
static int run_bcgs(Mf6Handle *h) {
  const Layout &L = h->L;
  hipStream_t s = h->stream;
  int nb = h->nblk;
  int g = grid_for(L.N);
  KState *st = h->d_state;
  HIPCHK(hipMemcpyAsync(h->dhat, h->d, sizeof(double) * L.NP, hipMemcpyDeviceToDevice, s));
  int batch = 4;
  for (int launched = 0; launched < h->s.iter1;) {
    int nb_it = batch;
    if (launched + nb_it > h->s.iter1) nb_it = h->s.iter1 - launched;
    for (int it = 0; it < nb_it; it++) {
      k_dot<double><<<nb, MF6_BLOCK, 0, s>>>(L.NP, h->dhat, h->d, h->d_part, st);
      k_s_bcgs_rho<<<1, MF6_BLOCK, 0, s>>>(st, h->d_part, nb);
      k_bcgs_p<double><<<g, MF6_BLOCK, 0, s>>>(L.NP, h->d, h->p, h->v, st);
      launch_ilu_apply<double, double, double>(h, h->A, h->invpiv, h->p, h->zf, h->phat, st);
      int slot;
      timed_spmv_begin(h, slot);
      k_spmv_dot<double, double><<<nb, MF6_BLOCK, 0, s>>>(L, h->A, h->phat, h->v, h->dhat, h->d_part, st);
      timed_spmv_end(h, slot);
      k_s_bcgs_alpha<<<1, MF6_BLOCK, 0, s>>>(st, h->d_part, nb);
      k_bcgs_q<double><<<g, MF6_BLOCK, 0, s>>>(L.NP, h->d, h->v, h->q, st);
      launch_ilu_apply<double, double, double>(h, h->A, h->invpiv, h->q, h->zf, h->qhat, st);
      k_bcgs_t<double, double><<<nb, MF6_BLOCK, 0, s>>>(L, h->A, h->qhat, h->t, h->q, h->d_part, st, nb);
      k_s_bcgs_omega<<<1, MF6_BLOCK, 0, s>>>(st, h->d_part, nb);
      k_bcgs_update<double, double><<<nb, MF6_BLOCK, 0, s>>>(L, h->x, h->phat, h->qhat, h->d, h->q, h->t,
                                             h->d_part, h->d_mpart, st, nb,
                                             h->s.iscl ? h->d_dscale : nullptr);
      k_s_iter_end<<<1, MF6_BLOCK, 0, s>>>(st, h->d_part, h->d_mpart, nb, h->d_hist,
                                           h->d_hloc, 1);
      if (h->s.north > 0) {
        k_residual<double, double><<<g, MF6_BLOCK, 0, s>>>(L, h->A, h->x, h->b, h->d, st);
        k_s_clear_north<<<1, 1, 0, s>>>(st);
      }
    }
    HIPCHK(hipGetLastError());
    launched += nb_it;
    CK(read_state(h));
    if (h->h_state->stop) break;
    if (batch < 32) batch *= 2;
  }
  return MF6HIP_OK;
}

// ---------------------------------------------------------------------------
// mixed precision: FP32 column-RB ILU0 + FP32 SpMV CG inside FP64 iterative
// refinement.  The FP64 true residual b - A x decides convergence.
// ---------------------------------------------------------------------------

// inner-solve state lives in a second KState
__global__ void k_s_inner_init(KState *st, const double *part, const MaxLoc *mpart, int nb,
                               double rtol, double rfloor) {
  __shared__ MaxLoc mlds[8];
  MaxLoc m = reduce_mparts(mpart, nb, mlds);
  if (threadIdx.x == 0) {
    st->iter = 0;
    st->stop = 0;
    st->icnvg = 0;
    st->rho0 = 0.0;
    st->inner_r0 = fabs(m.v);
    st->rclose = fmax(rtol * fabs(m.v), rfloor);
    st->dvclose = 0.0;
  }
}

// inner CG end: stop when max|r| <= rclose (FP32 inner target)
__global__ void k_s_inner_end(KState *st, const double *part, const MaxLoc *mpart, int nb,
                              int itmax) {
  __shared__ double lds[8];
  __shared__ MaxLoc mlds[8];
  if (st->stop) return;
  double ss = reduce_parts(part, nb, lds);
  MaxLoc mx = reduce_mparts(mpart, nb, mlds);
  MaxLoc mr = reduce_mparts(mpart + nb, nb, mlds);
  if (threadIdx.x != 0) return;
  (void)ss;
  st->iter = st->iter + 1;
  st->deltax = mx.v;
  st->xloc = mx.loc;
  st->rmax = mr.v;
  st->rloc = mr.loc;
  if (fabs(mr.v) <= st->rclose) { st->icnvg = 1; st->stop = 1; return; }
  if (st->rho == 0.0 || is_close_d(st->rho, st->rho0)) { st->stop = 1; return; }
  st->rho0 = st->rho;
  if (st->iter >= itmax) st->stop = 1;
}

// x += e (FP64) with max |e|, then e is reset
__global__ void __launch_bounds__(MF6_BLOCK) k_ir_correct(Layout L, double *__restrict__ x,
                                                          float *__restrict__ e) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < L.NP;
       p += gridDim.x * blockDim.x) {
    x[p] = x[p] + (double)e[p];
    e[p] = 0.0f;
  }
}

template <typename TP>
static int run_mixed_cg(Mf6Handle *h, const TP *piv32) {
  const Layout &L = h->L;
  hipStream_t s = h->stream;
  int nb = h->nblk;
  int g = grid_for(L.N);
  KState *ist = h->d_state + 1;   // inner state
  KState *hs = h->h_state;
  double rtol = h->s.inner_fp32_rtol > 0 ? h->s.inner_fp32_rtol : 1e-4;
  int inner_max = h->s.inner_fp32_max > 0 ? h->s.inner_fp32_max : h->s.iter1;
  int total = 0;
  int icnvg = 0;
  double last_dv = 1e300;
  HIPCHK(hipMemsetAsync(h->e32, 0, sizeof(float) * L.NP, s));
  MaxLoc *mres = (MaxLoc *)(h->d_part + 7 * MF6_MAXBLK);
  while (total < h->s.iter1) {
    // true FP64 residual and its max
    k_residual<double, double><<<g, MF6_BLOCK, 0, s>>>(L, h->A, h->x, h->b, h->d, nullptr);
    k_maxloc_ss<<<nb, MF6_BLOCK, 0, s>>>(L, h->d, h->d_part, h->d_mpart, 0.0);
    k_s_maxloc<<<1, MF6_BLOCK, 0, s>>>(h->d_mpart, nb, mres);
    MaxLoc hm;
    HIPCHK(hipMemcpyAsync(&hm, mres, sizeof(MaxLoc), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (fabs(hm.v) <= h->s.rclose && fabs(last_dv) <= h->s.dvclose) { icnvg = 1; break; }
    if (hm.v == 0.0) { icnvg = 1; break; }
    k_convert<double, float><<<g, MF6_BLOCK, 0, s>>>(L.NP, h->d, h->r32);
    k_s_inner_init<<<1, MF6_BLOCK, 0, s>>>(ist, h->d_part, h->d_mpart, nb, rtol,
                                           0.5 * h->s.rclose);
    float *pin = h->p32, *pout = h->p232;
    int batch = 4, done = 0;
    while (done < inner_max && total + done < h->s.iter1) {
      int nb_it = batch;
      for (int it = 0; it < nb_it; it++) {
        launch_ilu_apply<float, float, TP>(h, h->A32, piv32, h->r32, h->zf32, h->z32, ist);
        k_dot<float><<<nb, MF6_BLOCK, 0, s>>>(L.NP, h->r32, h->z32, h->d_part, ist);
        k_s_cg_rho<<<1, MF6_BLOCK, 0, s>>>(ist, h->d_part, nb);
        int slot;
        timed_spmv_begin(h, slot);
        k_cg_spmv<float, float><<<nb, MF6_BLOCK, 0, s>>>(L, h->A32, h->z32, pin, pout,
                                                         h->q32, h->d_part, ist);
        timed_spmv_end(h, slot);
        k_s_cg_alpha<<<1, MF6_BLOCK, 0, s>>>(ist, h->d_part, nb);
        k_cg_update<float, float><<<nb, MF6_BLOCK, 0, s>>>(L, h->e32, pout, h->r32, h->q32,
                                                           h->d_part, h->d_mpart, ist, nb);
        k_s_inner_end<<<1, MF6_BLOCK, 0, s>>>(ist, h->d_part, h->d_mpart, nb, inner_max);
        float *t = pin; pin = pout; pout = t;
      }
      HIPCHK(hipGetLastError());
      HIPCHK(hipMemcpyAsync(hs + 1, ist, sizeof(KState), hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      done += nb_it;
      if (hs[1].stop) break;
      if (batch < 64) batch *= 2;
    }
    int inner = hs[1].iter;
    total += inner;
    last_dv = hs[1].deltax;   // last inner step max |alpha p|
    k_ir_correct<<<g, MF6_BLOCK, 0, s>>>(L, h->x, h->e32);
    if (inner == 0) break;
  }
  hs->iter = total;
  hs->icnvg = icnvg;
  hs->deltax = last_dv;
  return MF6HIP_OK;
}

// inner BiCGSTAB end: stop when max|r| <= rclose (FP32 inner target), or at
// the breakdowns ims_base_bcgs exits on (ImsLinearBase.f90:475-505): rho,
// alpha or omega unchanged to within DSAME, rho * omega == 0
__global__ void k_s_inner_end_bcgs(KState *st, const MaxLoc *mpart, int nb, int itmax) {
  __shared__ MaxLoc mlds[8];
  if (st->stop) return;
  MaxLoc mx = reduce_mparts(mpart, nb, mlds);
  MaxLoc mr = reduce_mparts(mpart + nb, nb, mlds);
  if (threadIdx.x != 0) return;
  st->iter = st->iter + 1;
  st->deltax = mx.v;
  st->xloc = mx.loc;
  st->rmax = mr.v;
  st->rloc = mr.loc;
  if (fabs(mr.v) <= st->rclose) { st->icnvg = 1; st->stop = 1; return; }
  if (is_close_d(st->rho, st->rho0) || is_close_d(st->alpha, st->alpha0) ||
      is_close_d(st->omega, st->omega0) || st->rho * st->omega == 0.0) {
    st->stop = 1;
    return;
  }
  st->rho0 = st->rho;
  st->alpha0 = st->alpha;
  st->omega0 = st->omega;
  if (st->iter >= itmax) st->stop = 1;
}

// Mixed precision BiCGSTAB (nonsymmetric Newton systems): FP32 ILU0 +
// FP32 SpMV BiCGSTAB (ims_base_bcgs, ImsLinearBase.f90:249-549, ISCL = 0)
// inner solves inside FP64 iterative refinement.  Each refinement step
// restarts the inner BiCGSTAB from the FP64 true residual b - A x (FP32
// copy, also the shadow residual dhat) and reduces it by inner_fp32_rtol
// (default 1e-4) or to half of RCLOSE; the correction is added to the FP64
// heads.  Convergence is the FP64 true residual within RCLOSE with the last
// inner step's max |dx| within DVCLOSE, as for the mixed CG.
template <typename TP>
static int run_mixed_bcgs(Mf6Handle *h, const TP *piv32) {
  const Layout &L = h->L;
  hipStream_t s = h->stream;
  int nb = h->nblk;
  int g = grid_for(L.N);
  KState *ist = h->d_state + 1;
  KState *hs = h->h_state;
  double rtol = h->s.inner_fp32_rtol > 0 ? h->s.inner_fp32_rtol : 1e-4;
  int inner_max = h->s.inner_fp32_max > 0 ? h->s.inner_fp32_max : h->s.iter1;
  int total = 0, icnvg = 0;
  double last_dv = 1e300;
  float *r = h->r32, *p = h->p32, *v = h->p232, *q = h->q32, *phat = h->z32,
        *qhat = h->qhat32, *t = h->t32, *dhat = h->dhat32, *e = h->e32;
  HIPCHK(hipMemsetAsync(e, 0, sizeof(float) * L.NP, s));
  HIPCHK(hipMemsetAsync(v, 0, sizeof(float) * L.NP, s));
  MaxLoc *mres = (MaxLoc *)(h->d_part + 7 * MF6_MAXBLK);
  while (total < h->s.iter1) {
    k_residual<double, double><<<g, MF6_BLOCK, 0, s>>>(L, h->A, h->x, h->b, h->d, nullptr);
    k_maxloc_ss<<<nb, MF6_BLOCK, 0, s>>>(L, h->d, h->d_part, h->d_mpart, 0.0);
    k_s_maxloc<<<1, MF6_BLOCK, 0, s>>>(h->d_mpart, nb, mres);
    MaxLoc hm;
    HIPCHK(hipMemcpyAsync(&hm, mres, sizeof(MaxLoc), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (fabs(hm.v) <= h->s.rclose && fabs(last_dv) <= h->s.dvclose) { icnvg = 1; break; }
    if (hm.v == 0.0) { icnvg = 1; break; }
    k_convert<double, float><<<g, MF6_BLOCK, 0, s>>>(L.NP, h->d, r);
    HIPCHK(hipMemcpyAsync(dhat, r, sizeof(float) * L.NP, hipMemcpyDeviceToDevice, s));
    k_s_inner_init<<<1, MF6_BLOCK, 0, s>>>(ist, h->d_part, h->d_mpart, nb, rtol,
                                           0.5 * h->s.rclose);
    int batch = 4, done = 0;
    while (done < inner_max && total + done < h->s.iter1) {
      for (int it = 0; it < batch; it++) {
        k_dot<float><<<nb, MF6_BLOCK, 0, s>>>(L.NP, dhat, r, h->d_part, ist);
        k_s_bcgs_rho<<<1, MF6_BLOCK, 0, s>>>(ist, h->d_part, nb);
        k_bcgs_p<float><<<g, MF6_BLOCK, 0, s>>>(L.NP, r, p, v, ist);
        launch_ilu_apply<float, float, TP>(h, h->A32, piv32, p, h->zf32, phat, ist);
        int slot;
        timed_spmv_begin(h, slot);
        k_spmv_dot<float, float><<<nb, MF6_BLOCK, 0, s>>>(L, h->A32, phat, v, dhat, h->d_part,
                                                          ist);
        timed_spmv_end(h, slot);
        k_s_bcgs_alpha<<<1, MF6_BLOCK, 0, s>>>(ist, h->d_part, nb);
        k_bcgs_q<float><<<g, MF6_BLOCK, 0, s>>>(L.NP, r, v, q, ist);
        launch_ilu_apply<float, float, TP>(h, h->A32, piv32, q, h->zf32, qhat, ist);
        k_bcgs_t<float, float><<<nb, MF6_BLOCK, 0, s>>>(L, h->A32, qhat, t, q, h->d_part, ist,
                                                        nb);
        k_s_bcgs_omega<<<1, MF6_BLOCK, 0, s>>>(ist, h->d_part, nb);
        k_bcgs_update<float, float><<<nb, MF6_BLOCK, 0, s>>>(L, e, phat, qhat, r, q, t,
                                                             h->d_part, h->d_mpart, ist, nb, nullptr);
        k_s_inner_end_bcgs<<<1, MF6_BLOCK, 0, s>>>(ist, h->d_mpart, nb, inner_max);
      }
      HIPCHK(hipGetLastError());
      HIPCHK(hipMemcpyAsync(hs + 1, ist, sizeof(KState), hipMemcpyDeviceToHost, s));
      HIPCHK(hipStreamSynchronize(s));
      done += batch;
      if (hs[1].stop) break;
      if (batch < 32) batch *= 2;
    }
    int inner = hs[1].iter;
    total += inner;
    last_dv = hs[1].deltax;
    k_ir_correct<<<g, MF6_BLOCK, 0, s>>>(L, h->x, e);
    if (inner == 0) break;
  }
  hs->iter = total;
  hs->icnvg = icnvg;
  hs->deltax = last_dv;
  return MF6HIP_OK;
}

// Solve the device-resident system (h->A, h->b, h->x): factor, residual,
// Krylov.  Equivalent of imslinear_ap without scaling / reordering.
// ---------------------------------------------------------------------------
// IMS SCALING_METHOD (ImsLinearBase.f90:619-754 ims_base_scale) on the
// planes.  Slot s of row n (0 diagonal, 1 up, 2 back, 3 left, 4 right,
// 5 front, 6 down; CSR order) holds A(n, m) for the neighbour m in that
// direction; the same column entry seen from row m is m's slot 7 - s.
// ---------------------------------------------------------------------------

__device__ inline bool scl_has(const Layout &L, int k, int i, int j, int s) {
  switch (s) {
    case 0: return true;
    case 1: return k > 0;
    case 2: return i > 0;
    case 3: return j > 0;
    case 4: return j < L.ncol - 1;
    case 5: return i < L.nrow - 1;
    default: return k < L.nlay - 1;
  }
}

__device__ inline int scl_nb(const Layout &L, int p, int c, int k, int l, int s) {
  const int o = (1 - c) * L.H + k * L.HP;
  switch (s) {
    case 0: return p;
    case 1: return p - L.HP;
    case 2: return o + hn(l, -L.ncol);
    case 3: return o + hn(l, -1);
    case 4: return o + hn(l, 1);
    case 5: return o + hn(l, L.ncol);
    default: return p + L.HP;
  }
}

// iscl 1: ds1 = ds2 = 1 / sqrt|a_nn|.  iscl 2: ds1 = 1 / (row L2 norm), 1
// for a zero row, and the row scaled by it in place.
__global__ void k_scale_rows(Layout L, int iscl, double *__restrict__ A,
                             double *__restrict__ ds1, double *__restrict__ ds2) {
  const int N = L.NP;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < N; p += gridDim.x * blockDim.x) {
    int c, k, i, j, l;
    if (!pos_to_cell(L, p, c, k, i, j, l)) continue;
    if (iscl == 1) {
      double c1 = 1.0 / sqrt(fabs(A[p]));
      ds1[p] = c1;
      ds2[p] = c1;
    } else {
      double c1 = 0.0;
      for (int s = 0; s < MF6_NSLOT; s++)
        if (scl_has(L, k, i, j, s)) c1 = c1 + A[s * N + p] * A[s * N + p];
      c1 = sqrt(c1);
      c1 = c1 == 0.0 ? 1.0 : 1.0 / c1;
      ds1[p] = c1;
      for (int s = 0; s < MF6_NSLOT; s++)
        if (scl_has(L, k, i, j, s)) A[s * N + p] = c1 * A[s * N + p];
    }
  }
}

// iscl 2: ds2 = 1 / (column L2 norm of the row-scaled A), 1 for a zero
// column; the squares are summed in the reference's order, row by row
// (rows n - NRC, n - ncol, n - 1, n, n + 1, n + ncol, n + NRC).
__global__ void k_scale_cols(Layout L, const double *__restrict__ A, double *__restrict__ ds2) {
  const int N = L.NP;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < N; p += gridDim.x * blockDim.x) {
    int c, k, i, j, l;
    if (!pos_to_cell(L, p, c, k, i, j, l)) continue;
    double c2 = 0.0;
    for (int t = 0; t < MF6_NSLOT; t++) {
      const int s = t < 3 ? t + 1 : (t == 3 ? 0 : t);
      if (!scl_has(L, k, i, j, s)) continue;
      const double v = s == 0 ? A[p] : A[(7 - s) * N + scl_nb(L, p, c, k, l, s)];
      c2 = c2 + v * v;
    }
    ds2[p] = c2 == 0.0 ? 1.0 : 1.0 / sqrt(c2);
  }
}

// iopt 0: A = (ds1 A) ds2 (iscl 1) or ds2(col) A (iscl 2, rows already
// scaled), x = x / ds2, b = b ds1.  iopt 1: A = ((1 / ds1) A) (1 / ds2),
// x = x ds2, b = b / ds1.
__global__ void k_scale_apply(Layout L, int iopt, int iscl, double *__restrict__ A,
                              double *__restrict__ x, double *__restrict__ b,
                              const double *__restrict__ ds1, const double *__restrict__ ds2) {
  const int N = L.NP;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < N; p += gridDim.x * blockDim.x) {
    int c, k, i, j, l;
    if (!pos_to_cell(L, p, c, k, i, j, l)) continue;
    const double c1 = ds1[p];
    for (int s = 0; s < MF6_NSLOT; s++) {
      if (!scl_has(L, k, i, j, s)) continue;
      const double c2 = ds2[scl_nb(L, p, c, k, l, s)];
      double a = A[s * N + p];
      if (iopt == 0) a = iscl == 1 ? c1 * a * c2 : c2 * a;
      else a = (1.0 / c1) * a * (1.0 / c2);
      A[s * N + p] = a;
    }
    if (iopt == 0) {
      x[p] = x[p] / ds2[p];
      b[p] = b[p] * c1;
    } else {
      x[p] = x[p] * ds2[p];
      b[p] = b[p] / c1;
    }
  }
}

// Scales (iopt 0) or unscales (iopt 1) A, x and b on the device, as
// imslinear_ap does around the solve (ImsLinear.f90:646-652, :741-747).
static int ims_scale(Mf6Handle *h, int iopt) {
  const Layout &L = h->L;
  hipStream_t s = h->stream;
  const int g = grid_for(L.N);
  if (!h->d_dscale) CK(dalloc(&h->d_dscale, 2 * (size_t)L.NP));
  double *ds1 = h->d_dscale, *ds2 = h->d_dscale + L.NP;
  if (iopt == 0) {
    k_scale_rows<<<g, MF6_BLOCK, 0, s>>>(L, h->s.iscl, h->A, ds1, ds2);
    if (h->s.iscl == 2) k_scale_cols<<<g, MF6_BLOCK, 0, s>>>(L, h->A, ds2);
  }
  k_scale_apply<<<g, MF6_BLOCK, 0, s>>>(L, iopt, h->s.iscl, h->A, h->x, h->b, ds1, ds2);
  HIPCHK(hipGetLastError());
  return MF6HIP_OK;
}

int mf6hip_device_solve(Mf6Handle *h, int kstp, int kiter, int *icnvg, int *in_iter,
                        mf6hip_cnvg_out *out) {
  (void)kiter;
  const Layout &L = h->L;
  hipStream_t s = h->stream;
  int g = grid_for(L.N);
  h->last_spmv_n = 0;
  h->spmv_calls = 0;
  HIPCHK(hipEventRecord(h->ev0, s));
  if (h->s.iscl != 0) CK(ims_scale(h, 0));
  CK(ilu_factor(h));
  int sym = 0;
  if (h->s.ilinmeth == 1 && h->s.north == 0) CK(mf6hip_is_symmetric(h, &sym));
  if (h->s.preconditioner >= 2) {
    if (!sym) {
      mf6hip_set_error("the two-level preconditioner needs a symmetric system (CG)");
      return MF6HIP_EUNSUPPORTED;
    }
    CK(mf6hip_tl_setup(h));
  }
  if (h->s.precision != MF6HIP_FP64) {
    k_convert<double, float><<<grid_for(MF6_NSLOT * L.NP), MF6_BLOCK, 0, s>>>(
        MF6_NSLOT * L.NP, h->A, h->A32);
    CK(init_state(h, kstp));
    if (sym && h->s.icnvgopt == 0 && h->s.kernel_variant == 0) {
      int ic = 0, its = 0;
      CK(mf6hip_run_cg_reliable(h, &ic, &its));
      h->h_state->icnvg = ic;
      h->h_state->iter = its;
    } else if (h->s.ilinmeth == 2) {
      int rc;
      if (h->s.precision == MF6HIP_MIXED) rc = run_mixed_bcgs<float>(h, h->invpiv32);
      else rc = run_mixed_bcgs<unsigned short>(h, h->invpiv16);
      if (rc) return rc;
    } else {
      int rc;
      if (h->s.precision == MF6HIP_MIXED) rc = run_mixed_cg<float>(h, h->invpiv32);
      else rc = run_mixed_cg<unsigned short>(h, h->invpiv16);
      if (rc) return rc;
    }
  } else {
    k_residual<double, double><<<g, MF6_BLOCK, 0, s>>>(L, h->A, h->x, h->b, h->d, nullptr);
    CK(init_state(h, kstp));
    CK(compute_l2norm0(h));
    CK(read_state(h));
    if (!h->h_state->stop) {
      if (h->s.preconditioner == 1) CK(mf6hip_run_fcg(h));
      else if (h->s.ilinmeth == 1 && sym && h->s.kernel_variant == 0) CK(mf6hip_run_cg_fused(h));
      else if (h->s.ilinmeth == 1) CK(run_cg(h));
      else CK(run_bcgs(h));
    }
  }
  if (h->s.iscl != 0) CK(ims_scale(h, 1));
  HIPCHK(hipEventRecord(h->ev1, s));
  HIPCHK(hipEventSynchronize(h->ev1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  h->last_solve_ms = ms;
  if (h->timing && h->last_spmv_n > 0) {
    double tot = 0;
    for (int i = 0; i < h->last_spmv_n; i++) {
      float e = 0;
      hipEventElapsedTime(&e, h->spmv_ev[2 * i], h->spmv_ev[2 * i + 1]);
      tot += e;
    }
    h->last_spmv_ms = tot / h->last_spmv_n;
  }
  KState *hs = h->h_state;
  if (h->s.precision == MF6HIP_FP64) CK(read_state(h));
  int ic = hs->icnvg;
  if (ic < 0) ic = 0;   // ImsLinearBase.f90:238
  *icnvg = ic;
  *in_iter = hs->iter;
  if (out && out->capacity > 0) {
    int n = hs->iter < out->capacity ? hs->iter : out->capacity;
    if (n > h->hist_cap) n = h->hist_cap;
    if (h->s.precision != MF6HIP_FP64) n = 0;
    out->count = n;
    if (n > 0) {
      std::vector<double> hb(4 * (size_t)h->hist_cap);
      std::vector<int> hl(2 * (size_t)h->hist_cap);
      HIPCHK(hipMemcpy(hb.data(), h->d_hist, sizeof(double) * hb.size(), hipMemcpyDeviceToHost));
      HIPCHK(hipMemcpy(hl.data(), h->d_hloc, sizeof(int) * hl.size(), hipMemcpyDeviceToHost));
      for (int i = 0; i < n; i++) {
        if (out->dvmax) out->dvmax[i] = hb[i];
        if (out->rmax) out->rmax[i] = hb[h->hist_cap + i];
        if (out->alpha) out->alpha[i] = hb[2 * h->hist_cap + i];
        if (out->omega) out->omega[i] = hb[3 * h->hist_cap + i];
        if (out->locdv) out->locdv[i] = hl[i] + 1;
        if (out->locr) out->locr[i] = hl[h->hist_cap + i] + 1;
      }
    }
  }
  return MF6HIP_OK;
}

// ---------------------------------------------------------------------------
// C ABI: create / solve / destroy
// ---------------------------------------------------------------------------

static int check_settings(const mf6hip_ims_settings *s) {
  if (s->iscl < 0 || s->iscl > 2) { mf6hip_set_error("SCALING_METHOD must be 0 (none), 1 (symmetric) or 2 (L2 norm)"); return MF6HIP_EINVAL; }
  if (s->iscl != 0 && (s->precision != MF6HIP_FP64 || s->preconditioner != 0)) {
    mf6hip_set_error("SCALING_METHOD runs with FP64 precision and the ILU0/MILU0 preconditioner");
    return MF6HIP_EUNSUPPORTED;
  }
  if (s->iord != 0) { mf6hip_set_error("REORDERING_METHOD is not supported on the device path"); return MF6HIP_EUNSUPPORTED; }
  if (s->level > 0 || s->droptol > 0.0) { mf6hip_set_error("ILUT/MILUT (PRECONDITIONER_LEVELS / DROP_TOLERANCE) is not supported on the device path"); return MF6HIP_EUNSUPPORTED; }
  if (s->ilinmeth != 1 && s->ilinmeth != 2) { mf6hip_set_error("LINEAR_ACCELERATION must be CG (1) or BICGSTAB (2)"); return MF6HIP_EINVAL; }
  if (s->relax < 0.0 || s->relax > 1.0) { mf6hip_set_error("RELAXATION_FACTOR must be in [0, 1]"); return MF6HIP_EINVAL; }
  if (s->rclose == 0.0 && s->icnvgopt != 3) { mf6hip_set_error("INNER_RCLOSE must be > 0"); return MF6HIP_EINVAL; }
  if (s->north < 0) { mf6hip_set_error("NUMBER_ORTHOGONALIZATIONS must be >= 0"); return MF6HIP_EINVAL; }
  if (s->iter1 < 0) { mf6hip_set_error("INNER_MAXIMUM must be >= 0"); return MF6HIP_EINVAL; }
  if (s->precision < 0 || s->precision > 2) { mf6hip_set_error("unknown precision mode"); return MF6HIP_EINVAL; }
  if (s->ordering != MF6HIP_ORDER_RB) { mf6hip_set_error("unknown ordering"); return MF6HIP_EINVAL; }
  if (s->kernel_variant < 0 || s->kernel_variant > 1) { mf6hip_set_error("unknown kernel variant"); return MF6HIP_EINVAL; }
  if (s->preconditioner < 0 || s->preconditioner > 3) { mf6hip_set_error("unknown preconditioner"); return MF6HIP_EINVAL; }
  if (s->preconditioner >= 2 && (s->ilinmeth != 1 || s->kernel_variant != 0 || s->north != 0 ||
                                 (s->precision != MF6HIP_FP64 && s->icnvgopt != 0))) {
    mf6hip_set_error("the two-level preconditioner runs in the fused CG (LINEAR_ACCELERATION CG, no orthogonalization; mixed precision with ICNVGOPT 0)");
    return MF6HIP_EUNSUPPORTED;
  }
  if (s->coarse_size < 0 || s->coarse_size > 6144) { mf6hip_set_error("coarse_size must be in [0, 6144] (the coarse GEMV stages the coarse vector in LDS)"); return MF6HIP_EINVAL; }
  if (s->preconditioner == 1 && (s->ilinmeth != 1 || s->precision != MF6HIP_FP64)) {
    mf6hip_set_error("the learned preconditioner runs in FP64 flexible CG (LINEAR_ACCELERATION CG)");
    return MF6HIP_EUNSUPPORTED;
  }
  if (s->nlay < 1 || s->nrow < 1 || s->ncol < 1) { mf6hip_set_error("DIS shape nlay/nrow/ncol required"); return MF6HIP_EINVAL; }
  if (s->nlay > 32) { mf6hip_set_error("more than 32 layers is not supported on the device path"); return MF6HIP_EUNSUPPORTED; }
  return MF6HIP_OK;
}

// Verify ia/ja are the full DIS 7-point structure, diagonal first.
int mf6hip_check_structure(const Layout &L, int neq, int nja, const int *ia, const int *ja) {
  if (neq != L.N) { mf6hip_set_error("NEQ differs from nlay*nrow*ncol"); return MF6HIP_EINVAL; }
  long long e = 1;
  for (int n = 0; n < neq; n++) {
    int k = n / L.NRC, r = n - k * L.NRC, i = r / L.ncol, j = r - i * L.ncol;
    if (ia[n] != e) { mf6hip_set_error("IA does not match the DIS 7-point structure at node " + std::to_string(n + 1)); return MF6HIP_EINVAL; }
    int cols[7], c = 0;
    cols[c++] = n;
    if (k > 0) cols[c++] = n - L.NRC;
    if (i > 0) cols[c++] = n - L.ncol;
    if (j > 0) cols[c++] = n - 1;
    if (j < L.ncol - 1) cols[c++] = n + 1;
    if (i < L.nrow - 1) cols[c++] = n + L.ncol;
    if (k < L.nlay - 1) cols[c++] = n + L.NRC;
    for (int t = 0; t < c; t++) {
      if (e - 1 + t >= nja || ja[e - 1 + t] != cols[t] + 1) {
        mf6hip_set_error("JA does not match the DIS 7-point structure at node " + std::to_string(n + 1));
        return MF6HIP_EINVAL;
      }
    }
    e += c;
  }
  if (ia[neq] != e || nja != e - 1) { mf6hip_set_error("NJA does not match the DIS 7-point structure"); return MF6HIP_EINVAL; }
  return MF6HIP_OK;
}

// Bitwise symmetry of CSR values on the DIS 7-point structure (checked by
// mf6hip_check_structure): every coupling A(n, m) with m > n equals A(m, n).
// In a DIS row the slots are diagonal, up, back, left, right, front, down,
// each present when that neighbour exists.
int mf6hip_check_symmetric_csr(const Layout &L, const int *ia, const double *amat) {
  for (int n = 0; n < L.N; n++) {
    const int k = n / L.NRC, r = n - k * L.NRC, i = r / L.ncol, j = r - i * L.ncol;
    // entry index (0-based) of each present slot of row n
    int e = ia[n] - 1;
    int pos[MF6_NSLOT];
    pos[S_DIAG] = e++;
    pos[S_UP] = k > 0 ? e++ : -1;
    pos[S_BACK] = i > 0 ? e++ : -1;
    pos[S_LEFT] = j > 0 ? e++ : -1;
    pos[S_RIGHT] = j < L.ncol - 1 ? e++ : -1;
    pos[S_FRONT] = i < L.nrow - 1 ? e++ : -1;
    pos[S_DOWN] = k < L.nlay - 1 ? e++ : -1;
    // the lower slot of the neighbour that mirrors each upper slot: the
    // right neighbour's left, the front neighbour's back, the lower cell's up
    // (the neighbour's lower slot is its first entries after the diagonal)
    const int m[3] = {n + 1, n + L.ncol, n + L.NRC};
    const int us[3] = {S_RIGHT, S_FRONT, S_DOWN};
    for (int t = 0; t < 3; t++) {
      if (pos[us[t]] < 0) continue;
      const int mm = m[t];
      const int km = mm / L.NRC, rm = mm - km * L.NRC, im = rm / L.ncol;
      // slots of row mm before its left / back / up slot
      int em = ia[mm] - 1 + 1;   // after the diagonal
      if (us[t] == S_RIGHT) em += (km > 0) + (im > 0);   // up, back come first
      else if (us[t] == S_FRONT) em += (km > 0);         // up comes first
      if (!(amat[pos[us[t]]] == amat[em])) {
        mf6hip_set_error("the system is not symmetric at node " + std::to_string(n + 1));
        return MF6HIP_EUNSUPPORTED;
      }
    }
  }
  return MF6HIP_OK;
}

static int alloc_solver(Mf6Handle *h) {
  size_t N = h->L.NP;
  CK(dalloc(&h->A, MF6_NSLOT * N));
  CK(dalloc(&h->b, N));
  CK(dalloc(&h->x, N));
  CK(dalloc(&h->xtemp, N));
  CK(dalloc(&h->invpiv, N));
  CK(dalloc(&h->invpiv_tmp, N));
  CK(dalloc(&h->d_fail, 4));
  CK(dalloc(&h->d, N));
  CK(dalloc(&h->zf, N));
  CK(dalloc(&h->z, N));
  CK(dalloc(&h->q, N));
  CK(dalloc(&h->d_tmp_nat, N));
  if (h->s.precision == MF6HIP_FP64) {
    CK(dalloc(&h->p, N));
    CK(dalloc(&h->p2, N));
    // BiCGSTAB vectors; flexible CG uses t, v, dhat, phat as well
    if (h->s.ilinmeth == 2 || h->s.preconditioner == 1) {
      CK(dalloc(&h->t, N));
      CK(dalloc(&h->v, N));
      CK(dalloc(&h->dhat, N));
      CK(dalloc(&h->phat, N));
      CK(dalloc(&h->qhat, N));
    }
  } else {
    CK(dalloc(&h->A32, MF6_NSLOT * N));
    if (h->s.precision == MF6HIP_MIXED) CK(dalloc(&h->invpiv32, N));
    else {
      CK(dalloc(&h->invpiv16, N));
      CK(dalloc(&h->A16, MF6_NSLOT * N));
      CK(dalloc(&h->invpiv32s, N));
      CK(dalloc(&h->d_amax, 1));
    }
    CK(dalloc(&h->r32, N));
    CK(dalloc(&h->e32, N));
    if (h->s.ilinmeth == 2) {
      CK(dalloc(&h->dhat32, N));
      CK(dalloc(&h->qhat32, N));
      CK(dalloc(&h->t32, N));
    }
    CK(dalloc(&h->p32, N));
    CK(dalloc(&h->p232, N));
    CK(dalloc(&h->q32, N));
    CK(dalloc(&h->z32, N));
    CK(dalloc(&h->zf32, N));
  }
  CK(dalloc(&h->d_part, 8 * (size_t)MF6_MAXBLK));
  CK(dalloc(&h->d_mpart, 4 * (size_t)MF6_MAXBLK));
  CK(dalloc(&h->d_state, 4));
  h->hist_cap = h->s.iter1 > 0 ? h->s.iter1 : 1;
  CK(dalloc(&h->d_hist, 4 * (size_t)h->hist_cap));
  CK(dalloc(&h->d_hloc, 2 * (size_t)h->hist_cap));
  // every vector starts at zero: the padding slots of the layout must stay 0
  // and APC = 0 before the first factorisation (ImsLinear.f90:317)
  {
    double *dz[] = {h->A, h->b, h->x, h->xtemp, h->invpiv, h->invpiv_tmp, h->d, h->zf, h->z,
                    h->q, h->d_tmp_nat, h->p, h->p2, h->t, h->v, h->dhat, h->phat, h->qhat};
    size_t dn[] = {MF6_NSLOT * N, N, N, N, N, N, N, N, N, N, N, N, N, N, N, N, N, N};
    for (int t = 0; t < (int)(sizeof(dz) / sizeof(dz[0])); t++)
      if (dz[t]) HIPCHK(hipMemset(dz[t], 0, sizeof(double) * dn[t]));
    float *fz[] = {h->A32, h->invpiv32, h->r32, h->e32, h->p32, h->p232, h->q32, h->z32, h->zf32,
                   h->dhat32, h->qhat32, h->t32};
    size_t fn[] = {MF6_NSLOT * N, N, N, N, N, N, N, N, N, N, N, N};
    for (int t = 0; t < (int)(sizeof(fz) / sizeof(fz[0])); t++)
      if (fz[t]) HIPCHK(hipMemset(fz[t], 0, sizeof(float) * fn[t]));
    if (h->invpiv16) HIPCHK(hipMemset(h->invpiv16, 0, sizeof(unsigned short) * N));
  }
  return MF6HIP_OK;
}

extern "C" int mf6hip_create(int neq, int nja, const int *ia, const int *ja,
                             const mf6hip_ims_settings *settings, void **handle) {
  if (!settings || !handle || !ia || !ja) { mf6hip_set_error("null argument"); return MF6HIP_EINVAL; }
  *handle = nullptr;
  CK(check_settings(settings));
  Mf6Handle *h = new Mf6Handle();
  h->s = *settings;
  long long Nll = (long long)settings->nrow * settings->ncol * settings->nlay;
  if (Nll > (1LL << 30)) { delete h; mf6hip_set_error("grid too large for int32 indexing"); return MF6HIP_EUNSUPPORTED; }
  h->L = make_layout(settings->nlay, settings->nrow, settings->ncol);
  Layout &L = h->L;
  int rc = mf6hip_check_structure(L, neq, nja, ia, ja);
  if (rc) { delete h; return rc; }
  h->neq = neq;
  h->nja = nja;
  h->nblk = grid_for(L.NP);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    delete h;
    mf6hip_set_error("no HIP device available");
    return MF6HIP_EHIP;
  }
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    mf6hip_set_error("hipStreamCreate failed");
    return MF6HIP_EHIP;
  }
  rc = alloc_solver(h);
  if (!rc) {
    std::vector<int> ia0(neq + 1);
    for (int n = 0; n <= neq; n++) ia0[n] = ia[n] - 1;
    h->h_ia = ia0;
    rc = dalloc(&h->d_ia, (size_t)neq + 1);
    if (!rc && hipMemcpy(h->d_ia, ia0.data(), sizeof(int) * (neq + 1), hipMemcpyHostToDevice) != hipSuccess) rc = MF6HIP_EHIP;
    if (!rc) rc = dalloc(&h->d_amat_csr, (size_t)nja);
  }
  if (!rc) {
    if (hipHostMalloc((void **)&h->h_state, 4 * sizeof(KState), hipHostMallocDefault) != hipSuccess) rc = MF6HIP_EHIP;
  }
  if (!rc) {
    hipEventCreate(&h->ev0);
    hipEventCreate(&h->ev1);
    h->spmv_ev.resize(2 * 256);
    for (auto &e : h->spmv_ev) hipEventCreate(&e);
  }
  if (rc) {
    mf6hip_destroy(h);
    return rc;
  }
  *handle = h;
  return MF6HIP_OK;
}

extern "C" int mf6hip_solve(void *handle, int kstp, int kiter, const double *amat,
                            const double *rhs, double *x, int *icnvg, int *in_iter,
                            mf6hip_cnvg_out *out) {
  Mf6Handle *h = (Mf6Handle *)handle;
  if (!h || !amat || !rhs || !x || !icnvg || !in_iter) { mf6hip_set_error("null argument"); return MF6HIP_EINVAL; }
  const Layout &L = h->L;
  hipStream_t s = h->stream;
  int g = grid_for(L.N);
  HIPCHK(hipMemcpyAsync(h->d_amat_csr, amat, sizeof(double) * h->nja, hipMemcpyHostToDevice, s));
  k_csr_to_planes<<<g, MF6_BLOCK, 0, s>>>(L, h->d_ia, h->d_amat_csr, h->A);
  HIPCHK(hipMemcpyAsync(h->d_tmp_nat, rhs, sizeof(double) * L.N, hipMemcpyHostToDevice, s));
  k_nat_to_pos<double, double><<<g, MF6_BLOCK, 0, s>>>(L, h->d_tmp_nat, h->b);
  HIPCHK(hipMemcpyAsync(h->zf, x, sizeof(double) * L.N, hipMemcpyHostToDevice, s));
  k_nat_to_pos<double, double><<<g, MF6_BLOCK, 0, s>>>(L, h->zf, h->x);
  HIPCHK(hipGetLastError());
  CK(mf6hip_device_solve(h, kstp, kiter, icnvg, in_iter, out));
  k_pos_to_nat<double, double><<<g, MF6_BLOCK, 0, s>>>(L, h->x, h->d_tmp_nat);
  HIPCHK(hipMemcpyAsync(x, h->d_tmp_nat, sizeof(double) * L.N, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return MF6HIP_OK;
}

extern "C" int mf6hip_destroy(void *handle) {
  Mf6Handle *h = (Mf6Handle *)handle;
  if (!h) return MF6HIP_OK;
  if (h->stream) hipStreamSynchronize(h->stream);
  void *bufs[] = {h->d_ia, h->A, h->b, h->x, h->xtemp, h->invpiv, h->invpiv_tmp, h->d_fail,
                  h->d, h->p, h->p2, h->q, h->z, h->zf, h->t, h->v, h->dhat, h->phat,
                  h->qhat, h->A32, h->invpiv32, h->invpiv16, h->r32, h->e32, h->p32,
                  h->p232, h->q32, h->z32, h->zf32, h->d_part, h->d_mpart, h->d_state,
                  h->d_hist, h->d_dscale, h->d_hloc, h->d_tmp_nat, h->d_amat_csr, h->top, h->bot,
                  h->area, h->k11, h->k22, h->sat0, h->hn0, h->k33, h->ss, h->sy, h->delr, h->delc, h->icelltype,
                  h->iconvert, h->ibound, h->ibound0, h->condsat, h->xold, h->dxold,
                  h->wsave, h->hchold, h->deold, h->ibotnode, h->bnd_off, h->bnd_ent,
                  h->unet_feat, h->fcg_store, h->A16, h->invpiv32s, h->d_amax, h->sat,
                  h->flowja, h->strg, h->bsim, h->chd_flag, h->dhat32, h->qhat32, h->t32,
                  h->d_hfb_face, h->d_hfb_sav, h->d_gnc_np, h->d_gnc_mp, h->d_gnc_s,
                  h->d_gnc_jp, h->d_gnc_a, h->d_gnc_cond, h->d_gnc_off, h->d_gnc_ent,
                  h->d_hfbd_np, h->d_hfbd_mp, h->d_hfbd_s, h->d_hfbd_hy, h->d_hfbd_w,
                  h->d_hfbd_sav, h->d_hfbd_off, h->d_hfbd_ent};
  for (void *b : bufs)
    if (b) hipFree(b);
  if (h->cg_graph) (void)hipGraphExecDestroy(h->cg_graph);
  mf6hip_tl_free(h);
  for (int t = 0; t < 5; t++) {
    DevPackage &pk = h->pk[t];
    void *pb[] = {pk.nodes, pk.v1, pk.v2, pk.hcof, pk.brhs, pk.depth};
    for (void *b : pb)
      if (b) hipFree(b);
  }
  if (h->h_state) hipHostFree(h->h_state);
  if (h->ev0) hipEventDestroy(h->ev0);
  if (h->ev1) hipEventDestroy(h->ev1);
  for (auto &e : h->spmv_ev) hipEventDestroy(e);
  if (h->stream) hipStreamDestroy(h->stream);
  delete h;
  return MF6HIP_OK;
}

extern "C" int mf6hip_last_timing(void *handle, double *solve_ms, double *spmv_avg_ms,
                                  int *spmv_launches) {
  Mf6Handle *h = (Mf6Handle *)handle;
  if (!h) { mf6hip_set_error("null handle"); return MF6HIP_EINVAL; }
  if (solve_ms) *solve_ms = h->last_solve_ms;
  if (spmv_avg_ms) *spmv_avg_ms = h->last_spmv_ms;
  if (spmv_launches) *spmv_launches = h->spmv_calls;
  return MF6HIP_OK;
}

// Enable sampled HIP-event timing of the SpMV launches (every `sample`-th).
extern "C" int mf6hip_set_timing(void *handle, int enable, int sample) {
  Mf6Handle *h = (Mf6Handle *)handle;
  if (!h) { mf6hip_set_error("null handle"); return MF6HIP_EINVAL; }
  h->timing = enable != 0;
  if (sample > 0) h->spmv_sample = sample;
  return MF6HIP_OK;
}

// Debug / parity helpers: run single kernels on host vectors (natural order).
// op 0: y = A x (SpMV of the last system), op 1: y = M^-1 x (ILU apply of the
// last factor), op 2: y = inverse pivots (x ignored), op 3: y = M^-1 x of the
// two-level preconditioner (ILU0 plus the coarse correction; needs
// preconditioner = 2 or 3 and a system loaded by mf6hip_load_factor: the
// additive operator, whose coarse part the deflated form shares); ops 4-6:
// the fused CG's operators (mf6hip_fused_op: FP64 SpMV, FP32 SpMV, deflated
// two-level apply, ILU0 apply in FP64 and in the handle's mixed precision).
extern "C" int mf6hip_apply_op(void *handle, int op, const double *x, double *y) {
  Mf6Handle *h = (Mf6Handle *)handle;
  if (!h || !y || (op != 2 && !x)) { mf6hip_set_error("null argument"); return MF6HIP_EINVAL; }
  const Layout &L = h->L;
  hipStream_t s = h->stream;
  int g = grid_for(L.N);
  double *tx = h->q, *ty = h->z;
  if (op != 2) {
    HIPCHK(hipMemcpyAsync(h->d_tmp_nat, x, sizeof(double) * L.N, hipMemcpyHostToDevice, s));
    k_nat_to_pos<double, double><<<g, MF6_BLOCK, 0, s>>>(L, h->d_tmp_nat, tx);
  }
  if (op == 0) launch_spmv<double, double>(h, h->A, tx, ty);
  else if (op == 1) launch_ilu_apply<double, double, double>(h, h->A, h->invpiv, tx, h->zf, ty, nullptr);
  else if (op == 2) HIPCHK(hipMemcpyAsync(ty, h->invpiv, sizeof(double) * L.NP, hipMemcpyDeviceToDevice, s));
  else if (op == 3) {
    if (h->s.preconditioner < 2 || !h->tl_cmask) {
      mf6hip_set_error("op 3 needs preconditioner = 2 or 3 and a loaded system");
      return MF6HIP_EINVAL;
    }
    launch_ilu_apply<double, double, double>(h, h->A, h->invpiv, tx, h->zf, ty, nullptr);
    CK(mf6hip_tl_apply64(h, tx, ty));
  } else if (op >= 4 && op <= 8) {
    CK(mf6hip_fused_op(h, op, tx, ty));
  } else if (op == 9) {
    if (h->s.preconditioner != 3 || 2 * h->tl.nc > L.N) {
      mf6hip_set_error("op 9 needs preconditioner = 3 and at least twice as many cells as "
                       "aggregates");
      return MF6HIP_EINVAL;
    }
    HIPCHK(hipMemsetAsync(h->d_tmp_nat, 0, sizeof(double) * L.N, s));
    CK(mf6hip_tl_gather_check(h, tx, h->d_tmp_nat));
  } else { mf6hip_set_error("unknown op"); return MF6HIP_EINVAL; }
  if (op != 9) k_pos_to_nat<double, double><<<g, MF6_BLOCK, 0, s>>>(L, ty, h->d_tmp_nat);
  HIPCHK(hipMemcpyAsync(y, h->d_tmp_nat, sizeof(double) * L.N, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  HIPCHK(hipGetLastError());
  return MF6HIP_OK;
}

// Load a system (reference CSR values, natural-order rhs) and factor it
// without solving, for the parity helpers above.
extern "C" int mf6hip_load_factor(void *handle, const double *amat) {
  Mf6Handle *h = (Mf6Handle *)handle;
  if (!h || !amat) { mf6hip_set_error("null argument"); return MF6HIP_EINVAL; }
  const Layout &L = h->L;
  hipStream_t s = h->stream;
  HIPCHK(hipMemcpyAsync(h->d_amat_csr, amat, sizeof(double) * h->nja, hipMemcpyHostToDevice, s));
  k_csr_to_planes<<<grid_for(L.N), MF6_BLOCK, 0, s>>>(L, h->d_ia, h->d_amat_csr, h->A);
  CK(ilu_factor(h));
  if (h->s.preconditioner >= 2) CK(mf6hip_tl_setup(h));
  HIPCHK(hipStreamSynchronize(s));
  return MF6HIP_OK;
}

// Copy the device system back in the reference layout: CSR values (nja) and
// rhs in natural order.
int mf6hip_copy_system(Mf6Handle *h, double *amat_out, double *rhs_out) {
  const Layout &L = h->L;
  hipStream_t s = h->stream;
  int g = grid_for(L.N);
  if (amat_out) {
    k_planes_to_csr<<<g, MF6_BLOCK, 0, s>>>(L, h->d_ia, h->A, h->d_amat_csr);
    HIPCHK(hipMemcpyAsync(amat_out, h->d_amat_csr, sizeof(double) * h->nja,
                          hipMemcpyDeviceToHost, s));
  }
  if (rhs_out) {
    k_pos_to_nat<double, double><<<g, MF6_BLOCK, 0, s>>>(L, h->b, h->d_tmp_nat);
    HIPCHK(hipMemcpyAsync(rhs_out, h->d_tmp_nat, sizeof(double) * L.N,
                          hipMemcpyDeviceToHost, s));
  }
  HIPCHK(hipStreamSynchronize(s));
  HIPCHK(hipGetLastError());
  return MF6HIP_OK;
}

// Copy the system of the last sln_ls (after the row fixes) back, for tests.
extern "C" int mf6hip_get_system(void *handle, double *amat_out, double *rhs_out) {
  Mf6Handle *h = (Mf6Handle *)handle;
  if (!h) { mf6hip_set_error("null handle"); return MF6HIP_EINVAL; }
  return mf6hip_copy_system(h, amat_out, rhs_out);
}
