// mf6hip_noalloc.hip -- an entry point that allocates nothing, for callers
// that cannot call hipMalloc (code under HIP graph capture, tight coupling
// loops): the residual norms of mf6hip_gwf_residual_norms (sln_calc_residual
// + sln_l2norm, NumericalSolution.f90:2966-2982, :2855-2871) of a system the
// caller already holds on the device in the reference's CSR layout.
//
// The per-block partial sums live in `scratch`, a file-scope device array
// of fixed size that each GPU holds until the process ends.  Its device
// address is taken once per device with hipGetSymbolAddress (the array's
// name is a host-side stand-in in host code) and only that address reaches
// the kernels.  Every call on a GPU shares that GPU's array: each call
// overwrites every slot it later reads, so nothing needs clearing, but two
// calls that overlap on one GPU race on it.
#include <mutex>

#include "mf6hip_handle.h"

#define NA_BLOCK 256
#define NA_MAXBLK 2048   // partial slots per quantity (the grid's upper bound)
#define NA_MAXDEV 64
#define NA_MAXROWS (1 << 30)

// [0, B) sum r^2, [B, 2B) sum b^2, [2B, 3B) signed max r, [3B, 4B) its row
static __device__ __align__(256) double scratch[4 * NA_MAXBLK];

namespace {

std::mutex g_na_mu;
double *g_na_ptr[NA_MAXDEV] = {};

// device address of this device's scratch (looked up once per device)
int na_scratch(double **ptr) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= NA_MAXDEV) {
    mf6hip_set_error(std::string("hipGetDevice failed: ") + hipGetErrorString(e));
    return MF6HIP_EHIP;
  }
  std::lock_guard<std::mutex> lock(g_na_mu);
  if (!g_na_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      mf6hip_set_error(std::string("hipGetSymbolAddress(scratch) failed: ") +
                       hipGetErrorString(e));
      return MF6HIP_EHIP;
    }
    g_na_ptr[dev] = (double *)p;
  }
  *ptr = g_na_ptr[dev];
  return MF6HIP_OK;
}

__device__ inline MaxLoc na_pick(MaxLoc a, MaxLoc b) { return maxloc_pick(a, b); }

// r = A x - b (amux order, then + (-1) b as the plane kernels form it) on
// the rows with active >= 1 (all rows when active is NULL); per block: sum
// r^2, sum b^2 and the signed largest r with its (first) row
__global__ void __launch_bounds__(NA_BLOCK) k_na_resid(int neq, const int *__restrict__ ia,
                                                       const int *__restrict__ ja,
                                                       const double *__restrict__ a,
                                                       const double *__restrict__ b,
                                                       const double *__restrict__ x,
                                                       const int *__restrict__ active,
                                                       double *__restrict__ part, int nblk) {
  __shared__ double l0[NA_BLOCK / 64], l1[NA_BLOCK / 64];
  __shared__ MaxLoc lm[NA_BLOCK / 64];
  double ss = 0.0, bb = 0.0;
  MaxLoc m{0.0, -1};
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < neq; n += gridDim.x * blockDim.x) {
    if (active && active[n] < 1) continue;
    double t = 0.0;
    for (int k = ia[n] - 1; k < ia[n + 1] - 1; k++) t = t + a[k] * x[ja[k] - 1];
    const double r = t + (-1.0) * b[n];
    ss = ss + r * r;
    bb = bb + b[n] * b[n];
    m = na_pick(m, MaxLoc{r, n});
  }
  for (int o = 32; o > 0; o >>= 1) {
    ss = ss + __shfl_xor(ss, o);
    bb = bb + __shfl_xor(bb, o);
    MaxLoc q;
    q.v = __shfl_xor(m.v, o);
    q.loc = __shfl_xor(m.loc, o);
    m = na_pick(m, q);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    l0[w] = ss;
    l1[w] = bb;
    lm[w] = m;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double s0 = 0.0, s1 = 0.0;
  MaxLoc mm{0.0, -1};
  for (int v = 0; v < NA_BLOCK / 64; v++) {
    s0 = s0 + l0[v];
    s1 = s1 + l1[v];
    mm = na_pick(mm, lm[v]);
  }
  part[blockIdx.x] = s0;
  part[nblk + blockIdx.x] = s1;
  part[2 * nblk + blockIdx.x] = mm.v;
  part[3 * nblk + blockIdx.x] = (double)mm.loc;
}

// out[0] = ||r||_2, out[1] = ||b||_2, out[2] = the signed largest r,
// out[3] = its 1-based row (0 when no row is active)
__global__ void __launch_bounds__(NA_BLOCK) k_na_final(const double *__restrict__ part, int nblk,
                                                       double *__restrict__ out) {
  __shared__ double l0[NA_BLOCK / 64], l1[NA_BLOCK / 64];
  __shared__ MaxLoc lm[NA_BLOCK / 64];
  double s0 = 0.0, s1 = 0.0;
  MaxLoc m{0.0, -1};
  for (int t = threadIdx.x; t < nblk; t += blockDim.x) {
    s0 = s0 + part[t];
    s1 = s1 + part[nblk + t];
    m = na_pick(m, MaxLoc{part[2 * nblk + t], (int)part[3 * nblk + t]});
  }
  for (int o = 32; o > 0; o >>= 1) {
    s0 = s0 + __shfl_xor(s0, o);
    s1 = s1 + __shfl_xor(s1, o);
    MaxLoc q;
    q.v = __shfl_xor(m.v, o);
    q.loc = __shfl_xor(m.loc, o);
    m = na_pick(m, q);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    l0[w] = s0;
    l1[w] = s1;
    lm[w] = m;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  double a0 = 0.0, a1 = 0.0;
  MaxLoc mm{0.0, -1};
  for (int v = 0; v < NA_BLOCK / 64; v++) {
    a0 = a0 + l0[v];
    a1 = a1 + l1[v];
    mm = na_pick(mm, lm[v]);
  }
  out[0] = sqrt(a0);
  out[1] = sqrt(a1);
  out[2] = mm.v;
  out[3] = (double)(mm.loc + 1);
}

}  // namespace

extern "C" int mf6hip_csr_residual_norms_noalloc(int neq, int nja, const int *ia, const int *ja,
                                                 const double *amat, const double *rhs,
                                                 const double *x, const int *active,
                                                 double *out, void *stream) {
  if (!ia || !ja || !amat || !rhs || !x || !out) {
    mf6hip_set_error("null argument");
    return MF6HIP_EINVAL;
  }
  if (neq < 1 || neq > NA_MAXROWS) {
    mf6hip_set_error("neq must be in [1, 2^30] (mf6hip_csr_residual_norms_noalloc)");
    return MF6HIP_EINVAL;
  }
  if (nja < neq) {
    mf6hip_set_error("nja must be at least neq (diagonal first in every row)");
    return MF6HIP_EINVAL;
  }
  double *part = nullptr;
  int rc = na_scratch(&part);
  if (rc != MF6HIP_OK) return rc;
  hipStream_t s = (hipStream_t)stream;
  long long g = ((long long)neq + NA_BLOCK - 1) / NA_BLOCK;
  const int nb = g > NA_MAXBLK ? NA_MAXBLK : (int)g;
  k_na_resid<<<nb, NA_BLOCK, 0, s>>>(neq, ia, ja, amat, rhs, x, active, part, nb);
  k_na_final<<<1, NA_BLOCK, 0, s>>>(part, nb, out);
  HIPCHK(hipGetLastError());
  return MF6HIP_OK;
}
