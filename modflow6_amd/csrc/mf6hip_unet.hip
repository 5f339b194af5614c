// mf6hip_unet.hip -- the learned preconditioner network on gfx950: a 4-level
// 3D U-Net (channels c0..c3, 1x2x2 pooling, nearest 1x2x2 upsampling,
// concatenated skips, ReLU, 1x1x1 output convolution) with BF16 weights and
// activations and FP32 accumulation on MFMA (v_mfma_f32_16x16x16_bf16),
// hand-written forward and backward kernels, and an Adam optimiser.
//
// The network has no counterpart in the reference.  Tensors are
// channels-last: voxel v = (d * H + h) * W + w of a (D, H, W) grid, then the
// channel; a tensor slice is (pointer, ld = channel stride, off = first
// channel).  Every channel count the MFMA kernels see is a multiple of 16
// (the 3 input features are padded to 16 zero channels).
//
// 3x3x3 convolution as an implicit GEMM: one wave computes 16 consecutive
// voxels along W times all output channels; per tap and 16-channel chunk
// one MFMA per 16 output channels.  A-fragment (16 voxels x 16 input
// channels): lane l holds voxel l % 16, channels 4 (l / 16) .. +3 (one 8-byte
// load, channels-last).  B-fragment (16 input x 16 output channels): lane l
// holds output channel l % 16, input channels 4 (l / 16) .. +3, pre-packed
// in that order.  Accumulator: lane l holds output channel l % 16 of voxels
// 4 (l / 16) .. +3.
//
// The backward pass reuses the same kernel for the data gradient (weights
// flipped in the tap index and transposed in the channels), an MFMA
// weight-gradient kernel with split-K partials reduced in a fixed order, and
// elementwise kernels for ReLU masks, pooling, upsampling and the skips.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "mf6hip_handle.h"

namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned short bf16;

#define UCK(x)                        \
  do {                                \
    int rc_ = (x);                    \
    if (rc_ != MF6HIP_OK) return rc_; \
  } while (0)

// round to nearest even FP32 -> BF16 (finite inputs)
__host__ __device__ inline bf16 f2bf(float f) {
  unsigned u;
  memcpy(&u, &f, 4);
  unsigned r = u + 0x7fffu + ((u >> 16) & 1u);
  return (bf16)(r >> 16);
}
__host__ __device__ inline float bf2f(bf16 b) {
  unsigned u = ((unsigned)b) << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

struct Grid3 {
  int D, H, W;
  __host__ __device__ int V() const { return D * H * W; }
};

// ---------------------------------------------------------------------------
// 3x3x3 convolution (forward and data gradient)
// ---------------------------------------------------------------------------

// y[v, yoff + o] = act(bias[o] + sum_{t, c} x[v + off(t), xoff + c] W[t, c, o])
// act: relu when relu != 0; then y *= (mask[v, moff + o] > 0) when mask != 0.
// wp: packed fragments [27][cin / 16][NCO][64][4].  One wave per 16-voxel
// tile; tiles = D * H * ceil(W / 16).
template <int NCO>
__global__ void __launch_bounds__(256) k_conv3(Grid3 g, const bf16 *__restrict__ x, int ldx,
                                               int xoff, int cin, const s4 *__restrict__ wp,
                                               const float *__restrict__ bias, int relu,
                                               bf16 *__restrict__ y, int ldy, int yoff,
                                               const bf16 *__restrict__ mask, int ldm, int moff,
                                               int ntiles) {
  const int lane = threadIdx.x & 63;
  const int tile = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (tile >= ntiles) return;
  const int ntw = (g.W + 15) >> 4;
  const int w0 = (tile % ntw) << 4;
  const int hh = (tile / ntw) % g.H;
  const int dd = tile / (ntw * g.H);
  const int wl = w0 + (lane & 15);
  const int kq = lane >> 4;
  const int ncc = cin >> 4;
  f4 acc[NCO];
#pragma unroll
  for (int c = 0; c < NCO; c++) acc[c] = f4{0.f, 0.f, 0.f, 0.f};
  for (int t = 0; t < 27; t++) {
    const int zd = dd + t / 9 - 1, zh = hh + (t / 3) % 3 - 1, zw = wl + t % 3 - 1;
    const bool ok = zd >= 0 && zd < g.D && zh >= 0 && zh < g.H && zw >= 0 && zw < g.W;
    const bf16 *src = ok ? x + (size_t)((zd * g.H + zh) * g.W + zw) * ldx + xoff + 4 * kq : x;
    const s4 *wt = wp + (size_t)t * ncc * NCO * 64 + lane;
    for (int cc = 0; cc < ncc; cc++) {
      s4 a = s4{0, 0, 0, 0};
      if (ok) a = *reinterpret_cast<const s4 *>(src + 16 * cc);
#pragma unroll
      for (int c = 0; c < NCO; c++) {
        s4 b = wt[(cc * NCO + c) * 64];
        acc[c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[c], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int c = 0; c < NCO; c++) {
    const int o = c * 16 + (lane & 15);
    const float bo = bias ? bias[o] : 0.f;
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int w = w0 + 4 * kq + e;
      if (w >= g.W) continue;
      const size_t v = (size_t)(dd * g.H + hh) * g.W + w;
      float val = acc[c][e] + bo;
      if (relu && val < 0.f) val = 0.f;
      if (mask && !(bf2f(mask[v * ldm + moff + o]) > 0.f)) val = 0.f;
      y[v * ldy + yoff + o] = f2bf(val);
    }
  }
}

// pack W (fp32 [27][cin_real][cout]) into fragments of a cin-channel input
// (rows cin_real..cin-1 zero); flip != 0 packs the data-gradient weights
// Wt[t][o][c] = W[26 - t][c][o] (input / output swapped; cin_real == cin)
__global__ void k_pack_w(const float *__restrict__ w, int cin_real, int cin, int cout, int flip,
                         s4 *__restrict__ wp) {
  const int ci = flip ? cout : cin;   // fragment input channels
  const int co = flip ? cin : cout;   // fragment output channels
  const int nco = co >> 4, ncc = ci >> 4;
  const int total = 27 * ncc * nco * 64;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
    const int lane = i & 63;
    int r = i >> 6;
    const int c = r % nco;
    r /= nco;
    const int cc = r % ncc;
    const int t = r / ncc;
    s4 v;
    for (int e = 0; e < 4; e++) {
      const int kin = cc * 16 + 4 * (lane >> 4) + e;   // fragment input channel
      const int kout = c * 16 + (lane & 15);           // fragment output channel
      float f;
      if (flip) f = w[((size_t)(26 - t) * cin_real + kout) * cout + kin];
      else f = kin < cin_real ? w[((size_t)t * cin_real + kin) * cout + kout] : 0.f;
      v[e] = (short)f2bf(f);
    }
    wp[i] = v;
  }
}

// ---------------------------------------------------------------------------
// weight gradient: dW[t][c][o] = sum_v x[v + off(t), c] dy[v, o]
// ---------------------------------------------------------------------------

// block = (t, c-tile, o-tile, split); 4 waves; wave j of split s handles the
// 16-voxel chunks s * 4 + j, s * 4 + j + 4 S, ...  Partial tile written to
// part[(split * 27 + t) * cin * cout + c * cout + o] after the 4 waves are
// summed in order.
__global__ void __launch_bounds__(256) k_wgrad(Grid3 g, const bf16 *__restrict__ x, int ldx,
                                               int xoff, int cin, const bf16 *__restrict__ dy,
                                               int ldd, int doff, int cout, int nsplit,
                                               float *__restrict__ part) {
  __shared__ float red[4][256];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int b = blockIdx.x;
  const int split = b % nsplit;
  b /= nsplit;
  const int nco = cout >> 4, ncc = cin >> 4;
  const int ot = b % nco;
  b /= nco;
  const int ct = b % ncc;
  const int t = b / ncc;
  if (t >= 27) return;
  const int ti = t / 9 - 1, tj = (t / 3) % 3 - 1, tk = t % 3 - 1;
  const int ntw = (g.W + 15) >> 4;
  const int nchunk = g.D * g.H * ntw;
  const int kq = lane >> 4;
  const int ci = ct * 16 + (lane & 15);   // A row
  const int oc = ot * 16 + (lane & 15);   // B column
  f4 acc = f4{0.f, 0.f, 0.f, 0.f};
  for (int ch = split * 4 + wv; ch < nchunk; ch += 4 * nsplit) {
    const int w0 = (ch % ntw) << 4;
    const int hh = (ch / ntw) % g.H;
    const int dd = ch / (ntw * g.H);
    s4 a, bb;
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int w = w0 + 4 * kq + e;
      float xv = 0.f, dv = 0.f;
      if (w < g.W) {
        const size_t v = (size_t)(dd * g.H + hh) * g.W + w;
        dv = bf2f(dy[v * ldd + doff + oc]);
        const int zd = dd + ti, zh = hh + tj, zw = w + tk;
        if (zd >= 0 && zd < g.D && zh >= 0 && zh < g.H && zw >= 0 && zw < g.W)
          xv = bf2f(x[(size_t)((zd * g.H + zh) * g.W + zw) * ldx + xoff + ci]);
      }
      a[e] = (short)f2bf(xv);
      bb[e] = (short)f2bf(dv);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bb, acc, 0, 0, 0);
  }
  // lane holds column o = lane % 16, rows c = 4 (lane / 16) + e
#pragma unroll
  for (int e = 0; e < 4; e++) red[wv][(4 * kq + e) * 16 + (lane & 15)] = acc[e];
  __syncthreads();
  const int i = threadIdx.x;   // 256 = 16 x 16 tile entries
  const float s = ((red[0][i] + red[1][i]) + red[2][i]) + red[3][i];
  const int r = i >> 4, c = i & 15;
  part[((size_t)split * 27 + t) * cin * cout + (size_t)(ct * 16 + r) * cout + ot * 16 + c] = s;
}

// grad[i] (+)= sum over splits k of part[k * stride + i] (fixed order)
__global__ void k_wgrad_reduce(const float *__restrict__ part, int nsplit, int n, size_t stride,
                               float *__restrict__ grad, int accumulate) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    float s = 0.f;
    for (int k = 0; k < nsplit; k++) s = s + part[(size_t)k * stride + i];
    grad[i] = accumulate ? grad[i] + s : s;
  }
}

// bias gradient partials: block b sums voxels b, b + G, ... of channel
// threadIdx.x % cout (cout <= 256 divides 256)
__global__ void k_bgrad(int V, const bf16 *__restrict__ dy, int ldd, int doff, int cout,
                        float *__restrict__ part) {
  __shared__ float red[256];
  const int per = 256 / cout;
  const int o = threadIdx.x % cout, j = threadIdx.x / cout;
  float s = 0.f;
  if (j < per)
    for (int v = blockIdx.x * per + j; v < V; v += gridDim.x * per)
      s = s + bf2f(dy[(size_t)v * ldd + doff + o]);
  red[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x < cout) {
    float t = 0.f;
    for (int k = 0; k < per; k++) t = t + red[k * cout + threadIdx.x];
    part[blockIdx.x * cout + threadIdx.x] = t;
  }
}

// ---------------------------------------------------------------------------
// elementwise pieces
// ---------------------------------------------------------------------------

// Pooling and upsampling between a fine grid gf and its coarse grid gc with
// gc.H = ceil(gf.H / 2), gc.W = ceil(gf.W / 2): coarse cell (h, w) covers the
// fine cells (2h + a, 2w + b), a, b in {0, 1}, that exist (1, 2 or 4 cells
// on the last row / column of an odd extent).
__device__ __forceinline__ int pool_count(const Grid3 &gf, int hc, int wc) {
  const int nh = (2 * hc + 1 < gf.H) ? 2 : 1, nw = (2 * wc + 1 < gf.W) ? 2 : 1;
  return nh * nw;
}

// pooled[d, h, w, c] = mean of the fine cells of coarse cell (h, w); c < C
__global__ void k_pool(Grid3 gc, Grid3 gf, const bf16 *__restrict__ x, int ldx, int xoff, int C,
                       bf16 *__restrict__ y, int ldy, int yoff) {
  const int n = gc.V() * C;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int c = i % C, v = i / C;
    const int w = v % gc.W, h = (v / gc.W) % gc.H, d = v / (gc.W * gc.H);
    const bool hr = 2 * h + 1 < gf.H, wr = 2 * w + 1 < gf.W;
    const size_t b0 = ((size_t)(d * gf.H + 2 * h) * gf.W + 2 * w);
    float s = bf2f(x[b0 * ldx + xoff + c]);
    if (wr) s = s + bf2f(x[(b0 + 1) * ldx + xoff + c]);
    if (hr) {
      float t = bf2f(x[(b0 + gf.W) * ldx + xoff + c]);
      if (wr) t = t + bf2f(x[(b0 + gf.W + 1) * ldx + xoff + c]);
      s = s + t;
    }
    y[(size_t)v * ldy + yoff + c] = f2bf(s / (float)((hr ? 2 : 1) * (wr ? 2 : 1)));
  }
}

// dx (fine grid) (+)= dy (coarse) / count, the pooling gradient; then masked
// by (mask > 0) when mask != 0
__global__ void k_pool_bwd(Grid3 gf, Grid3 gc, const bf16 *__restrict__ dy, int ldd, int doff,
                           int C, const bf16 *__restrict__ add, int lda, int aoff,
                           const bf16 *__restrict__ mask, int ldm, int moff,
                           bf16 *__restrict__ dx, int ldx, int xoff) {
  const int n = gf.V() * C;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int c = i % C, v = i / C;
    const int w = v % gf.W, h = (v / gf.W) % gf.H, d = v / (gf.W * gf.H);
    const size_t vc = (size_t)(d * gc.H + h / 2) * gc.W + w / 2;
    float g = bf2f(dy[vc * ldd + doff + c]) / (float)pool_count(gf, h / 2, w / 2);
    if (add) g = bf2f(add[(size_t)v * lda + aoff + c]) + g;
    if (mask && !(bf2f(mask[(size_t)v * ldm + moff + c]) > 0.f)) g = 0.f;
    dx[(size_t)v * ldx + xoff + c] = f2bf(g);
  }
}

// nearest 1x2x2 upsampling of a coarse tensor into a fine slice
__global__ void k_up(Grid3 gf, Grid3 gc, const bf16 *__restrict__ x, int ldx, int xoff, int C,
                     bf16 *__restrict__ y, int ldy, int yoff) {
  const int n = gf.V() * C;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int c = i % C, v = i / C;
    const int w = v % gf.W, h = (v / gf.W) % gf.H, d = v / (gf.W * gf.H);
    const size_t vc = (size_t)(d * gc.H + h / 2) * gc.W + w / 2;
    y[(size_t)v * ldy + yoff + c] = x[vc * ldx + xoff + c];
  }
}

// coarse gradient of the upsampling: sum over the fine cells of the coarse
// cell, masked
__global__ void k_up_bwd(Grid3 gc, Grid3 gf, const bf16 *__restrict__ dy, int ldd, int doff,
                         int C, const bf16 *__restrict__ add, int lda, int aoff,
                         const bf16 *__restrict__ mask, int ldm, int moff,
                         bf16 *__restrict__ dx, int ldx, int xoff) {
  const int n = gc.V() * C;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int c = i % C, v = i / C;
    const int w = v % gc.W, h = (v / gc.W) % gc.H, d = v / (gc.W * gc.H);
    const bool hr = 2 * h + 1 < gf.H, wr = 2 * w + 1 < gf.W;
    const size_t b0 = ((size_t)(d * gf.H + 2 * h) * gf.W + 2 * w);
    float s = bf2f(dy[b0 * ldd + doff + c]);
    if (wr) s = s + bf2f(dy[(b0 + 1) * ldd + doff + c]);
    if (hr) {
      float t = bf2f(dy[(b0 + gf.W) * ldd + doff + c]);
      if (wr) t = t + bf2f(dy[(b0 + gf.W + 1) * ldd + doff + c]);
      s = s + t;
    }
    if (add) s = bf2f(add[(size_t)v * lda + aoff + c]) + s;
    if (mask && !(bf2f(mask[(size_t)v * ldm + moff + c]) > 0.f)) s = 0.f;
    dx[(size_t)v * ldx + xoff + c] = f2bf(s);
  }
}

// 1x1x1 output convolution y[v] = sum_c u[v, c] wo[c] + bo (fp32 out)
__global__ void k_out(int V, const bf16 *__restrict__ u, int ldu, int C,
                      const float *__restrict__ wo, const float *__restrict__ bo,
                      float *__restrict__ y) {
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
    float s = 0.f;
    for (int c = 0; c < C; c++) s = s + bf2f(u[(size_t)v * ldu + c]) * wo[c];
    y[v] = s + bo[0];
  }
}

// loss partials: sum (y - t)^2 and sum t^2 per block
__global__ void k_loss(int V, const float *__restrict__ y, const bf16 *__restrict__ t, int ldt,
                       int toff, float *__restrict__ part) {
  __shared__ float r1[256], r2[256];
  float a = 0.f, b = 0.f;
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
    const float tv = bf2f(t[(size_t)v * ldt + toff]);
    const float d = y[v] - tv;
    a = a + d * d;
    b = b + tv * tv;
  }
  r1[threadIdx.x] = a;
  r2[threadIdx.x] = b;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      r1[threadIdx.x] = r1[threadIdx.x] + r1[threadIdx.x + s];
      r2[threadIdx.x] = r2[threadIdx.x] + r2[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = r1[0];
    part[2 * blockIdx.x + 1] = r2[0];
  }
}

__global__ void k_loss_final(const float *__restrict__ part, int nb, float *__restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  float a = 0.f, b = 0.f;
  for (int i = 0; i < nb; i++) {
    a = a + part[2 * i];
    b = b + part[2 * i + 1];
  }
  out[0] = a / b;   // loss
  out[1] = b;       // sum t^2
}

// output-layer backward: g = 2 (y - t) / sum t^2 ; du[v, c] = g wo[c] masked
// by u > 0 ; partials of d wo[c] = sum_v g u[v, c] and d bo = sum_v g
__global__ void k_out_bwd(int V, const float *__restrict__ y, const bf16 *__restrict__ t,
                          int ldt, int toff, const float *__restrict__ lossv,
                          const bf16 *__restrict__ u, int ldu, int C,
                          const float *__restrict__ wo, bf16 *__restrict__ du,
                          float *__restrict__ part) {
  __shared__ float red[256][17];
  float acc[17];
  for (int c = 0; c <= C; c++) acc[c] = 0.f;
  const float sc = 2.f / lossv[1];
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x) {
    const float g = sc * (y[v] - bf2f(t[(size_t)v * ldt + toff]));
    for (int c = 0; c < C; c++) {
      const float uv = bf2f(u[(size_t)v * ldu + c]);
      acc[c] = acc[c] + g * uv;
      du[(size_t)v * C + c] = f2bf(uv > 0.f ? g * wo[c] : 0.f);
    }
    acc[C] = acc[C] + g;
  }
  for (int c = 0; c <= C; c++) red[threadIdx.x][c] = acc[c];
  __syncthreads();
  if ((int)threadIdx.x <= C) {
    float s = 0.f;
    for (int k = 0; k < (int)blockDim.x; k++) s = s + red[k][threadIdx.x];
    part[blockIdx.x * (C + 1) + threadIdx.x] = s;
  }
}

// sum of per-block partial rows: out[j] (+)= sum_b part[b * n + j]
__global__ void k_colsum(const float *__restrict__ part, int nb, int n, float *__restrict__ out,
                         int accumulate) {
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
    float s = 0.f;
    for (int b = 0; b < nb; b++) s = s + part[(size_t)b * n + j];
    out[j] = accumulate ? out[j] + s : s;
  }
}

// Adam (bias-corrected) on fp32 master parameters; grads scaled by gscale
__global__ void k_adam(int n, float *__restrict__ p, const float *__restrict__ g,
                       float *__restrict__ m, float *__restrict__ v, float lr, float b1,
                       float b2, float eps, float c1, float c2, float gscale) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float gi = g[i] * gscale;
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

// network input: channel 0..nin-1 from fp32 natural-order planes, rest 0
__global__ void k_make_input(int V, int nin, const float *__restrict__ src,
                             bf16 *__restrict__ x, int ldx) {
  const int n = V * ldx;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int c = i % ldx, v = i / ldx;
    x[i] = c < nin ? f2bf(src[(size_t)c * V + v]) : (bf16)0;
  }
}

// a training sample (bf16, [V][4]: 3 inputs + target) into the padded input
__global__ void k_expand_sample(int V, const bf16 *__restrict__ s, bf16 *__restrict__ x,
                                int ldx) {
  const int n = V * ldx;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int c = i % ldx, v = i / ldx;
    x[i] = c < 3 ? s[(size_t)v * 4 + c] : (bf16)0;
  }
}

__global__ void k_f32_to_bf16(int n, const float *__restrict__ a, bf16 *__restrict__ b) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    b[i] = f2bf(a[i]);
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

struct UConv {
  int cin, cout;        // cin as the kernels see it (padded to 16)
  int cin_real;         // parameter rows (the first layer: the real inputs)
  size_t w_off, b_off;  // offsets into the flat parameter vector
  s4 *wp = nullptr;     // forward fragments
  s4 *wt = nullptr;     // data-gradient fragments
};

// levels 0..nlev-1; conv layers in the order e0a e0b e1a e1b ... (two per
// level, encoder), then the decoder d_{nlev-2} ... d0, one per level
constexpr int MAXLEV = 8;
constexpr int MAXCONV = 3 * MAXLEV - 1;

struct UNetHandle {
  int nin = 3;
  int nlev = 4;
  int ch[MAXLEV];
  Grid3 g[MAXLEV];
  UConv conv[MAXCONV];
  size_t wo_off, bo_off, nparam;
  float *p = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr;
  float *part = nullptr;   // split-K and block partials
  size_t part_n = 0;
  // activations (bf16) per level
  bf16 *in0 = nullptr;                 // level-0 input, 16 channels
  bf16 *pin[MAXLEV] = {};              // pooled inputs (levels 1..nlev-1)
  bf16 *ha[MAXLEV] = {};               // first conv output
  bf16 *cat[MAXLEV] = {};              // decoder inputs [up | skip] (levels < nlev-1)
  bf16 *hbot = nullptr;                // bottom level output
  bf16 *u[MAXLEV] = {};                // decoder outputs (levels < nlev-1)
  float *y = nullptr;                  // fp32 output
  float *lossv = nullptr;              // [2]
  // gradients
  bf16 *gu[MAXLEV] = {}, *gcat[MAXLEV] = {}, *ghb[MAXLEV] = {}, *gha[MAXLEV] = {},
       *gpin[MAXLEV] = {};
  bf16 *data = nullptr;                // training samples [nsamples][V][4]
  int nsamples = 0;
  int adam_t = 0;
  hipStream_t stream = nullptr;
  int nconv() const { return 3 * nlev - 1; }
  // decoder conv of level l (< nlev - 1)
  int dconv(int l) const { return 2 * nlev + (nlev - 2 - l); }
  // channels of the upsampled part of cat[l]
  int cup(int l) const { return ch[l + 1]; }
};

namespace {

template <typename T>
int ualloc(T **p, size_t n) {
  if (hipMalloc((void **)p, sizeof(T) * (n ? n : 1)) != hipSuccess) {
    mf6hip_set_error("hipMalloc failed (learned preconditioner)");
    return MF6HIP_EHIP;
  }
  return MF6HIP_OK;
}

inline int nblk(size_t n) {
  size_t b = (n + 255) / 256;
  if (b > 8192) b = 8192;
  return (int)(b < 1 ? 1 : b);
}

int launch_conv(UNetHandle *u, const Grid3 &g, const bf16 *x, int ldx, int xoff, int cin,
                const s4 *wp, int cout, const float *bias, int relu, bf16 *y, int ldy, int yoff,
                const bf16 *mask, int ldm, int moff) {
  const int ntiles = g.D * g.H * ((g.W + 15) / 16);
  const int blocks = (ntiles + 3) / 4;
  hipStream_t s = u->stream;
  switch (cout / 16) {
    case 1: k_conv3<1><<<blocks, 256, 0, s>>>(g, x, ldx, xoff, cin, wp, bias, relu, y, ldy, yoff, mask, ldm, moff, ntiles); break;
    case 2: k_conv3<2><<<blocks, 256, 0, s>>>(g, x, ldx, xoff, cin, wp, bias, relu, y, ldy, yoff, mask, ldm, moff, ntiles); break;
    case 3: k_conv3<3><<<blocks, 256, 0, s>>>(g, x, ldx, xoff, cin, wp, bias, relu, y, ldy, yoff, mask, ldm, moff, ntiles); break;
    case 4: k_conv3<4><<<blocks, 256, 0, s>>>(g, x, ldx, xoff, cin, wp, bias, relu, y, ldy, yoff, mask, ldm, moff, ntiles); break;
    case 5: k_conv3<5><<<blocks, 256, 0, s>>>(g, x, ldx, xoff, cin, wp, bias, relu, y, ldy, yoff, mask, ldm, moff, ntiles); break;
    case 6: k_conv3<6><<<blocks, 256, 0, s>>>(g, x, ldx, xoff, cin, wp, bias, relu, y, ldy, yoff, mask, ldm, moff, ntiles); break;
    case 7: k_conv3<7><<<blocks, 256, 0, s>>>(g, x, ldx, xoff, cin, wp, bias, relu, y, ldy, yoff, mask, ldm, moff, ntiles); break;
    case 8: k_conv3<8><<<blocks, 256, 0, s>>>(g, x, ldx, xoff, cin, wp, bias, relu, y, ldy, yoff, mask, ldm, moff, ntiles); break;
    default:
      mf6hip_set_error("unsupported channel count " + std::to_string(cout));
      return MF6HIP_EUNSUPPORTED;
  }
  if (hipGetLastError() != hipSuccess) { mf6hip_set_error("conv launch failed"); return MF6HIP_EHIP; }
  return MF6HIP_OK;
}

// weight + bias gradients of one conv layer into u->grad (accumulate)
int conv_wgrad(UNetHandle *u, const Grid3 &g, const UConv &L, const bf16 *x, int ldx, int xoff,
               const bf16 *dy, int ldd, int doff, int accumulate) {
  hipStream_t s = u->stream;
  const int tiles = 27 * (L.cin / 16) * (L.cout / 16);
  int nsplit = 2048 / tiles;
  if (nsplit < 1) nsplit = 1;
  const int nchunk = g.D * g.H * ((g.W + 15) / 16);
  if (nsplit * 4 > nchunk) nsplit = (nchunk + 3) / 4;
  if (nsplit < 1) nsplit = 1;
  const size_t n = (size_t)27 * L.cin * L.cout;
  if ((size_t)nsplit * n > u->part_n) { mf6hip_set_error("partial buffer too small"); return MF6HIP_EINVAL; }
  k_wgrad<<<tiles * nsplit, 256, 0, s>>>(g, x, ldx, xoff, L.cin, dy, ldd, doff, L.cout, nsplit,
                                         u->part);
  // rows of the padded first layer beyond the real inputs are not
  // parameters: per tap keep rows 0..cin_real-1 ([27][cin][cout] partials
  // -> [27][cin_real][cout] gradient)
  if (L.cin_real == L.cin) {
    k_wgrad_reduce<<<nblk(n), 256, 0, s>>>(u->part, nsplit, (int)n, n, u->grad + L.w_off,
                                           accumulate);
  } else {
    const size_t nt = (size_t)L.cin_real * L.cout;
    for (int t = 0; t < 27; t++)
      k_wgrad_reduce<<<nblk(nt), 256, 0, s>>>(u->part + (size_t)t * L.cin * L.cout, nsplit,
                                              (int)nt, n, u->grad + L.w_off + t * nt, accumulate);
  }
  const int bb = nblk((size_t)g.V());
  const int nbb = bb > 1024 ? 1024 : bb;
  k_bgrad<<<nbb, 256, 0, s>>>(g.V(), dy, ldd, doff, L.cout, u->part);
  k_colsum<<<1, 256, 0, s>>>(u->part, nbb, L.cout, u->grad + L.b_off, accumulate);
  if (hipGetLastError() != hipSuccess) { mf6hip_set_error("wgrad launch failed"); return MF6HIP_EHIP; }
  return MF6HIP_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// network set-up, forward, backward, training
// ---------------------------------------------------------------------------

static int u_pack(UNetHandle *u) {
  hipStream_t s = u->stream;
  for (int i = 0; i < u->nconv(); i++) {
    UConv &L = u->conv[i];
    const int nf = 27 * (L.cin / 16) * (L.cout / 16) * 64;
    k_pack_w<<<nblk(nf), 256, 0, s>>>(u->p + L.w_off, L.cin_real, L.cin, L.cout, 0, L.wp);
    if (L.wt) k_pack_w<<<nblk(nf), 256, 0, s>>>(u->p + L.w_off, L.cin_real, L.cin, L.cout, 1, L.wt);
  }
  if (hipGetLastError() != hipSuccess) { mf6hip_set_error("pack launch failed"); return MF6HIP_EHIP; }
  return MF6HIP_OK;
}

// forward from the padded level-0 input u->in0 to u->y
static int u_forward(UNetHandle *u) {
  hipStream_t s = u->stream;
  const int *c = u->ch;
  const int nl = u->nlev, bot = nl - 1;
  for (int l = 0; l < nl; l++) {
    const Grid3 &g = u->g[l];
    const bf16 *xin;
    int ldx, cin;
    if (l == 0) { xin = u->in0; ldx = 16; cin = 16; }
    else {
      // pool hb_{l-1} (in cat[l-1] at offset cup(l-1)) into pin[l]
      const int cprev = c[l - 1], cu = u->cup(l - 1);
      k_pool<<<nblk((size_t)g.V() * cprev), 256, 0, s>>>(g, u->g[l - 1], u->cat[l - 1],
                                                          cu + cprev, cu, cprev, u->pin[l],
                                                          cprev, 0);
      xin = u->pin[l]; ldx = cprev; cin = cprev;
    }
    const UConv &A = u->conv[2 * l], &B = u->conv[2 * l + 1];
    UCK(launch_conv(u, g, xin, ldx, 0, cin, A.wp, A.cout, u->p + A.b_off, 1, u->ha[l], c[l], 0,
                    nullptr, 0, 0));
    bf16 *hb = l < bot ? u->cat[l] : u->hbot;
    const int ldh = l < bot ? u->cup(l) + c[l] : c[bot];
    const int offh = l < bot ? u->cup(l) : 0;
    UCK(launch_conv(u, g, u->ha[l], c[l], 0, c[l], B.wp, B.cout, u->p + B.b_off, 1, hb, ldh, offh,
                    nullptr, 0, 0));
  }
  for (int l = bot - 1; l >= 0; l--) {
    const Grid3 &g = u->g[l];
    const int cu = u->cup(l), ldc = cu + c[l];
    const bf16 *src = l == bot - 1 ? u->hbot : u->u[l + 1];
    k_up<<<nblk((size_t)g.V() * cu), 256, 0, s>>>(g, u->g[l + 1], src, cu, 0, cu, u->cat[l],
                                                   ldc, 0);
    const UConv &D = u->conv[u->dconv(l)];
    UCK(launch_conv(u, g, u->cat[l], ldc, 0, ldc, D.wp, D.cout, u->p + D.b_off, 1, u->u[l], c[l],
                    0, nullptr, 0, 0));
  }
  const Grid3 &g0 = u->g[0];
  k_out<<<nblk(g0.V()), 256, 0, s>>>(g0.V(), u->u[0], c[0], c[0], u->p + u->wo_off,
                                     u->p + u->bo_off, u->y);
  if (hipGetLastError() != hipSuccess) { mf6hip_set_error("forward launch failed"); return MF6HIP_EHIP; }
  return MF6HIP_OK;
}

// loss of u->y against the target (bf16 slice) and the parameter gradients
// (accumulated into u->grad when accumulate != 0); u_forward must have run
static int u_backward(UNetHandle *u, const bf16 *tgt, int ldt, int toff, int accumulate) {
  hipStream_t s = u->stream;
  const int *c = u->ch;
  const int nl = u->nlev, bot = nl - 1;
  const Grid3 &g0 = u->g[0];
  const int V0 = g0.V();
  const int nb = nblk(V0) > 1024 ? 1024 : nblk(V0);
  k_loss<<<nb, 256, 0, s>>>(V0, u->y, tgt, ldt, toff, u->part);
  k_loss_final<<<1, 64, 0, s>>>(u->part, nb, u->lossv);
  k_out_bwd<<<nb, 256, 0, s>>>(V0, u->y, tgt, ldt, toff, u->lossv, u->u[0], c[0], c[0],
                               u->p + u->wo_off, u->gu[0], u->part);
  k_colsum<<<1, 256, 0, s>>>(u->part, nb, c[0] + 1, u->grad + u->wo_off, accumulate);
  // decoder, top to bottom
  for (int l = 0; l < bot; l++) {
    const Grid3 &g = u->g[l];
    const int cu = u->cup(l), ldc = cu + c[l];
    const UConv &D = u->conv[u->dconv(l)];
    UCK(conv_wgrad(u, g, D, u->cat[l], ldc, 0, u->gu[l], c[l], 0, accumulate));
    UCK(launch_conv(u, g, u->gu[l], c[l], 0, c[l], D.wt, ldc, nullptr, 0, u->gcat[l], ldc, 0,
                    nullptr, 0, 0));
    const Grid3 &gc = u->g[l + 1];
    // gradient of the coarser decoder output (or of the bottom output)
    // through the upsampling, masked by that output > 0
    const bf16 *msk = l < bot - 1 ? u->u[l + 1] : u->hbot;
    bf16 *dst = l < bot - 1 ? u->gu[l + 1] : u->ghb[bot];
    k_up_bwd<<<nblk((size_t)gc.V() * cu), 256, 0, s>>>(gc, g, u->gcat[l], ldc, 0, cu, nullptr,
                                                        0, 0, msk, cu, 0, dst, cu, 0);
  }
  // encoder, bottom to top
  for (int l = bot; l >= 0; l--) {
    const Grid3 &g = u->g[l];
    if (l < bot) {
      // hb_l gradient: skip part of gcat[l] + pooled gradient from level l+1,
      // masked by hb_l > 0
      const int cu = u->cup(l), ldc = cu + c[l];
      k_pool_bwd<<<nblk((size_t)g.V() * c[l]), 256, 0, s>>>(g, u->g[l + 1], u->gpin[l + 1], c[l],
                                                             0, c[l], u->gcat[l], ldc, cu,
                                                             u->cat[l], ldc, cu, u->ghb[l], c[l],
                                                             0);
    }
    const UConv &A = u->conv[2 * l], &B = u->conv[2 * l + 1];
    UCK(conv_wgrad(u, g, B, u->ha[l], c[l], 0, u->ghb[l], c[l], 0, accumulate));
    UCK(launch_conv(u, g, u->ghb[l], c[l], 0, c[l], B.wt, c[l], nullptr, 0, u->gha[l], c[l], 0,
                    u->ha[l], c[l], 0));
    const bf16 *xin = l == 0 ? u->in0 : u->pin[l];
    const int ldx = l == 0 ? 16 : c[l - 1];
    UCK(conv_wgrad(u, g, A, xin, ldx, 0, u->gha[l], c[l], 0, accumulate));
    if (l > 0)
      UCK(launch_conv(u, g, u->gha[l], c[l], 0, c[l], A.wt, c[l - 1], nullptr, 0, u->gpin[l],
                      c[l - 1], 0, nullptr, 0, 0));
  }
  if (hipGetLastError() != hipSuccess) { mf6hip_set_error("backward launch failed"); return MF6HIP_EHIP; }
  return MF6HIP_OK;
}

static int u_adam(UNetHandle *u, float lr, float gscale) {
  u->adam_t++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - powf(b1, (float)u->adam_t), c2 = 1.f - powf(b2, (float)u->adam_t);
  k_adam<<<nblk(u->nparam), 256, 0, u->stream>>>((int)u->nparam, u->p, u->grad, u->m, u->v, lr,
                                                 b1, b2, 1e-8f, c1, c2, gscale);
  if (hipGetLastError() != hipSuccess) { mf6hip_set_error("adam launch failed"); return MF6HIP_EHIP; }
  return u_pack(u);
}

static void u_free(UNetHandle *u) {
  if (!u) return;
  if (u->stream) hipStreamSynchronize(u->stream);
  void *bufs[] = {u->p, u->grad, u->m, u->v, u->part, u->in0, u->hbot, u->y, u->lossv, u->data};
  for (void *b : bufs)
    if (b) hipFree(b);
  for (int l = 0; l < MAXLEV; l++) {
    void *b2[] = {u->pin[l], u->ha[l], u->ghb[l], u->gha[l], u->gpin[l],
                  u->cat[l], u->u[l], u->gu[l], u->gcat[l]};
    for (void *b : b2)
      if (b) hipFree(b);
  }
  for (int i = 0; i < MAXCONV; i++) {
    if (u->conv[i].wp) hipFree(u->conv[i].wp);
    if (u->conv[i].wt) hipFree(u->conv[i].wt);
  }
  if (u->stream) hipStreamDestroy(u->stream);
  delete u;
}

extern "C" int mf6hip_unet_create(int nlay, int nrow, int ncol, int nin, int nlev,
                                  const int *channels, void **handle) {
  if (!handle || !channels) { mf6hip_set_error("null argument"); return MF6HIP_EINVAL; }
  *handle = nullptr;
  if (nlev < 2 || nlev > MAXLEV) {
    mf6hip_set_error("the network has 2.." + std::to_string(MAXLEV) + " levels");
    return MF6HIP_EUNSUPPORTED;
  }
  if (nlay < 1 || nrow < 2 || ncol < 2) {
    mf6hip_set_error("the learned preconditioner needs nrow and ncol >= 2");
    return MF6HIP_EUNSUPPORTED;
  }
  if (nin < 1 || nin > 16) { mf6hip_set_error("1..16 input channels"); return MF6HIP_EINVAL; }
  for (int l = 0; l < nlev; l++)
    if (channels[l] < 16 || channels[l] > 64 || channels[l] % 16) {
      mf6hip_set_error("channels must be 16, 32, 48 or 64");
      return MF6HIP_EUNSUPPORTED;
    }
  if (channels[0] != 16) { mf6hip_set_error("level-0 channels must be 16"); return MF6HIP_EUNSUPPORTED; }
  if ((long long)nlay * nrow * ncol > (1LL << 28)) { mf6hip_set_error("grid too large"); return MF6HIP_EUNSUPPORTED; }
  UNetHandle *u = new UNetHandle();
  u->nin = nin;
  u->nlev = nlev;
  for (int l = 0; l < nlev; l++) {
    u->ch[l] = channels[l];
    u->g[l] = l == 0 ? Grid3{nlay, nrow, ncol}
                     : Grid3{nlay, (u->g[l - 1].H + 1) / 2, (u->g[l - 1].W + 1) / 2};
  }
  const int *c = u->ch;
  const int bot = nlev - 1;
  size_t off = 0;
  int rc = MF6HIP_OK;
  for (int i = 0; i < u->nconv(); i++) {
    UConv &L = u->conv[i];
    if (i < 2 * nlev) {
      const int l = i / 2;
      L.cin = i == 0 ? 16 : (i % 2 ? c[l] : c[l - 1]);
      L.cin_real = i == 0 ? nin : L.cin;
      L.cout = c[l];
    } else {
      const int l = bot - 1 - (i - 2 * nlev);   // decoder level
      L.cin = u->cup(l) + c[l];
      L.cin_real = L.cin;
      L.cout = c[l];
    }
    L.w_off = off;
    off += (size_t)27 * L.cin_real * L.cout;
    L.b_off = off;
    off += L.cout;
    const size_t nf = (size_t)27 * (L.cin / 16) * (L.cout / 16) * 64;
    if (!rc) rc = ualloc(&L.wp, nf);
    if (!rc && i != 0) rc = ualloc(&L.wt, nf);
  }
  u->wo_off = off;
  off += c[0];
  u->bo_off = off;
  off += 1;
  u->nparam = off;
  if (!rc) rc = ualloc(&u->p, off);
  if (!rc) rc = ualloc(&u->grad, off);
  if (!rc) rc = ualloc(&u->m, off);
  if (!rc) rc = ualloc(&u->v, off);
  // partials: the largest split-K buffer (2048 / tiles splits, at least 1)
  size_t pn = 1 << 20;
  for (int i = 0; i < u->nconv(); i++) {
    const UConv &L = u->conv[i];
    const size_t n = (size_t)27 * L.cin * L.cout;
    const int tiles = 27 * (L.cin / 16) * (L.cout / 16);
    int ns = 2048 / tiles;
    if (ns < 1) ns = 1;
    if ((size_t)ns * n > pn) pn = (size_t)ns * n;
  }
  u->part_n = pn;
  if (!rc) rc = ualloc(&u->part, pn);
  const size_t V0 = u->g[0].V();
  if (!rc) rc = ualloc(&u->in0, V0 * 16);
  if (!rc) rc = ualloc(&u->y, V0);
  if (!rc) rc = ualloc(&u->lossv, 4);
  for (int l = 0; l < nlev && !rc; l++) {
    const size_t V = u->g[l].V();
    if (l > 0) rc = ualloc(&u->pin[l], V * c[l - 1]);
    if (!rc) rc = ualloc(&u->ha[l], V * c[l]);
    if (!rc) rc = ualloc(&u->ghb[l], V * c[l]);
    if (!rc) rc = ualloc(&u->gha[l], V * c[l]);
    if (!rc && l > 0) rc = ualloc(&u->gpin[l], V * c[l - 1]);
    if (!rc && l < bot) {
      rc = ualloc(&u->cat[l], V * (u->cup(l) + c[l]));
      if (!rc) rc = ualloc(&u->u[l], V * c[l]);
      if (!rc) rc = ualloc(&u->gu[l], V * c[l]);
      if (!rc) rc = ualloc(&u->gcat[l], V * (u->cup(l) + c[l]));
    }
    if (!rc && l == bot) rc = ualloc(&u->hbot, V * c[bot]);
  }
  if (!rc && hipStreamCreateWithFlags(&u->stream, hipStreamNonBlocking) != hipSuccess) {
    mf6hip_set_error("hipStreamCreate failed");
    rc = MF6HIP_EHIP;
  }
  if (!rc) {
    hipMemsetAsync(u->p, 0, sizeof(float) * off, u->stream);
    hipMemsetAsync(u->m, 0, sizeof(float) * off, u->stream);
    hipMemsetAsync(u->v, 0, sizeof(float) * off, u->stream);
    hipMemsetAsync(u->grad, 0, sizeof(float) * off, u->stream);
    rc = u_pack(u);
  }
  if (rc) { u_free(u); return rc; }
  *handle = u;
  return MF6HIP_OK;
}

extern "C" int mf6hip_unet_destroy(void *handle) {
  u_free((UNetHandle *)handle);
  return MF6HIP_OK;
}

extern "C" int mf6hip_unet_num_params(void *handle, long long *n) {
  UNetHandle *u = (UNetHandle *)handle;
  if (!u || !n) { mf6hip_set_error("null argument"); return MF6HIP_EINVAL; }
  *n = (long long)u->nparam;
  return MF6HIP_OK;
}

extern "C" int mf6hip_unet_set_params(void *handle, const float *params) {
  UNetHandle *u = (UNetHandle *)handle;
  if (!u || !params) { mf6hip_set_error("null argument"); return MF6HIP_EINVAL; }
  if (hipMemcpyAsync(u->p, params, sizeof(float) * u->nparam, hipMemcpyHostToDevice, u->stream) != hipSuccess) {
    mf6hip_set_error("hipMemcpy failed");
    return MF6HIP_EHIP;
  }
  hipMemsetAsync(u->m, 0, sizeof(float) * u->nparam, u->stream);
  hipMemsetAsync(u->v, 0, sizeof(float) * u->nparam, u->stream);
  u->adam_t = 0;
  UCK(u_pack(u));
  if (hipStreamSynchronize(u->stream) != hipSuccess) { mf6hip_set_error("sync failed"); return MF6HIP_EHIP; }
  return MF6HIP_OK;
}

extern "C" int mf6hip_unet_get_params(void *handle, float *params) {
  UNetHandle *u = (UNetHandle *)handle;
  if (!u || !params) { mf6hip_set_error("null argument"); return MF6HIP_EINVAL; }
  if (hipMemcpyAsync(params, u->p, sizeof(float) * u->nparam, hipMemcpyDeviceToHost, u->stream) != hipSuccess ||
      hipStreamSynchronize(u->stream) != hipSuccess) {
    mf6hip_set_error("hipMemcpy failed");
    return MF6HIP_EHIP;
  }
  return MF6HIP_OK;
}

// x: nin planes of V0 fp32 values (natural order), y: V0 fp32
extern "C" int mf6hip_unet_forward(void *handle, const float *x, float *y) {
  UNetHandle *u = (UNetHandle *)handle;
  if (!u || !x || !y) { mf6hip_set_error("null argument"); return MF6HIP_EINVAL; }
  const int V0 = u->g[0].V();
  float *dx = nullptr;
  UCK(ualloc(&dx, (size_t)V0 * u->nin));
  int rc = MF6HIP_OK;
  if (hipMemcpyAsync(dx, x, sizeof(float) * V0 * u->nin, hipMemcpyHostToDevice, u->stream) != hipSuccess) rc = MF6HIP_EHIP;
  if (!rc) k_make_input<<<nblk((size_t)V0 * 16), 256, 0, u->stream>>>(V0, u->nin, dx, u->in0, 16);
  if (!rc) rc = u_forward(u);
  if (!rc && hipMemcpyAsync(y, u->y, sizeof(float) * V0, hipMemcpyDeviceToHost, u->stream) != hipSuccess) rc = MF6HIP_EHIP;
  if (hipStreamSynchronize(u->stream) != hipSuccess && !rc) rc = MF6HIP_EHIP;
  hipFree(dx);
  if (rc == MF6HIP_EHIP) mf6hip_set_error("unet forward failed");
  return rc;
}

// loss and parameter gradients of one sample: x as in forward, t: V0 fp32
extern "C" int mf6hip_unet_grad(void *handle, const float *x, const float *t, float *loss,
                                float *grad) {
  UNetHandle *u = (UNetHandle *)handle;
  if (!u || !x || !t || !loss || !grad) { mf6hip_set_error("null argument"); return MF6HIP_EINVAL; }
  const int V0 = u->g[0].V();
  float *dx = nullptr;
  bf16 *dt = nullptr;
  UCK(ualloc(&dx, (size_t)V0 * (u->nin + 1)));
  int rc = ualloc(&dt, (size_t)V0);
  if (!rc && hipMemcpyAsync(dx, x, sizeof(float) * V0 * u->nin, hipMemcpyHostToDevice, u->stream) != hipSuccess) rc = MF6HIP_EHIP;
  if (!rc && hipMemcpyAsync(dx + (size_t)V0 * u->nin, t, sizeof(float) * V0, hipMemcpyHostToDevice, u->stream) != hipSuccess) rc = MF6HIP_EHIP;
  if (!rc) {
    k_make_input<<<nblk((size_t)V0 * 16), 256, 0, u->stream>>>(V0, u->nin, dx, u->in0, 16);
    k_f32_to_bf16<<<nblk(V0), 256, 0, u->stream>>>(V0, dx + (size_t)V0 * u->nin, dt);
    rc = u_forward(u);
  }
  if (!rc) rc = u_backward(u, dt, 1, 0, 0);
  float lv[2] = {0.f, 0.f};
  if (!rc && hipMemcpyAsync(lv, u->lossv, sizeof(lv), hipMemcpyDeviceToHost, u->stream) != hipSuccess) rc = MF6HIP_EHIP;
  if (!rc && hipMemcpyAsync(grad, u->grad, sizeof(float) * u->nparam, hipMemcpyDeviceToHost, u->stream) != hipSuccess) rc = MF6HIP_EHIP;
  if (hipStreamSynchronize(u->stream) != hipSuccess && !rc) rc = MF6HIP_EHIP;
  *loss = lv[0];
  hipFree(dx);
  if (dt) hipFree(dt);
  if (rc == MF6HIP_EHIP) mf6hip_set_error("unet grad failed");
  return rc;
}

// training samples: n samples of V0 voxels x (nin inputs + 1 target), fp32
// planes [n][nin + 1][V0]; stored on the device in BF16 ([n][V0][4])
extern "C" int mf6hip_unet_set_data(void *handle, int n, const float *samples) {
  UNetHandle *u = (UNetHandle *)handle;
  if (!u || !samples || n < 1) { mf6hip_set_error("bad argument"); return MF6HIP_EINVAL; }
  if (u->nin != 3) { mf6hip_set_error("training data needs 3 input channels"); return MF6HIP_EUNSUPPORTED; }
  const size_t V0 = u->g[0].V();
  if (u->data) { hipFree(u->data); u->data = nullptr; }
  UCK(ualloc(&u->data, (size_t)n * V0 * 4));
  std::vector<bf16> tmp(V0 * 4);
  for (int i = 0; i < n; i++) {
    const float *sp = samples + (size_t)i * 4 * V0;
    for (size_t v = 0; v < V0; v++)
      for (int c = 0; c < 4; c++) tmp[v * 4 + c] = f2bf(sp[(size_t)c * V0 + v]);
    if (hipMemcpy(u->data + (size_t)i * V0 * 4, tmp.data(), sizeof(bf16) * V0 * 4, hipMemcpyHostToDevice) != hipSuccess) {
      mf6hip_set_error("hipMemcpy failed");
      return MF6HIP_EHIP;
    }
  }
  u->nsamples = n;
  return MF6HIP_OK;
}

// Adam training over the stored samples: nsteps steps of `batch` samples
// taken in the order given by order[nsteps * batch]; loss_out[nsteps]
// receives each step's mean sample loss.
extern "C" int mf6hip_unet_train(void *handle, int nsteps, int batch, const int *order, float lr,
                                 float *loss_out) {
  UNetHandle *u = (UNetHandle *)handle;
  if (!u || !order || !loss_out || nsteps < 1 || batch < 1) { mf6hip_set_error("bad argument"); return MF6HIP_EINVAL; }
  if (!u->data) { mf6hip_set_error("no training data"); return MF6HIP_EINVAL; }
  const size_t V0 = u->g[0].V();
  float *dl = nullptr;
  UCK(ualloc(&dl, (size_t)nsteps * batch));
  int rc = MF6HIP_OK;
  for (int st = 0; st < nsteps && !rc; st++) {
    for (int b = 0; b < batch && !rc; b++) {
      const int idx = order[st * batch + b];
      if (idx < 0 || idx >= u->nsamples) { mf6hip_set_error("sample index out of range"); rc = MF6HIP_EINVAL; break; }
      const bf16 *smp = u->data + (size_t)idx * V0 * 4;
      k_expand_sample<<<nblk(V0 * 16), 256, 0, u->stream>>>((int)V0, smp, u->in0, 16);
      rc = u_forward(u);
      if (!rc) rc = u_backward(u, smp, 4, 3, b > 0);
      if (!rc && hipMemcpyAsync(dl + (size_t)st * batch + b, u->lossv, sizeof(float), hipMemcpyDeviceToDevice, u->stream) != hipSuccess) rc = MF6HIP_EHIP;
    }
    if (!rc) rc = u_adam(u, lr, 1.f / batch);
  }
  std::vector<float> hl((size_t)nsteps * batch);
  if (!rc && hipMemcpyAsync(hl.data(), dl, sizeof(float) * hl.size(), hipMemcpyDeviceToHost, u->stream) != hipSuccess) rc = MF6HIP_EHIP;
  if (hipStreamSynchronize(u->stream) != hipSuccess && !rc) rc = MF6HIP_EHIP;
  hipFree(dl);
  if (!rc)
    for (int st = 0; st < nsteps; st++) {
      float s = 0.f;
      for (int b = 0; b < batch; b++) s += hl[(size_t)st * batch + b];
      loss_out[st] = s / batch;
    }
  if (rc == MF6HIP_EHIP) mf6hip_set_error("unet training failed");
  return rc;
}

// device-side application used by the solver: in0 already filled by the
// caller on the network's stream; result in u->y (fp32, natural order)
int mf6hip_unet_apply_device(void *handle, hipStream_t s) {
  UNetHandle *u = (UNetHandle *)handle;
  hipStream_t keep = u->stream;
  u->stream = s;
  int rc = u_forward(u);
  u->stream = keep;
  return rc;
}

unsigned short *mf6hip_unet_input(void *net) { return ((UNetHandle *)net)->in0; }
float *mf6hip_unet_output(void *net) { return ((UNetHandle *)net)->y; }
int mf6hip_unet_dims(void *net, int *nlay, int *nrow, int *ncol, int *nin) {
  UNetHandle *u = (UNetHandle *)net;
  if (!u) { mf6hip_set_error("null network"); return MF6HIP_EINVAL; }
  *nlay = u->g[0].D;
  *nrow = u->g[0].H;
  *ncol = u->g[0].W;
  *nin = u->nin;
  return MF6HIP_OK;
}
