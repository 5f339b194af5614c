// mf6hip_common.h -- shared definitions of libmf6hip (gfx950).
//
// Data layout in HBM ("column red-black" layout).  A DIS grid of
// nlay x nrow x ncol cells is split by the colour of its vertical cell
// column, c = (i + j) mod 2.  With l = i * ncol + j the cell's index in its
// layer, the position of cell (k, i, j) is
//
//     p = c * H + k * HP + (l >> 1),     HP = ceil(nrow * ncol / 2),
//                                        H  = nlay * HP
//
// (l >> 1 is the index of the column within its colour: for an even ncol
// every pair (2m, 2m+1) of a row holds one cell of each colour, for an odd
// ncol the colour equals l mod 2).  Adjacent columns of one colour are
// adjacent in memory, so thread-per-column and thread-per-cell kernels load
// coalesced; a vertical neighbour is +-HP away in the same colour and the
// four horizontal neighbours sit in the other colour at the same k, at
// (l + delta) >> 1 with delta in {-ncol, -1, +1, +ncol}.  When nrow * ncol is
// odd, colour 1 has one cell less per layer and its last slot of each layer
// is padding (kept at zero, skipped by every kernel).
//
// Matrices are 7 coefficient planes of NP values in this order of positions;
// plane s holds CSR slot s of the reference's DIS row (diag, up k-1, back
// i-1, left j-1, right j+1, front i+1, down k+1 --
// src/Model/ModelUtilities/Connections.f90:523-603).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MF6_NSLOT 7
enum { S_DIAG = 0, S_UP = 1, S_BACK = 2, S_LEFT = 3, S_RIGHT = 4, S_FRONT = 5, S_DOWN = 6 };

struct Layout {
  int nlay, nrow, ncol;
  int N;    // cells
  int NP;   // positions (2 * H; N + padding)
  int H;    // positions per colour
  int HP;   // positions per colour per layer
  int HP0;  // columns of colour 0
  int HP1;  // columns of colour 1
  int NRC;  // nrow * ncol
  int odd;  // ncol odd
};

inline Layout make_layout(int nlay, int nrow, int ncol) {
  Layout L;
  L.nlay = nlay;
  L.nrow = nrow;
  L.ncol = ncol;
  L.NRC = nrow * ncol;
  L.N = L.NRC * nlay;
  L.HP0 = (L.NRC + 1) / 2;
  L.HP1 = L.NRC / 2;
  L.HP = L.HP0;
  L.H = nlay * L.HP;
  L.NP = 2 * L.H;
  L.odd = ncol & 1;
  return L;
}

// l of the column h2 of colour c
__host__ __device__ inline int col_l(const Layout &L, int c, int h2) {
  if (L.odd) return 2 * h2 + c;
  int i = (2 * h2) / L.ncol;
  return 2 * h2 + ((i & 1) ^ c);
}

__host__ __device__ inline int ncols_of(const Layout &L, int c) { return c ? L.HP1 : L.HP0; }

// Cell of a position; false for a padding slot.
__host__ __device__ inline bool pos_to_cell(const Layout &L, int p, int &c, int &k, int &i,
                                            int &j, int &l) {
  c = p >= L.H;
  int r = p - c * L.H;
  k = r / L.HP;
  int h2 = r - k * L.HP;
  if (h2 >= ncols_of(L, c)) return false;
  l = col_l(L, c, h2);
  i = l / L.ncol;
  j = l - i * L.ncol;
  return true;
}

__host__ __device__ inline int cell_pos(const Layout &L, int k, int i, int j) {
  int l = i * L.ncol + j;
  int c = (i + j) & 1;
  return c * L.H + k * L.HP + (l >> 1);
}

__host__ __device__ inline int pos_of_natural(const Layout &L, int n) {
  int k = n / L.NRC;
  int l = n - k * L.NRC;
  int i = l / L.ncol;
  return cell_pos(L, k, i, l - i * L.ncol);
}

__host__ __device__ inline int natural(const Layout &L, int k, int i, int j) {
  return (k * L.nrow + i) * L.ncol + j;
}

// index, in the other colour's plane, of the horizontal neighbour of l
__host__ __device__ inline int hn(int l, int delta) { return (l + delta) >> 1; }

// Signed largest-magnitude value with its natural location; ties keep the
// smaller location (the reference scans nodes in ascending order with a
// strict '>' and so keeps the first occurrence).
struct MaxLoc {
  double v;
  int loc;
};

__device__ inline MaxLoc maxloc_pick(MaxLoc a, MaxLoc b) {
  double fa = fabs(a.v), fb = fabs(b.v);
  if (fb > fa || (fb == fa && b.loc < a.loc && fb > 0.0)) return b;
  return a;
}

// Device-side state of one Krylov solve (scalars of ims_base_cg / _bcgs).
struct KState {
  double rho, rho0, alpha, alpha0, omega, omega0, beta;
  double l2norm0, epfact, dvclose, rclose;
  double deltax, rmax, l2norm;
  double pq, tq, tt;
  int xloc, rloc;
  int iter;       // inner iterations done (INNERIT)
  int icnvg;      // ICNVG
  int stop;       // 1 once the reference loop would have exited
  int north_due;  // recompute d = b - A x before the next iteration
  int icnvgopt, north, itmax;
  int hist_cap;
  // mixed precision (iterative refinement) bookkeeping
  double inner_r0;
  int inner_iter;
  int inner_stop;
};

#define MF6_BLOCK 256
#define MF6_MAXBLK 8192

// Coarse space of the two-level preconditioner: aggregates of ha x ha grid
// columns (all layers) on an nai x naj aggregate grid; nc = nai naj coarse
// unknowns, dense rows of ld (a multiple of 64) entries.
struct TwoLevel {
  int ha, nai, naj, nc, ld;
};

// aggregate of the grid column l = i ncol + j
__host__ __device__ inline int tl_agg(const TwoLevel &T, const Layout &L, int l) {
  const int i = l / L.ncol, j = l - i * L.ncol;
  return (i / T.ha) * T.naj + j / T.ha;
}

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b runs on
// XCD b mod 8), and each XCD has its own L2.  Stencil kernels read the
// columns one grid row away (+-ncol/2 positions, a few workgroups away), so
// they map workgroups to work in XCD-contiguous ranges: XCD x takes the
// logical blocks [x*q + min(x, r), ...) of a grid of G = 8q + r workgroups.
// Any G gives a bijection of [0, G).
__device__ inline int xcd_block() {
  const int G = (int)gridDim.x, b = (int)blockIdx.x;
  const int x = b & 7, m = b >> 3, q = G >> 3, r = G & 7;
  return x * q + (x < r ? x : r) + m;
}
