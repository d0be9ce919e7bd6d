// MI355X (gfx950) multiscale Q1 basis construction: kernels and C-ABI.
//
// One coarse cell = one independent local problem (the per-cell QBasis of
// MsFEC, source/Q/q_basis.cc).  The batch is processed by three kernels:
//
//   q_assemble_kernel   fine Q1 stencil (lower half + diagonal) of every node,
//                       the load vector F and the Dirichlet-lifting right-hand
//                       sides -K_IB g_k (q_basis.cc:117-257)
//   q_band_kernel       banded LDL^T of the interior system K_II with the
//                       forward solve fused, then the backward solve
//                       (the condense + solve loop, q_basis.cc:545-565)
//   q_element_kernel    Psi = [g on the boundary, x inside], M = Psi^T K Psi,
//                       b = Psi^T F (assemble_global_element_matrix,
//                       q_basis.cc:262-300)
//
// The banded factorisation keeps a 64-slot sliding window of the active
// Schur complement in registers of one wavefront: lane (p,q) = (lane/8,
// lane%8) owns the 8x8 tile (8p..8p+7) x (8q..8q+7) of the window, slot
// s = index mod 64.  Each elimination step reads the pivot row through LDS,
// applies a rank-1 update to the register tiles, streams the finished column
// of L to HBM and recycles the pivot's slot for index j+64, whose (still
// unmodified) matrix row is inserted from the stencil table.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <array>
#include <new>
#include <string>
#include <mutex>
#include <vector>

#include "msfec.h"

namespace {

constexpr int kWin = 64;       // sliding-window slots (>= bandwidth + 1)
constexpr int kMaxNb = 8;      // bases per cell (2^dim)
constexpr int kMaxLow = 14;    // lower stencil entries incl. diagonal (3-D)
constexpr int kRowStride = 24; // row table: 16 lower-stencil slots + 8 lifting rhs
constexpr int kPStride = 80;   // per cell: P (8 x 8) | b_B (8) | |cell| / 2^dim

// Panel slots per cell of the band solvers: one per super-step of 4 pivots,
// rounded up to whole tile rows of 16 pivots (the matrix-core factorisation
// runs every super-step of its last tile row).
inline int panel_quads(int n) { return ((n + 15) / 16) * 4; }

// Tile rows (16 slots each) of the band window: the incoming rows enter the
// slots of the 4 pivots just eliminated, so the half-bandwidth must stay 4
// below the window (bw + 4 <= 16 NT).  requested: 16, 32, 64 slots, or 0 for
// the smallest window that holds the band; 0 is returned when none does.
inline int band_tile_rows(int bw, int requested) {
  for (int nt : {1, 2, 4}) {
    if (requested != 0 && requested != 16 * nt) continue;
    if (bw + 4 <= 16 * nt) return nt;
  }
  return 0;
}

struct QConf {
  int dim, N, S, nn, nfc, nv, ni, n, nst, nlow, center;
  int coef_mode, rhs_mode, reaction_mode;
  double rhs_const, reaction_const;
  int nt;                  // band window of the matrix-core kernels: 16 nt slots
  int off_full[kMaxLow];   // full-grid offset (>0) of lower neighbour t
  int off_int[kMaxLow];    // interior-grid offset (>0) of lower neighbour t
  int dir_low[kMaxLow];    // direction index e of lower neighbour t (t>0)
};

// ---------------------------------------------------------------------------
// Element matrix entries on an axis-aligned box (QGauss(2) is exact for the
// scalar-coefficient case, so the closed tensor-product form is used there).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double box_stiff(int dim, int a, int b, const double* h) {
  double s = 0.0;
  for (int d = 0; d < dim; ++d) {
    double t = ((((a ^ b) >> d) & 1) ? -1.0 : 1.0) / h[d];
    for (int e = 0; e < dim; ++e)
      if (e != d) t *= ((((a ^ b) >> e) & 1) ? (1.0 / 6.0) : (1.0 / 3.0)) * h[e];
    s += t;
  }
  return s;
}

__device__ __forceinline__ double box_mass(int dim, int a, int b, const double* h) {
  double t = 1.0;
  for (int e = 0; e < dim; ++e) t *= ((((a ^ b) >> e) & 1) ? (1.0 / 6.0) : (1.0 / 3.0)) * h[e];
  return t;
}

__device__ __constant__ double kGQ[2] = {0.5 - 0.28867513459481288225,
                                         0.5 + 0.28867513459481288225};

// grad phi_a at reference point xi (physical derivative on the box)
__device__ __forceinline__ void box_grad(int dim, int a, const double* xi, const double* h,
                                         double* g) {
  for (int d = 0; d < dim; ++d) {
    double v = (((a >> d) & 1) ? 1.0 : -1.0) / h[d];
    for (int e = 0; e < dim; ++e)
      if (e != d) v *= ((a >> e) & 1) ? xi[e] : 1.0 - xi[e];
    g[d] = v;
  }
}

__device__ __forceinline__ double box_val(int dim, int a, const double* xi) {
  double v = 1.0;
  for (int e = 0; e < dim; ++e) v *= ((a >> e) & 1) ? xi[e] : 1.0 - xi[e];
  return v;
}

// K_z(a,b) for fine cell z
__device__ double cell_entry(const QConf& c, const double* coef, const double* reac,
                             int z, int a, int b, const double* h) {
  double v;
  if (c.coef_mode == MSFEC_COEF_SCALAR_CELL) {
    v = coef[z] * box_stiff(c.dim, a, b, h);
  } else {
    const int nc = (c.dim == 2) ? 3 : 6;
    double vol = 1.0;
    for (int d = 0; d < c.dim; ++d) vol *= h[d];
    v = 0.0;
    for (int q = 0; q < c.nv; ++q) {
      double xi[3];
      double w = 1.0;
      for (int d = 0; d < c.dim; ++d) {
        xi[d] = kGQ[(q >> d) & 1];
        w *= 0.5;
      }
      double ga[3], gb[3];
      box_grad(c.dim, a, xi, h, ga);
      box_grad(c.dim, b, xi, h, gb);
      const double* s = coef + ((size_t)z * c.nv + q) * nc;
      double Ag[3];
      if (c.dim == 2) {
        Ag[0] = s[0] * gb[0] + s[1] * gb[1];
        Ag[1] = s[1] * gb[0] + s[2] * gb[1];
      } else {
        Ag[0] = s[0] * gb[0] + s[1] * gb[1] + s[2] * gb[2];
        Ag[1] = s[1] * gb[0] + s[3] * gb[1] + s[4] * gb[2];
        Ag[2] = s[2] * gb[0] + s[4] * gb[1] + s[5] * gb[2];
      }
      double t = 0.0;
      for (int d = 0; d < c.dim; ++d) t += ga[d] * Ag[d];
      v += t * w * vol;
    }
  }
  if (c.reaction_mode != MSFEC_REACTION_NONE) {
    const double r = (c.reaction_mode == MSFEC_REACTION_CONST) ? c.reaction_const : reac[z];
    v += r * box_mass(c.dim, a, b, h);
  }
  return v;
}

// K[n][n+dd] for node coordinates cn and direction dd (any of the 3^dim)
__device__ double pair_value(const QConf& c, const double* coef, const double* reac,
                             const int* cn, const int* dd, const double* h) {
  // cells z = cn - alpha with alpha_d fixed where dd_d != 0
  double s = 0.0;
  for (int al = 0; al < c.nv; ++al) {
    bool ok = true;
    int z = 0, mul = 1, a = 0, b = 0;
    for (int d = 0; d < c.dim; ++d) {
      const int ad = (al >> d) & 1;
      const int bd = ad + dd[d];
      if (bd < 0 || bd > 1) { ok = false; break; }
      const int zd = cn[d] - ad;
      if (zd < 0 || zd >= c.N) { ok = false; break; }
      z += zd * mul;
      mul *= c.N;
      a |= ad << d;
      b |= bd << d;
    }
    if (ok) s += cell_entry(c, coef, reac, z, a, b, h);
  }
  return s;
}

__device__ __forceinline__ void cell_box(const QConf& c, const double* verts, int cell,
                                         double* lo, double* h) {
  const double* v = verts + (size_t)cell * c.nv * c.dim;
  for (int d = 0; d < c.dim; ++d) {
    lo[d] = v[d];
    h[d] = (v[(c.nv - 1) * c.dim + d] - v[d]) / c.N;
  }
}

// g_k at fine node with grid coordinates m: trilinear coarse basis on the box
__device__ __forceinline__ double lift_value(const QConf& c, int k, const int* m) {
  double v = 1.0;
  for (int d = 0; d < c.dim; ++d) {
    const double xi = (double)m[d] / (double)c.N;
    v *= ((k >> d) & 1) ? xi : 1.0 - xi;
  }
  return v;
}

// ---------------------------------------------------------------------------
// K1: stencil table, load vector and lifting right-hand sides.
// Slo[cell][node][t]: t = 0 diagonal, t = 1+e lower direction e < center.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void q_assemble_kernel(
    QConf c, const double* __restrict__ verts, const double* __restrict__ coef_all,
    size_t coef_stride, const double* __restrict__ f_all, size_t f_stride,
    const double* __restrict__ reac_all, size_t reac_stride, double* __restrict__ Slo,
    double* __restrict__ Fv, double* __restrict__ Rhs) {
  const int cell = blockIdx.x;
  const double* coef = coef_all + (size_t)cell * coef_stride;
  const double* fcell = f_all ? f_all + (size_t)cell * f_stride : nullptr;
  const double* reac = reac_all ? reac_all + (size_t)cell * reac_stride : nullptr;
  double lo[3], h[3];
  cell_box(c, verts, cell, lo, h);
  double vol = 1.0;
  for (int d = 0; d < c.dim; ++d) vol *= h[d];
  double* slo = Slo + (size_t)cell * c.nn * c.nlow;
  double* fv = Fv + (size_t)cell * c.nn;
  double* rhs = Rhs + (size_t)cell * c.n * kMaxNb;

  for (int node = threadIdx.x; node < c.nn; node += blockDim.x) {
    int cn[3] = {0, 0, 0};
    {
      int r = node;
      for (int d = 0; d < c.dim; ++d) {
        cn[d] = r % c.S;
        r /= c.S;
      }
    }
    // lower stencil
    for (int t = 0; t < c.nlow; ++t) {
      int dd[3] = {0, 0, 0};
      if (t > 0) {
        int e = c.dir_low[t];
        for (int d = 0; d < c.dim; ++d) {
          dd[d] = e % 3 - 1;
          e /= 3;
        }
      }
      bool inside = true;
      for (int d = 0; d < c.dim; ++d) {
        const int m = cn[d] + dd[d];
        if (m < 0 || m > c.N) inside = false;
      }
      slo[(size_t)node * c.nlow + t] = inside ? pair_value(c, coef, reac, cn, dd, h) : 0.0;
    }
    // load vector: sum over the cells touching the node
    double fs = 0.0;
    for (int al = 0; al < c.nv; ++al) {
      int z = 0, mul = 1;
      bool ok = true;
      for (int d = 0; d < c.dim; ++d) {
        const int zd = cn[d] - ((al >> d) & 1);
        if (zd < 0 || zd >= c.N) { ok = false; break; }
        z += zd * mul;
        mul *= c.N;
      }
      if (!ok) continue;
      if (c.rhs_mode == MSFEC_RHS_QP) {
        for (int q = 0; q < c.nv; ++q) {
          double xi[3], w = 1.0;
          for (int d = 0; d < c.dim; ++d) {
            xi[d] = kGQ[(q >> d) & 1];
            w *= 0.5;
          }
          fs += box_val(c.dim, al, xi) * fcell[(size_t)z * c.nv + q] * w * vol;
        }
      } else {
        const double fz = (c.rhs_mode == MSFEC_RHS_CELL) ? fcell[z] : c.rhs_const;
        fs += fz * vol / c.nv;
      }
    }
    fv[node] = fs;
    // lifting right-hand sides for interior nodes
    bool interior = true;
    for (int d = 0; d < c.dim; ++d)
      if (cn[d] == 0 || cn[d] == c.N) interior = false;
    if (interior) {
      int ii = 0, mul = 1;
      for (int d = 0; d < c.dim; ++d) {
        ii += (cn[d] - 1) * mul;
        mul *= c.ni;
      }
      double acc[kMaxNb];
      for (int k = 0; k < kMaxNb; ++k) acc[k] = 0.0;
      for (int e = 0; e < c.nst; ++e) {
        if (e == c.center) continue;
        int dd[3] = {0, 0, 0}, m[3] = {0, 0, 0}, r = e;
        bool bnd = false;
        for (int d = 0; d < c.dim; ++d) {
          dd[d] = r % 3 - 1;
          r /= 3;
          m[d] = cn[d] + dd[d];
          if (m[d] == 0 || m[d] == c.N) bnd = true;
        }
        if (!bnd) continue;
        const double kv = pair_value(c, coef, reac, cn, dd, h);
        for (int k = 0; k < c.nv; ++k) acc[k] -= kv * lift_value(c, k, m);
      }
      for (int k = 0; k < kMaxNb; ++k) rhs[(size_t)ii * kMaxNb + k] = acc[k];
    }
  }
}

// ---------------------------------------------------------------------------
// K2: banded LDL^T in a register-resident sliding window (one wave per cell)
//
// Staging: the matrix rows and lifting right-hand sides that enter the window
// during the next 8 elimination steps are prepared in LDS once per group of 8
// steps, either computed from the LDS-resident fine-cell coefficients
// (scalar coefficient on a box: K[n][n+dd] = kv[mask(dd)] * sum of the
// coefficients of the fine cells shared by n and n+dd, plus the mass
// analogue for the reaction term) or read from the tables of
// q_assemble_kernel (tensor coefficients, quadrature-point loads).
// Synchronisation inside the wave uses wavefront-scope fences only, so the
// streamed stores of L never stall the elimination.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct BandArgs {
  int n, ni, N, S, nn, nfc;
  int nq;  // super-steps of 4 pivots (MFMA variant)
  int off_int[kMaxLow];
  // LDL^T rows (symmetric indefinite band): [n][16] row - column offsets of
  // the lower entries of each row, -1 where unused
  const int* loff;
  // general path tables (q_assemble_kernel)
  const double* Slo;      // [cell][nn][nlow]
  const double* Rhs;      // [cell][n][8]
  const int* inode;       // interior index -> node id
  const unsigned* imask;  // bit t: lower neighbour t is interior
  // fast path data
  const double* Rows;     // [cell][n][kRowStride]: lower stencil (16) | lifting rhs (8)
  const double* verts;    // [cell][2^dim][dim]
  const double* coef;     // [cell][nfc]
  const double* reac;     // [cell][nfc] or null
  const double* fcell;    // [cell][nfc] or null
  int reaction_mode, rhs_mode;
  double reaction_const, rhs_const;
  // scratch and outputs
  void* pan;              // matrix-core kernels: [cell][nq][4][64] panels (FP64 or FP32)
  void* zst;              // matrix-core kernels: [cell][nq][4][8] forward results
  double* Rres;           // mixed solver: [cell][n][8] refinement residual
  int* conv;              // mixed solver: [cell] 1 once the cell stops refining
  double* Pb;             // fast path: [cell][kPStride] P = G_B^T K_BB G_B | b_B | |cell| / 2^dim
  int last_pass;          // mixed solver: this refinement pass is the last one
  double* Lcol;  // [cell][n][64]  o = 0: pivot d_j, o > 0: L[j+o][j]
  double* Wf;    // [cell][n][8]   D^-1 L^-1 rhs
  double* X;     // [cell][n][8]   interior solution
  double* Psi;   // [cell][nv][nn]
  double* M;     // [cell][nv][nv]
  double* bvec;  // [cell][nv]
  int* status;   // [cell]
};

template <int DIM>
struct Dirs {
  static constexpr int NV = 1 << DIM;
  static constexpr int NST = (DIM == 3) ? 27 : 9;
  static constexpr int CENTER = (NST - 1) / 2;
  static constexpr int NLOW = CENTER + 1;
};

// LDS layouts pad every group of 8 doubles to 9 so that the 8 distinct
// groups read together by a wave fall into distinct banks.
__device__ __forceinline__ int pad8(int x) { return 9 * (x >> 3) + (x & 7); }

struct BandSmem {
  union {
    double ubuf[72];     // vector-ALU band kernel: pivot row, slot x at pad8(x)
    int stg_off[8][16];  // matrix-core LDL^T rows: offsets of the staged lower entries
  };
  double nbuf[72];   // incoming row, slot x at pad8(x)
  double ybuf[8];
  double kv[8];
  double mv[8];
  double stg_low[8][16];
  double stg_rhs[8][8];
  // phase-dependent pool:
  //   forward : RHS partials [64][9]
  //   backward: w of the group [8][8] | L columns of the group [8][72] | x window [64][9]
  //   epilogue: psi of 64 nodes [64][9] | (K psi)_B of 64 nodes [64][9] | F of 64 nodes [64]
  //   matrix-core backward: x ring [8][132] | z [4][8] | table rows [4][24] | x [4][8] | residual [4][8]
  double pool[1280];
};
#define SH_PART(sh) ((sh).pool)
#define SH_WBUF(sh) ((sh).pool)
#define SH_LBUF(sh) ((sh).pool + 64)
#define SH_XRING(sh) ((sh).pool + 64 + 576)
#define SH_PN(sh) ((sh).pool)
#define SH_KP(sh) ((sh).pool + 576)
#define SH_FS(sh) ((sh).pool + 1152)

// direction e -> component d in {-1,0,1}
template <int DIM>
__device__ __forceinline__ int dcomp(int e, int d) {
  return (d == 0) ? (e % 3) - 1 : (d == 1) ? ((e / 3) % 3) - 1 : (e / 9) - 1;
}

// Sum of the coefficients of the fine cells shared by node c and c+dd
// (cells clipped to the grid); cstride = (1, N, N^2).
template <int DIM>
__device__ __forceinline__ void shared_sums(const double* cf, const double* rf, bool has_rf,
                                            const int* c, const int* dd, int N, double& sa,
                                            double& sr, int& cnt) {
  sa = 0.0;
  sr = 0.0;
  cnt = 0;
#pragma unroll
  for (int al = 0; al < (1 << DIM); ++al) {
    bool ok = true;
    int z = 0, mul = 1;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const int ad = (al >> d) & 1;
      if (dd[d] == -1 && ad != 1) ok = false;
      if (dd[d] == 1 && ad != 0) ok = false;
      const int zd = c[d] - ad;
      if (zd < 0 || zd >= N) ok = false;
      z += zd * mul;
      mul *= N;
    }
    if (ok) {
      sa += cf[z];
      if (has_rf) sr += rf[z];
      ++cnt;
    }
  }
}

// g_k at a fine node with grid coordinates m (trilinear coarse basis on the box)
template <int DIM>
__device__ __forceinline__ double lift_k(int k, const int* m, double invN) {
  double v = 1.0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const double xi = (double)m[d] * invN;
    v *= ((k >> d) & 1) ? xi : 1.0 - xi;
  }
  return v;
}

// Sums over the fine cells shared by node c and each neighbour c + dd, for
// all 3^dim directions e (dd_d = e's base-3 digit d minus 1) at once: the
// coefficient (SA) and the reaction (SR, including the constant part).  The
// 2^dim cells around the node are loaded once (cells outside the grid count
// as zero); direction e sums the cells it shares with the node: dd_d = -1
// keeps the cells below the node in direction d, +1 those above, 0 both.
template <int DIM>
__device__ __forceinline__ void node_sums(const double* cf, const double* rf, bool has_rf,
                                          double rc, const int* c, int N, double* SA,
                                          double* SR) {
  constexpr int NV = 1 << DIM;
  constexpr int NST = (DIM == 3) ? 27 : 9;
  double ca[NV], cr[NV];
#pragma unroll
  for (int al = 0; al < NV; ++al) {
    bool ok = true;
    int z = 0, mul = 1;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const int zd = c[d] - ((al >> d) & 1);
      if (zd < 0 || zd >= N) ok = false;
      z += zd * mul;
      mul *= N;
    }
    ca[al] = ok ? cf[z] : 0.0;
    cr[al] = ok ? (has_rf ? rf[z] : 0.0) + rc : 0.0;
  }
  // dimension 0: al bit 0 -> three direction values per remaining bits
  double ta[NST], tr[NST];
#pragma unroll
  for (int e = 0; e < NST; ++e) {
    // al bits in dimension d: dd = -1 -> cell index bit 1, +1 -> bit 0, 0 -> both
    double sa = 0.0, sr = 0.0;
#pragma unroll
    for (int al = 0; al < NV; ++al) {
      bool take = true;
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        const int dd = (d == 0) ? (e % 3) - 1 : (d == 1) ? ((e / 3) % 3) - 1 : (e / 9) - 1;
        const int b = (al >> d) & 1;
        if ((dd == -1 && b != 1) || (dd == 1 && b != 0)) take = false;
      }
      if (take) {
        sa += ca[al];
        sr += cr[al];
      }
    }
    ta[e] = sa;
    tr[e] = sr;
  }
#pragma unroll
  for (int e = 0; e < NST; ++e) {
    SA[e] = ta[e];
    SR[e] = tr[e];
  }
}

// out[k] = sum_e w[e] g_k(c + dd_e) for the 2^dim coarse bases k, with the
// lift factorised per dimension (fl[d][dd_d + 1][k_d]): contracted one
// dimension at a time instead of forming g_k at every neighbour.
template <int DIM>
__device__ __forceinline__ void lift_contract(const double* w, const double (&fl)[3][3][2],
                                              double* out) {
  if (DIM == 2) {
    double t[3][2];  // [ey][kx]
#pragma unroll
    for (int ey = 0; ey < 3; ++ey)
#pragma unroll
      for (int kx = 0; kx < 2; ++kx)
        t[ey][kx] = w[3 * ey + 0] * fl[0][0][kx] + w[3 * ey + 1] * fl[0][1][kx] +
                    w[3 * ey + 2] * fl[0][2][kx];
#pragma unroll
    for (int ky = 0; ky < 2; ++ky)
#pragma unroll
      for (int kx = 0; kx < 2; ++kx)
        out[kx + 2 * ky] = t[0][kx] * fl[1][0][ky] + t[1][kx] * fl[1][1][ky] + t[2][kx] * fl[1][2][ky];
  } else {
    double t[9][2];  // [ey + 3 ez][kx]
#pragma unroll
    for (int q = 0; q < 9; ++q)
#pragma unroll
      for (int kx = 0; kx < 2; ++kx)
        t[q][kx] = w[3 * q + 0] * fl[0][0][kx] + w[3 * q + 1] * fl[0][1][kx] +
                   w[3 * q + 2] * fl[0][2][kx];
    double u[3][4];  // [ez][kx + 2 ky]
#pragma unroll
    for (int ez = 0; ez < 3; ++ez)
#pragma unroll
      for (int ky = 0; ky < 2; ++ky)
#pragma unroll
        for (int kx = 0; kx < 2; ++kx)
          u[ez][kx + 2 * ky] = t[3 * ez + 0][kx] * fl[1][0][ky] + t[3 * ez + 1][kx] * fl[1][1][ky] +
                               t[3 * ez + 2][kx] * fl[1][2][ky];
#pragma unroll
    for (int kz = 0; kz < 2; ++kz)
#pragma unroll
      for (int q = 0; q < 4; ++q)
        out[q + 4 * kz] = u[0][q] * fl[2][0][kz] + u[1][q] * fl[2][1][kz] + u[2][q] * fl[2][2][kz];
  }
}


// Element stiffness / mass of a box as functions of a^b (bit d set: the two
// vertices differ in direction d), for lanes x < 8.
template <int DIM>
__device__ __forceinline__ void box_tables(int x, const double* hh, double& ks, double& ms) {
  ks = 0.0;
  ms = 1.0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    double t = (((x >> d) & 1) ? -1.0 : 1.0) / hh[d];
#pragma unroll
    for (int e = 0; e < DIM; ++e)
      if (e != d) t *= (((x >> e) & 1) ? (1.0 / 6.0) : (1.0 / 3.0)) * hh[e];
    ks += t;
    ms *= (((x >> d) & 1) ? (1.0 / 6.0) : (1.0 / 3.0)) * hh[d];
  }
}

// ---------------------------------------------------------------------------
// K1f: interior row table for a scalar coefficient per fine cell (fast path).
// Row j of cell c: Rows[c][j][0..NLOW-1] = lower stencil (t = 0 diagonal,
// t = 1+e lower direction e, zero where the neighbour is on the boundary),
// Rows[c][j][16..23] = lifting right-hand sides -sum_{m in B} K[j][m] g_k(m).
// K[n][n+dd] = kv[mask(dd)] * (sum of the coefficients of the fine cells
// shared by n and n+dd) + the mass analogue for the reaction term.
//
// The same pass gives the boundary part of the element matrix and
// right-hand side and the boundary values of Psi: with g_k the lifting
// (coarse Q1 basis k at the boundary nodes, zero inside),
//   P[i][k] = sum_{m in B} g_i(m) (K_BB g_k)(m),   b_B[i] = sum_{m in B} g_i(m) F(m).
// One block per cell; the fine nodes are taken 256 at a time, one per thread,
// interior and boundary nodes alike: both need the node's coupling sums and
// one contraction of the couplings to boundary neighbours with the lifting
// (-K_IB g_k for an interior row, K_BB g_k at a boundary node), so a wave
// does that work once for whatever mix of nodes it holds.  Threads
// (i, k) < (2^dim)^2 then sum their entry of P over the chunk.
// ---------------------------------------------------------------------------
template <int DIM, bool REAC>
__global__ __launch_bounds__(256) void q_rows_kernel(BandArgs a) {
  using D = Dirs<DIM>;
  const int cell = blockIdx.x;
  extern __shared__ double dyn[];  // coefficient | reaction | load
  __shared__ double kv[8], mv[8];
  // chunk buffers of the boundary sums: g | K_BB g of 256 nodes (stride 9) | F
  __shared__ double rbuf[256 * 19];
  double* cf = dyn;
  double* rf = dyn + a.nfc;
  const bool has_rf = (a.reaction_mode == MSFEC_REACTION_CELL);
  double* ff = dyn + a.nfc * (has_rf ? 2 : 1);
  if (a.rhs_mode == MSFEC_RHS_CELL) {
    const double* fg = a.fcell + (size_t)cell * a.nfc;
    for (int z = threadIdx.x; z < a.nfc; z += blockDim.x) ff[z] = fg[z];
  }
  const double* cg = a.coef + (size_t)cell * a.nfc;
  for (int z = threadIdx.x; z < a.nfc; z += blockDim.x) cf[z] = cg[z];
  if (has_rf) {
    const double* rg = a.reac + (size_t)cell * a.nfc;
    for (int z = threadIdx.x; z < a.nfc; z += blockDim.x) rf[z] = rg[z];
  }
  double vol = 1.0;
  {
    const double* v = a.verts + (size_t)cell * D::NV * DIM;
    double hh[3] = {1.0, 1.0, 1.0};
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      hh[d] = (v[(D::NV - 1) * DIM + d] - v[d]) / a.N;
      vol *= hh[d];
    }
    if (threadIdx.x < 8) {
      double ks, ms;
      box_tables<DIM>(threadIdx.x, hh, ks, ms);
      kv[threadIdx.x] = ks;
      mv[threadIdx.x] = ms;
    }
  }
  __syncthreads();
  const double rc = (a.reaction_mode == MSFEC_REACTION_CONST) ? a.reaction_const : 0.0;
  const double invN = 1.0 / (double)a.N;
  const int tid = threadIdx.x;
  double* gs = rbuf;
  double* kps = rbuf + 256 * 9;
  double* fsh = rbuf + 256 * 18;
  double* psi_out = a.Psi + (size_t)cell * D::NV * a.nn;
  double* rows_out = const_cast<double*>(a.Rows) + (size_t)cell * a.n * kRowStride;
  // thread (quarter qt, pair (pi, pk)) sums its quarter of every chunk:
  // 4 independent chains over 64 nodes instead of one over 256
  const int qt = tid >> 6, pi = (tid >> 3) & 7, pk = tid & 7;
  const bool pthr = pi < D::NV && pk < D::NV;
  const bool bthr = pi == 0 && pk < D::NV;
  double pacc = 0.0, bacc = 0.0;
  for (int base = 0; base < a.nn; base += 256) {
    const int node = base + tid;
    const bool live = node < a.nn;
    int c[3] = {0, 0, 0};
    bool bnd = false;
    {
      int t = live ? node : 0;
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        c[d] = t % a.S;
        t /= a.S;
        if (c[d] == 0 || c[d] == a.N) bnd = true;
      }
    }
    // out: the contraction with the lifting (8 slots: the row table's
    // right-hand sides, zero beyond 2^dim)
    double g[D::NV], out[8], fsum = 0.0;
#pragma unroll
    for (int k = 0; k < D::NV; ++k) g[k] = 0.0;
#pragma unroll
    for (int k = 0; k < 8; ++k) out[k] = 0.0;
    if (live) {
      double SA[D::NST], SR[D::NST];
      node_sums<DIM>(cf, rf, has_rf, rc, c, a.N, SA, SR);
      double fl[3][3][2];
#pragma unroll
      for (int d = 0; d < DIM; ++d)
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          const double t = (double)(c[d] + q - 1) * invN;
          fl[d][q][0] = 1.0 - t;
          fl[d][q][1] = t;
        }
      // couplings to boundary neighbours inside the grid (the diagonal too at
      // a boundary node): -K_IB at an interior row, K_BB at a boundary node;
      // the lower stencil of an interior row from its interior neighbours
      double lo[16], wb[D::NST];
#pragma unroll
      for (int t = 0; t < 16; ++t) lo[t] = 0.0;
#pragma unroll
      for (int e = 0; e < D::NST; ++e) {
        int mask = 0;
        bool inside = true, mb = false;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
          const int dd = dcomp<DIM>(e, d);
          if (dd != 0) mask |= 1 << d;
          const int m = c[d] + dd;
          if (m < 0 || m > a.N) inside = false;
          if (m == 0 || m == a.N) mb = true;
        }
        const double kval = kv[mask] * SA[e] + (REAC ? mv[mask] * SR[e] : 0.0);
        const bool take = inside && mb && (bnd || e != D::CENTER);
        wb[e] = take ? (bnd ? kval : -kval) : 0.0;
        if (e == D::CENTER) lo[0] = kval;
        else if (!mb && e < D::CENTER) lo[1 + e] = kval;
      }
      lift_contract<DIM>(wb, fl, out);
      if (!bnd) {
        // slot 14: sum of the loads of the fine cells around the node (the
        // interior load is this times |cell| / 2^dim); slot 15: the node index
        double fj;
        if (a.rhs_mode == MSFEC_RHS_CELL) {
          fj = 0.0;
#pragma unroll
          for (int al = 0; al < D::NV; ++al) {
            int z = 0, mul = 1;
#pragma unroll
            for (int d = 0; d < DIM; ++d) {
              z += (c[d] - ((al >> d) & 1)) * mul;
              mul *= a.N;
            }
            fj += ff[z];
          }
        } else {
          fj = a.rhs_const * D::NV;
        }
        lo[14] = fj;
        lo[15] = (double)node;
        int jn = 0;
#pragma unroll
        for (int d = DIM - 1; d >= 0; --d) jn = jn * a.ni + (c[d] - 1);
        double2* dst = reinterpret_cast<double2*>(rows_out + (size_t)jn * kRowStride);
#pragma unroll
        for (int t = 0; t < 8; ++t) dst[t] = make_double2(lo[2 * t], lo[2 * t + 1]);
#pragma unroll
        for (int k = 0; k < 4; ++k) dst[8 + k] = make_double2(out[2 * k], out[2 * k + 1]);
#pragma unroll
        for (int k = 0; k < D::NV; ++k) out[k] = 0.0;  // no boundary contribution
      } else {
#pragma unroll
        for (int k = 0; k < D::NV; ++k) {
          g[k] = lift_k<DIM>(k, c, invN);
          psi_out[(size_t)k * a.nn + node] = g[k];
        }
        // F[node] = sum over the fine cells touching the node of f_z |z| / 2^dim
#pragma unroll
        for (int al = 0; al < D::NV; ++al) {
          bool ok = true;
          int z = 0, mul = 1;
#pragma unroll
          for (int d = 0; d < DIM; ++d) {
            const int zd = c[d] - ((al >> d) & 1);
            if (zd < 0 || zd >= a.N) ok = false;
            z += zd * mul;
            mul *= a.N;
          }
          if (ok) fsum += (a.rhs_mode == MSFEC_RHS_CELL) ? ff[z] : a.rhs_const;
        }
        fsum *= vol / D::NV;
      }
    }
    __syncthreads();  // the previous chunk has been summed
#pragma unroll
    for (int k = 0; k < D::NV; ++k) {
      gs[tid * 9 + k] = g[k];
      kps[tid * 9 + k] = out[k];
    }
    fsh[tid] = fsum;
    __syncthreads();
    if (pthr) {
      double p4[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 16
      for (int l = 64 * qt; l < 64 * qt + 64; ++l)
        p4[l & 3] = fma(gs[l * 9 + pi], kps[l * 9 + pk], p4[l & 3]);
      pacc += (p4[0] + p4[1]) + (p4[2] + p4[3]);
    }
    if (bthr) {
      for (int l = 64 * qt; l < 64 * qt + 64; ++l) bacc = fma(gs[l * 9 + pk], fsh[l], bacc);
    }
  }
  // the 4 quarters add up (the chunk buffers are free again)
  __syncthreads();
  gs[tid] = pacc;
  kps[tid] = bacc;
  __syncthreads();
  double* pb = a.Pb + (size_t)cell * kPStride;
  if (tid < 64 && pthr)
    pb[pi * 8 + pk] = (gs[tid] + gs[64 + tid]) + (gs[128 + tid] + gs[192 + tid]);
  if (tid < 64 && bthr)
    pb[64 + pk] = (kps[tid] + kps[64 + tid]) + (kps[128 + tid] + kps[192 + tid]);
  if (tid == 0) pb[72] = vol / D::NV;
}

// Rows jbase..jbase+7 of the row table are 192 contiguous doubles: every
// lane loads 3 of them into registers (issued one staging group ahead) and
// later stores them to the LDS staging area.
// LDL^T rows also carry the offsets of their lower entries (8 rows x 16 ints,
// two per lane).
struct RowPrefetch {
  double v[3];
  long long o;
};

template <bool LDL = false>
__device__ __forceinline__ void rows_load(const BandArgs& a, int cell, int jbase, RowPrefetch& pf) {
  const int lane = threadIdx.x;
  const double* src = a.Rows + ((size_t)cell * a.n + jbase) * kRowStride;
  const int lim = (a.n - jbase) * kRowStride;  // valid doubles from src on
#pragma unroll
  for (int u = 0; u < 3; ++u) {
    const int idx = lane + 64 * u;
    pf.v[u] = (idx < lim) ? src[idx] : 0.0;
  }
  if (LDL) {
    const long long* os = reinterpret_cast<const long long*>(a.loff + (size_t)jbase * 16);
    pf.o = (2 * lane < (a.n - jbase) * 16) ? os[lane] : -1LL;  // -1: two unused (-1) entries
  }
}

template <bool LDL = false>
__device__ __forceinline__ void rows_store(const RowPrefetch& pf, BandSmem& sh) {
  const int lane = threadIdx.x;
#pragma unroll
  for (int u = 0; u < 3; ++u) {
    const int idx = lane + 64 * u;
    const int r = idx / kRowStride, t = idx - kRowStride * r;
    if (t < 16) sh.stg_low[r][t] = pf.v[u];
    else sh.stg_rhs[r][t - 16] = pf.v[u];
  }
  if (LDL) reinterpret_cast<long long*>(&sh.stg_off[0][0])[lane] = pf.o;
  wave_sync();
}


// Stage the rows jbase..jbase+7 (lower stencil with zeros for non-interior
// neighbours, and the lifting right-hand sides) into LDS from the tables of
// q_assemble_kernel (general path).
template <int DIM>
__device__ __forceinline__ void stage_rows_general(const BandArgs& a, int cell, int jbase,
                                                   BandSmem& sh) {
  using D = Dirs<DIM>;
  const int lane = threadIdx.x;
  const int r = lane >> 3, s = lane & 7;
  const int jn = jbase + r;
  if (jn < a.n) {
    if (s == 0) {
      const unsigned msk = a.imask[jn];
      const double* src = a.Slo + ((size_t)cell * a.nn + a.inode[jn]) * D::NLOW;
#pragma unroll
      for (int t = 0; t < D::NLOW; ++t) sh.stg_low[r][t] = ((msk >> t) & 1u) ? src[t] : 0.0;
    }
    sh.stg_rhs[r][s] = a.Rhs[((size_t)cell * a.n + jn) * kMaxNb + s];
  }
  wave_sync();
}

// One staging group: the fast path stores the prefetched rows and issues the
// loads of the next group; the general path reads its tables directly.
template <int DIM, bool FAST>
__device__ __forceinline__ void stage_group(const BandArgs& a, int cell, int jbase, BandSmem& sh,
                                            RowPrefetch& pf) {
  if (FAST) {
    rows_store(pf, sh);
    rows_load(a, cell, jbase + 8, pf);
  } else {
    stage_rows_general<DIM>(a, cell, jbase, sh);
  }
}

// Insert index jn (staged as row I) into slot s = 8G+I of the window.
template <int DIM, int I>
__device__ __forceinline__ void band_insert(const BandArgs& a, int jn, int G, int my_off,
                                            double (&A)[8][8], double (&y)[8], BandSmem& sh) {
  using D = Dirs<DIM>;
  const int lane = threadIdx.x;
  const int p = lane >> 3, q = lane & 7;
  const int s = 8 * G + I;
  sh.nbuf[pad8(lane)] = 0.0;
  wave_sync();
  if (jn < a.n) {
    // only nonzero entries: they belong to distinct interior neighbours and
    // hence to distinct slots, while the offsets of non-neighbour directions
    // can coincide (e.g. with the diagonal when there is a single interior
    // node per direction)
    if (lane < D::NLOW) {
      const double v = sh.stg_low[I][lane];
      if (v != 0.0) sh.nbuf[pad8((s - my_off) & (kWin - 1))] = v;
    }
  } else if (lane == 0) {
    sh.nbuf[pad8(s)] = 1.0;  // past the end: an isolated unit diagonal
  }
  wave_sync();
  const int base = (p == G) ? 9 * q : 9 * p;
  double v[8];
#pragma unroll
  for (int b = 0; b < 8; ++b) v[b] = sh.nbuf[base + b];
  if (p == G) {
#pragma unroll
    for (int b = 0; b < 8; ++b) A[I][b] = v[b];
  }
  if (q == G) {
#pragma unroll
    for (int r = 0; r < 8; ++r) A[r][I] = v[r];
  }
  if (lane == s) {
    const bool live = jn < a.n;
#pragma unroll
    for (int k = 0; k < 8; ++k) y[k] = live ? sh.stg_rhs[I][k] : 0.0;
  }
}

template <int DIM, int I>
__device__ __forceinline__ void band_step(const BandArgs& a, int cell, int j, int G, int my_off,
                                          double (&A)[8][8], double (&y)[8], BandSmem& sh,
                                          int& bad, double& sacc) {
  const int lane = threadIdx.x;
  const int p = lane >> 3, q = lane & 7;
  const int s = 8 * G + I;
  if (p == G) {
#pragma unroll
    for (int b = 0; b < 8; ++b) sh.ubuf[9 * q + b] = A[I][b];
  }
  if (lane == s) {
#pragma unroll
    for (int k = 0; k < 8; ++k) sh.ybuf[k] = y[k];
  }
  wave_sync();
  const double d = sh.ubuf[pad8(s)];
  // keep the first failure: later pivots of a failed cell are meaningless
  if (!(d > 0.0) && bad == 0)
    bad = (d == d && d > -INFINITY) ? (int)MSFEC_CELL_NOT_SPD : (int)MSFEC_CELL_NONFINITE;
  const double rd = 1.0 / d;
  double l[8], u[8], ys[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) l[r] = sh.ubuf[9 * p + r] * rd;
#pragma unroll
  for (int b = 0; b < 8; ++b) u[b] = sh.ubuf[9 * q + b];
#pragma unroll
  for (int k = 0; k < 8; ++k) ys[k] = sh.ybuf[k];
  const double lm = sh.ubuf[pad8(lane)] * rd;
  const double wk = sh.ybuf[lane & 7] * rd;
  // Schur term of the element matrix: lane (i,k) accumulates z_j[i] z_j[k] / d_j
  sacc = fma(sh.ybuf[lane >> 3], wk, sacc);
  a.Lcol[((size_t)cell * a.n + j) * kWin + ((lane - s) & (kWin - 1))] = (lane == s) ? d : lm;
  if (lane < 8) a.Wf[((size_t)cell * a.n + j) * kMaxNb + lane] = wk;
#pragma unroll
  for (int r = 0; r < 8; ++r)
#pragma unroll
    for (int b = 0; b < 8; ++b) A[r][b] = fma(-l[r], u[b], A[r][b]);
  if (lane != s) {
#pragma unroll
    for (int k = 0; k < 8; ++k) y[k] = fma(-lm, ys[k], y[k]);
  }
  wave_sync();
  band_insert<DIM, I>(a, j + kWin, G, my_off, A, y, sh);
}

// Running interior coordinates of the backward sweep (j decreasing).
struct BackPos {
  int x, y, z;  // interior coordinates of j
};

template <int DIM, bool FAST, int R>
__device__ __forceinline__ void back_step(const BandArgs& a, int cell, int j, BandSmem& sh,
                                          BackPos& pos, double& bint, const double* ff) {
  using D = Dirs<DIM>;
  const int lane = threadIdx.x;
  const int k = lane & 7, m = lane >> 3;
  double part = 0.0;
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    const double lv = (m == 0 && t == 0) ? 0.0 : SH_LBUF(sh)[R * 72 + 9 * m + t];
    part = fma(lv, SH_XRING(sh)[9 * ((j + 8 * m + t) & (kWin - 1)) + k], part);
  }
  part += __shfl_xor(part, 8);
  part += __shfl_xor(part, 16);
  part += __shfl_xor(part, 32);
  const double xj = SH_WBUF(sh)[R * 8 + k] - part;
  if (m == 0) {
    SH_XRING(sh)[9 * (j & (kWin - 1)) + k] = xj;
    if (FAST) {
      // interior entry of Psi and its share of b = Psi^T F
      const int cx = pos.x + 1, cy = pos.y + 1, cz = (DIM == 3) ? pos.z + 1 : 0;
      const int node = cx + a.S * (cy + a.S * cz);
      if (k < D::NV) a.Psi[((size_t)cell * D::NV + k) * a.nn + node] = xj;
      double fj;
      if (a.rhs_mode == MSFEC_RHS_CELL) {
        double t = 0.0;
#pragma unroll
        for (int al = 0; al < D::NV; ++al) {
          const int zx = cx - (al & 1), zy = cy - ((al >> 1) & 1);
          const int zz = (DIM == 3) ? cz - ((al >> 2) & 1) : 0;
          t += ff[zx + a.N * (zy + a.N * zz)];
        }
        fj = t;
      } else {
        fj = a.rhs_const * D::NV;
      }
      bint = fma(xj, fj, bint);  // scaled by |z| / 2^dim at the end
    } else {
      a.X[((size_t)cell * a.n + j) * kMaxNb + k] = xj;
    }
  }
  // step to j-1 (uniform)
  if (--pos.x < 0) {
    pos.x = a.ni - 1;
    if (--pos.y < 0) {
      pos.y = a.ni - 1;
      --pos.z;
    }
  }
  wave_sync();
}

// Fast-path epilogue.  With R = -K_IB G_B the lifting right-hand sides and
// K_II = L D L^T, the element matrix psi_i^T K psi_j of the reference
// (q_basis.cc:280-293) is exactly
//   M = G_B^T K_BB G_B - R^T K_II^{-1} R = P - S,
// with S = sum_j z_j z_j^T / d_j (z = L^{-1} R, accumulated in the forward
// sweep) or, for the mixed solver, the energy form R^T X + X^T (R - K_II X).
// P, the boundary part b_B of b = Psi^T F and the boundary values of Psi
// come from q_rows_kernel; lane (i,k) holds S[i][k], lane k the interior
// part of b (before the factor |cell| / 2^dim).
template <int DIM>
__device__ __forceinline__ void fast_finish(const BandArgs& a, int cell, double sacc, double bint) {
  using D = Dirs<DIM>;
  const int lane = threadIdx.x;
  const int mi = lane >> 3, mk = lane & 7;
  const double* pb = a.Pb + (size_t)cell * kPStride;
  if (mi < D::NV && mk < D::NV)
    a.M[((size_t)cell * D::NV + mi) * D::NV + mk] = pb[mi * 8 + mk] - sacc;
  if (lane < D::NV) a.bvec[(size_t)cell * D::NV + lane] = pb[64 + lane] + bint * pb[72];
}

template <int DIM, bool FAST>
__global__ __launch_bounds__(64, 2) void q_band_kernel(BandArgs a) {
  using D = Dirs<DIM>;
  const int cell = blockIdx.x;
  const int lane = threadIdx.x;
  __shared__ BandSmem sh;
  extern __shared__ double dyn[];  // fast path: fine-cell coefficient | reaction | load
  const double* cf = dyn;
  const double* rf = dyn + a.nfc;
  const double* ff = dyn + a.nfc * ((a.reaction_mode == MSFEC_REACTION_CELL) ? 2 : 1);
  double vol = 1.0;
  if (FAST) {
    const double* cg = a.coef + (size_t)cell * a.nfc;
    for (int z = lane; z < a.nfc; z += 64) dyn[z] = cg[z];
    if (a.reaction_mode == MSFEC_REACTION_CELL) {
      const double* rg = a.reac + (size_t)cell * a.nfc;
      for (int z = lane; z < a.nfc; z += 64) dyn[a.nfc + z] = rg[z];
    }
    if (a.rhs_mode == MSFEC_RHS_CELL) {
      const double* fg = a.fcell + (size_t)cell * a.nfc;
      double* fd = dyn + a.nfc * ((a.reaction_mode == MSFEC_REACTION_CELL) ? 2 : 1);
      for (int z = lane; z < a.nfc; z += 64) fd[z] = fg[z];
    }
    const double* v = a.verts + (size_t)cell * D::NV * DIM;
    double hh[3] = {1.0, 1.0, 1.0};
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      hh[d] = (v[(D::NV - 1) * DIM + d] - v[d]) / a.N;
      vol *= hh[d];
    }
    if (lane < 8) {
      double ks, ms;
      box_tables<DIM>(lane, hh, ks, ms);
      sh.kv[lane] = ks;
      sh.mv[lane] = ms;
    }
    wave_sync();
  }

  // interior-grid offset of the lower stencil entry this lane scatters
  // (kept in a register: indexing the kernel argument by lane would reload it
  // from memory in every step)
  int my_off = 0;
#pragma unroll
  for (int t = 0; t < kMaxLow; ++t)
    if (lane == t) my_off = a.off_int[t];

  double A[8][8];
  double y[8];
#pragma unroll
  for (int r = 0; r < 8; ++r)
#pragma unroll
    for (int b = 0; b < 8; ++b) A[r][b] = 0.0;
#pragma unroll
  for (int k = 0; k < 8; ++k) y[k] = 0.0;

  // initial window: indices 0..63 in slots 0..63
  RowPrefetch pf;
  if (FAST) rows_load(a, cell, 0, pf);
  for (int G = 0; G < 8; ++G) {
    stage_group<DIM, FAST>(a, cell, 8 * G, sh, pf);
    band_insert<DIM, 0>(a, 8 * G + 0, G, my_off, A, y, sh);
    wave_sync();
    band_insert<DIM, 1>(a, 8 * G + 1, G, my_off, A, y, sh);
    wave_sync();
    band_insert<DIM, 2>(a, 8 * G + 2, G, my_off, A, y, sh);
    wave_sync();
    band_insert<DIM, 3>(a, 8 * G + 3, G, my_off, A, y, sh);
    wave_sync();
    band_insert<DIM, 4>(a, 8 * G + 4, G, my_off, A, y, sh);
    wave_sync();
    band_insert<DIM, 5>(a, 8 * G + 5, G, my_off, A, y, sh);
    wave_sync();
    band_insert<DIM, 6>(a, 8 * G + 6, G, my_off, A, y, sh);
    wave_sync();
    band_insert<DIM, 7>(a, 8 * G + 7, G, my_off, A, y, sh);
    wave_sync();
  }

  int bad = 0;
  double sacc = 0.0;
  const int n = a.n;
  for (int j0 = 0; j0 < n; j0 += 8) {
    const int G = (j0 >> 3) & 7;
    stage_group<DIM, FAST>(a, cell, j0 + kWin, sh, pf);
    if (j0 + 0 < n) band_step<DIM, 0>(a, cell, j0 + 0, G, my_off, A, y, sh, bad, sacc);
    if (j0 + 1 < n) band_step<DIM, 1>(a, cell, j0 + 1, G, my_off, A, y, sh, bad, sacc);
    if (j0 + 2 < n) band_step<DIM, 2>(a, cell, j0 + 2, G, my_off, A, y, sh, bad, sacc);
    if (j0 + 3 < n) band_step<DIM, 3>(a, cell, j0 + 3, G, my_off, A, y, sh, bad, sacc);
    if (j0 + 4 < n) band_step<DIM, 4>(a, cell, j0 + 4, G, my_off, A, y, sh, bad, sacc);
    if (j0 + 5 < n) band_step<DIM, 5>(a, cell, j0 + 5, G, my_off, A, y, sh, bad, sacc);
    if (j0 + 6 < n) band_step<DIM, 6>(a, cell, j0 + 6, G, my_off, A, y, sh, bad, sacc);
    if (j0 + 7 < n) band_step<DIM, 7>(a, cell, j0 + 7, G, my_off, A, y, sh, bad, sacc);
    wave_sync();
  }

  // backward solve x_j = w_j - sum_{o>=1} L[j+o][j] x_{j+o}; lane = (k, m):
  // RHS k = lane & 7, offsets 8m..8m+7 with m = lane >> 3.  The columns of
  // the next group of 8 steps are prefetched into registers.
  __threadfence_block();
  __syncthreads();
  for (int t = lane; t < 576; t += 64) SH_XRING(sh)[t] = 0.0;
  const double* Lc = a.Lcol + (size_t)cell * n * kWin;
  const double* Wc = a.Wf + (size_t)cell * n * kMaxNb;
  const int nchunk = (n + 7) >> 3;
  BackPos pos{a.ni - 1, a.ni - 1, (DIM == 3) ? a.ni - 1 : 0};
  double bint = 0.0;
  double lp[8], wp = 0.0;
  {
    const int jj = 8 * (nchunk - 1) + (lane >> 3);
#pragma unroll
    for (int t = 0; t < 8; ++t) lp[t] = (jj < n) ? Lc[(size_t)jj * kWin + 8 * (lane & 7) + t] : 0.0;
    wp = (jj < n) ? Wc[(size_t)jj * kMaxNb + (lane & 7)] : 0.0;
  }
  for (int cch = nchunk - 1; cch >= 0; --cch) {
    wave_sync();
#pragma unroll
    for (int t = 0; t < 8; ++t) SH_LBUF(sh)[(lane >> 3) * 72 + 9 * (lane & 7) + t] = lp[t];
    SH_WBUF(sh)[lane] = wp;
    wave_sync();
    if (cch > 0) {
      const int jj = 8 * (cch - 1) + (lane >> 3);
#pragma unroll
      for (int t = 0; t < 8; ++t) lp[t] = Lc[(size_t)jj * kWin + 8 * (lane & 7) + t];
      wp = Wc[(size_t)jj * kMaxNb + (lane & 7)];
    }
    const int jb = 8 * cch;
    if (jb + 7 < n) back_step<DIM, FAST, 7>(a, cell, jb + 7, sh, pos, bint, ff);
    if (jb + 6 < n) back_step<DIM, FAST, 6>(a, cell, jb + 6, sh, pos, bint, ff);
    if (jb + 5 < n) back_step<DIM, FAST, 5>(a, cell, jb + 5, sh, pos, bint, ff);
    if (jb + 4 < n) back_step<DIM, FAST, 4>(a, cell, jb + 4, sh, pos, bint, ff);
    if (jb + 3 < n) back_step<DIM, FAST, 3>(a, cell, jb + 3, sh, pos, bint, ff);
    if (jb + 2 < n) back_step<DIM, FAST, 2>(a, cell, jb + 2, sh, pos, bint, ff);
    if (jb + 1 < n) back_step<DIM, FAST, 1>(a, cell, jb + 1, sh, pos, bint, ff);
    back_step<DIM, FAST, 0>(a, cell, jb + 0, sh, pos, bint, ff);
  }
  int worst = bad;
  for (int o = 32; o > 0; o >>= 1) worst = max(worst, __shfl_xor(worst, o));
  if (lane == 0) a.status[cell] = worst;
  if (FAST) {
    wave_sync();
    fast_finish<DIM>(a, cell, sacc, bint);
  }
}

// ---------------------------------------------------------------------------
// K2 (MFMA variant): the same sliding-window band factorisation, eliminating
// 4 pivots per super-step with FP64 matrix-core updates.
//
// The 64-slot window is a symmetric 4x4 grid of 16x16 tiles of which the 10
// tiles (P >= Q) are kept, each as the 4-register accumulator of
// v_mfma_f64_16x16x4f64 (entry (i,j) in lane j + 16 (i mod 4), register
// i / 4).  The 4 pivots of a super-step occupy slots s0..s0+3 = rows 4g..4g+3
// of tile row T: in every tile of that row they are register g of all lanes.
// A super-step
//   1. copies the 4 pivot rows to LDS,
//   2. factors the 4x4 pivot block C C^T (every lane, redundantly),
//   3. forms the panel V = A21 C^-T, one value per lane and tile row, in the
//      A/B operand layout of the MFMA (lane i + 16k holds V[16P+i][k]),
//   4. updates the window T_PQ -= V_P V_Q^T (10 MFMAs), the right-hand sides
//      Y_P -= V_P Z with Z = C^-1 y_pivot (4 MFMAs) and the Schur term of the
//      element matrix S += Z^T Z (1 MFMA),
//   5. streams V and Z to HBM for the backward sweep and inserts the next 4
//      rows of the band into the freed slots.
// ---------------------------------------------------------------------------
typedef double mf_d4 __attribute__((ext_vector_type(4)));

struct MfmaSmem {
  double piv[4 * 72];   // pivot rows, slot x at pad8(x)
  double nrow[4 * 72];  // incoming rows, slot x at pad8(x)
  double yp[64];        // pivot right-hand sides [4][16]
};

#define MF_TILE(P, Q) ((P) * ((P) + 1) / 2 + (Q))

__device__ __forceinline__ mf_d4 mf_fma(double a, double b, mf_d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// pivot rows of the window -> LDS
template <int G>
__device__ __forceinline__ void mf_extract(const mf_d4 (&Tl)[10], int T, double* piv) {
  const int lane = threadIdx.x, lh = lane >> 4, ll = lane & 15;
#pragma unroll
  for (int P = 0; P < 4; ++P) {
#pragma unroll
    for (int Q = 0; Q <= P; ++Q) {
      if (P == T) piv[lh * 72 + pad8(16 * Q + ll)] = Tl[MF_TILE(P, Q)][G];
      if (Q == T && P > T && (ll >> 2) == G) {
        const int t = ll - 4 * G;
#pragma unroll
        for (int r = 0; r < 4; ++r) piv[t * 72 + pad8(16 * P + 4 * r + lh)] = Tl[MF_TILE(P, Q)][r];
      }
    }
  }
}

// incoming rows (LDS) -> window; rows and columns of slots s0..s0+3
template <int G>
__device__ __forceinline__ void mf_place(mf_d4 (&Tl)[10], int T, const double* nrow) {
  const int lane = threadIdx.x, lh = lane >> 4, ll = lane & 15;
#pragma unroll
  for (int P = 0; P < 4; ++P) {
#pragma unroll
    for (int Q = 0; Q <= P; ++Q) {
      if (P == T) Tl[MF_TILE(P, Q)][G] = nrow[lh * 72 + pad8(16 * Q + ll)];
      if (Q == T && (ll >> 2) == G) {
        const int t = ll - 4 * G;
#pragma unroll
        for (int r = 0; r < 4; ++r) Tl[MF_TILE(P, Q)][r] = nrow[t * 72 + pad8(16 * P + 4 * r + lh)];
      }
    }
  }
}

// Build the 4 incoming rows jbase..jbase+3 (staged as rows rb..rb+3) in LDS
// and place them, with their right-hand sides, into slots s0..s0+3.
template <int DIM, int G>
__device__ __forceinline__ void mf_insert(const BandArgs& a, int jbase, int rb, int T, int my_off,
                                          mf_d4 (&Tl)[10], mf_d4 (&Yt)[4], BandSmem& sh,
                                          MfmaSmem& ms) {
  using D = Dirs<DIM>;
  const int lane = threadIdx.x, lh = lane >> 4, ll = lane & 15;
  const int s0 = 16 * T + 4 * G;
#pragma unroll
  for (int t = 0; t < 4; ++t) ms.nrow[t * 72 + pad8(lane)] = 0.0;
  wave_sync();
  {
    const int t = lh, e = ll, jn = jbase + t;
    if (jn < a.n) {
      if (e < D::NLOW) {
        const double v = sh.stg_low[rb + t][e];
        if (v != 0.0) {
          ms.nrow[t * 72 + pad8((s0 + t - my_off) & (kWin - 1))] = v;
          // couplings between two incoming rows are needed in both rows
          if (e > 0 && my_off <= t) ms.nrow[(t - my_off) * 72 + pad8(s0 + t)] = v;
        }
      }
    } else if (e == 0) {
      ms.nrow[t * 72 + pad8(s0 + t)] = 1.0;  // past the end: isolated unit diagonal
    }
  }
  wave_sync();
  mf_place<G>(Tl, T, ms.nrow);
  const double yv = (ll < 8 && jbase + lh < a.n) ? sh.stg_rhs[rb + lh][ll & 7] : 0.0;
#pragma unroll
  for (int P = 0; P < 4; ++P)
    if (P == T) Yt[P][G] = yv;
  wave_sync();
}

template <int DIM, int G>
__device__ __forceinline__ void mf_superstep(const BandArgs& a, int cell, int j0, int T,
                                             mf_d4 (&Tl)[10], mf_d4 (&Yt)[4], mf_d4& Sacc,
                                             BandSmem& sh, MfmaSmem& ms, int& bad) {
  const int lane = threadIdx.x, lh = lane >> 4, ll = lane & 15;
  const int s0 = 16 * T + 4 * G;
  mf_extract<G>(Tl, T, ms.piv);
  // pivot right-hand sides: register G of the right-hand-side tile T
  {
    double yv = 0.0;
#pragma unroll
    for (int P = 0; P < 4; ++P)
      if (P == T) yv = Yt[P][G];
    ms.yp[lh * 16 + ll] = yv;
  }
  wave_sync();
  // 4x4 Cholesky of the pivot block and its inverse (all lanes)
  double a00 = ms.piv[0 * 72 + pad8(s0 + 0)];
  double a10 = ms.piv[1 * 72 + pad8(s0 + 0)], a11 = ms.piv[1 * 72 + pad8(s0 + 1)];
  double a20 = ms.piv[2 * 72 + pad8(s0 + 0)], a21 = ms.piv[2 * 72 + pad8(s0 + 1)];
  double a22 = ms.piv[2 * 72 + pad8(s0 + 2)];
  double a30 = ms.piv[3 * 72 + pad8(s0 + 0)], a31 = ms.piv[3 * 72 + pad8(s0 + 1)];
  double a32 = ms.piv[3 * 72 + pad8(s0 + 2)], a33 = ms.piv[3 * 72 + pad8(s0 + 3)];
  if (!(a00 > 0.0) && bad == 0) bad = (a00 == a00) ? (int)MSFEC_CELL_NOT_SPD : (int)MSFEC_CELL_NONFINITE;
  const double c00 = sqrt(a00), i00 = 1.0 / c00;
  const double c10 = a10 * i00, c20 = a20 * i00, c30 = a30 * i00;
  const double d1 = a11 - c10 * c10;
  if (!(d1 > 0.0) && bad == 0) bad = (d1 == d1) ? (int)MSFEC_CELL_NOT_SPD : (int)MSFEC_CELL_NONFINITE;
  const double c11 = sqrt(d1), i11 = 1.0 / c11;
  const double c21 = (a21 - c20 * c10) * i11, c31 = (a31 - c30 * c10) * i11;
  const double d2 = a22 - c20 * c20 - c21 * c21;
  if (!(d2 > 0.0) && bad == 0) bad = (d2 == d2) ? (int)MSFEC_CELL_NOT_SPD : (int)MSFEC_CELL_NONFINITE;
  const double c22 = sqrt(d2), i22 = 1.0 / c22;
  const double c32 = (a32 - c30 * c20 - c31 * c21) * i22;
  const double d3 = a33 - c30 * c30 - c31 * c31 - c32 * c32;
  if (!(d3 > 0.0) && bad == 0) bad = (d3 == d3) ? (int)MSFEC_CELL_NOT_SPD : (int)MSFEC_CELL_NONFINITE;
  const double c33 = sqrt(d3), i33 = 1.0 / c33;
  // C^-1 (lower): row k of the inverse
  const double i10 = -c10 * i00 * i11;
  const double i21 = -c21 * i11 * i22;
  const double i20 = -(c20 * i00 + c21 * i10) * i22;
  const double i32 = -c32 * i22 * i33;
  const double i31 = -(c31 * i11 + c32 * i21) * i33;
  const double i30 = -(c30 * i00 + c31 * i10 + c32 * i20) * i33;
  double ck0, ck1, ck2, ck3;  // row lh of C^-1
  if (lh == 0) {
    ck0 = i00; ck1 = 0.0; ck2 = 0.0; ck3 = 0.0;
  } else if (lh == 1) {
    ck0 = i10; ck1 = i11; ck2 = 0.0; ck3 = 0.0;
  } else if (lh == 2) {
    ck0 = i20; ck1 = i21; ck2 = i22; ck3 = 0.0;
  } else {
    ck0 = i30; ck1 = i31; ck2 = i32; ck3 = i33;
  }
  // panel V = A21 C^-T in operand layout, and Z = C^-1 y_pivot
  double V[4], nV[4];
#pragma unroll
  for (int P = 0; P < 4; ++P) {
    const int x = pad8(16 * P + ll);
    V[P] = ms.piv[0 * 72 + x] * ck0 + ms.piv[1 * 72 + x] * ck1 + ms.piv[2 * 72 + x] * ck2 +
           ms.piv[3 * 72 + x] * ck3;
    nV[P] = -V[P];
  }
  const double Z = ms.yp[0 * 16 + ll] * ck0 + ms.yp[1 * 16 + ll] * ck1 + ms.yp[2 * 16 + ll] * ck2 +
                   ms.yp[3 * 16 + ll] * ck3;
  // window, right-hand sides and Schur term
#pragma unroll
  for (int P = 0; P < 4; ++P)
#pragma unroll
    for (int Q = 0; Q <= P; ++Q) Tl[MF_TILE(P, Q)] = mf_fma(nV[P], V[Q], Tl[MF_TILE(P, Q)]);
#pragma unroll
  for (int P = 0; P < 4; ++P) Yt[P] = mf_fma(nV[P], Z, Yt[P]);
  Sacc = mf_fma(Z, Z, Sacc);
  // stream the panel (rows relative to s0; rows 0..3 are C) and Z to HBM
  const int qs = j0 >> 2;
  double* lst = a.Lcol + ((size_t)cell * a.nq + qs) * 256;
#pragma unroll
  for (int P = 0; P < 4; ++P) lst[((16 * P + ll - s0) & (kWin - 1)) * 4 + lh] = V[P];
  if (ll < 8) a.Wf[((size_t)cell * a.nq + qs) * 32 + lh * 8 + ll] = Z;
  wave_sync();
}

// Backward block step: x_blk = C^-T (z - V_rest^T x_window) for the 4 pivots
// of super-step qs; lane (k = lane>>4, r = (lane>>1)&7, h = lane&1) sums half
// of the 60 terms of (V_rest^T x)[k][r].
template <int DIM, bool FAST>
__device__ __forceinline__ void mf_back(const BandArgs& a, int cell, int qs, const double* vb,
                                        const double* zb, double* xr, double* tb, double& bint,
                                        const double* ff) {
  using D = Dirs<DIM>;
  const int lane = threadIdx.x;
  const int k = lane >> 4, r = (lane >> 1) & 7, h = lane & 1;
  const int j0 = 4 * qs;
  double part = 0.0;
#pragma unroll 6
  for (int m = 0; m < 30; ++m) {
    const int o = 4 + h + 2 * m;
    part = fma(vb[o * 5 + k], xr[9 * ((j0 + o) & (kWin - 1)) + r], part);
  }
  part += __shfl_xor(part, 1);
  if (h == 0) tb[k * 8 + r] = part;
  wave_sync();
  if (lane < 8) {
    const int rr = lane;
    const double c00 = vb[0 * 5 + 0], c10 = vb[1 * 5 + 0], c11 = vb[1 * 5 + 1];
    const double c20 = vb[2 * 5 + 0], c21 = vb[2 * 5 + 1], c22 = vb[2 * 5 + 2];
    const double c30 = vb[3 * 5 + 0], c31 = vb[3 * 5 + 1], c32 = vb[3 * 5 + 2], c33 = vb[3 * 5 + 3];
    const double r0 = zb[0 * 8 + rr] - tb[0 * 8 + rr], r1 = zb[1 * 8 + rr] - tb[1 * 8 + rr];
    const double r2 = zb[2 * 8 + rr] - tb[2 * 8 + rr], r3 = zb[3 * 8 + rr] - tb[3 * 8 + rr];
    const double x3 = r3 / c33;
    const double x2 = (r2 - c32 * x3) / c22;
    const double x1 = (r1 - c21 * x2 - c31 * x3) / c11;
    const double x0 = (r0 - c10 * x1 - c20 * x2 - c30 * x3) / c00;
    const double xs[4] = {x0, x1, x2, x3};
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int j = j0 + t;
      if (j >= a.n) continue;
      xr[9 * (j & (kWin - 1)) + rr] = xs[t];
      if (FAST) {
        int c[3] = {0, 0, 0}, tt = j;
#pragma unroll
        for (int d = 0; d < DIM; ++d) {
          c[d] = tt % a.ni + 1;
          tt /= a.ni;
        }
        const int node = c[0] + a.S * (c[1] + a.S * c[2]);
        if (rr < D::NV) a.Psi[((size_t)cell * D::NV + rr) * a.nn + node] = xs[t];
        double fj;
        if (a.rhs_mode == MSFEC_RHS_CELL) {
          double s = 0.0;
#pragma unroll
          for (int al = 0; al < D::NV; ++al) {
            const int zx = c[0] - (al & 1), zy = c[1] - ((al >> 1) & 1);
            const int zz = (DIM == 3) ? c[2] - ((al >> 2) & 1) : 0;
            s += ff[zx + a.N * (zy + a.N * zz)];
          }
          fj = s;
        } else {
          fj = a.rhs_const * D::NV;
        }
        bint = fma(xs[t], fj, bint);
      } else {
        a.X[((size_t)cell * a.n + j) * kMaxNb + rr] = xs[t];
      }
    }
  }
  wave_sync();
}

template <int DIM, bool FAST>
__global__ __launch_bounds__(64, 2) void q_band_mfma_kernel(BandArgs a) {
  using D = Dirs<DIM>;
  const int cell = blockIdx.x;
  const int lane = threadIdx.x;
  __shared__ BandSmem sh;
  __shared__ MfmaSmem ms;
  extern __shared__ double dyn[];  // fast path: fine-cell coefficient | reaction | load
  const double* cf = dyn;
  const double* rf = dyn + a.nfc;
  const double* ff = dyn + a.nfc * ((a.reaction_mode == MSFEC_REACTION_CELL) ? 2 : 1);
  double vol = 1.0;
  if (FAST) {
    const double* cg = a.coef + (size_t)cell * a.nfc;
    for (int z = lane; z < a.nfc; z += 64) dyn[z] = cg[z];
    if (a.reaction_mode == MSFEC_REACTION_CELL) {
      const double* rg = a.reac + (size_t)cell * a.nfc;
      for (int z = lane; z < a.nfc; z += 64) dyn[a.nfc + z] = rg[z];
    }
    if (a.rhs_mode == MSFEC_RHS_CELL) {
      const double* fg = a.fcell + (size_t)cell * a.nfc;
      double* fd = dyn + a.nfc * ((a.reaction_mode == MSFEC_REACTION_CELL) ? 2 : 1);
      for (int z = lane; z < a.nfc; z += 64) fd[z] = fg[z];
    }
    const double* v = a.verts + (size_t)cell * D::NV * DIM;
    double hh[3] = {1.0, 1.0, 1.0};
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      hh[d] = (v[(D::NV - 1) * DIM + d] - v[d]) / a.N;
      vol *= hh[d];
    }
    if (lane < 8) {
      const int x = lane;
      double ks = 0.0, msv = 1.0;
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        double t = (((x >> d) & 1) ? -1.0 : 1.0) / hh[d];
#pragma unroll
        for (int e = 0; e < DIM; ++e)
          if (e != d) t *= (((x >> e) & 1) ? (1.0 / 6.0) : (1.0 / 3.0)) * hh[e];
        ks += t;
        msv *= (((x >> d) & 1) ? (1.0 / 6.0) : (1.0 / 3.0)) * hh[d];
      }
      sh.kv[x] = ks;
      sh.mv[x] = msv;
    }
    wave_sync();
  }
  // interior-grid offset of the stencil entry scattered by this lane (lane & 15)
  int my_off = 0;
#pragma unroll
  for (int t = 0; t < kMaxLow; ++t)
    if ((lane & 15) == t) my_off = a.off_int[t];

  mf_d4 Tl[10];
  mf_d4 Yt[4];
  mf_d4 Sacc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int t = 0; t < 10; ++t) Tl[t] = (mf_d4){0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int P = 0; P < 4; ++P) Yt[P] = (mf_d4){0.0, 0.0, 0.0, 0.0};

  // initial window: indices 0..63 in slots 0..63
  RowPrefetch pf;
  if (FAST) rows_load(a, cell, 0, pf);
  for (int T = 0; T < 4; ++T) {
    stage_group<DIM, FAST>(a, cell, 16 * T, sh, pf);
    mf_insert<DIM, 0>(a, 16 * T + 0, 0, T, my_off, Tl, Yt, sh, ms);
    mf_insert<DIM, 1>(a, 16 * T + 4, 4, T, my_off, Tl, Yt, sh, ms);
    stage_group<DIM, FAST>(a, cell, 16 * T + 8, sh, pf);
    mf_insert<DIM, 2>(a, 16 * T + 8, 0, T, my_off, Tl, Yt, sh, ms);
    mf_insert<DIM, 3>(a, 16 * T + 12, 4, T, my_off, Tl, Yt, sh, ms);
  }

  int bad = 0;
  const int n = a.n;
  for (int j0 = 0; j0 < n; j0 += 16) {
    const int T = (j0 >> 4) & 3;
    stage_group<DIM, FAST>(a, cell, j0 + kWin, sh, pf);
    if (j0 < n) {
      mf_superstep<DIM, 0>(a, cell, j0, T, Tl, Yt, Sacc, sh, ms, bad);
      mf_insert<DIM, 0>(a, j0 + kWin, 0, T, my_off, Tl, Yt, sh, ms);
    }
    if (j0 + 4 < n) {
      mf_superstep<DIM, 1>(a, cell, j0 + 4, T, Tl, Yt, Sacc, sh, ms, bad);
      mf_insert<DIM, 1>(a, j0 + 4 + kWin, 4, T, my_off, Tl, Yt, sh, ms);
    }
    stage_group<DIM, FAST>(a, cell, j0 + 8 + kWin, sh, pf);
    if (j0 + 8 < n) {
      mf_superstep<DIM, 2>(a, cell, j0 + 8, T, Tl, Yt, Sacc, sh, ms, bad);
      mf_insert<DIM, 2>(a, j0 + 8 + kWin, 0, T, my_off, Tl, Yt, sh, ms);
    }
    if (j0 + 12 < n) {
      mf_superstep<DIM, 3>(a, cell, j0 + 12, T, Tl, Yt, Sacc, sh, ms, bad);
      mf_insert<DIM, 3>(a, j0 + 12 + kWin, 4, T, my_off, Tl, Yt, sh, ms);
    }
  }

  // Schur term of the element matrix: entry (i,j) of Sacc sits in lane
  // j + 16 (i mod 4), register i / 4; lane (i,k) of the epilogue needs (i,k)
  double sacc = 0.0;
  {
    double* sb = SH_PART(sh);
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      const int i = 4 * rr + (lane >> 4), jj = lane & 15;
      if (i < 8 && jj < 8) sb[i * 8 + jj] = Sacc[rr];
    }
    wave_sync();
    sacc = sb[lane];
    wave_sync();
  }

  // backward sweep over the super-steps, panels prefetched one step ahead
  __threadfence_block();
  __syncthreads();
  double* xr = SH_XRING(sh);
  double* vb = SH_LBUF(sh);        // [64][5]
  double* zb = SH_LBUF(sh) + 320;  // [4][8]
  double* tb = SH_LBUF(sh) + 352;  // [4][8]
  for (int t = lane; t < 576; t += 64) xr[t] = 0.0;
  const int nq = (n + 3) >> 2;
  const double* Lc = a.Lcol + (size_t)cell * a.nq * 256;
  const double* Zc = a.Wf + (size_t)cell * a.nq * 32;
  double vp[4], zp = 0.0;
#pragma unroll
  for (int t = 0; t < 4; ++t) vp[t] = Lc[(size_t)(nq - 1) * 256 + lane * 4 + t];
  if (lane < 32) zp = Zc[(size_t)(nq - 1) * 32 + lane];
  double bint = 0.0;
  for (int qs = nq - 1; qs >= 0; --qs) {
    wave_sync();
#pragma unroll
    for (int t = 0; t < 4; ++t) vb[lane * 5 + t] = vp[t];
    if (lane < 32) zb[lane] = zp;
    wave_sync();
    if (qs > 0) {
#pragma unroll
      for (int t = 0; t < 4; ++t) vp[t] = Lc[(size_t)(qs - 1) * 256 + lane * 4 + t];
      if (lane < 32) zp = Zc[(size_t)(qs - 1) * 32 + lane];
    }
    mf_back<DIM, FAST>(a, cell, qs, vb, zb, xr, tb, bint, ff);
  }
  int worst = bad;
  for (int o = 32; o > 0; o >>= 1) worst = max(worst, __shfl_xor(worst, o));
  if (lane == 0) a.status[cell] = worst;
  if (FAST) {
    wave_sync();
    // bint lives in lanes 0..7 (RHS r = lane); the epilogue reads it from lane k
    fast_finish<DIM>(a, cell, sacc, bint);
  }
}

// ---------------------------------------------------------------------------
// K2m: the band factorisation on the matrix cores with the tile rows of the
// window unrolled at compile time, for FP64 tiles (solver BAND_FP64_MFMA) or
// FP32 tiles (the factor of the mixed-precision solver).
//
// Layout.  The window is the lower half of a 4x4 grid of 16x16 tiles held as
// MFMA accumulators.  Window slot x = (row index mod 64) sits in tile row
// x / 16.  Inside a tile, accumulator register g of lane l holds row slot
// 4g + (l >> 4) and column slot cslot(l & 15); the same lane, as an A/B
// operand, holds row slot cslot(l & 15) and pivot k = l >> 4.  cslot is the
// identity for v_mfma_f64_16x16x4f64 (accumulator (i,j) in lane
// j + 16 (i mod 4), register i / 4) and 4 (x mod 4) + x / 4 for
// v_mfma_f32_16x16x4f32 (accumulator (i,j) in lane j + 16 (i / 4), register
// i mod 4): with it, the 4 pivots of a super-step are one register of every
// tile in both precisions.  Tile rows are template parameters, so every tile
// access is a fixed register.
//
// A super-step (tile row T, pivots 16T + 4G .. +3) copies the pivot rows to
// LDS, factors the 4x4 pivot block C C^T in FP64 (all lanes), forms the panel
// V = A21 C^-T in operand layout, updates the 10 window tiles, the 4
// right-hand-side tiles and the Schur term with 15 MFMAs, streams the panel
// (with C^-1 in place of its first 4 rows) and Z = C^-1 y to HBM, and
// inserts the next 4 rows of the band.  The backward sweep reads the panels
// back: x_blk = C^-T (z - sum_{o>=4} V[o]^T x[j0 + o]).
// ---------------------------------------------------------------------------
template <typename R>
struct Mf;
template <>
struct Mf<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // c - a b: the FP64 instruction negates its first operand itself (the
  // last modifier field is neg:[A,B,C] for this instruction)
  static __device__ __forceinline__ v4 mma_neg(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 1);
  }
  static __device__ __forceinline__ int cslot(int ll) { return ll; }
  // lane holding (row slot t, column slot cs) of a register of a tile
  static constexpr int lane_of(int t, int cs) { return 16 * t + cs; }
  // physical accumulator row of register g in lane group t
  static __device__ __forceinline__ int acc_row(int g, int t) { return 4 * g + t; }
};
template <>
struct Mf<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mma(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ v4 mma_neg(float a, float b, v4 c) { return mma(-a, b, c); }
  static __device__ __forceinline__ int cslot(int ll) { return 4 * (ll & 3) + (ll >> 2); }
  static constexpr int lane_of(int t, int cs) { return 16 * t + 4 * (cs & 3) + (cs >> 2); }
  static __device__ __forceinline__ int acc_row(int g, int t) { return 4 * t + g; }
};

// backward-sweep strides: panel rows padded to 68 (zeros at 64..67), x ring
// rows of 132 (copy at +64, zeros at 128..131)
constexpr int kPanRow = 68;
constexpr int kRing = 132;

// Row stride of the pivot and incoming rows in LDS (slot x of row t at
// t * stride + x).  A wave touches 16 consecutive slots of each of 4 rows
// (row parts) or one slot of each of 4 rows (column parts): FP64 rows 68
// apart (4 mod 16) and FP32 rows 80 apart (16 mod 32) keep both
// conflict-free.
template <typename R>
constexpr int bk_row() {
  return sizeof(R) == 8 ? 68 : 80;
}

template <typename R>
struct BkSmem {
  R piv[4 * 80];   // forward: pivot rows; backward: panel [4][68]
  R nrow[4 * 80 + 64];  // forward: incoming rows, then one spare slot per lane
  R yp[64];        // forward: pivot right-hand sides [4][16]
};

// What the backward sweep produces.
enum { BK_FINAL64 = 0, BK_X0 = 1, BK_REFINE = 2 };

// The row index as an opaque scalar: slot computations such as
// (row - offset) mod 64 stay inside the super-step instead of being hoisted
// out of the loop as one register per unrolled super-step.
__device__ __forceinline__ int bk_opaque(int x) {
  asm volatile("" : "+s"(x));
  return x;
}

// The lane index as an opaque value: the per-lane LDS addresses of a
// super-step are recomputed in it (a few integer operations) instead of
// being hoisted out of the loop as dozens of live registers.
__device__ __forceinline__ int bk_lane() {
  int x = threadIdx.x;
  asm volatile("" : "+v"(x));
  return x;
}

// Per-lane LDS offsets (in elements) of the super-step accesses, computed
// once per kernel: every access is one of these plus a compile-time constant
// (W = bk_row<R>()).
struct BkAddr {
  int lane, lh, ll, cs;
  int row;  // lh * W + cs: row part of tile (T, Q) at + 16 Q
  int col;  // (cs & 3) * W + lh: column part of tile (P, T) at + 16 P + 4 r
  int vrd;  // cs: panel reads at + W t + 16 P
  int zr;   // lane: zeroing of the incoming rows
  bool gcol[4];  // (cs >> 2) == G: this lane holds pivot column cs & 3 of group G
};

template <typename R>
__device__ __forceinline__ BkAddr bk_addr() {
  BkAddr ad;
  ad.lane = threadIdx.x;
  ad.lh = ad.lane >> 4;
  ad.ll = ad.lane & 15;
  ad.cs = Mf<R>::cslot(ad.ll);
  constexpr int W = bk_row<R>();
  ad.row = ad.lh * W + ad.cs;
  ad.col = (ad.cs & 3) * W + ad.lh;
  ad.vrd = ad.cs;
  ad.zr = ad.lane;
#pragma unroll
  for (int g = 0; g < 4; ++g) ad.gcol[g] = (ad.cs >> 2) == g;
  return ad;
}

template <typename R, int NT, int T, int G>
__device__ __forceinline__ void bk_extract(const BkAddr& ad, const typename Mf<R>::v4 (&Tl)[10],
                                           R* piv) {
#pragma unroll
  for (int Q = 0; Q <= T; ++Q) piv[ad.row + 16 * Q] = Tl[MF_TILE(T, Q)][G];
  if (ad.gcol[G]) {
#pragma unroll
    for (int P = T + 1; P < NT; ++P)
#pragma unroll
      for (int r = 0; r < 4; ++r) piv[ad.col + 16 * P + 4 * r] = Tl[MF_TILE(P, T)][r];
  }
}

// Insert the rows jbase..jbase+3 (staged as rows rb..rb+3) into slots
// 16T + 4G .. +3 in two parts.  The scatter writes their nonzeros into LDS
// in the slot layout: lane (t = lh, e = ll) places lower entry e of row
// jbase + t, and a coupling between two incoming rows also in the earlier
// row.  It has no branch: a lane with nothing to place writes into a slot
// of its own in the pool (unused during the elimination).  The placement
// then copies rows and columns into the window tiles.
template <typename R>
__device__ __forceinline__ double bk_scatter_load(const BkAddr& ad, int rb, BandSmem& sh,
                                                  BkSmem<R>& ms) {
#pragma unroll
  for (int t = 0; t < 4; ++t) ms.nrow[ad.zr + bk_row<R>() * t] = (R)0;
  return sh.stg_low[rb + ad.lh][ad.ll];
}

template <typename R, int DIM, int NT, int T, int G, bool LDL>
__device__ __forceinline__ void bk_scatter_store(const BandArgs& a, const BkAddr& ad, int jbase,
                                                 int rb, int my_off, double v, BandSmem& sh,
                                                 BkSmem<R>& ms) {
  using D = Dirs<DIM>;
  constexpr int nlow = LDL ? 16 : D::NLOW;
  constexpr int s0 = 16 * T + 4 * G;
  // lane-dependent addresses from the opaque lane index: recomputed here,
  // not hoisted out of the loop as one register per super-step
  const int lane = bk_lane();
  const int t = lane >> 4, e = lane & 15;
  if (LDL) my_off = sh.stg_off[rb + t][e];  // the offset of this row's entry e
  const bool live = jbase + t < a.n;
  // zeros stay unwritten: on a grid with one interior node per direction
  // some offsets coincide (or are 0), and their couplings are all zero
  const bool w1 = live && e < nlow && v != 0.0;
  const bool w2 = w1 && e > 0 && my_off <= t;
  const bool w3 = !live && e == 0;  // past the end: isolated unit diagonal
  constexpr int W = bk_row<R>();
  // Both stores go to ms.nrow at a selected index, a lane with nothing to
  // place writing its own spare slot past the rows: one address select and
  // one LDS write each, where pointers into two arrays became exec-mask
  // branches around every write.
  const int spare = 4 * 80 + lane;
  // slot of column jn - off (jbase = s0 mod 64)
  const int i1 = w1 ? t * W + ((bk_opaque(jbase) + t - my_off) & (16 * NT - 1))
                    : (w3 ? t * W + s0 + t : spare);
  ms.nrow[i1] = w1 ? (R)v : (R)1;
  ms.nrow[w2 ? max(t - my_off, 0) * W + s0 + t : spare] = (R)v;
}

template <typename R, int DIM, int NT, int T, int G>
__device__ __forceinline__ void bk_place(const BandArgs& a, const BkAddr& ad, int jbase, int rb,
                                         typename Mf<R>::v4 (&Tl)[10],
                                         typename Mf<R>::v4 (&Yt)[4], BandSmem& sh,
                                         BkSmem<R>& ms) {
#pragma unroll
  for (int Q = 0; Q <= T; ++Q) Tl[MF_TILE(T, Q)][G] = ms.nrow[ad.row + 16 * Q];
  if (ad.gcol[G]) {
#pragma unroll
    for (int P = T; P < NT; ++P)
#pragma unroll
      for (int r = 0; r < 4; ++r) Tl[MF_TILE(P, T)][r] = ms.nrow[ad.col + 16 * P + 4 * r];
  }
  Yt[T][G] = (ad.ll < 8 && jbase + ad.lh < a.n) ? (R)sh.stg_rhs[rb + ad.lh][ad.ll & 7] : (R)0;
}

template <typename R, int DIM, int NT, int T, int G, bool LDL>
__device__ __forceinline__ void bk_insert(const BandArgs& a, const BkAddr& ad, int jbase, int rb,
                                          int my_off, typename Mf<R>::v4 (&Tl)[10],
                                          typename Mf<R>::v4 (&Yt)[4], BandSmem& sh,
                                          BkSmem<R>& ms) {
  const double v = bk_scatter_load<R>(ad, rb, sh, ms);
  bk_scatter_store<R, DIM, NT, T, G, LDL>(a, ad, jbase, rb, my_off, v, sh, ms);
  wave_sync();
  bk_place<R, DIM, NT, T, G>(a, ad, jbase, rb, Tl, Yt, sh, ms);
  wave_sync();
}

// A value of one lane, read into every lane.
__device__ __forceinline__ double bk_readlane(double v, int l) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ float bk_readlane(float v, int l) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}

// 1/sqrt: the hardware estimate, refined by two Newton steps in FP64 (one
// step is not enough for FP64; the FP32 estimate is already within 1 ulp).
__device__ __forceinline__ double bk_rsqrt(double x) {
  double y = __builtin_amdgcn_rsq(x);
  y = y * fma(-0.5 * x * y, y, 1.5);
  y = y * fma(-0.5 * x * y, y, 1.5);
  return y;
}
__device__ __forceinline__ float bk_rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }

// Pivot monitor: the smallest pivot (NaN ignored) and the sum of all pivots.
// A non-positive pivot turns every later one into NaN, and with finite
// entries the first failing pivot is a number <= 0, so dmin <= 0 marks a
// non-SPD cell; otherwise a non-finite sum marks non-finite entries.
template <typename R>
struct BkMon {
  R dmin, dsum;
  __device__ __forceinline__ void init() {
    dmin = (R)1;
    dsum = (R)0;
  }
  __device__ __forceinline__ int status() const {
    if (!(dmin > (R)0)) return MSFEC_CELL_NOT_SPD;
    return isfinite(dsum) ? (int)MSFEC_CELL_OK : (int)MSFEC_CELL_NONFINITE;
  }
};

template <typename R>
__device__ __forceinline__ void bk_check(R d, BkMon<R>& mon) {
  mon.dmin = fmin(mon.dmin, d);
  mon.dsum += d;
}

// LDL^T pivots of an indefinite matrix may have either sign: a zero pivot
// (singular leading block) is what marks the cell.
template <typename R>
__device__ __forceinline__ void bk_check_ldl(R d, BkMon<R>& mon) {
  mon.dmin = fmin(mon.dmin, fabs(d));
  mon.dsum += d;
}

// 1/x: the hardware estimate refined by two Newton steps.
__device__ __forceinline__ double bk_rcp(double x) {
  double y = __builtin_amdgcn_rcp(x);
  y = fma(y, fma(-x, y, 1.0), y);
  y = fma(y, fma(-x, y, 1.0), y);
  return y;
}

// Order of the 15 updates of a super-step (tile row T, group G): first the
// tiles and the right-hand-side tile that the insertion of the next rows
// overwrites (row and column T), then, at the last group of the tile row,
// those that the next super-step extracts (row and column T + 1), then the
// rest, so that the next steps wait for as few matrix-core results as possible.
template <int NT, int T, int G>
__device__ constexpr int bk_prio(int P, int Q) {
  return (P == T || Q == T) ? 0 : (G == 3 && (P == ((T + 1) % NT) || Q == ((T + 1) % NT))) ? 1 : 2;
}

template <typename R, int DIM, int NT, int T, int G, bool LDL>
__device__ __forceinline__ void bk_superstep(const BandArgs& a, const BkAddr& ad, int cell, int j0,
                                             typename Mf<R>::v4 (&Tl)[10],
                                             typename Mf<R>::v4 (&Yt)[4],
                                             typename Mf<R>::v4& Sacc, BkSmem<R>& ms, BkMon<R>& bad,
                                             const R (&ek)[4]) {
  typedef Mf<R> M;
  const int lh = ad.lh, ll = ad.ll, cs = ad.cs;
  constexpr int s0 = 16 * T + 4 * G;
  bk_extract<R, NT, T, G>(ad, Tl, ms.piv);
  ms.yp[ad.lane] = Yt[T][G];
  wave_sync();
  const R* pv = ms.piv;
  // the pivot block (every lane), read from the lanes of the diagonal tile
  // that hold it
  const R pk = Tl[MF_TILE(T, T)][G];
  auto pe = [&](int t, int u) { return bk_readlane(pk, M::lane_of(t, 4 * G + u)); };
  R x0, x1, x2, x3;  // row lh of the inverse of the block's (unit) lower factor
  R xd, sc;          // panel diagonal entry of pivot lh, scale of the left operand
  if constexpr (!LDL) {
    // C C^T = pivot block; ci_t = 1 / C[t][t]
    const R a00 = pe(0, 0);
    bk_check(a00, bad);
    const R ci0 = bk_rsqrt(a00);
    const R c10 = pe(1, 0) * ci0, c20 = pe(2, 0) * ci0;
    const R c30 = pe(3, 0) * ci0;
    const R d1 = pe(1, 1) - c10 * c10;
    bk_check(d1, bad);
    const R ci1 = bk_rsqrt(d1);
    const R c21 = (pe(2, 1) - c20 * c10) * ci1;
    const R c31 = (pe(3, 1) - c30 * c10) * ci1;
    const R d2 = pe(2, 2) - c20 * c20 - c21 * c21;
    bk_check(d2, bad);
    const R ci2 = bk_rsqrt(d2);
    const R c32 = (pe(3, 2) - c30 * c20 - c31 * c21) * ci2;
    const R d3 = pe(3, 3) - c30 * c30 - c31 * c31 - c32 * c32;
    bk_check(d3, bad);
    const R ci3 = bk_rsqrt(d3);
    // row lh of C^-1: x C = e_lh solved from the last column backwards
    x3 = ek[3] * ci3;
    x2 = (ek[2] - x3 * c32) * ci2;
    x1 = (ek[1] - x2 * c21 - x3 * c31) * ci1;
    x0 = (ek[0] - x1 * c10 - x2 * c20 - x3 * c30) * ci0;
    xd = (lh == 0) ? x0 : (lh == 1) ? x1 : (lh == 2) ? x2 : x3;  // C^-1[lh][lh]
    sc = (R)1;
  } else {
    // L1 D1 L1^T = pivot block; u_ij = l_ij d_j, di_t = 1 / d_t
    const R d0 = pe(0, 0);
    bk_check_ldl(d0, bad);
    const R di0 = bk_rcp(d0);
    const R u10 = pe(1, 0), u20 = pe(2, 0), u30 = pe(3, 0);
    const R l10 = u10 * di0, l20 = u20 * di0, l30 = u30 * di0;
    const R d1 = pe(1, 1) - l10 * u10;
    bk_check_ldl(d1, bad);
    const R di1 = bk_rcp(d1);
    const R u21 = pe(2, 1) - l20 * u10, u31 = pe(3, 1) - l30 * u10;
    const R l21 = u21 * di1, l31 = u31 * di1;
    const R d2 = pe(2, 2) - l20 * u20 - l21 * u21;
    bk_check_ldl(d2, bad);
    const R di2 = bk_rcp(d2);
    const R u32 = pe(3, 2) - l30 * u20 - l31 * u21;
    const R l32 = u32 * di2;
    const R d3 = pe(3, 3) - l30 * u30 - l31 * u31 - l32 * u32;
    bk_check_ldl(d3, bad);
    const R di3 = bk_rcp(d3);
    // row lh of L1^-1: x L1 = e_lh
    x3 = ek[3];
    x2 = ek[2] - x3 * l32;
    x1 = ek[1] - x2 * l21 - x3 * l31;
    x0 = ek[0] - x1 * l10 - x2 * l20 - x3 * l30;
    xd = (lh == 0) ? di0 : (lh == 1) ? di1 : (lh == 2) ? di2 : di3;  // 1 / d_lh
    sc = xd;
  }
  // panel in operand layout: lane (row slot 16P + cs, pivot lh).  Cholesky:
  // V = A21 C^-T, update -V V^T.  LDL^T: V = A21 L1^-T, update -V D1^-1 V^T.
  R V[4], aV[4];  // aV: the left operand, V D1^-1 for LDL^T (negated by the instruction)
#pragma unroll
  for (int P = 0; P < NT; ++P) {
    constexpr int W = bk_row<R>();
    const int x = ad.vrd + 16 * P;
    V[P] = pv[0 * W + x] * x0 + pv[1 * W + x] * x1 + pv[2 * W + x] * x2 + pv[3 * W + x] * x3;
    aV[P] = LDL ? V[P] * sc : V[P];
  }
  const R Z = ms.yp[0 * 16 + ll] * x0 + ms.yp[1 * 16 + ll] * x1 + ms.yp[2 * 16 + ll] * x2 +
              ms.yp[3 * 16 + ll] * x3;
#pragma unroll
  for (int pr = 0; pr < 3; ++pr) {
#pragma unroll
    for (int P = 0; P < NT; ++P) {
#pragma unroll
      for (int Q = 0; Q <= P; ++Q)
        if (bk_prio<NT, T, G>(P, Q) == pr)
          Tl[MF_TILE(P, Q)] = M::mma_neg(aV[P], V[Q], Tl[MF_TILE(P, Q)]);
      if (bk_prio<NT, T, G>(P, P) == pr) Yt[P] = M::mma_neg(aV[P], Z, Yt[P]);
    }
  }
  if (!LDL) Sacc = M::mma(Z, Z, Sacc);
  // panel [k][o], o = offset from the first pivot.  Offsets 0..3 hold C
  // (lower part) with 1 / C[t][t] on the diagonal; for LDL^T, L1 D1 with
  // 1 / d_t on the diagonal: the backward sweep is the same for both.
  const int qs = j0 >> 2;
  R* pst = reinterpret_cast<R*>(a.pan) + ((size_t)cell * a.nq + qs) * 256;
  const int jr = bk_opaque(j0);  // = s0 mod 64
#pragma unroll
  for (int P = 0; P < NT; ++P) {
    const int o = (16 * P + cs - jr) & (16 * NT - 1);
    pst[lh * 64 + o] = (P == T && o == lh) ? xd : V[P];
  }
  if (ll < 8) reinterpret_cast<R*>(a.zst)[((size_t)cell * a.nq + qs) * 32 + lh * 8 + ll] = Z;
  wave_sync();
}

// Staging for the matrix-core kernels: two prefetch sets used alternately,
// so that rows are loaded 16 rows (4 super-steps) before they are stored.
template <bool LDL>
__device__ __forceinline__ void bk_stage(const BandArgs& a, int cell, int jbase, BandSmem& sh,
                                         RowPrefetch& pf) {
  rows_store<LDL>(pf, sh);
  rows_load<LDL>(a, cell, jbase + 16, pf);
}

// The 4 super-steps of tile row T (pivots j0 .. j0+15); the next 16 rows of
// the band enter the freed slots.  All 4 run even past the last row (rows
// >= n are isolated unit diagonals, their panels zero): with no branch
// between them, the waits for the row prefetch leave the panel stores of the
// last super-steps in flight instead of waiting for their completion.
template <typename R, int DIM, int NT, int T, bool LDL>
__device__ __forceinline__ void bk_tile_row(const BandArgs& a, const BkAddr& ad, int cell, int j0,
                                            int my_off,
                                            typename Mf<R>::v4 (&Tl)[10],
                                            typename Mf<R>::v4 (&Yt)[4],
                                            typename Mf<R>::v4& Sacc, BandSmem& sh,
                                            BkSmem<R>& ms, RowPrefetch (&pf)[2], BkMon<R>& bad,
                                            const R (&ek)[4]) {
  constexpr int W = 16 * NT;
  bk_stage<LDL>(a, cell, j0 + W, sh, pf[0]);
  bk_superstep<R, DIM, NT, T, 0, LDL>(a, ad, cell, j0, Tl, Yt, Sacc, ms, bad, ek);
  bk_insert<R, DIM, NT, T, 0, LDL>(a, ad, j0 + W, 0, my_off, Tl, Yt, sh, ms);
  bk_superstep<R, DIM, NT, T, 1, LDL>(a, ad, cell, j0 + 4, Tl, Yt, Sacc, ms, bad, ek);
  bk_insert<R, DIM, NT, T, 1, LDL>(a, ad, j0 + 4 + W, 4, my_off, Tl, Yt, sh, ms);
  bk_stage<LDL>(a, cell, j0 + 8 + W, sh, pf[1]);
  bk_superstep<R, DIM, NT, T, 2, LDL>(a, ad, cell, j0 + 8, Tl, Yt, Sacc, ms, bad, ek);
  bk_insert<R, DIM, NT, T, 2, LDL>(a, ad, j0 + 8 + W, 0, my_off, Tl, Yt, sh, ms);
  bk_superstep<R, DIM, NT, T, 3, LDL>(a, ad, cell, j0 + 12, Tl, Yt, Sacc, ms, bad, ek);
  bk_insert<R, DIM, NT, T, 3, LDL>(a, ad, j0 + 12 + W, 4, my_off, Tl, Yt, sh, ms);
}

// Initial fill of tile row T with the rows 16T .. 16T+15.
template <typename R, int DIM, int NT, int T, bool LDL>
__device__ __forceinline__ void bk_fill_row(const BandArgs& a, const BkAddr& ad, int cell,
                                            int my_off,
                                            typename Mf<R>::v4 (&Tl)[10],
                                            typename Mf<R>::v4 (&Yt)[4], BandSmem& sh,
                                            BkSmem<R>& ms, RowPrefetch (&pf)[2]) {
  bk_stage<LDL>(a, cell, 16 * T, sh, pf[0]);
  bk_insert<R, DIM, NT, T, 0, LDL>(a, ad, 16 * T + 0, 0, my_off, Tl, Yt, sh, ms);
  bk_insert<R, DIM, NT, T, 1, LDL>(a, ad, 16 * T + 4, 4, my_off, Tl, Yt, sh, ms);
  bk_stage<LDL>(a, cell, 16 * T + 8, sh, pf[1]);
  bk_insert<R, DIM, NT, T, 2, LDL>(a, ad, 16 * T + 8, 0, my_off, Tl, Yt, sh, ms);
  bk_insert<R, DIM, NT, T, 3, LDL>(a, ad, 16 * T + 12, 4, my_off, Tl, Yt, sh, ms);
}

// Wait for every outstanding vector memory operation (vmcnt(0); expcnt and
// lgkmcnt left at their maxima).
__device__ __forceinline__ void vm_wait_all() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// Sum over the 8 lanes of an aligned group (DPP: swap neighbours, swap
// pairs, mirror within the half row).
__device__ __forceinline__ double dpp_f64(double v, int ctrl_sel) {
  const long long b = __builtin_bit_cast(long long, v);
  int lo = (int)b, hi = (int)(b >> 32);
  if (ctrl_sel == 0) {
    lo = __builtin_amdgcn_mov_dpp(lo, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
    hi = __builtin_amdgcn_mov_dpp(hi, 0xB1, 0xF, 0xF, false);
  } else if (ctrl_sel == 1) {
    lo = __builtin_amdgcn_mov_dpp(lo, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
    hi = __builtin_amdgcn_mov_dpp(hi, 0x4E, 0xF, 0xF, false);
  } else if (ctrl_sel == 2) {
    lo = __builtin_amdgcn_mov_dpp(lo, 0x141, 0xF, 0xF, false);  // row_half_mirror
    hi = __builtin_amdgcn_mov_dpp(hi, 0x141, 0xF, 0xF, false);
  } else {
    lo = __builtin_amdgcn_mov_dpp(lo, 0x128, 0xF, 0xF, false);  // row_ror:8
    hi = __builtin_amdgcn_mov_dpp(hi, 0x128, 0xF, 0xF, false);
  }
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double sum8(double v) {
  v += dpp_f64(v, 0);
  v += dpp_f64(v, 1);
  v += dpp_f64(v, 2);
  return v;
}

// Backward sweep.  (V^T x)[k][r] over the 60 offsets of a panel: lane
// (k0 = 2 ((lane >> 3) & 1), r0 = 2 (lane >> 4), c = lane & 7) sums a 2 x 2
// block of it over every 8th offset, the 8 lanes of the block add up by DPP,
// the other pair of k comes from lanes l ^ 8, and every lane applies C^-T
// for its column r = r0 + (c & 1).  The steps go in groups of 4: at the
// start of a group one wait covers the data of the group (the panels, z;
// for the refinement sweep the 4 table rows of each step; for a load per
// fine cell the loads of the 4 rows), loaded during the previous group into
// the other of two register sets, and the loads of the next group are
// issued.  The steps issue no global stores:
// the solution rows stay in the LDS ring (64 rows) and are written out 64
// rows at a time, one row per lane, right after a group's wait, so no wait
// ever covers a store or load issued less than a group earlier.
// MODE BK_FINAL64: Psi and the interior part of b; BK_X0: the solution to X;
// BK_REFINE: the ring holds the correction, x = X + correction (kept in
// xfr [8][64]) goes to X, Psi and b, and lane (i,k) accumulates the Schur
// term in its energy form S = R^T X + X^T r (r = Rres = R - K_II X), exact
// up to the second order in the error of the first solution X:
// S = R^T x* - e^T K_II e for X = x* + e.
template <typename R>
struct BkPre {
  R pp[4], zp;
  double rp0, rp1, xp, qp;
};

template <typename R, int DIM, int MODE, int NT = 4>
__device__ __forceinline__ void bk_backward(const BandArgs& a, int cell, BandSmem& sh,
                                            BkSmem<R>& ms, double* xfr, double& bint,
                                            double& sacc, double& dmax, double& xmax) {
  using D = Dirs<DIM>;
  const int lane = threadIdx.x;
  const int c = lane & 7, k0 = 2 * ((lane >> 3) & 1), r0 = 2 * (lane >> 4);
  const int n = a.n, nq = (n + 3) >> 2;
  R* pb = ms.piv;                     // [4][68]
  double* xr = SH_PART(sh);           // [8][132]
  double* zb = xr + 8 * kRing;        // [4][8]
  double* rw = zb + 32;               // table rows of the step: [4][24] (REFINE) or loads [4]
  double* xb = rw + 96;               // [4][8] first solution X of the step (BK_REFINE)
  double* r0b = xb + 32;              // [4][8] its residual (BK_REFINE)
  // fine node of interior row j (the row table's slot 15)
  auto node_of = [&](int j) {
    const int c0 = j % a.ni + 1, t = j / a.ni;
    const int c1 = (DIM == 3 ? t % a.ni : t) + 1, c2 = (DIM == 3) ? t / a.ni + 1 : 0;
    return c0 + a.S * (c1 + a.S * c2);
  };
  // the load of row t of the step (slot 14): from the table for a load per
  // fine cell, else the constant load times 2^dim
  const bool cell_load = a.rhs_mode == MSFEC_RHS_CELL;
  const double const_load = a.rhs_const * D::NV;
  for (int t = lane; t < 8 * kRing; t += 64) xr[t] = 0.0;
  for (int t = lane; t < 4 * kPanRow; t += 64) pb[t] = (R)0;
  const R* Pc = reinterpret_cast<const R*>(a.pan) + (size_t)cell * a.nq * 256;
  const R* Zc = reinterpret_cast<const R*>(a.zst) + (size_t)cell * a.nq * 32;
  const double* Rc = a.Rows + (size_t)cell * n * kRowStride;
  auto load = [&](int q, BkPre<R>& f) {
    if (q < 0) return;
#pragma unroll
    for (int u = 0; u < 4; ++u)  // [k = u][o = lane]; offsets beyond the window are zero
      f.pp[u] = (NT == 4 || lane < 16 * NT) ? Pc[(size_t)q * 256 + u * 64 + lane] : (R)0;
    f.zp = (lane < 32) ? Zc[(size_t)q * 32 + lane] : (R)0;
    const int lim = (n - 4 * q) * kRowStride;  // valid doubles of the 4 rows
    if (MODE == BK_REFINE) {  // the 4 whole rows (the lifting right-hand sides too)
      f.rp0 = (lane < lim) ? Rc[(size_t)4 * q * kRowStride + lane] : 0.0;
      f.rp1 = (lane + 64 < min(lim, 96)) ? Rc[(size_t)4 * q * kRowStride + 64 + lane] : 0.0;
    } else if (MODE == BK_FINAL64 && cell_load) {  // the loads of the 4 rows: slot 14
      const int rr = 4 * q + lane;
      f.rp0 = (lane < 4 && rr < n) ? Rc[(size_t)rr * kRowStride + 14] : 0.0;
    }
    if (MODE == BK_REFINE) {
      const bool ok = lane < 32 && 4 * q + (lane >> 3) < n;
      const size_t xi = ((size_t)cell * n + 4 * q) * kMaxNb + (ok ? lane : 0);
      f.xp = ok ? a.X[xi] : 0.0;
      f.qp = ok ? a.Rres[xi] : 0.0;
    }
  };
  // rows [jlo, jhi) (at most 64, all still in the ring) to X and / or Psi;
  // lane = row
  auto flush = [&](int jlo, int jhi) {
    wave_sync();
    const int j = jlo + lane;
    if (j < jhi) {
      const int p = j & (kWin - 1);
      double v[kMaxNb];
#pragma unroll
      for (int rr = 0; rr < kMaxNb; ++rr)
        v[rr] = (MODE == BK_REFINE) ? xfr[rr * 64 + p] : xr[rr * kRing + p];
      if (MODE != BK_FINAL64) {
        double2* xd = reinterpret_cast<double2*>(a.X + ((size_t)cell * n + j) * kMaxNb);
#pragma unroll
        for (int rr = 0; rr < kMaxNb; rr += 2) xd[rr >> 1] = make_double2(v[rr], v[rr + 1]);
      }
      if (MODE != BK_X0) {
        const int node = node_of(j);
#pragma unroll
        for (int rr = 0; rr < D::NV; ++rr) a.Psi[((size_t)cell * D::NV + rr) * a.nn + node] = v[rr];
      }
    }
  };
  auto step = [&](int q, const BkPre<R>& f) {
    const int j0 = 4 * q;
    wave_sync();
#pragma unroll
    for (int v = 0; v < 4; ++v) pb[v * kPanRow + lane] = f.pp[v];
    if (lane < 32) zb[lane] = (double)f.zp;
    if (MODE == BK_REFINE) {
      rw[lane] = f.rp0;
      if (lane < 32) rw[64 + lane] = f.rp1;
    } else if (MODE == BK_FINAL64 && cell_load && lane < 4) {
      rw[lane] = f.rp0;
    }
    if (MODE == BK_REFINE && lane < 32) {
      xb[lane] = f.xp;
      r0b[lane] = f.qp;
    }
    wave_sync();
    if (MODE == BK_REFINE) {
      // energy-form Schur term: lane (i = lane >> 3, kk = lane & 7)
      const int i = lane >> 3, kk = lane & 7;
#pragma unroll
      for (int t = 0; t < 4; ++t)
        sacc += rw[t * kRowStride + 16 + i] * xb[t * 8 + kk] + xb[t * 8 + i] * r0b[t * 8 + kk];
    }
    // (V^T x)[k][r] over the offsets 4..67 (the panel is zero at 64..67):
    // lane (k0, r0, c) sums the 2 x 2 block k0 + {0,1}, r0 + {0,1} over the
    // offsets 4 + c + 8i, then the 8 lanes of the block add up.
    const double* xw = xr + r0 * kRing + (j0 & (kWin - 1)) + 4 + c;
    const R* vw = pb + k0 * kPanRow + 4 + c;
    double s00 = 0.0, s01 = 0.0, s10 = 0.0, s11 = 0.0;
#pragma unroll
    for (int i = 0; i < 2 * NT; ++i) {
      const double v0 = (double)vw[8 * i], v1 = (double)vw[kPanRow + 8 * i];
      const double x0 = xw[8 * i], x1 = xw[kRing + 8 * i];
      s00 = fma(v0, x0, s00);
      s01 = fma(v0, x1, s01);
      s10 = fma(v1, x0, s10);
      s11 = fma(v1, x1, s11);
    }
    s00 = sum8(s00);
    s01 = sum8(s01);
    s10 = sum8(s10);
    s11 = sum8(s11);
    // the other pair of k for the same pair of r sits in lanes l ^ 8 (a
    // rotation of the row of 16 by 8): every lane then has (V^T x)[k][r]
    // for all 4 k and solves for r = r0 + (c & 1), without a round trip
    // through LDS; the lanes c < 2 of the groups with k0 = 0 write
    const double q00 = dpp_f64(s00, 3), q01 = dpp_f64(s01, 3);
    const double q10 = dpp_f64(s10, 3), q11 = dpp_f64(s11, 3);
    const int rb = c & 1, rr = r0 + rb;
    const bool lowk = k0 == 0;
    const double oa = rb ? s01 : s00, ob = rb ? s11 : s10;  // own k0, k0 + 1
    const double qa = rb ? q01 : q00, qb = rb ? q11 : q10;  // the partner's
    {
      const double w0 = zb[0 * 8 + rr] - (lowk ? oa : qa), w1 = zb[1 * 8 + rr] - (lowk ? ob : qb);
      const double w2 = zb[2 * 8 + rr] - (lowk ? qa : oa), w3 = zb[3 * 8 + rr] - (lowk ? qb : ob);
      // C^T x = w with C (t,u) at pb[u * 68 + t] and 1 / C[t][t] at pb[t * 68 + t]
      const double x3 = w3 * (double)pb[3 * kPanRow + 3];
      const double x2 = (w2 - (double)pb[2 * kPanRow + 3] * x3) * (double)pb[2 * kPanRow + 2];
      const double x1 = (w1 - (double)pb[1 * kPanRow + 2] * x2 - (double)pb[1 * kPanRow + 3] * x3) *
                        (double)pb[1 * kPanRow + 1];
      const double x0 = (w0 - (double)pb[0 * kPanRow + 1] * x1 - (double)pb[0 * kPanRow + 2] * x2 -
                         (double)pb[0 * kPanRow + 3] * x3) *
                        (double)pb[0 * kPanRow + 0];
      const double xs[4] = {x0, x1, x2, x3};
      const bool wr = lowk && c < 2;
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int j = j0 + t;
        if (!wr || j >= n) continue;
        const int p = j & (kWin - 1);
        xr[rr * kRing + p] = xs[t];
        xr[rr * kRing + p + 64] = xs[t];
        if (MODE != BK_X0) {
          const double xf = (MODE == BK_REFINE) ? xb[t * 8 + rr] + xs[t] : xs[t];
          if (MODE == BK_REFINE) {
            xfr[rr * 64 + p] = xf;
            dmax = fmax(dmax, fabs(xs[t]));
            xmax = fmax(xmax, fabs(xf));
          }
          const double fj = (MODE == BK_REFINE) ? rw[t * kRowStride + 14]
                            : cell_load                ? rw[t]
                                                       : const_load;
          bint = fma(xf, fj, bint);
        }
      }
    }
  };
  int jflush = n;  // rows >= jflush are written out
  // group of the steps qg, qg - 1, .., qg - 3 with its data in cur
  auto group = [&](int qg, const BkPre<R> (&cur)[4], BkPre<R> (&nxt)[4]) {
    vm_wait_all();
    const int done = nq - 1 - qg;  // steps before this group
    if (done > 0 && (done & 15) == 0) {
      flush(4 * (qg + 1), jflush);
      jflush = 4 * (qg + 1);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) load(qg - 4 - u, nxt[u]);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (qg - u < 0) break;
      step(qg - u, cur[u]);
    }
  };
  BkPre<R> pa[4], pc[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) load(nq - 1 - u, pa[u]);
  for (int q0 = nq - 1; q0 >= 0; q0 -= 8) {
    group(q0, pa, pc);
    if (q0 - 4 < 0) break;
    group(q0 - 4, pc, pa);
  }
  if (jflush > 0) flush(0, jflush);
  wave_sync();
  // the interior part of b for basis r was summed by lane 16 (r >> 1) + (r & 1)
  bint = __shfl(bint, 16 * ((lane & 7) >> 1) + (lane & 1));
}

// Sacc (accumulator layout) -> lane (i = lane >> 3, k = lane & 7)
template <typename R>
__device__ __forceinline__ double bk_sacc(const typename Mf<R>::v4& Sacc, double* buf) {
  const int lane = threadIdx.x;
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    const int i = Mf<R>::acc_row(g, lane >> 4), j = lane & 15;
    if (i < 8 && j < 8) buf[i * 8 + j] = (double)Sacc[g];
  }
  wave_sync();
  const double v = buf[lane];
  wave_sync();
  return v;
}

// Factor + forward elimination of the lifting right-hand sides, then the
// backward sweep.  MODE BK_FINAL64 (FP64 tiles): Psi, M, b of the cell;
// MODE BK_X0 (FP32 tiles): the first solution X of the mixed solver.
// waves per SIMD the register budget is sized for: FP64 tiles need about
// 200 VGPRs, FP32 tiles half of that
template <typename R>
struct BkOcc {
  static constexpr int waves = sizeof(R) == 8 ? 2 : 3;
};

// LDL: the rows are those of a symmetric indefinite band matrix with the
// offsets of their lower entries in a.loff (factored as L D L^T, 16 lower
// entries per row); only with MODE BK_X0.
template <typename R, int DIM, int MODE, bool LDL = false, int NT = 4>
__global__ __launch_bounds__(64, BkOcc<R>::waves) void q_blk_factor_kernel(BandArgs a) {
  typedef typename Mf<R>::v4 V4;
  const int cell = blockIdx.x;
  const int lane = threadIdx.x;
  __shared__ BandSmem sh;
  __shared__ BkSmem<R> ms;
  int my_off = 0;
#pragma unroll
  for (int t = 0; t < kMaxLow; ++t)
    if ((lane & 15) == t) my_off = a.off_int[t];

  V4 Tl[10], Yt[4], Sacc = {0, 0, 0, 0};
#pragma unroll
  for (int t = 0; t < 10; ++t) Tl[t] = (V4){0, 0, 0, 0};
#pragma unroll
  for (int P = 0; P < 4; ++P) Yt[P] = (V4){0, 0, 0, 0};
  const BkAddr ad = bk_addr<R>();
  R ek[4];  // unit vector e_(lane >> 4)
#pragma unroll
  for (int t = 0; t < 4; ++t) ek[t] = ((lane >> 4) == t) ? (R)1 : (R)0;
  RowPrefetch pf[2];
  rows_load<LDL>(a, cell, 0, pf[0]);
  rows_load<LDL>(a, cell, 8, pf[1]);
  // the tile rows beyond NT are never instantiated (their template index
  // folds onto 0 and the branch is a compile-time false)
  constexpr int T1 = NT > 1 ? 1 : 0, T2 = NT > 2 ? 2 : 0, T3 = NT > 2 ? 3 : 0;
  bk_fill_row<R, DIM, NT, 0, LDL>(a, ad, cell, my_off, Tl, Yt, sh, ms, pf);
  if (NT > 1) bk_fill_row<R, DIM, NT, T1, LDL>(a, ad, cell, my_off, Tl, Yt, sh, ms, pf);
  if (NT > 2) bk_fill_row<R, DIM, NT, T2, LDL>(a, ad, cell, my_off, Tl, Yt, sh, ms, pf);
  if (NT > 2) bk_fill_row<R, DIM, NT, T3, LDL>(a, ad, cell, my_off, Tl, Yt, sh, ms, pf);
  BkMon<R> bad;
  bad.init();
  const int n = a.n;
  for (int j0 = 0; j0 < n; j0 += 16 * NT) {
    bk_tile_row<R, DIM, NT, 0, LDL>(a, ad, cell, j0, my_off, Tl, Yt, Sacc, sh, ms, pf, bad, ek);
    if (NT > 1 && j0 + 16 < n)
      bk_tile_row<R, DIM, NT, T1, LDL>(a, ad, cell, j0 + 16, my_off, Tl, Yt, Sacc, sh, ms, pf, bad, ek);
    if (NT > 2 && j0 + 32 < n)
      bk_tile_row<R, DIM, NT, T2, LDL>(a, ad, cell, j0 + 32, my_off, Tl, Yt, Sacc, sh, ms, pf, bad, ek);
    if (NT > 2 && j0 + 48 < n)
      bk_tile_row<R, DIM, NT, T3, LDL>(a, ad, cell, j0 + 48, my_off, Tl, Yt, Sacc, sh, ms, pf, bad, ek);
  }
  double sacc = (MODE == BK_FINAL64) ? bk_sacc<R>(Sacc, SH_PART(sh)) : 0.0;
  // panels of this wave must be visible to its own loads in the backward sweep
  __threadfence_block();
  __syncthreads();
  double bint = 0.0, dmax = 0.0, xmax = 0.0;
  bk_backward<R, DIM, MODE, NT>(a, cell, sh, ms, nullptr, bint, sacc, dmax, xmax);
  int worst = bad.status();
  for (int o = 32; o > 0; o >>= 1) worst = max(worst, __shfl_xor(worst, o));
  if (lane == 0) {
    a.status[cell] = worst;
    if (MODE == BK_X0 && !LDL) a.conv[cell] = (worst != 0) ? 1 : 0;  // failed cells do not refine
  }
  if (MODE == BK_FINAL64) fast_finish<DIM>(a, cell, sacc, bint);
}

// ---------------------------------------------------------------------------
// Mixed-precision solver, refinement residual: Rres = R - K_II X in FP64.
// One block per cell stages the cell's X (n x 8 doubles, at most 61.5 KB) in
// LDS; one thread per (interior row, right-hand side).  K_II comes from the
// row table: lower entries of row j, upper entries as the lower entries of
// the rows j + off (zero where the neighbour is not interior).
// ---------------------------------------------------------------------------
template <int DIM>
__global__ __launch_bounds__(256) void q_resid_kernel(BandArgs a, int stage_lo) {
  using D = Dirs<DIM>;
  const int cell = blockIdx.x;
  if (a.conv[cell]) return;  // uniform over the block
  extern __shared__ double xs[];  // [n][8] | lower stencil [n][NLOW] when stage_lo
  const int n = a.n;
  double* los = xs + (size_t)n * kMaxNb;
  const double* rows = a.Rows + (size_t)cell * n * kRowStride;
  const double* X = a.X + (size_t)cell * n * kMaxNb;
  for (int i = threadIdx.x; i < n * kMaxNb; i += blockDim.x) xs[i] = X[i];
  if (stage_lo)
    for (int i = threadIdx.x; i < n * D::NLOW; i += blockDim.x) {
      const int j = i / D::NLOW, t = i - j * D::NLOW;
      los[i] = rows[(size_t)j * kRowStride + t];
    }
  __syncthreads();
  for (int i = threadIdx.x; i < n * kMaxNb; i += blockDim.x) {
    const int j = i >> 3, r = i & (kMaxNb - 1);
    const double* rj = rows + (size_t)j * kRowStride;
    auto lo = [&](int jj, int t) {
      return stage_lo ? los[jj * D::NLOW + t] : rows[(size_t)jj * kRowStride + t];
    };
    double acc = rj[16 + r] - lo(j, 0) * xs[i];
    // no branches, so that the reads issue together: a lower coupling to a
    // row below 0 is zero in the table (its neighbour is on the boundary),
    // and an upper one beyond n - 1 is masked; indices are clamped
#pragma unroll
    for (int t = 1; t < D::NLOW; ++t) {
      const int off = a.off_int[t];
      const int jm = max(j - off, 0), jp = j + off;
      const bool pin = jp < n;
      const int jq = pin ? jp : n - 1;
      acc = fma(-lo(j, t), xs[jm * kMaxNb + r], acc);
      const double up = lo(jq, t);
      acc = fma(pin ? -up : 0.0, xs[jq * kMaxNb + r], acc);
    }
    a.Rres[((size_t)cell * n + j) * kMaxNb + r] = acc;
  }
}

// ---------------------------------------------------------------------------
// Mixed-precision solver, refinement sweep: forward elimination of Rres with
// the stored FP32 panels (4 MFMAs per super-step), backward sweep of the
// correction, x = X + correction, then the element matrix and Psi.
// ---------------------------------------------------------------------------
// Per-lane prefetch for one super-step of a sweep with the stored panels:
// the panel (4 values) and one value of the right-hand-side row entering the
// window (from a.Rres).
template <typename R>
struct RfPre {
  R pp[4];
  R rq;
};

template <typename R, int NT>
__device__ __forceinline__ void rf_load(const BandArgs& a, int cell, int q, RfPre<R>& f) {
  const int lane = threadIdx.x, lh = lane >> 4, ll = lane & 15;
  const int n = a.n;
  if (4 * q >= n) return;
  const R* Pn = reinterpret_cast<const R*>(a.pan) + ((size_t)cell * a.nq + q) * 256;
#pragma unroll
  for (int u = 0; u < 4; ++u)  // [k = u][o = lane]
    f.pp[u] = (NT == 4 || lane < 16 * NT) ? Pn[u * 64 + lane] : (R)0;
  const int jn = 4 * q + 16 * NT + lh;
  const R rv = (R)a.Rres[((size_t)cell * n + min(jn, n - 1)) * kMaxNb + (ll & 7)];
  f.rq = (ll < 8 && jn < n) ? rv : (R)0;
}

// One super-step of the forward elimination of a.Rres with the stored
// panels: Z = C^-1 y_pivot, Y_P -= V_P Z (4 MFMAs), Z to HBM, next rows in.
template <typename R, int DIM, int NT, int T, int G, bool LDL>
__device__ __forceinline__ void rf_step(const BandArgs& a, int cell, int j0,
                                        typename Mf<R>::v4 (&Yt)[4], BkSmem<R>& ms,
                                        RfPre<R>& f) {
  typedef Mf<R> M;
  const int lane = bk_lane(), lh = lane >> 4, ll = lane & 15, cs = M::cslot(ll);
  R* pb = ms.piv;  // [4][68]
  const int qs = j0 >> 2;
  // panel of this super-step and the pivot right-hand sides to LDS
#pragma unroll
  for (int u = 0; u < 4; ++u) pb[u * kPanRow + lane] = f.pp[u];
  ms.yp[lh * 16 + ll] = Yt[T][G];
  const R rin = f.rq;
  wave_sync();
  rf_load<R, NT>(a, cell, qs + 4, f);  // 4 super-steps ahead
  // Z = C^-1 y_pivot by forward substitution (C (t,u) at pb[u * 68 + t],
  // 1 / C[t][t] at pb[t * 68 + t]); lane (k = lh, rhs ll) keeps row lh
  const R y0 = ms.yp[0 * 16 + ll], y1 = ms.yp[1 * 16 + ll];
  const R y2 = ms.yp[2 * 16 + ll], y3 = ms.yp[3 * 16 + ll];
  const R z0 = y0 * pb[0 * kPanRow + 0];
  const R z1 = (y1 - pb[0 * kPanRow + 1] * z0) * pb[1 * kPanRow + 1];
  const R z2 = (y2 - pb[0 * kPanRow + 2] * z0 - pb[1 * kPanRow + 2] * z1) * pb[2 * kPanRow + 2];
  const R z3 = (y3 - pb[0 * kPanRow + 3] * z0 - pb[1 * kPanRow + 3] * z1 -
                pb[2 * kPanRow + 3] * z2) *
               pb[3 * kPanRow + 3];
  const R Z = (lh == 0) ? z0 : (lh == 1) ? z1 : (lh == 2) ? z2 : z3;
  // LDL^T panels hold C = L1 D1: the update takes D1^-1 L1^-1 y (= Z), the
  // backward sweep the unit-triangular L1^-1 y = D1 Z (1 / d_lh on the diagonal)
  const R Zst = LDL ? Z / pb[lh * kPanRow + lh] : Z;
  R nV[4];
  const int jr = bk_opaque(j0);  // = s0 mod 64
#pragma unroll
  for (int P = 0; P < NT; ++P) nV[P] = -pb[lh * kPanRow + ((16 * P + cs - jr) & (16 * NT - 1))];
#pragma unroll
  for (int P = 0; P < NT; ++P) Yt[P] = M::mma(nV[P], Z, Yt[P]);
  if (ll < 8) reinterpret_cast<R*>(a.zst)[((size_t)cell * a.nq + qs) * 32 + lh * 8 + ll] = Zst;
  // the next rows of the band enter the pivot slots
  Yt[T][G] = rin;
  wave_sync();
}

template <typename R, int DIM, int NT, int T, bool LDL>
__device__ __forceinline__ void rf_tile_row(const BandArgs& a, int cell, int j0,
                                            typename Mf<R>::v4 (&Yt)[4], BkSmem<R>& ms,
                                            RfPre<R> (&f)[4]) {
  const int n = a.n;
  rf_step<R, DIM, NT, T, 0, LDL>(a, cell, j0, Yt, ms, f[0]);
  if (j0 + 4 < n) rf_step<R, DIM, NT, T, 1, LDL>(a, cell, j0 + 4, Yt, ms, f[1]);
  if (j0 + 8 < n) rf_step<R, DIM, NT, T, 2, LDL>(a, cell, j0 + 8, Yt, ms, f[2]);
  if (j0 + 12 < n) rf_step<R, DIM, NT, T, 3, LDL>(a, cell, j0 + 12, Yt, ms, f[3]);
}

// Forward sweep of a.Rres over all super-steps of the cell.
template <typename R, int DIM, int NT = 4, bool LDL = false>
__device__ __forceinline__ void rf_forward(const BandArgs& a, int cell, BkSmem<R>& ms) {
  typedef typename Mf<R>::v4 V4;
  const int lane = threadIdx.x, lh = lane >> 4, ll = lane & 15;
  const int n = a.n;
  RfPre<R> pre[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) rf_load<R, NT>(a, cell, u, pre[u]);
  V4 Yt[4];
  // rows 0..16 NT - 1: slot 4g + lh of tile row P is register g
#pragma unroll
  for (int P = 0; P < NT; ++P)
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const int j = 16 * P + 4 * g + lh;
      const R rv = (R)a.Rres[((size_t)cell * n + min(j, n - 1)) * kMaxNb + (ll & 7)];
      Yt[P][g] = (ll < 8 && j < n) ? rv : (R)0;
    }
  for (int t = lane; t < 4 * kPanRow; t += 64) ms.piv[t] = (R)0;
  wave_sync();
  for (int j0 = 0; j0 < n; j0 += 16 * NT) {
    rf_tile_row<R, DIM, NT, 0, LDL>(a, cell, j0, Yt, ms, pre);
    if (NT > 1 && j0 + 16 < n)
      rf_tile_row<R, DIM, NT, (NT > 1 ? 1 : 0), LDL>(a, cell, j0 + 16, Yt, ms, pre);
    if (NT > 2 && j0 + 32 < n)
      rf_tile_row<R, DIM, NT, (NT > 2 ? 2 : 0), LDL>(a, cell, j0 + 32, Yt, ms, pre);
    if (NT > 2 && j0 + 48 < n)
      rf_tile_row<R, DIM, NT, (NT > 2 ? 3 : 0), LDL>(a, cell, j0 + 48, Yt, ms, pre);
  }
}

template <int DIM>
__global__ __launch_bounds__(64, 2) void q_blk_refine_kernel(BandArgs a) {
  const int cell = blockIdx.x;
  const int lane = threadIdx.x;
  if (a.conv[cell]) return;  // converged or failed: the whole wave leaves
  __shared__ BandSmem sh;
  __shared__ BkSmem<float> ms;
  __shared__ double xfr[kMaxNb * 64];  // refined solution of the ring rows
  rf_forward<float, DIM, 4>(a, cell, ms);
  __threadfence_block();
  __syncthreads();
  double bint = 0.0, sacc = 0.0, dmax = 0.0, xmax = 0.0;
  bk_backward<float, DIM, BK_REFINE>(a, cell, sh, ms, xfr, bint, sacc, dmax, xmax);
  for (int o = 32; o > 0; o >>= 1) {
    dmax = fmax(dmax, __shfl_xor(dmax, o));
    xmax = fmax(xmax, __shfl_xor(xmax, o));
  }
  // The error left after this step is about (|dx| / |x|)^2 |x| (the
  // contraction of the refinement is about |dx| / |x| itself); a cell stops
  // once that is below 1e-11, a tenth of the 1e-10 acceptance bound.
  const double q2 = (dmax * dmax) / fmax(xmax * xmax, 1e-300);
  if (lane == 0) {
    if (q2 <= 1e-11) {
      a.conv[cell] = 1;
    } else if (a.last_pass) {
      a.conv[cell] = 1;
      if (q2 > 1e-10 && a.status[cell] == MSFEC_CELL_OK) a.status[cell] = MSFEC_CELL_NOT_CONVERGED;
    }
  }
  fast_finish<DIM>(a, cell, sacc, bint);
}

// ---------------------------------------------------------------------------
// Solve K_II x = a.Rres (8 columns) with the panels stored by the last
// factorisation (FP64 or FP32), x to a.X.  Used by the adjoint for
// w = K_II^{-1} F_I.
// ---------------------------------------------------------------------------
template <typename R, int DIM, int NT>
__global__ __launch_bounds__(64, 2) void q_blk_solve_kernel(BandArgs a) {
  const int cell = blockIdx.x;
  __shared__ BandSmem sh;
  __shared__ BkSmem<R> ms;
  rf_forward<R, DIM, NT>(a, cell, ms);
  __threadfence_block();
  __syncthreads();
  double bint = 0.0, sacc = 0.0, dmax = 0.0, xmax = 0.0;
  bk_backward<R, DIM, BK_X0, NT>(a, cell, sh, ms, nullptr, bint, sacc, dmax, xmax);
}

// ---------------------------------------------------------------------------
// Mixed precision for the symmetric indefinite band systems (Q_Ned): FP32
// LDL^T on the matrix cores (q_blk_factor_kernel<float, 2, BK_X0, true>),
// then refinement passes, each an FP64 residual (ldl_resid_kernel), a
// correction solved with the stored FP32 panels (ldl_refine_kernel) and the
// update x += correction with the convergence test (ldl_update_kernel).
// ---------------------------------------------------------------------------
// Rres = R - K X in FP64 for rows [n][24] (16 lower entries with their
// offsets in loff, diagonal first; 8 right-hand sides).  The upper part of
// row j is read through utab [n][16]: (i << 4 | t) for the lower entry t of
// row i > j that couples to j, -1 unused.  One block per cell, X staged in LDS.
__global__ __launch_bounds__(256) void ldl_resid_kernel(BandArgs a, const int* __restrict__ utab) {
  const int cell = blockIdx.x;
  if (a.conv[cell]) return;  // uniform over the block
  extern __shared__ double xs[];  // [n][8]
  const int n = a.n;
  const double* rows = a.Rows + (size_t)cell * n * kRowStride;
  const double* X = a.X + (size_t)cell * n * kMaxNb;
  for (int i = threadIdx.x; i < n * kMaxNb; i += blockDim.x) xs[i] = X[i];
  __syncthreads();
  // one thread per row and all 8 right-hand sides: each matrix entry is
  // loaded once and used 8 times
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    const double* rj = rows + (size_t)j * kRowStride;
    const int* lo = a.loff + (size_t)j * 16;
    const int* up = utab + (size_t)j * 16;
    double acc[kMaxNb];
#pragma unroll
    for (int r = 0; r < kMaxNb; ++r) acc[r] = rj[16 + r];
    // 4 entries in flight: unrolled over all 32 the loads of every entry and
    // its 8 x values stay live at once and spill
#pragma unroll 4
    for (int t = 0; t < 16; ++t) {
      const int o = lo[t];
      const double v = (o >= 0) ? -rj[t] : 0.0;
      const double* xr = xs + ((o >= 0) ? j - o : j) * kMaxNb;
#pragma unroll
      for (int r = 0; r < kMaxNb; ++r) acc[r] = fma(v, xr[r], acc[r]);
    }
#pragma unroll 4
    for (int t = 0; t < 16; ++t) {
      const int u = up[t];
      const int ri = (u >= 0) ? (u >> 4) : j, ti = (u >= 0) ? (u & 15) : 0;
      const double v = (u >= 0) ? -rows[(size_t)ri * kRowStride + ti] : 0.0;
      const double* xr = xs + ri * kMaxNb;
#pragma unroll
      for (int r = 0; r < kMaxNb; ++r) acc[r] = fma(v, xr[r], acc[r]);
    }
    double* out = a.Rres + ((size_t)cell * n + j) * kMaxNb;
#pragma unroll
    for (int r = 0; r < kMaxNb; ++r) out[r] = acc[r];
  }
}

// The correction K d = Rres with the stored FP32 LDL^T panels, to a.X
// (which the host points at the correction buffer).
template <int NT>
__global__ __launch_bounds__(64, 2) void ldl_refine_kernel(BandArgs a) {
  const int cell = blockIdx.x;
  if (a.conv[cell]) return;  // converged or failed: the whole wave leaves
  __shared__ BandSmem sh;
  __shared__ BkSmem<float> ms;
  rf_forward<float, 2, NT, true>(a, cell, ms);
  __threadfence_block();
  __syncthreads();
  double bint = 0.0, sacc = 0.0, dmax = 0.0, xmax = 0.0;
  bk_backward<float, 2, BK_X0, NT>(a, cell, sh, ms, nullptr, bint, sacc, dmax, xmax);
}

// x += d for the cells still refining.  With q = |d| / |x| and q' its value
// in the pass before (1 in the first), the error left after the step is
// about q rho with the contraction rho = q / q' (the first pass: q itself,
// since d then is the first solution's error).  A cell stops once that is
// below 1e-13 of the largest unknown, or after the last pass, flagged
// MSFEC_CELL_NOT_CONVERGED above 1e-12.  The blocks of a mixed system differ
// in scale (sigma of Q_Ned is ~1/40 of the largest entry): 1e-13 of the
// largest entry keeps the refinement's share of sigma's error below 1e-11 of
// sigma, its FP64 floor.  One block per cell.
__global__ __launch_bounds__(256) void ldl_update_kernel(BandArgs a, const double* __restrict__ dX,
                                                         double* __restrict__ qprev) {
  const int cell = blockIdx.x;
  if (a.conv[cell]) return;
  __shared__ double red[2][4];
  const size_t base = (size_t)cell * a.n * kMaxNb;
  double dm = 0.0, xm = 0.0;
  for (int i = threadIdx.x; i < a.n * kMaxNb; i += blockDim.x) {
    const double x = a.X[base + i] + dX[base + i];
    a.X[base + i] = x;
    // a NaN turns the measure infinite (fmax alone would skip it)
    const double ad = fabs(dX[base + i]), ax = fabs(x);
    dm = (ad == ad) ? fmax(dm, ad) : INFINITY;
    xm = (ax == ax) ? fmax(xm, ax) : INFINITY;
  }
  for (int o = 32; o > 0; o >>= 1) {
    dm = fmax(dm, __shfl_xor(dm, o));
    xm = fmax(xm, __shfl_xor(xm, o));
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = dm;
    red[1][threadIdx.x >> 6] = xm;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    dm = fmax(fmax(red[0][0], red[0][1]), fmax(red[0][2], red[0][3]));
    xm = fmax(fmax(red[1][0], red[1][1]), fmax(red[1][2], red[1][3]));
    const double q = dm / fmax(xm, 1e-300);
    const double est = q * fmin(1.0, q / fmax(qprev[cell], 1e-300));
    qprev[cell] = q;
    if (!isfinite(est)) {
      a.conv[cell] = 1;
      if (a.status[cell] == MSFEC_CELL_OK) a.status[cell] = MSFEC_CELL_NONFINITE;
    } else if (est <= 1e-13) {
      a.conv[cell] = 1;
    } else if (a.last_pass) {
      a.conv[cell] = 1;
      if (est > 1e-12 && a.status[cell] == MSFEC_CELL_OK) a.status[cell] = MSFEC_CELL_NOT_CONVERGED;
    }
  }
}

// conv = 1 for the cells whose factorisation failed (they do not refine);
// the previous correction ratio starts at 1
__global__ void ldl_conv_init_kernel(const int* __restrict__ status, int* __restrict__ conv,
                                     double* __restrict__ qprev, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    conv[i] = (status[i] != MSFEC_CELL_OK) ? 1 : 0;
    qprev[i] = 1.0;
  }
}

// Right-hand side of w: column 0 = F_I (interior load), columns 1..7 = 0.
__global__ void q_adj_rhs_kernel(BandArgs a, int n_cells) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)n_cells * a.n * kMaxNb) return;
  const int r = (int)(i & (kMaxNb - 1));
  const long long cj = i >> 3;
  const int cell = (int)(cj / a.n);
  const double fj = a.Rows[(size_t)cj * kRowStride + 14] * a.Pb[(size_t)cell * kPStride + 72];
  a.Rres[i] = (r == 0) ? fj : 0.0;
}

// ---------------------------------------------------------------------------
// Adjoint gradient of J = 1/2 |u_H - u_target|^2 (free coarse nodes) with
// respect to the log-coefficient g (a = exp(g)) of every fine cell:
//   dJ/dg_e = -a_e vt_e^T Khat_e (ut_e + w_e),
// ut, vt the fine reconstructions of the coarse solution u and adjoint v,
// w = K_II^{-1} F_I (column 0 of a.X, zero on the cell boundary) and Khat_e
// the unit-coefficient element stiffness of fine cell e.  One block per
// coarse cell: the nodal fields go to LDS, then one thread per fine cell.
// ---------------------------------------------------------------------------
template <int DIM>
__global__ __launch_bounds__(256) void q_adj_grad_kernel(BandArgs a, int nc,
                                                         const double* __restrict__ u,
                                                         const double* __restrict__ v,
                                                         double* __restrict__ grad) {
  using D = Dirs<DIM>;
  const int cell = blockIdx.x;
  extern __shared__ double fld[];  // [nn] ut + w | [nn] vt
  double* su = fld;
  double* sv = fld + a.nn;
  __shared__ double kvu[8];
  double uk[D::NV], vk[D::NV];
  {
    int cijk[3] = {0, 0, 0}, r = cell;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      cijk[d] = r % nc;
      r /= nc;
    }
#pragma unroll
    for (int k = 0; k < D::NV; ++k) {
      int id = 0, mul = 1;
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        id += (cijk[d] + ((k >> d) & 1)) * mul;
        mul *= nc + 1;
      }
      uk[k] = u[id];
      vk[k] = v[id];
    }
  }
  if (threadIdx.x < 8) {
    const double* vt = a.verts + (size_t)cell * D::NV * DIM;
    double hh[3] = {1.0, 1.0, 1.0};
#pragma unroll
    for (int d = 0; d < DIM; ++d) hh[d] = (vt[(D::NV - 1) * DIM + d] - vt[d]) / a.N;
    double ks, ms;
    box_tables<DIM>(threadIdx.x, hh, ks, ms);
    kvu[threadIdx.x] = ks;
  }
  const double* psi = a.Psi + (size_t)cell * D::NV * a.nn;
  for (int node = threadIdx.x; node < a.nn; node += blockDim.x) {
    double ut = 0.0, vt = 0.0;
#pragma unroll
    for (int k = 0; k < D::NV; ++k) {
      const double p = psi[(size_t)k * a.nn + node];
      ut = fma(uk[k], p, ut);
      vt = fma(vk[k], p, vt);
    }
    int c[3] = {0, 0, 0}, t = node;
    bool inner = true;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      c[d] = t % a.S;
      t /= a.S;
      if (c[d] == 0 || c[d] == a.N) inner = false;
    }
    if (inner) {
      int j = 0, mul = 1;
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        j += (c[d] - 1) * mul;
        mul *= a.ni;
      }
      ut += a.X[((size_t)cell * a.n + j) * kMaxNb];
    }
    su[node] = ut;
    sv[node] = vt;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < a.nfc; e += blockDim.x) {
    int c[3] = {0, 0, 0}, t = e;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      c[d] = t % a.N;
      t /= a.N;
    }
    double ue[D::NV], ve[D::NV];
#pragma unroll
    for (int al = 0; al < D::NV; ++al) {
      int nd = 0, mul = 1;
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        nd += (c[d] + ((al >> d) & 1)) * mul;
        mul *= a.S;
      }
      ue[al] = su[nd];
      ve[al] = sv[nd];
    }
    double s = 0.0;
#pragma unroll
    for (int al = 0; al < D::NV; ++al) {
      double kr = 0.0;
#pragma unroll
      for (int be = 0; be < D::NV; ++be) kr = fma(kvu[al ^ be], ue[be], kr);
      s = fma(ve[al], kr, s);
    }
    grad[(size_t)cell * a.nfc + e] = -a.coef[(size_t)cell * a.nfc + e] * s;
  }
}

// ---------------------------------------------------------------------------
// K3: Psi, element matrix and element right-hand side
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void q_element_kernel(QConf c, const double* __restrict__ Slo,
                                                        const double* __restrict__ Fv,
                                                        const double* __restrict__ X,
                                                        double* __restrict__ Psi,
                                                        double* __restrict__ Mout,
                                                        double* __restrict__ bout) {
  extern __shared__ double smem[];
  const int cell = blockIdx.x;
  const int nn = c.nn, nv = c.nv;
  double* psi = smem;             // [nv][nn]
  double* kpsi = smem + nv * nn;  // [nv][nn]
  double* red = kpsi + nv * nn;   // [4][64 + 8]
  const double* slo = Slo + (size_t)cell * nn * c.nlow;
  const double* fv = Fv + (size_t)cell * nn;
  const double* xc = X + (size_t)cell * c.n * kMaxNb;
  double* psio = Psi + (size_t)cell * nv * nn;

  for (int node = threadIdx.x; node < nn; node += blockDim.x) {
    int cn[3] = {0, 0, 0}, r = node;
    bool interior = true;
    int ii = 0, mul = 1;
    for (int d = 0; d < c.dim; ++d) {
      cn[d] = r % c.S;
      r /= c.S;
      if (cn[d] == 0 || cn[d] == c.N) interior = false;
      ii += (cn[d] - 1) * mul;
      mul *= c.ni;
    }
    for (int k = 0; k < nv; ++k) {
      const double v = interior ? xc[(size_t)ii * kMaxNb + k] : lift_value(c, k, cn);
      psi[k * nn + node] = v;
      psio[(size_t)k * nn + node] = v;
    }
  }
  __syncthreads();
  // K psi with the symmetric stencil: lower entries from the own row, upper
  // entries from the neighbour's row
  for (int node = threadIdx.x; node < nn; node += blockDim.x) {
    int cn[3] = {0, 0, 0}, r = node;
    for (int d = 0; d < c.dim; ++d) {
      cn[d] = r % c.S;
      r /= c.S;
    }
    double acc[kMaxNb];
    for (int k = 0; k < kMaxNb; ++k) acc[k] = 0.0;
    for (int e = 0; e < c.nst; ++e) {
      int dd[3] = {0, 0, 0}, rr = e, off = 0, mm = 1;
      bool ok = true;
      for (int d = 0; d < c.dim; ++d) {
        dd[d] = rr % 3 - 1;
        rr /= 3;
        const int mcoord = cn[d] + dd[d];
        if (mcoord < 0 || mcoord > c.N) ok = false;
        off += dd[d] * mm;
        mm *= c.S;
      }
      if (!ok) continue;
      const int nb = node + off;
      double kv;
      if (e == c.center) kv = slo[(size_t)node * c.nlow];
      else if (e < c.center) kv = slo[(size_t)node * c.nlow + 1 + e];
      else kv = slo[(size_t)nb * c.nlow + 1 + (c.nst - 1 - e)];
      for (int k = 0; k < nv; ++k) acc[k] = fma(kv, psi[k * nn + nb], acc[k]);
    }
    for (int k = 0; k < nv; ++k) kpsi[k * nn + node] = acc[k];
  }
  __syncthreads();
  // M[i][j] = psi_i . (K psi_j): 64 (i,j) pairs x 4 node partitions
  {
    const int t = threadIdx.x;
    const int ij = t & 63, part = t >> 6;
    const int i = ij >> 3, j = ij & 7;
    double s = 0.0;
    if (i < nv && j < nv)
      for (int node = part; node < nn; node += 4) s = fma(psi[i * nn + node], kpsi[j * nn + node], s);
    red[part * 72 + ij] = s;
    double sb = 0.0;
    if (ij < nv)
      for (int node = part; node < nn; node += 4) sb = fma(psi[ij * nn + node], fv[node], sb);
    if (ij < 8) red[part * 72 + 64 + ij] = sb;
  }
  __syncthreads();
  if (threadIdx.x < 72) {
    const int t = threadIdx.x;
    const double s = (red[t] + red[72 + t]) + (red[144 + t] + red[216 + t]);
    if (t < 64) {
      const int i = t >> 3, j = t & 7;
      if (i < nv && j < nv) Mout[((size_t)cell * nv + i) * nv + j] = s;
    } else if (t - 64 < nv) {
      bout[(size_t)cell * nv + (t - 64)] = s;
    }
  }
}

__global__ void q_reconstruct_kernel(int nv, int nn, int n_cells, const double* __restrict__ Psi,
                                     const double* __restrict__ w, double* __restrict__ u) {
  const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (size_t)n_cells * nn) return;
  const size_t cell = idx / nn, node = idx % nn;
  double s = 0.0;
  for (int k = 0; k < nv; ++k) s = s + w[cell * nv + k] * Psi[(cell * nv + k) * nn + node];
  u[idx] = s;
}

// ---------------------------------------------------------------------------
// Coarse system: stencil assembly by gather (deterministic) and a
// single-workgroup Jacobi-PCG in FP64.
// ---------------------------------------------------------------------------
__global__ void coarse_assemble_kernel(int dim, int nc, const double* __restrict__ M,
                                       const double* __restrict__ b, double* __restrict__ K,
                                       double* __restrict__ rhs) {
  const int nst = (dim == 2) ? 9 : 27, nv = 1 << dim;
  int nnode = 1;
  for (int d = 0; d < dim; ++d) nnode *= nc + 1;
  // one thread per (stencil entry, node), nodes fastest: coalesced stores
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nnode * nst) return;
  const int e = (int)(t / nnode), node = (int)(t - (long long)e * nnode);
  int cn[3] = {0, 0, 0}, r = node;
  bool bnd = false;
  for (int d = 0; d < dim; ++d) {
    cn[d] = r % (nc + 1);
    r /= nc + 1;
    if (cn[d] == 0 || cn[d] == nc) bnd = true;
  }
  double rs = 0.0;
  {
    int dd[3] = {0, 0, 0}, rr = e;
    bool nbnd = false, inside = true;
    for (int d = 0; d < dim; ++d) {
      dd[d] = rr % 3 - 1;
      rr /= 3;
      const int m = cn[d] + dd[d];
      if (m < 0 || m > nc) inside = false;
      if (m == 0 || m == nc) nbnd = true;
    }
    double s = 0.0;
    if (inside && !bnd && !nbnd) {
      for (int al = 0; al < nv; ++al) {
        bool ok = true;
        int z = 0, mul = 1, a = 0, bb = 0;
        for (int d = 0; d < dim; ++d) {
          const int ad = (al >> d) & 1, bd = ad + dd[d];
          const int zd = cn[d] - ad;
          if (bd < 0 || bd > 1 || zd < 0 || zd >= nc) { ok = false; break; }
          z += zd * mul;
          mul *= nc;
          a |= ad << d;
          bb |= bd << d;
        }
        if (ok) s += M[((size_t)z * nv + a) * nv + bb];
      }
    } else if (bnd && e == (nst - 1) / 2) {
      s = 1.0;  // constrained node: unit row, zero right-hand side
    }
    K[(size_t)e * nnode + node] = s;  // [entry][node]
  }
  if (e == 0 && !bnd) {
    for (int al = 0; al < nv; ++al) {
      bool ok = true;
      int z = 0, mul = 1;
      for (int d = 0; d < dim; ++d) {
        const int zd = cn[d] - ((al >> d) & 1);
        if (zd < 0 || zd >= nc) { ok = false; break; }
        z += zd * mul;
        mul *= nc;
      }
      if (ok) rs += b[(size_t)z * nv + al];
    }
  }
  if (e == 0) rhs[node] = rs;
}

__device__ double block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  const int nw = blockDim.x >> 6;
  for (int i = 0; i < nw; ++i) s += red[i];
  return s;
}

__global__ __launch_bounds__(1024) void coarse_cg_kernel(int dim, int nc, const double* __restrict__ K,
                                                         const double* __restrict__ rhs,
                                                         double* __restrict__ u, double* r,
                                                         double* z, double* pv, double* qv,
                                                         double abs_tol, double rel_tol,
                                                         int max_iter, int* iters_out,
                                                         double* res_out) {
  __shared__ double red[16];
  const int nst = (dim == 2) ? 9 : 27, center = (nst - 1) / 2;
  int nnode = 1;
  for (int d = 0; d < dim; ++d) nnode *= nc + 1;
  const int S = nc + 1;
  double bb = 0.0;
  for (int i = threadIdx.x; i < nnode; i += blockDim.x) {
    u[i] = 0.0;
    r[i] = rhs[i];
    pv[i] = 0.0;
    bb += rhs[i] * rhs[i];
  }
  const double bnorm2 = block_sum(bb, red);
  const double stop = fmax(abs_tol * abs_tol, rel_tol * rel_tol * bnorm2);
  double rnorm2 = bnorm2, rz_old = 1.0;
  int it = 0;
  while (rnorm2 > stop && it < max_iter) {
    double rz = 0.0;
    for (int i = threadIdx.x; i < nnode; i += blockDim.x) {
      const double zi = r[i] / K[(size_t)center * nnode + i];
      z[i] = zi;
      rz += r[i] * zi;
    }
    rz = block_sum(rz, red);
    const double beta = (it == 0) ? 0.0 : rz / rz_old;
    rz_old = rz;
    for (int i = threadIdx.x; i < nnode; i += blockDim.x) pv[i] = z[i] + beta * pv[i];
    __syncthreads();
    double pq = 0.0;
    for (int i = threadIdx.x; i < nnode; i += blockDim.x) {
      int cn[3] = {0, 0, 0}, rr = i;
      for (int d = 0; d < dim; ++d) {
        cn[d] = rr % S;
        rr /= S;
      }
      double s = 0.0;
      for (int e = 0; e < nst; ++e) {
        int er = e, off = 0, mul = 1;
        bool ok = true;
        for (int d = 0; d < dim; ++d) {
          const int dd = er % 3 - 1;
          er /= 3;
          const int m = cn[d] + dd;
          if (m < 0 || m >= S) ok = false;
          off += dd * mul;
          mul *= S;
        }
        if (ok) s = fma(K[(size_t)e * nnode + i], pv[i + off], s);
      }
      qv[i] = s;
      pq += pv[i] * s;
    }
    pq = block_sum(pq, red);
    const double alpha = rz / pq;
    double rr2 = 0.0;
    for (int i = threadIdx.x; i < nnode; i += blockDim.x) {
      u[i] = fma(alpha, pv[i], u[i]);
      const double ri = fma(-alpha, qv[i], r[i]);
      r[i] = ri;
      rr2 += ri * ri;
    }
    rnorm2 = block_sum(rr2, red);
    ++it;
  }
  if (threadIdx.x == 0) {
    *iters_out = (rnorm2 > stop) ? -it : it;
    *res_out = sqrt(rnorm2);
  }
}

// Block sums of two values over the 16 waves of a 1024-thread workgroup;
// every thread gets both.  One barrier: a caller reuses a buffer only after a
// later barrier (two buffers in turn).
__device__ __forceinline__ void coarse_block_sum2(double& x, double& y, double (*red)[2]) {
  for (int o = 32; o > 0; o >>= 1) {
    x += __shfl_xor(x, o);
    y += __shfl_xor(y, o);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[w][0] = x;
    red[w][1] = y;
  }
  __syncthreads();
  double sx = 0.0, sy = 0.0;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    sx += red[i][0];
    sy += red[i][1];
  }
  x = sx;
  y = sy;
}

constexpr int kCoarseThreads = 1024;

// LDS bytes of coarse_pcg_kernel: the search direction over all nodes and the
// node ids of the interior nodes
inline size_t coarse_pcg_lds(int dim, int nc) {
  size_t nnode = 1, nint = 1;
  for (int d = 0; d < dim; ++d) {
    nnode *= (size_t)(nc + 1);
    nint *= (size_t)(nc - 1);
  }
  return nnode * sizeof(double) + nint * sizeof(int);
}

// Jacobi-PCG on the interior coarse nodes.  The boundary nodes carry unit
// rows and zero right-hand sides and no interior row couples to them, so
// their values stay 0 and only the interior is iterated.  One workgroup;
// thread t owns the interior nodes t, t + 1024, ...  The search direction
// over all nodes (zero on the boundary) and the node ids of the interior
// nodes sit in LDS; K is read in its [entry][node] layout, so the loads of
// one stencil entry by a wave are coalesced; the neighbour offsets are
// formed once.  u, r, z, q and the inverse diagonal (in pv) stay in global
// memory, indexed by node.  Two block sums per iteration.
template <int DIM>
__global__ __launch_bounds__(kCoarseThreads) void coarse_pcg_kernel(
    int nc, const double* __restrict__ K, const double* __restrict__ rhs, double* __restrict__ u,
    double* __restrict__ r, double* __restrict__ z, double* __restrict__ dinv,
    double* __restrict__ qv, double abs_tol, double rel_tol, int max_iter, int* iters_out,
    double* res_out) {
  constexpr int NST = (DIM == 3) ? 27 : 9, CEN = (NST - 1) / 2;
  extern __shared__ double dyn[];
  __shared__ double red[2][16][2];
  const int S = nc + 1, ni = nc - 1;
  const int nnode = (DIM == 3) ? S * S * S : S * S;
  const int nint = (DIM == 3) ? ni * ni * ni : ni * ni;
  double* P = dyn;
  int* nid = reinterpret_cast<int*>(dyn + nnode);
  const int tid = threadIdx.x;
  for (int i = tid; i < nnode; i += kCoarseThreads) {
    P[i] = 0.0;
    u[i] = 0.0;
  }
  for (int j = tid; j < nint; j += kCoarseThreads) {
    const int jx = j % ni, jr = j / ni;
    const int jy = (DIM == 3) ? jr % ni : jr, jz = (DIM == 3) ? jr / ni : 0;
    nid[j] = (jx + 1) + S * (jy + 1) + (DIM == 3 ? S * S * (jz + 1) : 0);
  }
  int off[NST];
#pragma unroll
  for (int e = 0; e < NST; ++e)
    off[e] = (e % 3 - 1) + S * ((e / 3) % 3 - 1) + (DIM == 3 ? S * S * (e / 9 - 1) : 0);
  __syncthreads();  // P zeroed, node ids
  // u = 0: r = rhs, p = z = D^-1 r
  double bb = 0.0, rz = 0.0;
  for (int j = tid; j < nint; j += kCoarseThreads) {
    const int i = nid[j];
    const double ri = rhs[i], di = 1.0 / K[(size_t)CEN * nnode + i];
    const double zi = di * ri;
    dinv[i] = di;
    r[i] = ri;
    P[i] = zi;
    bb = fma(ri, ri, bb);
    rz = fma(ri, zi, rz);
  }
  coarse_block_sum2(bb, rz, red[0]);  // also: p is in LDS
  const double stop = fmax(abs_tol * abs_tol, rel_tol * rel_tol * bb);
  double rr = bb;
  int it = 0;
  while (rr > stop && it < max_iter) {
    double pq = 0.0, dummy = 0.0;
    for (int j = tid; j < nint; j += kCoarseThreads) {
      const int i = nid[j];
      double q = 0.0;
#pragma unroll
      for (int e = 0; e < NST; ++e) q = fma(K[(size_t)e * nnode + i], P[i + off[e]], q);
      qv[i] = q;
      pq = fma(P[i], q, pq);
    }
    coarse_block_sum2(pq, dummy, red[1]);
    if (!(pq > 0.0)) break;  // not SPD: reported as not converged
    const double alpha = rz / pq;
    double rz1 = 0.0, rr1 = 0.0;
    for (int j = tid; j < nint; j += kCoarseThreads) {
      const int i = nid[j];
      u[i] = fma(alpha, P[i], u[i]);
      const double ri = fma(-alpha, qv[i], r[i]);
      const double zi = dinv[i] * ri;
      r[i] = ri;
      z[i] = zi;
      rz1 = fma(ri, zi, rz1);
      rr1 = fma(ri, ri, rr1);
    }
    coarse_block_sum2(rz1, rr1, red[0]);  // also: every thread is done reading p
    const double beta = rz1 / rz;
    rz = rz1;
    rr = rr1;
    for (int j = tid; j < nint; j += kCoarseThreads) {
      const int i = nid[j];
      P[i] = fma(beta, P[i], z[i]);
    }
    __syncthreads();
    ++it;
  }
  if (tid == 0) {
    *iters_out = (rr > stop) ? -it : it;
    *res_out = sqrt(rr);
  }
}

// The coarse Jacobi-PCG: coarse_pcg_kernel where its LDS holds the grid,
// the general one-workgroup kernel beyond (its vectors in global memory).
// Buffers as coarse_cg_kernel's (pv holds the inverse diagonal in the former).
hipError_t launch_coarse_cg(int dim, int nc, const double* K, const double* rhs, double* u,
                            double* r, double* z, double* pv, double* qv, double abs_tol,
                            double rel_tol, int max_iter, int* iters, double* res,
                            hipStream_t s) {
  const size_t lds = coarse_pcg_lds(dim, nc);
  if (nc >= 2 && lds <= 150 * 1024) {
    const void* fn = (dim == 3) ? (const void*)coarse_pcg_kernel<3> : (const void*)coarse_pcg_kernel<2>;
    if (lds > 48 * 1024) {
      const hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      if (e != hipSuccess) return e;
    }
    if (dim == 3)
      hipLaunchKernelGGL(coarse_pcg_kernel<3>, dim3(1), dim3(kCoarseThreads), lds, s, nc, K, rhs, u,
                         r, z, pv, qv, abs_tol, rel_tol, max_iter, iters, res);
    else
      hipLaunchKernelGGL(coarse_pcg_kernel<2>, dim3(1), dim3(kCoarseThreads), lds, s, nc, K, rhs, u,
                         r, z, pv, qv, abs_tol, rel_tol, max_iter, iters, res);
  } else {
    hipLaunchKernelGGL(coarse_cg_kernel, dim3(1), dim3(1024), 0, s, dim, nc, K, rhs, u, r, z, pv,
                       qv, abs_tol, rel_tol, max_iter, iters, res);
  }
  return hipGetLastError();
}

__global__ void coarse_weights_kernel(int dim, int nc, const double* __restrict__ u,
                                      double* __restrict__ w) {
  const int nv = 1 << dim;
  int ncell = 1;
  for (int d = 0; d < dim; ++d) ncell *= nc;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ncell * nv) return;
  const int cell = idx / nv, v = idx % nv;
  int r = cell, node = 0, mul = 1;
  for (int d = 0; d < dim; ++d) {
    node += (r % nc + ((v >> d) & 1)) * mul;
    r /= nc;
    mul *= nc + 1;
  }
  w[idx] = u[node];
}

}  // namespace

// ---------------------------------------------------------------------------
// Matrix-free CG (MSFEC_SOLVER_CG): the analogue of QBasis::solve_iterative
// (source/Q/q_basis.cc:335-357) with the Jacobi preconditioner in place of
// SSOR, for cells whose interior band does not fit the 64-row window.
// One workgroup of 512 threads per (cell, basis); thread t owns the interior
// rows t, t + 512, ... (at most 8).  The search direction lives in LDS over
// all fine nodes (zero on the boundary) so that K p is the 3^dim-point
// stencil of the neighbour sums of the fine-cell coefficients, recomputed
// from the coefficients in LDS at every iteration (no matrix is stored).
// The right-hand side and the diagonal come from the row table.
// ---------------------------------------------------------------------------
constexpr int kCgThreads = 512;
constexpr int kCgOwn = 8;
constexpr int kCgWaves = kCgThreads / 64;

struct CgArgs {
  double tol;
  int max_iter;
  int seg;     // rows per thread: segments of grid lines in x (cg_segment)
  int nsolve;  // bases solved per cell: 2^dim, or 2^dim - 1 when the last one
               // follows from the partition of unity (q_cg_unity_kernel)
  int* iters;  // [cells][NV][2]: iterations, state (0 converged, 1 not, 2 breakdown, 3 non-finite)
  const double* fnode;  // assembled-stencil path: nodal load [cells][nn]
};

// sum over the 64 lanes of the wave: DPP inside rows of 16, then across rows
__device__ __forceinline__ double cg_wave_sum(double v) {
  v += dpp_f64(v, 0);
  v += dpp_f64(v, 1);
  v += dpp_f64(v, 2);
  v += __shfl_xor(v, 8);
  v += __shfl_xor(v, 16);
  v += __shfl_xor(v, 32);
  return v;
}

// block sums of two values; every thread gets both.  One barrier: a caller
// reuses a buffer only after a later barrier (two buffers in turn).
__device__ __forceinline__ void cg_block_sum2(double& x, double& y, double (*red)[2]) {
  x = cg_wave_sum(x);
  y = cg_wave_sum(y);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[w][0] = x;
    red[w][1] = y;
  }
  __syncthreads();
  double sx = 0.0, sy = 0.0;
#pragma unroll
  for (int i = 0; i < kCgWaves; ++i) {
    sx += red[i][0];
    sy += red[i][1];
  }
  x = sx;
  y = sy;
}

// K v at the rows of one segment (nodes (x0 .. x0 + len - 1, y, z), all
// interior) from the nodal vector v in LDS.  Per node: the fine cells
// around it, each a_e sum_vt ks[al ^ vt] v(vertex vt) (+ the reaction mass
// when REAC), al the node's corner in the cell.  Along the segment the
// 3^dim neighbour values and the 2^dim cell coefficients slide in x: each
// node loads one new plane of each (3^(dim-1) values and 2^(dim-1)
// coefficients).  Every register index is static.
template <int DIM, bool REAC>
__device__ __forceinline__ void cg_line_apply(const double* cf, const double* rf, bool has_rf,
                                              double rc, const double (&kr)[8],
                                              const double (&mr)[8], int x0, int cy, int cz,
                                              int len, int N, int S, const double* v,
                                              double (&q)[kCgOwn]) {
  constexpr int NV = 1 << DIM;
  constexpr int NP = (DIM == 3) ? 9 : 3;  // neighbour values per x-plane
  constexpr int NC = (DIM == 3) ? 4 : 2;  // cells per x-plane
  auto plane = [&](int X, double (&w)[NP]) {
#pragma unroll
    for (int iz = 0; iz < (DIM == 3 ? 3 : 1); ++iz)
#pragma unroll
      for (int iy = 0; iy < 3; ++iy)
        w[iy + 3 * iz] = v[X + S * (cy - 1 + iy) + (DIM == 3 ? S * S * (cz - 1 + iz) : 0)];
  };
  auto cplane = [&](const double* src, int X, double (&c)[NC]) {
#pragma unroll
    for (int iz = 0; iz < (DIM == 3 ? 2 : 1); ++iz)
#pragma unroll
      for (int iy = 0; iy < 2; ++iy)
        c[iy + 2 * iz] = src[X + N * (cy - 1 + iy) + (DIM == 3 ? N * N * (cz - 1 + iz) : 0)];
  };
  // K[n][n+e] = ks[mask(e)] * (sum of the coefficients of the cells shared
  // by n and n+e): the y-z sums P of each x-plane of cells, then x.  The
  // sums of the plane behind a node are those of the plane ahead of the
  // node before it, so each node forms one plane's sums.
  auto ysums = [&](const double (&c)[NC], double (&P)[NP]) {
#pragma unroll
    for (int ez = 0; ez < (DIM == 3 ? 3 : 1); ++ez)
#pragma unroll
      for (int ey = 0; ey < 3; ++ey) {
        double v = 0.0;
#pragma unroll
        for (int iz = 0; iz < (DIM == 3 ? 2 : 1); ++iz)
#pragma unroll
          for (int iy = 0; iy < 2; ++iy) {
            // cell (y - 1 + iy) is shared with neighbour ey - 1 iff iy == ey / 2 or ey == 1
            const bool y_ok = (ey == 1) || (iy == ey / 2);
            const bool z_ok = (DIM == 2) || (ez == 1) || (iz == ez / 2);
            if (y_ok && z_ok) v += c[iy + 2 * iz];
          }
        P[ey + 3 * ez] = v;
      }
  };
  const bool crf = REAC && has_rf;
  double w0[NP], w1[NP], w2[NP], c0[NC], c1[NC], e0[NC], e1[NC], P0[NP], P1[NP];
  plane(x0 - 1, w0);
  plane(x0, w1);
  cplane(cf, x0 - 1, c0);
  if (!REAC) ysums(c0, P0);
#pragma unroll
  for (int t = 0; t < NC; ++t) e0[t] = 0.0;
  if (crf) cplane(rf, x0 - 1, e0);
#pragma unroll
  for (int sg = 0; sg < kCgOwn; ++sg) {
    if (sg < len) {
      const int X = x0 + sg;
      plane(X + 1, w2);
      cplane(cf, X, c1);
#pragma unroll
      for (int t = 0; t < NC; ++t) e1[t] = 0.0;
      if (crf) cplane(rf, X, e1);
      double acc = 0.0;
      if (!REAC) {
        ysums(c1, P1);
        double am[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) am[t] = 0.0;
#pragma unroll
        for (int e = 0; e < NP * 3; ++e) {
          const int ex = e % 3, eyz = e / 3, ey = eyz % 3, ez = eyz / 3;
          const int mask = (ex != 1 ? 1 : 0) | (ey != 1 ? 2 : 0) | (ez != 1 && DIM == 3 ? 4 : 0);
          const double sa = ex == 0 ? P0[eyz] : ex == 2 ? P1[eyz] : P0[eyz] + P1[eyz];
          const double nbv = (ex == 0 ? w0 : ex == 1 ? w1 : w2)[eyz];
          am[mask] = fma(sa, nbv, am[mask]);
        }
#pragma unroll
        for (int t = 0; t < (1 << DIM); ++t) acc = fma(kr[t], am[t], acc);
      }
#pragma unroll
      for (int al = 0; al < (REAC ? NV : 0); ++al) {
        const int ax = al & 1, ay = (al >> 1) & 1, az = (DIM == 3) ? (al >> 2) & 1 : 0;
        const int ci = (1 - ay) + (DIM == 3 ? 2 * (1 - az) : 0);  // cell (X - ax, y - ay, z - az)
        double sk = 0.0, sm = 0.0;
#pragma unroll
        for (int vt = 0; vt < NV; ++vt) {
          const int ex = (vt & 1) - ax + 1, ey = ((vt >> 1) & 1) - ay + 1;
          const int ez = (DIM == 3) ? ((vt >> 2) & 1) - az + 1 : 0;
          const double nbv = (ex == 0 ? w0 : ex == 1 ? w1 : w2)[ey + 3 * ez];
          sk = fma(kr[al ^ vt], nbv, sk);
          if (REAC) sm = fma(mr[al ^ vt], nbv, sm);
        }
        acc = fma((ax ? c0 : c1)[ci], sk, acc);
        if (REAC) acc = fma((ax ? e0 : e1)[ci] + rc, sm, acc);
      }
      q[sg] = acc;
#pragma unroll
      for (int t = 0; t < NP; ++t) {
        w0[t] = w1[t];
        w1[t] = w2[t];
      }
#pragma unroll
      for (int t = 0; t < NC; ++t) {
        c0[t] = c1[t];
        e0[t] = e1[t];
      }
#pragma unroll
      for (int t = 0; t < NP; ++t) P0[t] = P1[t];
    }
  }
}

// Segment length of the CG workgroup: the shortest (at most kCgOwn rows)
// with which the lines of the grid in x, cut into segments, need at most one
// thread each; 0 if there is none.
inline int cg_segment(int ni, int dim) {
  const int lines = (dim == 3) ? ni * ni : ni;
  for (int L = 1; L <= kCgOwn; ++L)
    if ((long long)lines * ((ni + L - 1) / L) <= kCgThreads) return L;
  return 0;
}

// interior row jn -> grid coordinates (1..N-1 per dimension)
template <int DIM>
__device__ __forceinline__ void cg_coords(int jn, int ni, int* c) {
  c[2] = 0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    c[d] = jn % ni + 1;
    jn /= ni;
  }
}

template <int DIM, bool REAC>
__global__ __launch_bounds__(kCgThreads) void q_cg_kernel(BandArgs a, CgArgs g) {
  using D = Dirs<DIM>;
  const int cell = blockIdx.x / g.nsolve, j = blockIdx.x - cell * g.nsolve;
  const int tid = threadIdx.x;
  // LDS: coefficient | reaction (per cell) | search direction over all nodes
  extern __shared__ double dyn[];
  __shared__ double kv[8], mv[8];
  // the two block sums of an iteration use one buffer each: the barrier of
  // the other sum (or the one closing the iteration) separates the reads of
  // a buffer from its next writes
  __shared__ double red[2][kCgWaves][2];
  const bool has_rf = (a.reaction_mode == MSFEC_REACTION_CELL);
  double* cf = dyn;
  double* rf = dyn + a.nfc;
  double* pv = dyn + a.nfc * (has_rf ? 2 : 1);
  const double* cg = a.coef + (size_t)cell * a.nfc;
  for (int z = tid; z < a.nfc; z += kCgThreads) cf[z] = cg[z];
  if (has_rf) {
    const double* rg = a.reac + (size_t)cell * a.nfc;
    for (int z = tid; z < a.nfc; z += kCgThreads) rf[z] = rg[z];
  }
  for (int z = tid; z < a.nn; z += kCgThreads) pv[z] = 0.0;
  if (tid < 8) {
    const double* v = a.verts + (size_t)cell * D::NV * DIM;
    double hh[3] = {1.0, 1.0, 1.0};
#pragma unroll
    for (int d = 0; d < DIM; ++d) hh[d] = (v[(D::NV - 1) * DIM + d] - v[d]) / a.N;
    double ks, ms;
    box_tables<DIM>(tid, hh, ks, ms);
    kv[tid] = ks;
    mv[tid] = ms;
  }
  const double rc = (a.reaction_mode == MSFEC_REACTION_CONST) ? a.reaction_const : 0.0;
  const double* rows = a.Rows + (size_t)cell * a.n * kRowStride;
  // the owned segment: interior rows (x0 .. x0 + len - 1, cy, cz), consecutive
  // in the row numbering (x fastest)
  const int ni = a.ni, L = g.seg, spl = (ni + L - 1) / L;
  const int lines = (DIM == 3) ? ni * ni : ni;
  // threads past the last segment take line 0 with no rows (its loads stay
  // inside the grid)
  const int line0 = tid / spl, sx = tid - line0 * spl;
  const int len = (line0 < lines) ? min(L, ni - sx * L) : 0;
  const int line = (line0 < lines) ? line0 : 0;
  const int x0 = 1 + sx * L, cy = 1 + line % ni, cz = (DIM == 3) ? 1 + line / ni : 0;
  const int jn0 = (x0 - 1) + ni * ((cy - 1) + (DIM == 3 ? ni * (cz - 1) : 0));
  const int node0 = x0 + a.S * cy + (DIM == 3 ? a.S * a.S * cz : 0);
  // owned rows: right-hand side, inverse diagonal, iterate, residual, direction.
  // The iteration starts from the coarse Q1 function of basis j at the
  // interior nodes, which is the solution for a constant coefficient: only
  // the coefficient's variation is left for CG to resolve.
  double b[kCgOwn], dinv[kCgOwn], x[kCgOwn], r[kCgOwn], p[kCgOwn];
  double bb = 0.0, rz = 0.0, rr = 0.0, bad = 0.0;
  double phi_yz = 1.0;  // the factors of phi_j in the other dimensions
  {
    const double invN = 1.0 / (double)a.N;
    const double ty = cy * invN, tz = cz * invN;
    phi_yz = ((j >> 1) & 1) ? ty : 1.0 - ty;
    if (DIM == 3) phi_yz *= ((j >> 2) & 1) ? tz : 1.0 - tz;
#pragma unroll
    for (int t = 0; t < kCgOwn; ++t) {
      const bool ok = t < len;
      const double tx = (x0 + t) * invN;
      b[t] = ok ? rows[(size_t)(jn0 + t) * kRowStride + 16 + j] : 0.0;
      dinv[t] = ok ? 1.0 / rows[(size_t)(jn0 + t) * kRowStride] : 0.0;
      x[t] = ok ? ((j & 1) ? tx : 1.0 - tx) * phi_yz : 0.0;
      if (ok && !(dinv[t] > 0.0)) bad = 1.0;  // a diagonal entry <= 0: not SPD
    }
  }
  __syncthreads();  // coefficients, box tables and the zeroed vector
  double kr[8], mr[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    kr[t] = kv[t];
    mr[t] = mv[t];
  }
#pragma unroll
  for (int t = 0; t < kCgOwn; ++t)
    if (t < len) pv[node0 + t] = x[t];
  __syncthreads();
  {
    double q[kCgOwn];
    cg_line_apply<DIM, REAC>(cf, rf, has_rf, rc, kr, mr, x0, cy, cz, len, a.N, a.S, pv, q);
#pragma unroll
    for (int t = 0; t < kCgOwn; ++t) {
      r[t] = (t < len) ? b[t] - q[t] : 0.0;
      p[t] = dinv[t] * r[t];
      bb = fma(b[t], b[t], bb);
      rr = fma(r[t], r[t], rr);
      rz = fma(r[t], p[t], rz);
    }
  }
  __syncthreads();  // every thread is done reading the initial iterate
#pragma unroll
  for (int t = 0; t < kCgOwn; ++t)
    if (t < len) pv[node0 + t] = p[t];
  // red[0] first: the loop's first sum writes red[0] again, which the
  // barrier of the red[1] sum orders after every read of this one
  cg_block_sum2(rr, bad, red[0]);  // also: p is in LDS
  cg_block_sum2(bb, rz, red[1]);
  const double stop = g.tol * g.tol * bb;
  int it = 0, state = (bad > 0.0) ? 2 : 0;
  while (state == 0 && rr > stop) {
    if (it >= g.max_iter) {
      state = 1;
      break;
    }
    double q[kCgOwn];
    cg_line_apply<DIM, REAC>(cf, rf, has_rf, rc, kr, mr, x0, cy, cz, len, a.N, a.S, pv, q);
    double pq = 0.0, dummy = 0.0;
#pragma unroll
    for (int t = 0; t < kCgOwn; ++t)
      if (t < len) pq = fma(p[t], q[t], pq);
    cg_block_sum2(pq, dummy, red[0]);
    if (!(pq > 0.0)) {
      state = 2;
      break;
    }
    const double alpha = rz / pq;
    double rz1 = 0.0, rr1 = 0.0;
#pragma unroll
    for (int t = 0; t < kCgOwn; ++t) {
      if (t < len) {
        x[t] = fma(alpha, p[t], x[t]);
        r[t] = fma(-alpha, q[t], r[t]);
        rz1 = fma(r[t], dinv[t] * r[t], rz1);
        rr1 = fma(r[t], r[t], rr1);
      }
    }
    cg_block_sum2(rz1, rr1, red[1]);  // also: every thread is done reading p
    const double beta = rz1 / rz;
    rz = rz1;
    rr = rr1;
#pragma unroll
    for (int t = 0; t < kCgOwn; ++t) {
      if (t < len) {
        p[t] = fma(beta, p[t], dinv[t] * r[t]);
        pv[node0 + t] = p[t];
      }
    }
    __syncthreads();
    ++it;
  }
  if (!isfinite(rr) || !isfinite(bb)) state = 3;
  // the true residual of the returned solution, for the energy form of S
  __syncthreads();  // an early exit leaves threads that may still read p
#pragma unroll
  for (int t = 0; t < kCgOwn; ++t)
    if (t < len) pv[node0 + t] = x[t];
  __syncthreads();
  double q[kCgOwn];
  cg_line_apply<DIM, REAC>(cf, rf, has_rf, rc, kr, mr, x0, cy, cz, len, a.N, a.S, pv, q);
  double* psi = a.Psi + ((size_t)cell * D::NV + j) * a.nn;
#pragma unroll
  for (int t = 0; t < kCgOwn; ++t) {
    if (t < len) {
      const size_t jn = (size_t)cell * a.n + jn0 + t;
      a.X[jn * kMaxNb + j] = x[t];
      a.Rres[jn * kMaxNb + j] = b[t] - q[t];
      psi[node0 + t] = x[t];
    }
  }
  if (tid == 0) {
    g.iters[((size_t)cell * D::NV + j) * 2] = it;
    g.iters[((size_t)cell * D::NV + j) * 2 + 1] = state;
  }
}

// The last basis of a cell from the others (no reaction term): the lifted
// values sum to 1 on the boundary and K 1 = 0, so sum_i psi_i = 1 and
//   x_last = 1 - sum_i x_i,   r_last = R_last - K x_last = -sum_i r_i
// (the right-hand sides sum to K_II 1).  One thread per interior row; the
// row's first thread also records the largest iteration count and the worst
// state of the solved bases for the last one.
template <int DIM>
__global__ __launch_bounds__(256) void q_cg_unity_kernel(BandArgs a, CgArgs g, int n_cells) {
  using D = Dirs<DIM>;
  constexpr int L = D::NV - 1;
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long long)n_cells * a.n) return;
  const int cell = (int)(idx / a.n), jn = (int)(idx - (long long)cell * a.n);
  const size_t base = ((size_t)cell * a.n + jn) * kMaxNb;
  double xs = 0.0, rs = 0.0;
#pragma unroll
  for (int i = 0; i < L; ++i) {
    xs += a.X[base + i];
    rs += a.Rres[base + i];
  }
  a.X[base + L] = 1.0 - xs;
  a.Rres[base + L] = -rs;
  int c[3];
  cg_coords<DIM>(jn, a.ni, c);
  const int node = c[0] + a.S * (c[1] + (DIM == 3 ? a.S * c[2] : 0));
  a.Psi[((size_t)cell * D::NV + L) * a.nn + node] = 1.0 - xs;
  if (jn == 0) {
    int it = 0, st = 0;
#pragma unroll
    for (int i = 0; i < L; ++i) {
      it = max(it, g.iters[((size_t)cell * D::NV + i) * 2]);
      st = max(st, g.iters[((size_t)cell * D::NV + i) * 2 + 1]);
    }
    g.iters[((size_t)cell * D::NV + L) * 2] = it;
    g.iters[((size_t)cell * D::NV + L) * 2 + 1] = st;
  }
}

// Element matrix and right-hand side of a CG cell: S in its energy form
// S[i][k] = R_i^T x_k + x_i^T r_k (r = R - K x, exact up to the second order
// in the error of x), symmetrised; M = P - S and b = b_B + sum_j x_j F_j as in
// the band kernels.  Thread (i = tid & 7, row group tid >> 3) sums row i of S
// and b_i over every 128th interior row: the 8 threads of a row read its
// 8-wide blocks of x, r and R once between them, and a wave covers 8 rows a
// step.  The row groups add up across lanes (shuffles) and waves (LDS).
constexpr int kFinWaves = 16;

template <int DIM>
__global__ __launch_bounds__(64 * kFinWaves) void q_cg_finish_kernel(BandArgs a, CgArgs g) {
  using D = Dirs<DIM>;
  constexpr int kGroups = 8 * kFinWaves;
  const int cell = blockIdx.x, tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6, i = tid & 7;
  __shared__ double part[kFinWaves][8][9];  // per wave: S[i][0..7] | b_i
  double s[8], bi = 0.0;
#pragma unroll
  for (int k = 0; k < 8; ++k) s[k] = 0.0;
  if (i < D::NV) {
    const double* rows = a.Rows + (size_t)cell * a.n * kRowStride;
    const double* X = a.X + (size_t)cell * a.n * kMaxNb;
    const double* Rr = a.Rres + (size_t)cell * a.n * kMaxNb;
#pragma unroll 2
    for (int jn = tid >> 3; jn < a.n; jn += kGroups) {
      const double Ri = rows[(size_t)jn * kRowStride + 16 + i];
      const double Fj = rows[(size_t)jn * kRowStride + 14];
      const double* xr = X + (size_t)jn * kMaxNb;
      const double* rr = Rr + (size_t)jn * kMaxNb;
      const double xi = xr[i];
#pragma unroll
      for (int k = 0; k < D::NV; ++k) s[k] = fma(xi, rr[k], fma(Ri, xr[k], s[k]));
      bi = fma(xi, Fj, bi);
    }
  }
  // the 8 row groups of a wave (lanes i, i + 8, ..., i + 56)
#pragma unroll
  for (int o = 8; o < 64; o <<= 1) {
#pragma unroll
    for (int k = 0; k < 8; ++k) s[k] += __shfl_xor(s[k], o);
    bi += __shfl_xor(bi, o);
  }
  if (lane < 8) {
#pragma unroll
    for (int k = 0; k < 8; ++k) part[w][lane][k] = s[k];
    part[w][lane][8] = bi;
  }
  __syncthreads();
  if (w == 0) {
    const int mi = lane >> 3, mk = lane & 7;  // lane of S[mi][mk] in fast_finish
    double sik = 0.0, ski = 0.0, bint = 0.0;
#pragma unroll
    for (int v = 0; v < kFinWaves; ++v) {
      sik += part[v][mi][mk];
      ski += part[v][mk][mi];
      bint += part[v][lane & 7][8];  // b_lane for the lanes < 2^dim
    }
    fast_finish<DIM>(a, cell, 0.5 * (sik + ski), bint);
    if (lane == 0) {
      int st = MSFEC_CELL_OK;
      for (int jj = 0; jj < D::NV; ++jj) {
        const int state = g.iters[((size_t)cell * D::NV + jj) * 2 + 1];
        if (state == 3) st = MSFEC_CELL_NONFINITE;
        else if (state == 2 && st != MSFEC_CELL_NONFINITE) st = MSFEC_CELL_NOT_SPD;
        else if (state == 1 && st == MSFEC_CELL_OK) st = MSFEC_CELL_NOT_CONVERGED;
      }
      a.status[cell] = st;
    }
  }
}

// ---------------------------------------------------------------------------
// CG on the assembled stencil: tensor coefficients at the Gauss points and /
// or loads per Gauss point (the reference's Diffusion_A and
// RightHandSideParsed evaluated at the QGauss(2) points, q_basis.cc:196-220),
// solved matrix-free beyond the band window.  q_assemble_kernel has written
// the lower stencil Slo of every node (diagonal first, then the lower
// directions t with full-grid offsets off_full[t]), the nodal load F and the
// lifting right-hand sides Rhs.  One workgroup per (cell, basis) as in
// q_cg_kernel: the search direction over all nodes in LDS (zero on the
// boundary), every thread a segment of a grid line in x; K p at node n is
// Slo[n][0] p[n] + sum_t Slo[n][t] p[n - off_t] + Slo[n + off_t][t] p[n + off_t]
// with the stencil rows read from global memory (L2) each iteration.
// ---------------------------------------------------------------------------
template <int DIM>
__device__ __forceinline__ double cg_stencil_row(const double* __restrict__ slo, const int* off,
                                                 int node, const double* v) {
  using D = Dirs<DIM>;
  double acc = slo[(size_t)node * D::NLOW] * v[node];
#pragma unroll
  for (int t = 1; t < D::NLOW; ++t) {
    const int o = off[t];
    acc = fma(slo[(size_t)node * D::NLOW + t], v[node - o], acc);
    acc = fma(slo[(size_t)(node + o) * D::NLOW + t], v[node + o], acc);
  }
  return acc;
}

template <int DIM>
__global__ __launch_bounds__(kCgThreads) void q_cg_stencil_kernel(BandArgs a, CgArgs g, QConf c) {
  using D = Dirs<DIM>;
  const int cell = blockIdx.x / g.nsolve, j = blockIdx.x - cell * g.nsolve;
  const int tid = threadIdx.x;
  extern __shared__ double dyn[];  // search direction over all nodes
  __shared__ double red[2][kCgWaves][2];
  __shared__ int off[D::NLOW];
  double* pv = dyn;
  if (tid < D::NLOW) off[tid] = c.off_full[tid];
  for (int z = tid; z < a.nn; z += kCgThreads) pv[z] = 0.0;
  const double* slo = a.Slo + (size_t)cell * a.nn * D::NLOW;
  const int ni = a.ni, L = g.seg, spl = (ni + L - 1) / L;
  const int lines = (DIM == 3) ? ni * ni : ni;
  const int line0 = tid / spl, sx = tid - line0 * spl;
  const int len = (line0 < lines) ? min(L, ni - sx * L) : 0;
  const int line = (line0 < lines) ? line0 : 0;
  const int x0 = 1 + sx * L, cy = 1 + line % ni, cz = (DIM == 3) ? 1 + line / ni : 0;
  const int jn0 = (x0 - 1) + ni * ((cy - 1) + (DIM == 3 ? ni * (cz - 1) : 0));
  const int node0 = x0 + a.S * cy + (DIM == 3 ? a.S * a.S * cz : 0);
  double b[kCgOwn], dinv[kCgOwn], x[kCgOwn], r[kCgOwn], p[kCgOwn];
  double bb = 0.0, rz = 0.0;
#pragma unroll
  for (int t = 0; t < kCgOwn; ++t) {
    const bool ok = t < len;
    b[t] = ok ? a.Rhs[((size_t)cell * a.n + jn0 + t) * kMaxNb + j] : 0.0;
    dinv[t] = ok ? 1.0 / slo[(size_t)(node0 + t) * D::NLOW] : 0.0;
    x[t] = 0.0;
    r[t] = b[t];
    p[t] = dinv[t] * r[t];
    bb = fma(b[t], b[t], bb);
    rz = fma(r[t], p[t], rz);
  }
  __syncthreads();  // zeroed vector, offsets
#pragma unroll
  for (int t = 0; t < kCgOwn; ++t)
    if (t < len) pv[node0 + t] = p[t];
  cg_block_sum2(bb, rz, red[1]);  // also: p is in LDS
  double rr = bb;
  const double stop = g.tol * g.tol * bb;
  int it = 0, state = 0;
  while (rr > stop) {
    if (it >= g.max_iter) {
      state = 1;
      break;
    }
    double q[kCgOwn];
#pragma unroll
    for (int t = 0; t < kCgOwn; ++t) q[t] = (t < len) ? cg_stencil_row<DIM>(slo, off, node0 + t, pv) : 0.0;
    double pq = 0.0, dummy = 0.0;
#pragma unroll
    for (int t = 0; t < kCgOwn; ++t)
      if (t < len) pq = fma(p[t], q[t], pq);
    cg_block_sum2(pq, dummy, red[0]);
    if (!(pq > 0.0)) {
      state = 2;
      break;
    }
    const double alpha = rz / pq;
    double rz1 = 0.0, rr1 = 0.0;
#pragma unroll
    for (int t = 0; t < kCgOwn; ++t) {
      if (t < len) {
        x[t] = fma(alpha, p[t], x[t]);
        r[t] = fma(-alpha, q[t], r[t]);
        rz1 = fma(r[t], dinv[t] * r[t], rz1);
        rr1 = fma(r[t], r[t], rr1);
      }
    }
    cg_block_sum2(rz1, rr1, red[1]);  // also: every thread is done reading p
    const double beta = rz1 / rz;
    rz = rz1;
    rr = rr1;
#pragma unroll
    for (int t = 0; t < kCgOwn; ++t) {
      if (t < len) {
        p[t] = fma(beta, p[t], dinv[t] * r[t]);
        pv[node0 + t] = p[t];
      }
    }
    __syncthreads();
    ++it;
  }
  if (!isfinite(rr) || !isfinite(bb)) state = 3;
  // the true residual (the last basis of a cell without reaction follows
  // from the partition of unity, q_cg_unity_kernel, which sums it)
  __syncthreads();  // an early exit leaves threads that may still read p
#pragma unroll
  for (int t = 0; t < kCgOwn; ++t)
    if (t < len) pv[node0 + t] = x[t];
  __syncthreads();
  double* psi = a.Psi + ((size_t)cell * D::NV + j) * a.nn;
#pragma unroll
  for (int t = 0; t < kCgOwn; ++t) {
    if (t < len) {
      const size_t jn = (size_t)cell * a.n + jn0 + t;
      a.X[jn * kMaxNb + j] = x[t];
      a.Rres[jn * kMaxNb + j] = b[t] - cg_stencil_row<DIM>(slo, off, node0 + t, pv);
      psi[node0 + t] = x[t];
    }
  }
  if (tid == 0) {
    g.iters[((size_t)cell * D::NV + j) * 2] = it;
    g.iters[((size_t)cell * D::NV + j) * 2 + 1] = state;
  }
}

// Element matrix of a cell solved on the assembled stencil, with the whole
// assembled matrix as the reference forms it (q_basis.cc:262-300):
// M[i][k] = psi_i^T K psi_k, b[i] = psi_i^T F.  Second order in the error of
// the interior solution: psi_k is K-harmonic at the interior rows and the
// error vanishes on the boundary.  Psi is completed on the boundary with the
// lifted coarse shape functions first; then every thread takes every 512th
// node, forms (K psi_k)(node) for the 2^dim bases and accumulates the 64
// products and the 8 loads, summed over the workgroup at the end.  512
// threads: the 72 accumulators need more than the 128 registers a thread of
// a 1024-thread workgroup may hold.
constexpr int kTfThreads = 512;

template <int DIM>
__global__ __launch_bounds__(kTfThreads) void q_cg_tensor_finish_kernel(BandArgs a, CgArgs g,
                                                                        QConf c) {
  using D = Dirs<DIM>;
  const int cell = blockIdx.x, tid = threadIdx.x;
  __shared__ double part[kTfThreads / 64][72];
  __shared__ int off[D::NLOW];
  if (tid < D::NLOW) off[tid] = c.off_full[tid];
  double* psi = a.Psi + (size_t)cell * D::NV * a.nn;
  const double* slo = a.Slo + (size_t)cell * a.nn * D::NLOW;
  // boundary values of every basis
  for (int node = tid; node < a.nn; node += kTfThreads) {
    int m[3] = {0, 0, 0}, rr = node;
    bool bnd = false;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      m[d] = rr % a.S;
      rr /= a.S;
      bnd = bnd || m[d] == 0 || m[d] == a.N;
    }
    if (bnd)
#pragma unroll
      for (int k = 0; k < D::NV; ++k) psi[(size_t)k * a.nn + node] = lift_value(c, k, m);
  }
  __threadfence_block();
  __syncthreads();
  double acc[D::NV * D::NV + D::NV];
#pragma unroll
  for (int t = 0; t < D::NV * D::NV + D::NV; ++t) acc[t] = 0.0;
  const double* F = g.fnode + (size_t)cell * a.nn;
  for (int node = tid; node < a.nn; node += kTfThreads) {
    int m[3] = {0, 0, 0}, rr = node;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      m[d] = rr % a.S;
      rr /= a.S;
    }
    double kp[D::NV];
#pragma unroll
    for (int k = 0; k < D::NV; ++k) kp[k] = slo[(size_t)node * D::NLOW] * psi[(size_t)k * a.nn + node];
    // the 3^dim - 1 neighbours inside the grid: lower ones from the own row,
    // upper ones from the neighbour's row
#pragma unroll
    for (int t = 1; t < D::NLOW; ++t) {
      const int o = off[t];
      const int e = c.dir_low[t];  // lower direction e: node - o is its neighbour
      bool lo_ok = true, up_ok = true;
      int ee = e;
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        const int dd = ee % 3 - 1;  // offset of the lower neighbour in direction d is dd
        ee /= 3;
        if (m[d] + dd < 0 || m[d] + dd > a.N) lo_ok = false;
        if (m[d] - dd < 0 || m[d] - dd > a.N) up_ok = false;
      }
      if (lo_ok) {
        const double kl = slo[(size_t)node * D::NLOW + t];
#pragma unroll
        for (int k = 0; k < D::NV; ++k) kp[k] = fma(kl, psi[(size_t)k * a.nn + node - o], kp[k]);
      }
      if (up_ok) {
        const double ku = slo[(size_t)(node + o) * D::NLOW + t];
#pragma unroll
        for (int k = 0; k < D::NV; ++k) kp[k] = fma(ku, psi[(size_t)k * a.nn + node + o], kp[k]);
      }
    }
    double pn[D::NV];
#pragma unroll
    for (int i = 0; i < D::NV; ++i) pn[i] = psi[(size_t)i * a.nn + node];
    const double fn = F[node];
#pragma unroll
    for (int i = 0; i < D::NV; ++i) {
#pragma unroll
      for (int k = 0; k < D::NV; ++k) acc[i * D::NV + k] = fma(pn[i], kp[k], acc[i * D::NV + k]);
      acc[D::NV * D::NV + i] = fma(pn[i], fn, acc[D::NV * D::NV + i]);
    }
  }
  const int w = tid >> 6, lane = tid & 63;
#pragma unroll
  for (int t = 0; t < D::NV * D::NV + D::NV; ++t) {
    const double v = cg_wave_sum(acc[t]);
    if (lane == 0) part[w][t] = v;
  }
  __syncthreads();
  if (tid < D::NV * D::NV + D::NV) {
    double v = 0.0;
#pragma unroll
    for (int u = 0; u < kTfThreads / 64; ++u) v += part[u][tid];
    if (tid < D::NV * D::NV) a.M[(size_t)cell * D::NV * D::NV + tid] = v;
    else a.bvec[(size_t)cell * D::NV + tid - D::NV * D::NV] = v;
  }
  if (tid == 0) {
    int st = MSFEC_CELL_OK;
    for (int jj = 0; jj < D::NV; ++jj) {
      const int state = g.iters[((size_t)cell * D::NV + jj) * 2 + 1];
      if (state == 3) st = MSFEC_CELL_NONFINITE;
      else if (state == 2 && st != MSFEC_CELL_NONFINITE) st = MSFEC_CELL_NOT_SPD;
      else if (state == 1 && st == MSFEC_CELL_OK) st = MSFEC_CELL_NOT_CONVERGED;
    }
    a.status[cell] = st;
  }
}

// ===========================================================================
// Host side: handle, error reporting, launches
// ===========================================================================
// the last error message of the calling thread, shared with the Q_Ned
// entry points (msfec_ned.hip)
namespace {
// q_blk_factor_kernel<double, DIM, BK_FINAL64> with a window of 16 nt slots
void launch_factor64(int dim, int nt, int n_cells, hipStream_t s, const BandArgs& a) {
#define F64(D, NT) \
  hipLaunchKernelGGL((q_blk_factor_kernel<double, D, BK_FINAL64, false, NT>), dim3(n_cells), dim3(64), 0, s, a)
  if (dim == 3) {
    if (nt == 1) F64(3, 1);
    else if (nt == 2) F64(3, 2);
    else F64(3, 4);
  } else {
    if (nt == 1) F64(2, 1);
    else if (nt == 2) F64(2, 2);
    else F64(2, 4);
  }
#undef F64
}
}  // namespace

namespace msfec_detail {
thread_local std::string g_last_error;

// Panel storage per cell of band_ldl_solve, in doubles.
size_t band_ldl_panel_doubles(int n) { return (size_t)panel_quads(n) * 256; }
size_t band_ldl_z_doubles(int n) { return (size_t)panel_quads(n) * 32; }

// The matrix-core band LDL^T of the Q_Ned interior systems: rows
// [cells][n][24] (16 lower entries, diagonal first, with their offsets in
// loff [n][16]; 8 right-hand sides), the solutions to X [cells][n][8], the
// cell status (MSFEC_CELL_NOT_SPD for a zero pivot) to status.
hipError_t band_ldl_solve(int n_cells, int n, int bw, const double* rows, const int* loff,
                          double* pan, double* zst, double* X, int* status, hipStream_t s) {
  const int nt = band_tile_rows(bw, 0);
  if (nt == 0) return hipErrorInvalidValue;
  BandArgs a{};
  a.n = n;
  a.nq = panel_quads(n);
  a.Rows = rows;
  a.loff = loff;
  a.pan = pan;
  a.zst = zst;
  a.X = X;
  a.status = status;
#define LDL(NT) \
  hipLaunchKernelGGL((q_blk_factor_kernel<double, 2, BK_X0, true, NT>), dim3(n_cells), dim3(64), 0, s, a)
  if (nt == 1) LDL(1);
  else if (nt == 2) LDL(2);
  else LDL(4);
#undef LDL
  return hipGetLastError();
}

// Scratch per cell of band_ldl_mixed_solve beyond band_ldl_solve's, in
// doubles: the residual and the correction ([n][8] each), the last correction
// ratio and the refinement flag.
size_t band_ldl_mixed_doubles(int n) { return (size_t)n * kMaxNb * 2 + 2; }

// The transposed offset table of ldl_resid_kernel from the lower offsets
// loff [n][16] (host side): entry (j, k) = (i << 4 | t) for the k-th lower
// entry t of a row i > j with i - loff[i][t] = j, -1 unused.  Empty if a row
// has more than 16 upper entries.
std::vector<int> band_upper_table(const std::vector<int>& loff, int n) {
  std::vector<int> ut((size_t)std::max(n, 1) * 16, -1);
  std::vector<int> cnt(std::max(n, 1), 0);
  for (int i = 0; i < n; ++i)
    for (int t = 0; t < 16; ++t) {
      const int o = loff[(size_t)i * 16 + t];
      if (o <= 0) continue;
      const int j = i - o;
      if (j < 0 || cnt[j] >= 16) return {};
      ut[(size_t)j * 16 + cnt[j]++] = (i << 4) | t;
    }
  return ut;
}

// band_ldl_solve with the factor in FP32 (panels pan and zst as FP32) and
// FP64 refinement: X to FP64 accuracy, cells that do not converge in
// max_refine passes flagged MSFEC_CELL_NOT_CONVERGED.  scratch: n_cells x
// band_ldl_mixed_doubles(n) doubles.
hipError_t band_ldl_mixed_solve(int n_cells, int n, int bw, const double* rows, const int* loff,
                                const int* utab, double* pan, double* zst, double* X,
                                double* scratch, int* status, int max_refine, hipStream_t s) {
  const int nt = band_tile_rows(bw, 0);
  if (nt == 0) return hipErrorInvalidValue;
  double* Rres = scratch;
  double* dX = Rres + (size_t)n_cells * n * kMaxNb;
  double* qprev = dX + (size_t)n_cells * n * kMaxNb;
  int* conv = reinterpret_cast<int*>(qprev + n_cells);
  BandArgs a{};
  a.n = n;
  a.nq = panel_quads(n);
  a.Rows = rows;
  a.loff = loff;
  a.pan = pan;
  a.zst = zst;
  a.X = X;
  a.Rres = Rres;
  a.conv = conv;
  a.status = status;
#define LDLF(NT) \
  hipLaunchKernelGGL((q_blk_factor_kernel<float, 2, BK_X0, true, NT>), dim3(n_cells), dim3(64), 0, s, a)
  if (nt == 1) LDLF(1);
  else if (nt == 2) LDLF(2);
  else LDLF(4);
#undef LDLF
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(ldl_conv_init_kernel, dim3((n_cells + 255) / 256), dim3(256), 0, s, status,
                     conv, qprev, n_cells);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  const size_t rsm = (size_t)n * kMaxNb * sizeof(double);
  const int passes = max_refine > 0 ? max_refine : 5;
  for (int pass = 1; pass <= passes; ++pass) {
    a.last_pass = (pass == passes) ? 1 : 0;
    a.X = X;
    hipLaunchKernelGGL(ldl_resid_kernel, dim3(n_cells), dim3(256), rsm, s, a, utab);
    BandArgs ar = a;
    ar.X = dX;  // the correction
#define LDLR(NT) hipLaunchKernelGGL((ldl_refine_kernel<NT>), dim3(n_cells), dim3(64), 0, s, ar)
    if (nt == 1) LDLR(1);
    else if (nt == 2) LDLR(2);
    else LDLR(4);
#undef LDLR
    hipLaunchKernelGGL(ldl_update_kernel, dim3(n_cells), dim3(256), 0, s, a, dX, qprev);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// Coarse assembly, Jacobi-PCG and the per-cell weights on the device for the
// element data of the nc^dim structured coarse grid (cells lexicographic):
// dM [cells][2^dim][2^dim], db [cells][2^dim].  Buffers: cK [3^dim][nodes],
// crhs [nodes], cvec [5 nodes + 1 + cells 2^dim] (u | r | z | p | q |
// final residual | weights), cit [1].  Enqueued on s.
hipError_t coarse_solve_device(int dim, int nc, const double* dM, const double* db, double* cK,
                               double* crhs, double* cvec, int* cit, double abs_tol,
                               double rel_tol, int max_iter, hipStream_t s) {
  int nnode = 1, ncell = 1;
  for (int d = 0; d < dim; ++d) {
    nnode *= nc + 1;
    ncell *= nc;
  }
  double* du = cvec;
  double* r = du + nnode;
  double* z = r + nnode;
  double* pv = z + nnode;
  double* qv = pv + nnode;
  double* res = qv + nnode;
  double* dw = res + 1;
  hipLaunchKernelGGL(coarse_assemble_kernel, dim3((nnode * (dim == 3 ? 27 : 9) + 255) / 256),
                     dim3(256), 0, s, dim, nc, dM,
                     db, cK, crhs);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = launch_coarse_cg(dim, nc, cK, crhs, du, r, z, pv, qv, abs_tol, rel_tol, max_iter, cit, res, s);
  if (e != hipSuccess) return e;
  const int nw = ncell << dim;
  hipLaunchKernelGGL(coarse_weights_kernel, dim3((nw + 255) / 256), dim3(256), 0, s, dim, nc, du, dw);
  return hipGetLastError();
}
}  // namespace msfec_detail

namespace {

using msfec_detail::g_last_error;

msfec_status fail(msfec_status st, const std::string& msg) {
  g_last_error = msg;
  return st;
}

#define MSFEC_HIP(call)                                                              \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                            \
      return fail(MSFEC_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

int ipow(int b, int e) {
  int r = 1;
  while (e-- > 0) r *= b;
  return r;
}

}  // namespace

struct msfec_q_batch {
  msfec_q_params p{};
  QConf conf{};
  int n_cells = 0;
  int bw = 0;
  size_t coef_count = 0, rhs_count = 0, reac_count = 0;
  // device buffers
  double* d_verts = nullptr;
  double* d_coef = nullptr;
  double* d_f = nullptr;
  double* d_reac = nullptr;
  double* d_rows = nullptr;
  double* d_rres = nullptr;
  double* d_pb = nullptr;
  int* d_conv = nullptr;
  int* d_cgit = nullptr;
  double* d_slo = nullptr;
  double* d_F = nullptr;
  double* d_rhs = nullptr;
  double* d_L = nullptr;
  double* d_W = nullptr;
  double* d_X = nullptr;
  double* d_psi = nullptr;
  double* d_M = nullptr;
  double* d_b = nullptr;
  int* d_status = nullptr;
  int* d_inode = nullptr;
  unsigned* d_imask = nullptr;
  size_t bytes = 0;
  bool coef_set = false, rhs_set = false, reac_set = false, computed = false;
  hipStream_t stream = nullptr;
  // per-compute kernel events (MSFEC_FLAG_TIMING): sets recorded since the
  // last msfec_q_kernel_times call, reused across calls
  std::vector<std::array<hipEvent_t, 6>> ev_sets;
  size_t ev_used = 0;
  // coarse scratch
  double* d_cK = nullptr;
  double* d_crhs = nullptr;
  double* d_cvec = nullptr;
  int* d_cit = nullptr;
  int coarse_nc = -1;
  // scalar coefficient and piecewise-constant load: one fused kernel per
  // batch; otherwise assembly tables + band kernel + element kernel
  bool fast = false;
};

namespace {

template <typename T>
msfec_status dalloc(msfec_q_batch* h, T** p, size_t count) {
  if (count == 0) count = 1;
  MSFEC_HIP(hipMalloc((void**)p, count * sizeof(T)));
  h->bytes += count * sizeof(T);
  return MSFEC_OK;
}

void free_all(msfec_q_batch* h) {
  void* ptrs[] = {h->d_pb,    h->d_conv,  h->d_cgit,  h->d_rres,  h->d_rows,  h->d_verts, h->d_coef, h->d_f,     h->d_reac,  h->d_slo,   h->d_F,
                  h->d_rhs,   h->d_L,    h->d_W,     h->d_X,     h->d_psi,   h->d_M,
                  h->d_b,     h->d_status, h->d_inode, h->d_imask, h->d_cK,  h->d_crhs,
                  h->d_cvec,  h->d_cit};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (auto& set : h->ev_sets)
    for (auto& e : set)
      if (e) (void)hipEventDestroy(e);
  h->ev_sets.clear();
}

// Inputs are copied on the handle's stream (that of its last compute call,
// the null stream before the first) and the call returns once the copy is
// done: a compute or graph replay still in flight on that stream reads the
// old buffer to the end, and device-resident data produced on that stream is
// read after its producer.
msfec_status copy_in(hipStream_t s, double* dst, const double* src, size_t count, int loc) {
  if (!src) return fail(MSFEC_ERR_INVALID, "null data pointer");
  MSFEC_HIP(hipMemcpyAsync(dst, src, count * sizeof(double),
                           loc == MSFEC_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
  MSFEC_HIP(hipStreamSynchronize(s));
  return MSFEC_OK;
}

// Outputs are copied on the handle's stream and waited for: a device
// destination holds the data when the call returns, whatever stream the
// caller reads it on next (a device-to-device hipMemcpy may return first).
msfec_status copy_out(hipStream_t s, double* dst, const double* src, size_t count, int loc) {
  MSFEC_HIP(hipMemcpyAsync(dst, src, count * sizeof(double),
                           loc == MSFEC_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
  MSFEC_HIP(hipStreamSynchronize(s));
  return MSFEC_OK;
}

}  // namespace

namespace {
BandArgs band_args(const msfec_q_batch* h) {
  const QConf& c = h->conf;
  BandArgs a{};
  a.n = c.n;
  a.ni = c.ni;
  a.N = c.N;
  a.S = c.S;
  a.nn = c.nn;
  a.nfc = c.nfc;
  for (int t = 0; t < kMaxLow; ++t) a.off_int[t] = c.off_int[t];
  a.Rows = h->d_rows;
  a.Slo = h->d_slo;
  a.Rhs = h->d_rhs;
  a.inode = h->d_inode;
  a.imask = h->d_imask;
  a.verts = h->d_verts;
  a.coef = h->d_coef;
  a.reac = h->reac_count ? h->d_reac : nullptr;
  a.fcell = h->rhs_count ? h->d_f : nullptr;
  a.reaction_mode = c.reaction_mode;
  a.rhs_mode = c.rhs_mode;
  a.reaction_const = c.reaction_const;
  a.rhs_const = c.rhs_const;
  a.Lcol = h->d_L;
  a.pan = h->d_L;
  a.zst = h->d_W;
  a.Rres = h->d_rres;
  a.conv = h->d_conv;
  a.Pb = h->d_pb;
  a.last_pass = 0;
  a.Wf = h->d_W;
  a.X = h->d_X;
  a.Psi = h->d_psi;
  a.M = h->d_M;
  a.bvec = h->d_b;
  a.status = h->d_status;
  a.nq = panel_quads(c.n);
  return a;
}

}  // namespace

namespace {
// The per-level constants of a batch: grid sizes and the stencil offsets of
// the lower neighbours on the full and the interior grid.
void fill_qconf(const msfec_q_params& p, QConf& c) {
  const int N = 1 << p.n_refine_local;
  const int ni = N - 1;
  const int nv = 1 << p.dim;
  c.dim = p.dim;
  c.N = N;
  c.S = N + 1;
  c.nn = ipow(c.S, p.dim);
  c.nfc = ipow(N, p.dim);
  c.nv = nv;
  c.ni = ni;
  c.n = ipow(ni, p.dim);
  c.nst = ipow(3, p.dim);
  c.center = (c.nst - 1) / 2;
  c.nlow = c.center + 1;
  c.coef_mode = p.coef_mode;
  c.rhs_mode = p.rhs_mode;
  c.reaction_mode = p.reaction_mode;
  c.rhs_const = p.rhs_const;
  c.reaction_const = p.reaction_const;
  c.off_full[0] = 0;
  c.off_int[0] = 0;
  c.dir_low[0] = -1;
  for (int e = 0; e < c.center; ++e) {
    int r = e, of = 0, oi = 0, mf = 1, mi = 1;
    for (int d = 0; d < p.dim; ++d) {
      const int dd = r % 3 - 1;
      r /= 3;
      of += dd * mf;
      oi += dd * mi;
      mf *= c.S;
      mi *= ni;
    }
    c.off_full[1 + e] = -of;
    c.off_int[1 + e] = -oi;
    c.dir_low[1 + e] = e;
  }
}
}  // namespace

// ---------------------------------------------------------------------------
// The allocation-free entry point (msfec_q_compute_noalloc).  Its kernels
// (q_rows_kernel + q_blk_factor_kernel<double, DIM, BK_FINAL64>, those of
// msfec_q_compute with MSFEC_SOLVER_BAND_FP64_MFMA) take their scratch -- the
// row table, P | b_B, the panels and the forward results -- from one
// file-scope device array instead of a handle's hipMalloc'd buffers.  Each
// GPU the library runs on holds its own copy of the array until the process
// ends.  Its host-side name is not a device address: the address is looked
// up once per device with hipGetSymbolAddress and only that is passed on.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = (size_t)128 << 20;  // 128 MiB per GPU
constexpr int kNoAllocMaxDevices = 64;
static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];

namespace {
std::mutex g_scratch_mu;
double* g_scratch_ptr[kNoAllocMaxDevices] = {};

// the device address of `scratch` on the current device, looked up once
msfec_status scratch_on_device(double** out) {
  int dev = -1;
  MSFEC_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocMaxDevices)
    return fail(MSFEC_ERR_UNSUPPORTED, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(MSFEC_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return MSFEC_OK;
}

inline size_t round_up32(size_t x) { return (x + 31) & ~(size_t)31; }
}  // namespace

extern "C" {

const char* msfec_last_error(void) { return g_last_error.c_str(); }

const char* msfec_version(void) { return "msfec-mi355x 0.1 (band-window FP64, gfx950)"; }

msfec_status msfec_q_create(const msfec_q_params* params, int n_cells, const double* vertices,
                            int vertices_loc, msfec_q_handle* out) {
  if (!params || !out || !vertices) return fail(MSFEC_ERR_INVALID, "null argument");
  *out = nullptr;
  const msfec_q_params& p = *params;
  if (p.dim != 2 && p.dim != 3) return fail(MSFEC_ERR_INVALID, "dim must be 2 or 3");
  if (n_cells <= 0) return fail(MSFEC_ERR_INVALID, "n_cells must be positive");
  if (p.n_refine_local < 1) return fail(MSFEC_ERR_INVALID, "n_refine_local must be >= 1");
  if (p.coef_mode != MSFEC_COEF_SCALAR_CELL && p.coef_mode != MSFEC_COEF_TENSOR_QP)
    return fail(MSFEC_ERR_INVALID, "unknown coef_mode");
  if (p.rhs_mode < MSFEC_RHS_CONST || p.rhs_mode > MSFEC_RHS_QP)
    return fail(MSFEC_ERR_INVALID, "unknown rhs_mode");
  if (p.reaction_mode < MSFEC_REACTION_NONE || p.reaction_mode > MSFEC_REACTION_CELL)
    return fail(MSFEC_ERR_INVALID, "unknown reaction_mode");
  if (p.solver != MSFEC_SOLVER_BAND_FP64 && p.solver != MSFEC_SOLVER_BAND_FP64_MFMA &&
      p.solver != MSFEC_SOLVER_BAND_MIXED && p.solver != MSFEC_SOLVER_CG)
    return fail(MSFEC_ERR_INVALID, "unknown solver");
  if (p.solver == MSFEC_SOLVER_CG && (p.cg_tol < 0.0 || p.cg_max_iter < 0))
    return fail(MSFEC_ERR_INVALID, "cg_tol and cg_max_iter must not be negative");
  if (p.max_refine < 0 || p.max_refine > 16)
    return fail(MSFEC_ERR_INVALID, "max_refine must be in 0..16");
  if (p.solver == MSFEC_SOLVER_BAND_MIXED &&
      (p.coef_mode != MSFEC_COEF_SCALAR_CELL || p.rhs_mode == MSFEC_RHS_QP))
    return fail(MSFEC_ERR_UNSUPPORTED,
                "the mixed-precision solver needs a scalar coefficient per fine cell and a "
                "constant or per-cell load");
  const int N = 1 << p.n_refine_local;
  const int ni = N - 1;
  const int bw = (p.dim == 2) ? ni + 1 : ni * ni + ni + 1;
  if (p.solver == MSFEC_SOLVER_CG) {
    const long long nint = (long long)ipow(ni, p.dim), nnod = (long long)ipow(N + 1, p.dim);
    const long long lds = ((long long)ipow(N, p.dim) *
                               (1 + (p.reaction_mode == MSFEC_REACTION_CELL ? 1 : 0)) +
                           nnod) * (long long)sizeof(double);
    if (nint > (long long)kCgThreads * kCgOwn || lds > 160 * 1024 || cg_segment(ni, p.dim) == 0)
      return fail(MSFEC_ERR_UNSUPPORTED,
                  "cell too fine for the CG workgroup (3-D: n_refine_local <= 4, 2-D: <= 6)");
  } else if (band_tile_rows(bw, 0) == 0)
    return fail(MSFEC_ERR_UNSUPPORTED,
                "interior half-bandwidth " + std::to_string(bw) +
                    " exceeds the 60 of the band solver (3-D: n_refine_local <= 3, 2-D: <= 5)");
  if (p.band_window != 0 && p.band_window != 16 && p.band_window != 32 && p.band_window != 64)
    return fail(MSFEC_ERR_INVALID, "band_window must be 0, 16, 32 or 64");
  if (p.solver != MSFEC_SOLVER_CG && band_tile_rows(bw, p.band_window) == 0)
    return fail(MSFEC_ERR_UNSUPPORTED, "a band window of " + std::to_string(p.band_window) +
                                           " slots does not hold the half-bandwidth " +
                                           std::to_string(bw) + " (it must be >= bw + 4)");
  // validate cells on the host (axis-aligned boxes)
  const int nv = 1 << p.dim;
  std::vector<double> hv((size_t)n_cells * nv * p.dim);
  if (vertices_loc == MSFEC_DEVICE) {
    if (hipMemcpy(hv.data(), vertices, hv.size() * sizeof(double), hipMemcpyDeviceToHost) !=
        hipSuccess)
      return fail(MSFEC_ERR_HIP, "copy of device vertices failed");
  } else {
    std::memcpy(hv.data(), vertices, hv.size() * sizeof(double));
  }
  for (int c = 0; c < n_cells; ++c) {
    const double* v = hv.data() + (size_t)c * nv * p.dim;
    for (int d = 0; d < p.dim; ++d) {
      const double lo = v[d], hi = v[(nv - 1) * p.dim + d];
      if (!(hi > lo) || !std::isfinite(lo) || !std::isfinite(hi))
        return fail(MSFEC_ERR_INVALID, "cell " + std::to_string(c) + " is degenerate");
      for (int k = 0; k < nv; ++k) {
        const double want = ((k >> d) & 1) ? hi : lo;
        if (std::fabs(v[k * p.dim + d] - want) > 1e-12 * (1.0 + std::fabs(want)))
          return fail(MSFEC_ERR_UNSUPPORTED,
                      "cell " + std::to_string(c) + " is not an axis-aligned box");
      }
    }
  }

  msfec_q_batch* h = new (std::nothrow) msfec_q_batch();
  if (!h) return fail(MSFEC_ERR_INVALID, "out of host memory");
  h->p = p;
  h->n_cells = n_cells;
  h->bw = bw;
  QConf& c = h->conf;
  fill_qconf(p, c);
  h->fast = (p.coef_mode == MSFEC_COEF_SCALAR_CELL) && (p.rhs_mode != MSFEC_RHS_QP);
  // the mixed solver's refinement sweep and the general path keep the 64-slot window
  c.nt = (p.solver == MSFEC_SOLVER_BAND_FP64_MFMA && h->fast) ? band_tile_rows(bw, p.band_window) : 4;
  const int nq = nv;
  h->coef_count = (p.coef_mode == MSFEC_COEF_SCALAR_CELL)
                      ? (size_t)n_cells * c.nfc
                      : (size_t)n_cells * c.nfc * nq * (p.dim == 2 ? 3 : 6);
  h->rhs_count = (p.rhs_mode == MSFEC_RHS_CONST) ? 0
                 : (p.rhs_mode == MSFEC_RHS_CELL) ? (size_t)n_cells * c.nfc
                                                  : (size_t)n_cells * c.nfc * nq;
  h->reac_count = (p.reaction_mode == MSFEC_REACTION_CELL) ? (size_t)n_cells * c.nfc : 0;

  msfec_status st = MSFEC_OK;
#define TRY(x)              \
  do {                      \
    st = (x);               \
    if (st != MSFEC_OK) {   \
      free_all(h);          \
      delete h;             \
      return st;            \
    }                       \
  } while (0)
  TRY(dalloc(h, &h->d_verts, hv.size()));
  TRY(dalloc(h, &h->d_coef, h->coef_count));
  TRY(dalloc(h, &h->d_f, h->rhs_count));
  TRY(dalloc(h, &h->d_reac, h->reac_count));
  if (h->fast) {
    TRY(dalloc(h, &h->d_rows, (size_t)n_cells * c.n * kRowStride));
    TRY(dalloc(h, &h->d_pb, (size_t)n_cells * kPStride));
  }
  if (p.solver == MSFEC_SOLVER_BAND_MIXED) {
    TRY(dalloc(h, &h->d_rres, (size_t)n_cells * c.n * kMaxNb));
    TRY(dalloc(h, &h->d_conv, (size_t)n_cells));
  }
  if (p.solver == MSFEC_SOLVER_CG) {
    TRY(dalloc(h, &h->d_rres, (size_t)n_cells * c.n * kMaxNb));
    TRY(dalloc(h, &h->d_cgit, (size_t)n_cells * nv * 2));
  }
  if (!h->fast) {
    TRY(dalloc(h, &h->d_slo, (size_t)n_cells * c.nn * c.nlow));
    TRY(dalloc(h, &h->d_F, (size_t)n_cells * c.nn));
    TRY(dalloc(h, &h->d_rhs, (size_t)n_cells * c.n * kMaxNb));
  }
  if (p.solver != MSFEC_SOLVER_CG) {
    // VALU variant: one column of 64 per pivot; MFMA variant: a 64 x 4 panel
    // per super-step of 4 pivots
    const size_t nq = (size_t)panel_quads(c.n);
    TRY(dalloc(h, &h->d_L, (size_t)n_cells * std::max((size_t)c.n * kWin, nq * 256)));
    TRY(dalloc(h, &h->d_W, (size_t)n_cells * std::max((size_t)c.n * kMaxNb, nq * 32)));
  }
  TRY(dalloc(h, &h->d_X, (size_t)n_cells * c.n * kMaxNb));
  TRY(dalloc(h, &h->d_psi, (size_t)n_cells * nv * c.nn));
  TRY(dalloc(h, &h->d_M, (size_t)n_cells * nv * nv));
  TRY(dalloc(h, &h->d_b, (size_t)n_cells * nv));
  TRY(dalloc(h, &h->d_status, (size_t)n_cells));
  TRY(dalloc(h, &h->d_inode, (size_t)c.n));
  TRY(dalloc(h, &h->d_imask, (size_t)c.n));
  // interior index -> node id and neighbour masks
  std::vector<int> inode(c.n);
  std::vector<unsigned> imask(c.n);
  for (int i = 0; i < c.n; ++i) {
    int r = i, node = 0, mul = 1, co[3] = {0, 0, 0};
    for (int d = 0; d < p.dim; ++d) {
      co[d] = r % ni;
      r /= ni;
      node += (co[d] + 1) * mul;
      mul *= c.S;
    }
    inode[i] = node;
    unsigned msk = 1u;  // diagonal
    for (int e = 0; e < c.center; ++e) {
      int rr = e;
      bool ok = true;
      for (int d = 0; d < p.dim; ++d) {
        const int dd = rr % 3 - 1;
        rr /= 3;
        const int m = co[d] + dd;
        if (m < 0 || m >= ni) ok = false;
      }
      if (ok) msk |= 1u << (1 + e);
    }
    imask[i] = msk;
  }
  if (hipMemcpy(h->d_inode, inode.data(), c.n * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->d_imask, imask.data(), c.n * sizeof(unsigned), hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(h->d_verts, hv.data(), hv.size() * sizeof(double), hipMemcpyHostToDevice) !=
          hipSuccess) {
    free_all(h);
    delete h;
    return fail(MSFEC_ERR_HIP, "upload of tables failed");
  }
#undef TRY
  *out = h;
  return MSFEC_OK;
}

msfec_status msfec_q_get_info(msfec_q_handle h, msfec_q_info* info) {
  if (!h || !info) return fail(MSFEC_ERR_INVALID, "null argument");
  info->n_cells = h->n_cells;
  info->dim = h->conf.dim;
  info->n_refine_local = h->p.n_refine_local;
  info->n_basis = h->conf.nv;
  info->n_fine_nodes = h->conf.nn;
  info->n_fine_cells = h->conf.nfc;
  info->n_interior = h->conf.n;
  info->bandwidth = h->bw;
  info->band_window = 16 * h->conf.nt;
  info->coef_count = h->coef_count;
  info->rhs_count = h->rhs_count;
  info->reaction_count = h->reac_count;
  info->device_bytes = h->bytes;
  return MSFEC_OK;
}

msfec_status msfec_q_set_coefficient(msfec_q_handle h, const double* data, size_t count, int loc) {
  if (!h) return fail(MSFEC_ERR_INVALID, "null handle");
  if (count != h->coef_count)
    return fail(MSFEC_ERR_INVALID, "coefficient count " + std::to_string(count) + " != expected " +
                                       std::to_string(h->coef_count));
  msfec_status st = copy_in(h->stream, h->d_coef, data, count, loc);
  if (st == MSFEC_OK) h->coef_set = true;
  return st;
}

msfec_status msfec_q_set_rhs(msfec_q_handle h, const double* data, size_t count, int loc) {
  if (!h) return fail(MSFEC_ERR_INVALID, "null handle");
  if (h->rhs_count == 0) return fail(MSFEC_ERR_STATE, "rhs_mode is constant; no data expected");
  if (count != h->rhs_count)
    return fail(MSFEC_ERR_INVALID, "rhs count mismatch (expected " + std::to_string(h->rhs_count) + ")");
  msfec_status st = copy_in(h->stream, h->d_f, data, count, loc);
  if (st == MSFEC_OK) h->rhs_set = true;
  return st;
}

msfec_status msfec_q_set_reaction(msfec_q_handle h, const double* data, size_t count, int loc) {
  if (!h) return fail(MSFEC_ERR_INVALID, "null handle");
  if (h->reac_count == 0) return fail(MSFEC_ERR_STATE, "reaction_mode takes no per-cell data");
  if (count != h->reac_count)
    return fail(MSFEC_ERR_INVALID, "reaction count mismatch (expected " + std::to_string(h->reac_count) + ")");
  msfec_status st = copy_in(h->stream, h->d_reac, data, count, loc);
  if (st == MSFEC_OK) h->reac_set = true;
  return st;
}

msfec_status msfec_q_compute(msfec_q_handle h, void* stream) {
  if (!h) return fail(MSFEC_ERR_INVALID, "null handle");
  if (!h->coef_set) return fail(MSFEC_ERR_STATE, "coefficient not set");
  if (h->rhs_count && !h->rhs_set) return fail(MSFEC_ERR_STATE, "rhs data not set");
  if (h->reac_count && !h->reac_set) return fail(MSFEC_ERR_STATE, "reaction data not set");
  hipStream_t s = (hipStream_t)stream;
  h->stream = s;
  const QConf& c = h->conf;
  const bool tm = (h->p.flags & MSFEC_FLAG_TIMING) != 0;
  hipEvent_t* ev = nullptr;
  if (tm) {
    if (h->ev_used == h->ev_sets.size()) {
      if (h->ev_sets.size() >= 4096) return fail(MSFEC_ERR_STATE, "too many timed computes pending");
      std::array<hipEvent_t, 6> set{};
      for (auto& e : set) MSFEC_HIP(hipEventCreate(&e));
      h->ev_sets.push_back(set);
    }
    ev = h->ev_sets[h->ev_used++].data();
  }
  const size_t cstride = h->coef_count / h->n_cells;
  const size_t fstride = h->rhs_count ? h->rhs_count / h->n_cells : 0;
  const size_t rstride = h->reac_count ? h->reac_count / h->n_cells : 0;
  const bool fast = h->fast;
  if (tm) MSFEC_HIP(hipEventRecord(ev[0], s));
  BandArgs a = band_args(h);
  if (fast) {
    const size_t rdyn = (size_t)c.nfc *
                        (1 + (c.reaction_mode == MSFEC_REACTION_CELL ? 1 : 0) +
                         (c.rhs_mode == MSFEC_RHS_CELL ? 1 : 0)) *
                        sizeof(double);
    const bool reac = c.reaction_mode != MSFEC_REACTION_NONE;
#define LAUNCH_ROWS(D, R)                                                                  \
  do {                                                                                     \
    if (rdyn > 48 * 1024)                                                                  \
      MSFEC_HIP(hipFuncSetAttribute((const void*)q_rows_kernel<D, R>,                      \
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)rdyn)); \
    hipLaunchKernelGGL((q_rows_kernel<D, R>), dim3(h->n_cells), dim3(256), rdyn, s, a);    \
  } while (0)
    if (c.dim == 3) {
      if (reac) LAUNCH_ROWS(3, true);
      else LAUNCH_ROWS(3, false);
    } else {
      if (reac) LAUNCH_ROWS(2, true);
      else LAUNCH_ROWS(2, false);
    }
#undef LAUNCH_ROWS
    MSFEC_HIP(hipGetLastError());
  }
  if (!fast) {
    hipLaunchKernelGGL(q_assemble_kernel, dim3(h->n_cells), dim3(256), 0, s, c, h->d_verts,
                       h->d_coef, cstride, h->rhs_count ? h->d_f : nullptr, fstride,
                       h->reac_count ? h->d_reac : nullptr, rstride, h->d_slo, h->d_F, h->d_rhs);
    MSFEC_HIP(hipGetLastError());
  }
  if (tm) MSFEC_HIP(hipEventRecord(ev[1], s));
  const bool mfma = (h->p.solver == MSFEC_SOLVER_BAND_FP64_MFMA);
  const bool mixed = (h->p.solver == MSFEC_SOLVER_BAND_MIXED);
  const size_t dyn = fast ? (size_t)c.nfc *
                                (1 + (c.reaction_mode == MSFEC_REACTION_CELL ? 1 : 0) +
                                 (c.rhs_mode == MSFEC_RHS_CELL ? 1 : 0)) *
                                sizeof(double)
                          : 0;
#define LAUNCH_BAND(K, D, F) hipLaunchKernelGGL((K<D, F>), dim3(h->n_cells), dim3(64), dyn, s, a)
  if (h->p.solver == MSFEC_SOLVER_CG && !fast) {
    // tensor coefficients / Gauss-point loads: CG on the assembled stencil
    CgArgs g{};
    g.tol = h->p.cg_tol > 0.0 ? h->p.cg_tol : 1e-13;
    g.max_iter = h->p.cg_max_iter > 0 ? h->p.cg_max_iter : 5000;
    g.iters = h->d_cgit;
    g.seg = cg_segment(c.ni, c.dim);
    g.fnode = h->d_F;
    const bool reac = c.reaction_mode != MSFEC_REACTION_NONE;
    g.nsolve = reac ? c.nv : c.nv - 1;
    const size_t cdyn = (size_t)c.nn * sizeof(double);
#define LAUNCH_CGS(D)                                                                     \
  do {                                                                                    \
    MSFEC_HIP(hipFuncSetAttribute((const void*)q_cg_stencil_kernel<D>,                    \
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)cdyn)); \
    hipLaunchKernelGGL((q_cg_stencil_kernel<D>), dim3(h->n_cells * g.nsolve), dim3(kCgThreads), \
                       cdyn, s, a, g, c);                                                 \
  } while (0)
    if (c.dim == 3) LAUNCH_CGS(3);
    else LAUNCH_CGS(2);
#undef LAUNCH_CGS
    MSFEC_HIP(hipGetLastError());
    if (!reac) {
      const unsigned ub = (unsigned)(((size_t)h->n_cells * c.n + 255) / 256);
      if (c.dim == 3)
        hipLaunchKernelGGL((q_cg_unity_kernel<3>), dim3(ub), dim3(256), 0, s, a, g, h->n_cells);
      else
        hipLaunchKernelGGL((q_cg_unity_kernel<2>), dim3(ub), dim3(256), 0, s, a, g, h->n_cells);
      MSFEC_HIP(hipGetLastError());
    }
    if (tm) MSFEC_HIP(hipEventRecord(ev[2], s));
    if (c.dim == 3)
      hipLaunchKernelGGL((q_cg_tensor_finish_kernel<3>), dim3(h->n_cells), dim3(kTfThreads), 0, s, a, g, c);
    else
      hipLaunchKernelGGL((q_cg_tensor_finish_kernel<2>), dim3(h->n_cells), dim3(kTfThreads), 0, s, a, g, c);
    MSFEC_HIP(hipGetLastError());
    if (tm) MSFEC_HIP(hipEventRecord(ev[3], s));
  } else if (h->p.solver == MSFEC_SOLVER_CG) {
    CgArgs g{};
    g.tol = h->p.cg_tol > 0.0 ? h->p.cg_tol : 1e-13;
    g.max_iter = h->p.cg_max_iter > 0 ? h->p.cg_max_iter : 5000;
    g.iters = h->d_cgit;
    g.seg = cg_segment(c.ni, c.dim);
    const size_t cdyn = ((size_t)c.nfc * (1 + (c.reaction_mode == MSFEC_REACTION_CELL ? 1 : 0)) +
                         (size_t)c.nn) * sizeof(double);
    const bool reac = c.reaction_mode != MSFEC_REACTION_NONE;
    g.nsolve = reac ? c.nv : c.nv - 1;
#define LAUNCH_CG(D, R)                                                                   \
  do {                                                                                    \
    MSFEC_HIP(hipFuncSetAttribute((const void*)q_cg_kernel<D, R>,                         \
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)cdyn)); \
    hipLaunchKernelGGL((q_cg_kernel<D, R>), dim3(h->n_cells * g.nsolve), dim3(kCgThreads), cdyn, \
                       s, a, g);                                                          \
  } while (0)
    if (c.dim == 3) {
      if (reac) LAUNCH_CG(3, true);
      else LAUNCH_CG(3, false);
    } else {
      if (reac) LAUNCH_CG(2, true);
      else LAUNCH_CG(2, false);
    }
#undef LAUNCH_CG
    MSFEC_HIP(hipGetLastError());
    if (!reac) {
      const unsigned ub = (unsigned)(((size_t)h->n_cells * c.n + 255) / 256);
      if (c.dim == 3)
        hipLaunchKernelGGL((q_cg_unity_kernel<3>), dim3(ub), dim3(256), 0, s, a, g, h->n_cells);
      else
        hipLaunchKernelGGL((q_cg_unity_kernel<2>), dim3(ub), dim3(256), 0, s, a, g, h->n_cells);
      MSFEC_HIP(hipGetLastError());
    }
    if (tm) MSFEC_HIP(hipEventRecord(ev[2], s));
    if (c.dim == 3)
      hipLaunchKernelGGL((q_cg_finish_kernel<3>), dim3(h->n_cells), dim3(64 * kFinWaves), 0, s, a, g);
    else
      hipLaunchKernelGGL((q_cg_finish_kernel<2>), dim3(h->n_cells), dim3(64 * kFinWaves), 0, s, a, g);
    MSFEC_HIP(hipGetLastError());
    if (tm) MSFEC_HIP(hipEventRecord(ev[3], s));
  } else if (mixed) {
    // FP32 factor and first solution, FP64 residual, FP32 correction sweep
    // X of the cell in LDS, and its lower stencil too when both fit in 64 KB
    const int stage_lo = ((size_t)c.n * (kMaxNb + c.nlow) * sizeof(double) <= 65536) ? 1 : 0;
    const size_t rsm = (size_t)c.n * (kMaxNb + (stage_lo ? c.nlow : 0)) * sizeof(double);
    if (c.dim == 3)
      hipLaunchKernelGGL((q_blk_factor_kernel<float, 3, BK_X0>), dim3(h->n_cells), dim3(64), 0, s, a);
    else
      hipLaunchKernelGGL((q_blk_factor_kernel<float, 2, BK_X0>), dim3(h->n_cells), dim3(64), 0, s, a);
    if (tm) MSFEC_HIP(hipEventRecord(ev[2], s));
    // refinement passes; cells that have converged leave each kernel at once
    const int passes = (h->p.max_refine > 0) ? h->p.max_refine : 3;
    for (int pass = 1; pass <= passes; ++pass) {
      a.last_pass = (pass == passes) ? 1 : 0;
      if (c.dim == 3)
        hipLaunchKernelGGL((q_resid_kernel<3>), dim3(h->n_cells), dim3(256), rsm, s, a, stage_lo);
      else
        hipLaunchKernelGGL((q_resid_kernel<2>), dim3(h->n_cells), dim3(256), rsm, s, a, stage_lo);
      if (tm && pass == 1) MSFEC_HIP(hipEventRecord(ev[3], s));
      if (c.dim == 3)
        hipLaunchKernelGGL((q_blk_refine_kernel<3>), dim3(h->n_cells), dim3(64), 0, s, a);
      else
        hipLaunchKernelGGL((q_blk_refine_kernel<2>), dim3(h->n_cells), dim3(64), 0, s, a);
    }
  } else if (mfma && fast) {
    launch_factor64(c.dim, c.nt, h->n_cells, s, a);
  } else if (mfma) {
    // tensor coefficients: stencil tables of q_assemble_kernel
    if (c.dim == 3) LAUNCH_BAND(q_band_mfma_kernel, 3, false);
    else LAUNCH_BAND(q_band_mfma_kernel, 2, false);
  } else {
    if (c.dim == 3) {
      if (fast) LAUNCH_BAND(q_band_kernel, 3, true);
      else LAUNCH_BAND(q_band_kernel, 3, false);
    } else {
      if (fast) LAUNCH_BAND(q_band_kernel, 2, true);
      else LAUNCH_BAND(q_band_kernel, 2, false);
    }
  }
#undef LAUNCH_BAND
  MSFEC_HIP(hipGetLastError());
  if (tm && !mixed && h->p.solver != MSFEC_SOLVER_CG) {
    MSFEC_HIP(hipEventRecord(ev[2], s));
    MSFEC_HIP(hipEventRecord(ev[3], s));
  }
  if (tm) MSFEC_HIP(hipEventRecord(ev[4], s));
  if (!fast && h->p.solver != MSFEC_SOLVER_CG) {
    const size_t shm = (size_t)(2 * c.nv * c.nn + 4 * 72) * sizeof(double);
    if (shm > 160 * 1024) return fail(MSFEC_ERR_UNSUPPORTED, "element kernel LDS too large");
    hipLaunchKernelGGL(q_element_kernel, dim3(h->n_cells), dim3(256), shm, s, c, h->d_slo,
                       h->d_F, h->d_X, h->d_psi, h->d_M, h->d_b);
    MSFEC_HIP(hipGetLastError());
  }
  if (tm) MSFEC_HIP(hipEventRecord(ev[5], s));
  h->computed = true;
  return MSFEC_OK;
}

msfec_status msfec_q_get_cg_iterations(msfec_q_handle h, int* iterations) {
  if (!h || !iterations) return fail(MSFEC_ERR_INVALID, "null argument");
  if (h->p.solver != MSFEC_SOLVER_CG) return fail(MSFEC_ERR_STATE, "not a CG handle");
  if (!h->computed) return fail(MSFEC_ERR_STATE, "compute not called");
  MSFEC_HIP(hipStreamSynchronize(h->stream));
  std::vector<int> raw((size_t)h->n_cells * h->conf.nv * 2);
  MSFEC_HIP(hipMemcpy(raw.data(), h->d_cgit, raw.size() * sizeof(int), hipMemcpyDeviceToHost));
  for (size_t t = 0; t < raw.size() / 2; ++t) iterations[t] = raw[2 * t];
  return MSFEC_OK;
}

msfec_status msfec_q_synchronize(msfec_q_handle h) {
  if (!h) return fail(MSFEC_ERR_INVALID, "null handle");
  MSFEC_HIP(hipStreamSynchronize(h->stream));
  return MSFEC_OK;
}

msfec_status msfec_q_kernel_times(msfec_q_handle h, float* ms, int max, int* n) {
  if (!h || !ms) return fail(MSFEC_ERR_INVALID, "null argument");
  if (h->ev_used == 0) return fail(MSFEC_ERR_STATE, "no timed compute (set MSFEC_FLAG_TIMING)");
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (size_t c = 0; c < h->ev_used; ++c) {
    MSFEC_HIP(hipEventSynchronize(h->ev_sets[c][5]));
    for (int i = 0; i < 5; ++i) {
      float t = 0.f;
      MSFEC_HIP(hipEventElapsedTime(&t, h->ev_sets[c][i], h->ev_sets[c][i + 1]));
      acc[i] += t;
    }
  }
  int k = 0;
  for (; k < 5 && k < max; ++k) ms[k] = (float)(acc[k] / (double)h->ev_used);
  if (n) *n = k;
  h->ev_used = 0;
  return MSFEC_OK;
}

msfec_status msfec_q_get_element(msfec_q_handle h, double* M, double* b, int loc) {
  if (!h) return fail(MSFEC_ERR_INVALID, "null handle");
  if (!h->computed) return fail(MSFEC_ERR_STATE, "compute not called");
  MSFEC_HIP(hipStreamSynchronize(h->stream));
  const size_t nv = h->conf.nv;
  if (M) {
    msfec_status st = copy_out(h->stream, M, h->d_M, (size_t)h->n_cells * nv * nv, loc);
    if (st) return st;
  }
  if (b) {
    msfec_status st = copy_out(h->stream, b, h->d_b, (size_t)h->n_cells * nv, loc);
    if (st) return st;
  }
  return MSFEC_OK;
}

msfec_status msfec_q_get_basis(msfec_q_handle h, double* Psi, int loc) {
  if (!h || !Psi) return fail(MSFEC_ERR_INVALID, "null argument");
  if (!h->computed) return fail(MSFEC_ERR_STATE, "compute not called");
  MSFEC_HIP(hipStreamSynchronize(h->stream));
  return copy_out(h->stream, Psi, h->d_psi, (size_t)h->n_cells * h->conf.nv * h->conf.nn, loc);
}

msfec_status msfec_q_get_cell_status(msfec_q_handle h, int* status) {
  if (!h || !status) return fail(MSFEC_ERR_INVALID, "null argument");
  if (!h->computed) return fail(MSFEC_ERR_STATE, "compute not called");
  MSFEC_HIP(hipStreamSynchronize(h->stream));
  MSFEC_HIP(hipMemcpy(status, h->d_status, h->n_cells * sizeof(int), hipMemcpyDeviceToHost));
  return MSFEC_OK;
}

msfec_status msfec_q_reconstruct(msfec_q_handle h, const double* weights, int weights_loc,
                                 double* u_fine, int out_loc) {
  if (!h || !weights || !u_fine) return fail(MSFEC_ERR_INVALID, "null argument");
  if (!h->computed) return fail(MSFEC_ERR_STATE, "compute not called");
  const size_t nw = (size_t)h->n_cells * h->conf.nv, nu = (size_t)h->n_cells * h->conf.nn;
  // host weights and outputs go through staging buffers that are freed on
  // every return path
  double* dw = nullptr;
  double* du = nullptr;
  const double* wsrc = weights;
  double* dst = u_fine;
  hipError_t e = hipSuccess;
  if (weights_loc != MSFEC_DEVICE) {
    e = hipMalloc((void**)&dw, nw * sizeof(double));
    if (e == hipSuccess)
      e = hipMemcpyAsync(dw, weights, nw * sizeof(double), hipMemcpyHostToDevice, h->stream);
    wsrc = dw;
  }
  if (e == hipSuccess && out_loc != MSFEC_DEVICE) {
    e = hipMalloc((void**)&du, nu * sizeof(double));
    dst = du;
  }
  if (e == hipSuccess) {
    const int threads = 256;
    const unsigned blocks = (unsigned)((nu + threads - 1) / threads);
    hipLaunchKernelGGL(q_reconstruct_kernel, dim3(blocks), dim3(threads), 0, h->stream,
                       h->conf.nv, h->conf.nn, h->n_cells, h->d_psi, wsrc, dst);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess && du) e = hipMemcpy(u_fine, du, nu * sizeof(double), hipMemcpyDeviceToHost);
  if (dw) (void)hipFree(dw);
  if (du) (void)hipFree(du);
  if (e != hipSuccess) return fail(MSFEC_ERR_HIP, std::string("reconstruct: ") + hipGetErrorString(e));
  return MSFEC_OK;
}

msfec_status msfec_coarse_solve(msfec_q_handle h, int nc, double abs_tol, double rel_tol,
                                int max_iter, double* u, int u_loc, double* weights,
                                int weights_loc, int* iterations, double* residual) {
  if (!h) return fail(MSFEC_ERR_INVALID, "null handle");
  if (!h->computed) return fail(MSFEC_ERR_STATE, "compute not called");
  const int dim = h->conf.dim, nv = h->conf.nv;
  if (nc <= 0 || ipow(nc, dim) != h->n_cells)
    return fail(MSFEC_ERR_INVALID, "nc^dim must equal the number of cells of the batch");
  const int nnode = ipow(nc + 1, dim), nst = ipow(3, dim);
  if (h->coarse_nc != nc) {
    if (h->d_cK) (void)hipFree(h->d_cK);
    if (h->d_crhs) (void)hipFree(h->d_crhs);
    if (h->d_cvec) (void)hipFree(h->d_cvec);
    if (h->d_cit) (void)hipFree(h->d_cit);
    h->d_cK = h->d_crhs = h->d_cvec = nullptr;
    h->d_cit = nullptr;
    MSFEC_HIP(hipMalloc((void**)&h->d_cK, (size_t)nnode * nst * sizeof(double)));
    MSFEC_HIP(hipMalloc((void**)&h->d_crhs, (size_t)nnode * sizeof(double)));
    MSFEC_HIP(hipMalloc((void**)&h->d_cvec, ((size_t)5 * nnode + 1 + (size_t)h->n_cells * nv) * sizeof(double)));
    MSFEC_HIP(hipMalloc((void**)&h->d_cit, sizeof(int)));
    h->coarse_nc = nc;
  }
  double* du = h->d_cvec;
  double* res = du + 5 * (size_t)nnode;
  double* dw = res + 1;
  MSFEC_HIP(msfec_detail::coarse_solve_device(dim, nc, h->d_M, h->d_b, h->d_cK, h->d_crhs,
                                              h->d_cvec, h->d_cit, abs_tol, rel_tol, max_iter,
                                              h->stream));
  const int nw = h->n_cells * nv;
  MSFEC_HIP(hipStreamSynchronize(h->stream));
  int it = 0;
  double rr = 0.0;
  MSFEC_HIP(hipMemcpy(&it, h->d_cit, sizeof(int), hipMemcpyDeviceToHost));
  MSFEC_HIP(hipMemcpy(&rr, res, sizeof(double), hipMemcpyDeviceToHost));
  if (iterations) *iterations = it;
  if (residual) *residual = rr;
  if (u) {
    msfec_status st = copy_out(h->stream, u, du, nnode, u_loc);
    if (st) return st;
  }
  if (weights) {
    msfec_status st = copy_out(h->stream, weights, dw, (size_t)nw, weights_loc);
    if (st) return st;
  }
  if (it < 0) return fail(MSFEC_ERR_CELL, "coarse CG did not converge");
  return MSFEC_OK;
}

msfec_status msfec_q_adjoint_gradient(msfec_q_handle h, int nc, const double* u_target,
                                      int target_loc, double rel_tol, int max_iter, double* J,
                                      double* grad, int grad_loc) {
  if (!h || !u_target || !J || !grad) return fail(MSFEC_ERR_INVALID, "null argument");
  if (!h->computed) return fail(MSFEC_ERR_STATE, "compute not called");
  if (!h->fast || h->conf.coef_mode != MSFEC_COEF_SCALAR_CELL)
    return fail(MSFEC_ERR_UNSUPPORTED, "the adjoint needs a scalar coefficient per fine cell");
  if (h->p.solver != MSFEC_SOLVER_BAND_FP64_MFMA && h->p.solver != MSFEC_SOLVER_BAND_MIXED)
    return fail(MSFEC_ERR_UNSUPPORTED,
                "the adjoint reuses the stored panels of BAND_FP64_MFMA or BAND_MIXED");
  const QConf& c = h->conf;
  const int dim = c.dim, nv = c.nv;
  if (nc <= 0 || ipow(nc, dim) != h->n_cells)
    return fail(MSFEC_ERR_INVALID, "nc^dim must equal the number of cells of the batch");
  hipStream_t s = h->stream;
  MSFEC_HIP(hipStreamSynchronize(s));
  {
    std::vector<int> st(h->n_cells);
    MSFEC_HIP(hipMemcpy(st.data(), h->d_status, st.size() * sizeof(int), hipMemcpyDeviceToHost));
    for (int v : st)
      if (v != MSFEC_CELL_OK) return fail(MSFEC_ERR_CELL, "a local solve failed; no gradient");
  }
  const int nnode = ipow(nc + 1, dim), nst = ipow(3, dim);
  if (h->coarse_nc != nc) {
    if (h->d_cK) (void)hipFree(h->d_cK);
    if (h->d_crhs) (void)hipFree(h->d_crhs);
    if (h->d_cvec) (void)hipFree(h->d_cvec);
    if (h->d_cit) (void)hipFree(h->d_cit);
    h->d_cK = h->d_crhs = h->d_cvec = nullptr;
    h->d_cit = nullptr;
    MSFEC_HIP(hipMalloc((void**)&h->d_cK, (size_t)nnode * nst * sizeof(double)));
    MSFEC_HIP(hipMalloc((void**)&h->d_crhs, (size_t)nnode * sizeof(double)));
    MSFEC_HIP(hipMalloc((void**)&h->d_cvec, ((size_t)5 * nnode + 1 + (size_t)h->n_cells * nv) * sizeof(double)));
    MSFEC_HIP(hipMalloc((void**)&h->d_cit, sizeof(int)));
    h->coarse_nc = nc;
  }
  double* du = h->d_cvec;
  double* r = du + nnode;
  double* z = r + nnode;
  double* pv = z + nnode;
  double* qv = pv + nnode;
  double* res = qv + nnode;
  // buffers of the adjoint: v, right-hand sides of w, the gradient (device)
  double *dv = nullptr, *dgrad = nullptr;
  MSFEC_HIP(hipMalloc((void**)&dv, (size_t)nnode * sizeof(double)));
  if (!h->d_rres) {
    if (hipMalloc((void**)&h->d_rres, (size_t)h->n_cells * c.n * kMaxNb * sizeof(double)) !=
        hipSuccess) {
      (void)hipFree(dv);
      return fail(MSFEC_ERR_HIP, "allocation of the adjoint right-hand sides failed");
    }
  }
  const size_t ngrad = (size_t)h->n_cells * c.nfc;
  if (grad_loc == MSFEC_DEVICE) {
    dgrad = grad;
  } else if (hipMalloc((void**)&dgrad, ngrad * sizeof(double)) != hipSuccess) {
    (void)hipFree(dv);
    return fail(MSFEC_ERR_HIP, "allocation of the gradient failed");
  }
  auto cleanup = [&]() {
    (void)hipFree(dv);
    if (grad_loc != MSFEC_DEVICE) (void)hipFree(dgrad);
  };
#define ADJ_HIP(call)                                 \
  do {                                                \
    hipError_t e_ = (call);                           \
    if (e_ != hipSuccess) {                           \
      cleanup();                                      \
      return fail(MSFEC_ERR_HIP, hipGetErrorString(e_)); \
    }                                                 \
  } while (0)
  // forward coarse problem
  hipLaunchKernelGGL(coarse_assemble_kernel, dim3((nnode * (dim == 3 ? 27 : 9) + 255) / 256),
                     dim3(256), 0, s, dim, nc,
                     h->d_M, h->d_b, h->d_cK, h->d_crhs);
  ADJ_HIP(hipGetLastError());
  ADJ_HIP(launch_coarse_cg(dim, nc, h->d_cK, h->d_crhs, du, r, z, pv, qv, 0.0, rel_tol, max_iter,
                           h->d_cit, res, s));
  int it = 0;
  ADJ_HIP(hipMemcpyAsync(&it, h->d_cit, sizeof(int), hipMemcpyDeviceToHost, s));
  std::vector<double> hu(nnode), ht(nnode), hd(nnode);
  ADJ_HIP(hipMemcpyAsync(hu.data(), du, nnode * sizeof(double), hipMemcpyDeviceToHost, s));
  ADJ_HIP(hipStreamSynchronize(s));
  if (it < 0) {
    cleanup();
    return fail(MSFEC_ERR_CELL, "coarse CG did not converge");
  }
  if (target_loc == MSFEC_DEVICE) {
    ADJ_HIP(hipMemcpy(ht.data(), u_target, nnode * sizeof(double), hipMemcpyDeviceToHost));
  } else {
    std::memcpy(ht.data(), u_target, nnode * sizeof(double));
  }
  // J and the adjoint right-hand side u_H - u_target on the free nodes
  double jsum = 0.0;
  for (int n = 0; n < nnode; ++n) {
    bool bnd = false;
    int t = n;
    for (int d = 0; d < dim; ++d) {
      const int cc = t % (nc + 1);
      t /= nc + 1;
      if (cc == 0 || cc == nc) bnd = true;
    }
    hd[n] = bnd ? 0.0 : hu[n] - ht[n];
    jsum += 0.5 * hd[n] * hd[n];
  }
  *J = jsum;
  ADJ_HIP(hipMemcpyAsync(h->d_crhs, hd.data(), nnode * sizeof(double), hipMemcpyHostToDevice, s));
  ADJ_HIP(launch_coarse_cg(dim, nc, h->d_cK, h->d_crhs, dv, r, z, pv, qv, 0.0, rel_tol, max_iter,
                           h->d_cit, res, s));
  // w = K_II^{-1} F_I with the stored panels
  BandArgs a = band_args(h);
  a.Rres = h->d_rres;
  {
    const long long nth = (long long)h->n_cells * c.n * kMaxNb;
    hipLaunchKernelGGL(q_adj_rhs_kernel, dim3((unsigned)((nth + 255) / 256)), dim3(256), 0, s, a,
                       h->n_cells);
    ADJ_HIP(hipGetLastError());
  }
  const bool f64 = (h->p.solver == MSFEC_SOLVER_BAND_FP64_MFMA);
  // the panels were stored with the handle's window (FP32 panels: always 64 slots)
#define SOLVE(R, D, NT) \
  hipLaunchKernelGGL((q_blk_solve_kernel<R, D, NT>), dim3(h->n_cells), dim3(64), 0, s, a)
  const int nt = f64 ? c.nt : 4;
  if (dim == 3) {
    if (!f64) SOLVE(float, 3, 4);
    else if (nt == 1) SOLVE(double, 3, 1);
    else if (nt == 2) SOLVE(double, 3, 2);
    else SOLVE(double, 3, 4);
  } else {
    if (!f64) SOLVE(float, 2, 4);
    else if (nt == 1) SOLVE(double, 2, 1);
    else if (nt == 2) SOLVE(double, 2, 2);
    else SOLVE(double, 2, 4);
#undef SOLVE
  }
  ADJ_HIP(hipGetLastError());
  const size_t gsm = (size_t)2 * c.nn * sizeof(double);
  if (dim == 3)
    hipLaunchKernelGGL((q_adj_grad_kernel<3>), dim3(h->n_cells), dim3(256), gsm, s, a, nc, du, dv, dgrad);
  else
    hipLaunchKernelGGL((q_adj_grad_kernel<2>), dim3(h->n_cells), dim3(256), gsm, s, a, nc, du, dv, dgrad);
  ADJ_HIP(hipGetLastError());
  ADJ_HIP(hipMemcpyAsync(&it, h->d_cit, sizeof(int), hipMemcpyDeviceToHost, s));
  if (grad_loc != MSFEC_DEVICE)
    ADJ_HIP(hipMemcpyAsync(grad, dgrad, ngrad * sizeof(double), hipMemcpyDeviceToHost, s));
  ADJ_HIP(hipStreamSynchronize(s));
#undef ADJ_HIP
  cleanup();
  if (it < 0) return fail(MSFEC_ERR_CELL, "adjoint coarse CG did not converge");
  return MSFEC_OK;
}

msfec_status msfec_q_compute_noalloc(const msfec_q_params* params, int n_cells,
                                     const double* vertices, const double* coef, double* Psi,
                                     double* M, double* b, int* status, void* stream) {
  if (!params || !vertices || !coef || !Psi || !M || !b || !status)
    return fail(MSFEC_ERR_INVALID, "null argument");
  const msfec_q_params& p = *params;
  if (p.dim != 2 && p.dim != 3) return fail(MSFEC_ERR_INVALID, "dim must be 2 or 3");
  if (n_cells <= 0) return fail(MSFEC_ERR_INVALID, "n_cells must be positive");
  if (p.n_refine_local < 1) return fail(MSFEC_ERR_INVALID, "n_refine_local must be >= 1");
  if (p.solver != MSFEC_SOLVER_BAND_FP64_MFMA || p.coef_mode != MSFEC_COEF_SCALAR_CELL ||
      p.rhs_mode != MSFEC_RHS_CONST ||
      (p.reaction_mode != MSFEC_REACTION_NONE && p.reaction_mode != MSFEC_REACTION_CONST))
    return fail(MSFEC_ERR_UNSUPPORTED,
                "msfec_q_compute_noalloc takes MSFEC_SOLVER_BAND_FP64_MFMA with a scalar "
                "coefficient per fine cell, a constant load and no or a constant reaction");
  const int ni = (1 << p.n_refine_local) - 1;
  const int bw = (p.dim == 2) ? ni + 1 : ni * ni + ni + 1;
  const int nt = band_tile_rows(bw, 0);
  if (nt == 0)
    return fail(MSFEC_ERR_UNSUPPORTED,
                "interior half-bandwidth " + std::to_string(bw) +
                    " exceeds the 60 of the band solver (3-D: n_refine_local <= 3, 2-D: <= 5)");
  QConf c{};
  fill_qconf(p, c);
  // scratch per cell: row table | P b_B | panels | forward results
  const size_t nq = (size_t)panel_quads(c.n);
  const size_t rows_d = (size_t)c.n * kRowStride, pb_d = kPStride, pan_d = nq * 256, z_d = nq * 32;
  const size_t per_cell = rows_d + pb_d + pan_d + z_d;
  const size_t cap_d = kNoAllocScratchBytes / sizeof(double);
  if (4 * 32 + per_cell > cap_d)
    return fail(MSFEC_ERR_UNSUPPORTED, "one cell needs more scratch than the 128 MiB array holds");
  const int chunk = (int)std::min<size_t>((size_t)n_cells, (cap_d - 4 * 32) / per_cell);
  double* scr = nullptr;
  msfec_status st = scratch_on_device(&scr);
  if (st != MSFEC_OK) return st;
  double* d_rows = scr;
  double* d_pb = d_rows + round_up32((size_t)chunk * rows_d);
  double* d_pan = d_pb + round_up32((size_t)chunk * pb_d);
  double* d_z = d_pan + round_up32((size_t)chunk * pan_d);
  hipStream_t s = (hipStream_t)stream;
  const int nv = c.nv;
  const size_t rdyn = (size_t)c.nfc * sizeof(double);
  const bool reac = p.reaction_mode != MSFEC_REACTION_NONE;
  // chunks of cells run one after another on the stream, each through the
  // same scratch: the kernels of a chunk are stream-ordered after those of
  // the chunk before
  for (int c0 = 0; c0 < n_cells; c0 += chunk) {
    const int nc = std::min(chunk, n_cells - c0);
    BandArgs a{};
    a.n = c.n;
    a.ni = c.ni;
    a.N = c.N;
    a.S = c.S;
    a.nn = c.nn;
    a.nfc = c.nfc;
    for (int t = 0; t < kMaxLow; ++t) a.off_int[t] = c.off_int[t];
    a.Rows = d_rows;
    a.verts = vertices + (size_t)c0 * nv * c.dim;
    a.coef = coef + (size_t)c0 * c.nfc;
    a.reac = nullptr;
    a.fcell = nullptr;
    a.reaction_mode = c.reaction_mode;
    a.rhs_mode = c.rhs_mode;
    a.reaction_const = c.reaction_const;
    a.rhs_const = c.rhs_const;
    a.pan = d_pan;
    a.zst = d_z;
    a.Pb = d_pb;
    a.Psi = Psi + (size_t)c0 * nv * c.nn;
    a.M = M + (size_t)c0 * nv * nv;
    a.bvec = b + (size_t)c0 * nv;
    a.status = status + c0;
    a.nq = (int)nq;
    if (c.dim == 3) {
      if (reac)
        hipLaunchKernelGGL((q_rows_kernel<3, true>), dim3(nc), dim3(256), rdyn, s, a);
      else
        hipLaunchKernelGGL((q_rows_kernel<3, false>), dim3(nc), dim3(256), rdyn, s, a);
      launch_factor64(3, nt, nc, s, a);
    } else {
      if (reac)
        hipLaunchKernelGGL((q_rows_kernel<2, true>), dim3(nc), dim3(256), rdyn, s, a);
      else
        hipLaunchKernelGGL((q_rows_kernel<2, false>), dim3(nc), dim3(256), rdyn, s, a);
      launch_factor64(2, nt, nc, s, a);
    }
    MSFEC_HIP(hipGetLastError());
  }
  return MSFEC_OK;
}

void msfec_q_destroy(msfec_q_handle h) {
  if (!h) return;
  free_all(h);
  delete h;
}

}  // extern "C"
