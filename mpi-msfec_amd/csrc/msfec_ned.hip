// Q_Ned: the mixed Q1 x lowest-order Nedelec multiscale basis in two
// dimensions on MI355X (gfx950), behind the msfec_ned_* entry points of
// include/msfec.h.
//
// Reference: source/Q_Ned/q_ned_basis.cc (three-dimensional QNedBasis).  Per
// coarse box cell, on its 2^l x 2^l fine grid:
//   (tau, B^-1 sigma) - (grad tau, u) = 0
//   (v, grad sigma) + (curl v, A curl u) + (v, R u) = (v, grad phi_n)  [vertex bases, full_rhs]
// (q_ned_basis.cc:483-491, 498-505), lifted from sigma = phi_n, u x n =
// grad phi_n (vertex bases, :223-294) or sigma = 0, u x n = N_E (edge bases,
// :296-357), solved directly (:570-625), then the element matrix of
// :841-939.
//
// Layout of the work:
//   ned_rows_kernel    one thread per interior row: the row of the interior
//                      system with the sigma rows negated (which makes it
//                      symmetric), in band order, its lower entries (at most
//                      16) and the 8 condensed right-hand sides, equilibrated
//                      (S K S, S = |diag K|^-1/2); the boundary values of the
//                      8 bases to Psi.
//   band_ldl_solve     (msfec_q.hip) one wave per cell: the matrix-core band
//                      factorisation of the Q path in its L D L^T form, 4
//                      pivots per step, with the forward elimination of the 8
//                      right-hand sides and the backward sweep.
//   ned_element_kernel one workgroup per cell: the interior of Psi (the row
//                      scale times the solution), then the 8 x 8 element
//                      matrix and the load vector from Psi staged in LDS.
// Band order: node row j (nodes (i,j)), interleaved with the horizontal
// edges of row j and the vertical edges between rows j-1 and j, every edge
// after its interior end nodes.  With that order every leading principal
// minor of the symmetric matrix [[-M, G^T], [G, C]] is non-singular (an
// edge field that is a discrete gradient of nodes already eliminated and
// orthogonal to their tests vanishes), so LDL^T needs no pivoting, and its
// 4 x 4 block form (the leading minors of every block are leading minors of
// the matrix) needs none either; the half-bandwidth is 3m + 3 (51 at l = 4).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <set>
#include <string>
#include <vector>

#include "msfec.h"

namespace msfec_detail {
extern thread_local std::string g_last_error;
// the matrix-core band LDL^T (msfec_q.hip)
size_t band_ldl_panel_doubles(int n);
size_t band_ldl_z_doubles(int n);
hipError_t band_ldl_solve(int n_cells, int n, int bw, const double* rows, const int* loff,
                          double* pan, double* zst, double* X, int* status, hipStream_t s);
// the same with an FP32 factor and FP64 refinement (msfec_q.hip)
size_t band_ldl_mixed_doubles(int n);
std::vector<int> band_upper_table(const std::vector<int>& loff, int n);
hipError_t band_ldl_mixed_solve(int n_cells, int n, int bw, const double* rows, const int* loff,
                                const int* utab, double* pan, double* zst, double* X,
                                double* scratch, int* status, int max_refine, hipStream_t s);
}

namespace {

constexpr int kNLow = 16;                // lower entries of a band row, diagonal first
constexpr int kMaxBand = 60;             // half-bandwidth of the matrix-core band solver
constexpr int kNR = 8;                   // bases per cell
constexpr int kRowLen = kNLow + kNR;     // row table: lower values, right-hand sides
constexpr int kNThreads = 256;
constexpr int kMaxRows = 1024;           // interior dofs of a cell (705 at l = 4)

struct NedArgs {
  int m, ns, nh, N;  // fine mesh: m cells per direction; nodes, horizontal edges, all dofs
  int n, p, full_rhs;
  const int* order;  // [n] band row -> fine dof
  const int* pos;    // [N] fine dof -> band row, -1 on the boundary
  const int* loff;   // [n][kNLow] offsets row - column of the lower entries, -1 unused
  const double* gbnd;        // [N][8] boundary values of the 8 bases (0 inside)
  const signed char* cslot;  // [n][4][8] per fine cell of the row and local dof: the
                             // lower slot of the coupling, -1 boundary dof, -2 upper
  const double* verts;
  const double* a;
  const double* binv;
  const double* r;
  const double* f;
  double* rows;  // [cells][n][kRowLen], equilibrated: S K S with S = |diag K|^-1/2
  double* scale; // [cells][n] the row scales s_r
  double* X;     // [cells][n][8] solution of the equilibrated system
  double* psi;   // [cells][8][N]
  double* M;     // [cells][64]
  double* b;     // [cells][8]
  int* status;
};

struct NedBox {
  double hx, hy;
};

__device__ __forceinline__ NedBox ned_box(const NedArgs& a, int cell) {
  const double* v = a.verts + (size_t)cell * 8;
  return NedBox{(v[2] - v[0]) / a.m, (v[5] - v[1]) / a.m};
}

// fine dof -> (type, i, j): 0 node (i,j); 1 horizontal edge (i,j)-(i+1,j);
// 2 vertical edge (i,j)-(i,j+1)
__host__ __device__ inline void ned_decode_dof(int ns, int nh, int m, int d, int& type, int& i,
                                               int& j) {
  if (d < ns) {
    type = 0;
    j = d / (m + 1);
    i = d - j * (m + 1);
  } else if (d < ns + nh) {
    type = 1;
    d -= ns;
    j = d / m;
    i = d - j * m;
  } else {
    type = 2;
    d -= ns + nh;
    j = d / (m + 1);
    i = d - j * (m + 1);
  }
}
__device__ __forceinline__ void ned_decode(const NedArgs& a, int d, int& type, int& i, int& j) {
  ned_decode_dof(a.ns, a.nh, a.m, d, type, i, j);
}

// dofs of fine cell (i, j): nodes in deal.II vertex order, then the lines in
// deal.II line order (x = left, x = right, y = bottom, y = top)
__device__ __forceinline__ void ned_cell_dofs(const NedArgs& a, int i, int j, int (&d)[8]) {
  const int m = a.m;
  d[0] = j * (m + 1) + i;
  d[1] = d[0] + 1;
  d[2] = d[0] + m + 1;
  d[3] = d[2] + 1;
  d[4] = a.ns + a.nh + j * (m + 1) + i;
  d[5] = d[4] + 1;
  d[6] = a.ns + j * m + i;
  d[7] = d[6] + m;
}

// Exact integrals of the mixed form over a fine box (hx, hy) with constant
// coefficients (the QGauss(3) rule of the reference integrates them
// exactly).  K[test][trial], local order as ned_cell_dofs.
//   Q1 mass           binv hx hy / 36 [4 2 2 1; 2 4 1 2; 2 1 4 2; 1 2 2 4]
//   G[e][k]           (N_e, grad phi_k)
//   curl N_e          s_e / (hx hy), s = (-1, 1, 1, -1)
//   Nedelec mass      hx / (3 hy), hx / (6 hy) for the vertical pair,
//                     hy / (3 hx), hy / (6 hx) for the horizontal pair
// The geometry factors of a fine box, formed once per coarse cell.
struct NedGeo {
  double m36, ihh, gx3, gx6, gy3, gy6;  // hx hy / 36, 1 / (hx hy), gx / 3, gx / 6, gy / 3, gy / 6
};
__device__ __forceinline__ NedGeo ned_geo(double hx, double hy) {
  const double gx = hx / hy, gy = hy / hx;
  return {hx * hy / 36.0, 1.0 / (hx * hy), gx / 3, gx / 6, gy / 3, gy / 6};
}

__device__ __forceinline__ void ned_local(const NedGeo& g, double ac, double bc, double rc,
                                          double (&K)[8][8]) {
  const double mq = bc * g.m36;
  const double Mq[4][4] = {{4, 2, 2, 1}, {2, 4, 1, 2}, {2, 1, 4, 2}, {1, 2, 2, 4}};
  const double G[4][4] = {{-g.gx3, -g.gx6, g.gx3, g.gx6},
                          {-g.gx6, -g.gx3, g.gx6, g.gx3},
                          {-g.gy3, g.gy3, -g.gy6, g.gy6},
                          {-g.gy6, g.gy6, -g.gy3, g.gy3}};
  const double s[4] = {-1, 1, 1, -1};
  const double cc = ac * g.ihh;
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int k = 0; k < 4; ++k) K[p][k] = mq * Mq[p][k];
#pragma unroll
  for (int e = 0; e < 4; ++e)
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      K[k][4 + e] = -G[e][k];
      K[4 + e][k] = G[e][k];
    }
#pragma unroll
  for (int e = 0; e < 4; ++e)
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      double mass = 0.0;
      if ((e >> 1) == (h >> 1))
        mass = (e >> 1) == 0 ? (e == h ? g.gx3 : g.gx6) : (e == h ? g.gy3 : g.gy6);
      K[4 + e][4 + h] = cc * s[e] * s[h] + rc * mass;
    }
}

// (N_e, grad phi_n) over fine cell (i, j) for the coarse bilinear phi_n:
// d_x phi_n depends on y only, d_y phi_n on x only, both linearly.
__device__ __forceinline__ double ned_source(double im, double cx, double cy, int i, int j,
                                             int n, int e) {
  // cx = hx / (6 Hy), cy = hy / (6 Hx); im = 1 / m (exact: m is a power of 2)
  auto lin = [](int hi, double t) { return hi ? t : 1.0 - t; };
  if (e < 2) {  // vertical line: (0, lin/hy) . grad phi_n -> d_y phi_n(x)
    const double c = (n >> 1) ? cx : -cx;
    const double h0 = lin(n & 1, i * im), h1 = lin(n & 1, (i + 1) * im);
    return e == 0 ? c * (2.0 * h0 + h1) : c * (h0 + 2.0 * h1);
  }
  const double c = (n & 1) ? cy : -cy;
  const double g0 = lin(n >> 1, j * im), g1 = lin(n >> 1, (j + 1) * im);
  return e == 2 ? c * (2.0 * g0 + g1) : c * (g0 + 2.0 * g1);
}

// Boundary value of basis b at boundary dof d (setup_basis_dofs_h1/_curl):
// vertex basis: phi_n at nodes, phi_n(end) - phi_n(start) on edges (the
// line integral of grad phi_n); edge basis E: 1/m on the fine edges of
// coarse edge E (the line integral of N_E), 0 elsewhere.  The same for every
// coarse cell of a refinement level: tabulated once per level on the host.
inline double ned_bval(int ns, int nh, int mi, int d, int b) {
  int type, i, j;
  ned_decode_dof(ns, nh, mi, d, type, i, j);
  const double m = mi;
  auto phi = [](int n, double X, double Y) {
    return ((n & 1) ? X : 1.0 - X) * ((n >> 1) ? Y : 1.0 - Y);
  };
  if (type == 0) return b < 4 ? phi(b, i / m, j / m) : 0.0;
  if (type == 1) {
    if (b < 4) return phi(b, (i + 1) / m, j / m) - phi(b, i / m, j / m);
    return ((b == 6 && j == 0) || (b == 7 && j == mi)) ? 1.0 / m : 0.0;
  }
  if (b < 4) return phi(b, i / m, (j + 1) / m) - phi(b, i / m, j / m);
  return ((b == 4 && i == 0) || (b == 5 && i == mi)) ? 1.0 / m : 0.0;
}

// The fine cells around fine dof d of type/(i, j) and the dof's local index
// in each (ned_cell_dofs order); returns their number.
__device__ __forceinline__ int ned_row_cells(int type, int i, int j, int (&ci)[4], int (&cj)[4],
                                             int (&lp)[4]) {
  if (type == 0) {
    ci[0] = i - 1; cj[0] = j - 1; lp[0] = 3;
    ci[1] = i;     cj[1] = j - 1; lp[1] = 2;
    ci[2] = i - 1; cj[2] = j;     lp[2] = 1;
    ci[3] = i;     cj[3] = j;     lp[3] = 0;
    return 4;
  }
  if (type == 1) {
    ci[0] = i; cj[0] = j - 1; lp[0] = 7;
    ci[1] = i; cj[1] = j;     lp[1] = 6;
    ci[2] = ci[3] = i; cj[2] = cj[3] = j; lp[2] = lp[3] = 6;
    return 2;
  }
  ci[0] = i - 1; cj[0] = j; lp[0] = 5;
  ci[1] = i;     cj[1] = j; lp[1] = 4;
  ci[2] = ci[3] = i; cj[2] = cj[3] = j; lp[2] = lp[3] = 4;
  return 2;
}

// Entry (p, k) of the local matrix, selected without dynamic register indexing.
__device__ __forceinline__ double ned_pick(const double (&K)[8][8], int p, int k) {
  double v = 0.0;
#pragma unroll
  for (int pp = 0; pp < 8; ++pp) v = (pp == p) ? K[pp][k] : v;
  return v;
}

__global__ __launch_bounds__(kNThreads) void ned_rows_kernel(NedArgs a) {
  const int cell = blockIdx.x;
  __shared__ double ss[kMaxRows];  // row scales
  __shared__ double srow[kNThreads][kNLow + 1];  // the lower entries of the thread's row
  double* sl = srow[threadIdx.x];
  const NedBox bx = ned_box(a, cell);
  const int m = a.m, nf = m * m;
  const double im = 1.0 / m, cx = bx.hx / (6.0 * bx.hy * m), cy = bx.hy / (6.0 * bx.hx * m);
  const NedGeo geo = ned_geo(bx.hx, bx.hy);
  const double* ca = a.a + (size_t)cell * nf;
  const double* cb = a.binv + (size_t)cell * nf;
  const double* cr = a.r + (size_t)cell * nf;
  // the row scales s_r = |K_rr|^-1/2 first (the diagonal summed in the order
  // the rows below accumulate it), so that every row is written once,
  // equilibrated: entry (r, r - o) times s_r s_(r-o), right-hand sides times s_r
  for (int r = threadIdx.x; r < a.n; r += blockDim.x) {
    int type, i, j, ci[4], cj[4], lp[4];
    ned_decode(a, a.order[r], type, i, j);
    const int nc = ned_row_cells(type, i, j, ci, cj, lp);
    double dg = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (q >= nc) break;
      const int fc = cj[q] * m + ci[q];
      double K[8][8];
      ned_local(geo, ca[fc], cb[fc], cr[fc], K);
      const int p = lp[q];
      const double sg = p < 4 ? -1.0 : 1.0;  // sigma rows negated
      double v = 0.0;
#pragma unroll
      for (int pp = 0; pp < 8; ++pp) v = (pp == p) ? K[pp][pp] : v;
      dg += sg * v;
    }
    dg = fabs(dg);
    ss[r] = dg > 0.0 ? 1.0 / sqrt(dg) : 1.0;
  }
  __syncthreads();
  for (int r = threadIdx.x; r < a.n; r += blockDim.x) {
    int type, i, j, ci[4], cj[4], lp[4];
    ned_decode(a, a.order[r], type, i, j);
    const int nc = ned_row_cells(type, i, j, ci, cj, lp);
    // slots of the couplings: byte 8q + k of the row's 32
    const int4* cw = reinterpret_cast<const int4*>(a.cslot + (size_t)r * 32);
    const int4 w0 = cw[0], w1 = cw[1];
    const int ws[8] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w};
    double rhs[kNR];
#pragma unroll
    for (int t = 0; t < kNLow; ++t) sl[t] = 0.0;
#pragma unroll
    for (int t = 0; t < kNR; ++t) rhs[t] = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (q >= nc) break;
      const int fc = cj[q] * m + ci[q];
      double K[8][8];
      ned_local(geo, ca[fc], cb[fc], cr[fc], K);
      int dofs[8];
      ned_cell_dofs(a, ci[q], cj[q], dofs);
      const int p = lp[q];
      const double sg = p < 4 ? -1.0 : 1.0;  // sigma rows negated
      double kr[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) kr[k] = sg * ned_pick(K, p, k);
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int t = (signed char)(ws[(8 * q + k) >> 2] >> (8 * ((8 * q + k) & 3)));
        if (t >= 0) {
          sl[t] += kr[k];  // lower coupling in slot t
        } else if (t == -1) {
#pragma unroll
          for (int bb = 0; bb < kNR; ++bb) rhs[bb] -= kr[k] * a.gbnd[(size_t)dofs[k] * kNR + bb];
        }
      }
      if (a.full_rhs && p >= 4)
#pragma unroll
        for (int n = 0; n < 4; ++n) rhs[n] += ned_source(im, cx, cy, ci[q], cj[q], n, p - 4);
    }
    const double sr = ss[r];
    double* out = a.rows + ((size_t)cell * a.n + r) * kRowLen;
    const int* lo = a.loff + (size_t)r * kNLow;
#pragma unroll
    for (int t = 0; t < kNLow; ++t) {
      const int o = lo[t];
      out[t] = o >= 0 ? sl[t] * (sr * ss[r - o]) : sl[t];
    }
#pragma unroll
    for (int t = 0; t < kNR; ++t) out[kNLow + t] = rhs[t] * sr;
    a.scale[(size_t)cell * a.n + r] = sr;
  }
  for (int d = threadIdx.x; d < a.N; d += blockDim.x)
    if (a.pos[d] < 0)
#pragma unroll
      for (int bb = 0; bb < kNR; ++bb) a.psi[((size_t)cell * kNR + bb) * a.N + d] = a.gbnd[(size_t)d * kNR + bb];
}

// Interior of Psi and the element matrix (q_ned_basis.cc:841-939):
// M[kt][kc] = psi_kt^T A psi_kc on the sigma block for vertex bases and the u
// block for edge bases; b[kt] = (psi_kt u, f) for the edge bases.  The
// interior of Psi (the row scale times the solution of the equilibrated
// system) goes to HBM and, with the boundary values, to LDS.  Thread
// (kt = tid & 7, group tid >> 3) sums row kt of the fine-cell contributions
// over every 32nd fine cell; the groups add up across lanes and waves.
// LDS slot of value d of basis bb in the element kernel, or -1: the element
// matrix reads only the sigma (node) values of the vertex bases and the u
// (edge) values of the edge bases, so LDS holds [4][ns] node values, then
// [4][N - ns] edge values (half of Psi).
__device__ __forceinline__ int ned_lds_slot(int bb, int d, int ns, int N) {
  if (bb < 4) return d < ns ? bb * ns + d : -1;
  return d >= ns ? 4 * ns + (bb - 4) * (N - ns) + (d - ns) : -1;
}

__global__ __launch_bounds__(kNThreads) void ned_element_kernel(NedArgs a) {
  extern __shared__ double ps[];  // [4][ns] node values | [4][N - ns] edge values
  __shared__ double part[kNThreads / 64][72];
  const int cell = blockIdx.x, tid = threadIdx.x;
  const int N = a.N, m = a.m, nf = m * m, n = a.n, ns = a.ns, nu = N - ns;
  double* pg = a.psi + (size_t)cell * kNR * N;
  for (int t = tid; t < kNR * N; t += kNThreads) {
    const int bb = t / N, d = t - bb * N;
    const int sl = ned_lds_slot(bb, d, ns, N);
    if (sl >= 0 && a.pos[d] < 0) ps[sl] = pg[t];  // boundary values (row kernel)
  }
  const double* x = a.X + (size_t)cell * n * kNR;
  const double* sc = a.scale + (size_t)cell * n;
  for (int i = tid; i < n * kNR; i += kNThreads) {
    const int r = i >> 3, bb = i & (kNR - 1);
    const int d = a.order[r];
    const double v = sc[r] * x[i];
    const int sl = ned_lds_slot(bb, d, ns, N);
    if (sl >= 0) ps[sl] = v;
    pg[bb * N + d] = v;
  }
  __syncthreads();
  const NedBox bx = ned_box(a, cell);
  const NedGeo geo = ned_geo(bx.hx, bx.hy);
  const int kt = tid & 7, grp = tid >> 3;
  const bool edge_t = kt >= 4;
  const int r0 = edge_t ? 4 : 0;
  double acc[kNR], bacc = 0.0;
#pragma unroll
  for (int k = 0; k < kNR; ++k) acc[k] = 0.0;
  for (int fc = grp; fc < nf; fc += kNThreads / 8) {
    const int i = fc % m, j = fc / m;
    int d[8];
    ned_cell_dofs(a, i, j, d);
    const size_t c = (size_t)cell * nf + fc;
    double K[8][8];
    ned_local(geo, a.a[c], a.binv[c], a.r[c], K);
    double u[4];
#pragma unroll
    for (int p = 0; p < 4; ++p)
      u[p] = edge_t ? ps[4 * ns + (kt - 4) * nu + d[4 + p] - ns] : ps[kt * ns + d[p]];
#pragma unroll
    for (int c0 = 0; c0 < 8; c0 += 4) {
      double w[4];  // row kt of psi^T K on the (r0, c0) block
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        double t = 0.0;
#pragma unroll
        for (int p = 0; p < 4; ++p) t += u[p] * (edge_t ? K[4 + p][c0 + q] : K[p][c0 + q]);
        w[q] = t;
      }
#pragma unroll
      for (int kc = c0; kc < c0 + 4; ++kc) {
        double t = 0.0;
#pragma unroll
        for (int q = 0; q < 4; ++q)
          t += w[q] * (c0 == 0 ? ps[kc * ns + d[q]] : ps[4 * ns + (kc - 4) * nu + d[4 + q] - ns]);
        acc[kc] += t;
      }
    }
    if (edge_t) {
      const double fx = a.f[2 * c], fy = a.f[2 * c + 1];
      const double lf[4] = {fy * bx.hx / 2, fy * bx.hx / 2, fx * bx.hy / 2, fx * bx.hy / 2};
#pragma unroll
      for (int q = 0; q < 4; ++q) bacc += u[q] * lf[q];
    }
  }
  // the 8 groups of a wave (lane bits 3..5), then the waves
#pragma unroll
  for (int k = 0; k < kNR; ++k)
#pragma unroll
    for (int o = 8; o < 64; o <<= 1) acc[k] += __shfl_xor(acc[k], o);
#pragma unroll
  for (int o = 8; o < 64; o <<= 1) bacc += __shfl_xor(bacc, o);
  const int wave = tid >> 6, lane = tid & 63;
  if (lane < 8) {
#pragma unroll
    for (int k = 0; k < kNR; ++k) part[wave][lane * 8 + k] = acc[k];
    part[wave][64 + lane] = bacc;
  }
  __syncthreads();
  if (tid < 72) {
    double v = 0.0;
#pragma unroll
    for (int w = 0; w < kNThreads / 64; ++w) v += part[w][tid];
    if (tid < 64) a.M[(size_t)cell * 64 + tid] = v;
    else a.b[(size_t)cell * kNR + tid - 64] = (tid - 64 < 4) ? 0.0 : v;  // vertex bases: no load
  }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
using msfec_detail::g_last_error;

msfec_status ned_fail(msfec_status st, const std::string& msg) {
  g_last_error = msg;
  return st;
}

#define NED_HIP(call)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return ned_fail(MSFEC_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// band order, its inverse, lower offsets and the pair list of one refinement level
struct NedConf {
  int l = 0, m = 0, ns = 0, nh = 0, nv = 0, N = 0, n = 0, p = 0;
  std::vector<int> order, pos, loff;
  std::vector<signed char> cslot;
  std::vector<double> gbnd;
};

bool ned_conf(int l, NedConf& c, std::string& why) {
  c.l = l;
  c.m = 1 << l;
  const int m = c.m;
  c.ns = (m + 1) * (m + 1);
  c.nh = m * (m + 1);
  c.nv = (m + 1) * m;
  c.N = c.ns + c.nh + c.nv;
  auto node = [&](int i, int j) { return j * (m + 1) + i; };
  auto hed = [&](int i, int j) { return c.ns + j * m + i; };
  auto ved = [&](int i, int j) { return c.ns + c.nh + j * (m + 1) + i; };
  c.order.clear();
  for (int j = 1; j < m; ++j)
    for (int i = 0; i <= m; ++i) {
      if (i >= 1 && i <= m - 1) c.order.push_back(node(i, j));
      if (i >= 1) c.order.push_back(hed(i - 1, j));
      if (i >= 1 && i <= m - 1) c.order.push_back(ved(i, j - 1));
    }
  for (int i = 1; i < m; ++i) c.order.push_back(ved(i, m - 1));
  c.n = (int)c.order.size();
  c.pos.assign(c.N, -1);
  for (int r = 0; r < c.n; ++r) c.pos[c.order[r]] = r;
  // dofs sharing a fine cell with each dof
  std::vector<std::set<int>> nb(c.N);
  for (int j = 0; j < m; ++j)
    for (int i = 0; i < m; ++i) {
      const int d[8] = {node(i, j),     node(i + 1, j),     node(i, j + 1), node(i + 1, j + 1),
                        ved(i, j),      ved(i + 1, j),      hed(i, j),      hed(i, j + 1)};
      for (int x : d)
        for (int y : d) nb[x].insert(y);
    }
  c.loff.assign((size_t)c.n * kNLow, -1);
  c.p = 0;
  for (int r = 0; r < c.n; ++r) {
    std::vector<int> offs;
    for (int g : nb[c.order[r]])
      if (c.pos[g] >= 0 && c.pos[g] <= r) offs.push_back(r - c.pos[g]);
    std::sort(offs.begin(), offs.end());
    if ((int)offs.size() > kNLow || offs.empty() || offs[0] != 0) {
      why = "band row structure exceeds the supported stencil";
      return false;
    }
    for (size_t t = 0; t < offs.size(); ++t) c.loff[(size_t)r * kNLow + t] = offs[t];
    c.p = std::max(c.p, offs.back());
  }
  // per band row, fine cell of the row (in the order of ned_rows_kernel) and
  // local dof (ned_cell_dofs): the lower slot of the coupling, -1 for a
  // boundary dof, -2 for an upper coupling
  c.cslot.assign((size_t)c.n * 32, (signed char)-2);
  for (int r = 0; r < c.n; ++r) {
    int d = c.order[r], type, i, j;
    if (d < c.ns) {
      type = 0, j = d / (m + 1), i = d % (m + 1);
    } else if (d < c.ns + c.nh) {
      d -= c.ns;
      type = 1, j = d / m, i = d % m;
    } else {
      d -= c.ns + c.nh;
      type = 2, j = d / (m + 1), i = d % (m + 1);
    }
    int ci[4], cj[4], nc;
    if (type == 0) {
      nc = 4;
      ci[0] = i - 1, cj[0] = j - 1, ci[1] = i, cj[1] = j - 1;
      ci[2] = i - 1, cj[2] = j, ci[3] = i, cj[3] = j;
    } else if (type == 1) {
      nc = 2;
      ci[0] = i, cj[0] = j - 1, ci[1] = i, cj[1] = j;
    } else {
      nc = 2;
      ci[0] = i - 1, cj[0] = j, ci[1] = i, cj[1] = j;
    }
    for (int q = 0; q < nc; ++q) {
      const int x = ci[q], y = cj[q];
      const int dd[8] = {node(x, y), node(x + 1, y), node(x, y + 1), node(x + 1, y + 1),
                         ved(x, y),  ved(x + 1, y),  hed(x, y),      hed(x, y + 1)};
      for (int k = 0; k < 8; ++k) {
        signed char sl = -1;
        const int pc = c.pos[dd[k]];
        if (pc >= 0) {
          sl = -2;
          if (pc <= r)  // a lower coupling (an upper one could match the -1 of an unused slot)
            for (int t = 0; t < kNLow; ++t)
              if (c.loff[(size_t)r * kNLow + t] == r - pc) sl = (signed char)t;
        }
        c.cslot[(size_t)r * 32 + 8 * q + k] = sl;
      }
    }
  }
  c.gbnd.assign((size_t)c.N * kNR, 0.0);
  for (int d = 0; d < c.N; ++d)
    if (c.pos[d] < 0)
      for (int b = 0; b < kNR; ++b) c.gbnd[(size_t)d * kNR + b] = ned_bval(c.ns, c.nh, m, d, b);
  if (c.p > kMaxBand || c.n > kMaxRows) {
    why = "half-bandwidth " + std::to_string(c.p) + " exceeds " + std::to_string(kMaxBand) +
          " (l <= 4)";
    return false;
  }
  return true;
}

}  // namespace

struct msfec_ned_batch {
  msfec_ned_params prm{};
  NedConf conf;
  int n_cells = 0;
  double* d_verts = nullptr;
  double* d_a = nullptr;
  double* d_binv = nullptr;
  double* d_r = nullptr;
  double* d_f = nullptr;
  double* d_rows = nullptr;
  double* d_pan = nullptr;    // band factor panels
  double* d_zst = nullptr;    // forward-eliminated right-hand sides
  double* d_X = nullptr;      // solution of the equilibrated system
  double* d_scale = nullptr;  // row scales
  double* d_psi = nullptr;
  double* d_M = nullptr;
  double* d_b = nullptr;
  int* d_status = nullptr;
  int* d_order = nullptr;
  int* d_pos = nullptr;
  int* d_loff = nullptr;
  signed char* d_cslot = nullptr;
  double* d_gbnd = nullptr;
  int* d_utab = nullptr;       // mixed solver: transposed offsets of the band rows
  double* d_mixed = nullptr;   // mixed solver: residual, correction, refinement flags
  bool coef_set = false, computed = false;
  hipStream_t stream = nullptr;
  std::vector<std::array<hipEvent_t, 4>> ev_sets;
  size_t ev_used = 0;
};

namespace {

template <typename T>
msfec_status ned_alloc(T** p, size_t count) {
  NED_HIP(hipMalloc((void**)p, std::max<size_t>(count, 1) * sizeof(T)));
  return MSFEC_OK;
}

void ned_free(msfec_ned_batch* h) {
  void* ptrs[] = {h->d_verts, h->d_a,   h->d_binv, h->d_r,   h->d_f,     h->d_rows,
                  h->d_pan,   h->d_zst, h->d_X,    h->d_scale, h->d_psi, h->d_M,
                  h->d_b,     h->d_status, h->d_order, h->d_pos, h->d_loff, h->d_cslot, h->d_gbnd,
                  h->d_utab,  h->d_mixed};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (auto& set : h->ev_sets)
    for (auto& e : set)
      if (e) (void)hipEventDestroy(e);
  h->ev_sets.clear();
}

// Copied on the handle's stream (that of its last compute call), waited for:
// work still in flight there reads the old buffer to the end.
msfec_status ned_in(hipStream_t s, double* dst, const double* src, size_t count, int loc) {
  if (!src) return ned_fail(MSFEC_ERR_INVALID, "null data pointer");
  NED_HIP(hipMemcpyAsync(dst, src, count * sizeof(double),
                         loc == MSFEC_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
  NED_HIP(hipStreamSynchronize(s));
  return MSFEC_OK;
}

// Copied on the handle's stream and waited for: a device destination holds
// the data when the call returns (a device-to-device hipMemcpy may return first).
msfec_status ned_out(hipStream_t s, double* dst, const double* src, size_t count, int loc) {
  if (!dst) return ned_fail(MSFEC_ERR_INVALID, "null output pointer");
  NED_HIP(hipMemcpyAsync(dst, src, count * sizeof(double),
                         loc == MSFEC_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
  NED_HIP(hipStreamSynchronize(s));
  return MSFEC_OK;
}

NedArgs ned_args(const msfec_ned_batch* h) {
  const NedConf& c = h->conf;
  NedArgs a{};
  a.m = c.m;
  a.ns = c.ns;
  a.nh = c.nh;
  a.N = c.N;
  a.n = c.n;
  a.p = c.p;
  a.full_rhs = h->prm.full_rhs ? 1 : 0;
  a.order = h->d_order;
  a.pos = h->d_pos;
  a.loff = h->d_loff;
  a.cslot = h->d_cslot;
  a.gbnd = h->d_gbnd;
  a.verts = h->d_verts;
  a.a = h->d_a;
  a.binv = h->d_binv;
  a.r = h->d_r;
  a.f = h->d_f;
  a.rows = h->d_rows;
  a.scale = h->d_scale;
  a.X = h->d_X;
  a.psi = h->d_psi;
  a.M = h->d_M;
  a.b = h->d_b;
  a.status = h->d_status;
  return a;
}

}  // namespace

extern "C" {

msfec_status msfec_ned_create(const msfec_ned_params* params, int n_cells, const double* vertices,
                              int vertices_loc, msfec_ned_handle* out) {
  if (!params || !vertices || !out) return ned_fail(MSFEC_ERR_INVALID, "null argument");
  *out = nullptr;
  if (n_cells <= 0) return ned_fail(MSFEC_ERR_INVALID, "n_cells must be positive");
  const int l = params->n_refine_local;
  if (l < 1 || l > 6) return ned_fail(MSFEC_ERR_INVALID, "n_refine_local must be in 1..6");
  if (vertices_loc != MSFEC_HOST && vertices_loc != MSFEC_DEVICE)
    return ned_fail(MSFEC_ERR_INVALID, "bad vertices location");
  if (params->solver != 0 && params->solver != MSFEC_SOLVER_BAND_FP64_MFMA &&
      params->solver != MSFEC_SOLVER_BAND_MIXED)
    return ned_fail(MSFEC_ERR_INVALID, "solver must be BAND_FP64_MFMA or BAND_MIXED");
  if (params->max_refine < 0 || params->max_refine > 16)
    return ned_fail(MSFEC_ERR_INVALID, "max_refine must be in 0..16");
  std::vector<double> hv((size_t)n_cells * 8);
  if (vertices_loc == MSFEC_DEVICE) {
    NED_HIP(hipMemcpy(hv.data(), vertices, hv.size() * sizeof(double), hipMemcpyDeviceToHost));
  } else {
    std::copy(vertices, vertices + hv.size(), hv.begin());
  }
  for (int c = 0; c < n_cells; ++c) {
    const double* v = &hv[(size_t)c * 8];
    const double Hx = v[2] - v[0], Hy = v[5] - v[1];
    const bool box = Hx > 0 && Hy > 0 && std::fabs(v[3] - v[1]) <= 1e-14 * Hy &&
                     std::fabs(v[4] - v[0]) <= 1e-14 * Hx && std::fabs(v[6] - v[2]) <= 1e-14 * Hx &&
                     std::fabs(v[7] - v[5]) <= 1e-14 * Hy;
    if (!box)
      return ned_fail(MSFEC_ERR_INVALID,
                      "cell " + std::to_string(c) + " is not an axis-aligned box in deal.II vertex order");
  }
  msfec_ned_batch* h = new msfec_ned_batch();
  h->prm = *params;
  h->n_cells = n_cells;
  std::string why;
  if (!ned_conf(l, h->conf, why)) {
    delete h;
    return ned_fail(MSFEC_ERR_UNSUPPORTED, why);
  }
  const NedConf& c = h->conf;
  const size_t nc = (size_t)n_cells, nf = (size_t)c.m * c.m;
  msfec_status st = MSFEC_OK;
#define NED_TRY(x)                 \
  do {                             \
    st = (x);                      \
    if (st != MSFEC_OK) {          \
      ned_free(h);                 \
      delete h;                    \
      return st;                   \
    }                              \
  } while (0)
  NED_TRY(ned_alloc(&h->d_verts, nc * 8));
  NED_TRY(ned_alloc(&h->d_a, nc * nf));
  NED_TRY(ned_alloc(&h->d_binv, nc * nf));
  NED_TRY(ned_alloc(&h->d_r, nc * nf));
  NED_TRY(ned_alloc(&h->d_f, nc * nf * 2));
  NED_TRY(ned_alloc(&h->d_rows, nc * c.n * kRowLen));
  NED_TRY(ned_alloc(&h->d_pan, nc * msfec_detail::band_ldl_panel_doubles(c.n)));
  NED_TRY(ned_alloc(&h->d_zst, nc * msfec_detail::band_ldl_z_doubles(c.n)));
  NED_TRY(ned_alloc(&h->d_X, nc * c.n * kNR));
  NED_TRY(ned_alloc(&h->d_scale, nc * c.n));
  NED_TRY(ned_alloc(&h->d_psi, nc * kNR * c.N));
  NED_TRY(ned_alloc(&h->d_M, nc * 64));
  NED_TRY(ned_alloc(&h->d_b, nc * kNR));
  NED_TRY(ned_alloc(&h->d_status, nc));
  NED_TRY(ned_alloc(&h->d_order, c.order.size()));
  NED_TRY(ned_alloc(&h->d_pos, c.pos.size()));
  NED_TRY(ned_alloc(&h->d_loff, c.loff.size()));
  NED_TRY(ned_alloc(&h->d_cslot, c.cslot.size()));
  NED_TRY(ned_alloc(&h->d_gbnd, c.gbnd.size()));
  std::vector<int> utab;
  if (h->prm.solver == MSFEC_SOLVER_BAND_MIXED) {
    utab = msfec_detail::band_upper_table(c.loff, c.n);
    if (utab.empty()) {
      ned_free(h);
      delete h;
      return ned_fail(MSFEC_ERR_UNSUPPORTED, "band rows with more than 16 upper entries");
    }
    NED_TRY(ned_alloc(&h->d_utab, utab.size()));
    NED_TRY(ned_alloc(&h->d_mixed, nc * msfec_detail::band_ldl_mixed_doubles(c.n)));
  }
#undef NED_TRY
  auto up = [&](void* dst, const void* src, size_t bytes) -> msfec_status {
    NED_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return MSFEC_OK;
  };
  msfec_status s2 = MSFEC_OK;
  if (s2 == MSFEC_OK) s2 = up(h->d_verts, hv.data(), hv.size() * sizeof(double));
  if (s2 == MSFEC_OK) s2 = up(h->d_order, c.order.data(), c.order.size() * sizeof(int));
  if (s2 == MSFEC_OK) s2 = up(h->d_pos, c.pos.data(), c.pos.size() * sizeof(int));
  if (s2 == MSFEC_OK) s2 = up(h->d_loff, c.loff.data(), c.loff.size() * sizeof(int));
  if (s2 == MSFEC_OK) s2 = up(h->d_cslot, c.cslot.data(), c.cslot.size());
  if (s2 == MSFEC_OK) s2 = up(h->d_gbnd, c.gbnd.data(), c.gbnd.size() * sizeof(double));
  if (s2 == MSFEC_OK && !utab.empty()) s2 = up(h->d_utab, utab.data(), utab.size() * sizeof(int));
  if (s2 != MSFEC_OK) {
    ned_free(h);
    delete h;
    return s2;
  }
  *out = h;
  return MSFEC_OK;
}

msfec_status msfec_ned_set_coefficients(msfec_ned_handle h, const double* a, const double* binv,
                                        const double* r, const double* f, int loc) {
  if (!h) return ned_fail(MSFEC_ERR_INVALID, "null handle");
  if (loc != MSFEC_HOST && loc != MSFEC_DEVICE) return ned_fail(MSFEC_ERR_INVALID, "bad location");
  const size_t cnt = (size_t)h->n_cells * h->conf.m * h->conf.m;
  msfec_status st = ned_in(h->stream, h->d_a, a, cnt, loc);
  if (st == MSFEC_OK) st = ned_in(h->stream, h->d_binv, binv, cnt, loc);
  if (st == MSFEC_OK) st = ned_in(h->stream, h->d_r, r, cnt, loc);
  if (st == MSFEC_OK) st = ned_in(h->stream, h->d_f, f, 2 * cnt, loc);
  if (st != MSFEC_OK) return st;
  h->coef_set = true;
  return MSFEC_OK;
}

msfec_status msfec_ned_compute(msfec_ned_handle h, void* stream) {
  if (!h) return ned_fail(MSFEC_ERR_INVALID, "null handle");
  if (!h->coef_set) return ned_fail(MSFEC_ERR_STATE, "coefficients not set");
  hipStream_t s = (hipStream_t)stream;
  h->stream = s;
  const NedArgs a = ned_args(h);
  const bool timing = (h->prm.flags & MSFEC_FLAG_TIMING) != 0;
  std::array<hipEvent_t, 4>* ev = nullptr;
  if (timing) {
    if (h->ev_used == h->ev_sets.size()) {
      std::array<hipEvent_t, 4> set{};
      for (auto& e : set) NED_HIP(hipEventCreate(&e));
      h->ev_sets.push_back(set);
    }
    ev = &h->ev_sets[h->ev_used++];
    NED_HIP(hipEventRecord((*ev)[0], s));
  }
  const size_t elem_lds = (size_t)4 * h->conf.N * sizeof(double);
  if (elem_lds > 160 * 1024) return ned_fail(MSFEC_ERR_UNSUPPORTED, "cell too large for LDS staging");
  hipLaunchKernelGGL(ned_rows_kernel, dim3(h->n_cells), dim3(kNThreads), 0, s, a);
  NED_HIP(hipGetLastError());
  if (ev) NED_HIP(hipEventRecord((*ev)[1], s));
  if (h->prm.solver == MSFEC_SOLVER_BAND_MIXED)
    NED_HIP(msfec_detail::band_ldl_mixed_solve(h->n_cells, h->conf.n, h->conf.p, h->d_rows,
                                               h->d_loff, h->d_utab, h->d_pan, h->d_zst, h->d_X,
                                               h->d_mixed, h->d_status, h->prm.max_refine, s));
  else
    NED_HIP(msfec_detail::band_ldl_solve(h->n_cells, h->conf.n, h->conf.p, h->d_rows, h->d_loff,
                                         h->d_pan, h->d_zst, h->d_X, h->d_status, s));
  if (ev) NED_HIP(hipEventRecord((*ev)[2], s));
  hipLaunchKernelGGL(ned_element_kernel, dim3(h->n_cells), dim3(kNThreads), elem_lds, s, a);
  NED_HIP(hipGetLastError());
  if (ev) NED_HIP(hipEventRecord((*ev)[3], s));
  h->computed = true;
  return MSFEC_OK;
}

msfec_status msfec_ned_synchronize(msfec_ned_handle h) {
  if (!h) return ned_fail(MSFEC_ERR_INVALID, "null handle");
  NED_HIP(hipStreamSynchronize(h->stream));
  return MSFEC_OK;
}

msfec_status msfec_ned_get_element(msfec_ned_handle h, double* M, double* b, int loc) {
  if (!h) return ned_fail(MSFEC_ERR_INVALID, "null handle");
  if (!h->computed) return ned_fail(MSFEC_ERR_STATE, "compute not called");
  NED_HIP(hipStreamSynchronize(h->stream));
  msfec_status st = MSFEC_OK;
  if (M) st = ned_out(h->stream, M, h->d_M, (size_t)h->n_cells * 64, loc);
  if (st == MSFEC_OK && b) st = ned_out(h->stream, b, h->d_b, (size_t)h->n_cells * kNR, loc);
  return st;
}

msfec_status msfec_ned_get_basis(msfec_ned_handle h, double* Psi, int loc) {
  if (!h) return ned_fail(MSFEC_ERR_INVALID, "null handle");
  if (!h->computed) return ned_fail(MSFEC_ERR_STATE, "compute not called");
  NED_HIP(hipStreamSynchronize(h->stream));
  return ned_out(h->stream, Psi, h->d_psi, (size_t)h->n_cells * kNR * h->conf.N, loc);
}

msfec_status msfec_ned_get_cell_status(msfec_ned_handle h, int* status) {
  if (!h || !status) return ned_fail(MSFEC_ERR_INVALID, "null argument");
  if (!h->computed) return ned_fail(MSFEC_ERR_STATE, "compute not called");
  NED_HIP(hipStreamSynchronize(h->stream));
  NED_HIP(hipMemcpy(status, h->d_status, (size_t)h->n_cells * sizeof(int), hipMemcpyDeviceToHost));
  return MSFEC_OK;
}

msfec_status msfec_ned_kernel_times(msfec_ned_handle h, float* ms, int max, int* n) {
  if (!h || !ms) return ned_fail(MSFEC_ERR_INVALID, "null argument");
  if (h->ev_used == 0) return ned_fail(MSFEC_ERR_STATE, "no timed compute (set MSFEC_FLAG_TIMING)");
  double acc[3] = {0.0, 0.0, 0.0};
  for (size_t c = 0; c < h->ev_used; ++c) {
    NED_HIP(hipEventSynchronize(h->ev_sets[c][3]));
    for (int i = 0; i < 3; ++i) {
      float t = 0.f;
      NED_HIP(hipEventElapsedTime(&t, h->ev_sets[c][i], h->ev_sets[c][i + 1]));
      acc[i] += t;
    }
  }
  int k = 0;
  for (; k < 3 && k < max; ++k) ms[k] = (float)(acc[k] / (double)h->ev_used);
  if (n) *n = k;
  h->ev_used = 0;
  return MSFEC_OK;
}

msfec_status msfec_ned_destroy(msfec_ned_handle h) {
  if (!h) return MSFEC_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  ned_free(h);
  delete h;
  return MSFEC_OK;
}

}  // extern "C"
