// RT_DQ: the mixed lowest-order Raviart-Thomas x DQ0 multiscale basis in two
// dimensions on MI355X (gfx950), behind the msfec_rt_* entry points of
// include/msfec.h.
//
// Reference: source/RT_DQ/rt_dq_basis.cc (three-dimensional RTDQBasis).  Per
// coarse box cell K and coarse face n, on the 2^l x 2^l fine grid:
//   (tau, A^-1 sigma) - (div tau, u) = 0,   (v, div sigma) = -+(v, 1/|K|)
// (rt_dq_basis.cc:495-538) with the normal flux of the coarse Raviart-Thomas
// function of face n on the boundary (:267-380), solved per face (:1114-1170),
// u replaced by 1 (set_u_to_std, :1039-1043), then the element matrix of
// :822-914.  Only the flux field reaches the outputs.
//
// The flux is found without the saddle-point system.  On an axis-aligned box
// the divergence-free Raviart-Thomas fields with zero boundary flux are
// exactly the curls (d_y psi, -d_x psi) of Q1 functions psi vanishing on the
// boundary, and the coarse function c_n already has the required divergence
// and boundary flux.  So sigma_n = c_n + curl psi_n, where psi_n minimises
// (c_n + curl psi, A^-1 (c_n + curl psi)):
//   (grad phi, A^-1 grad psi_n) = -(curl phi, A^-1 c_n)   for all interior phi,
// a symmetric positive definite Q1 problem with coefficient 1/a, (2^l - 1)^2
// unknowns and half-bandwidth 2^l, instead of the indefinite mixed system of
// 2 m (m - 1) + m^2 unknowns.  The first equation of the mixed problem holds
// for this sigma with some u because A^-1 sigma_n is orthogonal to every
// divergence-free test field; the u of the reference is replaced by 1 before
// it is used, so it is not formed.
//
// Layout of the work:
//   rt_rows_kernel     one workgroup per cell, one thread per interior node:
//                      its row of the Q1 system (diagonal, then the lower
//                      neighbours left, bottom-right, bottom, bottom-left) and
//                      the 4 right-hand sides, from exact integrals with 1/a
//                      staged in LDS, or, for a tensor A given at the
//                      QGauss(2) points, by that rule with A^-1 at each point.
//   band_ldl_solve     (msfec_q.hip) one wave per cell: the matrix-core band
//                      L D L^T, 4 pivots per step, with the forward
//                      elimination and the backward sweep.
//   rt_element_kernel  one workgroup per cell: psi in LDS, the fine fluxes of
//                      the 4 bases to HBM and the 5 x 5 element matrix and
//                      load from exact integrals per fine cell (QGauss(2)
//                      with A^-1 at each point for a tensor A).
// With a tensor A the curl form becomes (curl phi, A^-1 curl psi), still
// symmetric positive definite; both sides use the reference's quadrature.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
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
}

namespace {

constexpr int kNLow = 16;             // lower entries of a band row (layout of band_ldl_solve)
constexpr int kNR = 8;                // right-hand-side columns of band_ldl_solve
constexpr int kRowLen = kNLow + kNR;  // row table: lower values, right-hand sides
constexpr int kNB = 4;                // face bases per cell
constexpr int kNE = 25;               // element matrix entries (5 x 5)
constexpr int kNThreads = 256;
constexpr int kMaxL = 5;              // half-bandwidth 2^l must stay within the band window (60)

struct RtArgs {
  int m;     // fine cells per direction
  int n;     // interior nodes (m - 1)^2
  int ne;    // fine edges 2 m (m + 1)
  int tensor;  // 0: a I per fine cell; 1: the tensor (xx, xy, yy) at the 4 Gauss points
               // (the kernels are instantiated per mode)
  const int* loff;      // [n][kNLow] row - column offsets of the lower entries, -1 unused
  const double* verts;  // [cells][4][2], deal.II vertex order
  const double* a;      // [cells][m * m] diffusion per fine cell, or [cells][m * m][4][3]
  const double* f;      // [cells][m * m] load per fine cell
  double* rows;         // [cells][n][kRowLen]
  const double* X;      // [cells][n][kNR] psi of the 4 bases at the interior nodes
  double* sigma;        // [cells][4][ne] fine fluxes
  double* M;            // [cells][25]
  double* b;            // [cells][5]
  int* status;          // [cells]
};

struct RtBox {
  double hx, hy;
};

__device__ __forceinline__ RtBox rt_box(const RtArgs& a, int cell) {
  const double* v = a.verts + (size_t)cell * 8;
  return RtBox{(v[2] - v[0]) / a.m, (v[5] - v[1]) / a.m};
}

// Flux of coarse basis nb through fine vertical edge at x-node P (+x) and
// through fine horizontal edge at y-node Q (+y): the coarse functions are
// (1 - X, 0) / Hy, (X, 0) / Hy, (0, 1 - Y) / Hx, (0, Y) / Hx in the coarse
// reference coordinates (the Piola map of deal.II's unit-square functions,
// include/functions/basis_raviart_thomas.tpp), and a fine edge is 1/m of a
// coarse face.
__device__ __forceinline__ double rt_cflux_v(int nb, int P, int m) {
  const double X = (double)P / m;
  return nb == 0 ? (1.0 - X) / m : (nb == 1 ? X / m : 0.0);
}
__device__ __forceinline__ double rt_cflux_h(int nb, int Q, int m) {
  const double Y = (double)Q / m;
  return nb == 2 ? (1.0 - Y) / m : (nb == 3 ? Y / m : 0.0);
}

// QGauss(2) on [0, 1]: the points; each of the 4 points of a fine box
// carries the weight hx hy / 4.
__constant__ double kRtGauss[2] = {0.21132486540518711775, 0.78867513459481288225};

// A^-1 at Gauss point qp of fine cell fc (tensor mode): the inverse of the
// symmetric (xx, xy, yy), as DiffusionInverse_A gives it there
// (source/equation_data/eqn_coeff_A.cc:194-242).  B = (b00, b01, b11).
__device__ __forceinline__ void rt_ainv_qp(const double* ac, int fc, int qp, double& b00,
                                           double& b01, double& b11) {
  const double* t = ac + ((size_t)fc * 4 + qp) * 3;
  const double xx = t[0], xy = t[1], yy = t[2];
  const double id = 1.0 / (xx * yy - xy * xy);
  b00 = yy * id;
  b01 = -xy * id;
  b11 = xx * id;
}

// curl (d_y phi, -d_x phi) of the bilinear function of corner (ex, ey) at
// (s, t) of a fine box (hx, hy)
__device__ __forceinline__ void rt_corner_curl(int ex, int ey, double s, double t, double hx,
                                               double hy, double& cx, double& cy) {
  const double lx = ex ? s : 1.0 - s, ly = ey ? t : 1.0 - t;
  cx = (ey ? 1.0 : -1.0) * lx / hy;
  cy = -(ex ? 1.0 : -1.0) * ly / hx;
}

// Row of interior node (i, j) in the Q1 system with coefficient k = 1/a.
// Exact integrals on a fine box (hx, hy), al = hy / hx, be = hx / hy:
//   diagonal          sum of the 4 cells  k (al + be) / 3
//   left  (i-1, j)    (k00 + k01) (be / 6 - al / 3)
//   below (i, j-1)    (k00 + k10) (al / 6 - be / 3)
//   diagonal nbrs     k (al + be) / -6     (k00 below-left, k10 below-right)
// with k00, k10, k01, k11 the cells below-left, below-right, above-left and
// above-right of the node.  Right-hand side of basis nb:
// -(curl phi, k c_nb) summed over the 4 cells, with the 1-D exact integrals
// of a linear function against the hat function's two halves.
template <bool TENSOR>
__global__ __launch_bounds__(kNThreads) void rt_rows_kernel(RtArgs a) {
  extern __shared__ double kap[];  // [m * m] 1 / a (scalar mode)
  const int cell = blockIdx.x, m = a.m, nf = m * m;
  if (!TENSOR) {
    const double* ac = a.a + (size_t)cell * nf;
    for (int c = threadIdx.x; c < nf; c += blockDim.x) kap[c] = 1.0 / ac[c];
  }
  __syncthreads();
  const RtBox bx = rt_box(a, cell);
  const double al = bx.hy / bx.hx, be = bx.hx / bx.hy;
  const double Hx = bx.hx * m, Hy = bx.hy * m;
  const int mi = m - 1;
  if (TENSOR) {
    // the same row by QGauss(2) with A^-1 at each point: per cell around the
    // node and point, w curl(phi_node)^T B curl(phi_other) for the lower
    // neighbours and -w curl(phi_node)^T B c_n for the right-hand sides
    const double* ac = a.a + (size_t)cell * nf * 12;
    const double w = 0.25 * bx.hx * bx.hy;
    for (int r = threadIdx.x; r < a.n; r += blockDim.x) {
      const int i = r % mi + 1, j = r / mi + 1;
      double vd = 0.0, vl = 0.0, vbr = 0.0, vb = 0.0, vbl = 0.0;
      double rh[kNB] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int dy = 0; dy < 2; ++dy)
#pragma unroll
        for (int dx = 0; dx < 2; ++dx) {
          const int ci = i - 1 + dx, cj = j - 1 + dy, fc = cj * m + ci;
          const int bxn = 1 - dx, byn = 1 - dy;  // the node's corner in this cell
#pragma unroll
          for (int qp = 0; qp < 4; ++qp) {
            const double sq = kRtGauss[qp & 1], tq = kRtGauss[qp >> 1];
            double b00, b01, b11;
            rt_ainv_qp(ac, fc, qp, b00, b01, b11);
            double cx, cy;
            rt_corner_curl(bxn, byn, sq, tq, bx.hx, bx.hy, cx, cy);
            const double ux = w * (b00 * cx + b01 * cy), uy = w * (b01 * cx + b11 * cy);
#pragma unroll
            for (int ey = 0; ey < 2; ++ey)
#pragma unroll
              for (int ex = 0; ex < 2; ++ex) {
                double ox, oy;
                rt_corner_curl(ex, ey, sq, tq, bx.hx, bx.hy, ox, oy);
                const double v = ux * ox + uy * oy;
                const int di = ex - bxn, dj = ey - byn;
                if (di == 0 && dj == 0) vd += v;
                else if (di == -1 && dj == 0) vl += v;
                else if (di == 1 && dj == -1) vbr += v;
                else if (di == 0 && dj == -1) vb += v;
                else if (di == -1 && dj == -1) vbl += v;
              }
            const double X = (ci + sq) / m, Y = (cj + tq) / m;
            rh[0] -= ux * (1.0 - X) / Hy;
            rh[1] -= ux * X / Hy;
            rh[2] -= uy * (1.0 - Y) / Hx;
            rh[3] -= uy * Y / Hx;
          }
        }
      // the present couplings packed to the front of the row, written
      // straight to the table (a register array indexed by the running
      // count would live in scratch memory)
      double* dst = a.rows + ((size_t)cell * a.n + r) * kRowLen;
      dst[0] = vd;
      int t = 1;
      if (i > 1) dst[t++] = vl;
      if (j > 1 && i < mi) dst[t++] = vbr;
      if (j > 1) dst[t++] = vb;
      if (j > 1 && i > 1) dst[t++] = vbl;
      for (; t < kNLow; ++t) dst[t] = 0.0;
#pragma unroll
      for (int nb = 0; nb < kNB; ++nb) dst[kNLow + nb] = rh[nb];
#pragma unroll
      for (int q = kNLow + kNB; q < kRowLen; ++q) dst[q] = 0.0;
    }
    return;
  }
  for (int r = threadIdx.x; r < a.n; r += blockDim.x) {
    const int i = r % mi + 1, j = r / mi + 1;
    const double k00 = kap[(j - 1) * m + i - 1], k10 = kap[(j - 1) * m + i];
    const double k01 = kap[j * m + i - 1], k11 = kap[j * m + i];
    double* dst = a.rows + ((size_t)cell * a.n + r) * kRowLen;
    dst[0] = (k00 + k10 + k01 + k11) * (al + be) / 3.0;
    int t = 1;
    if (i > 1) dst[t++] = (k00 + k01) * (be / 6.0 - al / 3.0);          // left
    if (j > 1 && i < mi) dst[t++] = -k10 * (al + be) / 6.0;               // below-right
    if (j > 1) dst[t++] = (k00 + k10) * (al / 6.0 - be / 3.0);          // below
    if (j > 1 && i > 1) dst[t++] = -k00 * (al + be) / 6.0;                // below-left
    for (; t < kNLow; ++t) dst[t] = 0.0;
    // x bases: c = (g(X), 0), int g d_y phi over the 4 cells
    {
      const double gm = (double)(i - 1) / m, g0 = (double)i / m, gp = (double)(i + 1) / m;
      const double wl = bx.hx / 6.0 * (k00 - k01), wr = bx.hx / 6.0 * (k10 - k11);
      // basis 0: g = (1 - X) / Hy; basis 1: g = X / Hy
      dst[kNLow + 0] = -(wl * ((1.0 - gm) + 2.0 * (1.0 - g0)) + wr * (2.0 * (1.0 - g0) + (1.0 - gp))) / Hy;
      dst[kNLow + 1] = -(wl * (gm + 2.0 * g0) + wr * (2.0 * g0 + gp)) / Hy;
    }
    // y bases: c = (0, g(Y)), int -g d_x phi over the 4 cells
    {
      const double gm = (double)(j - 1) / m, g0 = (double)j / m, gp = (double)(j + 1) / m;
      const double wb = bx.hy / 6.0 * (k10 - k00), wt = bx.hy / 6.0 * (k11 - k01);
      dst[kNLow + 2] = -(wb * ((1.0 - gm) + 2.0 * (1.0 - g0)) + wt * (2.0 * (1.0 - g0) + (1.0 - gp))) / Hx;
      dst[kNLow + 3] = -(wb * (gm + 2.0 * g0) + wt * (2.0 * g0 + gp)) / Hx;
    }
#pragma unroll
    for (int q = kNLow + kNB; q < kRowLen; ++q) dst[q] = 0.0;
  }
}

// Fine fluxes and the element matrix (rt_dq_basis.cc:822-914 with u = 1):
//   M[i][j] = (sigma_i, A^-1 sigma_j)       i, j < 4
//   M[i][4] = -(div sigma_i, 1),  M[4][j] = (1, div sigma_j),  M[4][4] = 0
//   b = (0, 0, 0, 0, (1, f))
// Each thread takes every 256th fine cell: its 4 edge fluxes per basis from
// psi in LDS, the exact mass integral of the two linear flux components and
// the out-flux; the sums go across lanes and waves.
template <bool TENSOR>
__global__ __launch_bounds__(kNThreads) void rt_element_kernel(RtArgs a) {
  extern __shared__ double ps[];  // [4][(m + 1)^2] psi on every node, 0 on the boundary
  __shared__ double part[kNThreads / 64][16];
  const int cell = blockIdx.x, tid = threadIdx.x, m = a.m, mi = m - 1;
  const int nn = (m + 1) * (m + 1);
  const double* x = a.X + (size_t)cell * a.n * kNR;
  for (int t = tid; t < kNB * nn; t += kNThreads) {
    const int nb = t / nn, node = t - nb * nn;
    const int i = node % (m + 1), j = node / (m + 1);
    const bool in = i > 0 && i < m && j > 0 && j < m;
    ps[t] = in ? x[(size_t)((j - 1) * mi + i - 1) * kNR + nb] : 0.0;
  }
  __syncthreads();
  // fine fluxes: vertical edge (P, q) at q (m + 1) + P, horizontal (p, Q) at nv + Q m + p
  const int nv = m * (m + 1);
  double* sg = a.sigma + (size_t)cell * kNB * a.ne;
  for (int t = tid; t < kNB * a.ne; t += kNThreads) {
    const int nb = t / a.ne, e = t - nb * a.ne;
    const double* pb = ps + nb * nn;
    double v;
    if (e < nv) {
      const int P = e % (m + 1), q = e / (m + 1);
      v = rt_cflux_v(nb, P, m) + (pb[(q + 1) * (m + 1) + P] - pb[q * (m + 1) + P]);
    } else {
      const int p = (e - nv) % m, Q = (e - nv) / m;
      v = rt_cflux_h(nb, Q, m) - (pb[Q * (m + 1) + p + 1] - pb[Q * (m + 1) + p]);
    }
    sg[t] = v;
  }
  const RtBox bx = rt_box(a, cell);
  const double rx = bx.hx / bx.hy, ry = bx.hy / bx.hx;
  double acc[16];  // 10 mass entries (upper triangle), 4 out-fluxes, the load
#pragma unroll
  for (int k = 0; k < 16; ++k) acc[k] = 0.0;
  const int nf = m * m;
  for (int c = tid; c < nf; c += kNThreads) {
    const int p = c % m, q = c / m;
    const double kap = TENSOR ? 0.0 : 1.0 / a.a[(size_t)cell * nf + c];
    double FL[kNB], FR[kNB], FB[kNB], FT[kNB];
#pragma unroll
    for (int nb = 0; nb < kNB; ++nb) {
      const double* pb = ps + nb * nn;
      const double s00 = pb[q * (m + 1) + p], s10 = pb[q * (m + 1) + p + 1];
      const double s01 = pb[(q + 1) * (m + 1) + p], s11 = pb[(q + 1) * (m + 1) + p + 1];
      FL[nb] = rt_cflux_v(nb, p, m) + (s01 - s00);
      FR[nb] = rt_cflux_v(nb, p + 1, m) + (s11 - s10);
      FB[nb] = rt_cflux_h(nb, q, m) - (s10 - s00);
      FT[nb] = rt_cflux_h(nb, q + 1, m) - (s11 - s01);
    }
    if (TENSOR) {
      // QGauss(2) with A^-1 at each point: sigma_i there is
      // ((FL (1 - s) + FR s) / hy, (FB (1 - t) + FT t) / hx)
      const double* ac = a.a + (size_t)cell * nf * 12;
      const double w = 0.25 * bx.hx * bx.hy;
#pragma unroll
      for (int qp = 0; qp < 4; ++qp) {
        const double sq = kRtGauss[qp & 1], tq = kRtGauss[qp >> 1];
        double b00, b01, b11;
        rt_ainv_qp(ac, c, qp, b00, b01, b11);
        double sx[kNB], sy[kNB];
#pragma unroll
        for (int nb = 0; nb < kNB; ++nb) {
          sx[nb] = (FL[nb] * (1.0 - sq) + FR[nb] * sq) / bx.hy;
          sy[nb] = (FB[nb] * (1.0 - tq) + FT[nb] * tq) / bx.hx;
        }
        int k = 0;
#pragma unroll
        for (int i = 0; i < kNB; ++i) {
          const double ux = w * (b00 * sx[i] + b01 * sy[i]), uy = w * (b01 * sx[i] + b11 * sy[i]);
#pragma unroll
          for (int j = i; j < kNB; ++j) acc[k++] += ux * sx[j] + uy * sy[j];
        }
      }
    } else {
      int k = 0;
#pragma unroll
      for (int i = 0; i < kNB; ++i)
#pragma unroll
        for (int j = i; j < kNB; ++j) {
          const double xx = (FL[i] * FL[j] + FR[i] * FR[j]) / 3.0 + (FL[i] * FR[j] + FR[i] * FL[j]) / 6.0;
          const double yy = (FB[i] * FB[j] + FT[i] * FT[j]) / 3.0 + (FB[i] * FT[j] + FT[i] * FB[j]) / 6.0;
          acc[k++] += kap * (rx * xx + ry * yy);
        }
    }
#pragma unroll
    for (int nb = 0; nb < kNB; ++nb) acc[10 + nb] += (FR[nb] - FL[nb]) + (FT[nb] - FB[nb]);
    acc[14] += a.f[(size_t)cell * nf + c] * bx.hx * bx.hy;
  }
#pragma unroll
  for (int k = 0; k < 15; ++k)
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) acc[k] += __shfl_xor(acc[k], o);
  const int wave = tid >> 6, lane = tid & 63;
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < 15; ++k) part[wave][k] = acc[k];
  __syncthreads();
  if (tid == 0) {
    double s[15];
#pragma unroll
    for (int k = 0; k < 15; ++k) {
      double v = 0.0;
#pragma unroll
      for (int w = 0; w < kNThreads / 64; ++w) v += part[w][k];
      s[k] = v;
    }
    double Mc[kNE];
    int k = 0;
    for (int i = 0; i < kNB; ++i)
      for (int j = i; j < kNB; ++j) {
        Mc[i * 5 + j] = s[k];
        Mc[j * 5 + i] = s[k];
        ++k;
      }
    bool finite = true;
    for (int nb = 0; nb < kNB; ++nb) {
      Mc[nb * 5 + 4] = -s[10 + nb];
      Mc[4 * 5 + nb] = s[10 + nb];
    }
    Mc[24] = 0.0;
    double* Mo = a.M + (size_t)cell * kNE;
    for (int t = 0; t < kNE; ++t) {
      Mo[t] = Mc[t];
      finite = finite && isfinite(Mc[t]);
    }
    double* bo = a.b + (size_t)cell * 5;
    for (int t = 0; t < 4; ++t) bo[t] = 0.0;
    bo[4] = s[14];
    if (!finite && a.status[cell] == MSFEC_CELL_OK) a.status[cell] = MSFEC_CELL_NONFINITE;
  }
}

// set_global_weights (rt_dq_basis.cc:1011-1036): sigma = sum_i w_i sigma_i,
// u = w_4 on every fine cell (the u block of every basis is 1).
__global__ __launch_bounds__(kNThreads) void rt_reconstruct_kernel(const double* sigma,
                                                                   const double* w, int ne,
                                                                   int nf, double* sig_out,
                                                                   double* u_out) {
  const int cell = blockIdx.x;
  const double* wc = w + (size_t)cell * 5;
  const double* sg = sigma + (size_t)cell * kNB * ne;
  if (sig_out)
    for (int e = threadIdx.x; e < ne; e += blockDim.x) {
      double v = 0.0;
#pragma unroll
      for (int nb = 0; nb < kNB; ++nb) v += wc[nb] * sg[(size_t)nb * ne + e];
      sig_out[(size_t)cell * ne + e] = v;
    }
  if (u_out)
    for (int c = threadIdx.x; c < nf; c += blockDim.x) u_out[(size_t)cell * nf + c] = wc[4];
}

msfec_status rt_fail(msfec_status st, const std::string& msg) {
  msfec_detail::g_last_error = msg;
  return st;
}

#define RT_HIP(call)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return rt_fail(MSFEC_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

}  // namespace

struct msfec_rt_batch {
  msfec_rt_params prm{};
  int n_cells = 0, l = 0, m = 0, n = 0, ne = 0;
  double* d_verts = nullptr;
  double* d_a = nullptr;
  double* d_f = nullptr;
  double* d_rows = nullptr;
  double* d_pan = nullptr;
  double* d_zst = nullptr;
  double* d_X = nullptr;
  double* d_sigma = nullptr;
  double* d_M = nullptr;
  double* d_b = nullptr;
  double* d_w = nullptr;
  int* d_status = nullptr;
  int* d_loff = nullptr;
  int bw = 0;  // half-bandwidth of the reduced system (largest lower offset)
  bool coef_set = false, computed = false;
  hipStream_t stream = nullptr;
  std::vector<std::array<hipEvent_t, 4>> ev_sets;
  size_t ev_used = 0;
};

namespace {

template <typename T>
msfec_status rt_alloc(T** p, size_t count) {
  RT_HIP(hipMalloc((void**)p, std::max<size_t>(count, 1) * sizeof(T)));
  return MSFEC_OK;
}

void rt_free(msfec_rt_batch* h) {
  void* ptrs[] = {h->d_verts, h->d_a, h->d_f,     h->d_rows, h->d_pan,    h->d_zst, h->d_X,
                  h->d_sigma, h->d_M, h->d_b,     h->d_w,    h->d_status, h->d_loff};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (auto& set : h->ev_sets)
    for (auto& e : set)
      if (e) (void)hipEventDestroy(e);
  h->ev_sets.clear();
}

// Copied on the handle's stream (that of its last compute call), waited for:
// work still in flight there reads the old buffer to the end.
msfec_status rt_in(hipStream_t s, double* dst, const double* src, size_t count, int loc) {
  RT_HIP(hipMemcpyAsync(dst, src, count * sizeof(double),
                        loc == MSFEC_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
  RT_HIP(hipStreamSynchronize(s));
  return MSFEC_OK;
}

// Copied on the handle's stream and waited for: a device destination holds
// the data when the call returns (a device-to-device hipMemcpy may return first).
msfec_status rt_out(hipStream_t s, double* dst, const double* src, size_t count, int loc) {
  if (!dst) return rt_fail(MSFEC_ERR_INVALID, "null output pointer");
  RT_HIP(hipMemcpyAsync(dst, src, count * sizeof(double),
                        loc == MSFEC_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
  RT_HIP(hipStreamSynchronize(s));
  return MSFEC_OK;
}

RtArgs rt_args(const msfec_rt_batch* h) {
  RtArgs a{};
  a.m = h->m;
  a.n = h->n;
  a.ne = h->ne;
  a.tensor = h->prm.coef_mode == MSFEC_COEF_TENSOR_QP ? 1 : 0;
  a.loff = h->d_loff;
  a.verts = h->d_verts;
  a.a = h->d_a;
  a.f = h->d_f;
  a.rows = h->d_rows;
  a.X = h->d_X;
  a.sigma = h->d_sigma;
  a.M = h->d_M;
  a.b = h->d_b;
  a.status = h->d_status;
  return a;
}

// Row offsets of the lower entries in the order rt_rows_kernel writes them.
std::vector<int> rt_loff(int m) {
  const int mi = m - 1, n = mi * mi;
  std::vector<int> loff((size_t)std::max(n, 1) * kNLow, -1);
  for (int r = 0; r < n; ++r) {
    const int i = r % mi + 1, j = r / mi + 1;
    int* o = &loff[(size_t)r * kNLow];
    int t = 0;
    o[t++] = 0;
    if (i > 1) o[t++] = 1;
    if (j > 1 && i < mi) o[t++] = mi - 1;
    if (j > 1) o[t++] = mi;
    if (j > 1 && i > 1) o[t++] = mi + 1;
  }
  return loff;
}

}  // namespace

extern "C" {

msfec_status msfec_rt_create(const msfec_rt_params* params, int n_cells, const double* vertices,
                             int vertices_loc, msfec_rt_handle* out) {
  if (!params || !vertices || !out) return rt_fail(MSFEC_ERR_INVALID, "null argument");
  *out = nullptr;
  if (n_cells <= 0) return rt_fail(MSFEC_ERR_INVALID, "n_cells must be positive");
  const int l = params->n_refine_local;
  if (l < 1) return rt_fail(MSFEC_ERR_INVALID, "n_refine_local must be at least 1");
  if (l > kMaxL)
    return rt_fail(MSFEC_ERR_UNSUPPORTED,
                   "n_refine_local above 5: the band of the reduced system exceeds the band window");
  if (vertices_loc != MSFEC_HOST && vertices_loc != MSFEC_DEVICE)
    return rt_fail(MSFEC_ERR_INVALID, "bad vertices location");
  if (params->coef_mode != MSFEC_COEF_SCALAR_CELL && params->coef_mode != MSFEC_COEF_TENSOR_QP)
    return rt_fail(MSFEC_ERR_INVALID, "coef_mode must be MSFEC_COEF_SCALAR_CELL or MSFEC_COEF_TENSOR_QP");
  std::vector<double> hv((size_t)n_cells * 8);
  if (vertices_loc == MSFEC_DEVICE) {
    RT_HIP(hipMemcpy(hv.data(), vertices, hv.size() * sizeof(double), hipMemcpyDeviceToHost));
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
      return rt_fail(MSFEC_ERR_INVALID,
                     "cell " + std::to_string(c) + " is not an axis-aligned box in deal.II vertex order");
  }
  msfec_rt_batch* h = new msfec_rt_batch();
  h->prm = *params;
  h->n_cells = n_cells;
  h->l = l;
  h->m = 1 << l;
  h->n = (h->m - 1) * (h->m - 1);
  h->ne = 2 * h->m * (h->m + 1);
  const size_t nc = (size_t)n_cells, nf = (size_t)h->m * h->m;
  const std::vector<int> loff = rt_loff(h->m);
  h->bw = loff.empty() ? 0 : *std::max_element(loff.begin(), loff.end());
  msfec_status st = MSFEC_OK;
#define RT_TRY(x)         \
  do {                    \
    st = (x);             \
    if (st != MSFEC_OK) { \
      rt_free(h);         \
      delete h;           \
      return st;          \
    }                     \
  } while (0)
  RT_TRY(rt_alloc(&h->d_verts, nc * 8));
  RT_TRY(rt_alloc(&h->d_a, nc * nf * (params->coef_mode == MSFEC_COEF_TENSOR_QP ? 12 : 1)));
  RT_TRY(rt_alloc(&h->d_f, nc * nf));
  RT_TRY(rt_alloc(&h->d_rows, nc * h->n * kRowLen));
  RT_TRY(rt_alloc(&h->d_pan, nc * msfec_detail::band_ldl_panel_doubles(h->n)));
  RT_TRY(rt_alloc(&h->d_zst, nc * msfec_detail::band_ldl_z_doubles(h->n)));
  RT_TRY(rt_alloc(&h->d_X, nc * h->n * kNR));
  RT_TRY(rt_alloc(&h->d_sigma, nc * kNB * h->ne));
  RT_TRY(rt_alloc(&h->d_M, nc * kNE));
  RT_TRY(rt_alloc(&h->d_b, nc * 5));
  RT_TRY(rt_alloc(&h->d_w, nc * 5));
  RT_TRY(rt_alloc(&h->d_status, nc));
  RT_TRY(rt_alloc(&h->d_loff, loff.size()));
  RT_TRY(rt_in(h->stream, h->d_verts, hv.data(), hv.size(), MSFEC_HOST));
  {
    hipError_t e = hipMemcpy(h->d_loff, loff.data(), loff.size() * sizeof(int), hipMemcpyHostToDevice);
    if (e != hipSuccess) RT_TRY(rt_fail(MSFEC_ERR_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e)));
  }
#undef RT_TRY
  *out = h;
  return MSFEC_OK;
}

msfec_status msfec_rt_set_coefficients(msfec_rt_handle h, const double* a, const double* f,
                                       int loc) {
  if (!h || !a) return rt_fail(MSFEC_ERR_INVALID, "null argument");
  if (loc != MSFEC_HOST && loc != MSFEC_DEVICE) return rt_fail(MSFEC_ERR_INVALID, "bad location");
  const size_t cnt = (size_t)h->n_cells * h->m * h->m;
  const size_t acnt = cnt * (h->prm.coef_mode == MSFEC_COEF_TENSOR_QP ? 12 : 1);
  msfec_status st = rt_in(h->stream, h->d_a, a, acnt, loc);
  if (st != MSFEC_OK) return st;
  if (f) {
    st = rt_in(h->stream, h->d_f, f, cnt, loc);
  } else {
    const std::vector<double> ones(cnt, 1.0);
    st = rt_in(h->stream, h->d_f, ones.data(), cnt, MSFEC_HOST);
  }
  if (st != MSFEC_OK) return st;
  h->coef_set = true;
  return MSFEC_OK;
}

msfec_status msfec_rt_compute(msfec_rt_handle h, void* stream) {
  if (!h) return rt_fail(MSFEC_ERR_INVALID, "null handle");
  if (!h->coef_set) return rt_fail(MSFEC_ERR_STATE, "coefficients not set");
  hipStream_t s = (hipStream_t)stream;
  h->stream = s;
  const RtArgs a = rt_args(h);
  const bool timing = (h->prm.flags & MSFEC_FLAG_TIMING) != 0;
  std::array<hipEvent_t, 4>* ev = nullptr;
  if (timing) {
    if (h->ev_used == h->ev_sets.size()) {
      std::array<hipEvent_t, 4> set{};
      for (auto& e : set) RT_HIP(hipEventCreate(&e));
      h->ev_sets.push_back(set);
    }
    ev = &h->ev_sets[h->ev_used++];
    RT_HIP(hipEventRecord((*ev)[0], s));
  }
  const size_t rows_lds = (size_t)h->m * h->m * sizeof(double);
  const size_t elem_lds = (size_t)kNB * (h->m + 1) * (h->m + 1) * sizeof(double);
  if (h->n > 0) {
    if (a.tensor)
      hipLaunchKernelGGL(rt_rows_kernel<true>, dim3(h->n_cells), dim3(kNThreads), 0, s, a);
    else
      hipLaunchKernelGGL(rt_rows_kernel<false>, dim3(h->n_cells), dim3(kNThreads), rows_lds, s, a);
    RT_HIP(hipGetLastError());
  }
  if (ev) RT_HIP(hipEventRecord((*ev)[1], s));
  RT_HIP(msfec_detail::band_ldl_solve(h->n_cells, h->n, h->bw, h->d_rows, h->d_loff, h->d_pan, h->d_zst,
                                      h->d_X, h->d_status, s));
  if (ev) RT_HIP(hipEventRecord((*ev)[2], s));
  if (a.tensor)
    hipLaunchKernelGGL(rt_element_kernel<true>, dim3(h->n_cells), dim3(kNThreads), elem_lds, s, a);
  else
    hipLaunchKernelGGL(rt_element_kernel<false>, dim3(h->n_cells), dim3(kNThreads), elem_lds, s, a);
  RT_HIP(hipGetLastError());
  if (ev) RT_HIP(hipEventRecord((*ev)[3], s));
  h->computed = true;
  return MSFEC_OK;
}

msfec_status msfec_rt_synchronize(msfec_rt_handle h) {
  if (!h) return rt_fail(MSFEC_ERR_INVALID, "null handle");
  RT_HIP(hipStreamSynchronize(h->stream));
  return MSFEC_OK;
}

msfec_status msfec_rt_get_element(msfec_rt_handle h, double* M, double* b, int loc) {
  if (!h) return rt_fail(MSFEC_ERR_INVALID, "null handle");
  if (!h->computed) return rt_fail(MSFEC_ERR_STATE, "compute not called");
  RT_HIP(hipStreamSynchronize(h->stream));
  msfec_status st = MSFEC_OK;
  if (M) st = rt_out(h->stream, M, h->d_M, (size_t)h->n_cells * kNE, loc);
  if (st == MSFEC_OK && b) st = rt_out(h->stream, b, h->d_b, (size_t)h->n_cells * 5, loc);
  return st;
}

msfec_status msfec_rt_get_basis(msfec_rt_handle h, double* sigma, int loc) {
  if (!h) return rt_fail(MSFEC_ERR_INVALID, "null handle");
  if (!h->computed) return rt_fail(MSFEC_ERR_STATE, "compute not called");
  RT_HIP(hipStreamSynchronize(h->stream));
  return rt_out(h->stream, sigma, h->d_sigma, (size_t)h->n_cells * kNB * h->ne, loc);
}

msfec_status msfec_rt_reconstruct(msfec_rt_handle h, const double* w, int w_loc, double* sigma,
                                  double* u, int out_loc) {
  if (!h || !w) return rt_fail(MSFEC_ERR_INVALID, "null argument");
  if (!sigma && !u) return rt_fail(MSFEC_ERR_INVALID, "no output requested");
  if (w_loc != MSFEC_HOST && w_loc != MSFEC_DEVICE) return rt_fail(MSFEC_ERR_INVALID, "bad location");
  if (out_loc != MSFEC_HOST && out_loc != MSFEC_DEVICE)
    return rt_fail(MSFEC_ERR_INVALID, "bad location");
  if (!h->computed) return rt_fail(MSFEC_ERR_STATE, "compute not called");
  RT_HIP(hipStreamSynchronize(h->stream));
  msfec_status st = rt_in(h->stream, h->d_w, w, (size_t)h->n_cells * 5, w_loc);
  if (st != MSFEC_OK) return st;
  const size_t ns = (size_t)h->n_cells * h->ne, nu = (size_t)h->n_cells * h->m * h->m;
  double* ds = nullptr;
  double* du = nullptr;
  // device outputs are written in place; host outputs go through staging
  // buffers freed before returning
  if (sigma) {
    if (out_loc == MSFEC_DEVICE) ds = sigma;
    else RT_HIP(hipMalloc((void**)&ds, ns * sizeof(double)));
  }
  if (u) {
    if (out_loc == MSFEC_DEVICE) {
      du = u;
    } else if (hipMalloc((void**)&du, nu * sizeof(double)) != hipSuccess) {
      if (ds && out_loc != MSFEC_DEVICE) (void)hipFree(ds);
      return rt_fail(MSFEC_ERR_HIP, "hipMalloc of the reconstruction staging buffer failed");
    }
  }
  hipLaunchKernelGGL(rt_reconstruct_kernel, dim3(h->n_cells), dim3(kNThreads), 0, h->stream,
                     h->d_sigma, h->d_w, h->ne, h->m * h->m, ds, du);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
  if (e == hipSuccess && out_loc == MSFEC_HOST) {
    if (sigma) e = hipMemcpy(sigma, ds, ns * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess && u) e = hipMemcpy(u, du, nu * sizeof(double), hipMemcpyDeviceToHost);
  }
  if (out_loc == MSFEC_HOST) {
    if (ds) (void)hipFree(ds);
    if (du) (void)hipFree(du);
  }
  if (e != hipSuccess) return rt_fail(MSFEC_ERR_HIP, std::string("reconstruct: ") + hipGetErrorString(e));
  return MSFEC_OK;
}

msfec_status msfec_rt_get_cell_status(msfec_rt_handle h, int* status) {
  if (!h || !status) return rt_fail(MSFEC_ERR_INVALID, "null argument");
  if (!h->computed) return rt_fail(MSFEC_ERR_STATE, "compute not called");
  RT_HIP(hipStreamSynchronize(h->stream));
  RT_HIP(hipMemcpy(status, h->d_status, (size_t)h->n_cells * sizeof(int), hipMemcpyDeviceToHost));
  return MSFEC_OK;
}

msfec_status msfec_rt_kernel_times(msfec_rt_handle h, float* ms, int max, int* n) {
  if (!h || !ms) return rt_fail(MSFEC_ERR_INVALID, "null argument");
  if (h->ev_used == 0) return rt_fail(MSFEC_ERR_STATE, "no timed compute (set MSFEC_FLAG_TIMING)");
  double acc[3] = {0.0, 0.0, 0.0};
  for (size_t c = 0; c < h->ev_used; ++c) {
    RT_HIP(hipEventSynchronize(h->ev_sets[c][3]));
    for (int i = 0; i < 3; ++i) {
      float t = 0.f;
      RT_HIP(hipEventElapsedTime(&t, h->ev_sets[c][i], h->ev_sets[c][i + 1]));
      acc[i] += t;
    }
  }
  int k = 0;
  for (; k < 3 && k < max; ++k) ms[k] = (float)(acc[k] / (double)h->ev_used);
  if (n) *n = k;
  h->ev_used = 0;
  return MSFEC_OK;
}

msfec_status msfec_rt_destroy(msfec_rt_handle h) {
  if (!h) return MSFEC_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  rt_free(h);
  delete h;
  return MSFEC_OK;
}

}  // extern "C"
