// Ned_RT: the three-dimensional lowest-order Nedelec x Raviart-Thomas
// multiscale basis on MI355X (gfx950), behind the msfec_nrt_* entry points of
// include/msfec.h.
//
// Reference: source/Ned_RT/ned_rt_basis.cc (NedRTBasis).  Per coarse box K,
// on the 2^l x 2^l x 2^l fine grid, unknowns sigma on the fine edges and u on
// the fine faces:
//   (A^-1 sigma, tau) - (u, curl tau) = 0
//   (v, curl sigma) + (B div u, div v) + (R u, v) = (v, curl N_n)   (n < 12)
// (ned_rt_basis.cc:488-511) with the tangential trace of the coarse Nedelec
// function N_n (edge bases n < 12) or the normal trace of the coarse
// Raviart-Thomas function of face n - 12 (face bases) on the boundary
// (:225-359), one solve per basis (:1280-1352), then the 18 x 18 element
// matrix of :850-948.
//
// One workgroup per coarse cell does all of it (nrt_basis_kernel):
//   local matrices  the fine-cell blocks (edge mass with A^-1 at the 8 Gauss
//                   points, face div-div + reaction, the coefficient-free
//                   curl coupling) and the curl loads, in LDS;
//   assembly        the interior system in the symmetric form
//                   K' = [Me, -C^T; -C, -Du] (face rows negated), edge dofs
//                   first, held in registers as 16 x 16 FP64 matrix-core tiles
//                   (lower triangle), with the 18 right-hand sides (loads
//                   minus the boundary lift) appended as two more tile rows;
//   factorisation   L D L^T without pivoting, 4 pivots per step: the edge
//                   block Me is symmetric positive definite and its Schur
//                   complement -(Du + C Me^-1 C^T) negative definite, so every
//                   pivot has a known sign (a pivot of the wrong sign flags
//                   the cell).  Each step factors the 4 x 4 pivot block in
//                   every thread, writes the panel V = A21 L1^-T and
//                   V D1^-1 to LDS and updates all trailing tiles with one
//                   FP64 MFMA each; the right-hand-side rows ride along, so
//                   the forward elimination is free;
//   backward sweep  per tile column, right to left: the owners of the tiles
//                   below the diagonal form L^T x products on the matrix
//                   cores (their accumulator registers are the transposed A
//                   operand), the diagonal tile's owner solves 16 rows;
//   outputs         the full fine vectors of the 18 bases, then the element
//                   matrix by the reference's block rule from the local
//                   matrices (edge bases through their edge part, face bases
//                   through their face part) and the load (v, f) paired with
//                   the face part of the face bases.
// Tiles are spread over TN / 2 waves, wave w owning tile rows w and
// TN - 1 - w, which keeps the trailing work of every step balanced.
// Dense storage limits the interior system to 256 unknowns: l = 1 (18) and
// l = 2 (252).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <string>
#include <type_traits>
#include <vector>

#include "msfec.h"

namespace msfec_detail {
extern thread_local std::string g_last_error;
}

namespace {

constexpr int kNB = 18;   // 12 edge bases + 6 face bases
constexpr int kMaxL = 2;  // dense interior systems up to 256 unknowns

typedef double nrt_d4 __attribute__((ext_vector_type(4)));

template <int L>
struct Nd {
  static constexpr int m = 1 << L;
  static constexpr int nfc = m * m * m;                    // fine cells
  static constexpr int ns = 3 * m * (m - 1) * (m - 1);     // interior edges
  static constexpr int nu = 3 * (m - 1) * m * m;           // interior faces
  static constexpr int n = ns + nu;
  static constexpr int TN = ((n + 31) / 32) * 2;           // tile rows of the matrix
  static constexpr int NP = 16 * TN;                       // padded size
  static constexpr int NW = TN / 2;                        // waves
  static constexpr int NT = 64 * NW;                       // threads
  static constexpr int NR = NP + 32;                       // + two right-hand-side tile rows
  static constexpr int NS = TN + 5;                        // tile slots per wave
  static constexpr int XS = kNB;                           // row stride of x in LDS
  static constexpr int CH = nfc < 8 ? nfc : 8;             // fine cells per element chunk
  // LDS, in doubles.  Region A lives through the factorisation (pivot
  // columns, V, V D^-1), later the backward partial sums and the element
  // chunks; region B holds the coefficients, curl loads and dof table, later
  // x and the element chunk's basis values; region C the local matrices.
  static constexpr int cmax(int a, int b) { return a > b ? a : b; }
  static constexpr int oStage = 0, oV = NR * 4, oVD = 2 * NR * 4;
  static constexpr int oPart = 0;                          // NW x 16 x 18
  static constexpr int oLs = NW * 16 * 18;                 // 16 x 16
  static constexpr int oY = 0;                             // CH x 18 x 18
  static constexpr int szA = cmax(cmax(3 * NR * 4, NW * 16 * 18 + 256), CH * 324);
  static constexpr int oB = szA;
  static constexpr int oFc = oB;                           // 12 x nu
  static constexpr int oIc = oFc + 12 * nu;                // n x 4 ints
  static constexpr int oCf = oIc + 2 * n;                  // nfc x 8 x 8
  static constexpr int oRb = oCf;                          // 18 x n (after the local matrices)
  static constexpr int oX = oB;                            // NP x XS
  static constexpr int oXL = oB + NP * XS;                 // CH x 180
  static constexpr int szB = cmax(12 * nu + 2 * n + cmax(nfc * 64, kNB * n), NP * XS + CH * 180);
  static constexpr int oC = oB + szB;
  static constexpr int oMl = oC;                           // nfc x 78
  static constexpr int oDl = oMl + nfc * 78;               // nfc x 21
  static constexpr int oCc = oDl + nfc * 21;               // 6 x 12
  static constexpr int oFl = oCc + 72;                     // nfc x 6
  static constexpr int oPi = oFl + nfc * 6;                // NP / 4 pivot-block inverses, 4 x 4
  static constexpr int oSc = oPi + NP * 4;                 // n row scales
  static constexpr int oCd = oSc + n;                      // ints: nfc x 18 local dofs, then N f2i
  static constexpr int N = 3 * m * (m + 1) * (m + 1) + 3 * m * m * (m + 1);
  static constexpr int total = oCd + (nfc * 18 + N + 1) / 2;
  // phase-0 tables in region A: W[78][8], PP[21][8], DD[21], comp[78] (ints)
  static constexpr int oW = 0, oPP = 624, oDD = 624 + 168, oCmp = 624 + 168 + 21;
  static_assert(oCmp + 39 <= szA, "phase-0 tables exceed region A");
  static_assert(total * 8 <= 160 * 1024, "LDS budget");
  static_assert(oCc + 72 <= 0x8000, "local-matrix offsets must fit the 15 bits of the assembly map");
};

// deal.II's line order on the unit cube: axis of line L and the values of
// the two other coordinates (lower axis first)
__constant__ int kLineAx[12] = {1, 1, 0, 0, 1, 1, 0, 0, 2, 2, 2, 2};
__constant__ int kLineF1[12] = {0, 1, 0, 1, 0, 1, 0, 1, 0, 1, 0, 1};
__constant__ int kLineF2[12] = {0, 0, 0, 0, 1, 1, 1, 1, 0, 0, 1, 1};

__device__ __forceinline__ double gauss_pt(int b) {
  return b ? 0.78867513459481288225 : 0.21132486540518711775;
}
__device__ __forceinline__ double lag(int s, double t) { return s ? t : 1.0 - t; }
__device__ __forceinline__ double dlag(int s) { return s ? 1.0 : -1.0; }
__device__ __forceinline__ int oth1(int d) { return d == 0 ? 1 : 0; }
__device__ __forceinline__ int oth2(int d) { return d == 2 ? 1 : 2; }
// (e_a x e_b)_c
__device__ __forceinline__ double levi(int a, int b, int c) {
  if (a == b || b == c || a == c) return 0.0;
  return ((b - a + 3) % 3 == 1) ? 1.0 : -1.0;
}
// index of the (a, b) entry of a symmetric 3 x 3 stored (xx, xy, xz, yy, yz, zz)
__device__ __forceinline__ int sym6(int a, int b) {
  const int lo = min(a, b), hi = max(a, b);
  return lo == 0 ? hi : (lo == 1 ? 2 + hi : 5);
}
// packed lower triangle (i >= j)
__device__ __forceinline__ int tri(int i, int j) {
  const int hi = max(i, j), lo = min(i, j);
  return hi * (hi + 1) / 2 + lo;
}
__device__ __forceinline__ void tri_ij(int e, int& i, int& j) {
  i = 0;
  while ((i + 1) * (i + 2) / 2 <= e) ++i;
  j = e - i * (i + 1) / 2;
}

// a 3-vector of scalars: selecting a component by a run-time index stays a
// register select (an array indexed at run time would live in scratch)
struct V3 {
  double x, y, z;
};
__device__ __forceinline__ double sel3(V3 v, int i) {
  const double xy = i == 1 ? v.y : v.x;
  return i == 2 ? v.z : xy;
}

// Nedelec function of line L on a box of sizes h at reference point X (the
// covariant map of l_f1(X_o1) l_f2(X_o2) e_ax): its axis value
__device__ __forceinline__ double ned_val(int L, V3 X, V3 h) {
  const int ax = kLineAx[L];
  return lag(kLineF1[L], sel3(X, oth1(ax))) * lag(kLineF2[L], sel3(X, oth2(ax))) / sel3(h, ax);
}
// component c of its curl, grad(g / h_ax) x e_ax
__device__ __forceinline__ double ned_curl(int L, V3 X, V3 h, int c) {
  const int ax = kLineAx[L], o1 = oth1(ax), o2 = oth2(ax), f1 = kLineF1[L], f2 = kLineF2[L];
  const double ha = sel3(h, ax);
  const double g1 = dlag(f1) * lag(f2, sel3(X, o2)) / (ha * sel3(h, o1));
  const double g2 = lag(f1, sel3(X, o1)) * dlag(f2) / (ha * sel3(h, o2));
  return g1 * levi(o1, ax, c) + g2 * levi(o2, ax, c);
}
// Raviart-Thomas function of face F (axis F / 2, side F % 2): its axis
// value (contravariant map of l_s(X_ax) e_ax) and divergence
__device__ __forceinline__ double rt_val(int F, V3 X, V3 h) {
  const int ax = F >> 1;
  return lag(F & 1, sel3(X, ax)) / (sel3(h, oth1(ax)) * sel3(h, oth2(ax)));
}
__device__ __forceinline__ double rt_div(int F, V3 h) {
  return dlag(F & 1) / (h.x * h.y * h.z);
}
__device__ __forceinline__ V3 gauss_xyz(int q) {
  return V3{gauss_pt(q & 1), gauss_pt((q >> 1) & 1), gauss_pt(q >> 2)};
}

// f(integral_constant<int, s>) for s = B .. E - 1
template <int B, int E>
struct SlotLoop {
  template <typename F>
  __device__ __forceinline__ static void run(F&& f) {
    if constexpr (B < E) {
      f(std::integral_constant<int, B>{});
      SlotLoop<B + 1, E>::run(f);
    }
  }
};

// Opaque copies of the lane and wave index: the per-slot LDS addresses of a
// factorisation step are recomputed from them (a few integer operations)
// instead of being hoisted out of the loop as dozens of live registers.
__device__ __forceinline__ int opaque_v(int x) {
  asm volatile("" : "+v"(x));
  return x;
}
__device__ __forceinline__ int opaque_s(int x) {
  asm volatile("" : "+s"(x));
  return x;
}

__device__ __forceinline__ void wave_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// deal.II's line order as compile-time tables (for loops unrolled over the
// bases)
constexpr int kLAx[12] = {1, 1, 0, 0, 1, 1, 0, 0, 2, 2, 2, 2};
constexpr int kLF1[12] = {0, 1, 0, 1, 0, 1, 0, 1, 0, 1, 0, 1};
constexpr int kLF2[12] = {0, 0, 0, 0, 1, 1, 1, 1, 0, 0, 1, 1};

// Type (0..2 edges along x/y/z, 3..5 faces normal to x/y/z) and integer
// position of the full dof f: the start node of an edge, the level along
// its normal and the cell indices along the other axes of a face.
template <int M>
__device__ __forceinline__ void dof_pos(int f, int& t, int& p0, int& p1, int& p2) {
  constexpr int nx = M * (M + 1) * (M + 1), ne = 3 * nx, nf1 = M * M * (M + 1);
  int q, e0, e1;
  if (f < ne) {
    t = f / nx;
    q = f - t * nx;
    e0 = t == 0 ? M : M + 1;
    e1 = t == 1 ? M : M + 1;
  } else {
    const int tf = (f - ne) / nf1;
    q = (f - ne) - tf * nf1;
    e0 = tf == 0 ? M + 1 : M;
    e1 = tf == 1 ? M + 1 : M;
    t = 3 + tf;
  }
  p0 = q % e0;
  q /= e0;
  p1 = q % e1;
  p2 = q / e1;
}

// The fixed boundary value of basis r on the boundary dof f (full
// numbering): the line integral of the coarse Nedelec function r along a
// boundary edge, (1/m) l_f1(p_o1 / m) l_f2(p_o2 / m) for an edge along the
// line's axis, and the flux 1/m^2 of the coarse Raviart-Thomas function of
// face r - 12 through a fine face of that face.  Only called for dofs on the
// coarse boundary.
template <int M>
__device__ __forceinline__ double bnd_value(int f, int r) {
  int t, p0, p1, p2;
  dof_pos<M>(f, t, p0, p1, p2);
  if (t < 3) {
    if (r >= 12 || kLineAx[r] != t) return 0.0;
    const int po1 = t == 0 ? p1 : p0, po2 = t == 2 ? p1 : p2;
    return lag(kLineF1[r], (double)po1 / M) * lag(kLineF2[r], (double)po2 / M) / M;
  }
  const int d = t - 3, pd = d == 0 ? p0 : (d == 1 ? p1 : p2);
  return r == 12 + 2 * d + (pd == M ? 1 : 0) ? 1.0 / (M * M) : 0.0;
}

struct NrtArgs {
  const double* verts;  // n_cells x 8 x 3
  const double* ainv;   // n_cells x nfc x 8 x 6
  const double* bq;     // n_cells x nfc x 8
  const double* rq;     // n_cells x nfc x 8, or null (no reaction)
  const double* fq;     // n_cells x nfc x 8 x 3, or null (no load)
  const int* icell;     // n x 4: fine cell * 32 + local dof of each cell holding the interior dof, -1
  const int* cdof;      // nfc x 18: full dof of every local dof
  const int* f2i;       // N: interior index of a full dof, -1 on the boundary
  const unsigned short* amap;  // assembly map of the lower tiles (see nrt_tile)
  const unsigned char* tnz;    // the lower tiles with any entry
  double* X;            // n_cells x 18 x N
  double* M;            // n_cells x 18 x 18
  double* b;            // n_cells x 18
  int* status;          // n_cells
  int N;
};

// entry (a, b) of the local matrix of fine cell c: K' (face rows negated,
// symmetric) or, with orig, the assembled local matrix of the reference
// (row = test); Me, Du packed lower, C[F][L] = (psi_F, curl phi_L)
template <bool Orig>
__device__ __forceinline__ double kloc(const double* Ml, const double* Dl, const double* Cc, int c,
                                       int a, int b) {
  // one address and one sign, no branches: Me, Du packed lower, C[F][L]
  const bool ea = a < 12, eb = b < 12;
  const int hi = max(a, b), lo = min(a, b);
  const double* base = (ea && eb) ? Ml + c * 78 : ((!ea && !eb) ? Dl + c * 21 : Cc);
  const int off = (ea && eb) ? hi * (hi + 1) / 2 + lo
                             : ((!ea && !eb) ? (hi - 12) * (hi - 11) / 2 + (lo - 12) : (hi - 12) * 12 + lo);
  const double v = base[off];
  // K': every entry but the edge block negated; the assembled matrix: only
  // the (edge test, face trial) block, -C^T
  const bool neg = Orig ? (ea && !eb) : !(ea && eb);
  return neg ? -v : v;
}

// The entries of tile (P, Q) held by this lane (column 16 Q + ll, rows
// 16 P + 4 gg + lh): K'' for P < TN (identity on the padding), else the
// right-hand sides r = 16 (P - TN) + 4 gg + lh from rb.  An entry of K' is
// the sum of at most 4 local-matrix entries (the fine cells the two dofs
// share); amap lists them per tile, register and lane as 16-bit LDS offsets
// with the sign in the top bit (0xffff: none), and tnz flags the tiles that
// hold any.
template <int L>
__device__ __forceinline__ nrt_d4 nrt_tile(int P, int Q, int lane, const double* sm,
                                        const double* rb, const double* sc,
                                        const unsigned short* amap, const unsigned char* tnz) {
  typedef Nd<L> D;
  const int ll = lane & 15, lh = lane >> 4;
  nrt_d4 t = {0.0, 0.0, 0.0, 0.0};
  const int J = 16 * Q + ll;
  if (P < D::TN) {
    const int tix = P * (P + 1) / 2 + Q;
    if (J >= D::n) {
      for (int gg = 0; gg < 4; ++gg) t[gg] = (16 * P + 4 * gg + lh == J) ? 1.0 : 0.0;
      return t;
    }
    if (!tnz[tix]) return t;
    const double sj = sc[J];
    for (int gg = 0; gg < 4; ++gg) {
      const int I = 16 * P + 4 * gg + lh;
      const uint2 e = reinterpret_cast<const uint2*>(amap)[((size_t)tix * 4 + gg) * 64 + lane];
      const unsigned hs[4] = {e.x & 0xffffu, e.x >> 16, e.y & 0xffffu, e.y >> 16};
      double v = 0.0;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const unsigned h = hs[k];
        if (h != 0xffffu) {
          const double x = sm[h & 0x7fffu];
          v += (h & 0x8000u) ? -x : x;
        }
      }
      t[gg] = I < D::n ? v * sc[I] * sj : 0.0;
    }
    return t;
  }
  if (J >= D::n) return t;
  for (int gg = 0; gg < 4; ++gg) {
    const int r = 16 * (P - D::TN) + 4 * gg + lh;
    t[gg] = r < kNB ? rb[r * D::n + J] : 0.0;
  }
  return t;
}

template <int L>
__global__ __launch_bounds__(Nd<L>::NT) void nrt_basis_kernel(NrtArgs g) {
  typedef Nd<L> D;
  extern __shared__ double sm[];
  __shared__ int s_flag;
  const int cell = blockIdx.x, tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6, ll = lane & 15, lh = lane >> 4;
  const int N = g.N;

  // ---- coarse and fine box sizes
  V3 Hc, h;
  {
    const double* v = g.verts + (size_t)cell * 24;
    Hc = V3{v[3] - v[0], v[7] - v[1], v[14] - v[2]};
    h = V3{Hc.x / D::m, Hc.y / D::m, Hc.z / D::m};
  }
  const double vol = h.x * h.y * h.z;
  if (tid == 0) s_flag = MSFEC_CELL_OK;

  // ---- coefficients to LDS: per fine cell and point, A^-1 (6), B, R; the
  // dof tables; the point products of the shape functions
  double* cf = sm + D::oCf;
  for (int idx = tid; idx < D::nfc * 8; idx += D::NT) {
    const size_t gi = (size_t)cell * D::nfc * 8 + idx;
    double* o = cf + idx * 8;
#pragma unroll
    for (int k = 0; k < 6; ++k) o[k] = g.ainv[gi * 6 + k];
    o[6] = g.bq[gi];
    o[7] = g.rq ? g.rq[gi] : 0.0;
  }
  int* ic = reinterpret_cast<int*>(sm + D::oIc);
  for (int idx = tid; idx < D::n * 4; idx += D::NT) ic[idx] = g.icell[idx];
  int* cdl = reinterpret_cast<int*>(sm + D::oCd);
  int* f2l = cdl + D::nfc * 18;
  for (int idx = tid; idx < D::nfc * 18; idx += D::NT) cdl[idx] = g.cdof[idx];
  for (int idx = tid; idx < D::N; idx += D::NT) f2l[idx] = g.f2i[idx];
  double* Wt = sm + D::oW;
  double* PP = sm + D::oPP;
  double* DD = sm + D::oDD;
  int* cmp = reinterpret_cast<int*>(sm + D::oCmp);
  const double w8 = vol * 0.125;
  for (int idx = tid; idx < 78 * 8 + 21 * 8 + 21 + 78; idx += D::NT) {
    if (idx < 78 * 8) {
      const int e = idx >> 3, q = idx & 7;
      int i, j;
      tri_ij(e, i, j);
      const V3 X = gauss_xyz(q);
      Wt[idx] = w8 * ned_val(i, X, h) * ned_val(j, X, h);
    } else if (idx < 78 * 8 + 21 * 8) {
      const int e = (idx - 78 * 8) >> 3, q = (idx - 78 * 8) & 7;
      int i, j;
      tri_ij(e, i, j);
      const V3 X = gauss_xyz(q);
      PP[e * 8 + q] = (i >> 1) == (j >> 1) ? w8 * rt_val(i, X, h) * rt_val(j, X, h) : 0.0;
    } else if (idx < 78 * 8 + 21 * 8 + 21) {
      int i, j;
      tri_ij(idx - 78 * 8 - 21 * 8, i, j);
      DD[idx - 78 * 8 - 21 * 8] = w8 * rt_div(i, h) * rt_div(j, h);
    } else {
      const int e = idx - 78 * 8 - 21 * 8 - 21;
      int i, j;
      tri_ij(e, i, j);
      cmp[e] = sym6(kLineAx[i], kLineAx[j]);
    }
  }
  __syncthreads();

  // ---- local matrices, curl coupling, loads
  double* Ml = sm + D::oMl;
  double* Dl = sm + D::oDl;
  double* Cc = sm + D::oCc;
  double* Fl = sm + D::oFl;
  double* Fc = sm + D::oFc;
  for (int idx = tid; idx < D::nfc * 99; idx += D::NT) {
    const int c = idx / 99, e = idx % 99;
    const double* co = cf + c * 64;
    double acc = 0.0;
    if (e < 78) {
      const int ci = cmp[e];
#pragma unroll
      for (int q = 0; q < 8; ++q) acc += Wt[e * 8 + q] * co[q * 8 + ci];
      Ml[c * 78 + e] = acc;
    } else {
      const int e2 = e - 78;
      const double dd = DD[e2];
#pragma unroll
      for (int q = 0; q < 8; ++q) acc += co[q * 8 + 6] * dd + co[q * 8 + 7] * PP[e2 * 8 + q];
      Dl[c * 21 + e2] = acc;
    }
  }
  for (int idx = tid; idx < 72 + D::nfc * 6 + 12 * D::nu; idx += D::NT) {
    if (idx < 72) {
      const int F = idx / 12, Ln = idx % 12, ax = F >> 1;
      double acc = 0.0;
      for (int q = 0; q < 8; ++q) {
        const V3 X = gauss_xyz(q);
        acc += rt_val(F, X, h) * ned_curl(Ln, X, h, ax);
      }
      Cc[idx] = acc * w8;
    } else if (idx < 72 + D::nfc * 6) {
      const int c = (idx - 72) / 6, F = (idx - 72) % 6, ax = F >> 1;
      double acc = 0.0;
      if (g.fq) {
        const double* f = g.fq + ((size_t)cell * D::nfc + c) * 24;
        for (int q = 0; q < 8; ++q) {
          const V3 X = gauss_xyz(q);
          acc += rt_val(F, X, h) * f[q * 3 + ax];
        }
      }
      Fl[c * 6 + F] = acc * w8;
    } else {
      // (psi_I, curl N_n) over the two fine cells of interior face I: along
      // the face normal d both factors depend on the d-coordinate only, so
      // QGauss(2) reduces to its 2 points along d (4 points each)
      const int t = idx - 72 - D::nfc * 6, nb = t / D::nu, J = t % D::nu;
      double acc = 0.0;
      for (int k = 0; k < 2; ++k) {
        const int e = ic[(D::ns + J) * 4 + k];
        const int c = e >> 5, F = (e & 31) - 12, ax = F >> 1;
        const int c0 = c % D::m, c1 = (c / D::m) % D::m, c2 = c / (D::m * D::m);
        for (int qd = 0; qd < 2; ++qd) {
          const double gd = gauss_pt(qd);
          const V3 X{ax == 0 ? gd : 0.5, ax == 1 ? gd : 0.5, ax == 2 ? gd : 0.5};
          const V3 Xc{(c0 + X.x) / D::m, (c1 + X.y) / D::m, (c2 + X.z) / D::m};
          acc += rt_val(F, X, h) * ned_curl(nb, Xc, Hc, ax);
        }
      }
      Fc[nb * D::nu + J] = acc * 4.0 * w8;
    }
  }
  __syncthreads();

  // ---- right-hand sides b' = S F - K'_IB x_B of the 18 bases, rb[r][J], one
  // thread per interior dof J: the curl loads (face rows negated in K'), then
  // the lift of the boundary values over the fine cells of J (a boundary edge
  // carries the 4 edge bases along its axis, a boundary face one face basis)
  // symmetric equilibration: K'' = S K' S with S = |diag K'|^-1/2, so that
  // every pivot is of order one (the solution is x = S y)
  double* scl = sm + D::oSc;
  for (int I = tid; I < D::n; I += D::NT) {
    double dg = 0.0;
    for (int kc = 0; kc < 4; ++kc) {
      const int e = ic[I * 4 + kc];
      if (e >= 0) dg += kloc<false>(Ml, Dl, Cc, e >> 5, e & 31, e & 31);
    }
    const double a = fabs(dg);
    scl[I] = (a > 0.0 && isfinite(a)) ? 1.0 / sqrt(a) : 1.0;
  }
  __syncthreads();
  double* rb = sm + D::oRb;
  // four lanes per dof J, one per fine cell of J, summed by lane shuffles
  for (int base = 0; base < 4 * D::n; base += D::NT) {
    const int idx = base + tid, J = idx >> 2, kc = idx & 3;
    double y[kNB];
#pragma unroll
    for (int r = 0; r < kNB; ++r) y[r] = 0.0;
    const int e = J < D::n ? ic[J * 4 + kc] : -1;
    if (e >= 0) {
      const int c = e >> 5, lj = e & 31;
      for (int lb = 0; lb < 18; ++lb) {
        const int full = cdl[c * 18 + lb];
        if (f2l[full] >= 0) continue;
        const double kv = kloc<false>(Ml, Dl, Cc, c, lj, lb);
        int t, p0, p1, p2;
        dof_pos<D::m>(full, t, p0, p1, p2);
        if (t < 3) {
          const double x1 = (double)(t == 0 ? p1 : p0) / D::m;
          const double x2 = (double)(t == 2 ? p1 : p2) / D::m;
#pragma unroll
          for (int r = 0; r < 12; ++r)
            if (kLAx[r] == t) y[r] -= kv * (lag(kLF1[r], x1) * lag(kLF2[r], x2) / D::m);
        } else {
          const int d = t - 3, pd = d == 0 ? p0 : (d == 1 ? p1 : p2);
          const int F = 2 * d + (pd == D::m ? 1 : 0);
#pragma unroll
          for (int r = 0; r < 6; ++r)
            if (r == F) y[12 + r] -= kv * (1.0 / (D::m * D::m));
        }
      }
    }
#pragma unroll
    for (int r = 0; r < kNB; ++r) {
      y[r] += __shfl_xor(y[r], 1);
      y[r] += __shfl_xor(y[r], 2);
    }
    if (J < D::n && kc == 0) {
      const double sj = scl[J];
#pragma unroll
      for (int r = 0; r < kNB; ++r) {
        const double f0 = (r < 12 && J >= D::ns) ? -Fc[r * D::nu + (J - D::ns)] : 0.0;
        rb[r * D::n + J] = (f0 + y[r]) * sj;
      }
    }
  }
  __syncthreads();

  // ---- assembly into the tile registers
  // slot s of wave w: for s <= TN, tile (TN-1-w, s) while s <= TN-1-w, else
  // tile (w, TN-s) -- the tile column is a compile-time constant per branch;
  // slots TN+1..TN+4 are the right-hand-side tiles (TN, w), (TN, TN-1-w),
  // (TN+1, w), (TN+1, TN-1-w)
  const int Pb = D::TN - 1 - w;
  auto slot_pq = [&](int s, int& P, int& Q) {
    if (s <= D::TN) {
      const bool lo = s <= Pb;
      P = lo ? Pb : w;
      Q = lo ? s : D::TN - s;
    } else {
      const int r = s - D::TN - 1;
      P = D::TN + (r >> 1);
      Q = (r & 1) ? Pb : w;
    }
  };
  // every loop over the tile slots below runs through SlotLoop, so the slot
  // index is a compile-time constant and the tiles stay in registers
  nrt_d4 T[D::NS];
  SlotLoop<0, D::NS>::run([&](auto sc) {
    constexpr int s = decltype(sc)::value;
    int P, Q;
    slot_pq(s, P, Q);
    T[s] = nrt_tile<L>(P, Q, lane, sm, rb, scl, g.amap, g.tnz);
  });
  // first pivot columns (0..3) to LDS
  double* stg = sm + D::oStage;
  double* vb = sm + D::oV;
  double* vdb = sm + D::oVD;
  double* pinv = sm + D::oPi;
  SlotLoop<0, D::NS>::run([&](auto sc) {
    constexpr int s = decltype(sc)::value;
    int P, Q;
    slot_pq(s, P, Q);
    if (Q == 0 && ll < 4)
#pragma unroll
      for (int gg = 0; gg < 4; ++gg) stg[(16 * P + 4 * gg + lh) * 4 + ll] = T[s][gg];
  });

  // ---- block L D L^T, 4 pivots per step, right-hand-side rows alongside.
  // Phase A (every thread): the pivot block P (the current Schur complement,
  // staged in LDS) = L1 D1 L1^T, its inverse to LDS for the backward sweep,
  // and, for row i below it, V_i = A_i L1^-T and V_i D1^-1 to LDS.  Phase B
  // (the tile owners): T -= (V D1^-1) V^T, one MFMA per tile, then the next
  // pivot columns to LDS.  V vanishes on the pivot rows, so a pivot column
  // keeps its Schur-complement values A_i: the backward sweep works with
  // them and the stored inverses (block back substitution).
  for (int k = 0; k < D::NP; k += 4) {
    const int Qk = k >> 4;
    const int ll = opaque_v(tid) & 15, lh = (opaque_v(tid) >> 4) & 3;
    const int w = opaque_s(__builtin_amdgcn_readfirstlane(tid >> 6)), Pb = D::TN - 1 - w;
    auto slot_pq = [&](int s, int& P, int& Q) {
      if (s <= D::TN) {
        const bool lo = s <= Pb;
        P = lo ? Pb : w;
        Q = lo ? s : D::TN - s;
      } else {
        const int r = s - D::TN - 1;
        P = D::TN + (r >> 1);
        Q = (r & 1) ? Pb : w;
      }
    };
    __syncthreads();
    {
      double p[4][4];
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c <= r; ++c) p[r][c] = stg[(k + r) * 4 + c];
      double d[4], di[4], l1[4][4], u[4][4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        double dc = p[c][c];
#pragma unroll
        for (int t = 0; t < c; ++t) dc -= l1[c][t] * u[c][t];
        d[c] = dc;
        di[c] = 1.0 / dc;
#pragma unroll
        for (int r = c + 1; r < 4; ++r) {
          double v = p[r][c];
#pragma unroll
          for (int t = 0; t < c; ++t) v -= l1[r][t] * u[c][t];
          u[r][c] = v;
          l1[r][c] = v * di[c];
        }
      }
      // rows of L1^-1 (unit lower)
      double li[4][4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
#pragma unroll
        for (int c = 0; c < 4; ++c) li[r][c] = (r == c) ? 1.0 : 0.0;
#pragma unroll
        for (int c = 0; c < r; ++c) {
          double v = 0.0;
#pragma unroll
          for (int t = c; t < r; ++t) v -= l1[r][t] * li[t][c];
          li[r][c] = v;
        }
      }
      if (tid < 16) {
        // entry (r, c) of P^-1 = L1^-T D1^-1 L1^-1; the sign check of the pivots
        double pv = 0.0;
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            double v = 0.0;
#pragma unroll
            for (int t = 0; t < 4; ++t) v += li[t][r] * di[t] * li[t][c];
            pv = tid == r * 4 + c ? v : pv;
          }
        pinv[(k >> 2) * 16 + tid] = pv;
        if (tid < 4) {
          const double dc = tid == 0 ? d[0] : (tid == 1 ? d[1] : (tid == 2 ? d[2] : d[3]));
          const int idx = k + tid;
          const double sg = (idx < D::ns || idx >= D::n) ? 1.0 : -1.0;
          if (!(sg * dc > 0.0) || !isfinite(dc)) s_flag = MSFEC_CELL_NOT_SPD;
        }
      }
      for (int row = tid; row < D::NR; row += D::NT) {
        const bool act = row > k + 3;
        double a[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) a[c] = act ? stg[row * 4 + c] : 0.0;
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
          double v = 0.0;
#pragma unroll
          for (int c = 0; c <= kk; ++c) v += a[c] * li[kk][c];
          vb[row * 4 + kk] = v;
          vdb[row * 4 + kk] = act ? v * di[kk] : 0.0;
        }
      }
    }
    __syncthreads();
    {
      // operands, all read up front: V D^-1 of the wave's 4 tile rows, V of
      // every active tile column; updates by tile column, left to right
      const double aPb = vdb[(16 * Pb + ll) * 4 + lh], aPa = vdb[(16 * w + ll) * 4 + lh];
      const double aR0 = vdb[(16 * D::TN + ll) * 4 + lh];
      const double aR1 = vdb[(16 * (D::TN + 1) + ll) * 4 + lh];
      double bq[D::TN];
#pragma unroll
      for (int q = 0; q < D::TN; ++q) bq[q] = q >= Qk ? vb[(16 * q + ll) * 4 + lh] : 0.0;
      const double bR0 = w >= Qk ? vb[(16 * w + ll) * 4 + lh] : 0.0;
      const double bR1 = Pb >= Qk ? vb[(16 * Pb + ll) * 4 + lh] : 0.0;
      SlotLoop<0, D::TN>::run([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if (q >= Qk) {
          // tile (Pb, q) is slot q, tile (w, q) slot TN - q
          if (q <= Pb) T[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(aPb, bq[q], T[q], 0, 0, 1);
          if (q <= w)
            T[D::TN - q] = __builtin_amdgcn_mfma_f64_16x16x4f64(aPa, bq[q], T[D::TN - q], 0, 0, 1);
          if (q == w) {
            T[D::TN + 1] = __builtin_amdgcn_mfma_f64_16x16x4f64(aR0, bR0, T[D::TN + 1], 0, 0, 1);
            T[D::TN + 3] = __builtin_amdgcn_mfma_f64_16x16x4f64(aR1, bR0, T[D::TN + 3], 0, 0, 1);
          }
          if (q == Pb) {
            T[D::TN + 2] = __builtin_amdgcn_mfma_f64_16x16x4f64(aR0, bR1, T[D::TN + 2], 0, 0, 1);
            T[D::TN + 4] = __builtin_amdgcn_mfma_f64_16x16x4f64(aR1, bR1, T[D::TN + 4], 0, 0, 1);
          }
        }
      });
    }
    // next pivot columns to LDS (read after the barrier at the top)
    const int k2 = k + 4, Q2 = k2 >> 4, G2 = (k2 >> 2) & 3;
    if (k2 < D::NP) {
      SlotLoop<0, D::NS>::run([&](auto sc) {
        constexpr int s = decltype(sc)::value;
        int P, Q;
        slot_pq(s, P, Q);
        if (Q == Q2 && (ll >> 2) == G2)
#pragma unroll
          for (int gg = 0; gg < 4; ++gg) stg[(16 * P + 4 * gg + lh) * 4 + (ll & 3)] = T[s][gg];
      });
    }
  }
  __syncthreads();

  // ---- block back substitution, tile column by tile column from the right:
  // x_b = P_b^-1 (r_b - A_b^T x_below) for the 4 pivot blocks b of a tile
  // column, with r_b the right-hand sides' entries in the pivot columns and
  // A_b the Schur-complement columns below the block.  The owners of the
  // tiles below the diagonal form A^T x on the matrix cores (their
  // accumulator registers are the transposed A operand); the diagonal tile's
  // owner finishes the 4 blocks.
  double* xb = sm + D::oX;
  double* part = sm + D::oPart;
  double* Ls = sm + D::oLs;
  SlotLoop<D::TN + 1, D::NS>::run([&](auto sc) {
    constexpr int s = decltype(sc)::value;
    int P, Q;
    slot_pq(s, P, Q);
#pragma unroll
    for (int gg = 0; gg < 4; ++gg) {
      const int r = 16 * (P - D::TN) + 4 * gg + lh;
      if (r < kNB) xb[(16 * Q + ll) * D::XS + r] = T[s][gg];
    }
  });
  __syncthreads();
  for (int Qb = D::TN - 1; Qb >= 0; --Qb) {
    const int ll = opaque_v(tid) & 15, lh = (opaque_v(tid) >> 4) & 3;
    const int w = opaque_s(__builtin_amdgcn_readfirstlane(tid >> 6)), Pb = D::TN - 1 - w;
    auto slot_pq = [&](int s, int& P, int& Q) {
      if (s <= D::TN) {
        const bool lo = s <= Pb;
        P = lo ? Pb : w;
        Q = lo ? s : D::TN - s;
      } else {
        const int r = s - D::TN - 1;
        P = D::TN + (r >> 1);
        Q = (r & 1) ? Pb : w;
      }
    };
    nrt_d4 acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
    SlotLoop<0, D::TN + 1>::run([&](auto sc) {
      constexpr int s = decltype(sc)::value;
      int P, Q;
      slot_pq(s, P, Q);
      if (Q == Qb && P > Qb) {
#pragma unroll
        for (int sp = 0; sp < 4; ++sp) {
          const int rx = 16 * P + 4 * sp + lh;
          const double b0 = xb[rx * D::XS + ll];
          const double b1 = ll < 2 ? xb[rx * D::XS + 16 + ll] : 0.0;
          acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(T[s][sp], b0, acc0, 0, 0, 0);
          acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(T[s][sp], b1, acc1, 0, 0, 0);
        }
      }
    });
#pragma unroll
    for (int gg = 0; gg < 4; ++gg) {
      part[(w * 16 + 4 * gg + lh) * kNB + ll] = acc0[gg];
      if (ll < 2) part[(w * 16 + 4 * gg + lh) * kNB + 16 + ll] = acc1[gg];
    }
    __syncthreads();
    if (Qb == w || Qb == D::TN - 1 - w) {
      SlotLoop<0, D::TN + 1>::run([&](auto sc) {
        constexpr int s = decltype(sc)::value;
        int P, Q;
        slot_pq(s, P, Q);
        if (P == Qb && Q == Qb)
#pragma unroll
          for (int gg = 0; gg < 4; ++gg) Ls[(4 * gg + lh) * 16 + ll] = T[s][gg];
      });
      wave_fence();
      if (lane < kNB) {
        double x[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
          double v = xb[(16 * Qb + i) * D::XS + lane];
          for (int ww = 0; ww < D::NW; ++ww) v -= part[(ww * 16 + i) * kNB + lane];
          x[i] = v;
        }
#pragma unroll
        for (int b = 3; b >= 0; --b) {
          // r_b - A_b^T x_below (the rows of the later blocks of this tile)
#pragma unroll
          for (int c = 4 * b; c < 4 * b + 4; ++c)
#pragma unroll
            for (int r = 4 * b + 4; r < 16; ++r) x[c] -= Ls[r * 16 + c] * x[r];
          const double* pi = pinv + (4 * Qb + b) * 16;
          const double t0 = x[4 * b], t1 = x[4 * b + 1], t2 = x[4 * b + 2], t3 = x[4 * b + 3];
#pragma unroll
          for (int c = 0; c < 4; ++c)
            x[4 * b + c] = pi[c * 4] * t0 + pi[c * 4 + 1] * t1 + pi[c * 4 + 2] * t2 + pi[c * 4 + 3] * t3;
        }
#pragma unroll
        for (int i = 0; i < 16; ++i) xb[(16 * Qb + i) * D::XS + lane] = x[i];
      }
    }
    __syncthreads();
  }

  // x = S y
  for (int idx = tid; idx < D::n * kNB; idx += D::NT) {
    const int I = idx / kNB, r = idx - I * kNB;
    xb[I * D::XS + r] *= scl[I];
  }
  __syncthreads();

  // ---- the fine vectors of the bases: interior from x, boundary values fixed
  {
    double* Xo = g.X + (size_t)cell * kNB * N;
    for (int idx = tid; idx < kNB * N; idx += D::NT) {
      const int r = idx / N, f = idx - r * N;
      const int I = f2l[f];
      Xo[idx] = I >= 0 ? xb[I * D::XS + r] : bnd_value<D::m>(f, r);
    }
  }

  // ---- element matrix (ned_rt_basis.cc:850-948) and load, by fine-cell chunks
  constexpr int KP = (kNB * kNB + 6 + D::NT - 1) / D::NT;
  double macc[KP];
#pragma unroll
  for (int k = 0; k < KP; ++k) macc[k] = 0.0;
  double* XL = sm + D::oXL;  // [cl][144 edge values: local edge a, edge basis j | 36 face values]
  double* Yb = sm + D::oY;   // [cl][local row a][basis j]
  for (int c0 = 0; c0 < D::nfc; c0 += D::CH) {
    for (int idx = tid; idx < D::CH * 180; idx += D::NT) {
      const int cl = idx / 180, e = idx % 180, c = c0 + cl;
      int a, j;
      if (e < 144) {
        a = e / 12;
        j = e % 12;
      } else {
        a = 12 + (e - 144) / 6;
        j = 12 + (e - 144) % 6;
      }
      const int full = cdl[c * 18 + a];
      const int I = f2l[full];
      XL[idx] = I >= 0 ? xb[I * D::XS + j] : bnd_value<D::m>(full, j);
    }
    __syncthreads();
    // Y[cl][a][j] = (K_loc x_j)[a] over the block of basis j
    for (int idx = tid; idx < D::CH * 324; idx += D::NT) {
      const int cl = idx / 324, a = (idx % 324) / 18, j = idx % 18, c = c0 + cl;
      const double* xl = XL + cl * 180;
      double v = 0.0;
      if (j < 12) {
        for (int bb = 0; bb < 12; ++bb) v += kloc<true>(Ml, Dl, Cc, c, a, bb) * xl[bb * 12 + j];
      } else {
        for (int bb = 0; bb < 6; ++bb)
          v += kloc<true>(Ml, Dl, Cc, c, a, 12 + bb) * xl[144 + bb * 6 + (j - 12)];
      }
      Yb[idx] = v;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < KP; ++k) {
      const int pr = tid + k * D::NT;
      if (pr < kNB * kNB) {
        const int i = pr / kNB, j = pr % kNB;
        double v = 0.0;
        for (int cl = 0; cl < D::CH; ++cl) {
          const double* xl = XL + cl * 180;
          const double* y = Yb + cl * 324;
          if (i < 12) {
            for (int a = 0; a < 12; ++a) v += xl[a * 12 + i] * y[a * 18 + j];
          } else {
            for (int a = 0; a < 6; ++a) v += xl[144 + a * 6 + (i - 12)] * y[(12 + a) * 18 + j];
          }
        }
        macc[k] += v;
      } else if (pr < kNB * kNB + 6) {
        const int i = pr - kNB * kNB;
        double v = 0.0;
        for (int cl = 0; cl < D::CH; ++cl)
          for (int a = 0; a < 6; ++a) v += XL[cl * 180 + 144 + a * 6 + i] * Fl[(c0 + cl) * 6 + a];
        macc[k] += v;
      }
    }
    __syncthreads();
  }
  bool bad = false;
#pragma unroll
  for (int k = 0; k < KP; ++k) {
    const int pr = tid + k * D::NT;
    if (pr < kNB * kNB) {
      g.M[(size_t)cell * kNB * kNB + pr] = macc[k];
      bad |= !isfinite(macc[k]);
    } else if (pr < kNB * kNB + 6) {
      g.b[(size_t)cell * kNB + 12 + (pr - kNB * kNB)] = macc[k];
    }
    if (pr < 12) g.b[(size_t)cell * kNB + pr] = 0.0;
  }
  if (bad) atomicCAS(&s_flag, (int)MSFEC_CELL_OK, (int)MSFEC_CELL_NONFINITE);
  __syncthreads();
  if (tid == 0) g.status[cell] = s_flag;
}

// set_global_weights (ned_rt_basis.cc:1156-1181): sigma from the edge part of
// the edge bases, u from the face part of the face bases
__global__ __launch_bounds__(256) void nrt_reconstruct_kernel(const double* X, const double* w,
                                                               int N, int ne, double* sig,
                                                               double* u) {
  const int cell = blockIdx.x;
  const double* wc = w + (size_t)cell * kNB;
  const double* xc = X + (size_t)cell * kNB * N;
  if (sig)
    for (int e = threadIdx.x; e < ne; e += blockDim.x) {
      double v = 0.0;
#pragma unroll
      for (int k = 0; k < 12; ++k) v += wc[k] * xc[(size_t)k * N + e];
      sig[(size_t)cell * ne + e] = v;
    }
  if (u)
    for (int f = threadIdx.x; f < N - ne; f += blockDim.x) {
      double v = 0.0;
#pragma unroll
      for (int k = 12; k < kNB; ++k) v += wc[k] * xc[(size_t)k * N + ne + f];
      u[(size_t)cell * (N - ne) + f] = v;
    }
}

// ===========================================================================
// Beyond the dense limit (n_refine_local 3 and 4, or any l on request): the
// interior system is never formed.  nrt_it_local_kernel writes the local
// matrices of every fine cell to HBM once; nrt_it_diag_kernel,
// nrt_it_coarse_kernel and nrt_it_coarse_inv_kernel the preconditioner (a
// diagonal plus a coarse correction on the faces); nrt_minres_kernel solves
// each basis in a workgroup of its
// own by preconditioned MINRES (Paige-Saunders), applying K' from the local
// matrices of the fine cells (element by element at l = 4, row by row
// below); and
// nrt_it_element_kernel forms the element matrix from the assembled products.
// The reference solves these systems either directly (solve_direct,
// ned_rt_basis.cc:579-634) or by CG on the face Schur complement
// (solve_iterative, :636-846); MINRES on the whole symmetric form reaches the
// direct solution to a relative tolerance on the preconditioned residual.

constexpr int kItMaxL = 4;
constexpr int kItNT = 1024;  // threads of a MINRES workgroup
// The product K' v element by element (8 colour passes of one fine cell per
// thread) where a colour class fills the workgroup (l = 4: 512 cells), row
// by row below (l = 3: 64 cells per class would leave most threads idle).
// Measured on one cell at l = 4: 0.51 s against 0.71 s row by row.
template <int L>
constexpr bool kItCellProduct = L >= 4;
// The edge block of the preconditioner: an element additive Schwarz at
// l = 4 (64 cells of the reference example: 690 iterations and 1.93 s
// against 1 117 and 2.17 s with the diagonal), the diagonal below (l = 3:
// its 8 colour passes of 64 cells cost more than the 10 % fewer iterations
// save).
template <int L>
constexpr bool kItEdgeSchwarz = L >= 4;
constexpr int kItNW = kItNT / 64;
constexpr int kLM = 99;     // local-matrix entries per fine cell: Me (78), then Du (21)

struct NrtItArgs {
  const double* verts;
  const double* ainv;
  const double* bq;
  const double* rq;  // or null
  const double* fq;  // or null
  const int* idof;   // n: the full index of every interior dof
  double* LM;        // n_cells x 99 x nfc (entry-major, fine cell fastest)
  double* CC;        // n_cells x 72: C[F][L] = (psi_F, curl phi_L) of one fine cell
  double* FL;        // n_cells x nfc x 6: (psi_F, f) per fine cell
  double* PD;        // n_cells x N: the inverse preconditioner diagonal (0 on the boundary)
  double* WK;        // n_cells x 18 x 6 x N: MINRES vectors v, r1, r2, w, w2, z
  double* S0;        // n_cells x NC x NC: the coarse matrix, then its inverse
  double* EI;        // n_cells x 78 x nfc: per fine cell the inverse of its edge mass block
                     // over its interior edges (packed lower, entry-major, colour-major cells)
  const int* agg;    // n: the aggregate of every interior face dof, -1 for an edge
  double* X;         // n_cells x 18 x N
  double* M;         // n_cells x 18 x 18
  double* b;         // n_cells x 18
  int* its;          // n_cells x 18: MINRES iterations
  int* flg;          // n_cells x 18: MSFEC_CELL_* of every solve
  int* status;       // n_cells
  int n_cells, N, n, max_iter;
  double rtol;
};

__device__ __forceinline__ void it_sizes(const double* v, V3& Hc, V3& h, int m) {
  Hc = V3{v[3] - v[0], v[7] - v[1], v[14] - v[2]};
  h = V3{Hc.x / m, Hc.y / m, Hc.z / m};
}

// full dof of local dof b of fine cell (c0, c1, c2)
template <int M>
__device__ __forceinline__ int it_cdof(int c0, int c1, int c2, int b) {
  constexpr int nx = M * (M + 1) * (M + 1), ne = 3 * nx, nf1 = M * M * (M + 1);
  if (b < 12) {
    const int ax = kLAx[b], f1 = kLF1[b], f2 = kLF2[b];
    if (ax == 0) return ((c2 + f2) * (M + 1) + c1 + f1) * M + c0;
    if (ax == 1) return nx + ((c2 + f2) * M + c1) * (M + 1) + c0 + f1;
    return 2 * nx + (c2 * (M + 1) + c1 + f2) * (M + 1) + c0 + f1;
  }
  const int F = b - 12, d = F >> 1, s = F & 1;
  if (d == 0) return ne + (c2 * M + c1) * (M + 1) + c0 + s;
  if (d == 1) return ne + nf1 + (c2 * (M + 1) + c1 + s) * M + c0;
  return ne + 2 * nf1 + ((c2 + s) * M + c1) * M + c0;
}

// Fine cell k (0..3 for an edge, 0..1 for a face) of the dof of type t at
// (p0, p1, p2) (dof_pos) and the local index a of the dof in it; false when
// that cell lies outside the coarse cell.  An edge along axis ax sits in the
// cells below and above it along the two other axes (line (f1, f2) of those
// cells), a face in the cells on its two sides.
template <int M>
__device__ __forceinline__ bool it_cell(int t, int p0, int p1, int p2, int k, int& c0, int& c1,
                                        int& c2, int& a) {
  const int u = k & 1, v = k >> 1;
  if (t == 0) {
    c0 = p0; c1 = p1 - u; c2 = p2 - v; a = 2 + u + 4 * v;  // lines 2, 3, 6, 7
  } else if (t == 1) {
    c0 = p0 - u; c1 = p1; c2 = p2 - v; a = u + 4 * v;      // lines 0, 1, 4, 5
  } else if (t == 2) {
    c0 = p0 - u; c1 = p1 - v; c2 = p2; a = 8 + u + 2 * v;  // lines 8 .. 11
  } else {
    const int d = t - 3;
    c0 = p0 - (d == 0 ? u : 0); c1 = p1 - (d == 1 ? u : 0); c2 = p2 - (d == 2 ? u : 0);
    a = 12 + 2 * d + u;
    if (k >= 2) return false;
  }
  return c0 >= 0 && c0 < M && c1 >= 0 && c1 < M && c2 >= 0 && c2 < M;
}

template <int M>
__device__ __forceinline__ bool it_on_boundary(int t, int p0, int p1, int p2) {
  if (t == 0) return p1 == 0 || p1 == M || p2 == 0 || p2 == M;
  if (t == 1) return p0 == 0 || p0 == M || p2 == 0 || p2 == M;
  if (t == 2) return p0 == 0 || p0 == M || p1 == 0 || p1 == M;
  const int pd = t == 3 ? p0 : (t == 4 ? p1 : p2);
  return pd == 0 || pd == M;
}

// Storage index of fine cell (c0, c1, c2) in the HBM copy of the local
// matrices: cells grouped by the parity of their indices (8 colour classes,
// within which no two cells share a dof), so that a colour pass of the
// element-by-element product reads contiguous memory.
template <int M>
__device__ __forceinline__ int it_lmc(int c0, int c1, int c2) {
  constexpr int H = M / 2;
  const int col = (c0 & 1) + 2 * (c1 & 1) + 4 * (c2 & 1);
  return col * (H * H * H) + ((c2 >> 1) * H + (c1 >> 1)) * H + (c0 >> 1);
}

// whether local dof b of fine cell (c0, c1, c2) lies on the coarse boundary
template <int M>
__device__ __forceinline__ bool it_local_on_boundary(int c0, int c1, int c2, int b) {
  auto out = [](int q) { return q == 0 || q == M; };
  if (b < 12) {
    const int ax = kLAx[b], f1 = kLF1[b], f2 = kLF2[b];
    if (ax == 0) return out(c1 + f1) || out(c2 + f2);
    if (ax == 1) return out(c0 + f1) || out(c2 + f2);
    return out(c0 + f1) || out(c1 + f2);
  }
  const int F = b - 12, d = F >> 1, s = F & 1;
  return out((d == 0 ? c0 : (d == 1 ? c1 : c2)) + s);
}

// entry (a, b) of the local matrix of fine cell c (signs as kloc), from the
// entry-major HBM copy
template <bool Orig>
__device__ __forceinline__ double it_kloc(const double* lm, const double* Cc, int nfc, int c, int a,
                                          int b) {
  const bool ea = a < 12, eb = b < 12;
  const int hi = max(a, b), lo = min(a, b);
  double v;
  if (ea && eb) v = lm[(size_t)(hi * (hi + 1) / 2 + lo) * nfc + c];
  else if (!ea && !eb) v = lm[(size_t)(78 + (hi - 12) * (hi - 11) / 2 + (lo - 12)) * nfc + c];
  else v = Cc[(hi - 12) * 12 + lo];
  const bool neg = Orig ? (ea && !eb) : !(ea && eb);
  return neg ? -v : v;
}

// Row f of K' (or, with Orig, of the assembled matrix, row = test) times x,
// over the columns Cols selects (1: edges, 2: faces, 3: all)
template <int M, bool Orig, int Cols>
__device__ __forceinline__ double it_row(int f, const double* x, const double* lm, const double* Cc) {
  constexpr int nfc = M * M * M;
  int t, p0, p1, p2;
  dof_pos<M>(f, t, p0, p1, p2);
  double acc = 0.0;
  const int nk = t < 3 ? 4 : 2;
  for (int k = 0; k < nk; ++k) {
    int c0, c1, c2, a;
    if (!it_cell<M>(t, p0, p1, p2, k, c0, c1, c2, a)) continue;
    const int c = it_lmc<M>(c0, c1, c2);
#pragma unroll
    for (int b = 0; b < 18; ++b) {
      if ((b < 12 && !(Cols & 1)) || (b >= 12 && !(Cols & 2))) continue;
      acc += it_kloc<Orig>(lm, Cc, nfc, c, a, b) * x[it_cdof<M>(c0, c1, c2, b)];
    }
  }
  return acc;
}

// Row f of K' times x for an interior dof: all 4 (edge) or 2 (face) fine
// cells exist, so no bounds checks; one cell's 18 terms at a time.
template <int M>
__device__ __forceinline__ double it_row_interior(int f, const double* __restrict__ x,
                                                  const double* __restrict__ lm, const double* Cc) {
  constexpr int nfc = M * M * M;
  int t, p0, p1, p2;
  dof_pos<M>(f, t, p0, p1, p2);
  double acc = 0.0;
  const int nk = t < 3 ? 4 : 2;
#pragma unroll 1
  for (int k = 0; k < nk; ++k) {
    int c0, c1, c2, a;
    it_cell<M>(t, p0, p1, p2, k, c0, c1, c2, a);
    const int c = it_lmc<M>(c0, c1, c2);
#pragma unroll
    for (int b = 0; b < 18; ++b)
      acc += it_kloc<false>(lm, Cc, nfc, c, a, b) * x[it_cdof<M>(c0, c1, c2, b)];
  }
  return acc;
}

// y += K'_c v_c for fine cell (c0, c1, c2), element by element: the cell's
// 18 values of v gathered once into registers, then one row at a time (a
// run-time loop, so that only one row's entries are in flight) dotted with
// them and added into y unless the row's dof is on the boundary.  Cells of
// one colour class share no dof, so a class needs no atomics.
template <int M>
__device__ __forceinline__ void it_cell_product(int c0, int c1, int c2, const double* v, double* y,
                                                const double* lm, const double* Cc) {
  constexpr int nfc = M * M * M;
  // entry e of this cell is (lm + e nfc)[sc]: a uniform base plus one 32-bit
  // per-lane offset shared by all loads
  const unsigned sc = (unsigned)it_lmc<M>(c0, c1, c2);
  double xv[18];
#pragma unroll
  for (int b = 0; b < 18; ++b) xv[b] = v[it_cdof<M>(c0, c1, c2, b)];
#pragma unroll 1
  for (int i = 0; i < 12; ++i) {
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < 12; ++j) {
      const int e = i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i;
      acc += (lm + (size_t)e * nfc)[sc] * xv[j];
    }
#pragma unroll
    for (int F = 0; F < 6; ++F) acc -= Cc[F * 12 + i] * xv[12 + F];
    if (!it_local_on_boundary<M>(c0, c1, c2, i)) y[it_cdof<M>(c0, c1, c2, i)] += acc;
  }
#pragma unroll 1
  for (int F = 0; F < 6; ++F) {
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < 12; ++j) acc -= Cc[F * 12 + j] * xv[j];
#pragma unroll
    for (int G = 0; G < 6; ++G) {
      const int e = 78 + (F >= G ? F * (F + 1) / 2 + G : G * (G + 1) / 2 + F);
      acc -= (lm + (size_t)e * nfc)[sc] * xv[12 + G];
    }
    if (!it_local_on_boundary<M>(c0, c1, c2, 12 + F)) y[it_cdof<M>(c0, c1, c2, 12 + F)] += acc;
  }
}

// (psi_f, curl N_nb) of the face dof of type t at (p0, p1, p2), over its two
// fine cells (QGauss(2) reduced to its 2 points along the face normal, as
// both factors depend on that coordinate only)
template <int M>
__device__ double it_curl_load(int t, int p0, int p1, int p2, int nb, V3 h, V3 Hc) {
  double acc = 0.0;
  for (int k = 0; k < 2; ++k) {
    int c0, c1, c2, a;
    if (!it_cell<M>(t, p0, p1, p2, k, c0, c1, c2, a)) continue;
    const int F = a - 12, ax = F >> 1;
    for (int qd = 0; qd < 2; ++qd) {
      const double gd = gauss_pt(qd);
      const V3 X{ax == 0 ? gd : 0.5, ax == 1 ? gd : 0.5, ax == 2 ? gd : 0.5};
      const V3 Xc{(c0 + X.x) / M, (c1 + X.y) / M, (c2 + X.z) / M};
      acc += rt_val(F, X, h) * ned_curl(nb, Xc, Hc, ax);
    }
  }
  return acc * 4.0 * (h.x * h.y * h.z * 0.125);
}

// The local matrices (edge mass with A^-1 and face div-div + reaction at the
// 8 Gauss points, the formulas of nrt_basis_kernel), the curl coupling and
// the loads (psi_F, f) of every fine cell of one coarse cell.
template <int L>
__global__ __launch_bounds__(256) void nrt_it_local_kernel(NrtItArgs g) {
  constexpr int M = 1 << L, nfc = M * M * M;
  __shared__ double Wt[78 * 8], PP[21 * 8], DD[21];
  __shared__ int cmp[78];
  const int cell = blockIdx.x, tid = threadIdx.x;
  V3 Hc, h;
  it_sizes(g.verts + (size_t)cell * 24, Hc, h, M);
  const double w8 = h.x * h.y * h.z * 0.125;
  for (int idx = tid; idx < 78 * 8 + 21 * 8 + 21 + 78; idx += 256) {
    if (idx < 78 * 8) {
      const int e = idx >> 3, q = idx & 7;
      int i, j;
      tri_ij(e, i, j);
      const V3 X = gauss_xyz(q);
      Wt[idx] = w8 * ned_val(i, X, h) * ned_val(j, X, h);
    } else if (idx < 78 * 8 + 21 * 8) {
      const int e = (idx - 78 * 8) >> 3, q = (idx - 78 * 8) & 7;
      int i, j;
      tri_ij(e, i, j);
      const V3 X = gauss_xyz(q);
      PP[e * 8 + q] = (i >> 1) == (j >> 1) ? w8 * rt_val(i, X, h) * rt_val(j, X, h) : 0.0;
    } else if (idx < 78 * 8 + 21 * 8 + 21) {
      int i, j;
      tri_ij(idx - 78 * 8 - 21 * 8, i, j);
      DD[idx - 78 * 8 - 21 * 8] = w8 * rt_div(i, h) * rt_div(j, h);
    } else {
      const int e = idx - 78 * 8 - 21 * 8 - 21;
      int i, j;
      tri_ij(e, i, j);
      cmp[e] = sym6(kLineAx[i], kLineAx[j]);
    }
  }
  __syncthreads();
  double* lm = g.LM + (size_t)cell * kLM * nfc;
  for (int idx = tid; idx < nfc * kLM; idx += 256) {
    const int e = idx / nfc, c = idx - e * nfc;
    const size_t gi = (size_t)cell * nfc + c;
    double acc = 0.0;
    if (e < 78) {
      const int ci = cmp[e];
#pragma unroll
      for (int q = 0; q < 8; ++q) acc += Wt[e * 8 + q] * g.ainv[(gi * 8 + q) * 6 + ci];
    } else {
      const int e2 = e - 78;
      const double dd = DD[e2];
#pragma unroll
      for (int q = 0; q < 8; ++q)
        acc += g.bq[gi * 8 + q] * dd + (g.rq ? g.rq[gi * 8 + q] : 0.0) * PP[e2 * 8 + q];
    }
    lm[(size_t)e * nfc + it_lmc<M>(c % M, (c / M) % M, c / (M * M))] = acc;
  }
  for (int idx = tid; idx < 72 + nfc * 6; idx += 256) {
    if (idx < 72) {
      const int F = idx / 12, Ln = idx % 12, ax = F >> 1;
      double acc = 0.0;
      for (int q = 0; q < 8; ++q) {
        const V3 X = gauss_xyz(q);
        acc += rt_val(F, X, h) * ned_curl(Ln, X, h, ax);
      }
      g.CC[(size_t)cell * 72 + idx] = acc * w8;
    } else {
      const int c = (idx - 72) / 6, F = (idx - 72) % 6, ax = F >> 1;
      double acc = 0.0;
      if (g.fq) {
        const double* f = g.fq + ((size_t)cell * nfc + c) * 24;
        for (int q = 0; q < 8; ++q) {
          const V3 X = gauss_xyz(q);
          acc += rt_val(F, X, h) * f[q * 3 + ax];
        }
      }
      g.FL[((size_t)cell * nfc + c) * 6 + F] = acc * w8;
    }
  }
}

// diagonal of the edge block at full edge dof e: the sum over its fine cells
template <int M>
__device__ __forceinline__ double it_edge_diag(int e, const double* lm) {
  constexpr int nfc = M * M * M;
  int t, p0, p1, p2;
  dof_pos<M>(e, t, p0, p1, p2);
  double d = 0.0;
  for (int k = 0; k < 4; ++k) {
    int c0, c1, c2, a;
    if (it_cell<M>(t, p0, p1, p2, k, c0, c1, c2, a))
      d += lm[(size_t)(a * (a + 1) / 2 + a) * nfc + it_lmc<M>(c0, c1, c2)];
  }
  return d;
}

// The diagonal part of the MINRES preconditioner, the inverse of a positive
// diagonal: the edge mass diagonal on the edges, and on the faces the
// diagonal of the Schur complement D + C diag(Me)^-1 C^T (interior edges
// only).  Zero on the boundary.
template <int L>
__global__ __launch_bounds__(256) void nrt_it_diag_kernel(NrtItArgs g) {
  constexpr int M = 1 << L, nfc = M * M * M;
  const int cell = blockIdx.x;
  const double* lm = g.LM + (size_t)cell * kLM * nfc;
  const double* Cc = g.CC + (size_t)cell * 72;
  double* pd = g.PD + (size_t)cell * g.N;
  for (int f = threadIdx.x; f < g.N; f += 256) {
    int t, p0, p1, p2;
    dof_pos<M>(f, t, p0, p1, p2);
    if (it_on_boundary<M>(t, p0, p1, p2)) {
      pd[f] = 0.0;
      continue;
    }
    double d;
    if (t < 3) {
      d = it_edge_diag<M>(f, lm);
    } else {
      int cc[2][3], aa[2];
      bool in[2];
      d = 0.0;
      for (int k = 0; k < 2; ++k) {
        in[k] = it_cell<M>(t, p0, p1, p2, k, cc[k][0], cc[k][1], cc[k][2], aa[k]);
        if (in[k]) {
          const int a = aa[k] - 12;
          d += lm[(size_t)(78 + a * (a + 1) / 2 + a) * nfc + it_lmc<M>(cc[k][0], cc[k][1], cc[k][2])];
        }
      }
      // C_fe summed over the fine cells both dofs share
      for (int k = 0; k < 2; ++k) {
        if (!in[k]) continue;
        for (int Ln = 0; Ln < 12; ++Ln) {
          const int e = it_cdof<M>(cc[k][0], cc[k][1], cc[k][2], Ln);
          int te, q0, q1, q2;
          dof_pos<M>(e, te, q0, q1, q2);
          if (it_on_boundary<M>(te, q0, q1, q2)) continue;
          double cfe = Cc[(aa[k] - 12) * 12 + Ln];
          bool dup = false;
          const int o = 1 - k;
          if (in[o])
            for (int L2 = 0; L2 < 12; ++L2)
              if (it_cdof<M>(cc[o][0], cc[o][1], cc[o][2], L2) == e) {
                if (k == 1) dup = true;  // counted from cell 0
                cfe += Cc[(aa[o] - 12) * 12 + L2];
              }
          if (!dup) d += cfe * cfe / it_edge_diag<M>(e, lm);
        }
      }
    }
    pd[f] = (d > 0.0 && isfinite(d)) ? 1.0 / d : 0.0;
  }
}

// The coarse space of the MINRES preconditioner: the interior faces of each
// orientation aggregated in NB^3 blocks of B^3 fine cells (NC = 3 NB^3
// coarse dofs, 192 from l = 3 on).  The preconditioner is the additive
// two-level
//   z = diag^-1 r + R0^T S0^-1 R0 r_faces,  S0 = R0 (D + C diag(Me)^-1 C^T) R0^T,
// R0 summing the faces of each aggregate: symmetric positive definite, as
// MINRES needs, and it removes the smooth face modes that make the Schur
// block's condition number grow as h^-2.
template <int M>
struct ItCoarse {
  static constexpr int NB = M <= 4 ? M / 2 : 4;
  static constexpr int B = M / NB;
  static constexpr int NC = 3 * NB * NB * NB;
};

// aggregate of the face of orientation d at (q0, q1, q2) (its level along d,
// its cell indices along the other axes)
template <int M>
__device__ __forceinline__ int it_agg(int d, int q0, int q1, int q2) {
  typedef ItCoarse<M> C;
  const int a0 = min(q0 / C::B, C::NB - 1), a1 = min(q1 / C::B, C::NB - 1),
            a2 = min(q2 / C::B, C::NB - 1);
  return d * C::NB * C::NB * C::NB + (a2 * C::NB + a1) * C::NB + a0;
}

// whether local face G of fine cell (c0, c1, c2) is an interior face of
// aggregate J
template <int M>
__device__ __forceinline__ bool it_face_in(int c0, int c1, int c2, int G, int J) {
  const int d = G >> 1, s = G & 1;
  const int q0 = c0 + (d == 0 ? s : 0), q1 = c1 + (d == 1 ? s : 0), q2 = c2 + (d == 2 ? s : 0);
  const int lev = d == 0 ? q0 : (d == 1 ? q1 : q2);
  return lev > 0 && lev < M && it_agg<M>(d, q0, q1, q2) == J;
}

// The inclusive range of aggregate I along axis k (levels along its
// normal, interior only; cells along the others)
template <int M>
__device__ __forceinline__ void it_agg_range(int I, int k, int& lo, int& hi) {
  typedef ItCoarse<M> C;
  constexpr int NB3 = C::NB * C::NB * C::NB;
  const int d = I / NB3, r = I % NB3;
  const int a = k == 0 ? r % C::NB : (k == 1 ? (r / C::NB) % C::NB : r / (C::NB * C::NB));
  lo = a * C::B;
  hi = a == C::NB - 1 ? (k == d ? M - 1 : M - 1) : a * C::B + C::B - 1;
  if (k == d) lo = max(lo, 1);
}

// full dof of the face of orientation d at (q0, q1, q2)
template <int M>
__device__ __forceinline__ int it_face_dof(int d, int q0, int q1, int q2) {
  constexpr int nx = M * (M + 1) * (M + 1), ne = 3 * nx, nf1 = M * M * (M + 1);
  if (d == 0) return ne + (q2 * M + q1) * (M + 1) + q0;
  if (d == 1) return ne + nf1 + (q2 * (M + 1) + q1) * M + q0;
  return ne + 2 * nf1 + (q2 * M + q1) * M + q0;
}

// Column J of S0 for one coarse cell per workgroup: u = (D + C diag(Me)^-1
// C^T) 1_J on every interior face into LDS (faces too far from the
// aggregate are zero and skipped), then S0[I][J] = the sum of u over the
// faces of I, one thread per I in a fixed order (no atomics: the
// preconditioner, and so every result, is reproducible bit for bit).
template <int L>
__global__ __launch_bounds__(256) void nrt_it_coarse_kernel(NrtItArgs g) {
  constexpr int M = 1 << L, nfc = M * M * M, nx = M * (M + 1) * (M + 1), ne = 3 * nx;
  constexpr int NF = 3 * M * M * (M + 1);
  typedef ItCoarse<M> CS;
  extern __shared__ double uf[];  // NF
  const int cell = blockIdx.x / CS::NC, J = blockIdx.x % CS::NC, tid = threadIdx.x;
  const double* lm = g.LM + (size_t)cell * kLM * nfc;
  const double* Cc = g.CC + (size_t)cell * 72;
  // the cells with a face in J, widened by the reach of C diag(Me)^-1 C^T
  int jlo[3], jhi[3];
  for (int k = 0; k < 3; ++k) {
    it_agg_range<M>(J, k, jlo[k], jhi[k]);
    if (k == J / (CS::NB * CS::NB * CS::NB)) jlo[k] -= 1;  // the cell below the lowest level
    jlo[k] -= 2;
    jhi[k] += 2;
  }
  for (int f = tid; f < NF; f += 256) {
    int t, p0, p1, p2;
    dof_pos<M>(ne + f, t, p0, p1, p2);
    double u = 0.0;
    const bool near = p0 >= jlo[0] && p0 <= jhi[0] + 1 && p1 >= jlo[1] && p1 <= jhi[1] + 1 &&
                      p2 >= jlo[2] && p2 <= jhi[2] + 1;
    if (near && !it_on_boundary<M>(t, p0, p1, p2)) {
      for (int k = 0; k < 2; ++k) {
        int c0, c1, c2, a;
        it_cell<M>(t, p0, p1, p2, k, c0, c1, c2, a);
        const int F = a - 12;
        const int sc = it_lmc<M>(c0, c1, c2);
        for (int G = 0; G < 6; ++G)
          if (it_face_in<M>(c0, c1, c2, G, J)) {
            const int hi = max(F, G), lo = min(F, G);
            u += lm[(size_t)(78 + hi * (hi + 1) / 2 + lo) * nfc + sc];
          }
        for (int Ln = 0; Ln < 12; ++Ln) {
          const int e = it_cdof<M>(c0, c1, c2, Ln);
          int te, q0, q1, q2;
          dof_pos<M>(e, te, q0, q1, q2);
          if (it_on_boundary<M>(te, q0, q1, q2)) continue;
          double we = 0.0, me = 0.0;
          for (int k2 = 0; k2 < 4; ++k2) {
            int d0, d1, d2, a2;
            it_cell<M>(te, q0, q1, q2, k2, d0, d1, d2, a2);
            me += lm[(size_t)(a2 * (a2 + 1) / 2 + a2) * nfc + it_lmc<M>(d0, d1, d2)];
            for (int G = 0; G < 6; ++G)
              if (it_face_in<M>(d0, d1, d2, G, J)) we += Cc[G * 12 + a2];
          }
          if (we != 0.0) u += Cc[F * 12 + Ln] * we / me;
        }
      }
    }
    uf[f] = u;
  }
  __syncthreads();
  for (int I = tid; I < CS::NC; I += 256) {
    const int d = I / (CS::NB * CS::NB * CS::NB);
    int lo[3], hi[3];
    for (int k = 0; k < 3; ++k) it_agg_range<M>(I, k, lo[k], hi[k]);
    double sum = 0.0;
    for (int q2 = lo[2]; q2 <= hi[2]; ++q2)
      for (int q1 = lo[1]; q1 <= hi[1]; ++q1)
        for (int q0 = lo[0]; q0 <= hi[0]; ++q0) sum += uf[it_face_dof<M>(d, q0, q1, q2) - ne];
    g.S0[((size_t)cell * CS::NC + I) * CS::NC + J] = sum;
  }
}

// The edge part of the MINRES preconditioner, an element additive Schwarz:
// z_e = sum over fine cells of (Me_c restricted to its interior edges)^-1
// r_c.  One thread per fine cell inverts its 12 x 12 block in LDS by
// Gauss-Jordan (a boundary edge becomes a unit row and column, then zero).
template <int L>
__global__ __launch_bounds__(64) void nrt_it_edgeas_kernel(NrtItArgs g) {
  constexpr int M = 1 << L, nfc = M * M * M;
  constexpr int CPB = nfc < 64 ? nfc : 64;  // fine cells per workgroup
  __shared__ double A[64][145];
  const int tid = threadIdx.x;
  const int cell = blockIdx.x / (nfc / CPB), c = (blockIdx.x % (nfc / CPB)) * CPB + tid;
  if (tid >= CPB) return;
  const int c0 = c % M, c1 = (c / M) % M, c2 = c / (M * M);
  const int sc = it_lmc<M>(c0, c1, c2);
  const double* lm = g.LM + (size_t)cell * kLM * nfc;
  double* a = A[tid];
  bool in[12];
  for (int i = 0; i < 12; ++i) in[i] = !it_local_on_boundary<M>(c0, c1, c2, i);
  for (int i = 0; i < 12; ++i)
    for (int j = 0; j <= i; ++j) {
      const double v = (in[i] && in[j]) ? lm[(size_t)(i * (i + 1) / 2 + j) * nfc + sc]
                                         : (i == j ? 1.0 : 0.0);
      a[i * 12 + j] = v;
      a[j * 12 + i] = v;
    }
  for (int k = 0; k < 12; ++k) {
    const double p = 1.0 / a[k * 12 + k];
    for (int i = 0; i < 12; ++i)
      for (int j = 0; j < 12; ++j)
        if (i != k && j != k) a[i * 12 + j] -= a[i * 12 + k] * a[k * 12 + j] * p;
    for (int i = 0; i < 12; ++i)
      if (i != k) {
        a[i * 12 + k] *= -p;
        a[k * 12 + i] *= p;
      }
    a[k * 12 + k] = p;
  }
  double* ei = g.EI + (size_t)cell * 78 * nfc;
  for (int i = 0; i < 12; ++i)
    for (int j = 0; j <= i; ++j)
      ei[(size_t)(i * (i + 1) / 2 + j) * nfc + sc] =
          (in[i] && in[j]) ? 0.5 * (a[i * 12 + j] + a[j * 12 + i]) : 0.0;
}

// z += EI_c y_c on the interior edges of fine cell (c0, c1, c2); returns
// y_c . EI_c y_c, this cell's part of y . z
template <int M>
__device__ __forceinline__ double it_cell_edge_precond(int c0, int c1, int c2, const double* y,
                                                       double* z, const double* ei) {
  constexpr int nfc = M * M * M;
  const unsigned sc = (unsigned)it_lmc<M>(c0, c1, c2);
  double yv[12];
#pragma unroll
  for (int b = 0; b < 12; ++b) yv[b] = it_local_on_boundary<M>(c0, c1, c2, b) ? 0.0 : y[it_cdof<M>(c0, c1, c2, b)];
  double quad = 0.0;
#pragma unroll 1
  for (int i = 0; i < 12; ++i) {
    double acc = 0.0;
#pragma unroll
    for (int j = 0; j < 12; ++j) {
      const int e = i >= j ? i * (i + 1) / 2 + j : j * (j + 1) / 2 + i;
      acc += (ei + (size_t)e * nfc)[sc] * yv[j];
    }
    quad += yv[i] * acc;
    if (!it_local_on_boundary<M>(c0, c1, c2, i)) z[it_cdof<M>(c0, c1, c2, i)] += acc;
  }
  return quad;
}

// S0 <- S0^-1 in place, one coarse cell per workgroup: Gauss-Jordan without
// pivoting (S0 is symmetric positive definite), two barriers per pivot.
template <int L>
__global__ __launch_bounds__(1024) void nrt_it_coarse_inv_kernel(NrtItArgs g) {
  constexpr int M = 1 << L;
  constexpr int NC = ItCoarse<M>::NC;
  double* A = g.S0 + (size_t)blockIdx.x * NC * NC;
  for (int k = 0; k < NC; ++k) {
    const double p = 1.0 / A[k * NC + k];
    for (int idx = threadIdx.x; idx < NC * NC; idx += 1024) {
      const int i = idx / NC, j = idx % NC;
      if (i != k && j != k) A[idx] -= A[i * NC + k] * A[k * NC + j] * p;
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < 2 * NC; idx += 1024) {
      if (idx < NC) {
        if (idx != k) A[idx * NC + k] *= -p;
      } else if (idx - NC != k) {
        A[k * NC + idx - NC] *= p;
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) A[k * NC + k] = p;
    __syncthreads();
  }
}

// sum over the workgroup; red holds one value per wave and must not be
// written again before every thread has read it
__device__ __forceinline__ double it_block_sum(double v, double* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < kItNW; ++i) s += red[i];
  return s;
}

// cq = S0^-1 R0 y (y's face values; LDS q and cq of NC values): four
// threads per aggregate sum every fourth of its faces in a fixed order and
// combine by lane shuffles, then one thread per coarse row forms the product
// with the inverse.  Returns this thread's part of q . cq (the coarse part of
// y . M^-1 y).  Deterministic; ends with a barrier.
template <int M>
__device__ __forceinline__ double it_coarse_correct(const double* y, const double* S0i, double* q,
                                                    double* cq) {
  typedef ItCoarse<M> CS;
  static_assert(4 * CS::NC <= kItNT, "four threads per aggregate");
  const int tid = threadIdx.x, I = tid >> 2, part = tid & 3;
  double sum = 0.0;
  if (I < CS::NC) {
    const int d = I / (CS::NB * CS::NB * CS::NB);
    int lo[3], hi[3];
    for (int k = 0; k < 3; ++k) it_agg_range<M>(I, k, lo[k], hi[k]);
    int cnt = 0;
    for (int q2 = lo[2]; q2 <= hi[2]; ++q2)
      for (int q1 = lo[1]; q1 <= hi[1]; ++q1)
        for (int q0 = lo[0]; q0 <= hi[0]; ++q0, ++cnt)
          if ((cnt & 3) == part) sum += y[it_face_dof<M>(d, q0, q1, q2)];
  }
  sum += __shfl_xor(sum, 1);
  sum += __shfl_xor(sum, 2);
  if (I < CS::NC && part == 0) q[I] = sum;
  __syncthreads();
  double qcq = 0.0;
  if (tid < CS::NC) {
    const double* row = S0i + (size_t)tid * CS::NC;
    double c = 0.0;
    for (int J = 0; J < CS::NC; ++J) c += row[J] * q[J];
    cq[tid] = c;
    qcq = q[tid] * c;
  }
  __syncthreads();
  return qcq;
}

// One basis of one coarse cell per workgroup.  Workgroups are numbered so
// that those dispatched to the same XCD (block index mod 8) take the 18
// bases of one cell one after another, which keeps that cell's local
// matrices in the XCD's L2.
template <int L>
__global__ __launch_bounds__(kItNT) void nrt_minres_kernel(NrtItArgs g) {
  constexpr int M = 1 << L, nfc = M * M * M, ne = 3 * M * (M + 1) * (M + 1);
  constexpr int NC = ItCoarse<M>::NC;
  __shared__ double redA[kItNW], redB[kItNW];
  __shared__ double sq[NC], scq[NC];
  const int bid = blockIdx.x, slot = bid >> 3;
  const int cell = (slot / kNB) * 8 + (bid & 7), r = slot % kNB;
  if (cell >= g.n_cells) return;
  const int tid = threadIdx.x, N = g.N, n = g.n;
  V3 Hc, h;
  it_sizes(g.verts + (size_t)cell * 24, Hc, h, M);
  const double* lm = g.LM + (size_t)cell * kLM * nfc;
  const double* pd = g.PD + (size_t)cell * N;
  const double* gCc = g.CC + (size_t)cell * 72;  // uniform: read through the scalar cache
  const double* S0i = g.S0 + (size_t)cell * NC * NC;
  double* x = g.X + ((size_t)cell * kNB + r) * N;
  double* v = g.WK + ((size_t)cell * kNB + r) * 6 * N;
  double* P = v + N;   // r1
  double* Q = v + 2 * N;  // r2
  double* W0 = v + 3 * N;  // w
  double* W1 = v + 4 * N;  // w2
  double* Zb = v + 5 * N;  // z = M^-1 r2 on the edges
  const double* ei = g.EI + (size_t)cell * 78 * nfc;
  constexpr int H = M / 2;
  // z_e = the element additive Schwarz of y's edge values, colour class by
  // colour class (Zb zeroed on the interior edges beforehand); adds y_e . z_e
  auto edge_precond = [&](const double* y) {
    double quad = 0.0;
#pragma unroll 1
    for (int col = 0; col < 8; ++col) {
      for (int q = tid; q < H * H * H; q += kItNT)
        quad += it_cell_edge_precond<M>(2 * (q % H) + (col & 1), 2 * ((q / H) % H) + ((col >> 1) & 1),
                                        2 * (q / (H * H)) + (col >> 2), y, Zb, ei);
      __syncthreads();
    }
    return quad;
  };
  // x = the boundary values, zero inside; the work vectors zero
  for (int f = tid; f < N; f += kItNT) {
    int t, p0, p1, p2;
    dof_pos<M>(f, t, p0, p1, p2);
    x[f] = it_on_boundary<M>(t, p0, p1, p2) ? bnd_value<M>(f, r) : 0.0;
    v[f] = 0.0;
    W0[f] = 0.0;
    W1[f] = 0.0;
  }
  __syncthreads();
  // r1 = r2 = S F - K' x (the curl loads of the edge bases on the face rows)
  double acc = 0.0;
  for (int i = tid; i < n; i += kItNT) {
    const int f = g.idof[i];
    double rr = -it_row<M, false, 3>(f, x, lm, gCc);
    if (r < 12) {
      int t, p0, p1, p2;
      dof_pos<M>(f, t, p0, p1, p2);
      if (t >= 3) rr -= it_curl_load<M>(t, p0, p1, p2, r, h, Hc);
    }
    P[f] = 0.0;  // r1 enters the first iteration scaled by 0
    Q[f] = rr;
    if (kItEdgeSchwarz<L> && f < ne) Zb[f] = 0.0;
    else acc += rr * rr * pd[f];
  }
  __syncthreads();
  // r . M^-1 r: on the faces the diagonal part plus q . S0^-1 q of the
  // coarse correction; on the edges the diagonal part, or at l = 4 the
  // cells' quadratic forms of the additive Schwarz
  acc += it_coarse_correct<M>(Q, S0i, sq, scq);
  if constexpr (kItEdgeSchwarz<L>) acc += edge_precond(Q);
  const double beta1 = sqrt(it_block_sum(acc, redB));
  int flag = isfinite(beta1) ? MSFEC_CELL_OK : MSFEC_CELL_NONFINITE;
  int itn = 0;
  if (flag == MSFEC_CELL_OK && beta1 > 0.0) {
    for (int i = tid; i < n; i += kItNT) {
      const int f = g.idof[i], ag = g.agg[i];
      v[f] = (ag >= 0 ? pd[f] * Q[f] + scq[ag] : (kItEdgeSchwarz<L> ? Zb[f] : pd[f] * Q[f])) / beta1;
    }
    __syncthreads();
    double oldb = 0.0, beta = beta1, dbar = 0.0, epsln = 0.0, phibar = beta1, cs = -1.0, sn = 0.0;
    const double stop = g.rtol * beta1;
    flag = MSFEC_CELL_NOT_CONVERGED;
    while (itn < g.max_iter) {
      ++itn;
      // A: t = K' v - (beta / oldb) r1 into r1's slot (at l = 4 the scaled
      // r1, then the fine cells' products added colour class by colour
      // class; below, row by row); alfa = v . t
      const double c1 = itn > 1 ? beta / oldb : 0.0;
      acc = 0.0;
      double alfa;
      if constexpr (kItCellProduct<L>) {
        for (int i = tid; i < n; i += kItNT) {
          const int f = g.idof[i];
          P[f] = -c1 * P[f];
        }
        __syncthreads();
#pragma unroll 1
        for (int col = 0; col < 8; ++col) {
          for (int q = tid; q < H * H * H; q += kItNT)
            it_cell_product<M>(2 * (q % H) + (col & 1), 2 * ((q / H) % H) + ((col >> 1) & 1),
                               2 * (q / (H * H)) + (col >> 2), v, P, lm, gCc);
          __syncthreads();
        }
        for (int i = tid; i < n; i += kItNT) {
          const int f = g.idof[i];
          acc += v[f] * P[f];
        }
      } else {
        for (int i = tid; i < n; i += kItNT) {
          const int f = g.idof[i];
          const double t = it_row_interior<M>(f, v, lm, gCc) - c1 * P[f];
          P[f] = t;
          acc += v[f] * t;
        }
      }
      alfa = it_block_sum(acc, redA);
      // B: r2' = t - (alfa / beta) r2 into r1's slot, then r1 <-> r2;
      // beta' = |r2'| in the preconditioner's norm (its coarse part formed
      // between the two passes)
      const double c2 = alfa / beta;
      acc = 0.0;
      for (int i = tid; i < n; i += kItNT) {
        const int f = g.idof[i];
        const double y = P[f] - c2 * Q[f];
        P[f] = y;
        if (kItEdgeSchwarz<L> && f < ne) Zb[f] = 0.0;
        else acc += y * y * pd[f];
      }
      __syncthreads();
      acc += it_coarse_correct<M>(P, S0i, sq, scq);
      if constexpr (kItEdgeSchwarz<L>) acc += edge_precond(P);
      {
        double* tmp = P;
        P = Q;
        Q = tmp;
      }
      oldb = beta;
      beta = sqrt(it_block_sum(acc, redB));
      // the QR update of the Lanczos tridiagonal (Givens rotations)
      const double oldeps = epsln;
      const double delta = cs * dbar + sn * alfa;
      const double gbar = sn * dbar - cs * alfa;
      epsln = sn * beta;
      dbar = -cs * beta;
      double gamma = sqrt(gbar * gbar + beta * beta);
      gamma = fmax(gamma, 1e-300);
      cs = gbar / gamma;
      sn = beta / gamma;
      const double phi = cs * phibar;
      phibar = sn * phibar;
      const double denom = 1.0 / gamma;
      const double sb = beta > 0.0 ? 1.0 / beta : 0.0;
      // C: w' = (v - oldeps w2 - delta w) / gamma into w2's slot, then
      // w <-> w2; x += phi w'; the next v = M^-1 r2 / beta
      for (int i = tid; i < n; i += kItNT) {
        const int f = g.idof[i], ag = g.agg[i];
        const double wn = (v[f] - oldeps * W1[f] - delta * W0[f]) * denom;
        W1[f] = wn;
        x[f] += phi * wn;
        v[f] = (ag >= 0 ? pd[f] * Q[f] + scq[ag] : (kItEdgeSchwarz<L> ? Zb[f] : pd[f] * Q[f])) * sb;
      }
      {
        double* tmp = W0;
        W0 = W1;
        W1 = tmp;
      }
      __syncthreads();
      if (!isfinite(phibar) || !isfinite(beta)) {
        flag = MSFEC_CELL_NONFINITE;
        break;
      }
      if (phibar <= stop || beta == 0.0) {
        flag = MSFEC_CELL_OK;
        break;
      }
    }
  }
  // the assembled matrix times the block of this basis that enters the
  // element matrix (edges for an edge basis, faces for a face basis), over
  // every row, into r1's slot of the basis for nrt_it_element_kernel
  double* Z = g.WK + ((size_t)cell * kNB + r) * 6 * N + N;
  for (int f = tid; f < N; f += kItNT)
    Z[f] = r < 12 ? it_row<M, true, 1>(f, x, lm, gCc) : it_row<M, true, 2>(f, x, lm, gCc);
  if (tid == 0) {
    g.its[(size_t)cell * kNB + r] = itn;
    g.flg[(size_t)cell * kNB + r] = flag;
  }
}

// Row i of the element matrix and load of one coarse cell per workgroup
// (assemble_global_element_matrix, ned_rt_basis.cc:850-948): M(i, j) = the
// block of basis i times the assembled matrix times the block of basis j,
// the products of nrt_minres_kernel; b(i) = the face block of a face basis
// times the load (v, f).
template <int L>
__global__ __launch_bounds__(256) void nrt_it_element_kernel(NrtItArgs g) {
  constexpr int M = 1 << L, nfc = M * M * M;
  constexpr int nx = M * (M + 1) * (M + 1), ne = 3 * nx;
  __shared__ double red[4][kNB + 1];
  const int cell = blockIdx.x / kNB, i = blockIdx.x % kNB, tid = threadIdx.x;
  const int N = g.N;
  const double* xi = g.X + ((size_t)cell * kNB + i) * N;
  const double* fl = g.FL + (size_t)cell * nfc * 6;
  double acc[kNB + 1];
#pragma unroll
  for (int j = 0; j <= kNB; ++j) acc[j] = 0.0;
  const int r0 = i < 12 ? 0 : ne, r1 = i < 12 ? ne : N;
  for (int f = r0 + tid; f < r1; f += 256) {
    const double xv = xi[f];
#pragma unroll
    for (int j = 0; j < kNB; ++j) acc[j] += xv * g.WK[(((size_t)cell * kNB + j) * 6 + 1) * N + f];
    if (i >= 12) {
      int t, p0, p1, p2;
      dof_pos<M>(f, t, p0, p1, p2);
      double lf = 0.0;
      for (int k = 0; k < 2; ++k) {
        int c0, c1, c2, a;
        if (it_cell<M>(t, p0, p1, p2, k, c0, c1, c2, a))
          lf += fl[((c2 * M + c1) * M + c0) * 6 + a - 12];
      }
      acc[kNB] += xv * lf;
    }
  }
#pragma unroll
  for (int j = 0; j <= kNB; ++j) {
    double s = acc[j];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((tid & 63) == 0) red[tid >> 6][j] = s;
  }
  __syncthreads();
  if (tid <= kNB) {
    const double s = red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid];
    if (tid < kNB) g.M[((size_t)cell * kNB + i) * kNB + tid] = s;
    else g.b[(size_t)cell * kNB + i] = s;
  }
}

// Per-cell status from the 18 solves and the element matrix.
__global__ __launch_bounds__(256) void nrt_it_status_kernel(NrtItArgs g) {
  const int cell = blockIdx.x * 256 + threadIdx.x;
  if (cell >= g.n_cells) return;
  int st = MSFEC_CELL_OK;
  for (int r = 0; r < kNB; ++r) {
    const int f = g.flg[(size_t)cell * kNB + r];
    if (f != MSFEC_CELL_OK && st == MSFEC_CELL_OK) st = f;
  }
  for (int k = 0; k < kNB * kNB && st == MSFEC_CELL_OK; ++k)
    if (!isfinite(g.M[(size_t)cell * kNB * kNB + k])) st = MSFEC_CELL_NONFINITE;
  g.status[cell] = st;
}

// ===========================================================================
// RT_DQ in three dimensions (source/RT_DQ/rt_dq_basis.cc): the 6 face bases
// of the mixed Raviart-Thomas x DQ0 system.  As in two dimensions the flux
// of basis n is the coarse Raviart-Thomas function c_n (which has the
// prescribed divergence -+1/|K| and boundary flux) plus a divergence-free
// correction with zero boundary flux; in three dimensions those fields are
// the curls of the interior Nedelec fields (the discrete exact sequence), so
// sigma_n = c_n + D psi_n with D the face circulation of the edge values and
// psi_n solving the curl-curl system
//   D^T M_A D psi = -D^T M_A c_n   (M_A the face mass with A^-1)
// over the interior edges: the first mixed equation tested with the
// divergence-free fields, the second holding by construction.  The system
// is semi-definite (gradients are its kernel) and consistent; Jacobi-PCG
// converges on its range, and sigma does not see the kernel.  u is 1
// (set_u_to_std, :1039-1043).  No reaction term (ReactionRate is zero,
// eqn_coeff_R.cc:8-14).

struct Rt3Args {
  const double* verts;  // n_cells x 8 x 3
  const double* ainv;   // n_cells x nfc x 8 x 6
  const double* fq;     // n_cells x nfc x 8, or null
  double* MF;           // n_cells x 21 x nfc: face mass with A^-1 per fine cell (colour-major)
  double* PD;           // n_cells x ns: inverse Jacobi diagonal
  double* WK;           // n_cells x 6 x 4 x ns: x, r, p, q over the interior edges
  double* S;            // n_cells x 6 x nf: the fine fluxes
  double* M;            // n_cells x 7 x 7
  double* b;            // n_cells x 7
  int* its;             // n_cells x 6
  int* flg;             // n_cells x 6
  int* status;          // n_cells
  int n_cells, ns, max_iter;
  double rtol;
};

// Circulation sign of local line L in local face F of a fine cell (counter-
// clockwise seen from the face's +axis side), 0 when L is not on F: the
// coefficient of face F in the curl of the Nedelec function of L.
constexpr int rt3_inc(int F, int L) {
  const int ax = kLAx[L], o1 = ax == 0 ? 1 : 0, o2 = ax == 2 ? 1 : 2;
  const int f1 = kLF1[L], f2 = kLF2[L], d = F / 2, s = F % 2;
  if (d == ax) return 0;
  if ((d == o1 ? f1 : f2) != s) return 0;
  const int a = (d + 1) % 3, bb = (d + 2) % 3;
  if (ax == a) return (bb == o1 ? f1 : f2) == 0 ? 1 : -1;
  return (a == o1 ? f1 : f2) == 1 ? 1 : -1;
}

// the coarse Raviart-Thomas function of face n: its flux through local face
// F of fine cell (c0, c1, c2)
template <int M>
__device__ __forceinline__ double rt3_coarse_flux(int c0, int c1, int c2, int F, int n) {
  const int d = F >> 1;
  if (d != (n >> 1)) return 0.0;
  const int lev = (d == 0 ? c0 : (d == 1 ? c1 : c2)) + (F & 1);
  return lag(n & 1, (double)lev / M) / (M * M);
}

// fine face index (without the edge offset) of local face F
template <int M>
__device__ __forceinline__ int rt3_face(int c0, int c1, int c2, int F) {
  constexpr int nx = M * (M + 1) * (M + 1), ne = 3 * nx;
  return it_cdof<M>(c0, c1, c2, 12 + F) - ne;
}

// Index among the interior edges (x-, y-, then z-edges, each in the full
// numbering's order) of the edge along axis t starting at node (i, j, k),
// or -1 on the coarse boundary.
template <int M>
__device__ __forceinline__ int rt3_iidx(int t, int i, int j, int k) {
  constexpr int S1 = M * (M - 1) * (M - 1);
  auto in = [](int q) { return q >= 1 && q <= M - 1; };
  if (t == 0) return in(j) && in(k) ? ((k - 1) * (M - 1) + (j - 1)) * M + i : -1;
  if (t == 1) return in(i) && in(k) ? S1 + ((k - 1) * M + j) * (M - 1) + (i - 1) : -1;
  return in(i) && in(j) ? 2 * S1 + (k * (M - 1) + (j - 1)) * (M - 1) + (i - 1) : -1;
}
// ... of local line Ln of fine cell (c0, c1, c2)
template <int M>
__device__ __forceinline__ int rt3_line_iidx(int c0, int c1, int c2, int Ln) {
  const int ax = kLAx[Ln], f1 = kLF1[Ln], f2 = kLF2[Ln];
  if (ax == 0) return rt3_iidx<M>(0, c0, c1 + f1, c2 + f2);
  if (ax == 1) return rt3_iidx<M>(1, c0 + f1, c1, c2 + f2);
  return rt3_iidx<M>(2, c0 + f1, c1 + f2, c2);
}

// The face mass with A^-1 at the 8 Gauss points of every fine cell
// ((psi_F, A^-1 psi_G): psi_F has only its axis component), 21 packed
// entries, entry-major over colour-major cells.
template <int L>
__global__ __launch_bounds__(256) void rt3_local_kernel(Rt3Args g) {
  constexpr int M = 1 << L, nfc = M * M * M;
  const int cell = blockIdx.x;
  V3 Hc, h;
  it_sizes(g.verts + (size_t)cell * 24, Hc, h, M);
  const double w8 = h.x * h.y * h.z * 0.125;
  double* mf = g.MF + (size_t)cell * 21 * nfc;
  for (int idx = threadIdx.x; idx < nfc * 21; idx += 256) {
    const int e = idx / nfc, c = idx - e * nfc;
    int F, G;
    tri_ij(e, F, G);
    const int comp = sym6(F >> 1, G >> 1);
    const double* a = g.ainv + ((size_t)cell * nfc + c) * 48;
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const V3 X = gauss_xyz(q);
      acc += rt_val(F, X, h) * rt_val(G, X, h) * a[q * 6 + comp];
    }
    mf[(size_t)e * nfc + it_lmc<M>(c % M, (c / M) % M, c / (M * M))] = acc * w8;
  }
}

// the inverse Jacobi diagonal of D^T M_A D on the interior edges
template <int L>
__global__ __launch_bounds__(256) void rt3_diag_kernel(Rt3Args g) {
  constexpr int M = 1 << L, nfc = M * M * M, ne = 3 * M * (M + 1) * (M + 1);
  const int cell = blockIdx.x;
  const double* mf = g.MF + (size_t)cell * 21 * nfc;
  double* pd = g.PD + (size_t)cell * g.ns;
  for (int e = threadIdx.x; e < ne; e += 256) {
    int t, p0, p1, p2;
    dof_pos<M>(e, t, p0, p1, p2);
    const int ie = rt3_iidx<M>(t, p0, p1, p2);
    if (ie < 0) continue;
    double d = 0.0;
    for (int k = 0; k < 4; ++k) {
      int c0, c1, c2, a;
      it_cell<M>(t, p0, p1, p2, k, c0, c1, c2, a);
      const int sc = it_lmc<M>(c0, c1, c2);
      for (int F = 0; F < 6; ++F)
        for (int G = 0; G < 6; ++G) {
          const int sF = rt3_inc(F, a), sG = rt3_inc(G, a);
          if (sF && sG) {
            const int hi = max(F, G), lo = min(F, G);
            d += sF * sG * mf[(size_t)(hi * (hi + 1) / 2 + lo) * nfc + sc];
          }
        }
    }
    pd[ie] = d > 0.0 ? 1.0 / d : 0.0;
  }
}

// For fine cell (c0, c1, c2): the circulations D_c p of its 6 faces from
// its 12 edge values (plus the coarse flux cf when given), z = M_A,c D_c p,
// and y += +-D_c^T z on its interior edges (sign sg); returns (D_c p) . z.
template <int M>
__device__ __forceinline__ double rt3_cell_apply(int c0, int c1, int c2, const double* p, double* y,
                                                 const double* mf, int nb, double sg) {
  constexpr int nfc = M * M * M;
  const unsigned sc = (unsigned)it_lmc<M>(c0, c1, c2);
  double circ[6];
  if (nb < 0) {
    double pv[12];
#pragma unroll
    for (int Ln = 0; Ln < 12; ++Ln) {
      const int ie = rt3_line_iidx<M>(c0, c1, c2, Ln);
      pv[Ln] = ie >= 0 ? p[ie] : 0.0;
    }
#pragma unroll
    for (int F = 0; F < 6; ++F) {
      double s = 0.0;
#pragma unroll
      for (int Ln = 0; Ln < 12; ++Ln) {
        const int sgn = rt3_inc(F, Ln);
        if (sgn) s += sgn * pv[Ln];
      }
      circ[F] = s;
    }
  } else {
#pragma unroll
    for (int F = 0; F < 6; ++F) circ[F] = rt3_coarse_flux<M>(c0, c1, c2, F, nb);
  }
  double z[6];
  double quad = 0.0;
#pragma unroll
  for (int F = 0; F < 6; ++F) {
    double s = 0.0;
#pragma unroll
    for (int G = 0; G < 6; ++G) {
      const int e = F >= G ? F * (F + 1) / 2 + G : G * (G + 1) / 2 + F;
      s += (mf + (size_t)e * nfc)[sc] * circ[G];
    }
    z[F] = s;
    quad += circ[F] * s;
  }
#pragma unroll
  for (int Ln = 0; Ln < 12; ++Ln) {
    const int ie = rt3_line_iidx<M>(c0, c1, c2, Ln);
    if (ie < 0) continue;
    double s = 0.0;
#pragma unroll
    for (int F = 0; F < 6; ++F) {
      const int sgn = rt3_inc(F, Ln);
      if (sgn) s += sgn * z[F];
    }
    y[ie] += sg * s;
  }
  return quad;
}

// One face basis of one coarse cell per workgroup: Jacobi-PCG on the
// curl-curl system (the product element by element in 8 colour passes, its
// quadratic form p . A p accumulated on the way), then the fine fluxes
// sigma = c_n + D psi on every face.
template <int L>
__global__ __launch_bounds__(kItNT) void rt3_cg_kernel(Rt3Args g) {
  constexpr int M = 1 << L, nfc = M * M * M;
  constexpr int nf1 = M * M * (M + 1), nf = 3 * nf1, H = M / 2;
  constexpr int ns = 3 * M * (M - 1) * (M - 1);
  __shared__ double redA[kItNW], redB[2 * kItNW];
  const int bid = blockIdx.x, slot = bid >> 3;
  const int cell = (slot / 6) * 8 + (bid & 7), nb = slot % 6;
  if (cell >= g.n_cells) return;
  const int tid = threadIdx.x;
  const double* mf = g.MF + (size_t)cell * 21 * nfc;
  const double* pd = g.PD + (size_t)cell * ns;
  double* x = g.WK + ((size_t)cell * 6 + nb) * 4 * ns;
  double* r = x + ns;
  double* p = x + 2 * ns;
  double* q = x + 3 * ns;
  for (int e = tid; e < ns; e += kItNT) {
    x[e] = 0.0;
    r[e] = 0.0;
    q[e] = 0.0;
  }
  __syncthreads();
  // r = b = -D^T M_A c_n
#pragma unroll 1
  for (int col = 0; col < 8; ++col) {
    for (int cc = tid; cc < H * H * H; cc += kItNT)
      rt3_cell_apply<M>(2 * (cc % H) + (col & 1), 2 * ((cc / H) % H) + ((col >> 1) & 1),
                        2 * (cc / (H * H)) + (col >> 2), nullptr, r, mf, nb, -1.0);
    __syncthreads();
  }
  double rz = 0.0, rr = 0.0;
  for (int e = tid; e < ns; e += kItNT) {
    const double re = r[e], z = pd[e] * re;
    p[e] = z;
    rz += re * z;
    rr += re * re;
  }
  rz = it_block_sum(rz, redA);
  const double bnorm = sqrt(it_block_sum(rr, redB));
  int flag = isfinite(bnorm) ? MSFEC_CELL_OK : MSFEC_CELL_NONFINITE;
  int itn = 0;
  if (flag == MSFEC_CELL_OK && bnorm > 0.0) {
    const double stop = g.rtol * bnorm;
    flag = MSFEC_CELL_NOT_CONVERGED;
    __syncthreads();
    while (itn < g.max_iter) {
      ++itn;
      // q = A p (q cleared by the previous pass); p . A p from the cells
      double pq = 0.0;
#pragma unroll 1
      for (int col = 0; col < 8; ++col) {
        for (int cc = tid; cc < H * H * H; cc += kItNT)
          pq += rt3_cell_apply<M>(2 * (cc % H) + (col & 1), 2 * ((cc / H) % H) + ((col >> 1) & 1),
                                  2 * (cc / (H * H)) + (col >> 2), p, q, mf, -1, 1.0);
        if (col < 7) __syncthreads();
      }
      pq = it_block_sum(pq, redA);
      if (!(pq > 0.0)) {
        flag = isfinite(pq) ? MSFEC_CELL_NOT_SPD : MSFEC_CELL_NONFINITE;
        break;
      }
      const double alpha = rz / pq;
      double rzn = 0.0;
      rr = 0.0;
      for (int e = tid; e < ns; e += kItNT) {
        x[e] += alpha * p[e];
        const double re = r[e] - alpha * q[e];
        r[e] = re;
        rzn += re * pd[e] * re;
        rr += re * re;
      }
      // both sums in one reduction
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        rzn += __shfl_xor(rzn, off);
        rr += __shfl_xor(rr, off);
      }
      if ((tid & 63) == 0) {
        redB[tid >> 6] = rzn;
        redB[kItNW + (tid >> 6)] = rr;
      }
      __syncthreads();
      rzn = 0.0;
      rr = 0.0;
#pragma unroll
      for (int w = 0; w < kItNW; ++w) {
        rzn += redB[w];
        rr += redB[kItNW + w];
      }
      rr = sqrt(rr);
      if (!isfinite(rr)) {
        flag = MSFEC_CELL_NONFINITE;
        break;
      }
      if (rr <= stop) {
        flag = MSFEC_CELL_OK;
        break;
      }
      const double beta = rzn / rz;
      rz = rzn;
      for (int e = tid; e < ns; e += kItNT) {
        p[e] = pd[e] * r[e] + beta * p[e];
        q[e] = 0.0;
      }
      __syncthreads();
    }
  }
  __syncthreads();
  // sigma = c_n + D psi on every face (a face's circulation from its 4 edges)
  double* sg = g.S + ((size_t)cell * 6 + nb) * nf;
  auto X = [&](int t, int i, int j, int k) {
    const int ie = rt3_iidx<M>(t, i, j, k);
    return ie >= 0 ? x[ie] : 0.0;
  };
  for (int f = tid; f < nf; f += kItNT) {
    int t, i, j, k;
    dof_pos<M>(3 * M * (M + 1) * (M + 1) + f, t, i, j, k);
    double circ, c = 0.0;
    const int d = t - 3, lev = d == 0 ? i : (d == 1 ? j : k);
    if (d == 0) circ = X(1, i, j, k) + X(2, i, j + 1, k) - X(1, i, j, k + 1) - X(2, i, j, k);
    else if (d == 1) circ = X(2, i, j, k) + X(0, i, j, k + 1) - X(2, i + 1, j, k) - X(0, i, j, k);
    else circ = X(0, i, j, k) + X(1, i + 1, j, k) - X(0, i, j + 1, k) - X(1, i, j, k);
    if (d == (nb >> 1)) c = lag(nb & 1, (double)lev / M) / (M * M);
    sg[f] = c + circ;
  }
  if (tid == 0) {
    g.its[(size_t)cell * 6 + nb] = itn;
    g.flg[(size_t)cell * 6 + nb] = flag;
  }
}

// Row i of the 7 x 7 element matrix of one coarse cell per workgroup
// (assemble_global_element_matrix, rt_dq_basis.cc:822-914, with u = 1):
// the flux block sigma_i^T M_A sigma_j from the fine cells' face masses,
// the divergence entries -+ the net outflow, b_6 the integral of f.
template <int L>
__global__ __launch_bounds__(256) void rt3_element_kernel(Rt3Args g) {
  constexpr int M = 1 << L, nfc = M * M * M, nf = 3 * M * M * (M + 1);
  __shared__ double red[4][8];
  const int cell = blockIdx.x / 7, i = blockIdx.x % 7, tid = threadIdx.x;
  const double* mf = g.MF + (size_t)cell * 21 * nfc;
  const double* S = g.S + (size_t)cell * 6 * nf;
  double acc[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) acc[j] = 0.0;
  for (int c = tid; c < nfc; c += 256) {
    const int c0 = c % M, c1 = (c / M) % M, c2 = c / (M * M);
    const int sc = it_lmc<M>(c0, c1, c2);
    int fi[6];
#pragma unroll
    for (int F = 0; F < 6; ++F) fi[F] = rt3_face<M>(c0, c1, c2, F);
    if (i < 6) {
      double si[6];
#pragma unroll
      for (int F = 0; F < 6; ++F) si[F] = S[(size_t)i * nf + fi[F]];
      double ms[6];
#pragma unroll
      for (int F = 0; F < 6; ++F) {
        double s = 0.0;
#pragma unroll
        for (int G = 0; G < 6; ++G) {
          const int e = F >= G ? F * (F + 1) / 2 + G : G * (G + 1) / 2 + F;
          s += mf[(size_t)e * nfc + sc] * si[G];
        }
        ms[F] = s;
      }
#pragma unroll
      for (int j = 0; j < 6; ++j) {
        double s = 0.0;
#pragma unroll
        for (int F = 0; F < 6; ++F) s += ms[F] * S[(size_t)j * nf + fi[F]];
        acc[j] += s;
      }
      // net outflow of sigma_i from this fine cell
#pragma unroll
      for (int F = 0; F < 6; ++F) acc[6] += (F & 1 ? 1.0 : -1.0) * si[F];
    } else if (g.fq) {
      V3 Hc, h;
      it_sizes(g.verts + (size_t)cell * 24, Hc, h, M);
      const double* f = g.fq + ((size_t)cell * nfc + c) * 8;
      double s = 0.0;
#pragma unroll
      for (int q = 0; q < 8; ++q) s += f[q];
      acc[7] += s * h.x * h.y * h.z * 0.125;
    }
  }
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    double s = acc[j];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((tid & 63) == 0) red[tid >> 6][j] = s;
  }
  __syncthreads();
  if (tid == 0) {
    double t[8];
    for (int j = 0; j < 8; ++j) t[j] = red[0][j] + red[1][j] + red[2][j] + red[3][j];
    double* Mc = g.M + (size_t)cell * 49;
    double* bc = g.b + (size_t)cell * 7;
    if (i < 6) {
      for (int j = 0; j < 6; ++j) Mc[i * 7 + j] = t[j];
      Mc[i * 7 + 6] = -t[6];  // -(div sigma_i, 1)
      Mc[6 * 7 + i] = t[6];   // (1, div sigma_i)
      bc[i] = 0.0;
    } else {
      Mc[6 * 7 + 6] = 0.0;    // (R u, v) with R = 0
      bc[6] = t[7];
    }
  }
}

__global__ __launch_bounds__(256) void rt3_status_kernel(Rt3Args g) {
  const int cell = blockIdx.x * 256 + threadIdx.x;
  if (cell >= g.n_cells) return;
  int st = MSFEC_CELL_OK;
  for (int n = 0; n < 6; ++n) {
    const int f = g.flg[(size_t)cell * 6 + n];
    if (f != MSFEC_CELL_OK && st == MSFEC_CELL_OK) st = f;
  }
  for (int k = 0; k < 49 && st == MSFEC_CELL_OK; ++k)
    if (!isfinite(g.M[(size_t)cell * 49 + k])) st = MSFEC_CELL_NONFINITE;
  g.status[cell] = st;
}

msfec_status nrt_fail(msfec_status st, const std::string& msg) {
  msfec_detail::g_last_error = msg;
  return st;
}

#define NRT_HIP(call)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return nrt_fail(MSFEC_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// The fine dof tables of one coarse cell (they depend on l only).
struct NrtTables {
  int m = 0, nx = 0, ne = 0, nf1 = 0, N = 0, n = 0, ns = 0;
  std::vector<int> icell, cdof, f2i;
};

NrtTables nrt_tables(int l) {
  NrtTables t;
  const int m = 1 << l;
  t.m = m;
  t.nx = m * (m + 1) * (m + 1);
  t.ne = 3 * t.nx;
  t.nf1 = m * m * (m + 1);
  t.N = t.ne + 3 * t.nf1;
  auto ex = [&](int i, int j, int k) { return (k * (m + 1) + j) * m + i; };
  auto ey = [&](int i, int j, int k) { return t.nx + (k * m + j) * (m + 1) + i; };
  auto ez = [&](int i, int j, int k) { return 2 * t.nx + (k * (m + 1) + j) * (m + 1) + i; };
  auto fx = [&](int i, int j, int k) { return t.ne + (k * m + j) * (m + 1) + i; };
  auto fy = [&](int i, int j, int k) { return t.ne + t.nf1 + (k * (m + 1) + j) * m + i; };
  auto fz = [&](int i, int j, int k) { return t.ne + 2 * t.nf1 + (k * m + j) * m + i; };
  static const int ax[12] = {1, 1, 0, 0, 1, 1, 0, 0, 2, 2, 2, 2};
  static const int f1[12] = {0, 1, 0, 1, 0, 1, 0, 1, 0, 1, 0, 1};
  static const int f2[12] = {0, 0, 0, 0, 1, 1, 1, 1, 0, 0, 1, 1};
  const int nfc = m * m * m;
  t.cdof.assign((size_t)nfc * 18, 0);
  for (int c3 = 0; c3 < m; ++c3)
    for (int b = 0; b < m; ++b)
      for (int a = 0; a < m; ++a) {
        int* d = &t.cdof[(size_t)((c3 * m + b) * m + a) * 18];
        for (int L = 0; L < 12; ++L) {
          if (ax[L] == 0) d[L] = ex(a, b + f1[L], c3 + f2[L]);
          else if (ax[L] == 1) d[L] = ey(a + f1[L], b, c3 + f2[L]);
          else d[L] = ez(a + f1[L], b + f2[L], c3);
        }
        d[12] = fx(a, b, c3);
        d[13] = fx(a + 1, b, c3);
        d[14] = fy(a, b, c3);
        d[15] = fy(a, b + 1, c3);
        d[16] = fz(a, b, c3);
        d[17] = fz(a, b, c3 + 1);
      }
  // position of every full dof: type (0..2 edges along x/y/z, 3..5 faces
  // normal to x/y/z) and its integer coordinates; interior dofs are numbered
  // in full order (edges first)
  t.f2i.assign(t.N, -1);
  for (int f = 0; f < t.N; ++f) {
    int ty, p[3];
    if (f < t.ne) {
      ty = f / t.nx;
      int r = f % t.nx;
      int ext[3] = {m + 1, m + 1, m + 1};
      ext[ty] = m;
      p[0] = r % ext[0];
      r /= ext[0];
      p[1] = r % ext[1];
      p[2] = r / ext[1];
    } else {
      ty = 3 + (f - t.ne) / t.nf1;
      int r = (f - t.ne) % t.nf1;
      int ext[3] = {m, m, m};
      ext[ty - 3] = m + 1;
      p[0] = r % ext[0];
      r /= ext[0];
      p[1] = r % ext[1];
      p[2] = r / ext[1];
    }
    bool bnd;
    if (ty < 3) {
      bnd = false;
      for (int d = 0; d < 3; ++d)
        if (d != ty && (p[d] == 0 || p[d] == m)) bnd = true;
    } else {
      bnd = p[ty - 3] == 0 || p[ty - 3] == m;
    }
    if (!bnd) t.f2i[f] = t.n++;
  }
  t.ns = 0;
  for (int f = 0; f < t.ne; ++f)
    if (t.f2i[f] >= 0) ++t.ns;
  t.icell.assign((size_t)std::max(t.n, 1) * 4, -1);
  std::vector<int> cnt(t.n, 0);
  for (int c = 0; c < nfc; ++c)
    for (int a = 0; a < 18; ++a) {
      const int I = t.f2i[t.cdof[(size_t)c * 18 + a]];
      if (I < 0) continue;
      t.icell[(size_t)I * 4 + cnt[I]++] = c * 32 + a;
    }
  return t;
}

// The assembly map of nrt_tile: for lower tile (P, Q), register gg and lane
// (ll, lh) -- the entry (I, J) = (16 P + 4 gg + lh, 16 Q + ll) of K' --
// the LDS offsets of the local-matrix entries of the fine cells that I and J
// share, with the sign of K' (every block but the edge block negated).
template <int L>
void nrt_amap(const NrtTables& t, std::vector<unsigned short>& amap, std::vector<unsigned char>& tnz) {
  typedef Nd<L> D;
  const int ntiles = D::TN * (D::TN + 1) / 2;
  amap.assign((size_t)ntiles * 4 * 64 * 4, 0xffff);
  tnz.assign(ntiles, 0);
  auto tri = [](int i, int j) { const int hi = std::max(i, j), lo = std::min(i, j); return hi * (hi + 1) / 2 + lo; };
  for (int P = 0; P < D::TN; ++P)
    for (int Q = 0; Q <= P; ++Q) {
      const int tix = P * (P + 1) / 2 + Q;
      for (int gg = 0; gg < 4; ++gg)
        for (int lane = 0; lane < 64; ++lane) {
          const int I = 16 * P + 4 * gg + (lane >> 4), J = 16 * Q + (lane & 15);
          if (I >= t.n || J >= t.n) continue;
          unsigned short* out = &amap[(((size_t)tix * 4 + gg) * 64 + lane) * 4];
          int cnt = 0;
          for (int a = 0; a < 4; ++a) {
            const int ei = t.icell[(size_t)I * 4 + a];
            if (ei < 0) continue;
            for (int b = 0; b < 4; ++b) {
              const int ej = t.icell[(size_t)J * 4 + b];
              if (ej < 0 || (ej >> 5) != (ei >> 5)) continue;
              const int c = ei >> 5, li = ei & 31, lj = ej & 31;
              int off;
              if (li < 12 && lj < 12) off = D::oMl + c * 78 + tri(li, lj);
              else if (li >= 12 && lj >= 12) off = D::oDl + c * 21 + tri(li - 12, lj - 12);
              else off = D::oCc + (std::max(li, lj) - 12) * 12 + std::min(li, lj);
              const bool neg = !(li < 12 && lj < 12);
              out[cnt++] = (unsigned short)(off | (neg ? 0x8000 : 0));
              tnz[tix] = 1;
            }
          }
        }
    }
}

}  // namespace

struct msfec_nrt_batch {
  msfec_nrt_params prm{};
  int n_cells = 0, l = 0;
  NrtTables tab;
  double* d_verts = nullptr;
  double* d_ainv = nullptr;
  double* d_b = nullptr;
  double* d_r = nullptr;
  double* d_f = nullptr;
  int* d_icell = nullptr;
  int* d_cdof = nullptr;
  int* d_f2i = nullptr;
  unsigned short* d_amap = nullptr;
  unsigned char* d_tnz = nullptr;
  double* d_X = nullptr;
  double* d_M = nullptr;
  double* d_rhs = nullptr;
  double* d_w = nullptr;
  int* d_status = nullptr;
  // MINRES path
  bool iter = false;
  int* d_idof = nullptr;
  double* d_LM = nullptr;
  double* d_CC = nullptr;
  double* d_FL = nullptr;
  double* d_PD = nullptr;
  double* d_WK = nullptr;
  int* d_its = nullptr;
  int* d_flg = nullptr;
  double* d_S0 = nullptr;
  int* d_agg = nullptr;
  double* d_EI = nullptr;
  bool coef_set = false, has_r = false, has_f = false, computed = false;
  hipStream_t stream = nullptr;
  // per timed compute: start, after the local matrices and preconditioner,
  // after MINRES, end (the middle two only on the MINRES path)
  std::vector<std::array<hipEvent_t, 4>> ev_sets;
  size_t ev_used = 0;
};

namespace {

template <typename T>
msfec_status nrt_alloc(T** p, size_t count) {
  NRT_HIP(hipMalloc((void**)p, std::max<size_t>(count, 1) * sizeof(T)));
  return MSFEC_OK;
}

void nrt_free(msfec_nrt_batch* h) {
  void* ptrs[] = {h->d_verts, h->d_ainv, h->d_b,  h->d_r,   h->d_f,  h->d_icell, h->d_cdof,
                  h->d_f2i,   h->d_amap, h->d_tnz, h->d_X,  h->d_M,  h->d_rhs,   h->d_w,
                  h->d_status, h->d_idof, h->d_LM, h->d_CC, h->d_FL, h->d_PD,    h->d_WK,
                  h->d_its,   h->d_flg,  h->d_S0, h->d_agg, h->d_EI};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (auto& set : h->ev_sets)
    for (auto& e : set)
      if (e) (void)hipEventDestroy(e);
  h->ev_sets.clear();
}

// Copied on the handle's stream (that of its last compute call), waited for.
msfec_status nrt_in(hipStream_t s, void* dst, const void* src, size_t bytes, int loc) {
  NRT_HIP(hipMemcpyAsync(dst, src, bytes,
                         loc == MSFEC_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
  NRT_HIP(hipStreamSynchronize(s));
  return MSFEC_OK;
}

msfec_status nrt_out(hipStream_t s, double* dst, const double* src, size_t count, int loc) {
  if (!dst) return nrt_fail(MSFEC_ERR_INVALID, "null output pointer");
  if (loc != MSFEC_HOST && loc != MSFEC_DEVICE) return nrt_fail(MSFEC_ERR_INVALID, "bad location");
  NRT_HIP(hipMemcpyAsync(dst, src, count * sizeof(double),
                         loc == MSFEC_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
  NRT_HIP(hipStreamSynchronize(s));
  return MSFEC_OK;
}

template <int L>
msfec_status nrt_launch(msfec_nrt_batch* h, hipStream_t s) {
  typedef Nd<L> D;
  const size_t lds = (size_t)D::total * sizeof(double);
  NRT_HIP(hipFuncSetAttribute((const void*)nrt_basis_kernel<L>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  NrtArgs a{};
  a.verts = h->d_verts;
  a.ainv = h->d_ainv;
  a.bq = h->d_b;
  a.rq = h->has_r ? h->d_r : nullptr;
  a.fq = h->has_f ? h->d_f : nullptr;
  a.icell = h->d_icell;
  a.cdof = h->d_cdof;
  a.f2i = h->d_f2i;
  a.amap = h->d_amap;
  a.tnz = h->d_tnz;
  a.X = h->d_X;
  a.M = h->d_M;
  a.b = h->d_rhs;
  a.status = h->d_status;
  a.N = h->tab.N;
  hipLaunchKernelGGL(nrt_basis_kernel<L>, dim3(h->n_cells), dim3(D::NT), lds, s, a);
  NRT_HIP(hipGetLastError());
  return MSFEC_OK;
}

template <int L>
msfec_status nrt_it_launch(msfec_nrt_batch* h, hipStream_t s, std::array<hipEvent_t, 4>* ev) {
  NrtItArgs a{};
  a.verts = h->d_verts;
  a.ainv = h->d_ainv;
  a.bq = h->d_b;
  a.rq = h->has_r ? h->d_r : nullptr;
  a.fq = h->has_f ? h->d_f : nullptr;
  a.idof = h->d_idof;
  a.LM = h->d_LM;
  a.CC = h->d_CC;
  a.FL = h->d_FL;
  a.PD = h->d_PD;
  a.WK = h->d_WK;
  a.X = h->d_X;
  a.M = h->d_M;
  a.b = h->d_rhs;
  a.its = h->d_its;
  a.flg = h->d_flg;
  a.S0 = h->d_S0;
  a.agg = h->d_agg;
  a.EI = h->d_EI;
  a.status = h->d_status;
  a.n_cells = h->n_cells;
  a.N = h->tab.N;
  a.n = h->tab.n;
  a.max_iter = h->prm.max_iter > 0 ? h->prm.max_iter : 20000;
  a.rtol = h->prm.rtol > 0.0 ? h->prm.rtol : 1e-14;
  const int nc = h->n_cells;
  hipLaunchKernelGGL(nrt_it_local_kernel<L>, dim3(nc), dim3(256), 0, s, a);
  NRT_HIP(hipGetLastError());
  hipLaunchKernelGGL(nrt_it_diag_kernel<L>, dim3(nc), dim3(256), 0, s, a);
  NRT_HIP(hipGetLastError());
  {
    constexpr int M = 1 << L, NC = ItCoarse<M>::NC;
    const size_t lds = (size_t)3 * M * M * (M + 1) * sizeof(double);
    NRT_HIP(hipFuncSetAttribute((const void*)nrt_it_coarse_kernel<L>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(nrt_it_coarse_kernel<L>, dim3(nc * NC), dim3(256), lds, s, a);
    NRT_HIP(hipGetLastError());
    hipLaunchKernelGGL(nrt_it_coarse_inv_kernel<L>, dim3(nc), dim3(1024), 0, s, a);
    NRT_HIP(hipGetLastError());
    constexpr int nfc = M * M * M, CPB = nfc < 64 ? nfc : 64;
    if (kItEdgeSchwarz<L>) {
      hipLaunchKernelGGL(nrt_it_edgeas_kernel<L>, dim3(nc * (nfc / CPB)), dim3(64), 0, s, a);
      NRT_HIP(hipGetLastError());
    }
  }
  if (ev) NRT_HIP(hipEventRecord((*ev)[1], s));
  // cells in groups of 8 (one per XCD), 18 bases each
  const int groups = (nc + 7) / 8;
  hipLaunchKernelGGL(nrt_minres_kernel<L>, dim3(groups * 8 * kNB), dim3(kItNT), 0, s, a);
  NRT_HIP(hipGetLastError());
  if (ev) NRT_HIP(hipEventRecord((*ev)[2], s));
  hipLaunchKernelGGL(nrt_it_element_kernel<L>, dim3(nc * kNB), dim3(256), 0, s, a);
  NRT_HIP(hipGetLastError());
  hipLaunchKernelGGL(nrt_it_status_kernel, dim3((nc + 255) / 256), dim3(256), 0, s, a);
  NRT_HIP(hipGetLastError());
  return MSFEC_OK;
}

}  // namespace

extern "C" {

msfec_status msfec_nrt_create(const msfec_nrt_params* params, int n_cells, const double* vertices,
                              int vertices_loc, msfec_nrt_handle* out) {
  if (!params || !vertices || !out) return nrt_fail(MSFEC_ERR_INVALID, "null argument");
  *out = nullptr;
  if (n_cells <= 0) return nrt_fail(MSFEC_ERR_INVALID, "n_cells must be positive");
  const int l = params->n_refine_local;
  if (l < 1) return nrt_fail(MSFEC_ERR_INVALID, "n_refine_local must be at least 1");
  if (l > kItMaxL) return nrt_fail(MSFEC_ERR_UNSUPPORTED, "n_refine_local above 4");
  if (params->solver != 0 && params->solver != MSFEC_SOLVER_CG)
    return nrt_fail(MSFEC_ERR_INVALID, "solver must be 0 or MSFEC_SOLVER_CG");
  if (params->max_iter < 0 || !(params->rtol >= 0.0))
    return nrt_fail(MSFEC_ERR_INVALID, "max_iter and rtol must not be negative");
  if (vertices_loc != MSFEC_HOST && vertices_loc != MSFEC_DEVICE)
    return nrt_fail(MSFEC_ERR_INVALID, "bad vertices location");
  std::vector<double> hv((size_t)n_cells * 24);
  if (vertices_loc == MSFEC_DEVICE) {
    NRT_HIP(hipMemcpy(hv.data(), vertices, hv.size() * sizeof(double), hipMemcpyDeviceToHost));
  } else {
    std::copy(vertices, vertices + hv.size(), hv.begin());
  }
  for (int c = 0; c < n_cells; ++c) {
    const double* v = &hv[(size_t)c * 24];
    bool box = true;
    for (int d = 0; d < 3; ++d) {
      const double Hd = v[3 * (1 << d) + d] - v[d];
      if (!(Hd > 0)) box = false;
      for (int k = 0; k < 8 && box; ++k)
        if (std::fabs(v[3 * k + d] - (v[d] + ((k >> d) & 1) * Hd)) > 1e-14 * Hd) box = false;
    }
    if (!box)
      return nrt_fail(MSFEC_ERR_INVALID,
                      "cell " + std::to_string(c) + " is not an axis-aligned box in deal.II vertex order");
  }
  msfec_nrt_batch* h = new msfec_nrt_batch();
  h->prm = *params;
  h->n_cells = n_cells;
  h->l = l;
  h->tab = nrt_tables(l);
  h->iter = l > kMaxL || params->solver == MSFEC_SOLVER_CG;
  const size_t nc = (size_t)n_cells, nfc = (size_t)1 << (3 * l);
  msfec_status st = MSFEC_OK;
#define NRT_TRY(x)        \
  do {                    \
    st = (x);             \
    if (st != MSFEC_OK) { \
      nrt_free(h);        \
      delete h;           \
      return st;          \
    }                     \
  } while (0)
  NRT_TRY(nrt_alloc(&h->d_verts, nc * 24));
  NRT_TRY(nrt_alloc(&h->d_ainv, nc * nfc * 8 * 6));
  NRT_TRY(nrt_alloc(&h->d_b, nc * nfc * 8));
  NRT_TRY(nrt_alloc(&h->d_r, nc * nfc * 8));
  NRT_TRY(nrt_alloc(&h->d_f, nc * nfc * 8 * 3));
  NRT_TRY(nrt_alloc(&h->d_icell, h->tab.icell.size()));
  NRT_TRY(nrt_alloc(&h->d_cdof, h->tab.cdof.size()));
  NRT_TRY(nrt_alloc(&h->d_f2i, h->tab.f2i.size()));
  NRT_TRY(nrt_alloc(&h->d_X, nc * kNB * h->tab.N));
  NRT_TRY(nrt_alloc(&h->d_M, nc * kNB * kNB));
  NRT_TRY(nrt_alloc(&h->d_rhs, nc * kNB));
  NRT_TRY(nrt_alloc(&h->d_w, nc * kNB));
  NRT_TRY(nrt_alloc(&h->d_status, nc));
  NRT_TRY(nrt_in(h->stream, h->d_verts, hv.data(), hv.size() * sizeof(double), MSFEC_HOST));
  NRT_TRY(nrt_in(h->stream, h->d_icell, h->tab.icell.data(), h->tab.icell.size() * sizeof(int), MSFEC_HOST));
  NRT_TRY(nrt_in(h->stream, h->d_cdof, h->tab.cdof.data(), h->tab.cdof.size() * sizeof(int), MSFEC_HOST));
  NRT_TRY(nrt_in(h->stream, h->d_f2i, h->tab.f2i.data(), h->tab.f2i.size() * sizeof(int), MSFEC_HOST));
  if (h->iter) {
    const size_t N = (size_t)h->tab.N;
    std::vector<int> idof;
    for (int f = 0; f < h->tab.N; ++f)
      if (h->tab.f2i[f] >= 0) idof.push_back(f);
    NRT_TRY(nrt_alloc(&h->d_idof, idof.size()));
    NRT_TRY(nrt_in(h->stream, h->d_idof, idof.data(), idof.size() * sizeof(int), MSFEC_HOST));
    NRT_TRY(nrt_alloc(&h->d_LM, nc * kLM * nfc));
    NRT_TRY(nrt_alloc(&h->d_CC, nc * 72));
    NRT_TRY(nrt_alloc(&h->d_FL, nc * nfc * 6));
    NRT_TRY(nrt_alloc(&h->d_PD, nc * N));
    NRT_TRY(nrt_alloc(&h->d_WK, nc * kNB * 6 * N));
    NRT_TRY(nrt_alloc(&h->d_EI, nc * 78 * nfc));
    NRT_TRY(nrt_alloc(&h->d_its, nc * kNB));
    NRT_TRY(nrt_alloc(&h->d_flg, nc * kNB));
    // the coarse space: aggregate of every interior face dof (-1 for edges)
    const int m = 1 << l, NB = m <= 4 ? m / 2 : 4, B = m / NB, NC = 3 * NB * NB * NB;
    std::vector<int> agg;
    for (int f : idof) {
      if (f < h->tab.ne) {
        agg.push_back(-1);
        continue;
      }
      const int d = (f - h->tab.ne) / h->tab.nf1;
      int r = (f - h->tab.ne) % h->tab.nf1;
      int ext[3] = {m, m, m};
      ext[d] = m + 1;
      int q[3];
      q[0] = r % ext[0];
      r /= ext[0];
      q[1] = r % ext[1];
      q[2] = r / ext[1];
      int a[3];
      for (int k = 0; k < 3; ++k) a[k] = std::min(q[k] / B, NB - 1);
      agg.push_back(d * NB * NB * NB + (a[2] * NB + a[1]) * NB + a[0]);
    }
    NRT_TRY(nrt_alloc(&h->d_agg, agg.size()));
    NRT_TRY(nrt_in(h->stream, h->d_agg, agg.data(), agg.size() * sizeof(int), MSFEC_HOST));
    NRT_TRY(nrt_alloc(&h->d_S0, nc * NC * NC));
  } else {
    std::vector<unsigned short> amap;
    std::vector<unsigned char> tnz;
    if (l == 1) nrt_amap<1>(h->tab, amap, tnz);
    else nrt_amap<2>(h->tab, amap, tnz);
    NRT_TRY(nrt_alloc(&h->d_amap, amap.size()));
    NRT_TRY(nrt_alloc(&h->d_tnz, tnz.size()));
    NRT_TRY(nrt_in(h->stream, h->d_amap, amap.data(), amap.size() * sizeof(unsigned short), MSFEC_HOST));
    NRT_TRY(nrt_in(h->stream, h->d_tnz, tnz.data(), tnz.size(), MSFEC_HOST));
  }
#undef NRT_TRY
  *out = h;
  return MSFEC_OK;
}

msfec_status msfec_nrt_set_coefficients(msfec_nrt_handle h, const double* ainv, const double* b,
                                        const double* r, const double* f, int loc) {
  if (!h || !ainv || !b) return nrt_fail(MSFEC_ERR_INVALID, "null argument");
  if (loc != MSFEC_HOST && loc != MSFEC_DEVICE) return nrt_fail(MSFEC_ERR_INVALID, "bad location");
  const size_t pts = (size_t)h->n_cells * ((size_t)1 << (3 * h->l)) * 8;
  msfec_status st = nrt_in(h->stream, h->d_ainv, ainv, pts * 6 * sizeof(double), loc);
  if (st == MSFEC_OK) st = nrt_in(h->stream, h->d_b, b, pts * sizeof(double), loc);
  if (st == MSFEC_OK && r) st = nrt_in(h->stream, h->d_r, r, pts * sizeof(double), loc);
  if (st == MSFEC_OK && f) st = nrt_in(h->stream, h->d_f, f, pts * 3 * sizeof(double), loc);
  if (st != MSFEC_OK) return st;
  h->has_r = r != nullptr;
  h->has_f = f != nullptr;
  h->coef_set = true;
  return MSFEC_OK;
}

msfec_status msfec_nrt_compute(msfec_nrt_handle h, void* stream) {
  if (!h) return nrt_fail(MSFEC_ERR_INVALID, "null handle");
  if (!h->coef_set) return nrt_fail(MSFEC_ERR_STATE, "coefficients not set");
  hipStream_t s = (hipStream_t)stream;
  h->stream = s;
  std::array<hipEvent_t, 4>* ev = nullptr;
  if (h->prm.flags & MSFEC_FLAG_TIMING) {
    if (h->ev_used == h->ev_sets.size()) {
      std::array<hipEvent_t, 4> set{};
      for (auto& e : set) NRT_HIP(hipEventCreate(&e));
      h->ev_sets.push_back(set);
    }
    ev = &h->ev_sets[h->ev_used++];
    NRT_HIP(hipEventRecord((*ev)[0], s));
  }
  msfec_status st;
  if (h->iter) {
    switch (h->l) {
      case 1: st = nrt_it_launch<1>(h, s, ev); break;
      case 2: st = nrt_it_launch<2>(h, s, ev); break;
      case 3: st = nrt_it_launch<3>(h, s, ev); break;
      default: st = nrt_it_launch<4>(h, s, ev); break;
    }
  } else {
    st = h->l == 1 ? nrt_launch<1>(h, s) : nrt_launch<2>(h, s);
  }
  if (st != MSFEC_OK) return st;
  if (ev) {
    if (!h->iter) {
      NRT_HIP(hipEventRecord((*ev)[1], s));
      NRT_HIP(hipEventRecord((*ev)[2], s));
    }
    NRT_HIP(hipEventRecord((*ev)[3], s));
  }
  h->computed = true;
  return MSFEC_OK;
}

msfec_status msfec_nrt_synchronize(msfec_nrt_handle h) {
  if (!h) return nrt_fail(MSFEC_ERR_INVALID, "null handle");
  NRT_HIP(hipStreamSynchronize(h->stream));
  return MSFEC_OK;
}

msfec_status msfec_nrt_get_element(msfec_nrt_handle h, double* M, double* b, int loc) {
  if (!h) return nrt_fail(MSFEC_ERR_INVALID, "null handle");
  if (!h->computed) return nrt_fail(MSFEC_ERR_STATE, "compute not called");
  msfec_status st = MSFEC_OK;
  if (M) st = nrt_out(h->stream, M, h->d_M, (size_t)h->n_cells * kNB * kNB, loc);
  if (st == MSFEC_OK && b) st = nrt_out(h->stream, b, h->d_rhs, (size_t)h->n_cells * kNB, loc);
  return st;
}

msfec_status msfec_nrt_get_basis(msfec_nrt_handle h, double* X, int loc) {
  if (!h) return nrt_fail(MSFEC_ERR_INVALID, "null handle");
  if (!h->computed) return nrt_fail(MSFEC_ERR_STATE, "compute not called");
  return nrt_out(h->stream, X, h->d_X, (size_t)h->n_cells * kNB * h->tab.N, loc);
}

msfec_status msfec_nrt_reconstruct(msfec_nrt_handle h, const double* w, int w_loc, double* sigma,
                                   double* u, int out_loc) {
  if (!h || !w) return nrt_fail(MSFEC_ERR_INVALID, "null argument");
  if (!sigma && !u) return nrt_fail(MSFEC_ERR_INVALID, "no output requested");
  if (w_loc != MSFEC_HOST && w_loc != MSFEC_DEVICE) return nrt_fail(MSFEC_ERR_INVALID, "bad location");
  if (out_loc != MSFEC_HOST && out_loc != MSFEC_DEVICE)
    return nrt_fail(MSFEC_ERR_INVALID, "bad location");
  if (!h->computed) return nrt_fail(MSFEC_ERR_STATE, "compute not called");
  msfec_status st = nrt_in(h->stream, h->d_w, w, (size_t)h->n_cells * kNB * sizeof(double), w_loc);
  if (st != MSFEC_OK) return st;
  const int ne = h->tab.ne, nf = h->tab.N - h->tab.ne;
  const size_t ns = (size_t)h->n_cells * ne, nu = (size_t)h->n_cells * nf;
  double* ds = nullptr;
  double* du = nullptr;
  // device outputs are written in place; host outputs go through staging
  // buffers freed before returning
  if (sigma) {
    if (out_loc == MSFEC_DEVICE) ds = sigma;
    else NRT_HIP(hipMalloc((void**)&ds, ns * sizeof(double)));
  }
  if (u) {
    if (out_loc == MSFEC_DEVICE) {
      du = u;
    } else if (hipMalloc((void**)&du, nu * sizeof(double)) != hipSuccess) {
      if (ds && out_loc != MSFEC_DEVICE) (void)hipFree(ds);
      return nrt_fail(MSFEC_ERR_HIP, "hipMalloc of the reconstruction staging buffer failed");
    }
  }
  hipLaunchKernelGGL(nrt_reconstruct_kernel, dim3(h->n_cells), dim3(256), 0, h->stream, h->d_X,
                     h->d_w, h->tab.N, ne, ds, du);
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
  if (e != hipSuccess) return nrt_fail(MSFEC_ERR_HIP, std::string("reconstruct: ") + hipGetErrorString(e));
  return MSFEC_OK;
}

msfec_status msfec_nrt_get_cell_status(msfec_nrt_handle h, int* status) {
  if (!h || !status) return nrt_fail(MSFEC_ERR_INVALID, "null argument");
  if (!h->computed) return nrt_fail(MSFEC_ERR_STATE, "compute not called");
  NRT_HIP(hipStreamSynchronize(h->stream));
  NRT_HIP(hipMemcpy(status, h->d_status, (size_t)h->n_cells * sizeof(int), hipMemcpyDeviceToHost));
  return MSFEC_OK;
}

msfec_status msfec_nrt_get_iterations(msfec_nrt_handle h, int* iterations) {
  if (!h || !iterations) return nrt_fail(MSFEC_ERR_INVALID, "null argument");
  if (!h->iter) return nrt_fail(MSFEC_ERR_STATE, "the handle solves directly (no MINRES iterations)");
  if (!h->computed) return nrt_fail(MSFEC_ERR_STATE, "compute not called");
  NRT_HIP(hipStreamSynchronize(h->stream));
  NRT_HIP(hipMemcpy(iterations, h->d_its, (size_t)h->n_cells * kNB * sizeof(int), hipMemcpyDeviceToHost));
  return MSFEC_OK;
}

msfec_status msfec_nrt_kernel_times(msfec_nrt_handle h, float* ms, int max, int* n) {
  if (!h || !ms) return nrt_fail(MSFEC_ERR_INVALID, "null argument");
  if (h->ev_used == 0) return nrt_fail(MSFEC_ERR_STATE, "no timed compute (set MSFEC_FLAG_TIMING)");
  // the whole compute, then (MINRES path) local matrices + preconditioner,
  // MINRES, element matrix + status
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (size_t c = 0; c < h->ev_used; ++c) {
    const auto& e = h->ev_sets[c];
    NRT_HIP(hipEventSynchronize(e[3]));
    float t[4] = {0.f, 0.f, 0.f, 0.f};
    NRT_HIP(hipEventElapsedTime(&t[0], e[0], e[3]));
    NRT_HIP(hipEventElapsedTime(&t[1], e[0], e[1]));
    NRT_HIP(hipEventElapsedTime(&t[2], e[1], e[2]));
    NRT_HIP(hipEventElapsedTime(&t[3], e[2], e[3]));
    for (int k = 0; k < 4; ++k) acc[k] += t[k];
  }
  const int cnt = std::min(max, h->iter ? 4 : 1);
  for (int k = 0; k < cnt; ++k) ms[k] = (float)(acc[k] / (double)h->ev_used);
  if (n) *n = std::max(cnt, 0);
  h->ev_used = 0;
  return MSFEC_OK;
}

msfec_status msfec_nrt_destroy(msfec_nrt_handle h) {
  if (!h) return MSFEC_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  nrt_free(h);
  delete h;
  return MSFEC_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Host side of the 3-D RT_DQ basis (msfec_rt3_*).

struct msfec_rt3_batch {
  msfec_rt3_params prm{};
  int n_cells = 0, l = 0, ne = 0, nf = 0, ns = 0;
  double* d_verts = nullptr;
  double* d_ainv = nullptr;
  double* d_f = nullptr;
  double* d_MF = nullptr;
  double* d_PD = nullptr;
  double* d_WK = nullptr;
  double* d_S = nullptr;
  double* d_M = nullptr;
  double* d_b = nullptr;
  int* d_its = nullptr;
  int* d_flg = nullptr;
  int* d_status = nullptr;
  bool coef_set = false, has_f = false, computed = false;
  hipStream_t stream = nullptr;
  std::vector<std::array<hipEvent_t, 4>> ev_sets;
  size_t ev_used = 0;
};

namespace {

void rt3_free(msfec_rt3_batch* h) {
  void* ptrs[] = {h->d_verts, h->d_ainv, h->d_f,   h->d_MF,  h->d_PD,    h->d_WK,
                  h->d_S,     h->d_M,    h->d_b,   h->d_its, h->d_flg, h->d_status};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (auto& set : h->ev_sets)
    for (auto& e : set)
      if (e) (void)hipEventDestroy(e);
  h->ev_sets.clear();
}

template <int L>
msfec_status rt3_launch(msfec_rt3_batch* h, hipStream_t s, std::array<hipEvent_t, 4>* ev) {
  Rt3Args a{};
  a.verts = h->d_verts;
  a.ainv = h->d_ainv;
  a.fq = h->has_f ? h->d_f : nullptr;
  a.MF = h->d_MF;
  a.PD = h->d_PD;
  a.WK = h->d_WK;
  a.S = h->d_S;
  a.M = h->d_M;
  a.b = h->d_b;
  a.its = h->d_its;
  a.flg = h->d_flg;
  a.status = h->d_status;
  a.n_cells = h->n_cells;
  a.ns = h->ns;
  a.max_iter = h->prm.max_iter > 0 ? h->prm.max_iter : 20000;
  a.rtol = h->prm.rtol > 0.0 ? h->prm.rtol : 1e-13;
  const int nc = h->n_cells;
  hipLaunchKernelGGL(rt3_local_kernel<L>, dim3(nc), dim3(256), 0, s, a);
  NRT_HIP(hipGetLastError());
  hipLaunchKernelGGL(rt3_diag_kernel<L>, dim3(nc), dim3(256), 0, s, a);
  NRT_HIP(hipGetLastError());
  if (ev) NRT_HIP(hipEventRecord((*ev)[1], s));
  const int groups = (nc + 7) / 8;
  hipLaunchKernelGGL(rt3_cg_kernel<L>, dim3(groups * 8 * 6), dim3(kItNT), 0, s, a);
  NRT_HIP(hipGetLastError());
  if (ev) NRT_HIP(hipEventRecord((*ev)[2], s));
  hipLaunchKernelGGL(rt3_element_kernel<L>, dim3(nc * 7), dim3(256), 0, s, a);
  NRT_HIP(hipGetLastError());
  hipLaunchKernelGGL(rt3_status_kernel, dim3((nc + 255) / 256), dim3(256), 0, s, a);
  NRT_HIP(hipGetLastError());
  return MSFEC_OK;
}

}  // namespace

extern "C" {

msfec_status msfec_rt3_create(const msfec_rt3_params* params, int n_cells, const double* vertices,
                              int vertices_loc, msfec_rt3_handle* out) {
  if (!params || !vertices || !out) return nrt_fail(MSFEC_ERR_INVALID, "null argument");
  *out = nullptr;
  if (n_cells <= 0) return nrt_fail(MSFEC_ERR_INVALID, "n_cells must be positive");
  const int l = params->n_refine_local;
  if (l < 1) return nrt_fail(MSFEC_ERR_INVALID, "n_refine_local must be at least 1");
  if (l > 4) return nrt_fail(MSFEC_ERR_UNSUPPORTED, "n_refine_local above 4");
  if (params->max_iter < 0 || !(params->rtol >= 0.0))
    return nrt_fail(MSFEC_ERR_INVALID, "max_iter and rtol must not be negative");
  if (vertices_loc != MSFEC_HOST && vertices_loc != MSFEC_DEVICE)
    return nrt_fail(MSFEC_ERR_INVALID, "bad vertices location");
  std::vector<double> hv((size_t)n_cells * 24);
  if (vertices_loc == MSFEC_DEVICE) {
    NRT_HIP(hipMemcpy(hv.data(), vertices, hv.size() * sizeof(double), hipMemcpyDeviceToHost));
  } else {
    std::copy(vertices, vertices + hv.size(), hv.begin());
  }
  for (int c = 0; c < n_cells; ++c) {
    const double* v = &hv[(size_t)c * 24];
    bool box = true;
    for (int d = 0; d < 3; ++d) {
      const double Hd = v[3 * (1 << d) + d] - v[d];
      if (!(Hd > 0)) box = false;
      for (int k = 0; k < 8 && box; ++k)
        if (std::fabs(v[3 * k + d] - (v[d] + ((k >> d) & 1) * Hd)) > 1e-14 * Hd) box = false;
    }
    if (!box)
      return nrt_fail(MSFEC_ERR_INVALID,
                      "cell " + std::to_string(c) + " is not an axis-aligned box in deal.II vertex order");
  }
  msfec_rt3_batch* h = new msfec_rt3_batch();
  h->prm = *params;
  h->n_cells = n_cells;
  h->l = l;
  const NrtTables tab = nrt_tables(l);
  h->ne = tab.ne;
  h->nf = tab.N - tab.ne;
  const int m = 1 << l;
  h->ns = 3 * m * (m - 1) * (m - 1);  // interior edges
  const size_t nc = (size_t)n_cells, nfc = (size_t)1 << (3 * l);
  msfec_status st = MSFEC_OK;
#define RT3_TRY(x)        \
  do {                    \
    st = (x);             \
    if (st != MSFEC_OK) { \
      rt3_free(h);        \
      delete h;           \
      return st;          \
    }                     \
  } while (0)
  RT3_TRY(nrt_alloc(&h->d_verts, nc * 24));
  RT3_TRY(nrt_alloc(&h->d_ainv, nc * nfc * 8 * 6));
  RT3_TRY(nrt_alloc(&h->d_f, nc * nfc * 8));
  RT3_TRY(nrt_alloc(&h->d_MF, nc * 21 * nfc));
  RT3_TRY(nrt_alloc(&h->d_PD, nc * h->ns));
  RT3_TRY(nrt_alloc(&h->d_WK, nc * 6 * 4 * h->ns));
  RT3_TRY(nrt_alloc(&h->d_S, nc * 6 * h->nf));
  RT3_TRY(nrt_alloc(&h->d_M, nc * 49));
  RT3_TRY(nrt_alloc(&h->d_b, nc * 7));
  RT3_TRY(nrt_alloc(&h->d_its, nc * 6));
  RT3_TRY(nrt_alloc(&h->d_flg, nc * 6));
  RT3_TRY(nrt_alloc(&h->d_status, nc));
  RT3_TRY(nrt_in(h->stream, h->d_verts, hv.data(), hv.size() * sizeof(double), MSFEC_HOST));
#undef RT3_TRY
  *out = h;
  return MSFEC_OK;
}

msfec_status msfec_rt3_set_coefficients(msfec_rt3_handle h, const double* ainv, const double* f,
                                        int loc) {
  if (!h || !ainv) return nrt_fail(MSFEC_ERR_INVALID, "null argument");
  if (loc != MSFEC_HOST && loc != MSFEC_DEVICE) return nrt_fail(MSFEC_ERR_INVALID, "bad location");
  const size_t pts = (size_t)h->n_cells * ((size_t)1 << (3 * h->l)) * 8;
  msfec_status st = nrt_in(h->stream, h->d_ainv, ainv, pts * 6 * sizeof(double), loc);
  if (st == MSFEC_OK && f) st = nrt_in(h->stream, h->d_f, f, pts * sizeof(double), loc);
  if (st != MSFEC_OK) return st;
  h->has_f = f != nullptr;
  h->coef_set = true;
  return MSFEC_OK;
}

msfec_status msfec_rt3_compute(msfec_rt3_handle h, void* stream) {
  if (!h) return nrt_fail(MSFEC_ERR_INVALID, "null handle");
  if (!h->coef_set) return nrt_fail(MSFEC_ERR_STATE, "coefficients not set");
  hipStream_t s = (hipStream_t)stream;
  h->stream = s;
  std::array<hipEvent_t, 4>* ev = nullptr;
  if (h->prm.flags & MSFEC_FLAG_TIMING) {
    if (h->ev_used == h->ev_sets.size()) {
      std::array<hipEvent_t, 4> set{};
      for (auto& e : set) NRT_HIP(hipEventCreate(&e));
      h->ev_sets.push_back(set);
    }
    ev = &h->ev_sets[h->ev_used++];
    NRT_HIP(hipEventRecord((*ev)[0], s));
  }
  msfec_status st;
  switch (h->l) {
    case 1: st = rt3_launch<1>(h, s, ev); break;
    case 2: st = rt3_launch<2>(h, s, ev); break;
    case 3: st = rt3_launch<3>(h, s, ev); break;
    default: st = rt3_launch<4>(h, s, ev); break;
  }
  if (st != MSFEC_OK) return st;
  if (ev) NRT_HIP(hipEventRecord((*ev)[3], s));
  h->computed = true;
  return MSFEC_OK;
}

msfec_status msfec_rt3_synchronize(msfec_rt3_handle h) {
  if (!h) return nrt_fail(MSFEC_ERR_INVALID, "null handle");
  NRT_HIP(hipStreamSynchronize(h->stream));
  return MSFEC_OK;
}

msfec_status msfec_rt3_get_element(msfec_rt3_handle h, double* M, double* b, int loc) {
  if (!h) return nrt_fail(MSFEC_ERR_INVALID, "null handle");
  if (!h->computed) return nrt_fail(MSFEC_ERR_STATE, "compute not called");
  msfec_status st = MSFEC_OK;
  if (M) st = nrt_out(h->stream, M, h->d_M, (size_t)h->n_cells * 49, loc);
  if (st == MSFEC_OK && b) st = nrt_out(h->stream, b, h->d_b, (size_t)h->n_cells * 7, loc);
  return st;
}

msfec_status msfec_rt3_get_basis(msfec_rt3_handle h, double* sigma, int loc) {
  if (!h) return nrt_fail(MSFEC_ERR_INVALID, "null handle");
  if (!h->computed) return nrt_fail(MSFEC_ERR_STATE, "compute not called");
  return nrt_out(h->stream, sigma, h->d_S, (size_t)h->n_cells * 6 * h->nf, loc);
}

msfec_status msfec_rt3_get_cell_status(msfec_rt3_handle h, int* status) {
  if (!h || !status) return nrt_fail(MSFEC_ERR_INVALID, "null argument");
  if (!h->computed) return nrt_fail(MSFEC_ERR_STATE, "compute not called");
  NRT_HIP(hipStreamSynchronize(h->stream));
  NRT_HIP(hipMemcpy(status, h->d_status, (size_t)h->n_cells * sizeof(int), hipMemcpyDeviceToHost));
  return MSFEC_OK;
}

msfec_status msfec_rt3_get_iterations(msfec_rt3_handle h, int* iterations) {
  if (!h || !iterations) return nrt_fail(MSFEC_ERR_INVALID, "null argument");
  if (!h->computed) return nrt_fail(MSFEC_ERR_STATE, "compute not called");
  NRT_HIP(hipStreamSynchronize(h->stream));
  NRT_HIP(hipMemcpy(iterations, h->d_its, (size_t)h->n_cells * 6 * sizeof(int), hipMemcpyDeviceToHost));
  return MSFEC_OK;
}

msfec_status msfec_rt3_kernel_times(msfec_rt3_handle h, float* ms, int max, int* n) {
  if (!h || !ms) return nrt_fail(MSFEC_ERR_INVALID, "null argument");
  if (h->ev_used == 0) return nrt_fail(MSFEC_ERR_STATE, "no timed compute (set MSFEC_FLAG_TIMING)");
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (size_t c = 0; c < h->ev_used; ++c) {
    const auto& e = h->ev_sets[c];
    NRT_HIP(hipEventSynchronize(e[3]));
    float t[4] = {0.f, 0.f, 0.f, 0.f};
    NRT_HIP(hipEventElapsedTime(&t[0], e[0], e[3]));
    NRT_HIP(hipEventElapsedTime(&t[1], e[0], e[1]));
    NRT_HIP(hipEventElapsedTime(&t[2], e[1], e[2]));
    NRT_HIP(hipEventElapsedTime(&t[3], e[2], e[3]));
    for (int k = 0; k < 4; ++k) acc[k] += t[k];
  }
  const int cnt = std::min(max, 4);
  for (int k = 0; k < cnt; ++k) ms[k] = (float)(acc[k] / (double)h->ev_used);
  if (n) *n = std::max(cnt, 0);
  h->ev_used = 0;
  return MSFEC_OK;
}

msfec_status msfec_rt3_destroy(msfec_rt3_handle h) {
  if (!h) return MSFEC_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  rt3_free(h);
  delete h;
  return MSFEC_OK;
}

}  // extern "C"
