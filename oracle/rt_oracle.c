/*
 * FP64 CPU oracle for the 2-D Raviart-Thomas x DQ0 multiscale basis (test
 * infrastructure only; see rt_oracle.h for the reference files restated).
 *
 * This follows the reference's algorithm: assemble the mixed system with
 * QGauss(2) over the deal.II shape functions, impose the boundary fluxes,
 * solve the saddle-point system directly, replace u by 1 and form the element
 * matrix from the assembled matrix.  The GPU library solves an equivalent
 * reduced problem instead, so the two share no arithmetic.
 *
 * Pinning (tests/test_rt_oracle.py):
 *   - constant coefficient a: the fine basis is the coarse Raviart-Thomas
 *     function itself and the flux block of M is its mass matrix / a (closed
 *     form); a coefficient depending on x only leaves the two x bases
 *     unchanged as well (the flux is a gradient field);
 *   - any coefficient: every fine cell balances its fluxes against -+1/|K|,
 *     the boundary fluxes are those of the coarse function, and the u of the
 *     solve satisfies the first equation with the recomputed assembled
 *     matrix; the flux block equals the one of an independent stream-function
 *     solve written in the test; the complementary energy of each basis is at
 *     most that of the coarse function;
 *   - tensors at the Gauss points: a I given as a tensor reproduces the scalar
 *     mode, and random rotated anisotropic tensors match the test's
 *     stream-function solve with the same point values and satisfy the
 *     saddle-point equations.
 */
#include "rt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* 2-point Gauss-Legendre rule on [0, 1] (QGauss(degree + 2) with degree 0,
 * rt_dq_basis.cc:402) */
static const double kRtGx[2] = {0.21132486540518711775, 0.78867513459481288225};
static const double kRtGw[2] = {0.5, 0.5};

typedef struct {
  int m, nv, ne, nc, N;
  double x0, y0, Hx, Hy, hx, hy;
} rt_mesh;

static int rt_mesh_init(int l, const double *v, rt_mesh *g) {
  if (l < 1 || l > 6) return 1;
  g->m = 1 << l;
  g->nv = g->m * (g->m + 1);
  g->ne = 2 * g->nv;
  g->nc = g->m * g->m;
  g->N = g->ne + g->nc;
  /* deal.II vertex order: v0 (x0,y0), v1 (x1,y0), v2 (x0,y1), v3 (x1,y1) */
  g->x0 = v[0];
  g->y0 = v[1];
  g->Hx = v[2] - v[0];
  g->Hy = v[5] - v[1];
  if (!(g->Hx > 0) || !(g->Hy > 0)) return 1;
  if (fabs(v[3] - v[1]) > 1e-14 * g->Hy || fabs(v[4] - v[0]) > 1e-14 * g->Hx ||
      fabs(v[6] - v[2]) > 1e-14 * g->Hx || fabs(v[7] - v[5]) > 1e-14 * g->Hy)
    return 1;
  g->hx = g->Hx / g->m;
  g->hy = g->Hy / g->m;
  return 0;
}

static int vedge(const rt_mesh *g, int p, int q) { return q * (g->m + 1) + p; }
static int hedge(const rt_mesh *g, int p, int q) { return g->nv + q * g->m + p; }

/* local dofs of fine cell (p, q) in deal.II order: the faces x = left,
 * x = right, y = bottom, y = top, then the cell value */
static void rt_cell_dofs(const rt_mesh *g, int p, int q, int *d) {
  d[0] = vedge(g, p, q);
  d[1] = vedge(g, p + 1, q);
  d[2] = hedge(g, p, q);
  d[3] = hedge(g, p, q + 1);
  d[4] = g->ne + q * g->m + p;
}

/* deal.II's lowest-order Raviart-Thomas shape functions on the unit square
 * (face 0: (1-s, 0), face 1: (s, 0), face 2: (0, 1-t), face 3: (0, t)),
 * mapped to the fine box by the contravariant Piola transform J phi / det J,
 * and their divergence (reference divergence / det J). */
static void rt_shape(double s, double t, double hx, double hy, double (*phi)[2], double *div) {
  const double ref[4][2] = {{1.0 - s, 0.0}, {s, 0.0}, {0.0, 1.0 - t}, {0.0, t}};
  const double rdiv[4] = {-1.0, 1.0, -1.0, 1.0};
  const double det = hx * hy;
  for (int k = 0; k < 4; ++k) {
    phi[k][0] = hx * ref[k][0] / det;
    phi[k][1] = hy * ref[k][1] / det;
    div[k] = rdiv[k] / det;
  }
}

/* A^-1 at Gauss point qp of fine cell c: a I per fine cell (tensor == 0, a
 * holds one value per cell) or the inverse of the symmetric tensor
 * (xx, xy, yy) given at the point (tensor == 1, a holds 4 x 3 per cell),
 * which is what DiffusionInverse_A::value_list returns there
 * (source/equation_data/eqn_coeff_A.cc:194-242) */
static void rt_ainv(const double *a, int tensor, int c, int qp, double B[2][2]) {
  if (!tensor) {
    const double v = 1.0 / a[c];
    B[0][0] = v;
    B[1][1] = v;
    B[0][1] = B[1][0] = 0.0;
    return;
  }
  const double *t = a + ((size_t)c * 4 + qp) * 3;
  const double det = t[0] * t[2] - t[1] * t[1];
  B[0][0] = t[2] / det;
  B[1][1] = t[0] / det;
  B[0][1] = B[1][0] = -t[1] / det;
}

/* the local matrix (test, trial) of assemble_system on one fine cell
 * (rt_dq_basis.cc:495-503) and the local load (rt_dq_basis.cc:507) */
static void rt_local(const rt_mesh *g, const double *a, int tensor, int c, double f,
                     double K[5][5], double Fl[5]) {
  memset(K, 0, sizeof(double) * 25);
  memset(Fl, 0, sizeof(double) * 5);
  for (int qy = 0; qy < 2; ++qy)
    for (int qx = 0; qx < 2; ++qx) {
      double phi[4][2], div[4], B[2][2];
      rt_shape(kRtGx[qx], kRtGx[qy], g->hx, g->hy, phi, div);
      rt_ainv(a, tensor, c, qx + 2 * qy, B);
      const double jxw = kRtGw[qx] * kRtGw[qy] * g->hx * g->hy;
      /* local dof k < 4: sigma shape k, u value 0; k = 4: sigma 0, u 1 */
      for (int i = 0; i < 5; ++i) {
        const double si0 = i < 4 ? phi[i][0] : 0.0, si1 = i < 4 ? phi[i][1] : 0.0;
        const double di = i < 4 ? div[i] : 0.0, ui = i == 4 ? 1.0 : 0.0;
        for (int j = 0; j < 5; ++j) {
          const double sj0 = j < 4 ? phi[j][0] : 0.0, sj1 = j < 4 ? phi[j][1] : 0.0;
          const double dj = j < 4 ? div[j] : 0.0, uj = j == 4 ? 1.0 : 0.0;
          const double aa = si0 * (B[0][0] * sj0 + B[0][1] * sj1) + si1 * (B[1][0] * sj0 + B[1][1] * sj1);
          K[i][j] += (aa - di * uj + ui * dj) * jxw;
        }
        Fl[i] += ui * f * jxw;
      }
    }
}

static int rt_assemble(int l, const double *verts, const double *a, int tensor, const double *f,
                       double *A, double *F) {
  rt_mesh g;
  if (rt_mesh_init(l, verts, &g)) return 1;
  const int N = g.N;
  memset(A, 0, sizeof(double) * (size_t)N * N);
  memset(F, 0, sizeof(double) * N);
  for (int q = 0; q < g.m; ++q)
    for (int p = 0; p < g.m; ++p) {
      const int c = q * g.m + p;
      double K[5][5], Fl[5];
      rt_local(&g, a, tensor, c, f ? f[c] : 1.0, K, Fl);
      int d[5];
      rt_cell_dofs(&g, p, q, d);
      for (int i = 0; i < 5; ++i) {
        for (int j = 0; j < 5; ++j) A[(size_t)d[i] * N + d[j]] += K[i][j];
        F[d[i]] += Fl[i];
      }
    }
  return 0;
}

/* LU with partial pivoting of the n x n row-major matrix A, solving for nr
 * right-hand sides stored as B[n][nr]; returns 2 on a zero pivot */
static int rt_lu_solve(int n, double *A, int nr, double *B) {
  for (int k = 0; k < n; ++k) {
    int piv = k;
    double best = fabs(A[(size_t)k * n + k]);
    for (int i = k + 1; i < n; ++i) {
      const double v = fabs(A[(size_t)i * n + k]);
      if (v > best) {
        best = v;
        piv = i;
      }
    }
    if (!(best > 0.0)) return 2;
    if (piv != k) {
      for (int j = 0; j < n; ++j) {
        const double t = A[(size_t)k * n + j];
        A[(size_t)k * n + j] = A[(size_t)piv * n + j];
        A[(size_t)piv * n + j] = t;
      }
      for (int r = 0; r < nr; ++r) {
        const double t = B[(size_t)k * nr + r];
        B[(size_t)k * nr + r] = B[(size_t)piv * nr + r];
        B[(size_t)piv * nr + r] = t;
      }
    }
    const double inv = 1.0 / A[(size_t)k * n + k];
    for (int i = k + 1; i < n; ++i) {
      const double lik = A[(size_t)i * n + k] * inv;
      if (lik == 0.0) continue;
      double *ri = A + (size_t)i * n;
      const double *rk = A + (size_t)k * n;
      for (int j = k + 1; j < n; ++j) ri[j] -= lik * rk[j];
      for (int r = 0; r < nr; ++r) B[(size_t)i * nr + r] -= lik * B[(size_t)k * nr + r];
    }
  }
  for (int k = n - 1; k >= 0; --k) {
    const double inv = 1.0 / A[(size_t)k * n + k];
    for (int r = 0; r < nr; ++r) {
      double s = B[(size_t)k * nr + r];
      for (int j = k + 1; j < n; ++j) s -= A[(size_t)k * n + j] * B[(size_t)j * nr + r];
      B[(size_t)k * nr + r] = s * inv;
    }
  }
  return 0;
}

/* flux of the coarse Raviart-Thomas basis n through the fine boundary edge e
 * in its +x / +y direction: the coarse function of face n is
 * (1 - s, 0) / Hy, (s, 0) / Hy, (0, 1 - t) / Hx or (0, t) / Hx in the
 * coarse reference coordinates (BasisRaviartThomas, the Piola map of the unit
 * square functions), so a fine edge of its own face carries 1/m and every
 * other boundary edge nothing (project_boundary_values_div_conforming is exact
 * for these fluxes). */
static double rt_boundary_flux(const rt_mesh *g, int e, int n) {
  if (e < g->nv) {
    const int p = e % (g->m + 1);
    if (p == 0 && n == 0) return 1.0 / g->m;
    if (p == g->m && n == 1) return 1.0 / g->m;
  } else {
    const int q = (e - g->nv) / g->m;
    if (q == 0 && n == 2) return 1.0 / g->m;
    if (q == g->m && n == 3) return 1.0 / g->m;
  }
  return 0.0;
}

static int rt_on_boundary(const rt_mesh *g, int e) {
  if (e < g->nv) {
    const int p = e % (g->m + 1);
    return p == 0 || p == g->m;
  }
  const int q = (e - g->nv) / g->m;
  return q == 0 || q == g->m;
}

int oq_rt_assemble(int l, const double *verts, const double *a, const double *f, double *A,
                   double *F) {
  return rt_assemble(l, verts, a, 0, f, A, F);
}

int oq_rt_assemble_qp(int l, const double *verts, const double *a_qp, const double *f, double *A,
                      double *F) {
  return rt_assemble(l, verts, a_qp, 1, f, A, F);
}

static int rt_basis_cell(int l, const double *verts, const double *a, int tensor, const double *f,
                         double *M, double *b, double *Sigma, double *U) {
  rt_mesh g;
  if (rt_mesh_init(l, verts, &g)) return 1;
  const int N = g.N, ne = g.ne, nc = g.nc;
  double *A = (double *)malloc(sizeof(double) * (size_t)N * N);
  double *F = (double *)malloc(sizeof(double) * N);
  int *pos = (int *)malloc(sizeof(int) * N);
  double *sol = (double *)malloc(sizeof(double) * (size_t)N * 4);
  if (!A || !F || !pos || !sol) {
    free(A);
    free(F);
    free(pos);
    free(sol);
    return 3;
  }
  rt_assemble(l, verts, a, tensor, f, A, F);
  /* free unknowns: interior edges, then every cell but the last, whose
   * balance equation is the sum of the others (the pure Neumann problem
   * fixes u up to a constant; its value is pinned to 0 and the mean removed
   * afterwards, which is the Schur-complement solution of the reference) */
  int n = 0;
  for (int d = 0; d < N; ++d) pos[d] = (d < ne && rt_on_boundary(&g, d)) || d == N - 1 ? -1 : n++;
  double *S = (double *)calloc((size_t)n * n, sizeof(double));
  double *B = (double *)calloc((size_t)n * 4, sizeof(double));
  int st = (S && B) ? 0 : 3;
  if (!st) {
    const double vol = g.Hx * g.Hy;
    const double sgn[4] = {-1.0, 1.0, -1.0, 1.0};
    for (int r = 0; r < N; ++r) {
      const int pr = pos[r];
      if (pr < 0) continue;
      for (int c = 0; c < N; ++c) {
        const double v = A[(size_t)r * N + c];
        if (v == 0.0) continue;
        if (pos[c] >= 0) {
          S[(size_t)pr * n + pos[c]] = v;
        } else if (c < ne) {  /* condense the boundary fluxes */
          for (int k = 0; k < 4; ++k) B[(size_t)pr * 4 + k] -= v * rt_boundary_flux(&g, c, k);
        }
      }
      if (r >= ne) /* -+ (phi_u, 1/|K|): the cell area over |K| (rt_dq_basis.cc:516-538) */
        for (int k = 0; k < 4; ++k) B[(size_t)pr * 4 + k] += sgn[k] * g.hx * g.hy / vol;
    }
    st = rt_lu_solve(n, S, 4, B);
  }
  if (!st) {
    for (int d = 0; d < N; ++d)
      for (int k = 0; k < 4; ++k)
        sol[(size_t)d * 4 + k] = pos[d] >= 0 ? B[(size_t)pos[d] * 4 + k]
                                 : (d < ne ? rt_boundary_flux(&g, d, k) : 0.0);
    for (int k = 0; k < 4; ++k) {
      double mean = 0.0;
      for (int c = 0; c < nc; ++c) mean += sol[(size_t)(ne + c) * 4 + k];
      mean /= nc;
      for (int c = 0; c < nc; ++c) sol[(size_t)(ne + c) * 4 + k] -= mean;
    }
    if (Sigma)
      for (int k = 0; k < 4; ++k)
        for (int e = 0; e < ne; ++e) Sigma[(size_t)k * ne + e] = sol[(size_t)e * 4 + k];
    if (U)
      for (int k = 0; k < 4; ++k)
        for (int c = 0; c < nc; ++c) U[(size_t)k * nc + c] = sol[(size_t)(ne + c) * 4 + k];
    /* set_u_to_std, then assemble_global_element_matrix: test and trial
     * vectors are the flux block of basis i (i < 4) or the u block of basis
     * 0 (i = 4), which is now the constant 1; each product uses the block of
     * the assembled matrix that the two blocks select (rt_dq_basis.cc:842-912) */
    for (int it = 0; it < 5; ++it) {
      for (int jt = 0; jt < 5; ++jt) {
        double acc = 0.0;
        for (int r = (it < 4 ? 0 : ne); r < (it < 4 ? ne : N); ++r) {
          const double tv = it < 4 ? sol[(size_t)r * 4 + it] : 1.0;
          double row = 0.0;
          for (int c = (jt < 4 ? 0 : ne); c < (jt < 4 ? ne : N); ++c) {
            const double v = jt < 4 ? sol[(size_t)c * 4 + jt] : 1.0;
            row += A[(size_t)r * N + c] * v;
          }
          acc += tv * row;
        }
        M[it * 5 + jt] = acc;
      }
      double rb = 0.0;
      if (it == 4)
        for (int c = 0; c < nc; ++c) rb += F[ne + c];
      b[it] = rb;
    }
  }
  free(S);
  free(B);
  free(A);
  free(F);
  free(pos);
  free(sol);
  return st;
}

int oq_rt_basis_cell(int l, const double *verts, const double *a, const double *f, double *M,
                     double *b, double *Sigma, double *U) {
  return rt_basis_cell(l, verts, a, 0, f, M, b, Sigma, U);
}

int oq_rt_basis_cell_qp(int l, const double *verts, const double *a_qp, const double *f,
                        double *M, double *b, double *Sigma, double *U) {
  return rt_basis_cell(l, verts, a_qp, 1, f, M, b, Sigma, U);
}
