/*
 * FP64 CPU oracle for the 3-D Nedelec x Raviart-Thomas multiscale basis (test
 * infrastructure only; see nrt_oracle.h for the reference files restated).
 *
 * This follows the reference's algorithm: assemble the mixed system over all
 * fine dofs with QGauss(2) and the deal.II lowest-order shape functions, fix
 * the boundary dofs to the traces of the coarse function of each basis,
 * condense, solve directly (LU with partial pivoting and two steps of
 * iterative refinement here, UMFPACK there) and
 * form the element matrix from the assembled matrix by the reference's block
 * rule.  The GPU library solves the same systems with a different
 * factorisation (symmetric indefinite, tangential dofs first) and builds the
 * element matrix from per-cell local matrices, so the two share no
 * arithmetic beyond the shape-function formulas.
 *
 * Pinning (tests/test_nrt_oracle.py):
 *   - the discrete de Rham sequence: the coupling block applied to the edge
 *     values of the gradient of any nodal Q1 field vanishes; the face values
 *     of the curl of any edge field (circulations) have zero divergence
 *     blocks, and the coupling block applied to the edge field equals the
 *     face mass matrix applied to its circulations;
 *   - closed forms: edge and face mass matrices reproduce the integrals of
 *     constant fields (with a full tensor A^-1 for the edges), the
 *     divergence block the integral of a linear field, and the load (v, f)
 *     of a constant f;
 *   - the loads (v, curl N_n) equal the face mass applied to the
 *     circulations of the coarse Nedelec function's edge values;
 *   - a manufactured smooth solution converges at first order under
 *     refinement with the assembled matrix and load;
 *   - the bases satisfy the assembled interior equations with those loads and
 *     carry the coarse traces on the boundary, and the element matrix pairs
 *     the assembled blocks as the reference's block rule states.
 */
#include "nrt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* 2-point Gauss-Legendre rule on [0, 1] (QGauss(degree + 2) with degree 0,
 * ned_rt_basis.cc:373); 3-D points x fastest */
static const double kNrtG[2] = {0.21132486540518711775, 0.78867513459481288225};

typedef struct {
  int m;
  int nx, ne, nf1, nf, N;
  double x0[3], H[3], h[3];
} nrt_mesh;

static int nrt_mesh_init(int l, int lmax, const double *v, nrt_mesh *g) {
  if (l < 1 || l > lmax) return 1;
  g->m = 1 << l;
  const int m = g->m;
  g->nx = m * (m + 1) * (m + 1);
  g->ne = 3 * g->nx;
  g->nf1 = m * m * (m + 1);
  g->nf = 3 * g->nf1;
  g->N = g->ne + g->nf;
  /* deal.II vertex order: vertex n at (x_(n&1), y_((n>>1)&1), z_((n>>2)&1)) */
  for (int d = 0; d < 3; ++d) {
    g->x0[d] = v[d];
    g->H[d] = v[3 * (1 << d) + d] - v[d];
    if (!(g->H[d] > 0)) return 1;
    g->h[d] = g->H[d] / m;
  }
  for (int n = 0; n < 8; ++n)
    for (int d = 0; d < 3; ++d) {
      const double want = g->x0[d] + ((n >> d) & 1) * g->H[d];
      if (fabs(v[3 * n + d] - want) > 1e-14 * g->H[d]) return 1;
    }
  return 0;
}

static int ex(const nrt_mesh *g, int i, int j, int k) { return (k * (g->m + 1) + j) * g->m + i; }
static int ey(const nrt_mesh *g, int i, int j, int k) {
  return g->nx + (k * g->m + j) * (g->m + 1) + i;
}
static int ez(const nrt_mesh *g, int i, int j, int k) {
  return 2 * g->nx + (k * (g->m + 1) + j) * (g->m + 1) + i;
}
static int fx(const nrt_mesh *g, int i, int j, int k) {
  return g->ne + (k * g->m + j) * (g->m + 1) + i;
}
static int fy(const nrt_mesh *g, int i, int j, int k) {
  return g->ne + g->nf1 + (k * (g->m + 1) + j) * g->m + i;
}
static int fz(const nrt_mesh *g, int i, int j, int k) {
  return g->ne + 2 * g->nf1 + (k * g->m + j) * g->m + i;
}

/* deal.II's line order on the unit cube: the axis of line L and the values
 * (0 or 1) of the two other coordinates, lower axis first */
static const int kLine[12][3] = {{1, 0, 0}, {1, 1, 0}, {0, 0, 0}, {0, 1, 0}, {1, 0, 1}, {1, 1, 1},
                                 {0, 0, 1}, {0, 1, 1}, {2, 0, 0}, {2, 1, 0}, {2, 0, 1}, {2, 1, 1}};

static void other_axes(int d, int *o1, int *o2) {
  *o1 = d == 0 ? 1 : 0;
  *o2 = d == 2 ? 1 : 2;
}

/* local dofs of fine cell (a, b, c): the 12 lines, then the 6 faces, both in
 * deal.II order */
static void nrt_cell_dofs(const nrt_mesh *g, int a, int b, int c, int *d) {
  for (int L = 0; L < 12; ++L) {
    const int ax = kLine[L][0], f1 = kLine[L][1], f2 = kLine[L][2];
    if (ax == 0) d[L] = ex(g, a, b + f1, c + f2);
    else if (ax == 1) d[L] = ey(g, a + f1, b, c + f2);
    else d[L] = ez(g, a + f1, b + f2, c);
  }
  d[12] = fx(g, a, b, c);
  d[13] = fx(g, a + 1, b, c);
  d[14] = fy(g, a, b, c);
  d[15] = fy(g, a, b + 1, c);
  d[16] = fz(g, a, b, c);
  d[17] = fz(g, a, b, c + 1);
}

static double lag(int s, double t) { return s ? t : 1.0 - t; }
static double dlag(int s) { return s ? 1.0 : -1.0; }

/* The lowest-order Nedelec function of line L of a box of sizes h at the
 * box's reference point X in [0,1]^3, mapped covariantly (J^-T phi_ref,
 * basis_nedelec.tpp), and its curl (J curl_ref / det J,
 * basis_nedelec_curl.tpp): phi_ref = l_f1(X_o1) l_f2(X_o2) e_axis. */
static void ned_shape(int L, const double *X, const double *h, double *phi, double *curl) {
  const int ax = kLine[L][0], f1 = kLine[L][1], f2 = kLine[L][2];
  int o1, o2;
  other_axes(ax, &o1, &o2);
  const double v1 = lag(f1, X[o1]), v2 = lag(f2, X[o2]);
  phi[0] = phi[1] = phi[2] = 0.0;
  phi[ax] = v1 * v2 / h[ax];
  /* curl(f e_ax) = grad f x e_ax */
  double gr[3] = {0.0, 0.0, 0.0};
  gr[o1] = dlag(f1) * v2 / (h[ax] * h[o1]);
  gr[o2] = v1 * dlag(f2) / (h[ax] * h[o2]);
  double e[3] = {0.0, 0.0, 0.0};
  e[ax] = 1.0;
  curl[0] = gr[1] * e[2] - gr[2] * e[1];
  curl[1] = gr[2] * e[0] - gr[0] * e[2];
  curl[2] = gr[0] * e[1] - gr[1] * e[0];
}

/* The lowest-order Raviart-Thomas function of face F (axis F / 2, side F % 2)
 * of a box of sizes h, mapped contravariantly (J psi_ref / det J,
 * basis_raviart_thomas.tpp): psi_ref = l_side(X_axis) e_axis. */
static void rt_shape(int F, const double *X, const double *h, double *psi, double *div) {
  const int ax = F / 2, s = F % 2;
  const double det = h[0] * h[1] * h[2];
  psi[0] = psi[1] = psi[2] = 0.0;
  psi[ax] = h[ax] * lag(s, X[ax]) / det;
  *div = dlag(s) / det;
}

/* local matrix (test i, trial j) of one fine cell (ned_rt_basis.cc:488-498)
 * and its load (v, f) (:500) */
static void nrt_local(const nrt_mesh *g, int c, const double *ainv, const double *bq,
                      const double *rq, const double *fq, double K[18][18], double Fl[18]) {
  memset(K, 0, sizeof(double) * 18 * 18);
  memset(Fl, 0, sizeof(double) * 18);
  const double vol = g->h[0] * g->h[1] * g->h[2];
  for (int q = 0; q < 8; ++q) {
    const double X[3] = {kNrtG[q & 1], kNrtG[(q >> 1) & 1], kNrtG[(q >> 2) & 1]};
    const double w = vol / 8.0;
    double phi[12][3], cphi[12][3], psi[6][3], dv[6];
    for (int L = 0; L < 12; ++L) ned_shape(L, X, g->h, phi[L], cphi[L]);
    for (int F = 0; F < 6; ++F) rt_shape(F, X, g->h, psi[F], &dv[F]);
    const double *t = ainv + ((size_t)c * 8 + q) * 6;
    const double G[3][3] = {{t[0], t[1], t[2]}, {t[1], t[3], t[4]}, {t[2], t[4], t[5]}};
    const double B = bq[(size_t)c * 8 + q];
    const double R = rq ? rq[(size_t)c * 8 + q] : 0.0;
    for (int i = 0; i < 18; ++i)
      for (int j = 0; j < 18; ++j) {
        double v = 0.0;
        if (i < 12 && j < 12) {
          for (int r = 0; r < 3; ++r)
            for (int s = 0; s < 3; ++s) v += phi[i][r] * G[r][s] * phi[j][s];
        } else if (i < 12) {
          const double *p = psi[j - 12];
          v = -(cphi[i][0] * p[0] + cphi[i][1] * p[1] + cphi[i][2] * p[2]);
        } else if (j < 12) {
          const double *p = psi[i - 12];
          v = p[0] * cphi[j][0] + p[1] * cphi[j][1] + p[2] * cphi[j][2];
        } else {
          const double *pi = psi[i - 12], *pj = psi[j - 12];
          v = B * dv[i - 12] * dv[j - 12] + R * (pi[0] * pj[0] + pi[1] * pj[1] + pi[2] * pj[2]);
        }
        K[i][j] += v * w;
      }
    if (fq) {
      const double *f = fq + ((size_t)c * 8 + q) * 3;
      for (int F = 0; F < 6; ++F)
        Fl[12 + F] += (psi[F][0] * f[0] + psi[F][1] * f[1] + psi[F][2] * f[2]) * w;
    }
  }
}

static void nrt_assemble(const nrt_mesh *g, const double *ainv, const double *bq,
                         const double *rq, const double *fq, double *A, double *F) {
  const int N = g->N, m = g->m;
  memset(A, 0, sizeof(double) * (size_t)N * N);
  memset(F, 0, sizeof(double) * N);
  for (int c3 = 0; c3 < m; ++c3)
    for (int b = 0; b < m; ++b)
      for (int a = 0; a < m; ++a) {
        const int c = (c3 * m + b) * m + a;
        double K[18][18], Fl[18];
        int d[18];
        nrt_local(g, c, ainv, bq, rq, fq, K, Fl);
        nrt_cell_dofs(g, a, b, c3, d);
        for (int i = 0; i < 18; ++i) {
          for (int j = 0; j < 18; ++j) A[(size_t)d[i] * N + d[j]] += K[i][j];
          F[d[i]] += Fl[i];
        }
      }
}

/* (v, curl N_n) over all fine cells, N_n the coarse Nedelec function of line
 * n (BasisNedelecCurl at the fine quadrature points, ned_rt_basis.cc:441-450,
 * 503-511) */
static void nrt_curl_loads(const nrt_mesh *g, double *Fc) {
  const int N = g->N, m = g->m;
  memset(Fc, 0, sizeof(double) * 12 * (size_t)N);
  const double vol = g->h[0] * g->h[1] * g->h[2];
  for (int c3 = 0; c3 < m; ++c3)
    for (int b = 0; b < m; ++b)
      for (int a = 0; a < m; ++a) {
        int d[18];
        nrt_cell_dofs(g, a, b, c3, d);
        for (int q = 0; q < 8; ++q) {
          const double Xf[3] = {kNrtG[q & 1], kNrtG[(q >> 1) & 1], kNrtG[(q >> 2) & 1]};
          const double Xc[3] = {(a + Xf[0]) / m, (b + Xf[1]) / m, (c3 + Xf[2]) / m};
          double psi[6][3], dv[6];
          for (int F = 0; F < 6; ++F) rt_shape(F, Xf, g->h, psi[F], &dv[F]);
          for (int n = 0; n < 12; ++n) {
            double phi[3], curl[3];
            ned_shape(n, Xc, g->H, phi, curl);
            for (int F = 0; F < 6; ++F)
              Fc[(size_t)n * N + d[12 + F]] +=
                  (psi[F][0] * curl[0] + psi[F][1] * curl[1] + psi[F][2] * curl[2]) * vol / 8.0;
          }
        }
      }
}

/* Boundary dofs and their fixed values for basis n.  A fine edge on the
 * coarse boundary carries the line integral of the coarse Nedelec function
 * of the basis (the curl-conforming projection of ned_rt_basis.cc:261-267 is
 * exact: the coarse function lies in the fine space), a fine boundary face
 * the flux of the coarse Raviart-Thomas function (:337-343). */
static int nrt_on_boundary(const nrt_mesh *g, int dof, int *ax, int *p) {
  const int m = g->m;
  int i, j, k, t;
  if (dof < g->ne) {
    t = dof / g->nx;
    int r = dof - t * g->nx;
    if (t == 0) {
      i = r % m;
      r /= m;
      j = r % (m + 1);
      k = r / (m + 1);
    } else if (t == 1) {
      i = r % (m + 1);
      r /= (m + 1);
      j = r % m;
      k = r / m;
    } else {
      i = r % (m + 1);
      r /= (m + 1);
      j = r % (m + 1);
      k = r / (m + 1);
    }
  } else {
    t = 3 + (dof - g->ne) / g->nf1;
    int r = (dof - g->ne) % g->nf1;
    if (t == 3) {
      i = r % (m + 1);
      r /= (m + 1);
      j = r % m;
      k = r / m;
    } else if (t == 4) {
      i = r % m;
      r /= m;
      j = r % (m + 1);
      k = r / (m + 1);
    } else {
      i = r % m;
      r /= m;
      j = r % m;
      k = r / m;
    }
  }
  *ax = t;
  p[0] = i;
  p[1] = j;
  p[2] = k;
  if (t < 3) {
    /* an edge along axis t lies on the boundary when one of its two fixed
     * node coordinates is 0 or m */
    for (int d = 0; d < 3; ++d)
      if (d != t && (p[d] == 0 || p[d] == m)) return 1;
    return 0;
  }
  return p[t - 3] == 0 || p[t - 3] == m;
}

static double nrt_boundary_value(const nrt_mesh *g, int t, const int *p, int n) {
  const int m = g->m;
  if (n < 12) {
    const int ax = kLine[n][0];
    if (t != ax) return 0.0;
    int o1, o2;
    other_axes(ax, &o1, &o2);
    return lag(kLine[n][1], (double)p[o1] / m) * lag(kLine[n][2], (double)p[o2] / m) / m;
  }
  const int ax = (n - 12) / 2, s = (n - 12) % 2;
  if (t != 3 + ax || p[ax] != s * m) return 0.0;
  return 1.0 / ((double)m * m);
}

/* LU with partial pivoting of the n x n row-major matrix A in place (the
 * multipliers below the diagonal, the row exchanges in piv); returns 2 on a
 * zero pivot */
static int nrt_lu(int n, double *A, int *piv) {
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = fabs(A[(size_t)k * n + k]);
    for (int i = k + 1; i < n; ++i) {
      const double v = fabs(A[(size_t)i * n + k]);
      if (v > best) {
        best = v;
        p = i;
      }
    }
    if (!(best > 0.0)) return 2;
    piv[k] = p;
    if (p != k)
      for (int j = 0; j < n; ++j) {
        const double t = A[(size_t)k * n + j];
        A[(size_t)k * n + j] = A[(size_t)p * n + j];
        A[(size_t)p * n + j] = t;
      }
    const double inv = 1.0 / A[(size_t)k * n + k];
    const double *rk = A + (size_t)k * n;
    for (int i = k + 1; i < n; ++i) {
      double *ri = A + (size_t)i * n;
      const double lik = ri[k] * inv;
      ri[k] = lik;
      if (lik == 0.0) continue;
      for (int j = k + 1; j < n; ++j) ri[j] -= lik * rk[j];
    }
  }
  return 0;
}

/* x = A^-1 b with the factors of nrt_lu, in place on b (n values): the row
 * exchanges (whole rows, multipliers included, were swapped), then the unit
 * lower and the upper triangular solves */
static void nrt_lu_apply(int n, const double *LU, const int *piv, double *b) {
  for (int k = 0; k < n; ++k)
    if (piv[k] != k) {
      const double t = b[k];
      b[k] = b[piv[k]];
      b[piv[k]] = t;
    }
  for (int k = 0; k < n; ++k) {
    const double bk = b[k];
    if (bk != 0.0)
      for (int i = k + 1; i < n; ++i) b[i] -= LU[(size_t)i * n + k] * bk;
  }
  for (int k = n - 1; k >= 0; --k) {
    double s = b[k];
    for (int j = k + 1; j < n; ++j) s -= LU[(size_t)k * n + j] * b[j];
    b[k] = s / LU[(size_t)k * n + k];
  }
}

/* Solves S x = b for nr right-hand sides stored as B[n][nr]: LU with partial
 * pivoting, then two steps of iterative refinement with FP64 residuals of the
 * unfactored S (the pivot growth of the indefinite mixed systems otherwise
 * costs a digit or two on strongly varying coefficients); returns 2 on a zero
 * pivot, 3 when memory runs out */
static int nrt_solve(int n, const double *S, int nr, double *B) {
  double *LU = (double *)malloc(sizeof(double) * (size_t)n * n);
  int *piv = (int *)malloc(sizeof(int) * n);
  double *x = (double *)malloc(sizeof(double) * n);
  double *r = (double *)malloc(sizeof(double) * n);
  int st = (LU && piv && x && r) ? 0 : 3;
  if (!st) {
    memcpy(LU, S, sizeof(double) * (size_t)n * n);
    st = nrt_lu(n, LU, piv);
  }
  for (int k = 0; !st && k < nr; ++k) {
    for (int i = 0; i < n; ++i) x[i] = B[(size_t)i * nr + k];
    nrt_lu_apply(n, LU, piv, x);
    for (int it = 0; it < 2; ++it) {
      for (int i = 0; i < n; ++i) {
        double s = B[(size_t)i * nr + k];
        const double *si = S + (size_t)i * n;
        for (int j = 0; j < n; ++j) s -= si[j] * x[j];
        r[i] = s;
      }
      nrt_lu_apply(n, LU, piv, r);
      for (int i = 0; i < n; ++i) x[i] += r[i];
    }
    for (int i = 0; i < n; ++i) B[(size_t)i * nr + k] = x[i];
  }
  free(LU);
  free(piv);
  free(x);
  free(r);
  return st;
}

int oq_nrt_assemble(int l, const double *verts, const double *ainv, const double *bq,
                    const double *rq, const double *fq, double *A, double *F) {
  nrt_mesh g;
  if (nrt_mesh_init(l, 4, verts, &g)) return 1;
  nrt_assemble(&g, ainv, bq, rq, fq, A, F);
  return 0;
}

int oq_nrt_curl_loads(int l, const double *verts, double *Fc) {
  nrt_mesh g;
  if (nrt_mesh_init(l, 4, verts, &g)) return 1;
  nrt_curl_loads(&g, Fc);
  return 0;
}

int oq_nrt_assemble_coo(int l, const double *verts, const double *ainv, const double *bq,
                        const double *rq, const double *fq, int *rows, int *cols, double *vals,
                        double *F) {
  nrt_mesh g;
  if (nrt_mesh_init(l, 4, verts, &g)) return 1;
  const int m = g.m;
  memset(F, 0, sizeof(double) * g.N);
  size_t k = 0;
  for (int c3 = 0; c3 < m; ++c3)
    for (int b = 0; b < m; ++b)
      for (int a = 0; a < m; ++a) {
        const int c = (c3 * m + b) * m + a;
        double K[18][18], Fl[18];
        int d[18];
        nrt_local(&g, c, ainv, bq, rq, fq, K, Fl);
        nrt_cell_dofs(&g, a, b, c3, d);
        for (int i = 0; i < 18; ++i) {
          for (int j = 0; j < 18; ++j, ++k) {
            rows[k] = d[i];
            cols[k] = d[j];
            vals[k] = K[i][j];
          }
          F[d[i]] += Fl[i];
        }
      }
  return 0;
}

int oq_nrt_boundary(int l, const double *verts, double *G, int *on_boundary) {
  nrt_mesh g;
  if (nrt_mesh_init(l, 4, verts, &g)) return 1;
  for (int d = 0; d < g.N; ++d) {
    int t, p[3];
    on_boundary[d] = nrt_on_boundary(&g, d, &t, p);
    for (int n = 0; n < 18; ++n)
      G[(size_t)n * g.N + d] = on_boundary[d] ? nrt_boundary_value(&g, t, p, n) : 0.0;
  }
  return 0;
}

int oq_nrt_basis_cell(int l, const double *verts, const double *ainv, const double *bq,
                      const double *rq, const double *fq, double *M, double *b, double *X) {
  nrt_mesh g;
  if (nrt_mesh_init(l, 3, verts, &g)) return 1;
  const int N = g.N, ne = g.ne;
  double *A = (double *)malloc(sizeof(double) * (size_t)N * N);
  double *F = (double *)malloc(sizeof(double) * N);
  double *Fc = (double *)malloc(sizeof(double) * 12 * (size_t)N);
  double *xb = (double *)calloc((size_t)N * 18, sizeof(double));
  int *pos = (int *)malloc(sizeof(int) * N);
  int st = (A && F && Fc && xb && pos) ? 0 : 3;
  double *S = NULL, *B = NULL;
  int n = 0;
  if (!st) {
    nrt_assemble(&g, ainv, bq, rq, fq, A, F);
    nrt_curl_loads(&g, Fc);
    for (int d = 0; d < N; ++d) {
      int t, p[3];
      if (nrt_on_boundary(&g, d, &t, p)) {
        pos[d] = -1;
        for (int k = 0; k < 18; ++k) xb[(size_t)d * 18 + k] = nrt_boundary_value(&g, t, p, k);
      } else {
        pos[d] = n++;
      }
    }
    S = (double *)malloc(sizeof(double) * (size_t)n * n);
    B = (double *)calloc((size_t)n * 18, sizeof(double));
    if (!S || !B) st = 3;
  }
  if (!st) {
    /* condense: interior rows, boundary columns moved to the right-hand side
     * with the fixed values (AffineConstraints::condense + distribute) */
    for (int r = 0; r < N; ++r) {
      const int pr = pos[r];
      if (pr < 0) continue;
      for (int k = 0; k < 12; ++k) B[(size_t)pr * 18 + k] = Fc[(size_t)k * N + r];
      for (int c = 0; c < N; ++c) {
        const double v = A[(size_t)r * N + c];
        if (pos[c] >= 0) {
          S[(size_t)pr * n + pos[c]] = v;
        } else if (v != 0.0) {
          for (int k = 0; k < 18; ++k) B[(size_t)pr * 18 + k] -= v * xb[(size_t)c * 18 + k];
        }
      }
    }
    st = nrt_solve(n, S, 18, B);
  }
  if (!st) {
    for (int d = 0; d < N; ++d)
      if (pos[d] >= 0)
        for (int k = 0; k < 18; ++k) xb[(size_t)d * 18 + k] = B[(size_t)pos[d] * 18 + k];
    if (X)
      for (int k = 0; k < 18; ++k)
        for (int d = 0; d < N; ++d) X[(size_t)k * N + d] = xb[(size_t)d * 18 + k];
    /* assemble_global_element_matrix (ned_rt_basis.cc:850-948): an edge
     * basis enters through its sigma block (the edges), a face basis through
     * its u block (the faces); entry (i, j) is the test block of basis i
     * times the assembled block (test block, trial block) times the trial
     * block of basis j.  The load pairs the u block of face bases with (v, f). */
    for (int i = 0; i < 18; ++i) {
      const int r0 = i < 12 ? 0 : ne, r1 = i < 12 ? ne : N;
      for (int j = 0; j < 18; ++j) {
        const int c0 = j < 12 ? 0 : ne, c1 = j < 12 ? ne : N;
        double acc = 0.0;
        for (int r = r0; r < r1; ++r) {
          const double tv = xb[(size_t)r * 18 + i];
          if (tv == 0.0) continue;
          double row = 0.0;
          for (int c = c0; c < c1; ++c) row += A[(size_t)r * N + c] * xb[(size_t)c * 18 + j];
          acc += tv * row;
        }
        M[i * 18 + j] = acc;
      }
      double rb = 0.0;
      if (i >= 12)
        for (int r = ne; r < N; ++r) rb += xb[(size_t)r * 18 + i] * F[r];
      b[i] = rb;
    }
  }
  free(S);
  free(B);
  free(A);
  free(F);
  free(Fc);
  free(xb);
  free(pos);
  return st;
}

/* ---------------------------------------------------------------------------
 * RT_DQ in three dimensions (source/RT_DQ/rt_dq_basis.cc), on the same fine
 * grid: sigma on the fine faces (FE_RaviartThomas(0)), u per fine cell
 * (FE_DGQ(0)):
 *   (A^-1 sigma, tau) - (div tau, u) = 0,  (v, div sigma) + (R u, v) = -+(v, 1/|K|)
 * (:488-538; R = 0, as eqn_coeff_R.cc returns).  Basis n fixes the normal
 * trace of the coarse Raviart-Thomas function of face n on the boundary
 * (:267-380); u is pinned in the last cell (the pure Neumann problem fixes
 * it up to a constant) and its mean removed; set_u_to_std sets u = 1
 * (:1039-1043) before the 7 x 7 element matrix (:822-914).  Dofs: the fine
 * faces in the Ned_RT face numbering without the edge offset, then the fine
 * cells. */
static void rt3_local(const nrt_mesh *g, int c, const double *ainv, const double *fq,
                      double K[7][7], double Fl[7]) {
  memset(K, 0, sizeof(double) * 49);
  memset(Fl, 0, sizeof(double) * 7);
  const double vol = g->h[0] * g->h[1] * g->h[2];
  for (int q = 0; q < 8; ++q) {
    const double X[3] = {kNrtG[q & 1], kNrtG[(q >> 1) & 1], kNrtG[(q >> 2) & 1]};
    const double w = vol / 8.0;
    double psi[6][3], dv[6];
    for (int F = 0; F < 6; ++F) rt_shape(F, X, g->h, psi[F], &dv[F]);
    const double *t = ainv + ((size_t)c * 8 + q) * 6;
    const double G[3][3] = {{t[0], t[1], t[2]}, {t[1], t[3], t[4]}, {t[2], t[4], t[5]}};
    for (int i = 0; i < 6; ++i) {
      for (int j = 0; j < 6; ++j) {
        double v = 0.0;
        for (int r = 0; r < 3; ++r)
          for (int s = 0; s < 3; ++s) v += psi[i][r] * G[r][s] * psi[j][s];
        K[i][j] += v * w;
      }
      K[i][6] -= dv[i] * w; /* - (div tau, u) */
      K[6][i] += dv[i] * w; /* (v, div sigma) */
    }
    if (fq) Fl[6] += fq[(size_t)c * 8 + q] * w;
  }
}

static void rt3_assemble(const nrt_mesh *g, const double *ainv, const double *fq, double *A,
                         double *F) {
  const int m = g->m, nf = g->nf, N = nf + m * m * m;
  memset(A, 0, sizeof(double) * (size_t)N * N);
  memset(F, 0, sizeof(double) * N);
  for (int c3 = 0; c3 < m; ++c3)
    for (int b = 0; b < m; ++b)
      for (int a = 0; a < m; ++a) {
        const int c = (c3 * m + b) * m + a;
        double K[7][7], Fl[7];
        int d18[18], d[7];
        rt3_local(g, c, ainv, fq, K, Fl);
        nrt_cell_dofs(g, a, b, c3, d18);
        for (int i = 0; i < 6; ++i) d[i] = d18[12 + i] - g->ne;
        d[6] = nf + c;
        for (int i = 0; i < 7; ++i) {
          for (int j = 0; j < 7; ++j) A[(size_t)d[i] * N + d[j]] += K[i][j];
          F[d[i]] += Fl[i];
        }
      }
}

int oq_rt3_assemble(int l, const double *verts, const double *ainv, const double *fq, double *A,
                    double *F) {
  nrt_mesh g;
  if (nrt_mesh_init(l, 3, verts, &g)) return 1;
  rt3_assemble(&g, ainv, fq, A, F);
  return 0;
}

/* The assembled 3-D RT_DQ system as 49 triplets per fine cell (row = test
 * dof, duplicates summed by the caller) and the load, for a sparse solve at
 * refinements where the dense system does not fit. */
int oq_rt3_assemble_coo(int l, const double *verts, const double *ainv, const double *fq,
                        int *rows, int *cols, double *vals, double *F) {
  nrt_mesh g;
  if (nrt_mesh_init(l, 4, verts, &g)) return 1;
  const int m = g.m, nf = g.nf;
  memset(F, 0, sizeof(double) * (size_t)(nf + m * m * m));
  size_t k = 0;
  for (int c3 = 0; c3 < m; ++c3)
    for (int b = 0; b < m; ++b)
      for (int a = 0; a < m; ++a) {
        const int c = (c3 * m + b) * m + a;
        double K[7][7], Fl[7];
        int d18[18], d[7];
        rt3_local(&g, c, ainv, fq, K, Fl);
        nrt_cell_dofs(&g, a, b, c3, d18);
        for (int i = 0; i < 6; ++i) d[i] = d18[12 + i] - g.ne;
        d[6] = nf + c;
        for (int i = 0; i < 7; ++i) {
          for (int j = 0; j < 7; ++j, ++k) {
            rows[k] = d[i];
            cols[k] = d[j];
            vals[k] = K[i][j];
          }
          F[d[i]] += Fl[i];
        }
      }
  return 0;
}

int oq_rt3_basis_cell(int l, const double *verts, const double *ainv, const double *fq,
                      double *M, double *b, double *Sigma) {
  nrt_mesh g;
  if (nrt_mesh_init(l, 3, verts, &g)) return 1;
  const int nf = g.nf, nc = g.m * g.m * g.m, N = nf + nc;
  double *A = (double *)malloc(sizeof(double) * (size_t)N * N);
  double *F = (double *)malloc(sizeof(double) * N);
  double *fix = (double *)calloc((size_t)N * 6, sizeof(double));
  int *pos = (int *)malloc(sizeof(int) * N);
  int st = (A && F && fix && pos) ? 0 : 3;
  double *S = NULL, *B = NULL;
  int n = 0;
  if (!st) {
    rt3_assemble(&g, ainv, fq, A, F);
    for (int d = 0; d < N; ++d) {
      int t, p[3];
      if (d < nf && nrt_on_boundary(&g, g.ne + d, &t, p)) {
        pos[d] = -1;
        for (int k = 0; k < 6; ++k) fix[(size_t)d * 6 + k] = nrt_boundary_value(&g, t, p, 12 + k);
      } else if (d == N - 1) {
        pos[d] = -1; /* u of the last cell pinned to 0 */
      } else {
        pos[d] = n++;
      }
    }
    S = (double *)malloc(sizeof(double) * (size_t)n * n);
    B = (double *)calloc((size_t)n * 6, sizeof(double));
    if (!S || !B) st = 3;
  }
  if (!st) {
    const double volK = g.H[0] * g.H[1] * g.H[2], vol = g.h[0] * g.h[1] * g.h[2];
    for (int r = 0; r < N; ++r) {
      const int pr = pos[r];
      if (pr < 0) continue;
      for (int c = 0; c < N; ++c) {
        const double v = A[(size_t)r * N + c];
        if (pos[c] >= 0) S[(size_t)pr * n + pos[c]] = v;
        else if (v != 0.0)
          for (int k = 0; k < 6; ++k) B[(size_t)pr * 6 + k] -= v * fix[(size_t)c * 6 + k];
      }
      if (r >= nf) /* -+ (v, 1/|K|): minus for the lower face of an axis (:516-538) */
        for (int k = 0; k < 6; ++k) B[(size_t)pr * 6 + k] += (k % 2 ? 1.0 : -1.0) * vol / volK;
    }
    st = nrt_solve(n, S, 6, B);
  }
  if (!st) {
    double *sol = fix; /* reuse: the full solution, boundary values already there */
    for (int d = 0; d < N; ++d)
      if (pos[d] >= 0)
        for (int k = 0; k < 6; ++k) sol[(size_t)d * 6 + k] = B[(size_t)pos[d] * 6 + k];
    if (Sigma)
      for (int k = 0; k < 6; ++k)
        for (int f = 0; f < nf; ++f) Sigma[(size_t)k * nf + f] = sol[(size_t)f * 6 + k];
    /* set_u_to_std, then the block products: test and trial vectors are the
     * flux block of basis i (i < 6) or the u block of basis 0 (i = 6), now
     * the constant 1 */
    for (int it = 0; it < 7; ++it) {
      for (int jt = 0; jt < 7; ++jt) {
        double acc = 0.0;
        for (int r = (it < 6 ? 0 : nf); r < (it < 6 ? nf : N); ++r) {
          const double tv = it < 6 ? sol[(size_t)r * 6 + it] : 1.0;
          double row = 0.0;
          for (int c = (jt < 6 ? 0 : nf); c < (jt < 6 ? nf : N); ++c)
            row += A[(size_t)r * N + c] * (jt < 6 ? sol[(size_t)c * 6 + jt] : 1.0);
          acc += tv * row;
        }
        M[it * 7 + jt] = acc;
      }
      double rb = 0.0;
      if (it == 6)
        for (int c = 0; c < nc; ++c) rb += F[nf + c];
      b[it] = rb;
    }
  }
  free(S);
  free(B);
  free(A);
  free(F);
  free(fix);
  free(pos);
  return st;
}
