/*
 * FP64 CPU oracle for the 2-D Q1 x Nedelec multiscale basis (test
 * infrastructure only; see ned_oracle.h for the reference files restated).
 *
 * Pinning (tests/test_ned_oracle.py):
 *   - assembly: the curl-curl block annihilates discrete gradients, the
 *     (v, grad sigma) block equals the Nedelec mass matrix times the
 *     node-to-edge incidence (grad Q1 lies in the Nedelec space), the Q1
 *     mass matrix sums to the weighted area, the load sums to the integral;
 *   - edge bases: for constant coefficients and no reaction the solution is
 *     the coarse Nedelec function itself (sigma = 0, u = interpolant of N_E),
 *     a closed form;
 *   - element matrix: the vertex/edge blocks are minus each other's transpose,
 *     the diagonal blocks are symmetric positive semidefinite, and for
 *     constant coefficients the edge block is a/|K| s_E s_F;
 *   - solver: the residual of the interior system, recomputed in the test
 *     from the exported assembled matrix.
 */
#include "ned_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* 3-point Gauss-Legendre rule on [0, 1] (QGauss(degree + 2) with degree 1,
 * q_ned_basis.cc:371) */
static const double kGx[3] = {0.11270166537925831148, 0.5, 0.88729833462074168852};
static const double kGw[3] = {0.27777777777777777778, 0.44444444444444444444,
                              0.27777777777777777778};

typedef struct {
  int m, ns, nh, nv, nu, N;
  double x0, y0, Hx, Hy, hx, hy;
} ned_mesh;

static int ned_mesh_init(int l, const double *v, ned_mesh *g) {
  g->m = 1 << l;
  g->ns = (g->m + 1) * (g->m + 1);
  g->nh = g->m * (g->m + 1);
  g->nv = (g->m + 1) * g->m;
  g->nu = g->nh + g->nv;
  g->N = g->ns + g->nu;
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

static int node_id(const ned_mesh *g, int i, int j) { return j * (g->m + 1) + i; }
static int hedge_id(const ned_mesh *g, int i, int j) { return g->ns + j * g->m + i; }
static int vedge_id(const ned_mesh *g, int i, int j) { return g->ns + g->nh + j * (g->m + 1) + i; }

/* local dofs of fine cell (i, j): 4 nodes in deal.II vertex order, then the 4
 * lines in deal.II line order (x = left, x = right, y = bottom, y = top) */
static void cell_dofs(const ned_mesh *g, int i, int j, int *d) {
  d[0] = node_id(g, i, j);
  d[1] = node_id(g, i + 1, j);
  d[2] = node_id(g, i, j + 1);
  d[3] = node_id(g, i + 1, j + 1);
  d[4] = vedge_id(g, i, j);
  d[5] = vedge_id(g, i + 1, j);
  d[6] = hedge_id(g, i, j);
  d[7] = hedge_id(g, i, j + 1);
}

/* Reference-cell shape data at (s, t) in [0,1]^2, mapped to the fine box
 * (hx, hy): Q1 values and physical gradients; lowest-order Nedelec shape
 * functions of deal.II (line 0: (0, 1-s), line 1: (0, s), line 2: (1-t, 0),
 * line 3: (t, 0)), mapped covariantly (J^-T, BasisNedelec) and their curl
 * (reference curl / det J). */
static void shape_at(double s, double t, double hx, double hy, double *q, double (*gq)[2],
                     double (*nv)[2], double *curl) {
  q[0] = (1 - s) * (1 - t);
  q[1] = s * (1 - t);
  q[2] = (1 - s) * t;
  q[3] = s * t;
  gq[0][0] = -(1 - t) / hx; gq[0][1] = -(1 - s) / hy;
  gq[1][0] = (1 - t) / hx;  gq[1][1] = -s / hy;
  gq[2][0] = -t / hx;       gq[2][1] = (1 - s) / hy;
  gq[3][0] = t / hx;        gq[3][1] = s / hy;
  const double ref[4][2] = {{0, 1 - s}, {0, s}, {1 - t, 0}, {t, 0}};
  const double rcurl[4] = {-1, 1, 1, -1};
  for (int e = 0; e < 4; ++e) {
    nv[e][0] = ref[e][0] / hx;
    nv[e][1] = ref[e][1] / hy;
    curl[e] = rcurl[e] / (hx * hy);
  }
}

/* coarse bilinear basis n and its gradient at the physical point (x, y) */
static double coarse_phi(const ned_mesh *g, int n, double x, double y, double *grad) {
  const double X = (x - g->x0) / g->Hx, Y = (y - g->y0) / g->Hy;
  const double px[2] = {1 - X, X}, py[2] = {1 - Y, Y};
  const double dx[2] = {-1 / g->Hx, 1 / g->Hx}, dy[2] = {-1 / g->Hy, 1 / g->Hy};
  const int bx = n & 1, by = (n >> 1) & 1;
  if (grad) {
    grad[0] = dx[bx] * py[by];
    grad[1] = px[bx] * dy[by];
  }
  return px[bx] * py[by];
}

/* assemble_system restated (q_ned_basis.cc:420-544): full matrix A (row =
 * test), global load F = (v, f) and the vertex-basis sources S[n] = (v, grad phi_n) */
static void ned_assemble(const ned_mesh *g, const double *a, const double *binv, const double *r,
                         const double *f, double *A, double *F, double *S) {
  const int N = g->N, m = g->m;
  memset(A, 0, sizeof(double) * (size_t)N * N);
  if (F) memset(F, 0, sizeof(double) * N);
  if (S) memset(S, 0, sizeof(double) * 4 * N);
  const double jxw0 = g->hx * g->hy;
  for (int j = 0; j < m; ++j)
    for (int i = 0; i < m; ++i) {
      const int c = j * m + i;
      int d[8];
      cell_dofs(g, i, j, d);
      double loc[8][8] = {{0}}, lf[8] = {0}, ls[4][8] = {{0}};
      for (int qy = 0; qy < 3; ++qy)
        for (int qx = 0; qx < 3; ++qx) {
          const double s = kGx[qx], t = kGx[qy], w = kGw[qx] * kGw[qy] * jxw0;
          double q[4], gq[4][2], nv[4][2], curl[4];
          shape_at(s, t, g->hx, g->hy, q, gq, nv, curl);
          /* test i, trial j: tau B^-1 sigma - grad tau . u + v . grad sigma
           * + curl v A curl u + v R u (q_ned_basis.cc:483-491) */
          for (int p = 0; p < 4; ++p) {
            for (int k = 0; k < 4; ++k) loc[p][k] += q[p] * binv[c] * q[k] * w;
            for (int e = 0; e < 4; ++e)
              loc[p][4 + e] -= (gq[p][0] * nv[e][0] + gq[p][1] * nv[e][1]) * w;
          }
          for (int e = 0; e < 4; ++e) {
            for (int k = 0; k < 4; ++k)
              loc[4 + e][k] += (nv[e][0] * gq[k][0] + nv[e][1] * gq[k][1]) * w;
            for (int h = 0; h < 4; ++h)
              loc[4 + e][4 + h] += (curl[e] * a[c] * curl[h] +
                                    r[c] * (nv[e][0] * nv[h][0] + nv[e][1] * nv[h][1])) * w;
            lf[4 + e] += (nv[e][0] * f[2 * c] + nv[e][1] * f[2 * c + 1]) * w;
            const double x = g->x0 + (i + s) * g->hx, y = g->y0 + (j + t) * g->hy;
            for (int n = 0; n < 4; ++n) {
              double gp[2];
              coarse_phi(g, n, x, y, gp);
              ls[n][4 + e] += (nv[e][0] * gp[0] + nv[e][1] * gp[1]) * w;
            }
          }
        }
      for (int p = 0; p < 8; ++p) {
        for (int k = 0; k < 8; ++k) A[(size_t)d[p] * N + d[k]] += loc[p][k];
        if (F) F[d[p]] += lf[p];
        if (S)
          for (int n = 0; n < 4; ++n) S[(size_t)n * N + d[p]] += ls[n][p];
      }
    }
}

/* Boundary values of the 8 bases (setup_basis_dofs_h1/_curl):
 * vertex basis n: sigma = phi_n at the boundary nodes; the tangential
 * projection of grad phi_n onto the lowest-order edge trace is its line
 * integral phi_n(end) - phi_n(start);
 * edge basis E: sigma = 0; u = the line integral of N_E . t, which is
 * |e| / |E| = 1 / m on the fine edges of coarse edge E and 0 elsewhere. */
static void ned_boundary(const ned_mesh *g, double *G, unsigned char *isb) {
  const int N = g->N, m = g->m;
  memset(G, 0, sizeof(double) * 8 * N);
  memset(isb, 0, N);
  for (int j = 0; j <= m; ++j)
    for (int i = 0; i <= m; ++i)
      if (i == 0 || j == 0 || i == m || j == m) {
        const int id = node_id(g, i, j);
        isb[id] = 1;
        for (int n = 0; n < 4; ++n)
          G[(size_t)n * N + id] = coarse_phi(g, n, g->x0 + i * g->hx, g->y0 + j * g->hy, NULL);
      }
  for (int j = 0; j <= m; j += m)  /* horizontal boundary edges y = y0, y1 */
    for (int i = 0; i < m; ++i) {
      const int id = hedge_id(g, i, j);
      isb[id] = 1;
      const double y = g->y0 + j * g->hy;
      for (int n = 0; n < 4; ++n)
        G[(size_t)n * N + id] = coarse_phi(g, n, g->x0 + (i + 1) * g->hx, y, NULL) -
                                coarse_phi(g, n, g->x0 + i * g->hx, y, NULL);
      G[(size_t)(4 + (j == 0 ? 2 : 3)) * N + id] = 1.0 / m;
    }
  for (int j = 0; j < m; ++j)  /* vertical boundary edges x = x0, x1 */
    for (int i = 0; i <= m; i += m) {
      const int id = vedge_id(g, i, j);
      isb[id] = 1;
      const double x = g->x0 + i * g->hx;
      for (int n = 0; n < 4; ++n)
        G[(size_t)n * N + id] = coarse_phi(g, n, x, g->y0 + (j + 1) * g->hy, NULL) -
                                coarse_phi(g, n, x, g->y0 + j * g->hy, NULL);
      G[(size_t)(4 + (i == 0 ? 0 : 1)) * N + id] = 1.0 / m;
    }
}

/* dense LU with partial pivoting (the stand-in for SparseDirectUMFPACK,
 * q_ned_basis.cc:604-607): factor in place, row interchanges in piv */
static int lu_factor(int n, double *A, int *piv) {
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = fabs(A[(size_t)k * n + k]);
    for (int i = k + 1; i < n; ++i)
      if (fabs(A[(size_t)i * n + k]) > best) {
        best = fabs(A[(size_t)i * n + k]);
        p = i;
      }
    piv[k] = p;
    if (!(best > 0)) return 2;
    if (p != k)
      for (int c = 0; c < n; ++c) {
        const double tmp = A[(size_t)k * n + c];
        A[(size_t)k * n + c] = A[(size_t)p * n + c];
        A[(size_t)p * n + c] = tmp;
      }
    const double inv = 1.0 / A[(size_t)k * n + k];
    for (int i = k + 1; i < n; ++i) {
      const double lik = A[(size_t)i * n + k] * inv;
      A[(size_t)i * n + k] = lik;
      if (lik == 0.0) continue;
      for (int c = k + 1; c < n; ++c) A[(size_t)i * n + c] -= lik * A[(size_t)k * n + c];
    }
  }
  return 0;
}

/* B (n x nrhs, row-major) <- A^-1 B with the factors of lu_factor (whole
 * rows are interchanged there, so all interchanges come first) */
static void lu_apply(int n, const double *A, const int *piv, int nrhs, double *B) {
  for (int k = 0; k < n; ++k)
    if (piv[k] != k)
      for (int c = 0; c < nrhs; ++c) {
        const double tmp = B[(size_t)k * nrhs + c];
        B[(size_t)k * nrhs + c] = B[(size_t)piv[k] * nrhs + c];
        B[(size_t)piv[k] * nrhs + c] = tmp;
      }
  for (int k = 0; k < n; ++k) {
    for (int i = k + 1; i < n; ++i) {
      const double lik = A[(size_t)i * n + k];
      if (lik == 0.0) continue;
      for (int c = 0; c < nrhs; ++c) B[(size_t)i * nrhs + c] -= lik * B[(size_t)k * nrhs + c];
    }
  }
  for (int k = n - 1; k >= 0; --k)
    for (int c = 0; c < nrhs; ++c) {
      double s = B[(size_t)k * nrhs + c];
      for (int j = k + 1; j < n; ++j) s -= A[(size_t)k * n + j] * B[(size_t)j * nrhs + c];
      B[(size_t)k * nrhs + c] = s / A[(size_t)k * n + k];
    }
}

/* A x = B for nrhs columns: symmetric diagonal equilibration (the sigma and
 * u rows differ by powers of the fine mesh size), LU with partial pivoting,
 * then 3 steps of iterative refinement with FP64 residuals.  B is overwritten
 * with x. */
static int dense_solve(int n, const double *A, int nrhs, double *B) {
  double *s = malloc(sizeof(double) * n);
  double *F = malloc(sizeof(double) * (size_t)n * n);
  double *R = malloc(sizeof(double) * (size_t)n * nrhs);
  double *X = malloc(sizeof(double) * (size_t)n * nrhs);
  int *piv = malloc(sizeof(int) * n);
  for (int i = 0; i < n; ++i) {
    const double dii = fabs(A[(size_t)i * n + i]);
    s[i] = dii > 0 ? 1.0 / sqrt(dii) : 1.0;
  }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) F[(size_t)i * n + j] = s[i] * A[(size_t)i * n + j] * s[j];
  int st = lu_factor(n, F, piv);
  if (st == 0) {
    memset(X, 0, sizeof(double) * (size_t)n * nrhs);
    for (int pass = 0; pass < 4; ++pass) {
      /* R = s (B - A X) */
      for (int i = 0; i < n; ++i)
        for (int c = 0; c < nrhs; ++c) {
          double acc = B[(size_t)i * nrhs + c];
          for (int j = 0; j < n; ++j) acc -= A[(size_t)i * n + j] * X[(size_t)j * nrhs + c];
          R[(size_t)i * nrhs + c] = s[i] * acc;
        }
      lu_apply(n, F, piv, nrhs, R);
      for (int i = 0; i < n; ++i)
        for (int c = 0; c < nrhs; ++c) X[(size_t)i * nrhs + c] += s[i] * R[(size_t)i * nrhs + c];
    }
    memcpy(B, X, sizeof(double) * (size_t)n * nrhs);
  }
  free(piv);
  free(X);
  free(R);
  free(F);
  free(s);
  return st;
}

int oq_ned_assemble(int l, const double *verts, const double *a, const double *binv,
                    const double *r, const double *f, double *A, double *F) {
  ned_mesh g;
  if (ned_mesh_init(l, verts, &g)) return 1;
  ned_assemble(&g, a, binv, r, f, A, F, NULL);
  return 0;
}

int oq_ned_basis_cell(int l, const double *verts, const double *a, const double *binv,
                      const double *r, const double *f, int full_rhs, double *M, double *b,
                      double *Psi) {
  ned_mesh g;
  if (ned_mesh_init(l, verts, &g)) return 1;
  const int N = g.N;
  double *A = malloc(sizeof(double) * (size_t)N * N);
  double *F = malloc(sizeof(double) * N);
  double *S = malloc(sizeof(double) * 4 * N);
  double *G = malloc(sizeof(double) * 8 * N);
  unsigned char *isb = malloc(N);
  ned_assemble(&g, a, binv, r, f, A, F, S);
  ned_boundary(&g, G, isb);
  int *ii = malloc(sizeof(int) * N);
  int ni = 0;
  for (int k = 0; k < N; ++k)
    if (!isb[k]) ii[ni++] = k;
  /* condense (q_ned_basis.cc:1335-1362): rhs_I = S_n - A_IB g_B */
  double *AI = malloc(sizeof(double) * (size_t)ni * ni);
  double *B = malloc(sizeof(double) * (size_t)ni * 8);
  for (int p = 0; p < ni; ++p) {
    const double *row = A + (size_t)ii[p] * N;
    for (int q = 0; q < ni; ++q) AI[(size_t)p * ni + q] = row[ii[q]];
    for (int n = 0; n < 8; ++n) {
      double s = (n < 4 && full_rhs) ? S[(size_t)n * N + ii[p]] : 0.0;
      for (int k = 0; k < N; ++k)
        if (isb[k]) s -= row[k] * G[(size_t)n * N + k];
      B[(size_t)p * 8 + n] = s;
    }
  }
  const int st = dense_solve(ni, AI, 8, B);
  /* distribute: psi = g on the boundary, the solution inside */
  double *P = malloc(sizeof(double) * 8 * N);
  memcpy(P, G, sizeof(double) * 8 * N);
  for (int p = 0; p < ni; ++p)
    for (int n = 0; n < 8; ++n) P[(size_t)n * N + ii[p]] = B[(size_t)p * 8 + n];
  /* assemble_global_element_matrix (q_ned_basis.cc:841-939): one block of the
   * assembled matrix per pair, sigma block for vertex bases, u block for edge
   * bases; the right-hand side only for the edge bases */
  double *tmp = malloc(sizeof(double) * N);
  for (int kt = 0; kt < 8; ++kt) {
    const int r0 = kt < 4 ? 0 : g.ns, r1 = kt < 4 ? g.ns : N;
    for (int kc = 0; kc < 8; ++kc) {
      const int c0 = kc < 4 ? 0 : g.ns, c1 = kc < 4 ? g.ns : N;
      double s = 0.0;
      for (int ro = r0; ro < r1; ++ro) {
        double t = 0.0;
        for (int co = c0; co < c1; ++co) t += A[(size_t)ro * N + co] * P[(size_t)kc * N + co];
        s += P[(size_t)kt * N + ro] * t;
      }
      M[kt * 8 + kc] = s;
    }
    double sb = 0.0;
    if (kt >= 4)
      for (int ro = g.ns; ro < N; ++ro) sb += P[(size_t)kt * N + ro] * F[ro];
    b[kt] = sb;
  }
  if (Psi) memcpy(Psi, P, sizeof(double) * 8 * N);
  free(tmp);
  free(P);
  free(B);
  free(AI);
  free(ii);
  free(isb);
  free(G);
  free(S);
  free(F);
  free(A);
  return st;
}
