/*
 * FP64 CPU oracle for the multiscale Q1 basis construction -- TEST
 * INFRASTRUCTURE ONLY (see msfec_oracle.h for the reference files it restates
 * and for who may load it).
 */
#include "msfec_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OQ_PI 3.14159265358979323846

static int ipow(int b, int e) {
  int r = 1;
  while (e-- > 0) r *= b;
  return r;
}

/* ---------------------------------------------------------------------------
 * BasisQ1 -- include/functions/basis_q1.tpp:21-67 (constructor inverts the
 * vertex-monomial matrix with FullMatrix::invert) and :94-152 (value_list).
 * Pinned by: delta property at vertices, partition of unity, closed-form
 * tensor-product Lagrange values on boxes (tests/test_oracle.py).
 * ------------------------------------------------------------------------- */
static void monomials(int dim, const double *x, double *m) {
  if (dim == 2) {
    m[0] = 1.0;
    m[1] = x[0];
    m[2] = x[1];
    m[3] = x[0] * x[1];
  } else {
    m[0] = 1.0;
    m[1] = x[0];
    m[2] = x[1];
    m[3] = x[2];
    m[4] = x[0] * x[1];
    m[5] = x[1] * x[2];
    m[6] = x[0] * x[2];
    m[7] = x[0] * x[1] * x[2];
  }
}

/* Gauss-Jordan inversion with partial pivoting of an n x n row-major matrix. */
static int invert(int n, const double *A, double *Ainv) {
  double w[8 * 16];
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) {
      w[i * 2 * n + j] = A[i * n + j];
      w[i * 2 * n + n + j] = (i == j) ? 1.0 : 0.0;
    }
  }
  for (int c = 0; c < n; ++c) {
    int piv = c;
    for (int r = c + 1; r < n; ++r)
      if (fabs(w[r * 2 * n + c]) > fabs(w[piv * 2 * n + c])) piv = r;
    if (w[piv * 2 * n + c] == 0.0) return -1;
    if (piv != c)
      for (int j = 0; j < 2 * n; ++j) {
        double t = w[c * 2 * n + j];
        w[c * 2 * n + j] = w[piv * 2 * n + j];
        w[piv * 2 * n + j] = t;
      }
    double d = w[c * 2 * n + c];
    for (int j = 0; j < 2 * n; ++j) w[c * 2 * n + j] /= d;
    for (int r = 0; r < n; ++r) {
      if (r == c) continue;
      double fct = w[r * 2 * n + c];
      if (fct == 0.0) continue;
      for (int j = 0; j < 2 * n; ++j) w[r * 2 * n + j] -= fct * w[c * 2 * n + j];
    }
  }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) Ainv[i * n + j] = w[i * 2 * n + n + j];
  return 0;
}

int oq_basis_q1_coeffs(int dim, const double *verts, double *C) {
  const int nv = 1 << dim;
  double P[64];
  for (int i = 0; i < nv; ++i) monomials(dim, verts + i * dim, P + i * nv);
  /* coeff_matrix.invert(point_matrix): columns of C are the coefficients */
  return invert(nv, P, C);
}

double oq_basis_q1_value(int dim, const double *C, int index, const double *x) {
  const int nv = 1 << dim;
  double m[8];
  monomials(dim, x, m);
  double v = 0.0;
  for (int k = 0; k < nv; ++k) v += C[k * nv + index] * m[k];
  return v;
}

/* ---------------------------------------------------------------------------
 * Diffusion_A -- source/equation_data/eqn_coeff_A.cc:7-57 (rotation from the
 * fixed Euler angles alpha=pi/3, beta=pi/6, gamma=pi/4) and :163-191.
 * Pinned by: orthogonality of the rotation, eigenvalues equal to the diagonal
 * entries, rotate=false giving the diagonal (tests/test_oracle.py).
 * ------------------------------------------------------------------------- */
static void rotation(int rotate, double R[3][3]) {
  const double a = OQ_PI / 3, b = OQ_PI / 6, g = OQ_PI / 4;
  if (!rotate) {
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) R[i][j] = (i == j) ? 1.0 : 0.0;
    return;
  }
  R[0][0] = cos(a) * cos(g) - sin(a) * cos(b) * sin(g);
  R[0][1] = -cos(a) * sin(g) - sin(a) * cos(b) * cos(g);
  R[0][2] = sin(a) * sin(b);
  R[1][0] = sin(a) * cos(g) + cos(a) * cos(b) * sin(g);
  R[1][1] = -sin(a) * sin(g) + cos(a) * cos(b) * cos(g);
  R[1][2] = -cos(a) * sin(b);
  R[2][0] = sin(b) * sin(g);
  R[2][1] = sin(b) * cos(g);
  R[2][2] = cos(b);
}

void oq_diffusion_A(const oq_problem *p, const double *x, double *A) {
  double D[3][3] = {{0}};
  for (int d = 0; d < 3; ++d)
    D[d][d] = p->scale[d] * (1.0 - p->alpha[d] * sin(2 * OQ_PI * p->freq[d] * x[d]));
  if (!p->rotate) {
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) A[i * 3 + j] = D[i][j];
    return;
  }
  double R[3][3], T[3][3];
  rotation(1, R);
  /* value = rot * value * transpose(rot) */
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += R[i][k] * D[k][j];
      T[i][j] = s;
    }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += T[i][k] * R[j][k];
      A[i * 3 + j] = s;
    }
}

/* ---------------------------------------------------------------------------
 * Fine element matrix by QGauss(2) -- source/Q/q_basis.cc:177-237.
 * FEValues on an axis-aligned box: shape values are tensor products of 1-D
 * linear Lagrange functions; JxW = w_q * prod(h).  Pinned by the closed-form
 * Q1 stiffness (a h/12) S in 3-D and (a/6) S in 2-D, the Q1 mass matrix, and a
 * numpy brute-force 3-point Gauss assembly (tests/test_oracle.py and tests/golden/make_golden.py).
 * ------------------------------------------------------------------------- */
static const double GQ[2] = {0.5 - 0.28867513459481288225, 0.5 + 0.28867513459481288225};

static void sym_to_full(int dim, const double *s, double *A) {
  if (dim == 2) {
    A[0] = s[0]; A[1] = s[1];
    A[2] = s[1]; A[3] = s[2];
  } else {
    A[0] = s[0]; A[1] = s[1]; A[2] = s[2];
    A[3] = s[1]; A[4] = s[3]; A[5] = s[4];
    A[6] = s[2]; A[7] = s[4]; A[8] = s[5];
  }
}

void oq_element_matrix(int dim, const double *h, const double *Aq,
                       const double *fq, double rq, double *Ke, double *Fe) {
  const int nv = 1 << dim, nq = 1 << dim, nc = (dim == 2) ? 3 : 6;
  double vol = 1.0;
  for (int d = 0; d < dim; ++d) vol *= h[d];
  for (int i = 0; i < nv * nv; ++i) Ke[i] = 0.0;
  for (int i = 0; i < nv; ++i) Fe[i] = 0.0;
  for (int q = 0; q < nq; ++q) {
    double xi[3], w = 1.0;
    for (int d = 0; d < dim; ++d) {
      xi[d] = GQ[(q >> d) & 1];
      w *= 0.5;
    }
    const double JxW = w * vol;
    double val[8], grad[8][3];
    for (int a = 0; a < nv; ++a) {
      double v = 1.0;
      for (int d = 0; d < dim; ++d) v *= ((a >> d) & 1) ? xi[d] : 1.0 - xi[d];
      val[a] = v;
      for (int d = 0; d < dim; ++d) {
        double g = (((a >> d) & 1) ? 1.0 : -1.0) / h[d];
        for (int e = 0; e < dim; ++e)
          if (e != d) g *= ((a >> e) & 1) ? xi[e] : 1.0 - xi[e];
        grad[a][d] = g;
      }
    }
    double A[9];
    sym_to_full(dim, Aq + q * nc, A);
    for (int i = 0; i < nv; ++i) {
      for (int j = 0; j < nv; ++j) {
        double AgJ[3];
        for (int r = 0; r < dim; ++r) {
          double s = 0.0;
          for (int c = 0; c < dim; ++c) s += A[r * dim + c] * grad[j][c];
          AgJ[r] = s;
        }
        double s = 0.0;
        for (int r = 0; r < dim; ++r) s += grad[i][r] * AgJ[r];
        Ke[i * nv + j] += s * JxW + rq * val[i] * val[j] * JxW;
      }
      Fe[i] += val[i] * fq[q] * JxW;
    }
  }
}

/* ---------------------------------------------------------------------------
 * Per-cell basis construction -- source/Q/q_basis.cc:518-603.
 * ------------------------------------------------------------------------- */
typedef struct {
  int dim, N, S, nn, nv, nst;
  double lo[3], h[3];
  double *K;  /* stencil: nn x 3^dim, direction (dx+1)+3(dy+1)+9(dz+1) */
  double *F;  /* nn */
} oq_fine;

static int node_coord(const oq_fine *m, int node, int *ijk) {
  int r = node;
  for (int d = 0; d < m->dim; ++d) {
    ijk[d] = r % m->S;
    r /= m->S;
  }
  return 0;
}

static int is_boundary(const oq_fine *m, const int *ijk) {
  for (int d = 0; d < m->dim; ++d)
    if (ijk[d] == 0 || ijk[d] == m->N) return 1;
  return 0;
}

static int dir_index(int dim, const int *dd) {
  int e = 0, mul = 1;
  for (int d = 0; d < dim; ++d) {
    e += (dd[d] + 1) * mul;
    mul *= 3;
  }
  return e;
}

/* assemble_system, q_basis.cc:175-257: loop over fine cells, QGauss(2) cell
 * matrices, scatter into the (unconstrained) diffusion matrix and global_rhs. */
static int assemble(const oq_problem *p, oq_fine *m, const double *coef,
                    const double *f, const double *reac) {
  const int dim = m->dim, N = m->N, nv = m->nv, nq = nv;
  const int nc = (dim == 2) ? 3 : 6;
  const int ncell = ipow(N, dim);
  double Aq[8 * 6], fq[8], Ke[64], Fe[8];
  for (int e = 0; e < ncell; ++e) {
    int cijk[3] = {0, 0, 0}, r = e;
    for (int d = 0; d < dim; ++d) {
      cijk[d] = r % N;
      r /= N;
    }
    /* coefficient at the quadrature points */
    for (int q = 0; q < nq; ++q) {
      if (p->coef_mode == OQ_COEF_SCALAR_CELL) {
        const double a = coef[e];
        for (int c = 0; c < nc; ++c) Aq[q * nc + c] = 0.0;
        if (dim == 2) {
          Aq[q * nc + 0] = a;
          Aq[q * nc + 2] = a;
        } else {
          Aq[q * nc + 0] = a;
          Aq[q * nc + 3] = a;
          Aq[q * nc + 5] = a;
        }
      } else if (p->coef_mode == OQ_COEF_TENSOR_QP) {
        for (int c = 0; c < nc; ++c) Aq[q * nc + c] = coef[((size_t)e * nq + q) * nc + c];
      } else if (p->coef_mode == OQ_COEF_ANALYTIC) {
        if (dim != 3) return -1;
        double x[3], A[9];
        for (int d = 0; d < 3; ++d) x[d] = m->lo[d] + (cijk[d] + GQ[(q >> d) & 1]) * m->h[d];
        oq_diffusion_A(p, x, A);
        Aq[q * nc + 0] = A[0];
        Aq[q * nc + 1] = A[1];
        Aq[q * nc + 2] = A[2];
        Aq[q * nc + 3] = A[4];
        Aq[q * nc + 4] = A[5];
        Aq[q * nc + 5] = A[8];
      } else {
        return -1;
      }
      if (p->rhs_mode == OQ_RHS_CONST) fq[q] = p->rhs_const;
      else if (p->rhs_mode == OQ_RHS_CELL) fq[q] = f[e];
      else if (p->rhs_mode == OQ_RHS_QP) fq[q] = f[(size_t)e * nq + q];
      else return -1;
    }
    double rq = 0.0;
    if (p->reaction_mode == OQ_REACTION_CONST) rq = p->reaction_const;
    else if (p->reaction_mode == OQ_REACTION_CELL) rq = reac[e];
    oq_element_matrix(dim, m->h, Aq, fq, rq, Ke, Fe);
    for (int a = 0; a < nv; ++a) {
      int na = 0, mul = 1;
      for (int d = 0; d < dim; ++d) {
        na += (cijk[d] + ((a >> d) & 1)) * mul;
        mul *= m->S;
      }
      for (int b = 0; b < nv; ++b) {
        int dd[3];
        for (int d = 0; d < dim; ++d) dd[d] = ((b >> d) & 1) - ((a >> d) & 1);
        m->K[(size_t)na * m->nst + dir_index(dim, dd)] += Ke[a * nv + b];
      }
      m->F[na] += Fe[a];
    }
  }
  return 0;
}

/* y = K x with the full (unconstrained) stencil matrix */
static void stencil_apply(const oq_fine *m, const double *x, double *y) {
  const int dim = m->dim;
  for (int n = 0; n < m->nn; ++n) {
    int ijk[3] = {0, 0, 0};
    node_coord(m, n, ijk);
    double s = 0.0;
    for (int e = 0; e < m->nst; ++e) {
      int dd[3] = {0, 0, 0}, r = e, ok = 1, off = 0, mul = 1;
      for (int d = 0; d < dim; ++d) {
        dd[d] = r % 3 - 1;
        r /= 3;
        int c = ijk[d] + dd[d];
        if (c < 0 || c > m->N) ok = 0;
        off += dd[d] * mul;
        mul *= m->S;
      }
      if (ok) s += m->K[(size_t)n * m->nst + e] * x[n + off];
    }
    y[n] = s;
  }
}

/* Dense Cholesky of the interior block K_II (the solve_direct analogue,
 * q_basis.cc:304-330: UMFPACK of the condensed matrix yields the same
 * interior solution K_II x_I = -K_IB g_B). Pinned by scipy cho_solve golden
 * vectors and the A = aI closed form psi_i = phi_i. */
/* ... and, when w != NULL, w = K_II^{-1} F_I on the nodes (0 on the boundary),
 * the extra solve of the adjoint (the derivative of b = Psi^T F) */
static int solve_direct_w(const oq_fine *m, const double *g, double *Psi, double *w) {
  const int dim = m->dim, ni = m->N - 1, nb = m->nv;
  const int n = ipow(ni, dim);
  if (n == 0) {
    for (int k = 0; k < nb; ++k) memcpy(Psi + (size_t)k * m->nn, g + (size_t)k * m->nn, sizeof(double) * m->nn);
    if (w)
      for (int v = 0; v < m->nn; ++v) w[v] = 0.0;
    return 0;
  }
  int *inode = (int *)malloc(sizeof(int) * n);
  int *iidx = (int *)malloc(sizeof(int) * m->nn);
  double *L = (double *)calloc((size_t)n * n, sizeof(double));
  double *R = (double *)calloc((size_t)n * nb, sizeof(double));
  int status = 0;
  for (int v = 0; v < m->nn; ++v) iidx[v] = -1;
  for (int i = 0; i < n; ++i) {
    int r = i, node = 0, mul = 1;
    for (int d = 0; d < dim; ++d) {
      node += (r % ni + 1) * mul;
      r /= ni;
      mul *= m->S;
    }
    inode[i] = node;
    iidx[node] = i;
  }
  /* K_II and R = -K_IB g */
  for (int i = 0; i < n; ++i) {
    const int node = inode[i];
    int ijk[3] = {0, 0, 0};
    node_coord(m, node, ijk);
    for (int e = 0; e < m->nst; ++e) {
      int r = e, off = 0, mul = 1;
      for (int d = 0; d < dim; ++d) {
        off += (r % 3 - 1) * mul;
        r /= 3;
        mul *= m->S;
      }
      const int nb_node = node + off;
      const double kv = m->K[(size_t)node * m->nst + e];
      const int j = iidx[nb_node];
      if (j >= 0) {
        L[(size_t)i * n + j] += kv;
      } else {
        for (int k = 0; k < nb; ++k) R[(size_t)k * n + i] -= kv * g[(size_t)k * m->nn + nb_node];
      }
    }
  }
  /* Cholesky L L^T, lower triangle in place (row-major) */
  for (int j = 0; j < n && !status; ++j) {
    double s = L[(size_t)j * n + j];
    for (int k = 0; k < j; ++k) s -= L[(size_t)j * n + k] * L[(size_t)j * n + k];
    if (!(s > 0.0)) {
      status = 2;
      break;
    }
    const double djj = sqrt(s);
    L[(size_t)j * n + j] = djj;
    for (int i = j + 1; i < n; ++i) {
      double t = L[(size_t)i * n + j];
      const double *li = L + (size_t)i * n, *lj = L + (size_t)j * n;
      for (int k = 0; k < j; ++k) t -= li[k] * lj[k];
      L[(size_t)i * n + j] = t / djj;
    }
  }
  if (!status) {
    for (int k = 0; k < nb; ++k) {
      double *x = R + (size_t)k * n;
      for (int i = 0; i < n; ++i) {
        double t = x[i];
        for (int c = 0; c < i; ++c) t -= L[(size_t)i * n + c] * x[c];
        x[i] = t / L[(size_t)i * n + i];
      }
      for (int i = n - 1; i >= 0; --i) {
        double t = x[i];
        for (int r = i + 1; r < n; ++r) t -= L[(size_t)r * n + i] * x[r];
        x[i] = t / L[(size_t)i * n + i];
      }
      /* distribute: boundary entries = g, interior = x */
      for (int v = 0; v < m->nn; ++v)
        Psi[(size_t)k * m->nn + v] = (iidx[v] >= 0) ? x[iidx[v]] : g[(size_t)k * m->nn + v];
    }
    if (w) {
      double *x = (double *)malloc(sizeof(double) * n);
      for (int i = 0; i < n; ++i) x[i] = m->F[inode[i]];
      for (int i = 0; i < n; ++i) {
        double t = x[i];
        for (int c = 0; c < i; ++c) t -= L[(size_t)i * n + c] * x[c];
        x[i] = t / L[(size_t)i * n + i];
      }
      for (int i = n - 1; i >= 0; --i) {
        double t = x[i];
        for (int r = i + 1; r < n; ++r) t -= L[(size_t)r * n + i] * x[r];
        x[i] = t / L[(size_t)i * n + i];
      }
      for (int v = 0; v < m->nn; ++v) w[v] = (iidx[v] >= 0) ? x[iidx[v]] : 0.0;
      free(x);
    }
  }
  free(inode);
  free(iidx);
  free(L);
  free(R);
  return status;
}

/* CG + SSOR(omega) on the condensed full system -- q_basis.cc:335-357 with
 * SolverControl(1000, 1e-12) (absolute residual) and PreconditionSSOR 1.6.
 * AffineConstraints::condense keeps the constrained rows with their diagonal
 * entry and right-hand side diag*g (inhomogeneous Dirichlet data), and moves
 * the interior-boundary coupling to the right-hand side.  SSOR follows
 * SparseMatrix::precondition_SSOR: forward sweep, diagonal scaling by
 * omega(2-omega)a_ii, backward sweep.  DoFs are in lexicographic order here
 * (deal.II numbers them by cell traversal), so iteration counts may differ
 * slightly from the reference while the converged solution agrees to the
 * stopping tolerance. */
static int solve_cg_ssor(const oq_problem *p, const oq_fine *m, const double *g,
                         double *Psi, int *iters) {
  const int dim = m->dim, nn = m->nn, nst = m->nst, nb = m->nv;
  const int center = (nst - 1) / 2;
  const double om = p->ssor_omega;
  double *Kc = (double *)calloc((size_t)nn * nst, sizeof(double));
  int *offs = (int *)malloc(sizeof(int) * nst);
  int *valid = (int *)calloc((size_t)nn * nst, sizeof(int));
  double *rhs = (double *)malloc(sizeof(double) * nn);
  double *x = (double *)malloc(sizeof(double) * nn);
  double *r = (double *)malloc(sizeof(double) * nn);
  double *z = (double *)malloc(sizeof(double) * nn);
  double *pv = (double *)malloc(sizeof(double) * nn);
  double *q = (double *)malloc(sizeof(double) * nn);
  int status = 0, maxits = 0;
  for (int e = 0; e < nst; ++e) {
    int rr = e, off = 0, mul = 1;
    for (int d = 0; d < dim; ++d) {
      off += (rr % 3 - 1) * mul;
      rr /= 3;
      mul *= m->S;
    }
    offs[e] = off;
  }
  /* condensed matrix: boundary rows diagonal only; interior rows without
   * couplings to boundary DoFs */
  for (int n = 0; n < nn; ++n) {
    int ijk[3] = {0, 0, 0};
    node_coord(m, n, ijk);
    const int bn = is_boundary(m, ijk);
    for (int e = 0; e < nst; ++e) {
      int rr = e, ok = 1, nijk[3] = {0, 0, 0};
      for (int d = 0; d < dim; ++d) {
        int c = ijk[d] + rr % 3 - 1;
        rr /= 3;
        if (c < 0 || c > m->N) ok = 0;
        nijk[d] = c;
      }
      if (!ok) continue;
      valid[(size_t)n * nst + e] = 1;
      if (bn) {
        if (e == center) Kc[(size_t)n * nst + e] = m->K[(size_t)n * nst + e];
      } else if (!is_boundary(m, nijk)) {
        Kc[(size_t)n * nst + e] = m->K[(size_t)n * nst + e];
      }
    }
  }
  for (int k = 0; k < nb && !status; ++k) {
    const double *gk = g + (size_t)k * nn;
    for (int n = 0; n < nn; ++n) {
      int ijk[3] = {0, 0, 0};
      node_coord(m, n, ijk);
      if (is_boundary(m, ijk)) {
        rhs[n] = m->K[(size_t)n * nst + center] * gk[n];
      } else {
        double s = 0.0;
        for (int e = 0; e < nst; ++e) {
          if (!valid[(size_t)n * nst + e]) continue;
          int nbn = n + offs[e], nijk[3] = {0, 0, 0};
          node_coord(m, nbn, nijk);
          if (is_boundary(m, nijk)) s += m->K[(size_t)n * nst + e] * gk[nbn];
        }
        rhs[n] = -s;
      }
      x[n] = 0.0;
    }
    /* r = rhs - A x (x = 0) */
    double rnorm2 = 0.0;
    for (int n = 0; n < nn; ++n) {
      r[n] = rhs[n];
      rnorm2 += r[n] * r[n];
    }
    int it = 0;
    double rz_old = 0.0;
    while (sqrt(rnorm2) > p->cg_tol) {
      if (it >= p->cg_maxit) {
        status = 1;
        break;
      }
      /* z = SSOR^{-1} r */
      for (int n = 0; n < nn; ++n) {
        double s = 0.0;
        for (int e = 0; e < center; ++e)
          if (valid[(size_t)n * nst + e]) s += Kc[(size_t)n * nst + e] * z[n + offs[e]];
        z[n] = (r[n] - om * s) / Kc[(size_t)n * nst + center];
      }
      for (int n = 0; n < nn; ++n) z[n] *= om * (2.0 - om) * Kc[(size_t)n * nst + center];
      for (int n = nn - 1; n >= 0; --n) {
        double s = 0.0;
        for (int e = center + 1; e < nst; ++e)
          if (valid[(size_t)n * nst + e]) s += Kc[(size_t)n * nst + e] * z[n + offs[e]];
        z[n] = (z[n] - om * s) / Kc[(size_t)n * nst + center];
      }
      double rz = 0.0;
      for (int n = 0; n < nn; ++n) rz += r[n] * z[n];
      if (it == 0) {
        for (int n = 0; n < nn; ++n) pv[n] = z[n];
      } else {
        const double beta = rz / rz_old;
        for (int n = 0; n < nn; ++n) pv[n] = z[n] + beta * pv[n];
      }
      rz_old = rz;
      double pq = 0.0;
      for (int n = 0; n < nn; ++n) {
        double s = 0.0;
        for (int e = 0; e < nst; ++e)
          if (valid[(size_t)n * nst + e]) s += Kc[(size_t)n * nst + e] * pv[n + offs[e]];
        q[n] = s;
        pq += pv[n] * s;
      }
      const double alpha = rz / pq;
      rnorm2 = 0.0;
      for (int n = 0; n < nn; ++n) {
        x[n] += alpha * pv[n];
        r[n] -= alpha * q[n];
        rnorm2 += r[n] * r[n];
      }
      ++it;
    }
    if (it > maxits) maxits = it;
    /* distribute */
    for (int n = 0; n < nn; ++n) {
      int ijk[3] = {0, 0, 0};
      node_coord(m, n, ijk);
      Psi[(size_t)k * nn + n] = is_boundary(m, ijk) ? gk[n] : x[n];
    }
  }
  if (iters) *iters = maxits;
  free(Kc);
  free(offs);
  free(valid);
  free(rhs);
  free(x);
  free(r);
  free(z);
  free(pv);
  free(q);
  return status;
}

static int basis_cell(const oq_problem *p, const double *verts, const double *coef,
                      const double *f, const double *reac, int solver, double *M,
                      double *b, double *Psi, int *iters, double *w);

int oq_basis_cell(const oq_problem *p, const double *verts, const double *coef,
                  const double *f, const double *reac, int solver, double *M,
                  double *b, double *Psi, int *iters) {
  return basis_cell(p, verts, coef, f, reac, solver, M, b, Psi, iters, NULL);
}

static int basis_cell(const oq_problem *p, const double *verts, const double *coef,
                      const double *f, const double *reac, int solver, double *M,
                      double *b, double *Psi, int *iters, double *w) {
  const int dim = p->dim;
  if (dim != 2 && dim != 3) return -1;
  if (p->n_refine < 0 || p->n_refine > 6) return -1;
  oq_fine m;
  m.dim = dim;
  m.N = 1 << p->n_refine;
  m.S = m.N + 1;
  m.nn = ipow(m.S, dim);
  m.nv = 1 << dim;
  m.nst = ipow(3, dim);
  /* general_cell on an axis-aligned box: lower corner = vertex 0, upper
   * corner = last vertex (q_basis.cc:95-102); uniform refinement */
  for (int d = 0; d < dim; ++d) {
    m.lo[d] = verts[d];
    m.h[d] = (verts[(m.nv - 1) * dim + d] - verts[d]) / m.N;
    if (!(m.h[d] > 0.0)) return -1;
  }
  for (int v = 0; v < m.nv; ++v)
    for (int d = 0; d < dim; ++d) {
      const double want = ((v >> d) & 1) ? verts[(m.nv - 1) * dim + d] : verts[d];
      if (fabs(verts[v * dim + d] - want) > 1e-12 * (1.0 + fabs(want))) return -1;
    }
  m.K = (double *)calloc((size_t)m.nn * m.nst, sizeof(double));
  m.F = (double *)calloc((size_t)m.nn, sizeof(double));
  /* callers that only want M and b may pass Psi = NULL */
  double *psi_own = NULL;
  if (!Psi) {
    psi_own = (double *)malloc(sizeof(double) * (size_t)m.nv * m.nn);
    Psi = psi_own;
  }
  double *g = (double *)calloc((size_t)m.nv * m.nn, sizeof(double));
  double *tmp = (double *)malloc(sizeof(double) * m.nn);
  int status = assemble(p, &m, coef, f, reac);
  if (status == 0) {
    /* setup_system: interpolate_boundary_values of BasisQ1 index i onto the
     * fine boundary support points (q_basis.cc:131-149) */
    double C[64];
    if (oq_basis_q1_coeffs(dim, verts, C) != 0) status = -1;
    for (int n = 0; n < m.nn && status == 0; ++n) {
      int ijk[3] = {0, 0, 0};
      node_coord(&m, n, ijk);
      if (!is_boundary(&m, ijk)) continue;
      double x[3];
      for (int d = 0; d < dim; ++d) x[d] = m.lo[d] + ijk[d] * m.h[d];
      for (int k = 0; k < m.nv; ++k) g[(size_t)k * m.nn + n] = oq_basis_q1_value(dim, C, k, x);
    }
  }
  if (status == 0) {
    if (iters) *iters = 0;
    if (solver == OQ_SOLVER_DIRECT) status = solve_direct_w(&m, g, Psi, w);
    else status = solve_cg_ssor(p, &m, g, Psi, iters);
  }
  if (status == 0 || status == 1) {
    /* assemble_global_element_matrix (q_basis.cc:262-300): the reference
     * recomputes K psi_j for every i; one product per trial basis gives the
     * same numbers. */
    for (int j = 0; j < m.nv; ++j) {
      stencil_apply(&m, Psi + (size_t)j * m.nn, tmp);
      for (int i = 0; i < m.nv; ++i) {
        double s = 0.0;
        const double *pi = Psi + (size_t)i * m.nn;
        for (int n = 0; n < m.nn; ++n) s += pi[n] * tmp[n];
        M[i * m.nv + j] = s;
      }
    }
    for (int i = 0; i < m.nv; ++i) {
      double s = 0.0;
      const double *pi = Psi + (size_t)i * m.nn;
      for (int n = 0; n < m.nn; ++n) s += pi[n] * m.F[n];
      b[i] = s;
    }
  }
  free(m.K);
  free(m.F);
  free(g);
  free(tmp);
  free(psi_own);
  return status;
}

int oq_basis_batch(const oq_problem *p, int n_cells, const double *verts,
                   const double *coef, const double *f, const double *reac,
                   int solver, int n_threads, double *M, double *b, double *Psi,
                   int *iters, int *status) {
  const int dim = p->dim, nv = 1 << dim, N = 1 << p->n_refine;
  const int nfc = ipow(N, dim), nn = ipow(N + 1, dim), nq = nv;
  const size_t coef_stride = (p->coef_mode == OQ_COEF_SCALAR_CELL) ? (size_t)nfc
                             : (p->coef_mode == OQ_COEF_TENSOR_QP)
                                 ? (size_t)nfc * nq * ((dim == 2) ? 3 : 6)
                                 : 0;
  const size_t f_stride = (p->rhs_mode == OQ_RHS_CELL) ? (size_t)nfc
                          : (p->rhs_mode == OQ_RHS_QP) ? (size_t)nfc * nq
                                                        : 0;
  const size_t r_stride = (p->reaction_mode == OQ_REACTION_CELL) ? (size_t)nfc : 0;
  int worst = 0;
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#pragma omp parallel for schedule(dynamic, 1) reduction(max : worst)
#endif
  for (int c = 0; c < n_cells; ++c) {
    int it = 0;
    const int st = oq_basis_cell(
        p, verts + (size_t)c * nv * dim, coef ? coef + c * coef_stride : NULL,
        f ? f + c * f_stride : NULL, reac ? reac + c * r_stride : NULL, solver,
        M + (size_t)c * nv * nv, b + (size_t)c * nv,
        Psi ? Psi + (size_t)c * nv * nn : NULL, &it);
    if (status) status[c] = st;
    if (iters) iters[c] = it;
    if (st > worst) worst = st;
  }
  return worst;
}

/* ---------------------------------------------------------------------------
 * Coarse problem -- source/Q/q_global.cc.
 * Pinned by: A = aI gives the standard coarse Q1 FEM, checked against a dense
 * numpy solve of the closed-form coarse stiffness (tests/test_oracle.py).
 * ------------------------------------------------------------------------- */
static void coarse_cell_nodes(int dim, int nc, int c, int *nodes) {
  int cijk[3] = {0, 0, 0}, r = c;
  for (int d = 0; d < dim; ++d) {
    cijk[d] = r % nc;
    r /= nc;
  }
  const int nv = 1 << dim;
  for (int v = 0; v < nv; ++v) {
    int id = 0, mul = 1;
    for (int d = 0; d < dim; ++d) {
      id += (cijk[d] + ((v >> d) & 1)) * mul;
      mul *= nc + 1;
    }
    nodes[v] = id;
  }
}

static int coarse_is_boundary(int dim, int nc, int node) {
  for (int d = 0; d < dim; ++d) {
    const int c = node % (nc + 1);
    node /= nc + 1;
    if (c == 0 || c == nc) return 1;
  }
  return 0;
}

static int coarse_cg(int dim, int nc, const double *M, const double *b,
                     const double *rhs_nodes, double rtol, int maxit, double *u);

int oq_coarse_solve(int dim, int nc, const double *M, const double *b,
                    double rtol, int maxit, double *u) {
  return coarse_cg(dim, nc, M, b, NULL, rtol, maxit, u);
}

int oq_coarse_solve_rhs(int dim, int nc, const double *M, const double *rhs_nodes,
                        double rtol, int maxit, double *u) {
  return coarse_cg(dim, nc, M, NULL, rhs_nodes, rtol, maxit, u);
}

/* assembly with homogeneous Dirichlet constraints and Jacobi-PCG; the
 * right-hand side is assembled from b, or taken nodewise from rhs_nodes */
static int coarse_cg(int dim, int nc, const double *M, const double *b,
                     const double *rhs_nodes, double rtol, int maxit, double *u) {
  const int nv = 1 << dim, ncell = ipow(nc, dim), nn = ipow(nc + 1, dim);
  const int nst = ipow(3, dim);
  double *K = (double *)calloc((size_t)nn * nst, sizeof(double));
  double *rhs = (double *)calloc((size_t)nn, sizeof(double));
  double *r = (double *)malloc(sizeof(double) * nn);
  double *z = (double *)malloc(sizeof(double) * nn);
  double *pv = (double *)calloc((size_t)nn, sizeof(double));
  double *q = (double *)malloc(sizeof(double) * nn);
  int *offs = (int *)malloc(sizeof(int) * nst);
  int nodes[8];
  for (int e = 0; e < nst; ++e) {
    int rr = e, off = 0, mul = 1;
    for (int d = 0; d < dim; ++d) {
      off += (rr % 3 - 1) * mul;
      rr /= 3;
      mul *= nc + 1;
    }
    offs[e] = off;
  }
  /* distribute_local_to_global with homogeneous Dirichlet constraints */
  for (int c = 0; c < ncell; ++c) {
    coarse_cell_nodes(dim, nc, c, nodes);
    for (int a = 0; a < nv; ++a) {
      if (coarse_is_boundary(dim, nc, nodes[a])) continue;
      for (int bb = 0; bb < nv; ++bb) {
        if (coarse_is_boundary(dim, nc, nodes[bb])) continue;
        int dd[3];
        for (int d = 0; d < dim; ++d) dd[d] = ((bb >> d) & 1) - ((a >> d) & 1);
        K[(size_t)nodes[a] * nst + dir_index(dim, dd)] += M[((size_t)c * nv + a) * nv + bb];
      }
      if (b) rhs[nodes[a]] += b[(size_t)c * nv + a];
    }
  }
  if (rhs_nodes)
    for (int n = 0; n < nn; ++n) rhs[n] = rhs_nodes[n];
  const int center = (nst - 1) / 2;
  double bnorm2 = 0.0, rnorm2 = 0.0;
  for (int n = 0; n < nn; ++n) {
    u[n] = 0.0;
    r[n] = coarse_is_boundary(dim, nc, n) ? 0.0 : rhs[n];
    bnorm2 += r[n] * r[n];
  }
  rnorm2 = bnorm2;
  int it = 0;
  double rz_old = 0.0;
  const double stop = rtol * rtol * bnorm2;
  while (rnorm2 > stop && bnorm2 > 0.0) {
    if (it >= maxit) {
      it = -1;
      break;
    }
    for (int n = 0; n < nn; ++n)
      z[n] = coarse_is_boundary(dim, nc, n) ? 0.0 : r[n] / K[(size_t)n * nst + center];
    double rz = 0.0;
    for (int n = 0; n < nn; ++n) rz += r[n] * z[n];
    const double beta = (it == 0) ? 0.0 : rz / rz_old;
    for (int n = 0; n < nn; ++n) pv[n] = z[n] + beta * pv[n];
    rz_old = rz;
    double pq = 0.0;
    for (int n = 0; n < nn; ++n) {
      double s = 0.0;
      if (!coarse_is_boundary(dim, nc, n)) {
        int ijk[3] = {0, 0, 0}, rr = n;
        for (int d = 0; d < dim; ++d) {
          ijk[d] = rr % (nc + 1);
          rr /= nc + 1;
        }
        for (int e = 0; e < nst; ++e) {
          int er = e, ok = 1;
          for (int d = 0; d < dim; ++d) {
            const int cc = ijk[d] + er % 3 - 1;
            er /= 3;
            if (cc < 0 || cc > nc) ok = 0;
          }
          if (ok) s += K[(size_t)n * nst + e] * pv[n + offs[e]];
        }
      }
      q[n] = s;
      pq += pv[n] * s;
    }
    const double alpha = rz / pq;
    rnorm2 = 0.0;
    for (int n = 0; n < nn; ++n) {
      u[n] += alpha * pv[n];
      r[n] -= alpha * q[n];
      rnorm2 += r[n] * r[n];
    }
    ++it;
  }
  free(K);
  free(rhs);
  free(r);
  free(z);
  free(pv);
  free(q);
  free(offs);
  return it;
}

void oq_reconstruct(int dim, int nc, int n_refine, const double *u_coarse,
                    const double *Psi, double *u_fine) {
  const int nv = 1 << dim, ncell = ipow(nc, dim);
  const int nn = ipow((1 << n_refine) + 1, dim);
  int nodes[8];
  for (int c = 0; c < ncell; ++c) {
    coarse_cell_nodes(dim, nc, c, nodes);
    double *uf = u_fine + (size_t)c * nn;
    for (int n = 0; n < nn; ++n) uf[n] = 0.0;
    /* global_solution.sadd(1, w_k, psi_k) for k = 0..nv-1 */
    for (int k = 0; k < nv; ++k) {
      const double w = u_coarse[nodes[k]];
      const double *pk = Psi + ((size_t)c * nv + k) * nn;
      for (int n = 0; n < nn; ++n) uf[n] = uf[n] + w * pk[n];
    }
  }
}

/* ---------------------------------------------------------------------------
 * Adjoint gradient of the coarse misfit (no reference counterpart: the
 * differentiation target of the MI355X build).  Scalar coefficient a = exp(g)
 * per fine cell, structured nc^dim coarse grid as in oq_coarse_solve.
 *   J(g) = 1/2 sum over free coarse nodes (u_H - u_target)^2,
 *   K_H v = u_H - u_target (adjoint, same constraints),
 *   dJ/da_e = -vt_e^T Khat_e (ut_e + w_e),  dJ/dg_e = a_e dJ/da_e,
 * with ut = sum_k u_k psi_k and vt = sum_k v_k psi_k the fine reconstructions
 * of the coarse solution and adjoint, w = K_II^{-1} F_I (zero on the cell
 * boundary) and Khat_e the fine element stiffness of a unit coefficient.
 * (dM_ij/da_e = psi_i^T Khat_e psi_j because the bases are K-harmonic, and
 * db_i/da_e = -w^T Khat_e psi_i.)  Pinned by central finite differences of
 * oq_misfit (tests/test_oracle.py).
 * ------------------------------------------------------------------------- */
static int coarse_all(const oq_problem *p, int nc, const double *verts, const double *coef,
                      const double *f, double *M, double *b, double *Psi, double *W) {
  const int dim = p->dim, nv = 1 << dim, N = 1 << p->n_refine;
  const int nfc = ipow(N, dim), nn = ipow(N + 1, dim), ncell = ipow(nc, dim);
  int worst = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) reduction(max : worst)
#endif
  for (int c = 0; c < ncell; ++c) {
    int it = 0;
    const int st = basis_cell(p, verts + (size_t)c * nv * dim, coef + (size_t)c * nfc,
                              (p->rhs_mode == OQ_RHS_CELL && f) ? f + (size_t)c * nfc : NULL,
                              NULL, OQ_SOLVER_DIRECT, M + (size_t)c * nv * nv,
                              b + (size_t)c * nv, Psi + (size_t)c * nv * nn, &it,
                              W ? W + (size_t)c * nn : NULL);
    if (st > worst) worst = st;
  }
  return worst;
}

int oq_misfit(const oq_problem *p, int nc, const double *verts, const double *coef,
              const double *f, const double *u_target, double rtol, double *J, double *u) {
  if (p->coef_mode != OQ_COEF_SCALAR_CELL) return -1;
  const int dim = p->dim, nv = 1 << dim, N = 1 << p->n_refine;
  const int nn = ipow(N + 1, dim), ncell = ipow(nc, dim), ncn = ipow(nc + 1, dim);
  double *M = (double *)malloc(sizeof(double) * (size_t)ncell * nv * nv);
  double *b = (double *)malloc(sizeof(double) * (size_t)ncell * nv);
  double *Psi = (double *)malloc(sizeof(double) * (size_t)ncell * nv * nn);
  int st = coarse_all(p, nc, verts, coef, f, M, b, Psi, NULL);
  if (st == 0 && oq_coarse_solve(dim, nc, M, b, rtol, 100000, u) < 0) st = 1;
  double s = 0.0;
  for (int n = 0; n < ncn; ++n)
    if (!coarse_is_boundary(dim, nc, n)) s += 0.5 * (u[n] - u_target[n]) * (u[n] - u_target[n]);
  *J = s;
  free(M);
  free(b);
  free(Psi);
  return st;
}

int oq_adjoint_gradient(const oq_problem *p, int nc, const double *verts, const double *coef,
                        const double *f, const double *u_target, double rtol, double *J,
                        double *grad) {
  if (p->coef_mode != OQ_COEF_SCALAR_CELL || p->reaction_mode != OQ_REACTION_NONE) return -1;
  const int dim = p->dim, nv = 1 << dim, N = 1 << p->n_refine, S = N + 1;
  const int nfc = ipow(N, dim), nn = ipow(S, dim), ncell = ipow(nc, dim), ncn = ipow(nc + 1, dim);
  double *M = (double *)malloc(sizeof(double) * (size_t)ncell * nv * nv);
  double *b = (double *)malloc(sizeof(double) * (size_t)ncell * nv);
  double *Psi = (double *)malloc(sizeof(double) * (size_t)ncell * nv * nn);
  double *W = (double *)malloc(sizeof(double) * (size_t)ncell * nn);
  double *u = (double *)malloc(sizeof(double) * ncn);
  double *v = (double *)malloc(sizeof(double) * ncn);
  double *d = (double *)malloc(sizeof(double) * ncn);
  int st = coarse_all(p, nc, verts, coef, f, M, b, Psi, W);
  if (st == 0 && oq_coarse_solve(dim, nc, M, b, rtol, 100000, u) < 0) st = 1;
  double s = 0.0;
  for (int n = 0; n < ncn; ++n) {
    const int bd = coarse_is_boundary(dim, nc, n);
    d[n] = bd ? 0.0 : u[n] - u_target[n];
    s += 0.5 * d[n] * d[n];
  }
  *J = s;
  if (st == 0 && oq_coarse_solve_rhs(dim, nc, M, d, rtol, 100000, v) < 0) st = 1;
  if (st == 0) {
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int c = 0; c < ncell; ++c) {
      int nodes[8];
      coarse_cell_nodes(dim, nc, c, nodes);
      const double *vc = verts + (size_t)c * nv * dim;
      double h[3] = {1.0, 1.0, 1.0};
      for (int dd = 0; dd < dim; ++dd) h[dd] = (vc[(nv - 1) * dim + dd] - vc[dd]) / N;
      /* unit-coefficient element stiffness */
      double Aq[8 * 6] = {0}, fq[8] = {0}, Kh[64], Fe[8];
      const int ncomp = (dim == 2) ? 3 : 6;
      for (int q = 0; q < nv; ++q) {
        if (dim == 2) {
          Aq[q * ncomp + 0] = 1.0;
          Aq[q * ncomp + 2] = 1.0;
        } else {
          Aq[q * ncomp + 0] = 1.0;
          Aq[q * ncomp + 3] = 1.0;
          Aq[q * ncomp + 5] = 1.0;
        }
      }
      oq_element_matrix(dim, h, Aq, fq, 0.0, Kh, Fe);
      double *ut = (double *)calloc((size_t)nn, sizeof(double));
      double *vt = (double *)calloc((size_t)nn, sizeof(double));
      for (int k = 0; k < nv; ++k) {
        const double *pk = Psi + ((size_t)c * nv + k) * nn;
        for (int n = 0; n < nn; ++n) {
          ut[n] += u[nodes[k]] * pk[n];
          vt[n] += v[nodes[k]] * pk[n];
        }
      }
      const double *wc = W + (size_t)c * nn;
      for (int e = 0; e < nfc; ++e) {
        int cijk[3] = {0, 0, 0}, r = e;
        for (int dd = 0; dd < dim; ++dd) {
          cijk[dd] = r % N;
          r /= N;
        }
        double ve[8], ue[8];
        for (int a = 0; a < nv; ++a) {
          int na = 0, mul = 1;
          for (int dd = 0; dd < dim; ++dd) {
            na += (cijk[dd] + ((a >> dd) & 1)) * mul;
            mul *= S;
          }
          ve[a] = vt[na];
          ue[a] = ut[na] + wc[na];
        }
        double t = 0.0;
        for (int a = 0; a < nv; ++a)
          for (int bb = 0; bb < nv; ++bb) t += ve[a] * Kh[a * nv + bb] * ue[bb];
        grad[(size_t)c * nfc + e] = -coef[(size_t)c * nfc + e] * t;
      }
      free(ut);
      free(vt);
    }
  }
  free(M);
  free(b);
  free(Psi);
  free(W);
  free(u);
  free(v);
  free(d);
  return st;
}
