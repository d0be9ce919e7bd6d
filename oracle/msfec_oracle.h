/*
 * FP64 CPU oracle for the multiscale Q1 basis construction.
 *
 * TEST INFRASTRUCTURE ONLY.  This file and msfec_oracle.c restate, in plain C,
 * the reference algorithm of MsFEC's Q::QBasis / Q::QMultiscale so that the
 * HIP product path can be checked against it.  Only tests/, the smoke() entry
 * and the cpu_baseline leg of bench.py may load it.  The product library never
 * links or calls it.
 *
 * Reference files restated here (paths relative to the MsFEC repository):
 *   include/functions/basis_q1.tpp:43-152     BasisQ1 (monomial-matrix inverse)
 *   source/Q/q_basis.cc:95-102                make_grid (general_cell + refine)
 *   source/Q/q_basis.cc:117-170               setup_system (Dirichlet lifting)
 *   source/Q/q_basis.cc:175-257               assemble_system (QGauss(2))
 *   source/Q/q_basis.cc:304-357               solve_direct / solve_iterative
 *   source/Q/q_basis.cc:262-300               assemble_global_element_matrix
 *   source/Q/q_basis.cc:394-416               set_global_weights
 *   source/Q/q_basis.cc:518-603               run
 *   source/equation_data/eqn_coeff_A.cc:7-191 Diffusion_A (analytic tensor)
 *   source/Q/q_global.cc:92-193,198-297,327-392  coarse grid, assembly, CG
 *
 * Pinning (what checks each function against something other than itself) is
 * listed per function in msfec_oracle.c and exercised by tests/test_oracle.py.
 */
#ifndef MSFEC_ORACLE_H
#define MSFEC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
  OQ_COEF_SCALAR_CELL = 0, /* one scalar a per fine cell: A = a I            */
  OQ_COEF_TENSOR_QP = 1,   /* symmetric tensor per fine cell per Gauss point  */
  OQ_COEF_ANALYTIC = 2     /* Diffusion_A of eqn_coeff_A.cc (dim 3 only)      */
};
enum { OQ_RHS_CONST = 0, OQ_RHS_CELL = 1, OQ_RHS_QP = 2 };
enum { OQ_REACTION_NONE = 0, OQ_REACTION_CONST = 1, OQ_REACTION_CELL = 2 };
enum { OQ_SOLVER_DIRECT = 0, OQ_SOLVER_CG_SSOR = 1 };

typedef struct {
  int dim;            /* 2 or 3                                              */
  int n_refine;       /* local refinements l: 2^l fine cells per direction   */
  int coef_mode;      /* OQ_COEF_*                                           */
  int rhs_mode;       /* OQ_RHS_*                                            */
  double rhs_const;   /* f for OQ_RHS_CONST                                  */
  int reaction_mode;  /* OQ_REACTION_* (0 reproduces the reference)          */
  double reaction_const;
  /* Diffusion_A data (eqn_coeff_A.cc:60-140): frequency, scale, alpha, rotate */
  int freq[3];
  double scale[3];
  double alpha[3];
  int rotate;
  /* CG-SSOR controls (q_basis.cc:337-343): maxit 1000, abs tol 1e-12, omega 1.6 */
  int cg_maxit;
  double cg_tol;
  double ssor_omega;
} oq_problem;

/* Per-cell input layout (cell-major, fine cells lexicographic, x fastest):
 *   verts : 2^dim vertices x dim coordinates, deal.II vertex order (bit d of the
 *           vertex index selects the low/high side in direction d)
 *   coef  : SCALAR_CELL -> N^dim doubles; TENSOR_QP -> N^dim x 2^dim x ncomp with
 *           ncomp = 3 (xx,xy,yy) in 2-D or 6 (xx,xy,xz,yy,yz,zz) in 3-D
 *   f     : RHS_CELL -> N^dim; RHS_QP -> N^dim x 2^dim
 *   reac  : REACTION_CELL -> N^dim
 * Outputs per cell: M (nb x nb, row = test basis), b (nb), Psi (nb x (N+1)^dim;
 * NULL when not wanted),
 * nb = 2^dim, fine nodes lexicographic x fastest.
 * Returns 0 on success, 1 if CG did not converge, 2 if a pivot was not positive,
 * -1 on invalid input. */
int oq_basis_cell(const oq_problem *prob, const double *verts,
                  const double *coef, const double *f, const double *reac,
                  int solver, double *M, double *b, double *Psi, int *iters);

/* Batched driver over n_cells cells, OpenMP-parallel over cells with
 * n_threads threads (<=0: OpenMP default).  Strides follow the layout above.
 * status[c] receives the per-cell return code, iters[c] the max CG iteration
 * count over the 2^dim solves (0 for the direct solver). */
int oq_basis_batch(const oq_problem *prob, int n_cells, const double *verts,
                   const double *coef, const double *f, const double *reac,
                   int solver, int n_threads, double *M, double *b, double *Psi,
                   int *iters, int *status);

/* BasisQ1 coefficient matrix (basis_q1.tpp:43-67): C[m*nv + i] is the
 * coefficient of monomial m in phi_i, obtained by inverting the vertex-monomial
 * matrix.  Monomials: 2-D {1,x,y,xy}, 3-D {1,x,y,z,xy,yz,xz,xyz}. */
int oq_basis_q1_coeffs(int dim, const double *verts, double *C);
double oq_basis_q1_value(int dim, const double *C, int index, const double *x);

/* Diffusion_A::value_list at one point (eqn_coeff_A.cc:163-191), row-major 3x3 */
void oq_diffusion_A(const oq_problem *prob, const double *x, double *A);

/* Fine-cell element matrix and load on an axis-aligned box of size h[dim] by
 * QGauss(2) quadrature, exactly as q_basis.cc:222-237 (plus the optional
 * reaction term).  Aq: per Gauss point symmetric tensor (ncomp entries),
 * fq: per Gauss point load, rq: reaction. Ke is nv x nv row-major. */
void oq_element_matrix(int dim, const double *h, const double *Aq,
                       const double *fq, double rq, double *Ke, double *Fe);

/* ---------------- coarse problem (q_global.cc) ---------------------------- */
/* Structured coarse grid on [0,1]^dim with nc cells per direction, cells and
 * nodes lexicographic (x fastest).  Homogeneous Dirichlet data on all faces
 * (q_global.cc:123-139).  Assembles the element matrices/RHS (q_global.cc:
 * 226-296) and solves by Jacobi-PCG to relative residual rtol (the reference
 * uses AMG-CG to abs 1e-12, q_global.cc:327-363; both converge to the same
 * solution).  u receives (nc+1)^dim nodal values (0 on the boundary).
 * Returns the iteration count, or -1 on non-convergence. */
int oq_coarse_solve(int dim, int nc, const double *M, const double *b,
                    double rtol, int maxit, double *u);

/* Coarse vertex weights per cell (send_global_weights_to_cell, q_global.cc:
 * 368-392) followed by set_global_weights (q_basis.cc:394-416):
 * u_fine[c][node] = sum_k w_c[k] Psi_c[k][node]. */
void oq_reconstruct(int dim, int nc, int n_refine, const double *u_coarse,
                    const double *Psi, double *u_fine);

/* The same coarse system with a nodal right-hand side (boundary entries are
 * ignored): the adjoint solve. */
int oq_coarse_solve_rhs(int dim, int nc, const double *M, const double *rhs_nodes,
                        double rtol, int maxit, double *u);

/* ---------------- coarse misfit and its adjoint gradient ------------------ */
/* J = 1/2 sum over free coarse nodes (u_H - u_target)^2 for a scalar
 * coefficient per fine cell (cells of the nc^dim grid in lexicographic order,
 * direct local solves, coarse CG to rtol).  u receives u_H.  Returns 0 on
 * success. */
int oq_misfit(const oq_problem *p, int nc, const double *verts, const double *coef,
              const double *f, const double *u_target, double rtol, double *J, double *u);
/* dJ/dg per fine cell for a = exp(g) (layout of coef); J as in oq_misfit.
 * No reaction term.  See msfec_oracle.c for the formula. */
int oq_adjoint_gradient(const oq_problem *p, int nc, const double *verts, const double *coef,
                        const double *f, const double *u_target, double rtol, double *J,
                        double *grad);

#ifdef __cplusplus
}
#endif
#endif
