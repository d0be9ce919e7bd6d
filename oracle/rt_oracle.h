/*
 * FP64 CPU oracle for the 2-D multiscale Raviart-Thomas x DQ0 (RT_DQ) basis.
 *
 * TEST INFRASTRUCTURE ONLY, like msfec_oracle.h: only tests/, smoke() and the
 * cpu_baseline leg of bench.py load it, to check results against.
 *
 * The reference (source/RT_DQ/rt_dq_basis.cc) is three-dimensional.  This is
 * its two-dimensional restatement on one axis-aligned coarse box, lowest
 * order (FE_RaviartThomas(0) x FE_DGQ(0)):
 *   setup_grid            rt_dq_basis.cc:153-216   general_cell + refine_global(l)
 *   setup_basis_dofs_div  rt_dq_basis.cc:267-380   normal flux of the coarse RT
 *                                                  basis n on the boundary
 *   BasisRaviartThomas    include/functions/basis_raviart_thomas.tpp (Piola map)
 *   assemble_system       rt_dq_basis.cc:391-579   (sigma, A^-1 tau) - (div tau, u)
 *                                                  + (v, div sigma), QGauss(2);
 *                                                  local rhs -+1/|K| (is_laplace)
 *   run                   rt_dq_basis.cc:1114-1170 copy, condense, solve
 *   solve_direct          rt_dq_basis.cc:582-612
 *   set_u_to_std          rt_dq_basis.cc:1039-1043 u block of every basis := 1
 *   assemble_global_element_matrix  rt_dq_basis.cc:822-914
 * ReactionRate is zero (source/equation_data/eqn_coeff_R.cc:8-14).  A is
 * either a I on each fine cell (DiffusionInverse_A is then 1/a) or a
 * symmetric tensor given at the quadrature points.
 */
#ifndef RT_ORACLE_H
#define RT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Fine degrees of freedom of one cell, m = 2^l fine cells per direction:
 *   sigma: the flux through each fine edge in the +x (vertical edges) or +y
 *          (horizontal edges) direction, the node functional of
 *          FE_RaviartThomas(0) on an axis-aligned box;
 *          vertical edge (p, q)   [x = x0 + p hx, cell row q] at q (m + 1) + p,
 *          horizontal edge (p, q) [cell column p, y = y0 + q hy] at nv + q m + p,
 *          nv = m (m + 1), ne = 2 m (m + 1);
 *   u    : the value on fine cell (p, q) at ne + q m + p.
 * Bases 0..3 belong to the coarse faces in deal.II face order (x = x0,
 * x = x1, y = y0, y = y1); the element matrix is 5 x 5, index 4 the DQ0
 * basis, row = test basis.
 *
 * Inputs per fine cell (m x m, lexicographic, x fastest): a (diffusion) and f
 * (load).  Outputs: M (25), b (5), Sigma (4 x ne) and U (4 x m^2, the u of
 * the solve before set_u_to_std, shifted to mean zero like the reference's
 * Schur-complement solution); Sigma and U may be NULL.  Returns 0, 1 for a
 * cell that is not an axis-aligned box, 2 when the condensed system is
 * singular. */
int oq_rt_basis_cell(int l, const double *verts, const double *a, const double *f, double *M,
                     double *b, double *Sigma, double *U);

/* The same with the diffusion tensor A = (xx, xy, yy) at the 4 QGauss(2)
 * points of every fine cell (a_qp: m^2 x 4 x 3, points x fastest); the
 * solve uses its inverse at each point. */
int oq_rt_basis_cell_qp(int l, const double *verts, const double *a_qp, const double *f,
                        double *M, double *b, double *Sigma, double *U);

/* The assembled matrix of assemble_system over all fine dofs (ne + m^2 rows,
 * row = test function) and the load (v, f) (zero on the edges). */
int oq_rt_assemble(int l, const double *verts, const double *a, const double *f, double *A,
                   double *F);
int oq_rt_assemble_qp(int l, const double *verts, const double *a_qp, const double *f, double *A,
                      double *F);

#ifdef __cplusplus
}
#endif

#endif /* RT_ORACLE_H */
