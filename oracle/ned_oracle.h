/*
 * FP64 CPU oracle for the 2-D multiscale Q1 x Nedelec (Q_Ned) basis.
 *
 * TEST INFRASTRUCTURE ONLY, like msfec_oracle.h: only tests/, smoke() and the
 * cpu_baseline leg of bench.py load it, to check results against.
 *
 * The reference (source/Q_Ned/q_ned_basis.cc) is three-dimensional.  This is
 * its two-dimensional restatement on one axis-aligned coarse box:
 *   setup_grid            q_ned_basis.cc:164-175   general_cell + refine_global(l)
 *   setup_basis_dofs_h1   q_ned_basis.cc:223-294   sigma = phi_n, u x n = grad phi_n
 *   setup_basis_dofs_curl q_ned_basis.cc:296-357   sigma = 0,     u x n = N_E
 *   assemble_system       q_ned_basis.cc:360-567   mixed form, QGauss(degree + 2)
 *   solve_direct          q_ned_basis.cc:570-625   condense + direct solve
 *   assemble_global_element_matrix  q_ned_basis.cc:841-939
 *   BasisNedelec          include/functions/basis_nedelec.tpp (covariant map)
 * In 2-D the curl of a vector field is the scalar d_x u_y - d_y u_x and the
 * curl-curl coefficient A is a scalar per fine cell.
 */
#ifndef NED_ORACLE_H
#define NED_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Fine degrees of freedom of one cell, m = 2^l fine cells per direction:
 *   sigma: nodes (i, j), 0 <= i, j <= m, index j (m + 1) + i
 *   u    : edges, the line integral of the tangential component, tangent +x
 *          on horizontal and +y on vertical edges;
 *          horizontal edge (i, j) [nodes (i,j)-(i+1,j)] at ns + j m + i,
 *          vertical edge (i, j)   [nodes (i,j)-(i,j+1)] at ns + nh + j (m + 1) + i
 * with ns = (m+1)^2, nh = m (m+1), nv = (m+1) m.  Bases 0..3 are the vertex
 * bases (deal.II vertex order), 4..7 the edge bases (deal.II line order:
 * x = x0, x = x1, y = y0, y = y1).
 *
 * Inputs per fine cell (m x m, lexicographic, x fastest): a (curl-curl
 * coefficient), binv (B^-1), r (reaction), f (2 components).  full_rhs adds
 * the source (v, grad phi_n) to the vertex bases (q_ned_basis.cc:498-505).
 * Outputs: M (8 x 8, row = test basis), b (8), Psi (8 x (ns + nu); NULL when
 * not wanted).  Returns 0, or 2 when the interior system is singular.
 */
int oq_ned_basis_cell(int l, const double *verts, const double *a, const double *binv,
                      const double *r, const double *f, int full_rhs, double *M, double *b,
                      double *Psi);

/* The assembled full system (ns + nu)^2 (row = test function) and the
 * global right-hand side (v, f), for the assembly checks. */
int oq_ned_assemble(int l, const double *verts, const double *a, const double *binv,
                    const double *r, const double *f, double *A, double *F);

#ifdef __cplusplus
}
#endif

#endif
