/*
 * FP64 CPU oracle for the 3-D multiscale Nedelec x Raviart-Thomas (Ned_RT)
 * basis.
 *
 * TEST INFRASTRUCTURE ONLY, like msfec_oracle.h: only tests/, smoke() and the
 * cpu_baseline leg of bench.py load it, to check results against.
 *
 * Restated from source/Ned_RT/ned_rt_basis.cc on one axis-aligned coarse box,
 * lowest order (FE_Nedelec<3>(0) x FE_RaviartThomas<3>(0)):
 *   setup_grid             ned_rt_basis.cc:167-178  general_cell + refine_global(l)
 *   setup_basis_dofs_curl  ned_rt_basis.cc:225-293  sigma x n = coarse Nedelec
 *                                                   function n, u . n = 0
 *   setup_basis_dofs_div   ned_rt_basis.cc:295-359  sigma x n = 0, u . n = coarse
 *                                                   Raviart-Thomas function n
 *   assemble_system        ned_rt_basis.cc:361-577  (A^-1 sigma, tau) - (u, curl tau)
 *                                                   + (v, curl sigma) + (B div u, div v)
 *                                                   + (R u, v), QGauss(2); loads
 *                                                   (v, f) and, for the 12 edge
 *                                                   bases, (v, curl N_n) (full_rhs)
 *   run                    ned_rt_basis.cc:1280-1352 condense, solve_direct
 *   solve_direct           ned_rt_basis.cc:579-634
 *   assemble_global_element_matrix  ned_rt_basis.cc:850-948 (block rule in nrt_oracle.c)
 *   BasisNedelec / BasisNedelecCurl / BasisRaviartThomas
 *       include/functions/basis_nedelec.tpp, basis_nedelec_curl.tpp,
 *       basis_raviart_thomas.tpp (covariant, curl and contravariant maps)
 */
#ifndef NRT_ORACLE_H
#define NRT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Fine degrees of freedom of one coarse box, m = 2^l fine cells per
 * direction, fine cell (a, b, c) at (c m + b) m + a:
 *   edges (sigma): the line integral of the tangential component along +x,
 *     +y or +z (the node functional of FE_Nedelec(0) on an axis-aligned box);
 *     x-edge (i, j, k) [nodes (i,j,k)-(i+1,j,k)] at (k (m+1) + j) m + i,
 *     y-edge (i, j, k) [nodes (i,j,k)-(i,j+1,k)] at nx + (k m + j) (m+1) + i,
 *     z-edge (i, j, k) [nodes (i,j,k)-(i,j,k+1)] at 2 nx + (k (m+1) + j) (m+1) + i,
 *     nx = m (m+1)^2, ne = 3 nx;
 *   faces (u): the flux in the +x, +y or +z direction (the node functional of
 *     FE_RaviartThomas(0));
 *     x-face (i, j, k) [x = x0 + i hx, cell row (j, k)] at ne + (k m + j) (m+1) + i,
 *     y-face (i, j, k) [y = y0 + j hy]                   at ne + nf1 + (k (m+1) + j) m + i,
 *     z-face (i, j, k) [z = z0 + k hz]                   at ne + 2 nf1 + (k m + j) m + i,
 *     nf1 = m^2 (m+1), nf = 3 nf1.
 * Bases 0..11 are the edge bases in deal.II line order, 12..17 the face
 * bases in deal.II face order (x = x0, x = x1, y = y0, y = y1, z = z0,
 * z = z1).  The element matrix is 18 x 18, row = test basis.
 *
 * Inputs per fine cell and QGauss(2) point (8 points, x fastest):
 *   ainv: 6 values (xx, xy, xz, yy, yz, zz) of DiffusionInverse_A;
 *   bq:   Diffusion_B;
 *   rq:   ReactionRate (NULL: zero, as in eqn_coeff_R.cc:8-14);
 *   fq:   3 components of the right-hand side (NULL: zero).
 * Outputs: M (18 x 18), b (18), X (18 x (ne + nf), the full fine vectors of
 * the bases; NULL when not wanted).  Returns 0, 1 for a cell that is not an
 * axis-aligned box or l outside 1..3, 2 when the interior system is singular,
 * 3 when memory runs out. */
int oq_nrt_basis_cell(int l, const double *verts, const double *ainv, const double *bq,
                      const double *rq, const double *fq, double *M, double *b, double *X);

/* The assembled matrix of assemble_system over all fine dofs ((ne + nf)^2,
 * row = test function) and the load (v, f) (zero on the edges); l in 1..4
 * (l = 4: 5.8 GB). */
int oq_nrt_assemble(int l, const double *verts, const double *ainv, const double *bq,
                    const double *rq, const double *fq, double *A, double *F);

/* The loads (v, curl N_n) of the 12 edge bases (12 x (ne + nf), zero on the
 * edges), by the same quadrature as assemble_system; l in 1..4. */
int oq_nrt_curl_loads(int l, const double *verts, double *Fc);
/* The same assembly as (row, column, value) triplets, 18 x 18 per fine cell
 * in fine-cell order (duplicates summed by the reader), l in 1..4. */
int oq_nrt_assemble_coo(int l, const double *verts, const double *ainv, const double *bq,
                        const double *rq, const double *fq, int *rows, int *cols, double *vals,
                        double *F);
/* The boundary dofs (on_boundary, N ints) and the fixed values of the 18
 * bases on them (G, 18 x N, zero inside), l in 1..4. */
int oq_nrt_boundary(int l, const double *verts, double *G, int *on_boundary);
/* RT_DQ in three dimensions (source/RT_DQ/rt_dq_basis.cc) on the same grid:
 * dofs the fine faces (numbered as above without the edge offset, nf of
 * them), then the fine cells.  ainv as above; fq: the scalar load at the 8
 * points of every fine cell (NULL: zero); no reaction.  Outputs M (7 x 7: the
 * 6 face bases in deal.II face order, then the DQ0 basis), b (7) and Sigma
 * (6 x nf, the fine fluxes of the face bases; their u is 1).  l in 1..3. */
int oq_rt3_assemble(int l, const double *verts, const double *ainv, const double *fq, double *A,
                    double *F);
int oq_rt3_assemble_coo(int l, const double *verts, const double *ainv, const double *fq,
                        int *rows, int *cols, double *vals, double *F);
int oq_rt3_basis_cell(int l, const double *verts, const double *ainv, const double *fq,
                      double *M, double *b, double *Sigma);

#ifdef __cplusplus
}
#endif

#endif /* NRT_ORACLE_H */
